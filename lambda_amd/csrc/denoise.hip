// Film denoiser: the project's replacement for the OIDN filter of Denoise::SetData/Process
// (Denoise.cpp:13-67), which takes the colour image with albedo and normal auxiliary images.
// OIDN's trained network is not available, so the filter is the edge-avoiding A-Trous wavelet
// transform (Dammertz, Sewtz, Hanika, Lensch 2010): `iterations` passes of the 5x5 B3-spline
// kernel h = [1/16, 1/4, 3/8, 1/4, 1/16] (x) itself with taps 2^i pixels apart, each tap weighted
// by exp(-|dc|^2 / sc_i^2) exp(-|dn|^2 / sn^2) exp(-|da|^2 / sa^2) (colour, normal, albedo
// differences to the centre pixel; sc_i = sc 2^-i, a sigma of 0 drops that term). With an albedo
// image and `demodulate`, the filter runs on colour / albedo and multiplies the albedo back
// (Schied et al. 2017), so texture detail is not blurred. Each pass is an HBM-bound stencil:
// 16 B written per pixel, its 25 taps x 48 B read mostly from cache.
#include "host.h"

using namespace lgpu;

namespace {

constexpr float kAlbedoEps = 1e-3f;  // albedo channels below it are not divided out

__global__ void k_dn_mean(const double *sum, const uint32_t *count, float4 *out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = count[i];
  const double inv = c ? 1.0 / (double)c : 0.0;
  out[i] = make_float4((float)(sum[3 * i] * inv), (float)(sum[3 * i + 1] * inv), (float)(sum[3 * i + 2] * inv), 0.f);
}

__device__ inline float Demod(float c, float a) { return a > kAlbedoEps ? c / a : c; }
__device__ inline float Remod(float c, float a) { return a > kAlbedoEps ? c * a : c; }

__global__ void k_dn_demod(float4 *c, const float4 *a, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 v = c[i], al = a[i];
  c[i] = make_float4(Demod(v.x, al.x), Demod(v.y, al.y), Demod(v.z, al.z), 0.f);
}

struct AtrousParams {
  uint32_t w, h, step;
  float invSc2, invSn2, invSa2;  // 1 / sigma^2 of this pass, 0 = term off
  uint32_t hasN, hasA;
};

__device__ inline float Dist2(float4 a, float4 b) {
  const float x = a.x - b.x, y = a.y - b.y, z = a.z - b.z;
  return x * x + y * y + z * z;
}

// One A-Trous pass; taps outside the image are skipped (the weights renormalise).
__global__ void __launch_bounds__(256) k_dn_atrous(const float4 *__restrict__ cin, const float4 *__restrict__ alb,
                                                   const float4 *__restrict__ nrm, AtrousParams p, float4 *__restrict__ cout) {
  const uint32_t x = blockIdx.x * 16 + threadIdx.x % 16, y = blockIdx.y * 16 + threadIdx.x / 16;
  if (x >= p.w || y >= p.h) return;
  const float kw[3] = {0.375f, 0.25f, 0.0625f};
  const size_t pi = (size_t)y * p.w + x;
  const float4 cp = cin[pi];
  const float4 np = p.hasN ? nrm[pi] : make_float4(0.f, 0.f, 0.f, 0.f);
  const float4 ap = p.hasA ? alb[pi] : make_float4(0.f, 0.f, 0.f, 0.f);
  float sx = 0.f, sy = 0.f, sz = 0.f, sw = 0.f;
  for (int dy = -2; dy <= 2; ++dy) {
    const int qy = (int)y + dy * (int)p.step;
    if (qy < 0 || qy >= (int)p.h) continue;
    for (int dx = -2; dx <= 2; ++dx) {
      const int qx = (int)x + dx * (int)p.step;
      if (qx < 0 || qx >= (int)p.w) continue;
      const size_t qi = (size_t)qy * p.w + (size_t)qx;
      const float4 cq = cin[qi];
      float e = Dist2(cq, cp) * p.invSc2;
      if (p.hasN) e += Dist2(nrm[qi], np) * p.invSn2;
      if (p.hasA) e += Dist2(alb[qi], ap) * p.invSa2;
      const float wt = kw[dx < 0 ? -dx : dx] * kw[dy < 0 ? -dy : dy] * expf(-e);
      sx += wt * cq.x; sy += wt * cq.y; sz += wt * cq.z; sw += wt;
    }
  }
  // the centre tap has weight 9/64, so sw > 0
  cout[pi] = make_float4(sx / sw, sy / sw, sz / sw, 0.f);
}

__global__ void k_dn_to_film(const float4 *c, const float4 *a, uint32_t remod, uint32_t n, double *sum, uint32_t *count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 v = c[i];
  if (remod) {
    const float4 al = a[i];
    v = make_float4(Remod(v.x, al.x), Remod(v.y, al.y), Remod(v.z, al.z), 0.f);
  }
  sum[3 * (size_t)i] = v.x; sum[3 * (size_t)i + 1] = v.y; sum[3 * (size_t)i + 2] = v.z;
  count[i] = 1u;
}

}  // namespace

extern "C" {

lgpu_status lgpu_film_denoise(lgpu_film colour, lgpu_film albedo, lgpu_film normal, const lgpu_denoise_desc *d,
                              lgpu_film out) {
  if (!colour || !out || !d) return Fail(LGPU_ERR_INVALID_ARGUMENT, "film_denoise: null colour, output or descriptor");
  for (lgpu_film f : {albedo, normal, out}) {
    if (!f) continue;
    if (f->dev != colour->dev) return Fail(LGPU_ERR_INVALID_ARGUMENT, "film_denoise: films on different devices");
    if (f->width != colour->width || f->height != colour->height)
      return Fail(LGPU_ERR_INVALID_ARGUMENT, "film_denoise: films of different sizes");
  }
  if (d->iterations == 0 || d->iterations > 12) return Fail(LGPU_ERR_INVALID_ARGUMENT, "film_denoise: iterations must be 1..12");
  for (float s : {d->sigma_colour, d->sigma_normal, d->sigma_albedo})
    if (!(s >= 0.f) || !std::isfinite(s)) return Fail(LGPU_ERR_INVALID_ARGUMENT, "film_denoise: sigmas must be finite and >= 0");
  const uint32_t w = colour->width, h = colour->height, n = w * h;
  if (!n) return LGPU_OK;
  std::lock_guard<std::mutex> lk(colour->dev->mu);
  LGPU_HIP(hipSetDevice(colour->dev->device));
  (void)hipGetLastError();
  const hipStream_t st = colour->dev->stream;
  DevBuf ca, cb, ab, nb;
  LGPU_HIP(ca.alloc(16ull * n));
  LGPU_HIP(cb.alloc(16ull * n));
  const dim3 g1((n + 255) / 256), b1(256);
  hipLaunchKernelGGL(k_dn_mean, g1, b1, 0, st, colour->sum.as<double>(), colour->count.as<uint32_t>(), ca.as<float4>(), n);
  if (albedo) {
    LGPU_HIP(ab.alloc(16ull * n));
    hipLaunchKernelGGL(k_dn_mean, g1, b1, 0, st, albedo->sum.as<double>(), albedo->count.as<uint32_t>(), ab.as<float4>(), n);
  }
  if (normal) {
    LGPU_HIP(nb.alloc(16ull * n));
    hipLaunchKernelGGL(k_dn_mean, g1, b1, 0, st, normal->sum.as<double>(), normal->count.as<uint32_t>(), nb.as<float4>(), n);
  }
  const bool demod = albedo && d->demodulate;
  if (demod) hipLaunchKernelGGL(k_dn_demod, g1, b1, 0, st, ca.as<float4>(), ab.as<float4>(), n);
  float4 *src = ca.as<float4>(), *dst = cb.as<float4>();
  const dim3 g2((w + 15) / 16, (h + 15) / 16);
  for (uint32_t i = 0; i < d->iterations; ++i) {
    AtrousParams p;
    p.w = w; p.h = h; p.step = 1u << i;
    const float sc = d->sigma_colour * std::ldexp(1.f, -(int)i);
    p.invSc2 = sc > 0.f ? 1.f / (sc * sc) : 0.f;
    p.invSn2 = d->sigma_normal > 0.f ? 1.f / (d->sigma_normal * d->sigma_normal) : 0.f;
    p.invSa2 = d->sigma_albedo > 0.f ? 1.f / (d->sigma_albedo * d->sigma_albedo) : 0.f;
    p.hasN = normal ? 1u : 0u;
    p.hasA = albedo ? 1u : 0u;
    hipLaunchKernelGGL(k_dn_atrous, g2, dim3(256), 0, st, src, albedo ? ab.as<float4>() : nullptr,
                       normal ? nb.as<float4>() : nullptr, p, dst);
    std::swap(src, dst);
  }
  hipLaunchKernelGGL(k_dn_to_film, g1, b1, 0, st, src, demod ? ab.as<float4>() : nullptr, demod ? 1u : 0u, n,
                     out->sum.as<double>(), out->count.as<uint32_t>());
  const hipError_t le = hipGetLastError();
  const hipError_t se = hipStreamSynchronize(st);  // before the buffers leave scope
  LGPU_HIP(le);
  LGPU_HIP(se);
  return LGPU_OK;
}

}  // extern "C"
