// Post-processing of the film on the device and 8-bit PNG output: ConvolutionFilter
// (PostProcessing.cpp:16-45), ToneMap with the fitted ACES curve and clamp (ToneMap.cpp:5-73),
// the gamma and 8-bit packing of Texture::SaveToImageFile (Texture.cpp:19-32, 62-77), and a PNG
// writer (stored deflate blocks: no compression library needed). Every step is elementwise or a
// small stencil over the RGBA32F mean image: two HBM-bound kernels, 28 B read and 16 + 4 B written
// per pixel, plus the stencil's cached neighbour reads.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "host.h"

using namespace lgpu;

namespace {

// ACES fit input and output matrices (ToneMap.cpp:5-17), listed by rows.
struct Aces { float in[9], out[9]; };
__constant__ Aces kAces = {{0.59719f, 0.35458f, 0.04823f, 0.07600f, 0.90834f, 0.01566f, 0.02840f, 0.13383f, 0.83777f},
                           {1.60475f, -0.53108f, -0.07367f, -0.10208f, 1.10813f, -0.00605f, -0.00327f, -0.07276f, 1.07602f}};

// MulVec3Mat3x3 (ToneMap.cpp:19-24) sums the listed rows scaled by the vector's components,
// i.e. multiplies by the transpose (as written); corrected multiplies by the rows.
__device__ inline float3 MulMat(const float *m, float3 v, bool asWritten) {
  if (asWritten)
    return make_float3(m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z,
                       m[2] * v.x + m[5] * v.y + m[8] * v.z);
  return make_float3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
                     m[6] * v.x + m[7] * v.y + m[8] * v.z);
}
__device__ inline float Rrt(float v) {  // RRTAndOTDFit (ToneMap.cpp:29-33)
  const float a = v * (v + 0.0245786f) - 0.000090537f;
  const float b = v * (v * 0.983729f + 0.4329510f) + 0.238081f;
  return a / b;
}
// PackChannels (Texture.cpp:26-32). (int) of a float outside the int range (or NaN) is undefined;
// it is defined here as x86 computes it (INT_MIN), which the clamp then turns into 0.
__device__ inline int Pack(float c) {
  const float s = c * 256.f;
  if (!(s > -2147483648.f && s < 2147483648.f)) return 0;
  const int i = (int)s;
  return i < 0 ? 0 : (i > 255 ? 255 : i);
}

__global__ void k_film_mean(const double *sum, const uint32_t *count, float4 *out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = count[i];
  const double inv = c ? 1.0 / (double)c : 0.0;
  out[i] = make_float4((float)(sum[3 * i] * inv), (float)(sum[3 * i + 1] * inv), (float)(sum[3 * i + 2] * inv), 1.f);
}

struct PostParams {
  uint32_t w, h, kw, kh;
  float intensity;
  uint32_t tonemap, clamp, gamma, asWritten, conv;
};

__global__ void k_post(const float4 *img, const float *kern, PostParams p, float4 *rgba, uchar4 *rgba8) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.w * p.h) return;
  const uint32_t x = i % p.w, y = i / p.w;
  float3 c;
  if (p.conv) {  // ConvolutionFilter::Filter (PostProcessing.cpp:29-44)
    c = make_float3(0.f, 0.f, 0.f);
    const uint32_t kw2 = p.kw / 2, kh2 = p.kh / 2;
    for (uint32_t ky = 0; ky < p.kh; ++ky)
      for (uint32_t kx = 0; kx < p.kw; ++kx) {
        // unsigned coordinates as written: below zero they wrap and clamp to the far edge
        uint32_t sx = x + kx - kw2, sy = y + ky - kh2;
        if (p.asWritten) {
          sx = sx > p.w - 1 ? p.w - 1 : sx;
          sy = sy > p.h - 1 ? p.h - 1 : sy;
        } else {
          const int ix = (int)x + (int)kx - (int)kw2, iy = (int)y + (int)ky - (int)kh2;
          sx = (uint32_t)(ix < 0 ? 0 : (ix > (int)p.w - 1 ? (int)p.w - 1 : ix));
          sy = (uint32_t)(iy < 0 ? 0 : (iy > (int)p.h - 1 ? (int)p.h - 1 : iy));
        }
        const float wgt = kern[ky * p.kw + kx] * p.intensity;
        const float4 t = img[sy * p.w + sx];
        c.x += wgt * t.x;
        c.y += wgt * t.y;
        c.z += wgt * t.z;
      }
  } else {
    const float4 t = img[i];
    c = make_float3(t.x, t.y, t.z);
  }
  if (p.tonemap) {  // AcesFitted (ToneMap.cpp:53-58)
    c = MulMat(kAces.in, c, p.asWritten);
    c = make_float3(Rrt(c.x), Rrt(c.y), Rrt(c.z));
    c = MulMat(kAces.out, c, p.asWritten);
  }
  if (p.clamp) {  // ClampTextureChannels (ToneMap.cpp:45-52)
    c.x = fmaxf(fminf(c.x, 1.f), 0.f);
    c.y = fmaxf(fminf(c.y, 1.f), 0.f);
    c.z = fmaxf(fminf(c.z, 1.f), 0.f);
  }
  if (rgba) rgba[i] = make_float4(c.x, c.y, c.z, 1.f);
  if (rgba8) {
    if (p.gamma) {  // Gamma (Texture.cpp:19-23), exponent 1/2.2 as a float
      const float g = (float)(1.0 / 2.2);
      c = make_float3(powf(c.x, g), powf(c.y, g), powf(c.z, g));
    }
    rgba8[i] = make_uchar4((unsigned char)Pack(c.x), (unsigned char)Pack(c.y), (unsigned char)Pack(c.z), 255);
  }
}

uint32_t Crc32(const unsigned char *d, size_t n, uint32_t c = 0) {  // PNG CRC (ISO 3309 polynomial)
  c = ~c;
  for (size_t i = 0; i < n; ++i) {
    c ^= d[i];
    for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (0xEDB88320u & (0u - (c & 1u)));
  }
  return ~c;
}

void Be32(std::vector<unsigned char> &o, uint32_t v) {
  o.push_back((unsigned char)(v >> 24)); o.push_back((unsigned char)(v >> 16));
  o.push_back((unsigned char)(v >> 8)); o.push_back((unsigned char)v);
}

void Chunk(std::vector<unsigned char> &o, const char *type, const std::vector<unsigned char> &data) {
  Be32(o, (uint32_t)data.size());
  const size_t start = o.size();
  o.insert(o.end(), type, type + 4);
  o.insert(o.end(), data.begin(), data.end());
  Be32(o, Crc32(o.data() + start, o.size() - start));
}

}  // namespace

extern "C" {

lgpu_status lgpu_film_post(lgpu_film f, const lgpu_post_desc *d, float *rgba, uint8_t *rgba8) {
  if (!f || !d) return Fail(LGPU_ERR_INVALID_ARGUMENT, "film_post: null argument");
  if (d->tonemap > LGPU_TONEMAP_ACES_FITTED) return Fail(LGPU_ERR_INVALID_ARGUMENT, "film_post: unknown tone map");
  const bool conv = d->kernel != nullptr;
  if (conv && (d->kernel_w == 0 || d->kernel_h == 0 || d->kernel_w > 64 || d->kernel_h > 64))
    return Fail(LGPU_ERR_INVALID_ARGUMENT, "film_post: kernel size must be 1..64");
  if (!rgba && !rgba8) return LGPU_OK;
  std::lock_guard<std::mutex> lk(f->dev->mu);
  LGPU_HIP(hipSetDevice(f->dev->device));
  const uint32_t n = f->width * f->height;
  const hipStream_t st = f->dev->stream;
  DevBuf img, kern, o32, o8;
  LGPU_HIP(img.alloc(16ull * n));
  if (conv) LGPU_HIP(kern.upload(d->kernel, 4ull * d->kernel_w * d->kernel_h));
  if (rgba) LGPU_HIP(o32.alloc(16ull * n));
  if (rgba8) LGPU_HIP(o8.alloc(4ull * n));
  hipLaunchKernelGGL(k_film_mean, dim3((n + 255) / 256), dim3(256), 0, st, f->sum.as<double>(), f->count.as<uint32_t>(),
                     img.as<float4>(), n);
  PostParams p{f->width, f->height, conv ? d->kernel_w : 0u, conv ? d->kernel_h : 0u, d->kernel_intensity,
               d->tonemap, d->clamp, d->gamma, d->estimator_mode == LGPU_ESTIMATOR_FAITHFUL ? 1u : 0u, conv ? 1u : 0u};
  hipLaunchKernelGGL(k_post, dim3((n + 255) / 256), dim3(256), 0, st, img.as<float4>(), kern.as<float>(), p,
                     rgba ? o32.as<float4>() : nullptr, rgba8 ? o8.as<uchar4>() : nullptr);
  LGPU_HIP(hipGetLastError());
  if (rgba) LGPU_HIP(hipMemcpyAsync(rgba, o32.ptr, 16ull * n, hipMemcpyDeviceToHost, st));
  if (rgba8) LGPU_HIP(hipMemcpyAsync(rgba8, o8.ptr, 4ull * n, hipMemcpyDeviceToHost, st));
  LGPU_HIP(hipStreamSynchronize(st));
  return LGPU_OK;
}

lgpu_status lgpu_write_png(const char *path, const uint8_t *rgba8, uint32_t w, uint32_t h) {
  if (!path || !rgba8 || !w || !h) return Fail(LGPU_ERR_INVALID_ARGUMENT, "write_png: bad argument");
  // scanlines: filter byte 0, then RGBA8
  std::vector<unsigned char> raw;
  raw.reserve((size_t)h * (4ull * w + 1));
  for (uint32_t y = 0; y < h; ++y) {
    raw.push_back(0);
    raw.insert(raw.end(), rgba8 + 4ull * w * y, rgba8 + 4ull * w * (y + 1));
  }
  // zlib stream of stored deflate blocks (RFC 1950/1951)
  std::vector<unsigned char> z = {0x78, 0x01};
  uint32_t a = 1, b = 0;
  for (unsigned char c : raw) { a = (a + c) % 65521u; b = (b + a) % 65521u; }
  for (size_t off = 0; off < raw.size() || off == 0;) {
    const size_t len = std::min<size_t>(65535, raw.size() - off);
    const bool last = off + len >= raw.size();
    z.push_back(last ? 1 : 0);
    z.push_back((unsigned char)(len & 0xFF)); z.push_back((unsigned char)(len >> 8));
    z.push_back((unsigned char)(~len & 0xFF)); z.push_back((unsigned char)((~len >> 8) & 0xFF));
    z.insert(z.end(), raw.begin() + off, raw.begin() + off + len);
    off += len;
    if (last) break;
  }
  Be32(z, (b << 16) | a);
  std::vector<unsigned char> png = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
  std::vector<unsigned char> ihdr;
  Be32(ihdr, w);
  Be32(ihdr, h);
  ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8-bit RGBA, deflate, no filter set, no interlace
  Chunk(png, "IHDR", ihdr);
  Chunk(png, "IDAT", z);
  Chunk(png, "IEND", {});
  FILE *fp = std::fopen(path, "wb");
  if (!fp) return Fail(LGPU_ERR_INVALID_ARGUMENT, std::string("write_png: cannot open ") + path);
  const size_t wr = std::fwrite(png.data(), 1, png.size(), fp);
  std::fclose(fp);
  if (wr != png.size()) return Fail(LGPU_ERR_INVALID_ARGUMENT, std::string("write_png: short write to ") + path);
  return LGPU_OK;
}

}  // extern "C"
