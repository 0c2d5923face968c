// Host/device maths for the path tracer. Arithmetic order follows the
// reference renderer so the GPU reproduces its floating-point decisions:
// Vec3 / scalar multiplies by the reciprocal (vec3.h:146-148), Spectrum /
// scalar multiplies by the reciprocal (Spectrum.h:128-133).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define LGPU_HD __host__ __device__ inline

namespace lgpu {

constexpr float kPi2 = 2.f * 3.14159265358979323846264338f;
constexpr float kInvPi = 1.f / 3.14159265358979323846264338f;
constexpr float kInvPi2 = 1.f / (2.f * 3.14159265358979323846264338f);
constexpr float kInvPi4 = 1.f / (4.f * 3.14159265358979323846264338f);

struct V3 {
  float x, y, z;
};
LGPU_HD V3 v3(float a, float b, float c) { V3 r; r.x = a; r.y = b; r.z = c; return r; }
LGPU_HD V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
LGPU_HD V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
LGPU_HD V3 operator*(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
LGPU_HD V3 operator/(V3 a, V3 b) { return v3(a.x / b.x, a.y / b.y, a.z / b.z); }
LGPU_HD V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
LGPU_HD V3 operator/(V3 a, float s) { const float inv = 1.f / s; return v3(a.x * inv, a.y * inv, a.z * inv); }
LGPU_HD V3 operator-(V3 a) { return a * -1.f; }
LGPU_HD float Dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
LGPU_HD V3 Cross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
LGPU_HD float Mag(V3 v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }
LGPU_HD V3 Norm(V3 v) { return v / Mag(v); }
LGPU_HD float DistSq(V3 a, V3 b) { const V3 d = b - a; return d.x * d.x + d.y * d.y + d.z * d.z; }
LGPU_HD float Clampf(float v, float lo, float hi) { return fmaxf(fminf(v, hi), lo); }
LGPU_HD float Fract(float r) { return r - floorf(r); }
LGPU_HD V3 ToV3(float4 f) { return v3(f.x, f.y, f.z); }
LGPU_HD V3 ToV3(const float *p) { return v3(p[0], p[1], p[2]); }
LGPU_HD float4 ToF4(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
LGPU_HD V3 WorldToLocal(V3 v, V3 t, V3 n, V3 bt) { return v3(Dot(t, v), Dot(n, v), Dot(bt, v)); }
LGPU_HD V3 LocalToWorld(V3 v, V3 t, V3 n, V3 bt) { return t * v.x + n * v.y + bt * v.z; }
LGPU_HD V3 SphericalDirection(float s, float c, float phi) { return v3(s * cosf(phi), c, s * sinf(phi)); }
LGPU_HD float SphericalTheta(V3 v) { return acosf(fmaxf(fminf(v.y, 1.f), -1.f)); }
LGPU_HD float SphericalPhi(V3 v) {
  const float p = atan2f(v.z, v.x);
  return (p < 0) ? (float)((double)p + 3.14159265359 * 2.) : p;
}

// ---------------------------------------------------------------- spectrum: 4 lanes
// RGB mode uses lanes 0..2 (lane 3 is ignored); HERO4 uses all four.
struct Spec {
  float c[4];
};
LGPU_HD Spec S(float v) { Spec r; r.c[0] = r.c[1] = r.c[2] = r.c[3] = v; return r; }
LGPU_HD Spec S4(float4 f) { Spec r; r.c[0] = f.x; r.c[1] = f.y; r.c[2] = f.z; r.c[3] = f.w; return r; }
LGPU_HD float4 F4(Spec s) { return make_float4(s.c[0], s.c[1], s.c[2], s.c[3]); }
LGPU_HD Spec operator+(Spec a, Spec b) { Spec r; for (int i = 0; i < 4; ++i) r.c[i] = a.c[i] + b.c[i]; return r; }
LGPU_HD Spec operator-(Spec a, Spec b) { Spec r; for (int i = 0; i < 4; ++i) r.c[i] = a.c[i] - b.c[i]; return r; }
LGPU_HD Spec operator*(Spec a, Spec b) { Spec r; for (int i = 0; i < 4; ++i) r.c[i] = a.c[i] * b.c[i]; return r; }
LGPU_HD Spec operator/(Spec a, Spec b) { Spec r; for (int i = 0; i < 4; ++i) r.c[i] = a.c[i] / b.c[i]; return r; }
LGPU_HD Spec operator*(Spec a, float s) { Spec r; for (int i = 0; i < 4; ++i) r.c[i] = a.c[i] * s; return r; }
LGPU_HD Spec operator/(Spec a, float s) { const float inv = 1.f / s; Spec r; for (int i = 0; i < 4; ++i) r.c[i] = a.c[i] * inv; return r; }
LGPU_HD Spec operator+(Spec a, float s) { Spec r; for (int i = 0; i < 4; ++i) r.c[i] = a.c[i] + s; return r; }
LGPU_HD Spec operator-(Spec a, float s) { Spec r; for (int i = 0; i < 4; ++i) r.c[i] = a.c[i] - s; return r; }
LGPU_HD Spec SqrtS(Spec a) { Spec r; for (int i = 0; i < 4; ++i) r.c[i] = sqrtf(a.c[i]); return r; }
LGPU_HD bool IsBlack(Spec s, int mode) {
  const int n = mode ? 4 : 3;
  for (int i = 0; i < n; ++i)
    if (s.c[i] != 0.f) return false;
  return true;
}
// RGBSpectrum::y (Spectrum.h:328-331); hero mode: mean of the four lanes
LGPU_HD float Luminance(Spec s, int mode) {
  if (mode) return (s.c[0] + s.c[1] + s.c[2] + s.c[3]) * .25f;
  return 0.212671f * s.c[0] + 0.715160f * s.c[1] + 0.072169f * s.c[2];
}

// CIE 1931 analytic fit (Wyman, Sloan & Shirley 2013)
// kNative (device code only): t from the widths' reciprocals (folded at compile time) and the
// hardware exp2 of -t^2/2 log2(e); relative error about t^2 2^-23 + 1 ulp against the division
// and expf, for the film update where the fit dominates the cost.
template <bool kNative = false>
LGPU_HD float CmfLobe(float x, float mu, float s1, float s2) {
  if constexpr (kNative) {
    const float t = (x - mu) * (x < mu ? 1.f / s1 : 1.f / s2);
    return __expf(-.5f * t * t);
  } else {
    const float t = (x - mu) / (x < mu ? s1 : s2);
    return expf(-.5f * t * t);
  }
}
template <bool kNative = false>
LGPU_HD void CMF(float l, float xyz[3]) {
  xyz[0] = 1.056f * CmfLobe<kNative>(l, 599.8f, 37.9f, 31.0f) + 0.362f * CmfLobe<kNative>(l, 442.0f, 16.0f, 26.7f) -
           0.065f * CmfLobe<kNative>(l, 501.1f, 20.4f, 26.2f);
  xyz[1] = 0.821f * CmfLobe<kNative>(l, 568.8f, 46.9f, 40.5f) + 0.286f * CmfLobe<kNative>(l, 530.9f, 16.3f, 31.1f);
  xyz[2] = 1.217f * CmfLobe<kNative>(l, 437.0f, 11.8f, 36.0f) + 0.681f * CmfLobe<kNative>(l, 459.0f, 26.0f, 13.8f);
}
constexpr double kCieYIntegral = 106.856895;  // SPD_Data.h:22

// RGB -> spectral value at lambda: smoothstep partition of unity
// (blue/green edge 470-510 nm, green/red edge 570-610 nm).
LGPU_HD float Smooth01(float a, float b, float x) {
  float t = (x - a) / (b - a);
  t = Clampf(t, 0.f, 1.f);
  return t * t * (3.f - 2.f * t);
}
LGPU_HD float Uplift(float r, float g, float b, float l) {
  const float t1 = Smooth01(470.f, 510.f, l);
  const float t2 = Smooth01(570.f, 610.f, l);
  return r * t2 + g * (t1 - t2) + b * (1.f - t1);
}

// Planck, normalised at the Wien peak (GraphInputs.cpp:87-104)
LGPU_HD float Planck(float lambdaNm, float T) {
  const float c = 299792458.f, h = 6.62606957e-34f, kb = 1.3806488e-23f;
  const float l = lambdaNm * 1e-9f;
  const float l5 = l * l * l * l * l;
  return (2.f * h * c * c) / (l5 * (expf((h * c) / (l * kb * T)) - 1.f));
}
LGPU_HD float BlackbodyNormalised(float lambdaNm, float T) {
  const float lambdaMax = 2.8977721e-3f / T * 1e9f;
  return Planck(lambdaNm, T) / Planck(lambdaMax, T);
}

// ---------------------------------------------------------------- Philox4x32-10
LGPU_HD void Philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
  }
}
LGPU_HD float U32ToFloat(uint32_t u) { return (float)(u >> 8) * (1.f / 16777216.f); }

// Per-path sampler: dimension d is word (d & 3) of Philox(pixel, sample, d >> 2, 0; seed, 0).
struct Rng {
  uint32_t pixel, sample, dim, seed;
  uint32_t blk, c0, c1, c2, c3;  // the current Philox block (no indexed array: registers, not scratch)
  LGPU_HD void init(uint32_t p, uint32_t s, uint32_t d, uint32_t sd) {
    pixel = p; sample = s; dim = d; seed = sd; blk = 0xFFFFFFFFu;
  }
  LGPU_HD float Get1D() {
    const uint32_t b = dim >> 2;
    if (b != blk) {
      uint32_t c[4] = {pixel, sample, b, 0u};
      Philox(c, seed, 0u);
      c0 = c[0]; c1 = c[1]; c2 = c[2]; c3 = c[3];
      blk = b;
    }
    const uint32_t k = dim++ & 3;
    return U32ToFloat(k == 0 ? c0 : (k == 1 ? c1 : (k == 2 ? c2 : c3)));
  }
};
LGPU_HD float WavelengthU(uint32_t pixel, uint32_t sample, uint32_t seed) {
  uint32_t c[4] = {pixel, sample, 0x3FFFFFFFu, 0u};
  Philox(c, seed, 1u);
  return U32ToFloat(c[0]);
}
LGPU_HD void HeroWavelengths(float u, float lam[4]) {
  const float l0 = 360.f + 470.f * u;
  for (int i = 0; i < 4; ++i) {
    float l = l0 + 117.5f * (float)i;
    if (l >= 830.f) l -= 470.f;
    lam[i] = l;
  }
}

// ---------------------------------------------------------------- Piecewise1D in a flat pool
// layout at off: [n (float bits), integral, pdf[n], cdf[n+1], guide[n+1] (u32 bits)]
// (Piecewise.cpp:9-44; the guide table is this library's search accelerator)
LGPU_HD uint32_t AsU32(float f) { return __builtin_bit_cast(uint32_t, f); }
LGPU_HD float AsF32(uint32_t u) { return __builtin_bit_cast(float, u); }
LGPU_HD uint32_t PwN(const float *pool, uint32_t off) { return AsU32(pool[off]); }
template <class P>
LGPU_HD uint32_t FindInterval(int size, const P &pred) {  // maths.h:77-90
  int first = 0, len = size;
  while (len > 0) {
    const int half = len >> 1, middle = first + half;
    if (pred(middle)) { first = middle + 1; len -= half + 1; }
    else len = half;
  }
  int r = first - 1;
  if (r > size - 2) r = size - 2;
  if (r < 0) r = 0;
  return (uint32_t)r;
}
// FindInterval(n + 1, cdf[i] <= u) through the guide table: with c(u) = #{i : cdf[i] <= u} for a
// non-decreasing cdf, c(u) lies in [guide[k], guide[k + 1]] for the bin k = floor(u n) holding u.
// The bracket is checked against cdf itself (float binning may miss by one bin, u may be NaN) and
// widened to [0, n + 1] when the check fails, so the result equals the full search's.
LGPU_HD uint32_t PwFind(const float *cdf, uint32_t n, float u) {
  if (n == 0) return 0;
  const uint32_t *guide = reinterpret_cast<const uint32_t *>(cdf + n + 1);
  const float x = u * (float)n;
  const uint32_t k = x >= 0.f ? (x < (float)(n - 1) ? (uint32_t)x : n - 1) : 0u;
  uint32_t a = guide[k], b = guide[k + 1];
  if (!(a <= b && b <= n + 1 && (a == 0 || cdf[a - 1] <= u) && (b == n + 1 || cdf[b] > u))) { a = 0; b = n + 1; }
  int first = (int)a, len = (int)(b - a);
  while (len > 0) {
    const int half = len >> 1, middle = first + half;
    if (cdf[middle] <= u) { first = middle + 1; len -= half + 1; }
    else len = half;
  }
  int r = first - 1;
  if (r > (int)n - 1) r = (int)n - 1;
  if (r < 0) r = 0;
  return (uint32_t)r;
}
LGPU_HD uint32_t PwSampleDiscrete(const float *pool, uint32_t off, float u, float *pdf) {
  const uint32_t n = PwN(pool, off);
  const float integral = pool[off + 1];
  const float *pdfv = pool + off + 2;
  const float *cdf = pdfv + n;
  const uint32_t o = PwFind(cdf, n, u);
  if (pdf) *pdf = pdfv[o] / (integral * (float)n);
  return o;
}
LGPU_HD float PwSampleContinuous(const float *pool, uint32_t off, float u, float *pdf, uint32_t *offOut) {
  const uint32_t n = PwN(pool, off);
  const float integral = pool[off + 1];
  const float *pdfv = pool + off + 2;
  const float *cdf = pdfv + n;
  const uint32_t o = PwFind(cdf, n, u);
  if (offOut) *offOut = o;
  float du = u - cdf[o];
  if ((cdf[o + 1] - cdf[o]) > 0) du /= (cdf[o + 1] - cdf[o]);
  if (pdf) *pdf = pdfv[o] / integral;
  return ((float)o + du) / (float)n;
}
LGPU_HD float PwPdfRaw(const float *pool, uint32_t off, uint32_t i) { return pool[off + 2 + i]; }
LGPU_HD float PwPdfNormalised(const float *pool, uint32_t off, uint32_t i) {
  const uint32_t n = PwN(pool, off);
  return pool[off + 2 + i] / (pool[off + 1] * (float)n);
}
LGPU_HD float PwIntegral(const float *pool, uint32_t off) { return pool[off + 1]; }

}  // namespace lgpu
