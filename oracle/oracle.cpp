/*
 * oracle.cpp — CPU restatement of the Zoophish/Lambda render path.
 *
 * TEST INFRASTRUCTURE ONLY: this file is the checker that the HIP product is
 * compared against (tests/, __graft_entry__.smoke(), bench.py cpu_baseline).
 * The product library never includes or links it.
 *
 * Every function cites the reference file:line it restates. Floating-point
 * operation order follows the reference, including its SSE conventions:
 * Vec3 / scalar multiplies by the reciprocal (vec3.h:146-148), Spectrum /
 * scalar multiplies by the reciprocal (Spectrum.h:128-133), Spectrum / Spectrum
 * and the RR division divide lane-wise (Spectrum.h:62-66, 104-108).
 *
 * Two switches shared with the product:
 *   spectral mode  RGB   — the reference's RGBSpectrum (Spectrum.h:24)
 *                  HERO4 — 4 hero wavelengths (project-defined, see below)
 *   estimator mode FAITHFUL  — the reference's deterministic defects kept
 *                  CORRECTED — the defects listed at each `corrected` branch fixed
 *
 * Deviations that replace undefined or platform-defined behaviour:
 *   - Halton + blue-noise shifter (HaltonSampler.cpp, SampleShifter.cpp) are
 *     replaced by Philox4x32-10 keyed on (seed, stream) with counter
 *     (pixel, sample, dim/4, 0); one output word per dimension.
 *   - ManyLightSampler keeps triangle lights in an unordered_map
 *     (ManyLightSampler.cpp:273-281) and splits with std::partition
 *     (:215); both orders are implementation-defined. Here triangle lights are
 *     ordered by (mesh, triangle) and the split uses a stable partition.
 *   - float -> unsigned bucket conversion of NaN offsets (flat axes,
 *     ManyLightSampler.cpp:144,210) is UB; it is defined as bucket 0.
 *   - CIE 1931 colour matching uses the analytic multi-lobe fit of Wyman,
 *     Sloan and Shirley (JCGT 2013) instead of the 471-entry table
 *     (SPD_Data.cpp), sampled at the same 471 wavelengths.
 *
 * parity unpinned: the Embree hit convention (Ng orientation, barycentrics) is
 * not pinned by any reference test; this oracle defines Ng = (v1-v0)x(v2-v0)
 * and (u, v) weighting v1 and v2 (maths.h:57-60).
 */
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <thread>
#include <vector>

namespace orc {

static const float PI = 3.14159265358979323846264338f;
static const float PI2 = 2.f * PI;
static const float INV_PI = 1.f / PI;
static const float INV_PI2 = 1.f / (2.f * PI);
static const float INV_PI4 = 1.f / (4.f * PI);
static const float SURFACE_EPSILON = 1e-5f;  // ScatterEvent.h:12

// ---------------------------------------------------------------- maths (vec3.h, maths.h)
struct V2 { float x = 0, y = 0; V2() {} V2(float a, float b) : x(a), y(b) {} };
struct V3 {
  float x = 0, y = 0, z = 0;
  V3() {}
  V3(float a, float b, float c) : x(a), y(b), z(c) {}
  explicit V3(const float *p) : x(p[0]), y(p[1]), z(p[2]) {}
  V3 operator+(const V3 &o) const { return V3(x + o.x, y + o.y, z + o.z); }
  V3 operator-(const V3 &o) const { return V3(x - o.x, y - o.y, z - o.z); }
  V3 operator*(const V3 &o) const { return V3(x * o.x, y * o.y, z * o.z); }
  V3 operator/(const V3 &o) const { return V3(x / o.x, y / o.y, z / o.z); }
  V3 operator*(float s) const { return V3(x * s, y * s, z * s); }
  V3 operator/(float s) const { const float inv = 1.f / s; return V3(x * inv, y * inv, z * inv); }  // vec3.h:146
  V3 operator-() const { return *this * -1.f; }
  V3 &operator+=(const V3 &o) { x += o.x; y += o.y; z += o.z; return *this; }
  float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
static inline float Dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 Cross(const V3 &a, const V3 &b) {
  return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float Mag(const V3 &v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
static inline V3 Norm(const V3 &v) { return v / Mag(v); }
static inline float DistSq(const V3 &a, const V3 &b) { const V3 d = b - a; return d.x * d.x + d.y * d.y + d.z * d.z; }
static inline float Clampf(float v, float lo, float hi) { return std::max(std::min(v, hi), lo); }
static inline float Fract(float r) { return r - std::floor(r); }
// vec3.h:200-212, y is the normal axis
static inline V3 WorldToLocal(const V3 &v, const V3 &t, const V3 &n, const V3 &bt) { return V3(Dot(t, v), Dot(n, v), Dot(bt, v)); }
static inline V3 LocalToWorld(const V3 &v, const V3 &t, const V3 &n, const V3 &bt) { return t * v.x + n * v.y + bt * v.z; }
static inline V3 SphericalDirection(float sinT, float cosT, float phi) { return V3(sinT * std::cos(phi), cosT, sinT * std::sin(phi)); }
static inline float SphericalTheta(const V3 &v) { return std::acos(std::max(std::min(v.y, 1.f), -1.f)); }
static inline float SphericalPhi(const V3 &v) {
  const float p = std::atan2(v.z, v.x);
  return (p < 0) ? (float)(p + 3.14159265359 * 2.) : p;  // vec3.h:231-234 (double constant)
}
// vec3.h:237-241 (Rodrigues with sqrt(1-cos^2) and an unnormalised axis, as written)
static inline V3 Rotate(const V3 &v, const V3 &axis, float theta) {
  const float c = std::cos(theta);
  const float s = std::sqrt(1.f - c * c);
  return v * c + Cross(axis, v) * s + axis * Dot(axis, v) * (1.f - c);
}

struct Bounds {
  V3 mn, mx;  // default (0,0,0)-(0,0,0), bounds.h:61-64
  Bounds() {}
  explicit Bounds(const V3 &p) : mn(p), mx(p) {}
  Bounds(const V3 &a, const V3 &b) : mn(a), mx(b) {}
  float SurfaceArea() const { const V3 s = mx - mn; return (s.x * s.y + s.y * s.z + s.x * s.z) * 2.f; }
  float MaxLength() const { const V3 b = mx - mn; return std::max(std::max(b.x, b.y), b.z); }
  V3 Center() const { return mn + (mx - mn) * .5f; }
  float Diagonal() const { return Mag(mx - mn); }
  V3 Offset(const V3 &p) const { return (p - mn) / (mx - mn); }
};
static inline Bounds Union(const Bounds &a, const Bounds &b) {  // bounds.h:108-113
  return Bounds(V3(std::min(a.mn.x, b.mn.x), std::min(a.mn.y, b.mn.y), std::min(a.mn.z, b.mn.z)),
                V3(std::max(a.mx.x, b.mx.x), std::max(a.mx.y, b.mx.y), std::max(a.mx.z, b.mx.z)));
}
static inline Bounds Union(const Bounds &b, const V3 &p) {  // bounds.h:116-121
  return Bounds(V3(std::min(p.x, b.mn.x), std::min(p.y, b.mn.y), std::min(p.z, b.mn.z)),
                V3(std::max(p.x, b.mx.x), std::max(p.y, b.mx.y), std::max(p.z, b.mx.z)));
}

// maths.h:77-90
template <class P>
static inline unsigned FindInterval(int size, const P &pred) {
  int first = 0, len = size;
  while (len > 0) {
    int half = len >> 1, middle = first + half;
    if (pred(middle)) { first = middle + 1; len -= half + 1; }
    else len = half;
  }
  return (unsigned)std::max(std::min(first - 1, size - 2), 0);
}

// ---------------------------------------------------------------- spectrum (Spectrum.h:27-182, 306-374)
enum { RGB = 0, HERO4 = 1 };
struct Spec {
  float c[4];
  Spec(float v = 0.f) { c[0] = c[1] = c[2] = c[3] = v; }
  Spec operator+(const Spec &o) const { Spec r; for (int i = 0; i < 4; ++i) r.c[i] = c[i] + o.c[i]; return r; }
  Spec operator-(const Spec &o) const { Spec r; for (int i = 0; i < 4; ++i) r.c[i] = c[i] - o.c[i]; return r; }
  Spec operator*(const Spec &o) const { Spec r; for (int i = 0; i < 4; ++i) r.c[i] = c[i] * o.c[i]; return r; }
  Spec operator/(const Spec &o) const { Spec r; for (int i = 0; i < 4; ++i) r.c[i] = c[i] / o.c[i]; return r; }
  Spec operator*(float s) const { Spec r; for (int i = 0; i < 4; ++i) r.c[i] = c[i] * s; return r; }
  Spec operator/(float s) const { const float inv = 1.f / s; Spec r; for (int i = 0; i < 4; ++i) r.c[i] = c[i] * inv; return r; }
  Spec operator+(float s) const { Spec r; for (int i = 0; i < 4; ++i) r.c[i] = c[i] + s; return r; }
  Spec operator-(float s) const { Spec r; for (int i = 0; i < 4; ++i) r.c[i] = c[i] - s; return r; }
  Spec &operator+=(const Spec &o) { for (int i = 0; i < 4; ++i) c[i] += o.c[i]; return *this; }
  Spec &operator*=(const Spec &o) { for (int i = 0; i < 4; ++i) c[i] *= o.c[i]; return *this; }
};
static inline Spec SqrtS(const Spec &s) { Spec r; for (int i = 0; i < 4; ++i) r.c[i] = std::sqrt(s.c[i]); return r; }

struct Lanes { int mode; float lambda[4]; };
static inline int NL(const Lanes &l) { return l.mode == HERO4 ? 4 : 3; }
static inline bool IsBlack(const Spec &s, const Lanes &l) {
  for (int i = 0; i < NL(l); ++i) if (s.c[i] != 0) return false;
  return true;
}
// RGBSpectrum::y (Spectrum.h:328-331); HERO4: mean over the hero wavelengths
static inline float Luminance(const Spec &s, const Lanes &l) {
  if (l.mode == HERO4) return (s.c[0] + s.c[1] + s.c[2] + s.c[3]) * .25f;
  return 0.212671f * s.c[0] + 0.715160f * s.c[1] + 0.072169f * s.c[2];
}

// Analytic CIE 1931 fit (Wyman, Sloan, Shirley 2013, multi-lobe piecewise Gaussians)
static inline float Lobe(float x, float mu, float s1, float s2) {
  const float t = (x - mu) / (x < mu ? s1 : s2);
  return std::exp(-.5f * t * t);
}
static inline void CMF(float l, float xyz[3]) {
  xyz[0] = 1.056f * Lobe(l, 599.8f, 37.9f, 31.0f) + 0.362f * Lobe(l, 442.0f, 16.0f, 26.7f) - 0.065f * Lobe(l, 501.1f, 20.4f, 26.2f);
  xyz[1] = 0.821f * Lobe(l, 568.8f, 46.9f, 40.5f) + 0.286f * Lobe(l, 530.9f, 16.3f, 31.1f);
  xyz[2] = 1.217f * Lobe(l, 437.0f, 11.8f, 36.0f) + 0.681f * Lobe(l, 459.0f, 26.0f, 13.8f);
}
static const double CIE_Y_INTEGRAL = 106.856895;  // SPD_Data.h:22

// Project RGB -> spectrum uplift: smoothstep partition of unity (blue/green
// edge 470-510 nm, green/red edge 570-610 nm); white (1,1,1) -> 1 everywhere.
static inline float Smooth01(float a, float b, float x) {
  float t = (x - a) / (b - a);
  t = Clampf(t, 0.f, 1.f);
  return t * t * (3.f - 2.f * t);
}
static inline float Uplift(const float rgb[3], float l) {
  const float t1 = Smooth01(470.f, 510.f, l);
  const float t2 = Smooth01(570.f, 610.f, l);
  return rgb[0] * t2 + rgb[1] * (t1 - t2) + rgb[2] * (1.f - t1);
}

// GraphInputs.cpp:87-104
static float Planck(float lambdaNm, float T) {
  const float c = 299792458.f, h = 6.62606957e-34f, kb = 1.3806488e-23f;
  const float l = lambdaNm * 1e-9f;
  const float l5 = l * l * l * l * l;
  return (2.f * h * c * c) / (l5 * (std::exp((h * c) / (l * kb * T)) - 1.f));
}
static float BlackbodyNormalised(float lambdaNm, float T) {
  const float lambdaMax = 2.8977721e-3f / T * 1e9f;
  return Planck(lambdaNm, T) / Planck(lambdaMax, T);
}
// BlackbodyInput::MakeBlackbodySpectrum with Spectrum::nSamples = 3 (GraphInputs.cpp:106-116)
// then RGBSpectrum::FromSampled (Spectrum.h:353-373) over 471 CIE wavelengths.
static void SampledToRGB(const float *lam, const float *v, unsigned n, float out[3]) {
  float xyz[3] = {0, 0, 0};
  for (unsigned i = 0; i < 471; ++i) {
    const float li = 360.f + (float)i;
    float val;
    if (li <= lam[0]) val = v[0];
    else if (li >= lam[n - 1]) val = v[n - 1];
    else {
      const int off = (int)FindInterval((int)n, [&](int k) { return lam[k] <= li; });
      const float t = (li - lam[off]) / (lam[off + 1] - lam[off]);
      val = v[off] + (v[off + 1] - v[off]) * t;
    }
    float cmf[3];
    CMF(li, cmf);
    xyz[0] += val * cmf[0];
    xyz[1] += val * cmf[1];
    xyz[2] += val * cmf[2];
  }
  const float scale = (float)(830 - 360) / (float)(CIE_Y_INTEGRAL * 471);
  for (int k = 0; k < 3; ++k) xyz[k] *= scale;
  // XYZToRGB, Spectrum.h:194-198 (double coefficients)
  out[0] = (float)(3.240479 * xyz[0] - 1.537150 * xyz[1] - 0.498535 * xyz[2]);
  out[1] = (float)(-0.969256 * xyz[0] + 1.875991 * xyz[1] + 0.041556 * xyz[2]);
  out[2] = (float)(0.055648 * xyz[0] - 0.204043 * xyz[1] + 1.057311 * xyz[2]);
}
static void BlackbodyRGB(float T, float out[3]) {
  const unsigned n = 3;
  const float invSamples = 1.f / (float)n;
  float lam[3], v[3];
  for (unsigned i = 0; i < n; ++i) lam[i] = 400.f + 300.f * invSamples * (float)i;
  for (unsigned i = 0; i < n; ++i) v[i] = BlackbodyNormalised(lam[i], T);
  SampledToRGB(lam, v, n, out);
}

// ---------------------------------------------------------------- Philox4x32-10 sampler
static inline void Philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
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
static inline float U32ToFloat(uint32_t u) { return (float)(u >> 8) * (1.f / 16777216.f); }

// Sampler interface (Sampler.h:9-25): sequential per-path dimension counter.
struct Sampler {
  uint32_t pixel = 0, sample = 0, dim = 0, seed = 0;
  uint32_t cachedBlock = 0xFFFFFFFFu, cache[4];
  float Get1D() {
    const uint32_t blk = dim >> 2;
    if (blk != cachedBlock) {
      cache[0] = pixel; cache[1] = sample; cache[2] = blk; cache[3] = 0;
      Philox(cache, seed, 0u);
      cachedBlock = blk;
    }
    return U32ToFloat(cache[dim++ & 3]);
  }
  V2 Get2D() { const float a = Get1D(); const float b = Get1D(); return V2(a, b); }
};
// Wavelength stream: key word 1 = 1, dimension block 0x3FFFFFFF.
static inline float WavelengthU(uint32_t pixel, uint32_t sample, uint32_t seed) {
  uint32_t c[4] = {pixel, sample, 0x3FFFFFFFu, 0};
  Philox(c, seed, 1u);
  return U32ToFloat(c[0]);
}
static inline void HeroWavelengths(float u, float lam[4]) {
  const float l0 = 360.f + 470.f * u;
  for (int i = 0; i < 4; ++i) {
    float l = l0 + 117.5f * (float)i;
    if (l >= 830.f) l -= 470.f;
    lam[i] = l;
  }
}

// ---------------------------------------------------------------- Piecewise (Piecewise.cpp)
struct Piecewise1D {
  std::vector<float> pdf, cdf;
  float integral = 0;
  Piecewise1D() {}
  Piecewise1D(const float *d, unsigned n) : pdf(d, d + n), cdf(n + 1) {  // Piecewise.cpp:9-24
    cdf[0] = 0;
    for (unsigned i = 1; i < n + 1; ++i) cdf[i] = cdf[i - 1] + pdf[i - 1] / (float)n;
    integral = cdf[n];
    if (integral == 0) { for (unsigned i = 1; i < n + 1; ++i) cdf[i] = (float)i / (float)n; }
    else { for (unsigned i = 1; i < n + 1; ++i) cdf[i] /= integral; }
  }
  unsigned SampleDiscrete(float u, float *p) const {  // :38-44
    const unsigned off = FindInterval((int)cdf.size(), [&](int i) { return cdf[i] <= u; });
    if (p) *p = pdf[off] / (integral * (float)pdf.size());
    return off;
  }
  float SampleContinuous(float u, float *p, int *offOut) const {  // :26-35
    const unsigned off = FindInterval((int)cdf.size(), [&](int i) { return cdf[i] <= u; });
    if (offOut) *offOut = (int)off;
    float du = u - cdf[off];
    if ((cdf[off + 1] - cdf[off]) > 0) du /= (cdf[off + 1] - cdf[off]);
    if (p) *p = pdf[off] / integral;
    return ((float)off + du) / (float)pdf.size();
  }
  float PDF(unsigned i) const { return pdf[i]; }  // Piecewise.h:26-28 (raw weight, A28)
  float PDFNormalised(unsigned i) const { return pdf[i] / (integral * (float)pdf.size()); }
};
struct Piecewise2D {  // Piecewise.cpp:50-76
  std::vector<Piecewise1D> cond;
  Piecewise1D marg;
  Piecewise2D() {}
  Piecewise2D(const float *f, unsigned nu, unsigned nv) {
    for (unsigned v = 0; v < nv; ++v) cond.emplace_back(&f[v * nu], nu);
    std::vector<float> m;
    for (unsigned v = 0; v < nv; ++v) m.push_back(cond[v].integral);
    marg = Piecewise1D(m.data(), nv);
  }
  V2 SampleContinuous(const V2 &u, float *pdf) const {
    float p0, p1;
    int v;
    const float d1 = marg.SampleContinuous(u.y, &p1, &v);
    const float d0 = cond[v].SampleContinuous(u.x, &p0, nullptr);
    *pdf = p0 * p1;
    return V2(d0, d1);
  }
  float PDF(const V2 &uv) const {
    const int nu = (int)cond[0].pdf.size(), nv = (int)marg.pdf.size();
    const int iu = std::max(std::min((int)(uv.x * (float)nu), nu - 1), 0);
    const int iv = std::max(std::min((int)(uv.y * (float)nv), nv - 1), 0);
    return cond[iv].pdf[iu] / marg.integral;
  }
};

// ---------------------------------------------------------------- scene data
struct Texture {
  unsigned w = 0, h = 0, ch = 0;
  std::vector<float> data;
  // texture_t::GetPixelUVBilinear (Texture.h:71-81), RGB or scalar
  void Bilinear(float u, float v, float out[3]) const {
    const float fx = std::min((float)(w - 1), Clampf(u, 0.f, 1.f) * (float)w);
    const float fy = std::min((float)(h - 1), Clampf(v, 0.f, 1.f) * (float)h);
    const unsigned x = (unsigned)fx, y = (unsigned)fy;
    const unsigned x1 = std::min(x + 1, w - 1), y1 = std::min(y + 1, h - 1);
    const float tx = fx - (float)x, ty = fy - (float)y;
    for (unsigned c = 0; c < 3; ++c) {
      const unsigned cc = ch == 1 ? 0 : c;
      const float p00 = data[(y * w + x) * ch + cc], p10 = data[(y * w + x1) * ch + cc];
      const float p01 = data[(y1 * w + x) * ch + cc], p11 = data[(y1 * w + x1) * ch + cc];
      const float ya = p00 + (p10 - p00) * tx;
      const float yb = p01 + (p11 - p01) * tx;
      out[c] = ya + (yb - ya) * ty;
    }
  }
  // The same lookup with alpha: channel 3 of an RGBA texture, 1 for 1- and 3-channel textures.
  void Bilinear4(float u, float v, float out[4]) const {
    Bilinear(u, v, out);
    out[3] = 1.f;
    if (ch != 4) return;
    const float fx = std::min((float)(w - 1), Clampf(u, 0.f, 1.f) * (float)w);
    const float fy = std::min((float)(h - 1), Clampf(v, 0.f, 1.f) * (float)h);
    const unsigned x = (unsigned)fx, y = (unsigned)fy;
    const unsigned x1 = std::min(x + 1, w - 1), y1 = std::min(y + 1, h - 1);
    const float tx = fx - (float)x, ty = fy - (float)y;
    const float p00 = data[(y * w + x) * 4 + 3], p10 = data[(y * w + x1) * 4 + 3];
    const float p01 = data[(y1 * w + x) * 4 + 3], p11 = data[(y1 * w + x1) * 4 + 3];
    const float ya = p00 + (p10 - p00) * tx;
    const float yb = p01 + (p11 - p01) * tx;
    out[3] = ya + (yb - ya) * ty;
  }
};

struct Material {
  std::vector<lgpu_node> nodes;
  uint32_t bxdf = LGPU_NONE, emission = LGPU_NONE;
  float intensity = 1.f;
  uint32_t medIn = LGPU_NONE, medOut = LGPU_NONE;
  std::vector<float> bbRGB;  // per node: blackbody RGB for constant temperatures (RGB mode)
  std::vector<char> bbConst;
};

struct Proxy {  // InstanceProxy's mesh, as given
  uint32_t nv = 0, nt = 0, material = 0;
  std::vector<float> pos, nrm, tan, uv;
  std::vector<uint32_t> idx;
};
struct Mesh {
  std::vector<V3> p, n;
  std::vector<float> tan;  // 4 per vertex
  std::vector<V2> uv;
  std::vector<uint32_t> idx;
  uint32_t material = 0;
  uint32_t firstTri = 0;
  int meshLight = -1;  // index into lights
};

enum LightKind { L_MESH = 0, L_TRI = 1, L_POINT = 2, L_SPOT = 3, L_ENV = 4 };
struct Light {
  int kind = L_MESH;
  int mesh = -1;          // L_MESH / L_TRI
  uint32_t tri = 0;       // L_TRI: local triangle
  int parentMeshLight = -1;
  Piecewise1D triDist;    // L_MESH (MeshLight::InitDistribution, MeshLight.cpp:118-125)
  V3 position, axis;
  float cosCone = 0, cosFalloff = 0;
  float intensity = 1;
  float rgb[3] = {1, 1, 1};
  float temperature = 0;
  float bbRGB[3] = {0, 0, 0};
  // env
  Texture env;
  Piecewise2D envDist;
  V2 offset;
  float radius = 0;
  Bounds envBounds;
};

struct Hit {
  V3 point, normalG, normalS, tangent, bitangent;
  V2 uv;
  float tFar = INFINITY;
  uint32_t prim = 0;   // global triangle index
  uint32_t local = 0;  // RayHit::primId (mesh-local)
  int mesh = -1;
};

struct Sampler;
struct Event {  // ScatterEvent.h:18-40
  V3 wo, wi, woL, wiL;
  Hit *hit = nullptr;
  Sampler *sampler = nullptr;  // the integrator's sampler (LdMediumPoint draws from it)
  int sidedness = 1;
  float eta = 1.001f;
  int medium = -1;                 // ScatterEvent::medium
  bool mediumInteraction = false;
  Lanes lanes;
  V3 ToLocal(const V3 &v) const { return WorldToLocal(v, hit->tangent, hit->normalS, hit->bitangent); }
  V3 ToWorld(const V3 &v) const { return LocalToWorld(v, hit->tangent, hit->normalS, hit->bitangent); }
  void SurfaceLocalise() { woL = ToLocal(wo); wiL = ToLocal(wi); sidedness = (woL.y > 0) ? 1 : -1; }
};

// Light-tree node (ManyLightSampler.h:48-76)
struct Cone { V3 axis = V3(0, 1, 0); float thetaO = 0, thetaE = 0; };
struct LNode {
  LNode *child[2] = {nullptr, nullptr};
  LNode *parent = nullptr;
  unsigned first = 0, count = 0;
  Bounds bounds;
  Cone cone;
  float power = 0;
  int leafIndex = -1;
  bool IsLeaf() const { return !child[0] && !child[1]; }
};

struct BVHNode { Bounds b; uint32_t left = 0, right = 0, first = 0, count = 0; };

// HomogeneousMedium (HomogeneousMedium.cpp) or the density-grid medium (project extension:
// piecewise-constant voxels, sigma = density * sigma_base).
struct MediumO {
  int type = LGPU_MEDIUM_HOMOGENEOUS;
  float sa[3] = {0, 0, 0}, ss[3] = {0, 0, 0};
  float g = 0;
  uint32_t nx = 0, ny = 0, nz = 0;
  std::vector<float> dens;
  V3 bmin, bmax;
};

}  // namespace orc

using namespace orc;

struct oracle_scene_s {
  std::vector<Texture> textures;
  std::vector<MediumO> media;
  std::vector<Material> materials;
  std::vector<Mesh> meshes;
  std::vector<Proxy> proxies;
  std::vector<uint32_t> triMesh;  // global tri -> mesh
  std::vector<Light> lights;      // scene lights in add order (Scene::lights)
  int envLight = -1;
  uint32_t strategy = LGPU_LIGHTS_TREE;
  bool committed = false;
  // power sampler (LightSampler.cpp:14-35)
  Piecewise1D powerDist;
  float invTotalPower = 0;
  // tree (ManyLightSampler)
  std::vector<Light> tlights;           // tree light records
  std::vector<int> tOrder;              // permutation of tlights
  std::vector<std::unique_ptr<LNode>> tnodes;
  LNode *root = nullptr;
  std::vector<Piecewise1D> leafDist;
  std::vector<std::pair<LNode *, unsigned>> lightLeaf;  // per tlight index
  std::vector<int> triToTLight;         // global tri -> tlight, -1
  std::vector<int> sceneToTLight;       // scene light -> tlight (point/spot), -1
  float treePower = 0, infPower = 0;
  // BVH
  std::vector<BVHNode> bvh;
  std::vector<uint32_t> bvhTris;
  Bounds sceneBounds;
};

namespace orc {
typedef oracle_scene_s Scene;

// ---------------------------------------------------------------- geometry
static inline void TriVerts(const Scene &s, uint32_t tri, V3 &v0, V3 &v1, V3 &v2) {
  const Mesh &m = s.meshes[s.triMesh[tri]];
  const uint32_t l = tri - m.firstTri;
  v0 = m.p[m.idx[3 * l]]; v1 = m.p[m.idx[3 * l + 1]]; v2 = m.p[m.idx[3 * l + 2]];
}
// TriangleMesh::GetTriangleAreaAndNormal (TriangleMesh.h:60-64)
static inline void TriAreaNormal(const V3 &v0, const V3 &v1, const V3 &v2, float *area, V3 *normal) {
  const V3 c = Cross(v1 - v0, v2 - v0);
  *area = Mag(c) * .5f;
  if (normal) *normal = c / (*area * 2.f);
}
// Moller-Trumbore; the same arithmetic order is used by the product's kernel.
static inline bool RayTri(const V3 &o, const V3 &d, const V3 &v0, const V3 &v1, const V3 &v2, float tmin, float tmax,
                          float *t, float *u, float *v) {
  const V3 e1 = v1 - v0, e2 = v2 - v0;
  const V3 pv = Cross(d, e2);
  const float det = Dot(e1, pv);
  if (det == 0.f) return false;
  const float inv = 1.f / det;
  const V3 tv = o - v0;
  const float uu = Dot(tv, pv) * inv;
  if (uu < 0.f || uu > 1.f) return false;
  const V3 qv = Cross(tv, e1);
  const float vv = Dot(d, qv) * inv;
  if (vv < 0.f || uu + vv > 1.f) return false;
  const float tt = Dot(e2, qv) * inv;
  if (!(tt > tmin && tt < tmax)) return false;
  *t = tt; *u = uu; *v = vv;
  return true;
}

// Binned SAH build (16 bins on each axis, traversal cost 1 per node against 1 per triangle, leaves
// of at most 8 triangles); the hits it returns do not depend on the tree (closest t, ties to the
// smaller triangle index), only the time to find them. A node whose centroids coincide on every
// axis, or a best split that leaves one side empty, falls back to the median of the longest axis.
static void BuildBVH(Scene &s) {
  const uint32_t nt = (uint32_t)s.triMesh.size();
  s.bvh.clear();
  s.bvhTris.resize(nt);
  for (uint32_t i = 0; i < nt; ++i) s.bvhTris[i] = i;
  std::vector<V3> cen(nt);
  std::vector<Bounds> tb(nt);
  for (uint32_t i = 0; i < nt; ++i) {
    V3 a, b, c;
    TriVerts(s, i, a, b, c);
    tb[i] = Union(Union(Bounds(a), b), c);
    cen[i] = (a + b + c) * (1.f / 3.f);
  }
  if (nt == 0) return;
  constexpr int kBins = 16;
  struct Item { uint32_t node, first, count, depth; };
  s.bvh.push_back(BVHNode());
  std::vector<Item> stack{{0, 0, nt, 0}};
  while (!stack.empty()) {
    Item it = stack.back();
    stack.pop_back();
    Bounds b = tb[s.bvhTris[it.first]], cb(cen[s.bvhTris[it.first]]);
    for (uint32_t i = it.first; i < it.first + it.count; ++i) {
      b = Union(b, tb[s.bvhTris[i]]);
      cb = Union(cb, cen[s.bvhTris[i]]);
    }
    s.bvh[it.node].b = b;
    if (it.count <= 2) {
      s.bvh[it.node].first = it.first;
      s.bvh[it.node].count = it.count;
      continue;
    }
    // SAH over the bins of each axis
    const float leafCost = (float)it.count;
    float bestCost = INFINITY;
    int bestAx = -1, bestSplit = 0;
    for (int ax = 0; ax < 3; ++ax) {
      const float lo = cb.mn[ax], ext = cb.mx[ax] - lo;
      if (!(ext > 0)) continue;
      Bounds bb[kBins];
      uint32_t bn[kBins] = {0};
      const float scale = (float)kBins / ext;
      for (uint32_t i = it.first; i < it.first + it.count; ++i) {
        const uint32_t t = s.bvhTris[i];
        int k = (int)((cen[t][ax] - lo) * scale);
        k = std::min(std::max(k, 0), kBins - 1);
        bb[k] = bn[k] ? Union(bb[k], tb[t]) : tb[t];
        bn[k]++;
      }
      float rightA[kBins];
      uint32_t rightN[kBins];
      Bounds acc;
      uint32_t n = 0;
      for (int k = kBins - 1; k >= 1; --k) {
        if (bn[k]) { acc = n ? Union(acc, bb[k]) : bb[k]; n += bn[k]; }
        rightA[k] = n ? acc.SurfaceArea() : 0.f;
        rightN[k] = n;
      }
      n = 0;
      for (int k = 0; k < kBins - 1; ++k) {
        if (bn[k]) { acc = n ? Union(acc, bb[k]) : bb[k]; n += bn[k]; }
        if (!n || !rightN[k + 1]) continue;
        const float c = 1.f + (acc.SurfaceArea() * (float)n + rightA[k + 1] * (float)rightN[k + 1]) / b.SurfaceArea();
        if (c < bestCost) { bestCost = c; bestAx = ax; bestSplit = k; }
      }
    }
    if (it.count <= 8 && !(bestCost < leafCost)) {
      s.bvh[it.node].first = it.first;
      s.bvh[it.node].count = it.count;
      continue;
    }
    uint32_t mid;
    if (bestAx >= 0 && it.depth < 96) {
      const float lo = cb.mn[bestAx], scale = (float)kBins / (cb.mx[bestAx] - lo);
      auto *first = s.bvhTris.data() + it.first;
      auto *split = std::stable_partition(first, first + it.count, [&](uint32_t t) {
        int k = (int)((cen[t][bestAx] - lo) * scale);
        return std::min(std::max(k, 0), kBins - 1) <= bestSplit;
      });
      mid = it.first + (uint32_t)(split - first);
    } else {
      mid = it.first;  // no usable split: median below
    }
    if (mid == it.first || mid == it.first + it.count) {
      const V3 ext = cb.mx - cb.mn;
      const int ax = (ext.x >= ext.y && ext.x >= ext.z) ? 0 : (ext.y >= ext.z ? 1 : 2);
      mid = it.first + it.count / 2;
      std::nth_element(s.bvhTris.begin() + it.first, s.bvhTris.begin() + mid, s.bvhTris.begin() + it.first + it.count,
                       [&](uint32_t a, uint32_t c) { return cen[a][ax] < cen[c][ax] || (cen[a][ax] == cen[c][ax] && a < c); });
    }
    const uint32_t l = (uint32_t)s.bvh.size();
    s.bvh.push_back(BVHNode());
    s.bvh.push_back(BVHNode());
    s.bvh[it.node].left = l;
    s.bvh[it.node].right = l + 1;
    s.bvh[it.node].count = 0;
    stack.push_back({l, it.first, mid - it.first, it.depth + 1});
    stack.push_back({l + 1, mid, it.first + it.count - mid, it.depth + 1});
  }
  s.sceneBounds = s.bvh[0].b;
}

static inline bool BoxHit(const Bounds &b, const V3 &o, const V3 &invd, float tmin, float tmax) {
  float t0 = tmin, t1 = tmax;
  for (int a = 0; a < 3; ++a) {
    float tn = (b.mn[a] - o[a]) * invd[a], tf = (b.mx[a] - o[a]) * invd[a];
    if (tn > tf) std::swap(tn, tf);
    if (std::isnan(tn)) tn = -INFINITY;
    if (std::isnan(tf)) tf = INFINITY;
    t0 = std::max(t0, tn);
    t1 = std::min(t1, tf * 1.0000004f);
    if (t0 > t1) return false;
  }
  return true;
}

// Entry distance of a ray into a box within [tmin, tmax] (same slab arithmetic as BoxHit), or
// INFINITY when it misses.
static inline float BoxEntry(const Bounds &b, const V3 &o, const V3 &invd, float tmin, float tmax) {
  float t0 = tmin, t1 = tmax;
  for (int a = 0; a < 3; ++a) {
    float tn = (b.mn[a] - o[a]) * invd[a], tf = (b.mx[a] - o[a]) * invd[a];
    if (tn > tf) std::swap(tn, tf);
    if (std::isnan(tn)) tn = -INFINITY;
    if (std::isnan(tf)) tf = INFINITY;
    t0 = std::max(t0, tn);
    t1 = std::min(t1, tf * 1.0000004f);
    if (t0 > t1) return INFINITY;
  }
  return t0;
}

// Closest hit; ties in t resolve to the lower triangle index. Children are visited nearest
// first (the result does not depend on the order).
static bool ClosestHit(const Scene &s, const V3 &o, const V3 &d, float tmin, float tmax, float *t, float *u, float *v,
                       uint32_t *prim) {
  if (s.bvh.empty()) return false;
  const V3 invd(1.f / d.x, 1.f / d.y, 1.f / d.z);
  bool found = false;
  float bt = tmax;
  uint32_t stack[256];
  float stackT[256];
  int sp = 0;
  if (!BoxHit(s.bvh[0].b, o, invd, tmin, bt)) return false;
  stack[sp] = 0;
  stackT[sp++] = tmin;
  while (sp) {
    --sp;
    if (stackT[sp] > bt) continue;
    const BVHNode &n = s.bvh[stack[sp]];
    if (n.count) {
      for (uint32_t i = n.first; i < n.first + n.count; ++i) {
        const uint32_t tri = s.bvhTris[i];
        V3 a, b, c;
        TriVerts(s, tri, a, b, c);
        float tt, uu, vv;
        if (RayTri(o, d, a, b, c, tmin, tmax, &tt, &uu, &vv)) {
          if (!found || tt < bt || (tt == bt && tri < *prim)) {
            found = true; bt = tt; *t = tt; *u = uu; *v = vv; *prim = tri;
          }
        }
      }
    } else {
      const float tl = BoxEntry(s.bvh[n.left].b, o, invd, tmin, bt);
      const float tr = BoxEntry(s.bvh[n.right].b, o, invd, tmin, bt);
      const bool leftFirst = tl <= tr;
      const uint32_t nearC = leftFirst ? n.left : n.right, farC = leftFirst ? n.right : n.left;
      const float nearT = leftFirst ? tl : tr, farT = leftFirst ? tr : tl;
      if (farT < INFINITY) { stack[sp] = farC; stackT[sp++] = farT; }
      if (nearT < INFINITY) { stack[sp] = nearC; stackT[sp++] = nearT; }
    }
  }
  return found;
}
static bool AnyHit(const Scene &s, const V3 &o, const V3 &d, float tmin, float tmax) {
  if (s.bvh.empty()) return false;
  const V3 invd(1.f / d.x, 1.f / d.y, 1.f / d.z);
  uint32_t stack[256];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const BVHNode &n = s.bvh[stack[--sp]];
    if (!BoxHit(n.b, o, invd, tmin, tmax)) continue;
    if (n.count) {
      for (uint32_t i = n.first; i < n.first + n.count; ++i) {
        V3 a, b, c;
        TriVerts(s, s.bvhTris[i], a, b, c);
        float tt, uu, vv;
        if (RayTri(o, d, a, b, c, tmin, tmax, &tt, &uu, &vv)) return true;
      }
    } else {
      stack[sp++] = n.left;
      stack[sp++] = n.right;
    }
  }
  return false;
}

// Scene::Intersect (Scene.cpp:28-43) + Object::Hit (Object.h:27-35) + TriangleMesh::ProcessHit (TriangleMesh.cpp:80-114)
static bool Intersect(const Scene &s, const V3 &o, const V3 &d, Hit &h, bool corrected) {
  float t, u, v;
  uint32_t prim;
  if (!ClosestHit(s, o, d, 0.f, INFINITY, &t, &u, &v, &prim)) return false;
  if (!(t > 0 && t < INFINITY)) return false;
  const Mesh &m = s.meshes[s.triMesh[prim]];
  const uint32_t l = prim - m.firstTri;
  const uint32_t i0 = m.idx[3 * l], i1 = m.idx[3 * l + 1], i2 = m.idx[3 * l + 2];
  h.tFar = t;
  h.point = V3(o.x + d.x * t, o.y + d.y * t, o.z + d.z * t);
  h.normalG = Norm(Cross(m.p[i1] - m.p[i0], m.p[i2] - m.p[i0]));
  h.prim = prim;
  h.local = l;
  h.mesh = (int)s.triMesh[prim];
  if (!m.uv.empty()) {
    const V2 &a = m.uv[i0], &b = m.uv[i1], &c = m.uv[i2];
    h.uv = V2(a.x + (b.x - a.x) * u + (c.x - a.x) * v, a.y + (b.y - a.y) * u + (c.y - a.y) * v);
  } else h.uv = V2(0, 0);
  if (!m.n.empty() && !m.tan.empty()) {
    const V3 &n0 = m.n[i0], &n1 = m.n[i1], &n2 = m.n[i2];
    const V3 t0(&m.tan[4 * i0]), t1(&m.tan[4 * i1]), t2(&m.tan[4 * i2]);
    h.normalS = Norm(n0 + (n1 - n0) * u + (n2 - n0) * v);
    h.tangent = Norm(t0 + (t1 - t0) * u + (t2 - t0) * v);
    const V3 b0 = Cross(n0, t0) * m.tan[4 * i0 + 3], b1 = Cross(n1, t1) * m.tan[4 * i1 + 3], b2 = Cross(n2, t2) * m.tan[4 * i2 + 3];
    h.bitangent = Norm(b0 + (b1 - b0) * u + (b2 - b0) * v);
  } else {
    h.normalS = h.normalG;
    const V3 c1 = Cross(h.normalG, V3(0, 0, 1));
    const V3 c2 = Cross(h.normalG, V3(0, 1, 0));
    h.tangent = (Dot(c1, c1) > Dot(c2, c2) ? c1 : c2);
    if (corrected) h.tangent = Norm(h.tangent);  // A39: unnormalised tangent (TriangleMesh.cpp:108-111)
    h.bitangent = Cross(h.tangent, h.normalG);
  }
  return true;
}
// Scene::MutualVisibility (Scene.cpp:83-95): occlusion ray with tfar = |p2-p1| - 1e-5
static bool MutualVisibility(const Scene &s, const V3 &p1, const V3 &p2, V3 *w) {
  const V3 diff = p2 - p1;
  const float mag = Mag(diff);
  const V3 dir = diff / mag;
  const float tfar = (float)((double)mag - .00001);
  if (w) *w = dir;
  return !AnyHit(s, p1, dir, 0.f, tfar);
}
static bool RayEscapes(const Scene &s, const V3 &o, const V3 &d) { return !AnyHit(s, o, d, 0.f, INFINITY); }

// ---------------------------------------------------------------- shader graph (ShaderGraph.h:82-251)
enum VT { VT_SCALAR, VT_VEC2, VT_VEC3, VT_COLOUR, VT_SPECTRUM, VT_BXDF };
struct Val { int t = VT_SCALAR; float v[4] = {0, 0, 0, 0}; Spec s; };

static Val EvalNode(const Scene &sc, const Material &m, uint32_t i, const Event &ev);
static float AsReal(const Val &x, const Lanes &l) {
  switch (x.t) {
    case VT_SCALAR: return x.v[0];
    case VT_COLOUR: return x.v[0] * .3f + x.v[1] * .59f + x.v[2] * .11f;
    case VT_VEC2: return (x.v[0] + x.v[1]) * .5f;
    case VT_VEC3: return (x.v[0] + x.v[1] + x.v[2]) * (float).333333333333;
    case VT_SPECTRUM: return Luminance(x.s, l);
  }
  return 0;
}
static void AsColour(const Val &x, float c[3]) {
  switch (x.t) {
    case VT_COLOUR: case VT_VEC3: c[0] = x.v[0]; c[1] = x.v[1]; c[2] = x.v[2]; return;
    case VT_SCALAR: c[0] = c[1] = c[2] = x.v[0]; return;
    case VT_VEC2: c[0] = c[1] = c[2] = (x.v[0] + x.v[1]) * .5f; return;
    case VT_SPECTRUM: c[0] = x.s.c[0]; c[1] = x.s.c[1]; c[2] = x.s.c[2]; return;
  }
  c[0] = 1; c[1] = 0; c[2] = 1;
}
static V3 AsVec3(const Val &x) {
  switch (x.t) {
    case VT_VEC3: return V3(x.v[0], x.v[1], x.v[2]);
    case VT_VEC2: return V3(x.v[0], x.v[1], 0);
    case VT_SCALAR: return V3(x.v[0], x.v[0], x.v[0]);
  }
  return V3(0, 0, 0);
}
static Spec FromRGB(const float c[3], const Lanes &l) {
  Spec s;
  if (l.mode == HERO4) { for (int i = 0; i < 4; ++i) s.c[i] = Uplift(c, l.lambda[i]); }
  else { s.c[0] = c[0]; s.c[1] = c[1]; s.c[2] = c[2]; s.c[3] = 0; }
  return s;
}
static Spec AsSpectrum(const Val &x, const Lanes &l) {  // ShaderGraph.h:232-251
  switch (x.t) {
    case VT_COLOUR: { float c[3]; AsColour(x, c); return FromRGB(c, l); }
    case VT_SCALAR: return Spec(AsReal(x, l));
    case VT_SPECTRUM: return x.s;
  }
  return Spec(0);
}
// Alpha of a value read as a Colour: its own for colours, 1 for every conversion (Colour.h:22-33).
static float AlphaOf(const Val &x) { return x.t == VT_COLOUR ? x.v[3] : 1.f; }
static V2 AsVec2(const Val &x) {  // GetAs<Vec2> (ShaderGraph.h:176-199)
  switch (x.t) {
    case VT_VEC2: return V2(x.v[0], x.v[1]);
    case VT_VEC3: return V2(x.v[0], x.v[1]);
    case VT_SCALAR: return V2(x.v[0], x.v[0]);
  }
  return V2(0, 0);
}

// ---- procedural textures (GraphTexture.cpp:6-240). The hashes take unsigned arguments: a float
// lattice coordinate converts as x86-64 compiles float -> unsigned (through a 64-bit signed
// integer, so negative values wrap modulo 2^32).
static inline uint32_t LatticeU32(float f) { return (uint32_t)(int64_t)f; }
static inline uint32_t Hash21(uint32_t x, uint32_t y) {  // :21-25
  x *= 1597334677u;
  y *= 3812015801u;
  return (x ^ y) * 1597334677u;
}
static inline float HashUnit(uint32_t h) { return (float)((double)h * (1. / 4294967296.)); }  // Hash * inv32
static inline V2 Hash22(V2 v) {  // :27-30
  const float a = v.x * 127.1f + v.y * 311.7f, b = v.x * 269.5f + v.y * 183.3f;
  const float sa = std::sin(a) * 18.5453f, sb = std::sin(b) * 18.5453f;
  return V2(sa - std::floor(sa), sb - std::floor(sb));
}
static inline V3 Hash33(V3 p) {  // :33-37
  p = V3(p.x * .1031f, p.y * .1030f, p.z * .0973f);
  p = V3(p.x - std::floor(p.x), p.y - std::floor(p.y), p.z - std::floor(p.z));
  const float d = (p.x * (p.y + 33.33f) + p.y * (p.x + 33.33f)) + p.z * (p.z + 33.33f);
  p = V3(p.x + d, p.y + d, p.z + d);
  const V3 r((p.x + p.y) * p.z, (p.x + p.x) * p.y, (p.y + p.x) * p.z);
  return V3(r.x - std::floor(r.x), r.y - std::floor(r.y), r.z - std::floor(r.z));
}
static inline float Quintic(float x) { return x * x * x * (x * (x * 6.f - 15.f) + 10.f); }
static inline float CubicS(float x) { return x * x * (3 - 2 * x); }
static inline float LerpF(float a, float b, float r) { return a + (b - a) * r; }

static float Perlin2D(V2 t) {  // :66-86 (t already scaled)
  const float gx = std::floor(t.x), gy = std::floor(t.y);
  auto grad = [&](float px, float py) {
    float u = HashUnit(Hash21(LatticeU32(px), LatticeU32(py)));
    u *= PI2;
    return V2(std::cos(u), std::sin(u));
  };
  const V2 g00 = grad(gx, gy), g10 = grad(gx + 1, gy), g11 = grad(gx + 1, gy + 1), g01 = grad(gx, gy + 1);
  const float dg00 = (t.x - gx) * g00.x + (t.y - gy) * g00.y;
  const float dg10 = (t.x - (gx + 1)) * g10.x + (t.y - gy) * g10.y;
  const float dg11 = (t.x - (gx + 1)) * g11.x + (t.y - (gy + 1)) * g11.y;
  const float dg01 = (t.x - gx) * g01.x + (t.y - (gy + 1)) * g01.y;
  const float vx0 = LerpF(dg00, dg10, Quintic(t.x - gx));
  const float vx1 = LerpF(dg01, dg11, Quintic(t.x - gx));
  return LerpF(vx0, vx1, Quintic(t.y - gy)) + .5f;
}
static float Value2D(V2 t) {  // :127-138
  const float x0 = std::floor(t.x), y0 = std::floor(t.y);
  const V2 p(t.x - x0, t.y - y0);
  const float v00 = HashUnit(Hash21(LatticeU32(x0), LatticeU32(y0)));
  const float v10 = HashUnit(Hash21(LatticeU32(x0 + 1), LatticeU32(y0)));
  const float v11 = HashUnit(Hash21(LatticeU32(x0 + 1), LatticeU32(y0 + 1)));
  const float v01 = HashUnit(Hash21(LatticeU32(x0), LatticeU32(y0 + 1)));
  const float vx0 = LerpF(v00, v10, CubicS(p.x));
  const float vx1 = LerpF(v01, v11, CubicS(p.x));
  return LerpF(vx0, vx1, CubicS(p.y));
}
static float Checker2D(V2 t) {  // :105-108 (t already scaled)
  const float u = t.x - std::floor(t.x), v = t.y - std::floor(t.y);
  return (u > .5f) == (v > .5f) ? 0.f : 1.f;
}
static float Checker3D(V3 t) {  // :110-113: ((x > .5) == (y > .5)) == (z > .5)
  const float x = t.x - std::floor(t.x), y = t.y - std::floor(t.y), z = t.z - std::floor(t.z);
  return ((x > .5f) == (y > .5f)) == (z > .5f) ? 0.f : 1.f;
}
static float Voronoi2D(V2 t) {  // :189-202
  const float x0 = std::floor(t.x), y0 = std::floor(t.y);
  float m = INFINITY;
  for (int y = -1; y < 2; ++y)
    for (int x = -1; x < 2; ++x) {
      const V2 o(x0 + (float)x, y0 + (float)y);
      const V2 h = Hash22(o);
      const float rx = (o.x + h.x) - t.x, ry = (o.y + h.y) - t.y;
      const float d2 = rx * rx + ry * ry;
      if (d2 < m) m = d2;
    }
  return std::sqrt(m);
}
static float Voronoi3D(V3 t) {  // :204-220
  const V3 p0(std::floor(t.x), std::floor(t.y), std::floor(t.z));
  float m = INFINITY;
  for (int z = -1; z < 2; ++z)
    for (int y = -1; y < 2; ++y)
      for (int x = -1; x < 2; ++x) {
        const V3 o(p0.x + (float)x, p0.y + (float)y, p0.z + (float)z);
        const V3 h = Hash33(o);
        const V3 r((o.x + h.x) - t.x, (o.y + h.y) - t.y, (o.z + h.z) - t.z);
        const float d2 = (r.x * r.x + r.y * r.y) + r.z * r.z;
        if (d2 < m) m = d2;
      }
  return std::sqrt(m);
}
// One noise lookup: texture type, coordinates scaled by s0 then s1 then s2 (the octave node's
// scale, octave factor and the base node's scale). At a medium interaction the 3-D variants
// are used on the event point: Perlin and value noise return 0 there as written (:88-91, :140-142).
static float NoiseAt(uint32_t type, const Event &ev, float s0, float s1, float s2) {
  if (ev.mediumInteraction) {
    const V3 &p = ev.hit->point;
    const V3 t(p.x * s0 * s1 * s2, p.y * s0 * s1 * s2, p.z * s0 * s1 * s2);
    if (type == LGPU_NODE_CHECKER) return Checker3D(t);
    if (type == LGPU_NODE_VORONOI) return Voronoi3D(t);
    return 0.f;
  }
  const V2 t(ev.hit->uv.x * s0 * s1 * s2, ev.hit->uv.y * s0 * s1 * s2);
  switch (type) {
    case LGPU_NODE_CHECKER: return Checker2D(t);
    case LGPU_NODE_PERLIN: return Perlin2D(t);
    case LGPU_NODE_VALUE_NOISE: return Value2D(t);
    default: return Voronoi2D(t);
  }
}
// RGBSpectrum::FromSampled (Spectrum.h:353-373) of n samples at increasing wavelengths.
static void SampledToRGB(const float *lam, const float *v, unsigned n, float out[3]);
static float SpdAt(const Texture &t, float l0, float l1, float lambda) {  // InterpolateSpectrumSamples
  const unsigned n = t.w * t.h;
  const float step = (l1 - l0) / (float)(n - 1);
  if (lambda <= l0) return t.data[0];
  if (lambda >= l1) return t.data[(n - 1) * t.ch];
  unsigned k = (unsigned)((lambda - l0) / step);
  if (k > n - 2) k = n - 2;
  const float a = l0 + step * (float)k;
  const float f = (lambda - a) / step;
  return t.data[k * t.ch] + (t.data[(k + 1) * t.ch] - t.data[k * t.ch]) * f;
}
static Val Input(const Scene &sc, const Material &m, uint32_t node, int k, const Event &ev) {
  const uint32_t j = m.nodes[node].in[k];
  if (j == LGPU_NONE || j >= m.nodes.size()) return Val();
  return EvalNode(sc, m, j, ev);
}
static float FrDielectric(float cosI, float etaI, float etaT);
static Val EvalNode(const Scene &sc, const Material &m, uint32_t i, const Event &ev) {
  const lgpu_node &n = m.nodes[i];
  Val r;
  switch (n.type) {
    case LGPU_NODE_SCALAR: r.t = VT_SCALAR; r.v[0] = n.p[0]; break;
    case LGPU_NODE_VEC2: r.t = VT_VEC2; r.v[0] = n.p[0]; r.v[1] = n.p[1]; break;
    case LGPU_NODE_VEC3: r.t = VT_VEC3; r.v[0] = n.p[0]; r.v[1] = n.p[1]; r.v[2] = n.p[2]; break;
    case LGPU_NODE_RGB: r.t = VT_COLOUR; r.v[0] = n.p[0]; r.v[1] = n.p[1]; r.v[2] = n.p[2]; r.v[3] = 1; break;
    case LGPU_NODE_BLACKBODY: {  // GraphInputs.cpp:82-85
      r.t = VT_SPECTRUM;
      if (ev.lanes.mode != HERO4 && i < m.bbConst.size() && m.bbConst[i]) {
        r.s.c[0] = m.bbRGB[3 * i]; r.s.c[1] = m.bbRGB[3 * i + 1]; r.s.c[2] = m.bbRGB[3 * i + 2]; r.s.c[3] = 0;
        break;
      }
      const float T = AsReal(Input(sc, m, i, 0, ev), ev.lanes);
      if (ev.lanes.mode == HERO4) {
        for (int k = 0; k < 4; ++k) r.s.c[k] = BlackbodyNormalised(ev.lanes.lambda[k], T);
      } else {
        float rgb[3];
        BlackbodyRGB(T, rgb);
        r.s.c[0] = rgb[0]; r.s.c[1] = rgb[1]; r.s.c[2] = rgb[2]; r.s.c[3] = 0;
      }
      break;
    }
    case LGPU_NODE_IMAGE_COLOUR:
    case LGPU_NODE_IMAGE_SCALAR: {  // GraphInputs.cpp:128-146: Fract(uv) then bilinear
      const Texture &t = sc.textures[(uint32_t)n.p[0]];
      float c[3];
      t.Bilinear(Fract(ev.hit->uv.x), Fract(ev.hit->uv.y), c);
      if (n.type == LGPU_NODE_IMAGE_COLOUR) {
        float c4[4];
        t.Bilinear4(Fract(ev.hit->uv.x), Fract(ev.hit->uv.y), c4);
        r.t = VT_COLOUR; r.v[0] = c[0]; r.v[1] = c[1]; r.v[2] = c[2]; r.v[3] = c4[3];
      }
      else { r.t = VT_SCALAR; r.v[0] = c[0]; }
      break;
    }
    case LGPU_NODE_UV: r.t = VT_VEC2; r.v[0] = ev.hit->uv.x; r.v[1] = ev.hit->uv.y; break;
    case LGPU_NODE_FRESNEL: {  // FresnelInput::SchlickApprox (GraphInputs.cpp:188-198)
      const float n1 = ev.eta;
      const float n2 = n.in[0] == LGPU_NONE ? n.p[0] : AsReal(Input(sc, m, i, 0, ev), ev.lanes);
      float R0 = (n1 - n2) / (n1 + n2);
      R0 *= R0;
      const float cc = 1 - ev.woL.y;
      r.t = VT_SCALAR; r.v[0] = R0 + (1 - R0) * cc * cc * cc * cc * cc;
      break;
    }
    case LGPU_NODE_SPECTRAL: {  // SpectralInput (GraphInputs.cpp:55-71): a sampled SPD
      const Texture &t = sc.textures[(uint32_t)n.p[0]];
      r.t = VT_SPECTRUM;
      if (ev.lanes.mode == HERO4) {
        for (int k = 0; k < 4; ++k) r.s.c[k] = SpdAt(t, n.p[1], n.p[2], ev.lanes.lambda[k]);
      } else {
        const unsigned cnt = t.w * t.h;
        std::vector<float> lam(cnt), v(cnt);
        for (unsigned k = 0; k < cnt; ++k) { lam[k] = n.p[1] + (n.p[2] - n.p[1]) / (float)(cnt - 1) * (float)k; v[k] = t.data[k * t.ch]; }
        float rgb[3];
        SampledToRGB(lam.data(), v.data(), cnt, rgb);
        r.s.c[0] = rgb[0]; r.s.c[1] = rgb[1]; r.s.c[2] = rgb[2]; r.s.c[3] = 0;
      }
      break;
    }
    case LGPU_NODE_IMAGE_CHANNEL: {  // ImageTextureChannelInput (GraphInputs.cpp:142-151)
      const Texture &t = sc.textures[(uint32_t)n.p[0]];
      float c[4];
      t.Bilinear4(Fract(ev.hit->uv.x), Fract(ev.hit->uv.y), c);
      r.t = VT_SCALAR; r.v[0] = c[std::min(std::max((int)n.p[1], 0), 3)];
      break;
    }
    case LGPU_NODE_SPECTRAL_TEXTURE: {  // SpectralTextureInput (GraphInputs.cpp:157-165): RGB spectra
      const Texture &t = sc.textures[(uint32_t)n.p[0]];
      float c[3];
      t.Bilinear(Fract(ev.hit->uv.x), Fract(ev.hit->uv.y), c);
      r.t = VT_SPECTRUM; r.s = FromRGB(c, ev.lanes);
      break;
    }
    case LGPU_NODE_NORMAL:  // SurfaceInfoInput (GraphInputs.cpp:171-177): shading normal
      r.t = VT_VEC3; r.v[0] = ev.hit->normalS.x; r.v[1] = ev.hit->normalS.y; r.v[2] = ev.hit->normalS.z;
      break;
    case LGPU_NODE_CROSS: {  // CrossProductNode (GraphMaths.cpp:75-84)
      const V3 o = Cross(AsVec3(Input(sc, m, i, 0, ev)), AsVec3(Input(sc, m, i, 1, ev)));
      r.t = VT_VEC3; r.v[0] = o.x; r.v[1] = o.y; r.v[2] = o.z;
      break;
    }
    case LGPU_NODE_SEPARATE_XYZ: {  // SeparateXYZ (GraphConverters.cpp:8-29)
      const V3 v = AsVec3(Input(sc, m, i, 0, ev));
      const int k = std::min(std::max((int)n.p[0], 0), 2);
      r.t = VT_SCALAR; r.v[0] = k == 0 ? v.x : (k == 1 ? v.y : v.z);
      break;
    }
    case LGPU_NODE_MERGE_XYZ:  // MergeXYZ (GraphConverters.cpp:62-75)
      r.t = VT_VEC3;
      for (int k = 0; k < 3; ++k) r.v[k] = AsReal(Input(sc, m, i, k, ev), ev.lanes);
      break;
    case LGPU_NODE_SCALAR_TO_COLOUR: {  // ScalarToColour (GraphConverters.cpp:98-106)
      const float x = AsReal(Input(sc, m, i, 0, ev), ev.lanes);
      r.t = VT_COLOUR; r.v[0] = r.v[1] = r.v[2] = x; r.v[3] = 1;
      break;
    }
    case LGPU_NODE_PERLIN: case LGPU_NODE_VALUE_NOISE: case LGPU_NODE_VORONOI:
      r.t = VT_SCALAR; r.v[0] = NoiseAt(n.type, ev, n.p[0], 1.f, 1.f);
      break;
    case LGPU_NODE_OCTAVE_NOISE: {  // OctaveNoise (GraphTexture.cpp:157-175)
      float out = 0, amp = .5f, scl = 1;
      const unsigned oct = (unsigned)n.p[1];
      for (unsigned k = 0; k < oct; ++k) {
        out += NoiseAt((uint32_t)n.p[2], ev, n.p[0], scl, n.p[3]) * amp;
        scl *= 2;
        amp *= .5f;
      }
      r.t = VT_SCALAR; r.v[0] = out;
      break;
    }
    case LGPU_NODE_SCALAR_MATHS: {  // GraphMaths.cpp:17-30
      const float a = AsReal(Input(sc, m, i, 0, ev), ev.lanes), b = AsReal(Input(sc, m, i, 1, ev), ev.lanes);
      r.t = VT_SCALAR;
      switch ((int)n.p[0]) {
        case LGPU_OP_ADD: r.v[0] = a + b; break;
        case LGPU_OP_SUBTRACT: r.v[0] = a - b; break;
        case LGPU_OP_MULTIPLY: r.v[0] = a * b; break;
        default: r.v[0] = a / b; break;
      }
      break;
    }
    case LGPU_NODE_VECTOR_MATHS: {
      const V3 a = AsVec3(Input(sc, m, i, 0, ev)), b = AsVec3(Input(sc, m, i, 1, ev));
      V3 o;
      switch ((int)n.p[0]) {
        case LGPU_OP_ADD: o = a + b; break;
        case LGPU_OP_SUBTRACT: o = a - b; break;
        case LGPU_OP_MULTIPLY: o = a * b; break;
        default: o = a / b; break;
      }
      r.t = VT_VEC3; r.v[0] = o.x; r.v[1] = o.y; r.v[2] = o.z;
      break;
    }
    case LGPU_NODE_DOT: r.t = VT_SCALAR; r.v[0] = Dot(AsVec3(Input(sc, m, i, 0, ev)), AsVec3(Input(sc, m, i, 1, ev))); break;
    case LGPU_NODE_LENGTH: r.t = VT_SCALAR; r.v[0] = Mag(AsVec3(Input(sc, m, i, 0, ev))); break;
    case LGPU_NODE_MERGE_RGB:  // MergeRGBA (GraphConverters.cpp:79-95); alpha 1 when unconnected
      r.t = VT_COLOUR;
      for (int k = 0; k < 3; ++k) r.v[k] = AsReal(Input(sc, m, i, k, ev), ev.lanes);
      r.v[3] = n.in[3] == LGPU_NONE ? 1.f : AsReal(Input(sc, m, i, 3, ev), ev.lanes);
      break;
    case LGPU_NODE_SEPARATE_R: {  // SeparateRGBA (GraphConverters.cpp:32-59)
      const Val x = Input(sc, m, i, 0, ev);
      float c[4];
      AsColour(x, c);
      c[3] = AlphaOf(x);
      r.t = VT_SCALAR; r.v[0] = c[std::min(std::max((int)n.p[0], 0), 3)];
      break;
    }
    case LGPU_NODE_CHECKER: {  // GraphTexture.cpp:95-113
      r.t = VT_SCALAR; r.v[0] = NoiseAt(LGPU_NODE_CHECKER, ev, n.p[0], 1.f, 1.f);
      break;
    }
    default: r.t = VT_BXDF; break;
  }
  return r;
}

// ---------------------------------------------------------------- BxDFs
static inline bool SameHemisphere(const V3 &a, const V3 &b) { return a.y * b.y > 0; }
static inline float CosineHemispherePdf(const V3 &wo, const V3 &wi) { return SameHemisphere(wo, wi) ? std::abs(wi.y) * INV_PI : 0; }
static inline float Cos2Theta(const V3 &w) { return w.y * w.y; }
static inline float Sin2Theta(const V3 &w) { return std::max(0.f, 1.f - Cos2Theta(w)); }
static inline float SinTheta(const V3 &w) { return std::sqrt(Sin2Theta(w)); }
static inline float TanTheta(const V3 &w) { return SinTheta(w) / w.y; }
static inline float Tan2Theta(const V3 &w) { return Sin2Theta(w) / Cos2Theta(w); }
static inline float CosPhi(const V3 &w) { const float s = SinTheta(w); return (s < 0.000001f) ? 0 : Clampf(w.x / s, -1, 1); }
static inline float SinPhi(const V3 &w) { const float s = SinTheta(w); return (s < 0.000001f) ? 0 : Clampf(w.z / s, -1, 1); }
static inline float Cos2Phi(const V3 &w) { return CosPhi(w) * CosPhi(w); }
static inline float Sin2Phi(const V3 &w) { return SinPhi(w) * SinPhi(w); }
static inline V3 Reflect(const V3 &w, const V3 &n) { return w - n * 2.f * Dot(n, w); }
static inline bool Refract(const V3 &w, const V3 &n, float eta, V3 *wr) {  // ScatterEvent.h:52-63
  const float cosI = Dot(w, n);
  const float sin2I = std::max(0.f, 1.f - cosI * cosI);
  const float sin2T = eta * eta * sin2I;
  if (sin2T >= 1) { *wr = V3(-1, 1, -1) * w; return false; }
  const float cosT = std::sqrt(1 - sin2T);
  *wr = w * -eta + n * (eta * cosI - cosT);
  return true;
}
static inline V2 SampleUnitDisk(const V2 &u) {  // Sampling.h:9-13
  const float r = std::sqrt(u.x);
  const float theta = u.y * PI2;
  return V2(r * std::cos(theta), r * std::sin(theta));
}
static inline V3 SampleCosineHemisphere(const V2 &u) {  // Sampling.h:15-19
  const V2 d = SampleUnitDisk(u);
  const float up = std::sqrt(std::max(0.f, 1.f - d.x * d.x - d.y * d.y));
  return V3(d.x, up, d.y);
}

static float FrDielectric(float cosI, float etaI, float etaT) {  // Fresnel.cpp:5-20
  cosI = Clampf(cosI, -1, 1);
  if (cosI < 0) { std::swap(etaI, etaT); cosI = std::abs(cosI); }
  const float sinI = std::sqrt(std::max(0.f, 1 - cosI * cosI));
  const float sinT = etaI / etaT * sinI;
  if (sinT >= 1) return 1;
  const float cosT = std::sqrt(std::max(0.f, 1 - sinT * sinT));
  const float rPara = (etaT * cosI - etaI * cosT) / (etaT * cosI + etaI * cosT);
  const float rPerp = (etaI * cosI - etaT * cosT) / (etaI * cosI + etaT * cosT);
  return (rPara * rPara + rPerp * rPerp) * .5f;
}
static Spec FrConductor(float cosI, const Spec &etaI, const Spec &etaT, const Spec &k) {  // Fresnel.cpp:23-45
  cosI = Clampf(cosI, -1, 1);
  const Spec eta = etaT / etaI;
  const Spec etak = k / etaI;
  const float c2 = cosI * cosI;
  const float s2 = 1.f - c2;
  const Spec eta2 = eta * eta;
  const Spec etak2 = etak * etak;
  const Spec t0 = eta2 - etak2 - s2;
  const Spec a2pb2 = SqrtS(t0 * t0 + eta2 * 4.f * etak2);
  const Spec t1 = a2pb2 + c2;
  const Spec a = SqrtS((a2pb2 + t0) * .5f);
  const Spec t2 = a * cosI * 2.f;
  const Spec rS = (t1 - t2) / (t1 + t2);
  const Spec t3 = a2pb2 * c2 + s2 * s2;
  const Spec t4 = t2 * s2;
  const Spec rP = rS * (t3 - t4) / (t3 + t4);
  return (rP + rS) * .5f;
}

// MicrofacetDistribution.cpp:11-71
static float MfD(int ggx, const V3 &wh, float ax, float ay) {
  const float tan2 = Tan2Theta(wh);
  if (std::isinf(tan2)) return 0.;
  const float cos4 = Cos2Theta(wh) * Cos2Theta(wh);
  if (!ggx) return std::exp(-tan2 * (Cos2Phi(wh) / (ax * ax) + Sin2Phi(wh) / (ay * ay))) / (PI * ax * ay * cos4);
  const float e = (Cos2Phi(wh) / (ax * ax) + Sin2Phi(wh) / (ay * ay)) * tan2;
  return 1.f / (PI * ax * ay * cos4 * (1 + e) * (1 + e));
}
static float MfLambda(int ggx, const V3 &w, float ax, float ay) {
  const float absTan = std::abs(TanTheta(w));
  if (std::isinf(absTan)) return 0.;
  const float alpha = std::sqrt(Cos2Phi(w) * ax * ax + Sin2Phi(w) * ay * ay);
  if (!ggx) {
    const float a = 1 / (alpha * absTan);
    if (a >= 1.6f) return 0;
    return (1 - 1.259f * a + 0.396f * a * a) / (3.535f * a + 2.181f * a * a);
  }
  const float a2t2 = (alpha * absTan) * (alpha * absTan);
  return (-1 + std::sqrt(1.f + a2t2)) / 2;
}
static float MfG(int ggx, const V3 &wo, const V3 &wi, float ax, float ay) {
  return 1.f / (1.f + MfLambda(ggx, wo, ax, ay) + MfLambda(ggx, wi, ax, ay));
}
static float MfPdf(int ggx, const V3 &wh, float ax, float ay) { return MfD(ggx, wh, ax, ay) * std::abs(wh.y); }
static V3 MfSampleWh(int ggx, Sampler &smp, const V3 &wo, float ax) {
  const V2 u = smp.Get2D();
  if (!ggx) {
    const float lg = u.x == 1 ? 0 : std::log(1 - u.x);
    const float phi = PI2 * u.y;
    const float tan2 = -ax * ax * lg;
    const float cosT = 1 / std::sqrt(1 + tan2);
    const float sinT = std::sqrt(std::max(0.f, 1 - cosT * cosT));
    V3 h = SphericalDirection(sinT, cosT, phi);
    return !SameHemisphere(wo, h) ? -h : h;
  }
  const float phi = PI2 * u.y;
  const float tan2 = ax * ax * u.x / (1 - u.x);
  const float cosT = 1 / std::sqrt(1 + tan2);
  const float sinT = std::sqrt(std::max(0.f, 1.f - cosT * cosT));
  V3 wh = SphericalDirection(sinT, cosT, phi);
  if (!SameHemisphere(wo, wh)) wh = -wh;
  return wh;
}
static float ToAlpha(float r) {  // Microfacet.cpp:49-55
  r = std::max(r, 1e-3f);
  const float x = std::log(r);
  return 1.62142f + 0.819955f * x + 0.1734f * x * x + 0.0171201f * x * x * x + 0.000640711f * x * x * x * x;
}

struct Ctx {
  const Scene *sc;
  bool corrected;
  int mode;
  bool volumes = false;  // Scene::hasVolumes: never set in the reference (A42); corrected: scene has media
};

static Spec BxDF_f(const Ctx &c, const Material &m, uint32_t node, Event &ev);
static float BxDF_Pdf(const Ctx &c, const Material &m, uint32_t node, const V3 &wo, const V3 &wi, Event &ev);

static Spec SockSpec(const Ctx &c, const Material &m, uint32_t node, int k, const Event &ev) {
  return AsSpectrum(Input(*c.sc, m, node, k, ev), ev.lanes);
}
static float SockReal(const Ctx &c, const Material &m, uint32_t node, int k, const Event &ev) {
  return AsReal(Input(*c.sc, m, node, k, ev), ev.lanes);
}
static void MfAlpha(const Ctx &c, const Material &m, uint32_t node, const Event &ev, float *ax, float *ay) {
  const lgpu_node &n = m.nodes[node];
  const Val r = Input(*c.sc, m, node, 1, ev);
  float a, b;
  if (r.t == VT_VEC2) { a = r.v[0]; b = r.v[1]; }
  else { a = b = AsReal(r, ev.lanes); }
  if (n.p[9] == 0.f) { a = ToAlpha(a); b = ToAlpha(b); }
  *ax = a; *ay = b;
}
static Spec ConductorSpec(const float rgb[3], const Lanes &l) { return FromRGB(rgb, l); }

static float OrenNayarTerm(const Event &ev, float sigma) {  // OrenNayar.cpp:10-21
  const float thetaI = std::acos(Clampf(ev.wiL.y, -1, 1));
  const float thetaO = std::acos(Clampf(ev.woL.y, -1, 1));
  const float alpha = std::max(thetaI, thetaO);
  const float beta = std::min(thetaI, thetaO);
  const float phiI = std::acos(CosPhi(ev.wiL));
  const float phiO = std::acos(CosPhi(ev.woL));
  const float s2 = sigma * sigma;
  const float A = 1.f - s2 / (2.f * (s2 + .33f));
  const float B = (.45f * s2) / (s2 + .09f);
  return A + B * std::max(0.f, std::cos(phiI - phiO) * Clampf(std::sin(alpha), -1, 1) * Clampf(std::tan(beta), -1, 1));
}

// Rough dielectric (project extension; Walter et al. 2007 GGX, isotropic).
static void RoughDielectricEval(const V3 &wo, const V3 &wi, float alpha, float ior, float *fOut, float *pdfOut) {
  *fOut = 0; *pdfOut = 0;
  if (wo.y == 0 || wi.y == 0) return;
  const bool reflect = wo.y * wi.y > 0;
  const float etaI = wo.y > 0 ? 1.f : ior, etaT = wo.y > 0 ? ior : 1.f;
  V3 wh;
  if (reflect) wh = wo + wi;
  else wh = wo * etaI + wi * etaT;
  const float l = Mag(wh);
  if (l == 0) return;
  wh = wh / l;
  if (wh.y < 0) wh = -wh;
  const float dwo = Dot(wo, wh), dwi = Dot(wi, wh);
  // refraction needs wo and wi on opposite sides of the microfacet and wo in front of it (the
  // sampler reaches no refraction through a microfacet that faces away from wo)
  if (!reflect && (dwo * dwi >= 0 || dwo * wo.y <= 0)) return;
  const float F = FrDielectric(dwo, 1.f, ior);
  const float D = MfD(1, wh, alpha, alpha);
  const float G = MfG(1, wo, wi, alpha, alpha);
  const float pwh = D * std::abs(wh.y);
  if (reflect) {
    *fOut = D * G * F / (4.f * std::abs(wo.y) * std::abs(wi.y));
    *pdfOut = pwh * F / (4.f * std::abs(dwo));
  } else {
    const float sq = etaI * dwo + etaT * dwi;
    const float factor = (etaI * etaI) / (etaT * etaT);
    *fOut = (1.f - F) * D * G * etaT * etaT * std::abs(dwi) * std::abs(dwo) / (std::abs(wo.y) * std::abs(wi.y) * sq * sq) * factor;
    *pdfOut = pwh * (1.f - F) * etaT * etaT * std::abs(dwi) / (sq * sq);
  }
}

static Spec BxDF_f(const Ctx &c, const Material &m, uint32_t node, Event &ev) {
  const lgpu_node &n = m.nodes[node];
  switch (n.type) {
    case LGPU_NODE_LAMBERTIAN: return SockSpec(c, m, node, 0, ev) * INV_PI;  // BxDF.cpp:29-31
    case LGPU_NODE_OREN_NAYAR_BRDF:
    case LGPU_NODE_OREN_NAYAR_BTDF: {  // OrenNayar.cpp:23-27,44-48
      const float oN = OrenNayarTerm(ev, SockReal(c, m, node, 1, ev));
      const Spec alb = SockSpec(c, m, node, 0, ev);
      return alb * INV_PI * oN;
    }
    case LGPU_NODE_MICROFACET_BRDF: {  // Microfacet.cpp:12-25
      const float cosO = std::abs(ev.woL.y), cosI = std::abs(ev.wiL.y);
      V3 wh = ev.wiL + ev.woL;
      if (cosI == 0 || cosO == 0) return Spec(0);
      if (wh.x == 0 && wh.y == 0 && wh.z == 0) return Spec(0);
      wh = Norm(wh);
      const V3 wiL(ev.wiL.x, Clampf(ev.wiL.y, 0, 1), ev.wiL.z);
      const int ggx = n.p[0] != 0.f;
      Spec F;
      const float cosH = Dot(wiL, wh);
      if (n.p[1] != 0.f) {  // FresnelConductor::Evaluate(cos, etaT = 1e-5)  (A18)
        const Spec eta = ConductorSpec(&n.p[3], ev.lanes), k = ConductorSpec(&n.p[6], ev.lanes);
        if (c.corrected) F = FrConductor(cosH, Spec(1.f), eta, k);
        else F = FrConductor(cosH, eta, Spec(0.00001f), k);
      } else {  // FresnelDielectric::Evaluate(cos, etaI = 1e-5) -> FrDielectric(cos, etaT, 1e-5)
        if (c.corrected) F = Spec(FrDielectric(cosH, 1.f, n.p[2]));
        else F = Spec(FrDielectric(cosH, n.p[2], 0.00001f));
      }
      float ax, ay;
      MfAlpha(c, m, node, ev, &ax, &ay);
      const float D = MfD(ggx, wh, ax, ay);
      const float G = MfG(ggx, ev.woL, wiL, ax, ay);
      const Spec alb = SockSpec(c, m, node, 0, ev);
      return alb * D * G * F / (4 * cosI * cosO);
    }
    case LGPU_NODE_ROUGH_DIELECTRIC: {
      float f, pdf;
      RoughDielectricEval(ev.woL, ev.wiL, SockReal(c, m, node, 1, ev), n.p[0], &f, &pdf);
      return SockSpec(c, m, node, 0, ev) * f;
    }
    case LGPU_NODE_MIX_BSDF: {  // BxDF.cpp:45-48
      const float ratio = SockReal(c, m, node, 2, ev);
      return BxDF_f(c, m, n.in[0], ev) * (1.f - ratio) + BxDF_f(c, m, n.in[1], ev) * ratio;
    }
    default: return Spec(0);  // specular / ghost: f = 0
  }
}

static float BxDF_Pdf(const Ctx &c, const Material &m, uint32_t node, const V3 &wo, const V3 &wi, Event &ev) {
  const lgpu_node &n = m.nodes[node];
  switch (n.type) {
    case LGPU_NODE_LAMBERTIAN:
    case LGPU_NODE_OREN_NAYAR_BRDF:
    case LGPU_NODE_OREN_NAYAR_BTDF: return CosineHemispherePdf(wo, wi);  // BxDF.cpp:19-21
    case LGPU_NODE_MICROFACET_BRDF: {  // Microfacet.cpp:40-45
      if (!SameHemisphere(wo, wi)) return 0;
      const V3 wh = Norm(wo + wi);
      float ax, ay;
      MfAlpha(c, m, node, ev, &ax, &ay);
      return MfPdf(n.p[0] != 0.f, wh, ax, ay) / (4 * Dot(wo, wh));
    }
    case LGPU_NODE_ROUGH_DIELECTRIC: {
      float f, pdf;
      RoughDielectricEval(wo, wi, SockReal(c, m, node, 1, ev), n.p[0], &f, &pdf);
      return pdf;
    }
    case LGPU_NODE_MIX_BSDF: {  // BxDF.cpp:62-65
      const float ratio = SockReal(c, m, node, 2, ev);
      return BxDF_Pdf(c, m, n.in[0], wo, wi, ev) * (1.f - ratio) + BxDF_Pdf(c, m, n.in[1], wo, wi, ev) * ratio;
    }
    default: return 0;
  }
}

static bool IsSpecular(const Material &m, uint32_t node) {
  const uint32_t t = m.nodes[node].type;
  return t == LGPU_NODE_FRESNEL_BSDF || t == LGPU_NODE_SPECULAR_BRDF || t == LGPU_NODE_SPECULAR_BTDF || t == LGPU_NODE_GHOST_BTDF;
}

// Sample_f; *specOut reports a delta lobe for the corrected MIS weights.
static Spec BxDF_Sample(const Ctx &c, const Material &m, uint32_t node, Event &ev, Sampler &smp, float &pdf, bool *specOut) {
  const lgpu_node &n = m.nodes[node];
  Hit &h = *ev.hit;
  *specOut = false;
  switch (n.type) {
    case LGPU_NODE_LAMBERTIAN:
    case LGPU_NODE_OREN_NAYAR_BRDF: {  // BxDF.cpp:7-14
      ev.wiL = SampleCosineHemisphere(smp.Get2D());
      if (ev.woL.y < 0) ev.wiL.y *= -1;
      pdf = CosineHemispherePdf(ev.woL, ev.wiL);
      ev.wi = ev.ToWorld(ev.wiL);
      h.point += h.normalG * SURFACE_EPSILON * (float)ev.sidedness;
      return BxDF_f(c, m, node, ev);
    }
    case LGPU_NODE_OREN_NAYAR_BTDF: {  // OrenNayar.cpp:52-59
      ev.wiL = SampleCosineHemisphere(smp.Get2D());
      pdf = CosineHemispherePdf(ev.woL, ev.wiL);
      ev.wiL.y *= (float)-ev.sidedness;
      ev.wi = ev.ToWorld(ev.wiL);
      h.point += h.normalG * (float)-ev.sidedness * SURFACE_EPSILON;
      return BxDF_f(c, m, node, ev);
    }
    case LGPU_NODE_MICROFACET_BRDF: {  // Microfacet.cpp:27-38
      if (ev.woL.y == 0) { pdf = 0; return Spec(0); }
      float ax, ay;
      MfAlpha(c, m, node, ev, &ax, &ay);
      const int ggx = n.p[0] != 0.f;
      const V3 wh = Norm(MfSampleWh(ggx, smp, ev.woL, ax));
      ev.wiL = Reflect(-ev.woL, wh);
      if (!SameHemisphere(ev.woL, ev.wiL)) { pdf = 0; return Spec(0); }
      pdf = MfPdf(ggx, wh, ax, ay) / (4 * Dot(ev.woL, wh));
      ev.wi = ev.ToWorld(ev.wiL);
      h.point += h.normalG * SURFACE_EPSILON;
      return BxDF_f(c, m, node, ev);
    }
    case LGPU_NODE_ROUGH_DIELECTRIC: {
      const float alpha = SockReal(c, m, node, 1, ev), ior = n.p[0];
      const V2 u = smp.Get2D();
      const float uc = smp.Get1D();
      // GGX normal sampling in the hemisphere of wo (same warp as TrowbridgeReitz::Sample_wh)
      const float phi = PI2 * u.y;
      const float tan2 = alpha * alpha * u.x / (1 - u.x);
      const float cosT = 1 / std::sqrt(1 + tan2);
      const float sinT = std::sqrt(std::max(0.f, 1.f - cosT * cosT));
      V3 wh = SphericalDirection(sinT, cosT, phi);
      // reflect/refract chosen with the Fresnel of wo's side, the one RoughDielectricEval's pdf
      // uses: the half vector is still on the +y side here, so a wo inside the glass (negative
      // cosine) gets the glass-to-air reflectance (total internal reflection included)
      const float F = FrDielectric(Dot(ev.woL, wh), 1.f, ior);
      if (ev.woL.y < 0) wh = -wh;
      if (uc < F) {
        ev.wiL = Reflect(-ev.woL, wh);
        if (!SameHemisphere(ev.woL, ev.wiL)) { pdf = 0; return Spec(0); }
      } else {
        const float etaI = ev.woL.y > 0 ? 1.f : ior, etaT = ev.woL.y > 0 ? ior : 1.f;
        if (!(Dot(ev.woL, wh) > 0)) { pdf = 0; return Spec(0); }  // microfacet faces away from wo
        if (!Refract(ev.woL, wh, etaI / etaT, &ev.wiL)) { pdf = 0; return Spec(0); }
        if (SameHemisphere(ev.woL, ev.wiL) || ev.wiL.y == 0) { pdf = 0; return Spec(0); }
      }
      float f;
      RoughDielectricEval(ev.woL, ev.wiL, alpha, ior, &f, &pdf);
      ev.wi = ev.ToWorld(ev.wiL);
      // the next ray starts on the side wi points into (reflection: wo's side; refraction: across)
      h.point += h.normalG * (ev.wiL.y > 0 ? SURFACE_EPSILON : -SURFACE_EPSILON);
      return SockSpec(c, m, node, 0, ev) * f;
    }
    case LGPU_NODE_FRESNEL_BSDF: {  // Specular.cpp:18-45
      *specOut = true;
      const bool entering = ev.woL.y > 0;
      const float ior = SockReal(c, m, node, 1, ev);
      const float fr = FrDielectric(ev.woL.y, 1, ior);
      if (smp.Get1D() < fr) {
        ev.wiL = V3(-ev.woL.x, ev.woL.y, -ev.woL.z);
        ev.wi = ev.ToWorld(ev.wiL);
        const float cosT = std::abs(ev.wiL.y);
        h.point += h.normalG * (entering ? SURFACE_EPSILON : -SURFACE_EPSILON);
        pdf = fr;
        return SockSpec(c, m, node, 0, ev) * fr / cosT;
      }
      const float etaI = ev.eta = entering ? 1 : ior;
      const float etaT = entering ? ior : 1;
      const V3 nrm = Dot(V3(0, 1, 0), ev.woL) < 0 ? V3(0, -1, 0) : V3(0, 1, 0);
      if (!Refract(ev.woL, nrm, etaI / etaT, &ev.wiL)) {  // A17: TIR returns 0 with the pdf unset
        pdf = 0;
        return Spec(0);
      }
      Spec ft = SockSpec(c, m, node, 0, ev) * (1 - fr);
      ft = ft * ((etaI * etaI) / (etaT * etaT));
      const float cosT = std::abs(ev.wiL.y);
      ev.wi = ev.ToWorld(ev.wiL);
      h.point += h.normalG * (entering ? -SURFACE_EPSILON * 2 : SURFACE_EPSILON * 2);
      pdf = 1.f - fr;
      return ft / cosT;
    }
    case LGPU_NODE_SPECULAR_BRDF: {  // Specular.cpp:62-68
      *specOut = true;
      pdf = 1;
      ev.wiL = V3(-ev.woL.x, ev.woL.y, -ev.woL.z);
      ev.wi = ev.ToWorld(ev.wiL);
      h.point += h.normalG * SURFACE_EPSILON;
      return SockSpec(c, m, node, 0, ev) / std::abs(ev.wiL.y);
    }
    case LGPU_NODE_SPECULAR_BTDF: {  // Specular.cpp:76-95 (normalS flip and world wi.y kept as written)
      *specOut = true;
      pdf = 1;
      const V3 woL = ev.ToLocal(ev.wo);
      const bool outside = woL.y > 0;
      h.normalS = h.normalS * (outside ? 1.f : -1.f);
      const float etaT0 = n.p[0];
      const float etaI = outside ? ev.eta : etaT0;
      const float etaT = outside ? etaT0 : ev.eta;
      Refract(ev.wo, V3(0, 1, 0) * (outside ? 1.f : -1.f), etaI / etaT, &ev.wiL);
      // FresnelDielectric::Evaluate(cos, etaI) -> FrDielectric(cos, this->etaT, etaI)
      const float F = FrDielectric(std::abs(ev.wi.y), etaT0, etaT);
      Spec ft = SockSpec(c, m, node, 0, ev) * (Spec(1) - Spec(F));
      ft = ft * ((etaI * etaI) / (etaT * etaT));
      const float inv = 1.f / std::abs(ev.wiL.y);
      ev.wi = ev.ToWorld(ev.wiL);
      h.point += h.normalS * .0004f;
      return ft * inv;
    }
    case LGPU_NODE_GHOST_BTDF: {  // Ghost.cpp:13-28
      *specOut = true;
      pdf = 1;
      const float alpha = SockReal(c, m, node, 0, ev);
      const bool inside = ev.woL.y < 0;
      if (alpha > 0) {
        ev.wi = -ev.wo;
        ev.wiL = ev.ToLocal(ev.wi);
        h.point += h.normalG * SURFACE_EPSILON * 2.f * (inside ? 1.f : -1.f);
        return Spec(alpha / std::abs(ev.wiL.y));
      }
      pdf = 0;
      return Spec(0);
    }
    case LGPU_NODE_MIX_BSDF: {  // BxDF.cpp:50-60
      const float ratio = SockReal(c, m, node, 2, ev);
      bool sp;
      if (smp.Get1D() > ratio) {
        const Spec af = BxDF_Sample(c, m, n.in[0], ev, smp, pdf, &sp);
        if (c.corrected && !sp) pdf = BxDF_Pdf(c, m, node, ev.woL, ev.wiL, ev);  // A20
        else if (c.corrected) pdf *= (1.f - ratio);
        *specOut = sp;
        return af * (1.f - ratio) + BxDF_f(c, m, n.in[1], ev) * ratio;
      }
      const Spec bf = BxDF_Sample(c, m, n.in[1], ev, smp, pdf, &sp);
      if (c.corrected && !sp) pdf = BxDF_Pdf(c, m, node, ev.woL, ev.wiL, ev);
      else if (c.corrected) pdf *= ratio;
      *specOut = sp;
      return bf * ratio + BxDF_f(c, m, n.in[0], ev) * (1.f - ratio);
    }
  }
  pdf = 0;
  return Spec(0);
}

// ---------------------------------------------------------------- media (HomogeneousMedium.cpp, Media.h, Scene.cpp:45-81)
static inline int NSamples(const Lanes &l) { return l.mode == HERO4 ? 4 : 3; }
static inline Spec ExpS(const Spec &s) { Spec r; for (int i = 0; i < 4; ++i) r.c[i] = std::exp(s.c[i]); return r; }
static Spec SigmaT(const MediumO &m, const Lanes &l) { return FromRGB(m.sa, l) + FromRGB(m.ss, l); }
static Spec SigmaS(const MediumO &m, const Lanes &l) { return FromRGB(m.ss, l); }
// MediaBoundary::GetMedium (Media.h:53-55)
static int GetMedium(const Material &mat, const V3 &w, const V3 &n) {
  const uint32_t id = Dot(w, n) < 0 ? mat.medIn : mat.medOut;
  return id == LGPU_NONE ? -1 : (int)id;
}

// Grid traversal: visits the voxels of [t0, t1] along o + t d in order; fn(density, ta, tb) returns
// true to stop. Piecewise-constant (nearest-voxel) density.
template <class F>
static void GridWalk(const MediumO &m, const V3 &o, const V3 &d, float t0, float t1, const F &fn) {
  float tn = t0, tf = t1;
  for (int a = 0; a < 3; ++a) {
    const float inv = 1.f / d[a];
    float ta = (m.bmin[a] - o[a]) * inv, tb = (m.bmax[a] - o[a]) * inv;
    if (ta > tb) std::swap(ta, tb);
    if (ta > tn) tn = ta;
    if (tb < tf) tf = tb;
  }
  if (!(tn < tf)) return;
  const int n[3] = {(int)m.nx, (int)m.ny, (int)m.nz};
  float cs[3], tNext[3], tDelta[3];
  int cell[3], step[3];
  const V3 p = o + d * tn;
  for (int a = 0; a < 3; ++a) {
    cs[a] = (m.bmax[a] - m.bmin[a]) / (float)n[a];
    int c = (int)std::floor((p[a] - m.bmin[a]) / cs[a]);
    cell[a] = std::min(std::max(c, 0), n[a] - 1);
    if (d[a] > 0) { step[a] = 1; tNext[a] = (m.bmin[a] + (float)(cell[a] + 1) * cs[a] - o[a]) / d[a]; tDelta[a] = cs[a] / d[a]; }
    else if (d[a] < 0) { step[a] = -1; tNext[a] = (m.bmin[a] + (float)cell[a] * cs[a] - o[a]) / d[a]; tDelta[a] = -cs[a] / d[a]; }
    else { step[a] = 0; tNext[a] = INFINITY; tDelta[a] = INFINITY; }
  }
  float t = tn;
  for (int it = 0; it < n[0] + n[1] + n[2] + 3; ++it) {
    const int ax = (tNext[0] <= tNext[1] && tNext[0] <= tNext[2]) ? 0 : (tNext[1] <= tNext[2] ? 1 : 2);
    const float tc = std::max(t, std::min(tNext[ax], tf));  // entry rounding can put tNext before t
    const float rho = m.dens[((size_t)cell[2] * m.ny + cell[1]) * m.nx + cell[0]];
    if (fn(rho, t, tc)) return;
    t = tc;
    if (t >= tf) return;
    cell[ax] += step[ax];
    if (cell[ax] < 0 || cell[ax] >= n[ax]) return;
    tNext[ax] += tDelta[ax];
  }
}
// Optical-depth density integral over [0, tFar].
static float DensityIntegral(const MediumO &m, const V3 &o, const V3 &d, float tFar) {
  float acc = 0.f;
  GridWalk(m, o, d, 0.f, tFar, [&](float rho, float ta, float tb) { acc += rho * (tb - ta); return false; });
  return acc;
}
// Medium::Tr (HomogeneousMedium.cpp:12-15)
static Spec MediumTr(const MediumO &m, const V3 &o, const V3 &d, float tFar, const Lanes &l) {
  const Spec st = SigmaT(m, l);
  if (m.type == LGPU_MEDIUM_HOMOGENEOUS) return ExpS(st * -1.f * std::min(tFar, 3.40282347e38f));
  return ExpS(st * -1.f * DensityIntegral(m, o, d, tFar));
}
// HomogeneousMedium::SampleDistance (HomogeneousMedium.cpp:17-38); the grid walks voxels to the
// sampled optical depth of the chosen channel.
static Spec SampleDistanceU(const MediumO &m, const V3 &o, const V3 &d, float tFarHit, const V2 &u, const Lanes &l,
                            bool *interaction, float *tOut, float *pdfOut) {
  const Spec st = SigmaT(m, l), ss = SigmaS(m, l);
  const int ns = NSamples(l);
  const unsigned ch = std::min((unsigned)(u.x * (float)ns), (unsigned)ns - 1);
  float t, rhoAt = 1.f;
  float D = 0.f;
  if (m.type == LGPU_MEDIUM_HOMOGENEOUS) {
    const float dist = -std::log(1 - u.y) / st.c[ch];
    t = std::min(dist, tFarHit);
  } else {
    const float target = -std::log(1 - u.y);
    t = tFarHit;
    float acc = 0.f;
    bool found = false;
    GridWalk(m, o, d, 0.f, tFarHit, [&](float rho, float ta, float tb) {
      const float k = st.c[ch] * rho;
      const float seg = k * (tb - ta);
      if (k > 0 && acc + seg >= target) {
        t = ta + (target - acc) / (st.c[ch] * rho);
        D += rho * (t - ta);
        rhoAt = rho;
        found = true;
        return true;
      }
      acc += seg;
      D += rho * (tb - ta);
      return false;
    });
    if (!found) t = tFarHit;
  }
  *interaction = t < tFarHit;
  const Spec Tr = m.type == LGPU_MEDIUM_HOMOGENEOUS ? ExpS(st * -1.f * std::min(t, 3.40282347e38f)) : ExpS(st * -1.f * D);
  const Spec stl = st * rhoAt, ssl = ss * rhoAt;
  const Spec dens = *interaction ? stl * Tr : Tr;
  float pdf = 0;
  for (int i = 0; i < ns; ++i) pdf += dens.c[i];
  pdf *= 1 / (float)ns;
  if (pdf == 0) pdf = 1;
  *tOut = t;
  *pdfOut = pdf;
  return *interaction ? Tr * ssl / pdf : Tr / pdf;
}
static Spec SampleDistance(const MediumO &m, const V3 &o, const V3 &d, float tFarHit, Sampler &smp, const Lanes &l,
                           bool *interaction, float *tOut, float *pdfOut) {
  return SampleDistanceU(m, o, d, tFarHit, smp.Get2D(), l, interaction, tOut, pdfOut);
}
// HomogeneousMedium::PDFDistance (:40-47): the reference multiplies by `1 / Spectrum::nSamples`
// (integer division, 0); corrected returns the channel-averaged pdf.
static float PdfDistance(const MediumO &m, const V3 &o, const V3 &d, float t, const Lanes &l, bool corrected) {
  if (!corrected) return 0.f;
  const Spec st = SigmaT(m, l);
  float rho = 1.f, D = t;
  if (m.type != LGPU_MEDIUM_HOMOGENEOUS) {
    D = 0.f;
    rho = 0.f;
    GridWalk(m, o, d, 0.f, t, [&](float r, float ta, float tb) { D += r * (tb - ta); rho = r; return false; });
  }
  const Spec Tr = ExpS(st * -1.f * D);
  const int ns = NSamples(l);
  float pdf = 0;
  for (int i = 0; i < ns; ++i) pdf += st.c[i] * rho * Tr.c[i];
  return pdf * (1 / (float)ns);
}
// HomogeneousMedium::SampleEquiangular (:49-78). faithful: origin from the surface hit point
// (A14), t = D tan(theta) used as the ray distance, the unused Get2D kept; corrected: origin
// ray.o and t = l + D tan(theta).
static Spec SampleEquiangularU(const MediumO &m, const V3 &o, const V3 &d, const V3 &hitPoint, float tFarHit,
                               const V3 &lightPoint, float eps, const Lanes &l, bool corrected, bool *interaction,
                               float *tOut, float *pdfOut) {
  const V3 delta = lightPoint - (corrected ? o : hitPoint);
  const float lp = Dot(d, delta);
  const float D = std::sqrt(corrected ? std::max(0.f, Dot(delta, delta) - lp * lp) : Dot(delta, delta) - lp * lp);
  if (corrected && !(D > 0)) {  // light point on the ray's line: no equiangular sample
    *interaction = false;
    *tOut = tFarHit;
    *pdfOut = 0;
    return Spec(0);
  }
  const float thetaA = -std::atan2(lp, D);
  const float thetaB = std::atan2(tFarHit - lp, D);
  const float th = thetaB + (thetaA - thetaB) * eps;
  const float dist = D * std::tan(th);
  float t = corrected ? std::min(lp + dist, tFarHit) : std::min(dist, tFarHit);
  *interaction = t < tFarHit;
  const float x = corrected ? t - lp : t;
  const float pdf = D / ((thetaB - thetaA) * (D * D + x * x));
  *tOut = t;
  *pdfOut = pdf;
  const Spec st = SigmaT(m, l), ss = SigmaS(m, l);
  float Dn = t, rho = 1.f;
  if (m.type != LGPU_MEDIUM_HOMOGENEOUS) {
    Dn = 0.f;
    rho = 0.f;
    GridWalk(m, o, d, 0.f, t, [&](float r, float ta, float tb) { Dn += r * (tb - ta); rho = r; return false; });
  }
  const Spec Tr = ExpS(st * -1.f * std::min(Dn, 3.40282347e38f));
  return Tr * (ss * rho) / pdf;
}
static Spec SampleEquiangular(const MediumO &m, const V3 &o, const V3 &d, const V3 &hitPoint, float tFarHit,
                              const V3 &lightPoint, Sampler &smp, const Lanes &l, bool corrected, bool *interaction,
                              float *tOut, float *pdfOut) {
  const float eps = smp.Get1D();
  smp.Get2D();  // channel draw, unused (HomogeneousMedium.cpp:70)
  return SampleEquiangularU(m, o, d, hitPoint, tFarHit, lightPoint, eps, l, corrected, interaction, tOut, pdfOut);
}
// HomogeneousMedium::PDFEquiangular (:80-92)
static float PdfEquiangular(const V3 &o, const V3 &d, const V3 &lightPoint, float tFar, float t, bool corrected) {
  const V3 delta = lightPoint - o;
  const float lp = Dot(d, delta);
  const float D = std::sqrt(Dot(delta, delta) - lp * lp);
  const float thetaA = -std::atan2(lp, D);
  const float thetaB = std::atan2(tFar - lp, D);
  const float x = corrected ? t - lp : t;
  return D / ((thetaB - thetaA) * (D * D + x * x));
}
// Henyey-Greenstein (HenyeyGreenstein.cpp:8-36). faithful: frame from the unnormalised
// cross(-wo, y) (A21); corrected: orthonormal frame.
static inline float PhaseHG(float cosT, float g) {
  const float g2 = g * g;
  const float denom = 1 + g2 + 2 * g * cosT;
  return INV_PI4 * (1 - g2) / (denom * std::sqrt(denom));
}
static float HGSampleU(float g, const V3 &wo, V3 *wi, const V2 &u, bool corrected) {
  float cosT;
  if (std::abs(g) < 1e-3f) cosT = 1 - 2 * u.x;
  else {
    const float sq = (1 - g * g) / (1 - g + 2 * g * u.x);
    cosT = (1 + g * g - sq * sq) / (2 * g);
  }
  const float sinT = std::sqrt(std::max(0.f, 1 - cosT * cosT));
  const float phi = 2 * PI * u.y;
  const V3 woNeg = -wo;
  V3 v1 = Cross(woNeg, V3(0, 1, 0));
  if (corrected) {
    if (Dot(v1, v1) < 1e-12f) v1 = Cross(woNeg, V3(1, 0, 0));
    v1 = Norm(v1);
  }
  const V3 v2 = Cross(woNeg, v1);
  const V3 dd = SphericalDirection(sinT, cosT, phi);
  *wi = v1 * dd.x + woNeg * dd.y + v2 * dd.z;
  return PhaseHG(-cosT, g);
}
static float HGSample(float g, const V3 &wo, V3 *wi, Sampler &smp, bool corrected) {
  return HGSampleU(g, wo, wi, smp.Get2D(), corrected);
}

// Scene::IntersectTr (Scene.cpp:45-81, 280-300): skips surfaces without BxDF or light,
// multiplying the transmittance of the medium each segment crosses. maxT < 0: no limit.
static bool IntersectTr(const Scene &sc, V3 o, V3 d, Hit &h, int med, Spec *Tr, const Lanes &l, float maxT,
                        bool corrected, bool *leftAll) {
  float tFar = 0;
  if (leftAll) *leftAll = false;
  for (int k = 0; k < 256; ++k) {
    if (!Intersect(sc, o, d, h, corrected)) {
      if (leftAll) *leftAll = true;
      return false;
    }
    tFar += h.tFar;
    const bool terminate = maxT >= 0 && tFar > maxT;
    if (terminate) { const float delta = tFar - maxT; tFar -= delta; h.tFar -= delta; }
    if (med >= 0 && Tr) *Tr = *Tr * MediumTr(sc.media[med], o, d, h.tFar, l);
    const Material &mat = sc.materials[sc.meshes[h.mesh].material];
    if (terminate || mat.bxdf != LGPU_NONE || mat.emission != LGPU_NONE) { h.tFar = tFar; return true; }
    med = GetMedium(mat, d, h.normalG);
    o = h.point + h.normalG * (Dot(h.normalG, d) < 0 ? -SURFACE_EPSILON : SURFACE_EPSILON);
  }
  return false;
}
// Light::MutualVisibility / PointMutualVisibility / RayEscapes (Light.cpp:10-49) on the
// hasVolumes branch, reached only in corrected mode (A42). The ray starts in the medium the
// path is in at a medium vertex, or in the one GetMedium gives for the shading surface. A ray
// that leaves every boundary without another hit is unoccluded (the reference compares the
// last segment's tFar with minT there, A41).
static int NeeMedium(const Scene &sc, const Event &ev, const V3 &dir) {
  if (ev.mediumInteraction) return ev.medium;
  const Hit &h = *ev.hit;
  return GetMedium(sc.materials[sc.meshes[h.mesh].material], dir, h.normalG);
}
static bool VisibleTr(const Ctx &c, const V3 &p1, const V3 &p2, Event &ev, Spec *Tr, bool point) {
  const V3 diff = p2 - p1;
  const float mag = Mag(diff);
  const V3 dir = diff / mag;
  const float minT = mag - SURFACE_EPSILON * 10;
  ev.wi = dir;
  Hit h;
  bool left;
  IntersectTr(*c.sc, p1, dir, h, NeeMedium(*c.sc, ev, dir), Tr, ev.lanes, point ? mag : -1.f, true, &left);
  if (left) return true;
  return h.tFar > minT;
}
static bool EscapesTr(const Ctx &c, const V3 &o, const V3 &d, const Event &ev, Spec *Tr) {
  Hit h;
  return !IntersectTr(*c.sc, o, d, h, NeeMedium(*c.sc, ev, d), Tr, ev.lanes, -1.f, true, nullptr);
}

static Spec Emission(const Ctx &c, const Material &m, const Event &ev) {
  if (m.emission == LGPU_NONE) return Spec(0);
  return AsSpectrum(EvalNode(*c.sc, m, m.emission, ev), ev.lanes);
}
static Spec LightColour(const Light &L, const Lanes &l) {
  if (L.temperature > 0) {
    Spec s;
    if (l.mode == HERO4) { for (int i = 0; i < 4; ++i) s.c[i] = BlackbodyNormalised(l.lambda[i], L.temperature); }
    else { s.c[0] = L.bbRGB[0]; s.c[1] = L.bbRGB[1]; s.c[2] = L.bbRGB[2]; s.c[3] = 0; }
    return s;
  }
  return FromRGB(L.rgb, l);
}
static Spec EnvLookup(const Light &L, const V2 &uv, const Lanes &l) {  // TextureAdapter::GetUV
  float c[3];
  L.env.Bilinear(Fract(uv.x), Fract(uv.y), c);
  return FromRGB(c, l);
}
static Spec EnvLe(const Light &L, const V3 &w, const Lanes &l) {  // EnvironmentLight.h:44-47
  const V2 uv(Fract((SphericalPhi(w) - L.offset.x) * INV_PI2), Fract((SphericalTheta(w) - L.offset.y) * INV_PI));
  return EnvLookup(L, uv, l) * L.intensity;
}
static float EnvPdfLi(const Light &L, const V3 &wi) {  // EnvironmentLight.cpp:60-67
  const float theta = SphericalTheta(wi) - L.offset.y;
  const float phi = SphericalPhi(wi) - L.offset.x;
  const float sinT = std::sin(theta);
  if (sinT == 0) return 0;
  return L.envDist.PDF(V2(Fract(phi * INV_PI2), Fract(theta * INV_PI))) / (2.f * PI * PI * sinT);
}
static inline V3 SamplePointInTriangle(const V3 &v0, const V3 &v1, const V3 &v2, V2 u, bool corrected) {
  if (corrected && u.x + u.y > 1.f) { u = V2(1.f - u.x, 1.f - u.y); }  // A36: parallelogram -> fold
  return v0 + (v1 - v0) * u.x + (v2 - v0) * u.y;  // TriangleMesh.h:73-79
}

// Light::Sample_Li for each light kind. Returns Li, sets pdf and ev.wi/wiL.
static Spec SampleLi(const Ctx &c, const Light &L, Event &ev, Sampler &smp, float &pdf) {
  const Scene &sc = *c.sc;
  Hit &h = *ev.hit;
  switch (L.kind) {
    case L_MESH:
    case L_TRI: {  // MeshLight.cpp:22-41, 134-151
      const Mesh &m = sc.meshes[L.mesh];
      const Light &ML = L.kind == L_TRI ? sc.lights[L.parentMeshLight] : L;
      uint32_t tri;
      if (L.kind == L_MESH) tri = L.triDist.SampleDiscrete(smp.Get1D(), &pdf);
      else tri = L.tri;
      const V2 u = smp.Get2D();
      const V3 v0 = m.p[m.idx[3 * tri]], v1 = m.p[m.idx[3 * tri + 1]], v2 = m.p[m.idx[3 * tri + 2]];
      const V3 pL = SamplePointInTriangle(v0, v1, v2, u, c.corrected);
      const bool atMedium = c.volumes && ev.mediumInteraction;
      const V3 pS = atMedium ? h.point : h.point + h.normalG * SURFACE_EPSILON * (float)ev.sidedness;
      Spec Tr(1.f);
      if (c.volumes ? VisibleTr(c, pS, pL, ev, &Tr, false) : MutualVisibility(sc, pS, pL, &ev.wi)) {
        float area;
        V3 normal;
        TriAreaNormal(v0, v1, v2, &area, &normal);
        const float dn = Dot(normal, -ev.wi);
        const float cosT = std::abs(dn);
        // A29 corrected: one-sided emission, consistent with PDF_Li
        if (cosT > 0 && !(c.corrected && dn < 0)) {
          ev.wiL = ev.ToLocal(ev.wi);
          if (L.kind == L_MESH) pdf *= DistSq(h.point, pL) / (cosT * area);
          else if (c.corrected) pdf = DistSq(h.point, pL) / (cosT * area);
          else pdf = DistSq(h.point, pL) * (cosT * area);  // A6
          const Material &mat = sc.materials[m.material];
          return Emission(c, mat, ev) * Tr * ML.intensity * INV_PI;
        }
      }
      pdf = 0;
      return Spec(0);
    }
    case L_POINT:
    case L_SPOT: {  // PointLight.cpp:18-30, Spotlight.cpp:29-42
      const bool atMedium = c.volumes && ev.mediumInteraction;
      const V3 p = atMedium ? h.point : h.point + h.normalG * SURFACE_EPSILON * (float)ev.sidedness;
      Spec Tr(1.f);
      if (c.volumes ? VisibleTr(c, p, L.position, ev, &Tr, true) : MutualVisibility(sc, p, L.position, &ev.wi)) {
        pdf = 1;
        const float d2 = DistSq(L.position, h.point);
        ev.wiL = ev.ToLocal(ev.wi);
        Spec e = LightColour(L, ev.lanes) * Tr;
        if (L.kind == L_SPOT) {
          const V3 w = -ev.wi;
          const float cosT = Dot(w, L.axis);  // Spotlight.cpp:21-27
          float fall;
          if (cosT < L.cosCone) fall = 0;
          else if (cosT > L.cosFalloff) fall = 1;
          else { const float d = (cosT - L.cosCone) / (L.cosFalloff - L.cosCone); fall = (d * d) * (d * d); }
          e = e * fall;
        }
        return e * L.intensity / d2;
      }
      pdf = 0;
      return Spec(0);
    }
    case L_ENV: {  // EnvironmentLight.cpp:24-45
      const V2 uv = L.envDist.SampleContinuous(smp.Get2D(), &pdf);
      if (pdf == 0) return Spec(0);
      const float theta = uv.y * PI + L.offset.y;
      const float phi = uv.x * PI2 + L.offset.x;
      const float cosT = std::cos(theta), sinT = std::sin(theta);
      if (sinT == 0) { pdf = 0; return Spec(0); }
      ev.wi = SphericalDirection(sinT, cosT, phi);
      const bool atMedium = c.volumes && ev.mediumInteraction;
      const V3 pe = atMedium ? h.point : h.point + h.normalG * SURFACE_EPSILON;
      Spec Tr(1.f);
      if (c.volumes ? EscapesTr(c, pe, ev.wi, ev, &Tr) : RayEscapes(sc, pe, ev.wi)) {
        ev.wiL = ev.ToLocal(ev.wi);
        pdf /= 2 * PI * PI * sinT;
        return EnvLookup(L, uv, ev.lanes) * Tr * L.intensity;
      }
      pdf = 0;
      return Spec(0);
    }
  }
  pdf = 0;
  return Spec(0);
}

// Light::PDF_Li(event) for a hit on an emitter (MeshLight.cpp:56-64, 157-164)
static float PdfLiHit(const Ctx &c, const Light &L, const Event &ev, const Hit &hit) {
  if (Dot(ev.wi, hit.normalG) > 0) return 0;  // one-sided
  const float distSq = hit.tFar * hit.tFar;
  V3 v0, v1, v2;
  TriVerts(*c.sc, hit.prim, v0, v1, v2);
  float area;
  TriAreaNormal(v0, v1, v2, &area, nullptr);
  const float cosT = std::abs(Dot(hit.normalG, -ev.wi));
  if (L.kind == L_MESH) {
    const float triPdf = c.corrected ? L.triDist.PDFNormalised(hit.local) : L.triDist.PDF(hit.local);  // A28
    return triPdf * distSq / (cosT * area);
  }
  return distSq / (cosT * area);
}

static float LightPower(const Scene &sc, const Light &L);

// ---------------------------------------------------------------- light tree build (ManyLightSampler.cpp:37-303)
static Bounds TLBounds(const Scene &sc, const Light &L) {
  if (L.kind == L_TRI) {
    const Mesh &m = sc.meshes[L.mesh];
    const V3 p0 = m.p[m.idx[3 * L.tri]], p1 = m.p[m.idx[3 * L.tri + 1]], p2 = m.p[m.idx[3 * L.tri + 2]];
    return Union(Union(Bounds(p0), p1), p2);
  }
  return Bounds(L.position);
}
static V3 TLDir(const Scene &sc, const Light &L) {
  if (L.kind == L_TRI) {
    const Mesh &m = sc.meshes[L.mesh];
    float area;
    V3 n;
    TriAreaNormal(m.p[m.idx[3 * L.tri]], m.p[m.idx[3 * L.tri + 1]], m.p[m.idx[3 * L.tri + 2]], &area, &n);
    return n;
  }
  if (L.kind == L_SPOT) return L.axis;
  return V3(0, 1, 0);
}
static Cone MakeCone(const V3 &axis) { Cone c; c.axis = axis; c.thetaO = 0; c.thetaE = PI / 2.f; return c; }
// OrientationCone::Union swaps its (const&) arguments in place when b.thetaO > a.thetaO (:51).
static Cone ConeUnion(Cone &a, Cone &b) {
  if (b.thetaO > a.thetaO) std::swap(a, b);
  const float thetaE = std::max(a.thetaE, b.thetaE);
  const float thetaD = std::acos(Clampf(Dot(a.axis, b.axis), 0.f, 1.f));
  if (std::min(thetaD + b.thetaO, PI) <= a.thetaO) { Cone r; r.axis = a.axis; r.thetaO = a.thetaO; r.thetaE = thetaE; return r; }
  const float thetaO = (a.thetaO + thetaD + b.thetaO) * .5f;
  if (thetaO >= PI) { Cone r; r.axis = a.axis; r.thetaO = PI; r.thetaE = thetaE; return r; }
  const float thetaR = thetaO - a.thetaO;
  Cone r;
  r.axis = Rotate(a.axis, Cross(a.axis, b.axis), thetaR);
  r.thetaO = thetaO;
  r.thetaE = thetaE;
  return r;
}
static float MOmega(const Cone &c) {  // :71-78
  const float cosO = std::cos(c.thetaO);
  const float sinO = std::sqrt(1 - cosO * cosO);
  const float a = PI2 * (1.f - cosO);
  const float thetaW2 = std::min(c.thetaO + c.thetaE, PI) * 2.f;
  const float b = thetaW2 * sinO - std::cos(c.thetaO - thetaW2) - 2.f * c.thetaO * sinO + cosO;
  return a + PI * .5f * b;
}
struct Bucket { float power = 0; Bounds b; Cone cone; };
static unsigned BucketOf(float x) {  // float -> unsigned (NaN / negative -> 0)
  if (!(x >= 0.f)) return 0;
  unsigned b = (unsigned)x;
  return b;
}

struct TreeBuilder {
  Scene &sc;
  std::vector<int> &order;  // tlight indices
  std::vector<float> powers;
  explicit TreeBuilder(Scene &s) : sc(s), order(s.tOrder) {}
  float P(unsigned i) { return powers[order[i]]; }
  Bounds B(unsigned i) { return TLBounds(sc, sc.tlights[order[i]]); }
  V3 D(unsigned i) { return TLDir(sc, sc.tlights[order[i]]); }

  float SAOH(LNode &P0, const Bucket &R, const Bucket &L, unsigned axis) {  // :83-91
    const V3 box = P0.bounds.mx - P0.bounds.mn;
    const float lenMax = std::max(std::max(box.x, box.y), box.z);
    const float Kr = lenMax / box[axis];
    const float m = P0.bounds.SurfaceArea() * MOmega(P0.cone);
    const float mL = L.power * L.b.SurfaceArea() * MOmega(L.cone);
    const float mR = R.power * R.b.SurfaceArea() * MOmega(R.cone);
    return Kr * (mL + mR) / m;
  }
  bool SplitAxis(LNode *n, unsigned axis, float *minCost, unsigned *bIdx, Bucket *lb, Bucket *rb) {  // :136-185
    const unsigned NB = 12;
    Bucket buckets[12];
    bool init[12] = {false};
    for (unsigned i = n->first; i < n->first + n->count; ++i) {
      unsigned b = BucketOf((float)NB * n->bounds.Offset(B(i).Center())[axis]);
      if (b == NB) b = NB - 1;
      if (b > NB) b = 0;
      if (!init[b]) {
        buckets[b].power = P(i);
        buckets[b].b = B(i);
        buckets[b].cone = MakeCone(D(i));
        init[b] = true;
      } else {
        buckets[b].power += P(i);
        buckets[b].b = Union(buckets[b].b, B(i));
        Cone tmp = MakeCone(D(i));
        buckets[b].cone = ConeUnion(buckets[b].cone, tmp);
      }
    }
    bool cheaper = false;
    for (unsigned i = 0; i < NB - 1; ++i) {
      Bucket bL, bR;
      bL.power = 0; bL.b = buckets[0].b; bL.cone = buckets[0].cone;
      bR.power = 0; bR.b = buckets[i + 1].b; bR.cone = buckets[i + 1].cone;
      for (unsigned j = 0; j <= i; ++j) {
        bL.b = Union(bL.b, buckets[j].b);
        bL.cone = ConeUnion(bL.cone, buckets[j].cone);
        bL.power += buckets[j].power;
      }
      for (unsigned j = i + 1; j < NB; ++j) {
        bR.b = Union(bR.b, buckets[j].b);
        bR.cone = ConeUnion(bR.cone, buckets[j].cone);
        bR.power += buckets[j].power;
      }
      const float cost = SAOH(*n, bR, bL, axis);
      if (cost < *minCost) { *minCost = cost; *lb = bL; *rb = bR; *bIdx = i; cheaper = true; }
    }
    return cheaper;
  }
  void InitLeaf(LNode *n) {  // :125-134
    std::vector<float> d(n->count);
    for (unsigned i = 0; i < n->count; ++i) {
      d[i] = P(n->first + i);
      sc.lightLeaf[order[n->first + i]] = {n, i};
      sc.treePower += d[i];
    }
    n->leafIndex = (int)sc.leafDist.size();
    sc.leafDist.emplace_back(d.data(), n->count);
    n->child[0] = n->child[1] = nullptr;
  }
  void Build(LNode *n) {  // RecursiveBuild :187-248
    n->bounds = B(n->first);
    n->cone = MakeCone(D(n->first));
    for (unsigned i = 0; i < n->count; ++i) {
      n->bounds = Union(n->bounds, B(i + n->first));
      Cone tmp = MakeCone(D(i + n->first));
      n->cone = ConeUnion(n->cone, tmp);
    }
    Bucket lb, rb;
    float minCost = INFINITY;
    unsigned bIdx = 0;
    int axis = -1;
    for (int a = 0; a < 3; ++a)
      if (SplitAxis(n, (unsigned)a, &minCost, &bIdx, &lb, &rb)) axis = a;
    if (n->count > 1 && minCost < n->power && axis >= 0) {
      const unsigned NB = 12;
      auto pred = [&](int li) {
        unsigned b = BucketOf(n->bounds.Offset(TLBounds(sc, sc.tlights[li]).Center())[axis] * (float)NB);
        if (b == NB) b--;
        if (b > NB) b = 0;
        return b <= bIdx;
      };
      auto it = std::stable_partition(order.begin() + n->first, order.begin() + n->first + n->count, pred);
      const unsigned pivot = (unsigned)(it - order.begin());
      const unsigned ln = pivot - n->first, rn = n->count - ln;
      if (ln != 0 && rn != 0) {
        LNode *l = new LNode(), *r = new LNode();
        sc.tnodes.emplace_back(l);
        sc.tnodes.emplace_back(r);
        l->parent = n; l->first = n->first; l->count = ln; l->bounds = lb.b; l->cone = lb.cone; l->power = lb.power;
        r->parent = n; r->first = pivot; r->count = rn; r->bounds = rb.b; r->cone = rb.cone; r->power = rb.power;
        n->child[0] = l;
        n->child[1] = r;
        Build(l);
        Build(r);
      } else InitLeaf(n);
    } else InitLeaf(n);
  }
};

static float TriLightIrradiance(const Ctx &c, const Light &TL) {  // TriangleLight::Irradiance, MeshLight.cpp:203-221
  const Scene &sc = *c.sc;
  const Light &ML = sc.lights[TL.parentMeshLight];
  const Mesh &m = sc.meshes[TL.mesh];
  const Material &mat = sc.materials[m.material];
  const unsigned S = 3;
  const float inv = 1.f / (float)S;
  Hit fh;
  Event fe;
  fe.hit = &fh;
  fe.lanes.mode = RGB;
  float irr = 0;
  for (unsigned a = 0; a < S; ++a)
    for (unsigned b = 0; b < S; ++b) {
      if (!m.uv.empty()) {
        const V2 &u0 = m.uv[m.idx[3 * TL.tri]], &u1 = m.uv[m.idx[3 * TL.tri + 1]], &u2 = m.uv[m.idx[3 * TL.tri + 2]];
        const float ua = (float)a * inv, ub = (float)b * inv;
        fh.uv = V2(u0.x + (u1.x - u0.x) * ua + (u2.x - u0.x) * ub, u0.y + (u1.y - u0.y) * ua + (u2.y - u0.y) * ub);
      }
      irr += Luminance(Emission(c, mat, fe), fe.lanes);
    }
  return irr * inv * inv * ML.intensity;
}
static float LightPower(const Scene &sc, const Light &L) {  // Light::Power = Irradiance * Area
  switch (L.kind) {
    case L_MESH: {
      const Mesh &m = sc.meshes[L.mesh];
      float a = 0;
      for (size_t i = 0; i < m.idx.size() / 3; ++i) a += Mag(Cross(m.p[m.idx[3 * i + 1]] - m.p[m.idx[3 * i]], m.p[m.idx[3 * i + 2]] - m.p[m.idx[3 * i]]));
      return L.intensity * (a * .5f);
    }
    case L_POINT: return L.intensity * (4 * PI);
    case L_SPOT: return L.intensity * (PI2 - (1 - .5f * (L.cosFalloff + L.cosCone)));
    case L_ENV: return L.intensity * (PI * L.radius * L.radius);
  }
  return 0;
}

// ---------------------------------------------------------------- light samplers
// ManyLightSampler::ImportanceMeasure (ManyLightSampler.cpp:305-326). corrected: the cone angle
// is taken from the emitter towards the point (the reference measures it against the direction
// towards the light, giving emitters that face the point zero importance, A47) and the receiver
// cosine is |cos thetaI| (it goes negative below the horizon as written, A48).
static float TreeImportance(const LNode *nd, const V3 &point, const V3 &normalS, bool corrected) {
  V3 delta = nd->bounds.Center() - point;
  const float diam = nd->bounds.Diagonal();
  const float d2 = Dot(delta, delta);
  const float d = std::sqrt(d2);
  const float invD = 1.f / d;
  delta = delta * invD;
  const float theta = std::acos(corrected ? -Dot(delta, nd->cone.axis) : Dot(delta, nd->cone.axis));
  const float thetaU = std::atan((diam * .5f) * invD);
  const float thetaDash = std::max(theta - nd->cone.thetaO - thetaU, 0.f);
  if (thetaDash < nd->cone.thetaE) {
    const float E = nd->power;
    // a zero normal marks a medium vertex (corrected mode): no surface-cosine bound
    if (normalS.x == 0.f && normalS.y == 0.f && normalS.z == 0.f) return E / d2 * std::cos(thetaDash);
    // a cosine that rounding pushed past 1 (normal pointing straight at the node centre) is clamped:
    // acos would give NaN and a NaN pick probability (the device clamps the same way)
    const float thetaI = std::acos(Clampf(corrected ? std::abs(Dot(delta, normalS)) : Dot(delta, normalS), -1.f, 1.f));
    return (std::cos(std::max(thetaI - thetaU, 0.f)) * E) / d2 * std::cos(thetaDash);
  }
  return 0.f;
}

struct LightPick { int kind; int index; };  // kind 0: scene light index, 1: tree light index, -1 none

static LightPick SampleLightAt(const Ctx &c, const V3 &refP, const V3 &refN, Sampler &smp, float *pdf) {
  const Scene &sc = *c.sc;
  if (sc.strategy == LGPU_LIGHTS_POWER) {  // LightSampler.cpp:14-16
    if (sc.lights.empty()) { *pdf = 0; return {-1, -1}; }
    const unsigned i = sc.powerDist.SampleDiscrete(smp.Get1D(), pdf);
    return {0, (int)i};
  }
  // ManyLightSampler::Sample (ManyLightSampler.cpp:17-27)
  float eps = smp.Get1D();
  const float pTree = sc.treePower / (sc.infPower + sc.treePower);
  if (eps < pTree) {
    *pdf *= pTree;
    float e = smp.Get1D();  // A5: the rescaled epsilon is discarded
    LNode *n = sc.root;
    while (!n->IsLeaf()) {  // PickLight :355-378
      const float IL = TreeImportance(n->child[0], refP, refN, c.corrected);
      const float IR = TreeImportance(n->child[1], refP, refN, c.corrected);
      const float PL = (IL + IR == 0.f) ? .5f : IL / (IL + IR);  // A10
      const float PR = 1 - PL;
      if (e < PL) { e /= PL; *pdf *= PL; n = n->child[0]; }
      else { e = (e - PL) / PR; *pdf *= PR; n = n->child[1]; }
    }
    float lp = 1;
    const unsigned i = sc.leafDist[n->leafIndex].SampleDiscrete(e, &lp);
    *pdf *= lp;
    return {1, sc.tOrder[n->first + i]};
  }
  if (sc.envLight < 0) { *pdf = 0; return {-1, -1}; }
  *pdf *= 1 - pTree;
  return {0, sc.envLight};
}
static LightPick SampleLight(const Ctx &c, const Event &ev, Sampler &smp, float *pdf) {
  return SampleLightAt(c, ev.hit->point, ev.hit->normalS, smp, pdf);
}

// LightSampler::Pdf for the light hit by a BSDF sample.
// ManyLightSampler::Pdf / RecursivePDF (ManyLightSampler.cpp:29-35, 380-393);
// PowerLightSampler::Pdf (LightSampler.cpp:18-20).
static float LightPdf(const Ctx &c, const LightPick &lp, const V3 &point, const V3 &normalS) {
  const Scene &sc = *c.sc;
  if (sc.strategy == LGPU_LIGHTS_POWER) return LightPower(sc, sc.lights[lp.index]) * sc.invTotalPower;
  if (lp.kind == 0 && lp.index == sc.envLight) return sc.infPower / (sc.infPower + sc.treePower);
  const std::pair<LNode *, unsigned> &ref = sc.lightLeaf[lp.index];
  const Piecewise1D &ld = sc.leafDist[ref.first->leafIndex];
  float p = c.corrected ? ld.PDFNormalised(ref.second) : ld.PDF(ref.second);  // A28
  const LNode *child = ref.first;
  for (const LNode *n = ref.first->parent; n; child = n, n = n->parent) {
    const float IL = TreeImportance(n->child[0], point, normalS, c.corrected);
    const float IR = TreeImportance(n->child[1], point, normalS, c.corrected);
    const float p0 = (IL + IR == 0.f) ? .5f : IL / (IL + IR);
    const float p1 = c.corrected ? 1.f - p0 : 1.f - IR;  // A7
    p *= (child == n->child[0]) ? p0 : p1;
  }
  if (c.corrected) p *= sc.treePower / (sc.infPower + sc.treePower);
  return p;
}

static inline float PowerHeuristic(float f, float g) { return (f * f) / (f * f + g * g); }  // Integrator.h:22-29

// ---------------------------------------------------------------- integrators
struct RenderParams {
  uint32_t integrator, maxBounces, minBounces, mode, seed;
  bool corrected;
  float maxDepth = 1, power = 1, background[3] = {1, 0, 1};  // DepthPass
  bool normalG = false;                                      // NormalPass
};

static const Light &LightOf(const Scene &sc, const LightPick &lp) { return lp.kind == 1 ? sc.tlights[lp.index] : sc.lights[lp.index]; }
static bool SameLight(const LightPick &a, const LightPick &b) { return a.kind == b.kind && a.index == b.index; }

// PathIntegrator::Li (PathIntegrator.cpp:15-88)
static Spec PathLi(const Ctx &c, const RenderParams &rp, V3 ro, V3 rd, Sampler &smp, const Lanes &lanes) {
  const Scene &sc = *c.sc;
  Spec L(0), beta(1);
  Hit hit;
  Event ev;
  ev.hit = &hit;
  ev.lanes = lanes;
  ev.wo = -rd;
  bool scatterIntersect = false;
  int passThrough = 0;
  for (int bounces = 0; bounces < (int)rp.maxBounces; ++bounces) {
    if (bounces == 0 ? Intersect(sc, ro, rd, hit, c.corrected) : scatterIntersect) {
      const Material &mat = sc.materials[sc.meshes[hit.mesh].material];
      const bool isLight = mat.emission != LGPU_NONE;
      if (bounces == 0 && isLight) {
        const Mesh &m = sc.meshes[hit.mesh];
        if (!c.corrected || Dot(rd, hit.normalG) < 0)  // A29 corrected: front face only
          L += Emission(c, mat, ev) * sc.lights[m.meshLight].intensity * INV_PI;
      }
      if (mat.bxdf != LGPU_NONE) {
        ev.wo = -rd;
        ev.SurfaceLocalise();
        float lightDistPdf = 1;
        const V3 shadePoint = hit.point, shadeNormalS = hit.normalS;
        const LightPick l = SampleLight(c, ev, smp, &lightDistPdf);
        Spec Ld(0);
        float scatteringPDF = 0, lightPDF = 0;
        Spec f, Li(0);
        bool deltaLight = false;
        if (l.kind >= 0) {
          const Light &LL = LightOf(sc, l);
          deltaLight = LL.kind == L_POINT || LL.kind == L_SPOT;
          Li = SampleLi(c, LL, ev, smp, lightPDF);
        }
        lightPDF *= lightDistPdf;
        if (lightPDF > 0 && !IsBlack(Li, lanes)) {
          f = BxDF_f(c, mat, mat.bxdf, ev) * std::abs(ev.wiL.y);
          scatteringPDF = BxDF_Pdf(c, mat, mat.bxdf, ev.woL, ev.wiL, ev);
          if (!IsBlack(f, lanes) && scatteringPDF > 0) {
            const float w = (c.corrected && deltaLight) ? 1.f : PowerHeuristic(lightPDF, scatteringPDF);  // A38
            Ld += Li * f * w / lightPDF;
          }
        }
        bool specular;
        f = BxDF_Sample(c, mat, mat.bxdf, ev, smp, scatteringPDF, &specular);
        f = f * std::abs(ev.wiL.y);
        ro = hit.point;
        rd = ev.wi;
        scatterIntersect = Intersect(sc, ro, rd, hit, c.corrected);
        if (scatteringPDF > 0 && !IsBlack(f, lanes)) {
          LightPick nl{-1, -1};
          if (scatterIntersect) {
            const Mesh &hm = sc.meshes[hit.mesh];
            if (hm.meshLight >= 0) {
              if (sc.strategy == LGPU_LIGHTS_TREE) nl = {1, sc.triToTLight[hit.prim]};
              else nl = {0, hm.meshLight};
            }
          } else if (sc.envLight >= 0) nl = {0, sc.envLight};
          if (nl.kind >= 0) {
            // nl != l: the reference compares Light pointers; for the tree a
            // MeshLight never equals the sampled TriangleLight (:55).
            bool same = SameLight(nl, l);
            if (!c.corrected && sc.strategy == LGPU_LIGHTS_TREE && nl.kind == 1) same = false;
            if (!same) {
              // A8: faithful evaluates the tree at the hit on the light
              if (c.corrected) lightDistPdf = LightPdf(c, nl, shadePoint, shadeNormalS);
              else lightDistPdf = LightPdf(c, nl, hit.point, hit.normalS);
            }
            const Light &NL = LightOf(sc, nl);
            float pl;
            Spec Le;
            if (scatterIntersect) {
              // faithful: MeshLight::PDF_Li (mesh triangle distribution, A28) even under the tree
              if (!c.corrected && sc.strategy == LGPU_LIGHTS_TREE) pl = PdfLiHit(c, sc.lights[NL.parentMeshLight], ev, hit);
              else pl = PdfLiHit(c, NL, ev, hit);
              const Material &hmat = sc.materials[sc.meshes[hit.mesh].material];
              const int mlIdx = sc.meshes[hit.mesh].meshLight;
              Le = Emission(c, hmat, ev) * sc.lights[mlIdx].intensity * INV_PI;
              if (c.corrected && Dot(rd, hit.normalG) > 0) Le = Spec(0);
            } else {
              pl = EnvPdfLi(NL, ev.wi);
              Le = EnvLe(NL, rd, lanes);
            }
            lightPDF = lightDistPdf * pl;
            const float w = (c.corrected && specular) ? 1.f : PowerHeuristic(scatteringPDF, lightPDF);  // A40
            if (!IsBlack(Le, lanes) && (lightPDF > 0 || c.corrected)) Ld += Le * f * w / scatteringPDF;
          }
        } else {
          if (c.corrected) L += beta * Ld;  // A37: NEE kept when the BSDF sample fails
          break;
        }
        L += beta * Ld;
        beta = beta * (f / scatteringPDF);
      } else {
        // pass-through surface (PathIntegrator.cpp:69-74)
        ro = hit.point + hit.normalG * (Dot(rd, hit.normalG) < 0 ? -SURFACE_EPSILON : SURFACE_EPSILON);
        if (bounces > 0) scatterIntersect = Intersect(sc, ro, rd, hit, c.corrected);
        bounces--;
        if (++passThrough > 1024) break;  // A27 cap
        continue;
      }
      if (bounces > (int)rp.minBounces) {  // :76-80
        const float q = std::max(.05f, 1 - Luminance(beta, lanes));
        if (smp.Get1D() < q) break;
        beta = beta / Spec(1.f - q);
      }
    } else {
      if (bounces == 0 && sc.envLight >= 0) L += EnvLe(sc.lights[sc.envLight], rd, lanes);
      break;
    }
  }
  return L;
}

// ---------------------------------------------------------------- volumetric integrators
// Light reached by a BSDF- or phase-sampled ray (the MIS half of PathIntegrator.cpp:52-63,
// VolumetricPathIntegrator.cpp:127-139, MISVolumetricPathIntegrator.cpp:52-66). `h` is that
// ray's hit (si: it hit something). Returns Le * w; the caller applies f (or Tr) and 1/pdf.
// envNoLightPdf: the light technique cannot produce the environment (SamplePoint, corrected).
static Spec MisLight(const Ctx &c, const LightPick &l, float lightDistPdf, const V3 &refP, const V3 &refN, bool si,
                     Hit &h, Event &ev, const V3 &rd, float scatPdf, bool specular, bool envNoLightPdf) {
  const Scene &sc = *c.sc;
  LightPick nl{-1, -1};
  if (si) {
    const Mesh &hm = sc.meshes[h.mesh];
    if (hm.meshLight >= 0) nl = sc.strategy == LGPU_LIGHTS_TREE ? LightPick{1, sc.triToTLight[h.prim]} : LightPick{0, hm.meshLight};
  } else if (sc.envLight >= 0) {
    nl = {0, sc.envLight};
  }
  if (nl.kind < 0) return Spec(0);
  Hit *saved = ev.hit;
  ev.hit = &h;
  bool same = SameLight(nl, l);
  if (!c.corrected && sc.strategy == LGPU_LIGHTS_TREE && nl.kind == 1) same = false;  // MeshLight != TriangleLight
  float ldp = lightDistPdf;
  if (!same) ldp = c.corrected ? LightPdf(c, nl, refP, refN) : LightPdf(c, nl, h.point, h.normalS);  // A8
  const Light &NL = LightOf(sc, nl);
  float pl;
  Spec Le;
  if (si) {
    if (!c.corrected && sc.strategy == LGPU_LIGHTS_TREE) pl = PdfLiHit(c, sc.lights[NL.parentMeshLight], ev, h);
    else pl = PdfLiHit(c, NL, ev, h);
    const Material &hmat = sc.materials[sc.meshes[h.mesh].material];
    Le = Emission(c, hmat, ev) * sc.lights[sc.meshes[h.mesh].meshLight].intensity * INV_PI;
    if (c.corrected && Dot(rd, h.normalG) > 0) Le = Spec(0);
  } else {
    pl = envNoLightPdf ? 0.f : EnvPdfLi(NL, ev.wi);
    Le = EnvLe(NL, rd, ev.lanes);
  }
  ev.hit = saved;
  const float lightPDF = ldp * pl;
  const float w = (c.corrected && specular) ? 1.f : PowerHeuristic(scatPdf, lightPDF);
  if (!IsBlack(Le, ev.lanes) && (lightPDF > 0 || c.corrected)) return Le * w;
  return Spec(0);
}

// PartialLightSample (Light.h:107-113). pdf starts at 1: the reference leaves it uninitialised
// before ManyLightSampler::Sample multiplies into it (MISVolumetricPathIntegrator.cpp:95-96).
struct PartialLS {
  V3 point, normal;
  V2 uv;
  Spec Le;
  LightPick light{-1, -1};
  float pdf = 1, selPdf = 1;
};
// Light::SamplePoint (MeshLight.cpp:66-81, 166-180; PointLight.cpp:40-44; Spotlight.cpp:50-54;
// EnvironmentLight.cpp:69-74). Faithful writes the sampled uv into the event's hit as written.
static Spec SamplePointL(const Ctx &c, const Light &L, Event &ev, Sampler &smp, PartialLS &ls) {
  const Scene &sc = *c.sc;
  switch (L.kind) {
    case L_MESH:
    case L_TRI: {
      const Mesh &m = sc.meshes[L.mesh];
      const Light &ML = L.kind == L_TRI ? sc.lights[L.parentMeshLight] : L;
      uint32_t tri = L.tri;
      float area;
      if (L.kind == L_MESH) {
        float distPdf;
        tri = L.triDist.SampleDiscrete(smp.Get1D(), &distPdf);
        TriAreaNormal(m.p[m.idx[3 * tri]], m.p[m.idx[3 * tri + 1]], m.p[m.idx[3 * tri + 2]], &area, &ls.normal);
        ls.pdf *= distPdf / area;
      } else {
        TriAreaNormal(m.p[m.idx[3 * tri]], m.p[m.idx[3 * tri + 1]], m.p[m.idx[3 * tri + 2]], &area, &ls.normal);
        ls.pdf /= area;
      }
      const V2 u = smp.Get2D();
      const uint32_t i0 = m.idx[3 * tri], i1 = m.idx[3 * tri + 1], i2 = m.idx[3 * tri + 2];
      V2 uv(0, 0);
      if (!m.uv.empty()) {  // maths::BarycentricInterpolation (maths.h:58-60)
        uv.x = m.uv[i0].x + (m.uv[i1].x - m.uv[i0].x) * u.x + (m.uv[i2].x - m.uv[i0].x) * u.y;
        uv.y = m.uv[i0].y + (m.uv[i1].y - m.uv[i0].y) * u.x + (m.uv[i2].y - m.uv[i0].y) * u.y;
      }
      ls.uv = uv;
      ls.point = SamplePointInTriangle(m.p[i0], m.p[i1], m.p[i2], u, c.corrected);
      Hit tmp;
      Hit *saved = ev.hit;
      if (c.corrected) { tmp.uv = uv; ev.hit = &tmp; }
      else ev.hit->uv = uv;
      const Spec Le = Emission(c, sc.materials[m.material], ev) * ML.intensity * INV_PI;
      ev.hit = saved;
      return Le;
    }
    case L_POINT:
    case L_SPOT:
      ls.point = L.position;
      return LightColour(L, ev.lanes) * L.intensity;
    case L_ENV:
      ls.pdf *= 0;
      ls.point = V3(0, 0, 0);
      return Spec(0);
  }
  return Spec(0);
}
// Light::Visibility (MeshLight.cpp:83-92, 182-191; PointLight.cpp:46-55; Spotlight.cpp:56-65;
// EnvironmentLight.cpp:76-84): completes the pdf in ls and sets ev.wi. `eh` is the event's hit.
static Spec VisibilityL(const Ctx &c, const Light &L, const V3 &sp, Event &ev, Hit &eh, PartialLS &ls) {
  const Scene &sc = *c.sc;
  Spec Tr(1.f);
  switch (L.kind) {
    case L_MESH:
    case L_TRI: {
      if (c.volumes ? VisibleTr(c, sp, ls.point, ev, &Tr, false) : MutualVisibility(sc, sp, ls.point, &ev.wi)) {
        // TriangleLight uses the event hit's normal as written; corrected uses the light's
        const V3 n = (L.kind == L_TRI && !c.corrected) ? eh.normalG : ls.normal;
        const float dn = Dot(n, -ev.wi);
        const float cosT = std::abs(dn);
        if (c.corrected && !(dn > 0)) break;  // A29: one-sided emission
        ls.pdf *= DistSq(sp, ls.point) / cosT;
        return Tr;
      }
      break;
    }
    case L_POINT:
    case L_SPOT: {
      const V3 p = c.corrected ? sp : sp + eh.normalG * SURFACE_EPSILON * (float)ev.sidedness;
      if (c.volumes ? VisibleTr(c, p, ls.point, ev, &Tr, true) : MutualVisibility(sc, p, ls.point, &ev.wi)) {
        const float d2 = DistSq(L.position, eh.point);
        if (L.kind == L_SPOT) {
          const float cosT = Dot(-ev.wi, L.axis);
          float fall;
          if (cosT < L.cosCone) fall = 0;
          else if (cosT > L.cosFalloff) fall = 1;
          else { const float d = (cosT - L.cosCone) / (L.cosFalloff - L.cosCone); fall = (d * d) * (d * d); }
          return Tr * fall / d2;
        }
        return Tr / d2;
      }
      break;
    }
    case L_ENV: {
      const V3 wi = Norm(ls.point - sp);
      if (c.volumes ? EscapesTr(c, sp, wi, ev, &Tr) : RayEscapes(sc, sp, wi)) return Tr;
      break;
    }
  }
  ls.pdf = 0;
  return Spec(0);
}

// MISVolumetricPathIntegrator::LdMediumPoint (MISVolumetricPathIntegrator.cpp:23-69).
// faithful: `eh` is the path's hit (the reference's event.hit, overwritten here) and ls carries
// the pdf compounded by earlier calls; corrected works on copies.
static Spec LdMediumPoint(const Ctx &c, const MediumO &M, Event &ev, Hit &eh, const V3 &p, PartialLS &ls,
                          const V3 &refP, const V3 &refN, V3 *wiOut) {
  const Scene &sc = *c.sc;
  float partialLightPDF = c.corrected ? ls.selPdf : ls.pdf;
  eh.point = p;
  Hit *saved = ev.hit;
  ev.hit = &eh;
  Spec Li(0);
  bool delta = false;
  if (ls.light.kind >= 0) {
    const Light &LL = LightOf(sc, ls.light);
    delta = LL.kind == L_POINT || LL.kind == L_SPOT;
    Li = ls.Le * VisibilityL(c, LL, p, ev, eh, ls);
  } else {
    ls.pdf = 0;
  }
  float lightPDF = ls.pdf;
  Spec Ld(0);
  if (lightPDF > 0 && !IsBlack(Li, ev.lanes)) {
    const float ph = PhaseHG(Dot(ev.wo, ev.wi), M.g);
    if (ph > 0) {
      const float w = (c.corrected && delta) ? 1.f : PowerHeuristic(lightPDF, ph);
      Ld += Li * ph * w / lightPDF;
    }
  }
  V3 wi;
  const float scatPdf = HGSample(M.g, ev.wo, &wi, *ev.sampler, c.corrected);
  ev.wi = wi;
  *wiOut = wi;
  Spec Tr(1.f);
  const bool si = IntersectTr(sc, p, wi, eh, ev.medium, &Tr, ev.lanes, -1.f, c.corrected, nullptr);
  if (scatPdf > 0) {
    const Spec C = MisLight(c, ls.light, partialLightPDF, refP, refN, si, eh, ev, wi, scatPdf, false, c.corrected);
    Ld += c.corrected ? C * Tr : C * Tr / scatPdf;  // A43: the phase value is missing as written
  }
  ev.hit = saved;
  return Ld;
}

// VolumetricPathIntegrator::Li (VolumetricPathIntegrator.cpp:20-162) and
// MISVolumetricPathIntegrator::Li (MISVolumetricPathIntegrator.cpp:71-215).
static Spec VolLi(const Ctx &c, const RenderParams &rp, V3 ro, V3 rd, Sampler &smp, const Lanes &lanes, bool misVol) {
  const Scene &sc = *c.sc;
  Spec L(0), beta(1);
  Hit hit;
  Event ev;
  ev.hit = &hit;
  ev.lanes = lanes;
  ev.sampler = &smp;
  ev.wo = -rd;
  ev.medium = -1;  // InMedium returns nullptr (:15-18)
  bool scatterIntersect = false;
  int passThrough = 0;
  for (int bounces = 0; bounces < (int)rp.maxBounces; ++bounces) {
    if (!(bounces == 0 ? Intersect(sc, ro, rd, hit, c.corrected) : scatterIntersect)) {
      if (bounces == 0 && sc.envLight >= 0) L += EnvLe(sc.lights[sc.envLight], rd, lanes) * beta;
      break;
    }
    const Material &mat = sc.materials[sc.meshes[hit.mesh].material];
    if (bounces == 0 && mat.emission != LGPU_NONE && (!c.corrected || Dot(rd, hit.normalG) < 0)) {
      Spec Le = Emission(c, mat, ev) * sc.lights[sc.meshes[hit.mesh].meshLight].intensity * INV_PI;
      if (ev.medium >= 0) Le = Le * MediumTr(sc.media[ev.medium], ro, rd, hit.tFar, lanes);
      L += c.corrected ? beta * Le : Le;  // beta is taken as 1 as written
    }
    ev.mediumInteraction = false;
    if (ev.medium >= 0) {
      const MediumO &M = sc.media[ev.medium];
      if (!misVol) {
        bool inter;
        float tD, pdfD;
        beta = beta * SampleDistance(M, ro, rd, hit.tFar, smp, lanes, &inter, &tD, &pdfD);
        ev.mediumInteraction = inter;
        if (inter) {  // :99-148
          ev.wo = -rd;
          const V3 sp = ro + rd * tD;
          hit.point = sp;
          const V3 refP = sp, refN = c.corrected ? V3(0, 0, 0) : hit.normalS;
          float lightDistPdf = 1;
          const LightPick l = SampleLightAt(c, refP, refN, smp, &lightDistPdf);
          Spec Ld(0), Li(0);
          float lightPDF = 0;
          bool delta = false;
          if (l.kind >= 0) {
            const Light &LL = LightOf(sc, l);
            delta = LL.kind == L_POINT || LL.kind == L_SPOT;
            Li = SampleLi(c, LL, ev, smp, lightPDF);
          }
          lightPDF *= lightDistPdf;
          if (lightPDF > 0 && !IsBlack(Li, lanes)) {
            const float ph = PhaseHG(Dot(ev.wo, ev.wi), M.g);
            if (ph > 0) {
              const float w = (c.corrected && delta) ? 1.f : PowerHeuristic(lightPDF, ph);
              Ld += Li * ph * w / lightPDF;
            }
          }
          V3 wi;
          const float scatPdf = HGSample(M.g, ev.wo, &wi, smp, c.corrected);
          ev.wi = wi;
          Spec Tr(1.f);
          const bool si = IntersectTr(sc, sp, wi, hit, ev.medium, &Tr, lanes, -1.f, c.corrected, nullptr);
          if (scatPdf > 0) {
            const Spec C = MisLight(c, l, lightDistPdf, refP, refN, si, hit, ev, wi, scatPdf, false, false);
            Ld += c.corrected ? C * Tr : C * Tr / scatPdf;  // A43
          } else {
            if (c.corrected) L += beta * Ld;
            break;
          }
          ro = sp;
          rd = wi;
          scatterIntersect = Intersect(sc, ro, rd, hit, c.corrected);
          L += beta * Ld;
        }
      } else {  // :92-147
        PartialLS ls;
        const V3 refP = c.corrected ? ro : hit.point, refN = c.corrected ? V3(0, 0, 0) : hit.normalS;
        ls.light = SampleLightAt(c, refP, refN, smp, &ls.pdf);
        ls.selPdf = ls.pdf;
        if (ls.light.kind >= 0) ls.Le = SamplePointL(c, LightOf(sc, ls.light), ev, smp, ls);
        else { ls.pdf = 0; ls.Le = Spec(0); }
        const int choice = smp.Get1D() > .5f ? 0 : 1;  // 0 distance, 1 equiangular
        Spec mTr[2];
        bool inter[2];
        float tS[2], pdfS[2];
        mTr[0] = SampleDistance(M, ro, rd, hit.tFar, smp, lanes, &inter[0], &tS[0], &pdfS[0]);
        mTr[1] = SampleEquiangular(M, ro, rd, hit.point, hit.tFar, ls.point, smp, lanes, c.corrected, &inter[1], &tS[1], &pdfS[1]);
        if (!c.corrected) {
          ev.mediumInteraction = inter[0] && inter[1];
          if (!ev.mediumInteraction) beta = beta * mTr[0];
          else {
            ev.wo = -rd;
            const V3 mp0 = ro + rd * tS[0], mp1 = ro + rd * tS[1];
            const float pdfD2 = PdfDistance(M, ro, rd, tS[1], lanes, false);
            const float pdfE2 = PdfEquiangular(ro, rd, ls.point, hit.tFar, tS[0], false);
            const float wD = PowerHeuristic(pdfS[0], pdfE2), wE = PowerHeuristic(pdfS[1], pdfD2);
            V3 wi0, wi1;
            Spec Ld(0);
            Ld += LdMediumPoint(c, M, ev, hit, mp0, ls, refP, refN, &wi0) * mTr[0] * wD;
            Ld += LdMediumPoint(c, M, ev, hit, mp1, ls, refP, refN, &wi1) * mTr[1] * wE;
            ro = choice == 0 ? mp0 : mp1;
            rd = choice == 0 ? wi0 : wi1;
            scatterIntersect = Intersect(sc, ro, rd, hit, c.corrected);
            L += Ld;  // A15: beta missing as written
            beta = beta * mTr[choice];
          }
        } else {
          // both points estimate direct light unconditionally (power-heuristic MIS over the two
          // distance techniques); the path continues with the distance sample
          ev.wo = -rd;
          ev.mediumInteraction = true;  // NEE rays start in this medium
          const V3 mp0 = ro + rd * tS[0], mp1 = ro + rd * tS[1];
          Spec Ld(0);
          V3 wi0(0, 0, 0), wi1;
          if (inter[0]) {
            PartialLS l0 = ls;
            Hit eh = hit;
            const float wD = PowerHeuristic(pdfS[0], PdfEquiangular(ro, rd, ls.point, hit.tFar, tS[0], true));
            Ld += LdMediumPoint(c, M, ev, eh, mp0, l0, refP, refN, &wi0) * mTr[0] * wD;
          }
          if (inter[1]) {
            PartialLS l1 = ls;
            Hit eh = hit;
            const float wE = PowerHeuristic(pdfS[1], PdfDistance(M, ro, rd, tS[1], lanes, true));
            Ld += LdMediumPoint(c, M, ev, eh, mp1, l1, refP, refN, &wi1) * mTr[1] * wE;
          }
          L += beta * Ld;
          beta = beta * mTr[0];
          ev.mediumInteraction = inter[0];
          if (inter[0]) {
            ro = mp0;
            rd = wi0;
            scatterIntersect = Intersect(sc, ro, rd, hit, c.corrected);
          }
        }
        (void)choice;
      }
    }
    if (!ev.mediumInteraction) {
      if (mat.bxdf != LGPU_NONE) {  // :45-90 / :150-193
        ev.wo = -rd;
        ev.SurfaceLocalise();
        const V3 shadeP = hit.point, shadeN = hit.normalS;
        float lightDistPdf = 1;
        const LightPick l = SampleLight(c, ev, smp, &lightDistPdf);
        Spec Ld(0), Li(0);
        float lightPDF = 0, scatteringPDF = 0;
        bool delta = false;
        if (l.kind >= 0) {
          const Light &LL = LightOf(sc, l);
          delta = LL.kind == L_POINT || LL.kind == L_SPOT;
          Li = SampleLi(c, LL, ev, smp, lightPDF);
        }
        lightPDF *= lightDistPdf;
        if (lightPDF > 0 && !IsBlack(Li, lanes)) {
          const Spec f = BxDF_f(c, mat, mat.bxdf, ev) * std::abs(ev.wiL.y);
          scatteringPDF = BxDF_Pdf(c, mat, mat.bxdf, ev.woL, ev.wiL, ev);
          if (!IsBlack(f, lanes) && scatteringPDF > 0) {
            const float w = (c.corrected && delta) ? 1.f : PowerHeuristic(lightPDF, scatteringPDF);
            Ld += Li * f * w / lightPDF;
          }
        }
        bool specular;
        Spec f = BxDF_Sample(c, mat, mat.bxdf, ev, smp, scatteringPDF, &specular);
        f = f * std::abs(ev.wiL.y);
        ro = hit.point;
        rd = ev.wi;
        ev.medium = GetMedium(mat, ev.wi, hit.normalG);
        scatterIntersect = Intersect(sc, ro, rd, hit, c.corrected);
        if (scatteringPDF > 0 && !IsBlack(f, lanes)) {
          if (c.volumes) {  // corrected: the light behind pass-through surfaces, with Tr
            Hit th;
            Spec Tr(1.f);
            const bool si = IntersectTr(sc, ro, rd, th, ev.medium, &Tr, lanes, -1.f, true, nullptr);
            Ld += MisLight(c, l, lightDistPdf, shadeP, shadeN, si, th, ev, rd, scatteringPDF, specular, false) * f * Tr /
                  scatteringPDF;
          } else {
            Ld += MisLight(c, l, lightDistPdf, shadeP, shadeN, scatterIntersect, hit, ev, rd, scatteringPDF, specular, false) *
                  f / scatteringPDF;
          }
        } else {
          if (c.corrected) L += beta * Ld;
          break;
        }
        L += beta * Ld;
        beta = beta * (f / scatteringPDF);
      } else {  // pass-through (:91-97)
        ev.medium = GetMedium(mat, rd, hit.normalG);
        ro = hit.point + hit.normalG * (Dot(hit.normalG, rd) < 0 ? -SURFACE_EPSILON : SURFACE_EPSILON);
        if (bounces > 0) scatterIntersect = Intersect(sc, ro, rd, hit, c.corrected);
        bounces--;
        if (++passThrough > 1024) break;  // A27 cap
        continue;
      }
    }
    if (bounces > (int)rp.minBounces) {
      const float q = std::max(.05f, 1 - Luminance(beta, lanes));
      if (smp.Get1D() < q) break;
      beta = beta / Spec(1.f - q);
    }
  }
  return L;
}

// ---------------------------------------------------------------- direct lighting
// Light::PDF_Li(event, sampler) for the light chosen by SampleOneLight (MeshLight.cpp:44-55,
// :153-155, EnvironmentLight.cpp:44-58, PointLight/Spotlight: 0). Faithful fires the
// reference's own rays; corrected uses the MIS ray's hit (see DirectLi).
static float PdfLiRay(const Ctx &c, const LightPick &lp, const Event &ev) {
  const Scene &sc = *c.sc;
  const Light &L = LightOf(sc, lp);
  const Hit &h = *ev.hit;
  if (L.kind == L_POINT || L.kind == L_SPOT) return 0.f;
  if (L.kind == L_ENV) {
    const float theta = SphericalTheta(ev.wi) - L.offset.y;
    const float phi = SphericalPhi(ev.wi) - L.offset.x;
    const float sinT = std::sin(theta);
    if (sinT == 0) return 0.f;
    const V3 wiOffset = SphericalDirection(sinT, std::cos(theta), phi);
    if (!RayEscapes(sc, h.point, wiOffset)) return 0.f;
    return L.envDist.PDF(V2(Fract(phi * INV_PI2), Fract(theta * INV_PI))) / (2.f * PI * PI * sinT);
  }
  // MeshLight::PDF_Li(event, sampler); a TriangleLight forwards to its mesh light
  const Light &ML = L.kind == L_TRI ? sc.lights[L.parentMeshLight] : L;
  Hit lh;
  const V3 o = h.point + h.normalG * SURFACE_EPSILON * (float)ev.sidedness;
  if (!Intersect(sc, o, ev.wi, lh, c.corrected)) return 0.f;
  if (sc.meshes[lh.mesh].meshLight != (int)(&ML - &sc.lights[0])) return 0.f;
  V3 v0, v1, v2;
  TriVerts(sc, lh.prim, v0, v1, v2);
  float area;
  V3 normal;
  TriAreaNormal(v0, v1, v2, &area, &normal);
  const float cosT = std::abs(Dot(normal, -ev.wi));
  if (cosT > 0) return DistSq(h.point, lh.point) / (cosT * area);
  return 0.f;
}

// Integrator::SampleOneLight / EstimateDirect (Integrator.cpp:5-55) and
// DirectLightingIntegrator::Li (DirectLightingIntegrator.cpp:13-29). Corrected adds the
// emission of the first surface (front face), weighs delta lights and specular samples 1,
// offsets the MIS ray to the side of wi, and takes the MIS pdf and emission from that ray's
// own hit on the sampled light (triangle-selection pdf included).
static Spec DirectLi(const Ctx &c, V3 ro, V3 rd, Sampler &smp, const Lanes &lanes) {
  const Scene &sc = *c.sc;
  Hit hit;
  Event ev;
  ev.hit = &hit;
  ev.lanes = lanes;
  ev.sampler = &smp;
  Spec L(0);
  for (int pass = 0; pass <= 1024; ++pass) {
    if (!Intersect(sc, ro, rd, hit, c.corrected)) {
      if (sc.envLight >= 0) L += EnvLe(sc.lights[sc.envLight], rd, lanes);  // A3: no env light gives 0
      return L;
    }
    const Material &mat = sc.materials[sc.meshes[hit.mesh].material];
    if (c.corrected && mat.emission != LGPU_NONE && Dot(rd, hit.normalG) < 0)
      L += Emission(c, mat, ev) * sc.lights[sc.meshes[hit.mesh].meshLight].intensity * INV_PI;
    if (mat.bxdf == LGPU_NONE) {  // :24-26
      ro = hit.point + rd * .0001f;
      continue;
    }
    ev.wo = -rd;
    ev.SurfaceLocalise();
    float lightPdf = 1;
    const LightPick l = SampleLight(c, ev, smp, &lightPdf);
    if (l.kind < 0) return L;
    const Light &LL = LightOf(sc, l);
    const bool delta = LL.kind == L_POINT || LL.kind == L_SPOT;
    Spec Ld(0);
    float lightPDF = 0, scatteringPDF = 0;
    Spec Li = SampleLi(c, LL, ev, smp, lightPDF);
    if (lightPDF > 0 && !IsBlack(Li, lanes)) {
      const Spec f = BxDF_f(c, mat, mat.bxdf, ev) * std::abs(ev.wiL.y);
      scatteringPDF = BxDF_Pdf(c, mat, mat.bxdf, ev.woL, ev.wiL, ev);
      if (!IsBlack(f, lanes) && scatteringPDF > 0) {
        const float w = (c.corrected && delta) ? 1.f : PowerHeuristic(lightPDF, scatteringPDF);
        Ld += Li * f * w / lightPDF;
      }
    }
    bool specular;
    Spec f = BxDF_Sample(c, mat, mat.bxdf, ev, smp, scatteringPDF, &specular);
    f = f * std::abs(ev.wiL.y);
    if (scatteringPDF > 0 && !IsBlack(f, lanes)) {
      Spec Lb(0);
      float weight = 1;
      if (!c.corrected) {
        const float lp = PdfLiRay(c, l, ev);
        if (lp == 0) return L + Ld / lightPdf;
        weight = PowerHeuristic(scatteringPDF, lp);
        Hit lh;
        const V3 o = hit.point + hit.normalG * .00001f;
        if (Intersect(sc, o, ev.wi, lh, c.corrected)) {
          // material->light == &_light: a mesh's light against the sampled light (a TriangleLight
          // under the tree never equals it)
          const int ml = sc.meshes[lh.mesh].meshLight;
          if (ml >= 0 && l.kind == 0 && ml == l.index) {
            Hit *saved = ev.hit;
            ev.hit = &lh;
            Lb = Emission(c, sc.materials[sc.meshes[lh.mesh].material], ev) * sc.lights[ml].intensity * INV_PI;
            ev.hit = saved;
          }
        } else if (LL.kind == L_ENV) {
          Lb = EnvLe(LL, ev.wi, lanes);
        }
      } else if (!delta) {
        Hit lh;
        const V3 o = hit.point + hit.normalG * (Dot(ev.wi, hit.normalG) < 0 ? -SURFACE_EPSILON : SURFACE_EPSILON);
        float lp = 0;
        if (Intersect(sc, o, ev.wi, lh, true)) {
          const int ml = sc.meshes[lh.mesh].meshLight;
          const bool same = ml >= 0 && (l.kind == 1 ? sc.triToTLight[lh.prim] == l.index : ml == l.index);
          if (same && Dot(ev.wi, lh.normalG) < 0) {
            lp = PdfLiHit(c, LL, ev, lh);
            Hit *saved = ev.hit;
            ev.hit = &lh;
            Lb = Emission(c, sc.materials[sc.meshes[lh.mesh].material], ev) * sc.lights[ml].intensity * INV_PI;
            ev.hit = saved;
          }
        } else if (LL.kind == L_ENV) {
          lp = EnvPdfLi(LL, ev.wi);
          Lb = EnvLe(LL, ev.wi, lanes);
        }
        weight = specular ? 1.f : PowerHeuristic(scatteringPDF, lp);
        if (lp == 0 && !specular) Lb = Spec(0);
      }
      if (!IsBlack(Lb, lanes)) Ld += f * Lb * weight / scatteringPDF;
    }
    return L + Ld / lightPdf;
  }
  return L;
}

// Camera::GenerateRay (Camera.cpp:41-71) with the aperture draws of Aperture.cpp:10-27
static void GenerateRay(const lgpu_camera &cam, float u, float v, Sampler &smp, V3 *o, V3 *d, bool corrected) {
  const V3 origin(cam.origin), xh(cam.x_hat), yh(cam.y_hat), zh(cam.z_hat);
  const float tanFov = cam.tan_fov, tanFov2 = 2.f * cam.tan_fov, aspect = cam.aspect;
  if (cam.type == LGPU_CAMERA_SPHERICAL) {
    const float phi = PI2 * u + cam.offset_phi;
    const float theta = PI * v + cam.offset_theta;
    const V3 dd = Norm(SphericalDirection(std::sin(theta), std::cos(theta), phi));
    *o = origin;
    *d = xh * dd.x + yh * dd.y + zh * dd.z;
    return;
  }
  if (cam.type == LGPU_CAMERA_PINHOLE) {
    const V3 p = origin + xh * (u * -tanFov2 + tanFov) + yh * (v * -tanFov2 * aspect + tanFov * aspect) + zh;
    *o = origin;
    *d = Norm(p - origin);
    return;
  }
  const V3 fp = origin + (xh * (u * -tanFov2 + tanFov) + yh * (v * -tanFov2 * aspect + tanFov * aspect) + zh) * cam.focal_length;
  V2 ap(0, 0);
  if (cam.aperture == LGPU_APERTURE_CIRCULAR) {
    const V2 dk = SampleUnitDisk(smp.Get2D());
    const float s = cam.aperture_size;
    ap = V2(dk.x * s - s * .5f, dk.y * s - s * .5f);  // A24: offset disk (as written)
    if (corrected) ap = V2(dk.x * s, dk.y * s);
  } else if (cam.aperture == LGPU_APERTURE_BLADE) {
    const float dth = PI2 / (float)cam.blades;
    const float rnd = PI2 * smp.Get1D();
    const float tmin = std::floor(rnd / dth) * dth;
    const float tmax = tmin + dth;
    const V2 p1(std::cos(tmin), std::sin(tmin)), p2(std::cos(tmax), std::sin(tmax));
    const V2 uu = smp.Get2D();
    ap = V2((p1.x * uu.x + p2.x * uu.y) * cam.aperture_size, (p1.y * uu.x + p2.y * uu.y) * cam.aperture_size);
  }
  const V3 oo = origin + xh * ap.x + yh * ap.y;
  *o = oo;
  *d = Norm(fp - oo);
}

// TileRenderers::UniformSpp per-sample body (Render.cpp:53-60): returns the path's spectrum.
// UtilityIntegrator::Li (UtilityIntegrators.cpp:7-11) over DepthPass::P (:26-33), NormalPass::P
// (:46-53) and AOVPass::P (:66-77): a colour of the closest hit, as a reflectance spectrum.
static Spec UtilityLi(const Ctx &c, const RenderParams &rp, const V3 &ro, const V3 &rd, const Lanes &lanes) {
  Hit hit;
  float col[3] = {0, 0, 0};
  const bool h = Intersect(*c.sc, ro, rd, hit, c.corrected);
  if (rp.integrator == LGPU_INTEGRATOR_DEPTH) {
    if (h) {
      const float d = std::pow(Mag(hit.point - ro) / rp.maxDepth, rp.power);
      col[0] = col[1] = col[2] = d;
    } else {
      col[0] = rp.background[0]; col[1] = rp.background[1]; col[2] = rp.background[2];
    }
  } else if (rp.integrator == LGPU_INTEGRATOR_NORMAL && h) {
    const V3 n = rp.normalG ? hit.normalG : hit.normalS;
    col[0] = n.x * .5f + .5f;
    col[1] = n.y * .5f + (c.corrected ? .5f : 1.f);  // A50: the y channel is offset by 1 as written
    col[2] = n.z * .5f + .5f;
  }
  // AOVPass: the material AOV lookup is commented out (:72-75); the colour stays black.
  return FromRGB(col, lanes);
}

static Spec RenderSample(const Ctx &c, const RenderParams &rp, const lgpu_camera &cam, uint32_t W, uint32_t H,
                         uint32_t pixel, uint32_t sample, float *lambda0) {
  Sampler smp;
  smp.pixel = pixel; smp.sample = sample; smp.seed = rp.seed; smp.dim = 0;
  Lanes lanes;
  lanes.mode = (int)rp.mode;
  float l0 = 0;
  if (rp.mode == HERO4) { HeroWavelengths(WavelengthU(pixel, sample, rp.seed), lanes.lambda); l0 = lanes.lambda[0]; }
  else { lanes.lambda[0] = lanes.lambda[1] = lanes.lambda[2] = lanes.lambda[3] = 0; }
  const uint32_t x = pixel % W, y = pixel / W;
  const float xi = 1.f / (float)W, yi = 1.f / (float)H;
  const float u = xi * ((float)x + smp.Get1D() - .5f);
  const float v = yi * ((float)y + smp.Get1D() - .5f);
  V3 o, d;
  GenerateRay(cam, u, v, smp, &o, &d, c.corrected);
  if (lambda0) *lambda0 = l0;
  if (rp.integrator == LGPU_INTEGRATOR_DIRECT) return DirectLi(c, o, d, smp, lanes);
  if (rp.integrator >= LGPU_INTEGRATOR_DEPTH) return UtilityLi(c, rp, o, d, lanes);
  if (rp.integrator == LGPU_INTEGRATOR_VOLPATH) return VolLi(c, rp, o, d, smp, lanes, false);
  if (rp.integrator == LGPU_INTEGRATOR_MIS_VOLPATH) return VolLi(c, rp, o, d, smp, lanes, true);
  return PathLi(c, rp, o, d, smp, lanes);
}

// Spectral -> linear sRGB for the film (project Hero4 estimator, XYZToRGB Spectrum.h:194-198)
static void ToRGB(const Spec &L, int mode, const float lambda[4], double rgb[3]) {
  if (mode != HERO4) { rgb[0] = L.c[0]; rgb[1] = L.c[1]; rgb[2] = L.c[2]; return; }
  double xyz[3] = {0, 0, 0};
  for (int i = 0; i < 4; ++i) {
    float cm[3];
    CMF(lambda[i], cm);
    for (int k = 0; k < 3; ++k) xyz[k] += (double)L.c[i] * (double)cm[k];
  }
  const double s = 470.0 / 4.0 / CIE_Y_INTEGRAL;
  for (int k = 0; k < 3; ++k) xyz[k] *= s;
  rgb[0] = 3.240479 * xyz[0] - 1.537150 * xyz[1] - 0.498535 * xyz[2];
  rgb[1] = -0.969256 * xyz[0] + 1.875991 * xyz[1] + 0.041556 * xyz[2];
  rgb[2] = 0.055648 * xyz[0] - 0.204043 * xyz[1] + 1.057311 * xyz[2];
}

static void ParallelFor(uint32_t n, int threads, const std::function<void(uint32_t)> &fn) {
  if (threads <= 1 || n < 2) { for (uint32_t i = 0; i < n; ++i) fn(i); return; }
  std::atomic<uint32_t> next(0);
  std::vector<std::thread> ts;
  for (int t = 0; t < threads; ++t)
    ts.emplace_back([&]() { for (uint32_t i; (i = next.fetch_add(1)) < n;) fn(i); });
  for (auto &t : ts) t.join();
}

}  // namespace orc

// ======================================================================== C API
extern "C" {

oracle_scene oracle_scene_create(void) { return new oracle_scene_s(); }
void oracle_scene_destroy(oracle_scene s) { delete s; }

int oracle_scene_add_texture(oracle_scene s, const lgpu_texture_desc *d) {
  if (!s || !d || !d->data || !d->width || !d->height || (d->channels != 1 && d->channels != 3 && d->channels != 4)) return -1;
  Texture t;
  t.w = d->width; t.h = d->height; t.ch = d->channels;
  t.data.assign(d->data, d->data + (size_t)d->width * d->height * d->channels);
  s->textures.push_back(std::move(t));
  return (int)s->textures.size() - 1;
}
int oracle_scene_add_medium(oracle_scene s, const lgpu_medium_desc *d) {
  if (!s || !d) return -1;
  MediumO m;
  m.type = (int)d->type;
  for (int k = 0; k < 3; ++k) { m.sa[k] = d->sigma_a[k]; m.ss[k] = d->sigma_s[k]; }
  m.g = d->g;
  if (d->type == LGPU_MEDIUM_GRID) {
    if (!d->density || !d->nx || !d->ny || !d->nz) return -1;
    m.nx = d->nx; m.ny = d->ny; m.nz = d->nz;
    m.dens.assign(d->density, d->density + (size_t)d->nx * d->ny * d->nz);
    m.bmin = V3(d->bounds_min); m.bmax = V3(d->bounds_max);
  }
  s->media.push_back(std::move(m));
  return (int)s->media.size() - 1;
}
int oracle_scene_add_material(oracle_scene s, const lgpu_material_desc *d) {
  if (!s || !d || (d->num_nodes && !d->nodes)) return -1;
  Material m;
  m.nodes.assign(d->nodes, d->nodes + d->num_nodes);
  m.bxdf = d->bxdf_root; m.emission = d->emission_root; m.intensity = d->emission_intensity;
  m.medIn = d->medium_interior; m.medOut = d->medium_exterior;
  for (const lgpu_node &n : m.nodes) {  // texture references and noise parameters
    const bool tex = n.type == LGPU_NODE_IMAGE_COLOUR || n.type == LGPU_NODE_IMAGE_SCALAR || n.type == LGPU_NODE_SPECTRAL ||
                     n.type == LGPU_NODE_IMAGE_CHANNEL || n.type == LGPU_NODE_SPECTRAL_TEXTURE;
    if (tex && !(n.p[0] >= 0.f && (size_t)n.p[0] < s->textures.size())) return -1;
    if (n.type == LGPU_NODE_SPECTRAL) {
      const Texture &t = s->textures[(size_t)n.p[0]];
      if (t.w * t.h < 2 || !(n.p[2] > n.p[1])) return -1;
    }
    if (n.type == LGPU_NODE_OCTAVE_NOISE) {
      const uint32_t b = (uint32_t)n.p[2];
      if (!(n.p[1] >= 0.f && n.p[1] <= 32.f)) return -1;
      if (b != LGPU_NODE_CHECKER && b != LGPU_NODE_PERLIN && b != LGPU_NODE_VALUE_NOISE && b != LGPU_NODE_VORONOI) return -1;
    }
  }
  m.bbRGB.assign(3 * m.nodes.size(), 0.f);
  m.bbConst.assign(m.nodes.size(), 0);
  for (size_t i = 0; i < m.nodes.size(); ++i) {  // constant-temperature blackbody: evaluate once
    const lgpu_node &n = m.nodes[i];
    if (n.type == LGPU_NODE_BLACKBODY && n.in[0] < m.nodes.size() && m.nodes[n.in[0]].type == LGPU_NODE_SCALAR) {
      BlackbodyRGB(m.nodes[n.in[0]].p[0], &m.bbRGB[3 * i]);
      m.bbConst[i] = 1;
    }
  }
  s->materials.push_back(std::move(m));
  return (int)s->materials.size() - 1;
}
int oracle_scene_add_mesh(oracle_scene s, const lgpu_mesh_desc *d) {
  if (!s || !d || !d->positions || !d->indices || d->material >= s->materials.size()) return -1;
  Mesh m;
  for (uint32_t i = 0; i < d->num_vertices; ++i) m.p.push_back(V3(&d->positions[3 * i]));
  if (d->normals && d->tangents) {
    for (uint32_t i = 0; i < d->num_vertices; ++i) m.n.push_back(V3(&d->normals[3 * i]));
    m.tan.assign(d->tangents, d->tangents + 4 * (size_t)d->num_vertices);
  }
  if (d->uvs) for (uint32_t i = 0; i < d->num_vertices; ++i) m.uv.push_back(V2(d->uvs[2 * i], d->uvs[2 * i + 1]));
  m.idx.assign(d->indices, d->indices + 3 * (size_t)d->num_triangles);
  m.material = d->material;
  m.firstTri = (uint32_t)s->triMesh.size();
  const int mi = (int)s->meshes.size();
  for (uint32_t t = 0; t < d->num_triangles; ++t) s->triMesh.push_back((uint32_t)mi);
  const Material &mat = s->materials[d->material];
  if (mat.emission != LGPU_NONE) {  // MeshLight(mesh) + Scene::AddObject(_addLight) (Scene.cpp:114-121)
    Light L;
    L.kind = L_MESH;
    L.mesh = mi;
    L.intensity = mat.intensity;
    std::vector<float> areas(d->num_triangles);
    for (uint32_t t = 0; t < d->num_triangles; ++t) {
      float a;
      TriAreaNormal(m.p[m.idx[3 * t]], m.p[m.idx[3 * t + 1]], m.p[m.idx[3 * t + 2]], &a, nullptr);
      areas[t] = a;
    }
    L.triDist = Piecewise1D(areas.data(), d->num_triangles);
    m.meshLight = (int)s->lights.size();
    s->lights.push_back(std::move(L));
  }
  s->meshes.push_back(std::move(m));
  return mi;
}
// InstanceProxy (Instance.cpp:7-16): a mesh kept for instancing.
int oracle_scene_add_proxy(oracle_scene s, const lgpu_mesh_desc *d) {
  if (!s || !d || !d->positions || !d->indices || d->material >= s->materials.size()) return -1;
  Proxy p;
  p.nv = d->num_vertices; p.nt = d->num_triangles; p.material = d->material;
  p.pos.assign(d->positions, d->positions + 3 * (size_t)p.nv);
  p.idx.assign(d->indices, d->indices + 3 * (size_t)p.nt);
  if (d->normals && d->tangents) {
    p.nrm.assign(d->normals, d->normals + 3 * (size_t)p.nv);
    p.tan.assign(d->tangents, d->tangents + 4 * (size_t)p.nv);
  }
  if (d->uvs) p.uv.assign(d->uvs, d->uvs + 2 * (size_t)p.nv);
  s->proxies.push_back(std::move(p));
  return (int)s->proxies.size() - 1;
}
// Instance::Commit (Instance.cpp:22-34) with a world-space result: points through affine3's
// operator* (affine3.h:38-40), normals through the inverse transpose of the linear part (the
// columns of its cofactor matrix are b x c, c x a, a x b), tangents through the linear part.
int oracle_scene_add_instance(oracle_scene s, uint32_t proxy, const float x[12], uint32_t material) {
  if (!s || proxy >= s->proxies.size() || (material != LGPU_NONE && material >= s->materials.size())) return -1;
  const Proxy &p = s->proxies[proxy];
  const V3 a(x[0], x[1], x[2]), b(x[3], x[4], x[5]), c(x[6], x[7], x[8]), t(x[9], x[10], x[11]);
  const V3 k0 = Cross(b, c), k1 = Cross(c, a), k2 = Cross(a, b);
  const float det = a.x * k0.x + a.y * k0.y + a.z * k0.z;
  if (!(det != 0.f && std::isfinite(det))) return -1;
  const float sg = det < 0 ? -1.f : 1.f;
  std::vector<float> pos(3 * (size_t)p.nv), nrm(p.nrm.size()), tan(p.tan.size());
  for (uint32_t v = 0; v < p.nv; ++v) {
    const float px = p.pos[3 * v], py = p.pos[3 * v + 1], pz = p.pos[3 * v + 2];
    pos[3 * v] = a.x * px + b.x * py + c.x * pz + t.x;
    pos[3 * v + 1] = a.y * px + b.y * py + c.y * pz + t.y;
    pos[3 * v + 2] = a.z * px + b.z * py + c.z * pz + t.z;
    if (p.nrm.empty()) continue;
    const float nx = p.nrm[3 * v], ny = p.nrm[3 * v + 1], nz = p.nrm[3 * v + 2];
    float o[3] = {(k0.x * nx + k1.x * ny + k2.x * nz) * sg, (k0.y * nx + k1.y * ny + k2.y * nz) * sg,
                  (k0.z * nx + k1.z * ny + k2.z * nz) * sg};
    float l = std::sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
    for (int k = 0; k < 3; ++k) nrm[3 * v + k] = l > 0 ? o[k] / l : o[k];
    const float tx = p.tan[4 * v], ty = p.tan[4 * v + 1], tz = p.tan[4 * v + 2];
    float u[3] = {a.x * tx + b.x * ty + c.x * tz, a.y * tx + b.y * ty + c.y * tz, a.z * tx + b.z * ty + c.z * tz};
    l = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    for (int k = 0; k < 3; ++k) tan[4 * v + k] = l > 0 ? u[k] / l : u[k];
    tan[4 * v + 3] = p.tan[4 * v + 3] * sg;
  }
  lgpu_mesh_desc d;
  std::memset(&d, 0, sizeof(d));
  d.num_vertices = p.nv; d.num_triangles = p.nt;
  d.positions = pos.data(); d.indices = p.idx.data();
  d.normals = p.nrm.empty() ? nullptr : nrm.data();
  d.tangents = p.nrm.empty() ? nullptr : tan.data();
  d.uvs = p.uv.empty() ? nullptr : p.uv.data();
  d.material = material == LGPU_NONE ? p.material : material;
  return oracle_scene_add_mesh(s, &d);
}
int oracle_scene_add_light(oracle_scene s, const lgpu_light_desc *d) {
  if (!s || !d) return -1;
  Light L;
  L.intensity = d->intensity;
  L.rgb[0] = d->rgb[0]; L.rgb[1] = d->rgb[1]; L.rgb[2] = d->rgb[2];
  L.temperature = d->temperature;
  if (L.temperature > 0) BlackbodyRGB(L.temperature, L.bbRGB);
  L.position = V3(d->position);
  if (d->type == LGPU_LIGHT_POINT) L.kind = L_POINT;
  else if (d->type == LGPU_LIGHT_SPOT) {  // Spotlight.cpp:14-19
    L.kind = L_SPOT;
    L.axis = V3(d->axis);
    L.cosCone = std::cos(d->cone_angle);
    L.cosFalloff = std::cos(d->falloff_start);
  } else if (d->type == LGPU_LIGHT_ENVIRONMENT) {  // EnvironmentLight.cpp:7-22
    if (!d->env_rgb || !d->env_width || !d->env_height || s->envLight >= 0) return -1;
    L.kind = L_ENV;
    L.env.w = d->env_width; L.env.h = d->env_height; L.env.ch = 3;
    L.env.data.assign(d->env_rgb, d->env_rgb + 3 * (size_t)d->env_width * d->env_height);
    L.offset = V2(d->env_offset[0], d->env_offset[1]);
    const unsigned w = d->env_width, h = d->env_height;
    std::vector<float> img((size_t)w * h);
    for (unsigned y = 0; y < h; ++y) {
      const float vp = (float)y / (float)h;
      const float sinT = std::sin(PI * (float)(y + .5) / (float)h);
      for (unsigned x = 0; x < w; ++x) {
        const float up = (float)x / (float)w;
        float c[3];
        L.env.Bilinear(Fract(up), Fract(vp), c);
        img[x + y * w] = std::abs(0.212671f * c[0] + 0.715160f * c[1] + 0.072169f * c[2]);
        img[x + y * w] *= sinT;
      }
    }
    L.envDist = Piecewise2D(img.data(), w, h);
    s->envLight = (int)s->lights.size();
  } else return -1;
  s->lights.push_back(std::move(L));
  return (int)s->lights.size() - 1;
}

int oracle_scene_commit(oracle_scene s, uint32_t strategy) {
  if (!s || s->committed) return -1;
  s->strategy = strategy;
  BuildBVH(*s);
  Ctx c{s, false, RGB};
  if (s->envLight >= 0) {  // Scene::Commit (Scene.cpp:18-26)
    Light &E = s->lights[s->envLight];
    E.envBounds = s->sceneBounds;
    E.radius = s->sceneBounds.MaxLength();
  }
  if (strategy == LGPU_LIGHTS_POWER) {  // PowerLightSampler::Commit (LightSampler.cpp:22-35)
    const unsigned n = (unsigned)s->lights.size();
    if (n) {
      std::vector<float> imp(n);
      float total = 0;
      for (unsigned i = 0; i < n; ++i) { imp[i] = LightPower(*s, s->lights[i]); total += imp[i]; }
      s->powerDist = Piecewise1D(imp.data(), n);
      s->invTotalPower = 1.f / total;
    }
  } else {  // ManyLightSampler::Commit / InitLights (ManyLightSampler.cpp:37-44, 250-303)
    s->triToTLight.assign(s->triMesh.size(), -1);
    s->sceneToTLight.assign(s->lights.size(), -1);
    for (size_t li = 0; li < s->lights.size(); ++li) {
      const Light &L = s->lights[li];
      if (L.kind != L_MESH) continue;
      const Mesh &m = s->meshes[L.mesh];
      for (uint32_t t = 0; t < m.idx.size() / 3; ++t) {
        Light T;
        T.kind = L_TRI; T.mesh = L.mesh; T.tri = t; T.parentMeshLight = (int)li; T.intensity = L.intensity;
        s->triToTLight[m.firstTri + t] = (int)s->tlights.size();
        s->tlights.push_back(std::move(T));
      }
    }
    for (size_t li = 0; li < s->lights.size(); ++li) {
      const Light &L = s->lights[li];
      if (L.kind == L_POINT || L.kind == L_SPOT) {
        s->sceneToTLight[li] = (int)s->tlights.size();
        s->tlights.push_back(L);
      }
    }
    TreeBuilder tb(*s);
    tb.powers.resize(s->tlights.size());
    for (size_t i = 0; i < s->tlights.size(); ++i) {
      const Light &T = s->tlights[i];
      tb.powers[i] = T.kind == L_TRI ? TriLightIrradiance(c, T) * [&]() { float a; V3 v0, v1, v2; TriVerts(*s, s->meshes[T.mesh].firstTri + T.tri, v0, v1, v2); TriAreaNormal(v0, v1, v2, &a, nullptr); return a; }()
                                     : LightPower(*s, T);
    }
    s->tOrder.resize(s->tlights.size());
    for (size_t i = 0; i < s->tlights.size(); ++i) s->tOrder[i] = (int)i;
    s->lightLeaf.assign(s->tlights.size(), {nullptr, 0});
    s->treePower = 0;
    s->infPower = 0;
    if (s->envLight >= 0) s->infPower = LightPower(*s, s->lights[s->envLight]) * .1f;
    if (!s->tlights.empty()) {
      LNode *r = new LNode();
      s->tnodes.emplace_back(r);
      r->count = (unsigned)s->tlights.size();
      r->first = 0;
      r->bounds = tb.B(0);
      r->power = 0;
      r->cone = MakeCone(tb.D(0));
      for (unsigned i = 0; i < r->count; ++i) {
        r->bounds = Union(r->bounds, tb.B(i));
        Cone tmp = MakeCone(tb.D(i));
        r->cone = ConeUnion(r->cone, tmp);
        r->power += tb.P(i);
      }
      s->root = r;
      tb.Build(r);
    }
  }
  s->committed = true;
  return 0;
}

int oracle_light_tree(oracle_scene s, float *out, uint32_t max_nodes) {
  if (!s || !s->root) return 0;
  // Depth-first pre-order, left child first: the product's flattening order.
  std::vector<const LNode *> order;
  std::vector<const LNode *> st{s->root};
  while (!st.empty()) {
    const LNode *n = st.back();
    st.pop_back();
    order.push_back(n);
    if (!n->IsLeaf()) { st.push_back(n->child[1]); st.push_back(n->child[0]); }
  }
  auto idxOf = [&](const LNode *n) -> float { for (size_t i = 0; i < order.size(); ++i) if (order[i] == n) return (float)i; return -1.f; };
  for (size_t i = 0; i < order.size() && i < max_nodes; ++i) {
    const LNode *n = order[i];
    float *o = out + 15 * i;
    const V3 c = n->bounds.Center();
    o[0] = c.x; o[1] = c.y; o[2] = c.z; o[3] = n->bounds.Diagonal();
    o[4] = n->cone.axis.x; o[5] = n->cone.axis.y; o[6] = n->cone.axis.z;
    o[7] = n->cone.thetaO; o[8] = n->cone.thetaE; o[9] = n->power;
    o[10] = n->IsLeaf() ? -1.f : idxOf(n->child[0]);
    o[11] = n->IsLeaf() ? -1.f : idxOf(n->child[1]);
    o[12] = (float)n->first; o[13] = (float)n->count;
    o[14] = n->parent ? idxOf(n->parent) : -1.f;
  }
  return (int)order.size();
}

static RenderParams Params(const lgpu_render_desc *d) {
  RenderParams rp;
  rp.integrator = d->integrator; rp.maxBounces = d->max_bounces; rp.minBounces = d->min_bounces;
  rp.mode = d->spectral_mode; rp.seed = d->seed; rp.corrected = d->estimator_mode == LGPU_ESTIMATOR_CORRECTED;
  rp.maxDepth = d->pass_max_depth; rp.power = d->pass_power;
  for (int k = 0; k < 3; ++k) rp.background[k] = d->pass_background[k];
  rp.normalG = d->pass_geometric_normal != 0;
  return rp;
}

int oracle_trace_paths(oracle_scene s, const lgpu_render_desc *d, uint32_t width, uint32_t height, uint32_t n,
                       const uint32_t *pixels, const uint32_t *samples, float *out, int threads) {
  if (!s || !s->committed || !d || !width || !height) return -1;
  if (d->integrator > LGPU_INTEGRATOR_AOV) return -3;
  if (d->integrator == LGPU_INTEGRATOR_DEPTH && !(d->pass_max_depth > 0.f)) return -3;
  const RenderParams rp = Params(d);
  Ctx c{s, rp.corrected, (int)rp.mode};
  c.volumes = rp.corrected && !s->media.empty();
  ParallelFor(n, threads, [&](uint32_t i) {
    float l0;
    const Spec L = RenderSample(c, rp, d->camera, width, height, pixels[i], samples[i], &l0);
    for (int k = 0; k < 4; ++k) out[5 * i + k] = L.c[k];
    if (rp.mode != HERO4) out[5 * i + 3] = 0;
    out[5 * i + 4] = l0;
  });
  return 0;
}

int oracle_render(oracle_scene s, const lgpu_render_desc *d, uint32_t W, uint32_t H, uint32_t x0, uint32_t y0,
                  uint32_t w, uint32_t h, int threads, double *rgb, uint32_t *counts) {
  if (!s || !s->committed || !d || x0 + w > W || y0 + h > H) return -1;
  if (d->integrator > LGPU_INTEGRATOR_AOV) return -3;
  if (d->integrator == LGPU_INTEGRATOR_DEPTH && !(d->pass_max_depth > 0.f)) return -3;
  const RenderParams rp = Params(d);
  Ctx c{s, rp.corrected, (int)rp.mode};
  c.volumes = rp.corrected && !s->media.empty();
  const uint32_t TS = 16;
  const uint32_t tx = (w + TS - 1) / TS, ty = (h + TS - 1) / TS;
  std::memset(rgb, 0, sizeof(double) * 3 * w * h);
  std::memset(counts, 0, sizeof(uint32_t) * w * h);
  ParallelFor(tx * ty, threads, [&](uint32_t t) {
    const uint32_t bx = (t % tx) * TS, by = (t / tx) * TS;
    for (uint32_t y = by; y < std::min(by + TS, h); ++y)
      for (uint32_t x = bx; x < std::min(bx + TS, w); ++x) {
        const uint32_t pixel = (y + y0) * W + (x + x0);
        for (uint32_t smp = d->spp_begin; smp < d->spp_end; ++smp) {
          float lam[4];
          if (rp.mode == HERO4) HeroWavelengths(WavelengthU(pixel, smp, rp.seed), lam);
          float l0;
          const Spec L = RenderSample(c, rp, d->camera, W, H, pixel, smp, &l0);
          double o[3];
          ToRGB(L, (int)rp.mode, lam, o);
          double *px = rgb + 3 * ((size_t)y * w + x);
          px[0] += o[0]; px[1] += o[1]; px[2] += o[2];
          counts[(size_t)y * w + x]++;
        }
      }
  });
  return 0;
}

int oracle_intersect(oracle_scene s, uint32_t n, const float *rays, uint32_t *hits) {
  if (!s || !s->committed) return -1;
  for (uint32_t i = 0; i < n; ++i) {
    const float *r = rays + 8 * i;
    float t, u, v;
    uint32_t prim;
    if (ClosestHit(*s, V3(r[0], r[1], r[2]), V3(r[4], r[5], r[6]), r[3], r[7], &t, &u, &v, &prim)) {
      std::memcpy(&hits[4 * i], &t, 4); std::memcpy(&hits[4 * i + 1], &u, 4); std::memcpy(&hits[4 * i + 2], &v, 4);
      hits[4 * i + 3] = prim;
    } else {
      const float inf = INFINITY;
      std::memcpy(&hits[4 * i], &inf, 4); hits[4 * i + 1] = 0; hits[4 * i + 2] = 0; hits[4 * i + 3] = LGPU_NONE;
    }
  }
  return 0;
}
int oracle_occluded(oracle_scene s, uint32_t n, const float *rays, uint32_t *occ) {
  if (!s || !s->committed) return -1;
  for (uint32_t i = 0; i < n; ++i) {
    const float *r = rays + 8 * i;
    occ[i] = AnyHit(*s, V3(r[0], r[1], r[2]), V3(r[4], r[5], r[6]), r[3], r[7]) ? 1u : 0u;
  }
  return 0;
}

void oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  Philox(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = c[i];
}
float oracle_fr_dielectric(float c, float a, float b) { return FrDielectric(c, a, b); }
void oracle_fr_conductor(float cosI, const float ei[3], const float et[3], const float k[3], float out[3]) {
  Spec a, b, kk;
  for (int i = 0; i < 3; ++i) { a.c[i] = ei[i]; b.c[i] = et[i]; kk.c[i] = k[i]; }
  a.c[3] = b.c[3] = kk.c[3] = 1;
  const Spec r = FrConductor(cosI, a, b, kk);
  for (int i = 0; i < 3; ++i) out[i] = r.c[i];
}
float oracle_microfacet_d(int ggx, const float wh[3], float ax, float ay) { return MfD(ggx, V3(wh), ax, ay); }
float oracle_microfacet_lambda(int ggx, const float w[3], float ax, float ay) { return MfLambda(ggx, V3(w), ax, ay); }
void oracle_blackbody_rgb(float T, float out[3]) { BlackbodyRGB(T, out); }
float oracle_blackbody_normalised(float l, float T) { return BlackbodyNormalised(l, T); }
void oracle_cmf(float l, float out[3]) { CMF(l, out); }

// Medium closed forms (test hooks). lanes: mode 0 RGB, 1 HERO4 with lambda[4].
static Lanes MkLanes(int mode, const float *lambda) {
  Lanes l;
  l.mode = mode;
  for (int i = 0; i < 4; ++i) l.lambda[i] = lambda ? lambda[i] : 0.f;
  return l;
}
// ---------------------------------------------------------------- denoiser
// The library's replacement for OIDN's filter behind Denoise::SetData/Process (Denoise.cpp:13-67):
// the edge-avoiding A-Trous wavelet transform (Dammertz et al. 2010), restated per pixel on the
// film means (3 floats per pixel; albedo and normal may be null): `iterations` passes of the 5x5
// B3-spline kernel with taps 2^i apart, weight h(dx) h(dy) exp(-(|dc|^2/sc_i^2 + |dn|^2/sn^2 +
// |da|^2/sa^2)), sc_i = sigma_colour 2^-i, a zero sigma dropping its term; taps outside the image
// are skipped. Demodulation divides colour by albedo channels above 1e-3 and multiplies back.
void oracle_denoise(uint32_t w, uint32_t h, const float *colour, const float *albedo, const float *normal,
                    const lgpu_denoise_desc *d, float *out) {
  const size_t n = (size_t)w * h;
  const float eps = 1e-3f;
  const bool demod = albedo && d->demodulate;
  std::vector<float> a(colour, colour + 3 * n), b(3 * n);
  if (demod)
    for (size_t i = 0; i < 3 * n; ++i) a[i] = albedo[i] > eps ? a[i] / albedo[i] : a[i];
  const float kw[3] = {0.375f, 0.25f, 0.0625f};
  auto dist2 = [](const float *p, const float *q) {
    const float x = q[0] - p[0], y = q[1] - p[1], z = q[2] - p[2];
    return x * x + y * y + z * z;
  };
  for (uint32_t it = 0; it < d->iterations; ++it) {
    const int step = 1 << it;
    const float sc = d->sigma_colour * std::ldexp(1.f, -(int)it);
    const float isc = sc > 0.f ? 1.f / (sc * sc) : 0.f;
    const float isn = d->sigma_normal > 0.f ? 1.f / (d->sigma_normal * d->sigma_normal) : 0.f;
    const float isa = d->sigma_albedo > 0.f ? 1.f / (d->sigma_albedo * d->sigma_albedo) : 0.f;
    for (uint32_t y = 0; y < h; ++y)
      for (uint32_t x = 0; x < w; ++x) {
        const size_t pi = (size_t)y * w + x;
        float sx = 0.f, sy = 0.f, sz = 0.f, sw = 0.f;
        for (int dy = -2; dy <= 2; ++dy) {
          const int qy = (int)y + dy * step;
          if (qy < 0 || qy >= (int)h) continue;
          for (int dx = -2; dx <= 2; ++dx) {
            const int qx = (int)x + dx * step;
            if (qx < 0 || qx >= (int)w) continue;
            const size_t qi = (size_t)qy * w + (size_t)qx;
            float e = dist2(&a[3 * pi], &a[3 * qi]) * isc;
            if (normal) e += dist2(normal + 3 * pi, normal + 3 * qi) * isn;
            if (albedo) e += dist2(albedo + 3 * pi, albedo + 3 * qi) * isa;
            const float wt = kw[std::abs(dx)] * kw[std::abs(dy)] * std::exp(-e);
            sx += wt * a[3 * qi]; sy += wt * a[3 * qi + 1]; sz += wt * a[3 * qi + 2]; sw += wt;
          }
        }
        b[3 * pi] = sx / sw; b[3 * pi + 1] = sy / sw; b[3 * pi + 2] = sz / sw;
      }
    a.swap(b);
  }
  for (size_t i = 0; i < 3 * n; ++i) out[i] = demod && albedo[i] > eps ? a[i] * albedo[i] : a[i];
}

// ---------------------------------------------------------------- post-processing
// ConvolutionFilter (PostProcessing.cpp:16-45), AcesFitted + ClampTextureChannels (ToneMap.cpp:19-73),
// Gamma + PackChannels (Texture.cpp:19-32). in: RGBA32F film means; out32 (RGBA) and out8 (RGBA8)
// may be null. asWritten: ACES matrices by the transpose and the convolution's unsigned edge wrap.
static const float ACES_IN[3][3] = {{0.59719f, 0.35458f, 0.04823f}, {0.07600f, 0.90834f, 0.01566f}, {0.02840f, 0.13383f, 0.83777f}};
static const float ACES_OUT[3][3] = {{1.60475f, -0.53108f, -0.07367f}, {-0.10208f, 1.10813f, -0.00605f}, {-0.00327f, -0.07276f, 1.07602f}};
static V3 Mat3(const float m[3][3], const V3 &v, bool asWritten) {  // MulVec3Mat3x3 (ToneMap.cpp:19-24)
  if (asWritten) return V3(m[0][0] * v.x + m[1][0] * v.y + m[2][0] * v.z, m[0][1] * v.x + m[1][1] * v.y + m[2][1] * v.z,
                           m[0][2] * v.x + m[1][2] * v.y + m[2][2] * v.z);
  return V3(m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
            m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z);
}
static float RrtFit(float v) {  // RRTAndOTDFit (ToneMap.cpp:29-33)
  return (v * (v + 0.0245786f) - 0.000090537f) / (v * (v * 0.983729f + 0.4329510f) + 0.238081f);
}
static int PackChannel(float c) {  // out-of-range (int) conversion as x86 gives it: INT_MIN
  const float s = c * 256.f;
  const int i = (s > -2147483648.f && s < 2147483648.f) ? (int)s : INT32_MIN;
  return std::max(0, std::min(255, i));
}
void oracle_post(const float *in, uint32_t w, uint32_t h, const lgpu_post_desc *d, float *out32, uint8_t *out8) {
  const bool aw = d->estimator_mode == LGPU_ESTIMATOR_FAITHFUL;
  for (uint32_t y = 0; y < h; ++y)
    for (uint32_t x = 0; x < w; ++x) {
      V3 c;
      if (d->kernel) {
        c = V3(0, 0, 0);
        const uint32_t kw2 = d->kernel_w / 2, kh2 = d->kernel_h / 2;
        for (uint32_t ky = 0; ky < d->kernel_h; ++ky)
          for (uint32_t kx = 0; kx < d->kernel_w; ++kx) {
            uint32_t sx, sy;
            if (aw) {  // maths::Clamp(_x + x - kw2, 0u, w - 1) on unsigned values
              sx = std::min(x + kx - kw2, w - 1);
              sy = std::min(y + ky - kh2, h - 1);
            } else {
              sx = (uint32_t)std::max(0, std::min((int)w - 1, (int)x + (int)kx - (int)kw2));
              sy = (uint32_t)std::max(0, std::min((int)h - 1, (int)y + (int)ky - (int)kh2));
            }
            const float k = d->kernel[ky * d->kernel_w + kx] * d->kernel_intensity;
            const float *t = in + 4 * ((size_t)sy * w + sx);
            c.x += k * t[0]; c.y += k * t[1]; c.z += k * t[2];
          }
      } else {
        const float *t = in + 4 * ((size_t)y * w + x);
        c = V3(t[0], t[1], t[2]);
      }
      if (d->tonemap == LGPU_TONEMAP_ACES_FITTED) {
        c = Mat3(ACES_IN, c, aw);
        c = V3(RrtFit(c.x), RrtFit(c.y), RrtFit(c.z));
        c = Mat3(ACES_OUT, c, aw);
      }
      if (d->clamp) c = V3(Clampf(c.x, 0.f, 1.f), Clampf(c.y, 0.f, 1.f), Clampf(c.z, 0.f, 1.f));
      const size_t i = (size_t)y * w + x;
      if (out32) { out32[4 * i] = c.x; out32[4 * i + 1] = c.y; out32[4 * i + 2] = c.z; out32[4 * i + 3] = 1.f; }
      if (out8) {
        if (d->gamma) {
          const float g = (float)(1.0 / 2.2);
          c = V3(std::pow(c.x, g), std::pow(c.y, g), std::pow(c.z, g));
        }
        out8[4 * i] = (uint8_t)PackChannel(c.x); out8[4 * i + 1] = (uint8_t)PackChannel(c.y);
        out8[4 * i + 2] = (uint8_t)PackChannel(c.z); out8[4 * i + 3] = 255;
      }
    }
}

// One shader-node evaluation (test hook). ev[11] = {u, v, point xyz, shading normal xyz, eta,
// woL.y, medium interaction (0/1)}; out[4] receives the value (a spectrum's lanes, or the
// components of a scalar / vec / colour). Returns the value type (VT_*) or -1.
int oracle_eval_node(oracle_scene s, uint32_t material, uint32_t node, const float ev[11], int mode, const float *lambda,
                     float out[4]) {
  if (!s || material >= s->materials.size() || node >= s->materials[material].nodes.size()) return -1;
  Hit h;
  h.uv = V2(ev[0], ev[1]);
  h.point = V3(ev[2], ev[3], ev[4]);
  h.normalS = V3(ev[5], ev[6], ev[7]);
  Event e;
  e.hit = &h;
  e.eta = ev[8];
  e.woL = V3(0, ev[9], 0);
  e.mediumInteraction = ev[10] != 0.f;
  e.lanes = MkLanes(mode, lambda);
  const Val v = EvalNode(*s, s->materials[material], node, e);
  for (int i = 0; i < 4; ++i) out[i] = v.t == VT_SPECTRUM ? v.s.c[i] : v.v[i];
  return v.t;
}
// One BxDF of a material at a shading frame equal to the world frame (tangent x, normal y,
// bitangent z; normalG = normalS = +y), hit point at the origin. Test access for the closed-form
// pins of the lobes (sampling against pdf, energy, reciprocity). out per sample i (Philox pixel
// `pixel`, sample i): wiL[3], pdf, f[4] (f cos|wiL.y| is NOT applied), spec flag, offset side of
// the spawn point along normalG (+1, -1 or 0), eta after the sample.
static void LocalFrameHit(Hit &h) {
  h.point = V3(0, 0, 0);
  h.normalG = V3(0, 1, 0);
  h.normalS = V3(0, 1, 0);
  h.tangent = V3(1, 0, 0);
  h.bitangent = V3(0, 0, 1);
}
int oracle_bxdf_sample_n(oracle_scene s, uint32_t material, const float woL[3], int mode, const float *lambda,
                         int corrected, uint32_t pixel, uint32_t n, float eta, float *out) {
  if (!s || material >= s->materials.size() || s->materials[material].bxdf == LGPU_NONE) return -1;
  const Material &m = s->materials[material];
  Ctx c{s, corrected != 0, mode};
  for (uint32_t i = 0; i < n; ++i) {
    Hit h;
    LocalFrameHit(h);
    Sampler smp;
    smp.pixel = pixel; smp.sample = i; smp.dim = 0; smp.seed = 0;
    Event e;
    e.hit = &h;
    e.sampler = &smp;
    e.eta = eta;
    e.lanes = MkLanes(mode, lambda);
    e.wo = V3(woL[0], woL[1], woL[2]);
    e.wi = -e.wo;
    e.SurfaceLocalise();
    float pdf = 0;
    bool sp = false;
    const Spec f = BxDF_Sample(c, m, m.bxdf, e, smp, pdf, &sp);
    float *o = out + 11 * (size_t)i;
    o[0] = e.wiL.x; o[1] = e.wiL.y; o[2] = e.wiL.z; o[3] = pdf;
    for (int k = 0; k < 4; ++k) o[4 + k] = f.c[k];
    o[8] = sp ? 1.f : 0.f;
    o[9] = h.point.y > 0 ? 1.f : (h.point.y < 0 ? -1.f : 0.f);
    o[10] = e.eta;
  }
  return 0;
}
// f (4 lanes) and pdf of a material's BxDF for n (woL, wiL) pairs in the same frame: out 5 per pair.
int oracle_bxdf_eval_n(oracle_scene s, uint32_t material, uint32_t n, const float *woL, const float *wiL, int mode,
                       const float *lambda, int corrected, float *out) {
  if (!s || material >= s->materials.size() || s->materials[material].bxdf == LGPU_NONE) return -1;
  const Material &m = s->materials[material];
  Ctx c{s, corrected != 0, mode};
  for (uint32_t i = 0; i < n; ++i) {
    Hit h;
    LocalFrameHit(h);
    Event e;
    e.hit = &h;
    e.lanes = MkLanes(mode, lambda);
    e.wo = V3(woL + 3 * (size_t)i);
    e.wi = V3(wiL + 3 * (size_t)i);
    e.SurfaceLocalise();
    const Spec f = BxDF_f(c, m, m.bxdf, e);
    float *o = out + 5 * (size_t)i;
    for (int k = 0; k < 4; ++k) o[k] = f.c[k];
    o[4] = BxDF_Pdf(c, m, m.bxdf, e.woL, e.wiL, e);
  }
  return 0;
}
int oracle_medium_tr(oracle_scene s, uint32_t med, const float o[3], const float d[3], float tFar, int mode,
                     const float *lambda, float out[4]) {
  if (!s || med >= s->media.size()) return -1;
  const Spec r = MediumTr(s->media[med], V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2]), tFar, MkLanes(mode, lambda));
  for (int i = 0; i < 4; ++i) out[i] = r.c[i];
  return 0;
}
int oracle_medium_sample_distance(oracle_scene s, uint32_t med, const float o[3], const float d[3], float tFar,
                                  float u0, float u1, int mode, const float *lambda, float out[7]) {
  if (!s || med >= s->media.size()) return -1;
  bool inter;
  float t, pdf;
  const Spec w = SampleDistanceU(s->media[med], V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2]), tFar, V2(u0, u1),
                                 MkLanes(mode, lambda), &inter, &t, &pdf);
  out[0] = t; out[1] = pdf; out[2] = inter ? 1.f : 0.f;
  for (int i = 0; i < 4; ++i) out[3 + i] = w.c[i];
  return 0;
}
float oracle_medium_pdf_distance(oracle_scene s, uint32_t med, const float o[3], const float d[3], float t, int mode,
                                 const float *lambda, int corrected) {
  return PdfDistance(s->media[med], V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2]), t, MkLanes(mode, lambda), corrected != 0);
}
int oracle_medium_sample_equiangular(oracle_scene s, uint32_t med, const float o[3], const float d[3],
                                     const float hitPoint[3], float tFar, const float light[3], float eps, int mode,
                                     const float *lambda, int corrected, float out[7]) {
  if (!s || med >= s->media.size()) return -1;
  bool inter;
  float t, pdf;
  const Spec w = SampleEquiangularU(s->media[med], V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2]),
                                    V3(hitPoint[0], hitPoint[1], hitPoint[2]), tFar, V3(light[0], light[1], light[2]),
                                    eps, MkLanes(mode, lambda), corrected != 0, &inter, &t, &pdf);
  out[0] = t; out[1] = pdf; out[2] = inter ? 1.f : 0.f;
  for (int i = 0; i < 4; ++i) out[3 + i] = w.c[i];
  return 0;
}
float oracle_medium_pdf_equiangular(const float o[3], const float d[3], const float light[3], float tFar, float t,
                                    int corrected) {
  return PdfEquiangular(V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2]), V3(light[0], light[1], light[2]), tFar, t,
                        corrected != 0);
}
// Light-tree pick and pdf at a shading point (test hooks). pdfs: LightPdf for every tree light
// and, last, the environment light (0 when absent). Returns the number of tree lights.
int oracle_light_pdfs(oracle_scene s, const float p[3], const float n[3], int corrected, float *pdfs) {
  if (!s || !s->committed || s->strategy != LGPU_LIGHTS_TREE) return -1;
  const Ctx c{s, corrected != 0, 0};
  const V3 P(p[0], p[1], p[2]), N(n[0], n[1], n[2]);
  for (size_t i = 0; i < s->tlights.size(); ++i) pdfs[i] = LightPdf(c, LightPick{1, (int)i}, P, N);
  pdfs[s->tlights.size()] = s->envLight >= 0 ? LightPdf(c, LightPick{0, s->envLight}, P, N) : 0.f;
  return (int)s->tlights.size();
}
// SampleLight with the sampler of (pixel, sample): picked tree light index (-1 for the environment
// light, -2 none) and the pick pdf.
int oracle_light_sample(oracle_scene s, const float p[3], const float n[3], int corrected, uint32_t pixel,
                        uint32_t sample, float *pdf) {
  if (!s || !s->committed || s->strategy != LGPU_LIGHTS_TREE) return -3;
  const Ctx c{s, corrected != 0, 0};
  Sampler smp;
  smp.pixel = pixel; smp.sample = sample; smp.seed = 0; smp.dim = 0;
  float pd = 1;
  const LightPick lp = SampleLightAt(c, V3(p[0], p[1], p[2]), V3(n[0], n[1], n[2]), smp, &pd);
  *pdf = pd;
  if (lp.kind == 1) return lp.index;
  return lp.kind == 0 ? -1 : -2;
}

// Camera::GenerateRay for n image points (u, v) (test hook): sample i draws its aperture
// dimensions from Philox (pixel, sample i). out: o[3], d[3] per point.
int oracle_generate_ray_n(const lgpu_camera *cam, uint32_t n, const float *uv, uint32_t pixel, int corrected, float *out) {
  if (!cam) return -1;
  for (uint32_t i = 0; i < n; ++i) {
    Sampler smp;
    smp.pixel = pixel; smp.sample = i; smp.dim = 0; smp.seed = 0;
    V3 o, d;
    GenerateRay(*cam, uv[2 * (size_t)i], uv[2 * (size_t)i + 1], smp, &o, &d, corrected != 0);
    float *r = out + 6 * (size_t)i;
    r[0] = o.x; r[1] = o.y; r[2] = o.z; r[3] = d.x; r[4] = d.y; r[5] = d.z;
  }
  return 0;
}
// Light::Sample_Li of scene light `light` at point p with normal n (world frame = shading frame
// with normal n), n_samples draws from Philox (pixel, sample i): out wi[3], pdf, Li[4] per draw.
int oracle_light_sample_li_n(oracle_scene s, uint32_t light, const float p[3], const float nrm[3], uint32_t count,
                             int mode, const float *lambda, int corrected, uint32_t pixel, float *out) {
  if (!s || !s->committed || light >= s->lights.size()) return -1;
  const Ctx c{s, corrected != 0, mode};
  const V3 N = Norm(V3(nrm[0], nrm[1], nrm[2]));
  const V3 a = std::abs(N.x) > .5f ? V3(0, 1, 0) : V3(1, 0, 0);
  const V3 T = Norm(Cross(a, N)), B = Cross(T, N);
  for (uint32_t i = 0; i < count; ++i) {
    Hit h;
    h.point = V3(p[0], p[1], p[2]);
    h.normalG = N; h.normalS = N; h.tangent = T; h.bitangent = B;
    Sampler smp;
    smp.pixel = pixel; smp.sample = i; smp.dim = 0; smp.seed = 0;
    Event e;
    e.hit = &h;
    e.sampler = &smp;
    e.lanes = MkLanes(mode, lambda);
    e.wo = N;
    e.wi = -N;
    e.SurfaceLocalise();
    float pdf = 0;
    const Spec L = SampleLi(c, s->lights[light], e, smp, pdf);
    float *r = out + 8 * (size_t)i;
    r[0] = e.wi.x; r[1] = e.wi.y; r[2] = e.wi.z; r[3] = pdf;
    for (int k = 0; k < 4; ++k) r[4 + k] = L.c[k];
  }
  return 0;
}
// EnvironmentLight::PDF_Li for n directions (test hook).
int oracle_env_pdf_n(oracle_scene s, uint32_t light, uint32_t n, const float *wi, float *out) {
  if (!s || !s->committed || light >= s->lights.size() || s->lights[light].kind != L_ENV) return -1;
  for (uint32_t i = 0; i < n; ++i) out[i] = EnvPdfLi(s->lights[light], V3(wi + 3 * (size_t)i));
  return 0;
}

// ---------------------------------------------------------------- learned guiding (checker)
// The guiding network restated from its documented layout: 5 inputs in [0, 1] -> 80 frequency
// features (feature f: input f / 16, octave (f % 16) / 2, sin(2^octave pi x) for even f, cos for
// odd) -> 5 layers of 64 ReLU units -> 144 outputs (128 bin logits, the mixing logit at 128).
// Parameters: W1 (64 x 80), W2..W5 (64 x 64), W6 (144 x 64) row-major in layer order, then the
// biases. The objective for a record (inputs, bin k, BSDF pdf p_b, weight w) with s = softmax of
// the 128 logits, q = s_k / (4 pi / 128), c = 0.05 + 0.85 sigmoid(z_128):
//   loss = -sum (w / sum w) log(c q + (1 - c) p_b).
// Two restatements: `bf16` rounds what the device rounds (features, weights, activations and the
// output gradients to BF16, round to nearest even; products accumulated in double and rounded to
// float, where the device accumulates in float); and an exact double-precision one whose
// gradient is checked against finite differences of its loss.
namespace {
constexpr int GIn = 80, GHid = 64, GOut = 144, GBins = 128, GLayers = 6;
constexpr double GBinSr = 4.0 * 3.14159265358979323846 / GBins;
int GRows(int l) { return l == GLayers - 1 ? GOut : GHid; }
int GCols(int l) { return l == 0 ? GIn : GHid; }
int GWOff(int l) { int o = 0; for (int k = 0; k < l; ++k) o += GRows(k) * GCols(k); return o; }
int GBOff(int l) { int o = 0; for (int k = 0; k < l; ++k) o += GRows(k); return o + GWOff(GLayers); }
int GParams() { return GBOff(GLayers); }
float Bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}
double GFeature(const float *s, int f) {
  const int sc = f >> 4, w = f & 15, oct = w >> 1;
  const float t = Fract(s[sc] * (float)(1 << oct) * .5f);
  const float a = t * PI2;
  return (w & 1) ? std::cos(a) : std::sin(a);
}
// Forward of one query. x[l] = input of layer l (x[0]: features). rounded: the device's BF16 steps.
void GForward(const float *P, const float *in, bool rounded, std::vector<double> x[GLayers], std::vector<double> &z) {
  x[0].assign(GIn, 0.0);
  for (int f = 0; f < GIn; ++f) {
    const double v = GFeature(in, f);
    x[0][f] = rounded ? Bf16((float)v) : v;
  }
  for (int l = 0; l < GLayers; ++l) {
    const int R = GRows(l), Cc = GCols(l);
    std::vector<double> y(R);
    for (int i = 0; i < R; ++i) {
      double acc = 0.0;
      for (int k = 0; k < Cc; ++k) {
        const double w = rounded ? Bf16(P[GWOff(l) + i * Cc + k]) : (double)P[GWOff(l) + i * Cc + k];
        acc += w * x[l][k];
      }
      if (rounded) acc = (double)((float)acc + P[GBOff(l) + i]);
      else acc += (double)P[GBOff(l) + i];
      y[i] = acc;
    }
    if (l == GLayers - 1) { z = y; break; }
    x[l + 1].assign(R, 0.0);
    for (int i = 0; i < R; ++i) {
      const double h = y[i] > 0 ? y[i] : 0.0;
      x[l + 1][i] = rounded ? Bf16((float)h) : h;
    }
  }
}
// Gradient of the objective for n records (8 floats each) into g (GParams doubles); returns the loss.
double GGradient(const float *P, uint32_t n, const float *rec, bool rounded, std::vector<double> &g) {
  g.assign(GParams(), 0.0);
  double sw = 0.0;
  for (uint32_t r = 0; r < n; ++r) sw += rec[8 * r + 7];
  const double invW = sw > 0 ? 1.0 / sw : 0.0;
  double loss = 0.0;
  std::vector<double> x[GLayers], z;
  for (uint32_t r = 0; r < n; ++r) {
    const float *R = rec + 8 * (size_t)r;
    const double wn = R[7] * invW;
    if (!(wn > 0)) continue;
    const int kb = (int)R[5];
    const double pb = R[6];
    GForward(P, R, rounded, x, z);
    double zmax = -INFINITY;
    for (int j = 0; j < GBins; ++j) zmax = std::max(zmax, z[j]);
    std::vector<double> e(GBins);
    double sum = 0.0;
    for (int j = 0; j < GBins; ++j) { e[j] = std::exp(z[j] - zmax); sum += e[j]; }
    const double q = e[kb] / sum / GBinSr;
    const double sg = 1.0 / (1.0 + std::exp(-z[GBins]));
    const double c = 0.05 + 0.85 * sg;
    const double pmix = c * q + (1.0 - c) * pb;
    loss += -wn * std::log(pmix);
    const double coef = -wn / pmix;
    std::vector<double> dz(GOut, 0.0);
    for (int j = 0; j < GBins; ++j) dz[j] = coef * c * q * ((j == kb ? 1.0 : 0.0) - e[j] / sum);
    dz[GBins] = coef * (q - pb) * 0.85 * sg * (1.0 - sg);
    if (rounded)
      for (double &v : dz) v = Bf16((float)v);
    for (int l = GLayers - 1; l >= 0; --l) {
      const int Rr = GRows(l), Cc = GCols(l);
      for (int i = 0; i < Rr; ++i) {
        if (dz[i] == 0.0) continue;
        for (int k = 0; k < Cc; ++k) g[GWOff(l) + i * Cc + k] += dz[i] * x[l][k];
        g[GBOff(l) + i] += dz[i];
      }
      if (l == 0) break;
      std::vector<double> dn(Cc, 0.0);
      for (int k = 0; k < Cc; ++k) {
        double acc = 0.0;
        for (int i = 0; i < Rr; ++i) {
          const double w = rounded ? Bf16(P[GWOff(l) + i * Cc + k]) : (double)P[GWOff(l) + i * Cc + k];
          acc += w * dz[i];
        }
        const double v = x[l][k] > 0 ? acc : 0.0;
        dn[k] = rounded ? Bf16((float)v) : v;
      }
      dz = dn;
    }
  }
  return loss;
}
}  // namespace

int oracle_guide_num_params(void) { return GParams(); }
// Bin probabilities of the quantised table (bytes max(1, round(255 exp(z - z_max))), over their
// sum) and the mixing probability, per query: 129 floats.
void oracle_guide_forward(const float *params, uint32_t n, const float *inputs, float *out) {
  std::vector<double> x[GLayers], z;
  for (uint32_t i = 0; i < n; ++i) {
    GForward(params, inputs + 5 * (size_t)i, true, x, z);
    double zmax = -INFINITY;
    for (int j = 0; j < GBins; ++j) zmax = std::max(zmax, z[j]);
    float b[GBins], tot = 0.f;
    for (int j = 0; j < GBins; ++j) {
      const float e = (float)std::exp(z[j] - zmax);
      uint32_t v = (uint32_t)(255.f * e + .5f);
      v = v < 1u ? 1u : (v > 255u ? 255u : v);
      b[j] = (float)v;
      tot += b[j];
    }
    float *o = out + (size_t)(GBins + 1) * i;
    for (int j = 0; j < GBins; ++j) o[j] = b[j] / tot;
    o[GBins] = (float)(0.05 + 0.85 / (1.0 + std::exp(-z[GBins])));
  }
}
// rounded = 1: the device's BF16 arithmetic; 0: exact (double). grad: GParams floats.
float oracle_guide_gradient(const float *params, uint32_t n, const float *records, int rounded, float *grad) {
  std::vector<double> g;
  const double l = GGradient(params, n, records, rounded != 0, g);
  for (int i = 0; i < GParams(); ++i) grad[i] = (float)g[i];
  return (float)l;
}
// Loss of the exact restatement in double (for finite differences), parameters in double.
double oracle_guide_loss64(const double *params, uint32_t n, const float *records) {
  std::vector<float> P(GParams());
  // exact evaluation needs double parameters: run the forward on a double copy
  double sw = 0.0;
  for (uint32_t r = 0; r < n; ++r) sw += records[8 * r + 7];
  const double invW = sw > 0 ? 1.0 / sw : 0.0;
  double loss = 0.0;
  for (uint32_t r = 0; r < n; ++r) {
    const float *R = records + 8 * (size_t)r;
    const double wn = R[7] * invW;
    if (!(wn > 0)) continue;
    std::vector<double> x(GIn);
    for (int f = 0; f < GIn; ++f) x[f] = GFeature(R, f);
    for (int l = 0; l < GLayers; ++l) {
      const int Rr = GRows(l), Cc = GCols(l);
      std::vector<double> y(Rr);
      for (int i = 0; i < Rr; ++i) {
        double acc = params[GBOff(l) + i];
        for (int k = 0; k < Cc; ++k) acc += params[GWOff(l) + i * Cc + k] * x[k];
        y[i] = (l == GLayers - 1) ? acc : (acc > 0 ? acc : 0.0);
      }
      x = y;
    }
    double zmax = -INFINITY;
    for (int j = 0; j < GBins; ++j) zmax = std::max(zmax, x[j]);
    double sum = 0.0;
    for (int j = 0; j < GBins; ++j) sum += std::exp(x[j] - zmax);
    const double q = std::exp(x[(int)R[5]] - zmax) / sum / GBinSr;
    const double c = 0.05 + 0.85 / (1.0 + std::exp(-x[GBins]));
    loss += -wn * std::log(c * q + (1.0 - c) * R[6]);
  }
  return loss;
}

// TreeImportance of one node (test hook): node[] = center xyz, diagonal, axis xyz, thetaO, thetaE,
// power. The node's box is the cube of that diagonal around the center.
float oracle_tree_importance(const float node[10], const float p[3], const float n[3], int corrected) {
  LNode nd;
  const V3 c(node[0], node[1], node[2]);
  const float h = node[3] / (2.f * std::sqrt(3.f));
  nd.bounds.mn = c - V3(h, h, h);
  nd.bounds.mx = c + V3(h, h, h);
  nd.cone.axis = V3(node[4], node[5], node[6]);
  nd.cone.thetaO = node[7];
  nd.cone.thetaE = node[8];
  nd.power = node[9];
  return TreeImportance(&nd, V3(p[0], p[1], p[2]), V3(n[0], n[1], n[2]), corrected != 0);
}

float oracle_hg_sample(float g, const float wo[3], float u0, float u1, int corrected, float wi[3]) {
  V3 w;
  const float p = HGSampleU(g, V3(wo[0], wo[1], wo[2]), &w, V2(u0, u1), corrected != 0);
  wi[0] = w.x; wi[1] = w.y; wi[2] = w.z;
  return p;
}
float oracle_uplift(const float rgb[3], float l) { return Uplift(rgb, l); }
float oracle_hg_p(float cosT, float g) {  // HenyeyGreenstein.cpp:8-12
  const float d = 1 + g * g + 2 * g * cosT;
  return INV_PI4 * (1 - g * g) / (d * std::sqrt(d));
}
uint32_t oracle_piecewise1d_sample(const float *f, uint32_t n, float u, float *pdf) {
  Piecewise1D d(f, n);
  return d.SampleDiscrete(u, pdf);
}

}  // extern "C"
