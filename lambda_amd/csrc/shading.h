// Device-side shading: hit processing, BxDF lobes, light sampling and the
// light tree. Each function restates the reference function it cites with the
// same floating-point operation order.
#pragma once
#include "dmath.h"
#include "lgpu_internal.h"
#include "vm.h"

// The lobe functions and the Fresnel and light-pdf helpers they call are inlined into the
// shading kernels: out of line, their reference arguments (event, hit, sampler state, the
// shader's VM context, the scene record the shader points at) live in per-lane scratch and are
// written and re-read at every call, and the caller saves its live registers around each call.
// A kernel that defines LGPU_LOBE_INLINE empty before including this header leaves the choice to
// the compiler.
#ifndef LGPU_LOBE_INLINE
#define LGPU_LOBE_INLINE __attribute__((always_inline))
#endif

namespace lgpu {

struct Hit {
  V3 point, normalG, normalS, tangent, bitangent;
  float u, v;
  float tFar;
  uint32_t prim, local, mesh;
};
struct Ev {  // ScatterEvent.h:18-40
  V3 wo, wi, woL, wiL;
  int sidedness;
  float eta;
};
__device__ inline V3 ToLocal(const Hit &h, V3 v) { return WorldToLocal(v, h.tangent, h.normalS, h.bitangent); }
__device__ inline V3 ToWorld(const Hit &h, V3 v) { return LocalToWorld(v, h.tangent, h.normalS, h.bitangent); }

// Object::Hit + TriangleMesh::ProcessHit (Object.h:27-35, TriangleMesh.cpp:80-114).
// `bi` is the BVH-order triangle index recorded by the traversal.
__device__ inline void ProcessHit(const DevScene &sc, V3 o, V3 d, float t, float bu, float bv, uint32_t bi, bool corrected,
                                  Hit &h) {
  const float4 *tp = reinterpret_cast<const float4 *>(sc.tris + bi);
  const float4 w0 = tp[0], w1 = tp[1], w2 = tp[2];
  const uint32_t prim = __float_as_uint(w0.w), mesh = __float_as_uint(w1.w);
  const MeshRec m = sc.meshes[mesh];
  h.tFar = t;
  h.point = v3(o.x + d.x * t, o.y + d.y * t, o.z + d.z * t);
  h.normalG = Norm(Cross(v3(w1.x, w1.y, w1.z), v3(w2.x, w2.y, w2.z)));
  h.prim = prim;
  h.local = prim - m.first_tri;
  h.mesh = mesh;
  const float4 *at = sc.tri_attr + (size_t)bi * kTriAttrF4;
  if (m.flags & kMeshUV) {
    const float4 a0 = at[0], a1 = at[1], a2 = at[2], a6 = at[6];
    h.u = a0.w + (a2.w - a0.w) * bu + (a6.y - a0.w) * bv;
    h.v = a1.w + (a6.x - a1.w) * bu + (a6.z - a1.w) * bv;
  } else {
    h.u = 0.f;
    h.v = 0.f;
  }
  if (m.flags & kMeshSmooth) {
    const V3 n0 = ToV3(at[0]), n1 = ToV3(at[1]), n2 = ToV3(at[2]);
    const float4 t0 = at[3], t1 = at[4], t2 = at[5];
    const V3 T0 = ToV3(t0), T1 = ToV3(t1), T2 = ToV3(t2);
    h.normalS = Norm(n0 + (n1 - n0) * bu + (n2 - n0) * bv);
    h.tangent = Norm(T0 + (T1 - T0) * bu + (T2 - T0) * bv);
    const V3 b0 = Cross(n0, T0) * t0.w, b1 = Cross(n1, T1) * t1.w, b2 = Cross(n2, T2) * t2.w;
    h.bitangent = Norm(b0 + (b1 - b0) * bu + (b2 - b0) * bv);
  } else {
    h.normalS = h.normalG;
    const V3 c1 = Cross(h.normalG, v3(0, 0, 1));
    const V3 c2 = Cross(h.normalG, v3(0, 1, 0));
    h.tangent = (Dot(c1, c1) > Dot(c2, c2) ? c1 : c2);
    if (corrected) h.tangent = Norm(h.tangent);  // unnormalised in the reference (TriangleMesh.cpp:108-111)
    h.bitangent = Cross(h.tangent, h.normalG);
  }
}

// ---------------------------------------------------------------- BxDF helpers (ScatterEvent.h:44-98, Sampling.h)
__device__ inline bool SameHemisphere(V3 a, V3 b) { return a.y * b.y > 0; }
__device__ inline float CosineHemispherePdf(V3 wo, V3 wi) { return SameHemisphere(wo, wi) ? fabsf(wi.y) * kInvPi : 0.f; }
__device__ inline float Cos2Theta(V3 w) { return w.y * w.y; }
__device__ inline float Sin2Theta(V3 w) { return fmaxf(0.f, 1.f - Cos2Theta(w)); }
__device__ inline float SinTheta(V3 w) { return sqrtf(Sin2Theta(w)); }
__device__ inline float TanTheta(V3 w) { return SinTheta(w) / w.y; }
__device__ inline float Tan2Theta(V3 w) { return Sin2Theta(w) / Cos2Theta(w); }
__device__ inline float CosPhi(V3 w) { const float s = SinTheta(w); return (s < 0.000001f) ? 0.f : Clampf(w.x / s, -1.f, 1.f); }
__device__ inline float SinPhi(V3 w) { const float s = SinTheta(w); return (s < 0.000001f) ? 0.f : Clampf(w.z / s, -1.f, 1.f); }
__device__ inline float Cos2Phi(V3 w) { return CosPhi(w) * CosPhi(w); }
__device__ inline float Sin2Phi(V3 w) { return SinPhi(w) * SinPhi(w); }
__device__ inline V3 Reflect(V3 w, V3 n) { return w - n * 2.f * Dot(n, w); }
__device__ inline bool Refract(V3 w, V3 n, float eta, V3 *wr) {
  const float cosI = Dot(w, n);
  const float sin2I = fmaxf(0.f, 1.f - cosI * cosI);
  const float sin2T = eta * eta * sin2I;
  if (sin2T >= 1) { *wr = v3(-1, 1, -1) * w; return false; }
  const float cosT = sqrtf(1 - sin2T);
  *wr = w * -eta + n * (eta * cosI - cosT);
  return true;
}
__device__ inline void SampleUnitDisk(float ux, float uy, float *x, float *y) {
  const float r = sqrtf(ux);
  const float theta = uy * kPi2;
  *x = r * cosf(theta);
  *y = r * sinf(theta);
}
__device__ inline V3 SampleCosineHemisphere(float ux, float uy) {
  float dx, dy;
  SampleUnitDisk(ux, uy, &dx, &dy);
  const float up = sqrtf(fmaxf(0.f, 1.f - dx * dx - dy * dy));
  return v3(dx, up, dy);
}

__device__ LGPU_LOBE_INLINE inline Spec FrConductor(float cosI, Spec etaI, Spec etaT, Spec k) {  // Fresnel.cpp:23-45
  cosI = Clampf(cosI, -1.f, 1.f);
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

// MicrofacetDistribution.cpp:11-71 (ggx = TrowbridgeReitz)
__device__ inline float MfD(bool ggx, V3 wh, float ax, float ay) {
  const float tan2 = Tan2Theta(wh);
  if (isinf(tan2)) return 0.f;
  const float cos4 = Cos2Theta(wh) * Cos2Theta(wh);
  if (!ggx) return expf(-tan2 * (Cos2Phi(wh) / (ax * ax) + Sin2Phi(wh) / (ay * ay))) / (kPi * ax * ay * cos4);
  const float e = (Cos2Phi(wh) / (ax * ax) + Sin2Phi(wh) / (ay * ay)) * tan2;
  return 1.f / (kPi * ax * ay * cos4 * (1 + e) * (1 + e));
}
__device__ inline float MfLambda(bool ggx, V3 w, float ax, float ay) {
  const float absTan = fabsf(TanTheta(w));
  if (isinf(absTan)) return 0.f;
  const float alpha = sqrtf(Cos2Phi(w) * ax * ax + Sin2Phi(w) * ay * ay);
  if (!ggx) {
    const float a = 1 / (alpha * absTan);
    if (a >= 1.6f) return 0.f;
    return (1 - 1.259f * a + 0.396f * a * a) / (3.535f * a + 2.181f * a * a);
  }
  const float a2t2 = (alpha * absTan) * (alpha * absTan);
  return (-1 + sqrtf(1.f + a2t2)) / 2;
}
__device__ inline float MfG(bool ggx, V3 wo, V3 wi, float ax, float ay) {
  return 1.f / (1.f + MfLambda(ggx, wo, ax, ay) + MfLambda(ggx, wi, ax, ay));
}
__device__ inline float MfPdf(bool ggx, V3 wh, float ax, float ay) { return MfD(ggx, wh, ax, ay) * fabsf(wh.y); }
__device__ inline V3 MfSampleWh(bool ggx, float ux, float uy, V3 wo, float ax) {
  if (!ggx) {
    const float lg = ux == 1 ? 0.f : logf(1 - ux);
    const float phi = kPi2 * uy;
    const float tan2 = -ax * ax * lg;
    const float cosT = 1 / sqrtf(1 + tan2);
    const float sinT = sqrtf(fmaxf(0.f, 1 - cosT * cosT));
    const V3 h = SphericalDirection(sinT, cosT, phi);
    return !SameHemisphere(wo, h) ? -h : h;
  }
  const float phi = kPi2 * uy;
  const float tan2 = ax * ax * ux / (1 - ux);
  const float cosT = 1 / sqrtf(1 + tan2);
  const float sinT = sqrtf(fmaxf(0.f, 1.f - cosT * cosT));
  V3 wh = SphericalDirection(sinT, cosT, phi);
  if (!SameHemisphere(wo, wh)) wh = -wh;
  return wh;
}
__device__ inline float ToAlpha(float r) {  // Microfacet.cpp:49-55
  r = fmaxf(r, 1e-3f);
  const float x = logf(r);
  return 1.62142f + 0.819955f * x + 0.1734f * x * x + 0.0171201f * x * x * x + 0.000640711f * x * x * x * x;
}
__device__ inline float OrenNayarTerm(V3 wiL, V3 woL, float sigma) {  // OrenNayar.cpp:10-21
  const float thetaI = acosf(Clampf(wiL.y, -1.f, 1.f));
  const float thetaO = acosf(Clampf(woL.y, -1.f, 1.f));
  const float alpha = fmaxf(thetaI, thetaO);
  const float beta = fminf(thetaI, thetaO);
  const float phiI = acosf(CosPhi(wiL));
  const float phiO = acosf(CosPhi(woL));
  const float s2 = sigma * sigma;
  const float A = 1.f - s2 / (2.f * (s2 + .33f));
  const float B = (.45f * s2) / (s2 + .09f);
  return A + B * fmaxf(0.f, cosf(phiI - phiO) * Clampf(sinf(alpha), -1.f, 1.f) * Clampf(tanf(beta), -1.f, 1.f));
}
// Rough dielectric (project extension: Walter et al. 2007, isotropic GGX)
__device__ inline void RoughDielectricEval(V3 wo, V3 wi, float alpha, float ior, float *fOut, float *pdfOut) {
  *fOut = 0.f;
  *pdfOut = 0.f;
  if (wo.y == 0 || wi.y == 0) return;
  const bool reflect = wo.y * wi.y > 0;
  const float etaI = wo.y > 0 ? 1.f : ior, etaT = wo.y > 0 ? ior : 1.f;
  V3 wh = reflect ? wo + wi : wo * etaI + wi * etaT;
  const float l = Mag(wh);
  if (l == 0) return;
  wh = wh / l;
  if (wh.y < 0) wh = -wh;
  const float dwo = Dot(wo, wh), dwi = Dot(wi, wh);
  // refraction: wo and wi on opposite sides of the microfacet, wo in front of it
  if (!reflect && (dwo * dwi >= 0 || dwo * wo.y <= 0)) return;
  const float F = FrDielectric(dwo, 1.f, ior);
  const float D = MfD(true, wh, alpha, alpha);
  const float G = MfG(true, wo, wi, alpha, alpha);
  const float pwh = D * fabsf(wh.y);
  if (reflect) {
    *fOut = D * G * F / (4.f * fabsf(wo.y) * fabsf(wi.y));
    *pdfOut = pwh * F / (4.f * fabsf(dwo));
  } else {
    const float sq = etaI * dwo + etaT * dwi;
    const float factor = (etaI * etaI) / (etaT * etaT);
    *fOut = (1.f - F) * D * G * etaT * etaT * fabsf(dwi) * fabsf(dwo) / (fabsf(wo.y) * fabsf(wi.y) * sq * sq) * factor;
    *pdfOut = pwh * (1.f - F) * etaT * etaT * fabsf(dwi) / (sq * sq);
  }
}

// ---------------------------------------------------------------- lobes
// kPrograms: as for the socket readers (false only on scenes without program sockets).
template <bool kPrograms = true>
struct ShaderT {
  const DevScene *sc;
  VmCtx vm;
  bool corrected;
  __device__ Spec SockS(uint32_t i) const { return i == kNone ? S(0.f) : SocketSpectrum<kPrograms>(sc->sockets[i], vm); }
  __device__ LGPU_LOBE_INLINE float SockR(uint32_t i) const {
    return i == kNone ? 0.f : SocketReal<kPrograms>(sc->sockets[i], vm);
  }
  __device__ void Alpha(const Lobe &L, float *ax, float *ay) const {
    if (L.sock[1] == kNone) { *ax = *ay = 0.f; }
    else SocketRoughness<kPrograms>(sc->sockets[L.sock[1]], vm, ax, ay);
    if (L.p[9] == 0.f) { *ax = ToAlpha(*ax); *ay = ToAlpha(*ay); }
  }
  __device__ Spec Conductor(const float *rgb) const {
    if (vm.mode) return S4(UpliftV(rgb[0], rgb[1], rgb[2], vm));
    Spec s; s.c[0] = rgb[0]; s.c[1] = rgb[1]; s.c[2] = rgb[2]; s.c[3] = 0.f; return s;
  }

  // BxDF::f for a non-mix lobe
  __device__ LGPU_LOBE_INLINE Spec LeafF(const Lobe &L, const Ev &ev) const {
    switch (L.kind) {
      case kLambertian: return SockS(L.sock[0]) * kInvPi;  // BxDF.cpp:29-31
      case kOrenNayarR:
      case kOrenNayarT: {
        const float oN = OrenNayarTerm(ev.wiL, ev.woL, SockR(L.sock[1]));
        return SockS(L.sock[0]) * kInvPi * oN;
      }
      case kMicrofacet: {  // Microfacet.cpp:12-25
        const float cosO = fabsf(ev.woL.y), cosI = fabsf(ev.wiL.y);
        V3 wh = ev.wiL + ev.woL;
        if (cosI == 0 || cosO == 0) return S(0.f);
        if (wh.x == 0 && wh.y == 0 && wh.z == 0) return S(0.f);
        wh = Norm(wh);
        const V3 wiL = v3(ev.wiL.x, Clampf(ev.wiL.y, 0.f, 1.f), ev.wiL.z);
        const bool ggx = L.p[0] != 0.f;
        const float cosH = Dot(wiL, wh);
        Spec F;
        if (L.p[1] != 0.f) {
          const Spec eta = Conductor(&L.p[3]), k = Conductor(&L.p[6]);
          F = corrected ? FrConductor(cosH, S(1.f), eta, k) : FrConductor(cosH, eta, S(0.00001f), k);
        } else {
          F = corrected ? S(FrDielectric(cosH, 1.f, L.p[2])) : S(FrDielectric(cosH, L.p[2], 0.00001f));
        }
        float ax, ay;
        Alpha(L, &ax, &ay);
        const float D = MfD(ggx, wh, ax, ay);
        const float G = MfG(ggx, ev.woL, wiL, ax, ay);
        return SockS(L.sock[0]) * D * G * F / (4 * cosI * cosO);
      }
      case kRoughDielectric: {
        float f, pdf;
        RoughDielectricEval(ev.woL, ev.wiL, SockR(L.sock[1]), L.p[0], &f, &pdf);
        return SockS(L.sock[0]) * f;
      }
    }
    return S(0.f);
  }
  __device__ LGPU_LOBE_INLINE float LeafPdf(const Lobe &L, V3 wo, V3 wi) const {
    switch (L.kind) {
      case kLambertian: case kOrenNayarR: case kOrenNayarT: return CosineHemispherePdf(wo, wi);
      case kMicrofacet: {
        if (!SameHemisphere(wo, wi)) return 0.f;
        const V3 wh = Norm(wo + wi);
        float ax, ay;
        Alpha(L, &ax, &ay);
        return MfPdf(L.p[0] != 0.f, wh, ax, ay) / (4 * Dot(wo, wh));
      }
      case kRoughDielectric: {
        float f, pdf;
        RoughDielectricEval(wo, wi, SockR(L.sock[1]), L.p[0], &f, &pdf);
        return pdf;
      }
    }
    return 0.f;
  }
  __device__ LGPU_LOBE_INLINE Spec F(uint32_t li, const Ev &ev) const {
    const Lobe &L = sc->lobes[li];
    if (L.kind != kMix) return LeafF(L, ev);
    const float r = SockR(L.sock[2]);  // BxDF.cpp:45-48
    return LeafF(sc->lobes[L.child[0]], ev) * (1.f - r) + LeafF(sc->lobes[L.child[1]], ev) * r;
  }
  __device__ LGPU_LOBE_INLINE float Pdf(uint32_t li, V3 wo, V3 wi) const {
    const Lobe &L = sc->lobes[li];
    if (L.kind != kMix) return LeafPdf(L, wo, wi);
    const float r = SockR(L.sock[2]);  // BxDF.cpp:62-65
    return LeafPdf(sc->lobes[L.child[0]], wo, wi) * (1.f - r) + LeafPdf(sc->lobes[L.child[1]], wo, wi) * r;
  }

  // BxDF::Sample_f for a non-mix lobe; mutates hit point/normal as the reference does.
  __device__ LGPU_LOBE_INLINE Spec LeafSample(const Lobe &L, Ev &ev, Hit &h, Rng &rng, float &pdf, bool &spec) const {
    spec = false;
    switch (L.kind) {
      case kLambertian:
      case kOrenNayarR: {  // BxDF.cpp:7-14
        const float u0 = rng.Get1D(), u1 = rng.Get1D();
        ev.wiL = SampleCosineHemisphere(u0, u1);
        if (ev.woL.y < 0) ev.wiL.y *= -1;
        pdf = CosineHemispherePdf(ev.woL, ev.wiL);
        ev.wi = ToWorld(h, ev.wiL);
        h.point = h.point + h.normalG * kSurfaceEpsilon * (float)ev.sidedness;
        return LeafF(L, ev);
      }
      case kOrenNayarT: {  // OrenNayar.cpp:52-59
        const float u0 = rng.Get1D(), u1 = rng.Get1D();
        ev.wiL = SampleCosineHemisphere(u0, u1);
        pdf = CosineHemispherePdf(ev.woL, ev.wiL);
        ev.wiL.y *= (float)-ev.sidedness;
        ev.wi = ToWorld(h, ev.wiL);
        h.point = h.point + h.normalG * (float)-ev.sidedness * kSurfaceEpsilon;
        return LeafF(L, ev);
      }
      case kMicrofacet: {  // Microfacet.cpp:27-38
        if (ev.woL.y == 0) { pdf = 0.f; return S(0.f); }
        float ax, ay;
        Alpha(L, &ax, &ay);
        const bool ggx = L.p[0] != 0.f;
        const float u0 = rng.Get1D(), u1 = rng.Get1D();
        const V3 wh = Norm(MfSampleWh(ggx, u0, u1, ev.woL, ax));
        ev.wiL = Reflect(-ev.woL, wh);
        if (!SameHemisphere(ev.woL, ev.wiL)) { pdf = 0.f; return S(0.f); }
        pdf = MfPdf(ggx, wh, ax, ay) / (4 * Dot(ev.woL, wh));
        ev.wi = ToWorld(h, ev.wiL);
        h.point = h.point + h.normalG * kSurfaceEpsilon;
        return LeafF(L, ev);
      }
      case kRoughDielectric: {
        const float alpha = SockR(L.sock[1]), ior = L.p[0];
        const float ux = rng.Get1D(), uy = rng.Get1D();
        const float uc = rng.Get1D();
        const float phi = kPi2 * uy;
        const float tan2 = alpha * alpha * ux / (1 - ux);
        const float cosT = 1 / sqrtf(1 + tan2);
        const float sinT = sqrtf(fmaxf(0.f, 1.f - cosT * cosT));
        V3 wh = SphericalDirection(sinT, cosT, phi);
        // Fresnel of wo's side, as RoughDielectricEval takes it (the half vector on the +y side:
        // a negative cosine swaps the indices, so a wo inside gets the glass-to-air reflectance)
        const float Fr = FrDielectric(Dot(ev.woL, wh), 1.f, ior);
        if (ev.woL.y < 0) wh = -wh;
        if (uc < Fr) {
          ev.wiL = Reflect(-ev.woL, wh);
          if (!SameHemisphere(ev.woL, ev.wiL)) { pdf = 0.f; return S(0.f); }
        } else {
          const float etaI = ev.woL.y > 0 ? 1.f : ior, etaT = ev.woL.y > 0 ? ior : 1.f;
          if (!(Dot(ev.woL, wh) > 0)) { pdf = 0.f; return S(0.f); }  // microfacet faces away from wo
          if (!Refract(ev.woL, wh, etaI / etaT, &ev.wiL)) { pdf = 0.f; return S(0.f); }
          if (SameHemisphere(ev.woL, ev.wiL) || ev.wiL.y == 0) { pdf = 0.f; return S(0.f); }
        }
        float f;
        RoughDielectricEval(ev.woL, ev.wiL, alpha, ior, &f, &pdf);
        ev.wi = ToWorld(h, ev.wiL);
        // spawn on the side wi leaves into: reflections stay on wo's side, refractions cross
        h.point = h.point + h.normalG * (ev.wiL.y > 0 ? kSurfaceEpsilon : -kSurfaceEpsilon);
        return SockS(L.sock[0]) * f;
      }
      case kFresnelBsdf: {  // Specular.cpp:18-45
        spec = true;
        const bool entering = ev.woL.y > 0;
        const float ior = SockR(L.sock[1]);
        const float fr = FrDielectric(ev.woL.y, 1, ior);
        if (rng.Get1D() < fr) {
          ev.wiL = v3(-ev.woL.x, ev.woL.y, -ev.woL.z);
          ev.wi = ToWorld(h, ev.wiL);
          const float cosT = fabsf(ev.wiL.y);
          h.point = h.point + h.normalG * (entering ? kSurfaceEpsilon : -kSurfaceEpsilon);
          pdf = fr;
          return SockS(L.sock[0]) * fr / cosT;
        }
        const float etaI = ev.eta = entering ? 1.f : ior;
        const float etaT = entering ? ior : 1.f;
        const V3 nrm = Dot(v3(0, 1, 0), ev.woL) < 0 ? v3(0, -1, 0) : v3(0, 1, 0);
        if (!Refract(ev.woL, nrm, etaI / etaT, &ev.wiL)) { pdf = 0.f; return S(0.f); }
        Spec ft = SockS(L.sock[0]) * (1 - fr);
        ft = ft * ((etaI * etaI) / (etaT * etaT));
        const float cosT = fabsf(ev.wiL.y);
        ev.wi = ToWorld(h, ev.wiL);
        h.point = h.point + h.normalG * (entering ? -kSurfaceEpsilon * 2 : kSurfaceEpsilon * 2);
        pdf = 1.f - fr;
        return ft / cosT;
      }
      case kSpecularR: {  // Specular.cpp:62-68
        spec = true;
        pdf = 1.f;
        ev.wiL = v3(-ev.woL.x, ev.woL.y, -ev.woL.z);
        ev.wi = ToWorld(h, ev.wiL);
        h.point = h.point + h.normalG * kSurfaceEpsilon;
        return SockS(L.sock[0]) / fabsf(ev.wiL.y);
      }
      case kSpecularT: {  // Specular.cpp:76-95
        spec = true;
        pdf = 1.f;
        const V3 woL = ToLocal(h, ev.wo);
        const bool outside = woL.y > 0;
        h.normalS = h.normalS * (outside ? 1.f : -1.f);
        const float etaT0 = L.p[0];
        const float etaI = outside ? ev.eta : etaT0;
        const float etaT = outside ? etaT0 : ev.eta;
        Refract(ev.wo, v3(0, 1, 0) * (outside ? 1.f : -1.f), etaI / etaT, &ev.wiL);
        const float Fd = FrDielectric(fabsf(ev.wi.y), etaT0, etaT);
        Spec ft = SockS(L.sock[0]) * (S(1.f) - S(Fd));
        ft = ft * ((etaI * etaI) / (etaT * etaT));
        const float inv = 1.f / fabsf(ev.wiL.y);
        ev.wi = ToWorld(h, ev.wiL);
        h.point = h.point + h.normalS * .0004f;
        return ft * inv;
      }
      case kGhost: {  // Ghost.cpp:13-28
        spec = true;
        pdf = 1.f;
        const float alpha = SockR(L.sock[0]);
        const bool inside = ev.woL.y < 0;
        if (alpha > 0) {
          ev.wi = -ev.wo;
          ev.wiL = ToLocal(h, ev.wi);
          h.point = h.point + h.normalG * kSurfaceEpsilon * 2.f * (inside ? 1.f : -1.f);
          return S(alpha / fabsf(ev.wiL.y));
        }
        pdf = 0.f;
        return S(0.f);
      }
    }
    pdf = 0.f;
    return S(0.f);
  }
  __device__ LGPU_LOBE_INLINE Spec Sample(uint32_t li, Ev &ev, Hit &h, Rng &rng, float &pdf, bool &spec) const {
    const Lobe &L = sc->lobes[li];
    if (L.kind != kMix) return LeafSample(L, ev, h, rng, pdf, spec);
    const float r = SockR(L.sock[2]);  // BxDF.cpp:50-60
    const Lobe &A = sc->lobes[L.child[0]], &B = sc->lobes[L.child[1]];
    if (rng.Get1D() > r) {
      const Spec af = LeafSample(A, ev, h, rng, pdf, spec);
      if (corrected && !spec) pdf = Pdf(li, ev.woL, ev.wiL);
      else if (corrected) pdf *= (1.f - r);
      return af * (1.f - r) + LeafF(B, ev) * r;
    }
    const Spec bf = LeafSample(B, ev, h, rng, pdf, spec);
    if (corrected && !spec) pdf = Pdf(li, ev.woL, ev.wiL);
    else if (corrected) pdf *= r;
    return bf * r + LeafF(A, ev) * (1.f - r);
  }

  // The direction choice of LeafSample alone, for lobes without a delta component (Lambertian,
  // Oren-Nayar BRDF, microfacet BRDF, rough dielectric): sets ev.wiL and ev.wi, false when the
  // sample fails. The guided shade stage evaluates F and Pdf once for whichever strategy (guide
  // or BSDF) drew the direction, so a wave whose lanes chose differently runs the evaluation once.
  __device__ LGPU_LOBE_INLINE bool LeafSampleDir(const Lobe &L, Ev &ev, const Hit &h, Rng &rng) const {
    switch (L.kind) {
      case kLambertian:
      case kOrenNayarR: {
        const float u0 = rng.Get1D(), u1 = rng.Get1D();
        ev.wiL = SampleCosineHemisphere(u0, u1);
        if (ev.woL.y < 0) ev.wiL.y *= -1;
        break;
      }
      case kMicrofacet: {
        if (ev.woL.y == 0) return false;
        float ax, ay;
        Alpha(L, &ax, &ay);
        const float u0 = rng.Get1D(), u1 = rng.Get1D();
        const V3 wh = Norm(MfSampleWh(L.p[0] != 0.f, u0, u1, ev.woL, ax));
        ev.wiL = Reflect(-ev.woL, wh);
        if (!SameHemisphere(ev.woL, ev.wiL)) return false;
        break;
      }
      case kRoughDielectric: {
        const float alpha = SockR(L.sock[1]), ior = L.p[0];
        const float ux = rng.Get1D(), uy = rng.Get1D();
        const float uc = rng.Get1D();
        const float tan2 = alpha * alpha * ux / (1 - ux);
        const float cosT = 1 / sqrtf(1 + tan2);
        const float sinT = sqrtf(fmaxf(0.f, 1.f - cosT * cosT));
        V3 wh = SphericalDirection(sinT, cosT, kPi2 * uy);
        const float Fr = FrDielectric(Dot(ev.woL, wh), 1.f, ior);
        if (ev.woL.y < 0) wh = -wh;
        if (uc < Fr) {
          ev.wiL = Reflect(-ev.woL, wh);
          if (!SameHemisphere(ev.woL, ev.wiL)) return false;
        } else {
          const float etaI = ev.woL.y > 0 ? 1.f : ior, etaT = ev.woL.y > 0 ? ior : 1.f;
          if (!(Dot(ev.woL, wh) > 0)) return false;
          if (!Refract(ev.woL, wh, etaI / etaT, &ev.wiL)) return false;
          if (SameHemisphere(ev.woL, ev.wiL) || ev.wiL.y == 0) return false;
        }
        break;
      }
      default: return false;
    }
    ev.wi = ToWorld(h, ev.wiL);
    return true;
  }
  // Sample's direction choice for a guidable lobe or a mix of two (the child picked as Sample
  // picks it).
  __device__ LGPU_LOBE_INLINE bool SampleDir(uint32_t li, Ev &ev, const Hit &h, Rng &rng) const {
    const Lobe &L = sc->lobes[li];
    if (L.kind != kMix) return LeafSampleDir(L, ev, h, rng);
    const float r = SockR(L.sock[2]);
    return LeafSampleDir(sc->lobes[L.child[rng.Get1D() > r ? 0 : 1]], ev, h, rng);
  }
};
using Shader = ShaderT<true>;

// ---------------------------------------------------------------- light tree importance (ManyLightSampler.cpp:305-326)
// ManyLightSampler::ImportanceMeasure (ManyLightSampler.cpp:305-326) without inverse
// trigonometry: with c = cos(theta) = delta.axis, tan(thetaU) = r/d and the node's
// cos/sin(thetaO), cos(max(theta - thetaO - thetaU, 0)) = cos(theta - alpha) by the
// angle-difference identity (1 when theta <= alpha), the cone test thetaDash < thetaE
// is cos(thetaDash) > cos(thetaE), and cos(max(thetaI - thetaU, 0)) likewise.
// A |cos| above 1 (acos would give NaN) yields 0 as the reference's NaN comparison does.
// corrected: cone angle from the emitter towards the point and |cos thetaI| (see the CPU checker).
__device__ inline float TreeImportanceFast(const TreeImp &nd, V3 point, V3 normalS, bool corrected) {
  V3 delta = v3(nd.c[0], nd.c[1], nd.c[2]) - point;
  const float d2 = Dot(delta, delta);
  const float d = sqrtf(d2);
  const float invD = 1.f / d;
  delta = delta * invD;
  const float cA0 = Dot(delta, v3(nd.axis[0], nd.axis[1], nd.axis[2]));
  const float cT = corrected ? -cA0 : cA0;
  if (!(fabsf(cT) <= 1.f)) return 0.f;
  const float tU = nd.r * invD;
  const float cU = 1.f / sqrtf(1.f + tU * tU), sU = tU * cU;
  const float cA = nd.cosO * cU - nd.sinO * sU, sA = nd.sinO * cU + nd.cosO * sU;  // alpha = thetaO + thetaU
  float cD = 1.f;
  if (!(sA < 0.f) && cT < cA) cD = cT * cA + sqrtf(fmaxf(0.f, 1.f - cT * cT)) * sA;
  if (!(cD > nd.cosE)) return 0.f;
  const float E = nd.power;
  if (normalS.x == 0.f && normalS.y == 0.f && normalS.z == 0.f) return E / d2 * cD;
  const float cI = Clampf(corrected ? fabsf(Dot(delta, normalS)) : Dot(delta, normalS), -1.f, 1.f);
  float cIU = 1.f;
  if (cI < cU) cIU = cI * cU + sqrtf(fmaxf(0.f, 1.f - cI * cI)) * sU;
  return (cIU * E) / d2 * cD;
}


}  // namespace lgpu
