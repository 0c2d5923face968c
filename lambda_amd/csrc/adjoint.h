// Path-replay adjoint of the volumetric path integrator (Vicini, Speierer and Jakob 2021, "Path
// replay backpropagation"): the gradient pieces one shade launch scatters while it replays a path
// with the same random numbers as the primal render.
//
// Sampling is detached: directions, distances, light picks, MIS weights and Russian roulette are
// constants of the replayed path, so a contribution C that carries a factor phi(theta) adds
// delta . C dlog(phi)/dtheta to the gradient of theta (delta: the loss gradient of the path's
// pixel over its sample count). The contributions through one factor are known in the launch that
// evaluates it:
//   rest_j  = L_total - L_before_launch_j            (everything launch j and later gathers)
//   BSDF value at the sampled direction of vertex j:  rest_j - (what launch j gathers)
//   BSDF value at the NEE direction:                  that NEE contribution
//   distance-sampling weight of segment j:            rest_j - (launch j's light lookup for the
//                                                     previous vertex and camera emission)
//   transmittance of an NEE or light-lookup ray:      that contribution
//   emission intensity of a light:                    every contribution it emits
// The reference has no differentiation: this extends PathIntegrator::Li (PathIntegrator.cpp:15-88)
// and VolumetricPathIntegrator::Li (VolumetricPathIntegrator.cpp:20-162) with their adjoint.
#pragma once
#include "host.h"
#include "lights.h"
#include "media.h"
#include "shading.h"

namespace lgpu {

// Device view of one adjoint render.
struct AdjRefs {
  const float4 *ltot;  // per sample slot: the primal radiance of the path (RGB)
  const float *dimage; // loss gradient per pixel, w * h * 3 (RGB film values)
  float inv_spp;       // 1 / samples per pixel of the render
  float *gtex;         // texture gradients, laid out as DevScene::texels (user textures), or null
  float *gdens;        // density gradients, laid out as DevScene::density, or null
  float *gemit;        // per material: gradient of its emission intensity, or null
};

__device__ inline Spec AdjDelta(const AdjRefs &ar, uint32_t pixel) {
  const float *d = ar.dimage + 3 * (size_t)pixel;
  Spec s;
  s.c[0] = d[0] * ar.inv_spp; s.c[1] = d[1] * ar.inv_spp; s.c[2] = d[2] * ar.inv_spp; s.c[3] = 0.f;
  return s;
}
__device__ inline float Sum3(Spec a) { return a.c[0] + a.c[1] + a.c[2]; }

// Scatters comp[k] (gradient of the lookup's component k, TexBilinear's out[k]) of one bilinear
// lookup at (u, v) to the texels it read, with the lookup's weights.
__device__ inline void TexScatter(const DevScene &sc, float *g, uint32_t tex, float u, float v, const float comp[4]) {
  const uint4 T = sc.textures[tex];
  const uint32_t w = T.x, h = T.y, ch = T.z;
  const float fx = fminf((float)(w - 1), Clampf(Fract(u), 0.f, 1.f) * (float)w);
  const float fy = fminf((float)(h - 1), Clampf(Fract(v), 0.f, 1.f) * (float)h);
  const uint32_t x = (uint32_t)fx, y = (uint32_t)fy;
  const uint32_t x1 = x + 1 < w - 1 ? x + 1 : w - 1, y1 = y + 1 < h - 1 ? y + 1 : h - 1;
  const float tx = fx - (float)x, ty = fy - (float)y;
  const float w00 = (1.f - tx) * (1.f - ty), w10 = tx * (1.f - ty), w01 = (1.f - tx) * ty, w11 = tx * ty;
  float *base = g + T.w;
  const uint32_t nc = ch == 4 ? 4u : 3u;  // out[3] of a 1- or 3-channel texture is the constant 1
  for (uint32_t k = 0; k < nc; ++k) {
    const float gk = comp[k];
    if (gk == 0.f) continue;
    const uint32_t cc = ch == 1 ? 0 : k;
    atomicAdd(base + ((size_t)y * w + x) * ch + cc, gk * w00);
    atomicAdd(base + ((size_t)y * w + x1) * ch + cc, gk * w10);
    atomicAdd(base + ((size_t)y1 * w + x) * ch + cc, gk * w01);
    atomicAdd(base + ((size_t)y1 * w + x1) * ch + cc, gk * w11);
  }
}

// ---------------------------------------------------------------- microfacet derivatives
// d alpha / d roughness of Microfacet.cpp:49-55's polynomial in log(roughness).
__device__ inline float ToAlphaDeriv(float r) {
  if (!(r > 1e-3f)) return 0.f;  // clamped
  const float x = logf(r);
  return (0.819955f + 2.f * 0.1734f * x + 3.f * 0.0171201f * x * x + 4.f * 0.000640711f * x * x * x) / r;
}
// d log D / d alpha, isotropic (MfD with ax = ay = a)
__device__ inline float DLogDDAlpha(bool ggx, V3 wh, float a) {
  const float t = Tan2Theta(wh);
  if (isinf(t)) return 0.f;
  const float s = Cos2Phi(wh) + Sin2Phi(wh);  // MfD's e = s t / a^2
  if (!ggx) return 2.f * s * t / (a * a * a) - 2.f / a;
  const float e = s * t / (a * a);
  return -2.f / a + 4.f * e / (a * (1.f + e));
}
// d Lambda / d alpha, isotropic (MfLambda with ax = ay = a)
__device__ inline float DLambdaDAlpha(bool ggx, V3 w, float a) {
  const float absTan = fabsf(TanTheta(w));
  if (isinf(absTan)) return 0.f;
  const float s = sqrtf(Cos2Phi(w) + Sin2Phi(w));  // MfLambda's alpha = s a
  const float al = s * a;
  if (!ggx) {
    const float x = 1.f / (al * absTan);
    if (x >= 1.6f) return 0.f;
    const float num = 1.f - 1.259f * x + 0.396f * x * x, den = 3.535f * x + 2.181f * x * x;
    const float dL = ((-1.259f + 0.792f * x) * den - num * (3.535f + 4.362f * x)) / (den * den);
    return dL * (-x / al) * s;
  }
  const float a2t2 = (al * absTan) * (al * absTan);
  return al * absTan * absTan / (2.f * sqrtf(1.f + a2t2)) * s;
}

// Derivatives of one leaf lobe's f at ev: gA[c] = d f_c / d albedo_c, gR[c] = d f_c / d roughness
// (the roughness socket's value, before Microfacet.cpp:49-55's mapping when the lobe applies it).
template <bool kPrograms>
__device__ inline void LeafDerivs(const ShaderT<kPrograms> &shd, const Lobe &L, const Ev &ev, Spec *gA, Spec *gR) {
  *gA = S(0.f);
  *gR = S(0.f);
  if (L.kind == kLambertian) {
    *gA = S(kInvPi);
    return;
  }
  if (L.kind != kMicrofacet) return;
  // the factors of ShaderT::LeafF's microfacet case, in its order
  const float cosO = fabsf(ev.woL.y), cosI = fabsf(ev.wiL.y);
  V3 wh = ev.wiL + ev.woL;
  if (cosI == 0 || cosO == 0) return;
  if (wh.x == 0 && wh.y == 0 && wh.z == 0) return;
  wh = Norm(wh);
  const V3 wiL = v3(ev.wiL.x, Clampf(ev.wiL.y, 0.f, 1.f), ev.wiL.z);
  const bool ggx = L.p[0] != 0.f;
  const float cosH = Dot(wiL, wh);
  Spec F;
  if (L.p[1] != 0.f) {
    const Spec eta = shd.Conductor(&L.p[3]), k = shd.Conductor(&L.p[6]);
    F = shd.corrected ? FrConductor(cosH, S(1.f), eta, k) : FrConductor(cosH, eta, S(0.00001f), k);
  } else {
    F = shd.corrected ? S(FrDielectric(cosH, 1.f, L.p[2])) : S(FrDielectric(cosH, L.p[2], 0.00001f));
  }
  float r0, r1;
  if (L.sock[1] == kNone) { r0 = r1 = 0.f; }
  else SocketRoughness<kPrograms>(shd.sc->sockets[L.sock[1]], shd.vm, &r0, &r1);
  const bool mapped = L.p[9] == 0.f;
  const float a = mapped ? ToAlpha(r0) : r0;
  const float D = MfD(ggx, wh, a, a);
  const float lo = MfLambda(ggx, ev.woL, a, a), li = MfLambda(ggx, wiL, a, a);
  const float G = 1.f / (1.f + lo + li);
  const Spec K = F * (D * G / (4 * cosI * cosO));
  *gA = K;
  const float dlog = DLogDDAlpha(ggx, wh, a) -
                     (DLambdaDAlpha(ggx, ev.woL, a) + DLambdaDAlpha(ggx, wiL, a)) / (1.f + lo + li);
  const float dadr = mapped ? ToAlphaDeriv(r0) : 1.f;
  *gR = shd.SockS(L.sock[0]) * K * (dlog * dadr);
}

// Texture-gradient accumulator of one vertex: the lookup components' gradients of each leaf's
// albedo and roughness bindings (a vertex's NEE and continuation terms share its lookups).
struct VertexGrad {
  float alb[2][4];
  float rough[2][4];
  uint32_t leaf[2];  // lobe indices (kNone: unused)
  float mixw[2];
};
template <bool kPrograms>
__device__ inline void VertexGradInit(const ShaderT<kPrograms> &shd, uint32_t root, VertexGrad &vg) {
  for (int l = 0; l < 2; ++l)
    for (int k = 0; k < 4; ++k) { vg.alb[l][k] = 0.f; vg.rough[l][k] = 0.f; }
  const Lobe &L = shd.sc->lobes[root];
  if (L.kind == kMix) {
    const float r = shd.SockR(L.sock[2]);
    vg.leaf[0] = L.child[0]; vg.leaf[1] = L.child[1];
    vg.mixw[0] = 1.f - r; vg.mixw[1] = r;
  } else {
    vg.leaf[0] = root; vg.leaf[1] = kNone;
    vg.mixw[0] = 1.f; vg.mixw[1] = 0.f;
  }
}
// Adds the gradient of contributions X (delta-weighted, carrying f = the lobe tree's value at ev)
// through f.
template <bool kPrograms>
__device__ inline void VertexGradAdd(const ShaderT<kPrograms> &shd, const Ev &ev, Spec X, Spec f, VertexGrad &vg) {
  Spec q;  // X / f per channel
  for (int c = 0; c < 4; ++c) q.c[c] = (c < 3 && f.c[c] != 0.f) ? X.c[c] / f.c[c] : 0.f;
  if (q.c[0] == 0.f && q.c[1] == 0.f && q.c[2] == 0.f) return;
  for (int l = 0; l < 2; ++l) {
    if (vg.leaf[l] == kNone) continue;
    const AdjLobe A = shd.sc->adj_lobes[vg.leaf[l]];
    if (A.albedo_tex == kNone && A.rough_tex == kNone) continue;
    Spec gA, gR;
    LeafDerivs<kPrograms>(shd, shd.sc->lobes[vg.leaf[l]], ev, &gA, &gR);
    if (A.albedo_tex != kNone)
      for (int c = 0; c < 3; ++c) {
        const float s = q.c[c] * gA.c[c] * vg.mixw[l];
        for (int k = 0; k < 4; ++k) vg.alb[l][k] += s * A.albedo_w[c][k];
      }
    if (A.rough_tex != kNone) {
      const float s = (q.c[0] * gR.c[0] + q.c[1] * gR.c[1] + q.c[2] * gR.c[2]) * vg.mixw[l];
      for (int k = 0; k < 4; ++k) vg.rough[l][k] += s * A.rough_w[k];
    }
  }
}
__device__ inline void VertexGradFlush(const DevScene &sc, const AdjRefs &ar, const VmCtx &vm, const VertexGrad &vg) {
  if (!ar.gtex) return;
  for (int l = 0; l < 2; ++l) {
    if (vg.leaf[l] == kNone) continue;
    const AdjLobe A = sc.adj_lobes[vg.leaf[l]];
    if (A.albedo_tex != kNone) TexScatter(sc, ar.gtex, A.albedo_tex, vm.u, vm.v, vg.alb[l]);
    if (A.rough_tex != kNone) TexScatter(sc, ar.gtex, A.rough_tex, vm.u, vm.v, vg.rough[l]);
  }
}

// ---------------------------------------------------------------- media
// d/d density of MediumTr over [0, tFar] of a grid medium: -sigma_t len per voxel, weighted by X.
__device__ inline void TrGradSegment(const DevScene &sc, const AdjRefs &ar, const MediumRec &m, V3 o, V3 d, float tFar,
                                     Spec X, const VmCtx &vm) {
  if (m.type == LGPU_MEDIUM_HOMOGENEOUS || !ar.gdens) return;
  const Spec st = SigmaT(m, vm);
  const float k = -(X.c[0] * st.c[0] + X.c[1] * st.c[1] + X.c[2] * st.c[2]);
  if (k == 0.f) return;
  float *g = ar.gdens + m.dens_off;
  GridWalkV(sc, m, o, d, 0.f, tFar, [&](float, float ta, float tb, uint32_t vox) {
    const float len = tb - ta;
    if (len > 0.f) atomicAdd(g + vox, k * len);
    return false;
  });
}
// The transmittance gradient of an IntersectTr walk (Scene.cpp:45-81) from o along d starting in
// med: the same hits and segments, each crossed grid medium scattering its -sigma_t len terms.
__device__ inline void TrGradWalk(const DevScene &sc, const AdjRefs &ar, V3 o, V3 d, uint32_t med, float maxT, Spec X,
                                  const VmCtx &vm, const TraceCtx &tc) {
  if (!ar.gdens) return;
  float tFar = 0.f;
  Hit h;
  uint32_t bi;
  for (int k = 0; k < 256; ++k) {
    if (!ClosestHit(sc, o, d, tc, true, h, bi)) return;
    tFar += h.tFar;
    const bool terminate = maxT >= 0.f && tFar > maxT;
    if (terminate) {
      const float delta = tFar - maxT;
      tFar -= delta;
      h.tFar -= delta;
    }
    if (med != kNone) TrGradSegment(sc, ar, sc.media[med], o, d, h.tFar, X, vm);
    const MaterialRec mat = sc.materials[sc.meshes[h.mesh].material];
    if (terminate || mat.lobe != kNone || mat.emission != kNone) return;
    med = GetMedium(mat, d, h.normalG);
    o = h.point + h.normalG * (Dot(h.normalG, d) < 0 ? -kSurfaceEpsilon : kSurfaceEpsilon);
  }
}
// Gradient of SampleDistance's weight (Tr sigma_s rho / pdf, or Tr / pdf past the surface) in a
// grid medium: the same voxel walk and stopping voxel; -sigma_t len per voxel, and 1 / rho at the
// interaction's voxel (sigma_s scales with its density).
__device__ inline void DistanceGrad(const DevScene &sc, const AdjRefs &ar, const MediumRec &m, V3 o, V3 d, float tFarHit,
                                    float u0, float u1, bool inter, Spec X, const VmCtx &vm) {
  if (m.type == LGPU_MEDIUM_HOMOGENEOUS || !ar.gdens) return;
  const Spec st = SigmaT(m, vm);
  const int ns = NumLanes(vm);
  const uint32_t c0 = (uint32_t)(u0 * (float)ns);
  const uint32_t ch = c0 < (uint32_t)ns - 1 ? c0 : (uint32_t)ns - 1;
  const float stc = ch == 0 ? st.c[0] : (ch == 1 ? st.c[1] : (ch == 2 ? st.c[2] : st.c[3]));
  const float target = -logf(1 - u1);
  const float kt = -(X.c[0] * st.c[0] + X.c[1] * st.c[1] + X.c[2] * st.c[2]);
  const float xs = Sum3(X);
  float *g = ar.gdens + m.dens_off;
  float acc = 0.f;
  GridWalkV(sc, m, o, d, 0.f, tFarHit, [&](float rho, float ta, float tb, uint32_t vox) {
    const float k = stc * rho;
    const float seg = k * (tb - ta);
    if (k > 0 && acc + seg >= target) {
      const float t = ta + (target - acc) / k;
      float gv = kt * (t - ta);
      if (inter) gv += xs / rho;
      if (gv != 0.f) atomicAdd(g + vox, gv);
      return true;
    }
    acc += seg;
    const float len = tb - ta;
    if (len > 0.f && kt != 0.f) atomicAdd(g + vox, kt * len);
    return false;
  });
}

// ---------------------------------------------------------------- emission
// The material whose emission intensity scales a light's radiance (mesh and triangle lights), or
// kNone.
__device__ inline uint32_t EmitMaterial(const DevScene &sc, uint32_t pick) {
  if (pick == kNone) return kNone;
  const LightRec &L = LightOf(sc, pick);
  if (L.kind != kLightMesh && L.kind != kLightTri) return kNone;
  return sc.meshes[L.mesh].material;
}
// Per-thread emission gradient slots, flushed once per launch with a wave reduction: most lanes
// of a wave name the same few materials.
struct EmitAcc {
  uint32_t mat[2];
  float val[2];
};
__device__ inline void EmitAdd(const DevScene &sc, EmitAcc &e, int slot, uint32_t material, Spec X) {
  if (material == kNone) return;
  const float I = sc.materials[material].intensity;
  if (!(I > 0.f)) return;
  e.mat[slot] = material;
  e.val[slot] += Sum3(X) / I;
}
// All 64 lanes of the wave take part (the caller is converged).
__device__ inline void EmitFlush(const AdjRefs &ar, EmitAcc &e) {
  if (!ar.gemit) return;
  const uint32_t lane = __lane_id();
  for (int s = 0; s < 2; ++s) {
    for (int iter = 0; iter < 64; ++iter) {
      const unsigned long long pend = __ballot(e.mat[s] != kNone);
      if (!pend) break;
      const int leader = __ffsll((long long)pend) - 1;
      const uint32_t M = (uint32_t)__shfl((int)e.mat[s], leader);
      float v = e.mat[s] == M ? e.val[s] : 0.f;
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if ((int)lane == leader) atomicAdd(ar.gemit + M, v);
      if (e.mat[s] == M) e.mat[s] = kNone;
    }
  }
}

}  // namespace lgpu
