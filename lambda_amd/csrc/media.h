// Device-side participating media: homogeneous medium (HomogeneousMedium.cpp:12-92), the
// piecewise-constant density grid, Henyey-Greenstein (HenyeyGreenstein.cpp:8-36),
// MediaBoundary::GetMedium (Media.h:53-55) and Scene::IntersectTr (Scene.cpp:45-81).
// Same floating-point operation order as the CPU checker.
#pragma once
#include "dmath.h"
#include "lgpu_internal.h"
#include "shading.h"
#include "trace.h"
#include "vm.h"

namespace lgpu {

struct TraceCtx {  // LDS stack column + global spill of the calling thread
  uint32_t (*stk)[kBlock];
  uint32_t *spill;
};

// std::min / std::max semantics (NaN handling included), as the CPU checker uses them
__device__ inline float StdMin(float a, float b) { return (b < a) ? b : a; }
__device__ inline float StdMax(float a, float b) { return (a < b) ? b : a; }
__device__ inline Spec ExpS(Spec s) {
  Spec r;
  for (int i = 0; i < 4; ++i) r.c[i] = expf(s.c[i]);
  return r;
}
__device__ inline int NumLanes(const VmCtx &vm) { return vm.mode ? 4 : 3; }
__device__ inline Spec SigmaT(const MediumRec &m, const VmCtx &vm) {
  return S4(UpliftV(m.sa[0], m.sa[1], m.sa[2], vm)) + S4(UpliftV(m.ss[0], m.ss[1], m.ss[2], vm));
}
__device__ inline Spec SigmaS(const MediumRec &m, const VmCtx &vm) { return S4(UpliftV(m.ss[0], m.ss[1], m.ss[2], vm)); }
__device__ inline uint32_t GetMedium(const MaterialRec &mat, V3 w, V3 n) { return Dot(w, n) < 0 ? mat.med_in : mat.med_out; }

// ---------------------------------------------------------------- density grid traversal
__device__ inline void AxisSetup(float o, float d, float p, float bmin, float bmax, int n, float &cs, int &cell, int &step,
                                 float &tNext, float &tDelta) {
  cs = (bmax - bmin) / (float)n;
  const int c = (int)floorf((p - bmin) / cs);
  cell = c < 0 ? 0 : c;
  cell = cell < n - 1 ? cell : n - 1;
  if (d > 0) { step = 1; tNext = (bmin + (float)(cell + 1) * cs - o) / d; tDelta = cs / d; }
  else if (d < 0) { step = -1; tNext = (bmin + (float)cell * cs - o) / d; tDelta = -cs / d; }
  else { step = 0; tNext = INFINITY; tDelta = INFINITY; }
}
__device__ inline void SlabClip(float o, float d, float bmin, float bmax, float &tn, float &tf) {
  const float inv = 1.f / d;
  float ta = (bmin - o) * inv, tb = (bmax - o) * inv;
  if (ta > tb) { const float t = ta; ta = tb; tb = t; }
  if (ta > tn) tn = ta;
  if (tb < tf) tf = tb;
}
// Visits the voxels of [t0, t1] along o + t d in order (3-D DDA); fn(rho, ta, tb, voxel) returns
// true to stop (voxel: index into the medium's grid).
template <class F>
__device__ inline void GridWalkV(const DevScene &sc, const MediumRec &m, V3 o, V3 d, float t0, float t1, F &&fn) {
  float tn = t0, tf = t1;
  SlabClip(o.x, d.x, m.bmin[0], m.bmax[0], tn, tf);
  SlabClip(o.y, d.y, m.bmin[1], m.bmax[1], tn, tf);
  SlabClip(o.z, d.z, m.bmin[2], m.bmax[2], tn, tf);
  if (!(tn < tf)) return;
  const int nx = (int)m.nx, ny = (int)m.ny, nz = (int)m.nz;
  const V3 p = o + d * tn;
  float csx, csy, csz, tNx, tNy, tNz, dX, dY, dZ;
  int cx, cy, cz, sx, sy, sz;
  AxisSetup(o.x, d.x, p.x, m.bmin[0], m.bmax[0], nx, csx, cx, sx, tNx, dX);
  AxisSetup(o.y, d.y, p.y, m.bmin[1], m.bmax[1], ny, csy, cy, sy, tNy, dY);
  AxisSetup(o.z, d.z, p.z, m.bmin[2], m.bmax[2], nz, csz, cz, sz, tNz, dZ);
  const float *dens = sc.density + m.dens_off;
  float t = tn;
  const int iters = nx + ny + nz + 3;
  int it = 0;
  // One exit test per step (the stepping below is selects, not branches), so a wave whose lanes
  // leave at different steps carries one divergent branch per voxel instead of one per exit.
  bool go = true;
  while (go && it < iters) {
    ++it;
    const bool bx = tNx <= tNy && tNx <= tNz, by = !bx && tNy <= tNz, bz = !bx && !by;
    const float tnext = bx ? tNx : (by ? tNy : tNz);
    const float tc = StdMax(t, StdMin(tnext, tf));
    const uint32_t vox = ((uint32_t)cz * m.ny + (uint32_t)cy) * m.nx + (uint32_t)cx;
    const float rho = dens[vox];
    const bool stop = fn(rho, t, tc, vox);
    t = tc;
    cx += bx ? sx : 0;
    cy += by ? sy : 0;
    cz += bz ? sz : 0;
    const bool inside = cx >= 0 && cx < nx && cy >= 0 && cy < ny && cz >= 0 && cz < nz;
    tNx = bx ? tNx + dX : tNx;
    tNy = by ? tNy + dY : tNy;
    tNz = bz ? tNz + dZ : tNz;
    go = !stop && t < tf && inside;
  }
  if (sc.voxel_count) atomicAdd(sc.voxel_count, (unsigned long long)it);
}
// GridWalkV without the voxel index: fn(rho, ta, tb).
template <class F>
__device__ inline void GridWalk(const DevScene &sc, const MediumRec &m, V3 o, V3 d, float t0, float t1, F &&fn) {
  GridWalkV(sc, m, o, d, t0, t1, [&](float rho, float ta, float tb, uint32_t) { return fn(rho, ta, tb); });
}
struct DensAcc {
  float D, rho;
};
// Density integral over [0, t] and the density of the voxel holding t.
__device__ inline DensAcc DensityUpTo(const DevScene &sc, const MediumRec &m, V3 o, V3 d, float t) {
  DensAcc a{0.f, 0.f};
  GridWalk(sc, m, o, d, 0.f, t, [&](float r, float ta, float tb) { a.D += r * (tb - ta); a.rho = r; return false; });
  return a;
}

// ---------------------------------------------------------------- Medium interface
// Medium::Tr (HomogeneousMedium.cpp:12-15); grids integrate the density exactly.
__device__ inline Spec MediumTr(const DevScene &sc, const MediumRec &m, V3 o, V3 d, float tFar, const VmCtx &vm) {
  const Spec st = SigmaT(m, vm);
  if (m.type == LGPU_MEDIUM_HOMOGENEOUS) return ExpS(st * -1.f * StdMin(tFar, 3.40282347e38f));
  float acc = 0.f;
  GridWalk(sc, m, o, d, 0.f, tFar, [&](float rho, float ta, float tb) { acc += rho * (tb - ta); return false; });
  return ExpS(st * -1.f * acc);
}

struct MediumSample {
  Spec w;
  float t, pdf;
  bool inter;
  float D, rho;  // equiangular samples: density integral over [0, t] and the density at t
};
// HomogeneousMedium::SampleDistance (HomogeneousMedium.cpp:17-38); grids invert the chosen
// channel's optical depth voxel by voxel. DistWalk is that inversion as a per-voxel step, so a
// walk along the same ray can serve other integrals too (SampleDistanceEquiangular).
struct DistWalk {
  float stc, target, acc = 0.f, t, D = 0.f, rhoAt = 1.f;
  bool found = false;
  // one voxel segment [ta, tb] of density rho; true once the sample is placed
  __device__ __attribute__((always_inline)) bool Step(float rho, float ta, float tb) {
    const float k = stc * rho;
    const float seg = k * (tb - ta);
    if (k > 0 && acc + seg >= target) {
      t = ta + (target - acc) / k;
      D += rho * (t - ta);
      rhoAt = rho;
      found = true;
      return true;
    }
    acc += seg;
    D += rho * (tb - ta);
    return false;
  }
};
__device__ inline float DistanceChannelSigma(const MediumRec &m, float u0, const VmCtx &vm) {
  const Spec st = SigmaT(m, vm);
  const int ns = NumLanes(vm);
  const uint32_t c0 = (uint32_t)(u0 * (float)ns);
  const uint32_t ch = c0 < (uint32_t)ns - 1 ? c0 : (uint32_t)ns - 1;
  return ch == 0 ? st.c[0] : (ch == 1 ? st.c[1] : (ch == 2 ? st.c[2] : st.c[3]));
}
// The sample's weight and pdf from its distance t, the density integral D over [0, t] and the
// density rhoAt at t (grids; homogeneous media use t).
__device__ inline MediumSample DistanceFinish(const MediumRec &m, float t, float D, float rhoAt, float tFarHit,
                                              const VmCtx &vm) {
  const Spec st = SigmaT(m, vm), ss = SigmaS(m, vm);
  const int ns = NumLanes(vm);
  MediumSample r;
  r.inter = t < tFarHit;
  const Spec Tr = m.type == LGPU_MEDIUM_HOMOGENEOUS ? ExpS(st * -1.f * StdMin(t, 3.40282347e38f)) : ExpS(st * -1.f * D);
  const Spec stl = st * rhoAt, ssl = ss * rhoAt;
  const Spec dens = r.inter ? stl * Tr : Tr;
  float pdf = 0.f;
  for (int i = 0; i < ns; ++i) pdf += dens.c[i];
  pdf *= 1 / (float)ns;
  if (pdf == 0) pdf = 1;
  r.t = t;
  r.pdf = pdf;
  r.w = r.inter ? Tr * ssl / pdf : Tr / pdf;
  r.D = D;
  r.rho = rhoAt;
  return r;
}
__device__ inline MediumSample SampleDistance(const DevScene &sc, const MediumRec &m, V3 o, V3 d, float tFarHit, float u0,
                                              float u1, const VmCtx &vm) {
  const float stc = DistanceChannelSigma(m, u0, vm);
  if (m.type == LGPU_MEDIUM_HOMOGENEOUS) {
    const float dist = -logf(1 - u1) / stc;
    return DistanceFinish(m, StdMin(dist, tFarHit), 0.f, 1.f, tFarHit, vm);
  }
  DistWalk w;
  w.stc = stc;
  w.target = -logf(1 - u1);
  w.t = tFarHit;
  GridWalk(sc, m, o, d, 0.f, tFarHit, [&](float rho, float ta, float tb) { return w.Step(rho, ta, tb); });
  return DistanceFinish(m, w.found ? w.t : tFarHit, w.D, w.rhoAt, tFarHit, vm);
}
// HomogeneousMedium::PDFDistance (:40-47): 0 as written (integer 1/nSamples), channel average corrected.
// From the density integral D over [0, t] and the density rho at t (homogeneous: D = t, rho = 1).
__device__ inline float PdfDistanceAt(const MediumRec &m, float D, float rho, const VmCtx &vm, bool corrected) {
  if (!corrected) return 0.f;
  const Spec st = SigmaT(m, vm);
  const Spec Tr = ExpS(st * -1.f * D);
  const int ns = NumLanes(vm);
  float pdf = 0.f;
  for (int i = 0; i < ns; ++i) pdf += st.c[i] * rho * Tr.c[i];
  return pdf * (1 / (float)ns);
}
// HomogeneousMedium::SampleEquiangular (:49-78). faithful: origin at the surface hit, t = D tan(theta)
// taken as the ray distance; corrected: origin ray.o, t = l + D tan(theta). EquiangularSetup places
// the sample (no density needed; false: no sample), EquiangularFinish weights it from the density
// integral over [0, t] and the density at t (DensityUpTo for grids; t and 1 for homogeneous media).
__device__ inline bool EquiangularSetup(V3 o, V3 d, V3 hitPoint, float tFarHit, V3 lightPoint, float eps, bool corrected,
                                        MediumSample &r) {
  const V3 delta = lightPoint - (corrected ? o : hitPoint);
  const float lp = Dot(d, delta);
  const float q = Dot(delta, delta) - lp * lp;
  const float D = sqrtf(corrected ? StdMax(0.f, q) : q);
  if (corrected && !(D > 0)) {
    r.inter = false;
    r.t = tFarHit;
    r.pdf = 0.f;
    r.w = S(0.f);
    r.D = 0.f;
    r.rho = 1.f;
    return false;
  }
  const float thetaA = -atan2f(lp, D);
  const float thetaB = atan2f(tFarHit - lp, D);
  const float th = thetaB + (thetaA - thetaB) * eps;
  const float dist = D * tanf(th);
  const float tc = corrected ? lp + dist : dist;
  const float t = StdMin(tc, tFarHit);
  r.inter = t < tFarHit;
  const float x = corrected ? t - lp : t;
  r.t = t;
  r.pdf = D / ((thetaB - thetaA) * (D * D + x * x));
  return true;
}
__device__ inline void EquiangularFinish(const MediumRec &m, MediumSample &r, float Dn, float rho, const VmCtx &vm) {
  const Spec ss = SigmaS(m, vm);
  r.D = Dn;  // the distance pdf at this t reuses these (PdfDistanceAt)
  r.rho = rho;
  const Spec Tr = ExpS(SigmaT(m, vm) * -1.f * StdMin(Dn, 3.40282347e38f));
  r.w = Tr * (ss * rho) / r.pdf;
}
// Both samples of the MIS volumetric integrator along one ray, with one grid walk: the distance
// sample's inversion over [0, tFarHit] and the equiangular sample's density integral over
// [0, t_e] (t_e <= tFarHit) visit the same voxels in the same order with the same segment bounds
// (a segment reaching past t_e counts up to t_e, as DensityUpTo's walk clipped at t_e does), so both
// give the values of separate walks. The walk ends when the distance sample is placed and t_e passed.
__device__ inline void SampleDistanceEquiangular(const DevScene &sc, const MediumRec &m, V3 o, V3 d, float tFarHit,
                                                 float u0, float u1, V3 hitPoint, V3 lightPoint, float eps,
                                                 const VmCtx &vm, bool corrected, MediumSample &sD, MediumSample &sE) {
  const bool eq = EquiangularSetup(o, d, hitPoint, tFarHit, lightPoint, eps, corrected, sE);
  if (m.type == LGPU_MEDIUM_HOMOGENEOUS) {
    sD = SampleDistance(sc, m, o, d, tFarHit, u0, u1, vm);
    if (eq) EquiangularFinish(m, sE, sE.t, 1.f, vm);
    return;
  }
  DistWalk w;
  w.stc = DistanceChannelSigma(m, u0, vm);
  w.target = -logf(1 - u1);
  w.t = tFarHit;
  const float te = sE.t;
  float De = 0.f, rhoE = 0.f;  // DensityUpTo's start: no voxel before t_e leaves rho 0
  bool eDone = !eq;
  GridWalk(sc, m, o, d, 0.f, tFarHit, [&](float rho, float ta, float tb) {
    if (!eDone) {
      if (ta < te) {
        De += rho * (StdMin(tb, te) - ta);
        rhoE = rho;
      }
      eDone = !(tb < te);
    }
    const bool dDone = w.found || w.Step(rho, ta, tb);
    return dDone && eDone;
  });
  sD = DistanceFinish(m, w.found ? w.t : tFarHit, w.D, w.rhoAt, tFarHit, vm);
  if (eq) EquiangularFinish(m, sE, De, rhoE, vm);
}
// HomogeneousMedium::PDFEquiangular (:80-92)
__device__ inline float PdfEquiangular(V3 o, V3 d, V3 lightPoint, float tFar, float t, bool corrected) {
  const V3 delta = lightPoint - o;
  const float lp = Dot(d, delta);
  const float D = sqrtf(Dot(delta, delta) - lp * lp);
  const float thetaA = -atan2f(lp, D);
  const float thetaB = atan2f(tFar - lp, D);
  const float x = corrected ? t - lp : t;
  return D / ((thetaB - thetaA) * (D * D + x * x));
}

// ---------------------------------------------------------------- Henyey-Greenstein
__device__ inline float PhaseHG(float cosT, float g) {
  const float g2 = g * g;
  const float denom = 1 + g2 + 2 * g * cosT;
  return kInvPi4 * (1 - g2) / (denom * sqrtf(denom));
}
// HenyeyGreenstein::Sample_p; faithful frame from the unnormalised cross(-wo, y), corrected orthonormal.
__device__ inline float HGSample(float g, V3 wo, V3 *wi, float u0, float u1, bool corrected) {
  float cosT;
  if (fabsf(g) < 1e-3f) cosT = 1 - 2 * u0;
  else {
    const float sq = (1 - g * g) / (1 - g + 2 * g * u0);
    cosT = (1 + g * g - sq * sq) / (2 * g);
  }
  const float sinT = sqrtf(StdMax(0.f, 1 - cosT * cosT));
  const float phi = 2 * kPi * u1;
  const V3 woNeg = -wo;
  V3 v1 = Cross(woNeg, v3(0, 1, 0));
  if (corrected) {
    if (Dot(v1, v1) < 1e-12f) v1 = Cross(woNeg, v3(1, 0, 0));
    v1 = Norm(v1);
  }
  const V3 v2 = Cross(woNeg, v1);
  const V3 dd = SphericalDirection(sinT, cosT, phi);
  *wi = v1 * dd.x + woNeg * dd.y + v2 * dd.z;
  return PhaseHG(-cosT, g);
}

// ---------------------------------------------------------------- transmittance rays
// Scene::Intersect + hit processing; h is left untouched on a miss (as the reference's RayHit).
// The one-node-per-step loop (Trace4): these rays are traced inside divergent shading code.
__device__ inline bool ClosestHit(const DevScene &sc, V3 o, V3 d, const TraceCtx &tc, bool corrected, Hit &h, uint32_t &bi) {
  float t, u, v;
  uint32_t prim, b;
  if (!Trace4<false>(sc, o, d, 0.f, INFINITY, &t, &u, &v, &prim, &b, tc.stk, tc.spill)) return false;
  ProcessHit(sc, o, d, t, u, v, b, corrected, h);
  bi = b;
  return true;
}
// Scene::IntersectTr (Scene.cpp:45-81): skips surfaces without BxDF or emission, multiplying the
// transmittance of each crossed medium; maxT < 0: no limit. haveFirst: h/bi already hold the
// closest hit of (o, d). leftAll: the loop ended on a miss.
__device__ inline bool IntersectTr(const DevScene &sc, V3 o, V3 d, uint32_t med, float maxT, Spec *Tr, const VmCtx &vm,
                                   bool corrected, const TraceCtx &tc, Hit &h, uint32_t &bi, bool *leftAll, bool haveFirst) {
  float tFar = 0.f;
  if (leftAll) *leftAll = false;
  for (int k = 0; k < 256; ++k) {
    if (!(k == 0 && haveFirst) && !ClosestHit(sc, o, d, tc, corrected, h, bi)) {
      if (leftAll) *leftAll = true;
      return false;
    }
    tFar += h.tFar;
    const bool terminate = maxT >= 0.f && tFar > maxT;
    if (terminate) {
      const float delta = tFar - maxT;
      tFar -= delta;
      h.tFar -= delta;
    }
    if (med != kNone && Tr) *Tr = *Tr * MediumTr(sc, sc.media[med], o, d, h.tFar, vm);
    const MaterialRec mat = sc.materials[sc.meshes[h.mesh].material];
    if (terminate || mat.lobe != kNone || mat.emission != kNone) {
      h.tFar = tFar;
      return true;
    }
    med = GetMedium(mat, d, h.normalG);
    o = h.point + h.normalG * (Dot(h.normalG, d) < 0 ? -kSurfaceEpsilon : kSurfaceEpsilon);
  }
  return false;
}

// How an NEE ray is resolved inside the shading kernels: plain occlusion (Scene::MutualVisibility /
// RayEscapes) or, with volumes, an IntersectTr walk carrying the transmittance (Light.cpp:10-49,
// hasVolumes branch; a ray that leaves every boundary is unoccluded).
// The last transmittance query a VisCtx resolved (the adjoint render walks it again to scatter
// the gradient of its transmittance).
struct VisQuery {
  V3 o, d;
  float maxT;      // IntersectTr's limit (< 0: none)
  uint32_t med;    // start medium
  uint32_t valid;  // 1: a query through media was made
};
struct VisCtx {
  TraceCtx tc;
  bool volumes;
  bool atMedium;              // medium vertex: the ray starts in pathMed, without a surface offset
  uint32_t pathMed;
  const MaterialRec *surf;    // surface vertex: GetMedium(surf, dir, surfNg) gives the start medium
  V3 surfNg;
  VisQuery *rec = nullptr;    // when set, records the query
};
__device__ inline uint32_t NeeMedium(const VisCtx &vc, V3 dir) {
  return vc.atMedium ? vc.pathMed : GetMedium(*vc.surf, dir, vc.surfNg);
}
// Occlusion test of Light::MutualVisibility (point: PointMutualVisibility) from p1 towards dir over
// mag. Tr is multiplied with volumes.
__device__ inline bool Visible(const DevScene &sc, const VisCtx &vc, V3 p1, V3 dir, float mag, bool point, Spec *Tr,
                               const VmCtx &vm) {
  if (!vc.volumes) {
    float t, u, v;
    uint32_t prim, bi;
    const float tmax = (float)((double)mag - .00001);
    return !Trace<true, false>(sc, p1, dir, 0.f, tmax, &t, &u, &v, &prim, &bi, vc.tc.stk, vc.tc.spill, nullptr);
  }
  const float minT = mag - kSurfaceEpsilon * 10;
  Hit h;
  h.tFar = INFINITY;
  uint32_t bi;
  bool left;
  if (vc.rec) *vc.rec = VisQuery{p1, dir, point ? mag : -1.f, NeeMedium(vc, dir), 1u};
  IntersectTr(sc, p1, dir, NeeMedium(vc, dir), point ? mag : -1.f, Tr, vm, true, vc.tc, h, bi, &left, false);
  if (left) return true;
  return h.tFar > minT;
}
__device__ inline bool Escapes(const DevScene &sc, const VisCtx &vc, V3 o, V3 d, Spec *Tr, const VmCtx &vm) {
  if (!vc.volumes) {
    float t, u, v;
    uint32_t prim, bi;
    return !Trace<true, false>(sc, o, d, 0.f, INFINITY, &t, &u, &v, &prim, &bi, vc.tc.stk, vc.tc.spill, nullptr);
  }
  Hit h;
  uint32_t bi;
  if (vc.rec) *vc.rec = VisQuery{o, d, -1.f, NeeMedium(vc, d), 1u};
  return !IntersectTr(sc, o, d, NeeMedium(vc, d), -1.f, Tr, vm, true, vc.tc, h, bi, nullptr, false);
}

}  // namespace lgpu
