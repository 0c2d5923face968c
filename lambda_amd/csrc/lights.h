// Device-side light sampling: light selection (power CDF or light tree),
// Light::Sample_Li for each light kind, emitter pdfs and the environment map.
#pragma once
#include "dmath.h"
#include "lgpu_internal.h"
#include "media.h"
#include "shading.h"
#include "vm.h"

namespace lgpu {

// ---------------------------------------------------------------- lights
__device__ inline float PowerHeuristic(float f, float g) { return (f * f) / (f * f + g * g); }  // Integrator.h:22-29
constexpr uint32_t kTreeBit = 0x80000000u;

__device__ inline const LightRec &LightOf(const DevScene &sc, uint32_t pick) {
  return (pick & kTreeBit) ? sc.tlights[pick & ~kTreeBit] : sc.lights[pick];
}

// Copies the leading light-tree records into the block's LDS cache; every thread of the block
// must call it before any of them returns.
__device__ inline void LoadTreeLds(const DevScene &sc, TreeImp *lds) {
  const float4 *src = reinterpret_cast<const float4 *>(sc.timp);
  float4 *dst = reinterpret_cast<float4 *>(lds);
  const uint32_t n4 = sc.tree_lds * 4u;
  for (uint32_t k = threadIdx.x; k < n4; k += blockDim.x) dst[k] = src[k];
  __syncthreads();
}
__device__ inline TreeImp TreeRec(const DevScene &sc, const TreeImp *lds, uint32_t i) {
  return i < sc.tree_lds ? lds[i] : sc.timp[i];
}

// LightSampler::Sample: PowerLightSampler (LightSampler.cpp:14-16) or ManyLightSampler (ManyLightSampler.cpp:17-27).
// lds: the block's cached top of the light tree (LoadTreeLds).
__device__ inline uint32_t SampleLight(const DevScene &sc, V3 point, V3 normalS, Rng &rng, float *pdf, const TreeImp *lds,
                                       bool corrected) {
  if (sc.strategy == LGPU_LIGHTS_POWER) {
    if (!sc.num_lights) { *pdf = 0.f; return kNone; }
    return PwSampleDiscrete(sc.dist, sc.power_dist, rng.Get1D(), pdf);
  }
  const float eps = rng.Get1D();
  const float pTree = sc.tree_power / (sc.inf_power + sc.tree_power);
  if (eps < pTree) {
    *pdf *= pTree;
    float e = rng.Get1D();
    TreeImp nd = TreeRec(sc, lds, 0);
    while (nd.child[0] != kNone) {  // PickLight (ManyLightSampler.cpp:355-378)
      const TreeImp l = TreeRec(sc, lds, nd.child[0]), r = TreeRec(sc, lds, nd.child[1]);
      const float IL = TreeImportanceFast(l, point, normalS, corrected), IR = TreeImportanceFast(r, point, normalS, corrected);
      const float PL = (IL + IR == 0.f) ? .5f : IL / (IL + IR);
      const float PR = 1 - PL;
      if (e < PL) { e /= PL; *pdf *= PL; nd = l; }
      else { e = (e - PL) / PR; *pdf *= PR; nd = r; }
    }
    float lp = 1.f;
    const uint32_t i = PwSampleDiscrete(sc.dist, nd.leaf_dist, e, &lp);
    *pdf *= lp;
    return kTreeBit | sc.tree_order[nd.first + i];
  }
  if (!sc.has_env) { *pdf = 0.f; return kNone; }
  *pdf *= 1 - pTree;
  return sc.env_light;
}

// LightSampler::Pdf for the light hit by a BSDF sample (ManyLightSampler.cpp:29-35, 380-393; LightSampler.cpp:18-20)
__device__ LGPU_LOBE_INLINE inline float LightPdf(const DevScene &sc, uint32_t pick, V3 point, V3 normalS, bool corrected) {
  if (sc.strategy == LGPU_LIGHTS_POWER) return sc.lights[pick].power * sc.inv_total_power;
  if (!(pick & kTreeBit)) return sc.inf_power / (sc.inf_power + sc.tree_power);
  const uint32_t tl = pick & ~kTreeBit;
  const uint32_t leaf = sc.tlight_leaf[tl], slot = sc.tlight_slot[tl];
  const uint32_t off = sc.tree[leaf].leaf_dist;
  float p = corrected ? PwPdfNormalised(sc.dist, off, slot) : PwPdfRaw(sc.dist, off, slot);
  uint32_t child = leaf;
  for (uint32_t n = sc.tree[leaf].parent; n != kNone; child = n, n = sc.tree[n].parent) {
    const TreeNode nd = sc.tree[n];
    const float IL = TreeImportanceFast(sc.timp[sc.tree_bfs[nd.child[0]]], point, normalS, corrected);
    const float IR = TreeImportanceFast(sc.timp[sc.tree_bfs[nd.child[1]]], point, normalS, corrected);
    const float p0 = (IL + IR == 0.f) ? .5f : IL / (IL + IR);
    const float p1 = corrected ? 1.f - p0 : 1.f - IR;  // faithful: RecursivePDF :387
    p *= (child == nd.child[0]) ? p0 : p1;
  }
  if (corrected) p *= sc.tree_power / (sc.inf_power + sc.tree_power);
  return p;
}

__device__ inline Spec EnvLookup(const DevScene &sc, float u, float v, const VmCtx &vm) {
  float c[4];  // TexBilinear also writes out[3] (alpha)
  TexBilinear(sc.texels + sc.env.tex_off, sc.env.w, sc.env.h, 3, Fract(u), Fract(v), c);
  return S4(UpliftV(c[0], c[1], c[2], vm));
}
__device__ inline Spec EnvLe(const DevScene &sc, V3 w, const VmCtx &vm) {  // EnvironmentLight.h:44-47
  const float u = Fract((SphericalPhi(w) - sc.env.offset[0]) * kInvPi2);
  const float v = Fract((SphericalTheta(w) - sc.env.offset[1]) * kInvPi);
  return EnvLookup(sc, u, v, vm) * sc.env.intensity;
}
__device__ inline float Env2DPdf(const DevScene &sc, float u, float v) {  // Piecewise2D::PDF (Piecewise.cpp:70-76)
  const int nu = (int)sc.env.w, nv = (int)sc.env.h;
  int iu = (int)(u * (float)nu), iv = (int)(v * (float)nv);
  iu = iu < 0 ? 0 : (iu > nu - 1 ? nu - 1 : iu);
  iv = iv < 0 ? 0 : (iv > nv - 1 ? nv - 1 : iv);
  return PwPdfRaw(sc.dist, sc.env.cond_off + (uint32_t)iv * sc.env.cond_stride, (uint32_t)iu) / PwIntegral(sc.dist, sc.env.marg_off);
}
__device__ LGPU_LOBE_INLINE inline float EnvPdfLi(const DevScene &sc, V3 wi) {  // EnvironmentLight.cpp:60-67
  const float theta = SphericalTheta(wi) - sc.env.offset[1];
  const float phi = SphericalPhi(wi) - sc.env.offset[0];
  const float sinT = sinf(theta);
  if (sinT == 0) return 0.f;
  return Env2DPdf(sc, Fract(phi * kInvPi2), Fract(theta * kInvPi)) / (2.f * kPi * kPi * sinT);
}
__device__ inline V3 PointInTriangle(V3 v0, V3 v1, V3 v2, float ux, float uy, bool corrected) {
  if (corrected && ux + uy > 1.f) { ux = 1.f - ux; uy = 1.f - uy; }  // parallelogram in the reference (TriangleMesh.h:73-79)
  return v0 + (v1 - v0) * ux + (v2 - v0) * uy;
}
__device__ inline void TriAreaNormal(V3 v0, V3 v1, V3 v2, float *area, V3 *normal) {  // TriangleMesh.h:60-64
  const V3 c = Cross(v1 - v0, v2 - v0);
  *area = Mag(c) * .5f;
  if (normal) *normal = c / (*area * 2.f);
}
template <bool kPrograms = true>
__device__ inline Spec EmissionOf(const DevScene &sc, uint32_t mesh, const VmCtx &vm) {
  const MaterialRec m = sc.materials[sc.meshes[mesh].material];
  if (m.emission == kNone) return S(0.f);
  return SocketSpectrum<kPrograms>(sc.sockets[m.emission], vm);
}
__device__ inline Spec LightColour(const LightRec &L, const VmCtx &vm) {
  if (L.temperature > 0) {
    if (vm.mode) {
      Spec s;
      for (int i = 0; i < 4; ++i) s.c[i] = BlackbodyNormalised(vm.lam[i], L.temperature);
      return s;
    }
    Spec s; s.c[0] = L.rgb[0]; s.c[1] = L.rgb[1]; s.c[2] = L.rgb[2]; s.c[3] = 0.f; return s;
  }
  return S4(UpliftV(L.rgb[0], L.rgb[1], L.rgb[2], vm));
}

struct ShadowSpec {
  V3 o, d;
  float tmax;
};

// Light::Sample_Li for the light kinds (MeshLight.cpp:22-41,134-151; PointLight.cpp:18-30;
// Spotlight.cpp:29-42; EnvironmentLight.cpp:24-45). kInline = false: visibility is deferred to
// the shadow stage, `sh` receives the occlusion ray of Light::MutualVisibility / RayEscapes.
// kInline = true: the ray is resolved here through `vc` (volumetric paths), Li carries Tr.
// treeIdx (kInline = false only): the tree-light index of L when L is a tree light; its
// triangle record in sc.tri_geo is then fetched together with L.
template <bool kInline, bool kPrograms = true>
__device__ inline Spec SampleLi(const DevScene &sc, const LightRec &L, const Hit &h, Ev &ev, Rng &rng, const VmCtx &vm,
                                bool corrected, float &pdf, ShadowSpec &sh, const VisCtx *vc, uint32_t treeIdx = kNone) {
  const bool noOff = kInline && vc->volumes && vc->atMedium;  // medium vertex: no surface offset
  Spec Tr = S(1.f);
  const bool geo = !kInline && treeIdx != kNone;
  float4 g0 = make_float4(0.f, 0.f, 0.f, 0.f), g1 = g0, g2 = g0;
  if (geo) {
    const float4 *gp = sc.tri_geo + 3 * (size_t)treeIdx;
    g0 = gp[0]; g1 = gp[1]; g2 = gp[2];
  }
  switch (L.kind) {
    case kLightMesh:
    case kLightTri: {
      const bool fromGeo = geo && L.kind == kLightTri;
      uint32_t firstTri = 0;
      float mlIntensity = g1.w;
      uint32_t tri;
      if (!fromGeo) {
        firstTri = sc.meshes[L.mesh].first_tri;
        mlIntensity = (L.kind == kLightTri ? sc.lights[L.parent] : L).intensity;
      }
      if (L.kind == kLightMesh) tri = PwSampleDiscrete(sc.dist, L.cdf_off, rng.Get1D(), &pdf);
      else tri = L.tri;
      const float ux = rng.Get1D(), uy = rng.Get1D();
      V3 v0, v1, v2;
      if (fromGeo) {
        v0 = ToV3(g0); v1 = ToV3(g1); v2 = ToV3(g2);
      } else {
        const uint4 tv = sc.tri_vtx[firstTri + tri];
        v0 = ToV3(sc.vtx_pos[tv.x]); v1 = ToV3(sc.vtx_pos[tv.y]); v2 = ToV3(sc.vtx_pos[tv.z]);
      }
      const V3 pL = PointInTriangle(v0, v1, v2, ux, uy, corrected);
      const V3 pS = noOff ? h.point : h.point + h.normalG * kSurfaceEpsilon * (float)ev.sidedness;
      const V3 diff = pL - pS;  // Scene::MutualVisibility (Scene.cpp:83-95)
      const float mag = Mag(diff);
      ev.wi = diff / mag;
      sh.o = pS; sh.d = ev.wi; sh.tmax = (float)((double)mag - .00001);
      if (kInline && !Visible(sc, *vc, pS, ev.wi, mag, false, &Tr, vm)) { pdf = 0.f; return S(0.f); }
      float area;
      V3 normal;
      TriAreaNormal(v0, v1, v2, &area, &normal);
      const float dn = Dot(normal, -ev.wi);
      const float cosT = fabsf(dn);
      if (cosT > 0 && !(corrected && dn < 0)) {
        ev.wiL = ToLocal(h, ev.wi);
        if (L.kind == kLightMesh) pdf *= DistSq(h.point, pL) / (cosT * area);
        else if (corrected) pdf = DistSq(h.point, pL) / (cosT * area);
        else pdf = DistSq(h.point, pL) * (cosT * area);  // MeshLight.cpp:145 as written
        Spec e;
        if (fromGeo) {
          const uint32_t em = __float_as_uint(g0.w);
          e = em == kNone ? S(0.f) : SocketSpectrum<kPrograms>(sc.sockets[em], vm);
        } else {
          e = EmissionOf<kPrograms>(sc, L.mesh, vm);
        }
        return e * Tr * mlIntensity * kInvPi;
      }
      pdf = 0.f;
      return S(0.f);
    }
    case kLightPoint:
    case kLightSpot: {
      const V3 p = noOff ? h.point : h.point + h.normalG * kSurfaceEpsilon * (float)ev.sidedness;
      const V3 pos = ToV3(L.pos);
      const V3 diff = pos - p;
      const float mag = Mag(diff);
      ev.wi = diff / mag;
      sh.o = p; sh.d = ev.wi; sh.tmax = (float)((double)mag - .00001);
      if (kInline && !Visible(sc, *vc, p, ev.wi, mag, true, &Tr, vm)) { pdf = 0.f; return S(0.f); }
      pdf = 1.f;
      const float d2 = DistSq(pos, h.point);
      ev.wiL = ToLocal(h, ev.wi);
      Spec e = LightColour(L, vm) * Tr;
      if (L.kind == kLightSpot) {  // Spotlight::Falloff (Spotlight.cpp:21-27)
        const V3 w = -ev.wi;
        const float cosT = Dot(w, ToV3(L.axis));
        float fall;
        if (cosT < L.cos_cone) fall = 0.f;
        else if (cosT > L.cos_falloff) fall = 1.f;
        else { const float dd = (cosT - L.cos_cone) / (L.cos_falloff - L.cos_cone); fall = (dd * dd) * (dd * dd); }
        e = e * fall;
      }
      return e * L.intensity / d2;
    }
    case kLightEnv: {
      const float u0 = rng.Get1D(), u1 = rng.Get1D();
      // Piecewise2D::SampleContinuous (Piecewise.cpp:61-68)
      float p1, p0;
      uint32_t vi;
      const float d1 = PwSampleContinuous(sc.dist, sc.env.marg_off, u1, &p1, &vi);
      const float d0 = PwSampleContinuous(sc.dist, sc.env.cond_off + vi * sc.env.cond_stride, u0, &p0, nullptr);
      pdf = p0 * p1;
      if (pdf == 0) return S(0.f);
      const float theta = d1 * kPi + sc.env.offset[1];
      const float phi = d0 * kPi2 + sc.env.offset[0];
      const float cosT = cosf(theta), sinT = sinf(theta);
      if (sinT == 0) { pdf = 0.f; return S(0.f); }
      ev.wi = SphericalDirection(sinT, cosT, phi);
      sh.o = noOff ? h.point : h.point + h.normalG * kSurfaceEpsilon;
      sh.d = ev.wi;
      sh.tmax = INFINITY;
      if (kInline && !Escapes(sc, *vc, sh.o, ev.wi, &Tr, vm)) { pdf = 0.f; return S(0.f); }
      ev.wiL = ToLocal(h, ev.wi);
      pdf /= 2 * kPi * kPi * sinT;
      return EnvLookup(sc, d0, d1, vm) * Tr * sc.env.intensity;
    }
  }
  pdf = 0.f;
  return S(0.f);
}

// Light::PDF_Li(event) for a hit on an emitter (MeshLight.cpp:56-64,157-164)
__device__ inline float PdfLiHit(const DevScene &sc, const LightRec &L, V3 wi, const Hit &hit, uint32_t bi, bool corrected) {
  if (Dot(wi, hit.normalG) > 0) return 0.f;
  const float distSq = hit.tFar * hit.tFar;
  const float4 *tp = reinterpret_cast<const float4 *>(sc.tris + bi);
  const float4 w1 = tp[1], w2 = tp[2];
  const float area = Mag(Cross(v3(w1.x, w1.y, w1.z), v3(w2.x, w2.y, w2.z))) * .5f;  // TriangleMesh.h:60-64
  const float cosT = fabsf(Dot(hit.normalG, -wi));
  if (L.kind == kLightMesh) {
    const float triPdf = corrected ? PwPdfNormalised(sc.dist, L.cdf_off, hit.local) : PwPdfRaw(sc.dist, L.cdf_off, hit.local);
    return triPdf * distSq / (cosT * area);
  }
  return distSq / (cosT * area);
}

// Light reached by a BSDF- or phase-sampled ray: the MIS half of PathIntegrator.cpp:52-63,
// VolumetricPathIntegrator.cpp:127-139 and MISVolumetricPathIntegrator.cpp:52-66. `h`/`bi` is that
// ray's hit (si: it hit something); refP/refN is the corrected light-selection reference point.
// Returns Le * w; the caller applies f or Tr and 1/pdf. envNoLightPdf: the light technique
// cannot produce the environment (point sampling at medium vertices, corrected).
template <bool kPrograms = true>
__device__ inline Spec MisLight(const DevScene &sc, uint32_t l, float lightDistPdf, V3 refP, V3 refN, bool si, const Hit &h,
                                uint32_t bi, V3 rd, float scatPdf, bool specular, bool envNoLightPdf, VmCtx vm,
                                bool corrected) {
  uint32_t nl = kNone;
  if (si) {
    const uint32_t ml = sc.meshes[h.mesh].mesh_light;
    if (ml != kNone) nl = sc.strategy == LGPU_LIGHTS_TREE ? (kTreeBit | sc.tri_tlight[h.prim]) : ml;
  } else if (sc.has_env) {
    nl = sc.env_light;
  }
  if (nl == kNone) return S(0.f);
  bool same = nl == l;
  if (!corrected && sc.strategy == LGPU_LIGHTS_TREE && (nl & kTreeBit)) same = false;  // MeshLight != TriangleLight
  float ldp = lightDistPdf;
  if (!same) ldp = corrected ? LightPdf(sc, nl, refP, refN, true) : LightPdf(sc, nl, h.point, h.normalS, false);
  float pl;
  Spec Le;
  if (si) {
    const LightRec &NL = LightOf(sc, nl);
    if (!corrected && sc.strategy == LGPU_LIGHTS_TREE) pl = PdfLiHit(sc, sc.lights[NL.parent], rd, h, bi, false);
    else pl = PdfLiHit(sc, NL, rd, h, bi, corrected);
    VmAtHit(vm, h.u, h.v, h.point, h.normalS);
    Le = EmissionOf<kPrograms>(sc, h.mesh, vm) * sc.lights[sc.meshes[h.mesh].mesh_light].intensity * kInvPi;
    if (corrected && Dot(rd, h.normalG) > 0) Le = S(0.f);
  } else {
    pl = envNoLightPdf ? 0.f : EnvPdfLi(sc, rd);
    Le = EnvLe(sc, rd, vm);
  }
  const float lightPDF = ldp * pl;
  const float w = (corrected && specular) ? 1.f : PowerHeuristic(scatPdf, lightPDF);
  if (!IsBlack(Le, vm.mode) && (lightPDF > 0 || corrected)) return Le * w;
  return S(0.f);
}

// PartialLightSample (Light.h:107-113); pdf starts at 1 (left uninitialised by the reference).
struct PartialLS {
  V3 point, normal;
  Spec Le;
  uint32_t light;
  float pdf, selPdf;
};
// Light::SamplePoint (MeshLight.cpp:66-81,166-180; PointLight.cpp:40-44; Spotlight.cpp:50-54;
// EnvironmentLight.cpp:69-74). Faithful writes the sampled uv into the event hit (h, vm) as written.
template <bool kPrograms = true>
__device__ inline Spec SamplePointL(const DevScene &sc, const LightRec &L, Hit &h, Rng &rng, VmCtx &vm, bool corrected,
                                    PartialLS &ls) {
  switch (L.kind) {
    case kLightMesh:
    case kLightTri: {
      const MeshRec m = sc.meshes[L.mesh];
      const LightRec &ML = L.kind == kLightTri ? sc.lights[L.parent] : L;
      uint32_t tri = L.tri;
      float area;
      if (L.kind == kLightMesh) {
        float distPdf;
        tri = PwSampleDiscrete(sc.dist, L.cdf_off, rng.Get1D(), &distPdf);
        const uint4 tv = sc.tri_vtx[m.first_tri + tri];
        TriAreaNormal(ToV3(sc.vtx_pos[tv.x]), ToV3(sc.vtx_pos[tv.y]), ToV3(sc.vtx_pos[tv.z]), &area, &ls.normal);
        ls.pdf *= distPdf / area;
      } else {
        const uint4 tv = sc.tri_vtx[m.first_tri + tri];
        TriAreaNormal(ToV3(sc.vtx_pos[tv.x]), ToV3(sc.vtx_pos[tv.y]), ToV3(sc.vtx_pos[tv.z]), &area, &ls.normal);
        ls.pdf /= area;
      }
      const float ux = rng.Get1D(), uy = rng.Get1D();
      const uint4 tv = sc.tri_vtx[m.first_tri + tri];
      float uu = 0.f, uv = 0.f;
      if (m.flags & kMeshUV) {  // maths::BarycentricInterpolation (maths.h:58-60)
        const float2 a = sc.vtx_uv[tv.x], b = sc.vtx_uv[tv.y], c = sc.vtx_uv[tv.z];
        uu = a.x + (b.x - a.x) * ux + (c.x - a.x) * uy;
        uv = a.y + (b.y - a.y) * ux + (c.y - a.y) * uy;
      }
      ls.point = PointInTriangle(ToV3(sc.vtx_pos[tv.x]), ToV3(sc.vtx_pos[tv.y]), ToV3(sc.vtx_pos[tv.z]), ux, uy, corrected);
      VmCtx ve = vm;
      ve.u = uu;
      ve.v = uv;
      if (!corrected) { h.u = uu; h.v = uv; vm.u = uu; vm.v = uv; }
      return EmissionOf<kPrograms>(sc, L.mesh, ve) * ML.intensity * kInvPi;
    }
    case kLightPoint:
    case kLightSpot:
      ls.point = ToV3(L.pos);
      return LightColour(L, vm) * L.intensity;
    case kLightEnv:
      ls.pdf *= 0.f;
      ls.point = v3(0, 0, 0);
      return S(0.f);
  }
  return S(0.f);
}
// Light::Visibility (MeshLight.cpp:83-92,182-191; PointLight.cpp:46-55; Spotlight.cpp:56-65;
// EnvironmentLight.cpp:76-84): completes ls.pdf and sets ev.wi. `eh` is the event's hit.
__device__ inline Spec VisibilityL(const DevScene &sc, const LightRec &L, V3 sp, Ev &ev, const Hit &eh, PartialLS &ls,
                                   const VmCtx &vm, bool corrected, const VisCtx &vc) {
  Spec Tr = S(1.f);
  switch (L.kind) {
    case kLightMesh:
    case kLightTri: {
      const V3 diff = ls.point - sp;
      const float mag = Mag(diff);
      ev.wi = diff / mag;
      if (Visible(sc, vc, sp, ev.wi, mag, false, &Tr, vm)) {
        const V3 n = (L.kind == kLightTri && !corrected) ? eh.normalG : ls.normal;  // TriangleLight: event normal as written
        const float dn = Dot(n, -ev.wi);
        const float cosT = fabsf(dn);
        if (corrected && !(dn > 0)) break;
        ls.pdf *= DistSq(sp, ls.point) / cosT;
        return Tr;
      }
      break;
    }
    case kLightPoint:
    case kLightSpot: {
      const V3 p = corrected ? sp : sp + eh.normalG * kSurfaceEpsilon * (float)ev.sidedness;
      const V3 diff = ls.point - p;
      const float mag = Mag(diff);
      ev.wi = diff / mag;
      if (Visible(sc, vc, p, ev.wi, mag, true, &Tr, vm)) {
        const float d2 = DistSq(ToV3(L.pos), eh.point);
        if (L.kind == kLightSpot) {
          const float cosT = Dot(-ev.wi, ToV3(L.axis));
          float fall;
          if (cosT < L.cos_cone) fall = 0.f;
          else if (cosT > L.cos_falloff) fall = 1.f;
          else { const float dd = (cosT - L.cos_cone) / (L.cos_falloff - L.cos_cone); fall = (dd * dd) * (dd * dd); }
          return Tr * fall / d2;
        }
        return Tr / d2;
      }
      break;
    }
    case kLightEnv: {
      const V3 wi = Norm(ls.point - sp);
      if (Escapes(sc, vc, sp, wi, &Tr, vm)) return Tr;
      break;
    }
  }
  ls.pdf = 0.f;
  return S(0.f);
}

}  // namespace lgpu
