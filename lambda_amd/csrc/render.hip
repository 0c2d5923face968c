// Wavefront path tracer kernels and the render loop.
//
// One wave = up to `capacity` paths (pixel, sample) in flight. Per bounce:
//   extend  - closest-hit BVH2 traversal of the active queue (LDS stack)
//   shade   - emission + MIS of the previous BSDF sample, NEE light sample,
//             BSDF sample, Russian roulette; appends the next queue and shadow rays
//   shadow  - any-hit traversal; unoccluded rays add their contribution
// then accumulate adds each finished path to the FP64 film.
// Reference: TileRenderers::UniformSpp (Render.cpp:43-64) and
// PathIntegrator::Li (PathIntegrator.cpp:15-88).
#include <chrono>
#include <cstring>

#include "host.h"
#include "shading.h"

namespace lgpu {

constexpr int kBlock = 256;

// ---------------------------------------------------------------- traversal
struct TraceStats {
  uint32_t nodes, tris;
};

// Closest hit (kAny = false) or any hit (kAny = true) against the BVH2.
// The per-thread stack lives in LDS column `lane`; entries beyond kLdsStack
// spill to `spill` (kMaxBvhDepth - kLdsStack words per thread).
template <bool kAny, bool kCount>
__device__ inline bool Trace(const DevScene &sc, V3 o, V3 d, float tmin, float tmax, float *tOut, float *uOut,
                             float *vOut, uint32_t *primOut, uint32_t (*stk)[kBlock], uint32_t *spill, TraceStats *ts) {
  const int lane = threadIdx.x;
  const float ix = 1.f / d.x, iy = 1.f / d.y, iz = 1.f / d.z;
  const float ox = o.x * ix, oy = o.y * iy, oz = o.z * iz;
  float best = tmax;
  uint32_t bestPrim = kNone;
  float bu = 0.f, bv = 0.f;
  int sp = 0;
  uint32_t ref = 0;
  for (;;) {
    if (!(ref & kLeafBit)) {
      const float4 *np = reinterpret_cast<const float4 *>(sc.nodes + ref);
      const float4 a = np[0], b = np[1], c = np[2];
      const uint4 ch = reinterpret_cast<const uint4 *>(np)[3];
      if (kCount) ts->nodes++;
      const float bound = best;
      // child 0 slab test
      float x0 = a.x * ix - ox, x1 = a.y * ix - ox;
      float y0 = a.z * iy - oy, y1 = a.w * iy - oy;
      float z0 = c.x * iz - oz, z1 = c.y * iz - oz;
      const float lo0 = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
      const float hi0 = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), bound));
      x0 = b.x * ix - ox; x1 = b.y * ix - ox;
      y0 = b.z * iy - oy; y1 = b.w * iy - oy;
      z0 = c.z * iz - oz; z1 = c.w * iz - oz;
      const float lo1 = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
      const float hi1 = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), bound));
      const bool h0 = lo0 <= hi0, h1 = lo1 <= hi1;
      if (h0 && h1) {
        const bool first0 = lo0 <= lo1;
        const uint32_t nearRef = first0 ? ch.x : ch.y, farRef = first0 ? ch.y : ch.x;
        if (sp < kLdsStack) stk[sp][lane] = farRef;
        else spill[sp - kLdsStack] = farRef;
        ++sp;
        ref = nearRef;
        continue;
      }
      if (h0) { ref = ch.x; continue; }
      if (h1) { ref = ch.y; continue; }
    } else {
      const uint32_t first = ref & 0x03FFFFFFu, count = (ref >> 26) & 31u;
      for (uint32_t i = first; i < first + count; ++i) {
        const float4 *tp = reinterpret_cast<const float4 *>(sc.tris + i);
        const float4 t0 = tp[0], t1 = tp[1], t2 = tp[2];
        if (kCount) ts->tris++;
        // Moller-Trumbore, same operation order as the CPU checker
        const V3 v0 = v3(t0.x, t0.y, t0.z), e1 = v3(t1.x, t1.y, t1.z), e2 = v3(t2.x, t2.y, t2.z);
        const V3 pv = Cross(d, e2);
        const float det = Dot(e1, pv);
        if (det == 0.f) continue;
        const float inv = 1.f / det;
        const V3 tv = o - v0;
        const float uu = Dot(tv, pv) * inv;
        if (uu < 0.f || uu > 1.f) continue;
        const V3 qv = Cross(tv, e1);
        const float vv = Dot(d, qv) * inv;
        if (vv < 0.f || uu + vv > 1.f) continue;
        const float tt = Dot(e2, qv) * inv;
        if (!(tt > tmin && tt < tmax)) continue;
        const uint32_t prim = __float_as_uint(t0.w);
        if (kAny) return true;
        if (tt < best || (tt == best && prim < bestPrim)) {
          best = tt; bestPrim = prim; bu = uu; bv = vv;
        }
      }
    }
    if (sp == 0) break;
    --sp;
    ref = sp < kLdsStack ? stk[sp][lane] : spill[sp - kLdsStack];
  }
  if (kAny) return false;
  if (bestPrim == kNone) return false;
  *tOut = best; *uOut = bu; *vOut = bv; *primOut = bestPrim;
  return true;
}

// ---------------------------------------------------------------- camera (Camera.cpp:41-71, Render.cpp:53-57)
__device__ inline void GenerateRay(const CameraRec &cam, float u, float v, Rng &rng, bool corrected, V3 *o, V3 *d) {
  const V3 origin = ToV3(cam.origin), xh = ToV3(cam.x_hat), yh = ToV3(cam.y_hat), zh = ToV3(cam.z_hat);
  const float tanFov = cam.tan_fov, tanFov2 = 2.f * cam.tan_fov, aspect = cam.aspect;
  if (cam.type == LGPU_CAMERA_SPHERICAL) {
    const float phi = kPi2 * u + cam.offset_phi;
    const float theta = kPi * v + cam.offset_theta;
    const V3 dd = Norm(SphericalDirection(sinf(theta), cosf(theta), phi));
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
  float ax = 0.f, ay = 0.f;
  if (cam.aperture == LGPU_APERTURE_CIRCULAR) {  // Aperture.cpp:23-27
    const float u0 = rng.Get1D(), u1 = rng.Get1D();
    float dx, dy;
    SampleUnitDisk(u0, u1, &dx, &dy);
    const float s = cam.aperture_size;
    if (corrected) { ax = dx * s; ay = dy * s; }
    else { ax = dx * s - s * .5f; ay = dy * s - s * .5f; }
  } else if (cam.aperture == LGPU_APERTURE_BLADE) {  // Aperture.cpp:10-19
    const float dth = kPi2 / (float)cam.blades;
    const float rnd = kPi2 * rng.Get1D();
    const float tmn = floorf(rnd / dth) * dth;
    const float tmx = tmn + dth;
    const float p1x = cosf(tmn), p1y = sinf(tmn), p2x = cosf(tmx), p2y = sinf(tmx);
    const float u0 = rng.Get1D(), u1 = rng.Get1D();
    ax = (p1x * u0 + p2x * u1) * cam.aperture_size;
    ay = (p1y * u0 + p2y * u1) * cam.aperture_size;
  }
  const V3 oo = origin + xh * ax + yh * ay;
  *o = oo;
  *d = Norm(fp - oo);
}

__global__ void __launch_bounds__(kBlock) k_camera(RenderParams rp, PathState ps, uint32_t n, uint64_t base,
                                                   const uint32_t *listPix, const uint32_t *listSmp, uint32_t *queue) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t pixel, sample;
  if (listPix) {
    pixel = listPix[i];
    sample = listSmp[i];
  } else {
    const uint64_t g = base + i;
    const uint64_t npx = (uint64_t)rp.width * rp.height;
    pixel = (uint32_t)(g % npx);
    sample = rp.spp_begin + (uint32_t)(g / npx);
  }
  Rng rng;
  rng.init(pixel, sample, 0, rp.seed);
  float lambda0 = 0.f;
  if (rp.mode) {
    float lam[4];
    HeroWavelengths(WavelengthU(pixel, sample, rp.seed), lam);
    lambda0 = lam[0];
  }
  const uint32_t x = pixel % rp.width, y = pixel / rp.width;
  const float xi = 1.f / (float)rp.width, yi = 1.f / (float)rp.height;
  const float ju = rng.Get1D();
  const float jv = rng.Get1D();
  const float u = xi * ((float)x + ju - .5f);
  const float v = yi * ((float)y + jv - .5f);
  V3 o, d;
  GenerateRay(rp.cam, u, v, rng, rp.corrected != 0, &o, &d);
  ps.ray_o[i] = ToF4(o, 0.f);
  ps.ray_d[i] = ToF4(d, 0.f);
  ps.beta[i] = make_float4(1.f, 1.f, 1.f, 1.f);
  ps.L[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  ps.ids[i] = make_uint4(pixel, sample, rng.dim, kFlagCamera);
  ps.mis[i] = make_float4(0.f, 0.f, 0.f, lambda0);
  ps.nee_light[i] = kNone;
  ps.prev_p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  ps.prev_n[i] = make_float4(0.f, 0.f, 0.f, 1.001f);  // w: ScatterEvent::eta
  ps.pass[i] = 0;
  queue[i] = i;
}

template <bool kCount>
__global__ void __launch_bounds__(kBlock) k_extend(DevScene sc, PathState ps, const uint32_t *queue, uint32_t n,
                                                   uint32_t *spill, unsigned long long *counts) {
  __shared__ uint32_t stk[kLdsStack][kBlock];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = queue[i];
  const V3 o = ToV3(ps.ray_o[p]), d = ToV3(ps.ray_d[p]);
  float t, u, v;
  uint32_t prim;
  TraceStats ts{0, 0};
  const bool hit = Trace<false, kCount>(sc, o, d, 0.f, INFINITY, &t, &u, &v, &prim, stk,
                                        spill + (size_t)i * (kMaxBvhDepth - kLdsStack), &ts);
  ps.hit[p] = hit ? make_float4(t, u, v, __uint_as_float(prim)) : make_float4(INFINITY, 0.f, 0.f, __uint_as_float(kNone));
  if (kCount) {
    atomicAdd(&counts[0], (unsigned long long)ts.nodes);
    atomicAdd(&counts[1], (unsigned long long)ts.tris);
  }
}

__global__ void __launch_bounds__(kBlock) k_shadow(DevScene sc, PathState ps, const ShadowRay *rays, uint32_t n,
                                                   uint32_t *spill) {
  __shared__ uint32_t stk[kLdsStack][kBlock];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const ShadowRay r = rays[i];
  float t, u, v;
  uint32_t prim;
  const bool occ = Trace<true, false>(sc, ToV3(r.o), ToV3(r.d), 0.f, r.tmax, &t, &u, &v, &prim, stk,
                                      spill + (size_t)i * (kMaxBvhDepth - kLdsStack), nullptr);
  if (!occ) {
    float4 L = ps.L[r.path];
    L.x += r.contrib[0]; L.y += r.contrib[1]; L.z += r.contrib[2]; L.w += r.contrib[3];
    ps.L[r.path] = L;
  }
}

// ---------------------------------------------------------------- lights
__device__ inline float PowerHeuristic(float f, float g) { return (f * f) / (f * f + g * g); }  // Integrator.h:22-29
constexpr uint32_t kTreeBit = 0x80000000u;

__device__ inline const LightRec &LightOf(const DevScene &sc, uint32_t pick) {
  return (pick & kTreeBit) ? sc.tlights[pick & ~kTreeBit] : sc.lights[pick];
}

// LightSampler::Sample: PowerLightSampler (LightSampler.cpp:14-16) or ManyLightSampler (ManyLightSampler.cpp:17-27)
__device__ inline uint32_t SampleLight(const DevScene &sc, V3 point, V3 normalS, Rng &rng, float *pdf) {
  if (sc.strategy == LGPU_LIGHTS_POWER) {
    if (!sc.num_lights) { *pdf = 0.f; return kNone; }
    return PwSampleDiscrete(sc.dist, sc.power_dist, rng.Get1D(), pdf);
  }
  const float eps = rng.Get1D();
  const float pTree = sc.tree_power / (sc.inf_power + sc.tree_power);
  if (eps < pTree) {
    *pdf *= pTree;
    float e = rng.Get1D();
    uint32_t n = 0;
    TreeNode nd = sc.tree[0];
    while (nd.child[0] != kNone) {  // PickLight (ManyLightSampler.cpp:355-378)
      const TreeNode l = sc.tree[nd.child[0]], r = sc.tree[nd.child[1]];
      const float IL = TreeImportance(l, point, normalS), IR = TreeImportance(r, point, normalS);
      const float PL = (IL + IR == 0.f) ? .5f : IL / (IL + IR);
      const float PR = 1 - PL;
      if (e < PL) { e /= PL; *pdf *= PL; n = nd.child[0]; nd = l; }
      else { e = (e - PL) / PR; *pdf *= PR; n = nd.child[1]; nd = r; }
    }
    (void)n;
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
__device__ inline float LightPdf(const DevScene &sc, uint32_t pick, V3 point, V3 normalS, bool corrected) {
  if (sc.strategy == LGPU_LIGHTS_POWER) return sc.lights[pick].power * sc.inv_total_power;
  if (!(pick & kTreeBit)) return sc.inf_power / (sc.inf_power + sc.tree_power);
  const uint32_t tl = pick & ~kTreeBit;
  const uint32_t leaf = sc.tlight_leaf[tl], slot = sc.tlight_slot[tl];
  const uint32_t off = sc.tree[leaf].leaf_dist;
  float p = corrected ? PwPdfNormalised(sc.dist, off, slot) : PwPdfRaw(sc.dist, off, slot);
  uint32_t child = leaf;
  for (uint32_t n = sc.tree[leaf].parent; n != kNone; child = n, n = sc.tree[n].parent) {
    const TreeNode nd = sc.tree[n];
    const float IL = TreeImportance(sc.tree[nd.child[0]], point, normalS);
    const float IR = TreeImportance(sc.tree[nd.child[1]], point, normalS);
    const float p0 = (IL + IR == 0.f) ? .5f : IL / (IL + IR);
    const float p1 = corrected ? 1.f - p0 : 1.f - IR;  // faithful: RecursivePDF :387
    p *= (child == nd.child[0]) ? p0 : p1;
  }
  if (corrected) p *= sc.tree_power / (sc.inf_power + sc.tree_power);
  return p;
}

__device__ inline Spec EnvLookup(const DevScene &sc, float u, float v, const VmCtx &vm) {
  float c[3];
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
__device__ inline float EnvPdfLi(const DevScene &sc, V3 wi) {  // EnvironmentLight.cpp:60-67
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
__device__ inline Spec EmissionOf(const DevScene &sc, uint32_t mesh, const VmCtx &vm) {
  const MaterialRec m = sc.materials[sc.meshes[mesh].material];
  if (m.emission == kNone) return S(0.f);
  return SocketSpectrum(sc.sockets[m.emission], vm);
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
// Spotlight.cpp:29-42; EnvironmentLight.cpp:24-45). Visibility is deferred to the shadow stage:
// `sh` receives the occlusion ray of Light::MutualVisibility / RayEscapes.
__device__ inline Spec SampleLi(const DevScene &sc, const LightRec &L, const Hit &h, Ev &ev, Rng &rng, const VmCtx &vm,
                                bool corrected, float &pdf, ShadowSpec &sh) {
  switch (L.kind) {
    case kLightMesh:
    case kLightTri: {
      const MeshRec m = sc.meshes[L.mesh];
      const LightRec &ML = L.kind == kLightTri ? sc.lights[L.parent] : L;
      uint32_t tri;
      if (L.kind == kLightMesh) tri = PwSampleDiscrete(sc.dist, L.cdf_off, rng.Get1D(), &pdf);
      else tri = L.tri;
      const float ux = rng.Get1D(), uy = rng.Get1D();
      const uint4 tv = sc.tri_vtx[m.first_tri + tri];
      const V3 v0 = ToV3(sc.vtx_pos[tv.x]), v1 = ToV3(sc.vtx_pos[tv.y]), v2 = ToV3(sc.vtx_pos[tv.z]);
      const V3 pL = PointInTriangle(v0, v1, v2, ux, uy, corrected);
      const V3 pS = h.point + h.normalG * kSurfaceEpsilon * (float)ev.sidedness;
      const V3 diff = pL - pS;  // Scene::MutualVisibility (Scene.cpp:83-95)
      const float mag = Mag(diff);
      ev.wi = diff / mag;
      sh.o = pS; sh.d = ev.wi; sh.tmax = (float)((double)mag - .00001);
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
        return EmissionOf(sc, L.mesh, vm) * S(1.f) * ML.intensity * kInvPi;
      }
      pdf = 0.f;
      return S(0.f);
    }
    case kLightPoint:
    case kLightSpot: {
      const V3 p = h.point + h.normalG * kSurfaceEpsilon * (float)ev.sidedness;
      const V3 pos = ToV3(L.pos);
      const V3 diff = pos - p;
      const float mag = Mag(diff);
      ev.wi = diff / mag;
      sh.o = p; sh.d = ev.wi; sh.tmax = (float)((double)mag - .00001);
      pdf = 1.f;
      const float d2 = DistSq(pos, h.point);
      ev.wiL = ToLocal(h, ev.wi);
      Spec e = LightColour(L, vm) * S(1.f);
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
      sh.o = h.point + h.normalG * kSurfaceEpsilon;
      sh.d = ev.wi;
      sh.tmax = INFINITY;
      ev.wiL = ToLocal(h, ev.wi);
      pdf /= 2 * kPi * kPi * sinT;
      return EnvLookup(sc, d0, d1, vm) * S(1.f) * sc.env.intensity;
    }
  }
  pdf = 0.f;
  return S(0.f);
}

// Light::PDF_Li(event) for a hit on an emitter (MeshLight.cpp:56-64,157-164)
__device__ inline float PdfLiHit(const DevScene &sc, const LightRec &L, V3 wi, const Hit &hit, bool corrected) {
  if (Dot(wi, hit.normalG) > 0) return 0.f;
  const float distSq = hit.tFar * hit.tFar;
  const uint4 tv = sc.tri_vtx[hit.prim];
  float area;
  TriAreaNormal(ToV3(sc.vtx_pos[tv.x]), ToV3(sc.vtx_pos[tv.y]), ToV3(sc.vtx_pos[tv.z]), &area, nullptr);
  const float cosT = fabsf(Dot(hit.normalG, -wi));
  if (L.kind == kLightMesh) {
    const float triPdf = corrected ? PwPdfNormalised(sc.dist, L.cdf_off, hit.local) : PwPdfRaw(sc.dist, L.cdf_off, hit.local);
    return triPdf * distSq / (cosT * area);
  }
  return distSq / (cosT * area);
}

// ---------------------------------------------------------------- shade
struct Queues {
  uint32_t *next;
  ShadowRay *shadow;
  uint32_t *counters;  // [0] next count, [1] shadow count
};

__device__ inline void PushNext(const Queues &q, uint32_t p) {
  const uint32_t slot = atomicAdd(&q.counters[0], 1u);
  q.next[slot] = p;
}
__device__ inline void PushShadow(const Queues &q, const ShadowSpec &sh, uint32_t p, Spec c) {
  const uint32_t slot = atomicAdd(&q.counters[1], 1u);
  ShadowRay r;
  r.o[0] = sh.o.x; r.o[1] = sh.o.y; r.o[2] = sh.o.z; r.tmax = sh.tmax;
  r.d[0] = sh.d.x; r.d[1] = sh.d.y; r.d[2] = sh.d.z; r.path = p;
  r.contrib[0] = c.c[0]; r.contrib[1] = c.c[1]; r.contrib[2] = c.c[2]; r.contrib[3] = c.c[3];
  q.shadow[slot] = r;
}

__global__ void __launch_bounds__(kBlock) k_shade(DevScene sc, RenderParams rp, PathState ps, const uint32_t *queue,
                                                  uint32_t n, Queues q) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = queue[i];
  const bool corrected = rp.corrected != 0;
  const int mode = (int)rp.mode;
  uint4 ids = ps.ids[p];
  uint32_t flags = ids.w & 0xFFu;
  uint32_t depth = ids.w >> 8;
  const float4 mis = ps.mis[p];
  Spec L = S4(ps.L[p]);
  Spec beta = S4(ps.beta[p]);
  const V3 ro = ToV3(ps.ray_o[p]), rd = ToV3(ps.ray_d[p]);
  const float4 hr = ps.hit[p];
  const uint32_t prim = __float_as_uint(hr.w);
  const bool hit = prim != kNone;
  Rng rng;
  rng.init(ids.x, ids.y, ids.z, rp.seed);
  VmCtx vm;
  vm.mode = mode;
  vm.texels = sc.texels; vm.textures = sc.textures; vm.consts = sc.consts; vm.code = sc.code;
  if (mode) {
    const float l0 = mis.w;
    for (int k = 0; k < 4; ++k) {
      float l = l0 + 117.5f * (float)k;
      if (l >= 830.f) l -= 470.f;
      vm.lam[k] = l;
    }
  } else {
    vm.lam[0] = vm.lam[1] = vm.lam[2] = vm.lam[3] = 0.f;
  }
  vm.woLy = 0.f;
  Hit h;
  if (hit) ProcessHit(sc, ro, rd, hr.x, hr.y, hr.z, prim, corrected, h);
  vm.u = hit ? h.u : 0.f;
  vm.v = hit ? h.v : 0.f;
  bool alive = true;

  // ---- emission reaching this vertex
  if (flags & kFlagCamera) {
    if (!hit) {
      if (sc.has_env) L = L + EnvLe(sc, rd, vm);  // PathIntegrator.cpp:82-85
      alive = false;
    } else {
      const MeshRec m = sc.meshes[h.mesh];
      if (m.mesh_light != kNone && (!corrected || Dot(rd, h.normalG) < 0))  // :26
        L = L + EmissionOf(sc, h.mesh, vm) * sc.lights[m.mesh_light].intensity * kInvPi;
    }
  } else if (!(flags & kFlagNoEmit)) {
    // BSDF-sampled light with MIS (PathIntegrator.cpp:52-63)
    uint32_t nl = kNone;
    if (hit) {
      const uint32_t ml = sc.meshes[h.mesh].mesh_light;
      if (ml != kNone) nl = sc.strategy == LGPU_LIGHTS_TREE ? (kTreeBit | sc.tri_tlight[prim]) : ml;
    } else if (sc.has_env) {
      nl = sc.env_light;
    }
    if (nl != kNone) {
      const uint32_t l = ps.nee_light[p];
      bool same = nl == l;
      if (!corrected && sc.strategy == LGPU_LIGHTS_TREE && (nl & kTreeBit)) same = false;  // MeshLight != TriangleLight
      float lightDistPdf = mis.y;
      if (!same) {
        if (corrected) lightDistPdf = LightPdf(sc, nl, ToV3(ps.prev_p[p]), ToV3(ps.prev_n[p]), true);
        else lightDistPdf = LightPdf(sc, nl, h.point, h.normalS, false);  // evaluated at the light hit as written
      }
      float pl;
      Spec Le;
      if (hit) {
        const LightRec &NL = LightOf(sc, nl);
        if (!corrected && sc.strategy == LGPU_LIGHTS_TREE) pl = PdfLiHit(sc, sc.lights[NL.parent], rd, h, false);
        else pl = PdfLiHit(sc, NL, rd, h, corrected);
        Le = EmissionOf(sc, h.mesh, vm) * sc.lights[sc.meshes[h.mesh].mesh_light].intensity * kInvPi;
        if (corrected && Dot(rd, h.normalG) > 0) Le = S(0.f);
      } else {
        pl = EnvPdfLi(sc, rd);
        Le = EnvLe(sc, rd, vm);
      }
      const float lightPDF = lightDistPdf * pl;
      const float scatPdf = mis.x;
      const float w = (corrected && (flags & kFlagSpecular)) ? 1.f : PowerHeuristic(scatPdf, lightPDF);
      if (!IsBlack(Le, mode) && (lightPDF > 0 || corrected)) L = L + beta * (Le * w);
    }
    if ((flags & kFlagEmitOnly) || !hit) alive = false;
  } else if (!hit) {
    alive = false;
  }
  if (alive && (flags & kFlagRR)) beta = beta / S(1.f - mis.z);  // PathIntegrator.cpp:79

  if (alive) {
    const MaterialRec mat = sc.materials[sc.meshes[h.mesh].material];
    if (mat.lobe == kNone) {
      // pass-through surface (PathIntegrator.cpp:69-74)
      uint32_t pass = ps.pass[p] + 1;
      if (pass > 1024) {
        alive = false;
      } else {
        const V3 o2 = h.point + h.normalG * (Dot(rd, h.normalG) < 0 ? -kSurfaceEpsilon : kSurfaceEpsilon);
        ps.ray_o[p] = ToF4(o2, 0.f);
        ps.pass[p] = pass;
        flags = (flags & kFlagCamera) ? kFlagCamera : kFlagNoEmit;
        ps.beta[p] = F4(beta);
        ps.L[p] = F4(L);
        ps.ids[p] = make_uint4(ids.x, ids.y, rng.dim, (depth << 8) | flags);
        PushNext(q, p);
        return;
      }
    } else {
      // ---- shading vertex (PathIntegrator.cpp:28-80)
      Shader shd;
      shd.sc = &sc;
      shd.corrected = corrected;
      Ev ev;
      ev.wo = -rd;
      ev.wi = rd;
      ev.eta = ps.prev_n[p].w;
      ev.woL = ToLocal(h, ev.wo);
      ev.wiL = ToLocal(h, ev.wi);
      ev.sidedness = (ev.woL.y > 0) ? 1 : -1;
      vm.woLy = ev.woL.y;
      shd.vm = vm;
      const V3 shadeP = h.point, shadeN = h.normalS;
      float lightDistPdf = 1.f;
      const uint32_t pick = SampleLight(sc, h.point, h.normalS, rng, &lightDistPdf);
      float lightPDF = 0.f;
      Spec Li = S(0.f);
      ShadowSpec sh;
      bool delta = false;
      if (pick != kNone) {
        const LightRec &LL = LightOf(sc, pick);
        delta = LL.kind == kLightPoint || LL.kind == kLightSpot;
        Li = SampleLi(sc, LL, h, ev, rng, vm, corrected, lightPDF, sh);
      }
      lightPDF *= lightDistPdf;
      bool hasNee = false;
      Spec cNee = S(0.f);
      if (lightPDF > 0 && !IsBlack(Li, mode)) {
        const Spec f = shd.F(mat.lobe, ev) * fabsf(ev.wiL.y);
        const float sp = shd.Pdf(mat.lobe, ev.woL, ev.wiL);
        if (!IsBlack(f, mode) && sp > 0) {
          const float w = (corrected && delta) ? 1.f : PowerHeuristic(lightPDF, sp);
          cNee = Li * f * w / lightPDF;
          hasNee = true;
        }
      }
      float scatPdf = 0.f;
      bool spec = false;
      Spec f = shd.Sample(mat.lobe, ev, h, rng, scatPdf, spec);
      f = f * fabsf(ev.wiL.y);
      if (!(scatPdf > 0 && !IsBlack(f, mode))) {
        if (corrected && hasNee) PushShadow(q, sh, p, beta * cNee);  // NEE is dropped by `break` in the reference
        alive = false;
      } else {
        if (hasNee) PushShadow(q, sh, p, beta * cNee);
        beta = beta * (f / scatPdf);
        uint32_t nf = spec ? kFlagSpecular : 0u;
        float qrr = 0.f;
        if (depth > rp.min_bounces) {  // Russian roulette (PathIntegrator.cpp:76-80)
          qrr = fmaxf(.05f, 1 - Luminance(beta, mode));
          if (rng.Get1D() < qrr) nf |= kFlagEmitOnly;
          else nf |= kFlagRR;
        }
        if (depth + 1 >= rp.max_bounces) nf |= kFlagEmitOnly;
        ps.ray_o[p] = ToF4(h.point, 0.f);
        ps.ray_d[p] = ToF4(ev.wi, 0.f);
        ps.beta[p] = F4(beta);
        ps.L[p] = F4(L);
        ps.mis[p] = make_float4(scatPdf, lightDistPdf, qrr, mis.w);
        ps.nee_light[p] = pick;
        ps.prev_p[p] = ToF4(shadeP, 0.f);
        ps.prev_n[p] = ToF4(shadeN, ev.eta);
        ps.ids[p] = make_uint4(ids.x, ids.y, rng.dim, ((depth + 1) << 8) | nf);
        PushNext(q, p);
        return;
      }
    }
  }
  ps.L[p] = F4(L);
  ps.ids[p] = make_uint4(ids.x, ids.y, rng.dim, (depth << 8) | kFlagDead);
}

// Film::AddSample (Film.h:30-33) in FP64; hero wavelengths go through the CIE fit to XYZ then sRGB.
__global__ void __launch_bounds__(kBlock) k_accumulate(RenderParams rp, PathState ps, uint32_t n, double *film,
                                                       uint32_t *counts, float *out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 Lf = ps.L[i];
  const float l0 = ps.mis[i].w;
  if (out) {
    out[5 * (size_t)i + 0] = Lf.x;
    out[5 * (size_t)i + 1] = Lf.y;
    out[5 * (size_t)i + 2] = Lf.z;
    out[5 * (size_t)i + 3] = rp.mode ? Lf.w : 0.f;
    out[5 * (size_t)i + 4] = l0;
    return;
  }
  double rgb[3];
  if (rp.mode) {
    const float Lc[4] = {Lf.x, Lf.y, Lf.z, Lf.w};
    double xyz[3] = {0, 0, 0};
    for (int k = 0; k < 4; ++k) {
      float l = l0 + 117.5f * (float)k;
      if (l >= 830.f) l -= 470.f;
      float cm[3];
      CMF(l, cm);
      for (int c = 0; c < 3; ++c) xyz[c] += (double)Lc[k] * (double)cm[c];
    }
    const double s = 470.0 / 4.0 / kCieYIntegral;
    for (int c = 0; c < 3; ++c) xyz[c] *= s;
    rgb[0] = 3.240479 * xyz[0] - 1.537150 * xyz[1] - 0.498535 * xyz[2];
    rgb[1] = -0.969256 * xyz[0] + 1.875991 * xyz[1] + 0.041556 * xyz[2];
    rgb[2] = 0.055648 * xyz[0] - 0.204043 * xyz[1] + 1.057311 * xyz[2];
  } else {
    rgb[0] = Lf.x; rgb[1] = Lf.y; rgb[2] = Lf.z;
  }
  const uint32_t px = ps.ids[i].x;
  unsafeAtomicAdd(&film[3 * (size_t)px + 0], rgb[0]);
  unsafeAtomicAdd(&film[3 * (size_t)px + 1], rgb[1]);
  unsafeAtomicAdd(&film[3 * (size_t)px + 2], rgb[2]);
  atomicAdd(&counts[px], 1u);
}

__global__ void k_philox(uint32_t n, const uint4 *ctr, uint32_t k0, uint32_t k1, uint4 *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint4 c = ctr[i];
  uint32_t w[4] = {c.x, c.y, c.z, c.w};
  Philox(w, k0, k1);
  out[i] = make_uint4(w[0], w[1], w[2], w[3]);
}

__global__ void __launch_bounds__(kBlock) k_query(DevScene sc, const float4 *rays, uint32_t n, uint4 *hits,
                                                  uint32_t *occ, uint32_t *spill) {
  __shared__ uint32_t stk[kLdsStack][kBlock];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 a = rays[2 * i], b = rays[2 * i + 1];
  float t, u, v;
  uint32_t prim;
  uint32_t *sp = spill + (size_t)i * (kMaxBvhDepth - kLdsStack);
  if (occ) {
    occ[i] = Trace<true, false>(sc, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), a.w, b.w, &t, &u, &v, &prim, stk, sp, nullptr) ? 1u : 0u;
  } else if (Trace<false, false>(sc, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), a.w, b.w, &t, &u, &v, &prim, stk, sp, nullptr)) {
    hits[i] = make_uint4(__float_as_uint(t), __float_as_uint(u), __float_as_uint(v), prim);
  } else {
    hits[i] = make_uint4(__float_as_uint(INFINITY), 0u, 0u, kNone);
  }
}

// ---------------------------------------------------------------- host side
struct Wave {
  DevBuf ray_o, ray_d, beta, L, hit, ids, mis, nee, prev_p, prev_n, pass, q0, q1, shadow, counters, spill, tcount;
  uint32_t cap = 0;
  PathState ps{};
  lgpu_status alloc(uint32_t n) {
    if (n <= cap) return LGPU_OK;
    LGPU_HIP(ray_o.alloc(16ull * n)); LGPU_HIP(ray_d.alloc(16ull * n)); LGPU_HIP(beta.alloc(16ull * n));
    LGPU_HIP(L.alloc(16ull * n)); LGPU_HIP(hit.alloc(16ull * n)); LGPU_HIP(ids.alloc(16ull * n));
    LGPU_HIP(mis.alloc(16ull * n)); LGPU_HIP(nee.alloc(4ull * n)); LGPU_HIP(prev_p.alloc(16ull * n));
    LGPU_HIP(prev_n.alloc(16ull * n)); LGPU_HIP(pass.alloc(4ull * n)); LGPU_HIP(q0.alloc(4ull * n));
    LGPU_HIP(q1.alloc(4ull * n)); LGPU_HIP(shadow.alloc(sizeof(ShadowRay) * (size_t)n));
    LGPU_HIP(counters.alloc(64)); LGPU_HIP(tcount.alloc(64));
    LGPU_HIP(spill.alloc(4ull * (kMaxBvhDepth - kLdsStack) * n));
    cap = n;
    ps.ray_o = ray_o.as<float4>(); ps.ray_d = ray_d.as<float4>(); ps.beta = beta.as<float4>(); ps.L = L.as<float4>();
    ps.hit = hit.as<float4>(); ps.ids = ids.as<uint4>(); ps.mis = mis.as<float4>(); ps.nee_light = nee.as<uint32_t>();
    ps.prev_p = prev_p.as<float4>(); ps.prev_n = prev_n.as<float4>(); ps.pass = pass.as<uint32_t>();
    return LGPU_OK;
  }
};

static inline uint32_t Blocks(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

static lgpu_status ValidateRender(const lgpu_render_desc *d) {
  if (!d || !d->scene) return Fail(LGPU_ERR_INVALID_ARGUMENT, "render: null descriptor or scene");
  if (!d->scene->committed) return Fail(LGPU_ERR_STATE, "render: scene not committed");
  if (d->integrator != LGPU_INTEGRATOR_PATH) return Fail(LGPU_ERR_UNSUPPORTED, "render: only the path integrator is built");
  if (d->spectral_mode > LGPU_SPECTRAL_HERO4 || d->estimator_mode > LGPU_ESTIMATOR_CORRECTED)
    return Fail(LGPU_ERR_INVALID_ARGUMENT, "render: bad spectral or estimator mode");
  if (d->max_bounces == 0) return Fail(LGPU_ERR_INVALID_ARGUMENT, "render: max_bounces must be > 0");
  if (d->camera.type > LGPU_CAMERA_SPHERICAL) return Fail(LGPU_ERR_INVALID_ARGUMENT, "render: bad camera type");
  if (d->camera.aperture == LGPU_APERTURE_BLADE && d->camera.blades == 0)
    return Fail(LGPU_ERR_INVALID_ARGUMENT, "render: blade aperture needs blades > 0");
  return LGPU_OK;
}

static RenderParams MakeParams(const lgpu_render_desc *d, uint32_t w, uint32_t h) {
  RenderParams rp;
  std::memset(&rp, 0, sizeof(rp));
  const lgpu_camera &c = d->camera;
  rp.cam.type = c.type;
  for (int k = 0; k < 3; ++k) {
    rp.cam.origin[k] = c.origin[k]; rp.cam.x_hat[k] = c.x_hat[k]; rp.cam.y_hat[k] = c.y_hat[k]; rp.cam.z_hat[k] = c.z_hat[k];
  }
  rp.cam.tan_fov = c.tan_fov; rp.cam.aspect = c.aspect; rp.cam.focal_length = c.focal_length;
  rp.cam.aperture = c.aperture; rp.cam.aperture_size = c.aperture_size; rp.cam.blades = c.blades;
  rp.cam.offset_phi = c.offset_phi; rp.cam.offset_theta = c.offset_theta;
  rp.width = w; rp.height = h;
  rp.spp_begin = d->spp_begin;
  rp.max_bounces = d->max_bounces; rp.min_bounces = d->min_bounces;
  rp.mode = d->spectral_mode; rp.corrected = d->estimator_mode == LGPU_ESTIMATOR_CORRECTED;
  rp.seed = d->seed; rp.integrator = d->integrator;
  return rp;
}

// Renders `total` paths; list mode when listPix != nullptr (trace_paths).
static lgpu_status RunWaves(const lgpu_render_desc *d, uint32_t w, uint32_t h, uint64_t total, const uint32_t *listPix,
                            const uint32_t *listSmp, double *film, uint32_t *counts, float *out, lgpu_render_stats *stats) {
  lgpu_scene s = d->scene;
  lgpu_device dev = s->dev;
  LGPU_HIP(hipSetDevice(dev->device));
  (void)hipGetLastError();  // clear a sticky error left by an earlier call
  const auto t0 = std::chrono::steady_clock::now();
  uint64_t cap = d->path_capacity ? d->path_capacity : (1u << 22);
  if (cap > total) cap = total;
  if (cap == 0) cap = 1;
  if (!dev->wave) {  // path pool kept on the device object and reused across calls
    dev->wave = new Wave();
    dev->wave_free = [](void *p) { delete static_cast<Wave *>(p); };
  }
  Wave &W = *static_cast<Wave *>(dev->wave);
  lgpu_status st = W.alloc((uint32_t)cap);
  if (st != LGPU_OK) return st;
  const RenderParams rp = MakeParams(d, w, h);
  const bool countTrav = (d->flags & LGPU_RENDER_COUNT_TRAVERSAL) != 0;
  hipStream_t stream = dev->stream;
  hipEvent_t *ev = dev->ev;
  lgpu_render_stats S;
  std::memset(&S, 0, sizeof(S));
  if (countTrav) LGPU_HIP(hipMemsetAsync(W.tcount.ptr, 0, 16, stream));
  for (uint64_t base = 0; base < total; base += cap) {
    const uint32_t n = (uint32_t)std::min<uint64_t>(cap, total - base);
    uint32_t *qa = W.q0.as<uint32_t>(), *qb = W.q1.as<uint32_t>();
    LGPU_HIP(hipEventRecord(ev[0], stream));
    hipLaunchKernelGGL(k_camera, dim3(Blocks(n)), dim3(kBlock), 0, stream, rp, W.ps, n, base,
                       listPix ? listPix + base : nullptr, listSmp ? listSmp + base : nullptr, qa);
    LGPU_HIP(hipGetLastError());
    LGPU_HIP(hipEventRecord(ev[1], stream));
    LGPU_HIP(hipStreamSynchronize(stream));
    float ms;
    LGPU_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    S.ms_camera += ms;
    uint32_t count = n;
    uint32_t iter = 0;
    while (count > 0 && iter < 4 * (d->max_bounces + 2) + 1100) {
      ++iter;
      LGPU_HIP(hipMemsetAsync(W.counters.ptr, 0, 8, stream));
      LGPU_HIP(hipEventRecord(ev[2], stream));
      if (countTrav)
        hipLaunchKernelGGL(k_extend<true>, dim3(Blocks(count)), dim3(kBlock), 0, stream, s->dsc, W.ps, qa, count,
                           W.spill.as<uint32_t>(), W.tcount.as<unsigned long long>());
      else
        hipLaunchKernelGGL(k_extend<false>, dim3(Blocks(count)), dim3(kBlock), 0, stream, s->dsc, W.ps, qa, count,
                           W.spill.as<uint32_t>(), W.tcount.as<unsigned long long>());
      LGPU_HIP(hipGetLastError());
      LGPU_HIP(hipEventRecord(ev[3], stream));
      Queues q{qb, W.shadow.as<ShadowRay>(), W.counters.as<uint32_t>()};
      hipLaunchKernelGGL(k_shade, dim3(Blocks(count)), dim3(kBlock), 0, stream, s->dsc, rp, W.ps, qa, count, q);
      LGPU_HIP(hipGetLastError());
      LGPU_HIP(hipEventRecord(ev[4], stream));
      uint32_t cnt[2];
      LGPU_HIP(hipMemcpyAsync(cnt, W.counters.ptr, 8, hipMemcpyDeviceToHost, stream));
      LGPU_HIP(hipStreamSynchronize(stream));
      S.extension_rays += count;
      S.launches_extend++;
      S.launches_shade++;
      LGPU_HIP(hipEventElapsedTime(&ms, ev[2], ev[3]));
      S.ms_extend += ms;
      LGPU_HIP(hipEventElapsedTime(&ms, ev[3], ev[4]));
      S.ms_shade += ms;
      if (cnt[1]) {
        LGPU_HIP(hipEventRecord(ev[5], stream));
        hipLaunchKernelGGL(k_shadow, dim3(Blocks(cnt[1])), dim3(kBlock), 0, stream, s->dsc, W.ps, W.shadow.as<ShadowRay>(),
                           cnt[1], W.spill.as<uint32_t>());
        LGPU_HIP(hipGetLastError());
        LGPU_HIP(hipEventRecord(ev[6], stream));
        LGPU_HIP(hipEventSynchronize(ev[6]));
        LGPU_HIP(hipEventElapsedTime(&ms, ev[5], ev[6]));
        S.ms_shadow += ms;
        S.shadow_rays += cnt[1];
        S.launches_shadow++;
      }
      count = cnt[0];
      std::swap(qa, qb);
    }
    if (count > 0) return Fail(LGPU_ERR_STATE, "render: path wave did not terminate");
    hipLaunchKernelGGL(k_accumulate, dim3(Blocks(n)), dim3(kBlock), 0, stream, rp, W.ps, n, film, counts,
                       out ? out + 5 * base : nullptr);
    LGPU_HIP(hipGetLastError());
    S.waves++;
  }
  LGPU_HIP(hipStreamSynchronize(stream));
  if (countTrav) {
    unsigned long long tc[2];
    LGPU_HIP(hipMemcpy(tc, W.tcount.ptr, 16, hipMemcpyDeviceToHost));
    S.nodes_visited = tc[0];
    S.tris_tested = tc[1];
  }
  S.samples = total;
  S.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = S;
  return LGPU_OK;
}

}  // namespace lgpu

using namespace lgpu;

extern "C" lgpu_status lgpu_render(const lgpu_render_desc *d, lgpu_render_stats *stats) {
  lgpu_status st = ValidateRender(d);
  if (st != LGPU_OK) return st;
  if (!d->film) return Fail(LGPU_ERR_INVALID_ARGUMENT, "render: null film");
  if (d->spp_end <= d->spp_begin) return Fail(LGPU_ERR_INVALID_ARGUMENT, "render: empty spp range");
  lgpu_film f = d->film;
  const uint64_t total = (uint64_t)f->width * f->height * (d->spp_end - d->spp_begin);
  return RunWaves(d, f->width, f->height, total, nullptr, nullptr, f->sum.as<double>(), f->count.as<uint32_t>(), nullptr, stats);
}

extern "C" lgpu_status lgpu_trace_paths(const lgpu_render_desc *d, uint32_t num, const uint32_t *pixels,
                                        const uint32_t *samples, float *out) {
  lgpu_status st = ValidateRender(d);
  if (st != LGPU_OK) return st;
  if (!d->film) return Fail(LGPU_ERR_INVALID_ARGUMENT, "trace_paths: film gives the image size and may not be null");
  if (!num) return LGPU_OK;
  if (!pixels || !samples || !out) return Fail(LGPU_ERR_INVALID_ARGUMENT, "trace_paths: null array");
  lgpu_film f = d->film;
  const uint64_t npx = (uint64_t)f->width * f->height;
  for (uint32_t i = 0; i < num; ++i)
    if (pixels[i] >= npx) return Fail(LGPU_ERR_INVALID_ARGUMENT, "trace_paths: pixel index out of range");
  LGPU_HIP(hipSetDevice(d->scene->dev->device));
  DevBuf dp, ds, dout;
  LGPU_HIP(dp.upload(pixels, 4ull * num));
  LGPU_HIP(ds.upload(samples, 4ull * num));
  LGPU_HIP(dout.alloc(20ull * num));
  st = RunWaves(d, f->width, f->height, num, dp.as<uint32_t>(), ds.as<uint32_t>(), nullptr, nullptr, dout.as<float>(), nullptr);
  (void)hipStreamSynchronize(d->scene->dev->stream);  // no kernel may outlive the list buffers
  if (st != LGPU_OK) return st;
  LGPU_HIP(hipMemcpy(out, dout.ptr, 20ull * num, hipMemcpyDeviceToHost));
  return LGPU_OK;
}

static lgpu_status Query(lgpu_scene s, uint32_t n, const float *rays, uint32_t *hits, uint32_t *occ) {
  if (!s || !rays || (!hits && !occ)) return Fail(LGPU_ERR_INVALID_ARGUMENT, "query: null argument");
  if (!s->committed) return Fail(LGPU_ERR_STATE, "query: scene not committed");
  if (!n) return LGPU_OK;
  LGPU_HIP(hipSetDevice(s->dev->device));
  (void)hipGetLastError();
  DevBuf dr, dh, dsp;
  LGPU_HIP(dr.upload(rays, 32ull * n));
  LGPU_HIP(dh.alloc(occ ? 4ull * n : 16ull * n));
  LGPU_HIP(dsp.alloc(4ull * (kMaxBvhDepth - kLdsStack) * n));
  hipLaunchKernelGGL(k_query, dim3(Blocks(n)), dim3(kBlock), 0, s->dev->stream, s->dsc, dr.as<float4>(), n,
                     occ ? nullptr : dh.as<uint4>(), occ ? dh.as<uint32_t>() : nullptr, dsp.as<uint32_t>());
  const hipError_t le = hipGetLastError();
  const hipError_t se = hipStreamSynchronize(s->dev->stream);  // before the buffers leave scope
  LGPU_HIP(le);
  LGPU_HIP(se);
  LGPU_HIP(hipMemcpy(occ ? (void *)occ : (void *)hits, dh.ptr, occ ? 4ull * n : 16ull * n, hipMemcpyDeviceToHost));
  return LGPU_OK;
}
extern "C" lgpu_status lgpu_intersect(lgpu_scene s, uint32_t n, const float *rays, uint32_t *hits) {
  return Query(s, n, rays, hits, nullptr);
}
extern "C" lgpu_status lgpu_occluded(lgpu_scene s, uint32_t n, const float *rays, uint32_t *occ) {
  return Query(s, n, rays, nullptr, occ);
}

extern "C" lgpu_status lgpu_philox(lgpu_device dev, uint32_t n, const uint32_t *counters, const uint32_t key[2], uint32_t *out) {
  if (!dev || !counters || !key || !out) return Fail(LGPU_ERR_INVALID_ARGUMENT, "philox: null argument");
  if (!n) return LGPU_OK;
  LGPU_HIP(hipSetDevice(dev->device));
  DevBuf dc, dout;
  LGPU_HIP(dc.upload(counters, 16ull * n));
  LGPU_HIP(dout.alloc(16ull * n));
  hipLaunchKernelGGL(k_philox, dim3((n + 255) / 256), dim3(256), 0, dev->stream, n, dc.as<uint4>(), key[0], key[1], dout.as<uint4>());
  const hipError_t le = hipGetLastError();
  const hipError_t se = hipStreamSynchronize(dev->stream);
  LGPU_HIP(le);
  LGPU_HIP(se);
  LGPU_HIP(hipMemcpy(out, dout.ptr, 16ull * n, hipMemcpyDeviceToHost));
  return LGPU_OK;
}

namespace lgpu {
__global__ void __launch_bounds__(kBlock) k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) dst[i] = src[i];
}
}  // namespace lgpu

extern "C" lgpu_status lgpu_copy_bench(lgpu_device dev, uint64_t bytes, uint32_t iters, double *ms) {
  if (!dev || !ms || bytes < 16 || !iters) return Fail(LGPU_ERR_INVALID_ARGUMENT, "copy_bench: bad argument");
  LGPU_HIP(hipSetDevice(dev->device));
  const uint64_t n = bytes / 16;
  DevBuf a, b;
  LGPU_HIP(a.alloc(16 * n));
  LGPU_HIP(b.alloc(16 * n));
  LGPU_HIP(hipMemsetAsync(a.ptr, 0, 16 * n, dev->stream));
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + kBlock - 1) / kBlock, 256ull * 64);
  hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(kBlock), 0, dev->stream, a.as<float4>(), b.as<float4>(), n);
  LGPU_HIP(hipEventRecord(dev->ev[0], dev->stream));
  for (uint32_t i = 0; i < iters; ++i)
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(kBlock), 0, dev->stream, a.as<float4>(), b.as<float4>(), n);
  LGPU_HIP(hipEventRecord(dev->ev[1], dev->stream));
  const hipError_t le = hipGetLastError();
  const hipError_t se = hipStreamSynchronize(dev->stream);
  LGPU_HIP(le);
  LGPU_HIP(se);
  float t;
  LGPU_HIP(hipEventElapsedTime(&t, dev->ev[0], dev->ev[1]));
  *ms = t / iters;
  return LGPU_OK;
}
