// Direct-lighting shade stage: DirectLightingIntegrator::Li (DirectLightingIntegrator.cpp:13-29)
// with Integrator::SampleOneLight / EstimateDirect (Integrator.cpp:5-55). One launch shades the
// camera hits of the live paths: surfaces without a BxDF are crossed (the continuation goes back
// to the extend stage), a shading surface estimates direct light from one light chosen by the
// light sampler (light sample + BSDF sample, power-heuristic MIS) and ends the path. The
// visibility, PDF_Li and MIS rays are traced inline.
#include "host.h"
#include "lights.h"
#include "media.h"
#include "shading.h"
#include "trace.h"

namespace lgpu {

namespace {

template <bool kPrograms>
__device__ inline Spec EmissionAt(const DevScene &sc, const Hit &lh, VmCtx vm) {
  VmAtHit(vm, lh.u, lh.v, lh.point, lh.normalS);
  return EmissionOf<kPrograms>(sc, lh.mesh, vm) * sc.lights[sc.meshes[lh.mesh].mesh_light].intensity * kInvPi;
}

// Light::PDF_Li(event, sampler) as written (MeshLight.cpp:44-55, :153-155; EnvironmentLight.cpp:44-58;
// point and spot lights: 0): the reference fires its own ray for it.
__device__ inline float PdfLiRay(const DevScene &sc, uint32_t pick, const Hit &h, const Ev &ev, const TraceCtx &tc) {
  const LightRec &L = LightOf(sc, pick);
  if (L.kind == kLightPoint || L.kind == kLightSpot) return 0.f;
  if (L.kind == kLightEnv) {
    const float theta = SphericalTheta(ev.wi) - sc.env.offset[1];
    const float phi = SphericalPhi(ev.wi) - sc.env.offset[0];
    const float sinT = sinf(theta);
    if (sinT == 0) return 0.f;
    const V3 wiOffset = SphericalDirection(sinT, cosf(theta), phi);
    float t, u, v;
    uint32_t prim, bi;
    if (Trace<true, false>(sc, h.point, wiOffset, 0.f, INFINITY, &t, &u, &v, &prim, &bi, tc.stk, tc.spill, nullptr)) return 0.f;
    return Env2DPdf(sc, Fract(phi * kInvPi2), Fract(theta * kInvPi)) / (2.f * kPi * kPi * sinT);
  }
  const uint32_t meshLight = L.kind == kLightTri ? L.parent : pick;
  Hit lh;
  uint32_t lbi;
  const V3 o = h.point + h.normalG * kSurfaceEpsilon * (float)ev.sidedness;
  if (!ClosestHit(sc, o, ev.wi, tc, false, lh, lbi)) return 0.f;
  if (sc.meshes[lh.mesh].mesh_light != meshLight) return 0.f;
  const float4 *tp = reinterpret_cast<const float4 *>(sc.tris + lbi);
  const float4 w1 = tp[1], w2 = tp[2];
  const V3 c = Cross(v3(w1.x, w1.y, w1.z), v3(w2.x, w2.y, w2.z));  // TriangleMesh::GetTriangleAreaAndNormal
  const float area = Mag(c) * .5f;
  const V3 normal = c / (area * 2.f);
  const float cosT = fabsf(Dot(normal, -ev.wi));
  if (cosT > 0) return DistSq(h.point, lh.point) / (cosT * area);
  return 0.f;
}

// kPrograms: the scene has program sockets (DevScene::has_programs); without them the kernel is
// compiled without the shader VM.
template <bool kPrograms>
__global__ void __launch_bounds__(kBlock) k_shade_direct(DevScene sc, RenderParams rp, PathState ps, SampleState ss,
                                                         uint32_t n, Queues q, uint32_t *spillBase) {
  __shared__ uint32_t stk[kLdsStack][kBlock];
  __shared__ TreeImp treeLds[kTreeLds];
  LoadTreeLds(sc, treeLds);
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const TraceCtx tc{stk, spillBase + (size_t)i * sc.spill_words};
  const bool corrected = rp.corrected != 0;
  const int mode = (int)rp.mode;
  const uint4 ids = ps.ids[i];
  const uint4 aux = ps.aux[i];
  const uint32_t sid = aux.x;
  const float4 ro4 = ps.ray_o[i], rd4 = ps.ray_d[i];
  const V3 ro = ToV3(ro4), rd = ToV3(rd4);
  const float4 hr = ps.hit[i];
  const uint32_t bi = __float_as_uint(hr.w);
  Rng rng;
  rng.init(ids.x, ids.y, ids.z, rp.seed);
  VmCtx vm;
  vm.mode = mode;
  vm.texels = sc.texels; vm.textures = sc.textures; vm.consts = sc.consts; vm.code = sc.code;
  const float l0 = ro4.w;
  if (mode) {
    for (int k = 0; k < 4; ++k) {
      float l = l0 + 117.5f * (float)k;
      if (l >= 830.f) l -= 470.f;
      vm.lam[k] = l;
    }
  } else {
    vm.lam[0] = vm.lam[1] = vm.lam[2] = vm.lam[3] = 0.f;
  }
  vm.woLy = 0.f;
  vm.eta = rd4.w;
  vm.medium = 0;
  VmAtHit(vm, 0.f, 0.f, v3(0.f, 0.f, 0.f), v3(0.f, 0.f, 0.f));
  Spec Lacc = S(0.f);
  if (bi == kNone) {  // :28 (no environment light: 0)
    if (sc.has_env) Lacc = EnvLe(sc, rd, vm);
  } else {
    Hit h;
    ProcessHit(sc, ro, rd, hr.x, hr.y, hr.z, bi, corrected, h);
    VmAtHit(vm, h.u, h.v, h.point, h.normalS);
    const MeshRec m = sc.meshes[h.mesh];
    const MaterialRec mat = sc.materials[m.material];
    if (corrected && m.mesh_light != kNone && Dot(rd, h.normalG) < 0)  // first-surface emission (corrected)
      Lacc = EmissionOf<kPrograms>(sc, h.mesh, vm) * sc.lights[m.mesh_light].intensity * kInvPi;
    if (mat.lobe == kNone) {  // :24-26
      const uint32_t pass = aux.z + 1;
      if (pass <= 1024) {
        const uint32_t j = atomicAdd(&q.counters[0], 1u);
        q.next.ray_o[j] = ToF4(h.point + rd * .0001f, l0);
        q.next.ray_d[j] = rd4;
        q.next.beta[j] = ps.beta[i];
        q.next.ids[j] = make_uint4(ids.x, ids.y, rng.dim, kFlagCamera);
        q.next.aux[j] = make_uint4(sid, kNone, pass, kNone);
      }
    } else {
      ShaderT<kPrograms> shd;
      shd.sc = &sc;
      shd.corrected = corrected;
      Ev ev;
      ev.wo = -rd;
      ev.wi = rd;
      ev.eta = rd4.w;
      ev.woL = ToLocal(h, ev.wo);
      ev.wiL = ToLocal(h, ev.wi);
      ev.sidedness = (ev.woL.y > 0) ? 1 : -1;
      vm.woLy = ev.woL.y;
      shd.vm = vm;
      float lightPdf = 1.f;
      const uint32_t pick = SampleLight(sc, h.point, h.normalS, rng, &lightPdf, treeLds, corrected);
      if (pick != kNone) {
        const LightRec &LL = LightOf(sc, pick);
        const bool delta = LL.kind == kLightPoint || LL.kind == kLightSpot;
        const VisCtx vc{tc, false, false, kNone, &mat, h.normalG};
        float lightPDF = 0.f;
        ShadowSpec sh;
        const Spec Li = SampleLi<true, kPrograms>(sc, LL, h, ev, rng, vm, corrected, lightPDF, sh, &vc);
        Spec Ld = S(0.f);
        if (lightPDF > 0 && !IsBlack(Li, mode)) {
          const Spec f = shd.F(mat.lobe, ev) * fabsf(ev.wiL.y);
          const float sp = shd.Pdf(mat.lobe, ev.woL, ev.wiL);
          if (!IsBlack(f, mode) && sp > 0) {
            const float w = (corrected && delta) ? 1.f : PowerHeuristic(lightPDF, sp);
            Ld = Ld + Li * f * w / lightPDF;
          }
        }
        float scatPdf = 0.f;
        bool spec = false;
        Spec f = shd.Sample(mat.lobe, ev, h, rng, scatPdf, spec);
        f = f * fabsf(ev.wiL.y);
        bool done = false;
        if (scatPdf > 0 && !IsBlack(f, mode)) {
          Spec Lb = S(0.f);
          float weight = 1.f;
          if (!corrected) {
            const float lp = PdfLiRay(sc, pick, h, ev, tc);
            if (lp == 0) {
              done = true;
            } else {
              weight = PowerHeuristic(scatPdf, lp);
              Hit lh;
              uint32_t lbi;
              if (ClosestHit(sc, h.point + h.normalG * .00001f, ev.wi, tc, false, lh, lbi)) {
                // material->light == &_light: under the tree a TriangleLight never equals it
                const uint32_t ml = sc.meshes[lh.mesh].mesh_light;
                if (ml != kNone && !(pick & kTreeBit) && ml == pick) Lb = EmissionAt<kPrograms>(sc, lh, vm);
              } else if (LL.kind == kLightEnv) {
                Lb = EnvLe(sc, ev.wi, vm);
              }
            }
          } else if (!delta) {
            Hit lh;
            uint32_t lbi;
            float lp = 0.f;
            const V3 o = h.point + h.normalG * (Dot(ev.wi, h.normalG) < 0 ? -kSurfaceEpsilon : kSurfaceEpsilon);
            if (ClosestHit(sc, o, ev.wi, tc, true, lh, lbi)) {
              const uint32_t ml = sc.meshes[lh.mesh].mesh_light;
              const bool same = ml != kNone && ((pick & kTreeBit) ? (kTreeBit | sc.tri_tlight[lh.prim]) == pick : ml == pick);
              if (same && Dot(ev.wi, lh.normalG) < 0) {
                lp = PdfLiHit(sc, LL, ev.wi, lh, lbi, true);
                Lb = EmissionAt<kPrograms>(sc, lh, vm);
              }
            } else if (LL.kind == kLightEnv) {
              lp = EnvPdfLi(sc, ev.wi);
              Lb = EnvLe(sc, ev.wi, vm);
            }
            weight = spec ? 1.f : PowerHeuristic(scatPdf, lp);
            if (lp == 0 && !spec) Lb = S(0.f);
          }
          if (!done && !IsBlack(Lb, mode)) Ld = Ld + f * Lb * weight / scatPdf;
        }
        Lacc = Lacc + Ld / lightPdf;
      }
    }
  }
  if (!IsBlack(Lacc, 1)) {
    float4 L = ss.L[sid];
    L.x += Lacc.c[0]; L.y += Lacc.c[1]; L.z += Lacc.c[2]; L.w += Lacc.c[3];
    ss.L[sid] = L;
  }
}

}  // namespace

void LaunchShadeDirect(uint32_t blocks, hipStream_t stream, const DevScene &sc, const RenderParams &rp, PathState ps,
                       SampleState ss, uint32_t n, Queues q, uint32_t *spill) {
  if (sc.has_programs)
    hipLaunchKernelGGL(k_shade_direct<true>, dim3(blocks), dim3(kBlock), 0, stream, sc, rp, ps, ss, n, q, spill);
  else
    hipLaunchKernelGGL(k_shade_direct<false>, dim3(blocks), dim3(kBlock), 0, stream, sc, rp, ps, ss, n, q, spill);
}

}  // namespace lgpu
