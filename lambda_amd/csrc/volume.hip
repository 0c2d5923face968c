// Volumetric shade stage of the wavefront: one launch processes one iteration of
// VolumetricPathIntegrator::Li (VolumetricPathIntegrator.cpp:20-162) or
// MISVolumetricPathIntegrator::Li (MISVolumetricPathIntegrator.cpp:71-215) for every live path,
// after the extend stage has traced the path's ray.
//
// Per path and launch:
//   1. the light reached by the previous vertex's BSDF / phase sample (MIS), deferred to here
//      because that sample's ray is the continuation ray the extend stage just traced;
//      with kFlagTrMIS the lookup continues through pass-through surfaces multiplying Tr;
//   2. bounce-0 emission (camera rays, also after pass-through surfaces at depth 0);
//   3. inside a medium: distance sampling (VOLPATH) or the light-point / distance /
//      equiangular triple with both medium points' direct light (MIS_VOLPATH);
//   4. otherwise the surface vertex (NEE + BSDF sample) or a pass-through boundary;
//   5. Russian roulette, then the continuation is appended to the next queue.
// NEE visibility and the medium points' phase-sampled light lookups are traced inline.
#include "adjoint.h"
#include "host.h"
#include "lights.h"
#include "media.h"
#include "shading.h"
#include "trace.h"

namespace lgpu {

namespace {

struct PathOut {  // continuation written to the next queue
  V3 o, d;
  float eta;
  Spec beta;
  uint32_t flags, depth, pick, pass, med;
  float4 mis;
  V3 prevP, prevN;
  bool medInter = false;  // the vertex is a medium interaction (its deferred light lookup reads 3-D textures)
};

__device__ inline void Push(const Queues &q, const uint4 &ids, uint32_t sid, uint32_t dim, float l0, const PathOut &p) {
  const uint32_t j = atomicAdd(&q.counters[0], 1u);
  q.next.ray_o[j] = ToF4(p.o, l0);
  q.next.ray_d[j] = ToF4(p.d, p.eta);
  q.next.beta[j] = F4(p.beta);
  q.next.ids[j] = make_uint4(ids.x, ids.y, dim, (p.depth << 8) | p.flags);
  q.next.aux[j] = make_uint4(sid, p.pick, p.pass, p.med);
  q.next.mis[j] = p.mis;
  q.next.prev_p[j] = ToF4(p.prevP, 0.f);
  q.next.prev_n[j] = ToF4(p.prevN, p.medInter ? 1.f : 0.f);
}

// MISVolumetricPathIntegrator::LdMediumPoint (MISVolumetricPathIntegrator.cpp:23-69). `eh` is the
// event's hit: the path's hit in faithful mode (overwritten here as written), a copy in corrected.
template <bool kPrograms>
__device__ inline Spec LdMediumPoint(const DevScene &sc, const MediumRec &M, uint32_t med, Ev &ev, Hit &eh, V3 p,
                                     PartialLS &ls, V3 refP, V3 refN, V3 *wiOut, Rng &rng, const VmCtx &vm,
                                     bool corrected, const VisCtx &vc) {
  const float partialLightPDF = corrected ? ls.selPdf : ls.pdf;
  VmCtx vmm = vm;  // the phase-sampled light is looked up at a medium interaction
  vmm.medium = 1;
  eh.point = p;
  Spec Li = S(0.f);
  bool delta = false;
  if (ls.light != kNone) {
    const LightRec &LL = LightOf(sc, ls.light);
    delta = LL.kind == kLightPoint || LL.kind == kLightSpot;
    Li = ls.Le * VisibilityL(sc, LL, p, ev, eh, ls, vm, corrected, vc);
  } else {
    ls.pdf = 0.f;
  }
  const float lightPDF = ls.pdf;
  Spec Ld = S(0.f);
  if (lightPDF > 0 && !IsBlack(Li, vm.mode)) {
    const float ph = PhaseHG(Dot(ev.wo, ev.wi), M.g);
    if (ph > 0) {
      const float w = (corrected && delta) ? 1.f : PowerHeuristic(lightPDF, ph);
      Ld = Ld + Li * ph * w / lightPDF;
    }
  }
  const float u0 = rng.Get1D(), u1 = rng.Get1D();
  V3 wi;
  const float scatPdf = HGSample(M.g, ev.wo, &wi, u0, u1, corrected);
  ev.wi = wi;
  *wiOut = wi;
  // corrected (the caller's hit record is a copy): without emissive meshes and an environment the
  // phase-sampled ray reaches no light, its term adds zero, and its transmittance walk is skipped
  if (corrected && !(sc.has_mesh_lights || sc.has_env)) return Ld;
  Spec Tr = S(1.f);
  uint32_t tbi = kNone;
  const bool si = IntersectTr(sc, p, wi, med, -1.f, &Tr, vm, corrected, vc.tc, eh, tbi, nullptr, false);
  if (scatPdf > 0) {
    const Spec C = MisLight<kPrograms>(sc, ls.light, partialLightPDF, refP, refN, si, eh, tbi, wi, scatPdf, false, corrected, vmm,
                            corrected);
    Ld = Ld + (corrected ? C * Tr : C * Tr / scatPdf);  // the phase value is missing as written
  }
  return Ld;
}

// kPrograms: the scene has program sockets (DevScene::has_programs); without them the kernel is
// compiled without the shader VM.
// kAdjoint: the replay of the adjoint render (corrected estimator, RGB, VOLPATH): the same path
// decisions and radiance as the primal launch, plus the gradient pieces of adjoint.h.
// np: when set, the item count is read from this device counter (the grid covers an upper bound).
template <bool kPrograms, bool kAdjoint>
__global__ void __launch_bounds__(kBlock) k_shade_vol(DevScene sc, RenderParams rp, PathState ps, SampleState ss,
                                                      uint32_t n, const uint32_t *np, Queues q, uint32_t *spillBase,
                                                      AdjRefs ar) {
  __shared__ uint32_t stk[kLdsStack][kBlock];
  __shared__ TreeImp treeLds[kTreeLds];
  if (np) n = *np;
  LoadTreeLds(sc, treeLds);
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  EmitAcc eacc{{kNone, kNone}, {0.f, 0.f}};
  if (i < n) {  // the adjoint's emission flush below needs every lane of the wave
  const TraceCtx tc{stk, spillBase + (size_t)i * sc.spill_words};
  const bool corrected = rp.corrected != 0;
  const bool volumes = corrected && sc.num_media > 0;  // Scene::hasVolumes: never set as written
  const bool misVol = rp.integrator == LGPU_INTEGRATOR_MIS_VOLPATH;
  const int mode = (int)rp.mode;
  const uint4 ids = ps.ids[i];
  const uint4 aux = ps.aux[i];
  const uint32_t sid = aux.x;
  const uint32_t flags = ids.w & 0xFFu;
  const uint32_t depth = ids.w >> 8;
  const uint32_t sideFlag = flags & kFlagSideNeg;
  const float4 ro4 = ps.ray_o[i], rd4 = ps.ray_d[i];
  const V3 ro = ToV3(ro4), rd = ToV3(rd4);
  const float4 hr = ps.hit[i];
  uint32_t bi = __float_as_uint(hr.w);
  const bool hit = bi != kNone;
  const uint32_t med = aux.w;
  Spec beta = S4(ps.beta[i]);
  Spec Lacc = S(0.f);
  Spec adjd = S(0.f), rest = S(0.f);  // adjoint: loss gradient of the sample, radiance still to come
  VisQuery vq{};
  if constexpr (kAdjoint) {
    adjd = AdjDelta(ar, ss.pixel[sid]);
    rest = S4(ar.ltot[sid]) - S4(ss.L[sid]);
  }
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
  Hit h;
  if (hit) ProcessHit(sc, ro, rd, hr.x, hr.y, hr.z, bi, corrected, h);
  if (hit) VmAtHit(vm, h.u, h.v, h.point, h.normalS);
  else VmAtHit(vm, 0.f, 0.f, v3(0.f, 0.f, 0.f), v3(0.f, 0.f, 0.f));
  bool alive = true;

  // ---- 1. light reached by the previous vertex's sampled direction. Without emissive meshes and
  // an environment no light can be reached (MisLight finds none), so the term and its transmittance
  // walk are skipped: they add zero.
  if (!(flags & kFlagCamera) && !(flags & kFlagNoEmit) && !(sc.has_mesh_lights || sc.has_env)) {
    if ((flags & kFlagEmitOnly) || !hit) alive = false;
  } else if (!(flags & kFlagCamera) && !(flags & kFlagNoEmit)) {
    const float4 mis = ps.mis[i];
    bool si = hit;
    Hit th = h;
    uint32_t tbi = bi;
    Spec Tr = S(1.f);
    if ((flags & kFlagTrMIS) && hit) si = IntersectTr(sc, ro, rd, med, -1.f, &Tr, vm, corrected, tc, th, tbi, nullptr, true);
    const float4 pn = ps.prev_n[i];
    VmCtx vmm = vm;
    vmm.medium = pn.w != 0.f;  // looked up for a medium interaction
    const Spec C = MisLight<kPrograms>(sc, aux.y, mis.y, ToV3(ps.prev_p[i]), ToV3(pn), si, th, tbi, rd, mis.x,
                            (flags & kFlagSpecular) != 0, false, vmm, corrected);
    const Spec C1 = beta * (C * Tr * mis.w);
    Lacc = Lacc + C1;
    if constexpr (kAdjoint) {  // the light's intensity; the transmittance of the lookup's walk
      const Spec X = adjd * C1;
      if (si && sc.meshes[th.mesh].mesh_light != kNone) EmitAdd(sc, eacc, 0, sc.meshes[th.mesh].material, X);
      if ((flags & kFlagTrMIS) && hit) TrGradWalk(sc, ar, ro, rd, med, -1.f, X, vm, tc);
    }
    if ((flags & kFlagEmitOnly) || !hit) alive = false;
  } else if (!(flags & kFlagCamera) && !hit) {
    alive = false;
  }
  if (alive && (flags & kFlagRR)) beta = beta / S(1.f - ps.mis[i].z);  // VolumetricPathIntegrator.cpp:153

  // ---- 2. bounce-0 emission (VolumetricPathIntegrator.cpp:32-39, 156-159)
  if (alive && (flags & kFlagCamera)) {
    if (!hit) {
      if (sc.has_env) Lacc = Lacc + EnvLe(sc, rd, vm) * beta;
      alive = false;
    } else {
      const MeshRec m = sc.meshes[h.mesh];
      if (m.mesh_light != kNone && (!corrected || Dot(rd, h.normalG) < 0)) {
        Spec Le = EmissionOf<kPrograms>(sc, h.mesh, vm) * sc.lights[m.mesh_light].intensity * kInvPi;
        if (med != kNone) Le = Le * MediumTr(sc, sc.media[med], ro, rd, h.tFar, vm);
        Lacc = Lacc + (corrected ? beta * Le : Le);  // beta taken as 1 as written
        if constexpr (kAdjoint) {
          const Spec X = adjd * (beta * Le);
          EmitAdd(sc, eacc, 0, m.material, X);
          if (med != kNone) TrGradSegment(sc, ar, sc.media[med], ro, rd, h.tFar, X, vm);
        }
      }
    }
  }

  PathOut out;
  bool push = false;
  bool rrCheck = false;  // a scattering vertex happened: Russian roulette applies
  if (alive) {
    bool medInter = false;
    // ---- 3. medium
    if (med != kNone) {
      const MediumRec M = sc.media[med];
      Ev ev;
      ev.wo = -rd;
      ev.wi = rd;
      ev.eta = rd4.w;
      ev.woL = ev.wo;
      ev.wiL = ev.wi;
      ev.sidedness = sideFlag ? -1 : 1;
      const VisCtx vc{tc, volumes, true, med, nullptr, v3(0, 0, 0), kAdjoint ? &vq : nullptr};
      if (!misVol) {
        const float u0 = rng.Get1D(), u1 = rng.Get1D();
        const MediumSample ds = SampleDistance(sc, M, ro, rd, h.tFar, u0, u1, vm);
        // every contribution from here on carries the distance sample's weight
        if constexpr (kAdjoint) DistanceGrad(sc, ar, M, ro, rd, h.tFar, u0, u1, ds.inter, adjd * (rest - Lacc), vm);
        beta = beta * ds.w;
        if (ds.inter) {  // VolumetricPathIntegrator.cpp:99-148
          medInter = true;
          const V3 sp = ro + rd * ds.t;
          h.point = sp;
          vm.medium = 1;
          vm.px = sp.x; vm.py = sp.y; vm.pz = sp.z;
          const V3 refP = sp, refN = corrected ? v3(0, 0, 0) : h.normalS;
          float lightDistPdf = 1.f;
          const uint32_t pick = SampleLight(sc, refP, refN, rng, &lightDistPdf, treeLds, corrected);
          Spec Li = S(0.f);
          float lightPDF = 0.f;
          bool delta = false;
          if (pick != kNone) {
            const LightRec &LL = LightOf(sc, pick);
            delta = LL.kind == kLightPoint || LL.kind == kLightSpot;
            ShadowSpec sh;
            Li = SampleLi<true, kPrograms>(sc, LL, h, ev, rng, vm, corrected, lightPDF, sh, &vc);
          }
          lightPDF *= lightDistPdf;
          Spec Ld = S(0.f);
          if (lightPDF > 0 && !IsBlack(Li, mode)) {
            const float ph = PhaseHG(Dot(ev.wo, ev.wi), M.g);
            if (ph > 0) {
              const float w = (corrected && delta) ? 1.f : PowerHeuristic(lightPDF, ph);
              Ld = Ld + Li * ph * w / lightPDF;
            }
          }
          if constexpr (kAdjoint) {  // the NEE contribution's light intensity and transmittance
            const Spec X = adjd * (beta * Ld);
            EmitAdd(sc, eacc, 1, EmitMaterial(sc, pick), X);
            if (vq.valid) TrGradWalk(sc, ar, vq.o, vq.d, vq.med, vq.maxT, X, vm, tc);
          }
          const float hu0 = rng.Get1D(), hu1 = rng.Get1D();
          V3 wi;
          const float scatPdf = HGSample(M.g, ev.wo, &wi, hu0, hu1, corrected);
          if (!(scatPdf > 0)) {
            if (corrected) Lacc = Lacc + beta * Ld;
          } else {
            Lacc = Lacc + beta * Ld;
            out.o = sp; out.d = wi; out.eta = rd4.w;
            out.flags = kFlagTrMIS | sideFlag;
            out.pick = pick;
            out.mis = make_float4(scatPdf, lightDistPdf, 0.f, corrected ? 1.f : 1.f / scatPdf);
            out.prevP = refP; out.prevN = refN;
            out.medInter = true;
            out.med = med;
            push = true;
            rrCheck = true;
          }
        }
      } else {  // MISVolumetricPathIntegrator.cpp:92-147
        const V3 refP = corrected ? ro : h.point, refN = corrected ? v3(0, 0, 0) : h.normalS;
        PartialLS ls;
        ls.pdf = 1.f;
        ls.light = SampleLight(sc, refP, refN, rng, &ls.pdf, treeLds, corrected);
        ls.selPdf = ls.pdf;
        ls.point = v3(0, 0, 0);
        ls.normal = v3(0, 0, 0);
        if (ls.light != kNone) ls.Le = SamplePointL<kPrograms>(sc, LightOf(sc, ls.light), h, rng, vm, corrected, ls);
        else { ls.pdf = 0.f; ls.Le = S(0.f); }
        const int choice = rng.Get1D() > .5f ? 0 : 1;  // 0 distance, 1 equiangular
        const float du0 = rng.Get1D(), du1 = rng.Get1D();
        const float eps = rng.Get1D();
        rng.Get1D();  // channel draw, unused (HomogeneousMedium.cpp:70)
        rng.Get1D();
        MediumSample sD, sE;  // one grid walk for both samples
        SampleDistanceEquiangular(sc, M, ro, rd, h.tFar, du0, du1, h.point, ls.point, eps, vm, corrected, sD, sE);
        const V3 mp0 = ro + rd * sD.t, mp1 = ro + rd * sE.t;
        if (!corrected) {
          if (!(sD.inter && sE.inter)) {
            beta = beta * sD.w;
          } else {
            medInter = true;
            const float pdfD2 = PdfDistanceAt(M, sE.D, sE.rho, vm, false);
            const float pdfE2 = PdfEquiangular(ro, rd, ls.point, h.tFar, sD.t, false);
            const float wD = PowerHeuristic(sD.pdf, pdfE2), wE = PowerHeuristic(sE.pdf, pdfD2);
            V3 wi0, wi1;
            Spec Ld = S(0.f);
            Ld = Ld + LdMediumPoint<kPrograms>(sc, M, med, ev, h, mp0, ls, refP, refN, &wi0, rng, vm, false, vc) * sD.w * wD;
            Ld = Ld + LdMediumPoint<kPrograms>(sc, M, med, ev, h, mp1, ls, refP, refN, &wi1, rng, vm, false, vc) * sE.w * wE;
            Lacc = Lacc + Ld;  // beta missing as written
            beta = beta * (choice == 0 ? sD.w : sE.w);
            out.o = choice == 0 ? mp0 : mp1;
            out.d = choice == 0 ? wi0 : wi1;
            push = true;
          }
        } else {
          // both points estimate direct light (power-heuristic MIS over the two distance
          // techniques); the path continues with the distance sample
          Spec Ld = S(0.f);
          V3 wi0 = v3(0, 0, 0), wi1;
          if (sD.inter) {
            PartialLS l0 = ls;
            Hit eh = h;
            const float wD = PowerHeuristic(sD.pdf, PdfEquiangular(ro, rd, ls.point, h.tFar, sD.t, true));
            Ld = Ld + LdMediumPoint<kPrograms>(sc, M, med, ev, eh, mp0, l0, refP, refN, &wi0, rng, vm, true, vc) * sD.w * wD;
          }
          if (sE.inter) {
            PartialLS l1 = ls;
            Hit eh = h;
            const float wE = PowerHeuristic(sE.pdf, PdfDistanceAt(M, sE.D, sE.rho, vm, true));
            Ld = Ld + LdMediumPoint<kPrograms>(sc, M, med, ev, eh, mp1, l1, refP, refN, &wi1, rng, vm, true, vc) * sE.w * wE;
          }
          Lacc = Lacc + beta * Ld;
          beta = beta * sD.w;
          if (sD.inter) {
            medInter = true;
            out.o = mp0;
            out.d = wi0;
            push = true;
          }
        }
        if (push) {
          out.eta = rd4.w;
          out.flags = kFlagNoEmit | sideFlag;  // the phase-sampled light was looked up above
          out.pick = kNone;
          out.mis = make_float4(0.f, 0.f, 0.f, 0.f);
          out.prevP = v3(0, 0, 0); out.prevN = v3(0, 0, 0);
          out.med = med;
          rrCheck = true;
        }
      }
    }

    // ---- 4. surface
    if (!medInter) {
      const MaterialRec mat = sc.materials[sc.meshes[h.mesh].material];
      if (mat.lobe == kNone) {  // pass-through boundary (VolumetricPathIntegrator.cpp:91-97)
        const uint32_t pass = aux.z + 1;
        if (pass <= 1024) {
          out.o = h.point + h.normalG * (Dot(h.normalG, rd) < 0 ? -kSurfaceEpsilon : kSurfaceEpsilon);
          out.d = rd; out.eta = rd4.w;
          out.flags = ((flags & kFlagCamera) ? kFlagCamera : kFlagNoEmit) | sideFlag;
          out.pick = aux.y;
          out.pass = pass;
          out.med = GetMedium(mat, rd, h.normalG);
          out.mis = ps.mis[i];
          out.prevP = ToV3(ps.prev_p[i]); out.prevN = ToV3(ps.prev_n[i]);
          out.beta = beta;
          out.depth = depth;
          Push(q, ids, sid, rng.dim, l0, out);
        }
        push = false;
      } else {  // surface vertex (VolumetricPathIntegrator.cpp:45-90)
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
        const V3 shadeP = h.point, shadeN = h.normalS;
        const VisCtx vc{tc, volumes, false, med, &mat, h.normalG, kAdjoint ? &vq : nullptr};
        VertexGrad vg;
        if constexpr (kAdjoint) VertexGradInit(shd, mat.lobe, vg);
        float lightDistPdf = 1.f;
        const uint32_t pick = SampleLight(sc, h.point, h.normalS, rng, &lightDistPdf, treeLds, corrected);
        float lightPDF = 0.f;
        Spec Li = S(0.f);
        bool delta = false;
        if (pick != kNone) {
          const LightRec &LL = LightOf(sc, pick);
          delta = LL.kind == kLightPoint || LL.kind == kLightSpot;
          ShadowSpec sh;
          Li = SampleLi<true, kPrograms>(sc, LL, h, ev, rng, vm, corrected, lightPDF, sh, &vc);
        }
        lightPDF *= lightDistPdf;
        Spec cNee = S(0.f);
        if (lightPDF > 0 && !IsBlack(Li, mode)) {
          const Spec f0 = shd.F(mat.lobe, ev);
          const Spec f = f0 * fabsf(ev.wiL.y);
          const float sp = shd.Pdf(mat.lobe, ev.woL, ev.wiL);
          if (!IsBlack(f, mode) && sp > 0) {
            const float w = (corrected && delta) ? 1.f : PowerHeuristic(lightPDF, sp);
            cNee = Li * f * w / lightPDF;
            if constexpr (kAdjoint) {  // NEE: the BSDF at the light direction, intensity, transmittance
              const Spec X = adjd * (beta * cNee);
              VertexGradAdd(shd, ev, X, f0, vg);
              EmitAdd(sc, eacc, 1, EmitMaterial(sc, pick), X);
              if (vq.valid) TrGradWalk(sc, ar, vq.o, vq.d, vq.med, vq.maxT, X, vm, tc);
            }
          }
        }
        float scatPdf = 0.f;
        bool spec = false;
        const Spec fS = shd.Sample(mat.lobe, ev, h, rng, scatPdf, spec);
        const Spec f = fS * fabsf(ev.wiL.y);
        const uint32_t side = ev.sidedness < 0 ? kFlagSideNeg : 0u;
        if (!(scatPdf > 0 && !IsBlack(f, mode))) {
          if (corrected) Lacc = Lacc + beta * cNee;  // NEE is dropped by `break` as written
          if constexpr (kAdjoint) VertexGradFlush(sc, ar, vm, vg);
        } else {
          Lacc = Lacc + beta * cNee;
          if constexpr (kAdjoint) {  // continuation: everything later launches gather
            if (!spec) VertexGradAdd(shd, ev, adjd * (rest - Lacc), fS, vg);
            VertexGradFlush(sc, ar, vm, vg);
          }
          beta = beta * (f / scatPdf);
          out.o = h.point; out.d = ev.wi; out.eta = ev.eta;
          out.flags = (spec ? kFlagSpecular : 0u) | (volumes ? kFlagTrMIS : 0u) | side;
          out.pick = pick;
          out.mis = make_float4(scatPdf, lightDistPdf, 0.f, 1.f);
          out.prevP = shadeP; out.prevN = shadeN;
          out.med = GetMedium(mat, ev.wi, h.normalG);
          push = true;
          rrCheck = true;
        }
      }
    }
  }

  // ---- 5. Russian roulette (VolumetricPathIntegrator.cpp:150-154) and the continuation
  if (push) {
    out.beta = beta;
    out.depth = depth + 1;
    out.pass = aux.z;
    const bool deferred = !(out.flags & kFlagNoEmit);  // light lookup pending at the next hit
    bool kill = false;
    if (rrCheck && depth > rp.min_bounces) {
      const float qrr = fmaxf(.05f, 1 - Luminance(beta, mode));
      if (rng.Get1D() < qrr) kill = true;
      else out.flags |= kFlagRR;
      out.mis.z = qrr;
    }
    if (depth + 1 >= rp.max_bounces) kill = true;
    if (kill) {
      out.flags |= kFlagEmitOnly;
      out.flags &= ~kFlagRR;
    }
    if (!kill || deferred) Push(q, ids, sid, rng.dim, l0, out);
  }
  if (!IsBlack(Lacc, 1)) {
    float4 L = ss.L[sid];
    L.x += Lacc.c[0]; L.y += Lacc.c[1]; L.z += Lacc.c[2]; L.w += Lacc.c[3];
    ss.L[sid] = L;
  }
  }  // i < n
  if constexpr (kAdjoint) EmitFlush(ar, eacc);
}

}  // namespace

void LaunchShadeVol(uint32_t blocks, hipStream_t stream, const DevScene &sc, const RenderParams &rp, PathState ps,
                    SampleState ss, uint32_t n, Queues q, uint32_t *spill, const AdjRefs *adj, const uint32_t *np) {
  const AdjRefs ar = adj ? *adj : AdjRefs{};
  const dim3 g(blocks), b(kBlock);
  if (adj) {
    if (sc.has_programs)
      hipLaunchKernelGGL((k_shade_vol<true, true>), g, b, 0, stream, sc, rp, ps, ss, n, np, q, spill, ar);
    else
      hipLaunchKernelGGL((k_shade_vol<false, true>), g, b, 0, stream, sc, rp, ps, ss, n, np, q, spill, ar);
  } else if (sc.has_programs) {
    hipLaunchKernelGGL((k_shade_vol<true, false>), g, b, 0, stream, sc, rp, ps, ss, n, np, q, spill, ar);
  } else {
    hipLaunchKernelGGL((k_shade_vol<false, false>), g, b, 0, stream, sc, rp, ps, ss, n, np, q, spill, ar);
  }
}

}  // namespace lgpu
