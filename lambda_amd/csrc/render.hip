// Wavefront path tracer kernels and the render loop.
//
// One wave = up to `capacity` paths (pixel, sample) in flight; the camera stage lays the image
// out in 16x16 pixel tiles. Per bounce:
//   extend  - closest-hit traversal of the 4-wide BVH for the active queue (LDS stack)
//   shade   - paths grouped by shading family within the block, then emission + MIS of the
//             previous BSDF sample, NEE light sample, BSDF sample, Russian roulette; appends
//             the next queue and the shadow rays
//   shadow  - any-hit traversal; unoccluded rays add their contribution
// then accumulate adds each pixel's samples of the wave to the FP64 film.
// Reference: TileRenderers::UniformSpp (Render.cpp:43-64) and
// PathIntegrator::Li (PathIntegrator.cpp:15-88).
#include <chrono>
#include <cstring>

#include "adjoint.h"
#include "guide.h"
#include "host.h"
#include "lights.h"
#include "shading.h"
#include "trace.h"

namespace lgpu {



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

// Image-order slot -> pixel: a bijection on [0, w h) that walks the image in 16x16 tiles (row-major
// tiles, each four 8x8 quarters in row-major order, row-major inside a quarter) over the largest
// multiple-of-16 rectangle, then the rest of the pixels in scanline order. A wave's 64 camera rays
// then cover one 8x8 square and a 256-path block one 16x16 square instead of runs of a scanline,
// so they take similar paths through the BVH.
__device__ inline uint32_t TiledPixel(uint32_t j, uint32_t w, uint32_t h) {
  const uint32_t wt = w & ~15u, ht = h & ~15u, nt = wt * ht;
  if (j < nt) {
    const uint32_t tile = j >> 8, quarter = (j >> 6) & 3u, in = j & 63u, tilesX = wt >> 4;
    const uint32_t x = (tile % tilesX) * 16 + (quarter & 1u) * 8 + (in & 7u);
    const uint32_t y = (tile / tilesX) * 16 + (quarter >> 1) * 8 + (in >> 3);
    return y * w + x;
  }
  j -= nt;  // pixels right of the tiled rectangle (rows < ht), then the rows below it
  const uint32_t rw = w - wt;
  if (j < rw * ht) return (j / rw) * w + wt + j % rw;
  j -= rw * ht;
  return ht * w + j;
}

__global__ void __launch_bounds__(kBlock) k_camera(RenderParams rp, PathState ps, SampleState ss, uint32_t n,
                                                   uint64_t base, const uint32_t *listPix, const uint32_t *listSmp) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t pixel, sample;
  if (listPix) {
    pixel = listPix[i];
    sample = listSmp[i];
  } else {
    const uint64_t g = base + i;
    const uint64_t npx = (uint64_t)rp.width * rp.height;
    pixel = TiledPixel((uint32_t)(g % npx), rp.width, rp.height);
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
  ps.ray_o[i] = ToF4(o, lambda0);
  ps.ray_d[i] = ToF4(d, 1.001f);  // w: ScatterEvent::eta initial value (ScatterEvent.h:24)
  ps.beta[i] = make_float4(1.f, 1.f, 1.f, 1.f);
  ps.ids[i] = make_uint4(pixel, sample, rng.dim, kFlagCamera);
  ps.aux[i] = make_uint4(i, kNone, 0u, kNone);
  ss.L[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  ss.pixel[i] = pixel;
  ss.lambda0[i] = lambda0;
}

// Guided renders: the extension stage also lists the queue slots whose hit is a guided vertex
// (guidable material, path still scattering, depth below the guide's limit), for the guide's
// inference before shading. One device atomic per block allocates the block's entries, taken by
// the block's last wave to finish its rays, so no wave waits at a barrier for the others.
struct GuideList {
  uint32_t *list;
  uint32_t *count;  // zeroed before the extension launch
  uint32_t max_depth;
};

struct GuideListLds {
  uint64_t mask[kBlock / 64];  // per wave: its lanes that take an entry
  uint32_t arrived;            // waves done (zeroed, then a barrier, before any wave arrives)
};

__device__ inline void GuideListAppend(bool take, uint32_t blockFirst, const GuideList &gl, GuideListLds &lds) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint64_t m = __ballot(take);
  uint32_t last = 0;
  if (lane == 0) {
    lds.mask[wv] = m;
    __threadfence_block();  // the mask is in LDS before the arrival is counted
    last = atomicAdd(&lds.arrived, 1u) == (uint32_t)(kBlock / 64 - 1);
  }
  if (!__shfl(last, 0)) return;
  __threadfence_block();
  uint64_t mk[kBlock / 64];
  uint32_t total = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    mk[w] = lds.mask[w];
    total += (uint32_t)__popcll(mk[w]);
  }
  if (!total) return;
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(gl.count, total);
  base = __shfl(base, 0);
  const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    if ((mk[w] >> lane) & 1ull) gl.list[base + (uint32_t)__popcll(mk[w] & below)] = blockFirst + w * 64 + lane;
    base += (uint32_t)__popcll(mk[w]);
  }
}

// Kernels of the path loop take their item count either as a value or, when `np` is set, from a
// device counter written by the previous stage (the grid then covers an upper bound of it).
template <bool kCount, bool kList = false>
__global__ void __launch_bounds__(kBlock) k_extend(DevScene sc, PathState ps, uint32_t n, const uint32_t *np,
                                                   uint32_t *spill, unsigned long long *counts, GuideList gl) {
  __shared__ uint32_t stk[kLdsStack][kBlock];
  __shared__ GuideListLds listLds;
  if (np) n = *np;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  // with the list, every wave of a partial block arrives at it; whole idle blocks leave
  if (kList ? blockIdx.x * kBlock >= n : i >= n) return;
  if (kList) {
    if (threadIdx.x == 0) listLds.arrived = 0u;
    __syncthreads();
  }
  bool take = false;
  if (!kList || i < n) {
    const V3 o = ToV3(ps.ray_o[i]), d = ToV3(ps.ray_d[i]);
    float t, u, v;
    uint32_t prim, bi;
    TraceStats ts{0, 0};
    const bool hit = Trace<false, kCount>(sc, o, d, 0.f, INFINITY, &t, &u, &v, &prim, &bi, stk,
                                          spill + (size_t)i * sc.spill_words, &ts);
    ps.hit[i] = hit ? make_float4(t, u, v, __uint_as_float(bi)) : make_float4(INFINITY, 0.f, 0.f, __uint_as_float(kNone));
    if (kList && hit && (sc.tris[bi].key & kKeyGuidable)) {
      const uint32_t w = ps.ids[i].w;
      take = !(w & kFlagEmitOnly) && (w >> 8) < gl.max_depth;
    }
    if (kCount) {
      atomicAdd(&counts[0], (unsigned long long)ts.nodes);
      atomicAdd(&counts[1], (unsigned long long)ts.tris);
      // the same ray through the 4-wide tree the render walks: node and triangle records fetched
      TraceStats t4{0, 0};
      float t2, u2, v2;
      uint32_t p2, b2;
      Trace4Closest<true>(sc, o, d, 0.f, INFINITY, &t2, &u2, &v2, &p2, &b2, stk, spill + (size_t)i * sc.spill_words, &t4);
      atomicAdd(&counts[3], (unsigned long long)t4.nodes);
      atomicAdd(&counts[4], (unsigned long long)t4.tris);
    }
  }
  if (kList) GuideListAppend(take, blockIdx.x * kBlock, gl, listLds);
}

__global__ void __launch_bounds__(kBlock) k_shadow(DevScene sc, SampleState ss, const ShadowRay *rays, uint32_t n,
                                                   const uint32_t *np, uint32_t *spill) {
  __shared__ uint32_t stk[kLdsStack][kBlock];
  if (np) n = *np;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const ShadowRay r = rays[i];
  float t, u, v;
  uint32_t prim, bi;
  const bool occ = Trace<true, false>(sc, ToV3(r.o), ToV3(r.d), 0.f, r.tmax, &t, &u, &v, &prim, &bi, stk,
                                      spill + (size_t)i * sc.spill_words, nullptr);
  if (!occ) {  // one shadow ray per path per bounce: no other writer of this slot runs concurrently
    float4 L = ss.L[r.sid];
    L.x += r.contrib[0]; L.y += r.contrib[1]; L.z += r.contrib[2]; L.w += r.contrib[3];
    ss.L[r.sid] = L;
  }
}

// ---------------------------------------------------------------- shade
__device__ inline void PushShadow(const Queues &q, const ShadowSpec &sh, uint32_t sid, Spec c) {
  const uint32_t slot = atomicAdd(&q.counters[1], 1u);
  ShadowRay r;
  r.o[0] = sh.o.x; r.o[1] = sh.o.y; r.o[2] = sh.o.z; r.tmax = sh.tmax;
  r.d[0] = sh.d.x; r.d[1] = sh.d.y; r.d[2] = sh.d.z; r.sid = sid;
  r.contrib[0] = c.c[0]; r.contrib[1] = c.c[1]; r.contrib[2] = c.c[2]; r.contrib[3] = c.c[3];
  q.shadow[slot] = r;
}

// Queue slots are handed out per block: a device-scope atomic on one counter serialises across
// the whole GPU (about 11 ns each on MI355X, measured), so one per block and queue instead of one
// per wave cuts that cost fourfold. The block's pushes are collected (the extension ray in
// registers, the shadow ray staged in LDS) and written after one allocation at the block's end.
struct ExtPush {
  float4 o, d, beta, mis, pp, pn;
  uint4 ids, aux;
};
__device__ inline void StageShadow(ShadowRay &r, const ShadowSpec &sh, uint32_t sid, Spec c) {
  r.o[0] = sh.o.x; r.o[1] = sh.o.y; r.o[2] = sh.o.z; r.tmax = sh.tmax;
  r.d[0] = sh.d.x; r.d[1] = sh.d.y; r.d[2] = sh.d.z; r.sid = sid;
  r.contrib[0] = c.c[0]; r.contrib[1] = c.c[1]; r.contrib[2] = c.c[2]; r.contrib[3] = c.c[3];
}
// Block-wide allocation of the block's pushes to the two queues; every thread of the block calls
// it (a barrier inside). cnt/bases: LDS scratch of kBlock / 64 + 1 words per queue. q.counters
// must be 8-byte aligned.
__device__ inline void BlockPush(const Queues &q, bool pushE, const ExtPush &e, bool pushS, const ShadowRay *stage,
                                 uint32_t (*cnt)[kBlock / 64], uint32_t *bases) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint64_t mE = __ballot(pushE), mS = __ballot(pushS);
  if (lane == 0) {
    cnt[0][wv] = (uint32_t)__popcll(mE);
    cnt[1][wv] = (uint32_t)__popcll(mS);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t te = 0, ts = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      const uint32_t ce = cnt[0][w], cs = cnt[1][w];
      cnt[0][w] = te;
      cnt[1][w] = ts;
      te += ce;
      ts += cs;
    }
    // both counts in one 64-bit atomic: counters[0] is the low word, counters[1] the high one (a
    // count stays below 2^32, so nothing carries between them)
    if (te | ts) {
      const unsigned long long b =
          atomicAdd(reinterpret_cast<unsigned long long *>(q.counters), ((unsigned long long)ts << 32) | te);
      bases[0] = (uint32_t)b;
      bases[1] = (uint32_t)(b >> 32);
    }
  }
  __syncthreads();
  const uint64_t below = (1ull << lane) - 1ull;
  if (pushE) {
    const uint32_t j = bases[0] + cnt[0][wv] + (uint32_t)__popcll(mE & below);
    q.next.ray_o[j] = e.o;
    q.next.ray_d[j] = e.d;
    q.next.beta[j] = e.beta;
    q.next.ids[j] = e.ids;
    q.next.aux[j] = e.aux;
    q.next.mis[j] = e.mis;
    q.next.prev_p[j] = e.pp;
    q.next.prev_n[j] = e.pn;
  }
  if (pushS) q.shadow[bases[1] + cnt[1][wv] + (uint32_t)__popcll(mS & below)] = stage[threadIdx.x];
}

// kCorrected: the estimator mode (rp.corrected) as a template argument, so each instantiation
// carries only its own mode's branches.
// kPrograms: the scene has program sockets (DevScene::has_programs); without them the kernel is
// compiled without the shader VM, spills little, and runs best at four waves per SIMD (with the
// VM's call frames and call-site spills, six).
// Guide table of one path slot: GuideBin / sampling over the 128 bin weights (sum `tot`).
__device__ inline float GuideBinPdf(const uint8_t *tab, float tot, uint32_t bin) {
  return (float)tab[bin] / (tot * kGuideBinSr);
}
// The bin whose cumulative weight first exceeds target (in [0, tot)): 16 bins per 16-byte load,
// byte sums on v_dot4_u32_u8; the sums are exact integers, so the comparisons are those of a
// running float sum.
__device__ inline uint32_t GuideSampleBin(const uint32_t *tab, float target) {
  const uint4 *t4 = reinterpret_cast<const uint4 *>(tab);  // 128-byte tables
  uint32_t acc = 0;
  for (uint32_t k = 0; k < kGuideBins / 16; ++k) {
    const uint4 q = t4[k];
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
    uint32_t s[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) s[j] = __builtin_amdgcn_udot4(w[j], 0x01010101u, 0u, false);
    if ((float)(acc + s[0] + s[1] + s[2] + s[3]) > target) {
      uint32_t a = acc;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        if ((float)(a + s[j]) > target) {
          const uint32_t b0 = w[j] & 255u, b1 = (w[j] >> 8) & 255u, b2 = (w[j] >> 16) & 255u;
          const uint32_t bin = 16 * k + 4 * j;
          if ((float)(a + b0) > target) return bin;
          if ((float)(a + b0 + b1) > target) return bin + 1;
          if ((float)(a + b0 + b1 + b2) > target) return bin + 2;
          return bin + 3;
        }
        a += s[j];
      }
      const uint32_t b0 = w[3] & 255u, b1 = (w[3] >> 8) & 255u, b2 = (w[3] >> 16) & 255u;
      if ((float)(a + b0) > target) return 16 * k + 12;
      if ((float)(a + b0 + b1) > target) return 16 * k + 13;
      if ((float)(a + b0 + b1 + b2) > target) return 16 * k + 14;
      return 16 * k + 15;
    }
    acc += s[0] + s[1] + s[2] + s[3];
  }
  return kGuideBins - 1;  // target rounded up to tot
}
__device__ inline uint32_t Hash3(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u ^ (c + 0x165667B1u) * 0xC2B2AE3Du;
  h ^= h >> 15; h *= 0x2C1B3C6Du; h ^= h >> 12; h *= 0x297A2D39u; h ^= h >> 15;
  return h;
}

// kGuided (corrected estimator only): guidable vertices sample one-sample MIS of the learned
// guide (the path slot's bin table, inferred after extension) and the BSDF, evaluate both
// strategies' pdfs for MIS with next-event estimation, and append training records.
template <bool kCorrected, bool kPrograms, bool kGuided>
__global__ void __launch_bounds__(kBlock, kPrograms ? 6 : 4) k_shade(const DevScene *__restrict__ scp, RenderParams rp, PathState ps, SampleState ss,
                                                  uint32_t n, const uint32_t *np, Queues q, GuideRefs gr) {
  const DevScene &sc = *scp;
  __shared__ TreeImp treeLds[kTreeLds];
  __shared__ uint32_t keyStart[kShadeKeys];
  __shared__ uint16_t perm[kBlock];
  __shared__ ShadowRay shStage[kBlock];
  __shared__ uint32_t pushCnt[2][kBlock / 64];
  __shared__ uint32_t pushBase[2];
  if (np) n = *np;
  const uint32_t base = blockIdx.x * kBlock;
  if (base >= n) return;  // whole block idle: skip the LDS tree load
  const uint32_t m = min(n - base, (uint32_t)kBlock);
  if (threadIdx.x < kShadeKeys) keyStart[threadIdx.x] = 0u;
  LoadTreeLds(sc, treeLds);  // ends with a barrier
  // Group the block's paths by shading family (counting sort in LDS): each thread then shades the
  // path at its sorted position, so a wave mostly runs one family's BSDF and shader code. The
  // gathers stay inside the block's own 256 consecutive paths.
  uint32_t key = 0, rank = 0;
  if (threadIdx.x < m) {
    const uint32_t hb = __float_as_uint(ps.hit[base + threadIdx.x].w);
    key = hb == kNone ? 0u : (sc.tris[hb].key & (kShadeKeys - 1));
    rank = atomicAdd(&keyStart[key], 1u);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (uint32_t k = 0; k < kShadeKeys; ++k) {
      const uint32_t c = keyStart[k];
      keyStart[k] = acc;
      acc += c;
    }
  }
  __syncthreads();
  if (threadIdx.x < m) perm[keyStart[key] + rank] = (uint16_t)threadIdx.x;
  __syncthreads();
  bool pushE = false, pushS = false;
  ExtPush ext;
  // the path's shading; `break` ends it, every thread then takes part in the block's queue pushes
  do {
  if (threadIdx.x >= m) break;
  const uint32_t i = base + perm[threadIdx.x];
  constexpr bool corrected = kCorrected;
  const int mode = (int)rp.mode;
  const uint4 ids = ps.ids[i];
  const uint4 aux = ps.aux[i];
  const uint32_t sid = aux.x;
  uint32_t flags = ids.w & 0xFFu;
  const uint32_t depth = ids.w >> 8;
  const float4 ro4 = ps.ray_o[i], rd4 = ps.ray_d[i];
  const V3 ro = ToV3(ro4), rd = ToV3(rd4);
  const float4 hr = ps.hit[i];
  const uint32_t bi = __float_as_uint(hr.w);
  const bool hit = bi != kNone;
  Spec beta = S4(ps.beta[i]);
  Spec Ladd = S(0.f);
  bool addL = false;
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

  // ---- emission reaching this vertex
  if (flags & kFlagCamera) {
    if (!hit) {
      if (sc.has_env) { Ladd = EnvLe(sc, rd, vm); addL = true; }  // PathIntegrator.cpp:82-85
      alive = false;
    } else {
      const MeshRec m = sc.meshes[h.mesh];
      if (m.mesh_light != kNone && (!corrected || Dot(rd, h.normalG) < 0)) {  // :26
        Ladd = EmissionOf<kPrograms>(sc, h.mesh, vm) * sc.lights[m.mesh_light].intensity * kInvPi;
        addL = true;
      }
    }
  } else if (!(flags & kFlagNoEmit)) {
    // BSDF-sampled light with MIS (PathIntegrator.cpp:52-63)
    const float4 mis = ps.mis[i];
    uint32_t nl = kNone;
    if (hit) {
      const uint32_t ml = sc.meshes[h.mesh].mesh_light;
      if (ml != kNone) nl = sc.strategy == LGPU_LIGHTS_TREE ? (kTreeBit | sc.tri_tlight[h.prim]) : ml;
    } else if (sc.has_env) {
      nl = sc.env_light;
    }
    if (nl != kNone) {
      const uint32_t l = aux.y;
      bool same = nl == l;
      if (!corrected && sc.strategy == LGPU_LIGHTS_TREE && (nl & kTreeBit)) same = false;  // MeshLight != TriangleLight
      float lightDistPdf = mis.y;
      if (!same) {
        if (corrected) lightDistPdf = LightPdf(sc, nl, ToV3(ps.prev_p[i]), ToV3(ps.prev_n[i]), true);
        else lightDistPdf = LightPdf(sc, nl, h.point, h.normalS, false);  // evaluated at the light hit as written
      }
      float pl;
      Spec Le;
      if (hit) {
        const LightRec &NL = LightOf(sc, nl);
        if (!corrected && sc.strategy == LGPU_LIGHTS_TREE) pl = PdfLiHit(sc, sc.lights[NL.parent], rd, h, bi, false);
        else pl = PdfLiHit(sc, NL, rd, h, bi, corrected);
        Le = EmissionOf<kPrograms>(sc, h.mesh, vm) * sc.lights[sc.meshes[h.mesh].mesh_light].intensity * kInvPi;
        if (corrected && Dot(rd, h.normalG) > 0) Le = S(0.f);
      } else {
        pl = EnvPdfLi(sc, rd);
        Le = EnvLe(sc, rd, vm);
      }
      const float lightPDF = lightDistPdf * pl;
      const float scatPdf = mis.x;
      const float w = (corrected && (flags & kFlagSpecular)) ? 1.f : PowerHeuristic(scatPdf, lightPDF);
      if (!IsBlack(Le, mode) && (lightPDF > 0 || corrected)) { Ladd = beta * (Le * w); addL = true; }
    }
    if ((flags & kFlagEmitOnly) || !hit) alive = false;
    if (alive && (flags & kFlagRR)) beta = beta / S(1.f - mis.z);  // PathIntegrator.cpp:79
  } else if (!hit) {
    alive = false;
  }
  if (kGuided && !(flags & kFlagCamera)) {
    // the previous guided vertex's record: radiance so far, before anything along its wi lands
    const uint32_t r = __float_as_uint(ps.mis[i].w);
    if (r < gr.rec_cap) gr.rec[r].c.x = Luminance(S4(ss.L[sid]), mode);
  }
  if (addL) {
    float4 L = ss.L[sid];
    L.x += Ladd.c[0]; L.y += Ladd.c[1]; L.z += Ladd.c[2]; L.w += Ladd.c[3];
    ss.L[sid] = L;
  }
  if (!alive) break;

  const MaterialRec mat = sc.materials[sc.meshes[h.mesh].material];
  if (mat.lobe == kNone) {
    // pass-through surface (PathIntegrator.cpp:69-74)
    const uint32_t pass = aux.z + 1;
    if (pass > 1024) break;
    const V3 o2 = h.point + h.normalG * (Dot(rd, h.normalG) < 0 ? -kSurfaceEpsilon : kSurfaceEpsilon);
    flags = (flags & kFlagCamera) ? kFlagCamera : kFlagNoEmit;
    ext.o = ToF4(o2, l0);
    ext.d = rd4;
    ext.beta = F4(beta);
    ext.ids = make_uint4(ids.x, ids.y, rng.dim, (depth << 8) | flags);
    ext.aux = make_uint4(sid, aux.y, pass, aux.w);
    float4 pm = ps.mis[i];
    if (kGuided) pm.w = __uint_as_float(kNone);  // record closed above
    ext.mis = pm;
    ext.pp = ps.prev_p[i];
    ext.pn = ps.prev_n[i];
    pushE = true;
    break;
  }

  // ---- shading vertex (PathIntegrator.cpp:28-80)
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
  const bool guided = kGuided && (mat.pad[0] & 1u) && depth < gr.max_depth;
  float gc = 0.f, gtot = 1.f;
  const uint8_t *gtab = nullptr;
  if (guided) {
    const float2 gh = gr.hdr[i];
    gc = gh.x;
    gtot = gh.y;
    gtab = reinterpret_cast<const uint8_t *>(gr.bins + (size_t)i * 8);
  }
  float lightDistPdf = 1.f;
  const uint32_t pick = SampleLight(sc, h.point, h.normalS, rng, &lightDistPdf, treeLds, corrected);
  float lightPDF = 0.f;
  Spec Li = S(0.f);
  ShadowSpec sh;
  bool delta = false;
  if (pick != kNone) {
    const LightRec &LL = LightOf(sc, pick);
    delta = LL.kind == kLightPoint || LL.kind == kLightSpot;
    Li = SampleLi<false, kPrograms>(sc, LL, h, ev, rng, vm, corrected, lightPDF, sh, nullptr, (pick & kTreeBit) ? (pick & ~kTreeBit) : kNone);
  }
  lightPDF *= lightDistPdf;
  bool hasNee = false;
  Spec cNee = S(0.f);
  if (lightPDF > 0 && !IsBlack(Li, mode)) {
    const Spec f = shd.F(mat.lobe, ev) * fabsf(ev.wiL.y);
    const float sp = shd.Pdf(mat.lobe, ev.woL, ev.wiL);
    if (!IsBlack(f, mode) && sp > 0) {
      // with a guide the BSDF-side strategy is the guide/BSDF mixture
      const float spm = guided ? gc * GuideBinPdf(gtab, gtot, GuideBin(ev.wi)) + (1.f - gc) * sp : sp;
      const float w = (corrected && delta) ? 1.f : PowerHeuristic(lightPDF, spm);
      cNee = Li * f * w / lightPDF;
      hasNee = true;
    }
  }
  // Corrected: the NEE estimate stands whatever the BSDF sample does, so its shadow ray leaves
  // now (its state is then dead during the BSDF sample). Faithful: the reference's `break` on a
  // failed BSDF sample drops it, so it waits for the sample.
  if (corrected && hasNee) {
    StageShadow(shStage[threadIdx.x], sh, sid, beta * cNee);
    pushS = true;
  }
  float scatPdf = 0.f;
  bool spec = false;
  Spec f;
  uint32_t gbin = kNone;
  float pbsdf = 0.f;
  if (guided) {
    // one-sample MIS: the guide with probability c, else the BSDF; pdf c q + (1 - c) p_bsdf.
    // Directions the BSDF cannot sample (p_bsdf = 0) are rejected, as BSDF sampling never
    // reaches them either. Either strategy only draws the direction; the BSDF value and pdf are
    // evaluated once after the two branches join.
    bool drawn = true;
    if (rng.Get1D() < gc) {
      gbin = GuideSampleBin(reinterpret_cast<const uint32_t *>(gtab), rng.Get1D() * gtot);
      const float u0 = rng.Get1D(), u1 = rng.Get1D();
      ev.wi = GuideBinDirection(gbin, u0, u1);
      ev.wiL = ToLocal(h, ev.wi);
    } else {
      drawn = shd.SampleDir(mat.lobe, ev, h, rng);
      gbin = GuideBin(ev.wi);
    }
    pbsdf = drawn ? shd.Pdf(mat.lobe, ev.woL, ev.wiL) : 0.f;
    f = pbsdf > 0 ? shd.F(mat.lobe, ev) : S(0.f);
    h.point = h.point + h.normalG * (Dot(ev.wi, h.normalG) > 0 ? kSurfaceEpsilon : -kSurfaceEpsilon);
    scatPdf = pbsdf > 0 ? gc * GuideBinPdf(gtab, gtot, gbin) + (1.f - gc) * pbsdf : 0.f;
  } else {
    f = shd.Sample(mat.lobe, ev, h, rng, scatPdf, spec);
  }
  f = f * fabsf(ev.wiL.y);
  if (!(scatPdf > 0 && !IsBlack(f, mode))) break;
  if (!corrected && hasNee) {
    StageShadow(shStage[threadIdx.x], sh, sid, beta * cNee);
    pushS = true;
  }
  uint32_t recId = kNone;
  if (guided && Hash3(ids.x, ids.y, depth) <= gr.rec_threshold) {
    const uint32_t slot = atomicAdd(gr.rec_count, 1u);
    if (slot < gr.rec_cap) {
      float s5[kGuideScalars];
      GuideInputs(shadeP, -rd, gr.bmin, gr.binv, s5);
      GuideRecord R;
      R.a = make_float4(s5[0], s5[1], s5[2], s5[3]);
      R.b = make_float4(s5[4], __uint_as_float(gbin), pbsdf, __uint_as_float(sid));
      R.c = make_float4(__uint_as_float(0x7FC00000u), Luminance(beta, mode), 0.f, 0.f);
      gr.rec[slot] = R;
      recId = slot;
    }
  }
  beta = beta * (f / scatPdf);
  uint32_t nf = spec ? kFlagSpecular : 0u;
  float qrr = 0.f;
  if (depth > rp.min_bounces) {  // Russian roulette (PathIntegrator.cpp:76-80)
    qrr = fmaxf(.05f, 1 - Luminance(beta, mode));
    if (rng.Get1D() < qrr) nf |= kFlagEmitOnly;
    else nf |= kFlagRR;
  }
  if (depth + 1 >= rp.max_bounces) nf |= kFlagEmitOnly;
  ext.o = ToF4(h.point, l0);
  ext.d = ToF4(ev.wi, ev.eta);
  ext.beta = F4(beta);
  ext.ids = make_uint4(ids.x, ids.y, rng.dim, ((depth + 1) << 8) | nf);
  ext.aux = make_uint4(sid, pick, aux.z, aux.w);
  ext.mis = make_float4(scatPdf, lightDistPdf, qrr, kGuided ? __uint_as_float(recId) : 0.f);
  ext.pp = ToF4(shadeP, 0.f);
  ext.pn = ToF4(shadeN, 0.f);
  pushE = true;
  } while (false);
  BlockPush(q, pushE, ext, pushS, shStage, pushCnt, pushBase);
}

// Film::AddSample (Film.h:30-33) in FP64; hero wavelengths go through the CIE fit to XYZ then sRGB.
__device__ inline void SampleRGB(const RenderParams &rp, float4 Lf, float l0, double rgb[3]) {
  if (rp.mode) {
    const float Lc[4] = {Lf.x, Lf.y, Lf.z, Lf.w};
    double xyz[3] = {0, 0, 0};
    for (int k = 0; k < 4; ++k) {
      float l = l0 + 117.5f * (float)k;
      if (l >= 830.f) l -= 470.f;
      float cm[3];
      CMF<true>(l, cm);
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
}

// Per-sample film update for list-mode waves (any pixel may repeat): FP64 atomics, or the raw
// per-sample record when `out` is set.
__global__ void __launch_bounds__(kBlock) k_accumulate(RenderParams rp, SampleState ss, uint32_t n, double *film,
                                                       uint32_t *counts, float *out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 Lf = ss.L[i];
  const float l0 = ss.lambda0[i];
  if (out) {
    out[5 * (size_t)i + 0] = Lf.x;
    out[5 * (size_t)i + 1] = Lf.y;
    out[5 * (size_t)i + 2] = Lf.z;
    out[5 * (size_t)i + 3] = rp.mode ? Lf.w : 0.f;
    out[5 * (size_t)i + 4] = l0;
    return;
  }
  double rgb[3];
  SampleRGB(rp, Lf, l0, rgb);
  const uint32_t px = ss.pixel[i];
  unsafeAtomicAdd(&film[3 * (size_t)px + 0], rgb[0]);
  unsafeAtomicAdd(&film[3 * (size_t)px + 1], rgb[1]);
  unsafeAtomicAdd(&film[3 * (size_t)px + 2], rgb[2]);
  atomicAdd(&counts[px], 1u);
}

// Film update of an image-order wave: slot i holds global sample g = base + i, pixel g mod npx,
// so a pixel's samples in the wave sit npx slots apart. Thread t owns slot t's pixel, sums that
// pixel's samples of the wave in sample order and updates its film entry once: each pixel has
// one owner per launch (t < min(n, npx)), so no atomics are needed and the sum order is fixed.
__global__ void __launch_bounds__(kBlock) k_accumulate_px(RenderParams rp, SampleState ss, uint32_t n, uint32_t npx,
                                                          double *film, uint32_t *counts) {
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= n || t >= npx) return;
  double acc[3] = {0, 0, 0};
  uint32_t c = 0;
  for (uint32_t i = t; i < n; i += npx) {
    double rgb[3];
    SampleRGB(rp, ss.L[i], ss.lambda0[i], rgb);
    acc[0] += rgb[0]; acc[1] += rgb[1]; acc[2] += rgb[2];
    ++c;
    if (npx > n - i) break;  // next slot would pass n (also guards i + npx overflow)
  }
  const uint32_t px = ss.pixel[t];
  double *f = film + 3 * (size_t)px;
  f[0] += acc[0]; f[1] += acc[1]; f[2] += acc[2];
  counts[px] += c;
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
  uint32_t prim, bi;
  uint32_t *sp = spill + (size_t)i * sc.spill_words;
  if (occ) {
    occ[i] = Trace<true, false>(sc, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), a.w, b.w, &t, &u, &v, &prim, &bi, stk, sp, nullptr) ? 1u : 0u;
  } else if (Trace<false, false>(sc, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), a.w, b.w, &t, &u, &v, &prim, &bi, stk, sp, nullptr)) {
    hits[i] = make_uint4(__float_as_uint(t), __float_as_uint(u), __float_as_uint(v), prim);
  } else {
    hits[i] = make_uint4(__float_as_uint(INFINITY), 0u, 0u, kNone);
  }
}

// ---------------------------------------------------------------- host side
static void LaunchShade(bool corrected, bool programs, const GuideRefs *gr, uint32_t blocks, hipStream_t stream,
                        const DevScene *scp, const RenderParams &rp, const PathState &ps, const SampleState &ss, uint32_t n,
                        const uint32_t *np, const Queues &q) {
  const dim3 g(blocks), b(kBlock);
  const GuideRefs none{};
  if (gr && programs) hipLaunchKernelGGL((k_shade<true, true, true>), g, b, 0, stream, scp, rp, ps, ss, n, np, q, *gr);
  else if (gr) hipLaunchKernelGGL((k_shade<true, false, true>), g, b, 0, stream, scp, rp, ps, ss, n, np, q, *gr);
  else if (corrected && programs) hipLaunchKernelGGL((k_shade<true, true, false>), g, b, 0, stream, scp, rp, ps, ss, n, np, q, none);
  else if (corrected) hipLaunchKernelGGL((k_shade<true, false, false>), g, b, 0, stream, scp, rp, ps, ss, n, np, q, none);
  else if (programs) hipLaunchKernelGGL((k_shade<false, true, false>), g, b, 0, stream, scp, rp, ps, ss, n, np, q, none);
  else hipLaunchKernelGGL((k_shade<false, false, false>), g, b, 0, stream, scp, rp, ps, ss, n, np, q, none);
}

// ---------------------------------------------------------------- batched renders (batch.hip)
// Counter slots of one job's chunk: slot 0 = the camera's path count, the rest cleared (inside a
// captured graph, where memset nodes are avoided).
__global__ void k_batch_init(uint32_t *ctr, uint32_t words, uint32_t n) {
  for (uint32_t k = threadIdx.x; k < words; k += blockDim.x) ctr[k] = k == 0 ? n : 0u;
}
// Paths still queued after a chunk's last captured iteration (must stay 0).
__global__ void k_batch_trunc(const uint32_t *cnt, unsigned long long *acc) {
  if (threadIdx.x == 0 && *cnt) atomicAdd(acc, (unsigned long long)*cnt);
}
void BatchLaunchInit(hipStream_t s, uint32_t *ctr, uint32_t words, uint32_t n) {
  hipLaunchKernelGGL(k_batch_init, dim3(1), dim3(64), 0, s, ctr, words, n);
}
void BatchLaunchTrunc(hipStream_t s, const uint32_t *cnt, unsigned long long *acc) {
  hipLaunchKernelGGL(k_batch_trunc, dim3(1), dim3(64), 0, s, cnt, acc);
}
void BatchLaunchCamera(hipStream_t s, const RenderParams &rp, const PathState &ps, const SampleState &ss, uint32_t n,
                       uint64_t base) {
  hipLaunchKernelGGL(k_camera, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, rp, ps, ss, n, base, nullptr, nullptr);
}
void BatchLaunchExtend(hipStream_t s, const DevScene &dsc, const PathState &ps, uint32_t ub, const uint32_t *np,
                       uint32_t *spill) {
  hipLaunchKernelGGL(k_extend<false>, dim3((ub + kBlock - 1) / kBlock), dim3(kBlock), 0, s, dsc, ps, ub, np, spill, nullptr,
                     GuideList{});
}
void BatchLaunchShade(hipStream_t s, const DevScene *scp, const RenderParams &rp, const PathState &ps,
                      const SampleState &ss, uint32_t ub, const uint32_t *np, const Queues &q, bool programs) {
  LaunchShade(true, programs, nullptr, (ub + kBlock - 1) / kBlock, s, scp, rp, ps, ss, ub, np, q);
}
void BatchLaunchShadow(hipStream_t s, const DevScene &dsc, const SampleState &ss, const ShadowRay *rays, uint32_t ub,
                       const uint32_t *np, uint32_t *spill) {
  hipLaunchKernelGGL(k_shadow, dim3((ub + kBlock - 1) / kBlock), dim3(kBlock), 0, s, dsc, ss, rays, ub, np, spill);
}
void BatchLaunchAccumulate(hipStream_t s, const RenderParams &rp, const SampleState &ss, uint32_t n, uint32_t npx,
                           double *film, uint32_t *counts) {
  const uint32_t m = n < npx ? n : npx;
  hipLaunchKernelGGL(k_accumulate_px, dim3((m + kBlock - 1) / kBlock), dim3(kBlock), 0, s, rp, ss, n, npx, film, counts);
}

// Guiding hooks (guide.hip).
lgpu_status GuidePrepareWave(lgpu_guide g, const float bounds[6], uint64_t expectedVertices, GuideRefs *refs,
                             const uint4 *bins, const float2 *hdr, hipStream_t s);
lgpu_status GuideInfer(lgpu_guide g, const GuideRefs &refs, const PathState &ps, uint32_t ub, const uint32_t *list,
                       const uint32_t *count, uint4 *bins, float2 *hdr, hipStream_t s);
lgpu_status GuideFinishWave(lgpu_guide g, const SampleState &ss, int mode, uint32_t trainSteps, uint64_t vertices,
                            hipStream_t s, uint32_t *records);
uint64_t GuideExpectedVertices(lgpu_guide g);
bool GuideOnDevice(lgpu_guide g, lgpu_device dev);

struct Wave {
  // two path-state buffers (ping-pong), per-sample radiance, shadow queue
  DevBuf st[2][9], L, pix, lam, shadow, counters, spill, tcount, scene, ring;
  DevBuf gbins, ghdr;  // guided renders: per path slot, the guide's bin table and {c, sum}
  DevBuf ltot;         // adjoint renders: per path slot, the primal pass's radiance
  DevBuf glist, gcount; // guided renders: the bounce's queue slots that shade a guided vertex
  uint32_t cap = 0;
  PathState ps[2]{};
  SampleState ss{};
  // the loop's views of the buffers: into the DevBufs above (alloc) or into a caller-provided
  // region (bind, the no-allocation entry point)
  ShadowRay *shadow_p = nullptr;
  uint32_t *counters_p = nullptr, *spill_p = nullptr;
  unsigned long long *tcount_p = nullptr;
  DevScene *scene_p = nullptr;
  // pipelined path loop: device counter slots, their pinned host copies, per-slot stage events
  static constexpr int kSlots = 4;
  uint32_t *hring = nullptr;
  hipEvent_t pev[kSlots][6] = {};
  ~Wave() {
    if (hring) (void)hipHostFree(hring);
    for (auto &row : pev)
      for (hipEvent_t &e : row)
        if (e) (void)hipEventDestroy(e);
  }
  // Device bytes one path in flight holds (both state buffers, sample state, shadow slot, spill;
  // guided: the bin table and header).
  static uint64_t BytesPerPath(uint32_t spillWords, bool guided = false, bool adjoint = false) {
    return 2ull * 9 * 16 + 16 + 4 + 4 + sizeof(ShadowRay) + 4ull * spillWords + (guided ? 140ull : 0ull) +
           (adjoint ? 16ull : 0ull);
  }
  uint64_t bytes() const {
    uint64_t b = L.bytes + pix.bytes + lam.bytes + shadow.bytes + spill.bytes + gbins.bytes + ghdr.bytes + glist.bytes +
                 ltot.bytes;
    for (const auto &row : st)
      for (const DevBuf &x : row) b += x.bytes;
    return b;
  }
  lgpu_status alloc(uint32_t n, uint32_t spillWords, bool guided = false) {
    LGPU_HIP(spill.alloc(std::max<size_t>(256, 4ull * spillWords * n)));
    if (guided) {
      LGPU_HIP(gbins.alloc(128ull * n));
      LGPU_HIP(ghdr.alloc(8ull * n));
      LGPU_HIP(glist.alloc(4ull * n));
      LGPU_HIP(gcount.alloc(256));
    }
    if (!hring) {
      LGPU_HIP(hipHostMalloc((void **)&hring, 16 * kSlots, hipHostMallocDefault));
      LGPU_HIP(ring.alloc(16 * kSlots));
      for (auto &row : pev)
        for (hipEvent_t &e : row) LGPU_HIP(hipEventCreate(&e));
    }
    if (n <= cap) return LGPU_OK;
    for (int b = 0; b < 2; ++b) {
      for (int k = 0; k < 9; ++k) LGPU_HIP(st[b][k].alloc(16ull * n));
      PathState &p = ps[b];
      p.ray_o = st[b][0].as<float4>(); p.ray_d = st[b][1].as<float4>(); p.beta = st[b][2].as<float4>();
      p.hit = st[b][3].as<float4>(); p.ids = st[b][4].as<uint4>(); p.aux = st[b][5].as<uint4>();
      p.mis = st[b][6].as<float4>(); p.prev_p = st[b][7].as<float4>(); p.prev_n = st[b][8].as<float4>();
    }
    LGPU_HIP(L.alloc(16ull * n)); LGPU_HIP(pix.alloc(4ull * n)); LGPU_HIP(lam.alloc(4ull * n));
    LGPU_HIP(shadow.alloc(sizeof(ShadowRay) * (size_t)n));
    LGPU_HIP(counters.alloc(64)); LGPU_HIP(tcount.alloc(64)); LGPU_HIP(scene.alloc(sizeof(DevScene)));
    ss.L = L.as<float4>(); ss.pixel = pix.as<uint32_t>(); ss.lambda0 = lam.as<float>();
    shadow_p = shadow.as<ShadowRay>(); counters_p = counters.as<uint32_t>(); spill_p = spill.as<uint32_t>();
    tcount_p = tcount.as<unsigned long long>(); scene_p = scene.as<DevScene>();
    cap = n;
    return LGPU_OK;
  }
  // Bytes a wave of n paths needs in one region (bind): 256-B aligned parts.
  static uint64_t RegionBytes(uint64_t n, uint32_t spillWords) {
    auto up = [](uint64_t b) { return (b + 255) & ~255ull; };
    return 18 * up(16 * n) + up(16 * n) + 2 * up(4 * n) + up(sizeof(ShadowRay) * n) + up(4ull * spillWords * n) +
           3 * 256 + up(sizeof(DevScene));
  }
  // Lays a wave of n paths out in [base, base + RegionBytes(n)) without allocating; base is a
  // device address (256-B aligned).
  void bind(uint8_t *base, uint32_t n, uint32_t spillWords) {
    auto up = [](uint64_t b) { return (b + 255) & ~255ull; };
    uint8_t *p = base;
    auto take = [&](uint64_t b) { uint8_t *r = p; p += up(b); return r; };
    for (int b = 0; b < 2; ++b) {
      PathState &q = ps[b];
      q.ray_o = (float4 *)take(16ull * n); q.ray_d = (float4 *)take(16ull * n); q.beta = (float4 *)take(16ull * n);
      q.hit = (float4 *)take(16ull * n); q.ids = (uint4 *)take(16ull * n); q.aux = (uint4 *)take(16ull * n);
      q.mis = (float4 *)take(16ull * n); q.prev_p = (float4 *)take(16ull * n); q.prev_n = (float4 *)take(16ull * n);
    }
    ss.L = (float4 *)take(16ull * n); ss.pixel = (uint32_t *)take(4ull * n); ss.lambda0 = (float *)take(4ull * n);
    shadow_p = (ShadowRay *)take(sizeof(ShadowRay) * (uint64_t)n);
    spill_p = (uint32_t *)take(4ull * spillWords * n);
    counters_p = (uint32_t *)take(256);
    tcount_p = (unsigned long long *)take(256);
    (void)take(256);  // spare
    scene_p = (DevScene *)take(sizeof(DevScene));
    cap = n;
  }
};

static inline uint32_t Blocks(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

static lgpu_status ValidateRender(const lgpu_render_desc *d) {
  if (!d || !d->scene) return Fail(LGPU_ERR_INVALID_ARGUMENT, "render: null descriptor or scene");
  if (!d->scene->committed) return Fail(LGPU_ERR_STATE, "render: scene not committed");
  if (d->integrator > LGPU_INTEGRATOR_AOV) return Fail(LGPU_ERR_UNSUPPORTED, "render: unknown integrator");
  if (d->integrator == LGPU_INTEGRATOR_DEPTH && !(d->pass_max_depth > 0.f))
    return Fail(LGPU_ERR_INVALID_ARGUMENT, "render: depth pass needs pass_max_depth > 0");
  if (d->spectral_mode > LGPU_SPECTRAL_HERO4 || d->estimator_mode > LGPU_ESTIMATOR_CORRECTED)
    return Fail(LGPU_ERR_INVALID_ARGUMENT, "render: bad spectral or estimator mode");
  if (d->max_bounces == 0) return Fail(LGPU_ERR_INVALID_ARGUMENT, "render: max_bounces must be > 0");
  if (d->guide && (d->integrator != LGPU_INTEGRATOR_PATH || d->estimator_mode != LGPU_ESTIMATOR_CORRECTED))
    return Fail(LGPU_ERR_UNSUPPORTED, "render: guiding needs the path integrator and the corrected estimator");
  if (d->guide && !GuideOnDevice(d->guide, d->scene->dev))
    return Fail(LGPU_ERR_INVALID_ARGUMENT, "render: the guide belongs to another (or a destroyed) device");
  if (d->camera.type > LGPU_CAMERA_SPHERICAL) return Fail(LGPU_ERR_INVALID_ARGUMENT, "render: bad camera type");
  if (d->camera.aperture == LGPU_APERTURE_BLADE && d->camera.blades == 0)
    return Fail(LGPU_ERR_INVALID_ARGUMENT, "render: blade aperture needs blades > 0");
  return LGPU_OK;
}

RenderParams MakeParams(const lgpu_render_desc *d, uint32_t w, uint32_t h) {
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
  rp.pass_max_depth = d->pass_max_depth; rp.pass_power = d->pass_power;
  for (int k = 0; k < 3; ++k) rp.pass_bg[k] = d->pass_background[k];
  rp.pass_normal_g = d->pass_geometric_normal;
  return rp;
}

// Renders `total` paths; list mode when listPix != nullptr (trace_paths). `bound`: a wave already
// laid out in caller memory (no-allocation entry point): no device or pinned allocation happens and
// the bounce loop synchronises per bounce instead of running ahead through pinned counter copies.
// `adj`: adjoint render (image-order waves, volumetric integrator): every wave is traced twice with
// the same random numbers, the primal pass (film accumulation, per-path radiance kept) and the
// adjoint replay that scatters the gradients.
static lgpu_status RunWaves(const lgpu_render_desc *d, uint32_t w, uint32_t h, uint64_t total, const uint32_t *listPix,
                            const uint32_t *listSmp, double *film, uint32_t *counts, float *out, lgpu_render_stats *stats,
                            Wave *bound = nullptr, const AdjRefs *adj = nullptr) {
  lgpu_scene s = d->scene;
  lgpu_device dev = s->dev;
  LGPU_HIP(hipSetDevice(dev->device));
  (void)hipGetLastError();  // clear a sticky error left by an earlier call
  const auto t0 = std::chrono::steady_clock::now();
  uint64_t cap;
  if (bound) {
    cap = bound->cap;
  } else {
    if (!dev->wave) {  // path pool kept on the device object and reused across calls
      dev->wave = new Wave();
      dev->wave_free = [](void *p) { delete static_cast<Wave *>(p); };
    }
    Wave &W0 = *static_cast<Wave *>(dev->wave);
    // Paths in flight: the caller's figure, else up to kDefaultPaths, held to a quarter of the device
    // memory that is free or already in the pool. Fewer, larger waves spend less time in the
    // launch-bound tails of their bounce loops.
    cap = d->path_capacity;
    if (!cap) {
      constexpr uint64_t kDefaultPaths = 1ull << 27;
      cap = kDefaultPaths;
      size_t freeB = 0, totalB = 0;
      if (hipMemGetInfo(&freeB, &totalB) == hipSuccess) {
        const uint64_t fit =
            (freeB + W0.bytes()) / 4 / Wave::BytesPerPath(s->dsc.spill_words, d->guide != nullptr, adj != nullptr);
        cap = std::min<uint64_t>(cap, std::max<uint64_t>(fit, 1ull << 16));
      }
    }
    // the default pool holds no more than this render's samples; an explicit capacity sizes the
    // pool as asked, so a small render can prepare it for a larger one
    if (!d->path_capacity && cap > total) cap = total;
    if (cap == 0) cap = 1;
    lgpu_status st = W0.alloc((uint32_t)cap, s->dsc.spill_words, d->guide != nullptr);
    if (st != LGPU_OK) return st;
    if (adj && W0.ltot.bytes < 16ull * cap) LGPU_HIP(W0.ltot.alloc(16ull * cap));
  }
  Wave &W = bound ? *bound : *static_cast<Wave *>(dev->wave);
  AdjRefs adjW{};
  if (adj) {
    if (bound) return Fail(LGPU_ERR_UNSUPPORTED, "render: no adjoint in caller memory");
    adjW = *adj;
    adjW.ltot = W.ltot.as<float4>();
  }
  const RenderParams rp = MakeParams(d, w, h);
  const bool countTrav = (d->flags & LGPU_RENDER_COUNT_TRAVERSAL) != 0;
  DevScene dsc = s->dsc;
  if (countTrav) dsc.voxel_count = W.tcount_p + 2;
  hipStream_t stream = dev->stream;
  // The shading kernels read the scene record through a device pointer: a by-value kernel argument
  // whose address reaches out-of-line functions is copied into every lane's scratch.
  LGPU_HIP(hipMemcpyAsync(W.scene_p, &dsc, sizeof(DevScene), hipMemcpyHostToDevice, stream));
  const DevScene *dscp = W.scene_p;
  const bool progs = dsc.has_programs != 0;
  hipEvent_t *ev = dev->ev;
  lgpu_render_stats S;
  std::memset(&S, 0, sizeof(S));
  if (countTrav) LGPU_HIP(hipMemsetAsync(W.tcount_p, 0, 40, stream));
  lgpu_guide G = d->guide;
  GuideRefs gref{};
  for (uint64_t base = 0; base < total; base += cap) {
    const uint32_t n = (uint32_t)std::min<uint64_t>(cap, total - base);
    int cur = 0;
    uint64_t waveRays = 0;
    if (G) {
      const uint64_t expect = GuideExpectedVertices(G) ? GuideExpectedVertices(G) * n / std::max<uint64_t>(1, cap)
                                                         : 2ull * n;
      lgpu_status gs = GuidePrepareWave(G, s->bounds, expect, &gref, W.gbins.as<uint4>(), W.ghdr.as<float2>(), stream);
      if (gs != LGPU_OK) return gs;
    }
    for (int rep = 0; rep < (adj ? 2 : 1); ++rep) {  // adjoint: primal pass, then the replay
    const AdjRefs *ap = rep == 1 ? &adjW : nullptr;
    cur = 0;
    LGPU_HIP(hipEventRecord(ev[0], stream));
    hipLaunchKernelGGL(k_camera, dim3(Blocks(n)), dim3(kBlock), 0, stream, rp, W.ps[0], W.ss, n, base,
                       listPix ? listPix + base : nullptr, listSmp ? listSmp + base : nullptr);
    LGPU_HIP(hipGetLastError());
    LGPU_HIP(hipEventRecord(ev[1], stream));
    LGPU_HIP(hipStreamSynchronize(stream));
    float ms;
    LGPU_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    S.ms_camera += ms;
    const uint32_t maxIter = 4 * (d->max_bounces + 2) + 1100;
    if (rp.integrator == LGPU_INTEGRATOR_PATH && !bound) {
      // Pipelined bounces: each stage reads its item count from the device counter the previous
      // stage wrote, the grids cover the last count the host has seen (counts never grow), and the
      // host stays up to kLag bounces ahead, retiring them through pinned copies and events. No
      // host round trip sits between two kernels.
      constexpr int kLag = 3;
      static_assert(kLag < Wave::kSlots, "a counter slot is reused only after its bounce retired");
      uint32_t *ring = W.ring.as<uint32_t>();
      LGPU_HIP(hipMemsetD32Async((hipDeviceptr_t)(ring + 4 * (Wave::kSlots - 1)), (int)n, 1, stream));  // bounce 0's input count
      uint32_t ub = n, live = n, issued = 0, retired = 0;
      bool done = false;
      while (!(done && retired == issued)) {
        if (!done && issued < maxIter && issued - retired < (uint32_t)kLag) {
          const int sl = (int)(issued % Wave::kSlots), prev = (int)((issued + Wave::kSlots - 1) % Wave::kSlots);
          uint32_t *cin = ring + 4 * prev, *cout = ring + 4 * sl;
          hipEvent_t *e = W.pev[sl];
          LGPU_HIP(hipMemsetAsync(cout, 0, 8, stream));
          // guided: the extension stage lists the bounce's guided vertices for the inference
          const GuideList gl{W.glist.as<uint32_t>(), W.gcount.as<uint32_t>(), G ? gref.max_depth : 0u};
          if (G) LGPU_HIP(hipMemsetAsync(gl.count, 0, 4, stream));
          LGPU_HIP(hipEventRecord(e[0], stream));
          if (countTrav && G)
            hipLaunchKernelGGL((k_extend<true, true>), dim3(Blocks(ub)), dim3(kBlock), 0, stream, dsc, W.ps[cur], ub, cin,
                               W.spill_p, W.tcount_p, gl);
          else if (countTrav)
            hipLaunchKernelGGL(k_extend<true>, dim3(Blocks(ub)), dim3(kBlock), 0, stream, dsc, W.ps[cur], ub, cin,
                               W.spill_p, W.tcount_p, gl);
          else if (G)
            hipLaunchKernelGGL((k_extend<false, true>), dim3(Blocks(ub)), dim3(kBlock), 0, stream, dsc, W.ps[cur], ub, cin,
                               W.spill_p, W.tcount_p, gl);
          else
            hipLaunchKernelGGL(k_extend<false>, dim3(Blocks(ub)), dim3(kBlock), 0, stream, dsc, W.ps[cur], ub, cin,
                               W.spill_p, W.tcount_p, gl);
          LGPU_HIP(hipGetLastError());
          LGPU_HIP(hipEventRecord(e[1], stream));
          if (G) {  // the guide's distribution for the queue's guided vertices (before shading reads it)
            lgpu_status gs = GuideInfer(G, gref, W.ps[cur], ub, gl.list, gl.count, W.gbins.as<uint4>(),
                                        W.ghdr.as<float2>(), stream);
            if (gs != LGPU_OK) return gs;
          }
          LGPU_HIP(hipEventRecord(e[5], stream));
          Queues q{W.ps[cur ^ 1], W.shadow_p, cout};
          LaunchShade(rp.corrected != 0, progs, G ? &gref : nullptr, Blocks(ub), stream, dscp, rp, W.ps[cur], W.ss, ub, cin, q);
          LGPU_HIP(hipGetLastError());
          LGPU_HIP(hipEventRecord(e[2], stream));
          hipLaunchKernelGGL(k_shadow, dim3(Blocks(ub)), dim3(kBlock), 0, stream, dsc, W.ss, W.shadow_p, ub,
                             cout + 1, W.spill_p);
          LGPU_HIP(hipGetLastError());
          LGPU_HIP(hipEventRecord(e[3], stream));
          LGPU_HIP(hipMemcpyAsync(W.hring + 4 * sl, cout, 8, hipMemcpyDeviceToHost, stream));
          LGPU_HIP(hipEventRecord(e[4], stream));
          cur ^= 1;
          ++issued;
          continue;
        }
        if (retired == issued) break;  // bounce limit reached without termination
        const int sl = (int)(retired % Wave::kSlots);
        hipEvent_t *e = W.pev[sl];
        LGPU_HIP(hipEventSynchronize(e[4]));
        ++retired;
        if (done) continue;  // issued after the wave ended: processed no paths
        const uint32_t c0 = W.hring[4 * sl], c1 = W.hring[4 * sl + 1];
        float ms;
        S.extension_rays += live;
        waveRays += live;
        S.launches_extend++;
        S.launches_shade++;
        LGPU_HIP(hipEventElapsedTime(&ms, e[0], e[1]));
        S.ms_extend += ms;
        LGPU_HIP(hipEventElapsedTime(&ms, e[1], e[5]));
        S.ms_guide_infer += ms;
        LGPU_HIP(hipEventElapsedTime(&ms, e[5], e[2]));
        S.ms_shade += ms;
        if (c1) {
          LGPU_HIP(hipEventElapsedTime(&ms, e[2], e[3]));
          S.ms_shadow += ms;
          S.shadow_rays += c1;
          S.launches_shadow++;
        }
        if (c0 > live || c1 > live) return Fail(LGPU_ERR_STATE, "render: queue overflow");
        live = ub = c0;
        if (c0 == 0) done = true;
      }
      if (!done) return Fail(LGPU_ERR_STATE, "render: path wave did not terminate");
      if (G) {  // close the wave's records, then train on them before the next wave infers
        const auto tg = std::chrono::steady_clock::now();
        uint32_t recs = 0;
        lgpu_status gs = GuideFinishWave(G, W.ss, (int)rp.mode, d->guide_train_steps, waveRays, stream, &recs);
        if (gs != LGPU_OK) return gs;
        S.guide_records += recs;
        S.ms_guide_train += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tg).count();
      }
    } else {
      uint32_t count = n;
      uint32_t iter = 0;
      while (count > 0 && iter < maxIter) {
        ++iter;
        LGPU_HIP(hipMemsetAsync(W.counters_p, 0, 8, stream));
        LGPU_HIP(hipEventRecord(ev[2], stream));
        if (countTrav)
          hipLaunchKernelGGL(k_extend<true>, dim3(Blocks(count)), dim3(kBlock), 0, stream, dsc, W.ps[cur], count, nullptr,
                             W.spill_p, W.tcount_p, GuideList{});
        else
          hipLaunchKernelGGL(k_extend<false>, dim3(Blocks(count)), dim3(kBlock), 0, stream, dsc, W.ps[cur], count, nullptr,
                             W.spill_p, W.tcount_p, GuideList{});
        LGPU_HIP(hipGetLastError());
        LGPU_HIP(hipEventRecord(ev[3], stream));
        Queues q{W.ps[cur ^ 1], W.shadow_p, W.counters_p};
        if (rp.integrator == LGPU_INTEGRATOR_PATH)
          LaunchShade(rp.corrected != 0, progs, nullptr, Blocks(count), stream, dscp, rp, W.ps[cur], W.ss, count, nullptr, q);
        else if (rp.integrator == LGPU_INTEGRATOR_DIRECT)
          LaunchShadeDirect(Blocks(count), stream, dsc, rp, W.ps[cur], W.ss, count, q, W.spill_p);
        else if (rp.integrator >= LGPU_INTEGRATOR_DEPTH)
          LaunchShadeUtility(Blocks(count), stream, dsc, rp, W.ps[cur], W.ss, count);
        else
          LaunchShadeVol(Blocks(count), stream, dsc, rp, W.ps[cur], W.ss, count, q, W.spill_p, ap);
        LGPU_HIP(hipGetLastError());
        LGPU_HIP(hipEventRecord(ev[4], stream));
        uint32_t cnt[2];
        LGPU_HIP(hipMemcpyAsync(cnt, W.counters_p, 8, hipMemcpyDeviceToHost, stream));
        LGPU_HIP(hipStreamSynchronize(stream));
        S.extension_rays += count;
        S.launches_extend++;
        S.launches_shade++;
        LGPU_HIP(hipEventElapsedTime(&ms, ev[2], ev[3]));
        S.ms_extend += ms;
        LGPU_HIP(hipEventElapsedTime(&ms, ev[3], ev[4]));
        S.ms_shade += ms;
        if (cnt[0] > count || cnt[1] > count) return Fail(LGPU_ERR_STATE, "render: queue overflow");
        if (cnt[1]) {
          LGPU_HIP(hipEventRecord(ev[5], stream));
          hipLaunchKernelGGL(k_shadow, dim3(Blocks(cnt[1])), dim3(kBlock), 0, stream, dsc, W.ss, W.shadow_p,
                             cnt[1], nullptr, W.spill_p);
          LGPU_HIP(hipGetLastError());
          LGPU_HIP(hipEventRecord(ev[6], stream));
          LGPU_HIP(hipEventSynchronize(ev[6]));
          LGPU_HIP(hipEventElapsedTime(&ms, ev[5], ev[6]));
          S.ms_shadow += ms;
          S.shadow_rays += cnt[1];
          S.launches_shadow++;
        }
        count = cnt[0];
        cur ^= 1;
      }
      if (count > 0) return Fail(LGPU_ERR_STATE, "render: path wave did not terminate");
    }
    if (rep == 1) break;  // the replay's radiance equals the primal pass's: accumulated once
    if (adj) LGPU_HIP(hipMemcpyAsync(W.ltot.as<float4>(), W.ss.L, 16ull * n, hipMemcpyDeviceToDevice, stream));
    const uint32_t npx = w * h;
    if (!listPix && !out && npx > 0)
      hipLaunchKernelGGL(k_accumulate_px, dim3(Blocks(std::min(n, npx))), dim3(kBlock), 0, stream, rp, W.ss, n, npx, film,
                         counts);
    else
      hipLaunchKernelGGL(k_accumulate, dim3(Blocks(n)), dim3(kBlock), 0, stream, rp, W.ss, n, film, counts,
                         out ? out + 5 * base : nullptr);
    LGPU_HIP(hipGetLastError());
    }  // rep
    S.waves++;
  }
  LGPU_HIP(hipStreamSynchronize(stream));
  if (countTrav) {
    unsigned long long tc[5];
    LGPU_HIP(hipMemcpy(tc, W.tcount_p, 40, hipMemcpyDeviceToHost));
    S.nodes_visited = tc[0];
    S.tris_tested = tc[1];
    S.voxel_steps = tc[2];
    S.nodes4_fetched = tc[3];
    S.tris4_fetched = tc[4];
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
  if (d->film->dev != d->scene->dev) return Fail(LGPU_ERR_INVALID_ARGUMENT, "render: film and scene on different devices");
  std::lock_guard<std::mutex> lk(d->scene->dev->mu);
  lgpu_film f = d->film;
  const uint64_t total = (uint64_t)f->width * f->height * (d->spp_end - d->spp_begin);
  return RunWaves(d, f->width, f->height, total, nullptr, nullptr, f->sum.as<double>(), f->count.as<uint32_t>(), nullptr, stats);
}

// Adjoint render: the primal samples go to the film as in lgpu_render, and the gradient of
// sum(d_image * film mean of this call's samples) is added to the parameter gradients.
extern "C" lgpu_status lgpu_render_adjoint(const lgpu_render_desc *d, const lgpu_adjoint_desc *a,
                                           lgpu_render_stats *stats) {
  lgpu_status st = ValidateRender(d);
  if (st != LGPU_OK) return st;
  if (!a || !a->d_image) return Fail(LGPU_ERR_INVALID_ARGUMENT, "render_adjoint: null adjoint descriptor or d_image");
  if (!d->film) return Fail(LGPU_ERR_INVALID_ARGUMENT, "render_adjoint: null film");
  if (d->spp_end <= d->spp_begin) return Fail(LGPU_ERR_INVALID_ARGUMENT, "render_adjoint: empty spp range");
  if (d->film->dev != d->scene->dev) return Fail(LGPU_ERR_INVALID_ARGUMENT, "render_adjoint: film and scene on different devices");
  if (d->integrator != LGPU_INTEGRATOR_VOLPATH || d->estimator_mode != LGPU_ESTIMATOR_CORRECTED ||
      d->spectral_mode != LGPU_SPECTRAL_RGB || d->guide)
    return Fail(LGPU_ERR_UNSUPPORTED,
                "render_adjoint: needs the volumetric path integrator, the corrected estimator, RGB spectra and no guide");
  lgpu_scene s = d->scene;
  if (a->grad_textures && s->adj_texture_shared)
    return Fail(LGPU_ERR_UNSUPPORTED, "render_adjoint: a texture-bound albedo or roughness texture is also read elsewhere");
  std::lock_guard<std::mutex> lk(s->dev->mu);
  lgpu_film f = d->film;
  const uint64_t npx = (uint64_t)f->width * f->height;
  const uint64_t total = npx * (d->spp_end - d->spp_begin);
  size_t nTex = 0, nDens = 0;
  for (const auto &t : s->textures) nTex += t.size();
  for (const auto &g : s->media_density) nDens += g.size();
  const size_t nMat = s->materials.size();
  LGPU_HIP(hipSetDevice(s->dev->device));
  AdjRefs ar{};
  ar.inv_spp = 1.f / (float)(d->spp_end - d->spp_begin);
  const bool onDevice = (a->flags & LGPU_ADJOINT_DEVICE_POINTERS) != 0;
  DevBuf bImg, bTex, bDens, bEmit;
  if (onDevice) {
    ar.dimage = a->d_image;
    ar.gtex = nTex ? a->grad_textures : nullptr;
    ar.gdens = nDens ? a->grad_density : nullptr;
    ar.gemit = nMat ? a->grad_emission : nullptr;
  } else {
    LGPU_HIP(bImg.upload(a->d_image, 12ull * npx));
    ar.dimage = bImg.as<float>();
    auto zeroed = [&](DevBuf &b, size_t count, float *host) -> float * {
      if (!host || !count) return nullptr;
      if (b.alloc(4ull * count) != hipSuccess || hipMemset(b.ptr, 0, 4ull * count) != hipSuccess) return nullptr;
      return b.as<float>();
    };
    ar.gtex = zeroed(bTex, nTex, a->grad_textures);
    ar.gdens = zeroed(bDens, nDens, a->grad_density);
    ar.gemit = zeroed(bEmit, nMat, a->grad_emission);
    if ((a->grad_textures && nTex && !ar.gtex) || (a->grad_density && nDens && !ar.gdens) ||
        (a->grad_emission && nMat && !ar.gemit))
      return Fail(LGPU_ERR_OUT_OF_MEMORY, "render_adjoint: gradient buffers");
    LGPU_HIP(hipDeviceSynchronize());  // the clears land before the render stream reads them
  }
  st = RunWaves(d, f->width, f->height, total, nullptr, nullptr, f->sum.as<double>(), f->count.as<uint32_t>(), nullptr,
                stats, nullptr, &ar);
  if (st != LGPU_OK || onDevice) return st;
  // host pointers: add the device sums into the caller's arrays
  auto addBack = [&](const DevBuf &b, size_t count, float *host) -> hipError_t {
    if (!host || !count) return hipSuccess;
    std::vector<float> t(count);
    const hipError_t e = hipMemcpy(t.data(), b.ptr, 4ull * count, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    for (size_t k = 0; k < count; ++k) host[k] += t[k];
    return hipSuccess;
  };
  LGPU_HIP(addBack(bTex, nTex, a->grad_textures));
  LGPU_HIP(addBack(bDens, nDens, a->grad_density));
  LGPU_HIP(addBack(bEmit, nMat, a->grad_emission));
  return LGPU_OK;
}

namespace lgpu {
// No-allocation render: the path pool lives in this fixed device array, one copy per GPU, held
// until the process ends. Its host-side name is a stand-in, not a device address: host code uses
// only the address hipGetSymbolAddress gives for the current device (kept on the device object),
// for kernel arguments and memsets alike.
constexpr uint64_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

// Called once per device object by lgpu_device_create, with that device current.
lgpu_status ResolveScratch(lgpu_device dev) {
  void *p = nullptr;
  LGPU_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p || (reinterpret_cast<uintptr_t>(p) & 255u)) return Fail(LGPU_ERR_HIP, "device_create: scratch array unaligned");
  dev->scratch = p;
  return LGPU_OK;
}

// Most paths a wave may hold in the scratch array for a scene with `spillWords` spill words.
static uint64_t ScratchPaths(uint32_t spillWords) {
  uint64_t n = kScratchBytes / Wave::BytesPerPath(spillWords);
  n &= ~255ull;
  while (n && Wave::RegionBytes(n, spillWords) > kScratchBytes) n -= 256;
  return n;
}
}  // namespace lgpu

extern "C" lgpu_status lgpu_render_noalloc(const lgpu_render_desc *d, lgpu_render_stats *stats) {
  lgpu_status st = ValidateRender(d);
  if (st != LGPU_OK) return st;
  if (!d->film) return Fail(LGPU_ERR_INVALID_ARGUMENT, "render_noalloc: null film");
  if (d->spp_end <= d->spp_begin) return Fail(LGPU_ERR_INVALID_ARGUMENT, "render_noalloc: empty spp range");
  if (d->film->dev != d->scene->dev) return Fail(LGPU_ERR_INVALID_ARGUMENT, "render_noalloc: film and scene on different devices");
  if (d->guide) return Fail(LGPU_ERR_UNSUPPORTED, "render_noalloc: guiding allocates its tables; use lgpu_render");
  lgpu_device dev = d->scene->dev;
  std::lock_guard<std::mutex> lk(dev->mu);
  if (!dev->scratch) return Fail(LGPU_ERR_STATE, "render_noalloc: device has no scratch array");
  const uint64_t maxPaths = ScratchPaths(d->scene->dsc.spill_words);
  if (maxPaths == 0)
    return Fail(LGPU_ERR_INVALID_ARGUMENT, "render_noalloc: this scene's per-path state does not fit the 64 MiB scratch array");
  if (d->path_capacity > maxPaths)
    return Fail(LGPU_ERR_INVALID_ARGUMENT, "render_noalloc: path_capacity " + std::to_string(d->path_capacity) +
                                               " exceeds the " + std::to_string(maxPaths) +
                                               " paths the 64 MiB scratch array holds for this scene");
  lgpu_film f = d->film;
  const uint64_t total = (uint64_t)f->width * f->height * (d->spp_end - d->spp_begin);
  uint64_t n = d->path_capacity ? d->path_capacity : maxPaths;
  if (n > total) n = total;
  Wave W;
  W.bind(static_cast<uint8_t *>(dev->scratch), (uint32_t)n, d->scene->dsc.spill_words);
  return RunWaves(d, f->width, f->height, total, nullptr, nullptr, f->sum.as<double>(), f->count.as<uint32_t>(), nullptr,
                  stats, &W);
}

extern "C" lgpu_status lgpu_trace_paths(const lgpu_render_desc *d, uint32_t num, const uint32_t *pixels,
                                        const uint32_t *samples, float *out) {
  lgpu_status st = ValidateRender(d);
  if (st != LGPU_OK) return st;
  if (!d->film) return Fail(LGPU_ERR_INVALID_ARGUMENT, "trace_paths: film gives the image size and may not be null");
  if (!num) return LGPU_OK;
  if (!pixels || !samples || !out) return Fail(LGPU_ERR_INVALID_ARGUMENT, "trace_paths: null array");
  lgpu_film f = d->film;
  std::lock_guard<std::mutex> lk(d->scene->dev->mu);
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
  std::lock_guard<std::mutex> lk(s->dev->mu);
  LGPU_HIP(hipSetDevice(s->dev->device));
  (void)hipGetLastError();
  DevBuf dr, dh, dsp;
  LGPU_HIP(dr.upload(rays, 32ull * n));
  LGPU_HIP(dh.alloc(occ ? 4ull * n : 16ull * n));
  LGPU_HIP(dsp.alloc(std::max<size_t>(256, 4ull * s->dsc.spill_words * n)));
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
