// BVH traversal: closest hit and any hit with a per-thread LDS stack. Rendering
// walks the 4-wide quantized BVH; traversal-count passes walk the binary BVH the
// algorithmic byte counts are defined on. Both return the same hit: the closest
// t, ties broken by the smaller primitive index.
#pragma once
#include "dmath.h"
#include "lgpu_internal.h"

namespace lgpu {

constexpr int kBlock = 256;

// ---------------------------------------------------------------- traversal
struct TraceStats {
  uint32_t nodes, tris;
};

// Moller-Trumbore against one BVH-order triangle, same operation order as the CPU checker.
// Returns true when the triangle is hit inside (tmin, tmax); t/u/v/prim are then set.
__device__ inline bool HitTri(const DevScene &sc, uint32_t i, V3 o, V3 d, float tmin, float tmax, float *t, float *u,
                              float *v, uint32_t *prim) {
  const float4 *tp = reinterpret_cast<const float4 *>(sc.tris + i);
  const float4 t0 = tp[0], t1 = tp[1], t2 = tp[2];
  const V3 v0 = v3(t0.x, t0.y, t0.z), e1 = v3(t1.x, t1.y, t1.z), e2 = v3(t2.x, t2.y, t2.z);
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
  *t = tt; *u = uu; *v = vv; *prim = __float_as_uint(t0.w);
  return true;
}

__device__ inline float ByteF(uint32_t w, int k) { return (float)((w >> (8 * k)) & 0xFFu); }

// Box tests use 1/d; a zero (or denormal) direction component would give inf and, through
// inf - inf, NaN slab planes that no longer cull, so an axis-parallel ray would visit the whole
// tree. Such components are replaced by +-1e-20 for the box tests only (triangle tests use d).
__device__ inline float SafeDir(float c) { return fabsf(c) < 1e-20f ? copysignf(1e-20f, c) : c; }

// Closest hit (kAny = false) or any hit (kAny = true) against the 4-wide quantized BVH, one
// inner node or one leaf per step. Children are visited nearest first; the per-thread stack
// lives in LDS column `lane` and entries beyond kLdsStack spill to `spill`
// (DevScene::spill_words per thread). Shadow rays use this loop: a lane stops at its first
// occluder, which parking leaves (Trace4Closest) would delay.
template <bool kAny>
__device__ inline bool Trace4(const DevScene &sc, V3 o, V3 d, float tmin, float tmax, float *tOut, float *uOut,
                              float *vOut, uint32_t *primOut, uint32_t *biOut, uint32_t (*stk)[kBlock], uint32_t *spill) {
  const int lane = threadIdx.x;
  const float ix = 1.f / SafeDir(d.x), iy = 1.f / SafeDir(d.y), iz = 1.f / SafeDir(d.z);
  const float ox = o.x * ix, oy = o.y * iy, oz = o.z * iz;
  const bool nx = ix < 0.f, ny = iy < 0.f, nz = iz < 0.f;  // near plane = hi bound on negative axes
  float best = tmax;
  uint32_t bestPrim = kNone, bestBi = kNone;
  float bu = 0.f, bv = 0.f;
  int sp = 0;
  uint32_t ref = 0;
  for (;;) {
    if (!(ref & kLeafBit)) {
      const float4 *np = reinterpret_cast<const float4 *>(sc.nodes4 + ref);
      const float4 a = np[0], b = np[1];
      const uint4 qc = reinterpret_cast<const uint4 *>(np)[2], ch = reinterpret_cast<const uint4 *>(np)[3];
      // t = (p + q s) / d - o / d = q (s / d) + (p / d - o / d)
      const float Ax = a.w * ix, Bx = a.x * ix - ox;
      const float Ay = b.x * iy, By = a.y * iy - oy;
      const float Az = b.y * iz, Bz = a.z * iz - oz;
      const uint32_t qxl = __float_as_uint(b.z), qxh = __float_as_uint(b.w);
      const uint32_t qxn = nx ? qxh : qxl, qxf = nx ? qxl : qxh;
      const uint32_t qyn = ny ? qc.y : qc.x, qyf = ny ? qc.x : qc.y;
      const uint32_t qzn = nz ? qc.w : qc.z, qzf = nz ? qc.z : qc.w;
      float tn[4];
      uint32_t cr[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float lo = fmaxf(fmaxf(fmaf(ByteF(qxn, k), Ax, Bx), fmaf(ByteF(qyn, k), Ay, By)),
                               fmaxf(fmaf(ByteF(qzn, k), Az, Bz), tmin));
        const float hi = fminf(fminf(fmaf(ByteF(qxf, k), Ax, Bx), fmaf(ByteF(qyf, k), Ay, By)),
                               fminf(fmaf(ByteF(qzf, k), Az, Bz), best));
        const uint32_t r = k == 0 ? ch.x : (k == 1 ? ch.y : (k == 2 ? ch.z : ch.w));
        const bool h = lo <= hi && r != kNone;
        tn[k] = h ? lo : INFINITY;
        cr[k] = h ? r : kNone;
      }
      // sort the four (t, ref) pairs nearest first (5 compare-exchanges)
#define LGPU_CX(i, j)                                                          \
  if (tn[j] < tn[i]) {                                                         \
    const float tt_ = tn[i]; tn[i] = tn[j]; tn[j] = tt_;                        \
    const uint32_t rr_ = cr[i]; cr[i] = cr[j]; cr[j] = rr_;                     \
  }
      LGPU_CX(0, 1) LGPU_CX(2, 3) LGPU_CX(0, 2) LGPU_CX(1, 3) LGPU_CX(1, 2)
#undef LGPU_CX
      if (cr[0] != kNone) {
#pragma unroll
        for (int k = 3; k >= 1; --k) {
          if (cr[k] != kNone) {
            if (sp < kLdsStack) stk[sp][lane] = cr[k];
            else spill[sp - kLdsStack] = cr[k];
            ++sp;
          }
        }
        ref = cr[0];
        continue;
      }
    } else {
      const uint32_t first = ref & 0x03FFFFFFu, count = (ref >> 26) & 31u;
      for (uint32_t i = first; i < first + count; ++i) {
        float tt, uu, vv;
        uint32_t prim;
        if (!HitTri(sc, i, o, d, tmin, tmax, &tt, &uu, &vv, &prim)) continue;
        if (kAny) return true;
        if (tt < best || (tt == best && prim < bestPrim)) {
          best = tt; bestPrim = prim; bestBi = i; bu = uu; bv = vv;
        }
      }
    }
    if (sp == 0) break;
    --sp;
    ref = sp < kLdsStack ? stk[sp][lane] : spill[sp - kLdsStack];
  }
  if (kAny) return false;
  if (bestPrim == kNone) return false;
  *tOut = best; *uOut = bu; *vOut = bv; *primOut = bestPrim; *biOut = bestBi;
  return true;
}

// Closest hit against the 4-wide quantized BVH; same stack and child order as Trace4.
// Loop shape (while-while with one postponed leaf, Aila and Laine 2009): a lane that reaches a
// leaf while its wave is still walking inner nodes parks it and keeps walking; the inner-node
// loop ends when every lane still in it holds a parked leaf, then the wave tests leaves
// together. Inner-node and leaf work thus run in separate wave-wide phases instead of
// alternating under divergent masks every step. The closest hit does not depend on the visit
// order: ties at equal t go to the smaller primitive index.
// kCount: count the 64-B node records and 48-B triangle records this walk fetches (into ts).
template <bool kCount = false>
__device__ inline bool Trace4Closest(const DevScene &sc, V3 o, V3 d, float tmin, float tmax, float *tOut, float *uOut,
                              float *vOut, uint32_t *primOut, uint32_t *biOut, uint32_t (*stk)[kBlock], uint32_t *spill,
                              TraceStats *ts = nullptr) {
  const int lane = threadIdx.x;
  const float ix = 1.f / SafeDir(d.x), iy = 1.f / SafeDir(d.y), iz = 1.f / SafeDir(d.z);
  const float ox = o.x * ix, oy = o.y * iy, oz = o.z * iz;
  const bool nx = ix < 0.f, ny = iy < 0.f, nz = iz < 0.f;  // near plane = hi bound on negative axes
  float best = tmax;
  uint32_t bestPrim = kNone, bestBi = kNone;
  float bu = 0.f, bv = 0.f;
  int sp = 0;
  uint32_t ref = 0;        // next node or leaf to visit; kNone once the stack is empty
  uint32_t parked = kNone; // postponed leaf
#define LGPU_POP() (sp == 0 ? kNone : (--sp, sp < kLdsStack ? stk[sp][lane] : spill[sp - kLdsStack]))
  for (;;) {
    while (ref != kNone && !(ref & kLeafBit)) {
      if (kCount) ts->nodes++;
      const float4 *np = reinterpret_cast<const float4 *>(sc.nodes4 + ref);
      const float4 a = np[0], b = np[1];
      const uint4 qc = reinterpret_cast<const uint4 *>(np)[2], ch = reinterpret_cast<const uint4 *>(np)[3];
      // t = (p + q s) / d - o / d = q (s / d) + (p / d - o / d)
      const float Ax = a.w * ix, Bx = a.x * ix - ox;
      const float Ay = b.x * iy, By = a.y * iy - oy;
      const float Az = b.y * iz, Bz = a.z * iz - oz;
      const uint32_t qxl = __float_as_uint(b.z), qxh = __float_as_uint(b.w);
      const uint32_t qxn = nx ? qxh : qxl, qxf = nx ? qxl : qxh;
      const uint32_t qyn = ny ? qc.y : qc.x, qyf = ny ? qc.x : qc.y;
      const uint32_t qzn = nz ? qc.w : qc.z, qzf = nz ? qc.z : qc.w;
      float tn[4];
      uint32_t cr[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float lo = fmaxf(fmaxf(fmaf(ByteF(qxn, k), Ax, Bx), fmaf(ByteF(qyn, k), Ay, By)),
                               fmaxf(fmaf(ByteF(qzn, k), Az, Bz), tmin));
        const float hi = fminf(fminf(fmaf(ByteF(qxf, k), Ax, Bx), fmaf(ByteF(qyf, k), Ay, By)),
                               fminf(fmaf(ByteF(qzf, k), Az, Bz), best));
        const uint32_t r = k == 0 ? ch.x : (k == 1 ? ch.y : (k == 2 ? ch.z : ch.w));
        const bool h = lo <= hi && r != kNone;
        tn[k] = h ? lo : INFINITY;
        cr[k] = h ? r : kNone;
      }
      // sort the four (t, ref) pairs nearest first (5 compare-exchanges)
#define LGPU_CX(i, j)                                                          \
  if (tn[j] < tn[i]) {                                                         \
    const float tt_ = tn[i]; tn[i] = tn[j]; tn[j] = tt_;                        \
    const uint32_t rr_ = cr[i]; cr[i] = cr[j]; cr[j] = rr_;                     \
  }
      LGPU_CX(0, 1) LGPU_CX(2, 3) LGPU_CX(0, 2) LGPU_CX(1, 3) LGPU_CX(1, 2)
#undef LGPU_CX
      if (cr[0] != kNone) {
#pragma unroll
        for (int k = 3; k >= 1; --k) {
          if (cr[k] != kNone) {
            if (sp < kLdsStack) stk[sp][lane] = cr[k];
            else spill[sp - kLdsStack] = cr[k];
            ++sp;
          }
        }
        ref = cr[0];
      } else {
        ref = LGPU_POP();
      }
      if (ref != kNone && (ref & kLeafBit) && parked == kNone) {  // park the leaf, keep walking
        parked = ref;
        ref = LGPU_POP();
      }
      if (__all(parked != kNone)) break;
    }
    // leaves: the parked one, then `ref` while it is a leaf
    while (parked != kNone || (ref != kNone && (ref & kLeafBit))) {
      uint32_t leaf;
      if (parked != kNone) { leaf = parked; parked = kNone; }
      else { leaf = ref; ref = LGPU_POP(); }
      const uint32_t first = leaf & 0x03FFFFFFu, count = (leaf >> 26) & 31u;
      if (kCount) ts->tris += count;
      for (uint32_t i = first; i < first + count; ++i) {
        float tt, uu, vv;
        uint32_t prim;
        if (!HitTri(sc, i, o, d, tmin, tmax, &tt, &uu, &vv, &prim)) continue;
        if (tt < best || (tt == best && prim < bestPrim)) {
          best = tt; bestPrim = prim; bestBi = i; bu = uu; bv = vv;
        }
      }
    }
    if (ref == kNone) break;
  }
#undef LGPU_POP
  if (bestPrim == kNone) return false;
  *tOut = best; *uOut = bu; *vOut = bv; *primOut = bestPrim; *biOut = bestBi;
  return true;
}

// Closest hit (kAny = false) or any hit (kAny = true) against the binary BVH, counting
// node fetches and triangle tests (kCount). Stack as Trace4 (depth <= kMaxBvhDepth).
template <bool kAny, bool kCount>
__device__ inline bool Trace2(const DevScene &sc, V3 o, V3 d, float tmin, float tmax, float *tOut, float *uOut,
                              float *vOut, uint32_t *primOut, uint32_t *biOut, uint32_t (*stk)[kBlock], uint32_t *spill,
                              TraceStats *ts) {
  const int lane = threadIdx.x;
  const float ix = 1.f / SafeDir(d.x), iy = 1.f / SafeDir(d.y), iz = 1.f / SafeDir(d.z);
  const float ox = o.x * ix, oy = o.y * iy, oz = o.z * iz;
  float best = tmax;
  uint32_t bestPrim = kNone, bestBi = kNone;
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
          best = tt; bestPrim = prim; bestBi = i; bu = uu; bv = vv;
        }
      }
    }
    if (sp == 0) break;
    --sp;
    ref = sp < kLdsStack ? stk[sp][lane] : spill[sp - kLdsStack];
  }
  if (kAny) return false;
  if (bestPrim == kNone) return false;
  *tOut = best; *uOut = bu; *vOut = bv; *primOut = bestPrim; *biOut = bestBi;
  return true;
}

// Rendering traversal (4-wide BVH); kCount = true walks the binary BVH and counts.
template <bool kAny, bool kCount>
__device__ inline bool Trace(const DevScene &sc, V3 o, V3 d, float tmin, float tmax, float *tOut, float *uOut,
                             float *vOut, uint32_t *primOut, uint32_t *biOut, uint32_t (*stk)[kBlock], uint32_t *spill,
                             TraceStats *ts) {
  if (kCount) return Trace2<kAny, true>(sc, o, d, tmin, tmax, tOut, uOut, vOut, primOut, biOut, stk, spill, ts);
  if (kAny) return Trace4<true>(sc, o, d, tmin, tmax, tOut, uOut, vOut, primOut, biOut, stk, spill);
  return Trace4Closest(sc, o, d, tmin, tmax, tOut, uOut, vOut, primOut, biOut, stk, spill);
}

}  // namespace lgpu
