// BVH2 traversal: closest hit and any hit with a per-thread LDS stack.
#pragma once
#include "dmath.h"
#include "lgpu_internal.h"

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
                             float *vOut, uint32_t *primOut, uint32_t *biOut, uint32_t (*stk)[kBlock], uint32_t *spill,
                             TraceStats *ts) {
  const int lane = threadIdx.x;
  const float ix = 1.f / d.x, iy = 1.f / d.y, iz = 1.f / d.z;
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

}  // namespace lgpu
