// Shader bytecode interpreter (host and device). A material's node network
// (ShaderGraph.h:72-327) is compiled into Socket records: constants folded
// on the host, event-dependent sockets as straight-line programs over float4
// registers. Socket conversions follow ShaderGraph.h:90-251.
#pragma once
#include "dmath.h"
#include "lgpu_internal.h"

namespace lgpu {

// texture_t::GetPixelUVBilinear (Texture.h:71-81); 1-, 3- or 4-channel FP32. out[3] is the
// alpha channel of an RGBA texture and 1 otherwise.
LGPU_HD void TexBilinear(const float *data, uint32_t w, uint32_t h, uint32_t ch, float u, float v, float out[4]) {
  const float fx = fminf((float)(w - 1), Clampf(u, 0.f, 1.f) * (float)w);
  const float fy = fminf((float)(h - 1), Clampf(v, 0.f, 1.f) * (float)h);
  const uint32_t x = (uint32_t)fx, y = (uint32_t)fy;
  const uint32_t x1 = x + 1 < w - 1 ? x + 1 : w - 1, y1 = y + 1 < h - 1 ? y + 1 : h - 1;
  const float tx = fx - (float)x, ty = fy - (float)y;
  out[3] = 1.f;
  for (uint32_t c = 0; c < (ch == 4 ? 4u : 3u); ++c) {
    const uint32_t cc = ch == 1 ? 0 : c;
    const float p00 = data[(y * w + x) * ch + cc], p10 = data[(y * w + x1) * ch + cc];
    const float p01 = data[(y1 * w + x) * ch + cc], p11 = data[(y1 * w + x1) * ch + cc];
    const float ya = p00 + (p10 - p00) * tx;
    const float yb = p01 + (p11 - p01) * tx;
    out[c] = ya + (yb - ya) * ty;
  }
}

// ---- procedural textures (GraphTexture.cpp:6-240). Lattice coordinates reach the hashes as
// unsigned integers through a 64-bit signed conversion (negative values wrap modulo 2^32).
LGPU_HD uint32_t LatticeU32(float f) { return (uint32_t)(int64_t)f; }
LGPU_HD uint32_t Hash21(uint32_t x, uint32_t y) {
  x *= 1597334677u;
  y *= 3812015801u;
  return (x ^ y) * 1597334677u;
}
LGPU_HD float HashUnit(uint32_t h) { return (float)h * 2.3283064365386963e-10f; }  // h / 2^32, rounded once
LGPU_HD float Quintic(float x) { return x * x * x * (x * (x * 6.f - 15.f) + 10.f); }
LGPU_HD float CubicS(float x) { return x * x * (3.f - 2.f * x); }
LGPU_HD float LerpF(float a, float b, float r) { return a + (b - a) * r; }
LGPU_HD float Perlin2D(float tx, float ty) {
  const float gx = floorf(tx), gy = floorf(ty);
  float gxs[4], gys[4];
  const float cx[4] = {gx, gx + 1, gx + 1, gx}, cy[4] = {gy, gy, gy + 1, gy + 1};
#pragma unroll 1
  for (int k = 0; k < 4; ++k) {
    float u = HashUnit(Hash21(LatticeU32(cx[k]), LatticeU32(cy[k])));
    u *= 6.28318530717958647692f;
    gxs[k] = cosf(u);
    gys[k] = sinf(u);
  }
  const float d00 = (tx - cx[0]) * gxs[0] + (ty - cy[0]) * gys[0];
  const float d10 = (tx - cx[1]) * gxs[1] + (ty - cy[1]) * gys[1];
  const float d11 = (tx - cx[2]) * gxs[2] + (ty - cy[2]) * gys[2];
  const float d01 = (tx - cx[3]) * gxs[3] + (ty - cy[3]) * gys[3];
  const float qx = Quintic(tx - gx);
  const float vx0 = LerpF(d00, d10, qx), vx1 = LerpF(d01, d11, qx);
  return LerpF(vx0, vx1, Quintic(ty - gy)) + .5f;
}
LGPU_HD float Value2D(float tx, float ty) {
  const float x0 = floorf(tx), y0 = floorf(ty);
  const float px = tx - x0, py = ty - y0;
  const float v00 = HashUnit(Hash21(LatticeU32(x0), LatticeU32(y0)));
  const float v10 = HashUnit(Hash21(LatticeU32(x0 + 1), LatticeU32(y0)));
  const float v11 = HashUnit(Hash21(LatticeU32(x0 + 1), LatticeU32(y0 + 1)));
  const float v01 = HashUnit(Hash21(LatticeU32(x0), LatticeU32(y0 + 1)));
  const float cx = CubicS(px);
  return LerpF(LerpF(v00, v10, cx), LerpF(v01, v11, cx), CubicS(py));
}
LGPU_HD float Voronoi2D(float tx, float ty) {
  const float x0 = floorf(tx), y0 = floorf(ty);
  float m = INFINITY;
#pragma unroll 1
  for (int y = -1; y < 2; ++y)
#pragma unroll 1
    for (int x = -1; x < 2; ++x) {
      const float ox = x0 + (float)x, oy = y0 + (float)y;
      const float a = sinf(ox * 127.1f + oy * 311.7f) * 18.5453f, b = sinf(ox * 269.5f + oy * 183.3f) * 18.5453f;
      const float hx = a - floorf(a), hy = b - floorf(b);
      const float rx = (ox + hx) - tx, ry = (oy + hy) - ty;
      const float d2 = rx * rx + ry * ry;
      if (d2 < m) m = d2;
    }
  return sqrtf(m);
}
LGPU_HD float Voronoi3D(float tx, float ty, float tz) {
  const float x0 = floorf(tx), y0 = floorf(ty), z0 = floorf(tz);
  float m = INFINITY;
#pragma unroll 1
  for (int z = -1; z < 2; ++z)
#pragma unroll 1
    for (int y = -1; y < 2; ++y)
#pragma unroll 1
      for (int x = -1; x < 2; ++x) {
        const float ox = x0 + (float)x, oy = y0 + (float)y, oz = z0 + (float)z;
        float px = ox * .1031f, py = oy * .1030f, pz = oz * .0973f;  // Hash33
        px = px - floorf(px); py = py - floorf(py); pz = pz - floorf(pz);
        const float d = (px * (py + 33.33f) + py * (px + 33.33f)) + pz * (pz + 33.33f);
        px += d; py += d; pz += d;
        float hx = (px + py) * pz, hy = (px + px) * py, hz = (py + px) * pz;
        hx = hx - floorf(hx); hy = hy - floorf(hy); hz = hz - floorf(hz);
        const float rx = (ox + hx) - tx, ry = (oy + hy) - ty, rz = (oz + hz) - tz;
        const float d2 = (rx * rx + ry * ry) + rz * rz;
        if (d2 < m) m = d2;
      }
  return sqrtf(m);
}

LGPU_HD float FrDielectric(float cosI, float etaI, float etaT) {  // Fresnel.cpp:5-20
  cosI = Clampf(cosI, -1.f, 1.f);
  if (cosI < 0) { const float t = etaI; etaI = etaT; etaT = t; cosI = fabsf(cosI); }
  const float sinI = sqrtf(fmaxf(0.f, 1.f - cosI * cosI));
  const float sinT = etaI / etaT * sinI;
  if (sinT >= 1) return 1;
  const float cosT = sqrtf(fmaxf(0.f, 1.f - sinT * sinT));
  const float rPara = (etaT * cosI - etaI * cosT) / (etaT * cosI + etaI * cosT);
  const float rPerp = (etaI * cosI - etaT * cosT) / (etaI * cosI + etaT * cosT);
  return (rPara * rPara + rPerp * rPerp) * .5f;
}

struct VmCtx {
  float u, v;       // surface uv
  float woLy;       // local wo.y (Fresnel input node)
  float eta;        // the path's eta (Fresnel input node)
  float px, py, pz; // event point (3-D noise at a medium interaction)
  float nx, ny, nz; // shading normal (surface-information node)
  int medium;       // event is a medium interaction
  float lam[4];     // hero wavelengths
  int mode;         // 0 RGB, 1 HERO4
  const float *texels;
  const uint4 *textures;
  const float4 *consts;
  const Instr *code;
};

LGPU_HD float4 Lane4(float x) { return make_float4(x, x, x, x); }

// Points the context at a hit: uv, point and shading normal (the fields of ScatterEvent::hit a
// socket may read).
LGPU_HD void VmAtHit(VmCtx &vm, float u, float v, V3 p, V3 n) {
  vm.u = u; vm.v = v;
  vm.px = p.x; vm.py = p.y; vm.pz = p.z;
  vm.nx = n.x; vm.ny = n.y; vm.nz = n.z;
}

LGPU_HD float AsRealV(float4 r, uint32_t t, int mode) {
  switch (t) {
    case kScalar: return r.x;
    case kColour: return r.x * .3f + r.y * .59f + r.z * .11f;
    case kVec2: return (r.x + r.y) * .5f;
    case kVec3: return (r.x + r.y + r.z) * (float).333333333333;
    case kSpectrum: {
      Spec s = S4(r);
      return Luminance(s, mode);
    }
  }
  return 0.f;
}
// GetAs<Colour> (ShaderGraph.h:131-168); w is alpha: a colour's own, 1 after a conversion.
LGPU_HD float4 AsColourV(float4 r, uint32_t t) {
  switch (t) {
    case kColour: return r;
    case kVec3: return make_float4(r.x, r.y, r.z, 1.f);
    case kScalar: return make_float4(r.x, r.x, r.x, 1.f);
    case kVec2: { const float m = (r.x + r.y) * .5f; return make_float4(m, m, m, 1.f); }
    case kSpectrum: return make_float4(r.x, r.y, r.z, 1.f);
  }
  return make_float4(1.f, 0.f, 1.f, 1.f);
}
LGPU_HD float4 AsVec3V(float4 r, uint32_t t) {
  switch (t) {
    case kVec3: return make_float4(r.x, r.y, r.z, 0.f);
    case kVec2: return make_float4(r.x, r.y, 0.f, 0.f);
    case kScalar: return make_float4(r.x, r.x, r.x, 0.f);
  }
  return make_float4(0.f, 0.f, 0.f, 0.f);
}
LGPU_HD float4 UpliftV(float r, float g, float b, const VmCtx &c) {
  if (c.mode) return make_float4(Uplift(r, g, b, c.lam[0]), Uplift(r, g, b, c.lam[1]), Uplift(r, g, b, c.lam[2]), Uplift(r, g, b, c.lam[3]));
  return make_float4(r, g, b, 0.f);
}
LGPU_HD float4 AsSpectrumV(float4 r, uint32_t t, const VmCtx &c) {
  switch (t) {
    case kColour: { const float4 col = AsColourV(r, t); return UpliftV(col.x, col.y, col.z, c); }
    case kScalar: return Lane4(r.x);
    case kSpectrum: return r;
  }
  return Lane4(0.f);
}
LGPU_HD float Comp(float4 r, uint32_t i) { return i == 0 ? r.x : (i == 1 ? r.y : (i == 2 ? r.z : r.w)); }

// One procedural-texture lookup of node type `type`: the coordinates are multiplied by s0,
// s1, s2 in turn (the octave node's scale, octave factor, base node scale). A medium
// interaction takes the 3-D variant on the event point (0 for Perlin and value noise).
// Out of line and with rolled loops: the bytecode interpreter's register budget is shared by every
// shading kernel that may call it.
static __host__ __device__ __attribute__((noinline)) float NoiseAt(uint32_t type, const VmCtx &c, float s0, float s1,
                                                                   float s2) {
  if (c.medium) {
    const float tx = c.px * s0 * s1 * s2, ty = c.py * s0 * s1 * s2, tz = c.pz * s0 * s1 * s2;
    if (type == LGPU_NODE_CHECKER) {
      const float x = tx - floorf(tx), y = ty - floorf(ty), z = tz - floorf(tz);
      return ((x > .5f) == (y > .5f)) == (z > .5f) ? 0.f : 1.f;
    }
    if (type == LGPU_NODE_VORONOI) return Voronoi3D(tx, ty, tz);
    return 0.f;
  }
  const float tx = c.u * s0 * s1 * s2, ty = c.v * s0 * s1 * s2;
  switch (type) {
    case LGPU_NODE_CHECKER: {
      const float x = tx - floorf(tx), y = ty - floorf(ty);
      return (x > .5f) == (y > .5f) ? 0.f : 1.f;
    }
    case LGPU_NODE_PERLIN: return Perlin2D(tx, ty);
    case LGPU_NODE_VALUE_NOISE: return Value2D(tx, ty);
  }
  return Voronoi2D(tx, ty);
}

// The ops that depend on the wavelengths or touch textures and tables, out of line: inlined into
// the interpreter loop, their per-wavelength constants were hoisted and kept live across the whole
// loop, which raised the register count of every shading kernel that calls the interpreter.
static __host__ __device__ __attribute__((noinline)) float4 RunHeavyOp(const VmCtx &c, const Instr in, float4 a) {
  float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
  switch (in.op) {
    case kOpTex: case kOpTexChannel: case kOpSpecTex: {  // GraphInputs.cpp:128-165
      const uint4 t = c.textures[in.imm];
      float px[4];
      TexBilinear(c.texels + t.w, t.x, t.y, t.z, Fract(c.u), Fract(c.v), px);
      if (in.op == kOpTex) o = make_float4(px[0], px[1], px[2], px[3]);
      else if (in.op == kOpTexChannel) o.x = px[in.c];
      else o = UpliftV(px[0], px[1], px[2], c);
      break;
    }
    case kOpSpectral: {  // GraphInputs.cpp:55-71
      const float4 k = c.consts[in.imm];
      if (!c.mode) { o = c.consts[in.imm + 1]; break; }
      const uint4 t = c.textures[(uint32_t)k.z];
      const float *d = c.texels + t.w;
      const uint32_t n = t.x * t.y;
      const float step = (k.y - k.x) / (float)(n - 1);
      float lv[4];
      for (int j = 0; j < 4; ++j) {
        const float l = c.lam[j];
        if (l <= k.x) { lv[j] = d[0]; continue; }
        if (l >= k.y) { lv[j] = d[(n - 1) * t.z]; continue; }
        uint32_t q = (uint32_t)((l - k.x) / step);
        if (q > n - 2) q = n - 2;
        const float f = (l - (k.x + step * (float)q)) / step;
        lv[j] = d[q * t.z] + (d[(q + 1) * t.z] - d[q * t.z]) * f;
      }
      o = make_float4(lv[0], lv[1], lv[2], lv[3]);
      break;
    }
    case kOpNoise: {  // GraphTexture.cpp:56-220
      const float4 k = c.consts[in.imm];
      o.x = NoiseAt(in.c, c, k.x, k.y, k.z);
      break;
    }
    case kOpBlackbody: {
      const float T = a.x;
      if (c.mode) {
        o = make_float4(BlackbodyNormalised(c.lam[0], T), BlackbodyNormalised(c.lam[1], T),
                        BlackbodyNormalised(c.lam[2], T), BlackbodyNormalised(c.lam[3], T));
      } else {
        // 3-sample spectrum at 400/500/600 nm through the XYZ->RGB weights in consts[imm..imm+2]
        const float v0 = BlackbodyNormalised(400.f, T), v1 = BlackbodyNormalised(500.f, T), v2 = BlackbodyNormalised(600.f, T);
        const float4 w0 = c.consts[in.imm], w1 = c.consts[in.imm + 1], w2 = c.consts[in.imm + 2];
        o = make_float4(v0 * w0.x + v1 * w0.y + v2 * w0.z, v0 * w1.x + v1 * w1.y + v2 * w1.z,
                        v0 * w2.x + v1 * w2.y + v2 * w2.z, 0.f);
      }
      break;
    }
    case kOpToReal: o.x = AsRealV(a, in.c, c.mode); break;
    case kOpToSpectrum: o = AsSpectrumV(a, in.c, c); break;
  }
  return o;
}

// Straight-line program; returns r[res]. Kept out of line: its register file is an indexed array
// (scratch), which must not burden the shading kernels' common path where every socket is a
// folded constant.
static __host__ __device__ __attribute__((noinline)) float4 RunProgram(const VmCtx &c, uint32_t off, uint32_t len,
                                                                        uint32_t res) {
  float4 r[kMaxRegs];
  for (uint32_t pc = off; pc < off + len; ++pc) {
    const Instr in = c.code[pc];
    float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
    switch (in.op) {
      case kOpConst: o = c.consts[in.imm]; break;
      case kOpUV: o = make_float4(c.u, c.v, 0.f, 0.f); break;
      case kOpTex: case kOpTexChannel: case kOpSpecTex: case kOpSpectral: case kOpNoise: case kOpBlackbody:
      case kOpToReal: case kOpToSpectrum:
        o = RunHeavyOp(c, in, r[in.a & (kMaxRegs - 1)]);
        break;
      case kOpFresnel: {  // GraphInputs.cpp:188-198
        const float n1 = c.eta, n2 = r[in.a].x;
        float R0 = (n1 - n2) / (n1 + n2);
        R0 *= R0;
        const float cc = 1.f - c.woLy;
        o.x = R0 + (1.f - R0) * cc * cc * cc * cc * cc;
        break;
      }
      case kOpNormal: o = make_float4(c.nx, c.ny, c.nz, 0.f); break;
      case kOpCross: {
        const float4 a = r[in.a], b = r[in.b];
        o = make_float4(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, 0.f);
        break;
      }
      case kOpAlpha: o = r[in.a]; o.w = r[in.b].x; break;
      case kOpSMaths: {
        const float a = r[in.a].x, b = r[in.b].x;
        o.x = in.c == 0 ? a + b : (in.c == 1 ? a - b : (in.c == 2 ? a * b : a / b));
        break;
      }
      case kOpVMaths: {
        const float4 a = r[in.a], b = r[in.b];
        if (in.c == 0) o = make_float4(a.x + b.x, a.y + b.y, a.z + b.z, 0.f);
        else if (in.c == 1) o = make_float4(a.x - b.x, a.y - b.y, a.z - b.z, 0.f);
        else if (in.c == 2) o = make_float4(a.x * b.x, a.y * b.y, a.z * b.z, 0.f);
        else o = make_float4(a.x / b.x, a.y / b.y, a.z / b.z, 0.f);
        break;
      }
      case kOpDot: { const float4 a = r[in.a], b = r[in.b]; o.x = a.x * b.x + a.y * b.y + a.z * b.z; break; }
      case kOpLength: { const float4 a = r[in.a]; o.x = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); break; }
      case kOpMerge: o = make_float4(r[in.a].x, r[in.b].x, r[in.c].x, 1.f); break;  // w: alpha (MergeRGBA) or 0 (MergeXYZ)
      case kOpSeparate: o.x = Comp(r[in.a], in.c); break;
      case kOpToColour: o = AsColourV(r[in.a], in.c); break;
      case kOpToVec3: o = AsVec3V(r[in.a], in.c); break;
    }
    r[in.dst] = o;
  }
  return r[res];
}

// Socket readers. kPrograms = false compiles them without the VM: for shading kernels launched
// only on scenes without program sockets (DevScene::has_programs == 0), where the out-of-line
// RunProgram calls would otherwise shape the whole kernel's register allocation.
// Socket read as a spectrum (GetAsSpectrum).
template <bool kPrograms = true>
LGPU_HD Spec SocketSpectrum(const Socket &s, const VmCtx &c) {
  switch (s.kind) {
    case kSockScalar: return S(s.value[0]);
    case kSockColour: return S4(UpliftV(s.value[0], s.value[1], s.value[2], c));
    case kSockBlackbody: {
      if (c.mode) {
        const float T = s.value[3];
        Spec o;
        for (int i = 0; i < 4; ++i) o.c[i] = BlackbodyNormalised(c.lam[i], T);
        return o;
      }
      Spec o;
      o.c[0] = s.value[0]; o.c[1] = s.value[1]; o.c[2] = s.value[2]; o.c[3] = 0.f;
      return o;
    }
    case kSockVec2: return S(0.f);
    default:
      if constexpr (kPrograms) return S4(RunProgram(c, s.prog_off, s.prog_len, s.res));
      else return S(0.f);
  }
}
// Socket read as a real (GetAs<Real>).
template <bool kPrograms = true>
LGPU_HD float SocketReal(const Socket &s, const VmCtx &c) {
  switch (s.kind) {
    case kSockScalar: return s.value[0];
    case kSockVec2: return (s.value[0] + s.value[1]) * .5f;
    case kSockProgram:
      if constexpr (kPrograms) return RunProgram(c, s.prog_off, s.prog_len, s.res).x;
      else return 0.f;
  }
  return s.value[0];
}
// Microfacet roughness: scalar (isotropic) or vec2 (anisotropic) (Microfacet.cpp:57-69).
template <bool kPrograms = true>
LGPU_HD void SocketRoughness(const Socket &s, const VmCtx &c, float *a, float *b) {
  if (s.kind == kSockVec2) { *a = s.value[0]; *b = s.value[1]; return; }
  if constexpr (kPrograms) {
    if (s.kind == kSockProgram) {
      const float4 v = RunProgram(c, s.prog_off, s.prog_len, s.res);
      if (s.res_type == kVec2) { *a = v.x; *b = v.y; }
      else { *a = v.x; *b = v.x; }
      return;
    }
  }
  *a = *b = s.value[0];
}

}  // namespace lgpu
