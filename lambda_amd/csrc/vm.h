// Shader bytecode interpreter (host and device). A material's node network
// (ShaderGraph.h:72-327) is compiled into Socket records: constants folded
// on the host, event-dependent sockets as straight-line programs over float4
// registers. Socket conversions follow ShaderGraph.h:90-251.
#pragma once
#include "dmath.h"
#include "lgpu_internal.h"

namespace lgpu {

// texture_t::GetPixelUVBilinear (Texture.h:71-81); 1- or 3-channel FP32.
LGPU_HD void TexBilinear(const float *data, uint32_t w, uint32_t h, uint32_t ch, float u, float v, float out[3]) {
  const float fx = fminf((float)(w - 1), Clampf(u, 0.f, 1.f) * (float)w);
  const float fy = fminf((float)(h - 1), Clampf(v, 0.f, 1.f) * (float)h);
  const uint32_t x = (uint32_t)fx, y = (uint32_t)fy;
  const uint32_t x1 = x + 1 < w - 1 ? x + 1 : w - 1, y1 = y + 1 < h - 1 ? y + 1 : h - 1;
  const float tx = fx - (float)x, ty = fy - (float)y;
  for (uint32_t c = 0; c < 3; ++c) {
    const uint32_t cc = ch == 1 ? 0 : c;
    const float p00 = data[(y * w + x) * ch + cc], p10 = data[(y * w + x1) * ch + cc];
    const float p01 = data[(y1 * w + x) * ch + cc], p11 = data[(y1 * w + x1) * ch + cc];
    const float ya = p00 + (p10 - p00) * tx;
    const float yb = p01 + (p11 - p01) * tx;
    out[c] = ya + (yb - ya) * ty;
  }
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
  float lam[4];     // hero wavelengths
  int mode;         // 0 RGB, 1 HERO4
  const float *texels;
  const uint4 *textures;
  const float4 *consts;
  const Instr *code;
};

LGPU_HD float4 Lane4(float x) { return make_float4(x, x, x, x); }

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
LGPU_HD float4 AsColourV(float4 r, uint32_t t) {
  switch (t) {
    case kColour: case kVec3: return make_float4(r.x, r.y, r.z, 0.f);
    case kScalar: return make_float4(r.x, r.x, r.x, 0.f);
    case kVec2: { const float m = (r.x + r.y) * .5f; return make_float4(m, m, m, 0.f); }
    case kSpectrum: return make_float4(r.x, r.y, r.z, 0.f);
  }
  return make_float4(1.f, 0.f, 1.f, 0.f);
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

// Straight-line program; returns r[res]. Kept out of line: its register file
// is an indexed array (scratch), which must not burden the shading kernels'
// common path where every socket is a folded constant.
static __host__ __device__ __attribute__((noinline)) float4 RunProgram(const VmCtx &c, uint32_t off, uint32_t len,
                                                                        uint32_t res) {
  float4 r[kMaxRegs];
  for (uint32_t pc = off; pc < off + len; ++pc) {
    const Instr in = c.code[pc];
    float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
    switch (in.op) {
      case kOpConst: o = c.consts[in.imm]; break;
      case kOpUV: o = make_float4(c.u, c.v, 0.f, 0.f); break;
      case kOpTex: {
        const uint4 t = c.textures[in.imm];
        float px[3];
        TexBilinear(c.texels + t.w, t.x, t.y, t.z, Fract(c.u), Fract(c.v), px);
        o = make_float4(px[0], px[1], px[2], 0.f);
        break;
      }
      case kOpChecker: {  // GraphTexture.cpp:105-108
        const float s = c.consts[in.imm].x;
        const float u = Fract(c.u * s), v = Fract(c.v * s);
        o.x = (u > .5f) == (v > .5f) ? 0.f : 1.f;
        break;
      }
      case kOpFresnel: o.x = FrDielectric(c.woLy, 1.f, c.consts[in.imm].x); break;
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
      case kOpMerge: o = make_float4(r[in.a].x, r[in.b].x, r[in.c].x, 1.f); break;
      case kOpSeparate: o.x = Comp(r[in.a], in.c); break;
      case kOpBlackbody: {
        const float T = r[in.a].x;
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
      case kOpToReal: o.x = AsRealV(r[in.a], in.c, c.mode); break;
      case kOpToColour: o = AsColourV(r[in.a], in.c); break;
      case kOpToVec3: o = AsVec3V(r[in.a], in.c); break;
      case kOpToSpectrum: o = AsSpectrumV(r[in.a], in.c, c); break;
    }
    r[in.dst] = o;
  }
  return r[res];
}

// Socket read as a spectrum (GetAsSpectrum).
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
    default: return S4(RunProgram(c, s.prog_off, s.prog_len, s.res));
  }
}
// Socket read as a real (GetAs<Real>).
LGPU_HD float SocketReal(const Socket &s, const VmCtx &c) {
  switch (s.kind) {
    case kSockScalar: return s.value[0];
    case kSockVec2: return (s.value[0] + s.value[1]) * .5f;
    case kSockProgram: return RunProgram(c, s.prog_off, s.prog_len, s.res).x;
  }
  return s.value[0];
}
// Microfacet roughness: scalar (isotropic) or vec2 (anisotropic) (Microfacet.cpp:57-69).
LGPU_HD void SocketRoughness(const Socket &s, const VmCtx &c, float *a, float *b) {
  if (s.kind == kSockVec2) { *a = s.value[0]; *b = s.value[1]; return; }
  if (s.kind == kSockProgram) {
    const float4 v = RunProgram(c, s.prog_off, s.prog_len, s.res);
    if (s.res_type == kVec2) { *a = v.x; *b = v.y; }
    else { *a = v.x; *b = v.x; }
    return;
  }
  *a = *b = s.value[0];
}

}  // namespace lgpu
