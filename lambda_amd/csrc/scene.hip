// Scene commit: BVH build, material compilation to bytecode, light samplers
// (power CDF and light tree), environment CDFs and the upload to HBM.
// Reference: Scene::Commit (Scene.cpp:18-26), ManyLightSampler::Commit
// (ManyLightSampler.cpp:37-303), PowerLightSampler::Commit (LightSampler.cpp:22-35),
// EnvironmentLight ctor (EnvironmentLight.cpp:7-22).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>

#include "bvh_build.h"
#include "host.h"
#include "vm.h"

namespace lgpu {
namespace {

// ---------------------------------------------------------------- Piecewise1D pool
struct Pool {
  std::vector<float> f;
  // Piecewise.cpp:9-24
  uint32_t add(const float *d, uint32_t n) {
    const uint32_t off = (uint32_t)f.size();
    f.push_back(AsF32(n));
    f.push_back(0.f);
    for (uint32_t i = 0; i < n; ++i) f.push_back(d[i]);
    std::vector<float> cdf(n + 1);
    cdf[0] = 0;
    for (uint32_t i = 1; i < n + 1; ++i) cdf[i] = cdf[i - 1] + d[i - 1] / (float)n;
    const float integral = cdf[n];
    if (integral == 0) {
      for (uint32_t i = 1; i < n + 1; ++i) cdf[i] = (float)i / (float)n;
    } else {
      for (uint32_t i = 1; i < n + 1; ++i) cdf[i] /= integral;
    }
    f[off + 1] = integral;
    for (uint32_t i = 0; i < n + 1; ++i) f.push_back(cdf[i]);
    // Guide table: guide[k] = number of cdf entries <= k / n, k = 0..n (the search in PwFind
    // starts from the bracket [guide[k], guide[k + 1]] of the bin k that holds u). A CDF that is
    // not non-decreasing gets no guide (all ones), so its searches are the plain binary search.
    bool monotone = true;
    for (uint32_t i = 1; i < n + 1; ++i) monotone &= cdf[i] >= cdf[i - 1];
    uint32_t c = 0;
    for (uint32_t k = 0; k < n + 1; ++k) {
      const float t = (float)k / (float)n;
      while (c < n + 1 && cdf[c] <= t) ++c;
      f.push_back(AsF32(monotone ? c : 0xFFFFFFFFu));
    }
    return off;
  }
};

// ---------------------------------------------------------------- light tree (faithful build)
struct Cone { V3 axis = v3(0, 1, 0); float thetaO = 0, thetaE = 0; };
struct BoundsF {
  V3 mn = v3(0, 0, 0), mx = v3(0, 0, 0);
  float SurfaceArea() const { const V3 s = mx - mn; return (s.x * s.y + s.y * s.z + s.x * s.z) * 2.f; }
  V3 Center() const { return mn + (mx - mn) * .5f; }
  float Diagonal() const { return Mag(mx - mn); }
  V3 Offset(V3 p) const { return (p - mn) / (mx - mn); }
};
inline float Comp3(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
inline BoundsF BPoint(V3 p) { BoundsF b; b.mn = p; b.mx = p; return b; }
inline BoundsF BUnion(const BoundsF &a, const BoundsF &b) {  // bounds.h:108-113
  BoundsF r;
  r.mn = v3(std::min(a.mn.x, b.mn.x), std::min(a.mn.y, b.mn.y), std::min(a.mn.z, b.mn.z));
  r.mx = v3(std::max(a.mx.x, b.mx.x), std::max(a.mx.y, b.mx.y), std::max(a.mx.z, b.mx.z));
  return r;
}
inline BoundsF BUnionP(const BoundsF &b, V3 p) {  // bounds.h:116-121
  BoundsF r;
  r.mn = v3(std::min(p.x, b.mn.x), std::min(p.y, b.mn.y), std::min(p.z, b.mn.z));
  r.mx = v3(std::max(p.x, b.mx.x), std::max(p.y, b.mx.y), std::max(p.z, b.mx.z));
  return r;
}
inline V3 RotateR(V3 v, V3 axis, float theta) {  // vec3.h:237-241
  const float c = std::cos(theta);
  const float s = std::sqrt(1.f - c * c);
  return v * c + Cross(axis, v) * s + axis * Dot(axis, v) * (1.f - c);
}
inline Cone MakeCone(V3 axis) { Cone c; c.axis = axis; c.thetaO = 0; c.thetaE = kPi / 2.f; return c; }
// OrientationCone::Union, including its in-place swap of the arguments (ManyLightSampler.cpp:50-64)
inline Cone ConeUnion(Cone &a, Cone &b) {
  if (b.thetaO > a.thetaO) std::swap(a, b);
  const float thetaE = std::max(a.thetaE, b.thetaE);
  const float thetaD = std::acos(Clampf(Dot(a.axis, b.axis), 0.f, 1.f));
  Cone r;
  if (std::min(thetaD + b.thetaO, kPi) <= a.thetaO) { r.axis = a.axis; r.thetaO = a.thetaO; r.thetaE = thetaE; return r; }
  const float thetaO = (a.thetaO + thetaD + b.thetaO) * .5f;
  if (thetaO >= kPi) { r.axis = a.axis; r.thetaO = kPi; r.thetaE = thetaE; return r; }
  const float thetaR = thetaO - a.thetaO;
  r.axis = RotateR(a.axis, Cross(a.axis, b.axis), thetaR);
  r.thetaO = thetaO;
  r.thetaE = thetaE;
  return r;
}
inline float MOmega(const Cone &c) {  // :71-78
  const float cosO = std::cos(c.thetaO);
  const float sinO = std::sqrt(1 - cosO * cosO);
  const float a = kPi2 * (1.f - cosO);
  const float thetaW2 = std::min(c.thetaO + c.thetaE, kPi) * 2.f;
  const float b = thetaW2 * sinO - std::cos(c.thetaO - thetaW2) - 2.f * c.thetaO * sinO + cosO;
  return a + kPi * .5f * b;
}
inline unsigned BucketOf(float x) { return (x >= 0.f) ? (unsigned)x : 0u; }

struct TLight { BoundsF b; V3 dir; float power; uint32_t rec; };
struct BNode {
  int child[2] = {-1, -1};
  int parent = -1;
  unsigned first = 0, count = 0;
  BoundsF bounds;
  Cone cone;
  float power = 0;
  int leaf = -1;
};
struct Bucket { float power = 0; BoundsF b; Cone cone; };

struct TreeBuilder {
  std::vector<TLight> &L;
  std::vector<uint32_t> order;
  std::vector<BNode> nodes;
  std::vector<std::vector<float>> leafPowers;
  std::vector<std::pair<int, unsigned>> lightLeaf;
  float treePower = 0;
  explicit TreeBuilder(std::vector<TLight> &l) : L(l) {}
  const TLight &At(unsigned i) const { return L[order[i]]; }

  float SAOH(const BNode &P, const Bucket &R, const Bucket &Lb, unsigned axis) {  // :83-91
    const V3 box = P.bounds.mx - P.bounds.mn;
    const float lenMax = std::max(std::max(box.x, box.y), box.z);
    const float Kr = lenMax / Comp3(box, (int)axis);
    const float m = P.bounds.SurfaceArea() * MOmega(P.cone);
    const float mL = Lb.power * Lb.b.SurfaceArea() * MOmega(Lb.cone);
    const float mR = R.power * R.b.SurfaceArea() * MOmega(R.cone);
    return Kr * (mL + mR) / m;
  }
  bool SplitAxis(int ni, unsigned axis, float *minCost, unsigned *bIdx, Bucket *lb, Bucket *rb) {  // :136-185
    const unsigned NB = 12;
    Bucket buckets[12];
    bool init[12] = {false};
    const BNode &n = nodes[ni];
    for (unsigned i = n.first; i < n.first + n.count; ++i) {
      unsigned b = BucketOf((float)NB * Comp3(n.bounds.Offset(At(i).b.Center()), (int)axis));
      if (b == NB) b = NB - 1;
      if (b > NB) b = 0;
      if (!init[b]) {
        buckets[b].power = At(i).power;
        buckets[b].b = At(i).b;
        buckets[b].cone = MakeCone(At(i).dir);
        init[b] = true;
      } else {
        buckets[b].power += At(i).power;
        buckets[b].b = BUnion(buckets[b].b, At(i).b);
        Cone tmp = MakeCone(At(i).dir);
        buckets[b].cone = ConeUnion(buckets[b].cone, tmp);
      }
    }
    bool cheaper = false;
    for (unsigned i = 0; i < NB - 1; ++i) {
      Bucket bL, bR;
      bL.power = 0; bL.b = buckets[0].b; bL.cone = buckets[0].cone;
      bR.power = 0; bR.b = buckets[i + 1].b; bR.cone = buckets[i + 1].cone;
      for (unsigned j = 0; j <= i; ++j) {
        bL.b = BUnion(bL.b, buckets[j].b);
        bL.cone = ConeUnion(bL.cone, buckets[j].cone);
        bL.power += buckets[j].power;
      }
      for (unsigned j = i + 1; j < NB; ++j) {
        bR.b = BUnion(bR.b, buckets[j].b);
        bR.cone = ConeUnion(bR.cone, buckets[j].cone);
        bR.power += buckets[j].power;
      }
      const float cost = SAOH(n, bR, bL, axis);
      if (cost < *minCost) { *minCost = cost; *lb = bL; *rb = bR; *bIdx = i; cheaper = true; }
    }
    return cheaper;
  }
  void InitLeaf(int ni) {  // :125-134
    BNode &n = nodes[ni];
    std::vector<float> d(n.count);
    for (unsigned i = 0; i < n.count; ++i) {
      d[i] = At(n.first + i).power;
      lightLeaf[order[n.first + i]] = {ni, i};
      treePower += d[i];
    }
    n.leaf = (int)leafPowers.size();
    leafPowers.push_back(d);
    n.child[0] = n.child[1] = -1;
  }
  void Build(int ni) {  // RecursiveBuild :187-248
    {
      BNode &n = nodes[ni];
      n.bounds = At(n.first).b;
      n.cone = MakeCone(At(n.first).dir);
      for (unsigned i = 0; i < n.count; ++i) {
        n.bounds = BUnion(n.bounds, At(i + n.first).b);
        Cone tmp = MakeCone(At(i + n.first).dir);
        n.cone = ConeUnion(n.cone, tmp);
      }
    }
    Bucket lb, rb;
    float minCost = INFINITY;
    unsigned bIdx = 0;
    int axis = -1;
    for (int a = 0; a < 3; ++a)
      if (SplitAxis(ni, (unsigned)a, &minCost, &bIdx, &lb, &rb)) axis = a;
    const BNode n = nodes[ni];
    if (n.count > 1 && minCost < n.power && axis >= 0) {
      const unsigned NB = 12;
      auto pred = [&](uint32_t li) {
        unsigned b = BucketOf(Comp3(n.bounds.Offset(L[li].b.Center()), axis) * (float)NB);
        if (b == NB) b--;
        if (b > NB) b = 0;
        return b <= bIdx;
      };
      auto it = std::stable_partition(order.begin() + n.first, order.begin() + n.first + n.count, pred);
      const unsigned pivot = (unsigned)(it - order.begin());
      const unsigned ln = pivot - n.first, rn = n.count - ln;
      if (ln != 0 && rn != 0) {
        const int l = (int)nodes.size();
        nodes.push_back(BNode());
        nodes.push_back(BNode());
        const int r = l + 1;
        nodes[l].parent = ni; nodes[l].first = n.first; nodes[l].count = ln;
        nodes[l].bounds = lb.b; nodes[l].cone = lb.cone; nodes[l].power = lb.power;
        nodes[r].parent = ni; nodes[r].first = pivot; nodes[r].count = rn;
        nodes[r].bounds = rb.b; nodes[r].cone = rb.cone; nodes[r].power = rb.power;
        nodes[ni].child[0] = l;
        nodes[ni].child[1] = r;
        Build(l);
        Build(r);
        return;
      }
    }
    InitLeaf(ni);
  }
};

// ---------------------------------------------------------------- material compiler
struct Compiler {
  std::vector<Socket> &sockets;
  std::vector<Lobe> &lobes;
  std::vector<Instr> &code;
  std::vector<float4> &consts;
  const std::vector<lgpu_node> &nodes;
  const std::vector<uint32_t> &texIds;  // texture id -> index in DevScene textures
  const std::vector<std::vector<float>> &texData;  // host texels per texture id
  const std::vector<uint32_t> &texDims;            // w, h, channels per texture id
  std::string err;
  uint32_t nextReg = 0;

  uint32_t OutType(uint32_t t) const {
    switch (t) {
      case LGPU_NODE_SCALAR: case LGPU_NODE_IMAGE_SCALAR: case LGPU_NODE_FRESNEL: case LGPU_NODE_SCALAR_MATHS:
      case LGPU_NODE_DOT: case LGPU_NODE_LENGTH: case LGPU_NODE_SEPARATE_R: case LGPU_NODE_CHECKER:
      case LGPU_NODE_IMAGE_CHANNEL: case LGPU_NODE_SEPARATE_XYZ: case LGPU_NODE_PERLIN: case LGPU_NODE_VALUE_NOISE:
      case LGPU_NODE_OCTAVE_NOISE: case LGPU_NODE_VORONOI: return kScalar;
      case LGPU_NODE_VEC2: case LGPU_NODE_UV: return kVec2;
      case LGPU_NODE_VEC3: case LGPU_NODE_VECTOR_MATHS: case LGPU_NODE_NORMAL: case LGPU_NODE_CROSS:
      case LGPU_NODE_MERGE_XYZ: return kVec3;
      case LGPU_NODE_RGB: case LGPU_NODE_IMAGE_COLOUR: case LGPU_NODE_MERGE_RGB: case LGPU_NODE_SCALAR_TO_COLOUR: return kColour;
      case LGPU_NODE_BLACKBODY: case LGPU_NODE_SPECTRAL: case LGPU_NODE_SPECTRAL_TEXTURE: return kSpectrum;
    }
    return kBxdf;
  }
  // Constant (event-independent, spectral-mode independent) nodes.
  bool IsConst(uint32_t i, int depth = 0) const {
    if (i >= nodes.size() || depth > 64) return false;
    const lgpu_node &n = nodes[i];
    switch (n.type) {
      case LGPU_NODE_SCALAR: case LGPU_NODE_VEC2: case LGPU_NODE_VEC3: case LGPU_NODE_RGB: return true;
      case LGPU_NODE_SCALAR_MATHS: case LGPU_NODE_VECTOR_MATHS: case LGPU_NODE_DOT: case LGPU_NODE_CROSS:
        return IsConst(n.in[0], depth + 1) && IsConst(n.in[1], depth + 1);
      case LGPU_NODE_LENGTH: case LGPU_NODE_SEPARATE_R: case LGPU_NODE_SEPARATE_XYZ: case LGPU_NODE_SCALAR_TO_COLOUR:
        return IsConst(n.in[0], depth + 1);
      case LGPU_NODE_MERGE_XYZ:
        return IsConst(n.in[0], depth + 1) && IsConst(n.in[1], depth + 1) && IsConst(n.in[2], depth + 1);
      case LGPU_NODE_MERGE_RGB:
        return IsConst(n.in[0], depth + 1) && IsConst(n.in[1], depth + 1) && IsConst(n.in[2], depth + 1) &&
               (n.in[3] == LGPU_NONE || IsConst(n.in[3], depth + 1));
    }
    return false;
  }
  // Host evaluation of a constant node: value and its type (mirrors the pull evaluation).
  float4 EvalConst(uint32_t i, uint32_t *type) const {
    const lgpu_node &n = nodes[i];
    *type = OutType(n.type);
    float4 r = make_float4(0, 0, 0, 0);
    uint32_t ta, tb, tc;
    switch (n.type) {
      case LGPU_NODE_SCALAR: r.x = n.p[0]; break;
      case LGPU_NODE_VEC2: r.x = n.p[0]; r.y = n.p[1]; break;
      case LGPU_NODE_VEC3: r = make_float4(n.p[0], n.p[1], n.p[2], 0); break;
      case LGPU_NODE_RGB: r = make_float4(n.p[0], n.p[1], n.p[2], 1); break;
      case LGPU_NODE_SCALAR_MATHS: {
        const float a = AsRealV(EvalConst(n.in[0], &ta), ta, 0), b = AsRealV(EvalConst(n.in[1], &tb), tb, 0);
        const int op = (int)n.p[0];
        r.x = op == 0 ? a + b : (op == 1 ? a - b : (op == 2 ? a * b : a / b));
        break;
      }
      case LGPU_NODE_VECTOR_MATHS: {
        const float4 a = AsVec3V(EvalConst(n.in[0], &ta), ta), b = AsVec3V(EvalConst(n.in[1], &tb), tb);
        const int op = (int)n.p[0];
        if (op == 0) r = make_float4(a.x + b.x, a.y + b.y, a.z + b.z, 0);
        else if (op == 1) r = make_float4(a.x - b.x, a.y - b.y, a.z - b.z, 0);
        else if (op == 2) r = make_float4(a.x * b.x, a.y * b.y, a.z * b.z, 0);
        else r = make_float4(a.x / b.x, a.y / b.y, a.z / b.z, 0);
        break;
      }
      case LGPU_NODE_DOT: {
        const float4 a = AsVec3V(EvalConst(n.in[0], &ta), ta), b = AsVec3V(EvalConst(n.in[1], &tb), tb);
        r.x = a.x * b.x + a.y * b.y + a.z * b.z;
        break;
      }
      case LGPU_NODE_LENGTH: {
        const float4 a = AsVec3V(EvalConst(n.in[0], &ta), ta);
        r.x = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
        break;
      }
      case LGPU_NODE_MERGE_RGB: {
        uint32_t td;
        r = make_float4(AsRealV(EvalConst(n.in[0], &ta), ta, 0), AsRealV(EvalConst(n.in[1], &tb), tb, 0),
                        AsRealV(EvalConst(n.in[2], &tc), tc, 0), 1);
        if (n.in[3] != LGPU_NONE) r.w = AsRealV(EvalConst(n.in[3], &td), td, 0);
        break;
      }
      case LGPU_NODE_MERGE_XYZ:
        r = make_float4(AsRealV(EvalConst(n.in[0], &ta), ta, 0), AsRealV(EvalConst(n.in[1], &tb), tb, 0),
                        AsRealV(EvalConst(n.in[2], &tc), tc, 0), 0);
        break;
      case LGPU_NODE_SEPARATE_R: {
        const float4 c = AsColourV(EvalConst(n.in[0], &ta), ta);
        r.x = Comp(c, (uint32_t)std::min(std::max((int)n.p[0], 0), 3));
        break;
      }
      case LGPU_NODE_SEPARATE_XYZ: {
        const float4 a = AsVec3V(EvalConst(n.in[0], &ta), ta);
        r.x = Comp(a, (uint32_t)std::min(std::max((int)n.p[0], 0), 2));
        break;
      }
      case LGPU_NODE_SCALAR_TO_COLOUR: {
        const float x = AsRealV(EvalConst(n.in[0], &ta), ta, 0);
        r = make_float4(x, x, x, 1);
        break;
      }
      case LGPU_NODE_CROSS: {
        const float4 a = AsVec3V(EvalConst(n.in[0], &ta), ta), b = AsVec3V(EvalConst(n.in[1], &tb), tb);
        r = make_float4(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, 0);
        break;
      }
    }
    return r;
  }
  uint32_t AddConst(float4 v) { consts.push_back(v); return (uint32_t)consts.size() - 1; }
  uint32_t NewReg() {
    if (nextReg >= (uint32_t)kMaxRegs) { err = "shader program needs more than 16 registers"; return 0; }
    return nextReg++;
  }
  void Emit(uint32_t op, uint32_t dst, uint32_t a = 0, uint32_t b = 0, uint32_t c = 0, uint32_t imm = 0) {
    Instr in;
    in.op = op; in.dst = dst; in.a = a; in.b = b; in.c = c; in.imm = imm; in.pad[0] = in.pad[1] = 0;
    code.push_back(in);
  }
  // Emits code computing node i; returns its register, *type its ValType.
  uint32_t EmitNode(uint32_t i, uint32_t *type, int depth = 0) {
    if (i >= nodes.size() || depth > 64) { err = "shader input references a missing node"; *type = kScalar; return 0; }
    const lgpu_node &n = nodes[i];
    *type = OutType(n.type);
    if (IsConst(i)) {
      uint32_t t;
      const float4 v = EvalConst(i, &t);
      const uint32_t r = NewReg();
      Emit(kOpConst, r, 0, 0, 0, AddConst(v));
      return r;
    }
    auto asReal = [&](uint32_t in) {
      uint32_t t;
      const uint32_t r = EmitNode(in, &t, depth + 1);
      if (t == kScalar) return r;
      const uint32_t o = NewReg();
      Emit(kOpToReal, o, r, 0, t);
      return o;
    };
    auto asVec3 = [&](uint32_t in) {
      uint32_t t;
      const uint32_t r = EmitNode(in, &t, depth + 1);
      const uint32_t o = NewReg();
      Emit(kOpToVec3, o, r, 0, t);
      return o;
    };
    const uint32_t dst = NewReg();
    switch (n.type) {
      case LGPU_NODE_BLACKBODY: {
        const uint32_t a = asReal(n.in[0]);
        Emit(kOpBlackbody, dst, a, 0, 0, BlackbodyWeights());
        break;
      }
      case LGPU_NODE_IMAGE_COLOUR:
      case LGPU_NODE_IMAGE_SCALAR: {
        const uint32_t t = (uint32_t)n.p[0];
        if (t >= texIds.size()) { err = "image node names a missing texture"; return dst; }
        Emit(kOpTex, dst, 0, 0, 0, texIds[t]);
        break;
      }
      case LGPU_NODE_UV: Emit(kOpUV, dst); break;
      case LGPU_NODE_NORMAL: Emit(kOpNormal, dst); break;
      case LGPU_NODE_FRESNEL: {
        uint32_t a;
        if (n.in[0] == LGPU_NONE) { a = NewReg(); Emit(kOpConst, a, 0, 0, 0, AddConst(make_float4(n.p[0], 0, 0, 0))); }
        else a = asReal(n.in[0]);
        Emit(kOpFresnel, dst, a);
        break;
      }
      case LGPU_NODE_IMAGE_CHANNEL:
      case LGPU_NODE_SPECTRAL_TEXTURE: {
        const uint32_t t = (uint32_t)n.p[0];
        if (t >= texIds.size()) { err = "image node names a missing texture"; return dst; }
        if (n.type == LGPU_NODE_IMAGE_CHANNEL) Emit(kOpTexChannel, dst, 0, 0, (uint32_t)std::min(std::max((int)n.p[1], 0), 3), texIds[t]);
        else Emit(kOpSpecTex, dst, 0, 0, 0, texIds[t]);
        break;
      }
      case LGPU_NODE_SPECTRAL: {
        const uint32_t t = (uint32_t)n.p[0];
        if (t >= texIds.size()) { err = "spectral node names a missing texture"; return dst; }
        const std::vector<float> &d = texData[t];
        const uint32_t ch = texDims[3 * t + 2], cnt = (uint32_t)(d.size() / ch);
        if (cnt < 2 || !(n.p[2] > n.p[1])) { err = "spectral node needs >= 2 samples over an increasing range"; return dst; }
        std::vector<float> lam(cnt), v(cnt);
        for (uint32_t k = 0; k < cnt; ++k) { lam[k] = n.p[1] + (n.p[2] - n.p[1]) / (float)(cnt - 1) * (float)k; v[k] = d[k * ch]; }
        float rgb[3];
        SampledRGBHost(lam.data(), v.data(), cnt, rgb);
        const uint32_t k0 = AddConst(make_float4(n.p[1], n.p[2], (float)texIds[t], 0));
        AddConst(make_float4(rgb[0], rgb[1], rgb[2], 0));
        Emit(kOpSpectral, dst, 0, 0, 0, k0);
        break;
      }
      case LGPU_NODE_CHECKER: case LGPU_NODE_PERLIN: case LGPU_NODE_VALUE_NOISE: case LGPU_NODE_VORONOI:
        Emit(kOpNoise, dst, 0, 0, n.type, AddConst(make_float4(n.p[0], 1, 1, 0)));
        break;
      case LGPU_NODE_OCTAVE_NOISE: {  // OctaveNoise::Get2D/3D (GraphTexture.cpp:157-175), unrolled
        const uint32_t base = (uint32_t)n.p[2];
        if (base != LGPU_NODE_CHECKER && base != LGPU_NODE_PERLIN && base != LGPU_NODE_VALUE_NOISE && base != LGPU_NODE_VORONOI) {
          err = "octave noise needs a checker, Perlin, value-noise or Voronoi base"; return dst;
        }
        if (!(n.p[1] >= 0.f && n.p[1] <= 32.f)) { err = "octave noise: 0..32 octaves"; return dst; }
        const uint32_t oct = (uint32_t)n.p[1];
        Emit(kOpConst, dst, 0, 0, 0, AddConst(make_float4(0, 0, 0, 0)));
        float amp = .5f, scl = 1.f;
        for (uint32_t k = 0; k < oct; ++k) {
          const uint32_t nr = NewReg(), ar = NewReg(), pr = NewReg();
          Emit(kOpNoise, nr, 0, 0, base, AddConst(make_float4(n.p[0], scl, n.p[3], 0)));
          Emit(kOpConst, ar, 0, 0, 0, AddConst(make_float4(amp, 0, 0, 0)));
          Emit(kOpSMaths, pr, nr, ar, LGPU_OP_MULTIPLY);
          Emit(kOpSMaths, dst, dst, pr, LGPU_OP_ADD);
          nextReg -= 3;
          scl *= 2.f;
          amp *= .5f;
        }
        break;
      }
      case LGPU_NODE_CROSS: { const uint32_t a = asVec3(n.in[0]), b = asVec3(n.in[1]); Emit(kOpCross, dst, a, b); break; }
      case LGPU_NODE_SEPARATE_XYZ: {
        const uint32_t a = asVec3(n.in[0]);
        Emit(kOpSeparate, dst, a, 0, (uint32_t)std::min(std::max((int)n.p[0], 0), 2));
        break;
      }
      case LGPU_NODE_MERGE_XYZ: {
        const uint32_t a = asReal(n.in[0]), b = asReal(n.in[1]), c = asReal(n.in[2]);
        Emit(kOpMerge, dst, a, b, c);
        const uint32_t z = NewReg();
        Emit(kOpConst, z, 0, 0, 0, AddConst(make_float4(0, 0, 0, 0)));
        Emit(kOpAlpha, dst, dst, z);
        break;
      }
      case LGPU_NODE_SCALAR_TO_COLOUR: {
        const uint32_t a = asReal(n.in[0]);
        Emit(kOpToColour, dst, a, 0, kScalar);
        break;
      }
      case LGPU_NODE_SCALAR_MATHS: {
        const uint32_t a = asReal(n.in[0]), b = asReal(n.in[1]);
        Emit(kOpSMaths, dst, a, b, (uint32_t)n.p[0]);
        break;
      }
      case LGPU_NODE_VECTOR_MATHS: {
        const uint32_t a = asVec3(n.in[0]), b = asVec3(n.in[1]);
        Emit(kOpVMaths, dst, a, b, (uint32_t)n.p[0]);
        break;
      }
      case LGPU_NODE_DOT: { const uint32_t a = asVec3(n.in[0]), b = asVec3(n.in[1]); Emit(kOpDot, dst, a, b); break; }
      case LGPU_NODE_LENGTH: { const uint32_t a = asVec3(n.in[0]); Emit(kOpLength, dst, a); break; }
      case LGPU_NODE_MERGE_RGB: {
        const uint32_t a = asReal(n.in[0]), b = asReal(n.in[1]), c = asReal(n.in[2]);
        Emit(kOpMerge, dst, a, b, c);
        if (n.in[3] != LGPU_NONE) { const uint32_t al = asReal(n.in[3]); Emit(kOpAlpha, dst, dst, al); }
        break;
      }
      case LGPU_NODE_SEPARATE_R: {
        uint32_t t;
        const uint32_t r = EmitNode(n.in[0], &t, depth + 1);
        const uint32_t c = NewReg();
        Emit(kOpToColour, c, r, 0, t);
        Emit(kOpSeparate, dst, c, 0, (uint32_t)std::min(std::max((int)n.p[0], 0), 3));
        break;
      }
      default: err = "node type cannot feed a socket"; break;
    }
    // The inputs' registers are dead once this node's value is written: a node keeps one
    // register, so a program needs about (depth x fan-in) registers, not one per node.
    nextReg = dst + 1;
    return dst;
  }
  // XYZ->RGB weights of the 3-sample blackbody spectrum (400/500/600 nm) for RGB mode.
  uint32_t bbw = kNone;
  uint32_t BlackbodyWeights() {
    if (bbw != kNone) return bbw;
    double W[3][3] = {{0}};  // W[xyz][sample]
    for (unsigned i = 0; i < 471; ++i) {
      const float li = 360.f + (float)i;
      float w[3] = {0, 0, 0};
      if (li <= 400.f) w[0] = 1;
      else if (li >= 600.f) w[2] = 1;
      else if (li < 500.f) { const float t = (li - 400.f) / 100.f; w[0] = 1 - t; w[1] = t; }
      else { const float t = (li - 500.f) / 100.f; w[1] = 1 - t; w[2] = t; }
      float cmf[3];
      CMF(li, cmf);
      for (int k = 0; k < 3; ++k)
        for (int s = 0; s < 3; ++s) W[k][s] += (double)w[s] * cmf[k];
    }
    const double scale = 470.0 / (kCieYIntegral * 471.0);
    const double M[3][3] = {{3.240479, -1.537150, -0.498535}, {-0.969256, 1.875991, 0.041556}, {0.055648, -0.204043, 1.057311}};
    float R[3][3];
    for (int r = 0; r < 3; ++r)
      for (int s = 0; s < 3; ++s) {
        double acc = 0;
        for (int k = 0; k < 3; ++k) acc += M[r][k] * W[k][s] * scale;
        R[r][s] = (float)acc;
      }
    bbw = AddConst(make_float4(R[0][0], R[0][1], R[0][2], 0));
    AddConst(make_float4(R[1][0], R[1][1], R[1][2], 0));
    AddConst(make_float4(R[2][0], R[2][1], R[2][2], 0));
    return bbw;
  }

  enum Want { kWantReal, kWantSpectrum, kWantRoughness };
  uint32_t CompileSocket(uint32_t i, Want want) {
    Socket s;
    std::memset(&s, 0, sizeof(s));
    s.kind = kSockScalar;
    if (i == LGPU_NONE) {  // unconnected SocketRef: default value 0
      sockets.push_back(s);
      return (uint32_t)sockets.size() - 1;
    }
    if (i >= nodes.size()) { err = "socket references a missing node"; sockets.push_back(s); return (uint32_t)sockets.size() - 1; }
    const lgpu_node &n = nodes[i];
    if (IsConst(i)) {
      uint32_t t;
      const float4 v = EvalConst(i, &t);
      if (want == kWantReal || (want == kWantRoughness && t != kVec2)) {
        s.kind = kSockScalar;
        s.value[0] = AsRealV(v, t, 0);
      } else if (want == kWantRoughness) {
        s.kind = kSockVec2; s.value[0] = v.x; s.value[1] = v.y;
      } else {  // spectrum (ShaderGraph.h:232-251)
        if (t == kColour) { const float4 c = AsColourV(v, t); s.kind = kSockColour; s.value[0] = c.x; s.value[1] = c.y; s.value[2] = c.z; }
        else if (t == kScalar) { s.kind = kSockScalar; s.value[0] = v.x; }
        else { s.kind = kSockScalar; s.value[0] = 0.f; }
      }
    } else if (want == kWantSpectrum && n.type == LGPU_NODE_BLACKBODY && IsConst(n.in[0])) {
      uint32_t t;
      const float T = AsRealV(EvalConst(n.in[0], &t), t, 0);
      s.kind = kSockBlackbody;
      float rgb[3];
      BlackbodyRGBHost(T, rgb);
      s.value[0] = rgb[0]; s.value[1] = rgb[1]; s.value[2] = rgb[2]; s.value[3] = T;
    } else {
      nextReg = 0;
      s.kind = kSockProgram;
      s.prog_off = (uint32_t)code.size();
      uint32_t t;
      uint32_t r = EmitNode(i, &t);
      if (want == kWantReal && t != kScalar) { const uint32_t o = NewReg(); Emit(kOpToReal, o, r, 0, t); r = o; t = kScalar; }
      if (want == kWantRoughness && t != kVec2 && t != kScalar) { const uint32_t o = NewReg(); Emit(kOpToReal, o, r, 0, t); r = o; t = kScalar; }
      if (want == kWantSpectrum) { const uint32_t o = NewReg(); Emit(kOpToSpectrum, o, r, 0, t); r = o; t = kSpectrum; }
      s.prog_len = (uint32_t)code.size() - s.prog_off;
      s.res = r;
      s.res_type = t;
    }
    sockets.push_back(s);
    return (uint32_t)sockets.size() - 1;
  }

  // BlackbodyInput::MakeBlackbodySpectrum (3 samples) + RGBSpectrum::FromSampled (Spectrum.h:353-373)
  static void BlackbodyRGBHost(float T, float out[3]) {
    const unsigned n = 3;
    const float invSamples = 1.f / (float)n;
    float lam[3], v[3];
    for (unsigned i = 0; i < n; ++i) lam[i] = 400.f + 300.f * invSamples * (float)i;
    for (unsigned i = 0; i < n; ++i) v[i] = BlackbodyNormalised(lam[i], T);
    SampledRGBHost(lam, v, n, out);
  }
  // RGBSpectrum::FromSampled (Spectrum.h:353-373): n samples at increasing wavelengths -> RGB.
  static void SampledRGBHost(const float *lam, const float *v, unsigned n, float out[3]) {
    float xyz[3] = {0, 0, 0};
    for (unsigned i = 0; i < 471; ++i) {
      const float li = 360.f + (float)i;
      float val;
      if (li <= lam[0]) val = v[0];
      else if (li >= lam[n - 1]) val = v[n - 1];
      else {
        const int off = (int)FindInterval((int)n, [&](int k) { return lam[k] <= li; });
        const float t = (li - lam[off]) / (lam[off + 1] - lam[off]);
        val = v[off] + (v[off + 1] - v[off]) * t;
      }
      float cmf[3];
      CMF(li, cmf);
      xyz[0] += val * cmf[0];
      xyz[1] += val * cmf[1];
      xyz[2] += val * cmf[2];
    }
    const float scale = (float)(830 - 360) / (float)(kCieYIntegral * 471);
    for (int k = 0; k < 3; ++k) xyz[k] *= scale;
    out[0] = (float)(3.240479 * xyz[0] - 1.537150 * xyz[1] - 0.498535 * xyz[2]);
    out[1] = (float)(-0.969256 * xyz[0] + 1.875991 * xyz[1] + 0.041556 * xyz[2]);
    out[2] = (float)(0.055648 * xyz[0] - 0.204043 * xyz[1] + 1.057311 * xyz[2]);
  }

  uint32_t CompileLobe(uint32_t i, int depth = 0) {
    if (i >= nodes.size() || depth > 32) { err = "BxDF input references a missing node"; return kNone; }
    const lgpu_node &n = nodes[i];
    Lobe L;
    std::memset(&L, 0, sizeof(L));
    L.kind = n.type;
    L.sock[0] = L.sock[1] = L.sock[2] = kNone;
    L.child[0] = L.child[1] = kNone;
    for (int k = 0; k < 10; ++k) L.p[k] = n.p[k];
    switch (n.type) {
      case LGPU_NODE_LAMBERTIAN: case LGPU_NODE_SPECULAR_BRDF: case LGPU_NODE_SPECULAR_BTDF:
        L.sock[0] = CompileSocket(n.in[0], kWantSpectrum);
        break;
      case LGPU_NODE_OREN_NAYAR_BRDF: case LGPU_NODE_OREN_NAYAR_BTDF: case LGPU_NODE_FRESNEL_BSDF:
      case LGPU_NODE_ROUGH_DIELECTRIC:
        L.sock[0] = CompileSocket(n.in[0], kWantSpectrum);
        L.sock[1] = CompileSocket(n.in[1], kWantReal);
        break;
      case LGPU_NODE_MICROFACET_BRDF:
        L.sock[0] = CompileSocket(n.in[0], kWantSpectrum);
        L.sock[1] = CompileSocket(n.in[1], kWantRoughness);
        break;
      case LGPU_NODE_GHOST_BTDF: L.sock[0] = CompileSocket(n.in[0], kWantReal); break;
      case LGPU_NODE_MIX_BSDF:
        L.child[0] = CompileLobe(n.in[0], depth + 1);
        L.child[1] = CompileLobe(n.in[1], depth + 1);
        L.sock[2] = CompileSocket(n.in[2], kWantReal);
        break;
      default: err = "bxdf_root is not a BxDF node"; return kNone;
    }
    lobes.push_back(L);
    return (uint32_t)lobes.size() - 1;
  }
};

template <class T>
lgpu_status Upload(lgpu_scene s, const std::vector<T> &v, const T **dst) {
  s->bufs.emplace_back();
  DevBuf &b = s->bufs.back();
  if (v.empty()) {
    T zero{};
    LGPU_HIP(b.upload(&zero, sizeof(T)));
  } else {
    LGPU_HIP(b.upload(v.data(), v.size() * sizeof(T)));
  }
  *dst = b.as<const T>();
  return LGPU_OK;
}

}  // namespace

lgpu_status CommitScene(lgpu_scene s, uint32_t strategy) {
  LGPU_HIP(hipSetDevice(s->dev->device));
  // ------------------------------------------------ flatten geometry
  std::vector<float4> pos, nrm, tan;
  std::vector<float2> uv;
  std::vector<uint4> triVtx;
  std::vector<MeshRec> meshes;
  for (size_t m = 0; m < s->meshes.size(); ++m) {
    const HostMesh &hm = s->meshes[m];
    MeshRec r;
    std::memset(&r, 0, sizeof(r));
    r.material = hm.material;
    r.first_tri = (uint32_t)triVtx.size();
    r.num_tris = (uint32_t)hm.idx.size() / 3;
    r.first_vtx = (uint32_t)pos.size();
    r.flags = (!hm.nrm.empty() && !hm.tan.empty() ? kMeshSmooth : 0u) | (!hm.uv.empty() ? kMeshUV : 0u);
    r.mesh_light = kNone;
    const size_t nv = hm.pos.size() / 3;
    for (size_t v = 0; v < nv; ++v) {
      pos.push_back(make_float4(hm.pos[3 * v], hm.pos[3 * v + 1], hm.pos[3 * v + 2], 0.f));
      if (r.flags & kMeshSmooth) {
        nrm.push_back(make_float4(hm.nrm[3 * v], hm.nrm[3 * v + 1], hm.nrm[3 * v + 2], 0.f));
        tan.push_back(make_float4(hm.tan[4 * v], hm.tan[4 * v + 1], hm.tan[4 * v + 2], hm.tan[4 * v + 3]));
      } else {
        nrm.push_back(make_float4(0, 0, 0, 0));
        tan.push_back(make_float4(0, 0, 0, 0));
      }
      uv.push_back((r.flags & kMeshUV) ? make_float2(hm.uv[2 * v], hm.uv[2 * v + 1]) : make_float2(0, 0));
    }
    for (uint32_t t = 0; t < r.num_tris; ++t)
      triVtx.push_back(make_uint4(r.first_vtx + hm.idx[3 * t], r.first_vtx + hm.idx[3 * t + 1], r.first_vtx + hm.idx[3 * t + 2], (uint32_t)m));
    meshes.push_back(r);
  }
  const uint32_t nt = (uint32_t)triVtx.size();
  s->num_tris = nt;
  if (nt >= (1u << 26)) return Fail(LGPU_ERR_UNSUPPORTED, "more than 2^26 triangles");
  auto P = [&](uint32_t vi) { return v3(pos[vi].x, pos[vi].y, pos[vi].z); };

  // ------------------------------------------------ BVH
  BvhBuilder bb;
  bb.tb.resize(nt);
  bb.cen.resize(3 * (size_t)nt);
  bb.idx.resize(nt);
  Box scene;
  for (uint32_t t = 0; t < nt; ++t) {
    const V3 a = P(triVtx[t].x), b = P(triVtx[t].y), c = P(triVtx[t].z);
    Box &bx = bb.tb[t];
    const float pa[3] = {a.x, a.y, a.z}, pb[3] = {b.x, b.y, b.z}, pc[3] = {c.x, c.y, c.z};
    bx.grow(pa); bx.grow(pb); bx.grow(pc);
    scene.grow(bx);
    for (int k = 0; k < 3; ++k) bb.cen[3 * t + k] = (bx.lo[k] + bx.hi[k]) * .5f;
    bb.idx[t] = t;
  }
  float ext = 0;
  for (int k = 0; k < 3 && nt; ++k) ext = std::max(ext, std::max(std::fabs(scene.lo[k]), std::fabs(scene.hi[k])));
  bb.pad = ext * 1e-7f;
  bb.nodes.reserve(nt ? 2 * nt / 2 + 1 : 1);
  bb.nodes.push_back(BvhNode());  // root
  bb.cbox.emplace_back();
  if (nt == 0) {
    Box empty;
    bb.nodes[0].child[0] = MakeLeaf(0, 0);
    bb.nodes[0].child[1] = MakeLeaf(0, 0);
    bb.storeChild(bb.nodes[0], 0, empty);
    bb.storeChild(bb.nodes[0], 1, empty);
  } else {
    const uint32_t r = bb.build(0, nt, 0);
    if (r & kLeafBit) {  // whole scene is one leaf: root holds it next to an empty child
      bb.nodes[0].child[0] = r;
      bb.nodes[0].child[1] = MakeLeaf(0, 0);
      bb.cbox[0][0] = bb.boundsOf(0, nt);
      bb.cbox[0][1] = Box();
      bb.storeChild(bb.nodes[0], 0, bb.boundsOf(0, nt));
      bb.storeChild(bb.nodes[0], 1, Box());
    } else {
      // build() appended the real root at index 1; copy it into slot 0
      bb.nodes[0] = bb.nodes[r];
      bb.cbox[0] = bb.cbox[r];
    }
  }
  s->num_bvh_nodes = (uint32_t)bb.nodes.size();
  s->bvh_depth = bb.maxDepth;
  // The binary traversal stack holds at most one entry per level (LDS + spill = kMaxBvhDepth).
  if (bb.maxDepth + 1 >= (uint32_t)kMaxBvhDepth) return Fail(LGPU_ERR_UNSUPPORTED, "BVH deeper than the traversal stack");
  // 4-wide tree: node 0 is the root (a scene that is one leaf sits in slot 0 of it).
  if (nt == 0) {
    Bvh4Node root;
    std::memset(&root, 0, sizeof(root));
    for (int k = 0; k < 4; ++k) root.child[k] = kNone;
    bb.nodes4.push_back(root);
  } else {
    bb.collapse(0, 0);
  }
  // each 4-wide level pushes at most three entries
  if (3 * (bb.depth4 + 1) + 1 > (uint32_t)kMaxStack) return Fail(LGPU_ERR_UNSUPPORTED, "4-wide BVH deeper than the traversal stack");
  // Spill words a traversal can reach: the larger of the 4-wide bound and the binary bound
  // (traversal-count renders walk the binary tree).
  const uint32_t stackNeed = std::max(3 * (bb.depth4 + 1) + 1, bb.maxDepth + 1);
  const uint32_t spillWords = stackNeed > (uint32_t)kLdsStack ? stackNeed - (uint32_t)kLdsStack : 0u;
  std::vector<BvhTri> btris(nt);
  bool anyAttr = false;
  for (const MeshRec &m : meshes) anyAttr |= (m.flags & (kMeshSmooth | kMeshUV)) != 0;
  std::vector<float4> attr(anyAttr ? (size_t)nt * kTriAttrF4 : 0);
  for (uint32_t i = 0; i < nt; ++i) {
    const uint32_t t = bb.idx[i];
    const uint4 tv = triVtx[t];
    const V3 a = P(tv.x), b = P(tv.y), c = P(tv.z);
    const V3 e1 = b - a, e2 = c - a;
    BvhTri &o = btris[i];
    o.v0[0] = a.x; o.v0[1] = a.y; o.v0[2] = a.z; o.prim = t;
    o.e1[0] = e1.x; o.e1[1] = e1.y; o.e1[2] = e1.z; o.mesh = tv.w;
    o.e2[0] = e2.x; o.e2[1] = e2.y; o.e2[2] = e2.z; o.key = 0;  // set once materials are compiled
    if (anyAttr) {
      float4 *r = &attr[(size_t)i * kTriAttrF4];
      const float2 u0 = uv[tv.x], u1 = uv[tv.y], u2 = uv[tv.z];
      r[0] = make_float4(nrm[tv.x].x, nrm[tv.x].y, nrm[tv.x].z, u0.x);
      r[1] = make_float4(nrm[tv.y].x, nrm[tv.y].y, nrm[tv.y].z, u0.y);
      r[2] = make_float4(nrm[tv.z].x, nrm[tv.z].y, nrm[tv.z].z, u1.x);
      r[3] = tan[tv.x];
      r[4] = tan[tv.y];
      r[5] = tan[tv.z];
      r[6] = make_float4(u1.y, u2.x, u2.y, 0.f);
    }
  }

  // ------------------------------------------------ textures
  std::vector<float> texels;
  std::vector<uint4> texRecs;
  std::vector<uint32_t> texIds;
  for (size_t t = 0; t < s->textures.size(); ++t) {
    texIds.push_back((uint32_t)texRecs.size());
    texRecs.push_back(make_uint4(s->tex_dims[3 * t], s->tex_dims[3 * t + 1], s->tex_dims[3 * t + 2], (uint32_t)texels.size()));
    texels.insert(texels.end(), s->textures[t].begin(), s->textures[t].end());
  }

  // ------------------------------------------------ materials
  std::vector<Socket> sockets;
  std::vector<Lobe> lobes;
  std::vector<Instr> code;
  std::vector<float4> consts;
  std::vector<MaterialRec> mats;
  for (const HostMaterial &hm : s->materials) {
    Compiler c{sockets, lobes, code, consts, hm.nodes, texIds, s->textures, s->tex_dims};
    MaterialRec r;
    std::memset(&r, 0, sizeof(r));
    r.lobe = hm.desc.bxdf_root == LGPU_NONE ? kNone : c.CompileLobe(hm.desc.bxdf_root);
    r.emission = hm.desc.emission_root == LGPU_NONE ? kNone : c.CompileSocket(hm.desc.emission_root, Compiler::kWantSpectrum);
    r.intensity = hm.desc.emission_intensity;
    r.med_in = hm.desc.medium_interior;
    r.med_out = hm.desc.medium_exterior;
    if (!c.err.empty()) return Fail(LGPU_ERR_GRAPH, "material " + std::to_string(mats.size()) + ": " + c.err);
    for (uint32_t id : {r.med_in, r.med_out})
      if (id != kNone && id >= s->media.size())
        return Fail(LGPU_ERR_INVALID_ARGUMENT, "material " + std::to_string(mats.size()) + ": unknown medium id");
    mats.push_back(r);
  }
  // Guidable: no delta lobe and a pdf that matches the sampler (Lambertian, Oren-Nayar BRDF,
  // microfacet BRDF, rough dielectric, mixes of those): the guided shade stage may then sample
  // the lobe's directions from the guide as well (MaterialRec::pad[0] bit 0).
  {
    std::function<bool(uint32_t)> guidable = [&](uint32_t li) -> bool {
      const Lobe &L = lobes[li];
      switch (L.kind) {
        case kLambertian: case kOrenNayarR: case kMicrofacet: case kRoughDielectric: return true;
        case kMix: return guidable(L.child[0]) && guidable(L.child[1]);
        default: return false;
      }
    };
    for (MaterialRec &r : mats)
      if (r.lobe != kNone && guidable(r.lobe)) r.pad[0] |= 1u;
  }
  // Adjoint texture bindings (AdjLobe): the albedo socket of Lambertian and microfacet lobes and
  // the roughness socket of isotropic microfacet lobes, when the socket is one bilinear lookup of
  // an image texture through at most one conversion. A texture that is also read anywhere else is
  // flagged: the adjoint render refuses texture gradients for such scenes (that other read would
  // miss its gradient).
  std::vector<AdjLobe> adjLobes(lobes.size());
  {
    std::vector<uint32_t> refs(texRecs.size(), 0), bound(texRecs.size(), 0);
    for (const Instr &in : code)
      if ((in.op == kOpTex || in.op == kOpTexChannel || in.op == kOpSpecTex) && in.imm < refs.size()) refs[in.imm]++;
    // weights of the lookup's components for a socket, or false if it is not a bound lookup
    auto decode = [&](uint32_t si, bool spectrum, uint32_t *tex, float w[3][4]) -> bool {
      if (si == kNone) return false;
      const Socket &sk = sockets[si];
      if (sk.kind != kSockProgram || sk.prog_len == 0 || sk.prog_len > 2) return false;
      const Instr &a = code[sk.prog_off];
      if ((a.op != kOpTex && a.op != kOpTexChannel) || a.imm >= texRecs.size()) return false;
      float comp[3][4] = {};  // value channel (3 for a spectrum, 1 otherwise) from lookup component
      uint32_t t = kScalar;   // type of the lookup's register: a colour (kOpTex) or one channel
      float base[4][4] = {};  // register lanes x..w as weights of the components
      if (a.op == kOpTex) {
        for (int k = 0; k < 4; ++k) base[k][k] = 1.f;
        t = kColour;
      } else {
        base[0][std::min<uint32_t>(a.c, 3)] = 1.f;
        t = kScalar;
      }
      uint32_t reg = a.dst;
      if (sk.prog_len == 2) {
        const Instr &b = code[sk.prog_off + 1];
        if (b.a != reg) return false;
        if (b.op == kOpToSpectrum && spectrum) {
          if (b.c == kColour) {  // RGB mode: the colour's lanes x, y, z
            for (int c = 0; c < 3; ++c) for (int k = 0; k < 4; ++k) comp[c][k] = base[c][k];
          } else if (b.c == kScalar) {
            for (int c = 0; c < 3; ++c) for (int k = 0; k < 4; ++k) comp[c][k] = base[0][k];
          } else {
            return false;
          }
        } else if (b.op == kOpToReal && !spectrum) {
          if (b.c == kColour) {
            const float cw[3] = {.3f, .59f, .11f};
            for (int k = 0; k < 4; ++k) comp[0][k] = cw[0] * base[0][k] + cw[1] * base[1][k] + cw[2] * base[2][k];
          } else if (b.c == kScalar) {
            for (int k = 0; k < 4; ++k) comp[0][k] = base[0][k];
          } else {
            return false;
          }
        } else {
          return false;
        }
        reg = b.dst;
      } else {
        if (spectrum || !(sk.res_type == kScalar)) return false;
        for (int k = 0; k < 4; ++k) comp[0][k] = base[0][k];
      }
      if (reg != sk.res) return false;
      (void)t;
      *tex = a.imm;
      for (int c = 0; c < 3; ++c) for (int k = 0; k < 4; ++k) w[c][k] = comp[c][k];
      return true;
    };
    for (size_t li = 0; li < lobes.size(); ++li) {
      AdjLobe &A = adjLobes[li];
      std::memset(&A, 0, sizeof(A));
      A.albedo_tex = A.rough_tex = kNone;
      const Lobe &L = lobes[li];
      if (L.kind != kLambertian && L.kind != kMicrofacet) continue;
      uint32_t t;
      float w[3][4];
      if (decode(L.sock[0], true, &t, w)) {
        A.albedo_tex = t;
        std::memcpy(A.albedo_w, w, sizeof(w));
        bound[t]++;
      }
      if (L.kind == kMicrofacet && decode(L.sock[1], false, &t, w)) {
        A.rough_tex = t;
        std::memcpy(A.rough_w, w[0], sizeof(A.rough_w));
        bound[t]++;
      }
    }
    s->adj_texture_shared = false;
    for (size_t t = 0; t < refs.size(); ++t)
      if (bound[t] > 0 && refs[t] > bound[t]) s->adj_texture_shared = true;
    // textures an emission socket reads feed the light sampler's power estimates: updating them
    // in place would leave those stale
    s->tex_emissive.assign(s->textures.size(), 0);
    for (const MaterialRec &r : mats) {
      if (r.emission == kNone || sockets[r.emission].kind != kSockProgram) continue;
      const Socket &sk = sockets[r.emission];
      for (uint32_t pc = sk.prog_off; pc < sk.prog_off + sk.prog_len; ++pc) {
        const Instr &in = code[pc];
        if ((in.op == kOpTex || in.op == kOpTexChannel || in.op == kOpSpecTex) && in.imm < s->tex_emissive.size())
          s->tex_emissive[in.imm] = 1;
      }
    }
  }
  for (BvhTri &t : btris) {
    const uint32_t root = mats[meshes[t.mesh].material].lobe;
    t.key = root == kNone ? 1u : 2u + (lobes[root].kind - (uint32_t)kLambertian);
    if (t.key >= kShadeKeys) t.key = kShadeKeys - 1;
    if (mats[meshes[t.mesh].material].pad[0] & 1u) t.key |= kKeyGuidable;
  }
  // scene bounds (guiding inputs are positions normalised to them)
  {
    float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (const BvhTri &t : btris) {
      const float px[3][3] = {{t.v0[0], t.v0[1], t.v0[2]},
                              {t.v0[0] + t.e1[0], t.v0[1] + t.e1[1], t.v0[2] + t.e1[2]},
                              {t.v0[0] + t.e2[0], t.v0[1] + t.e2[1], t.v0[2] + t.e2[2]}};
      for (const auto &q : px)
        for (int a = 0; a < 3; ++a) { b[a] = std::min(b[a], q[a]); b[3 + a] = std::max(b[3 + a], q[a]); }
    }
    if (btris.empty()) { for (int a = 0; a < 3; ++a) { b[a] = -1.f; b[3 + a] = 1.f; } }
    for (int a = 0; a < 6; ++a) s->bounds[a] = b[a];
  }

  // ------------------------------------------------ media
  std::vector<MediumRec> media;
  std::vector<float> density;
  for (size_t k = 0; k < s->media.size(); ++k) {
    const lgpu_medium_desc &md = s->media[k];
    MediumRec r;
    std::memset(&r, 0, sizeof(r));
    r.type = md.type;
    for (int a = 0; a < 3; ++a) {
      r.sa[a] = md.sigma_a[a]; r.ss[a] = md.sigma_s[a];
      r.bmin[a] = md.bounds_min[a]; r.bmax[a] = md.bounds_max[a];
    }
    r.g = md.g;
    if (md.type == LGPU_MEDIUM_GRID) {
      r.nx = md.nx; r.ny = md.ny; r.nz = md.nz;
      r.dens_off = (uint32_t)density.size();
      density.insert(density.end(), s->media_density[k].begin(), s->media_density[k].end());
    }
    media.push_back(r);
  }
  if (density.empty()) density.push_back(0.f);

  // ------------------------------------------------ lights
  Pool pool;
  std::vector<LightRec> lights;
  EnvRec env;
  std::memset(&env, 0, sizeof(env));
  uint32_t hasEnv = 0, envLight = kNone;
  Box sb;
  for (uint32_t t = 0; t < nt; ++t) sb.grow(bb.tb[t]);
  const float sceneMaxLen = nt ? std::max(std::max(sb.hi[0] - sb.lo[0], sb.hi[1] - sb.lo[1]), sb.hi[2] - sb.lo[2]) : 0.f;
  VmCtx hostVm;
  std::memset(&hostVm, 0, sizeof(hostVm));
  hostVm.eta = 1.001f;  // a default ScatterEvent
  hostVm.texels = texels.data();
  hostVm.textures = texRecs.data();
  hostVm.consts = consts.data();
  hostVm.code = code.data();
  auto LightPower = [&](const LightRec &L) -> float {
    switch (L.kind) {
      case kLightMesh: {
        const MeshRec &m = meshes[L.mesh];
        float a = 0;
        for (uint32_t i = 0; i < m.num_tris; ++i) {
          const uint4 tv = triVtx[m.first_tri + i];
          a += Mag(Cross(P(tv.y) - P(tv.x), P(tv.z) - P(tv.x)));
        }
        return L.intensity * (a * .5f);
      }
      case kLightPoint: return L.intensity * (4 * kPi);
      case kLightSpot: return L.intensity * (kPi2 - (1 - .5f * (L.cos_falloff + L.cos_cone)));
      case kLightEnv: return L.intensity * (kPi * env.radius * env.radius);
    }
    return 0.f;
  };
  std::vector<float> texelsEnv;
  for (size_t li = 0; li < s->lights.size(); ++li) {
    const HostLight &hl = s->lights[li];
    LightRec L;
    std::memset(&L, 0, sizeof(L));
    L.intensity = hl.desc.intensity;
    L.temperature = hl.desc.temperature;
    L.mesh = kNone; L.parent = kNone;
    if (hl.mesh >= 0) {  // MeshLight (MeshLight.cpp:9-13,118-125)
      L.kind = kLightMesh;
      L.mesh = (uint32_t)hl.mesh;
      MeshRec &m = meshes[hl.mesh];
      m.mesh_light = (uint32_t)li;
      L.intensity = s->materials[m.material].desc.emission_intensity;
      std::vector<float> areas(m.num_tris);
      for (uint32_t i = 0; i < m.num_tris; ++i) {
        const uint4 tv = triVtx[m.first_tri + i];
        areas[i] = Mag(Cross(P(tv.y) - P(tv.x), P(tv.z) - P(tv.x))) * .5f;
      }
      L.cdf_off = pool.add(areas.data(), m.num_tris);
      L.count = m.num_tris;
    } else if (hl.desc.type == LGPU_LIGHT_POINT || hl.desc.type == LGPU_LIGHT_SPOT) {
      L.kind = hl.desc.type == LGPU_LIGHT_POINT ? kLightPoint : kLightSpot;
      for (int k = 0; k < 3; ++k) { L.pos[k] = hl.desc.position[k]; L.axis[k] = hl.desc.axis[k]; }
      L.cos_cone = std::cos(hl.desc.cone_angle);
      L.cos_falloff = std::cos(hl.desc.falloff_start);
      if (hl.desc.temperature > 0) Compiler::BlackbodyRGBHost(hl.desc.temperature, L.rgb);
      else for (int k = 0; k < 3; ++k) L.rgb[k] = hl.desc.rgb[k];
    } else {  // EnvironmentLight (EnvironmentLight.cpp:7-22)
      L.kind = kLightEnv;
      const uint32_t w = hl.desc.env_width, h = hl.desc.env_height;
      env.w = w; env.h = h;
      env.tex_off = (uint32_t)texels.size();
      env.offset[0] = hl.desc.env_offset[0]; env.offset[1] = hl.desc.env_offset[1];
      env.intensity = hl.desc.intensity;
      env.radius = sceneMaxLen;  // Scene::Commit: radius = bounds.MaxLength()
      env.light = (uint32_t)li;
      std::vector<float> img((size_t)w * h);
      for (uint32_t y = 0; y < h; ++y) {
        const float vp = (float)y / (float)h;
        const float sinT = std::sin(kPi * (float)(y + .5) / (float)h);
        for (uint32_t x = 0; x < w; ++x) {
          const float up = (float)x / (float)w;
          float c[4];  // TexBilinear also writes out[3] (alpha)
          TexBilinear(hl.env.data(), w, h, 3, Fract(up), Fract(vp), c);
          img[x + y * w] = std::fabs(0.212671f * c[0] + 0.715160f * c[1] + 0.072169f * c[2]);
          img[x + y * w] *= sinT;
        }
      }
      // Piecewise2D (Piecewise.cpp:50-59): conditionals per row, marginal over row integrals
      env.cond_off = (uint32_t)pool.f.size();
      std::vector<float> marg(h);
      for (uint32_t v = 0; v < h; ++v) {
        const uint32_t off = pool.add(&img[(size_t)v * w], w);
        marg[v] = pool.f[off + 1];
        if (v == 1) env.cond_stride = off - env.cond_off;
      }
      if (h == 1) env.cond_stride = 0;
      env.marg_off = pool.add(marg.data(), h);
      texelsEnv = hl.env;
      hasEnv = 1;
      envLight = (uint32_t)li;
    }
    lights.push_back(L);
  }
  if (hasEnv) {
    env.tex_off = (uint32_t)texels.size();
    texels.insert(texels.end(), texelsEnv.begin(), texelsEnv.end());
  }
  for (LightRec &L : lights) L.power = LightPower(L);

  DevScene &d = s->dsc;
  std::memset(&d, 0, sizeof(d));
  d.strategy = strategy;
  d.num_lights = (uint32_t)lights.size();
  d.num_media = (uint32_t)media.size();
  d.spill_words = spillWords;
  d.has_programs = 0;
  for (const Socket &k : sockets)
    if (k.kind != kSockScalar && k.kind != kSockColour && k.kind != kSockBlackbody && k.kind != kSockVec2) d.has_programs = 1;
  d.has_env = hasEnv;
  d.has_mesh_lights = 0;  // as MisLight reads it: some mesh record names a light
  for (const MeshRec &m : meshes)
    if (m.mesh_light != kNone) d.has_mesh_lights = 1;
  d.env_light = envLight;
  d.env = env;
  d.num_tris = nt;
  std::vector<LightRec> tl;
  std::vector<uint32_t> triTL(nt, kNone), tlLeaf, tlSlot, treeOrder;
  std::vector<TreeNode> tree;
  if (strategy == LGPU_LIGHTS_POWER) {
    if (!lights.empty()) {  // PowerLightSampler::Commit
      std::vector<float> imp(lights.size());
      float total = 0;
      for (size_t i = 0; i < lights.size(); ++i) { imp[i] = lights[i].power; total += imp[i]; }
      d.power_dist = pool.add(imp.data(), (uint32_t)imp.size());
      d.inv_total_power = 1.f / total;
    }
  } else {
    // ManyLightSampler::InitLights: triangle lights (scene order, triangle order), then point/spot lights
    std::vector<TLight> T;
    for (size_t li = 0; li < lights.size(); ++li) {
      const LightRec &L = lights[li];
      if (L.kind != kLightMesh) continue;
      const MeshRec &m = meshes[L.mesh];
      const Socket *em = mats[m.material].emission == kNone ? nullptr : &sockets[mats[m.material].emission];
      for (uint32_t t = 0; t < m.num_tris; ++t) {
        LightRec R = L;
        R.kind = kLightTri;
        R.tri = t;
        R.parent = (uint32_t)li;
        const uint4 tv = triVtx[m.first_tri + t];
        const V3 p0 = P(tv.x), p1 = P(tv.y), p2 = P(tv.z);
        // TriangleLight::Irradiance (MeshLight.cpp:203-221): 3x3 barycentric samples of emission.y()
        const float inv = 1.f / 3.f;
        float irr = 0;
        for (unsigned a = 0; a < 3; ++a)
          for (unsigned b = 0; b < 3; ++b) {
            if (m.flags & kMeshUV) {
              const float2 u0 = uv[tv.x], u1 = uv[tv.y], u2 = uv[tv.z];
              const float ua = (float)a * inv, ub = (float)b * inv;
              hostVm.u = u0.x + (u1.x - u0.x) * ua + (u2.x - u0.x) * ub;
              hostVm.v = u0.y + (u1.y - u0.y) * ua + (u2.y - u0.y) * ub;
            } else { hostVm.u = 0; hostVm.v = 0; }
            hostVm.mode = 0;
            const Spec e = em ? SocketSpectrum(*em, hostVm) : S(0.f);
            irr += Luminance(e, 0);
          }
        const float area = Mag(Cross(p1 - p0, p2 - p0)) * .5f;
        TLight x;
        x.b = BUnionP(BUnionP(BPoint(p0), p1), p2);
        const V3 cr = Cross(p1 - p0, p2 - p0);
        const float ar = Mag(cr) * .5f;
        x.dir = cr / (ar * 2.f);
        x.power = (irr * inv * inv * L.intensity) * area;
        R.power = x.power;
        x.rec = (uint32_t)tl.size();
        triTL[m.first_tri + t] = (uint32_t)tl.size();
        tl.push_back(R);
        T.push_back(x);
      }
    }
    for (size_t li = 0; li < lights.size(); ++li) {
      const LightRec &L = lights[li];
      if (L.kind != kLightPoint && L.kind != kLightSpot) continue;
      TLight x;
      x.b = BPoint(v3(L.pos[0], L.pos[1], L.pos[2]));
      x.dir = L.kind == kLightSpot ? v3(L.axis[0], L.axis[1], L.axis[2]) : v3(0, 1, 0);
      x.power = L.power;
      x.rec = (uint32_t)tl.size();
      tl.push_back(L);
      T.push_back(x);
    }
    TreeBuilder tbld(T);
    tbld.order.resize(T.size());
    for (size_t i = 0; i < T.size(); ++i) tbld.order[i] = (uint32_t)i;
    tbld.lightLeaf.assign(T.size(), {-1, 0});
    d.inf_power = hasEnv ? lights[envLight].power * .1f : 0.f;
    if (!T.empty()) {
      tbld.nodes.push_back(BNode());
      BNode &r = tbld.nodes[0];
      r.count = (unsigned)T.size();
      r.first = 0;
      r.bounds = T[0].b;
      r.power = 0;
      r.cone = MakeCone(T[0].dir);
      for (unsigned i = 0; i < r.count; ++i) {
        r.bounds = BUnion(r.bounds, T[i].b);
        Cone tmp = MakeCone(T[i].dir);
        r.cone = ConeUnion(r.cone, tmp);
        r.power += T[i].power;
      }
      tbld.Build(0);
    }
    d.tree_power = tbld.treePower;
    // Flatten depth-first, left child first.
    std::vector<int> pre, newIdx(tbld.nodes.size(), -1);
    std::vector<int> st;
    if (!tbld.nodes.empty()) st.push_back(0);
    while (!st.empty()) {
      const int n = st.back();
      st.pop_back();
      newIdx[n] = (int)pre.size();
      pre.push_back(n);
      if (tbld.nodes[n].child[0] >= 0) { st.push_back(tbld.nodes[n].child[1]); st.push_back(tbld.nodes[n].child[0]); }
    }
    std::vector<uint32_t> leafOff(tbld.leafPowers.size());
    for (size_t l = 0; l < tbld.leafPowers.size(); ++l) leafOff[l] = pool.add(tbld.leafPowers[l].data(), (uint32_t)tbld.leafPowers[l].size());
    for (int n : pre) {
      const BNode &b = tbld.nodes[n];
      TreeNode t;
      std::memset(&t, 0, sizeof(t));
      const V3 c = b.bounds.Center();
      t.center[0] = c.x; t.center[1] = c.y; t.center[2] = c.z;
      t.diag = b.bounds.Diagonal();
      t.axis[0] = b.cone.axis.x; t.axis[1] = b.cone.axis.y; t.axis[2] = b.cone.axis.z;
      t.thetaO = b.cone.thetaO; t.thetaE = b.cone.thetaE;
      t.power = b.power;
      t.child[0] = b.child[0] >= 0 ? (uint32_t)newIdx[b.child[0]] : kNone;
      t.child[1] = b.child[1] >= 0 ? (uint32_t)newIdx[b.child[1]] : kNone;
      t.first = b.first; t.count = b.count;
      t.parent = b.parent >= 0 ? (uint32_t)newIdx[b.parent] : kNone;
      t.leaf_dist = b.leaf >= 0 ? leafOff[b.leaf] : kNone;
      tree.push_back(t);
    }
    treeOrder = tbld.order;
    tlLeaf.resize(T.size());
    tlSlot.resize(T.size());
    for (size_t i = 0; i < T.size(); ++i) {
      tlLeaf[i] = tbld.lightLeaf[i].first >= 0 ? (uint32_t)newIdx[tbld.lightLeaf[i].first] : kNone;
      tlSlot[i] = tbld.lightLeaf[i].second;
    }
    d.num_tree_nodes = (uint32_t)tree.size();
  }
  s->tree = tree;
  s->num_tlights = (uint32_t)tl.size();

  // ------------------------------------------------ upload
  lgpu_status st;
#define UP(vec, field) if ((st = Upload(s, vec, &d.field)) != LGPU_OK) return st
  UP(bb.nodes, nodes);
  UP(bb.nodes4, nodes4);
  UP(btris, tris);
  UP(attr, tri_attr);
  UP(triVtx, tri_vtx);
  UP(pos, vtx_pos);
  UP(nrm, vtx_nrm);
  UP(tan, vtx_tan);
  UP(uv, vtx_uv);
  UP(meshes, meshes);
  UP(mats, materials);
  UP(lobes, lobes);
  UP(sockets, sockets);
  UP(code, code);
  UP(consts, consts);
  UP(texels, texels);
  UP(texRecs, textures);
  UP(lights, lights);
  UP(tl, tlights);
  // Triangle lights' vertices, emission socket and intensity by tree-light index, so light
  // sampling reads them in one step instead of through mesh, vertex-index and material records.
  std::vector<float4> triGeo(3 * tl.size(), make_float4(0.f, 0.f, 0.f, 0.f));
  for (size_t i = 0; i < tl.size(); ++i) {
    const LightRec &R = tl[i];
    if (R.kind != kLightTri) continue;
    const MeshRec &m = meshes[R.mesh];
    const uint4 tv = triVtx[m.first_tri + R.tri];
    const float4 p0 = pos[tv.x], p1 = pos[tv.y], p2 = pos[tv.z];
    triGeo[3 * i + 0] = make_float4(p0.x, p0.y, p0.z, AsF32(mats[m.material].emission));
    triGeo[3 * i + 1] = make_float4(p1.x, p1.y, p1.z, lights[R.parent].intensity);
    triGeo[3 * i + 2] = make_float4(p2.x, p2.y, p2.z, 0.f);
  }
  UP(triGeo, tri_geo);
  UP(treeOrder, tree_order);
  UP(tree, tree);
  // pick-walk records in breadth-first order: the top levels form a prefix that
  // shading blocks cache in LDS
  std::vector<uint32_t> bfsOf(tree.size(), kNone), bfsOrder;
  if (!tree.empty()) {
    bfsOrder.push_back(0);
    for (size_t q = 0; q < bfsOrder.size(); ++q) {
      const TreeNode &t = tree[bfsOrder[q]];
      for (int c = 0; c < 2; ++c)
        if (t.child[c] != kNone) bfsOrder.push_back(t.child[c]);
    }
    for (size_t q = 0; q < bfsOrder.size(); ++q) bfsOf[bfsOrder[q]] = (uint32_t)q;
  }
  std::vector<TreeImp> timp(bfsOrder.size());
  for (size_t q = 0; q < bfsOrder.size(); ++q) {
    const TreeNode &t = tree[bfsOrder[q]];
    TreeImp &m = timp[q];
    std::memset(&m, 0, sizeof(m));
    for (int a = 0; a < 3; ++a) { m.c[a] = t.center[a]; m.axis[a] = t.axis[a]; }
    m.r = t.diag * .5f;
    m.power = t.power;
    m.cosO = std::cos(t.thetaO); m.sinO = std::sin(t.thetaO); m.cosE = std::cos(t.thetaE);
    for (int c = 0; c < 2; ++c) m.child[c] = t.child[c] == kNone ? kNone : bfsOf[t.child[c]];
    m.first = t.first; m.leaf_dist = t.leaf_dist;
  }
  d.tree_lds = (uint32_t)std::min<size_t>(timp.size(), kTreeLds);
  UP(timp, timp);
  UP(bfsOf, tree_bfs);
  UP(triTL, tri_tlight);
  UP(tlLeaf, tlight_leaf);
  UP(tlSlot, tlight_slot);
  UP(pool.f, dist);
  UP(media, media);
  UP(density, density);
  UP(adjLobes, adj_lobes);
  s->density_off.assign(s->media.size(), 0);
  for (size_t k = 0; k < media.size(); ++k) s->density_off[k] = media[k].dens_off;
  s->texel_off.assign(s->textures.size(), 0);
  for (size_t t = 0; t < s->textures.size(); ++t) s->texel_off[t] = texRecs[texIds[t]].w;
#undef UP
  s->committed = true;
  return LGPU_OK;
}

}  // namespace lgpu
