// Internal data layouts shared by the host scene builder (C++) and the HIP
// kernels. Everything here lives in HBM as flat arrays; the kernels read them
// through the DevScene pointer bundle, passed by value (the path shade kernel
// reads a device copy of it through a pointer).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lgpu {

constexpr float kPi = 3.14159265358979323846264338f;
constexpr float kSurfaceEpsilon = 1e-5f;  // ScatterEvent.h:12
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kLdsStack = 16;             // per-thread BVH stack entries held in LDS
constexpr int kMaxBvhDepth = 64;          // binary-BVH depth bound (builder guarantees)
constexpr int kMaxStack = 96;             // 4-wide traversal stack entries per thread (LDS + spill)
constexpr int kSpillWords = kMaxStack - kLdsStack;  // per-thread global spill, worst case over trees

// ---------------------------------------------------------------- BVH2
// Two child boxes per node (64 B, one node = 4 x 16-B loads):
//   a = {c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y}
//   b = {c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y}
//   c = {c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z}
//   d = {child0, child1, -, -}; a child >= 0 is an inner node index, a child
//       with the top bit set is a leaf: bits 0..25 first triangle (BVH order),
//       bits 26..30 triangle count.
struct BvhNode {
  float a[4], b[4], c[4];
  uint32_t child[2], pad[2];
};
constexpr uint32_t kLeafBit = 0x80000000u;
inline uint32_t MakeLeaf(uint32_t first, uint32_t count) { return kLeafBit | (count << 26) | first; }

// 4-wide node with 8-bit quantized child boxes (64 B = 4 x 16-B loads). Child k's
// box on axis a is [p_a + qlo_a[k] * s_a, p_a + qhi_a[k] * s_a] with s_a a power
// of two, so the decode q * s is exact; boxes are quantized outwards from
// padded child bounds. q[2a] / q[2a+1] hold the lo / hi bytes of the four
// children (byte k = child k). child[k]: inner node index, leaf ref (as
// BvhNode), or kNone for an empty slot.
//   load 0 = {p.x, p.y, p.z, s.x}, load 1 = {s.y, s.z, qlo.x, qhi.x},
//   load 2 = {qlo.y, qhi.y, qlo.z, qhi.z}, load 3 = {child0..3}
struct Bvh4Node {
  float p[3];
  float s[3];
  uint32_t q[6];
  uint32_t child[4];
};

// Triangles in BVH order: {v0.xyz, prim}, {e1.xyz, mesh}, {e2.xyz, shade key} with
// e1 = v1 - v0, e2 = v2 - v0 (48 B). Ng = e1 x e2 is bit-identical to
// (v1 - v0) x (v2 - v0).
struct BvhTri {
  float v0[3];
  uint32_t prim;
  float e1[3];
  uint32_t mesh;
  float e2[3];
  uint32_t key;  // ShadeKey of the triangle's material
};

// Shading-code family of a path vertex; the path shade stage groups a block's paths by it so
// that a wave runs one material family's code. 0 = ray left the scene, 1 = pass-through
// surface, 2 + (root lobe kind - kLambertian) otherwise.
constexpr uint32_t kShadeKeys = 16;
constexpr uint32_t kKeyGuidable = 0x80000000u;  // key bit: the material may be guided (no delta lobe)
// Shading attributes per triangle in BVH order (7 x float4 = 112 B), present
// for meshes with smooth normals or uvs:
//   {n0, uv0.x} {n1, uv0.y} {n2, uv1.x} {t0} {t1} {t2} {uv1.y, uv2.x, uv2.y, 0}
constexpr int kTriAttrF4 = 7;

// ---------------------------------------------------------------- geometry for shading
struct MeshRec {
  uint32_t material;
  uint32_t first_tri;   // global triangle index of local triangle 0
  uint32_t num_tris;
  uint32_t flags;       // bit0: smooth normals+tangents, bit1: uvs
  uint32_t mesh_light;  // scene light index or kNone
  uint32_t first_vtx;
  uint32_t pad[2];
};
constexpr uint32_t kMeshSmooth = 1u, kMeshUV = 2u;

// ---------------------------------------------------------------- lights
enum LightKind : uint32_t { kLightMesh = 0, kLightTri = 1, kLightPoint = 2, kLightSpot = 3, kLightEnv = 4 };
struct LightRec {
  uint32_t kind;
  uint32_t mesh;        // mesh index (mesh / triangle lights)
  uint32_t tri;         // triangle light: local triangle index
  uint32_t parent;      // triangle light: scene index of its MeshLight
  uint32_t cdf_off;     // mesh light: offset of its triangle-area Piecewise1D in dist pool
  uint32_t count;       // mesh light: number of triangles
  float intensity;
  float temperature;
  float rgb[4];         // point/spot colour (RGB-mode blackbody precomputed)
  float pos[4];
  float axis[4];
  float cos_cone, cos_falloff, power, pad;
};

// Piecewise1D (Piecewise.cpp:9-44) in a flat pool: [n, integral, pdf[n], cdf[n+1]]
// stored as floats starting at an offset; n stored as float bits.

// Light tree (ManyLightSampler.h:48-76), flattened depth-first, left child first.
struct TreeNode {
  float center[3];
  float diag;
  float axis[3];
  float thetaO;
  float thetaE;
  float power;
  uint32_t child[2];    // kNone for leaves
  uint32_t first;       // leaf: first position in the tree light order
  uint32_t count;
  uint32_t parent;      // kNone at the root
  uint32_t leaf_dist;   // leaf: offset of its Piecewise1D in the dist pool
};

// Light-tree node as the pick walk reads it (64 B): bounds centre and half diagonal,
// cone axis and power, cos/sin of the cone's thetaO and cos of thetaE (the importance
// measure without inverse trigonometry), children and leaf data.
struct TreeImp {
  float c[3], r;
  float axis[3], power;
  float cosO, sinO, cosE, pad;
  uint32_t child[2];    // kNone for leaves
  uint32_t first;       // leaf: first position in the tree light order
  uint32_t leaf_dist;   // leaf: offset of its Piecewise1D in the dist pool
};

constexpr int kTreeLds = 255;  // light-tree records cached in LDS per shading block (8 levels)

struct EnvRec {
  uint32_t w, h;
  uint32_t tex_off;     // float offset of RGB data in the texel pool
  uint32_t marg_off;    // Piecewise1D offset of the marginal
  uint32_t cond_off;    // offset of conditional 0; conditional v at cond_off + v * cond_stride
  uint32_t cond_stride;
  float offset[2];
  float intensity;
  float radius;
  uint32_t light;       // scene light index
  uint32_t pad;
};

// ---------------------------------------------------------------- materials / bytecode
// Socket value types (ShaderGraph SocketType)
enum ValType : uint32_t { kScalar = 0, kVec2 = 1, kVec3 = 2, kColour = 3, kSpectrum = 4, kBxdf = 5 };

// VM instructions: 16 B. Registers are float4; programs are straight-line.
enum Op : uint32_t {
  kOpConst = 1,      // r[dst] = consts[imm]
  kOpUV = 2,         // r[dst] = (u, v, 0, 0)
  kOpTex = 3,        // r[dst] = bilinear(texture imm, fract(uv))
  kOpChecker = 4,    // r[dst].x = checker(uv * consts[imm].x)
  kOpFresnel = 5,    // r[dst].x = Schlick(eta, r[a].x, woL.y)
  kOpSMaths = 6,     // r[dst].x = r[a].x op r[b].x ; op in c
  kOpVMaths = 7,     // r[dst].xyz = r[a] op r[b] ; op in c
  kOpDot = 8,
  kOpLength = 9,
  kOpMerge = 10,     // r[dst] = (r[a].x, r[b].x, r[c].x)
  kOpSeparate = 11,  // r[dst].x = r[a][c]
  kOpBlackbody = 12, // r[dst] = normalised Planck at the hero wavelengths of T = r[a].x
  kOpToReal = 13,    // r[dst].x = AsReal(r[a], type c)
  kOpToColour = 14,  // r[dst] = AsColour(r[a], type c)
  kOpToVec3 = 15,
  kOpToSpectrum = 16, // r[dst] = AsSpectrum(r[a], type c)
  kOpNormal = 17,    // r[dst].xyz = shading normal
  kOpTexChannel = 18, // r[dst].x = channel c of bilinear(texture imm, fract(uv))
  kOpCross = 19,     // r[dst].xyz = r[a] x r[b]
  kOpNoise = 20,     // r[dst].x = noise c (node type) at uv (or the point at a medium
                     //   interaction) scaled by consts[imm].x, .y, .z in turn
  kOpSpectral = 21,  // r[dst] = sampled SPD of texture consts[imm].z over [consts[imm].x, .y] nm
                     //   at the hero wavelengths; RGB mode: consts[imm + 1]
  kOpSpecTex = 22,   // r[dst] = RGB texel of texture imm as a spectrum
  kOpAlpha = 23      // r[dst] = (r[a].xyz, r[b].x)
};
struct Instr {
  uint32_t op, dst, a, b, c, imm, pad[2];
};
constexpr int kMaxRegs = 16;

// A socket read by a lobe or an emitter, already converted to the type the
// reader requests (GetAs<Real>, GetAs<Vec2> or GetAsSpectrum).
enum SocketKind : uint32_t {
  kSockScalar = 0,      // value.x (as spectrum: broadcast)
  kSockColour = 1,      // value.rgb; as spectrum: RGB in RGB mode, uplift at the hero wavelengths
  kSockBlackbody = 2,   // value.rgb = RGB-mode blackbody, value.w = temperature
  kSockVec2 = 3,        // value.xy (anisotropic roughness)
  kSockProgram = 4      // run code[prog_off .. +prog_len); result in r[res]
};
struct Socket {
  float value[4];
  uint32_t kind;
  uint32_t prog_off, prog_len, res;
  uint32_t res_type;    // ValType of r[res] for program sockets
  uint32_t pad[3];
};

enum LobeKind : uint32_t {
  kLambertian = 64, kOrenNayarR = 65, kOrenNayarT = 66, kMicrofacet = 67, kFresnelBsdf = 68,
  kSpecularR = 69, kSpecularT = 70, kGhost = 71, kMix = 72, kRoughDielectric = 73
};
struct Lobe {
  uint32_t kind;
  uint32_t sock[3];     // socket indices (albedo, roughness/sigma/ior/alpha/ratio ...)
  uint32_t child[2];    // mix children (lobe indices)
  float p[10];          // constant parameters (microfacet p[0..9], ior, etaT)
};

// Differentiable texture bindings of one lobe (the adjoint render). A socket bound to a texture
// reads one bilinear lookup of it through at most one conversion, so its value is linear in the
// lookup's four components (TexBilinear's out[0..3]):
//   albedo (spectrum socket, RGB mode): albedo_c = sum_k albedo_w[c][k] out[k]
//   roughness (scalar socket):          r        = sum_k rough_w[k] out[k]
// kNone: the socket is not texture-bound (its gradient is not taken).
struct AdjLobe {
  uint32_t albedo_tex;   // DevScene::textures index
  uint32_t rough_tex;
  uint32_t pad[2];
  float albedo_w[3][4];
  float rough_w[4];
};

struct MaterialRec {
  uint32_t lobe;        // root lobe index, kNone = pass-through
  uint32_t emission;    // socket index, kNone = not emissive
  float intensity;
  uint32_t med_in, med_out;
  uint32_t pad[3];
};

// Medium (HomogeneousMedium.h; the grid is piecewise constant over [bmin, bmax], z-major).
struct MediumRec {
  uint32_t type;
  float sa[3], ss[3];
  float g;
  uint32_t nx, ny, nz;
  uint32_t dens_off;  // offset into DevScene::density
  float bmin[3], bmax[3];
};

// ---------------------------------------------------------------- path state (SoA)
enum PathFlags : uint32_t {
  kFlagCamera = 1u,        // ray came from the camera (or pass-through at depth 0)
  kFlagEmitOnly = 2u,      // terminated (RR / max depth): gather emission at the next hit only
  kFlagSpecular = 4u,      // last BSDF sample was a delta lobe
  kFlagNoEmit = 8u,        // pass-through at depth > 0: no emission at the next hit
  kFlagRR = 16u,           // RR survived: divide beta by (1 - q) at the next vertex
  kFlagDead = 32u,
  kFlagSideNeg = 64u,      // ScatterEvent::sidedness == -1 (kept along volumetric paths)
  kFlagTrMIS = 128u        // the sampled light is looked up through pass-through surfaces with Tr
};

// Active paths are compacted: slot i of the current buffer holds the i-th
// live path; shade appends survivors to the other buffer. Radiance lives in a
// per-sample array indexed by the sample slot `sid` of the wave.
struct PathState {
  float4 *ray_o;     // xyz, w = lambda0 (hero mode)
  float4 *ray_d;     // xyz, w = ScatterEvent::eta (persists along the path)
  float4 *beta;
  float4 *hit;       // {t, u, v, BVH-order triangle index bits}
  uint4 *ids;        // {pixel, sample, dim, (depth << 8) | flags}
  uint4 *aux;        // {sid, NEE light of the previous vertex, pass-through count, medium}
  float4 *mis;       // {scatter pdf, light dist pdf, q, weight factor (volumetric paths)}
  float4 *prev_p;    // shading point of the previous vertex
  float4 *prev_n;    // shading normal of the previous vertex
};
struct SampleState {
  float4 *L;         // radiance per sample slot
  uint32_t *pixel;
  float *lambda0;
};

struct ShadowRay {
  float o[3];
  float tmax;
  float d[3];
  uint32_t sid;
  float contrib[4];
};

struct DevScene {
  const BvhNode *nodes;       // binary BVH (traversal-count passes)
  const Bvh4Node *nodes4;     // 4-wide quantized BVH (rendering)
  const BvhTri *tris;
  const float4 *tri_attr;    // kTriAttrF4 per BVH-order triangle (smooth / uv meshes)
  const uint4 *tri_vtx;      // {i0, i1, i2, mesh} global vertex indices
  const float4 *vtx_pos;
  const float4 *vtx_nrm;
  const float4 *vtx_tan;
  const float2 *vtx_uv;
  const MeshRec *meshes;
  const MaterialRec *materials;
  const Lobe *lobes;
  const Socket *sockets;
  const Instr *code;
  const float4 *consts;
  const float *texels;
  const uint4 *textures;     // {w, h, channels, offset}
  const LightRec *lights;    // scene lights (add order)
  const LightRec *tlights;   // tree lights
  const float4 *tri_geo;     // 3 per tree light: triangle lights' {v0, emission socket}, {v1, intensity}, {v2, 0}
  const uint32_t *tree_order;// leaf position -> tree light index
  const TreeNode *tree;
  const TreeImp *timp;       // same nodes, pick-walk layout, breadth-first order
  const uint32_t *tree_bfs;  // tree (depth-first) index -> timp (breadth-first) index
  uint32_t tree_lds;         // leading timp records a shading block caches in LDS
  const uint32_t *tri_tlight;// global triangle -> tree light (kNone)
  const uint32_t *tlight_leaf; // tree light -> (leaf node index)
  const uint32_t *tlight_slot; // tree light -> position inside its leaf
  const float *dist;         // Piecewise1D pool
  EnvRec env;
  uint32_t has_env;
  uint32_t has_mesh_lights;  // some mesh is emissive (a BSDF- or phase-sampled ray can reach a light)
  uint32_t env_light;        // scene light index of the environment light
  uint32_t strategy;         // 0 power, 1 tree
  uint32_t num_lights;
  uint32_t power_dist;       // Piecewise1D offset (power sampler)
  float inv_total_power;
  float tree_power, inf_power;
  uint32_t num_tree_nodes;
  uint32_t num_tris;
  const MediumRec *media;
  const float *density;      // grid densities of all media
  uint32_t num_media;
  uint32_t spill_words;      // per-thread traversal-stack words beyond the LDS part this tree can need
  uint32_t has_programs;     // some socket runs VM code (shading kernels specialise on it)
  unsigned long long *voxel_count;  // non-null in counting renders: voxels visited by grid walks
  const AdjLobe *adj_lobes;  // per lobe: texture bindings for the adjoint render
};

struct CameraRec {
  uint32_t type;
  float origin[3], x_hat[3], y_hat[3], z_hat[3];
  float tan_fov, aspect, focal_length;
  uint32_t aperture;
  float aperture_size;
  uint32_t blades;
  float offset_phi, offset_theta;
};

// Output queues of a shade launch.
struct Queues {
  PathState next;      // compacted survivors
  ShadowRay *shadow;
  uint32_t *counters;  // [0] next count, [1] shadow count
};

struct RenderParams {
  CameraRec cam;
  uint32_t width, height;
  uint32_t spp_begin;
  uint32_t max_bounces, min_bounces;
  uint32_t mode;        // 0 RGB, 1 HERO4
  uint32_t corrected;   // 0 faithful, 1 corrected
  uint32_t seed;
  uint32_t integrator;
  float pass_max_depth, pass_power;  // utility passes
  float pass_bg[3];
  uint32_t pass_normal_g;
};

// Batched renders (batch.hip) record the single-render kernels through these (render.hip).
void BatchLaunchInit(hipStream_t s, uint32_t *ctr, uint32_t words, uint32_t n);
void BatchLaunchTrunc(hipStream_t s, const uint32_t *cnt, unsigned long long *acc);
void BatchLaunchCamera(hipStream_t s, const RenderParams &rp, const PathState &ps, const SampleState &ss, uint32_t n,
                       uint64_t base);
void BatchLaunchExtend(hipStream_t s, const DevScene &dsc, const PathState &ps, uint32_t ub, const uint32_t *np,
                       uint32_t *spill);
void BatchLaunchShade(hipStream_t s, const DevScene *scp, const RenderParams &rp, const PathState &ps,
                      const SampleState &ss, uint32_t ub, const uint32_t *np, const Queues &q, bool programs);
void BatchLaunchShadow(hipStream_t s, const DevScene &dsc, const SampleState &ss, const ShadowRay *rays, uint32_t ub,
                       const uint32_t *np, uint32_t *spill);
void BatchLaunchAccumulate(hipStream_t s, const RenderParams &rp, const SampleState &ss, uint32_t n, uint32_t npx,
                           double *film, uint32_t *counts);

// Volumetric shade stage (volume.hip): VolumetricPathIntegrator / MISVolumetricPathIntegrator;
// with `adj`, the adjoint replay (adjoint.h).
struct AdjRefs;
void LaunchShadeVol(uint32_t blocks, hipStream_t stream, const DevScene &sc, const RenderParams &rp, PathState ps,
                    SampleState ss, uint32_t n, Queues q, uint32_t *spill, const AdjRefs *adj = nullptr,
                    const uint32_t *np = nullptr);
// Utility-pass shade stage (utility.hip): DepthPass, NormalPass, AOVPass.
void LaunchShadeUtility(uint32_t blocks, hipStream_t stream, const DevScene &sc, const RenderParams &rp, PathState ps,
                        SampleState ss, uint32_t n);
// Direct-lighting shade stage (direct.hip): DirectLightingIntegrator.
void LaunchShadeDirect(uint32_t blocks, hipStream_t stream, const DevScene &sc, const RenderParams &rp, PathState ps,
                       SampleState ss, uint32_t n, Queues q, uint32_t *spill);

}  // namespace lgpu
