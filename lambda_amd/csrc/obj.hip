// Wavefront OBJ ingestion (host only): the geometry path of AssetImporter::Import with
// aiProcess_Triangulate and aiProcess_CalcTangentSpace (AssetImporter.cpp:5-15) followed by
// MeshImport::LoadMeshVertexBuffers (MeshImport.cpp:11-92). Assimp (absent here, version
// unpinned) is restated from its published behaviour:
//   - one mesh per (object, material) run of faces, `o` and `usemtl` starting a new one;
//   - every face corner becomes its own vertex (the OBJ importer does not share vertices);
//   - quads are split from their concave corner if they have one, else from corner 0; larger
//     polygons are fanned from corner 0 (convex polygons only);
//   - tangents and bitangents per face from the uv derivatives, projected off the vertex normal,
//     then averaged over corners at the same position whose normals agree (cos >= 0.9999) and
//     whose tangents and bitangents are within 45 degrees; no normals or no uvs: no tangents.
// MeshImport then stores the bitangent's handedness sign(dot(cross(n, t), b)) in the tangent's w
// (:39-49), flips both uv coordinates (u, v) -> (1 - u, 1 - v) (:80-86) and shades smoothly only
// when tangents exist (:50). A MTL file named by `mtllib` is read for Kd and Ke.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "host.h"

using namespace lgpu;

namespace {

struct P3 { float x, y, z; };
P3 Sub(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
P3 Add(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
P3 Mul(P3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
float Dot(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
P3 Cross(P3 a, P3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
P3 Norm(P3 a) {
  const float l = std::sqrt(Dot(a, a));
  return l > 0.f ? Mul(a, 1.f / l) : a;
}

struct Corner { int v = -1, t = -1, n = -1; };

struct ObjMesh {
  std::string name;
  uint32_t material = 0;
  std::vector<float> pos, nrm, tan, uv;
  std::vector<uint32_t> idx;
};

}  // namespace

struct lgpu_obj_s {
  std::vector<ObjMesh> meshes;
  std::vector<std::string> mat_names;
  std::vector<lgpu_obj_material> mats;
};

namespace {

int Resolve(int i, size_t n) { return i > 0 ? i - 1 : (i < 0 ? (int)n + i : -1); }

bool ParseCorner(const std::string &tok, Corner &c, size_t nv, size_t nt, size_t nn) {
  int a = 0, b = 0, d = 0;
  const size_t s1 = tok.find('/');
  a = std::atoi(tok.substr(0, s1).c_str());
  if (s1 != std::string::npos) {
    const size_t s2 = tok.find('/', s1 + 1);
    const std::string t = tok.substr(s1 + 1, s2 == std::string::npos ? std::string::npos : s2 - s1 - 1);
    if (!t.empty()) b = std::atoi(t.c_str());
    if (s2 != std::string::npos) d = std::atoi(tok.substr(s2 + 1).c_str());
  }
  c.v = Resolve(a, nv);
  c.t = b ? Resolve(b, nt) : -1;
  c.n = d ? Resolve(d, nn) : -1;
  return c.v >= 0 && (size_t)c.v < nv && (c.t < 0 || (size_t)c.t < nt) && (c.n < 0 || (size_t)c.n < nn);
}

// Assimp's CalcTangentsProcess, per mesh of unshared corners (3 per triangle, in order).
void Tangents(ObjMesh &m) {
  const size_t nv = m.pos.size() / 3;
  std::vector<P3> T(nv), B(nv);
  auto P = [&](size_t i) { return P3{m.pos[3 * i], m.pos[3 * i + 1], m.pos[3 * i + 2]}; };
  auto N = [&](size_t i) { return P3{m.nrm[3 * i], m.nrm[3 * i + 1], m.nrm[3 * i + 2]}; };
  for (size_t f = 0; f < m.idx.size() / 3; ++f) {
    const uint32_t i0 = m.idx[3 * f], i1 = m.idx[3 * f + 1], i2 = m.idx[3 * f + 2];
    const P3 v = Sub(P(i1), P(i0)), w = Sub(P(i2), P(i0));
    float sx = m.uv[2 * i1] - m.uv[2 * i0], sy = m.uv[2 * i1 + 1] - m.uv[2 * i0 + 1];
    float tx = m.uv[2 * i2] - m.uv[2 * i0], ty = m.uv[2 * i2 + 1] - m.uv[2 * i0 + 1];
    const float dir = (tx * sy - ty * sx) < 0.f ? -1.f : 1.f;
    if (sx * ty == sy * tx) { sx = 0.f; sy = 1.f; tx = 1.f; ty = 0.f; }
    const P3 t = Mul(Sub(Mul(w, sy), Mul(v, ty)), dir);
    const P3 b = Mul(Sub(Mul(w, sx), Mul(v, tx)), dir);
    for (uint32_t i : {i0, i1, i2}) {
      const P3 n = N(i);
      T[i] = Norm(Sub(t, Mul(n, Dot(t, n))));
      B[i] = Norm(Sub(b, Mul(n, Dot(b, n))));
    }
  }
  // smoothing over corners at one position (exact match of the parsed coordinates)
  std::map<std::tuple<float, float, float>, std::vector<uint32_t>> at;
  for (uint32_t i = 0; i < nv; ++i) at[std::make_tuple(m.pos[3 * i], m.pos[3 * i + 1], m.pos[3 * i + 2])].push_back(i);
  const float lim = std::cos(45.f * 3.14159265358979f / 180.f);
  std::vector<char> done(nv, 0);
  std::vector<P3> Ts = T, Bs = B;
  for (uint32_t i = 0; i < nv; ++i) {
    if (done[i]) continue;
    const auto &cand = at[std::make_tuple(m.pos[3 * i], m.pos[3 * i + 1], m.pos[3 * i + 2])];
    std::vector<uint32_t> close;
    for (uint32_t j : cand) {
      if (done[j]) continue;
      if (Dot(N(j), N(i)) < .9999f || Dot(T[j], T[i]) < lim || Dot(B[j], B[i]) < lim) continue;
      close.push_back(j);
    }
    P3 st{0, 0, 0}, sb{0, 0, 0};
    for (uint32_t j : close) { st = Add(st, T[j]); sb = Add(sb, B[j]); }
    st = Norm(st);
    sb = Norm(sb);
    for (uint32_t j : close) { Ts[j] = st; Bs[j] = sb; done[j] = 1; }
  }
  m.tan.assign(4 * nv, 0.f);
  for (uint32_t i = 0; i < nv; ++i) {
    const P3 t{std::isfinite(Ts[i].x) ? Ts[i].x : 0.f, std::isfinite(Ts[i].y) ? Ts[i].y : 0.f,
               std::isfinite(Ts[i].z) ? Ts[i].z : 0.f};  // MeshImport::CheckReal
    const P3 b{std::isfinite(Bs[i].x) ? Bs[i].x : 0.f, std::isfinite(Bs[i].y) ? Bs[i].y : 0.f,
               std::isfinite(Bs[i].z) ? Bs[i].z : 0.f};
    m.tan[4 * i] = t.x; m.tan[4 * i + 1] = t.y; m.tan[4 * i + 2] = t.z;
    m.tan[4 * i + 3] = Dot(Cross(N(i), t), b) < 0.f ? -1.f : 1.f;
  }
}

void ReadMtl(const std::string &path, lgpu_obj_s &o) {
  std::ifstream f(path);
  if (!f) return;
  std::string line;
  int cur = -1;
  while (std::getline(f, line)) {
    std::istringstream ls(line);
    std::string k;
    if (!(ls >> k)) continue;
    if (k == "newmtl") {
      std::string n;
      ls >> n;
      cur = -1;
      for (size_t i = 0; i < o.mat_names.size(); ++i) if (o.mat_names[i] == n) cur = (int)i;
      if (cur < 0) {
        o.mat_names.push_back(n);
        lgpu_obj_material m;
        std::memset(&m, 0, sizeof(m));
        o.mats.push_back(m);
        cur = (int)o.mats.size() - 1;
      }
    } else if (cur >= 0 && (k == "Kd" || k == "Ke")) {
      float v[3] = {0, 0, 0};
      ls >> v[0] >> v[1] >> v[2];
      lgpu_obj_material &m = o.mats[cur];
      float *dst = k == "Kd" ? m.kd : m.ke;
      for (int i = 0; i < 3; ++i) dst[i] = v[i];
      (k == "Kd" ? m.has_kd : m.has_ke) = 1;
    }
  }
}

}  // namespace

extern "C" {

lgpu_status lgpu_obj_load(const char *path, lgpu_obj *out) {
  if (!path || !out) return Fail(LGPU_ERR_INVALID_ARGUMENT, "obj_load: null argument");
  std::ifstream f(path);
  if (!f) return Fail(LGPU_ERR_INVALID_ARGUMENT, std::string("obj_load: cannot open ") + path);
  std::string dir = path;
  const size_t slash = dir.find_last_of('/');
  dir = slash == std::string::npos ? std::string() : dir.substr(0, slash + 1);
  std::vector<P3> V, N;
  std::vector<float> TC;
  auto *o = new lgpu_obj_s;
  std::string objName = "default";
  int curMat = -1;
  // faces of the current run; a new run starts at `o` and `usemtl`
  struct Run { std::string name; int mat; std::vector<std::vector<Corner>> faces; };
  std::vector<Run> runs(1, Run{objName, -1, {}});
  auto matIndex = [&](const std::string &n) {
    for (size_t i = 0; i < o->mat_names.size(); ++i) if (o->mat_names[i] == n) return (int)i;
    o->mat_names.push_back(n);
    lgpu_obj_material m;
    std::memset(&m, 0, sizeof(m));
    o->mats.push_back(m);
    return (int)o->mats.size() - 1;
  };
  std::vector<std::string> mtllibs;
  std::string line;
  size_t lineNo = 0;
  while (std::getline(f, line)) {
    ++lineNo;
    std::istringstream ls(line);
    std::string k;
    if (!(ls >> k) || k[0] == '#') continue;
    if (k == "v") { P3 p{0, 0, 0}; ls >> p.x >> p.y >> p.z; V.push_back(p); }
    else if (k == "vn") { P3 p{0, 0, 0}; ls >> p.x >> p.y >> p.z; N.push_back(p); }
    else if (k == "vt") { float u = 0, v = 0; ls >> u >> v; TC.push_back(u); TC.push_back(v); }
    else if (k == "o") { ls >> objName; runs.push_back(Run{objName, curMat, {}}); }
    else if (k == "usemtl") { std::string n; ls >> n; curMat = matIndex(n); runs.push_back(Run{objName, curMat, {}}); }
    else if (k == "mtllib") { std::string n; ls >> n; mtllibs.push_back(n); }
    else if (k == "f") {
      std::vector<Corner> face;
      std::string tok;
      while (ls >> tok) {
        Corner c;
        if (!ParseCorner(tok, c, V.size(), TC.size() / 2, N.size())) {
          delete o;
          return Fail(LGPU_ERR_INVALID_ARGUMENT, "obj_load: bad face index at line " + std::to_string(lineNo));
        }
        face.push_back(c);
      }
      if (face.size() >= 3) runs.back().faces.push_back(std::move(face));
    }
  }
  for (const std::string &lib : mtllibs) ReadMtl(dir + lib, *o);
  for (Run &r : runs) {
    if (r.faces.empty()) continue;
    ObjMesh m;
    m.name = r.name;
    m.material = r.mat < 0 ? (uint32_t)matIndex("") : (uint32_t)r.mat;
    bool allN = true, allT = true;
    for (const auto &fc : r.faces)
      for (const Corner &c : fc) { allN = allN && c.n >= 0; allT = allT && c.t >= 0; }
    for (const auto &fc : r.faces) {
      const uint32_t base = (uint32_t)(m.pos.size() / 3);
      for (const Corner &c : fc) {  // one vertex per face corner
        m.pos.insert(m.pos.end(), {V[c.v].x, V[c.v].y, V[c.v].z});
        if (allN) m.nrm.insert(m.nrm.end(), {N[c.n].x, N[c.n].y, N[c.n].z});
        if (allT) m.uv.insert(m.uv.end(), {TC[2 * c.t], TC[2 * c.t + 1]});
      }
      const size_t n = fc.size();
      if (n == 4) {  // TriangulateProcess: split a quad from its concave corner
        uint32_t start = 0;
        for (uint32_t i = 0; i < 4; ++i) {
          const P3 v = V[fc[i].v];
          const P3 left = Norm(Sub(V[fc[(i + 3) % 4].v], v)), diag = Norm(Sub(V[fc[(i + 2) % 4].v], v)),
                   right = Norm(Sub(V[fc[(i + 1) % 4].v], v));
          const float ang = std::acos(std::fmax(-1.f, std::fmin(1.f, Dot(left, diag)))) +
                            std::acos(std::fmax(-1.f, std::fmin(1.f, Dot(right, diag))));
          if (ang > 3.14159265358979f) { start = i; break; }
        }
        m.idx.insert(m.idx.end(), {base + start, base + (start + 1) % 4, base + (start + 2) % 4});
        m.idx.insert(m.idx.end(), {base + start, base + (start + 2) % 4, base + (start + 3) % 4});
      } else {
        for (size_t i = 1; i + 1 < n; ++i) m.idx.insert(m.idx.end(), {base, base + (uint32_t)i, base + (uint32_t)i + 1});
      }
    }
    if (allN && allT) Tangents(m);
    else m.nrm.clear();  // no tangent space: flat shading (MeshImport.cpp:50)
    for (size_t i = 0; i < m.uv.size(); ++i) m.uv[i] = 1.f - m.uv[i];  // MeshImport.cpp:84-85
    o->meshes.push_back(std::move(m));
  }
  if (o->meshes.empty()) {
    delete o;
    return Fail(LGPU_ERR_INVALID_ARGUMENT, std::string("obj_load: no faces in ") + path);
  }
  for (size_t i = 0; i < o->mats.size(); ++i) o->mats[i].name = o->mat_names[i].c_str();
  *out = o;
  return LGPU_OK;
}

lgpu_status lgpu_obj_counts(lgpu_obj o, uint32_t *num_meshes, uint32_t *num_materials) {
  if (!o) return Fail(LGPU_ERR_INVALID_ARGUMENT, "obj_counts: null object");
  if (num_meshes) *num_meshes = (uint32_t)o->meshes.size();
  if (num_materials) *num_materials = (uint32_t)o->mats.size();
  return LGPU_OK;
}

lgpu_status lgpu_obj_mesh(lgpu_obj o, uint32_t i, lgpu_mesh_desc *d, const char **name) {
  if (!o || !d || i >= o->meshes.size()) return Fail(LGPU_ERR_INVALID_ARGUMENT, "obj_mesh: bad object or index");
  const ObjMesh &m = o->meshes[i];
  std::memset(d, 0, sizeof(*d));
  d->num_vertices = (uint32_t)(m.pos.size() / 3);
  d->num_triangles = (uint32_t)(m.idx.size() / 3);
  d->positions = m.pos.data();
  d->indices = m.idx.data();
  d->normals = m.nrm.empty() ? nullptr : m.nrm.data();
  d->tangents = m.tan.empty() || m.nrm.empty() ? nullptr : m.tan.data();
  d->uvs = m.uv.empty() ? nullptr : m.uv.data();
  d->material = m.material;
  if (name) *name = m.name.c_str();
  return LGPU_OK;
}

lgpu_status lgpu_obj_material_info(lgpu_obj o, uint32_t i, lgpu_obj_material *out) {
  if (!o || !out || i >= o->mats.size()) return Fail(LGPU_ERR_INVALID_ARGUMENT, "obj_material_info: bad object or index");
  *out = o->mats[i];
  return LGPU_OK;
}

lgpu_status lgpu_obj_destroy(lgpu_obj o) {
  delete o;
  return LGPU_OK;
}

}  // extern "C"
