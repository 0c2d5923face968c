// C ABI entry points: device, scene assembly, film, camera helpers, errors.
// The render entry points live in render.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "dmath.h"
#include "host.h"

namespace lgpu {
static thread_local std::string g_error;
void SetError(const std::string &msg) { g_error = msg; }
lgpu_status Fail(lgpu_status s, const std::string &msg) {
  g_error = msg;
  return s;
}
}  // namespace lgpu

using namespace lgpu;

#define CHECK_ARG(cond, msg) \
  if (!(cond)) return Fail(LGPU_ERR_INVALID_ARGUMENT, msg)

extern "C" {

const char *lgpu_last_error(void) { return g_error.c_str(); }
uint32_t lgpu_abi_version(void) { return LGPU_ABI_VERSION; }

// lambdaCreateDevice / lambdaReleaseDevice (Lambda.cpp:148-162)
lgpu_status lgpu_device_create(int hip_device, lgpu_device *out) {
  CHECK_ARG(out, "device_create: null out");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return Fail(LGPU_ERR_NO_DEVICE, "device_create: no HIP device");
  CHECK_ARG(hip_device >= 0 && hip_device < n, "device_create: device index out of range");
  LGPU_HIP(hipSetDevice(hip_device));
  lgpu_device d = new lgpu_device_s();
  d->device = hip_device;
  if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
    delete d;
    return Fail(LGPU_ERR_HIP, "device_create: stream creation failed");
  }
  for (auto &e : d->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      delete d;
      return Fail(LGPU_ERR_HIP, "device_create: event creation failed");
    }
  }
  if (ResolveScratch(d) != LGPU_OK) {
    const std::string msg = lgpu_last_error();
    for (auto &e : d->ev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(d->stream);
    delete d;
    return Fail(LGPU_ERR_HIP, msg);
  }
  *out = d;
  return LGPU_OK;
}

lgpu_status lgpu_device_destroy(lgpu_device d) {
  if (!d) return LGPU_OK;
  (void)hipSetDevice(d->device);
  (void)hipStreamSynchronize(d->stream);
  if (d->wave && d->wave_free) d->wave_free(d->wave);
  if (d->batch && d->batch_free) d->batch_free(d->batch);
  // guides still alive lose their device memory here; their handles stay valid for destroy
  if (d->guide_orphan)
    for (void *g : d->guides) d->guide_orphan(g);
  for (auto &e : d->ev)
    if (e) (void)hipEventDestroy(e);
  if (d->stream) (void)hipStreamDestroy(d->stream);
  delete d;
  return LGPU_OK;
}

// lambdaCreateScene (Lambda.cpp:168-173)
lgpu_status lgpu_scene_create(lgpu_device dev, lgpu_scene *out) {
  CHECK_ARG(dev && out, "scene_create: null argument");
  lgpu_scene s = new lgpu_scene_s();
  s->dev = dev;
  *out = s;
  return LGPU_OK;
}

lgpu_status lgpu_scene_destroy(lgpu_scene s) {
  if (!s) return LGPU_OK;
  (void)hipSetDevice(s->dev->device);
  delete s;
  return LGPU_OK;
}

lgpu_status lgpu_scene_add_texture(lgpu_scene s, const lgpu_texture_desc *d, uint32_t *id) {
  CHECK_ARG(s && d && d->data, "add_texture: null argument");
  CHECK_ARG(!s->committed, "add_texture: scene already committed");
  CHECK_ARG(d->width && d->height && (d->channels == 1 || d->channels == 3 || d->channels == 4), "add_texture: bad size or channel count");
  s->textures.emplace_back(d->data, d->data + (size_t)d->width * d->height * d->channels);
  s->tex_dims.push_back(d->width);
  s->tex_dims.push_back(d->height);
  s->tex_dims.push_back(d->channels);
  if (id) *id = (uint32_t)s->textures.size() - 1;
  return LGPU_OK;
}

lgpu_status lgpu_scene_add_medium(lgpu_scene s, const lgpu_medium_desc *d, uint32_t *id) {
  CHECK_ARG(s && d, "add_medium: null argument");
  CHECK_ARG(!s->committed, "add_medium: scene already committed");
  CHECK_ARG(d->type == LGPU_MEDIUM_HOMOGENEOUS || d->type == LGPU_MEDIUM_GRID, "add_medium: bad type");
  lgpu_medium_desc c = *d;
  s->media_density.emplace_back();
  if (d->type == LGPU_MEDIUM_GRID) {
    CHECK_ARG(d->density && d->nx && d->ny && d->nz, "add_medium: grid needs density and dimensions");
    s->media_density.back().assign(d->density, d->density + (size_t)d->nx * d->ny * d->nz);
  }
  c.density = nullptr;
  s->media.push_back(c);
  if (id) *id = (uint32_t)s->media.size() - 1;
  return LGPU_OK;
}

// In-place parameter updates of a committed scene (inverse rendering loops): the device copy and
// the host copy change together; nothing derived from the texture or grid needs rebuilding.
lgpu_status lgpu_scene_update_texture(lgpu_scene s, uint32_t id, const float *texels) {
  CHECK_ARG(s && texels, "update_texture: null argument");
  if (!s->committed) return Fail(LGPU_ERR_STATE, "update_texture: scene not committed");
  CHECK_ARG(id < s->textures.size(), "update_texture: unknown texture id");
  if (s->tex_emissive[id])
    return Fail(LGPU_ERR_UNSUPPORTED, "update_texture: an emission socket reads this texture (re-commit the scene)");
  std::vector<float> &t = s->textures[id];
  std::copy(texels, texels + t.size(), t.begin());
  LGPU_HIP(hipSetDevice(s->dev->device));
  LGPU_HIP(hipMemcpy(const_cast<float *>(s->dsc.texels) + s->texel_off[id], t.data(), t.size() * sizeof(float),
                     hipMemcpyHostToDevice));
  return LGPU_OK;
}

lgpu_status lgpu_scene_update_density(lgpu_scene s, uint32_t id, const float *density) {
  CHECK_ARG(s && density, "update_density: null argument");
  if (!s->committed) return Fail(LGPU_ERR_STATE, "update_density: scene not committed");
  CHECK_ARG(id < s->media.size(), "update_density: unknown medium id");
  CHECK_ARG(s->media[id].type == LGPU_MEDIUM_GRID, "update_density: the medium has no density grid");
  std::vector<float> &g = s->media_density[id];
  std::copy(density, density + g.size(), g.begin());
  LGPU_HIP(hipSetDevice(s->dev->device));
  LGPU_HIP(hipMemcpy(const_cast<float *>(s->dsc.density) + s->density_off[id], g.data(), g.size() * sizeof(float),
                     hipMemcpyHostToDevice));
  return LGPU_OK;
}

// lambdaCreateShaderNode / Material (Material.cpp:94-195): stored here, compiled at commit.
lgpu_status lgpu_scene_add_material(lgpu_scene s, const lgpu_material_desc *d, uint32_t *id) {
  CHECK_ARG(s && d, "add_material: null argument");
  CHECK_ARG(!s->committed, "add_material: scene already committed");
  CHECK_ARG(d->num_nodes == 0 || d->nodes, "add_material: null nodes");
  CHECK_ARG(d->bxdf_root == LGPU_NONE || d->bxdf_root < d->num_nodes, "add_material: bxdf_root out of range");
  CHECK_ARG(d->emission_root == LGPU_NONE || d->emission_root < d->num_nodes, "add_material: emission_root out of range");
  HostMaterial m;
  m.nodes.assign(d->nodes, d->nodes + d->num_nodes);
  m.desc = *d;
  m.desc.nodes = nullptr;
  // nested Mix BSDFs are not supported by the lobe evaluator (one mix level)
  for (const lgpu_node &n : m.nodes) {
    if (n.type != LGPU_NODE_MIX_BSDF) continue;
    for (int k = 0; k < 2; ++k) {
      CHECK_ARG(n.in[k] < d->num_nodes, "add_material: mix input out of range");
      if (m.nodes[n.in[k]].type == LGPU_NODE_MIX_BSDF) return Fail(LGPU_ERR_UNSUPPORTED, "add_material: nested Mix BSDF");
    }
  }
  s->materials.push_back(std::move(m));
  if (id) *id = (uint32_t)s->materials.size() - 1;
  return LGPU_OK;
}

// Checks a mesh descriptor and copies it (TriangleMesh, lambdaTriangleMeshSetBuffer Lambda.cpp:211-244).
static lgpu_status CopyMesh(lgpu_scene s, const lgpu_mesh_desc *d, const char *what, HostMesh &m) {
  if (!(s && d && d->positions && d->indices)) return Fail(LGPU_ERR_INVALID_ARGUMENT, std::string(what) + ": null argument");
  if (s->committed) return Fail(LGPU_ERR_INVALID_ARGUMENT, std::string(what) + ": scene already committed");
  if (d->material >= s->materials.size()) return Fail(LGPU_ERR_INVALID_ARGUMENT, std::string(what) + ": material id out of range");
  if (!(d->num_triangles > 0 && d->num_vertices > 0)) return Fail(LGPU_ERR_INVALID_ARGUMENT, std::string(what) + ": empty mesh");
  for (size_t i = 0; i < 3ull * d->num_triangles; ++i)
    if (d->indices[i] >= d->num_vertices) return Fail(LGPU_ERR_INVALID_ARGUMENT, std::string(what) + ": index out of range");
  m.pos.assign(d->positions, d->positions + 3ull * d->num_vertices);
  m.idx.assign(d->indices, d->indices + 3ull * d->num_triangles);
  if (d->normals && d->tangents) {
    m.nrm.assign(d->normals, d->normals + 3ull * d->num_vertices);
    m.tan.assign(d->tangents, d->tangents + 4ull * d->num_vertices);
  }
  if (d->uvs) m.uv.assign(d->uvs, d->uvs + 2ull * d->num_vertices);
  m.material = d->material;
  return LGPU_OK;
}

// Scene::AddObject (Scene.cpp:114-121); an emissive material also adds its MeshLight.
static uint32_t AddObject(lgpu_scene s, HostMesh &&m) {
  const uint32_t material = m.material;
  const int mi = (int)s->meshes.size();
  s->meshes.push_back(std::move(m));
  if (s->materials[material].desc.emission_root != LGPU_NONE) {  // MeshLight(mesh) + AddLight
    HostLight L;
    std::memset(&L.desc, 0, sizeof(L.desc));
    L.mesh = mi;
    L.desc.intensity = s->materials[material].desc.emission_intensity;
    s->lights.push_back(std::move(L));
  }
  return (uint32_t)mi;
}

// Scene::AddObject(TriangleMesh) (Scene.cpp:114-121); lambdaTriangleMeshSetBuffer (Lambda.cpp:211-244)
lgpu_status lgpu_scene_add_mesh(lgpu_scene s, const lgpu_mesh_desc *d, uint32_t *object_id) {
  HostMesh m;
  const lgpu_status st = CopyMesh(s, d, "add_mesh", m);
  if (st != LGPU_OK) return st;
  const uint32_t id = AddObject(s, std::move(m));
  if (object_id) *object_id = id;
  return LGPU_OK;
}

// InstanceProxy(object) (Instance.cpp:7-16)
lgpu_status lgpu_scene_add_proxy(lgpu_scene s, const lgpu_mesh_desc *d, uint32_t *proxy_id) {
  HostMesh m;
  const lgpu_status st = CopyMesh(s, d, "add_proxy", m);
  if (st != LGPU_OK) return st;
  s->proxies.push_back(std::move(m));
  if (proxy_id) *proxy_id = (uint32_t)s->proxies.size() - 1;
  return LGPU_OK;
}

// Instance::Commit (Instance.cpp:22-34) flattened: the proxy's vertices through the affine
// transform (affine3::operator*, affine3.h:38-40), normals through the inverse transpose of its
// linear part, tangents through the linear part (bitangent sign flipped by a mirroring transform).
lgpu_status lgpu_scene_add_instance(lgpu_scene s, uint32_t proxy_id, const float x[12], uint32_t material,
                                    uint32_t *object_id) {
  CHECK_ARG(s && x, "add_instance: null argument");
  CHECK_ARG(!s->committed, "add_instance: scene already committed");
  CHECK_ARG(proxy_id < s->proxies.size(), "add_instance: proxy id out of range");
  CHECK_ARG(material == LGPU_NONE || material < s->materials.size(), "add_instance: material id out of range");
  for (int k = 0; k < 12; ++k) CHECK_ARG(std::isfinite(x[k]), "add_instance: transform is not finite");
  // cofactor matrix of the linear part M (columns x[0..2], x[3..5], x[6..8]): M^-T = cof / det
  const float c00 = x[4] * x[8] - x[5] * x[7], c01 = x[5] * x[6] - x[3] * x[8], c02 = x[3] * x[7] - x[4] * x[6];
  const float c10 = x[7] * x[2] - x[8] * x[1], c11 = x[8] * x[0] - x[6] * x[2], c12 = x[6] * x[1] - x[7] * x[0];
  const float c20 = x[1] * x[5] - x[2] * x[4], c21 = x[2] * x[3] - x[0] * x[5], c22 = x[0] * x[4] - x[1] * x[3];
  const float det = x[0] * c00 + x[1] * c01 + x[2] * c02;
  CHECK_ARG(det != 0.f && std::isfinite(det), "add_instance: singular transform");
  const float sg = det < 0.f ? -1.f : 1.f;
  const HostMesh &p = s->proxies[proxy_id];
  HostMesh m = p;
  if (material != LGPU_NONE) m.material = material;
  const size_t nv = p.pos.size() / 3;
  for (size_t v = 0; v < nv; ++v) {
    const float px = p.pos[3 * v], py = p.pos[3 * v + 1], pz = p.pos[3 * v + 2];
    m.pos[3 * v] = x[0] * px + x[3] * py + x[6] * pz + x[9];
    m.pos[3 * v + 1] = x[1] * px + x[4] * py + x[7] * pz + x[10];
    m.pos[3 * v + 2] = x[2] * px + x[5] * py + x[8] * pz + x[11];
  }
  if (!p.nrm.empty()) {
    for (size_t v = 0; v < nv; ++v) {
      const float nx = p.nrm[3 * v], ny = p.nrm[3 * v + 1], nz = p.nrm[3 * v + 2];
      // column j of cof holds the cofactors of column j of M: n' = cof * n
      float ox = (c00 * nx + c10 * ny + c20 * nz) * sg, oy = (c01 * nx + c11 * ny + c21 * nz) * sg,
            oz = (c02 * nx + c12 * ny + c22 * nz) * sg;
      float l = std::sqrt(ox * ox + oy * oy + oz * oz);
      if (l > 0.f) { ox /= l; oy /= l; oz /= l; }
      m.nrm[3 * v] = ox; m.nrm[3 * v + 1] = oy; m.nrm[3 * v + 2] = oz;
      const float tx = p.tan[4 * v], ty = p.tan[4 * v + 1], tz = p.tan[4 * v + 2];
      float ux = x[0] * tx + x[3] * ty + x[6] * tz, uy = x[1] * tx + x[4] * ty + x[7] * tz,
            uz = x[2] * tx + x[5] * ty + x[8] * tz;
      l = std::sqrt(ux * ux + uy * uy + uz * uz);
      if (l > 0.f) { ux /= l; uy /= l; uz /= l; }
      m.tan[4 * v] = ux; m.tan[4 * v + 1] = uy; m.tan[4 * v + 2] = uz; m.tan[4 * v + 3] = p.tan[4 * v + 3] * sg;
    }
  }
  const uint32_t id = AddObject(s, std::move(m));
  if (object_id) *object_id = id;
  return LGPU_OK;
}

// Scene::AddLight (Scene.cpp:109-112)
lgpu_status lgpu_scene_add_light(lgpu_scene s, const lgpu_light_desc *d, uint32_t *id) {
  CHECK_ARG(s && d, "add_light: null argument");
  CHECK_ARG(!s->committed, "add_light: scene already committed");
  CHECK_ARG(d->type >= LGPU_LIGHT_POINT && d->type <= LGPU_LIGHT_ENVIRONMENT, "add_light: bad light type");
  HostLight L;
  L.desc = *d;
  if (d->type == LGPU_LIGHT_ENVIRONMENT) {
    CHECK_ARG(d->env_rgb && d->env_width && d->env_height, "add_light: environment needs an image");
    for (const HostLight &o : s->lights)
      if (o.mesh < 0 && o.desc.type == LGPU_LIGHT_ENVIRONMENT)
        return Fail(LGPU_ERR_UNSUPPORTED, "add_light: one environment light per scene");
    L.env.assign(d->env_rgb, d->env_rgb + 3ull * d->env_width * d->env_height);
  }
  L.desc.env_rgb = nullptr;
  s->lights.push_back(std::move(L));
  if (id) *id = (uint32_t)s->lights.size() - 1;
  return LGPU_OK;
}

// Scene::Commit + LightSampler::Commit (Scene.cpp:18-26)
lgpu_status lgpu_scene_commit(lgpu_scene s, uint32_t strategy) {
  CHECK_ARG(s, "scene_commit: null scene");
  CHECK_ARG(!s->committed, "scene_commit: already committed");
  CHECK_ARG(strategy == LGPU_LIGHTS_POWER || strategy == LGPU_LIGHTS_TREE, "scene_commit: bad light strategy");
  CHECK_ARG(!s->meshes.empty(), "scene_commit: scene has no geometry");
  return CommitScene(s, strategy);
}

lgpu_status lgpu_scene_info(lgpu_scene s, uint32_t *nt, uint32_t *nb, uint32_t *nl, uint32_t *ntn) {
  CHECK_ARG(s, "scene_info: null scene");
  CHECK_ARG(s->committed, "scene_info: scene not committed");
  if (nt) *nt = s->num_tris;
  if (nb) *nb = s->num_bvh_nodes;
  if (nl) *nl = (uint32_t)s->lights.size();
  if (ntn) *ntn = (uint32_t)s->tree.size();
  return LGPU_OK;
}

lgpu_status lgpu_scene_light_tree(lgpu_scene s, float *out, uint32_t max_nodes) {
  CHECK_ARG(s && out, "scene_light_tree: null argument");
  CHECK_ARG(s->committed, "scene_light_tree: scene not committed");
  for (size_t i = 0; i < s->tree.size() && i < max_nodes; ++i) {
    const TreeNode &t = s->tree[i];
    float *o = out + 15 * i;
    o[0] = t.center[0]; o[1] = t.center[1]; o[2] = t.center[2]; o[3] = t.diag;
    o[4] = t.axis[0]; o[5] = t.axis[1]; o[6] = t.axis[2];
    o[7] = t.thetaO; o[8] = t.thetaE; o[9] = t.power;
    o[10] = t.child[0] == kNone ? -1.f : (float)t.child[0];
    o[11] = t.child[1] == kNone ? -1.f : (float)t.child[1];
    o[12] = (float)t.first; o[13] = (float)t.count;
    o[14] = t.parent == kNone ? -1.f : (float)t.parent;
  }
  return LGPU_OK;
}

// Camera (Camera.cpp:7-35)
void lgpu_camera_init(lgpu_camera *c, uint32_t type, const float origin[3], float width, float height) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->type = type;
  if (origin) for (int k = 0; k < 3; ++k) c->origin[k] = origin[k];
  c->aspect = height / width;
  lgpu_camera_set_fov(c, 1.1f);
  lgpu_camera_set_rotation(c, 0.f, 0.f);
}
void lgpu_camera_set_rotation(lgpu_camera *c, float phi, float theta) {
  if (!c) return;
  V3 z = v3(std::cos(theta) * std::sin(phi), std::sin(theta), std::cos(theta) * std::cos(phi));
  z = Norm(z);
  const V3 x = Cross(z, v3(0, 1, 0));
  const V3 y = Cross(x, z);
  c->z_hat[0] = z.x; c->z_hat[1] = z.y; c->z_hat[2] = z.z;
  c->x_hat[0] = x.x; c->x_hat[1] = x.y; c->x_hat[2] = x.z;
  c->y_hat[0] = y.x; c->y_hat[1] = y.y; c->y_hat[2] = y.z;
}
void lgpu_camera_set_forwards(lgpu_camera *c, const float f[3]) {
  if (!c || !f) return;
  const V3 z = Norm(v3(f[0], f[1], f[2]));
  const V3 x = Cross(z, v3(0, 1, 0));
  const V3 y = Cross(z, x);  // Camera::SetForwards (Camera.cpp:26-30) as written
  c->z_hat[0] = z.x; c->z_hat[1] = z.y; c->z_hat[2] = z.z;
  c->x_hat[0] = x.x; c->x_hat[1] = x.y; c->x_hat[2] = x.z;
  c->y_hat[0] = y.x; c->y_hat[1] = y.y; c->y_hat[2] = y.z;
}
void lgpu_camera_set_fov(lgpu_camera *c, float fov) {
  if (!c) return;
  c->tan_fov = (float)std::tan(.5 * (double)fov);
}

// Film (Film.h:19-45, Film.cpp:6-30; lambdaCreateFilm Lambda.h:121)
lgpu_status lgpu_film_create(lgpu_device dev, uint32_t w, uint32_t h, lgpu_film *out) {
  CHECK_ARG(dev && out && w && h, "film_create: bad argument");
  LGPU_HIP(hipSetDevice(dev->device));
  lgpu_film f = new lgpu_film_s();
  f->dev = dev;
  f->width = w;
  f->height = h;
  if (f->sum.alloc(24ull * w * h) != hipSuccess || f->count.alloc(4ull * w * h) != hipSuccess) {
    delete f;
    return Fail(LGPU_ERR_OUT_OF_MEMORY, "film_create: allocation failed");
  }
  *out = f;
  return lgpu_film_clear(f);
}
lgpu_status lgpu_film_destroy(lgpu_film f) {
  if (!f) return LGPU_OK;
  (void)hipSetDevice(f->dev->device);
  delete f;
  return LGPU_OK;
}
lgpu_status lgpu_film_clear(lgpu_film f) {
  CHECK_ARG(f, "film_clear: null film");
  std::lock_guard<std::mutex> lk(f->dev->mu);
  LGPU_HIP(hipSetDevice(f->dev->device));
  LGPU_HIP(hipMemsetAsync(f->sum.ptr, 0, f->sum.bytes, f->dev->stream));
  LGPU_HIP(hipMemsetAsync(f->count.ptr, 0, f->count.bytes, f->dev->stream));
  LGPU_HIP(hipStreamSynchronize(f->dev->stream));
  return LGPU_OK;
}
lgpu_status lgpu_film_read(lgpu_film f, double *rgb, uint32_t *counts) {
  CHECK_ARG(f, "film_read: null film");
  std::lock_guard<std::mutex> lk(f->dev->mu);
  LGPU_HIP(hipSetDevice(f->dev->device));
  LGPU_HIP(hipStreamSynchronize(f->dev->stream));
  if (rgb) LGPU_HIP(hipMemcpy(rgb, f->sum.ptr, 24ull * f->width * f->height, hipMemcpyDeviceToHost));
  if (counts) LGPU_HIP(hipMemcpy(counts, f->count.ptr, 4ull * f->width * f->height, hipMemcpyDeviceToHost));
  return LGPU_OK;
}
// The inverse of lgpu_film_read: the film's sums and counts from host arrays (a saved film to
// resume, or an image to post-process or denoise, as sums with counts of 1). NULL counts set 1.
lgpu_status lgpu_film_write(lgpu_film f, const double *rgb, const uint32_t *counts) {
  CHECK_ARG(f && rgb, "film_write: null argument");
  std::lock_guard<std::mutex> lk(f->dev->mu);
  LGPU_HIP(hipSetDevice(f->dev->device));
  LGPU_HIP(hipStreamSynchronize(f->dev->stream));
  const size_t n = (size_t)f->width * f->height;
  LGPU_HIP(hipMemcpy(f->sum.ptr, rgb, 24ull * n, hipMemcpyHostToDevice));
  if (counts) {
    LGPU_HIP(hipMemcpy(f->count.ptr, counts, 4ull * n, hipMemcpyHostToDevice));
  } else {
    std::vector<uint32_t> ones(n, 1u);
    LGPU_HIP(hipMemcpy(f->count.ptr, ones.data(), 4ull * n, hipMemcpyHostToDevice));
  }
  return LGPU_OK;
}
// Film::ToRGBTexture (Film.cpp:11-24): mean per pixel; pixels with no samples read 0.
lgpu_status lgpu_film_read_rgba(lgpu_film f, float *rgba) {
  CHECK_ARG(f && rgba, "film_read_rgba: null argument");
  std::lock_guard<std::mutex> lk(f->dev->mu);
  return FilmToRgba(f, rgba);
}

}  // extern "C"
