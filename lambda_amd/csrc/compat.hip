// The reference's lambda* C API (include/lambda/Lambda.h:37-202, src/api/Lambda.cpp) over the
// lgpu_* entry points: objects keep the caller's description until lambdaCommitScene builds an
// lgpu scene from the attached instances; a render directive becomes an lgpu_render_desc and the
// progressive renderer an lgpu_progressive. Host code only.
#include <cmath>
#include <cstring>
#include <set>
#include <string>
#include <vector>

#include "../../include/lambda_compat.h"
#include "../../include/lambda_gpu.h"

struct LAMBDA_TriangleMesh {
  LAMBDA_Device *dev = nullptr;
  size_t nv = 0, nt = 0;
  const float *vertices = nullptr, *normals = nullptr, *tangents = nullptr, *uvs = nullptr;  // 4, 4, 4, 2 floats
  const unsigned *triangles = nullptr;
};
struct LAMBDA_Proxy {
  LAMBDA_Device *dev = nullptr;
  LAMBDA_TriangleMesh *mesh = nullptr;
};
struct LAMBDA_Instance {
  LAMBDA_Device *dev = nullptr;
  LAMBDA_Proxy *proxy = nullptr;
  float xfm[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
};
struct LAMBDA_Scene {
  LAMBDA_Device *dev = nullptr;
  std::vector<LAMBDA_Instance *> objects;
  lgpu_scene scene = nullptr;
  uint32_t strategy = LGPU_LIGHTS_TREE;
  bool committed = false;
};
struct LAMBDA_Film {
  LAMBDA_Device *dev = nullptr;
  lgpu_film film = nullptr;
  int width = 0, height = 0;
};
struct LAMBDA_Camera {
  LAMBDA_Device *dev = nullptr;
  LAMBDA_CameraType type = LAMBDA_CAMERA_THIN_LENS;
  lgpu_camera cam{};
};
struct LAMBDA_RenderDirective {
  LAMBDA_Device *dev = nullptr;
  LAMBDA_Scene *scene = nullptr;
  LAMBDA_Film *film = nullptr;
  LAMBDA_Camera *camera = nullptr;
  LAMBDA_RenderProperties props{};
};
struct LAMBDA_ProgressiveRenderer {
  LAMBDA_Device *dev = nullptr;
  lgpu_progressive p = nullptr;
  void (*callback)() = nullptr;
  std::vector<float> rgba;
  int width = 0, height = 0;
};
struct LAMBDA_Device {
  lgpu_device dev = nullptr;
  // objects created on the device, freed with it (each release removes its entry)
  std::set<LAMBDA_Scene *> scenes;
  std::set<LAMBDA_TriangleMesh *> meshes;
  std::set<LAMBDA_Proxy *> proxies;
  std::set<LAMBDA_Instance *> instances;
  std::set<LAMBDA_Film *> films;
  std::set<LAMBDA_Camera *> cameras;
  std::set<LAMBDA_RenderDirective *> directives;
  std::set<LAMBDA_ProgressiveRenderer *> renderers;
  std::set<LAMBDA_RenderProperties *> properties;
  std::set<std::pair<std::string, int>> names;  // the resource map's keys (Lambda.cpp:22-63)
};

namespace {

template <class T>
void Drop(std::set<T *> &owner, T *p) {
  owner.erase(p);
  delete p;
}

void FreeScene(LAMBDA_Scene *s) {
  if (s->scene) lgpu_scene_destroy(s->scene);
  s->scene = nullptr;
}

// Builds the lgpu scene: the default material (white Lambertian, Lambda.cpp:136-145), each
// attached instance's proxy mesh once, the instances with their matrices.
lgpu_status Build(LAMBDA_Scene *s, uint32_t strategy) {
  FreeScene(s);
  lgpu_status st = lgpu_scene_create(s->dev->dev, &s->scene);
  if (st != LGPU_OK) return st;
  lgpu_node nodes[2];
  std::memset(nodes, 0, sizeof(nodes));
  nodes[0].type = LGPU_NODE_RGB;
  for (int k = 0; k < 4; ++k) nodes[0].in[k] = nodes[1].in[k] = LGPU_NONE;
  nodes[0].p[0] = nodes[0].p[1] = nodes[0].p[2] = 1.f;
  nodes[1].type = LGPU_NODE_LAMBERTIAN;
  nodes[1].in[0] = 0;
  lgpu_material_desc md{2, nodes, 1, LGPU_NONE, 0.f, LGPU_NONE, LGPU_NONE};
  uint32_t mat;
  if ((st = lgpu_scene_add_material(s->scene, &md, &mat)) != LGPU_OK) return st;
  std::vector<std::pair<LAMBDA_Proxy *, uint32_t>> proxyIds;
  for (LAMBDA_Instance *inst : s->objects) {
    LAMBDA_Proxy *px = inst->proxy;
    if (!px || !px->mesh || !px->mesh->vertices || !px->mesh->triangles || !px->mesh->nv || !px->mesh->nt) continue;
    uint32_t pid = LGPU_NONE;
    for (auto &e : proxyIds)
      if (e.first == px) pid = e.second;
    if (pid == LGPU_NONE) {
      const LAMBDA_TriangleMesh &m = *px->mesh;
      std::vector<float> pos(3 * m.nv), nrm, tan, uv;
      for (size_t v = 0; v < m.nv; ++v)
        for (int a = 0; a < 3; ++a) pos[3 * v + a] = m.vertices[4 * v + a];
      if (m.normals) {
        nrm.resize(3 * m.nv);
        tan.resize(4 * m.nv);
        for (size_t v = 0; v < m.nv; ++v) {
          const float n[3] = {m.normals[4 * v], m.normals[4 * v + 1], m.normals[4 * v + 2]};
          for (int a = 0; a < 3; ++a) nrm[3 * v + a] = n[a];
          float t[3];
          if (m.tangents) {
            for (int a = 0; a < 3; ++a) t[a] = m.tangents[4 * v + a];
          } else {  // any unit vector orthogonal to the normal
            const float h[3] = {std::fabs(n[0]) < .9f ? 1.f : 0.f, std::fabs(n[0]) < .9f ? 0.f : 1.f, 0.f};
            t[0] = h[1] * n[2] - h[2] * n[1]; t[1] = h[2] * n[0] - h[0] * n[2]; t[2] = h[0] * n[1] - h[1] * n[0];
            const float l = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
            for (float &c : t) c = l > 0 ? c / l : 0.f;
          }
          for (int a = 0; a < 3; ++a) tan[4 * v + a] = t[a];
          tan[4 * v + 3] = 1.f;
        }
      }
      if (m.uvs) uv.assign(m.uvs, m.uvs + 2 * m.nv);
      lgpu_mesh_desc d{(uint32_t)m.nv, (uint32_t)m.nt, pos.data(), (const uint32_t *)m.triangles,
                       nrm.empty() ? nullptr : nrm.data(), tan.empty() ? nullptr : tan.data(),
                       uv.empty() ? nullptr : uv.data(), mat};
      if ((st = lgpu_scene_add_proxy(s->scene, &d, &pid)) != LGPU_OK) return st;
      proxyIds.emplace_back(px, pid);
    }
    uint32_t obj;
    if ((st = lgpu_scene_add_instance(s->scene, pid, inst->xfm, LGPU_NONE, &obj)) != LGPU_OK) return st;
  }
  if ((st = lgpu_scene_commit(s->scene, strategy)) != LGPU_OK) return st;
  s->strategy = strategy;
  s->committed = true;
  return LGPU_OK;
}

// affine3::operator* (affine3.h:54-69): 3x3 parts multiplied as written there, translations added.
void AffineMul(const float a[12], const float b[12], float out[12]) {
  float r[12];
  for (int row = 0; row < 3; ++row)
    for (int col = 0; col < 3; ++col)
      r[3 * row + col] = a[3 * row] * b[col] + a[3 * row + 1] * b[3 + col] + a[3 * row + 2] * b[6 + col];
  for (int k = 0; k < 3; ++k) r[9 + k] = a[9 + k] + b[9 + k];
  std::memcpy(out, r, sizeof(r));
}
// affine3::GetRotationX/Y/Z (affine3.h:75-89): cos taken as sqrt(1 - sin^2), as there.
void Rotation(int axis, float theta, float out[12]) {
  const float s = std::sin(theta), c = std::sqrt(1.f - s * s);
  const float x[9] = {1, 0, 0, 0, c, -s, 0, s, c};
  const float y[9] = {c, 0, s, 0, 1, 0, -s, 0, c};
  const float z[9] = {c, -s, 0, s, c, 0, 0, 0, 1};
  std::memcpy(out, axis == 0 ? x : (axis == 1 ? y : z), sizeof(x));
  out[9] = out[10] = out[11] = 0.f;
}
void EulerMatrix(const float e[3], float out[12]) {
  float rx[12], ry[12], rz[12], t[12];
  Rotation(0, e[0], rx);
  Rotation(1, e[1], ry);
  Rotation(2, e[2], rz);
  AffineMul(rx, ry, t);
  AffineMul(t, rz, out);
}

lgpu_integrator Integrator(LAMBDA_Integrator i) {
  switch (i) {
    case LAMBDA_INTEGRATOR_VOLPATH: return LGPU_INTEGRATOR_VOLPATH;
    case LAMBDA_INTEGRATOR_DIRECT: return LGPU_INTEGRATOR_DIRECT;
    case LAMBDA_INTEGRATOR_NORMAL: return LGPU_INTEGRATOR_NORMAL;
    case LAMBDA_INTEGRATOR_DEPTH: return LGPU_INTEGRATOR_DEPTH;
    default: return LGPU_INTEGRATOR_PATH;  // AOV included (Lambda.cpp:369-371)
  }
}

void Trampoline(void *user) {
  LAMBDA_ProgressiveRenderer *r = static_cast<LAMBDA_ProgressiveRenderer *>(user);
  if (r->callback) r->callback();
}

}  // namespace

extern "C" {

LAMBDA_Device *lambdaCreateDevice(void) {
  lgpu_device d = nullptr;
  if (lgpu_device_create(0, &d) != LGPU_OK) return nullptr;
  LAMBDA_Device *dev = new LAMBDA_Device();
  dev->dev = d;
  return dev;
}

void lambdaReleaseDevice(LAMBDA_Device *_device) {
  if (!_device) return;
  for (auto *r : _device->renderers) { if (r->p) lgpu_progressive_destroy(r->p); delete r; }
  for (auto *x : _device->directives) delete x;
  for (auto *x : _device->properties) delete x;
  for (auto *s : _device->scenes) { FreeScene(s); delete s; }
  for (auto *f : _device->films) { if (f->film) lgpu_film_destroy(f->film); delete f; }
  for (auto *x : _device->cameras) delete x;
  for (auto *x : _device->instances) delete x;
  for (auto *x : _device->proxies) delete x;
  for (auto *x : _device->meshes) delete x;
  lgpu_device_destroy(_device->dev);
  delete _device;
}

void lambdaReleaseResource(LAMBDA_Device *_device, char *_name, LAMBDA_Type _type) {
  if (_device && _name) _device->names.erase({std::string(_name), (int)_type});
}

LAMBDA_Scene *lambdaCreateScene(LAMBDA_Device *_device, char *_name) {
  LAMBDA_Scene *s = new LAMBDA_Scene();
  s->dev = _device;
  _device->scenes.insert(s);
  if (_name) _device->names.insert({std::string(_name), (int)LAMBDA_SCENE});
  return s;
}

void lambdaCommitScene(LAMBDA_Scene *_scene) {
  if (_scene) (void)Build(_scene, _scene->strategy);
}

void lambdaReleaseScene(LAMBDA_Scene *_scene) {
  if (!_scene) return;
  FreeScene(_scene);
  Drop(_scene->dev->scenes, _scene);
}

void lambdaAttachObject(LAMBDA_Scene *_scene, LAMBDA_Instance *_instance) {
  if (_scene && _instance) _scene->objects.push_back(_instance);
}

void lambdaDetachObject(LAMBDA_Scene *_scene, LAMBDA_Instance *_instance) {
  if (!_scene) return;
  auto &o = _scene->objects;
  for (size_t k = 0; k < o.size(); ++k)
    if (o[k] == _instance) { o.erase(o.begin() + (long)k); break; }
}

LAMBDA_TriangleMesh *lambdaCreateTriangleMesh(LAMBDA_Device *_device, char *_name) {
  LAMBDA_TriangleMesh *m = new LAMBDA_TriangleMesh();
  m->dev = _device;
  _device->meshes.insert(m);
  if (_name) _device->names.insert({std::string(_name), (int)LAMBDA_TRIANGLE_MESH});
  return m;
}

void lambdaReleaseTriangleMesh(LAMBDA_TriangleMesh *_mesh) {
  if (_mesh) Drop(_mesh->dev->meshes, _mesh);
}

void lambdaTriangleMeshAllocData(LAMBDA_TriangleMesh *_mesh, size_t _numVertices, size_t numTriangles) {
  if (!_mesh) return;
  _mesh->nv = _numVertices;
  _mesh->nt = numTriangles;
}

void lambdaTriangleMeshSetBuffer(LAMBDA_TriangleMesh *_mesh, LAMBDA_Buffer _bufferType, void *_ptr, size_t _num) {
  if (!_mesh) return;
  const size_t want = _bufferType == LAMBDA_BUFFER_TRIANGLE ? _mesh->nt : _mesh->nv;
  if (_num != want) return;  // the reference throws on a length mismatch (Lambda.cpp:208-240)
  switch (_bufferType) {
    case LAMBDA_BUFFER_VERTEX: _mesh->vertices = (const float *)_ptr; break;
    case LAMBDA_BUFFER_TRIANGLE: _mesh->triangles = (const unsigned *)_ptr; break;
    case LAMBDA_BUFFER_VERTEX_NORMAL: _mesh->normals = (const float *)_ptr; break;
    case LAMBDA_BUFFER_VERTEX_TANGENT: _mesh->tangents = (const float *)_ptr; break;
    case LAMBDA_BUFFER_TEXTURE_COORDINATE: _mesh->uvs = (const float *)_ptr; break;
  }
}

LAMBDA_Proxy *lambdaCreateProxy(LAMBDA_Device *_device, LAMBDA_TriangleMesh *_mesh) {
  LAMBDA_Proxy *p = new LAMBDA_Proxy();
  p->dev = _device;
  p->mesh = _mesh;
  _device->proxies.insert(p);
  return p;
}

void lambdaReleaseProxy(LAMBDA_Proxy *_proxy) {
  if (_proxy) Drop(_proxy->dev->proxies, _proxy);
}

LAMBDA_Instance *lambdaCreateInstance(LAMBDA_Device *_device, LAMBDA_Proxy *_proxy) {
  LAMBDA_Instance *i = new LAMBDA_Instance();
  i->dev = _device;
  i->proxy = _proxy;
  _device->instances.insert(i);
  return i;
}

void lambdaReleaseInstance(LAMBDA_Instance *_instance) {
  if (_instance) Drop(_instance->dev->instances, _instance);
}

void lambdaMakeAffineTransform(float _xfm[12], float _position[3], float _scale[3], float _eulerAngles[3]) {
  float x[12] = {_scale[0], 0, 0, 0, _scale[1], 0, 0, 0, _scale[2], _position[0], _position[1], _position[2]};
  float r[12];
  EulerMatrix(_eulerAngles, r);
  AffineMul(x, r, _xfm);
}

void lambdaSetTransform(LAMBDA_Instance *_instance, float *_xfm) {
  if (_instance && _xfm) std::memcpy(_instance->xfm, _xfm, sizeof(_instance->xfm));
}

void lambdaSetPosition(LAMBDA_Instance *_instance, float _position[3]) {
  for (int k = 0; k < 3; ++k) _instance->xfm[9 + k] = _position[k];
}

void lambdaSetScale(LAMBDA_Instance *_instance, float _scale[3]) {
  _instance->xfm[0] = _scale[0];
  _instance->xfm[4] = _scale[1];
  _instance->xfm[8] = _scale[2];
}

void lambdaSetEulerAngles(LAMBDA_Instance *_instance, float _eulerAngles[3]) {
  EulerMatrix(_eulerAngles, _instance->xfm);
}

LAMBDA_Film *lambdaCreateFilm(LAMBDA_Device *_device, int _width, int _height) {
  if (_width <= 0 || _height <= 0) return nullptr;
  LAMBDA_Film *f = new LAMBDA_Film();
  f->dev = _device;
  f->width = _width;
  f->height = _height;
  if (lgpu_film_create(_device->dev, (uint32_t)_width, (uint32_t)_height, &f->film) != LGPU_OK) {
    delete f;
    return nullptr;
  }
  _device->films.insert(f);
  return f;
}

void lambdaReleaseFilm(LAMBDA_Film *_film) {
  if (!_film) return;
  if (_film->film) lgpu_film_destroy(_film->film);
  Drop(_film->dev->films, _film);
}

LAMBDA_Camera *lambdaCreateCamera(LAMBDA_Device *_device, LAMBDA_CameraType _cameraType, float _pos[3], float _phi,
                                  float _theta) {
  LAMBDA_Camera *c = new LAMBDA_Camera();
  c->dev = _device;
  c->type = _cameraType;
  const uint32_t t = _cameraType == LAMBDA_CAMERA_SPHERICAL ? LGPU_CAMERA_SPHERICAL : LGPU_CAMERA_THIN_LENS;
  lgpu_camera_init(&c->cam, t, _pos, 1.f, 1.f);
  lgpu_camera_set_rotation(&c->cam, _phi, _theta);
  c->cam.focal_length = 1.f;
  c->cam.aperture = LGPU_APERTURE_CIRCULAR;
  c->cam.aperture_size = 0.f;
  _device->cameras.insert(c);
  return c;
}

void lambdaReleaseCamera(LAMBDA_Camera *_camera) {
  if (_camera) Drop(_camera->dev->cameras, _camera);
}

void lambdaSetCameraPosition(LAMBDA_Camera *_camera, float _pos[3]) {
  for (int k = 0; k < 3; ++k) _camera->cam.origin[k] = _pos[k];
}

void lambdaSetCameraAngles(LAMBDA_Camera *_camera, float _phi, float _theta) {
  lgpu_camera_set_rotation(&_camera->cam, _phi, _theta);
}

void lambdaSetCameraAxes(LAMBDA_Camera *_camera, float _xHat[3], float _yHat[3], float _zHat[3]) {
  for (int k = 0; k < 3; ++k) {
    _camera->cam.x_hat[k] = _xHat[k];
    _camera->cam.y_hat[k] = _yHat[k];
    _camera->cam.z_hat[k] = _zHat[k];
  }
}

void lambdaSetCameraFocalLength(LAMBDA_Camera *_camera, float _focalLength) {
  if (_camera->type == LAMBDA_CAMERA_THIN_LENS) _camera->cam.focal_length = _focalLength;
}

void lambdaSetCameraApertureSize(LAMBDA_Camera *_camera, float _size) { _camera->cam.aperture_size = _size; }

LAMBDA_RenderProperties *lambdaCreateRenderProperties(void) {
  // no device argument (Lambda.h:171): freed with the process, or with the device a directive
  // made from it belongs to
  LAMBDA_RenderProperties *p = new LAMBDA_RenderProperties();
  p->spp = 1;
  p->tileSizeX = p->tileSizeY = 16;
  p->numThreads = 0;
  p->integrator = LAMBDA_INTEGRATOR_PATH;
  p->lightStrategy = LAMBDA_LIGHT_STRATEGY_TREE;
  return p;
}

LAMBDA_RenderDirective *lambdaCreateRenderDirective(LAMBDA_Device *_device, LAMBDA_Scene *_scene, LAMBDA_Film *_film,
                                                    LAMBDA_Camera *_camera, LAMBDA_RenderProperties *_properties) {
  LAMBDA_RenderDirective *d = new LAMBDA_RenderDirective();
  d->dev = _device;
  d->scene = _scene;
  d->film = _film;
  d->camera = _camera;
  if (_properties) {
    d->props = *_properties;
    if (_device->properties.count(_properties) == 0) _device->properties.insert(_properties);
  } else {
    LAMBDA_RenderProperties *p = lambdaCreateRenderProperties();
    d->props = *p;
    delete p;
  }
  _device->directives.insert(d);
  return d;
}

void lambdaReleaseRenderDirective(LAMBDA_RenderDirective *_directive) {
  if (_directive) Drop(_directive->dev->directives, _directive);
}

void lambdaSetCamera(LAMBDA_RenderDirective *_directive, LAMBDA_Camera *_camera) { _directive->camera = _camera; }

void lambdaSetScene(LAMBDA_RenderDirective *_directive, LAMBDA_Scene *_scene) { _directive->scene = _scene; }

LAMBDA_ProgressiveRenderer *lambdaCreateProgressiveRenderer(LAMBDA_Device *_device, LAMBDA_RenderDirective *_directive) {
  if (!_directive || !_directive->scene || !_directive->film || !_directive->camera) return nullptr;
  LAMBDA_Scene *s = _directive->scene;
  const uint32_t strategy =
      _directive->props.lightStrategy == LAMBDA_LIGHT_STRATEGY_POWER ? LGPU_LIGHTS_POWER : LGPU_LIGHTS_TREE;
  if (!s->committed || s->strategy != strategy) {  // the light sampler is the directive's (Lambda.cpp:379-392)
    if (Build(s, strategy) != LGPU_OK) return nullptr;
  }
  lgpu_render_desc d;
  std::memset(&d, 0, sizeof(d));
  d.scene = s->scene;
  d.film = _directive->film->film;
  d.camera = _directive->camera->cam;
  d.camera.aspect = 1.f;  // cameras are made with unit width and height (Lambda.cpp:302-307)
  d.integrator = Integrator(_directive->props.integrator);
  d.spp_begin = 0;
  d.spp_end = _directive->props.spp ? _directive->props.spp : 1;
  const bool direct = d.integrator == LGPU_INTEGRATOR_DIRECT;
  d.max_bounces = direct ? 4 : 128;  // DirectLightingIntegrator.h:8, PathIntegrator.h:10
  d.min_bounces = 3;
  d.spectral_mode = LGPU_SPECTRAL_RGB;       // the reference's RGBSpectrum
  d.estimator_mode = LGPU_ESTIMATOR_FAITHFUL;  // its estimator as written
  d.pass_max_depth = 100.f;                  // DepthPass(sampler, 100) (Lambda.cpp:365)
  d.pass_power = 1.f;
  d.pass_background[0] = 1.f; d.pass_background[2] = 1.f;
  LAMBDA_ProgressiveRenderer *r = new LAMBDA_ProgressiveRenderer();
  r->dev = _device;
  r->width = _directive->film->width;
  r->height = _directive->film->height;
  if (lgpu_progressive_create(&d, &r->p) != LGPU_OK) {
    delete r;
    return nullptr;
  }
  lgpu_progressive_set_callback(r->p, Trampoline, r);
  _device->renderers.insert(r);
  return r;
}

void lambdaReleaseProgressiveRenderer(LAMBDA_ProgressiveRenderer *_renderer) {
  if (!_renderer) return;
  if (_renderer->p) lgpu_progressive_destroy(_renderer->p);
  Drop(_renderer->dev->renderers, _renderer);
}

void lambdaSetProgressiveRendererCallback(LAMBDA_ProgressiveRenderer *_renderer, void (*_callback)()) {
  if (_renderer) _renderer->callback = _callback;
}

void *lambdaGetProgressiveRendererData(LAMBDA_ProgressiveRenderer *_renderer, int *_width, int *_height) {
  if (!_renderer) return nullptr;
  _renderer->rgba.resize((size_t)4 * _renderer->width * _renderer->height);
  lgpu_progressive_read(_renderer->p, _renderer->rgba.data(), nullptr);
  if (_width) *_width = _renderer->width;
  if (_height) *_height = _renderer->height;
  return _renderer->rgba.data();
}

void lambdaStartProgressiveRenderer(LAMBDA_ProgressiveRenderer *_renderer) {
  if (_renderer) lgpu_progressive_start(_renderer->p);
}

void lambdaStopProgressiveRenderer(LAMBDA_ProgressiveRenderer *_renderer) {
  if (_renderer) lgpu_progressive_stop(_renderer->p);
}

}  // extern "C"
