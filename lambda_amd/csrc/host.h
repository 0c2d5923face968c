// Host-side objects behind the opaque C-ABI handles.
#pragma once
#include <hip/hip_runtime.h>

#include <deque>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lambda_gpu.h"
#include "lgpu_internal.h"

namespace lgpu {

void SetError(const std::string &msg);
lgpu_status Fail(lgpu_status s, const std::string &msg);
#define LGPU_HIP(call)                                                                             \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess) return ::lgpu::Fail(LGPU_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

// A device allocation owned by the host object (move-only).
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); ptr = o.ptr; bytes = o.bytes; o.ptr = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  hipError_t upload(const void *src, size_t n) {
    release();
    if (!n) return hipSuccess;
    hipError_t e = hipMalloc(&ptr, n);
    if (e != hipSuccess) return e;
    bytes = n;
    return hipMemcpy(ptr, src, n, hipMemcpyHostToDevice);
  }
  hipError_t alloc(size_t n) {
    if (n <= bytes && ptr) return hipSuccess;
    release();
    hipError_t e = hipMalloc(&ptr, n);
    if (e == hipSuccess) bytes = n;
    return e;
  }
  template <class T> T *as() const { return static_cast<T *>(ptr); }
};

struct HostMaterial {
  std::vector<lgpu_node> nodes;
  lgpu_material_desc desc;
};
struct HostMesh {
  std::vector<float> pos, nrm, tan, uv;
  std::vector<uint32_t> idx;
  uint32_t material;
};
struct HostLight {
  lgpu_light_desc desc;
  std::vector<float> env;
  int mesh = -1;  // mesh lights
};

// Film::ToRGBTexture into host memory (4 floats per pixel); the caller holds the device mutex.
lgpu_status FilmToRgba(lgpu_film f, float *rgba);

}  // namespace lgpu

struct lgpu_device_s {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;  // one caller at a time on the stream and the wave pool (renders, film reads)
  hipEvent_t ev[8] = {};
  void *wave = nullptr;               // path-state pool of the render loop
  void (*wave_free)(void *) = nullptr;
  void *batch = nullptr;              // capture streams and events of batched renders
  void (*batch_free)(void *) = nullptr;
  void *scratch = nullptr;            // device address of this GPU's no-allocation scratch array
  std::vector<void *> guides;         // live guiding networks (released with the device)
  void (*guide_orphan)(void *) = nullptr;
};

struct lgpu_film_s {
  lgpu_device dev = nullptr;
  uint32_t width = 0, height = 0;
  lgpu::DevBuf sum;    // double x3 per pixel
  lgpu::DevBuf count;  // uint32 per pixel
};

struct lgpu_scene_s {
  lgpu_device dev = nullptr;
  bool committed = false;
  std::vector<lgpu::HostMaterial> materials;
  std::vector<lgpu::HostMesh> meshes;
  std::vector<lgpu::HostMesh> proxies;  // InstanceProxy objects: instanced, not in the scene themselves
  std::vector<lgpu::HostLight> lights;  // scene light order (mesh lights interleaved)
  std::vector<std::vector<float>> textures;
  std::vector<uint32_t> tex_dims;        // w, h, ch per texture
  std::vector<lgpu_medium_desc> media;
  std::vector<std::vector<float>> media_density;

  // host copies kept for introspection
  std::vector<lgpu::TreeNode> tree;
  uint32_t num_bvh_nodes = 0;
  uint32_t bvh_depth = 0;
  uint32_t num_tris = 0;
  uint32_t num_tlights = 0;

  float bounds[6] = {0, 0, 0, 0, 0, 0};  // scene box (min xyz, max xyz) of all triangles
  // adjoint render and in-place parameter updates
  bool adj_texture_shared = false;       // a texture-bound socket's texture is also read elsewhere
  std::vector<uint8_t> tex_emissive;     // per texture: read by an emission socket
  std::vector<uint32_t> texel_off;       // per texture: offset of its texels in DevScene::texels
  std::vector<uint32_t> density_off;     // per medium: offset of its grid in DevScene::density
  lgpu::DevScene dsc{};
  std::deque<lgpu::DevBuf> bufs;  // owns every array referenced by dsc (stable addresses)
};

namespace lgpu {
lgpu_status CommitScene(lgpu_scene s, uint32_t strategy);
lgpu_status ResolveScratch(lgpu_device dev);  // render.hip: the no-allocation scratch array's address
RenderParams MakeParams(const lgpu_render_desc *d, uint32_t w, uint32_t h);  // render.hip
}
