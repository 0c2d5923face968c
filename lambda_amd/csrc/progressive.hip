// Progressive rendering (ProgressiveRender.cpp:13-62; lambdaCreate/Start/Stop/...ProgressiveRenderer,
// Lambda.cpp:431-458) and the film's RGBA32F readback (Film::ToRGBTexture, Film.cpp:11-24).
//
// A worker thread renders passes of `n` samples per pixel (sample indices [begin + k n,
// begin + (k + 1) n)) with the render call, converts the film to RGBA32F on the device, copies it
// into the back one of two pinned host buffers and swaps it to the front under a mutex, then
// calls the update callback. Readers copy the front buffer out under the same mutex, so they never
// see a half-written image (the reference hands out a pointer the worker keeps rewriting).
// Because the random numbers are keyed on the sample index, the film after k passes equals one
// render of [begin, begin + k n).
#include <atomic>
#include <cstring>
#include <thread>

#include "host.h"

namespace lgpu {

namespace {

__global__ void k_film_rgba(const double *sum, const uint32_t *count, float4 *out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = count[i];
  const double inv = c ? 1.0 / (double)c : 0.0;  // pixels without samples read 0 (Film.cpp:16 divides by 0)
  out[i] = make_float4((float)(sum[3 * i] * inv), (float)(sum[3 * i + 1] * inv), (float)(sum[3 * i + 2] * inv), 1.f);
}

}  // namespace

// Film::ToRGBTexture into host memory (rgba: 4 floats per pixel). The caller holds the device mutex.
lgpu_status FilmToRgba(lgpu_film f, float *rgba) {
  const uint32_t n = f->width * f->height;
  LGPU_HIP(hipSetDevice(f->dev->device));
  DevBuf out;
  LGPU_HIP(out.alloc(16ull * n));
  hipLaunchKernelGGL(k_film_rgba, dim3((n + 255) / 256), dim3(256), 0, f->dev->stream, f->sum.as<double>(),
                     f->count.as<uint32_t>(), out.as<float4>(), n);
  LGPU_HIP(hipGetLastError());
  LGPU_HIP(hipMemcpyAsync(rgba, out.ptr, 16ull * n, hipMemcpyDeviceToHost, f->dev->stream));
  LGPU_HIP(hipStreamSynchronize(f->dev->stream));
  return LGPU_OK;
}

}  // namespace lgpu

using namespace lgpu;

struct lgpu_progressive_s {
  lgpu_render_desc desc;      // scene, film, camera, integrator...; spp range = the pass size
  uint32_t pass_spp = 1;
  lgpu_progressive_callback callback = nullptr;
  void *user = nullptr;
  std::thread worker;
  std::atomic<bool> running{false}, stop{false}, clear{false};
  std::mutex data_mu;         // guards front, passes_front, error
  float *buf[2] = {nullptr, nullptr};  // pinned host RGBA32F images
  int front = 0;
  uint32_t passes = 0, passes_front = 0;
  lgpu_status error = LGPU_OK;
  std::string error_msg;

  void Run() {
    const lgpu_film f = desc.film;
    while (!stop.load()) {
      if (clear.exchange(false)) {
        const lgpu_status st = lgpu_film_clear(f);
        if (st != LGPU_OK) { Record(st); break; }
        passes = 0;
      }
      lgpu_render_desc d = desc;
      d.spp_begin = desc.spp_begin + passes * pass_spp;
      d.spp_end = d.spp_begin + pass_spp;
      lgpu_status st = lgpu_render(&d, nullptr);  // ProgressiveRender::RunPass: every tile, one increment
      if (st == LGPU_OK) {
        std::lock_guard<std::mutex> lk(f->dev->mu);
        st = FilmToRgba(f, buf[front ^ 1]);  // ProgressiveRender::UpdateOutputTexture
      }
      if (st != LGPU_OK) { Record(st); break; }
      ++passes;
      {
        std::lock_guard<std::mutex> lk(data_mu);
        front ^= 1;
        passes_front = passes;
      }
      if (callback) callback(user);
    }
    running.store(false);
  }
  void Record(lgpu_status st) {
    std::lock_guard<std::mutex> lk(data_mu);
    error = st;
    error_msg = lgpu_last_error();
  }
  void Join() {
    stop.store(true);
    if (worker.joinable()) worker.join();
    stop.store(false);
  }
};

extern "C" {

lgpu_status lgpu_progressive_create(const lgpu_render_desc *d, lgpu_progressive *out) {
  if (!d || !out) return Fail(LGPU_ERR_INVALID_ARGUMENT, "progressive_create: null argument");
  if (!d->scene || !d->film) return Fail(LGPU_ERR_INVALID_ARGUMENT, "progressive_create: needs a scene and a film");
  if (!d->scene->committed) return Fail(LGPU_ERR_STATE, "progressive_create: scene not committed");
  if (d->spp_end <= d->spp_begin) return Fail(LGPU_ERR_INVALID_ARGUMENT, "progressive_create: empty pass (spp range)");
  auto *p = new lgpu_progressive_s;
  p->desc = *d;
  p->pass_spp = d->spp_end - d->spp_begin;
  const size_t bytes = 16ull * d->film->width * d->film->height;
  hipError_t e = hipSetDevice(d->scene->dev->device);
  for (int k = 0; k < 2 && e == hipSuccess; ++k) e = hipHostMalloc((void **)&p->buf[k], bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    for (float *b : p->buf) if (b) (void)hipHostFree(b);
    delete p;
    return Fail(LGPU_ERR_OUT_OF_MEMORY, std::string("progressive_create: pinned buffers: ") + hipGetErrorString(e));
  }
  std::memset(p->buf[0], 0, bytes);
  std::memset(p->buf[1], 0, bytes);
  *out = p;
  return LGPU_OK;
}

lgpu_status lgpu_progressive_destroy(lgpu_progressive p) {
  if (!p) return LGPU_OK;
  p->Join();
  for (float *b : p->buf) if (b) (void)hipHostFree(b);
  delete p;
  return LGPU_OK;
}

lgpu_status lgpu_progressive_set_callback(lgpu_progressive p, lgpu_progressive_callback cb, void *user) {
  if (!p) return Fail(LGPU_ERR_INVALID_ARGUMENT, "progressive_set_callback: null renderer");
  if (p->running.load()) return Fail(LGPU_ERR_STATE, "progressive_set_callback: renderer is running");
  p->callback = cb;
  p->user = user;
  return LGPU_OK;
}

lgpu_status lgpu_progressive_start(lgpu_progressive p) {
  if (!p) return Fail(LGPU_ERR_INVALID_ARGUMENT, "progressive_start: null renderer");
  if (p->running.load()) return LGPU_OK;  // ProgressiveRender::Init runs once
  if (p->worker.joinable()) p->worker.join();
  {
    std::lock_guard<std::mutex> lk(p->data_mu);
    p->error = LGPU_OK;
    p->error_msg.clear();
  }
  p->running.store(true);
  p->worker = std::thread([p] { p->Run(); });
  return LGPU_OK;
}

lgpu_status lgpu_progressive_stop(lgpu_progressive p) {
  if (!p) return Fail(LGPU_ERR_INVALID_ARGUMENT, "progressive_stop: null renderer");
  p->Join();  // the pass in flight completes first
  return LGPU_OK;
}

lgpu_status lgpu_progressive_clear(lgpu_progressive p) {
  if (!p) return Fail(LGPU_ERR_INVALID_ARGUMENT, "progressive_clear: null renderer");
  if (p->running.load()) {
    p->clear.store(true);  // applied by the worker between passes
    return LGPU_OK;
  }
  const lgpu_status st = lgpu_film_clear(p->desc.film);
  if (st != LGPU_OK) return st;
  p->passes = 0;
  std::lock_guard<std::mutex> lk(p->data_mu);
  p->passes_front = 0;
  std::memset(p->buf[p->front], 0, 16ull * p->desc.film->width * p->desc.film->height);
  return LGPU_OK;
}

lgpu_status lgpu_progressive_read(lgpu_progressive p, float *rgba, uint32_t *passes) {
  if (!p) return Fail(LGPU_ERR_INVALID_ARGUMENT, "progressive_read: null renderer");
  std::lock_guard<std::mutex> lk(p->data_mu);
  if (rgba) std::memcpy(rgba, p->buf[p->front], 16ull * p->desc.film->width * p->desc.film->height);
  if (passes) *passes = p->passes_front;
  if (p->error != LGPU_OK) return Fail(p->error, "progressive: " + p->error_msg);
  return LGPU_OK;
}

int lgpu_progressive_running(lgpu_progressive p) { return p && p->running.load() ? 1 : 0; }

}  // extern "C"
