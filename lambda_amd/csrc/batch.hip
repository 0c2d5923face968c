// Batched renders: many (scene, film, camera) jobs rendered together from one shared path pool.
// Each chunk of the batch is one hipGraph holding every job's whole wave: counter init, camera,
// a fixed number of bounce iterations (extend, shade, shadow; item counts read from device
// counters, grids over the job's segment) and the film update. Jobs are recorded on several
// capture streams, so the graph holds independent branches that the device may overlap: small
// scenes run beside large ones instead of after them. The kernels are those of lgpu_render; a
// graph launch replaces the per-bounce host round trips of the single-scene loops. Replaces the
// reference's tile dispatch (MosaicRenderer.cpp:22-80, Render.cpp:11-39), one scene at a time.
//
// A job's segment is sized in proportion to its paths, so jobs finish in about the same number of
// chunks. The autotuner times one-sample trials over pool capacities and branch counts and keeps
// the fastest per batch signature.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "host.h"
#include "trace.h"

namespace lgpu {

namespace {

constexpr uint32_t kMaxBranches = 16;

struct BatchPool {
  DevBuf st[2][9], L, pix, lam, shadow, spill, scenes, ctr, trunc;
};

lgpu_status Validate(const lgpu_render_desc *d, uint32_t count) {
  if (!d || count == 0) return Fail(LGPU_ERR_INVALID_ARGUMENT, "render_batch: no jobs");
  lgpu_device dev = d[0].scene ? d[0].scene->dev : nullptr;
  for (uint32_t j = 0; j < count; ++j) {
    const lgpu_render_desc &x = d[j];
    const std::string at = "render_batch: job " + std::to_string(j) + ": ";
    if (!x.scene || !x.film) return Fail(LGPU_ERR_INVALID_ARGUMENT, at + "null scene or film");
    if (!x.scene->committed) return Fail(LGPU_ERR_STATE, at + "scene not committed");
    if (x.scene->dev != dev || x.film->dev != dev) return Fail(LGPU_ERR_INVALID_ARGUMENT, at + "jobs on different devices");
    if (x.spp_end <= x.spp_begin) return Fail(LGPU_ERR_INVALID_ARGUMENT, at + "empty spp range");
    if (x.max_bounces == 0) return Fail(LGPU_ERR_INVALID_ARGUMENT, at + "max_bounces must be > 0");
    if (x.integrator != LGPU_INTEGRATOR_PATH && x.integrator != LGPU_INTEGRATOR_VOLPATH &&
        x.integrator != LGPU_INTEGRATOR_MIS_VOLPATH)
      return Fail(LGPU_ERR_UNSUPPORTED, at + "batches take the path and volumetric path integrators");
    if (x.estimator_mode != LGPU_ESTIMATOR_CORRECTED || x.guide || x.spectral_mode > LGPU_SPECTRAL_HERO4)
      return Fail(LGPU_ERR_UNSUPPORTED, at + "batches take the corrected estimator without a guide");
    if (x.camera.type > LGPU_CAMERA_SPHERICAL || (x.camera.aperture == LGPU_APERTURE_BLADE && x.camera.blades == 0))
      return Fail(LGPU_ERR_INVALID_ARGUMENT, at + "bad camera");
    for (uint32_t k = 0; k < j; ++k)
      if (d[k].film == x.film) return Fail(LGPU_ERR_INVALID_ARGUMENT, at + "film shared with job " + std::to_string(k));
  }
  return LGPU_OK;
}

// Streams the capture forks into (default priority), kept per device object's lifetime.
struct Branches {
  hipStream_t s[kMaxBranches] = {};
  hipEvent_t fork = nullptr, join[kMaxBranches] = {};
  ~Branches() {
    for (auto x : s)
      if (x) (void)hipStreamDestroy(x);
    for (auto e : join)
      if (e) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
  }
  lgpu_status init() {
    if (fork) return LGPU_OK;
    for (auto &x : s) LGPU_HIP(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
    for (auto &e : join) LGPU_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    LGPU_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    return LGPU_OK;
  }
};
// The device object owns them (released with the device; calls on one device are serialised).
Branches &BranchesOf(lgpu_device dev) {
  if (!dev->batch) {
    dev->batch = new Branches();
    dev->batch_free = [](void *p) { delete static_cast<Branches *>(p); };
  }
  return *static_cast<Branches *>(dev->batch);
}

// One batched render of spp [spp_begin, spp_begin + trialSpp) of every job (trialSpp 0: the
// full ranges) from a pool of `capacity` paths on `branches` capture streams; accumulate = false
// leaves the films untouched.
lgpu_status RunBatch(const lgpu_render_desc *d, uint32_t count, uint64_t capacity, uint32_t branches,
                     uint32_t trialSpp, bool accumulate, uint32_t iterations, lgpu_batch_stats *out) {
  lgpu_device dev = d[0].scene->dev;
  LGPU_HIP(hipSetDevice(dev->device));
  (void)hipGetLastError();
  const auto t0 = std::chrono::steady_clock::now();
  const uint32_t J = count;
  branches = std::max<uint32_t>(1, std::min<uint32_t>(branches, kMaxBranches));
  std::vector<uint64_t> total(J);
  uint64_t sumTotal = 0;
  uint32_t maxB = 0;
  for (uint32_t j = 0; j < J; ++j) {
    const uint64_t npx = (uint64_t)d[j].film->width * d[j].film->height;
    const uint32_t spp = trialSpp ? std::min(trialSpp, d[j].spp_end - d[j].spp_begin) : d[j].spp_end - d[j].spp_begin;
    total[j] = npx * spp;
    sumTotal += total[j];
    maxB = std::max(maxB, d[j].max_bounces);
  }
  // bounce iterations per chunk: bounces plus pass-through crossings (which do not count as bounces)
  const uint32_t iters = iterations ? iterations : 2 * maxB + 8;
  const uint32_t ctrWords = 2 * (iters + 1);
  // segments in proportion to the jobs' paths, at least one block each
  std::vector<uint32_t> cap(J);
  std::vector<uint64_t> off(J + 1, 0), spillOff(J + 1, 0);
  for (uint32_t j = 0; j < J; ++j) {
    uint64_t c = (uint64_t)((double)capacity * (double)total[j] / (double)sumTotal + 0.5);
    c = std::max<uint64_t>(c, (uint64_t)kBlock);
    c = std::min<uint64_t>(c, total[j]);
    c = std::min<uint64_t>(c, 1ull << 30);
    cap[j] = (uint32_t)c;
    off[j + 1] = off[j] + c;
    spillOff[j + 1] = spillOff[j] + c * d[j].scene->dsc.spill_words;
  }
  const uint64_t N = off[J];
  BatchPool P;
  for (int b = 0; b < 2; ++b)
    for (int k = 0; k < 9; ++k) LGPU_HIP(P.st[b][k].alloc(16 * N));
  LGPU_HIP(P.L.alloc(16 * N));
  LGPU_HIP(P.pix.alloc(4 * N));
  LGPU_HIP(P.lam.alloc(4 * N));
  LGPU_HIP(P.shadow.alloc(sizeof(ShadowRay) * N));
  LGPU_HIP(P.spill.alloc(std::max<uint64_t>(256, 4 * spillOff[J])));
  std::vector<DevScene> scenes(J);
  for (uint32_t j = 0; j < J; ++j) scenes[j] = d[j].scene->dsc;
  LGPU_HIP(P.scenes.upload(scenes.data(), sizeof(DevScene) * J));
  LGPU_HIP(P.ctr.alloc(4ull * ctrWords * J));
  LGPU_HIP(P.trunc.alloc(8));
  LGPU_HIP(hipMemset(P.trunc.ptr, 0, 8));
  LGPU_HIP(hipDeviceSynchronize());

  struct Job {
    RenderParams rp;
    PathState ps[2];
    SampleState ss;
    ShadowRay *shadow;
    uint32_t *spill, *ctr;
    const DevScene *scp;
    bool path, programs;
  };
  std::vector<Job> jobs(J);
  for (uint32_t j = 0; j < J; ++j) {
    Job &b = jobs[j];
    b.rp = MakeParams(&d[j], d[j].film->width, d[j].film->height);
    for (int q = 0; q < 2; ++q) {
      PathState &p = b.ps[q];
      p.ray_o = P.st[q][0].as<float4>() + off[j]; p.ray_d = P.st[q][1].as<float4>() + off[j];
      p.beta = P.st[q][2].as<float4>() + off[j]; p.hit = P.st[q][3].as<float4>() + off[j];
      p.ids = P.st[q][4].as<uint4>() + off[j]; p.aux = P.st[q][5].as<uint4>() + off[j];
      p.mis = P.st[q][6].as<float4>() + off[j]; p.prev_p = P.st[q][7].as<float4>() + off[j];
      p.prev_n = P.st[q][8].as<float4>() + off[j];
    }
    b.ss.L = P.L.as<float4>() + off[j];
    b.ss.pixel = P.pix.as<uint32_t>() + off[j];
    b.ss.lambda0 = P.lam.as<float>() + off[j];
    b.shadow = P.shadow.as<ShadowRay>() + off[j];
    b.spill = P.spill.as<uint32_t>() + spillOff[j];
    b.ctr = P.ctr.as<uint32_t>() + (size_t)ctrWords * j;
    b.scp = P.scenes.as<DevScene>() + j;
    b.path = d[j].integrator == LGPU_INTEGRATOR_PATH;
    b.programs = scenes[j].has_programs != 0;
  }
  Branches &br = BranchesOf(dev);
  { lgpu_status s = br.init(); if (s != LGPU_OK) return s; }

  hipStream_t origin = dev->stream;
  uint64_t chunks = 0;
  for (uint32_t j = 0; j < J; ++j) chunks = std::max<uint64_t>(chunks, (total[j] + cap[j] - 1) / cap[j]);
  uint32_t nodes = 0;
  double msCapture = 0.0;
  std::vector<hipGraphExec_t> execs;
  auto cleanup = [&]() {
    for (hipGraphExec_t e : execs) (void)hipGraphExecDestroy(e);
    execs.clear();
  };
  const auto t1 = std::chrono::steady_clock::now();
  for (uint64_t c = 0; c < chunks; ++c) {
    // capture chunk c: every job with paths left, round-robin over the branches
    const auto tc = std::chrono::steady_clock::now();
    hipError_t e = hipStreamBeginCapture(origin, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) { cleanup(); return Fail(LGPU_ERR_HIP, std::string("render_batch: capture: ") + hipGetErrorString(e)); }
    (void)hipEventRecord(br.fork, origin);
    for (uint32_t k = 0; k < branches; ++k) (void)hipStreamWaitEvent(br.s[k], br.fork, 0);
    for (uint32_t j = 0; j < J; ++j) {
      const uint64_t base = c * (uint64_t)cap[j];
      if (base >= total[j]) continue;
      const uint32_t n = (uint32_t)std::min<uint64_t>(cap[j], total[j] - base);
      const Job &b = jobs[j];
      hipStream_t s = br.s[j % branches];
      const uint32_t w = d[j].film->width, h = d[j].film->height;
      BatchLaunchInit(s, b.ctr, ctrWords, n);
      BatchLaunchCamera(s, b.rp, b.ps[0], b.ss, n, base);
      nodes += 2;
      for (uint32_t it = 0; it < iters; ++it) {
        const uint32_t cur = it & 1u;
        uint32_t *cin = b.ctr + 2 * it, *cout = b.ctr + 2 * (it + 1);
        BatchLaunchExtend(s, scenes[j], b.ps[cur], n, cin, b.spill);
        const Queues q{b.ps[cur ^ 1u], b.shadow, cout};
        if (b.path) {
          BatchLaunchShade(s, b.scp, b.rp, b.ps[cur], b.ss, n, cin, q, b.programs);
          BatchLaunchShadow(s, scenes[j], b.ss, b.shadow, n, cout + 1, b.spill);
          nodes += 3;
        } else {
          LaunchShadeVol((n + kBlock - 1) / kBlock, s, scenes[j], b.rp, b.ps[cur], b.ss, n, q, b.spill, nullptr, cin);
          nodes += 2;
        }
      }
      BatchLaunchTrunc(s, b.ctr + 2 * iters, P.trunc.as<unsigned long long>());
      if (accumulate)
        BatchLaunchAccumulate(s, b.rp, b.ss, n, w * h, d[j].film->sum.as<double>(), d[j].film->count.as<uint32_t>());
      nodes += accumulate ? 2 : 1;
    }
    for (uint32_t k = 0; k < branches; ++k) {
      (void)hipEventRecord(br.join[k], br.s[k]);
      (void)hipStreamWaitEvent(origin, br.join[k], 0);
    }
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(origin, &graph);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      cleanup();
      return Fail(LGPU_ERR_HIP, std::string("render_batch: capture: ") + hipGetErrorString(e));
    }
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      cleanup();
      return Fail(LGPU_ERR_HIP, std::string("render_batch: instantiate: ") + hipGetErrorString(e));
    }
    execs.push_back(exec);
    msCapture += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tc).count();
    e = hipGraphLaunch(exec, origin);  // runs while the next chunk is captured
    if (e != hipSuccess) {
      cleanup();
      return Fail(LGPU_ERR_HIP, std::string("render_batch: launch: ") + hipGetErrorString(e));
    }
  }
  const hipError_t se = hipStreamSynchronize(origin);
  const auto t2 = std::chrono::steady_clock::now();
  cleanup();
  if (se != hipSuccess) return Fail(LGPU_ERR_HIP, std::string("render_batch: ") + hipGetErrorString(se));
  unsigned long long trunc = 0;
  LGPU_HIP(hipMemcpy(&trunc, P.trunc.ptr, 8, hipMemcpyDeviceToHost));
  if (trunc)
    return Fail(LGPU_ERR_STATE, "render_batch: " + std::to_string(trunc) +
                                    " paths outlived the captured iterations (raise max_iterations)");
  if (out) {
    std::memset(out, 0, sizeof(*out));
    out->samples = sumTotal;
    out->ms_render = std::chrono::duration<double, std::milli>(t2 - t1).count();
    out->ms_capture = msCapture;
    out->ms_total = std::chrono::duration<double, std::milli>(t2 - t0).count();
    out->chunks = (uint32_t)chunks;
    out->graph_nodes = nodes;
    out->path_capacity = (uint32_t)std::min<uint64_t>(N, 0xFFFFFFFFu);
    out->branches = branches;
    out->iterations = iters;
    out->jobs = J;
  }
  return LGPU_OK;
}

std::string Signature(const lgpu_render_desc *d, uint32_t count) {
  std::string s;
  for (uint32_t j = 0; j < count; ++j) {
    char buf[160];
    std::snprintf(buf, sizeof(buf), "%p:%ux%u:%u:%u:%u:%u;", (void *)d[j].scene, d[j].film->width, d[j].film->height,
                  d[j].max_bounces, d[j].integrator, d[j].spectral_mode, d[j].spp_end - d[j].spp_begin);
    s += buf;
  }
  return s;
}

struct TuneChoice {
  uint64_t capacity;
  uint32_t branches;
};
std::map<std::string, TuneChoice> &TuneCache() {
  static std::map<std::string, TuneChoice> m;
  return m;
}

}  // namespace

}  // namespace lgpu

using namespace lgpu;

extern "C" lgpu_status lgpu_render_batch(const lgpu_render_desc *descs, uint32_t count, const lgpu_batch_options *opt,
                                         lgpu_batch_stats *stats) {
  lgpu_status st = Validate(descs, count);
  if (st != LGPU_OK) return st;
  lgpu_device dev = descs[0].scene->dev;
  std::lock_guard<std::mutex> lk(dev->mu);
  lgpu_batch_options o{};
  if (opt) o = *opt;
  uint64_t capacity = o.path_capacity ? o.path_capacity : (1ull << 24);
  uint32_t branches = o.branches ? std::min<uint32_t>(o.branches, kMaxBranches) : 8u;
  double tuneMs = 0.0;
  if (o.flags & LGPU_BATCH_AUTOTUNE) {
    const std::string key = Signature(descs, count);
    auto it = TuneCache().find(key);
    if (it == TuneCache().end()) {
      const auto t0 = std::chrono::steady_clock::now();
      double best = -1.0;
      TuneChoice pick{capacity, branches};
      // Candidate pools from 2^23 paths up to the first that holds the whole batch (one chunk),
      // within a quarter of the free device memory; each trial renders the whole batch, since the
      // chunk count (and with it the graph's node count) depends on the real sample ranges.
      uint64_t paths = 0, spillWords = 0;
      for (uint32_t j = 0; j < count; ++j) {
        paths += (uint64_t)descs[j].film->width * descs[j].film->height * (descs[j].spp_end - descs[j].spp_begin);
        spillWords = std::max<uint64_t>(spillWords, descs[j].scene->dsc.spill_words);
      }
      const uint64_t bytesPerPath = 2 * 9 * 16 + 16 + 4 + 4 + sizeof(ShadowRay) + 4 * spillWords;
      size_t freeB = 0, totalB = 0;
      LGPU_HIP(hipMemGetInfo(&freeB, &totalB));
      for (uint64_t c = 1ull << 23; c <= (1ull << 27); c <<= 1) {
        if (c > (1ull << 23) && c * bytesPerPath > freeB / 4) break;
        for (uint32_t b : {4u, 8u, 16u}) {
          lgpu_batch_stats t;
          st = RunBatch(descs, count, c, b, 0, false, o.max_iterations, &t);
          if (st != LGPU_OK) return st;
          const double rate = (double)t.samples / std::max(1e-6, t.ms_render);
          if (rate > best) { best = rate; pick = TuneChoice{c, b}; }
        }
        if (c >= paths) break;  // one chunk already
      }
      tuneMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      it = TuneCache().emplace(key, pick).first;
    }
    capacity = it->second.capacity;
    branches = it->second.branches;
  }
  st = RunBatch(descs, count, capacity, branches, 0, true, o.max_iterations, stats);
  if (st == LGPU_OK && stats) stats->ms_autotune = tuneMs;
  return st;
}
