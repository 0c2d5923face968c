"""Benchmark: Msamples/s of the wavefront render call on one MI355X.

Workload (default): the many-light glossy mesh scene — 1,204,514 triangles
(200 displaced level-4 icospheres + 297x297-quad ground + 2,048 emissive quads
= 4,096 emissive triangles sampled through the light tree), 1024x512 FP32
procedural sky, thin-lens camera at 1920x1080, max depth 16, hero-wavelength
spectral sampling, corrected estimator. One step = one sample per pixel over
the whole frame (2,073,600 paths); --steps K times K such passes in one render
call. Synthetic procedural data (no assets).

Prints one JSON line with the metric, a roofline entry for the dominant kernel
(BVH extension traversal) and a CPU baseline (the CPU checker on the host
cores over a bounded crop of the same workload).

Other workloads (--workload): "media" (128^3 FP32 fBm density grid in a 2x2x2
pass-through box, 1 spotlight + 64 point lights, MIS distance + equiangular
volumetric integrator, 1024x1024, 128 spp, depth 32) and "cornell" (36 triangles,
256x256, depth 8). --relmse measures time-to-relMSE <= 1e-3 of progressive
rendering against an FP64 GPU reference at --ref-spp with another seed.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


DEFAULT_STEPS = {"many-light": 256, "media": 128, "cornell": 64, "inverse": 4, "batch": 4}


def build_workload(name):
    from lambda_amd import scene
    if name == "cornell":
        sd = scene.cornell_box()
        W, H = 256, 256
        cam = scene.cornell_camera(W, H)
        cfg = {"workload": "cornell-box 36 triangles, pinhole 256x256, max depth 8",
               "width": W, "height": H, "max_bounces": 8, "triangles": sd.n_tris, "integrator": "path"}
        return sd, cam, cfg
    if name == "media":
        sd = scene.media_scene(res=128)
        W, H = 1024, 1024
        cam = scene.media_camera(W, H)
        cfg = {"workload": "heterogeneous media: 128^3 FP32 fBm density grid (sigma_t 4 d, albedo 0.9, HG g 0.6) "
                           "in a pass-through 2x2x2 box, 1 spotlight + 64 point lights, MIS distance + equiangular "
                           "volumetric integrator, pinhole 1024x1024, depth 32",
               "width": W, "height": H, "max_bounces": 32, "triangles": sd.n_tris, "integrator": "mis_volpath"}
        return sd, cam, cfg
    sd = scene.many_light_scene()
    W, H = 1920, 1080
    cam = scene.many_light_camera(W, H)
    cfg = {"workload": "many-light glossy mesh: 1.2M triangles (200 displaced icospheres + ground), 4096 "
                       "emissive triangles via the light tree, 1024x512 HDR sky, thin lens 1920x1080, depth 16",
           "width": W, "height": H, "max_bounces": 16, "triangles": sd.n_tris, "integrator": "path"}
    return sd, cam, cfg


def measured_traffic(workload):
    """HBM bytes per unit of work measured with FETCH_SIZE / WRITE_SIZE (copy-kernel calibrated)
    for this workload: the `traffic` block of its committed counter summary, or {}."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        {"many-light": "hbm_traffic_many_light_1080p_pmc.md",
                         "media": "hbm_traffic_media_1024_pmc.md"}.get(workload, ""))
    out = {}
    if not os.path.isfile(path):
        return out
    block = False
    for line in open(path):
        if line.startswith("```traffic"):
            block = True
            continue
        if block and line.startswith("```"):
            break
        if block:
            f = line.split()
            if len(f) == 6 and f[2] == "read" and f[4] == "write":
                out[f[0]] = {"read": float(f[3]), "write": float(f[5]), "unit": f[1],
                             "source": "profiles/" + os.path.basename(path)}
    return out


def measured_valu_busy(workload):
    """VALU busy share per kernel (SQ_ACTIVE_INST_VALU over SIMD cycles) from the committed counter
    summary, for this workload: {kernel base name: fraction}, or {}."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "valu_lds_utilisation.md")
    out = {}
    if not os.path.isfile(path):
        return out
    for line in open(path):
        f = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(f) == 10 and f[0] == workload and f[1].startswith("k_") and f[6].endswith("%"):
            out.setdefault(f[1].split("<")[0], float(f[6].rstrip(" %")) / 100.0)
    return out


def measured_time_to_relmse():
    """Plain and guided time-to-relMSE <= 1e-3 on the many-light workload from the committed
    measurement (bench.py --relmse --guided), or None."""
    import re
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "relmse_many_light_1080p_plain_vs_guided.md")
    if not os.path.isfile(path):
        return None
    m = re.search(r"\*\*plain ([0-9.]+) s\*\* \(([0-9]+) spp.*?\*\*guided ([0-9.]+) s\*\* \(([0-9]+) spp",
                  open(path).read())
    if not m:
        return None
    p, g = float(m.group(1)), float(m.group(3))
    return {"plain_s": p, "plain_spp": int(m.group(2)), "guided_s": g, "guided_spp": int(m.group(4)),
            "guided_over_plain": round(g / p, 3), "source": "profiles/" + os.path.basename(path)}


def guided_throughput(desc, film, W, H, dev):
    """The guided estimator on the same workload: a fresh guide trained online as in the
    time-to-relMSE measurement (256 spp in 8-spp waves with 16 Adam steps each, untimed), then 256
    timed spp in 64-spp waves with 4 steps each (wall time of the render calls, inference and
    training included). Guided vertices: path depth 0 and 1."""
    from lambda_amd import gpu
    g = gpu.Guide(dev, seed=7, max_depth=2)
    try:
        gpu.render(desc(0, 1, guide=g, capacity=W * H * 64))  # sizes the guided path pool (untimed)
        for b in range(0, 256, 8):
            gpu.render(desc(b, b + 8, guide=g, steps=16, capacity=W * H * 8))
        film.clear()
        kern = {"ms_extend": 0.0, "ms_shade": 0.0, "ms_shadow": 0.0, "ms_guide_infer": 0.0, "ms_guide_train": 0.0}
        t0 = time.perf_counter()
        for b in range(256, 512, 64):
            st = gpu.render(desc(b, b + 64, guide=g, steps=4, capacity=W * H * 64))
            for k in kern:
                kern[k] += getattr(st, k)
        dt = time.perf_counter() - t0
    finally:
        g.close()
    out = {"msamples_per_s": round(W * H * 256 / dt / 1e6, 1), "spp": 256, "seconds": round(dt, 4),
           "schedule": "after 256 spp of online training in 8-spp waves: 64-spp waves, 4 Adam steps each, "
                       "guided vertices at path depth 0 and 1"}
    out.update({k: round(v, 1) for k, v in kern.items()})
    return out


def integrator_id(cfg):
    from lambda_amd import abi
    return {"path": abi.INTEGRATOR_PATH, "volpath": abi.INTEGRATOR_VOLPATH,
            "mis_volpath": abi.INTEGRATOR_MIS_VOLPATH, "direct": abi.INTEGRATOR_DIRECT}[cfg["integrator"]]


def cpu_baseline(sd, cam, cfg, spectral, estimator, budget_s):
    """The CPU checker timed on the host cores over a centre crop of the same workload."""
    import oracle
    from lambda_amd import abi
    threads = max(1, min(os.cpu_count() or 1, 16))
    o = oracle.OracleScene()
    keep = sd.push(o, abi.LIGHTS_TREE)
    W, H = cfg["width"], cfg["height"]
    cw = min(256, W)
    ch = min(256, H)
    win = ((W - cw) // 2, (H - ch) // 2, cw, ch)
    d = abi.RenderDesc()
    d.camera = cam
    d.integrator, d.max_bounces, d.min_bounces = integrator_id(cfg), cfg["max_bounces"], 3
    d.spectral_mode, d.estimator_mode, d.seed = spectral, estimator, 0
    d.spp_begin, d.spp_end = 0, 1
    t0 = time.perf_counter()
    o.render(d, W, H, window=win, threads=threads)
    probe = time.perf_counter() - t0
    spp = max(1, min(256, int(budget_s / max(probe, 1e-3))))
    d.spp_begin, d.spp_end = 1, 1 + spp
    t0 = time.perf_counter()
    o.render(d, W, H, window=win, threads=threads)
    dt = time.perf_counter() - t0
    samples = cw * ch * spp
    del keep
    return {"value": samples / dt / 1e6, "unit": "Msamples/s", "cores": threads, "kind": "port",
            "sample": "%dx%d centre crop at %d spp (%d samples, %.1f s) of the same workload and estimator"
                      % (cw, ch, spp, samples, dt), "host": host_cpu()}


def host_cpu():
    """The host the CPU baseline ran on: logical CPUs, this process's affinity, the CPU model.
    The baseline uses at most 16 threads: the CPU share one GPU's job gets on these machines."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    return {"os_cpu_count": os.cpu_count(), "affinity_cpus": aff, "model": model, "threads_cap": 16}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="spp timed (default: 256 many-light, 128 media, 64 cornell)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="many-light", choices=["many-light", "media", "cornell", "inverse", "batch"],
                    help="inverse: one differentiable-rendering iteration (forward + path-replay adjoint); "
                         "batch: 64 uneven scenes in one batched graph render vs one render call each")
    ap.add_argument("--integrator", default="", choices=["", "path", "direct", "volpath", "mis_volpath"],
                    help="render the workload with another integrator (default: the workload's own)")
    ap.add_argument("--spectral", default="hero4", choices=["hero4", "rgb"])
    ap.add_argument("--estimator", default="corrected", choices=["corrected", "faithful"])
    ap.add_argument("--capacity", type=int, default=0)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-guided", action="store_true",
                    help="many-light: skip the guided estimator's throughput beside the plain one")
    ap.add_argument("--no-count", action="store_true",
                    help="skip the traversal-counting pass (counter runs: every dispatch is then the timed render's)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (HBM counter calibration) and exit")
    ap.add_argument("--save-image", default="")
    ap.add_argument("--relmse", action="store_true", help="measure time-to-relMSE <= 1e-3 instead")
    ap.add_argument("--ref-spp", type=int, default=16384)
    ap.add_argument("--cpu-reference", action="store_true",
                    help="relMSE against the CPU checker's render at --ref-spp (FP64 film sums, 16 host threads)")
    ap.add_argument("--ref-guided", action="store_true",
                    help="with --relmse: render the reference with the guided estimator (its own guide, trained "
                         "online in the reference's first 256 spp): lower variance than a plain reference")
    ap.add_argument("--guided", action="store_true",
                    help="with --relmse: also measure the guided estimator (online-trained guiding network)")
    ap.add_argument("--guide-steps", type=int, default=16,
                    help="Adam steps after each guided wave of the warm-up (waves of 8 spp)")
    ap.add_argument("--guide-warm-spp", default="256",
                    help="guided samples per pixel rendered in 8-spp waves with --guide-steps of training each; "
                         "a comma-separated list measures each")
    ap.add_argument("--guide-late-steps", default="4",
                    help="Adam steps after each later guided wave (waves of --guide-late-wave spp); a "
                         "comma-separated list measures each")
    ap.add_argument("--no-plain", action="store_true",
                    help="with --relmse --guided: skip the plain estimator's run")
    ap.add_argument("--guide-late-wave", default="64",
                    help="spp per later guided wave; a comma list runs one guided measurement per value")
    ap.add_argument("--guide-depth", default="2",
                    help="guide vertices below this path depth only (0: all; 2: the first two scattering "
                         "vertices); a comma-separated list measures each")
    args = ap.parse_args()
    if not args.steps:
        args.steps = DEFAULT_STEPS[args.workload]
    if args.gpus != 1:
        print("bench.py measures one GPU: this build renders on a single MI355X "
              "(multi-GPU sample partitioning is a design note, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from lambda_amd import abi, gpu
    if args.copy_calibration:
        dev = gpu.Device(0)
        ms = dev.copy_bench(args.copy_calibration, 20)
        print(json.dumps({"copy_bytes": args.copy_calibration, "ms": ms,
                          "GBps_read_plus_write": 2 * args.copy_calibration / ms / 1e6}), flush=True)
        dev.close()
        return
    spectral = abi.SPECTRAL_HERO4 if args.spectral == "hero4" else abi.SPECTRAL_RGB
    estimator = abi.ESTIMATOR_CORRECTED if args.estimator == "corrected" else abi.ESTIMATOR_FAITHFUL
    if args.workload == "inverse":
        adjoint_run(args)
        return
    if args.workload == "batch":
        batch_run(args)
        return

    t0 = time.perf_counter()
    sd, cam, cfg = build_workload(args.workload)
    if args.integrator:
        cfg["integrator"] = args.integrator
        cfg["note"] = "workload rendered with the %s integrator" % args.integrator
    log("workload generated in %.1f s (%d triangles)" % (time.perf_counter() - t0, sd.n_tris))
    dev = gpu.Device(0)
    s = gpu.Scene(dev)
    t0 = time.perf_counter()
    keep = sd.push(s, abi.LIGHTS_TREE)
    commit_s = time.perf_counter() - t0
    info = s.info()
    log("scene committed in %.2f s: %s" % (commit_s, info))
    W, H = cfg["width"], cfg["height"]
    film = gpu.Film(dev, W, H)

    def desc(b, e, flags=0, f=None, seed=0, guide=None, steps=0, capacity=None):
        return gpu.render_desc(s, f or film, cam, b, e, max_bounces=cfg["max_bounces"], min_bounces=3,
                               spectral=spectral, estimator=estimator,
                               capacity=args.capacity if capacity is None else capacity, flags=flags,
                               integrator=integrator_id(cfg), seed=seed, guide=guide, guide_train_steps=steps)

    if args.relmse:
        relmse_run(args, sd, cam, cfg, spectral, estimator, desc, film, dev)
        del keep
        film.close()
        s.close()
        dev.close()
        return

    # traversal work per extension ray (untimed counting pass, 1 spp): 4-wide node and triangle
    # records the render's walk fetches, density voxels per path vertex
    if args.no_count:
        nodes4_per_ray = tris4_per_ray = voxels_per_vertex = None
    else:
        cst = gpu.render(desc(10000, 10001, abi.RENDER_COUNT_TRAVERSAL))
        nodes4_per_ray = cst.nodes4_fetched / max(1, cst.extension_rays)
        tris4_per_ray = cst.tris4_fetched / max(1, cst.extension_rays)
        voxels_per_vertex = cst.voxel_steps / max(1, cst.extension_rays)
        film.clear()
        log("counting pass: %.2f 4-wide node records, %.2f triangle records, %.1f voxel steps per extension ray"
            % (nodes4_per_ray, tris4_per_ray, voxels_per_vertex))

    # the path pool of the timed render (as the library would size it by default), allocated by the
    # warm-up so that the timed region measures the steady state, not the pool's growth
    pool = args.capacity or min(W * H * args.steps, 1 << 27)
    if args.warmup:
        gpu.render(desc(0, args.warmup, capacity=pool))
        log("warmup %d spp done" % args.warmup)
    film.clear()
    t0 = time.perf_counter()
    st = gpu.render(desc(args.warmup, args.warmup + args.steps, capacity=pool))
    dt = time.perf_counter() - t0
    samples = W * H * args.steps
    assert st.samples == samples
    value = samples / dt / 1e6
    log("timed %d spp in %.3f s: %.1f Msamples/s; kernel ms extend %.1f shade %.1f shadow %.1f camera %.1f; "
        "rays ext %d shadow %d" % (args.steps, dt, value, st.ms_extend, st.ms_shade, st.ms_shadow, st.ms_camera,
                                   st.extension_rays, st.shadow_rays))

    # Rooflines (HIP-event kernel times on the render stream).
    # extend: algorithmic bytes per ray = 64 B per 4-wide node record + 48 B per triangle record
    #   the walk fetches (counting pass) + 32 B ray read + 16 B hit write.
    # shade: algorithmic bytes per path vertex = 2 x 96 B path state (read + write), plus 4 B per
    #   density-grid voxel visited by the vertex's medium walks.
    # traffic: the HBM bytes per unit that FETCH_SIZE / WRITE_SIZE measured for this kernel and
    #   workload (calibrated against a copy kernel), read from the committed counter summary.
    bytes_per_ray = None if nodes4_per_ray is None else 64.0 * nodes4_per_ray + 48.0 * tris4_per_ray + 48.0
    bytes_per_vertex = None if voxels_per_vertex is None else 192.0 + 4.0 * voxels_per_vertex
    shade_name = {"path": "k_shade", "direct": "k_shade_direct"}.get(cfg["integrator"], "k_shade_vol")
    measured = measured_traffic(args.workload)
    valu = measured_valu_busy(args.workload)
    ext_s = st.ms_extend / 1000.0
    sh_s = st.ms_shade / 1000.0

    def roof(kernel, alg, units, secs, launches, unit):
        r = {"bound": "hbm", "achieved": None if alg is None else round(alg * units / secs / 1e9, 1),
             "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": None, "kernel": kernel, "launches": launches,
             "avg_launch_ms": round(secs * 1000.0 / max(1, launches), 4),
             "alg_bytes_per_unit": None if alg is None else round(alg, 1), "unit_of_work": unit,
             "units_per_launch": round(units / max(1, launches), 1), "units_per_s": round(units / secs, 1)}
        r["frac"] = None if alg is None else round(r["achieved"] / HBM_PEAK_GBS, 4)
        m = measured.get(kernel)
        if m:
            per = m["read"] + m["write"]
            r["traffic"] = round(per * units / max(1, launches), 1)
            r["traffic_per_unit"] = {"read": m["read"], "write": m["write"], "source": m["source"]}
            r["hbm_achieved_GBps"] = round(per * units / secs / 1e9, 1)
        if r["frac"] is not None and r["frac"] > 1.0:
            # the algorithmic bytes are mostly served by L2 and MALL, so HBM does not bound this
            # kernel: report it as VALU-bound with the measured VALU busy share, and its HBM share
            # from the measured traffic
            r["bound"] = "valu"
            r["frac"] = valu.get(kernel)
            r["valu_busy_source"] = "profiles/valu_lds_utilisation.md"
            if "hbm_achieved_GBps" in r:
                r["hbm_frac"] = round(r["hbm_achieved_GBps"] / HBM_PEAK_GBS, 4)
        return r

    ext = roof("k_extend", bytes_per_ray, st.extension_rays, ext_s, st.launches_extend, "extension ray")
    shd = roof(shade_name, bytes_per_vertex, st.extension_rays, sh_s, st.launches_shade, "path vertex")
    roofline = dict(shd if st.ms_shade >= st.ms_extend else ext)
    roofline["other_kernels"] = [shd if roofline["kernel"] == "k_extend" else ext]
    if args.workload != "many-light":
        cfg["note"] = "secondary workload; the headline line is --workload many-light"
    if args.save_image:
        rgba = film.read_rgba()
        import numpy as np
        np.save(args.save_image, rgba)

    guided = None
    if args.workload == "many-light" and not args.no_guided and not args.integrator and args.estimator == "corrected":
        log("guided estimator ...")
        guided = guided_throughput(desc, film, W, H, dev)
        guided["time_to_relmse"] = measured_time_to_relmse()
        log("guided: %s" % guided)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(sd, cam, cfg, spectral, estimator, args.cpu_budget)
        log("cpu baseline: %s" % cpu)

    cfg.update({"spp": args.steps, "spectral": args.spectral, "estimator": args.estimator,
                "light_strategy": "tree", "bvh_nodes": info["bvh_nodes"], "tree_nodes": info["tree_nodes"],
                "scene_commit_s": round(commit_s, 2)})
    line = {"metric": "Msamples/s (plain wavefront render, 1 MI355X)", "value": round(value, 3),
            "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(dt * 1000.0 / args.steps, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic (procedural scene, no assets)",
            "config": cfg, "roofline": roofline, "cpu_baseline": cpu, "guided": guided,
            "stats": {"extension_rays": st.extension_rays, "shadow_rays": st.shadow_rays,
                      "ms_extend": round(st.ms_extend, 2), "ms_shade": round(st.ms_shade, 2),
                      "ms_shadow": round(st.ms_shadow, 2), "ms_camera": round(st.ms_camera, 2),
                      "waves": st.waves}}
    print(json.dumps(line), flush=True)
    del keep
    film.close()
    s.close()
    dev.close()


def batch_run(args):
    """64 seeded scenes of uneven size (about 10k to 2M triangles, 16 to 4096 emissive triangles,
    64^2 to 1024^2 pixels, one in four with a grid-medium cloud and the volumetric integrator),
    `steps` spp each: one lgpu_render_batch call (autotuned; the tuning trials are untimed) against
    one lgpu_render call per scene. `value` is the batch's throughput."""
    from lambda_amd import abi, gpu, scene
    t0 = time.perf_counter()
    dev = gpu.Device(0)
    jobs = []
    for k in range(64):
        sd, cam, W, H, integ, maxb = scene.batch_job(k)
        s = gpu.Scene(dev)
        keep = sd.push(s, abi.LIGHTS_TREE)
        jobs.append(dict(sd_tris=sd.n_tris, s=s, keep=keep, cam=cam, W=W, H=H, integ=integ, maxb=maxb,
                         film=gpu.Film(dev, W, H)))
        if k % 8 == 7:
            log("batch workload: %d scenes generated and committed (%.1f s)" % (k + 1, time.perf_counter() - t0))
    spp = args.steps

    def desc(j, seed):
        return gpu.render_desc(j["s"], j["film"], j["cam"], 0, spp, max_bounces=j["maxb"], min_bounces=3,
                               spectral=abi.SPECTRAL_HERO4, estimator=abi.ESTIMATOR_CORRECTED, seed=seed,
                               integrator=j["integ"])
    samples = sum(j["W"] * j["H"] for j in jobs) * spp
    # warm-up and autotuning (cached for the timed call)
    for j in jobs:
        j["film"].clear()
    st = gpu.render_batch([desc(j, 1) for j in jobs], autotune=True)
    log("autotune: capacity %d, %d branches (trials %.0f ms)" % (st.path_capacity, st.branches, st.ms_autotune))
    for j in jobs:
        gpu.render(desc(j, 1))
    for j in jobs:
        j["film"].clear()
    t1 = time.perf_counter()
    for j in jobs:
        gpu.render(desc(j, 2))
    seq_s = time.perf_counter() - t1
    log("one render call per scene: %.3f s (%.1f Msamples/s)" % (seq_s, samples / seq_s / 1e6))
    for j in jobs:
        j["film"].clear()
    t1 = time.perf_counter()
    st = gpu.render_batch([desc(j, 2) for j in jobs], autotune=True)
    batch_s = time.perf_counter() - t1
    value = samples / batch_s / 1e6
    log("batched: %.3f s (%.1f Msamples/s), %d chunks, %d graph nodes, capture %.0f ms"
        % (batch_s, value, st.chunks, st.graph_nodes, st.ms_capture))
    line = {"metric": "Msamples/s (64-scene batched render, 1 MI355X)", "value": round(value, 3),
            "unit": "Msamples/s", "n_gpus": 1, "steps": spp, "warmup": 1, "ms_per_step": round(batch_s * 1000 / spp, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (procedural scenes, no assets)",
            "config": {"workload": "64 seeded scenes of uneven size (many-light field patches, about 10k to 2M "
                                   "triangles, 16 to 4096 emissive triangles, 64^2 to 1024^2, one in four with a "
                                   "grid-medium cloud), hero spectra, corrected estimator, depth 8",
                       "triangles": sum(j["sd_tris"] for j in jobs), "pixels": sum(j["W"] * j["H"] for j in jobs),
                       "spp": spp, "note": "secondary workload; the headline line is --workload many-light"},
            "stats": {"sequential_s": round(seq_s, 4), "sequential_Msamples_per_s": round(samples / seq_s / 1e6, 3),
                      "speedup_vs_sequential": round(seq_s / batch_s, 3), "chunks": st.chunks,
                      "graph_nodes": st.graph_nodes, "path_capacity": st.path_capacity, "branches": st.branches,
                      "ms_capture": round(st.ms_capture, 1)}}
    print(json.dumps(line), flush=True)
    for j in jobs:
        j["film"].close()
        j["s"].close()
    dev.close()


def adjoint_run(args):
    """Differentiable inverse rendering (path-replay adjoint): the target is rendered at 1024 spp
    from perturbed parameters (albedo + U[-0.2, 0.2], roughness x U[0.5, 2], density x U[0.5, 2]);
    each timed step is one iteration of the L2 image loss at 64 spp: forward render, then the
    adjoint render with d_image = 2 (I - T) / N. `value` is the adjoint render's throughput."""
    import numpy as np
    from lambda_amd import abi, gpu, scene
    W = H = 512
    spp, bounces = 64, 6
    t0 = time.perf_counter()
    sd, ids = scene.inverse_scene()
    tex = [o[1] for o in sd.ops if o[0] == "texture"]
    dens = [o[1]["density"] for o in sd.ops if o[0] == "medium"][0]
    rng = np.random.default_rng(1)
    albedo_t = np.clip(tex[ids["albedo"]] + rng.uniform(-0.2, 0.2, tex[ids["albedo"]].shape), 0.02, 0.98)
    rough_t = np.clip(tex[ids["roughness"]] * rng.uniform(0.5, 2.0, tex[ids["roughness"]].shape), 0.05, 1.0)
    dens_t = dens * rng.uniform(0.5, 2.0, dens.shape)
    cam = scene.inverse_camera(W, H)
    log("inverse workload generated in %.1f s (%d triangles)" % (time.perf_counter() - t0, sd.n_tris))
    dev = gpu.Device(0)
    s = gpu.Scene(dev)
    keep = sd.push(s, abi.LIGHTS_TREE)
    film = gpu.Film(dev, W, H)
    n_tex = sum(t.size for t in tex)

    def desc(e, seed):
        return gpu.render_desc(s, film, cam, 0, e, max_bounces=bounces, min_bounces=2, spectral=abi.SPECTRAL_RGB,
                               estimator=abi.ESTIMATOR_CORRECTED, seed=seed, integrator=abi.INTEGRATOR_VOLPATH)

    def image(e, seed):
        film.clear()
        gpu.render(desc(e, seed))
        rgb, cnt = film.read()
        return (rgb / cnt[..., None]).astype(np.float32)

    def set_params(a, r, d):
        s.update_texture(ids["albedo"], a.reshape(-1))
        s.update_texture(ids["roughness"], r.reshape(-1))
        s.update_density(ids["medium"], d.reshape(-1))

    set_params(albedo_t, rough_t, dens_t)
    t0 = time.perf_counter()
    target = image(1024, 99)
    log("target: 1024 spp in %.2f s" % (time.perf_counter() - t0))
    set_params(tex[ids["albedo"]], tex[ids["roughness"]], dens)

    def iteration(it):
        img = image(spp, 2 * it)
        dimg = 2.0 * (img - target) / img.size
        film.clear()
        g = gpu.render_adjoint(desc(spp, 2 * it + 1), dimg, n_tex, dens.size, 0)
        return float(((img - target) ** 2).mean()), g

    for it in range(args.warmup):
        iteration(it)
    times, adj_ms, st = [], [], None
    loss0 = None
    for it in range(args.steps):
        t0 = time.perf_counter()
        loss, (gt, gd, _, st) = iteration(100 + it)
        times.append(time.perf_counter() - t0)
        adj_ms.append(st.ms_total)
        loss0 = loss if loss0 is None else loss0
    dt = sum(times) / len(times)
    adj_s = sum(adj_ms) / len(adj_ms) / 1000.0
    samples = W * H * spp
    value = samples / adj_s / 1e6
    log("iteration %.1f ms (adjoint render %.1f ms: %.1f Msamples/s, primal pass + replay); loss %.4e"
        % (dt * 1000, adj_s * 1000, value, loss0))
    line = {"metric": "Msamples/s (path-replay adjoint render, 1 MI355X)", "value": round(value, 3),
            "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(dt * 1000.0, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic (procedural scene, no assets)",
            "config": {"workload": "differentiable inverse rendering: 50k-triangle patch of the many-light field, "
                                   "512^2 FP32 albedo and roughness textures, 64^3 density grid, 512x512, "
                                   "L2 image loss, adjoint at 64 spp per iteration (RGB, volumetric path "
                                   "integrator, depth 6)", "triangles": sd.n_tris, "width": W, "height": H,
                       "spp": spp, "note": "secondary workload; the headline line is --workload many-light"},
            "stats": {"ms_adjoint_render": round(adj_s * 1000, 2), "ms_shade": round(st.ms_shade, 2),
                      "ms_extend": round(st.ms_extend, 2), "extension_rays": st.extension_rays,
                      "grad_nonzero": {"albedo_roughness": int(np.count_nonzero(gt)), "density": int(np.count_nonzero(gd))}}}
    print(json.dumps(line), flush=True)
    del keep
    film.close()
    s.close()
    dev.close()


def relmse_run(args, sd, cam, cfg, spectral, estimator, desc, film, dev):
    """time-to-relMSE <= 1e-3: progressive rendering (2^k spp checkpoints, resumable sample
    ranges, cumulative render wall time) against an FP64 reference at --ref-spp rendered with
    seed 1; the reference is checked against the CPU checker on a 64x64 centre crop."""
    import numpy as np
    from lambda_amd import gpu
    W, H = cfg["width"], cfg["height"]
    ref = gpu.Film(dev, W, H)
    chunk = 256
    t0 = time.perf_counter()
    rguide = gpu.Guide(dev, seed=101) if args.ref_guided else None
    b = 0
    cpu_ref = None
    if args.cpu_reference:
        import oracle
        from lambda_amd import abi
        o = oracle.OracleScene()
        okeep = sd.push(o, abi.LIGHTS_TREE)
        d = abi.RenderDesc()
        d.camera = cam
        d.integrator, d.max_bounces, d.min_bounces = integrator_id(cfg), cfg["max_bounces"], 3
        d.spectral_mode, d.estimator_mode, d.seed = spectral, estimator, 1
        acc, cnt = np.zeros((H, W, 3)), np.zeros((H, W))
        threads = max(1, min(os.cpu_count() or 1, 16))
        per = 256
        while b < args.ref_spp:
            d.spp_begin, d.spp_end = b, min(args.ref_spp, b + per)
            crgb, ccnt = o.render(d, W, H, threads=threads)
            acc += crgb
            cnt += ccnt
            b = d.spp_end
            log("CPU reference %d / %d spp (%.0f s)" % (b, args.ref_spp, time.perf_counter() - t0))
        cpu_ref = acc / cnt[..., None]
        del okeep
        b = args.ref_spp  # no GPU reference
    while b < args.ref_spp:
        if rguide is not None:
            per = 8 if b < 256 else 64
            e = min(args.ref_spp, b + per)
            gpu.render(desc(b, e, f=ref, seed=1, guide=rguide, steps=16 if b < 256 else 2, capacity=W * H * per))
        else:
            e = min(args.ref_spp, b + chunk)
            gpu.render(desc(b, e, f=ref, seed=1))
        b = e
        if b % 256 == 0 or b == args.ref_spp:
            log("reference %d / %d spp (%.0f s)" % (b, args.ref_spp, time.perf_counter() - t0))
    if rguide is not None:
        rguide.close()
    ref_s = time.perf_counter() - t0
    if cpu_ref is not None:
        r = cpu_ref
    else:
        rsum, rcnt = ref.read()
        r = rsum / rcnt[..., None]
    # untimed: the path pool the plain checkpoints' largest call needs (as the timed bench's warm-up)
    gpu.render(desc(0, 1, capacity=min(W * H * max(1, args.ref_spp // 8), 1 << 27)))

    def progressive(guide=None, late_steps=4, late_wave=64, warm_spp=256):
        """Checkpoints of one progressive render (cumulative wall time of the render calls,
        training included for the guided estimator); stops at relMSE <= 1e-3."""
        film.clear()
        checkpoints, hit = [], None
        spent, done = 0.0, 0
        kern = {"ms_extend": 0.0, "ms_shade": 0.0, "ms_shadow": 0.0, "ms_guide_infer": 0.0, "ms_guide_train": 0.0}

        def tally(st):
            for key in kern:
                kern[key] += getattr(st, key)

        # checkpoints at 2^k, 1.25, 1.5 and 1.75 x 2^k spp (integers), up to half the reference's
        targets = sorted({(m << j) // 4 for j in range(31) for m in (4, 5, 6, 7)
                          if (m << j) % 4 == 0 and 0 < (m << j) // 4 <= args.ref_spp // 2})
        for target in targets:
            t = time.perf_counter()
            if guide is None:
                tally(gpu.render(desc(done, target)))
            else:
                # online training between waves: 8-spp waves with --guide-steps each during the
                # warm-up, then larger waves with --guide-late-steps
                b = done
                while b < target:
                    warm = b < warm_spp
                    per = min(target - b, 8 if warm else late_wave)
                    steps = args.guide_steps if warm else late_steps
                    tally(gpu.render(desc(b, b + per, guide=guide, steps=steps, capacity=W * H * per)))
                    b += per
            spent += time.perf_counter() - t
            done = target
            xs, xc = film.read()
            x = xs / xc[..., None]
            rel = float(np.mean((x - r) ** 2 / (r * r + 0.01)))
            checkpoints.append({"spp": target, "seconds": round(spent, 4), "relmse": rel,
                                "msamples_per_s": round(W * H * target / spent / 1e6, 1),
                                "kernel_ms": {key: round(v, 1) for key, v in kern.items()}})
            log("%s spp %d: %.3f s, relMSE %.3e; ms %s" % ("guided" if guide else "plain", target, spent, rel,
                                                            {key: round(v, 1) for key, v in kern.items()}))
            if rel <= 1e-3:
                hit = dict(checkpoints[-1])
                if len(checkpoints) > 1:
                    # where the curve crosses 1e-3 between the last two checkpoints (log-log
                    # interpolation of relMSE against spp, time linear in spp between them)
                    a, b = checkpoints[-2], checkpoints[-1]
                    if a["relmse"] > 1e-3 > 0 and b["relmse"] > 0:
                        f = math.log(a["relmse"] / 1e-3) / math.log(a["relmse"] / b["relmse"])
                        spp_x = a["spp"] * (b["spp"] / a["spp"]) ** f
                        hit["interpolated_spp"] = round(spp_x, 1)
                        hit["interpolated_s"] = round(a["seconds"] + (b["seconds"] - a["seconds"]) *
                                                      (spp_x - a["spp"]) / (b["spp"] - a["spp"]), 3)
                break
        return checkpoints, hit

    checkpoints, hit = ([], None) if (args.guided and args.no_plain) else progressive()
    guided = None
    if args.guided:
        guided = []
        for depth in [int(x) for x in args.guide_depth.split(",")]:
            for late, wave, wspp in [(int(x), int(y), int(z)) for x in args.guide_late_steps.split(",")
                                     for y in args.guide_late_wave.split(",") for z in args.guide_warm_spp.split(",")]:
                g = gpu.Guide(dev, seed=7, max_depth=depth)
                gck, ghit = progressive(g, late, wave, wspp)
                guided.append({"value": ghit["seconds"] if ghit else None,
                               "spp_at_target": ghit["spp"] if ghit else None,
                               "interpolated_s": ghit.get("interpolated_s") if ghit else None,
                               "checkpoints": gck, "guide": g.stats(),
                               "schedule": {"warm_spp": wspp, "warm_wave_spp": 8,
                                            "warm_steps": args.guide_steps, "late_wave_spp": wave,
                                            "late_steps": late, "guided_depth": depth or None}})
                g.close()
    # CPU validation of the reference on a 64x64 centre crop
    valid = None
    if not args.no_cpu and cpu_ref is None:
        import oracle
        from lambda_amd import abi
        o = oracle.OracleScene()
        keep = sd.push(o, abi.LIGHTS_TREE)
        cw = 64
        win = ((W - cw) // 2, (H - cw) // 2, cw, cw)
        d = abi.RenderDesc()
        d.camera = cam
        d.integrator, d.max_bounces, d.min_bounces = integrator_id(cfg), cfg["max_bounces"], 3
        d.spectral_mode, d.estimator_mode, d.seed = spectral, estimator, 2
        d.spp_begin, d.spp_end = 0, 1024
        t = time.perf_counter()
        crgb, ccnt = o.render(d, W, H, window=win, threads=max(1, min(os.cpu_count() or 1, 16)))
        c = crgb / ccnt[..., None]
        rc = r[win[1]:win[1] + cw, win[0]:win[0] + cw]
        valid = {"crop": "64x64 centre", "cpu_spp": 1024, "cpu_s": round(time.perf_counter() - t, 1),
                 "cpu_mean_rgb": [round(float(v), 6) for v in c.mean(axis=(0, 1))],
                 "gpu_ref_mean_rgb": [round(float(v), 6) for v in rc.mean(axis=(0, 1))],
                 "relmse_cpu_vs_ref": float(np.mean((c - rc) ** 2 / (rc * rc + 0.01)))}
        log("reference check: %s" % valid)
        del keep
    line = {"metric": "time-to-relMSE<=1e-3 (plain wavefront, 1 MI355X)",
            "value": hit["seconds"] if hit else None, "unit": "s", "n_gpus": 1, "higher_is_better": False,
            "spp_at_target": hit["spp"] if hit else None,
            "interpolated_s": hit.get("interpolated_s") if hit else None, "dtype": "f32", "data": "synthetic (procedural scene)",
            "config": dict(cfg, spectral=args.spectral, estimator=args.estimator, reference_spp=args.ref_spp,
                           reference_seed=1, reference_s=round(ref_s, 1),
                           reference_estimator="guided" if args.ref_guided else "plain",
                           reference_renderer="CPU checker (oracle/, FP64 film sums, %d host threads)"
                                              % max(1, min(os.cpu_count() or 1, 16)) if cpu_ref is not None
                                              else "this GPU renderer"),
            "checkpoints": checkpoints, "reference_check": valid, "guided": guided}
    print(json.dumps(line), flush=True)
    ref.close()


if __name__ == "__main__":
    main()
