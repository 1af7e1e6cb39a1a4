"""Throughput benchmark of the MI355X path tracer (one JSON line on stdout).

A step is one dispatch of the plain tracer over one frame of the BVH-stress workload:
65,536 random triangles + one emissive sphere, 320x180, 4 spp, 4 bounces, median-split
BVH (<= 4 primitives per leaf), black background, frame seed 0 (the reference's static
benchmark mode, path_tracer.rs:602-609), per-pixel PCG streams as in the reference
kernel.  Scene data is resident in HBM before timing; value = paths traced per second.
The default schedule is the speculative one (two octets per pixel: the pixel's next
sample traced beside the current one from a foreseen stream state), whose image equals
the sequential kernels' bit for bit.

roofline.achieved uses algorithmic bytes per launch: the reference-layout bytes the
reference traversal reads for this frame (48 B per node visited, 8 B per primitive
reference, 64/32 B per triangle/sphere test, 48+80 B per closest hit, 96 B per light
record scanned, 16 B per output pixel), counted by the oracle on the same seeds, divided
by the kernel's mean duration from HIP events around each launch.  roofline.nearest_first
prices the node visits of a nearer-child-first walk (the GPU's ordered-traversal
counters on the same frame), the order the 4-wide kernels visit in.  parity: the timed
frame against the oracle's image of the same frame (in the cpu_baseline leg, outside the
timed region); a frame beyond tolerance makes the run exit with status 1.

equal_time compares the guided tracer with the plain tracer on the indirect-lit terrain
(262,144 triangles, 1280x720, 8 bounces): a plain reference at --reference-spp, the
plain tracer at 64 and 256 spp, and the guided tracer at 256 spp trained online from
scratch (all training inside the timed region), five times with independent guides
and sample streams.  From the two plain points, relRMSE^2 = C / spp + E (E = the
reference's own noise); the plain time to reach a guided run's error is
T_plain(256) * (C / (relRMSE_guided^2 - E)) / 256, and time_ratio is the guided time over
that, averaged over the five runs.  Ablations with the estimator held fixed: the same
ratio against the plain estimator without the environment sample (what the guided default
drops at all but a path's last vertex), the mixture estimator around the untrained network
(relRMSE at 256 spp), and the guided tracer keeping the environment sample.  mfma: the
network's inference and training launches timed by event pairs in a separate default run,
their algorithmic FLOP/s against the 2.5 PFLOP/s dense BF16 peak.

reference_suite times the reference's own benchmark scenes (scenes/benchmarks/001-007)
as its benchmark mode does; batched renders 16 scenes x 4 frames in one persistent
launch beside one launch per frame; scene_update times a frame of moved geometry on the
headline scene, rebuilt and uploaded as the reference does against refitted on the GPU.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    """Progress on stderr (stdout carries the one JSON line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(c, pixels, accumulated=False):
    b = (48 * c["nodes_visited"] + 8 * c["primref_tests"] + 64 * c["triangle_tests"] +
         32 * c["sphere_tests"] + 48 * c["closest_hits"] + 80 * c["triangle_hits"] +
         96 * c["lights_scanned"] + 4 * 16 * c["env_lookups"] + 16 * 19 * c["env_samples"] +
         16 * pixels)
    if accumulated:
        b += 16 * pixels
    return b


def _native_oracle():
    """Builds the oracle with -O3 -march=native for this host into a temporary directory
    (the in-tree build is portable); returns (path, flags) or (None, reason)."""
    import subprocess
    import tempfile
    src = os.path.join(ROOT, "oracle", "oracle.cpp")
    out = os.path.join(tempfile.mkdtemp(prefix="oracle_native_"), "liboracle.so")
    flags = ["-O3", "-march=native", "-std=c++17", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared"]
    try:
        subprocess.run(["g++"] + flags + ["-o", out, src], check=True, capture_output=True, timeout=240)
    except Exception as e:  # no compiler on the host: the portable build is used
        return None, f"native build failed ({type(e).__name__})"
    return out, " ".join(flags)


def _cpu_identity():
    info = {}
    try:
        import subprocess
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "CPU(s)", "Thread(s) per core", "Core(s) per socket", "Socket(s)",
                             "CPU family", "Model", "Stepping", "CPU max MHz", "L3 cache"):
                info[k.strip()] = v.strip()
    except Exception:
        pass
    if "Model name" not in info:
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        info["Model name"] = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    return info


def cpu_baseline(sc, gpu_image, budget_s=8.0):
    """The oracle (C++ restatement, pcg mode, built -O3 -march=native for this host) on
    this host's cores: full frames of the headline workload until the budget is spent,
    and one frame on one thread.  Its first frame is the headline frame (frame seed 0):
    the GPU image is compared with it here, outside the timed region (|gpu - oracle| <=
    1e-4 (1 + |oracle|) per channel on >= 99.5 % of pixels).  Also returns its counters
    for the algorithmic-byte figure."""
    import numpy as np
    import oracle
    native, flags = _native_oracle()
    if native:
        oracle.LIB_PATH = native
    oc = oracle.OracleScene(sc)
    threads = max(1, min(os.cpu_count() or 1, 16))
    ref, counters = oc.render(threads=threads, counters=True)
    close = np.abs(gpu_image - ref) <= 1e-4 * (1.0 + np.abs(ref))
    parity = {"within_tolerance": round(float(np.mean(np.all(close, axis=-1))), 6),
              "bit_identical": round(float(np.mean(np.all(gpu_image == ref, axis=-1))), 6),
              "tolerance": "|gpu - oracle| <= 1e-4 (1 + |oracle|) per channel, >= 99.5 % of pixels",
              "image": "headline frame, frame seed 0"}
    parity["ok"] = parity["within_tolerance"] >= 0.995
    frames, t0 = 0, time.perf_counter()
    while True:
        oc.render(threads=threads)
        frames += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    # single-thread figure on one full frame (the reference CPU renderer is single-threaded)
    t1 = time.perf_counter()
    oc.render(threads=1)
    dt1 = time.perf_counter() - t1
    oc.close()
    paths = sc.paths_per_frame
    ident = _cpu_identity()
    return {
        "value": round(frames * paths / dt / 1e6, 4), "unit": "Mpaths/s", "cores": threads,
        "kind": "port",
        "sample": f"{frames} full frames of the same workload (oracle, pcg mode, OpenMP over pixels)",
        "single_thread_mpaths_s": round(sc.paths_per_frame / dt1 / 1e6, 4),
        "build": flags if native else "in-tree oracle/liboracle.so (" + flags + ")",
        "cpu_model": ident.get("Model name", ""), "lscpu": ident, "nproc": os.cpu_count(),
    }, counters, parity


def scene_update_leg(ctx, sc, frames=10):
    """Moved geometry per frame on the headline scene (every other triangle rotated about
    the cloud's centre, the light shifted): the reference's way (host repack with a new
    median-split BVH + upload of every array, path_tracer.rs:639-720) against a device
    refit of the uploaded tree (host repack without BVH + refit), median ms over
    `frames` frames, each followed by one render of the moved frame."""
    import dataclasses
    import numpy as np
    from toaster_amd import GpuScene, RenderOptions, _abi as A
    from toaster_amd.scene import pack
    opts = RenderOptions(traversal=A.TH_TRAVERSE_WIDE, schedule=A.TH_SCHEDULE_SPECULATIVE)
    v0 = sc.triangles["v"].astype(np.float64)
    c = v0.reshape(-1, 3).mean(axis=0)
    sel = np.arange(len(v0)) % 2 == 0

    def frame(k):
        a = 0.02 * k
        R = np.array([[np.cos(a), 0.0, np.sin(a)], [0.0, 1.0, 0.0], [-np.sin(a), 0.0, np.cos(a)]])
        v = v0.copy()
        v[sel] = (v[sel] - c) @ R.T + c
        tris = sc.triangles.copy()
        tris["v"] = v.astype(np.float32)
        sph = sc.spheres.copy()
        sph["center"] = sph["center"] + np.float32(0.01 * k)
        return dataclasses.replace(sc, triangles=tris, spheres=sph)

    moved = [frame(k) for k in range(1, frames + 1)]
    out = {}
    everything = (A.TH_CHANGE_CAMERA | A.TH_CHANGE_SPHERES | A.TH_CHANGE_TRIANGLES | A.TH_CHANGE_LIGHTS
                  | A.TH_CHANGE_BVH | A.TH_CHANGE_MATERIALS)
    refit = (A.TH_CHANGE_CAMERA | A.TH_CHANGE_SPHERES | A.TH_CHANGE_TRIANGLES | A.TH_CHANGE_LIGHTS
             | A.TH_CHANGE_REFIT)
    for mode in ("rebuild", "refit"):
        gs = GpuScene.from_scene(ctx, sc)
        upd, dev, ren = [], [], []
        for m in moved:
            ctx.synchronize()
            t0 = time.perf_counter()
            packed = pack(m, bvh=(mode == "rebuild"))  # the host's share: records (+ tree)
            t1 = time.perf_counter()
            gs.update(packed, everything if mode == "rebuild" else refit)  # upload (+ device refit)
            ctx.synchronize()
            t2 = time.perf_counter()
            gs.render(options=opts, frame_index=0, accumulated_samples=0)
            ctx.synchronize()
            t3 = time.perf_counter()
            upd.append((t2 - t0) * 1e3)
            dev.append((t2 - t1) * 1e3)
            ren.append((t3 - t2) * 1e3)
        out[mode] = {"update_ms": round(float(np.median(upd)), 3),
                     "upload_and_device_ms": round(float(np.median(dev)), 3),
                     "render_ms": round(float(np.median(ren)), 4)}
        if mode == "refit":  # summed node-box area of the refitted tree over the built one's
            out[mode]["area_ratio_after"] = round(gs.refit_quality(), 4)
        gs.close()
    out["workload"] = (f"{len(v0)} triangles, every other one rotated per frame, {frames} frames, "
                       "render with the speculative kernel")
    return out


def reference_suite(warmup=5, runs=21):
    """The reference's benchmark scenes (scenes/benchmarks/001-007) as its benchmark mode
    runs them (frame seed 0, median over measured frames): dispatch_wait = render +
    wait, the span of the published `dispatch_wait_ms` (dispatch.rs:44-57: submit, the
    GPU-side copy of the output, poll); readback = the copy of the linear-HDR output to
    host memory, which the reference reports separately."""
    import numpy as np
    from toaster_amd import GpuContext, GpuScene, RenderOptions, _abi as A
    from toaster_amd.loader import PUBLISHED_DISPATCH_WAIT_MS, load_scene, suite_paths
    ctx = GpuContext(0)
    out = {}
    for path in suite_paths():
        sc = load_scene(path)
        gs = GpuScene.from_scene(ctx, sc)
        row = {}
        # reference RNG: the 4-lane, octet and speculative schedules give the same image;
        # all are timed and the fastest is reported (short traversals favour the 4-lane one)
        for mode, opts in (("reference_rng_octet", RenderOptions(traversal=A.TH_TRAVERSE_WIDE,
                                                                 schedule=A.TH_SCHEDULE_OCTET)),
                           ("reference_rng_speculative", RenderOptions(traversal=A.TH_TRAVERSE_WIDE,
                                                                       schedule=A.TH_SCHEDULE_SPECULATIVE)),
                           ("reference_rng_quad", RenderOptions(traversal=A.TH_TRAVERSE_WIDE)),
                           ("per_sample_rng", RenderOptions(traversal=A.TH_TRAVERSE_WIDE,
                                                            rng_mode=A.TH_RNG_PER_SAMPLE))):
            times, kms, rb = [], [], []
            for i in range(warmup + runs):
                ctx.synchronize()
                ctx.reset_stats()
                t = time.perf_counter()
                gs.render(options=opts, frame_index=0, accumulated_samples=0)
                ctx.synchronize()
                t1 = time.perf_counter()
                gs.readback()
                t2 = time.perf_counter()
                if i >= warmup:
                    times.append((t1 - t) * 1e3)
                    rb.append((t2 - t1) * 1e3)
                    kms.append(ctx.stats()["total_kernel_ms"])
            med = float(np.median(times))
            row[mode] = {"dispatch_wait_ms": round(med, 4), "kernel_ms": round(float(np.median(kms)), 4),
                         "readback_ms": round(float(np.median(rb)), 4),
                         "mpaths_s": round(sc.paths_per_frame / med / 1e3, 2)}
        timed = {"octet": row.pop("reference_rng_octet"), "speculative": row.pop("reference_rng_speculative"),
                 "static": row.pop("reference_rng_quad")}
        sched = min(timed, key=lambda k: timed[k]["dispatch_wait_ms"])
        row["reference_rng"] = dict(timed[sched], schedule=sched,
                                    other_schedules_dispatch_wait_ms={k: v["dispatch_wait_ms"] for k, v in timed.items()
                                                                      if k != sched})
        pub = PUBLISHED_DISPATCH_WAIT_MS[sc.name]
        row["rtx2080ti_dispatch_wait_ms"] = pub
        row["speedup_vs_rtx2080ti"] = round(pub / row["reference_rng"]["dispatch_wait_ms"], 3)
        out[sc.name] = row
        gs.close()
    ctx.close()
    return out


MLP_MACS = 48 * 64 + 64 * 64 + 64 * 64 + 64 * 40  # 13,824 per evaluation
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA


def equal_time_guided(reference_spp):
    """Equal-time relRMSE of guided vs plain tracing on the terrain workload, with
    ablations that hold the estimator fixed: the plain estimator without the environment
    sample (what the guided default drops), the same mixture estimator around the
    untrained network, and the guided tracer keeping the environment sample; plus the
    network's MFMA throughput from a separately timed run."""
    from toaster_amd import GpuContext, GpuScene, RenderOptions, _abi as A, scenes
    from toaster_amd.guide import GUIDED_DEFAULTS, Guide, rel_rmse, render_guided
    sc = scenes.terrain(1280, 720, samples=16, max_bounces=8)
    ctx = GpuContext(0)
    gs = GpuScene.from_scene(ctx, sc)
    px = sc.width * sc.height

    def timed(fn):
        ctx.synchronize()
        t = time.perf_counter()
        r = fn()
        ctx.synchronize()
        return time.perf_counter() - t, r

    # the plain estimator (per-sample streams) through each plain kernel; the fastest is
    # the baseline and renders the reference
    g0 = Guide(ctx, seed=1)
    plain = {
        "quad_megakernel": lambda n, f: gs.render(samples=n, frame_index=f, accumulated_samples=0,
                                                  options=RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE,
                                                                        traversal=A.TH_TRAVERSE_WIDE)),
        "wavefront": lambda n, f: gs.render(samples=n, frame_index=f, accumulated_samples=0,
                                            options=RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE,
                                                                  traversal=A.TH_TRAVERSE_WIDE,
                                                                  kernel=A.TH_KERNEL_WAVEFRONT)),
        "wavefront_1spp_passes": lambda n, f: (setattr(gs, "frame_index", f),
                                               render_guided(gs, g0, n, alpha=0.0, train=False,
                                                             progressive=False, warmup_steps=0)),
    }
    speed = {}
    for name, fn in plain.items():
        fn(4, 1)
        speed[name] = timed(lambda: fn(16, 2))[0] / 16
    best = min(speed, key=speed.get)
    fn = plain[best]
    chunk = 256
    t_ref = 0.0
    log(f"equal_time: plain kernel {best}; reference {reference_spp} spp")
    for c in range(max(1, reference_spp // chunk)):  # progressive chunks into one average
        if c % 16 == 0:
            log(f"equal_time: reference {c * chunk}/{reference_spp} spp")
        if best == "wavefront_1spp_passes":
            gs.frame_index = 1000 + c
            gs.accumulated_samples = c * chunk
            t_ref += timed(lambda: render_guided(gs, g0, chunk, alpha=0.0, train=False, warmup_steps=0))[0]
        else:
            t_ref += timed(lambda: gs.render(samples=chunk, frame_index=1000 + c, accumulated_samples=c * chunk,
                                             options=RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE,
                                                                   traversal=A.TH_TRAVERSE_WIDE,
                                                                   kernel=(A.TH_KERNEL_WAVEFRONT if best == "wavefront"
                                                                           else A.TH_KERNEL_MEGA))))[0]
    ref = gs.readback()

    def curve(render):
        """Two points (64, 256 spp) of an estimator: relRMSE^2 = C / spp + E, where E is
        the reference's own noise plus any difference in what the estimator converges to."""
        pts = {}
        for n in (64, 256):
            dt, _ = timed(lambda: render(n, 7 + n))
            pts[n] = (dt, rel_rmse(gs.readback(), ref))
        (t64, r64), (t256, r256) = pts[64], pts[256]
        C = (r64 ** 2 - r256 ** 2) / (1 / 64 - 1 / 256)
        E = max(r256 ** 2 - C / 256, 0.0)
        return {"t256": t256, "r256": r256, "r64": r64, "C": C, "E": E}

    def t_equal(cv, rel):
        """Time the estimator of curve `cv` needs to reach relRMSE `rel`."""
        return cv["t256"] * (cv["C"] / max(rel ** 2 - cv["E"], 1e-12)) / 256

    log("equal_time: plain curves, guided runs, ablations")
    base = curve(fn)
    no_nee = curve(lambda n, f: (setattr(gs, "frame_index", f),
                                 render_guided(gs, g0, n, alpha=0.0, train=False, env_nee=False, progressive=False,
                                               warmup_steps=0)))
    runs = []
    for seed in (1, 2, 3, 4, 5):  # independent guides and sample streams: the ratio's spread
        g = Guide(ctx, seed=seed)
        gs.frame_index = 2 + seed
        dtg, st = timed(lambda: render_guided(gs, g, 256, progressive=False, **GUIDED_DEFAULTS))
        relg = rel_rmse(gs.readback(), ref)
        t_eq, t_eq_nn = t_equal(base, relg), t_equal(no_nee, relg)
        runs.append({"seconds": round(dtg, 4), "rel_rmse": round(relg, 5),
                     "plain_seconds_to_guided_error": round(t_eq, 4), "time_ratio": round(dtg / t_eq, 4),
                     "time_ratio_vs_plain_without_env_sample": round(dtg / t_eq_nn, 4),
                     "train_steps": st["train_steps"], "records": st["records"]})
        g.close()
    # the same mixture estimator around the untrained network (no training, alpha from the
    # first pass): what the trained weights add
    gu = Guide(ctx, seed=1)
    kw_untrained = dict(GUIDED_DEFAULTS, alpha_ramp=0, warmup_steps=0)
    gs.frame_index = 3
    dtu, _ = timed(lambda: render_guided(gs, gu, 256, train=False, progressive=False, alpha_from_start=True,
                                         **kw_untrained))
    relu = rel_rmse(gs.readback(), ref)
    gu.close()
    # the guided tracer keeping the reference's environment sample at every vertex
    nee_runs = []
    for seed in (1, 2):
        g = Guide(ctx, seed=seed)
        gs.frame_index = 2 + seed
        dtn, _ = timed(lambda: render_guided(gs, g, 256, progressive=False, **dict(GUIDED_DEFAULTS, env_nee=True)))
        reln = rel_rmse(gs.readback(), ref)
        nee_runs.append({"seconds": round(dtn, 4), "rel_rmse": round(reln, 5),
                         "time_ratio": round(dtn / t_equal(base, reln), 4)})
        g.close()
    # the network's launches timed on their own (event pairs around every inference and
    # training launch of a default guided run; not part of the timed runs above)
    gt = Guide(ctx, seed=1)
    gs.frame_index = 3
    st = render_guided(gs, gt, 256, progressive=False, time_network=True, **GUIDED_DEFAULTS)
    gt.close()
    infer_tf = st["inferred_rows"] * 2 * MLP_MACS / max(st["infer_ms"] * 1e-3, 1e-12) / 1e12
    train_tf = st["trained_samples"] * 3 * 2 * MLP_MACS / max(st["train_ms"] * 1e-3, 1e-12) / 1e12
    mfma = {"inference": {"rows": st["inferred_rows"], "ms": round(st["infer_ms"], 3),
                          "tflops": round(infer_tf, 2), "utilization": round(infer_tf / BF16_PEAK_TFLOPS, 5)},
            "training": {"samples": st["trained_samples"], "steps": st["train_steps"], "ms": round(st["train_ms"], 3),
                         "tflops": round(train_tf, 2), "utilization": round(train_tf / BF16_PEAK_TFLOPS, 5)},
            "flops": "2 x 13,824 per evaluation; 3 x that per trained sample (forward, two backward products); "
                     "peak 2.5 PFLOP/s dense BF16"}
    g0.close()
    gs.close()
    ctx.close()
    mean = lambda rs, k: sum(r[k] for r in rs) / len(rs)  # noqa: E731
    return {
        "workload": "terrain_262144: 512x256-cell fBm height field under a roof, sky + sun HDR "
                    "(sun blocked), 1280x720, 8 bounces",
        "reference_spp": reference_spp, "reference_seconds": round(t_ref, 2),
        "plain_kernel": best, "plain_ms_per_spp": {k: round(v * 1e3, 3) for k, v in speed.items()},
        "plain": {"spp": 256, "seconds": round(base["t256"], 4), "rel_rmse": round(base["r256"], 5),
                  "mpaths_s": round(256 * px / base["t256"] / 1e6, 2), "rel_rmse_64spp": round(base["r64"], 5)},
        "plain_without_env_sample": {"spp": 256, "seconds": round(no_nee["t256"], 4),
                                     "rel_rmse": round(no_nee["r256"], 5),
                                     "rel_rmse_64spp": round(no_nee["r64"], 5)},
        "guided": {"spp": 256, "runs": runs, "settings": GUIDED_DEFAULTS,
                   "mpaths_s": round(256 * px / mean(runs, "seconds") / 1e6, 2)},
        "guided_untrained_network": {"spp": 256, "seconds": round(dtu, 4), "rel_rmse": round(relu, 5),
                                     "trained_over_untrained_rel_rmse": round(mean(runs, "rel_rmse") / relu, 4)},
        "guided_with_env_sample": {"spp": 256, "runs": nee_runs,
                                   "time_ratio": round(mean(nee_runs, "time_ratio"), 4)},
        "mfma": mfma,
        "reference_noise_rel_rmse": round(base["E"] ** 0.5, 5),
        "time_ratio": round(mean(runs, "time_ratio"), 4), "target_time_ratio": 0.5,
        "time_ratio_vs_plain_without_env_sample": round(mean(runs, "time_ratio_vs_plain_without_env_sample"), 4),
    }


def workloads_leg(runs=3):
    """Workloads 1 (Cornell box, 320x180, 4 spp, 4 bounces) and 3 (random triangles and
    spheres on a ground quad under the 1024x512 sky + sun HDR, 1280x720, 64 spp, 8
    bounces) at their full shapes: device ms per frame and Mpaths/s of the 4-wide kernels
    (reference RNG: static, octet and speculative schedules, the wavefront) and of the
    per-sample quad kernel; the mean of `runs` frames after one untimed frame."""
    from toaster_amd import GpuContext, GpuScene, RenderOptions, _abi as A, scenes
    W = A.TH_TRAVERSE_WIDE
    legs = (("quad", RenderOptions(traversal=W)),
            ("octet", RenderOptions(traversal=W, schedule=A.TH_SCHEDULE_OCTET)),
            ("speculative", RenderOptions(traversal=W, schedule=A.TH_SCHEDULE_SPECULATIVE)),
            ("wavefront", RenderOptions(traversal=W, kernel=A.TH_KERNEL_WAVEFRONT)),
            ("quad_per_sample_rng", RenderOptions(traversal=W, rng_mode=A.TH_RNG_PER_SAMPLE)))
    ctx = GpuContext(0)
    out = {}
    for name, sc in (("cornell_320x180_4spp", scenes.cornell()),
                     ("mixed_environment_1280x720_64spp", scenes.mixed_environment())):
        gs = GpuScene.from_scene(ctx, sc)
        res = {"paths_per_frame": sc.paths_per_frame, "extracted_primitives": gs.traversal_info()[0]}
        for lname, o in legs:
            gs.render(options=o, frame_index=0, accumulated_samples=0)
            ctx.synchronize()
            ctx.reset_stats()
            for _ in range(runs):
                gs.render(options=o, frame_index=0, accumulated_samples=0)
            st = ctx.stats()
            kms = st["total_kernel_ms"] / max(1, st["launches"])
            res[lname] = {"kernel_ms": round(kms, 4), "mpaths_s": round(sc.paths_per_frame / kms / 1e3, 2)}
        gs.close()
        out[name] = res
    ctx.close()
    return out


def batched_leg(runs=3):
    """The batched workload: 16 scenes (128 - 262,144 triangles, 0 - 256 spheres, three
    resolutions) x 4 rigid-motion frames, 16 spp, 6 bounces, per-sample RNG, rendered by
    one persistent launch (th_render_batch) with the library's default launch shapes and
    with the autotuner's (per scene: quads per pixel, LDS treelet, tile width, tuned on
    its first frame; per batch: blocks per compute unit), beside the same 64 frames
    rendered one th_render each."""
    from toaster_amd import GpuContext, RenderOptions, _abi as A, scenes
    from toaster_amd.batch import BatchRenderer
    workload = scenes.batched_workload()
    ctx = GpuContext(0)
    br = BatchRenderer(ctx)
    for k, per_scene in enumerate(workload):
        for sc in per_scene:
            br.add(sc, group=k)
    o = RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE, traversal=A.TH_TRAVERSE_WIDE)

    def timed_batch(use_tuning):
        br.render(o, use_tuning=use_tuning)
        ctx.synchronize()
        ctx.reset_stats()
        t = time.perf_counter()
        for _ in range(runs):
            br.render(o, use_tuning=use_tuning)
        ctx.synchronize()
        return (time.perf_counter() - t) / runs, ctx.stats()["total_kernel_ms"] / runs

    dt, kms = timed_batch(False)
    t0 = time.perf_counter()
    tuning = br.autotune(o)
    bpc, _ = br.autotune_batch(o)
    tune_s = time.perf_counter() - t0
    dt_t, kms_t = timed_batch(True)
    # the same frames, one launch each
    for gs in br.scenes:
        gs.render(options=o, frame_index=0, accumulated_samples=0)
    ctx.synchronize()
    t = time.perf_counter()
    for gs in br.scenes:
        gs.render(options=o, frame_index=0, accumulated_samples=0)
    ctx.synchronize()
    dseq = time.perf_counter() - t
    paths = br.paths
    chosen = sorted({(g, c, tw) for g, c, _, tw in tuning})
    br.close()
    ctx.close()
    return {"workload": "16 scenes x 4 rigid-motion frames, 128 - 262,144 triangles, 0 - 256 spheres, "
                        "320x180 / 640x360 / 1280x720, 16 spp, 6 bounces, per-sample RNG",
            "frames": len(br.groups), "paths": paths, "launch_wall_ms": round(dt * 1e3, 2),
            "launch_kernel_ms": round(kms, 2), "gpaths_s": round(paths / dt / 1e9, 3),
            "tuned": {"launch_wall_ms": round(dt_t * 1e3, 2), "launch_kernel_ms": round(kms_t, 2),
                      "gpaths_s": round(paths / dt_t / 1e9, 3), "tuning_seconds": round(tune_s, 2),
                      "blocks_per_cu": bpc,
                      "shapes_quads_cached_tilewidth": [list(x) for x in chosen]},
            "one_launch_per_frame_wall_ms": round(dseq * 1e3, 2),
            "one_launch_per_frame_gpaths_s": round(paths / dseq / 1e9, 3)}


def measured_peaks(ctx):
    """The device's reachable ceilings, beside the published peaks the roofline fractions
    use: streaming copy and streaming read of 1 GiB (th_measure_copy; GB/s moved), and
    dense BF16 MFMA throughput of the inference (32x32x16) and training (16x16x32) shapes
    (th_measure_mfma; an MFMA result that is not exact reports null)."""
    out = {"copy_gbs": round(ctx.measure_copy(1 << 30, 8), 1),
           "read_gbs": round(ctx.measure_copy(1 << 30, 8, read_only=True), 1)}
    for shape, waves, name in ((1, 4, "mfma_bf16_32x32x16_tflops"), (0, 8, "mfma_bf16_16x16x32_tflops")):
        tf, exact = ctx.measure_mfma(shape, waves, 16384)
        out[name] = round(tf, 1) if exact else None
    out["published"] = {"hbm_gbs": HBM_PEAK_GBS, "bf16_dense_tflops": BF16_PEAK_TFLOPS}
    return out


def copy_calibration(mib=1024, copies=10):
    """A plain device-to-device copy of `mib` MiB, `copies` times (torch's copy kernel on
    device memory); each copy reads and writes exactly mib MiB, so the profiler's
    FETCH_SIZE / WRITE_SIZE of that kernel over mib * 1024 KiB is its scale factor."""
    import torch
    n = mib * (1 << 20)
    a = torch.ones(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(copies):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copy calibration: {copies} copies of {n} bytes ({mib * 1024} KiB each)", file=sys.stderr)


def profile_plain_wavefront(passes):
    """The plain estimator as the equal-time leg's fastest plain kernel runs it (the 1-spp
    guided wavefront at alpha 0 with the environment sample, no network) on the terrain
    workload, for a profiler.  First one counting render of the same rays (th_render,
    per-sample streams, one sample, frame seed 0, nearer-first scalar traversal; the
    alpha-0 wavefront equals it bit for bit), whose closest-ray counters give the
    reference-layout bytes the extend kernel's rays read per pass; printed on stderr."""
    from toaster_amd import GpuContext, GpuScene, RenderOptions, _abi as A, scenes
    from toaster_amd.guide import Guide, render_guided
    ctx = GpuContext(0)
    sc = scenes.terrain(1280, 720, samples=1, max_bounces=8)
    gs = GpuScene.from_scene(ctx, sc)
    gs.render(samples=1, frame_index=0, accumulated_samples=0,
              options=RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE, traversal=A.TH_TRAVERSE_ORDERED,
                                    collect_stats=True))
    c = ctx.stats()
    trav = (48 * c["closest_nodes_visited"] + 8 * c["closest_primref_tests"] + 64 * c["closest_triangle_tests"]
            + 32 * c["closest_sphere_tests"])
    io = 40 * c["closest_rays"]  # the wavefront's ray state: origin + direction read, hit record written
    g = Guide(ctx, seed=1)
    t = time.perf_counter()
    render_guided(gs, g, passes, alpha=0.0, train=False, progressive=False, warmup_steps=0)
    ctx.synchronize()
    dt = time.perf_counter() - t
    print(json.dumps({"passes": passes, "seconds": round(dt, 4), "closest_rays_per_pass": c["closest_rays"],
                      "closest_nodes_per_pass": c["closest_nodes_visited"],
                      "extend_traversal_bytes_per_pass": trav, "extend_ray_io_bytes_per_pass": io}),
          file=sys.stderr)
    g.close()
    gs.close()
    ctx.close()


def profile_guided(passes):
    """The guided tracer alone, as the equal-time leg runs it (fresh guide, default
    settings, one sample per pixel per pass) on the terrain workload."""
    from toaster_amd import GpuContext, GpuScene, scenes
    from toaster_amd.guide import GUIDED_DEFAULTS, Guide, render_guided
    ctx = GpuContext(0)
    gs = GpuScene.from_scene(ctx, scenes.terrain(1280, 720, samples=16, max_bounces=8))
    g = Guide(ctx, seed=1)
    t = time.perf_counter()
    st = render_guided(gs, g, passes, progressive=False, **GUIDED_DEFAULTS)
    ctx.synchronize()
    print(f"guided: {passes} passes in {time.perf_counter() - t:.3f} s, {st['train_steps']} training steps",
          file=sys.stderr)
    g.close()
    gs.close()
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--triangles", type=int, default=65536)
    ap.add_argument("--traversal", choices=["reference", "ordered", "wide"], default="wide")
    ap.add_argument("--rng", choices=["reference", "per_sample"], default="reference")
    ap.add_argument("--schedule", choices=["static", "dynamic", "octet", "speculative"], default="speculative",
                    help="pixel-to-lane assignment of the wide kernel with the reference RNG")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-guided", action="store_true")
    ap.add_argument("--no-suite", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-workloads", action="store_true")
    ap.add_argument("--reference-spp", type=int, default=16384)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 1 GiB device buffer 10 times (a copy kernel of known size for the "
                         "profiler's FETCH_SIZE / WRITE_SIZE calibration); prints no JSON line")
    ap.add_argument("--profile-plain-wavefront", type=int, default=0, metavar="PASSES",
                    help="only render PASSES plain 1-spp wavefront passes on the terrain workload (for a "
                         "profiler), after one counting render of the same rays; prints no JSON line")
    ap.add_argument("--profile-guided", type=int, default=0, metavar="PASSES",
                    help="only render PASSES guided passes on the terrain workload (for a profiler); "
                         "prints no JSON line")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.profile_guided:
        profile_guided(args.profile_guided)
        return
    if args.profile_plain_wavefront:
        profile_plain_wavefront(args.profile_plain_wavefront)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)

    from toaster_amd import GpuContext, GpuScene, RenderOptions, _abi as A, scenes
    sc = scenes.bvh_stress(args.triangles)
    trav = {"reference": A.TH_TRAVERSE_REFERENCE, "ordered": A.TH_TRAVERSE_ORDERED,
            "wide": A.TH_TRAVERSE_WIDE}[args.traversal]
    rng = A.TH_RNG_REFERENCE if args.rng == "reference" else A.TH_RNG_PER_SAMPLE
    sched = {"static": A.TH_SCHEDULE_STATIC, "dynamic": A.TH_SCHEDULE_DYNAMIC,
             "octet": A.TH_SCHEDULE_OCTET, "speculative": A.TH_SCHEDULE_SPECULATIVE}[args.schedule]
    opts = RenderOptions(rng_mode=rng, traversal=trav, schedule=sched)
    if args.traversal == "wide" and args.rng == "reference":
        kernel_name = {"static": "k_render_quad", "dynamic": "k_render_quad_dyn",
                       "octet": "k_render_octet", "speculative": "k_render_spec"}[args.schedule]
    elif args.traversal == "wide":
        kernel_name = "k_render_quad"
    else:
        kernel_name = "k_render_pixel" if args.rng == "reference" else "k_render_sample"
    ctx = GpuContext(0)
    gs = GpuScene.from_scene(ctx, sc)
    # one counting pass (outside the timed region) for rays per frame; rays do not
    # depend on the traversal order, so the scalar ordered traversal counts them
    gs.render(options=RenderOptions(rng_mode=rng, traversal=A.TH_TRAVERSE_ORDERED,
                                    collect_stats=True))
    st = ctx.stats()
    rays_per_frame = st["rays"]
    ordered_counters = dict(st)  # nearer-child-first traversal: the bytes a nearest-first walk reads
    for _ in range(args.warmup):
        gs.render(options=opts, frame_index=0, accumulated_samples=0)
    ctx.synchronize()
    ctx.reset_stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gs.render(options=opts, frame_index=0, accumulated_samples=0)
    ctx.synchronize()
    t1 = time.perf_counter()
    st = ctx.stats()
    headline_image = gs.readback().copy()  # the timed frame (seed 0, no accumulation)
    # other kernel variants on the same frame (short runs, reported beside the headline)
    variants = {}
    for name, o in (("megakernel_reference_traversal", RenderOptions(rng_mode=rng, traversal=A.TH_TRAVERSE_REFERENCE)),
                    ("megakernel_ordered", RenderOptions(rng_mode=rng, traversal=A.TH_TRAVERSE_ORDERED)),
                    ("quad_wide", RenderOptions(rng_mode=rng, traversal=A.TH_TRAVERSE_WIDE)),
                    ("octet_wide", RenderOptions(rng_mode=rng, traversal=A.TH_TRAVERSE_WIDE,
                                                 schedule=A.TH_SCHEDULE_OCTET)),
                    ("speculative_wide", RenderOptions(rng_mode=rng, traversal=A.TH_TRAVERSE_WIDE,
                                                       schedule=A.TH_SCHEDULE_SPECULATIVE)),
                    ("quad_wide_dynamic_schedule", RenderOptions(rng_mode=rng, traversal=A.TH_TRAVERSE_WIDE,
                                                                 schedule=A.TH_SCHEDULE_DYNAMIC)),
                    ("quad_wide_per_sample_rng", RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE,
                                                               traversal=A.TH_TRAVERSE_WIDE)),
                    ("megakernel_ordered_per_sample_rng", RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE,
                                                                        traversal=A.TH_TRAVERSE_ORDERED)),
                    ("wavefront_wide", RenderOptions(rng_mode=rng, traversal=A.TH_TRAVERSE_WIDE,
                                                     kernel=A.TH_KERNEL_WAVEFRONT)),
                    ("wavefront_wide_per_sample_rng", RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE,
                                                                    traversal=A.TH_TRAVERSE_WIDE,
                                                                    kernel=A.TH_KERNEL_WAVEFRONT)),
                    ("wavefront_ordered", RenderOptions(rng_mode=rng, traversal=A.TH_TRAVERSE_ORDERED,
                                                        kernel=A.TH_KERNEL_WAVEFRONT))):
        gs.render(options=o, frame_index=0, accumulated_samples=0)
        ctx.synchronize()
        ctx.reset_stats()
        n = 10
        for _ in range(n):
            gs.render(options=o, frame_index=0, accumulated_samples=0)
        vs = ctx.stats()
        kms = vs["total_kernel_ms"] / max(1, vs["launches"])
        variants[name] = {"kernel_ms": round(kms, 4), "mpaths_s": round(sc.paths_per_frame / kms / 1e3, 2)}
    gs.close()
    scene_update = scene_update_leg(ctx, sc)
    log("measured ceilings")
    peaks = measured_peaks(ctx)
    ctx.close()
    dt = t1 - t0
    paths = sc.paths_per_frame
    ms_step = dt * 1e3 / args.steps
    kernel_ms = st["total_kernel_ms"] / max(1, st["launches"])

    log("headline done; equal_time leg")
    equal_time = None if args.no_guided else equal_time_guided(args.reference_spp)
    log("reference suite leg")
    suite = None if args.no_suite else reference_suite()
    log("batched leg")
    batched = None if args.no_batch else batched_leg()
    log("workloads 1 and 3 leg")
    workloads = None if args.no_workloads else workloads_leg()
    cpu, counters, parity = (None, None, None)
    if not args.no_cpu:
        log("cpu baseline leg (native oracle build, parity check)")
        cpu, counters, parity = cpu_baseline(sc, headline_image)
    roofline = None
    if counters is not None:
        bytes_launch = algorithmic_bytes(counters, sc.width * sc.height)
        achieved = bytes_launch / (kernel_ms * 1e-3) / 1e9
        bytes_nf = algorithmic_bytes(ordered_counters, sc.width * sc.height)
        traffic = None
        tpath = os.path.join(ROOT, "profiles", f"traffic_bvh_stress_{args.triangles}.json")
        if os.path.exists(tpath):
            with open(tpath) as f:
                traffic = json.load(f).get("hbm_bytes_per_launch")
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                    "algorithmic_bytes_per_launch": bytes_launch,
                    "kernel": kernel_name,
                    "kernel_ms": round(kernel_ms, 5),
                    # the same record sizes over the node visits of a nearer-child-first
                    # walk (the GPU's ordered traversal counters), the order the 4-wide
                    # kernels visit in: what the headline kernel actually has to read
                    "nearest_first": {"algorithmic_bytes_per_launch": bytes_nf,
                                      "achieved": round(bytes_nf / (kernel_ms * 1e-3) / 1e9, 2),
                                      "frac": round(bytes_nf / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
                                      "nodes_visited": ordered_counters["nodes_visited"],
                                      "reference_order_nodes_visited": counters["nodes_visited"]},
                    # against the streaming read this device measurably reaches
                    "measured_read_gbs": peaks["read_gbs"],
                    "frac_of_measured_read": round(achieved / peaks["read_gbs"], 5) if peaks["read_gbs"] else None}
    if equal_time is not None and "mfma" in equal_time:
        for leg, shape in (("inference", "mfma_bf16_32x32x16_tflops"), ("training", "mfma_bf16_16x16x32_tflops")):
            m = equal_time["mfma"].get(leg)
            if m is not None and peaks.get(shape):
                m["utilization_of_measured"] = round(m["tflops"] / peaks[shape], 5)
    line = {
        "metric": "Mpaths/s + equal-time relRMSE: guided vs plain HIP tracer, 1 MI355X",
        "value": round(args.steps * paths / dt / 1e6, 3),
        "unit": "Mpaths/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"bvh_stress_{args.triangles}: {args.triangles} random triangles + "
                               "1 emissive sphere, 320x180, 4 spp, 4 bounces, median-split BVH",
                   "triangles": args.triangles, "width": sc.width, "height": sc.height,
                   "spp": sc.samples, "max_bounces": sc.max_bounces, "paths_per_step": paths,
                   "rng": args.rng, "traversal": args.traversal, "schedule": args.schedule,
                   "kernel": kernel_name},
        "rays_per_s": round(args.steps * rays_per_frame / dt, 1),
        "kernel_ms_mean": round(kernel_ms, 5),
        "roofline": roofline,
        "measured_peaks": peaks,
        "variants": variants,
        "scene_update": scene_update,
        "equal_time": equal_time,
        "reference_suite": suite,
        "batched": batched,
        "workloads": workloads,
        "cpu_baseline": cpu,
        "parity": parity,
        "host": platform.node(),
    }
    print(json.dumps(line))
    if parity is not None and not parity["ok"]:
        print("bench.py: the headline image differs from the oracle beyond tolerance", file=sys.stderr)
        sys.exit(1)


if __name__ == "__main__":
    main()
