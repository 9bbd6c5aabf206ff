"""Benchmark of the HIP rendering core.

Headline workload, bvh_stress (one "step" = one full frame): the seeded
bvh_stress scene -- a displaced geodesic icosphere of 1,048,820 triangles plus
4,096 random spheres (1,052,916 primitives, reference-topology SAH BVH) -- at
1920x1080, 1 spp, Whitted depth 1 with one hard point light: one primary
(closest-hit) ray and one shadow (any-hit) ray per pixel that hits.  Scene,
BVH and output buffer are resident in HBM before the timed region.

value = algorithmic Mrays/s of the plain per-pixel megakernel (closest-hit +
any-hit queries per second, whole frame, host wall clock around K
synchronous frames).  The batched wavefront tracer runs the same frame and is
reported beside it with its image RMSE against the plain one.  The
reference's own convention (trace() calls only, no shadow rays) is reported
too.

roofline: the dominant kernel is the render megakernel.  Its algorithmic bytes
per launch are sum over queries of (32 B ray + 8 B hit) + 112 B per 4-wide
node entered (four child boxes + four refs) + the primitive's bytes per test (sphere 16, triangle 48, cube 16,
plane 32, cylinder 48, cone 48), counted by an instrumented launch of the
same frame; divided by the kernel's mean HIP-event duration.

cpu_baseline: the reference renderer itself -- its parser, BVH and
Renderer::render, driven by oracle/_ref/ref_harness (built from the reference
sources by oracle/Makefile) on the same scene written as .pecl + .obj -- with
every host core this process may run on (os.sched_getaffinity), render()
timed alone as src/pencil.cpp:311-313 does (parse and BVH build untimed),
the median of 5 renders after one warm-up in the same process.  The CPU
model and core count are reported with it.  Without the reference build the
baseline is reported as missing (never replaced by the oracle).

Every leg carries its own "roofline" (algorithmic bytes of its queries from
an instrumented launch, over the leg's kernel time) and "cpu_baseline"
(the reference on a bounded sample of the same workload).

Further workloads ("legs", each reported under "legs" in the JSON line):
  cornell_pt   Cornell box (pencil_amd/scenes/cornell_c5.pecl) with the
               1,048,820-triangle mesh inside, 1920x1080, 256 spp, path
               tracing with 8 bounces and Russian roulette: one frame, plain
               megakernel and batched wavefront, Mrays/s and RMSE.
  batch64      64 seeded scenes (mesh of 16..500k triangles, log-uniform,
               plus spheres with a diffuse / metal / glass mix) at 256^2,
               512^2 or 1024^2, 16 spp, Whitted depth 6: one batched launch
               against 64 sequential plain renders.
  inverse      Cornell box at 512x512, 64 spp, 4 bounces, 48 parameters
               (8 materials x albedo + emission): target from the true
               parameters, 200 Adam steps from a seeded U[0.2, 0.8] start;
               gradient steps per second (one pass renders the image and
               its gradient), loss curve ends,
               and the gradient against central differences at the start.
"""
import argparse
import json
import os
import shutil
import statistics
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from pencil_amd import workloads as wl  # noqa: E402

PRIM_BYTES = [16, 48, 16, 32, 48, 48]   # sphere, triangle, cube, plane, cylinder, cone
NODE_BYTES = 112                         # a 4-wide node read: 6 x 16 B of child boxes + 16 B of refs
HBM_PEAK_GBS = 8000.0
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
PROFILES = os.path.join(ROOT, "profiles")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cfg_dict(c):
    return {"slots": c.slots, "blocks_per_cu": c.blocks_per_cu, "stack_k": c.stack_k,
            "iters_per_graph": c.iters_per_graph, "drain_below": c.drain_below,
            "inline_shadows": c.inline_shadows, "streams": c.streams}


def cfg_str(c):
    return (f"slots {c.slots}, blocks/CU {c.blocks_per_cu}, stack {c.stack_k}, iters/graph {c.iters_per_graph}, "
            f"drain below {c.drain_below}, inline shadows {c.inline_shadows}, streams {c.streams}")


def host_cpu():
    """(threads this process may use, CPU model)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return n, model


def oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    return oracle_lib


def rmse(a, b):
    return float(np.sqrt(np.mean((a.astype(np.float64) - b) ** 2)))


def rounded(d):
    if isinstance(d, dict):
        return {k: rounded(v) for k, v in d.items()}
    if isinstance(d, float):
        return round(d, 4) if abs(d) >= 1e-3 or d == 0 else float(f"{d:.4g}")
    return d


def alg_bytes(st, wc):
    """Algorithmic bytes of a launch's queries: 32 B ray + 8 B hit per query, 112 B per
    4-wide node entered, the primitive's bytes per test."""
    q = st.closest_queries + st.anyhit_queries
    return q * (32 + 8) + NODE_BYTES * wc.node_visits + sum(b * wc.prim_tests[k] for k, b in enumerate(PRIM_BYTES))


def profile_traffic(name):
    """HBM bytes per launch (or per job) from a committed rocprofv3 PMC summary, or None."""
    path = os.path.join(PROFILES, name + ".traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        t = json.load(f)
    return int(round(t["hbm_bytes_per_dispatch"])), t


def _pmc_run(counter, args, tmp, name):
    """One rocprofv3 PMC pass (one counter) over `python3 bench.py <args>` in a child
    process; returns the rows of its counter CSV."""
    import csv
    out = os.path.join(tmp, name)
    cmd = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", out, "-o", "run", "--",
           sys.executable, os.path.join(ROOT, "bench.py")] + args
    env = dict(os.environ, TMPDIR="/tmp")
    r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True, timeout=240)
    files = [os.path.join(dp, f) for dp, _, fs in os.walk(out) for f in fs if f.endswith("counter_collection.csv")]
    if not files:
        raise RuntimeError(f"rocprofv3 --pmc {counter}: no counter file (rc {r.returncode}): {r.stderr[-300:]}")
    with open(files[0]) as f:
        return list(csv.DictReader(f))


def live_traffic(batched_cfg):
    """HBM bytes of the headline frame measured in this run: FETCH_SIZE and WRITE_SIZE
    of the plain kernel's two timed frames and of the batched frames (every wavefront
    kernel, per frame), from child runs of this script under rocprofv3, each counter
    scaled by what it reads for a 256 MiB copy kernel.  Returns (plain, batched)
    dicts, or (None, None) when a pass fails (the committed profiles are used then)."""
    tmp = tempfile.mkdtemp(prefix="pcl_pmc_", dir="/tmp")
    try:
        args = ["--no-cpu", "--no-live-counters", "--legs", "none", "--steps", "2", "--warmup", "0",
                "--batched-config", batched_cfg]
        scale = {}
        for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
            rows = _pmc_run(ctr, ["--copy-probe"], tmp, "copy_" + ctr)
            v = [float(r["Counter_Value"]) for r in rows if "AUnaryFunctor" in r["Kernel_Name"]]
            scale[ctr] = 262144.0 / (sum(v) / len(v))   # 256 MiB per dispatch, counters in KiB
        got = {"plain": {}, "batched": {}}
        for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
            rows = _pmc_run(ctr, args, tmp, "head_" + ctr)
            plain = sorted({int(r["Dispatch_Id"]): float(r["Counter_Value"]) for r in rows
                            if "k_render_p<false, 2, 0>" in r["Kernel_Name"]}.items())[:2]
            wf = [float(r["Counter_Value"]) for r in rows if "pcl::wf::" in r["Kernel_Name"]]
            if len(plain) != 2 or not wf:
                raise RuntimeError(f"unexpected dispatches under rocprofv3 ({len(plain)} plain)")
            got["plain"][ctr] = sum(v for _, v in plain) / 2 * 1024 * scale[ctr]
            got["batched"][ctr] = sum(wf) / 2 * 1024 * scale[ctr]
        res = tuple({"read_bytes": g["FETCH_SIZE"], "write_bytes": g["WRITE_SIZE"],
                     "hbm_bytes": g["FETCH_SIZE"] + g["WRITE_SIZE"], "fetch_scale": scale["FETCH_SIZE"],
                     "write_scale": scale["WRITE_SIZE"]} for g in (got["plain"], got["batched"]))
        return res
    except Exception as e:   # no profiler, a failed pass: fall back to the committed profiles
        log(f"live HBM counters not measured ({e}); the rooflines keep the committed profiles' traffic")
        return None, None
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def roofline(bytes_per_launch, kernel_ms, traffic_name, scope, limiter, live=None):
    """Algorithmic-byte roofline of a kernel (or a job of kernels), with the
    counter-measured HBM traffic, L2 hit rate and VALU issue beside it.
    `frac` is algorithmic-byte throughput over the HBM peak; `limiter` says
    what the counters show actually bounds the kernel (none of these is
    bandwidth-bound: the BVH sits in the 256 MB MALL)."""
    achieved = bytes_per_launch / (kernel_ms / 1000.0) / 1e9
    traffic, t = profile_traffic(traffic_name)
    source = f"profiles/{traffic_name}.traffic.json"
    if live:   # counters of this run
        traffic, source = int(round(live["hbm_bytes"])), "rocprofv3 FETCH_SIZE / WRITE_SIZE passes in this run"
    r = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": source, "scope": scope,
         "algorithmic_bytes": int(bytes_per_launch), "limiter": limiter}
    if traffic:
        # the rate the counters see, over the kernel time measured here
        r["measured_hbm_gbs"] = round(traffic / (kernel_ms / 1000.0) / 1e9, 1)
        r["measured_frac"] = round(traffic / (kernel_ms / 1000.0) / 1e9 / HBM_PEAK_GBS, 4)
        r["l2_hit"] = round(t["l2_hit"], 3)
        r["valu_issue"] = round(t["valu_issue"], 3)
        r["traffic_profile"] = f"profiles/{traffic_name}.md"
    return r


# ---------------------------------------------------------------------------
# the reference renderer on the host cores (CPU baseline)
# ---------------------------------------------------------------------------

def ref_render(pecl, opts, threads, rows=None, repeat=6):
    """Renders with the reference (parse and BVH build once, untimed), `repeat`
    times; returns (median of the renders after the first, all render seconds,
    reference-convention rays, HDR image or band)."""
    out = pecl + ".bin"
    cmd = [HARNESS, "render", pecl, out] + opts + ["--threads", str(threads), "--repeat", str(repeat), "--no-tag"]
    if rows:
        cmd += ["--rows", str(rows[0]), str(rows[1])]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        raise RuntimeError(f"reference render failed ({r.returncode}): {r.stderr[-500:]}")
    secs, rays = None, None
    for line in r.stdout.splitlines():
        if line.startswith("seconds"):
            secs = [float(x) for x in line.split()[1:]]
        elif line.startswith("rays"):
            rays = int(line.split()[1])
    raw = open(out, "rb").read()
    W, H = np.frombuffer(raw[:8], dtype=np.int32)
    img = np.frombuffer(raw[8:8 + W * H * 12], dtype=np.float32).reshape(H, W, 3).copy()
    timed = secs[1:] if len(secs) > 1 else secs
    return statistics.median(timed), secs, rays, img


def cpu_entry(value, unit, threads, model, sample, secs):
    return {"value": round(value, 6) if value < 1 else round(value, 3), "unit": unit, "cores": threads,
            "kind": "reference", "cpu_model": model, "sample": sample,
            "render_s_runs": [round(x, 4) for x in secs]}


# ---------------------------------------------------------------------------
# headline: bvh_stress
# ---------------------------------------------------------------------------

def leg_stress(core, pa, args):
    t0 = time.time()
    hs, st = wl.stress(pa)
    desc, cam = hs.desc(), hs.camera()
    log(f"scene: {desc.n_prims} primitives, generated in {time.time() - t0:.2f} s")
    sid = core.upload(desc)
    info = core.scene_info(sid)
    log(f"bvh: {info.n_nodes} nodes, depth {info.depth}, {info.device_bytes / 1e6:.1f} MB on device, "
        f"build {info.build_ms:.0f} ms, upload {info.upload_ms:.0f} ms")
    W, H = st.width, st.height
    d_out = core.alloc(W * H * 3 * 4)

    # instrumented frame: work counts for the bytes model (not timed)
    cst, wc = core.render_counts(sid, cam, st, d_out)
    queries = cst.closest_queries + cst.anyhit_queries
    bytes_per_launch = alg_bytes(cst, wc)
    log(f"counts: {cst.closest_queries} closest + {cst.anyhit_queries} any-hit queries, "
        f"{wc.node_visits} node visits, tests {[wc.prim_tests[k] for k in range(6)]}, "
        f"{bytes_per_launch / 1e9:.2f} GB algorithmic per frame")

    for _ in range(args.warmup):
        core.render_device(sid, cam, st, d_out)
    kms = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        s = core.render_device(sid, cam, st, d_out)
        kms.append(s.kernel_ms)
    t_end = time.perf_counter()
    ms_per_step = (t_end - t_start) * 1000.0 / args.steps
    kernel_ms = sum(kms) / len(kms)
    value = queries / (ms_per_step / 1000.0) / 1e6
    log(f"timed: {args.steps} frames, {ms_per_step:.3f} ms/frame wall, {kernel_ms:.3f} ms kernel (HIP events)")

    batched = None
    if not args.plain_only:
        d_wf = core.alloc(W * H * 3 * 4)
        job = [(sid, cam, st, d_wf)]
        if args.batched_config:
            best = pa.WfConfig(*[int(x) for x in args.batched_config.split(",")])
            tune_ms = 0.0
        else:
            best, tune_ms = core.autotune(job)
        log(f"autotune: {cfg_str(best)} -> {tune_ms:.3f} ms")
        for _ in range(args.warmup):
            core.render_batch(job, best)
        wkms = []
        t_start = time.perf_counter()
        for _ in range(args.steps):
            bs = core.render_batch(job, best)
            wkms.append(bs.kernel_ms)
        t_end = time.perf_counter()
        a = np.empty((H, W, 3), dtype=np.float32)
        b = np.empty((H, W, 3), dtype=np.float32)
        core.d2h(a, d_out)
        core.d2h(b, d_wf)
        wq = bs.closest_queries + bs.anyhit_queries
        wms = (t_end - t_start) * 1000.0 / args.steps
        wk = sum(wkms) / len(wkms)
        batched = {"ms_per_step": wms, "kernel_ms": wk, "queries": wq,
                   "mrays": wq / (wms / 1000.0) / 1e6, "rmse_vs_plain": rmse(a, b),
                   "bit_identical_to_plain": bool(np.array_equal(a.view(np.uint32), b.view(np.uint32))),
                   "iterations": bs.iterations, "config": cfg_dict(bs.used),
                   "roofline": roofline(bytes_per_launch, wk, "wavefront_bvh_stress_1080p", "job",
                                         "one iteration: the closest-hit kernel's traversal latency, then resolve")}
        log(f"batched: {wms:.3f} ms/frame wall, {wk:.3f} ms on the stream, "
            f"{batched['mrays']:.1f} Mrays/s, RMSE vs plain {batched['rmse_vs_plain']:.3g}")
        core.free(d_wf)
    # the device BVH builder (use_bvh = 2) on the same scene: build time and frame time
    desc.use_bvh = 2
    t0 = time.perf_counter()
    sid2 = core.upload(desc)
    up_ms = (time.perf_counter() - t0) * 1000.0
    info2 = core.scene_info(sid2)
    for _ in range(args.warmup):
        core.render_device(sid2, cam, st, d_out)
    t_start = time.perf_counter()
    for _ in range(args.steps):
        s2 = core.render_device(sid2, cam, st, d_out)
    lb_ms = (time.perf_counter() - t_start) * 1000.0 / args.steps
    lbvh = {"upload_ms_total": up_ms, "host_prep_ms": info2.build_ms, "device_build_and_upload_ms": info2.upload_ms,
            "depth": info2.depth, "ms_per_step": lb_ms,
            "mrays": (s2.closest_queries + s2.anyhit_queries) / (lb_ms / 1000.0) / 1e6,
            "host_sah_build_ms": info.build_ms}
    log(f"device LBVH: upload {up_ms:.1f} ms (host prep {info2.build_ms:.1f}, device build + copy "
        f"{info2.upload_ms:.1f}), depth {info2.depth}, {lb_ms:.3f} ms/frame")
    core.free_scene(sid2)
    core.free(d_out)
    core.free_scene(sid)
    return dict(value=value, ms_per_step=ms_per_step, kernel_ms=kernel_ms, bytes=bytes_per_launch, queries=queries,
                cst=cst, info=info, desc_n=desc.n_prims, W=W, H=H, batched=batched, lbvh=lbvh)


def cpu_stress(queries, threads, model):
    """The reference on the whole headline frame."""
    tmp = tempfile.mkdtemp(prefix="pcl_bench_")
    try:
        import pencil_amd as pa
        pa.HostScene.write_stress(tmp, "stress", **wl.STRESS)
        t0 = time.time()
        med, secs, rays, _ = ref_render(os.path.join(tmp, "stress.pecl"), [], threads)
        wall = time.time() - t0
        return cpu_entry(queries / med / 1e6, "Mrays/s", threads, model,
                         f"the full 1920x1080 frame: Renderer::render of the reference build, median of 5 renders "
                         f"after a warm-up ({med:.3f} s; {rays:,} reference-convention rays; the {queries:,} "
                         f"queries counted by the core on the same frame; {wall:.1f} s wall incl. .pecl/.obj "
                         f"parse and BVH build)", secs)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


# ---------------------------------------------------------------------------
# glass_defer: the Cornell + mesh frame, glass vertices evaluated in place
# and deferred, plain and batched, soft and hard light (not run by default)
# ---------------------------------------------------------------------------

def leg_glass_defer(core, pa, args, threads, model):
    """Glass vertices evaluated in place vs deferred to a second launch, for
    the plain megakernel and the batched tracer's finish launch, on the
    path-traced Cornell + mesh frame as configured (soft point light) and with
    the light's radius set to 0.  Every image must be the same bit for bit."""
    cfg_s = args.batched_config or "2073600,6,16,4,0,-1,3"
    cfg = pa.WfConfig(*[int(x) for x in cfg_s.split(",")])
    out, imgs = {}, []
    for light in ("soft", "hard"):
        hs, st = wl.cornell_pt(pa, samples=args.pt_samples)
        desc, cam = hs.desc(), hs.camera()
        if light == "hard":
            for i in range(desc.n_lights):
                desc.lights[i].radius = 0.0
        W, H = st.width, st.height
        res = {}
        # the bench's own context, toggled: a second context in the process
        # would take a hardware queue from the batched tracer's three streams
        sid = core.upload(desc)
        d = core.alloc(W * H * 12)
        try:
            for name, enable in (("in_place", False), ("deferred", True)):
                core.set_glass_deferral(enable)
                r = {}
                for path in ("plain", "batched"):
                    run = (lambda: core.render_device(sid, cam, st, d)) if path == "plain" else \
                        (lambda: core.render_batch([(sid, cam, st, d)], cfg))
                    run()   # warm-up (buffers, graph)
                    t0 = time.perf_counter()
                    s = run()
                    ms = (time.perf_counter() - t0) * 1000.0
                    img = np.empty((H, W, 3), dtype=np.float32)
                    core.d2h(img, d)
                    imgs.append(img)
                    q = s.closest_queries + s.anyhit_queries
                    r[path] = {"ms": ms, "kernel_ms": s.kernel_ms, "mrays_s": q / ms / 1e3, "queries": q,
                               "deferred_glass_vertices": s.deferred_glass}
                    log(f"glass_defer {light} {path} {name}: {ms:.1f} ms, {q / ms / 1e3:.1f} Mrays/s, "
                        f"{s.deferred_glass} deferred")
                res[name] = r
        finally:
            core.set_glass_deferral(True)
            core.free(d)
            core.free_scene(sid)
        res["speedup_deferred_plain"] = res["in_place"]["plain"]["ms"] / res["deferred"]["plain"]["ms"]
        res["speedup_deferred_batched"] = res["in_place"]["batched"]["ms"] / res["deferred"]["batched"]["ms"]
        out[light] = res
    out["workload"] = (f"Cornell box + {desc.n_prims - 14}-tri mesh, {W}x{H}, {st.samples ** 2} spp, path tracing, "
                       f"8 bounces, Russian roulette after 3; the point light as configured (radius 0.5: soft, "
                       f"jittered) and with its radius set to 0 (hard); plain megakernel and batched tracer "
                       f"({cfg_s})")
    out["bit_identical"] = bool(all(np.array_equal(imgs[k].view(np.uint32), imgs[k - k % 4].view(np.uint32))
                                    for k in range(len(imgs))))
    return rounded(out)


# ---------------------------------------------------------------------------
# cornell_pt: path tracing, plain vs batched
# ---------------------------------------------------------------------------

def leg_cornell_pt(core, pa, args, threads, model):
    hs, st = wl.cornell_pt(pa, samples=args.pt_samples)
    desc, cam = hs.desc(), hs.camera()
    sid = core.upload(desc)
    W, H = st.width, st.height
    d_a = core.alloc(W * H * 12)
    d_b = core.alloc(W * H * 12)
    try:
        cst, wc = core.render_counts(sid, cam, st, d_b)   # instrumented: the bytes model's units
        nbytes = alg_bytes(cst, wc)
        t0 = time.perf_counter()
        ps = core.render_device(sid, cam, st, d_a)
        plain_ms = (time.perf_counter() - t0) * 1000.0
        q = ps.closest_queries + ps.anyhit_queries
        log(f"cornell_pt plain: {plain_ms:.1f} ms, {q} queries, {q / plain_ms / 1e3:.1f} Mrays/s, "
            f"{nbytes / 1e9:.1f} GB algorithmic")
        if args.batched_config:
            cfg = pa.WfConfig(*[int(x) for x in args.batched_config.split(",")])
        else:
            # tuned on the same frame at 2x2 spp (a few seconds), rendered at full spp
            st_t = pa.Settings.from_buffer_copy(st)
            st_t.samples = 2
            cfg, tune_ms = core.autotune([(sid, cam, st_t, d_b)])
            log(f"cornell_pt autotune at 2x2 spp: {cfg_str(cfg)} -> {tune_ms:.1f} ms")
        core.render_batch([(sid, cam, st, d_b)], cfg)   # warm-up at full spp (the drain's buffers)
        t0 = time.perf_counter()
        bs = core.render_batch([(sid, cam, st, d_b)], cfg)
        wf_ms = (time.perf_counter() - t0) * 1000.0
        wq = bs.closest_queries + bs.anyhit_queries
        log(f"cornell_pt batched: {wf_ms:.1f} ms, {wq} queries, {wq / wf_ms / 1e3:.1f} Mrays/s, "
            f"{bs.iterations} iterations, {bs.finished_pixels} pixels finished after the drain ({cfg_str(bs.used)})")
        a = np.empty((H, W, 3), dtype=np.float32)
        b = np.empty((H, W, 3), dtype=np.float32)
        core.d2h(a, d_a)
        core.d2h(b, d_b)
    finally:
        core.free(d_a)
        core.free(d_b)
        core.free_scene(sid)
    out = {"workload": f"Cornell box + {desc.n_prims - 14}-tri mesh, {W}x{H}, {st.samples ** 2} spp, path tracing, "
                       f"8 bounces, Russian roulette after 3",
           "primitives": desc.n_prims, "queries": q, "ref_convention_rays": ps.rays_traced,
           "plain_ms": plain_ms, "plain_kernel_ms": ps.kernel_ms, "plain_mrays_s": q / plain_ms / 1e3,
           "batched_ms": wf_ms, "batched_mrays_s": wq / wf_ms / 1e3, "batched_kernel_ms": bs.kernel_ms,
           "batched_iterations": bs.iterations, "batched_config": cfg_dict(bs.used),
           "batched_pixels_finished_after_drain": bs.finished_pixels,
           "speedup_batched_vs_plain": plain_ms / wf_ms, "rmse_batched_vs_plain": rmse(a, b),
           "bit_identical": bool(np.array_equal(a.view(np.uint32), b.view(np.uint32))),
           "image_mean": float(a.mean()),
           "roofline": roofline(nbytes, ps.kernel_ms, "k_render_cornell_pt_1080p_256spp", "kernel",
                                "VALU issue: divergent lanes (path lengths, glass subtrees)"),
           "batched_roofline": roofline(nbytes, bs.kernel_ms, "wavefront_cornell_pt_1080p_256spp", "job",
                                        "traversal latency, iteration barriers and the drain tail")}
    if not args.no_cpu and os.path.exists(HARNESS):
        rows = (H // 2 - 2, H // 2 + 2)
        tmp = tempfile.mkdtemp(prefix="pcl_bench_")
        try:
            pa.HostScene.write_cornell_mesh(tmp, "cornell_pt", wl.CORNELL, **wl.MESH_IN_BOX)
            opts = ["--width", str(W), "--height", str(H), "--samples", str(st.samples), "--pathtrace", "--bounces",
                    str(st.pathtrace_bounces)]
            med, secs, rays, band = ref_render(os.path.join(tmp, "cornell_pt.pecl"), opts, threads, rows=rows)
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
        # the oracle checks the sample (bit for bit: same pixels, same rays) and counts its queries
        ol = oracle()
        img, c = ol.OracleScene(desc).render(cam, st, threads=threads, rows=rows)
        same = bool(np.array_equal(img[rows[0]:rows[1]], band)) and c.rays_traced == rays
        cq = c.closest_queries + c.anyhit_queries
        out["cpu_baseline"] = cpu_entry(
            cq / med / 1e6, "Mrays/s", threads, model,
            f"rows {rows[0]}-{rows[1] - 1} (the middle of the frame) at {st.samples ** 2} spp: render_tile_region of "
            f"the reference build, median of 5 after a warm-up ({med:.3f} s; {rays:,} reference-convention rays, "
            f"{cq:,} queries as the oracle counts them on the same rows, its pixels "
            f"{'identical' if same else 'NOT identical'} to the reference's)", secs)
    return rounded(out)


# ---------------------------------------------------------------------------
# batch64: 64 uneven scenes in one batched job vs 64 plain renders
# ---------------------------------------------------------------------------

def leg_batch64(core, pa, args, threads, model):
    specs = wl.batch64_specs()
    t0 = time.time()
    scenes = wl.batch64(pa)
    sids = [core.upload(hs.desc()) for hs, _ in scenes]
    n_prims = sum(hs.desc().n_prims for hs, _ in scenes)
    log(f"batch64: {len(scenes)} scenes, {n_prims} primitives, built and uploaded in {time.time() - t0:.1f} s")
    outs = [core.alloc(st.width * st.height * 12) for _, st in scenes]
    outs2 = [core.alloc(st.width * st.height * 12) for _, st in scenes]
    per_scene_q = []
    try:
        nbytes = 0
        for sid, (hs, st), d in zip(sids, scenes, outs2):
            cst, wc = core.render_counts(sid, hs.camera(), st, d)
            nbytes += alg_bytes(cst, wc)
        job = [(sid, hs.camera(), st, d) for sid, (hs, st), d in zip(sids, scenes, outs)]
        if args.batched_config:
            cfg = pa.WfConfig(*[int(x) for x in args.batched_config.split(",")])
        else:
            # tuned on the same 64 scenes at 2x2 spp, rendered at full spp
            tune_job = []
            for sid, cam, st, d in job:
                st_t = pa.Settings.from_buffer_copy(st)
                st_t.samples = 2
                tune_job.append((sid, cam, st_t, d))
            cfg, tune_ms = core.autotune(tune_job)
            log(f"batch64 autotune at 2x2 spp: {cfg_str(cfg)} -> {tune_ms:.1f} ms")
        core.render_batch(job, cfg)   # warm-up (graph capture, buffers)
        t0 = time.perf_counter()
        bs = core.render_batch(job, cfg)
        wf_ms = (time.perf_counter() - t0) * 1000.0
        wq = bs.closest_queries + bs.anyhit_queries
        for sid, (hs, st), d in zip(sids, scenes, outs2):   # warm-up
            core.render_device(sid, hs.camera(), st, d)
        t0 = time.perf_counter()
        pq, pk = 0, 0.0
        for sid, (hs, st), d in zip(sids, scenes, outs2):
            s = core.render_device(sid, hs.camera(), st, d)
            pq += s.closest_queries + s.anyhit_queries
            pk += s.kernel_ms
            per_scene_q.append(s.closest_queries + s.anyhit_queries)
        plain_ms = (time.perf_counter() - t0) * 1000.0
        err, same = 0.0, True
        for (hs, st), d, d2 in zip(scenes, outs, outs2):
            a = np.empty((st.height, st.width, 3), np.float32)
            b = np.empty_like(a)
            core.d2h(a, d)
            core.d2h(b, d2)
            err = max(err, rmse(a, b))
            same = same and bool(np.array_equal(a.view(np.uint32), b.view(np.uint32)))
    finally:
        for d in outs + outs2:
            core.free(d)
        for sid in sids:
            core.free_scene(sid)
    log(f"batch64: batched {wf_ms:.1f} ms ({wq / wf_ms / 1e3:.1f} Mrays/s), 64 plain renders {plain_ms:.1f} ms "
        f"({pq / plain_ms / 1e3:.1f} Mrays/s)")
    out = {"workload": "64 seeded scenes (mesh 16..500k tris log-uniform + 8..256 spheres, diffuse/metal/glass), "
                       "256^2/512^2/1024^2, 16 spp, Whitted depth 6, soft point light",
           "primitives": n_prims, "pixels": int(sum(st.width * st.height for _, st in scenes)),
           "queries": wq, "batched_ms": wf_ms, "batched_mrays_s": wq / wf_ms / 1e3,
           "batched_kernel_ms": bs.kernel_ms, "batched_iterations": bs.iterations,
           "batched_config": cfg_dict(bs.used), "batched_pixels_finished_after_drain": bs.finished_pixels,
           "plain_sequential_ms": plain_ms, "plain_sequential_kernel_ms": pk,
           "plain_sequential_mrays_s": pq / plain_ms / 1e3, "speedup_batched_vs_plain": plain_ms / wf_ms,
           "max_rmse_batched_vs_plain": err, "bit_identical": same,
           "roofline": roofline(nbytes, bs.kernel_ms, "wavefront_batch64_16spp", "job",
                                "traversal latency and iteration barriers"),
           "plain_roofline": roofline(nbytes, pk, "k_render_batch64_16spp", "64 kernels",
                                      "fill: a small image cannot occupy the chip")}
    if not args.no_cpu and os.path.exists(HARNESS):
        # a bounded sample: the job's first 4 scenes, each rendered whole by the reference
        sample = list(range(4))
        tmp = tempfile.mkdtemp(prefix="pcl_bench_")
        meds, all_secs = [], []
        try:
            for k in sample:
                pa.HostScene.write_stress(tmp, f"b{k}", mat_mix=1, **specs[k])
                med, secs, _, _ = ref_render(os.path.join(tmp, f"b{k}.pecl"),
                                             ["--samples", str(wl.BATCH64["samples"]), "--depth",
                                              str(wl.BATCH64["depth"])], threads)
                meds.append(med)
                all_secs.append(secs)
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
        cq = sum(per_scene_q[k] for k in sample)
        desc_s = ", ".join(f"{specs[k]['width']}^2/{20 * specs[k]['freq'] ** 2} tris" for k in sample)
        out["cpu_baseline"] = cpu_entry(
            cq / sum(meds) / 1e6, "Mrays/s", threads, model,
            f"scenes {sample} of the job ({desc_s}) rendered whole, one after another, by Renderer::render of the "
            f"reference build: sum of per-scene medians of 5 after a warm-up ({sum(meds):.3f} s; {cq:,} queries "
            f"as the core counts them on the same scenes)", [round(sum(x), 4) for x in zip(*all_secs)])
    return rounded(out)


# ---------------------------------------------------------------------------
# inverse: adjoint gradient steps
# ---------------------------------------------------------------------------

def leg_inverse(core, pa, args, threads, model, steps=200):
    hs, st = wl.inverse(pa)
    desc, cam = hs.desc(), hs.camera()
    size, spp_side, bounces = st.width, st.samples, st.pathtrace_bounces
    sid = core.upload(desc)
    mats, n = pa.materials_of(desc)
    true = pa.material_params(mats, n)
    n_px = size * size * 3
    d_tgt = core.alloc(n_px * 4)
    try:
        cst, wc = core.render_counts(sid, cam, st, d_tgt)    # queries of one pass (the adjoint replays them)
        nbytes = alg_bytes(cst, wc)
        core.render_device(sid, cam, st, d_tgt)   # target from the true parameters
        x = np.random.default_rng(7).uniform(0.2, 0.8, size=true.shape)
        pa.set_material_params(mats, n, x)
        core.set_materials(sid, mats, n)
        # gradient against central differences at the start: emission with
        # h = 0.5 (the loss is quadratic in it: exact), albedo with h = 1e-3
        g0, l0, _ = core.adjoint_device(sid, cam, st, d_tgt)
        tgt = np.empty((size, size, 3), np.float32)
        core.d2h(tgt, d_tgt)
        tgt64 = tgt.astype(np.float64)
        fd_err = {}
        for (m, k, h) in [(2, 0, 1e-3), (3, 1, 1e-3), (6, 2, 1e-3), (0, 3, 0.5), (5, 4, 0.5), (4, 5, 0.5)]:
            vals = []
            for sgn in (1, -1):
                p = x.copy()
                p[m, k] = np.float32(x[m, k]) + sgn * np.float32(h)
                pa.set_material_params(mats, n, p)
                core.set_materials(sid, mats, n)
                img, _ = core.render(sid, cam, st)
                vals.append(float(np.sum((img.astype(np.float64) - tgt64) ** 2)))
            fd = (vals[0] - vals[1]) / (2 * h)
            fd_err[f"m{m}_{'albedo' if k < 3 else 'emission'}_{'rgb'[k % 3]}"] = {
                "adjoint": round(float(g0[m, k]), 4), "fd": round(fd, 4),
                "rel_err": float(abs(g0[m, k] - fd) / max(abs(fd), 1e-12))}
        # Which parameters does the image determine?  Every albedo and emission
        # term adds non-negative radiance, so the VJP with an all-ones image
        # adjoint (the column sums of dI/dp) at the true parameters is zero
        # exactly for a parameter that no pixel depends on.
        pa.set_material_params(mats, n, true)
        core.set_materials(sid, mats, n)
        d_one = core.alloc(n_px * 4)
        core.h2d(d_one, np.ones(n_px, np.float32))
        sens, _ = core.render_vjp_device(sid, cam, st, d_one)
        core.free(d_one)
        determined = sens != 0.0
        cond = conditioning(core, pa, sid, cam, st, mats, n, true) if args.inverse_conditioning else None
        pa.set_material_params(mats, n, x)
        core.set_materials(sid, mats, n)
        # projected Adam on the 48 parameters: albedos kept in [0, 1], emissions
        # >= 0 (physical ranges).  Steps 0.3 (albedo) and 1.0 (emission) with a
        # cosine decay to 1 % over the 200 steps: the best of the schedules
        # measured on this workload (constant 0.02 / 0.1 left the loss at 2.0e4)
        b1, b2, eps = 0.9, 0.999, 1e-8
        mom, vel = np.zeros_like(x), np.zeros_like(x)
        lr = np.empty_like(x)
        losses = []
        step_ms = []
        core.adjoint_device(sid, cam, st, d_tgt)   # warm-up
        t_start = time.perf_counter()
        for it in range(1, steps + 1):
            g, loss, s = core.adjoint_device(sid, cam, st, d_tgt)
            losses.append(loss)
            step_ms.append(s.kernel_ms)
            decay = 0.01 + 0.99 * 0.5 * (1.0 + np.cos(np.pi * (it - 1) / steps))
            lr[:, :3], lr[:, 3:] = 0.3 * decay, 1.0 * decay
            mom = b1 * mom + (1 - b1) * g
            vel = b2 * vel + (1 - b2) * g * g
            x = x - lr * (mom / (1 - b1 ** it)) / (np.sqrt(vel / (1 - b2 ** it)) + eps)
            x[:, :3] = np.clip(x[:, :3], 0.0, 1.0)
            x[:, 3:] = np.maximum(x[:, 3:], 0.0)
            pa.set_material_params(mats, n, x)
            core.set_materials(sid, mats, n)
        total = time.perf_counter() - t_start
        g_end, loss_end, _ = core.adjoint_device(sid, cam, st, d_tgt)
    finally:
        core.free(d_tgt)
        core.free_scene(sid)
    ms_step = total * 1000.0 / steps
    gpu_ms = float(np.mean(step_ms))
    log(f"inverse: {steps} Adam steps, {ms_step:.2f} ms/step ({steps / total:.2f} steps/s), "
        f"loss {losses[0]:.4g} -> {losses[-1]:.4g}")
    out = {"workload": f"Cornell box, {size}x{size}, {spp_side ** 2} spp, path tracing, {bounces} bounces, "
                       f"{n} materials x (albedo rgb + emission rgb) = {n * 6} parameters, L2 loss, projected Adam "
                       f"(albedo step 0.3 in [0, 1], emission step 1.0, >= 0, cosine decay to 1 %)",
           "steps": steps, "grad_steps_per_s": steps / total, "ms_per_step": ms_step,
           "gpu_ms_per_step": gpu_ms,
           "loss_first": losses[0], "loss_last": losses[-1], "loss_after_last_step": loss_end,
           "loss_at_100": losses[min(99, steps - 1)],
           "param_rmse_start_vs_true": None, "fd_check_at_start": fd_err,
           "fd_check_max_rel_err": max(v["rel_err"] for v in fd_err.values()),
           "roofline": roofline(nbytes, gpu_ms, "k_adjoint_cornell_512_64spp", "kernel",
                                "VALU issue and traversal latency of the replayed paths")}
    x0 = np.random.default_rng(7).uniform(0.2, 0.8, true.shape)
    out["param_rmse_last_vs_true"] = float(np.sqrt(np.mean((x - true) ** 2)))
    out["param_rmse_start_vs_true"] = float(np.sqrt(np.mean((x0 - true) ** 2)))
    # albedos (columns 0-2) and emissions (3-5) apart, over all parameters and
    # over the parameters the image determines
    names = [f"m{m}_{'albedo' if k < 3 else 'emission'}_{'rgb'[k % 3]}" for m in range(n) for k in range(6)]
    out["undetermined_params"] = [names[i] for i in np.flatnonzero(~determined.ravel())]
    out["undetermined_grad_zero_at_start_and_end"] = bool(np.all(g0[~determined] == 0) and
                                                          np.all(g_end[~determined] == 0))
    for name, cols in (("albedo", slice(0, 3)), ("emission", slice(3, 6))):
        e0, e1, dm = x0[:, cols] - true[:, cols], x[:, cols] - true[:, cols], determined[:, cols]
        out[f"{name}_rmse_start_vs_true"] = float(np.sqrt(np.mean(e0 ** 2)))
        out[f"{name}_rmse_last_vs_true"] = float(np.sqrt(np.mean(e1 ** 2)))
        out[f"determined_{name}_rmse_start"] = float(np.sqrt(np.mean(e0[dm] ** 2)))
        out[f"determined_{name}_rmse_last"] = float(np.sqrt(np.mean(e1[dm] ** 2)))
    out["abs_err_last_per_material"] = {f"m{m}": {"albedo": round(float(np.abs(x[m, :3] - true[m, :3]).max()), 5),
                                                  "emission": round(float(np.abs(x[m, 3:] - true[m, 3:]).max()), 5)}
                                        for m in range(n)}
    if cond is not None:
        out["conditioning_at_truth"] = cond
    if not args.no_cpu and os.path.exists(HARNESS):
        rows = (size // 2 - 8, size // 2 + 8)
        opts = ["--width", str(size), "--height", str(size), "--samples", str(spp_side), "--pathtrace", "--bounces",
                str(bounces)]
        med, secs, rays, _ = ref_render(wl.CORNELL, opts, threads, rows=rows)
        full = med * size / (rows[1] - rows[0])
        per_grad = full * (2 * n * 6 + 1)
        out["cpu_baseline"] = cpu_entry(
            1.0 / per_grad, "grad-steps/s", threads, model,
            f"rows {rows[0]}-{rows[1] - 1} of the {size}^2 {spp_side ** 2}-spp frame: render_tile_region of the "
            f"reference build, median of 5 after a warm-up ({med:.3f} s), scaled to the frame ({full:.2f} s) and to "
            f"the {2 * n * 6 + 1} renders of a central-difference gradient (one image + 2 per parameter)", secs)
    return rounded(out)


def conditioning(core, pa, sid, cam, st, mats, n, true):
    """Gauss-Newton matrix H = J^T J of the image at the true parameters, J from
    forward differences of renders with common random numbers (albedo step
    1e-2, emission step 1; channel c of the image depends only on the
    albedo / emission channel c, so H is three 2n x 2n blocks).  Reports per
    parameter sqrt(diag(pinv(H))) -- how far the parameter can move along the
    least-determined direction per unit of image error -- and each block's
    eigenvalue range."""
    N = st.width * st.height

    def render(p):
        pa.set_material_params(mats, n, p)
        core.set_materials(sid, mats, n)
        img, _ = core.render(sid, cam, st)
        return img.reshape(N, 3).astype(np.float64)

    base = render(true)
    J = np.zeros((3, N, 2 * n), np.float32)
    for m in range(n):
        for k in range(6):
            h = 1e-2 if k < 3 else 1.0
            p = true.copy()
            p[m, k] += h
            c = k % 3
            J[c, :, m if k < 3 else n + m] = (render(p)[:, c] - base[:, c]) / h
    out = {"method": "J^T J from forward differences at the truth (albedo h 1e-2, emission h 1), per channel"}
    crb = np.zeros((n, 6))
    for c in range(3):
        H = J[c].astype(np.float64).T @ J[c].astype(np.float64)
        w = np.linalg.eigvalsh(H)
        sd = np.sqrt(np.maximum(np.diag(np.linalg.pinv(H, rcond=1e-12)), 0.0))
        crb[:, c], crb[:, 3 + c] = sd[:n], sd[n:]
        out[f"eig_{'rgb'[c]}"] = [float(w.min()), float(w.max())]
    out["sqrt_diag_pinv_H"] = {f"m{m}": {"albedo": [round(float(v), 5) for v in crb[m, :3]],
                                         "emission": [round(float(v), 5) for v in crb[m, 3:]]} for m in range(n)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--plain-only", action="store_true", help="skip the batched leg of the headline workload")
    ap.add_argument("--legs", default="cornell_pt,batch64,inverse",
                    help="further workloads to run after the headline one (comma list, or 'none'); "
                         "glass_defer (not run by default) times glass vertices in place vs deferred")
    ap.add_argument("--pt-samples", type=int, default=wl.CORNELL_PT["samples"],
                    help="samples per side for the cornell_pt leg (16 = the workload's 256 spp; less for profiling)")
    ap.add_argument("--batched-config", default="",
                    help="slots,blocks_per_cu,stack_k,iters_per_graph[,drain_below[,inline_shadows[,streams]]] for the batched legs "
                         "instead of autotuning")
    ap.add_argument("--inverse-conditioning", type=int, default=1,
                    help="1: the inverse leg also reports the Gauss-Newton conditioning at the true parameters")
    ap.add_argument("--skip-headline", action="store_true",
                    help="run only the legs (profiling runs); prints the legs' JSON, not the bench line")
    ap.add_argument("--no-live-counters", action="store_true",
                    help="skip the rocprofv3 FETCH_SIZE / WRITE_SIZE passes that measure the headline frame's HBM "
                         "traffic in this run (the rooflines then keep the committed profiles' traffic)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain 256 MiB read + 256 MiB write elementwise kernel (rocprof PMC calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        x = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        log("copy probe: 3 x (268435456 B read + 268435456 B written)")
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(the path would shard pixels or scenes across GPUs, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import pencil_amd as pa
    core = pa.Core(0)
    threads, model = host_cpu()
    log(f"host: {threads} cores usable ({os.cpu_count()} online), {model}")
    r = None if args.skip_headline else leg_stress(core, pa, args)
    legs = {}
    for name in [x for x in args.legs.split(",") if x and x != "none"]:
        fn = {"cornell_pt": leg_cornell_pt, "batch64": leg_batch64, "inverse": leg_inverse,
              "glass_defer": leg_glass_defer}[name]
        t0 = time.time()
        legs[name] = fn(core, pa, args, threads, model)
        log(f"leg {name}: {time.time() - t0:.1f} s")
    core.close()
    if r is None:
        print(json.dumps({"legs": legs}), flush=True)
        return
    cpu = None
    if not args.no_cpu:
        if os.path.exists(HARNESS):
            t0 = time.time()
            cpu = cpu_stress(r["queries"], threads, model)
            log(f"cpu baseline (reference): {cpu['value']} Mrays/s, {time.time() - t0:.1f} s")
        else:
            log(f"cpu baseline: {HARNESS} not built (make -C oracle ref); reported as missing")

    value, cst, info, batched = r["value"], r["cst"], r["info"], r["batched"]
    live_plain = live_batched = None
    if not args.no_live_counters:
        t0 = time.time()
        bc = batched["config"] if batched else {"slots": 0, "blocks_per_cu": 0, "stack_k": 0, "iters_per_graph": 0,
                                                 "drain_below": 0, "inline_shadows": 0, "streams": 0}
        cfg = ",".join(str(bc[k]) for k in ("slots", "blocks_per_cu", "stack_k", "iters_per_graph", "drain_below",
                                            "inline_shadows", "streams"))
        live_plain, live_batched = live_traffic(cfg)
        if live_plain:
            log(f"live HBM counters: plain frame {live_plain['hbm_bytes'] / 1e6:.1f} MB, batched frame "
                f"{live_batched['hbm_bytes'] / 1e6:.1f} MB ({time.time() - t0:.1f} s)")
    if batched and live_batched:
        batched["roofline"] = rounded(roofline(r["bytes"], batched["kernel_ms"], "wavefront_bvh_stress_1080p", "job",
                                               batched["roofline"]["limiter"], live=live_batched))
    line = {
        "metric": "Mrays/s at 1080p 1M-tri (plain vs batched); adjoint grad-steps/s at FD tol",
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded bvh_stress scene)",
        "config": {"workload": "bvh_stress: 1,048,820-tri displaced icosphere + 4,096 spheres, 1920x1080, "
                               "1 spp, depth 1, primary + 1 shadow ray (plain per-pixel megakernel)",
                   "primitives": r["desc_n"], "width": r["W"], "height": r["H"], "spp": 1,
                   "queries_per_frame": r["queries"], "ref_convention_rays_per_frame": cst.rays_traced,
                   "ref_convention_mrays_s": round(cst.rays_traced / (r["ms_per_step"] / 1000.0) / 1e6, 3),
                   "kernel_ms": round(r["kernel_ms"], 4), "bvh_build_ms": round(info.build_ms, 1),
                   "path": "plain",
                   "plain_mrays_s": round(value, 3),
                   "batched": rounded(batched) if batched else None,
                   "device_bvh_build": rounded(r["lbvh"]),
                   "adjoint_grad_steps_per_s": (round(legs["inverse"]["grad_steps_per_s"], 3)
                                                if "inverse" in legs else None)},
        "roofline": roofline(r["bytes"], r["kernel_ms"], "k_render_bvh_stress_1080p", "kernel",
                             "dependent node-fetch latency of the traversal", live=live_plain),
        "cpu_baseline": cpu,
        "legs": legs,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
