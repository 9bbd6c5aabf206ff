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

cpu_baseline: the reference renderer (oracle/_ref/pencil, built from the
reference sources) on the same scene written as .pecl + .obj, all host cores
used (capped at 16), render() time as the reference prints it; falls back to
the oracle port if the reference binary was not built.

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
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PRIM_BYTES = [16, 48, 16, 32, 48, 48]   # sphere, triangle, cube, plane, cylinder, cone
NODE_BYTES = 112                         # a 4-wide node read: 6 x 16 B of child boxes + 16 B of refs
HBM_PEAK_GBS = 8000.0
STRESS = dict(freq=229, n_spheres=4096, seed=1234, width=1920, height=1080)
CORNELL = os.path.join(ROOT, "pencil_amd", "scenes", "cornell_c5.pecl")
MESH_IN_BOX = dict(freq=229, seed=1234, scale=0.25, center=(0.0, 1.75, 0.5))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cfg_dict(c):
    return {"slots": c.slots, "blocks_per_cu": c.blocks_per_cu, "stack_k": c.stack_k,
            "iters_per_graph": c.iters_per_graph, "drain_below": c.drain_below,
            "inline_shadows": c.inline_shadows}


def cfg_str(c):
    return (f"slots {c.slots}, blocks/CU {c.blocks_per_cu}, stack {c.stack_k}, iters/graph {c.iters_per_graph}, "
            f"drain below {c.drain_below}, inline shadows {c.inline_shadows}")


def cpu_threads():
    return min(16, os.cpu_count() or 1)


def oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    return oracle_lib


def cpu_baseline(queries, threads):
    """Reference renderer on the bvh_stress frame; returns dict or None."""
    ref = os.path.join(ROOT, "oracle", "_ref", "pencil")
    import pencil_amd as pa
    tmp = tempfile.mkdtemp(prefix="pcl_bench_")
    try:
        if os.path.exists(ref):
            pa.HostScene.write_stress(tmp, "stress", **STRESS)
            out = os.path.join(tmp, "out.ppm")
            t0 = time.time()
            r = subprocess.run([ref, os.path.join(tmp, "stress.pecl"), out, "--threads", str(threads)],
                               capture_output=True, text=True, timeout=600)
            wall = time.time() - t0
            m = re.search(r"rendered in ([0-9.]+) s \(([0-9,]+) rays", r.stdout)
            if r.returncode != 0 or not m:
                log("reference run failed:", r.returncode, r.stderr[-500:])
                return None
            sec = float(m.group(1))
            return {"value": round(queries / sec / 1e6, 3), "unit": "Mrays/s", "cores": threads,
                    "kind": "reference",
                    "sample": f"full 1920x1080 frame, reference binary built from the reference sources, "
                              f"render() {sec:.3f} s as it prints it ({m.group(2)} reference-convention rays; "
                              f"{wall:.1f} s wall incl. .pecl/.obj parse and BVH build)"}
        ol = oracle()
        hs = pa.HostScene.stress(**STRESS)
        osc = ol.OracleScene(hs.desc())
        s = hs.settings()
        rows = (0, 270)   # a quarter of the frame
        t0 = time.time()
        _, c = osc.render(hs.camera(), s, threads=threads, rows=rows)
        sec = time.time() - t0
        q = c.closest_queries + c.anyhit_queries
        return {"value": round(q / sec / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
                "sample": f"rows 0-269 of the 1920x1080 frame through the oracle ({q} queries, {sec:.2f} s)"}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def rmse(a, b):
    return float(np.sqrt(np.mean((a.astype(np.float64) - b) ** 2)))


def rounded(d):
    return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items()}


# ---------------------------------------------------------------------------
# headline: bvh_stress
# ---------------------------------------------------------------------------

def leg_stress(core, pa, args):
    t0 = time.time()
    hs = pa.HostScene.stress(**STRESS)
    desc, cam, st = hs.desc(), hs.camera(), hs.settings()
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
    tests = [wc.prim_tests[k] for k in range(6)]
    bytes_per_launch = queries * (32 + 8) + NODE_BYTES * wc.node_visits + sum(b * t for b, t in zip(PRIM_BYTES, tests))
    log(f"counts: {cst.closest_queries} closest + {cst.anyhit_queries} any-hit queries, "
        f"{wc.node_visits} node visits, tests {tests}, {bytes_per_launch / 1e9:.2f} GB algorithmic per frame")

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
    achieved = bytes_per_launch / (kernel_ms / 1000.0) / 1e9
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
        batched = {"ms_per_step": wms, "kernel_ms": sum(wkms) / len(wkms), "queries": wq,
                   "mrays": wq / (wms / 1000.0) / 1e6, "rmse_vs_plain": rmse(a, b), "iterations": bs.iterations,
                   "config": cfg_dict(bs.used)}
        log(f"batched: {wms:.3f} ms/frame wall, {batched['kernel_ms']:.3f} ms on the stream, "
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
    return dict(value=value, ms_per_step=ms_per_step, kernel_ms=kernel_ms, achieved=achieved, queries=queries,
                cst=cst, info=info, desc_n=desc.n_prims, W=W, H=H, batched=batched, lbvh=lbvh)


# ---------------------------------------------------------------------------
# cornell_pt: path tracing, plain vs batched
# ---------------------------------------------------------------------------

def leg_cornell_pt(core, pa, args, threads):
    spp_side = args.pt_samples
    hs = pa.HostScene.cornell_mesh(CORNELL, width=1920, height=1080, **MESH_IN_BOX)
    desc, cam, st = hs.desc(), hs.camera(), hs.settings()
    st.samples, st.pathtrace, st.pathtrace_bounces = spp_side, 1, 8
    sid = core.upload(desc)
    W, H = st.width, st.height
    d_a = core.alloc(W * H * 12)
    d_b = core.alloc(W * H * 12)
    try:
        t0 = time.perf_counter()
        ps = core.render_device(sid, cam, st, d_a)
        plain_ms = (time.perf_counter() - t0) * 1000.0
        q = ps.closest_queries + ps.anyhit_queries
        log(f"cornell_pt plain: {plain_ms:.1f} ms, {q} queries, {q / plain_ms / 1e3:.1f} Mrays/s")
        if args.batched_config:
            cfg = pa.WfConfig(*[int(x) for x in args.batched_config.split(",")])
        else:
            # tuned on the same frame at 2x2 spp (a few seconds), rendered at full spp
            st_t = pa.Settings.from_buffer_copy(st)
            st_t.samples = 2
            cfg, tune_ms = core.autotune([(sid, cam, st_t, d_b)])
            log(f"cornell_pt autotune at 2x2 spp: {cfg_str(cfg)} -> {tune_ms:.1f} ms")
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
    out = {"workload": f"Cornell box + {desc.n_prims - 14}-tri mesh, {W}x{H}, {spp_side ** 2} spp, path tracing, "
                       f"8 bounces, Russian roulette after 3",
           "primitives": desc.n_prims, "queries": q, "ref_convention_rays": ps.rays_traced,
           "plain_ms": plain_ms, "plain_mrays_s": q / plain_ms / 1e3,
           "batched_ms": wf_ms, "batched_mrays_s": wq / wf_ms / 1e3, "batched_kernel_ms": bs.kernel_ms,
           "batched_iterations": bs.iterations, "batched_config": cfg_dict(bs.used),
           "batched_pixels_finished_after_drain": bs.finished_pixels,
           "speedup_batched_vs_plain": plain_ms / wf_ms, "rmse_batched_vs_plain": rmse(a, b),
           "image_mean": float(a.mean())}
    if not args.no_cpu:
        ol = oracle()
        osc = ol.OracleScene(desc)
        rows = (H // 2 - 2, H // 2 + 2)
        t0 = time.time()
        _, c = osc.render(cam, st, threads=threads, rows=rows)
        sec = time.time() - t0
        cq = c.closest_queries + c.anyhit_queries
        out["cpu_baseline"] = {"value": round(cq / sec / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
                               "sample": f"rows {rows[0]}-{rows[1] - 1} (the middle of the frame) at 256 spp "
                                         f"through the oracle ({cq} queries, {sec:.2f} s)"}
    return rounded(out)


# ---------------------------------------------------------------------------
# batch64: 64 uneven scenes in one batched job vs 64 plain renders
# ---------------------------------------------------------------------------

def batch64_specs(n=64, seed=2024):
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(n):
        tris = float(np.exp(rng.uniform(np.log(16), np.log(500_000))))
        freq = max(1, int(round(np.sqrt(tris / 20.0))))
        res = int(rng.choice([256, 512, 1024]))
        specs.append(dict(freq=freq, n_spheres=int(rng.integers(8, 257)), seed=int(seed * 1000 + k),
                          width=res, height=res))
    return specs


def leg_batch64(core, pa, args, threads):
    specs = batch64_specs()
    t0 = time.time()
    scenes = []
    for sp in specs:
        hs = pa.HostScene.mixed(**sp)
        st = hs.settings()
        st.samples, st.max_depth, st.pathtrace = 4, 6, 0
        scenes.append((hs, st))
    sids = [core.upload(hs.desc()) for hs, _ in scenes]
    n_prims = sum(hs.desc().n_prims for hs, _ in scenes)
    log(f"batch64: {len(scenes)} scenes, {n_prims} primitives, built and uploaded in {time.time() - t0:.1f} s")
    outs = [core.alloc(st.width * st.height * 12) for _, st in scenes]
    outs2 = [core.alloc(st.width * st.height * 12) for _, st in scenes]
    try:
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
        pq = 0
        for sid, (hs, st), d in zip(sids, scenes, outs2):
            s = core.render_device(sid, hs.camera(), st, d)
            pq += s.closest_queries + s.anyhit_queries
        plain_ms = (time.perf_counter() - t0) * 1000.0
        err = 0.0
        for (hs, st), d, d2 in zip(scenes, outs, outs2):
            a = np.empty((st.height, st.width, 3), np.float32)
            b = np.empty_like(a)
            core.d2h(a, d)
            core.d2h(b, d2)
            err = max(err, rmse(a, b))
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
           "plain_sequential_ms": plain_ms, "plain_sequential_mrays_s": pq / plain_ms / 1e3,
           "max_rmse_batched_vs_plain": err}
    if not args.no_cpu:
        ol = oracle()
        small = [i for i, (_, st) in enumerate(scenes) if st.width == 256][:2]
        t0 = time.time()
        cq = 0
        for i in small:
            hs, st = scenes[i]
            _, c = ol.OracleScene(hs.desc()).render(hs.camera(), st, threads=threads)
            cq += c.closest_queries + c.anyhit_queries
        sec = time.time() - t0
        out["cpu_baseline"] = {"value": round(cq / sec / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
                               "sample": f"scenes {small} (256^2) rendered whole through the oracle, incl. its BVH "
                                         f"build ({cq} queries, {sec:.2f} s)"}
    return rounded(out)


# ---------------------------------------------------------------------------
# inverse: adjoint gradient steps
# ---------------------------------------------------------------------------

def leg_inverse(core, pa, args, threads, steps=200):
    size, spp_side, bounces = 512, 8, 4
    hs = pa.HostScene.from_pecl(CORNELL, size, size, spp_side, 4)
    desc, cam, st = hs.desc(), hs.camera(), hs.settings()
    st.pathtrace, st.pathtrace_bounces = 1, bounces
    sid = core.upload(desc)
    mats, n = pa.materials_of(desc)
    true = pa.material_params(mats, n)
    n_px = size * size * 3
    d_tgt = core.alloc(n_px * 4)
    try:
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
        pa.set_material_params(mats, n, x)
        core.set_materials(sid, mats, n)
        # projected Adam on the 48 parameters: albedos kept in [0, 1], emissions
        # >= 0 (physical ranges); emission steps of 0.1 so the panel's 15 is
        # reachable in 200 steps, albedo steps of 0.02
        lr = np.empty_like(x)
        lr[:, :3], lr[:, 3:] = 0.02, 0.1
        b1, b2, eps = 0.9, 0.999, 1e-8
        mom, vel = np.zeros_like(x), np.zeros_like(x)
        losses = []
        step_ms = []
        core.adjoint_device(sid, cam, st, d_tgt)   # warm-up
        t_start = time.perf_counter()
        for it in range(1, steps + 1):
            g, loss, s = core.adjoint_device(sid, cam, st, d_tgt)
            losses.append(loss)
            step_ms.append(s.kernel_ms)
            mom = b1 * mom + (1 - b1) * g
            vel = b2 * vel + (1 - b2) * g * g
            x = x - lr * (mom / (1 - b1 ** it)) / (np.sqrt(vel / (1 - b2 ** it)) + eps)
            x[:, :3] = np.clip(x[:, :3], 0.0, 1.0)
            x[:, 3:] = np.maximum(x[:, 3:], 0.0)
            pa.set_material_params(mats, n, x)
            core.set_materials(sid, mats, n)
        total = time.perf_counter() - t_start
    finally:
        core.free(d_tgt)
        core.free_scene(sid)
    ms_step = total * 1000.0 / steps
    log(f"inverse: {steps} Adam steps, {ms_step:.2f} ms/step ({steps / total:.2f} steps/s), "
        f"loss {losses[0]:.4g} -> {losses[-1]:.4g}")
    out = {"workload": f"Cornell box, {size}x{size}, {spp_side ** 2} spp, path tracing, {bounces} bounces, "
                       f"{n} materials x (albedo rgb + emission rgb) = {n * 6} parameters, L2 loss, projected Adam "
                       f"(albedo step 0.02 in [0, 1], emission step 0.1, >= 0)",
           "steps": steps, "grad_steps_per_s": steps / total, "ms_per_step": ms_step,
           "gpu_ms_per_step": float(np.mean(step_ms)),
           "loss_first": losses[0], "loss_last": losses[-1], "loss_at_100": losses[min(99, steps - 1)],
           "param_rmse_start_vs_true": None, "fd_check_at_start": fd_err,
           "fd_check_max_rel_err": max(v["rel_err"] for v in fd_err.values())}
    x0 = np.random.default_rng(7).uniform(0.2, 0.8, true.shape)
    out["param_rmse_last_vs_true"] = float(np.sqrt(np.mean((x - true) ** 2)))
    out["param_rmse_start_vs_true"] = float(np.sqrt(np.mean((x0 - true) ** 2)))
    # albedos (columns 0-2) and emissions (3-5) apart
    for name, cols in (("albedo", slice(0, 3)), ("emission", slice(3, 6))):
        out[f"{name}_rmse_start_vs_true"] = float(np.sqrt(np.mean((x0[:, cols] - true[:, cols]) ** 2)))
        out[f"{name}_rmse_last_vs_true"] = float(np.sqrt(np.mean((x[:, cols] - true[:, cols]) ** 2)))
    if not args.no_cpu:
        ol = oracle()
        osc = ol.OracleScene(desc)
        rows = (size // 2 - 8, size // 2 + 8)
        t0 = time.time()
        osc.render(cam, st, threads=threads, rows=rows)
        sec = time.time() - t0
        full = sec * size / (rows[1] - rows[0])
        per_grad = full * (2 * n * 6 + 1)
        out["cpu_baseline"] = {"value": round(1.0 / per_grad, 6), "unit": "grad-steps/s", "cores": threads,
                               "kind": "port",
                               "sample": f"rows {rows[0]}-{rows[1] - 1} of one 512^2 64-spp forward render through "
                                         f"the oracle ({sec:.2f} s), scaled to the frame and to the "
                                         f"{2 * n * 6 + 1} renders of a central-difference gradient"}
    return rounded(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--plain-only", action="store_true", help="skip the batched leg of the headline workload")
    ap.add_argument("--legs", default="cornell_pt,batch64,inverse",
                    help="further workloads to run after the headline one (comma list, or 'none')")
    ap.add_argument("--pt-samples", type=int, default=16,
                    help="samples per side for the cornell_pt leg (16 = the workload's 256 spp; less for profiling)")
    ap.add_argument("--batched-config", default="",
                    help="slots,blocks_per_cu,stack_k,iters_per_graph[,drain_below[,inline_shadows]] for the batched legs "
                         "instead of autotuning")
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
    threads = cpu_threads()
    r = leg_stress(core, pa, args)
    legs = {}
    for name in [x for x in args.legs.split(",") if x and x != "none"]:
        fn = {"cornell_pt": leg_cornell_pt, "batch64": leg_batch64, "inverse": leg_inverse}[name]
        t0 = time.time()
        legs[name] = fn(core, pa, args, threads)
        log(f"leg {name}: {time.time() - t0:.1f} s")
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(r["queries"], threads)
        except Exception as e:   # the baseline is reported, never required
            log("cpu baseline failed:", e)
    core.close()

    traffic = None   # HBM bytes per dispatch from the committed rocprofv3 PMC summary of this kernel
    tpath = os.path.join(ROOT, "profiles", "k_render_bvh_stress_1080p.traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = round(json.load(f)["hbm_bytes_per_dispatch"])

    value, cst, info, batched = r["value"], r["cst"], r["info"], r["batched"]
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
        "roofline": {"bound": "hbm", "achieved": round(r["achieved"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(r["achieved"] / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
        "legs": legs,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
