"""The batched wavefront tracer against the per-pixel megakernel and the reference.

Both paths run the same per-pixel program with the same float operations in
the same order, so their images must be identical bit for bit; against the
reference fixtures the same bars as the megakernel hold.
"""
import json

import numpy as np
import pytest

import pencil_amd as pa
from conftest import SCENES, golden, img_stats
from test_gpu_parity import assert_reference_bar
from test_oracle import HDR_FIXTURES, render_fixture_inputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fx_name", HDR_FIXTURES)
def test_wavefront_equals_megakernel_and_reference(core, fx_name):
    fx, hs, s = render_fixture_inputs(fx_name)
    sid = core.upload(hs.desc())
    try:
        mk, mst = core.render(sid, hs.camera(), s)
        wf, wst = core.render_wavefront(sid, hs.camera(), s)
    finally:
        core.free_scene(sid)
    assert np.array_equal(wf, mk)
    assert wst.rays_traced == mst.rays_traced
    assert wst.closest_queries == mst.closest_queries and wst.anyhit_queries == mst.anyhit_queries
    assert_reference_bar(fx_name, img_stats(wf, fx["hdr"]), path=fx_name.startswith(("pt_", "mis_")))


@pytest.mark.parametrize("cfg", [dict(slots=1000, blocks_per_cu=1, stack_k=8, iters_per_graph=1),
                                 dict(slots=4096, blocks_per_cu=16, stack_k=32, iters_per_graph=3),
                                 dict(slots=1000, blocks_per_cu=2, stack_k=16, iters_per_graph=2, drain_below=-1),
                                 dict(slots=1000, blocks_per_cu=2, stack_k=16, iters_per_graph=1,
                                      drain_below=1 << 30),
                                 dict(slots=2000, blocks_per_cu=4, stack_k=8, iters_per_graph=2, inline_shadows=-1),
                                 dict(slots=2000, blocks_per_cu=4, stack_k=32, iters_per_graph=2, inline_shadows=32)])
def test_wavefront_knobs_do_not_change_results(core, cfg):
    hs = pa.HostScene.from_pecl(golden("glass.pecl"), width=64, height=48, samples=2)
    sid = core.upload(hs.desc())
    try:
        mk, _ = core.render(sid, hs.camera(), hs.settings())
        wf, st = core.render_wavefront(sid, hs.camera(), hs.settings(), pa.WfConfig(**cfg))
        assert st.used.slots == min(cfg["slots"], 64 * 48) and st.used.stack_k == cfg["stack_k"]
    finally:
        core.free_scene(sid)
    assert np.array_equal(wf, mk)


@pytest.mark.parametrize("slots,drain,inline", [(1920, -1, -1), (1920, 700, 1), (1920, 1 << 30, -1),
                                                (500, 300, 1), (500, 1 << 30, 32)])
def test_drain_finishes_pixels_bit_for_bit(core, slots, drain, inline):
    """Path tracing with Russian roulette (uneven sample chains).  Whatever
    the drain threshold -- also with fewer slots than pixels, where the drain
    waits until every pixel is taken -- the pixels the finish launch
    completes give the megakernel's image and counts."""
    hs = pa.HostScene.from_pecl(golden("cornell_box.pecl"), width=48, height=40, samples=2)
    s = hs.settings()
    s.pathtrace, s.pathtrace_bounces = 1, 8
    sid = core.upload(hs.desc())
    try:
        mk, mst = core.render(sid, hs.camera(), s)
        cfg = pa.WfConfig(slots=slots, blocks_per_cu=2, stack_k=16, iters_per_graph=1, drain_below=drain,
                          inline_shadows=inline)
        wf, wst = core.render_wavefront(sid, hs.camera(), s, cfg)
    finally:
        core.free_scene(sid)
    assert np.array_equal(wf, mk)
    assert wst.rays_traced == mst.rays_traced
    assert (wst.closest_queries, wst.anyhit_queries) == (mst.closest_queries, mst.anyhit_queries)
    assert (wst.used.drain_below, wst.used.inline_shadows) == (drain, inline)
    if drain < 0:
        assert wst.finished_pixels == 0
    elif drain > slots:
        assert 0 < wst.finished_pixels <= slots


@pytest.mark.parametrize("streams", [1, 2, 3, 4])
def test_concurrent_parts_do_not_change_results(core, streams):
    """The job's tiles dealt to 1..4 engines on their own streams: path tracing
    with Russian roulette and a drain in every part, plus a one-tile item, so
    that some parts get no tile of it and a one-tile job runs as one part."""
    hs = pa.HostScene.from_pecl(golden("cornell_box.pecl"), width=72, height=40, samples=2)
    s = hs.settings()
    s.pathtrace, s.pathtrace_bounces = 1, 8
    small = pa.HostScene.from_pecl(golden("shapes.pecl"), width=8, height=1, samples=2)
    sids, bufs = [], []
    try:
        jobs, refs, stats = [], [], []
        for h, st in ((hs, s), (small, small.settings())):
            sid = core.upload(h.desc())
            sids.append(sid)
            img, mst = core.render(sid, h.camera(), st)
            d = core.alloc(st.width * st.height * 12)
            bufs.append((d, st))
            jobs.append((sid, h.camera(), st, d))
            refs.append(img)
            stats.append(mst)
        cfg = pa.WfConfig(slots=1500, blocks_per_cu=2, stack_k=16, iters_per_graph=2, drain_below=900,
                          streams=streams)
        bs = core.render_batch(jobs, cfg)
        assert bs.used.streams == streams
        assert bs.rays_traced == sum(m.rays_traced for m in stats)
        assert bs.closest_queries == sum(m.closest_queries for m in stats)
        assert bs.anyhit_queries == sum(m.anyhit_queries for m in stats)
        for (d, st), ref in zip(bufs, refs):
            out = np.empty((st.height, st.width, 3), dtype=np.float32)
            core.d2h(out, d)
            assert np.array_equal(out, ref)
        one, _ = core.render_wavefront(sids[1], small.camera(), small.settings(), pa.WfConfig(streams=streams))
        assert np.array_equal(one, refs[1])
        _, ost = core.render_wavefront(sids[1], small.camera(), small.settings(), pa.WfConfig(streams=streams))
        assert ost.used.streams == 1   # one tile: one part
    finally:
        for d, _ in bufs:
            core.free(d)
        for sid in sids:
            core.free_scene(sid)


def test_streams_out_of_range_is_an_error(core):
    hs = pa.HostScene.from_pecl(golden("shapes.pecl"), width=16, height=16, samples=1)
    sid = core.upload(hs.desc())
    try:
        for bad in (-1, 5):
            with pytest.raises(pa.CoreError, match="streams"):
                core.render_wavefront(sid, hs.camera(), hs.settings(), pa.WfConfig(streams=bad))
    finally:
        core.free_scene(sid)


def test_default_parts_follow_the_job(core):
    """streams 0: one part for a one-sample job, two for Whitted with more
    samples, three for path tracing with >= 16 spp."""
    hs = pa.HostScene.from_pecl(golden("three_spheres.pecl"), width=256, height=64, samples=1)
    sid = core.upload(hs.desc())
    try:
        for samples, pt, want in ((1, 0, 1), (2, 0, 2), (4, 0, 2), (2, 1, 2), (4, 1, 3)):
            s = hs.settings()
            s.samples, s.pathtrace, s.pathtrace_bounces = samples, pt, 3
            mk, _ = core.render(sid, hs.camera(), s)
            wf, st = core.render_wavefront(sid, hs.camera(), s)
            assert st.used.streams == want, (samples, pt)
            assert np.array_equal(wf, mk)
    finally:
        core.free_scene(sid)


def test_batch_of_uneven_scenes_equals_individual_renders(core):
    """Several scenes of different size, resolution and mode in one job."""
    jobs, refs, bufs, sids = [], [], [], []
    specs = [("three_spheres", dict(width=80, height=60, samples=2), False),
             ("shapes", dict(width=33, height=17, samples=1), False),
             ("cornell_box", dict(width=64, height=64, samples=2), True),
             ("glass", dict(width=50, height=40, samples=1), False)]
    stress = pa.HostScene.stress(freq=30, n_spheres=200, seed=5, width=96, height=54)
    try:
        for name, kw, pt in specs:
            hs = pa.HostScene.from_pecl(golden(name + ".pecl"), **kw)
            s = hs.settings()
            if pt:
                s.pathtrace, s.pathtrace_bounces = 1, 6
            sid = core.upload(hs.desc())
            sids.append(sid)
            img, _ = core.render(sid, hs.camera(), s)
            d = core.alloc(s.width * s.height * 12)
            bufs.append((d, s))
            jobs.append((sid, hs.camera(), s, d))
            refs.append(img)
        sid = core.upload(stress.desc())
        sids.append(sid)
        s = stress.settings()
        img, _ = core.render(sid, stress.camera(), s)
        d = core.alloc(s.width * s.height * 12)
        bufs.append((d, s))
        jobs.append((sid, stress.camera(), s, d))
        refs.append(img)
        bs = core.render_batch(jobs)
        assert bs.iterations > 0
        for (d, s), ref in zip(bufs, refs):
            out = np.empty((s.height, s.width, 3), dtype=np.float32)
            core.d2h(out, d)
            assert np.array_equal(out, ref)
    finally:
        for d, _ in bufs:
            core.free(d)
        for sid in sids:
            core.free_scene(sid)


@pytest.mark.parametrize("inline", [1, 4, 32])
@pytest.mark.parametrize("depth", [1, 6])
def test_pixels_finished_in_the_closest_hit_kernel(core, inline, depth):
    """In one-sample Whitted jobs with inline shadow rays the closest-hit
    kernel finishes a pixel whose vertex ends it (no child, no frame waiting,
    every shadow ray inline) and writes it out; the others go through
    k_resolve.  Several scenes in one job (hard and soft lights, metal, glass,
    DOF, a stress mesh) and the stress scene alone, depth 1 and 6: images and
    counts are the megakernel's."""
    jobs, refs, bufs, sids, mstats = [], [], [], [], []
    specs = [("three_spheres", dict(width=40, height=30, samples=1)), ("hard_shadows", dict(width=36, height=20, samples=1)),
             ("soft_shadows", dict(width=30, height=24, samples=1)), ("dof", dict(width=24, height=24, samples=1)),
             ("shapes", dict(width=33, height=17, samples=1))]
    scenes = [pa.HostScene.from_pecl(golden(n + ".pecl"), **kw) for n, kw in specs]
    scenes.append(pa.HostScene.stress(freq=30, n_spheres=200, seed=5, width=96, height=54))
    try:
        for hs in scenes:
            s = hs.settings()
            s.max_depth = depth
            sid = core.upload(hs.desc())
            sids.append(sid)
            img, mst = core.render(sid, hs.camera(), s)
            d = core.alloc(s.width * s.height * 12)
            bufs.append((d, s))
            jobs.append((sid, hs.camera(), s, d))
            refs.append(img)
            mstats.append(mst)
        for one in (False, True):
            sel = range(len(jobs)) if not one else [len(jobs) - 1]
            cfg = pa.WfConfig(slots=20000, blocks_per_cu=2, stack_k=16, iters_per_graph=1, inline_shadows=inline)
            bs = core.render_batch([jobs[i] for i in sel], cfg)
            assert bs.used.inline_shadows == inline
            assert bs.rays_traced == sum(mstats[i].rays_traced for i in sel)
            assert bs.closest_queries == sum(mstats[i].closest_queries for i in sel)
            assert bs.anyhit_queries == sum(mstats[i].anyhit_queries for i in sel)
            for i in sel:
                d, s = bufs[i]
                out = np.empty((s.height, s.width, 3), dtype=np.float32)
                core.d2h(out, d)
                assert np.array_equal(out.view(np.uint32), refs[i].view(np.uint32)), specs[i][0] if i < len(specs) else "stress"
    finally:
        for d, _ in bufs:
            core.free(d)
        for sid in sids:
            core.free_scene(sid)


def test_autotune_returns_a_working_config(core):
    hs = pa.HostScene.from_pecl(golden("shapes.pecl"), width=96, height=72, samples=1)
    sid = core.upload(hs.desc())
    d = core.alloc(96 * 72 * 12)
    try:
        best, ms = core.autotune([(sid, hs.camera(), hs.settings(), d)])
        assert best.stack_k in (8, 16, 32) and best.slots > 0 and ms > 0
        assert best.drain_below == 0 and best.streams == 0   # left to the rules
        mk, _ = core.render(sid, hs.camera(), hs.settings())
        wf, _ = core.render_wavefront(sid, hs.camera(), hs.settings(), best)
        assert np.array_equal(wf, mk)
    finally:
        core.free(d)
        core.free_scene(sid)


def test_wavefront_stress_full_frame_matches_megakernel(core):
    fx = np.load(golden("bvh_stress.npz"))
    p = json.loads(str(fx["full_params"]))
    hs = pa.HostScene.stress(**p)
    sid = core.upload(hs.desc())
    try:
        mk, mst = core.render(sid, hs.camera(), hs.settings())
        wf, wst = core.render_wavefront(sid, hs.camera(), hs.settings())
        assert np.array_equal(wf, mk)
        assert (wst.closest_queries, wst.anyhit_queries) == (mst.closest_queries, mst.anyhit_queries)
    finally:
        core.free_scene(sid)
