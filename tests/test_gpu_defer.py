"""Deferred glass vertices in the path-traced megakernel and in the batched
tracer's finish launch.

A path that hits glass records the vertex as a task with its generator
state, advances the generator past the draws of the Whitted subtree (soft
lights: counted by walking the subtree's closest hits) and goes on with its
next sample; a second launch evaluates the subtrees of all tasks and a third
adds each deferred pixel's samples in sample order.  The image and every
count must be the ones of the in-place program bit for bit: the same
samples, the same draws, the same additions in the same order.  The task
array can fill up, in which case the remaining glass vertices are evaluated
in place, also bit for bit.  Scenes run as parsed (point lights at the
default radius 0.5: soft, 4 jittered samples each) and with every point
light's radius set to 0 (Whitted then draws nothing); the deferred_glass
count shows that the deferred program ran.
"""
import os

import numpy as np
import pytest

import pencil_amd as pa
from conftest import golden

pytestmark = pytest.mark.gpu

CORNELL = os.path.join(os.path.dirname(pa.__file__), "scenes", "cornell_c5.pecl")


def _core(**env):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update({k: str(v) for k, v in env.items()})
    try:
        return pa.Core(0)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.fixture(scope="module")
def cores():
    c_off, c_on, c_small = _core(PCL_DEFER_GLASS=0), _core(PCL_DEFER_GLASS=1), \
        _core(PCL_DEFER_GLASS=1, PCL_DEFER_TASKS=37)
    yield c_off, c_on, c_small
    for c in (c_off, c_on, c_small):
        c.close()


def _hard(hs):
    """Point lights of the host scene set to radius 0 (in place)."""
    d = hs.desc()
    for i in range(d.n_lights):
        d.lights[i].radius = 0.0
    return hs


def _scene(name, w, h, side, pathtrace, bounces, hard):
    path = CORNELL if name == "cornell_c5" else golden(name + ".pecl")
    hs = pa.HostScene.from_pecl(path, width=w, height=h, samples=side)
    if hard:
        _hard(hs)
    st = hs.settings()
    st.pathtrace, st.pathtrace_bounces = pathtrace, bounces
    return hs, st


def _render(core, hs, st):
    sid = core.upload(hs.desc())
    try:
        return core.render(sid, hs.camera(), st)
    finally:
        core.free_scene(sid)


def _same(a, sa, b, sb):
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert (sa.rays_traced, sa.closest_queries, sa.anyhit_queries) == \
        (sb.rays_traced, sb.closest_queries, sb.anyhit_queries)


CASES = [("cornell_c5", 96, 72, 4, 1, 8), ("cornell_c5", 96, 72, 4, 2, 8), ("cornell_c5", 80, 60, 3, 1, 3),
         ("three_spheres", 64, 48, 3, 1, 8), ("glass", 64, 48, 2, 1, 6), ("glass", 64, 48, 2, 2, 5)]


@pytest.mark.parametrize("hard", [True, False], ids=["hard", "soft"])
@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-{c[1]}x{c[2]}-{c[3]}x{c[3]}spp-pt{c[4]}-b{c[5]}" for c in CASES])
def test_deferred_glass_is_bit_identical(cores, case, hard):
    c_off, c_on, c_small = cores
    hs, st = _scene(*case, hard)
    ref, sr = _render(c_off, hs, st)
    assert sr.deferred_glass == 0
    img, s = _render(c_on, hs, st)
    _same(img, s, ref, sr)
    assert s.deferred_glass > 37
    img, s = _render(c_small, hs, st)   # a 37-task array: the rest in place
    _same(img, s, ref, sr)
    assert s.deferred_glass == 37


@pytest.mark.parametrize("hard", [True, False], ids=["hard", "soft"])
def test_deferred_glass_on_the_1m_triangle_cornell_frame(cores, hard):
    """The path-traced Cornell + mesh workload's scene at 320x180, 4 spp,
    8 bounces with Russian roulette."""
    from pencil_amd import workloads as wl
    c_off, c_on, _ = cores
    hs, st = wl.cornell_pt(pa, width=320, height=180, samples=2)
    if hard:
        _hard(hs)
    ref, sr = _render(c_off, hs, st)
    img, s = _render(c_on, hs, st)
    _same(img, s, ref, sr)
    assert s.deferred_glass > 0


@pytest.mark.parametrize("soft_samples", [1, 3])
def test_mixed_lights_and_soft_sample_counts(cores, soft_samples):
    """One hard and one soft point light, 1 or 3 jittered samples per soft
    light: the generator skip counts only the soft light's draws."""
    c_off, c_on, c_small = cores
    hs = pa.HostScene.from_pecl(golden("glass.pecl"), width=48, height=36, samples=3)
    st = hs.settings()
    st.pathtrace, st.pathtrace_bounces, st.soft_shadow_samples = 1, 6, soft_samples
    d = hs.desc()
    assert d.n_lights >= 2
    d.lights[0].radius = 0.0
    d.lights[1].radius = 0.3
    ref, sr = _render(c_off, hs, st)
    for core in (c_on, c_small):
        img, s = _render(core, hs, st)
        _same(img, s, ref, sr)
        assert s.deferred_glass > 0


@pytest.mark.parametrize("hard", [False, True], ids=["soft", "hard"])
def test_deferred_glass_persistent_grid_and_many_tasks(hard):
    """The persistent grid (forced) with deferral, on enough samples that the
    glass launch strides over its grid more than once and its ordering sort
    runs over many blocks (> 100 k deferred vertices; keys: the walk's counts
    under the soft light, the depth limits under the hard one): the in-place
    persistent render's image and counts."""
    from pencil_amd import workloads as wl
    c_off = _core(PCL_DEFER_GLASS=0, PCL_RENDER_KERNEL=1)
    c_on = _core(PCL_DEFER_GLASS=1, PCL_RENDER_KERNEL=1)
    try:
        hs, st = wl.cornell_pt(pa, width=640, height=360, samples=4)
        if hard:
            _hard(hs)
        ref, sr = _render(c_off, hs, st)
        img, s = _render(c_on, hs, st)
    finally:
        c_off.close()
        c_on.close()
    _same(img, s, ref, sr)
    assert s.deferred_glass > 100_000


def _batch_case(core, items, cfg):
    """Renders items [(HostScene, Settings)] as one batched job; returns the
    images and the job's stats."""
    sids, bufs, jobs = [], [], []
    try:
        for hs, st in items:
            sid = core.upload(hs.desc())
            sids.append(sid)
            d = core.alloc(st.width * st.height * 12)
            bufs.append(d)
            jobs.append((sid, hs.camera(), st, d))
        bs = core.render_batch(jobs, cfg)
        imgs = []
        for (_, st), d in zip(items, bufs):
            out = np.empty((st.height, st.width, 3), dtype=np.float32)
            core.d2h(out, d)
            imgs.append(out)
        return imgs, bs
    finally:
        for d in bufs:
            core.free(d)
        for sid in sids:
            core.free_scene(sid)


def _batch_items(hard, whitted):
    """hard: True, False, or "mixed" (the first and last scene hard, the middle one soft)."""
    items = []
    for n, (name, w, h, side, pt, b) in enumerate((("cornell_c5", 96, 72, 3, 1, 8), ("glass", 64, 48, 2, 2, 6),
                                                   ("three_spheres", 40, 30, 4, 1, 8))):
        items.append(_scene(name, w, h, side, pt, b, (n != 1) if hard == "mixed" else hard))
    if whitted:
        hs = pa.HostScene.from_pecl(golden("glass.pecl"), width=48, height=36, samples=2)
        items.append((hs, hs.settings()))
    return items


@pytest.mark.parametrize("drain", [900, 1 << 30])
@pytest.mark.parametrize("streams", [1, 3])
@pytest.mark.parametrize("hard", [True, False, "mixed"], ids=["hard", "soft", "mixed"])
def test_batched_finish_defers_glass_bit_for_bit(cores, hard, streams, drain):
    """The batched tracer's drain: its finish launch defers the glass vertices
    of path-traced jobs (three scenes of different sizes and sample counts,
    path and MIS) to a second launch as the megakernel does.  Images and
    counts equal the plain in-place renders bit for bit, with the task array
    holding every vertex and holding 37; a drain at once (threshold 2^30)
    leaves nearly every sample to the finish launch."""
    c_off, c_on, c_small = cores
    items = _batch_items(hard, whitted=False)
    refs, rst = [], []
    for hs, st in items:
        img, s = _render(c_off, hs, st)
        refs.append(img)
        rst.append(s)
    cfg = pa.WfConfig(slots=6000, blocks_per_cu=2, stack_k=16, iters_per_graph=1, drain_below=drain,
                      streams=streams)
    deferred = []
    for core in (c_off, c_on, c_small):
        imgs, bs = _batch_case(core, items, cfg)
        for img, ref in zip(imgs, refs):
            assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
        assert bs.rays_traced == sum(s.rays_traced for s in rst)
        assert bs.closest_queries == sum(s.closest_queries for s in rst)
        assert bs.anyhit_queries == sum(s.anyhit_queries for s in rst)
        assert bs.finished_pixels > 0
        deferred.append(bs.deferred_glass)
    # in place; every vertex; at most 37 per part (each part's finish launch)
    assert deferred[0] == 0
    assert deferred[1] > 37 * streams
    assert 0 < deferred[2] <= 37 * streams


def test_batched_mixed_job_finishes_in_place(cores):
    """A job that mixes Whitted and path-traced items finishes its drain in
    place (deferral is for path-traced jobs); the images are the plain ones."""
    c_off, c_on, _ = cores
    items = _batch_items(False, whitted=True)
    refs = [_render(c_off, hs, st)[0] for hs, st in items]
    cfg = pa.WfConfig(slots=6000, blocks_per_cu=2, stack_k=16, iters_per_graph=1, drain_below=1 << 30, streams=1)
    imgs, bs = _batch_case(c_on, items, cfg)
    for img, ref in zip(imgs, refs):
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
    assert bs.finished_pixels > 0 and bs.deferred_glass == 0


def test_runtime_toggle_equals_the_environment():
    """pcl_set_glass_deferral on one context: off, on, on with a 37-task cap,
    for plain and batched renders; images bit-identical, counts as set; a
    negative cap is refused."""
    hs, st = _scene("cornell_c5", 96, 72, 3, 1, 8, False)
    core = pa.Core(0)
    try:
        sid = core.upload(hs.desc())
        d = core.alloc(st.width * st.height * 12)
        cfg = pa.WfConfig(slots=6000, blocks_per_cu=2, stack_k=16, iters_per_graph=1, drain_below=1 << 30,
                          streams=1)
        imgs, counts = [], []
        for enable, cap in ((False, 0), (True, 0), (True, 37)):
            core.set_glass_deferral(enable, cap)
            img, s = core.render(sid, hs.camera(), st)
            bs = core.render_batch([(sid, hs.camera(), st, d)], cfg)
            bimg = np.empty_like(img)
            core.d2h(bimg, d)
            imgs += [img, bimg]
            counts.append((s.deferred_glass, bs.deferred_glass))
        for im in imgs[1:]:
            assert np.array_equal(im.view(np.uint32), imgs[0].view(np.uint32))
        assert counts[0] == (0, 0)
        assert counts[1][0] > 37 and counts[1][1] > 37
        assert counts[2] == (37, 37)
        with pytest.raises(pa.CoreError):
            core.set_glass_deferral(True, -1)
        core.free(d)
        core.free_scene(sid)
    finally:
        core.close()


def test_huge_light_sample_counts_render_in_place(cores):
    """Above 2^15 generator draws per shaded node (here 2 soft lights x 20,000
    jittered samples) a subtree's draws could pass 2^31, so glass vertices are
    evaluated in place; the image is the in-place program's."""
    c_off, c_on, _ = cores
    hs = pa.HostScene.from_pecl(golden("glass.pecl"), width=16, height=12, samples=1)
    st = hs.settings()
    st.pathtrace, st.pathtrace_bounces, st.soft_shadow_samples = 1, 4, 20000
    d = hs.desc()
    assert d.n_lights >= 2
    d.lights[0].radius = d.lights[1].radius = 0.3
    ref, sr = _render(c_off, hs, st)
    img, s = _render(c_on, hs, st)
    _same(img, s, ref, sr)
    assert s.deferred_glass == 0
