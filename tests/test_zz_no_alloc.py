"""pcl_adjoint_device_noalloc: the adjoint without any allocation.

Through the C ABI: two calls in a row, the second on a smaller image, each
against pcl_adjoint_device on the same inputs -- gradient, loss, image and
counts bit for bit (the two share their kernels and the reduction order is
fixed) -- and the size limit of the fixed scratch array, refused on the host
before anything is enqueued.
"""
import ctypes as C

import numpy as np
import pytest

import pencil_amd as pa
from conftest import golden

pytestmark = pytest.mark.gpu


def _case(core, sid, hs, W, H, samples, mode):
    s = hs.settings()
    s.width, s.height, s.samples = W, H, samples
    s.pathtrace, s.pathtrace_bounces = mode, 4
    cam = hs.camera()
    cam.aspect = W / H
    n = W * H * 3
    rng = np.random.default_rng(W * 1000 + H)
    tgt = rng.uniform(0.0, 1.0, size=n).astype(np.float32)
    d_tgt, d_a, d_b = core.alloc(n * 4), core.alloc(n * 4), core.alloc(n * 4)
    try:
        core.h2d(d_tgt, tgt)
        g0, l0, s0 = core.adjoint_device(sid, cam, s, d_tgt, d_a)
        g1, l1, s1 = core.adjoint_device_noalloc(sid, cam, s, d_tgt, d_b)
        a = np.empty(n, np.float32)
        b = np.empty(n, np.float32)
        core.d2h(a, d_a)
        core.d2h(b, d_b)
    finally:
        for d in (d_tgt, d_a, d_b):
            core.free(d)
    assert np.array_equal(g0.view(np.uint64), g1.view(np.uint64))
    assert np.float64(l0).view(np.uint64) == np.float64(l1).view(np.uint64)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert (s0.rays_traced, s0.closest_queries, s0.anyhit_queries) == \
        (s1.rays_traced, s1.closest_queries, s1.anyhit_queries)
    assert s1.rays_traced > 0 and np.abs(g1).sum() > 0


@pytest.mark.parametrize("mode", [1, 0])
def test_noalloc_adjoint_equals_adjoint_device(core, mode):
    hs = pa.HostScene.from_pecl(golden("adj_init.pecl"))
    sid = core.upload(hs.desc())
    try:
        _case(core, sid, hs, 64, 48, 2, mode)
        _case(core, sid, hs, 24, 16, 1, mode)    # smaller: stale partials past the end must not matter
    finally:
        core.free_scene(sid)


def test_noalloc_adjoint_refuses_work_beyond_its_scratch(core):
    hs = pa.HostScene.from_pecl(golden("adj_init.pecl"))
    sid = core.upload(hs.desc())
    try:
        s = hs.settings()
        s.width, s.height, s.samples, s.pathtrace = 4096, 4096, 1, 1   # 262,144 tiles
        grad = np.zeros((hs.n_materials(), 6), np.float64)
        loss = C.c_double()
        d_tgt = core.alloc(s.width * s.height * 12)   # a real target: nothing would read past it
        try:
            rc = core.lib.pcl_adjoint_device_noalloc(core.ctx, sid, C.byref(hs.camera()), C.byref(s),
                                                     C.c_void_p(d_tgt), grad.ctypes.data_as(C.c_void_p),
                                                     C.byref(loss), None, None)
        finally:
            core.free(d_tgt)
        assert rc == -5   # PCL_ERR_LIMIT, refused before any launch
        assert b"scratch" in core.lib.pcl_last_error(core.ctx)
    finally:
        core.free_scene(sid)
