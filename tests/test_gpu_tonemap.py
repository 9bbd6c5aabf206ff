"""The device tone map + 8-bit encode against the host's color_to_rgb8.

The device scales by 2^exposure and applies ACES with the reference's float
operations and encodes the gamma step through the host-built threshold table
(pinned exhaustively on the CPU), so its bytes must equal the host encode's
exactly: 100 % of channels, on the reference HDR fixtures, on edge values, on
odd pixel counts (the per-lane tail), and through pcl_render_rgb8.
"""
import numpy as np
import pytest

import pencil_amd as pa
from conftest import SCENES, golden
from test_oracle import HDR_FIXTURES, load

pytestmark = pytest.mark.gpu


def device_encode(core, hdr, exposure=0.0):
    hdr = np.ascontiguousarray(hdr, dtype=np.float32)
    n_px = hdr.size // 3
    d_in, d_out = core.alloc(max(16, hdr.nbytes)), core.alloc(max(4, n_px * 3 + 4))
    try:
        core.h2d(d_in, hdr)
        core.tonemap_device(d_in, n_px, d_out, exposure)
        out = np.empty(hdr.shape, dtype=np.uint8)
        core.d2h(out, d_out)
    finally:
        core.free(d_in)
        core.free(d_out)
    return out


@pytest.mark.parametrize("fx_name", HDR_FIXTURES)
def test_fixture_encode_equals_host(core, fx_name):
    hdr = np.load(golden(f"hdr_{fx_name}.npz"))["hdr"]
    assert np.array_equal(device_encode(core, hdr), pa.tonemap_rgb8(hdr))


@pytest.mark.parametrize("n_px", [1, 2, 3, 5, 4097, 123457])
@pytest.mark.parametrize("exposure", [0.0, 1.25, -3.0])
def test_edge_values_and_tails(core, n_px, exposure):
    rng = np.random.default_rng(n_px)
    hdr = rng.exponential(0.6, (n_px, 3)).astype(np.float32)
    special = np.array([0.0, -0.0, -2.0, 1e-38, 3e38, np.inf, -np.inf, np.nan, 1.0, 0.5], np.float32)
    hdr.flat[: min(hdr.size, special.size)] = special[: min(hdr.size, special.size)]
    assert np.array_equal(device_encode(core, hdr, exposure), pa.tonemap_rgb8(hdr, exposure))


@pytest.mark.parametrize("scene", SCENES)
def test_render_rgb8_equals_host_encode(core, scene):
    hs = load(scene)
    s = hs.settings()
    sid = core.upload(hs.desc())
    try:
        hdr, _ = core.render(sid, hs.camera(), s)
        rgb, st = core.render_rgb8(sid, hs.camera(), s)
    finally:
        core.free_scene(sid)
    assert np.array_equal(rgb, pa.tonemap_rgb8(hdr))
    assert st.rays_traced > 0


def test_misaligned_buffers_are_refused(core):
    d = core.alloc(64)
    try:
        with pytest.raises(pa.CoreError, match="aligned"):
            core.tonemap_device(d + 4, 1, d + 32)
    finally:
        core.free(d)
