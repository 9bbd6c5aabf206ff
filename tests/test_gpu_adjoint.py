"""The adjoint (path-replay backprop) pass against finite differences.

Scene: tests/golden/adj_init.pecl -- pencil_amd/scenes/cornell_c5.pecl at a
seeded initial point, 8 materials x (albedo rgb + emission rgb) = 48
parameters -- at 24x24, 2x2 spp, path tracing with 4 bounces (Russian
roulette never fires, so the image is a polynomial in the parameters for
fixed RNG streams), the same with MIS light sampling (trace_path_mis), and
Whitted with depth 4.  tests/golden/adj_cornell.npz
holds, from the reference renderer (oracle/_ref/ref_harness): the target
images (true parameters), the L2 loss at the initial point and its central
differences, h = 1e-3.

Bars:
  - the loss equals the reference's within 1e-4 relative (the forward image
    is the reference's but for a few path-traced pixels);
  - against central differences with h = 1e-3 (the reference's, and this
    core's own): ||g - fd|| <= 1e-3 ||fd||, and every entry within 1e-3
    relative plus FD_FLOOR = 0.25, the noise of such differences here (float32
    pixels of up to 15 over 1728 channels: ~1e-6 rounding per pixel, summed
    with random signs and divided by 2h);
  - emission entries much tighter: the image is linear in every emission, so
    the loss is quadratic in it and a central difference is exact for any
    step; with h = 0.5 the rounding noise shrinks 500-fold: 1e-5 relative;
  - the gradient is bit-reproducible; the VJP entry point and a
    torch.autograd.Function built on it give the same gradient.
"""
import json

import numpy as np
import pytest

import pencil_amd as pa
from conftest import golden

pytestmark = pytest.mark.gpu

FIX = np.load(golden("adj_cornell.npz"))
W, H = int(FIX["w"]), int(FIX["h"])
MODES = ["pt", "wh", "mis"]
REL = 1e-3
FD_FLOOR = 0.25


def scene(mode):
    hs = pa.HostScene.from_pecl(golden("adj_init.pecl"), W, H, 2, 4)
    st = hs.settings()
    if mode in ("pt", "mis"):
        st.pathtrace, st.pathtrace_bounces = (1 if mode == "pt" else 2), 4
    else:
        st.pathtrace, st.max_depth = 0, 4
    return hs, st


def check_fd(g, fd):
    assert g.shape == fd.shape
    assert np.linalg.norm(g - fd) <= REL * np.linalg.norm(fd)
    bad = np.abs(g - fd) > REL * np.abs(fd) + FD_FLOOR
    assert not bad.any(), f"max |g-fd| {np.abs(g - fd).max():.3g} at {np.argwhere(bad)[:5].tolist()}"


@pytest.fixture(scope="module", params=MODES)
def case(core, request):
    mode = request.param
    hs, st = scene(mode)
    desc = hs.desc()
    sid = core.upload(desc)
    mats, n = pa.materials_of(desc)
    grad, loss, img, stats = core.adjoint(sid, hs.camera(), st, FIX["target_" + mode])
    yield dict(mode=mode, hs=hs, st=st, sid=sid, mats=mats, n=n, grad=grad, loss=loss, img=img, stats=stats)
    core.free_scene(sid)


def test_loss_matches_reference(case):
    # a few path-traced pixels take another bounce where glibc's cosf/sinf
    # round differently from the device's correctly rounded ones
    ref = float(FIX["loss_" + case["mode"]])
    assert abs(case["loss"] - ref) <= 1e-4 * ref


def test_grad_matches_reference_fd(case):
    check_fd(case["grad"], FIX["fd_" + case["mode"]])


def test_grad_matches_own_fd(core, case):
    """Central differences of this core's forward renders: h = 1e-3 for
    albedo, h = 0.5 for emission (exact for the quadratic loss)."""
    sid, mats, n, st = case["sid"], case["mats"], case["n"], case["st"]
    cam = case["hs"].camera()
    tgt = FIX["target_" + case["mode"]].astype(np.float64)
    base = pa.material_params(mats, n)
    h = 1e-3

    def loss_at(params):
        pa.set_material_params(mats, n, params)
        core.set_materials(sid, mats, n)
        img, _ = core.render(sid, cam, st)
        return float(np.sum((img.astype(np.float64) - tgt) ** 2))

    fd = np.zeros_like(base)
    try:
        for m in range(n):
            for k in range(6):
                hk = h if k < 3 else 0.5
                p = base.copy()
                p[m, k] = np.float32(base[m, k]) + np.float32(hk)
                lp = loss_at(p)
                p[m, k] = np.float32(base[m, k]) - np.float32(hk)
                lm = loss_at(p)
                fd[m, k] = (lp - lm) / (2 * hk)
    finally:
        pa.set_material_params(mats, n, base)
        core.set_materials(sid, mats, n)
    g = case["grad"]
    check_fd(g[:, :3], fd[:, :3])
    e, fe = g[:, 3:], fd[:, 3:]
    assert np.all(np.abs(e - fe) <= 1e-5 * np.abs(fe) + 1e-3), np.abs(e - fe).max()


def test_grad_is_reproducible(core, case):
    g2, l2, _, _ = core.adjoint(case["sid"], case["hs"].camera(), case["st"], FIX["target_" + case["mode"]])
    assert np.array_equal(g2, case["grad"]) and l2 == case["loss"]


# The adjoint pass renders the image itself: the forward render's bits, from
# the forward render's queries.
def test_image_is_the_forward_render(core, case):
    img, st = core.render(case["sid"], case["hs"].camera(), case["st"])
    assert np.array_equal(img, case["img"])
    a = case["stats"]
    assert (a.rays_traced, a.closest_queries, a.anyhit_queries) == (st.rays_traced, st.closest_queries,
                                                                     st.anyhit_queries)


def test_device_entry_without_image_output(core, case):
    tgt = FIX["target_" + case["mode"]].astype(np.float32)
    d = core.alloc(tgt.nbytes)
    try:
        core.h2d(d, tgt)
        g, l, _ = core.adjoint_device(case["sid"], case["hs"].camera(), case["st"], d)
    finally:
        core.free(d)
    assert np.array_equal(g, case["grad"]) and l == case["loss"]


def test_vjp_equals_adjoint(core, case):
    adj = (2.0 * (case["img"] - FIX["target_" + case["mode"]])).astype(np.float32)
    d = core.alloc(adj.nbytes)
    try:
        core.h2d(d, adj)
        g, _ = core.render_vjp_device(case["sid"], case["hs"].camera(), case["st"], d)
    finally:
        core.free(d)
    assert np.array_equal(g, case["grad"])


def test_vjp_is_linear(core, case):
    rng = np.random.default_rng(3)
    a1 = rng.standard_normal((H, W, 3)).astype(np.float32)
    a2 = rng.standard_normal((H, W, 3)).astype(np.float32)
    out = []
    d = core.alloc(a1.nbytes)
    try:
        for adj in (a1, a2, (a1 + 2 * a2).astype(np.float32)):
            core.h2d(d, adj)
            out.append(core.render_vjp_device(case["sid"], case["hs"].camera(), case["st"], d)[0])
    finally:
        core.free(d)
    lin = out[0] + 2 * out[1]
    assert np.allclose(out[2], lin, rtol=1e-5, atol=1e-6 * np.abs(lin).max())


def test_zero_residual_gives_zero_gradient(core, case):
    g, loss, _, _ = core.adjoint(case["sid"], case["hs"].camera(), case["st"], case["img"])
    assert loss == 0.0 and not g.any()


def test_torch_autograd_function(core, case):
    """A torch.autograd.Function over the render and its VJP (test harness only)."""
    torch = pytest.importorskip("torch")
    sid, mats, n, st = case["sid"], case["mats"], case["n"], case["st"]
    cam = case["hs"].camera()
    base = pa.material_params(mats, n)

    class Render(torch.autograd.Function):
        @staticmethod
        def forward(ctx, params):
            pa.set_material_params(mats, n, params.detach().cpu().numpy())
            core.set_materials(sid, mats, n)
            img, _ = core.render(sid, cam, st)
            return torch.from_numpy(img)

        @staticmethod
        def backward(ctx, grad_img):
            adj = np.ascontiguousarray(grad_img.detach().cpu().numpy(), dtype=np.float32)
            d = core.alloc(adj.nbytes)
            try:
                core.h2d(d, adj)
                g, _ = core.render_vjp_device(sid, cam, st, d)
            finally:
                core.free(d)
            return torch.from_numpy(g)

    params = torch.tensor(base, dtype=torch.float64, requires_grad=True)
    try:
        img = Render.apply(params)
        loss = ((img - torch.from_numpy(FIX["target_" + case["mode"]])) ** 2).sum()
        loss.backward()
    finally:
        pa.set_material_params(mats, n, base)
        core.set_materials(sid, mats, n)
    assert np.array_equal(params.grad.numpy(), case["grad"])


# A lens camera (aperture, polygonal bokeh) spends RNG draws on every primary
# ray: the pass must replay them to render the forward image and follow the
# same paths.  Emission entries against exact central differences.
@pytest.mark.parametrize("mode", MODES)
def test_lens_camera(core, mode):
    hs, st = scene(mode)
    cam = hs.camera()
    cam.aperture, cam.focus_dist, cam.bokeh_blades = 0.15, 5.0, 6
    sid = core.upload(hs.desc())
    mats, n = pa.materials_of(hs.desc())
    tgt = FIX["target_" + mode]
    try:
        g, loss, img, _ = core.adjoint(sid, cam, st, tgt)
        ref, _ = core.render(sid, cam, st)
        assert np.array_equal(img, ref)
        plain, _ = core.render(sid, hs.camera(), st)
        assert not np.array_equal(img, plain)   # the lens changed the image
        d = (ref - tgt).astype(np.float64)   # the residual in float32, squared and summed in FP64
        assert abs(loss - float(np.sum(d * d))) <= 1e-9 * loss
        base = pa.material_params(mats, n)
        t64 = tgt.astype(np.float64)
        try:
            for m, k in [(0, 3), (2, 4), (4, 5), (6, 3)]:
                vals = []
                for sgn in (1, -1):
                    p = base.copy()
                    p[m, k] = np.float32(base[m, k]) + sgn * np.float32(0.5)
                    pa.set_material_params(mats, n, p)
                    core.set_materials(sid, mats, n)
                    im, _ = core.render(sid, cam, st)
                    vals.append(float(np.sum((im.astype(np.float64) - t64) ** 2)))
                fd = (vals[0] - vals[1]) / 1.0
                assert abs(g[m, k] - fd) <= 1e-5 * abs(fd) + 1e-3, (m, k, g[m, k], fd)
        finally:
            pa.set_material_params(mats, n, base)
            core.set_materials(sid, mats, n)
    finally:
        core.free_scene(sid)


def test_limits(core):
    hs = pa.HostScene.stress(freq=4, n_spheres=80, seed=1, width=16, height=16)   # 81 materials
    desc = hs.desc()
    sid = core.upload(desc)
    try:
        with pytest.raises(pa.CoreError, match="materials"):
            core.adjoint(sid, hs.camera(), hs.settings(), np.zeros((16, 16, 3), np.float32))
        mats, n = pa.materials_of(desc)
        with pytest.raises(pa.CoreError, match="expected"):
            core.set_materials(sid, mats, n - 1)
    finally:
        core.free_scene(sid)
