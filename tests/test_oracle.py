"""Pins the CPU oracle (oracle/oracle.cpp) before it is trusted as the checker.

Checks, each independent of the oracle's own formulas:
  * the reference's golden PNGs (tests/*.png of the reference, rendered by the
    reference with each scene's own settings): 8-bit ACES output bit-exact,
    cornell_box within 1 LSB on a handful of pixels (that golden predates the
    current reference code by one rounding change -- the reference itself shows
    the same 7-pixel difference);
  * fixtures from oracle/ref_harness (the reference's own parser, BVH and
    Renderer): HDR images bit-exact, reference-convention ray counts, primary
    hit t / primitive id / point / normal / material bit-exact, BVH leaf order
    and split points equal, including the 1,052,916-primitive stress scene;
  * closed forms: xorshift32's first outputs by hand, a ray against a unit
    sphere / plane, brute force (linear scan) vs BVH on random rays.
"""
import json

import numpy as np
import pytest
from PIL import Image

import pencil_amd as pa
import oracle_lib as ol
from conftest import INGEST_SCENES, RANDOM_SCENES, SCENES, golden, load_ingest, load_random


def load(name, **kw):
    return pa.HostScene.from_pecl(golden(name + ".pecl"), **kw)


@pytest.mark.parametrize("name", SCENES)
def test_oracle_matches_reference_golden_png(name):
    hs = load(name)
    osc = ol.OracleScene(hs.desc())
    img, _ = osc.render(hs.camera(), hs.settings())
    rgb = ol.tonemap_rgb8(img)
    ref = np.asarray(Image.open(golden(name + ".png")).convert("RGB"))
    assert rgb.shape == ref.shape
    diff = np.abs(rgb.astype(np.int16) - ref.astype(np.int16))
    assert diff.max() <= 1
    if name == "cornell_box":
        assert int((diff > 0).any(-1).sum()) <= 16
    else:
        assert diff.max() == 0


def _opts_to_kw(opts):
    o = opts.split()
    kw, st = {}, {"bounces": 4}
    vals = {o[i]: o[i + 1] for i in range(len(o) - 1) if o[i].startswith("--") and not o[i + 1].startswith("--")}
    for k, name in (("--width", "width"), ("--height", "height"), ("--samples", "samples")):
        if k in vals:
            kw[name] = int(vals[k])
    if "--bounces" in vals:
        st["bounces"] = int(vals["--bounces"])
    st["pathtrace"] = "--pathtrace" in o
    st["mis"] = "--mis" in o
    return kw, st


def render_fixture_inputs(fx_name):
    fx = np.load(golden(f"hdr_{fx_name}.npz"))
    kw, st = _opts_to_kw(str(fx["opts"]))
    hs = load(str(fx["scene"]), **kw)
    s = hs.settings()
    if st["pathtrace"]:
        s.pathtrace = 2 if st["mis"] else 1
        s.pathtrace_bounces = st["bounces"]
    return fx, hs, s


HDR_FIXTURES = SCENES + ["pt_cornell", "pt_cornell_rr", "pt_three_spheres", "pt_shapes", "mis_cornell",
                         "mis_soft_shadows"]


@pytest.mark.parametrize("fx_name", HDR_FIXTURES)
def test_oracle_hdr_matches_reference_harness(fx_name):
    fx, hs, s = render_fixture_inputs(fx_name)
    img, c = ol.OracleScene(hs.desc()).render(hs.camera(), s)
    assert np.array_equal(img, fx["hdr"])
    assert c.rays_traced == int(fx["rays"])


@pytest.mark.parametrize("name", SCENES)
def test_oracle_primary_hits_and_bvh_match_reference(name):
    hs = load(name)
    osc = ol.OracleScene(hs.desc())
    fx = np.load(golden(f"hits_{name}.npz"))
    rays = ol.camera_rays(hs.camera(), int(fx["width"]), int(fx["height"]))
    hits = osc.trace_rays(rays)
    rec = osc.hit_records(rays)
    assert np.array_equal(hits["prim_id"], fx["prim_id"])
    assert np.array_equal(hits["t"], fx["t"])
    hit = fx["prim_id"] >= 0
    assert np.array_equal(rec[hit, 0:3], fx["p"][hit])
    assert np.array_equal(rec[hit, 3:6], fx["normal"][hit])
    assert np.array_equal(rec[hit, 7].astype(np.int32), fx["mat"][hit])
    assert np.array_equal(rec[hit, 8].astype(np.int32), fx["front"][hit])
    b = np.load(golden(f"bvh_{name}.npz"))
    assert np.array_equal(osc.leaf_order(), b["leaves"])
    assert np.array_equal(osc.split_points(), b["splits"])


def test_oracle_stress_small_matches_reference():
    fx = np.load(golden("bvh_stress.npz"))
    p = json.loads(str(fx["params"]))
    hs = pa.HostScene.stress(**p)
    osc = ol.OracleScene(hs.desc())
    assert np.array_equal(osc.leaf_order(), fx["leaves"])
    assert np.array_equal(osc.split_points(), fx["splits"])
    rays = ol.camera_rays(hs.camera(), p["width"], p["height"])
    hits = osc.trace_rays(rays)
    assert np.array_equal(hits["prim_id"], fx["hit_prim_id"])
    assert np.array_equal(hits["t"], fx["hit_t"])
    img, c = osc.render(hs.camera(), hs.settings())
    assert np.array_equal(img, fx["hdr"])
    assert c.rays_traced == int(fx["rays"])


def test_oracle_stress_full_topology_matches_reference():
    import hashlib
    fx = np.load(golden("bvh_stress.npz"))
    p = json.loads(str(fx["full_params"]))
    hs = pa.HostScene.stress(**p)
    osc = ol.OracleScene(hs.desc())
    leaves, splits = osc.leaf_order(), osc.split_points()
    assert len(leaves) == int(fx["full_n"]) == 1048820 + 4096
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(leaves, dtype="<i4").tobytes())
    h.update(np.ascontiguousarray(splits, dtype="<i4").tobytes())
    assert h.hexdigest() == str(fx["full_hash"])
    rays = ol.camera_rays(hs.camera(), int(fx["full_hit_width"]), int(fx["full_hit_height"]))
    hits = osc.trace_rays(rays)
    assert np.array_equal(hits["prim_id"], fx["full_hit_prim_id"])
    assert np.array_equal(hits["t"], fx["full_hit_t"])


@pytest.mark.parametrize("name", INGEST_SCENES)
def test_ingest_scenes_match_reference(name):
    """mesh{} (OBJ loader, fan triangulation, fit_mesh) and include through the
    host parser: the reference's BVH leaf order and split points (so the same
    triangles in the same insertion order), its primary hits and its HDR image,
    through the oracle and through the core's host BVH build."""
    fx, hs, s = load_ingest(name)
    d = hs.desc()
    assert d.n_prims == len(fx["leaves"])
    osc = ol.OracleScene(d)
    assert np.array_equal(osc.leaf_order(), fx["leaves"])
    assert np.array_equal(osc.split_points(), fx["splits"])
    order, info = pa.bvh_leaf_order_host(d)
    assert np.array_equal(order, fx["leaves"])
    rays = ol.camera_rays(hs.camera(), int(fx["hit_width"]), int(fx["hit_height"]))
    hits = osc.trace_rays(rays)
    assert np.array_equal(hits["prim_id"], fx["hit_prim_id"])
    assert np.array_equal(hits["t"], fx["hit_t"])
    rec = osc.hit_records(rays)
    hit = fx["hit_prim_id"] >= 0
    assert np.array_equal(rec[hit, 0:3], fx["hit_p"][hit])
    assert np.array_equal(rec[hit, 3:6], fx["hit_normal"][hit])
    assert np.array_equal(rec[hit, 7].astype(np.int32), fx["hit_mat"][hit])
    img, c = osc.render(hs.camera(), s)
    assert np.array_equal(img, fx["hdr"])
    assert c.rays_traced == int(fx["rays"])


@pytest.mark.parametrize("name", RANDOM_SCENES)
def test_random_scenes_match_reference(name):
    """24 seeded random scenes over every primitive, material and light kind,
    Whitted depths 0..7, path tracing and MIS: the reference's HDR image and
    ray count, bit for bit."""
    fx, hs, s = load_random(name)
    assert hs.unsupported() == ""
    img, c = ol.OracleScene(hs.desc()).render(hs.camera(), s)
    assert np.array_equal(img, fx["hdr"])
    assert c.rays_traced == int(fx["rays"])


def test_random_scenes_are_not_empty():
    """The random scenes show their geometry: most pixels differ from the
    background in most scenes."""
    lit = []
    for name in RANDOM_SCENES:
        fx, hs, _ = load_random(name)
        bg = np.array(hs.desc().background, dtype=np.float32)
        lit.append(float(np.any(fx["hdr"] != bg, axis=-1).mean()))
    assert np.median(lit) > 0.3, lit


PRIM_TEST_FIXTURES = HDR_FIXTURES + RANDOM_SCENES


def count_inputs(name):
    """Scene and settings of a primitive-test count fixture (tests/golden/prim_tests.json)."""
    if name.startswith("rand_"):
        _, hs, s = load_random(name)
    else:
        _, hs, s = render_fixture_inputs(name)
    return hs, s


@pytest.mark.parametrize("name", PRIM_TEST_FIXTURES)
def test_primitive_tests_match_reference(name):
    """The bytes model's primitive-test counts: the oracle's traversal makes
    exactly the reference's Primitive::hit calls, per primitive kind."""
    ref = json.load(open(golden("prim_tests.json")))[name]
    hs, s = count_inputs(name)
    _, c = ol.OracleScene(hs.desc()).render(hs.camera(), s)
    assert [c.prim_tests[k] for k in range(6)] == ref


def test_oracle_cornell_pt_bands_match_reference():
    """The path-traced workload at full size (1920x1080, 256 spp, 8 bounces,
    Russian roulette, 1,048,820-triangle mesh): three row bands of the
    reference's own render, bit for bit, with the reference's ray counts."""
    from pencil_amd import workloads as wl
    fx = np.load(golden("cfg_cornell_pt.npz"))
    hs, st = wl.cornell_pt(pa)
    assert str(fx["opts"]).split()[1::2][:3] == [str(st.width), str(st.height), str(st.samples)]
    osc = ol.OracleScene(hs.desc())
    for k, (r0, r1) in enumerate(fx["bands"]):
        img, c = osc.render(hs.camera(), st, rows=(int(r0), int(r1)))
        assert np.array_equal(img[r0:r1], fx["hdr"][k])
        assert c.rays_traced == int(fx["rays"][k])


def test_oracle_batch64_bands_match_reference():
    """Three scenes of the 64-scene batched workload (meshes, metal / glass
    spheres, a soft light, 16 spp, Whitted depth 6) on their metal- and
    glass-rich row bands: the reference's output bit for bit."""
    from pencil_amd import workloads as wl
    fx = np.load(golden("cfg_batch64.npz"))
    specs = json.loads(str(fx["specs"]))
    assert [wl.batch64_specs()[int(k)] for k in fx["picks"]] == specs
    for k, sp in enumerate(specs):
        hs, st = wl.batch64_scene(pa, sp)
        r0, r1 = (int(x) for x in fx[f"band{k}"])
        img, c = ol.OracleScene(hs.desc()).render(hs.camera(), st, rows=(r0, r1))
        assert np.array_equal(img[r0:r1], fx[f"hdr{k}"])
        assert c.rays_traced == int(fx[f"rays{k}"])


def test_oracle_trig_mode_changes_only_rounding():
    """trig_mode(1) (cos / sin / pow evaluated in double, rounded once) moves
    a soft-shadow scene by rounding only: most pixels identical, none by an
    8-bit step; mode 0 is restored on exit."""
    hs = load("soft_shadows", width=64, height=48, samples=2)
    osc = ol.OracleScene(hs.desc())
    a, _ = osc.render(hs.camera(), hs.settings())
    with ol.trig_mode(1):
        b, _ = osc.render(hs.camera(), hs.settings())
    c, _ = osc.render(hs.camera(), hs.settings())
    assert np.array_equal(a, c)
    assert not np.array_equal(a, b)
    assert np.all(a == b, axis=-1).mean() >= 0.99
    assert np.array_equal(ol.tonemap_rgb8(a), ol.tonemap_rgb8(b))


def test_xorshift32_known_answers():
    # Marsaglia (2003) xorshift32 with shifts 13, 17, 5, worked by hand:
    # x=1: x^=x<<13 -> 8193; x^=x>>17 -> 8193; x^=x<<5 -> 8193^262176 = 270369
    assert ol.lib().orc_rng_u32(1, 0) == 270369
    x = 270369
    for _ in range(4):
        x ^= (x << 13) & 0xFFFFFFFF
        x ^= x >> 17
        x ^= (x << 5) & 0xFFFFFFFF
    assert ol.lib().orc_rng_u32(1, 4) == x


def _desc_from(prims, mats=None, lights=()):
    import ctypes as C
    P = (pa.Prim * len(prims))()
    for i, (kind, mid, vals) in enumerate(prims):
        P[i].kind, P[i].material_id = kind, mid
        for k, v in enumerate(vals):
            P[i].p[k] = v
    mats = mats or [pa.Material(kind=0)]
    M = (pa.Material * len(mats))(*mats)
    L = (pa.Light * max(1, len(lights)))(*lights) if lights else (pa.Light * 1)()
    d = pa.SceneDesc()
    d.prims, d.n_prims = C.cast(P, C.POINTER(pa.Prim)), len(prims)
    d.materials, d.n_materials = C.cast(M, C.POINTER(pa.Material)), len(mats)
    d.lights, d.n_lights = C.cast(L, C.POINTER(pa.Light)), len(lights)
    d.use_bvh = 1
    d._keep = (P, M, L)
    return d


def _rays(o, d, tmin=1e-4, tmax=1e9):
    r = np.zeros(len(o), dtype=pa.RAY_DTYPE)
    r["origin"], r["dir"], r["tmin"], r["tmax"] = o, d, tmin, tmax
    return r


def test_closed_form_sphere_and_plane():
    d = _desc_from([(pa.PRIM_SPHERE, 0, [0, 0, 5, 1]), (pa.PRIM_PLANE, 0, [0, -2, 0, 0, 1, 0])])
    osc = ol.OracleScene(d)
    hits = osc.trace_rays(_rays([[0, 0, 0], [0, 0, 0]], [[0, 0, 1], [0, -1, 0]]))
    assert hits["prim_id"].tolist() == [0, 1]
    assert hits["t"].tolist() == [4.0, 2.0]     # exact in float: |5-1| and |-2|


def test_bvh_equals_linear_scan_on_random_rays():
    rng = np.random.default_rng(3)
    prims = []
    for _ in range(300):   # spheres + triangles (no cones: a cone's answer depends on t_max)
        c = rng.uniform(-5, 5, 3)
        if rng.random() < 0.5:
            prims.append((pa.PRIM_SPHERE, 0, list(c) + [float(rng.uniform(0.05, 0.4))]))
        else:
            prims.append((pa.PRIM_TRIANGLE, 0, list((c + rng.normal(0, 0.5, (3, 3))).ravel())))
    d = _desc_from(prims)
    bvh = ol.OracleScene(d)
    d.use_bvh = 0
    lin = ol.OracleScene(d)
    o = rng.uniform(-8, 8, (4000, 3))
    dirs = rng.normal(size=(4000, 3))
    r = _rays(o, dirs)
    hb, hl = bvh.trace_rays(r), lin.trace_rays(r)
    assert np.array_equal(hb["t"], hl["t"])
    same = hb["prim_id"] == hl["prim_id"]
    # equal-t ties may resolve to a different primitive; none here
    assert same.mean() == 1.0
    ab, al = bvh.trace_rays(r, any_hit=True), lin.trace_rays(r, any_hit=True)
    assert np.array_equal(ab["prim_id"], al["prim_id"])


# ---- adjoint fixture: the oracle recomputes the reference's loss and FD ----

ADJ = np.load(golden("adj_cornell.npz"))


def _adj_scene(mode):
    hs = pa.HostScene.from_pecl(golden("adj_init.pecl"), int(ADJ["w"]), int(ADJ["h"]), 2, 4)
    s = hs.settings()
    if mode in ("pt", "mis"):
        s.pathtrace, s.pathtrace_bounces = (1 if mode == "pt" else 2), 4
    else:
        s.pathtrace, s.max_depth = 0, 4
    return hs, s


def _oracle_loss(hs, s, target):
    img, _ = ol.OracleScene(hs.desc()).render(hs.camera(), s)
    return float(np.sum((img.astype(np.float64) - target) ** 2))


@pytest.mark.parametrize("mode", ["pt", "wh", "mis"])
def test_adjoint_fixture_loss_and_fd_match_oracle(mode):
    """The fixture's loss and a few of its central differences, recomputed
    through the oracle (the same renders, parameters edited on the host)."""
    hs, s = _adj_scene(mode)
    tgt = ADJ["target_" + mode].astype(np.float64)
    assert abs(_oracle_loss(hs, s, tgt) - float(ADJ["loss_" + mode])) <= 1e-9 * float(ADJ["loss_" + mode])
    mats, n = pa.materials_of(hs.desc())
    base = pa.material_params(mats, n)
    h = np.float32(1e-3)
    for m, k in [(0, 0), (3, 4), (5, 1), (6, 2)]:
        vals = []
        for sign in (1, -1):
            p = base[m].astype(np.float32).copy()
            p[k] = np.float32(p[k]) + sign * h
            hs.set_material_color(m, 0, p[:3])
            hs.set_material_color(m, 1, p[3:])
            vals.append(_oracle_loss(hs, s, tgt))
        hs.set_material_color(m, 0, base[m, :3].astype(np.float32))
        hs.set_material_color(m, 1, base[m, 3:].astype(np.float32))
        fd = (vals[0] - vals[1]) / (2 * 1e-3)
        ref = ADJ["fd_" + mode][m, k]
        assert abs(fd - ref) <= 1e-6 * max(1.0, abs(ref)), (m, k, fd, ref)


def test_adjoint_fixture_structure():
    """Emission of the glass material cannot reach the path-traced image (its
    emission is added and taken back, A5) nor the Whitted one (the glass
    node's local term is discarded): the reference's FD reads zero there."""
    assert np.abs(ADJ["fd_wh"][7, 3:]).max() == 0.0
    assert np.abs(ADJ["fd_pt"][7, 3:]).max() <= 2e-3   # the FD noise floor of a ~5e3 loss
    assert np.abs(ADJ["fd_mis"][7, 3:]).max() <= 2e-3
    # MIS reweights the light term only (a point light here: w_light near 1),
    # so its differences stay near the plain path tracer's, but its albedo
    # entries move by more than the gradient tests' bar (0.25 + 1e-3 |fd|):
    # an adjoint that ignored the MIS weights would fail them
    assert np.allclose(ADJ["fd_mis"], ADJ["fd_pt"], rtol=0.05, atol=5.0)
    d = np.abs(ADJ["fd_mis"][:, :3] - ADJ["fd_pt"][:, :3])
    assert (d > 0.25 + 1e-3 * np.abs(ADJ["fd_mis"][:, :3])).sum() >= 3


def test_terrain_blocks_match_reference():
    """Terrain blocks (tern, src/pecl_parser.hpp:1180-1207, 1541-1576), parsed
    here and rendered by the oracle, equal the reference's render of the same
    file through its own parser, bit for bit.  This covers the grid, the fBm
    heights, a name seed and a number seed, and the end-of-block reset
    (:1847-1851): the second block's scale and octaves fall back to
    1.0 and 4, while its grid would survive.  tests/golden/make_terrain_fixture.py
    wrote the fixture with the reference build."""
    T = np.load(golden("terrain_ref.npz"))
    hs = pa.HostScene.from_pecl(golden("terrain.pecl"), int(T["w"]), int(T["h"]), 2, 3)
    d = hs.desc()
    assert d.n_prims == 2 * 28 * 28 + 2 * 9 * 9 + 1 and hs.unsupported() == ""
    img, _ = ol.OracleScene(d).render(hs.camera(), hs.settings())
    assert np.array_equal(img, T["hdr"])

