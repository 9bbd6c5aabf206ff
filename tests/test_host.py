"""Host-side logic of the product, on the CPU: the .pecl parser, the core's
reference-topology BVH build (pcl_bvh_build_host, no device needed), the
host tone map, the synthetic stress scene, and the C ABI's exported symbols."""
import ctypes as C
import json
import os
import re
import subprocess

import numpy as np
import pytest

import pencil_amd as pa
import oracle_lib as ol
from conftest import ROOT, SCENES, golden


@pytest.mark.parametrize("name", SCENES)
def test_core_bvh_topology_matches_reference(name):
    hs = pa.HostScene.from_pecl(golden(name + ".pecl"))
    order, info = pa.bvh_leaf_order_host(hs.desc())
    b = np.load(golden(f"bvh_{name}.npz"))
    assert np.array_equal(order, b["leaves"])
    assert info.n_prims == len(b["leaves"])
    assert info.n_nodes == max(len(b["leaves"]) - 1, 0)
    # the 4-wide collapse: the oracle's independent marking of cut heads
    assert info.n_wide_nodes == ol.OracleScene(hs.desc()).wide_heads()
    assert 0 <= info.stack_need <= 112


def test_core_bvh_topology_stress_scenes():
    import hashlib
    fx = np.load(golden("bvh_stress.npz"))
    hs = pa.HostScene.stress(**json.loads(str(fx["params"])))
    order, _ = pa.bvh_leaf_order_host(hs.desc())
    assert np.array_equal(order, fx["leaves"])
    full = pa.HostScene.stress(**json.loads(str(fx["full_params"])))
    order, info = pa.bvh_leaf_order_host(full.desc())
    # the leaf order plus the oracle's split points (both pinned to the reference)
    osc = ol.OracleScene(full.desc())
    assert np.array_equal(order, osc.leaf_order())
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(order, dtype="<i4").tobytes())
    h.update(np.ascontiguousarray(osc.split_points(), dtype="<i4").tobytes())
    assert h.hexdigest() == str(fx["full_hash"])
    assert info.n_prims == 1052916 and info.n_nodes == 1052915
    assert 0 < info.depth < 112
    assert info.n_wide_nodes == osc.wide_heads()
    # a 4-wide node has 2..4 children, so between a third and all of the binary nodes
    assert info.n_nodes / 3 <= info.n_wide_nodes <= info.n_nodes and 0 < info.stack_need <= 112


def test_parser_counts_and_settings():
    hs = pa.HostScene.from_pecl(golden("three_spheres.pecl"))
    d, s, cam = hs.desc(), hs.settings(), hs.camera()
    assert (d.n_prims, d.n_materials, d.n_lights) == (4, 4, 2)
    assert (s.width, s.height, s.samples, s.max_depth) == (640, 480, 4, 8)
    assert abs(cam.aspect - 640 / 480) < 1e-7 and cam.vfov_deg == 55.0
    m = [d.materials[i] for i in range(4)]
    assert [x.kind for x in m] == [pa.MAT_DIFFUSE, pa.MAT_DIFFUSE, pa.MAT_DIELECTRIC, pa.MAT_METAL]
    assert m[3].reflectivity == np.float32(1.0) - np.float32(0.02)   # fuz forces metal, 1 - fuzz
    assert m[0].albedo[2] == np.float32(55.0) / np.float32(100.0)
    assert d.lights[0].radius == np.float32(0.5)                     # default soft radius
    hs2 = pa.HostScene.from_pecl(golden("three_spheres.pecl"), width=320, samples=2, depth=3)
    s2 = hs2.settings()
    assert (s2.width, s2.height, s2.samples, s2.max_depth) == (320, 480, 2, 3)
    assert abs(hs2.camera().aspect - 320 / 480) < 1e-7


def test_parser_quad_emissive_and_light_types(tmp_path):
    p = tmp_path / "s.pecl"
    p.write_text("""
bg: <1, 2, 3>
mat { clr: <10, 20, 30>, emt: <150, 0, 0> }
sphere { pos: <1, 2, 3>, rad: 2 }
quad { a: <0,0,0>, b: <1,0,0>, c: <1,1,0>, d: <0,1,0> }
light { type: dir, dir: <0, -2, 0>, clr: <50, 50, 50> }
light { pos: <0, 5, 0>, radius: 0 }
cone { apex: <0, 1, 0>, rad: 0.5, h: 1 }
""")
    hs = pa.HostScene.from_pecl(str(p))
    d = hs.desc()
    assert d.n_prims == 4 and d.n_materials == 3
    assert [d.prims[i].kind for i in range(4)] == [pa.PRIM_SPHERE, pa.PRIM_TRIANGLE, pa.PRIM_TRIANGLE, pa.PRIM_CONE]
    assert d.prims[1].material_id == d.prims[2].material_id == 1
    # emissive sphere -> is_emissive point light at rad/2 (pecl_parser.hpp:1690-1693)
    assert d.n_lights == 3 and d.lights[0].is_emissive == 1 and d.lights[0].radius == 1.0
    assert d.lights[1].kind == pa.LIGHT_DIRECTIONAL and list(d.lights[1].pos_or_dir) == [0.0, -1.0, 0.0]
    assert d.lights[2].radius == 0.0
    assert d.background[0] == np.float32(0.01)
    assert d.prims[3].p[3:6] == [0.0, 1.0, 0.0]   # default cone axis


def test_parser_flags_features_outside_the_core(tmp_path):
    p = tmp_path / "f.pecl"
    p.write_text("fog { density: 0.1 }\nmat { clr: <50,50,50>, metallic: 0.5 }\nsphere { pos: <0,0,0> }\n")
    hs = pa.HostScene.from_pecl(str(p))
    assert "fog" in hs.unsupported()
    assert hs.desc().materials[0].kind == pa.MAT_PRINCIPLED
    with pytest.raises(RuntimeError):
        pa.bvh_leaf_order_host(hs.desc())


def test_parser_reports_errors(tmp_path):
    p = tmp_path / "bad.pecl"
    p.write_text("sphere { pos: <0, 0> }\n")
    with pytest.raises(RuntimeError, match="line 1"):
        pa.HostScene.from_pecl(str(p))


def test_out_of_range_number_is_an_error_not_a_crash(tmp_path):
    """std::stof throws on 1e99, as in the reference's parser; the C entry
    point turns that into an error instead of letting it cross the boundary."""
    p = tmp_path / "huge.pecl"
    p.write_text("sphere { pos: <0, 0, 1e99>, rad: 1 }\n")
    with pytest.raises(RuntimeError, match="stof"):
        pa.HostScene.from_pecl(str(p))


def test_host_tonemap_matches_oracle():
    rng = np.random.default_rng(0)
    hdr = np.concatenate([rng.uniform(0, 3, (5000, 3)), rng.uniform(0, 1e-3, (500, 3)),
                          np.array([[0, 0, 0], [1e6, 50, 1]])]).astype(np.float32)
    assert np.array_equal(pa.tonemap_rgb8(hdr), ol.tonemap_rgb8(hdr))


def test_empty_scene_builds():
    d = pa.SceneDesc()
    order, info = pa.bvh_leaf_order_host(d)
    assert len(order) == 0 and info.n_nodes == 0


def test_invalid_material_id_rejected():
    P = (pa.Prim * 1)()
    P[0].kind, P[0].material_id = pa.PRIM_SPHERE, 3
    d = pa.SceneDesc()
    d.prims, d.n_prims = C.cast(P, C.POINTER(pa.Prim)), 1
    with pytest.raises(RuntimeError):
        pa.bvh_leaf_order_host(d)


def test_c_abi_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "pencil_hip.h")).read()
    header = re.sub(r"/\*.*?\*/", "", header, flags=re.S)
    declared = set(re.findall(r"\b(pcl_[a-z0-9_]+)\s*\(", header))
    assert len(declared) >= 15
    out = subprocess.run(["nm", "-D", "--defined-only", pa.HIP_LIB], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = declared - exported
    assert not missing, missing
    lib = pa.hip_lib()
    assert lib.pcl_abi_version() == 8
    for name in declared:
        assert hasattr(lib, name)


def test_cli_refuses_cpu_backend_and_unknown_features():
    r = subprocess.run([pa.CLI, golden("three_spheres.pecl"), "/tmp/x.ppm", "--backend", "cpu"],
                       capture_output=True, text=True)
    assert r.returncode == 2 and "reference" in r.stderr
    r = subprocess.run([pa.CLI, golden("three_spheres.pecl"), "/tmp/x.ppm", "--bdpt"], capture_output=True, text=True)
    assert r.returncode == 2


# ---- device tone map: the threshold table the encode uses ------------------

def _encode_with_thresholds(hdr, exposure, t):
    """The device algorithm in numpy float32: scale, ACES, clamp, count of t[1..] <= m."""
    x = hdr.astype(np.float32) * np.float32(2.0 ** exposure)
    a, b, c, d, e = (np.float32(v) for v in (2.51, 0.03, 2.43, 0.59, 0.14))
    with np.errstate(all="ignore"):
        y = (x * (a * x + b)) / (x * (c * x + d) + e)
    y = np.where(np.float32(0.0) < y, y, np.float32(0.0))
    m = np.clip(y, np.float32(0.0), np.float32(1.0))
    return np.searchsorted(t[1:], m, side="right").astype(np.uint8)


def test_tonemap_thresholds_exhaustive():
    """Every float in [0, 1] (1,065,353,217 of them) encodes through the
    table exactly as color_to_rgb8's powf step does."""
    t = pa.tonemap_thresholds()
    assert t[0] == -np.inf and np.all(np.diff(t[1:]) > 0)
    assert ol.check_thresholds(t) == 0


def test_threshold_encode_matches_host_tonemap():
    t = pa.tonemap_thresholds()
    rng = np.random.default_rng(11)
    vals = np.concatenate([rng.exponential(0.5, 300000), rng.uniform(0, 40, 100000),
                           [0.0, -0.0, -1.0, 1e-30, 1e30, np.inf, -np.inf, np.nan]]).astype(np.float32)
    vals = vals[: len(vals) // 3 * 3].reshape(-1, 3)
    for exposure in (0.0, 1.5, -2.0):
        assert np.array_equal(_encode_with_thresholds(vals, exposure, t), pa.tonemap_rgb8(vals, exposure))
