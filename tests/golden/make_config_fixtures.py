"""Reference fixtures for the path-traced and batched workloads (run where /root/reference exists).

Drives oracle/_ref/ref_harness (our driver compiled against the reference's
own headers, oracle/Makefile target `ref`) on the scenes of
pencil_amd/workloads.py written as .pecl + .obj files the reference parses
(float-exact: see pencil_amd/host/stress_scene.hpp), and stores row bands of
the reference's HDR output through its public render_tile_region:

  cfg_cornell_pt.npz   Cornell box + 1,048,820-triangle mesh, 1920x1080,
                       256 spp, path tracing, 8 bounces, Russian roulette:
                       rows 196-199, 538-541, 880-883 and their
                       reference-convention ray counts
  cfg_batch64.npz      three scenes of the 64-scene job (the largest mesh,
                       a 1024^2 scene, a 256^2 scene), 16 spp, Whitted
                       depth 6: in each, the 4-row band whose camera rays
                       hit metal or glass most often

Usage: python tests/golden/make_config_fixtures.py [--threads T]
"""
from __future__ import annotations

import json
import os
import platform
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import pencil_amd as pa  # noqa: E402
from pencil_amd import workloads as wl  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
C3_BANDS = [(196, 200), (538, 542), (880, 884)]


def c4_picks(specs):
    """Indices of the largest mesh, the first other 1024^2 scene and the first 256^2 scene."""
    big = max(range(len(specs)), key=lambda k: specs[k]["freq"])
    r1024 = next(k for k, s in enumerate(specs) if s["width"] == 1024 and k != big)
    r256 = next(k for k, s in enumerate(specs) if s["width"] == 256 and k != big)
    return [big, r1024, r256]


def busiest_band(hs, rows=4):
    """The `rows`-row band whose pixel-center camera rays hit metal or glass most
    often (so the band exercises the Whitted recursion), from the oracle's hits."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    d, cam, st = hs.desc(), hs.camera(), hs.settings()
    W, H = st.width, st.height
    rec = ol.OracleScene(d).hit_records(ol.camera_rays(cam, W, H))
    kinds = np.array([d.materials[m].kind for m in range(d.n_materials)])
    mat = rec[:, 7].astype(np.int64)
    special = (rec[:, 6] > 0) & (kinds[mat] != pa.MAT_DIFFUSE)
    per_row = special.reshape(H, W).sum(1)
    win = np.convolve(per_row, np.ones(rows, dtype=np.int64), mode="valid")
    r0 = int(np.argmax(win))
    return (r0, r0 + rows), int(win[r0])


def render_band(pecl, opts, band, threads):
    out = pecl + f".{band[0]}.bin"
    r = subprocess.run([HARNESS, "render", pecl, out] + opts +
                       ["--rows", str(band[0]), str(band[1]), "--threads", str(threads)],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"harness failed on {pecl} {band}: {r.stderr[-2000:]}")
    raw = open(out, "rb").read()
    W, H = np.frombuffer(raw[:8], dtype=np.int32)
    img = np.frombuffer(raw[8:8 + W * H * 12], dtype=np.float32).reshape(H, W, 3).copy()
    rays = int(np.frombuffer(raw[8 + W * H * 12:], dtype=np.uint64)[0])
    return img, rays


def main():
    if not os.path.exists(HARNESS):
        sys.exit(f"{HARNESS} missing: make -C oracle ref")
    threads = int(sys.argv[sys.argv.index("--threads") + 1]) if "--threads" in sys.argv else (os.cpu_count() or 1)
    tmp = tempfile.mkdtemp()
    meta_path = os.path.join(HERE, "meta.json")
    with open(meta_path) as f:
        meta = json.load(f)

    # path-traced Cornell box + mesh
    pa.HostScene.write_cornell_mesh(tmp, "cornell_pt", wl.CORNELL, **wl.MESH_IN_BOX)
    c = wl.CORNELL_PT
    opts = ["--width", str(c["width"]), "--height", str(c["height"]), "--samples", str(c["samples"]),
            "--pathtrace", "--bounces", str(c["bounces"])]
    hdr, rays = [], []
    for band in C3_BANDS:
        img, n = render_band(os.path.join(tmp, "cornell_pt.pecl"), opts, band, threads)
        hdr.append(img)
        rays.append(n)
        print("cornell_pt rows", band, img.shape, n, flush=True)
    np.savez_compressed(os.path.join(HERE, "cfg_cornell_pt.npz"), bands=np.array(C3_BANDS, dtype=np.int32),
                        hdr=np.stack(hdr), rays=np.array(rays, dtype=np.uint64), opts=" ".join(opts))

    # three scenes of the 64-scene job
    specs = wl.batch64_specs()
    picks = c4_picks(specs)
    b = wl.BATCH64
    out = {"picks": np.array(picks, dtype=np.int32), "specs": json.dumps([specs[k] for k in picks])}
    for n_k, k in enumerate(picks):
        sp = specs[k]
        pa.HostScene.write_stress(tmp, f"b64_{k}", mat_mix=1, **sp)
        band, n_special = busiest_band(pa.HostScene.mixed(**sp))
        print("batch64 scene", k, "band", band, "metal/glass primary hits", n_special, flush=True)
        img, n = render_band(os.path.join(tmp, f"b64_{k}.pecl"),
                             ["--samples", str(b["samples"]), "--depth", str(b["depth"])], band, threads)
        out[f"hdr{n_k}"] = img
        out[f"band{n_k}"] = np.array(band, dtype=np.int32)
        out[f"rays{n_k}"] = np.uint64(n)
        print("batch64 scene", k, sp, band, n, flush=True)
    np.savez_compressed(os.path.join(HERE, "cfg_batch64.npz"), **out)
    meta["config_fixtures"] = {"harness": "oracle/ref_harness.cpp (render --rows)", "glibc": platform.libc_ver()[1],
                               "cornell_pt": {"bands": C3_BANDS, "opts": opts, "rays": rays},
                               "batch64": {"picks": picks}}
    with open(meta_path, "w") as f:
        json.dump(meta, f, indent=1)
    print("done")


if __name__ == "__main__":
    main()
