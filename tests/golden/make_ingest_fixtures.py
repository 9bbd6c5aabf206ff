"""Reference fixtures for scene ingest through mesh{} (OBJ) and include (run where /root/reference exists).

The reference's own example scenes examples/mesh_test.pecl (diablo3_pose.obj,
5,022 triangles), examples/mesh_bench.pecl (african_head.obj) and
examples/include_demo.pecl (which includes examples/common_mat.pecl) are
copied with their OBJ models under tests/golden/examples/ (their paths are
relative to the working directory, as the reference resolves mesh files), and
oracle/_ref/ref_harness -- our driver compiled against the reference's
headers -- is run from tests/golden/ on them:

  ingest_<scene>.npz   BVH leaf order + split points, pixel-center primary
                       hits (t, prim id, p, normal, material, front face) on
                       a 96x120 grid, and an HDR render at reduced size with
                       its reference-convention ray count

The scenes' golden PNGs in the reference's tests/ are stale (every pixel
differs from the current build's render), so they are not used.

Usage: python tests/golden/make_ingest_fixtures.py
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
sys.path.insert(0, HERE)
from make_fixtures import HARNESS, read_bvh, read_hits, read_render  # noqa: E402

# scene, render options (reduced size, the scene's own aspect)
SCENES = [("mesh_test", ["--width", "128", "--height", "160", "--samples", "2"]),
          ("mesh_bench", ["--width", "128", "--height", "160", "--samples", "2"]),
          ("include_demo", ["--width", "160", "--height", "120", "--samples", "2"])]


def run(args):
    r = subprocess.run([HARNESS] + args, capture_output=True, text=True, cwd=HERE)
    if r.returncode != 0:
        raise RuntimeError(f"harness {args} failed: {r.stderr}")
    return r.stdout


def main():
    if not os.path.exists(HARNESS):
        sys.exit(f"{HARNESS} missing: make -C oracle ref")
    tmp = tempfile.mkdtemp()
    meta_path = os.path.join(HERE, "meta.json")
    with open(meta_path) as f:
        meta = json.load(f)
    meta["ingest_fixtures"] = {"harness": "oracle/ref_harness.cpp (run from tests/golden)",
                               "glibc": platform.libc_ver()[1], "scenes": {}}
    for name, opts in SCENES:
        pecl = os.path.join("examples", name + ".pecl")
        out = {}
        run(["bvh", pecl, os.path.join(tmp, name + ".bvh")])
        out["leaves"], out["splits"] = read_bvh(os.path.join(tmp, name + ".bvh"))
        run(["hits", pecl, "96", "120", os.path.join(tmp, name + ".hits")])
        out.update({"hit_" + k: v for k, v in read_hits(os.path.join(tmp, name + ".hits")).items()})
        run(["render", pecl, os.path.join(tmp, name + ".bin")] + opts)
        img, rays = read_render(os.path.join(tmp, name + ".bin"))
        out["hdr"], out["rays"], out["opts"] = img, np.uint64(rays), " ".join(opts)
        np.savez_compressed(os.path.join(HERE, f"ingest_{name}.npz"), **out)
        meta["ingest_fixtures"]["scenes"][name] = {"prims": int(len(out["leaves"])), "opts": opts, "rays": rays}
        print(name, len(out["leaves"]), "prims", img.shape, rays, flush=True)
    with open(meta_path, "w") as f:
        json.dump(meta, f, indent=1)
    print("done")


if __name__ == "__main__":
    main()
