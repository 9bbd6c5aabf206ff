"""Generate the reference fixtures under tests/golden/ (run where /root/reference exists).

Drives oracle/_ref/ref_harness -- our harness compiled against the reference's
own headers (oracle/Makefile target `ref`) -- and stores small .npz fixtures:

  hdr_<scene>.npz      HDR float32 image + reference-convention ray count
                       (Whitted at reduced size, the goldens' 8 scenes; path
                       tracing on cornell_box / three_spheres, with and
                       without Russian roulette)
  hits_<scene>.npz     pixel-center primary rays: t, prim id (insertion
                       index), p, normal, material, front_face
  bvh_<scene>.npz      BVH leaf order (insertion ids) and preorder split points
  bvh_stress.npz       the same for a small seeded stress scene, plus a
                       SHA-256 of the leaf order / splits of the full
                       1,048,820-triangle stress scene
  meta.json            glibc version, sizes

Usage: python tests/golden/make_fixtures.py [--skip-full] [--only NAME,NAME]
(--only regenerates just those HDR fixtures and leaves everything else alone)
"""
from __future__ import annotations

import hashlib
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

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
SCENES = ["three_spheres", "shapes", "cornell_box", "hard_shadows", "soft_shadows", "emissive", "dof", "glass"]

# (fixture name, scene, options)
RENDERS = [(s, s, ["--width", "128", "--height", "96", "--samples", "2"]) for s in SCENES] + [
    ("pt_cornell", "cornell_box", ["--width", "96", "--height", "96", "--samples", "2", "--pathtrace", "--bounces", "4"]),
    ("pt_cornell_rr", "cornell_box", ["--width", "64", "--height", "64", "--samples", "2", "--pathtrace", "--bounces", "8"]),
    ("pt_three_spheres", "three_spheres", ["--width", "96", "--height", "72", "--samples", "2", "--pathtrace", "--bounces", "6"]),
    ("pt_shapes", "shapes", ["--width", "96", "--height", "72", "--samples", "2", "--pathtrace", "--bounces", "5"]),
    ("mis_cornell", "cornell_box", ["--width", "96", "--height", "96", "--samples", "2", "--pathtrace", "--mis",
                                    "--bounces", "4"]),
    ("mis_soft_shadows", "soft_shadows", ["--width", "96", "--height", "72", "--samples", "2", "--pathtrace", "--mis",
                                          "--bounces", "6"]),
]


def run(args):
    r = subprocess.run([HARNESS] + args, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"harness {args} failed: {r.stderr}")
    return r.stdout


def read_render(path):
    raw = open(path, "rb").read()
    W, H = np.frombuffer(raw[:8], dtype=np.int32)
    img = np.frombuffer(raw[8:8 + W * H * 12], dtype=np.float32).reshape(H, W, 3)
    rays = int(np.frombuffer(raw[8 + W * H * 12:], dtype=np.uint64)[0])
    return img, rays


def read_hits(path):
    raw = open(path, "rb").read()
    W, H = np.frombuffer(raw[:8], dtype=np.int32)
    rec = np.frombuffer(raw[8:], dtype=np.float32).reshape(H * W, 10)
    ids = rec[:, 1].copy().view(np.int32)
    return dict(t=rec[:, 0].copy(), prim_id=ids, p=rec[:, 2:5].copy(), normal=rec[:, 5:8].copy(),
                mat=rec[:, 8].astype(np.int32), front=rec[:, 9].astype(np.int32), width=W, height=H)


def read_bvh(path):
    raw = open(path, "rb").read()
    n, ns = np.frombuffer(raw[:8], dtype=np.int32)
    leaves = np.frombuffer(raw[8:8 + 4 * n], dtype=np.int32).copy()
    splits = np.frombuffer(raw[8 + 4 * n:8 + 4 * n + 4 * ns], dtype=np.int32).copy()
    return leaves, splits


def topo_hash(leaves, splits):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(leaves, dtype="<i4").tobytes())
    h.update(np.ascontiguousarray(splits, dtype="<i4").tobytes())
    return h.hexdigest()


def main():
    if not os.path.exists(HARNESS):
        sys.exit(f"{HARNESS} missing: make -C oracle ref")
    tmp = tempfile.mkdtemp()
    only = None
    if "--only" in sys.argv:
        only = set(sys.argv[sys.argv.index("--only") + 1].split(","))
        with open(os.path.join(HERE, "meta.json")) as f:
            meta = json.load(f)
    else:
        meta = {"glibc": platform.libc_ver()[1], "harness": "oracle/ref_harness.cpp", "renders": {}}
    for name, scene, opts in RENDERS:
        if only is not None and name not in only:
            continue
        out = os.path.join(tmp, name + ".bin")
        run(["render", os.path.join(HERE, scene + ".pecl"), out] + opts)
        img, rays = read_render(out)
        np.savez_compressed(os.path.join(HERE, f"hdr_{name}.npz"), hdr=img, rays=np.uint64(rays),
                            scene=scene, opts=" ".join(opts))
        meta["renders"][name] = {"scene": scene, "opts": opts, "rays": rays}
        print("render", name, img.shape, rays)
    if only is not None:
        with open(os.path.join(HERE, "meta.json"), "w") as f:
            json.dump(meta, f, indent=1)
        print("done")
        return
    for scene in SCENES:
        out = os.path.join(tmp, scene + ".hits")
        run(["hits", os.path.join(HERE, scene + ".pecl"), "160", "120", out])
        np.savez_compressed(os.path.join(HERE, f"hits_{scene}.npz"), **read_hits(out))
        out = os.path.join(tmp, scene + ".bvh")
        run(["bvh", os.path.join(HERE, scene + ".pecl"), out])
        leaves, splits = read_bvh(out)
        np.savez_compressed(os.path.join(HERE, f"bvh_{scene}.npz"), leaves=leaves, splits=splits)
        print("hits/bvh", scene, len(leaves))
    # stress scenes through the reference parser (.pecl + .obj)
    small = dict(freq=24, n_spheres=256, seed=7, width=192, height=108)
    pa.HostScene.write_stress(tmp, "stress_small", **small)
    sp = os.path.join(tmp, "stress_small.pecl")
    out = os.path.join(tmp, "stress_small.bvh")
    run(["bvh", sp, out])
    leaves, splits = read_bvh(out)
    out_h = os.path.join(tmp, "stress_small.hits")
    run(["hits", sp, "192", "108", out_h])
    hits = read_hits(out_h)
    out_r = os.path.join(tmp, "stress_small.bin")
    run(["render", sp, out_r])
    img, rays = read_render(out_r)
    fixture = dict(leaves=leaves, splits=splits, hdr=img, rays=np.uint64(rays),
                   params=json.dumps(small), **{"hit_" + k: v for k, v in hits.items()})
    if "--skip-full" not in sys.argv:
        full = dict(freq=229, n_spheres=4096, seed=1234, width=1920, height=1080)
        pa.HostScene.write_stress(tmp, "stress_full", **full)
        outf = os.path.join(tmp, "stress_full.bvh")
        run(["bvh", os.path.join(tmp, "stress_full.pecl"), outf])
        fl, fs = read_bvh(outf)
        fixture["full_params"] = json.dumps(full)
        fixture["full_hash"] = topo_hash(fl, fs)
        fixture["full_n"] = np.int64(len(fl))
        # hits on a 96x54 grid of pixel-center rays over the 1080p frame
        outfh = os.path.join(tmp, "stress_full.hits")
        run(["hits", os.path.join(tmp, "stress_full.pecl"), "96", "54", outfh])
        fh = read_hits(outfh)
        fixture.update({"full_hit_" + k: v for k, v in fh.items()})
        print("stress full", len(fl), fixture["full_hash"][:16])
    np.savez_compressed(os.path.join(HERE, "bvh_stress.npz"), **fixture)
    with open(os.path.join(HERE, "meta.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print("done")


if __name__ == "__main__":
    main()
