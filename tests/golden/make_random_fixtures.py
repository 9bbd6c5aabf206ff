"""Seeded random scenes and their reference renders (run where /root/reference exists).

Writes tests/golden/random/rand_NN.pecl -- small scenes drawn from every
primitive, material and light the core renders: spheres, cubes, planes,
triangles, quads, cylinders and cones (random axes), diffuse / Phong / metal
(rfl, fuz) / glass (ior, incl. total internal reflection) / emissive
materials (emissive spheres add their automatic light), hard and soft point
lights and directional lights, depth of field with round and polygonal
apertures -- and renders each with oracle/_ref/ref_harness (our driver
compiled against the reference's headers) in Whitted mode at depths 0..7,
path tracing (with and without Russian roulette) or MIS path tracing:

  random/rand_NN.npz   HDR float32 image, reference-convention ray count and
                       the harness options

Usage: python tests/golden/make_random_fixtures.py [N]
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_fixtures import HARNESS, read_render  # noqa: E402

OUT = os.path.join(HERE, "random")
W, H = 48, 36


def v3(a):
    return "<" + ", ".join(f"{x:.4g}" for x in a) + ">"


def unit(rng):
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def material(rng):
    clr = rng.integers(5, 96, size=3)
    kind = rng.choice(["diffuse", "phong", "metal", "fuzz", "glass", "emit"], p=[.3, .2, .15, .1, .15, .1])
    s = f"clr: {v3(clr)}"
    if kind == "phong":
        s += f", spc: {int(rng.integers(5, 60))}, shn: {int(rng.integers(8, 128))}"
    elif kind == "metal":
        s += f", rfl: {int(rng.integers(20, 100))}"
    elif kind == "fuzz":
        s += f", fuz: {int(rng.integers(0, 40))}"
    elif kind == "glass":
        s += f", ior: {rng.uniform(1.1, 2.4):.3f}"
    elif kind == "emit":
        s += f", emt: {v3(rng.integers(50, 600, size=3))}"
    return "mat { " + s + " }\n"


def primitive(rng, kind):
    c = rng.uniform(-2.5, 2.5, size=3)
    if kind == "sphere":
        return f"sphere {{ pos: {v3(c)}, rad: {rng.uniform(0.2, 1.2):.3f} }}\n"
    if kind == "cube":
        return f"cube {{ pos: {v3(c)}, size: {rng.uniform(0.2, 0.9):.3f} }}\n"
    if kind == "plane":
        n = unit(rng) if rng.uniform() < 0.3 else np.array([0.0, 1.0, 0.0])
        return f"plane {{ pos: {v3([0, rng.uniform(-2, -0.5), 0])}, nrm: {v3(n)} }}\n"
    if kind == "tri":
        p = [c + rng.uniform(-1.2, 1.2, size=3) for _ in range(3)]
        return f"tri {{\n  v0: {v3(p[0])}\n  v1: {v3(p[1])}\n  v2: {v3(p[2])}\n}}\n"
    if kind == "quad":
        u, w = unit(rng) * rng.uniform(0.5, 1.5), unit(rng) * rng.uniform(0.5, 1.5)
        a, b, cc, d = c, c + u, c + u + w, c + w
        return f"quad {{\n  a: {v3(a)}\n  b: {v3(b)}\n  c: {v3(cc)}\n  d: {v3(d)}\n}}\n"
    ax = unit(rng) if rng.uniform() < 0.6 else np.array([0.0, 1.0, 0.0])
    if kind == "cylinder":
        return (f"cylinder {{\n  pos: {v3(c)}\n  ax: {v3(ax)}\n  rad: {rng.uniform(0.2, 0.8):.3f}\n"
                f"  h: {rng.uniform(0.4, 2.0):.3f}\n}}\n")
    return (f"cone {{\n  apex: {v3(c)}\n  ax: {v3(ax)}\n  rad: {rng.uniform(0.2, 0.9):.3f}\n"
            f"  h: {rng.uniform(0.4, 2.0):.3f}\n}}\n")


def scene(k):
    rng = np.random.default_rng(1000 + k)
    s = f"# random scene {k} (tests/golden/make_random_fixtures.py)\n"
    s += f"bg: {v3(rng.integers(0, 40, size=3))}\namb: {v3(rng.integers(0, 15, size=3))}\n"
    eye = unit(rng) * rng.uniform(6.0, 9.0)
    eye[1] = abs(eye[1]) + 0.5
    s += f"cmra {{\n  pos: {v3(eye)}\n  aim: {v3(rng.uniform(-0.5, 0.5, size=3))}\n  up: <0, 1, 0>\n"
    s += f"  fov: {rng.uniform(35, 65):.2f}\n"
    if rng.uniform() < 0.3:
        s += f"  aperture: {rng.uniform(0.05, 0.3):.3f}\n  focus: {rng.uniform(4, 9):.3f}\n"
        if rng.uniform() < 0.5:
            s += f"  bd: {int(rng.integers(3, 8))}\n"
    s += "}\n"
    kinds = ["sphere", "cube", "plane", "tri", "quad", "cylinder", "cone"]
    n = int(rng.integers(3, 11))
    planes = 0
    for _ in range(n):
        kd = str(rng.choice(kinds))
        if kd == "plane":
            planes += 1
            if planes > 1:
                kd = "sphere"
        s += material(rng) + primitive(rng, kd)
    for _ in range(int(rng.integers(1, 4))):
        if rng.uniform() < 0.25:
            d = unit(rng)
            d[1] = -abs(d[1]) - 0.2
            s += f"light {{\n  type: directional\n  dir: {v3(d)}\n  clr: {v3(rng.integers(30, 100, size=3))}\n}}\n"
        else:
            r = 0.0 if rng.uniform() < 0.4 else rng.uniform(0.2, 0.8)
            pos = rng.uniform(-4, 4, size=3)
            pos[1] = abs(pos[1]) + 2.0
            s += (f"light {{\n  type: point\n  pos: {v3(pos)}\n  clr: {v3(rng.integers(30, 100, size=3))}\n"
                  f"  radius: {r:.3f}\n}}\n")
    mode = rng.choice(["whitted", "path", "mis"], p=[.5, .3, .2])
    s += f"render {{\n  size: {W} {H}\n  spp: 2\n  depth: {int(rng.integers(0, 8))}\n}}\n"
    opts = []
    if mode != "whitted":
        opts += ["--pathtrace", "--bounces", str(int(rng.integers(1, 9)))]
        if mode == "mis":
            opts.append("--mis")
    return s, opts


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 24
    if not os.path.exists(HARNESS):
        sys.exit(f"{HARNESS} missing: make -C oracle ref")
    os.makedirs(OUT, exist_ok=True)
    tmp = tempfile.mkdtemp()
    index = []
    for k in range(n):
        text, opts = scene(k)
        path = os.path.join(OUT, f"rand_{k:02d}.pecl")
        with open(path, "w") as f:
            f.write(text)
        out = os.path.join(tmp, f"{k}.bin")
        r = subprocess.run([HARNESS, "render", path, out] + opts, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"harness failed on {path}: {r.stderr}")
        img, rays = read_render(out)
        np.savez_compressed(os.path.join(OUT, f"rand_{k:02d}.npz"), hdr=img, rays=np.uint64(rays),
                            opts=" ".join(opts))
        index.append({"scene": f"rand_{k:02d}", "opts": opts, "rays": rays})
        print(k, opts, rays, flush=True)
    with open(os.path.join(OUT, "index.json"), "w") as f:
        json.dump(index, f, indent=1)


if __name__ == "__main__":
    main()
