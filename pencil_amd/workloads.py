"""The benchmark workloads as scene builders, shared by the bench and the tests.

Each builder returns host scenes on the host side of the API
(``pencil_amd.HostScene``) with the render settings of its workload, so a
test renders exactly the frame the bench times.

  stress()       1,048,820-triangle displaced icosphere + 4,096 spheres,
                 1920x1080, 1 spp, Whitted depth 1, one hard point light
                 (primary + one shadow ray per pixel).
  cornell_pt()   Cornell box (scenes/cornell_c5.pecl) with the same mesh
                 scaled inside, 1920x1080, 16x16 = 256 spp, path tracing
                 with 8 bounces (Russian roulette after bounce 3).
  batch64()      64 seeded scenes, meshes of 16..500k triangles
                 (log-uniform) plus 8..256 spheres with a diffuse / metal /
                 glass mix, 256^2 / 512^2 / 1024^2, 4x4 = 16 spp, Whitted
                 depth 6, one soft point light.
  inverse()      Cornell box at 512x512, 8x8 = 64 spp, path tracing with 4
                 bounces; 8 materials x (albedo rgb + emission rgb).
"""
from __future__ import annotations

import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
CORNELL = os.path.join(_HERE, "scenes", "cornell_c5.pecl")
STRESS = dict(freq=229, n_spheres=4096, seed=1234, width=1920, height=1080)
MESH_IN_BOX = dict(freq=229, seed=1234, scale=0.25, center=(0.0, 1.75, 0.5))
CORNELL_PT = dict(width=1920, height=1080, samples=16, bounces=8)
BATCH64 = dict(n=64, seed=2024, samples=4, depth=6)
INVERSE = dict(size=512, samples=8, bounces=4)


def stress(pa):
    hs = pa.HostScene.stress(**STRESS)
    return hs, hs.settings()


def cornell_pt(pa, width=CORNELL_PT["width"], height=CORNELL_PT["height"], samples=CORNELL_PT["samples"]):
    hs = pa.HostScene.cornell_mesh(CORNELL, width=width, height=height, **MESH_IN_BOX)
    st = hs.settings()
    st.samples, st.pathtrace, st.pathtrace_bounces = samples, 1, CORNELL_PT["bounces"]
    return hs, st


def batch64_specs(n=BATCH64["n"], seed=BATCH64["seed"]):
    """Per-scene generator arguments of the 64-scene job (seeded)."""
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(n):
        tris = float(np.exp(rng.uniform(np.log(16), np.log(500_000))))
        freq = max(1, int(round(np.sqrt(tris / 20.0))))
        res = int(rng.choice([256, 512, 1024]))
        specs.append(dict(freq=freq, n_spheres=int(rng.integers(8, 257)), seed=int(seed * 1000 + k),
                          width=res, height=res))
    return specs


def batch64_scene(pa, spec):
    hs = pa.HostScene.mixed(**spec)
    st = hs.settings()
    st.samples, st.max_depth, st.pathtrace = BATCH64["samples"], BATCH64["depth"], 0
    return hs, st


def batch64(pa):
    return [batch64_scene(pa, sp) for sp in batch64_specs()]


def inverse(pa):
    size = INVERSE["size"]
    hs = pa.HostScene.from_pecl(CORNELL, size, size, INVERSE["samples"], 4)
    st = hs.settings()
    st.pathtrace, st.pathtrace_bounces = 1, INVERSE["bounces"]
    return hs, st
