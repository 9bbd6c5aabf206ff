"""Primitive-test counts of the reference's traversal (run where /root/reference exists).

oracle/_ref/ref_harness wraps every primitive of the reference's scene in a
delegating subclass; its `render` prints how often the reference's BVH
traversal called Primitive::hit per primitive kind (closest-hit and occlusion
queries).  These counts are the primitive-test terms of the algorithmic-bytes
model, and the core's instrumented launch must reproduce them.  Stored for the
HDR fixtures of the 8 golden scenes and the 24 random scenes:

  prim_tests.json   {fixture: [sphere, triangle, cube, plane, cylinder, cone]}

Usage: python tests/golden/make_count_fixtures.py
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
from make_fixtures import HARNESS, RENDERS  # noqa: E402


def tests_of(pecl, opts, tmp):
    r = subprocess.run([HARNESS, "render", pecl, os.path.join(tmp, "x.bin")] + opts, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"harness failed on {pecl}: {r.stderr}")
    line = next(x for x in r.stdout.splitlines() if x.startswith("tests"))
    return [int(v) for v in line.split()[1:]]


def main():
    if not os.path.exists(HARNESS):
        sys.exit(f"{HARNESS} missing: make -C oracle ref")
    tmp = tempfile.mkdtemp()
    out = {}
    for name, scene, opts in RENDERS:
        out[name] = tests_of(os.path.join(HERE, scene + ".pecl"), opts, tmp)
        print(name, out[name], flush=True)
    with open(os.path.join(HERE, "random", "index.json")) as f:
        for e in json.load(f):
            out[e["scene"]] = tests_of(os.path.join(HERE, "random", e["scene"] + ".pecl"), e["opts"], tmp)
            print(e["scene"], out[e["scene"]], flush=True)
    with open(os.path.join(HERE, "prim_tests.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
