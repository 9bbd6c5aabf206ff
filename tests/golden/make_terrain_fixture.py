"""Generate tests/golden/terrain_ref.npz (run where oracle/_ref/ref_harness is built).

The reference renderer (oracle/_ref/ref_harness, compiled from the reference
sources) renders tests/golden/terrain.pecl, parsed by the reference's own
parser, at 96x72, 2x2 spp, Whitted depth 3, hard shadows:
  hdr   the HDR image (float32, H x W x 3)

Usage: python tests/golden/make_terrain_fixture.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
SCENE = os.path.join(HERE, "terrain.pecl")
W, H = 96, 72
OPTS = ["--width", str(W), "--height", str(H), "--samples", "2", "--depth", "3"]


def main():
    if not os.path.exists(HARNESS):
        sys.exit(f"{HARNESS} missing: make -C oracle ref")
    out = os.path.join(tempfile.mkdtemp(), "terrain.bin")
    r = subprocess.run([HARNESS, "render", SCENE, out] + OPTS, capture_output=True, text=True)
    if r.returncode != 0:
        sys.exit(f"harness failed: {r.stderr}")
    raw = open(out, "rb").read()
    img = np.frombuffer(raw[8:8 + W * H * 12], dtype=np.float32).reshape(H, W, 3).copy()
    np.savez_compressed(os.path.join(HERE, "terrain_ref.npz"), hdr=img, w=np.int32(W), h=np.int32(H))
    print("done", float(img.mean()))


if __name__ == "__main__":
    main()
