import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(TESTS, "golden")
SCENES = ["three_spheres", "shapes", "cornell_box", "hard_shadows", "soft_shadows", "emissive", "dof", "glass"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the HIP core")


def golden(name):
    return os.path.join(GOLDEN, name)


INGEST_SCENES = ["mesh_test", "mesh_bench", "include_demo"]


def load_ingest(name):
    """One of the reference's mesh{} / include example scenes (tests/golden/examples),
    parsed from tests/golden as the reference resolves their relative paths from
    its working directory, at the fixture's render options.  Returns the
    fixture, the host scene and its settings."""
    import numpy as np
    import pencil_amd as pa
    fx = np.load(golden(f"ingest_{name}.npz"))
    o = str(fx["opts"]).split()
    kw = {k[2:]: int(v) for k, v in zip(o[::2], o[1::2])}
    cwd = os.getcwd()
    try:
        os.chdir(GOLDEN)
        hs = pa.HostScene.from_pecl(os.path.join("examples", name + ".pecl"), **kw)
    finally:
        os.chdir(cwd)
    return fx, hs, hs.settings()


RANDOM_SCENES = [f"rand_{k:02d}" for k in range(24)]


def load_random(name):
    """A seeded random scene of tests/golden/random (see make_random_fixtures.py)
    with the settings its reference render used: (fixture, host scene, settings)."""
    import numpy as np
    import pencil_amd as pa
    fx = np.load(os.path.join(GOLDEN, "random", name + ".npz"))
    hs = pa.HostScene.from_pecl(os.path.join(GOLDEN, "random", name + ".pecl"))
    s = hs.settings()
    o = str(fx["opts"]).split()
    if "--pathtrace" in o:
        s.pathtrace = 2 if "--mis" in o else 1
        s.pathtrace_bounces = int(o[o.index("--bounces") + 1])
    return fx, hs, s


def img_stats(img, ref):
    """Parity figures of an HDR image against a reference: fraction of pixels
    bit-identical in all channels, RMSE, and the largest 8-bit difference and
    fraction of pixels within one 8-bit step after the CLI's encode (ACES,
    gamma 2.2)."""
    import numpy as np
    import pencil_amd as pa
    exact = float(np.all(img == ref, axis=-1).mean())
    rmse = float(np.sqrt(np.mean((img.astype(np.float64) - ref) ** 2)))
    d8 = np.abs(pa.tonemap_rgb8(img).astype(int) - pa.tonemap_rgb8(ref).astype(int))
    return {"exact": exact, "rmse": rmse, "max8": int(d8.max()) if d8.size else 0,
            "within1": float((d8 <= 1).all(-1).mean()) if d8.size else 1.0}


@pytest.fixture(scope="session")
def core():
    """One pcl_ctx on device 0 for the whole GPU session."""
    import pencil_amd as pa
    c = pa.Core(0)
    yield c
    c.close()


_PARITY = []


@pytest.fixture(scope="session")
def parity_log():
    """Collects per-case parity figures (exact fraction, RMSE, max 8-bit
    difference ...).  With PCL_PARITY_OUT set, they are written there as JSON
    at the end of the session."""
    def record(case, **figures):
        row = {"case": case}
        row.update({k: (float(v) if isinstance(v, (float, int)) and not isinstance(v, bool) else v)
                    for k, v in figures.items()})
        _PARITY.append(row)
        print(json.dumps(row), flush=True)
    return record


def pytest_sessionfinish(session, exitstatus):
    out = os.environ.get("PCL_PARITY_OUT")
    if out and _PARITY:
        os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
        with open(out, "w") as f:
            json.dump(_PARITY, f, indent=1)
