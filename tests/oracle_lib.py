"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: a restatement of the reference renderer's
hot path used as the checker for the HIP core.  Only tests, smoke() and the
bench's cpu_baseline leg load it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

import pencil_amd as pa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")


class Counts(C.Structure):
    _fields_ = [("rays_traced", C.c_uint64), ("closest_queries", C.c_uint64), ("anyhit_queries", C.c_uint64),
                ("node_visits", C.c_uint64), ("prim_tests", C.c_uint64 * 6)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"{ORACLE_LIB} missing: make -C oracle")
        L = C.CDLL(ORACLE_LIB)
        P = C.c_void_p
        for name, res, args in [
            ("orc_scene_create", P, [C.POINTER(pa.SceneDesc)]),
            ("orc_scene_destroy", None, [P]),
            ("orc_leaf_order", C.c_int, [P, P, C.c_int32]),
            ("orc_split_points", C.c_int, [P, P, C.c_int32]),
            ("orc_wide_heads", C.c_int, [P]),
            ("orc_trace_rays", C.c_int, [P, P, C.c_int32, C.c_int32, P, C.POINTER(Counts), C.c_int32]),
            ("orc_hit_records", C.c_int, [P, P, C.c_int32, P, C.c_int32]),
            ("orc_camera_rays", C.c_int, [C.POINTER(pa.Camera), C.c_int32, C.c_int32, P]),
            ("orc_render", C.c_int, [P, C.POINTER(pa.Camera), C.POINTER(pa.Settings), P, C.POINTER(Counts),
                                     C.c_int32, C.c_int32, C.c_int32]),
            ("orc_tonemap_rgb8", C.c_int, [P, C.c_int64, C.c_float, P]),
            ("orc_rng_u32", C.c_uint32, [C.c_uint32, C.c_int32]),
            ("orc_check_thresholds", C.c_int64, [P, C.c_int32]),
            ("orc_set_trig_mode", C.c_int, [C.c_int]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class trig_mode:
    """Context manager: 1 = the oracle evaluates cos / sin / pow / atan2 in
    double and rounds once (the HIP core's arithmetic), 0 = the float libm
    functions the reference calls (the default)."""

    def __init__(self, mode):
        self.mode = mode

    def __enter__(self):
        self.prev = lib().orc_set_trig_mode(self.mode)
        return self

    def __exit__(self, *exc):
        lib().orc_set_trig_mode(self.prev)
        return False


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class OracleScene:
    def __init__(self, desc: pa.SceneDesc):
        self.h = lib().orc_scene_create(C.byref(desc))
        if not self.h:
            raise RuntimeError("oracle rejected the scene")
        self.n = desc.n_prims

    def leaf_order(self):
        out = np.empty(max(self.n, 1), dtype=np.int32)
        n = lib().orc_leaf_order(self.h, _p(out), out.size)
        return out[:n]

    def split_points(self):
        out = np.empty(max(self.n, 1), dtype=np.int32)
        n = lib().orc_split_points(self.h, _p(out), out.size)
        return out[:n]

    def wide_heads(self):
        """Binary nodes that head a 4-wide node under the surface-area collapse."""
        return int(lib().orc_wide_heads(self.h))

    def trace_rays(self, rays, any_hit=False, threads=os.cpu_count() or 1, counts=False):
        rays = np.ascontiguousarray(rays, dtype=pa.RAY_DTYPE)
        hits = np.empty(rays.shape[0], dtype=pa.HIT_DTYPE)
        c = Counts()
        lib().orc_trace_rays(self.h, _p(rays), rays.shape[0], int(any_hit), _p(hits),
                             C.byref(c) if counts else None, threads)
        return (hits, c) if counts else hits

    def hit_records(self, rays, threads=os.cpu_count() or 1):
        rays = np.ascontiguousarray(rays, dtype=pa.RAY_DTYPE)
        out = np.empty((rays.shape[0], 9), dtype=np.float32)
        lib().orc_hit_records(self.h, _p(rays), rays.shape[0], _p(out), threads)
        return out

    def render(self, cam, settings, threads=os.cpu_count() or 1, rows=None):
        img = np.zeros((settings.height, settings.width, 3), dtype=np.float32)
        c = Counts()
        r0, r1 = rows if rows else (0, 0)
        lib().orc_render(self.h, C.byref(cam), C.byref(settings), _p(img), C.byref(c), threads, r0, r1)
        return img, c

    def close(self):
        if self.h:
            lib().orc_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def camera_rays(cam, W, H):
    out = np.empty(W * H, dtype=pa.RAY_DTYPE)
    lib().orc_camera_rays(C.byref(cam), W, H, _p(out))
    return out


def check_thresholds(t, threads=os.cpu_count() or 1):
    """Mismatches of a gamma threshold table against the encode on every float in [0, 1]."""
    t = np.ascontiguousarray(t, dtype=np.float32)
    return int(lib().orc_check_thresholds(_p(t), threads))


def tonemap_rgb8(hdr, exposure=0.0):
    hdr = np.ascontiguousarray(hdr, dtype=np.float32)
    out = np.empty(hdr.shape, dtype=np.uint8)
    lib().orc_tonemap_rgb8(_p(hdr), hdr.size // 3, exposure, _p(out))
    return out
