"""pencil_amd - Python bindings of the MI355X (gfx950) rendering core for Pencil.

The product is the C ABI in ``include/pencil_hip.h`` (``libpencil_hip.so``:
hand-written HIP kernels plus the host-side BVH build).  This module binds it
with ctypes, together with ``libpencil_host.so`` (the host-side Pencil API:
``.pecl`` parser, synthetic scenes, tone mapping) for tests and benchmarks.

There is no fallback: if the built libraries are missing, importing a binding
raises.  Build with ``python -c "import __graft_entry__ as g; g.build()"``
(or ``make -C pencil_amd``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PENCIL_AMD_HIP_LIB selects another build of the core (other compile options).
HIP_LIB = os.environ.get("PENCIL_AMD_HIP_LIB") or os.path.join(_HERE, "libpencil_hip.so")
HOST_LIB = os.path.join(_HERE, "libpencil_host.so")
CLI = os.path.join(_HERE, "bin", "pencil")

PCL_OK = 0
PRIM_SPHERE, PRIM_TRIANGLE, PRIM_CUBE, PRIM_PLANE, PRIM_CYLINDER, PRIM_CONE = range(6)
PRIM_NAMES = ["sphere", "triangle", "cube", "plane", "cylinder", "cone"]
MAT_DIFFUSE, MAT_METAL, MAT_DIELECTRIC, MAT_PRINCIPLED = range(4)
LIGHT_POINT, LIGHT_DIRECTIONAL = range(2)


class Prim(C.Structure):
    _fields_ = [("kind", C.c_int32), ("material_id", C.c_int32), ("p", C.c_float * 12)]


class Material(C.Structure):
    _fields_ = [("kind", C.c_int32), ("albedo", C.c_float * 3), ("reflectivity", C.c_float), ("ior", C.c_float),
                ("specular", C.c_float), ("shininess", C.c_float), ("specular_color", C.c_float * 3),
                ("emission", C.c_float * 3)]


class Light(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pos_or_dir", C.c_float * 3), ("intensity", C.c_float * 3),
                ("radius", C.c_float), ("is_emissive", C.c_int32)]


class SceneDesc(C.Structure):
    _fields_ = [("prims", C.POINTER(Prim)), ("n_prims", C.c_int32), ("materials", C.POINTER(Material)),
                ("n_materials", C.c_int32), ("lights", C.POINTER(Light)), ("n_lights", C.c_int32),
                ("background", C.c_float * 3), ("ambient", C.c_float * 3), ("use_bvh", C.c_int32)]


class Camera(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("look", C.c_float * 3), ("up", C.c_float * 3), ("vfov_deg", C.c_float),
                ("aspect", C.c_float), ("aperture", C.c_float), ("focus_dist", C.c_float),
                ("bokeh_blades", C.c_int32)]


class Settings(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("samples", C.c_int32), ("max_depth", C.c_int32),
                ("shadow_bias", C.c_float), ("soft_shadow_samples", C.c_int32), ("pathtrace", C.c_int32),
                ("pathtrace_bounces", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("rays_traced", C.c_uint64), ("closest_queries", C.c_uint64), ("anyhit_queries", C.c_uint64),
                ("kernel_ms", C.c_double), ("total_ms", C.c_double), ("deferred_glass", C.c_uint64)]


class WorkCounts(C.Structure):
    _fields_ = [("node_visits", C.c_uint64), ("prim_tests", C.c_uint64 * 6)]


class BatchItem(C.Structure):
    _fields_ = [("scene", C.c_int32), ("pad", C.c_int32), ("camera", Camera), ("settings", Settings),
                ("d_out_rgb", C.c_void_p)]


class WfConfig(C.Structure):
    _fields_ = [("slots", C.c_int32), ("blocks_per_cu", C.c_int32), ("stack_k", C.c_int32),
                ("iters_per_graph", C.c_int32), ("drain_below", C.c_int32),
                ("inline_shadows", C.c_int32), ("streams", C.c_int32)]


class BatchStats(C.Structure):
    _fields_ = [("rays_traced", C.c_uint64), ("closest_queries", C.c_uint64), ("anyhit_queries", C.c_uint64),
                ("iterations", C.c_uint64), ("kernel_ms", C.c_double), ("total_ms", C.c_double),
                ("used", WfConfig), ("finished_pixels", C.c_uint64),
                ("deferred_glass", C.c_uint64)]


class SceneInfo(C.Structure):
    _fields_ = [("n_prims", C.c_int32), ("n_nodes", C.c_int32), ("depth", C.c_int32), ("root", C.c_int32),
                ("device_bytes", C.c_uint64), ("build_ms", C.c_double), ("upload_ms", C.c_double),
                ("n_wide_nodes", C.c_int32), ("stack_need", C.c_int32)]


# pcl_ray / pcl_hit as numpy records (32 B / 8 B)
RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("dir", "<f4", 3), ("tmin", "<f4"), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("prim_id", "<i4")])

_P = C.c_void_p


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_hip = None
_host = None


def hip_lib():
    """The C-ABI library; raises if it was not built."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB):
            raise RuntimeError(f"{HIP_LIB} is missing: build the HIP core first (make -C pencil_amd)")
        lib = C.CDLL(HIP_LIB, mode=C.RTLD_GLOBAL)
        _sig(lib, "pcl_abi_version", C.c_int, [])
        _sig(lib, "pcl_init", C.c_int, [C.c_int, C.POINTER(_P)])
        _sig(lib, "pcl_destroy", None, [_P])
        _sig(lib, "pcl_last_error", C.c_char_p, [_P])
        _sig(lib, "pcl_scene_upload", C.c_int, [_P, C.POINTER(SceneDesc), C.POINTER(C.c_int32)])
        _sig(lib, "pcl_scene_free", C.c_int, [_P, C.c_int32])
        _sig(lib, "pcl_scene_info_get", C.c_int, [_P, C.c_int32, C.POINTER(SceneInfo)])
        _sig(lib, "pcl_scene_leaf_order", C.c_int, [_P, C.c_int32, _P, C.c_int32])
        _sig(lib, "pcl_bvh_build_host", C.c_int, [C.POINTER(SceneDesc), C.POINTER(SceneInfo), _P, C.c_int32])
        _sig(lib, "pcl_trace_rays", C.c_int, [_P, C.c_int32, _P, C.c_int32, C.c_int32, _P, C.POINTER(Stats)])
        _sig(lib, "pcl_trace_rays_device", C.c_int, [_P, C.c_int32, _P, C.c_int32, C.c_int32, _P, C.POINTER(Stats)])
        _sig(lib, "pcl_trace_rays_counts", C.c_int,
             [_P, C.c_int32, _P, C.c_int32, C.c_int32, _P, C.POINTER(WorkCounts)])
        _sig(lib, "pcl_render", C.c_int, [_P, C.c_int32, C.POINTER(Camera), C.POINTER(Settings), _P, C.POINTER(Stats)])
        _sig(lib, "pcl_render_device", C.c_int,
             [_P, C.c_int32, C.POINTER(Camera), C.POINTER(Settings), _P, C.POINTER(Stats)])
        _sig(lib, "pcl_render_counts", C.c_int,
             [_P, C.c_int32, C.POINTER(Camera), C.POINTER(Settings), _P, C.POINTER(Stats), C.POINTER(WorkCounts)])
        _sig(lib, "pcl_render_batch", C.c_int,
             [_P, C.POINTER(BatchItem), C.c_int32, C.POINTER(WfConfig), C.POINTER(BatchStats)])
        _sig(lib, "pcl_set_glass_deferral", C.c_int, [_P, C.c_int32, C.c_int64])
        _sig(lib, "pcl_wf_autotune", C.c_int,
             [_P, C.POINTER(BatchItem), C.c_int32, C.POINTER(WfConfig), C.POINTER(C.c_double)])
        _sig(lib, "pcl_scene_set_materials", C.c_int, [_P, C.c_int32, C.POINTER(Material), C.c_int32])
        _sig(lib, "pcl_tonemap_rgb8_device", C.c_int, [_P, _P, C.c_int64, C.c_float, _P])
        _sig(lib, "pcl_render_rgb8", C.c_int,
             [_P, C.c_int32, C.POINTER(Camera), C.POINTER(Settings), C.c_float, _P, C.POINTER(Stats)])
        _sig(lib, "pcl_tonemap_thresholds", C.c_int, [_P])
        _sig(lib, "pcl_render_vjp_device", C.c_int,
             [_P, C.c_int32, C.POINTER(Camera), C.POINTER(Settings), _P, _P, C.POINTER(Stats)])
        _sig(lib, "pcl_adjoint", C.c_int,
             [_P, C.c_int32, C.POINTER(Camera), C.POINTER(Settings), _P, _P, _P, _P, C.POINTER(Stats)])
        _sig(lib, "pcl_adjoint_device", C.c_int,
             [_P, C.c_int32, C.POINTER(Camera), C.POINTER(Settings), _P, _P, _P, _P, C.POINTER(Stats)])
        _sig(lib, "pcl_adjoint_device_noalloc", C.c_int,
             [_P, C.c_int32, C.POINTER(Camera), C.POINTER(Settings), _P, _P, _P, _P, C.POINTER(Stats)])
        _sig(lib, "pcl_device_alloc", C.c_int, [_P, C.c_size_t, C.POINTER(_P)])
        _sig(lib, "pcl_device_free", C.c_int, [_P, _P])
        _sig(lib, "pcl_memcpy_h2d", C.c_int, [_P, _P, _P, C.c_size_t])
        _sig(lib, "pcl_memcpy_d2h", C.c_int, [_P, _P, _P, C.c_size_t])
        _hip = lib
    return _hip


def host_lib():
    """The host-side Pencil API (parser, synthetic scenes, tone map)."""
    global _host
    if _host is None:
        hip_lib()
        if not os.path.exists(HOST_LIB):
            raise RuntimeError(f"{HOST_LIB} is missing: build with make -C pencil_amd")
        lib = C.CDLL(HOST_LIB)
        _sig(lib, "pclh_load_pecl", _P, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_int])
        _sig(lib, "pclh_make_stress", _P, [C.c_int, C.c_int, C.c_uint32, C.c_int, C.c_int])
        _sig(lib, "pclh_write_stress", C.c_int,
             [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_int])
        _sig(lib, "pclh_write_cornell_mesh", C.c_int,
             [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_float,
              C.c_char_p, C.c_int])
        _sig(lib, "pclh_make_cornell_mesh", _P, [C.c_char_p, C.c_int, C.c_uint32, C.c_float, C.c_float, C.c_float,
                                                  C.c_float, C.c_int, C.c_int, C.c_char_p, C.c_int])
        _sig(lib, "pclh_make_mixed", _P, [C.c_int, C.c_int, C.c_uint32, C.c_int, C.c_int])
        _sig(lib, "pclh_renderer_render", C.c_int,
             [_P, C.POINTER(Settings), C.c_int, _P, C.POINTER(C.c_uint64), C.c_char_p, C.c_int])
        _sig(lib, "pclh_free", None, [_P])
        _sig(lib, "pclh_desc", C.c_int, [_P, C.POINTER(SceneDesc)])
        _sig(lib, "pclh_camera", C.c_int, [_P, C.POINTER(Camera)])
        _sig(lib, "pclh_settings", C.c_int, [_P, C.POINTER(Settings)])
        _sig(lib, "pclh_unsupported", C.c_int, [_P, C.c_char_p, C.c_int])
        _sig(lib, "pclh_n_materials", C.c_int, [_P])
        _sig(lib, "pclh_set_material_color", C.c_int, [_P, C.c_int, C.c_int, C.POINTER(C.c_float)])
        _sig(lib, "pclh_tonemap_rgb8", C.c_int, [_P, C.c_int64, C.c_float, _P])
        _host = lib
    return _host


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class HostScene:
    """A scene on the host side of the API: parsed .pecl or synthetic."""

    def __init__(self, handle):
        if not handle:
            raise RuntimeError("scene construction failed")
        self._h = handle

    @classmethod
    def from_pecl(cls, path: str, width=0, height=0, samples=0, depth=0):
        err = C.create_string_buffer(4096)
        h = host_lib().pclh_load_pecl(path.encode(), width, height, samples, depth, err, len(err))
        if not h:
            raise RuntimeError(err.value.decode())
        return cls(h)

    @classmethod
    def stress(cls, freq=229, n_spheres=4096, seed=1234, width=1920, height=1080):
        return cls(host_lib().pclh_make_stress(freq, n_spheres, seed, width, height))

    @classmethod
    def cornell_mesh(cls, pecl: str, freq=229, seed=1234, scale=0.25, center=(0.0, 0.8, 0.3), width=1920,
                     height=1080):
        """A Cornell-box .pecl with the displaced icosphere inside (the path-tracing workload)."""
        err = C.create_string_buffer(4096)
        h = host_lib().pclh_make_cornell_mesh(pecl.encode(), freq, seed, scale, *center, width, height, err, len(err))
        if not h:
            raise RuntimeError(err.value.decode())
        return cls(h)

    @classmethod
    def mixed(cls, freq, n_spheres, seed, width, height):
        """Icosphere + seeded diffuse / metal / glass spheres (one batched-workload scene)."""
        return cls(host_lib().pclh_make_mixed(freq, n_spheres, seed, width, height))

    @staticmethod
    def write_stress(directory: str, name: str, freq=229, n_spheres=4096, seed=1234, width=1920, height=1080,
                     mat_mix=0):
        """The stress scene (mat_mix=1: a batched-workload scene, as `mixed`) as .pecl + .obj files."""
        err = C.create_string_buffer(1024)
        if host_lib().pclh_write_stress(directory.encode(), name.encode(), freq, n_spheres, seed, width, height,
                                        mat_mix, err, len(err)) != 0:
            raise RuntimeError(err.value.decode())

    @staticmethod
    def write_cornell_mesh(directory: str, name: str, pecl: str, freq=229, seed=1234, scale=0.25,
                           center=(0.0, 0.8, 0.3)):
        """`cornell_mesh`'s scene as .pecl + .obj files the reference parses."""
        err = C.create_string_buffer(1024)
        if host_lib().pclh_write_cornell_mesh(pecl.encode(), directory.encode(), name.encode(), freq, seed, scale,
                                              *center, err, len(err)) != 0:
            raise RuntimeError(err.value.decode())

    def desc(self) -> SceneDesc:
        d = SceneDesc()
        host_lib().pclh_desc(self._h, C.byref(d))
        return d

    def camera(self) -> Camera:
        c = Camera()
        host_lib().pclh_camera(self._h, C.byref(c))
        return c

    def settings(self) -> Settings:
        s = Settings()
        host_lib().pclh_settings(self._h, C.byref(s))
        return s

    def render_api(self, settings: Settings = None, bokeh_blades: int = 0):
        """Renders through pencil::Renderer::render (the drop-in entry point,
        backend hip): (W x H x 3 float32 image, (rays, closest, any-hit))."""
        s = settings if settings is not None else self.settings()
        img = np.empty((s.height, s.width, 3), dtype=np.float32)
        counts = (C.c_uint64 * 3)()
        err = C.create_string_buffer(4096)
        if host_lib().pclh_renderer_render(self._h, C.byref(s), bokeh_blades, _ptr(img), counts, err, len(err)) != 0:
            raise RuntimeError("Renderer::render failed: " + err.value.decode())
        return img, tuple(int(x) for x in counts)

    def unsupported(self) -> str:
        buf = C.create_string_buffer(4096)
        host_lib().pclh_unsupported(self._h, buf, len(buf))
        return buf.value.decode()

    def n_materials(self) -> int:
        return host_lib().pclh_n_materials(self._h)

    def set_material_color(self, m: int, which: int, rgb):
        arr = (C.c_float * 3)(*[float(x) for x in rgb])
        if host_lib().pclh_set_material_color(self._h, m, which, arr) != 0:
            raise ValueError("bad material edit")

    def close(self):
        if self._h:
            host_lib().pclh_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def materials_of(desc: SceneDesc):
    """A copy of a scene description's material table."""
    arr = (Material * max(1, desc.n_materials))()
    for i in range(desc.n_materials):
        C.memmove(C.byref(arr[i]), C.byref(desc.materials[i]), C.sizeof(Material))
    return arr, desc.n_materials


def material_params(mats, n) -> np.ndarray:
    """(n, 6) float64: albedo rgb, emission rgb per material."""
    return np.array([list(mats[i].albedo) + list(mats[i].emission) for i in range(n)], dtype=np.float64)


def set_material_params(mats, n, params: np.ndarray):
    """Writes (n, 6) albedo / emission values into a material table (as float32)."""
    for i in range(n):
        for c in range(3):
            mats[i].albedo[c] = float(np.float32(params[i, c]))
            mats[i].emission[c] = float(np.float32(params[i, 3 + c]))


def tonemap_rgb8(hdr: np.ndarray, exposure: float = 0.0) -> np.ndarray:
    """Host ACES + gamma 2.2 + 8-bit encode (the CLI's write_image path)."""
    hdr = np.ascontiguousarray(hdr, dtype=np.float32)
    out = np.empty(hdr.shape, dtype=np.uint8)
    host_lib().pclh_tonemap_rgb8(_ptr(hdr), hdr.size // 3, exposure, _ptr(out))
    return out


def tonemap_thresholds() -> np.ndarray:
    """The device encode's 256-entry gamma threshold table (host only)."""
    t = np.empty(256, dtype=np.float32)
    if hip_lib().pcl_tonemap_thresholds(_ptr(t)) != PCL_OK:
        raise RuntimeError("pcl_tonemap_thresholds failed")
    return t


def bvh_leaf_order_host(desc: SceneDesc):
    """The core's host BVH build alone (no device): (leaf order, SceneInfo)."""
    info = SceneInfo()
    out = np.empty(max(desc.n_prims, 1), dtype=np.int32)
    rc = hip_lib().pcl_bvh_build_host(C.byref(desc), C.byref(info), _ptr(out), out.size)
    if rc != PCL_OK:
        raise RuntimeError(f"pcl_bvh_build_host failed ({rc})")
    return out[:desc.n_prims], info


class CoreError(RuntimeError):
    pass


class Core:
    """One pcl_ctx: a HIP device and stream."""

    def __init__(self, device: int = 0):
        self.lib = hip_lib()
        self.ctx = C.c_void_p()
        rc = self.lib.pcl_init(device, C.byref(self.ctx))
        if rc != PCL_OK:
            raise CoreError(f"pcl_init({device}) failed with {rc}")

    def _check(self, rc, what):
        if rc != PCL_OK:
            raise CoreError(f"{what} failed ({rc}): {self.lib.pcl_last_error(self.ctx).decode()}")

    def upload(self, desc: SceneDesc) -> int:
        sid = C.c_int32(-1)
        self._check(self.lib.pcl_scene_upload(self.ctx, C.byref(desc), C.byref(sid)), "pcl_scene_upload")
        self._nmat = getattr(self, "_nmat", {})
        self._nmat[sid.value] = desc.n_materials
        return sid.value

    def tonemap_device(self, d_hdr: int, n_px: int, d_rgb8: int, exposure: float = 0.0):
        self._check(self.lib.pcl_tonemap_rgb8_device(self.ctx, C.c_void_p(d_hdr), n_px, exposure, C.c_void_p(d_rgb8)),
                    "pcl_tonemap_rgb8_device")

    def render_rgb8(self, sid: int, cam: Camera, settings: Settings, exposure: float = 0.0):
        out = np.empty((settings.height, settings.width, 3), dtype=np.uint8)
        st = Stats()
        self._check(self.lib.pcl_render_rgb8(self.ctx, sid, C.byref(cam), C.byref(settings), exposure, _ptr(out),
                                             C.byref(st)), "pcl_render_rgb8")
        return out, st

    def set_materials(self, sid: int, mats, n: int):
        self._check(self.lib.pcl_scene_set_materials(self.ctx, sid, mats, n), "pcl_scene_set_materials")

    def adjoint(self, sid: int, cam: Camera, settings: Settings, target: np.ndarray):
        """L2 loss against target and its gradient: (grad (n_mat, 6), loss, image, Stats)."""
        target = np.ascontiguousarray(target, dtype=np.float32)
        n = self._nmat[sid]
        grad = np.zeros((max(n, 1), 6), dtype=np.float64)
        loss = C.c_double()
        img = np.empty((settings.height, settings.width, 3), dtype=np.float32)
        st = Stats()
        self._check(self.lib.pcl_adjoint(self.ctx, sid, C.byref(cam), C.byref(settings), _ptr(target), _ptr(grad),
                                         C.byref(loss), _ptr(img), C.byref(st)), "pcl_adjoint")
        return grad[:n], loss.value, img, st

    def adjoint_device(self, sid: int, cam: Camera, settings: Settings, d_target: int, d_out: int = 0):
        n = self._nmat[sid]
        grad = np.zeros((max(n, 1), 6), dtype=np.float64)
        loss = C.c_double()
        st = Stats()
        self._check(self.lib.pcl_adjoint_device(self.ctx, sid, C.byref(cam), C.byref(settings), C.c_void_p(d_target),
                                                _ptr(grad), C.byref(loss), C.c_void_p(d_out or None), C.byref(st)),
                    "pcl_adjoint_device")
        return grad[:n], loss.value, st

    def adjoint_device_noalloc(self, sid: int, cam: Camera, settings: Settings, d_target: int, d_out: int = 0):
        """adjoint_device through pcl_adjoint_device_noalloc (work buffers in the
        library's fixed per-device scratch; nothing allocated)."""
        n = self._nmat[sid]
        grad = np.zeros((max(n, 1), 6), dtype=np.float64)
        loss = C.c_double()
        st = Stats()
        self._check(self.lib.pcl_adjoint_device_noalloc(self.ctx, sid, C.byref(cam), C.byref(settings),
                                                        C.c_void_p(d_target), _ptr(grad), C.byref(loss),
                                                        C.c_void_p(d_out or None), C.byref(st)),
                    "pcl_adjoint_device_noalloc")
        return grad[:n], loss.value, st

    def render_vjp_device(self, sid: int, cam: Camera, settings: Settings, d_adj: int):
        """grad (n_mat, 6) of sum(adj * image) for a device adjoint image."""
        n = self._nmat[sid]
        grad = np.zeros((max(n, 1), 6), dtype=np.float64)
        st = Stats()
        self._check(self.lib.pcl_render_vjp_device(self.ctx, sid, C.byref(cam), C.byref(settings),
                                                   C.c_void_p(d_adj), _ptr(grad), C.byref(st)),
                    "pcl_render_vjp_device")
        return grad[:n], st

    def free_scene(self, sid: int):
        self._check(self.lib.pcl_scene_free(self.ctx, sid), "pcl_scene_free")

    def scene_info(self, sid: int) -> SceneInfo:
        info = SceneInfo()
        self._check(self.lib.pcl_scene_info_get(self.ctx, sid, C.byref(info)), "pcl_scene_info_get")
        return info

    def leaf_order(self, sid: int) -> np.ndarray:
        n = self.scene_info(sid).n_prims
        out = np.empty(max(n, 1), dtype=np.int32)
        self._check(self.lib.pcl_scene_leaf_order(self.ctx, sid, _ptr(out), out.size), "pcl_scene_leaf_order")
        return out[:n]

    def trace_rays(self, sid: int, rays: np.ndarray, any_hit: bool = False):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.empty(rays.shape[0], dtype=HIT_DTYPE)
        st = Stats()
        self._check(self.lib.pcl_trace_rays(self.ctx, sid, _ptr(rays), rays.shape[0], int(any_hit), _ptr(hits),
                                            C.byref(st)), "pcl_trace_rays")
        return hits, st

    def render(self, sid: int, cam: Camera, settings: Settings):
        img = np.empty((settings.height, settings.width, 3), dtype=np.float32)
        st = Stats()
        self._check(self.lib.pcl_render(self.ctx, sid, C.byref(cam), C.byref(settings), _ptr(img), C.byref(st)),
                    "pcl_render")
        return img, st

    # device buffers, for timing with inputs resident in HBM
    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        self._check(self.lib.pcl_device_alloc(self.ctx, nbytes, C.byref(p)), "pcl_device_alloc")
        return p.value

    def free(self, dptr: int):
        self._check(self.lib.pcl_device_free(self.ctx, C.c_void_p(dptr)), "pcl_device_free")

    def h2d(self, dptr: int, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        self._check(self.lib.pcl_memcpy_h2d(self.ctx, C.c_void_p(dptr), _ptr(arr), arr.nbytes), "pcl_memcpy_h2d")

    def d2h(self, arr: np.ndarray, dptr: int):
        self._check(self.lib.pcl_memcpy_d2h(self.ctx, _ptr(arr), C.c_void_p(dptr), arr.nbytes), "pcl_memcpy_d2h")

    def trace_rays_device(self, sid: int, d_rays: int, n: int, d_hits: int, any_hit=False) -> Stats:
        st = Stats()
        self._check(self.lib.pcl_trace_rays_device(self.ctx, sid, C.c_void_p(d_rays), n, int(any_hit),
                                                   C.c_void_p(d_hits), C.byref(st)), "pcl_trace_rays_device")
        return st

    def trace_counts(self, sid: int, d_rays: int, n: int, d_hits: int, any_hit=False) -> WorkCounts:
        wc = WorkCounts()
        self._check(self.lib.pcl_trace_rays_counts(self.ctx, sid, C.c_void_p(d_rays), n, int(any_hit),
                                                   C.c_void_p(d_hits), C.byref(wc)), "pcl_trace_rays_counts")
        return wc

    def render_device(self, sid: int, cam: Camera, settings: Settings, d_out: int) -> Stats:
        st = Stats()
        self._check(self.lib.pcl_render_device(self.ctx, sid, C.byref(cam), C.byref(settings), C.c_void_p(d_out),
                                               C.byref(st)), "pcl_render_device")
        return st

    def render_counts(self, sid: int, cam: Camera, settings: Settings, d_out: int):
        st, wc = Stats(), WorkCounts()
        self._check(self.lib.pcl_render_counts(self.ctx, sid, C.byref(cam), C.byref(settings), C.c_void_p(d_out),
                                               C.byref(st), C.byref(wc)), "pcl_render_counts")
        return st, wc

    def set_glass_deferral(self, enable: bool, max_tasks: int = 0):
        """Glass vertices of path-traced renders deferred (the default) or in place."""
        self._check(self.lib.pcl_set_glass_deferral(self.ctx, int(bool(enable)), int(max_tasks)),
                    "pcl_set_glass_deferral")

    def render_batch(self, jobs, cfg: WfConfig = None):
        """jobs: list of (scene id, Camera, Settings, device output pointer)."""
        arr = (BatchItem * max(1, len(jobs)))()
        for i, (sid, cam, st, dptr) in enumerate(jobs):
            arr[i].scene, arr[i].camera, arr[i].settings, arr[i].d_out_rgb = sid, cam, st, dptr
        bs = BatchStats()
        self._check(self.lib.pcl_render_batch(self.ctx, arr, len(jobs), C.byref(cfg) if cfg else None, C.byref(bs)),
                    "pcl_render_batch")
        return bs

    def autotune(self, jobs):
        arr = (BatchItem * max(1, len(jobs)))()
        for i, (sid, cam, st, dptr) in enumerate(jobs):
            arr[i].scene, arr[i].camera, arr[i].settings, arr[i].d_out_rgb = sid, cam, st, dptr
        best, ms = WfConfig(), C.c_double()
        self._check(self.lib.pcl_wf_autotune(self.ctx, arr, len(jobs), C.byref(best), C.byref(ms)), "pcl_wf_autotune")
        return best, ms.value

    def render_wavefront(self, sid: int, cam: Camera, settings: Settings, cfg: WfConfig = None):
        """One item through the batched tracer, image copied back to the host."""
        n = settings.width * settings.height * 3
        d = self.alloc(n * 4)
        try:
            bs = self.render_batch([(sid, cam, settings, d)], cfg)
            img = np.empty((settings.height, settings.width, 3), dtype=np.float32)
            self.d2h(img, d)
        finally:
            self.free(d)
        return img, bs

    def close(self):
        if self.ctx:
            self.lib.pcl_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
