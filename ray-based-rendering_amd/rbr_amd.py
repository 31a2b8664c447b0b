"""Python host binding of the MI355X path (ctypes over lib/librbr.so, include/rbr.h).

This is the host-side mirror the tests and the benchmark drive.  It mirrors the reference's
`Accelerator` interface for the traversal path (build / intersect / intersectP /
worldBounds, include/render/acceleration/accelerator.h:19-34) in batched form, and the
`Integrator::render` entry for the path tracer.  All compute runs in the HIP library; there is
no CPU fallback here — if the library is missing, importing this module fails.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_PKG, "lib")

RAY_DTYPE = np.dtype([("o", "<f4", 3), ("d", "<f4", 3), ("tmin", "<f4"), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("id", "<u4"), ("t", "<f4")])
MISS = 0xFFFFFFFF

ERRORS = {1: "invalid argument", 2: "HIP error", 3: "bad state", 4: "out of memory", 5: "unsupported",
          6: "traversal watchdog", 7: "file error"}
ERR_WATCHDOG = 6
ERR_IO = 7


class RbrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"rbr error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


def _load(name: str) -> C.CDLL:
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built: run ray-based-rendering_amd/build.py (or __graft_entry__.build())")
    return C.CDLL(path)


_lib = _load("librbr.so")
_wl = _load("librbr_workloads.so")

_vp = C.c_void_p
_u32 = C.c_uint32
_sz = C.c_size_t


class BvhStats(C.Structure):
    _fields_ = [("width", C.c_uint32), ("node_count", C.c_uint32), ("leaf_count", C.c_uint32),
                ("prim_count", C.c_uint32), ("unbounded", C.c_uint32), ("max_depth", C.c_uint32),
                ("node_bytes", C.c_uint64), ("prim_bytes", C.c_uint64), ("sah_cost", C.c_double),
                ("build_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


def _sig(fn, args, res=C.c_int):
    fn.argtypes = args
    fn.restype = res
    return fn


_sig(_lib.rbr_ctx_create, [C.c_int, C.POINTER(_vp)])
_sig(_lib.rbr_ctx_destroy, [_vp], None)
_sig(_lib.rbr_last_error, [_vp], C.c_char_p)
_sig(_lib.rbr_device_info, [_vp, C.c_char_p, _sz, C.POINTER(C.c_int)])
_sig(_lib.rbr_device_alloc, [_vp, _sz, C.POINTER(_vp)])
_sig(_lib.rbr_device_free, [_vp, _vp])
_sig(_lib.rbr_copy_to_device, [_vp, _vp, _vp, _sz])
_sig(_lib.rbr_copy_to_host, [_vp, _vp, _vp, _sz])
_sig(_lib.rbr_synchronize, [_vp])
_sig(_lib.rbr_calib_copy, [_vp, _vp, _vp, _sz])
_sig(_lib.rbr_scene_create, [_vp, C.POINTER(_vp)])
_sig(_lib.rbr_scene_destroy, [_vp], None)
_sig(_lib.rbr_scene_set_geometry, [_vp, _vp, _u32, _vp, _vp, _u32, _vp, _vp, _u32, _vp, _vp, _u32, _vp])
_sig(_lib.rbr_scene_build_bvh, [_vp, C.c_int, C.c_int])
_sig(_lib.rbr_scene_build_bvh_device, [_vp, C.c_int, C.c_int])
_sig(_lib.rbr_scene_bvh_stats, [_vp, C.POINTER(BvhStats)])
_sig(_lib.rbr_scene_device_nodes, [_vp, _vp, _sz, C.POINTER(_sz)])
_sig(_lib.rbr_scene_bvh_arrays, [_vp, _vp, _sz, _vp, _sz, C.POINTER(_sz), C.POINTER(_sz)])
BVH_QUANTIZED = 0x100
BVH_GREEDY = 0x200
BVH_PLOC = 0x400   # device build: clustering tree instead of the radix tree
_sig(_lib.rbr_scene_world_bounds, [_vp, _vp, _vp])
_sig(_lib.rbr_trace_closest, [_vp, _vp, _sz, _vp])
_sig(_lib.rbr_trace_any, [_vp, _vp, _sz, _vp])
_sig(_lib.rbr_trace_closest_device, [_vp, _vp, _sz, _vp])
_sig(_lib.rbr_trace_any_device, [_vp, _vp, _sz, _vp])
_sig(_lib.rbr_set_trace_tuning, [_vp, C.c_int, C.c_int, C.c_int])
_sig(_lib.rbr_set_trace_step_cap, [_vp, C.c_uint32])
_sig(_lib.rbr_ctx_set_stream, [_vp, _vp])
_sig(_lib.rbr_trace_resident_waves, [_vp, C.c_int, C.POINTER(C.c_int)])
_sig(_lib.rbr_time_trace, [_vp, C.c_int, _vp, _sz, _vp, C.c_int, C.POINTER(C.c_float)])
_sig(_lib.rbr_host_trace_closest, [_vp, _vp, _sz, _vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)])
_sig(_lib.rbr_host_trace_any, [_vp, _vp, _sz, _vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)])

_sig(_lib.rbr_obj_parse, [C.c_char_p, _sz, _vp, _sz, C.POINTER(_sz), C.POINTER(_sz), C.POINTER(_sz)])
_sig(_lib.rbr_obj_parse_normals, [C.c_char_p, _sz, _vp, _sz, C.POINTER(_sz)])
_sig(_lib.rbr_scene_set_tri_normals, [_vp, _vp, _sz])
_sig(_lib.rbr_triangle_shading_normal, [_vp, _vp, _vp, _vp, _vp])
_sig(_lib.rbr_film_to_rgb8, [_vp, C.c_int, C.c_int, C.c_float, C.c_int, C.c_float, _vp])
_sig(_lib.rbr_film_write_png, [C.c_char_p, _vp, C.c_int, C.c_int])
_sig(_lib.rbr_film_write_pfm, [C.c_char_p, _vp, C.c_int, C.c_int])
_sig(_lib.rbr_film_read_pfm, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), _vp, _sz])

_sig(_wl.rbrw_icosphere, [C.c_int, C.c_float, C.c_float, C.c_uint64, _vp], _sz)
_sig(_wl.rbrw_noise3, [C.c_float, C.c_float, C.c_float, C.c_uint64], C.c_float)
_sig(_wl.rbrw_rays_uniform, [_vp, _vp, C.c_float, _sz, C.c_uint64, C.c_float, C.c_float, _vp], None)
_sig(_wl.rbrw_rays_shadow, [_vp, _vp, _sz, C.c_uint64, C.c_float, _vp], None)
_sig(_wl.rbrw_u01, [C.c_uint64, C.c_uint64], C.c_float)


def lib_path() -> str:
    return os.path.join(LIB_DIR, "librbr.so")


def _ptr(a) -> int | None:
    if a is None:
        return None
    return a.ctypes.data


def _f32(a, cols: int) -> np.ndarray | None:
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1, cols)
    return a


def _ids(a) -> np.ndarray | None:
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=np.uint32)


def rays_array(o, d, tmin=1e-4, tmax=np.inf) -> np.ndarray:
    """Packs origins/directions (n,3) into the 32-byte ray layout (directions used as given)."""
    o = np.asarray(o, np.float32).reshape(-1, 3)
    d = np.asarray(d, np.float32).reshape(-1, 3)
    r = np.empty(len(o), RAY_DTYPE)
    r["o"] = o
    r["d"] = d
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


class Context:
    """A HIP device + stream (rbr_ctx)."""

    def __init__(self, device: int = 0):
        h = _vp()
        rc = _lib.rbr_ctx_create(device, C.byref(h))
        if rc != 0:
            raise RbrError(rc, f"rbr_ctx_create(device={device}) failed")
        self._h = h
        self._scenes: list = []

    def check(self, rc: int) -> None:
        if rc != 0:
            raise RbrError(rc, _lib.rbr_last_error(self._h).decode())

    @property
    def handle(self):
        return self._h

    def device_info(self) -> tuple[str, int]:
        buf = C.create_string_buffer(256)
        cu = C.c_int(0)
        self.check(_lib.rbr_device_info(self._h, buf, 256, C.byref(cu)))
        return buf.value.decode(), cu.value

    def alloc(self, nbytes: int) -> "DeviceBuffer":
        return DeviceBuffer(self, nbytes)

    def synchronize(self) -> None:
        self.check(_lib.rbr_synchronize(self._h))

    def set_stream(self, stream) -> None:
        """Makes a caller's hipStream_t (an int or c_void_p handle) the context stream; None
        returns to the context's own stream (rbr_ctx_set_stream)."""
        self.check(_lib.rbr_ctx_set_stream(self._h, None if stream is None else C.c_void_p(int(stream))))

    def set_trace_step_cap(self, cap: int) -> None:
        """Traversal watchdog cap (node visits + primitive tests per ray); 0 = the BVH's own bound."""
        self.check(_lib.rbr_set_trace_step_cap(self._h, int(cap)))

    def close(self) -> None:
        if getattr(self, "_h", None):
            for s in self._scenes:
                s.close()
            _lib.rbr_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """A device allocation owned by a context (inputs kept resident in HBM)."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = _vp()
        ctx.check(_lib.rbr_device_alloc(ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = p.value

    def upload(self, arr: np.ndarray) -> None:
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        self.ctx.check(_lib.rbr_copy_to_device(self.ctx.handle, self.ptr, arr.ctypes.data, arr.nbytes))

    def download(self, out: np.ndarray) -> np.ndarray:
        assert out.flags.c_contiguous and out.nbytes <= self.nbytes
        self.ctx.check(_lib.rbr_copy_to_host(self.ctx.handle, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            self.ctx.check(_lib.rbr_device_free(self.ctx.handle, self.ptr))
            self.ptr = None


class Scene:
    """Geometry + wide BVH on the device (rbr_scene): the batched Accelerator."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = _vp()
        ctx.check(_lib.rbr_scene_create(ctx.handle, C.byref(h)))
        self._h = h
        ctx._scenes.append(self)

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.rbr_scene_destroy(self._h)
            self._h = None

    def set_geometry(self, tris=None, tri_ids=None, spheres=None, sphere_ids=None, quads=None, quad_ids=None,
                     planes=None, plane_ids=None) -> None:
        t, s, q, p = _f32(tris, 9), _f32(spheres, 4), _f32(quads, 12), _f32(planes, 4)
        ti, si, qi, pi = _ids(tri_ids), _ids(sphere_ids), _ids(quad_ids), _ids(plane_ids)
        self._keep = (t, s, q, p, ti, si, qi, pi)
        n = lambda a: 0 if a is None else len(a)
        self.ctx.check(_lib.rbr_scene_set_geometry(self._h, _ptr(t), n(t), _ptr(ti), _ptr(s), n(s), _ptr(si),
                                                   _ptr(q), n(q), _ptr(qi), _ptr(p), n(p), _ptr(pi)))

    def set_tri_normals(self, normals) -> None:
        """Vertex normals (n, 9) for the triangles of set_geometry (None removes them)."""
        nm = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 9)
        self._keep_normals = nm
        self.ctx.check(_lib.rbr_scene_set_tri_normals(self._h, _ptr(nm), 0 if nm is None else len(nm)))

    def build_bvh(self, width: int = 4, leaf_max: int = 4, device: bool = False) -> None:
        """Host binned-SAH build, or (device=True) the GPU Morton-order build."""
        fn = _lib.rbr_scene_build_bvh_device if device else _lib.rbr_scene_build_bvh
        self.ctx.check(fn(self._h, width, leaf_max))

    def device_nodes(self) -> np.ndarray:
        """The node array as the kernels read it, as uint32 words."""
        n = _sz(0)
        self.ctx.check(_lib.rbr_scene_device_nodes(self._h, None, 0, C.byref(n)))
        out = np.zeros(n.value // 4, np.uint32)
        self.ctx.check(_lib.rbr_scene_device_nodes(self._h, out.ctypes.data, n.value, C.byref(n)))
        return out

    def float_nodes(self) -> np.ndarray:
        """The float wide nodes (bvh.h layout) as uint32 words."""
        nb, pb = _sz(0), _sz(0)
        self.ctx.check(_lib.rbr_scene_bvh_arrays(self._h, None, 0, None, 0, C.byref(nb), C.byref(pb)))
        out = np.zeros(nb.value // 4, np.uint32)
        self.ctx.check(_lib.rbr_scene_bvh_arrays(self._h, out.ctypes.data, nb.value, None, 0, C.byref(nb), C.byref(pb)))
        return out

    def prim_records(self) -> np.ndarray:
        """Leaf-ordered primitive records as float32 (n, 3, 4); word 3 of row 0 holds the id bits."""
        nb, pb = _sz(0), _sz(0)
        self.ctx.check(_lib.rbr_scene_bvh_arrays(self._h, None, 0, None, 0, C.byref(nb), C.byref(pb)))
        out = np.zeros(pb.value // 4, np.float32)
        self.ctx.check(_lib.rbr_scene_bvh_arrays(self._h, None, 0, out.ctypes.data, pb.value, C.byref(nb), C.byref(pb)))
        return out.reshape(-1, 3, 4)

    def stats(self) -> dict:
        st = BvhStats()
        self.ctx.check(_lib.rbr_scene_bvh_stats(self._h, C.byref(st)))
        return st.as_dict()

    def world_bounds(self) -> tuple[np.ndarray, np.ndarray]:
        lo = np.zeros(3, np.float32)
        hi = np.zeros(3, np.float32)
        self.ctx.check(_lib.rbr_scene_world_bounds(self._h, lo.ctypes.data, hi.ctypes.data))
        return lo, hi

    def trace_closest(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.empty(len(rays), HIT_DTYPE)
        self.ctx.check(_lib.rbr_trace_closest(self._h, rays.ctypes.data, len(rays), hits.ctypes.data))
        return hits

    def trace_any(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        occ = np.empty(len(rays), np.uint8)
        self.ctx.check(_lib.rbr_trace_any(self._h, rays.ctypes.data, len(rays), occ.ctypes.data))
        return occ

    def trace_closest_device(self, d_rays: DeviceBuffer, n: int, d_hits: DeviceBuffer) -> None:
        self.ctx.check(_lib.rbr_trace_closest_device(self._h, d_rays.ptr, n, d_hits.ptr))

    def trace_any_device(self, d_rays: DeviceBuffer, n: int, d_occ: DeviceBuffer) -> None:
        self.ctx.check(_lib.rbr_trace_any_device(self._h, d_rays.ptr, n, d_occ.ptr))

    def time_trace(self, kind: int, d_rays: DeviceBuffer, n: int, d_out: DeviceBuffer, iters: int) -> float:
        ms = C.c_float(0.0)
        self.ctx.check(_lib.rbr_time_trace(self._h, kind, d_rays.ptr, n, d_out.ptr, iters, C.byref(ms)))
        return ms.value

    def resident_waves(self, any_hit: bool = False) -> int:
        """Traversal waves per CU resident at once under the context's current tuning."""
        w = C.c_int(0)
        self.ctx.check(_lib.rbr_trace_resident_waves(self._h, 1 if any_hit else 0, C.byref(w)))
        return w.value

    def host_trace_closest(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.empty(len(rays), HIT_DTYPE)
        nv, pt = C.c_uint64(0), C.c_uint64(0)
        self.ctx.check(_lib.rbr_host_trace_closest(self._h, rays.ctypes.data, len(rays), hits.ctypes.data,
                                                   C.byref(nv), C.byref(pt)))
        return hits, nv.value, pt.value

    def host_trace_any(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        occ = np.empty(len(rays), np.uint8)
        nv, pt = C.c_uint64(0), C.c_uint64(0)
        self.ctx.check(_lib.rbr_host_trace_any(self._h, rays.ctypes.data, len(rays), occ.ctypes.data,
                                               C.byref(nv), C.byref(pt)))
        return occ, nv.value, pt.value


# ---------------------------------------------------------------- workloads

def parse_obj(text: bytes) -> np.ndarray:
    """Triangles (n, 9) of Wavefront OBJ text, as the reference's ObjImporter builds them."""
    n, nn, nu = _sz(0), _sz(0), _sz(0)
    if _lib.rbr_obj_parse(text, len(text), None, 0, C.byref(n), C.byref(nn), C.byref(nu)):
        raise ValueError("OBJ import failed: a face names a missing vertex, or no triangle")
    out = np.zeros((n.value, 9), np.float32)
    if _lib.rbr_obj_parse(text, len(text), out.ctypes.data, n.value, C.byref(n), C.byref(nn), C.byref(nu)):
        raise ValueError("OBJ import failed")
    return out


def parse_obj_normals(text: bytes) -> np.ndarray | None:
    """Vertex normals (n, 9) of the triangles parse_obj returns, as written in the file, or
    None when the reference's importer would drop them (some corner without a normal)."""
    n = _sz(0)
    if _lib.rbr_obj_parse_normals(text, len(text), None, 0, C.byref(n)):
        raise ValueError("OBJ import failed: a face names a missing vertex, or no triangle")
    if n.value == 0:
        return None
    out = np.zeros((n.value, 9), np.float32)
    if _lib.rbr_obj_parse_normals(text, len(text), out.ctypes.data, n.value, C.byref(n)):
        raise ValueError("OBJ import failed")
    return out


def load_obj(path: str, with_normals: bool = False):
    """Triangles of an OBJ file; with_normals: (triangles, vertex normals or None)."""
    with open(path, "rb") as f:
        text = f.read()
    return (parse_obj(text), parse_obj_normals(text)) if with_normals else parse_obj(text)


def triangle_shading_normal(verts, normals, o, d) -> np.ndarray:
    """Triangle::intersect's shading normal for a triangle with vertex normals (host)."""
    v, n = np.ascontiguousarray(verts, np.float32).reshape(9), np.ascontiguousarray(normals, np.float32).reshape(9)
    oo, dd = np.ascontiguousarray(o, np.float32).reshape(3), np.ascontiguousarray(d, np.float32).reshape(3)
    out = np.zeros(3, np.float32)
    if _lib.rbr_triangle_shading_normal(v.ctypes.data, n.ctypes.data, oo.ctypes.data, dd.ctypes.data, out.ctypes.data):
        raise ValueError("invalid argument")
    return out


def _film(film) -> np.ndarray:
    f = np.ascontiguousarray(film, np.float32)
    if f.ndim != 3 or f.shape[2] != 3 or f.shape[0] < 1 or f.shape[1] < 1:
        raise ValueError("film must be (H, W, 3)")
    return f


def film_to_rgb8(film, gamma: float = 2.2, tone_map: bool = False, exposure: float = 1.0) -> np.ndarray:
    """Film::toRGB8 / toRGB8ToneMapped (film.cpp:27-63) of an (H, W, 3) mean-radiance film."""
    f = _film(film)
    out = np.empty(f.shape, np.uint8)
    if _lib.rbr_film_to_rgb8(_ptr(f), f.shape[1], f.shape[0], gamma, int(tone_map), exposure, _ptr(out)):
        raise ValueError("rbr_film_to_rgb8: invalid arguments")
    return out


def write_png(path: str, rgb8) -> None:
    """8-bit RGB PNG, ImageExporter::writeRGB8's default format (image_exporter.cpp:13-39)."""
    a = np.ascontiguousarray(rgb8, np.uint8)
    if a.ndim != 3 or a.shape[2] != 3:
        raise ValueError("image must be (H, W, 3) uint8")
    if _lib.rbr_film_write_png(os.fsencode(path), _ptr(a), a.shape[1], a.shape[0]):
        raise OSError(f"cannot write {path}")


def write_pfm(path: str, film) -> None:
    """HDR film as PFM (FP32, no reference counterpart)."""
    f = _film(film)
    if _lib.rbr_film_write_pfm(os.fsencode(path), _ptr(f), f.shape[1], f.shape[0]):
        raise OSError(f"cannot write {path}")


def read_pfm(path: str) -> np.ndarray:
    w, h = C.c_int(0), C.c_int(0)
    if _lib.rbr_film_read_pfm(os.fsencode(path), C.byref(w), C.byref(h), None, 0):
        raise OSError(f"cannot read {path} as an RGB PFM")
    out = np.empty((h.value, w.value, 3), np.float32)
    if _lib.rbr_film_read_pfm(os.fsencode(path), C.byref(w), C.byref(h), _ptr(out), out.size):
        raise OSError(f"truncated PFM {path}")
    return out


def icosphere(n: int = 224, radius: float = 100.0, amp: float = 0.1, seed: int = 0xC1) -> np.ndarray:
    """Displaced geodesic icosphere, 20*n*n triangles, (T, 9) float32."""
    count = _wl.rbrw_icosphere(n, radius, amp, seed, None)
    out = np.empty((count, 9), np.float32)
    _wl.rbrw_icosphere(n, radius, amp, seed, out.ctypes.data)
    return out


def rays_uniform(lo, hi, n: int, seed: int, scale: float = 1.1, tmin: float = 1e-4, tmax: float = np.inf) -> np.ndarray:
    lo = np.ascontiguousarray(lo, np.float32)
    hi = np.ascontiguousarray(hi, np.float32)
    r = np.empty(n, RAY_DTYPE)
    _wl.rbrw_rays_uniform(lo.ctypes.data, hi.ctypes.data, scale, n, seed, tmin, tmax, r.ctypes.data)
    return r


def rays_shadow(lo, hi, n: int, seed: int, tmin: float = 1e-4) -> np.ndarray:
    lo = np.ascontiguousarray(lo, np.float32)
    hi = np.ascontiguousarray(hi, np.float32)
    r = np.empty(n, RAY_DTYPE)
    _wl.rbrw_rays_shadow(lo.ctypes.data, hi.ctypes.data, n, seed, tmin, r.ctypes.data)
    return r


def mesh_bounds(tris: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    v = tris.reshape(-1, 3)
    return v.min(0).astype(np.float32), v.max(0).astype(np.float32)


# ---------------------------------------------------------------- rendering (rbr.h structs)

class RbrMaterial(C.Structure):
    _fields_ = [("base_color", C.c_float * 3), ("metallic", C.c_float), ("roughness", C.c_float), ("ior", C.c_float),
                ("transmission", C.c_float), ("emission", C.c_float * 3)]


class RbrLight(C.Structure):
    _fields_ = [("type", C.c_int32), ("color", C.c_float * 3), ("intensity", C.c_float), ("position", C.c_float * 3),
                ("u", C.c_float * 3), ("v", C.c_float * 3), ("normal", C.c_float * 3), ("direction", C.c_float * 3)]


class RbrMedium(C.Structure):
    _fields_ = [("type", C.c_int32), ("sigma_a", C.c_float * 3), ("sigma_s", C.c_float * 3), ("falloff", C.c_float),
                ("base_height", C.c_float), ("g", C.c_float)]


class RbrCamera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("forward", C.c_float * 3), ("right", C.c_float * 3), ("up", C.c_float * 3),
                ("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float), ("width", C.c_int32),
                ("height", C.c_int32)]


class RbrRenderConfig(C.Structure):
    _fields_ = [("integrator", C.c_int32), ("spp", C.c_int32), ("max_depth", C.c_int32), ("use_nee", C.c_int32),
                ("use_mis", C.c_int32), ("guided", C.c_int32), ("seed", C.c_uint64), ("sample_offset", C.c_uint32),
                ("guide_depth", C.c_uint32), ("ao_samples", C.c_int32), ("ao_distance", C.c_float),
                ("depth_near", C.c_float), ("depth_far", C.c_float), ("ambient", C.c_float * 3)]


class RbrJob(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("camera", RbrCamera), ("config", RbrRenderConfig),
                ("film_rgb", C.POINTER(C.c_float))]


PATH, VOLPATH, NORMAL, DEPTH, AO, WHITTED = 0, 1, 2, 3, 4, 5

_sig(_lib.rbr_scene_set_shading, [_vp, C.POINTER(RbrMaterial), _u32, _vp, _u32, C.POINTER(RbrLight), _u32,
                                  C.POINTER(RbrMedium), _vp])
_sig(_lib.rbr_render, [_vp, C.POINTER(RbrJob), C.c_int])
_sig(_lib.rbr_render_moments, [_vp, C.POINTER(RbrJob), _vp, _vp])
_sig(_lib.rbr_last_render_timing, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_double)])
_sig(_lib.rbr_set_render_tuning, [_vp, C.c_uint32, C.c_int, C.c_int])


def set_render_tuning(ctx: Context, pool_paths: int = 1 << 24, graphs: bool = True, sort: bool = True) -> None:
    """Queue capacity, hipGraph replay of each round and material-ordered shading."""
    ctx.check(_lib.rbr_set_render_tuning(ctx.handle, pool_paths, int(graphs), int(sort)))


def _f3(x):
    return (C.c_float * 3)(*[float(v) for v in x])


def c_materials(mats: list) -> C.Array:
    arr = (RbrMaterial * max(1, len(mats)))()
    for i, m in enumerate(mats):
        arr[i] = RbrMaterial(_f3(m["base_color"]), m["metallic"], m["roughness"], m["ior"], m["transmission"],
                             _f3(m["emission"]))
    return arr


def c_lights(lights: list) -> C.Array:
    arr = (RbrLight * max(1, len(lights)))()
    for i, l in enumerate(lights):
        arr[i] = RbrLight(l["type"], _f3(l["color"]), l["intensity"], _f3(l["position"]), _f3(l["u"]), _f3(l["v"]),
                          _f3(l["normal"]), _f3(l["direction"]))
    return arr


def c_medium(m: dict | None) -> RbrMedium:
    if not m:
        return RbrMedium()
    return RbrMedium(m["type"], _f3(m["sigma_a"]), _f3(m["sigma_s"]), m["falloff"], m["base_height"], m["g"])


def c_camera(cam) -> RbrCamera:
    return RbrCamera(_f3(cam.position), _f3(cam.forward), _f3(cam.right), _f3(cam.up), cam.fx, cam.fy, cam.cx, cam.cy,
                     cam.width, cam.height)


def c_config(spp: int, depth: int, seed: int = 1, integrator: int = PATH, nee: bool = True, mis: bool = True,
             sample_offset: int = 0, guided: int = 0, guide_depth: int = 0, ao_samples: int = 0, ao_distance: float = 0.0,
             depth_near: float = 0.0, depth_far: float = 0.0, ambient=(0.1, 0.1, 0.1)) -> RbrRenderConfig:
    """ambient: the Whitted integrator's Scene::ambientLight (the reference's default 0.1)."""
    return RbrRenderConfig(integrator, spp, depth, int(nee), int(mis), int(guided), seed, sample_offset, guide_depth,
                           ao_samples, ao_distance, depth_near, depth_far, (C.c_float * 3)(*map(float, ambient)))


def _job_config(j: dict) -> RbrRenderConfig:
    return c_config(j["spp"], j["depth"], j.get("seed", 1), j.get("integrator", PATH), j.get("nee", True),
                    j.get("mis", True), j.get("sample_offset", 0), j.get("guided", 0), j.get("guide_depth", 0),
                    j.get("ao_samples", 0), j.get("ao_distance", 0.0), j.get("depth_near", 0.0), j.get("depth_far", 0.0),
                    j.get("ambient", (0.1, 0.1, 0.1)))


def scene_from_desc(ctx: Context, desc, width: int = 4 | 0x100, leaf: int = 4, device_build: bool = False) -> Scene:
    """Uploads a scenes.SceneDesc: geometry, wide BVH (quantized BVH4 by default; host SAH
    build, or the GPU build with device_build) and shading tables."""
    sc = Scene(ctx)
    sc.set_geometry(tris=desc.tris if len(desc.tris) else None, spheres=desc.spheres if len(desc.spheres) else None,
                    quads=desc.quads if len(desc.quads) else None, planes=desc.planes if len(desc.planes) else None)
    if getattr(desc, "tri_normals", None) is not None:
        sc.set_tri_normals(desc.tri_normals)
    sc.build_bvh(width, leaf, device=device_build)
    pm = np.ascontiguousarray(desc.prim_material, np.int32)
    mats, lights, med = c_materials(desc.materials), c_lights(desc.lights), c_medium(desc.medium)
    bg = np.asarray(desc.background, np.float32)
    ctx.check(_lib.rbr_scene_set_shading(sc.handle, mats, len(desc.materials), pm.ctypes.data, len(pm), lights,
                                         len(desc.lights), C.byref(med), bg.ctypes.data))
    sc.desc = desc
    return sc


PROGRESS_FN = C.CFUNCTYPE(None, C.c_float, C.c_void_p)
_sig(_lib.rbr_render_progress, [_vp, C.POINTER(RbrJob), C.c_int, PROGRESS_FN, _vp])


def render(ctx: Context, jobs: list, progress=None) -> list[np.ndarray]:
    """Integrator::render for one or many jobs in one launch sequence.  Each job is a dict
    with scene (a Scene from scene_from_desc), spp, depth and optional seed, integrator, nee,
    mis, camera, sample_offset, guided (1: sample the context's guide, 3: also record training
    samples, 2: record training samples from plain BSDF-sampled paths), guide_depth.  Returns one (H, W, 3) float32 film per job (per-pixel mean radiance).
    progress(fraction): called after each round of the launch (rbr_render_progress)."""
    arr = (RbrJob * len(jobs))()
    films = []
    for i, j in enumerate(jobs):
        cam = j.get("camera") or j["scene"].desc.camera
        film = np.zeros((cam.height, cam.width, 3), np.float32)
        films.append(film)
        arr[i] = RbrJob(j["scene"].handle.value, c_camera(cam), _job_config(j),
                        film.ctypes.data_as(C.POINTER(C.c_float)))
    if progress is None:
        ctx.check(_lib.rbr_render(ctx.handle, arr, len(jobs)))
    else:
        cb = PROGRESS_FN(lambda f, _user: progress(f))
        ctx.check(_lib.rbr_render_progress(ctx.handle, arr, len(jobs), cb, None))
    return films


def render_moments(ctx: Context, job: dict):
    """Per-pixel sum and sum of squares of the single-sample radiance (float64)."""
    cam = job.get("camera") or job["scene"].desc.camera
    film = np.zeros((cam.height, cam.width, 3), np.float32)
    s1 = np.zeros((cam.height, cam.width, 3), np.float64)
    s2 = np.zeros_like(s1)
    j = RbrJob(job["scene"].handle.value, c_camera(cam), _job_config(job), film.ctypes.data_as(C.POINTER(C.c_float)))
    ctx.check(_lib.rbr_render_moments(ctx.handle, C.byref(j), s1.ctypes.data, s2.ctypes.data))
    return film, s1, s2


class RbrRenderStats(C.Structure):
    _fields_ = [("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("shadow_empty", C.c_uint64),
                ("sampled_extension", C.c_uint64), ("sampled_shadow", C.c_uint64),
                ("ext_nodes_per_ray", C.c_double), ("ext_prims_per_ray", C.c_double),
                ("shadow_nodes_per_ray", C.c_double), ("shadow_prims_per_ray", C.c_double)]


_sig(_lib.rbr_set_render_census, [_vp, C.c_uint32])
_sig(_lib.rbr_last_render_stats, [_vp, C.POINTER(RbrRenderStats)])


def set_render_census(ctx: Context, rays_per_queue: int) -> None:
    """Traversal census of later path renders (0: off): sampled rays of every bounce's queues
    are traced by the host traversal to count node visits and primitive tests."""
    ctx.check(_lib.rbr_set_render_census(ctx.handle, rays_per_queue))


def last_render_stats(ctx: Context) -> dict:
    st = RbrRenderStats()
    ctx.check(_lib.rbr_last_render_stats(ctx.handle, C.byref(st)))
    return {f: getattr(st, f) for f, _ in RbrRenderStats._fields_}


def last_render_ms(ctx: Context) -> float:
    t = C.c_double(0)
    st = (C.c_double * 4)()
    ctx.check(_lib.rbr_last_render_timing(ctx.handle, C.byref(t), st))
    return t.value


# ---------------------------------------------------------------- path guiding

_sig(_lib.rbr_guide_param_count, [])
_sig(_lib.rbr_guide_init, [_vp, C.c_uint64, _vp, _vp])
_sig(_lib.rbr_guide_get_params, [_vp, _vp, _sz])
_sig(_lib.rbr_guide_set_params, [_vp, _vp, _sz])
_sig(_lib.rbr_guide_save, [_vp, C.c_char_p])
_sig(_lib.rbr_guide_load, [_vp, C.c_char_p])
_sig(_lib.rbr_guide_eval, [_vp, _vp, _sz, _vp])
_sig(_lib.rbr_guide_loss_grad, [_vp, _vp, _sz, C.POINTER(C.c_double), _vp])
_sig(_lib.rbr_guide_train_step, [_vp, _vp, _sz, C.c_float, C.POINTER(C.c_double)])

_sig(_lib.rbr_guide_sample, [_vp, _vp, _sz, _sz, _vp, _vp, _vp])
_sig(_lib.rbr_guide_reserve, [_vp, C.c_uint32])
_sig(_lib.rbr_guide_time, [_vp, C.c_int, _sz, C.c_int, C.POINTER(C.c_float)])
_sig(_lib.rbr_guide_collected, [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_double)])
_sig(_lib.rbr_guide_train_collected, [_vp, C.c_int, C.c_uint32, C.c_float, C.POINTER(C.c_double)])
_sig(_lib.rbr_guide_clear_collected, [_vp])

GUIDE_PARAMS = _lib.rbr_guide_param_count()


class Guide:
    """The context's path-guiding network (rbr_guide_*)."""

    def __init__(self, ctx: Context, lo, hi, seed: int = 1):
        self.ctx = ctx
        lo = np.ascontiguousarray(lo, np.float32)
        hi = np.ascontiguousarray(hi, np.float32)
        ctx.check(_lib.rbr_guide_init(ctx.handle, seed, lo.ctypes.data, hi.ctypes.data))
        self.lo, self.hi = lo, hi

    @classmethod
    def load(cls, ctx: Context, path: str) -> "Guide":
        """The guide saved at `path` (rbr_guide_load): parameters, scene box, Adam state."""
        g = cls.__new__(cls)
        g.ctx = ctx
        ctx.check(_lib.rbr_guide_load(ctx.handle, os.fsencode(path)))
        g.lo = g.hi = None
        return g

    def save(self, path: str) -> None:
        self.ctx.check(_lib.rbr_guide_save(self.ctx.handle, os.fsencode(path)))

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(GUIDE_PARAMS, np.float32)
        self.ctx.check(_lib.rbr_guide_get_params(self.ctx.handle, p.ctypes.data, GUIDE_PARAMS))
        return p

    @params.setter
    def params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (GUIDE_PARAMS,)
        self.ctx.check(_lib.rbr_guide_set_params(self.ctx.handle, p.ctypes.data, GUIDE_PARAMS))

    def eval(self, queries) -> np.ndarray:
        """queries (n, 9): position, normal, wo -> (n, 8, 5) lobes (mu 3, kappa, weight)."""
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 9)
        out = np.zeros((len(q), 8, 5), np.float32)
        self.ctx.check(_lib.rbr_guide_eval(self.ctx.handle, q.ctypes.data, len(q), out.ctypes.data))
        return out

    def sample(self, queries, u) -> tuple[np.ndarray, np.ndarray]:
        """Directions drawn from each query's mixture: u (nq, m, 3) uniforms -> dirs (nq, m, 3),
        pdfs (nq, m), as the renderer samples them."""
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 9)
        uu = np.ascontiguousarray(u, np.float32).reshape(len(q), -1, 3)
        m = uu.shape[1]
        dirs = np.zeros((len(q), m, 3), np.float32)
        pdfs = np.zeros((len(q), m), np.float32)
        self.ctx.check(_lib.rbr_guide_sample(self.ctx.handle, q.ctypes.data, len(q), m, uu.ctypes.data,
                                             dirs.ctypes.data, pdfs.ctypes.data))
        return dirs, pdfs

    @staticmethod
    def _samples14(samples) -> np.ndarray:
        """(n, 14) float32 samples; (n, 13) ones get a BSDF density of -1 (none)."""
        s = np.asarray(samples, np.float32)
        s = s.reshape(-1, s.shape[-1] if s.ndim == 2 else 14)
        if s.shape[1] == 13:
            s = np.concatenate([s, np.full((len(s), 1), -1.0, np.float32)], axis=1)
        if s.shape[1] != 14:
            raise ValueError("guide samples have 13 or 14 columns")
        return np.ascontiguousarray(s)

    def loss_grad(self, samples) -> tuple[float, np.ndarray]:
        """samples (n, 14): position, normal, wo, direction, weight, BSDF density of the
        direction (<= 0: none; 13 columns: none) -> (mean loss, gradient)."""
        s = self._samples14(samples)
        g = np.zeros(GUIDE_PARAMS, np.float32)
        loss = C.c_double(0)
        self.ctx.check(_lib.rbr_guide_loss_grad(self.ctx.handle, s.ctypes.data, len(s), C.byref(loss), g.ctypes.data))
        return loss.value, g

    def train_step(self, samples, lr: float = 1e-3) -> float:
        s = self._samples14(samples)
        loss = C.c_double(0)
        self.ctx.check(_lib.rbr_guide_train_step(self.ctx.handle, s.ctypes.data, len(s), lr, C.byref(loss)))
        return loss.value

    def time(self, kind: int, n: int, iters: int = 10) -> float:
        """Mean ms of one inference launch (kind 0) or training step (kind 1) on n resident
        synthetic inputs (rbr_guide_time)."""
        ms = C.c_float(0)
        self.ctx.check(_lib.rbr_guide_time(self.ctx.handle, kind, n, iters, C.byref(ms)))
        return ms.value

    def reserve(self, capacity: int) -> None:
        self.ctx.check(_lib.rbr_guide_reserve(self.ctx.handle, capacity))

    def collected(self) -> tuple[int, float]:
        n = C.c_uint64(0)
        cs = C.c_double(0)
        self.ctx.check(_lib.rbr_guide_collected(self.ctx.handle, C.byref(n), C.byref(cs)))
        return n.value, cs.value

    def train_collected(self, steps: int, batch: int = 1 << 18, lr: float = 1e-3) -> float:
        loss = C.c_double(0)
        self.ctx.check(_lib.rbr_guide_train_collected(self.ctx.handle, steps, batch, lr, C.byref(loss)))
        return loss.value

    def clear_collected(self) -> None:
        self.ctx.check(_lib.rbr_guide_clear_collected(self.ctx.handle))
