"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

liboracle.so : the repository's restatement of the reference (orc_*.cpp)
libref.so    : the reference's own sources compiled from where they lie (ref_shim.cpp)

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this module,
and only as the checker / the baseline — never as the measured or shipped path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")

RAY_DTYPE = np.dtype([("o", "<f4", 3), ("d", "<f4", 3), ("tmin", "<f4"), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("id", "<u4"), ("t", "<f4")])

_vp = C.c_void_p
_u32 = C.c_uint32
_sz = C.c_size_t


def _sig(fn, args, res=None):
    fn.argtypes = args
    fn.restype = res
    return fn


def _load_oracle():
    if not os.path.exists(ORACLE_SO):
        raise ImportError(f"{ORACLE_SO} not built (make -C oracle)")
    lib = C.CDLL(ORACLE_SO)
    _sig(lib.orc_scene_create, [_vp, _u32, _vp, _vp, _u32, _vp, _vp, _u32, _vp, _vp, _u32, _vp], _vp)
    _sig(lib.orc_scene_destroy, [_vp])
    _sig(lib.orc_scene_build_bvh, [_vp])
    for name in ("orc_closest_brute", "orc_closest_bvh", "orc_any_brute", "orc_any_bvh"):
        _sig(getattr(lib, name), [_vp, _vp, _sz, _vp, C.c_int])
    _sig(lib.orc_triangle_t, [_vp, _vp, C.POINTER(C.c_float)], C.c_int)
    _sig(lib.orc_max_threads, [], C.c_int)
    return lib


_orc = _load_oracle()
_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def _load_ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise ImportError(f"{REF_SO} not built (make -C oracle ref, where the reference tree exists)")
        lib = C.CDLL(REF_SO)
        _sig(lib.ref_world_create, [_vp, _u32, _vp, _u32, _vp, _u32, _vp, _u32], _vp)
        _sig(lib.ref_world_destroy, [_vp])
        _sig(lib.ref_bvh_build, [_vp], C.c_double)
        _sig(lib.ref_bvh_closest, [_vp, _vp, _sz, _vp, C.c_int], C.c_double)
        _sig(lib.ref_bvh_any, [_vp, _vp, _sz, _vp, C.c_int], C.c_double)
        _sig(lib.ref_brute_closest, [_vp, _vp, _sz, _vp, C.c_int])
        _sig(lib.ref_brute_any, [_vp, _vp, _sz, _vp, C.c_int])
        _sig(lib.ref_max_threads, [], C.c_int)
        _sig(lib.ref_film_to_rgb8, [_vp, C.c_int, C.c_int, C.c_float, C.c_int, C.c_float, _vp])
        _sig(lib.ref_obj_parse, [C.c_char_p, _sz, _vp, _sz], C.c_long)
        _sig(lib.ref_obj_parse_normals, [C.c_char_p, _sz, _vp, _sz], C.c_long)
        _sig(lib.ref_triangle_normal, [_vp, _vp, _vp, _vp, _vp, _vp, _vp], C.c_int)
        _ref = lib
    return _ref


def ref_film_to_rgb8(film, gamma: float = 2.2, tone_map: bool = False, exposure: float = 1.0):
    """The reference's Film::toRGB8 / toRGB8ToneMapped (film.cpp:33-63) on an (H, W, 3) film."""
    import numpy as np
    f = np.ascontiguousarray(film, np.float32)
    out = np.empty(f.shape, np.uint8)
    _load_ref().ref_film_to_rgb8(f.ctypes.data, f.shape[1], f.shape[0], gamma, int(tone_map), exposure, out.ctypes.data)
    return out


def ref_obj_parse(text: bytes):
    """The reference's ObjImporter::parse: (n, 9) float32 triangles, or None when it fails."""
    import numpy as np
    lib = _load_ref()
    n = lib.ref_obj_parse(text, len(text), None, 0)
    if n < 0:
        return None
    out = np.zeros((n, 9), np.float32)
    lib.ref_obj_parse(text, len(text), out.ctypes.data, n)
    return out


def ref_obj_parse_normals(text: bytes):
    """The vertex normals the reference's ObjImporter keeps, (n, 9) per triangle corner, None
    when it keeps none; raises ValueError when the import fails."""
    import numpy as np
    lib = _load_ref()
    n = lib.ref_obj_parse_normals(text, len(text), None, 0)
    if n < 0:
        raise ValueError("reference import failed")
    if n == 0:
        return None
    out = np.zeros((n, 9), np.float32)
    lib.ref_obj_parse_normals(text, len(text), out.ctypes.data, n)
    return out


def ref_triangle_normal(verts, normals, o, d):
    """The reference's Triangle::intersect with vertex normals: (normal faced towards the ray,
    frontFace, the normalized direction it used), or None on a miss."""
    import numpy as np
    v, n = np.ascontiguousarray(verts, np.float32).reshape(9), np.ascontiguousarray(normals, np.float32).reshape(9)
    oo, dd = np.ascontiguousarray(o, np.float32).reshape(3), np.ascontiguousarray(d, np.float32).reshape(3)
    dn, out = np.zeros(3, np.float32), np.zeros(3, np.float32)
    front = C.c_int(0)
    if not _load_ref().ref_triangle_normal(v.ctypes.data, n.ctypes.data, oo.ctypes.data, dd.ctypes.data, dn.ctypes.data,
                                           out.ctypes.data, C.byref(front)):
        return None
    return out, bool(front.value), dn


def _arr(a, cols):
    if a is None:
        return None, 0
    a = np.ascontiguousarray(a, np.float32).reshape(-1, cols)
    return a, len(a)


def _p(a):
    return None if a is None else a.ctypes.data


def max_threads() -> int:
    return _orc.orc_max_threads()


class OracleScene:
    """The oracle's primitive set: lowest-id closest hit by brute force or via its BVH."""

    def __init__(self, tris=None, tri_ids=None, spheres=None, sphere_ids=None, quads=None, quad_ids=None,
                 planes=None, plane_ids=None):
        t, nt = _arr(tris, 9)
        s, ns = _arr(spheres, 4)
        q, nq = _arr(quads, 12)
        p, np_ = _arr(planes, 4)
        ids = [None if x is None else np.ascontiguousarray(x, np.uint32) for x in (tri_ids, sphere_ids, quad_ids, plane_ids)]
        self._keep = (t, s, q, p, ids)
        self._h = _orc.orc_scene_create(_p(t), nt, _p(ids[0]), _p(s), ns, _p(ids[1]), _p(q), nq, _p(ids[2]),
                                        _p(p), np_, _p(ids[3]))
        self._bvh = False

    def close(self):
        if self._h:
            _orc.orc_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def build_bvh(self):
        _orc.orc_scene_build_bvh(self._h)
        self._bvh = True

    def closest(self, rays, brute: bool = False, threads: int = 0) -> np.ndarray:
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.empty(len(rays), HIT_DTYPE)
        fn = _orc.orc_closest_brute if (brute or not self._bvh) else _orc.orc_closest_bvh
        fn(self._h, rays.ctypes.data, len(rays), hits.ctypes.data, threads)
        return hits

    def any(self, rays, brute: bool = False, threads: int = 0) -> np.ndarray:
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        occ = np.empty(len(rays), np.uint8)
        fn = _orc.orc_any_brute if (brute or not self._bvh) else _orc.orc_any_bvh
        fn(self._h, rays.ctypes.data, len(rays), occ.ctypes.data, threads)
        return occ


def triangle_t(tri9, ray) -> float | None:
    tri = np.ascontiguousarray(tri9, np.float32).reshape(9)
    r = np.ascontiguousarray(np.asarray(ray, RAY_DTYPE).reshape(1))
    t = C.c_float(0.0)
    ok = _orc.orc_triangle_t(tri.ctypes.data, r.ctypes.data, C.byref(t))
    return t.value if ok else None


class RefWorld:
    """The reference's own Triangle/Sphere/Quad/Plane + BVH, compiled from its sources.
    Primitive ids follow the input order (tris, spheres, quads, planes)."""

    def __init__(self, tris=None, spheres=None, quads=None, planes=None):
        lib = _load_ref()
        t, nt = _arr(tris, 9)
        s, ns = _arr(spheres, 4)
        q, nq = _arr(quads, 12)
        p, np_ = _arr(planes, 4)
        self._keep = (t, s, q, p)
        self._lib = lib
        self._h = lib.ref_world_create(_p(t), nt, _p(s), ns, _p(q), nq, _p(p), np_)

    def close(self):
        if self._h:
            self._lib.ref_world_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def build_bvh(self) -> float:
        return self._lib.ref_bvh_build(self._h)

    def bvh_closest(self, rays, threads: int = 0):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.empty(len(rays), HIT_DTYPE)
        sec = self._lib.ref_bvh_closest(self._h, rays.ctypes.data, len(rays), hits.ctypes.data, threads)
        return hits, sec

    def bvh_any(self, rays, threads: int = 0):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        occ = np.empty(len(rays), np.uint8)
        sec = self._lib.ref_bvh_any(self._h, rays.ctypes.data, len(rays), occ.ctypes.data, threads)
        return occ, sec

    def brute_closest(self, rays, threads: int = 0):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.empty(len(rays), HIT_DTYPE)
        self._lib.ref_brute_closest(self._h, rays.ctypes.data, len(rays), hits.ctypes.data, threads)
        return hits

    def brute_any(self, rays, threads: int = 0):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        occ = np.empty(len(rays), np.uint8)
        self._lib.ref_brute_any(self._h, rays.ctypes.data, len(rays), occ.ctypes.data, threads)
        return occ


# ---------------------------------------------------------------- rendering

def _render_structs():
    import rbr_amd  # the C ABI struct layouts (rbr.h) shared by the oracle entry points

    return rbr_amd


_sig(_orc.orc_rscene_create, [_vp, _u32, _vp, _vp, _u32, _vp, _vp, _u32, _vp, _vp, _u32, _vp,
                              _vp, _u32, _vp, _u32, _vp, _u32, _vp, _vp, _vp], _vp)
_sig(_orc.orc_rscene_destroy, [_vp])
_sig(_orc.orc_render, [_vp, _vp, _vp, _vp, C.c_int])
_sig(_orc.orc_render_moments, [_vp, _vp, _vp, _vp, _sz, _vp, _vp, C.c_int])


def _geom(desc):
    t, nt = _arr(desc.tris if len(desc.tris) else None, 9)
    s, ns = _arr(desc.spheres if len(desc.spheres) else None, 4)
    q, nq = _arr(desc.quads if len(desc.quads) else None, 12)
    p, np_ = _arr(desc.planes if len(desc.planes) else None, 4)
    return (t, nt), (s, ns), (q, nq), (p, np_)


class OracleRender:
    """The oracle's restated Integrator::render over a scenes.SceneDesc (lowest-id BVH)."""

    def __init__(self, desc):
        R = _render_structs()
        (t, nt), (s, ns), (q, nq), (p, np_) = _geom(desc)
        pm = np.ascontiguousarray(desc.prim_material, np.int32)
        mats, lights, med = R.c_materials(desc.materials), R.c_lights(desc.lights), R.c_medium(desc.medium)
        bg = np.asarray(desc.background, np.float32)
        tn = getattr(desc, "tri_normals", None)
        tn = None if tn is None else np.ascontiguousarray(tn, np.float32).reshape(-1, 9)
        self._keep = (t, s, q, p, pm, mats, lights, med, bg, tn)
        self.desc = desc
        self._R = R
        self._h = _orc.orc_rscene_create(_p(t), nt, None, _p(s), ns, None, _p(q), nq, None, _p(p), np_, None,
                                         C.addressof(mats), len(desc.materials), pm.ctypes.data, len(pm),
                                         C.addressof(lights), len(desc.lights), C.addressof(med), bg.ctypes.data,
                                         _p(tn))

    def close(self):
        if self._h:
            _orc.orc_rscene_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def render(self, spp: int, depth: int, integrator: int = 0, nee: bool = True, mis: bool = True, camera=None,
               threads: int = 0) -> np.ndarray:
        """Integrator::render with the reference's per-row sampler seeding (y * width)."""
        cam = camera or self.desc.camera
        c = self._R.c_camera(cam)
        cfg = self._R.c_config(spp, depth, 0, integrator, nee, mis)
        film = np.zeros((cam.height, cam.width, 3), np.float32)
        _orc.orc_render(self._h, C.addressof(c), C.addressof(cfg), film.ctypes.data, threads)
        return film

    def moments(self, spp: int, depth: int, seed: int = 7, integrator: int = 0, nee: bool = True, mis: bool = True,
                pixels=None, camera=None, threads: int = 0):
        """Per-pixel sum and sum of squares (float64) of `spp` single-sample estimates."""
        cam = camera or self.desc.camera
        c = self._R.c_camera(cam)
        cfg = self._R.c_config(spp, depth, seed, integrator, nee, mis)
        s1 = np.zeros((cam.height * cam.width, 3), np.float64)
        s2 = np.zeros_like(s1)
        px = None if pixels is None else np.ascontiguousarray(pixels, np.uint32)
        _orc.orc_render_moments(self._h, C.addressof(c), C.addressof(cfg), _p(px), 0 if px is None else len(px),
                                s1.ctypes.data, s2.ctypes.data, threads)
        return s1, s2


class RefRender:
    """The reference's own Scene + PathTracingIntegrator / VolumetricPathIntegrator."""

    def __init__(self, desc):
        lib = _load_ref()
        if not hasattr(lib, "_render_sig"):
            _sig(lib.ref_scene_create, [_vp, _u32, _vp, _u32, _vp, _u32, _vp, _u32, _vp, _u32, _vp, _vp, _u32, _vp, _vp, _vp],
                 _vp)
            _sig(lib.ref_scene_destroy, [_vp])
            _sig(lib.ref_render, [_vp, _vp, _vp, _vp, C.c_int], C.c_double)
            lib._render_sig = True
        R = _render_structs()
        (t, nt), (s, ns), (q, nq), (p, np_) = _geom(desc)
        pm = np.ascontiguousarray(desc.prim_material, np.int32)
        mats, lights, med = R.c_materials(desc.materials), R.c_lights(desc.lights), R.c_medium(desc.medium)
        bg = np.asarray(desc.background, np.float32)
        tn = getattr(desc, "tri_normals", None)
        tn = None if tn is None else np.ascontiguousarray(tn, np.float32).reshape(-1, 9)
        self._keep = (t, s, q, p, pm, mats, lights, med, bg, tn)
        self._lib, self._R, self.desc = lib, R, desc
        self._h = lib.ref_scene_create(_p(t), nt, _p(s), ns, _p(q), nq, _p(p), np_, C.addressof(mats),
                                       len(desc.materials), pm.ctypes.data, C.addressof(lights), len(desc.lights),
                                       C.addressof(med), bg.ctypes.data, _p(tn))

    def close(self):
        if self._h:
            self._lib.ref_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def render(self, spp: int, depth: int, integrator: int = 0, nee: bool = True, mis: bool = True, camera=None,
               threads: int = 0):
        cam = camera or self.desc.camera
        c = self._R.c_camera(cam)
        cfg = self._R.c_config(spp, depth, 0, integrator, nee, mis)
        film = np.zeros((cam.height, cam.width, 3), np.float32)
        sec = self._lib.ref_render(self._h, C.addressof(c), C.addressof(cfg), film.ctypes.data, threads)
        return film, sec
