"""Seeded test worlds and ray sets modelled on the reference's accelerator test
(tests/test_accelerators.cpp:37-55 worlds, :84-93 rays) plus the C1 icosphere workload."""
from __future__ import annotations

import numpy as np

RAY_DTYPE = np.dtype([("o", "<f4", 3), ("d", "<f4", 3), ("tmin", "<f4"), ("tmax", "<f4")])


def normalize32(d: np.ndarray) -> np.ndarray:
    """Vec3::normalized in FP32 (vector.h:203-209): d * (1 / sqrt((x*x + y*y) + z*z))."""
    d = np.asarray(d, np.float32)
    x, y, z = d[..., 0], d[..., 1], d[..., 2]
    ln = np.sqrt((x * x + y * y) + z * z).astype(np.float32)
    inv = (np.float32(1.0) / ln).astype(np.float32)
    return (d * inv[..., None]).astype(np.float32)


def make_world(rng: np.random.Generator, n_spheres: int, n_tris: int, with_plane: bool) -> dict:
    """makeWorld of test_accelerators.cpp:37-55: random spheres, jittered triangles, 20
    axis-aligned flat triangles and optionally the plane y = -12."""
    pos = lambda *s: rng.uniform(-10.0, 10.0, s).astype(np.float32)
    size = lambda *s: rng.uniform(0.1, 1.5, s).astype(np.float32)
    spheres = np.concatenate([pos(n_spheres, 3), size(n_spheres, 1)], axis=1)
    c = pos(n_tris, 1, 3)
    tris = (c + (size(n_tris, 3, 3) - np.float32(0.8))).reshape(n_tris, 9)
    flat = []
    for _ in range(20):
        z, x, y = pos(3)
        flat.append([x, y, z, x + 2, y, z, x, y + 2, z])
    tris = np.concatenate([tris, np.asarray(flat, np.float32)]).astype(np.float32)
    planes = np.asarray([[0, 1, 0, 12.0]], np.float32) if with_plane else None
    return {"tris": tris, "spheres": spheres, "planes": planes}


def world_ids(w: dict) -> dict:
    """Primitive ids in the reference's insertion order (spheres first, then triangles,
    then the plane), as makeWorld adds them."""
    ns, nt = len(w["spheres"]), len(w["tris"])
    ids = {"sphere_ids": np.arange(ns, dtype=np.uint32), "tri_ids": np.arange(ns, ns + nt, dtype=np.uint32)}
    if w["planes"] is not None:
        ids["plane_ids"] = np.asarray([ns + nt], np.uint32)
    return ids


def make_rays(rng: np.random.Generator, n: int) -> np.ndarray:
    """Ray set of test_accelerators.cpp:84-93: a third axis-aligned, some in the xz plane,
    a quarter with a finite tMax; directions normalized as Ray's constructor does."""
    o = rng.uniform(-14.0, 14.0, (n, 3)).astype(np.float32)
    d = rng.uniform(-1.0, 1.0, (n, 3)).astype(np.float32)
    idx = np.arange(n)
    ax = idx % 3 == 0
    d[ax] = 0.0
    d[ax, 2] = np.where(idx[ax] % 2 == 1, 1.0, -1.0)
    xz = (~ax) & (idx % 7 == 0)
    d[xz, 1] = 0.0
    ok = np.sqrt((d.astype(np.float64) ** 2).sum(1)) >= 1e-3
    o, d, idx = o[ok], d[ok], idx[ok]
    r = np.empty(len(o), RAY_DTYPE)
    r["o"] = o
    r["d"] = normalize32(d)
    r["tmin"] = np.float32(1e-4)
    tmax = np.full(len(o), np.inf, np.float32)
    fin = idx % 4 == 0
    tmax[fin] = rng.uniform(0.5, 40.0, fin.sum()).astype(np.float32)
    r["tmax"] = tmax
    return r


def geometry_kwargs(w: dict) -> dict:
    kw = {"tris": w["tris"], "spheres": w["spheres"]}
    if w["planes"] is not None:
        kw["planes"] = w["planes"]
    kw.update(world_ids(w))
    return kw


def _quad_row(origin, u, v) -> np.ndarray:
    """(origin, u, v, normal) with Quad's normal = cross(u, v).normalized() (quad.cpp:6-10),
    in FP32 without contraction."""
    origin, u, v = (np.asarray(a, np.float32) for a in (origin, u, v))
    n = np.array([u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]], np.float32)
    return np.concatenate([origin, u, v, normalize32(n)]).astype(np.float32)


def make_quad_world(rng: np.random.Generator, n_random: int) -> np.ndarray:
    """Quads as Quad (quad.cpp:12-42) meets them in the reference's scenes: a closed box of
    axis-aligned walls that share edges (Cornell-style), two small rotated boxes whose faces
    meet at edges, and randomly oriented quads of random extent (some long and thin)."""
    rows = []
    s = np.float32(12.0)
    for o, u, v in (((-s, -s, -s), (2 * s, 0, 0), (0, 0, 2 * s)), ((-s, s, -s), (2 * s, 0, 0), (0, 0, 2 * s)),
                    ((-s, -s, s), (2 * s, 0, 0), (0, 2 * s, 0)), ((-s, -s, -s), (0, 2 * s, 0), (0, 0, 2 * s)),
                    ((s, -s, -s), (0, 0, 2 * s), (0, 2 * s, 0))):
        rows.append(_quad_row(o, u, v))
    for size, ang, off in (((3.3, 3.3, 3.3), -18.0, (-4.0, -12.0, -2.0)), ((3.3, 6.6, 3.3), 15.0, (2.0, -12.0, 3.0))):
        a = np.radians(ang)
        c, sn = np.float32(np.cos(a)), np.float32(np.sin(a))
        rot = lambda p: np.array([c * p[0] + sn * p[2], p[1], -sn * p[0] + c * p[2]], np.float32)
        sx, sy, sz = size
        off = np.asarray(off, np.float32)
        for o, u, v in (((0, 0, 0), (sx, 0, 0), (0, sy, 0)), ((0, 0, sz), (sx, 0, 0), (0, sy, 0)),
                        ((0, 0, 0), (0, 0, sz), (0, sy, 0)), ((sx, 0, 0), (0, 0, sz), (0, sy, 0)),
                        ((0, 0, 0), (sx, 0, 0), (0, 0, sz)), ((0, sy, 0), (sx, 0, 0), (0, 0, sz))):
            rows.append(_quad_row(rot(o) + off, rot(u), rot(v)))
    for k in range(n_random):
        o = rng.uniform(-10.0, 10.0, 3)
        u = rng.normal(size=3) * rng.uniform(0.2, 4.0)
        v = rng.normal(size=3) * rng.uniform(0.2, 4.0) * (0.05 if k % 9 == 0 else 1.0)
        rows.append(_quad_row(o, u, v))
    return np.asarray(rows, np.float32)


def make_quad_rays(rng: np.random.Generator, quads: np.ndarray, n: int) -> np.ndarray:
    """Rays aimed at quad points with (u, v) coordinates on the edges and corners (0, 1),
    just inside and outside them, and random; plus near-grazing rays whose |normal . d| is
    around the reference's EPSILON = 1e-4 cut, and the random set of make_rays."""
    n_aim = n // 2
    q = quads[rng.integers(0, len(quads), n_aim)]
    grid = np.array([0.0, 1.0, 0.5, 1e-7, 1 - 1e-7, -1e-6, 1 + 1e-6], np.float32)
    cu = np.where(rng.random(n_aim) < 0.6, grid[rng.integers(0, len(grid), n_aim)], rng.random(n_aim)).astype(np.float32)
    cv = np.where(rng.random(n_aim) < 0.6, grid[rng.integers(0, len(grid), n_aim)], rng.random(n_aim)).astype(np.float32)
    target = q[:, 0:3] + cu[:, None] * q[:, 3:6] + cv[:, None] * q[:, 6:9]
    o = rng.uniform(-11.5, 11.5, (n_aim, 3)).astype(np.float32)
    d = (target - o).astype(np.float32)
    graze = np.arange(n_aim) % 5 == 0
    # Tilt the direction into the quad's plane: keep only a tiny normal component.
    nrm = q[:, 9:12]
    dn = (d * nrm).sum(1, keepdims=True)
    tang = d - dn * nrm
    eps = rng.choice(np.array([0.5e-4, 0.99e-4, 1.01e-4, 2e-4, 1e-3], np.float32), n_aim)
    tl = np.linalg.norm(tang, axis=1, keepdims=True) + 1e-12
    d_g = tang / tl + (np.sign(dn) + (dn == 0)) * eps[:, None] * nrm
    d[graze] = d_g[graze]
    ok = np.linalg.norm(d.astype(np.float64), axis=1) > 1e-3
    r = np.empty(int(ok.sum()), RAY_DTYPE)
    r["o"], r["d"] = o[ok], normalize32(d[ok])
    r["tmin"] = np.float32(1e-4)
    r["tmax"] = np.inf
    return np.concatenate([r, make_rays(rng, n - n_aim)])
