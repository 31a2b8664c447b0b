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
