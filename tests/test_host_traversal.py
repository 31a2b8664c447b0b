"""CPU tests of the product library without a GPU: the C ABI loads and exports every
symbol include/*.h declares, and the wide-BVH build + host traversal (the path GpuBVH's
per-ray intersect takes) returns exactly the oracle's lowest-id closest hit."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle_py as O
import rbr_amd as R
import worlds

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header: str) -> list[str]:
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s+((?:rbr|rbrw)_[a-z_0-9]+)\s*\(", text, re.M)))


@pytest.mark.parametrize("header,lib", [("rbr.h", "librbr.so"), ("rbr_workloads.h", "librbr_workloads.so")])
def test_library_exports_every_declared_symbol(header, lib):
    names = declared_functions(header)
    assert len(names) >= 5
    so = ctypes.CDLL(os.path.join(R.LIB_DIR, lib))
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_host_context_rejects_kernels():
    ctx = R.Context(-1)
    sc = R.Scene(ctx)
    sc.set_geometry(spheres=np.array([[0, 0, 0, 1]], np.float32))
    sc.build_bvh()
    with pytest.raises(R.RbrError):
        sc.trace_closest(R.rays_array([[0, 0, 5]], [[0, 0, -1]]))
    with pytest.raises(R.RbrError):
        R.Scene(ctx).trace_closest(R.rays_array([[0, 0, 5]], [[0, 0, -1]]))  # no BVH yet


def test_invalid_arguments():
    ctx = R.Context(-1)
    sc = R.Scene(ctx)
    sc.set_geometry(tris=np.zeros((1, 9), np.float32))
    for w, leaf in ((3, 4), (4, 0), (4, 9), (16, 4)):
        with pytest.raises(R.RbrError):
            sc.build_bvh(w, leaf)
    with pytest.raises(R.RbrError):
        sc.set_geometry(tris=np.zeros((1, 9), np.float32), tri_ids=np.array([0xFFFFFFFF], np.uint32))


def test_empty_and_bounds():
    ctx = R.Context(-1)
    sc = R.Scene(ctx)
    sc.set_geometry()
    sc.build_bvh()
    lo, hi = sc.world_bounds()
    assert np.all(lo > hi)  # empty box, as BVH::worldBounds (bvh.cpp:43-45)
    h, _, _ = sc.host_trace_closest(R.rays_array([[0, 0, 5]], [[0, 0, -1]]))
    assert h["id"][0] == R.MISS


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("leaf", [1, 4, 8])
def test_host_wide_bvh_equals_oracle(width, leaf):
    rng = np.random.default_rng(99 + width + leaf)
    w = worlds.make_world(rng, 60, 400, True)
    rays = worlds.make_rays(rng, 4000)
    kw = worlds.geometry_kwargs(w)
    ref = O.OracleScene(**kw).closest(rays, brute=True)
    ctx = R.Context(-1)
    sc = R.Scene(ctx)
    sc.set_geometry(**kw)
    sc.build_bvh(width, leaf)
    h, nodes, prims = sc.host_trace_closest(rays)
    assert np.array_equal(h["id"], ref["id"]) and np.array_equal(h["t"], ref["t"])
    occ, _, _ = sc.host_trace_any(rays)
    assert np.array_equal(occ, O.OracleScene(**kw).any(rays, brute=True))
    assert nodes > 0 and prims > 0


def test_host_traversal_icosphere_mesh():
    tris = R.icosphere(24, 100.0, 0.1, 0xC1)
    lo, hi = R.mesh_bounds(tris)
    rays = R.rays_uniform(lo, hi, 20000, 3)
    o = O.OracleScene(tris=tris)
    o.build_bvh()
    ref = o.closest(rays)
    ctx = R.Context(-1)
    sc = R.Scene(ctx)
    sc.set_geometry(tris=tris)
    sc.build_bvh(4, 4)
    h, _, _ = sc.host_trace_closest(rays)
    assert np.array_equal(h["id"], ref["id"]) and np.array_equal(h["t"], ref["t"])
    assert sc.stats()["prim_count"] == len(tris) == 20 * 24 * 24


def test_icosphere_is_watertight_and_seeded():
    a = R.icosphere(6, 100.0, 0.1, 5)
    b = R.icosphere(6, 100.0, 0.1, 5)
    c = R.icosphere(6, 100.0, 0.1, 6)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    # Every edge is shared by exactly two triangles with bit-identical endpoints.
    v = a.reshape(-1, 3, 3)
    keys = {}
    for t in v:
        for i in range(3):
            e = tuple(sorted([tuple(t[i]), tuple(t[(i + 1) % 3])]))
            keys[e] = keys.get(e, 0) + 1
    assert set(keys.values()) == {2}
    r = np.linalg.norm(v.reshape(-1, 3), axis=1)
    assert r.min() > 89.9 and r.max() < 110.1
