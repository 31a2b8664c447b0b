"""The device BVH build (rbr_scene_build_bvh_device): its trees give the same bit-exact hits as
brute force / the host-built tree (hits do not depend on the tree), its structure obeys the
layout's invariants (every primitive in exactly one leaf, padded child boxes that contain
their primitives, leaves of at most leaf_max primitives), its quantized nodes decode to boxes
that contain its float boxes, rendering on it gives the host tree's images, and the host copies
of its arrays (made on first host use) serve concurrent host traversals and repeated builds."""
import sys
import threading

import numpy as np
import pytest

import oracle_py as O
import rbr_amd as R
import scenes as S
import worlds
from test_gpu_traversal import WIDTHS, assert_same

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = R.Context(0)
    yield c
    c.close()


def dev_scene(ctx, width=4, leaf=4, **kw):
    sc = R.Scene(ctx)
    sc.set_geometry(**kw)
    sc.build_bvh(width, leaf, device=True)
    return sc


@pytest.mark.parametrize("collapse", [0, R.BVH_GREEDY, R.BVH_PLOC, R.BVH_PLOC | R.BVH_GREEDY])
@pytest.mark.parametrize("width", WIDTHS)
@pytest.mark.parametrize("cfg", [(30, 30, False, 3000), (400, 1500, True, 6000), (0, 2, False, 500),
                                 (1, 0, False, 200)])
def test_device_build_hits_match_brute_force(ctx, width, cfg, collapse):
    """Device-built trees (SAH-optimal or greedy collapse) give the brute-force hits."""
    ns, nt, plane, nr = cfg
    say(f"case: device build world {cfg} width {width} collapse {collapse}")
    rng = np.random.default_rng(77 + ns + nt)
    w = worlds.make_world(rng, ns, nt, plane)
    rays = worlds.make_rays(rng, nr)
    kw = worlds.geometry_kwargs(w)
    o = O.OracleScene(**kw)
    sc = dev_scene(ctx, width | collapse, 4, **kw)
    assert_same(sc.trace_closest(rays), o.closest(rays, brute=True))
    assert np.array_equal(sc.trace_any(rays), o.any(rays, brute=True))
    # the host traversal runs on the downloaded copy of the same tree
    assert_same(sc.host_trace_closest(rays)[0], o.closest(rays, brute=True))


def test_duplicates_and_leaf_sizes(ctx):
    """Identical triangles (equal Morton codes) and every leaf size: lowest-id ties hold."""
    say("case: device build duplicates")
    tri = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32)
    tris = np.concatenate([tri] * 37)
    ids = np.random.default_rng(3).permutation(np.arange(100, 137)).astype(np.uint32)
    r = R.rays_array([[0.2, 0.2, 1.0]], [[0, 0, -1]])
    for leaf in range(1, 9):
        for width in WIDTHS:
            for tree in (0, R.BVH_PLOC):
                h = dev_scene(ctx, width | tree, leaf, tris=tris, tri_ids=ids).trace_closest(r)
                assert h["id"][0] == 100 and h["t"][0] == np.float32(1.0), (leaf, width, tree)


@pytest.mark.parametrize("collapse", [0, R.BVH_GREEDY, R.BVH_PLOC])
@pytest.mark.parametrize("width", [4, 8])
def test_structure_and_quantization(ctx, width, collapse):
    """Quantized nodes carry the float nodes' refs and decode (fma(1024 + q, 2^e, p), one
    rounding, as the kernels evaluate them) to boxes that contain the float child boxes; every primitive
    sits in exactly one leaf of at most leaf_max primitives, inside its padded child box."""
    say(f"case: device build structure width {width} collapse {collapse}")
    tris = R.icosphere(48, 100.0, 0.1, 0xC1)
    leaf = 4
    sc = dev_scene(ctx, width | R.BVH_QUANTIZED | collapse, leaf, tris=tris)
    st = sc.stats()
    assert st["prim_count"] == len(tris) and st["node_count"] >= 1
    q4 = width // 4
    nq = sc.device_nodes().reshape(-1, 16 if width == 4 else 32)
    nf = sc.float_nodes().reshape(-1, 32 if width == 4 else 64)
    assert nq.shape[0] == nf.shape[0] == st["node_count"]
    refs = nf[:, 6 * width:7 * width]
    assert np.array_equal(nq[:, 4 + 6 * q4:4 + 6 * q4 + width], refs)
    fb = nf[:, :6 * width].view(np.float32).reshape(-1, 6, width)   # lo_x hi_x lo_y hi_y lo_z hi_z
    live = refs != 0xFFFFFFFF
    p = nq[:, 0:3].view(np.float32)
    for ax in range(3):
        if width == 4:   # scales stored as floats (powers of two) in dwords 3, 14, 15
            word = nq[:, (3, 14, 15)[ax]]
            assert np.all(word & 0x807FFFFF == 0)
            e = ((word >> 23) & 0xFF).astype(np.int64) - 127
        else:            # biased exponent bytes in dword 3
            e = ((nq[:, 3] >> (8 * ax)) & 0xFF).astype(np.int64) - 127
        s = np.ldexp(np.float32(1.0), e).astype(np.float32)
        for k in range(width):
            word = nq[:, 4 + ax * (width // 2) + (k >> 1)]   # lo, lo, hi, hi of children 2g, 2g+1
            qlo = 1024 + ((word >> (8 * (k & 1))) & 0xFF).astype(np.float64)
            qhi = 1024 + ((word >> (8 * (2 + (k & 1)))) & 0xFF).astype(np.float64)
            # (1024 + q) * 2^e is exact; float64 sum then float32 rounding stands for the fma
            dlo = (qlo * s + p[:, ax].astype(np.float64)).astype(np.float32)
            dhi = (qhi * s + p[:, ax].astype(np.float64)).astype(np.float32)
            m = live[:, k]
            assert (dlo[m] <= fb[m, 2 * ax, k]).all() and (dhi[m] >= fb[m, 2 * ax + 1, k]).all()
    lo_t = np.minimum(np.minimum(tris[:, 0:3], tris[:, 3:6]), tris[:, 6:9])
    hi_t = np.maximum(np.maximum(tris[:, 0:3], tris[:, 3:6]), tris[:, 6:9])
    order = sc.prim_records()[:, 0, 3].view(np.uint32)   # id of each leaf-ordered record
    seen = np.zeros(len(tris), np.int32)
    for ni, k in zip(*np.nonzero(live & ((refs & 0x80000000) != 0))):
        r = int(refs[ni, k])
        cnt, start = ((r >> 27) & 7) + 1, r & ((1 << 27) - 1)
        assert cnt <= leaf
        ids = order[start:start + cnt]
        seen[ids] += 1
        assert (lo_t[ids] >= fb[ni, [0, 2, 4], k]).all() and (hi_t[ids] <= fb[ni, [1, 3, 5], k]).all()
    assert (seen == 1).all()


def test_c1_size_build_and_hits(ctx):
    """C1's 1.0M-triangle mesh: build time, node count, and hits on 2^20 rays equal to the
    host-built tree's."""
    say("case: device build C1")
    tris = R.icosphere(224, 100.0, 0.1, 0xC1)
    lo, hi = R.mesh_bounds(tris)
    rays = R.rays_uniform(lo, hi, 1 << 20, 0xC1 + 1)
    d = dev_scene(ctx, 4 | R.BVH_QUANTIZED, 4, tris=tris)
    h = R.Scene(ctx)
    h.set_geometry(tris=tris)
    h.build_bvh(4 | R.BVH_QUANTIZED, 4)
    sd, sh = d.stats(), h.stats()
    say(f"  device build {sd['build_ms']:.1f} ms, {sd['node_count']} nodes, depth {sd['max_depth']}; "
        f"host {sh['build_ms']:.1f} ms, {sh['node_count']} nodes, depth {sh['max_depth']}")
    ref = h.trace_closest(rays)
    assert_same(d.trace_closest(rays), ref)
    c = dev_scene(ctx, 4 | R.BVH_QUANTIZED | R.BVH_PLOC, 4, tris=tris)
    sc = c.stats()
    say(f"  clustering build {sc['build_ms']:.1f} ms, {sc['node_count']} nodes, depth {sc['max_depth']}")
    assert_same(c.trace_closest(rays), ref)


def test_render_on_device_built_tree(ctx):
    """Images depend only on the hits, so they are bit-identical on either tree."""
    say("case: render on device-built BVH")
    desc = S.cornell_c0(40, 30)
    a = R.scene_from_desc(ctx, desc)
    b = R.scene_from_desc(ctx, desc, device_build=True)
    fa = R.render(ctx, [dict(scene=a, spp=8, depth=8, seed=3)])[0]
    fb = R.render(ctx, [dict(scene=b, spp=8, depth=8, seed=3)])[0]
    assert np.array_equal(fa, fb)


def test_host_copies_on_first_use_concurrent_and_rebuilt(ctx):
    """A device build leaves the host arrays to be copied on first host use: eight threads that
    start host traversals at once on a fresh tree all get the device's hits; a rebuild of the same
    scene (temporaries reused from the context) with another layout, and a two-scene render
    (whose node arena is assembled on the host), still agree with the host-built tree."""
    say("case: host copies of a device-built tree, concurrent first use")
    rng = np.random.default_rng(91)
    w = worlds.make_world(rng, 200, 3000, False)
    rays = worlds.make_rays(rng, 4096)
    kw = worlds.geometry_kwargs(w)
    sc = dev_scene(ctx, 4 | R.BVH_QUANTIZED, 4, **kw)
    st = sc.stats()
    ref = sc.trace_closest(rays)
    out = [None] * 8

    def work(k):
        out[k] = sc.host_trace_closest(rays)[0]

    threads = [threading.Thread(target=work, args=(k,)) for k in range(8)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for o in out:
        assert_same(o, ref)
    assert st["node_bytes"] == sc.device_nodes().nbytes
    say("case: rebuild on the device with another layout")
    sc.build_bvh(8, 2, device=True)
    assert_same(sc.trace_closest(rays), ref)
    assert_same(sc.host_trace_closest(rays)[0], ref)
    say("case: two-scene render on device-built trees")
    desc = S.cornell_c0(32, 24)
    a1, a2 = R.scene_from_desc(ctx, desc), R.scene_from_desc(ctx, S.cornell_c0(24, 32))
    b1, b2 = R.scene_from_desc(ctx, desc, device_build=True), R.scene_from_desc(ctx, S.cornell_c0(24, 32), device_build=True)
    jobs = lambda x, y: [dict(scene=x, spp=4, depth=6, seed=5), dict(scene=y, spp=4, depth=6, seed=6)]
    fa = R.render(ctx, jobs(a1, a2))
    fb = R.render(ctx, jobs(b1, b2))
    for x, y in zip(fa, fb):
        assert np.array_equal(x, y)
