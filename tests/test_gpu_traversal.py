"""GPU parity of the traversal kernels through the C ABI: closest-hit ids and t must be
bit-identical to the oracle's lowest-id brute force / BVH, any-hit flags identical."""
import os
import sys

import numpy as np
import pytest

import oracle_py as O
import rbr_amd as R
import worlds

pytestmark = pytest.mark.gpu
GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "traversal_small.npz"))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = R.Context(0)
    yield c
    c.close()


# BVH4 / BVH8 with float boxes and with quantized 8-bit child boxes
WIDTHS = [4, 8, 4 | R.BVH_QUANTIZED, 8 | R.BVH_QUANTIZED]


def gpu_scene(ctx, width=4, leaf=4, **kw):
    sc = R.Scene(ctx)
    sc.set_geometry(**kw)
    sc.build_bvh(width, leaf)
    return sc


def assert_same(h, ref):
    bad = np.nonzero((h["id"] != ref["id"]) | (h["t"] != ref["t"]))[0]
    assert len(bad) == 0, f"{len(bad)} mismatches, first {bad[:5]}: {h[bad[:5]]} vs {ref[bad[:5]]}"


def test_empty_and_single_sphere(ctx):
    say("case: empty + single sphere")
    sc = gpu_scene(ctx)
    r = R.rays_array([[0, 0, 5]], [[0, 0, -1]])
    assert sc.trace_closest(r)["id"][0] == R.MISS
    assert sc.trace_any(r)[0] == 0
    sc = gpu_scene(ctx, spheres=np.array([[0, 0, 0, 1]], np.float32))
    h = sc.trace_closest(r)
    assert h["id"][0] == 0 and h["t"][0] == np.float32(4.0)
    assert sc.trace_any(r)[0] == 1


@pytest.mark.parametrize("width", WIDTHS)
def test_golden_fixture(ctx, width):
    say(f"case: golden fixture width {width}")
    rays = GOLDEN["rays"].view(worlds.RAY_DTYPE).reshape(-1)
    sc = gpu_scene(ctx, width, 4, tris=GOLDEN["tris"], spheres=GOLDEN["spheres"], planes=GOLDEN["planes"])
    h = sc.trace_closest(rays)
    ref = O.OracleScene(tris=GOLDEN["tris"], spheres=GOLDEN["spheres"], planes=GOLDEN["planes"]).closest(rays, brute=True)
    assert_same(h, ref)
    # Against the reference's own answers: identical t, ids identical up to exact ties.
    rid, rt = GOLDEN["ref_brute_id"], GOLDEN["ref_brute_t"]
    hit = rid != R.MISS
    assert np.array_equal(h["id"] == R.MISS, ~hit) and np.array_equal(h["t"][hit], rt[hit])
    assert np.array_equal(sc.trace_any(rays), GOLDEN["ref_any"])
    mrays = GOLDEN["mesh_rays"].view(worlds.RAY_DTYPE).reshape(-1)
    ms = gpu_scene(ctx, width, 4, tris=GOLDEN["mesh"])
    mh = ms.trace_closest(mrays)
    assert np.array_equal(mh["t"], np.where(GOLDEN["mesh_ref_id"] == R.MISS, np.float32(0), GOLDEN["mesh_ref_t"]))


@pytest.mark.parametrize("width", WIDTHS)
@pytest.mark.parametrize("cfg", [(30, 30, False, 3000), (400, 1500, False, 6000), (60, 200, True, 3000)])
def test_reference_worlds(ctx, width, cfg):
    ns, nt, plane, nr = cfg
    say(f"case: reference world {cfg} width {width}")
    rng = np.random.default_rng(4242 + ns)
    w = worlds.make_world(rng, ns, nt, plane)
    rays = worlds.make_rays(rng, nr)
    kw = worlds.geometry_kwargs(w)
    o = O.OracleScene(**kw)
    sc = gpu_scene(ctx, width, 4, **kw)
    assert_same(sc.trace_closest(rays), o.closest(rays, brute=True))
    assert np.array_equal(sc.trace_any(rays), o.any(rays, brute=True))


def test_rebuild_replaces_contents(ctx):
    say("case: rebuild")
    rng = np.random.default_rng(1)
    big = worlds.make_world(rng, 400, 1500, False)
    small = worlds.make_world(rng, 30, 30, False)
    sc = gpu_scene(ctx, **worlds.geometry_kwargs(big))
    sc.set_geometry(**worlds.geometry_kwargs(small))
    sc.build_bvh()
    r = R.rays_array([[0, 0, 50]], [[0, 0, -1]])
    assert_same(sc.trace_closest(r), O.OracleScene(**worlds.geometry_kwargs(small)).closest(r, brute=True))


def test_leaf_sizes_and_ties(ctx):
    say("case: leaf sizes and ties")
    tri = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32)
    tris = np.concatenate([tri] * 9)
    ids = np.array([40, 12, 33, 7, 90, 8, 11, 70, 9], np.uint32)
    r = R.rays_array([[0.2, 0.2, 1.0]], [[0, 0, -1]])
    for leaf in range(1, 9):
        for width in WIDTHS:
            h = gpu_scene(ctx, width, leaf, tris=tris, tri_ids=ids).trace_closest(r)
            assert h["id"][0] == 7 and h["t"][0] == np.float32(1.0)


def test_device_pointer_path_matches_host_path(ctx):
    say("case: device pointers")
    tris = R.icosphere(32, 100.0, 0.1, 0xC1)
    lo, hi = R.mesh_bounds(tris)
    rays = R.rays_uniform(lo, hi, 50000, 11)
    sc = gpu_scene(ctx, tris=tris)
    d_rays = ctx.alloc(rays.nbytes)
    d_hits = ctx.alloc(len(rays) * 8)
    d_rays.upload(rays)
    sc.trace_closest_device(d_rays, len(rays), d_hits)
    out = np.empty(len(rays), R.HIT_DTYPE)
    d_hits.download(out)
    assert_same(out, sc.trace_closest(rays))
    d_rays.free()
    d_hits.free()


@pytest.mark.parametrize("width", WIDTHS)
def test_c1_full_size(ctx, width):
    """~1.0M-triangle displaced icosphere (20*224^2) and 2^22 rays: bit-exact against the
    oracle BVH (itself equal to brute force), plus the 2^22 shadow batch."""
    say(f"case: C1 full size width {width}")
    tris = R.icosphere(224, 100.0, 0.1, 0xC1)
    lo, hi = R.mesh_bounds(tris)
    rays = R.rays_uniform(lo, hi, 1 << 22, 0xC1 + 1)
    shadow = R.rays_shadow(lo, hi, 1 << 22, 0xC1 + 2)
    sc = gpu_scene(ctx, width, 4, tris=tris)
    h = sc.trace_closest(rays)
    occ = sc.trace_any(shadow)
    if "c1" not in _ORACLE:   # the oracle's answer does not depend on the GPU layout
        say("  gpu done; oracle")
        o = O.OracleScene(tris=tris)
        o.build_bvh()
        _ORACLE["c1"] = (o.closest(rays), o.any(shadow))
    ref_h, ref_occ = _ORACLE["c1"]
    assert_same(h, ref_h)
    assert np.array_equal(occ, ref_occ)
    assert 0.3 < (h["id"] != R.MISS).mean() < 1.0


_ORACLE: dict = {}


@pytest.mark.parametrize("mode,depth,waves", [(0, 8, 1), (0, 12, 32), (0, 32, 16), (1, 8, 16), (1, 24, 16),
                                               (2, 8, 32), (3, 8, 1), (3, 12, 16), (4, 32, 16),
                                               (5, 8, 32), (6, 8, 1), (6, 12, 16), (7, 32, 16), (8, 8, 24),
                                               (9, 8, 32), (10, 12, 1), (11, 8, 32), (12, 32, 16),
                                               (0 | (11 + 1) << 8, 8, 32), (11 | (0 + 1) << 8, 12, 16),
                                               (7, 10, 32), (11, 10, 24), (3, 10, 16)])
def test_launch_tunings_identical(ctx, mode, depth, waves):
    """Every launch configuration (persistent waves with tiny/large LDS stacks and so
    deep global spills, one ray per lane, persistent waves whose idle lanes fetch new rays,
    in rounds or in single node / primitive steps, and a different mode for any-hit
    launches) returns the same bit-exact hits."""
    say(f"case: tuning mode {mode} lds {depth} waves {waves}")
    tris = R.icosphere(48, 100.0, 0.1, 0xC1)
    lo, hi = R.mesh_bounds(tris)
    rays = R.rays_uniform(lo, hi, 100000, 21)
    o = O.OracleScene(tris=tris)
    o.build_bvh()
    ref = o.closest(rays)
    ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, mode, depth, waves))
    try:
        for width in WIDTHS:
            sc = gpu_scene(ctx, width, 4, tris=tris)
            assert_same(sc.trace_closest(rays), ref)
            assert np.array_equal(sc.trace_any(rays), o.any(rays))
    finally:
        ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, 0, 8, 32))


@pytest.mark.parametrize("mode", [0, 3, 5, 11])
def test_planes_only_every_kernel(ctx, mode):
    """A scene of unbounded planes only (no BVH, so the kernels take their no-node path) in
    the persistent and single-step kernels: the same hits as brute force."""
    say(f"case: planes only mode {mode}")
    rng = np.random.default_rng(77 + mode)
    planes = np.concatenate([rng.normal(size=(5, 3)), rng.uniform(-20, 20, (5, 1))], 1).astype(np.float32)
    planes[:, :3] /= np.linalg.norm(planes[:, :3], axis=1, keepdims=True)
    rays = worlds.make_rays(rng, 5000)
    o = O.OracleScene(planes=planes)
    ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, mode, 8, 32))
    try:
        sc = gpu_scene(ctx, 4 | R.BVH_QUANTIZED, 4, planes=planes)
        h = sc.trace_closest(rays)
        assert_same(h, o.closest(rays, brute=True))
        assert np.array_equal(sc.trace_any(rays), o.any(rays, brute=True))
        assert (h["id"] != R.MISS).any()
    finally:
        ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, 0, 8, 32))


@pytest.mark.parametrize("mode", [0, 1, 3, 11])
def test_watchdog_trips_on_tiny_cap_and_recovers(ctx, mode):
    """A step cap far below what a ray needs makes every kernel family give up on its rays
    (the kernels still end normally) and the synchronous calls report the traversal watchdog;
    the device-pointer launch reports it at the next synchronize.  Back at the default cap
    (the BVH's own node + primitive bound, which a correct traversal never exceeds) the hits
    are bit-exact again."""
    say(f"case: watchdog mode {mode}")
    tris = R.icosphere(24, 100.0, 0.1, 0xC1)
    lo, hi = R.mesh_bounds(tris)
    rays = R.rays_uniform(lo, hi, 20000, 5)
    o = O.OracleScene(tris=tris)
    ref = o.closest(rays, brute=True)
    ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, mode, 8, 32))
    sc = gpu_scene(ctx, 4 | R.BVH_QUANTIZED, 4, tris=tris)
    try:
        ctx.set_trace_step_cap(3)
        for call in (sc.trace_closest, sc.trace_any):
            with pytest.raises(R.RbrError) as e:
                call(rays)
            assert e.value.code == R.ERR_WATCHDOG
        d_rays, d_hits = ctx.alloc(rays.nbytes), ctx.alloc(len(rays) * 8)
        d_rays.upload(rays)
        sc.trace_closest_device(d_rays, len(rays), d_hits)
        with pytest.raises(R.RbrError) as e:
            ctx.synchronize()
        assert e.value.code == R.ERR_WATCHDOG
        ctx.synchronize()   # the flag was cleared by the report
        ctx.set_trace_step_cap(0)
        assert_same(sc.trace_closest(rays), ref)
        assert np.array_equal(sc.trace_any(rays), o.any(rays, brute=True))
        sc.trace_closest_device(d_rays, len(rays), d_hits)
        ctx.synchronize()
        d_rays.free()
        d_hits.free()
    finally:
        ctx.set_trace_step_cap(0)
        ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, 0, 8, 32))


@pytest.mark.parametrize("mode", [5, 8, 9, 11])
def test_step_modes_mixed_primitives(ctx, mode):
    """The single-step kernel (primitive steps every iteration or deferred) on a world of
    spheres, triangles and an unbounded plane, for every layout: the same hits as brute force
    (closest with lowest-id ties, and any hit)."""
    say(f"case: step mode {mode} mixed world")
    rng = np.random.default_rng(500 + mode)
    w = worlds.make_world(rng, 60, 600, True)
    rays = worlds.make_rays(rng, 20000)
    kw = worlds.geometry_kwargs(w)
    o = O.OracleScene(**kw)
    ref_c, ref_a = o.closest(rays, brute=True), o.any(rays, brute=True)
    ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, mode, 8, 32))
    try:
        for width in WIDTHS:
            sc = gpu_scene(ctx, width, 4, **kw)
            assert_same(sc.trace_closest(rays), ref_c)
            assert np.array_equal(sc.trace_any(rays), ref_a)
    finally:
        ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, 0, 8, 32))


@pytest.mark.parametrize("width", WIDTHS)
def test_quads_bit_exact(ctx, width):
    """Quad::intersect (quad.cpp:12-42) through the traversal kernels: Cornell-style walls
    sharing edges, rotated box faces and random parallelograms, with rays aimed at edges and
    corners and near-grazing rays (|n.d| around the 1e-4 cut).  Ids and t equal the oracle's
    lowest-id brute force, for the default kernel and the tuned step kernel."""
    say(f"case: quad world width {width}")
    rng = np.random.default_rng(77)
    q = worlds.make_quad_world(rng, 300)
    rays = worlds.make_quad_rays(rng, q, 20000)
    o = O.OracleScene(quads=q)
    ref_c, ref_a = o.closest(rays, brute=True), o.any(rays, brute=True)
    # with triangles of the same world in front: mixed leaves
    tris = worlds.make_world(rng, 0, 200, False)["tris"]
    mixed_kw = dict(tris=tris, tri_ids=np.arange(len(q), len(q) + len(tris), dtype=np.uint32), quads=q,
                    quad_ids=np.arange(len(q), dtype=np.uint32))
    om = O.OracleScene(**mixed_kw)
    ref_mc = om.closest(rays, brute=True)
    for mode in (0, 11):
        ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, mode, 8, 32))
        try:
            sc = gpu_scene(ctx, width, 4, quads=q)
            assert_same(sc.trace_closest(rays), ref_c)
            assert np.array_equal(sc.trace_any(rays), ref_a)
            assert_same(gpu_scene(ctx, width, 4, **mixed_kw).trace_closest(rays), ref_mc)
        finally:
            ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, 0, 8, 32))


@pytest.mark.parametrize("mode", [7, 11])
def test_step_kernel_staged_results_every_count(ctx, mode):
    """The step kernels stage each 64-ray range's results in LDS and write the range out when
    its last ray is done (a range that finds its slot busy writes directly).  Ray counts below,
    at and around multiples of 64, a device-memory count, and a grid of one wave (every range
    of the launch through one wave's slots): every result is written, bit-exact, and nothing
    past the count is touched."""
    say(f"case: staged results mode {mode}")
    tris = R.icosphere(24, 100.0, 0.1, 0xC1)
    lo, hi = R.mesh_bounds(tris)
    rays = R.rays_uniform(lo, hi, 4097, 9)
    o = O.OracleScene(tris=tris)
    ref_c, ref_a = o.closest(rays, brute=True), o.any(rays, brute=True)
    sc = gpu_scene(ctx, 4 | R.BVH_QUANTIZED, 4, tris=tris)
    try:
        for waves in (32, 1):
            ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, mode, 8, waves))
            for n in (1, 5, 63, 64, 65, 127, 128, 1000, 4097):
                assert_same(sc.trace_closest(rays[:n]), ref_c[:n])
                assert np.array_equal(sc.trace_any(rays[:n]), ref_a[:n])
            # device pointers with a guard region after the count
            n = 1000
            d_rays, d_hits = ctx.alloc(rays.nbytes), ctx.alloc((n + 64) * 8)
            d_rays.upload(rays)
            guard = np.full((n + 64) * 8, 0xAB, np.uint8)
            d_hits.upload(guard)
            sc.trace_closest_device(d_rays, n, d_hits)
            ctx.synchronize()
            out = np.empty((n + 64) * 8, np.uint8)
            d_hits.download(out)
            assert_same(out[: n * 8].view(R.HIT_DTYPE), ref_c[:n])
            assert (out[n * 8:] == 0xAB).all()
            d_rays.free()
            d_hits.free()
    finally:
        ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, 0, 8, 32))


def test_resident_waves_query(ctx):
    """rbr_trace_resident_waves reports the waves per CU the launch keeps resident (the HIP
    occupancy calculator on the kernel the tuning selects): at most the grid asked for, at
    least one, and no more for the step kernel with a deep LDS stack (16 KiB of stack and 3 KiB
    of staged results per wave leave room for 8 of them in 160 KiB) than with a shallow one."""
    tris = R.icosphere(16, 100.0, 0.1, 0xC1)
    sc = gpu_scene(ctx, 4 | R.BVH_QUANTIZED, 4, tris=tris)
    try:
        ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, 11, 8, 32))
        w8 = sc.resident_waves(False)
        say(f"resident waves: mode 11, lds 8: {w8}")
        assert 1 <= w8 <= 32
        ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, 11, 32, 32))
        w32 = sc.resident_waves(False)
        assert 1 <= w32 <= min(w8, 160 * 1024 // (32 * 512 + 3072))
        ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, 11, 8, 4))
        assert sc.resident_waves(True) == 4
    finally:
        ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, 0, 8, 32))


def test_caller_stream(ctx):
    """rbr_ctx_set_stream: the device-pointer launches run on a stream the caller created (the
    asynchronous variant), the synchronous calls wait on it, and NULL returns to the context's
    own stream before the caller destroys its stream; results are those of the own stream."""
    import ctypes as C
    # the HIP runtime librbr.so runs on (a second copy of the runtime would have no device)
    with open("/proc/self/maps") as maps:
        libs = sorted({ln.split()[-1] for ln in maps if "libamdhip64.so" in ln})
    assert libs, "libamdhip64 not loaded"
    hip = C.CDLL(libs[0])
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    tris = R.icosphere(24, 100.0, 0.1, 0xC1)
    lo, hi = R.mesh_bounds(tris)
    rays = R.rays_uniform(lo, hi, 20000, 13)
    o = O.OracleScene(tris=tris)
    ref = o.closest(rays, brute=True)
    sc = gpu_scene(ctx, 4 | R.BVH_QUANTIZED, 4, tris=tris)
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    try:
        ctx.set_stream(stream.value)
        d_rays, d_hits = ctx.alloc(rays.nbytes), ctx.alloc(len(rays) * 8)
        d_rays.upload(rays)
        for mode in (0, 11):
            ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, mode, 8, 32))
            sc.trace_closest_device(d_rays, len(rays), d_hits)
            assert hip.hipStreamSynchronize(stream) == 0   # the caller's own wait
            h = np.empty(len(rays), R.HIT_DTYPE)
            d_hits.download(h)
            assert_same(h, ref)
            assert_same(sc.trace_closest(rays), ref)
        d_rays.free()
        d_hits.free()
    finally:
        ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, 0, 8, 32))
        ctx.set_stream(None)
        assert hip.hipStreamDestroy(stream) == 0
    assert_same(sc.trace_closest(rays), ref)
