"""rbr_trace_closest_device_noalloc (the closest-hit entry point that allocates nothing; its
counters and stack spill rows live in a static scratch array per GPU) against
rbr_trace_closest_device: bit-identical hits on two calls in a row, the second with a
smaller input, under the default launch and the tuned step kernel with a deep LDS stack
(whose 32 waves per CU the scratch cannot all hold, so the grid is lowered)."""
import ctypes as C
import sys

import numpy as np
import pytest

import rbr_amd as R

pytestmark = pytest.mark.gpu

R._lib.rbr_trace_closest_device_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
R._lib.rbr_trace_closest_device_noalloc.restype = C.c_int


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def trace_both(ctx, sc, rays, d_rays, d_a, d_b):
    d_rays.upload(rays)
    sc.trace_closest_device(d_rays, len(rays), d_a)
    ctx.check(R._lib.rbr_trace_closest_device_noalloc(sc.handle, d_rays.ptr, len(rays), d_b.ptr))
    a = np.empty(len(rays), R.HIT_DTYPE)
    b = np.empty(len(rays), R.HIT_DTYPE)
    d_a.download(a)
    d_b.download(b)
    return a, b


@pytest.mark.parametrize("mode,lds,waves", [(0, 8, 32), (11, 12, 32), (1, 8, 16)])
def test_noalloc_equals_device_entry_point(mode, lds, waves):
    say(f"case: no-alloc closest hit, mode {mode} lds {lds} waves {waves}")
    tris = R.icosphere(96, 100.0, 0.1, 0xC1)
    lo, hi = R.mesh_bounds(tris)
    big = R.rays_uniform(lo, hi, 1 << 20, 31)
    small = R.rays_uniform(lo, hi, 50000, 32)
    with R.Context(0) as ctx:
        ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, mode, lds, waves))
        sc = R.Scene(ctx)
        sc.set_geometry(tris=tris)
        sc.build_bvh(4 | R.BVH_QUANTIZED, 4)
        d_rays = ctx.alloc(big.nbytes)
        d_a = ctx.alloc(len(big) * 8)
        d_b = ctx.alloc(len(big) * 8)
        for rays in (big, small):
            a, b = trace_both(ctx, sc, rays, d_rays, d_a, d_b)
            assert np.array_equal(a["id"], b["id"]) and np.array_equal(a["t"], b["t"])
            assert 0.3 < (a["id"] != R.MISS).mean() < 1.0
        for d in (d_rays, d_a, d_b):
            d.free()


def test_noalloc_rejects_oversized_batch():
    """n beyond 2^31 is refused on the host before anything is launched."""
    with R.Context(0) as ctx:
        sc = R.Scene(ctx)
        sc.set_geometry(spheres=np.array([[0, 0, 0, 1]], np.float32))
        sc.build_bvh()
        d = ctx.alloc(64)
        rc = R._lib.rbr_trace_closest_device_noalloc(sc.handle, d.ptr, (1 << 31) + 1, d.ptr)
        assert rc == 1
        assert b"too large" in R._lib.rbr_last_error(ctx.handle)
        d.free()
