"""The autotuner (ray-based-rendering_amd/autotune.py) on small workloads: every traversal
configuration it sweeps returns bit-identical hits (the sweep raises otherwise), each record
names a resident-wave count the launch can hold, the chosen configuration round-trips through
the tuning file and applies to a context; the render sweep keeps images equal across pools,
material sorting, graphs, waves and traversal modes."""
import json
import os
import sys

import numpy as np
import pytest

import rbr_amd as R
import scenes as S

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ray-based-rendering_amd"))
import autotune as A  # noqa: E402

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = R.Context(0)
    yield c
    c.close()


def test_traversal_sweep_identical_and_cache_round_trip(ctx, tmp_path):
    say("case: traversal autotune sweep, 20k tris, 2^15 rays")
    tris = R.icosphere(32, 100.0, 0.1, 0xC1)
    lo, hi = R.mesh_bounds(tris)
    rays = R.rays_uniform(lo, hi, 1 << 15, 3)
    res = A.sweep(ctx, tris, rays, 0, widths=[4, 4 | R.BVH_QUANTIZED], leaves=[4], modes=[0, 1, 7, 11],
                  depths=[8, 12], waves=[8, 32], iters=2, log=lambda s: None)
    # mode 1 is swept at one wave count only
    assert len(res) == 2 * (3 * 2 * 2 + 2)
    for r in res:
        assert r["identical"]
        assert r["ms"] > 0 and r["mrays_s"] > 0
        if r["mode"] != 1:
            assert 1 <= r["resident_waves_per_cu"] <= r["waves_per_cu"]
    best = min(res, key=lambda x: x["ms"])
    path = tmp_path / "traversal.json"
    path.write_text(json.dumps({"closest": best}))
    assert A.best_config("closest", str(path)) == best
    assert A.best_config("any", str(path)) is None
    A.apply(ctx, best)
    sc = R.Scene(ctx)
    sc.set_geometry(tris=tris)
    sc.build_bvh(best["width"], best["leaf"])
    if best["mode"] != 1:
        assert sc.resident_waves(False) == best["resident_waves_per_cu"]
    ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, 0, 8, 32))


def test_any_hit_sweep_identical(ctx):
    say("case: any-hit autotune sweep")
    tris = R.icosphere(24, 100.0, 0.1, 0xC1)
    lo, hi = R.mesh_bounds(tris)
    rays = R.rays_shadow(lo, hi, 1 << 14, 4)
    res = A.sweep(ctx, tris, rays, 1, widths=[8 | R.BVH_QUANTIZED], leaves=[2, 8], modes=[0, 3, 12],
                  depths=[10], waves=[16], iters=2, log=lambda s: None)
    assert len(res) == 6 and all(r["identical"] for r in res)


def test_render_sweep_keeps_images(ctx, tmp_path):
    say("case: render autotune sweep, Cornell 64x64")
    sc = R.scene_from_desc(ctx, S.cornell_c0(64, 64))
    rj = [dict(scene=sc, spp=4, depth=6, seed=3)]
    res = A.render_sweep(ctx, rj, 64 * 64 * 4, pools=[1 << 16, 1 << 20], sorts=[False, True], graphs=[False, True],
                         waves=[16], modes=[0, 11], reps=1, log=lambda s: None)
    assert len(res) == 16
    best = min(res, key=lambda x: x["ms"])
    path = tmp_path / "render.json"
    path.write_text(json.dumps({"workloads": {"c0": {"best": best}}}))
    assert A.render_config("c0", str(path)) == best
    assert A.render_config("c2", str(path)) is None
    ref = R.render(ctx, rj)[0]
    A.apply_render(ctx, best)
    img = R.render(ctx, rj)[0]
    A.apply_render(ctx, None)
    assert np.allclose(img, ref, rtol=2e-6, atol=1e-7)
