"""Guided path tracing on the GPU.  Guiding changes only how continuation directions are
drawn (one-sample MIS of the trained vMF mixture with the BSDF); NEE and its weights are the
plain integrator's, so a guided image must converge to the plain GPU image -- itself checked
against the CPU oracle (test_gpu_render.py) -- and the per-pixel z-test must show no bias.
These tests also cover the on-device training records and the variance gain of a trained
guide over the plain tracer."""
import sys

import numpy as np
import pytest

import rbr_amd as R
import scenes as S
from test_gpu_render import check_unbiased, zscores

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = R.Context(0)
    yield c
    c.close()


def train(ctx, sc, g, passes, steps, batch=1 << 16, depth=8, seed0=1000):
    losses = []
    for it in range(passes):
        R.render(ctx, [dict(scene=sc, spp=1, depth=depth, seed=seed0 + it, guided=3)])
        losses.append(g.train_collected(steps, batch, 1e-3))
        g.clear_collected()
    return losses


def relvar(s1, s2, n):
    m = s1 / n
    v = np.maximum(s2 / n - m * m, 0.0)
    lum = np.array([0.2126, 0.7152, 0.0722])
    return float(np.mean((v @ lum) / ((m @ lum) ** 2 + 1e-2)))


@pytest.mark.parametrize("trained", [False, True])
def test_guided_cornell_unbiased(ctx, trained):
    say(f"case: guided cornell 48x48 trained={trained}")
    sc = R.scene_from_desc(ctx, S.cornell_c0(48, 48))
    lo, hi = sc.world_bounds()
    g = R.Guide(ctx, lo, hi, seed=4)
    if trained:
        train(ctx, sc, g, 16, 8)
    _, p1, p2 = R.render_moments(ctx, dict(scene=sc, spp=1024, depth=8, seed=21))
    _, g1, g2 = R.render_moments(ctx, dict(scene=sc, spp=1024, depth=8, seed=22, guided=1))
    z, mg, mp, bz, zi = zscores(g1, g2, 1024, p1, p2, 1024)
    say(f"  mean z {z.mean():+.3f}  |z|>4.5 {(np.abs(z) > 4.5).mean():.4f}  image z {zi:+.2f}  "
        f"means {mg.mean():.5f} vs {mp.mean():.5f}")
    check_unbiased(z, mg, mp, bz, zi)


def test_guided_materials_unbiased(ctx):
    """Delta lobes (mirror, glass), rough transmission and mixed lobes: delta samples keep
    their discrete probability (1 - s) x p_bsdf with s the learned selection probability,
    all-delta vertices are not guided."""
    say("case: guided materials 40x30")
    sc = R.scene_from_desc(ctx, S.materials_scene(40, 30))
    lo, hi = sc.world_bounds()
    g = R.Guide(ctx, lo, hi, seed=5)
    train(ctx, sc, g, 8, 8)
    _, p1, p2 = R.render_moments(ctx, dict(scene=sc, spp=2048, depth=8, seed=31))
    _, g1, g2 = R.render_moments(ctx, dict(scene=sc, spp=2048, depth=8, seed=32, guided=1))
    z, mg, mp, bz, zi = zscores(g1, g2, 2048, p1, p2, 2048)
    say(f"  mean z {z.mean():+.3f}  |z|>4.5 {(np.abs(z) > 4.5).mean():.4f}  image z {zi:+.2f}  "
        f"means {mg.mean():.5f} vs {mp.mean():.5f}")
    check_unbiased(z, mg, mp, bz, zi, frac_tail=0.01)


def test_records_training_and_clear(ctx):
    say("case: guide records")
    sc = R.scene_from_desc(ctx, S.cornell_c0(64, 64))
    lo, hi = sc.world_bounds()
    g = R.Guide(ctx, lo, hi, seed=6)
    g.clear_collected()
    assert g.collected() == (0, 0.0)
    # guided=1 samples the guide but records nothing
    R.render(ctx, [dict(scene=sc, spp=1, depth=8, seed=1, guided=1)])
    assert g.collected()[0] == 0
    R.render(ctx, [dict(scene=sc, spp=2, depth=8, seed=1, guided=3)])
    n, cs = g.collected()
    say(f"  records {n}, mean c {cs / max(n, 1):.4g}")
    # a closed box: every camera path hits a diffuse surface and most continue
    assert 64 * 64 * 2 <= n <= 64 * 64 * 2 * 7
    assert cs > 0
    l0 = g.train_collected(1, 4096, 1e-3)
    assert np.isfinite(l0) and l0 > 0
    # more steps on the same records lower the loss
    for _ in range(20):
        last = g.train_collected(4, 4096, 1e-3)
    say(f"  loss {l0:.4f} -> {last:.4f}")
    assert last < l0
    g.clear_collected()
    assert g.collected() == (0, 0.0)
    # capacity: records beyond it are dropped, the count stays at the capacity
    g.reserve(1000)
    R.render(ctx, [dict(scene=sc, spp=1, depth=8, seed=2, guided=3)])
    assert g.collected()[0] == 1000
    g.reserve(1 << 22)


def test_record_only_is_plain_image(ctx):
    """guided = 2 records training samples from the plain integrator's paths: the film is the
    plain one bit for bit, and the records train the guide like guided = 3 records do."""
    say("case: record-only cornell 48x48")
    sc = R.scene_from_desc(ctx, S.cornell_c0(48, 48))
    lo, hi = sc.world_bounds()
    g = R.Guide(ctx, lo, hi, seed=9)
    g.clear_collected()
    plain = R.render(ctx, [dict(scene=sc, spp=4, depth=8, seed=5)])[0]
    rec = R.render(ctx, [dict(scene=sc, spp=4, depth=8, seed=5, guided=2)])[0]
    assert np.array_equal(plain, rec)
    n, cs = g.collected()
    say(f"  records {n}, mean c {cs / max(n, 1):.4g}")
    assert 48 * 48 * 4 <= n <= 48 * 48 * 4 * 7 and cs > 0
    l0 = g.train_collected(1, 4096, 1e-3)
    for _ in range(20):
        last = g.train_collected(4, 4096, 1e-3)
    say(f"  loss {l0:.4f} -> {last:.4f}")
    assert np.isfinite(l0) and last < l0
    g.clear_collected()


def test_guided_errors(ctx):
    sc = R.scene_from_desc(ctx, S.cornell_c0(16, 16))
    with R.Context(0) as c2:
        sc2 = R.scene_from_desc(c2, S.cornell_c0(16, 16))
        with pytest.raises(R.RbrError) as e:   # no rbr_guide_init on this context
            R.render(c2, [dict(scene=sc2, spp=1, depth=2, guided=1)])
        assert e.value.code == 3
    lo, hi = sc.world_bounds()
    R.Guide(ctx, lo, hi, seed=1)
    with pytest.raises(R.RbrError) as e:
        R.render(ctx, [dict(scene=sc, spp=1, depth=2, guided=1, integrator=R.VOLPATH)])
    assert e.value.code == 5
    with pytest.raises(R.RbrError):
        R.render(ctx, [dict(scene=sc, spp=1, depth=2, guided=4)])


def test_trained_guide_reduces_variance(ctx):
    """What the learned part gains: on the Cornell box (81% of the variance is indirect light)
    a trained guide lowers the per-pixel variance of the plain NEE/MIS tracer at equal spp."""
    say("case: variance gain cornell 128x128")
    sc = R.scene_from_desc(ctx, S.cornell_c0(128, 128))
    lo, hi = sc.world_bounds()
    g = R.Guide(ctx, lo, hi, seed=7)
    _, p1, p2 = R.render_moments(ctx, dict(scene=sc, spp=256, depth=8, seed=41))
    _, u1, u2 = R.render_moments(ctx, dict(scene=sc, spp=256, depth=8, seed=42, guided=1))
    losses = train(ctx, sc, g, 64, 16)
    _, g1, g2 = R.render_moments(ctx, dict(scene=sc, spp=256, depth=8, seed=43, guided=1))
    vp, vu, vg = relvar(p1, p2, 256), relvar(u1, u2, 256), relvar(g1, g2, 256)
    say(f"  relative variance: plain {vp:.4g}, untrained guide {vu:.4g}, trained guide {vg:.4g} "
        f"(ratio {vp / vg:.3f}); loss {losses[0]:.3f} -> {losses[-1]:.3f}")
    assert losses[-1] < losses[0]
    assert vg < vu
    assert vp / vg > 1.1


def test_guided_graph_replay_matches_eager(ctx):
    """Guided rounds (inference + guided shading + records) replayed as hipGraphs give the
    eager images; training between replays is picked up (weights are read at run time)."""
    say("case: guided graph replay")
    sc = R.scene_from_desc(ctx, S.cornell_c0(32, 32))
    lo, hi = sc.world_bounds()
    g = R.Guide(ctx, lo, hi, seed=8)
    job = dict(scene=sc, spp=4, depth=8, seed=9, guided=1)
    try:
        for rep in range(3):
            R.set_render_tuning(ctx, 1 << 23, False, False)
            eager = R.render(ctx, [job])[0]
            R.set_render_tuning(ctx, 1 << 23, True, rep == 2)
            replay = R.render(ctx, [job])[0]
            assert np.array_equal(eager, replay), rep
            train(ctx, sc, g, 1, 4, seed0=77 + rep)
    finally:
        R.set_render_tuning(ctx)   # defaults
