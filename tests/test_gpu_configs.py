"""Image parity of the wavefront path tracer on the benchmark configurations.

  C0  Cornell box 256x256, depth 8: the relMSE against the converged 4096-spp image of the
      CPU oracle (tests/golden/c0_oracle_256_4096spp.npz, FP64 accumulation) must sit at the
      Monte Carlo bound at equal spp, and a per-pixel z-test against it must show no bias.
  C2  terrain at 1920x1080 camera geometry: a stratified subset of pixels, z-tested against
      the oracle rendering those pixels on the host cores.
  C3  the 16 fog jobs in ONE rbr_render call: each job bit-equal to its own single-job render
      and z-tested against its own oracle.
  C4  guided rendering on the C2 scene: a trained guide's image z-tested against the plain
      GPU image and against the oracle's pixel subset.

The oracle (oracle/orc_render.cpp) gives the reference's own integrators' images bit for bit
on the same random stream (test_oracle_render.py); here the GPU's independent stream makes
parity statistical."""
import os
import sys

import numpy as np
import pytest

import oracle_py as O
import rbr_amd as R
import scenes as S
from test_gpu_render import check_unbiased, zscores

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "c0_oracle_256_4096spp.npz")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = R.Context(0)
    yield c
    c.close()


def relmse(x, ref):
    """mean over pixels and channels of (x - ref)^2 / (ref^2 + 1e-2)."""
    x = np.asarray(x, np.float64).reshape(-1, 3)
    ref = np.asarray(ref, np.float64).reshape(-1, 3)
    return float(((x - ref) ** 2 / (ref * ref + 1e-2)).mean())


def mc_bound(var, ref, n, n_ref):
    """Expected relMSE of an unbiased n-spp image against an n_ref-spp reference, from the
    per-pixel variance of one sample."""
    var = np.asarray(var, np.float64).reshape(-1, 3)
    ref = np.asarray(ref, np.float64).reshape(-1, 3)
    return float((var * (1.0 / n + 1.0 / n_ref) / (ref * ref + 1e-2)).mean())


def stratified_pixels(w, h, nx, ny, seed):
    """One random pixel in each cell of an nx x ny grid over the film (indices y * w + x)."""
    rng = np.random.default_rng(seed)
    xs = (np.arange(nx)[None, :] * w) // nx + rng.integers(0, max(1, w // nx), (ny, nx))
    ys = (np.arange(ny)[:, None] * h) // ny + rng.integers(0, max(1, h // ny), (ny, nx))
    return np.sort((np.minimum(ys, h - 1) * w + np.minimum(xs, w - 1)).reshape(-1)).astype(np.uint32)


def check_blocks(g1, g2, ng, r1, r2, nr, w, h, bs, zmax=4.5):
    """Bias test over bs x bs pixel blocks: each block's summed mean radiance against the
    other side's, z = (S_a - S_b) / sqrt(sum var_a / n_a + var_b / n_b).  A block averages
    bs^2 * n samples, so its z is close to normal even where single pixels are heavy-tailed
    or the two estimators differ in skew (a per-pixel mean z is then biased without any bias
    in the images)."""
    def per_block(a1, a2, n):
        m = a1.reshape(h, w, 3) / n
        v = np.maximum(a2.reshape(h, w, 3) / n - m * m, 0.0) / n
        hb, wb = h // bs, w // bs
        m = m[:hb * bs, :wb * bs].reshape(hb, bs, wb, bs, 3).sum((1, 3))
        v = v[:hb * bs, :wb * bs].reshape(hb, bs, wb, bs, 3).sum((1, 3))
        return m, v
    ma, va = per_block(g1, g2, ng)
    mb, vb = per_block(r1, r2, nr)
    z = (ma - mb) / np.sqrt(va + vb + 1e-30)
    nb = z.shape[0] * z.shape[1]
    say(f"  {nb} blocks of {bs}x{bs}: mean z {z.mean():+.3f}, max |z| {np.abs(z).max():.2f}")
    assert abs(float(z.mean())) < 5.0 / np.sqrt(nb)
    assert float(np.abs(z).max()) < zmax + np.sqrt(2 * np.log(max(nb, 2)))


# ---------------------------------------------------------------------------------- C0

@pytest.mark.parametrize("spp", [16, 256])
def test_c0_relmse_at_mc_bound(ctx, spp):
    """North-star image criterion: at equal spp the GPU image's relMSE against the converged
    oracle image equals the Monte Carlo bound computed from the oracle's per-pixel variance
    (an unbiased estimator with the reference's variance lands at ratio 1; oracle renders
    with five seeds gave 0.98-1.03).  A 1 % brightness bias at 256 spp would add about 10 %."""
    f = np.load(GOLDEN)
    ref, var, n_ref = f["mean"], f["var"], int(f["spp"])
    sc = R.scene_from_desc(ctx, S.cornell_c0(256, 256))
    img = R.render(ctx, [dict(scene=sc, spp=spp, depth=int(f["depth"]), seed=1234 + spp)])[0]
    e, b = relmse(img, ref), mc_bound(var, ref, spp, n_ref)
    say(f"case: C0 256x256 {spp} spp: relMSE {e:.6f}, MC bound {b:.6f}, ratio {e / b:.4f}")
    assert 0.9 < e / b < 1.1


def test_c0_unbiased_vs_converged_oracle(ctx):
    """Per-pixel z-test of a 4096-spp GPU render against the 4096-spp oracle image (equal
    sample counts: with skewed per-pixel radiance a z-score of unequal sample counts has a
    nonzero mean under no bias), and the same test over 16x16 blocks."""
    say("case: C0 256x256 4096 spp z-test vs the converged oracle")
    f = np.load(GOLDEN)
    n_ref = int(f["spp"])
    ref, var = f["mean"].astype(np.float64), f["var"].astype(np.float64)
    r1 = ref * n_ref
    r2 = (var * (n_ref - 1) + ref * ref * n_ref)
    sc = R.scene_from_desc(ctx, S.cornell_c0(256, 256))
    _, g1, g2 = R.render_moments(ctx, dict(scene=sc, spp=n_ref, depth=int(f["depth"]), seed=77))
    z, mg, mr, bz, zi = zscores(g1, g2, n_ref, r1, r2, n_ref)
    say(f"  mean z {z.mean():+.4f}  |z|>4.5 {(np.abs(z) > 4.5).mean():.5f}  image z {zi:+.2f}  "
        f"means {mg.mean():.6f} vs {mr.mean():.6f}")
    check_unbiased(z, mg, mr, bz, zi, frac_tail=0.002, mean_tol=0.01)
    check_blocks(g1, g2, n_ref, r1, r2, n_ref, 256, 256, 16)


# ---------------------------------------------------------------------------------- C2

_C2 = {}


def c2_oracle(desc, pixels, spp):
    key = (len(pixels), spp)
    if key not in _C2:
        _C2[key] = O.OracleRender(desc).moments(spp, 8, seed=0xC2, pixels=pixels)
    return _C2[key]


def test_c2_terrain_pixel_subset_vs_oracle(ctx):
    """C2 at 1920x1080 (524,288 terrain triangles, 64 spheres of random GGX / metal / glass
    materials, 4 area lights + a sun), depth 8: the GPU's full frame at 1024 spp against the
    oracle on 4,096 stratified pixels at 1024 spp."""
    say("case: C2 terrain 1080p, 4096 stratified pixels x 1024 spp")
    desc = S.terrain_c2(1920, 1080)
    px = stratified_pixels(1920, 1080, 64, 64, 0xC2)
    sc = R.scene_from_desc(ctx, desc)
    _, g1, g2 = R.render_moments(ctx, dict(scene=sc, spp=1024, depth=8, seed=0xC2C2))
    say("  gpu done; oracle")
    r1, r2 = c2_oracle(desc, px, 1024)
    g1, g2 = g1.reshape(-1, 3)[px], g2.reshape(-1, 3)[px]
    r1, r2 = r1[px], r2[px]
    z, mg, mr, bz, zi = zscores(g1, g2, 1024, r1, r2, 1024)
    say(f"  mean z {z.mean():+.4f}  |z|>4.5 {(np.abs(z) > 4.5).mean():.5f}  image z {zi:+.2f}  "
        f"means {mg.mean():.5f} vs {mr.mean():.5f}")
    # Caustics through the glass and mirror spheres make a few pixels heavy-tailed.
    check_unbiased(z, mg, mr, bz, zi, frac_tail=0.01, mean_tol=0.03)


# ---------------------------------------------------------------------------------- C3

def test_c3_sixteen_jobs_one_launch(ctx):
    """The 16 C3 fog jobs (uneven resolutions, spp, depths 4-12, 10k-300k triangles, homogeneous
    and exponential fog) share ONE rbr_render call and one ray pool.  Each job's film is bit
    for bit its single-job render, whose samples (rbr_render_moments, same seed, so the same
    random stream) are z-tested against the job's own oracle at equal spp (16x the job's spp,
    so the test has power).  The reference's homogeneous scatter weight sigma_s*Tr/avg^2
    (medium.cpp:22-27) makes fog radiance heavy-tailed, so the per-pixel statistics are pooled
    over all jobs and each job is judged by its image-mean z."""
    say("case: C3 16 jobs in one launch (scale 0.05)")
    jobs = S.fog_jobs_c3(scale=0.05)
    gjobs = []
    for j in jobs:
        sc = R.scene_from_desc(ctx, j["scene"])
        gjobs.append(dict(scene=sc, spp=16 * j["spp"], depth=j["depth"], seed=j["seed"], integrator=R.VOLPATH))
    films = R.render(ctx, gjobs)
    zs = []
    for k, (j, gj, film) in enumerate(zip(jobs, gjobs, films)):
        alone = R.render(ctx, [gj])[0]
        assert np.array_equal(alone, film), f"job {k}: multi-job film differs from its single-job render"
        n = gj["spp"]
        fm, g1, g2 = R.render_moments(ctx, gj)
        assert np.array_equal(fm, film), f"job {k}: moments render differs from the film"
        r1, r2 = O.OracleRender(j["scene"]).moments(n, gj["depth"], seed=500 + k, integrator=R.VOLPATH)
        z, mg, mr, bz, zi = zscores(g1, g2, n, r1, r2, n)
        assert not bz.any(), f"job {k}: a zero-variance pixel differs"
        w, h = j["scene"].camera.width, j["scene"].camera.height
        say(f"  job {k:2d} {w}x{h} spp {n} depth {gj['depth']} medium {j['scene'].medium['type']}: "
            f"mean z {z.mean():+.3f} |z|>4.5 {(np.abs(z) > 4.5).mean():.4f} image z {zi:+.2f}")
        assert abs(zi) < 5.0, k
        zs.append(z.reshape(-1, 3))
    z = np.concatenate(zs)
    say(f"  all jobs: {len(z)} pixels, mean z {z.mean():+.4f}, |z|>4.5 {(np.abs(z) > 4.5).mean():.4f}")
    assert abs(float(z.mean())) < 5.0 / np.sqrt(len(z))
    assert float((np.abs(z) > 4.5).mean()) < 0.01


# ---------------------------------------------------------------------------------- C4

def test_c4_guided_terrain_unbiased(ctx):
    """C4 on the C2 scene (480x270 film of the same camera): the guide is trained on records
    of plain paths, then the guided image (one-sample MIS, guide fraction 0.5) is z-tested
    against the plain GPU image at equal spp and against the oracle's pixel subset."""
    say("case: C4 guided terrain 480x270")
    desc = S.terrain_c2(480, 270)
    sc = R.scene_from_desc(ctx, desc)
    lo, hi = sc.world_bounds()
    g = R.Guide(ctx, lo, hi, seed=0xC4)
    g.reserve(1 << 22)
    losses = []
    for it in range(6):
        R.render(ctx, [dict(scene=sc, spp=2, depth=8, seed=900 + it, guided=2)])
        losses.append(g.train_collected(8, 1 << 16, 1e-3))
        g.clear_collected()
    say(f"  training loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    _, p1, p2 = R.render_moments(ctx, dict(scene=sc, spp=512, depth=8, seed=41))
    _, g1, g2 = R.render_moments(ctx, dict(scene=sc, spp=512, depth=8, seed=42, guided=1))
    z, mg, mp, bz, zi = zscores(g1, g2, 512, p1, p2, 512)
    say(f"  guided vs plain: mean z {z.mean():+.4f}  |z|>4.5 {(np.abs(z) > 4.5).mean():.5f}  image z {zi:+.2f}  "
        f"means {mg.mean():.5f} vs {mp.mean():.5f}")
    # The guided and plain estimators differ in skew, so the per-pixel mean z is not the bias
    # statistic here; the 15x15-pixel blocks and the image mean are.
    check_unbiased(z, mg, mp, bz, zi, frac_tail=0.01, mean_tol=0.03, zmean=False)
    check_blocks(g1, g2, 512, p1, p2, 512, 480, 270, 15)
    px = stratified_pixels(480, 270, 48, 48, 0xC4)
    r1, r2 = O.OracleRender(desc).moments(512, 8, seed=0xC4, pixels=px)
    z, mg, mr, bz, zi = zscores(g1.reshape(-1, 3)[px], g2.reshape(-1, 3)[px], 512, r1[px], r2[px], 512)
    say(f"  guided vs oracle subset: mean z {z.mean():+.4f}  |z|>4.5 {(np.abs(z) > 4.5).mean():.5f}  image z {zi:+.2f}")
    check_unbiased(z, mg, mr, bz, zi, frac_tail=0.01, mean_tol=0.03)


def relvar(s1, s2, n):
    """Mean over pixels of the luminance variance of one sample over (mean^2 + 1e-2)."""
    m = s1.reshape(-1, 3) / n
    v = np.maximum(s2.reshape(-1, 3) / n - m * m, 0.0)
    lum = np.array([0.2126, 0.7152, 0.0722])
    return float(np.mean((v @ lum) / ((m @ lum) ** 2 + 1e-2)))


def test_c4_variance_decomposition_bounds_guiding(ctx):
    """What guiding can gain on the C2 scene.  A guide changes only how continuation
    directions are drawn; the camera vertex and its NEE (a depth-1 render: emission seen
    directly, the area lights and the sun sampled at the first hit) are the same random
    variables with or without it.  The share of the per-sample variance they carry bounds
    the variance reduction of any guide at equal spp: relvar(full) / relvar(depth 1) is
    about 2.3-2.6 here (97.7 % of the image is that depth-1 light), so with the guide's inference and training on top, a 2x sooner
    time-to-relMSE is out of reach on this scene.  A trained guide (first bounce, the best
    measured schedule) is checked to stay above that floor."""
    say("case: C4 variance decomposition, terrain 480x270")
    sc = R.scene_from_desc(ctx, S.terrain_c2(480, 270))
    _, d1, d2 = R.render_moments(ctx, dict(scene=sc, spp=256, depth=1, seed=71))
    _, f1, f2 = R.render_moments(ctx, dict(scene=sc, spp=256, depth=8, seed=72))
    lo, hi = sc.world_bounds()
    g = R.Guide(ctx, lo, hi, seed=0xC4)
    g.reserve(1 << 22)
    for it in range(8):
        R.render(ctx, [dict(scene=sc, spp=2, depth=8, seed=900 + it, guided=2)])
        g.train_collected(16, 1 << 16, 1e-3)
        g.clear_collected()
    _, g1, g2 = R.render_moments(ctx, dict(scene=sc, spp=256, depth=8, seed=73, guided=1, guide_depth=1))
    vd, vf, vg = relvar(d1, d2, 256), relvar(f1, f2, 256), relvar(g1, g2, 256)
    mean_d, mean_f = float(d1.mean() / 256), float(f1.mean() / 256)
    say(f"  relvar: depth-1 {vd:.4g}, full {vf:.4g} (bound on the gain {vf / vd:.3f}), guided {vg:.4g}; "
        f"depth-1 share of the image mean {mean_d / mean_f:.3f}")
    assert 1.0 < vf / vd < 3.0
    assert vg > 0.9 * vd
