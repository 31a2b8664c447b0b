"""GPU guiding network (BF16 matrix-core kernels) vs the FP64 oracle (oracle/guide_mlp.py,
pinned by tests/test_oracle_guide.py).  Tolerances: the BF16-emulating oracle rounds at the
same points as the device, so only float32 summation order and rare one-ulp rounding flips
separate them; the exact FP64 network differs by BF16 quantisation (~2^-9 relative per
rounding point)."""
import sys

import numpy as np
import pytest

import guide_mlp as G
import rbr_amd as R
from test_oracle_guide import random_params, random_samples

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = R.Context(0)
    yield c
    c.close()


def test_param_layout_and_init_match_oracle(ctx):
    assert R.GUIDE_PARAMS == G.N_PARAMS
    say("guide init seed 5")
    g = R.Guide(ctx, [0, 0, 0], [1, 1, 1], seed=5)
    np.testing.assert_array_equal(g.params, G.init_params(5))


def fp16(a):
    return np.asarray(a, np.float32).astype(np.float16).astype(np.float64)


@pytest.mark.parametrize("n", [1, 15, 16, 17, 63, 64, 65, 1000, 4099])
def test_eval_matches_oracle(ctx, n):
    params = random_params(100 + n).astype(np.float32)
    samples, lo, hi = random_samples(200 + n, n)
    g = R.Guide(ctx, lo, hi, seed=1)
    g.params = params
    say(f"guide eval n={n}")
    out = g.eval(samples[:, :9]).astype(np.float64)
    # the record's weights carry the selection probability (they sum to it)
    mu, kappa, w = G.eval_record(params.astype(np.float64), samples[:, :9], lo, hi, quant=True)
    assert np.all(np.isfinite(out))
    # The record is FP16 (2^-11 relative).  Beyond that, most queries agree to float32
    # rounding; the hardware sine and angle doubling (error ~1e-6) sometimes put an encoded
    # feature on the other side of a BF16 rounding boundary, which moves that query's outputs
    # by ~1e-3.  A wrong operand layout or feature order moves every query by O(1).
    e_q = np.maximum.reduce([np.abs(out[:, :, :3] - fp16(mu)).max(axis=(1, 2)),
                             np.abs(np.log(out[:, :, 3]) - np.log(fp16(kappa))).max(axis=1),
                             np.abs(out[:, :, 4] - fp16(w)).max(axis=1)])
    flipped = int(np.sum(e_q > 2e-3))
    say(f"  per-query max error: median {np.median(e_q):.2e} max {e_q.max():.2e}; {flipped}/{n} above 2e-3")
    assert np.median(e_q) < 1e-3
    assert flipped <= max(1, int(0.15 * n))
    # a flipped query's unit mean moves further when its raw mean vector is short
    assert e_q.max() < 0.2
    # against the unquantised network: BF16 error budget
    mu_x, kappa_x, w_x = G.eval_record(params.astype(np.float64), samples[:, :9], lo, hi, quant=False)
    assert np.median(np.abs(out[:, :, :3] - mu_x)) < 2e-2
    assert np.median(np.abs(out[:, :, 4] - w_x)) < 5e-3
    # the mixture density, the quantity the sampler uses
    rng = np.random.default_rng(n)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pg = G.mixture_pdf(out[:, :, :3], out[:, :, 3], out[:, :, 4], d)[:, 0]
    po = G.mixture_pdf(mu, kappa, w, d)[:, 0]
    assert np.quantile(np.abs(pg / po - 1.0), 0.99) < 2e-2


def sphere_bins(dirs, nz=32, nphi=64):
    """Equal-area bins: uniform in z and in azimuth."""
    z = np.clip(dirs[..., 2], -1.0, 1.0 - 1e-12)
    phi = np.mod(np.arctan2(dirs[..., 1], dirs[..., 0]), 2 * np.pi)
    iz = np.minimum(((z + 1.0) * 0.5 * nz).astype(int), nz - 1)
    ip = np.minimum((phi / (2 * np.pi) * nphi).astype(int), nphi - 1)
    return iz * nphi + ip


def test_sampler_matches_its_density(ctx):
    """Directions drawn by the renderer's guide sampler follow the mixture density it reports
    (chi-square over 2048 equal-area bins, expected counts by quadrature of the oracle's
    mixture density at the record's FP16 lobes), and the reported density equals that mixture
    at the drawn direction.  A sampler/density mismatch would bias guided rendering."""
    nq, m = 3, 1 << 20
    # concentrations kept to kappa <~ 100 so that 8x8 sub-cell quadrature resolves every lobe
    p64 = random_params(77, scale=2.0)
    Ws, bs = G.unpack(p64)
    prng = np.random.default_rng(76)
    for k in range(8):
        Ws[4][5 * k + 3] *= 0.2
        bs[4][5 * k + 3] = prng.uniform(-1.0, 3.5)
    params = G.pack(Ws, bs).astype(np.float32)
    samples, lo, hi = random_samples(78, nq)
    g = R.Guide(ctx, lo, hi, seed=1)
    g.params = params
    lobes = g.eval(samples[:, :9]).astype(np.float64)
    say(f"guide sampler: kappa range {lobes[:, :, 3].min():.3g}..{lobes[:, :, 3].max():.3g}")
    rng = np.random.default_rng(79)
    u = rng.random((nq, m, 3), dtype=np.float32)
    dirs, pdfs = g.sample(samples[:, :9], u)
    assert np.all(np.isfinite(dirs)) and np.all(pdfs > 0)
    np.testing.assert_allclose(np.linalg.norm(dirs, axis=2), 1.0, atol=1e-5)
    mu = lobes[:, :, :3] / np.linalg.norm(lobes[:, :, :3], axis=2, keepdims=True)
    kap, w = lobes[:, :, 3], lobes[:, :, 4] / lobes[:, :, 4].sum(axis=1, keepdims=True)
    pref = G.mixture_pdf(mu, kap, w, dirs.astype(np.float64))
    assert np.quantile(np.abs(pdfs / pref - 1.0), 0.999) < 1e-3
    # bin probabilities by 8x8 sub-cell quadrature
    nz, nphi, sub = 32, 64, 8
    zc = -1.0 + (np.arange(nz * sub) + 0.5) * 2.0 / (nz * sub)
    pc = (np.arange(nphi * sub) + 0.5) * 2 * np.pi / (nphi * sub)
    Z, P = np.meshgrid(zc, pc, indexing="ij")
    rxy = np.sqrt(1 - Z * Z)
    grid = np.stack([rxy * np.cos(P), rxy * np.sin(P), Z], -1).reshape(-1, 3)
    cell = 4 * np.pi / grid.shape[0]
    for qi in range(nq):
        dens = G.mixture_pdf(mu[qi:qi + 1], kap[qi:qi + 1], w[qi:qi + 1], grid[None])[0]
        prob = (dens * cell).reshape(nz, sub, nphi, sub).sum(axis=(1, 3)).reshape(-1)
        prob /= prob.sum()
        counts = np.bincount(sphere_bins(dirs[qi].astype(np.float64)), minlength=nz * nphi)
        expct = prob * m
        ok = expct >= 20
        chi2 = float(np.sum((counts[ok] - expct[ok]) ** 2 / expct[ok]))
        dof = int(ok.sum()) - 1
        say(f"  query {qi}: chi2/dof {chi2 / dof:.3f} over {dof} bins, mass outside {1 - prob[ok].sum():.2e}")
        assert chi2 / dof < 1.0 + 6.0 * np.sqrt(2.0 / dof)


def test_eval_empty_and_errors(ctx):
    g = R.Guide(ctx, [0, 0, 0], [1, 1, 1], seed=2)
    assert g.eval(np.zeros((0, 9), np.float32)).shape == (0, 8, 5)
    # a parameter vector of the wrong size is rejected
    with pytest.raises(R.RbrError):
        ctx.check(R._lib.rbr_guide_set_params(ctx.handle, np.zeros(3, np.float32).ctypes.data, 3))


@pytest.mark.parametrize("n", [16, 333, 5000])
def test_loss_grad_matches_fp64_oracle(ctx, n):
    params = random_params(300 + n).astype(np.float32)
    samples, lo, hi = random_samples(400 + n, n)
    g = R.Guide(ctx, lo, hi, seed=1)
    g.params = params
    say(f"guide loss_grad n={n}")
    loss, grad = g.loss_grad(samples)
    p64 = params.astype(np.float64)
    lq, gq = G.loss_grad(p64, samples, lo, hi, quant=True)
    lx, gx = G.loss_grad(p64, samples, lo, hi, quant=False)
    assert np.all(np.isfinite(grad))
    assert abs(loss - lq) <= 2e-3 * abs(lq) + 1e-6
    assert abs(loss - lx) <= 2e-2 * abs(lx) + 1e-5

    def cos(a, b):
        return float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))

    rel_q = np.linalg.norm(grad - gq) / np.linalg.norm(gq)
    rel_x = np.linalg.norm(grad - gx) / np.linalg.norm(gx)
    say(f"  loss {loss:.6f} oracle {lx:.6f}; grad cos(q) {cos(grad, gq):.6f} rel(q) {rel_q:.2e} "
        f"cos(x) {cos(grad, gx):.6f} rel(x) {rel_x:.2e}")
    assert cos(grad, gq) > 0.9999 and rel_q < 1e-2
    # against exact FP64: no further from it than BF16 rounding itself puts the gradient
    # (0.12 relative at 16 samples, 0.05 at 333, 0.01 at 5000 -- averaging over samples)
    rel_qx = np.linalg.norm(gq - gx) / np.linalg.norm(gx)
    assert rel_x <= rel_qx + 1e-3 and cos(grad, gx) > 0.99
    # per-layer agreement (a transposed or misplaced block fails here even if the total is close)
    offs = np.cumsum([0] + [a * b for a, b in G.W_SHAPES])
    for l in range(5):
        a, b = grad[offs[l]:offs[l + 1]], gq[offs[l]:offs[l + 1]]
        assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(b) + 1e-7, l
    np.testing.assert_allclose(grad[G.N_WEIGHTS:], gq[G.N_WEIGHTS:], atol=2e-2 * np.abs(gq[G.N_WEIGHTS:]).max())


def test_training_fits_a_known_distribution(ctx):
    """Directions drawn from a position-dependent vMF: Adam steps must drive the loss towards
    the target's entropy (the minimum of E[-log p])."""
    rng = np.random.default_rng(7)
    lo, hi = np.zeros(3), np.ones(3)
    kap_t = 20.0

    def batch(n):
        pos = rng.random((n, 3))
        ang = 2 * np.pi * pos[:, 0]
        mu = np.stack([np.cos(ang), np.sin(ang), np.zeros(n)], axis=1)
        # sample vMF(mu, kappa) (Wood's method for the sphere)
        u = rng.random(n)
        wz = 1.0 + np.log(u + (1.0 - u) * np.exp(-2.0 * kap_t)) / kap_t
        phi = 2 * np.pi * rng.random(n)
        rxy = np.sqrt(np.maximum(0.0, 1.0 - wz * wz))
        # frame around mu
        t1 = np.stack([-np.sin(ang), np.cos(ang), np.zeros(n)], axis=1)
        t2 = np.cross(mu, t1)
        d = wz[:, None] * mu + rxy[:, None] * (np.cos(phi)[:, None] * t1 + np.sin(phi)[:, None] * t2)
        nrm = np.tile([0.0, 0.0, 1.0], (n, 1))
        return np.concatenate([pos, nrm, nrm, d, np.ones((n, 1))], axis=1)

    # E[-log f] = -log C(kappa) - kappa (E[mu.w] - 1), E[mu.w] = coth(kappa) - 1/kappa
    e_cos = 1.0 / np.tanh(kap_t) - 1.0 / kap_t
    entropy = -(np.log(kap_t) - np.log(2 * np.pi) - np.log1p(-np.exp(-2 * kap_t)) + kap_t * (e_cos - 1.0))
    g = R.Guide(ctx, lo, hi, seed=3)
    test = batch(4096)
    l0, _ = g.loss_grad(test)
    say(f"guide training: initial loss {l0:.4f}, target entropy {entropy:.4f}")
    for it in range(600):
        g.train_step(batch(4096), lr=3e-3)
    l1, _ = g.loss_grad(test)
    say(f"  after 600 steps: {l1:.4f}")
    assert l1 < l0 - 0.5 * (l0 - entropy)
    assert l1 < entropy + 0.5


def test_save_load_resumes_training(ctx, tmp_path):
    """rbr_guide_save / rbr_guide_load: a loaded guide has the saved parameters and scene box
    (identical lobes) and, with the Adam moments and step restored, the same next training steps
    as the guide that kept going; a file of another kind is refused with RBR_ERR_IO."""
    samples, lo, hi = random_samples(91, 4096)
    g = R.Guide(ctx, lo, hi, seed=3)
    for _ in range(3):
        g.train_step(samples, lr=1e-3)
    path = str(tmp_path / "guide.bin")
    g.save(path)
    q = samples[:64, :9]
    before = g.eval(q)
    with R.Context(0) as c2:
        h = R.Guide.load(c2, path)
        np.testing.assert_array_equal(h.params, g.params)
        np.testing.assert_array_equal(h.eval(q), before)
        la = [g.train_step(samples, lr=1e-3) for _ in range(2)]
        lb = [h.train_step(samples, lr=1e-3) for _ in range(2)]
        np.testing.assert_allclose(la, lb, rtol=1e-12)   # losses are summed with atomics
        np.testing.assert_array_equal(h.params, g.params)
        bad = tmp_path / "bad.bin"
        bad.write_bytes(b"not a guide file" * 8)
        with pytest.raises(R.RbrError) as e:
            R.Guide.load(c2, str(bad))
        assert e.value.code == R.ERR_IO
    # a guide saved before any training step loads without Adam state
    g2 = R.Guide(ctx, lo, hi, seed=4)
    g2.save(path)
    with R.Context(0) as c3:
        np.testing.assert_array_equal(R.Guide.load(c3, path).params, g2.params)
