"""Pins the FP64 guiding-network oracle (oracle/guide_mlp.py) by checks that do not retype it:
numpy's sin/cos evaluated directly, torch float64 layers and autograd, central finite
differences, and quadrature of the vMF mixture over the sphere."""
import numpy as np
import pytest
import torch

import guide_mlp as G


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = G.init_params(seed).astype(np.float64)
    # non-zero biases and a spread of log kappa / logits so every term of the loss matters
    Ws, bs = G.unpack(p)
    for b in bs:
        b[:] = rng.normal(0.0, 0.1 * scale, b.shape)
    Ws[4][:] = rng.normal(0.0, 0.3 * scale, Ws[4].shape)
    Ws[4][41:] = 0.0   # row 40 is the selection logit
    Ws[0][:, G.padding_features()] = 0.0
    return G.pack(Ws, bs)


def random_samples(seed, n, lo=(-2.0, -1.0, 0.0), hi=(3.0, 2.0, 5.0)):
    rng = np.random.default_rng(seed)
    lo, hi = np.asarray(lo), np.asarray(hi)
    pos = lo + (hi - lo) * rng.random((n, 3))
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    wo = rng.normal(size=(n, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c = rng.exponential(1.0, n)
    c[rng.random(n) < 0.1] = 0.0
    # BSDF density of the direction; a quarter of the samples carry none (the guide alone)
    pb = rng.exponential(0.3, n)
    pb[rng.random(n) < 0.25] = -1.0
    # the device takes float32 samples: keep values representable
    samples = np.concatenate([pos, nrm, wo, d, c[:, None], pb[:, None]], axis=1).astype(np.float32).astype(np.float64)
    return samples, lo.astype(np.float32), hi.astype(np.float32)


def torch_loss(params, samples, lo, hi):
    """Independent float64 restatement with torch layers; autograd gives the gradient."""
    p = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    s9, dirs, c, pb = G.split_samples(samples, lo, hi)
    x = torch.tensor(s9)
    feats = []
    for jk in G.feature_pairs():
        if jk is None:
            feats += [torch.zeros(len(x), dtype=torch.float64)] * 2
        else:
            ang = (2.0 ** jk[1]) * np.pi * x[:, jk[0]]
            feats += [torch.sin(ang), torch.cos(ang)]
    X = torch.stack(feats, dim=1)
    o = 0
    Ws = []
    for (r, cc) in G.W_SHAPES:
        Ws.append(p[o:o + r * cc].view(r, cc))
        o += r * cc
    bs = []
    for nb in G.B_SIZES:
        bs.append(p[o:o + nb])
        o += nb
    h = X
    for l in range(5):
        if l == 4:
            h4 = h
        h = torch.nn.functional.linear(h, Ws[l], bs[l])
        if l < 4:
            h = torch.relu(h)
    out = h[:, :40].reshape(-1, 8, 5)
    mu = torch.nn.functional.normalize(out[:, :, :3], dim=2, eps=1e-6)
    kap = torch.exp(torch.clamp(out[:, :, 3], G.LOG_KAPPA_MIN, G.LOG_KAPPA_MAX))
    logw = torch.log_softmax(out[:, :, 4], dim=1)
    d = torch.tensor(dirs)
    cos = (mu * d[:, None, :]).sum(dim=2)
    lognorm = torch.log(kap) - np.log(2 * np.pi) - torch.log(1.0 - torch.exp(-2.0 * kap))
    logp = torch.logsumexp(logw + lognorm + kap * (cos - 1.0), dim=1)
    # one-sample MIS with the BSDF: q = s p + (1 - s) pb where the sample carries pb > 0; the
    # lobes learn -log p, the selection logit -log q with p held constant
    # the selection row reads the hidden features but sends them no gradient
    sel = G.SEL_MIN + (G.SEL_MAX - G.SEL_MIN) * torch.sigmoid(h4.detach() @ Ws[4][40] + bs[4][40])
    pbt = torch.tensor(pb)
    ct = torch.tensor(c)
    pbs = torch.where(pbt > 0, pbt, torch.ones_like(pbt))
    second = ct * ct * pbs / (sel * torch.exp(logp.detach()) + (1.0 - sel) * pbs)
    surrogate = (-(ct * logp) + torch.where(pbt > 0, second, torch.zeros_like(second))).sum() / len(c)
    surrogate.backward()
    with torch.no_grad():
        logq = torch.where(pbt > 0, torch.log(sel * torch.exp(logp) + (1.0 - sel) * torch.clamp(pbt, min=1e-300)), logp)
        loss = -(ct * logq).sum() / len(c)
    return loss.item(), p.grad.numpy()


def test_feature_order_covers_every_scalar_and_frequency_once():
    pairs = G.feature_pairs()
    assert len(pairs) == 112
    real = [jk for jk in pairs if jk is not None]
    assert sorted(real) == [(j, k) for j in range(9) for k in range(12)]
    # lane group g of a k-block holds pairs p with (p % 8) // 2 == g: its frequencies are 3g..3g+2
    for p, jk in enumerate(pairs):
        if jk is not None:
            assert jk[1] // 3 == (p % 8) // 2
    assert len(G.padding_features()) == 8


def test_encoding_matches_direct_sin_cos():
    rng = np.random.default_rng(3)
    s9 = rng.uniform(-1.5, 1.5, (64, 9))
    X = G.encode(s9)
    for p, jk in enumerate(G.feature_pairs()):
        if jk is None:
            assert np.all(X[:, 2 * p:2 * p + 2] == 0.0)
            continue
        j, k = jk
        np.testing.assert_allclose(X[:, 2 * p], np.sin(2.0 ** k * np.pi * s9[:, j]), atol=1e-9)
        np.testing.assert_allclose(X[:, 2 * p + 1], np.cos(2.0 ** k * np.pi * s9[:, j]), atol=1e-9)


def test_forward_and_gradient_match_torch_autograd():
    params = random_params(11)
    samples, lo, hi = random_samples(12, 96)
    l_or, g_or = G.loss_grad(params, samples, lo, hi)
    l_t, g_t = torch_loss(params, samples, lo, hi)
    assert abs(l_or - l_t) <= 1e-10 * max(1.0, abs(l_t))
    # padding parameters get zero gradient in both
    np.testing.assert_allclose(g_or, g_t, rtol=1e-8, atol=1e-11)


def test_gradient_matches_finite_differences():
    params = random_params(21)
    samples, lo, hi = random_samples(22, 24)
    _, g = G.loss_grad(params, samples, lo, hi)
    rng = np.random.default_rng(23)
    offs = np.cumsum([0] + [a * b for a, b in G.W_SHAPES])
    # a few entries of every weight matrix and bias vector, largest-gradient ones included
    idx = list(np.argsort(-np.abs(g))[:12])
    for l, (r, c) in enumerate(G.W_SHAPES):
        real = np.setdiff1d(np.arange(c), G.padding_features()) if l == 0 else np.arange(c)
        for _ in range(4):
            idx.append(offs[l] + rng.integers(r if l < 4 else 41) * c + rng.choice(real))
    # the selection logit's row and bias
    idx += [offs[4] + 40 * 64 + int(rng.integers(64)) for _ in range(3)]
    b0 = G.N_WEIGHTS
    for nb in G.B_SIZES:
        idx.append(b0 + rng.integers(min(nb, 40)))
        b0 += nb
    idx.append(G.N_WEIGHTS + 4 * 64 + 40)
    h = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd = (G.surrogate(p1, samples, lo, hi, params) - G.surrogate(p2, samples, lo, hi, params)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_selection_term_off_without_bsdf_density():
    """Samples without a BSDF density (column 13 <= 0, or 13 columns) train the guide alone:
    the loss is -c log p and the selection logit gets no gradient.  With a density, the lobes'
    gradient is unchanged (they learn -c log p either way) and only the selection row moves."""
    params = random_params(25)
    samples, lo, hi = random_samples(26, 64)
    none = samples.copy()
    none[:, 13] = -1.0
    l13, g13 = G.loss_grad(params, samples[:, :13], lo, hi)
    l14, g14 = G.loss_grad(params, none, lo, hi)
    assert l13 == l14 and np.array_equal(g13, g14)
    offs = np.cumsum([0] + [a * b for a, b in G.W_SHAPES])
    assert not g14[offs[4] + 40 * 64:offs[4] + 41 * 64].any() and g14[G.N_WEIGHTS + 4 * 64 + 40] == 0.0
    s9, dirs, c, _ = G.split_samples(samples, lo, hi)
    _, O = G.forward(params, G.encode(s9))
    mu, kappa, w = G.decode(O)
    p = G.mixture_pdf(mu, kappa, w, dirs)[:, 0]
    assert abs(l13 - float(np.sum(-c * np.log(p)) / len(c))) <= 1e-12 * abs(l13)
    # with the densities the gradient differs from g14 in the selection row and bias only: the
    # lobes learn -c log p either way and the hidden layers get nothing from the selection term
    _, gs = G.loss_grad(params, samples, lo, hi)
    sel_idx = set(range(offs[4] + 40 * 64, offs[4] + 41 * 64)) | {G.N_WEIGHTS + 4 * 64 + 40}
    assert set(np.nonzero(gs != g14)[0].tolist()) <= sel_idx
    assert gs[offs[4] + 40 * 64:offs[4] + 41 * 64].any()


def test_selection_optimum_matches_closed_form():
    """For one query (all samples share the network outputs), the selection row minimises
    sum_i c_i^2 pb_i / (s p_i + (1 - s) pb_i), the second moment of the one-sample MIS
    estimator for directions drawn from the BSDF; its derivative in s vanishes at the optimum:
    the oracle's gradient w.r.t. the selection bias is zero at the s that a 1-D root finder puts
    there (and has the sign of the derivative on either side)."""
    from scipy.optimize import brentq
    params = random_params(27)
    samples, lo, hi = random_samples(28, 40)
    samples[:, :9] = samples[0, :9]          # one query: same network outputs for all
    s9, dirs, c, _ = G.split_samples(samples, lo, hi)
    _, O = G.forward(params, G.encode(s9))
    mu, kappa, w = G.decode(O)
    p = G.mixture_pdf(mu, kappa, w, dirs)[:, 0]
    # half the samples favour the guide, half the BSDF, so the optimum is inside (0, 1)
    samples[:, 13] = (p * np.where(np.arange(len(p)) % 2 == 0, 0.01, 1.5)).astype(np.float32)
    s9, dirs, c, pb = G.split_samples(samples, lo, hi)
    dL = lambda s: float(np.sum(-c * c * pb * (p - pb) / (s * p + (1 - s) * pb) ** 2))
    lo_s, hi_s = G.SEL_MIN + 1e-9, G.SEL_MAX - 1e-9
    assert dL(lo_s) < 0 < dL(hi_s)
    s_star = brentq(dL, lo_s, hi_s, xtol=1e-14)
    Ws, bs = G.unpack(params)
    z_now = O[0, 40]
    ib = G.N_WEIGHTS + 4 * 64 + 40
    logit = lambda s: np.log((s - G.SEL_MIN) / (G.SEL_MAX - s))
    for s_t, sign in ((s_star, 0), (0.5 * (lo_s + s_star), -1), (0.5 * (s_star + hi_s), 1)):
        bs2 = [b.copy() for b in bs]
        bs2[4][40] += logit(s_t) - z_now
        _, g = G.loss_grad(G.pack(Ws, bs2), samples, lo, hi)
        if sign == 0:
            assert abs(g[ib]) < 1e-9
        else:
            assert np.sign(g[ib]) == sign


def fibonacci_sphere(m):
    k = np.arange(m) + 0.5
    z = 1.0 - 2.0 * k / m
    phi = np.pi * (1.0 + 5 ** 0.5) * k
    rxy = np.sqrt(1.0 - z * z)
    return np.stack([rxy * np.cos(phi), rxy * np.sin(phi), z], axis=1)


def test_mixture_density_integrates_to_one():
    params = random_params(31, scale=2.0)
    samples, lo, hi = random_samples(32, 6)
    mu, kappa, w = G.eval_lobes(params, samples[:, :9], lo, hi)
    assert np.allclose(w.sum(axis=1), 1.0)
    assert np.allclose(np.linalg.norm(mu, axis=2), 1.0)
    dirs = fibonacci_sphere(400_000)
    pdf = G.mixture_pdf(mu, np.minimum(kappa, 200.0), w, np.broadcast_to(dirs, (6,) + dirs.shape))
    integral = pdf.mean(axis=1) * 4.0 * np.pi
    np.testing.assert_allclose(integral, 1.0, rtol=2e-3)


def test_single_lobe_normaliser_closed_form():
    # kappa/(2 pi (1 - e^{-2 kappa})) == kappa / (4 pi sinh kappa) * e^{kappa}
    for kap in [np.exp(-5.0), 0.01, 0.5, 1.0, 7.0, 100.0, np.exp(8.0)]:
        want = np.log(kap) - np.log(4 * np.pi) - np.log(np.sinh(kap)) + kap if kap < 500 else None
        got = G.vmf_log_norm(np.array(kap))
        if want is not None:
            assert abs(got - want) < 1e-12 * max(1.0, abs(want))


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e38, 1e-40], np.float32)
    got = G.bf16(x)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5], np.float64)
    np.testing.assert_array_equal(got[:5], want)
    assert np.isfinite(got[5])
    # bfloat16 keeps float32's exponent range: tiny values keep their order of magnitude
    assert 0.0 < got[6] < 2e-40


def test_quantised_gradient_close_to_exact():
    params = random_params(41)
    samples, lo, hi = random_samples(42, 512)
    _, g = G.loss_grad(params, samples, lo, hi)
    _, gq = G.loss_grad(params, samples, lo, hi, quant=True)
    cos = g @ gq / (np.linalg.norm(g) * np.linalg.norm(gq))
    assert cos > 0.995
