"""FP64 restatement of the path-guiding network -- test infrastructure only.

Only tests/, the smoke check and the benchmark's CPU-baseline leg may import this module; it
is the checker of the GPU guiding kernels, never a fallback for them.

The network has no counterpart in the reference (its integrators sample the BSDF only,
path_tracing_integrator.cpp:40-60), so there is nothing of the reference to follow here.  What
pins this restatement instead (tests/test_oracle_guide.py):
  * the encoding against numpy's sin/cos of 2^k * pi * x evaluated directly,
  * the forward pass against torch.nn.functional.linear/relu in float64,
  * the hand-written backward against torch autograd in float64 and against central finite
    differences of the loss,
  * the vMF mixture density against quadrature over the sphere (it integrates to 1) and the
    closed-form single-lobe normaliser.

Network (all shapes as the device stores them):
  input  : 9 scalars (position scaled to [0,1]^3 by the box lo..hi, outward normal, wo)
           108 (sin, cos) pairs sin/cos(2^k pi s_j), j < 9, k < 12, in the order of
           feature_pairs(), plus 4 zero pairs: 224 features
  layers : W0[64][224] W1..W3[64][64] W4[48][64], biases b0..b3[64], b4[48]; ReLU after
           layers 0..3; the padding features and outputs 40..47 carry no information
  output : 8 lobes x (mean 3, log kappa, logit) in outputs 0..39; mean normalised, log kappa
           clamped to [-5, 8] (zero gradient outside), weights = softmax(logits); output 40
           sets the selection probability s = 0.1 + 0.8 sigmoid(o_40) with which the
           renderer draws a direction from the guide rather than the BSDF (one-sample MIS)
  density: p(w) = sum_k w_k kappa_k / (2 pi (1 - e^{-2 kappa_k})) e^{kappa_k (mu_k . w - 1)}
  record : what inference hands the renderer: mu, kappa and the weights times s, so the
           stored weights sum to s
  loss   : (1/n) sum_i -c_i log q_i(d_i) with q = s p + (1 - s) p_bsdf when the sample carries
           the BSDF density p_bsdf > 0 of its direction, q = p when it does not
  gradient: of the surrogate (1/n) sum_i [-c_i log p_i(d_i) + c_i^2 g_i / q^_i] with
           q^ = s p^ + (1 - s) p_bsdf, p^ = p held constant and g the density the direction was
           drawn with (the BSDF's for these samples): the lobes learn the guide's own
           cross-entropy (as without a selection term, so a small s does not starve them of
           gradient), and the selection row alone learns to lower the second moment of the
           one-sample MIS estimator, int f^2 / q, estimated per sample as c^2 g / q (c = f / g
           for a direction drawn with density g); its derivative is
           d/d o_40 = -c^2 g (p - p_bsdf) / q^2 * ds/d o_40; the backward pass leaves row 40 of
           W4 out of the hidden layers' gradient

quant=True reproduces the device's rounding points: encoded inputs, weights, hidden
activations and back-propagated activation gradients are rounded to bfloat16 (round to
nearest even); products are exact and sums are carried in float64 here (float32 on the
device), and the outputs stay unrounded.
"""
from __future__ import annotations

import numpy as np

N_IN, N_IN_PAD, N_HID, N_OUT, N_OUT_PAD, LOBES, FREQ = 216, 224, 64, 40, 48, 8, 12
SEL = 40   # output index of the selection logit
SEL_MIN, SEL_MAX = 0.1, 0.9   # s = SEL_MIN + (SEL_MAX - SEL_MIN) sigmoid(o_40)
W_SHAPES = [(64, 224), (64, 64), (64, 64), (64, 64), (48, 64)]
B_SIZES = [64, 64, 64, 64, 48]
N_WEIGHTS = sum(a * b for a, b in W_SHAPES)
N_PARAMS = N_WEIGHTS + sum(B_SIZES)
LOG_KAPPA_MIN, LOG_KAPPA_MAX = -5.0, 8.0
LOG_2PI = np.log(2.0 * np.pi)


def unpack(params: np.ndarray) -> tuple[list[np.ndarray], list[np.ndarray]]:
    p = np.asarray(params, np.float64)
    assert p.shape == (N_PARAMS,)
    Ws, bs, o = [], [], 0
    for (r, c) in W_SHAPES:
        Ws.append(p[o:o + r * c].reshape(r, c))
        o += r * c
    for n in B_SIZES:
        bs.append(p[o:o + n])
        o += n
    return Ws, bs


def pack(Ws: list[np.ndarray], bs: list[np.ndarray]) -> np.ndarray:
    return np.concatenate([w.reshape(-1) for w in Ws] + [b.reshape(-1) for b in bs])


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even) via the float32 bit pattern; returns float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def scale_queries(q9: np.ndarray, lo, hi) -> np.ndarray:
    """(n, 9) raw queries (position, normal, wo) -> network scalars.

    The scaling is specified in float32, as the device computes it: s = (p - lo) * inv with
    inv = 1 / (hi - lo) rounded to float32 and each operation rounded.  The encoding's
    highest frequency (2^11 pi) turns a float32-vs-float64 difference of the position into a
    visible phase change, so the oracle keeps float32 here."""
    q = np.asarray(q9, np.float32).reshape(-1, 9).copy()
    lo = np.asarray(lo, np.float32)
    ext = np.asarray(hi, np.float32) - lo
    inv = np.where(ext > 0, np.float32(1.0) / np.where(ext > 0, ext, np.float32(1.0)), np.float32(1.0)).astype(np.float32)
    q[:, :3] = (q[:, :3] - lo) * inv
    return q.astype(np.float64)


def feature_pairs() -> list:
    """Scalar j and frequency k of each of the 112 (sin, cos) feature pairs, None for the 4
    zero padding pairs.  The order is the device's operand layout (guide.h guidePair): pair p
    sits in k-block p // 16 and slot group g = (p % 8) // 2, and lane group g owns the local
    pairs q = 4 (p // 16) + 2 ((p % 16) // 8) + p % 2 = scalar q // 3 at frequency 3g + q % 3."""
    out = []
    for p in range(N_IN_PAD // 2):
        kb, rem = divmod(p, 16)
        half, g, t = rem // 8, (rem % 8) // 2, rem % 2
        q = 4 * kb + 2 * half + t
        out.append(None if q >= 27 else (q // 3, 3 * g + q % 3))
    return out


def encode(s9: np.ndarray) -> np.ndarray:
    """(n, 9) scalars -> (n, 224) features: for pair p = (j, k), x[2p] = sin(2^k pi s_j),
    x[2p+1] = cos(2^k pi s_j), evaluated as sin/cos(2 pi frac(2^(k-1) s_j))."""
    s = np.asarray(s9, np.float64)
    n = s.shape[0]
    X = np.zeros((n, N_IN_PAD))
    for p, jk in enumerate(feature_pairs()):
        if jk is None:
            continue
        j, k = jk
        r = s[:, j] * 2.0 ** (k - 1)
        f = r - np.floor(r)
        X[:, 2 * p] = np.sin(2.0 * np.pi * f)
        X[:, 2 * p + 1] = np.cos(2.0 * np.pi * f)
    return X


def padding_features() -> np.ndarray:
    """Indices of the 8 zero input features."""
    return np.array([2 * p + t for p, jk in enumerate(feature_pairs()) if jk is None for t in (0, 1)])


def forward(params: np.ndarray, X: np.ndarray, quant: bool = False):
    """Returns (A, O): A[l] is the input of layer l (A[0] = X), O the (n, 48) outputs."""
    Ws, bs = unpack(params)
    q = bf16 if quant else (lambda v: v)
    A = [q(X)]
    h = A[0]
    for l in range(5):
        z = h @ q(Ws[l]).T + bs[l]
        if l < 4:
            h = q(np.maximum(z, 0.0))
            A.append(h)
        else:
            O = z
    return A, O


def decode(O: np.ndarray):
    """(n, 48) outputs -> mu (n, 8, 3), kappa (n, 8), weight (n, 8)."""
    o = np.asarray(O, np.float64)[:, :N_OUT].reshape(-1, LOBES, 5)
    m = o[:, :, :3]
    ln = np.maximum(np.linalg.norm(m, axis=2, keepdims=True), 1e-6)
    mu = m / ln
    kappa = np.exp(np.clip(o[:, :, 3], LOG_KAPPA_MIN, LOG_KAPPA_MAX))
    z = o[:, :, 4]
    z = z - z.max(axis=1, keepdims=True)
    w = np.exp(z) / np.exp(z).sum(axis=1, keepdims=True)
    return mu, kappa, w


def selection(O: np.ndarray) -> np.ndarray:
    """(n, 48) outputs -> selection probability s = SEL_MIN + (SEL_MAX - SEL_MIN) sigmoid(o_40)."""
    return SEL_MIN + (SEL_MAX - SEL_MIN) / (1.0 + np.exp(-np.asarray(O, np.float64)[:, SEL]))


def decode_record(O: np.ndarray):
    """What inference stores: mu, kappa and the weights times the selection probability."""
    mu, kappa, w = decode(O)
    return mu, kappa, w * selection(O)[:, None]


def log1mexp(x):
    """log(1 - e^{-x}), x > 0."""
    x = np.asarray(x, np.float64)
    return np.where(x < 0.693, np.log(-np.expm1(-x)), np.log1p(-np.exp(-x)))


def vmf_log_norm(kappa):
    """log of kappa / (2 pi (1 - e^{-2 kappa}))."""
    return np.log(kappa) - LOG_2PI - log1mexp(2.0 * kappa)


def mixture_pdf(mu, kappa, w, dirs) -> np.ndarray:
    """Density of unit directions dirs (n, 3) [or (n, m, 3)] under each query's mixture."""
    d = np.asarray(dirs, np.float64)
    if d.ndim == 2:
        d = d[:, None, :]
    cos = np.einsum("nkc,nmc->nmk", mu, d)
    lp = np.log(w)[:, None, :] + vmf_log_norm(kappa)[:, None, :] + kappa[:, None, :] * (cos - 1.0)
    return np.exp(lp).sum(axis=2)


def _loss_and_dO(O: np.ndarray, dirs: np.ndarray, c: np.ndarray, pb: np.ndarray | None = None, n_mean: int = 1):
    """Per-sample loss -c log q and its gradient w.r.t. the 48 outputs: q = s p + (1 - s) pb
    where pb > 0, else q = p (module docstring)."""
    n = O.shape[0]
    o = O[:, :N_OUT].reshape(n, LOBES, 5)
    m = o[:, :, :3]
    ln = np.maximum(np.linalg.norm(m, axis=2, keepdims=True), 1e-6)
    mu = m / ln
    lraw = o[:, :, 3]
    l = np.clip(lraw, LOG_KAPPA_MIN, LOG_KAPPA_MAX)
    kap = np.exp(l)
    z = o[:, :, 4]
    zmax = z.max(axis=1, keepdims=True)
    lse = zmax + np.log(np.exp(z - zmax).sum(axis=1, keepdims=True))
    wlog = z - lse
    dcos = np.einsum("nkc,nc->nk", mu, dirs)
    a = wlog + l - LOG_2PI - log1mexp(2.0 * kap) + kap * (dcos - 1.0)
    amax = a.max(axis=1, keepdims=True)
    logp = (amax + np.log(np.exp(a - amax).sum(axis=1, keepdims=True)))[:, 0]
    r = np.exp(a - logp[:, None])                       # responsibilities
    # Selection (module docstring): the guide terms below are those of -log p; the selection
    # logit gets d(-log q)/d o_40 = -(p - pb) / q * s (1 - s) with p held constant.
    if pb is None:
        pb = np.full(n, -1.0)
    pb = np.asarray(pb, np.float64)
    sel = pb > 0.0
    sg = 1.0 / (1.0 + np.exp(-O[:, SEL]))
    s_ = SEL_MIN + (SEL_MAX - SEL_MIN) * sg
    log_pb = np.log(np.where(sel, pb, 1.0))
    logq = np.where(sel, np.logaddexp(np.log(s_) + logp, np.log(1.0 - s_) + log_pb), logp)
    # second-moment gradient: c here is c_i / n_mean, so c^2 g / q^2 per mean sample is
    # c * c * n_mean * g / q^2 with g = pb (the directions were drawn from the BSDF)
    q_ = np.exp(logq)
    dsel = np.where(sel, -(c * c * n_mean) * np.where(sel, pb, 0.0) * (np.exp(logp) - np.where(sel, pb, 0.0)) / (q_ * q_)
                    * (SEL_MAX - SEL_MIN) * sg * (1.0 - sg), 0.0)
    cc = c[:, None]
    x = 2.0 * kap
    # x / (e^x - 1), written as x e^-x / (1 - e^-x) so that large x underflows to 0 instead of
    # overflowing expm1; the series 1 - x/2 below 1e-4.
    xs = np.maximum(x, 1e-300)
    ratio = np.where(x < 1e-4, 1.0 - 0.5 * x, xs * np.exp(-xs) / -np.expm1(-xs))
    dl = -cc * r * (1.0 - ratio + kap * (dcos - 1.0))
    dl = np.where((lraw < LOG_KAPPA_MIN) | (lraw > LOG_KAPPA_MAX), 0.0, dl)
    gm = (-cc * r * kap / ln[:, :, 0])[:, :, None] * (dirs[:, None, :] - dcos[:, :, None] * mu)
    dz = cc * (np.exp(wlog) - r)
    dO = np.zeros((n, N_OUT_PAD))
    g = np.concatenate([gm, dl[:, :, None], dz[:, :, None]], axis=2).reshape(n, N_OUT)
    dO[:, :N_OUT] = g
    dO[:, SEL] = dsel
    return -c * logq, dO


def split_samples(samples: np.ndarray, lo, hi):
    """(n, 13) or (n, 14) samples: position, normal, wo, direction, weight c and, in column 13,
    the BSDF density of the direction (<= 0 or absent: none)."""
    s = np.asarray(samples, np.float64)
    s = s.reshape(-1, s.shape[-1] if s.ndim == 2 else 13)
    s9 = scale_queries(s[:, :9], lo, hi)
    dirs = s[:, 9:12]
    c = np.maximum(np.where(np.isfinite(s[:, 12]), s[:, 12], 0.0), 0.0)
    pb = s[:, 13] if s.shape[1] > 13 else np.full(len(c), -1.0)
    pb = np.where(np.isfinite(pb), pb, -1.0)
    return s9, dirs, c, pb


def loss(params: np.ndarray, samples: np.ndarray, lo, hi, quant: bool = False) -> float:
    s9, dirs, c, pb = split_samples(samples, lo, hi)
    _, O = forward(params, encode(s9), quant)
    l, _ = _loss_and_dO(O, dirs, c, pb)
    return float(l.sum() / len(c))


def surrogate(params: np.ndarray, samples: np.ndarray, lo, hi, at: np.ndarray) -> float:
    """The function loss_grad differentiates, evaluated at params with p^ and the hidden
    features under the selection row taken at `at`: (1/n) sum [-c log p + c^2 pb / (s p^ +
    (1 - s) pb)] (the second term only where pb > 0).  Its derivative at params = at is
    loss_grad's gradient (finite-difference checks)."""
    s9, dirs, c, pb = split_samples(samples, lo, hi)
    X = encode(s9)
    _, O = forward(params, X)
    Aa, Oa = forward(at, X)
    mu, kappa, w = decode(O)
    p = mixture_pdf(mu, kappa, w, dirs)[:, 0]
    mua, kappaa, wa = decode(Oa)
    pa = mixture_pdf(mua, kappaa, wa, dirs)[:, 0]
    # the selection row on the hidden features at `at` (they get no gradient from it)
    Ws, bs = unpack(params)
    Osel = Oa.copy()
    Osel[:, SEL] = Aa[4] @ Ws[4][SEL] + bs[4][SEL]
    s_ = selection(Osel)
    pbs = np.where(pb > 0.0, pb, 1.0)
    second = np.where(pb > 0.0, c * c * pbs / (s_ * pa + (1.0 - s_) * pbs), 0.0)
    return float(np.sum(-c * np.log(p) + second) / len(c))


def loss_grad(params: np.ndarray, samples: np.ndarray, lo, hi, quant: bool = False):
    """Mean loss and the gradient (of surrogate(); module docstring) w.r.t. params (float64),
    by hand-written backpropagation."""
    s9, dirs, c, pb = split_samples(samples, lo, hi)
    n = len(c)
    Ws, bs = unpack(params)
    q = bf16 if quant else (lambda v: v)
    A, O = forward(params, encode(s9), quant)
    l, dO = _loss_and_dO(O, dirs, c / n, pb, n)
    gW, gb = [None] * 5, [None] * 5
    dz = q(dO)
    for layer in range(4, -1, -1):
        gW[layer] = dz.T @ A[layer]
        gb[layer] = dz.sum(axis=0)
        if layer > 0:
            Wb = Ws[layer].copy()
            if layer == 4:
                Wb[N_OUT:] = 0.0   # the selection row stays out of the hidden layers' gradient
            da = dz @ q(Wb)
            dz = q(np.where(A[layer] > 0.0, da, 0.0))
    return float(l.sum()), pack(gW, gb)


def eval_lobes(params: np.ndarray, queries: np.ndarray, lo, hi, quant: bool = False):
    s9 = scale_queries(queries, lo, hi)
    _, O = forward(params, encode(s9), quant)
    return decode(O)


def eval_record(params: np.ndarray, queries: np.ndarray, lo, hi, quant: bool = False):
    """The lobes as inference stores them (weights times the selection probability)."""
    s9 = scale_queries(queries, lo, hi)
    _, O = forward(params, encode(s9), quant)
    return decode_record(O)


def init_params(seed: int) -> np.ndarray:
    """The device's He-uniform initialisation (splitmix64 stream), as float32 values."""
    mask = (1 << 64) - 1
    s = (seed * 0x9E3779B97F4A7C15 + 1) & mask
    p = np.zeros(N_PARAMS, np.float32)
    ins = [N_IN, 64, 64, 64, 64]
    outs = [64, 64, 64, 64, N_OUT]
    offs = np.cumsum([0] + [a * b for a, b in W_SHAPES])
    real0 = [2 * p + t for p, jk in enumerate(feature_pairs()) if jk is not None for t in (0, 1)]
    for l in range(5):
        lim = np.float32(np.sqrt(np.float32(6.0) / np.float32(ins[l]))) * np.float32(0.1 if l == 4 else 1.0)
        K = W_SHAPES[l][1]
        cols = real0 if l == 0 else list(range(ins[l]))
        for o in range(outs[l]):
            for i in cols:
                s = (s + 0x9E3779B97F4A7C15) & mask
                z = s
                z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
                z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
                z ^= z >> 31
                u = np.float32(z >> 40) * np.float32(1.0 / 16777216.0)
                p[offs[l] + o * K + i] = (np.float32(2.0) * u - np.float32(1.0)) * lim
    return p
