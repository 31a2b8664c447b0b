"""Learned path guiding vs the plain NEE/MIS tracer: time to a target relMSE.

  python ray-based-rendering_amd/bench_guiding.py [--width 1920 --height 1080] [--ref-spp 1024]

Scene: the terrain workload (524,288 triangles, 64 spheres, 4 area lights + sun), depth 8.

* Reference: the plain GPU tracer at --ref-spp samples per pixel, accumulated in FP64 on the
  device (rbr_render_moments).  Its own error is estimated from its per-pixel variance and
  reported ("ref_relmse").
* relMSE(I) = mean over pixels and RGB channels of (I - R)^2 / (R^2 + 1e-2).
* Plain: progressive passes of --spp-per-call samples; the image is the running mean.
* Guided: the guide is trained online while rendering: --train-passes passes of one sample
  each that sample the guide and record training samples (guided = 3), each followed by
  --train-steps Adam steps of --batch records; then guided passes (guided = 1).  Every pass,
  training passes included, is unbiased and enters the running mean.  Training time is
  part of the guided time.
* Times are host wall-clock around the render and training calls (each render call returns
  its film, so they include the film download); relMSE evaluation is not timed.

Prints one JSON line with both convergence curves and the time to each target relMSE.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, PKG)

import rbr_amd as R  # noqa: E402
import scenes as S  # noqa: E402

EPS = 1e-2


def relmse(img: np.ndarray, ref: np.ndarray) -> float:
    d = img.astype(np.float64) - ref
    return float(np.mean(d * d / (ref * ref + EPS)))


def time_to(curve, target):
    """Time at which the relMSE curve first reaches target (log-linear interpolation between
    checkpoints), or None."""
    prev = None
    for t, e in curve:
        if e <= target:
            if prev is None:
                return t
            (t0, e0) = prev
            if e0 <= e or e0 <= 0:
                return t
            a = (np.log(e0) - np.log(target)) / (np.log(e0) - np.log(e))
            return t0 + a * (t - t0)
        prev = (t, e)
    return None


def progressive(ctx, sc, ref, max_spp, per_call, seed, guided=0, pre=None, log=None, guide_depth=0):
    """Runs passes until max_spp; returns the [(seconds, relMSE)] curve and spp reached."""
    acc = None
    n = 0
    t = 0.0
    curve = []
    if pre is not None:   # training passes: list of (callable returning (film, spp, seconds))
        for step in pre:
            film, k, dt = step()
            t += dt
            acc = film.astype(np.float64) * k if acc is None else acc + film.astype(np.float64) * k
            n += k
            curve.append((t, relmse(acc / n, ref)))
    call = 0
    while n < max_spp:
        k = min(per_call, max_spp - n)
        t0 = time.perf_counter()
        film = R.render(ctx, [dict(scene=sc, spp=k, depth=8, seed=seed + call, guided=guided,
                                   guide_depth=guide_depth)])[0]
        t += time.perf_counter() - t0
        call += 1
        acc = film.astype(np.float64) * k if acc is None else acc + film.astype(np.float64) * k
        n += k
        e = relmse(acc / n, ref)
        curve.append((t, e))
        if log:
            log(f"  spp {n:5d}  t {t:8.3f}s  relMSE {e:.5f}")
    return curve, n


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--ref-spp", type=int, default=1024)
    ap.add_argument("--max-spp", type=int, default=256)
    ap.add_argument("--spp-per-call", type=int, default=4)
    ap.add_argument("--train-passes", type=int, default=16)
    ap.add_argument("--train-steps", type=int, default=8)
    ap.add_argument("--batch", type=int, default=1 << 18)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--guide-depth", type=int, default=0, help="bounces that draw from the guide (0 = all)")
    ap.add_argument("--targets", default="0.01,0.003,0.001")
    ap.add_argument("--quiet", action="store_true")
    args = ap.parse_args()
    log = None if args.quiet else (lambda m: print(m, file=sys.stderr, flush=True))

    ctx = R.Context(0)
    desc = S.terrain_c2(args.width, args.height)
    sc = R.scene_from_desc(ctx, desc)
    R.render(ctx, [dict(scene=sc, spp=1, depth=8, seed=1)])   # warm-up (code objects, pool)

    t0 = time.perf_counter()
    _, s1, s2 = R.render_moments(ctx, dict(scene=sc, spp=args.ref_spp, depth=8, seed=0x5EED))
    ref_s = time.perf_counter() - t0
    ref = s1 / args.ref_spp
    var = np.maximum(s2 / args.ref_spp - ref * ref, 0.0)
    ref_err = float(np.mean(var / args.ref_spp / (ref * ref + EPS)))
    if log:
        log(f"reference: {args.ref_spp} spp in {ref_s:.1f} s, estimated relMSE {ref_err:.2e}")

    if log:
        log("plain:")
    plain, _ = progressive(ctx, sc, ref, args.max_spp, args.spp_per_call, 1000, 0, log=log)

    lo, hi = sc.world_bounds()
    g = R.Guide(ctx, lo, hi, seed=1)
    g.clear_collected()
    losses = []

    def train_pass(i):
        def step():
            t0 = time.perf_counter()
            film = R.render(ctx, [dict(scene=sc, spp=1, depth=8, seed=50000 + i, guided=3,
                                       guide_depth=args.guide_depth)])[0]
            losses.append(g.train_collected(args.train_steps, args.batch, args.lr))
            g.clear_collected()
            return film, 1, time.perf_counter() - t0
        return step

    if log:
        log("guided:")
    guided, _ = progressive(ctx, sc, ref, args.max_spp, args.spp_per_call, 90000, 1,
                            pre=[train_pass(i) for i in range(args.train_passes)], log=log,
                            guide_depth=args.guide_depth)
    targets = [float(x) for x in args.targets.split(",")]
    tt = {}
    for tgt in targets:
        tp, tg = time_to(plain, tgt), time_to(guided, tgt)
        tt[str(tgt)] = {"plain_s": tp, "guided_s": tg,
                        "speedup": (tp / tg) if (tp is not None and tg is not None and tg > 0) else None}
    out = {
        "workload": f"c4 guiding terrain {args.width}x{args.height} depth 8",
        "metric": "time-to-relMSE guided vs plain GPU PT",
        "reference": {"spp": args.ref_spp, "accumulation": "fp64", "ref_relmse": ref_err},
        "training": {"passes": args.train_passes, "steps_per_pass": args.train_steps, "batch": args.batch,
                     "lr": args.lr, "guide_fraction": 0.5, "guide_depth": args.guide_depth,
                     "loss_first": losses[0] if losses else None,
                     "loss_last": losses[-1] if losses else None},
        "max_spp": args.max_spp,
        "time_to_target": tt,
        "final_relmse": {"plain": plain[-1][1], "guided": guided[-1][1]},
        "final_time_s": {"plain": plain[-1][0], "guided": guided[-1][0]},
        "curve_plain": [[round(t, 4), e] for t, e in plain],
        "curve_guided": [[round(t, 4), e] for t, e in guided],
    }
    print(json.dumps(out), flush=True)
    sc.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
