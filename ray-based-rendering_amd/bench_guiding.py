"""Learned path guiding vs the plain NEE/MIS tracer: time to a target relMSE.

  python ray-based-rendering_amd/bench_guiding.py [--scene terrain|cornell_indirect]
      [--width 1920 --height 1080] [--ref-spp 1024] [--schedules '[{...}, ...]']

Scenes (depth 8):
  terrain           the C2 workload: 524,288 triangles, 64 spheres, 4 area lights + sun
  cornell_indirect  the Cornell box lit only indirectly (light panel facing the ceiling)

* Reference: the plain GPU tracer at --ref-spp samples per pixel, accumulated in FP64 on the
  device (rbr_render_moments).  Its own error is estimated from its per-pixel variance and
  reported ("ref_relmse").
* relMSE(I) = mean over pixels and RGB channels of (I - R)^2 / (R^2 + 1e-2).
* Plain: progressive passes of --spp-per-call samples; the image is the running mean.
* Guided, one run per schedule (a dict; missing keys take the defaults below):
    train_passes  passes that record training samples, each followed by train_steps Adam
                  steps of --batch records (then the records are cleared)
    train_spp     samples per pixel of a training pass
    train_mode    "guided": the guide is sampled while recording (guided = 3);
                  "plain": records come from the plain integrator's paths (guided = 2)
    online_steps  after the training passes, every guided pass also records and is followed
                  by this many Adam steps (0: guided passes only sample, guided = 1)
    guide_depth   bounces that draw from the guide (0 = all)
  Every pass, training passes included, is unbiased and enters the running mean; training
  time is part of the guided time.  The guide starts from the same initialisation each run.
  The record capacity (--reserve) holds a whole 1080p pass, so every bounce is recorded.
* Times are host wall-clock around the render and training calls (each render call returns
  its film, so they include the film download); relMSE evaluation is not timed.

Prints one JSON line per schedule with both convergence curves and the time to each target.
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
DEFAULT_SCHEDULE = {"train_passes": 8, "train_spp": 2, "train_mode": "plain", "train_steps": 16,
                    "online_steps": 0, "guide_depth": 0}
SCENES = {"terrain": S.terrain_c2, "cornell_indirect": S.cornell_indirect}


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


class Progressive:
    """Running mean of passes with its (seconds, relMSE) curve."""

    def __init__(self, ref):
        self.ref, self.acc, self.n, self.t, self.curve = ref, None, 0, 0.0, []

    def add(self, film, k, dt):
        self.t += dt
        f = film.astype(np.float64) * k
        self.acc = f if self.acc is None else self.acc + f
        self.n += k
        e = relmse(self.acc / self.n, self.ref)
        self.curve.append((self.t, e))
        return e


def run_plain(ctx, sc, ref, args, log):
    p = Progressive(ref)
    call = 0
    while p.n < args.max_spp:
        k = min(args.spp_per_call, args.max_spp - p.n)
        t0 = time.perf_counter()
        film = R.render(ctx, [dict(scene=sc, spp=k, depth=8, seed=1000 + call)])[0]
        e = p.add(film, k, time.perf_counter() - t0)
        call += 1
        if log:
            log(f"  plain spp {p.n:5d}  t {p.t:8.3f}s  relMSE {e:.5f}")
    return p


def run_guided(ctx, sc, ref, args, sched, log):
    lo, hi = sc.world_bounds()
    g = R.Guide(ctx, lo, hi, seed=1)
    g.reserve(args.reserve)
    g.clear_collected()
    losses = []
    p = Progressive(ref)
    gd = sched["guide_depth"]
    mode = 3 if sched["train_mode"] == "guided" else 2
    for i in range(sched["train_passes"]):
        k = sched["train_spp"]
        t0 = time.perf_counter()
        film = R.render(ctx, [dict(scene=sc, spp=k, depth=8, seed=50000 + i, guided=mode, guide_depth=gd)])[0]
        losses.append(g.train_collected(sched["train_steps"], args.batch, args.lr))
        g.clear_collected()
        e = p.add(film, k, time.perf_counter() - t0)
        if log:
            log(f"  train spp {p.n:5d}  t {p.t:8.3f}s  relMSE {e:.5f}  loss {losses[-1]:.4f}")
    online = sched["online_steps"]
    call = 0
    while p.n < args.max_spp:
        k = min(args.spp_per_call, args.max_spp - p.n)
        t0 = time.perf_counter()
        film = R.render(ctx, [dict(scene=sc, spp=k, depth=8, seed=90000 + call, guided=3 if online else 1,
                                   guide_depth=gd)])[0]
        if online:
            losses.append(g.train_collected(online, args.batch, args.lr))
            g.clear_collected()
        e = p.add(film, k, time.perf_counter() - t0)
        call += 1
        if log:
            log(f"  guided spp {p.n:5d}  t {p.t:8.3f}s  relMSE {e:.5f}")
    return p, losses


def c4_summary(ctx, width: int = 1920, height: int = 1080, ref_spp: int = 1024, max_spp: int = 64,
               spp_per_call: int = 4, target: float = 0.01, schedule: dict | None = None, batch: int = 1 << 18,
               reserve: int = 1 << 24, log=None) -> dict:
    """Time to relMSE <= target on the C2 terrain for the plain GPU NEE/MIS tracer and the
    guided tracer (one schedule), against a ref_spp-sample plain reference accumulated in FP64,
    for the benchmark's JSON line."""
    args = argparse.Namespace(max_spp=max_spp, spp_per_call=spp_per_call, batch=batch, lr=1e-3, reserve=reserve)
    sched = {**DEFAULT_SCHEDULE, **(schedule or {})}
    sc = R.scene_from_desc(ctx, S.terrain_c2(width, height))
    R.render(ctx, [dict(scene=sc, spp=1, depth=8, seed=1)])
    _, s1, s2 = R.render_moments(ctx, dict(scene=sc, spp=ref_spp, depth=8, seed=0x5EED))
    ref = s1 / ref_spp
    var = np.maximum(s2 / ref_spp - ref * ref, 0.0)
    ref_err = float(np.mean(var / ref_spp / (ref * ref + EPS)))
    plain = run_plain(ctx, sc, ref, args, log)
    g0 = R.Guide(ctx, *sc.world_bounds(), seed=1)   # warm-up of the guided code paths
    R.render(ctx, [dict(scene=sc, spp=1, depth=8, seed=7, guided=3, guide_depth=sched["guide_depth"])])
    g0.train_collected(1, batch, 1e-3)
    g0.clear_collected()
    guided, losses = run_guided(ctx, sc, ref, args, sched, log)
    tp, tg = time_to(plain.curve, target), time_to(guided.curve, target)
    sc.close()
    return {"workload": f"c4 guiding, terrain {width}x{height}, depth 8", "target_relmse": target,
            "plain_s": tp, "guided_s": tg, "speedup": (tp / tg) if (tp and tg) else None,
            "reference": {"spp": ref_spp, "accumulation": "fp64", "ref_relmse": ref_err},
            "schedule": sched, "loss_first": losses[0] if losses else None, "loss_last": losses[-1] if losses else None,
            "final": {"spp": max_spp, "plain_relmse": plain.curve[-1][1], "guided_relmse": guided.curve[-1][1],
                      "plain_s": plain.curve[-1][0], "guided_s": guided.curve[-1][0]}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--scene", choices=sorted(SCENES), default="terrain")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--ref-spp", type=int, default=1024)
    ap.add_argument("--max-spp", type=int, default=256)
    ap.add_argument("--spp-per-call", type=int, default=4)
    ap.add_argument("--batch", type=int, default=1 << 18)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--reserve", type=int, default=1 << 24,
                    help="record capacity: records of a pass beyond it are dropped, and records are "
                         "appended bounce by bounce, so it must hold a whole pass to cover deep vertices")
    ap.add_argument("--schedules", default=json.dumps([DEFAULT_SCHEDULE]),
                    help="JSON list of guided schedules (keys as in the module docstring)")
    ap.add_argument("--targets", default="0.01,0.003,0.001")
    ap.add_argument("--quiet", action="store_true")
    args = ap.parse_args()
    log = None if args.quiet else (lambda m: print(m, file=sys.stderr, flush=True))
    schedules = [{**DEFAULT_SCHEDULE, **s} for s in json.loads(args.schedules)]

    ctx = R.Context(0)
    sc = R.scene_from_desc(ctx, SCENES[args.scene](args.width, args.height))
    R.render(ctx, [dict(scene=sc, spp=1, depth=8, seed=1)])   # warm-up (code objects, pool)

    t0 = time.perf_counter()
    _, s1, s2 = R.render_moments(ctx, dict(scene=sc, spp=args.ref_spp, depth=8, seed=0x5EED))
    ref_s = time.perf_counter() - t0
    ref = s1 / args.ref_spp
    var = np.maximum(s2 / args.ref_spp - ref * ref, 0.0)
    ref_err = float(np.mean(var / args.ref_spp / (ref * ref + EPS)))
    if log:
        log(f"reference: {args.ref_spp} spp in {ref_s:.1f} s, estimated relMSE {ref_err:.2e}")
    plain = run_plain(ctx, sc, ref, args, log)
    targets = [float(x) for x in args.targets.split(",")]
    for sched in schedules:
        # warm-up of the guided code paths (graphs, inference) outside the timed run
        g0 = R.Guide(ctx, *sc.world_bounds(), seed=1)
        R.render(ctx, [dict(scene=sc, spp=1, depth=8, seed=7, guided=3, guide_depth=sched["guide_depth"])])
        g0.train_collected(1, args.batch, args.lr)
        g0.clear_collected()
        if log:
            log(f"guided {sched}:")
        guided, losses = run_guided(ctx, sc, ref, args, sched, log)
        tt = {}
        for tgt in targets:
            tp, tg = time_to(plain.curve, tgt), time_to(guided.curve, tgt)
            tt[str(tgt)] = {"plain_s": tp, "guided_s": tg,
                            "speedup": (tp / tg) if (tp is not None and tg is not None and tg > 0) else None}
        out = {
            "workload": f"c4 guiding {args.scene} {args.width}x{args.height} depth 8",
            "metric": "time-to-relMSE guided vs plain GPU PT",
            "reference": {"spp": args.ref_spp, "accumulation": "fp64", "ref_relmse": ref_err},
            "schedule": sched,
            "training": {"batch": args.batch, "lr": args.lr,
                         "guide_selection": "learned per vertex, s in [0.1, 0.9]",
                         "loss_first": losses[0] if losses else None,
                         "loss_last": losses[-1] if losses else None},
            "max_spp": args.max_spp,
            "time_to_target": tt,
            "final_relmse": {"plain": plain.curve[-1][1], "guided": guided.curve[-1][1]},
            "final_time_s": {"plain": plain.curve[-1][0], "guided": guided.curve[-1][0]},
            "curve_plain": [[round(t, 4), e] for t, e in plain.curve],
            "curve_guided": [[round(t, 4), e] for t, e in guided.curve],
        }
        print(json.dumps(out), flush=True)
    sc.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
