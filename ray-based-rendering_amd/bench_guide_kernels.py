"""Guiding-network kernels against the matrix-core roofline.

  python ray-based-rendering_amd/bench_guide_kernels.py

One JSON line per measurement (inputs resident in HBM, HIP events on the context stream; the
best of three timed runs of 20 launches, 10 for training steps):
  inference  n queries:  encoding + 5 layers + lobe decode (guideInferKernel)
  training   n samples:  forward + loss gradient + backward + weight/bias gradients + Adam
Algorithmic FLOP (real layer sizes, no padding):
  forward  2 * (216*64 + 3*64*64 + 64*40)            = 57,344 per query
  training forward + backward-data 2 * (40*64 + 3*64*64) + weight gradient 57,344
                                                      = 144,384 per sample
Peak: 2.5 PFLOP/s dense BF16 (MI355X).
"""
from __future__ import annotations

import json
import os
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, PKG)

import rbr_amd as R  # noqa: E402

FWD_FLOP = 2 * (216 * 64 + 3 * 64 * 64 + 64 * 40)
TRAIN_FLOP = FWD_FLOP + 2 * (40 * 64 + 3 * 64 * 64) + FWD_FLOP
PEAK = 2.5e15


def main() -> int:
    ctx = R.Context(0)
    g = R.Guide(ctx, [0, 0, 0], [1, 1, 1], seed=1)
    for n in (1 << 18, 1 << 20, 1 << 22):
        ms = min(g.time(0, n, 20) for _ in range(3))
        fl = FWD_FLOP * n / (ms * 1e-3)
        print(json.dumps({"kernel": "guide inference", "n": n, "ms": ms, "queries_s": n / (ms * 1e-3),
                          "tflop_s": fl / 1e12, "frac_bf16_peak": fl / PEAK}), flush=True)
    for n in (1 << 16, 1 << 18):
        ms = min(g.time(1, n, 10) for _ in range(3))
        fl = TRAIN_FLOP * n / (ms * 1e-3)
        print(json.dumps({"kernel": "guide training step", "n": n, "ms": ms, "samples_s": n / (ms * 1e-3),
                          "tflop_s": fl / 1e12, "frac_bf16_peak": fl / PEAK}), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
