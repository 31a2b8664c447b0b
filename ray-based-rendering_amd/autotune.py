"""Autotuner.

Traversal (default): sweeps BVH width, leaf size, launch mode, LDS stack depth and persistent
waves per CU on a resident ray batch, checks that every configuration returns bit-identical
hits, and records the fastest in tuning/traversal.json (read by `best_config()`; bench.py
uses it when present).

  python ray-based-rendering_amd/autotune.py [--quick] [--out PATH]

Render: for each path-tracing workload, sweeps the path pool (queue capacity), material-
ordered shading, hipGraph replay and the persistent traversal waves per CU; checks that every
configuration gives the same images (bit-identical at equal pool, float summation order
otherwise) and records the fastest per workload in tuning/render.json (read by
`render_config()`; bench_paths.py uses it when present).

  python ray-based-rendering_amd/autotune.py render [--only c0,c2,c3] [--quick]
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import sys
import time

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, PKG)
import rbr_amd as R  # noqa: E402

CACHE = os.path.join(PKG, "tuning", "traversal.json")
RENDER_CACHE = os.path.join(PKG, "tuning", "render.json")




def sweep(ctx: R.Context, tris: np.ndarray, rays: np.ndarray, kind: int, widths, leaves, modes, depths, waves,
          iters: int = 10, log=print) -> list[dict]:
    n = len(rays)
    d_rays = ctx.alloc(rays.nbytes)
    d_out = ctx.alloc(n * (8 if kind == 0 else 1))
    d_rays.upload(rays)
    results = []
    ref = None
    for width, leaf in itertools.product(widths, leaves):
        sc = R.Scene(ctx)
        sc.set_geometry(tris=tris)
        sc.build_bvh(width, leaf)
        for mode, depth, wpc in itertools.product(modes, depths, waves):
            if mode == 1 and wpc != waves[0]:
                continue  # waves per CU only matters for the persistent kernel
            ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, mode, depth, wpc))
            sc.time_trace(kind, d_rays, n, d_out, 2)
            # best of three timed runs: single runs of a few launches vary by 2-3 %
            ms = min(sc.time_trace(kind, d_rays, n, d_out, iters) for _ in range(3))
            out = np.empty(n, R.HIT_DTYPE if kind == 0 else np.uint8)
            d_out.download(out)
            same = True
            if ref is None:
                ref = out.copy()
            else:
                same = bool(np.array_equal(out, ref))
            # Waves resident at once (registers and LDS of the kernel); a larger grid is still
            # correct (its extra waves start as others end), so such points stay in the sweep.
            rec = {"width": width, "leaf": leaf, "mode": mode, "lds_depth": depth, "waves_per_cu": wpc,
                   "resident_waves_per_cu": sc.resident_waves(kind == 1) if mode != 1 else None,
                   "ms": ms, "mrays_s": n / ms / 1e3, "identical": same}
            log(json.dumps(rec))
            if not same:
                raise RuntimeError(f"configuration {rec} changed the results")
            results.append(rec)
        sc.close()
    ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, 0, 8, 32))
    d_rays.free()
    d_out.free()
    return results


def best_config(kind: str = "closest", path: str = CACHE) -> dict | None:
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(kind)


def apply(ctx: R.Context, cfg: dict | None) -> None:
    if cfg:
        ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, cfg["mode"], cfg["lds_depth"], cfg["waves_per_cu"]))


# ---------------------------------------------------------------- render tuning

def render_sweep(ctx: R.Context, rj: list, paths: int, pools, sorts, graphs, waves, modes=(0,), reps: int = 2,
                 log=print) -> list[dict]:
    results = []
    ref = None
    for pool, sort, graph, wpc, mode in itertools.product(pools, sorts, graphs, waves, modes):
        R.set_render_tuning(ctx, pool, graph, sort)
        ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, mode, 8, wpc))
        films = R.render(ctx, rj)   # warm-up; records the round graphs when on
        times = []
        for _ in range(reps):
            R.render(ctx, rj)
            times.append(R.last_render_ms(ctx))
        ms = min(times)
        if ref is None:
            ref = films
        else:
            for a, b in zip(films, ref):
                if not np.allclose(a, b, rtol=2e-6, atol=1e-7):
                    raise RuntimeError(f"pool={pool} sort={sort} graphs={graph} waves={wpc} mode={mode} changed the image")
        rec = {"pool_paths": pool, "sort_materials": int(sort), "use_graphs": int(graph), "waves_per_cu": wpc,
               "trace_mode": mode, "ms": ms, "mpaths_s": paths / ms / 1e3}
        log(json.dumps(rec))
        results.append(rec)
    R.set_render_tuning(ctx)
    ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, 0, 8, 32))
    return results


def render_config(workload: str, path: str = RENDER_CACHE) -> dict | None:
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("workloads", {}).get(workload, {}).get("best")


def apply_render(ctx: R.Context, cfg: dict | None) -> None:
    """Applies a render_config() entry (None: library defaults)."""
    if cfg:
        R.set_render_tuning(ctx, cfg["pool_paths"], bool(cfg["use_graphs"]), bool(cfg["sort_materials"]))
        ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, cfg.get("trace_mode", 0), 8, cfg["waves_per_cu"]))
    else:
        R.set_render_tuning(ctx)
        ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, 0, 8, 32))


def main_render(argv) -> int:
    import bench_paths as B
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default="c0,c2,c3")
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--out", default=RENDER_CACHE)
    args = ap.parse_args(argv)
    ctx = R.Context(0)
    name, cus = ctx.device_info()
    out = {"device": name, "cus": cus, "time": time.strftime("%Y-%m-%dT%H:%M:%S"), "workloads": {}}
    if os.path.exists(args.out):
        with open(args.out) as f:
            out["workloads"] = json.load(f).get("workloads", {})
    if args.quick:
        grid = dict(pools=[1 << 24, 1 << 25], sorts=[False, True], graphs=[True], waves=[16, 32], modes=[0, 11])
    else:
        grid = dict(pools=[1 << 22, 1 << 23, 1 << 24, 1 << 25], sorts=[False, True], graphs=[False, True],
                    waves=[16, 32], modes=[0, 6, 8, 11])
    for key in args.only.split(","):
        jobs = B.workload_jobs(key)
        scs = [R.scene_from_desc(ctx, j["desc"]) for j in jobs]
        rj = [dict(scene=s, spp=j["spp"], depth=j["depth"], seed=j.get("seed", 1), integrator=j.get("integrator", R.PATH))
              for s, j in zip(scs, jobs)]
        paths = sum(j["desc"].camera.width * j["desc"].camera.height * j["spp"] for j in jobs)
        print(f"workload {key}: {B.WORKLOADS[key]}", flush=True)
        res = render_sweep(ctx, rj, paths, **grid, log=lambda s: print("  " + s, flush=True))
        best = min(res, key=lambda x: x["ms"])
        out["workloads"][key] = {"name": B.WORKLOADS[key], "best": best}
        print(f"  best: {best}", flush=True)
        for s in scs:
            s.close()
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    ctx.close()
    return 0


def main() -> int:
    if len(sys.argv) > 1 and sys.argv[1] == "render":
        return main_render(sys.argv[2:])
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--out", default=CACHE)
    ap.add_argument("--n-sub", type=int, default=224)
    ap.add_argument("--rays", type=int, default=1 << 22)
    args = ap.parse_args()
    tris = R.icosphere(args.n_sub, 100.0, 0.1, 0xC1)
    lo, hi = R.mesh_bounds(tris)
    rays = R.rays_uniform(lo, hi, args.rays, 0xC1 + 1)
    shadow = R.rays_shadow(lo, hi, args.rays, 0xC1 + 2)
    ctx = R.Context(0)
    name, cus = ctx.device_info()
    if args.quick:
        grid = dict(widths=[4, 8, 4 | R.BVH_QUANTIZED, 8 | R.BVH_QUANTIZED], leaves=[4], modes=[0, 2, 6, 7, 8, 11, 12],
                    depths=[8, 10, 12], waves=[16, 24, 32])
    else:
        grid = dict(widths=[4, 8, 4 | R.BVH_QUANTIZED, 8 | R.BVH_QUANTIZED], leaves=[2, 4, 8], modes=list(range(13)),
                    depths=[8, 10, 12, 16, 24, 32], waves=[4, 8, 16, 24, 32])
    out = {"device": name, "cus": cus, "workload": f"icosphere n={args.n_sub} ({len(tris)} tris), {args.rays} rays",
           "time": time.strftime("%Y-%m-%dT%H:%M:%S")}
    for kind, r in (("closest", rays), ("any", shadow)):
        res = sweep(ctx, tris, r, 0 if kind == "closest" else 1, **grid, log=lambda s: print(s, flush=True))
        best = min(res, key=lambda x: x["ms"])
        out[kind] = best
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print("best closest:", out["closest"], "\nbest any:", out["any"])
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
