#!/usr/bin/env python3
"""Closest-hit throughput on the ~1.0M-triangle displaced icosphere with 2^22 rays (one
MI355X), next to the reference's own BVH timed on the host cores.

One "step" = one closest-hit traversal launch over the whole 2^22-ray batch, rays and BVH
already resident in HBM.  Prints ONE JSON line (rank 0).  `--gpus N` with N > 1 exits: this
path runs on one GPU (the design for several is a replicated BVH with sharded rays).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ray-based-rendering_amd"))

HBM_PEAK_GBS = 8000.0
SEED = 0xC1
N_SUB = 224            # 20 * 224^2 = 1,003,520 triangles
N_RAYS = 1 << 22


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(tris, rays, gpu_hits, seconds: float) -> dict:
    """The reference's BVH::build + BVH::intersect (compiled from its sources into
    oracle/_ref/libref.so), OpenMP over rays on every host core, on a bounded ray sample
    sized to about `seconds` of CPU work.  Falls back to the oracle restatement ("port")
    where the reference build is absent.  Also checks the GPU ids on that sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O

    threads = O.max_threads()
    if O.ref_available():
        w = O.RefWorld(tris=tris)
        build_ms = w.build_bvh()
        kind = "reference"

        def run(r):
            h, sec = w.bvh_closest(r, threads=threads)
            return h, sec
    else:
        w = O.OracleScene(tris=tris)
        t0 = time.perf_counter()
        w.build_bvh()
        build_ms = (time.perf_counter() - t0) * 1e3
        kind = "port"

        def run(r):
            t0 = time.perf_counter()
            h = w.closest(r, threads=threads)
            return h, time.perf_counter() - t0
    probe = min(len(rays), 1 << 16)
    _, sec = run(rays[:probe])
    n = int(min(len(rays), max(probe, probe * seconds / max(sec, 1e-6))))
    h, sec = run(rays[:n])
    g = gpu_hits[:n]
    ids_equal = int((h["id"] == g["id"]).sum())
    miss_c, miss_g = h["id"] == 0xFFFFFFFF, g["id"] == 0xFFFFFFFF
    hit = ~miss_c
    t_equal = bool(np.array_equal(miss_c, miss_g) and np.array_equal(h["t"][hit], g["t"][hit]))
    # Every id difference must be an exact tie: the GPU's (lower) id and the CPU's id both
    # give the same t under the Moller-Trumbore test of the oracle.
    diff = np.nonzero(h["id"] != g["id"])[0]
    ties = 0
    for i in diff[:1000]:
        if miss_c[i] or miss_g[i]:
            continue
        tg = O.triangle_t(tris[g["id"][i]], rays[i])
        tc = O.triangle_t(tris[h["id"][i]], rays[i])
        ties += int(tg is not None and tc is not None and tg == tc == float(g["t"][i]) and g["id"][i] < h["id"][i])
    return {
        "value": n / sec / 1e6, "unit": "Mrays/s", "cores": threads, "kind": kind,
        "sample": f"first {n} of the {len(rays)} closest-hit rays, same mesh; BVH build {build_ms:.0f} ms (serial)",
        "cpu": platform.processor() or platform.machine(),
        "parity_on_sample": {"rays": n, "ids_equal": ids_equal, "hit_sets_equal_and_t_bit_equal": t_equal,
                             "id_differences": int(len(diff)), "id_differences_that_are_exact_ties": ties},
    }


def layout_name(width: int) -> str:
    """bvh4 / bvh8 (float child boxes) or bvh4q / bvh8q (quantized child boxes)."""
    return f"bvh{width & 0xFF}" + ("q" if width & 0x100 else "")


def tuned_layout() -> tuple[int, int, dict | None]:
    """BVH width code, leaf size and launch tuning of the fastest closest-hit configuration in
    the traversal tuning cache (autotune.py), else quantized BVH4 with the library defaults."""
    sys.path.insert(0, os.path.join(ROOT, "ray-based-rendering_amd"))
    import autotune
    cfg = autotune.best_config("closest")
    if not cfg:
        return 4 | 0x100, 4, None
    return int(cfg["width"]), int(cfg["leaf"]), cfg


def profile_summary(width: int) -> dict:
    """The committed rocprofv3 --pmc summary of this workload for this BVH layout (empty when
    there is none): calibrated HBM bytes per closest-hit launch (FETCH_SIZE + WRITE_SIZE, scaled
    by the same counters' reading for a copy kernel of known size) and the pipeline counters."""
    path = os.path.join(ROOT, "profiles", f"traversal_closest_c1_{layout_name(width)}_pmc.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def limiter_text(pm: dict) -> str | None:
    """What the counters say limits the traversal, from the same summary."""
    need = ("l1_to_l2_latency_cycles", "TD_TD_BUSY_sum_frac", "TA_TA_BUSY_sum_frac", "valu_issue_frac")
    if not all(k in pm for k in need):
        return None
    return (f"vector memory pipeline: about one L1 miss per node or primitive fetch at a "
            f"~{pm['l1_to_l2_latency_cycles']:.0f}-cycle L2 latency (texture data unit busy "
            f"{100 * pm['TD_TD_BUSY_sum_frac']:.0f} %, address unit {100 * pm['TA_TA_BUSY_sum_frac']:.0f} %, "
            f"VALU issue {100 * pm['valu_issue_frac']:.0f} % of SIMD quad-cycles), from the TA/TD/TCP/SQ "
            f"counters in profiles/{pm.get('name', '')}.md")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--width", type=int, default=None,
                    help="BVH width 4/8, +256 for quantized child boxes (default: the tuned layout)")
    ap.add_argument("--leaf", type=int, default=None)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a 256 MiB streaming copy kernel 3x (byte-counter calibration under rocprofv3 --pmc)")
    ap.add_argument("--no-paths", action="store_true", help="skip the path-tracing lines (C0, C2, C3)")
    ap.add_argument("--no-c4", action="store_true", help="skip guided vs plain time-to-relMSE (C4)")
    ap.add_argument("--path-cpu-seconds", type=float, default=5.0)
    ap.add_argument("--no-device-build", action="store_true",
                    help="skip the same mesh built on the GPU (build time, its trace rate, hits against the host tree)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this path runs on one GPU; --gpus {args.gpus} is not supported "
              "(several GPUs would replicate the BVH and shard the rays; see DESIGN.md)", file=sys.stderr)
        return 2

    import rbr_amd as R

    log("bench: generating mesh and rays")
    tris = R.icosphere(N_SUB, 100.0, 0.1, SEED)
    lo, hi = R.mesh_bounds(tris)
    rays = R.rays_uniform(lo, hi, N_RAYS, SEED + 1)
    shadow = R.rays_shadow(lo, hi, N_RAYS, SEED + 2)

    ctx = R.Context(0)
    dev_name, cus = ctx.device_info()
    width, leaf, tuning = tuned_layout()
    if args.width is not None:
        width, tuning = args.width, None
    if args.leaf is not None:
        leaf = args.leaf
    args.width, args.leaf = width, leaf
    if tuning:
        ctx.check(R._lib.rbr_set_trace_tuning(ctx.handle, tuning["mode"], tuning["lds_depth"], tuning["waves_per_cu"]))
    sc = R.Scene(ctx)
    sc.set_geometry(tris=tris)
    log("bench: building BVH")
    sc.build_bvh(args.width, args.leaf)
    st = sc.stats()

    d_rays = ctx.alloc(rays.nbytes)
    d_shadow = ctx.alloc(shadow.nbytes)
    d_hits = ctx.alloc(N_RAYS * 8)
    d_occ = ctx.alloc(N_RAYS)
    d_rays.upload(rays)
    d_shadow.upload(shadow)

    log("bench: warmup")
    for _ in range(args.warmup):
        sc.trace_closest_device(d_rays, N_RAYS, d_hits)
        sc.trace_any_device(d_shadow, N_RAYS, d_occ)
    ctx.synchronize()

    log("bench: timed closest-hit")
    ctx.synchronize()
    t0 = time.perf_counter()
    ms_closest = sc.time_trace(0, d_rays, N_RAYS, d_hits, args.steps)
    ctx.synchronize()
    wall_s = time.perf_counter() - t0
    ms_any = sc.time_trace(1, d_shadow, N_RAYS, d_occ, max(1, args.steps // 2))

    hits = np.empty(N_RAYS, R.HIT_DTYPE)
    d_hits.download(hits)
    if args.calib_copy:
        nb = 256 << 20
        src, dst = ctx.alloc(nb), ctx.alloc(nb)
        for _ in range(3):
            ctx.check(R._lib.rbr_calib_copy(ctx.handle, dst.ptr, src.ptr, nb))
        ctx.synchronize()
        src.free()
        dst.free()

    # Algorithmic bytes per ray from the host traversal of the same layout on a sample.
    sample = rays[:: max(1, N_RAYS // (1 << 16))]
    _, nv, pt = sc.host_trace_closest(sample)
    node_b = st["node_bytes"] / max(1, st["node_count"])   # 64/128 B (quantized) or 128/256 B
    nodes_per_ray, prims_per_ray = nv / len(sample), pt / len(sample)
    bytes_compulsory = 32 + 8
    # A triangle test reads its three vertices (36 B); the leaf record the kernel loads is 48 B
    # (v0, e1, e2 and the id/kind words), reported as the layout figure.
    bytes_per_ray = bytes_compulsory + nodes_per_ray * node_b + prims_per_ray * 36
    bytes_per_ray_layout = bytes_compulsory + nodes_per_ray * node_b + prims_per_ray * 48
    rays_per_s = N_RAYS / (ms_closest * 1e-3)
    achieved = rays_per_s * bytes_per_ray / 1e9
    pm = profile_summary(args.width)
    traffic = pm.get("traffic_bytes_per_launch")

    rec = {
        "metric": "closest-hit Mrays/s (1.0M-tri displaced icosphere, 2^22 incoherent rays)",
        "value": rays_per_s / 1e6,
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_closest,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded icosphere + seeded rays)",
        "config": {
            "workload": "closest-hit, icosphere 20*224^2 = %d tris r=100 displaced 0.1r, 2^22 rays origins in bbox*1.1, "
                        "dirs uniform on S^2, tmin 1e-4, tmax inf" % len(tris),
            "bvh": f"{layout_name(args.width).upper()}, leaf<={args.leaf}, {node_b:.0f} B nodes", "nodes": st["node_count"],
            "max_depth": st["max_depth"],
            "launch": ({**{k: tuning[k] for k in ("mode", "lds_depth", "waves_per_cu")},
                        "resident_waves_per_cu": sc.resident_waves(False)}
                       if tuning
                       else "library defaults (persistent, LDS 8, 32 waves/CU)"),
            "seed": SEED,
        },
        "wall_s_timed_region": wall_s,
        "any_hit_mrays_s": N_RAYS / (ms_any * 1e-3) / 1e6,
        "hit_fraction": float((hits["id"] != R.MISS).mean()),
        "device": dev_name, "cus": cus,
        "bvh_build_ms": st["build_ms"],
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_frac": (traffic / (ms_closest * 1e-3) / 1e9 / HBM_PEAK_GBS) if traffic else None,
            "limiter": limiter_text(pm),
            "bytes_per_ray": bytes_per_ray, "bytes_per_ray_layout": bytes_per_ray_layout,
            "frac_layout_bytes": rays_per_s * bytes_per_ray_layout / 1e9 / HBM_PEAK_GBS,
            "nodes_per_ray": nodes_per_ray, "prims_per_ray": prims_per_ray,
            "frac_compulsory_only": rays_per_s * bytes_compulsory / 1e9 / HBM_PEAK_GBS,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        log("bench: CPU baseline")
        rec["cpu_baseline"] = cpu_baseline(tris, rays, hits, args.cpu_seconds)
    if not args.no_device_build:
        # The same mesh and layout built on the GPU (rbr_scene_build_bvh_device), outside the timed
        # steps: build time (best of three), closest-hit rate on its tree, and its hits, which must
        # equal the host tree's (hits do not depend on the tree).
        log("bench: device BVH build")
        dsc = R.Scene(ctx)
        dsc.set_geometry(tris=tris)
        build_ms = []
        for _ in range(3):
            dsc.build_bvh(args.width, args.leaf, device=True)
            build_ms.append(dsc.stats()["build_ms"])
        for _ in range(args.warmup):
            dsc.trace_closest_device(d_rays, N_RAYS, d_hits)
        ms_dev = dsc.time_trace(0, d_rays, N_RAYS, d_hits, max(1, args.steps // 2))
        dh = np.empty(N_RAYS, R.HIT_DTYPE)
        d_hits.download(dh)
        dst = dsc.stats()
        rec["device_build"] = {
            "build_ms": min(build_ms), "nodes": dst["node_count"], "max_depth": dst["max_depth"],
            "closest_mrays_s": N_RAYS / (ms_dev * 1e-3) / 1e6,
            "hits_equal_host_tree": bool(np.array_equal(dh["id"], hits["id"]) and np.array_equal(dh["t"], hits["t"])),
        }
        dsc.close()
    for b in (d_rays, d_shadow, d_hits, d_occ):
        b.free()
    sc.close()
    # The other workloads of the metric, outside the timed region of the closest-hit steps.
    if not args.no_paths:
        import bench_paths as BP
        rec["paths"] = {}
        for key in ("c0", "c2", "c3"):
            log(f"bench: path tracing {key}")
            rec["paths"][key] = BP.measure(ctx, key, cpu_seconds=0.0 if args.no_cpu else args.path_cpu_seconds, log=log)
    if not args.no_c4:
        import bench_guiding as BG
        log("bench: c4 guided vs plain time to relMSE 0.01")
        rec["c4"] = BG.c4_summary(ctx)
        rec["time_to_relmse_0.01_plain_s"] = rec["c4"]["plain_s"]
        rec["time_to_relmse_0.01_guided_s"] = rec["c4"]["guided_s"]
        rec["time_to_relmse_0.01_speedup"] = rec["c4"]["speedup"]
    print(json.dumps(rec), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
