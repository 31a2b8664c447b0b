"""Path-tracing throughput on the path-tracing workloads, one JSON line each:

  c0  Cornell 555, 256x256, 16 spp, depth 8, NEE + MIS
  c2  terrain (524,288 triangles) + 64 spheres, 1920x1080, 64 spp, depth 8
  c3  16 uneven fog jobs (volumetric path tracer) in one launch sequence
  wm  Whitted integrator on the material scene (mirror, glass, metal), 640x480, 16 spp, depth 5
  ao  ambient occlusion (16 AO rays per sample) on the Cornell box, 256x256, 16 spp

  python ray-based-rendering_amd/bench_paths.py [--only c0,c2,c3,wm,ao] [--cpu-seconds 10] [--no-cpu]

Path samples/s = pixels * spp / GPU time (HIP events around the whole render).  The CPU
figure is the reference's own Integrator::render (compiled from its sources, oracle/_ref)
on every host core over a bounded sample of the same workload, as stated in "sample".
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
sys.path.insert(0, PKG)

import rbr_amd as R  # noqa: E402
import scenes as S  # noqa: E402


def cpu_reference(desc, spp, depth, integrator, seconds, cam=None):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O

    if not O.ref_available():
        return None
    ref = O.RefRender(desc)
    cam = cam or desc.camera
    # Probe one sample per pixel, then size the sample to about `seconds` of CPU work.
    _, sec = ref.render(1, depth, integrator, camera=cam)
    paths = cam.width * cam.height
    n_spp = max(1, min(spp, int(seconds / max(sec, 1e-6))))
    if n_spp > 1:
        _, sec = ref.render(n_spp, depth, integrator, camera=cam)
    return {"value": paths * n_spp / sec / 1e6, "unit": "Mpaths/s", "cores": O.max_threads(), "kind": "reference",
            "sample": f"{cam.width}x{cam.height} x {n_spp} spp of the same scene ({sec:.1f} s)"}


def shrink(cam, w, h):
    """The same pinhole camera at a smaller film (focal length scaled with the height)."""
    s = h / cam.height
    return S.Camera(cam.position, cam.forward, cam.right, cam.up, cam.fx * s, cam.fy * s, 0.5 * w, 0.5 * h, w, h)


def run(ctx, name, jobs, cpu, seconds, tuning=None):
    scs = [R.scene_from_desc(ctx, j["desc"]) for j in jobs]
    rj = [dict(scene=s, spp=j["spp"], depth=j["depth"], seed=j.get("seed", 1), integrator=j.get("integrator", R.PATH))
          for s, j in zip(scs, jobs)]
    R.render(ctx, rj)   # warm-up (allocations, code objects)
    t0 = time.perf_counter()
    films = R.render(ctx, rj)
    wall = time.perf_counter() - t0
    ms = R.last_render_ms(ctx)
    paths = sum(j["desc"].camera.width * j["desc"].camera.height * j["spp"] for j in jobs)
    rec = {"workload": name, "value": paths / ms / 1e3, "unit": "Mpaths/s (path samples/s / 1e6)",
           "gpu_ms": ms, "wall_s": wall, "paths": paths, "jobs": len(jobs),
           "mean_radiance": [float(sum(f.mean() for f in films) / len(films))], "cpu_baseline": None,
           "tuning": tuning or "library defaults"}
    if cpu:
        j0 = jobs[0]
        cam = j0["desc"].camera
        if cam.width * cam.height > 400_000:
            cam = shrink(cam, 480, 270)
        rec["cpu_baseline"] = cpu_reference(j0["desc"], j0["spp"], j0["depth"], j0.get("integrator", 0), seconds, cam)
    print(json.dumps(rec), flush=True)
    for s in scs:
        s.close()


WORKLOADS = {
    "c0": "c0 cornell 256x256 16spp depth 8",
    "c2": "c2 terrain 1920x1080 64spp depth 8",
    "c3": "c3 16 fog jobs (volumetric)",
    "wm": "whitted material scene 640x480 16spp depth 5",
    "ao": "ao cornell 256x256 16spp (16 AO rays each)",
}


def workload_jobs(key: str) -> list:
    """The path-tracing workloads as job lists (scene descriptions + settings)."""
    if key == "c0":
        return [dict(desc=S.cornell_c0(), spp=16, depth=8)]
    if key == "c2":
        return [dict(desc=S.terrain_c2(), spp=64, depth=8)]
    if key == "c3":
        return [dict(desc=j["scene"], spp=j["spp"], depth=j["depth"], seed=j["seed"], integrator=R.VOLPATH)
                for j in S.fog_jobs_c3()]
    if key == "wm":
        return [dict(desc=S.materials_scene(640, 480), spp=16, depth=5, integrator=R.WHITTED)]
    if key == "ao":
        return [dict(desc=S.cornell_c0(), spp=16, depth=1, integrator=R.AO)]
    raise KeyError(key)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default="c0,c2,c3")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tuning", action="store_true", help="library defaults instead of tuning/render.json")
    args = ap.parse_args()
    want = set(args.only.split(","))
    ctx = R.Context(0)
    cpu = not args.no_cpu
    import autotune  # noqa: E402  (tuning cache reader)
    for key in WORKLOADS:
        if key not in want:
            continue
        cfg = None if args.no_tuning else autotune.render_config(key)
        autotune.apply_render(ctx, cfg)
        run(ctx, WORKLOADS[key], workload_jobs(key), cpu, args.cpu_seconds, tuning=cfg)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
