"""Path-tracing throughput on the path-tracing workloads, one JSON record each:

  c0  Cornell 555, 256x256, 16 spp, depth 8, NEE + MIS
  c2  terrain (524,288 triangles) + 64 spheres, 1920x1080, 64 spp, depth 8
  c3  16 uneven fog jobs (volumetric path tracer) in one launch sequence
  wm  Whitted integrator on the material scene (mirror, glass, metal), 640x480, 16 spp, depth 5
  ao  ambient occlusion (16 AO rays per sample) on the Cornell box, 256x256, 16 spp

  python ray-based-rendering_amd/bench_paths.py [--only c0,c2,c3,wm,ao] [--cpu-seconds 5] [--no-cpu]

* value: path samples/s = pixels * spp / GPU time of the whole render (HIP events on the
  context stream around every launch of the render; scenes and BVHs resident).
* roofline (c0, c2, c3): algorithmic HBM bytes of the render / its GPU time, against 8 TB/s.
  Per extension segment (camera ray or continuation): ray 32 + hit 8 + path state read and
  written 2 x 48 = 136 B; per shadow ray: ray 32 + occlusion 1 = 33 B; plus, for every ray,
  its traversal: node visits x node bytes + primitive tests x 36 B (the triangle's three
  vertices).  Ray counts and the mean node visits / primitive tests per ray come from a
  census render of the same jobs (rbr_set_render_census: a strided sample of every bounce's
  queues traced again by the host traversal of the same BVH).
* cpu_baseline: the reference's own Integrator::render (compiled from its sources,
  oracle/_ref) on every host core over a bounded sample of the same workload, as "sample"
  states (c3: all 16 jobs one after the other).
* relmse_vs_oracle: the GPU image against the FP64 CPU oracle (oracle/orc_render.cpp, the
  reference's integrators restated) -- c0: the committed converged 4096-spp image; c2 and
  c3: a stratified pixel subset rendered by the oracle at high spp -- next to the Monte Carlo
  bound of an unbiased image at the GPU's spp (ratio ~1 when unbiased).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
sys.path.insert(0, PKG)

import rbr_amd as R  # noqa: E402
import scenes as S  # noqa: E402

HBM_PEAK_GBS = 8000.0
SEG_BYTES = 32 + 8 + 2 * 48    # extension segment: ray, hit, path state read + written
SHADOW_BYTES = 32 + 1          # shadow ray: ray, occlusion flag
PRIM_BYTES = 36                # a triangle test reads three vertices
EPS = 1e-2
C0_GOLDEN = os.path.join(ROOT, "tests", "golden", "c0_oracle_256_4096spp.npz")


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    return O


def shrink(cam, w, h):
    """The same pinhole camera at a smaller film (focal length scaled with the height)."""
    s = h / cam.height
    return S.Camera(cam.position, cam.forward, cam.right, cam.up, cam.fx * s, cam.fy * s, 0.5 * w, 0.5 * h, w, h)


def cpu_reference(jobs: list, seconds: float, port: bool = False, film_scale: float = 1.0) -> dict | None:
    """The reference's Integrator::render on all host cores, every job of the workload one
    after the other (films above 400k pixels at 480x270 of the same camera, or scaled by
    film_scale), spp scaled so that the whole sample takes about `seconds`.  port: the
    oracle's restatement instead of the reference build (its integrators give the reference's
    images bit for bit, and its exponential-fog march is closed-form: the reference's march
    re-uses one random number and, for a small one, advances by less than an ulp of t and
    never returns)."""
    O = _oracle()
    if not port and not O.ref_available():
        return None
    refs = []
    for j in jobs:
        cam = j["desc"].camera
        if film_scale != 1.0:
            cam = shrink(cam, max(8, int(cam.width * film_scale)), max(8, int(cam.height * film_scale)))
        elif cam.width * cam.height > 400_000:
            cam = shrink(cam, 480, 270)
        refs.append((O.OracleRender(j["desc"]) if port else O.RefRender(j["desc"]), cam, j))

    def run(ref, cam, j, n):
        integ = j.get("integrator", 0)
        if port:
            t0 = time.perf_counter()
            ref.moments(n, j["depth"], seed=0xBA5E, integrator=integ, camera=cam)
            return time.perf_counter() - t0
        return ref.render(n, j["depth"], integ, camera=cam)[1]

    probe = sum(run(ref, cam, j, 1) for ref, cam, j in refs)   # one sample per pixel of every job
    k = max(1, int(seconds / max(probe, 1e-6)))
    paths, sec = 0, 0.0
    for ref, cam, j in refs:
        n = min(k, j["spp"] * 8)
        sec += run(ref, cam, j, n)
        paths += cam.width * cam.height * n
    films = "; ".join(sorted({f"{c.width}x{c.height}" for _, c, _ in refs}))[:120]
    return {"value": paths / sec / 1e6, "unit": "Mpaths/s", "cores": O.max_threads(), "kind": "port" if port else "reference",
            "sample": f"{len(refs)} job(s) one after the other, up to {k} spp each (films {films}); "
                      f"{paths} paths in {sec:.1f} s" + ("; the oracle's restated integrator" if port else "")}


def relmse(x, ref):
    x = np.asarray(x, np.float64).reshape(-1, 3)
    ref = np.asarray(ref, np.float64).reshape(-1, 3)
    return float(((x - ref) ** 2 / (ref * ref + EPS)).mean())


def stratified(w, h, n, seed):
    rng = np.random.default_rng(seed)
    g = max(1, int(np.sqrt(n)))
    xs = (np.arange(g)[None, :] * w) // g + rng.integers(0, max(1, w // g), (g, g))
    ys = (np.arange(g)[:, None] * h) // g + rng.integers(0, max(1, h // g), (g, g))
    return np.unique((np.minimum(ys, h - 1) * w + np.minimum(xs, w - 1)).reshape(-1)).astype(np.uint32)


def oracle_error(key: str, jobs: list, films: list, pixels: int, oracle_spp: int) -> dict | None:
    """relMSE of the GPU films against the FP64 oracle, with the Monte Carlo bound."""
    if key == "c0":
        if not os.path.exists(C0_GOLDEN):
            return None
        f = np.load(C0_GOLDEN)
        ref, var, n_ref = f["mean"], f["var"], int(f["spp"])
        n = jobs[0]["spp"]
        e = relmse(films[0], ref)
        b = float((var.reshape(-1, 3).astype(np.float64) * (1 / n + 1 / n_ref)
                   / (ref.reshape(-1, 3).astype(np.float64) ** 2 + EPS)).mean())
        return {"relmse": e, "mc_bound": b, "ratio": e / b, "oracle": f"committed {n_ref}-spp FP64 image, all pixels"}
    O = _oracle()
    groups: dict = {}
    per_job = max(16, pixels // len(jobs))
    for k, (j, film) in enumerate(zip(jobs, films)):
        cam = j["desc"].camera
        px = stratified(cam.width, cam.height, per_job, 0xE55 + k)
        orc = O.OracleRender(j["desc"])
        integ = j.get("integrator", 0)
        s1, s2 = orc.moments(oracle_spp, j["depth"], seed=0xE55, integrator=integ, pixels=px)
        m = s1[px] / oracle_spp
        v = np.maximum(s2[px] / oracle_spp - m * m, 0.0) * oracle_spp / (oracle_spp - 1)
        n = j["spp"]
        b = float((v * (1 / n + 1 / oracle_spp) / (m * m + EPS)).mean())
        e = relmse(film.reshape(-1, 3)[px], m)
        rec = {"relmse": e, "mc_bound": b, "pixels": len(px)}
        if key == "c3":
            # Control: the oracle's own image at the job's spp against the same reference.  The
            # reference's homogeneous scatter weight sigma_s*Tr/avg^2 (medium.cpp:22-27) makes
            # radiance so heavy-tailed that a relMSE at 8-32 spp often lies far above the bound
            # estimated from 256 samples, for the oracle itself as for the GPU.
            c1, _ = orc.moments(n, j["depth"], seed=0xC0C0 + k, integrator=integ, pixels=px)
            rec["oracle_self_relmse"] = relmse(c1[px] / n, m)
        gname = "all" if key != "c3" else ("homogeneous" if j["desc"].medium["type"] == 1 else "exponential")
        groups.setdefault(gname, []).append(rec)
    out = {"oracle": f"FP64 CPU oracle, stratified pixels at {oracle_spp} spp"}
    for gname, recs in groups.items():
        w = np.array([r["pixels"] for r in recs], np.float64)
        e = float(np.dot(w, [r["relmse"] for r in recs]) / w.sum())
        b = float(np.dot(w, [r["mc_bound"] for r in recs]) / w.sum())
        g = {"relmse": e, "mc_bound": b, "ratio": e / b, "pixels": int(w.sum())}
        if "oracle_self_relmse" in recs[0]:
            c = float(np.dot(w, [r["oracle_self_relmse"] for r in recs]) / w.sum())
            g.update(oracle_self_relmse=c, oracle_self_ratio=c / b)
        if gname == "all":
            out.update(g)
        else:
            out[gname] = g
    return out


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


def measure(ctx, key: str, reps: int = 3, census: int = 4096, cpu_seconds: float = 5.0, error: bool = True,
            tuning: bool = True, log=None) -> dict:
    """Times one workload (mean GPU ms of `reps` renders after a warm-up) and returns its record."""
    import autotune

    jobs = workload_jobs(key)
    cfg = autotune.render_config(key) if tuning else None
    autotune.apply_render(ctx, cfg)
    scs = [R.scene_from_desc(ctx, j["desc"]) for j in jobs]
    rj = [dict(scene=s, spp=j["spp"], depth=j["depth"], seed=j.get("seed", 1), integrator=j.get("integrator", R.PATH))
          for s, j in zip(scs, jobs)]
    R.render(ctx, rj)   # warm-up (allocations, code objects, round graphs)
    times = []
    films = None
    for _ in range(reps):
        films = R.render(ctx, rj)
        times.append(R.last_render_ms(ctx))
    ms = float(np.mean(times))
    paths = sum(j["desc"].camera.width * j["desc"].camera.height * j["spp"] for j in jobs)
    rec = {"workload": WORKLOADS[key], "value": paths / ms / 1e3, "unit": "Mpaths/s (path samples/s / 1e6)",
           "gpu_ms": ms, "gpu_ms_reps": times, "paths": paths, "jobs": len(jobs),
           "tuning": cfg or "library defaults"}
    if census and key in ("c0", "c2", "c3"):
        R.set_render_census(ctx, census)
        try:
            R.render(ctx, rj)
            st = R.last_render_stats(ctx)
        finally:
            R.set_render_census(ctx, 0)
        node_b = scs[0].stats()["node_bytes"] / max(1, scs[0].stats()["node_count"])
        ext_b = SEG_BYTES + st["ext_nodes_per_ray"] * node_b + st["ext_prims_per_ray"] * PRIM_BYTES
        sh_b = SHADOW_BYTES + st["shadow_nodes_per_ray"] * node_b + st["shadow_prims_per_ray"] * PRIM_BYTES
        total = st["extension_rays"] * ext_b + st["shadow_rays"] * sh_b
        achieved = total / (ms * 1e-3) / 1e9
        rec["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                           "bytes_per_path": total / paths, "bytes_per_extension_ray": ext_b,
                           "bytes_per_shadow_ray": sh_b, "node_bytes": node_b, "census": st}
    if error:
        if log:
            log(f"bench: {key} oracle error")
        rec["relmse_vs_oracle"] = (oracle_error(key, jobs, films, 1024, 256) if key == "c3"
                                   else oracle_error(key, jobs, films, 4096, 1024))
    if cpu_seconds > 0:
        if log:
            log(f"bench: {key} CPU baseline")
        # C3: all 16 jobs at a quarter of their film size, by the oracle's restatement (the
        # reference's exponential-fog march can stall forever on a small random number).
        rec["cpu_baseline"] = (cpu_reference(jobs, cpu_seconds, port=True, film_scale=0.25) if key == "c3"
                               else cpu_reference(jobs, cpu_seconds))
    else:
        rec["cpu_baseline"] = None
    for s in scs:
        s.close()
    autotune.apply_render(ctx, None)
    return rec


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default="c0,c2,c3")
    ap.add_argument("--cpu-seconds", type=float, default=5.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--no-tuning", action="store_true", help="library defaults instead of tuning/render.json")
    args = ap.parse_args()
    want = args.only.split(",")
    ctx = R.Context(0)
    for key in WORKLOADS:
        if key not in want:
            continue
        rec = measure(ctx, key, cpu_seconds=0.0 if args.no_cpu else args.cpu_seconds, error=not args.no_error,
                      tuning=not args.no_tuning, log=lambda m: print(m, file=sys.stderr, flush=True))
        print(json.dumps(rec), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
