"""Condenses rocprofv3 CSV output into the short summaries kept under profiles/.

  python ray-based-rendering_amd/profile_summary.py --stats DIR/x_kernel_stats.csv \
      [--fetch DIR/x_counter_collection.csv] [--write DIR/x_counter_collection.csv] \
      [--l2 DIR/x_counter_collection.csv] [--sq CSV] [--mem CSV] [--rays N] --name NAME \
      --command "..." [--kernel SUBSTR] [--copy-bytes 268435456]

Writes profiles/NAME.md (kernel table, counters) and, for traversal, profiles/NAME_pmc.json
with the calibrated HBM bytes per launch: FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are
scaled by the ratio of the known bytes of the calibration copy kernel to what the same
counter reads for it.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def counters(path: str) -> dict:
    agg = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        agg[(r["Kernel_Name"], r["Counter_Name"])].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in agg.items()}


def short(name: str) -> str:
    return name.replace("void rbr::(anonymous namespace)::", "").replace("(rbr::TraceArgs, rbr::PersistArgs)", "")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--stats", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--l2")
    ap.add_argument("--sq", help="counter CSV of an SQ_* run (instruction counts, wave cycles)")
    ap.add_argument("--rays", type=float, default=0.0, help="rays per launch of --kernel (per-64-ray figures)")
    ap.add_argument("--name", required=True)
    ap.add_argument("--command", required=True)
    ap.add_argument("--kernel", default="traceKernelPersistent<4, true, 8, false")
    ap.add_argument("--copy-bytes", type=float, default=256 * 1024 * 1024)
    ap.add_argument("--mem", help="counter CSV of a TA/TD/TCP/TCC/GRBM run (vector memory pipeline)")
    ap.add_argument("--xcds", type=float, default=8.0, help="GRBM_GUI_ACTIVE is summed over this many XCDs")
    ap.add_argument("--simds", type=float, default=1024.0, help="SIMDs of the device (4 per CU)")
    args = ap.parse_args()
    rows = list(csv.DictReader(open(args.stats)))
    lines = [f"# {args.name}", "", "Command (from the repository root, on one MI355X):", "", "```", args.command, "```", "",
             "| kernel | calls | avg us | min us | max us | % time |", "|---|---|---|---|---|---|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"]))[:12]:
        lines.append(f"| `{short(r['Name'])[:90]}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                     f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} |")
    out = {"name": args.name, "command": args.command}
    tgt = [r for r in rows if args.kernel in r["Name"]]
    if tgt:
        out["kernel"] = short(tgt[0]["Name"])
        out["avg_us"] = float(tgt[0]["AverageNs"]) / 1e3
    if args.fetch and args.write:
        f = counters(args.fetch)
        w = counters(args.write)
        cf = [v for (k, c), v in f.items() if "calibCopy" in k]
        cw = [v for (k, c), v in w.items() if "calibCopy" in k]
        kib = args.copy_bytes / 1024.0
        sf = kib / cf[0] if cf else 1.0
        sw = kib / cw[0] if cw else 1.0
        lines += ["", "Byte counters (KiB per dispatch, then calibrated with the 256 MiB copy kernel: "
                  f"FETCH x{sf:.3f}, WRITE x{sw:.3f}):", "", "| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM MB/launch |",
                  "|---|---|---|---|"]
        for (k, c), v in sorted(f.items()):
            wv = w.get((k, "WRITE_SIZE"), 0.0)
            mb = (v * sf + wv * sw) * 1024 / 1e6
            lines.append(f"| `{short(k)[:90]}` | {v:.0f} | {wv:.0f} | {mb:.1f} |")
            if args.kernel in k:
                out["fetch_kib"] = v
                out["write_kib"] = wv
                out["fetch_scale"] = sf
                out["write_scale"] = sw
                out["traffic_bytes_per_launch"] = (v * sf + wv * sw) * 1024
    if args.l2:
        l2 = counters(args.l2)
        lines += ["", "| kernel | TCC_HIT | TCC_MISS | L2 hit rate |", "|---|---|---|---|"]
        kernels = sorted({k for k, _ in l2})
        for k in kernels:
            h, m = l2.get((k, "TCC_HIT_sum"), 0.0), l2.get((k, "TCC_MISS_sum"), 0.0)
            if h + m:
                lines.append(f"| `{short(k)[:90]}` | {h:.0f} | {m:.0f} | {h / (h + m):.3f} |")
                if args.kernel in k:
                    out["l2_hit_rate"] = h / (h + m)
    if args.sq:
        sq = counters(args.sq)
        names = sorted({c for k, c in sq if args.kernel in k})
        kern = [k for k, _ in sq if args.kernel in k]
        if kern:
            k0 = kern[0]
            lines += ["", f"SQ counters of `{short(k0)[:90]}` (per launch):", "", "| counter | value |", "|---|---|"]
            for c in names:
                lines.append(f"| {c} | {sq[(k0, c)]:.4g} |")
                out[c] = sq[(k0, c)]
            if args.rays and (k0, "SQ_INSTS_VALU") in sq:
                per = 64.0 / args.rays
                out["valu_per_64_rays"] = sq[(k0, "SQ_INSTS_VALU")] * per
                out["salu_per_64_rays"] = sq.get((k0, "SQ_INSTS_SALU"), 0.0) * per
                cyc = max(sq.get((k0, "SQ_WAVE_CYCLES"), 0.0), 1.0)   # wave quad-cycles
                txt = [f"VALU wave-instructions per 64 rays: {out['valu_per_64_rays']:.0f} (SALU "
                       f"{out['salu_per_64_rays']:.0f})"]
                for c, what in (("SQ_WAIT_ANY", "waiting (dependencies, memory)"),
                                ("SQ_WAIT_INST_ANY", "ready but waiting for issue"),
                                ("SQ_ACTIVE_INST_ANY", "issuing")):
                    if (k0, c) in sq:
                        out[c + "_frac"] = sq[(k0, c)] / cyc
                        txt.append(f"{what}: {out[c + '_frac']:.0%} of wave residency")
                grbm = sq.get((k0, "GRBM_GUI_ACTIVE"), 0.0)
                if grbm and (k0, "SQ_ACTIVE_INST_VALU2") in sq:
                    quads = grbm / args.xcds / 4.0 * args.simds   # SIMD quad-cycles of the launch
                    issue = sq[(k0, "SQ_INSTS_VALU")] - sq[(k0, "SQ_ACTIVE_INST_VALU2")]
                    out["valu_issue_frac"] = issue / quads
                    txt.append(f"quad-cycles of a SIMD with a VALU issue: {out['valu_issue_frac']:.0%} "
                               "(two VALU issued together counted once)")
                lines += ["", "; ".join(txt) + "."]
    if args.mem:
        mem = counters(args.mem)
        kern = [k for k, _ in mem if args.kernel in k]
        if kern:
            k0 = kern[0]
            g = lambda c: mem.get((k0, c), 0.0)
            cyc = g("GRBM_GUI_ACTIVE") / args.xcds   # clock cycles of the launch
            cus = args.simds / 4
            lines += ["", f"Vector memory pipeline of `{short(k0)[:90]}` (per launch; per-CU figures over "
                      f"{cyc:.3g} cycles):", "", "| counter | value | per CU, of the launch |", "|---|---|---|"]
            for c in ("TA_TA_BUSY_sum", "TA_ADDR_STALLED_BY_TC_CYCLES_sum", "TD_TD_BUSY_sum", "TD_TC_STALL_sum",
                      "TCP_PENDING_STALL_CYCLES_sum"):
                if (k0, c) in mem:
                    out[c + "_frac"] = g(c) / cus / max(cyc, 1.0)
                    lines.append(f"| {c} | {g(c):.4g} | {out[c + '_frac']:.0%} |")
            req, lat = g("TCP_TCC_READ_REQ_sum"), g("TCP_TCC_READ_REQ_LATENCY_sum")
            acc = g("TCP_TOTAL_CACHE_ACCESSES_sum")
            if req:
                out["l1_to_l2_latency_cycles"] = lat / req
                out["l1_misses_in_flight_per_cu"] = lat / cus / max(cyc, 1.0)
                lines.append(f"| TCP_TCC_READ_REQ_sum | {req:.4g} | L1->L2 latency {lat / req:.0f} cycles, "
                             f"{out['l1_misses_in_flight_per_cu']:.0f} in flight |")
                if args.rays:
                    out["l1_misses_per_ray"] = req / args.rays
            if acc:
                out["l1_hit_rate"] = 1.0 - req / acc
                lines.append(f"| TCP_TOTAL_CACHE_ACCESSES_sum | {acc:.4g} | L1 hit rate {out['l1_hit_rate']:.2f} |")
            h, m = g("TCC_HIT_sum"), g("TCC_MISS_sum")
            if h + m:
                out["l2_hit_rate_mem_run"] = h / (h + m)
                lines.append(f"| TCC_HIT_sum / TCC_MISS_sum | {h:.4g} / {m:.4g} | L2 hit rate {h / (h + m):.2f} |")
    os.makedirs(os.path.join(ROOT, "profiles"), exist_ok=True)
    with open(os.path.join(ROOT, "profiles", args.name + ".md"), "w") as fh:
        fh.write("\n".join(lines) + "\n")
    if "traffic_bytes_per_launch" in out:
        with open(os.path.join(ROOT, "profiles", args.name + "_pmc.json"), "w") as fh:
            json.dump(out, fh, indent=1)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
