"""Profiling helpers for the traversal kernels (run under rocprofv3).

  python3 raptor_amd/profiling.py copy [--mib 1024]
      one device-to-device copy of a known size: the calibration for the
      FETCH_SIZE / WRITE_SIZE counters (KiB per dispatch).
  python3 raptor_amd/profiling.py summarize --stats DIR --fetch DIR --write DIR
                                          --copy-fetch DIR --copy-write DIR
                                          --bench JSON --out PREFIX
      reads rocprofv3 CSV output and writes PREFIX.md (kernel summary) and
      PREFIX.json (HBM bytes per launch of the dominant kernel), which
      bench.py reports as roofline.traffic.
  python3 raptor_amd/profiling.py kernels --stats DIR --match 'SUBSTR[|SUBSTR]' 
                                          --title TEXT --command TEXT --out FILE
      writes FILE (markdown): the rocprofv3 kernel statistics of the kernels
      whose names contain one of the substrings.
  python3 raptor_amd/profiling.py counters --dirs DIR[,DIR...] --title TEXT
                                           --command TEXT --out FILE
      writes FILE (markdown): per trace stage of the frame, the averages of
      VALUBusy, VALUUtilization, OccupancyPercent, MeanOccupancyPerActiveCU,
      the L2 hit rate (TCC_HIT_sum, TCC_MISS_sum) and instruction counts
      collected by one rocprofv3 --pmc pass per directory.
"""
import argparse
import csv
import glob
import json
import os
import sys


def cmd_copy(args):
    import torch

    n = args.mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4}))


def _rows(d, name):
    files = glob.glob(os.path.join(d, "**", name), recursive=True)
    if not files:
        raise SystemExit(f"no {name} under {d}")
    with open(files[0]) as f:
        return list(csv.DictReader(f))


def _counter(d, counter, match, largest=False):
    vals = [float(r["Counter_Value"]) for r in _rows(d, "*counter_collection.csv")
            if r["Counter_Name"] == counter and match(r["Kernel_Name"])]
    if not vals:
        return None
    # The calibration copy is the largest copy dispatch of its run.
    return max(vals) if largest else sum(vals) / len(vals)


def cmd_summarize(args):
    with open(args.bench) as f:
        bench = json.loads(f.read().strip().splitlines()[-1])
    kernel = bench["roofline"]["kernel"].split(" (")[0]          # e.g. k_trace<4, 2, false, 8>
    stage = bench["roofline"]["kernel"].split("(")[-1].rstrip(")")
    match = lambda name: kernel.replace(" ", "") in name.replace(" ", "")
    is_copy = lambda name: "copy" in name.lower() or "elementwise" in name.lower()
    fetch = _counter(args.fetch, "FETCH_SIZE", match)
    write = _counter(args.write, "WRITE_SIZE", match)
    copy_bytes = args.copy_bytes
    cf = _counter(args.copy_fetch, "FETCH_SIZE", is_copy, largest=True)
    cw = _counter(args.copy_write, "WRITE_SIZE", is_copy, largest=True)
    # Counters are KiB per dispatch; the copy moves copy_bytes each way.
    fetch_scale = copy_bytes / (cf * 1024.0)
    write_scale = copy_bytes / (cw * 1024.0)
    traffic = fetch * 1024.0 * fetch_scale + write * 1024.0 * write_scale
    stats = _rows(args.stats, "*kernel_stats.csv")
    lines = [f"# Kernel summary: {bench['config']['workload']}", "",
             "Command (from the repository root, TMPDIR=/tmp):", "",
             "    " + args.command, "",
             "| kernel | calls | avg us | total % |", "|---|---|---|---|"]
    for r in stats[:16]:
        name = r["Name"].replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "")
        lines.append(f"| `{name}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | {float(r['Percentage']):.2f} |")
    dom = [r for r in stats if match(r["Name"])]
    dom_us = float(dom[0]["AverageNs"]) / 1e3 if dom else None
    lines += ["", f"Dominant kernel `{kernel}` ({stage}): rocprof average {dom_us:.2f} us, bench HIP-event "
              f"average {bench['roofline']['launch_ms'] * 1e3:.2f} us.", "",
              "HBM traffic per launch (rocprofv3 --pmc, separate runs):", "",
              f"* FETCH_SIZE {fetch:.1f} KiB x calibration {fetch_scale:.4f} "
              f"(copy of {copy_bytes} B read {cf:.1f} KiB)",
              f"* WRITE_SIZE {write:.1f} KiB x calibration {write_scale:.4f} "
              f"(copy of {copy_bytes} B wrote {cw:.1f} KiB)",
              f"* traffic = {traffic / 1e6:.2f} MB per launch over {bench['roofline']['launch_ms'] * 1e3:.1f} us "
              f"= {traffic / (bench['roofline']['launch_ms'] * 1e-3) / 1e9:.0f} GB/s, "
              f"{traffic / (bench['roofline']['launch_ms'] * 1e-3) / 1e9 / bench['roofline']['peak']:.3f} of "
              f"{bench['roofline']['peak']:.0f} GB/s",
              f"* per-visit demand model (every node and triangle fetch priced as HBM bytes, no cache reuse): "
              f"{bench['roofline']['visit_model']['bytes_per_launch'] / 1e6:.2f} MB per launch",
              f"* compulsory bytes (BVH + triangles read once, rays and hits in and out): "
              f"{bench['roofline']['compulsory']['bytes_per_launch'] / 1e6:.2f} MB per launch"]
    with open(args.out + ".md", "w") as f:
        f.write("\n".join(lines) + "\n")
    out = {"kernel": kernel, "kernel_stage": stage, "width": bench["config"]["bvh_width"],
           "lds_stack": bench["config"].get("lds_stack"), "leaf_size": bench["config"]["leaf_size"],
           "sah_bins": bench["config"].get("sah_bins", 0),
           "bytes_per_launch": round(traffic), "fetch_kib": fetch, "write_kib": write,
           "fetch_scale": fetch_scale, "write_scale": write_scale, "rocprof_avg_us": dom_us}
    with open(args.out + ".json", "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


def cmd_kernels(args):
    stats = _rows(args.stats, "*kernel_stats.csv")
    keys = [k.strip() for k in args.match.split("|") if k.strip()]
    lines = [f"# Kernel summary: {args.title}", "", "Command (from the repository root, TMPDIR=/tmp):", "",
             "    " + args.command, "", "| kernel | calls | avg us | min us | max us |", "|---|---|---|---|---|"]
    for r in stats:
        name = r["Name"].replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "")
        if not any(k in name for k in keys):
            continue
        lines.append(f"| `{name}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                     f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} |")
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def _trace_kernel(name):
    """'rt::k_trace<4, 2, false, 8>' style key of a trace-kernel dispatch, or
    None for other kernels and for the visit-counting variants."""
    k = name.replace("void ", "").replace("(anonymous namespace)::", "").split("(rt::TraceArgs")[0].strip()
    if not (k.startswith("rt::k_trace<") or k.startswith("rt::k_trace_pers<")) or "true" in k:
        return None
    return k


def cmd_counters(args):
    import collections

    res = collections.defaultdict(lambda: collections.defaultdict(list))
    res_regs = {}
    for d in args.dirs.split(","):
        for r in _rows(d, "*counter_collection.csv"):
            k = _trace_kernel(r["Kernel_Name"])
            if k is None:
                continue
            res[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
            res_regs[k] = (r.get("VGPR_Count"), r.get("SGPR_Count"), r.get("LDS_Block_Size"), r.get("Scratch_Size"),
                           r.get("Workgroup_Size"))
    names = {0: "primary closest", 1: "shadow any hit", 2: "bounce closest", 3: "bounce-shadow any hit"}
    lines = [f"# Counters: {args.title}", "", "Command (one rocprofv3 pass per counter group):", "",
             "    " + args.command, "", "Averages over the frame launches of each kernel (the counting frame excluded).",
             "", "| kernel | stage | VALUBusy % | VALUUtilization % | Occupancy % | waves / active CU | L2 hit % | "
             "SQ_INSTS_VALU | SQ_INSTS_SALU | SQ_INSTS_LDS | VALU instr / wave |",
             "|---|---|---|---|---|---|---|---|---|---|---|"]
    for k in sorted(res):
        d = {c: sum(v) / len(v) for c, v in res[k].items()}
        st = int(k.split(",")[1])
        hit = 100 * d.get("TCC_HIT_sum", 0.0) / max(1.0, d.get("TCC_HIT_sum", 0.0) + d.get("TCC_MISS_sum", 0.0))
        per_wave = d.get("SQ_INSTS_VALU", 0.0) / max(1.0, d.get("SQ_WAVES", 0.0))
        lines.append(f"| `{k}` | {names.get(st, st)} | {d.get('VALUBusy', 0):.1f} | {d.get('VALUUtilization', 0):.1f} | "
                     f"{d.get('OccupancyPercent', 0):.1f} | {d.get('MeanOccupancyPerActiveCU', 0):.2f} | {hit:.1f} | "
                     f"{d.get('SQ_INSTS_VALU', 0):.3g} | {d.get('SQ_INSTS_SALU', 0):.3g} | "
                     f"{d.get('SQ_INSTS_LDS', 0):.3g} | {per_wave:.0f} |")
    lines += ["", "Resources per kernel (from the same dispatch records):", "",
              "| kernel | VGPRs | SGPRs | LDS bytes / workgroup | scratch bytes / lane | workgroup |",
              "|---|---|---|---|---|---|"]
    for k in sorted(res_regs):
        v, s, lds, scr, wg = res_regs[k]
        lines.append(f"| `{k}` | {v} | {s} | {lds} | {scr} | {wg} |")
    lines += ["", "VALUBusy: share of the kernel's time the VALUs issue. VALUUtilization: share of the 64 lanes",
              "active in the VALU instructions issued (SIMD efficiency). VALU instr / wave: SQ_INSTS_VALU / SQ_WAVES",
              "(for the block-local form a wave holds many rays over its life)."]
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--mib", type=int, default=1024)
    s = sub.add_parser("summarize")
    for k in ("stats", "fetch", "write", "copy_fetch", "copy_write", "bench", "out"):
        s.add_argument("--" + k.replace("_", "-"), dest=k, required=True)
    s.add_argument("--copy-bytes", type=int, default=1024 * (1 << 20))
    s.add_argument("--command", default="")
    k = sub.add_parser("kernels")
    for a in ("stats", "match", "title", "command", "out"):
        k.add_argument("--" + a, required=True)
    c2 = sub.add_parser("counters")
    for a in ("dirs", "title", "command", "out"):
        c2.add_argument("--" + a, required=True)
    args = ap.parse_args()
    if args.cmd == "counters":
        cmd_counters(args)
    elif args.cmd == "copy":
        cmd_copy(args)
    elif args.cmd == "kernels":
        cmd_kernels(args)
    else:
        cmd_summarize(args)


if __name__ == "__main__":
    sys.exit(main())
