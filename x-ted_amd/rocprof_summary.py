"""Summarises rocprofv3 output of bench.py runs into the short text kept
under profiles/.

  kernels:  python x-ted_amd/rocprof_summary.py kernels <kernel_trace.csv> [--bench bench.json]
            per kernel: dispatches, average and total duration; with the
            bench JSON line, the algorithmic bytes per launch of each class
            (4 bytes of TD per interior cell) and the rate they imply.
  counters: python x-ted_amd/rocprof_summary.py counters <sq.csv> <lds.csv> [--occ occ.csv] [--valu valu.csv]
            per kernel from the SQ counter runs: waves, VALU / LDS / VMEM
            instructions per wave, VALU-active and waiting shares of wave
            cycles, LDS bank-conflict share of LDS-active cycles; with the
            derived-counter runs, mean occupancy per CU, VALU busy and SIMD
            utilisation.
  resources: python x-ted_amd/rocprof_summary.py resources [csrc/forest.hip]
            registers, scratch, static LDS and waves per SIMD of every
            kernel as the compiler reports them for gfx950.
  traffic:  python x-ted_amd/rocprof_summary.py traffic KERNEL <fetch.csv> <write.csv>
                <copy_fetch.csv> <copy_write.csv> --copy-bytes N
            FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of the dispatches of
            KERNEL, each scaled by what the counter reads per KiB of a plain
            copy of N bytes, and their sum as HBM bytes per launch.
"""
import argparse
import csv
import json
import os
import sys
from collections import defaultdict


def short(name):
    """'rowscan_w4_c4', 'xl_bw1024', 'rowscan_w4_c4_cost' (the variant under a
    cost model), 'leaf_kernel', ... from a demangled kernel name."""
    base = name.split("(")[0].replace("void ", "").replace("xted::", "")
    if base.startswith("rowscan_kernel<"):
        parts = [p.strip() for p in base[len("rowscan_kernel<"):-1].split(",")]
        w, c = int(parts[0]), int(parts[1])
        cost = "_cost" if len(parts) > 3 and parts[3] == "true" else ""
        if len(parts) > 2 and parts[2] == "true":
            return f"xl_bw{64 * w * c}{cost}"
        return f"rowscan_w{w}_c{c}{cost}"
    if base.startswith("tiny_kernel<"):
        parts = [p.strip() for p in base[len("tiny_kernel<"):-1].split(",")]
        return f"tiny_s{parts[0]}" + ("_cost" if parts[-1] == "true" else "")
    return base


# Resource columns of a rocprofv3 kernel trace (names differ between
# versions: the first one present is used).
RESOURCES = [("vgpr", ("Arch_VGPR_Count", "VGPR_Count")),
             ("agpr", ("Accum_VGPR_Count",)),
             ("sgpr", ("SGPR_Count",)),
             ("lds_B", ("LDS_Block_Size", "Group_Segment_Size")),
             ("scratch_B", ("Scratch_Size", "Private_Segment_Size"))]


def kernels(path, bench=None):
    """Per kernel: dispatches, average and total duration, the registers,
    LDS and scratch of its dispatches (largest over them: LDS is dynamic),
    and with the bench line the algorithmic bytes per launch. The bench
    line's class cells come from its event-timed steps: `classes_1_stream`
    when present (the run traced with one level stream), else `classes`,
    plus the 20,000-node pair's column-blocked classes (`xl_pair.classes`)."""
    per = defaultdict(list)
    res = defaultdict(dict)
    with open(path) as f:
        for r in csv.DictReader(f):
            k = short(r["Kernel_Name"])
            per[k].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
            for name, cols in RESOURCES:
                for col in cols:
                    if r.get(col) not in (None, ""):
                        res[k][name] = max(res[k].get(name, 0), int(float(r[col])))
                        break
    classes = {}
    if bench:
        with open(bench) as f:
            for line in f:
                if line.startswith("{"):
                    j = json.loads(line)
                    if "roofline" in j and "classes" not in j and "classes_1_stream" not in j:
                        continue
                    classes = dict(j.get("classes_1_stream") or j.get("classes", {}))
                    cm = j.get("cost_model") or {}
                    classes.update({k + "_cost": v for k, v in cm.get("classes", {}).items()})
                    classes.update((j.get("xl_pair") or {}).get("classes", {}))
    rows = sorted(per.items(), key=lambda kv: -sum(kv[1]))
    head = f"{'kernel':<26}{'dispatches':>11}{'avg_us':>11}{'total_ms':>11}"
    rcols = [name for name, _ in RESOURCES if any(name in v for v in res.values())]
    head += "".join(f"{c:>10}" for c in rcols)
    out = [head + f"{'alg_MB/launch':>15}{'alg_GB/s':>10}"]
    for k, d in rows:
        avg = sum(d) / len(d) / 1e3
        line = f"{k:<26}{len(d):>11}{avg:>11.1f}{sum(d) / 1e6:>11.2f}"
        line += "".join(f"{res[k].get(c, ''):>10}" for c in rcols)
        c = classes.get(k)
        if c and (k.startswith("rowscan") or k.startswith("xl_bw")):
            b = 4 * c["cells"] / c["launches"]
            line += f"{b / 1e6:>15.1f}{b / (avg * 1e3):>10.1f}"
        out.append(line)
    return "\n".join(out)


def resources(src, hipcc="/opt/rocm/bin/hipcc", arch="gfx950"):
    """Registers, scratch, LDS and occupancy of every kernel of a HIP source
    as the compiler reports them (-Rpass-analysis=kernel-resource-usage),
    compiled with the library's device flags. rocprofv3's kernel-trace
    register columns count allocation granules and its LDS column leaves out
    dynamic LDS; these are the compiler's own counts."""
    import re
    import subprocess
    here = os.path.dirname(os.path.abspath(__file__))
    cmd = [hipcc, f"--offload-arch={arch}", "-O3", "-std=c++17", "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops", "-mllvm", "-pragma-unroll-threshold=1000000",
           "--cuda-device-only", "-c", src, "-o", os.devnull,
           "-Rpass-analysis=kernel-resource-usage"]
    r = subprocess.run(cmd, cwd=here, capture_output=True, text=True, check=True)
    rows, cur = [], None
    for line in r.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            cur = {"name": m.group(1)}
            rows.append(cur)
            continue
        for key, pat in (("vgpr", r" VGPRs: (\d+)"), ("agpr", r"AGPRs: (\d+)"),
                         ("sgpr", r"TotalSGPRs: (\d+)"), ("scratch_B", r"ScratchSize \[bytes/lane\]: (\d+)"),
                         ("occupancy", r"Occupancy \[waves/SIMD\]: (\d+)"),
                         ("lds_static_B", r"LDS Size \[bytes/block\]: (\d+)")):
            m = re.search(pat, line)
            if m and cur is not None:
                cur[key] = int(m.group(1))

    def pretty(mangled):
        m = re.search(r"rowscan_kernelILi(\d+)ELi(\d+)ELb([01])ELb([01])", mangled)
        if m:
            w, c, xl, cost = (int(v) for v in m.groups())
            base = f"xl_bw{64 * w * c}" if xl else f"rowscan_w{w}_c{c}"
            return base + ("_cost" if cost else "")
        m = re.search(r"N4xted\d+(\w+?)E", mangled) or re.search(r"4xted(\d+)(\w+)", mangled)
        for name in ("leaf_cost_kernel", "leaf_kernel", "tiny_kernel", "plan_kernel", "tree_prep_kernel",
                     "cost_prefix_kernel", "xl_layout_kernel", "xl_zero_kernel",
                     "plan_scan_kernel", "plan_clear_kernel", "gather_dist_kernel"):
            if name in mangled:
                if name == "tiny_kernel":
                    m = re.search(r"tiny_kernelILi(\d+)ELb([01])E", mangled)
                    return f"tiny_s{m.group(1)}" + ("_cost" if m.group(2) == "1" else "")
                if name == "plan_kernel":
                    return "plan_kernel<true>" if "ILb1E" in mangled else "plan_kernel<false>"
                return name
        return mangled
    head = (f"{'kernel':<24}{'vgpr':>6}{'agpr':>6}{'sgpr':>6}{'scratch_B':>10}"
            f"{'waves/SIMD':>11}{'lds_static_B':>13}")
    out = [head]
    for r in rows:
        out.append(f"{pretty(r['name']):<24}{r.get('vgpr', 0):>6}{r.get('agpr', 0):>6}"
                   f"{r.get('sgpr', 0):>6}{r.get('scratch_B', 0):>10}{r.get('occupancy', 0):>11}"
                   f"{r.get('lds_static_B', 0):>13}")
    return "\n".join(out)


def _sums(path):
    per = defaultdict(lambda: defaultdict(float))
    with open(path) as f:
        for r in csv.DictReader(f):
            per[short(r["Kernel_Name"])][r["Counter_Name"]] += float(r["Counter_Value"])
    return per


def _weighted(path):
    """Per kernel and derived counter: the average over dispatches weighted
    by each dispatch's duration (for ratios such as MeanOccupancyPerCU)."""
    acc = defaultdict(lambda: defaultdict(lambda: [0.0, 0.0]))
    with open(path) as f:
        for r in csv.DictReader(f):
            d = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
            a = acc[short(r["Kernel_Name"])][r["Counter_Name"]]
            a[0] += float(r["Counter_Value"]) * d
            a[1] += d
    return {k: {c: v[0] / max(v[1], 1) for c, v in cs.items()} for k, cs in acc.items()}


def counters(sq_path, lds_path, occ_path=None, valu_path=None):
    """SQ / LDS counter sums per kernel; with the occupancy and VALU-busy
    runs, also the duration-weighted MeanOccupancyPerCU (waves resident per
    CU), VALUBusy (% of SIMD cycles issuing VALU) and SIMD_UTILIZATION
    (share of time CUs hold any wave)."""
    sq, lds = _sums(sq_path), _sums(lds_path)
    occ = _weighted(occ_path) if occ_path else {}
    valu = _weighted(valu_path) if valu_path else {}
    head = (f"{'kernel':<22}{'waves':>12}{'valu/wave':>11}{'lds/wave':>10}{'vmem/wave':>10}"
            f"{'valu_active':>12}{'waiting':>9}{'lds_conflict':>13}")
    if occ or valu:
        head += f"{'occ/CU':>8}{'VALUBusy%':>10}{'SIMD_util':>10}"
    out = [head]
    for k in sorted(sq, key=lambda k: -sq[k]["SQ_BUSY_CYCLES"]):
        a, b = sq[k], lds.get(k, {})
        w = max(a["SQ_WAVES"], 1)
        cyc = max(a["SQ_WAVE_CYCLES"], 1)
        line = (f"{k:<22}{a['SQ_WAVES']:>12.0f}{a['SQ_INSTS_VALU'] / w:>11.0f}"
                f"{a['SQ_INSTS_LDS'] / w:>10.0f}{a['SQ_INSTS_VMEM_RD'] / w:>10.0f}"
                f"{a['SQ_ACTIVE_INST_VALU'] / cyc:>12.3f}{a['SQ_WAIT_INST_ANY'] / cyc:>9.3f}"
                f"{b.get('SQ_LDS_BANK_CONFLICT', 0) / max(b.get('SQ_LDS_IDX_ACTIVE', 0), 1):>13.3f}")
        if occ or valu:
            o, v = occ.get(k, {}), valu.get(k, {})
            line += (f"{o.get('MeanOccupancyPerCU', float('nan')):>8.2f}"
                     f"{v.get('VALUBusy', float('nan')):>10.1f}"
                     f"{v.get('SIMD_UTILIZATION', float('nan')):>10.3f}")
        out.append(line)
    return "\n".join(out)


def _per_dispatch(path, kernel):
    vals = []
    with open(path) as f:
        for r in csv.DictReader(f):
            if kernel is None or short(r["Kernel_Name"]) == kernel:
                vals.append(float(r["Counter_Value"]))
    return vals


def traffic(kernel, fetch, write, cfetch, cwrite, copy_bytes):
    f = _per_dispatch(fetch, kernel)
    w = _per_dispatch(write, kernel)
    # The copy run's largest dispatch is the copy itself.
    cf = max(_per_dispatch(cfetch, None))
    cw = max(_per_dispatch(cwrite, None))
    kf = cf * 1024 / copy_bytes  # counter KiB per KiB read
    kw = cw * 1024 / copy_bytes
    fk, wk = sum(f) / len(f), sum(w) / len(w)
    hbm = (fk / kf + wk / kw) * 1024
    return "\n".join([
        f"kernel: {kernel}",
        f"dispatches: {len(f)}",
        f"fetch_kib_per_dispatch: {fk:.1f}",
        f"write_kib_per_dispatch: {wk:.1f}",
        f"copy_calibration: {copy_bytes} bytes copied; FETCH_SIZE {cf:.1f} KiB "
        f"(factor {kf:.4f}), WRITE_SIZE {cw:.1f} KiB (factor {kw:.4f})",
        f"hbm_bytes_per_launch: {hbm:.0f}",
    ])


def levels(bench, key):
    """Per-level table of a bench.py JSON line's level_profile (the batch's
    at the top level, or under key, e.g. xl_pair)."""
    with open(bench) as f:
        line = [x for x in f.read().splitlines() if x.startswith("{")][-1]
    d = json.loads(line)
    lp = (d[key] if key else d)["level_profile"]
    out = [f"levels: {lp['levels']}",
           f"critical_path_ms (sum over levels of the longest launch, one stream): {lp['critical_path_ms']}",
           f"level_spans_ms_4_streams: {lp['level_spans_ms_4_streams']}",
           f"dp_span_ms_4_streams: {lp['dp_span_ms_4_streams']}",
           f"thin_levels (fewer workgroups than half the CUs): {lp['thin_levels']}, "
           f"{lp['thin_level_ms_4_streams']} ms, fraction {lp['thin_fraction']}",
           f"{'level':>6}{'tables':>12}{'longest_1s_ms':>15}{'span_4s_ms':>12}"]
    for L, (t, lo, sp) in enumerate(zip(lp["per_level_tables"], lp["per_level_longest_ms_1_stream"],
                                        lp["per_level_ms_4_streams"])):
        out.append(f"{L:>6}{t:>12}{lo:>15.3f}{sp:>12.3f}")
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("trace")
    k.add_argument("--bench")
    c = sub.add_parser("counters")
    c.add_argument("sq")
    c.add_argument("lds")
    c.add_argument("--occ", help="run with --pmc MeanOccupancyPerCU")
    c.add_argument("--valu", help="run with --pmc VALUBusy SIMD_UTILIZATION")
    rs = sub.add_parser("resources")
    rs.add_argument("source", nargs="?", default="csrc/forest.hip",
                    help="HIP source, relative to x-ted_amd/")
    t = sub.add_parser("traffic")
    t.add_argument("kernel")
    t.add_argument("fetch")
    t.add_argument("write")
    t.add_argument("copy_fetch")
    t.add_argument("copy_write")
    t.add_argument("--copy-bytes", type=int, required=True)
    lv = sub.add_parser("levels", help="per-level table from a bench.py JSON line")
    lv.add_argument("bench")
    lv.add_argument("--key", default=None, help="e.g. xl_pair for the 20k pair's profile")
    a = ap.parse_args()
    if a.cmd == "levels":
        print(levels(a.bench, a.key))
    elif a.cmd == "resources":
        print(resources(a.source))
    elif a.cmd == "kernels":
        print(kernels(a.trace, a.bench))
    elif a.cmd == "counters":
        print(counters(a.sq, a.lds, a.occ, a.valu))
    else:
        print(traffic(a.kernel, a.fetch, a.write, a.copy_fetch, a.copy_write, a.copy_bytes))


if __name__ == "__main__":
    sys.exit(main())
