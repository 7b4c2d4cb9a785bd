"""Summarises rocprofv3 output of bench.py runs into the short text kept
under profiles/.

  kernels:  python x-ted_amd/rocprof_summary.py kernels <kernel_trace.csv> [--bench bench.json]
            per kernel: dispatches, average and total duration; with the
            bench JSON line, the algorithmic bytes per launch of each class
            (4 bytes of TD per interior cell) and the rate they imply.
  counters: python x-ted_amd/rocprof_summary.py counters <sq.csv> <lds.csv>
            per kernel from the SQ counter runs: waves, VALU / LDS / VMEM
            instructions per wave, VALU-active and waiting shares of wave
            cycles, LDS bank-conflict share of LDS-active cycles.
  traffic:  python x-ted_amd/rocprof_summary.py traffic KERNEL <fetch.csv> <write.csv>
                <copy_fetch.csv> <copy_write.csv> --copy-bytes N
            FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of the dispatches of
            KERNEL, each scaled by what the counter reads per KiB of a plain
            copy of N bytes, and their sum as HBM bytes per launch.
"""
import argparse
import csv
import json
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
        return "tiny_kernel_cost" if base == "tiny_kernel<true>" else "tiny_kernel"
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


def _sums(path):
    per = defaultdict(lambda: defaultdict(float))
    with open(path) as f:
        for r in csv.DictReader(f):
            per[short(r["Kernel_Name"])][r["Counter_Name"]] += float(r["Counter_Value"])
    return per


def counters(sq_path, lds_path):
    sq, lds = _sums(sq_path), _sums(lds_path)
    out = [f"{'kernel':<22}{'waves':>12}{'valu/wave':>11}{'lds/wave':>10}{'vmem/wave':>10}"
           f"{'valu_active':>12}{'waiting':>9}{'lds_conflict':>13}"]
    for k in sorted(sq, key=lambda k: -sq[k]["SQ_BUSY_CYCLES"]):
        a, b = sq[k], lds.get(k, {})
        w = max(a["SQ_WAVES"], 1)
        cyc = max(a["SQ_WAVE_CYCLES"], 1)
        out.append(f"{k:<22}{a['SQ_WAVES']:>12.0f}{a['SQ_INSTS_VALU'] / w:>11.0f}"
                   f"{a['SQ_INSTS_LDS'] / w:>10.0f}{a['SQ_INSTS_VMEM_RD'] / w:>10.0f}"
                   f"{a['SQ_ACTIVE_INST_VALU'] / cyc:>12.3f}{a['SQ_WAIT_INST_ANY'] / cyc:>9.3f}"
                   f"{b.get('SQ_LDS_BANK_CONFLICT', 0) / max(b.get('SQ_LDS_IDX_ACTIVE', 0), 1):>13.3f}")
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


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("trace")
    k.add_argument("--bench")
    c = sub.add_parser("counters")
    c.add_argument("sq")
    c.add_argument("lds")
    t = sub.add_parser("traffic")
    t.add_argument("kernel")
    t.add_argument("fetch")
    t.add_argument("write")
    t.add_argument("copy_fetch")
    t.add_argument("copy_write")
    t.add_argument("--copy-bytes", type=int, required=True)
    a = ap.parse_args()
    if a.cmd == "kernels":
        print(kernels(a.trace, a.bench))
    elif a.cmd == "counters":
        print(counters(a.sq, a.lds))
    else:
        print(traffic(a.kernel, a.fetch, a.write, a.copy_fetch, a.copy_write, a.copy_bytes))


if __name__ == "__main__":
    sys.exit(main())
