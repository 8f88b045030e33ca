"""Text summary of a rocprofv3 kernel-trace run and its FETCH_SIZE / WRITE_SIZE counter runs.

Usage: python -m dgm_in_rust_amd.rocprof_summary OUT "title" "bench arguments" > summary.txt
       python -m dgm_in_rust_amd.rocprof_summary --trace-only OUT "title" "command" > summary.txt

OUT holds trace/kt_kernel_stats.csv (rocprofv3 --kernel-trace --stats --output-format csv -o kt),
fetch/pmc_counter_collection.csv and write/pmc_counter_collection.csv (rocprofv3 --pmc FETCH_SIZE /
--pmc WRITE_SIZE --output-format csv -o pmc) of the same command. The counters are per-dispatch KiB;
each is scaled by what it reads for the largest plain device copy in the run (bench.py --copy-probe),
whose true read and write sizes are known, and the summary states both factors. The traffic column,
(FETCH * f_fetch + WRITE * f_write) per dispatch, is what bench.py reports as roofline.traffic.
"""
from __future__ import annotations

import collections
import csv
import os
import re
import sys

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak bandwidth, GB/s


def kernel_key(name: str) -> str:
    """Kernel name without namespace, argument list and spaces: k_surface<float,10,16,5>."""
    n = re.sub(r"(dgm::)?\(anonymous namespace\)::", "", name)
    n = re.sub(r"\(dgm::.*|\(.*", "", n).replace("void ", "").replace(" ", "")
    return n[:64]


def read_stats(path):
    rows = {}
    for r in csv.DictReader(open(path)):
        rows[kernel_key(r["Name"])] = dict(calls=int(r["Calls"]), avg=float(r["AverageNs"]) / 1e3,
                                           mn=float(r["MinNs"]) / 1e3, mx=float(r["MaxNs"]) / 1e3,
                                           tot=float(r["TotalDurationNs"]) / 1e6)
    return rows


def read_counter(path, counter):
    per = collections.defaultdict(list)
    grid = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        k = kernel_key(r["Kernel_Name"])
        per[k].append(float(r["Counter_Value"]))
        grid[k].append(int(r["Grid_Size"]))
    return per, grid


def copy_scale(per, grid, true_kib):
    """Scale factor of a counter from the largest __amd_rocclr_copyBuffer dispatch."""
    vals = per.get("__amd_rocclr_copyBuffer")
    if not vals:
        return None, None
    i = max(range(len(vals)), key=lambda j: grid["__amd_rocclr_copyBuffer"][j])
    return true_kib / vals[i], vals[i]


def summarise(out, title, args, copy_mib=256):
    stats = read_stats(os.path.join(out, "trace", "kt_kernel_stats.csv"))
    fetch, fgrid = read_counter(os.path.join(out, "fetch", "pmc_counter_collection.csv"), "FETCH_SIZE")
    write, wgrid = read_counter(os.path.join(out, "write", "pmc_counter_collection.csv"), "WRITE_SIZE")
    kib = copy_mib * 1024.0
    ff, fraw = copy_scale(fetch, fgrid, kib)
    fw, wraw = copy_scale(write, wgrid, kib)
    # the scale factors are rounded to the exact ratio they approximate (the counters are block-granular)
    ff = round(ff * 4) / 4 if ff else None
    fw = round(fw * 4) / 4 if fw else None
    lines = [f"# {title}",
             "# commands (cd /tmp; export TMPDIR=/tmp; R = repository root):",
             f"#   rocprofv3 --kernel-trace --stats --output-format csv -d OUT/trace -o kt -- python3 R/bench.py {args}",
             f"#   rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT/fetch -o pmc -- python3 R/bench.py {args}",
             f"#   rocprofv3 --pmc WRITE_SIZE --output-format csv -d OUT/write -o pmc -- python3 R/bench.py {args}",
             f"#   python -m dgm_in_rust_amd.rocprof_summary OUT",
             f"# counter scale from the {copy_mib} MiB device copy (reads and writes {kib:.0f} KiB): "
             f"FETCH_SIZE read {fraw:.0f} KiB -> x{ff:.2f}; WRITE_SIZE read {wraw:.0f} KiB -> x{fw:.2f}",
             "# avg/min/max in microseconds per dispatch; fetch/write: raw KiB per dispatch (mean);",
             "# traffic: (fetch * fetch_scale + write * write_scale) in MB (1e6 bytes) per dispatch;",
             f"# gb_s: traffic / avg_us; hbm_frac: gb_s / {HBM_PEAK_GBS:.0f} GB/s (MI355X HBM3E peak)",
             "%-40s %6s %9s %9s %9s %9s %11s %11s %10s %8s %8s" % (
                 "kernel", "calls", "avg_us", "min_us", "max_us", "total_ms", "fetch_kib", "write_kib", "traffic_mb",
                 "gb_s", "hbm_frac")]
    for k, r in sorted(stats.items(), key=lambda kv: -kv[1]["tot"]):
        f = sum(fetch[k]) / len(fetch[k]) if fetch.get(k) else float("nan")
        w = sum(write[k]) / len(write[k]) if write.get(k) else float("nan")
        t = (f * ff + w * fw) * 1024.0 / 1e6
        gbs = t * 1e6 / (r["avg"] * 1e-6) / 1e9 if r["avg"] > 0 else float("nan")
        lines.append("%-40s %6d %9.2f %9.2f %9.2f %9.3f %11.1f %11.1f %10.2f %8.0f %8.3f" %
                     (k, r["calls"], r["avg"], r["mn"], r["mx"], r["tot"], f, w, t, gbs, gbs / HBM_PEAK_GBS))
    sq = os.path.join(out, "sq", "pmc_counter_collection.csv")
    if os.path.exists(sq):
        lines += instruction_table(sq, args)
    return "\n".join(lines) + "\n"


SQ_COUNTERS = ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_VALU_MFMA_F32", "SQ_INSTS_SALU")


def instruction_table(path, args):
    """Per-kernel instruction counts of an SQ counter run (rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU
    SQ_INSTS_VALU_MFMA_F32 SQ_INSTS_SALU): wavefronts per dispatch and instructions per wavefront.
    The rows start with "sq" so that read_traffic() skips them."""
    cols = {c: read_counter(path, c)[0] for c in SQ_COUNTERS}
    keys = sorted(set().union(*[set(v) for v in cols.values()]))
    lines = ["#",
             "# instruction counts from a fourth run of the same command:",
             f"#   rocprofv3 --pmc {' '.join(SQ_COUNTERS)} --output-format csv -d OUT/sq -o pmc -- python3 R/bench.py {args}",
             "# waves: wavefronts per dispatch; valu, mfma, salu: instructions per wavefront (dispatch means)",
             "%-43s %10s %9s %8s %8s" % ("sq kernel", "waves", "valu", "mfma", "salu")]

    def mean(c, k):
        v = cols[c].get(k)
        return sum(v) / len(v) if v else float("nan")

    for k in keys:
        w = mean("SQ_WAVES", k)
        if not (w > 0):
            continue
        lines.append("sq %-40s %10.0f %9.1f %8.1f %8.1f" % (k, w, mean("SQ_INSTS_VALU", k) / w,
                                                          mean("SQ_INSTS_VALU_MFMA_F32", k) / w,
                                                          mean("SQ_INSTS_SALU", k) / w))
    return lines


def trace_summary(out, title, command):
    """Per-kernel durations of a kernel-trace run alone (rocprofv3 --kernel-trace --stats --output-format
    csv -d OUT/trace -o kt -- COMMAND), for runs whose kernels are bound by latency, not by traffic."""
    stats = read_stats(os.path.join(out, "trace", "kt_kernel_stats.csv"))
    lines = [f"# {title}",
             "# command (cd /tmp; export TMPDIR=/tmp; R = repository root):",
             f"#   rocprofv3 --kernel-trace --stats --output-format csv -d OUT/trace -o kt -- {command}",
             f"#   python -m dgm_in_rust_amd.rocprof_summary --trace-only OUT",
             "# avg/min/max in microseconds per dispatch",
             "%-48s %7s %10s %10s %10s %10s" % ("kernel", "calls", "avg_us", "min_us", "max_us", "total_ms")]
    for k, r in sorted(stats.items(), key=lambda kv: -kv[1]["tot"]):
        lines.append("%-48s %7d %10.2f %10.2f %10.2f %10.3f" % (k, r["calls"], r["avg"], r["mn"], r["mx"], r["tot"]))
    return "\n".join(lines) + "\n"


def read_traffic(path, prefix):
    """Mean traffic (bytes per dispatch) and mean duration (us) of the kernels whose key starts with
    prefix, from a summary written by summarise(); None when the file or the kernels are absent."""
    if not os.path.exists(path):
        return None
    tr, us = [], []
    for line in open(path):
        if line.startswith("#"):
            continue
        f = line.split()
        if len(f) in (9, 11) and f[0].startswith(prefix) and f[8] != "nan":
            tr.append(float(f[8]) * 1e6)
            us.append(float(f[2]))
    if not tr:
        return None
    return sum(tr) / len(tr), sum(us) / len(us)


if __name__ == "__main__":
    if sys.argv[1] == "--trace-only":
        sys.stdout.write(trace_summary(sys.argv[2], sys.argv[3] if len(sys.argv) > 3 else "rocprofv3 summary",
                                       sys.argv[4] if len(sys.argv) > 4 else "python3 ..."))
        sys.exit(0)
    out = sys.argv[1]
    title = sys.argv[2] if len(sys.argv) > 2 else "rocprofv3 summary"
    args = sys.argv[3] if len(sys.argv) > 3 else "--steps 200 --warmup 20 --no-cpu --copy-probe 256"
    sys.stdout.write(summarise(out, title, args))
