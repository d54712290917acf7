"""Summaries of rocprofv3 runs of bench.py (kernel trace and HBM counters).

    python 4d_lung_amd/profiling.py TRACE_DIR FETCH_DIR WRITE_DIR WORKLOAD OUT_DIR [BINS_PER_LAUNCH]
    python 4d_lung_amd/profiling.py raw RD_DIR WR_DIR TRACE_DIR WORKLOAD OUT_DIR [BINS_PER_LAUNCH]

TRACE_DIR holds ``run_kernel_trace.csv`` of ``rocprofv3 --kernel-trace
--stats -f csv -o run``; FETCH_DIR / WRITE_DIR hold ``run_counter_collection.csv``
of ``rocprofv3 --pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE`` runs of
``bench.py --calibrate-copy``.  Kernels are named by CG stage (the forward
strided passes run y then x, the inverse ones x then y; a fused gridding
launch, k_fft_xgrid, is the inverse x pass of its transform).  The counters are
scaled by what they read for the calibration copy (1 GiB read and 1 GiB
written by the runtime's copy kernel).  Per stage the median over the run's
launches is reported (setup launches of the same kernels are outnumbered by
the CG iterations).  Writes
``OUT_DIR/kernels_WORKLOAD.md`` (durations; a description above the tables of
an existing file is kept) and ``OUT_DIR/traffic_WORKLOAD.txt`` (bytes per
launch, read back by bench.py for ``roofline.traffic``).  With
BINS_PER_LAUNCH (binned workloads profiled with every launch covering that
many bins) the traffic is written per bin.
"""
import collections
import csv
import os
import re
import statistics
import sys

CAL_BYTES = 1 << 30
TABLE_MARK = "Per CG stage (kernel trace):"


def _short(name):
    return re.sub(r"\(.*\)$", "", name).replace("void ", "").strip()


def stage_of(name, counters):
    """CG-stage name of a dispatch (None for setup-only kernels)."""
    n = _short(name)
    if "k_fft_zpre" in n:
        return "fft_z_fwd"
    if "k_fft_zpost" in n:
        return "fft_z_inv"
    if "k_fft_xgrid" in n:
        # fused gridding + inverse x pass: the inverse transform's next plain
        # launch is its y pass
        counters["xgrid"] = True
        return "fft_x_inv_grid"
    m = re.search(r"k_fft_plain<(\d+), (\d+), (-?1)", n)
    if m:
        d = int(m.group(3))
        if d > 0 and counters.pop("xgrid", False):
            return "fft_y_inv"
        k = counters[d]
        counters[d] += 1
        if d < 0:
            return "fft_y_fwd" if k % 2 == 0 else "fft_x_fwd"
        return "fft_x_inv" if k % 2 == 0 else "fft_y_inv"
    for key, st in (("k_degrid", "degrid"), ("k_grid<", "grid"), ("k_cg_update_xr", "cg_update_xr"),
                    ("k_cg_update_p", "cg_update_p"), ("k_cg_alpha", "cg_alpha"), ("k_cg_beta", "cg_beta")):
        if key in n:
            return st
    return None


def _rows(path):
    with open(path) as f:
        return list(csv.DictReader(f))


def kernel_durations(trace_dir):
    per = collections.defaultdict(list)
    names = {}
    cnt = {-1: 0, 1: 0}
    for r in _rows(os.path.join(trace_dir, "run_kernel_trace.csv")):
        st = stage_of(r["Kernel_Name"], cnt)
        if st:
            per[st].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
            names[st] = _short(r["Kernel_Name"])
    return per, names


def counter_bytes(pmc_dir, counter):
    rows = [r for r in _rows(os.path.join(pmc_dir, "run_counter_collection.csv")) if r["Counter_Name"] == counter]
    copies = [float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"]]
    if not copies:
        raise SystemExit("no calibration copy found in %s (run bench.py with --calibrate-copy)" % pmc_dir)
    scale = CAL_BYTES / (max(copies) * 1024.0)
    per = collections.defaultdict(list)
    cnt = {-1: 0, 1: 0}
    for r in rows:
        st = stage_of(r["Kernel_Name"], cnt)
        if st:
            per[st].append(float(r["Counter_Value"]) * 1024.0 * scale)
    return per, max(copies), scale


# best streaming rates measured on the box (4 GiB buffers): a copy (read +
# write) and a write-only fill
COPY_GBS = 5605.0
FILL_GBS = 6300.0

RAW_READ = ("TCC_EA0_RDREQ_sum", "TCC_EA0_RDREQ_32B_sum", "TCC_BUBBLE_sum", "TCC_EA0_RDREQ_DRAM_sum")
RAW_WRITE = ("TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum", "TCC_EA0_WRREQ_DRAM_sum")


def raw_counters(pmc_dir, names):
    """Per stage and for the 1 GiB calibration copy (its largest dispatch):
    {stage: {counter: median per dispatch}}."""
    disp = collections.OrderedDict()
    for r in _rows(os.path.join(pmc_dir, "run_counter_collection.csv")):
        d = disp.setdefault(r["Dispatch_Id"], {"name": r["Kernel_Name"]})
        d[r["Counter_Name"]] = float(r["Counter_Value"])
    per = collections.defaultdict(lambda: collections.defaultdict(list))
    copy = collections.defaultdict(float)
    cnt = {-1: 0, 1: 0}
    for d in disp.values():
        if "copyBuffer" in d["name"]:
            for c in names:
                copy[c] = max(copy[c], d.get(c, 0.0))
            continue
        st = stage_of(d["name"], cnt)
        if st:
            for c in names:
                per[st][c].append(d.get(c, 0.0))
    out = {st: {c: statistics.median(v) for c, v in cs.items()} for st, cs in per.items()}
    out["calibration copy"] = dict(copy)
    return out


def raw_main(argv):
    """raw RD_DIR WR_DIR TRACE_DIR WORKLOAD OUT_DIR [BINS_PER_LAUNCH]: table of
    the TCC request counters per stage (profiles/counters_WORKLOAD_requests.md)."""
    rd_dir, wr_dir, trace_dir, workload, out_dir = argv[2:7]
    nb = int(argv[7]) if len(argv) > 7 else 1
    rd = raw_counters(rd_dir, RAW_READ)
    wr = raw_counters(wr_dir, RAW_WRITE)
    dur, _ = kernel_durations(trace_dir)
    med = statistics.median
    path = os.path.join(out_dir, "counters_%s_requests.md" % workload)
    with open(path, "w") as f:
        f.write("# TCC request counters of the CG-iteration kernels, workload %s\n\n" % workload)
        f.write("Commands (on the GPU box, from `/tmp` with `TMPDIR=/tmp`, `$R` = repository root, `$A` = the\n"
                "bench.py arguments of the kernel trace beside this file):\n\n```\n"
                "rocprofv3 --pmc %s -f csv -d $R/gpurun_out/pmc_rawrd_%s -o run -- python3 $R/bench.py $A\n"
                "rocprofv3 --pmc %s -f csv -d $R/gpurun_out/pmc_rawwr_%s -o run -- python3 $R/bench.py $A\n"
                "python3 4d_lung_amd/profiling.py raw gpurun_out/pmc_rawrd_%s gpurun_out/pmc_rawwr_%s "
                "gpurun_out/ktrace_%s %s profiles%s\n```\n\n"
                % (" ".join(RAW_READ), workload, " ".join(RAW_WRITE), workload, workload, workload, workload,
                   workload, " %d" % nb if nb > 1 else ""))
        f.write("Median per dispatch%s; the calibration copy moves 1 GiB each way.  Read bytes are\n"
                "given for 64-byte and for 128-byte requests (the counters cannot tell them apart here,\n"
                "see below); writes are 64-byte requests.\n\n" % (" (launches of %d bins)" % nb if nb > 1 else ""))
        f.write("| stage | RDREQ | RDREQ_32B | BUBBLE | RDREQ_DRAM | read GB (64 B / 128 B) | WRREQ | WRREQ_64B |"
                " WRREQ_DRAM | write GB | median us | GB/s at 128 B reads |\n")
        f.write("|---|---|---|---|---|---|---|---|---|---|---|---|\n")
        for st in ["calibration copy"] + sorted(dur, key=lambda k: -med(dur[k])):
            if st not in rd or st not in wr:
                continue
            r, w = rd[st], wr[st]
            r64 = r["TCC_EA0_RDREQ_sum"] * 64 / 1e9
            wgb = (w["TCC_EA0_WRREQ_64B_sum"] * 64 + (w["TCC_EA0_WRREQ_sum"] - w["TCC_EA0_WRREQ_64B_sum"]) * 32) / 1e9
            us = med(dur[st]) if st in dur else None
            rate = "%.0f" % ((2 * r64 + wgb) / (us * 1e-6)) if us else ""
            f.write("| %s | %.4g | %.4g | %.4g | %.4g | %.3f / %.3f | %.4g | %.4g | %.4g | %.3f | %s | %s |\n" % (
                st, r["TCC_EA0_RDREQ_sum"], r["TCC_EA0_RDREQ_32B_sum"], r["TCC_BUBBLE_sum"],
                r["TCC_EA0_RDREQ_DRAM_sum"], r64, 2 * r64, w["TCC_EA0_WRREQ_sum"], w["TCC_EA0_WRREQ_64B_sum"],
                w["TCC_EA0_WRREQ_DRAM_sum"], wgb, "%.1f" % us if us else "", rate))
        f.write(RAW_NOTE)
    return 0


RAW_NOTE = """
The 1 GiB calibration copy issues 8.39 M read requests, 128 bytes each, yet
TCC_BUBBLE (128-byte read requests) and RDREQ_32B read 0, so FETCH_SIZE
(128 BUBBLE + 64 (RDREQ - BUBBLE - RDREQ_32B) + 32 RDREQ_32B bytes) counts
them as 64 bytes: 0.5 GiB.  Writes are counted exactly (WRREQ_64B x 64 bytes =
1 GiB).  A read request of these counters is therefore 64 or 128 bytes, and
the counters do not say which.  The strided FFT passes issue half as many read
as write requests, i.e. 128-byte reads, and at 128 bytes their reads equal
their algorithmic bytes.  RDREQ_DRAM equals RDREQ everywhere: hits in the
memory-side cache are not told apart from HBM reads.  The traffic summaries
(FETCH_SIZE x 2 + WRITE_SIZE) are therefore the L2-to-memory bytes at
128-byte reads: exact for kernels that read whole 128-byte lines, an upper
bound (up to 2x on reads) for kernels that issue 64-byte reads, and inclusive
of memory-side-cache hits.
"""


def main(argv):
    if len(argv) > 1 and argv[1] == "raw":
        return raw_main(argv)
    trace_dir, fetch_dir, write_dir, workload, out_dir = argv[1:6]
    nb = int(argv[6]) if len(argv) > 6 else 1
    dur, names = kernel_durations(trace_dir)
    fetch, fcal, fs = counter_bytes(fetch_dir, "FETCH_SIZE")
    write, wcal, ws = counter_bytes(write_dir, "WRITE_SIZE")
    med = statistics.median
    order = sorted(dur, key=lambda k: -med(dur[k]))
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, "traffic_%s.txt" % workload), "w") as f:
        if nb > 1:
            f.write("# HBM bytes per bin of the CG-iteration kernels (launches of %d bins), workload %s\n"
                    % (nb, workload))
        else:
            f.write("# HBM bytes per launch of the CG-iteration kernels, workload %s\n" % workload)
        f.write("# FETCH_SIZE / WRITE_SIZE (KiB per dispatch), median over the run's launches,\n")
        f.write("# scaled by the 1 GiB calibration copy: it read %.0f KiB (x%.4f) and wrote %.0f KiB (x%.4f)\n"
                % (fcal, fs, wcal, ws))
        f.write("# (the copy's reads are 128-byte requests that FETCH_SIZE counts as 64 bytes: the read\n"
                "# figures assume the same of every kernel and are an upper bound where a kernel issues\n"
                "# 64-byte read requests; they include what the memory-side cache serves)\n")
        f.write("# stage fetch_bytes write_bytes implied_GBps (launch bytes / median kernel-trace time)\n")
        for st in order:
            if st in fetch and st in write:
                fb = med(fetch[st]) / nb
                wb = med(write[st]) / nb
                gbs = (fb + wb) * nb / (med(dur[st]) * 1e-6) / 1e9
                # a stage that mostly writes is held to the fill rate, others to the copy rate
                kind, bound = ("fill", FILL_GBS) if fb < 0.25 * wb else ("copy", COPY_GBS)
                note = ("  # above the best %s rate (%.0f GB/s): 64-byte reads or memory-side cache hits"
                        % (kind, bound) if gbs > bound else "")
                f.write("%s %.6g %.6g %.0f%s\n" % (st, fb, wb, gbs, note))
    kpath = os.path.join(out_dir, "kernels_%s.md" % workload)
    head = ""
    if os.path.exists(kpath):   # the description above the tables is kept
        with open(kpath) as f:
            head = f.read().split(TABLE_MARK)[0]
    with open(kpath, "w") as f:
        f.write(head)
        f.write(TABLE_MARK + "\n\n")
        f.write("| stage | kernel | launches | median us | min us | max us |\n|---|---|---|---|---|---|\n")
        for st in order:
            d = dur[st]
            f.write("| %s | `%s` | %d | %.1f | %.1f | %.1f |\n" % (st, names[st], len(d), med(d), min(d), max(d)))
        stats = os.path.join(trace_dir, "run_kernel_stats.csv")
        if os.path.exists(stats):
            f.write("\nrocprofv3 --stats (all kernels of the run, by total time):\n\n")
            f.write("| kernel | calls | average us | % of GPU time |\n|---|---|---|---|\n")
            for r in _rows(stats)[:16]:
                f.write("| `%s` | %s | %.1f | %.2f |\n" % (_short(r["Name"]), r["Calls"], float(r["AverageNs"]) / 1e3,
                                                          float(r["Percentage"])))
    return 0


def read_traffic(path):
    """({stage: (fetch_bytes, write_bytes)}, per_bin) from a traffic_*.txt
    summary; per_bin when the figures are per bin of a binned workload."""
    out = {}
    per_bin = False
    if not os.path.exists(path):
        return out, per_bin
    with open(path) as f:
        for line in f:
            if line.startswith("# HBM bytes per bin"):
                per_bin = True
            if line.startswith("#") or not line.strip():
                continue
            st, fb, wb = line.split("#")[0].split()[:3]
            out[st] = (float(fb), float(wb))
    return out, per_bin


if __name__ == "__main__":
    sys.exit(main(sys.argv))
