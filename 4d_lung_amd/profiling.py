"""Summaries of rocprofv3 runs of bench.py (kernel trace and HBM counters).

    python 4d_lung_amd/profiling.py TRACE_DIR FETCH_DIR WRITE_DIR WORKLOAD OUT_DIR [BINS_PER_LAUNCH]

TRACE_DIR holds ``run_kernel_trace.csv`` of ``rocprofv3 --kernel-trace
--stats -f csv -o run``; FETCH_DIR / WRITE_DIR hold ``run_counter_collection.csv``
of ``rocprofv3 --pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE`` runs of
``bench.py --calibrate-copy``.  Kernels are named by CG stage (the forward
strided passes run y then x, the inverse ones x then y).  The counters are
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
    m = re.search(r"k_fft_plain<(\d+), (\d+), (-?1)", n)
    if m:
        d = int(m.group(3))
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


def main(argv):
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
        f.write("# stage fetch_bytes write_bytes\n")
        for st in order:
            if st in fetch and st in write:
                fb = med(fetch[st]) / nb
                wb = med(write[st]) / nb
                f.write("%s %.6g %.6g\n" % (st, fb, wb))
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
            st, fb, wb = line.split()
            out[st] = (float(fb), float(wb))
    return out, per_bin


if __name__ == "__main__":
    sys.exit(main(sys.argv))
