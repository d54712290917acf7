"""Host-side pieces of the measurement and tuning code (no GPU): the bin
grouping candidates of the autotuner, the grid cells the gridding kernels
touch (against a brute-force count), the algorithmic byte counts (against the
per-iteration formula they add up to) and the rocprofv3 traffic summaries."""
import importlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

tune = importlib.import_module("4d_lung_amd.autotune")
prof = importlib.import_module("4d_lung_amd.profiling")


def test_group_candidates_are_equal_splits():
    assert tune.group_candidates(20, 10) == [10, 5, 3, 2, 1]
    assert tune.group_candidates(8, 8) == [8, 4, 2, 1]
    assert tune.group_candidates(20, 7) == [7, 4, 2, 1]
    assert tune.group_candidates(1, 1) == [1]
    for n in range(1, 30):
        for g in range(1, n + 1):
            for e in tune.group_candidates(n, g):
                ngroups = -(-n // e)
                assert e <= g and -(-n // ngroups) == e   # an equal split of n bins


def test_touched_cells_brute_force():
    rng = np.random.default_rng(1)
    for grid in [(12, 10, 9), (8, 16, 6)]:
        k = rng.uniform(0, 1, (40, 3))
        k[:3] = [[0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [0.0, 1.0, 0.5]]   # windows clamped at the edges
        cells = set()
        for p in k:
            n = [int(np.floor((grid[a] - 1) * p[a] + 0.5)) for a in range(3)]
            for o in np.ndindex(5, 5, 5):
                cells.add(tuple(min(max(n[a] + o[a] - 2, 0), grid[a] - 1) for a in range(3)))
        assert bench.touched_cells(k, grid, 5) == len(cells)


def test_algorithmic_bytes_add_up_to_the_iteration_formula():
    """Per iteration and bin without pruning: C(56 Ng + ...) of the design's
    byte model, here (112C + 104) Ng + M (16C + 292) with the FP64 tables."""
    ng, C, m, vs = 1000, 16, 300, 16
    stages = ["fft_z_fwd", "fft_y_fwd", "fft_x_fwd", "degrid", "grid", "fft_x_inv", "fft_y_inv", "fft_z_inv",
              "cg_update_xr", "cg_update_p", "cg_alpha", "cg_beta"]
    tot = sum(bench.algorithmic_bytes(s, ng, C, m, vs) for s in stages)
    assert tot == (112 * C + 104) * ng + m * (16 * C + 292)
    # pruning removes the skipped groups' reads and writes from the forward x
    # pass, their reads from the inverse one, and the skipped tiles' writes
    full = {s: bench.algorithmic_bytes(s, ng, C, m, vs) for s in stages}
    part = {s: bench.algorithmic_bytes(s, ng, C, m, vs, None, 0.25, 0.5) for s in stages}
    assert part["fft_x_fwd"] == 0.25 * full["fft_x_fwd"]
    assert part["fft_x_inv"] == 0.625 * full["fft_x_inv"]
    assert full["grid"] - part["grid"] == 0.5 * 8 * C * ng
    # degridding reads only the touched cells
    assert bench.algorithmic_bytes("degrid", ng, C, m, vs, 100) == 8 * C * 100 + m * (16 + 8 * vs + 4 + 8 * C)


def test_traffic_summary_round_trip(tmp_path):
    p = tmp_path / "traffic_x.txt"
    p.write_text("# HBM bytes per bin of the CG-iteration kernels (launches of 5 bins), workload x\n"
                 "# stage fetch_bytes write_bytes\nfft_y_fwd 2e9 2e9\ngrid 1e8 7e9\n")
    t, per_bin = prof.read_traffic(str(p))
    assert per_bin and t == {"fft_y_fwd": (2e9, 2e9), "grid": (1e8, 7e9)}
    t, per_bin = prof.read_traffic(str(tmp_path / "missing.txt"))
    assert t == {} and not per_bin
    t, per_bin = prof.read_traffic(os.path.join(ROOT, "profiles", "traffic_single_160x160x112.txt"))
    assert not per_bin and set(t) >= {"grid", "degrid", "fft_x_fwd", "fft_z_inv"}


def test_stage_names_of_kernel_trace():
    cnt = {-1: 0, 1: 0}
    names = ["void cgs::k_fft_plain<320, 16, -1, 3, 0>(cgs::PlainArgs)", "cgs::k_fft_plain<320, 16, -1, 4, 0>",
             "cgs::k_fft_plain<320, 16, 1, 4, 0>", "cgs::k_fft_plain<320, 16, 1, 3, 0>",
             "cgs::k_fft_zpre<224, 8, 256, true, true>", "cgs::k_fft_zpost<224, 8, 256>",
             "cgs::k_grid<5, 4, 8, 8, 16, 2>", "cgs::k_degrid<5, 16>", "cgs::k_copy_rhs"]
    got = [prof.stage_of(n, cnt) for n in names]
    assert got == ["fft_y_fwd", "fft_x_fwd", "fft_x_inv", "fft_y_inv", "fft_z_fwd", "fft_z_inv", "grid", "degrid",
                   None]
    # fused gridding: the x pass of the inverse transform is k_fft_xgrid, so
    # the one plain inverse launch after it is the y pass; a later unfused
    # transform alternates x, y again
    cnt = {-1: 0, 1: 0}
    names = ["cgs::k_fft_xgrid<256, 16, 4, 2, 5>", "cgs::k_fft_plain<256, 32, 1, 6, 0>",
             "cgs::k_fft_zpost_mb<192, 8, 256>", "cgs::k_fft_xgrid<256, 16, 4, 2, 5>",
             "cgs::k_fft_plain<256, 32, 1, 6, 0>", "cgs::k_grid<5, 4, 8, 8, 16, 2>",
             "cgs::k_fft_plain<256, 16, 1, 4, 0>", "cgs::k_fft_plain<256, 32, 1, 6, 0>"]
    got = [prof.stage_of(n, cnt) for n in names]
    assert got == ["fft_x_inv_grid", "fft_y_inv", "fft_z_inv", "fft_x_inv_grid", "fft_y_inv", "grid", "fft_x_inv",
                   "fft_y_inv"]


def test_kernel_summary_keeps_its_description(tmp_path):
    """Rewriting kernels_<workload>.md from a new trace keeps the text above
    the tables (workload description and commands) and replaces the tables."""
    tr = tmp_path / "trace"
    tr.mkdir()
    (tr / "run_kernel_trace.csv").write_text(
        "Kernel_Name,Start_Timestamp,End_Timestamp\n"
        "\"cgs::k_grid<5, 4, 8, 8, 16, 2>(cgs::GridArgs)\",0,1000\n"
        "\"cgs::k_degrid<5, 16>(cgs::GridArgs)\",2000,2500\n")
    for d, c in (("fetch", "FETCH_SIZE"), ("write", "WRITE_SIZE")):
        (tmp_path / d).mkdir()
        (tmp_path / d / "run_counter_collection.csv").write_text(
            "Kernel_Name,Counter_Name,Counter_Value\n"
            "__amd_rocclr_copyBuffer,%s,1048576\n"
            "\"cgs::k_grid<5, 4, 8, 8, 16, 2>(cgs::GridArgs)\",%s,100\n"
            "\"cgs::k_degrid<5, 16>(cgs::GridArgs)\",%s,200\n" % (c, c, c))
    out = tmp_path / "out"
    out.mkdir()
    md = out / "kernels_w.md"
    md.write_text("# title\n\nworkload text\n\n" + prof.TABLE_MARK + "\n\n| old table |\n")
    args = ["x", str(tr), str(tmp_path / "fetch"), str(tmp_path / "write"), "w", str(out)]
    assert prof.main(args) == 0
    text = md.read_text()
    assert text.startswith("# title\n\nworkload text\n\n" + prof.TABLE_MARK)
    assert "old table" not in text and "`cgs::k_grid<5, 4, 8, 8, 16, 2>`" in text
    assert text.count(prof.TABLE_MARK) == 1
    t, _ = prof.read_traffic(str(out / "traffic_w.txt"))
    assert set(t) == {"grid", "degrid"}


def test_bench_refuses_more_than_one_gpu():
    """The benchmark measures one GPU; asked for more it exits with a message
    before touching a device."""
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_survey_model_totals():
    """The per-iteration byte model that counts each 3D FFT once gives the
    totals quoted for the workloads: 23.63 GB per single-frame iteration,
    101.3 GB per batched iteration of the 8-bin 4D workload (whatever the
    split of its 786,432 samples per coil over the bins)."""
    synth = importlib.import_module("4d_lung_amd.synth")
    c2 = bench.survey_model(synth.CONFIGS["single_160x160x112"], [955733], 1.0)
    assert abs(c2["bytes_per_batched_iteration"] / 1e9 - 23.63) < 0.005
    parts = [50000, 120000, 90000, 100000, 146432, 80000, 100000, 100000]
    assert sum(parts) == 786432
    c3 = bench.survey_model(synth.CONFIGS["4d_128x128x96"], parts, 8.0)
    assert abs(c3["bytes_per_batched_iteration"] / 1e9 - 101.3) < 0.05
    # rate 8 bin-iterations/s = one batched iteration per second
    assert abs(c3["gbs"] - c3["bytes_per_batched_iteration"] / 1e9) < 1e-6
