#!/usr/bin/env python
"""CG-SENSE bin-iterations/s on one MI355X.

Workload (default): the batched 4D respiratory-resolved case -- 128x128x96
image on a 256x256x192 grid, 16 coils, R=2x2 wave-CAIPI lines (3072 of 12288,
256 samples per line, 3 wave cycles) assigned to 8 respiratory bins by a
seeded respiratory trace (6-20 % of the lines per bin), per-bin Voronoi
density weights, k-space synthesized from a seeded phantom x smooth coil maps
+ complex noise, FP32 gridding/FFT with FP64 CG scalars, 15 CG iterations.
Every bin is its own CG-SENSE solve (the reference's per-state loop,
CGSense.cpp:200-282).  A step is one CG-SENSE iteration of every bin: one
E^H W E application over all coils plus the CG update (Reco::cgSenseDynamic
loop body, Reco.cpp:230-311), counted once per bin.  Inputs are resident in
HBM before the timed region; the right-hand sides E^H W y are formed before
it.  The plain path (bins one after another, host reads the residual every
iteration, like the reference's loop) is timed beside it.

The JSON line carries the measured rate, the relative L2 error of every bin's
image against the FP64 CPU oracle after the workload's iteration count, the
roofline of the dominant kernel (HIP-event timing on the library's stream,
algorithmic bytes per launch) and the CPU baseline (the oracle's FP32
statement of the reference loops timed on this host).  --workload selects
the other configurations (single_160x160x112, plain_64x64x48,
4d_256x256x160).
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def algorithmic_bytes(name, ng, C, m, vs, touched=None, groups=1.0, tiles=1.0):
    """HBM bytes one launch must move at minimum (complex64 = 8 B).  touched:
    grid cells within the KB window of at least one sample, the only ones
    degridding reads (the whole grid when not given).  groups / tiles: the
    fractions of the x pass's column groups that column pruning transforms
    and of the tile columns gridding writes (the forward x pass touches only
    those groups, the inverse one reads them and writes every column;
    without pruning gridding writes every cell, zeros included, since the
    inverse FFT reads the whole grid)."""
    nt = ng if touched is None else touched
    per = {
        "fft_z_fwd": (16 * C + 12) * ng,             # p, S_c, kft in; coil grids out
        "fft_y_fwd": 16 * C * ng, "fft_x_fwd": 16 * C * ng * groups,
        "fft_x_inv": 8 * C * ng * (1.0 + groups), "fft_y_inv": 16 * C * ng,
        # inverse x pass with gridding fused in: the samples in, every column out
        "fft_x_inv_grid": 8 * C * ng + m * (16 + 8 * vs + 8 * C),
        "fft_z_inv": (16 * C + 20) * ng,             # grids, S_c, kft, p in; q out
        "degrid": 8 * C * nt + m * (16 + 8 * vs + 4 + 8 * C),
        "grid": 8 * C * ng * tiles + m * (16 + 8 * vs + 8 * C),
        "cg_update_xr": 48 * ng,
        "cg_update_p": 24 * ng,
        "cg_alpha": 0, "cg_beta": 0,
    }
    return per.get(name, 0)


def touched_cells(k, grid, kw):
    """Grid cells the gridding kernels touch for samples k (m x 3 in [0, 1]):
    the kw cells per axis around each sample's nearest cell, (int)((N-1)k+0.5)
    (Reco.cpp:705-712), clamped to the grid."""
    half = kw // 2
    near = [np.floor((n - 1) * k[:, a] + 0.5).astype(np.int64) for a, n in enumerate(grid)]
    mark = np.zeros(int(np.prod(grid)), bool)
    for ox in range(-half, half + 1):
        ix = np.clip(near[0] + ox, 0, grid[0] - 1) * grid[1]
        for oy in range(-half, half + 1):
            ixy = (ix + np.clip(near[1] + oy, 0, grid[1] - 1)) * grid[2]
            for oz in range(-half, half + 1):
                mark[ixy + np.clip(near[2] + oz, 0, grid[2] - 1)] = True
    return int(mark.sum())


def survey_model(cfg, counts, rate):
    """The per-iteration byte model that counts every 3D FFT once (one read
    and one write of the coil grids): per bin and iteration
    C*(56*Ng + 48*Ni + 44*M_b) + 72*Ni bytes (pad/expand 8 Ng, FFT 16 Ng,
    degrid 8 Ng + 24 M_b, grid 20 M_b + 8 Ng, inverse FFT 16 Ng, crop / conj /
    coil sum 32 Ni per coil, CG vector work 72 Ni).  rate: bin-iterations/s."""
    C = cfg["coils"]
    ng = int(np.prod(cfg["grid"]))
    ni = int(np.prod(cfg["image"]))
    per_bin = [C * (56 * ng + 48 * ni + 44 * m) + 72 * ni for m in counts]
    avg = float(np.mean(per_bin))
    gbs = avg * rate / 1e9
    return {"bytes_per_batched_iteration": float(sum(per_bin)), "gbs": gbs, "frac": gbs / HBM_PEAK_GBS,
            "ceiling_bin_iterations_per_s": HBM_PEAK_GBS * 1e9 / avg}


def stage_rates(rep, n_iter, bytes_per_iter):
    """Per-stage launch count, average launch ms and algorithmic GB/s from a
    profile report of n_iter iterations (bytes_per_iter: stage -> bytes one
    iteration moves at minimum over all bins)."""
    out = {}
    for name, (cnt, ms) in rep.items():
        b = bytes_per_iter.get(name, 0) * n_iter
        out[name] = dict(count=cnt, avg_ms=ms / cnt, bytes=b / cnt if b else 0, n_iter=n_iter,
                         gbs=(b / (ms * 1e-3) / 1e9) if b else None)
    return out


def roofline_of(kernels, workload, nbins=1):
    """Roofline object of the dominant stage (largest total time).  The HBM
    traffic comes from the committed rocprofv3 summary of the workload
    (per bin for binned workloads, scaled to the bins of one launch here)."""
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"] * kernels[k]["count"])
    prof_mod = importlib.import_module("4d_lung_amd.profiling")
    traffic_src = os.path.join("profiles", "traffic_%s.txt" % workload)
    traffic, per_bin = prof_mod.read_traffic(os.path.join(ROOT, traffic_src))
    dk = kernels[dom]
    t = None
    if dom in traffic:
        t = traffic[dom][0] + traffic[dom][1]
        if per_bin:
            t *= nbins * kernels[dom]["n_iter"] / kernels[dom]["count"]
    return {"bound": "hbm", "kernel": dom, "achieved": dk["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (dk["gbs"] / HBM_PEAK_GBS) if dk["gbs"] else None,
            "traffic": t, "traffic_source": traffic_src if t is not None else None,
            "algorithmic_bytes": dk["bytes"], "avg_ms": dk["avg_ms"]}


def bench_binned(args, cfg, pkg, torch):
    """Respiratory-resolved workload: every bin is its own CG-SENSE solve with
    its own Voronoi density weights; all bins share trajectory, maps and KB
    tables.  Bins are held in compact form (acquired sample indices, values and
    weights).  A batched step (groups of bins per launch, device-side loop
    control, captured in a hipGraph) counts one iteration per bin; the plain
    path runs the bins one after another with a host read of the residual per
    iteration, as the reference's per-state loop does (CGSense.cpp:200-282)."""
    synth = pkg.synth
    tune = importlib.import_module("4d_lung_amd.autotune")
    grid, samp, image, C, B = cfg["grid"], cfg["samp"], cfg["image"], cfg["coils"], cfg["bins"]
    ng, ns = int(np.prod(grid)), int(np.prod(samp))
    K, W = args.steps, args.warmup
    wl = synth.binned_workload(pkg, cfg, log=log, pruning=args.pruning)
    ctx, traj, maps, bins, voronoi_s = wl["ctx"], wl["traj"], wl["maps"], wl["bins"], wl["voronoi_seconds"]
    voronoi = voronoi_s is not None
    counts = [int(bb[0].size) for bb in bins]
    ctx.set_grid_lists(args.grid_lists)
    if args.zpass_bins > 0:
        ctx.set_zpass_bins(args.zpass_bins)
    ctx.set_fused_gridding(args.fused_gridding)
    ctx.set_num_bins(B)

    def load_bins():
        for b, (idx, yb, w) in enumerate(bins):
            ctx.set_bin_samples(b, idx, yb, w)

    # autotune: gridding tile on one batched iteration, then bins per launch
    if args.tile < 0:
        tcfg, _ = tune.tune_tile(ctx, load_bins, log)
    else:
        tcfg = args.tile
        ctx.set_tile_config(tcfg)
        load_bins()
    if args.bin_group < 0:
        group, _ = tune.tune_group(ctx, 2, log)
    else:
        group = args.bin_group
        ctx.set_bin_group(group)

    ctx.cg_prepare(0.0, K - 1)
    ctx.cg_iterate(pkg.SOLVE_BATCHED, max(1, W))
    ctx.cg_prepare(0.0, K - 1)
    torch.cuda.synchronize()
    ts = time.perf_counter()
    ctx.cg_iterate(pkg.SOLVE_BATCHED, K)
    torch.cuda.synchronize()
    secs = time.perf_counter() - ts
    x, res, its = ctx.cg_results(want_x=(args.check_iters is None or args.check_iters == K))
    assert all(int(i) == K for i in its), its
    value = B * K / secs
    log("batched: %d iterations of %d bins (%d per launch) in %.3f s: %.2f bin-iterations/s"
        % (K, B, ctx.bin_group(), secs, value))
    plain_value = None
    if not args.no_plain:
        ctx.cg_prepare(0.0, K - 1)
        torch.cuda.synchronize()
        tp = time.perf_counter()
        ctx.cg_iterate(pkg.SOLVE_PLAIN, K)
        torch.cuda.synchronize()
        plain_value = B * K / (time.perf_counter() - tp)
        log("plain (bin after bin, host residual read per iteration): %.2f bin-iterations/s, batched/plain %.2fx"
            % (plain_value, value / plain_value))
    # per-stage HIP-event timing over 2 batched iterations
    n_prof = 2
    ctx.cg_prepare(0.0, K - 1)
    ctx.profile(True)
    ctx.profile_reset()
    ctx.cg_iterate(pkg.SOLVE_BATCHED, n_prof)
    rep = ctx.profile_report()
    ctx.profile(False)
    kw = ctx.kernel_info()[0]
    vs = (3 * kw + 1) & ~1
    touched = [touched_cells(traj[idx], grid, kw) for idx, _, _ in bins]
    pruned = [ctx.prune_fractions(b) for b in range(B)]
    per_iter = {name: sum(algorithmic_bytes(name, ng, C, m, vs, t, gf, tf)
                          for m, t, (gf, tf) in zip(counts, touched, pruned)) for name in rep}
    kernels = stage_rates(rep, n_prof, per_iter)
    for name in sorted(kernels, key=lambda k: -kernels[k]["avg_ms"] * kernels[k]["count"]):
        kk = kernels[name]
        log("  %-14s %8.3f ms per launch (%d launches)  %s" % (
            name, kk["avg_ms"], kk["count"], "%.0f GB/s" % kk["gbs"] if kk["gbs"] else ""))
    total_bytes = sum(per_iter.values())
    out = {
        "metric": "CG-SENSE bin-iterations/sec (E^H E applies/s) at ≤1e-4 rel-L2, 1 MI355X, batched vs plain",
        "value": value,
        "unit": "bin-iterations/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * secs / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 CG scalars)",
        "data": "synthetic (seeded phantom x smooth coil maps + noise, wave-CAIPI, seeded respiratory bins)",
        "config": {"workload": args.workload, "image": list(image), "grid": list(grid), "coils": C, "bins": B,
                   "samples_per_bin": counts, "acceleration": list(cfg["R"]), "cg_iterations": K,
                   "gridding_tile": list(ctx.tile_config()[1]), "tile_config": int(ctx.tile_config()[0]),
                   "bins_per_launch": ctx.bin_group(), "density": "voronoi per bin" if voronoi else "none",
                   "column_pruning": ctx.pruning(),
                   "x_groups_transformed": round(float(np.mean([p[0] for p in pruned])), 4),
                   "grid_tiles_written": round(float(np.mean([p[1] for p in pruned])), 4)},
        "plain_path_value": plain_value,
        "batched_over_plain": value / plain_value if plain_value else None,
        "roofline": roofline_of(kernels, args.workload, B),
        "algorithmic_gbs_per_iteration": total_bytes * value / B / 1e9,
        "survey_model": survey_model(cfg, counts, value),
        "voronoi_seconds": voronoi_s,
        "kernels": {k: {"avg_ms": round(v["avg_ms"], 4), "launches_per_iteration": v["count"] // n_prof,
                        "gbs": round(v["gbs"], 1) if v["gbs"] else None} for k, v in kernels.items()},
    }
    ci = K if args.check_iters is None else args.check_iters
    if args.check_all_bins or args.check_bins == "all":
        check_bins = list(range(B))
    else:
        check_bins = [int(v) for v in args.check_bins.split(",") if v]
    if not args.no_check and ci != K:
        # the FP64 check runs the workload's own iteration count (or what the
        # oracle's time allows): the GPU solves every bin again for that many
        ctx.cg_prepare(0.0, ci - 1)
        ctx.cg_iterate(pkg.SOLVE_BATCHED, ci)
        x, _, its = ctx.cg_results()
        assert all(int(i) == ci for i in its), its
    del ctx
    from oracle import oracle as orc

    orc.set_verbose(1)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    def oracle_for(b):
        idx, yb, w = bins[b]
        d = np.zeros((C, ns), np.complex64)
        d[:, idx] = yb
        wf = None
        if w is not None:
            wf = np.zeros(ns, np.float32)
            wf[idx] = w
        o = orc.OracleReco(grid, samp, C, nthreads=threads)
        o.set_trajectory(traj)
        o.calculate_weights(d)
        o.calculate_kernel_ft()
        o.set_maps(maps)
        o.set_density(wf)
        return o, d

    ref0 = None
    if not args.no_check:
        errs = {}
        for b in check_bins:
            tc = time.time()
            o, d = oracle_for(b)
            xr, _, itr = o.cg_sense(d, 0.0, ci - 1, prec=64)
            assert itr == ci, (b, itr)
            errs[b] = float(np.linalg.norm(x[b].ravel() - xr.ravel()) / np.linalg.norm(xr.ravel()))
            log("FP64 oracle check bin %d after %d iterations: rel-L2 %.3e (%.0fs)" % (b, ci, errs[b], time.time() - tc))
            if b == 0:
                ref0 = xr
            del o, d, xr
        out["rel_l2_vs_fp64_oracle_per_bin"] = errs
        out["rel_l2_vs_fp64_oracle_max"] = max(errs.values())
        out["check_iterations"] = ci
        out["checked_bins"] = len(errs)
        out["within_tolerance"] = bool(max(errs.values()) <= 1e-4)
    del x
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(orc, oracle_for, ci, ref0, threads, "bin 0 of the same workload",
                                           "bin-iterations/s")
    return out


def cpu_baseline(orc, oracle_for, ci, ref0, threads, what, unit, budget_s=30.0):
    """The reference's CPU loop (the oracle's FP32 statement of
    Reco.cpp:175-553: per-sample tap lists, coil-parallel OpenMP, float
    atomic coil sum as Reco.cpp:120-129, own CPU FFT) on one dataset of the
    workload: a warm-up solve of one iteration, then a timed solve of the
    check's iteration count when it fits the time budget (else of as many
    iterations as fit); the median of the per-iteration times.  With the
    full iteration count its image's relative L2 error against the FP64
    oracle is reported too."""
    o, d = oracle_for(0)
    o.set_atomic_coil_sum(True)
    _, _, _, warm_s, warm_it = o.cg_sense(d, 0.0, 0, prec=32, timed=True, iter_times=True)
    per_it = float(warm_it[0])
    n = ci if per_it * ci <= budget_s else max(2, int(budget_s / per_it))
    x32, _, itc, loop_s, per = o.cg_sense(d, 0.0, n - 1, prec=32, timed=True, iter_times=True)
    med = float(np.median(per))
    rec = {"value": 1.0 / med, "unit": unit, "cores": threads, "kind": "port",
           "sample": "median of %d timed CG iterations (after a 1-iteration warm-up solve) of %s: FP32 "
                     "statement of the reference loop Reco.cpp:175-553 (per-sample tap lists, coil-parallel "
                     "OpenMP, float omp-atomic coil sum as Reco.cpp:120-129, own CPU FFT; FFTW is not available, "
                     "so this understates the reference's FFT speed)" % (len(per), what),
           "iterations": int(itc), "seconds": float(loop_s)}
    if ref0 is not None and n == ci:
        rec["rel_l2_vs_fp64_oracle"] = float(np.linalg.norm(x32.ravel() - ref0.ravel()) / np.linalg.norm(ref0.ravel()))
    log("CPU baseline: %.4f %s on %d threads (median of %d iterations, %.1f s)%s" % (
        rec["value"], unit, threads, len(per), loop_s,
        ", rel-L2 vs FP64 %.2e" % rec["rel_l2_vs_fp64_oracle"] if "rel_l2_vs_fp64_oracle" in rec else ""))
    del o, d
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed CG iterations (default: the workload's own)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="4d_128x128x96")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 oracle accuracy check")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--tile", type=int, default=-1, help="gridding tile config (-1: autotune)")
    ap.add_argument("--check-all-bins", action="store_true", help="binned workloads: oracle check of every bin")
    ap.add_argument("--check-bins", default=None,
                    help="binned workloads: bins checked against the oracle, comma separated or 'all' "
                         "(default: the workload's own setting)")
    ap.add_argument("--check-iters", type=int, default=None,
                    help="binned workloads: iterations of the oracle check (default: the workload's own, else "
                         "--steps); with fewer than --steps the checked bins are solved again for that many")
    ap.add_argument("--bin-group", type=int, default=-1, help="binned workloads: bins per launch (-1: autotune)")
    ap.add_argument("--pruning", type=int, default=-1, choices=(-1, 0, 1, 2),
                    help="column pruning of the x passes: 0 off, 1 automatic, 2 on (-1: the library's default)")
    ap.add_argument("--no-plain", action="store_true", help="binned workloads: skip timing the plain path")
    ap.add_argument("--grid-lists", type=int, default=1, choices=(0, 1),
                    help="per-tile gridding candidate lists built at packing (1, the library's default) or a scan "
                         "in every gridding launch (0)")
    ap.add_argument("--fused-gridding", type=int, default=1, choices=(0, 1),
                    help="gridding fused into the inverse x pass where supported (1, the library's default) or a "
                         "separate gridding pass (0)")
    ap.add_argument("--zpass-bins", type=int, default=0,
                    help="bins per workgroup of the batched inverse z pass (0: the library's default, 64)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one device copy of 1 GiB (the reference kernel that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters of a rocprofv3 --pmc run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(bins shard across GPUs without communication; see DESIGN.md)", file=sys.stderr)
        sys.exit(1)

    import torch  # device-wide synchronisation around the timed region

    if args.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: %d bytes read, %d bytes written" % (src.numel() * 4, dst.numel() * 4))
        del src, dst

    pkg = importlib.import_module("4d_lung_amd")
    synth = pkg.synth
    cfg = synth.CONFIGS[args.workload]
    if args.steps is None:
        args.steps = cfg["iters"]
    if args.check_iters is None:
        args.check_iters = cfg.get("check_iters")
    if args.check_bins is None:
        args.check_bins = cfg.get("check_bins", "0")
    if cfg["bins"] != 1:
        print(json.dumps(bench_binned(args, cfg, pkg, torch)), flush=True)
        return
    grid, samp, image, C = cfg["grid"], cfg["samp"], cfg["image"], cfg["coils"]
    ng, ns = int(np.prod(grid)), int(np.prod(samp))
    K, W = args.steps, args.warmup
    tune = importlib.import_module("4d_lung_amd.autotune")

    t0 = time.time()
    traj = synth.wave_caipi_trajectory(grid, samp, cfg["cycles"], cfg["amp"])
    lines = synth.caipi_mask(samp[1], samp[2], cfg["R"], cfg["shift"])
    smask = synth.sample_mask_from_lines(lines, samp[0])
    maps = synth.coil_maps(grid, C, cfg["seed"] + 1)
    rho = synth.phantom(grid, image, cfg["seed"] + 2).reshape(-1)
    log("synthetic inputs %.1fs (%d lines acquired of %d)" % (time.time() - t0, lines.sum(), lines.size))

    ctx = pkg.Context(grid, samp, C)
    if args.pruning >= 0:
        ctx.set_pruning(args.pruning)
    ctx.set_grid_lists(args.grid_lists)
    ctx.set_fused_gridding(args.fused_gridding)
    ctx.set_trajectory(traj)
    ctx.calculate_kernel_ft()
    ctx.set_maps(maps)
    # k-space of the phantom: E applied on the acquired samples, then noise
    mask_data = np.zeros((1, ns), np.complex64)
    mask_data[0, smask] = 1
    ctx.exclude_missing_data(mask_data)
    img = maps * rho[None, :]
    y = ctx.nufft_fwd(img)
    del img
    data = synth.add_noise(y, smask, 1e-3, cfg["seed"] + 3)
    del y
    ctx.exclude_missing_data(data)
    ctx.set_bin_data(0, data)
    m = ctx.bin_count(0)
    kw = ctx.kernel_info()[0]
    vs = (3 * kw + 1) & ~1
    log("data synthesized, %d active samples per coil, %.1fs" % (m, time.time() - t0))

    # gridding tile autotune: fastest of the shapes on one E^H W E application
    if args.tile < 0:
        best = None
        p = np.ascontiguousarray(rho.astype(np.complex64))
        for t in range(pkg.cgs.lib().cgs_num_tile_configs()):
            ctx.set_tile_config(t)
            ctx.set_bin_data(0, data)
            ctx.apply_normal(p)
            ctx.profile(True)
            ctx.profile_reset()
            ctx.apply_normal(p)
            rep = ctx.profile_report()
            ctx.profile(False)
            ms = tune.tile_stages_ms(rep)
            log("tile config %d %s: gridding+degridding %.3f ms" % (t, ctx.tile_config()[1], ms))
            if best is None or ms < best[1]:
                best = (t, ms)
        tcfg = best[0]
    else:
        tcfg = args.tile
    ctx.set_tile_config(tcfg)
    ctx.set_bin_data(0, data)

    # warm-up (also instantiates the captured iteration graph)
    ctx.cg_prepare(0.0, K - 1)
    ctx.cg_iterate(pkg.SOLVE_BATCHED, max(1, W))
    # timed region: exactly K iterations
    ctx.cg_prepare(0.0, K - 1)
    torch.cuda.synchronize()
    ts = time.perf_counter()
    ctx.cg_iterate(pkg.SOLVE_BATCHED, K)
    torch.cuda.synchronize()
    te = time.perf_counter()
    x, res, its = ctx.cg_results()
    assert int(its[0]) == K, its
    secs = te - ts
    value = K / secs
    log("timed %d iterations: %.2f ms/iter, %.1f it/s" % (K, 1e3 * secs / K, value))

    # plain single-frame path (host reads the residual every iteration)
    ctx.cg_prepare(0.0, K - 1)
    torch.cuda.synchronize()
    tp = time.perf_counter()
    ctx.cg_iterate(pkg.SOLVE_PLAIN, K)
    torch.cuda.synchronize()
    plain_value = K / (time.perf_counter() - tp)

    # per-kernel HIP-event timing over 3 profiled iterations
    ctx.cg_prepare(0.0, K - 1)
    ctx.profile(True)
    ctx.profile_reset()
    ctx.cg_iterate(pkg.SOLVE_BATCHED, 3)
    rep = ctx.profile_report()
    ctx.profile(False)
    kernels = {}
    touched = touched_cells(traj[np.flatnonzero(data[0] != 0)], grid, kw)
    gfrac, tfrac = ctx.prune_fractions(0)
    for name, (cnt, ms) in rep.items():
        avg = ms / cnt
        b = algorithmic_bytes(name, ng, C, m, vs, touched, gfrac, tfrac)
        kernels[name] = dict(count=cnt, avg_ms=avg, gbs=(b / (avg * 1e-3) / 1e9) if b else None, bytes=b)
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"] * kernels[k]["count"])
    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE runs of
    # this workload (profiles/traffic_<workload>.txt, scaled by a copy kernel)
    prof_mod = importlib.import_module("4d_lung_amd.profiling")
    traffic_src = os.path.join("profiles", "traffic_%s.txt" % args.workload)
    traffic, _ = prof_mod.read_traffic(os.path.join(ROOT, traffic_src))
    dk = kernels[dom]
    for name in sorted(kernels, key=lambda k: -kernels[k]["avg_ms"]):
        kk = kernels[name]
        log("  %-14s %8.3f ms  %s" % (name, kk["avg_ms"], "%.0f GB/s" % kk["gbs"] if kk["gbs"] else ""))

    out = {
        "metric": "CG-SENSE iterations/sec (E^H E applies/s) at ≤1e-4 rel-L2, 1 MI355X vs plain",
        "value": value,
        "unit": "iterations/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * secs / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 CG scalars)",
        "data": "synthetic (seeded phantom x smooth coil maps + noise, wave-CAIPI)",
        "config": {"workload": args.workload, "image": list(image), "grid": list(grid), "coils": C,
                   "samples_per_coil": int(m), "acceleration": list(cfg["R"]), "cg_iterations": K,
                   "column_pruning": ctx.pruning(),
                   "gridding_tile": list(ctx.tile_config()[1]), "tile_config": int(ctx.tile_config()[0])},
        "plain_path_value": plain_value,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": dk["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (dk["gbs"] / HBM_PEAK_GBS) if dk["gbs"] else None,
                     "traffic": traffic.get(dom, (None, None))[0] + traffic[dom][1] if dom in traffic else None,
                     "traffic_source": traffic_src if dom in traffic else None,
                     "algorithmic_bytes": dk["bytes"], "avg_ms": dk["avg_ms"]},
        # every stage runs once per iteration: the iteration's algorithmic
        # bytes times the iteration rate
        "algorithmic_gbs_per_iteration": sum(v["bytes"] or 0 for v in kernels.values()) * value / 1e9,
        "survey_model": survey_model(cfg, [m], value),
        "kernels": {k: {"avg_ms": round(v["avg_ms"], 4), "gbs": round(v["gbs"], 1) if v["gbs"] else None}
                    for k, v in kernels.items()},
    }
    del ctx

    from oracle import oracle as orc

    orc.set_verbose(1)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if not args.no_check:
        tc = time.time()
        o = orc.OracleReco(grid, samp, C, nthreads=threads)
        o.set_trajectory(traj)
        o.calculate_weights(data)
        o.calculate_kernel_ft()
        o.set_maps(maps)
        xr, resr, itr = o.cg_sense(data, 0.0, K - 1, prec=64)
        err = float(np.linalg.norm(x[0].ravel() - xr.ravel()) / np.linalg.norm(xr.ravel()))
        out["rel_l2_vs_fp64_oracle"] = err
        out["within_tolerance"] = bool(err <= 1e-4)
        log("FP64 oracle check: rel-L2 %.3e (%.0fs)" % (err, time.time() - tc))
        ref0 = xr
        del o

    def oracle_for(_b):
        o = orc.OracleReco(grid, samp, C, nthreads=threads)
        o.set_trajectory(traj)
        o.calculate_weights(data)
        o.calculate_kernel_ft()
        o.set_maps(maps)
        return o, data

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(orc, oracle_for, K, None if args.no_check else ref0, threads,
                                           "the same workload", "iterations/s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
