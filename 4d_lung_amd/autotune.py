"""Autotuner of the CG-SENSE solve: gridding tile shape and bins per launch.

Both choices are measured on the caller's own data, not modelled:

* tile shape -- every gridding tile configuration runs one batched CG
  iteration over all bins; the sum of the kernel times that depend on it
  (gridding or the inverse x pass with gridding fused in, and degridding;
  HIP events on the library's stream) decides;
* bins per launch -- the batched launches cover groups of bins (as many as
  device memory holds coil grids for, or fewer); each candidate group size
  runs a few captured iterations, timed end to end.

The reference has no counterpart: it solves one respiratory state after the
other with fixed loops (CGSense.cpp:200-282).
"""
import math
import time

from . import cgs


def tile_stages_ms(rep):
    """Time of the stages the gridding tile shape affects in a profile
    report: degridding (samples in tile order) and gridding, or the inverse
    x pass that has it fused in."""
    ms = rep["degrid"][1]
    for name in ("grid", "fft_x_inv_grid"):
        if name in rep:
            ms += rep[name][1]
    return ms


def tune_tile(ctx, reload, log=None):
    """Selects the fastest gridding tile configuration and returns
    ``(config, {config: grid+degrid ms})``.

    ``reload()`` must send every bin's samples again: a tile change re-sorts
    the samples into the new tiles.  The context is left on the chosen
    configuration with its bins reloaded.
    """
    timings = {}
    for t in range(cgs.lib().cgs_num_tile_configs()):
        ctx.set_tile_config(t)
        reload()
        ctx.cg_prepare(0.0, 2)
        ctx.profile(True)
        ctx.profile_reset()
        ctx.cg_iterate(cgs.SOLVE_BATCHED, 1)
        rep = ctx.profile_report()
        ctx.profile(False)
        timings[t] = tile_stages_ms(rep)
        if log:
            log("tile config %d %s: gridding+degridding %.3f ms per iteration" % (t, ctx.tile_config()[1], timings[t]))
    best = min(timings, key=timings.get)
    ctx.set_tile_config(best)
    reload()
    return best, timings


def group_candidates(nbins, largest):
    """Bins per launch worth trying: the largest group the memory allows,
    then halving down to one bin (each as the equal split it produces)."""
    out = []
    g = largest
    while True:
        ngroups = (nbins + g - 1) // g
        eff = (nbins + ngroups - 1) // ngroups
        if eff not in out:
            out.append(eff)
        if g == 1:
            break
        g = max(1, math.ceil(g / 2))
    return out


def tune_group(ctx, n_iter=2, log=None):
    """Selects the number of bins per batched launch and returns
    ``(bins_per_launch, {bins_per_launch: ms per iteration})``.  The context is
    left with that cap set (cg_prepare must follow)."""
    ctx.set_bin_group(0)
    ctx.cg_prepare(0.0, n_iter + 1)
    largest = ctx.bin_group()
    timings = {}
    for g in group_candidates(ctx.nbins, largest):
        ctx.set_bin_group(g)
        ctx.cg_prepare(0.0, n_iter + 1)
        ctx.cg_iterate(cgs.SOLVE_BATCHED, 1)          # captures the iteration graph
        t0 = time.perf_counter()
        ctx.cg_iterate(cgs.SOLVE_BATCHED, n_iter)     # returns after the device finished
        timings[g] = 1e3 * (time.perf_counter() - t0) / n_iter
        if log:
            log("bins per launch %d: %.2f ms per batched iteration" % (g, timings[g]))
    best = min(timings, key=timings.get)
    ctx.set_bin_group(best)
    return best, timings
