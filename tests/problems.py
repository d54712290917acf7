"""Seeded test problems shared by the CPU and GPU tests.

Data are synthesized with the FP64 oracle's forward operator (so both the
oracle and the HIP path see identical float32 inputs), undersampled by a CAIPI
line mask and corrupted with complex Gaussian noise (sigma = 1e-3 max|y|).
"""
import importlib

import numpy as np

synth = importlib.import_module("4d_lung_amd.synth")


def small_problem(orc, grid=(32, 32, 24), image=(16, 16, 12), samp=(32, 16, 12), ncoil=4, R=(2, 2), cycles=2,
                  amp=2.0, seed=7, dcf=False, nthreads=0):
    """Returns a dict with trajectory, maps, phantom, data and an oracle object
    whose weights already exclude the missing samples."""
    traj = synth.wave_caipi_trajectory(grid, samp, cycles, amp, border=3.0)
    lines = synth.caipi_mask(samp[1], samp[2], R, 1)
    smask = synth.sample_mask_from_lines(lines, samp[0])
    maps = synth.coil_maps(grid, ncoil, seed + 1, coarse=4)
    rho = synth.phantom(grid, image, seed + 2, n_ellipsoids=5).reshape(-1)
    o = orc.OracleReco(grid, samp, ncoil, nthreads=nthreads)
    o.set_trajectory(traj)
    o.calculate_weights()
    o.calculate_kernel_ft()
    o.set_maps(maps)
    img = maps.astype(np.complex128) * rho[None, :].astype(np.complex128)
    y = o.nufft_fwd(img, prec=64)
    data = synth.add_noise(y, smask, 1e-3, seed + 3)
    o.exclude_missing_data(data)
    d = None
    if dcf:
        rng = np.random.default_rng(seed + 4)
        d = rng.uniform(0.5, 1.5, size=int(np.prod(samp))).astype(np.float32)
        o.set_density(d)
    return dict(grid=grid, samp=samp, ncoil=ncoil, traj=traj, maps=maps, rho=rho, data=data, smask=smask,
                lines=lines, dcf=d, oracle=o)


def config_problem(orc, name, nthreads=0, with_oracle=True):
    """The named workload of synth.CONFIGS (single-bin ones)."""
    cfg = synth.CONFIGS[name]
    grid, samp, image = cfg["grid"], cfg["samp"], cfg["image"]
    traj = synth.wave_caipi_trajectory(grid, samp, cfg["cycles"], cfg["amp"])
    lines = synth.caipi_mask(samp[1], samp[2], cfg["R"], cfg["shift"])
    smask = synth.sample_mask_from_lines(lines, samp[0])
    maps = synth.coil_maps(grid, cfg["coils"], cfg["seed"] + 1)
    rho = synth.phantom(grid, image, cfg["seed"] + 2).reshape(-1)
    o = orc.OracleReco(grid, samp, cfg["coils"], nthreads=nthreads)
    o.set_trajectory(traj)
    o.calculate_weights(np.where(smask, 1, 0).astype(np.complex64)[None, :])
    o.calculate_kernel_ft()
    o.set_maps(maps)
    img = maps.astype(np.complex128) * rho[None, :].astype(np.complex128)
    y = o.nufft_fwd(img, prec=64)
    data = synth.add_noise(y, smask, 1e-3, cfg["seed"] + 3)
    o.exclude_missing_data(data)
    return dict(grid=grid, samp=samp, ncoil=cfg["coils"], traj=traj, maps=maps, rho=rho, data=data, smask=smask,
                lines=lines, dcf=None, oracle=o, iters=cfg["iters"])


def gpu_context(pkg, P, tcfg=None):
    ctx = pkg.Context(P["grid"], P["samp"], P["ncoil"])
    if tcfg is not None:
        ctx.set_tile_config(tcfg)
    ctx.set_trajectory(P["traj"])
    ctx.calculate_weights()
    ctx.exclude_missing_data(P["data"])
    ctx.calculate_kernel_ft()
    ctx.set_density(P["dcf"])
    ctx.set_maps(P["maps"])
    return ctx
