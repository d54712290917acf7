"""Seeded synthetic inputs for the CG-SENSE workloads.

The reference reconstructs measured wave-CAIPI lung data that is not available
offline (Readme:5), so every workload here is synthetic and written in the
reference's in-memory layouts:

* trajectory: ``kSpaceTraj`` of Trajectory.h:32, ``nS x 3`` doubles in [0, 1]
  per axis (grid point g sits at g/(N-1), Reco.cpp:712-714), sample index
  ``i = ix*(sY*sZ) + iy*sZ + iz`` with the readout ix slowest
  (Trajectory.cpp:167-175);
* k-space data: ``C x nS`` complex64, zero-filled where a line is not acquired
  (the zeros are how excludeMissingData, Reco.cpp:681-696, sees undersampling
  and bin membership);
* coil maps: ``C x Ng`` complex64 on the oversampled grid (Reco.cpp:450-456);
* images: ``Ng`` complex64 ``[x][y][z]`` with z fastest.

Positions are in grid cells before normalisation: the readout runs along x
with one sample per grid cell-ish spacing, phase-encode lines sit on a
Cartesian (ky, kz) lattice, and the wave-CAIPI helix adds ``amp*sin`` to ky and
``amp*cos`` to kz along the readout (``cycles`` periods per readout).  A border
of ``border + amp`` cells keeps every kernel footprint on the grid, like the
42-point border of Trajectory.cpp:144-164.
"""
import numpy as np

# Workload table.  image = reconstructed FOV, grid = 2x oversampled grid,
# samp = (readout samples, phase-encode lines y, phase-encode lines z).
CONFIGS = {
    # Plain CG-SENSE, CPU-runnable: 64x64x48 image, 8 coils, R=2x2 CAIPI shift 1,
    # 128 samples/line, 3 wave cycles, amplitude 4, 10 CG iterations.
    "plain_64x64x48": dict(image=(64, 64, 48), grid=(128, 128, 96), samp=(128, 64, 48), coils=8,
                           R=(2, 2), shift=1, cycles=3, amp=4.0, iters=10, bins=1, seed=1001),
    # Single-frame full size: 160x160x112 image, 16 coils, R=3x2, 320 samples/line,
    # 6 wave cycles, 20 CG iterations.
    "single_160x160x112": dict(image=(160, 160, 112), grid=(320, 320, 224), samp=(320, 160, 112), coils=16,
                               R=(3, 2), shift=1, cycles=6, amp=4.0, iters=20, bins=1, seed=1002),
    # Batched 4D respiratory-resolved: 128x128x96 image, 16 coils, R=2x2, 8 bins
    # holding 6-20 % of lines each, 15 CG iterations.
    # Every bin is checked against the FP64 oracle at the workload's own 15
    # iterations (about 20 s of oracle time per bin on 16 host cores).
    "4d_128x128x96": dict(image=(128, 128, 96), grid=(256, 256, 192), samp=(256, 128, 96), coils=16,
                          R=(2, 2), shift=1, cycles=3, amp=4.0, iters=15, bins=8, bin_frac=(0.06, 0.20),
                          seed=1003, dcf="voronoi", check_iters=15, check_bins="all"),
    # Differentiable unrolled recon: 96x96x64 image, 8 coils, R=2x2, 5 unrolled iterations.
    "unrolled_96x96x64": dict(image=(96, 96, 64), grid=(192, 192, 128), samp=(192, 96, 64), coils=8,
                              R=(2, 2), shift=1, cycles=3, amp=4.0, iters=5, bins=1, seed=1004),
    # Large 4D precision stress: 256x256x160 image, 32 coils, 20 bins of 4-9 % of
    # lines each, R=3x3, 512 samples/line, per-bin Voronoi DCF, 30 CG
    # iterations.  check_iters: the FP64 oracle's solve of a bin takes about
    # half a minute per iteration on 16 host cores, so its accuracy check runs
    # 4 iterations on both sides.
    "4d_256x256x160": dict(image=(256, 256, 160), grid=(512, 512, 320), samp=(512, 256, 160), coils=32,
                           R=(3, 3), shift=1, cycles=6, amp=4.0, iters=30, bins=20, bin_frac=(0.04, 0.09),
                           seed=1005, check_iters=4, check_bins="0", dcf="voronoi"),
}


def wave_caipi_trajectory(grid, samp, cycles, amp, border=4.0, phase_y=0.0, phase_z=0.0):
    """Wave-CAIPI sample positions, ``nS x 3`` float64 in [0, 1], memory order.

    Returns the trajectory for every phase-encode line (acquired or not), as
    the reference's WaveCaipi::calculateTraj does (Trajectory.cpp:122-127).
    """
    nx, ny, nz = grid
    sx, sy, sz = samp
    ix = np.arange(sx, dtype=np.float64)
    iy = np.arange(sy, dtype=np.float64)
    iz = np.arange(sz, dtype=np.float64)
    bx = float(border)
    by = float(border) + float(amp)
    bz = float(border) + float(amp)
    gx = bx + ix * (nx - 1 - 2 * bx) / (sx - 1)
    py = by + iy * (ny - 1 - 2 * by) / (sy - 1)
    pz = bz + iz * (nz - 1 - 2 * bz) / (sz - 1)
    ph = 2.0 * np.pi * cycles * ix / sx
    wy = amp * np.sin(ph + phase_y)
    wz = amp * np.cos(ph + phase_z)
    k = np.empty((sx, sy, sz, 3), np.float64)
    k[..., 0] = (gx / (nx - 1))[:, None, None]
    k[..., 1] = ((py[None, :] + wy[:, None]) / (ny - 1))[:, :, None]
    k[..., 2] = ((pz[None, :] + wz[:, None]) / (nz - 1))[:, None, :]
    return k.reshape(-1, 3)


def caipi_mask(sy, sz, R, shift):
    """Acquired (ky, kz) lines of a 2D-CAIPIRINHA pattern with acceleration
    R = (Ry, Rz): every Rz-th kz line, and on it every Ry-th ky line, shifted by
    ``shift`` per acquired kz line."""
    ry, rz = R
    iy = np.arange(sy)[:, None]
    iz = np.arange(sz)[None, :]
    return (iz % rz == 0) & ((iy - (iz // rz) * shift) % ry == 0)


def phantom(grid, image, seed, n_ellipsoids=10):
    """Seeded complex 3D ellipsoid phantom centred in the grid (complex64)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = grid
    fx, fy, fz = image
    x = (np.arange(nx, dtype=np.float32) - nx // 2) / (fx / 2)
    y = (np.arange(ny, dtype=np.float32) - ny // 2) / (fy / 2)
    z = (np.arange(nz, dtype=np.float32) - nz // 2) / (fz / 2)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij", sparse=True)
    mag = np.zeros(grid, np.float32)
    for e in range(n_ellipsoids):
        c = rng.uniform(-0.35, 0.35, 3).astype(np.float32)
        ax = rng.uniform(0.08, 0.45, 3).astype(np.float32)
        if e == 0:
            c[:] = 0
            ax[:] = (0.85, 0.8, 0.85)
        th = rng.uniform(0, np.pi, 3)
        cz, sz_ = np.cos(th[0]), np.sin(th[0])
        cy, sy_ = np.cos(th[1]), np.sin(th[1])
        Rz = np.array([[cz, -sz_, 0], [sz_, cz, 0], [0, 0, 1]])
        Ry = np.array([[cy, 0, sy_], [0, 1, 0], [-sy_, 0, cy]])
        Rm = (Rz @ Ry).astype(np.float32)
        dx, dy, dz = X - c[0], Y - c[1], Z - c[2]
        u = (Rm[0, 0] * dx + Rm[0, 1] * dy + Rm[0, 2] * dz) / ax[0]
        v = (Rm[1, 0] * dx + Rm[1, 1] * dy + Rm[1, 2] * dz) / ax[1]
        w = (Rm[2, 0] * dx + Rm[2, 1] * dy + Rm[2, 2] * dz) / ax[2]
        inside = (u * u + v * v + w * w) <= 1.0
        amp = 1.0 if e == 0 else rng.uniform(-0.4, 0.6)
        mag += np.float32(amp) * inside
    ph = rng.uniform(-1.0, 1.0, 3).astype(np.float32)
    phase = ph[0] * X + ph[1] * Y + ph[2] * Z
    img = (mag * np.exp(1j * phase)).astype(np.complex64)
    return img


def _interp_matrix(n_out, n_in):
    """Linear interpolation matrix from n_in coarse samples spanning the
    grid to n_out fine samples."""
    pos = np.linspace(0, n_in - 1, n_out)
    i0 = np.clip(np.floor(pos).astype(int), 0, n_in - 2)
    f = pos - i0
    M = np.zeros((n_out, n_in), np.float32)
    M[np.arange(n_out), i0] = 1 - f
    M[np.arange(n_out), i0 + 1] += f
    return M


def coil_maps(grid, ncoil, seed, coarse=8):
    """Gaussian-smoothed random complex coil maps normalised so that
    sum_c |S_c|^2 = 1 at every voxel; complex64 ``C x Ng``.

    Complex Gaussian noise on a ``coarse^3`` lattice is smoothed with a
    Gaussian (sigma = coarse/6 lattice points, i.e. about N/6 voxels) and
    interpolated linearly onto the grid.
    """
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    nx, ny, nz = grid
    Mx, My, Mz = (_interp_matrix(n, coarse) for n in (nx, ny, nz))
    maps = np.empty((ncoil,) + tuple(grid), np.complex64)
    for c in range(ncoil):
        g = rng.standard_normal((coarse,) * 3) + 1j * rng.standard_normal((coarse,) * 3)
        g = gaussian_filter(g.real, coarse / 6.0, mode="nearest") + 1j * gaussian_filter(
            g.imag, coarse / 6.0, mode="nearest")
        g = g.astype(np.complex64)
        t = np.tensordot(Mx, g, axes=(1, 0))            # nx, c, c
        t = np.tensordot(t, My, axes=(1, 1))            # nx, c, ny
        t = np.tensordot(t, Mz, axes=(1, 1))            # nx, ny, nz
        maps[c] = t
    ss = np.zeros(tuple(grid), np.float32)
    for c in range(ncoil):
        ss += maps[c].real ** 2 + maps[c].imag ** 2
    inv = (1.0 / np.sqrt(ss)).astype(np.float32)
    for c in range(ncoil):
        maps[c] *= inv
    return maps.reshape(ncoil, -1)


def respiratory_bins(line_mask, nbins, frac_range, seed, max_tries=1000):
    """Assign every acquired line to a respiratory bin.

    Lines are acquired in a seeded random order, one every TR; a 0.25 Hz
    sinusoidal trace plus Gaussian jitter is sampled at each line, and bin
    edges are amplitude quantiles at Dirichlet-drawn cumulative fractions,
    redrawn until every bin holds a fraction of the lines inside
    ``frac_range``.  Returns an int32 ``(sY, sZ)`` array: bin index, or -1 for
    lines that are not acquired.
    """
    rng = np.random.default_rng(seed)
    acq = np.flatnonzero(line_mask.ravel())
    order = rng.permutation(acq)
    tr = 0.01
    t = np.arange(order.size) * tr
    trace = np.sin(2 * np.pi * 0.25 * t) + 0.15 * rng.standard_normal(order.size)
    sorted_tr = np.sort(trace)
    lo, hi = frac_range
    # Dirichlet concentration: 8 for up to 8 bins; for more bins the fractions
    # must cluster tighter around 1/nbins to fit a narrow range
    conc = 8.0 if nbins <= 8 else 40.0
    for _ in range(max_tries):
        fr = rng.dirichlet(np.full(nbins, conc))
        if fr.min() < lo or fr.max() > hi:
            continue
        cuts = np.cumsum(fr)[:-1]
        edges = sorted_tr[np.minimum((cuts * order.size).astype(int), order.size - 1)]
        b = np.searchsorted(edges, trace, side="right")
        counts = np.bincount(b, minlength=nbins) / order.size
        if counts.min() >= lo and counts.max() <= hi:
            out = np.full(line_mask.size, -1, np.int32)
            out[order] = b
            return out.reshape(line_mask.shape)
    raise RuntimeError("could not draw respiratory bins within %s" % (frac_range,))


def sample_mask_from_lines(line_sel, sx):
    """Expand a (sY, sZ) boolean line selection to the nS sample mask."""
    return np.broadcast_to(line_sel[None, :, :], (sx,) + line_sel.shape).reshape(-1)


def add_noise(y, sample_mask, sigma_rel, seed):
    """Zero unacquired samples and add complex Gaussian noise of
    sigma = sigma_rel * max|y| on the acquired ones (complex64 ``C x nS``)."""
    rng = np.random.default_rng(seed)
    y = np.array(y, dtype=np.complex64, copy=True)
    m = np.asarray(sample_mask, bool)
    sig = sigma_rel * float(np.abs(y[:, m]).max())
    n = m.sum()
    for c in range(y.shape[0]):
        noise = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) * sig
        y[c, m] += noise.astype(np.complex64)
        y[c, ~m] = 0
    # An acquired sample must not be exactly zero in coil 0, or it would be
    # taken for a missing one (Reco.cpp:687).
    z = m & (y[0] == 0)
    y[0, z] = np.complex64(sig)
    return y


def binned_workload(pkg, cfg, log=None, pruning=-1):
    """Inputs of a respiratory-resolved workload on the device.

    The phantom's k-space is formed with the library's own forward operator E
    on every acquired sample (a context ``ctx`` of the workload's grid,
    samples and coils with trajectory, kernelFT and maps set), noise is added,
    and the acquired lines are split into the seeded respiratory bins; every
    bin keeps its samples in compact form (original indices, ``C x m``
    values) and, for ``dcf="voronoi"``, the Voronoi weights of its own
    samples (cgs_voronoi_density_rule on the device with the clipped rule:
    cells reaching beyond the bin's bounding box get 0).  Returns a dict with ``ctx``,
    ``traj``, ``maps``, ``labels`` and ``bins`` = [(idx, values, weights)] and
    ``voronoi_seconds`` (None without density weights).
    """
    import time

    say = log or (lambda _m: None)
    grid, samp, image, C, B = cfg["grid"], cfg["samp"], cfg["image"], cfg["coils"], cfg["bins"]
    ns = int(np.prod(samp))
    t0 = time.time()
    traj = wave_caipi_trajectory(grid, samp, cfg["cycles"], cfg["amp"])
    lines = caipi_mask(samp[1], samp[2], cfg["R"], cfg["shift"])
    labels = respiratory_bins(lines, B, cfg["bin_frac"], cfg["seed"] + 5)
    smask = sample_mask_from_lines(lines, samp[0])
    maps = coil_maps(grid, C, cfg["seed"] + 1)
    rho = phantom(grid, image, cfg["seed"] + 2).reshape(-1)
    say("synthetic image and maps %.1fs (%d lines in %d bins)" % (time.time() - t0, int(lines.sum()), B))
    ctx = pkg.Context(grid, samp, C)
    if pruning >= 0:
        ctx.set_pruning(pruning)
    ctx.set_trajectory(traj)
    ctx.calculate_kernel_ft()
    ctx.set_maps(maps)
    mask_data = np.zeros((1, ns), np.complex64)
    mask_data[0, smask] = 1
    ctx.exclude_missing_data(mask_data)
    del mask_data
    y = ctx.nufft_fwd(maps * rho[None, :])
    data = add_noise(y, smask, 1e-3, cfg["seed"] + 3)
    del y
    say("k-space synthesized %.1fs" % (time.time() - t0))
    bins = []
    tv = time.time()
    voronoi = cfg.get("dcf", "voronoi") == "voronoi"
    for b in range(B):
        idx = np.flatnonzero(sample_mask_from_lines(labels == b, samp[0])).astype(np.int32)
        w = pkg.cgs.voronoi_density(traj[idx], rule="clipped") if voronoi else None
        bins.append((idx, np.ascontiguousarray(data[:, idx]), w))
        if voronoi:
            say("bin %d: %d samples, Voronoi weights after %.1fs" % (b, idx.size, time.time() - tv))
    counts = [int(bb[0].size) for bb in bins]
    voronoi_s = time.time() - tv if voronoi else None
    if voronoi:
        say("per-bin Voronoi density weights: %.1fs for %d samples (bins of %d..%d)"
            % (voronoi_s, sum(counts), min(counts), max(counts)))
    else:
        say("bins of %d..%d samples, no density weights" % (min(counts), max(counts)))
    return dict(ctx=ctx, traj=traj, maps=maps, labels=labels, bins=bins, voronoi_seconds=voronoi_s)
