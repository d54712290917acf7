"""Voronoi density compensation (estimateDensityV3D, functions.h:161-258).

CPU: the oracle's clipping statement against Qhull (scipy) restating the
reference's Delaunay-dual sum, and invariants (regular lattice, translation
and scale).  GPU: the device kernel against the oracle bit for bit, and the
per-bin DCF inside CG."""
import importlib

import numpy as np
import pytest

from conftest import rel_l2

synth = importlib.import_module("4d_lung_amd.synth")


def lattice(n=8):
    g = np.arange(n) / (n - 1.0)
    return np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)


def test_lattice_interior_cells_are_unit_cubes(orc):
    p = lattice(7)
    vol = orc.voronoi_volumes(p)
    interior = np.all((p > 1e-9) & (p < 1 - 1e-9), axis=1)
    np.testing.assert_allclose(vol[interior], (1 / 6.0) ** 3, rtol=1e-12)
    assert np.all(vol[~interior] == 0)          # hull points: 0 (functions.h:148-150)
    d = orc.voronoi_density(p)
    assert np.all(d[interior] == 1.0) and np.all(d[~interior] == 0)


def test_clipping_matches_qhull_dual_sum(orc):
    from oracle import voronoi_ref

    rng = np.random.default_rng(0)
    p = rng.uniform(0, 1, (2500, 3))
    a = orc.voronoi_volumes(p)
    b = voronoi_ref.voronoi_volumes(p)
    both = (a > 0) & (b > 0)
    np.testing.assert_allclose(a[both], b[both], rtol=1e-10)
    # every hull point (unbounded cell) is 0 here too; the extra zeros are the
    # near-boundary cells that reach beyond the widened bounding box
    assert np.all(a[b == 0] == 0)
    extra = (a == 0) & (b > 0)
    assert np.all(np.min(np.minimum(p, 1 - p)[extra], 1) < 0.15)


def test_wave_caipi_interior_matches_qhull(orc):
    from oracle import voronoi_ref

    grid, samp = (64, 48, 40), (64, 24, 20)
    traj = synth.wave_caipi_trajectory(grid, samp, 3, 4.0)
    sm = synth.sample_mask_from_lines(synth.caipi_mask(samp[1], samp[2], (2, 2), 1), samp[0])
    p = traj[sm]
    a = orc.voronoi_volumes(p)
    b = voronoi_ref.voronoi_volumes(p)
    both = (a > 0) & (b > 0)
    assert both.sum() > 0.85 * len(p)
    assert np.all(a[b == 0] == 0)
    np.testing.assert_allclose(a[both], b[both], rtol=1e-10)


def sparse_bin_points(n_lines=None):
    """Samples of the sparsest respiratory bin of the 128x128x96 4D workload
    (7 % of the R=2x2 lines): long, thin cells with many faces, which outgrow
    the small cell limits and are rebuilt with the large ones."""
    cfg = synth.CONFIGS["4d_128x128x96"]
    grid, samp = cfg["grid"], cfg["samp"]
    traj = synth.wave_caipi_trajectory(grid, samp, cfg["cycles"], cfg["amp"])
    lines = synth.caipi_mask(samp[1], samp[2], cfg["R"], cfg["shift"])
    labels = synth.respiratory_bins(lines, cfg["bins"], cfg["bin_frac"], cfg["seed"] + 5)
    b = int(np.argmin(np.bincount(labels[labels >= 0])))
    return traj[synth.sample_mask_from_lines(labels == b, samp[0])]


def test_dense_cluster_interior_matches_qhull(orc):
    # a dense cluster inside a uniform background: the cluster's interior
    # cells outgrow the small limits while still reaching the bounding box
    # early in their construction; built again with the large limits they
    # must come out closed and equal to Qhull's volumes
    from oracle import voronoi_ref

    rng = np.random.default_rng(0)
    p = np.concatenate([rng.normal(0.5, 0.05, (5000, 3)), rng.uniform(0, 1, (3000, 3))])
    a = orc.voronoi_volumes(p)
    b = voronoi_ref.voronoi_volumes(p)
    both = (a > 0) & (b > 0)
    np.testing.assert_allclose(a[both], b[both], rtol=1e-10)
    assert np.all(a[b == 0] == 0)
    core = np.linalg.norm(p - 0.5, axis=1) < 0.1
    assert np.all(a[core] > 0)


def test_sparse_bin_cells_fit_the_large_limits(orc):
    p = sparse_bin_points()
    vol = orc.voronoi_volumes(p)          # raises if a closed cell outgrew the limits
    assert (vol > 0).mean() > 0.75


def test_density_invariant_to_translation_and_scale(orc):
    rng = np.random.default_rng(1)
    p = rng.uniform(0, 1, (1500, 3))
    d0 = orc.voronoi_density(p)
    d1 = orc.voronoi_density(p * 0.5 + np.array([0.1, 0.2, 0.3]))
    np.testing.assert_allclose(d1, d0, rtol=1e-9, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["uniform", "lattice", "wave", "sparse_bin"])
def test_device_voronoi_bit_exact(pkg, orc, case):
    if case == "uniform":
        p = np.random.default_rng(2).uniform(0, 1, (20000, 3))
    elif case == "lattice":
        p = lattice(10)
    elif case == "sparse_bin":
        p = sparse_bin_points()
    else:
        grid, samp = (128, 128, 96), (128, 64, 48)
        traj = synth.wave_caipi_trajectory(grid, samp, 3, 4.0)
        p = traj[synth.sample_mask_from_lines(synth.caipi_mask(samp[1], samp[2], (2, 2), 1), samp[0])]
    got = pkg.cgs.voronoi_volumes(p)
    ref = orc.voronoi_volumes(p)
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(pkg.cgs.voronoi_density(p), orc.voronoi_density(p))


@pytest.mark.gpu
def test_bin_voronoi_dcf_in_cg(pkg, orc):
    from problems import gpu_context, small_problem

    P = small_problem(orc, seed=41)
    ctx = gpu_context(pkg, P)
    ctx.set_bin_data(0, P["data"])
    ctx.set_bin_voronoi(0)
    x, _, _ = ctx.cg_sense_bins(mode=1, thresh=0.0, max_iter=7)
    act = np.flatnonzero(P["data"][0] != 0)
    w = np.zeros(P["data"].shape[1], np.float32)
    w[act] = orc.voronoi_density(P["traj"][act])
    P["oracle"].set_density(w)
    xr, _, _ = P["oracle"].cg_sense(P["data"], thresh=0.0, max_iter=7, prec=64)
    P["oracle"].set_density(None)
    assert rel_l2(x[0], xr) < 1e-4


def test_density_averaging_closed_form(orc):
    """The averaging and normalisation of estimateDensityV3D
    (functions.h:212-246) on volumes exact in float: every sample of class
    k = i % sX has volume k + 1, so the class sums are n_k (k + 1) exactly;
    each sample takes its class's sum (classes above (sX - 1) * cutoff take
    the cutoff class's), over the total, and after the division by the
    maximum the weights are s[k'] / max s[k'] up to float rounding.  Without
    averaging the weights are the volumes over their maximum."""
    sx = 8
    i = np.arange(8 * 37 + 5)                 # ragged: the first 5 classes have one sample more
    k = i % sx
    vol = (k + 1).astype(np.float64)
    s = np.bincount(k, minlength=sx) * (np.arange(sx) + 1.0)
    for cutoff in (1.0, 0.5):
        kk = np.minimum(k, int((sx - 1) * cutoff))
        d = orc.density_from_volumes(vol, average=True, sx=sx, cutoff=cutoff)
        np.testing.assert_allclose(d, s[kk] / s[kk].max(), rtol=3e-7)
        assert d.max() == 1.0
    d = orc.density_from_volumes(vol)
    np.testing.assert_allclose(d, vol / vol.max(), rtol=1e-7)
    assert d.max() == 1.0


@pytest.mark.gpu
def test_device_density_averaging_bit_exact(pkg, orc):
    """estimateDensityV3D with per-readout averaging on the device
    (cgs_voronoi_density) equals the oracle's statement bit for bit, on every
    point of a wave-CAIPI trajectory (the reference's own input: all lines)."""
    grid, samp = (64, 64, 48), (64, 32, 24)
    p = synth.wave_caipi_trajectory(grid, samp, 3, 4.0)
    vol = orc.voronoi_volumes(p)
    for cutoff in (1.0, 0.75):
        print("density averaging: cutoff %.2f" % cutoff, flush=True)
        got = pkg.cgs.voronoi_density(p, average=True, sx=samp[0], cutoff=cutoff)
        np.testing.assert_array_equal(got, orc.density_from_volumes(vol, True, samp[0], cutoff))
