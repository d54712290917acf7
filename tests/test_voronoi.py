"""Voronoi density compensation (estimateDensityV3D, functions.h:161-258).

CPU: the oracle's clipping statement against Qhull (scipy) restating the
reference's Delaunay-dual sum, and invariants (regular lattice, translation
and scale), for both rules: "hull" (the reference's: 0 only on the convex
hull) and "clipped" (the per-bin extension: 0 also for cells reaching beyond
the points' widened bounding box).  GPU: the device kernel against the oracle
bit for bit, and the per-bin DCF inside CG."""
import importlib

import numpy as np
import pytest

from conftest import rel_l2

synth = importlib.import_module("4d_lung_amd.synth")


def lattice(n=8):
    g = np.arange(n) / (n - 1.0)
    return np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)


@pytest.mark.parametrize("rule", ["hull", "clipped"])
def test_lattice_interior_cells_are_unit_cubes(orc, rule):
    p = lattice(7)
    vol = orc.voronoi_volumes(p, rule=rule)
    interior = np.all((p > 1e-9) & (p < 1 - 1e-9), axis=1)
    np.testing.assert_allclose(vol[interior], (1 / 6.0) ** 3, rtol=1e-12)
    assert np.all(vol[~interior] == 0)          # hull points: 0 (functions.h:148-150)
    d = orc.voronoi_density(p, rule=rule)
    assert np.all(d[interior] == 1.0) and np.all(d[~interior] == 0)


def _agree_with_qhull(a, b, rtol):
    """Hull rule against the Qhull dual sum: the same zero set (the points
    on the convex hull) and the same finite volumes."""
    np.testing.assert_array_equal(a == 0, b == 0)
    np.testing.assert_allclose(a[b > 0], b[b > 0], rtol=rtol)


def test_clipping_matches_qhull_dual_sum(orc):
    from oracle import voronoi_ref

    rng = np.random.default_rng(0)
    p = rng.uniform(0, 1, (2500, 3))
    b = voronoi_ref.voronoi_volumes(p)
    _agree_with_qhull(orc.voronoi_volumes(p, rule="hull"), b, 1e-9)
    a = orc.voronoi_volumes(p, rule="clipped")
    both = (a > 0) & (b > 0)
    np.testing.assert_allclose(a[both], b[both], rtol=1e-10)
    # clipped rule: every hull point (unbounded cell) is 0 here too; the
    # extra zeros are the near-boundary cells that reach beyond the widened
    # bounding box
    assert np.all(a[b == 0] == 0)
    extra = (a == 0) & (b > 0)
    assert np.all(np.min(np.minimum(p, 1 - p)[extra], 1) < 0.15)


def test_wave_caipi_interior_matches_qhull(orc):
    from oracle import voronoi_ref

    grid, samp = (64, 48, 40), (64, 24, 20)
    traj = synth.wave_caipi_trajectory(grid, samp, 3, 4.0)
    sm = synth.sample_mask_from_lines(synth.caipi_mask(samp[1], samp[2], (2, 2), 1), samp[0])
    p = traj[sm]
    b = voronoi_ref.voronoi_volumes(p)
    _agree_with_qhull(orc.voronoi_volumes(p, rule="hull"), b, 1e-9)
    a = orc.voronoi_volumes(p, rule="clipped")
    both = (a > 0) & (b > 0)
    assert both.sum() > 0.85 * len(p)
    assert np.all(a[b == 0] == 0)
    np.testing.assert_allclose(a[both], b[both], rtol=1e-10)


@pytest.mark.parametrize("grid,samp", [((64, 64, 48), (64, 32, 24)), ((48, 48, 40), (48, 24, 20))])
def test_full_trajectory_hull_rule_matches_qhull(orc, grid, samp):
    """The reference's own Voronoi input (every point of a wave-CAIPI
    trajectory, acquired or not, CGSense.cpp:171-174) with its rule: zeros
    exactly on Qhull's unbounded (convex-hull) cells, every other volume
    finite -- up to about 300 times the points' extent for the cells just
    inside the hull -- and equal to Qhull's dual sum; the normalised weights
    (divided by the largest of those near-hull cells, functions.h:234-245)
    equal the restatement's.  The agreement is 1e-9 relative rather than
    1e-10: the trajectory's cospherical lattice points give Qhull
    near-duplicate circumcentres."""
    from oracle import voronoi_ref

    p = synth.wave_caipi_trajectory(grid, samp, 3, 4.0)
    a = orc.voronoi_volumes(p, rule="hull")
    b = voronoi_ref.voronoi_volumes(p)
    _agree_with_qhull(a, b, 1e-9)
    assert 0.03 < (a == 0).mean() < 0.07
    # the cells just inside the hull are kept, not zeroed as by the clipped rule
    kept = (a > 0) & (orc.voronoi_volumes(p, rule="clipped") == 0)
    assert kept.mean() > 0.01 and a[kept].min() > np.median(a[a > 0])
    np.testing.assert_allclose(orc.voronoi_density(p, rule="hull"), voronoi_ref.normalise(b), rtol=2e-7, atol=0)


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
    _agree_with_qhull(orc.voronoi_volumes(p, rule="hull"), voronoi_ref.voronoi_volumes(p), 1e-9)
    a = orc.voronoi_volumes(p, rule="clipped")
    b = voronoi_ref.voronoi_volumes(p)
    both = (a > 0) & (b > 0)
    np.testing.assert_allclose(a[both], b[both], rtol=1e-10)
    assert np.all(a[b == 0] == 0)
    core = np.linalg.norm(p - 0.5, axis=1) < 0.1
    assert np.all(a[core] > 0)


def test_sparse_bin_cells_fit_the_large_limits(orc):
    p = sparse_bin_points()
    vol = orc.voronoi_volumes(p, rule="clipped")   # raises if a closed cell outgrew the limits
    assert (vol > 0).mean() > 0.75


@pytest.mark.parametrize("rule", ["hull", "clipped"])
def test_density_invariant_to_translation_and_scale(orc, rule):
    rng = np.random.default_rng(1)
    p = rng.uniform(0, 1, (1500, 3))
    d0 = orc.voronoi_density(p, rule=rule)
    d1 = orc.voronoi_density(p * 0.5 + np.array([0.1, 0.2, 0.3]), rule=rule)
    np.testing.assert_allclose(d1, d0, rtol=1e-9, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("case,rule", [("uniform", "clipped"), ("lattice", "clipped"), ("wave", "clipped"),
                                       ("sparse_bin", "clipped"), ("uniform", "hull"), ("lattice", "hull"),
                                       ("wave", "hull"), ("full_trajectory", "hull")])
def test_device_voronoi_bit_exact(pkg, orc, case, rule):
    if case == "uniform":
        p = np.random.default_rng(2).uniform(0, 1, (20000, 3))
    elif case == "lattice":
        p = lattice(10)
    elif case == "sparse_bin":
        p = sparse_bin_points()
    elif case == "full_trajectory":
        p = synth.wave_caipi_trajectory((64, 64, 48), (64, 32, 24), 3, 4.0)
    else:
        grid, samp = (128, 128, 96), (128, 64, 48)
        traj = synth.wave_caipi_trajectory(grid, samp, 3, 4.0)
        p = traj[synth.sample_mask_from_lines(synth.caipi_mask(samp[1], samp[2], (2, 2), 1), samp[0])]
    print("device Voronoi %s, %s rule, %d points" % (case, rule, len(p)), flush=True)
    got = pkg.cgs.voronoi_volumes(p, rule=rule)
    ref = orc.voronoi_volumes(p, rule=rule)
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(pkg.cgs.voronoi_density(p, rule=rule), orc.voronoi_density(p, rule=rule))


@pytest.mark.gpu
@pytest.mark.parametrize("tier", [1, 2])
def test_device_voronoi_every_tier_bit_exact(pkg, orc, tier, monkeypatch):
    """Cells built from the start with the medium (LDS) or the large (device
    memory) limits give the same bits as the oracle: a cell's bits do not
    depend on the limits it is built with."""
    monkeypatch.setenv("CGS_VORONOI_FIRST_TIER", str(tier))
    for case, p, rule in (("sparse_bin", sparse_bin_points()[:6000], "clipped"), ("lattice", lattice(8), "hull")):
        print("device Voronoi %s, %s rule, %d points, first tier %d" % (case, rule, len(p), tier), flush=True)
        np.testing.assert_array_equal(pkg.cgs.voronoi_volumes(p, rule=rule), orc.voronoi_volumes(p, rule=rule))


@pytest.mark.gpu
@pytest.mark.parametrize("wl,b", [("4d_128x128x96", 7), ("4d_256x256x160", 0)])
def test_device_voronoi_workload_bin_bit_exact(pkg, orc, wl, b):
    """Per-bin weights of the respiratory-resolved workloads (the clipped rule
    on one bin's acquired samples): long thin cells, a tenth of them built
    again with the medium limits on the 256x256x160 workload."""
    cfg = synth.CONFIGS[wl]
    grid, samp = cfg["grid"], cfg["samp"]
    traj = synth.wave_caipi_trajectory(grid, samp, cfg["cycles"], cfg["amp"])
    lines = synth.caipi_mask(samp[1], samp[2], cfg["R"], cfg["shift"])
    labels = synth.respiratory_bins(lines, cfg["bins"], cfg["bin_frac"], cfg["seed"] + 5)
    p = np.ascontiguousarray(traj[synth.sample_mask_from_lines(labels == b, samp[0])])
    print("device Voronoi %s bin %d, %d points" % (wl, b, len(p)), flush=True)
    got = pkg.cgs.voronoi_volumes(p, rule="clipped")
    np.testing.assert_array_equal(got, orc.voronoi_volumes(p, rule="clipped"))
    assert 0.7 < (got > 0).mean() < 0.95


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
    w[act] = orc.voronoi_density(P["traj"][act], rule="clipped")   # the per-bin rule
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
    for rule in ("hull", "clipped"):
        vol = orc.voronoi_volumes(p, rule=rule)
        for cutoff in (1.0, 0.75):
            print("density averaging: %s rule, cutoff %.2f" % (rule, cutoff), flush=True)
            got = pkg.cgs.voronoi_density(p, average=True, sx=samp[0], cutoff=cutoff, rule=rule)
            np.testing.assert_array_equal(got, orc.density_from_volumes(vol, True, samp[0], cutoff))
