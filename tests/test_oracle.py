"""Pins the CPU oracle against independent computations (no GPU needed).

The reference ships no tests, fixtures or golden outputs, so every oracle
function is checked against mathematics that fixes it: scipy's I0, numpy's
FFT and a brute-force DFT, closed forms of the KB kernel, a brute-force scan
of every grid point for the tap lists, dense matrices for E and E^H, the
adjoint identity and a textbook CG on the dense normal matrix.
"""
import numpy as np
import pytest
import scipy.special as sp

from conftest import rel_l2

W, B, ORDER = 4.25, 14.9086, 5


def f32(x):
    return np.float32(x)


# ----------------------------------------------------------------- I0 / KB
def test_i0_matches_scipy(orc):
    xs = np.concatenate([np.linspace(0, 16, 2001), [0.0, 1e-3, 14.9086, 20.0, 30.0]])
    got = np.array([orc.i0(x) for x in xs])
    np.testing.assert_allclose(got, sp.i0(xs), rtol=4e-15, atol=0)
    np.testing.assert_allclose(got, [orc.i0_std(x) for x in xs], rtol=4e-15, atol=0)
    assert orc.i0(0.0) == 1.0


def _kb_axis_numpy(dxf, n, w=W, b=B):
    """Reco.h:221-240 evaluated with numpy float32 casts and scipy's I0."""
    dkf = f32(1.0 / float(f32(n - 1)))
    wHalf = f32(float(f32(w)) / 2.0)
    if f32(dxf) > f32(wHalf * dkf):
        return 0.0
    q = f32(f32(2.0) * f32(dxf)) / f32(f32(w) * dkf)
    q = f32(q)
    arg = float(f32(b)) * np.sqrt(1.0 - float(q) * float(q))
    return sp.i0(arg) / float(f32(f32(f32(100.0) * f32(w)) * dkf))


@pytest.mark.parametrize("n", [24, 96, 128, 320, 512])
def test_kb_axis_closed_form(orc, n):
    dkf = float(f32(1.0 / float(f32(n - 1))))
    denom = float(f32(f32(f32(100.0) * f32(W)) * f32(dkf)))
    # peak: I0(b) / float(100 w dk)
    assert orc.kb_axis(0.0, n, W, B) == pytest.approx(sp.i0(float(f32(B))) / denom, rel=4e-15)
    # just inside the cut-off the square root vanishes: I0(0) / denom
    edge = float(f32(f32(2.125) * f32(dkf)))
    assert orc.kb_axis(edge, n, W, B) == pytest.approx(1.0 / denom, rel=1e-6)
    # beyond the cut-off: 0
    assert orc.kb_axis(float(np.nextafter(f32(edge), f32(1.0))), n, W, B) == 0.0
    assert orc.kb_axis(3.0 * dkf, n, W, B) == 0.0
    for d in np.linspace(0, 2.2, 23) * dkf:
        assert orc.kb_axis(d, n, W, B) == pytest.approx(_kb_axis_numpy(d, n), rel=1e-13, abs=0)


def test_conv_function_is_product_of_axes(orc):
    grid = (128, 96, 64)
    rng = np.random.default_rng(0)
    for _ in range(50):
        d = rng.uniform(0, 2.5, 3) / (np.array(grid) - 1)
        v = [orc.kb_axis(float(f32(d[a])), grid[a], W, B) for a in range(3)]
        assert orc.conv_function(d, grid, W, B) == f32((v[0] * v[1]) * v[2])


def test_sqrt_of_square_is_abs():
    """sqrt((k - g)^2) of Reco.cpp:767-771 equals |k - g| exactly, which the
    GPU setup kernel uses."""
    rng = np.random.default_rng(1)
    e = rng.uniform(-0.05, 0.05, 200000)
    e = np.concatenate([e, np.ldexp(rng.uniform(0.5, 1, 1000), -50)])
    np.testing.assert_array_equal(np.sqrt(e * e), np.abs(e))


# ----------------------------------------------------------- tap lists
def _bruteforce_taps(traj, grid, thr):
    """Every grid point of a small grid, distance and kernel as Reco.cpp:767-778."""
    nx, ny, nz = grid
    dk = [f32(1.0 / float(f32(n - 1))) for n in grid]
    out = []
    for k in traj:
        taps, wts = [], []
        ax = []
        for a, n in enumerate(grid):
            g = np.arange(n)
            e = k[a] - (g.astype(np.float32) * dk[a]).astype(np.float64)
            ax.append([_kb_axis_numpy(float(f32(np.sqrt(v * v))), n) for v in e])
        for x in range(nx):
            for y in range(ny):
                for z in range(nz):
                    wv = f32((ax[0][x] * ax[1][y]) * ax[2][z])
                    if float(wv) > thr:
                        taps.append(z + nz * y + ny * nz * x)
                        wts.append(wv)
        out.append((np.array(taps, np.int64), np.array(wts, np.float32)))
    return out


def test_weights_match_bruteforce_scan(orc):
    grid = (14, 12, 10)
    samp = (3, 2, 2)
    rng = np.random.default_rng(2)
    traj = rng.uniform(0.0, 1.0, (12, 3))
    traj[0] = [0.0, 0.0, 0.0]
    traj[1] = [1.0, 1.0, 1.0]
    traj[2] = [0.5, 0.5, 0.5]
    o = orc.OracleReco(grid, samp, 1)
    o.set_trajectory(traj)
    o.calculate_weights()
    counts, idx, w = o.get_weights()
    thr = o.kernel_threshold()
    zero = _kb_axis_numpy(0.0, grid[0]) * _kb_axis_numpy(0.0, grid[1]) * _kb_axis_numpy(0.0, grid[2])
    assert thr == pytest.approx(float(f32(zero)) * 1e-7, rel=1e-6)
    ref = _bruteforce_taps(traj, grid, thr)
    starts = np.concatenate([[0], np.cumsum(counts)])
    for i, (t, wt) in enumerate(ref):
        np.testing.assert_array_equal(idx[starts[i]:starts[i + 1]], t)
        np.testing.assert_allclose(w[starts[i]:starts[i + 1]], wt, rtol=2e-7)


def test_taps_per_interior_sample(orc):
    """60..84 kept taps per sample away from the grid edge for w=4.25, b=14.9086."""
    grid = (64, 64, 64)
    rng = np.random.default_rng(3)
    traj = rng.uniform(0.2, 0.8, (200, 3))
    o = orc.OracleReco(grid, (200, 1, 1), 1)
    o.set_trajectory(traj)
    o.calculate_weights()
    counts, _, _ = o.get_weights()
    assert counts.min() >= 60 and counts.max() <= 125
    assert 60 <= counts.mean() <= 84


@pytest.mark.parametrize("grid", [(128, 128, 96), (320, 320, 224), (256, 256, 192), (192, 192, 128),
                                  (512, 512, 320), (32, 32, 24)])
def test_flat_index_decomposition_is_exact(grid):
    """(int)(p * (1/(Ny*Nz))) of Reco.cpp:763-765 recovers (x, y, z) for every
    cell, so separable per-axis tables reproduce the reference's taps."""
    nx, ny, nz = grid
    invYZ = 1.0 / float(ny * nz)
    invZ = 1.0 / float(nz)
    yz = ny * nz
    for x0 in range(0, nx, 64):
        x = np.arange(x0, min(nx, x0 + 64))[:, None, None]
        y = np.arange(ny)[None, :, None]
        z = np.arange(nz)[None, None, :]
        p = (z + nz * y + yz * x).astype(np.int64)
        i0 = (p * invYZ).astype(np.int64)
        i1 = ((p - yz * i0) * invZ).astype(np.int64)
        i2 = p - nz * i1 - yz * i0
        assert np.all(i0 == x) and np.all(i1 == y) and np.all(i2 == z)


# --------------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [2, 3, 5, 7, 8, 12, 20, 28, 96, 224, 320, 11])
def test_fft1_matches_numpy(orc, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    np.testing.assert_allclose(orc.fft1(a, -1), np.fft.fft(a), rtol=0, atol=1e-13 * np.abs(a).sum())
    np.testing.assert_allclose(orc.fft1(a, +1), np.fft.ifft(a) * n, rtol=0, atol=1e-13 * np.abs(a).sum())


def test_centered_fft_bruteforce(orc):
    shape = (4, 6, 4)
    rng = np.random.default_rng(4)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    idx = np.indices(shape).reshape(3, -1).T
    s = (-1.0) ** idx.sum(1)
    # M F M with M = diag(-(-1)^(x+y+z)) applied on both sides (Fourier.cpp:45-72)
    ph = np.exp(-2j * np.pi * sum(np.outer(idx[:, k], idx[:, k]) / shape[k] for k in range(3)))
    ref = s * (ph @ (s * a.ravel()))
    np.testing.assert_allclose(orc.fft3_centered(a, -1).ravel(), ref, atol=1e-12)
    refb = s * (ph.conj() @ (s * a.ravel())) / np.sqrt(a.size)
    np.testing.assert_allclose(orc.fft3_centered(a, +1).ravel(), refb, atol=1e-12)


@pytest.mark.parametrize("shape", [(8, 12, 16), (16, 8, 24), (128, 128, 96)])
def test_centered_fft_is_shifted_fft(orc, shape):
    """With every N/2 even the centred transform is fftshift(fftn(ifftshift(.)))."""
    rng = np.random.default_rng(5)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    ref = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(a)))
    assert rel_l2(orc.fft3_centered(a, -1), ref) < 1e-14
    back = orc.fft3_centered(orc.fft3_centered(a, -1), +1)          # BWD(FWD(a)) = sqrt(N) a
    assert rel_l2(back, np.sqrt(a.size) * a) < 1e-14


def test_fp32_fft_mode_close_to_fp64(orc):
    shape = (32, 24, 40)
    rng = np.random.default_rng(6)
    a = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    assert rel_l2(orc.fft3_centered(a, -1), orc.fft3_centered(a.astype(np.complex128), -1)) < 2e-6


# ----------------------------------------------------------- deapodization
def test_kernel_ft_matches_numpy_idft(orc):
    grid = (16, 12, 20)
    o = orc.OracleReco(grid, (1, 1, 1), 1)
    o.calculate_kernel_ft()
    kft = o.get_kernel_ft()
    # sampled kernel on the grid about N/2 (Reco.cpp:571-585), float 3D product as the reference stores it
    ax = []
    for a, n in enumerate(grid):
        dk = 1.0 / (n - 1)
        m = np.arange(n) - n // 2
        ax.append(np.array([_kb_axis_numpy(float(f32(np.sqrt(dk * dk * v * v))), n) for v in m]))
    k3 = (ax[0][:, None, None] * ax[1][None, :, None] * ax[2][None, None, :]).astype(np.float32).astype(np.float64)
    # centred inverse DFT without scaling (Reco.cpp:589): fftshift(ifftn(ifftshift(.))) * N
    X = np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(k3))) * k3.size
    np.testing.assert_allclose(kft, 1.0 / X.real, rtol=2e-6)


# ------------------------------------------------------- NUFFT operators
def _dense_problem(orc, grid=(8, 8, 4), samp=(10, 4, 4), ncoil=2, seed=8):
    rng = np.random.default_rng(seed)
    ns = int(np.prod(samp))
    traj = rng.uniform(0.15, 0.85, (ns, 3))
    o = orc.OracleReco(grid, samp, ncoil)
    o.set_trajectory(traj)
    o.calculate_weights()
    o.calculate_kernel_ft()
    ng = int(np.prod(grid))
    maps = (rng.standard_normal((ncoil, ng)) + 1j * rng.standard_normal((ncoil, ng))).astype(np.complex64)
    o.set_maps(maps)
    # dense gridding matrix from the independent brute-force scan
    Gm = np.zeros((ns, ng))
    for i, (t, w) in enumerate(_bruteforce_taps(traj, grid, o.kernel_threshold())):
        Gm[i, t] = w.astype(np.float64)
    kft = o.get_kernel_ft().astype(np.float64).ravel()

    def F(v):   # centred unnormalized forward transform of one grid
        return np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(v.reshape(grid)))).ravel()

    def Fh(v):  # centred backward transform times 1/sqrt(N)
        return np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(v.reshape(grid)))).ravel() * ng / np.sqrt(ng)

    return o, traj, maps.astype(np.complex128), Gm, kft, F, Fh, ns, ng


def test_nufft_fwd_bwd_match_dense_operators(orc):
    o, traj, maps, Gm, kft, F, Fh, ns, ng = _dense_problem(orc)
    rng = np.random.default_rng(9)
    img = rng.standard_normal((2, ng)) + 1j * rng.standard_normal((2, ng))
    y = o.nufft_fwd(img, prec=64)
    for c in range(2):
        np.testing.assert_allclose(y[c], Gm @ F(kft * img[c]), atol=1e-10 * np.abs(y[c]).max())
    d = rng.standard_normal((2, ns)) + 1j * rng.standard_normal((2, ns))
    im = o.nufft_bwd(d, prec=64)
    for c in range(2):
        ref = kft * Fh(Gm.T @ d[c]) * 1e-3
        np.testing.assert_allclose(im[c], ref, atol=1e-10 * np.abs(ref).max())


def test_adjoint_identity(orc):
    """<E x, y> * 1e-3 / sqrt(Ng) = <x, E^H y> (scale convention of Reco.cpp:506-515)
    holds when the deapodization is real, as kernelFT is."""
    o, traj, maps, Gm, kft, F, Fh, ns, ng = _dense_problem(orc, seed=10)
    rng = np.random.default_rng(11)
    x = rng.standard_normal((2, ng)) + 1j * rng.standard_normal((2, ng))
    y = rng.standard_normal((2, ns)) + 1j * rng.standard_normal((2, ns))
    Ex = o.nufft_fwd(x, prec=64)
    Ehy = o.nufft_bwd(y, prec=64)
    lhs = np.vdot(y, Ex) * 1e-3 / np.sqrt(ng)
    rhs = np.vdot(Ehy, x)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_cg_matches_textbook_cg_on_dense_matrix(orc):
    o, traj, maps, Gm, kft, F, Fh, ns, ng = _dense_problem(orc, seed=12)
    ncoil = 2
    rng = np.random.default_rng(13)
    dcf = rng.uniform(0.5, 1.5, ns).astype(np.float32)
    o.set_density(dcf)
    data = (rng.standard_normal((ncoil, ns)) + 1j * rng.standard_normal((ncoil, ns))).astype(np.complex64)

    def E(v):
        return np.stack([Gm @ F(kft * maps[c] * v) for c in range(ncoil)])

    def Eh(y):
        return sum(np.conj(maps[c]) * kft * Fh(Gm.T @ y[c]) * 1e-3 for c in range(ncoil))

    Wd = dcf.astype(np.float64)
    b = Eh(Wd * data.astype(np.complex128))
    # textbook CG (Hestenes-Stiefel) with the reference's complex scalars
    x = np.zeros(ng, complex)
    r = b.copy()
    p = b.copy()
    res = []
    for it in range(12):
        q = Eh(Wd * E(p))
        rr = np.vdot(r, r)
        al = rr / np.vdot(p, q)
        x = x + al * p
        r = r - al * q
        res.append(np.linalg.norm(r))
        be = np.vdot(r, r) / rr
        p = r + be * p
    xo, reso, ito = o.cg_sense(data, thresh=0.0, max_iter=11, prec=64)
    assert ito == 12
    # float32 inputs (data * dcf rounded to float as performDensityCorrection
    # does) and last-ulp differences of the brute-force weights bound the match
    assert rel_l2(xo.ravel(), x) < 1e-6
    np.testing.assert_allclose(reso, res, rtol=1e-6)


def test_cg_converges_to_normal_equation_solution(orc):
    """On a small well-posed problem CG reaches A x = b (numpy solve)."""
    o, traj, maps, Gm, kft, F, Fh, ns, ng = _dense_problem(orc, grid=(8, 4, 4), samp=(30, 6, 6), ncoil=3, seed=14)
    ncoil = 3
    rng = np.random.default_rng(15)
    data = (rng.standard_normal((ncoil, ns)) + 1j * rng.standard_normal((ncoil, ns))).astype(np.complex64)
    A = np.zeros((ng, ng), complex)
    for j in range(ng):
        e = np.zeros(ng, complex)
        e[j] = 1
        A[:, j] = sum(np.conj(maps[c]) * kft * Fh(Gm.T @ (Gm @ F(kft * maps[c] * e))) * 1e-3 for c in range(ncoil))
    b = sum(np.conj(maps[c]) * kft * Fh(Gm.T @ data[c].astype(np.complex128)) * 1e-3 for c in range(ncoil))
    xs = np.linalg.solve(A, b)
    xo, reso, ito = o.cg_sense(data, thresh=0.0, max_iter=400, prec=64)
    assert rel_l2(xo.ravel(), xs) < 1e-6


def test_fp32_reference_mode_close_to_fp64(orc):
    from problems import small_problem

    P = small_problem(orc)
    x64, _, _ = P["oracle"].cg_sense(P["data"], 0.0, 9, prec=64)
    x32, _, _ = P["oracle"].cg_sense(P["data"], 0.0, 9, prec=32)
    assert rel_l2(x32, x64) < 1e-4


def test_atomic_coil_sum_is_the_same_sum(orc):
    """The reference's loop shape for the coil sum (float atomics, thread
    order, Reco.cpp:120-129, 257-262) adds the same terms as the fixed coil
    order: one application agrees to float rounding, in FP64 exactly up to
    the order of C terms; the CG image stays within 1e-4 of FP64; the
    per-iteration times cover the loop."""
    from problems import small_problem

    P = small_problem(orc, nthreads=4)
    o = P["oracle"]
    rng = np.random.default_rng(1)
    p = rng.standard_normal(o.ng) + 1j * rng.standard_normal(o.ng)
    q64 = o.apply_normal(p, prec=64)
    q32 = o.apply_normal(p.astype(np.complex64), prec=32)
    x64, _, _ = o.cg_sense(P["data"], 0.0, 9, prec=64)
    o.set_atomic_coil_sum(True)
    try:
        qa64 = o.apply_normal(p, prec=64)
        qa32 = o.apply_normal(p.astype(np.complex64), prec=32)
        xa, _, ita, loop_s, per = o.cg_sense(P["data"], 0.0, 9, prec=32, timed=True, iter_times=True)
    finally:
        o.set_atomic_coil_sum(False)
    assert rel_l2(qa64, q64) < 1e-14
    assert rel_l2(qa32, q32) < 1e-6
    assert rel_l2(xa, x64) < 1e-4
    assert ita == 10 and per.shape == (10,) and np.all(per > 0)
    assert abs(per.sum() - loop_s) <= 1e-6 + 1e-3 * loop_s


def test_exclude_missing_data_drops_zero_samples(orc):
    from problems import small_problem

    P = small_problem(orc)
    counts, _, _ = P["oracle"].get_weights()
    assert np.all(counts[P["data"][0] == 0] == 0)
    assert np.all(counts[P["data"][0] != 0] > 0)


def test_cg_on_empty_data_makes_one_nan_pass(orc):
    """Reco::cgSenseDynamic on a dataset without acquired samples: the
    right-hand side is 0, so rr = pq = 0 and complexDivision's 0/0
    (Reco.cpp:145-151) makes alpha NaN; x and r become NaN after the first
    pass and `NaN > threshold` is false, so the loop stops there
    (Reco.cpp:230, 276-309).  Both precisions of the oracle follow it."""
    from problems import small_problem

    P = small_problem(orc, seed=41)
    o = P["oracle"]
    d = np.zeros_like(P["data"])
    o.calculate_weights(d)
    counts, _, _ = o.get_weights()
    assert not np.any(counts)
    for prec in (64, 32):
        x, res, it = o.cg_sense(d, thresh=0.0, max_iter=6, prec=prec)
        assert it == 1 and res.shape == (1,) and np.isnan(res[0]) and np.isnan(x).all()
