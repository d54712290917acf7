"""Differentiable operators (4d_lung_amd.differentiable): E / E^H on device
tensors with gradients in the image, the per-sample k-space coordinates and
the wave-CAIPI parameters, and the unrolled, checkpointed CG-SENSE.

Gradients are checked against FP64 central finite differences of the oracle
(step 1e-4 in the coordinates and wave parameters) and, for the image,
against the E^H-based gradient 2 E^*(E x - y) computed by the oracle.  The
reference has no gradients; its E, E^H and CG (Reco.cpp:175-553) define the
functions being differentiated.
"""
import numpy as np
import pytest
import scipy.special

from conftest import rel_l2

GRID = (32, 32, 24)
SAMP = (32, 16, 12)
NCOIL = 4
CYCLES, AMP, BORDER = 2, 2.0, 3.0


def kb_closed(d, w, b, n):
    """One axis factor of convFunction in FP64 (no float casts) and its
    derivative in d by the closed form used on the device."""
    dk = 1.0 / (n - 1)
    W = w * dk
    q = 2.0 * d / W
    beta = b * np.sqrt(1.0 - q * q)
    denom = 100.0 * w * dk
    f = scipy.special.i0(beta) / denom
    i1x = np.where(beta > 1e-12, scipy.special.i1(beta) / np.maximum(beta, 1e-300), 0.5)
    df = -(2.0 * b * b * q / W) * i1x / denom
    return f, df


def test_kb_derivative_closed_form_matches_fd():
    # d/dd of I0(b sqrt(1 - (2d/W)^2)) / denom against FP64 central differences,
    # including the window edge where beta -> 0
    for n in (96, 128, 320):
        d = np.linspace(0.0, 0.999, 400) * 2.125 / (n - 1)
        h = 1e-9
        f, df = kb_closed(d, 4.25, 14.9086, n)
        fp, _ = kb_closed(d + h, 4.25, 14.9086, n)
        fm, _ = kb_closed(np.abs(d - h), 4.25, 14.9086, n)
        fd = (fp - fm) / (2 * h)
        inner = d > 2 * h
        np.testing.assert_allclose(df[inner], fd[inner], rtol=2e-5, atol=1e-6 * np.abs(df).max())


def test_kb_closed_form_matches_oracle_factor():
    # the FP64 closed form equals the oracle's (reference-rounded) factor to float accuracy
    from oracle import oracle as orc

    n = 128
    for d in np.linspace(0.0, 2.0, 17) / (n - 1):
        f, _ = kb_closed(d, 4.25, 14.9086, n)
        assert abs(orc.kb_axis(np.float32(d), n, 4.25, 14.9086) - f) <= 2e-6 * f + 1e-12


def test_smooth_kernel_mode_agrees_with_reference_weights():
    # the oracle's smooth-kernel mode (finite-difference checks) keeps the
    # reference's tap sets and matches its float weights to float accuracy
    from oracle import oracle as orc

    pkg, lines, smask, k32, maps, rho = _setup(orc)
    a = _point_oracle(orc, maps, k32)
    b = _point_oracle(orc, maps, k32, smooth=True)
    ca, ia, wa = a.get_weights()
    cb, ib, wb = b.get_weights()
    np.testing.assert_array_equal(ca, cb)
    np.testing.assert_array_equal(ia, ib)
    np.testing.assert_allclose(wb, wa, rtol=2e-6, atol=1e-7 * wa.max())


def test_wave_coords_match_synth_and_differentiate():
    import torch
    from importlib import import_module

    pkg = import_module("4d_lung_amd")
    from importlib import import_module as im

    D = im("4d_lung_amd.differentiable")
    lines = pkg.synth.caipi_mask(SAMP[1], SAMP[2], (2, 2), 1)
    smask = pkg.synth.sample_mask_from_lines(lines, SAMP[0])
    for amp, py, pz in ((2.0, 0.0, 0.0), (1.5, 0.3, -0.7)):
        ref = pkg.synth.wave_caipi_trajectory(GRID, SAMP, CYCLES, amp, border=BORDER, phase_y=py, phase_z=pz)[smask]
        k = D.wave_caipi_coords(GRID, SAMP, lines, CYCLES, amp, py, pz, border=BORDER, device="cpu")
        np.testing.assert_allclose(k.numpy(), ref.astype(np.float32), rtol=0, atol=1e-7)
    # d k / d (amp, phase_y, phase_z) against central differences of the numpy generator
    wts = np.random.default_rng(0).standard_normal(ref.shape)
    a = torch.tensor(1.5, dtype=torch.float64, requires_grad=True)
    py = torch.tensor(0.3, dtype=torch.float64, requires_grad=True)
    pz = torch.tensor(-0.7, dtype=torch.float64, requires_grad=True)
    k = D.wave_caipi_coords(GRID, SAMP, lines, CYCLES, a, py, pz, border=BORDER, device="cpu")
    (k.double() * torch.from_numpy(wts)).sum().backward()
    h = 1e-6

    def f(amp, py_, pz_):
        t = pkg.synth.wave_caipi_trajectory(GRID, SAMP, CYCLES, amp, border=BORDER, phase_y=py_, phase_z=pz_)
        return float((t[smask] * wts).sum())

    fd = [(f(1.5 + h, 0.3, -0.7) - f(1.5 - h, 0.3, -0.7)) / (2 * h),
          (f(1.5, 0.3 + h, -0.7) - f(1.5, 0.3 - h, -0.7)) / (2 * h),
          (f(1.5, 0.3, -0.7 + h) - f(1.5, 0.3, -0.7 - h)) / (2 * h)]
    got = [a.grad.item(), py.grad.item(), pz.grad.item()]
    # the generator rounds to float32 at the end; its gradient is the FP64 one
    np.testing.assert_allclose(got, fd, rtol=1e-5, atol=1e-6)


# ---------------------------------------------------------------- GPU part
def _setup(orc, seed=5, R=(2, 2)):
    from importlib import import_module

    pkg = import_module("4d_lung_amd")
    lines = pkg.synth.caipi_mask(SAMP[1], SAMP[2], R, 1)
    smask = pkg.synth.sample_mask_from_lines(lines, SAMP[0])
    k64 = pkg.synth.wave_caipi_trajectory(GRID, SAMP, CYCLES, AMP, border=BORDER)[smask]
    k32 = k64.astype(np.float32)
    maps = pkg.synth.coil_maps(GRID, NCOIL, seed, coarse=4)
    rho = pkg.synth.phantom(GRID, (16, 16, 12), seed + 1, n_ellipsoids=5).reshape(-1)
    return pkg, lines, smask, k32, maps, rho


def _point_oracle(orc, maps, k, dcf=None, smooth=False):
    """The oracle on the points k (samples = points, all active).  smooth:
    FP64 kernel weights without the reference's float roundings, so that
    central differences in k measure the kernel's derivative."""
    o = orc.OracleReco(GRID, (len(k), 1, 1), NCOIL)
    o.set_trajectory(np.asarray(k, np.float64))
    if smooth:
        o.set_exact_weights(True)
    o.calculate_weights()
    o.calculate_kernel_ft()
    o.set_maps(maps)
    if dcf is not None:
        o.set_density(dcf)
    return o


def _E_orc(o, maps, x):
    return o.nufft_fwd(maps.astype(np.complex128) * np.asarray(x, np.complex128)[None, :]).T   # (m, C)


def _EH_orc(o, maps, y):
    img = o.nufft_bwd(np.ascontiguousarray(np.asarray(y, np.complex128).T))
    return (np.conj(maps.astype(np.complex128)) * img).sum(0)


def _rand_c(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.fixture(scope="module")
def dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


@pytest.mark.gpu
def test_point_set_operators_match_oracle_and_are_adjoint(orc, dev):
    import torch

    pkg, lines, smask, k32, maps, rho = _setup(orc)
    from importlib import import_module

    D = import_module("4d_lung_amd.differentiable")
    enc = D.Encoding(GRID, maps)
    rng = np.random.default_rng(1)
    x = _rand_c(rng, enc.ng)
    y = _rand_c(rng, (len(k32), NCOIL))
    kt = torch.from_numpy(k32).to(dev)
    enc.set_coords(kt)
    Ex = enc.forward_raw(torch.from_numpy(x).to(dev)).cpu().numpy()
    EHy = enc.adjoint_raw(torch.from_numpy(y).to(dev)).cpu().numpy().reshape(-1)
    o = _point_oracle(orc, maps, k32)
    assert rel_l2(Ex, _E_orc(o, maps, x)) < 1e-5
    assert rel_l2(EHy, _EH_orc(o, maps, y)) < 1e-5
    lhs = np.vdot(Ex.astype(np.complex128), y) * enc.scale
    rhs = np.vdot(x.astype(np.complex128), EHy)
    assert abs(lhs - rhs) <= 1e-5 * abs(rhs)
    enc.close()


@pytest.mark.gpu
def test_image_gradient_matches_EH_reference(orc, dev):
    import torch
    from importlib import import_module

    D = import_module("4d_lung_amd.differentiable")
    pkg, lines, smask, k32, maps, rho = _setup(orc)
    enc = D.Encoding(GRID, maps)
    rng = np.random.default_rng(2)
    y0 = _rand_c(rng, (len(k32), NCOIL)) * 0.01
    x = torch.from_numpy(rho.astype(np.complex64)).to(dev).requires_grad_(True)
    kt = torch.from_numpy(k32).to(dev)
    loss = (enc.E(x, kt) - torch.from_numpy(y0).to(dev)).abs().pow(2).sum()
    loss.backward()
    o = _point_oracle(orc, maps, k32)
    res = _E_orc(o, maps, rho) - y0
    ref = 2.0 * _EH_orc(o, maps, res) / enc.scale          # 2 E^*(E x - y)
    assert rel_l2(x.grad.cpu().numpy(), ref) < 1e-5
    enc.close()


def _fd_dir(f, k64, d, h=1e-4):
    return (f(k64 + h * d) - f(k64 - h * d)) / (2 * h)


@pytest.mark.gpu
def test_coordinate_gradients_match_fp64_finite_differences(orc, dev):
    import torch
    from importlib import import_module

    D = import_module("4d_lung_amd.differentiable")
    pkg, lines, smask, k32, maps, rho = _setup(orc)
    enc = D.Encoding(GRID, maps)
    rng = np.random.default_rng(3)
    m = len(k32)
    x = rho.astype(np.complex64)
    y0 = _rand_c(rng, (m, NCOIL)) * 0.01
    y1 = _rand_c(rng, (m, NCOIL))
    x0 = _rand_c(rng, enc.ng) * 1e-6
    d = rng.standard_normal((m, 3))
    d /= np.linalg.norm(d)
    k64 = k32.astype(np.float64)

    # through E: L = |E_k x - y0|^2
    kt = torch.from_numpy(k32).to(dev).requires_grad_(True)
    loss = (enc.E(torch.from_numpy(x).to(dev), kt) - torch.from_numpy(y0).to(dev)).abs().pow(2).sum()
    loss.backward()
    g = kt.grad.double().cpu().numpy()

    def fE(k):
        o = _point_oracle(orc, maps, k, smooth=True)
        return float(np.sum(np.abs(_E_orc(o, maps, x) - y0) ** 2))

    fd = _fd_dir(fE, k64, d)
    assert abs(np.sum(g * d) - fd) <= 1e-3 * abs(fd), (np.sum(g * d), fd)

    # through E^H: L = |E^H_k y1 - x0|^2
    kt = torch.from_numpy(k32).to(dev).requires_grad_(True)
    loss = (enc.EH(torch.from_numpy(y1).to(dev), kt).reshape(-1) - torch.from_numpy(x0).to(dev)).abs().pow(2).sum()
    loss.backward()
    g = kt.grad.double().cpu().numpy()

    def fEH(k):
        o = _point_oracle(orc, maps, k, smooth=True)
        return float(np.sum(np.abs(_EH_orc(o, maps, y1) - x0) ** 2))

    fd = _fd_dir(fEH, k64, d)
    assert abs(np.sum(g * d) - fd) <= 1e-3 * abs(fd), (np.sum(g * d), fd)

    # per-sample gradient: a handful of single-sample, single-axis directions
    kt = torch.from_numpy(k32).to(dev).requires_grad_(True)
    loss = (enc.E(torch.from_numpy(x).to(dev), kt) - torch.from_numpy(y0).to(dev)).abs().pow(2).sum()
    loss.backward()
    g = kt.grad.double().cpu().numpy()
    top = np.argsort(-np.abs(g).ravel())[:4]
    for t in top:
        e = np.zeros_like(k64)
        e.ravel()[t] = 1.0
        fd = _fd_dir(fE, k64, e)
        assert abs(g.ravel()[t] - fd) <= 2e-3 * abs(fd), (t, g.ravel()[t], fd)
    enc.close()


@pytest.mark.gpu
def test_wave_parameter_gradients_match_fp64_finite_differences(orc, dev):
    import torch
    from importlib import import_module

    D = import_module("4d_lung_amd.differentiable")
    pkg, lines, smask, k32, maps, rho = _setup(orc)
    enc = D.Encoding(GRID, maps)
    rng = np.random.default_rng(4)
    x = rho.astype(np.complex64)
    y0 = _rand_c(rng, (len(k32), NCOIL)) * 0.01
    a = torch.tensor(AMP, dtype=torch.float64, device=dev, requires_grad=True)
    py = torch.tensor(0.2, dtype=torch.float64, device=dev, requires_grad=True)
    pz = torch.tensor(-0.4, dtype=torch.float64, device=dev, requires_grad=True)
    k = D.wave_caipi_coords(GRID, SAMP, lines, CYCLES, a, py, pz, border=BORDER, device=dev)
    loss = (enc.E(torch.from_numpy(x).to(dev), k) - torch.from_numpy(y0).to(dev)).abs().pow(2).sum()
    loss.backward()

    def f(amp, p_y, p_z):
        kk = pkg.synth.wave_caipi_trajectory(GRID, SAMP, CYCLES, amp, border=BORDER, phase_y=p_y, phase_z=p_z)[smask]
        o = _point_oracle(orc, maps, kk, smooth=True)
        return float(np.sum(np.abs(_E_orc(o, maps, x) - y0) ** 2))

    h = 1e-4
    fd = [(f(AMP + h, 0.2, -0.4) - f(AMP - h, 0.2, -0.4)) / (2 * h),
          (f(AMP, 0.2 + h, -0.4) - f(AMP, 0.2 - h, -0.4)) / (2 * h),
          (f(AMP, 0.2, -0.4 + h) - f(AMP, 0.2, -0.4 - h)) / (2 * h)]
    got = [a.grad.item(), py.grad.item(), pz.grad.item()]
    for gv, fv in zip(got, fd):
        assert abs(gv - fv) <= 2e-3 * max(abs(fv), 1e-3 * max(abs(v) for v in fd)), (got, fd)
    enc.close()


@pytest.mark.gpu
def test_unrolled_cg_matches_oracle_and_its_gradients_fd(orc, dev):
    import torch
    from importlib import import_module

    D = import_module("4d_lung_amd.differentiable")
    pkg, lines, smask, k32, maps, rho = _setup(orc)
    enc = D.Encoding(GRID, maps)
    rng = np.random.default_rng(6)
    m = len(k32)
    k64 = k32.astype(np.float64)
    dcf = rng.uniform(0.5, 1.5, m).astype(np.float32)
    o = _point_oracle(orc, maps, k32, dcf)
    y = (_E_orc(o, maps, rho) + 1e-3 * _rand_c(rng, (m, NCOIL))).astype(np.complex64)
    n_iter = 5

    # forward: the unrolled loop equals the oracle's CG-SENSE (max_iter + 1 passes)
    kt = torch.from_numpy(k32).to(dev).requires_grad_(True)
    yt = torch.from_numpy(y).to(dev).requires_grad_(True)
    x = D.cg_sense_unrolled(enc, yt, kt, torch.from_numpy(dcf).to(dev), n_iter=n_iter)
    xr, _, its = o.cg_sense(np.ascontiguousarray(y.T), 0.0, n_iter - 1, prec=64)
    assert its == n_iter
    assert rel_l2(x.detach().cpu().numpy(), xr.reshape(-1)) < 1e-4
    target = torch.from_numpy(rho.astype(np.complex64)).to(dev)
    loss = (x - target).abs().pow(2).sum()
    loss.backward()

    def f(k, yy):
        oo = _point_oracle(orc, maps, k, dcf, smooth=True)
        xx, _, _ = oo.cg_sense(np.ascontiguousarray(yy.T), 0.0, n_iter - 1, prec=64)
        return float(np.sum(np.abs(xx.reshape(-1) - rho) ** 2))

    d = rng.standard_normal((m, 3))
    d /= np.linalg.norm(d)
    fd = _fd_dir(lambda k: f(k, y), k64, d)
    g = float(np.sum(kt.grad.double().cpu().numpy() * d))
    assert abs(g - fd) <= 5e-3 * abs(fd), (g, fd)
    # gradient in the measured data: the oracle takes complex64 data, so the
    # perturbation is relative (1e-3 of every sample) to stay far above the
    # rounding of the large k-space-centre samples
    u = (rng.uniform(-1, 1, (m, NCOIL)) + 1j * rng.uniform(-1, 1, (m, NCOIL)))
    v = y.astype(np.complex128) * u
    h = 1e-3
    fdy = (f(k64, y + h * v) - f(k64, y - h * v)) / (2 * h)
    gy = yt.grad.cpu().numpy().astype(np.complex128)
    gdy = float(np.sum(gy.real * v.real + gy.imag * v.imag))
    assert abs(gdy - fdy) <= 5e-3 * abs(fdy), (gdy, fdy)

    # gradient in the input image through the synthesized data y = E rho + noise
    noise = (y - _E_orc(o, maps, rho)).astype(np.complex128)
    rt = torch.from_numpy(rho.astype(np.complex64)).to(dev).requires_grad_(True)
    kt = torch.from_numpy(k32).to(dev)
    yy = enc.E(rt, kt) + torch.from_numpy(noise.astype(np.complex64)).to(dev)
    x = D.cg_sense_unrolled(enc, yy, kt, torch.from_numpy(dcf).to(dev), n_iter=n_iter)
    (x - target).abs().pow(2).sum().backward()
    vr = rho.astype(np.complex128) * rng.uniform(-1, 1, rho.shape)

    def fr(r):
        return f(k64, _E_orc(o, maps, r) + noise)

    fdr = (fr(rho + h * vr) - fr(rho - h * vr)) / (2 * h)
    gr = rt.grad.cpu().numpy().astype(np.complex128)
    gdr = float(np.sum(gr.real * vr.real + gr.imag * vr.imag))
    assert abs(gdr - fdr) <= 5e-3 * abs(fdr), (gdr, fdr)
    enc.close()


@pytest.mark.gpu
def test_unrolled_cg_checkpointing_gives_the_same_gradient(orc, dev):
    import torch
    from importlib import import_module

    D = import_module("4d_lung_amd.differentiable")
    pkg, lines, smask, k32, maps, rho = _setup(orc)
    enc = D.Encoding(GRID, maps)
    rng = np.random.default_rng(8)
    m = len(k32)
    o = _point_oracle(orc, maps, k32)
    y = (_E_orc(o, maps, rho) + 1e-3 * _rand_c(rng, (m, NCOIL))).astype(np.complex64)
    target = torch.from_numpy(rho.astype(np.complex64)).to(dev)
    grads = []
    for ck in (True, False):
        kt = torch.from_numpy(k32).to(dev).requires_grad_(True)
        x = D.cg_sense_unrolled(enc, torch.from_numpy(y).to(dev), kt, None, n_iter=3, checkpoint=ck)
        (x - target).abs().pow(2).sum().backward()
        grads.append(kt.grad.cpu().numpy())
    np.testing.assert_array_equal(grads[0], grads[1])
    enc.close()
