"""Host-side checks that need no GPU: the C-ABI library loads and exports
every function include/cgs.h declares, the argument checks of the binding,
and the synthetic workload generator."""
import ctypes
import importlib

import numpy as np
import pytest

synth = importlib.import_module("4d_lung_amd.synth")


def test_library_exports_every_declared_symbol(pkg):
    L = pkg.cgs.lib()
    names = pkg.cgs.exported_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing


def test_create_rejects_bad_sizes(pkg):
    L = pkg.cgs.lib()
    h = ctypes.c_void_p()
    grid = (ctypes.c_int * 3)(128, 128, 97)      # 97 = prime: no radix plan
    samp = (ctypes.c_int * 3)(4, 4, 4)
    assert L.cgs_create(grid, samp, 8, ctypes.byref(h)) == pkg.cgs.CGS_ERR_ARG
    grid = (ctypes.c_int * 3)(128, 128, 96)
    assert L.cgs_create(grid, samp, 0, ctypes.byref(h)) == pkg.cgs.CGS_ERR_ARG
    assert L.cgs_create(grid, samp, 257, ctypes.byref(h)) == pkg.cgs.CGS_ERR_ARG   # above CGS_MAX_COILS


def test_fft_rejects_unsupported_length(pkg):
    with pytest.raises(pkg.CgsError):
        pkg.Fourier((13, 16, 16))


def test_null_context_is_an_error_not_a_crash(pkg):
    L = pkg.cgs.lib()
    assert L.cgs_calculate_weights(None) == pkg.cgs.CGS_ERR_ARG
    assert L.cgs_set_maps(None, None) == pkg.cgs.CGS_ERR_ARG
    assert L.cgs_last_error(None) == b"null context"


@pytest.mark.parametrize("name", list(synth.CONFIGS))
def test_workload_table_is_consistent(name):
    cfg = synth.CONFIGS[name]
    assert all(g == 2 * i for g, i in zip(cfg["grid"], cfg["image"]))
    assert cfg["samp"][0] == cfg["grid"][0]
    assert cfg["samp"][1:] == cfg["image"][1:]


def test_wave_caipi_trajectory_stays_on_grid():
    grid, samp = (128, 128, 96), (128, 64, 48)
    k = synth.wave_caipi_trajectory(grid, samp, 3, 4.0)
    assert k.shape == (np.prod(samp), 3)
    pos = k * (np.array(grid) - 1)
    assert pos.min() >= 4.0 - 1e-9 and np.all(pos.max(0) <= np.array(grid) - 1 - 4.0 + 1e-9)
    # readout is the slowest index: kx constant over the phase-encode block of one readout point
    kk = k.reshape(samp + (3,))
    assert np.all(kk[3, :, :, 0] == kk[3, 0, 0, 0])
    # wave: ky - (line position) is a sinusoid of amplitude 4 cells along the readout
    wy = (kk[:, 5, 0, 1] - kk[0, 5, 0, 1]) * (grid[1] - 1)
    ph = 2 * np.pi * 3 * np.arange(samp[0]) / samp[0]
    np.testing.assert_allclose(wy, 4.0 * np.sin(ph), atol=1e-9)


@pytest.mark.parametrize("R,sy,sz", [((2, 2), 64, 48), ((3, 2), 160, 112), ((3, 3), 256, 160)])
def test_caipi_mask_fraction(R, sy, sz):
    m = synth.caipi_mask(sy, sz, R, 1)
    assert abs(m.mean() - 1.0 / (R[0] * R[1])) < 2.0 / sy
    # every acquired kz line holds ky lines shifted by one per kz step
    kz = np.flatnonzero(m.any(0))
    assert np.all(kz % R[1] == 0)


def test_respiratory_bins_fractions():
    m = synth.caipi_mask(128, 96, (2, 2), 1)
    b = synth.respiratory_bins(m, 8, (0.06, 0.20), 1234)
    assert np.all(b[~m] == -1) and np.all(b[m] >= 0)
    fr = np.bincount(b[m], minlength=8) / m.sum()
    assert fr.min() >= 0.06 and fr.max() <= 0.20


def test_coil_maps_are_normalised():
    maps = synth.coil_maps((16, 12, 8), 4, 5, coarse=4)
    ss = (np.abs(maps) ** 2).sum(0)
    np.testing.assert_allclose(ss, 1.0, rtol=1e-5)


def test_add_noise_keeps_zero_fill():
    y = np.ones((2, 10), np.complex64)
    m = np.zeros(10, bool)
    m[::2] = True
    out = synth.add_noise(y, m, 1e-3, 0)
    assert np.all(out[:, ~m] == 0) and np.all(out[0, m] != 0)
