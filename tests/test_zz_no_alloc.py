"""cgs_fft_device_noalloc: the centred 3D FFT on a device array with its
twiddle tables in the library's fixed device array, against the allocating
entry points cgs_fft_forward / cgs_fft_backward (Fourier::transformFWD /
transformBWD, Fourier.cpp:19-42) on the same input, for two calls in a row,
the second on a smaller grid, in both directions."""
import ctypes

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

FFT_TOL = 2e-6       # FP32 transforms with tables rounded from FP64, relative L2


def _grid(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.mark.parametrize("direction,rescale", [(-1, False), (1, True), (1, False)])
def test_noalloc_fft_matches_allocating_entry_point(pkg, direction, rescale):
    import torch

    for k, shape in enumerate([(64, 48, 56), (16, 24, 20)]):
        print("no-alloc FFT: call %d dims %s dir %d rescale %d" % (k, shape, direction, rescale), flush=True)
        a = _grid(shape, 10 * k + direction + 3)
        F = pkg.Fourier(shape)
        ref = F.forward(a) if direction < 0 else F.backward(a, rescale=rescale)
        d = torch.from_numpy(a).to("cuda:0")
        stream = torch.cuda.current_stream().cuda_stream
        pkg.cgs.fft_device_noalloc(shape, d.data_ptr(), direction, rescale, stream)
        got = d.cpu().numpy()
        err = rel_l2(got, ref)
        print("no-alloc FFT: rel-L2 %.2e vs the allocating entry point" % err, flush=True)
        assert err < FFT_TOL


def test_noalloc_fft_refuses_unsupported_sizes(pkg):
    import torch

    d = torch.zeros(8, dtype=torch.complex64, device="cuda:0")
    L = pkg.cgs.lib()
    for dims in [(640, 8, 8), (8, 8, 1024), (8, 9, 8)]:
        st = L.cgs_fft_device_noalloc((ctypes.c_int * 3)(*dims), ctypes.c_void_p(d.data_ptr()), -1, 0, None)
        assert st == pkg.cgs.CGS_ERR_ARG, dims
    st = L.cgs_fft_device_noalloc((ctypes.c_int * 3)(8, 8, 8), ctypes.c_void_p(d.data_ptr()), 2, 0, None)
    assert st == pkg.cgs.CGS_ERR_ARG
