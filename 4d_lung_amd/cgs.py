"""Python binding of libcgs_hip.so (the C-ABI declared in include/cgs.h).

Mirrors the reference's Reco / Fourier interface for the CG-SENSE hot path
(Reco.h:13-214, Fourier.h:12-54) with numpy arrays in the reference's layouts:

* images and coil maps: complex64, ``(C, Nx, Ny, Nz)`` or flat ``(C, Ng)``;
* k-space data: complex64 ``(C, nS)`` zero-filled, sample index
  ``ix*(sY*sZ) + iy*sZ + iz`` (readout slowest);
* trajectory: float64 ``(nS, 3)`` in [0, 1];
* density weights: float32 ``(nS,)``.

Errors come back as :class:`CgsError` carrying the library's status and
message; nothing here falls back to a CPU implementation: if the shared
library is missing the import fails.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcgs_hip.so")

CGS_OK = 0
CGS_ERR_ARG = 1
CGS_ERR_HIP = 2
CGS_ERR_STATE = 3
CGS_ERR_NONFINITE = 4
CGS_ERR_MEMORY = 5
SOLVE_PLAIN = 0
SOLVE_BATCHED = 1

_lib = None


class CgsError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("cgs status %d: %s" % (status, msg))
        self.status = status


def lib():
    """Load libcgs_hip.so (built by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libcgs_hip.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
    # torch bundles its own HIP runtime (SONAME libamdhip64.so.7, loaded by
    # the name libamdhip64.so).  Loaded after the library, torch would bring
    # a second runtime into the process, and the second one to open the GPU
    # finds none.  With torch loaded first the library binds to torch's copy
    # by SONAME and the process holds one runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    I = ctypes.c_int
    I3 = ctypes.POINTER(ctypes.c_int)
    D = ctypes.c_double
    LL = ctypes.c_longlong
    sig = {
        "cgs_last_error": (ctypes.c_char_p, [P]),
        "cgs_create": (I, [I3, I3, I, ctypes.POINTER(P)]),
        "cgs_destroy": (I, [P]),
        "cgs_set_kernel": (I, [P, ctypes.c_float, ctypes.c_float, I]),
        "cgs_set_img_range": (I, [P, I3, I3]),
        "cgs_set_tile_config": (I, [P, I]),
        "cgs_num_tile_configs": (I, []),
        "cgs_get_tile_config": (I, [P, ctypes.POINTER(I), I3]),
        "cgs_set_trajectory": (I, [P, P]),
        "cgs_calculate_weights": (I, [P]),
        "cgs_exclude_missing_data": (I, [P, P]),
        "cgs_calculate_kernel_ft": (I, [P]),
        "cgs_get_kernel_ft": (I, [P, P]),
        "cgs_set_density": (I, [P, P]),
        "cgs_set_maps": (I, [P, P]),
        "cgs_nufft_fwd": (I, [P, P, P]),
        "cgs_nufft_bwd": (I, [P, P, P]),
        "cgs_apply_normal": (I, [P, P, P]),
        "cgs_cg_sense_images": (I, [P, P, D, I, P, P, P]),
        "cgs_set_num_bins": (I, [P, I]),
        "cgs_set_bin_group": (I, [P, I]),
        "cgs_set_pruning": (I, [P, I]),
        "cgs_set_grid_lists": (I, [P, I]),
        "cgs_set_zpass_bins": (I, [P, I]),
        "cgs_set_fused_gridding": (I, [P, I]),
        "cgs_fused_gridding_active": (I, [P, P]),
        "cgs_get_pruning": (I, [P, ctypes.POINTER(ctypes.c_int)]),
        "cgs_get_prune_fractions": (I, [P, I, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
        "cgs_get_bin_group": (I, [P, ctypes.POINTER(ctypes.c_int)]),
        "cgs_set_bin_data": (I, [P, I, P, P]),
        "cgs_set_bin_samples": (I, [P, I, LL, P, P, P]),
        "cgs_cg_prepare": (I, [P, D, I]),
        "cgs_cg_iterate": (I, [P, I, I]),
        "cgs_cg_results": (I, [P, P, P, P]),
        "cgs_cg_history_length": (I, [P, ctypes.POINTER(ctypes.c_int)]),
        "cgs_cg_sense_bins": (I, [P, I, D, I, P, P, P]),
        "cgs_bin_count": (I, [P, I, ctypes.POINTER(LL)]),
        "cgs_bin_tables": (I, [P, I, P, P, P, P, P]),
        "cgs_kernel_info": (I, [P, ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(D)]),
        "cgs_voronoi_volumes": (I, [P, LL, P]),
        "cgs_voronoi_density": (I, [P, LL, I, I, D, P]),
        "cgs_voronoi_volumes_rule": (I, [P, LL, I, P]),
        "cgs_voronoi_density_rule": (I, [P, LL, I, I, I, D, P]),
        "cgs_set_bin_voronoi": (I, [P, I]),
        "cgs_fft_create": (I, [I3, ctypes.POINTER(P)]),
        "cgs_fft_destroy": (I, [P]),
        "cgs_fft_forward": (I, [P, P]),
        "cgs_fft_backward": (I, [P, P, I]),
        "cgs_fft_device_noalloc": (I, [I3, P, I, I, P]),
        "cgs_profile_enable": (I, [P, I]),
        "cgs_profile_report": (I, [P, ctypes.c_char_p, I]),
        "cgs_profile_reset": (I, [P]),
        "cgs_device_bytes": (LL, [P]),
        "cgs_set_stream": (I, [P, P]),
        "cgs_pts_create": (I, [P, ctypes.POINTER(P)]),
        "cgs_pts_destroy": (I, [P]),
        "cgs_pts_set_coords": (I, [P, LL, P]),
        "cgs_pts_count": (I, [P, ctypes.POINTER(LL)]),
        "cgs_pts_forward": (I, [P, P, P]),
        "cgs_pts_adjoint": (I, [P, P, P]),
        "cgs_pts_grad_coords_fwd": (I, [P, P, P, P]),
        "cgs_pts_grad_coords_adj": (I, [P, P, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols():
    """Names of every function declared in include/cgs.h."""
    import re

    hdr = os.path.join(os.path.dirname(_HERE), "include", "cgs.h")
    text = open(hdr).read()
    return sorted(set(re.findall(r"\b(cgs_[a-z0-9_]+)\s*\(", text)))


def _i3(v):
    return (ctypes.c_int * 3)(*[int(x) for x in v])


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _c64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.complex64)
    if n is not None and a.size != n:
        raise ValueError("expected %d complex values, got %d" % (n, a.size))
    return a


def _f32(a, n, what):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.size != n:
        raise ValueError("%s: expected %d values, got %d" % (what, n, a.size))
    return a


class Context:
    """Device state of one reconstruction (the reference's Reco object)."""

    def __init__(self, grid, samp, ncoil):
        self._L = lib()
        self.grid = tuple(int(v) for v in grid)
        self.samp = tuple(int(v) for v in samp)
        self.ncoil = int(ncoil)
        self.ng = int(np.prod(self.grid))
        self.ns = int(np.prod(self.samp))
        self.nbins = 1
        h = ctypes.c_void_p()
        st = self._L.cgs_create(_i3(self.grid), _i3(self.samp), self.ncoil, ctypes.byref(h))
        if st != CGS_OK:
            raise CgsError(st, "cgs_create failed for grid %s samples %s coils %d" % (self.grid, self.samp, ncoil))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self._L.cgs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, st):
        if st != CGS_OK:
            raise CgsError(st, self._L.cgs_last_error(self.h).decode())

    # -- configuration (Reco / ConvKernel setters)
    def set_kernel(self, w=4.25, b=14.9086, order=5):
        self._ck(self._L.cgs_set_kernel(self.h, w, b, order))

    def set_img_range(self, rng, offc=(0, 0, 0)):
        self._ck(self._L.cgs_set_img_range(self.h, _i3(rng), _i3(offc)))

    def set_tile_config(self, tcfg):
        self._ck(self._L.cgs_set_tile_config(self.h, int(tcfg)))

    def tile_config(self):
        t = ctypes.c_int()
        d = (ctypes.c_int * 3)()
        self._ck(self._L.cgs_get_tile_config(self.h, ctypes.byref(t), d))
        return t.value, tuple(d)

    def set_trajectory(self, traj):
        t = np.ascontiguousarray(traj, dtype=np.float64)
        if t.shape != (self.ns, 3):
            raise ValueError("trajectory must be (%d, 3)" % self.ns)
        self._ck(self._L.cgs_set_trajectory(self.h, _ptr(t)))

    def calculate_weights(self):
        self._ck(self._L.cgs_calculate_weights(self.h))

    def exclude_missing_data(self, data):
        """Only coil 0 is read: data may be (C, nS) or just (1, nS)."""
        d = _c64(data)
        if d.size not in (self.ns, self.ncoil * self.ns):
            raise ValueError("data must hold nS or C*nS complex values")
        self._ck(self._L.cgs_exclude_missing_data(self.h, _ptr(d)))

    def calculate_kernel_ft(self):
        self._ck(self._L.cgs_calculate_kernel_ft(self.h))

    def kernel_ft(self):
        out = np.zeros(self.grid, np.float32)
        self._ck(self._L.cgs_get_kernel_ft(self.h, _ptr(out)))
        return out

    def set_density(self, dcf):
        if dcf is None:
            self._ck(self._L.cgs_set_density(self.h, None))
        else:
            d = _f32(dcf, self.ns, "density weights")
            self._ck(self._L.cgs_set_density(self.h, _ptr(d)))

    def set_maps(self, maps):
        m = _c64(maps, self.ncoil * self.ng)
        self._ck(self._L.cgs_set_maps(self.h, _ptr(m)))

    # -- operators
    def nufft_fwd(self, img):
        i = _c64(img, self.ncoil * self.ng)
        out = np.zeros((self.ncoil, self.ns), np.complex64)
        self._ck(self._L.cgs_nufft_fwd(self.h, _ptr(i), _ptr(out)))
        return out

    def nufft_bwd(self, data):
        d = _c64(data, self.ncoil * self.ns)
        out = np.zeros((self.ncoil, self.ng), np.complex64)
        self._ck(self._L.cgs_nufft_bwd(self.h, _ptr(d), _ptr(out)))
        return out

    def apply_normal(self, p):
        pp = _c64(p, self.ng)
        q = np.zeros(self.ng, np.complex64)
        self._ck(self._L.cgs_apply_normal(self.h, _ptr(pp), _ptr(q)))
        return q

    def cg_sense_images(self, img, thresh=0.0, max_iter=15):
        i = _c64(img, self.ncoil * self.ng)
        x = np.zeros(self.ng, np.complex64)
        res = np.zeros(max_iter + 1, np.float64)
        it = ctypes.c_int()
        self._ck(self._L.cgs_cg_sense_images(self.h, _ptr(i), float(thresh), int(max_iter), _ptr(x), _ptr(res),
                                             ctypes.byref(it)))
        return x.reshape(self.grid), res[: it.value], it.value

    # -- respiratory bins
    def set_num_bins(self, n):
        self._ck(self._L.cgs_set_num_bins(self.h, int(n)))
        self.nbins = int(n)

    def set_bin_group(self, max_bins):
        """Cap on the bins one batched launch covers (0: device memory decides)."""
        self._ck(self._L.cgs_set_bin_group(self.h, int(max_bins)))

    def bin_group(self):
        """Bins per batched launch chosen by the last cg_prepare."""
        g = ctypes.c_int()
        self._ck(self._L.cgs_get_bin_group(self.h, ctypes.byref(g)))
        return g.value

    def set_pruning(self, mode):
        """Column pruning of the bins' x passes: 0 off, 1 automatic, 2 on."""
        self._ck(self._L.cgs_set_pruning(self.h, int(mode)))

    def set_grid_lists(self, on):
        """Per-tile gridding candidate lists built when the bins are packed
        (True, the default) or found by a scan in every gridding launch."""
        self._ck(self._L.cgs_set_grid_lists(self.h, int(bool(on))))

    def set_fused_gridding(self, on):
        """Gridding fused into the inverse x pass for the packed bins (True,
        the default, where supported) or a separate gridding pass."""
        self._ck(self._L.cgs_set_fused_gridding(self.h, int(bool(on))))

    def fused_gridding(self):
        """Whether the packed bins use the fused gridding."""
        v = ctypes.c_int(0)
        self._ck(self._L.cgs_fused_gridding_active(self.h, ctypes.byref(v)))
        return bool(v.value)

    def set_zpass_bins(self, bins):
        """Bins per workgroup of the batched inverse z pass (1: one
        workgroup per bin and cell line; default 64)."""
        self._ck(self._L.cgs_set_zpass_bins(self.h, int(bins)))

    def pruning(self):
        """Whether the packed bins use column pruning."""
        a = ctypes.c_int()
        self._ck(self._L.cgs_get_pruning(self.h, ctypes.byref(a)))
        return bool(a.value)

    def prune_fractions(self, b=0):
        """(x-pass column groups transformed, gridding tile columns written)
        as fractions, for bin b of the packed bins."""
        g, t = ctypes.c_double(), ctypes.c_double()
        self._ck(self._L.cgs_get_prune_fractions(self.h, int(b), ctypes.byref(g), ctypes.byref(t)))
        return g.value, t.value

    def set_bin_data(self, b, data, dcf=None):
        d = _c64(data, self.ncoil * self.ns)
        w = None if dcf is None else _f32(dcf, self.ns, "density weights")
        self._ck(self._L.cgs_set_bin_data(self.h, int(b), _ptr(d), _ptr(w)))

    def set_bin_samples(self, b, idx, y, dcf=None):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        y = _c64(y, self.ncoil * idx.size)
        w = None if dcf is None else _f32(dcf, idx.size, "density weights")
        self._ck(self._L.cgs_set_bin_samples(self.h, int(b), idx.size, _ptr(idx), _ptr(y), _ptr(w)))

    def cg_prepare(self, thresh=0.0, max_iter=15):
        self._ck(self._L.cgs_cg_prepare(self.h, float(thresh), int(max_iter)))

    def cg_iterate(self, mode, n_iter):
        self._ck(self._L.cgs_cg_iterate(self.h, int(mode), int(n_iter)))

    def cg_history_length(self):
        """Residuals per bin that cg_results returns (max_iter + 1 of the
        last cg_prepare or solve)."""
        n = ctypes.c_int()
        self._ck(self._L.cgs_cg_history_length(self.h, ctypes.byref(n)))
        return n.value

    def cg_results(self, want_x=True):
        x = np.zeros((self.nbins,) + self.grid, np.complex64) if want_x else None
        res = np.zeros((self.nbins, self.cg_history_length()), np.float64)
        its = np.zeros(self.nbins, np.int32)
        self._ck(self._L.cgs_cg_results(self.h, _ptr(x), _ptr(res), _ptr(its)))
        return x, res, its

    def cg_sense_bins(self, mode=SOLVE_BATCHED, thresh=0.0, max_iter=15):
        self.cg_prepare(thresh, max_iter)
        self.cg_iterate(mode, max_iter + 1)
        return self.cg_results()

    def set_bin_voronoi(self, b):
        """Voronoi weights of the bin's active samples become its DCF."""
        self._ck(self._L.cgs_set_bin_voronoi(self.h, int(b)))

    # -- introspection
    def bin_count(self, b=0):
        m = ctypes.c_longlong()
        self._ck(self._L.cgs_bin_count(self.h, int(b), ctypes.byref(m)))
        return m.value

    def kernel_info(self):
        kw, nt, thr = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self._ck(self._L.cgs_kernel_info(self.h, ctypes.byref(kw), ctypes.byref(nt), ctypes.byref(thr)))
        return kw.value, nt.value, thr.value

    def bin_tables(self, b=0, nearest=True):
        m = self.bin_count(b)
        kw, nt, _ = self.kernel_info()
        ne = np.zeros((self.ns, 3), np.int32) if nearest else None
        perm = np.zeros(m, np.int32)
        toff = np.zeros(nt + 1, np.int32)
        base = np.zeros((m, 3), np.int32)
        vals = np.zeros((m, 3 * kw), np.float64)
        self._ck(self._L.cgs_bin_tables(self.h, int(b), _ptr(ne), _ptr(perm), _ptr(toff), _ptr(base), _ptr(vals)))
        return dict(nearest=ne, perm=perm, tile_off=toff, base=base, vals=vals)

    # -- timing
    def profile(self, on=True):
        self._ck(self._L.cgs_profile_enable(self.h, int(bool(on))))

    def profile_reset(self):
        self._ck(self._L.cgs_profile_reset(self.h))

    def profile_report(self):
        n = self._L.cgs_profile_report(self.h, None, 0)
        buf = ctypes.create_string_buffer(max(n, 1))
        self._L.cgs_profile_report(self.h, buf, len(buf))
        out = {}
        for line in buf.value.decode().splitlines():
            name, cnt, ms = line.split()
            out[name] = (int(cnt), float(ms))
        return out

    def device_bytes(self):
        return int(self._L.cgs_device_bytes(self.h))


VORONOI_RULES = {"hull": 0, "clipped": 1}


def voronoi_volumes(points, rule="hull"):
    """Voronoi cell volumes on the device.  rule "hull" (the reference's,
    functions.h:148-150): 0 only for points on the convex hull; "clipped":
    0 also for cells reaching more than one hash cell beyond the points'
    bounding box (the per-bin weights of respiratory bins)."""
    p = np.ascontiguousarray(points, dtype=np.float64)
    vol = np.zeros(len(p), np.float64)
    st = lib().cgs_voronoi_volumes_rule(_ptr(p), len(p), VORONOI_RULES[rule], _ptr(vol))
    if st:
        raise CgsError(st, "cgs_voronoi_volumes_rule failed")
    return vol


def voronoi_density(points, average=False, sx=0, cutoff=1.0, rule="hull"):
    """estimateDensityV3D (functions.h:161-258) on the device."""
    p = np.ascontiguousarray(points, dtype=np.float64)
    d = np.zeros(len(p), np.float32)
    st = lib().cgs_voronoi_density_rule(_ptr(p), len(p), VORONOI_RULES[rule], int(bool(average)), int(sx),
                                        float(cutoff), _ptr(d))
    if st:
        raise CgsError(st, "cgs_voronoi_density_rule failed")
    return d


class Fourier:
    """Centred 3D FFT of one complex64 grid (Fourier.h:12-54)."""

    def __init__(self, dims):
        self._L = lib()
        self.dims = tuple(int(v) for v in dims)
        h = ctypes.c_void_p()
        st = self._L.cgs_fft_create(_i3(self.dims), ctypes.byref(h))
        if st != CGS_OK:
            raise CgsError(st, "unsupported FFT size %s" % (self.dims,))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            self._L.cgs_fft_destroy(self.h)
            self.h = None

    def forward(self, a):
        a = np.array(a, dtype=np.complex64, copy=True, order="C")
        st = self._L.cgs_fft_forward(self.h, _ptr(a))
        if st:
            raise CgsError(st, "transformFWD failed")
        return a

    def backward(self, a, rescale=True):
        a = np.array(a, dtype=np.complex64, copy=True, order="C")
        st = self._L.cgs_fft_backward(self.h, _ptr(a), int(bool(rescale)))
        if st:
            raise CgsError(st, "transformBWD failed")
        return a


def fft_device_noalloc(dims, a_ptr, direction, rescale=True, stream=None):
    """Centred 3D FFT in place on a device array at address ``a_ptr``
    (complex64, ``dims`` = (Nx, Ny, Nz)) on ``stream`` (a hipStream_t
    handle as an int, or None), without any device allocation: direction -1
    is Fourier::transformFWD, +1 transformBWD (with ``rescale``).  Returns
    without synchronising."""
    st = lib().cgs_fft_device_noalloc(_i3(dims), ctypes.c_void_p(int(a_ptr)), int(direction), int(bool(rescale)),
                                      None if stream is None else ctypes.c_void_p(int(stream)))
    if st:
        raise CgsError(st, "cgs_fft_device_noalloc failed for dims %s" % (tuple(dims),))
