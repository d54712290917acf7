"""ctypes binding of the CPU oracle (oracle/cgs_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.  The oracle restates
the reference's CGSense hot path (Reco.cpp, Reco.h, Fourier.cpp) on the CPU;
every function of the C++ file cites the reference line it follows.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_cgs.so")
_lib = None


def build():
    """Compile the oracle with its Makefile (g++, no GPU needed)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        I3 = ctypes.POINTER(ctypes.c_int)
        L.orc_i0.restype = ctypes.c_double
        L.orc_i0.argtypes = [ctypes.c_double]
        L.orc_kb_axis.restype = ctypes.c_double
        L.orc_kb_axis.argtypes = [ctypes.c_float, ctypes.c_int, ctypes.c_float, ctypes.c_float]
        L.orc_conv_function.restype = ctypes.c_float
        L.orc_conv_function.argtypes = [ctypes.POINTER(ctypes.c_double), I3, ctypes.c_float, ctypes.c_float]
        L.orc_fft3.restype = ctypes.c_int
        L.orc_fft3.argtypes = [P, I3, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_fft1.restype = ctypes.c_int
        L.orc_fft1.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.orc_create.restype = P
        L.orc_create.argtypes = [I3, I3, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int]
        L.orc_destroy.argtypes = [P]
        L.orc_set_img_range.argtypes = [P, I3, I3]
        L.orc_set_trajectory.argtypes = [P, P]
        L.orc_calculate_weights.argtypes = [P, P]
        L.orc_set_exact_weights.argtypes = [P, ctypes.c_int]
        L.orc_exclude_missing_data.argtypes = [P, P]
        L.orc_num_taps.restype = ctypes.c_long
        L.orc_num_taps.argtypes = [P]
        L.orc_get_weights.argtypes = [P, P, P, P]
        L.orc_kernel_threshold.restype = ctypes.c_double
        L.orc_kernel_threshold.argtypes = [P]
        L.orc_calculate_kernel_ft.argtypes = [P]
        L.orc_get_kernel_ft.argtypes = [P, P]
        L.orc_set_density.argtypes = [P, P]
        L.orc_set_maps.argtypes = [P, P]
        L.orc_density_correct.argtypes = [P, P]
        L.orc_nufft_fwd.restype = ctypes.c_int
        L.orc_nufft_fwd.argtypes = [P, ctypes.c_int, P, P]
        L.orc_nufft_bwd.restype = ctypes.c_int
        L.orc_nufft_bwd.argtypes = [P, ctypes.c_int, P, P]
        L.orc_apply_normal.restype = ctypes.c_int
        L.orc_apply_normal.argtypes = [P, ctypes.c_int, P, P]
        L.orc_cg_sense.restype = ctypes.c_int
        L.orc_cg_sense.argtypes = [P, ctypes.c_int, P, ctypes.c_double, ctypes.c_int, P, P, P]
        L.orc_cg_sense_timed.restype = ctypes.c_int
        L.orc_cg_sense_timed.argtypes = [P, ctypes.c_int, P, ctypes.c_double, ctypes.c_int, P, P, P,
                                         ctypes.POINTER(ctypes.c_double), P]
        L.orc_set_atomic_coil_sum.argtypes = [P, ctypes.c_int]
        L.orc_set_verbose.argtypes = [ctypes.c_int]
        L.orc_i0_std.restype = ctypes.c_double
        L.orc_i0_std.argtypes = [ctypes.c_double]
        L.orc_voronoi_volumes.restype = ctypes.c_long
        L.orc_voronoi_volumes.argtypes = [P, ctypes.c_long, P, ctypes.c_int]
        L.orc_voronoi_density.restype = ctypes.c_long
        L.orc_voronoi_density.argtypes = [P, ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_double, P,
                                          ctypes.c_int]
        L.orc_voronoi_volumes_rule.restype = ctypes.c_long
        L.orc_voronoi_volumes_rule.argtypes = [P, ctypes.c_long, ctypes.c_int, P, ctypes.c_int]
        L.orc_voronoi_density_rule.restype = ctypes.c_long
        L.orc_voronoi_density_rule.argtypes = [P, ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_double, P, ctypes.c_int]
        L.orc_density_from_volumes.restype = None
        L.orc_density_from_volumes.argtypes = [P, ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_double, P]
        _lib = L
    return _lib


def _i3(v):
    return (ctypes.c_int * 3)(*[int(x) for x in v])


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def set_verbose(v):
    lib().orc_set_verbose(int(v))


def i0(x):
    return lib().orc_i0(float(x))


def i0_std(x):
    return lib().orc_i0_std(float(x))


VORONOI_RULES = {"hull": 0, "clipped": 1}


def voronoi_volumes(points, nthreads=0, rule="hull"):
    """Cell volumes by half-space clipping (the device algorithm on the CPU).
    rule "hull": 0 only for unbounded cells (points on the convex hull), as
    the reference (functions.h:148-150); "clipped": 0 also for cells that
    reach more than one hash cell beyond the points' bounding box."""
    p = np.ascontiguousarray(points, dtype=np.float64)
    vol = np.zeros(len(p), np.float64)
    over = lib().orc_voronoi_volumes_rule(_ptr(p), len(p), VORONOI_RULES[rule], _ptr(vol), int(nthreads))
    if over:
        raise RuntimeError("%d Voronoi cells exceeded the face/vertex limits" % over)
    return vol


def voronoi_density(points, average=False, sx=0, cutoff=1.0, nthreads=0, rule="hull"):
    """estimateDensityV3D (functions.h:161-258) with the clipped volumes of
    the given rule."""
    p = np.ascontiguousarray(points, dtype=np.float64)
    d = np.zeros(len(p), np.float32)
    over = lib().orc_voronoi_density_rule(_ptr(p), len(p), VORONOI_RULES[rule], int(bool(average)), int(sx),
                                          float(cutoff), _ptr(d), int(nthreads))
    if over:
        raise RuntimeError("%d Voronoi cells exceeded the face/vertex limits" % over)
    return d


def density_from_volumes(vol, average=False, sx=0, cutoff=1.0):
    """The averaging and normalisation of estimateDensityV3D
    (functions.h:212-246) applied to given cell volumes."""
    v = np.ascontiguousarray(vol, dtype=np.float64)
    d = np.zeros(len(v), np.float32)
    lib().orc_density_from_volumes(_ptr(v), len(v), int(bool(average)), int(sx), float(cutoff), _ptr(d))
    return d


def kb_axis(dxf, n, w, b):
    return lib().orc_kb_axis(float(dxf), int(n), float(w), float(b))


def conv_function(dist, grid, w, b):
    d = (ctypes.c_double * 3)(*[float(x) for x in dist])
    return lib().orc_conv_function(d, _i3(grid), float(w), float(b))


def fft3_centered(a, direction, rescale=True):
    """Fourier::transformFWD (direction=-1) / transformBWD (+1), in place
    semantics on a copy; complex128 -> FP64, complex64 -> FP32."""
    a = np.ascontiguousarray(a).copy()
    prec = 64 if a.dtype == np.complex128 else 32
    assert a.dtype in (np.complex128, np.complex64) and a.ndim == 3
    st = lib().orc_fft3(_ptr(a), _i3(a.shape), prec, int(direction), int(bool(rescale)))
    assert st == 0
    return a


def fft1(a, sign):
    a = np.ascontiguousarray(a, dtype=np.complex128).copy()
    lib().orc_fft1(_ptr(a), a.size, int(sign))
    return a


class OracleReco:
    """The oracle's Reco state: grid, sampling, coils, KB constants
    (Reco.cpp:374-400 defaults w=4.25, b=14.9086, order 5)."""

    def __init__(self, grid, samp, ncoil, w=4.25, b=14.9086, order=5, nthreads=0):
        self.grid = tuple(int(x) for x in grid)
        self.samp = tuple(int(x) for x in samp)
        self.ncoil = int(ncoil)
        self.ng = int(np.prod(self.grid))
        self.ns = int(np.prod(self.samp))
        self.h = lib().orc_create(_i3(grid), _i3(samp), ncoil, w, b, order, nthreads)
        self._keep = []

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set_img_range(self, rng, offc=(0, 0, 0)):
        lib().orc_set_img_range(self.h, _i3(rng), _i3(offc))

    def set_trajectory(self, traj):
        traj = np.ascontiguousarray(traj, dtype=np.float64)
        assert traj.shape == (self.ns, 3)
        lib().orc_set_trajectory(self.h, _ptr(traj))

    def set_exact_weights(self, on=True):
        """Smooth-kernel mode: prec-64 operators use FP64 weights without the
        reference's float roundings (finite-difference gradient checks)."""
        lib().orc_set_exact_weights(self.h, int(bool(on)))

    def calculate_weights(self, data=None):
        if data is None:
            lib().orc_calculate_weights(self.h, None)
        else:
            d = np.ascontiguousarray(data, dtype=np.complex64)
            lib().orc_calculate_weights(self.h, _ptr(d))

    def exclude_missing_data(self, data):
        d = np.ascontiguousarray(data, dtype=np.complex64)
        lib().orc_exclude_missing_data(self.h, _ptr(d))

    def get_weights(self):
        n = lib().orc_num_taps(self.h)
        counts = np.zeros(self.ns, np.int32)
        idx = np.zeros(n, np.int32)
        w = np.zeros(n, np.float32)
        lib().orc_get_weights(self.h, _ptr(counts), _ptr(idx), _ptr(w))
        return counts, idx, w

    def kernel_threshold(self):
        return lib().orc_kernel_threshold(self.h)

    def calculate_kernel_ft(self):
        lib().orc_calculate_kernel_ft(self.h)

    def get_kernel_ft(self):
        out = np.zeros(self.grid, np.float32)
        lib().orc_get_kernel_ft(self.h, _ptr(out))
        return out

    def set_density(self, d):
        if d is None:
            lib().orc_set_density(self.h, None)
        else:
            d = np.ascontiguousarray(d, dtype=np.float32)
            self._keep.append(d)
            lib().orc_set_density(self.h, _ptr(d))

    def set_maps(self, maps):
        m = np.ascontiguousarray(maps, dtype=np.complex64)
        assert m.size == self.ncoil * self.ng
        lib().orc_set_maps(self.h, _ptr(m))

    def nufft_fwd(self, img, prec=64):
        dt = np.complex128 if prec == 64 else np.complex64
        img = np.ascontiguousarray(img, dtype=dt).reshape(self.ncoil, self.ng)
        out = np.zeros((self.ncoil, self.ns), dt)
        assert lib().orc_nufft_fwd(self.h, prec, _ptr(img), _ptr(out)) == 0
        return out

    def nufft_bwd(self, data, prec=64):
        dt = np.complex128 if prec == 64 else np.complex64
        data = np.ascontiguousarray(data, dtype=dt).reshape(self.ncoil, self.ns)
        out = np.zeros((self.ncoil, self.ng), dt)
        assert lib().orc_nufft_bwd(self.h, prec, _ptr(data), _ptr(out)) == 0
        return out

    def apply_normal(self, p, prec=64):
        dt = np.complex128 if prec == 64 else np.complex64
        p = np.ascontiguousarray(p, dtype=dt).reshape(self.ng)
        q = np.zeros(self.ng, dt)
        st = lib().orc_apply_normal(self.h, prec, _ptr(p), _ptr(q))
        if st:
            raise RuntimeError("oracle apply_normal status %d" % st)
        return q

    def set_atomic_coil_sum(self, on=True):
        """Coil sum in the reference's loop shape: float omp atomic adds of
        every coil into q (Reco.cpp:120-129, 257-262), order left to the
        threads; the default sums the coils in a fixed order."""
        lib().orc_set_atomic_coil_sum(self.h, int(bool(on)))

    def cg_sense(self, data, thresh=0.0, max_iter=15, prec=64, timed=False, iter_times=False):
        """Reco::cgSenseDynamic after density correction and nufftBWD of data.
        With timed=True also returns the seconds spent in the iteration loop;
        with iter_times=True, after that, the seconds of every iteration."""
        dt = np.complex128 if prec == 64 else np.complex64
        d = np.ascontiguousarray(data, dtype=np.complex64).reshape(self.ncoil, self.ns)
        x = np.zeros(self.ng, dt)
        res = np.zeros(max_iter + 2, np.float64)
        per = np.zeros(max_iter + 2, np.float64)
        it = ctypes.c_int(0)
        secs = ctypes.c_double(0.0)
        st = lib().orc_cg_sense_timed(self.h, prec, _ptr(d), float(thresh), int(max_iter), _ptr(x), _ptr(res),
                                      ctypes.byref(it), ctypes.byref(secs), _ptr(per))
        if st:
            raise RuntimeError("oracle cg_sense status %d" % st)
        out = (x.reshape(self.grid), res[: it.value], it.value)
        if timed:
            out = out + (secs.value,)
        if iter_times:
            out = out + (per[: it.value],)
        return out
