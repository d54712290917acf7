"""Differentiable encoding operators on device tensors (torch.autograd).

E (Reco::nufftFWD, Reco.cpp:521-553) and E^H (Reco::nufftBWD,
Reco.cpp:468-519) of one context, as an adjoint pair of
``torch.autograd.Function`` s over the C-ABI point-set calls
(``cgs_pts_*`` in include/cgs.h).  Their backward passes return the gradient
with respect to the input (the other operator of the pair, rescaled) and with
respect to the per-sample k-space coordinates (hand-written kernel over the
analytic derivative of the Kaiser-Bessel kernel, Reco.h:218-242).  On top:

* :func:`wave_caipi_coords` -- the sinusoidal wave-CAIPI sample positions of
  WaveCaipi::calculateTraj (Trajectory.cpp:18-184) in closed form as torch
  operations, so a loss can be differentiated with respect to the wave
  amplitude and phases;
* :func:`cg_sense_unrolled` -- the CG-SENSE loop of Reco::cgSenseDynamic
  (Reco.cpp:175-321) for a fixed number of iterations, with every iteration
  checkpointed (recomputed in the backward pass instead of stored).

Scale convention (the reference's): ``<E x, y> * 1e-3/sqrt(Ng) = <x, E^H y>``.
Tensors: images ``(Ng,)`` or ``(Nx, Ny, Nz)`` complex64, samples ``(m, C)``
complex64 (coil fastest), coordinates ``(m, 3)`` float32 in [0, 1], all on
the GPU.  There is no CPU path: without the HIP library the import of
:mod:`cgs` fails.
"""
import ctypes
import math

import numpy as np
import torch

from .cgs import CGS_OK, CgsError, Context, lib


class Encoding:
    """E and E^H of one grid / coil set / map set on a changing point set.

    ``maps``: complex64 ``(C, Nx, Ny, Nz)`` (host array, fixed).  The kernel
    defaults to the reference's constants (Reco.cpp:378-380).
    """

    def __init__(self, grid, maps, kernel=(4.25, 14.9086, 5), tile=0):
        self._L = lib()
        self.grid = tuple(int(v) for v in grid)
        maps = np.ascontiguousarray(maps, np.complex64)
        self.ncoil = int(maps.shape[0])
        self.ng = int(np.prod(self.grid))
        self.ctx = Context(self.grid, (1, 1, 1), self.ncoil)
        self.ctx.set_kernel(*kernel)
        self.ctx.set_tile_config(tile)
        self.ctx.calculate_kernel_ft()
        self.ctx.set_maps(maps.reshape(self.ncoil, -1))
        p = ctypes.c_void_p()
        self._ck(self._L.cgs_pts_create(self.ctx.h, ctypes.byref(p)))
        self.p = p
        self.scale = 1e-3 / math.sqrt(self.ng)
        self._stream = None
        self._coords = None
        self.m = 0

    def close(self):
        if getattr(self, "p", None):
            self._L.cgs_pts_destroy(self.p)
            self.p = None
        if getattr(self, "ctx", None) is not None:
            self.ctx.close()
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, st):
        if st != CGS_OK:
            raise CgsError(st, self._L.cgs_last_error(self.ctx.h).decode())

    def _stream_sync(self):
        s = torch.cuda.current_stream().cuda_stream
        if s != self._stream:
            self._ck(self._L.cgs_set_stream(self.ctx.h, ctypes.c_void_p(s)))
            self._stream = s

    def set_coords(self, k):
        """Bin the ``(m, 3)`` float32 coordinates (kept until coordinates of
        other values are given)."""
        if k.dtype != torch.float32 or not k.is_cuda or k.dim() != 2 or k.shape[1] != 3:
            raise ValueError("coordinates must be a (m, 3) float32 CUDA tensor")
        kd = k.detach()
        # The binning (sort, KB tables, a stream synchronisation) is redone
        # only when the coordinates change.  A private copy of the last binned
        # coordinates is compared with the caller's tensor by value, so the
        # cache also holds for the detached copies that autograd and
        # checkpointing hand back, and an in-place change is never missed.
        self._stream_sync()
        if self._coords is not None and self._coords.shape == kd.shape and torch.equal(kd, self._coords):
            return
        kc = kd.contiguous()
        self._ck(self._L.cgs_pts_set_coords(self.p, kc.shape[0], ctypes.c_void_p(kc.data_ptr())))
        self.m = int(kc.shape[0])
        self._coords = kc.clone() if kc.data_ptr() == k.data_ptr() else kc

    @staticmethod
    def _cplx(t, n):
        if t.dtype != torch.complex64 or not t.is_cuda or t.numel() != n:
            raise ValueError("expected a complex64 CUDA tensor with %d elements" % n)
        return t.contiguous()

    # raw device calls (no autograd)
    def forward_raw(self, x):
        x = self._cplx(x, self.ng)
        self._stream_sync()
        y = torch.empty((self.m, self.ncoil), dtype=torch.complex64, device=x.device)
        self._ck(self._L.cgs_pts_forward(self.p, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr())))
        return y

    def adjoint_raw(self, y):
        y = self._cplx(y, self.m * self.ncoil)
        self._stream_sync()
        x = torch.empty(self.grid, dtype=torch.complex64, device=y.device)
        self._ck(self._L.cgs_pts_adjoint(self.p, ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(x.data_ptr())))
        return x

    def grad_coords_fwd(self, x, gy):
        x = self._cplx(x, self.ng)
        gy = self._cplx(gy, self.m * self.ncoil)
        self._stream_sync()
        gk = torch.empty((self.m, 3), dtype=torch.float32, device=x.device)
        self._ck(self._L.cgs_pts_grad_coords_fwd(self.p, ctypes.c_void_p(x.data_ptr()),
                                                  ctypes.c_void_p(gy.data_ptr()), ctypes.c_void_p(gk.data_ptr())))
        return gk

    def grad_coords_adj(self, y, gx):
        y = self._cplx(y, self.m * self.ncoil)
        gx = self._cplx(gx, self.ng)
        self._stream_sync()
        gk = torch.empty((self.m, 3), dtype=torch.float32, device=y.device)
        self._ck(self._L.cgs_pts_grad_coords_adj(self.p, ctypes.c_void_p(y.data_ptr()),
                                                  ctypes.c_void_p(gx.data_ptr()), ctypes.c_void_p(gk.data_ptr())))
        return gk

    # differentiable operators
    def E(self, x, k):
        """y = E_k x, differentiable in x and k."""
        return _EncodeFn.apply(x, k, self)

    def EH(self, y, k):
        """x = E_k^H y, differentiable in y and k."""
        return _AdjointFn.apply(y, k, self)


class _EncodeFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, k, enc):
        enc.set_coords(k)
        y = enc.forward_raw(x.detach())
        ctx.save_for_backward(x, k)
        ctx.enc = enc
        ctx.shape = x.shape
        return y

    @staticmethod
    def backward(ctx, gy):
        x, k = ctx.saved_tensors
        enc = ctx.enc
        enc.set_coords(k)
        gy = gy.contiguous()
        gx = gk = None
        if ctx.needs_input_grad[0]:
            # E^* = E^H / scale
            gx = (enc.adjoint_raw(gy) * (1.0 / enc.scale)).reshape(ctx.shape)
        if ctx.needs_input_grad[1]:
            gk = enc.grad_coords_fwd(x.detach(), gy)
        return gx, gk, None


class _AdjointFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y, k, enc):
        enc.set_coords(k)
        x = enc.adjoint_raw(y.detach())
        ctx.save_for_backward(y, k)
        ctx.enc = enc
        return x

    @staticmethod
    def backward(ctx, gx):
        y, k = ctx.saved_tensors
        enc = ctx.enc
        enc.set_coords(k)
        gx = gx.contiguous()
        gy = gk = None
        if ctx.needs_input_grad[0]:
            # (E^H)^H = scale * E
            gy = enc.forward_raw(gx) * enc.scale
        if ctx.needs_input_grad[1]:
            gk = enc.grad_coords_adj(y.detach(), gx)
        return gy, gk, None


def wave_caipi_coords(grid, samp, lines, cycles, amp, phase_y=0.0, phase_z=0.0, border=4.0, device="cuda"):
    """Acquired wave-CAIPI sample positions ``(m, 3)`` float32 in [0, 1].

    The closed-form sinusoidal waves of WaveCaipi::calculateTraj
    (Trajectory.cpp:18-184) in grid units: readout ``ix`` at
    ``border + ix (Nx-1-2 border)/(sX-1)``, phase-encode line ``(iy, iz)`` at
    its Cartesian position plus ``amp sin(2 pi cycles ix/sX + phase_y)`` (y) and
    ``amp cos(2 pi cycles ix/sX + phase_z)`` (z), border ``border + amp`` in y
    and z.  ``lines``: boolean ``(sY, sZ)`` mask of acquired lines; samples
    come in the reference's memory order (readout slowest).  ``amp``,
    ``phase_y``, ``phase_z`` may be 0-dim tensors that require gradients.
    """
    nx, ny, nz = grid
    sx, sy, sz = samp
    f64 = dict(dtype=torch.float64, device=device)
    amp = torch.as_tensor(amp, **f64)
    phase_y = torch.as_tensor(phase_y, **f64)
    phase_z = torch.as_tensor(phase_z, **f64)
    ix = torch.arange(sx, **f64)
    iy = torch.arange(sy, **f64)
    iz = torch.arange(sz, **f64)
    bx = float(border)
    by = border + amp
    bz = border + amp
    gx = bx + ix * (nx - 1 - 2 * bx) / (sx - 1)
    py = by + iy * (ny - 1 - 2 * by) / (sy - 1)
    pz = bz + iz * (nz - 1 - 2 * bz) / (sz - 1)
    ph = 2.0 * math.pi * cycles * ix / sx
    wy = amp * torch.sin(ph + phase_y)
    wz = amp * torch.cos(ph + phase_z)
    li = torch.as_tensor(np.flatnonzero(np.asarray(lines).reshape(-1)), device=device)
    ly, lz = li // sz, li % sz
    kx = (gx / (nx - 1))[:, None].expand(sx, li.numel())
    ky = (py[ly][None, :] + wy[:, None]) / (ny - 1)
    kz = (pz[lz][None, :] + wz[:, None]) / (nz - 1)
    return torch.stack([kx, ky, kz], dim=-1).reshape(-1, 3).to(torch.float32)


def _cdot(a, b):
    """sum conj(a) b: complex64 products summed in FP64 (conjScalarProduct,
    Reco.cpp:131-141)."""
    return (a.conj() * b).sum(dtype=torch.complex128)


def cg_sense_unrolled(enc, y, k, dcf=None, n_iter=5, checkpoint=True):
    """``n_iter`` iterations of Reco::cgSenseDynamic (Reco.cpp:175-321) on
    ``E^H W E x = E^H W y`` from x = 0, differentiable in ``y``, ``k`` (and
    ``dcf``).  Scalars alpha and beta are complex FP64 (complexDivision,
    Reco.cpp:143-173); vectors stay complex64.  With ``checkpoint`` every
    iteration keeps only its inputs and is recomputed in the backward pass.
    Returns the image ``(Ng,)``."""
    w = None if dcf is None else dcf.to(torch.float32)[:, None]

    def A(v):
        t = enc.E(v, k)
        if w is not None:
            t = t * w
        return enc.EH(t, k).reshape(-1)

    yw = y if w is None else y * w
    b = enc.EH(yw, k).reshape(-1)
    x = torch.zeros_like(b)
    r = b
    p = b
    rr = _cdot(r, r)

    def step(x, r, p, rr):
        q = A(p)
        alpha = rr / _cdot(p, q)
        x = x + alpha.to(torch.complex64) * p
        r = r - alpha.to(torch.complex64) * q
        rr_new = _cdot(r, r)
        beta = rr_new / rr
        p = r + beta.to(torch.complex64) * p
        return x, r, p, rr_new

    for _ in range(n_iter):
        if checkpoint:
            x, r, p, rr = torch.utils.checkpoint.checkpoint(step, x, r, p, rr, use_reentrant=False)
        else:
            x, r, p, rr = step(x, r, p, rr)
    return x
