"""MI355X-native CG-SENSE engine for respiratory-resolved 3D wave-CAIPI MRI.

The hot path of expRad/4d_lung's CGSense program -- the encoding operator E,
its adjoint E^H and the CG-SENSE loop -- as hand-written HIP kernels for
gfx950 behind the C-ABI of include/cgs.h.  ``cgs`` is the Python binding of
that ABI, ``synth`` the seeded synthetic workload generator and
``differentiable`` (imported on demand, it needs torch) the autograd binding
of E / E^H with coordinate gradients and the unrolled CG-SENSE.
"""
from . import cgs, synth  # noqa: F401
from .cgs import CGS_ERR_NONFINITE, CgsError, Context, Fourier, SOLVE_BATCHED, SOLVE_PLAIN  # noqa: F401
