"""Facies probabilities over the posterior from facies-conditional samples
(the FaciesProb step fed by rplib, reference src/faciesprob.cpp:544-930 and
rplib/distributionsrockdem.cpp:125-156): densities of (ln vp, ln vs, ln rho)
per facies on a box, smoothed by the posterior uncertainty, then Bayes'
rule per voxel.  Samples come from the surrogate with its rplib-backed
error gate (crava_amd.surrogate) or from crv_rock_forward directly."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .hip import Context, FFTGrid

FACIES = ("shale", "brine_sand", "gas_sand")


def make_box(lo, hi) -> L.FaciesBox:
    b = L.FaciesBox()
    for d in range(3):
        b.lo[d], b.hi[d] = float(lo[d]), float(hi[d])
    return b


def density(ctx: Context, d_samples, box: L.FaciesBox, smooth_cov, bins=(100, 100, 50),
            kernel=L.CRV_SMOOTH_FOURIER) -> FFTGrid:
    """Smoothed density of device samples (n, 3) float on the box (a torch
    tensor: torch's work on it ends before the library's stream reads it, and
    the library is done with it on return).  kernel CRV_SMOOTH_FOURIER: the
    Gaussian's transform, normalised to unit integral; CRV_SMOOTH_SAMPLED:
    the reference's sampled kernel, a mass per bin summing to 1."""
    import torch
    g = FFTGrid(ctx, bins[0], bins[1], bins[2], precision=32, cubetype=L.CRV_OPERATOR)
    S = np.ascontiguousarray(smooth_cov, np.float64).reshape(3, 3)
    torch.cuda.synchronize()
    ctx.check(ctx.lib.crv_facies_density_kernel(ctx.h, C.c_void_p(d_samples.data_ptr()), int(d_samples.shape[0]),
                                                C.byref(box), S.ctypes.data_as(C.POINTER(C.c_double)), int(kernel),
                                                g.h), "facies_density")
    ctx.synchronize()
    return g


def probabilities(ctx: Context, densities, prior, box: L.FaciesBox, mean, p_undefined=0.0, with_undefined=False):
    """P(facies | posterior mean) per voxel, one real FP32 grid per facies.
    p_undefined > 0 adds FaciesProb's undefined class (0.01 in the
    reference's defaults); with_undefined also returns its probability grid."""
    m0 = mean[0]
    out = [FFTGrid(ctx, m0.nxp, m0.nyp, m0.nzp, precision=32, cubetype=L.CRV_PARAMETER) for _ in densities]
    und = FFTGrid(ctx, m0.nxp, m0.nyp, m0.nzp, precision=32, cubetype=L.CRV_PARAMETER) if with_undefined else None
    pr = np.ascontiguousarray(prior, np.float64)
    ds = (C.c_void_p * len(densities))(*[g.h.value for g in densities])
    ms = (C.c_void_p * 3)(*[g.h.value for g in mean])
    ps = (C.c_void_p * len(out))(*[g.h.value for g in out])
    ctx.check(ctx.lib.crv_facies_prob_undef(ctx.h, len(densities), ds, pr.ctypes.data_as(C.POINTER(C.c_double)),
                                            float(p_undefined), C.byref(box), ms, ps,
                                            und.h if und is not None else None), "facies_prob")
    return (out, und) if with_undefined else out
