"""Helpers shared by the GPU tests: the reference pipeline on the GPU
(crv_* through the Python view) and the same pipeline in the FP64 / FP32
oracle, on identical seeded inputs."""
from __future__ import annotations

import math

import numpy as np

from crava_amd import model
from oracle import oracle as O


class Case:
    def __init__(self, s: model.Survey):
        self.s = s
        self.syn = model.make_synthetic(s)
        self.pin = model.posterior_inputs(s, self.syn.data_var)
        self.cov_real = model.prior_cov_grids(s, np.float64)
        self.err_real = model.err_corr_grid(s, np.float64)


def gpu_posterior(ctx, case: Case, prec=32, mode=0, separable=False, skip_mean_inverse=False, **kw):
    from crava_amd.hip import FFTGrid, computePostMeanResidAndFFTCov
    s = case.s
    mk = lambda ct: FFTGrid(ctx, s.nxp, s.nyp, s.nzp, precision=prec, cubetype=ct)
    mean = [mk(FFTGrid.PARAMETER) for _ in range(3)]
    data = [mk(FFTGrid.DATA) for _ in range(s.n)]
    cov = [mk(FFTGrid.COVARIANCE) for _ in range(6)]
    err = None if separable else mk(FFTGrid.COVARIANCE)
    for g, a in zip(mean, case.syn.mean):
        g.upload_real(a)
    for g, a in zip(data, case.syn.data):
        g.upload_real(a)
    if not separable:
        for g, a in zip(cov, case.cov_real):
            g.upload_real(a)
        err.upload_real(case.err_real)
    p = case.pin
    st = computePostMeanResidAndFFTCov(ctx, mean, data, err, cov, A=p.A, theta=p.theta, tables=p.tables,
                                       nz=s.nz, lz=s.lz, low_cut=s.low_cut, high_cut=s.high_cut, wnc=s.wnc,
                                       sigma0=p.sigma0, mode=mode, separable=separable, cxy=p.cxy, ct=p.ct,
                                       skip_mean_inverse=skip_mean_inverse, **kw)
    return dict(mean=mean, data=data, cov=cov, stats=st)


def oracle_posterior(case: Case, dtype=np.float64, nthreads=4, separable=False):
    """Oracle pipeline: FP64 transforms (numpy) + the restated loop in dtype.
    Returns real posterior means, residual spectra and covariance spectra.
    separable: the prior covariance and error spectra are the exact products
    Sigma0[a][b] cxy ct and cxy (what a separable plan uses) instead of the
    transforms of the float real-space grids, whose rounding leaves a noise
    floor of about 1e-8 of the peak in the spectra."""
    s = case.s
    mean = [O.fft_r2c(m, s.nxp, 2) for m in case.syn.mean]
    data = [O.fft_r2c(d, s.nxp, 1) for d in case.syn.data]
    if separable:
        cxy, ct = model.prior_spectra(s)
        base = ct[:, None, None] * cxy[None, :, :]
        S0 = case.pin.sigma0
        cov = [S0[a, b] * base + 0j for a, b in ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))]
        err = np.broadcast_to(cxy[None, :, :], base.shape) + 0j
    else:
        err = O.fft_r2c(case.err_real, s.nxp, 3)
        cov = [O.fft_r2c(c, s.nxp, 3) for c in case.cov_real]
    t = case.pin.tables
    pm, pd, pc, nf = O.posterior_loop(mean, data, err, cov, kwd=t.kwd, em1=t.em1, em2=t.em2, em3=t.em3,
                                      A=case.pin.A, theta=case.pin.theta, sigma0=case.pin.sigma0,
                                      invert=case.pin.invert, wnc=s.wnc, dtype=dtype, nthreads=nthreads)
    mreal = [O.fft_c2r(m, s.nxp, 2) for m in pm]
    return dict(mean_spec=pm, mean=mreal, data=pd, cov=pc, nfail=nf)
