"""Synthetic surveys generated on the GPU (stand-in for the SEG-Y cubes).

True model: the prior means plus a draw from the prior, simulated exactly as
Crava::simulate draws from a covariance spectrum (here the prior's), with the
reference noise stream from RandomGen(seed_truth).  Seismic: the linearised
Aki-Richards forward model conj(W_l) D1 A_l m (Crava::computeSyntSeismic,
reference src/crava.cpp:1767-1894) plus white noise at the survey's S/N,
where CRAVA's S/N is (|signal|^2 + |noise|^2) / |noise|^2.  The data
variance and errThetaCov then follow Crava::computeDataVariance and
setupErrorCorrelation (src/crava.cpp:407-487).
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib as L
from . import hip, model


class GpuSurvey:
    """Inputs of one inversion, resident on the GPU in FP64 (`mean`, `data`)."""

    def __init__(self, ctx: hip.Context, s: model.Survey, truth_noise: str = "reference"):
        self.ctx, self.s = ctx, s
        mk = lambda ct: hip.FFTGrid(ctx, s.nxp, s.nyp, s.nzp, precision=64, cubetype=ct, nx=s.nx, ny=s.ny, nz=s.nz)
        self.cxy, self.ct = model.prior_spectra(s)
        self.sigma0 = s.sigma0()
        self.A = model.reflection_matrix(s.angles_deg, s.vsvp)
        self.tables = model.kz_tables(s)
        self.invert = model.invert_mask(s)
        # prior draw -> true model
        cov = [mk(L.CRV_COVARIANCE) for _ in range(6)]
        hip.fill_prior_spectra(ctx, self.sigma0, self.cxy, self.ct, cov)
        self.mean = [mk(L.CRV_PARAMETER) for _ in range(3)]
        for g, mu in zip(self.mean, s.mu):
            hip.grid_fill(g, mu)
        truth = [mk(L.CRV_PARAMETER) for _ in range(3)]
        if truth_noise == "reference":
            hip.simulate(ctx, cov, self.mean, truth, noise_mode=L.CRV_NOISE_REFERENCE, lcg_state=s.seed_truth)
        else:
            hip.simulate(ctx, cov, self.mean, truth, noise_mode=L.CRV_NOISE_COUNTER, seed=s.seed_truth)
        for g in cov:
            g.close()
        self.truth = truth
        spec = [mk(L.CRV_PARAMETER) for _ in range(3)]
        for sp, t in zip(spec, truth):
            sp.fftFrom(t)
        # seismic = K m + white noise
        self.data = [mk(L.CRV_DATA) for _ in range(s.n)]
        kwd = self.tables.as_interleaved()[0]
        hip.forward_model(ctx, self.A, kwd, spec, self.data)
        for sp in spec:
            sp.close()
        self.data_var = []
        for l, g in enumerate(self.data):
            g.invFFTInPlace()
            _, s2, n = hip.moments(g)
            sig2 = s2 / max(n, 1)
            hip.add_white_noise(g, math.sqrt(sig2 / max(s.sn[l] - 1.0, 1e-6)), s.seed_noise, l)
            _, s2, n = hip.moments(g)
            self.data_var.append(float(np.float32(s2 / max(n, 1))))
        self.theta = model.err_theta_cov(s, self.data_var)

    def posterior_kwargs(self) -> dict:
        s = self.s
        return dict(A=self.A, theta=self.theta, tables=self.tables, nz=s.nz, lz=s.lz, low_cut=s.low_cut,
                    high_cut=s.high_cut, wnc=s.wnc, sigma0=self.sigma0, cxy=self.cxy, ct=self.ct)

    def close(self):
        for g in self.mean + self.data + self.truth:
            g.close()
