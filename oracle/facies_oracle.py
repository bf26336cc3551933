"""CPU oracle of the facies-probability step (test infrastructure only).

Restates crv_facies_density / crv_facies_prob (include/crava_hip.h) in
float64 numpy: histogram of (ln vp, ln vs, ln rho) samples on a box of
nx x ny x nz bins (bin centres lo + (i + 1/2) w), circular smoothing by the
Gaussian exp(-2 pi^2 f^T S f) in the Fourier domain, clamping at zero and
normalisation to unit integral; per voxel P(f | m) = pi_f p_f(m) /
sum_g pi_g p_g(m) with trilinear interpolation (edge bins beyond the box).
This is the density/classification of FaciesProb (reference
src/faciesprob.cpp:544-930); the reference keeps no outputs to compare
with, so the oracle is pinned by closed forms in tests/test_oracle.py: a
Gaussian sample cloud smoothed by S gives the Gaussian of covariance C + S
(plus the bin width's uniform variance), and two Gaussian facies with equal
covariance give logistic probabilities linear in m.
"""
from __future__ import annotations

import numpy as np


def density(samples, lo, hi, smooth_cov, bins):
    """(nz, ny, nx) density: x along ln vp, y along ln vs, z along ln rho."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    nb = np.asarray(bins)
    w = (hi - lo) / nb
    s = np.asarray(samples, np.float64)
    idx = np.floor((s - lo) / w).astype(np.int64)
    ok = np.all((idx >= 0) & (idx < nb), axis=1)
    h = np.zeros((bins[2], bins[1], bins[0]))
    np.add.at(h, (idx[ok, 2], idx[ok, 1], idx[ok, 0]), 1.0)
    H = np.fft.rfftn(h)
    L = hi - lo
    fz = np.fft.fftfreq(bins[2], d=1.0 / bins[2])[:, None, None] / L[2]
    fy = np.fft.fftfreq(bins[1], d=1.0 / bins[1])[None, :, None] / L[1]
    fx = np.arange(bins[0] // 2 + 1)[None, None, :] / L[0]
    S = np.asarray(smooth_cov, np.float64).reshape(3, 3)
    q = (S[0, 0] * fx * fx + S[1, 1] * fy * fy + S[2, 2] * fz * fz
         + 2.0 * (S[0, 1] * fx * fy + S[0, 2] * fx * fz + S[1, 2] * fy * fz))
    d = np.fft.irfftn(H * np.exp(-2.0 * np.pi ** 2 * q), s=h.shape)
    d = np.maximum(d, 0.0)
    cell = np.prod(w)
    return d / (d.sum() * cell)


def trilinear(d, lo, hi, bins, m):
    """Density at points m (n, 3)."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    nb = np.asarray(bins)
    t = (np.asarray(m, np.float64) - lo) / ((hi - lo) / nb) - 0.5
    t = np.clip(t, 0.0, nb - 1)
    i0 = np.minimum(np.floor(t).astype(np.int64), np.maximum(nb - 2, 0))
    f = t - i0
    out = np.zeros(t.shape[0])
    for c in range(8):
        dx, dy, dz = c & 1, (c >> 1) & 1, c >> 2
        x = np.minimum(i0[:, 0] + dx, nb[0] - 1)
        y = np.minimum(i0[:, 1] + dy, nb[1] - 1)
        z = np.minimum(i0[:, 2] + dz, nb[2] - 1)
        wt = ((f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1])
              * (f[:, 2] if dz else 1 - f[:, 2]))
        out += wt * d[z, y, x]
    return out


def probabilities(dens, prior, lo, hi, bins, m):
    pr = np.asarray(prior, np.float64)
    pr = pr / pr.sum()
    q = np.stack([p * trilinear(d, lo, hi, bins, m) for p, d in zip(pr, dens)], 0)
    s = q.sum(0)
    return np.where(s > 0, q / np.where(s > 0, s, 1.0), 1.0 / len(dens))


def probabilities_undef(dens, prior, lo, hi, bins, m, p_undefined):
    """With FaciesProb's undefined class (reference src/faciesprob.cpp:
    2008-2056): the denominator gains undef = p_undefined / (number of bins)
    for densities summing to 1 over the box -- in general p_undefined times
    the mean bin value of the first density; returns (P per facies, P of the
    undefined class)."""
    pr = np.asarray(prior, np.float64)
    pr = pr / pr.sum()
    d0 = np.asarray(dens[0], np.float64)
    undef = p_undefined * d0.sum() / d0.size
    q = np.stack([p * trilinear(d, lo, hi, bins, m) for p, d in zip(pr, dens)], 0)
    s = q.sum(0) + undef
    return q / s, undef / s


def density_pdf3d(samples, lo, hi, sigma, bins):
    """PosteriorElasticPDF3D's density (reference
    src/posteriorelasticpdf3d.cpp:12-131, 369-420), (nz, ny, nx): the
    histogram of the samples (bin floor((d - lo) / w)) times 1/n, circularly
    convolved with exp(-x^T sigma^-1 x / 2) sampled at the cyclic bin lags
    (j for j <= n/2, j - n above) as float and divided by its float running
    sum in the reference's loop order (z, y outer, x fastest).  Samples
    outside the box are dropped (the reference writes past its grid)."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    nb = np.asarray(bins)
    w = (hi - lo) / nb
    s = np.asarray(samples, np.float64)
    idx = np.floor((s - lo) / w).astype(np.int64)
    ok = np.all((idx >= 0) & (idx < nb), axis=1)
    h = np.zeros((bins[2], bins[1], bins[0]))
    np.add.at(h, (idx[ok, 2], idx[ok, 1], idx[ok, 0]), 1.0)
    h = h.astype(np.float32) * np.float32(1.0 / len(s))
    si = np.linalg.inv(np.asarray(sigma, np.float64).reshape(3, 3))
    lag = lambda n: np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n)
    z, y, x = np.meshgrid(lag(bins[2]) * w[2], lag(bins[1]) * w[1], lag(bins[0]) * w[0], indexing="ij")
    q = (x * x * si[0, 0] + y * y * si[1, 1] + z * z * si[2, 2]
         + 2 * x * y * si[1, 0] + 2 * x * z * si[2, 0] + 2 * y * z * si[2, 1])
    k = np.exp(-0.5 * q).astype(np.float32)
    ksum = np.cumsum(k.ravel(), dtype=np.float32)[-1]
    k = k / ksum
    return np.fft.irfftn(np.fft.rfftn(h.astype(np.float64)) * np.fft.rfftn(k.astype(np.float64)), s=h.shape)
