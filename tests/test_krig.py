"""Prediction kriging to wells (Crava::doPostKriging, reference
src/crava.cpp:1736-1765; CKrigingAdmin, src/krigingadmin.cpp).

CPU: the oracle against what kriging fixes whatever its planner decides --
the kriged value at a data cell is the datum, a single datum gives
trend + gamma(d) / gamma(0) r (through the cross-covariance for the other
parameters), cells beyond half a padded grid keep the trend -- and the
ranges of a known exponential covariance.  GPU: crv_krig_post against the
oracle (the planner's integers equal, the kriged means to float rounding)
on a posterior from the inversion with wells drawn from the true model."""
import math

import numpy as np
import pytest

from oracle import krig_oracle as KO

M = float(KO.MISSING)
SCALES = (1.0, 0.5, 0.3, 0.4, 0.2, 0.1)   # aa bb rr ab ar br


def _exp_cov(nxp, nyp, nzp, rx, ry, rz, scale):
    """scale exp(-3 sqrt((dx/rx)^2 + (dy/ry)^2)) exp(-3 |dz| / rz) on the
    periodic padded grid, as a real (nzp, nyp, nxp) array."""
    dx = np.minimum(np.arange(nxp), nxp - np.arange(nxp))
    dy = np.minimum(np.arange(nyp), nyp - np.arange(nyp))
    dz = np.minimum(np.arange(nzp), nzp - np.arange(nzp))
    h = np.sqrt((dx[None, :] / rx) ** 2 + (dy[:, None] / ry) ** 2)
    return (scale * np.exp(-3 * h)[None] * np.exp(-3 * dz / rz)[:, None, None]).astype(np.float32)


def _covs(shape, r=(6.0, 5.0, 4.0)):
    return [_exp_cov(*shape, *r, s) for s in SCALES]


def test_ranges_of_exponential_covariance():
    c = KO.CovSep(_exp_cov(40, 36, 32, 7.0, 5.0, 9.0, 2.0), 40, 36, 32)
    rx, ry, rz = c.ranges()
    assert rx == math.floor(7.0 * math.log(20) / 3)
    assert ry == math.floor(5.0 * math.log(20) / 3)
    assert rz == math.ceil(9.0 * math.log(20) / 3) - 1


@pytest.mark.parametrize("target", [10000, 12])
def test_kriging_interpolates_the_data(target):
    """Every datum reproduced at its cell, in one block (target above the
    data count) and with the planner's blocks (target 12)."""
    nxp, nyp, nzp = 32, 28, 24
    dims = (30, 26, 20)
    rng = np.random.default_rng(3)
    trend = [np.full((nzp, nyp, nxp), m, np.float32) + rng.normal(0, 0.01, (nzp, nyp, nxp)).astype(np.float32)
             for m in (8.0, 7.3, 7.7)]
    pts = []
    for (i, j) in ((4, 5), (20, 9), (11, 21)):
        for k in range(0, dims[2], 2):
            v = [trend[a][k, j, i] + rng.normal(0, 0.05) for a in range(3)]
            if k % 6 == 2:
                v[1] = M
            pts.append([i, j, k, *v])
    out, st = KO.krig_post(trend, _covs((nxp, nyp, nzp)), pts, target, dims, (nxp, nyp, nzp))
    for p in pts:
        i, j, k = p[:3]
        for a in range(3):
            if p[3 + a] != M:
                assert abs(out[a][k, j, i] - np.float32(p[3 + a])) < 2e-6, (a, p)
    if target < len(pts):
        assert st["block"][0] < dims[0] or st["block"][1] < dims[1] or st["block"][2] < dims[2]
    assert st["data_points"] == len(pts)


def test_one_datum_closed_form():
    """One point with only log Vp: alpha kriged by gamma_aa(c - d) / gamma_aa(0)
    r, beta and rho by the cross-covariances gamma_ab, gamma_ar from the
    cell to the datum over gamma_aa(0); cells half a padded grid away or more
    keep the trend."""
    nxp, nyp, nzp = 32, 28, 24
    dims = (32, 28, 24)
    trend = [np.full((nzp, nyp, nxp), m, np.float32) for m in (8.0, 7.3, 7.7)]
    d = (9, 12, 6)
    r = 0.07
    pts = [[*d, 8.0 + r, M, M]]
    covs = [KO.CovSep(c, nxp, nyp, nzp) for c in _covs((nxp, nyp, nzp))]
    out, st = KO.krig_with(trend, covs, pts, 100, dims)
    g0 = covs[0].g2(*d, *d)
    rr = np.float32(np.float32(8.0 + r) - np.float32(8.0))
    kk, jj, ii = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    for a, q, flip in ((0, 0, False), (1, 3, True), (2, 4, True)):
        sg = -1 if flip else 1
        g = covs[q].g2_many(sg * (d[0] - ii), sg * (d[1] - jj), sg * (d[2] - kk)).astype(np.float64)
        want = (trend[a] + (g * (float(rr) / float(g0))).astype(np.float32)).astype(np.float32)
        assert np.allclose(out[a], want, rtol=0, atol=2e-6)
        far = (np.abs(ii - d[0]) >= nxp // 2) | (np.abs(jj - d[1]) >= nyp // 2) | (np.abs(kk - d[2]) >= nzp // 2)
        assert far.any() and np.array_equal(out[a][far], trend[a][far])
    assert st["max_block_data"] == [1, 1, 1]


def _wells(truth, dims, cols, step=2, beta_gap=7):
    pts = []
    for (i, j) in cols:
        for k in range(0, dims[2], step):
            v = [float(truth[a][k, j, i]) for a in range(3)]
            if beta_gap and k % beta_gap == 3:
                v[1] = M
            pts.append([i, j, k, *v])
    return pts


WELLS3 = [(6, 7), (27, 18), (15, 30)]
WELLS8 = [(18, 18), (30, 34), (1, 5), (32, 34), (9, 11), (34, 15), (10, 29), (10, 14)]


@pytest.mark.gpu
@pytest.mark.parametrize("cols,target,gap", [(WELLS3, 100000, 7), (WELLS8, 20, None)])
def test_krig_post_matches_oracle(ctx, cols, target, gap):
    """The posterior means of a 40x36x30 inversion kriged to wells of the
    true model, sampled at every other cell (at every cell the band-limited
    posterior covariance makes the system indefinite, where the reference's
    KrigBlock throws as well): three wells in one block with log Vs missing
    at some samples, and eight wells in the planner's blocks of about 20
    data:
    crv_krig_post's planner integers equal the oracle's, its kriged means
    agree to float rounding, and every datum is reproduced at its cell."""
    from conftest import rel_rms
    from crava_amd import hip, model
    from pipeline import Case, gpu_posterior
    case = Case(model.config_small(40, 36, 30, 3))
    s = case.s
    r = gpu_posterior(ctx, case, prec=32, mode=0, separable=True)
    for g in r["cov"]:
        g.invFFTInPlace()
    dims = (s.nxp, s.nyp, s.nzp)
    trend = [g.download_real()[:, :, :s.nxp].copy() for g in r["mean"]]
    covs = [g.download_real()[:, :, :s.nxp].copy() for g in r["cov"]]
    pts = _wells(case.syn.truth, dims, cols, beta_gap=gap)
    st = hip.krig_post(ctx, r["mean"], r["cov"], pts, target)
    want, ost = KO.krig_post(trend, covs, pts, target, dims, (s.nxp, s.nyp, s.nzp))
    print("gpu", st)
    print("oracle", ost)
    for key in ("block", "block_ext", "range"):
        assert tuple(st[key]) == tuple(ost[key]), key
    assert st["data_points"] == ost["data_points"] and st["blocks"] == ost["blocks"]
    assert (st["blocks"] > 1) == (gap is None)
    assert list(st["empty_blocks"]) == ost["empty_blocks"]
    assert list(st["max_block_data"]) == ost["max_block_data"]
    assert list(st["cells_updated"]) == ost["cells_updated"]
    for a in range(3):
        got = r["mean"][a].download_real()[:, :, :s.nxp]
        assert np.max(np.abs(got - want[a])) < 4e-6
        assert rel_rms(got - trend[a], want[a] - trend[a]) < 1e-5
        assert np.max(np.abs(got - trend[a])) > 1e-3
        for p in pts:
            if p[3 + a] != M:
                assert abs(got[p[2], p[1], p[0]] - np.float32(p[3 + a])) < 4e-6


@pytest.mark.gpu
def test_krig_post_rejects_smoothing(ctx):
    from crava_amd import hip
    g = [hip.FFTGrid(ctx, 16, 12, 10) for _ in range(9)]
    with pytest.raises(hip.CravaError, match="not built"):
        hip.krig_post(ctx, g[:3], g[3:], [[1, 1, 1, 8.0, 7.0, 7.5]], 10, smooth=True)
