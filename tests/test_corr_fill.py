"""The general prior's correlation grids (SeismicParametersHolder::
initializeCorrelations, reference src/seismicparametersholder.cpp:197-216;
FFTGrid::fillInParamCorr / fillInErrCorr, src/fftgrid.cpp:1018-1117).

CPU: the oracle's fills against closed forms -- no gradient gives the
separable product corrXY(i, j) circCorrT(k); a gradient of one sample per
cell shifts the temporal correlation by the lateral offset; half a sample
averages two neighbours; the error grid holds corrXY on plane 0 only, and
with a gradient the two planes around the shifted position share the
weight, 1 - f and f, where the shift is positive (for a negative shift f
in (-1, 0) the reference's truncation toward zero gives plane 0 the weight
1 + |f|, a quirk the fill keeps).  GPU: crv_fill_param_corr and
crv_fill_err_corr equal the oracle bit for bit, and the general posterior
over grids with gradients matches the FP64 oracle."""
import numpy as np
import pytest

from oracle import oracle as O

NX, NY, NZ = 12, 10, 16
RNX = 2 * (NX // 2 + 1)
S6 = (1.0, 0.5, 0.3, 0.4, 0.2, 0.1)


def _inputs(nx=NX, ny=NY, nz=NZ):
    rng = np.random.default_rng(5)
    cxy = rng.uniform(0.2, 1.0, (ny, nx))
    ct = rng.uniform(-0.3, 1.0, nz).astype(np.float32)
    return cxy, ct


def test_param_corr_without_gradient_is_the_product():
    cxy, ct = _inputs()
    g = O.param_corr_grids(cxy, ct, 0.0, 0.0, S6, NX, NY, NZ, RNX)
    want = (ct[:, None, None] * cxy.astype(np.float32)[None]).astype(np.float32)
    for q, s in enumerate(S6):
        assert np.array_equal(g[q][:, :, :NX], (want * np.float32(s)).astype(np.float32))
        assert np.all(g[q][:, :, NX:] == np.float32(-99999.0) * np.float32(s))


def test_param_corr_integer_and_half_gradients():
    cxy, ct = _inputs()
    ci = np.where(np.arange(NX) < NX - np.arange(NX), np.arange(NX), np.arange(NX) - NX)
    g1 = O.param_corr_grids(cxy, ct, 1.0, 0.0, (1.0,) * 6, NX, NY, NZ, RNX)[0]
    k = np.arange(NZ)[:, None, None]
    shifted = ct[(k + ci[None, None, :]) % NZ] * cxy.astype(np.float32)[None]
    assert np.array_equal(g1[:, :, :NX], np.broadcast_to(shifted, g1[:, :, :NX].shape).astype(np.float32))
    gh = O.param_corr_grids(cxy, ct, 0.0, 0.5, (1.0,) * 6, NX, NY, NZ, RNX)[0]
    cj = np.where(np.arange(NY) < NY - np.arange(NY), np.arange(NY), np.arange(NY) - NY)
    j = 3  # cj = 3: shift 1.5 samples
    assert cj[j] == 3
    lo, hi = (np.arange(NZ) + 1) % NZ, (np.arange(NZ) + 2) % NZ
    want = (0.5 * ct[lo] + 0.5 * ct[hi])[:, None] * cxy[j][None, :]
    assert np.allclose(gh[:, j, :NX], want, rtol=2e-7, atol=1e-7)


def test_err_corr_closed_forms():
    cxy, _ = _inputs()
    c32 = cxy.astype(np.float32)
    e0 = O.err_corr_grid(cxy, 0.0, 0.0, NX, NY, NZ, RNX)
    assert np.array_equal(e0[0, :, :NX], c32) and not np.any(e0[1:, :, :NX])
    assert np.all(e0[:, :, NX:] == np.float32(-99999.0))
    e = O.err_corr_grid(cxy, 0.3, 0.0, NX, NY, NZ, RNX)
    c = lambda i: c32[0, i]
    # column 1 (offset +1, shift +0.3): plane 0 at s = 0.3 weighs 0.7, plane
    # nzp - 1 at s = nzp - 0.7 (next to plane nzp = 0) weighs 0.3
    assert np.isclose(e[0, 0, 1], np.float32(0.7) * c(1), rtol=1e-6)
    assert np.isclose(e[NZ - 1, 0, 1], np.float32(0.3) * c(1), rtol=1e-6)
    assert not np.any(e[1:NZ - 1, 0, 1])
    # column nxp - 1 (offset -1, shift -0.3): plane 1 at s = 0.7 weighs 0.3;
    # plane 0 at s = -0.3 keeps the fraction -0.3 after int() and weighs
    # 1.3 (not 0.7), the reference's truncation toward zero
    assert np.isclose(e[1, 0, NX - 1], np.float32(0.3) * c(NX - 1), rtol=1e-6)
    assert np.isclose(e[0, 0, NX - 1], np.float32(1.3) * c(NX - 1), rtol=1e-6)
    assert e[NZ - 1, 0, NX - 1] == 0
    # column nxp - 2 (offset -2, shift -0.6): plane 1 weighs 0.6, plane 0 1.6
    assert np.isclose(e[1, 0, NX - 2], np.float32(0.6) * c(NX - 2), rtol=1e-6)
    assert np.isclose(e[0, 0, NX - 2], np.float32(1.6) * c(NX - 2), rtol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("grads", [(0.0, 0.0), (0.37, -0.21), (1.0, 0.5), (-2.25, 0.125)])
def test_gpu_fills_equal_oracle(ctx, prec, grads):
    from crava_amd import hip
    cxy, ct = _inputs(20, 18, 24)
    nx, ny, nz = 20, 18, 24
    rnx = 2 * (nx // 2 + 1)
    cov = [hip.FFTGrid(ctx, nx, ny, nz, precision=prec, cubetype=3) for _ in range(6)]
    err = hip.FFTGrid(ctx, nx, ny, nz, precision=prec, cubetype=3)
    hip.fill_param_corr(ctx, cov, cxy, ct, *grads, S6)
    hip.fill_err_corr(ctx, err, cxy, *grads)
    want = O.param_corr_grids(cxy, ct, *grads, S6, nx, ny, nz, rnx)
    for g, w in zip(cov, want):
        assert np.array_equal(g.download_real(), w.astype(g.download_real().dtype))
    we = O.err_corr_grid(cxy, *grads, nx, ny, nz, rnx)
    assert np.array_equal(err.download_real(), we.astype(err.download_real().dtype))


@pytest.mark.gpu
def test_general_posterior_with_correlation_gradients(ctx):
    """A non-separable prior: the six covariance grids and errCorr filled on
    the device with gradients (0.3, -0.2), then the general posterior (FP64
    and FP32 refine), against the FP64 oracle on the same grids."""
    from conftest import rel_rms
    from crava_amd import hip, model
    from pipeline import Case, gpu_posterior, oracle_posterior
    s = model.config_small(24, 20, 32, 3)
    case = Case(s)
    cxy = model.corr_xy(s)
    ct = model.circ_corr_t(s).astype(np.float32)
    gi, gj = 0.3, -0.2
    S0 = s.sigma0()
    case.cov_real = [g.astype(np.float64) for g in O.param_corr_grids(cxy, ct, gi, gj, S0, s.nxp, s.nyp, s.nzp,
                                                                         s.rnxp)]
    case.err_real = O.err_corr_grid(cxy, gi, gj, s.nxp, s.nyp, s.nzp, s.rnxp).astype(np.float64)
    cov = [hip.FFTGrid(ctx, s.nxp, s.nyp, s.nzp, precision=32, cubetype=3) for _ in range(6)]
    hip.fill_param_corr(ctx, cov, cxy, ct, gi, gj, S0)
    for g, w in zip(cov, case.cov_real):
        assert np.array_equal(g.download_real(), w.astype(np.float32))
    o = oracle_posterior(case, np.float64)
    for prec, mode in ((64, 0), (32, 1)):
        r = gpu_posterior(ctx, case, prec=prec, mode=mode, separable=False)
        tol = 1e-9 if prec == 64 else 1e-4
        for a, (g, m) in enumerate(zip(r["mean"], o["mean"])):
            mu = s.mu[a]
            assert rel_rms(g.download_real()[:, :, :s.nxp] - mu, m[:, :, :s.nxp] - mu) < tol
        for g, c in zip(r["cov"], o["cov"]):
            assert rel_rms(g.download_complex(), c) < tol
