"""Crava::computePostMeanResidAndFFTCov on the GPU against the reference's
own float loop (golden vectors from the compiled lib_matr), the FP64 oracle,
and the closed-form special cases."""
import numpy as np
import pytest

from conftest import cplx, rel_rms
from crava_amd import model
from pipeline import Case, gpu_posterior, oracle_posterior

pytestmark = pytest.mark.gpu


def _golden_run(ctx, golden, prec, mode=0):
    from crava_amd.hip import FFTGrid, computePostMeanResidAndFFTCov
    from crava_amd.model import KzTables
    s = model.config_small(12, 10, 16, 3)
    mk = lambda ct: FFTGrid(ctx, s.nxp, s.nyp, s.nzp, precision=prec, cubetype=ct)
    mean = [mk(2) for _ in range(3)]
    data = [mk(1) for _ in range(3)]
    cov = [mk(3) for _ in range(6)]
    err = mk(3)
    for g, a in zip(mean, golden["post_in_mean"]):
        g.upload_complex(cplx(a))
    for g, a in zip(data, golden["post_in_data"]):
        g.upload_complex(cplx(a))
    for g, a in zip(cov, golden["post_in_cov"]):
        g.upload_complex(cplx(a))
    err.upload_complex(cplx(golden["post_in_err"][0]))
    tab = KzTables(cplx(golden["post_kwd"]), cplx(golden["post_em1"]), cplx(golden["post_em2"]),
                   cplx(golden["post_em3"]))
    st = computePostMeanResidAndFFTCov(ctx, mean, data, err, cov, A=golden["post_A"], theta=golden["post_theta"],
                                       tables=tab, nz=s.nz, lz=s.lz, sigma0=golden["post_sigma0"], mode=mode,
                                       skip_mean_inverse=True)
    return mean, data, cov, st


@pytest.mark.parametrize("prec,mode,tol", [(32, 0, 2e-5), (64, 0, 2e-5), (32, 1, 2e-5)])
def test_matches_reference_float_loop(ctx, golden, prec, mode, tol):
    """Same spectra in, compared with the reference's compiled float loop."""
    mean, data, cov, st = _golden_run(ctx, golden, prec, mode)
    assert st.n_chol_fail == int(golden["post_nfail"])
    assert st.n_inverted == int(golden["post_invert"].sum()) * 10 * 7
    for g, r in zip(mean, golden["post_out_mean"]):
        assert rel_rms(g.download_complex(), cplx(r)) < tol
    for g, r in zip(data, golden["post_out_data"]):
        assert rel_rms(g.download_complex(), cplx(r)) < tol
    for g, r in zip(cov, golden["post_out_cov"]):
        assert rel_rms(g.download_complex(), cplx(r)) < tol


@pytest.fixture(scope="module")
def case_small():
    return Case(model.config_small(20, 18, 30, 5))


@pytest.fixture(scope="module")
def oracle_small(case_small):
    return oracle_posterior(case_small, np.float64)


def test_fp64_gpu_equals_fp64_oracle(ctx, case_small, oracle_small):
    r = gpu_posterior(ctx, case_small, prec=64)
    o = oracle_small
    for g, m in zip(r["mean"], o["mean"]):
        assert rel_rms(g.download_real()[:, :, :20], m[:, :, :20]) < 1e-12
    for g, d in zip(r["data"], o["data"]):
        assert rel_rms(g.download_complex(), d) < 1e-11
    for g, c in zip(r["cov"], o["cov"]):
        assert rel_rms(g.download_complex(), c) < 1e-11
    assert r["stats"].n_chol_fail == o["nfail"] == 0


@pytest.mark.parametrize("prec", [32, 64])
def test_separable_prior_equals_grids(ctx, case_small, prec):
    a = gpu_posterior(ctx, case_small, prec=prec)
    b = gpu_posterior(ctx, case_small, prec=prec, separable=True)
    # the covariance grids hold float-rounded correlations (as fillInParamCorr
    # builds them in float); the separable spectra are exact
    tol = 2e-6 if prec == 32 else 5e-7
    for x, y in zip(a["mean"], b["mean"]):
        assert rel_rms(x.download_real()[:, :, :20], y.download_real()[:, :, :20]) < tol
    for x, y in zip(a["cov"], b["cov"]):
        assert rel_rms(x.download_complex(), y.download_complex()) < tol


def test_fp32_refine_close_to_fp64(ctx, case_small, oracle_small):
    r = gpu_posterior(ctx, case_small, prec=32, mode=1)
    o = oracle_small
    for g, m in zip(r["mean"], o["mean"]):
        assert rel_rms(g.download_real()[:, :, :20], m[:, :, :20]) < 1e-6
    for g, c in zip(r["cov"], o["cov"]):
        assert rel_rms(g.download_complex(), c) < 1e-5
    st = r["stats"]
    assert st.n_inverted > 0 and st.max_cond > 1.0


def test_non_inverted_band_keeps_cleaned_prior(ctx, case_small):
    """Outside the frequency band the means/residuals are untouched and the
    covariance is the cleaned prior Sigma0 |Re c / Sigma0| (im = 0)."""
    s = case_small.s
    r = gpu_posterior(ctx, case_small, prec=64, skip_mean_inverse=True)
    inv = case_small.pin.invert.astype(bool)
    from oracle import oracle as O
    data0 = O.fft_r2c(case_small.syn.data[0], s.nxp, 1)
    cov0 = O.fft_r2c(case_small.cov_real[0], s.nxp, 3)
    d = r["data"][0].download_complex()
    c = r["cov"][0].download_complex()
    assert np.allclose(d[~inv], data0[~inv], rtol=1e-12, atol=1e-12)
    want = case_small.pin.sigma0[0, 0] * np.abs(cov0[~inv].real / case_small.pin.sigma0[0, 0])
    assert np.allclose(c[~inv].real, want, rtol=1e-12) and np.all(c[~inv].imag == 0)


@pytest.mark.parametrize("prec,mode", [(64, 0), (32, 0), (32, 1)])
def test_forced_cholesky_failure_is_prior(ctx, prec, mode):
    """wnc = 0, a zero error spectrum and a zero prior on one kz plane make
    Re M00 = 0 there: the factorisation is rejected, so posterior = prior and
    residual = data on that plane (crava.cpp:1421-1437)."""
    from crava_amd.hip import FFTGrid, computePostMeanResidAndFFTCov
    from oracle import oracle as O
    case = Case(model.config_small(8, 6, 16, 3))
    s = case.s
    k = int(np.nonzero(case.pin.invert)[0][0])
    mk = lambda ct: FFTGrid(ctx, s.nxp, s.nyp, s.nzp, precision=prec, cubetype=ct)
    mean = [mk(2) for _ in range(3)]
    data = [mk(1) for _ in range(3)]
    cov = [mk(3) for _ in range(6)]
    err = mk(3)
    data0 = [O.fft_r2c(d, s.nxp, 1) for d in case.syn.data]
    for g, a in zip(mean, case.syn.mean):
        g.upload_complex(O.fft_r2c(a, s.nxp, 2))
    for g, a in zip(data, data0):
        g.upload_complex(a)
    for g, c in zip(cov, case.cov_real):
        spec = O.fft_r2c(c, s.nxp, 3)
        spec[k] = 0.0
        g.upload_complex(spec)
    err.upload_complex(np.zeros_like(data0[0]))
    p = case.pin
    st = computePostMeanResidAndFFTCov(ctx, mean, data, err, cov, A=p.A, theta=p.theta, tables=p.tables, nz=s.nz,
                                       lz=s.lz, wnc=0.0, sigma0=p.sigma0, mode=mode, skip_mean_inverse=True)
    assert st.n_chol_fail >= s.cnxp * s.nyp
    tol = 1e-12 if prec == 64 else 1e-6
    for g, d in zip(data, data0):
        got = g.download_complex()[k]
        assert np.allclose(got, d[k], rtol=tol, atol=tol * np.abs(d[k]).max())
    for g in cov:
        assert np.all(g.download_complex()[k] == 0)


def test_c0_end_to_end_fp32_vs_fp64_oracle(ctx):
    """Config 0 (64x64x128, 3 angles): the FP32 GPU pipeline (plain, the
    reference's precision) against the FP64 oracle."""
    case = Case(model.config0())
    o = oracle_posterior(case, np.float64, nthreads=8)
    r = gpu_posterior(ctx, case, prec=32)
    for a, (g, m) in enumerate(zip(r["mean"], o["mean"])):
        mu = case.s.mu[a]
        assert rel_rms(g.download_real()[:, :, :64] - mu, m[:, :, :64] - mu) < 1e-4
    for g, c in zip(r["cov"], o["cov"]):
        assert rel_rms(g.download_complex(), c) < 1e-4


@pytest.mark.parametrize("n", [1, 2, 4, 6, 7, 8])
def test_angle_counts_against_fp64_oracle(ctx, n):
    """Angle counts other than the configs' 3 and 5: 1..6 run the per-plane
    kernels in FP32 (the eigen-basis kernel with refinement; 6 fills its
    largest record), 7 and 8 the general one.  FP64 equals the FP64
    oracle; FP32 with refinement stays within the precision target; the
    separable prior equals the grid path."""
    case = Case(model.config_small(16, 12, 20, n))
    o = oracle_posterior(case, np.float64)
    r64 = gpu_posterior(ctx, case, prec=64)
    for g, m in zip(r64["mean"], o["mean"]):
        assert rel_rms(g.download_real()[:, :, :16], m[:, :, :16]) < 1e-11
    for g, c in zip(r64["cov"], o["cov"]):
        assert rel_rms(g.download_complex(), c) < 1e-10
    for g, d in zip(r64["data"], o["data"]):
        assert rel_rms(g.download_complex(), d) < 1e-10
    r32 = gpu_posterior(ctx, case, prec=32, mode=1, separable=True)
    for g, m in zip(r32["mean"], o["mean"]):
        assert rel_rms(g.download_real()[:, :, :16], m[:, :, :16]) < 1e-5
    for g, c in zip(r32["cov"], o["cov"]):
        assert rel_rms(g.download_complex(), c) < 1e-4
    assert r32["stats"].n_chol_fail == 0 and r32["stats"].n_inverted > 0


@pytest.mark.parametrize("separable", [True, False])
def test_fp64_list_takes_every_wavenumber(ctx, separable):
    """FP32 refine mode with a condition threshold below 1 (every estimate is
    at least 1) queues every inverted wavenumber for the FP64 solve with one
    refinement step: the count equals the inverted count, and the results,
    FP64 solves of the FP32 spectra, stay within FP32 rounding of the FP64
    oracle."""
    case = Case(model.config_small(16, 12, 20, 5))
    o = oracle_posterior(case, np.float64)
    r = gpu_posterior(ctx, case, prec=32, mode=1, separable=separable, cond_threshold=0.5)
    st = r["stats"]
    assert st.n_fp64 == st.n_inverted > 0 and st.n_chol_fail == 0
    for g, m in zip(r["mean"], o["mean"]):
        assert rel_rms(g.download_real()[:, :, :16], m[:, :, :16]) < 1e-6
    for g, c in zip(r["cov"], o["cov"]):
        assert rel_rms(g.download_complex(), c) < 1e-5
    for g, d in zip(r["data"], o["data"]):
        assert rel_rms(g.download_complex(), d) < 1e-4


def test_separable_refine_without_white_noise(ctx):
    """wnc = 0 leaves the noise diagonal D at zero, so the plan has no
    eigen-basis records (M = lam E has no D to scale by) and runs the
    Cholesky record kernel; the results still match the FP64 oracle."""
    import dataclasses
    case = Case(dataclasses.replace(model.config_small(16, 12, 20, 3), wnc=0.0))
    o = oracle_posterior(case, np.float64)
    r = gpu_posterior(ctx, case, prec=32, mode=1, separable=True)
    for g, m in zip(r["mean"], o["mean"]):
        assert rel_rms(g.download_real()[:, :, :16], m[:, :, :16]) < 1e-5
    for g, c in zip(r["cov"], o["cov"]):
        assert rel_rms(g.download_complex(), c) < 1e-4
    assert r["stats"].n_inverted > 0
