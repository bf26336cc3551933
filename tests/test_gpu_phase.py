"""Parity with per-angle wavelets that differ in peak frequency (35 -> 28 Hz)
and in phase: the error covariance is complex Hermitian, the per-plane
matrices of the eigen-basis posterior are complex, and -- with error-norm
wavelets of their own phase, as Wavelet3D's average wavelet gives (reference
src/wavelet3D.cpp:449-460) -- the posterior covariance is complex too.  Every
posterior kernel (FP64 general, FP32 general, FP32 eigen-basis, FP32 Cholesky
records, the FP64 list with refinement) and both realization paths (the
simulation kernel and crv_invert's fused realization) against the FP64
oracle on the same inputs."""
import dataclasses

import numpy as np
import pytest

from conftest import rel_rms
from crava_amd import model
from oracle import oracle as O
from pipeline import Case, gpu_posterior, oracle_posterior

pytestmark = pytest.mark.gpu

CASES = [(20, 18, 30, False), (20, 18, 30, True), (64, 48, 40, True)]


@pytest.fixture(scope="module", params=CASES, ids=lambda c: f"{c[0]}x{c[1]}x{c[2]}{'-own' if c[3] else ''}")
def phase_case(request):
    nx, ny, nz, own = request.param
    case = Case(model.config_phase(nx, ny, nz, 5, own_error_wavelets=own))
    return case, oracle_posterior(case, np.float64), own


def _imag_frac(cov):
    """max |Im| / max |.| over the three off-diagonal covariance spectra."""
    return max(float(np.abs(c.imag).max() / np.abs(c).max()) for c in cov[3:])


def test_oracle_covariance_phase_structure(phase_case):
    case, o, own = phase_case
    f = _imag_frac(o["cov"])
    if own:
        assert f > 1e-3
    else:
        assert f < 1e-9


def test_fp64_general_kernel_equals_oracle(ctx, phase_case):
    case, o, _ = phase_case
    nx = case.s.nxp
    r = gpu_posterior(ctx, case, prec=64)
    for g, m in zip(r["mean"], o["mean"]):
        assert rel_rms(g.download_real()[:, :, :nx], m[:, :, :nx]) < 1e-11
    for g, d in zip(r["data"], o["data"]):
        assert rel_rms(g.download_complex(), d) < 1e-10
    for g, c in zip(r["cov"], o["cov"]):
        got = g.download_complex()
        assert rel_rms(got, c) < 1e-10
        assert np.abs(got.imag - c.imag).max() <= 1e-10 * np.abs(c).max()
    assert r["stats"].n_chol_fail == o["nfail"] == 0


# (label, keyword arguments): the FP32 general kernel (grids, plain), the
# FP32 general kernel with refinement (grids), the eigen-basis kernel
# (separable, refine), and every inverted wavenumber through the FP64 list
FP32_PATHS = [("general", dict(mode=0)), ("general-refine", dict(mode=1)),
              ("eigen", dict(mode=1, separable=True)),
              ("fp64-list", dict(mode=1, separable=True, cond_threshold=0.5))]


@pytest.mark.parametrize("label,kw", FP32_PATHS, ids=[p[0] for p in FP32_PATHS])
def test_fp32_kernels_vs_oracle(ctx, phase_case, label, kw):
    """Means, residuals and covariances within the FP32 targets, and the
    imaginary parts of the covariance within 1e-5 of its largest value."""
    case, o, _ = phase_case
    nx = case.s.nxp
    r = gpu_posterior(ctx, case, prec=32, **kw)
    for a, (g, m) in enumerate(zip(r["mean"], o["mean"])):
        mu = case.s.mu[a]
        assert rel_rms(g.download_real()[:, :, :nx] - mu, m[:, :, :nx] - mu) < 1e-4
    for g, d in zip(r["data"], o["data"]):
        assert rel_rms(g.download_complex(), d) < 1e-4
    for g, c in zip(r["cov"], o["cov"]):
        got = g.download_complex()
        assert rel_rms(got, c) < 1e-4
        assert np.abs(got.imag - c.imag).max() <= 1e-5 * np.abs(c).max()
    st = r["stats"]
    assert st.n_chol_fail == 0 and st.n_inverted > 0
    if label == "fp64-list":
        assert st.n_fp64 == st.n_inverted


def test_cholesky_record_kernel_vs_oracle(ctx, phase_case):
    """wnc = 0 leaves the noise diagonal at zero, so the separable plan runs
    the Cholesky record kernel instead of the eigen-basis one."""
    case, _, own = phase_case
    s = dataclasses.replace(case.s, wnc=0.0)
    c0 = Case(s)
    o = oracle_posterior(c0, np.float64)
    r = gpu_posterior(ctx, c0, prec=32, mode=1, separable=True)
    nx = s.nxp
    for a, (g, m) in enumerate(zip(r["mean"], o["mean"])):
        mu = s.mu[a]
        assert rel_rms(g.download_real()[:, :, :nx] - mu, m[:, :, :nx] - mu) < 1e-4
    for g, c in zip(r["cov"], o["cov"]):
        got = g.download_complex()
        assert rel_rms(got, c) < 1e-4
        assert np.abs(got.imag - c.imag).max() <= 1e-5 * np.abs(c).max()


def _oracle_realization(case, o, seed, realization):
    """FP64 realization: the counter noise (FP64 restatement), scaled by the
    Cholesky factor of the oracle's posterior covariance (oracle simulate
    scaling; wavenumbers whose covariance is zero -- the temporal spectrum
    is cut below the low-frequency limit -- are rejected and get no noise,
    as in the reference), inverse transformed, plus the oracle's posterior
    mean."""
    s = case.s
    z = [O.counter_noise(s.nxp, s.nyp, s.nzp, seed, realization, g) for g in range(3)]
    sc, nf = O.simulate_scale(o["cov"], z, dtype=np.float64)
    assert nf < o["cov"][0].size // 4
    return [O.fft_c2r(sc[a], s.nxp, 2)[:, :, :s.nxp] + o["mean"][a][:, :, :s.nxp] for a in range(3)]


def test_realizations_vs_oracle(ctx, phase_case):
    """crv_invert's fused realization (eigen-basis posterior, FP32 refine)
    and the simulation kernel on the GPU posterior (FP64 3x3 factors) against
    the oracle's FP64 realization of the same counter noise."""
    from crava_amd.hip import PostPlan, grid_batch, simulate_batch
    case, _, _ = phase_case
    s = case.s
    nx = s.nxp
    seed, real0 = 31, 2
    # the realization goes with the square root of the covariance: the
    # oracle takes the exact separable prior spectra, as the GPU plan does
    o = oracle_posterior(case, np.float64, separable=True)
    want = _oracle_realization(case, o, seed, real0)
    cxy, ct = model.prior_spectra(s)
    p = case.pin
    kw = dict(A=p.A, theta=p.theta, tables=p.tables, nz=s.nz, lz=s.lz, sigma0=p.sigma0, separable=True, cxy=cxy,
              ct=ct, skip_mean_inverse=True)
    plan = PostPlan(ctx, s.nxp, s.nyp, s.nzp, 32, s.n, mode=1, **kw)
    batch = lambda n, ct_: grid_batch(ctx, n, s.nxp, s.nyp, s.nzp, 32, ct_)
    mean, data, cov, sims = batch(3, 2), batch(s.n, 1), batch(6, 3), batch(3, 2)
    for g, a in zip(mean, case.syn.mean):
        g.upload_real(a)
    for g, a in zip(data, case.syn.data):
        g.upload_real(a)
    plan.invert(mean, data, cov, sims, seed=seed, realization0=real0)
    for a in range(3):
        mu = s.mu[a]
        assert rel_rms(sims[a].download_real()[:, :, :nx] - mu, want[a] - mu) < 1e-4
    # the simulation kernel on the one-call posterior's covariance spectra
    r = gpu_posterior(ctx, case, prec=32, mode=1, separable=True)
    out = batch(3, 2)
    simulate_batch(ctx, r["cov"], r["mean"], out, seed=seed, realization0=real0, fp64_chol=True)
    for a in range(3):
        mu = s.mu[a]
        assert rel_rms(out[a].download_real()[:, :, :nx] - mu, want[a] - mu) < 1e-4
    plan.close()
