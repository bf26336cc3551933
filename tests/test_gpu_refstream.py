"""The reference noise stream (RandomGen + fillInComplexNoise, reference
libs/lib/random.cpp:38-119 and src/fftgrid.cpp:1191-1240) in the one-call
inversion: crv_invert_noise draws the realizations' noise from an LCG
state as Crava::simulate does (seed0, seed1, seed2 of each realization in
turn, crava.cpp:1598-1603), so its realizations equal crv_simulate's with
the same state, one realization after another, and it returns the same
final state."""
import numpy as np
import pytest

from conftest import rel_rms
from crava_amd import model
from pipeline import Case

pytestmark = pytest.mark.gpu


def _kw(case):
    s, p = case.s, case.pin
    cxy, ct = model.prior_spectra(s)
    return dict(A=p.A, theta=p.theta, tables=p.tables, nz=s.nz, lz=s.lz, sigma0=p.sigma0, separable=True, cxy=cxy,
                ct=ct, skip_mean_inverse=True)


@pytest.mark.parametrize("prec,chunk_rows,dims,n", [(64, 0, (32, 24, 20), 3), (32, 5, (64, 48, 40), 5),
                                                    (32, 0, (64, 48, 40), 5), (32, 0, (21, 18, 30), 3)])
def test_invert_reference_stream_matches_simulate(ctx, prec, chunk_rows, dims, n):
    """Two realizations from state 12345: the FP64 plan (general kernels) and
    the FP32 refine plan by chunks of y rows equal the sequence of
    crv_posterior_exec and two crv_simulate calls bit for bit; the
    whole-grid FP32 call (simulation fused into the posterior) differs only
    by the out-of-band prior factor's rounding (rel RMS below 1e-6)."""
    from crava_amd.hip import PostPlan, grid_batch, simulate
    case = Case(model.config_small(*dims, n))
    s = case.s
    plan = PostPlan(ctx, s.nxp, s.nyp, s.nzp, prec, s.n, mode=0 if prec == 64 else 1, **_kw(case))
    batch = lambda k, ct: grid_batch(ctx, k, s.nxp, s.nyp, s.nzp, prec, ct)

    def inputs():
        mean, data = batch(3, 2), batch(s.n, 1)
        for g, a in zip(mean, case.syn.mean):
            g.upload_real(a)
        for g, a in zip(data, case.syn.data):
            g.upload_real(a)
        return mean, data

    s0, nreal = 12345, 2
    mean, data = inputs()
    cov, sims = batch(6, 3), batch(3 * nreal, 2)
    for g in data + mean:
        g.fftInPlace()
    plan.exec(mean, data, None, cov)
    for g in mean:
        g.invFFTInPlace()
    st = s0
    for r in range(nreal):
        st = simulate(ctx, cov, mean, sims[3 * r:3 * r + 3], noise_mode=0, lcg_state=st,
                      fp64_chol=2 if prec == 32 else 0)
    for g in cov:
        g.invFFTInPlace()
    ctx.synchronize()
    want = [g.download_real() for g in mean + cov + sims]
    mean2, data2 = inputs()
    cov2, sims2 = batch(6, 3), batch(3 * nreal, 2)
    st2 = plan.invert(mean2, data2, cov2, sims2, lcg_state=s0, chunk_rows=chunk_rows)
    assert st2 == st
    got = [g.download_real() for g in mean2 + cov2 + sims2]
    nx = s.nxp
    for q, (w, g) in enumerate(zip(want, got)):
        if q >= 9 and prec == 32 and chunk_rows == 0:
            assert rel_rms(g[:, :, :nx], w[:, :, :nx]) < 1e-6
        else:
            assert np.array_equal(w[:, :, :nx], g[:, :, :nx]), f"grid {q}"
    plan.close()


def test_reference_and_counter_realizations_agree_in_distribution(ctx):
    """The same posterior through both streams: each realization's spectrum
    divided by the posterior variance spectrum is unit complex white noise,
    so its mean square over wavenumbers agrees between the streams within
    sampling error (about 0.3 % for 4 realizations of 64x48x40)."""
    from crava_amd.hip import PostPlan, grid_batch
    case = Case(model.config_small(64, 48, 40, 3))
    s = case.s
    plan = PostPlan(ctx, s.nxp, s.nyp, s.nzp, 32, s.n, mode=1, **_kw(case))
    batch = lambda k, ct: grid_batch(ctx, k, s.nxp, s.nyp, s.nzp, 32, ct)
    nx, nreal = s.nxp, 4
    ms = []
    for ref in (True, False):
        mean, data = batch(3, 2), batch(s.n, 1)
        for g, a in zip(mean, case.syn.mean):
            g.upload_real(a)
        for g, a in zip(data, case.syn.data):
            g.upload_real(a)
        cov, sims = batch(6, 3), batch(3 * nreal, 2)
        if ref:
            plan.invert(mean, data, cov, sims, lcg_state=777)
        else:
            plan.invert(mean, data, cov, sims, seed=777)
        m = [g.download_real()[:, :, :nx].astype(np.float64) for g in mean]
        out = []
        for a in range(3):
            var = np.abs(np.fft.rfftn(cov[a].download_real()[:, :, :nx].astype(np.float64)))
            keep = var > 1e-6 * var.max()
            z = [np.abs(np.fft.rfftn(sims[3 * r + a].download_real()[:, :, :nx] - m[a])) ** 2 / np.where(keep, var, 1)
                 for r in range(nreal)]
            out.append(np.mean([zz[keep].mean() for zz in z]))
        ms.append(np.array(out))
    ratio = ms[0] / ms[1]
    print("reference / counter mean square per parameter:", ratio)
    assert np.all(np.abs(ratio - 1) < 0.03), ratio
    plan.close()
