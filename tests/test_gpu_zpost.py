"""crv_invert's fused z pass (the z forward transform of the data and the
means, the eigen-basis posterior of the means and the residuals, and the z
inverse transform of the means, one tile of pencils at a time) against the
same call without it (CRV_TUNE_ZPOST = 0, the default: separate z forward
pass, posterior kernel and z inverse pass): every output equal bit
for bit, on z lengths of several two-stage radix pairs, 1 to 6 angles,
phase-rotated wavelets, odd x sizes, the counter and the reference noise."""
import numpy as np
import pytest

from crava_amd import _lib as L
from crava_amd import model
from pipeline import Case

pytestmark = pytest.mark.gpu

CASES = [
    # (survey, realizations, reference noise stream)
    (lambda: model.config_small(64, 48, 128, 5), 2, False),   # 16 x 8
    (lambda: model.config_small(21, 18, 120, 3), 1, False),   # 15 x 8, odd nxp
    (lambda: model.config_phase(30, 20, 84, 6), 1, False),    # 12 x 7, complex Sigma_e
    (lambda: model.config_small(32, 24, 108, 1), 0, False),   # 12 x 9, no realization
    (lambda: model.config_phase(24, 20, 200, 2, own_error_wavelets=True), 1, True),  # 25 x 8
    (lambda: model.config_phase(18, 16, 320, 5), 1, False),   # 20 x 16 (the production z length)
]


@pytest.mark.parametrize("which", range(len(CASES)))
def test_fused_z_pass_matches_unfused(ctx, which):
    from crava_amd.hip import PostPlan, grid_batch
    make, nreal, refstream = CASES[which]
    case = Case(make())
    s = case.s
    p = case.pin
    cxy, ct = model.prior_spectra(s)
    plan = PostPlan(ctx, s.nxp, s.nyp, s.nzp, 32, s.n, mode=1, A=p.A, theta=p.theta, tables=p.tables, nz=s.nz,
                    lz=s.lz, sigma0=p.sigma0, separable=True, cxy=cxy, ct=ct)

    def run(zpost):
        ctx.tune(L.CRV_TUNE_ZPOST, zpost)
        mean = grid_batch(ctx, 3, s.nxp, s.nyp, s.nzp, 32, 2)
        data = grid_batch(ctx, s.n, s.nxp, s.nyp, s.nzp, 32, 1)
        for g, a in zip(mean, case.syn.mean):
            g.upload_real(a)
        for g, a in zip(data, case.syn.data):
            g.upload_real(a)
        cov = grid_batch(ctx, 6, s.nxp, s.nyp, s.nzp, 32, 3)
        sims = grid_batch(ctx, 3 * nreal, s.nxp, s.nyp, s.nzp) if nreal else []
        ctx.profile(True)
        lcg = plan.invert(mean, data, cov, sims, seed=11, realization0=2, lcg_state=777 if refstream else None)
        ctx.synchronize()
        rep = ctx.profile_report()
        ctx.profile(False)
        nx = s.nxp  # real grids: the pad columns are not part of the result
        out = ([g.download_real()[:, :, :nx] for g in mean], [g.download_complex() for g in data],
               [g.download_real()[:, :, :nx] for g in cov], [g.download_real()[:, :, :nx] for g in sims])
        return out, plan.stats(), rep, lcg

    try:
        got, st1, rep1, lcg1 = run(1)
        want, st0, rep0, lcg0 = run(0)
    finally:
        ctx.tune(L.CRV_TUNE_ZPOST, 0)
    assert "fft_z_post" in rep1 and "fft_z_post" not in rep0
    assert "fft_z_fwd" not in rep1 and "fft_z_fwd" in rep0
    # largest difference of each output family (complex for the residuals)
    diffs = [max((float(np.max(np.abs(w.astype(np.complex128) - g))) for w, g in zip(want[part], got[part])),
                 default=0.0) for part in range(4)]
    assert diffs == [0.0, 0.0, 0.0, 0.0], diffs
    assert st1.n_fp64 == st0.n_fp64 == 0 and st1.n_chol_fail == st0.n_chol_fail
    assert st1.max_cond == st0.max_cond > 0
    assert lcg1 == lcg0
    plan.close()
