"""The z inverse pass's persistent kernel that loads each tile's inputs while
the previous tile is transformed (CRV_TUNE_ZINV: by default for the
covariances' generating z inverse of crv_invert, 1 for every z inverse)
against the one-tile-per-block kernels (CRV_TUNE_ZINV = 2): invFFTInPlace
over z lengths of several two-stage radix pairs and odd x sizes (a last tile
narrower than the others), and crv_invert, whose covariances' z inverse
generates the planes outside the inverted band; every output equal bit for
bit."""
import numpy as np
import pytest

from crava_amd import _lib as L
from crava_amd import model
from pipeline import Case

pytestmark = pytest.mark.gpu

SHAPES = [(21, 18, 320), (64, 12, 200), (35, 10, 84), (16, 9, 120), (30, 8, 128), (20, 6, 450), (18, 6, 225)]


def _inverse(ctx, shape, x, zinv):
    from crava_amd.hip import FFTGrid
    ctx.tune(L.CRV_TUNE_ZINV, zinv)
    g = FFTGrid(ctx, *shape, precision=32, cubetype=L.CRV_PARAMETER)
    g.upload_real(x)
    g.fftInPlace()
    g.invFFTInPlace()
    out = g.download_real()
    g.close()
    return out


@pytest.mark.parametrize("shape", SHAPES)
def test_inverse_transform_matches_one_tile_per_block(ctx, shape):
    nx, ny, nz = shape
    rnx = 2 * (nx // 2 + 1)
    x = np.random.default_rng(nx * ny + nz).standard_normal((nz, ny, rnx)).astype(np.float32)
    try:
        got = _inverse(ctx, shape, x, 1)[:, :, :nx]
        want = _inverse(ctx, shape, x, 2)[:, :, :nx]
    finally:
        ctx.tune(L.CRV_TUNE_ZINV, 0)
    assert np.array_equal(got, want)
    assert np.max(np.abs(got - x[:, :, :nx])) < 1e-4 * np.max(np.abs(x))


CASES = [
    (lambda: model.config_small(64, 48, 128, 5), 2, False),   # 16 x 8
    (lambda: model.config_phase(30, 20, 84, 6), 1, False),    # 12 x 7, complex Sigma_e
    (lambda: model.config_phase(24, 20, 200, 2, own_error_wavelets=True), 1, True),  # 25 x 8
    (lambda: model.config_phase(18, 16, 320, 5), 1, False),   # 20 x 16 (the production z length)
    (lambda: model.config_small(21, 18, 120, 3), 1, False),   # 15 x 8, odd nxp
]


@pytest.mark.parametrize("zinv", [0, 1])
@pytest.mark.parametrize("which", range(len(CASES)))
def test_invert_matches_one_tile_per_block(ctx, which, zinv):
    from crava_amd.hip import PostPlan, grid_batch
    make, nreal, refstream = CASES[which]
    case = Case(make())
    s = case.s
    p = case.pin
    cxy, ct = model.prior_spectra(s)
    plan = PostPlan(ctx, s.nxp, s.nyp, s.nzp, 32, s.n, mode=1, A=p.A, theta=p.theta, tables=p.tables, nz=s.nz,
                    lz=s.lz, sigma0=p.sigma0, separable=True, cxy=cxy, ct=ct)

    def run(zinv):
        ctx.tune(L.CRV_TUNE_ZINV, zinv)
        mean = grid_batch(ctx, 3, s.nxp, s.nyp, s.nzp, 32, 2)
        data = grid_batch(ctx, s.n, s.nxp, s.nyp, s.nzp, 32, 1)
        for g, a in zip(mean, case.syn.mean):
            g.upload_real(a)
        for g, a in zip(data, case.syn.data):
            g.upload_real(a)
        cov = grid_batch(ctx, 6, s.nxp, s.nyp, s.nzp, 32, 3)
        sims = grid_batch(ctx, 3 * nreal, s.nxp, s.nyp, s.nzp) if nreal else []
        plan.invert(mean, data, cov, sims, seed=5, realization0=1, lcg_state=4242 if refstream else None)
        ctx.synchronize()
        nx = s.nxp
        return ([g.download_real()[:, :, :nx] for g in mean], [g.download_complex() for g in data],
                [g.download_real()[:, :, :nx] for g in cov], [g.download_real()[:, :, :nx] for g in sims])

    try:
        got = run(zinv)
        want = run(2)
    finally:
        ctx.tune(L.CRV_TUNE_ZINV, 0)
    for part in range(4):
        for w, g in zip(want[part], got[part]):
            assert np.array_equal(w, g), part
    plan.close()
