"""crv_fill_prior_spectra_noalloc, the allocation-free variant of
crv_fill_prior_spectra (its staging of the lateral and temporal spectra
lives in a fixed device array instead of a per-call allocation): two calls
in a row, the second on a smaller grid, each equal bit for bit to the
allocating entry point on the same inputs; grids beyond the staging area are
refused before any work is queued."""
import ctypes as C

import numpy as np
import pytest

from crava_amd import model

pytestmark = pytest.mark.gpu


def _grids(ctx, s, prec):
    from crava_amd.hip import FFTGrid
    return [FFTGrid(ctx, s.nxp, s.nyp, s.nzp, precision=prec, cubetype=FFTGrid.COVARIANCE) for _ in range(6)]


@pytest.mark.parametrize("prec", [32, 64])
def test_noalloc_prior_spectra_equals_allocating_call(ctx, prec):
    from crava_amd import hip
    big = model.config3_realizations()          # 256x256x200: 129*256 + 200 staged doubles
    small = model.config_small(20, 18, 30, 3)   # 11*18 + 30
    specs = {}
    for s in (big, small):
        cxy, ct = model.prior_spectra(s)
        specs[s.nxp] = (s, cxy, ct)
    # the allocation-free entry point twice in a row, the second call smaller
    got = {}
    for key in (big.nxp, small.nxp):
        s, cxy, ct = specs[key]
        g = _grids(ctx, s, prec)
        hip.fill_prior_spectra(ctx, s.sigma0(), cxy, ct, g, noalloc=True)
        got[key] = g
    for key in (big.nxp, small.nxp):
        s, cxy, ct = specs[key]
        want = _grids(ctx, s, prec)
        hip.fill_prior_spectra(ctx, s.sigma0(), cxy, ct, want)
        for a, b in zip(got[key], want):
            x, y = a.download_complex(), b.download_complex()
            assert x.shape == (s.nzp, s.nyp, s.cnxp)
            assert np.array_equal(x, y)
            assert a.info().is_transformed == 1 and a.info().cubetype == 3
    # a third call after the smaller one: the big grid again, still equal
    s, cxy, ct = specs[big.nxp]
    g = _grids(ctx, s, prec)
    hip.fill_prior_spectra(ctx, s.sigma0(), cxy, ct, g, noalloc=True)
    want = _grids(ctx, s, prec)
    hip.fill_prior_spectra(ctx, s.sigma0(), cxy, ct, want)
    assert np.array_equal(g[5].download_complex(), want[5].download_complex())


def test_noalloc_prior_spectra_rejects_oversized_grid(ctx):
    """cnxp*nyp + nzp = 513*1024 + 1 doubles exceeds the 524288 of the
    staging area: CRV_E_UNSUPPORTED, and the grids keep their state."""
    from crava_amd.hip import FFTGrid
    nxp, nyp, nzp = 1024, 1024, 1
    g = [FFTGrid(ctx, nxp, nyp, nzp, precision=32, cubetype=FFTGrid.COVARIANCE) for _ in range(6)]
    cxy = np.ones((nyp, nxp // 2 + 1))
    ct = np.ones(nzp)
    lib = ctx.lib
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    s0 = np.eye(3)
    covs = (C.c_void_p * 6)(*[x.h.value for x in g])
    rc = lib.crv_fill_prior_spectra_noalloc(ctx.h, dp(s0), dp(cxy), dp(ct), covs)
    assert rc == -5
    assert b"staging area" in lib.crv_last_error(ctx.h)
    assert all(x.info().is_transformed == 0 for x in g)
