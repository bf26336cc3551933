"""Model set-up around the path (A-matrix, wavelets, per-kz operators,
correlations, band mask, padding) against closed forms.  CPU only."""
import math

import numpy as np
import pytest

from crava_amd import model


def test_reflection_matrix_closed_form():
    # PP Aki-Richards rows for vs/vp = 0.5 (values of the 3-angle survey)
    A = model.reflection_matrix((0.0, 15.0, 30.0), 0.5)
    want = np.array([[0.5, 0.0, 0.5], [0.53590, -0.06699, 0.46651], [2 / 3, -0.25, 0.375]])
    assert np.allclose(A, want, atol=5e-5)
    assert A.dtype == np.float32


@pytest.mark.parametrize("cfg,count", [(model.config0(), 52), (model.config1(), 128),
                                       (model.Survey(256, 256, 200), 78)])
def test_inverted_band(cfg, count):
    inv = model.invert_mask(cfg)
    assert int(inv.sum()) == count
    f = np.arange(cfg.nzp) / (cfg.nzp * cfg.dt * 1e-3)
    f = np.minimum(f, 1000.0 / cfg.dt - f)
    assert np.all((f[inv == 1] > 5.0) & (f[inv == 1] < 55.0))


def test_ricker_closed_form():
    w = model.wavelet_fft_order(128, 4.0, 30.0)
    assert w[0] == pytest.approx(1.0)
    t = 8.0
    c = (math.pi * 30.0 * t * 1e-3) ** 2
    assert w[2] == pytest.approx((1 - 2 * c) * math.exp(-c), rel=1e-6)
    assert w[-2] == w[2]                      # FFT order, symmetric
    assert np.all(w[30:99] == 0.0)            # beyond the +-119 ms support


def test_d1_spectrum_closed_form():
    nzp = 64
    D1 = model.full_spectrum(model.d1_operator(nzp))
    k = np.arange(nzp)
    assert np.allclose(D1, -1.0 + np.exp(2j * np.pi * k / nzp), atol=1e-12)


def test_kz_tables_structure():
    s = model.config0()
    t = model.kz_tables(s)
    # K scalar = D1 conj(W); W real and even -> conj(W) = W (real)
    W = model.full_spectrum(t.wavelets[0])
    assert np.allclose(W.imag, 0, atol=1e-9)
    assert np.allclose(t.em1[:, 0] * model.norm_time(t.wavelets[0]), np.conj(W))
    assert np.allclose(t.em3, 1.0)
    # Hermitian in kz: value at nzp-k is the conjugate
    k = 7
    assert np.allclose(t.kwd[s.nzp - k], np.conj(t.kwd[k]))


def test_circular_temporal_correlation():
    s = model.config0()
    c = model.circ_corr_t(s)
    assert c[0] == pytest.approx(1.0)
    assert np.allclose(c[1:], c[1:][::-1], atol=1e-7)          # even, circular
    spec = np.fft.fft(c.astype(np.float64)).real
    low = int(math.floor(np.float32(s.low_cut) * (s.nzp * 0.001 * np.float32(s.dt))))
    assert np.allclose(spec[: low + 1], 0.0, atol=1e-5)        # zeroed below the low cut
    assert np.all(spec > -1e-5)                                 # positive definite


def test_spherical_lateral_correlation():
    s = model.Survey(128, 96, 16)
    c = model.corr_xy(s)
    assert c[0, 0] == 1.0
    assert c[0, 40] == 0.0 and c[0, 60] == 0.0                  # lag >= 1000 m = range
    lag = 20 * s.dx
    d = lag / s.lateral_range
    assert c[0, 20] == pytest.approx(1 + d ** 3 / 2 - 1.5 * d, rel=1e-6)
    assert c[0, 128 - 20] == c[0, 20]                          # cyclic lag
    assert c[96 - 7, 0] == c[7, 0]


def test_prior_spectra_separable_equals_grid_fft():
    s = model.config_small(12, 10, 16, 3)
    cxy, ct = model.prior_spectra(s)
    g = model.prior_cov_grids(s, np.float64)[0][:, :, : s.nxp]
    G = np.fft.rfftn(g)
    want = s.sigma0()[0, 0] * ct[:, None, None] * cxy[None]
    # the grids hold float32 values; the spectra agree to float rounding
    assert np.abs(G.real - want).max() < 1e-6 * np.abs(want).max()
    assert np.abs(G.imag).max() < 1e-9 * np.abs(want).max()


def test_padding_rule_c3_volumes():
    # every un-padded size of the uneven-volume batch pads to a 2,3,5,7-smooth size >= n
    for n in (64, 96, 128, 160, 192, 240, 320):
        p, _, _ = model.estimate_padding(n, n, 100, 25.0, 25.0, 4.0)
        m = p
        for f in (2, 3, 5, 7):
            while m % f == 0:
                m //= f
        assert m == 1 and p >= n
    assert model.padding_size(320, 0.0625) == 343


def test_err_theta_cov():
    s = model.config0()
    th = model.err_theta_cov(s, [2.0, 3.0, 4.0])
    ev = np.array([2.0, 3.0, 4.0]) / 5.0
    r = math.radians(10.0)
    assert th[0, 0] == pytest.approx(ev[0], rel=1e-6)
    assert th[0, 1] == pytest.approx(math.sqrt(ev[0] * ev[1]) * math.exp(-3 * math.radians(15) / r), rel=1e-5)
    assert np.allclose(th, th.T)


def test_circ_corr_t_library_equals_model():
    """crv_circ_corr_t (the library's computeCircCorrT + makeCircCorrTPosDef)
    against the numpy restatement, on the configurations' temporal
    correlation and grids."""
    from crava_amd import hip
    for s in (model.config0(), model.config1(), model.config3_realizations(), model.config_small(12, 10, 16)):
        ncorr = s.nzp // 2 + 1 if s.nzp % 2 == 0 else s.nzp // 2
        lags = np.arange(ncorr)
        prior = np.exp(-3.0 * (lags * s.dt / s.temporal_range) ** 2).astype(np.float32)
        cut = int(math.floor(np.float32(s.low_cut) * (s.nzp * 0.001 * np.float32(s.dt))))
        got = hip.circ_corr_t(prior, s.nzp, cut)
        assert np.allclose(got, model.circ_corr_t(s), rtol=0, atol=2e-7)


def test_circ_corr_t_closed_forms():
    """A correlation whose circular spectrum is already positive: the
    modulus changes nothing, so with only the DC bin cut the result is
    (c - mean c) / (c0 - mean c); the output is even, 1 at lag 0, has a
    non-negative spectrum that is 0 up to the cut, and a second pass
    returns it unchanged."""
    from crava_amd import hip
    nzp = 64
    k = np.arange(nzp)
    lag = np.minimum(k, nzp - k)
    c = np.exp(-np.abs(lag) / 3.0)          # exponential: positive circular spectrum
    F = np.fft.fft(c).real
    assert F.min() > 0
    got = hip.circ_corr_t(c[: nzp // 2 + 1].astype(np.float32), nzp, 0).astype(np.float64)
    cf = c.astype(np.float32).astype(np.float64)
    want = (cf - cf.mean()) / (cf[0] - cf.mean())
    assert np.allclose(got, want, rtol=0, atol=1e-6)
    for cut in (0, 2, 5):
        x = hip.circ_corr_t(c[:20].astype(np.float32), nzp, cut).astype(np.float64)
        assert x[0] == 1.0 and np.allclose(x, x[(-k) % nzp], atol=1e-7)
        S = np.fft.rfft(x).real
        assert S.min() > -1e-5 and np.all(np.abs(S[: cut + 1]) < 1e-5)
        y = hip.circ_corr_t(x[: nzp // 2 + 1].astype(np.float32), nzp, cut)
        assert np.allclose(y, x, atol=2e-6)
    # a single lag (white) with the DC cut is -1/(nzp-1) off lag 0, still defined
    w = hip.circ_corr_t(np.array([1.0], np.float32), nzp, 0)
    assert np.allclose(w[1:], -1.0 / (nzp - 1), atol=1e-7)
    # every bin cut: undefined, as the reference aborts
    from crava_amd.hip import CravaError
    with pytest.raises(CravaError):
        hip.circ_corr_t(c[:20].astype(np.float32), nzp, nzp // 2)


def test_error_shape_operator_spectrum():
    """em2 = conj(E3)/|e3| with e3 = central(backward(forward(w))): the
    transform of the composed cyclic differences is
    i sin(t) (1 - e^{-it}) (e^{it} - 1) W = -4 i sin(t) sin^2(t/2) W, t = 2 pi k / nzp,
    and |e3| is its Parseval norm (wavelet.cpp:790-808, crava.cpp:1288-1323)."""
    for s in (model.config0(), model.config_phase(12, 10, 64, 3)):
        t = model.kz_tables(s)
        th = 2 * np.pi * np.arange(s.nzp) / s.nzp
        for l in range(s.n):
            w = t.wavelets[l].astype(np.float64)
            W = np.fft.fft(w)
            E3 = -4j * np.sin(th) * np.sin(th / 2) ** 2 * W
            n3 = math.sqrt(float(np.sum(np.abs(E3) ** 2)) / s.nzp)
            got = t.em2[:, l]
            assert np.allclose(got * n3, np.conj(E3), rtol=0, atol=2e-5 * np.abs(E3).max())
            assert np.allclose(np.abs(got), np.abs(E3) / n3, rtol=0, atol=2e-5 * np.abs(E3).max() / n3)


def test_data_variance_zeros_and_padding_excluded():
    """computeDataVariance (crava.cpp:407-449): mean of d^2 over the un-padded
    box, exact zeros excluded, float squares summed in double."""
    rng = np.random.default_rng(31)
    s = model.Survey(40, 24, 25, nx=37, ny=22, nz=19)
    d = np.zeros((s.nzp, s.nyp, s.rnxp), np.float32)
    d[:, :, : s.nxp] = rng.standard_normal((s.nzp, s.nyp, s.nxp)) * 2
    box = d[: s.nz, : s.ny, : s.nx]
    box[rng.random(box.shape) < 0.3] = 0.0
    d[:, :, s.nxp:] = 50.0
    v = box[box != 0].astype(np.float64)
    assert model.data_variance(s, d) == pytest.approx(float(np.float32(np.mean(v ** 2))), rel=1e-7)
    assert model.data_variance(s, np.zeros_like(d)) == 0.0
