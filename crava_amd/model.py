"""Model set-up around the hot path: the inputs CRAVA builds before it calls
Crava::computePostMeanResidAndFFTCov and Crava::simulate.

Everything here runs once per inversion on the host and is small (tables of
nzp x n, 2D lateral spectra) except the synthetic input generator, which
stands in for reading SEG-Y cubes.

Reference behaviour restated (file:line in the reference tree):
  * Ricker wavelet (src/wavelet.cpp:1354-1358), resampled to dt in FFT order
    (src/wavelet.cpp:416-463; the kriging interpolation of getWaveletValue is
    exact at the 1 ms nodes the configurations sample, so the resampled value
    is the Ricker value at t = k dt, and 0 beyond the +-119 ms support of the
    wavelet, src/wavelet.cpp:140-160);
  * difference operators (src/wavelet1D.cpp:367-478), norms
    (src/wavelet.cpp:790-808), getCAmp (309-324);
  * per-kz multipliers of Crava::fillErrMultVectors for constant thickness
    (src/crava.cpp:1336-1364, fillkW 1471-1479, fillkWNorm 1481-1489);
  * PP Aki-Richards rows (src/modelavodynamic.cpp:877-902);
  * data variance, error variance and errThetaCov
    (src/crava.cpp:407-487), with the generalised-exponential angular
    correlation (src/vario.cpp:81-94);
  * spherical lateral correlation (src/vario.cpp:60-73) on cyclic lags
    (src/modelgeneral.cpp:4194-4232);
  * circular temporal correlation made positive definite
    (src/seismicparametersholder.cpp:430-498), lowIntCut
    (src/modelgeneral.cpp:4081);
  * prior covariance grids (src/fftgrid.cpp:1072-1117 with zero gradients)
    and the error-correlation grid (src/fftgrid.cpp:1018-1069);
  * padding sizes (src/modelgeneral.cpp:616-622, 1426-1516).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

RMISSING = -99999.0


# ----------------------------------------------------------------------------
@dataclass
class Survey:
    """A synthetic survey: grid, angles, wavelets, prior and noise levels."""
    nxp: int
    nyp: int
    nzp: int
    dx: float = 25.0
    dy: float = 25.0
    dt: float = 4.0                      # ms
    angles_deg: tuple = (0.0, 15.0, 30.0)
    peak_hz: tuple = (30.0, 30.0, 30.0)
    sn: tuple = (5.0, 5.0, 5.0)
    vsvp: float = 0.5
    mu: tuple = (math.log(3000.0), math.log(1500.0), math.log(2300.0))
    stds: tuple = (0.10, 0.10, 0.05)
    corr: tuple = (0.7, 0.5, 0.5)        # (vp,vs), (vp,rho), (vs,rho)
    lateral_range: float = 1000.0        # m, spherical
    temporal_range: float = 20.0         # ms, Gaussian (exp(-3 (t/r)^2))
    low_cut: float = 5.0
    high_cut: float = 55.0
    wnc: float = 0.1
    angular_range_deg: float = 10.0      # GenExp(power 1) default
    seed_truth: int = 1
    seed_noise: int = 2
    seed_sim: int = 0
    phase_deg: tuple = ()                # per-angle constant phase rotation (empty: zero phase)
    err_phase_deg: tuple = ()            # phase of the error-norm wavelets (empty: same as the forward ones)
    nx: int = 0
    ny: int = 0
    nz: int = 0

    def __post_init__(self):
        if not self.nx:
            self.nx = self.nxp
        if not self.ny:
            self.ny = self.nyp
        if not self.nz:
            self.nz = self.nzp

    @property
    def n(self) -> int:
        return len(self.angles_deg)

    @property
    def cnxp(self) -> int:
        return self.nxp // 2 + 1

    @property
    def rnxp(self) -> int:
        return 2 * self.cnxp

    @property
    def voxels(self) -> int:
        return self.nxp * self.nyp * self.nzp

    @property
    def lz(self) -> float:
        return self.nz * self.dt

    def sigma0(self) -> np.ndarray:
        s = np.array(self.stds, dtype=np.float64)
        c = np.eye(3)
        c[0, 1] = c[1, 0] = self.corr[0]
        c[0, 2] = c[2, 0] = self.corr[1]
        c[1, 2] = c[2, 1] = self.corr[2]
        return c * np.outer(s, s)


def config0() -> Survey:
    """3 angles 0/15/30, padded 64x64x128, 30 Hz Ricker, S/N 5."""
    return Survey(64, 64, 128)


def config1() -> Survey:
    """5 angles 5..33, padded 450x450x320, Ricker 35->28 Hz, S/N 4."""
    ang = (5.0, 12.0, 19.0, 26.0, 33.0)
    peaks = tuple(35.0 - 7.0 * i / 4.0 for i in range(5))
    return Survey(450, 450, 320, angles_deg=ang, peak_hz=peaks, sn=(4.0,) * 5)


def config3_realizations() -> Survey:
    """Batched simulation: the config-1 survey setup on a 256x256x200 padded
    grid (64 realizations, counter-stream seeds 0..63)."""
    ang = (5.0, 12.0, 19.0, 26.0, 33.0)
    peaks = tuple(35.0 - 7.0 * i / 4.0 for i in range(5))
    return Survey(256, 256, 200, angles_deg=ang, peak_hz=peaks, sn=(4.0,) * 5)


def config3_volumes(count: int = 12, seed: int = 3) -> list:
    """Uneven volumes: nx, ny from {64, 96, 128, 160, 192, 240, 320}, nz from
    {100, 128, 160, 200, 256, 320, 400}, padded with the reference's default
    pad fractions, prior stds scaled by U(0.7, 1.3) around config 0."""
    rng = np.random.default_rng(seed)
    xs = (64, 96, 128, 160, 192, 240, 320)
    zs = (100, 128, 160, 200, 256, 320, 400)
    base = config0()
    out = []
    for v in range(count):
        nx, ny, nz = int(rng.choice(xs)), int(rng.choice(xs)), int(rng.choice(zs))
        nxp, nyp, nzp = estimate_padding(nx, ny, nz, base.dx, base.dy, base.dt, base.lateral_range)
        f = rng.uniform(0.7, 1.3, 3)
        stds = tuple(float(a * b) for a, b in zip(base.stds, f))
        out.append(Survey(nxp, nyp, nzp, nx=nx, ny=ny, nz=nz, stds=stds, seed_truth=100 + v, seed_noise=200 + v))
    return out


def config_small(nxp=12, nyp=10, nzp=16, n=3) -> Survey:
    ang = tuple(np.linspace(0.0, 30.0, n))
    return Survey(nxp, nyp, nzp, angles_deg=ang, peak_hz=(30.0,) * n, sn=(5.0,) * n)


PHASES_DEG = (0.0, 25.0, -40.0, 60.0, 90.0, -75.0, 15.0, -120.0)


ERR_PHASES_DEG = (10.0, -30.0, 45.0, 20.0, -60.0, 80.0, -5.0, 35.0)


def config_phase(nxp=20, nyp=18, nzp=30, n=5, own_error_wavelets=False) -> Survey:
    """Wavelets that differ between angles in peak frequency (35 -> 28 Hz
    from near to far, as the production survey) and in phase (each Ricker
    rotated by its own constant phase, as wavelets estimated from wells
    are): the forward operators and the error covariance of different
    angles then differ in phase (Im Sigma_e != 0 off the diagonal).  With one
    wavelet per angle the phases cancel from the posterior covariance
    (K = conj(W) D1 A and e1 = conj(W)/|W| share W), which stays real; with
    own_error_wavelets the error terms use wavelets of other phases, as
    Wavelet3D's error norm uses the average wavelet instead of the local one
    (reference src/wavelet3D.cpp:449-460), and the posterior covariance is
    complex as well."""
    ang = tuple(np.linspace(5.0, 33.0, n))
    peaks = tuple(35.0 - 7.0 * i / max(n - 1, 1) for i in range(n))
    return Survey(nxp, nyp, nzp, angles_deg=ang, peak_hz=peaks, sn=(4.0,) * n, phase_deg=PHASES_DEG[:n],
                  err_phase_deg=ERR_PHASES_DEG[:n] if own_error_wavelets else ())


# ----------------------------------------------------------------------------
# padding (src/fftgrid.cpp:2857-2885, src/modelgeneral.cpp:616-622)
def find_closest_factorable(leastint: int) -> int:
    f32 = np.float32
    lg = float(np.log(f32(leastint)))
    m2 = int(math.ceil(lg / float(np.log(f32(2.0)))))
    m3 = int(math.ceil(lg / float(np.log(f32(3.0)))))
    m5 = int(math.ceil(lg / float(np.log(f32(5.0)))))
    m7 = int(math.ceil(lg / float(np.log(f32(7.0)))))
    best = 2 ** m2
    for i in range(m2 + 1):
        for j in range(m3 + 1):
            for k in range(m5 + 1):
                for l in range(m7 + 1):
                    f = 2 ** i * 3 ** j * 5 ** k * 7 ** l
                    if leastint <= f < best and f <= 2147483647:
                        best = f
    return best


def padding_size(n: int, pad_fac: float) -> int:
    least = int(math.ceil(n * (1.0 + float(pad_fac))))
    return find_closest_factorable(least)


def estimate_padding(nx: int, ny: int, nz: int, dx: float, dy: float, dt: float,
                     lateral_range: float = 1000.0, wavelet_length_ms: float = 200.0):
    """ModelGeneral::estimateXYPaddingSizes / estimateZPaddingSize with the
    default estimated pad fractions (src/modelgeneral.cpp:1426-1516)."""
    lx, ly, lz = nx * dx, ny * dy, nz * dt
    xpad = max(0.5 * lateral_range, dx)
    ypad = max(0.5 * lateral_range, dy)
    fx, fy = min(1.0, xpad / lx), min(1.0, ypad / ly)
    fz = min(1.0, wavelet_length_ms / lz)
    return padding_size(nx, fx), padding_size(ny, fy), padding_size(nz, fz)


# ----------------------------------------------------------------------------
# wavelets and per-kz operators
def ricker(t_ms, peak_hz):
    c = math.pi * math.pi * peak_hz * peak_hz * np.asarray(t_ms, dtype=np.float64) ** 2 * 1e-6
    return (1.0 - 2.0 * c) * np.exp(-c)


def wavelet_fft_order(nzp: int, dt: float, peak_hz: float, phase_deg: float = 0.0) -> np.ndarray:
    """Resampled Ricker in FFT order (float32), zero outside +-119 ms,
    optionally rotated by a constant phase: w cos(p) - H[w] sin(p) with H the
    discrete Hilbert transform (multiplier -i sgn(f), 0 at DC and Nyquist), so
    the rotated wavelet is real and its spectrum is W(f) e^{i p sgn(f)} away
    from DC and Nyquist."""
    k = np.arange(nzp)
    t = np.where(k < nzp // 2 + 1, dt * k, dt * (k - nzp)).astype(np.float32).astype(np.float64)
    w = ricker(t, peak_hz)
    w[np.abs(t) > 119.0] = 0.0
    if phase_deg:
        p = math.radians(phase_deg)
        f = np.fft.fftfreq(nzp)
        mult = math.cos(p) + 1j * math.sin(p) * np.sign(f)
        if nzp % 2 == 0:
            mult[nzp // 2] = math.cos(p)
        w = np.fft.ifft(np.fft.fft(w) * mult).real
    return w.astype(np.float32)


def norm_time(w: np.ndarray) -> float:
    return float(np.float32(math.sqrt(float(np.sum(w.astype(np.float64) ** 2)))))


def diff_forward(w):  # r[i] = w[i+1] - w[i] (cyclic)
    return (np.roll(w, -1) - w).astype(np.float32)


def diff_backward(w):  # r[i] = w[i] - w[i-1] (cyclic)
    return (w - np.roll(w, 1)).astype(np.float32)


def diff_central(w):  # r[i] = 0.5 (w[i+1] - w[i-1]) (cyclic)
    return (0.5 * (np.roll(w, -1).astype(np.float64) - np.roll(w, 1))).astype(np.float32)


def full_spectrum(w: np.ndarray) -> np.ndarray:
    """Unnormalised DFT (sign -1) of a real operator, all nzp bins; the upper
    half equals getCAmp's conjugate of the mirrored bin."""
    return np.fft.fft(w.astype(np.float64))


def d1_operator(nzp: int) -> np.ndarray:
    r = np.zeros(nzp, dtype=np.float32)
    r[0] = -1.0
    r[nzp - 1] = 1.0
    return r


def reflection_matrix(angles_deg, vsvp: float) -> np.ndarray:
    """PP rows (1+tan^2)/2, -4 g^2 sin^2, (1-4 g^2 sin^2)/2 as float32."""
    A = np.zeros((len(angles_deg), 3), dtype=np.float32)
    g2 = vsvp * vsvp
    for i, a in enumerate(angles_deg):
        th = math.radians(a)
        s2 = math.sin(th) ** 2
        t2 = math.tan(th) ** 2
        A[i] = [(1.0 + t2) / 2.0, -4.0 * g2 * s2, (1.0 - 4.0 * g2 * s2) / 2.0]
    return A


@dataclass
class KzTables:
    kwd: np.ndarray   # (nzp, n) complex128: kD * conj(W)
    em1: np.ndarray   # (nzp, n) complex128
    em2: np.ndarray
    em3: np.ndarray
    wavelets: list = field(default_factory=list)   # time-domain wavelets (float32)

    def as_interleaved(self):
        f = lambda a: np.ascontiguousarray(np.stack([a.real, a.imag], axis=-1).astype(np.float64))
        return f(self.kwd), f(self.em1), f(self.em2), f(self.em3)


def kz_tables(s: Survey) -> KzTables:
    nzp, n = s.nzp, s.n
    D1 = full_spectrum(d1_operator(nzp))
    kwd = np.zeros((nzp, n), complex)
    em1 = np.zeros((nzp, n), complex)
    em2 = np.zeros((nzp, n), complex)
    em3 = np.ones((nzp, n), complex)
    wl = []
    for l in range(n):
        w = wavelet_fft_order(nzp, s.dt, s.peak_hz[l], s.phase_deg[l] if s.phase_deg else 0.0)
        wl.append(w)
        W = full_spectrum(w)
        kwd[:, l] = D1 * np.conj(W)
        # the error norm's wavelet (createWavelet1DForErrorNorm): the same one,
        # or for own error wavelets one of another phase
        we = w if not s.err_phase_deg else wavelet_fft_order(nzp, s.dt, s.peak_hz[l], s.err_phase_deg[l])
        We = full_spectrum(we)
        nw = norm_time(we)
        e3 = diff_central(diff_backward(diff_forward(we)))
        E3 = full_spectrum(e3)
        n3 = norm_time(e3)
        em1[:, l] = np.conj(We) / nw
        em2[:, l] = np.conj(E3) / n3
    return KzTables(kwd, em1, em2, em3, wl)


def invert_mask(s: Survey) -> np.ndarray:
    """crava.cpp:1104-1105 in the reference's mixed float/double arithmetic."""
    nzp = s.nzp
    k = np.arange(nzp)
    base = float(np.float32(s.nz * np.float32(1000.0))) / (s.lz * nzp)
    fr = (base * np.minimum(k, nzp - k).astype(np.float64)).astype(np.float32)
    lo = np.float64(np.float32(s.low_cut)) * 1.0
    return ((fr.astype(np.float64) > lo) & (fr < np.float32(s.high_cut))).astype(np.int32)


# ----------------------------------------------------------------------------
# correlations and prior / error spectra
def spherical_corr(dx, dy, r):
    d = np.sqrt((np.asarray(dx, np.float64) / r) ** 2 + (np.asarray(dy, np.float64) / r) ** 2)
    return np.where(d >= 1.0, 0.0, 1.0 + d ** 3 / 2.0 - 1.5 * d).astype(np.float32)


def genexp_corr(d, r, power=1.0):
    return np.float32(math.exp(-3.0 * (abs(d) / r) ** power))


def corr_xy(s: Survey) -> np.ndarray:
    """(nyp, nxp) lateral correlation on cyclic lags."""
    i = np.arange(s.nxp)
    j = np.arange(s.nyp)
    ri = np.where(i < s.nxp // 2 + 1, i, i - s.nxp).astype(np.float32) * np.float32(s.dx)
    rj = np.where(j < s.nyp // 2 + 1, j, j - s.nyp).astype(np.float32) * np.float32(s.dy)
    return spherical_corr(ri[None, :], rj[:, None], s.lateral_range)


def circ_corr_t(s: Survey) -> np.ndarray:
    """Circular temporal correlation after makeCircCorrTPosDef (nzp,)."""
    nzp = s.nzp
    ncorr = nzp // 2 + 1 if nzp % 2 == 0 else nzp // 2
    lags = np.arange(ncorr)
    prior = np.exp(-3.0 * (lags * s.dt / s.temporal_range) ** 2).astype(np.float32)
    k = np.arange(nzp)
    refk = np.where(k < nzp // 2 + 1, k, nzp - k)
    c = np.where(refk < ncorr, prior[np.minimum(refk, ncorr - 1)], 0.0).astype(np.float64)
    low_int_cut = int(math.floor(np.float32(s.low_cut) * (nzp * 0.001 * np.float32(s.dt))))
    F = np.fft.rfft(c)
    a = np.abs(F)
    a[: low_int_cut + 1] = 0.0
    c2 = np.fft.irfft(a, nzp)
    if c2[0] <= 1e-5:
        raise ValueError("circular temporal correlation undefined: too few layers")
    return (c2 / c2[0]).astype(np.float32)


def prior_spectra(s: Survey):
    """Separable prior: cxy (nyp, cnxp) = 2D DFT of the lateral correlation,
    ct (nzp,) = DFT of the circular temporal correlation (both real)."""
    cxy = np.fft.rfft2(corr_xy(s).astype(np.float64)).real
    ct = np.fft.fft(circ_corr_t(s).astype(np.float64)).real
    return np.ascontiguousarray(cxy), np.ascontiguousarray(ct)


def prior_cov_grids(s: Survey, dtype=np.float32) -> list:
    """The six real COVARIANCE grids (aa bb rr ab ar br) in the padded layout
    (nzp, nyp, rnxp): Sigma0[a][b] * corrT(k) * corrXY(i, j), pad = RMISSING."""
    cxy = corr_xy(s).astype(np.float32)
    ct = circ_corr_t(s)
    S0 = s.sigma0()
    base = (ct[:, None, None] * cxy[None, :, :]).astype(np.float32)
    out = []
    for a, b in ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)):
        g = np.full((s.nzp, s.nyp, s.rnxp), RMISSING, dtype=np.float32)
        g[:, :, : s.nxp] = base * np.float32(S0[a, b])
        out.append(g.astype(dtype))
    return out


def err_corr_grid(s: Survey, dtype=np.float32) -> np.ndarray:
    """FFTGrid::fillInErrCorr with zero gradients: corrXY at k = 0, else 0."""
    g = np.zeros((s.nzp, s.nyp, s.rnxp), dtype=np.float32)
    g[:, :, s.nxp:] = RMISSING
    g[0, :, : s.nxp] = corr_xy(s)
    return g.astype(dtype)


def err_theta_cov(s: Survey, data_var) -> np.ndarray:
    """errThetaCov[i][j] = sqrt(ev_i) sqrt(ev_j) corr(theta_i - theta_j)."""
    ev = [np.float32(np.float32(dv) / np.float32(sn)) for dv, sn in zip(data_var, s.sn)]
    rng = math.radians(s.angular_range_deg)
    th = np.zeros((s.n, s.n))
    for i in range(s.n):
        for j in range(s.n):
            dth = np.float32(math.radians(s.angles_deg[i])) - np.float32(math.radians(s.angles_deg[j]))
            th[i, j] = float(np.float32(np.sqrt(ev[i]) * np.sqrt(ev[j]) * genexp_corr(float(dth), rng)))
    return th


def data_variance(s: Survey, data_real: np.ndarray) -> float:
    """Crava::computeDataVariance over the un-padded box, zeros excluded."""
    box = data_real[: s.nz, : s.ny, : s.nx].astype(np.float64)
    nz = box != 0.0
    cnt = int(nz.sum())
    return float(np.float32((box[nz] ** 2).sum() / cnt)) if cnt else 0.0


# ----------------------------------------------------------------------------
# synthetic inputs (stand-in for the SEG-Y cubes and the true model)
def _hermitian_noise(rng: np.random.Generator, s: Survey) -> np.ndarray:
    """Complex white noise of unit variance whose inverse real transform is a
    real field (rfftn of real white noise scaled to unit variance)."""
    x = rng.standard_normal((s.nzp, s.nyp, s.nxp))
    return np.fft.rfftn(x) / math.sqrt(s.voxels)


@dataclass
class Synthetic:
    mean: list        # 3 prior mean grids (nzp, nyp, rnxp) float64 (constant)
    data: list        # n seismic grids (nzp, nyp, rnxp) float64
    truth: list       # 3 true model grids
    data_var: list


def make_synthetic(s: Survey) -> Synthetic:
    """True model = mu + a draw from the prior; seismic = conj(W) D1 A m + white
    noise at the survey's S/N (CRAVA's S/N = (|signal|^2 + |noise|^2)/|noise|^2)."""
    rng = np.random.default_rng(s.seed_truth)
    cxy, ct = prior_spectra(s)
    sc = np.sqrt(np.abs(ct[:, None, None] * cxy[None, :, :]))
    L = np.linalg.cholesky(s.sigma0())
    z = [_hermitian_noise(rng, s) for _ in range(3)]
    N = s.voxels
    mhat = []
    for a in range(3):
        spec = sum(L[a, b] * z[b] for b in range(3)) * sc
        mhat.append(spec)  # PARAMETER-scaled spectrum of the anomaly
    truth = []
    for a in range(3):
        m = np.fft.irfftn(mhat[a], s=(s.nzp, s.nyp, s.nxp)) * math.sqrt(N) + s.mu[a]
        truth.append(m)
    A = reflection_matrix(s.angles_deg, s.vsvp).astype(np.float64)
    tab = kz_tables(s)
    rng_e = np.random.default_rng(s.seed_noise)
    data, dvar = [], []
    for l in range(s.n):
        shat = tab.kwd[:, l][:, None, None] * sum(A[l, m] * mhat[m] for m in range(3))
        sig = np.fft.irfftn(shat, s=(s.nzp, s.nyp, s.nxp)) * math.sqrt(N)
        sv = float(np.mean(sig ** 2))
        ev = sv / max(s.sn[l] - 1.0, 1e-6)
        d = sig + rng_e.standard_normal(sig.shape) * math.sqrt(ev)
        data.append(d)
    def pad(g):
        out = np.zeros((s.nzp, s.nyp, s.rnxp))
        out[:, :, : s.nxp] = g
        return out
    data_p = [pad(d) for d in data]
    dvar = [data_variance(s, d) for d in data_p]
    mean = [np.full((s.nzp, s.nyp, s.rnxp), s.mu[a]) for a in range(3)]
    for m in mean:
        m[:, :, s.nxp:] = 0.0
    return Synthetic(mean, data_p, [pad(t) for t in truth], dvar)


@dataclass
class PosteriorInputs:
    """Everything crv_posterior needs besides the grids."""
    A: np.ndarray
    theta: np.ndarray
    tables: KzTables
    invert: np.ndarray
    sigma0: np.ndarray
    cxy: np.ndarray
    ct: np.ndarray


def posterior_inputs(s: Survey, data_var) -> PosteriorInputs:
    cxy, ct = prior_spectra(s)
    return PosteriorInputs(reflection_matrix(s.angles_deg, s.vsvp), err_theta_cov(s, data_var),
                           kz_tables(s), invert_mask(s), s.sigma0(), cxy, ct)
