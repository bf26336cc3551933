"""ORACLE -- test infrastructure, not product code.

Python side of the CPU oracle: ctypes access to liboracle.so (the plain-C
restatement in crava_oracle.c) and, where it was built, to
_ref/libcravaref.so (the reference's own FFTW 2.1.2 / lib_matr / RandomGen
compiled from its sources), plus FP64 transforms on CRAVA's padded layout
through numpy's pocketfft (an independent FFT implementation).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, and only to check or to time the reference algorithm.

Reference conventions restated here:
  * FFTGrid::fftInPlace / invFFTInPlace scaling by grid type
    (src/fftgrid.cpp:1783-1841): forward x 1/sqrt(N) unless COVARIANCE or
    OPERATOR; inverse x 1/N for those, 1/sqrt(N) otherwise; the half-complex
    inverse ignores the imaginary parts of the kx = 0 and kx = nxp/2 planes.
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SO = HERE / "_ref" / "libcravaref.so"

COVARIANCE, OPERATOR = 3, 5


def build() -> None:
    """Compile liboracle.so (and _ref when the reference tree is present)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", str(HERE), "-j8"], check=True)


_O = None
_R = None


def olib() -> C.CDLL:
    global _O
    if _O is None:
        if not ORACLE_SO.exists():
            build()
        o = C.CDLL(str(ORACLE_SO))
        fp = C.POINTER(C.c_float)
        dp = C.POINTER(C.c_double)
        o.or_rnorm_seq.argtypes = [C.POINTER(C.c_uint32), C.c_long, dp]
        o.or_fill_complex_noise.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint32), fp]
        o.or_fill_complex_noise_d.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint32), dp]
        o.or_noise_draws.restype = C.c_long
        o.or_noise_draws.argtypes = [C.c_int] * 3
        o.or_chol_flat_f.argtypes = [C.c_int, fp]
        o.or_chol_flat_d.argtypes = [C.c_int, dp]
        o.or_axeqb_flat_f.argtypes = [C.c_int, fp, fp, C.c_int]
        for nm, pt in (("or_posterior_f", fp), ("or_posterior_d", dp)):
            f = getattr(o, nm)
            f.restype = C.c_long
            f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(pt), C.POINTER(pt), pt, C.POINTER(pt),
                          pt, pt, pt, pt, fp, dp, dp, C.POINTER(C.c_int), C.c_float, C.c_int]
        for nm, pt in (("or_simulate_scale_f", fp), ("or_simulate_scale_d", dp)):
            f = getattr(o, nm)
            f.restype = C.c_long
            f.argtypes = [C.c_long, C.POINTER(pt), C.POINTER(pt)]
        o.or_find_closest_factorable.argtypes = [C.c_int]
        o.or_find_closest_factorable.restype = C.c_int
        _O = o
    return _O


def ref_available() -> bool:
    return REF_SO.exists()


def rlib() -> C.CDLL:
    global _R
    if _R is None:
        if not REF_SO.exists():
            raise RuntimeError("oracle/_ref/libcravaref.so not built (needs the reference tree)")
        r = C.CDLL(str(REF_SO))
        fp = C.POINTER(C.c_float)
        dp = C.POINTER(C.c_double)
        r.ref_rnorm_seq.restype = C.c_uint
        r.ref_rnorm_seq.argtypes = [C.c_uint, C.c_long, dp]
        r.ref_unif_seq.argtypes = [C.c_uint, C.c_long, dp]
        r.ref_fill_complex_noise.restype = C.c_uint
        r.ref_fill_complex_noise.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint, fp]
        r.ref_chol_flat.argtypes = [C.c_int, fp]
        r.ref_axeqb_flat.argtypes = [C.c_int, fp, fp, C.c_int]
        r.ref_posterior.restype = C.c_long
        r.ref_posterior.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(fp), C.POINTER(fp), fp,
                                    C.POINTER(fp), fp, fp, fp, fp, fp, dp, dp, C.POINTER(C.c_int), C.c_float, C.c_int]
        r.ref_simulate_scale.restype = C.c_long
        r.ref_simulate_scale.argtypes = [C.c_long, C.POINTER(fp), C.POINTER(fp)]
        r.ref_fft3d_r2c.argtypes = [C.c_int, C.c_int, C.c_int, fp, C.c_int]
        r.ref_fft3d_c2r.argtypes = [C.c_int, C.c_int, C.c_int, fp, C.c_int]
        _R = r
    return _R


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


# ------------------------------------------------------------------ RNG
def rnorm_seq(seed: int, n: int):
    """(normals, final state) of the restated RandomGen."""
    st = C.c_uint32(seed & 0xFFFFFFFF)
    out = np.empty(n, np.float64)
    olib().or_rnorm_seq(C.byref(st), n, _p(out, C.c_double))
    return out, st.value


def fill_complex_noise(nxp, nyp, nzp, seed: int, dtype=np.float32):
    cnxp = nxp // 2 + 1
    out = np.empty((nzp, nyp, cnxp, 2), dtype)
    st = C.c_uint32(seed & 0xFFFFFFFF)
    if dtype == np.float32:
        olib().or_fill_complex_noise(nxp, nyp, nzp, C.byref(st), _p(out, C.c_float))
    else:
        olib().or_fill_complex_noise_d(nxp, nyp, nzp, C.byref(st), _p(out, C.c_double))
    return out[..., 0] + 1j * out[..., 1], st.value


def noise_draws(nxp, nyp, nzp) -> int:
    return int(olib().or_noise_draws(nxp, nyp, nzp))


def chol(a: np.ndarray, dtype=np.float32):
    """lib_matrCholCpx restated on an n x n complex matrix; (flag, L)."""
    n = a.shape[0]
    buf = np.ascontiguousarray(np.stack([a.real, a.imag], -1), dtype=dtype)
    if dtype == np.float32:
        flag = olib().or_chol_flat_f(n, _p(buf, C.c_float))
    else:
        flag = olib().or_chol_flat_d(n, _p(buf, C.c_double))
    return flag, buf[..., 0] + 1j * buf[..., 1]


def find_closest_factorable(n: int) -> int:
    return int(olib().or_find_closest_factorable(n))


# ------------------------------------------------------------ transforms
def fft_r2c(real: np.ndarray, nxp: int, cubetype: int) -> np.ndarray:
    """FP64 FFTGrid::fftInPlace on a (nzp, nyp, rnxp) real grid."""
    x = np.asarray(real, np.float64)[:, :, :nxp]
    N = x.size
    s = 1.0 if cubetype in (COVARIANCE, OPERATOR) else 1.0 / math.sqrt(N)
    return np.fft.rfftn(x * s)


def fft_c2r(spec: np.ndarray, nxp: int, cubetype: int) -> np.ndarray:
    """FP64 FFTGrid::invFFTInPlace; returns (nzp, nyp, rnxp) with zero pads."""
    nzp, nyp, cnxp = spec.shape
    N = nxp * nyp * nzp
    sp = np.array(spec, dtype=np.complex128)
    # complex inverse over (z, y) first, then the half-complex inverse along
    # x, which keeps only the real parts of the kx = 0 and kx = nxp/2 bins
    t = np.fft.ifft(np.fft.ifft(sp, axis=0), axis=1) * (nzp * nyp)
    t[:, :, 0] = t[:, :, 0].real
    if nxp % 2 == 0:
        t[:, :, nxp // 2] = t[:, :, nxp // 2].real
    x = np.fft.irfft(t, n=nxp, axis=2) * nxp
    s = 1.0 / N if cubetype in (COVARIANCE, OPERATOR) else 1.0 / math.sqrt(N)
    out = np.zeros((nzp, nyp, 2 * cnxp))
    out[:, :, :nxp] = x * s
    return out


def ref_fft_r2c(real: np.ndarray, nxp: int, cubetype: int) -> np.ndarray:
    """The reference's own FFTW 2.1.2 float transform (needs _ref)."""
    nzp, nyp, rnxp = real.shape
    buf = np.ascontiguousarray(real, np.float32).copy()
    rlib().ref_fft3d_r2c(nxp, nyp, nzp, _p(buf, C.c_float), cubetype)
    v = buf.reshape(nzp, nyp, rnxp // 2, 2)
    return v[..., 0] + 1j * v[..., 1]


def ref_fft_c2r(spec: np.ndarray, nxp: int, cubetype: int) -> np.ndarray:
    nzp, nyp, cnxp = spec.shape
    buf = np.ascontiguousarray(np.stack([spec.real, spec.imag], -1), np.float32).copy()
    rlib().ref_fft3d_c2r(nxp, nyp, nzp, _p(buf, C.c_float), cubetype)
    return buf.reshape(nzp, nyp, 2 * cnxp)


# ------------------------------------------------------------- posterior
def _inter(a, dtype):
    return np.ascontiguousarray(np.stack([a.real, a.imag], -1), dtype=dtype)


def posterior_loop(mean, data, errcorr, cov, *, kwd, em1, em2, em3, A, theta, sigma0, invert, wnc,
                   dtype=np.float32, nthreads=1, backend="oracle"):
    """Run the per-wavenumber loop on half spectra (complex arrays of shape
    (nzp, nyp, cnxp)).  Returns (mean, data(residual), cov, nfail) as complex.
    backend: "oracle" (restated lib_matr) or "reference" (compiled lib_matr)."""
    nzp, nyp, cnxp = mean[0].shape
    nxp = 2 * (cnxp - 1)
    n = len(data)
    ct = C.c_float if dtype == np.float32 else C.c_double
    pt = C.POINTER(ct)
    M = [_inter(m, dtype) for m in mean]
    D = [_inter(d, dtype) for d in data]
    E = _inter(errcorr, dtype)
    V = [_inter(c, dtype) for c in cov]
    tabs = [_inter(t, dtype) for t in (kwd, em1, em2, em3)]
    A_ = np.ascontiguousarray(A, np.float32)
    th = np.ascontiguousarray(theta, np.float64)
    s0 = np.ascontiguousarray(sigma0, np.float64)
    inv = np.ascontiguousarray(invert, np.int32)
    mp = (pt * 3)(*[_p(m, ct) for m in M])
    dp_ = (pt * n)(*[_p(d, ct) for d in D])
    vp = (pt * 6)(*[_p(v, ct) for v in V])
    if backend == "reference":
        if dtype != np.float32:
            raise ValueError("the reference loop is float only")
        fn = rlib().ref_posterior
    else:
        fn = olib().or_posterior_f if dtype == np.float32 else olib().or_posterior_d
    nfail = fn(nxp, nyp, nzp, n, mp, dp_, _p(E, ct), vp, *[_p(t, ct) for t in tabs], _p(A_, C.c_float),
               _p(th, C.c_double), _p(s0, C.c_double), _p(inv, C.c_int), float(np.float32(wnc)), nthreads)
    cz = lambda a: a[..., 0] + 1j * a[..., 1]
    return [cz(m) for m in M], [cz(d) for d in D], [cz(v) for v in V], int(nfail)


def simulate_scale(cov, seeds, dtype=np.float32, backend="oracle"):
    """Crava::simulate's per-wavenumber Cholesky scaling on half spectra."""
    ct = C.c_float if dtype == np.float32 else C.c_double
    pt = C.POINTER(ct)
    V = [_inter(c, dtype) for c in cov]
    S = [_inter(s, dtype) for s in seeds]
    vp = (pt * 6)(*[_p(v, ct) for v in V])
    sp = (pt * 3)(*[_p(s, ct) for s in S])
    csize = V[0].size // 2
    if backend == "reference":
        nfail = rlib().ref_simulate_scale(csize, vp, sp)
    else:
        fn = olib().or_simulate_scale_f if dtype == np.float32 else olib().or_simulate_scale_d
        nfail = fn(csize, vp, sp)
    cz = lambda a: a[..., 0] + 1j * a[..., 1]
    return [cz(s) for s in S], int(nfail)


def gaussian_update(Sm, K, Se, mu, d):
    """Closed-form Gaussian posterior (FP64): mu + Sm K^H (K Sm K^H + Se)^-1 (d - K mu),
    Sm - Sm K^H (K Sm K^H + Se)^-1 K Sm; residual d - K mu_post."""
    M = K @ Sm @ K.conj().T + Se
    G = np.linalg.solve(M, K @ Sm)            # M^-1 K Sm
    mp = mu + G.conj().T @ (d - K @ mu)
    Sp = Sm - (K @ Sm).conj().T @ G
    return mp, Sp, d - K @ mp


# ----------------------------------------------- counter-based noise (FP64)
M32 = 0xFFFFFFFF


def philox10(c0, c1, c2, c3, k0, k1):
    """Philox-4x32-10 (Salmon et al., SC 2011) on numpy uint64 arrays."""
    c0, c1, c2, c3 = (np.asarray(v, np.uint64) & M32 for v in (c0, c1, c2, c3))
    k0 = np.uint64(k0 & M32)
    k1 = np.uint64(k1 & M32)
    for _ in range(10):
        p0 = np.uint64(0xD2511F53) * c0
        p1 = np.uint64(0xCD9E8D57) * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & np.uint64(M32)
        hi1, lo1 = p1 >> np.uint64(32), p1 & np.uint64(M32)
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
        k0 = (k0 + np.uint64(0x9E3779B9)) & np.uint64(M32)
        k1 = (k1 + np.uint64(0xBB67AE85)) & np.uint64(M32)
    return c0, c1, c2, c3


def counter_noise(nxp, nyp, nzp, seed, realization, grid):
    """Counter-stream simulation noise of one grid (0..2) laid out with the
    reference Hermitian rules (FFTGrid::fillInComplexNoise,
    src/fftgrid.cpp:1191-1240)."""
    cnxp = nxp // 2 + 1
    n = cnxp * nyp * nzp
    i = np.arange(n, dtype=np.int64)
    x = i % cnxp
    jk = i // cnxp
    j = jk % nyp
    k = jk // nyp
    jcc = np.where(j == 0, 0, nyp - j)
    kcc = np.where(k == 0, 0, nzp - k)
    jkcc = jcc + kcc * nyp
    special = (x == 0) | ((x == cnxp - 1) & (i % 2 == 1))
    src = np.where(special & (jkcc < jk), jkcc * cnxp + x, i)

    def pair(idx):
        # one call per element; six 21-bit fields: the top 21 bits of each
        # word, then low0 2^10 + low1 / 2 and low2 2^10 + low3 / 2; grid g
        # takes fields 2g, 2g + 1 as the uniforms (2 f + 1) 2^-22
        hi = np.uint64(realization >> 32) & np.uint64(0xFFFFFF)
        c = philox10(idx.astype(np.uint64) & np.uint64(M32), idx.astype(np.uint64) >> np.uint64(32),
                     np.full(idx.shape, realization & M32, np.uint64), np.full(idx.shape, hi, np.uint64),
                     seed & M32, (seed >> 32) & M32)
        lo = [w & np.uint64(0x7FF) for w in c]
        f = [w >> np.uint64(11) for w in c]
        f.append((lo[0] << np.uint64(10)) | (lo[1] >> np.uint64(1)))
        f.append((lo[2] << np.uint64(10)) | (lo[3] >> np.uint64(1)))
        fa, fb = f[2 * grid], f[2 * grid + 1]
        u1 = (fa * np.uint64(2) + np.uint64(1)).astype(np.float64) / 4194304.0
        u2 = (fb * np.uint64(2) + np.uint64(1)).astype(np.float64) / 4194304.0
        r = np.sqrt(-2.0 * np.log(u1))
        return r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)

    n0, n1 = pair(src)
    sd = float(np.float32(1.0 / math.sqrt(2.0)))
    re = (sd * n0).astype(np.float32).astype(np.float64)
    im = (sd * n1).astype(np.float32).astype(np.float64)
    selfc = special & (jkcc == jk)
    re = np.where(selfc, n0.astype(np.float32).astype(np.float64), re)
    im = np.where(selfc, 0.0, im)
    copy = special & (jkcc < jk)
    im = np.where(copy, -im, im)
    return (re + 1j * im).reshape(nzp, nyp, cnxp)


# ---------------------------------------------------------------- rplib chain
class _RockParams(C.Structure):
    _fields_ = [("kq", C.c_double), ("gq", C.c_double), ("rhoq", C.c_double), ("kc", C.c_double),
                ("gc", C.c_double), ("rhoc", C.c_double), ("temp_c", C.c_double), ("salinity", C.c_double),
                ("gas_gravity", C.c_double), ("gas", C.c_int)]


def rock_forward(x, params):
    """or_rock_forward_batch (rplib_oracle.c): (n, 5) inputs -> (n, 3) ln vp,
    ln vs, ln rho and (n,) status; params use the crv_rock_params names."""
    o = olib()
    x = np.ascontiguousarray(x, np.float64)
    n = x.shape[0]
    out = np.zeros((n, 3), np.float64)
    st = np.zeros(n, np.int32)
    p = _RockParams(params["k_quartz"], params["g_quartz"], params["rho_quartz"], params["k_clay"],
                    params["g_clay"], params["rho_clay"], params["temperature"], params["salinity"],
                    params["gas_gravity"], 0)
    dp = C.POINTER(C.c_double)
    o.or_rock_forward_batch.argtypes = [C.POINTER(_RockParams), C.c_long, dp, dp, C.POINTER(C.c_int)]
    o.or_rock_forward_batch(C.byref(p), n, x.ctypes.data_as(dp), out.ctypes.data_as(dp),
                            st.ctypes.data_as(C.POINTER(C.c_int)))
    return out, st


# ---- prior and error correlation grids of the general prior ----------------

def _cycle(n):
    """The offsets the fills give index i of an axis of n: i, or i - n when
    that is nearer zero (ties to i)."""
    i = np.arange(n)
    return np.where(i < -(i - n), i, i - n)


def param_corr_grids(corr_xy, circ_corr_t, grad_i, grad_j, sigma0, nxp, nyp, nzp, rnxp):
    """SeismicParametersHolder::initializeCorrelations (reference
    src/seismicparametersholder.cpp:197-216): FFTGrid::fillInParamCorr
    (src/fftgrid.cpp:1072-1117) then multiplyByScalar(float(Sigma0)) over
    every element, in numpy float32 with the reference's operation order.
    Six (nzp, nyp, rnxp) float32 grids (aa bb rr ab ar br)."""
    f = np.float32
    ci = _cycle(nxp).astype(np.float32)[None, None, :]
    cj = _cycle(nyp).astype(np.float32)[None, :, None]
    k = np.arange(nzp).astype(np.float32)[:, None, None]
    subk = (k + ci * f(grad_i)).astype(np.float32)
    subk = (subk + cj * f(grad_j)).astype(np.float32)
    base = np.floor(subk).astype(np.int64)
    frac = (subk - base.astype(np.float32)).astype(np.float32)
    base = base % nzp
    ct = np.asarray(circ_corr_t, np.float32)
    nxt = np.where(base != nzp - 1, base + 1, 0)
    v = ((f(1) - frac) * ct[base]).astype(np.float32)
    v = (v + frac * ct[nxt]).astype(np.float32)
    v = (v * np.asarray(corr_xy, np.float64).astype(np.float32)[None]).astype(np.float32)
    full = np.full((nzp, nyp, rnxp), f(-99999.0), np.float32)
    full[:, :, :nxp] = v
    S = np.asarray(sigma0, np.float64)
    s6 = S.reshape(-1) if S.size == 6 else [S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2]]
    return [(full * f(s)).astype(np.float32) for s in s6]


def err_corr_grid(corr_xy, grad_i, grad_j, nxp, nyp, nzp, rnxp):
    """FFTGrid::fillInErrCorr (reference src/fftgrid.cpp:1018-1069) in numpy
    float32: within one sample of plane 0 (or of plane nzp, cyclically) the
    weight 1 - (shift minus its truncation toward zero), with the
    reference's int() truncation: a shift f in (-1, 0) keeps fraction f and
    weighs 1 - f = 1 + |f|; (nzp, nyp, rnxp) float32."""
    f = np.float32
    ci = _cycle(nxp).astype(np.float32)[None, None, :]
    cj = _cycle(nyp).astype(np.float32)[None, :, None]
    k = np.arange(nzp).astype(np.float32)[:, None, None]
    subk = (k + ci * f(grad_i)).astype(np.float32)
    subk = (subk + cj * f(grad_j)).astype(np.float32)
    near = (np.abs(subk) < f(1)) | (np.abs((subk - f(nzp)).astype(np.float32)) < f(1))
    base = np.trunc(subk).astype(np.int64)
    frac = (subk - base.astype(np.float32)).astype(np.float32)
    base = (base + 1) % nzp - 1        # the reference's cyclic moves into [-1, 1)
    frac = np.where(base < 0, (f(1) - frac).astype(np.float32), frac)
    v = ((f(1) - frac) * np.asarray(corr_xy, np.float64).astype(np.float32)[None]).astype(np.float32)
    v = np.where(near, v, f(0)).astype(np.float32)
    full = np.full((nzp, nyp, rnxp), f(-99999.0), np.float32)
    full[:, :, :nxp] = v
    return full
