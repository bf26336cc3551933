"""Python view of the C ABI with the reference's names.

``FFTGrid`` mirrors the reference class for the hot path (src/fftgrid.h):
``fftInPlace`` / ``invFFTInPlace`` / ``add`` / ``multiplyByScalar`` /
``fillInComplexNoise`` / ``getIsTransformed``; ``computePostMeanResidAndFFTCov``
and ``simulate`` mirror the Crava members (src/crava.h:43-64, 154).  Errors
raise ``CravaError`` carrying the library's message; the reference asserts
instead (same conditions).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L


class CravaError(RuntimeError):
    pass


class Context:
    def __init__(self, device: int = 0):
        self.lib = L.lib()
        h = C.c_void_p()
        rc = self.lib.crv_ctx_create(device, C.byref(h))
        if rc != 0:
            raise CravaError(f"crv_ctx_create failed ({rc}): no usable GPU")
        self.h = h

    def check(self, rc: int, what: str = ""):
        if rc != 0:
            msg = self.lib.crv_last_error(self.h)
            raise CravaError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def synchronize(self):
        self.check(self.lib.crv_ctx_synchronize(self.h), "synchronize")

    def set_stream(self, stream_handle: int | None):
        """Run on an external HIP stream (e.g. torch.cuda.current_stream().cuda_stream)."""
        self.check(self.lib.crv_ctx_set_stream(self.h, stream_handle), "set_stream")

    def tune(self, key: int, value: int):
        self.check(self.lib.crv_ctx_tune(self.h, key, value), "tune")

    def profile(self, on: bool):
        self.check(self.lib.crv_ctx_profile(self.h, int(on)), "profile")

    def profile_report(self) -> dict:
        """{kernel: (launches, total_ms, min_ms, max_ms)} of the recorded launches."""
        buf = C.create_string_buffer(1 << 16)
        n = self.lib.crv_ctx_profile_report(self.h, buf, len(buf))
        if n < 0:
            self.check(n, "profile_report")
        out = {}
        for line in buf.value.decode().splitlines():
            name, cnt, tot, mn, mx = line.split()
            out[name] = (int(cnt), float(tot), float(mn), float(mx))
        return out

    def close(self):
        if self.h:
            self.lib.crv_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FFTGrid:
    DATA, PARAMETER, COVARIANCE, OPERATOR = L.CRV_DATA, L.CRV_PARAMETER, L.CRV_COVARIANCE, L.CRV_OPERATOR

    def __init__(self, ctx: Context, nxp, nyp, nzp, precision=32, cubetype=L.CRV_PARAMETER, nx=None, ny=None, nz=None,
                 handle=None):
        self.ctx = ctx
        if handle is None:
            h = C.c_void_p()
            ctx.check(ctx.lib.crv_grid_create(ctx.h, nx or nxp, ny or nyp, nz or nzp, nxp, nyp, nzp, precision,
                                              cubetype, C.byref(h)), "crv_grid_create")
        else:
            h = C.c_void_p(handle)
        self.h = h
        self.nxp, self.nyp, self.nzp = nxp, nyp, nzp
        self.cnxp = nxp // 2 + 1
        self.rnxp = 2 * self.cnxp
        self.precision = precision
        self.dtype = np.float64 if precision == 64 else np.float32

    # --- state -------------------------------------------------------------
    def info(self) -> L.GridInfo:
        i = L.GridInfo()
        self.ctx.check(self.ctx.lib.crv_grid_info_get(self.h, C.byref(i)), "info")
        return i

    def getIsTransformed(self) -> bool:
        return bool(self.info().is_transformed)

    def getType(self) -> int:
        return self.info().cubetype

    def setType(self, t: int):
        self.ctx.check(self.ctx.lib.crv_grid_set_cubetype(self.h, t), "set_cubetype")

    def setTransformedStatus(self, t: bool):
        self.ctx.check(self.ctx.lib.crv_grid_set_transformed(self.h, int(t)), "set_transformed")

    # --- data movement -------------------------------------------------------
    def upload_real(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype).reshape(-1)
        self.ctx.check(self.ctx.lib.crv_grid_upload_real(self.h, a.ctypes.data, a.size), "upload_real")

    def upload_complex(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        if np.iscomplexobj(a):
            a = np.stack([a.real, a.imag], axis=-1)
        a = np.ascontiguousarray(a, dtype=self.dtype).reshape(-1)
        self.ctx.check(self.ctx.lib.crv_grid_upload_complex(self.h, a.ctypes.data, a.size), "upload_complex")

    def download_real(self) -> np.ndarray:
        out = np.empty((self.nzp, self.nyp, self.rnxp), dtype=self.dtype)
        self.ctx.check(self.ctx.lib.crv_grid_download(self.h, out.ctypes.data, out.size), "download")
        return out

    def download_complex(self) -> np.ndarray:
        out = np.empty((self.nzp, self.nyp, self.cnxp, 2), dtype=self.dtype)
        self.ctx.check(self.ctx.lib.crv_grid_download(self.h, out.ctypes.data, out.size), "download")
        return out[..., 0] + 1j * out[..., 1]

    # --- reference operations ------------------------------------------------
    def fftInPlace(self):
        self.ctx.check(self.ctx.lib.crv_fft_r2c(self.h), "fftInPlace")

    def invFFTInPlace(self):
        self.ctx.check(self.ctx.lib.crv_fft_c2r(self.h), "invFFTInPlace")

    def fftFrom(self, src: "FFTGrid"):
        self.ctx.check(self.ctx.lib.crv_fft_r2c_from(self.h, src.h), "fft_r2c_from")

    def invFFTAdd(self, src: "FFTGrid", add: "FFTGrid | None"):
        self.ctx.check(self.ctx.lib.crv_fft_c2r_add(self.h, src.h, add.h if add else None), "fft_c2r_add")

    def add(self, other: "FFTGrid"):
        self.ctx.check(self.ctx.lib.crv_grid_add(self.h, other.h), "add")

    def multiplyByScalar(self, s: float):
        self.ctx.check(self.ctx.lib.crv_grid_multiply_by_scalar(self.h, float(s)), "multiplyByScalar")

    def copyFrom(self, other: "FFTGrid"):
        self.ctx.check(self.ctx.lib.crv_grid_copy(self.h, other.h), "copy")

    def fillInComplexNoise(self, state: int) -> int:
        st = C.c_uint(state & 0xFFFFFFFF)
        self.ctx.check(self.ctx.lib.crv_fill_complex_noise(self.h, C.byref(st)), "fillInComplexNoise")
        return st.value

    def close(self):
        # a grid outliving its context (closed first) is left to the process
        # exit: its memory went with the context's device
        if self.h and self.ctx.h:
            self.ctx.lib.crv_grid_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _arr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


@dataclass
class PostResult:
    n_wavenumbers: int
    n_inverted: int
    n_chol_fail: int
    n_fp64: int
    max_cond: float


def _post_params(n, *, A, theta, tables, nz, lz, low_cut=5.0, high_cut=55.0, wnc=0.1, min_rel_thick=1.0, sigma0,
                 mode=L.CRV_MODE_PLAIN, cond_threshold=1e5, lowk_fraction=0.02, separable=False, cxy=None, ct=None,
                 skip_mean_inverse=False):
    """crv_post_params and the arrays it points into (kept alive by the caller)."""
    keep = []
    A_ = np.ascontiguousarray(A, dtype=np.float32); keep.append(A_)
    th = np.ascontiguousarray(theta, dtype=np.float64); keep.append(th)
    kwd, em1, em2, em3 = tables.as_interleaved(); keep += [kwd, em1, em2, em3]
    s0 = np.ascontiguousarray(sigma0, dtype=np.float64); keep.append(s0)
    p = L.PostParams()
    p.ntheta = n; p.nz = int(nz); p.lz = float(lz); p.min_rel_thick = float(min_rel_thick)
    p.low_cut = float(np.float32(low_cut)); p.high_cut = float(np.float32(high_cut)); p.wnc = float(np.float32(wnc))
    p.A = _arr(A_, C.c_float); p.err_theta_cov = _arr(th, C.c_double)
    p.kwd = _arr(kwd, C.c_double); p.em1 = _arr(em1, C.c_double)
    p.em2 = _arr(em2, C.c_double); p.em3 = _arr(em3, C.c_double)
    p.mode = mode; p.cond_threshold = cond_threshold; p.lowk_fraction = lowk_fraction
    p.sigma0 = _arr(s0, C.c_double)
    p.separable = int(separable)
    if separable:
        cx = np.ascontiguousarray(cxy, dtype=np.float64); c_t = np.ascontiguousarray(ct, dtype=np.float64)
        keep += [cx, c_t]
        p.cxy = _arr(cx, C.c_double); p.ct = _arr(c_t, C.c_double)
    p.skip_mean_inverse = int(skip_mean_inverse)
    return p, keep


def computePostMeanResidAndFFTCov(ctx: Context, mean, data, err_corr, cov, *, want_stats=True,
                                  **kw) -> PostResult | None:
    """Crava::computePostMeanResidAndFFTCov over FFTGrid objects (crava.h:43-64).
    Keywords: A, theta, tables, nz, lz, low_cut, high_cut, wnc, min_rel_thick,
    sigma0, mode, cond_threshold, lowk_fraction, separable, cxy, ct,
    skip_mean_inverse."""
    n = len(data)
    p, keep = _post_params(n, **kw)
    means = (C.c_void_p * 3)(*[g.h.value for g in mean])
    datas = (C.c_void_p * n)(*[g.h.value for g in data])
    covs = (C.c_void_p * 6)(*[g.h.value for g in cov])
    st = L.PostStats()
    ctx.check(ctx.lib.crv_posterior(ctx.h, C.byref(p), means, datas, err_corr.h if err_corr is not None else None,
                                    covs, C.byref(st) if want_stats else None), "computePostMeanResidAndFFTCov")
    if not want_stats:
        return None
    return PostResult(st.n_wavenumbers, st.n_inverted, st.n_chol_fail, st.n_fp64, st.max_cond)


def simulate(ctx: Context, cov, post_mean, out, *, noise_mode=L.CRV_NOISE_COUNTER, lcg_state=None, seed=0,
             realization=0, host_noise=None, fp64_chol=False):
    """One realization of Crava::simulate (crava.cpp:1599-1708).  Returns the
    updated LCG state for the reference stream."""
    spec = L.NoiseSpec()
    spec.mode = noise_mode
    st = C.c_uint((lcg_state or 0) & 0xFFFFFFFF)
    spec.lcg_state = C.pointer(st)
    spec.seed = seed
    spec.realization = realization
    keep = []
    if host_noise is not None:
        for g in range(3):
            a = np.ascontiguousarray(np.stack([host_noise[g].real, host_noise[g].imag], -1),
                                     dtype=cov[0].dtype)
            keep.append(a)
            spec.host_noise[g] = a.ctypes.data
    covs = (C.c_void_p * 6)(*[g.h.value for g in cov])
    pm = (C.c_void_p * 3)(*[g.h.value if g is not None else None for g in post_mean]) if post_mean else None
    outs = (C.c_void_p * 3)(*[g.h.value for g in out])
    ctx.check(ctx.lib.crv_simulate(ctx.h, covs, pm, C.byref(spec), int(fp64_chol), outs), "simulate")
    return st.value


def fill_param_corr(ctx: Context, cov, corr_xy, circ_corr_t, grad_i, grad_j, sigma0):
    """crv_fill_param_corr: the six real prior covariance grids of the
    general prior (fillInParamCorr times Sigma0).  corr_xy (nyp, nxp),
    circ_corr_t (nzp,), sigma0 (3, 3) or its six entries aa bb rr ab ar br."""
    cx = np.ascontiguousarray(corr_xy, np.float64)
    ct = np.ascontiguousarray(circ_corr_t, np.float32)
    S = np.asarray(sigma0, np.float64)
    s6 = S.reshape(-1) if S.size == 6 else np.array([S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2]])
    s6 = np.ascontiguousarray(s6, np.float64)
    covs = (C.c_void_p * 6)(*[g.h.value for g in cov])
    ctx.check(ctx.lib.crv_fill_param_corr(covs, cx.ctypes.data_as(C.POINTER(C.c_double)),
                                          ct.ctypes.data_as(C.POINTER(C.c_float)), float(grad_i), float(grad_j),
                                          s6.ctypes.data_as(C.POINTER(C.c_double))), "fill_param_corr")


def fill_err_corr(ctx: Context, err, corr_xy, grad_i, grad_j):
    """crv_fill_err_corr: the real error correlation grid (fillInErrCorr)."""
    cx = np.ascontiguousarray(corr_xy, np.float64)
    ctx.check(ctx.lib.crv_fill_err_corr(err.h, cx.ctypes.data_as(C.POINTER(C.c_double)), float(grad_i),
                                        float(grad_j)), "fill_err_corr")


def krig_post(ctx: Context, mean, cov, points, data_target, smooth=False) -> dict:
    """crv_krig_post (Crava::doPostKriging): the real posterior means mean[3]
    kriged in place to the well points (rows i, j, k, log Vp, log Vs,
    log rho; L.CRV_WELL_MISSING where absent) with the six real posterior
    covariance grids cov[6].  Returns the planner's statistics."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 6)
    arr = (L.WellPoint * max(1, len(pts)))()
    for n, r in enumerate(pts):
        arr[n] = L.WellPoint(int(r[0]), int(r[1]), int(r[2]), float(r[3]), float(r[4]), float(r[5]))
    st = L.KrigStats()
    means = (C.c_void_p * 3)(*[g.h.value for g in mean])
    covs = (C.c_void_p * 6)(*[g.h.value for g in cov])
    ctx.check(ctx.lib.crv_krig_post(means, covs, arr, len(pts), int(data_target), int(bool(smooth)), C.byref(st)),
              "krig_post")
    return dict(block=tuple(st.block), block_ext=tuple(st.block_ext), range=tuple(st.range),
                data_points=st.data_points, blocks=st.blocks, empty_blocks=tuple(st.empty_blocks),
                max_block_data=tuple(st.max_block_data), cells_updated=tuple(st.cells_updated),
                cells_missing=tuple(st.cells_missing))


def noise_guard(ctx: Context, scale=0.0, force_table=False):
    """crv_noise_guard: set the reference decode's guard scale (0 keeps it)
    and whether the next decode tables every normal near a margin; returns
    the last decode's (fix passes, host-evaluated, patched, runs)."""
    info = (C.c_longlong * 4)()
    ctx.check(ctx.lib.crv_noise_guard(ctx.h, float(scale), int(bool(force_table)), info), "noise_guard")
    return tuple(int(v) for v in info)


def fill_prior_spectra(ctx: Context, sigma0, cxy, ct, cov, noalloc=False):
    """crv_fill_prior_spectra, or with noalloc its allocation-free variant."""
    s0 = np.ascontiguousarray(sigma0, np.float64)
    cx = np.ascontiguousarray(cxy, np.float64)
    c_t = np.ascontiguousarray(ct, np.float64)
    covs = (C.c_void_p * 6)(*[g.h.value for g in cov])
    fn = ctx.lib.crv_fill_prior_spectra_noalloc if noalloc else ctx.lib.crv_fill_prior_spectra
    ctx.check(fn(ctx.h, _arr(s0, C.c_double), _arr(cx, C.c_double), _arr(c_t, C.c_double), covs),
              "fill_prior_spectra")


def grid_fill(g: FFTGrid, value: float):
    g.ctx.check(g.ctx.lib.crv_grid_fill(g.h, float(value)), "grid_fill")


def forward_model(ctx: Context, A, kwd_interleaved, model, data):
    A_ = np.ascontiguousarray(A, np.float32)
    K = np.ascontiguousarray(kwd_interleaved, np.float64)
    ms = (C.c_void_p * 3)(*[g.h.value for g in model])
    ds = (C.c_void_p * len(data))(*[g.h.value for g in data])
    ctx.check(ctx.lib.crv_forward_model(ctx.h, len(data), _arr(A_, C.c_float), _arr(K, C.c_double), ms, ds),
              "forward_model")


def add_white_noise(g: FFTGrid, sigma: float, seed: int, stream: int):
    g.ctx.check(g.ctx.lib.crv_add_white_noise(g.h, float(sigma), seed, stream), "add_white_noise")


def moments(g: FFTGrid):
    s, s2, n = C.c_double(), C.c_double(), C.c_longlong()
    g.ctx.check(g.ctx.lib.crv_grid_moments(g.h, C.byref(s), C.byref(s2), C.byref(n)), "moments")
    return s.value, s2.value, n.value


def noise_draws(nxp, nyp, nzp) -> int:
    return int(L.lib().crv_noise_draws(nxp, nyp, nzp))


def find_closest_factorable(n: int) -> int:
    return int(L.lib().crv_find_closest_factorable(n))


def padding_size(n: int, frac: float) -> int:
    return int(L.lib().crv_padding_size(n, frac))


def circ_corr_t(corr, nzp: int, low_int_cut: int) -> np.ndarray:
    """SeismicParametersHolder::computeCircCorrT + makeCircCorrTPosDef
    (crv_circ_corr_t): nzp floats; raises when the correlation is undefined."""
    c = np.ascontiguousarray(corr, np.float32)
    out = np.empty(nzp, np.float32)
    rc = L.lib().crv_circ_corr_t(c.ctypes.data_as(C.POINTER(C.c_float)), c.size, nzp, int(low_int_cut),
                                 out.ctypes.data_as(C.POINTER(C.c_float)))
    if rc != 0:
        raise CravaError(f"circ_corr_t failed ({rc}): circular temporal correlation undefined")
    return out


# rock physics constants of the learned-surrogate workload: quartz and clay
# moduli in kPa, densities in g/cm3, fluids at 80 C
ROCK_DEFAULTS = dict(k_quartz=37.0e6, g_quartz=44.0e6, rho_quartz=2.65, k_clay=21.0e6, g_clay=7.0e6,
                     rho_clay=2.58, temperature=80.0, salinity=0.035, gas_gravity=0.6)


def rock_forward_device(ctx: Context, params: dict, d_in: int, d_out: int, n: int, out_fp32: bool = False,
                        d_status: int = 0):
    """crv_rock_forward on device buffers (addresses as ints)."""
    p = L.RockParams(**params)
    ctx.check(ctx.lib.crv_rock_forward(ctx.h, C.byref(p), int(n), C.c_void_p(d_in), C.c_void_p(d_out),
                                       int(out_fp32), C.c_void_p(d_status) if d_status else None),
              "rock_forward")


def rock_forward(ctx: Context, x: np.ndarray, params: dict | None = None, out_fp32: bool = False):
    """(n, 5) inputs (porosity, clay, Sw, aspect ratio, effective pressure MPa)
    -> ((n, 3) ln vp, ln vs, ln rho; (n,) status).  Device memory via torch."""
    import torch
    x = np.ascontiguousarray(x, np.float64)
    n = x.shape[0]
    dev = torch.device("cuda", 0)
    d_in = torch.from_numpy(x.ravel()).to(dev)
    d_out = torch.empty(n * 3, dtype=torch.float32 if out_fp32 else torch.float64, device=dev)
    d_st = torch.zeros(n, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    rock_forward_device(ctx, params or ROCK_DEFAULTS, d_in.data_ptr(), d_out.data_ptr(), n, out_fp32,
                        d_st.data_ptr())
    ctx.synchronize()
    return d_out.cpu().numpy().reshape(n, 3), d_st.cpu().numpy()


class PostPlan:
    """crv_post_plan: the posterior's per-kz tables on the device for one grid
    size; exec launches kernels only (capturable in a Graph)."""

    def __init__(self, ctx: Context, nxp, nyp, nzp, precision, n, **kw):
        self.ctx = ctx
        self.n = n
        p, keep = _post_params(n, **kw)
        h = C.c_void_p()
        ctx.check(ctx.lib.crv_post_plan_create(ctx.h, C.byref(p), nxp, nyp, nzp, precision, C.byref(h)),
                  "post_plan_create")
        self.h = h

    def exec(self, mean, data, err_corr, cov):
        means = (C.c_void_p * 3)(*[g.h.value for g in mean])
        datas = (C.c_void_p * len(data))(*[g.h.value for g in data])
        covs = (C.c_void_p * 6)(*[g.h.value for g in cov])
        self.ctx.check(self.ctx.lib.crv_posterior_exec(self.h, means, datas,
                                                       err_corr.h if err_corr is not None else None, covs),
                       "posterior_exec")

    def invert(self, mean, data, cov, sims=None, *, mean_src=None, data_src=None, seed=0, realization0=0,
               chunk_rows=0, lcg_state=None):
        """crv_invert: forward transforms of the real means and data (or of
        mean_src / data_src into them), the posterior, len(sims) // 3
        realizations and the inverse transforms, pipelined by chunks of y
        rows.  With lcg_state the realizations draw the reference stream
        from that state (crv_invert_noise) and the updated state is
        returned."""
        arr = lambda gs: (C.c_void_p * len(gs))(*[g.h.value for g in gs]) if gs else None
        sims = sims or []
        if lcg_state is None:
            self.ctx.check(self.ctx.lib.crv_invert(self.h, arr(mean), arr(data), arr(mean_src), arr(data_src),
                                                   arr(cov), seed, realization0, len(sims) // 3, arr(sims),
                                                   chunk_rows), "invert")
            return None
        spec = L.NoiseSpec()
        spec.mode = L.CRV_NOISE_REFERENCE
        st = C.c_uint(int(lcg_state) & 0xFFFFFFFF)
        spec.lcg_state = C.pointer(st)
        self.ctx.check(self.ctx.lib.crv_invert_noise(self.h, arr(mean), arr(data), arr(mean_src), arr(data_src),
                                                     arr(cov), C.byref(spec), len(sims) // 3, arr(sims),
                                                     chunk_rows), "invert_noise")
        return st.value

    def stats(self) -> PostResult:
        st = L.PostStats()
        self.ctx.check(self.ctx.lib.crv_post_plan_stats(self.h, C.byref(st)), "post_plan_stats")
        return PostResult(st.n_wavenumbers, st.n_inverted, st.n_chol_fail, st.n_fp64, st.max_cond)

    def close(self):
        if self.h and self.ctx.h:
            self.ctx.lib.crv_post_plan_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(ctx: Context, nxp, nyp, nzp, precision=32, reps=3) -> dict:
    """crv_fft_autotune: the launch shapes chosen for this grid size."""
    sh = L.FftShape()
    ctx.check(ctx.lib.crv_fft_autotune(ctx.h, nxp, nyp, nzp, precision, reps, C.byref(sh)), "fft_autotune")
    return dict(rows_per_tile=sh.rows_per_tile, tx_y=sh.tx_y, tx_z=sh.tx_z, radix_order=list(sh.radix_order),
                ms_r2c=sh.ms_r2c, ms_c2r=sh.ms_c2r)


def grid_batch(ctx: Context, count, nxp, nyp, nzp, precision=32, cubetype=L.CRV_PARAMETER, nx=None, ny=None,
               nz=None) -> list:
    """count FFTGrids in one contiguous allocation (crv_grid_create_batch)."""
    hs = (C.c_void_p * count)()
    ctx.check(ctx.lib.crv_grid_create_batch(ctx.h, nx or nxp, ny or nyp, nz or nzp, nxp, nyp, nzp, precision,
                                            cubetype, count, hs), "grid_create_batch")
    return [FFTGrid(ctx, nxp, nyp, nzp, precision, cubetype, handle=hs[i]) for i in range(count)]


def simulate_batch(ctx: Context, cov, post_mean, out, *, seed=0, realization0=0, fp64_chol=False):
    """len(out) // 3 realizations with the counter stream; out from grid_batch.
    fp64_chol: 0 float factor, 1 FP64 with refinement, 2 float with FP64
    where the float factor is rejected or ill-conditioned."""
    nreal = len(out) // 3
    covs = (C.c_void_p * 6)(*[g.h.value for g in cov])
    pm = (C.c_void_p * 3)(*[g.h.value for g in post_mean]) if post_mean else None
    outs = (C.c_void_p * len(out))(*[g.h.value for g in out])
    ctx.check(ctx.lib.crv_simulate_batch(ctx.h, covs, pm, seed, realization0, nreal, int(fp64_chol), outs),
              "simulate_batch")


class Graph:
    """A HIP graph captured from the context's stream:

        g = Graph.begin(ctx); ...launches...; g.end(); g.launch()
    """

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.h = None

    @classmethod
    def begin(cls, ctx: Context) -> "Graph":
        g = cls(ctx)
        ctx.check(ctx.lib.crv_graph_begin(ctx.h), "graph_begin")
        return g

    def end(self) -> "Graph":
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.crv_graph_end(self.ctx.h, C.byref(h)), "graph_end")
        self.h = h
        return self

    def launch(self):
        self.ctx.check(self.ctx.lib.crv_graph_launch(self.h), "graph_launch")

    def close(self):
        if self.h and self.ctx.h:
            self.ctx.lib.crv_graph_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fft_r2c_batch(grids, src=None):
    """FFTCovGrids over one contiguous batch (grid_batch); with src, the
    spectra of the src batch are written into grids."""
    ctx = grids[0].ctx
    d = (C.c_void_p * len(grids))(*[g.h.value for g in grids])
    if src is None:
        ctx.check(ctx.lib.crv_fft_r2c_batch(d, len(grids)), "fft_r2c_batch")
    else:
        s_ = (C.c_void_p * len(src))(*[g.h.value for g in src])
        ctx.check(ctx.lib.crv_fft_r2c_batch_from(d, s_, len(grids)), "fft_r2c_batch_from")


def fft_c2r_batch(grids):
    """invFFTCovGrids over one contiguous batch (grid_batch)."""
    ctx = grids[0].ctx
    d = (C.c_void_p * len(grids))(*[g.h.value for g in grids])
    ctx.check(ctx.lib.crv_fft_c2r_batch(d, len(grids)), "fft_c2r_batch")


# ---- output: derived parameters and CRAVA's binary grid format --------------
def derived(ctx: Context, kind: int, alpha=None, beta=None, rho=None, out=None) -> FFTGrid:
    """crv_grid_derived (ParameterOutput's per-voxel formulas) into out, or a
    new real grid shaped like the first input."""
    g0 = next(g for g in (alpha, beta, rho, out) if g is not None)
    if out is None:
        out = FFTGrid(ctx, g0.nxp, g0.nyp, g0.nzp, precision=g0.precision)
    h = lambda g: g.h if g is not None else None
    ctx.check(ctx.lib.crv_grid_derived(ctx.h, kind, h(alpha), h(beta), h(rho), out.h), "derived")
    return out


def crava_header(sb: "L.SimboxInfo", rnxp: int, nyp: int, nzp: int) -> bytes:
    """The CRAVA binary header (crv_crava_header); needs no GPU."""
    lib = L.lib()
    n = lib.crv_crava_header(C.byref(sb), rnxp, nyp, nzp, None, 0)
    buf = C.create_string_buffer(n)
    if lib.crv_crava_header(C.byref(sb), rnxp, nyp, nzp, buf, n) != n:
        raise CravaError("crava_header failed")
    return buf.raw[:n]


def write_crava(g: FFTGrid, path: str, sb: "L.SimboxInfo"):
    """FFTGrid::writeCravaFile through crv_grid_write_crava."""
    g.ctx.check(g.ctx.lib.crv_grid_write_crava(g.h, str(path).encode(), C.byref(sb)), "write_crava")


def storm_header(si: "L.StormInfo", cubetype: int, nx: int, ny: int, nz: int, ascii: bool = False) -> bytes:
    """The STORM header (crv_storm_header); needs no GPU."""
    lib = L.lib()
    n = lib.crv_storm_header(C.byref(si), cubetype, nx, ny, nz, int(ascii), None, 0)
    if n < 0:
        raise CravaError("storm_header: invalid arguments")
    buf = C.create_string_buffer(n)
    if lib.crv_storm_header(C.byref(si), cubetype, nx, ny, nz, int(ascii), buf, n) != n:
        raise CravaError("storm_header failed")
    return buf.raw[:n]


def write_storm(g: FFTGrid, path: str, si: "L.StormInfo", ascii: bool = False, padding: bool = False,
                scientific: bool = False):
    """FFTGrid::writeStormFile through crv_grid_write_storm (path complete,
    suffix included)."""
    g.ctx.check(g.ctx.lib.crv_grid_write_storm(g.h, str(path).encode(), C.byref(si), int(ascii), int(padding),
                                               int(scientific)), "write_storm")


def read_crava(g: FFTGrid, path: str) -> "L.SimboxInfo":
    """FFTGrid::readCravaFile through crv_grid_read_crava; returns the geometry."""
    sb = L.SimboxInfo()
    g.ctx.check(g.ctx.lib.crv_grid_read_crava(g.h, str(path).encode(), C.byref(sb)), "read_crava")
    return sb
