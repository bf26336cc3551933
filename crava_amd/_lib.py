"""ctypes binding of libcrava_hip.so (include/crava_hip.h).

The library is the only compute path: if it is missing, importing this
module raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIBPATH = Path(os.environ.get("CRAVA_HIP_LIB", _PKG / "libcrava_hip.so"))

CRV_OK = 0
CRV_FP32, CRV_FP64 = 32, 64
CRV_CTMISSING, CRV_DATA, CRV_PARAMETER, CRV_COVARIANCE, CRV_VELOCITY, CRV_OPERATOR = range(6)
CRV_MODE_PLAIN, CRV_MODE_FP32_REFINE = 0, 1
CRV_NOISE_REFERENCE, CRV_NOISE_COUNTER, CRV_NOISE_HOST = 0, 1, 2

# every symbol declared in include/crava_hip.h
EXPORTS = [
    "crv_ctx_create", "crv_ctx_destroy", "crv_last_error", "crv_ctx_synchronize", "crv_ctx_set_stream",
    "crv_version", "crv_grid_create", "crv_grid_destroy", "crv_grid_info_get", "crv_grid_set_cubetype",
    "crv_grid_set_transformed", "crv_grid_device_ptr", "crv_grid_upload_real", "crv_grid_upload_complex",
    "crv_grid_download", "crv_grid_copy", "crv_grid_add", "crv_grid_multiply_by_scalar", "crv_fft_r2c",
    "crv_fft_c2r", "crv_fft_r2c_from", "crv_fft_c2r_add", "crv_posterior", "crv_simulate",
    "crv_fill_complex_noise", "crv_noise_draws", "crv_find_closest_factorable", "crv_padding_size",
    "crv_ctx_profile", "crv_ctx_profile_report", "crv_fill_prior_spectra", "crv_grid_fill",
    "crv_forward_model", "crv_add_white_noise", "crv_grid_moments", "crv_ctx_tune", "crv_rock_forward",
    "crv_mlp_param_count", "crv_mlp_create", "crv_mlp_destroy", "crv_mlp_set_params", "crv_mlp_get_params",
    "crv_mlp_gradients", "crv_mlp_train_step", "crv_mlp_predict", "crv_mlp_flops", "crv_rock_refine",
    "crv_post_plan_create", "crv_post_plan_destroy", "crv_posterior_exec", "crv_post_plan_stats", "crv_invert",
    "crv_fft_autotune", "crv_grid_create_batch", "crv_simulate_batch", "crv_graph_begin", "crv_graph_end",
    "crv_graph_launch", "crv_graph_destroy", "crv_facies_density", "crv_facies_prob", "crv_facies_prob_undef",
    "crv_grid_derived", "crv_crava_header", "crv_grid_write_crava", "crv_grid_read_crava",
    "crv_storm_header", "crv_grid_write_storm",
    "crv_fft_r2c_batch", "crv_fft_r2c_batch_from", "crv_fft_c2r_batch",
    "crv_fill_prior_spectra_noalloc", "crv_circ_corr_t", "crv_facies_density_kernel",
    "crv_invert_noise", "crv_noise_guard", "crv_krig_post", "crv_fill_param_corr", "crv_fill_err_corr",
]
CRV_SMOOTH_FOURIER, CRV_SMOOTH_SAMPLED = 0, 1
CRV_TUNE_CHUNK_MB, CRV_TUNE_TX_Y, CRV_TUNE_TX_Z, CRV_TUNE_ROWS, CRV_TUNE_POST_BLOCKS = 1, 2, 3, 4, 5
CRV_TUNE_BATCH_MB = 6
CRV_TUNE_ORDER = 7
CRV_TUNE_ZPOST = 8
CRV_TUNE_ZINV = 9


class GridInfo(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("nxp", C.c_int), ("nyp", C.c_int), ("nzp", C.c_int),
                ("cnxp", C.c_int), ("rnxp", C.c_int),
                ("precision", C.c_int), ("cubetype", C.c_int), ("is_transformed", C.c_int),
                ("rsize", C.c_size_t), ("csize", C.c_size_t)]


_dp = C.POINTER(C.c_double)


class PostParams(C.Structure):
    _fields_ = [("ntheta", C.c_int), ("nz", C.c_int), ("lz", C.c_double), ("min_rel_thick", C.c_double),
                ("low_cut", C.c_double), ("high_cut", C.c_double), ("wnc", C.c_double),
                ("A", C.POINTER(C.c_float)), ("err_theta_cov", _dp),
                ("kwd", _dp), ("em1", _dp), ("em2", _dp), ("em3", _dp),
                ("mode", C.c_int), ("cond_threshold", C.c_double), ("lowk_fraction", C.c_double),
                ("separable", C.c_int), ("sigma0", _dp), ("cxy", _dp), ("ct", _dp),
                ("skip_mean_inverse", C.c_int)]


class PostStats(C.Structure):
    _fields_ = [("n_wavenumbers", C.c_longlong), ("n_inverted", C.c_longlong),
                ("n_chol_fail", C.c_longlong), ("n_fp64", C.c_longlong), ("max_cond", C.c_double)]


class RockParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("k_quartz", "g_quartz", "rho_quartz", "k_clay", "g_clay", "rho_clay",
                                           "temperature", "salinity", "gas_gravity")]


class FaciesBox(C.Structure):
    _fields_ = [("lo", C.c_double * 3), ("hi", C.c_double * 3)]


class SimboxInfo(C.Structure):
    _fields_ = [("x0", C.c_double), ("y0", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
                ("nx", C.c_int), ("ny", C.c_int), ("il0", C.c_double), ("xl0", C.c_double),
                ("il_step_x", C.c_double), ("il_step_y", C.c_double), ("xl_step_x", C.c_double),
                ("xl_step_y", C.c_double), ("angle", C.c_double)]


class StormInfo(C.Structure):
    _fields_ = [("xmin", C.c_double), ("lx", C.c_double), ("ymin", C.c_double), ("ly", C.c_double),
                ("lz", C.c_double), ("angle", C.c_double), ("top_name", C.c_char_p), ("bot_name", C.c_char_p),
                ("flat", C.c_int)]


CRV_OUT_EXP, CRV_OUT_AI, CRV_OUT_SI, CRV_OUT_VPVS, CRV_OUT_POISSON = 0, 1, 2, 3, 4
CRV_OUT_LAME_MU, CRV_OUT_LAME_LAMBDA, CRV_OUT_MU_RHO, CRV_OUT_LAMBDA_RHO = 5, 6, 7, 8


class FftShape(C.Structure):
    _fields_ = [("rows_per_tile", C.c_int), ("tx_y", C.c_int), ("tx_z", C.c_int), ("radix_order", C.c_int * 3),
                ("ms_r2c", C.c_double), ("ms_c2r", C.c_double)]


class MlpNorm(C.Structure):
    _fields_ = [("in_lo", C.c_float * 5), ("in_scale", C.c_float * 5), ("in_log_mask", C.c_int),
                ("out_mean", C.c_float * 3), ("out_std", C.c_float * 3)]


class WellPoint(C.Structure):
    _fields_ = [("i", C.c_int), ("j", C.c_int), ("k", C.c_int), ("alpha", C.c_float), ("beta", C.c_float),
                ("rho", C.c_float)]


class KrigStats(C.Structure):
    _fields_ = [("block", C.c_int * 3), ("block_ext", C.c_int * 3), ("range", C.c_int * 3),
                ("data_points", C.c_int), ("blocks", C.c_int), ("empty_blocks", C.c_int * 3),
                ("max_block_data", C.c_int * 3), ("cells_updated", C.c_longlong * 3),
                ("cells_missing", C.c_longlong * 3)]


CRV_WELL_MISSING = -99999.0


class NoiseSpec(C.Structure):
    _fields_ = [("mode", C.c_int), ("lcg_state", C.POINTER(C.c_uint)), ("seed", C.c_ulonglong),
                ("realization", C.c_ulonglong), ("host_noise", C.c_void_p * 3)]


def load(path: Path | None = None) -> C.CDLL:
    p = Path(path) if path else LIBPATH
    if not p.exists():
        raise RuntimeError(f"libcrava_hip.so not found at {p}: run `python -m crava_amd.build` first")
    # torch (device memory for the host-side drivers) brings its own HIP
    # runtime; loading it first lets this library bind to the same one, so
    # both see the device whichever is used first
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(p))
    vp = C.c_void_p
    i = C.c_int
    sig = {
        "crv_ctx_create": (i, [i, C.POINTER(vp)]),
        "crv_ctx_destroy": (i, [vp]),
        "crv_last_error": (C.c_char_p, [vp]),
        "crv_ctx_synchronize": (i, [vp]),
        "crv_ctx_set_stream": (i, [vp, vp]),
        "crv_version": (C.c_char_p, []),
        "crv_grid_create": (i, [vp, i, i, i, i, i, i, i, i, C.POINTER(vp)]),
        "crv_grid_destroy": (i, [vp]),
        "crv_grid_info_get": (i, [vp, C.POINTER(GridInfo)]),
        "crv_grid_set_cubetype": (i, [vp, i]),
        "crv_grid_set_transformed": (i, [vp, i]),
        "crv_grid_device_ptr": (i, [vp, C.POINTER(vp)]),
        "crv_grid_upload_real": (i, [vp, vp, C.c_size_t]),
        "crv_grid_upload_complex": (i, [vp, vp, C.c_size_t]),
        "crv_grid_download": (i, [vp, vp, C.c_size_t]),
        "crv_grid_copy": (i, [vp, vp]),
        "crv_grid_add": (i, [vp, vp]),
        "crv_grid_multiply_by_scalar": (i, [vp, C.c_double]),
        "crv_fft_r2c": (i, [vp]),
        "crv_fft_c2r": (i, [vp]),
        "crv_fft_r2c_from": (i, [vp, vp]),
        "crv_fft_c2r_add": (i, [vp, vp, vp]),
        "crv_posterior": (i, [vp, C.POINTER(PostParams), C.POINTER(vp), C.POINTER(vp), vp, C.POINTER(vp),
                              C.POINTER(PostStats)]),
        "crv_simulate": (i, [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(NoiseSpec), i, C.POINTER(vp)]),
        "crv_fill_complex_noise": (i, [vp, C.POINTER(C.c_uint)]),
        "crv_noise_draws": (C.c_longlong, [i, i, i]),
        "crv_find_closest_factorable": (i, [i]),
        "crv_padding_size": (i, [i, C.c_double]),
        "crv_ctx_profile": (i, [vp, i]),
        "crv_ctx_profile_report": (i, [vp, C.c_char_p, C.c_size_t]),
        "crv_fill_prior_spectra": (i, [vp, _dp, _dp, _dp, C.POINTER(vp)]),
        "crv_fill_prior_spectra_noalloc": (i, [vp, _dp, _dp, _dp, C.POINTER(vp)]),
        "crv_circ_corr_t": (i, [C.POINTER(C.c_float), i, i, i, C.POINTER(C.c_float)]),
        "crv_grid_fill": (i, [vp, C.c_double]),
        "crv_forward_model": (i, [vp, i, C.POINTER(C.c_float), _dp, C.POINTER(vp), C.POINTER(vp)]),
        "crv_add_white_noise": (i, [vp, C.c_double, C.c_ulonglong, C.c_ulonglong]),
        "crv_grid_moments": (i, [vp, _dp, _dp, C.POINTER(C.c_longlong)]),
        "crv_ctx_tune": (i, [vp, i, i]),
        "crv_rock_forward": (i, [vp, C.POINTER(RockParams), C.c_longlong, vp, vp, i, vp]),
        "crv_mlp_param_count": (C.c_longlong, []),
        "crv_mlp_create": (i, [vp, C.POINTER(MlpNorm), C.c_longlong, C.POINTER(vp)]),
        "crv_mlp_destroy": (i, [vp]),
        "crv_mlp_set_params": (i, [vp, C.POINTER(C.c_float)]),
        "crv_mlp_get_params": (i, [vp, C.POINTER(C.c_float)]),
        "crv_mlp_gradients": (i, [vp, vp, vp, C.c_longlong, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "crv_mlp_train_step": (i, [vp, vp, vp, C.c_longlong, C.c_float, i, C.POINTER(C.c_float)]),
        "crv_mlp_predict": (i, [vp, vp, C.c_longlong, vp]),
        "crv_mlp_flops": (i, [vp, _dp, _dp]),
        "crv_rock_refine": (i, [vp, C.POINTER(RockParams), C.c_longlong, vp, vp, C.POINTER(MlpNorm), vp, i,
                                C.c_float, C.POINTER(C.c_longlong)]),
        "crv_post_plan_create": (i, [vp, C.POINTER(PostParams), i, i, i, i, C.POINTER(vp)]),
        "crv_post_plan_destroy": (i, [vp]),
        "crv_posterior_exec": (i, [vp, C.POINTER(vp), C.POINTER(vp), vp, C.POINTER(vp)]),
        "crv_post_plan_stats": (i, [vp, C.POINTER(PostStats)]),
        "crv_invert": (i, [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp),
                           C.c_ulonglong, C.c_ulonglong, i, C.POINTER(vp), i]),
        "crv_invert_noise": (i, [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp),
                                 C.POINTER(NoiseSpec), i, C.POINTER(vp), i]),
        "crv_noise_guard": (i, [vp, C.c_double, i, C.POINTER(C.c_longlong)]),
        "crv_krig_post": (i, [C.POINTER(vp), C.POINTER(vp), C.POINTER(WellPoint), i, i, i, C.POINTER(KrigStats)]),
        "crv_fill_param_corr": (i, [C.POINTER(vp), C.POINTER(C.c_double), C.POINTER(C.c_float), C.c_float, C.c_float,
                                    C.POINTER(C.c_double)]),
        "crv_fill_err_corr": (i, [vp, C.POINTER(C.c_double), C.c_float, C.c_float]),
        "crv_fft_autotune": (i, [vp, i, i, i, i, i, C.POINTER(FftShape)]),
        "crv_grid_create_batch": (i, [vp, i, i, i, i, i, i, i, i, i, C.POINTER(vp)]),
        "crv_simulate_batch": (i, [vp, C.POINTER(vp), C.POINTER(vp), C.c_ulonglong, C.c_ulonglong, i, i,
                                   C.POINTER(vp)]),
        "crv_graph_begin": (i, [vp]),
        "crv_graph_end": (i, [vp, C.POINTER(vp)]),
        "crv_graph_launch": (i, [vp]),
        "crv_graph_destroy": (i, [vp]),
        "crv_facies_density": (i, [vp, vp, C.c_longlong, C.POINTER(FaciesBox), _dp, vp]),
        "crv_facies_density_kernel": (i, [vp, vp, C.c_longlong, C.POINTER(FaciesBox), _dp, i, vp]),
        "crv_fft_r2c_batch": (i, [C.POINTER(vp), i]),
        "crv_fft_r2c_batch_from": (i, [C.POINTER(vp), C.POINTER(vp), i]),
        "crv_fft_c2r_batch": (i, [C.POINTER(vp), i]),
        "crv_facies_prob": (i, [vp, i, C.POINTER(vp), _dp, C.POINTER(FaciesBox), C.POINTER(vp), C.POINTER(vp)]),
        "crv_facies_prob_undef": (i, [vp, i, C.POINTER(vp), _dp, C.c_double, C.POINTER(FaciesBox), C.POINTER(vp),
                                      C.POINTER(vp), vp]),
        "crv_grid_derived": (i, [vp, i, vp, vp, vp, vp]),
        "crv_crava_header": (i, [C.POINTER(SimboxInfo), i, i, i, C.c_char_p, C.c_size_t]),
        "crv_grid_write_crava": (i, [vp, C.c_char_p, C.POINTER(SimboxInfo)]),
        "crv_grid_read_crava": (i, [vp, C.c_char_p, C.POINTER(SimboxInfo)]),
        "crv_storm_header": (i, [C.POINTER(StormInfo), i, i, i, i, i, C.c_char_p, C.c_size_t]),
        "crv_grid_write_storm": (i, [vp, C.c_char_p, C.POINTER(StormInfo), i, i, i]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_LIB: C.CDLL | None = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB
