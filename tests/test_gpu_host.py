"""The host C++ classes (FFTGrid, SeismicParametersHolder, Crava) driven
through libcrava_host.so give the same inversion as the C ABI driven from
Python, and their model set-up (Ricker tables, reflection rows,
errThetaCov) matches the Python set-up."""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_rms
from crava_amd import PKG_DIR, model
from pipeline import Case, gpu_posterior

pytestmark = pytest.mark.gpu


def test_host_classes_match_abi(ctx):
    s = model.config_small(20, 16, 32, 3)
    case = Case(s)
    host = C.CDLL(str(PKG_DIR / "libcrava_host.so"))
    fp = C.POINTER(C.c_float)
    rs = s.nzp * s.nyp * s.rnxp
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    means = [f32(m) for m in case.syn.mean]
    datas = [f32(d) for d in case.syn.data]
    covs = [f32(c) for c in case.cov_real]
    err = f32(case.err_real)
    arr = lambda lst: (fp * len(lst))(*[a.ctypes.data_as(fp) for a in lst])
    pm_out = [np.zeros(rs, np.float32) for _ in range(3)]
    cov_out = [np.zeros(2 * s.cnxp * s.nyp * s.nzp, np.float32) for _ in range(6)]
    real_out = [np.zeros(rs, np.float32) for _ in range(3)]
    dp = lambda a: np.ascontiguousarray(a, np.float64).ctypes.data_as(C.POINTER(C.c_double))
    S0 = np.ascontiguousarray(s.sigma0())
    ang = np.array(s.angles_deg, np.float64)
    pk = np.array(s.peak_hz, np.float64)
    sn = np.array(s.sn, np.float64)
    ebuf = C.create_string_buffer(512)
    rc = host.crvh_invert(s.nxp, s.nyp, s.nzp, s.n, arr(means), arr(datas), arr(covs), err.ctypes.data_as(fp),
                          S0.ctypes.data_as(C.POINTER(C.c_double)), ang.ctypes.data_as(C.POINTER(C.c_double)),
                          pk.ctypes.data_as(C.POINTER(C.c_double)), sn.ctypes.data_as(C.POINTER(C.c_double)),
                          C.c_double(s.dt), C.c_double(s.vsvp), C.c_double(s.low_cut), C.c_double(s.high_cut),
                          C.c_double(s.wnc), 0, C.c_uint(0), arr(pm_out), arr(cov_out), arr(real_out), ebuf, 512)
    assert rc == 0, ebuf.value
    r = gpu_posterior(ctx, case, prec=32)
    for a in range(3):
        got = pm_out[a].reshape(s.nzp, s.nyp, s.rnxp)[:, :, : s.nxp] - s.mu[a]
        want = r["mean"][a].download_real()[:, :, : s.nxp] - s.mu[a]
        assert rel_rms(got, want) < 1e-5
    for q in range(6):
        got = cov_out[q].reshape(s.nzp, s.nyp, s.cnxp, 2)
        assert rel_rms(got[..., 0] + 1j * got[..., 1], r["cov"][q].download_complex()) < 1e-5
    for a in range(3):
        assert np.all(np.isfinite(real_out[a]))


def test_host_fftgrid_output_path(ctx, tmp_path):
    """Host FFTGrid::expTransf + writeCravaFile + readCravaFile: the file is
    CRAVA's binary format of exp(grid) and reads back unchanged; the same
    grid through writeStormFile is the STORM binary file of the padded grid."""
    nxp, nyp, nzp = 30, 20, 16
    rng = np.random.default_rng(5)
    x = (np.log(2500.0) + 0.1 * rng.standard_normal((nzp, nyp, 2 * (nxp // 2 + 1)))).astype(np.float32)
    host = C.CDLL(str(PKG_DIR / "libcrava_host.so"))
    out = np.zeros_like(x)
    ebuf = C.create_string_buffer(512)
    base = str(tmp_path / "vp")
    fp = C.POINTER(C.c_float)
    rc = host.crvh_write_read_crava(nxp, nyp, nzp, x.ctypes.data_as(fp), base.encode(), out.ctypes.data_as(fp),
                                    ebuf, 512)
    assert rc == 0, ebuf.value
    want = np.exp(x.astype(np.float64)).astype(np.float32)
    assert np.max(np.abs(out - want) / want) <= 1.2e-7
    raw = (tmp_path / "vp.crava").read_bytes()
    assert raw.startswith(b"crava_fftgrid_binary\n")
    assert raw[-out.size * 4:] == out.astype(">f4").tobytes()
    # FFTGrid::writeStormFile of the same grid (binary, padded, flat)
    from crava_amd import _lib as L
    from crava_amd.hip import storm_header
    si = L.StormInfo(xmin=0.0, lx=25.0 * nxp, ymin=0.0, ly=25.0 * nyp, lz=4.0 * nzp, angle=0.0, flat=1)
    head = storm_header(si, L.CRV_PARAMETER, nxp, nyp, nzp)
    storm = (tmp_path / "vp.storm").read_bytes()
    assert storm == head + out[:, :, :nxp].astype(">f4").tobytes() + b"0\n"
