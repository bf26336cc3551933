"""The C ABI library loads without a GPU and exports every entry point that
include/crava_hip.h declares; the host-only entry points work on the CPU."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    txt = (ROOT / "include" / "crava_hip.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(crv_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    from crava_amd import _lib
    assert sorted(_lib.EXPORTS) == _declared()


def test_header_constants_match_the_binding():
    """Every CRV_* integer constant the header defines has the same value in
    the Python binding (tuning keys, modes, grid types, error codes)."""
    from crava_amd import _lib
    txt = (ROOT / "include" / "crava_hip.h").read_text()
    consts = dict(re.findall(r"^#define (CRV_[A-Z0-9_]+)\s+(-?\d+)\s*$", txt, flags=re.M))
    assert any(k.startswith("CRV_TUNE_") for k in consts)
    for name, value in consts.items():
        if hasattr(_lib, name):
            assert getattr(_lib, name) == int(value), name
    tune = [k for k in consts if k.startswith("CRV_TUNE_")]
    assert all(hasattr(_lib, k) for k in tune), [k for k in tune if not hasattr(_lib, k)]


def test_library_exports_every_symbol():
    from crava_amd import _lib
    lib = ctypes.CDLL(str(_lib.LIBPATH))
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_host_entry_points_without_gpu():
    from crava_amd import _lib
    from oracle import oracle as O
    lib = _lib.lib()
    assert lib.crv_version().startswith(b"crava_hip")
    for n in (1, 7, 11, 13, 97, 450, 451, 1000, 4097):
        assert lib.crv_find_closest_factorable(n) == O.find_closest_factorable(n)
    for dims in ((8, 6, 10), (9, 4, 6), (64, 64, 128), (450, 450, 320)):
        assert lib.crv_noise_draws(*dims) == O.noise_draws(*dims)
    assert lib.crv_padding_size(320, 0.0625) == 343


def test_crava_header_layout():
    """The CRAVA binary header (FFTGrid::writeCravaFile, reference
    src/fftgrid.cpp:2411-2440): the type line, then 4 doubles, 2 ints, 7
    doubles and rnxp, nyp, nzp, big-endian IEEE as NRLib writes by default
    (libs/nrlib/iotools/fileio.hpp:266-268) -- checked against numpy's
    big-endian encoders."""
    import numpy as np
    from crava_amd import _lib as L
    from crava_amd.hip import crava_header
    sb = L.SimboxInfo(x0=456000.5, y0=6.7e6, dx=25.0, dy=12.5, nx=450, ny=451, il0=1001.0, xl0=2002.0,
                      il_step_x=0.5, il_step_y=-0.25, xl_step_x=1.0 / 3.0, xl_step_y=-2.0, angle=0.3)
    got = crava_header(sb, 452, 450, 320)
    want = (b"crava_fftgrid_binary\n" + np.array([456000.5, 6.7e6, 25.0, 12.5], ">f8").tobytes()
            + np.array([450, 451], ">i4").tobytes()
            + np.array([1001.0, 2002.0, 0.5, -0.25, 1.0 / 3.0, -2.0, 0.3], ">f8").tobytes()
            + np.array([452, 450, 320], ">i4").tobytes())
    assert got == want


def test_storm_header_text():
    """The STORM header (Simbox::getStormHeader, reference src/simbox.cpp:
    408-458; reals through NRLib::ToString(x, 6), i.e. fixed notation with 6
    decimals): named surfaces, binary; flat, text."""
    from crava_amd import _lib as L
    from crava_amd.hip import storm_header
    si = L.StormInfo(xmin=456000.5, lx=11250.0, ymin=6700000.25, ly=5625.0, lz=1280.0, angle=0.3,
                     top_name=b"top.irap", bot_name=b"base.irap", flat=0)
    want = ("storm_petro_binary\n0 3 -99999.000000\nFFTGrid\n"
            "456000.500000 11250.000000 6700000.250000 5625.000000 top.irap base.irap 0.0 0.0\n"
            "1280.000000 17.188734\n\n450 451 320\n")
    assert storm_header(si, 3, 450, 451, 320) == want.encode()
    si.flat = 1
    want = ("storm_petro_ascii\n0 2 -99999.000000\nFFTGrid\n"
            "456000.500000 11250.000000 6700000.250000 5625.000000 0.0 1280.000000 0.0 0.0\n"
            "1280.000000 17.188734\n\n64 64 128\n")
    assert storm_header(si, 2, 64, 64, 128, ascii=True) == want.encode()


def test_header_encoders_reject_bad_arguments():
    """The header encoders validate before writing: no simbox, a non-flat
    STORM header without surface names, empty dimensions, or a buffer that
    is too short give CRV_E_INVALID and leave the buffer alone."""
    import ctypes as C
    from crava_amd import _lib as L
    lib = L.lib()
    inval = -1  # CRV_E_INVALID (include/crava_hip.h)
    si = L.StormInfo(xmin=0.0, lx=1.0, ymin=0.0, ly=1.0, lz=1.0, angle=0.0, flat=0)
    assert lib.crv_storm_header(C.byref(si), 2, 4, 4, 4, 0, None, 0) == inval      # no surface names
    si.flat = 1
    assert lib.crv_storm_header(C.byref(si), 2, 0, 4, 4, 0, None, 0) == inval      # empty grid
    n = lib.crv_storm_header(C.byref(si), 2, 4, 4, 4, 0, None, 0)
    buf = C.create_string_buffer(b"x" * n, n)
    assert lib.crv_storm_header(C.byref(si), 2, 4, 4, 4, 0, buf, n - 1) == inval  # too short
    assert buf.raw == b"x" * n
    assert lib.crv_storm_header(None, 2, 4, 4, 4, 0, None, 0) == inval
    sb = L.SimboxInfo()
    m = lib.crv_crava_header(C.byref(sb), 10, 4, 4, None, 0)
    assert m > 0 and lib.crv_crava_header(C.byref(sb), 10, 4, 4, C.create_string_buffer(m), m - 1) == inval
    assert lib.crv_crava_header(None, 10, 4, 4, None, 0) == inval


def test_context_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from crava_amd import hip
    with pytest.raises(hip.CravaError):
        hip.Context(0)


def test_product_does_not_import_oracle():
    for p in (ROOT / "crava_amd").rglob("*.py"):
        src = p.read_text()
        assert "import oracle" not in src and "from oracle" not in src, p
    for p in list((ROOT / "crava_amd" / "csrc").rglob("*")):
        if p.is_file():
            assert "oracle" not in p.read_text(errors="ignore").lower(), p
