"""FFTGrid::fftInPlace / invFFTInPlace on the GPU against the reference's
FFTW 2.1.2 output (golden vectors), the FP64 oracle transform and a direct
DFT, including odd and radix-7 sizes and the 450x450x320 grid."""
import math

import numpy as np
import pytest

from conftest import cplx, rel_rms
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLD_SHAPES = [(8, 6, 10), (9, 7, 12), (12, 10, 16), (6, 5, 7), (16, 9, 15)]


def _grid(ctx, shape, prec, ct):
    from crava_amd.hip import FFTGrid
    nx, ny, nz = shape
    return FFTGrid(ctx, nx, ny, nz, precision=prec, cubetype=ct)


@pytest.mark.parametrize("shape", GOLD_SHAPES)
@pytest.mark.parametrize("ct", [1, 3])
def test_r2c_matches_fftw(ctx, golden, shape, ct):
    nx, ny, nz = shape
    x = golden[f"r2c_{nx}x{ny}x{nz}_{ct}_in"]
    ref = cplx(golden[f"r2c_{nx}x{ny}x{nz}_{ct}_out"])
    g = _grid(ctx, shape, 32, ct)
    g.upload_real(x)
    g.fftInPlace()
    assert g.getIsTransformed()
    assert rel_rms(g.download_complex(), ref) < 1e-6        # FP32 vs FFTW float
    g64 = _grid(ctx, shape, 64, ct)
    g64.upload_real(x.astype(np.float64))
    g64.fftInPlace()
    assert rel_rms(g64.download_complex(), O.fft_r2c(x, nx, ct)) < 1e-14


@pytest.mark.parametrize("shape", GOLD_SHAPES)
@pytest.mark.parametrize("ct", [2, 3])
def test_c2r_matches_fftw(ctx, golden, shape, ct):
    nx, ny, nz = shape
    spec = cplx(golden[f"c2r_{nx}x{ny}x{nz}_{ct}_in"])
    ref = golden[f"c2r_{nx}x{ny}x{nz}_{ct}_out"][:, :, :nx]
    g = _grid(ctx, shape, 32, ct)
    g.upload_complex(spec.astype(np.complex64))
    g.invFFTInPlace()
    assert not g.getIsTransformed()
    assert rel_rms(g.download_real()[:, :, :nx], ref) < 1e-6
    g64 = _grid(ctx, shape, 64, ct)
    g64.upload_complex(spec)
    g64.invFFTInPlace()
    assert rel_rms(g64.download_real()[:, :, :nx], O.fft_c2r(spec, nx, ct)[:, :, :nx]) < 1e-14


SHAPES = [(64, 64, 128), (450, 6, 10), (343, 7, 5), (14, 49, 25), (1, 1, 1), (2, 3, 5), (30, 1, 7),
          (96, 84, 150), (160, 240, 100), (256, 20, 200), (7, 9, 343),
          # the pairs 12 x 5, 7 x 6, 15 x 5 (x half lengths 60, 42, 75; y 42, 75), 15 x 14 and 21 x 15
          # (z and y 210, z 315)
          (120, 42, 210), (84, 75, 315), (150, 210, 7)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("prec", [32, 64])
def test_fft_vs_fp64_oracle(ctx, shape, prec):
    nx, ny, nz = shape
    rng = np.random.default_rng(nx * 131 + ny * 7 + nz)
    x = np.zeros((nz, ny, 2 * (nx // 2 + 1)))
    x[:, :, :nx] = rng.standard_normal((nz, ny, nx))
    g = _grid(ctx, shape, prec, 1)
    g.upload_real(x)
    g.fftInPlace()
    X = g.download_complex()
    tol = 2e-7 * math.log2(max(2, nx * ny * nz)) if prec == 32 else 1e-14
    assert rel_rms(X, O.fft_r2c(x, nx, 1)) < tol
    g.invFFTInPlace()
    y = g.download_real()[:, :, :nx]
    assert rel_rms(y, x[:, :, :nx]) < tol


@pytest.mark.parametrize("shape", [(24, 18, 20), (21, 18, 20), (343, 6, 5), (450, 8, 6)])
@pytest.mark.parametrize("prec", [32, 64])
def test_out_of_place_and_fused_add(ctx, prec, shape):
    """Out-of-place R2C and the C2R that adds a real grid in its last pass,
    over even and odd x lengths (odd lengths take the C2R pass's own
    branch for the added values)."""
    nx, ny, nz = shape
    rng = np.random.default_rng(5)
    x = np.zeros((nz, ny, 2 * (nx // 2 + 1)))
    x[:, :, :nx] = rng.standard_normal((nz, ny, nx))
    src = _grid(ctx, shape, prec, 2)
    dst = _grid(ctx, shape, prec, 2)
    add = _grid(ctx, shape, prec, 2)
    src.upload_real(x)
    a = rng.standard_normal(x.shape)
    add.upload_real(a)
    dst.fftFrom(src)
    assert not src.getIsTransformed() and dst.getIsTransformed()
    assert np.array_equal(src.download_real(), x.astype(src.dtype))
    out = _grid(ctx, shape, prec, 2)
    out.invFFTAdd(dst, add)
    tol = 1e-6 if prec == 32 else 1e-14
    assert rel_rms(out.download_real()[:, :, :nx], (x + a)[:, :, :nx]) < tol


def test_state_errors(ctx):
    from crava_amd.hip import CravaError
    g = _grid(ctx, (8, 8, 8), 32, 1)
    g.upload_real(np.zeros((8, 8, 10)))
    with pytest.raises(CravaError):
        g.invFFTInPlace()                     # not transformed
    g.fftInPlace()
    with pytest.raises(CravaError):
        g.fftInPlace()                        # already transformed
    g2 = _grid(ctx, (8, 8, 8), 32, 0)
    g2.upload_real(np.zeros((8, 8, 10)))
    with pytest.raises(CravaError):
        g2.fftInPlace()                       # CTMISSING


def test_unsupported_size(ctx):
    from crava_amd.hip import CravaError, FFTGrid
    with pytest.raises(CravaError):
        FFTGrid(ctx, 22, 4, 4)                # 11 is not a supported factor


def test_full_size_450x450x320(ctx):
    """Size-independent properties at the production grid: Parseval, round
    trip, and 24 wavenumbers against a direct DFT."""
    nx, ny, nz = 450, 450, 320
    rng = np.random.default_rng(11)
    x = np.zeros((nz, ny, 452), np.float32)
    x[:, :, :nx] = rng.standard_normal((nz, ny, nx), dtype=np.float32)
    g = _grid(ctx, (nx, ny, nz), 32, 1)
    g.upload_real(x)
    g.fftInPlace()
    X = g.download_complex()
    # Parseval with 1/sqrt(N): sum |x|^2 = |X0|^2 + 2 sum_{0<kx<nx/2} |X|^2 + |X_{nx/2}|^2
    e_t = float(np.sum(x[:, :, :nx].astype(np.float64) ** 2))
    w = np.full(X.shape[2], 2.0)
    w[0] = w[-1] = 1.0
    e_f = float(np.sum(np.abs(X.astype(np.complex128)) ** 2 * w))
    assert abs(e_f / e_t - 1) < 1e-5
    xs = x[:, :, :nx].astype(np.float64)
    s = 1 / math.sqrt(nx * ny * nz)
    for _ in range(24):
        kz, ky, kx = rng.integers(nz), rng.integers(ny), rng.integers(nx // 2 + 1)
        ex = np.exp(-2j * np.pi * kx * np.arange(nx) / nx)
        ey = np.exp(-2j * np.pi * ky * np.arange(ny) / ny)
        ez = np.exp(-2j * np.pi * kz * np.arange(nz) / nz)
        d = ((xs @ ex) @ ey) @ ez * s
        assert abs(X[kz, ky, kx] - d) < 2e-5 * math.sqrt(e_t / X.size) + 1e-6 * abs(d)
    g.invFFTInPlace()
    assert rel_rms(g.download_real()[:, :, :nx], x[:, :, :nx]) < 1e-6
