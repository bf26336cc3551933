"""Output next to the path: derived parameters (ParameterOutput's formulas,
reference src/parameteroutput.cpp:14-420) against numpy in float64, and the
CRAVA binary grid format written and read back through the C ABI, its bytes
equal to the header plus the grid's big-endian floats."""
import numpy as np
import pytest

from crava_amd import _lib as L

pytestmark = pytest.mark.gpu


def _grids(ctx, shape=(24, 18, 20), seed=0):
    from crava_amd.hip import FFTGrid
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    vals = [np.log(3000.0) + 0.1 * rng.standard_normal((nz, ny, 2 * (nx // 2 + 1))),
            np.log(1500.0) + 0.1 * rng.standard_normal((nz, ny, 2 * (nx // 2 + 1))),
            np.log(2300.0) + 0.05 * rng.standard_normal((nz, ny, 2 * (nx // 2 + 1)))]
    vals = [v.astype(np.float32) for v in vals]
    gs = []
    for v in vals:
        g = FFTGrid(ctx, nx, ny, nz)
        g.upload_real(v)
        gs.append(g)
    return gs, vals


def _formula(kind, A, B, R):
    e = np.exp
    return {
        L.CRV_OUT_EXP: e(A),
        L.CRV_OUT_AI: e(A + R),
        L.CRV_OUT_SI: e(B + R),
        L.CRV_OUT_VPVS: e(A - B),
        L.CRV_OUT_POISSON: 0.5 * (e(2 * (A - B)) - 2) / (e(2 * (A - B)) - 1),
        L.CRV_OUT_LAME_MU: e(R + 2 * B - 13.81551),
        L.CRV_OUT_LAME_LAMBDA: e(R) * (e(2 * A - 13.81551) - 2 * e(2 * B - 13.81551)),
        L.CRV_OUT_MU_RHO: e(2 * (B + R) - 13.81551),
        L.CRV_OUT_LAMBDA_RHO: e(2 * (A + R) - 13.81551) - 2 * e(2 * (B + R) - 13.81551),
    }[kind]


@pytest.mark.parametrize("kind", list(range(9)))
def test_derived_parameters(ctx, kind):
    from crava_amd.hip import derived
    (a, b, r), (A, B, R) = _grids(ctx)
    out = derived(ctx, kind, a, b, r)
    got = out.download_real()
    want = _formula(kind, A.astype(np.float64), B.astype(np.float64), R.astype(np.float64)).astype(np.float32)
    # double evaluation rounded to float: equal up to the last float bit
    assert np.max(np.abs(got - want) / np.abs(want)) <= 1.2e-7


def test_exp_keeps_rmissing_in_place(ctx):
    from crava_amd.hip import derived
    (a, _, _), (A, _, _) = _grids(ctx)
    A = A.copy()
    A[0, 0, :5] = -99999.0
    a.upload_real(A)
    derived(ctx, L.CRV_OUT_EXP, a, out=a)
    got = a.download_real()
    assert np.all(got[0, 0, :5] == np.float32(-99999.0))
    assert np.allclose(got[1], np.exp(A[1].astype(np.float64)), rtol=2e-7)


def test_crava_file_round_trip(ctx, tmp_path):
    from crava_amd.hip import FFTGrid, crava_header, read_crava, write_crava
    (a, _, _), (A, _, _) = _grids(ctx, shape=(450, 64, 100))   # 11.6 MB: one chunk
    big = FFTGrid(ctx, 450, 450, 96)                          # 78 MB: five 16 MB chunks
    Xb = np.random.default_rng(3).standard_normal((96, 450, 452)).astype(np.float32)
    big.upload_real(Xb)
    sb = L.SimboxInfo(x0=1.0, y0=2.0, dx=25.0, dy=25.0, nx=400, ny=60, il0=1.0, xl0=1.0, il_step_x=1.0,
                      il_step_y=0.0, xl_step_x=0.0, xl_step_y=1.0, angle=0.1)
    for g, X in ((a, A), (big, Xb)):
        p = tmp_path / "grid.crava"
        write_crava(g, p, sb)
        raw = p.read_bytes()
        head = crava_header(sb, g.rnxp, g.nyp, g.nzp)
        assert raw[:len(head)] == head
        assert raw[len(head):] == X.astype(">f4").tobytes()
        h = FFTGrid(ctx, g.nxp, g.nyp, g.nzp)
        sb2 = read_crava(h, p)
        assert np.array_equal(h.download_real(), X)
        assert (sb2.nx, sb2.ny, sb2.angle) == (400, 60, 0.1)
    # dimensions are checked
    wrong = FFTGrid(ctx, 450, 64, 98)
    with pytest.raises(Exception):
        read_crava(wrong, tmp_path / "grid.crava")


@pytest.mark.parametrize("ascii,padding,scientific", [(False, False, False), (False, True, False),
                                                      (True, False, False), (True, True, True)])
def test_storm_file(ctx, tmp_path, ascii, padding, scientific):
    """FFTGrid::writeStormFile (reference src/fftgrid.cpp:2117-2176): the
    header, the grid (padded or not) x fastest as big-endian floats, or as
    text (fixed with 6 decimals, or scientific) with one line per x row,
    then the line "0"."""
    from crava_amd.hip import FFTGrid, storm_header, write_storm
    nx, ny, nz, nxp, nyp, nzp = 12, 10, 7, 15, 12, 9
    g = FFTGrid(ctx, nxp, nyp, nzp, cubetype=L.CRV_PARAMETER, nx=nx, ny=ny, nz=nz)
    X = (np.random.default_rng(5).standard_normal((nzp, nyp, g.rnxp)) * 100).astype(np.float32)
    g.upload_real(X)
    si = L.StormInfo(xmin=10.0, lx=300.0, ymin=-20.0, ly=250.0, lz=28.0, angle=0.0, top_name=b"top",
                     bot_name=b"base", flat=0)
    p = tmp_path / ("g.dat" if ascii else "g.storm")
    write_storm(g, p, si, ascii=ascii, padding=padding, scientific=scientific)
    NX, NY, NZ = (nxp, nyp, nzp) if padding else (nx, ny, nz)
    raw = p.read_bytes()
    head = storm_header(si, L.CRV_PARAMETER, NX, NY, NZ, ascii)
    assert raw[:len(head)] == head
    body = raw[len(head):]
    vals = X[:NZ, :NY, :NX]
    if ascii:
        fmt = "%e" if scientific else "%.6f"
        want = "".join(" ".join(fmt % v for v in row) + "\n" for row in vals.reshape(-1, NX).astype(np.float64))
        assert body.decode() == want + "0\n"
    else:
        assert body[-2:] == b"0\n"
        assert body[:-2] == vals.astype(">f4").tobytes()
