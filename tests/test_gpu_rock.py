"""crv_rock_forward (DEM dry frame + Batzle-Wang fluids + Gassmann, one
thread per sample in FP64) against the C oracle of the same rplib chain."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _samples(n, seed):
    rng = np.random.default_rng(seed)
    return np.column_stack([rng.uniform(0.05, 0.35, n), rng.beta(2, 5, n), rng.uniform(0, 1, n),
                            10 ** rng.uniform(-2, 0, n), rng.uniform(10, 40, n)])


@pytest.mark.parametrize("out_fp32", [False, True])
def test_rock_forward_vs_oracle(ctx, out_fp32):
    from crava_amd.hip import ROCK_DEFAULTS, rock_forward
    x = _samples(4000, 11)
    got, st = rock_forward(ctx, x, ROCK_DEFAULTS, out_fp32=out_fp32)
    want, st_o = O.rock_forward(x, ROCK_DEFAULTS)
    assert np.array_equal(st, st_o)
    ok = st == 0
    assert ok.mean() > 0.99
    # float outputs: one rounding of values near 8 (ulp 9.5e-7)
    tol = 1e-6 if out_fp32 else 1e-11
    assert np.max(np.abs(got[ok] - want[ok])) < tol
    # vs < vp < 6.5 km/s and rho between the lightest (gas-filled, 35 %
    # porosity) and densest (tight) rock; with crack-like pores at high
    # porosity the DEM frame nearly loses its rigidity, so vs has no floor
    vp, vs, rho = np.exp(want[ok]).T
    assert np.all((vp < 6500) & (vs < vp) & (rho > 1700) & (rho < 2700))


def test_rock_forward_edge_inputs(ctx):
    """Endpoints of the input box: tiny and unit aspect ratio, dry and fully
    saturated pores, lowest and highest porosity and pressure."""
    from crava_amd.hip import ROCK_DEFAULTS, rock_forward
    x = np.array([[0.05, 0.0, 0.0, 0.01, 10.0], [0.35, 0.9, 1.0, 1.0, 40.0], [0.2, 0.3, 0.5, 0.01, 40.0],
                  [0.35, 0.0, 1.0, 0.01, 10.0], [0.05, 1.0, 0.0, 1.0, 25.0]])
    got, st = rock_forward(ctx, x, ROCK_DEFAULTS)
    want, st_o = O.rock_forward(x, ROCK_DEFAULTS)
    assert np.array_equal(st, st_o)
    ok = st == 0
    assert np.max(np.abs(got[ok] - want[ok])) < 1e-11
