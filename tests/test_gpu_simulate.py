"""Crava::simulate on the GPU: the reference noise stream (RandomGen +
fillInComplexNoise) decoded in parallel against the reference's own stream,
the Cholesky scaling and inverse transforms against the oracle, and the
counter-based stream against its FP64 restatement."""
import math

import numpy as np
import pytest

from conftest import cplx, rel_rms
from crava_amd import model
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(8, 6, 10), (10, 7, 12), (9, 4, 6)])
def test_noise_fill_matches_reference(ctx, golden, shape):
    from crava_amd.hip import FFTGrid
    nx, ny, nz = shape
    key = f"noise_{nx}x{ny}x{nz}"
    g = FFTGrid(ctx, nx, ny, nz, precision=32)
    st = g.fillInComplexNoise(int(golden[key + "_seed"]))
    assert st == int(golden[key + "_state"])
    assert g.getIsTransformed() and g.getType() == FFTGrid.PARAMETER
    assert np.array_equal(g.download_complex().astype(np.complex64), cplx(golden[key]).astype(np.complex64))


def test_noise_fill_c0_matches_reference(ctx, golden):
    from crava_amd.hip import FFTGrid
    g = FFTGrid(ctx, 64, 64, 128, precision=32)
    st = g.fillInComplexNoise(0)
    assert st == int(golden["noise_c0_state"])
    c = g.download_complex()
    flat = np.stack([c.real, c.imag], -1).reshape(-1, 2).astype(np.float32)
    assert np.array_equal(flat[golden["noise_c0_idx"]], golden["noise_c0_val"])


@pytest.mark.parametrize("shape,seed", [((256, 256, 200), 7), ((450, 30, 64), 231171), ((35, 21, 9), 4)])
def test_noise_fill_long_stream(ctx, shape, seed):
    """Millions of draws: every value equal to the oracle's sequential stream
    and the final LCG state identical (the decode is exact)."""
    from crava_amd.hip import FFTGrid
    nx, ny, nz = shape
    g = FFTGrid(ctx, nx, ny, nz, precision=32)
    st = g.fillInComplexNoise(seed)
    ref, st_ref = O.fill_complex_noise(nx, ny, nz, seed)
    assert st == st_ref
    got = g.download_complex()
    mism = np.count_nonzero(got.astype(np.complex64) != ref.astype(np.complex64))
    print(f"noise {shape} seed {seed}: {mism} of {ref.size} values differ")
    assert mism == 0


@pytest.mark.parametrize("force", [False, True])
def test_noise_guard_host_resolution(ctx, force):
    """The decode's guard sends normals whose decisions or float roundings
    lie near where the device and host libm could differ to the host.
    With the margins widened 1e9 times hundreds of normals take that path:
    their values are patched from the host, or (force) the decode runs
    again with them in its table.  Either way the stream equals the
    oracle's sequential one and ends in the same state."""
    from crava_amd.hip import FFTGrid, noise_guard
    nx, ny, nz, seed = 64, 48, 40, 99
    ref, st_ref = O.fill_complex_noise(nx, ny, nz, seed)
    try:
        noise_guard(ctx, 1e9, force)
        g = FFTGrid(ctx, nx, ny, nz, precision=32)
        st = g.fillInComplexNoise(seed)
        passes, host, patched, runs = noise_guard(ctx)
    finally:
        noise_guard(ctx, 1.0, False)
    print(f"guard x1e9 force={force}: {passes} fix passes, {host} host normals, {patched} patched, {runs} runs")
    assert st == st_ref
    assert np.array_equal(g.download_complex().astype(np.complex64), ref.astype(np.complex64))
    assert host > 100
    if force:
        assert runs == 2 and patched == 0
    else:
        assert patched > 50
    g2 = FFTGrid(ctx, nx, ny, nz, precision=32)
    assert g2.fillInComplexNoise(seed) == st_ref
    assert noise_guard(ctx)[1] == 0


def test_noise_fill_c1_nyquist_plane(ctx):
    """The reference stream at the 450x450x320 grid (cnxp 226, even): every
    point of the x = 225 plane is "special" under the global-parity rule
    (fftgrid.cpp:1201), so the plane is exactly Hermitian and the fill draws
    N normals: the final LCG state equals the oracle stream's after N draws."""
    from crava_amd.hip import FFTGrid
    nx, ny, nz = 450, 450, 320
    g = FFTGrid(ctx, nx, ny, nz, precision=32)
    st = g.fillInComplexNoise(231171)
    _, st_ref = O.rnorm_seq(231171, nx * ny * nz)
    assert st == st_ref
    c = g.download_complex()
    p = c[:, :, -1]
    pc = np.conj(p[(-np.arange(nz)) % nz][:, (-np.arange(ny)) % ny])
    assert np.array_equal(p, pc)
    p0 = c[:, :, 0]
    assert np.array_equal(p0, np.conj(p0[(-np.arange(nz)) % nz][:, (-np.arange(ny)) % ny]))


def _sim_case():
    s = model.config_small(16, 12, 20, 3)
    from pipeline import Case, oracle_posterior
    case = Case(s)
    return case, oracle_posterior(case, np.float64)


def test_simulate_reference_noise_vs_oracle(ctx):
    """One realization with the reference stream from seed 0: seed0, seed1,
    seed2 filled in order, scaled by chol(P), inverse transformed, plus the
    posterior mean."""
    from crava_amd.hip import FFTGrid, simulate
    from pipeline import gpu_posterior
    case, o = _sim_case()
    s = case.s
    r = gpu_posterior(ctx, case, prec=64)
    out = [FFTGrid(ctx, s.nxp, s.nyp, s.nzp, precision=64) for _ in range(3)]
    st = simulate(ctx, r["cov"], r["mean"], out, noise_mode=0, lcg_state=0)
    # oracle
    seeds, state = [], 0
    for _ in range(3):
        z, state = O.fill_complex_noise(s.nxp, s.nyp, s.nzp, state, np.float64)
        seeds.append(z)
    assert st == state
    sc, nf = O.simulate_scale(o["cov"], seeds, dtype=np.float64)
    for a in range(3):
        want = O.fft_c2r(sc[a], s.nxp, 2)[:, :, : s.nxp] + o["mean"][a][:, :, : s.nxp]
        got = out[a].download_real()[:, :, : s.nxp]
        mu = s.mu[a]
        assert rel_rms(got - mu, want - mu) < 1e-11


def test_simulate_host_noise_fp32_vs_reference_scaling(ctx, golden):
    """Host-supplied noise through the FP32 kernel against the reference's
    compiled Cholesky scaling (golden), then the inverse transform."""
    from crava_amd.hip import FFTGrid, simulate
    nx, ny, nz = 12, 10, 16
    cov = [FFTGrid(ctx, nx, ny, nz, precision=32, cubetype=3) for _ in range(6)]
    for g, a in zip(cov, golden["post_out_cov"]):
        g.upload_complex(cplx(a))
    out = [FFTGrid(ctx, nx, ny, nz, precision=32) for _ in range(3)]
    seeds = [cplx(a) for a in golden["sim_in_seed"]]
    simulate(ctx, cov, None, out, noise_mode=2, host_noise=seeds)
    for a in range(3):
        want = O.fft_c2r(cplx(golden["sim_out"][a]), nx, 2)[:, :, :nx]
        assert rel_rms(out[a].download_real()[:, :, :nx], want) < 2e-6


@pytest.mark.parametrize("prec", [32, 64])
def test_counter_noise_vs_fp64_restatement(ctx, prec):
    from crava_amd.hip import FFTGrid, simulate
    nx, ny, nz = 18, 10, 14
    # identity covariance: the realization is the inverse transform of the noise
    cov = [FFTGrid(ctx, nx, ny, nz, precision=prec, cubetype=3) for _ in range(6)]
    cn = nx // 2 + 1
    for q, g in enumerate(cov):
        g.upload_complex(np.full((nz, ny, cn), 1.0 if q < 3 else 0.0, complex))
    out = [FFTGrid(ctx, nx, ny, nz, precision=prec) for _ in range(3)]
    simulate(ctx, cov, None, out, noise_mode=1, seed=123456789, realization=3)
    for g in range(3):
        z = O.counter_noise(nx, ny, nz, 123456789, 3, g)
        want = O.fft_c2r(z, nx, 2)[:, :, :nx]
        tol = 1e-6 if prec == 32 else 1e-12
        assert rel_rms(out[g].download_real()[:, :, :nx], want) < tol


def test_simulation_covariance(ctx):
    """Empirical covariance over realizations matches the posterior: with a
    diagonal spectrum P(k) = p, each real voxel has variance mean_k p(k)."""
    from crava_amd.hip import FFTGrid, simulate
    nx, ny, nz = 10, 8, 8
    cn = nx // 2 + 1
    P = np.array([[2.0, 0.6, 0.3], [0.6, 1.0, 0.2], [0.3, 0.2, 0.5]])
    cov = [FFTGrid(ctx, nx, ny, nz, precision=64, cubetype=3) for _ in range(6)]
    for q, (a, b) in enumerate(((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))):
        cov[q].upload_complex(np.full((nz, ny, cn), P[a, b], complex))
    out = [FFTGrid(ctx, nx, ny, nz, precision=64) for _ in range(3)]
    acc = np.zeros((3, 3))
    nreal = 400
    for r in range(nreal):
        simulate(ctx, cov, None, out, noise_mode=1, seed=5, realization=r)
        v = np.stack([o.download_real()[:, :, :nx].ravel() for o in out])
        acc += v @ v.T / v.shape[1]
    emp = acc / nreal
    assert np.allclose(emp, P, atol=0.05 * P.max())
