"""The uneven-volume batch: the 12 inversions of model.config3_volumes()
(nx, ny, nz drawn from {64..320} x {100..400}, padded to 2^a 3^b 5^c 7^d by
the reference's rule; radix-7 and odd lengths among them) captured in one HIP
graph equal the same crv_invert calls made eagerly, bit for bit, on every
replay; and on two of the sizes the one-call inversion matches the FP64
oracle (posterior means, covariances and the realization)."""
import numpy as np
import pytest

from conftest import rel_rms
from crava_amd import model
from oracle import oracle as O
from pipeline import Case, oracle_posterior

pytestmark = pytest.mark.gpu


def _volume(ctx, v):
    from crava_amd import _lib as L
    from crava_amd import hip
    from crava_amd.survey import GpuSurvey
    sv = GpuSurvey(ctx, v, truth_noise="counter")
    mk = lambda ct: hip.FFTGrid(ctx, v.nxp, v.nyp, v.nzp, precision=32, cubetype=ct, nx=v.nx, ny=v.ny, nz=v.nz)
    ins = dict(mean=[mk(L.CRV_PARAMETER) for _ in range(3)], data=[mk(L.CRV_DATA) for _ in range(v.n)])
    for g, src in zip(ins["mean"], sv.mean):
        g.copyFrom(src)
    for g, src in zip(ins["data"], sv.data):
        g.copyFrom(src)
    plan = hip.PostPlan(ctx, v.nxp, v.nyp, v.nzp, 32, v.n, mode=L.CRV_MODE_FP32_REFINE, separable=True,
                        skip_mean_inverse=True, **sv.posterior_kwargs())
    sv.close()
    return ins, plan


def _outputs(ctx, v):
    from crava_amd import _lib as L
    from crava_amd import hip
    mk = lambda ct: hip.FFTGrid(ctx, v.nxp, v.nyp, v.nzp, precision=32, cubetype=ct, nx=v.nx, ny=v.ny, nz=v.nz)
    return dict(mean=[mk(L.CRV_PARAMETER) for _ in range(3)], data=[mk(L.CRV_DATA) for _ in range(v.n)],
                cov=[mk(L.CRV_COVARIANCE) for _ in range(6)],
                sims=hip.grid_batch(ctx, 3, v.nxp, v.nyp, v.nzp, 32, L.CRV_PARAMETER, v.nx, v.ny, v.nz))


def _read(o):
    """Every output buffer of one volume as stored (FP32): real means,
    covariances and realization, residual spectra (interleaved re, im)."""
    return [g.download_real() for g in o["mean"] + o["cov"] + o["sims"] + o["data"]]


def test_twelve_volumes_in_one_graph_equal_eager_calls(ctx):
    from crava_amd.hip import FFTGrid, Graph
    vols = model.config3_volumes()
    assert len(vols) == 12 and len({(v.nxp, v.nyp, v.nzp) for v in vols}) >= 10
    work = [(v, *_volume(ctx, v)) for v in vols]

    def run(outs):
        for (v, ins, plan), o in zip(work, outs):
            plan.invert(o["mean"], o["data"], o["cov"], o["sims"], mean_src=ins["mean"], data_src=ins["data"],
                        seed=v.seed_sim, realization0=0)

    eager = [_outputs(ctx, v) for v in vols]
    run(eager)
    ctx.synchronize()
    want = [_read(o) for o in eager]
    for o in eager:
        for g in sum(o.values(), []):
            g.close()
    cap = [_outputs(ctx, v) for v in vols]
    graph = Graph.begin(ctx)
    run(cap)
    graph.end()
    for rep in range(2):
        junk = FFTGrid(ctx, 120, 150, 250, cubetype=2)
        junk.upload_real(np.full((250, 150, 122), float(rep + 1), np.float32))
        junk.fftInPlace()
        ctx.synchronize()
        graph.launch()
        ctx.synchronize()
        for v, o, w in zip(vols, cap, want):
            got = _read(o)
            for a, b in zip(got, w):
                assert np.array_equal(a, b), f"volume {v.nxp}x{v.nyp}x{v.nzp} replay {rep}"
        junk.close()
    graph.close()
    for (v, ins, plan), o in zip(work, cap):
        plan.close()
        for g in sum(o.values(), []) + ins["mean"] + ins["data"]:
            g.close()


@pytest.mark.parametrize("idx", [2, 5])
def test_uneven_volume_invert_vs_fp64_oracle(ctx, idx):
    """Volumes 84x84x250 and 150x84x210 (radix-7 and odd-prime-power
    lengths) with their own prior standard deviations: the one-call FP32
    refine inversion against the FP64 oracle with the same separable prior,
    and its realization against the oracle's FP64 realization of the same
    counter noise."""
    from crava_amd import _lib as L
    from crava_amd import hip
    v = model.config3_volumes()[idx]
    case = Case(v)
    s, p = case.s, case.pin
    o = oracle_posterior(case, np.float64, separable=True)
    cxy, ct = model.prior_spectra(s)
    plan = hip.PostPlan(ctx, s.nxp, s.nyp, s.nzp, 32, s.n, mode=L.CRV_MODE_FP32_REFINE, separable=True,
                        skip_mean_inverse=True, A=p.A, theta=p.theta, tables=p.tables, nz=s.nz, lz=s.lz,
                        sigma0=p.sigma0, cxy=cxy, ct=ct)
    out = _outputs(ctx, s)
    for g, a in zip(out["mean"], case.syn.mean):
        g.upload_real(a)
    for g, a in zip(out["data"], case.syn.data):
        g.upload_real(a)
    seed, real0 = 7, 0
    plan.invert(out["mean"], out["data"], out["cov"], out["sims"], seed=seed, realization0=real0)
    nx = s.nxp
    for a, (g, m) in enumerate(zip(out["mean"], o["mean"])):
        mu = s.mu[a]
        assert rel_rms(g.download_real()[:, :, :nx] - mu, m[:, :, :nx] - mu) < 1e-4
    for g, c in zip(out["cov"], o["cov"]):
        want = O.fft_c2r(c, nx, 3)[:, :, :nx]
        assert rel_rms(g.download_real()[:, :, :nx], want) < 1e-4
    for g, d in zip(out["data"], o["data"]):
        assert rel_rms(g.download_complex(), d) < 1e-4
    z = [O.counter_noise(s.nxp, s.nyp, s.nzp, seed, real0, gi) for gi in range(3)]
    sc, _ = O.simulate_scale(o["cov"], z, dtype=np.float64)
    for a in range(3):
        want = O.fft_c2r(sc[a], nx, 2)[:, :, :nx] + o["mean"][a][:, :, :nx]
        mu = s.mu[a]
        assert rel_rms(out["sims"][a].download_real()[:, :, :nx] - mu, want - mu) < 1e-4
    plan.close()
