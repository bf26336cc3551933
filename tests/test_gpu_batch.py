"""Batched and graph-captured execution (the many-realization / many-volume
workload): the plan/exec posterior, the batched simulation and a captured
graph reproduce the one-call entry points exactly, and the autotuner's
shapes give the same transforms."""
import numpy as np
import pytest

from conftest import rel_rms
from crava_amd import model
from pipeline import Case

pytestmark = pytest.mark.gpu


def _inputs(ctx, case, prec=32):
    """Transformed means/data of a small case plus separable prior keywords."""
    from crava_amd.hip import FFTGrid
    from oracle import oracle as O
    s = case.s
    mk = lambda ct: FFTGrid(ctx, s.nxp, s.nyp, s.nzp, precision=prec, cubetype=ct)
    mean = [mk(2) for _ in range(3)]
    data = [mk(1) for _ in range(s.n)]
    for g, a in zip(mean, case.syn.mean):
        g.upload_complex(O.fft_r2c(a, s.nxp, 2))
    for g, a in zip(data, case.syn.data):
        g.upload_complex(O.fft_r2c(a, s.nxp, 1))
    p = case.pin
    cxy, ct = model.prior_spectra(s)
    kw = dict(A=p.A, theta=p.theta, tables=p.tables, nz=s.nz, lz=s.lz, sigma0=p.sigma0, separable=True, cxy=cxy,
              ct=ct, skip_mean_inverse=True)
    return mean, data, kw


def _copy(ctx, grids):
    from crava_amd.hip import FFTGrid
    out = []
    for g in grids:
        c = FFTGrid(ctx, g.nxp, g.nyp, g.nzp, precision=g.precision, cubetype=g.getType())
        c.copyFrom(g)
        out.append(c)
    return out


@pytest.mark.parametrize("mode", [0, 1])
def test_posterior_exec_matches_one_call(ctx, mode):
    from crava_amd.hip import FFTGrid, PostPlan, computePostMeanResidAndFFTCov
    case = Case(model.config_small(20, 18, 30, 5))
    s = case.s
    mean, data, kw = _inputs(ctx, case)
    mean2, data2 = _copy(ctx, mean), _copy(ctx, data)
    cov = [FFTGrid(ctx, s.nxp, s.nyp, s.nzp, cubetype=3) for _ in range(6)]
    cov2 = [FFTGrid(ctx, s.nxp, s.nyp, s.nzp, cubetype=3) for _ in range(6)]
    st = computePostMeanResidAndFFTCov(ctx, mean, data, None, cov, mode=mode, **kw)
    plan = PostPlan(ctx, s.nxp, s.nyp, s.nzp, 32, s.n, mode=mode, **kw)
    plan.exec(mean2, data2, None, cov2)
    st2 = plan.stats()
    for a, b in zip(mean + data + cov, mean2 + data2 + cov2):
        assert np.array_equal(a.download_complex(), b.download_complex())
    assert (st.n_inverted, st.n_chol_fail, st.n_fp64) == (st2.n_inverted, st2.n_chol_fail, st2.n_fp64)


def _posterior_cov(ctx, case):
    from crava_amd.hip import FFTGrid, computePostMeanResidAndFFTCov
    s = case.s
    mean, data, kw = _inputs(ctx, case)
    cov = [FFTGrid(ctx, s.nxp, s.nyp, s.nzp, cubetype=3) for _ in range(6)]
    kw["skip_mean_inverse"] = False
    computePostMeanResidAndFFTCov(ctx, mean, data, None, cov, mode=1, **kw)
    return mean, cov


@pytest.fixture
def batch_mb(ctx, request):
    """CRV_TUNE_BATCH_MB for one test, restored afterwards (1 MB splits the
    batches of these small grids into sub-batches)."""
    from crava_amd import _lib as L
    ctx.tune(L.CRV_TUNE_BATCH_MB, request.param)
    yield request.param
    ctx.tune(L.CRV_TUNE_BATCH_MB, 0)


@pytest.mark.parametrize("batch_mb", [0, 1], indirect=True)
def test_batched_transforms_match_single_grids(ctx, batch_mb):
    from crava_amd import _lib as L
    from crava_amd.hip import FFTGrid, fft_c2r_batch, fft_r2c_batch, grid_batch
    nx, ny, nz, cnt = 64, 48, 40, 7
    rng = np.random.default_rng(2)
    xs = [rng.standard_normal((nz, ny, 2 * (nx // 2 + 1))).astype(np.float32) for _ in range(cnt)]
    src = grid_batch(ctx, cnt, nx, ny, nz, 32, L.CRV_DATA)
    dst = grid_batch(ctx, cnt, nx, ny, nz, 32, L.CRV_DATA)
    inp = grid_batch(ctx, cnt, nx, ny, nz, 32, L.CRV_DATA)
    for g, h, x in zip(src, inp, xs):
        g.upload_real(x)
        h.upload_real(x)
    fft_r2c_batch(dst, src=src)
    fft_r2c_batch(inp)
    ctx.synchronize()
    one = []
    for x, g, d, h in zip(xs, src, dst, inp):
        a = FFTGrid(ctx, nx, ny, nz, cubetype=L.CRV_DATA)
        a.upload_real(x)
        a.fftInPlace()
        want = a.download_complex()
        assert np.array_equal(d.download_complex(), want)
        assert np.array_equal(h.download_complex(), want)
        assert np.array_equal(g.download_real()[:, :, :nx], x[:, :, :nx])   # source untouched
        one.append(a)
    fft_c2r_batch(dst)
    for a, d in zip(one, dst):
        a.invFFTInPlace()
        assert np.array_equal(d.download_real()[:, :, :nx], a.download_real()[:, :, :nx])


@pytest.mark.parametrize("batch_mb", [0, 1], indirect=True)
@pytest.mark.parametrize("fp64_chol", [0, 1, 2])
def test_simulate_batch_matches_single_calls(ctx, fp64_chol, batch_mb):
    from crava_amd import _lib as L
    from crava_amd.hip import FFTGrid, grid_batch, simulate, simulate_batch
    case = Case(model.config_small(64, 48, 40, 5))
    s = case.s
    mean, cov = _posterior_cov(ctx, case)
    nreal = 4
    out = grid_batch(ctx, 3 * nreal, s.nxp, s.nyp, s.nzp)
    simulate_batch(ctx, cov, mean, out, seed=99, realization0=10, fp64_chol=fp64_chol)
    for r in range(nreal):
        one = [FFTGrid(ctx, s.nxp, s.nyp, s.nzp) for _ in range(3)]
        simulate(ctx, cov, mean, one, noise_mode=L.CRV_NOISE_COUNTER, seed=99, realization=10 + r,
                 fp64_chol=fp64_chol)
        for g in range(3):
            assert np.array_equal(out[3 * r + g].download_real()[:, :, :s.nxp], one[g].download_real()[:, :, :s.nxp])
    # different realizations differ
    assert not np.array_equal(out[0].download_real(), out[3].download_real())


def test_graph_replays_match_eager(ctx):
    """Capture forward transforms, the posterior exec, a batch of
    realizations and the covariance inverse transforms; every replay (with
    eager work and synchronisations between replays) equals the eager run."""
    from crava_amd.hip import FFTGrid, Graph, PostPlan, grid_batch, simulate_batch
    case = Case(model.config_small(20, 18, 30, 5))
    s = case.s
    mk = lambda ct: FFTGrid(ctx, s.nxp, s.nyp, s.nzp, cubetype=ct)
    mean_in = [mk(2) for _ in range(3)]
    data_in = [mk(1) for _ in range(s.n)]
    for g, a in zip(mean_in, case.syn.mean):
        g.upload_real(a)
    for g, a in zip(data_in, case.syn.data):
        g.upload_real(a)
    _, _, kw = _inputs(ctx, case)

    def pipeline_grids():
        return ([mk(2) for _ in range(3)], [mk(1) for _ in range(s.n)], [mk(3) for _ in range(6)],
                grid_batch(ctx, 6, s.nxp, s.nyp, s.nzp))

    plan = PostPlan(ctx, s.nxp, s.nyp, s.nzp, 32, s.n, mode=1, **kw)

    def run(mean, data, cov, sims):
        for w, src in zip(data, data_in):
            w.fftFrom(src)
        for w, src in zip(mean, mean_in):
            w.fftFrom(src)
        plan.exec(mean, data, None, cov)
        simulate_batch(ctx, cov, None, sims, seed=5, realization0=0, fp64_chol=True)
        for c in cov:
            c.invFFTInPlace()

    eager = pipeline_grids()
    run(*eager)
    ctx.synchronize()
    want = [g.download_complex() if i < 3 + s.n else g.download_real()
            for i, g in enumerate(eager[0] + eager[1] + eager[2] + eager[3])]
    cap = pipeline_grids()
    graph = Graph.begin(ctx)
    run(*cap)
    graph.end()
    for rep in range(3):
        # eager work on other grids between replays
        junk = mk(2)
        junk.upload_real(np.full((s.nzp, s.nyp, s.rnxp), float(rep), np.float32))
        junk.fftInPlace()
        ctx.synchronize()
        graph.launch()
        ctx.synchronize()
        got = [g.download_complex() if i < 3 + s.n else g.download_real()
               for i, g in enumerate(cap[0] + cap[1] + cap[2] + cap[3])]
        for a, b in zip(got, want):
            assert np.array_equal(a, b)
    graph.close()


def test_invert_fp64_matches_sequence(ctx):
    """crv_invert on an FP64 plan (no per-plane records: the general
    posterior kernel, the simulation kernel, every covariance plane written)
    equals the FP64 one-call sequence bit for bit, realizations included."""
    from crava_amd.hip import FFTGrid, PostPlan, grid_batch, simulate_batch
    case = Case(model.config_small(32, 24, 20, 3))
    s = case.s
    _, _, kw = _inputs(ctx, case)
    plan = PostPlan(ctx, s.nxp, s.nyp, s.nzp, 64, s.n, mode=0, **kw)
    batch = lambda n, ct: grid_batch(ctx, n, s.nxp, s.nyp, s.nzp, 64, ct)

    def inputs():
        mean, data = batch(3, 2), batch(s.n, 1)
        for g, a in zip(mean, case.syn.mean):
            g.upload_real(a)
        for g, a in zip(data, case.syn.data):
            g.upload_real(a)
        return mean, data

    mean, data = inputs()
    cov, sims = batch(6, 3), batch(3, 2)
    for g in data + mean:
        g.fftInPlace()
    plan.exec(mean, data, None, cov)
    for g in mean:
        g.invFFTInPlace()
    simulate_batch(ctx, cov, mean, sims, seed=4, realization0=1, fp64_chol=False)
    for g in cov:
        g.invFFTInPlace()
    ctx.synchronize()
    want = ([g.download_real() for g in mean + cov + sims], [g.download_complex() for g in data])
    mean2, data2 = inputs()
    cov2, sims2 = batch(6, 3), batch(3, 2)
    plan.invert(mean2, data2, cov2, sims2, seed=4, realization0=1)
    got = ([g.download_real() for g in mean2 + cov2 + sims2], [g.download_complex() for g in data2])
    for w, g in zip(want[0], got[0]):
        assert np.array_equal(w[:, :, :s.nxp], g[:, :, :s.nxp])
    for w, g in zip(want[1], got[1]):
        assert np.array_equal(w, g)


def test_invert_graph_replays_match_eager(ctx):
    """crv_invert (one call: transforms, eigen-basis posterior with the
    realizations fused, generated prior planes, inverse transforms) captured
    in a graph: every replay, with eager work and synchronisations between
    replays, equals the eager call bit for bit."""
    from crava_amd.hip import FFTGrid, Graph, PostPlan, grid_batch
    case = Case(model.config_small(20, 18, 30, 5))
    s = case.s
    _, _, kw = _inputs(ctx, case)
    plan = PostPlan(ctx, s.nxp, s.nyp, s.nzp, 32, s.n, mode=1, **kw)
    batch = lambda n, ct: grid_batch(ctx, n, s.nxp, s.nyp, s.nzp, 32, ct)
    mean_in, data_in = batch(3, 2), batch(s.n, 1)
    for g, a in zip(mean_in, case.syn.mean):
        g.upload_real(a)
    for g, a in zip(data_in, case.syn.data):
        g.upload_real(a)

    def grids():
        return batch(3, 2), batch(s.n, 1), batch(6, 3), batch(6, 2)

    def run(mean, data, cov, sims):
        plan.invert(mean, data, cov, sims, mean_src=mean_in, data_src=data_in, seed=9, realization0=2)

    def read(mean, data, cov, sims):
        return ([g.download_real() for g in mean] + [g.download_complex() for g in data]
                + [g.download_real() for g in cov + sims])

    eager = grids()
    run(*eager)
    ctx.synchronize()
    want = read(*eager)
    cap = grids()
    graph = Graph.begin(ctx)
    run(*cap)
    graph.end()
    for rep in range(3):
        junk = FFTGrid(ctx, s.nxp, s.nyp, s.nzp, cubetype=2)
        junk.upload_real(np.full((s.nzp, s.nyp, s.rnxp), float(rep), np.float32))
        junk.fftInPlace()
        ctx.synchronize()
        graph.launch()
        ctx.synchronize()
        for a, b in zip(read(*cap), want):
            assert np.array_equal(a, b)
    graph.close()


@pytest.mark.parametrize("dims", [(80, 60, 50), (450, 450, 320)])
def test_autotuned_shapes_keep_the_transform(ctx, dims):
    from crava_amd.hip import FFTGrid, autotune
    nx, ny, nz = dims
    rng = np.random.default_rng(1)
    a = FFTGrid(ctx, nx, ny, nz)
    x = rng.standard_normal((nz, ny, 2 * (nx // 2 + 1))).astype(np.float32)
    a.upload_real(x)
    a.fftInPlace()
    before = a.download_complex()
    shape = autotune(ctx, nx, ny, nz, 32, reps=2)
    assert shape["ms_r2c"] > 0 and shape["ms_c2r"] > 0
    assert shape["rows_per_tile"] >= 1 and shape["tx_y"] >= 1 and shape["tx_z"] >= 1
    b = FFTGrid(ctx, nx, ny, nz)
    b.upload_real(x)
    b.fftInPlace()
    assert rel_rms(b.download_complex(), before) < 1e-6
    b.invFFTInPlace()
    assert rel_rms(b.download_real()[:, :, :nx], x[:, :, :nx]) < 1e-6


@pytest.mark.parametrize("chunk_rows,batched,src,n,dims",
                         [(0, False, False, 5, (64, 48, 40)), (7, False, False, 5, (64, 48, 40)),
                          (1, True, False, 5, (64, 48, 40)), (0, True, True, 5, (64, 48, 40)),
                          (7, True, True, 5, (64, 48, 40)), (5, False, True, 5, (64, 48, 40)),
                          (0, True, False, 6, (64, 48, 40)), (3, True, False, 2, (64, 48, 40)),
                          (0, True, False, 5, (21, 18, 30)), (4, False, True, 3, (21, 18, 30))])
def test_invert_matches_sequence(ctx, chunk_rows, batched, src, n, dims):
    """crv_invert (transforms, posterior, realizations and inverse transforms,
    whole grids or pipelined by chunks of y rows) equals the one-call
    sequence bit for bit, but for the realizations of the whole-grid call:
    outside the inverted band its fused simulation scales the noise by
    sqrt(lam) times the plane's factor of the prior S, where the simulation
    kernel factors the stored FP32 product lam S (rel RMS below 1e-6)."""
    from crava_amd.hip import FFTGrid, PostPlan, fft_c2r_batch, grid_batch, simulate_batch
    case = Case(model.config_small(*dims, n))
    s = case.s
    _, _, kw = _inputs(ctx, case)
    plan = PostPlan(ctx, s.nxp, s.nyp, s.nzp, 32, s.n, mode=1, **kw)

    def grids(ct, count):
        if batched:
            return grid_batch(ctx, count, s.nxp, s.nyp, s.nzp, 32, ct)
        return [FFTGrid(ctx, s.nxp, s.nyp, s.nzp, cubetype=ct) for _ in range(count)]

    def inputs():
        mean, data = grids(2, 3), grids(1, s.n)
        for g, a in zip(mean, case.syn.mean):
            g.upload_real(a)
        for g, a in zip(data, case.syn.data):
            g.upload_real(a)
        return mean, data

    nreal = 2
    # the sequence of one-call entry points
    mean, data = inputs()
    cov = grids(3, 6)
    sims = grid_batch(ctx, 3 * nreal, s.nxp, s.nyp, s.nzp)
    for g in data + mean:
        g.fftInPlace()
    plan.exec(mean, data, None, cov)
    st_seq = plan.stats()
    for g in mean:
        g.invFFTInPlace()
    simulate_batch(ctx, cov, mean, sims, seed=17, realization0=3, fp64_chol=2)
    for g in cov:
        g.invFFTInPlace()
    ctx.synchronize()
    want = ([g.download_real() for g in mean], [g.download_complex() for g in data],
            [g.download_real() for g in cov], [g.download_real() for g in sims])
    # the pipelined call, in place or from source grids that stay unchanged
    if src:
        msrc, dsrc = inputs()
        mean2, data2 = grids(2, 3), grids(1, s.n)
    else:
        msrc = dsrc = None
        mean2, data2 = inputs()
    cov2 = grids(3, 6)
    sims2 = grid_batch(ctx, 3 * nreal, s.nxp, s.nyp, s.nzp)
    plan.invert(mean2, data2, cov2, sims2, mean_src=msrc, data_src=dsrc, seed=17, realization0=3,
                chunk_rows=chunk_rows)
    if src:
        for g, a in zip(msrc + dsrc, list(case.syn.mean) + list(case.syn.data)):
            assert np.array_equal(g.download_real()[:, :, :s.nxp], a[:, :, :s.nxp].astype(np.float32))
    st = plan.stats()
    got = ([g.download_real() for g in mean2], [g.download_complex() for g in data2],
           [g.download_real() for g in cov2], [g.download_real() for g in sims2])
    nx = s.nxp
    for w, g in zip(want[0] + want[2], got[0] + got[2]):
        assert np.array_equal(w[:, :, :nx], g[:, :, :nx])
    for w, g in zip(want[3], got[3]):
        if chunk_rows > 0:
            assert np.array_equal(w[:, :, :nx], g[:, :, :nx])
        else:
            assert 0 < rel_rms(g[:, :, :nx], w[:, :, :nx]) < 1e-6
    for w, g in zip(want[1], got[1]):
        assert np.array_equal(w, g)
    assert st.n_fp64 == st_seq.n_fp64 and st.n_chol_fail == st_seq.n_chol_fail
    assert all(not g.getIsTransformed() for g in mean2 + cov2 + sims2)
    assert all(g.getIsTransformed() for g in data2)


@pytest.mark.parametrize("n,dims", [(3, (20, 84, 30)), (5, (24, 120, 40)), (2, (18, 108, 84))])
def test_invert_band_y_covariances_match_sequence(ctx, n, dims):
    """crv_invert with a two-stage y length: the covariances' y inverse runs on
    the band's planes only (generated there) and their z inverse generates
    the planes outside the band from lam's y transform, so the inverse passes
    run in the other order than the sequence's: the covariances agree with
    the sequence to rounding (rel RMS below 1e-6); means and residuals stay
    bit-equal."""
    from crava_amd.hip import PostPlan, fft_c2r_batch, grid_batch, simulate_batch
    case = Case(model.config_phase(*dims, n))
    s = case.s
    p = case.pin
    cxy, ct = model.prior_spectra(s)
    plan = PostPlan(ctx, s.nxp, s.nyp, s.nzp, 32, s.n, mode=1, A=p.A, theta=p.theta, tables=p.tables, nz=s.nz,
                    lz=s.lz, sigma0=p.sigma0, separable=True, cxy=cxy, ct=ct)
    grids = lambda ct_, k: grid_batch(ctx, k, s.nxp, s.nyp, s.nzp, 32, ct_)

    def inputs():
        mean, data = grids(2, 3), grids(1, s.n)
        for g, a in zip(mean, case.syn.mean):
            g.upload_real(a)
        for g, a in zip(data, case.syn.data):
            g.upload_real(a)
        return mean, data

    mean, data = inputs()
    cov = grids(3, 6)
    for g in data + mean:
        g.fftInPlace()
    plan.exec(mean, data, None, cov)
    for g in mean:
        g.invFFTInPlace()
    fft_c2r_batch(cov)
    ctx.synchronize()
    nx = s.nxp
    want = ([g.download_real()[:, :, :nx] for g in mean], [g.download_complex() for g in data],
            [g.download_real()[:, :, :nx] for g in cov])
    mean2, data2 = inputs()
    cov2 = grids(3, 6)
    ctx.profile(True)
    plan.invert(mean2, data2, cov2, grids(2, 3), seed=3)
    ctx.synchronize()
    rep = ctx.profile_report()
    ctx.profile(False)
    assert "fft_y_inv_band" in rep
    got = ([g.download_real()[:, :, :nx] for g in mean2], [g.download_complex() for g in data2],
           [g.download_real()[:, :, :nx] for g in cov2])
    for w, g in zip(want[0] + want[1], got[0] + got[1]):
        assert np.array_equal(w, g)
    for q, (w, g) in enumerate(zip(want[2], got[2])):
        e = rel_rms(g, w)
        assert e < 1e-6, (q, e)
        assert np.max(np.abs(g - w)) <= 1e-5 * np.max(np.abs(w)), q
    plan.close()
