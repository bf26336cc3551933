"""Facies densities and probabilities on the GPU against the float64
oracle (oracle/facies_oracle.py), and the surrogate's gain: facies
probabilities over a posterior from surrogate samples agree with those from
rplib samples while the samples cost less to make."""
import time

import numpy as np
import pytest

from oracle import facies_oracle as F

pytestmark = pytest.mark.gpu


def _cloud(n, seed):
    rng = np.random.default_rng(seed)
    mu = np.array([8.0, 7.3, 7.7])
    C = np.array([[0.010, 0.006, 0.001], [0.006, 0.020, 0.0015], [0.001, 0.0015, 0.002]])
    return rng.multivariate_normal(mu, C, n).astype(np.float32), mu, C


@pytest.mark.parametrize("bins", [(40, 36, 30), (100, 100, 50)])
def test_density_matches_oracle(ctx, bins):
    import torch
    from crava_amd import facies
    x, mu, C = _cloud(200_000, 1)
    lo, hi = mu - 6 * np.sqrt(np.diag(C)), mu + 6 * np.sqrt(np.diag(C))
    S = 0.3 * C
    box = facies.make_box(lo, hi)
    d_x = torch.from_numpy(x).to(torch.device("cuda", 0))
    torch.cuda.synchronize()
    g = facies.density(ctx, d_x, box, S, bins)
    got = g.download_real()[:, :, :bins[0]]
    want = F.density(x, lo, hi, S, bins)
    assert np.max(np.abs(got - want)) / want.max() < 2e-4


@pytest.mark.parametrize("bins", [(40, 36, 30), (100, 100, 50)])
def test_sampled_kernel_density_matches_reference_restatement(ctx, bins):
    """PosteriorElasticPDF3D's density (histogram / n convolved with the
    sampled, sum-normalised Gaussian kernel; reference
    src/posteriorelasticpdf3d.cpp:12-131, 369-420) against its CPU
    restatement.  The restatement keeps the reference's float running sum of
    the kernel, the library sums it in double: a common scale factor that
    differs by less than 1e-4.  The result is a mass per bin: its sum is 1."""
    import torch
    from crava_amd import _lib as L, facies
    x, mu, C = _cloud(200_000, 4)
    lo, hi = mu - 6 * np.sqrt(np.diag(C)), mu + 6 * np.sqrt(np.diag(C))
    S = 0.3 * C
    box = facies.make_box(lo, hi)
    d_x = torch.from_numpy(x).to(torch.device("cuda", 0))
    torch.cuda.synchronize()
    g = facies.density(ctx, d_x, box, S, bins, kernel=L.CRV_SMOOTH_SAMPLED)
    got = g.download_real()[:, :, :bins[0]].astype(np.float64)
    want = F.density_pdf3d(x, lo, hi, S, bins)
    assert np.max(np.abs(got - want)) / want.max() < 1e-4
    assert got.sum() == pytest.approx(1.0, abs=1e-4)
    # against the Fourier-Gaussian mode: the same smoothing up to the
    # kernel's sampling (its width is 2-3 bins here) and the unit change
    f = facies.density(ctx, d_x, box, S, bins).download_real()[:, :, :bins[0]].astype(np.float64)
    cell = np.prod((hi - lo) / np.array(bins))
    assert np.max(np.abs(f * cell - got)) / got.max() < 2e-2


def test_probabilities_match_oracle(ctx):
    import torch
    from crava_amd import facies
    from crava_amd.hip import FFTGrid
    bins = (48, 40, 32)
    rng = np.random.default_rng(3)
    clouds = []
    for k, shift in enumerate((np.zeros(3), np.array([0.1, 0.2, 0.03]), np.array([-0.1, 0.25, -0.04]))):
        x, mu, C = _cloud(100_000, 10 + k)
        clouds.append((x + shift).astype(np.float32))
    lo, hi = np.array([7.5, 6.7, 7.45]), np.array([8.5, 7.9, 7.95])
    S = np.diag([0.004, 0.008, 0.001])
    box = facies.make_box(lo, hi)
    dens = [facies.density(ctx, torch.from_numpy(c).to(torch.device("cuda", 0)), box, S, bins) for c in clouds]
    nx, ny, nz = 20, 18, 16
    m = [np.ascontiguousarray(np.pad(rng.uniform(lo[q], hi[q], (nz, ny, nx)), ((0, 0), (0, 0), (0, 2))),
                              dtype=np.float32) for q in range(3)]
    mean = []
    for a in m:
        g = FFTGrid(ctx, nx, ny, nz)
        g.upload_real(a)
        mean.append(g)
    prior = [0.5, 0.3, 0.2]
    P = facies.probabilities(ctx, dens, prior, box, mean)
    got = np.stack([p.download_real()[:, :, :nx].ravel() for p in P])
    pts = np.column_stack([a[:, :, :nx].ravel() for a in m])
    dens_o = [F.density(c, lo, hi, S, bins) for c in clouds]
    want = F.probabilities(dens_o, prior, lo, hi, bins, pts)
    assert np.allclose(got.sum(0), 1.0, atol=1e-5)
    # Far in the tails every density is at the FP32 transform's round-off
    # floor (about 1e-6 of its peak) and the ratio carries no information;
    # compare where the evidence is at least 1e-3 of the largest peak.
    pr = np.asarray(prior) / np.sum(prior)
    ev = sum(p * F.trilinear(d, lo, hi, bins, pts) for p, d in zip(pr, dens_o))
    sig = ev >= 1e-3 * max(p * d.max() for p, d in zip(pr, dens_o))
    assert sig.mean() > 0.05
    assert np.max(np.abs(got - want)[:, sig]) < 2e-3
    # with the reference's undefined class (p_undef 0.01) the tails are
    # bounded: compare every voxel
    Pd, Pu = facies.probabilities(ctx, dens, prior, box, mean, p_undefined=0.01, with_undefined=True)
    got_d = np.stack([p.download_real()[:, :, :nx].ravel() for p in Pd])
    got_u = Pu.download_real()[:, :, :nx].ravel()
    want_d, want_u = F.probabilities_undef(dens_o, prior, lo, hi, bins, pts, 0.01)
    assert np.allclose(got_d.sum(0) + got_u, 1.0, atol=1e-5)
    err = max(np.max(np.abs(got_d - want_d)), np.max(np.abs(got_u - want_u)))
    print(f"facies probabilities with the undefined class: max error over all voxels {err:.2e}, "
          f"undefined-class share {got_u.mean():.3f}")
    assert err < 2e-3


def test_surrogate_gain_in_facies_probabilities(ctx):
    """Three facies x 1e7 samples: the surrogate with its rplib-backed error
    gate gives the facies probabilities of rplib itself (max difference
    0.03 over a posterior cube) and generates the samples faster."""
    import torch
    from crava_amd import facies, surrogate as S
    from crava_amd.hip import FFTGrid
    dev = torch.device("cuda", 0)
    r = S.train(ctx, n_train=1_000_000, n_val=2_000_000, epochs=400, batch=16384, lr=3e-3, seed=5, vp_weight=3.0)
    table = S.ErrorTable.build(ctx, r.mlp, r.val_x, r.val_y, nb=6)
    rng = np.random.default_rng(21)
    xs = [torch.from_numpy(S.sample_inputs(10_000_000, rng, facies=f)).to(dev) for f in facies.FACIES]
    S.rplib_targets(ctx, xs[0][:65536], 65536)
    S.predict_checked(ctx, r.mlp, table, xs[0][:65536], thresh=0.005)
    ctx.synchronize()
    t0 = time.perf_counter()
    ys_rp = [S.rplib_targets(ctx, x, x.shape[0])[0] for x in xs]
    ctx.synchronize()
    t_rplib = time.perf_counter() - t0
    t0 = time.perf_counter()
    outs = [S.predict_checked(ctx, r.mlp, table, x, thresh=0.005) for x in xs]
    ctx.synchronize()
    t_nn = time.perf_counter() - t0
    ys_nn = [o[0] for o in outs]
    lo, hi = np.array([6.5, 4.5, 7.3]), np.array([8.9, 8.4, 8.0])
    box = facies.make_box(lo, hi)
    Sm = np.diag([0.003, 0.01, 0.0008])
    keep = lambda y: y[torch.isfinite(y).all(1)].contiguous()
    d_rp = [facies.density(ctx, keep(y), box, Sm) for y in ys_rp]
    d_nn = [facies.density(ctx, keep(y), box, Sm) for y in ys_nn]
    nx, ny, nz = 64, 48, 40
    rng2 = np.random.default_rng(4)
    mean = []
    for q in range(3):
        g = FFTGrid(ctx, nx, ny, nz)
        g.upload_real(np.pad(rng2.uniform(lo[q] + 0.3 * (hi[q] - lo[q]), hi[q] - 0.3 * (hi[q] - lo[q]), (nz, ny, nx)),
                             ((0, 0), (0, 0), (0, 2))).astype(np.float32))
        mean.append(g)
    prior = [0.4, 0.4, 0.2]
    P_rp = facies.probabilities(ctx, d_rp, prior, box, mean)
    P_nn = facies.probabilities(ctx, d_nn, prior, box, mean)
    # compare where the rplib evidence is at least 1e-3 of its largest peak
    bins = (100, 100, 50)
    pts = np.column_stack([g.download_real()[:, :, :nx].ravel() for g in mean])
    dh = [d.download_real()[:, :, :bins[0]].astype(np.float64) for d in d_rp]
    pr = np.asarray(prior) / np.sum(prior)
    ev = sum(p * F.trilinear(d, lo, hi, bins, pts) for p, d in zip(pr, dh))
    sig = ev >= 1e-3 * max(p * d.max() for p, d in zip(pr, dh))
    assert sig.mean() > 0.05
    diff = max(np.max(np.abs(a.download_real()[:, :, :nx].ravel() - b.download_real()[:, :, :nx].ravel())[sig])
               for a, b in zip(P_rp, P_nn))
    refined = [o[1] / 10_000_000 for o in outs]
    # every voxel, with the reference's undefined class (p_undef 0.01),
    # which keeps the facies probabilities of voxels far from every facies
    # near 0 instead of ratios of vanishing densities
    Pu_rp, U_rp = facies.probabilities(ctx, d_rp, prior, box, mean, p_undefined=0.01, with_undefined=True)
    Pu_nn, U_nn = facies.probabilities(ctx, d_nn, prior, box, mean, p_undefined=0.01, with_undefined=True)
    diff_all = max(np.max(np.abs(a.download_real()[:, :, :nx] - b.download_real()[:, :, :nx]))
                   for a, b in zip(Pu_rp + [U_rp], Pu_nn + [U_nn]))
    print(f"facies probabilities: max |surrogate - rplib| = {diff:.2e} (significant evidence), {diff_all:.2e} "
          f"(every voxel, undefined class 0.01); samples: rplib {t_rplib * 1e3:.1f} ms, "
          f"surrogate + gate {t_nn * 1e3:.1f} ms (refined {', '.join(f'{f:.1%}' for f in refined)})")
    assert diff < 0.03
    assert diff_all < 0.03
    assert t_nn < t_rplib
