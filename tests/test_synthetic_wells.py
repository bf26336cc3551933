"""Synthetic well logs of FaciesProb::GenerateSyntWellData (reference
src/faciesprob.cpp:2494-2559) with the correlated rock updates of
DEMTools::UpdateU (rplib/demmodelling.cpp:820-850), as crava_amd.surrogate
restates them: the run lengths, facies choice, correlation within runs and
marginals are checked against what the reference's formulas fix."""
import math

import numpy as np
from scipy import stats
from scipy.special import ndtri

from crava_amd import surrogate as S


def _wells(seed=1, nwells=400, length=300, dz=2.0, corr=0.5):
    return S.synthetic_wells(nwells, length, dz, np.random.default_rng(seed), corr=corr)


def test_run_lengths_are_geometric_with_mean_10_over_dz():
    """ceil(-ln U / lambda) with lambda = -ln(1 - dz/10) is geometric with
    success probability p = dz / 10: mean 10 / dz, P(L = 1) = p."""
    dz = 2.0
    f, x, w, r = _wells(dz=dz)
    L = np.bincount(r)
    p = dz / 10.0
    mean, sd = 1.0 / p, math.sqrt(1.0 - p) / p
    assert abs(L.mean() - mean) < 4 * sd / math.sqrt(len(L))
    ones = np.mean(L == 1)
    assert abs(ones - p) < 4 * math.sqrt(p * (1 - p) / len(L))


def test_facies_of_runs_uniform_and_wells_long_enough():
    length = 300
    f, x, w, r = _wells(length=length)
    first = np.r_[0, np.nonzero(np.diff(r))[0] + 1]
    counts = np.bincount(f[first], minlength=3)
    assert stats.chisquare(counts).pvalue > 1e-3
    # a well holds at least `length` samples, and stops at the first run
    # that reaches it
    for wi in range(10):
        sel = w == wi
        n = int(sel.sum())
        last = r[sel][-1]
        assert n >= length and n - int(np.sum(r[sel] == last)) < length
    # runs never mix facies
    for ri in range(0, r.max(), 97):
        assert len(np.unique(f[r == ri])) == 1


def _porosity_scores(f, x):
    """Normal score of each sample's porosity quantile (porosity is uniform
    on a facies-dependent interval)."""
    lo = np.where(f == 0, 0.05, 0.15)
    hi = np.where(f == 0, 0.2, 0.35)
    u = (x[:, 0].astype(np.float64) - lo) / (hi - lo)
    return u, ndtri(np.clip(u, 1e-12, 1 - 1e-12))


def test_updates_keep_the_marginals_and_correlate_within_runs():
    """u' = Phi(corr Phi^-1(u) + sqrt(1 - corr^2) Z) leaves U(0,1) invariant
    and makes consecutive normal scores of a run correlated by corr; runs
    are independent of each other."""
    for corr in (0.5, 0.9):
        f, x, w, r = _wells(seed=3, corr=corr)
        u, z = _porosity_scores(f, x)
        assert stats.kstest(u, "uniform").pvalue > 1e-3
        same = r[1:] == r[:-1]
        c_in = np.corrcoef(z[:-1][same], z[1:][same])[0, 1]
        c_out = np.corrcoef(z[:-1][~same], z[1:][~same])[0, 1]
        assert abs(c_in - corr) < 0.02, (corr, c_in)
        assert abs(c_out) < 0.03


def test_injected_uniforms_give_the_reference_formulas():
    """With the facies and length uniforms injected: facies floor(nf U),
    length ceil(-ln U / lambda)."""
    dz = 1.0
    lam = -math.log(1.0 - dz / 10.0)
    seq = iter([0.1, math.exp(-lam * 2.5), 0.99, math.exp(-lam * 0.2), 0.5, math.exp(-lam * 7.0)])
    f, x, w, r = S.synthetic_wells(1, 11, dz, np.random.default_rng(0), uniforms=lambda: next(seq))
    # runs: facies 0 of 3 samples, facies 2 of 1, facies 1 of 7 (reaches 11)
    assert list(f) == [0] * 3 + [2] + [1] * 7
    assert list(np.bincount(r)) == [3, 1, 7]


def test_training_mix_follows_the_facies_weights():
    """Half the training inputs come from the broad box, the other half from
    the facies in proportion to the weights; Sw = 1 exactly marks shale and
    brine sand (the box draws Sw uniformly), clay >= 0.5 among those marks
    shale (brine sand's clay is at most 0.25)."""
    n = 60000
    x = S.training_inputs(n, np.random.default_rng(4), weights=(3.0, 1.0, 1.0))
    assert x.shape == (n, 5)
    sat = x[:, 2] == 1.0
    assert int(sat.sum()) == 18000 + 6000
    assert int((sat & (x[:, 1] >= 0.5)).sum()) == 18000
    # equal weights: the thirds of the original mix
    y = S.training_inputs(n, np.random.default_rng(4))
    sat = y[:, 2] == 1.0
    assert int(sat.sum()) == 20000 and int((sat & (y[:, 1] >= 0.5)).sum()) == 10000
