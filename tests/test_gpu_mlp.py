"""The rock-physics surrogate's BF16 matrix-core kernels against the float64
oracle (oracle/mlp_oracle.py): forward values, loss and gradients of every
parameter, then training and the rplib-backed error gate."""
import numpy as np
import pytest

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _setup(ctx, n, seed):
    import torch
    from crava_amd import surrogate as S
    rng = np.random.default_rng(seed)
    x = S.sample_inputs(n, rng)
    t = np.column_stack([rng.normal(8.0, 0.2, n), rng.normal(7.2, 0.5, n),
                         rng.normal(7.7, 0.05, n)]).astype(np.float32)
    nm = S.make_norm([8.0, 7.2, 7.7], [0.2, 0.5, 0.05])
    mlp = S.MLP(ctx, nm, n)
    p = S._init_params(seed)
    # non-zero biases so that the bias columns are exercised
    mask = M.trainable_mask().astype(bool)
    b = np.zeros_like(p)
    for off, rows, ld, col in ((M.OFF_W1, 128, 16, 5), (M.OFF_W2, 128, 144, 128), (M.OFF_W3, 128, 144, 128),
                               (M.OFF_W4, 3, 144, 128)):
        b[off + np.arange(rows) * ld + col] = rng.normal(0, 0.1, rows)
    p = (p + b * mask).astype(np.float32)
    mlp.set_params(p)
    dev = torch.device("cuda", 0)
    d_x = torch.from_numpy(x).to(dev)
    d_t = torch.from_numpy(t).to(dev)
    torch.cuda.synchronize()
    return mlp, p, x, t, d_x, d_t, S.norm_dict(nm)


# 70001: more 128-sample tiles than blocks (each block walks several, the
# next one's inputs loaded ahead) and a partial last tile
@pytest.mark.parametrize("n", [4096, 1000, 70001])
def test_forward_matches_oracle(ctx, n):
    import torch
    mlp, p, x, t, d_x, d_t, nd = _setup(ctx, n, 5)
    d_y = torch.empty((n, 3), dtype=torch.float32, device=d_x.device)
    mlp.predict(d_x.data_ptr(), n, d_y.data_ptr())
    ctx.synchronize()
    y = d_y.cpu().numpy()
    yo, _ = M.forward(p, x, nd, mode="bf16")
    yo = yo * np.asarray(nd["out_std"]) + np.asarray(nd["out_mean"])
    # the kernel accumulates in FP32 and rounds activations to BF16 as the
    # oracle does; differences come from FP32 vs float64 sums and tanh
    assert np.max(np.abs(y - yo) / np.asarray(nd["out_std"])) < 2e-2
    assert np.sqrt(np.mean(((y - yo) / np.asarray(nd["out_std"])) ** 2)) < 2e-3


@pytest.mark.parametrize("n", [4096, 3000])
def test_gradients_match_oracle(ctx, n):
    mlp, p, x, t, d_x, d_t, nd = _setup(ctx, n, 11)
    loss, g = mlp.gradients(d_x.data_ptr(), d_t.data_ptr(), n)
    lo, go = M.loss_grad(p, x, t, nd, mode="bf16")
    le, ge = M.loss_grad(p, x, t, nd, mode="exact")
    assert loss == pytest.approx(lo, rel=1e-3)
    assert loss == pytest.approx(le, rel=1e-2)
    mask = M.trainable_mask().astype(bool)
    assert np.all(g[~mask] == 0)
    for a, b in ((0, M.OFF_W2), (M.OFF_W2, M.OFF_W3), (M.OFF_W3, M.OFF_W4), (M.OFF_W4, M.NPARAMS)):
        gl, ol, el = g[a:b][mask[a:b]], go[a:b][mask[a:b]], ge[a:b][mask[a:b]]
        # against the oracle with the kernels' BF16 roundings, per layer
        assert np.linalg.norm(gl - ol) / np.linalg.norm(ol) < 1e-2
        # against exact float64 arithmetic
        assert np.linalg.norm(gl - el) / np.linalg.norm(el) < 5e-2
        assert np.dot(gl, el) / (np.linalg.norm(gl) * np.linalg.norm(el)) > 0.999


def test_training_fits_rplib(ctx):
    """3840 Adam steps on 524288 rplib samples lower the loss by an order of
    magnitude and the network's ln vp error on held-out samples is small."""
    from crava_amd import surrogate as S
    r = S.train(ctx, n_train=524288, n_val=65536, epochs=60, batch=8192, lr=3e-3, seed=3, vp_weight=3.0)
    first, last = r.history[0][1], r.history[-1][1]
    assert last < 0.1 * first
    err = S.validation_errors(ctx, r.mlp, r.val_x, r.val_y)
    assert np.median(np.abs(err[:, 0])) < 0.01
    assert r.mfma_tflops > 0


def test_error_gate_bounds_the_error(ctx):
    """After the rplib fallback the ln vp error exceeds the gate threshold
    for no more samples than the table's 0.999 quantile allows, and the gate
    trusts part of the samples (the rest equal rplib)."""
    import torch
    from crava_amd import surrogate as S
    r = S.train(ctx, n_train=1_000_000, n_val=1_000_000, epochs=200, batch=16384, lr=3e-3, seed=4, vp_weight=3.0)
    table = S.ErrorTable.build(ctx, r.mlp, r.val_x, r.val_y, nb=5)
    rng = np.random.default_rng(9)
    x = S.sample_inputs(200_000, rng)
    d_x = torch.from_numpy(x).to(torch.device("cuda", 0))
    y_gate, nref = S.predict_checked(ctx, r.mlp, table, d_x, thresh=0.005)
    y_ref, st = S.rplib_targets(ctx, d_x, x.shape[0])
    ok = (st == 0).cpu().numpy()
    y_gate, y_ref = y_gate.cpu().numpy(), y_ref.cpu().numpy()
    print(f"gate: refined {nref / x.shape[0]:.1%}")
    assert 0 < nref < 0.9 * x.shape[0]
    err = np.abs(y_gate[ok, 0] - y_ref[ok, 0])
    assert np.mean(err > 0.005) < 0.003
    assert np.sum(err == 0.0) >= nref - np.sum(~ok)
