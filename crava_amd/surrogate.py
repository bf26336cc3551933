"""Rock-physics surrogate: training on the GPU, error-gated inference with
the rplib kernel as fallback, and facies-conditional sample generation.

The workload (5 -> 128 -> 128 -> 128 -> 3 tanh MLP of porosity, clay
fraction, water saturation, pore aspect ratio and effective pressure ->
ln vp, ln vs, ln rho) stands in for the per-sample rplib chain that
DistributionsRockDEM runs when faciesprob generates synthetic elastic
samples (reference rplib/distributionsrockdem.cpp:125-156,
src/faciesprob.cpp:2494-2559).  crv_rock_forward, the GPU port of that chain,
produces the training targets, measures the network's error and re-evaluates
every sample the error gate does not trust.

Device memory is held in torch tensors (plumbing); every computation is a
crv_* kernel.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .hip import ROCK_DEFAULTS, Context, rock_forward_device

# input box of the workload: porosity U(0.05, 0.35), clay Beta(2, 5),
# Sw U(0, 1), aspect ratio log-uniform (0.01, 1), effective pressure U(10, 40) MPa
IN_LO = (0.05, 0.0, 0.0, -2.0, 10.0)
IN_HI = (0.35, 1.0, 1.0, 0.0, 40.0)
IN_LOG_MASK = 1 << 3


def sample_inputs(n: int, rng: np.random.Generator, facies: str | None = None) -> np.ndarray:
    """(n, 5) float32 inputs.  facies None draws the training box; 'shale',
    'brine_sand' and 'gas_sand' narrow clay and saturation to the facies."""
    phi = rng.uniform(0.05, 0.35, n)
    clay = rng.beta(2, 5, n)
    sw = rng.uniform(0, 1, n)
    asp = 10.0 ** rng.uniform(-2, 0, n)
    pe = rng.uniform(10, 40, n)
    if facies == "shale":
        clay = 0.5 + 0.5 * rng.beta(4, 2, n)
        sw = np.ones(n)
        phi = rng.uniform(0.05, 0.2, n)
        asp = 10.0 ** rng.uniform(-2, -1, n)
    elif facies == "brine_sand":
        clay = 0.25 * rng.beta(2, 5, n)
        sw = np.ones(n)
        phi = rng.uniform(0.15, 0.35, n)
        asp = 10.0 ** rng.uniform(-0.7, 0, n)
    elif facies == "gas_sand":
        clay = 0.25 * rng.beta(2, 5, n)
        sw = rng.uniform(0.05, 0.4, n)
        phi = rng.uniform(0.15, 0.35, n)
        asp = 10.0 ** rng.uniform(-0.7, 0, n)
    return np.column_stack([phi, clay, sw, asp, pe]).astype(np.float32)


FACIES_INPUTS = ("shale", "brine_sand", "gas_sand")


def _quantile_inputs(u: np.ndarray, facies: str | None) -> np.ndarray:
    """The inputs of sample_inputs as functions of five uniforms u (n, 5):
    each parameter is its distribution's quantile at its own u (the
    reference draws a rock from such quantiles, DistributionsRockDEM::
    GenerateSamplePrivate, rplib/distributionsrockdem.cpp:125-156)."""
    from scipy.special import betaincinv
    uni = lambda q, a, b: a + (b - a) * q
    phi, clay, sw, lasp, pe = (uni(u[:, 0], 0.05, 0.35), betaincinv(2, 5, u[:, 1]), uni(u[:, 2], 0.0, 1.0),
                               uni(u[:, 3], -2.0, 0.0), uni(u[:, 4], 10.0, 40.0))
    if facies == "shale":
        clay = 0.5 + 0.5 * betaincinv(4, 2, u[:, 1])
        sw = np.ones(len(u))
        phi = uni(u[:, 0], 0.05, 0.2)
        lasp = uni(u[:, 3], -2.0, -1.0)
    elif facies in ("brine_sand", "gas_sand"):
        clay = 0.25 * betaincinv(2, 5, u[:, 1])
        sw = np.ones(len(u)) if facies == "brine_sand" else uni(u[:, 2], 0.05, 0.4)
        phi = uni(u[:, 0], 0.15, 0.35)
        lasp = uni(u[:, 3], -0.7, 0.0)
    return np.column_stack([phi, clay, sw, 10.0 ** lasp, pe]).astype(np.float32)


def synthetic_wells(nwells: int, length: int, dz: float, rng: np.random.Generator, facies=FACIES_INPUTS,
                    corr: float = 0.5, uniforms=None):
    """Synthetic well logs as FaciesProb::GenerateSyntWellData makes them
    (reference src/faciesprob.cpp:2494-2559): a well is a sequence of facies
    runs until it holds at least `length` samples (the last run is kept
    whole); each run's facies is floor(nf U), its length ceil(-ln U / lambda)
    with lambda = -ln(1 - dz / 10) (geometric, mean 10 / dz samples; dz must
    be below 10), and its samples come from DistributionsRock::
    GenerateWellSample (rplib/distributionsrock.cpp:42-59) with corr = 0.5:
    the first from fresh uniforms, each next one by DEMTools::UpdateU
    (rplib/demmodelling.cpp:820-850), u' = Phi(corr Phi^-1(u) +
    sqrt(1 - corr^2) Z) per parameter.  The uniforms and normals come from
    rng (PCG64), not the reference's dSFMT stream; uniforms (a callable
    returning the next uniform) replaces rng's for the facies and lengths.

    Returns (facies index per sample, inputs (n, 5) float32, well index per
    sample, run index per sample)."""
    from scipy.special import ndtr, ndtri
    if not 0.0 < dz < 10.0:
        raise ValueError("synthetic wells: dz must be in (0, 10)")
    lam = -math.log(1.0 - dz / 10.0)
    unif = uniforms if uniforms is not None else rng.random
    nf = len(facies)
    fac, us, well, run = [], [], [], []
    eps = math.sqrt(1.0 - corr * corr)
    r = 0
    for w in range(nwells):
        k = 0
        while k < length:
            f = int(math.floor(nf * unif()))
            L = int(math.ceil(-math.log(unif()) / lam))
            u = np.empty((L, 5))
            u[0] = rng.random(5)
            for i in range(1, L):
                u[i] = ndtr(corr * ndtri(u[i - 1]) + eps * rng.standard_normal(5))
            fac.append(np.full(L, f))
            us.append((f, u))
            well.append(np.full(L, w))
            run.append(np.full(L, r))
            r += 1
            k += L
    x = np.concatenate([_quantile_inputs(u, facies[f]) for f, u in us]) if us else np.zeros((0, 5), np.float32)
    return np.concatenate(fac), x, np.concatenate(well), np.concatenate(run)


def training_inputs(n: int, rng: np.random.Generator, weights=(1.0, 1.0, 1.0)) -> np.ndarray:
    """(n, 5) training inputs: half from the broad box, the other half split
    over the facies distributions that the sampling step queries in the
    proportions `weights` (FACIES_INPUTS order), shuffled (a surrogate is
    trained on the distribution it will see)."""
    nb = n // 2
    w = np.asarray(weights, float) / float(np.sum(weights))
    counts = np.floor(w * (n - nb)).astype(int)
    counts[-1] = n - nb - int(counts[:-1].sum())
    parts = [sample_inputs(nb, rng)]
    for f, m in zip(FACIES_INPUTS, counts):
        parts.append(sample_inputs(int(m), rng, facies=f))
    x = np.concatenate(parts)
    return x[rng.permutation(n)]


def make_norm(out_mean, out_std) -> L.MlpNorm:
    nm = L.MlpNorm()
    for f in range(5):
        nm.in_lo[f] = IN_LO[f]
        nm.in_scale[f] = 2.0 / (IN_HI[f] - IN_LO[f])
    nm.in_log_mask = IN_LOG_MASK
    for c in range(3):
        nm.out_mean[c] = float(out_mean[c])
        nm.out_std[c] = float(out_std[c])
    return nm


def norm_dict(nm: L.MlpNorm) -> dict:
    return dict(in_lo=list(nm.in_lo), in_scale=list(nm.in_scale), in_log_mask=nm.in_log_mask,
                out_mean=list(nm.out_mean), out_std=list(nm.out_std))


def _dev():
    import torch
    return torch.device("cuda", 0)


class MLP:
    """crv_mlp handle."""

    def __init__(self, ctx: Context, norm: L.MlpNorm, max_batch: int):
        self.ctx = ctx
        self.lib = ctx.lib
        self.norm = norm
        self.max_batch = int(max_batch)
        h = C.c_void_p()
        ctx.check(self.lib.crv_mlp_create(ctx.h, C.byref(norm), self.max_batch, C.byref(h)), "mlp_create")
        self.h = h
        self.nparams = int(self.lib.crv_mlp_param_count())

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        self.ctx.check(self.lib.crv_mlp_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))), "mlp_set_params")

    def get_params(self) -> np.ndarray:
        p = np.empty(self.nparams, np.float32)
        self.ctx.check(self.lib.crv_mlp_get_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))), "mlp_get_params")
        return p

    def gradients(self, d_x, d_t, n: int):
        g = np.empty(self.nparams, np.float32)
        loss = C.c_float()
        self.ctx.check(self.lib.crv_mlp_gradients(self.h, C.c_void_p(d_x), C.c_void_p(d_t), int(n),
                                                  g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss)),
                       "mlp_gradients")
        return float(loss.value), g

    def train_step(self, d_x, d_t, n: int, lr: float, step: int, want_loss: bool = False):
        loss = C.c_float()
        self.ctx.check(self.lib.crv_mlp_train_step(self.h, C.c_void_p(d_x), C.c_void_p(d_t), int(n), float(lr),
                                                   int(step), C.byref(loss) if want_loss else None), "mlp_train_step")
        return float(loss.value) if want_loss else None

    def predict(self, d_x, n: int, d_y):
        self.ctx.check(self.lib.crv_mlp_predict(self.h, C.c_void_p(d_x), int(n), C.c_void_p(d_y)), "mlp_predict")

    def flops(self):
        f, b = C.c_double(), C.c_double()
        self.ctx.check(self.lib.crv_mlp_flops(self.h, C.byref(f), C.byref(b)), "mlp_flops")
        return f.value, b.value

    def close(self):
        if self.h and self.ctx.h:
            self.lib.crv_mlp_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rplib_targets(ctx: Context, d_x, n: int, params: dict | None = None):
    """ln vp, ln vs, ln rho of device inputs through crv_rock_forward (float
    outputs) and the per-sample status."""
    import torch
    x64 = d_x.to(torch.float64)
    y = torch.empty((n, 3), dtype=torch.float32, device=_dev())
    st = torch.zeros(n, dtype=torch.int32, device=_dev())
    torch.cuda.synchronize()
    rock_forward_device(ctx, params or ROCK_DEFAULTS, x64.data_ptr(), y.data_ptr(), n, True, st.data_ptr())
    ctx.synchronize()
    return y, st


@dataclass
class TrainResult:
    mlp: MLP
    history: list = field(default_factory=list)
    seconds: float = 0.0
    steps: int = 0
    train_rows: int = 0
    mfma_tflops: float = 0.0
    mfma_util: float = 0.0
    step_kernels_ms: dict = field(default_factory=dict)   # one training step, per kernel


def train(ctx: Context, n_train: int = 1_000_000, epochs: int = 40, batch: int = 16384, lr: float = 2e-3,
          seed: int = 7, n_val: int = 200_000, params: dict | None = None, log=None,
          vp_weight: float = 1.0, mix_facies: bool = True, facies_weights=(1.0, 1.0, 1.0)) -> TrainResult:
    """Generate n_train + n_val rplib samples on the GPU (crv_rock_forward)
    (training_inputs, or the broad box only without mix_facies) and fit the MLP with Adam, cosine learning-rate decay and reshuffled
    mini-batches.  vp_weight > 1 shrinks the normalisation scale of ln vp,
    which weights its errors by vp_weight^2 in the loss (the error gate is
    on vp).  Returns the trained network with loss history, wall time and
    the matrix-core rate of the training kernels."""
    import torch
    dev = _dev()
    rng = np.random.default_rng(seed)
    x = (training_inputs(n_train + n_val, rng, facies_weights) if mix_facies
         else sample_inputs(n_train + n_val, rng))
    d_x = torch.from_numpy(x).to(dev)
    d_y, st = rplib_targets(ctx, d_x, x.shape[0], params)
    keep = (st == 0)
    d_x, d_y = d_x[keep].contiguous(), d_y[keep].contiguous()
    n_all = d_x.shape[0]
    n_tr = n_all - n_val
    xtr, ytr = d_x[:n_tr].contiguous(), d_y[:n_tr].contiguous()
    yh = ytr.double().cpu().numpy()
    sd = yh.std(0)
    sd[0] /= vp_weight
    nm = make_norm(yh.mean(0), sd)
    mlp = MLP(ctx, nm, batch)
    mlp.set_params(_init_params(seed))
    per_epoch = n_tr // batch
    total = max(1, epochs * per_epoch)
    res = TrainResult(mlp=mlp)
    step = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for ep in range(epochs):
        # the previous epoch's steps read xs/ys on the library's stream:
        # finish them before torch may reuse that memory
        ctx.synchronize()
        perm = torch.randperm(n_tr, device=dev)
        xs, ys = xtr[perm].contiguous(), ytr[perm].contiguous()
        torch.cuda.synchronize()
        loss = None
        for b in range(per_epoch):
            step += 1
            lr_t = 1e-5 + (lr - 1e-5) * 0.5 * (1.0 + math.cos(math.pi * step / total))
            last = b == per_epoch - 1
            loss = mlp.train_step(xs[b * batch:].data_ptr(), ys[b * batch:].data_ptr(), batch, lr_t, step,
                                  want_loss=last)
        res.history.append((ep, loss))
        if log:
            log(f"[surrogate] epoch {ep:3d} loss {loss:.3e}")
    ctx.synchronize()
    res.seconds = time.perf_counter() - t0
    res.steps = step
    res.train_rows = step * batch
    # matrix-core rate of one profiled step
    ctx.profile(True)
    mlp.train_step(xtr.data_ptr(), ytr.data_ptr(), batch, 1e-5, step + 1)
    ctx.profile(False)
    rep = ctx.profile_report()
    res.step_kernels_ms = {k: v[1] for k, v in rep.items()}
    ms = sum(v[1] for k, v in rep.items() if k.startswith("mlp_") and k not in ("mlp_reduce", "mlp_reduce_adam"))
    f, b_ = mlp.flops()
    if ms > 0:
        res.mfma_tflops = (f + b_) / (ms * 1e-3) / 1e12
        res.mfma_util = res.mfma_tflops / 2500.0
    res.val_x, res.val_y = d_x[n_tr:].contiguous(), d_y[n_tr:].contiguous()
    return res


def _init_params(seed: int) -> np.ndarray:
    """Glorot-uniform weights, zero biases and padding, in the augmented layout."""
    rng = np.random.default_rng(seed)
    HID, KA, K1, NO = 128, 144, 16, 32
    W1, W2, W3, W4 = (np.zeros((HID, K1)), np.zeros((HID, KA)), np.zeros((HID, KA)), np.zeros((NO, KA)))
    W1[:, :5] = rng.uniform(-1, 1, (HID, 5)) * math.sqrt(6.0 / (5 + HID))
    W2[:, :HID] = rng.uniform(-1, 1, (HID, HID)) * math.sqrt(6.0 / (2 * HID))
    W3[:, :HID] = rng.uniform(-1, 1, (HID, HID)) * math.sqrt(6.0 / (2 * HID))
    W4[:3, :HID] = rng.uniform(-1, 1, (3, HID)) * math.sqrt(6.0 / (HID + 3))
    return np.concatenate([W1.ravel(), W2.ravel(), W3.ravel(), W4.ravel()]).astype(np.float32)


def predict(ctx: Context, mlp: MLP, d_x):
    """Network outputs (device (n, 3) float) for device inputs (inference
    needs no per-batch scratch: one launch for any n)."""
    import torch
    n = d_x.shape[0]
    d_y = torch.empty((n, 3), dtype=torch.float32, device=d_x.device)
    # the library's stream does not order itself after torch's: torch's work
    # on these tensors ends first, and the outputs are complete on return
    torch.cuda.synchronize()
    mlp.predict(d_x.data_ptr(), n, d_y.data_ptr())
    ctx.synchronize()
    return d_y


def validation_errors(ctx: Context, mlp: MLP, d_x, d_y) -> np.ndarray:
    """(n, 3) network minus rplib on held-out samples (host)."""
    p = predict(ctx, mlp, d_x)
    ctx.synchronize()
    return (p - d_y).cpu().numpy().astype(np.float64)


def input_cells(x: np.ndarray, nb: int) -> np.ndarray:
    """Error-table cell of each (n, 5) input, as the gate kernel forms it."""
    x = np.asarray(x, np.float32).copy()
    cell = np.zeros(x.shape[0], np.int64)
    for f in range(5):
        v = x[:, f]
        if IN_LOG_MASK & (1 << f):
            v = np.log10(v)
        lo, sc = np.float32(IN_LO[f]), np.float32(2.0 / (IN_HI[f] - IN_LO[f]))
        t = ((v - lo) * sc) * np.float32(0.5)
        b = np.clip(np.floor(t * np.float32(nb)).astype(np.int64), 0, nb - 1)
        cell = cell * nb + b
    return cell


class ErrorTable:
    """A high quantile of the network's error on held-out samples (ln vp by
    default), per cell of an nb^5 grid over the normalised inputs; +inf
    where fewer than min_count samples fell."""

    def __init__(self, values: np.ndarray, nb: int):
        import torch
        self.nb = nb
        self.values = np.asarray(values, np.float32)
        self.d = torch.from_numpy(self.values).to(_dev())

    @classmethod
    def build(cls, ctx: Context, mlp: MLP, d_x, d_y, nb: int = 6, min_count: int = 20,
              quantile: float = 0.999, outputs=(0,)) -> "ErrorTable":
        """Per cell, the `quantile` of the held-out errors (the largest error
        when quantile >= 1); +inf below min_count samples.  A sample's error
        is the largest |error| over `outputs` (0 ln vp, 1 ln vs, 2 ln rho)."""
        err = np.abs(validation_errors(ctx, mlp, d_x, d_y)[:, list(outputs)]).max(axis=1)
        cell = input_cells(d_x.cpu().numpy(), nb)
        order = np.lexsort((err, cell))
        cs, es = cell[order], err[order]
        ncell = nb ** 5
        start = np.searchsorted(cs, np.arange(ncell), side="left")
        stop = np.searchsorted(cs, np.arange(ncell), side="right")
        cnt = stop - start
        vals = np.full(ncell, np.inf)
        ok = cnt >= max(1, min_count)
        # index of the quantile inside each cell's sorted errors
        q = np.minimum(stop - 1, start + np.floor(quantile * (cnt - 1)).astype(np.int64))
        vals[ok] = es[q[ok]]
        return cls(vals, nb)


def predict_checked(ctx: Context, mlp: MLP, table: ErrorTable, d_x, thresh: float = 0.005,
                    params: dict | None = None):
    """Network outputs with every sample whose table cell predicts an ln vp
    error above thresh recomputed by the rplib kernel.  Returns the device
    outputs and the number of recomputed samples."""
    import torch
    n = d_x.shape[0]
    d_y = predict(ctx, mlp, d_x)
    p = L.RockParams(**(params or ROCK_DEFAULTS))
    nref = C.c_longlong()
    ctx.check(ctx.lib.crv_rock_refine(ctx.h, C.byref(p), n, C.c_void_p(d_x.data_ptr()), C.c_void_p(d_y.data_ptr()),
                                      C.byref(mlp.norm), C.c_void_p(table.d.data_ptr()), table.nb, float(thresh),
                                      C.byref(nref)), "rock_refine")
    ctx.synchronize()
    return d_y, int(nref.value)
