"""Learned-path form of the MLP (include/pct_net.h).

A small network (12 -> 64 -> 64 -> 8, tanh, BF16 MFMA with FP32 accumulation)
predicts the offsets of the exact Schulte MLP from the cubic Hermite path at
four interior knots of the normalised depth, for t and v; the path is then
stepped in FP32 as Hermite + clamped cubic spline (pct_amd/csrc/pct_net_basis.h)
and voxelised like the precision form. Histories whose inputs are outside the
trained range, or whose (u-span, bending) cell failed the calibration check,
take the exact FP64 walk (.cu:5279-5501).

Everything runs in libpct.so: training data (exact MLP at the knots), the
training kernels (forward + backward + Adam), the surrogate walk and the
calibration. This module drives them: it builds the training set from
random-ellipsoid phantoms, trains, installs and calibrates a network, and
reads the fallback counters. There is no CPU path.
"""
import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib, configs
from ._lib import NET_IN, NET_OUT, NET_PARAMS, PctNetCalib, PctNetTrainCfg

FEATURES = ("dt*2", "dv*2", "dtheta_t*20", "dtheta_v*20", "u2/8-1.5", "wepl/8-1.5", "t_in/10", "t_out/10",
            "v_in/4", "v_out/4", "xz_in*10", "xz_out*10")
OUT_SCALE_CM = 0.01          # one network output unit in cm


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def init_params(seed=0, device="cuda"):
    """Xavier-uniform weights, zero biases (FP32 master copy, layout of pct_net.h)"""
    import torch
    rng = np.random.default_rng(seed)

    def xavier(fi, fo):
        lim = math.sqrt(6.0 / (fi + fo))
        return rng.uniform(-lim, lim, size=(fi, fo))
    p = np.concatenate([xavier(NET_IN, 64).ravel(), np.zeros(64), xavier(64, 64).ravel(), np.zeros(64),
                        xavier(64, NET_OUT).ravel(), np.zeros(NET_OUT)]).astype(np.float32)
    assert p.size == NET_PARAMS
    return torch.tensor(p, device=device)


def features(ctx, hist, start=0, count=None, targets=True):
    """network inputs [count][12] and exact-MLP targets [count][8] (or None)"""
    torch = ctx.torch
    count = hist.n - start if count is None else count
    F = torch.empty((max(count, 1), NET_IN), dtype=torch.float32, device=ctx.device)
    Y = torch.empty((max(count, 1), NET_OUT), dtype=torch.float32, device=ctx.device) if targets else None
    hs = hist.soa()
    ctx._check(ctx.L.pct_net_features(ctx._ctx, ctypes.byref(hs), start, count, _ptr(F), _ptr(Y)), "pct_net_features")
    return F[:count], (Y[:count] if targets else None)


def gradient(ctx, params, F, Y):
    """(gradient[5512], loss) of the mean squared error over the rows of F, Y"""
    torch = ctx.torch
    g = torch.empty(NET_PARAMS, dtype=torch.float32, device=ctx.device)
    loss = ctypes.c_double(0.0)
    ctx._check(ctx.L.pct_net_gradient(ctx._ctx, _ptr(params), _ptr(F), _ptr(Y), F.shape[0], _ptr(g),
                                      ctypes.byref(loss)), "pct_net_gradient")
    return g, loss.value


def train(ctx, params, F, Y, epochs=12, batch=8192, lr=3e-3, lr_min=1e-5, seed=1, betas=(0.9, 0.999), eps=1e-8):
    """Adam on the device; params updated in place; returns the per-epoch loss"""
    cfg = PctNetTrainCfg(epochs, batch, lr, lr_min, betas[0], betas[1], eps, seed)
    out = (ctypes.c_double * epochs)()
    ctx._check(ctx.L.pct_net_train(ctx._ctx, _ptr(params), _ptr(F), _ptr(Y), F.shape[0], ctypes.byref(cfg), out),
               "pct_net_train")
    return list(out)


def input_range(F, margin=0.0):
    """per-input [lo, hi] of a training set (host lists), widened by margin * span"""
    lo = F.min(dim=0).values.cpu().numpy().astype(np.float64)
    hi = F.max(dim=0).values.cpu().numpy().astype(np.float64)
    span = hi - lo
    return list(lo - margin * span), list(hi + margin * span)


def install(ctx, params, lo, hi):
    """make params the network of ctx's surrogate walk (clears its calibration)"""
    a = (ctypes.c_float * NET_IN)(*lo)
    b = (ctypes.c_float * NET_IN)(*hi)
    ctx._check(ctx.L.pct_net_set(ctx._ctx, _ptr(params), a, b), "pct_net_set")


def install_hermite(ctx, lo, hi):
    """the surrogate walk without the network: the Hermite path alone"""
    a = (ctypes.c_float * NET_IN)(*lo)
    b = (ctypes.c_float * NET_IN)(*hi)
    ctx._check(ctx.L.pct_net_set_hermite(ctx._ctx, a, b), "pct_net_set_hermite")


def forward(ctx, F):
    """outputs [n][8] of the installed network (the surrogate walk's kernel)"""
    torch = ctx.torch
    out = torch.empty((max(F.shape[0], 1), NET_OUT), dtype=torch.float32, device=ctx.device)
    ctx._check(ctx.L.pct_net_forward(ctx._ctx, _ptr(F), F.shape[0], _ptr(out)), "pct_net_forward")
    return out[:F.shape[0]]


def calibrate(ctx, hist, stride=64, tol=0.0, min_samples=8, start=0, count=None, rule="deviation",
              threshold=None):
    """calibration check on every stride-th history; rule "deviation": a cell
    falls back when its 99th-percentile path deviation exceeds tol cm (tol <= 0:
    measure only); rule "lists": when the share of its sampled histories whose
    voxel list differs from the exact walk's exceeds `threshold` (0 = any
    disagreement; None or < 0: measure only). Returns a dict of the results."""
    count = hist.n - start if count is None else count
    c = PctNetCalib()
    hs = hist.soa()
    if rule == "deviation":
        code, thr = _lib.PCT_NET_RULE_DEVIATION, (tol if tol > 0 else -1.0)
    elif rule == "lists":
        code, thr = _lib.PCT_NET_RULE_LISTS, (-1.0 if threshold is None else threshold)
    else:
        raise ValueError(rule)
    ctx._check(ctx.L.pct_net_calibrate_rule(ctx._ctx, ctypes.byref(hs), start, count, stride, code, float(thr),
                                            int(min_samples), ctypes.byref(c)), "pct_net_calibrate_rule")
    return {"samples": c.samples, "list_mismatch": c.list_mismatch, "in_range": c.in_range,
            "list_mismatch_rate": c.list_mismatch / max(c.samples, 1),
            "surrogate": c.surrogate, "surrogate_mismatch": c.surrogate_mismatch,
            "fallback_rate": 1.0 - c.surrogate / max(c.samples, 1),
            "dev_median_cm": c.dev_median, "dev_p99_cm": c.dev_p99, "dev_max_cm": c.dev_max,
            "cells_fallback": c.cells_fallback, "cell_samples": list(c.cell_samples),
            "cell_p99_cm": [float(v) for v in c.cell_p99], "cell_fallback": list(c.cell_fallback),
            "cell_list_mismatch": [float(v) for v in c.cell_list_mismatch]}


def fallback_stats(ctx):
    """(fallbacks, histories) of the surrogate MLP passes since the last call"""
    a, b = ctypes.c_int64(0), ctypes.c_int64(0)
    ctx._check(ctx.L.pct_net_fallback_stats(ctx._ctx, ctypes.byref(a), ctypes.byref(b)), "pct_net_fallback_stats")
    return a.value, b.value


# ---------------------------------------------------------------------------
# training set: exact-MLP paths of random-ellipsoid phantoms
# ---------------------------------------------------------------------------
def training_scans(n_scans=8, histories_per_scan=1_500_000, seed=configs.SEED_BASE + 5):
    """random-ellipsoid phantom scans of the batched configuration's family
    (grids 64..256 columns, 8..48 slices, RSP in U[0.2, 1.8]) with a seed of
    their own, so that no evaluation scan is among them"""
    scans = configs.batch_scans(seed=seed)[:n_scans]
    for sc in scans:
        sc.synth.n_histories = histories_per_scan
    return scans


def scan_histories(ctx, synth):
    """generate -> recon-volume intersections -> cylinder hull -> MLP endpoints:
    the histories the MLP walks"""
    torch = ctx.torch
    hist, missed = ctx.recon_volume_intersections(ctx.generate(synth))
    hist = ctx.compact(hist, (missed == 0).to(torch.uint8))
    hull = ctx.hull_init()
    return ctx.compact(hist, ctx.mlp_endpoints(hist, hull))


def training_set(n_paths=10_000_000, scans=None, log=None):
    """(features, targets) of n_paths exact-MLP paths (device tensors)"""
    import torch
    from . import Context
    scans = training_scans() if scans is None else scans
    Fs, Ys, have = [], [], 0
    for sc in scans:
        if have >= n_paths:
            break
        ctx = Context(sc.params)
        try:
            hist = scan_histories(ctx, sc.synth)
            take = min(hist.n, n_paths - have)
            F, Y = features(ctx, hist, 0, take)
            ctx.sync()
            Fs.append(F.clone())
            Ys.append(Y.clone())
            have += take
            if log:
                log(f"training set: {sc.name}: {take} paths")
            del hist, F, Y
        finally:
            ctx.close()
    if have < n_paths:
        raise RuntimeError(f"training scans gave {have} paths, {n_paths} wanted")
    return torch.cat(Fs), torch.cat(Ys)


@dataclass
class PathNetwork:
    params: object                      # device float32 [5512]
    lo: list = field(default_factory=list)
    hi: list = field(default_factory=list)
    epoch_loss: list = field(default_factory=list)

    def install(self, ctx):
        install(ctx, self.params, self.lo, self.hi)

    def save(self, path):
        np.savez(path, params=self.params.cpu().numpy(), lo=np.array(self.lo), hi=np.array(self.hi),
                 epoch_loss=np.array(self.epoch_loss))

    @staticmethod
    def load(path, device="cuda"):
        import torch
        d = np.load(path, allow_pickle=False)
        return PathNetwork(torch.tensor(d["params"], device=device), list(d["lo"]), list(d["hi"]),
                           list(d["epoch_loss"]))


def train_network(ctx, F, Y, epochs=12, batch=8192, lr=3e-3, seed=0, range_margin=0.02):
    """a network trained on (F, Y) with the library's kernels"""
    params = init_params(seed, ctx.device)
    losses = train(ctx, params, F, Y, epochs=epochs, batch=batch, lr=lr, seed=seed + 1)
    lo, hi = input_range(F, range_margin)
    return PathNetwork(params, lo, hi, losses)


def main(argv=None):
    """Train a path network on exact-MLP paths of random-ellipsoid phantoms and
    save it:  python -m pct_amd.learned --paths 10000000 --epochs 20 --save net.npz"""
    import argparse
    import time
    import torch
    from . import Context
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--paths", type=int, default=configs.LEARNED_TRAIN_PATHS)
    ap.add_argument("--scans", type=int, default=configs.LEARNED_TRAIN_SCANS)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--save", default=None)
    a = ap.parse_args(argv)
    per_scan = int(a.paths / a.scans * 1.25) + 1
    t0 = time.time()
    F, Y = training_set(a.paths, training_scans(a.scans, per_scan), log=print)
    torch.cuda.synchronize()
    t1 = time.time()
    ctx = Context(configs.params("cfg2"))
    net = train_network(ctx, F, Y, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed)
    torch.cuda.synchronize()
    t2 = time.time()
    print(f"training set {t1 - t0:.2f} s, training {t2 - t1:.2f} s, loss per epoch {net.epoch_loss}", flush=True)
    if a.save:
        net.save(a.save)
    ctx.close()


if __name__ == "__main__":
    main()
