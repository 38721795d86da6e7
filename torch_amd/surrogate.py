"""Cooling surrogate: the MLP of torchgpu_mlp_* and its training loop.

The network (8 -> 4 x 128 SiLU -> 3, bf16 weights, fp32 accumulation) is
trained on the GPU by the library's own BF16-MFMA forward/backward kernels
and Adam (torchgpu_mlp_train_step) on samples the FP64 subcycled integrator
(torchgpu_thermo_cells) generates; the cooling operator then uses it through
torchgpu_thermo_cells_surrogate with the FP64 integrator as gate and
fallback.  This module only sequences those calls.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from . import Cooling, TorchGpuError, default_stream, lib

NIN, NH, NOUT = 8, 128, 3
NW = NH * NIN + 3 * NH * NH + NOUT * NH
NB = 4 * NH + NOUT


class SurrogateCfg(ctypes.Structure):
    _fields_ = [("mu", ctypes.c_float * 8), ("isd", ctypes.c_float * 8), ("lo", ctypes.c_double * 7),
                ("hi", ctypes.c_double * 7), ("gate", ctypes.c_double), ("nexact", ctypes.c_int), ("pad", ctypes.c_int)]


_bound = False


def _bind():
    global _bound
    if _bound:
        return
    L = lib()
    vp, i, f, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_int64
    L.torchgpu_mlp_create.argtypes = [i, vp, ctypes.POINTER(vp)]
    L.torchgpu_mlp_destroy.argtypes = [vp]
    L.torchgpu_mlp_destroy.restype = None
    L.torchgpu_mlp_set_params.argtypes = [vp, vp, vp]
    L.torchgpu_mlp_get_params.argtypes = [vp, vp, vp]
    L.torchgpu_mlp_forward.argtypes = [vp, vp, i64, vp]
    L.torchgpu_mlp_loss_grad.argtypes = [vp, vp, vp, i, vp, i, ctypes.POINTER(f), vp, vp]
    L.torchgpu_mlp_train_step.argtypes = [vp, vp, vp, i, vp, i, f, f, ctypes.POINTER(f)]
    L.torchgpu_cooling_set_surrogate.argtypes = [vp, vp, ctypes.POINTER(SurrogateCfg)]
    L.torchgpu_thermo_cells_surrogate.argtypes = [vp, i64, vp, vp, vp, ctypes.POINTER(i64)]
    L.torchgpu_surrogate_training_data.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, ctypes.POINTER(i64)]
    L.torchgpu_thermo_subcycles.argtypes = [vp, i64, vp, vp]
    _bound = True


def _check(rc, what):
    if rc:
        raise TorchGpuError(rc, what)


def init_params(seed: int = 0):
    """He-normal weights, zero biases."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dims = [NIN, NH, NH, NH, NH, NOUT]
    Ws, bs = [], []
    for L in range(1, 6):
        fan_in = dims[L - 1]
        W = rng.normal(0.0, math.sqrt(2.0 / fan_in), (dims[L], dims[L - 1]))
        Ws.append(W)
        bs.append(np.zeros(dims[L]))
    return pack(Ws, bs)


def pack(Ws, bs):
    return (np.concatenate([w.ravel() for w in Ws]).astype(np.float32),
            np.concatenate([b.ravel() for b in bs]).astype(np.float32))


def unpack(W, b):
    dims = [NIN, NH, NH, NH, NH, NOUT]
    Ws, bs, ow, ob = [], [], 0, 0
    for L in range(1, 6):
        n = dims[L] * dims[L - 1]
        Ws.append(np.asarray(W[ow:ow + n], dtype=np.float64).reshape(dims[L], dims[L - 1]))
        bs.append(np.asarray(b[ob:ob + dims[L]], dtype=np.float64))
        ow += n
        ob += dims[L]
    return Ws, bs


class MLP:
    def __init__(self, max_batch: int = 65536, stream: int | None = None):
        _bind()
        self._h = ctypes.c_void_p()
        _check(lib().torchgpu_mlp_create(max_batch, ctypes.c_void_p(default_stream(stream)), ctypes.byref(self._h)),
               "mlp create")

    def close(self):
        if self._h:
            lib().torchgpu_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, W: np.ndarray, b: np.ndarray):
        W = np.ascontiguousarray(W, dtype=np.float32)
        b = np.ascontiguousarray(b, dtype=np.float32)
        assert W.size == NW and b.size == NB
        _check(lib().torchgpu_mlp_set_params(self._h, W.ctypes.data, b.ctypes.data), "set_params")

    def get_params(self):
        W = np.empty(NW, np.float32)
        b = np.empty(NB, np.float32)
        _check(lib().torchgpu_mlp_get_params(self._h, W.ctypes.data, b.ctypes.data), "get_params")
        return W, b

    def forward(self, feat, out=None):
        import torch

        B = feat.shape[0]
        if out is None:
            out = torch.empty((B, NOUT), dtype=torch.float32, device=feat.device)
        _check(lib().torchgpu_mlp_forward(self._h, feat.data_ptr(), B, out.data_ptr()), "forward")
        return out

    def loss_grad(self, feat, target, weights=(1.0, 1.0, 1.0), err_head=False):
        w = (ctypes.c_float * 3)(*weights)
        loss = ctypes.c_float()
        gW = np.empty(NW, np.float32)
        gb = np.empty(NB, np.float32)
        _check(lib().torchgpu_mlp_loss_grad(self._h, feat.data_ptr(), target.data_ptr(), feat.shape[0], w,
                                            int(err_head), ctypes.byref(loss), gW.ctypes.data, gb.ctypes.data),
               "loss_grad")
        return loss.value, gW, gb

    def train_step(self, feat, target, lr, weights=(1.0, 1.0, 1.0), err_head=True, weight_decay=0.0,
                   want_loss=False):
        w = (ctypes.c_float * 3)(*weights)
        loss = ctypes.c_float()
        _check(lib().torchgpu_mlp_train_step(self._h, feat.data_ptr(), target.data_ptr(), feat.shape[0], w,
                                             int(err_head), lr, weight_decay,
                                             ctypes.byref(loss) if want_loss else None), "train_step")
        return loss.value if want_loss else None


def set_surrogate(cool: Cooling, mlp: MLP, cfg: SurrogateCfg):
    _bind()
    _check(lib().torchgpu_cooling_set_surrogate(cool._h, mlp._h, ctypes.byref(cfg)), "set_surrogate")


def make_cfg(mu, isd, lo, hi, gate=0.01, nexact=4) -> SurrogateCfg:
    """Configuration with the training envelope [lo, hi] over the first seven
    raw features; bounds past the given ones (a file with six) are open."""
    c = SurrogateCfg()
    for i in range(8):
        c.mu[i] = float(mu[i]) if i < len(mu) else 0.0
        c.isd[i] = float(isd[i]) if i < len(isd) else 0.0
    for i in range(7):
        c.lo[i] = float(lo[i]) if i < len(lo) else -math.inf
        c.hi[i] = float(hi[i]) if i < len(hi) else math.inf
    c.gate = gate
    c.nexact = nexact
    return c


def thermo_cells_surrogate(cool: Cooling, inp, out=None, path=None):
    """Surrogate cooling step on [9][n] device inputs; returns (out [5][n],
    path uint8 [n], (n_surrogate, n_fallback))."""
    import torch

    _bind()
    n = inp.shape[1]
    if out is None:
        out = torch.empty((5, n), dtype=torch.float64, device=inp.device)
    if path is None:
        path = torch.empty(n, dtype=torch.uint8, device=inp.device)
    counts = (ctypes.c_int64 * 2)()
    _check(lib().torchgpu_thermo_cells_surrogate(cool._h, n, inp.data_ptr(), out.data_ptr(), path.data_ptr(), counts),
           "thermo_cells_surrogate")
    return out, path, (counts[0], counts[1])


def subcycles(cool: Cooling, inp):
    """Per-cell subcycle count (device int64 [n]) of [9][n] device inputs:
    the explicit steps the FP64 integrator would take after the first, 0 for
    a cell that is not subcycled, -1 past the int range
    (torchgpu_thermo_subcycles; nothing is integrated)."""
    import torch

    _bind()
    n = inp.shape[1]
    assert inp.dtype == torch.float64 and inp.is_cuda and inp.is_contiguous() and inp.shape[0] == 9
    cnt = torch.empty(n, dtype=torch.int64, device=inp.device)
    _check(lib().torchgpu_thermo_subcycles(cool._h, n, inp.data_ptr(), cnt.data_ptr()), "thermo_subcycles")
    return cnt


def training_data(cool: Cooling, inp, out_fp64):
    """Features/targets of the cells past nexact subcycles (device tensors)."""
    import torch

    _bind()
    n = inp.shape[1]
    dev = inp.device
    lst = torch.empty(n, dtype=torch.int64, device=dev)
    feat = torch.empty((n, 8), dtype=torch.float32, device=dev)
    tgt = torch.empty((n, 3), dtype=torch.float32, device=dev)
    raw = torch.empty((n, 8), dtype=torch.float32, device=dev)
    cnt = ctypes.c_int64()
    _check(lib().torchgpu_surrogate_training_data(cool._h, n, inp.data_ptr(), out_fp64.data_ptr(), lst.data_ptr(),
                                                  feat.data_ptr(), tgt.data_ptr(), raw.data_ptr(), ctypes.byref(cnt)),
           "training_data")
    k = cnt.value
    return lst[:k], feat[:k], tgt[:k], raw[:k]


def residual_scale(log10_nsub):
    """Scale of the network's output 0 (log10 units) for subcycle counts
    10^log10_nsub: 0.01 / sqrt(n) + 2e-4, as the cooling kernels use it."""
    import torch

    return 0.01 * torch.pow(10.0, -0.5 * log10_nsub.double()) + 2.0e-4


def validate(mlp: MLP, feat, target, sigma, gate: float = 0.01) -> dict:
    """Relative T error |10^(sigma (y0 - t0)) - 1| of the network on held-out
    cells (sigma: their residual scales), and the fraction its error head
    accepts at `gate` (estimated error sigma 10^y2)."""
    import torch

    y = mlp.forward(feat)
    torch.cuda.synchronize()
    err = (torch.pow(10.0, sigma * (y[:, 0] - target[:, 0]).double()) - 1.0).abs()
    acc = sigma * torch.pow(10.0, y[:, 2].double()) <= gate
    q = torch.quantile(err, torch.tensor([0.5, 0.9, 0.99], dtype=torch.float64, device=err.device))
    out = {"median": float(q[0]), "p90": float(q[1]), "p99": float(q[2]), "accepted": float(acc.double().mean())}
    if int(acc.sum()):
        out["accepted_p99"] = float(torch.quantile(err[acc], 0.99))
    return out


def train(mlp: MLP, feat, target, epochs: int = 10, batch: int = 65536, lr: float = 2e-3, seed: int = 0,
          err_head_from: int = 1, weights=(1.0, 1.0, 0.05), log=None, sample_weight=None, val=None,
          val_every: int = 10):
    """Adam over mini-batches with a cosine learning-rate decay.  Each epoch
    draws n samples: a permutation, or, with `sample_weight`, samples drawn
    with replacement in proportion to it (inverse-CDF).  The error head is
    trained from epoch `err_head_from` on."""
    import torch

    n = feat.shape[0]
    g = torch.Generator(device=feat.device)
    g.manual_seed(seed)
    cdf = None
    if sample_weight is not None:
        cdf = torch.cumsum(sample_weight.double(), 0)
        cdf = cdf / cdf[-1]
    steps_per_epoch = max(1, n // batch)
    total = epochs * steps_per_epoch
    step = 0
    t0 = time.perf_counter()
    hist = []
    for ep in range(epochs):
        if cdf is None:
            perm = torch.randperm(n, generator=g, device=feat.device)
        else:
            u = torch.rand(n, generator=g, device=feat.device, dtype=torch.float64)
            perm = torch.searchsorted(cdf, u).clamp_max(n - 1)
        losses = []
        for s in range(steps_per_epoch):
            idx = perm[s * batch:(s + 1) * batch]
            fb = feat[idx].contiguous()
            tb = target[idx].contiguous()
            cur = lr * 0.5 * (1.0 + math.cos(math.pi * step / total))
            want = (s == steps_per_epoch - 1)
            w = weights if ep >= err_head_from else (weights[0], weights[1], 0.0)
            l = mlp.train_step(fb, tb, cur, weights=w, err_head=True, want_loss=want)
            if want:
                losses.append(l)
            step += 1
        hist.append(losses[-1] if losses else float("nan"))
        if log:
            msg = f"epoch {ep}: loss {hist[-1]:.3e} ({time.perf_counter() - t0:.1f}s)"
            if val is not None and (ep % val_every == val_every - 1 or ep == epochs - 1):
                msg += " val " + " ".join(f"{k} {v:.4g}" for k, v in validate(mlp, *val).items())
            log(msg)
    return hist


def load(path: str):
    """(W, b, cfg fields) of a file written by torch_amd.train_surrogate."""
    f = np.load(path, allow_pickle=False)
    return {k: f[k] for k in f.files}


def attach(cool: Cooling, path: str, gate: float = 0.01, mlp: MLP | None = None, max_batch: int = 65536) -> MLP:
    """Loads trained parameters into an MLP handle and makes it the cooling
    operator's surrogate (estimated-error gate `gate`).  Returns the handle,
    which must outlive its use by `cool`."""
    d = load(path)
    if mlp is None:
        mlp = MLP(max_batch=max_batch)
    mlp.set_params(d["W"], d["b"])
    mu = list(d["mu"]) + [0.0] * (8 - len(d["mu"]))
    isd = list(d["isd"]) + [0.0] * (8 - len(d["isd"]))
    set_surrogate(cool, mlp, make_cfg(mu, isd, d["lo"], d["hi"], gate=gate, nexact=int(d["nexact"])))
    return mlp


def compare(cool: Cooling, inp, return_fp64: bool = False):
    """Runs the FP64 subcycled step and the surrogate step on the same cells
    ([9][n] device inputs) and returns their event-timed durations and the
    surrogate's relative temperature error against FP64 (and the FP64 output
    when `return_fp64`).  The temperature
    ratio equals the ratio of the subcycled pressures (same density and
    ionisation), out[4]."""
    import torch

    s = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record(s)
    ref, nsub = cool.thermo_cells(inp)
    ev[1].record(s)
    ev[2].record(s)
    out, path, (n_mlp, n_fb) = thermo_cells_surrogate(cool, inp)
    ev[3].record(s)
    torch.cuda.synchronize()
    rel = ((out[4] / ref[4]) - 1.0).abs()
    p = path
    acc = p == 1
    q = torch.tensor([0.5, 0.99], dtype=torch.float64, device=rel.device)

    def quant(x):
        if x.numel() == 0:
            return [0.0, 0.0]
        # torch.quantile is limited to 2^24 elements
        if x.numel() > (1 << 24):
            x = x[torch.randperm(x.numel(), device=x.device)[: 1 << 24]]
        return [float(v) for v in torch.quantile(x, q)]

    all_q, acc_q = quant(rel), quant(rel[acc])
    n = inp.shape[1]
    res = {
        "cells": int(n),
        "fp64_ms": ev[0].elapsed_time(ev[1]),
        "surrogate_ms": ev[2].elapsed_time(ev[3]),
        "fp64_subcycles": int(nsub),
        "exact": int((p == 0).sum()), "mlp_accepted": int(acc.sum()), "fallback": int((p == 2).sum()),
        "mlp_queued": int(n_mlp),
        "rel_T_err_all": {"median": all_q[0], "p99": all_q[1], "max": float(rel.max())},
        "rel_T_err_mlp_cells": {"median": acc_q[0], "p99": acc_q[1],
                                "max": float(rel[acc].max()) if int(acc.sum()) else 0.0},
    }
    return (res, ref) if return_fp64 else res
