"""Trains the cooling surrogate on the GPU and writes its parameters.

    python -m torch_amd.train_surrogate [--samples N] [--chunk C] [--epochs E]
                                        [--out torch_amd/data/surrogate_cooling.npz]

1. Samples: cells of the cooling workload (torch_amd.cooling_inputs: density
   log-uniform 1e-24..1e-19 g/cm^3, T 10..1e5 K, x in [0, 1], ionising flux
   1e6..1e14 cm^-2 s^-1, tau in [0, 10], dt 1e7..1e11 s) drawn with seeds
   1000 + chunk, disjoint from the evaluation seeds; with --workload cloud,
   cells of the operator-split cloud's conditions (sample_cgs_cloud: densities
   to 3e-16 g/cm^3, unlit cells, optical depths to 3e3, dt to 1e13 s) drawn
   with seeds 5000 + chunk.  --max-subcycles drops the samples whose
   subcycle count (torchgpu_thermo_subcycles, known before integrating)
   exceeds it, and the envelope's seventh bound keeps such cells on the FP64
   fallback at run time.
2. Targets: the FP64 subcycled integrator (torchgpu_thermo_cells) on the
   same GPU; the cells with more than `nexact` subcycles (the ones the
   network replaces) are extracted on the device with their raw features by
   torchgpu_surrogate_training_data.  Output 0's target is the residual
   log10(T_sub / Tc) of the prepass's coarse end-temperature estimate Tc,
   divided by the residual scale sigma = 0.01 / sqrt(subcycles) + 2e-4.
3. Features: log10 rho, log10 T, x, log10 F_FUV, A_V, log10 dt and, from
   the FP64 prepass, log10 of the subcycle count and log10(Tc / T0);
   normalisation: per-feature mean and standard deviation of
   the training set; the training envelope is the [min, max] of the first
   seven.  Batches oversample cells in proportion to 1 + w sqrt(subcycles)
   (an FP64 fallback of a long cell costs its whole subcycle); the last
   2^18 queued cells are held out for validation.
4. Training: Adam with a cosine learning-rate decay over shuffled batches,
   every step a torchgpu_mlp_train_step (the library's BF16-MFMA forward and
   backward kernels with FP32 accumulation).  Output 2 learns the log10 of
   the current |relative T error| of output 0 from epoch --err-head-from on; the cooling
   step gates on it.

The output file holds W, b (fp32), mu, isd, lo, hi, nexact and the training
record; torch_amd.surrogate.load() reads it.
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = {"c3": os.path.join(HERE, "data", "surrogate_cooling.npz"),
               "cloud": os.path.join(HERE, "data", "surrogate_cloud.npz")}
SEED0 = {"c3": 1000, "cloud": 5000}


def log(*a):
    print(*a, flush=True)


def workload_inputs(n: int, seed: int, device="cuda", workload: str = "c3"):
    """[9][n] float64 device tensor of cooling-workload cells in code units
    (workload "c3": the 2^24-cell cooling workload's distribution; "cloud":
    the operator-split cloud's conditions)."""
    import torch

    import torch_amd as ta
    from torch_amd import cooling_inputs as ci

    scales = (ci.DENSITY_SCALE, ci.PRESSURE_SCALE, ci.TIME_SCALE)
    tc = lambda v, m, l, t: ta.units(scales, v, m, l, t)  # noqa: E731
    x = ci.to_code_inputs(ci.SAMPLERS[workload](n, seed=seed), tc)
    return torch.from_numpy(np.ascontiguousarray(x.T)).to(device)


def generate(cool, mlp, samples: int, chunk: int, nexact: int, seed0: int = 1000, workload: str = "c3",
             max_subcycles: int = 0):
    """Raw features [k][8] and targets [k][3] (device fp32) of the cells past
    nexact subcycles in `samples` workload cells (those with more than
    max_subcycles subcycles dropped first, when it is > 0)."""
    import torch

    from torch_amd import surrogate as sg

    big = [1e30] * 7
    sg.set_surrogate(cool, mlp, sg.make_cfg([0.0] * 8, [1.0] * 8, [-x for x in big], big, gate=0.0, nexact=nexact))
    raws, tgts = [], []
    done, i = 0, 0
    while done < samples:
        n = min(chunk, samples - done)
        t0 = time.perf_counter()
        inp = workload_inputs(n, seed0 + i, workload=workload)
        dropped = 0
        if max_subcycles > 0:
            cnt = sg.subcycles(cool, inp)
            keep = (cnt >= 0) & (cnt <= max_subcycles)
            dropped = int(n - keep.sum())
            inp = inp[:, keep].contiguous()
        out, nsub = cool.thermo_cells(inp)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        _, _, tgt, raw = sg.training_data(cool, inp, out)
        raws.append(raw.clone())
        tgts.append(tgt.clone())
        log(f"chunk {i}: {n} cells, {dropped} past {max_subcycles} subcycles dropped, {tgt.shape[0]} past "
            f"{nexact} subcycles, {nsub:.3e} subcycles, integrator {t1 - t0:.1f}s")
        del inp, out
        done += n
        i += 1
    return torch.cat(raws), torch.cat(tgts)


def main():
    import torch

    import torch_amd as ta
    from torch_amd import surrogate as sg

    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--samples", type=int, default=1 << 24)
    ap.add_argument("--chunk", type=int, default=1 << 24)
    ap.add_argument("--epochs", type=int, default=100)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--nexact", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--err-head-from", type=int, default=5)
    ap.add_argument("--long-weight", type=float, default=0.05,
                    help="sampling weight 1 + w sqrt(subcycles): oversamples long subcycles")
    ap.add_argument("--workload", choices=sorted(DEFAULT_OUT), default="c3")
    ap.add_argument("--max-subcycles", type=int, default=0,
                    help="drop samples with more subcycles (0: keep all); bounds the integrator's longest cell")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    out_path = a.out or DEFAULT_OUT[a.workload]

    t_start = time.perf_counter()
    cool = ta.Cooling(ta.thermo_params())
    mlp = sg.MLP(max_batch=a.batch)
    raw, tgt = generate(cool, mlp, a.samples, a.chunk, a.nexact, seed0=SEED0[a.workload], workload=a.workload,
                        max_subcycles=a.max_subcycles)
    k = raw.shape[0]
    mu = raw.mean(0).double()
    sd = raw.double().std(0).clamp_min(1e-3)
    lo = raw[:, :7].min(0).values.double()
    hi = raw[:, :7].max(0).values.double()
    feat = ((raw.double() - mu) / sd).float().contiguous()
    # long subcycles weigh more: they are the cells the surrogate saves time on
    # (an FP64 fallback of one of them costs its whole subcycle)
    sample_weight = 1.0 + a.long_weight * torch.pow(10.0, 0.5 * raw[:, 6].double())
    # held-out cells: the last chunk's first 2^18 queued cells, long ones
    # (>= 1e6 subcycles) reported separately
    nv = min(1 << 18, k // 8)
    vidx = torch.arange(k - nv, k, device=feat.device)
    long_v = vidx[raw[k - nv:, 6] >= 6.0]
    sample_weight[k - nv:] = 0.0
    sigma = sg.residual_scale(raw[:, 6])
    val_sets = (feat[vidx].contiguous(), tgt[vidx].contiguous(), sigma[vidx])
    val_long = (feat[long_v].contiguous(), tgt[long_v].contiguous(), sigma[long_v]) if long_v.numel() else None
    del raw
    log(f"training set: {k} cells; mu {mu.tolist()} sd {sd.tolist()}")
    W, b = sg.init_params(a.seed)
    mlp.set_params(W, b)
    t0 = time.perf_counter()
    hist = sg.train(mlp, feat, tgt.contiguous(), epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed, log=log,
                    sample_weight=sample_weight, err_head_from=a.err_head_from, val=val_sets)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    val_all = sg.validate(mlp, *val_sets)
    val_l = sg.validate(mlp, *val_long) if val_long is not None else None
    log(f"held-out: all {val_all}; >=1e6 subcycles ({0 if val_long is None else val_long[0].shape[0]}) {val_l}")
    W, b = mlp.get_params()
    os.makedirs(os.path.dirname(os.path.abspath(out_path)), exist_ok=True)
    record = {"workload": a.workload, "max_subcycles": a.max_subcycles,
              "samples": a.samples, "train_cells": int(k), "epochs": a.epochs, "batch": a.batch, "lr": a.lr,
              "seed": a.seed, "long_weight": a.long_weight, "err_head_from": a.err_head_from,
              "held_out": val_all, "held_out_long": val_l, "loss_history": hist, "train_seconds": t_train,
              "total_seconds": time.perf_counter() - t_start}
    np.savez(out_path, W=W, b=b, mu=mu.cpu().numpy(), isd=(1.0 / sd).cpu().numpy(), lo=lo.cpu().numpy(),
             hi=hi.cpu().numpy(), nexact=np.int32(a.nexact), record=np.array(json.dumps(record)))
    log(json.dumps(record))


if __name__ == "__main__":
    main()
