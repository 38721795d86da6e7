"""Plain device-to-device copies of a known size, used to calibrate the
rocprofv3 FETCH_SIZE / WRITE_SIZE counters against bytes actually moved.

    python -m torch_amd.hbm_probe [--mib 1024] [--reps 5]

Each repetition copies a float64 tensor of `mib` MiB into another one, so a
dispatch reads and writes exactly that many bytes; the printed GB/s is the
event-timed copy rate.
"""
from __future__ import annotations

import argparse
import json


def main():
    import torch

    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=5)
    a = ap.parse_args()
    n = a.mib * (1 << 20) // 8
    src = torch.arange(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(a.reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.reps
    nbytes = n * 8
    print(json.dumps({"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms,
                      "GBs_read_plus_write": 2 * nbytes / (ms * 1e-3) / 1e9}), flush=True)


if __name__ == "__main__":
    main()
