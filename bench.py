"""Benchmark of the mapstencil hot path on one MI355X.

Default workload: Game of Life on an 8192 x 8192 UInt8 grid (Bernoulli(0.3),
seed 2), Moore(1), Wrap, Halo padding, B3/S23, as a device-resident
SwitchingStencilArray.  One step is one generation (one `mapstencil!` call);
the timed steps run as replays of a captured hipGraph.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "stencils.jl_amd"))
sys.path.insert(0, str(REPO / "oracle"))

METRIC = "cells/s + % HBM roofline, batched/served vs plain mapstencil; p99 latency"
HBM_PEAK_GBS = 8000.0


def life_grid(n: int, seed: int = 2, p: float = 0.3) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    return (rng.random((n, n)) < p).astype(np.uint8)


def device_switching(A: np.ndarray, sh, torch):
    dev = torch.device("cuda", torch.cuda.current_device())
    t = sh.colmajor_empty(A.shape, torch.uint8, dev)
    t.copy_(torch.from_numpy(np.ascontiguousarray(A)).to(dev))
    return sh.SwitchingStencilArray(t, sh.Moore(1), boundary=sh.Wrap(), padding=sh.Halo())


def time_steps(sh, torch, S, rule, steps: int, graph: bool):
    """Wall time and HIP-event time of `steps` generations (inputs resident)."""
    stream = torch.cuda.current_stream()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record(stream)
    S = sh.mapstencil_steps(rule, S, steps, graph=graph, flags=sh._abi.ASYNC)
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return S, wall, start.elapsed_time(end) / 1e3


def cpu_baseline(A: np.ndarray, target_s: float = 10.0):
    """The oracle (plain-C restatement of the reference) on all host cores,
    on a bounded number of generations of the same grid."""
    import oracle
    import stencils_hip as sh
    cores = oracle.host_cores()
    S = sh.SwitchingStencilArray(np.asfortranarray(A), sh.Moore(1), boundary=sh.Wrap(),
                                 padding=sh.Halo())
    t0 = time.perf_counter()
    S = oracle.mapstencil_steps(sh.LifeRule(), S, 1, nthreads=cores)
    t1 = time.perf_counter() - t0
    steps = int(max(2, min(200, target_s / max(t1, 1e-3))))
    t0 = time.perf_counter()
    oracle.mapstencil_steps(sh.LifeRule(), S, steps, nthreads=cores)
    dt = time.perf_counter() - t0
    cells = A.size * steps
    return {"value": cells / dt, "unit": "cells/s", "cores": cores, "kind": "port",
            "sample": f"{steps} generations of the same {A.shape[0]}x{A.shape[1]} grid "
                      f"({dt:.1f} s, oracle/stencil_oracle.c on {cores} threads)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported "
              "(the batched and served workloads shard without exchange; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import stencils_hip as sh
    sh.load_library()
    torch.cuda.set_device(int(os.environ.get("LOCAL_RANK", 0)))

    n = args.n
    A = life_grid(n)
    rule = sh.LifeRule()
    cells = n * n

    S = device_switching(A, sh, torch)
    S, _, _ = time_steps(sh, torch, S, rule, args.warmup, graph=True)
    S, wall, ev = time_steps(sh, torch, S, rule, args.steps, graph=True)
    per_step_ev = ev / args.steps
    bytes_per_step = 2 * cells  # 1 B read + 1 B written per interior cell
    achieved = bytes_per_step / per_step_ev / 1e9

    # measured stream-copy bandwidth on the same bytes (one read + one write)
    src_t = torch.empty(cells, dtype=torch.uint8, device="cuda").fill_(1)
    dst_t = torch.empty_like(src_t)
    for _ in range(5):
        dst_t.copy_(src_t)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(50):
        dst_t.copy_(src_t)
    e1.record()
    torch.cuda.synchronize()
    copy_gbs = 2 * cells / (e0.elapsed_time(e1) / 1e3 / 50) / 1e9
    del src_t, dst_t

    # the same steps as plain per-step launches (no graph)
    S, _, _ = time_steps(sh, torch, S, rule, 10, graph=False)
    S, wall_plain, _ = time_steps(sh, torch, S, rule, args.steps, graph=False)

    out = {
        "metric": METRIC,
        "value": cells * args.steps / wall,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": f"game-of-life {n}x{n} UInt8 Bernoulli(0.3) seed 2, Moore(1), Wrap, "
                               "Halo, B3/S23; SwitchingStencilArray resident in HBM; hipGraph-replayed "
                               "steps", "cells": cells, "step": "one generation"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "life_kernel<16>", "bytes_per_launch": bytes_per_step,
                     "avg_launch_us": per_step_ev * 1e6, "stream_copy_gbs": copy_gbs,
                     "frac_of_copy": achieved / copy_gbs},
        "plain_per_step": {"value": cells * args.steps / wall_plain, "unit": "cells/s",
                           "ms_per_step": wall_plain / args.steps * 1e3,
                           "graph_speedup": wall_plain / wall},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(A, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
