"""Benchmark of the batched thermochemistry hot path on one MI355X.

One step = one pass of the TSR-GP-24 constant-volume reactor RHS plus its analytic Jacobian over
2^20 cells resident in HBM (structure-of-arrays, FP64): inputs rho_1..rho_24, T (200 B/cell),
outputs wdot_1..24, dT/dt and the 25x25 Jacobian (5,200 B/cell).

    python bench.py [--steps K] [--warmup W] [--cells N] [--no-cpu]

Prints one JSON line.  The cpu_baseline leg times the C oracle (same algorithm, OpenMP over
cells) on a bounded sample on this host's cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12  # B/s, MI355X HBM3E


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cells", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="also run a plain device-to-device copy of this many bytes (counter calibration)")
    return ap.parse_args()


def cpu_baseline(md, Z, seconds):
    """Repeated passes of the C oracle over a fixed sample until about `seconds` of wall time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    o = O.Oracle(md)
    Zs = Z.copy()
    n = Zs.shape[1]
    o.jac(Zs[:, :1024])  # warm-up
    reps = 0
    t = time.perf_counter()
    while True:
        o.jac(Zs)
        reps += 1
        dt = time.perf_counter() - t
        if dt >= seconds:
            break
    return dict(value=reps * n / dt, unit="cells/s", cores=o.num_threads(), kind="port",
                sample=f"{reps} passes over the first {n} cells of the same workload, C oracle "
                       f"(OpenMP over cells), wall {dt:.1f} s")


def traffic_from_profile():
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if present."""
    p = os.path.join(ROOT, "profiles", "rhs_jac_tsr_gp_24_1M.traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (cells shard with no "
              "exchange; see the sharding paragraph of the design notes). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from flint_amd import states as S
    from flint_amd.mechdata import mech_data
    from flint_amd.native import Mechanism

    dev = torch.device("cuda:0")
    md = mech_data("TSR-GP-24")
    mech = Mechanism.from_data(md)
    n = args.cells
    nz = md.ns + 1
    st = S.radical_states(md.species, md.W, 2024, 0, n)
    Z = np.vstack([S.soa(st["rho"]), st["T"][None, :]])
    Zd = torch.from_numpy(Z).to(dev)
    F = torch.empty((nz, n), dtype=torch.float64, device=dev)
    J = torch.empty((nz * nz, n), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()

    if args.copy_bytes:
        src = torch.ones(args.copy_bytes // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for _ in range(args.warmup):
        mech.rhs(Zd, F, J, stream=stream)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        mech.rhs(Zd, F, J, stream=stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = ev0.elapsed_time(ev1)
    ms_step = ms_total / args.steps
    cells_per_s = n / (ms_step * 1e-3)

    bytes_per_cell = nz * 8 + (nz + nz * nz) * 8
    achieved = bytes_per_cell * n / (ms_step * 1e-3)
    out = {
        "metric": "omega_dot evals/s (TSR-GP-24 source terms + analytic Jacobian, FP64)",
        "value": cells_per_s,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic seeded reactor states (CH4/O2 reactant-product blend + radicals)",
        "config": {"workload": "TSR-GP-24 rhs_native + analytic Jacobian, 2^20 cells SoA, FP64",
                   "mechanism": "TSR-GP-24", "cells": n, "T_range_K": [900, 2800], "p_range_atm": [1, 50],
                   "bytes_per_cell": bytes_per_cell, "wall_s": wall},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM, "traffic": traffic_from_profile()},
    }
    if not args.no_cpu:
        ns_cpu = min(n, 1 << 18)
        out["cpu_baseline"] = cpu_baseline(md, Z[:, :ns_cpu], args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
