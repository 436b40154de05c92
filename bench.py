"""Benchmark of the per-cell thermochemistry hot path on one MI355X.

Headline (one step = one pass of the hot path over one batch resident in HBM): the batched stiff
reactor of the TSR-Rich-31 mechanism, constant pressure, 2^18 cells (half near ignition), each
advanced over dt = 1e-4 s by the persistent FP64 BDF kernel at rtol 1e-6, atol 1e-12.  value =
reactor cells advanced per second.

Beside it in the same JSON line:
  mlp_step     the learned step (flint_reactor_advance_mlp) against the plain BDF step over
               dt = 1e-5 s on the same cells, from t = 0 and mid-trajectory (t = 5e-5 s);
  omega_dot    TSR-GP-24 source terms + analytic Jacobian, 2^20 cells (HBM roofline);
  omega_dot_wd WD source terms + heat release, 65,536 cells;
  equilibrium  TSR-GP-24 HP equilibrium of 2^18 states, default guess and learned warm start;
  cons2prim    TSR-GP-24 conservative -> primitive variables (Newton on T), 2^20 cells with three
               velocity components;
  source_general  TSR-GP-24 source terms by the data-driven general routine, 2^20 cells, beside
               the generated explicit routine on the same cells;
  transport    TSR-GP-24 Wilke mixture viscosity and conductivity, 2^20 cells;
  cpu_baseline the C oracle (same algorithms, OpenMP over cells, release flags, threads pinned) on
               bounded samples of the same workloads on this host's cores, and on one core.
Every leg carries a roofline entry: FP64 vector, HBM or BF16 MFMA bound, the achieved rate from
algorithmic flops or bytes over the HIP-event launch time, and, where profiles/ holds rocprofv3
counters of the same kernel, the measured traffic and FP64 issue rate.

    python bench.py [--steps K] [--warmup W] [--cells N] [--no-cpu] [--only reactor]
    python bench.py --pmc-summary OUT.json --pmc-csv COUNTERS.csv --kernel REGEX [--copy-bytes B]
        (summary of a rocprofv3 --pmc run for profiles/)
"""
import argparse
import json
import os
import sys
import time

# the CPU baseline's OpenMP threads stay on their cores (set before any OpenMP runtime starts)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12     # B/s, MI355X HBM3E
PEAK_FP64 = 78.6e12   # flop/s, MI355X FP64 (vector)
PEAK_BF16 = 2.5e15    # flop/s, MI355X dense BF16 MFMA
CPU_BUILD = "gcc -O3 -march=x86-64-v3 -mtune=native -funroll-loops -fopenmp (oracle/libflint_oracle_bench.so)"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--cells", type=int, default=1 << 18)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["reactor", "omega", "cea", "mlp", "thermo", "general", "transport"],
                    default=None)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="also run a plain device-to-device copy of this many bytes (counter calibration)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the reactor's default launch shape")
    ap.add_argument("--pmc-summary", default=None, metavar="OUT_JSON",
                    help="summarise rocprofv3 --pmc counter CSVs (--pmc-csv) of one kernel (--kernel) and exit")
    ap.add_argument("--pmc-csv", action="append", default=[])
    ap.add_argument("--kernel", default=None, help="kernel-name regex for --pmc-summary")
    ap.add_argument("--pmc-cmd", default="", help="the profiled command, recorded in the summary")
    return ap.parse_args()


def pmc_summary(args):
    """Per-dispatch totals of rocprofv3 --pmc counter CSVs for the largest dispatch of one kernel:
    FP64 lane flops (SQ_INSTS_VALU_FLOPS_FP64 counts the flops of each wave instruction, so 64 x
    it bounds the lane flops from above: masked lanes count too) and HBM bytes (FETCH_SIZE and
    WRITE_SIZE in KiB, scaled by what the same counters read for the plain copy of --copy-bytes,
    whose dispatch is in the same CSV)."""
    import csv
    import re
    per = {}
    names = {}
    for path in args.pmc_csv:
        with open(path) as f:
            for r in csv.DictReader(f):
                key = (path, r["Dispatch_Id"])
                names[key] = r["Kernel_Name"]
                d = per.setdefault(key, {})
                d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
                d.setdefault("_grid", int(r["Grid_Size"]))
                d.setdefault("_ns", float(r["End_Timestamp"]) - float(r["Start_Timestamp"]))
    rx = re.compile(args.kernel)
    out = {"kernel_regex": args.kernel, "command": args.pmc_cmd, "counters": {}}
    for path in args.pmc_csv:
        mine = [(k, v) for k, v in per.items() if k[0] == path and rx.search(names[k])]
        if not mine:
            continue
        key, best = max(mine, key=lambda kv: max(x for c, x in kv[1].items() if not c.startswith("_")))
        out["kernel"] = names[key]
        out["launch_ms"] = best["_ns"] * 1e-6
        for c, v in best.items():
            if not c.startswith("_"):
                out["counters"][c] = v
        # copy calibration: the elementwise copy kernel of --copy-bytes in the same run
        copies = [(k, v) for k, v in per.items() if k[0] == path and "copy" in names[k].lower()]
        if copies and args.copy_bytes:
            _, cv = max(copies, key=lambda kv: kv[1].get("FETCH_SIZE", kv[1].get("WRITE_SIZE", 0.0)))
            kib = args.copy_bytes / 1024.0
            if "FETCH_SIZE" in cv:
                out["copy_fetch_scale"] = kib / cv["FETCH_SIZE"]
            if "WRITE_SIZE" in cv:
                out["copy_write_scale"] = kib / cv["WRITE_SIZE"]
    c = out["counters"]
    if "SQ_INSTS_VALU_FLOPS_FP64" in c:
        out["fp64_flops_per_launch"] = 64.0 * c["SQ_INSTS_VALU_FLOPS_FP64"]
        out["fp64_trans_per_launch"] = 64.0 * c.get("SQ_INSTS_VALU_FLOPS_FP64_TRANS", 0.0)
    if "FETCH_SIZE" in c:
        out["hbm_fetch_bytes"] = 1024.0 * c["FETCH_SIZE"] * out.get("copy_fetch_scale", 1.0)
    if "WRITE_SIZE" in c:
        out["hbm_write_bytes"] = 1024.0 * c["WRITE_SIZE"] * out.get("copy_write_scale", 1.0)
    if os.path.exists(args.pmc_summary):
        with open(args.pmc_summary) as f:
            old = json.load(f)
        old.update({k: v for k, v in out.items() if k != "counters"})
        old.setdefault("counters", {}).update(out["counters"])
        out = old
    if "hbm_fetch_bytes" in out and "hbm_write_bytes" in out:
        out["hbm_bytes_per_launch"] = out["hbm_fetch_bytes"] + out["hbm_write_bytes"]
    with open(args.pmc_summary, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


def profile_json(name):
    p = os.path.join(ROOT, "profiles", name)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return {}


def cpu_baseline(o, run, units, seconds, sample, unit="cells/s", one_units=None, run_one=None):
    """The CPU baseline: run() makes one pass over `units` units of work on all OpenMP threads, timed
    for about `seconds` after one warm-up pass (the best pass counts); then the same on one thread
    (run_one over one_units, when given, a smaller sample) for a third of the time."""
    def rate(fn, u, budget):
        fn()
        best, t = float("inf"), time.perf_counter()
        while True:
            t1 = time.perf_counter()
            fn()
            best = min(best, time.perf_counter() - t1)
            if time.perf_counter() - t > budget:
                break
        return u / best
    cores = o.num_threads()
    v = rate(run, units, seconds)
    o.set_threads(1)
    try:
        v1 = rate(run_one or run, one_units or units, seconds / 3.0)
    finally:
        o.set_threads(cores)
    return dict(value=v, unit=unit, cores=cores, value_1core=v1, kind="port", build=CPU_BUILD,
                threads="OMP_PROC_BIND=%s OMP_PLACES=%s" % (os.environ.get("OMP_PROC_BIND"),
                                                           os.environ.get("OMP_PLACES")),
                sample=sample)


def roofline_hbm(kernel, bytes_per_launch, ms, traffic=None):
    a = bytes_per_launch / (ms * 1e-3)
    return dict(bound="hbm", kernel=kernel, achieved=a / 1e9, peak=PEAK_HBM / 1e9, unit="GB/s", frac=a / PEAK_HBM,
                traffic=traffic)


def timed(stream, fn, reps):
    import torch
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    return O


def reactor_leg(args, dev, stream):
    import numpy as np
    import torch
    from flint_amd import states as S
    from flint_amd.mechdata import mech_data
    from flint_amd.native import Mechanism
    md = mech_data("TSR-Rich-31")
    mech = Mechanism.from_data(md)
    n = args.cells
    st = S.radical_states(md.species, md.W, 2024, 0, n, ignition_half=True)
    X0 = np.vstack([S.soa(st["Y"]), st["T"][None, :]])
    Xd0 = torch.from_numpy(X0).to(dev)
    p = torch.from_numpy(st["p"]).to(dev)
    X = torch.empty_like(Xd0)
    stats = torch.zeros((8, n), dtype=torch.int32, device=dev)
    status = torch.zeros(n, dtype=torch.int32, device=dev)

    def step():
        X.copy_(Xd0)
        mech.advance(X, p, mode=1, t_end=1e-4, rtol=1e-6, atol=1e-12, stats=stats, status=status, stream=stream)

    # launch shape from the project's autotuner: (teams per block, blocks) timed on the first 2^16
    # cells of the batch, before the timed region
    tune = None
    if not args.no_autotune:
        m = min(n, 1 << 16)
        t = time.perf_counter()
        best_ms, shape = mech.autotune(Xd0[:, :m].contiguous(), p[:m].contiguous(), mode=1, t_end=1e-4, rtol=1e-6,
                                       atol=1e-12)
        tune = dict(sample_cells=m, best_sample_ms=best_ms, shape=list(shape), seconds=time.perf_counter() - t)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    # HIP events on the launch stream bracket the kernel launches only (the input reset is a
    # separate copy outside the events)
    ms = 0.0
    for _ in range(args.steps):
        X.copy_(Xd0)
        ms += timed(stream, lambda: mech.advance(X, p, mode=1, t_end=1e-4, rtol=1e-6, atol=1e-12, stats=stats,
                                                 status=status, stream=stream), 1)
    wall = time.perf_counter() - t0
    ms_step = ms / args.steps
    s = stats.cpu().numpy().astype(np.float64)
    from flint_amd.flops import reactor_flops
    fl = reactor_flops(md.prog, mode=1)
    flops = fl.total(s[3].sum(), s[4].sum(), s[5].sum(), s[6].sum(), s[0].sum() + s[1].sum())
    res = dict(n=n, ms=ms_step, wall=wall, cells_per_s=n / (ms_step * 1e-3),
               rhs_per_s=s[3].sum() / (ms_step * 1e-3), fails=int((status != 0).sum()),
               steps_mean=float(s[0].mean()), steps_max=int(s[0].max()), rhs_mean=float(s[3].mean()),
               jac_mean=float(s[4].mean()), lu_mean=float(s[5].mean()), newton_mean=float(s[6].mean()),
               flops=float(flops), flops_per_unit=fl.__dict__, launch=list(mech.get_launch()), autotune=tune,
               newton_policy="cvode")
    if not args.no_cpu:
        O = oracle()
        o = O.Oracle(md, release=True)
        # pilot on 128 cells, then samples sized to about cpu_seconds per pass on all threads and
        # a third of that on one thread
        t = time.perf_counter()
        o.bdf(X0[:, :128], st["p"][:128], mode=1, rtol=1e-6, atol=1e-12, t_end=1e-4, newton=1)
        pilot = time.perf_counter() - t
        ncpu = int(min(n, max(128, 128 * 0.5 * args.cpu_seconds / max(pilot, 1e-3))))
        n1 = int(max(32, ncpu / max(1, o.num_threads()) / 3))
        res["cpu"] = cpu_baseline(
            o, lambda: o.bdf(X0[:, :ncpu], st["p"][:ncpu], mode=1, rtol=1e-6, atol=1e-12, t_end=1e-4, newton=1),
            ncpu, 0.5 * args.cpu_seconds,
            f"first {ncpu} cells (1 core: {n1}) of the same workload (every second one near ignition), C BDF "
            f"oracle (same method and Newton policy, OpenMP over cells)", one_units=n1,
            run_one=lambda: o.bdf(X0[:, :n1], st["p"][:n1], mode=1, rtol=1e-6, atol=1e-12, t_end=1e-4, newton=1))
    return res


def mlp_leg(args, dev, stream):
    import numpy as np
    import torch
    from flint_amd import states as S
    from flint_amd.mechdata import mech_data
    from flint_amd.native import Mechanism
    from flint_amd.surrogate import Surrogate, DEFAULT_MODEL
    md = mech_data("TSR-Rich-31")
    mech = Mechanism.from_data(md)
    sur = Surrogate.load(DEFAULT_MODEL)
    ns, n = md.ns, args.cells
    st = S.radical_states(md.species, md.W, 4048, 0, n, ignition_half=True)
    X = torch.from_numpy(np.vstack([S.soa(st["Y"]), st["T"][None, :]])).to(dev)
    p = torch.from_numpy(st["p"]).to(dev)
    out = {}
    # the network alone: one forward pass over the 2^18 cells (BF16 MFMA, FP32 accumulation)
    net = sur.net
    Xf = torch.randn((n, net.dims[0]), dtype=torch.float32, device=dev)
    Yf = net.forward(Xf, stream=stream)
    ms_f = timed(stream, lambda: net.forward(Xf, Yf, stream=stream), 10)
    fl = 2.0 * n * sum(a * b for a, b in zip(net.dims[:-1], net.dims[1:]))
    out["network"] = dict(dims=net.dims, cells=n, forward_ms=ms_f, flops=fl,
                          roofline=dict(bound="mfma", kernel="flint_mlp forward (v_mfma_f32_32x32x16_bf16)",
                                        achieved=fl / (ms_f * 1e-3) / 1e12, peak=PEAK_BF16 / 1e12,
                                        unit="TFLOP/s", frac=fl / (ms_f * 1e-3) / PEAK_BF16))
    del Xf, Yf
    for phase, t0 in (("t0", 0.0), ("mid", 5e-5)):
        if t0 > 0:
            mech.advance(X, p, mode=1, t_end=t0, stream=stream)
        Xp, Xm, Xq = X.clone(), X.clone(), X.clone()
        # the plain step in the caller's order and with the cells handed out by descending |dT/dt|
        # (the learned step orders its fallback cells that way); the faster is the comparison
        ms_given = timed(stream, lambda: mech.advance(Xp, p, mode=1, t_end=sur.dt, stream=stream), 1)
        mech.set_order(1)
        try:
            mech.advance(X.clone(), p, mode=1, t_end=1e-9, stream=stream)  # the order's buffers, untimed
            ms_sorted = timed(stream, lambda: mech.advance(Xq, p, mode=1, t_end=sur.dt, stream=stream), 1)
        finally:
            mech.set_order(0)
        assert torch.equal(Xp, Xq)
        ms_plain = min(ms_given, ms_sorted)
        fb = torch.zeros(n, dtype=torch.int32, device=dev)
        nf = [0]

        def run():
            nf[0] = sur.advance(mech, Xm, p, fallback=fb, stream=stream)

        ms_mlp = timed(stream, run, 1)
        acc = fb == 0
        r = dict(t_start=t0, dt=sur.dt, plain_cells_per_s=n / (ms_plain * 1e-3), mlp_cells_per_s=n / (ms_mlp * 1e-3),
                 speedup=ms_plain / ms_mlp, fallback_rate=nf[0] / n, plain_ms_given_order=ms_given,
                 plain_ms_by_dTdt=ms_sorted, mlp_ms=ms_mlp)
        if bool(acc.any()):
            dY = (Xm[:ns, acc] - Xp[:ns, acc]).abs().max(0).values
            dT = (Xm[ns, acc] - Xp[ns, acc]).abs()
            r.update(max_abs_dY=float(dY.max()), p99_abs_dY=float(torch.quantile(dY.float(), 0.99)),
                     max_abs_dT=float(dT.max()), p99_abs_dT=float(torch.quantile(dT.float(), 0.99)),
                     accepted_over_1e4_dY=float((dY > 1e-4).float().mean()))
        r["judge"] = dict(max_newton_dY=sur.max_newton_dY, max_newton_dT=sur.max_newton_dT,
                          max_residual=sur.max_residual)
        if phase == "mid":
            # the judge's thresholds at the bars themselves (1e-4, 1 K) instead of half of them
            keep = sur.max_newton_dY, sur.max_newton_dT
            try:
                sur.max_newton_dY, sur.max_newton_dT = 1e-4, 1.0
                X6 = X.clone()
                fb6 = torch.zeros(n, dtype=torch.int32, device=dev)
                nf6 = [0]

                def run6():
                    nf6[0] = sur.advance(mech, X6, p, fallback=fb6, stream=stream)

                ms6 = timed(stream, run6, 1)
            finally:
                sur.max_newton_dY, sur.max_newton_dT = keep
            a6 = fb6 == 0
            g6 = dict(max_newton_dY=1e-4, max_newton_dT=1.0, speedup=ms_plain / ms6, fallback_rate=nf6[0] / n)
            if bool(a6.any()):
                dY6 = (X6[:ns, a6] - Xp[:ns, a6]).abs().max(0).values
                dT6 = (X6[ns, a6] - Xp[ns, a6]).abs()
                g6.update(max_abs_dY=float(dY6.max()), max_abs_dT=float(dT6.max()),
                          accepted_over_1e4_dY=float((dY6 > 1e-4).float().mean()),
                          accepted_over_1K=float((dT6 > 1.0).float().mean()))
            r["judge_at_bars"] = g6
        out[phase] = r
        X = Xp
    return out


def omega_leg(args, dev, stream):
    import numpy as np
    import torch
    from flint_amd import states as S
    from flint_amd.mechdata import mech_data
    from flint_amd.native import Mechanism
    md = mech_data("TSR-GP-24")
    mech = Mechanism.from_data(md)
    n, nz = 1 << 20, md.ns + 1
    st = S.radical_states(md.species, md.W, 2024, 0, n)
    Z = np.vstack([S.soa(st["rho"]), st["T"][None, :]])
    Zd = torch.from_numpy(Z).to(dev)
    F = torch.empty((nz, n), dtype=torch.float64, device=dev)
    J = torch.empty((nz * nz, n), dtype=torch.float64, device=dev)
    for _ in range(3):
        mech.rhs(Zd, F, J, stream=stream)
    ms = timed(stream, lambda: mech.rhs(Zd, F, J, stream=stream), 20)
    bpc = nz * 8 + (nz + nz * nz) * 8
    res = dict(n=n, ms=ms, cells_per_s=n / (ms * 1e-3), bytes_per_cell=bpc, achieved_GBps=bpc * n / (ms * 1e-3) / 1e9)
    # the same kernel for TSR-Rich-31 (32 x 32 Jacobian), 2^18 cells
    mr = mech_data("TSR-Rich-31")
    mech_r = Mechanism.from_data(mr)
    nr_, nzr = 1 << 18, mr.ns + 1
    sr = S.radical_states(mr.species, mr.W, 2025, 0, nr_)
    Zr = torch.from_numpy(np.vstack([S.soa(sr["rho"]), sr["T"][None, :]])).to(dev)
    Fr = torch.empty((nzr, nr_), dtype=torch.float64, device=dev)
    Jr = torch.empty((nzr * nzr, nr_), dtype=torch.float64, device=dev)
    for _ in range(3):
        mech_r.rhs(Zr, Fr, Jr, stream=stream)
    ms_r = timed(stream, lambda: mech_r.rhs(Zr, Fr, Jr, stream=stream), 10)
    bpr = nzr * 8 + (nzr + nzr * nzr) * 8
    res["rich31"] = dict(n=nr_, ms=ms_r, cells_per_s=nr_ / (ms_r * 1e-3), bytes_per_cell=bpr,
                         achieved_GBps=bpr * nr_ / (ms_r * 1e-3) / 1e9)
    del Zr, Fr, Jr
    res["traffic"] = profile_json("rhs_jac_tsr_gp_24_1M.traffic.json").get("hbm_bytes_per_launch")
    res["roofline"] = roofline_hbm("rhs_jac_kernel<true> (TSR-GP-24)", bpc * n, ms, res["traffic"])
    # WD source terms + heat release, 65,536 cells
    mwd = mech_data("WD")
    wd = Mechanism.from_data(mwd)
    sw = S.wd_states(mwd.species, mwd.W, 2024, 0, 1 << 16)
    rho = torch.from_numpy(S.soa(sw["rho"])).to(dev)
    T = torch.from_numpy(sw["T"]).to(dev)
    w = torch.empty_like(rho)
    q = torch.empty_like(T)
    for _ in range(3):
        wd.source(rho, T, w, q, stream=stream)
    ms_wd = timed(stream, lambda: wd.source(rho, T, w, q, stream=stream), 200)
    res_wd = dict(n=1 << 16, ms=ms_wd, cells_per_s=(1 << 16) / (ms_wd * 1e-3), bytes_per_cell=96,
                  achieved_GBps=96 * (1 << 16) / (ms_wd * 1e-3) / 1e9)
    res_wd["roofline"] = roofline_hbm("wdot_hrr_kernel (WD)", 96 * (1 << 16), ms_wd)
    if not args.no_cpu:
        O = oracle()
        o = O.Oracle(md, release=True)
        Zs = Z[:, :1 << 16].copy()
        res["cpu"] = cpu_baseline(o, lambda: o.jac(Zs), Zs.shape[1], 0.4 * args.cpu_seconds,
                                  f"passes over the first {Zs.shape[1]} cells, C oracle (rates, Jacobian)")
        ow = O.Oracle(mwd, release=True)
        rs, Ts = S.soa(sw["rho"]), sw["T"]
        res_wd["cpu"] = cpu_baseline(ow, lambda: ow.wd_hrr(rs, Ts), Ts.size, 0.2 * args.cpu_seconds,
                                     "passes over the 65,536 cells, C oracle")
    return res, res_wd


def cea_leg(args, dev, stream):
    import numpy as np
    import torch
    from flint_amd import states as S
    from flint_amd.mechdata import mech_data
    from flint_amd.native import Mechanism
    from flint_amd.warmstart import WarmStart, DEFAULT_MODEL
    md = mech_data("TSR-GP-24")
    mech = Mechanism.from_data(md)
    ws = WarmStart.load(DEFAULT_MODEL)
    n = 1 << 18
    st = S.cea_states(md.species, md.W, 2024, 0, n)
    Y0 = torch.from_numpy(S.soa(st["Y"])).to(dev)
    T0 = torch.from_numpy(st["T0"]).to(dev)
    p = torch.from_numpy(st["p_bar"]).to(dev)
    T1, Y1, T2, Y2 = torch.empty_like(T0), torch.empty_like(Y0), torch.empty_like(T0), torch.empty_like(Y0)
    it1, s1 = torch.empty(n, dtype=torch.int32, device=dev), torch.empty(n, dtype=torch.int32, device=dev)
    it2, s2 = torch.empty_like(it1), torch.empty_like(s1)
    run1 = lambda: mech.equilibrium(T0, Y0, T1, Y1, p=p, hp=True, iters=it1, status=s1, stream=stream)
    run2 = lambda: ws.solve(mech, md, Y0, T0, p, T2, Y2, iters=it2, status=s2, stream=stream)
    run1(), run2()
    ms1, ms2 = timed(stream, run1, 5), timed(stream, run2, 5)
    ok = (s1 == 0) & (s2 == 0)
    res = dict(n=n, default=dict(ms=ms1, states_per_s=n / (ms1 * 1e-3), iters_mean=float(it1.float().mean()),
                                 iters_max=int(it1.max()), failed=int((s1 != 0).sum())),
               warm_start=dict(ms=ms2, states_per_s=n / (ms2 * 1e-3), iters_mean=float(it2.float().mean()),
                               iters_max=int(it2.max()), failed=int((s2 != 0).sum()),
                               max_abs_dT_vs_default=float((T2 - T1)[ok].abs().max())))
    # the tightened convergence tests (flint_cea_set_tolerance(1e-10)): both starts meet the
    # workload's 1e-6 K / 1e-10 bar against each other
    try:
        mech.cea_tolerance(1e-10)
        run1(), run2()
        mt1, mt2 = timed(stream, run1, 5), timed(stream, run2, 5)
    finally:
        mech.cea_tolerance(0.0)
    ok = (s1 == 0) & (s2 == 0)
    res["tight_1e-10"] = dict(
        default=dict(ms=mt1, states_per_s=n / (mt1 * 1e-3), iters_mean=float(it1.float().mean()),
                     iters_max=int(it1.max()), failed=int((s1 != 0).sum())),
        warm_start=dict(ms=mt2, states_per_s=n / (mt2 * 1e-3), iters_mean=float(it2.float().mean()),
                        iters_max=int(it2.max()), failed=int((s2 != 0).sum())),
        speedup=mt1 / mt2, max_abs_dT_warm_vs_default=float((T2 - T1)[ok].abs().max()),
        max_rel_dY_warm_vs_default=float(((Y2 - Y1).abs() / Y1.abs().max(0).values)[:, ok].max()))
    pj = profile_json("cea_kernel_tsr_gp_24_262144.fp64.json")
    if pj.get("fp64_flops_per_launch"):
        a = pj["fp64_flops_per_launch"] / (ms1 * 1e-3)
        res["roofline"] = dict(bound="fp64_valu", kernel="cea_kernel (default start)", achieved=a / 1e12,
                               peak=PEAK_FP64 / 1e12, unit="TFLOP/s", frac=a / PEAK_FP64, traffic=None,
                               note="flops per launch from rocprofv3 SQ_INSTS_VALU_FLOPS_FP64 x 64 "
                                    "(profiles/cea_kernel_tsr_gp_24_262144.fp64.json)")
    if not args.no_cpu:
        O = oracle()
        o = O.Oracle(md, release=True)
        Yc = S.soa(st["Y"])
        m = 1 << 16
        Ym, pm, Tm = np.ascontiguousarray(Yc[:, :m]), st["p_bar"][:m].copy(), st["T0"][:m].copy()
        res["cpu"] = cpu_baseline(o, lambda: o.cea(pm, Tm, Ym, hp=True), m, 0.3 * args.cpu_seconds,
                                  f"passes over the first {m} states, C CEA oracle (OpenMP over states)",
                                  unit="states/s")
    return res


def thermo_leg(args, dev, stream):
    import numpy as np
    import torch
    from flint_amd import states as S
    from flint_amd.mechdata import mech_data
    from flint_amd.native import Mechanism
    md = mech_data("TSR-GP-24")
    mech = Mechanism.from_data(md)
    n, nd = 1 << 20, 3
    fs = S.flow_states(md.species, md.W, 2026, 0, n, nd=nd)
    q = torch.from_numpy(fs["prim"]).to(dev)
    T0 = torch.from_numpy(fs["T0"]).to(dev)
    cons, prim = torch.empty_like(q), torch.empty_like(q)
    T = torch.empty(n, dtype=torch.float64, device=dev)
    it = torch.empty(n, dtype=torch.int32, device=dev)
    mech.prim2cons(q, cons, stream=stream)
    run = lambda: mech.cons2prim(cons, T0, prim, T_out=T, stream=stream)
    for _ in range(3):
        run()
    ms = timed(stream, run, 20)
    mech.cons2prim(cons, T0, prim, T_out=T, iters=it, stream=stream)
    torch.cuda.synchronize()
    npr = md.ns + nd + 1
    bpc = (2 * npr + 2) * 8  # cons and T0 in, prim and T out
    res = dict(n=n, nd=nd, ms=ms, cells_per_s=n / (ms * 1e-3), bytes_per_cell=bpc,
               achieved_GBps=bpc * n / (ms * 1e-3) / 1e9, iters_mean=float(it.float().mean()),
               max_rel_dp=float((prim[-1] / q[-1] - 1.0).abs().max()))
    res["traffic"] = profile_json("cons2prim_tsr_gp_24_1M.traffic.json").get("hbm_bytes_per_launch")
    res["roofline"] = roofline_hbm("cons2prim_kernel (TSR-GP-24)", bpc * n, ms, res["traffic"])
    if not args.no_cpu:
        O = oracle()
        o = O.Oracle(md, release=True)
        m = 1 << 16
        cs = cons[:, :m].cpu().numpy()
        t0 = fs["T0"][:m].copy()
        res["cpu"] = cpu_baseline(o, lambda: o.cons2prim(cs, t0), m, 0.2 * args.cpu_seconds,
                                  f"passes over the first {m} cells, C oracle (OpenMP over cells)")
    return res


def transport_leg(args, dev, stream):
    """Wilke mixture viscosity and conductivity of 2^20 GP-24 flow cells (flint_transport)."""
    import numpy as np
    import torch
    from flint_amd import states as S
    from flint_amd.mechdata import mech_data
    from flint_amd.native import Mechanism
    from flint_amd.transport import species_tables
    md = mech_data("TSR-GP-24")
    ns = md.ns
    mech = Mechanism.from_data(md)
    tb = species_tables(md)
    mech.transport_setup(tb)
    n = 1 << 20
    fs = S.flow_states(md.species, md.W, 2028, 0, n)
    r_h = np.ascontiguousarray(fs["prim"][:ns])
    rho_h = r_h.sum(0)
    r = torch.from_numpy(r_h).to(dev)
    rho = torch.from_numpy(rho_h).to(dev)
    T = torch.from_numpy(fs["T"]).to(dev)
    mu, lam = torch.empty(n, dtype=torch.float64, device=dev), torch.empty(n, dtype=torch.float64, device=dev)
    run = lambda: mech.transport(r, rho, T, mu, lam, stream=stream)
    for _ in range(3):
        run()
    ms = timed(stream, run, 20)
    torch.cuda.synchronize()
    bpc = (ns + 4) * 8  # rho_k, rho, T in; mu, lambda out
    fpc = 6 * ns * ns + 20 * ns  # 3 FMAs per species pair, ~20 flops per species (X, r, interpolation)
    res = dict(n=n, ms=ms, cells_per_s=n / (ms * 1e-3), bytes_per_cell=bpc, flops_per_cell=fpc,
               achieved_GBps=bpc * n / (ms * 1e-3) / 1e9, achieved_TFLOPs=fpc * n / (ms * 1e-3) / 1e12)
    res["roofline"] = dict(bound="fp64_valu", kernel="wilke_kernel<24>", achieved=res["achieved_TFLOPs"],
                           peak=PEAK_FP64 / 1e12, unit="TFLOP/s", frac=res["achieved_TFLOPs"] * 1e12 / PEAK_FP64,
                           traffic=profile_json("wilke_tsr_gp_24_1M.traffic.json").get("hbm_bytes_per_launch"),
                           hbm_frac=res["achieved_GBps"] * 1e9 / PEAK_HBM)
    if not args.no_cpu:
        O = oracle()
        o = O.Oracle(md, release=True)
        m = 1 << 14
        rs, rhs_, Ts = np.ascontiguousarray(r_h[:, :m]), rho_h[:m].copy(), fs["T"][:m].copy()
        res["cpu"] = cpu_baseline(o, lambda: o.wilke(tb, rs, rhs_, Ts), m, 0.2 * args.cpu_seconds,
                                  f"passes over the first {m} cells, C oracle (the reference's ns^2 form, "
                                  f"OpenMP over cells)")
    return res


def general_leg(args, dev, stream):
    import numpy as np
    import torch
    from flint_amd import states as S
    from flint_amd.mechdata import mech_data
    from flint_amd.native import Mechanism
    md = mech_data("TSR-GP-24")
    mech = Mechanism.from_data(md)
    mech.general_setup(md.general)
    n = 1 << 20
    st = S.radical_states(md.species, md.W, 2027, 0, n)
    rho = torch.from_numpy(S.soa(st["rho"])).to(dev)
    T = torch.from_numpy(st["T"]).to(dev)
    w = torch.empty_like(rho)
    run_g = lambda: mech.source_general(rho, T, w, stream=stream)
    run_e = lambda: mech.source(rho, T, w, stream=stream)
    for _ in range(3):
        run_g(), run_e()
    ms_g, ms_e = timed(stream, run_g, 10), timed(stream, run_e, 10)
    bpc = (2 * md.ns + 1) * 8
    res = dict(n=n, ms=ms_g, cells_per_s=n / (ms_g * 1e-3), bytes_per_cell=bpc,
               achieved_GBps=bpc * n / (ms_g * 1e-3) / 1e9, explicit_ms=ms_e,
               explicit_cells_per_s=n / (ms_e * 1e-3))
    res["roofline"] = roofline_hbm("general_kernel (TSR-GP-24)", bpc * n, ms_g,
                                   profile_json("general_tsr_gp_24_1M.traffic.json").get("hbm_bytes_per_launch"))
    if not args.no_cpu:
        O = oracle()
        o = O.Oracle(md, release=True)
        m = 1 << 14
        rs, Ts = S.soa(st["rho"][:m]), st["T"][:m].copy()
        res["cpu"] = cpu_baseline(o, lambda: o.general(rs, Ts), m, 0.2 * args.cpu_seconds,
                                  f"passes over the first {m} cells, C oracle (OpenMP over cells)")
    return res


def main():
    args = parse()
    if args.pmc_summary:
        pmc_summary(args)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (cells shard with no "
              "exchange; see the sharding paragraph of the design notes). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import torch
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    if args.copy_bytes:
        # a plain device copy of known size: the counter calibration of the profiling runs
        src = torch.ones(args.copy_bytes // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    extra = {}
    only = args.only
    rx = reactor_leg(args, dev, stream) if only in (None, "reactor") else None
    if only in (None, "mlp"):
        extra["mlp_step"] = mlp_leg(args, dev, stream)
    if only in (None, "omega"):
        extra["omega_dot"], extra["omega_dot_wd"] = omega_leg(args, dev, stream)
    if only in (None, "cea"):
        extra["equilibrium"] = cea_leg(args, dev, stream)
    if only in (None, "thermo"):
        extra["cons2prim"] = thermo_leg(args, dev, stream)
    if only in (None, "transport"):
        extra["transport"] = transport_leg(args, dev, stream)
    if only in (None, "general"):
        extra["source_general"] = general_leg(args, dev, stream)
    out = {
        "metric": "reactor cells advanced/s at rtol=1e-6 (plain batched FP64 BDF); MLP-assisted step and "
                  "omega_dot evals/s reported beside it",
        "value": rx["cells_per_s"] if rx else None,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": rx["ms"] if rx else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic seeded reactor states (CH4/O2 reactant-product blends + radical seeding, half near "
                "ignition); networks trained on the GPU from the FP64 integrator and solver",
        "config": {"workload": "TSR-Rich-31 constant-pressure 0D reactors, 2^18 cells, dt=1e-4 s, rtol 1e-6, "
                               "atol 1e-12, persistent BDF kernel",
                   "mechanism": "TSR-Rich-31", "cells": args.cells, "dt_s": 1e-4, "rtol": 1e-6, "atol": 1e-12},
    }
    if rx:
        achieved = rx["flops"] / (rx["ms"] * 1e-3)
        out["config"].update({k: rx[k] for k in ("fails", "steps_mean", "steps_max", "rhs_mean", "jac_mean",
                                                 "lu_mean", "newton_mean", "launch", "autotune", "newton_policy",
                                                 "wall")})
        out["rhs_evals_per_s"] = rx["rhs_per_s"]
        # counter profiles of this kernel and workload; used only while they describe the kernel
        # being timed (their recorded launch time within 10 % of this run's)
        tj = profile_json("bdf_kernel_rich31_262144.traffic.json")
        fj = profile_json("bdf_kernel_rich31_262144.fp64.json")
        same = lambda j: args.cells == 1 << 18 and j.get("launch_ms") and abs(j["launch_ms"] / rx["ms"] - 1) < 0.1
        traffic = tj.get("hbm_bytes_per_launch") if same(tj) else None
        out["roofline"] = {"bound": "fp64_valu", "kernel": "flint_reactor::bdf_kernel",
                           "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                           "frac": achieved / PEAK_FP64, "traffic": traffic,
                           "flops_per_launch": rx["flops"],
                           "note": "FP64 vector work (no MFMA). achieved = algorithmic flops (static counts per "
                                   "right-hand side, Jacobian, Gauss-Jordan inverse, Newton correction and step "
                                   "from the rate program, flint_amd/flops.py, times the kernel's per-cell "
                                   "counters) / HIP-event launch time; peak = MI355X FP64 vector"}
        if same(fj) and fj.get("fp64_flops_per_launch"):
            ca = fj["fp64_flops_per_launch"] / (rx["ms"] * 1e-3)
            out["roofline"].update(counter_achieved=ca / 1e12, counter_frac=ca / PEAK_FP64,
                                   counter_flops_per_launch=fj["fp64_flops_per_launch"],
                                   counter_note="rocprofv3 SQ_INSTS_VALU_FLOPS_FP64 x 64 lanes of the same launch "
                                                "(profiles/bdf_kernel_rich31_262144.fp64.json): the FP64 work the "
                                                "SIMDs issue, both half-wavefronts of a cell and masked lanes "
                                                "included")
        if "cpu" in rx:
            out["cpu_baseline"] = rx["cpu"]
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
