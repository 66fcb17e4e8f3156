"""Benchmark of the CDKLM16 step on one MI355X.

Default workload: the 2048x1024 double-jet member (dx=dy=1 km, periodic in x,
walls in y, f=1e-4, dt = 0.7 x CFL), FP32, RK2.  A step is one RK2 time step
(two launches of the substep kernel) over the whole domain; the state is
resident in HBM.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c1|c0]
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_STEP = 68.0   # RK2: 28 B (substep 0) + 40 B (substep 1) per interior cell
BYTES_SUBSTEP = (28.0, 40.0)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def read_traffic(workload):
    """HBM bytes per launch from the committed rocprofv3 counter summary, if any."""
    path = os.path.join(ROOT, "profiles", f"cdklm_march_{workload}_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch")


def cpu_baseline(kw, workload, budget_s=15.0):
    """FP32 oracle (OpenMP, all host threads) on a bounded number of steps."""
    import numpy as np
    from oracle.oracle import OracleCDKLM16, num_threads
    ora = OracleCDKLM16(**kw)
    dt = np.float32(ora.dt)
    cells = kw["nx"] * kw["ny"]
    t0 = time.perf_counter()
    ora.run_steps(1, dt)
    one = time.perf_counter() - t0
    n = max(1, min(200, int(budget_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    ora.run_steps(n, dt)
    el = time.perf_counter() - t0
    return {"value": cells * n / el, "unit": "cell-steps/s", "cores": num_threads(), "kind": "port",
            "sample": f"{workload} {kw['nx']}x{kw['ny']}, {n} RK2 steps of the FP32 oracle "
                      f"({el:.1f} s), wall and periodic ghosts in C"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="c1", choices=["c1", "c0"])
    ap.add_argument("--fast", action="store_true",
                    help="use the fast kernel build (not bit-compatible with the oracle)")
    ap.add_argument("--chunk-rows", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(members or domains would shard across GPUs, see the design notes).", file=sys.stderr)
        sys.exit(2)

    import ctypes
    import numpy as np
    import torch
    import cases
    from gpuocean_amd import CDKLM16, Common
    from gpuocean_amd import _lib as L

    kw = cases.config_c1() if args.workload == "c1" else cases.config_c0()
    ctx = Common.DeviceContext(0)
    sim = CDKLM16(ctx, **kw, exact_arithmetic=not args.fast, chunk_rows=args.chunk_rows or None)
    dt = np.float32(sim.dt)
    nx, ny = kw["nx"], kw["ny"]
    cells = nx * ny
    stream = sim.gpu_stream
    log(f"workload {args.workload} {nx}x{ny} dt={dt:.4f} s, warmup {args.warmup}, steps {args.steps}")

    # warmup
    sim._rk2_steps(args.warmup, dt)
    stream.synchronize()

    # timed region: K RK2 steps, bracketed by synchronisation
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t_wall0 = time.perf_counter()
    ev0.record(stream)
    sim._rk2_steps(args.steps, dt)
    ev1.record(stream)
    stream.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall0
    gpu_ms = ev0.elapsed_time(ev1)
    ms_per_step = gpu_ms / args.steps
    value = cells * args.steps / (gpu_ms / 1e3)
    log(f"timed: {gpu_ms:.2f} ms GPU ({wall * 1e3:.2f} ms wall) for {args.steps} steps -> "
        f"{ms_per_step * 1e3:.2f} us/step, {value:.4g} cell-steps/s")

    # per-launch durations of the substep kernel, events on the kernel's stream
    n_prof = min(args.steps, 200)
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(2 * n_prof + 1)]
    p = sim.params
    h0, h1 = sim.h0, sim.h1
    evs[0].record(stream)
    k = 1
    for _ in range(n_prof):
        sim._substep(h0, h1, None, dt, 0, None)
        evs[k].record(stream)
        sim._substep(h1, h0, h0, dt, 1, None)
        evs[k + 1].record(stream)
        k += 2
    stream.synchronize()
    d0 = [evs[2 * i].elapsed_time(evs[2 * i + 1]) for i in range(n_prof)]
    d1 = [evs[2 * i + 1].elapsed_time(evs[2 * i + 2]) for i in range(n_prof)]
    t0_us = 1e3 * float(np.mean(d0))
    t1_us = 1e3 * float(np.mean(d1))
    avg_launch_us = 0.5 * (t0_us + t1_us)
    bytes_per_launch = 0.5 * BYTES_PER_CELL_STEP * cells
    achieved = bytes_per_launch / (avg_launch_us * 1e-6) / 1e9
    log(f"substep 0: {t0_us:.2f} us ({BYTES_SUBSTEP[0] * cells / t0_us / 1e3:.0f} GB/s), "
        f"substep 1: {t1_us:.2f} us ({BYTES_SUBSTEP[1] * cells / t1_us / 1e3:.0f} GB/s)")

    # sanity: the state stays finite
    eta = sim.download()[0]
    assert np.all(np.isfinite(eta)), "non-finite state"

    result = {
        "metric": "member-cell-steps/s (cell-steps/s of one member); % HBM roofline of the fused step",
        "value": value,
        "unit": "member-cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded double jet over Gaussian-filtered random bathymetry)",
        "config": {"workload": f"CDKLM16 {args.workload.upper()} single member {nx}x{ny}, RK2, "
                               f"{'periodic-x walls-y' if args.workload == 'c1' else 'walls'}",
                   "nx": nx, "ny": ny, "dt": float(dt), "members": 1,
                   "kernel_build": "fast" if args.fast else "exact"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(args.workload),
                     "kernel": "cdklm_march", "bytes_per_cell_per_launch": BYTES_PER_CELL_STEP / 2,
                     "avg_launch_us": avg_launch_us},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        log("cpu baseline (FP32 oracle)...")
        result["cpu_baseline"] = cpu_baseline(kw, args.workload)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
