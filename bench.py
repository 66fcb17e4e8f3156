"""Benchmark of the CDKLM16 step on one MI355X.

Default workload (c2): the batched stochastic ensemble of 256 members of
seeded uneven sizes (nx in 128..512, ny in 128..384, multiples of 16), each the
double jet scaled to its domain, SOAR model error q0=5e-3, L=0.75 dx, its own
Philox seed and its own dt (0.7 x CFL).  A step is one RK2 time step of every
member plus its model error (random field, SOAR, geostrophic balance), from a
captured hipGraph; the states are resident in HBM.  value = member-cell-steps/s
= sum of interior cells x steps / wall time.  The same members run one by one
through the per-member path give the sequential figure the speedup is quoted
against, and the 2048x1024 single member (c1) is measured alongside.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c1|c0|c3|c4|da|copy]

Prints one JSON line.
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
BYTES_NOISE = 28.0           # model error: eta/hu/hv read-modify-write + bathymetry per interior cell


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def read_traffic(kernel, workload):
    """HBM bytes per launch (FETCH + WRITE, calibrated on a copy kernel) from
    the committed rocprofv3 counter summary profiles/<kernel>_<workload>_pmc.txt
    (a line 'hbm_bytes_per_launch: N'), if present."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_pmc.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("hbm_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def event_pair(stream):
    import torch
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    return a, b


# ---------------------------------------------------------------------------- CPU baselines

def cpu_baseline_single(kw, workload, budget_s=15.0, precision="f32"):
    """The oracle (OpenMP, all host threads; FP32, or FP64 for the double-precision
    build) on a bounded number of steps."""
    import numpy as np
    from oracle.oracle import OracleCDKLM16, num_threads
    ora = OracleCDKLM16(**kw, precision=precision)
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
            "sample": f"{workload} {kw['nx']}x{kw['ny']}, {n} RK2 steps of the {precision.upper()} oracle "
                      f"({el:.1f} s), wall and periodic ghosts in C"}


def cpu_baseline_ensemble(members_kw, budget_s=15.0, steps_per_member=8):
    """FP32 oracle on the first members of the ensemble, one after another:
    RK2 step + SOAR + geostrophic balance per step (the normal field is drawn
    with numpy outside the timed calls)."""
    import numpy as np
    from oracle import oracle as O
    from gpuocean_amd.OceanStateNoise import soar_weights
    total_cells = 0
    steps_done = 0
    el = 0.0
    used = 0
    rng = np.random.default_rng(0)
    for kw in members_kw:
        ora = O.OracleCDKLM16(**kw)
        nx, ny = kw["nx"], kw["ny"]
        dt = np.float32(ora.dt)
        W = soar_weights(np.float32(kw["dx"]), np.float32(kw["dy"]), np.float32(5e-3), np.float32(0.75 * kw["dx"]))
        xi = rng.standard_normal((ny + 8, nx)).astype(np.float32)  # periodic in x, walls in y
        for _ in range(steps_per_member):
            t0 = time.perf_counter()
            ora.run_steps(1, dt)
            coarse = O.soar(nx, ny, W, 1.0, False, True, xi)
            e, u, v = ora.h0
            O.geostrophic_balance(nx, ny, np.float32(kw["dx"]), np.float32(kw["dy"]), np.float32(kw["g"]),
                                  np.float32(kw["f"]), 0.0, 0.0, 1.0, None, coarse, e, u, v, ora.Hi)
            el += time.perf_counter() - t0
            total_cells += nx * ny
            steps_done += 1
        used += 1
        if el > budget_s:
            break
    return {"value": total_cells / el, "unit": "member-cell-steps/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"first {used} of the {len(members_kw)} members, {steps_per_member} stochastic RK2 steps each "
                      f"(step + SOAR + balance) of the FP32 oracle, one member after another ({el:.1f} s)"}


# ---------------------------------------------------------------------------- single member

def run_single(kw, workload, steps, warmup, fast, chunk_rows, waves_per_simd=0):
    import numpy as np
    import torch
    from gpuocean_amd import CDKLM16, Common

    ctx = Common.DeviceContext(0)
    sim = CDKLM16(ctx, **kw, exact_arithmetic=not fast, chunk_rows=chunk_rows or None,
                  waves_per_simd=waves_per_simd or None)
    dt = np.float32(sim.dt)
    nx, ny = kw["nx"], kw["ny"]
    cells = nx * ny
    stream = sim.gpu_stream
    log(f"[{workload}] {nx}x{ny} dt={dt:.4f} s, warmup {warmup}, steps {steps}")
    sim._rk2_steps(warmup, dt)
    stream.synchronize()

    torch.cuda.synchronize()
    ev0, ev1 = event_pair(stream)
    t_wall0 = time.perf_counter()
    ev0.record(stream)
    sim._rk2_steps(steps, dt)
    ev1.record(stream)
    stream.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall0
    gpu_ms = ev0.elapsed_time(ev1)
    value = cells * steps / wall
    log(f"[{workload}] timed: {gpu_ms:.2f} ms GPU, {wall * 1e3:.2f} ms wall for {steps} steps -> "
        f"{1e6 * wall / steps:.2f} us/step, {value:.4g} cell-steps/s")

    # The timed region holds nothing but back-to-back launches of the substep
    # kernel on one stream: its average launch duration is the GPU time / 2K.
    avg_launch_us = 1e3 * gpu_ms / (2 * steps)
    achieved = 0.5 * BYTES_PER_CELL_STEP * cells / (avg_launch_us * 1e-6) / 1e9
    # each substep alone: blocks of back-to-back launches between two events
    n_prof = min(steps, 200)
    h0, h1 = sim.h0, sim.h1
    part_us = []
    for rk in (0, 1):
        ev0, ev1 = event_pair(stream)
        ev0.record(stream)
        for _ in range(n_prof):
            if rk == 0:
                sim._substep(h0, h1, None, dt, 0, None)
            else:
                sim._substep(h1, h0, h0, dt, 1, None)
        ev1.record(stream)
        stream.synchronize()
        part_us.append(1e3 * ev0.elapsed_time(ev1) / n_prof)
    t0_us, t1_us = part_us
    log(f"[{workload}] average launch {avg_launch_us:.2f} us ({achieved:.0f} GB/s algorithmic); "
        f"substep 0: {t0_us:.2f} us ({BYTES_SUBSTEP[0] * cells / t0_us / 1e3:.0f} GB/s), "
        f"substep 1: {t1_us:.2f} us ({BYTES_SUBSTEP[1] * cells / t1_us / 1e3:.0f} GB/s)")
    assert np.all(np.isfinite(sim.download()[0])), "non-finite state"
    return {
        "value": value, "unit": "cell-steps/s", "ms_per_step": 1e3 * wall / steps, "steps": steps,
        "config": {"workload": f"CDKLM16 single member {nx}x{ny}, RK2, "
                               f"{'periodic-x walls-y double jet' if workload.startswith('c1') else 'walls, Gaussian bump'}",
                   "nx": nx, "ny": ny, "dt": float(dt), "members": 1,
                   "kernel_build": "fast" if fast else "exact"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic("cdklm_march", workload),
                     "kernel": "cdklm_march", "bytes_per_cell_per_launch": BYTES_PER_CELL_STEP / 2,
                     "cells_per_launch": cells, "avg_launch_us": avg_launch_us},
    }


# ---------------------------------------------------------------------------- batched ensemble

def run_ensemble(args):
    import ctypes
    import numpy as np
    import torch
    import cases
    from gpuocean_amd import BatchedEnsemble, CDKLM16, Common
    from gpuocean_amd import _lib as L

    ctx = Common.DeviceContext(0)
    t0 = time.perf_counter()
    members_kw = cases.config_c2(args.members)
    sims = []
    for i, kw in enumerate(members_kw):
        sim = CDKLM16(ctx, **kw, local_particle_id=i, exact_arithmetic=not args.fast)
        sim.setSOARModelError(5e-3, xorwow_seed=1_000_003 + i)
        sims.append(sim)
    torch.cuda.synchronize()
    cells = sum(int(s.nx) * int(s.ny) for s in sims)
    log(f"[c2] {len(sims)} members, {cells / 1e6:.2f} M interior cells, built in {time.perf_counter() - t0:.1f} s")

    t0 = time.perf_counter()
    ens = BatchedEnsemble(sims, autotune=not args.no_autotune, chunk_rows=args.chunk_rows or None,
                          waves_per_simd=args.waves_per_simd or None)
    n_work, n_blocks = ens.info()
    log(f"[c2] chunk_rows={ens.chunk_rows} waves_per_simd={ens.waves_per_simd} work items={n_work} "
        f"workgroups={n_blocks} (tuned in {time.perf_counter() - t0:.1f} s)")
    G = args.graph_steps

    ens.step(args.warmup, graph_steps=G)
    torch.cuda.synchronize()
    t_wall0 = time.perf_counter()
    ens.step(args.steps, graph_steps=G)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall0
    value = cells * args.steps / wall
    log(f"[c2] batched: {args.steps} steps in {wall * 1e3:.1f} ms -> {1e3 * wall / args.steps:.3f} ms/step, "
        f"{value:.4g} member-cell-steps/s")

    # per-launch durations of each part of a step: blocks of back-to-back
    # launches of one part between two events on the ensemble stream
    st = ens.stream
    sh = ctypes.c_void_p(st.cuda_stream)
    n_prof = min(args.steps, 50)
    ens._sync_in()
    part_us = []
    for part in range(3):
        ev0, ev1 = event_pair(st)
        ev0.record(st)
        for _ in range(n_prof):
            L.check(L.lib().gpo_ensemble_part(ens.handle, part, sh), "gpo_ensemble_part")
        ev1.record(st)
        st.synchronize()
        part_us.append(1e3 * ev0.elapsed_time(ev1) / n_prof)
    for s in sims:   # the model error advanced every member's draw index n_prof times
        s.model_error.draw += n_prof
        s._apply_bc(s.h0)
    ens._draws = [d + n_prof for d in ens._draws]
    avg_launch_us = 0.5 * (part_us[0] + part_us[1])
    achieved = 0.5 * BYTES_PER_CELL_STEP * cells / (avg_launch_us * 1e-6) / 1e9
    noise_gbs = BYTES_NOISE * cells / (part_us[2] * 1e-6) / 1e9
    log(f"[c2] substep 0: {part_us[0]:.1f} us, substep 1: {part_us[1]:.1f} us "
        f"({achieved:.0f} GB/s algorithmic), model error: {part_us[2]:.1f} us ({noise_gbs:.0f} GB/s algorithmic)")

    # the same members one by one through the per-member path
    torch.cuda.synchronize()
    n_seq = args.seq_steps
    t_wall0 = time.perf_counter()
    for s in sims:
        for _ in range(n_seq):
            s.step(np.float32(s.dt))
    torch.cuda.synchronize()
    wall_seq = time.perf_counter() - t_wall0
    value_seq = cells * n_seq / wall_seq
    log(f"[c2] sequential per-member path: {n_seq} steps of every member in {wall_seq:.2f} s -> "
        f"{value_seq:.4g} member-cell-steps/s; batched speedup {value / value_seq:.2f}x")
    for s in sims[:4]:
        assert np.all(np.isfinite(s.download()[0])), "non-finite state"
    ens.destroy()
    observe_r = run_observe(sims)

    sizes = [(int(s.nx), int(s.ny)) for s in sims]
    return members_kw, {
        "metric": "member-cell-steps/s at 256 members; % HBM roofline; adjoint/forward time",
        "value": value,
        "unit": "member-cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded double jets over Gaussian-filtered random bathymetry, Philox model error)",
        "config": {"workload": f"CDKLM16 batched stochastic ensemble, {len(sims)} uneven members "
                               f"(nx 128..512, ny 128..384), RK2 + SOAR q0=5e-3 L=0.75dx, hipGraph of {G} steps",
                   "members": len(sims), "interior_cells": cells,
                   "nx_range": [min(a for a, _ in sizes), max(a for a, _ in sizes)],
                   "ny_range": [min(b for _, b in sizes), max(b for _, b in sizes)],
                   "chunk_rows": ens.chunk_rows, "waves_per_simd": ens.waves_per_simd,
                   "kernel_build": "fast" if args.fast else "exact"},
        "sequential": {"value": value_seq, "unit": "member-cell-steps/s", "steps": n_seq,
                       "speedup_batched": value / value_seq,
                       "path": "each member's own CDKLM16.step, one member after another"},
        "parts_us": {"substep0": part_us[0], "substep1": part_us[1], "model_error": part_us[2]},
        "observe": observe_r,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic("cdklm_march_batched", "c2"),
                     "kernel": "cdklm_march_batched", "bytes_per_cell_per_launch": BYTES_PER_CELL_STEP / 2,
                     "cells_per_launch": cells, "avg_launch_us": avg_launch_us},
    }


def run_coarse_ensemble(n_members=128, factor=3, steps=50, seq_steps=4, graph_steps=10):
    """Batched ensemble whose model error lives on a grid `factor` times
    coarser (the double-jet data-assimilation set-up): uneven members with nx
    a multiple of 16 * factor (periodic in x), the step rate, the model-error
    part alone, and the per-member path for comparison."""
    import ctypes
    import numpy as np
    import torch
    import cases
    from gpuocean_amd import BatchedEnsemble, CDKLM16, Common
    from gpuocean_amd import _lib as L

    ctx = Common.DeviceContext(0)
    q = 16 * factor
    sims = []
    for i, (nx, ny) in enumerate(cases.c2_member_sizes(n_members, seed=77)):
        nx = q * max(3, int(round(nx / q)))
        sim = CDKLM16(ctx, **cases.config_c2_member(nx, ny, 5000 + i), local_particle_id=i)
        sim.setSOARModelError(5e-3, small_scale_perturbation_interpolation_factor=factor, xorwow_seed=900 + i)
        sims.append(sim)
    cells = sum(int(s.nx) * int(s.ny) for s in sims)
    ens = BatchedEnsemble(sims)
    ens.step(graph_steps, graph_steps=graph_steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.step(steps, graph_steps=graph_steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    value = cells * steps / wall
    st = ens.stream
    sh = ctypes.c_void_p(st.cuda_stream)
    ens._sync_in()
    n_prof = 20
    ev0, ev1 = event_pair(st)
    ev0.record(st)
    for _ in range(n_prof):
        L.check(L.lib().gpo_ensemble_part(ens.handle, 4, sh), "gpo_ensemble_part")
    ev1.record(st)
    st.synchronize()
    noise_us = 1e3 * ev0.elapsed_time(ev1) / n_prof
    for s in sims:
        s.model_error.draw += n_prof
    ens._draws = [d + n_prof for d in ens._draws]
    ens.destroy()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in sims:
        for _ in range(seq_steps):
            s.step(np.float32(s.dt))
    torch.cuda.synchronize()
    value_seq = cells * seq_steps / (time.perf_counter() - t0)
    log(f"[coarse ensemble] {len(sims)} members, {cells / 1e6:.2f} M cells, model error {factor}x coarser: "
        f"{1e3 * wall / steps:.3f} ms/step, {value:.4g} member-cell-steps/s (model error {noise_us:.1f} us); "
        f"per member {value_seq:.4g}, {value / value_seq:.1f}x")
    for s in sims:
        s.cleanUp()
    return {"value": value, "unit": "member-cell-steps/s", "members": len(sims), "interior_cells": cells,
            "interpolation_factor": factor, "steps": steps, "ms_per_step": 1e3 * wall / steps,
            "model_error_us": noise_us, "per_member_value": value_seq, "speedup_batched": value / value_seq,
            "workload": f"batched stochastic ensemble, {len(sims)} uneven members (nx multiples of {q}), "
                        f"SOAR q0=5e-3 on a grid {factor}x coarser, bicubic interpolation, hipGraph of "
                        f"{graph_steps} steps"}


def run_coarse_noise(n=2048, factor=3, reps=100):
    """Model error generated on a grid `factor` times coarser (2048 x 2048,
    walls): the bicubic interpolation + balance kernel timed alone with HIP
    events, and one whole perturbation (random field, SOAR, interpolation)."""
    import ctypes
    import numpy as np
    import torch
    import cases
    from gpuocean_amd import CDKLM16, Common
    from gpuocean_amd import _lib as L

    ctx = Common.DeviceContext(0)
    kw = cases.config_c0(n=n)
    sim = CDKLM16(ctx, **kw)
    sim.setSOARModelError(5e-3, small_scale_perturbation_interpolation_factor=factor, xorwow_seed=5)
    me = sim.model_error
    st = sim.gpu_stream
    sim.perturbState()
    st.synchronize()
    ev0, ev1 = event_pair(st)
    ev0.record(st)
    for _ in range(reps):
        sim.perturbState()
    ev1.record(st)
    st.synchronize()
    whole_us = 1e3 * ev0.elapsed_time(ev1) / reps
    args = [int(sim.nx), int(sim.ny), float(sim.dx), float(sim.dy), int(me.coarse_nx), int(me.coarse_ny),
            float(me.coarse_dx), float(me.coarse_dy), 0, 0, float(sim.g), float(sim.f), 0.0,
            float(sim.params.north_x), float(sim.params.north_y), None, 0,
            ctypes.c_void_p(me.coarse_buffer.data_ptr()), 4 * (int(me.coarse_nx) + 4),
            ctypes.c_void_p(sim.h0[0].data_ptr()), ctypes.c_void_p(sim.h0[1].data_ptr()),
            ctypes.c_void_p(sim.h0[2].data_ptr()), sim.h0[0].pitch, ctypes.c_void_p(sim.Hi.data_ptr()), sim.Hi.pitch,
            float(sim.mask_value), ctypes.c_void_p(st.cuda_stream)]
    lib = L.lib()
    ev0, ev1 = event_pair(st)
    ev0.record(st)
    for _ in range(reps):
        L.check(lib.gpo_bicubic_interpolation(*args), "gpo_bicubic_interpolation")
    ev1.record(st)
    st.synchronize()
    kernel_us = 1e3 * ev0.elapsed_time(ev1) / reps
    cells = n * n
    achieved = BYTES_NOISE * cells / (kernel_us * 1e-6) / 1e9
    assert np.all(np.isfinite(sim.download()[0])), "non-finite state"
    log(f"[coarse model error] {n}x{n}, factor {factor}: interpolation + balance {kernel_us:.1f} us "
        f"({achieved:.0f} GB/s algorithmic, {BYTES_NOISE} B/cell), whole perturbation {whole_us:.1f} us")
    sim.cleanUp()
    return {"workload": f"{n}x{n} walls, SOAR on a grid {factor} times coarser, bicubic interpolation",
            "interpolation_us": kernel_us, "perturbation_us": whole_us,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "kernel": "bicubic_balance",
                         "bytes_per_cell_per_launch": BYTES_NOISE, "cells_per_launch": cells}}


def run_observe(sims, n_points=1000, reps=5):
    """hu, hv of every member at the same n_points positions: the library's
    one-launch observation operator against the reference's way (each
    member's interior state downloaded, then numpy indexing), both ending
    with the values on the host."""
    import numpy as np
    import torch
    from gpuocean_amd import observe

    nx = min(int(s.nx) for s in sims)
    ny = min(int(s.ny) for s in sims)
    rng = np.random.default_rng(17)
    pts = np.stack([rng.uniform(0, nx * float(sims[0].dx), n_points),
                    rng.uniform(0, ny * float(sims[0].dy), n_points)], axis=1)
    table = observe._Table()
    observe.observe_members(sims, observe.FLOW, positions=pts, table=table)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        got = observe.observe_members(sims, observe.FLOW, positions=pts, table=table)
    ms = 1e3 * (time.perf_counter() - t0) / reps
    ix = np.floor(pts[:, 0] / float(sims[0].dx)).astype(int)
    iy = np.floor(pts[:, 1] / float(sims[0].dy)).astype(int)
    t0 = time.perf_counter()
    ref = np.empty((len(sims), n_points, 2))
    for p, s in enumerate(sims):
        _, hu, hv = s.download(interior_domain_only=True)
        ref[p, :, 0] = hu[iy, ix]
        ref[p, :, 1] = hv[iy, ix]
    ms_ref = 1e3 * (time.perf_counter() - t0)
    assert np.array_equal(got, ref), "observation operator differs from the download path"
    log(f"[observe] {len(sims)} members x {n_points} points: {ms:.2f} ms (one launch, one read-back) vs "
        f"{ms_ref:.1f} ms downloading every member; {ms_ref / ms:.0f}x")
    return {"members": len(sims), "points": n_points, "ms": ms, "ms_download_path": ms_ref,
            "speedup": ms_ref / ms}


def run_da(n_members=64, cycles=2, obs_interval=300.0):
    """OceanModelEnsemble.stepToObservation over uneven stochastic members
    (model time step 60 s, observations every obs_interval s): the batched
    cadence against the members' own dataAssimilationStep, one after another."""
    import numpy as np
    import torch
    import cases
    from gpuocean_amd import Common, OceanModelEnsemble

    ctx = Common.DeviceContext(0)
    sizes = cases.c2_member_sizes(n_members, 2024)

    def build(batched):
        data = []
        for i, (nx, ny) in enumerate(sizes):
            kw = cases.config_c2_member(nx, ny, 2025 + i)
            data.append({k: kw[k] for k in ("eta0", "hu0", "hv0", "H", "nx", "ny", "dx", "dy",
                                             "boundary_conditions")})
        sim_args = {"dt": 0.0, "g": 9.81, "f": 1e-4, "r": 0.0, "courant_number": 0.7, "model_time_step": 60.0}
        return OceanModelEnsemble(ctx, sim_args, data, {"small_scale_perturbation_amplitude": 5e-3,
                                                        "xorwow_seed": 11}, numParticles=n_members,
                                  batched=batched)

    cells = sum(nx * ny for nx, ny in sizes)
    out = {}
    for batched in (True, False):
        ens = build(batched)
        ens.stepToObservation(obs_interval)   # warm-up cycle
        torch.cuda.synchronize()
        steps0 = sum(s.num_iterations for s in ens.particles)
        t0 = time.perf_counter()
        for c in range(cycles):
            ens.stepToObservation(obs_interval * (c + 2))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        member_steps = sum(s.num_iterations for s in ens.particles) - steps0
        mean_cells = cells / n_members
        out["batched" if batched else "sequential"] = {
            "seconds": wall, "member_steps": member_steps,
            "value": member_steps * mean_cells / wall}
        log(f"[da] {'batched' if batched else 'per member'}: {cycles} observation cycles of {obs_interval:.0f} s, "
            f"{member_steps} member steps in {wall:.2f} s -> {member_steps * mean_cells / wall:.4g} member-cell-steps/s")
        ens.cleanUp()
    out["speedup"] = out["batched"]["value"] / out["sequential"]["value"]
    out["config"] = (f"{n_members} uneven members (c2 sizes), SOAR model error, model time step 60 s, "
                     f"observation every {obs_interval:.0f} s, dt updated every 5 model steps")
    log(f"[da] batched speedup {out['speedup']:.2f}x")
    return out


def run_copy(n_bytes=256 << 20, reps=5):
    """A plain device-to-device copy of n_bytes, reps times: the reference
    kernel whose FETCH_SIZE / WRITE_SIZE counters calibrate the HBM traffic
    figures of the step kernels."""
    import torch
    src = torch.ones(n_bytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    ev0, ev1 = event_pair(None)
    ev0.record()
    for _ in range(reps):
        dst.copy_(src)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    print(json.dumps({"copy_bytes": n_bytes, "reps": reps, "ms_per_copy": ms,
                      "GB_per_s_read_plus_write": 2 * n_bytes / (ms * 1e-3) / 1e9}), flush=True)


def c4_member_sizes(n_members=128, seed=4048):
    """Seeded coastal member sizes: nx = 16*round(U{256..768}/16), ny = 16*round(U{256..512}/16)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    nx = 16 * np.round(rng.integers(256, 769, n_members) / 16).astype(int)
    ny = 16 * np.round(rng.integers(256, 513, n_members) / 16).astype(int)
    return [(int(a), int(b)) for a, b in zip(nx, ny)]


def run_coastal(args):
    """Batched coastal ensemble: 128 members (nx 256..768, ny 256..512), land
    in the western 10 %, flow relaxation (10 cells) on the other sides, SOAR
    q0 = 5e-3, 1000 drifters per member; batched steps against the same
    members stepped one after another."""
    import numpy as np
    import torch
    import cases
    from gpuocean_amd import BatchedEnsemble, CDKLM16, Common, GPUDrifterCollection

    ctx = Common.DeviceContext(0)
    t0 = time.perf_counter()
    sims = []
    for i, (nx, ny) in enumerate(c4_member_sizes(args.members)):
        kw = cases.config_c4_member(nx, ny, 7000 + i)
        sim = CDKLM16(ctx, **kw, local_particle_id=i)
        sim.setSOARModelError(5e-3, xorwow_seed=2_000_003 + i)
        d = GPUDrifterCollection(ctx, 1000, boundaryConditions=sim.boundary_conditions,
                                 domain_size_x=nx * sim.dx, domain_size_y=ny * sim.dy, gpu_stream=sim.gpu_stream)
        d.setDrifterPositions(cases.c4_drifter_positions(nx, ny, sim.dx, sim.dy, 1000, 7000 + i))
        sim.attachDrifters(d)
        sims.append(sim)
    torch.cuda.synchronize()
    cells = sum(int(s.nx) * int(s.ny) for s in sims)
    log(f"[c4] {len(sims)} members, {cells / 1e6:.2f} M interior cells, built in {time.perf_counter() - t0:.1f} s")
    ens = BatchedEnsemble(sims, autotune=not args.no_autotune, chunk_rows=args.chunk_rows or None,
                          waves_per_simd=args.waves_per_simd or None)
    G = args.graph_steps
    ens.step(args.warmup, graph_steps=G)   # includes the first step, taken member by member
    torch.cuda.synchronize()
    t_wall0 = time.perf_counter()
    ens.step(args.steps, graph_steps=G)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall0
    value = cells * args.steps / wall
    log(f"[c4] batched: {args.steps} steps in {wall * 1e3:.1f} ms -> {1e3 * wall / args.steps:.3f} ms/step, "
        f"{value:.4g} member-cell-steps/s (chunk_rows={ens.chunk_rows}, waves_per_simd={ens.waves_per_simd})")
    n_seq = args.seq_steps
    torch.cuda.synchronize()
    t_wall0 = time.perf_counter()
    for s in sims:
        for _ in range(n_seq):
            s.step(np.float32(s.dt))
    torch.cuda.synchronize()
    wall_seq = time.perf_counter() - t_wall0
    value_seq = cells * n_seq / wall_seq
    log(f"[c4] sequential per-member path: {n_seq} steps of every member in {wall_seq:.2f} s -> "
        f"{value_seq:.4g} member-cell-steps/s; batched speedup {value / value_seq:.2f}x")
    for s in sims[:4]:
        assert np.all(np.isfinite(s.download()[0])), "non-finite state"
        assert np.all(np.isfinite(s.drifters.getDrifterPositions())), "non-finite drifter"
    ens.destroy()
    return {"value": value, "unit": "member-cell-steps/s", "steps": args.steps, "ms_per_step": 1e3 * wall / args.steps,
            "sequential": {"value": value_seq, "steps": n_seq, "speedup_batched": value / value_seq},
            "config": {"workload": "CDKLM16 batched coastal ensemble: land (western 10 %), flow relaxation 10 cells "
                                   "N/E/S, wall W, SOAR q0=5e-3, 1000 drifters per member",
                       "members": len(sims), "interior_cells": cells, "nx_range": [256, 768],
                       "ny_range": [256, 512], "chunk_rows": ens.chunk_rows,
                       "waves_per_simd": ens.waves_per_simd}}


def run_adjoint_batch(n_members=128, n_steps=300, k=20, reps=2, n_seq=4):
    """C4's batched gradient: the 128 coastal members (nx 256..768, ny
    256..512; land in the western 10 %, flow relaxation on the other sides,
    SOAR model error q0 = 5e-3 with fixed normals, 1000 drifters each), 300
    RK2 steps, checkpoints every 20; J = per-member drifter-position misfit.
    Forward = the batched checkpointed run; gradient = batched recomputation
    and reverse sweep.  Against that, the first n_seq members' gradients one
    after another (step-by-step path), scaled to all members by cells."""
    import numpy as np
    import torch
    import cases
    from gpuocean_amd import CDKLM16, Common
    from gpuocean_amd.adjoint import DifferentiableCDKLM16, DifferentiableEnsemble

    ctx = Common.DeviceContext(0)
    t_build = time.perf_counter()
    sims, dpos, dobs, x0 = [], [], [], []
    for i, (nx, ny) in enumerate(c4_member_sizes(n_members)):
        sim = CDKLM16(ctx, **cases.config_c4_member(nx, ny, 7000 + i), local_particle_id=i)
        sim.setSOARModelError(5e-3, xorwow_seed=2_000_003 + i)
        p = cases.c4_drifter_positions(nx, ny, sim.dx, sim.dy, 1000, 7000 + i)
        rng = np.random.default_rng(9000 + i)
        sims.append(sim)
        dpos.append(p)
        dobs.append((p + rng.normal(0.0, 500.0, p.shape)).astype(np.float32))
        x0.append([torch.from_numpy(a).cuda() for a in sim.download()])
    cells = sum(int(s.nx) * int(s.ny) for s in sims)
    none = [np.zeros((0, 2), np.float32)] * len(sims)
    ens = DifferentiableEnsemble(sims, none, none, n_steps, checkpoint_every=k, model_error=True,
                                 drifter_positions=dpos, drifter_observed=dobs)
    log(f"[c4-adjoint] {len(sims)} members, {cells / 1e6:.2f} M cells, 1000 drifters each, built in "
        f"{time.perf_counter() - t_build:.1f} s")
    tf, tb = [], []
    for _ in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J = ens.forward(x0)
        t1 = time.perf_counter()
        g = ens.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        tf.append(t1 - t0)
        tb.append(t2 - t1)
        log(f"[c4-adjoint] forward {1e3 * (t1 - t0):.1f} ms, gradient {1e3 * (t2 - t1):.1f} ms")
    f_ms, b_ms = 1e3 * float(np.median(tf[1:])), 1e3 * float(np.median(tb[1:]))
    assert ens.flag_value() == 0 and all(torch.isfinite(a).all().item() for a in g[0]) and np.all(np.isfinite(J))
    state_gb = sum(b.numel() * b.element_size() for b, _ in ens._bufs) / 1e9
    slots_all = ens.n_slots
    ens.destroy()
    del ens, g
    # binomial checkpointing with 4 state slots per member instead of one per segment
    ens = DifferentiableEnsemble(sims, none, none, n_steps, checkpoint_every=k, model_error=True,
                                 drifter_positions=dpos, drifter_observed=dobs, checkpoint_slots=4)
    tb4 = []
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ens.forward(x0)
        t1 = time.perf_counter()
        ens.backward()
        torch.cuda.synchronize()
        tb4.append((t1 - t0, time.perf_counter() - t1))
    f4_ms, b4_ms = 1e3 * tb4[-1][0], 1e3 * tb4[-1][1]
    state4_gb = sum(b.numel() * b.element_size() for b, _ in ens._bufs) / 1e9
    ens.destroy()
    del ens
    log(f"[c4-adjoint] binomial checkpointing, 4 slots: forward {f4_ms:.1f} ms, gradient {b4_ms:.1f} ms "
        f"(gradient/forward {b4_ms / f4_ms:.2f}); state buffers {state4_gb:.1f} GB against {state_gb:.1f} GB")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    seq_cells = 0
    for s, p, o, x in list(zip(sims, dpos, dobs, x0))[:n_seq]:
        one = DifferentiableCDKLM16(s, np.zeros((0, 2), np.float32), np.zeros((0, 2), np.float32), n_steps,
                                    checkpoint_every=k, model_error=True, drifter_positions=p, drifter_observed=o)
        one.forward(*x)
        one.backward()
        seq_cells += int(s.nx) * int(s.ny)
    torch.cuda.synchronize()
    seq_ms = 1e3 * (time.perf_counter() - t0) * cells / seq_cells
    log(f"[c4-adjoint] {n_steps} steps: batched forward {f_ms:.1f} ms, gradient {b_ms:.1f} ms (gradient/forward "
        f"{b_ms / f_ms:.2f}); members one after another (forward + gradient, {n_seq} members scaled by cells) "
        f"{seq_ms:.1f} ms; batched speedup {seq_ms / (f_ms + b_ms):.2f}x")
    return {"members": len(sims), "interior_cells": cells, "steps": n_steps, "checkpoint_every": k,
            "drifters_per_member": 1000, "misfit": "drifter positions after the last step",
            "model_error": "SOAR q0=5e-3, fixed normals", "forward_ms": f_ms, "gradient_ms": b_ms,
            "ratio": b_ms / f_ms, "sequential_ms": seq_ms, "sequential_sample_members": n_seq,
            "speedup_batched": seq_ms / (f_ms + b_ms), "state_slots": slots_all, "state_buffers_gb": state_gb,
            "binomial_4_slots": {"forward_ms": f4_ms, "gradient_ms": b4_ms, "ratio": b4_ms / f4_ms,
                                 "state_buffers_gb": state4_gb}}


def run_adjoint(reps=3):
    """Adjoint/forward time on the differentiable 512x512 member: 400 RK2 steps,
    200 observation points, checkpoints every 20 steps; forward = FP32 run with
    checkpoints + misfit, gradient = recompute of every segment + adjoint sweep."""
    import numpy as np
    import torch
    import cases
    from gpuocean_amd import CDKLM16, Common
    from gpuocean_amd.adjoint import DifferentiableCDKLM16

    ctx = Common.DeviceContext(0)
    kw = cases.config_c3(n=512, seed=3)
    bg, truth = cases.c3_background_and_truth(kw, seed=3)
    sim = CDKLM16(ctx, **kw)
    pos = cases.c3_drifters(512, 512, 200.0, 200.0, n=200, seed=3)
    probe = DifferentiableCDKLM16(sim, pos, np.zeros_like(pos), 400, checkpoint_every=20)
    probe.forward(*[torch.from_numpy(a).cuda() for a in truth])
    model = DifferentiableCDKLM16(sim, pos, probe.observe(), 400, checkpoint_every=20)
    x = [torch.from_numpy(a).cuda() for a in bg]
    fw, bw = [], []
    for _ in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J, _ = model.forward(*x)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        g = model.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        fw.append(t1 - t0)
        bw.append(t2 - t1)
    f_ms, b_ms = 1e3 * float(np.median(fw[1:])), 1e3 * float(np.median(bw[1:]))
    assert all(np.isfinite(a.cpu().numpy()).all() for a in g) and model.flag_value() == 0
    log(f"[c3] forward {f_ms:.1f} ms, gradient {b_ms:.1f} ms, gradient/forward {b_ms / f_ms:.2f}")
    return {"forward_ms": f_ms, "gradient_ms": b_ms, "ratio": b_ms / f_ms, "J": J,
            "config": "512x512, 400 RK2 steps, 200 observation points, checkpoints every 20 steps, FP32"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="c2", choices=["c2", "c1", "c1big", "c0", "c3", "c4", "da", "copy"],
                    help="c3 = the adjoint/forward timing alone; c4 = the batched coastal ensemble (land, flow "
                         "relaxation, drifters); da = OceanModelEnsemble.stepToObservation batched against "
                         "per member; copy = HBM copy of 256 MiB, the calibration kernel for the traffic counters")
    ap.add_argument("--members", type=int, default=256)
    ap.add_argument("--graph-steps", type=int, default=10)
    ap.add_argument("--seq-steps", type=int, default=10)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--fast", action="store_true",
                    help="use the fast kernel build (not bit-compatible with the oracle)")
    ap.add_argument("--chunk-rows", type=int, default=0)
    ap.add_argument("--waves-per-simd", type=int, default=0, help="persistent grid of the batched launch")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c1", action="store_true", help="skip the single-member line of the c2 run")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward timing of the c2 run")
    ap.add_argument("--no-c4", action="store_true",
                    help="skip the coastal ensemble (C4) forward and batched-gradient lines of the c2 run")
    ap.add_argument("--adjoint-steps", type=int, default=300,
                    help="RK2 steps of the C4 batched gradient (300 is the stated configuration; fewer for profiling)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(ensemble members would be split across GPUs by member, with no exchange while stepping).",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "copy":
        run_copy()
        return

    import cases

    if args.workload == "c4":
        args.members = 128 if args.members == 256 else args.members
        args.steps = 300 if args.steps is None else args.steps
        args.warmup = 10 if args.warmup is None else args.warmup
        args.seq_steps = 3 if args.seq_steps == 10 else args.seq_steps
        out = {"coastal_c4": run_coastal(args)}
        if not args.no_adjoint:
            out["adjoint_batched_c4"] = run_adjoint_batch(n_steps=args.adjoint_steps)
        print(json.dumps(out), flush=True)
        return

    if args.workload == "da":
        print(json.dumps({"step_to_observation": run_da()}), flush=True)
        return

    if args.workload == "c3":
        print(json.dumps({"adjoint_c3": run_adjoint(reps=3 if args.steps is None else max(1, args.steps))}),
              flush=True)
        return

    if args.workload == "c2":
        args.steps = 500 if args.steps is None else args.steps
        args.warmup = 20 if args.warmup is None else args.warmup
        members_kw, result = run_ensemble(args)
        if not args.no_c1:
            c1 = run_single(cases.config_c1(), "c1", 2000, 50, args.fast, 0)
            result["single_member_c1"] = c1
            big = run_single(cases.config_c1_large(), "c1big", 200, 10, args.fast, 0)
            big.pop("config", None)
            result["single_member_8192x4096"] = big
            result["coarse_model_error"] = run_coarse_noise()
            result["coarse_model_error_ensemble"] = run_coarse_ensemble()
        if not args.no_adjoint:
            result["adjoint_c3"] = run_adjoint()
        if not args.no_c4:
            c4 = argparse.Namespace(**vars(args))
            c4.members, c4.steps, c4.warmup, c4.seq_steps = 128, 300, 10, 3
            c4.chunk_rows = c4.waves_per_simd = 0
            result["coastal_c4"] = run_coastal(c4)
            if not args.no_adjoint:
                result["adjoint_batched_c4"] = run_adjoint_batch(n_steps=args.adjoint_steps)
        result["cpu_baseline"] = None
        if not args.no_cpu:
            log("cpu baseline (FP32 oracle, ensemble members one after another)...")
            result["cpu_baseline"] = cpu_baseline_ensemble(members_kw)
            log("cpu baseline of the FP64 oracle (C1, bounded sample)...")
            result["cpu_baseline_f64_c1"] = cpu_baseline_single(cases.config_c1(), "c1", budget_s=5.0,
                                                                precision="f64")
        print(json.dumps(result), flush=True)
        return

    args.steps = 2000 if args.steps is None else args.steps
    args.warmup = 50 if args.warmup is None else args.warmup
    kw = {"c1": cases.config_c1, "c1big": cases.config_c1_large}.get(args.workload, cases.config_c0)()
    r = run_single(kw, args.workload, args.steps, args.warmup, args.fast, args.chunk_rows, args.waves_per_simd)
    result = {
        "metric": "member-cell-steps/s at 256 members; % HBM roofline; adjoint/forward time",
        "value": r["value"], "unit": "member-cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded double jet over Gaussian-filtered random bathymetry)"
        if args.workload == "c1" else "synthetic (Gaussian bump)",
        "config": r["config"], "roofline": r["roofline"], "cpu_baseline": None,
    }
    if not args.no_cpu:
        log("cpu baseline (FP32 oracle)...")
        result["cpu_baseline"] = cpu_baseline_single(kw, args.workload)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
