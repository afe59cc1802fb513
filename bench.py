"""Benchmark of the ideal-MHD step on one MI355X.

Default workload: 3D Alfvénic turbulence on a periodic 256^3 grid (rho = 1, p = 1,
gamma = 5/3, B0 = (1, 0, 0), seeded velocity and field perturbations of rms 0.1
over modes 1 <= |n| <= 4), minmod, CFL 0.3.  A step is one full SSP-RK2 (Heun)
step including its CFL reduction; inputs are resident in HBM.  The script times
the FP64 path and the mixed-precision path, reports the L1 density error of the
mixed path against the FP64 path after 200 steps of both at the FP64 path's stable
step, the roofline of the stage kernel measured with HIP events on the stream it
runs on, a CPU baseline (the NumPy oracle on one core: one step of the same 256^3
workload) and a summary of the refined learned-indicator 3D step (``c5``).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cells 256]
    python bench.py --workload ot2d|blast2d-amr|harris2d   (the other workloads)

Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
STATE_BYTES_3D = 64  # 5 FP64 cell fields + 3 FP64 edge-potential components per cell
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "stage3d_turbulence256_traffic.md")


def profiled_traffic(precision: str, nxs: int, waves: int, cells: int):
    """HBM bytes per stage launch measured by the FETCH_SIZE / WRITE_SIZE runs of the
    committed profile (copy-calibrated), if it covers this workload and launch."""
    if cells != 256**3 or not os.path.exists(TRAFFIC_PROFILE):
        return None
    want = f"{precision} nxs={nxs}" + (f" waves={waves}" if precision == "mixed" else "")
    with open(TRAFFIC_PROFILE) as fh:
        for line in fh:
            if line.startswith("traffic_bytes_per_launch:"):
                body = line.split(":", 1)[1].strip()
                spec, value = body.rsplit(" ", 1)
                if spec == want:
                    return float(value)
    return None


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=40)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--cells", type=int, default=256)
    p.add_argument("--fp64-kernels", default="fused", choices=("split", "fused"))
    p.add_argument("--cpu-seconds", type=float, default=5.0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-tune", action="store_true")
    p.add_argument("--nxs", type=int, nargs=2, default=None, metavar=("FP64", "MIXED"),
                   help="cells per workgroup along x (skips the autotuner)")
    p.add_argument("--waves", type=int, default=None, help="waves/SIMD of the mixed stage")
    p.add_argument("--calibrate", action="store_true",
                   help="also run one device copy of 1 GiB (the traffic counters' reference)")
    p.add_argument("--accuracy-steps", type=int, default=200,
                   help="steps of the same-dt mixed-vs-FP64 error run (the config's 200)")
    p.add_argument("--with-c5", dest="with_c5", action="store_true", default=True,
                   help="add the refined learned-indicator step's summary (default)")
    p.add_argument("--no-c5", dest="with_c5", action="store_false")
    p.add_argument("--workload", default="turbulence3d",
                   choices=("turbulence3d", "ot2d", "blast2d-amr", "harris2d", "blast3d-amr",
                            "harris2d-resistive", "plasmoid2d-amr"))
    return p.parse_args()


def time_model(model, steps, warmup):
    """Warm up, then time ``steps`` full steps; returns (seconds, stage-kernel ms per launch)."""
    import torch

    for _ in range(warmup):
        model.advance(1)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    events = []
    t0 = time.perf_counter()
    for _ in range(steps):
        marks = model.advance(1, events=True)
        events.extend(marks)
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t0
    stage_ms = [a.elapsed_time(b) for a, b in events]
    del stream
    return seconds, sum(stage_ms) / max(1, len(stage_ms))


def cpu_baseline(cells_sample: int, budget: float, max_steps: int = 100, problem=None,
                 warmup: bool = True):
    """The NumPy oracle step (one core) on a turbulence box of ``cells_sample``^3 (or on
    ``problem``, the GPU run's own set-up, when given)."""
    from tests.helpers import oracle3d
    from wellbalanced_amd import initial

    if problem is None:
        problem = initial.turbulence_3d(cells_sample, seed=0)
    oracle = oracle3d(problem)
    if warmup:
        oracle.step()
    steps, t0 = 0, time.perf_counter()
    while True:
        oracle.step()
        steps += 1
        elapsed = time.perf_counter() - t0
        if elapsed >= budget or steps >= max_steps:
            break
    return cells_sample**3 * steps / elapsed, steps, elapsed


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(the multi-GPU sharding is a design in DESIGN.md, not built)"}))
        sys.exit(2)
    runs = {"turbulence3d": run_turbulence3d, "ot2d": run_ot2d, "blast2d-amr": run_blast2d_amr,
            "harris2d": run_harris2d, "blast3d-amr": run_blast3d_amr,
            "harris2d-resistive": run_harris2d_resistive, "plasmoid2d-amr": run_plasmoid2d_amr}
    print(json.dumps(runs[args.workload](args)), flush=True)


def host_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu": model}


def run_turbulence3d(args):
    import torch

    from wellbalanced_amd import MHD3D, initial

    torch.cuda.set_device(0)
    n = args.cells
    problem = initial.turbulence_3d(n, seed=0)
    cells = n**3

    fp64 = MHD3D.from_problem(problem, kernels=args.fp64_kernels, precision="fp64")
    mixed_kernels = "fused"
    try:
        mixed = MHD3D.from_problem(problem, kernels=mixed_kernels, precision="mixed")
    except (NotImplementedError, ValueError):
        mixed = None
    tuned = {}
    if args.nxs is not None:
        fp64.nxs = args.nxs[0]
        if mixed is not None:
            mixed.nxs = args.nxs[1]
            mixed.waves = args.waves or 2
        tuned = {"fp64": {"nxs": fp64.nxs}, "mixed": {"nxs": args.nxs[1], "waves": args.waves or 2}}
    elif not args.no_tune:
        from wellbalanced_amd import autotune

        tuned["fp64"] = autotune.tune(fp64, steps=2)
        if mixed is not None:
            tuned["mixed"] = autotune.tune(mixed, steps=2)
    if args.calibrate:
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    fp64.check_interval = 0
    sec64, stage64 = time_model(fp64, args.steps, args.warmup)
    rate64 = cells * args.steps / sec64
    result_model, sec, stage_ms, precision = fp64, sec64, stage64, "fp64"
    l1 = linf = 0.0
    accuracy = {}
    if mixed is not None:
        mixed.check_interval = 0
        secm, stagem = time_model(mixed, args.steps, args.warmup)
        result_model, sec, stage_ms, precision = mixed, secm, stagem, "mixed"
        fp64.check()
        mixed.check()
        accuracy = same_dt_error_3d(problem, args.accuracy_steps, tuned)
        l1, linf = accuracy["l1"], accuracy["linf"]
    else:
        fp64.check()
    value = cells * args.steps / sec
    ms_per_step = 1e3 * sec / args.steps
    # stage kernel roofline: algorithmic bytes per launch = 2.5 S per cell (stage 1
    # reads U^n, writes U^1: 2S; stage 2 reads U^1 and U^n, writes U^n+1: 3S)
    bytes_per_launch = 2.5 * STATE_BYTES_3D * cells
    achieved = bytes_per_launch / (stage_ms * 1e-3) / 1e9
    launch = tuned.get(precision, {})
    traffic = profiled_traffic(precision, launch.get("nxs", result_model.nxs),
                               launch.get("waves", getattr(result_model, "waves", 0)), cells)
    out = {
        "metric": "zone-updates/s on 1 MI355X, mixed vs FP64 at stated L1 error; % HBM peak",
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-flux/f64-state" if precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"3D Alfvenic turbulence {n}^3 periodic, minmod, CFL 0.3, SSP-RK2",
            "cells": cells,
            "precision": precision,
            "fp64_kernels": args.fp64_kernels,
            "launch": tuned,
        },
        "fp64_zone_updates_per_s": rate64,
        "speedup_vs_fp64": value / rate64,
        "l1_density_vs_fp64": l1,
        "linf_density_vs_fp64": linf,
        "error_run": accuracy.get("note"),
        "hbm_fraction_algorithmic": value * 5 * STATE_BYTES_3D / (HBM_PEAK_GBS * 1e9),
        # the compulsory lower bound: each step reads U^n and writes U^n+1 once (2S)
        "hbm_fraction_compulsory": value * 2 * STATE_BYTES_3D / (HBM_PEAK_GBS * 1e9),
        "fp64_stage_ms": stage64,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": (os.path.relpath(TRAFFIC_PROFILE, ROOT) if traffic is not None
                               else None),
            "kernel": "stage (both Heun stages)",
            "stage_ms": stage_ms,
        },
    }
    if not args.no_cpu:
        # one oracle step of the timed workload itself (the same set-up and size), and a
        # small box beside it to show how the CPU's rate depends on the size
        cpu_rate, steps, elapsed = cpu_baseline(n, 0.0, max_steps=1, problem=problem,
                                                warmup=False)
        small = 48
        small_rate, small_steps, small_elapsed = cpu_baseline(small, args.cpu_seconds)
        out["cpu_baseline"] = {
            "value": cpu_rate,
            "unit": "zone-updates/s",
            "cores": 1,
            "kind": "port",
            "sample": f"NumPy FP64 oracle, one step of the timed workload (turbulence {n}^3) "
                      f"in {elapsed:.1f} s",
            "small_box": {"cells": small**3, "steps": small_steps, "seconds": small_elapsed,
                          "zone_updates_per_s": small_rate},
            "host": host_info(),
        }
    if args.with_c5:
        out["c5"] = c5_summary(args)
    return out


def same_dt_error_3d(problem, steps, tuned):
    """L1 / L-inf of the mixed density against FP64 after ``steps`` steps of both at the
    FP64 path's stable step (so that no time offset enters the error)."""
    import numpy as np

    from wellbalanced_amd import MHD3D

    pair = {}
    for precision in ("fp64", "mixed"):
        m = MHD3D.from_problem(problem, kernels="fused", precision=precision)
        for key, value in tuned.get(precision, {}).items():
            setattr(m, key, value)
        m.check_interval = 0
        pair[precision] = m
    for _ in range(steps):
        dt = pair["fp64"].stable_step()
        pair["fp64"].step(dt)
        pair["mixed"].step(dt)
    pair["fp64"].check()
    pair["mixed"].check()
    a = pair["fp64"].dataset()["density"]
    b = pair["mixed"].dataset()["density"]
    return {"l1": float(np.mean(np.abs(a - b))), "linf": float(np.max(np.abs(a - b))),
            "note": f"mixed vs FP64 density after {steps} steps of both at the FP64 stable step "
                    f"(t = {pair['fp64'].time:.6g})"}


def c5_summary(args):
    """The refined 3D step of the learned-indicator workload (see run_blast3d_amr), in
    short: its mixed and FP64 zone-updates/s, their ratio, the L1 error and the block
    stage's fraction of the HBM roofline."""
    line = run_blast3d_amr(args, judge=False)
    return {k: line[k] for k in ("value", "fp64_zone_updates_per_s", "speedup_vs_fp64",
                                 "l1_density_vs_fp64", "linf_density_vs_fp64", "ms_per_step",
                                 "hbm_fraction_algorithmic")} | {
        "stage_frac": line["roofline"]["frac"], "stage_ms": line["roofline"]["stage_ms"],
        "workload": line["config"]["workload"], "cells": line["config"]["cells"]}


STATE_BYTES_2D = 56  # 6 FP64 cell fields + 1 FP64 node (A_z) per cell


def _line(metric_value, rate64, steps, args, sec, workload, cells, precision, l1, linf, stage_ms,
          state_bytes, extra):
    """The JSON line of a non-default workload (same contract as the default one)."""
    achieved = 2.5 * state_bytes * cells / (stage_ms * 1e-3) / 1e9 if stage_ms else None
    out = {
        "metric": "zone-updates/s on 1 MI355X, mixed vs FP64 at stated L1 error; % HBM peak",
        "value": metric_value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * sec / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-flux/f64-state" if precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": workload, "cells": cells, "precision": precision, **extra},
        "fp64_zone_updates_per_s": rate64,
        "speedup_vs_fp64": metric_value / rate64,
        "l1_density_vs_fp64": l1,
        "linf_density_vs_fp64": linf,
        "hbm_fraction_algorithmic": metric_value * 5 * state_bytes / (HBM_PEAK_GBS * 1e9),
        "hbm_fraction_compulsory": metric_value * 2 * state_bytes / (HBM_PEAK_GBS * 1e9),
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": None,
            "kernel": "stage (both Heun stages)", "stage_ms": stage_ms,
        },
    }
    return out


def _timed_steps(m, args, graph_steps=0):
    """Seconds and steps of the timed region: ``args.steps`` eager steps after
    ``args.warmup``, or replays of a HIP graph of ``graph_steps`` steps (one replay
    first, untimed).  Returns (seconds, steps)."""
    import torch

    from wellbalanced_amd.graph import StepGraph

    if graph_steps:
        graph = StepGraph(m, steps=graph_steps)
        reps = max(1, args.steps // graph_steps)
        graph.replay(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        graph.replay(reps)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, reps * graph_steps
    for _ in range(args.warmup):
        m.advance(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m.advance(args.steps)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, args.steps


def _alternating_steps(models, args, graph_steps, rounds=5):
    """Seconds and steps per model of replays of a HIP graph of ``graph_steps`` steps,
    timed in ``rounds`` alternating rounds (each model's ``args.steps`` steps once per
    round, one untimed replay of each first): the median round of each model, so that a
    clock drift over the run weighs on all models alike.  Returns {name: (seconds,
    steps)}."""
    import statistics

    import torch

    from wellbalanced_amd.graph import StepGraph

    graphs = {k: StepGraph(m, steps=graph_steps) for k, m in models.items()}
    reps = max(1, args.steps // graph_steps)
    for g in graphs.values():
        g.replay(1)
    torch.cuda.synchronize()
    times = {k: [] for k in graphs}
    for _ in range(rounds):
        for k, g in graphs.items():
            t0 = time.perf_counter()
            g.replay(reps)
            torch.cuda.synchronize()
            times[k].append(time.perf_counter() - t0)
    return {k: (statistics.median(v), reps * graph_steps) for k, v in times.items()}


def _pair_2d(problem, args, graph_steps=0):
    """FP64 and mixed fused 2D models, tuned and timed; returns the line's pieces."""
    import numpy as np
    import torch

    from wellbalanced_amd import MHD2D, autotune

    res = {}
    models = {}
    for precision in ("fp64", "mixed"):
        m = MHD2D.from_problem(problem, kernels="fused", precision=precision)
        launch = autotune.tune(m, steps=2) if not args.no_tune else {}
        m.check_interval = 0
        _, stage_ms = time_model(m, 3, 1)
        sec, steps = _timed_steps(m, args, graph_steps)
        m.check()
        res[precision] = (m.nx * m.ny * steps / sec, sec, steps, stage_ms, launch)
        models[precision] = m
    a = models["fp64"].dataset()["density"]
    b = models["mixed"].dataset()["density"]
    l1, linf = float(np.mean(np.abs(a - b))), float(np.max(np.abs(a - b)))
    div = max(models["fp64"].divergence(), models["mixed"].divergence())
    return res, l1, linf, div, models


def run_ot2d(args):
    from wellbalanced_amd import initial

    n = 256
    problem = initial.orszag_tang(n, limiter="minmod")
    res, l1, linf, div, models = _pair_2d(problem, args, graph_steps=10)
    rate, sec, steps, stage_ms, launch = res["mixed"]
    return _line(rate, res["fp64"][0], steps, args, sec,
                 f"2D Orszag-Tang {n}^2 periodic, minmod, CFL 0.4, SSP-RK2, steps replayed "
                 "from HIP graphs of 10 steps", n * n, "mixed",
                 l1, linf, stage_ms, STATE_BYTES_2D,
                 {"launch": {"fp64": res["fp64"][4], "mixed": launch}, "divergence": div,
                  "time": models["mixed"].time})


def run_harris2d(args):
    from wellbalanced_amd import initial

    problem = initial.harris_sheet(2048, 1024)
    res, l1, linf, div, models = _pair_2d(problem, args, graph_steps=10)
    rate, sec, steps, stage_ms, launch = res["mixed"]
    return _line(rate, res["fp64"][0], steps, args, sec,
                 "2D Harris sheet 2048x1024, walls in y, minmod, CFL 0.4, SSP-RK2, "
                 "steps replayed from HIP graphs of 10 steps", 2048 * 1024, "mixed", l1, linf,
                 stage_ms, STATE_BYTES_2D,
                 {"launch": {"fp64": res["fp64"][4], "mixed": launch}, "divergence": div,
                  "time": models["mixed"].time})


def run_harris2d_resistive(args):
    """Resistive double Harris sheets with conduction on 256 x 256 cells: the workload
    of the reference's published step time (63 ms, CHANGELOG.md:32-33, CPU)."""
    from wellbalanced_amd import initial
    from wellbalanced_amd.hierarchy import AMRHierarchy

    h = AMRHierarchy((256, 256), ((0.0, 2.0), (0.0, 2.0)), periodic=(True, True))
    state = initial.harris_sheets(h, width=0.08, perturbation=0.05)
    problem = initial.Problem(h, {"resistivity": 2e-3, "conduction": (0.01, 0.001),
                                  "limiter": "vanleer"}, state)
    res, l1, linf, div, models = _pair_2d(problem, args, graph_steps=10)
    rate, sec, steps, stage_ms, launch = res["mixed"]
    return _line(rate, res["fp64"][0], steps, args, sec,
                 "2D double Harris sheets 256x256 periodic on [0,2]^2 (width 0.08, "
                 "perturbation 0.05), eta 2e-3, conduction (0.01, 0.001), van Leer, CFL 0.4, "
                 "SSP-RK2, steps replayed from HIP graphs of 10 steps", 256 * 256, "mixed",
                 l1, linf, stage_ms, STATE_BYTES_2D,
                 {"launch": {"fp64": res["fp64"][4], "mixed": launch}, "divergence": div,
                  "time": models["mixed"].time,
                  "reference_cpu_zone_updates_per_s": 256 * 256 / 0.063})


def run_plasmoid2d_amr(args):
    """The reference's plasmoid example at its production geometry (examples/
    plasmoid_reconnection.py:39-48,75-91): 256 x 128 base on [0,4] x [0,2], 3 levels,
    width 0.02, eta 1e-4, regrid on |J| dx > 0.08 every 10 steps (buffer 2, patches of
    at most 32 cells). The timed steps include the regrids."""
    import torch

    from wellbalanced_amd import initial
    from wellbalanced_amd.amr import AMRMHD2D
    from wellbalanced_amd.diagnostics import sheet_diagnostics
    from wellbalanced_amd.hierarchy import AMRHierarchy

    levels, every = 3, 10
    res, models = {}, {}
    for precision in ("fp64", "mixed"):
        mesh = AMRHierarchy((256, 128), ((0.0, 4.0), (0.0, 2.0)), periodic=(True, True))
        state = initial.harris_sheets(mesh, width=0.02, perturbation=1e-3, noise=1e-4)
        m = AMRMHD2D(mesh, resistivity=1e-4, limiter="vanleer", precision=precision)
        m.set_state(**state)

        def regrid():
            m.regrid(m.current_tags(0.08, levels=levels), buffer=2, max_patch_size=32)

        for _ in range(levels):
            regrid()
            m.set_state(**state)
        m.check_interval = 0
        for _ in range(args.warmup):
            m.advance(1)
        torch.cuda.synchronize()
        zones, t_regrid = 0, 0.0
        t0 = time.perf_counter()
        for k in range(args.steps):
            if k and k % every == 0:
                r0 = time.perf_counter()
                regrid()
                t_regrid += time.perf_counter() - r0
            m.advance(1)
            zones += m.cells
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        m.check()
        _, stage_ms = time_model(m, 2, 0)
        res[precision] = (zones / sec, sec, stage_ms, t_regrid)
        models[precision] = m
    a, b = models["fp64"], models["mixed"]
    same = [p.box for p in a.hierarchy.patches] == [p.box for p in b.hierarchy.patches]
    l1, linf = _leaf_l1(a, b) if same else (None, None)
    rate, sec, stage_ms, t_regrid = res["mixed"]
    leaf = b.leaf_values()
    return _line(rate, res["fp64"][0], args.steps, args, sec,
                 "2D double Harris sheets, 256x128 base on [0,4]x[0,2], 3 levels regridded "
                 "every 10 steps on |J| dx > 0.08, eta 1e-4, van Leer, CFL 0.4, SSP-RK2 "
                 "(regrids inside the timed steps)", b.cells, "mixed", l1, linf, stage_ms,
                 STATE_BYTES_2D,
                 {"patches": len(b.layouts), "leaf_cells": int(leaf["density"].size),
                  "levels": b.hierarchy.max_level + 1, "regrid_seconds": t_regrid,
                  "fp64_regrid_seconds": res["fp64"][3], "same_hierarchy": same,
                  "divergence": b.divergence(), "composite_divergence": b.composite_divergence(),
                  "sheet": sheet_diagnostics(b, leaf)})


def _leaf_l1(a_model, b_model):
    """Volume-weighted mean and max |rho_a - rho_b| over the leaf cells of two
    hierarchy models on the same geometry (the domain is the unit square)."""
    import numpy as np

    from wellbalanced_amd.amr_maps import HALO

    buf_a, buf_b = a_model.buffer(), b_model.buffer()
    maps = a_model.maps
    covered = np.zeros(maps.total, bool)
    for L in range(a_model.nlev - 1):
        rows = maps.average[L]
        if len(rows):
            covered[rows[:, 0]] = True
    total, area, worst = 0.0, 0.0, 0.0
    for lay in maps.layouts:
        dx, dy = a_model.hierarchy.spacing(lay.level)
        i, j = np.meshgrid(np.arange(lay.nx) + HALO, np.arange(lay.ny) + HALO, indexing="ij")
        offs = lay.cell(i, j)[~covered[lay.cell(i, j)]]
        d = np.abs(buf_a[offs] - buf_b[offs])
        total += float(d.sum()) * dx * dy
        area += offs.size * dx * dy
        worst = max(worst, float(d.max()) if d.size else 0.0)
    return total / area, worst


def run_blast2d_amr(args):
    import torch

    from wellbalanced_amd import autotune
    from wellbalanced_amd.amr import AMRMHD2D, blast_hierarchy, blast_state

    h = blast_hierarchy(512, 3)
    res, models = {}, {}
    for precision in ("fp64", "mixed"):
        m = AMRMHD2D(h, limiter="minmod", mean_field=(2**-0.5, 2**-0.5), precision=precision)
        m.set_state(**blast_state())
        launch = autotune.tune(m, steps=2) if not args.no_tune else {}
        m.check_interval = 0
        _, stage_ms = time_model(m, 3, 1)
        sec, steps = _timed_steps(m, args, graph_steps=10)
        m.check()
        res[precision] = (m.cells * steps / sec, sec, stage_ms, launch, steps)
        models[precision] = m
    l1, linf = _leaf_l1(models["fp64"], models["mixed"])
    m = models["mixed"]
    leaf = int(m.cells - sum(len(m.maps.average[L]) for L in range(m.nlev - 1)))
    rate, sec, stage_ms, launch, steps = res["mixed"]
    return _line(rate, res["fp64"][0], steps, args, sec,
                 "2D blast on a 3-level hierarchy (512^2 base, patches of 16..128 cells), "
                 "minmod, CFL 0.4, SSP-RK2, one persistent launch per stage, steps replayed "
                 "from HIP graphs of 10 steps", m.cells, "mixed",
                 l1, linf, stage_ms, STATE_BYTES_2D,
                 {"launch": {"fp64": res["fp64"][3], "mixed": launch}, "patches": len(m.layouts),
                  "leaf_cells": leaf, "divergence": m.divergence(),
                  "composite_divergence": m.composite_divergence()})



def _footprint(h, level: int = 1):
    """Level-0 cells covered by the blocks of ``level``."""
    import numpy as np

    n0 = h.level_shape(0)[0]
    mask = np.zeros((n0,) * 3, bool)
    side = h.block // 2**level
    for b in h.levels[level] if len(h.levels) > level else ():
        sl = tuple(slice(v * side, (v + 1) * side) for v in b)
        mask[sl] = True
    return mask


def _restrict_arrays(cons, pots):
    """Haloed uniform arrays (``cons (5, n+4, n+4, n+4)``, edge potentials) restricted to
    half the resolution: each coarse cell the mean of its 8 children, each coarse edge
    the mean of the two fine edges along it (halos zero)."""
    import numpy as np

    n = cons.shape[1] - 4
    m = n // 2
    inner = cons[:, 2:-2, 2:-2, 2:-2].reshape(5, m, 2, m, 2, m, 2)
    out = np.zeros((5, m + 4, m + 4, m + 4))
    out[:, 2:-2, 2:-2, 2:-2] = inner.mean(axis=(2, 4, 6))
    coarse = []
    for c, p in enumerate(pots):
        ext = [n + (a != c) for a in range(3)]
        e = p[tuple(slice(2, 2 + ext[a]) for a in range(3))]
        e = e[tuple(slice(None) if a == c else slice(None, None, 2) for a in range(3))]
        lo = e[tuple(slice(0, None, 2) if a == c else slice(None) for a in range(3))]
        hi = e[tuple(slice(1, None, 2) if a == c else slice(None) for a in range(3))]
        q = np.zeros(tuple(m + 4 + (a != c) for a in range(3)))
        q[tuple(slice(2, 2 + m + (a != c)) for a in range(3))] = 0.5 * lo + 0.5 * hi
        coarse.append(q)
    return out, coarse


def _hierarchy_buffer(model, levels):
    """A host buffer of an `AMRMHD3D` whose every block copies its own level of the
    haloed uniform arrays ``levels[L] = (cons, pots)``."""
    import numpy as np

    from wellbalanced_amd.amr3d_maps import block_views

    buf = np.zeros(model.tables.total)
    b = model.hierarchy.block
    for row in model.tables.blocks:
        cons, pots = levels[int(row["level"])]
        cells, views = block_views(buf, int(row["off"]), b)
        o = [int(row[a]) * b for a in ("bx", "by", "bz")]
        cells[:, 2:-2, 2:-2, 2:-2] = cons[:, 2 + o[0] : 2 + o[0] + b, 2 + o[1] : 2 + o[1] + b,
                                          2 + o[2] : 2 + o[2] + b]
        for c in range(3):
            sl = tuple(slice(2, 2 + b + (a != c)) for a in range(3))
            src = tuple(slice(2 + o[a], 2 + o[a] + b + (a != c)) for a in range(3))
            views[c][sl] = pots[c][src]
    return buf


def judge_errors(problem, n0, hierarchies, end_time, factor: int = 4):
    """The plain estimator's question turned on the solution.  A uniform FP64 run
    ``factor`` times finer than the base (the finest level of the hierarchies) is the
    reference; every other run starts from its initial state restricted to its own
    levels (cell means, edge means), so that a refined block starts from the
    reference's data, and runs in FP64 to ``end_time``: each hierarchy, the unrefined
    base grid and a uniform grid of twice the base.  The level-0 density (covered
    cells averaged down) against the reference's, averaged onto the base cells: L1
    over all base cells and over those where the reference departs from 1 by more
    than 0.05 (the blast and its shell)."""
    import numpy as np
    import torch

    from wellbalanced_amd import MHD3D, initial
    from wellbalanced_amd.amr3d import AMRMHD3D

    ref = MHD3D.from_problem(initial.blast_turbulence_3d(factor * n0, seed=0), kernels="fused",
                             precision="fp64")
    levels = {int(np.log2(factor)): ref.arrays()}
    for L in range(int(np.log2(factor)) - 1, -1, -1):
        levels[L] = _restrict_arrays(*levels[L + 1])
    ref.check_interval = 0
    ref_steps = ref.run(end_time)
    r = ref.dataset()["density"].reshape(n0, factor, n0, factor, n0, factor).mean(axis=(1, 3, 5))
    del ref
    torch.cuda.empty_cache()
    blast = np.abs(r - 1.0) > 0.05
    out = {"reference": f"uniform {factor * n0}^3 FP64 run ({ref_steps} steps), averaged onto "
                        f"the {n0}^3 base; every run starts from its initial state restricted",
           "end_time": end_time, "blast_cells": int(blast.sum())}

    def errors(rho):
        d = np.abs(rho - r)
        return {"l1": float(d.mean()), "l1_blast": float(d[blast].mean())}

    opts = dict(gamma=problem.options.get("gamma", 5.0 / 3.0),
                mean_field=problem.options["mean_field"],
                limiter=problem.options.get("limiter", "minmod"), precision="fp64")
    for name, h in hierarchies.items():
        m = AMRMHD3D(h, **opts)
        m.load_buffer(_hierarchy_buffer(m, levels))
        m.check_interval = 0
        steps = 0
        while m.time < end_time * (1.0 - 1e-14):
            m.step(min(m.stable_step(), end_time - m.time))
            steps += 1
        m.check()
        out[name] = {**errors(m.level_state(0)[0][0][2:-2, 2:-2, 2:-2]), "steps": steps,
                     "leaf_cells": h.leaf_cells(), "blocks": [len(s) for s in h.levels]}
        del m
        torch.cuda.empty_cache()
    for name, L in (("unrefined", 0), ("uniform_2x", 1)):
        u = MHD3D.from_problem(initial.blast_turbulence_3d(n0 * 2**L, seed=0), kernels="fused",
                               precision="fp64")
        u.load_arrays(*levels[L])
        u.check_interval = 0
        steps = u.run(end_time)
        u.check()
        rho = u.dataset()["density"]
        if L:
            rho = rho.reshape(n0, 2, n0, 2, n0, 2).mean(axis=(1, 3, 5))
        out[name] = {**errors(rho), "steps": steps, "cells": (n0 * 2**L) ** 3}
        del u
    return out


def run_blast3d_amr(args, judge: bool = True):
    """The learned-indicator refined step: a blast in Alfvenic turbulence on a 128^3
    base with 16^3 blocks and 3 levels, refined by the CNN indicator (trained on the
    GPU on the Loehner flags 8 steps ahead) and, for comparison, by the plain Loehner
    flags; the hierarchies start from the uniform run at step 12.  The mixed-vs-FP64
    error is taken after 20 steps of both at the FP64 stable step; with ``judge`` the
    solution error of both hierarchies is judged against a finer uniform run."""
    import numpy as np
    import torch

    from wellbalanced_amd import MHD3D, autotune, indicator, initial
    from wellbalanced_amd.amr3d import AMRMHD3D

    n0, lead, start_step = 128, 8, 12
    problem = initial.blast_turbulence_3d(n0, seed=0)
    net, losses, train_s = indicator.train_lookahead(problem, lead=lead)
    # the start state and the flags each indicator raises there; the truth 8 steps on
    start = MHD3D.from_problem(problem, kernels="fused", precision="fp64")
    start.check_interval = 0
    start.advance(start_step)
    est = indicator.lohner(start)
    logit = net.predict(start)
    later = MHD3D.from_problem(problem, kernels="fused", precision="fp64")
    later.check_interval = 0
    later.advance(start_step + lead)
    truth = (indicator.lohner(later) > indicator.CUTOFF).cpu().numpy()
    end_time = later.time
    del later
    est_np, logit_np = est.cpu().numpy(), logit.cpu().numpy()
    hierarchies = {
        "learned": indicator.block_hierarchy(n0, logit_np > 0.0, logit_np > 3.0),
        "plain": indicator.block_hierarchy(n0, est_np > indicator.CUTOFF, est_np > 0.95),
    }
    coverage = {k: [float(_footprint(h, L)[truth].mean()) for L in (1, 2)]
                for k, h in hierarchies.items()}
    # the refined step on the learned hierarchy, mixed and FP64
    h = hierarchies["learned"]
    res, models, stage = {}, {}, {}
    launch = {}
    for precision in ("fp64", "mixed"):
        m = AMRMHD3D.from_uniform(start, h, precision=precision)
        m.check_interval = 0
        launch[precision] = autotune.tune(m, steps=2) if not args.no_tune else {}
        _, stage[precision] = time_model(m, 2, 1)
        models[precision] = m
    # both steps timed in alternating rounds (their ratio is the point of the workload)
    for precision, (sec, steps) in _alternating_steps(models, args, graph_steps=10).items():
        m = models[precision]
        m.check()
        res[precision] = (m.cells * steps / sec, sec, stage[precision], steps)
    # the error of the mixed step: both from the start state at the FP64 stable step
    pair = {p: AMRMHD3D.from_uniform(start, h, precision=p) for p in ("fp64", "mixed")}
    for _ in range(20):
        dt = pair["fp64"].stable_step()
        for m in pair.values():
            m.step(dt)
    a, b = pair["fp64"].level_state(0)[0][0], pair["mixed"].level_state(0)[0][0]
    d = np.abs(a - b)[2:-2, 2:-2, 2:-2]
    del pair
    m = models["mixed"]
    rate, sec, stage_ms, steps = res["mixed"]
    extra = {"blocks_per_level": [len(s) for s in h.levels], "launch": launch,
             "leaf_cells": h.leaf_cells(), "divergence": m.divergence(),
             "indicator": {"train_seconds": train_s, "loss_first": losses[0],
                           "loss_last": losses[-1],
                           "flags_8_steps_on_covered_by_levels_1_2": coverage,
                           "flagged_now": {"learned": int((logit_np > 0).sum()),
                                           "plain": int((est_np > indicator.CUTOFF).sum())},
                           "blocks_plain": [len(s) for s in hierarchies["plain"].levels]},
             "l1_note": "level-0 (averaged-down) density, mixed vs FP64 after 20 steps of "
                        "both at the FP64 stable step"}
    del models
    torch.cuda.empty_cache()
    if judge:
        extra["judge"] = judge_errors(problem, n0, hierarchies, end_time)
    return _line(rate, res["fp64"][0], steps, args, sec,
                 "3D blast in Alfvenic turbulence, 128^3 base, 3 levels of 16^3 blocks refined "
                 "by the learned indicator, minmod, CFL 0.3, SSP-RK2, steps replayed from HIP "
                 "graphs of 10 steps", m.cells, "mixed",
                 float(d.mean()), float(d.max()), stage_ms, STATE_BYTES_3D, extra)


if __name__ == "__main__":
    main()
