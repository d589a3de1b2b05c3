#!/usr/bin/env python3
"""Benchmark: FP64 sediment column-steps/s on one MI355X.

Workload (shelf-sea configuration): 300 x 300 columns, 40 % random-blob land
mask, 30 geometric layers (1 mm top, 0.2 m total), inputs spatially smoothed
(Gaussian, 10 cells), daily bottom-water forcing (seasonal sine plus seeded
noise), adaptive EMP with tolerance 1e-6 and per-column dt in [1, 3600] s.

A step is one simulated day of every wet column: one persistent launch of
the adaptive-EMP kernel (sedcol_integrate over [day, day+1)).  `--warmup`
days run first, untimed (they include the start-up transient from detritus-
free sediment), then `--steps` days are timed.  The value is accepted
column-steps per second of wall time over the timed days, inputs resident in
HBM.  The roofline entry prices the dominant kernel (k_emp_adaptive) with the
algorithmic FP64 operation counts of sedcol/flops.py over its average launch
time measured with HIP events on the library's stream.  The CPU baseline runs
the plain-C oracle on a bounded sample of the same columns and days on the
host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "mossco-code_amd"))

import numpy as np  # noqa: E402

METRIC = "FP64 sediment column-steps/s on 1 MI355X; spin-up wall time surrogate vs cold"
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
DAY = 86400.0

#: HBM bytes per k_emp_adaptive launch of the default workload (300 x 300,
#: seed 0, fast form, EMP, 5 untimed days then 20 timed days), from
#: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of that same command over
#: its 20 timed launches: mean FETCH_SIZE 458 544 KiB x 2.00 + WRITE_SIZE
#: 273 073 KiB x 1.00, the factors being what the counters read for a plain
#: 256 MiB copy (FETCH_SIZE 131 086 KiB, WRITE_SIZE 262 144 KiB).  Reported
#: for that window only.
DEFAULT_TRAFFIC_BYTES = (458544 * 2.0 + 273073 * 1.0) * 1024


def cpu_baseline(dom, state0, day0, ndays, seconds, nthreads, cols_per_thread=64):
    """Oracle (plain C, OpenMP over columns) on a sample of wet columns,
    starting from the same state; stops after `seconds` of work."""
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle as O
    from sedcol import cases

    wet = np.flatnonzero(dom.mask == 0)
    ncols = min(wet.size, cols_per_thread * nthreads)
    cols = wet[np.linspace(0, wet.size - 1, ncols).astype(np.int64)]
    sub = dom.subset(cols)
    off = dom.off
    x = np.concatenate([state0[:, off[c]:off[c + 1]] for c in cols], axis=1)
    x = np.ascontiguousarray(x)
    orc = O.Oracle(sub, sub.params.si())
    t0 = time.perf_counter()
    days = 0
    for d in range(ndays):
        spans = cases.record_spans((day0 + d) * DAY, (day0 + d + 1) * DAY, sub.nrec, sub.t0, sub.dt_rec)
        orc.integrate_series(x, (day0 + d) * DAY, (day0 + d + 1) * DAY, spans, nthreads=nthreads)
        days += 1
        if time.perf_counter() - t0 >= seconds:
            break
    el = time.perf_counter() - t0
    acc = int(orc.diag["accepted"].sum())
    return {
        "value": acc / el,
        "unit": "column-steps/s",
        "cores": nthreads,
        "kind": "port",
        "sample": f"{ncols} wet columns x {days} simulated days (days {day0}-{day0 + days - 1}), "
                  f"{acc} accepted column-steps in {el:.1f} s, oracle/sedcol_oracle.c -O3 OpenMP",
    }


def rk4_bench(args):
    """The plain RK4 case: 4 x 4 columns, K = 30, dt = 360 s, 30 simulated
    days (7200 steps) in one persistent launch, reference operation order.
    Beside it the reference's own ode_solver (solver_library.F90 compiled
    unmodified into oracle/_ref) with the oracle's right-hand side, on one
    core, from the same state; the final states are compared bitwise."""
    import torch

    from sedcol import cases, capi
    from sedcol.params import RUNGE_KUTTA_4

    torch.cuda.init()
    dom = cases.plain_columns(seed=args.seed)
    dom.params.arith = 0
    dom.params.clamp = 0  # ode_solver alone; the caller's clamp is not part of it
    nsteps = 7200
    g = capi.SedimentColumns.from_domain(dom)
    for _ in range(max(args.warmup, 1)):  # untimed: library and kernel start-up
        g.set_state(dom.conc0)
        g.run(360.0, nsteps, RUNGE_KUTTA_4)
    g.set_state(dom.conc0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.run(360.0, nsteps, RUNGE_KUTTA_4)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms, _ = g.last_timing()
    xg = g.get_state()
    g.close()
    colsteps = dom.ncol * nsteps
    out = {
        "metric": "FP64 RK4 column-steps/s on 1 MI355X (4x4 columns, K=30, dt 360 s, 30 days)",
        "value": colsteps / wall,
        "unit": "column-steps/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain RK4 4x4x30, dt 360 s, 7200 steps in one launch, no clamp", "seed": args.seed,
                   "arith": "reference"},
        "wall_s": wall,
        "kernel_ms": ms,
    }
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle as O

    if O.ref_available():
        x = np.ascontiguousarray(dom.conc0.copy())
        orc = O.Oracle(dom, dom.params.si())
        t0 = time.perf_counter()
        done = 0
        while done < nsteps:  # in chunks of 1 simulated day
            assert orc.ref_ode_solver(x, 360.0, RUNGE_KUTTA_4, nsteps=240) == 0
            done += 240
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": dom.ncol * nsteps / el, "unit": "column-steps/s", "cores": 1, "kind": "reference",
            "sample": f"the whole case, {dom.ncol} columns x {nsteps} steps in {el:.2f} s: solver_library.F90 "
                      "compiled from the reference (oracle/_ref), with the oracle's right-hand side",
        }
        out["results_bitwise_equal"] = bool(np.array_equal(xg, x))
    print(json.dumps(out), flush=True)
    return 0


def ragged_bench(args):
    """Ragged batch: 1e5 columns, K in 10..60, depth 0.1-0.5 m, one forcing
    record.  After an untimed warm-up the same span is integrated from the
    same checkpoint by each schedule (one persistent launch per layer bin,
    hipGraph-captured windows of --window attempts, one launch per attempt),
    first with the default kernel shape, then with the autotuned one; the
    final states must be bitwise equal across schedules (equal accuracy)."""
    import hashlib

    import torch

    from sedcol import cases, capi, flops

    torch.cuda.init()
    t = time.perf_counter()
    dom = cases.ragged_columns(ncol=args.ncol, seed=args.seed)
    dom.params.arith = args.arith
    dom.params.emp_weights = args.emp_weights
    gen_s = time.perf_counter() - t
    K = dom.nlayers
    span = args.span_hours * 3600.0
    t_w = args.warmup * DAY
    g = capi.SedimentColumns.from_domain(dom)
    if t_w > 0:
        g.integrate(0.0, t_w)
    ck = g.save_checkpoint()
    state_w = g.get_state()
    schedules = [
        ("persistent", dict(window=0, graph_windows=0, check_every=1, concurrent_bins=0)),
        ("persistent-concurrent-bins", dict(window=0, graph_windows=0, check_every=1, concurrent_bins=1)),
        (f"graph-window-{args.window}", dict(window=args.window, graph_windows=args.graph_windows, check_every=1,
                                              concurrent_bins=1)),
        ("launch-per-attempt", dict(window=1, graph_windows=0, check_every=64, concurrent_bins=0)),
    ]

    def run_all(label):
        rows = []
        for name, sch in schedules:
            g.load_checkpoint(ck)
            g.set_schedule(**sch)
            d0 = g.diagnostics()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            g.integrate(t_w, t_w + span)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            ms, nl = g.last_timing()
            d1 = g.diagnostics()
            acc = d1["accepted"] - d0["accepted"]
            att = acc + (d1["rejected"] - d0["rejected"])
            newt = d1["newton"] - d0["newton"]
            x = g.get_state()
            rows.append({
                "schedule": name, "shape": label, "wall_s": wall, "kernel_ms": ms, "launches": nl,
                "column_steps_per_s": float(acc.sum()) / wall,
                "tflops": flops.emp_flops(K, att, newt, weights=args.emp_weights) / wall / 1e12,
                "state_sha1": hashlib.sha1(x.tobytes()).hexdigest()[:16],
            })
            print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
        return rows, int(acc.sum())

    rows, colsteps = run_all("default")
    g.load_checkpoint(ck)
    g.set_schedule(window=0)
    t = time.perf_counter()
    best, table = g.autotune(t_w, args.tune_span, repeats=2)
    tune_s = time.perf_counter() - t
    rows_t, _ = run_all("autotuned")
    g.close()
    allrows = rows + rows_t
    same = len({r["state_sha1"] for r in allrows}) == 1
    best_row = max(allrows, key=lambda r: r["column_steps_per_s"])
    step_row = [r for r in rows_t if r["schedule"] == "launch-per-attempt"][0]
    out = {
        "metric": "FP64 sediment column-steps/s on 1 MI355X (ragged batch)",
        "value": best_row["column_steps_per_s"],
        "unit": "column-steps/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"ragged {args.ncol} columns, K U{{10..60}}, depth U[0.1,0.5] m, adaptive EMP tol 1e-6, "
                        f"dt in [1,3600] s; {args.warmup} warm-up day(s), then {args.span_hours} h timed",
            "cells": int(dom.ncell), "seed": args.seed, "arith": "fast" if args.arith else "reference",
            "emp_weights": "geometric" if args.emp_weights else "product",
        },
        "best": {"schedule": best_row["schedule"], "shape": best_row["shape"]},
        "speedup_vs_launch_per_attempt": best_row["column_steps_per_s"] / step_row["column_steps_per_s"],
        "results_bitwise_equal": same,
        "accepted_column_steps": colsteps,
        "autotune": {"tuning": best, "seconds": tune_s, "span_s": args.tune_span,
                     "ms_table_bin_layers_per_lane_lanes_workgroup": np.round(table, 3).tolist()},
        "rows": allrows,
        "input_generation_s": gen_s,
    }
    if not args.no_cpu:
        # the oracle on a sample of the same ragged columns, same span from
        # the same warmed state, on up to 16 host cores and on one
        nthreads = min(16, os.cpu_count() or 1)
        ndays = max(1, int(round(span / DAY)))
        # a 10 240-column subset (one day is a few seconds of oracle work)
        out["cpu_baseline"] = cpu_baseline(dom, state_w, args.warmup, ndays, args.cpu_seconds, nthreads,
                                           cols_per_thread=10240 // nthreads)
        one = cpu_baseline(dom, state_w, args.warmup, ndays, args.cpu_seconds / 2, 1, cols_per_thread=600)
        out["cpu_baseline"]["one_core"] = {k: one[k] for k in ("value", "unit", "cores", "sample")}
    print(json.dumps(out), flush=True)
    return 0 if same else 1


def spinup_bench(args):
    """Spin-up wall time, surrogate warm start against the plain cold
    spin-up (the second half of the metric): sedcol.surrogate.evaluate on
    --spin-columns drawn columns (layers U{20..40}, 5 % of the inputs outside
    the training ranges), cold = the reference's presimulation of
    --spin-years per column, warm = prediction + integration retiring each
    column at residual < 1e-6 with every top flux within 1 % of the cold
    result, fallback to the cold spin-up after 30 days; one stepper for all
    (--spin-method, steps up to one day)."""
    import torch

    from sedcol import surrogate

    torch.cuda.init()
    model = surrogate.Model.load(REPO / "mossco-code_amd" / "sedcol" / "models" / "spinup_mlp.npz")
    r = surrogate.evaluate(model, args.spin_columns, args.seed if args.seed else 7, years=args.spin_years,
                           warm_days=30.0, adaptive_method=args.spin_method, dt_max=86400.0,
                           log=lambda m: print(m, file=sys.stderr, flush=True))
    out = {
        "metric": "spin-up wall time to steady state, plain cold spin-up / surrogate warm start",
        "value": r["speedup_wall"],
        "unit": "x",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f64 integration, bf16 network",
        "data": "synthetic",
        "config": {"workload": f"{args.spin_columns} columns, K U{{20..40}}, 5% inputs outside the training "
                               f"ranges; cold = {args.spin_years} y presimulation per column; warm = BF16 MLP "
                               "prediction + integration to residual < 1e-6 and fluxes within 1% of cold, "
                               "fallback to cold after 30 d",
                   "adaptive_method": args.spin_method, "seed": args.seed if args.seed else 7},
        **{k: v for k, v in r.items() if k != "per_variable"},
        "per_variable": r["per_variable"],
    }
    print(json.dumps(out), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed simulated days")
    ap.add_argument("--warmup", type=int, default=5, help="untimed simulated days first")
    ap.add_argument("--n", type=int, default=300, help="grid is n x n columns")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--arith", type=int, default=1, help="0 reference operation order, 1 fast form")
    ap.add_argument("--emp-weights", type=int, default=0,
                    help="Patankar factor of adaptive EMP: 0 product (default), 1 geometric mean")
    ap.add_argument("--workload", choices=("shelf", "ragged", "rk4", "spinup"), default="shelf",
                    help="shelf: the benchmark line; ragged: schedule/autotuner comparison on ragged columns; "
                         "rk4: the plain 4x4 RK4 case beside the reference's own stepper on the CPU; "
                         "spinup: surrogate warm start against cold spin-up (wall-time ratio)")
    ap.add_argument("--spin-columns", type=int, default=100000,
                    help="spinup: columns (1e5: the run is bound by its total work, about 11 minutes; "
                         "4000 fits on the GPU at once, so the slowest column sets the wall time)")
    ap.add_argument("--spin-years", type=float, default=2.0, help="spinup: cold presimulation (years)")
    ap.add_argument("--spin-method", type=int, default=7, help="spinup: 6 adaptive EMP, 7 linearly implicit")
    ap.add_argument("--adaptive-method", type=int, default=6,
                    help="shelf: adaptive stepper, 6 EMP (the configuration's) or 7 linearly implicit")
    ap.add_argument("--ncol", type=int, default=100000, help="ragged: columns")
    ap.add_argument("--span-hours", type=float, default=24.0, help="ragged: timed span per schedule")
    ap.add_argument("--window", type=int, default=100, help="ragged: attempts per column per window")
    ap.add_argument("--graph-windows", type=int, default=8, help="ragged: windows per captured graph")
    ap.add_argument("--tune-span", type=float, default=3600.0, help="ragged: autotuner trial span (s)")
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(columns are independent: a multi-GPU run would give each process its own share of the "
              "wet columns).", file=sys.stderr)
        return 2
    if args.workload == "ragged":
        return ragged_bench(args)
    if args.workload == "rk4":
        return rk4_bench(args)
    if args.workload == "spinup":
        return spinup_bench(args)

    import torch

    from sedcol import budget, cases, capi, flops

    torch.cuda.init()
    days = args.warmup + args.steps
    t = time.perf_counter()
    dom = cases.shelf_sea(n=args.n, days=days, seed=args.seed)
    dom.params.arith = args.arith
    dom.params.emp_weights = args.emp_weights
    dom.params.adaptive_method = args.adaptive_method
    gen_s = time.perf_counter() - t
    wet = dom.mask == 0
    K = dom.nlayers
    g = capi.SedimentColumns.from_domain(dom)
    inv_start = g.inventory()
    if args.warmup:
        g.integrate(0.0, args.warmup * DAY)
    state_w = g.get_state()
    d0 = g.diagnostics()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kernel_ms = 0.0
    launches = 0
    for d in range(args.steps):
        day = args.warmup + d
        g.integrate(day * DAY, (day + 1) * DAY)
        ms, nl = g.last_timing()
        kernel_ms += ms
        launches += nl
        if (d + 1) % 30 == 0:
            print(f"bench: {d + 1} of {args.steps} days", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    d1 = g.diagnostics()
    wall = t1 - t0
    acc = (d1["accepted"] - d0["accepted"])[wet]
    rej = (d1["rejected"] - d0["rejected"])[wet]
    newt = (d1["newton"] - d0["newton"])[wet]
    colsteps = int(acc.sum())
    value = colsteps / wall
    ros = args.adaptive_method == 7
    fl = flops.ros2_flops(K[wet], acc + rej) if ros else flops.emp_flops(K[wet], acc + rej, newt, weights=args.emp_weights)
    avg_launch_s = kernel_ms * 1e-3 / max(launches, 1)
    achieved = fl / max(launches, 1) / avg_launch_s / 1e12
    xs = g.get_state()
    land_cells = np.repeat(dom.mask, K) != 0
    min_state = float(xs[:, ~land_cells].min())
    # C/N/P closure over the whole run (warm-up included), outside the timing
    res = budget.closure(inv_start, g.inventory(), g.budget())
    closure_max = {e: float(r[wet].max()) for e, r in res.items()}
    g.close()
    default_workload = (args.n == 300 and args.seed == 0 and args.arith == 1 and args.emp_weights == 0
                        and not ros and args.steps == 20 and args.warmup == 5)
    traffic = DEFAULT_TRAFFIC_BYTES if default_workload else None

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "column-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"shelf-sea {args.n}x{args.n} columns, 40% land, K=30 geometric (1 mm top, 0.2 m), "
                        "adaptive EMP tol 1e-6, dt in [1,3600] s, daily forcing; step = 1 simulated day",
            "columns": int(dom.ncol),
            "wet_columns": int(wet.sum()),
            "layers": 30,
            "timed_days": args.steps,
            "seed": args.seed,
            "arith": "fast" if args.arith else "reference",
            "emp_weights": "geometric" if args.emp_weights else "product",
            "adaptive_method": "linearly implicit ROS2 (7)" if ros else "EMP (6)",
            "forcing_phase": f"days {args.warmup}-{args.warmup + args.steps - 1} of a seasonal cycle whose "
                             "coldest, most oxygenated bottom water is on day 0 (sin(2 pi d/365 - pi/2))",
        },
        "roofline": {
            "bound": "fp64-valu",
            "pipe": "FP64 VALU (MI355X FP64 vector peak 78.6 TFLOP/s; no matrix work on this path)",
            "kernel": "k_ros_adaptive<32>" if ros else "k_emp_adaptive<32>",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": traffic,
            "traffic_unit": "bytes/launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated)" if traffic else None,
            "algorithmic_bytes_per_launch": flops.launch_bytes(K[wet], records=1),
            "algorithmic_flop_per_launch": fl / max(launches, 1),
            "avg_launch_ms": avg_launch_s * 1e3,
        },
        "accepted_column_steps": colsteps,
        "simulated_column_days_per_s": float(wet.sum()) * args.steps / wall,
        "rejected_attempts": int(rej.sum()),
        "newton_iterations": int(newt.sum()),
        "min_state": min_state,
        "budget_closure_max": closure_max,
        "input_generation_s": gen_s,
    }
    if not args.no_cpu:
        nthreads = min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(dom, state_w, args.warmup, args.steps, args.cpu_seconds, nthreads)
        # the same oracle on one host core, a smaller sample of the same days
        one = cpu_baseline(dom, state_w, args.warmup, args.steps, args.cpu_seconds / 2, 1, cols_per_thread=48)
        out["cpu_baseline"]["one_core"] = {k: one[k] for k in ("value", "unit", "cores", "sample")}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
