"""Benchmark of the device RKHEVI step (one MI355X).

Workload: icosahedral grid res 7 (163842 cells, 491520 edges) x 26 layers,
moist (6 constituents), flat surface, the analytic balanced state with seeded
noise, dt = 120 s -- the configuration the simulated-days-per-wall-hour metric
is quoted on.  A step is one RKHEVI time step (manage_rkhevi plus the driver's
state copy) on inputs resident in HBM.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--res 7] [--no-cpu]
"""
import argparse
import json
import math
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, vendor figure


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(dims, moist, orography):
    """Minimal-pass traffic of one step in bytes: the phase-schedule model
    (FP64 doubles per cell-layer per step: 172 dry flat, 202 dry with
    orography, 262 moist flat) times the number of cell-layers."""
    per_cell_layer = 262 if moist else (202 if orography != "flat" else 172)
    return per_cell_layer * 8.0 * dims.s


def kernel_bytes(name, dims):
    """Algorithmic HBM bytes of one launch of a kernel: every distinct array
    element the kernel must read or write once (double = 8 B, int = 4 B;
    neighbour tables shared by all layers counted once per launch)."""
    S, V, NSH, NVH, L, NT = dims.s, dims.v, dims.nsh, dims.nvh, dims.n_layers, dims.nt
    E = L * NVH                     # edge values per vector field
    W_ = (L - 1) * NSH              # interior w values
    nc = dims.n_const
    moist = dims.n_const > 1
    table = {
        # per column: reads rho_gas, rhotheta_v, theta_v_pert, exner_pert of the old
        # and new state, the tendencies of rho_gas, rhotheta_v and w, w old, z_scalar,
        # exner_bg, volume, area at w, damping; writes rho_gas, rhotheta_v,
        # theta_v_pert, exner_pert and w
        "k_ver_waves": 8 * (S * (4 + 4 + 2 + 4) + W_ * (1 + 1 + 1 + 1) + W_ * 1 + S * 4 + W_ * 1),
        # all tracers of a column: rho_k old, tendency and new (5 x 3 S), w old and new,
        # area at w, volume, temperature, sediment heat
        "k_gen_densities": 8 * ((nc - 1) * 3 * S + 3 * W_ + 3 * S),
        # TRSK flux: mass flux (edges + w), PV (tangential + rhombus), TRSK tables once, ipw slots 6/7, output
        "k_vorticity_flux_e": 8 * (E + (L + 1) * NSH + (L + 1) * NVH + L * NVH + 2 * S + E) + 4 * 20 * NVH + 8 * 10 * NVH,
        "k_vorticity_flux_w": 8 * (E + dims.curl // 2 + 6 * S + (L + 1) * NSH) + 4 * 6 * NSH,
        "k_pot_vort": 8 * (dims.curl + L * NT + E + (L + 1) * NSH + S + dims.curl // 2) + 4 * 8 * NVH + 8 * 4 * NVH,
        "k_rel_vort_tri": 8 * (E + E + L * NT + L * NT) + 4 * 6 * NT,
        # all constituents: rho_k, rhotheta_v, wind at edges, area, volume, tendencies
        # (read + write); moist adds phase rates, temperature, heating terms, exner
        "k_scalar_tend": 8 * (S * (nc + 1) + 2 * E + S + 2 * S * (nc + 1) + (S * (nc + 6) if moist else 0)) + 4 * 12 * NSH,
        "k_vector_tend": 8 * (V * 6 + V + 2 * E),
        "k_pgrad_e": 8 * (E * 5 + S * (3 + nc)) + 4 * 2 * NVH,
        "k_pgrad_w": 8 * (W_ * 6 + S * (3 + nc)),
        "k_grad_e": 8 * (S + E + E) + 4 * 2 * NVH,
        "k_grad_w": 8 * (S + 2 * W_),
        "k_inner_product": 8 * (V + 8 * S + S) + 4 * 6 * NSH,
        "k_mass_flux": 8 * (S + 2 * (E + W_)) + 4 * 2 * NVH,
        "k_phase_trans": 8 * (S * (nc + 1) + S * 6),
        "k_temperature": 8 * (S * 5 + S),
        "k_lincomb_scalars": 8 * (S * (nc + 3) * 2 + S * (nc + 3)),
        "k_lincomb_vector": 8 * (2 * V + V),
    }
    return table.get(name)


def cpu_info():
    model = platform.processor() or ""
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    return model


def cpu_baseline(grid, state0, dt, steps):
    """The oracle (C/OpenMP restatement of the reference step, bit-identical to
    it) timed on the host cores; bounded sample of `steps` steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    om = pyoracle.OracleModel(grid, state0, dt)
    om.run(1)                        # warm-up step
    t0 = time.perf_counter()
    om.run(steps)
    sec = time.perf_counter() - t0
    sec_per_step = sec / steps
    return {"value": round((dt / 86400.0) / (sec_per_step / 3600.0), 3), "unit": "sim-days/wall-hour",
            "cores": threads, "kind": "port",
            "sample": f"{steps} steps (after 1 warm-up) of the same res {grid.dims.res_id} x "
                      f"{grid.dims.n_layers} moist workload, {sec_per_step*1e3:.1f} ms/step, "
                      f"OMP_NUM_THREADS={threads}, {cpu_info()}"}


def copy_probe(mib):
    """A plain device-to-device copy of known size: the reference point that
    scales the profiler's FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    n = mib * 2**20 // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8, "launches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--res", type=int, default=7)
    ap.add_argument("--layers", type=int, default=26)
    ap.add_argument("--dry", action="store_true")
    ap.add_argument("--dt", default="auto",
                    help="seconds, or 'auto' for the reference rule 1.61e-3*eff_hor_res (src/coordinator.c:155-171)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--lloyd", type=int, default=2000)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates HBM byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU design is horizontal domain decomposition, not built here)", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        copy_probe(args.copy_probe)
        return

    from game_archived_amd.grid import Grid
    from game_archived_amd.rkhevi import DeviceModel, NEW
    from game_archived_amd.state import initial_state

    moist = not args.dry
    t0 = time.time()
    n_oro = 23 if args.layers == 26 else int(round(args.layers * 23 / 26))
    grid = Grid(args.res, n_layers=args.layers, n_oro_layers=n_oro, moisture_on=1 if moist else 0,
                n_lloyd=args.lloyd)
    log(f"grid res {args.res} x {args.layers}: {grid.dims.nsh} cells, {grid.dims.nvh} edges, "
        f"built in {time.time() - t0:.1f} s")
    d = grid.dims
    if args.dt == "auto":
        # src/coordinator.c:155-171: mean w-point area over all levels
        a = grid.g["area"]
        idx = (np.arange(d.n_levels)[:, None] * d.vpl + np.arange(d.nsh)[None, :]).ravel()
        eff_hor_res = math.sqrt(a[idx].sum() / (d.n_levels * d.nsh))
        args.dt = 1.61e-3 * eff_hor_res
    args.dt = float(args.dt)
    dm = DeviceModel(grid)
    t0 = time.time()
    state0 = initial_state(grid, dm.balance_terms(), moist=moist)
    log(f"initial state in {time.time() - t0:.1f} s; device memory {dm.bytes_allocated() / 2**30:.2f} GiB")
    dm.start(state0)
    i0 = dm.integrals()
    dm.run(args.warmup, args.dt)
    dm.synchronize()
    dm.reset_timing()
    dm.enable_timing(True)
    t0 = time.perf_counter()
    dm.run(args.steps, args.dt)
    dm.synchronize()
    wall = time.perf_counter() - t0
    dm.enable_timing(False)
    stats = dm.kernel_stats()
    i1 = dm.integrals()
    final = dm.get_state(NEW)
    dm.close()
    finite = bool(np.isfinite(final.rho).all() and np.isfinite(final.wind).all())
    ms_per_step = wall / args.steps * 1e3
    sdph = (args.dt / 86400.0) / (ms_per_step / 1e3 / 3600.0)
    d = grid.dims
    # per-kernel table and the dominant kernel's roofline
    total_ms = sum(ms for _, ms in stats.values())
    rows = sorted(stats.items(), key=lambda kv: -kv[1][1])
    log(f"{'kernel':24s} {'launches':>8s} {'avg us':>9s} {'share':>6s} {'GB/s':>8s}")
    for name, (n, ms) in rows:
        b = kernel_bytes(name, d)
        gbs = (b / (ms / n * 1e-3) / 1e9) if b else float("nan")
        log(f"{name:24s} {n:8d} {ms / n * 1e3:9.1f} {ms / total_ms * 100:5.1f}% {gbs:8.1f}")
    dom_name, (dom_n, dom_ms) = rows[0]
    dom_bytes = kernel_bytes(dom_name, d)
    dom_avg_s = dom_ms / dom_n * 1e-3
    achieved = dom_bytes / dom_avg_s / 1e9
    step_bytes = algorithmic_bytes(d, moist, "flat")
    step_gbs = step_bytes / (ms_per_step * 1e-3) / 1e9
    mass_drift = (i1[d.n_cond] - i0[d.n_cond]) / i0[d.n_cond]
    dry_drift = ((i1[d.n_cond] - (i1[d.n_cond + 1] if moist else 0.0))
                 - (i0[d.n_cond] - (i0[d.n_cond + 1] if moist else 0.0))) / (i0[d.n_cond] - (i0[d.n_cond + 1] if moist else 0.0))
    out = {
        "metric": "sim-days/wall-hour res7x26L mixed vs FP64 (1 GPU); adjoint/forward cost ratio",
        "value": round(sdph, 2),
        "unit": "sim-days/wall-hour",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"RKHEVI step, icosahedral res {args.res} x {args.layers} layers, "
                        f"{'moist (6 constituents)' if moist else 'dry'}, flat, dt {args.dt:g} s, FP64 form",
            "cells": d.nsh, "edges": d.nvh, "layers": d.n_layers, "dt_s": round(args.dt, 3),
            "dt_rule": "reference auto time step 1.61e-3*eff_hor_res (src/coordinator.c:171)",
            "precision_form": "fp64", "lloyd_iterations": args.lloyd,
            "column_updates_per_s": round(d.nsh * args.steps / wall, 1),
            "step_algorithmic_GBps": round(step_gbs, 1),
            "step_roofline_frac": round(step_gbs / HBM_PEAK_GBS, 4),
            "gas_mass_rel_drift": float(mass_drift),
            "dry_mass_rel_drift": float(dry_drift),
            "finite": finite,
        },
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "avg_launch_us": round(dom_avg_s * 1e6, 2), "algorithmic_bytes": dom_bytes},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(grid, state0, args.dt, args.cpu_steps)
    print(json.dumps(out), flush=True)
    if not finite:
        sys.exit(3)


if __name__ == "__main__":
    main()
