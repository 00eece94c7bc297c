"""Throughput of the DSMC step (move + sort + collide) on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload roofline|precision|cpu]

Workload (default "roofline"): 256^3 voxel periodic box of polydisperse
spheres (radius log-uniform 3..10 voxels, porosity 0.45, seed 3), 64^3
collision cells of 4^3 voxels, 1e8 hard-sphere argon particles at Kn 0.05
against the voxel edge, dt = 0.2 mean collision times, diffuse walls at
300 K, body force along x.  State is synthetic and resident in HBM before
timing starts.  One step = one System::step (accelerate, move with wall
reflection, stable counting sort into cells, NTC collisions), per-cell FP64
moment sampling included.

Prints one JSON line (the driver contract) with a `roofline` object for the
dominant kernel (algorithmic bytes: 56 B per particle = the 28-byte state
read and written once) and a `cpu_baseline` object: the reference solver
(oracle/_ref, compiled from the reference sources) when present, else the
FP64 oracle, timed on this host's cores on a bounded sub-box sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
STATE_BYTES = 28


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    from master_amd import config as C
    from master_amd import dsmc
    if name == "roofline":
        p = C.ROOFLINE
        w = dsmc.WorldData.poly_spheres((p["n_vox"],) * 3, p["rmin"], p["rmax"], p["porosity"], p["geom_seed"])
    elif name == "precision":
        p = C.PRECISION
        w = dsmc.WorldData.packed_spheres(p["n_vox"], p["radius_vox"], p["porosity"], p["geom_seed"])
    else:
        p = C.CPU_SIZE
        w = dsmc.WorldData.packed_spheres(p["n_vox"], p["radius_vox"], p["porosity"], p["geom_seed"])
    ini = C.dsmc_settings(p["n_vox"], w.porosity, p["n_particles"], p["kn"], p["cells"], seed=1)
    return p, w, ini


def cpu_baseline(p, budget_s=20.0):
    """Reference solver (or oracle) on a bounded sub-box of the same workload:
    the same geometry statistics, Kn and particles per pore voxel in a 64^3
    box, stepped until ~budget_s of CPU time.  Single core."""
    import numpy as np
    from master_amd import config as C
    from oracle import binding as B
    n_sub = 64
    if "rmin" in p:
        wd = B.poly_spheres((n_sub,) * 3, p["rmin"], p["rmax"], p["porosity"], p["geom_seed"])
    else:
        wd = B.packed_spheres(n_sub, p["radius_vox"], p["porosity"], p["geom_seed"])
    nsub = int(p["n_particles"] * (n_sub / p["n_vox"]) ** 3)
    ini = C.dsmc_settings(n_sub, wd["porosity"], nsub, p["kn"], n_sub // 4, seed=1)
    kind = "reference" if B.ref_available() else "port"
    try:
        sysobj = B.RefSystem(ini, wd) if kind == "reference" else None
    except Exception as e:  # pragma: no cover - reported, then the port is used
        log("reference harness unavailable:", e)
        kind, sysobj = "port", None
    if sysobj is None:
        sysobj = B.OracleSystem(B.params_from_settings(ini, wd, rng_mode=B.RAN1), wd)
        sysobj.seed()
    n = sysobj.n
    steps, t0 = 0, time.process_time()
    w0 = time.time()
    while time.process_time() - t0 < budget_s and steps < 2000:
        sysobj.step()
        steps += 1
    el = time.process_time() - t0
    wall = time.time() - w0
    return {"value": n * steps / el, "unit": "particle-steps/s", "cores": 1, "kind": kind,
            "sample": f"{n_sub}^3 sub-box of the workload ({n} particles, same Kn, porosity and "
                      f"sphere radii), {steps} steps, {el:.1f} s CPU ({wall:.1f} s wall), FP64"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="roofline", choices=["roofline", "precision", "cpu"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--lds-cap", type=int, default=0, help="particles per cell held in LDS (0: automatic)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device-to-device copy (rocprof counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 1 << 28                                 # 2^28 float32 = 1 GiB
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the DSMC step here is a single-GPU path "
              "(multi-GPU sharding is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from master_amd import dsmc
    t_setup = time.time()
    p, w, ini = workload(args.workload)
    log(f"[bench] world {w.n} porosity {w.porosity:.4f} built in {time.time() - t_setup:.1f}s")
    g = dsmc.System(ini, w)
    n = g.initialize()
    g.set_sampling(1)
    if args.lds_cap:
        g.set_option(dsmc.OPT_LDS_CAPACITY, args.lds_cap)
    log(f"[bench] {n} particles seeded+sorted ({time.time() - t_setup:.1f}s), "
        f"max cell occupancy {g.counters()['max_cell_occupancy']}")

    # per-kernel device times (HIP events on the handle's stream), eager launches
    g.set_option(dsmc.OPT_TIMING, 1)
    g.step(args.profile_steps)
    ktimes = g.kernel_times()
    g.set_option(dsmc.OPT_TIMING, 0)
    log(f"[bench] kernel ms/launch: {ktimes}")

    g.set_option(dsmc.OPT_USE_GRAPH, 1)
    g.step(args.warmup)
    g.synchronize()
    log("[bench] warmup done")
    t0 = time.perf_counter()
    g.step(args.steps)
    g.synchronize()
    el = time.perf_counter() - t0
    c = g.counters()
    if c["errors"]:
        log(f"[bench] device error counter {c['errors']}")
    value = n * args.steps / el
    ms = 1000 * el / args.steps
    log(f"[bench] {ms:.3f} ms/step, {value:.4g} particle-steps/s, collisions/step "
        f"{c['collisions'] / max(c['steps'], 1):.0f}, overflow cells {c['overflow_cells']}")

    dom = max(ktimes, key=ktimes.get)
    achieved = 2 * STATE_BYTES * n / (ktimes[dom] * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "bytes_per_particle": 2 * STATE_BYTES,
            "kernel_ms": {k: round(v, 4) for k, v in ktimes.items()},
            "step_bytes_per_particle_pipeline": 4 * STATE_BYTES,
            "step_frac": round(4 * STATE_BYTES * n / (el / args.steps) / 1e9 / HBM_PEAK_GBS, 4)}
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(p, args.cpu_budget)
        except Exception as e:  # reported, never hides the GPU number
            log(f"[bench] cpu baseline failed: {e}")
    out = {
        "metric": "particle-steps/s (move+sort+collide)",
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {w.n[0]}^3 voxels, {g.ncells} cells, {n} particles, "
                               f"Kn {p['kn']}, porosity {w.porosity:.3f}",
                   "particles": n, "voxels": list(w.n), "cells": g.ncells, "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
