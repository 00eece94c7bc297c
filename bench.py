"""Throughput of the DSMC step (move + sort + collide) on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload roofline|precision|cpu|ensemble]

Workload (default "roofline"): 256^3 voxel periodic box of polydisperse
spheres (radius log-uniform 3..10 voxels, porosity 0.45, seed 3), 64^3
collision cells of 4^3 voxels, 1e8 hard-sphere argon particles at Kn 0.05
against the voxel edge, dt = 0.2 mean collision times, diffuse walls at
300 K, body force along x.  State is synthetic and resident in HBM before
timing starts.  One step = one System::step (accelerate, move with wall
reflection, stable counting sort into cells, NTC collisions), per-cell FP64
moment sampling included.

Prints one JSON line (the driver contract) with
* `roofline`: the dominant kernel's algorithmic bytes per launch over its
  mean launch time (HIP events on the library's stream), against the 8 TB/s
  HBM peak.  Algorithmic bytes per particle: k_move 56 (28-byte state read
  and written), k_gather 56 (state read from the old order, written in cell
  order), k_collide 24 (velocities read and written), k_gather_collide 56
  (the fused gather + collide pass: the collide adds no traffic) and
  k_gather_collide_move 56 (the same pass also moving its cells for the next
  step, which the library runs for all but the last step of a dsmc_step
  call: the move works on the data the pass has just written); the step's
  reference count is 140 (sort row 60 = gather 56 + the 4-byte cell key,
  read inside k_move; collide 24).  The eager per-kernel profile is one
  dsmc_step call, so k_move runs on its first step only.
  `traffic` is the HBM bytes one launch of that kernel really moves, from two
  rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) over the same workload,
  each scaled by what the counter reads for a plain 1 GiB copy in the same
  process; null when rocprofv3 is unavailable.
* `cpu_baseline`: the reference solver (oracle/_ref, compiled from the
  reference sources) when present, else the FP64 oracle, timed on one host
  core on a bounded sub-box sample of the same workload.
"""
import argparse
import glob
import json
import os
import re
import shutil
import sqlite3
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per particle and launch
KERNEL_BYTES = {"k_move": 56, "k_gather": 56, "k_collide": 24, "k_gather_collide": 56, "k_gather_collide_move": 56}
STEP_BYTES = 140
COPY_BYTES = 1 << 30


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


def ensemble_system(args):
    """64 independent members (edges 32..128, porosity 0.3..0.8, radius 2..8,
    Kn 0.01..1, Mach 0.01..0.1, seeds 100..163) in one handle, autotuned."""
    from master_amd import config as C
    from master_amd import dsmc
    t_setup = time.time()
    specs = C.ensemble_members()
    inis, worlds = [], []
    for sp in specs:
        e, r = sp["edge"], sp["radius"]
        w = dsmc.WorldData.poly_spheres((e, e, e), r, r, sp["porosity"], sp["seed"])
        inis.append(C.member_settings(sp, w.porosity))
        worlds.append(w)
    g = dsmc.System(inis, worlds)
    n = g.initialize()
    g.set_sampling(1)
    log(f"[bench] ensemble of {len(specs)} members, {n} particles, {g.ncells} cells "
        f"({time.time() - t_setup:.1f}s)")
    tune = g.autotune(8)
    log(f"[bench] autotune: {tune}")
    p = {"kn": "0.01-1 (64 members)"}
    return p, worlds, g, n, tune


def make_system(args):
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
    return p, w, g, n


class _StdoutToStderr:
    """The reference solver prints progress on stdout; keep the bench's
    stdout to the one JSON line by pointing fd 1 at stderr meanwhile."""

    def __enter__(self):
        import ctypes
        self._libc = ctypes.CDLL(None)
        sys.stdout.flush()
        self._saved = os.dup(1)
        os.dup2(2, 1)
        return self

    def __exit__(self, *exc):
        sys.stdout.flush()
        self._libc.fflush(None)
        os.dup2(self._saved, 1)
        os.close(self._saved)
        return False


def cpu_baseline(p, budget_s=20.0):
    with _StdoutToStderr():
        return _cpu_baseline(p, budget_s)


def _cpu_baseline(p, budget_s):
    """Reference solver (or oracle) on a bounded sub-box of the same workload:
    the same geometry statistics, Kn and particles per pore voxel in a 64^3
    box, stepped until ~budget_s of CPU time.  Single core."""
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


def pmc_pass(args):
    """Body of one counter pass (run under rocprofv3 by measure_traffic):
    a plain 1 GiB device copy three times (calibration), then the workload's
    kernels for a few eager steps."""
    import torch
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    torch.cuda.empty_cache()
    _, _, g, n = make_system(args)
    g.step(args.pmc_steps)
    g.synchronize()
    log(f"[pmc pass] {args.pmc_steps} steps done, {n} particles")


def _short(name):
    m = re.search(r"\b(k_\w+)(<[^>]*>)?\(", name)
    if not m:
        return name
    # the fused pass's instantiation with the next step's move
    if m.group(1) == "k_gather_collide" and m.group(2) and "true" in m.group(2):
        return "k_gather_collide_move"
    return m.group(1)


def _read_counter(db, counter, steps):
    """Mean counter value (KiB) per dispatch: the last `steps` dispatches of
    each library kernel, and the 1 GiB copy dispatches (by duration)."""
    con = sqlite3.connect(db)
    rows = con.execute("select name, dispatch_id, duration, counter_value from pmc_events "
                       "where counter_name = ? order by dispatch_id", (counter,)).fetchall()
    con.close()
    per, copy = {}, []
    for name, _, dur, val in rows:
        if "copyBuffer" in name and dur > 50_000:            # ns; the 1 GiB copies only
            copy.append(val)
        k = _short(name)
        if k in KERNEL_BYTES:
            per.setdefault(k, []).append(val)
    out = {k: sum(v[-steps:]) / len(v[-steps:]) for k, v in per.items()}
    return out, (sum(copy) / len(copy) if copy else None)


def measure_traffic(args):
    """HBM bytes per launch of each hot kernel from two rocprofv3 counter
    passes (a child process each), scaled by the same counter's reading of a
    1 GiB copy.  Returns (per-kernel bytes, scale factors) or (None, reason)."""
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None, "rocprofv3 not found"
    base = tempfile.mkdtemp(prefix="dsmc_pmc_")
    env = dict(os.environ, TMPDIR=base)
    vals, scale = {}, {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = os.path.join(base, counter)
        cmd = ["timeout", "-s", "KILL", str(args.pmc_timeout), prof, "--pmc", counter, "-d", d, "-o", "pass",
               "--", sys.executable, os.path.abspath(__file__), "--pmc-pass", "--workload", args.workload,
               "--lds-cap", str(args.lds_cap), "--pmc-steps", str(args.pmc_steps)]
        log(f"[bench] counter pass {counter}")
        r = subprocess.run(cmd, cwd=base, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        dbs = glob.glob(os.path.join(d, "**", "*.db"), recursive=True)
        if r.returncode != 0 or not dbs:
            log(r.stdout[-2000:])
            return None, f"{counter} pass exited {r.returncode}"
        per, copy_kib = _read_counter(dbs[0], counter, args.pmc_steps)
        if not copy_kib:
            return None, f"{counter}: calibration copy not found"
        scale[counter] = COPY_BYTES / (copy_kib * 1024.0)
        for k, v in per.items():
            vals.setdefault(k, {})[counter] = v * 1024.0 * scale[counter]
    shutil.rmtree(base, ignore_errors=True)
    traffic = {k: v.get("FETCH_SIZE", 0.0) + v.get("WRITE_SIZE", 0.0) for k, v in vals.items()}
    return {"bytes": traffic, "read": {k: v.get("FETCH_SIZE") for k, v in vals.items()},
            "written": {k: v.get("WRITE_SIZE") for k, v in vals.items()}}, scale


def bench_ensemble(args):
    from master_amd import dsmc
    p, worlds, g, n, tune = ensemble_system(args)
    g.set_option(dsmc.OPT_USE_GRAPH, 1)
    g.step(args.warmup)
    g.synchronize()
    t0 = time.perf_counter()
    g.step(args.steps)
    g.synchronize()
    el = time.perf_counter() - t0
    c = g.counters()
    value = n * args.steps / el
    out = {"metric": "particle-steps/s (move+sort+collide), 64-member ensemble", "value": round(value, 1),
           "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(1000 * el / args.steps, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"ensemble: 64 members, edges 32-128, {n} particles, {g.ncells} cells",
                      "particles": n, "cells": g.ncells, "members": len(worlds), "parallelism": "single GPU"},
           "autotune": tune, "errors": c["errors"]}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="roofline", choices=["roofline", "precision", "cpu", "ensemble"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch configuration")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--lds-cap", type=int, default=0, help="particles per cell held in LDS (0: automatic)")
    ap.add_argument("--statistics", type=int, default=0,
                    help="also run the statistics sampler every N steps (100 y bins); off by default")
    ap.add_argument("--pmc-pass", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--pmc-steps", type=int, default=3, help=argparse.SUPPRESS)
    ap.add_argument("--pmc-timeout", type=int, default=150, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the DSMC step here is a single-GPU path "
              "(multi-GPU sharding is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.pmc_pass:
        pmc_pass(args)
        return

    from master_amd import dsmc
    if args.workload == "ensemble":
        return bench_ensemble(args)
    p, w, g, n = make_system(args)
    if args.statistics:
        g.set_statistics(args.statistics, 100, 1_000_000)

    # per-kernel device times (HIP events on the handle's stream), eager launches
    g.set_option(dsmc.OPT_TIMING, 1)
    g.step(args.profile_steps)
    ktimes = g.kernel_times()
    g.set_option(dsmc.OPT_TIMING, 0)
    log(f"[bench] kernel ms/launch: {ktimes}")
    if args.statistics and ktimes.get("k_stats"):
        # a sample reads oy, vox, vx, vy, vz: 20 B per particle
        log(f"[bench] statistics sample: {ktimes['k_stats']:.3f} ms, "
            f"{20 * n / (ktimes['k_stats'] * 1e-3) / 1e9:.0f} GB/s of particle state read")

    g.set_option(dsmc.OPT_USE_GRAPH, 1)
    tune = None
    if not args.no_autotune:
        tune = g.autotune(8)
        log(f"[bench] autotune: {tune}")
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
    g.close()

    traffic, scale = (None, "skipped") if args.no_traffic else measure_traffic(args)
    if traffic is None:
        log(f"[bench] no traffic counters: {scale}")

    per_kernel = {}
    for k, t in ktimes.items():
        e = {"ms": round(t, 4)}
        if k in KERNEL_BYTES and t > 0:
            gbs = KERNEL_BYTES[k] * n / (t * 1e-3) / 1e9
            e.update(bytes_per_particle=KERNEL_BYTES[k], achieved_gbs=round(gbs, 1),
                     frac=round(gbs / HBM_PEAK_GBS, 4))
            if traffic and k in traffic["bytes"]:
                e["traffic_bytes"] = round(traffic["bytes"][k])
                e["traffic_per_particle"] = round(traffic["bytes"][k] / n, 2)
        per_kernel[k] = e
    dom = max((k for k in ktimes if k in KERNEL_BYTES), key=lambda k: ktimes[k])
    achieved = per_kernel[dom]["achieved_gbs"]
    step_gbs = STEP_BYTES * n / (el / args.steps) / 1e9
    roof = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": per_kernel[dom].get("traffic_bytes"),
            "algorithmic_bytes": KERNEL_BYTES[dom] * n,
            "launch_ms": per_kernel[dom]["ms"],
            "traffic_scale": scale if traffic else None,
            "kernels": per_kernel,
            "step": {"bytes_per_particle": STEP_BYTES, "achieved_gbs": round(step_gbs, 1),
                     "frac": round(step_gbs / HBM_PEAK_GBS, 4)}}
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
        "autotune": tune,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
