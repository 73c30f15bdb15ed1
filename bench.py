"""Benchmark: FLIP dam break on the two-level sparse grid, one MI355X.

Workload (default): 256^3 cells, dx = 1/256, closed solid walls, water block x<0.5, y<0.5, full z,
8 jittered particles per cell (33,554,431 particles), FP32, gravity -9.8, dt = 1/240, FLIP 0.95,
pressure PCG to 1e-5 relative residual, 8^3 blocks.  A "step" is one FLIPMethod::advanceOneStep
(binning, P2G, divergence, PCG, pressure apply, G2P/advection, removal) with the particles resident
in HBM.  The timed region runs `--steps` consecutive steps of one simulation started from rest.

Prints one JSON line (driver contract) with roofline and cpu_baseline objects.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/sec + p50/p99 step latency at 1e-5 rel divergence, vs plain GPU"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--block", type=int, default=8)
    ap.add_argument("--ppc", type=int, default=8)
    ap.add_argument("--solver", choices=["relative", "reference"], default="relative")
    ap.add_argument("--precond", choices=["none", "jacobi", "mg"], default="mg")
    ap.add_argument("--tol", type=float, default=1e-5)
    ap.add_argument("--cpu-baseline", choices=["on", "off"], default="on")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-kernel-timing", action="store_true")
    return ap.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (the multi-GPU "
              "sharding design is described in DESIGN.md)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    from vare_amd import _lib as L

    n = a.n
    dt = 1.0 / 240.0
    kw = dict(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / n,
              solver_mode=L.SOLVER_RELATIVE if a.solver == "relative" else L.SOLVER_REFERENCE,
              cg_rel_tol=a.tol, max_iterations=20000,
              preconditioner={"none": L.PRECOND_NONE, "jacobi": L.PRECOND_JACOBI, "mg": L.PRECOND_MG}[a.precond])
    t0 = time.time()
    pos = L.seed_box(n, n, n, 0.5 * n, 0.5 * n, float(n), a.ppc, 1)
    vel = np.zeros_like(pos)
    npart = pos.shape[0]
    print("bench: seeded %d particles in %.1fs on %s" % (npart, time.time() - t0, L.device_name()),
          file=sys.stderr, flush=True)
    B = a.block
    nblk = (n // B) ** 3
    h = L.FlipHandle(L.grid_desc(n, n, n, block_size=B), L.default_params(**kw),
                     particle_capacity=npart, block_capacity=nblk)

    # warm-up: run the first steps of the simulation, then restart from the initial state
    h.set_particles(pos, vel)
    for _ in range(a.warmup):
        rc = h.step_rc(dt)
        if rc not in (L.VARE_OK, L.VARE_ENOCONV):
            L.check(rc)
    h.set_particles(pos, vel)
    if not a.no_kernel_timing:
        h.set_kernel_timing(["spmv"])
        h.kernel_timing(reset=True)
    L.device_synchronize()

    step_ms = []
    iters = []
    relres = []
    nonconv = 0
    stats = None
    t_start = time.perf_counter()
    for s in range(a.steps):
        ts = time.perf_counter()
        rc = h.step_rc(dt)
        if rc == L.VARE_ENOCONV:
            nonconv += 1
        elif rc != L.VARE_OK:
            L.check(rc)
        step_ms.append((time.perf_counter() - ts) * 1e3)
        stats = h.stats()
        iters.append(stats["cg_iterations"])
        if stats["b_norm"] > 0:
            relres.append(stats["residual_norm"] / stats["b_norm"])
        if (s + 1) % 50 == 0:
            print("bench: step %d/%d %.2f ms cg=%d" % (s + 1, a.steps, step_ms[-1], iters[-1]),
                  file=sys.stderr, flush=True)
    L.device_synchronize()
    total_s = time.perf_counter() - t_start
    kt = h.kernel_timing(reset=True) if not a.no_kernel_timing else None

    value = npart * a.steps / total_s
    sm = np.array(step_ms)
    roofline = None
    if kt is not None and kt["spmv"][1] > 0:
        nc = stats["num_cells"]
        # algorithmic bytes of one A*p launch: read p (4 B) + stencil descriptor (2 B) + write Ap (4 B)
        # per active cell; neighbour values of p are re-reads of the same vector.
        bytes_per_launch = 10.0 * nc
        avg_ms = kt["spmv"][0] / kt["spmv"][1]
        achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
        roofline = {"kernel": "k_cg_spmv", "bound": "hbm", "achieved": round(achieved, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": None, "avg_launch_us": round(avg_ms * 1e3, 2), "launches": kt["spmv"][1],
                    "bytes_per_launch": bytes_per_launch}

    cpu = None
    if a.cpu_baseline == "on":
        cpu = cpu_baseline(h, n, B, kw, dt, a.cpu_steps)

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(total_s * 1e3 / a.steps, 3),
        "p50_ms": round(float(np.percentile(sm, 50)), 3),
        "p99_ms": round(float(np.percentile(sm, 99)), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "dam break %d^3, water x<0.5 y<0.5, %d ppc, closed walls, dt=1/240, FLIP 0.95, "
                               "%s PCG %s to %.0e relative, %d^3 blocks, %d steps from rest"
                               % (n, a.ppc, a.precond, a.solver, a.tol, B, a.steps),
                   "particles": npart, "cells": n ** 3, "active_blocks_last": stats["num_blocks"],
                   "fluid_cells_last": stats["num_fluid"]},
        "cg_iterations_mean": round(float(np.mean(iters)), 1),
        "cg_iterations_max": int(np.max(iters)),
        "rel_residual_max": float(np.max(relres)) if relres else 0.0,
        "nonconverged_steps": nonconv,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(h, n, B, kw, dt, steps):
    """Time the oracle (the FP32 restatement of the reference step, one host thread) on the same
    workload: `steps` step(s) of the 256^3 dam break taken from the GPU's current particle state,
    with the reference's own solver (unpreconditioned CG, |r| < 1e-3 or 202 products)."""
    from oracle import oracle as O
    d = h.get_particles()
    ref_kw = dict(kw, solver_mode=0, preconditioner=0)
    o = O.Oracle((n, n, n), B=B, prm=O.params(**ref_kw))
    o.set_particles(d["pos"], d["vel"])
    npart = d["pos"].shape[0]
    t = time.perf_counter()
    for _ in range(steps):
        o.step(dt)
    el = time.perf_counter() - t
    st = o.stats()
    o.close()
    return {"value": round(npart * steps / el, 1), "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": "%d step(s) of the %d^3 dam break (%d particles) from the GPU run's final state, "
                      "reference CG rule (%d iterations), %.1f s" % (steps, n, npart, st["cg_iters"], el)}


if __name__ == "__main__":
    main()
