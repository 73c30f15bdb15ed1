"""Benchmark of the FLIP step on the two-level sparse grid, one MI355X.

Default (config 2 of the benchmark set): FLIP dam break, 256^3 cells, dx = 1/256, closed solid walls,
water block x<0.5, y<0.5, full z, 8 jittered particles per cell (33,554,431 particles), FP32,
gravity -9.8, dt = 1/240, FLIP ratio 0.95, pressure PCG to 1e-5 relative residual (multigrid
preconditioner), 8^3 blocks, 200 steps from rest.  A "step" is one FLIPMethod::advanceOneStep
(binning, P2G, divergence, PCG, pressure apply, G2P/advection, removal) with the particles resident
in HBM.

Other configurations (--config):
  1  the same dam break at 64^3, water x<0.4, y<0.6, 100 steps; the CPU baseline runs all 100 steps;
  3  mixed precision at 512^3: a pool of depth 0.3 plus a sphere of water (radius 0.2, centre
     (0.5, 0.7, 0.5)), 8 particles per cell, velocities stored in FP16, FP32 PCG with FP64 dot products
     and FP64 iterative refinement to 1e-8 relative, 10 steps;
  4  served batch: independent tanks (edges log-uniform 32..128, 4..8 particles per cell, water
     fractions 0.2..0.7, 10 steps each) arriving as a Poisson process, continuously batched by the
     library's server (every resident tank stepped in one round of shared kernel launches), compared
     with the same requests solved one at a time ("plain GPU"), with the same arrivals stepped one
     stream per tank without shared launches, and with all requests waiting at t=0 (capacity);
  5  sparse wave on an obstacle: 1024x512x1024 cells (dx = 1/512, world 2x1x2), a wall of water
     x<0.25, y<0.3, full z moving +x at 4 m/s, a static box obstacle x in [0.5,0.7], y in [0,0.5],
     z in [0.6,1.4] given as a signed distance field, FP16 velocities, FP64 refinement to 1e-6, 100
     steps at dt = 1/1000 (4 m/s is 8.5 cells per step at 1/240; 1/1000 keeps it near 2).
The timed region is K consecutive steps of the configuration's run (--steps, default the whole run).
A shorter region is centred on the run: the simulation first advances untimed to step (run - K) / 2
(--start-step), so a 20-step sample of the 200-step dam break times steps 90..109.

"vs plain GPU" (configurations 1 and 2, --plain): the same timed steps are run again with the plain
FP32 GPU step -- FP32 multigrid PCG to the target on its recursive residual, no FP64 refinement --
and both solvers take one step from the same final particle state, which reports each one's FP64
relative residual and the divergence ratio D = |div_c v_post| / |div_c v_pre| (the reference's
central-difference divergence; VARE_FLAG_REPORT_DIVERGENCE).  `vs_plain_gpu` is the ratio of the two
rates; `vs_baseline` stays null because the reference publishes no number for this metric.

The CPU baseline is the OpenMP build of the oracle (the FP32 restatement of the reference step) on the
host of the machine the GPU sits in: all threads of the host's share and one thread under the reference's own CG rule, and
all threads under Jacobi-PCG to 1e-5 relative.  For configs 3 and 5 it times steps of the same scene
at a quarter of the resolution per axis (a bounded sample; the rate is per particle-step).

Prints one JSON line with `roofline` and `cpu_baseline` objects.
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
# HBM bytes per processed cell of one launch, measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
# on this benchmark (20 steps) and calibrated against plain 1 GiB copies (FETCH_SIZE x 2.0,
# WRITE_SIZE x 1.0); profiles/hbm_traffic_dam_break_256.md.
PSPMV_TRAFFIC_PER_CELL = 17.10    # k_pupdate_spmv: (2 x 22302.6 + 32056.6) KiB per 4.590 M cells
UPD_TRAFFIC_PER_CELL = 17.60     # k_update_smooth2: (2 x 23206.2 + 32417.6) KiB per 4.587 M cells
MG0_TRAFFIC_PER_CELL = 12.51      # k_mg0_prolong_smooth: (2 x 20663.5 + 14684.8) KiB per 4.585 M cells
TIMING_STRIDE = 4                  # the dominant kernel is timed on every 4th step of the timed region
OTHER_TIMING_STEPS = 10            # steps after the timed region that time the secondary kernels
DT = 1.0 / 240.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3, 4, 5])
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--block", type=int, default=8)
    ap.add_argument("--ppc", type=int, default=8)
    ap.add_argument("--solver", choices=["relative", "reference"], default="relative")
    ap.add_argument("--precond", choices=["none", "jacobi", "mg"], default="mg")
    ap.add_argument("--tol", type=float, default=None,
                    help="target |b-Ax|/|b|, met in FP64 through iterative refinement (see --no-refine); "
                         "default 1e-5 (1e-8 for config 3, 1e-6 for config 5)")
    ap.add_argument("--inner-tol", type=float, default=1e-3,
                    help="relative tolerance of each FP32 PCG solve inside the refinement")
    ap.add_argument("--no-refine", action="store_true",
                    help="plain FP32 PCG to --tol (recursive residual; its FP64 residual is reported)")
    ap.add_argument("--mg-sweeps", type=int, default=2)
    ap.add_argument("--mg-coarse-sweeps", type=int, default=20, help="Jacobi sweeps on the coarsest multigrid grid")
    ap.add_argument("--mg-omega-pair", type=float, nargs=2, default=None, metavar=("A", "B"),
                    help="weights of the first and second sweep of each smoothing pair (default: the "
                         "library's Chebyshev pair; 0 0: weighted Jacobi, 0.8 for both)")
    ap.add_argument("--cpu-baseline", choices=["on", "off"], default="on")
    ap.add_argument("--cpu-steps", type=int, default=None)
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--requests", type=int, default=256)
    ap.add_argument("--slots", type=int, default=32)
    ap.add_argument("--start-step", type=int, default=None,
                    help="untimed steps of the simulation before the warm-up and the timed region; default: "
                         "centre a shorter timed region on the configuration's full run")
    ap.add_argument("--plain", choices=["on", "off"], default=None,
                    help="also time the plain FP32 GPU step (FP32 PCG to --tol, no refinement) on the same "
                         "steps; default on for configurations 1 and 2")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy a buffer of MIB mebibytes once on the GPU (torch) and exit: the reference "
                         "kernel that scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters to bytes")
    ap.add_argument("--cpu-threads", type=int, default=None,
                    help="threads of the multi-core CPU baseline (default: OMP_NUM_THREADS, else the host's "
                         "cores up to 16)")
    return ap.parse_args()


def solver_kw(a, L, n):
    """Step parameters for a scene with voxel size 1/n (world units)."""
    refine = a.solver == "relative" and not a.no_refine
    return dict(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / n,
                solver_mode=L.SOLVER_RELATIVE if a.solver == "relative" else L.SOLVER_REFERENCE,
                cg_rel_tol=a.inner_tol if refine else a.tol, refine_tol=a.tol if refine else 0.0,
                refine_max=16, max_iterations=20000,
                preconditioner={"none": L.PRECOND_NONE, "jacobi": L.PRECOND_JACOBI, "mg": L.PRECOND_MG}[a.precond],
                mg_sweeps=a.mg_sweeps, mg_coarse_sweeps=a.mg_coarse_sweeps, **({} if a.mg_omega_pair is None else
                                          dict(mg_omega_a=a.mg_omega_pair[0], mg_omega_b=a.mg_omega_pair[1])))


def solver_desc(a, cfg=2, kw=None):
    if a.solver != "relative":
        return "reference CG rule"
    if (kw is not None and kw["refine_tol"] == 0.0) or (kw is None and a.no_refine):
        return "FP32 %s-PCG to %.0e relative (recursive residual)" % (a.precond, a.tol)
    return ("FP32 %s-PCG (%.0e per solve) + FP64 iterative refinement to %.0e relative FP64 residual"
            % (a.precond, a.inner_tol, a.tol))


def smoothing_desc(a):
    if a.precond != "mg" or a.mg_sweeps != 2:
        return ""
    if a.mg_omega_pair is None:
        return ", Chebyshev-weighted Jacobi smoothing pairs"
    if a.mg_omega_pair[0] > 0.0 and a.mg_omega_pair[1] > 0.0:
        return ", smoothing pairs weighted %g and %g" % tuple(a.mg_omega_pair)
    return ", Jacobi smoothing (0.8)"


def plain_kw(a, L, n):
    """The plain GPU step: FP32 PCG to --tol on its recursive residual, no refinement."""
    kw = solver_kw(a, L, n)
    kw.update(cg_rel_tol=a.tol, refine_tol=0.0)
    return kw


def plain_comparison(a, L, cfg, steps, kw, state, h, value):
    """The plain GPU step (FP32 PCG, no refinement) on the same timed steps, at two tolerances: the
    target on its recursive residual, and the loosest recursive tolerance whose pressure meets the
    target in FP64 on the final state of the timed run (equal post-projection divergence), found by
    halving from the target.  Also both solvers' FP64 residual and divergence ratio D on that state."""
    res_n = scene(cfg)["res"]
    d = h.get_particles()
    final = (d["pos"], d["vel"])
    tol = a.tol
    calib = []
    equal_kw = None
    for _ in range(8):
        pkw = dict(plain_kw(a, L, res_n), cg_rel_tol=tol)
        r = same_state_divergence(a, L, cfg, final, {"plain": pkw})["plain"]
        calib.append({"cg_rel_tol": tol, "fp64_rel_residual": r["fp64_rel_residual"], "cg_iterations": r["cg_iterations"]})
        if r["fp64_rel_residual"] <= a.tol:
            equal_kw = pkw
            break
        tol *= 0.5
    out = {"plain_calibration": calib}
    lines = {"plain_gpu": plain_kw(a, L, res_n)}
    if equal_kw is not None:
        lines["plain_gpu_equal_divergence"] = equal_kw
    for name, pkw in lines.items():
        hp, _, pres, _ = run_scene(a, L, cfg, steps, kw=pkw, timing=False, state=state)
        hp.close()
        out[name] = {"value": pres["value"], "ms_per_step": pres["ms_per_step"], "p50_ms": pres["p50_ms"],
                     "p99_ms": pres["p99_ms"], "cg_iterations_mean": pres["cg_iterations_mean"],
                     "fp64_rel_residual_max": pres["fp64_rel_residual_max"],
                     "fp64_rel_residual_mean": pres["fp64_rel_residual_mean"],
                     "solver": "FP32 %s-PCG to %.2g relative (recursive residual), no refinement"
                               % (a.precond, pkw["cg_rel_tol"])}
        print("bench: %s %.3g particle-steps/s, FP64 residual max %.2e" % (
            name, pres["value"], pres["fp64_rel_residual_max"]), file=sys.stderr, flush=True)
    if equal_kw is None:
        out["plain_gpu_equal_divergence"] = None
        out["note"] = ("no recursive tolerance down to %.2g brought the plain FP32 step's FP64 residual to %.0e"
                       % (calib[-1]["cg_rel_tol"], a.tol))
    out["vs_plain_gpu"] = round(value / max(out["plain_gpu"]["value"], 1e-9), 3)
    if equal_kw is not None:
        out["vs_plain_gpu_equal_divergence"] = round(value / max(out["plain_gpu_equal_divergence"]["value"], 1e-9), 3)
    kws = {"mixed": kw, "plain": lines["plain_gpu"]}
    if equal_kw is not None:
        kws["plain_equal_divergence"] = equal_kw
    out["divergence_same_state"] = same_state_divergence(a, L, cfg, final, kws)
    return out


def same_state_divergence(a, L, cfg, state, kws):
    """One step from the same particle state per solver setting: the FP64 relative residual of the
    pressure and D = |div_c v_post| / |div_c v_pre| (the reference's central-difference divergence)."""
    sc = scene(cfg)
    out = {}
    for name, kw in kws.items():
        h = scene_handle(L, sc, dict(kw, flags=L.FLAG_REPORT_DIVERGENCE), a.block, state[0].shape[0])
        h.set_particles(state[0], state[1])
        step_ok(L, h, sc["dt"])
        st = h.stats()
        out[name] = {"fp64_rel_residual": st["true_rel_residual"], "D": st["div_ratio"],
                     "cg_iterations": st["cg_iterations"], "refine_steps": st["refine_steps"]}
        h.close()
    return out


def scene(cfg, scale=1):
    """Scene of a benchmark configuration; `scale` divides the resolution per axis (CPU sample)."""
    if cfg in (1, 2):
        n, wx, wy = (64, 0.4, 0.6) if cfg == 1 else (256, 0.5, 0.5)
        n //= scale
        return dict(dims=(n, n, n), world=(1.0, 1.0, 1.0), res=n, kind="box", water=(wx, wy, 1.0), vel=(0, 0, 0),
                    obstacle=None, fp16=False, steps=100 if cfg == 1 else 200, dt=DT,
                    desc="dam break %d^3, water x<%.1f y<%.1f, closed walls" % (n, wx, wy))
    if cfg == 3:
        n = 512 // scale
        return dict(dims=(n, n, n), world=(1.0, 1.0, 1.0), res=n, kind="pool_sphere",
                    pool=(0.3, (0.5, 0.7, 0.5), 0.2), vel=(0, 0, 0), obstacle=None, fp16=True, steps=10, dt=DT,
                    desc="pool depth 0.3 + sphere r=0.2 at (0.5,0.7,0.5), %d^3, closed walls" % n)
    if cfg == 5:
        r = 512 // scale
        return dict(dims=(2 * r, r, 2 * r), world=(2.0, 1.0, 2.0), res=r, kind="box", water=(0.25, 0.3, 2.0),
                    vel=(4.0, 0.0, 0.0), obstacle=((0.5, 0.0, 0.6), (0.7, 0.5, 1.4)), fp16=True, steps=100,
                    dt=1.0 / 1000.0,
                    desc="wave x<0.25 y<0.3 at +4 m/s on a static box [0.5,0.7]x[0,0.5]x[0.6,1.4], %dx%dx%d (dx=1/%d)"
                         % (2 * r, r, 2 * r, r))
    raise ValueError(cfg)


def seed_scene(L, sc, ppc):
    """Particles (voxel space) and their initial velocities (voxel units per second)."""
    nx, ny, nz = sc["dims"]
    r = sc["res"]
    if sc["kind"] == "box":
        wx, wy, wz = sc["water"]
        pos = L.seed_box(nx, ny, nz, wx * r, wy * r, wz * r, ppc, 1)
    else:
        depth, c, rad = sc["pool"]
        pos = L.seed_pool_sphere(nx, ny, nz, depth * r, [x * r for x in c], rad * r, ppc, 1)
    vel = np.empty_like(pos)
    vel[:] = np.array(sc["vel"], np.float32) * np.float32(r)
    return pos, vel


def scene_handle(L, sc, kw, B, npart):
    nx, ny, nz = sc["dims"]
    kw = dict(kw, velocity_fp16=1 if sc["fp16"] else 0)
    h = L.FlipHandle(L.grid_desc(nx, ny, nz, (0.0, 0.0, 0.0), sc["world"], block_size=B), L.default_params(**kw),
                     particle_capacity=npart, block_capacity=(nx // B) * (ny // B) * (nz // B))
    if sc["obstacle"] is not None:
        h.set_collision_box(sc["obstacle"][0], sc["obstacle"][1], (0.0, 0.0, 0.0), 3.0)
    return h


def step_ok(L, h, dt):
    rc = h.step_rc(dt)
    if rc not in (L.VARE_OK, L.VARE_ENOCONV):
        L.check(rc)
    return rc


def start_step(a, cfg, steps):
    """First timed step of the simulation: a timed region shorter than the configuration's run is
    centred on that run, so its steps are representative of the whole run."""
    if a.start_step is not None:
        return max(0, a.start_step)
    full = scene(cfg)["steps"]
    return max(0, (full - steps) // 2)


def run_scene(a, L, cfg, steps, kw=None, timing=True, state=None):
    """Time `steps` steps of configuration `cfg`.  `state` = (pos, vel, start) reuses a seeded state
    (pos, vel after `start` untimed steps) instead of seeding and pre-rolling."""
    sc = scene(cfg)
    kw = kw or solver_kw(a, L, sc["res"])
    B = a.block
    if state is None:
        t0 = time.time()
        pos, vel = seed_scene(L, sc, a.ppc)
        print("bench: seeded %d particles in %.1fs on %s" % (pos.shape[0], time.time() - t0, L.device_name()),
              file=sys.stderr, flush=True)
        start = start_step(a, cfg, steps)
        npart = pos.shape[0]
        h = scene_handle(L, sc, kw, B, npart)
        if start:
            # advance the simulation to the first timed step (untimed), from rest
            h.set_particles(pos, vel)
            for _ in range(start):
                step_ok(L, h, sc["dt"])
            d = h.get_particles()
            pos, vel = d["pos"], d["vel"]
            print("bench: advanced %d untimed steps" % start, file=sys.stderr, flush=True)
    else:
        pos, vel, start = state
        npart = pos.shape[0]
        h = scene_handle(L, sc, kw, B, npart)
    state = (pos, vel, start)

    # warm-up: run the next steps, then restart from the first timed step's state
    h.set_particles(pos, vel)
    for _ in range(a.warmup):
        step_ok(L, h, sc["dt"])
    h.set_particles(pos, vel)
    timing = timing and not a.no_kernel_timing
    # HIP events around a launch cost it about 5 us, so inside the timed region only the dominant
    # kernel is timed, on every TIMING_STRIDE-th step (a uniform sample of its launches over the
    # region); the other kernels' lines come from a short run after it
    if timing:
        h.kernel_timing(reset=True)
    L.device_synchronize()

    step_ms, iters, relres, truerel, refines, nps = [], [], [], [], [], []
    fluid_launch_cells = 0                  # sum over the timed update launches of the step's fluid cells
    nonconv = 0
    stats = None
    t_start = time.perf_counter()
    for s in range(steps):
        if timing:
            h.set_kernel_timing(["update"] if s % TIMING_STRIDE == 0 else [])
        ts = time.perf_counter()
        nps.append(h.num_particles())
        rc = h.step_rc(sc["dt"])
        if rc == L.VARE_ENOCONV:
            nonconv += 1
        elif rc != L.VARE_OK:
            L.check(rc)
        step_ms.append((time.perf_counter() - ts) * 1e3)
        stats = h.stats()
        iters.append(stats["cg_iterations"])
        if timing and s % TIMING_STRIDE == 0:
            # one update launch ran per PCG iteration of the step
            fluid_launch_cells += stats["cg_iterations"] * stats["num_fluid"]
        if stats["b_norm"] > 0:
            relres.append(stats["residual_norm"] / stats["b_norm"])
            truerel.append(stats["true_rel_residual"])
            refines.append(stats["refine_steps"])
        if (s + 1) % 10 == 0:
            print("bench: step %d/%d %.2f ms cg=%d" % (s + 1, steps, step_ms[-1], iters[-1]),
                  file=sys.stderr, flush=True)
    L.device_synchronize()
    total_s = time.perf_counter() - t_start
    kt = h.kernel_timing(reset=True) if timing else None
    if kt is not None:
        # untimed: OTHER_TIMING_STEPS more steps with events on the other reported kernels
        h.set_kernel_timing(["spmv", "mg0", "p2g"])
        h.kernel_timing(reset=True)
        for _ in range(OTHER_TIMING_STEPS):
            rc = h.step_rc(sc["dt"])
            if rc not in (L.VARE_OK, L.VARE_ENOCONV):
                L.check(rc)
        L.device_synchronize()
        ko = h.kernel_timing(reset=True)
        h.set_kernel_timing([])
        kt["spmv"], kt["mg0"], kt["p2g"] = ko["spmv"], ko["mg0"], ko["p2g"]
        sampled = "every %d-th step of the timed region" % TIMING_STRIDE
        after = "%d untimed steps after the timed region" % OTHER_TIMING_STEPS

    roofline = None
    if kt is not None and kt["mg0"][1] > 0 and kt["spmv"][1] > 0 and kt["update"][1] > 0:
        def line(kernel, cls, bytes_per_cell, traffic_per_cell):
            ms, launches, cells = kt[cls]
            bytes_total = bytes_per_cell * cells
            achieved = bytes_total / (ms * 1e-3) / 1e9
            return {"kernel": kernel, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": round(traffic_per_cell * cells / launches, 1) if traffic_per_cell else None,
                    "avg_launch_us": round(ms * 1e3 / launches, 2), "launches": launches,
                    "bytes_per_launch": round(bytes_total / launches, 1)}
        # dominant kernel (the most time per step): the CG's r update fused with the V-cycle's first
        # level-0 pass (x += alpha p runs in the same iteration's bottom V-cycle launch).  Algorithmic
        # bytes per solver cell: read r, Ap 8 + descriptor 2, write r, z 8.  Neighbour values across a
        # block face are re-reads.
        roofline = line("k_update_smooth2", "update", 18.0, UPD_TRAFFIC_PER_CELL)
        roofline["timed_on"] = sampled
        # the same launches counted on fluid cells only (the solver runs over every cell of a block that
        # holds a fluid cell; the others are no-ops of the stencil descriptor)
        if kt["update"][2] > 0 and fluid_launch_cells > 0:
            share = fluid_launch_cells / kt["update"][2]
            roofline["fluid_cell_share"] = round(share, 4)
            roofline["frac_on_fluid_cells"] = round(roofline["frac"] * share, 4)
        roofline["others"] = [
            # the level-0 prolongation + Jacobi pass: read z 4 + coarse z 0.5 + r 4 + descriptor 2, write z 4
            line("k_mg0_prolong_smooth", "mg0", 14.5, MG0_TRAFFIC_PER_CELL),
            # the direction update fused with A*p: read old p 4 + z 4 + descriptor 2, write p 4 + Ap 4
            line("k_pupdate_spmv", "spmv", 18.0, PSPMV_TRAFFIC_PER_CELL)]
        if kt["p2g"][1] > 0:
            # both P2G passes between one pair of events; the library counts their algorithmic bytes
            # (24 per particle + 33 per active cell)
            ms, launches, nbytes = kt["p2g"]
            achieved = nbytes / (ms * 1e-3) / 1e9
            roofline["others"].append({
                "kernel": "k_p2g_own + k_p2g_faces", "bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "avg_launch_us": round(ms * 1e3 / launches, 2), "launches": launches,
                "bytes_per_launch": round(nbytes / launches, 1)})
        for o in roofline["others"]:
            o["timed_on"] = after
    sm = np.array(step_ms)
    nx, ny, nz = sc["dims"]
    res = {
        "value": round(float(np.sum(nps)) / total_s, 1),
        "ms_per_step": round(total_s * 1e3 / steps, 3),
        "p50_ms": round(float(np.percentile(sm, 50)), 3),
        "p99_ms": round(float(np.percentile(sm, 99)), 3),
        "config": {"workload": "%s, %d ppc, dt=1/%d, FLIP 0.95, %s%s, %d^3 blocks, steps %d..%d of the %d-step run"
                               % (sc["desc"], a.ppc, round(1.0 / sc["dt"]), solver_desc(a, cfg, kw) + smoothing_desc(a),
                                  ", FP16 particle velocities" if sc["fp16"] else "", B, start, start + steps - 1,
                                  max(sc["steps"], start + steps)),
                   "particles": npart, "particles_last": int(nps[-1]), "cells": nx * ny * nz,
                   "first_timed_step": start,
                   "active_blocks_last": stats["num_blocks"], "frame_blocks": (nx // B) * (ny // B) * (nz // B),
                   "fluid_cells_last": stats["num_fluid"]},
        "cg_iterations_mean": round(float(np.mean(iters)), 1),
        "cg_iterations_max": int(np.max(iters)),
        "rel_residual_max": float(np.max(relres)) if relres else 0.0,
        "fp64_rel_residual_max": float(np.max(truerel)) if truerel else 0.0,
        "refine_steps_mean": round(float(np.mean(refines)), 2) if refines else 0.0,
        "nonconverged_steps": nonconv,
        "fp64_rel_residual_mean": float(np.mean(truerel)) if truerel else 0.0,
        "roofline": roofline,
    }
    return h, kw, res, state


def host_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return os.cpu_count() or 1, model


def cpu_threads(a):
    if a.cpu_threads:
        return a.cpu_threads
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return max(1, min(16, os.cpu_count() or 1))


def cpu_baseline(a, L, cfg, h, kw):
    """Time the oracle's OpenMP build (the FP32 restatement of the reference step; loops over
    particles and cells in parallel, P2G by z-slabs two apart) on the GPU machine's host: with every
    thread of the host's share and with one thread, under the reference's own solver rule
    (unpreconditioned CG, |r| < 1e-3 or 202 products), and with every thread also under the configured
    rule (Jacobi-PCG to 1e-5 relative).  A bounded sample of the workload: config 1 all its steps from
    rest with all threads and its first 10 with one; config 2 one step from the GPU run's final particle
    state; configs 3 and 5 two steps of the same scene at a quarter of the resolution per axis, after
    four untimed ones."""
    from oracle import oracle as O
    nproc, model = host_info()
    nth = cpu_threads(a)
    if cfg == 1:
        sc = scene(1)
        p, v = seed_scene(L, sc, a.ppc)
        steps, start = a.cpu_steps or sc["steps"], "rest"
    elif cfg == 2:
        sc = scene(2)
        d = h.get_particles()
        p, v = d["pos"], d["vel"]
        steps, start = a.cpu_steps or 1, "the GPU run's final state"
    else:
        sc = scene(cfg, scale=4)
        p, v = seed_scene(L, sc, a.ppc)
        steps, start = a.cpu_steps or 2, "the state after 4 untimed steps"
    base_kw = dict(ext_force=kw["ext_force"], wall_level=kw["wall_level"], voxel_size=1.0 / sc["res"])

    def timed(threads, nsteps, **solver):
        O.set_threads(threads)
        o = O.Oracle(sc["dims"], (0.0, 0.0, 0.0), sc["world"], B=a.block, prm=O.params(**base_kw, **solver),
                     lib_path=O.OMP_LIB_PATH)
        if sc["obstacle"] is not None:
            o.set_collision_box(sc["obstacle"][0], sc["obstacle"][1], (0.0, 0.0, 0.0), 3.0)
        o.set_particles(p, v)
        if cfg in (3, 5):
            for _ in range(4):          # leave the trivial first steps from rest out of the sample
                o.step(sc["dt"])
        its, counts = [], []
        t = time.perf_counter()
        for _ in range(nsteps):
            counts.append(o.num_particles())
            its.append(o.step(sc["dt"]))
        el = time.perf_counter() - t
        o.close()
        return {"value": round(float(np.sum(counts)) / el, 1), "steps": nsteps, "seconds": round(el, 2),
                "cg_iterations_mean": round(float(np.mean(its)), 1)}

    ref_all = timed(nth, steps)
    ref_one = timed(1, min(steps, 10))
    rel_all = timed(nth, min(steps, 10), solver_mode=1, cg_rel_tol=1e-5, max_iterations=20000, preconditioner=1)
    print("bench: CPU baseline %d threads %.3g, 1 thread %.3g, relative 1e-5 %.3g particle-steps/s" % (
        nth, ref_all["value"], ref_one["value"], rel_all["value"]), file=sys.stderr, flush=True)
    return {"value": ref_all["value"], "unit": "particle-steps/s", "cores": nth, "kind": "port",
            "sample": "%d step(s) of %s (%d particles, %d ppc) from %s, reference CG rule (mean %.0f iterations), "
                      "%.1f s on %d threads of the OpenMP build of the oracle"
                      % (steps, sc["desc"], p.shape[0], a.ppc, start, ref_all["cg_iterations_mean"],
                         ref_all["seconds"], nth),
            "host_nproc": nproc, "host_cpu": model,
            "one_core": dict(ref_one, rule="reference CG rule"),
            "relative_1e-5": dict(rel_all, cores=nth, rule="Jacobi-PCG to 1e-5 relative (FP32)")}


def run_served(a, L):
    """Config 4: a Poisson stream of independent tank solves through the continuous-batching server,
    against the same requests solved one at a time."""
    rng = np.random.default_rng(2024)
    nreq = a.requests
    reqs = []
    for k in range(nreq):
        edge = int(8 * round(np.exp(rng.uniform(np.log(32), np.log(128))) / 8))
        ppc = int(rng.integers(4, 9))
        fx, fy = rng.uniform(0.2, 0.7, 2)
        pos = L.seed_box(edge, edge, edge, fx * edge, fy * edge, float(edge), ppc, 100 + k)
        kw = solver_kw(a, L, edge)
        reqs.append((L.grid_desc(edge, edge, edge, block_size=a.block), L.default_params(**kw), pos,
                     np.zeros_like(pos)))
    total_particles = sum(r[2].shape[0] for r in reqs)
    print("bench: %d requests, %d particles in total" % (nreq, total_particles), file=sys.stderr, flush=True)
    budget = 200 << 30
    # isolated solve time (sequential server, no arrival gaps) sets the Poisson rate
    # (a first pass loads the kernels and creates the device states; the second one is timed)
    for timed in (False, True):
        warm = L.Server(budget, 1, batching=False)
        for g, p, pos, vel in reqs[:8]:
            warm.submit(g, p, pos, vel, steps=10, dt=DT)
        t = time.perf_counter()
        warm.run()
        iso = (time.perf_counter() - t) / 8
        warm.close()
    mean_gap = 0.5 * iso
    arrivals = np.cumsum(rng.exponential(mean_gap, nreq))
    arrivals -= arrivals[0]

    def serve(slots, with_arrivals, batching=True):
        srv = L.Server(budget, slots, batching=batching)
        for k, (g, p, pos, vel) in enumerate(reqs):
            srv.submit(g, p, pos, vel, steps=10, dt=DT, arrival_s=float(arrivals[k]) if with_arrivals else 0.0)
        srv.run()
        return srv.stats()

    def brief(st, note):
        return {"value": round(st["particle_steps_per_s"], 1), "p50_ms": round(st["step_p50_ms"], 3),
                "p99_ms": round(st["step_p99_ms"], 3), "wall_s": round(st["wall_s"], 3),
                "max_batch": int(st["max_batch"]), "note": note}

    plain = serve(1, False, batching=False)   # one tank at a time, back to back, vare_flip_step
    L.batch_stats(reset=True)
    served = serve(a.slots, True)             # Poisson arrivals, batched rounds
    bstats = L.batch_stats()
    streams = serve(a.slots, True, batching=False)   # Poisson arrivals, one stream per resident tank
    L.batch_stats(reset=True)
    saturated = serve(a.slots, False)         # every request waiting at t=0: the server's capacity
    sstats = L.batch_stats()
    streams_sat = serve(a.slots, False, batching=False)   # all waiting at t=0, one stream per tank
    return {
        "value": round(served["particle_steps_per_s"], 1),
        "ms_per_step": round(served["step_mean_ms"], 3),
        "p50_ms": round(served["step_p50_ms"], 3),
        "p99_ms": round(served["step_p99_ms"], 3),
        "steps": int(served["steps"]),
        "config": {"workload": "served batch: %d tank solves, edges log-uniform 32..128 (multiple of 8), 4..8 ppc, "
                               "water fractions U(0.2,0.7), 10 steps each at dt=1/240, %s, Poisson "
                               "arrivals (mean gap %.1f ms = 0.5 x isolated solve), %d slots, 200 GB budget, "
                               "resident tanks stepped in batched rounds (shared kernel launches)"
                               % (nreq, solver_desc(a, 4), 1e3 * mean_gap, a.slots),
                   "particles_total": total_particles, "max_resident": served["max_resident"],
                   "handles_created": served["handles_created"], "handles_reused": served["handles_reused"],
                   "peak_reserved_gb": round(served["peak_bytes"] / 2 ** 30, 2), "wall_s": round(served["wall_s"], 3)},
        "batching": {"rounds": int(served["rounds"]), "max_batch": int(served["max_batch"]),
                     "merged_launches": bstats["launches"], "tank_launches": bstats["tank_launches"],
                     "sync_points": bstats["flushes"], "flush_host_s": round(bstats["flush_ms"] / 1e3, 3),
                     "flush_wait_s": round(bstats["wait_ms"] / 1e3, 3),
                     "saturated": {"rounds": int(saturated["rounds"]), "merged_launches": sstats["launches"],
                                   "tank_launches": sstats["tank_launches"], "sync_points": sstats["flushes"],
                                   "flush_host_s": round(sstats["flush_ms"] / 1e3, 3),
                                   "flush_wait_s": round(sstats["wait_ms"] / 1e3, 3),
                                   "between_flushes_s": round(sstats["gap_ms"] / 1e3, 3),
                                   "in_batched_steps_s": round(sstats["step_ms"] / 1e3, 3)}},
        "plain_gpu": brief(plain, "same requests, one tank at a time (vare_flip_step)"),
        "streams": brief(streams, "same arrivals, each resident tank stepped on its own stream, no shared launches"),
        "saturated": brief(saturated, "batched rounds with all %d requests waiting at t=0 (capacity)" % nreq),
        "streams_saturated": brief(streams_sat, "one stream per tank with all %d requests waiting at t=0" % nreq),
        "vs_plain_gpu": round(served["particle_steps_per_s"] / max(plain["particle_steps_per_s"], 1e-9), 3),
        "vs_streams": round(served["particle_steps_per_s"] / max(streams["particle_steps_per_s"], 1e-9), 3),
        "saturated_vs_plain_gpu": round(saturated["particle_steps_per_s"] / max(plain["particle_steps_per_s"], 1e-9), 3),
        "saturated_vs_streams": round(saturated["particle_steps_per_s"] / max(streams_sat["particle_steps_per_s"], 1e-9), 3),
        "roofline": None,
    }


def copy_calibration(mib):
    import torch
    n = (mib << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 4}), flush=True)


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (the step is not sharded "
              "across GPUs)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    from vare_amd import _lib as L

    if a.tol is None:
        a.tol = {3: 1e-8, 5: 1e-6}.get(a.config, 1e-5)
    if a.config == 4:
        res = run_served(a, L)
        steps, cpu = res.pop("steps"), None
    else:
        steps = a.steps if a.steps is not None else scene(a.config)["steps"]
        h, kw, res, state = run_scene(a, L, a.config, steps)
        plain_on = (a.plain or ("on" if a.config in (1, 2) else "off")) == "on"
        if plain_on and a.solver == "relative" and kw["refine_tol"] > 0.0:
            res.update(plain_comparison(a, L, a.config, steps, kw, state, h, res["value"]))
        cpu = cpu_baseline(a, L, a.config, h, kw) if a.cpu_baseline == "on" else None
    out = {
        "metric": METRIC,
        "value": res.pop("value"),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": res.pop("ms_per_step"),
        "p50_ms": res.pop("p50_ms"),
        "p99_ms": res.pop("p99_ms"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": res.pop("config"),
    }
    out.update(res)
    out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
