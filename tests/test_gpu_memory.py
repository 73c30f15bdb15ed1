"""GPU tests of the handle's device-memory accounting and of its recovery from a failed allocation.

* vare_flip_plan_bytes (what the server charges a request) covers the bytes a real handle reaches
  for create, upload, steps and download, for several grids and parameter sets;
* an allocation that fails (injected through vare_debug_inject_alloc_failure) leaves the handle
  with its previous capacity and particles: the call returns VARE_ENOMEM, the next call works and
  gives exactly what a fresh handle gives;
* kernel timing counts only the solver iterations that ran (launches enqueued past convergence
  return at the device-side done flag and are dropped).
"""
import numpy as np
import pytest

from vare_amd import _lib as L

pytestmark = pytest.mark.gpu

DT = 1.0 / 240.0


def scene(n, fx=0.5, fy=0.5, ppc=4, seed=1):
    pos = L.seed_box(n, n, n, fx * n, fy * n, float(n), ppc, seed)
    rng = np.random.default_rng(seed)
    return pos, (rng.standard_normal(pos.shape) * 5).astype(np.float32)


def params(n, **kw):
    base = dict(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / n)
    base.update(kw)
    return L.default_params(**base)


@pytest.mark.parametrize("n,B,kw", [
    (32, 4, {}),
    (48, 8, dict(solver_mode=1, preconditioner=L.PRECOND_MG)),
    (40, 8, dict(solver_mode=1, preconditioner=L.PRECOND_MG, refine_tol=1e-6, velocity_fp16=1)),
    (64, 8, dict(solver_mode=1, preconditioner=L.PRECOND_JACOBI)),
])
def test_plan_bytes_cover_the_handle(n, B, kw):
    pos, vel = scene(n)
    g = L.grid_desc(n, n, n, block_size=B)
    prm = params(n, **kw)
    planned = L.Server.request_bytes(g, prm, pos.shape[0])
    assert planned == L.plan_bytes(g, prm, pos.shape[0])
    h = L.FlipHandle(g, prm, particle_capacity=pos.shape[0])
    h.set_particles(pos, vel)
    for _ in range(3):
        h.step(DT)
    h.get_particles()
    live, peak = h.device_bytes()
    print("n=%d B=%d planned=%d peak=%d live=%d" % (n, B, planned, peak, live), flush=True)
    assert peak <= planned
    assert peak >= 0.9 * planned          # the plan is the same allocation sequence, not a guess
    h.close()


def _same(a, b):
    return all(np.array_equal(a[k], b[k]) for k in ("pos", "vel", "adj"))


def test_failed_particle_growth_keeps_the_handle():
    n = 32
    pos, vel = scene(n, 0.3, 0.3, 2)
    extra, evel = scene(n, 0.6, 0.6, 2, seed=7)
    g = L.grid_desc(n, n, n, block_size=4)
    h = L.FlipHandle(g, params(n), particle_capacity=pos.shape[0])
    h.set_particles(pos, vel)
    try:
        for skip in (0, 5, 17):             # fail the first, a middle and a late array of the growth
            L.inject_alloc_failure(skip, 1)
            rc = L.load().vare_flip_add_particles(h.h, L.ptr(extra), L.ptr(evel), extra.shape[0], L.SPACE_VOXEL,
                                                  L.MEM_HOST)
            assert rc == L.VARE_ENOMEM, rc
            assert h.num_particles() == pos.shape[0]
    finally:
        L.inject_alloc_failure(0, 0)
    h.step(DT)
    ref = L.FlipHandle(g, params(n))
    ref.set_particles(pos, vel)
    ref.step(DT)
    assert _same(h.get_particles(), ref.get_particles())
    # the handle grows once the allocations succeed again
    h.add_particles(extra, evel)
    assert h.num_particles() == pos.shape[0] + extra.shape[0]
    h.step(DT)


@pytest.mark.parametrize("skip", [0, 3, 20])
def test_failed_block_growth_keeps_the_handle(skip):
    n = 48
    pos, vel = scene(n, 0.5, 0.4, 4)
    g = L.grid_desc(n, n, n, block_size=8)
    prm = params(n, solver_mode=1, preconditioner=L.PRECOND_MG)
    h = L.FlipHandle(g, prm, block_capacity=1)      # the first step must grow the cell arrays
    h.set_particles(pos, vel)
    try:
        L.inject_alloc_failure(skip, 1)
        assert h.step_rc(DT) == L.VARE_ENOMEM
    finally:
        L.inject_alloc_failure(0, 0)
    assert h.num_particles() == pos.shape[0]
    h.step(DT)
    ref = L.FlipHandle(g, prm)
    ref.set_particles(pos, vel)
    ref.step(DT)
    assert _same(h.get_particles(), ref.get_particles())


def test_kernel_timing_counts_only_iterations_that_ran():
    n = 48
    pos, vel = scene(n, 0.5, 0.5, 4)
    g = L.grid_desc(n, n, n, block_size=8)
    h = L.FlipHandle(g, params(n, solver_mode=1, preconditioner=L.PRECOND_MG, refine_tol=1e-5))
    h.set_particles(pos, vel)
    h.step(DT)                                       # sizes the enqueue-ahead chunks
    h.set_kernel_timing(("update", "mg0", "spmv"))
    total = 0
    for _ in range(4):
        h.step(DT)
        total += h.stats()["cg_iterations"]
    t = h.kernel_timing(reset=True)
    print("timed launches", {k: t[k][1] for k in ("update", "mg0", "spmv")}, "iterations", total, flush=True)
    # one update pass per iteration; one level-0 prolongation per V-cycle: the set-up V-cycle of each
    # solve plus every iteration but the last
    assert t["update"][1] == total
    assert t["mg0"][1] == total
    assert 0 < t["spmv"][1] < total
