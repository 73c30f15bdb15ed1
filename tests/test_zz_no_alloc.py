"""GPU test of vare_flip_step_noalloc, the allocation-free variant of vare_flip_step.

Two handles on the same input, one stepped by vare_flip_step and one by vare_flip_step_noalloc, must
give bit-identical particles and solver statistics (every kernel of the step is deterministic and
the variant only moves the reductions' partials and the scans' tile sums into the static device
scratch).  Two calls in a row, the second on a smaller particle set; the no-allocation handle's
device bytes do not change across its calls.  Also its refusals: kernel timing on, a block capacity
the step outgrows.
"""
import numpy as np
import pytest

from vare_amd import _lib as L

pytestmark = pytest.mark.gpu

DT = 1.0 / 240.0


def params(n, **kw):
    base = dict(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / n)
    base.update(kw)
    return L.default_params(**base)


def scene(n, fx, fy, ppc, seed):
    pos = L.seed_box(n, n, n, fx * n, fy * n, float(n), ppc, seed)
    rng = np.random.default_rng(seed)
    return pos, (rng.standard_normal(pos.shape) * 4).astype(np.float32)


def same(a, b):
    for k in ("pos", "vel", "adj", "vort", "curv"):
        assert np.array_equal(a[k], b[k]), k


CASES = [
    (48, 8, dict(solver_mode=1, preconditioner=L.PRECOND_MG, refine_tol=1e-6)),
    (32, 4, dict()),                                                  # reference CG, parity block size
    (40, 8, dict(solver_mode=1, preconditioner=L.PRECOND_JACOBI, velocity_fp16=1)),
]


@pytest.mark.parametrize("n,B,kw", CASES)
def test_noalloc_step_equals_step_twice(n, B, kw):
    g = L.grid_desc(n, n, n, block_size=B)
    prm = params(n, **kw)
    big = scene(n, 0.5, 0.6, 6, 3)
    small = scene(n, 0.3, 0.4, 4, 4)
    ref = L.FlipHandle(g, prm, particle_capacity=big[0].shape[0])
    nal = L.FlipHandle(g, prm, particle_capacity=big[0].shape[0])
    for pos, vel in (big, small):
        print("noalloc case n=%d B=%d np=%d" % (n, B, pos.shape[0]), flush=True)
        ref.set_particles(pos, vel)
        nal.set_particles(pos, vel)
        before = nal.device_bytes()
        ref.step(DT)
        nal.step_noalloc(DT)
        assert nal.device_bytes() == before           # the variant allocated nothing
        same(ref.get_particles(), nal.get_particles())
        sr, sn = ref.stats(), nal.stats()
        for k in ("num_particles", "num_blocks", "num_fluid", "cg_iterations", "refine_steps", "converged"):
            assert sr[k] == sn[k], k
        assert sr["true_rel_residual"] == sn["true_rel_residual"]
    ref.close()
    nal.close()


def test_noalloc_step_refusals_leave_the_handle_usable():
    n = 48
    g = L.grid_desc(n, n, n, block_size=8)
    prm = params(n, solver_mode=1, preconditioner=L.PRECOND_MG)
    pos, vel = scene(n, 0.5, 0.5, 4, 5)
    h = L.FlipHandle(g, prm, block_capacity=1)
    h.set_particles(pos, vel)
    assert h.step_noalloc_rc(DT) == L.VARE_ENOMEM        # the active blocks outgrow the capacity
    h.set_kernel_timing(("update",))
    assert h.step_noalloc_rc(DT) == L.VARE_EINVAL        # timing creates events
    h.set_kernel_timing(())
    h.step(DT)                                           # the allocating step grows the handle
    ref = L.FlipHandle(g, prm)
    ref.set_particles(pos, vel)
    ref.step(DT)
    same(ref.get_particles(), h.get_particles())
    # a handle with the default capacity (the whole frame) takes the no-allocation step from there
    nal = L.FlipHandle(g, prm)
    d = ref.get_particles()
    nal.set_particles(d["pos"], d["vel"], vort=d["vort"], curv=d["curv"], adj=d["adj"])
    ref.step(DT)
    nal.step_noalloc(DT)
    same(ref.get_particles(), nal.get_particles())
