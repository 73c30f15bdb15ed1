"""GPU tests of vare_flip_step_batched: many independent tanks stepped with shared kernel launches give
bit for bit the result of stepping each tank alone (vare_flip_step), for tanks of different sizes, block
sizes, solvers, velocity storage and collision sources, and the launches are shared."""
import numpy as np
import pytest

from vare_amd import _lib as L

pytestmark = pytest.mark.gpu

DT = 1.0 / 240.0


def dam(nx, ny, nz, fx, fy, ppc, seed):
    pos = L.seed_box(nx, ny, nz, fx * nx, fy * ny, float(nz), ppc, seed)
    rng = np.random.default_rng(seed)
    vel = (rng.standard_normal(pos.shape) * 0.5).astype(np.float32)
    return pos, vel


def relative(n, **kw):
    args = dict(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / n, solver_mode=1, cg_rel_tol=1e-5,
                preconditioner=L.PRECOND_MG)
    args.update(kw)
    return L.default_params(**args)


# (name, grid, params, particles, collision box or None)
def scenes():
    out = []
    n = 48
    p, v = dam(n, n, n, 0.4, 0.6, 6, 1)
    out.append(("mg-refine 48^3 B8", L.grid_desc(n, n, n, block_size=8),
                relative(n, cg_rel_tol=1e-3, refine_tol=1e-5), (p, v), None))
    n = 32
    p, v = dam(n, n, n, 0.5, 0.5, 8, 2)
    out.append(("reference CG 32^3 B4", L.grid_desc(n, n, n, block_size=4),
                L.default_params(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / n), (p, v), None))
    n = 40
    p, v = dam(n, n, n, 0.3, 0.5, 5, 3)
    out.append(("fp16 + box obstacle 40^3 B8", L.grid_desc(n, n, n, block_size=8),
                relative(n, velocity_fp16=1), (p, v), ((0.5, 0.0, 0.25), (0.65, 0.5, 0.75))))
    p, v = dam(64, 32, 64, 0.25, 0.6, 4, 4)
    out.append(("jacobi 64x32x64 B8", L.grid_desc(64, 32, 64, bmax=(1.0, 0.5, 1.0), block_size=8),
                L.default_params(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / 64, solver_mode=1,
                                 cg_rel_tol=1e-5, preconditioner=L.PRECOND_JACOBI), (p, v), None))
    return out


def make(sc):
    name, g, prm, (p, v), box = sc
    h = L.FlipHandle(g, prm)
    h.set_particles(p, v)
    if box is not None:
        h.set_collision_box(box[0], box[1])
    return h


def same(a, b, what):
    for k in ("pos", "vel", "adj"):
        assert a[k].shape == b[k].shape, (what, k)
        assert np.array_equal(a[k], b[k]), (what, k, np.abs(a[k] - b[k]).max())


def test_batched_equals_alone_mixed_tanks():
    sc = scenes()
    steps = 4
    alone = []
    for s in sc:
        print("alone:", s[0], flush=True)
        h = make(s)
        its = []
        for _ in range(steps):
            h.step(DT)
            its.append(h.stats()["cg_iterations"])
        alone.append((h.get_particles(), its))
        h.close()
    hs = [make(s) for s in sc]
    print("batched: %d tanks" % len(hs), flush=True)
    L.batch_stats(reset=True)
    its = [[] for _ in hs]
    for _ in range(steps):
        st = L.step_batched(hs, DT)
        assert st == [0] * len(hs)
        for i, h in enumerate(hs):
            its[i].append(h.stats()["cg_iterations"])
    bs = L.batch_stats()
    print("batch stats", bs, flush=True)
    assert bs["steps"] == steps and bs["tanks"] == steps * len(hs)
    for i, s in enumerate(sc):
        same(hs[i].get_particles(), alone[i][0], s[0])
        assert its[i] == alone[i][1], (s[0], its[i], alone[i][1])
    for h in hs:
        h.close()


def test_batched_shares_launches():
    # twenty-four tanks of one configuration and uneven sizes (more than sixteen, so a merged launch
    # holds up to 32 tanks): the launches are merged (far fewer merged launches than the tanks would
    # make alone), and each result equals its tank alone
    sizes = [32, 40, 48, 56, 64, 32, 40, 48, 56, 64, 32, 48, 64, 40, 56, 72, 32, 40, 48, 56, 64, 40, 48, 32]
    tanks = []
    for k, n in enumerate(sizes):
        p, v = dam(n, n, n, 0.2 + 0.03 * k, 0.7 - 0.02 * k, 4 + k % 5, 100 + k)
        tanks.append(("tank %d (%d^3)" % (k, n), L.grid_desc(n, n, n, block_size=8),
                      relative(n, cg_rel_tol=1e-3, refine_tol=1e-5), (p, v), None))
    steps = 3
    hs = [make(s) for s in tanks]
    L.batch_stats(reset=True)
    for _ in range(steps):
        assert L.step_batched(hs, DT) == [0] * len(hs)
    bs = L.batch_stats()
    print("batch stats", bs, flush=True)
    assert bs["tank_launches"] > 4 * bs["launches"], bs
    for i in (0, 7, 15, 23):
        h = make(tanks[i])
        for _ in range(steps):
            h.step(DT)
        same(hs[i].get_particles(), h.get_particles(), tanks[i][0])
        h.close()
    for h in hs:
        h.close()


def test_batched_argument_errors():
    n = 32
    p, v = dam(n, n, n, 0.5, 0.5, 4, 7)
    a = make(("a", L.grid_desc(n, n, n, block_size=8), relative(n), (p, v), None))
    b = make(("b", L.grid_desc(n, n, n, block_size=8), relative(n), (p, v), None))
    with pytest.raises(L.VareError) as e:
        L.step_batched([a, a], DT)
    assert e.value.code == L.VARE_EINVAL
    b.set_kernel_timing(["update"])
    with pytest.raises(L.VareError) as e:
        L.step_batched([a, b], DT)
    assert e.value.code == L.VARE_EINVAL
    b.set_kernel_timing([])
    # one tank alone in a batch still steps (and matches the plain step of an identical tank)
    assert L.step_batched([a], DT) == [0]
    b.step(DT)
    same(a.get_particles(), b.get_particles(), "one-tank batch")
    a.close()
    b.close()
