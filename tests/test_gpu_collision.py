"""GPU parity of the collision source (CollisionSource.h; FLIPMethod.cu:116-164, :242-260) against
the oracle: the sampled field bit for bit, then one step stage by stage with a box obstacle."""
import numpy as np
import pytest

from oracle import oracle as O
import vare_amd._lib as L

from test_gpu_parity import DT, make_scene, pair, rel

pytestmark = pytest.mark.gpu

BOX_LO, BOX_HI, BOX_V = (0.45, 0.0, 0.25), (0.7, 0.35, 0.75), (0.0, 0.5, 0.0)


@pytest.mark.parametrize("B", [4, 8])
def test_collision_box_field_bit_exact(B):
    n = 32
    pos, vel = make_scene(n, seed=21)
    h, o32, _ = pair(n, B, pos, vel)
    print("case collision field B=%d" % B, flush=True)
    h.set_collision_box(BOX_LO, BOX_HI, BOX_V, 3.0)
    o32.set_collision_box(BOX_LO, BOX_HI, BOX_V, 3.0)
    gp, gphi, gvel, gnrm, gbg = h.collision()
    op, ophi, ovel, onrm, obg = o32.collision()
    np.testing.assert_array_equal(gp, op)
    assert gphi.size > 0 and gbg == obg
    np.testing.assert_array_equal(gphi, ophi.astype(np.float32))
    np.testing.assert_array_equal(gvel, ovel.astype(np.float32))
    np.testing.assert_array_equal(gnrm, onrm.astype(np.float32))


def test_collision_sampler_callback_bit_exact():
    # a sphere through the general seeds + sampler entry point, same callback on both sides
    n, B = 16, 4
    pos, vel = make_scene(n, seed=22)
    h, o32, _ = pair(n, B, pos, vel)
    c, r = np.array([0.5, 0.3, 0.5]), 0.2

    def sample(p):
        d = p.astype(np.float64) - c
        dist = float(np.sqrt(d @ d))
        return np.float32(dist - r), (0.1, 0.0, -0.2), d / max(dist, 1e-12)

    ii = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    wp = ((ii + 0.5) / n).astype(np.float32)
    seeds = wp[np.abs(np.linalg.norm(wp - c, axis=1) - r) < 2.0 / n]
    print("case collision callback seeds=%d" % len(seeds), flush=True)
    h.set_collision(seeds, 2.0 / n, sample)
    o32.set_collision(seeds, 2.0 / n, sample)
    g, o = h.collision(), o32.collision()
    np.testing.assert_array_equal(g[0], o[0])
    for a, b in zip(g[1:4], o[1:4]):
        np.testing.assert_array_equal(a, b.astype(np.float32))


@pytest.mark.parametrize("B", [4, 8])
def test_collision_step_stagewise(B):
    # dam break hitting a box obstacle that also pushes upward: solid cells, collision velocity,
    # solid stencil terms, projection and removal
    n = 32
    pos, vel = make_scene(n, wx=0.5, wy=0.6, seed=23, vel_scale=10.0)
    kw = dict(solver_mode=1, cg_rel_tol=1e-6, max_iterations=3000, col_vel_scale=1.0)
    h, o32, o64 = pair(n, B, pos, vel, **kw)
    print("case collision stagewise B=%d np=%d" % (B, pos.shape[0]), flush=True)
    for x in (h, o32, o64):
        x.set_collision_box(BOX_LO, BOX_HI, BOX_V, 3.0)
    for s in range(2):
        h.stage(s, DT)
        o32.stage(s, DT)
        o64.stage(s, DT)
    # cell classes (air 0, solid 1, fluid 4; fluid ranks are numbered in the next stage)
    mark = np.minimum(h.cell_field(L.FIELD_MARK), 4)
    np.testing.assert_array_equal(mark, o32.field(O.FIELD_MARK).astype(np.uint64))
    assert (mark == 1).sum() > 100                     # solid cells present
    assert rel(h.cell_field(L.FIELD_VELOCITY), o32.field(O.FIELD_VELO)) < 1e-6
    h.stage(2, DT)
    o32.stage(2, DT)
    o64.stage(2, DT)
    np.testing.assert_array_equal(h.cell_field(L.FIELD_MARK), o32.field(O.FIELD_MARK).astype(np.uint64))
    fluid = o32.field(O.FIELD_FLUID).astype(np.int64)
    assert rel(h.cell_field(L.FIELD_DIV)[fluid], o32.field(O.FIELD_B)) < 1e-5
    h.stage(3, DT)
    o64.stage(3, DT)
    assert rel(h.cell_field(L.FIELD_PRESSURE), o64.field(O.FIELD_PRES)) < 1e-4
    assert rel(h.cell_field(L.FIELD_COMP), o64.field(O.FIELD_COMP)) < 1e-4
    h.stage(4, DT)
    o64.stage(4, DT)
    g = h.get_particles()
    o = o64.get_particles()
    assert rel(g["pos"], o["pos"]) < 1e-6
    assert rel(g["vel"], o["vel"]) < 1e-4
    gm, om = h.removed_mask(), o64.removed()
    # particles projected onto the surface are removed when their final distance reads < 0, which
    # can flip with the last bit of the position; everything else agrees exactly
    assert om.sum() > 0
    assert (gm != om).sum() <= max(2, om.sum() // 50)


def test_collision_clear_restores_plain_step():
    n, B = 32, 8
    pos, vel = make_scene(n, seed=24)
    out = []
    for mode in ("none", "cleared"):
        g = L.grid_desc(n, n, n, block_size=B)
        h = L.FlipHandle(g, L.default_params(ext_force=(0, -9.8, 0), wall_level=1e30, voxel_size=1 / n,
                                             solver_mode=1, cg_rel_tol=1e-6, preconditioner=L.PRECOND_MG))
        if mode == "cleared":
            h.set_collision_box(BOX_LO, BOX_HI, BOX_V, 3.0)
            h.clear_collision()
        h.set_particles(pos, vel)
        h.step(DT)
        out.append(h.get_particles()["pos"])
    np.testing.assert_array_equal(out[0], out[1])


def test_collision_bad_arguments():
    n = 16
    g = L.grid_desc(n, n, n, block_size=4)
    h = L.FlipHandle(g, L.default_params(voxel_size=1 / n))
    f3 = L.C.c_float * 3
    assert h.lib.vare_flip_set_collision_box(h.h, f3(0.5, 0, 0), f3(0.4, 1, 1), f3(0, 0, 0), 3.0) == L.VARE_EINVAL
    assert h.lib.vare_flip_set_collision_box(h.h, f3(0, 0, 0), f3(1, 1, 1), f3(0, 0, 0), 0.0) == L.VARE_EINVAL
    # a box outside the domain activates nothing: queries read the background (+band voxels)
    h.set_collision_box((5, 5, 5), (6, 6, 6), (0, 0, 0), 3.0)
    p, phi, _, _, bg = h.collision()
    assert phi.size == 0 and np.all(p == -1) and bg == pytest.approx(3.0)
