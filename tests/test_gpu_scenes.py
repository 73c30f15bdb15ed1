"""GPU parity on the scenes the benchmark configurations use, at sizes the oracle finishes in seconds.

* the wave-on-obstacle scene (the 1024x512x1024 configuration at 64x32x64, world 2x1x2): a non-cubic
  frame (fx != fy, so the block activation's index aliasing borrows across a different row length),
  a wall of water moving +x at 4 m/s, an inflow slab emitted through vare_flip_add_particles, a static
  box obstacle, FP16 particle velocities and FP64 refinement to 1e-6 -- stage by stage against the
  FP32 and FP64 oracles;
* G2P with collisions and FP16 storage on identical grid fields: positions, adjacency and the removal
  mask bit for bit, velocities equal to the oracle's rounded to FP16;
* the water-tank mode (FLIPMethod.h:57, SparseFrame.cu:163-173, FLIPMethod.cu:102-114): blocks below
  the tank level always active, wall faces below and open faces above it, damping near the side faces;
* a hydrostatic column against its closed form (linear pressure, cancelled velocity);
* 100-step aggregates of the dam break against the oracle run.

Bars as in test_gpu_parity.py: integer outputs bit-exact; P2G 1e-6 and divergence 1e-5 relative L2
against the FP32 oracle; pressure and projected velocity within max(2 x the FP32 oracle's own error,
1e-5) of the FP64 oracle; particle velocities within 2e-3 with FP16 storage (unit roundoff 2^-11).
"""
import numpy as np
import pytest

from vare_amd import _lib as L
from oracle import oracle as O

from test_gpu_parity import rel

pytestmark = pytest.mark.gpu

# wave on an obstacle, reduced: 64x32x64 cells, dx = 1/32
WDIMS = (64, 32, 64)
WWORLD = (2.0, 1.0, 2.0)
WRES = 32
WDT = 1.0 / 240.0
WAVE_V = 4.0                                 # m/s along +x
BOX = ((0.34, 0.0, 0.6), (0.54, 0.5, 1.4))   # world


def wave_params(**kw):
    base = dict(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / WRES, solver_mode=1,
                preconditioner=L.PRECOND_MG, cg_rel_tol=1e-3, refine_tol=1e-6, refine_max=16, max_iterations=4000,
                velocity_fp16=1)
    base.update(kw)
    return base


def oracle_params(**kw):
    base = dict(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / WRES, solver_mode=1,
                cg_rel_tol=1e-6, max_iterations=20000, preconditioner=1)
    base.update(kw)
    return O.params(**base)


def inflow_slab():
    """Particles of the inflow slab x < 4 dt (world), y < 0.3, all z, moving +x at 4 m/s (voxel units)."""
    nx, ny, nz = WDIMS
    pos = L.seed_box(nx, ny, nz, WAVE_V * WDT * WRES, 0.3 * WRES, float(nz), 8, 77)
    vel = np.zeros_like(pos)
    vel[:, 0] = np.float32(WAVE_V * WRES)                  # 128 voxels/s, exact in FP16
    return pos, vel


def wave_state(steps=6):
    """The wave after `steps` GPU steps with the inflow slab emitted before each one (voxel space)."""
    nx, ny, nz = WDIMS
    pos = L.seed_box(nx, ny, nz, 0.3 * WRES, 0.3 * WRES, float(nz), 8, 5)
    vel = np.zeros_like(pos)
    vel[:, 0] = np.float32(WAVE_V * WRES)
    slab_p, slab_v = inflow_slab()
    g = L.grid_desc(nx, ny, nz, (0, 0, 0), WWORLD, block_size=8)
    h = L.FlipHandle(g, L.default_params(**wave_params()), particle_capacity=pos.shape[0] + (steps + 2) * slab_p.shape[0])
    h.set_collision_box(BOX[0], BOX[1], (0.0, 0.0, 0.0), 3.0)
    h.set_particles(pos, vel)
    for _ in range(steps):
        h.add_particles(slab_p, slab_v)
        h.step(WDT)
    return h.get_particles()


def cpu_binning_dims(pos, pointers, B, f, dims):
    """Stable counting sort by the storage index of the P2G base cell (bin Nc outside the domain)."""
    fx, fy, fz = f
    nc = int((pointers >= 0).sum()) * B ** 3
    inside = np.all((pos >= 0) & (pos <= np.asarray(dims, np.float32)), axis=1)
    base = np.maximum(np.trunc(pos - np.float32(0.5)).astype(np.int64), 0)
    base = np.where(inside[:, None], base, 0)
    blk = base // B
    loc = base % B
    ptr = pointers[blk[:, 0] + fx * (blk[:, 1] + fy * blk[:, 2])]
    key = np.where(inside & (ptr >= 0), ptr + loc[:, 0] + B * (loc[:, 1] + B * loc[:, 2]), nc)
    perm = np.argsort(key, kind="stable").astype(np.int32)
    return perm, np.searchsorted(key[perm], np.arange(nc + 2)).astype(np.int32)


def wave_pair(d, slab_p, slab_v, vort):
    nx, ny, nz = WDIMS
    g = L.grid_desc(nx, ny, nz, (0, 0, 0), WWORLD, block_size=8)
    n0 = d["pos"].shape[0]
    h = L.FlipHandle(g, L.default_params(**wave_params()), particle_capacity=n0 + slab_p.shape[0])
    h.set_particles(d["pos"], d["vel"], vort=vort[:n0], curv=d["curv"], adj=d["adj"])
    h.add_particles(slab_p, slab_v)                        # the emitter: appended in reference order
    allp = np.concatenate([d["pos"], slab_p])
    allv = np.concatenate([d["vel"], slab_v])
    zeros = np.zeros(slab_p.shape[0], np.float32)
    o32 = O.Oracle(WDIMS, (0, 0, 0), WWORLD, B=8, prm=oracle_params(), fp64=False)
    o64 = O.Oracle(WDIMS, (0, 0, 0), WWORLD, B=8, prm=oracle_params(cg_rel_tol=1e-12), fp64=True)
    for o in (o32, o64):
        o.set_particles(allp, allv, vort=vort, curv=np.concatenate([d["curv"], zeros]),
                        adj=np.concatenate([d["adj"], zeros]))
    for x in (h, o32, o64):
        x.set_collision_box(BOX[0], BOX[1], (0.0, 0.0, 0.0), 3.0)
    return h, o32, o64, allp


def test_wave_obstacle_noncubic_stagewise():
    d = wave_state()
    slab_p, slab_v = inflow_slab()
    n_all = d["pos"].shape[0] + slab_p.shape[0]
    vort = np.arange(n_all, dtype=np.float32)               # rides along: identifies each particle
    h, o32, o64, allp = wave_pair(d, slab_p, slab_v, vort)
    print("case wave 64x32x64 np=%d (inflow %d)" % (n_all, slab_p.shape[0]), flush=True)
    assert h.num_particles() == n_all

    # stage 0: block activation on the non-cubic frame and the binning permutation -- bit-exact
    h.stage(0)
    for o in (o32, o64):
        o.stage(0)
    gp, gc = h.block_table()
    op, oc = o32.block_table()
    np.testing.assert_array_equal(gp, op)
    np.testing.assert_array_equal(gc, oc)
    fr = h.frame()
    assert (fr.fx, fr.fy, fr.fz) == (8, 4, 8)
    perm, bstart = h.binning()
    cperm, cbstart = cpu_binning_dims(allp, gp, 8, (8, 4, 8), WDIMS)
    np.testing.assert_array_equal(bstart, cbstart)
    np.testing.assert_array_equal(perm, cperm)

    # stage 1: P2G, count field, classification with the obstacle's solid cells
    h.stage(1)
    for o in (o32, o64):
        o.stage(1)
    np.testing.assert_array_equal(h.cell_field(L.FIELD_COUNT), o32.field(O.FIELD_COUNT))
    mark = np.minimum(h.cell_field(L.FIELD_MARK), 4)
    np.testing.assert_array_equal(mark, o32.field(O.FIELD_MARK).astype(np.uint64))
    assert (mark == 1).sum() > 100
    assert rel(h.cell_field(L.FIELD_MASS), o32.field(O.FIELD_MASS)) < 1e-6
    assert rel(h.cell_field(L.FIELD_VELOCITY), o32.field(O.FIELD_VELO)) < 1e-6

    # stage 2: fluid numbering bit-exact, divergence
    h.stage(2)
    for o in (o32, o64):
        o.stage(2)
    np.testing.assert_array_equal(h.cell_field(L.FIELD_MARK), o32.field(O.FIELD_MARK).astype(np.uint64))
    fluid = o32.field(O.FIELD_FLUID).astype(np.int64)
    assert rel(h.cell_field(L.FIELD_DIV)[fluid], o32.field(O.FIELD_B)) < 1e-5

    # stage 3: FP32 MG-PCG + FP64 refinement to 1e-6 against the FP64 oracle's converged solve
    h.stage(3)
    for o in (o32, o64):
        o.stage(3)
    st = h.stats()
    assert st["converged"] == 1 and st["true_rel_residual"] <= 1e-6 * 1.0001
    e_p = rel(h.cell_field(L.FIELD_PRESSURE), o64.field(O.FIELD_PRES))
    e_p32 = rel(o32.field(O.FIELD_PRES), o64.field(O.FIELD_PRES))
    e_c = rel(h.cell_field(L.FIELD_COMP), o64.field(O.FIELD_COMP))
    e_c32 = rel(o32.field(O.FIELD_COMP), o64.field(O.FIELD_COMP))
    print("pressure err gpu=%.3g oracle32=%.3g  velocity err gpu=%.3g oracle32=%.3g  refine=%d iters=%d" % (
        e_p, e_p32, e_c, e_c32, st["refine_steps"], st["cg_iterations"]), flush=True)
    assert e_p < max(2 * e_p32, 1e-5)
    assert e_c < max(2 * e_c32, 1e-5)

    # stage 4: G2P with the obstacle and FP16 velocities (k_g2p_tile<8, true, true>)
    h.stage(4, WDT)
    for o in (o32, o64):
        o.stage(4, WDT)
    g = h.get_particles()
    r64 = o64.get_particles()
    r32 = o32.get_particles()
    assert rel(g["pos"], r64["pos"]) < 1e-6
    ev, ev32 = rel(g["vel"], r64["vel"]), rel(r32["vel"], r64["vel"])
    print("particle velocity err gpu=%.3g oracle32=%.3g" % (ev, ev32), flush=True)
    assert ev < max(2 * ev32, 2e-3)
    np.testing.assert_array_equal(g["vel"], g["vel"].astype(np.float16).astype(np.float32))
    # removal: the flags agree except where the oracle's final position lies within 1e-3 voxels of a
    # threshold (signed distance 0 or 3 after a projection, or a domain face)
    gm, om = h.removed_mask(), o32.removed()
    assert om.sum() > 0
    diff = np.nonzero(gm != om)[0]
    if diff.size:
        phi, _, _ = o32.collision_query(r32["pos"][diff])
        edge = np.min(np.abs(np.concatenate([r32["pos"][diff], np.asarray(WDIMS, np.float64) - r32["pos"][diff]], 1)),
                      axis=1)
        borderline = (np.abs(phi) < 1e-3) | (np.abs(phi - 3.0) < 1e-3) | (edge < 1e-3)
        print("removal flags differing: %d of %d, all borderline: %s" % (diff.size, om.sum(), borderline.all()),
              flush=True)
        assert borderline.all()

    # stage 5: Array::remove on the GPU's own mask is the reference permutation
    h.stage(5)
    g5 = h.get_particles()
    assert g5["pos"].shape[0] == n_all - int(gm.sum())
    # vorticity identifies each particle: its index, 0 for the emitted ones (newParticle gives them 0)
    ident = np.concatenate([vort[:d["pos"].shape[0]], np.zeros(slab_p.shape[0], np.float32)])
    np.testing.assert_array_equal(g5["vort"], ident[O.remove_perm(gm)])


def test_g2p_collision_fp16_bit_exact_on_identical_fields():
    """FLIPMethod::updateParticles (FLIPMethod.cu:54-175) with the obstacle and FP16 storage, given
    the GPU's own grid fields: the same FP32 arithmetic in the same order, so positions, adjacency
    and removal flags are bit-identical and velocities equal the oracle's rounded to FP16."""
    d = wave_state(5)
    slab_p, slab_v = inflow_slab()
    n_all = d["pos"].shape[0] + slab_p.shape[0]
    h, o32, _, _ = wave_pair(d, slab_p, slab_v, np.zeros(n_all, np.float32))
    for s in range(4):
        h.stage(s, WDT)
        o32.stage(s, WDT)
    o32.set_field(O.FIELD_VELO, h.cell_field(L.FIELD_VELOCITY).astype(np.float64))
    o32.set_field(O.FIELD_COMP, h.cell_field(L.FIELD_COMP).astype(np.float64))
    o32.set_field(O.FIELD_COUNT, h.cell_field(L.FIELD_COUNT).astype(np.float64))
    h.stage(4, WDT)
    o32.stage(4, WDT)
    g, r = h.get_particles(), o32.get_particles()
    gm, om = h.removed_mask(), o32.removed()
    print("identical fields: np=%d removed gpu=%d oracle=%d" % (n_all, gm.sum(), om.sum()), flush=True)
    assert om.sum() > 0
    np.testing.assert_array_equal(gm, om)
    np.testing.assert_array_equal(g["pos"], r["pos"].astype(np.float32))
    np.testing.assert_array_equal(g["adj"], r["adj"].astype(np.float32))
    np.testing.assert_array_equal(g["vel"], r["vel"].astype(np.float32).astype(np.float16).astype(np.float32))


@pytest.mark.parametrize("B", [4, 8])
def test_water_tank_stagewise(B):
    """enableWaterTank: every block below the tank level is active (SparseFrame.cu:163-173), domain
    faces are walls below the level and open above it, particles are clamped only below it and damped
    within dampingBand of the side faces (FLIPMethod.cu:102-114, :135)."""
    n = 32
    pool = O.seed_box(n, n, n, float(n), 0.3 * n, float(n), 4, 31)
    blob = O.seed_box(n, n, n, 0.3 * n, 0.8 * n, 0.4 * n, 4, 32)
    blob = blob[blob[:, 1] > 0.5 * n]
    pos = np.concatenate([pool, blob]).astype(np.float32)
    rng = np.random.default_rng(33)
    vel = (rng.standard_normal(pos.shape) * 30).astype(np.float32)
    vel[pool.shape[0]:, 0] -= 400.0                       # part of the blob leaves through the open x=0 face
    kw = dict(ext_force=(0.0, -9.8, 0.0), wall_level=0.35, enable_water_tank=1, damping_band=0.1,
              voxel_size=1.0 / n, solver_mode=1, cg_rel_tol=1e-6, max_iterations=3000)
    g = L.grid_desc(n, n, n, block_size=B)
    h = L.FlipHandle(g, L.default_params(**kw))
    h.set_particles(pos, vel)
    o32 = O.Oracle((n, n, n), B=B, prm=O.params(**kw), fp64=False)
    o64 = O.Oracle((n, n, n), B=B, prm=O.params(**kw), fp64=True)
    for o in (o32, o64):
        o.set_particles(pos, vel)
    print("case water tank B=%d np=%d" % (B, pos.shape[0]), flush=True)
    h.stage(0)
    for o in (o32, o64):
        o.stage(0)
    gp, gc = h.block_table()
    op, oc = o32.block_table()
    np.testing.assert_array_equal(gp, op)
    np.testing.assert_array_equal(gc, oc)
    # the tank marks every block row below the level, also where no particle is
    tall = int(np.float32(0.35 * n)) // B
    f = n // B
    assert np.all(gp.reshape(f, f, f)[:, :tall, :] >= 0)
    for s in (1, 2):
        h.stage(s)
        for o in (o32, o64):
            o.stage(s)
    np.testing.assert_array_equal(h.cell_field(L.FIELD_COUNT), o32.field(O.FIELD_COUNT))
    np.testing.assert_array_equal(h.cell_field(L.FIELD_MARK), o32.field(O.FIELD_MARK).astype(np.uint64))
    fluid = o32.field(O.FIELD_FLUID).astype(np.int64)
    assert rel(h.cell_field(L.FIELD_DIV)[fluid], o32.field(O.FIELD_B)) < 1e-5
    h.stage(3)
    for o in (o32, o64):
        o.stage(3)
    e_p = rel(h.cell_field(L.FIELD_PRESSURE), o64.field(O.FIELD_PRES))
    e_p32 = rel(o32.field(O.FIELD_PRES), o64.field(O.FIELD_PRES))
    print("tank pressure err gpu=%.3g oracle32=%.3g" % (e_p, e_p32), flush=True)
    assert e_p < max(2 * e_p32, 1e-5)
    h.stage(4, WDT)
    for o in (o32, o64):
        o.stage(4, WDT)
    gm, om = h.removed_mask(), o32.removed()
    np.testing.assert_array_equal(gm, om)
    assert om.sum() > 0                                   # particles left through the open face
    gpt, r32, r64 = h.get_particles(), o32.get_particles(), o64.get_particles()
    np.testing.assert_array_equal(gpt["adj"], r32["adj"].astype(np.float32))
    assert rel(gpt["pos"], r64["pos"]) < 1e-6
    assert rel(gpt["vel"], r64["vel"]) < max(2 * rel(r32["vel"], r64["vel"]), 1e-5)
    h.stage(5)
    for o in (o32, o64):
        o.stage(5)
    assert h.num_particles() == o32.num_particles() == pos.shape[0] - int(om.sum())


@pytest.mark.parametrize("B", [4, 8])
def test_hydrostatic_column_closed_form(B):
    """A column of water at rest under gravity in a closed box: the cells j <= H are fluid (every
    stencil corner of a particle with y < H), the right-hand side is g dt on the bottom row and the
    pressure the hydrostatic p_j = g dt (H + 1 - j - 1/2) (zero at the free surface, the face above
    the top fluid row); the projection cancels the velocity g dt of every fluid cell."""
    n, H = 32, 10
    pos = L.seed_box(n, n, n, float(n), float(H), float(n), 8, 41)
    gv = np.float32(9.8 * n)                              # |gravity| in voxel units
    c = np.float32(gv / np.float32(240.0))
    vel = np.zeros_like(pos)
    vel[:, 1] = -c
    g = L.grid_desc(n, n, n, block_size=B)
    h = L.FlipHandle(g, L.default_params(ext_force=(0, -9.8, 0), wall_level=1e30, voxel_size=1.0 / n, solver_mode=1,
                                         preconditioner=L.PRECOND_MG, cg_rel_tol=1e-7, max_iterations=2000))
    h.set_particles(pos, vel)
    for s in range(4):
        h.stage(s, 1.0 / 240.0)
    pointers, coords = h.block_table()
    nc = coords.shape[0] * B ** 3
    loc = np.arange(nc) % B ** 3
    ijk = coords[np.arange(nc) // B ** 3] * B + np.stack([loc % B, (loc // B) % B, loc // (B * B)], 1)
    mark = h.cell_field(L.FIELD_MARK)
    fl = mark >= 4
    np.testing.assert_array_equal(fl, ijk[:, 1] <= H)
    j = ijk[fl, 1]
    b = h.cell_field(L.FIELD_DIV)[fl]
    np.testing.assert_allclose(b, np.where(j == 0, float(c), 0.0), rtol=0, atol=2e-6 * float(c))
    p = h.cell_field(L.FIELD_PRESSURE)[fl].astype(np.float64)
    p_exact = float(c) * (H + 1 - j - 0.5)
    print("hydrostatic B=%d: pressure rel err %.2e, max |v| %.2e (g dt = %.4f)" % (
        B, rel(p, p_exact), np.abs(h.cell_field(L.FIELD_COMP)[fl]).max(), c), flush=True)
    assert rel(p, p_exact) < 1e-5
    assert np.abs(h.cell_field(L.FIELD_COMP)[fl]).max() < 1e-4 * float(c)


def test_dam_break_100_step_aggregates():
    """100 steps of the 32^3 dam break against the FP32 oracle run: the trajectories diverge from
    summation-order differences, so the comparison is on aggregates -- the particle count exactly,
    kinetic energy within 2%, the centre of mass within 0.05 voxels and the front (largest x) and the
    highest particle within 0.5 voxels."""
    n = 32
    pos = O.seed_box(n, n, n, 0.4 * n, 0.6 * n, float(n), 8, 1)
    vel = np.zeros_like(pos)
    kw = dict(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / n)
    h = L.FlipHandle(L.grid_desc(n, n, n, block_size=4), L.default_params(**kw))
    h.set_particles(pos, vel)
    o = O.Oracle((n, n, n), B=4, prm=O.params(**kw))
    o.set_particles(pos, vel)
    for s in range(100):
        h.step(1.0 / 240.0)
        o.step(1.0 / 240.0)
    g, r = h.get_particles(), o.get_particles()
    ke_g = float((g["vel"].astype(np.float64) ** 2).sum())
    ke_r = float((r["vel"] ** 2).sum())
    com_g, com_r = g["pos"].astype(np.float64).mean(0), r["pos"].mean(0)
    front_g, front_r = float(g["pos"][:, 0].max()), float(r["pos"][:, 0].max())
    top_g, top_r = float(g["pos"][:, 1].max()), float(r["pos"][:, 1].max())
    print("100 steps: KE %.6g / %.6g  com %s / %s  front %.3f / %.3f  top %.3f / %.3f" % (
        ke_g, ke_r, com_g, com_r, front_g, front_r, top_g, top_r), flush=True)
    assert g["pos"].shape[0] == r["pos"].shape[0]
    assert abs(ke_g - ke_r) / ke_r < 2e-2
    assert np.abs(com_g - com_r).max() < 0.05
    assert abs(front_g - front_r) < 0.5 and abs(top_g - top_r) < 0.5
