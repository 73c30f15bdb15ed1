"""GPU parity of the HIP step against the CPU oracle, stage by stage, through the C ABI.

Bars (stated per check):
  * integer / index / byte outputs -- block table, count field, cell marks and fluid numbering,
    adjacency, removal mask and permutation: bit-exact vs the FP32 oracle on the same input;
  * floating outputs: relative L2 vs the FP32 oracle of 1e-6 (P2G), 1e-5 (divergence, particles),
    and for the pressure solve the GPU's error against the FP64 oracle must stay within
    max(2 x the FP32 oracle's own error, 1e-5).  Differences come from float-atomic P2G summation
    order and the dot-product reduction order.
"""
import numpy as np
import pytest

from vare_amd import _lib as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DT = 1.0 / 240.0


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def make_scene(n=32, wx=0.4, wy=0.6, ppc=8, seed=1, vel_scale=20.0):
    pos = O.seed_box(n, n, n, wx * n, wy * n, n, ppc, seed)
    rng = np.random.default_rng(seed)
    vel = (rng.standard_normal(pos.shape) * vel_scale).astype(np.float32)
    return pos, vel


def pair(n, B, pos, vel, vort=None, **kw):
    base = dict(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / n)
    base.update(kw)
    g = L.grid_desc(n, n, n, (0, 0, 0), (1, 1, 1), B)
    h = L.FlipHandle(g, L.default_params(**base))
    h.set_particles(pos, vel, vort=vort)
    o32 = O.Oracle((n, n, n), B=B, prm=O.params(**base), fp64=False)
    o64 = O.Oracle((n, n, n), B=B, prm=O.params(**base), fp64=True)
    for o in (o32, o64):
        o.set_particles(pos, vel, vort=vort)
    return h, o32, o64


@pytest.mark.parametrize("B", [4, 8])
def test_stagewise_parity(B):
    n = 32
    pos, vel = make_scene(n)
    h, o32, o64 = pair(n, B, pos, vel, solver_mode=1, cg_rel_tol=1e-6, max_iterations=3000)
    print("case stagewise B=%d n=%d np=%d" % (B, n, pos.shape[0]), flush=True)

    # stage 0: binning + block activation -- bit-exact
    h.stage(0)
    for o in (o32, o64):
        o.stage(0)
    gp, gc = h.block_table()
    op, oc = o32.block_table()
    np.testing.assert_array_equal(gp, op)
    np.testing.assert_array_equal(gc, oc)

    # stage 1: P2G + count + classification
    h.stage(1)
    for o in (o32, o64):
        o.stage(1)
    np.testing.assert_array_equal(h.cell_field(L.FIELD_COUNT), o32.field(O.FIELD_COUNT))
    assert rel(h.cell_field(L.FIELD_MASS), o32.field(O.FIELD_MASS)) < 1e-6
    assert rel(h.cell_field(L.FIELD_VELOCITY), o32.field(O.FIELD_VELO)) < 1e-6
    assert np.all((h.cell_field(L.FIELD_MASS) > 0) == (o32.field(O.FIELD_MASS) > 0))

    # stage 2: fluid numbering (bit-exact) + divergence
    h.stage(2)
    for o in (o32, o64):
        o.stage(2)
    np.testing.assert_array_equal(h.cell_field(L.FIELD_MARK), o32.field(O.FIELD_MARK).astype(np.uint64))
    fluid = o32.field(O.FIELD_FLUID).astype(np.int64)
    b_gpu = h.cell_field(L.FIELD_DIV)
    assert rel(b_gpu[fluid], o32.field(O.FIELD_B)) < 1e-5
    nonfluid = np.ones(b_gpu.size, bool)
    nonfluid[fluid] = False
    assert np.all(b_gpu[nonfluid] == 0)

    # stage 3: solve + pressure apply
    h.stage(3)
    for o in (o32, o64):
        o.stage(3)
    st = h.stats()
    assert st["num_fluid"] == fluid.size
    e_gpu = rel(h.cell_field(L.FIELD_PRESSURE), o64.field(O.FIELD_PRES))
    e_o32 = rel(o32.field(O.FIELD_PRES), o64.field(O.FIELD_PRES))
    print("pressure err gpu=%.3g oracle32=%.3g iters gpu=%d o32=%d o64=%d" % (
        e_gpu, e_o32, st["cg_iterations"], o32.stats()["cg_iters"], o64.stats()["cg_iters"]), flush=True)
    assert st["converged"] == 1
    assert abs(st["cg_iterations"] - o64.stats()["cg_iters"]) <= max(3, 0.05 * o64.stats()["cg_iters"])
    assert e_gpu < max(2 * e_o32, 1e-5)
    c_gpu = rel(h.cell_field(L.FIELD_COMP), o64.field(O.FIELD_COMP))
    c_o32 = rel(o32.field(O.FIELD_COMP), o64.field(O.FIELD_COMP))
    assert c_gpu < max(2 * c_o32, 1e-5)

    # stage 4: G2P + advection
    h.stage(4, DT)
    for o in (o32, o64):
        o.stage(4, DT)
    np.testing.assert_array_equal(h.removed_mask(), o32.removed())
    g = h.get_particles()
    r = o32.get_particles()
    r64 = o64.get_particles()
    np.testing.assert_array_equal(g["adj"], r["adj"])
    assert rel(g["pos"], r64["pos"]) < 1e-6
    ev_gpu = rel(g["vel"], r64["vel"])
    ev_o32 = rel(r["vel"], r64["vel"])
    print("velocity err gpu=%.3g oracle32=%.3g" % (ev_gpu, ev_o32), flush=True)
    assert ev_gpu < max(2 * ev_o32, 1e-5)

    # stage 5: removal (none in a closed box)
    h.stage(5)
    for o in (o32, o64):
        o.stage(5)
    assert h.num_particles() == o32.num_particles() == pos.shape[0]


@pytest.mark.parametrize("B", [4, 8])
def test_reference_mode_step(B):
    # reference solver: unpreconditioned CG, |r| < 1e-3 absolute or 202 products
    n = 32
    pos, vel = make_scene(n, seed=2)
    h, o32, o64 = pair(n, B, pos, vel)
    h.step(DT)
    it32 = o32.step(DT)
    o64.step(DT)
    st = h.stats()
    print("reference mode B=%d iters gpu=%d o32=%d" % (B, st["cg_iterations"], it32), flush=True)
    assert abs(st["cg_iterations"] - it32) <= 3
    g = h.get_particles()
    r = o32.get_particles()
    r64 = o64.get_particles()
    np.testing.assert_array_equal(g["adj"], r["adj"])
    assert rel(g["vel"], r64["vel"]) < max(2 * rel(r["vel"], r64["vel"]), 1e-5)


def test_removal_order_bit_exact():
    # open domain (wallLevel = -1e30: no clamping, FLIPMethod.cu:135) with particles falling out
    # of the bottom: removal flags and the Array::remove permutation must match exactly.
    n = 32
    pos, vel = make_scene(n, wx=0.5, wy=0.3, seed=3, vel_scale=5.0)
    low = pos[:, 1] < 2.0
    vel[low, 1] = -900.0                            # leaves through y < 0 within one step
    vort = np.arange(pos.shape[0], dtype=np.float32)  # rides along: identifies each particle
    h, o32, _ = pair(n, 4, pos, vel, vort=vort, wall_level=-1e30)
    h.step(DT)
    o32.step(DT)
    mask_g = h.removed_mask()
    mask_o = o32.removed()
    np.testing.assert_array_equal(mask_g, mask_o)
    assert mask_o.sum() > 100
    g = h.get_particles()
    r = o32.get_particles()
    assert g["vort"].shape[0] == pos.shape[0] - int(mask_o.sum())
    np.testing.assert_array_equal(g["vort"].astype(np.int64), r["vort"].astype(np.int64))
    np.testing.assert_array_equal(g["vort"].astype(np.int64), O.remove_perm(mask_o))


def test_multistep_aggregates():
    n = 32
    pos, _ = make_scene(n, vel_scale=0.0)
    vel = np.zeros_like(pos)
    h, o32, _ = pair(n, 4, pos, vel)
    for _ in range(10):
        h.step(DT)
        o32.step(DT)
    g = h.get_particles()
    r = o32.get_particles()
    assert g["pos"].shape == r["pos"].shape
    ke_g = float((g["vel"].astype(np.float64) ** 2).sum())
    ke_r = float((r["vel"] ** 2).sum())
    assert abs(ke_g - ke_r) / ke_r < 1e-3
    np.testing.assert_allclose(g["pos"].mean(0), r["pos"].mean(0), rtol=1e-4)
    assert abs(g["pos"][:, 0].max() - r["pos"][:, 0].max()) < 1e-2


def test_block_size_invariance():
    # B=4 and B=8 give the same fluid set and physics up to summation order
    n = 32
    pos, vel = make_scene(n, seed=4)
    out = {}
    for B in (4, 8):
        g = L.grid_desc(n, n, n, (0, 0, 0), (1, 1, 1), B)
        h = L.FlipHandle(g, L.default_params(ext_force=(0, -9.8, 0), wall_level=1e30, voxel_size=1 / n,
                                             solver_mode=1, cg_rel_tol=1e-6, max_iterations=3000))
        h.set_particles(pos, vel)
        h.step(DT)
        out[B] = (h.get_particles(), h.stats())
    assert out[4][1]["num_fluid"] == out[8][1]["num_fluid"]
    np.testing.assert_array_equal(out[4][0]["adj"], out[8][0]["adj"])
    assert rel(out[4][0]["vel"], out[8][0]["vel"]) < 1e-4


def test_world_frame_api_matches_oracle():
    # advanceOneFrame in world space: toVoxel, step(params.dt), toWorld (FLIPMethod.cu:177-206)
    from vare_amd import FLIPMethod, Grid, Particles
    n = 24
    pos_v, vel_v = make_scene(n, seed=5, vel_scale=3.0)
    fm = FLIPMethod(block_size=4)
    fm.params.extForce = (0.0, -9.8, 0.0)
    fm.params.wallLevel = 10.0
    fm.params.voxelSize = 1.0 / n
    fm.params.dt = DT
    fm.initialize(Grid(n, n, n, (0, 0, 0), (1, 1, 1)))
    pts = Particles(position=(pos_v / n).astype(np.float32), velocity=(vel_v / n).astype(np.float32),
                    vorticity=np.zeros(pos_v.shape[0], np.float32), curvature=np.zeros(pos_v.shape[0], np.float32),
                    adjacency=np.zeros(pos_v.shape[0], np.float32))
    fm.advanceOneFrame(pts)
    o = O.Oracle((n, n, n), B=4, prm=O.params(ext_force=(0, -9.8, 0), wall_level=10.0, voxel_size=1 / n))
    # the oracle steps in voxel space: apply Grid::voxelPoint/voxelVector (Grid.h:445-456) in FP32
    s = np.float32(1.0) / np.float32(n)
    o.set_particles((pos_v / n).astype(np.float32) / s, (vel_v / n).astype(np.float32) / s)
    o.step(DT)
    r = o.get_particles()
    assert pts.position.shape[0] == r["pos"].shape[0]
    assert rel(pts.position, r["pos"] / n) < 1e-5
    assert rel(pts.velocity, r["vel"] / n) < 1e-4


def test_empty_and_error_paths():
    n = 16
    g = L.grid_desc(n, n, n, block_size=4)
    h = L.FlipHandle(g, L.default_params(voxel_size=1 / n))
    h.step(DT)                                                    # zero particles
    assert h.stats()["num_particles"] == 0
    bad = np.array([[1.0, np.nan, 2.0]], np.float32)
    with pytest.raises(L.VareError) as e:
        h.set_particles(bad, np.zeros_like(bad))
    assert e.value.code == L.VARE_ENONFINITE
    # more than 2^30 particles is refused before any input is read
    one = np.zeros((1, 3), np.float32)
    rc = h.lib.vare_flip_set_particles(h.h, L.ptr(one), L.ptr(one), None, None, None, (1 << 30) + 1,
                                       L.SPACE_VOXEL, L.MEM_HOST)
    assert rc == L.VARE_ENOMEM
    assert h.num_particles() == 0
    pos, vel = make_scene(n, seed=6)
    h2 = L.FlipHandle(g, L.default_params(voxel_size=1 / n, ext_force=(0, -9.8, 0), wall_level=1e30,
                                          solver_mode=1, cg_rel_tol=1e-9, max_iterations=2))
    h2.set_particles(pos, vel)
    rc = h2.step_rc(DT)
    assert rc == L.VARE_ENOCONV
    assert h2.stats()["cg_iterations"] == 2


def test_mid_size_properties():
    # 128^3 dam break (B=8, relative PCG): size-independent invariants of one step
    n = 128
    pos = L.seed_box(n, n, n, 0.5 * n, 0.5 * n, n, 8, 1)
    g = L.grid_desc(n, n, n, block_size=8)
    h = L.FlipHandle(g, L.default_params(ext_force=(0, -9.8, 0), wall_level=1e30, voxel_size=1 / n,
                                         solver_mode=1, cg_rel_tol=1e-5, max_iterations=4000))
    h.set_particles(pos, np.zeros_like(pos))
    h.step(DT)
    h.step(DT)
    st = h.stats()
    assert st["converged"] == 1 and st["residual_norm"] <= 1e-5 * st["b_norm"] * 1.0001
    assert st["num_particles"] == pos.shape[0] and st["num_removed"] == 0
    d = h.get_particles()
    assert np.all(d["pos"] >= 0) and np.all(d["pos"] <= n)
    assert np.all(np.isfinite(d["vel"]))
    cnt = h.cell_field(L.FIELD_COUNT)
    assert cnt.sum() >= 8 * pos.shape[0] * 0.99 and np.all(cnt == np.round(cnt))


@pytest.mark.parametrize("B,nu", [(8, 2), (8, 1), (4, 2)])
def test_multigrid_pcg_solution(B, nu):
    # MG-preconditioned CG solves the same system: pressure and projected velocity agree with the
    # FP64 oracle's converged solve, in far fewer iterations than unpreconditioned CG.
    n = 48
    pos, vel = make_scene(n, wx=0.5, wy=0.5, seed=7)
    base = dict(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / n, solver_mode=1,
                cg_rel_tol=1e-6, max_iterations=4000)
    g = L.grid_desc(n, n, n, (0, 0, 0), (1, 1, 1), B)
    h = L.FlipHandle(g, L.default_params(preconditioner=L.PRECOND_MG, mg_sweeps=nu, **base))
    h.set_particles(pos, vel)
    o64 = O.Oracle((n, n, n), B=B, prm=O.params(**dict(base, cg_rel_tol=1e-10)), fp64=True)
    o64.set_particles(pos, vel)
    for s in range(4):
        h.stage(s, DT)
        o64.stage(s, DT)
    st = h.stats()
    print("mg B=%d nu=%d iters=%d (oracle CG to 1e-10: %d)" % (B, nu, st["cg_iterations"], o64.stats()["cg_iters"]),
          flush=True)
    assert st["converged"] == 1
    assert st["cg_iterations"] < 60
    assert st["residual_norm"] <= 1e-6 * st["b_norm"] * 1.0001
    assert rel(h.cell_field(L.FIELD_PRESSURE), o64.field(O.FIELD_PRES)) < 2e-5
    assert rel(h.cell_field(L.FIELD_COMP), o64.field(O.FIELD_COMP)) < 2e-5


def test_step_bit_reproducible():
    # gather P2G, fixed-order reductions and gather-form restriction: two runs of whole steps give
    # identical bits (pressure, particles, iteration counts)
    n = 32
    pos, vel = make_scene(n, seed=8)
    outs = []
    for _ in range(2):
        g = L.grid_desc(n, n, n, block_size=8)
        h = L.FlipHandle(g, L.default_params(ext_force=(0, -9.8, 0), wall_level=1e30, voxel_size=1 / n,
                                             solver_mode=1, cg_rel_tol=1e-6, preconditioner=L.PRECOND_MG))
        h.set_particles(pos, vel)
        iters = []
        for _ in range(3):
            h.step(DT)
            iters.append(h.stats()["cg_iterations"])
        pt = h.get_particles()
        outs.append((h.cell_field(L.FIELD_PRESSURE), pt["pos"], pt["vel"], iters))
    for a, b in zip(outs[0][:3], outs[1][:3]):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    assert outs[0][3] == outs[1][3]


@pytest.mark.parametrize("B,fp16", [(8, 0), (8, 1), (4, 1)])
def test_step_equals_stages(B, fp16):
    # vare_flip_step moves the particles into bin order inside P2G's first pass (and the outside-domain
    # bin with a small strided kernel); the stage API moves them in the binning stage with the separate
    # reorder pass.  Whole steps both ways give identical bits, with 2% of the particles starting above
    # the domain (the outside bin; with the wall level at half height they are not clamped back and are
    # removed) and with FP16 velocities.
    n = 40
    pos, vel = make_scene(n, seed=21)
    sel = np.random.default_rng(5).random(len(pos)) < 0.02
    pos[sel, 1] += np.float32(n + 2)
    outs = []
    for staged in (False, True):
        g = L.grid_desc(n, n, n, block_size=B)
        h = L.FlipHandle(g, L.default_params(ext_force=(0, -9.8, 0), wall_level=0.5, voxel_size=1 / n,
                                             solver_mode=1, cg_rel_tol=1e-3, refine_tol=1e-5,
                                             preconditioner=L.PRECOND_MG, velocity_fp16=fp16))
        h.set_particles(pos, vel)
        removed = []
        for _ in range(3):
            if staged:
                for s in range(6):
                    h.stage(s, DT)
            else:
                h.step(DT)
            removed.append(h.num_particles())
        pt = h.get_particles()
        outs.append((pt["pos"], pt["vel"], pt["adj"], h.cell_field(L.FIELD_PRESSURE), removed))
        h.close()
    print("particles after each step:", outs[0][4], flush=True)
    assert outs[0][4] == outs[1][4] and outs[0][4][0] < len(pos)
    for a, b in zip(outs[0][:4], outs[1][:4]):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def cpu_binning(pos, pointers, B, f, n):
    """Stable counting sort of particle indices by bin: the storage index of the P2G base cell
    max((int)(p-0.5), 0) for particles inside [0,n]^3, Nc (= active cells) for the rest."""
    fx, fy, fz = f
    nc = int((pointers >= 0).sum()) * B ** 3
    inside = np.all((pos >= 0) & (pos <= np.float32(n)), axis=1)
    base = np.maximum(np.trunc(pos - np.float32(0.5)).astype(np.int64), 0)
    base = np.where(inside[:, None], base, 0)
    blk = base // B
    loc = base % B
    lin = blk[:, 0] + fx * (blk[:, 1] + fy * blk[:, 2])
    ptr = pointers[lin]
    key = np.where(inside & (ptr >= 0), ptr + loc[:, 0] + B * (loc[:, 1] + B * loc[:, 2]), nc)
    perm = np.argsort(key, kind="stable").astype(np.int32)
    bstart = np.searchsorted(key[perm], np.arange(nc + 2)).astype(np.int32)
    return perm, bstart


@pytest.mark.parametrize("B", [4, 8])
def test_binning_permutation_bit_exact(B):
    n = 32
    pos, vel = make_scene(n, seed=9)
    rng = np.random.default_rng(9)
    shuffle = rng.permutation(pos.shape[0])           # arbitrary input order
    pos, vel = pos[shuffle], vel[shuffle]
    pos[:50] += np.float32(40.0)                      # some particles outside the domain (bin Nc)
    pos[50:90] = np.float32([5.7, 3.2, 9.9])          # one bin longer than the register sort
    pos[90:93] = np.float32([0.2, 0.3, 0.1])          # base cell clamped to 0
    pos[93:96] = np.float32(n)                        # on the far faces (still inside)
    h, o32, _ = pair(n, B, pos, vel)
    h.stage(0)
    pointers, _ = h.block_table()
    perm, bstart = h.binning()
    cperm, cbstart = cpu_binning(pos, pointers, B, (n // B,) * 3, n)
    np.testing.assert_array_equal(bstart, cbstart)
    np.testing.assert_array_equal(perm, cperm)


@pytest.mark.parametrize("B", [4, 8])
def test_p2g_bit_reproducible_and_clustered(B):
    # the gather P2G sums in a fixed order: two runs agree bit for bit, also with a crowded cell
    n = 32
    pos, vel = make_scene(n, seed=12)
    pos[:3000] = np.float32([7.3, 4.6, 11.2]) + np.random.default_rng(3).random((3000, 3), np.float32) * 0.4
    out = []
    for _ in range(2):
        h, o32, _ = pair(n, B, pos, vel)
        h.stage(0)
        h.stage(1)
        out.append((h.cell_field(L.FIELD_MASS), h.cell_field(L.FIELD_VELOCITY), h.cell_field(L.FIELD_COUNT)))
    o32.stage(0)
    o32.stage(1)
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    np.testing.assert_array_equal(out[0][2], o32.field(O.FIELD_COUNT))
    assert rel(out[0][0], o32.field(O.FIELD_MASS)) < 1e-6
    assert rel(out[0][1], o32.field(O.FIELD_VELO)) < 1e-6


@pytest.mark.parametrize("B", [4, 8])
def test_binned_path_matches_direct_path(B):
    # the LDS-tiled P2G/G2P (particle_path 0) and the per-particle path (1) compute the same step
    n = 32
    pos, vel = make_scene(n, seed=10)
    out = []
    for path in (0, 1):
        g = L.grid_desc(n, n, n, block_size=B)
        h = L.FlipHandle(g, L.default_params(ext_force=(0, -9.8, 0), wall_level=1e30, voxel_size=1 / n,
                                             particle_path=path))
        h.set_particles(pos, vel)
        h.stage(0); h.stage(1)
        fields = [h.cell_field(L.FIELD_MASS), h.cell_field(L.FIELD_VELOCITY), h.cell_field(L.FIELD_COUNT)]
        for s in (2, 3, 4, 5):
            h.stage(s, DT)
        out.append((fields, h.get_particles(), h.stats()))
    (f0, p0, s0), (f1, p1, s1) = out
    np.testing.assert_array_equal(f0[2], f1[2])
    assert rel(f0[0], f1[0]) < 1e-6 and rel(f0[1], f1[1]) < 1e-6
    np.testing.assert_array_equal(p0["adj"], p1["adj"])
    assert rel(p0["vel"], p1["vel"]) < 1e-5 and rel(p0["pos"], p1["pos"]) < 1e-7


def test_iterative_refinement_fp64():
    # FP32 MG-PCG (inner 1e-4) + FP64 residual recompute: the FP64 residual of the refined pressure
    # reaches 1e-8 relative; the FP32 pressure handed to the gradient matches the FP64 oracle's
    # converged solve to FP32 rounding.
    n = 48
    pos, vel = make_scene(n, wx=0.5, wy=0.5, seed=11)
    base = dict(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / n, solver_mode=1)
    g = L.grid_desc(n, n, n, (0, 0, 0), (1, 1, 1), 8)
    h = L.FlipHandle(g, L.default_params(preconditioner=L.PRECOND_MG, cg_rel_tol=1e-4, refine_tol=1e-8,
                                         max_iterations=500, **base))
    h.set_particles(pos, vel)
    o64 = O.Oracle((n, n, n), B=8, prm=O.params(cg_rel_tol=1e-12, max_iterations=10000, **base), fp64=True)
    o64.set_particles(pos, vel)
    for s in range(4):
        h.stage(s, DT)
        o64.stage(s, DT)
    st = h.stats()
    print("refinement steps=%d inner iters=%d fp64 rel residual=%.3g (fp32 pressure: %.3g)" % (
        st["refine_steps"], st["cg_iterations"], st["residual_norm"] / st["b_norm"], st["true_rel_residual"]),
        flush=True)
    assert st["converged"] == 1 and st["refine_steps"] >= 2
    assert st["residual_norm"] <= 1e-8 * st["b_norm"]
    assert rel(h.cell_field(L.FIELD_PRESSURE), o64.field(O.FIELD_PRES)) < 1e-6
    assert rel(h.cell_field(L.FIELD_COMP), o64.field(O.FIELD_COMP)) < 1e-6


def test_fp16_velocity_storage():
    # particle velocities stored in FP16 (positions FP32): the step matches the FP32-storage step
    # to FP16 rounding of the velocities (unit roundoff 2^-11 ~ 4.9e-4), and the FP16 values come
    # back exactly representable
    n = 32
    pos, vel = make_scene(n, seed=12)
    out = {}
    for half in (0, 1):
        g = L.grid_desc(n, n, n, block_size=8)
        h = L.FlipHandle(g, L.default_params(ext_force=(0, -9.8, 0), wall_level=1e30, voxel_size=1 / n,
                                             solver_mode=1, cg_rel_tol=1e-6, preconditioner=L.PRECOND_MG,
                                             velocity_fp16=half))
        h.set_particles(pos, vel)
        for _ in range(3):
            h.step(DT)
        out[half] = h.get_particles()
    v16 = out[1]["vel"]
    np.testing.assert_array_equal(v16, v16.astype(np.float16).astype(np.float32))
    assert rel(v16, out[0]["vel"]) < 2e-3
    assert rel(out[1]["pos"], out[0]["pos"]) < 1e-5


def test_multigrid_deep_hierarchy_matches_jacobi():
    # 160^3 with 8^3 blocks: the edge-2 and edge-1 levels are too large for the single-workgroup
    # bottom V-cycle, which takes the two dense levels (10^3, 5^3) below them.  The multigrid
    # preconditioner must give the same pressure as Jacobi-PCG on the same step, in few iterations.
    n = 160
    pos, vel = make_scene(n, wx=0.5, wy=0.5, seed=11)
    g = L.grid_desc(n, n, n, (0, 0, 0), (1, 1, 1), 8)
    base = dict(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / n, solver_mode=1,
                cg_rel_tol=1e-7, max_iterations=20000)
    h0 = L.FlipHandle(g, L.default_params(preconditioner=L.PRECOND_MG, **base))
    h0.set_particles(pos, vel)
    for _ in range(3):                          # a few steps so the surface is no longer flat
        h0.step(DT)
    d = h0.get_particles()
    out = {}
    for name, pc in (("mg", L.PRECOND_MG), ("jacobi", L.PRECOND_JACOBI)):
        h = L.FlipHandle(g, L.default_params(preconditioner=pc, **base))
        h.set_particles(d["pos"], d["vel"])
        h.step(DT)
        st = h.stats()
        out[name] = (h.cell_field(L.FIELD_PRESSURE), st)
        print(name, "iterations", st["cg_iterations"], "converged", st["converged"], flush=True)
        assert st["converged"] == 1
    pm, sm = out["mg"]
    pj, sj = out["jacobi"]
    assert sm["cg_iterations"] < 40 < sj["cg_iterations"]
    assert rel(pm, pj) < 1e-5


def test_multigrid_smoothing_pairs():
    # the two sweeps of each smoothing pair: the default Chebyshev pair (the roots of the degree-2
    # Chebyshev polynomial on [0.4, 2]) and weighted Jacobi (0.8 twice, mg_omega_a = mg_omega_b = 0)
    # both precondition the same step to the same pressure, the Chebyshev pair in fewer iterations
    n = 96
    pos, vel = make_scene(n, wx=0.5, wy=0.5, seed=12, vel_scale=3.0)
    g = L.grid_desc(n, n, n, (0, 0, 0), (1, 1, 1), 8)
    base = dict(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / n, solver_mode=1,
                cg_rel_tol=1e-7, max_iterations=4000, preconditioner=L.PRECOND_MG)
    assert L.default_params().mg_omega_a == pytest.approx(1.576514) and L.default_params().mg_omega_b == pytest.approx(0.566352)
    out = {}
    for name, kw in (("chebyshev", {}), ("jacobi", dict(mg_omega_a=0.0, mg_omega_b=0.0))):
        h = L.FlipHandle(g, L.default_params(**base, **kw))
        h.set_particles(pos, vel)
        h.step(DT)
        st = h.stats()
        out[name] = (h.cell_field(L.FIELD_PRESSURE), st)
        print(name, "iterations", st["cg_iterations"], flush=True)
        assert st["converged"] == 1
    assert out["chebyshev"][1]["cg_iterations"] < out["jacobi"][1]["cg_iterations"]
    assert rel(out["chebyshev"][0], out["jacobi"][0]) < 1e-5
    with pytest.raises(L.VareError):
        L.FlipHandle(g, L.default_params(**base, mg_omega_a=2.5))


@pytest.mark.parametrize("nu", [1, 2])
def test_multigrid_solve_independent_of_earlier_solves(nu):
    # 160^3 with 8^3 blocks runs the edge-2 and edge-1 levels through the per-level kernels.  A
    # V-cycle must be a fixed linear operator: every value it reads is written earlier in the same
    # cycle, so a handle that first solved another scene reproduces a fresh handle's step bit for bit
    # (with one sweep per level this covers the fine cells under an air coarse cell, whose first-sweep
    # value the restriction stores for the prolongation).
    n = 160
    x_pos, x_vel = make_scene(n, wx=0.5, wy=0.5, seed=11)
    y_pos, y_vel = make_scene(n, wx=0.3, wy=0.7, seed=3)
    g = L.grid_desc(n, n, n, (0, 0, 0), (1, 1, 1), 8)
    prm = L.default_params(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / n, solver_mode=1,
                           cg_rel_tol=1e-6, max_iterations=4000, preconditioner=L.PRECOND_MG, mg_sweeps=nu)
    fresh = L.FlipHandle(g, prm)
    fresh.set_particles(x_pos, x_vel)
    fresh.step(DT)
    used = L.FlipHandle(g, prm)
    used.set_particles(y_pos, y_vel)
    used.step(DT)
    used.set_particles(x_pos, x_vel)
    used.step(DT)
    sa, sb = fresh.stats(), used.stats()
    print("nu=%d iterations %d / %d" % (nu, sa["cg_iterations"], sb["cg_iterations"]), flush=True)
    assert sa["converged"] == 1 and sb["converged"] == 1
    assert sa["cg_iterations"] == sb["cg_iterations"]
    assert np.array_equal(fresh.cell_field(L.FIELD_PRESSURE), used.cell_field(L.FIELD_PRESSURE))
    assert np.array_equal(fresh.get_particles()["vel"], used.get_particles()["vel"])
