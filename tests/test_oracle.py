"""Pins of the CPU oracle against closed forms, hand-derived known answers, invariants and brute force.

The reference ships no test or golden vector for any stage of the FLIP path (its unit tests cover
Array/Vector/Json only, base/unittest/test.cpp:16-27), so each stage is pinned here by mathematics
chosen so that a plausible restatement mistake (dropped term, wrong sign/index, transposed operand)
fails at least one check.
"""
import numpy as np
import pytest

from oracle import oracle as O


def dam_params(n, **kw):
    base = dict(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / n)
    base.update(kw)
    return O.params(**base)


# ---------------------------------------------------------------- A15 Rand (Random.h:18-27)
def test_rand_zero_seed_point():
    # (seed ^ 12345391) == 0  =>  every later step keeps 0  =>  Rand == 0 exactly
    assert O.rand(12345391) == 0.0


def test_rand_range_and_uniformity():
    v = np.array([O.rand(s) for s in range(20000)], dtype=np.float64)
    assert v.min() >= 0.0 and v.max() <= 1.0
    hist, _ = np.histogram(v, bins=10, range=(0, 1))
    assert abs(v.mean() - 0.5) < 0.01
    assert hist.min() > 1800 and hist.max() < 2200


def test_rand_golden(oracle_mod):
    import json, os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "rand_kat.json")))
    for s, expect in zip(g["seeds"], g["values"]):
        assert O.rand(s) == np.float32(expect)


# ---------------------------------------------------------------- A2 frame geometry
def test_voxel_size_truncation_kat():
    # Grid.h:145-147: nx = (size_t)(float(max-min) / voxelSize), float division then truncation:
    # 1.3f/0.1f = 12.999999f -> 12;  0.9f/0.3f = 2.9999998f -> 2;  1.0f/0.1f rounds to 10.0f -> 10
    assert list(O.grid_from_voxel_size(0.1, (0, 0, 0), (1.3, 1.0, 1.0))) == [12, 10, 10]
    assert list(O.grid_from_voxel_size(0.3, (0, 0, 0), (0.9, 0.9, 0.9))) == [2, 2, 2]
    assert list(O.grid_from_voxel_size(1.0 / 64, (0, 0, 0), (1, 1, 1))) == [64, 64, 64]


def test_frame_padding_and_recentre():
    # SparseFrame.cu:81-103: 10 cells, B=4 -> 3 blocks = 12 cells, h = 0.1, centred on 0.5
    g = O.geom((10, 10, 10), (0, 0, 0), (1, 1, 1), 4)
    assert list(g["f"]) == [3, 3, 3]
    assert list(g["dims"]) == [12, 12, 12]
    np.testing.assert_allclose(g["translate"], [-0.1, -0.1, -0.1], atol=1e-6)
    np.testing.assert_allclose(g["scale"], [0.1, 0.1, 0.1], atol=1e-7)
    # anisotropic cell sizes use h = min(dx,dy,dz)
    g = O.geom((8, 4, 8), (0, 0, 0), (1, 1, 2), 4)
    assert list(g["dims"]) == [8, 4, 8]
    np.testing.assert_allclose(g["scale"], [0.125, 0.125, 0.125], atol=1e-7)
    np.testing.assert_allclose(g["translate"], [0.0, 0.25, 0.5], atol=1e-6)


# ---------------------------------------------------------------- A3 binning / activation
def test_block_activation_aliasing_kat():
    # hand-derived known answer: p = (0.5, 5, 5), B = 4  ->  c = (0, 1, 1); the i = -1 plane
    # aliases into x = fx-1 of the previous row/slab (int index arithmetic), (-1,0,0) is dropped.
    n = 32
    o = O.Oracle((n, n, n), B=4, prm=dam_params(n))
    o.set_particles(np.array([[0.5, 5.0, 5.0]], np.float32), np.zeros((1, 3), np.float32))
    o.stage(0)
    fx = fy = fz = n // 4
    m = o.markers().reshape(fz, fy, fx)
    active = {(i, j, k) for k in range(fz) for j in range(fy) for i in range(fx) if m[k, j, i] > 0}
    direct = {(i, j, k) for i in (0, 1) for j in (0, 1, 2) for k in (0, 1, 2)}
    aliased = {(fx - 1, 0, 0), (fx - 1, 1, 0), (fx - 1, fy - 1, 0), (fx - 1, 0, 1), (fx - 1, 1, 1),
               (fx - 1, fy - 1, 1), (fx - 1, 0, 2), (fx - 1, 1, 2)}
    assert active == direct | aliased
    # each direct block is marked once
    assert all(m[k, j, i] == 1 for (i, j, k) in direct)


def _brute_active(pos, f, B):
    """Brute-force activation: signed linear index of every 3x3x3 neighbour, dropped when < 0."""
    fx, fy, fz = f
    mark = np.zeros(fx * fy * fz, np.int64)
    for p in pos:
        c = [max(int(np.float32(p[a])) // B if int(np.float32(p[a])) >= 0 else -((-int(np.float32(p[a]))) // B), 0)
             for a in range(3)]
        for k in range(c[2] - 1, min(c[2] + 1, fz - 1) + 1):
            for j in range(c[1] - 1, min(c[1] + 1, fy - 1) + 1):
                for i in range(c[0] - 1, min(c[0] + 1, fx - 1) + 1):
                    t = k * fx * fy + j * fx + i
                    if t >= 0:
                        mark[t] += 1
    return mark


@pytest.mark.parametrize("B", [4, 8])
def test_block_activation_brute_force(B):
    rng = np.random.default_rng(7)
    n = 32
    pos = (rng.random((300, 3)) * n).astype(np.float32)
    pos[:20, 0] = rng.random(20) * 2          # near the low faces: aliasing cases
    pos[20:40, 1] = rng.random(20) * 2
    o = O.Oracle((n, n, n), B=B, prm=dam_params(n))
    o.set_particles(pos, np.zeros_like(pos))
    o.stage(0)
    f = (n // B,) * 3
    np.testing.assert_array_equal(o.markers(), _brute_active(pos, f, B))
    pointers, coords = o.block_table()
    act = np.nonzero(o.markers() > 0)[0]
    # pointers strictly increasing by B^3 over active blocks in ascending linear order
    np.testing.assert_array_equal(pointers[act], np.arange(act.size) * B ** 3)
    assert np.all(pointers[o.markers() == 0] == -1)
    lin = coords[:, 0] + f[0] * (coords[:, 1] + f[1] * coords[:, 2])
    np.testing.assert_array_equal(lin, act)
    # every occupied block is active
    occ = (pos.astype(np.int32) // B)
    occ_lin = np.unique(occ[:, 0] + f[0] * (occ[:, 1] + f[1] * occ[:, 2]))
    assert np.all(o.markers()[occ_lin] > 0)


# ---------------------------------------------------------------- A5/A6 P2G and count field
def _single(n, p, v, B=4):
    o = O.Oracle((n, n, n), B=B, prm=dam_params(n))
    o.set_particles(np.array([p], np.float32), np.array([v], np.float32))
    o.stage(0)
    o.stage(1)
    return o


def _cell_lookup(o, B):
    pointers, coords = o.block_table()
    f = o.geom["f"]

    def idx(i, j, k):
        t = (k // B) * f[0] * f[1] + (j // B) * f[0] + i // B
        s = pointers[t]
        return None if s < 0 else s + (k % B) * B * B + (j % B) * B + i % B
    return idx


def test_p2g_single_particle_closed_form():
    n, B = 16, 4
    p = (5.3, 6.8, 7.55)
    v = (1.5, -2.0, 0.25)
    o = _single(n, p, v, B)
    idx = _cell_lookup(o, B)
    mass = o.field(O.FIELD_MASS)
    velo = o.field(O.FIELD_VELO)
    cnt = o.field(O.FIELD_COUNT)
    total = 0.0
    for k in (7, 8):
        for j in (6, 7):
            for i in (4, 5):
                w = 1.0
                for a, c in zip(p, (i, j, k)):
                    w *= max(0.0, 1.0 - abs(a - (c + 0.5)))
                c = idx(i, j, k)
                assert mass[c] == pytest.approx(w + 1e-5, abs=2e-7)
                np.testing.assert_allclose(velo[c], v, rtol=1e-6)      # uniform velocity reproduced
                assert cnt[c] == 1.0
                total += mass[c]
    assert total == pytest.approx(1.0 + 8e-5, abs=1e-6)                 # partition of unity
    assert np.count_nonzero(cnt) == 8 and cnt.sum() == 8.0


def test_p2g_low_face_truncation_kat():
    # (int)(p-0.5) truncates toward 0: p.x = 0.3 -> cells 0 and 1; cell 1's x-weight is 0,
    # so it receives only the 1e-5 floor (and still becomes fluid).
    n, B = 16, 4
    o = _single(n, (0.3, 5.5, 5.5), (0, 0, 0), B)
    idx = _cell_lookup(o, B)
    mass = o.field(O.FIELD_MASS)
    assert mass[idx(1, 5, 5)] == pytest.approx(1e-5, rel=1e-6)
    # (5.5 is a cell centre: y/z weights 1 for cells 5 and 0 for cells 6)
    assert mass[idx(0, 5, 5)] == pytest.approx(0.8 + 1e-5, rel=1e-6)
    assert mass[idx(0, 6, 5)] == pytest.approx(1e-5, rel=1e-5)
    mark = o.field(O.FIELD_MARK)
    assert mark[idx(1, 5, 5)] == 4 and mark[idx(1, 6, 6)] == 4


def test_p2g_momentum_conservation():
    rng = np.random.default_rng(3)
    n, B = 16, 4
    pos = (2 + rng.random((500, 3)) * 10).astype(np.float32)
    vel = rng.standard_normal((500, 3)).astype(np.float32)
    o = O.Oracle((n, n, n), B=B, prm=dam_params(n), fp64=True)
    o.set_particles(pos, vel)
    o.stage(0)
    o.stage(1)
    mass = o.field(O.FIELD_MASS)
    velo = o.field(O.FIELD_VELO)
    # sum of weights times particle velocity, computed independently
    expect_m = 0.0
    expect_mv = np.zeros(3)
    for p, v in zip(pos.astype(np.float64), vel.astype(np.float64)):
        c0 = np.maximum(np.floor(p - 0.5).astype(int), 0)
        for dk in (0, 1):
            for dj in (0, 1):
                for di in (0, 1):
                    c = c0 + (di, dj, dk)
                    w = np.prod(np.maximum(0.0, 1.0 - np.abs(p - (c + 0.5)))) + 1e-5
                    expect_m += w
                    expect_mv += w * v
    assert mass.sum() == pytest.approx(expect_m, rel=1e-12)
    np.testing.assert_allclose((mass[:, None] * velo).sum(0), expect_mv, rtol=1e-10, atol=1e-9)


# ---------------------------------------------------------------- A8 divergence / system
def _scene(n, B=4, fp64=False, wx=0.5, wy=0.5, wz=1.0, ppc=8, **kw):
    pos = O.seed_box(n, n, n, wx * n, wy * n, wz * n, ppc, 1)
    o = O.Oracle((n, n, n), B=B, prm=dam_params(n, **kw), fp64=fp64)
    o.set_particles(pos, np.zeros_like(pos))
    return o, pos


def test_divergence_linear_field_closed_form():
    # comp = (x, 0, 0)  =>  b = -0.5*((x+1) - (x-1)) = -1 at interior fluid cells
    n, B = 16, 4
    o, _ = _scene(n, B, fp64=True)
    o.stage(0)
    o.stage(1)
    pointers, coords = o.block_table()
    nc = coords.shape[0] * B ** 3
    loc = np.arange(nc) % B ** 3
    blk = coords[np.arange(nc) // B ** 3]
    ijk = blk * B + np.stack([loc % B, (loc // B) % B, loc // (B * B)], 1)
    comp = np.zeros((nc, 3))
    comp[:, 0] = ijk[:, 0] + 0.5
    o.set_field(O.FIELD_COMP, comp)
    o.stage(2)
    b = o.field(O.FIELD_B)
    fluid = o.field(O.FIELD_FLUID).astype(np.int64)
    aval = o.field(O.FIELD_AVAL).reshape(-1, 7)
    interior = aval[:, 1:].sum(1) == -6          # all six neighbours fluid
    assert interior.sum() > 50
    np.testing.assert_allclose(b[interior], -1.0, atol=1e-12)


def test_system_symmetric_and_diagonal():
    n, B = 16, 4
    o, _ = _scene(n, B, fp64=True, wall_level=0.3)     # mix of wall and open faces
    o.step(1 / 240)                                     # give the fluid a velocity
    o.stage(0); o.stage(1); o.stage(2)
    aval = o.field(O.FIELD_AVAL).reshape(-1, 7)
    aind = o.field(O.FIELD_AIND).reshape(-1, 7).astype(np.int64)
    nf = aval.shape[0]
    A = np.zeros((nf, nf))
    for r in range(nf):
        for q in range(7):
            if aind[r, q] >= 0:
                A[r, aind[r, q]] += aval[r, q]
    np.testing.assert_array_equal(A, A.T)
    np.testing.assert_array_equal(aind[:, 0], np.arange(nf))
    # diagonal = 6 + nAir - nSolidOrWall in {0..12}; off-diagonals are -1 per fluid neighbour
    offd = (aind[:, 1:] >= 0).sum(1)
    np.testing.assert_array_equal(-aval[:, 1:].sum(1), offd)
    assert np.all((aval[:, 0] >= 0) & (aval[:, 0] <= 12))
    # with open sides the system is positive definite
    assert np.linalg.eigvalsh(A).min() > 0


# ---------------------------------------------------------------- A9 CG / PCG
def _system(o):
    aval = o.field(O.FIELD_AVAL).reshape(-1, 7)
    aind = o.field(O.FIELD_AIND).reshape(-1, 7).astype(np.int64)
    nf = aval.shape[0]
    A = np.zeros((nf, nf))
    for r in range(nf):
        for q in range(7):
            if aind[r, q] >= 0:
                A[r, aind[r, q]] += aval[r, q]
    return A, o.field(O.FIELD_B)


@pytest.mark.parametrize("precond", [0, 1])
def test_pcg_matches_dense_solve(precond):
    n, B = 12, 4
    o, _ = _scene(n, B, fp64=True, wx=0.6, wy=0.6, solver_mode=1, cg_rel_tol=1e-12,
                  max_iterations=5000, preconditioner=precond)
    o.step(1 / 240)
    o.stage(0); o.stage(1); o.stage(2)
    A, b = _system(o)
    x_ref = np.linalg.solve(A, b)
    o.stage(3)
    x = o.field(O.FIELD_X)
    np.testing.assert_allclose(x, x_ref, rtol=1e-9, atol=1e-9 * np.abs(x_ref).max())


def test_reference_cg_stopping_rule():
    # LinearSystemSolver.cpp:27-46: stop at |r| < 1e-3 returning i, or after max+2 products
    # returning max+1.  A cap of 3 must return 4 when the tolerance is not met.
    n, B = 16, 4
    o, _ = _scene(n, B, proj_iteration=3)
    o.step(1 / 240)
    it = o.step(1 / 240)
    assert it == 4
    o2, _ = _scene(n, B, proj_iteration=200)
    o2.step(1 / 240)
    it2 = o2.step(1 / 240)
    s = o2.stats()
    assert it2 < 200 and s["cg_residual"] < 1e-3


# ---------------------------------------------------------------- A12 G2P / advection
def test_g2p_uniform_field_and_advection():
    # all particles moving with the same velocity in a closed box: the grid velocity equals the
    # particle velocity everywhere it is sampled, so FLIP leaves it unchanged before gravity
    n, B = 16, 4
    rng = np.random.default_rng(5)
    pos = (4 + rng.random((2000, 3)) * 8).astype(np.float32)
    v = np.array([3.0, -1.0, 2.0], np.float32)
    o = O.Oracle((n, n, n), B=B, prm=dam_params(n, ext_force=(0, 0, 0)), fp64=True)
    o.set_particles(pos, np.tile(v, (2000, 1)))
    o.stage(0); o.stage(1)
    o.stage(2)
    # skip the projection: comp = velo (stage 3 would modify comp); run G2P directly
    o.set_field(O.FIELD_COMP, o.field(O.FIELD_VELO))
    dt = 1 / 240
    o.stage(4, dt)
    d = o.get_particles()
    np.testing.assert_allclose(d["vel"], np.tile(v, (2000, 1)), rtol=1e-9)
    np.testing.assert_allclose(d["pos"], pos.astype(np.float64) + v * dt, rtol=0, atol=1e-6)


def test_wall_clamp_kat():
    # FLIPMethod.cu:135-144: below the wall level positions clamp to [0, n] and the outward
    # velocity component is zeroed; gravity is added after the clamp.
    n, B = 16, 4
    o = O.Oracle((n, n, n), B=B, prm=dam_params(n, ext_force=(0, 0, 0)), fp64=True)
    pos = np.array([[0.01, 8.0, 8.0], [8.0, 15.995, 8.0]], np.float32)
    vel = np.array([[-100.0, 0.0, 0.0], [0.0, 100.0, 0.0]], np.float32)
    o.set_particles(pos, vel)
    o.stage(0); o.stage(1); o.stage(2)
    o.set_field(O.FIELD_COMP, o.field(O.FIELD_VELO))
    o.stage(4, 1 / 240)
    d = o.get_particles()
    assert d["pos"][0, 0] == 0.0 and d["vel"][0, 0] == 0.0
    assert d["pos"][1, 1] == 16.0 and d["vel"][1, 1] == 0.0
    assert np.all(o.removed() == 0)


# ---------------------------------------------------------------- A14 removal permutation
def _remove_closed_form(mask):
    """k-th removed slot below newN (ascending) <- k-th survivor at or above newN (descending)."""
    n = mask.size
    newn = n - int(mask.sum())
    out = np.arange(newn)
    holes = np.nonzero(mask[:newn])[0]
    surv = np.nonzero(mask[newn:] == 0)[0][::-1] + newn
    assert holes.size == surv.size
    out[holes] = surv
    return out


@pytest.mark.parametrize("seed", range(6))
def test_remove_permutation_closed_form(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 400))
    mask = (rng.random(n) < rng.random()).astype(np.uint8)
    np.testing.assert_array_equal(O.remove_perm(mask), _remove_closed_form(mask))


def test_remove_permutation_edges():
    for mask in ([1], [0], [1, 1, 1], [0, 0, 1], [1, 0, 0], [1, 0, 1, 0, 1, 0]):
        m = np.array(mask, np.uint8)
        np.testing.assert_array_equal(O.remove_perm(m), _remove_closed_form(m))


# ---------------------------------------------------------------- FP32 vs FP64 instances
def test_fp32_fp64_instances_agree_one_step():
    n, B = 24, 4
    o32, pos = _scene(n, B, wx=0.4, wy=0.6)
    o64, _ = _scene(n, B, fp64=True, wx=0.4, wy=0.6)
    for _ in range(3):
        o32.step(1 / 240)
        o64.step(1 / 240)
    a, b = o32.get_particles(), o64.get_particles()
    assert np.linalg.norm(a["vel"] - b["vel"]) / np.linalg.norm(b["vel"]) < 1e-4
    np.testing.assert_array_equal(a["adj"], b["adj"])


# ---------------------------------------------------------------- collision source (CollisionSource.h)
def _box_sdf(p, lo, hi):
    """Exact box signed distance in float64 (negative inside); corners as the FP32 values passed."""
    lo = np.asarray(lo, np.float32).astype(np.float64)
    hi = np.asarray(hi, np.float32).astype(np.float64)
    c, e = (lo + hi) / 2, (hi - lo) / 2
    q = np.abs(np.asarray(p, np.float64) - c) - e
    return np.linalg.norm(np.maximum(q, 0), axis=-1) + np.minimum(q.max(-1), 0)


def _storage_ijk(pointers, f, B):
    """(i,j,k) of every cell of a sparse field in storage order."""
    act = np.nonzero(pointers >= 0)[0]
    order = np.argsort(pointers[act])
    blk = act[order]
    bi, bj, bk = blk % f[0], (blk // f[0]) % f[1], blk // (f[0] * f[1])
    l = np.arange(B ** 3)
    li, lj, lk = l % B, (l // B) % B, l // (B * B)
    return np.stack([(bi[:, None] * B + li).ravel(), (bj[:, None] * B + lj).ravel(),
                     (bk[:, None] * B + lk).ravel()], 1)


@pytest.mark.parametrize("B", [4, 8])
def test_collision_box_field_closed_form(B):
    n, band = 32, 3.0
    lo, hi, v = (0.3, 0.1, 0.25), (0.55, 0.4, 0.7), (0.5, 0.0, -0.25)
    o = O.Oracle((n, n, n), B=B, prm=dam_params(n))
    o.set_collision_box(lo, hi, v, band)
    pointers, phi, vel, nrm, bg = o.collision()
    h = 1.0 / n
    f = o.geom["f"]
    assert bg == band
    # activation = 3x3x3 block neighbourhoods of the cells within `band` voxels of the surface
    ii = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    ctr = (ii + 0.5).astype(np.float32)
    sdf32 = _box_sdf(ctr * np.float32(h), lo, hi).astype(np.float32)
    seeds = ctr[np.abs(sdf32 / np.float32(h)) < band]
    np.testing.assert_array_equal(pointers >= 0, _brute_active(seeds, f, B) > 0)
    ijk = _storage_ijk(pointers, f, B)
    wp = (ijk + 0.5) * h
    np.testing.assert_allclose(phi, _box_sdf(wp, lo, hi) / h, rtol=0, atol=2e-5)
    np.testing.assert_allclose(vel, np.tile(np.array(v) / h, (phi.size, 1)), rtol=1e-6)
    # normals: unit gradient of the exact SDF
    lo64, hi64 = np.float32(lo).astype(np.float64), np.float32(hi).astype(np.float64)
    c = (lo64 + hi64) / 2
    e = (hi64 - lo64) / 2
    d = wp - c
    q = np.abs(d) - e
    out = q.max(1) > 0
    g = np.sign(d) * np.maximum(q, 0)
    g[out] /= np.linalg.norm(g[out], axis=1)[:, None]
    ax = np.argmax(q, 1)
    gi = np.zeros_like(g)
    gi[np.arange(len(ax)), ax] = np.where(d[np.arange(len(ax)), ax] < 0, -1.0, 1.0)
    g[~out] = gi[~out]
    np.testing.assert_allclose(nrm, g, atol=1e-6)


def _scene_particles(n, seed=4):
    pos = O.seed_box(n, n, n, 0.5 * n, 0.5 * n, n, 8, 1)
    vel = (np.random.default_rng(seed).standard_normal(pos.shape) * 5).astype(np.float32)
    return pos, vel


def test_collision_classification_and_velocity():
    # FLIPMethod.cu:242-260: solid = no mass and phi < 0 at the cell centre; where phi < 0 the
    # collision velocity times col_vel_scale is added to the P2G velocity (fluid cells included)
    n, B, scale = 16, 4, 0.5
    pos, vel = _scene_particles(n)
    vbox = np.array([0.3, -0.2, 0.1])
    with_col = O.Oracle((n, n, n), B=B, prm=dam_params(n, col_vel_scale=scale), fp64=True)
    no_col = O.Oracle((n, n, n), B=B, prm=dam_params(n, col_vel_scale=scale), fp64=True)
    with_col.set_collision_box((0.3, 0.05, 0.2), (0.8, 0.45, 0.6), vbox, 3.0)
    for o in (with_col, no_col):
        o.set_particles(pos, vel)
        o.stage(0)
        o.stage(1)
    pointers, _ = with_col.block_table()
    ijk = _storage_ijk(pointers, with_col.geom["f"], B)
    phi, _, _ = with_col.collision_query((ijk + 0.5).astype(np.float32))
    mass = with_col.field(O.FIELD_MASS)
    expect = np.where(mass > 0, 4, np.where(phi < 0, 1, 0))
    np.testing.assert_array_equal(with_col.field(O.FIELD_MARK), expect)
    assert (expect == 1).sum() > 50 and ((expect == 4) & (phi < 0)).sum() > 50
    dv = with_col.field(O.FIELD_VELO) - no_col.field(O.FIELD_VELO)
    np.testing.assert_allclose(dv, (phi < 0)[:, None] * (vbox * n * scale)[None, :], atol=1e-9)
    np.testing.assert_array_equal(with_col.field(O.FIELD_COMP), with_col.field(O.FIELD_VELO))


def test_collision_projection_and_push_kat():
    # FLIPMethod.cu:116-164 with zero grid velocity and no gravity: a particle inside the box is
    # projected along the (exact, axis-aligned) normal onto the face; a particle outside is left
    # alone; a moving box pushes an inside particle by v*dt first (here out of the box, so no
    # projection and no removal).
    n, B = 16, 4
    dt = 1 / 240
    o = O.Oracle((n, n, n), B=B, prm=dam_params(n, ext_force=(0, 0, 0)), fp64=True)
    pos = np.array([[5.3, 8.2, 7.9],       # 1.3 voxels inside the -x face (x = 4)
                    [2.0, 8.0, 8.0],       # outside
                    [8.1, 11.4, 7.6],      # 0.6 inside the +y face (y = 12)
                    [11.95, 8.3, 8.4]], np.float32)   # 0.05 inside +x; pushed out by the moving box
    o.set_collision_box((0.25, 0.25, 0.25), (0.75, 0.75, 0.75), (0.0, 0.0, 0.0), 3.0)
    o.set_particles(pos, np.zeros_like(pos))
    o.stage(0); o.stage(1); o.stage(2)
    o.set_field(O.FIELD_COMP, o.field(O.FIELD_VELO))
    o.stage(4, dt)
    p = o.get_particles()["pos"]
    assert abs(p[0, 0] - 4.0) < 1e-9 and p[0, 1] == np.float32(8.2) and p[0, 2] == np.float32(7.9)
    np.testing.assert_array_equal(p[1], pos[1].astype(np.float64))
    assert abs(p[2, 1] - 12.0) < 1e-9 and p[2, 0] == np.float32(8.1) and p[2, 2] == np.float32(7.6)
    assert abs(p[3, 0] - 12.0) < 1e-9
    phi_after, _, _ = o.collision_query(p.astype(np.float32))
    rem = o.removed()
    assert rem[1] == 0
    # a particle that collided is removed iff its final distance reads < 0 or > 3
    assert all(rem[q] == (phi_after[q] < 0 or phi_after[q] > 3) for q in (0, 2, 3))
    # the same box moving at +2 world units/s in x: the +x-face particle is pushed out first
    # (col_vel_scale 0 keeps the box velocity out of the grid velocity, so the particle's own
    # velocity stays 0 and the push is the only motion)
    o2 = O.Oracle((n, n, n), B=B, prm=dam_params(n, ext_force=(0, 0, 0), col_vel_scale=0.0), fp64=True)
    o2.set_collision_box((0.25, 0.25, 0.25), (0.75, 0.75, 0.75), (2.0, 0.0, 0.0), 3.0)
    o2.set_particles(pos[3:], np.zeros_like(pos[3:]))
    o2.stage(0); o2.stage(1); o2.stage(2)
    o2.set_field(O.FIELD_COMP, o2.field(O.FIELD_VELO))
    o2.stage(4, dt)
    p2 = o2.get_particles()["pos"]
    assert abs(p2[0, 0] - (np.float64(np.float32(11.95)) + 2.0 * n * np.float64(np.float32(dt)))) < 1e-9
    assert o2.removed()[0] == 0


def test_collision_sampler_callback_matches_box():
    # the general (seeds + sampler) entry reproduces the built-in box when fed the same seeds
    n, B = 16, 4
    lo, hi, v = (0.3, 0.2, 0.25), (0.6, 0.5, 0.7), (0.1, 0.2, 0.3)
    a = O.Oracle((n, n, n), B=B, prm=dam_params(n))
    a.set_collision_box(lo, hi, v, 2.0)
    h = np.float32(1.0 / n)
    ii = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    ctr = (ii + 0.5).astype(np.float32)
    wp = ctr * h
    sdf = _box_sdf(wp, lo, hi).astype(np.float32)
    seeds = wp[np.abs(sdf / h) < 2.0]
    b = O.Oracle((n, n, n), B=B, prm=dam_params(n))
    lo64, hi64 = np.float32(lo).astype(np.float64), np.float32(hi).astype(np.float64)
    c = 0.5 * (lo64 + hi64)
    e = 0.5 * (hi64 - lo64)

    def sample(p):
        d = p.astype(np.float64) - c
        q = np.abs(d) - e
        o = np.maximum(q, 0)
        out = np.sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2])
        qm = max(q[0], max(q[1], q[2]))
        s = np.float32(out + min(qm, 0.0))
        if out > 0:
            g = [np.float32((-o[x] if d[x] < 0 else o[x]) / out) for x in range(3)]
        else:
            ax = 0 if (q[0] >= q[1] and q[0] >= q[2]) else (1 if q[1] >= q[2] else 2)
            g = [0.0, 0.0, 0.0]
            g[ax] = -1.0 if d[ax] < 0 else 1.0
        return s, v, g

    b.set_collision(seeds, 2.0 / n, sample)
    ra, rb = a.collision(), b.collision()
    np.testing.assert_array_equal(ra[0], rb[0])
    for x, y in zip(ra[1:4], rb[1:4]):
        np.testing.assert_array_equal(x, y)
    assert ra[4] == rb[4]
