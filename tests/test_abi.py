"""CPU tests of the C ABI: the library loads, exports every symbol include/vare_flip.h declares,
and its host-side logic (parameter defaults, grid truncation, seeding hash, error paths that fail
before any device call) agrees with the reference / oracle.  No kernel is launched here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from vare_amd import _lib as L
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "vare_flip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(vare_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = header_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []
    # and the Python binding table covers them all
    assert sorted(L.SIGNATURES) == names


def test_params_default_mirror_reference():
    p = L.default_params()
    # FLIPMethod.h:53-66
    assert list(p.ext_force) == [0.0, pytest.approx(-9.80665), 0.0]
    assert p.max_substeps == 1 and p.pts_per_cell == 8 and p.proj_iteration == 200
    assert p.dt == pytest.approx(1 / 24) and p.wall_level == pytest.approx(-1e30)
    assert p.flip_ratio == pytest.approx(0.95) and p.cg_abs_tol == pytest.approx(1e-3)


def test_grid_from_voxel_size_matches_oracle():
    lib = L.load()
    for voxel, mx in [(0.1, (1.3, 1.0, 1.0)), (0.3, (0.9, 0.9, 0.9)), (1 / 64, (1, 1, 1)), (0.15, (0.7, 2.3, 3.3))]:
        g = L.GridDesc()
        arr = lambda v: (C.c_float * 3)(*v)
        L.check(lib.vare_grid_from_voxel_size(voxel, arr((0, 0, 0)), arr(mx), 4, C.byref(g)))
        assert [g.nx, g.ny, g.nz] == list(O.grid_from_voxel_size(voxel, (0, 0, 0), mx))


def test_rand_matches_oracle_and_golden():
    import json
    lib = L.load()
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "rand_kat.json")))
    for s, v in zip(g["seeds"], g["values"]):
        assert lib.vare_rand(s) == np.float32(v) == O.rand(s)


@pytest.mark.parametrize("dims,w,ppc", [((16, 16, 16), (6.4, 9.6, 16), 8), ((20, 12, 8), (10, 6, 8), 5)])
def test_seed_box_matches_oracle(dims, w, ppc):
    a = L.seed_box(*dims, *w, ppc, seed=3)
    b = O.seed_box(*dims, *w, ppc, seed=3)
    np.testing.assert_array_equal(a, b)
    assert np.all(a[:, 0] < w[0]) and np.all(a[:, 1] < w[1])


def test_seed_box_counts_config_sizes():
    # 256^3 dam break x<0.5, y<0.5, 8 ppc: 128*128*256 cells = 33,554,432 draws; exactly one draw
    # has Rand == 1.0f (float rounding of i/2^32) on the x=128 face and is dropped by the strict
    # region test
    n = C.c_size_t()
    L.check(L.load().vare_seed_box(256, 256, 256, 128.0, 128.0, 256.0, 8, 1, None, 0, C.byref(n)))
    assert n.value == 33554431


def test_create_rejects_bad_grid_before_device_use():
    lib = L.load()
    h = C.c_void_p()
    g = L.grid_desc(0, 8, 8)
    assert lib.vare_flip_create(C.byref(g), None, 0, 0, C.byref(h)) == L.VARE_EINVAL      # Grid.h:107
    g = L.grid_desc(8, 8, 8, block_size=5)
    assert lib.vare_flip_create(C.byref(g), None, 0, 0, C.byref(h)) == L.VARE_EINVAL
    p = L.default_params(enable_curl=1)
    g = L.grid_desc(8, 8, 8)
    assert lib.vare_flip_create(C.byref(g), C.byref(p), 0, 0, C.byref(h)) == L.VARE_ENOTIMPL
    assert b"curl" in lib.vare_last_error()


def test_seed_pool_sphere_region():
    n = 32
    pos = L.seed_pool_sphere(n, n, n, 0.3 * n, (0.5 * n, 0.7 * n, 0.5 * n), 0.2 * n, 4, seed=2)
    c = np.array([0.5, 0.7, 0.5]) * n
    inside = (pos[:, 1] < 0.3 * n) | (np.linalg.norm(pos - c, axis=1) < 0.2 * n)
    assert inside.all()
    # the pool part is exactly the box seeder's output for the same cells
    box = L.seed_box(n, n, n, float(n), 0.3 * n, float(n), 4, seed=2)
    pool = pos[pos[:, 1] < 0.3 * n]
    assert pool.shape[0] >= box.shape[0]
    assert {tuple(p) for p in box} <= {tuple(p) for p in pool}


def test_plan_bytes_counts_every_array_without_a_gpu():
    # vare_flip_plan_bytes runs the allocation sequence of create + first steps on a planning handle
    # (no device call); per particle it must count at least the 9 attribute arrays, their 7 reorder
    # buffers, the binning/removal arrays and the 7-float download staging (about 130 bytes)
    g = L.grid_desc(64, 64, 64, block_size=8)
    prm = L.default_params(solver_mode=1, preconditioner=L.PRECOND_MG)
    a = L.plan_bytes(g, prm, 100000)
    b = L.plan_bytes(g, prm, 1100000)
    assert (b - a) / 1e6 >= 125
    # velocity storage in FP16 saves 2 bytes per component in the arrays and their reorder buffers
    h16 = L.default_params(solver_mode=1, preconditioner=L.PRECOND_MG, velocity_fp16=1)
    assert L.plan_bytes(g, prm, 1100000) - L.plan_bytes(g, h16, 1100000) == 2 * 2 * 3 * 1100000
    # the multigrid levels are counted only with the multigrid preconditioner
    cg = L.default_params(solver_mode=1, preconditioner=L.PRECOND_JACOBI)
    assert L.plan_bytes(g, cg, 100000) < a
    # a full-frame block capacity of 8^3 blocks: at least the 14 FP32 cell fields per cell
    assert a >= 14 * 4 * 64 ** 3
    # the server charges exactly the plan
    assert L.Server.request_bytes(g, prm, 100000) == a
