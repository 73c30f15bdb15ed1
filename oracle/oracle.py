"""ctypes bindings of the CPU oracle (``oracle/lib/libvare_oracle.so``) -- TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of ``bench.py`` import this
module, and only as the checker.  See ``flip_oracle.cpp`` for the restated reference stages.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# VARE_ORACLE_LIB points the checker at another build of the same source (the mutation test of
# tests/test_oracle_mutants.py loads deliberately broken copies this way)
LIB_PATH = os.environ.get("VARE_ORACLE_LIB") or os.path.join(HERE, "lib", "libvare_oracle.so")
# the OpenMP build of the same source: the multi-core CPU baseline of bench.py
OMP_LIB_PATH = os.path.join(HERE, "lib", "libvare_oracle_omp.so")

FIELD_MASS, FIELD_VELO, FIELD_COMP, FIELD_MARK, FIELD_PRES, FIELD_COUNT = 0, 1, 2, 3, 4, 5
FIELD_B, FIELD_X, FIELD_FLUID, FIELD_AVAL, FIELD_AIND = 6, 7, 8, 9, 10


class OrcParams(C.Structure):
    _fields_ = [("ext_force", C.c_float * 3), ("max_substeps", C.c_int), ("dt", C.c_float),
                ("wall_level", C.c_float), ("enable_water_tank", C.c_int), ("enable_curl", C.c_int),
                ("damping_band", C.c_float), ("voxel_size", C.c_float), ("col_vel_scale", C.c_float),
                ("pts_per_cell", C.c_int), ("vort_pres_rate", C.c_float), ("adjacency_under", C.c_float),
                ("proj_iteration", C.c_int), ("redist_iteration", C.c_int), ("flip_ratio", C.c_float),
                ("solver_mode", C.c_int), ("cg_abs_tol", C.c_float), ("cg_rel_tol", C.c_float),
                ("max_iterations", C.c_int), ("preconditioner", C.c_int)]


class OrcStats(C.Structure):
    _fields_ = [("num_particles", C.c_int64), ("num_blocks", C.c_int64), ("num_cells", C.c_int64),
                ("num_fluid", C.c_int64), ("cg_iters", C.c_int), ("pad", C.c_int),
                ("cg_residual", C.c_double), ("b_norm", C.c_double)]


_libs = {}
_V = C.c_void_p
# sampler(user, world_xyz[3], *sdf, vel[3], grad[3]) -- world units
SAMPLER = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                      C.POINTER(C.c_float))


def load(path=None):
    """The checker build, or another build of the same source (`path`, e.g. OMP_LIB_PATH)."""
    path = path or LIB_PATH
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise RuntimeError("oracle not built at %s; run oracle/build.py" % path)
    lib = C.CDLL(path)
    sig = {
        "orc_rand": (C.c_float, [C.c_uint32]),
        "orc_dot_f32": (C.c_float, [C.c_void_p, C.c_void_p, C.c_size_t]),
        "orc_dot_f64": (C.c_double, [C.c_void_p, C.c_void_p, C.c_size_t]),
        "orc_set_threads": (C.c_int, [C.c_int]),
        "orc_geom": (None, [_V, _V, _V, C.c_int, _V, _V, _V, _V]),
        "orc_grid_from_voxel_size": (None, [C.c_float, _V, _V, _V]),
        "orc_create": (_V, [_V, _V, _V, C.c_int, C.POINTER(OrcParams), C.c_int]),
        "orc_destroy": (None, [_V]),
        "orc_set_particles": (None, [_V, _V, _V, _V, _V, _V, C.c_size_t]),
        "orc_get_particles": (None, [_V, _V, _V, _V, _V, _V]),
        "orc_num_particles": (C.c_size_t, [_V]),
        "orc_step": (C.c_int, [_V, C.c_float]),
        "orc_stage": (None, [_V, C.c_int, C.c_float]),
        "orc_get_block_table": (C.c_size_t, [_V, _V, _V]),
        "orc_get_markers": (None, [_V, _V]),
        "orc_get_field": (C.c_size_t, [_V, C.c_int, _V]),
        "orc_set_field": (None, [_V, C.c_int, _V]),
        "orc_apply_pressure": (None, [_V]),
        "orc_post_process": (None, [_V]),
        "orc_post_part": (None, [_V, C.c_int, C.c_int]),
        "orc_get_removed": (C.c_size_t, [_V, _V]),
        "orc_get_stats": (None, [_V, C.POINTER(OrcStats)]),
        "orc_remove_perm": (C.c_size_t, [_V, C.c_size_t, _V]),
        "orc_seed_box": (C.c_size_t, [C.c_int64, C.c_int64, C.c_int64, C.c_float, C.c_float, C.c_float, C.c_int,
                                      C.c_uint32, _V, C.c_size_t]),
        "orc_set_collision_box": (None, [_V, _V, _V, _V, C.c_float]),
        "orc_set_collision": (None, [_V, _V, C.c_size_t, C.c_float, SAMPLER, _V]),
        "orc_get_collision": (C.c_size_t, [_V, _V, _V, _V, _V, _V]),
        "orc_collision_query": (None, [_V, _V, C.c_size_t, _V, _V, _V]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _libs[path] = lib
    return lib


def set_threads(n, path=OMP_LIB_PATH):
    """Threads of the OpenMP build (returns the count in use)."""
    return load(path).orc_set_threads(int(n))


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def params(**kw) -> OrcParams:
    """Defaults of FLIPMethod::Params (FLIPMethod.h:53-66) + reference constants, overridable."""
    p = OrcParams()
    p.ext_force[:] = [0.0, -9.80665, 0.0]
    p.max_substeps = 1
    p.dt = 1.0 / 24.0
    p.wall_level = -1e30
    p.enable_water_tank = 0
    p.enable_curl = 0
    p.damping_band = 0.0
    p.voxel_size = 0.1
    p.col_vel_scale = 1.0
    p.pts_per_cell = 8
    p.vort_pres_rate = 0.2
    p.adjacency_under = 2.0
    p.proj_iteration = 200
    p.redist_iteration = 80
    p.flip_ratio = 0.95
    p.solver_mode = 0
    p.cg_abs_tol = 1e-3
    p.cg_rel_tol = 1e-5
    p.max_iterations = 2000
    p.preconditioner = 0
    for k, v in kw.items():
        if k == "ext_force":
            p.ext_force[:] = [float(x) for x in v]
        else:
            setattr(p, k, v)
    return p


def dot(a, c):
    """The oracle solver's dot product (BLAS::dot) on float32 or float64 vectors."""
    if a.dtype == np.float32:
        a, c = np.ascontiguousarray(a, np.float32), np.ascontiguousarray(c, np.float32)
        return load().orc_dot_f32(_p(a), _p(c), a.size)
    a, c = np.ascontiguousarray(a, np.float64), np.ascontiguousarray(c, np.float64)
    return load().orc_dot_f64(_p(a), _p(c), a.size)


def rand(seed: int) -> float:
    return load().orc_rand(seed)


def geom(dims, bmin, bmax, B):
    d = np.asarray(dims, np.int64)
    mn = np.asarray(bmin, np.float32)
    mx = np.asarray(bmax, np.float32)
    od = np.zeros(3, np.int64)
    of = np.zeros(3, np.int32)
    ot = np.zeros(3, np.float32)
    os_ = np.zeros(3, np.float32)
    load().orc_geom(_p(d), _p(mn), _p(mx), B, _p(od), _p(of), _p(ot), _p(os_))
    return {"dims": od, "f": of, "translate": ot, "scale": os_}


def grid_from_voxel_size(voxel, bmin, bmax):
    mn = np.asarray(bmin, np.float32)
    mx = np.asarray(bmax, np.float32)
    od = np.zeros(3, np.int64)
    load().orc_grid_from_voxel_size(voxel, _p(mn), _p(mx), _p(od))
    return od


def remove_perm(mask):
    """Array::remove (Array.h:995-1024) applied to the identity: source index of each kept slot."""
    m = np.ascontiguousarray(mask, np.uint8)
    out = np.zeros(m.shape[0], np.int64)
    n = load().orc_remove_perm(_p(m), m.shape[0], _p(out))
    return out[:n]


def seed_box(nx, ny, nz, wx, wy, wz, ppc, seed=1):
    lib = load()
    n = lib.orc_seed_box(nx, ny, nz, wx, wy, wz, ppc, seed, None, 0)
    pos = np.empty((n, 3), np.float32)
    lib.orc_seed_box(nx, ny, nz, wx, wy, wz, ppc, seed, _p(pos), n)
    return pos


class Oracle:
    """One restated FLIPMethod + its particles (voxel space)."""

    def __init__(self, dims, bmin=(0, 0, 0), bmax=(1, 1, 1), B=4, prm: OrcParams = None, fp64=False, lib_path=None):
        self.lib = load(lib_path)
        self.dims = np.asarray(dims, np.int64)
        mn = np.asarray(bmin, np.float32)
        mx = np.asarray(bmax, np.float32)
        self.prm = prm if prm is not None else params()
        self.h = self.lib.orc_create(_p(self.dims), _p(mn), _p(mx), B, C.byref(self.prm), 1 if fp64 else 0)
        self.B = B
        self.geom = geom(dims, bmin, bmax, B)

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_particles(self, pos, vel, vort=None, curv=None, adj=None):
        pos = np.ascontiguousarray(pos, np.float32)
        vel = np.ascontiguousarray(vel, np.float32)
        conv = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (vort, curv, adj)]
        self.lib.orc_set_particles(self.h, _p(pos), _p(vel), _p(conv[0]), _p(conv[1]), _p(conv[2]), pos.shape[0])

    def num_particles(self):
        return self.lib.orc_num_particles(self.h)

    def get_particles(self):
        n = self.num_particles()
        pos = np.empty((n, 3))
        vel = np.empty((n, 3))
        adj = np.empty(n)
        vort = np.empty(n)
        curv = np.empty(n)
        self.lib.orc_get_particles(self.h, _p(pos), _p(vel), _p(adj), _p(vort), _p(curv))
        return {"pos": pos, "vel": vel, "adj": adj, "vort": vort, "curv": curv}

    def step(self, dt):
        return self.lib.orc_step(self.h, dt)

    def stage(self, s, dt=0.0):
        self.lib.orc_stage(self.h, s, dt)

    def block_table(self):
        f = self.geom["f"]
        pointers = np.empty(int(f[0]) * int(f[1]) * int(f[2]), np.int64)
        nb = self.lib.orc_get_block_table(self.h, None, None)
        coords = np.empty((nb, 3), np.int32)
        self.lib.orc_get_block_table(self.h, _p(pointers), _p(coords))
        return pointers, coords

    def markers(self):
        f = self.geom["f"]
        out = np.empty(int(f[0]) * int(f[1]) * int(f[2]), np.int32)
        self.lib.orc_get_markers(self.h, _p(out))
        return out

    def field(self, which):
        n = self.lib.orc_get_field(self.h, which, None)
        out = np.empty(n)
        self.lib.orc_get_field(self.h, which, _p(out))
        if which in (FIELD_VELO, FIELD_COMP):
            out = out.reshape(-1, 3)
        return out

    def set_field(self, which, values):
        v = np.ascontiguousarray(values, np.float64).reshape(-1)
        self.lib.orc_set_field(self.h, which, _p(v))

    def apply_pressure(self):
        """K10-K12 (FluidProjection.cu:297-378) on the x set with set_field(FIELD_X, ...)."""
        self.lib.orc_apply_pressure(self.h)

    def post_process(self):
        """FLIPMethod::postProcess (FLIPMethod.cu:300-405) on the voxel-space particles."""
        self.lib.orc_post_process(self.h)

    def post_part(self, part, arg=0):
        """One part of the post-process on the current fields: 0 splat, 1 tank flood fill, 2 one
        redistancing iteration (arg: tank), 3 curl velo->comp, 4 normal gradient surf->comp,
        5 divergence comp->pres, 6 the whole post-process after the splat (FLIPMethod.cu:320-404)."""
        self.lib.orc_post_part(self.h, int(part), int(arg))

    def removed(self):
        out = np.empty(self.lib.orc_get_removed(self.h, None), np.uint8)
        self.lib.orc_get_removed(self.h, _p(out))
        return out

    def stats(self):
        s = OrcStats()
        self.lib.orc_get_stats(self.h, C.byref(s))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def set_collision_box(self, lo, hi, vel=(0.0, 0.0, 0.0), band=3.0):
        """Box obstacle [lo,hi] (world), moving with `vel` (world units); narrow band of `band` voxels."""
        lo, hi, v = (np.ascontiguousarray(a, np.float32) for a in (lo, hi, vel))
        self.lib.orc_set_collision_box(self.h, _p(lo), _p(hi), _p(v), float(band))

    def set_collision(self, seeds, background, fn):
        """General source: seeds (n,3 world) and fn(world_xyz) -> (sdf, vel[3], grad[3]) in world units."""
        seeds = np.ascontiguousarray(seeds, np.float32).reshape(-1)

        def cb(user, wp, sdf, vel, grad):
            d, v, g = fn(np.array([wp[0], wp[1], wp[2]], np.float32))
            sdf[0] = d
            for a in range(3):
                vel[a] = v[a]
                grad[a] = g[a]

        self._sampler = SAMPLER(cb)
        self.lib.orc_set_collision(self.h, _p(seeds), seeds.size // 3, float(background), self._sampler, None)

    def collision(self):
        """(pointers, phi, vel, nrm, background) of the collision source's sparse field."""
        f = self.geom["f"]
        pointers = np.empty(int(f[0]) * int(f[1]) * int(f[2]), np.int64)
        n = self.lib.orc_get_collision(self.h, None, None, None, None, None)
        phi, vel, nrm, bg = np.empty(n), np.empty((n, 3)), np.empty((n, 3)), np.empty(1)
        self.lib.orc_get_collision(self.h, _p(pointers), _p(phi), _p(vel), _p(nrm), _p(bg))
        return pointers, phi, vel, nrm, float(bg[0])

    def collision_query(self, pts):
        """Signed distance, velocity and normal lerps at voxel-space points."""
        pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
        n = pts.shape[0]
        phi, vel, nrm = np.empty(n), np.empty((n, 3)), np.empty((n, 3))
        self.lib.orc_collision_query(self.h, _p(pts), n, _p(phi), _p(vel), _p(nrm))
        return phi, vel, nrm
