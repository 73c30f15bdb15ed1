"""ctypes bindings of ``libvareflip.so`` (the C ABI declared in ``include/vare_flip.h``).

The library must be built (``python -m vare_amd.build`` or ``__graft_entry__.build()``); loading
fails loudly otherwise -- there is no CPU fallback for any entry point.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libvareflip.so")

VARE_OK = 0
VARE_EINVAL = -1
VARE_ENOMEM = -2
VARE_EBUDGET = -3
VARE_ENOCONV = -4
VARE_EHIP = -5
VARE_ENOTIMPL = -6
VARE_ENONFINITE = -7

MEM_HOST = 0
MEM_DEVICE = 1
SPACE_WORLD = 0
SPACE_VOXEL = 1
SOLVER_REFERENCE = 0
SOLVER_RELATIVE = 1
PRECOND_NONE = 0
PRECOND_JACOBI = 1
PRECOND_MG = 2
FLAG_REPORT_DIVERGENCE = 1
FLAG_FP32_APPLY = 2

FIELD_MASS = 0
FIELD_VELOCITY = 1
FIELD_COMP = 2
FIELD_MARK = 3
FIELD_PRESSURE = 4
FIELD_COUNT = 5
FIELD_SURFACE = 5
FIELD_DIV = 6

KERNEL_NAMES = ["spmv", "update", "p2g", "g2p", "bin", "div", "apply", "pupdate", "mg0"]
KERNEL_COUNT = 9

ERROR_NAMES = {
    VARE_EINVAL: "VARE_EINVAL",
    VARE_ENOMEM: "VARE_ENOMEM",
    VARE_EBUDGET: "VARE_EBUDGET",
    VARE_ENOCONV: "VARE_ENOCONV",
    VARE_EHIP: "VARE_EHIP",
    VARE_ENOTIMPL: "VARE_ENOTIMPL",
    VARE_ENONFINITE: "VARE_ENONFINITE",
}


class VareError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (ERROR_NAMES.get(code, str(code)), msg))
        self.code = code


class GridDesc(C.Structure):
    _fields_ = [("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
                ("min", C.c_float * 3), ("max", C.c_float * 3), ("block_size", C.c_int)]


class FlipParams(C.Structure):
    _fields_ = [("ext_force", C.c_float * 3), ("max_substeps", C.c_int), ("dt", C.c_float),
                ("wall_level", C.c_float), ("enable_water_tank", C.c_int), ("enable_curl", C.c_int),
                ("damping_band", C.c_float), ("voxel_size", C.c_float), ("col_vel_scale", C.c_float),
                ("pts_per_cell", C.c_int), ("vort_pres_rate", C.c_float), ("adjacency_under", C.c_float),
                ("proj_iteration", C.c_int), ("redist_iteration", C.c_int), ("flip_ratio", C.c_float),
                ("solver_mode", C.c_int), ("cg_abs_tol", C.c_float), ("cg_rel_tol", C.c_float),
                ("max_iterations", C.c_int), ("preconditioner", C.c_int), ("mg_sweeps", C.c_int),
                ("mg_omega", C.c_float), ("mg_coarse_sweeps", C.c_int), ("particle_path", C.c_int),
                ("refine_tol", C.c_float), ("refine_max", C.c_int), ("velocity_fp16", C.c_int),
                ("flags", C.c_int), ("mg_omega_a", C.c_float), ("mg_omega_b", C.c_float)]


class FrameInfo(C.Structure):
    _fields_ = [("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
                ("fx", C.c_int), ("fy", C.c_int), ("fz", C.c_int), ("block_size", C.c_int),
                ("translate", C.c_float * 3), ("scale", C.c_float * 3), ("voxel_wall_level", C.c_float),
                ("voxel_ext_force", C.c_float * 3)]


class StepStats(C.Structure):
    _fields_ = [("num_particles", C.c_int64), ("num_removed", C.c_int64), ("num_blocks", C.c_int64),
                ("num_cells", C.c_int64), ("num_fluid", C.c_int64), ("cg_iterations", C.c_int),
                ("converged", C.c_int), ("residual_norm", C.c_double), ("b_norm", C.c_double),
                ("step_ms", C.c_double), ("solve_ms", C.c_double), ("refine_steps", C.c_int), ("pad", C.c_int),
                ("true_rel_residual", C.c_double), ("div_ratio", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TankRequest(C.Structure):
    _fields_ = [("grid", GridDesc), ("params", FlipParams), ("pos", C.c_void_p), ("vel", C.c_void_p),
                ("n", C.c_size_t), ("steps", C.c_int), ("dt", C.c_float), ("arrival_s", C.c_double)]


class RequestStats(C.Structure):
    _fields_ = [("arrival_s", C.c_double), ("admit_s", C.c_double), ("finish_s", C.c_double),
                ("step_ms_mean", C.c_double), ("step_ms_max", C.c_double), ("steps_done", C.c_int),
                ("status", C.c_int), ("num_particles", C.c_int64), ("bytes_reserved", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(C.Structure):
    _fields_ = [("requests", C.c_int64), ("steps", C.c_int64), ("wall_s", C.c_double),
                ("step_p50_ms", C.c_double), ("step_p99_ms", C.c_double), ("step_mean_ms", C.c_double),
                ("particle_steps_per_s", C.c_double), ("max_resident", C.c_int), ("pad", C.c_int),
                ("peak_bytes", C.c_int64), ("handles_created", C.c_int64), ("handles_reused", C.c_int64),
                ("rounds", C.c_int64), ("max_batch", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class BatchStats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("tanks", C.c_int64), ("flushes", C.c_int64), ("launches", C.c_int64),
                ("tank_launches", C.c_int64), ("flush_ms", C.c_double), ("wait_ms", C.c_double),
                ("gap_ms", C.c_double), ("step_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_P = C.c_void_p
_FP = C.POINTER(C.c_float)
# vare_collision_sampler(user, wp[3], *sdf, vel[3], grad[3]) -- world units
COLLISION_SAMPLER = C.CFUNCTYPE(None, C.c_void_p, _FP, _FP, _FP, _FP)

# name -> (restype, argtypes); every symbol declared in include/vare_flip.h
SIGNATURES = {
    "vare_flip_params_default": (C.c_int, [C.POINTER(FlipParams)]),
    "vare_grid_from_voxel_size": (C.c_int, [C.c_float, C.c_float * 3, C.c_float * 3, C.c_int, C.POINTER(GridDesc)]),
    "vare_flip_create": (C.c_int, [C.POINTER(GridDesc), C.POINTER(FlipParams), C.c_size_t, C.c_size_t, C.POINTER(_P)]),
    "vare_flip_destroy": (C.c_int, [_P]),
    "vare_flip_plan_bytes": (C.c_int, [C.POINTER(GridDesc), C.POINTER(FlipParams), C.c_size_t, C.c_size_t,
                                       C.POINTER(C.c_size_t)]),
    "vare_flip_get_device_bytes": (C.c_int, [_P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "vare_debug_inject_alloc_failure": (C.c_int, [C.c_long, C.c_long]),
    "vare_flip_set_params": (C.c_int, [_P, C.POINTER(FlipParams)]),
    "vare_flip_get_frame": (C.c_int, [_P, C.POINTER(FrameInfo)]),
    "vare_flip_add_particles": (C.c_int, [_P, _P, _P, C.c_size_t, C.c_int, C.c_int]),
    "vare_flip_set_particles": (C.c_int, [_P, _P, _P, _P, _P, _P, C.c_size_t, C.c_int, C.c_int]),
    "vare_flip_get_particles": (C.c_int, [_P, _P, _P, _P, _P, _P, C.c_size_t, C.POINTER(C.c_size_t), C.c_int, C.c_int]),
    "vare_flip_num_particles": (C.c_int, [_P, C.POINTER(C.c_size_t)]),
    "vare_flip_particles_to_voxel": (C.c_int, [_P]),
    "vare_flip_particles_to_world": (C.c_int, [_P]),
    "vare_flip_step": (C.c_int, [_P, C.c_float]),
    "vare_flip_step_noalloc": (C.c_int, [_P, C.c_float]),
    "vare_flip_step_batched": (C.c_int, [C.POINTER(_P), C.c_int, _FP, C.POINTER(C.c_int)]),
    "vare_batch_get_stats": (C.c_int, [C.c_int, C.POINTER(BatchStats), C.c_int]),
    "vare_debug_batch_plan": (C.c_int, [_P, _P, C.c_int, _P, _P, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "vare_flip_advance_frame": (C.c_int, [_P]),
    "vare_flip_post_process": (C.c_int, [_P]),
    "vare_flip_export_surface": (C.c_int, [_P, _P, _P, C.c_size_t, C.POINTER(C.c_size_t), _FP, _FP]),
    "vare_flip_export_velocity": (C.c_int, [_P, _P, _P, C.c_size_t, C.POINTER(C.c_size_t), _FP]),
    "vare_flip_stage": (C.c_int, [_P, C.c_int, C.c_float]),
    "vare_flip_get_stats": (C.c_int, [_P, C.POINTER(StepStats)]),
    "vare_flip_get_block_table": (C.c_int, [_P, _P, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "vare_flip_get_cell_field": (C.c_int, [_P, C.c_int, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "vare_flip_get_removed_mask": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "vare_flip_set_collision": (C.c_int, [_P, _P, C.c_size_t, C.c_float, COLLISION_SAMPLER, _P]),
    "vare_flip_set_collision_box": (C.c_int, [_P, C.c_float * 3, C.c_float * 3, C.c_float * 3, C.c_float]),
    "vare_flip_clear_collision": (C.c_int, [_P]),
    "vare_flip_get_collision": (C.c_int, [_P, _P, _P, _P, _P, C.c_size_t, C.POINTER(C.c_size_t), _FP]),
    "vare_flip_get_binning": (C.c_int, [_P, _P, _P, C.c_size_t, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "vare_seed_box": (C.c_int, [C.c_int64, C.c_int64, C.c_int64, C.c_float, C.c_float, C.c_float, C.c_int,
                                C.c_uint32, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "vare_seed_pool_sphere": (C.c_int, [C.c_int64, C.c_int64, C.c_int64, C.c_float, C.c_float, C.c_float, C.c_float,
                                        C.c_float, C.c_int, C.c_uint32, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "vare_rand": (C.c_float, [C.c_uint32]),
    "vare_device_synchronize": (C.c_int, []),
    "vare_device_name": (C.c_int, [C.c_char_p, C.c_size_t]),
    "vare_flip_set_kernel_timing": (C.c_int, [_P, C.c_int]),
    "vare_flip_get_kernel_timing": (C.c_int, [_P, C.POINTER(C.c_double * KERNEL_COUNT),
                                              C.POINTER(C.c_int64 * KERNEL_COUNT),
                                              C.POINTER(C.c_int64 * KERNEL_COUNT), C.c_int]),
    "vare_last_error": (C.c_char_p, []),
    "vare_server_create": (C.c_int, [C.c_size_t, C.c_int, C.POINTER(_P)]),
    "vare_server_destroy": (C.c_int, [_P]),
    "vare_server_set_batching": (C.c_int, [_P, C.c_int]),
    "vare_server_submit": (C.c_int, [_P, C.POINTER(TankRequest), C.POINTER(C.c_int64)]),
    "vare_server_run": (C.c_int, [_P]),
    "vare_server_result": (C.c_int, [_P, C.c_int64, _P, _P, C.c_size_t, C.POINTER(C.c_size_t),
                                     C.POINTER(RequestStats)]),
    "vare_server_get_stats": (C.c_int, [_P, C.POINTER(ServerStats)]),
    "vare_server_request_bytes": (C.c_int, [C.POINTER(TankRequest), C.POINTER(C.c_size_t)]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libvareflip.so not built at %s; run __graft_entry__.build()" % path)
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != VARE_OK:
        raise VareError(rc, load().vare_last_error().decode())
    return rc


def ptr(a):
    """ctypes void pointer of a contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)


def default_params(**kw) -> FlipParams:
    p = FlipParams()
    check(load().vare_flip_params_default(C.byref(p)))
    for k, v in kw.items():
        if k == "ext_force":
            p.ext_force[:] = [float(x) for x in v]
        else:
            setattr(p, k, v)
    return p


def grid_desc(nx, ny, nz, bmin=(0.0, 0.0, 0.0), bmax=(1.0, 1.0, 1.0), block_size=4) -> GridDesc:
    g = GridDesc()
    g.nx, g.ny, g.nz = int(nx), int(ny), int(nz)
    g.min[:] = [float(x) for x in bmin]
    g.max[:] = [float(x) for x in bmax]
    g.block_size = int(block_size)
    return g


def seed_box(nx, ny, nz, wx, wy, wz, ppc, seed=1):
    """Dam-break particles in voxel space (see vare_seed_box)."""
    lib = load()
    n = C.c_size_t()
    check(lib.vare_seed_box(nx, ny, nz, wx, wy, wz, ppc, seed, None, 0, C.byref(n)))
    pos = np.empty((n.value, 3), np.float32)
    check(lib.vare_seed_box(nx, ny, nz, wx, wy, wz, ppc, seed, ptr(pos), n.value, C.byref(n)))
    return pos


def seed_pool_sphere(nx, ny, nz, depth, centre, radius, ppc, seed=1):
    """Pool of depth `depth` plus a sphere, voxel space (see vare_seed_pool_sphere)."""
    lib = load()
    n = C.c_size_t()
    args = (nx, ny, nz, depth, centre[0], centre[1], centre[2], radius, ppc, seed)
    check(lib.vare_seed_pool_sphere(*args, None, 0, C.byref(n)))
    pos = np.empty((n.value, 3), np.float32)
    check(lib.vare_seed_pool_sphere(*args, ptr(pos), n.value, C.byref(n)))
    return pos


class FlipHandle:
    """Owning wrapper of one ``vare_flip`` handle."""

    def __init__(self, grid: GridDesc, params: FlipParams, particle_capacity=0, block_capacity=0):
        self.lib = load()
        self.h = C.c_void_p()
        check(self.lib.vare_flip_create(C.byref(grid), C.byref(params), particle_capacity, block_capacity,
                                        C.byref(self.h)))
        self.grid = grid
        self.params = params

    def close(self):
        if self.h:
            self.lib.vare_flip_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_bytes(self):
        """(live, peak) device bytes of the handle."""
        live, peak = C.c_size_t(), C.c_size_t()
        check(self.lib.vare_flip_get_device_bytes(self.h, C.byref(live), C.byref(peak)))
        return live.value, peak.value

    def set_params(self, params: FlipParams):
        check(self.lib.vare_flip_set_params(self.h, C.byref(params)))
        self.params = params

    def frame(self) -> FrameInfo:
        f = FrameInfo()
        check(self.lib.vare_flip_get_frame(self.h, C.byref(f)))
        return f

    def set_particles(self, pos, vel, vort=None, curv=None, adj=None, space=SPACE_VOXEL):
        pos = np.ascontiguousarray(pos, np.float32)
        vel = np.ascontiguousarray(vel, np.float32)
        n = pos.shape[0]
        conv = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (vort, curv, adj)]
        check(self.lib.vare_flip_set_particles(self.h, ptr(pos), ptr(vel), ptr(conv[0]), ptr(conv[1]), ptr(conv[2]),
                                               n, space, MEM_HOST))

    def add_particles(self, pos, vel, space=SPACE_VOXEL):
        pos = np.ascontiguousarray(pos, np.float32)
        vel = np.ascontiguousarray(vel, np.float32)
        check(self.lib.vare_flip_add_particles(self.h, ptr(pos), ptr(vel), pos.shape[0], space, MEM_HOST))

    def num_particles(self) -> int:
        n = C.c_size_t()
        check(self.lib.vare_flip_num_particles(self.h, C.byref(n)))
        return n.value

    def get_particles(self, space=SPACE_VOXEL):
        n = self.num_particles()
        pos = np.empty((n, 3), np.float32)
        vel = np.empty((n, 3), np.float32)
        adj = np.empty(n, np.float32)
        vort = np.empty(n, np.float32)
        curv = np.empty(n, np.float32)
        got = C.c_size_t()
        check(self.lib.vare_flip_get_particles(self.h, ptr(pos), ptr(vel), ptr(adj), ptr(vort), ptr(curv), n,
                                               C.byref(got), space, MEM_HOST))
        return {"pos": pos, "vel": vel, "adj": adj, "vort": vort, "curv": curv}

    def step(self, dt):
        check(self.lib.vare_flip_step(self.h, float(dt)))

    def step_rc(self, dt):
        return self.lib.vare_flip_step(self.h, float(dt))

    def step_noalloc(self, dt):
        """vare_flip_step_noalloc: the same step with no device allocation (static scratch)."""
        check(self.lib.vare_flip_step_noalloc(self.h, float(dt)))

    def step_noalloc_rc(self, dt):
        return self.lib.vare_flip_step_noalloc(self.h, float(dt))

    def stage(self, s, dt=0.0):
        check(self.lib.vare_flip_stage(self.h, int(s), float(dt)))

    def advance_frame(self):
        check(self.lib.vare_flip_advance_frame(self.h))

    def post_process(self):
        """vare_flip_post_process: FLIPMethod::postProcess on the voxel-space particles."""
        check(self.lib.vare_flip_post_process(self.h))

    def export_surface(self):
        """(ijk[n,3], phi[n], background, xform[4]) of the surface VDB export."""
        n = C.c_size_t()
        bg = C.c_float()
        xf = (C.c_float * 4)()
        check(self.lib.vare_flip_export_surface(self.h, None, None, 0, C.byref(n), C.byref(bg), xf))
        ijk = np.empty((n.value, 3), np.int32)
        phi = np.empty(n.value, np.float32)
        check(self.lib.vare_flip_export_surface(self.h, ptr(ijk), ptr(phi), n.value, C.byref(n), C.byref(bg), xf))
        return ijk, phi, bg.value, np.array(list(xf), np.float32)

    def export_velocity(self):
        """(ijk[n,3], vel[n,3] world units, xform[4]) of the velocity VDB export."""
        n = C.c_size_t()
        xf = (C.c_float * 4)()
        check(self.lib.vare_flip_export_velocity(self.h, None, None, 0, C.byref(n), xf))
        ijk = np.empty((n.value, 3), np.int32)
        vel = np.empty((n.value, 3), np.float32)
        check(self.lib.vare_flip_export_velocity(self.h, ptr(ijk), ptr(vel), n.value, C.byref(n), xf))
        return ijk, vel, np.array(list(xf), np.float32)

    def to_voxel(self):
        check(self.lib.vare_flip_particles_to_voxel(self.h))

    def to_world(self):
        check(self.lib.vare_flip_particles_to_world(self.h))

    def stats(self) -> dict:
        s = StepStats()
        check(self.lib.vare_flip_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    def block_table(self):
        f = self.frame()
        nb = C.c_size_t()
        check(self.lib.vare_flip_get_block_table(self.h, None, None, 0, C.byref(nb)))
        pointers = np.empty(f.fx * f.fy * f.fz, np.int64)
        coords = np.empty((nb.value, 3), np.int32)
        check(self.lib.vare_flip_get_block_table(self.h, ptr(pointers), ptr(coords), nb.value, C.byref(nb)))
        return pointers, coords

    def cell_field(self, which):
        n = C.c_size_t()
        check(self.lib.vare_flip_get_cell_field(self.h, which, None, 0, C.byref(n)))
        nc = n.value
        if which in (FIELD_VELOCITY, FIELD_COMP):
            out = np.empty((nc, 3), np.float32)
        elif which == FIELD_MARK:
            out = np.empty(nc, np.uint64)
        else:
            out = np.empty(nc, np.float32)
        check(self.lib.vare_flip_get_cell_field(self.h, which, ptr(out), out.nbytes, C.byref(n)))
        return out

    def binning(self):
        n = C.c_size_t()
        nbk = C.c_size_t()
        check(self.lib.vare_flip_get_binning(self.h, None, None, 0, C.byref(n), C.byref(nbk)))
        perm = np.empty(n.value, np.int32)
        bstart = np.empty(nbk.value + 1, np.int32)
        check(self.lib.vare_flip_get_binning(self.h, ptr(perm), ptr(bstart), n.value, C.byref(n), C.byref(nbk)))
        return perm, bstart

    def removed_mask(self):
        n = C.c_size_t()
        check(self.lib.vare_flip_get_removed_mask(self.h, None, 0, C.byref(n)))
        out = np.empty(n.value, np.uint8)
        check(self.lib.vare_flip_get_removed_mask(self.h, ptr(out), n.value, C.byref(n)))
        return out

    def set_collision_box(self, lo, hi, vel=(0.0, 0.0, 0.0), band=3.0):
        """Box obstacle [lo,hi] (world) moving with `vel` (world units/s); narrow band in voxels."""
        f3 = C.c_float * 3
        check(self.lib.vare_flip_set_collision_box(self.h, f3(*lo), f3(*hi), f3(*vel), float(band)))

    def set_collision(self, seeds, background, fn):
        """General source: seeds (n,3 world) and fn(world_xyz) -> (sdf, vel[3], grad[3]) in world units."""
        seeds = np.ascontiguousarray(seeds, np.float32).reshape(-1)

        def cb(user, wp, sdf, vel, grad):
            d, v, g = fn(np.array([wp[0], wp[1], wp[2]], np.float32))
            sdf[0] = d
            for a in range(3):
                vel[a] = v[a]
                grad[a] = g[a]

        sampler = COLLISION_SAMPLER(cb)
        check(self.lib.vare_flip_set_collision(self.h, ptr(seeds), seeds.size // 3, float(background), sampler, None))

    def clear_collision(self):
        check(self.lib.vare_flip_clear_collision(self.h))

    def collision(self):
        """(pointers, phi, vel, nrm, background) of the collision source's sparse field."""
        fr = self.frame()
        nblk = fr.fx * fr.fy * fr.fz
        n = C.c_size_t()
        bg = C.c_float()
        pointers = np.empty(nblk, np.int64)
        check(self.lib.vare_flip_get_collision(self.h, ptr(pointers), None, None, None, 0, C.byref(n), C.byref(bg)))
        m = n.value
        phi, vel, nrm = np.empty(m, np.float32), np.empty((m, 3), np.float32), np.empty((m, 3), np.float32)
        check(self.lib.vare_flip_get_collision(self.h, None, ptr(phi), ptr(vel), ptr(nrm), m, C.byref(n), C.byref(bg)))
        return pointers, phi, vel, nrm, bg.value

    def set_kernel_timing(self, kernels=()):
        mask = 0
        for k in kernels:
            mask |= 1 << KERNEL_NAMES.index(k)
        check(self.lib.vare_flip_set_kernel_timing(self.h, mask))

    def kernel_timing(self, reset=False):
        """{class: (total ms, launches, work items)} since the last reset."""
        ms = (C.c_double * KERNEL_COUNT)()
        cnt = (C.c_int64 * KERNEL_COUNT)()
        units = (C.c_int64 * KERNEL_COUNT)()
        check(self.lib.vare_flip_get_kernel_timing(self.h, C.byref(ms), C.byref(cnt), C.byref(units),
                                                   1 if reset else 0))
        return {KERNEL_NAMES[i]: (ms[i], cnt[i], units[i]) for i in range(KERNEL_COUNT)}


class Server:
    """Owning wrapper of a ``vare_server`` (continuous batching of independent tank solves)."""

    def __init__(self, budget_bytes, max_slots, batching=True):
        self.lib = load()
        self.s = C.c_void_p()
        check(self.lib.vare_server_create(int(budget_bytes), int(max_slots), C.byref(self.s)))
        check(self.lib.vare_server_set_batching(self.s, 1 if batching else 0))
        self._keep = []

    def close(self):
        if self.s:
            self.lib.vare_server_destroy(self.s)
            self.s = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def make_request(grid, params, pos, vel, steps, dt, arrival_s=0.0):
        pos = np.ascontiguousarray(pos, np.float32)
        vel = np.ascontiguousarray(vel, np.float32)
        r = TankRequest()
        r.grid = grid
        r.params = params
        r.pos = pos.ctypes.data_as(C.c_void_p)
        r.vel = vel.ctypes.data_as(C.c_void_p)
        r.n = pos.shape[0]
        r.steps = int(steps)
        r.dt = float(dt)
        r.arrival_s = float(arrival_s)
        return r, (pos, vel)

    def submit(self, grid, params, pos, vel, steps, dt, arrival_s=0.0) -> int:
        r, keep = self.make_request(grid, params, pos, vel, steps, dt, arrival_s)
        rid = C.c_int64()
        check(self.lib.vare_server_submit(self.s, C.byref(r), C.byref(rid)))
        return rid.value

    def run(self):
        check(self.lib.vare_server_run(self.s))

    def result(self, rid):
        n = C.c_size_t()
        st = RequestStats()
        check(self.lib.vare_server_result(self.s, rid, None, None, 0, C.byref(n), C.byref(st)))
        pos = np.empty((n.value, 3), np.float32)
        vel = np.empty((n.value, 3), np.float32)
        check(self.lib.vare_server_result(self.s, rid, ptr(pos), ptr(vel), n.value, C.byref(n), C.byref(st)))
        return pos, vel, st.as_dict()

    def stats(self):
        st = ServerStats()
        check(self.lib.vare_server_get_stats(self.s, C.byref(st)))
        return st.as_dict()

    @staticmethod
    def request_bytes(grid, params, n, steps=1, dt=1 / 240):
        r = TankRequest()
        r.grid = grid
        r.params = params
        r.n = int(n)
        r.steps = steps
        r.dt = dt
        b = C.c_size_t()
        check(load().vare_server_request_bytes(C.byref(r), C.byref(b)))
        return b.value


def step_batched(handles, dt):
    """vare_flip_step_batched: one step of every handle, the kernels of all tanks launched together.
    dt is one value or one per handle.  Returns the per-tank status codes (raises on the first error)."""
    n = len(handles)
    hs = (_P * n)(*[h.h for h in handles])
    dts = (C.c_float * n)(*([float(dt)] * n if np.isscalar(dt) else [float(x) for x in dt]))
    st = (C.c_int * n)()
    check(load().vare_flip_step_batched(hs, n, dts, st))
    return list(st)


def batch_plan(seqs):
    """vare_debug_batch_plan: the issue order of recorded operation sequences (kernel ids >= 0, -1 a
    readback, -2 an event).  Returns (tank, position, merged item) per issued operation."""
    keys = np.ascontiguousarray(np.concatenate([np.asarray(q, np.int32) for q in seqs]) if seqs else [], np.int32)
    lens = np.ascontiguousarray([len(q) for q in seqs], np.int32)
    n = C.c_size_t()
    check(load().vare_debug_batch_plan(ptr(keys), ptr(lens), len(seqs), None, None, None, 0, C.byref(n)))
    t = np.empty(n.value, np.int32)
    p = np.empty(n.value, np.int32)
    it = np.empty(n.value, np.int32)
    check(load().vare_debug_batch_plan(ptr(keys), ptr(lens), len(seqs), ptr(t), ptr(p), ptr(it), n.value, C.byref(n)))
    return t, p, it


def batch_stats(device=0, reset=False):
    """Counters of the batched steps: calls, tank steps, sync points, merged and per-tank launches."""
    s = BatchStats()
    check(load().vare_batch_get_stats(int(device), C.byref(s), 1 if reset else 0))
    return s.as_dict()


def plan_bytes(grid, params, particle_capacity=0, block_capacity=0):
    """Device bytes vare_flip_create + the first steps allocate (vare_flip_plan_bytes; no GPU needed)."""
    b = C.c_size_t()
    check(load().vare_flip_plan_bytes(C.byref(grid), C.byref(params), particle_capacity, block_capacity,
                                      C.byref(b)))
    return b.value


def inject_alloc_failure(skip, count):
    """Test hook: after `skip` device allocations, fail the next `count` (count <= 0: off)."""
    check(load().vare_debug_inject_alloc_failure(int(skip), int(count)))


def device_synchronize():
    check(load().vare_device_synchronize())


def device_name() -> str:
    buf = C.create_string_buffer(256)
    check(load().vare_device_name(buf, 256))
    return buf.value.decode()
