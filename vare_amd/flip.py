"""Host-side mirror of the reference solver API, ``Bora::FLIPMethod`` (base/header/FLIPMethod.h:13-126).

Same method names, argument meaning and in-place semantics as the reference: the caller owns a
``Particles`` object and every call updates its arrays in place (the reference appends to and
removes from ``Particles`` arrays it does not own, FLIPMethod.h:75-99 and FLIPMethod.cu:293-297).
All compute runs in ``libvareflip.so`` on the GPU; this module only moves arrays across the C ABI.
"""
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L


class Grid:
    """Subset of ``Bora::Grid`` (base/header/Grid.h:97-158): cell counts and world bounding box."""

    def __init__(self, nx, ny, nz, bmin=(0.0, 0.0, 0.0), bmax=(1.0, 1.0, 1.0)):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.bmin = tuple(float(x) for x in bmin)
        self.bmax = tuple(float(x) for x in bmax)

    @classmethod
    def from_voxel_size(cls, voxel_size, bmin, bmax):
        """``Grid::initialize(voxelSize, min, max)`` (Grid.h:143-158): nx = (size_t)((max-min)/voxelSize)."""
        g = L.GridDesc()
        arr = lambda v: (L.C.c_float * 3)(*[float(x) for x in v])
        L.check(L.load().vare_grid_from_voxel_size(float(voxel_size), arr(bmin), arr(bmax), 4, L.C.byref(g)))
        return cls(g.nx, g.ny, g.nz, bmin, bmax)

    def numVoxels(self):
        return self.nx * self.ny * self.nz


@dataclass
class Particles:
    """``Bora::Particles`` SoA subset used by the FLIP path (Particles.h:24-36)."""
    position: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.float32))
    velocity: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.float32))
    vorticity: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    curvature: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    adjacency: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))

    def size(self):
        return self.position.shape[0]


@dataclass
class Params:
    """``FLIPMethod::Params`` (FLIPMethod.h:51-67) with the reference's names and defaults, plus the
    constants the reference hard-codes (FLIP ratio, CG stopping rule) and the solver options."""
    extForce: tuple = (0.0, -9.80665, 0.0)
    maxsubsteps: int = 1
    dt: float = 1.0 / 24.0
    wallLevel: float = -1e30
    enableWaterTank: bool = False
    enableCurl: bool = False
    dampingBand: float = 0.0
    voxelSize: float = 0.1
    colVelScale: float = 1.0
    ptsPerCell: int = 8
    vortPresRate: float = 0.2
    adjacencyUnder: float = 2.0
    projIteration: int = 200
    redistIteration: int = 80
    flipRatio: float = 0.95
    solverMode: int = L.SOLVER_REFERENCE
    cgAbsTol: float = 1e-3
    cgRelTol: float = 1e-5
    maxIterations: int = 2000
    preconditioner: int = L.PRECOND_NONE

    def to_c(self) -> L.FlipParams:
        return L.default_params(
            ext_force=self.extForce, max_substeps=int(self.maxsubsteps), dt=float(self.dt),
            wall_level=float(self.wallLevel), enable_water_tank=int(bool(self.enableWaterTank)),
            enable_curl=int(bool(self.enableCurl)), damping_band=float(self.dampingBand),
            voxel_size=float(self.voxelSize), col_vel_scale=float(self.colVelScale),
            pts_per_cell=int(self.ptsPerCell), vort_pres_rate=float(self.vortPresRate),
            adjacency_under=float(self.adjacencyUnder), proj_iteration=int(self.projIteration),
            redist_iteration=int(self.redistIteration), flip_ratio=float(self.flipRatio),
            solver_mode=int(self.solverMode), cg_abs_tol=float(self.cgAbsTol), cg_rel_tol=float(self.cgRelTol),
            max_iterations=int(self.maxIterations), preconditioner=int(self.preconditioner))


class FLIPMethod:
    """Drop-in mirror of ``Bora::FLIPMethod`` on the gfx950 library."""

    def __init__(self, block_size: int = 4):
        self.params = Params()
        self.block_size = int(block_size)
        self._h = None
        self._grid = None
        self._pending = []          # newParticle appends before the next upload

    # FLIPMethod::initialize (FLIPMethod.cu:16-47)
    def initialize(self, grid: Grid):
        if grid.numVoxels() == 0:
            return                                       # FLIPMethod.cu:19
        if self._h is not None:
            self._h.close()
        gd = L.grid_desc(grid.nx, grid.ny, grid.nz, grid.bmin, grid.bmax, self.block_size)
        self._h = L.FlipHandle(gd, self.params.to_c())
        f = self._h.frame()
        n = (f.nx, f.ny, f.nz)
        self._grid = Grid(f.nx, f.ny, f.nz, tuple(f.translate[a] for a in range(3)),
                          tuple(f.translate[a] + n[a] * f.scale[a] for a in range(3)))

    def _sync_params(self):
        self._h.set_params(self.params.to_c())

    def _upload(self, pts: Particles):
        self._sync_params()
        self._h.set_particles(pts.position, pts.velocity, pts.vorticity, pts.curvature, pts.adjacency,
                              space=L.SPACE_VOXEL)

    def _download(self, pts: Particles):
        d = self._h.get_particles(space=L.SPACE_VOXEL)
        pts.position, pts.velocity = d["pos"], d["vel"]
        pts.adjacency, pts.vorticity, pts.curvature = d["adj"], d["vort"], d["curv"]

    # FLIPMethod::newParticle (FLIPMethod.h:75-99)
    def newParticle(self, pts: Particles, p, v):
        pts.position = np.concatenate([pts.position, np.asarray(p, np.float32).reshape(1, 3)])
        pts.velocity = np.concatenate([pts.velocity, np.asarray(v, np.float32).reshape(1, 3)])
        for name in ("vorticity", "curvature", "adjacency"):
            setattr(pts, name, np.concatenate([getattr(pts, name), np.zeros(1, np.float32)]))

    # FLIPMethod::advanceOneStep (FLIPMethod.cu:208-298), voxel space
    def advanceOneStep(self, pts: Particles, dt: float):
        self._upload(pts)
        self._h.step(dt)
        self._download(pts)
        return self._h.stats()

    # FLIPMethod::advanceOneFrame (FLIPMethod.cu:177-206): toVoxel, step, postProcess, toWorld
    def advanceOneFrame(self, pts: Particles):
        if pts.size() == 0:
            return None                                  # FLIPMethod.cu:179
        self._upload(pts)
        self._h.advance_frame()
        self._download(pts)
        return self._h.stats()

    # FLIPMethod::postProcess (FLIPMethod.cu:300-405), voxel space
    def postProcess(self, pts: Particles):
        self._upload(pts)
        self._h.post_process()
        self._download(pts)

    # FLIPMethod::prepareNextStep (FLIPMethod.cu:49-52)
    def prepareNextStep(self, pts: Particles):
        self._upload(pts)
        self._h.stage(0)

    # FLIPMethod::particlesToWorld / particlesToVoxel (FLIPMethod.cu:480-514)
    def particlesToWorld(self, pts: Particles):
        self._upload(pts)
        self._h.to_world()
        self._download(pts)

    def particlesToVoxel(self, pts: Particles):
        self._upload(pts)
        self._h.to_voxel()
        self._download(pts)

    def getGrid(self) -> Grid:
        return self._grid

    # accessors of the sparse fields of the last step (FLIPMethod.h:123-125), storage order
    def getVelocityField(self):
        return self._h.cell_field(L.FIELD_VELOCITY)

    def getDensityField(self):
        return self._h.cell_field(L.FIELD_MASS)

    def getSurfaceField(self):
        """The level set after a frame (postProcess); the count field after a bare step."""
        return self._h.cell_field(L.FIELD_SURFACE)

    # the SOP's VDB export (SOP_FlipWater.cpp:251-275 -> FieldCast.h convertSurfaceVDB/convertVelocityVDB)
    def exportSurfaceVDB(self):
        return self._h.export_surface()

    def exportVelocityVDB(self):
        return self._h.export_velocity()

    @property
    def handle(self) -> "L.FlipHandle":
        return self._h
