/* vare_flip.h -- C ABI of the MI355X (gfx950) FLIP step on the two-level sparse grid.
 *
 * One handle = one Bora::FLIPMethod instance (reference: base/header/FLIPMethod.h:13-126) plus the
 * Particles it advances.  Plain C types only: no HIP, torch or C++ types cross this boundary.
 * Pointers passed in are host pointers unless a `mem` argument says VARE_MEM_DEVICE.
 *
 * Conventions
 *   - Every entry point returns an int status (VARE_OK on success).
 *   - The library owns all device state; the caller owns the buffers it passes.
 *   - Thread-safe across handles, not within one handle.
 *   - Inside the step everything is in voxel space (reference FLIPMethod.cu:191-200).
 */
#ifndef VARE_FLIP_H
#define VARE_FLIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference has none: every FLIPMethod call returns void) ---- */
#define VARE_OK        0
#define VARE_EINVAL   -1   /* bad argument (null handle, zero-sized grid, bad block size ...) */
#define VARE_ENOMEM   -2   /* device allocation failed or capacity exceeded */
#define VARE_EBUDGET  -3   /* serving: request exceeds the HBM budget */
#define VARE_ENOCONV  -4   /* solver hit its iteration cap without meeting its tolerance */
#define VARE_EHIP     -5   /* a HIP runtime call failed */
#define VARE_ENOTIMPL -6   /* feature outside the built path */
#define VARE_ENONFINITE -7 /* NaN/Inf found in particle input */

#define VARE_MEM_HOST   0
#define VARE_MEM_DEVICE 1

#define VARE_SPACE_WORLD 0
#define VARE_SPACE_VOXEL 1

/* solver modes */
#define VARE_SOLVER_REFERENCE 0   /* CG, stop at |r|_2 < cg_abs_tol or after proj_iteration+2
                                     iterations (LinearSystemSolver.cpp:13-50) */
#define VARE_SOLVER_RELATIVE  1   /* PCG, stop at |r|_2 <= cg_rel_tol*|b|_2 or max_iterations */

/* preconditioners for VARE_SOLVER_RELATIVE */
#define VARE_PRECOND_NONE   0
#define VARE_PRECOND_JACOBI 1
#define VARE_PRECOND_MG     2   /* geometric multigrid V-cycle on the sparse block layout */

/* Grid description.  Mirrors Bora::Grid (base/header/Grid.h:97-158): cell counts and the world
 * bounding box [min, max].  block_size is the sparse-frame block edge: 4 is the reference's
 * _block_res (SparseFrame.h:19), 8 is the performance layout. */
typedef struct vare_grid_desc {
    int64_t nx, ny, nz;
    float min[3];
    float max[3];
    int block_size;
} vare_grid_desc;

/* Mirrors FLIPMethod::Params (FLIPMethod.h:51-67) plus the constants the reference hard-codes. */
typedef struct vare_flip_params {
    float ext_force[3];      /* world units, default (0,-9.80665,0) */
    int max_substeps;        /* default 1 */
    float dt;                /* frame dt, default 1/24 */
    float wall_level;        /* world y; default -1e30 (open domain faces) */
    int enable_water_tank;   /* default 0 */
    int enable_curl;         /* default 0; 1 is rejected with VARE_ENOTIMPL */
    float damping_band;      /* default 0 */
    float voxel_size;        /* default 0.1 */
    float col_vel_scale;     /* default 1 */
    int pts_per_cell;        /* default 8 */
    float vort_pres_rate;    /* default 0.2 (post-process only) */
    float adjacency_under;   /* default 2 */
    int proj_iteration;      /* default 200 (CG cap in reference mode is this + 2) */
    int redist_iteration;    /* default 80 (post-process only) */
    /* reference constants, made explicit */
    float flip_ratio;        /* FLIPMethod.cu:99, 0.95 */
    int solver_mode;         /* VARE_SOLVER_REFERENCE (default) | VARE_SOLVER_RELATIVE */
    float cg_abs_tol;        /* LinearSystemSolver.cpp:39, 1e-3 */
    float cg_rel_tol;        /* relative mode tolerance, default 1e-5 */
    int max_iterations;      /* relative mode cap, default 2000 */
    int preconditioner;      /* relative mode: VARE_PRECOND_* */
    int mg_sweeps;           /* multigrid: Jacobi sweeps before and after the coarse correction (1|2) */
    float mg_omega;          /* multigrid: Jacobi weight of the coarsest grid's sweeps (and of the
                                smoothing sweeps when mg_omega_a/b are 0, or with one sweep), default 0.8 */
    int mg_coarse_sweeps;    /* multigrid: sweeps on the coarsest grid, default 20 */
    int particle_path;       /* 0: block-binned P2G/G2P through LDS tiles (default); 1: one thread per
                                particle with global atomics / gathers */
    float refine_tol;        /* > 0: FP64 iterative refinement of the FP32 PCG until the FP64
                                residual |b - A x|/|b| <= refine_tol (pressure kept in FP64) */
    int refine_max;          /* refinement step cap, default 8 */
    int velocity_fp16;       /* 1: particle velocities stored in FP16 (positions stay FP32); fixed at
                                vare_flip_create */
    int flags;               /* VARE_FLAG_* bits, default 0 */
    float mg_omega_a;        /* multigrid, two sweeps: weights of the first and second sweep of each */
    float mg_omega_b;        /* smoothing pair, each in [0, 2); 0 for either: mg_omega for both.
                                Default 1.576514 and 0.566352, the roots of the degree-2 Chebyshev
                                polynomial on [0.4, 2] (Jacobi-scaled 7-point operator: spectrum in
                                [0, 2], its high-frequency band from 1/3): they damp [1/3, 2] to at most
                                0.39 against 0.54 for two sweeps of 0.8.  The two sweeps commute, so the
                                V-cycle stays symmetric.  The coarsest grid keeps mg_omega. */
} vare_flip_params;

/* vare_flip_params.flags */
#define VARE_FLAG_REPORT_DIVERGENCE 1  /* after the pressure apply, also compute the reference's
                                          central-difference divergence b of the projected velocity
                                          (FluidProjection.cu:220-235 on the same cell types) and report
                                          D = |b(v_post)| / |b(v_pre)| in vare_step_stats.div_ratio */
#define VARE_FLAG_FP32_APPLY        2  /* with refinement: round the FP64 pressure to FP32 first and apply
                                          the air extrapolation and gradient in FP32 (default: in FP64
                                          from the refined pressure) */

/* Frame geometry after SparseFrame::initialize (SparseFrame.cu:72-117): padded dims, block
 * counts, and the voxel<->world transform (Grid.h:431-457). */
typedef struct vare_frame_info {
    int64_t nx, ny, nz;
    int fx, fy, fz;
    int block_size;
    float translate[3];
    float scale[3];
    float voxel_wall_level;
    float voxel_ext_force[3];
} vare_frame_info;

/* Per-step record (the reference returns nothing; FLIPMethod.cu:287 drops the CG count). */
typedef struct vare_step_stats {
    int64_t num_particles;       /* after removal */
    int64_t num_removed;
    int64_t num_blocks;          /* active blocks */
    int64_t num_cells;           /* active cells = num_blocks * B^3 */
    int64_t num_fluid;           /* fluid cells (rows of the Poisson system) */
    int cg_iterations;           /* the value LinearSystemSolver::cg returns */
    int converged;
    double residual_norm;        /* recursive |r|_2 at exit */
    double b_norm;               /* |b|_2 */
    double step_ms;              /* device time of the whole step (HIP events) */
    double solve_ms;             /* device time of the Poisson solve */
    int refine_steps;            /* FP64 iterative-refinement steps (0 when refinement is off) */
    int pad;
    double true_rel_residual;    /* |b - A p|_2 / |b|_2 recomputed in FP64 from the final pressure */
    double div_ratio;            /* D = |div_c v_post| / |div_c v_pre| (VARE_FLAG_REPORT_DIVERGENCE), else 0 */
} vare_step_stats;

/* cell fields for vare_flip_get_cell_field (storage order = block-major, x fastest) */
#define VARE_FIELD_MASS     0   /* float   (_sMassField) */
#define VARE_FIELD_VELOCITY 1   /* float*3 (_sVeloField, pre-projection) */
#define VARE_FIELD_COMP     2   /* float*3 (_sCompField, projected) */
#define VARE_FIELD_MARK     3   /* uint64  (_sMarkField: 0 air, 1 solid, 4+fluid rank) */
#define VARE_FIELD_PRESSURE 4   /* float   (_sPresField) */
#define VARE_FIELD_COUNT    5   /* float   (_sSurfField after PointsToCountField) */
#define VARE_FIELD_SURFACE  5   /* float   (_sSurfField after vare_flip_post_process: the level set) */
#define VARE_FIELD_DIV      6   /* float   right-hand side b per cell (0 off the fluid set) */

typedef struct vare_flip vare_flip;

/* Defaults of FLIPMethod::Params plus the reference constants. */
int vare_flip_params_default(vare_flip_params* p);

/* Grid::initialize(voxelSize, min, max) (Grid.h:143-158): nx = (size_t)((max-min)/voxelSize). */
int vare_grid_from_voxel_size(float voxel_size, const float min[3], const float max[3], int block_size,
                              vare_grid_desc* out);

/* FLIPMethod::initialize(const Grid&) (FLIPMethod.cu:16-47).  particle_capacity and
 * block_capacity reserve device memory (0 = choose from the grid); both grow on demand
 * outside of graph capture.  A handle holds at most 2^30 particles (the kernels address particle
 * arrays by 32-bit byte offsets); adding or setting more returns VARE_ENOMEM. */
int vare_flip_create(const vare_grid_desc* grid, const vare_flip_params* params, size_t particle_capacity,
                     size_t block_capacity, vare_flip** out);
int vare_flip_destroy(vare_flip* h);
/* Device bytes vare_flip_create(grid, params, particle_capacity, block_capacity) allocates, plus what
 * the first steps add (multigrid levels, scan tiles) and the staging of one host upload or download of
 * the particles: the peak of the same allocation sequence, counted on the host without a GPU.  No
 * counterpart in the reference (it allocates managed memory as it goes); the server admits requests
 * with it. */
int vare_flip_plan_bytes(const vare_grid_desc* grid, const vare_flip_params* params, size_t particle_capacity,
                         size_t block_capacity, size_t* bytes);
/* Device bytes the handle holds now (live) and the most it has held (peak).  Any pointer may be NULL. */
int vare_flip_get_device_bytes(const vare_flip* h, size_t* live, size_t* peak);
int vare_flip_set_params(vare_flip* h, const vare_flip_params* params);
int vare_flip_get_frame(const vare_flip* h, vare_frame_info* out);

/* FLIPMethod::newParticle (FLIPMethod.h:75-99), batched: appends n particles with zero
 * vorticity, curvature and adjacency.  pos/vel are xyz-interleaved float[3n]. */
int vare_flip_add_particles(vare_flip* h, const float* pos, const float* vel, size_t n, int space, int mem);
/* Replace the whole particle set (checkpoint resume).  vort/curv/adj may be NULL (zeros).
 * `space` (here, in add_particles and get_particles) names the space of the caller's positions and
 * velocities relative to voxel-space storage: VARE_SPACE_WORLD converts world -> voxel on the way in
 * (voxel -> world on the way out), VARE_SPACE_VOXEL copies as is.  The handle does not track which
 * space its particles are in: particles_to_voxel / particles_to_world and advance_frame convert them
 * in place, as FLIPMethod does its Particles, so a handle driven by advance_frame takes and returns
 * world coordinates with VARE_SPACE_VOXEL (no conversion). */
int vare_flip_set_particles(vare_flip* h, const float* pos, const float* vel, const float* vort,
                            const float* curv, const float* adj, size_t n, int space, int mem);
/* Read particles back in the reference's order (Array::remove order, Array.h:995-1024).
 * Any output pointer may be NULL. */
int vare_flip_get_particles(vare_flip* h, float* pos, float* vel, float* adj, float* vort, float* curv,
                            size_t cap, size_t* n, int space, int mem);
int vare_flip_num_particles(const vare_flip* h, size_t* n);

/* FLIPMethod::particlesToVoxel / particlesToWorld (FLIPMethod.cu:480-514). */
int vare_flip_particles_to_voxel(vare_flip* h);
int vare_flip_particles_to_world(vare_flip* h);

/* FLIPMethod::advanceOneStep(pts, dt) (FLIPMethod.cu:208-298), voxel space, including
 * prepareNextStep (:49-52).  Inside the step the adjacency is recomputed for every particle by the
 * G2P stage and not carried through the binning; after a step that returned an error other than
 * VARE_ENOCONV the particles' adjacency is undefined until the next step. */
int vare_flip_step(vare_flip* h, float dt);

/* vare_flip_step without any device allocation, with bit-identical results.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The step's scratch -- the reductions'
 * per-workgroup partials and the scans' tile sums -- comes from one static device array of
 * VARE_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends and shared by every call on that
 * GPU: calls that overlap on one GPU race.  Nothing in it has to start at zero.  The handle is never
 * grown: VARE_ENOMEM when the active blocks outgrow the block capacity reserved at create, or when
 * the step needs more scratch than the array holds (16 bytes per two blocks of block capacity, at
 * least 32 KiB, plus 4 bytes per 4096 cells or particles); VARE_EINVAL while kernel timing is on.
 * The multigrid levels are allocated by vare_flip_create / vare_flip_set_params. */
#define VARE_NOALLOC_SCRATCH_BYTES (4u << 20)
int vare_flip_step_noalloc(vare_flip* h, float dt);
/* vare_flip_step on n independent tanks (distinct handles on one device), their kernels launched
 * together: FLIPMethod::advanceOneStep (FLIPMethod.cu:208-298) per tank, as the reference advances
 * one solver per call; a caller that runs many solvers (a wedge of tanks) calls this instead of one
 * vare_flip_step per handle.  Each kernel of the step is launched once for all the tanks that reach
 * it together, over a list of (tank, workgroup) ranges, so the launch count of a step does not grow
 * with the number of tanks.  Every tank's result is bit-identical to vare_flip_step on its handle
 * alone.  dt[i] and status[i] (may be NULL) belong to hs[i]; the return value is the first error.
 * Kernel timing must be off on every handle (VARE_EINVAL). */
int vare_flip_step_batched(vare_flip* const* hs, int n, const float* dt, int* status);

typedef struct vare_batch_stats {
    int64_t steps;            /* vare_flip_step_batched calls */
    int64_t tanks;            /* tank steps over those calls */
    int64_t flushes;          /* sync points at which the recorded work was issued */
    int64_t launches;         /* merged kernel launches */
    int64_t tank_launches;    /* the kernel launches the tanks would have made on their own */
    double flush_ms;          /* host wall time of the flushes: merging, issuing and waiting */
    double wait_ms;           /* of that, waiting for the merged work to finish on the device */
    double gap_ms;            /* host time between consecutive flushes of a step (the tanks' host code) */
    double step_ms;           /* wall time inside vare_flip_step_batched */
} vare_batch_stats;
/* Counters of the batched steps on `device` since the last reset. */
int vare_batch_get_stats(int device, vare_batch_stats* out, int reset);
/* Test hook (no GPU needed): the merge plan vare_flip_step_batched makes for ntank recorded operation
 * sequences.  keys holds the sequences one after the other (lens[t] entries for tank t): a kernel as
 * its id (>= 0; equal ids merge), a readback as -1, an event record as -2.  For the n-th operation
 * issued, out_tank/out_pos give its tank and its position in that tank's sequence and out_item the
 * merged launch (or single copy or event) it belongs to.  With the outputs NULL only *nout is set. */
int vare_debug_batch_plan(const int32_t* keys, const int32_t* lens, int ntank, int32_t* out_tank, int32_t* out_pos,
                          int32_t* out_item, size_t cap, size_t* nout);

/* FLIPMethod::advanceOneFrame (FLIPMethod.cu:177-206): toVoxel, step(params.dt), postProcess,
 * toWorld. */
int vare_flip_advance_frame(vare_flip* h);
/* FLIPMethod::postProcess (FLIPMethod.cu:300-405) on the voxel-space particles: rebuild the frame
 * from them, splat the surface level set and the velocity (Rasterization.cu:311-393), with the water
 * tank fill the enclosed air below its level (host-side, as the reference), redistance the surface
 * redist_iteration times (Rasterization.cu:395-490), and update each particle's vorticity (curl of
 * the splatted velocity) and curvature (divergence of the surface normals).  Afterwards
 * VARE_FIELD_SURFACE holds the level set, VARE_FIELD_VELOCITY the splatted velocity and
 * VARE_FIELD_COMP the normals.  The reference's splat sums with float atomics in no fixed order;
 * this one sums each cell in a fixed order, so the results repeat from run to run. */
int vare_flip_post_process(vare_flip* h);

/* Sparse -> VDB export (FieldCast.h:14-47, :362-440, as SOP_FlipWater.cpp:251-275 calls it after a
 * frame): the voxels a caller writes with openvdb accessor setValue, and the grid transform.
 *   surface:  every active cell, phi = level set * dx (world units); background = 10 * dx; with the
 *             water tank also the reference's 3-cell shell outside the domain below the level
 *             (min(world cell centre y - wall_level, value)).
 *   velocity: every active cell whose world velocity has length >= 1e-30, in world units.
 * xform[4] = {dx, translate xyz}: openvdb Transform::createLinearTransform(dx) post-translated by
 * the frame's min corner + dx/2.  With ijk and values NULL only *n is written.  Valid after
 * vare_flip_post_process (VARE_EINVAL before). */
int vare_flip_export_surface(vare_flip* h, int32_t* ijk, float* phi, size_t cap, size_t* n, float* background,
                             float xform[4]);
int vare_flip_export_velocity(vare_flip* h, int32_t* ijk, float* vel, size_t cap, size_t* n, float xform[4]);

/* Stages of the step, each equal to the named part of FLIPMethod::advanceOneStep:
 *   0 prepareNextStep / SparseFrame::buildFromPoints   (SparseFrame.cu:127-199)
 *   1 P2G + count field + cell classification          (Rasterization.cu:121-230, FLIPMethod.cu:242-280)
 *   2 FluidProjection::buildLinearSystem (divergence)  (FluidProjection.cu:166-274)
 *   3 FluidProjection::solve (CG + pressure apply)     (FluidProjection.cu:277-381)
 *   4 FLIPMethod::updateParticles                      (FLIPMethod.cu:54-175)
 *   5 Array::remove on the five particle arrays         (FLIPMethod.cu:293-297)
 * Running 0..5 in order with the same dt equals vare_flip_step. */
int vare_flip_stage(vare_flip* h, int stage, float dt);

int vare_flip_get_stats(const vare_flip* h, vare_step_stats* out);

/* Debug/parity accessors. pointers[fx*fy*fz] receives the active slot start (cell offset) or -1
 * (SparseFrame::_pointers with NULL_MAX as -1); coords[3*nb] the active block coordinates in
 * ascending linear order (SparseFrame::_coords). */
int vare_flip_get_block_table(vare_flip* h, int64_t* pointers, int32_t* coords, size_t cap, size_t* nb);
/* Copy a cell field (VARE_FIELD_*) in storage order; n = number of cells. */
int vare_flip_get_cell_field(vare_flip* h, int which, void* out, size_t cap_bytes, size_t* n);
/* Binning permutation of the last prepare stage (particle_path 0): perm[n] lists particle indices
 * grouped by bin -- bin c < Nc is the cell with storage index c, holding the particles whose P2G
 * base cell max((int)(p-0.5),0) it is; bin Nc holds the particles outside [0,n]^3 -- ascending
 * inside a bin: a stable counting sort by bin.  bstart[Nc+2] holds the bin starts.  Either pointer
 * may be NULL; n and nbucket (= Nc+1) are always written.  No counterpart in the reference (its
 * particles are never binned). */
int vare_flip_get_binning(vare_flip* h, int32_t* perm, int32_t* bstart, size_t cap, size_t* n, size_t* nbucket);
/* Per-particle removal flags of the last step, in the order the particles had before removal. */
int vare_flip_get_removed_mask(vare_flip* h, uint8_t* out, size_t cap, size_t* n);

/* ---- Collision source (CollisionSource, base/header/CollisionSource.h:13-150) ----
 * Replaces FLIPMethod::updateCollisionSource(openvdb SDF grid, openvdb velocity grid)
 * (base/header/FLIPMethod.h:105-108 -> CollisionSource::update, CollisionSource.h:66-126).  The
 * source is a sparse field over the simulation frame: the 3x3x3 block neighbourhood of the block
 * of every seed (the world positions of the source's active voxels, as the reference feeds the
 * VDB's active voxels to SparseFrame::buildFromPoints) is activated, and every active cell centre
 * is sampled once.  The step then uses it exactly where the reference does: solid cells and the
 * collision velocity in the cell classification (FLIPMethod.cu:242-260), and the push, projection
 * and removal of particles in G2P (FLIPMethod.cu:116-164).  Until a source is set every query
 * returns signed distance 5 (no collision), as after CollisionSource::initialize. */

/* Sampler standing in for the OpenVDB grids (BoxSampler reads of the SDF, velocity and SDF-gradient
 * grids): at world point wp return the signed distance (negative inside), the collision velocity
 * and the SDF gradient, all in world units.  Called on the host, once per active cell. */
typedef void (*vare_collision_sampler)(void* user, const float wp[3], float* sdf, float vel[3], float grad[3]);
/* Set the source from seeds (nseed world xyz triples) and a sampler; background_world is the
 * signed distance of cells outside the source's active blocks (the VDB background). */
int vare_flip_set_collision(vare_flip* h, const float* seeds_world, size_t nseed, float background_world,
                            vare_collision_sampler fn, void* user);
/* Built-in source: the axis-aligned box [lo,hi] (world) moving with vel (world units/s, NULL = 0),
 * exact box SDF and gradient.  Seeds = cells whose centre lies within band_voxels of the surface
 * (a narrow-band level set of that half-width); background = +band_voxels.  Sampled on host
 * threads. */
int vare_flip_set_collision_box(vare_flip* h, const float lo[3], const float hi[3], const float vel[3],
                                float band_voxels);
/* Remove the source (queries return signed distance 5 again). */
int vare_flip_clear_collision(vare_flip* h);
/* Debug/parity accessor: pointers[fx*fy*fz] = cell offset of each block in the source's field or -1;
 * phi[nc] (voxel units), vel[3*nc] (voxel units), nrm[3*nc] (unit gradient) in storage order;
 * *nc and *background (voxel units) always written.  Field pointers may be NULL. */
int vare_flip_get_collision(vare_flip* h, int64_t* pointers, float* phi, float* vel, float* nrm, size_t cap,
                            size_t* nc, float* background);

/* Seed a box of water in voxel space (the dam-break input of the benchmark): for every cell
 * whose centre is inside [0,wx)x[0,wy)x[0,wz), ppc points at cell + Rand(s+3q+a), a=0,1,2,
 * q = ppc*(i + nx*(j + ny*k)) + n (Rand = Random.h:18-27); only points inside the box are kept.
 * pos may be NULL to count.  Host-side generator. */
int vare_seed_box(int64_t nx, int64_t ny, int64_t nz, float wx, float wy, float wz, int ppc, uint32_t seed,
                  float* pos, size_t cap, size_t* n);
/* Seed a pool of depth `depth` plus a sphere (centre c, radius r), all in voxel units, with the same
 * per-cell rule as vare_seed_box (a cell is seeded when its centre is inside the union; a point is
 * kept when it is inside).  The mixed-precision benchmark scene. */
int vare_seed_pool_sphere(int64_t nx, int64_t ny, int64_t nz, float depth, float cx, float cy, float cz, float r,
                          int ppc, uint32_t seed, float* pos, size_t cap, size_t* n);
/* Rand(seed) of Random.h:18-27 (for checking the generator). */
float vare_rand(uint32_t seed);

/* Device helpers for harnesses. */
int vare_device_synchronize(void);
int vare_device_name(char* buf, size_t cap);

/* Kernel timing.  `mask` is a bitmask of kernel classes (1 << VARE_KERNEL_*); each launch of a
 * selected class is bracketed by HIP events on the handle's stream (the stream the kernel runs
 * on) and the durations are accumulated until reset.  0 disables. */
#define VARE_KERNEL_SPMV     0   /* CG: p = z + beta p fused with A*p + p.Ap (iterations after a solve's first) */
#define VARE_KERNEL_UPDATE   1   /* CG: x,r update + r.r (with multigrid: fused with the first level-0 sweeps) */
#define VARE_KERNEL_P2G      2
#define VARE_KERNEL_G2P      3
#define VARE_KERNEL_BIN      4
#define VARE_KERNEL_DIV      5
#define VARE_KERNEL_APPLY    6
#define VARE_KERNEL_PUPDATE  7   /* unused: the direction update is part of VARE_KERNEL_SPMV */
#define VARE_KERNEL_MG0      8   /* multigrid: level-0 prolongation + Jacobi sweep */
#define VARE_KERNEL_COUNT    9
int vare_flip_set_kernel_timing(vare_flip* h, int mask);
/* Totals per class since the last reset: milliseconds, launches and work items (cells or particles
 * the timed launches processed; filled for VARE_KERNEL_SPMV, VARE_KERNEL_MG0 and, with multigrid and two
 * sweeps, VARE_KERNEL_UPDATE, whose events are
 * recorded on the kernel's own dispatch and so time its execution alone).  For VARE_KERNEL_P2G
 * (particle_path 0: both P2G passes between two events) the units are the algorithmic bytes: 24 per
 * particle (18 with FP16 velocities) + 33 per active cell.  Any array may be NULL. */
int vare_flip_get_kernel_timing(vare_flip* h, double ms_total[VARE_KERNEL_COUNT], int64_t launches[VARE_KERNEL_COUNT],
                                int64_t units[VARE_KERNEL_COUNT], int reset);

/* ---- Serving: many independent tank solves, continuously batched on one GPU ----
 * No counterpart in the reference (one Houdini cook drives one FLIPMethod).  A server admits
 * submitted requests in arrival order whenever a slot and enough of the HBM budget are free; each
 * admitted tank gets its own handle (a pooled one when one fits), and the resident tanks are stepped
 * together in rounds of vare_flip_step_batched (every kernel launched once for all of them); with
 * vare_server_set_batching(s, 0) each tank is stepped on its own stream instead.  A finished tank
 * releases its slot and memory at once and the next arrived request is admitted (continuous
 * batching).  A request's admit_s is stamped when it is admitted, so admit_s follows arrival order. */
typedef struct vare_server vare_server;

typedef struct vare_tank_request {
    vare_grid_desc grid;
    vare_flip_params params;
    const float* pos;        /* voxel-space xyz, host memory, copied at submit */
    const float* vel;
    size_t n;
    int steps;
    float dt;
    double arrival_s;        /* arrival time in seconds after vare_server_run starts */
} vare_tank_request;

typedef struct vare_request_stats {
    double arrival_s, admit_s, finish_s;   /* seconds after run start */
    double step_ms_mean, step_ms_max;
    int steps_done;
    int status;                            /* VARE_OK or the first error of the request */
    int64_t num_particles;
    int64_t bytes_reserved;
} vare_request_stats;

typedef struct vare_server_stats {
    int64_t requests, steps;
    double wall_s;                         /* first arrival to last finish */
    double step_p50_ms, step_p99_ms, step_mean_ms;
    double particle_steps_per_s;
    int max_resident;                      /* most tanks resident at once */
    int pad;
    int64_t peak_bytes;                    /* peak reserved HBM (resident tanks + pooled handles) */
    int64_t handles_created;               /* device states allocated by the last run */
    int64_t handles_reused;                /* requests served by a pooled device state of the same grid */
    int64_t rounds;                        /* batched rounds (one vare_flip_step_batched call each) */
    int64_t max_batch;                     /* most tanks stepped in one round */
} vare_server_stats;

int vare_server_create(size_t hbm_budget_bytes, int max_slots, vare_server** out);
/* Batched stepping (on = 1, the default): every round advances all resident tanks by one step with
 * vare_flip_step_batched.  Off: each resident tank steps on its own handle's stream (vare_flip_step),
 * overlapping with the others only as concurrent streams. */
int vare_server_set_batching(vare_server* s, int on);
int vare_server_destroy(vare_server* s);
/* Queue a request (copies the particle arrays); returns its id. */
int vare_server_submit(vare_server* s, const vare_tank_request* req, int64_t* id);
/* Run until every submitted request has finished, honouring arrival times. */
int vare_server_run(vare_server* s);
/* Final particles of a finished request (voxel space) and its record; any pointer may be NULL. */
int vare_server_result(vare_server* s, int64_t id, float* pos, float* vel, size_t cap, size_t* n,
                       vare_request_stats* st);
int vare_server_get_stats(vare_server* s, vare_server_stats* out);
/* Estimated device bytes a request reserves (what admission charges against the budget). */
int vare_server_request_bytes(const vare_tank_request* req, size_t* bytes);

/* Test hook (process-wide): after the next `skip` device allocations of any handle, the following
 * `count` fail with VARE_ENOMEM as if hipMalloc had failed; count <= 0 switches it off.  A failed
 * allocation leaves a handle usable with its previous capacity. */
int vare_debug_inject_alloc_failure(long skip, long count);

/* Last error message of the calling thread. */
const char* vare_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* VARE_FLIP_H */
