// flip_oracle.cpp -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// CPU restatement of the reference FLIP step (Bora FLIPMethod::advanceOneStep) used only as the
// checker by tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py.  The product
// path (vare_amd/csrc) never includes, links or calls anything in this directory.
//
// Written from the reference sources read as text; every stage cites the file:line it follows.
// Two instances are compiled from one template:
//   * R = float  : reference-faithful FP32 (same operation order, no FMA contraction -- build with
//                  -ffp-contract=off), sequential P2G in particle order.
//   * R = double : the same algorithm in FP64, the error yardstick for the GPU results.
//
// Built a second time with -fopenmp (oracle/lib/libvare_oracle_omp.so) as the multi-core CPU baseline
// that bench.py times: the same stages with the loops over particles and cells run in parallel, and
// P2G / the count field scattering z-slabs of particles two apart at a time (no two threads write a
// cell at once, no atomics), so its FP32 sums come in another order.  The checker build has no
// OpenMP and keeps the sequential order.
//
// Documented deviations from the reference (both unavoidable on a CPU, neither changes physics):
//   * K1 block marking (SparseFrame.cu:149-157) writes out of bounds of _markers when the signed
//     int index k*fx*fy+j*fx+i is negative; those writes are dropped here.  In-bounds aliasing
//     (a -1 coordinate borrowing from the next axis) is reproduced.
//   * BLAS::dot (BLAS.cu:98-104) is thrust::inner_product in FP32 with an unspecified tree order,
//     which cannot be reproduced; here it accumulates in double and rounds once to R, within n ulps
//     of the exactly rounded sum (pinned by tests/test_oracle_solver.py).  The order changes the
//     reference-mode CG's iteration count by a few at most.
#include <cstdint>
#include <cstddef>
#include <cstring>
#include <cmath>
#include <climits>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

static const uint64_t NULL_MAX = 0xffffffffffffffffull;   // Definitions.h:23
enum { kAir = 0, kSolid = 1, kWall = 2, kMark = 3, kFluid = 4 };  // FluidProjection.h:32-39

// Random.h:18-27 -- integer hash to [0,1].
static inline float rand01(uint32_t seed) {
    uint32_t i = (seed ^ 12345391u) * 2654435769u;
    i ^= (i << 6) ^ (i >> 26);
    i *= 2654435769u;
    i += (i << 5) ^ (i >> 12);
    return (float)i / (float)UINT_MAX;
}

template <class T> static inline T Min(const T& a, const T& b) { return (a < b) ? a : b; }   // MathUtils.h:174-178
template <class T> static inline T Max(const T& a, const T& b) { return (a > b) ? a : b; }   // MathUtils.h:195-199
template <class T> static inline T Abs(const T& x) { return (x < 0) ? -x : x; }              // MathUtils.h:129-132
template <class T> static inline T Clamp(const T x, const T lo, const T hi) {                 // MathUtils.h:286-292
    if (x < lo) return lo;
    if (x > hi) return hi;
    return x;
}

template <class R> struct V3 {
    R x, y, z;
};
struct Idx3 { int64_t i, j, k; };

struct Params {            // mirrors FLIPMethod::Params (FLIPMethod.h:51-67) + hard-coded constants
    float ext_force[3];
    int max_substeps;
    float dt;
    float wall_level;
    int enable_water_tank;
    int enable_curl;
    float damping_band;
    float voxel_size;
    float col_vel_scale;
    int pts_per_cell;
    float vort_pres_rate;
    float adjacency_under;
    int proj_iteration;
    int redist_iteration;
    float flip_ratio;       // FLIPMethod.cu:99 (0.95)
    int solver_mode;        // 0 = reference CG (LinearSystemSolver.cpp:13-50), 1 = relative-tolerance PCG
    float cg_abs_tol;       // LinearSystemSolver.cpp:39 (1e-3)
    float cg_rel_tol;       // target tolerance for solver_mode 1
    int max_iterations;     // solver_mode 1 cap
    int preconditioner;     // solver_mode 1: 0 none, 1 Jacobi
};

struct Geom {              // SparseFrame::initialize (SparseFrame.cu:72-117)
    int B;
    int fx, fy, fz;
    int64_t nx, ny, nz;     // padded dims
    float t[3], s[3];       // frame grid translate / scale
};

// SparseFrame.cu:76-116 + Grid.h:123-138: pad to a block multiple, isotropic h, recentre.
static Geom make_geom(int64_t gx, int64_t gy, int64_t gz, const float bmin[3], const float bmax[3], int B) {
    Geom g;
    g.B = B;
    g.fx = (int)((gx / B) + ((gx % B) ? 1 : 0));
    g.fy = (int)((gy / B) + ((gy % B) ? 1 : 0));
    g.fz = (int)((gz / B) + ((gz % B) ? 1 : 0));
    g.nx = (int64_t)g.fx * B;
    g.ny = (int64_t)g.fy * B;
    g.nz = (int64_t)g.fz * B;
    float cen[3], d[3];
    for (int a = 0; a < 3; ++a) cen[a] = (bmin[a] + bmax[a]) * 0.5f;
    d[0] = (bmax[0] - bmin[0]) / (float)gx;
    d[1] = (bmax[1] - bmin[1]) / (float)gy;
    d[2] = (bmax[2] - bmin[2]) / (float)gz;
    const float h = Min(Min(d[0], d[1]), d[2]);
    const float l[3] = {(float)g.nx * h, (float)g.ny * h, (float)g.nz * h};
    float mn[3], mx[3];
    for (int a = 0; a < 3; ++a) { mn[a] = cen[a] - l[a] * 0.5f; mx[a] = mn[a] + l[a]; }
    const int64_t n[3] = {g.nx, g.ny, g.nz};
    for (int a = 0; a < 3; ++a) {
        float L = mx[a] - mn[a];
        if (a == 2 && n[2] == 1) L = 1.f;
        g.t[a] = mn[a];
        g.s[a] = L / (float)n[a];
    }
    return g;
}

template <class R> struct State {
    Geom g;
    Params prm;
    R vox_force[3];
    R vox_wall;

    // particles (Particles.h:24-36), voxel space inside the step
    std::vector<V3<R>> pos, vel;
    std::vector<R> adj, vort, curv;
    std::vector<char> removed;

    // two-level frame (SparseFrame.h:18-26)
    std::vector<int> markers;
    std::vector<uint64_t> pointers;
    std::vector<Idx3> coords;

    // sparse fields (FLIPMethod.h:39-45)
    std::vector<uint64_t> mark;     // background 0
    std::vector<R> mass;            // background 0
    std::vector<R> surf;            // background 10 (count field)
    std::vector<R> pres;            // background 0
    std::vector<V3<R>> velo;        // background 0 (pre-projection P2G velocity)
    std::vector<V3<R>> comp;        // background 0 (projected velocity)
    R surf_bg = R(10);

    // linear system (FluidProjection.h:17-22)
    std::vector<uint64_t> fluid;    // fluid cell buffer (storage index per rank)
    std::vector<R> A_val;           // ELL, 7 per row
    std::vector<uint64_t> A_ind;
    std::vector<R> x, b;

    // collision source (CollisionSource.h:13-150): its own sparse frame on the same geometry, with
    // signed distance (voxel units, background col_bg), velocity and normal fields (background 0).
    // Before the first update the fields are empty and signedDistance returns 5 (CollisionSource.h:56,
    // :128-131).
    std::vector<uint64_t> col_pointers;
    std::vector<Idx3> col_coords;
    std::vector<R> col_phi;
    std::vector<V3<R>> col_vel, col_nrm;
    R col_bg = R(5);

    // stats
    int cg_iters = 0;
    double cg_res = 0;              // last recursive residual norm
    double b_norm = 0;

    // ---- geometry helpers ----
    int64_t nblk() const { return (int64_t)g.fx * g.fy * g.fz; }
    int64_t B3() const { return (int64_t)g.B * g.B * g.B; }

    // SparseField.h:89-95 (findIndex); NULL_MAX when the block is inactive
    uint64_t find_index(int64_t i, int64_t j, int64_t k) const {
        const int B = g.B;
        const uint64_t p = pointers[(k / B) * ((int64_t)g.fx * g.fy) + (j / B) * g.fx + (i / B)];
        if (p == NULL_MAX) return NULL_MAX;
        return p + (k % B) * B * B + (j % B) * B + (i % B);
    }
    // SparseField.h:103-112 (findIndices)
    void find_indices(uint64_t n, int64_t& i, int64_t& j, int64_t& k) const {
        const int B = g.B;
        const Idx3& c = coords[n / B3()];
        const uint64_t b = n % B3();
        i = c.i * B + (b) % B;
        j = c.j * B + (b / B) % B;
        k = c.k * B + (b / (B * B)) % B;
    }
    template <class T> T at(const std::vector<T>& f, const T& bg, int64_t i, int64_t j, int64_t k) const {
        const uint64_t n = find_index(i, j, k);
        return n == NULL_MAX ? bg : f[n];
    }
    // Grid.h:407-417 (inside, voxel space)
    bool inside(const V3<R>& p) const {
        if (p.x < R(0)) return false;
        if (p.x > (R)(float)g.nx) return false;
        if (p.y < R(0)) return false;
        if (p.y > (R)(float)g.ny) return false;
        if (p.z < R(0)) return false;
        if (p.z > (R)(float)g.nz) return false;
        return true;
    }

    // ---- FLIPMethod::initialize (FLIPMethod.cu:16-47) ----
    void init(const Params& p) {
        prm = p;
        for (int a = 0; a < 3; ++a) vox_force[a] = (R)prm.ext_force[a] / (R)g.s[a];   // Grid.h:452-456
        vox_wall = ((R)prm.wall_level - (R)g.t[1]) / (R)g.s[1];                         // Grid.h:445-450
    }

    // ---- SparseFrame::buildFromPoints (SparseFrame.cu:127-199) ----
    void build_from_points() {
        const int fx = g.fx, fy = g.fy, fz = g.fz, B = g.B;
        markers.assign((size_t)nblk(), 0);
#pragma omp parallel for
        for (size_t n = 0; n < pos.size(); ++n) {
            const V3<R>& p = pos[n];
            const int ci = Max((int)(p.x) / B, 0);
            const int cj = Max((int)(p.y) / B, 0);
            const int ck = Max((int)(p.z) / B, 0);
            for (int k = ck - 1; k <= Min(ck + 1, fz - 1); ++k)
                for (int j = cj - 1; j <= Min(cj + 1, fy - 1); ++j)
                    for (int i = ci - 1; i <= Min(ci + 1, fx - 1); ++i) {
                        const int64_t t = (int64_t)(k * fx * fy + j * fx + i);   // int arithmetic as in the reference
                        if (t < 0) continue;                                        // reference writes out of bounds here
#pragma omp atomic
                        markers[(size_t)t] += 1;
                    }
        }
        if (prm.enable_water_tank) {   // SparseFrame.cu:163-173
            const int tall = Clamp((int)((float)vox_wall) / B, 0, fy);
            for (int k = 0; k < fz; ++k)
                for (int j = 0; j < tall; ++j)
                    for (int i = 0; i < fx; ++i) markers[(size_t)k * fx * fy + (size_t)j * fx + i] = 1;
        }
        pointers.assign((size_t)nblk(), NULL_MAX);
        coords.clear();
        uint64_t count = 0;
        for (size_t n = 0; n < markers.size(); ++n) {
            if (markers[n] > 0) {
                pointers[n] = count;
                coords.push_back(Idx3{(int64_t)(n % fx), (int64_t)((n / fx) % fy), (int64_t)((n / ((size_t)fx * fy)) % fz)});
                count += (uint64_t)B3();
            }
        }
        // SparseField::build (SparseField.h:221-226): realloc + fill background
        const size_t nc = coords.size() * (size_t)B3();
        mark.assign(nc, 0);
        mass.assign(nc, R(0));
        surf.assign(nc, surf_bg);
        pres.assign(nc, R(0));
        velo.assign(nc, V3<R>{0, 0, 0});
        comp.assign(nc, V3<R>{0, 0, 0});
    }

    // ---- Rasterization::PointsToVelField, sparse (Rasterization.cu:121-190) ----
    // Run scatter(n) for every particle.  Sequentially in particle order, or in the OpenMP build by
    // z-slabs of kSlab cell layers of the P2G base cell: a particle writes cells of its own slab and
    // the first layer of the next, so slabs two apart never share a cell; even slabs run in parallel,
    // then odd ones, each slab's particles in particle order.
    template <class F> void for_particles_scatter(F scatter) {
#ifdef _OPENMP
        constexpr int kSlab = 4;
        const int nsl = (int)((g.nz + kSlab - 1) / kSlab);
        std::vector<int> slab(pos.size());
        std::vector<size_t> start((size_t)nsl + 1, 0);
        for (size_t n = 0; n < pos.size(); ++n) {
            const int ck = Max((int)(pos[n].z - R(0.5)), 0);
            slab[n] = inside(pos[n]) ? Min(ck / kSlab, nsl - 1) : -1;
            if (slab[n] >= 0) start[(size_t)slab[n] + 1]++;
        }
        for (int q = 0; q < nsl; ++q) start[(size_t)q + 1] += start[(size_t)q];
        std::vector<size_t> order(start[(size_t)nsl]), fill(start.begin(), start.end() - 1);
        for (size_t n = 0; n < pos.size(); ++n)
            if (slab[n] >= 0) order[fill[(size_t)slab[n]]++] = n;
        for (int parity = 0; parity < 2; ++parity) {
#pragma omp parallel for schedule(dynamic, 1)
            for (int q = parity; q < nsl; q += 2)
                for (size_t t = start[(size_t)q]; t < start[(size_t)q + 1]; ++t) scatter(order[t]);
        }
#else
        for (size_t n = 0; n < pos.size(); ++n) scatter(n);
#endif
    }

    // ---- Rasterization::PointsToVelField, sparse (Rasterization.cu:121-190) ----
    void p2g() {
        std::fill(mass.begin(), mass.end(), R(0));
        std::fill(velo.begin(), velo.end(), V3<R>{0, 0, 0});
        const int EI = (int)g.nx - 1, EJ = (int)g.ny - 1, EK = (int)g.nz - 1;
        for_particles_scatter([&](size_t n) {
            const V3<R>& p = pos[n];
            const V3<R>& v = vel[n];
            if (!inside(p)) return;
            const int ci = Max((int)(p.x - R(0.5)), 0);
            const int cj = Max((int)(p.y - R(0.5)), 0);
            const int ck = Max((int)(p.z - R(0.5)), 0);
            for (int k = ck; k <= Min(ck + 1, EK); ++k)
                for (int j = cj; j <= Min(cj + 1, EJ); ++j)
                    for (int i = ci; i <= Min(ci + 1, EI); ++i) {
                        const uint64_t idx = find_index(i, j, k);
                        if (idx == NULL_MAX) continue;
                        const R cx = (R)i + R(0.5), cy = (R)j + R(0.5), cz = (R)k + R(0.5);   // Grid.h:275-278
                        const R dx = p.x - cx, dy = p.y - cy, dz = p.z - cz;
                        const R xw = Max(R(0), R(1) - Abs(dx));
                        const R yw = Max(R(0), R(1) - Abs(dy));
                        const R zw = Max(R(0), R(1) - Abs(dz));
                        const R w = xw * yw * zw + R(1e-05);
                        mass[idx] += w;
                        velo[idx].x += w * v.x;
                        velo[idx].y += w * v.y;
                        velo[idx].z += w * v.z;
                    }
        });
#pragma omp parallel for
        for (size_t n = 0; n < mass.size(); ++n) {    // Rasterization.cu:175-188
            if (mass[n] > R(0)) {
                velo[n].x /= mass[n];
                velo[n].y /= mass[n];
                velo[n].z /= mass[n];
            } else {
                mass[n] = R(0);
                velo[n] = V3<R>{0, 0, 0};
            }
        }
    }

    // ---- Rasterization::PointsToCountField, sparse (Rasterization.cu:198-230) ----
    void count_field() {
        std::fill(surf.begin(), surf.end(), R(0));
        const int EI = (int)g.nx - 1, EJ = (int)g.ny - 1, EK = (int)g.nz - 1;
        for_particles_scatter([&](size_t n) {
            const V3<R>& p = pos[n];
            if (!inside(p)) return;
            const int ci = Max((int)(p.x - R(0.5)), 0);
            const int cj = Max((int)(p.y - R(0.5)), 0);
            const int ck = Max((int)(p.z - R(0.5)), 0);
            for (int k = ck; k <= Min(ck + 1, EK); ++k)
                for (int j = cj; j <= Min(cj + 1, EJ); ++j)
                    for (int i = ci; i <= Min(ci + 1, EI); ++i) {
                        const uint64_t idx = find_index(i, j, k);
                        if (idx == NULL_MAX) continue;
                        surf[idx] += R(1);
                    }
        });
    }

    // ---- cell classification, FLIPMethod.cu:242-280 (no collision source: phi = 5, v = 0,
    //      CollisionSource.h:128-144 with an empty field) ----
    // ---- FLIPMethod::advanceOneStep mm_kernel (FLIPMethod.cu:242-280), curl disabled ----
    void classify() {
        const R scale = (R)prm.col_vel_scale;
#pragma omp parallel for
        for (size_t n = 0; n < mark.size(); ++n) {
            int64_t i, j, k;
            find_indices(n, i, j, k);
            const V3<R> cell{(R)((float)i + 0.5f), (R)((float)j + 0.5f), (R)((float)k + 0.5f)};   // Grid.h:274-277
            const R phi = signed_distance(cell);
            const V3<R> v = col_velocity(cell);
            if (mass[n] > R(0)) mark[n] = kFluid;
            else mark[n] = (phi < R(0)) ? kSolid : kAir;
            if (phi < R(0)) {
                velo[n].x += v.x * scale;
                velo[n].y += v.y * scale;
                velo[n].z += v.z * scale;
            }
            comp[n] = velo[n];
        }
    }

    // ---- collision source queries (CollisionSource.h:128-144) ----
    uint64_t col_index(int64_t i, int64_t j, int64_t k) const {
        const int B = g.B;
        const uint64_t p = col_pointers[(k / B) * ((int64_t)g.fx * g.fy) + (j / B) * g.fx + (i / B)];
        if (p == NULL_MAX) return NULL_MAX;
        return p + (k % B) * B * B + (j % B) * B + (i % B);
    }
    R signed_distance(const V3<R>& p) const {
        if (col_phi.empty()) return col_bg;
        int64_t i, j, k;
        R x, y, z;
        lerp_corner(p, i, j, k, x, y, z);
        auto F = [&](int64_t a, int64_t b2, int64_t c) {
            const uint64_t n = col_index(a, b2, c);
            return n == NULL_MAX ? col_bg : col_phi[n];
        };
        const R c00 = F(i, j, k) * (R(1) - x) + F(i + 1, j, k) * x;
        const R c01 = F(i, j, k + 1) * (R(1) - x) + F(i + 1, j, k + 1) * x;
        const R c10 = F(i, j + 1, k) * (R(1) - x) + F(i + 1, j + 1, k) * x;
        const R c11 = F(i, j + 1, k + 1) * (R(1) - x) + F(i + 1, j + 1, k + 1) * x;
        const R c0 = c00 * (R(1) - y) + c10 * y;
        const R c1 = c01 * (R(1) - y) + c11 * y;
        return c0 * (R(1) - z) + c1 * z;
    }
    V3<R> col_lerp(const std::vector<V3<R>>& f, const V3<R>& p) const {
        if (f.empty()) return V3<R>{0, 0, 0};
        int64_t i, j, k;
        R x, y, z;
        lerp_corner(p, i, j, k, x, y, z);
        auto F = [&](int64_t a, int64_t b2, int64_t c) {
            const uint64_t n = col_index(a, b2, c);
            return n == NULL_MAX ? V3<R>{0, 0, 0} : f[n];
        };
        auto L = [&](const V3<R>& a, const V3<R>& c, R t) {
            return V3<R>{a.x * (R(1) - t) + c.x * t, a.y * (R(1) - t) + c.y * t, a.z * (R(1) - t) + c.z * t};
        };
        const V3<R> c00 = L(F(i, j, k), F(i + 1, j, k), x);
        const V3<R> c01 = L(F(i, j, k + 1), F(i + 1, j, k + 1), x);
        const V3<R> c10 = L(F(i, j + 1, k), F(i + 1, j + 1, k), x);
        const V3<R> c11 = L(F(i, j + 1, k + 1), F(i + 1, j + 1, k + 1), x);
        return L(L(c00, c10, y), L(c01, c11, y), z);
    }
    V3<R> col_velocity(const V3<R>& p) const { return col_lerp(col_vel, p); }
    V3<R> col_normal(const V3<R>& p) const { return col_lerp(col_nrm, p); }
    // Vector3::normalized (Vector3.h:354-358): d = 1/Sqrt(x^2+y^2+z^2+EPSILON), EPSILON = 1e-30f
    static V3<R> normalized(const V3<R>& v) {
        const R d = R(1) / std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z + R(1e-30f));
        return V3<R>{v.x * d, v.y * d, v.z * d};
    }

    // CollisionSource::update (CollisionSource.h:66-126).  The OpenVDB inputs are replaced by
    // `seeds` (world positions of the source's active voxels) and a sampler returning the signed
    // distance, velocity and gradient (world units) at a world point.
    template <class Sampler> void set_collision(const std::vector<float>& seeds, float bg_world, Sampler sample) {
        const float h = g.s[0];                                  // CollisionSource.h:53-54 (Lx / nx)
        const int fx = g.fx, fy = g.fy, fz = g.fz, B = g.B;
        std::vector<int> mk((size_t)nblk(), 0);
        for (size_t q = 0; q + 2 < seeds.size(); q += 3) {
            float vp[3];                                          // Grid::voxelPoint (Grid.h:445-450)
            for (int a = 0; a < 3; ++a) vp[a] = (seeds[q + a] - g.t[a]) / g.s[a];
            const int ci = Max((int)(vp[0]) / B, 0), cj = Max((int)(vp[1]) / B, 0), ck = Max((int)(vp[2]) / B, 0);
            for (int k = ck - 1; k <= Min(ck + 1, fz - 1); ++k)       // SparseFrame.cu:145-157
                for (int j = cj - 1; j <= Min(cj + 1, fy - 1); ++j)
                    for (int i = ci - 1; i <= Min(ci + 1, fx - 1); ++i) {
                        const int64_t t = (int64_t)(k * fx * fy + j * fx + i);
                        if (t < 0) continue;
                        mk[(size_t)t] += 1;
                    }
        }
        col_pointers.assign((size_t)nblk(), NULL_MAX);
        col_coords.clear();
        uint64_t count = 0;
        for (size_t n = 0; n < mk.size(); ++n)
            if (mk[n] > 0) {
                col_pointers[n] = count;
                col_coords.push_back(Idx3{(int64_t)(n % fx), (int64_t)((n / fx) % fy), (int64_t)(n / ((size_t)fx * fy))});
                count += (uint64_t)B3();
            }
        col_bg = (R)(bg_world / h);                               // CollisionSource.h:90
        const size_t nc = col_coords.size() * (size_t)B3();
        col_phi.assign(nc, col_bg);
        col_vel.assign(nc, V3<R>{0, 0, 0});
        col_nrm.assign(nc, V3<R>{0, 0, 0});
        for (size_t n = 0; n < nc; ++n) {
            const Idx3& c = col_coords[n / B3()];
            const uint64_t b = n % B3();
            const int64_t i = c.i * B + b % B, j = c.j * B + (b / B) % B, k = c.k * B + b / (B * B);
            const float ctr[3] = {(float)i + 0.5f, (float)j + 0.5f, (float)k + 0.5f};
            float wp[3];                                          // Grid::worldPoint (Grid.h:431-436)
            for (int a = 0; a < 3; ++a) wp[a] = ctr[a] * g.s[a] + g.t[a];
            float sdf = 0.f, v[3] = {0, 0, 0}, gr[3] = {0, 0, 0};
            sample(wp, &sdf, v, gr);
            col_phi[n] = (R)sdf / (R)h;                           // :108
            col_vel[n] = V3<R>{(R)v[0] / (R)g.s[0], (R)v[1] / (R)g.s[1], (R)v[2] / (R)g.s[2]};   // :113 voxelVector
            col_nrm[n] = normalized(V3<R>{(R)gr[0] / (R)g.s[0], (R)gr[1] / (R)g.s[1], (R)gr[2] / (R)g.s[2]});   // :120
        }
    }

    // ---- FluidProjection::buildLinearSystem, sparse (FluidProjection.cu:166-274) ----
    void build_linear_system() {
        fluid.clear();
        uint64_t count = 0;
        for (size_t n = 0; n < mark.size(); ++n) {
            if (mark[n] >= kFluid) {
                mark[n] = count + kFluid;
                fluid.push_back(n);
                count++;
            }
        }
        const int64_t EI = g.nx - 1, EJ = g.ny - 1, EK = g.nz - 1;
        A_val.assign(count * 7, R(0));
        A_ind.assign(count * 7, NULL_MAX);
        x.assign(count, R(0));
        b.assign(count, R(0));
        const float wallLv = (float)vox_wall;
        auto T = [&](int64_t i, int64_t j, int64_t k) { return at(mark, (uint64_t)0, i, j, k); };
        auto Vf = [&](int64_t i, int64_t j, int64_t k) { return at(comp, V3<R>{0, 0, 0}, i, j, k); };
#pragma omp parallel for
        for (uint64_t n = 0; n < count; ++n) {
            int64_t i, j, k;
            find_indices(fluid[n], i, j, k);
            const uint64_t wall = ((float)j + 0.5f <= wallLv + 1.f) ? (uint64_t)kWall : (uint64_t)kAir;
            const uint64_t t_cc = T(i, j, k);
            const uint64_t t_i0 = i == 0 ? wall : T(i - 1, j, k);
            const uint64_t t_i1 = i == EI ? wall : T(i + 1, j, k);
            const uint64_t t_j0 = j == 0 ? wall : T(i, j - 1, k);
            const uint64_t t_j1 = j == EJ ? wall : T(i, j + 1, k);
            const uint64_t t_k0 = k == 0 ? wall : T(i, j, k - 1);
            const uint64_t t_k1 = k == EK ? wall : T(i, j, k + 1);
            const V3<R> v = Vf(i, j, k);
            const R i0_j_k = i == 0 ? R(2) * v.x - Vf(i + 1, j, k).x : Vf(i - 1, j, k).x;
            const R i1_j_k = i == EI ? R(2) * v.x - Vf(i - 1, j, k).x : Vf(i + 1, j, k).x;
            const R i_j0_k = j == 0 ? R(2) * v.y - Vf(i, j + 1, k).y : Vf(i, j - 1, k).y;
            const R i_j1_k = j == EJ ? R(2) * v.y - Vf(i, j - 1, k).y : Vf(i, j + 1, k).y;
            const R i_j_k0 = k == 0 ? R(2) * v.z - Vf(i, j, k + 1).z : Vf(i, j, k - 1).z;
            const R i_j_k1 = k == EK ? R(2) * v.z - Vf(i, j, k - 1).z : Vf(i, j, k + 1).z;
            const R v_i0 = t_i0 == kAir ? R(2) * v.x - i1_j_k : t_i0 == kWall ? Abs(v.x) : i0_j_k;
            const R v_i1 = t_i1 == kAir ? R(2) * v.x - i0_j_k : t_i1 == kWall ? -Abs(v.x) : i1_j_k;
            const R v_j0 = t_j0 == kAir ? R(2) * v.y - i_j1_k : t_j0 == kWall ? Abs(v.y) : i_j0_k;
            const R v_j1 = t_j1 == kAir ? R(2) * v.y - i_j0_k : t_j1 == kWall ? -Abs(v.y) : i_j1_k;
            const R v_k0 = t_k0 == kAir ? R(2) * v.z - i_j_k1 : t_k0 == kWall ? Abs(v.z) : i_j_k0;
            const R v_k1 = t_k1 == kAir ? R(2) * v.z - i_j_k0 : t_k1 == kWall ? -Abs(v.z) : i_j_k1;
            R p = R(6), pn[6] = {0, 0, 0, 0, 0, 0};
            const R d = -((v_i1 - v_i0) * R(0.5) + (v_j1 - v_j0) * R(0.5) + (v_k1 - v_k0) * R(0.5));
            const uint64_t tn[6] = {t_i0, t_i1, t_j0, t_j1, t_k0, t_k1};
            for (int q = 0; q < 6; ++q) {
                if (tn[q] == kAir) p += R(1);
                else if (tn[q] == kSolid || tn[q] == kWall) p -= R(1);
                else if (tn[q] >= kFluid) pn[q] -= R(1);
            }
            R* vals = &A_val[n * 7];
            uint64_t* inds = &A_ind[n * 7];
            vals[0] = p;
            for (int q = 0; q < 6; ++q) vals[q + 1] = pn[q];
            inds[0] = t_cc - kFluid;
            for (int q = 0; q < 6; ++q) inds[q + 1] = tn[q] >= kFluid ? tn[q] - kFluid : NULL_MAX;
            x[n] = at(pres, R(0), i, j, k);
            b[n] = d;
        }
    }

    // ---- BLAS (BLAS.cu:12-209) ----
    void spmv(const std::vector<R>& p, std::vector<R>& out) const {     // BLAS::mul, BLAS.cu:12-40
        const size_t rows = p.size();
#pragma omp parallel for
        for (size_t r = 0; r < rows; ++r) {
            R acc = R(0);
            for (int c = 0; c < 7; ++c) {
                const uint64_t n = A_ind[r * 7 + c];
                if (n != NULL_MAX) acc += A_val[r * 7 + c] * p[n];
            }
            out[r] = acc;
        }
    }
    static R dot(const std::vector<R>& a, const std::vector<R>& c) {  // BLAS::dot, BLAS.cu:98-104
        double s = 0.0;
#pragma omp parallel for reduction(+ : s)
        for (size_t n = 0; n < a.size(); ++n) s += (double)a[n] * (double)c[n];
        return (R)s;
    }

    // LinearSystemSolver::cg (LinearSystemSolver.cpp:13-50)
    int cg_reference(int max) {
        const size_t N = b.size();
        std::vector<R> r(N), p(N), Ap(N);
        spmv(x, Ap);
#pragma omp parallel for
        for (size_t n = 0; n < N; ++n) r[n] = b[n] - Ap[n];      // BLAS::submul(b,A,x,r)
        p = r;                                                    // BLAS::equ
        R rsold = dot(r, r);
        b_norm = std::sqrt((double)dot(b, b));
        cg_res = std::sqrt((double)rsold);
        if (std::sqrt(rsold) < (R)1e-30f) return 0;
        const int iter = (int)N;
        int count = 0;
        for (int i = 0; i < iter + 5; ++i, ++count) {
            spmv(p, Ap);
            const R alpha = rsold / dot(p, Ap);
#pragma omp parallel for
            for (size_t n = 0; n < N; ++n) x[n] = x[n] + (alpha * p[n]);
#pragma omp parallel for
            for (size_t n = 0; n < N; ++n) r[n] = r[n] - (alpha * Ap[n]);
            const R rsnew = dot(r, r);
            cg_res = std::sqrt((double)rsnew);
            if (std::sqrt(rsnew) < (R)prm.cg_abs_tol) break;
            const R beta = rsnew / rsold;
#pragma omp parallel for
            for (size_t n = 0; n < N; ++n) p[n] = r[n] + (beta * p[n]);
            rsold = rsnew;
            if (count > max) break;
        }
        return count;
    }

    // Relative-tolerance PCG (the build's target mode; LinearSystemSolver.cpp:52-105 is the
    // reference's dead pcg()).  Returns the number of A*p products performed.
    int pcg_relative(int max) {
        const size_t N = b.size();
        std::vector<R> r(N), p(N), z(N), Ap(N), dinv(N);
        for (size_t n = 0; n < N; ++n) dinv[n] = prm.preconditioner == 1 ? R(1) / A_val[n * 7] : R(1);
        spmv(x, Ap);
#pragma omp parallel for
        for (size_t n = 0; n < N; ++n) r[n] = b[n] - Ap[n];
        b_norm = std::sqrt((double)dot(b, b));
        R rr = dot(r, r);
        cg_res = std::sqrt((double)rr);
        const R tol = (R)prm.cg_rel_tol * (R)b_norm;
        if (std::sqrt(rr) <= tol || b_norm == 0.0) return 0;
#pragma omp parallel for
        for (size_t n = 0; n < N; ++n) z[n] = dinv[n] * r[n];
        p = z;
        R rz = dot(r, z);
        int it = 0;
        while (it < max) {
            spmv(p, Ap);
            const R alpha = rz / dot(p, Ap);
#pragma omp parallel for
            for (size_t n = 0; n < N; ++n) x[n] = x[n] + (alpha * p[n]);
#pragma omp parallel for
            for (size_t n = 0; n < N; ++n) r[n] = r[n] - (alpha * Ap[n]);
            ++it;
            rr = dot(r, r);
            cg_res = std::sqrt((double)rr);
            if (std::sqrt(rr) <= tol) break;
#pragma omp parallel for
            for (size_t n = 0; n < N; ++n) z[n] = dinv[n] * r[n];
            const R rz_new = dot(r, z);
            const R beta = rz_new / rz;
#pragma omp parallel for
            for (size_t n = 0; n < N; ++n) p[n] = z[n] + (beta * p[n]);
            rz = rz_new;
        }
        return it;
    }

    // ---- FluidProjection::solve, sparse (FluidProjection.cu:277-381) ----
    void solve_and_apply() {
        cg_iters = prm.solver_mode == 0 ? cg_reference(prm.proj_iteration) : pcg_relative(prm.max_iterations);
        apply_pressure();
    }

    // K10-K12 on the current solution x (FluidProjection.cu:297-378)
    void apply_pressure() {
#pragma omp parallel for
        for (size_t n = 0; n < fluid.size(); ++n) {               // K10, :297-306
            const uint64_t idx = fluid[n];
            pres[idx] = x[mark[idx] - kFluid];
        }
        const int64_t EI = g.nx - 1, EJ = g.ny - 1, EK = g.nz - 1;
        auto T = [&](int64_t i, int64_t j, int64_t k) { return at(mark, (uint64_t)0, i, j, k); };
        auto P = [&](int64_t i, int64_t j, int64_t k) { return at(pres, R(0), i, j, k); };
#pragma omp parallel for
        for (size_t n = 0; n < mark.size(); ++n) {                // K11, :314-335
            if (mark[n] != kAir) continue;
            int64_t i, j, k;
            find_indices(n, i, j, k);
            R press = R(0);
            int cnt = 0;
            if (i > 0 && T(i - 1, j, k) >= kFluid) { press += P(i - 1, j, k); cnt++; }
            if (i < EI && T(i + 1, j, k) >= kFluid) { press += P(i + 1, j, k); cnt++; }
            if (j > 0 && T(i, j - 1, k) >= kFluid) { press += P(i, j - 1, k); cnt++; }
            if (j < EJ && T(i, j + 1, k) >= kFluid) { press += P(i, j + 1, k); cnt++; }
            if (k > 0 && T(i, j, k - 1) >= kFluid) { press += P(i, j, k - 1); cnt++; }
            if (k < EK && T(i, j, k + 1) >= kFluid) { press += P(i, j, k + 1); cnt++; }
            if (cnt > 0) pres[n] = press / (R)cnt;
        }
        const float wallLv = (float)vox_wall;
#pragma omp parallel for
        for (size_t t = 0; t < fluid.size(); ++t) {               // K12, :342-378
            const uint64_t n = fluid[t];
            int64_t i, j, k;
            find_indices(n, i, j, k);
            const uint64_t wall = ((float)j + 0.5f <= wallLv + 1.f) ? (uint64_t)kSolid : (uint64_t)kAir;
            const R p = pres[n];
            const uint64_t t_i0 = i == 0 ? wall : T(i - 1, j, k);
            const uint64_t t_i1 = i == EI ? wall : T(i + 1, j, k);
            const uint64_t t_j0 = j == 0 ? wall : T(i, j - 1, k);
            const uint64_t t_j1 = j == EJ ? wall : T(i, j + 1, k);
            const uint64_t t_k0 = k == 0 ? wall : T(i, j, k - 1);
            const uint64_t t_k1 = k == EK ? wall : T(i, j, k + 1);
            const R i0_j_k = i == 0 ? R(2) * p - P(i + 1, j, k) : P(i - 1, j, k);
            const R i1_j_k = i == EI ? R(2) * p - P(i - 1, j, k) : P(i + 1, j, k);
            const R i_j0_k = j == 0 ? R(2) * p - P(i, j + 1, k) : P(i, j - 1, k);
            const R i_j1_k = j == EJ ? R(2) * p - P(i, j - 1, k) : P(i, j + 1, k);
            const R i_j_k0 = k == 0 ? R(2) * p - P(i, j, k + 1) : P(i, j, k - 1);
            const R i_j_k1 = k == EK ? R(2) * p - P(i, j, k - 1) : P(i, j, k + 1);
            const R p_i0 = t_i0 == kSolid ? R(2) * p - i1_j_k : t_i0 == kAir ? -i0_j_k : i0_j_k;
            const R p_i1 = t_i1 == kSolid ? R(2) * p - i0_j_k : t_i1 == kAir ? -i1_j_k : i1_j_k;
            const R p_j0 = t_j0 == kSolid ? R(2) * p - i_j1_k : t_j0 == kAir ? -i_j0_k : i_j0_k;
            const R p_j1 = t_j1 == kSolid ? R(2) * p - i_j0_k : t_j1 == kAir ? -i_j1_k : i_j1_k;
            const R p_k0 = t_k0 == kSolid ? R(2) * p - i_j_k1 : t_k0 == kAir ? -i_j_k0 : i_j_k0;
            const R p_k1 = t_k1 == kSolid ? R(2) * p - i_j_k0 : t_k1 == kAir ? -i_j_k1 : i_j_k1;
            comp[n].x -= (p_i1 - p_i0) * R(0.5);
            comp[n].y -= (p_j1 - p_j0) * R(0.5);
            comp[n].z -= (p_k1 - p_k0) * R(0.5);
        }
    }

    // SparseField::lerp / lerpMin (SparseField.h:150-196)
    void lerp_corner(const V3<R>& p, int64_t& i, int64_t& j, int64_t& k, R& x, R& y, R& z) const {
        i = Clamp((int)(p.x - R(0.5)), 0, (int)g.nx - 2);
        j = Clamp((int)(p.y - R(0.5)), 0, (int)g.ny - 2);
        k = Clamp((int)(p.z - R(0.5)), 0, (int)g.nz - 2);
        x = Clamp(p.x - ((R)i + R(0.5)), R(0), R(1));
        y = Clamp(p.y - ((R)j + R(0.5)), R(0), R(1));
        z = Clamp(p.z - ((R)k + R(0.5)), R(0), R(1));
    }
    V3<R> lerp_vec(const std::vector<V3<R>>& f, const V3<R>& p) const {
        int64_t i, j, k;
        R x, y, z;
        lerp_corner(p, i, j, k, x, y, z);
        const V3<R> bg{0, 0, 0};
        auto F = [&](int64_t a, int64_t b2, int64_t c) { return at(f, bg, a, b2, c); };
        auto L = [&](const V3<R>& a, const V3<R>& c, R t) {
            return V3<R>{a.x * (R(1) - t) + c.x * t, a.y * (R(1) - t) + c.y * t, a.z * (R(1) - t) + c.z * t};
        };
        const V3<R> c00 = L(F(i, j, k), F(i + 1, j, k), x);
        const V3<R> c01 = L(F(i, j, k + 1), F(i + 1, j, k + 1), x);
        const V3<R> c10 = L(F(i, j + 1, k), F(i + 1, j + 1, k), x);
        const V3<R> c11 = L(F(i, j + 1, k + 1), F(i + 1, j + 1, k + 1), x);
        const V3<R> c0 = L(c00, c10, y);
        const V3<R> c1 = L(c01, c11, y);
        return L(c0, c1, z);
    }
    R lerp_min(const std::vector<R>& f, R bg, const V3<R>& p) const {
        int64_t i, j, k;
        R x, y, z;
        lerp_corner(p, i, j, k, x, y, z);
        auto F = [&](int64_t a, int64_t b2, int64_t c) { return at(f, bg, a, b2, c); };
        R v = Min(F(i, j, k), F(i + 1, j, k));
        v = Min(v, Min(F(i, j, k + 1), F(i + 1, j, k + 1)));
        v = Min(v, Min(F(i, j + 1, k), F(i + 1, j + 1, k)));
        v = Min(v, Min(F(i, j + 1, k + 1), F(i + 1, j + 1, k + 1)));
        return v;
    }

    // ---- FLIPMethod::updateParticles (FLIPMethod.cu:54-175) ----
    void update_particles(R dt) {
        const R lx = (R)((float)g.nx + 0.5f), lz = (R)((float)g.nz + 0.5f);
        const bool tank = prm.enable_water_tank != 0;
        const R band = (R)(prm.damping_band / prm.voxel_size);
        const R wall = vox_wall;
        // FLIPMethod.cu:99 uses the literal pair 0.95f / 0.05f
        const R fr = prm.flip_ratio == 0.95f ? R(0.95) : (R)prm.flip_ratio;
        const R pr = prm.flip_ratio == 0.95f ? R(0.05) : R(1) - fr;
#pragma omp parallel for
        for (size_t ix = 0; ix < pos.size(); ++ix) {
            V3<R>& ps = pos[ix];
            V3<R>& v = vel[ix];
            adj[ix] = lerp_min(surf, surf_bg, ps);
            {
                const V3<R> gv = lerp_vec(comp, ps);
                const V3<R> pv = lerp_vec(velo, ps);
                const V3<R> f{gv.x - pv.x, gv.y - pv.y, gv.z - pv.z};
                v.x += f.x; v.y += f.y; v.z += f.z;
                v.x = v.x * fr + gv.x * pr;
                v.y = v.y * fr + gv.y * pr;
                v.z = v.z * fr + gv.z * pr;
            }
            if (tank && ps.y <= wall) {
                R w = R(1);
                if (ps.x < band) w = Min(w, Clamp((ps.x - R(0)) / band, R(0), R(1)));
                if (ps.x > lx - band) w = Min(w, Clamp((lx - ps.x) / band, R(0), R(1)));
                if (ps.z < band) w = Min(w, Clamp((ps.z - R(0)) / band, R(0), R(1)));
                if (ps.z > lz - band) w = Min(w, Clamp((lz - ps.z) / band, R(0), R(1)));
                v.x *= w; v.y *= w; v.z *= w;
            }
            if (signed_distance(ps) < R(0)) {                      // :116-119
                const V3<R> cv = col_velocity(ps);
                ps.x += cv.x * dt;
                ps.y += cv.y * dt;
                ps.z += cv.z * dt;
            }
            {
                const R len = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
                const int iter = 1 + Min((int)(len * dt * R(1.5)), Max(prm.max_substeps, 1));
                const R sub_dt = dt / (R)iter;
                bool col = false;
                for (int s = 0; s < iter; ++s) {
                    ps.x += v.x * sub_dt;
                    ps.y += v.y * sub_dt;
                    ps.z += v.z * sub_dt;
                    if (ps.y <= wall) {
                        if (ps.x < R(0)) { ps.x = R(0); if (v.x < R(0)) v.x = R(0); }
                        if (ps.y < R(0)) { ps.y = R(0); if (v.y < R(0)) v.y = R(0); }
                        if (ps.z < R(0)) { ps.z = R(0); if (v.z < R(0)) v.z = R(0); }
                        if (ps.x > (R)(float)g.nx) { ps.x = (R)(float)g.nx; if (v.x > R(0)) v.x = R(0); }
                        if (ps.y > (R)(float)g.ny) { ps.y = (R)(float)g.ny; if (v.y > R(0)) v.y = R(0); }
                        if (ps.z > (R)(float)g.nz) { ps.z = (R)(float)g.nz; if (v.z > R(0)) v.z = R(0); }
                    }
                    const R phi = signed_distance(ps);                  // :146-157
                    if (phi < R(0)) {
                        const V3<R> nor = normalized(col_normal(ps));
                        ps.x -= nor.x * phi;
                        ps.y -= nor.y * phi;
                        ps.z -= nor.z * phi;
                        col = true;
                    }
                }
                if (col) {                                              // :161-164
                    const R phi = signed_distance(ps);
                    if (phi > R(3) || phi < R(0)) removed[ix] = 1;
                }
                if (!inside(ps)) removed[ix] = 1;
            }
            v.x += vox_force[0] * dt;
            v.y += vox_force[1] * dt;
            v.z += vox_force[2] * dt;
        }
    }

    // ---- Array::remove (Array.h:995-1024), literal loop ----
    template <class T> static void remove_literal(std::vector<T>& data, const std::vector<char>& mask) {
        size_t size = data.size();
        if (size == 0) return;
        size_t end = size - 1;
        size_t del = 0;
        for (size_t n = 0; n < mask.size(); ++n) {
            if (n >= size) break;
            if (mask[n]) {
                del++;
                for (size_t m = end; m >= n + 1; --m) {
                    if (mask[m] == 0) {
                        data[n] = data[m];
                        end = m - 1;
                        break;
                    }
                }
            }
        }
        data.resize(size - del);
    }

    // ---- FLIPMethod::postProcess (FLIPMethod.cu:300-405) ----
    // Scalar lerp with a background (SparseField::lerp, SparseField.h:150-173)
    R lerp_scalar(const std::vector<R>& f, R bg, const V3<R>& p) const {
        int64_t i, j, k;
        R x, y, z;
        lerp_corner(p, i, j, k, x, y, z);
        auto F = [&](int64_t a, int64_t b2, int64_t c) { return at(f, bg, a, b2, c); };
        const R c00 = F(i, j, k) * (R(1) - x) + F(i + 1, j, k) * x;
        const R c01 = F(i, j, k + 1) * (R(1) - x) + F(i + 1, j, k + 1) * x;
        const R c10 = F(i, j + 1, k) * (R(1) - x) + F(i + 1, j + 1, k) * x;
        const R c11 = F(i, j + 1, k + 1) * (R(1) - x) + F(i + 1, j + 1, k + 1) * x;
        const R c0 = c00 * (R(1) - y) + c10 * y;
        const R c1 = c01 * (R(1) - y) + c11 * y;
        return c0 * (R(1) - z) + c1 * z;
    }

    std::vector<R> splat_w;          // the weight sums of the splat (the reference keeps them in _sPresField)

    // Rasterization::ParticlesToSurfaceAndVelocity (Rasterization.cu:311-393), influence 1.5, droplet
    // 0.5: surf <- |weighted mean position - cell centre| - r (5 where no weight), velo <- weighted
    // mean velocity (0 where no weight); the weighted mean position lands in comp.
    void splat_surface_velocity(R influence, R droplet) {
        const size_t nc = surf.size();
        std::vector<R> sw(nc, R(0));
        std::fill(surf.begin(), surf.end(), R(0));
        std::fill(comp.begin(), comp.end(), V3<R>{0, 0, 0});
        std::fill(velo.begin(), velo.end(), V3<R>{0, 0, 0});
        const int off = (int)influence;                           // neighborCells(p, min, max, (int)h)
        for (size_t n = 0; n < pos.size(); ++n) {
            const V3<R>& p = pos[n];
            const V3<R>& v = vel[n];
            if (!inside(p)) continue;
            const int64_t mn[3] = {Clamp((int)p.x - off, 0, (int)g.nx - 1), Clamp((int)p.y - off, 0, (int)g.ny - 1),
                                   Clamp((int)p.z - off, 0, (int)g.nz - 1)};
            const int64_t mx[3] = {Clamp((int)p.x + off, 0, (int)g.nx - 1), Clamp((int)p.y + off, 0, (int)g.ny - 1),
                                   Clamp((int)p.z + off, 0, (int)g.nz - 1)};
            for (int64_t k = mn[2]; k <= mx[2]; ++k)
                for (int64_t j = mn[1]; j <= mx[1]; ++j)
                    for (int64_t i = mn[0]; i <= mx[0]; ++i) {
                        const R cx = (R)((float)i + 0.5f), cy = (R)((float)j + 0.5f), cz = (R)((float)k + 0.5f);
                        const R dx = p.x - cx, dy = p.y - cy, dz = p.z - cz;
                        const R dist = std::sqrt(dx * dx + dy * dy + dz * dz);     // Vector3::length
                        const R sr = dist / influence;
                        const R t = R(1) - sr * sr;
                        const R w = Max(R(0), t * t * t);                          // Pow3
                        const uint64_t idx = find_index(i, j, k);
                        if (idx == NULL_MAX) continue;
                        sw[idx] += w;
                        velo[idx].x += w * v.x;
                        velo[idx].y += w * v.y;
                        velo[idx].z += w * v.z;
                        surf[idx] += w;
                        comp[idx].x += w * p.x;
                        comp[idx].y += w * p.y;
                        comp[idx].z += w * p.z;
                    }
        }
        for (size_t n = 0; n < nc; ++n) {                          // kernel2, :375-392
            if (surf[n] > R(1e-30f)) {
                int64_t i, j, k;
                find_indices(n, i, j, k);
                comp[n].x /= surf[n];
                comp[n].y /= surf[n];
                comp[n].z /= surf[n];
                const R ex = comp[n].x - (R)((float)i + 0.5f), ey = comp[n].y - (R)((float)j + 0.5f),
                        ez = comp[n].z - (R)((float)k + 0.5f);
                surf[n] = std::sqrt(ex * ex + ey * ey + ez * ez) - droplet;       // Vector3::distanceTo
            } else {
                surf[n] = R(5);
            }
            if (sw[n] > R(1e-30f)) {
                velo[n].x /= sw[n];
                velo[n].y /= sw[n];
                velo[n].z /= sw[n];
            } else {
                velo[n] = V3<R>{0, 0, 0};
            }
        }
        splat_w = sw;
    }

    // FLIPMethod.cu:320-363 (water tank): air cells of the surface field below the tank level that
    // are not connected to the level's layer (flood fill from it, host-serial in the reference as
    // here) become liquid (-5).
    void tank_flood_fill() {
        const int64_t nx = g.nx, ny = g.ny, nz = g.nz;
        std::vector<char> scan((size_t)(nx * ny * nz), 0);
        auto S = [&](int64_t i, int64_t j, int64_t k) -> char& { return scan[(size_t)(i + nx * (j + ny * k))]; };
        auto F = [&](int64_t i, int64_t j, int64_t k) { return at(surf, surf_bg, i, j, k); };
        std::vector<Idx3> cells;
        cells.push_back(Idx3{nx - 1, ny - 1, nz - 1});
        const int64_t wall = (int64_t)(size_t)(float)vox_wall;
        if (wall >= 0 && wall < ny)                               // (the reference reads out of range otherwise)
            for (int64_t k = 0; k < nz; ++k)
                for (int64_t i = 0; i < nx; ++i)
                    if (F(i, wall, k) > R(0)) cells.push_back(Idx3{i, wall, k});
        while (!cells.empty()) {
            const Idx3 c = cells.back();
            cells.pop_back();
            S(c.i, c.j, c.k) = 1;
            if (c.j > wall) continue;
            const int64_t d[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
            for (int q = 0; q < 6; ++q) {
                const int64_t i = c.i + d[q][0], j = c.j + d[q][1], k = c.k + d[q][2];
                if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) continue;
                if (F(i, j, k) > R(0) && S(i, j, k) == 0) cells.push_back(Idx3{i, j, k});
            }
        }
        for (size_t n = 0; n < surf.size(); ++n) {                // mm_kernel, :351-360
            int64_t i, j, k;
            find_indices(n, i, j, k);
            if (j > wall) continue;
            if (surf[n] > R(0) && S(i, j, k) == 0) surf[n] = R(-5);
        }
    }

    // Rasterization::ReinitializeSurface (Rasterization.cu:395-490): `iter` Godunov steps of
    // d(phi)/dt = S(phi) (1 - |grad phi|), dt 0.3, both fields with background 10.
    void reinitialize_surface(int iter, bool tank, R waterlevel) {
        const int64_t EI = g.nx - 1, EJ = g.ny - 1, EK = g.nz - 1;
        const R dt = R(0.3f);
        std::vector<R> tmp(surf.size());
        for (int it = 0; it < iter; ++it) {
            auto F = [&](int64_t i, int64_t j, int64_t k) { return at(surf, R(10), i, j, k); };
            for (size_t n = 0; n < surf.size(); ++n) {
                int64_t i, j, k;
                find_indices(n, i, j, k);
                const R v = surf[n];
                const R sign = v / std::sqrt(v * v + R(1));
                R v_i0, v_i1, v_j0, v_j1, v_k0, v_k1;
                if (tank) {
                    const R ll = (R)((float)j + 0.5f) - waterlevel;
                    v_j0 = j == 0 ? ll - R(1) : F(i, j - 1, k);
                    v_j1 = j == EJ ? ll + R(1) : F(i, j + 1, k);
                    v_i0 = i == 0 ? ll : F(i - 1, j, k);
                    v_i1 = i == EI ? ll : F(i + 1, j, k);
                    v_k0 = k == 0 ? ll : F(i, j, k - 1);
                    v_k1 = k == EK ? ll : F(i, j, k + 1);
                } else {
                    v_i0 = i == 0 ? R(2) * v - F(i + 1, j, k) : F(i - 1, j, k);
                    v_i1 = i == EI ? R(2) * v - F(i - 1, j, k) : F(i + 1, j, k);
                    v_j0 = j == 0 ? R(2) * v - F(i, j + 1, k) : F(i, j - 1, k);
                    v_j1 = j == EJ ? R(2) * v - F(i, j - 1, k) : F(i, j + 1, k);
                    v_k0 = k == 0 ? R(2) * v - F(i, j, k + 1) : F(i, j, k - 1);
                    v_k1 = k == EK ? R(2) * v - F(i, j, k - 1) : F(i, j, k + 1);
                }
                const R dx0 = v - v_i0, dx1 = v_i1 - v;
                const R dy0 = v - v_j0, dy1 = v_j1 - v;
                const R dz0 = v - v_k0, dz1 = v_k1 - v;
                auto P2 = [](R a) { return a * a; };
                const R gp = std::sqrt(P2(Max(dx0, R(0))) + P2(Min(dx1, R(0))) + P2(Max(dy0, R(0))) + P2(Min(dy1, R(0))) +
                                       P2(Max(dz0, R(0))) + P2(Min(dz1, R(0))));
                const R gm = std::sqrt(P2(Min(dx0, R(0))) + P2(Max(dx1, R(0))) + P2(Min(dy0, R(0))) + P2(Max(dy1, R(0))) +
                                       P2(Min(dz0, R(0))) + P2(Max(dz1, R(0))));
                tmp[n] = v - dt * Max(sign, R(0)) * (gp - R(1)) - dt * Min(sign, R(0)) * (gm - R(1));
            }
            surf.swap(tmp);
        }
    }

    // FieldOP::curl (FieldOP.cu:338-369): central differences, domain faces extrapolated (2v - v_in).
    void curl_field(const std::vector<V3<R>>& vf, std::vector<V3<R>>& out) const {
        const int64_t EI = g.nx - 1, EJ = g.ny - 1, EK = g.nz - 1;
        const V3<R> bg{0, 0, 0};
        auto F = [&](int64_t i, int64_t j, int64_t k) { return at(vf, bg, i, j, k); };
        auto ext = [](const V3<R>& v, const V3<R>& o) { return V3<R>{v.x * R(2) - o.x, v.y * R(2) - o.y, v.z * R(2) - o.z}; };
        out.assign(vf.size(), bg);
        for (size_t n = 0; n < vf.size(); ++n) {
            int64_t i, j, k;
            find_indices(n, i, j, k);
            const V3<R> v = vf[n];
            const V3<R> x0 = i == 0 ? ext(v, F(i + 1, j, k)) : F(i - 1, j, k);
            const V3<R> x1 = i == EI ? ext(v, F(i - 1, j, k)) : F(i + 1, j, k);
            const V3<R> y0 = j == 0 ? ext(v, F(i, j + 1, k)) : F(i, j - 1, k);
            const V3<R> y1 = j == EJ ? ext(v, F(i, j - 1, k)) : F(i, j + 1, k);
            const V3<R> z0 = k == 0 ? ext(v, F(i, j, k + 1)) : F(i, j, k - 1);
            const V3<R> z1 = k == EK ? ext(v, F(i, j, k - 1)) : F(i, j, k + 1);
            out[n] = V3<R>{(y1.z - y0.z) * R(0.5) - (z1.y - z0.y) * R(0.5), (z1.x - z0.x) * R(0.5) - (x1.z - x0.z) * R(0.5),
                           (x1.y - x0.y) * R(0.5) - (y1.x - y0.x) * R(0.5)};
        }
    }

    // FieldOP::normalGradient (FieldOP.cu:80-110): normalised central-difference gradient.
    void normal_gradient(const std::vector<R>& sf, R bg, std::vector<V3<R>>& out) const {
        const int64_t EI = g.nx - 1, EJ = g.ny - 1, EK = g.nz - 1;
        auto F = [&](int64_t i, int64_t j, int64_t k) { return at(sf, bg, i, j, k); };
        out.assign(sf.size(), V3<R>{0, 0, 0});
        for (size_t n = 0; n < sf.size(); ++n) {
            int64_t i, j, k;
            find_indices(n, i, j, k);
            const R v = sf[n];
            const R i0 = i == 0 ? R(2) * v - F(i + 1, j, k) : F(i - 1, j, k);
            const R i1 = i == EI ? R(2) * v - F(i - 1, j, k) : F(i + 1, j, k);
            const R j0 = j == 0 ? R(2) * v - F(i, j + 1, k) : F(i, j - 1, k);
            const R j1 = j == EJ ? R(2) * v - F(i, j - 1, k) : F(i, j + 1, k);
            const R k0 = k == 0 ? R(2) * v - F(i, j, k + 1) : F(i, j, k - 1);
            const R k1 = k == EK ? R(2) * v - F(i, j, k - 1) : F(i, j, k + 1);
            out[n] = normalized(V3<R>{(i1 - i0) * R(0.5), (j1 - j0) * R(0.5), (k1 - k0) * R(0.5)});
        }
    }

    // FieldOP::curvature (FieldOP.cu:273-301): central-difference divergence of a vector field.
    void divergence_of(const std::vector<V3<R>>& nf, std::vector<R>& out) const {
        const int64_t EI = g.nx - 1, EJ = g.ny - 1, EK = g.nz - 1;
        const V3<R> bg{0, 0, 0};
        auto F = [&](int64_t i, int64_t j, int64_t k) { return at(nf, bg, i, j, k); };
        out.assign(nf.size(), R(0));
        for (size_t n = 0; n < nf.size(); ++n) {
            int64_t i, j, k;
            find_indices(n, i, j, k);
            const V3<R> s = nf[n];
            const R i0 = i == 0 ? s.x * R(2) - F(i + 1, j, k).x : F(i - 1, j, k).x;
            const R i1 = i == EI ? s.x * R(2) - F(i - 1, j, k).x : F(i + 1, j, k).x;
            const R j0 = j == 0 ? s.y * R(2) - F(i, j + 1, k).y : F(i, j - 1, k).y;
            const R j1 = j == EJ ? s.y * R(2) - F(i, j - 1, k).y : F(i, j + 1, k).y;
            const R k0 = k == 0 ? s.z * R(2) - F(i, j, k + 1).z : F(i, j, k - 1).z;
            const R k1 = k == EK ? s.z * R(2) - F(i, j, k - 1).z : F(i, j, k + 1).z;
            out[n] = (i1 - i0) * R(0.5) + (j1 - j0) * R(0.5) + (k1 - k0) * R(0.5);
        }
    }

    // FLIPMethod::postProcess (FLIPMethod.cu:300-405), voxel space; dt = params.dt.
    void post_process() {
        build_from_points();                                        // prepareNextStep, :302
        splat_surface_velocity(R(1.5f), R(0.5f));                   // :318
        post_after_splat();
    }

    // FLIPMethod.cu:320-404: everything after the splat, on the current surf and velo fields.
    void post_after_splat() {
        if (prm.enable_water_tank) tank_flood_fill();               // :320-363
        reinitialize_surface(prm.redist_iteration, prm.enable_water_tank != 0, vox_wall);   // :365
        curl_field(velo, comp);                                     // :367
        const R dis = (R)std::pow(1.f - prm.vort_pres_rate, prm.dt); // :373
        for (size_t n = 0; n < pos.size(); ++n) {                   // vort_kernel, :375-385
            const V3<R> c = lerp_vec(comp, pos[n]);
            const V3<R> wv{c.x * (R)g.s[0], c.y * (R)g.s[1], c.z * (R)g.s[2]};   // Grid::worldVector
            R v = std::sqrt(wv.x * wv.x + wv.y * wv.y + wv.z * wv.z);
            v = Max(R(0), v - R(0.2f));
            vort[n] -= vort[n] * dis;
            vort[n] = Max(v, vort[n]);
        }
        normal_gradient(surf, surf_bg, comp);                       // :390
        divergence_of(comp, pres);                                  // :391
        for (size_t n = 0; n < pos.size(); ++n) curv[n] = lerp_scalar(pres, R(10), pos[n]);   // :393-398, pres bg 10
        std::fill(pres.begin(), pres.end(), R(0));                  // :402-404
    }

    // ---- FLIPMethod::advanceOneStep (FLIPMethod.cu:208-298) ----
    int step(R dt) {
        build_from_points();                        // prepareNextStep, :49-52 / :210
        removed.assign(pos.size(), 0);              // :231-232
        p2g();                                      // :234
        count_field();                              // :235
        classify();                                 // :242-281
        build_linear_system();                      // :286
        solve_and_apply();                          // :287
        update_particles(dt);                       // :290
        remove_literal(pos, removed);               // :293-297
        remove_literal(vel, removed);
        remove_literal(vort, removed);
        remove_literal(curv, removed);
        remove_literal(adj, removed);
        return cg_iters;
    }
};

struct Handle {
    int fp64;
    State<float>* s32;
    State<double>* s64;
};

}  // namespace orc

using namespace orc;

// ------------------------------------------------------------------------------------------------
// C ABI for the test harness (ctypes).  Field ids: 0 mass, 1 velo(3), 2 comp(3), 3 mark (as the
// reference's size_t mark: 0 air, 1 solid, 4+rank fluid), 4 pres, 5 surf/count, 6 b per fluid rank,
// 7 x per fluid rank, 8 fluid cell buffer (storage index per rank).
extern "C" {

typedef orc::Params orc_params;

struct orc_stats {
    int64_t num_particles, num_blocks, num_cells, num_fluid;
    int cg_iters;
    int pad;
    double cg_residual, b_norm;
};

float orc_rand(uint32_t seed) { return rand01(seed); }

// The solver's dot product (BLAS::dot, BLAS.cu:98-104) of the FP32 and FP64 instances, for its pins.
float orc_dot_f32(const float* a, const float* c, size_t n) {
    return orc::State<float>::dot(std::vector<float>(a, a + n), std::vector<float>(c, c + n));
}
double orc_dot_f64(const double* a, const double* c, size_t n) {
    return orc::State<double>::dot(std::vector<double>(a, a + n), std::vector<double>(c, c + n));
}

// Threads of the OpenMP build (returns the count in use; 1 in the checker build).
int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

void orc_geom(const int64_t dims[3], const float bmin[3], const float bmax[3], int B, int64_t out_dims[3], int out_f[3],
              float out_t[3], float out_s[3]) {
    Geom g = make_geom(dims[0], dims[1], dims[2], bmin, bmax, B);
    out_dims[0] = g.nx; out_dims[1] = g.ny; out_dims[2] = g.nz;
    out_f[0] = g.fx; out_f[1] = g.fy; out_f[2] = g.fz;
    for (int a = 0; a < 3; ++a) { out_t[a] = g.t[a]; out_s[a] = g.s[a]; }
}

// Grid::initialize(voxelSize, min, max) (Grid.h:143-158): truncating cell count.
void orc_grid_from_voxel_size(float voxel, const float mn[3], const float mx[3], int64_t out_dims[3]) {
    out_dims[0] = (int64_t)(size_t)((mx[0] - mn[0]) / voxel);
    out_dims[1] = (int64_t)(size_t)((mx[1] - mn[1]) / voxel);
    out_dims[2] = (int64_t)(size_t)((mx[2] - mn[2]) / voxel);
}

void* orc_create(const int64_t dims[3], const float bmin[3], const float bmax[3], int B, const orc_params* prm, int fp64) {
    Handle* h = new Handle();
    h->fp64 = fp64;
    h->s32 = nullptr;
    h->s64 = nullptr;
    Geom g = make_geom(dims[0], dims[1], dims[2], bmin, bmax, B);
    if (fp64) { h->s64 = new State<double>(); h->s64->g = g; h->s64->init(*prm); }
    else { h->s32 = new State<float>(); h->s32->g = g; h->s32->init(*prm); }
    return h;
}

void orc_destroy(void* hp) {
    Handle* h = (Handle*)hp;
    delete h->s32;
    delete h->s64;
    delete h;
}

extern "C++" {
template <class R>
static void set_particles_t(State<R>* s, const float* pos, const float* vel, const float* vort, const float* curv,
                            const float* adj, size_t n) {
    s->pos.resize(n); s->vel.resize(n); s->vort.assign(n, R(0)); s->curv.assign(n, R(0)); s->adj.assign(n, R(0));
    for (size_t i = 0; i < n; ++i) {
        s->pos[i] = V3<R>{(R)pos[3 * i], (R)pos[3 * i + 1], (R)pos[3 * i + 2]};
        s->vel[i] = V3<R>{(R)vel[3 * i], (R)vel[3 * i + 1], (R)vel[3 * i + 2]};
        if (vort) s->vort[i] = (R)vort[i];
        if (curv) s->curv[i] = (R)curv[i];
        if (adj) s->adj[i] = (R)adj[i];
    }
}
}  // extern "C++"

void orc_set_particles(void* hp, const float* pos, const float* vel, const float* vort, const float* curv, const float* adj,
                       size_t n) {
    Handle* h = (Handle*)hp;
    if (h->fp64) set_particles_t(h->s64, pos, vel, vort, curv, adj, n);
    else set_particles_t(h->s32, pos, vel, vort, curv, adj, n);
}

extern "C++" {
template <class R>
static void get_particles_t(State<R>* s, double* pos, double* vel, double* adj, double* vort, double* curv) {
    for (size_t i = 0; i < s->pos.size(); ++i) {
        if (pos) { pos[3 * i] = s->pos[i].x; pos[3 * i + 1] = s->pos[i].y; pos[3 * i + 2] = s->pos[i].z; }
        if (vel) { vel[3 * i] = s->vel[i].x; vel[3 * i + 1] = s->vel[i].y; vel[3 * i + 2] = s->vel[i].z; }
        if (adj) adj[i] = s->adj[i];
        if (vort) vort[i] = s->vort[i];
        if (curv) curv[i] = s->curv[i];
    }
}
}  // extern "C++"

void orc_get_particles(void* hp, double* pos, double* vel, double* adj, double* vort, double* curv) {
    Handle* h = (Handle*)hp;
    if (h->fp64) get_particles_t(h->s64, pos, vel, adj, vort, curv);
    else get_particles_t(h->s32, pos, vel, adj, vort, curv);
}

size_t orc_num_particles(void* hp) {
    Handle* h = (Handle*)hp;
    return h->fp64 ? h->s64->pos.size() : h->s32->pos.size();
}

int orc_step(void* hp, float dt) {
    Handle* h = (Handle*)hp;
    return h->fp64 ? h->s64->step((double)dt) : h->s32->step(dt);
}

// Stage-level entry points (same order as orc_step); 0 prepare, 1 p2g+count+classify,
// 2 build system, 3 solve+apply, 4 update particles, 5 remove.
void orc_stage(void* hp, int stage, float dt) {
    Handle* h = (Handle*)hp;
    auto run = [&](auto* s) {
        using R = typename std::remove_reference<decltype(s->mass[0])>::type;
        switch (stage) {
            case 0: s->build_from_points(); s->removed.assign(s->pos.size(), 0); break;
            case 1: s->p2g(); s->count_field(); s->classify(); break;
            case 2: s->build_linear_system(); break;
            case 3: s->solve_and_apply(); break;
            case 4: s->update_particles((R)dt); break;
            case 5:
                s->remove_literal(s->pos, s->removed); s->remove_literal(s->vel, s->removed);
                s->remove_literal(s->vort, s->removed); s->remove_literal(s->curv, s->removed);
                s->remove_literal(s->adj, s->removed);
                break;
        }
    };
    if (h->fp64) run(h->s64); else run(h->s32);
}

extern "C++" {
template <class R>
static size_t block_table_t(State<R>* s, int64_t* pointers, int32_t* coords) {
    if (pointers)
        for (size_t n = 0; n < s->pointers.size(); ++n)
            pointers[n] = s->pointers[n] == NULL_MAX ? -1 : (int64_t)s->pointers[n];
    if (coords)
        for (size_t n = 0; n < s->coords.size(); ++n) {
            coords[3 * n] = (int32_t)s->coords[n].i;
            coords[3 * n + 1] = (int32_t)s->coords[n].j;
            coords[3 * n + 2] = (int32_t)s->coords[n].k;
        }
    return s->coords.size();
}
}  // extern "C++"

size_t orc_get_block_table(void* hp, int64_t* pointers, int32_t* coords) {
    Handle* h = (Handle*)hp;
    return h->fp64 ? block_table_t(h->s64, pointers, coords) : block_table_t(h->s32, pointers, coords);
}

void orc_get_markers(void* hp, int32_t* out) {
    Handle* h = (Handle*)hp;
    const std::vector<int>& m = h->fp64 ? h->s64->markers : h->s32->markers;
    std::memcpy(out, m.data(), m.size() * sizeof(int));
}

extern "C++" {
template <class R>
static size_t field_t(State<R>* s, int which, double* out) {
    switch (which) {
        case 0: if (out) for (size_t n = 0; n < s->mass.size(); ++n) out[n] = s->mass[n]; return s->mass.size();
        case 1: if (out) for (size_t n = 0; n < s->velo.size(); ++n) { out[3*n] = s->velo[n].x; out[3*n+1] = s->velo[n].y; out[3*n+2] = s->velo[n].z; } return s->velo.size() * 3;
        case 2: if (out) for (size_t n = 0; n < s->comp.size(); ++n) { out[3*n] = s->comp[n].x; out[3*n+1] = s->comp[n].y; out[3*n+2] = s->comp[n].z; } return s->comp.size() * 3;
        case 3: if (out) for (size_t n = 0; n < s->mark.size(); ++n) out[n] = (double)s->mark[n]; return s->mark.size();
        case 4: if (out) for (size_t n = 0; n < s->pres.size(); ++n) out[n] = s->pres[n]; return s->pres.size();
        case 5: if (out) for (size_t n = 0; n < s->surf.size(); ++n) out[n] = s->surf[n]; return s->surf.size();
        case 6: if (out) for (size_t n = 0; n < s->b.size(); ++n) out[n] = s->b[n]; return s->b.size();
        case 7: if (out) for (size_t n = 0; n < s->x.size(); ++n) out[n] = s->x[n]; return s->x.size();
        case 8: if (out) for (size_t n = 0; n < s->fluid.size(); ++n) out[n] = (double)s->fluid[n]; return s->fluid.size();
        case 9: if (out) for (size_t n = 0; n < s->A_val.size(); ++n) out[n] = s->A_val[n]; return s->A_val.size();
        case 10: if (out) for (size_t n = 0; n < s->A_ind.size(); ++n) out[n] = s->A_ind[n] == NULL_MAX ? -1.0 : (double)s->A_ind[n]; return s->A_ind.size();
    }
    return 0;
}
}  // extern "C++"

size_t orc_get_field(void* hp, int which, double* out) {
    Handle* h = (Handle*)hp;
    return h->fp64 ? field_t(h->s64, which, out) : field_t(h->s32, which, out);
}

// Overwrite a field (testing a stage on given inputs): 0 mass, 1 velo, 2 comp, 3 mark (cell types
// 0 air / 1 solid / 4 fluid before build_linear_system numbers the fluid cells), 4 pres, 5 surf (the
// count field), 7 x (the solution per fluid rank, for apply_pressure).
void orc_set_field(void* hp, int which, const double* in) {
    Handle* h = (Handle*)hp;
    auto run = [&](auto* s) {
        using R = typename std::remove_reference<decltype(s->mass[0])>::type;
        if (which == 1) for (size_t n = 0; n < s->velo.size(); ++n) s->velo[n] = {(R)in[3*n], (R)in[3*n+1], (R)in[3*n+2]};
        if (which == 2) for (size_t n = 0; n < s->comp.size(); ++n) s->comp[n] = {(R)in[3*n], (R)in[3*n+1], (R)in[3*n+2]};
        if (which == 0) for (size_t n = 0; n < s->mass.size(); ++n) s->mass[n] = (R)in[n];
        if (which == 3) for (size_t n = 0; n < s->mark.size(); ++n) s->mark[n] = (uint64_t)in[n];
        if (which == 4) for (size_t n = 0; n < s->pres.size(); ++n) s->pres[n] = (R)in[n];
        if (which == 5) for (size_t n = 0; n < s->surf.size(); ++n) s->surf[n] = (R)in[n];
        if (which == 7) for (size_t n = 0; n < s->x.size(); ++n) s->x[n] = (R)in[n];
    };
    if (h->fp64) run(h->s64); else run(h->s32);
}

// FLIPMethod::postProcess (FLIPMethod.cu:300-405) on the current (voxel-space) particles.
void orc_post_process(void* hp) {
    Handle* h = (Handle*)hp;
    if (h->fp64) h->s64->post_process(); else h->s32->post_process();
}

// Single post-process parts on the current fields (pins): 0 splat, 1 tank flood fill, 2 one
// redistancing iteration (arg = tank flag), 3 curl velo -> comp, 4 normal gradient surf -> comp,
// 5 divergence comp -> pres, 6 the whole post-process after the splat (FLIPMethod.cu:320-404).
void orc_post_part(void* hp, int part, int arg) {
    Handle* h = (Handle*)hp;
    auto run = [&](auto* s) {
        using R = typename std::remove_reference<decltype(s->mass[0])>::type;
        switch (part) {
            case 0: s->splat_surface_velocity(R(1.5f), R(0.5f)); break;
            case 1: s->tank_flood_fill(); break;
            case 2: s->reinitialize_surface(1, arg != 0, s->vox_wall); break;
            case 3: { std::vector<V3<R>> o; s->curl_field(s->velo, o); s->comp = o; } break;
            case 4: { std::vector<V3<R>> o; s->normal_gradient(s->surf, s->surf_bg, o); s->comp = o; } break;
            case 5: s->divergence_of(s->comp, s->pres); break;
            case 6: s->post_after_splat(); break;
        }
    };
    if (h->fp64) run(h->s64); else run(h->s32);
}

// Pressure scatter, air extrapolation and gradient subtract (FluidProjection.cu:297-378) on the
// current x, without solving (x set by orc_set_field(7) after build_linear_system).
void orc_apply_pressure(void* hp) {
    Handle* h = (Handle*)hp;
    if (h->fp64) h->s64->apply_pressure(); else h->s32->apply_pressure();
}

size_t orc_get_removed(void* hp, uint8_t* out) {
    Handle* h = (Handle*)hp;
    const std::vector<char>& r = h->fp64 ? h->s64->removed : h->s32->removed;
    if (out) for (size_t n = 0; n < r.size(); ++n) out[n] = (uint8_t)r[n];
    return r.size();
}

void orc_get_stats(void* hp, orc_stats* st) {
    Handle* h = (Handle*)hp;
    auto run = [&](auto* s) {
        st->num_particles = (int64_t)s->pos.size();
        st->num_blocks = (int64_t)s->coords.size();
        st->num_cells = (int64_t)s->mass.size();
        st->num_fluid = (int64_t)s->fluid.size();
        st->cg_iters = s->cg_iters;
        st->cg_residual = s->cg_res;
        st->b_norm = s->b_norm;
    };
    if (h->fp64) run(h->s64); else run(h->s32);
}

// ---- collision source ----
// Axis-aligned box [lo,hi] (world): exact signed distance (negative inside) and its gradient,
// evaluated in FP64 and rounded once: outside, |max(q,0)| with q = |p-c| - e per axis and the
// gradient along sign(p-c)*max(q,0); inside, max(q) (<= 0) with the gradient along the axis of the
// largest q (first of x, y, z on ties).
static void box_sample(const float lo[3], const float hi[3], const float p[3], float* sdf, float grad[3]) {
    double d[3], q[3], o[3];
    for (int a = 0; a < 3; ++a) {
        const double c = 0.5 * ((double)lo[a] + (double)hi[a]);
        const double e = 0.5 * ((double)hi[a] - (double)lo[a]);
        d[a] = (double)p[a] - c;
        q[a] = std::fabs(d[a]) - e;
        o[a] = q[a] > 0.0 ? q[a] : 0.0;
    }
    const double out = std::sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
    const double qm = std::max(q[0], std::max(q[1], q[2]));
    *sdf = (float)(out + std::min(qm, 0.0));
    if (out > 0.0) {
        for (int a = 0; a < 3; ++a) grad[a] = (float)((d[a] < 0.0 ? -o[a] : o[a]) / out);
    } else {
        const int ax = q[0] >= q[1] && q[0] >= q[2] ? 0 : (q[1] >= q[2] ? 1 : 2);
        for (int a = 0; a < 3; ++a) grad[a] = a == ax ? (d[a] < 0.0 ? -1.f : 1.f) : 0.f;
    }
}

// Box obstacle moving with velocity vel (world units): the source's active voxels are the cells
// whose centre lies within `band` voxels of the surface (|sdf|/h < band, as in a narrow-band level
// set of half-width `band`); the background is +band voxels.
void orc_set_collision_box(void* hp, const float lo[3], const float hi[3], const float vel[3], float band) {
    Handle* h = (Handle*)hp;
    auto run = [&](auto* s) {
        const Geom& g = s->g;
        const float hv = g.s[0];
        std::vector<float> seeds;
        for (int64_t k = 0; k < g.nz; ++k)
            for (int64_t j = 0; j < g.ny; ++j)
                for (int64_t i = 0; i < g.nx; ++i) {
                    const float ctr[3] = {(float)i + 0.5f, (float)j + 0.5f, (float)k + 0.5f};
                    float wp[3], sdf, gr[3];
                    for (int a = 0; a < 3; ++a) wp[a] = ctr[a] * g.s[a] + g.t[a];
                    box_sample(lo, hi, wp, &sdf, gr);
                    if (std::fabs(sdf / hv) < band)
                        for (int a = 0; a < 3; ++a) seeds.push_back(wp[a]);
                }
        s->set_collision(seeds, band * hv, [&](const float* wp, float* sdf, float* v, float* gr) {
            box_sample(lo, hi, wp, sdf, gr);
            for (int a = 0; a < 3; ++a) v[a] = vel[a];
        });
    };
    if (h->fp64) run(h->s64); else run(h->s32);
}

// General source: seeds (world xyz, n points) and a sampler callback.
typedef void (*orc_sampler)(void* user, const float* wp, float* sdf, float* vel, float* grad);
void orc_set_collision(void* hp, const float* seeds, size_t n, float bg_world, orc_sampler fn, void* user) {
    Handle* h = (Handle*)hp;
    std::vector<float> sd(seeds, seeds + 3 * n);
    auto cb = [&](const float* wp, float* sdf, float* v, float* gr) { fn(user, wp, sdf, v, gr); };
    if (h->fp64) h->s64->set_collision(sd, bg_world, cb); else h->s32->set_collision(sd, bg_world, cb);
}

// Collision frame + fields: pointers[fx*fy*fz] (-1 inactive), phi[nc], vel[3nc], nrm[3nc]; returns nc.
size_t orc_get_collision(void* hp, int64_t* pointers, double* phi, double* vel, double* nrm, double* bg) {
    Handle* h = (Handle*)hp;
    auto run = [&](auto* s) {
        if (pointers)
            for (size_t n = 0; n < s->col_pointers.size(); ++n)
                pointers[n] = s->col_pointers[n] == NULL_MAX ? -1 : (int64_t)s->col_pointers[n];
        for (size_t n = 0; n < s->col_phi.size(); ++n) {
            if (phi) phi[n] = s->col_phi[n];
            if (vel) { vel[3 * n] = s->col_vel[n].x; vel[3 * n + 1] = s->col_vel[n].y; vel[3 * n + 2] = s->col_vel[n].z; }
            if (nrm) { nrm[3 * n] = s->col_nrm[n].x; nrm[3 * n + 1] = s->col_nrm[n].y; nrm[3 * n + 2] = s->col_nrm[n].z; }
        }
        if (bg) *bg = s->col_bg;
        return s->col_phi.size();
    };
    return h->fp64 ? run(h->s64) : run(h->s32);
}

// Signed distance, velocity and normal lerps at voxel-space points (CollisionSource.h:128-144).
void orc_collision_query(void* hp, const float* pts, size_t n, double* phi, double* vel, double* nrm) {
    Handle* h = (Handle*)hp;
    auto run = [&](auto* s) {
        using R = typename std::remove_reference<decltype(s->mass.data()[0])>::type;
        for (size_t q = 0; q < n; ++q) {
            const V3<R> p{(R)pts[3 * q], (R)pts[3 * q + 1], (R)pts[3 * q + 2]};
            if (phi) phi[q] = s->signed_distance(p);
            const V3<R> v = s->col_velocity(p), nr = s->col_normal(p);
            if (vel) { vel[3 * q] = v.x; vel[3 * q + 1] = v.y; vel[3 * q + 2] = v.z; }
            if (nrm) { nrm[3 * q] = nr.x; nrm[3 * q + 1] = nr.y; nrm[3 * q + 2] = nr.z; }
        }
    };
    if (h->fp64) run(h->s64); else run(h->s32);
}

// Array::remove applied to the identity: out_src[n] = original index now stored at slot n.
size_t orc_remove_perm(const uint8_t* mask, size_t n, int64_t* out_src) {
    std::vector<int64_t> idx(n);
    for (size_t i = 0; i < n; ++i) idx[i] = (int64_t)i;
    std::vector<char> m(mask, mask + n);
    State<float>::remove_literal(idx, m);
    std::memcpy(out_src, idx.data(), idx.size() * sizeof(int64_t));
    return idx.size();
}

// Dam-break seeding in voxel space: for every cell whose centre lies inside the box
// [0,wx)x[0,wy)x[0,wz) (voxel units), ppc points cell + (Rand(s+3q), Rand(s+3q+1), Rand(s+3q+2)),
// q = ppc*(i + nx*(j + ny*k)) + n; a point is kept only if it lies inside the box.
size_t orc_seed_box(int64_t nx, int64_t ny, int64_t nz, float wx, float wy, float wz, int ppc, uint32_t seed,
                    float* pos, size_t cap) {
    size_t cnt = 0;
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                if (!((float)i + 0.5f < wx && (float)j + 0.5f < wy && (float)k + 0.5f < wz)) continue;
                const uint64_t cell = (uint64_t)i + (uint64_t)nx * ((uint64_t)j + (uint64_t)ny * (uint64_t)k);
                for (int n = 0; n < ppc; ++n) {
                    const uint32_t q = (uint32_t)((uint64_t)ppc * cell + (uint64_t)n);
                    const float px = (float)i + rand01(seed + 3u * q);
                    const float py = (float)j + rand01(seed + 3u * q + 1u);
                    const float pz = (float)k + rand01(seed + 3u * q + 2u);
                    if (!(px < wx && py < wy && pz < wz)) continue;
                    if (pos && cnt < cap) { pos[3 * cnt] = px; pos[3 * cnt + 1] = py; pos[3 * cnt + 2] = pz; }
                    cnt++;
                }
            }
    return cnt;
}

}  // extern "C"
