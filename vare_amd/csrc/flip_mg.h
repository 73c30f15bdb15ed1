// flip_mg.h -- geometric multigrid V-cycle used as the preconditioner of the pressure PCG
// (VARE_PRECOND_MG).  The reference's only preconditioner is a host-serial MIC(0) that is never
// built (BLAS.cu:211-331, FluidProjection.cu:273); this one is designed for the GPU layout.
//
// Level hierarchy (all levels discretise the same 7-point operator as the fine system):
//   * sparse levels: the active blocks of the frame with cell edge E = B, B/2, ..., 1 per block,
//     stored slot-major like the fine cells (cell (slot, local) at slot*E^3 + local);
//   * dense levels: the block frame itself halved repeatedly ((fx+1)/2, ...), one array each;
//   * the coarsest dense level (<= kMgCoarseMax cells) gets a fixed number of Jacobi sweeps, and the
//     last dense levels (at most kMgBottomCells cells together) run inside one workgroup
//     (k_mg_bottom).
// A coarse cell is fluid when at least half of its children are (k_mg_flag); its diagonal is 6 + nAir - nWall counted on the
// coarse grid with the fine level's boundary rule (FluidProjection.cu:208: a domain face is a wall
// below the wall level).  Each level solves stencil_l z_l = rt_l where rt_{l+1} = 0.5 * sum of the
// 8 children's residuals (restriction R = P^T/8 times the 4x operator scaling), prolongation is
// piecewise constant, smoothing is weighted Jacobi with the same weight and sweep count before and
// after the coarse correction, so the V-cycle is a symmetric positive-definite preconditioner.
#pragma once
#include "flip_device.h"

namespace vare {

constexpr int kMgCoarseMax = 512;

struct MgLevelK {
    int dense;          // 0: sparse level of edge 1<<sh; 1: dense grid
    int sh;             // sparse: log2(edge)
    int dx, dy, dz;     // dense: grid dims; sparse: block frame dims (fx, fy, fz)
    int ncell;
    int scale_sh;       // log2 of this level's cell size in fine cells
    float* r;           // right-hand side rt_l (level 0: the CG residual)
    uint16_t* info;     // kInfoFluid | fluid-neighbour mask | diag << 8
};

struct MgFrame {
    const int* slot;    // [fx*fy*fz]
    const int* coord;   // [Nb]
    const int* nbr;     // [Nb*6]
    float wall_level;
};

// ---- indexing ----
__device__ __forceinline__ void mg_coords(const MgLevelK& L, const MgFrame& F, int c, int& i, int& j, int& k) {
    if (L.dense) {
        i = c % L.dx;
        j = (c / L.dx) % L.dy;
        k = c / (L.dx * L.dy);
    } else {
        const int sh = L.sh;
        const int s = c >> (3 * sh);
        const int l = c & ((1 << (3 * sh)) - 1);
        const int t = F.coord[s];
        const int bi = t % L.dx, bj = (t / L.dx) % L.dy, bk = t / (L.dx * L.dy);
        const int m = (1 << sh) - 1;
        i = (bi << sh) + (l & m);
        j = (bj << sh) + ((l >> sh) & m);
        k = (bk << sh) + (l >> (2 * sh));
    }
}

// Neighbour index in direction q: >= 0 cell, -1 inactive (non-fluid), -2 outside the domain.
__device__ __forceinline__ int mg_nbr(const MgLevelK& L, const MgFrame& F, int c, int q) {
    const int axis = q >> 1;
    const int dir = (q & 1) ? 1 : -1;
    if (L.dense) {
        int i = c % L.dx, j = (c / L.dx) % L.dy, k = c / (L.dx * L.dy);
        if (axis == 0) i += dir; else if (axis == 1) j += dir; else k += dir;
        if (i < 0 || i >= L.dx || j < 0 || j >= L.dy || k < 0 || k >= L.dz) return -2;
        return (k * L.dy + j) * L.dx + i;
    }
    const int sh = L.sh;
    const int E = 1 << sh;
    const int s = c >> (3 * sh);
    const int l = c & ((1 << (3 * sh)) - 1);
    const int lc = axis == 0 ? (l & (E - 1)) : (axis == 1 ? ((l >> sh) & (E - 1)) : (l >> (2 * sh)));
    const int stride = axis == 0 ? 1 : (axis == 1 ? E : E * E);
    const int nl = lc + dir;
    if (nl >= 0 && nl < E) return c + dir * stride;
    const int ns = F.nbr[s * 6 + q];
    if (ns < 0) return ns;
    return (ns << (3 * sh)) + l - dir * (E - 1) * stride;
}

// Parent (coarse level C) of fine cell c (level L).
__device__ __forceinline__ int mg_parent(const MgLevelK& L, const MgLevelK& C, const MgFrame& F, int c) {
    if (!L.dense && L.sh > 0) {
        const int sh = L.sh, csh = sh - 1, E = 1 << sh;
        const int s = c >> (3 * sh);
        const int l = c & ((1 << (3 * sh)) - 1);
        const int li = (l & (E - 1)) >> 1, lj = ((l >> sh) & (E - 1)) >> 1, lk = (l >> (2 * sh)) >> 1;
        return (s << (3 * csh)) + (((lk << csh) + lj) << csh) + li;
    }
    int i, j, k;
    if (L.dense) { i = c % L.dx; j = (c / L.dx) % L.dy; k = c / (L.dx * L.dy); }
    else { const int t = F.coord[c]; i = t % L.dx; j = (t / L.dx) % L.dy; k = t / (L.dx * L.dy); }
    return ((k >> 1) * C.dy + (j >> 1)) * C.dx + (i >> 1);
}

// Child a (0..7, x fastest) of coarse cell cc (level C) on fine level L; -1 if absent.
__device__ __forceinline__ int mg_child(const MgLevelK& L, const MgLevelK& C, const MgFrame& F, int cc, int a) {
    const int ai = a & 1, aj = (a >> 1) & 1, ak = a >> 2;
    if (!L.dense && L.sh > 0) {
        const int sh = L.sh, csh = sh - 1, Ec = 1 << csh;
        const int s = cc >> (3 * csh);
        const int l = cc & ((1 << (3 * csh)) - 1);
        const int li = l & (Ec - 1), lj = (l >> csh) & (Ec - 1), lk = l >> (2 * csh);
        return (s << (3 * sh)) + ((((2 * lk + ak) << sh) + (2 * lj + aj)) << sh) + (2 * li + ai);
    }
    const int I = cc % C.dx, J = (cc / C.dx) % C.dy, K = cc / (C.dx * C.dy);
    const int i = 2 * I + ai, j = 2 * J + aj, k = 2 * K + ak;
    if (i >= L.dx || j >= L.dy || k >= L.dz) return -1;
    if (L.dense) return (k * L.dy + j) * L.dx + i;
    return F.slot[(k * L.dy + j) * L.dx + i];      // sparse edge-1 level: cell index = block slot
}

__device__ __forceinline__ bool mg_fluid(unsigned info) { return (info & kInfoFluid) != 0; }
__device__ __forceinline__ float mg_diag(unsigned info) { return (float)((info >> 8) & 0x1f); }

// Jacobi weight omega / D of a cell, as omega times the hardware reciprocal of its integer diagonal.
// The smoother belongs to the preconditioner, not to the discretised system, so it only has to be a
// fixed symmetric operator: the reciprocal is deterministic and the same at every use of a cell, so
// the V-cycle stays symmetric, and it costs one instruction where a correctly rounded division
// costs about ten.
__device__ __forceinline__ float mg_wd(unsigned info, float omega) { return omega * __builtin_amdgcn_rcpf(mg_diag(info)); }

// Whether child a of coarse cell cc lies inside the fine level's grid (an inactive block inside the
// frame is inside: it is air; a position past an odd-sized grid's last cell is outside the domain).
__device__ __forceinline__ bool mg_child_inside(const MgLevelK& L, const MgLevelK& C, int cc, int a) {
    if (!L.dense && L.sh > 0) return true;
    const int I = cc % C.dx, J = (cc / C.dx) % C.dy, K = cc / (C.dx * C.dy);
    return 2 * I + (a & 1) < L.dx && 2 * J + ((a >> 1) & 1) < L.dy && 2 * K + (a >> 2) < L.dz;
}

// ---- set-up (once per solve) ----
// Coarse fluid flag: at least half of the children inside the grid are fluid.  Coarse cells that
// are mostly air stay air (Dirichlet), so the coarse free surface lies where the fine one does; a
// fluid cell whose parent is air is corrected by the smoother alone, which converges fast next to
// a Dirichlet boundary.  (Marking a coarse cell fluid when any child is fluid pushes the coarse
// surface outwards by up to a coarse cell per level and, measured on 256^3 dam-break systems,
// needs 2.5x the PCG iterations.)
struct k_mg_flag_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, MgLevelK L, MgLevelK C, MgFrame F) {
        const int cc = vbid * WG + threadIdx.x;
        if (cc >= C.ncell) return;
        int nin = 0, nfl = 0;
    #pragma unroll
        for (int a = 0; a < 8; ++a) {
            if (!mg_child_inside(L, C, cc, a)) continue;
            ++nin;
            const int ch = mg_child(L, C, F, cc, a);
            if (ch >= 0 && mg_fluid(L.info[ch])) ++nfl;
        }
        C.info[cc] = (nfl > 0 && 2 * nfl >= nin) ? kInfoFluid : 0;
    }
};
__global__ void k_mg_flag(MgLevelK L, MgLevelK C, MgFrame F) {
    k_mg_flag_f::run(blockIdx.x, gridDim.x, L, C, F);
}

// Stencil descriptor of a coarse level from its fluid flags.
struct k_mg_info_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, MgLevelK C, MgFrame F) {
        const int c = vbid * WG + threadIdx.x;
        if (c >= C.ncell) return;
        const uint16_t me = C.info[c];
        if (!mg_fluid(me)) return;
        int i, j, k;
        mg_coords(C, F, c, i, j, k);
        const bool wall = ((float)(j << C.scale_sh) + 0.5f <= F.wall_level + 1.f);
        int diag = 6;
        uint16_t mask = 0;
    #pragma unroll
        for (int q = 0; q < 6; ++q) {
            const int n = mg_nbr(C, F, c, q);
            if (n == -2) diag += wall ? -1 : 1;
            else if (n >= 0 && mg_fluid(C.info[n])) mask |= (uint16_t)(1u << q);
            else diag += 1;
        }
        C.info[c] = (uint16_t)(kInfoFluid | mask | (diag << 8));
    }
};
__global__ void k_mg_info(MgLevelK C, MgFrame F) {
    k_mg_info_f::run(blockIdx.x, gridDim.x, C, F);
}

// ---- V-cycle kernels of the levels below the block-aligned ones ----
// These levels hold few cells, so a launch costs the latency of its chain of dependent loads rather than
// bandwidth.  Each kernel issues every load whose address it knows at once (the done flag, the cell's
// descriptor and values, its neighbour-table row), then every load that needs a neighbour index, and
// tests the done flag only before its store: two round trips to memory where a load per step of the
// chain (flag, descriptor, table, neighbour) took four or five.  The neighbour loads are unconditional,
// from the cell itself where the descriptor has no fluid neighbour, so they sit in one basic block and
// issue together; a skipped term subtracts +0, which leaves the sum bit-identical.  mg_pin makes a value
// live before the done branch, so the compiler cannot sink the loads behind it.  When the solver is done
// the kernels store nothing.  The level's kind is a template argument, so each instance is straight-line
// code with all of its loads in one basic block.
__device__ __forceinline__ void mg_pin(float v) { asm volatile("" ::"v"(v)); }

// Kinds of level: sparse with edge 2 or more (a cell's children and parent lie in its block), sparse
// with edge 1 (cell index = block slot; the next level is dense), dense grid.
enum MgKind : int { kMgBlk = 0, kMgEdge1 = 1, kMgDense = 2 };
inline int mg_kind(const MgLevelK& L) { return L.dense ? kMgDense : (L.sh == 0 ? kMgEdge1 : kMgBlk); }

// z at cell c for the "first sweep from zero" case: omega * r / diag.
__device__ __forceinline__ float mg_z0(unsigned in, float r, float omega) {
    return mg_fluid(in) ? r * mg_wd(in, omega) : 0.f;
}

// The neighbour-table row of cell c (mg_nbr's table entries for its block; on a sparse level of edge 1
// the cell index is the block slot).  A dense level has no table.
template <int KIND>
__device__ __forceinline__ void mg_row(const MgLevelK& L, const MgFrame& F, int c, int t[6]) {
    const int sh = KIND == kMgEdge1 ? 0 : L.sh;
    const int s = c >> (3 * sh);
#pragma unroll
    for (int q = 0; q < 6; ++q) t[q] = KIND == kMgDense ? 0 : F.nbr[s * 6 + q];
}

// Makes the done flag, a descriptor and six indices complete at this point, so their loads are in
// flight together before it and nothing formed from them is reordered ahead of it.
// The values are read-write operands of the (empty) asm statement, so whatever is computed from them
// comes after it.
__device__ __forceinline__ void mg_pin7(int& dn, unsigned& a, int t[6]) {
    asm volatile("" : "+s"(dn), "+v"(a), "+v"(t[0]), "+v"(t[1]), "+v"(t[2]), "+v"(t[3]), "+v"(t[4]), "+v"(t[5]));
}
// The same for six loaded values (a neighbour round): issued together, complete here, so the arithmetic
// that consumes them is not interleaved with waits for individual loads.
template <class T> __device__ __forceinline__ void mg_pin6(T v[6]) {
    asm volatile("" : "+v"(v[0]), "+v"(v[1]), "+v"(v[2]), "+v"(v[3]), "+v"(v[4]), "+v"(v[5]));
}
template <class T> __device__ __forceinline__ void mg_pin1(T& v) { asm volatile("" : "+v"(v)); }

// mg_nbr for all six directions from the cell's table row t.
template <int KIND>
__device__ __forceinline__ void mg_nbrs(const MgLevelK& L, int c, const int t[6], int nb[6]) {
    if (KIND == kMgDense) {
        const int i = c % L.dx, j = (c / L.dx) % L.dy, k = c / (L.dx * L.dy);
        const int sy = L.dx, sz = L.dx * L.dy;
        nb[0] = i > 0 ? c - 1 : -2;
        nb[1] = i + 1 < L.dx ? c + 1 : -2;
        nb[2] = j > 0 ? c - sy : -2;
        nb[3] = j + 1 < L.dy ? c + sy : -2;
        nb[4] = k > 0 ? c - sz : -2;
        nb[5] = k + 1 < L.dz ? c + sz : -2;
        return;
    }
    if (KIND == kMgEdge1) {
#pragma unroll
        for (int q = 0; q < 6; ++q) nb[q] = t[q];
        return;
    }
    const int sh = L.sh, E = 1 << sh;
    const int l = c & ((1 << (3 * sh)) - 1);
    const int lc[3] = {l & (E - 1), (l >> sh) & (E - 1), l >> (2 * sh)};
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const int axis = q >> 1, dir = (q & 1) ? 1 : -1;
        const int stride = 1 << (axis * sh);
        const int nl = lc[axis] + dir;
        nb[q] = (nl >= 0 && nl < E) ? c + dir * stride : (t[q] < 0 ? t[q] : (t[q] << (3 * sh)) + l - dir * (E - 1) * stride);
    }
}

// The index to load for a neighbour: the neighbour when it exists, else the cell itself (always valid).
// The value is used only where the descriptor marks the neighbour fluid; choosing the index without the
// descriptor keeps these loads off the descriptor's round trip.
__device__ __forceinline__ int mg_ldx(int n, int c) { return n >= 0 ? n : c; }

// Two Jacobi sweeps from zero in one pass: zout = z1 + omega2 (r - A z1)/D with z1 = omega D^-1 r.
template <int KIND>
struct k_mg_smooth_first2_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, MgLevelK L, MgFrame F, float omega, float omega2, float* __restrict__ zout,
                                       const int* done) {
        int dn = *done;
        const int c = vbid * WG + threadIdx.x;
        if (c >= L.ncell) return;
        unsigned in = L.info[c];
        float rc = L.r[c];
        int t[6], nb[6];
        mg_row<KIND>(L, F, c, t);
        mg_pin7(dn, in, t);
        mg_pin1(rc);
        mg_nbrs<KIND>(L, c, t, nb);
        unsigned ni[6];
        float nr[6], zn[6];
    #pragma unroll
        for (int q = 0; q < 6; ++q) {
            const int n = mg_ldx(nb[q], c);
            ni[q] = L.info[n];
            nr[q] = L.r[n];
        }
        mg_pin6(ni);
        mg_pin6(nr);
    #pragma unroll
        for (int q = 0; q < 6; ++q) zn[q] = mg_z0(ni[q], nr[q], omega);
        const float d = mg_diag(in), wd = mg_wd(in, omega);
        const float z1 = rc * wd;
        float az = d * z1;
    #pragma unroll
        for (int q = 0; q < 6; ++q) az -= (in & (1u << q)) ? zn[q] : 0.f;
        const float out = mg_fluid(in) ? z1 + (rc - az) * (omega2 == omega ? wd : mg_wd(in, omega2)) : 0.f;
        mg_pin(out);
        if (dn) return;
        zout[c] = out;
    }
};
template <int KIND>
__global__ void k_mg_smooth_first2(MgLevelK L, MgFrame F, float omega, float omega2, float* __restrict__ zout,
                                   const int* done) {
    k_mg_smooth_first2_f<KIND>::run(blockIdx.x, gridDim.x, L, F, omega, omega2, zout, done);
}

// One Jacobi sweep: zout = z + omega (r - A z)/D.
template <int KIND>
struct k_mg_smooth_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, MgLevelK L, MgFrame F, float omega, const float* __restrict__ z, float* __restrict__ zout,
                                const int* done) {
        int dn = *done;
        const int c = vbid * WG + threadIdx.x;
        if (c >= L.ncell) return;
        unsigned in = L.info[c];
        float rc = L.r[c], zc = z[c];
        int t[6], nb[6];
        mg_row<KIND>(L, F, c, t);
        mg_pin7(dn, in, t);
        mg_pin1(rc);
        mg_pin1(zc);
        mg_nbrs<KIND>(L, c, t, nb);
        float zn[6];
    #pragma unroll
        for (int q = 0; q < 6; ++q) zn[q] = z[mg_ldx(nb[q], c)];
        mg_pin6(zn);
        float az = mg_diag(in) * zc;
    #pragma unroll
        for (int q = 0; q < 6; ++q) az -= (in & (1u << q)) ? zn[q] : 0.f;
        const float out = mg_fluid(in) ? zc + (rc - az) * mg_wd(in, omega) : 0.f;
        mg_pin(out);
        if (dn) return;
        zout[c] = out;
    }
};
template <int KIND>
__global__ void k_mg_smooth(MgLevelK L, MgFrame F, float omega, const float* __restrict__ z, float* __restrict__ zout,
                            const int* done) {
    k_mg_smooth_f<KIND>::run(blockIdx.x, gridDim.x, L, F, omega, z, zout, done);
}

// Restriction of the residual: rC[cc] = 0.5 * sum over fluid children of (r - A z).  With FIRST the
// fine z is the first sweep from zero (omega D^-1 r), computed on the fly and stored to zf for every
// fluid fine cell (a fluid cell under an air coarse cell too: the prolongation reads it).  One thread
// per (coarse cell, child) pair: the 8 lanes of a coarse cell each compute one child's residual and
// are summed with three butterfly shuffles, (c0+c1)+(c2+c3) ... in a fixed order, so the result is
// deterministic.  The child (mg_child's value) and its six neighbours come in one round of loads: on
// a sparse fine level of edge 2 or more the child is arithmetic and its block's table row is loaded;
// on one of edge 1 the child's slot and its neighbours' slots are read from the frame's slot table at
// their coordinates, not through the child's table row, which would add a round trip.
template <int KIND, bool FIRST>
struct k_mg_restrict_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, MgLevelK L, MgLevelK C, MgFrame F, float omega, float* __restrict__ zf,
                                  const int* done) {
        int dn = *done;
        const int tid = vbid * WG + threadIdx.x;
        const int cc = tid >> 3, a = tid & 7;
        const bool valid = cc < C.ncell;
        const int ccv = valid ? cc : 0;
        const unsigned cin = C.info[ccv];
        int ch, t[6], nb[6];
        unsigned in0;
        float rch, zch0;
        if (KIND == kMgBlk) {                                // mg_child's in-block case
            const int sh = L.sh, csh = sh - 1, Ec = 1 << csh;
            const int s = ccv >> (3 * csh);
            const int l = ccv & ((1 << (3 * csh)) - 1);
            const int li = l & (Ec - 1), lj = (l >> csh) & (Ec - 1), lk = l >> (2 * csh);
            ch = (s << (3 * sh)) + ((((2 * lk + (a >> 2)) << sh) + (2 * lj + ((a >> 1) & 1))) << sh) + (2 * li + (a & 1));
            mg_row<KIND>(L, F, ch, t);
            in0 = L.info[ch];
            rch = L.r[ch];
            zch0 = FIRST ? 0.f : zf[ch];
            mg_pin7(dn, in0, t);
            mg_nbrs<KIND>(L, ch, t, nb);
        } else {
            const int I = ccv % C.dx, J = (ccv / C.dx) % C.dy, K = ccv / (C.dx * C.dy);
            const int i = 2 * I + (a & 1), j = 2 * J + ((a >> 1) & 1), k = 2 * K + (a >> 2);
            const bool inside = i < L.dx && j < L.dy && k < L.dz;
            if (KIND == kMgDense) {
                ch = inside ? (k * L.dy + j) * L.dx + i : -1;
                mg_nbrs<KIND>(L, inside ? ch : 0, t, nb);
            } else {
                // slot-table entries at clamped coordinates, all loaded unconditionally, then masked
                auto slot_raw = [&](int x, int y, int z) -> int {
                    x = x < 0 ? 0 : (x >= L.dx ? L.dx - 1 : x);
                    y = y < 0 ? 0 : (y >= L.dy ? L.dy - 1 : y);
                    z = z < 0 ? 0 : (z >= L.dz ? L.dz - 1 : z);
                    return F.slot[(z * L.dy + y) * L.dx + x];
                };
                auto in_frame = [&](int x, int y, int z) {
                    return x >= 0 && y >= 0 && z >= 0 && x < L.dx && y < L.dy && z < L.dz;
                };
                unsigned cs = (unsigned)slot_raw(i, j, k);
                t[0] = slot_raw(i - 1, j, k);
                t[1] = slot_raw(i + 1, j, k);
                t[2] = slot_raw(i, j - 1, k);
                t[3] = slot_raw(i, j + 1, k);
                t[4] = slot_raw(i, j, k - 1);
                t[5] = slot_raw(i, j, k + 1);
                mg_pin7(dn, cs, t);
                nb[0] = in_frame(i - 1, j, k) ? t[0] : -2;        // -2 outside the frame, else slot or -1
                nb[1] = in_frame(i + 1, j, k) ? t[1] : -2;
                nb[2] = in_frame(i, j - 1, k) ? t[2] : -2;
                nb[3] = in_frame(i, j + 1, k) ? t[3] : -2;
                nb[4] = in_frame(i, j, k - 1) ? t[4] : -2;
                nb[5] = in_frame(i, j, k + 1) ? t[5] : -2;
                ch = (inside && (int)cs >= 0) ? (int)cs : -1;
            }
            const int cv = ch >= 0 ? ch : 0;
            in0 = L.info[cv];
            rch = L.r[cv];
            zch0 = FIRST ? 0.f : zf[cv];
        }
        const int chv = ch >= 0 ? ch : 0;
        const unsigned in = ch >= 0 ? in0 : 0u;
        float zch, zn[6];
        if (FIRST) {
            unsigned ni[6];
            float nr[6];
    #pragma unroll
            for (int q = 0; q < 6; ++q) {
                const int n = mg_ldx(nb[q], chv);
                ni[q] = L.info[n];
                nr[q] = L.r[n];
            }
            mg_pin6(ni);
            mg_pin6(nr);
            zch = rch * mg_wd(in, omega);
    #pragma unroll
            for (int q = 0; q < 6; ++q) zn[q] = mg_z0(ni[q], nr[q], omega);
        } else {
            zch = zch0;
    #pragma unroll
            for (int q = 0; q < 6; ++q) zn[q] = zf[mg_ldx(nb[q], chv)];
            mg_pin6(zn);
        }
        mg_pin1(rch);
        float az = mg_diag(in) * zch;
    #pragma unroll
        for (int q = 0; q < 6; ++q) az -= (in & (1u << q)) ? zn[q] : 0.f;
        const bool cfl = valid && mg_fluid(cin);
        const bool fine = valid && mg_fluid(in);
        float res = (cfl && fine) ? rch - az : 0.f;
        res += __shfl_xor(res, 1, 64);
        res += __shfl_xor(res, 2, 64);
        res += __shfl_xor(res, 4, 64);
        mg_pin(res);
        if (dn) return;
        if (FIRST && fine) zf[ch] = zch;
        if (valid && a == 0) C.r[cc] = cfl ? 0.5f * res : 0.f;
    }
};
template <int KIND, bool FIRST>
__global__ void k_mg_restrict(MgLevelK L, MgLevelK C, MgFrame F, float omega, float* __restrict__ zf,
                              const int* done) {
    k_mg_restrict_f<KIND, FIRST>::run(blockIdx.x, gridDim.x, L, C, F, omega, zf, done);
}

// Prolongation (piecewise constant) fused with one Jacobi sweep:
// zp = z + zC[parent]; zout = zp + omega (r - A zp)/D.  On a sparse level of edge 2 or more a cell's
// parent is arithmetic in its index; otherwise (edge 1, dense) the parents of the cell and of its
// neighbours follow from the cell's coordinates, so no parent needs its neighbour's index first.
template <int KIND>
struct k_mg_prolong_smooth_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, MgLevelK L, MgLevelK C, MgFrame F, float omega, const float* __restrict__ z,
                                        const float* __restrict__ zc, float* __restrict__ zout, const int* done) {
        int dn = *done;
        const int c = vbid * WG + threadIdx.x;
        if (c >= L.ncell) return;
        constexpr bool inblk = KIND == kMgBlk;
        auto parent_ijk = [&](int x, int y, int w) { return ((w >> 1) * C.dy + (y >> 1)) * C.dx + (x >> 1); };
        auto parent_blk = [&](int x) {                       // mg_parent's in-block case
            const int sh = L.sh, csh = sh - 1, E = 1 << sh;
            const int s = x >> (3 * sh);
            const int l = x & ((1 << (3 * sh)) - 1);
            const int li = (l & (E - 1)) >> 1, lj = ((l >> sh) & (E - 1)) >> 1, lk = (l >> (2 * sh)) >> 1;
            return (s << (3 * csh)) + (((lk << csh) + lj) << csh) + li;
        };
        unsigned in = L.info[c];
        float rc = L.r[c], zcc = z[c];
        int i = 0, j = 0, k = 0;
        float zpc = 0.f;
        if (KIND == kMgEdge1) {
            const int tc = F.coord[c];
            i = tc % L.dx; j = (tc / L.dx) % L.dy; k = tc / (L.dx * L.dy);
        } else if (KIND == kMgDense) {
            i = c % L.dx; j = (c / L.dx) % L.dy; k = c / (L.dx * L.dy);
            zpc = zc[parent_ijk(i, j, k)];
        } else {
            zpc = zc[parent_blk(c)];
        }
        int t[6], nb[6];
        mg_row<KIND>(L, F, c, t);
        mg_pin7(dn, in, t);
        mg_pin1(rc);
        mg_pin1(zcc);
        if (KIND == kMgEdge1) zpc = zc[parent_ijk(i, j, k)];
        mg_nbrs<KIND>(L, c, t, nb);
        float za[6], zb[6], zn[6];
    #pragma unroll
        for (int q = 0; q < 6; ++q) {
            const int n = mg_ldx(nb[q], c);
            int pn;
            if (inblk) {
                pn = parent_blk(n);
            } else {
                const int dir = (q & 1) ? 1 : -1;
                const bool use = nb[q] >= 0;                 // else the cell itself, like n
                pn = parent_ijk(i + (use && q < 2 ? dir : 0), j + (use && (q >> 1) == 1 ? dir : 0),
                                k + (use && q >= 4 ? dir : 0));
            }
            za[q] = z[n];
            zb[q] = zc[pn];
        }
        mg_pin6(za);
        mg_pin6(zb);
        mg_pin1(zpc);
    #pragma unroll
        for (int q = 0; q < 6; ++q) zn[q] = za[q] + zb[q];
        const float zp = zcc + zpc;
        float az = mg_diag(in) * zp;
    #pragma unroll
        for (int q = 0; q < 6; ++q) az -= (in & (1u << q)) ? zn[q] : 0.f;
        const float out = mg_fluid(in) ? zp + (rc - az) * mg_wd(in, omega) : 0.f;
        mg_pin(out);
        if (dn) return;
        zout[c] = out;
    }
};
template <int KIND>
__global__ void k_mg_prolong_smooth(MgLevelK L, MgLevelK C, MgFrame F, float omega, const float* __restrict__ z,
                                    const float* __restrict__ zc, float* __restrict__ zout, const int* done) {
    k_mg_prolong_smooth_f<KIND>::run(blockIdx.x, gridDim.x, L, C, F, omega, z, zc, zout, done);
}

// ---- bottom of the hierarchy: the last levels, at most kMgBottomCells cells, in one workgroup ----
// The V-cycle on these levels (down: two Jacobi sweeps from zero and the restriction; the
// coarsest level: `coarse_sweeps` Jacobi sweeps from zero; up: prolongation and the post-sweeps)
// runs in one 1024-thread workgroup with every level's r, z and descriptors in LDS, replacing four
// launches per level.  Each level's arithmetic is that of k_mg_smooth_first2 / k_mg_restrict /
// k_mg_prolong_smooth / k_mg_smooth, in the same order, and like those kernels the first sweep from
// zero and the prolonged values are formed on the fly, so a level costs two barrier-separated passes
// down and one (nu = 1) or two up.  The coarsest level gets `coarse_sweeps` Jacobi sweeps from zero.
constexpr int kMgBottomLevels = 6;                     // levels handled at most
constexpr int kMgBottomCells = 5120;                   // cells of those levels together (LDS: 14 B each)
constexpr int kMgBottomThreads = 1024;

// A sparse level (edge 2^sh per block) enters as its dense view: the block frame at that level's
// resolution, dims (fx, fy, fz) << sh, whose cell (i, j, k) is stored at slot*E^3 + local for an
// active block and is air (not fluid) in an inactive one.
struct MgBottom {
    int nlev;                                          // levels handled; the last is the coarsest
    int dx[kMgBottomLevels], dy[kMgBottomLevels], dz[kMgBottomLevels];
    int off[kMgBottomLevels];                          // first cell of each level in the LDS arrays
    int sh[kMgBottomLevels];                           // sparse level: log2 of its block edge; dense: -1
    const uint16_t* info[kMgBottomLevels];
    const int* slot;                                   // frame slot table (sparse levels)
    int fx, fy;                                        // frame dims
    const float* r0;                                   // right-hand side of the first level
    float* z0;                                         // its correction (output)
};

// The CG's x += alpha p over the solver's blocks (block list blk, 2^sh3 cells each), applied by the
// extra workgroups of k_mg_bottom; x == nullptr: none.  A workgroup takes kXUpdPasses consecutive
// groups of blocks, one block per 2^sh3 threads (whole waves, so a block's slot is a uniform scalar
// load), and issues all of its x and p loads at once.
struct XTerm {
    float* x;
    const float* p;                                    // this iteration's direction
    const float* alpha;                                // this iteration's alpha (solver state)
    const int* blk;
    int sh3;                                           // log2 of the cells per block (>= 6)
    int nblk;                                          // listed blocks
};
constexpr int kXUpdPasses = 16;

// Storage index of dense-view cell c of bottom level l (-1: inactive block).
__device__ __forceinline__ int bottom_index(const MgBottom& bt, int l, int c) {
    const int sh = bt.sh[l];
    if (sh < 0) return c;
    const int dx = bt.dx[l], dy = bt.dy[l];
    const int i = c % dx, j = (c / dx) % dy, k = c / (dx * dy);
    const int s = bt.slot[((k >> sh) * bt.fy + (j >> sh)) * bt.fx + (i >> sh)];
    if (s < 0) return -1;
    const int m = (1 << sh) - 1;
    return (s << (3 * sh)) + ((((k & m) << sh) + (j & m)) << sh) + (i & m);
}

// diag * z - sum of the fluid neighbours' z on a dense level held in LDS (q order -x,+x,-y,+y,-z,+z).
// The six neighbour offsets are read unconditionally, from the cell itself where the descriptor has no
// fluid neighbour, so the seven reads issue together; a skipped term subtracts +0 (bit-identical).
__device__ __forceinline__ void bottom_nbrs(unsigned in, int c, int sy, int sz, int nb[6]) {
    nb[0] = (in & 1u) ? c - 1 : c;
    nb[1] = (in & 2u) ? c + 1 : c;
    nb[2] = (in & 4u) ? c - sy : c;
    nb[3] = (in & 8u) ? c + sy : c;
    nb[4] = (in & 16u) ? c - sz : c;
    nb[5] = (in & 32u) ? c + sz : c;
}
__device__ __forceinline__ float bottom_az(unsigned in, const float* z, int c, int sy, int sz) {
    int nb[6];
    bottom_nbrs(in, c, sy, sz, nb);
    float v[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) v[q] = z[nb[q]];
    float az = mg_diag(in) * z[c];
#pragma unroll
    for (int q = 0; q < 6; ++q) az -= (in & (1u << q)) ? v[q] : 0.f;
    return az;
}

// One Jacobi sweep at cell c of a bottom level held in LDS: zout = z + omega (r - A z)/D (k_mg_smooth).
__device__ __forceinline__ float bottom_sweep(unsigned in, float r, const float* z, int c, int sy, int sz, float omega) {
    return mg_fluid(in) ? z[c] + (r - bottom_az(in, z, c, sy, sz)) * mg_wd(in, omega) : 0.f;
}

// Workgroup 0 runs the bottom V-cycle; workgroups 1.. (when xt.x is set) apply the CG's x += alpha p
// meanwhile (see XTerm).  omega weighs the coarsest grid's sweeps, wa and wb the first and second
// sweep of each smoothing pair on the levels above it.
struct k_mg_bottom_f {
    static constexpr int kLB = kMgBottomThreads;
    static constexpr int kWPE = 1;
    // bt by reference: its per-level arrays are indexed by a run-time level, which reads them where they
    // lie (the kernel-argument segment) instead of copying the struct to private memory
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, const MgBottom& bt, float omega,
                                               float wa, float wb, int nu, int coarse_sweeps, const int* done, XTerm xt) {
        if (*done) return;
        __shared__ float sr[kMgBottomCells];
        __shared__ float sz[kMgBottomCells];
        __shared__ float st[kMgBottomCells];
        __shared__ uint16_t si[kMgBottomCells];
        const int tid = threadIdx.x;
        constexpr int T = kMgBottomThreads;
        if (vbid > 0) {
            const float a = *xt.alpha;
            const int bpp = T >> xt.sh3;                   // blocks per pass
            const int l = tid & ((1 << xt.sh3) - 1);
            const int b0 = (vbid - 1) * bpp * kXUpdPasses + (tid >> xt.sh3);
            int c[kXUpdPasses];
            float xv[kXUpdPasses], pv[kXUpdPasses];
    #pragma unroll
            for (int u = 0; u < kXUpdPasses; ++u) {
                const int b = __builtin_amdgcn_readfirstlane(b0 + u * bpp);     // uniform: whole waves per block
                c[u] = b < xt.nblk ? (xt.blk[b] << xt.sh3) | l : -1;
            }
    #pragma unroll
            for (int u = 0; u < kXUpdPasses; ++u) {
                const int cs = c[u] >= 0 ? c[u] : 0;
                xv[u] = xt.x[cs];
                pv[u] = xt.p[cs];
            }
    #pragma unroll
            for (int u = 0; u < kXUpdPasses; ++u)
                if (c[u] >= 0) xt.x[c[u]] = xv[u] + (a * pv[u]);
            return;
        }
        constexpr int kPer = (kMgBottomCells + T - 1) / T;      // cells per thread of any one level, at most
        const int L = bt.nlev;
        // load: a level has at most kPer cells per thread; their indices, then all of their descriptors
        // and right-hand sides, are loaded in unrolled rounds (safe index 0 where there is no cell), so a
        // level costs one round trip to memory, two for a sparse level's slot lookup, not one per cell
        for (int l = 0; l < L; ++l) {
            const int n = bt.dx[l] * bt.dy[l] * bt.dz[l];
            int g[kPer];
            uint16_t iv[kPer];
            float rv[kPer];
            const int sh = bt.sh[l];
            if (sh < 0) {
    #pragma unroll
                for (int k = 0; k < kPer; ++k) g[k] = tid + k * T < n ? tid + k * T : -1;
            } else {                                         // bottom_index with the slot loads batched
                const int dx = bt.dx[l], dy = bt.dy[l], m = (1 << sh) - 1;
                int raw[kPer];
    #pragma unroll
                for (int k = 0; k < kPer; ++k) {
                    const int c = tid + k * T < n ? tid + k * T : 0;
                    const int i = c % dx, j = (c / dx) % dy, kk = c / (dx * dy);
                    raw[k] = bt.slot[((kk >> sh) * bt.fy + (j >> sh)) * bt.fx + (i >> sh)];
                }
    #pragma unroll
                for (int k = 0; k < kPer; ++k) {
                    const int c = tid + k * T;
                    const int i = c % dx, j = (c / dx) % dy, kk = c / (dx * dy);
                    g[k] = (c < n && raw[k] >= 0) ? (raw[k] << (3 * sh)) + ((((kk & m) << sh) + (j & m)) << sh) + (i & m) : -1;
                }
            }
    #pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int gs = g[k] >= 0 ? g[k] : 0;
                iv[k] = bt.info[l][gs];
                rv[k] = l == 0 ? bt.r0[gs] : 0.f;
            }
    #pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int c = tid + k * T;
                if (c < n) {
                    const uint16_t in = g[k] >= 0 ? iv[k] : (uint16_t)0;
                    si[bt.off[l] + c] = in;
                    if (l == 0) sr[c] = mg_fluid(in) ? rv[k] : 0.f;
                }
            }
        }
        __syncthreads();
        // down: the pre-smoothed z of level l goes to sz + off[l]
        for (int l = 0; l + 1 < L; ++l) {
            const int o = bt.off[l], dx = bt.dx[l], dy = bt.dy[l], n = dx * dy * bt.dz[l], sy = dx, szz = dx * dy;
            const float* r = sr + o;
            float* z = sz + o;
            const uint16_t* in = si + o;
            // nu = 1: z = omega r / D; nu = 2: a second sweep from that z, whose neighbour values are formed
            // on the fly (k_mg_smooth_first2)
            for (int c = tid; c < n; c += T) {
                const unsigned ic = in[c];
                const float wd = mg_wd(ic, wa);
                const float z1 = r[c] * wd;
                float out = z1;
                if (nu == 2) {
                    int nb[6];
                    bottom_nbrs(ic, c, sy, szz, nb);
                    float v[6];
    #pragma unroll
                    for (int q = 0; q < 6; ++q) v[q] = r[nb[q]] * mg_wd(in[nb[q]], wa);
                    float az = mg_diag(ic) * z1;
    #pragma unroll
                    for (int q = 0; q < 6; ++q) az -= (ic & (1u << q)) ? v[q] : 0.f;
                    out = z1 + (r[c] - az) * (wb == wa ? wd : mg_wd(ic, wb));
                }
                z[c] = mg_fluid(ic) ? out : 0.f;
            }
            __syncthreads();
            // restriction: rC = 0.5 * sum of the fluid children's residuals r - A z, children summed as
            // ((c0+c1)+(c2+c3))+((c4+c5)+(c6+c7)) like k_mg_restrict's butterflies
            const int oc = bt.off[l + 1], cdx = bt.dx[l + 1], cdy = bt.dy[l + 1], nc = cdx * cdy * bt.dz[l + 1];
            for (int cc = tid; cc < nc; cc += T) {
                float acc = 0.f;
                if (mg_fluid(si[oc + cc])) {
                    const int I = cc % cdx, J = (cc / cdx) % cdy, K = cc / (cdx * cdy);
                    float ch[8];
    #pragma unroll
                    for (int a = 0; a < 8; ++a) {
                        const int i = 2 * I + (a & 1), j = 2 * J + ((a >> 1) & 1), k = 2 * K + (a >> 2);
                        const bool inside = i < dx && j < dy && k < bt.dz[l];
                        const int c = inside ? (k * dy + j) * dx + i : 0;
                        const unsigned ic = inside ? in[c] : 0u;
                        const float res = r[c] - bottom_az(ic, z, c, sy, szz);
                        ch[a] = mg_fluid(ic) ? res : 0.f;
                    }
                    acc = 0.5f * (((ch[0] + ch[1]) + (ch[2] + ch[3])) + ((ch[4] + ch[5]) + (ch[6] + ch[7])));
                }
                sr[oc + cc] = acc;
            }
            __syncthreads();
        }
        // coarsest level: `coarse_sweeps` Jacobi sweeps from zero (the first is omega r / D), alternating
        // between sz and st; zfin = where the last one is
        float* zfin;
        {
            const int l = L - 1;
            const int o = bt.off[l], dx = bt.dx[l], dy = bt.dy[l], n = dx * dy * bt.dz[l], sy = dx, szz = dx * dy;
            const float* r = sr + o;
            const uint16_t* in = si + o;
            float* cur = sz + o;
            float* nxt = st + o;
            for (int c = tid; c < n; c += T) cur[c] = (coarse_sweeps > 0 && mg_fluid(in[c])) ? r[c] * mg_wd(in[c], omega) : 0.f;
            if (n <= 64) {
                // a coarsest grid of at most 64 cells (one per lane of wave 0, which also wrote its first
                // sweep above) is swept by that wave alone: LDS reads and writes of one wave complete in
                // order, so its sweeps need no workgroup barrier; the other waves wait at the barrier below
                if (tid < 64) {
                    for (int s = 1; s < coarse_sweeps; ++s) {
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                        if (tid < n) nxt[tid] = bottom_sweep(in[tid], r[tid], cur, tid, sy, szz, omega);
                        float* tmp = cur; cur = nxt; nxt = tmp;
                    }
                }
            } else {
                for (int s = 1; s < coarse_sweeps; ++s) {
                    __syncthreads();
                    for (int c = tid; c < n; c += T) nxt[c] = bottom_sweep(in[c], r[c], cur, c, sy, szz, omega);
                    float* tmp = cur; cur = nxt; nxt = tmp;
                }
            }
            zfin = ((coarse_sweeps < 1 ? 1 : coarse_sweeps) & 1) ? sz + o : st + o;
            __syncthreads();
        }
        // up: level l's post-smoothed z ends in st (nu = 1) or sz (nu = 2) + off[l]
        for (int l = L - 2; l >= 0; --l) {
            const int o = bt.off[l], dx = bt.dx[l], dy = bt.dy[l], n = dx * dy * bt.dz[l], sy = dx, szz = dx * dy;
            const int cdx = bt.dx[l + 1], cdy = bt.dy[l + 1];
            const float* r = sr + o;
            float* z = sz + o;
            float* zt = st + o;
            const uint16_t* in = si + o;
            const float* zc = zfin;
            // prolongation fused with the first post-sweep: zp = z + zC[parent] on the cell and its fluid
            // neighbours, formed on the fly (k_mg_prolong_smooth), into zt
            for (int c = tid; c < n; c += T) {
                const unsigned ic = in[c];
                const int i = c % dx, j = (c / dx) % dy, k = c / (dx * dy);
                auto par = [&](int x, int y, int w) { return ((w >> 1) * cdy + (y >> 1)) * cdx + (x >> 1); };
                int nb[6], pn[6];
                bottom_nbrs(ic, c, sy, szz, nb);
                pn[0] = par(i - ((ic & 1u) ? 1 : 0), j, k);
                pn[1] = par(i + ((ic & 2u) ? 1 : 0), j, k);
                pn[2] = par(i, j - ((ic & 4u) ? 1 : 0), k);
                pn[3] = par(i, j + ((ic & 8u) ? 1 : 0), k);
                pn[4] = par(i, j, k - ((ic & 16u) ? 1 : 0));
                pn[5] = par(i, j, k + ((ic & 32u) ? 1 : 0));
                float v[6];
    #pragma unroll
                for (int q = 0; q < 6; ++q) v[q] = z[nb[q]] + zc[pn[q]];
                const float zpc = z[c] + zc[par(i, j, k)];
                float az = mg_diag(ic) * zpc;
    #pragma unroll
                for (int q = 0; q < 6; ++q) az -= (ic & (1u << q)) ? v[q] : 0.f;
                zt[c] = mg_fluid(ic) ? zpc + (r[c] - az) * mg_wd(ic, wa) : 0.f;
            }
            __syncthreads();
            zfin = zt;
            if (nu == 2) {                                   // second post-sweep (k_mg_smooth), into z
                for (int c = tid; c < n; c += T) z[c] = bottom_sweep(in[c], r[c], zt, c, sy, szz, wb);
                __syncthreads();
                zfin = z;
            }
        }
        {
            const int n = bt.dx[0] * bt.dy[0] * bt.dz[0];
    #pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int c = tid + k * T;
                const int g = c < n ? bottom_index(bt, 0, c) : -1;
                if (g >= 0) bt.z0[g] = zfin[c];
            }
        }
    }
};
__global__ void __launch_bounds__(kMgBottomThreads) k_mg_bottom(MgBottom bt, float omega, float wa, float wb, int nu, int coarse_sweeps,
                                                                const int* done, XTerm xt) {
    k_mg_bottom_f::run(blockIdx.x, gridDim.x, bt, omega, wa, wb, nu, coarse_sweeps, done, xt);
}

}  // namespace vare
