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
__global__ void k_mg_flag(MgLevelK L, MgLevelK C, MgFrame F) {
    const int cc = blockIdx.x * WG + threadIdx.x;
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

// Stencil descriptor of a coarse level from its fluid flags.
__global__ void k_mg_info(MgLevelK C, MgFrame F) {
    const int c = blockIdx.x * WG + threadIdx.x;
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

// ---- V-cycle kernels; each returns at once when the solver is done ----
// z at cell c for the "first sweep from zero" case: omega * r / diag.
__device__ __forceinline__ float mg_z0(const MgLevelK& L, int c, float omega) {
    const uint16_t in = L.info[c];
    return mg_fluid(in) ? L.r[c] * mg_wd(in, omega) : 0.f;
}

// Two Jacobi sweeps from zero in one pass: zout = z1 + omega (r - A z1)/D with z1 = omega D^-1 r.
__global__ void k_mg_smooth_first2(MgLevelK L, MgFrame F, float omega, float* __restrict__ zout, const int* done) {
    if (*done) return;
    const int c = blockIdx.x * WG + threadIdx.x;
    if (c >= L.ncell) return;
    const uint16_t in = L.info[c];
    float out = 0.f;
    if (mg_fluid(in)) {
        const float d = mg_diag(in), wd = mg_wd(in, omega);
        const float z1 = L.r[c] * wd;
        float az = d * z1;
#pragma unroll
        for (int q = 0; q < 6; ++q)
            if (in & (1u << q)) az -= mg_z0(L, mg_nbr(L, F, c, q), omega);
        out = z1 + (L.r[c] - az) * wd;
    }
    zout[c] = out;
}

// One Jacobi sweep: zout = z + omega (r - A z)/D.
__global__ void k_mg_smooth(MgLevelK L, MgFrame F, float omega, const float* __restrict__ z, float* __restrict__ zout,
                            const int* done) {
    if (*done) return;
    const int c = blockIdx.x * WG + threadIdx.x;
    if (c >= L.ncell) return;
    const uint16_t in = L.info[c];
    float out = 0.f;
    if (mg_fluid(in)) {
        const float d = mg_diag(in);
        const float zc = z[c];
        float az = d * zc;
#pragma unroll
        for (int q = 0; q < 6; ++q)
            if (in & (1u << q)) az -= z[mg_nbr(L, F, c, q)];
        out = zc + (L.r[c] - az) * mg_wd(in, omega);
    }
    zout[c] = out;
}

// Restriction of the residual: rC[cc] = 0.5 * sum over fluid children of (r - A z).  With
// first != 0 the fine z is the first sweep from zero (omega D^-1 r), computed on the fly and
// stored to zf.  One thread per (coarse cell, child) pair: the 8 lanes of a coarse cell each
// compute one child's residual and are summed with three butterfly shuffles, (c0+c1)+(c2+c3) ...
// in a fixed order, so the result is deterministic.
__global__ void k_mg_restrict(MgLevelK L, MgLevelK C, MgFrame F, float omega, int first, float* __restrict__ zf,
                              const int* done) {
    if (*done) return;
    const int t = blockIdx.x * WG + threadIdx.x;
    const int cc = t >> 3, a = t & 7;
    const bool valid = cc < C.ncell;
    float res = 0.f;
    const bool cfl = valid && mg_fluid(C.info[cc]);
    if (cfl) {
        const int ch = mg_child(L, C, F, cc, a);
        const uint16_t in = ch >= 0 ? L.info[ch] : (uint16_t)0;
        if (mg_fluid(in)) {
            const float d = mg_diag(in);
            const float zc = first ? L.r[ch] * mg_wd(in, omega) : zf[ch];
            if (first) zf[ch] = zc;
            float az = d * zc;
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                if (!(in & (1u << q))) continue;
                const int n = mg_nbr(L, F, ch, q);
                az -= first ? mg_z0(L, n, omega) : zf[n];
            }
            res = L.r[ch] - az;
        }
    }
    res += __shfl_xor(res, 1, 64);
    res += __shfl_xor(res, 2, 64);
    res += __shfl_xor(res, 4, 64);
    if (valid && a == 0) C.r[cc] = cfl ? 0.5f * res : 0.f;
}

// Prolongation (piecewise constant) fused with one Jacobi sweep:
// zp = z + zC[parent]; zout = zp + omega (r - A zp)/D.
__global__ void k_mg_prolong_smooth(MgLevelK L, MgLevelK C, MgFrame F, float omega, const float* __restrict__ z,
                                    const float* __restrict__ zc, float* __restrict__ zout, const int* done) {
    if (*done) return;
    const int c = blockIdx.x * WG + threadIdx.x;
    if (c >= L.ncell) return;
    const uint16_t in = L.info[c];
    float out = 0.f;
    if (mg_fluid(in)) {
        const float d = mg_diag(in);
        const float zp = z[c] + zc[mg_parent(L, C, F, c)];
        float az = d * zp;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            if (!(in & (1u << q))) continue;
            const int n = mg_nbr(L, F, c, q);
            az -= z[n] + zc[mg_parent(L, C, F, n)];
        }
        out = zp + (L.r[c] - az) * mg_wd(in, omega);
    }
    zout[c] = out;
}

// ---- bottom of the hierarchy: the last levels, at most kMgBottomCells cells, in one workgroup ----
// The V-cycle on these levels (down: two Jacobi sweeps from zero and the restriction; the
// coarsest level: `coarse_sweeps` Jacobi sweeps from zero; up: prolongation and the post-sweeps)
// runs in one 1024-thread workgroup with every level's r, z and descriptors in LDS, replacing four
// launches per level.  Each level's arithmetic is that of k_mg_smooth_first2 / k_mg_restrict /
// k_mg_prolong_smooth / k_mg_smooth, in the same order; the coarsest level gets `coarse_sweeps`
// Jacobi sweeps from zero.
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
__device__ __forceinline__ float bottom_az(unsigned in, const float* z, int c, int sy, int sz) {
    float az = mg_diag(in) * z[c];
    if (in & 1u) az -= z[c - 1];
    if (in & 2u) az -= z[c + 1];
    if (in & 4u) az -= z[c - sy];
    if (in & 8u) az -= z[c + sy];
    if (in & 16u) az -= z[c - sz];
    if (in & 32u) az -= z[c + sz];
    return az;
}

__global__ void __launch_bounds__(kMgBottomThreads) k_mg_bottom(MgBottom bt, float omega, int nu, int coarse_sweeps,
                                                                const int* done) {
    if (*done) return;
    __shared__ float sr[kMgBottomCells];
    __shared__ float sz[kMgBottomCells];
    __shared__ float st[kMgBottomCells];
    __shared__ uint16_t si[kMgBottomCells];
    const int tid = threadIdx.x;
    constexpr int T = kMgBottomThreads;
    const int L = bt.nlev;
    for (int l = 0; l < L; ++l) {
        const int n = bt.dx[l] * bt.dy[l] * bt.dz[l];
        for (int c = tid; c < n; c += T) {
            const int g = bottom_index(bt, l, c);
            const uint16_t in = g >= 0 ? bt.info[l][g] : (uint16_t)0;
            si[bt.off[l] + c] = in;
            if (l == 0) sr[c] = mg_fluid(in) ? bt.r0[g] : 0.f;
        }
    }
    __syncthreads();
    // down
    for (int l = 0; l + 1 < L; ++l) {
        const int o = bt.off[l], dx = bt.dx[l], dy = bt.dy[l], n = dx * dy * bt.dz[l], sy = dx, szz = dx * dy;
        float* r = sr + o;
        float* z = sz + o;
        float* zt = st + o;
        const uint16_t* in = si + o;
        // first sweep from zero: omega r / D (k_mg_smooth_first2's z1, or the whole pre-smoothing for nu = 1)
        for (int c = tid; c < n; c += T) (nu == 1 ? z : zt)[c] = mg_fluid(in[c]) ? r[c] * mg_wd(in[c], omega) : 0.f;
        __syncthreads();
        if (nu == 2) {
            for (int c = tid; c < n; c += T) {
                float out = 0.f;
                if (mg_fluid(in[c])) out = zt[c] + (r[c] - bottom_az(in[c], zt, c, sy, szz)) * mg_wd(in[c], omega);
                z[c] = out;
            }
            __syncthreads();
        }
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
                    ch[a] = 0.f;
                    if (i < dx && j < dy && k < bt.dz[l]) {
                        const int c = (k * dy + j) * dx + i;
                        if (mg_fluid(in[c])) ch[a] = r[c] - bottom_az(in[c], z, c, sy, szz);
                    }
                }
                acc = 0.5f * (((ch[0] + ch[1]) + (ch[2] + ch[3])) + ((ch[4] + ch[5]) + (ch[6] + ch[7])));
            }
            sr[oc + cc] = acc;
        }
        __syncthreads();
    }
    // coarsest level: Jacobi sweeps from zero
    {
        const int l = L - 1;
        const int o = bt.off[l], dx = bt.dx[l], dy = bt.dy[l], n = dx * dy * bt.dz[l], sy = dx, szz = dx * dy;
        float* cur = sz + o;
        float* nxt = st + o;
        for (int c = tid; c < n; c += T) cur[c] = 0.f;
        __syncthreads();
        for (int s = 0; s < coarse_sweeps; ++s) {
            for (int c = tid; c < n; c += T) {
                const unsigned in = si[o + c];
                float out = 0.f;
                if (mg_fluid(in)) out = cur[c] + (sr[o + c] - bottom_az(in, cur, c, sy, szz)) * mg_wd(in, omega);
                nxt[c] = out;
            }
            __syncthreads();
            float* t = cur; cur = nxt; nxt = t;
        }
        if (cur != sz + o) {
            for (int c = tid; c < n; c += T) sz[o + c] = cur[c];
            __syncthreads();
        }
    }
    // up
    for (int l = L - 2; l >= 0; --l) {
        const int o = bt.off[l], dx = bt.dx[l], dy = bt.dy[l], n = dx * dy * bt.dz[l], sy = dx, szz = dx * dy;
        const int oc = bt.off[l + 1], cdx = bt.dx[l + 1], cdy = bt.dy[l + 1];
        float* r = sr + o;
        float* z = sz + o;
        float* zt = st + o;
        const uint16_t* in = si + o;
        // zp = z + zC[parent] on every cell (k_mg_prolong_smooth adds the same two terms per cell)
        for (int c = tid; c < n; c += T) {
            const int i = c % dx, j = (c / dx) % dy, k = c / (dx * dy);
            zt[c] = z[c] + sz[oc + ((k >> 1) * cdy + (j >> 1)) * cdx + (i >> 1)];
        }
        __syncthreads();
        // one sweep from zp; nu = 2 adds a second (k_mg_smooth)
        for (int c = tid; c < n; c += T) {
            float out = 0.f;
            if (mg_fluid(in[c])) out = zt[c] + (r[c] - bottom_az(in[c], zt, c, sy, szz)) * mg_wd(in[c], omega);
            z[c] = out;
        }
        __syncthreads();
        if (nu == 2) {
            for (int c = tid; c < n; c += T) {
                float out = 0.f;
                if (mg_fluid(in[c])) out = z[c] + (r[c] - bottom_az(in[c], z, c, sy, szz)) * mg_wd(in[c], omega);
                zt[c] = out;
            }
            __syncthreads();
            for (int c = tid; c < n; c += T) z[c] = zt[c];
            __syncthreads();
        }
    }
    {
        const int n = bt.dx[0] * bt.dy[0] * bt.dz[0];
        for (int c = tid; c < n; c += T) {
            const int g = bottom_index(bt, 0, c);
            if (g >= 0) bt.z0[g] = sz[c];
        }
    }
}

}  // namespace vare
