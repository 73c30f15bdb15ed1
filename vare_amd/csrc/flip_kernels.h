// flip_kernels.h -- gfx950 kernels of one FLIP step on the two-level sparse grid.
//
// Each kernel cites the reference stage it replaces.  All kernels are bandwidth-bound integer /
// stencil / scatter work (no MFMA).  Floating-point expressions keep the reference's operation
// order; the library is compiled with -ffp-contract=off so no FMA is formed behind them.
#pragma once
#include "flip_device.h"
#include "flip_collision.h"

namespace vare {

// ------------------------------------------------------------------------------------------------
// Generic fill kernels (graph-safe replacement for memset).
struct k_fill_u32_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, uint32_t* __restrict__ a, uint32_t v, size_t n) {
        const size_t i = (size_t)vbid * WG + threadIdx.x;
        if (i < n) a[i] = v;
    }
};
__global__ void k_fill_u32(uint32_t* __restrict__ a, uint32_t v, size_t n) {
    k_fill_u32_f::run(blockIdx.x, gridDim.x, a, v, n);
}

// ------------------------------------------------------------------------------------------------
// Binning + block activation: SparseFrame::buildFromPoints K1 (SparseFrame.cu:141-161).
// Each particle marks the 3x3x3 block neighbourhood starting at c-1 with c = max((int)p/B, 0).
// The linear index uses the reference's signed int arithmetic, so a -1 coordinate aliases into
// the next axis exactly as in the reference; indices below zero (out of bounds in the reference)
// are dropped.  Only "count > 0" matters downstream, so a plain store replaces the atomicAdd.
//
// Two passes.  k_bin_occ records which blocks c hold a particle (occ, stamped with the step
// number, on the frame extended by one block per axis: a particle with c = f still marks the
// layer c-1 = f-1, one with c > f marks nothing); lanes whose block equals the previous lane's skip
// the store, and particles stored in bin order give one store per wave.  k_bin_dilate then marks the
// neighbourhood of every occupied block once, so a block's 27 flags are written once per step
// rather than once per wave of its particles.
// A workgroup covers kOccSlices consecutive slices of WG particles, one particle per thread per slice,
// with every slice's loads issued before any is used.
constexpr int kOccSlices = 4;
template <int B>
struct k_bin_occ_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
                              int np, int fx, int fy, int fz, int stamp, int* __restrict__ occ) {
        constexpr int J = kOccSlices;
        float x[J], y[J], z[J];
        int idx[J];
    #pragma unroll
        for (int j = 0; j < J; ++j) {
            idx[j] = (vbid * J + j) * WG + threadIdx.x;
            const int i = idx[j] < np ? idx[j] : 0;
            x[j] = px[i]; y[j] = py[i]; z[j] = pz[i];
        }
        const int lane = threadIdx.x & 63;
    #pragma unroll
        for (int j = 0; j < J; ++j) {
            int e = -1;
            if (idx[j] < np) {
                const int ci = rmax((int)(x[j]) / B, 0);
                const int cj = rmax((int)(y[j]) / B, 0);
                const int ck = rmax((int)(z[j]) / B, 0);
                if (ci <= fx && cj <= fy && ck <= fz) e = (ck * (fy + 1) + cj) * (fx + 1) + ci;
            }
            const int prev = __shfl_up(e, 1, 64);
            if (e >= 0 && (lane == 0 || prev != e)) occ[e] = stamp;
        }
    }
};
template <int B>
__global__ void k_bin_occ(const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
                          int np, int fx, int fy, int fz, int stamp, int* __restrict__ occ) {
    k_bin_occ_f<B>::run(blockIdx.x, gridDim.x, px, py, pz, np, fx, fy, fz, stamp, occ);
}

struct k_bin_dilate_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int fx, int fy, int fz, int stamp, const int* __restrict__ occ, int* __restrict__ flag) {
        const int e = vbid * WG + threadIdx.x;
        if (e >= (fx + 1) * (fy + 1) * (fz + 1)) return;
        if (occ[e] != stamp) return;
        const int ci = e % (fx + 1), cj = (e / (fx + 1)) % (fy + 1), ck = e / ((fx + 1) * (fy + 1));
        const int ke = rmin(ck + 1, fz - 1), je = rmin(cj + 1, fy - 1), ie = rmin(ci + 1, fx - 1);
        for (int k = ck - 1; k <= ke; ++k)
            for (int j = cj - 1; j <= je; ++j)
                for (int ii = ci - 1; ii <= ie; ++ii) {
                    const int t = k * fx * fy + j * fx + ii;
                    if (t >= 0) flag[t] = 1;
                }
    }
};
__global__ void k_bin_dilate(int fx, int fy, int fz, int stamp, const int* __restrict__ occ, int* __restrict__ flag) {
    k_bin_dilate_f::run(blockIdx.x, gridDim.x, fx, fy, fz, stamp, occ, flag);
}

// Water-tank activation (SparseFrame.cu:163-173): every block below the wall level is active.
struct k_bin_tank_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int fx, int fy, int fz, int tall, int* __restrict__ flag) {
        const int t = vbid * WG + threadIdx.x;
        if (t >= fx * fy * fz) return;
        const int j = (t / fx) % fy;
        if (j < tall) flag[t] = 1;
    }
};
__global__ void k_bin_tank(int fx, int fy, int fz, int tall, int* __restrict__ flag) {
    k_bin_tank_f::run(blockIdx.x, gridDim.x, fx, fy, fz, tall, flag);
}

// ------------------------------------------------------------------------------------------------
// Three-phase workgroup scan (reduce tiles -> scan tile sums -> scan tiles and emit).  A tile is
// 256 threads x ITEMS consecutive elements.  Op supplies value(i) and emit(i, prefix, value).
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = WG * SCAN_ITEMS;

// The tile's total only: any order of the sum, so the loads are coalesced (element q*WG + t of the
// tile for thread t) and all issued at once.
template <class Op>
struct k_scan_reduce_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, Op op, int n, int* __restrict__ tile_sums) {
        __shared__ int sh[4];
        const int tile0 = vbid * SCAN_TILE;
        int v[SCAN_ITEMS];
    #pragma unroll
        for (int q = 0; q < SCAN_ITEMS; ++q) {
            const int e = tile0 + q * WG + threadIdx.x;
            v[q] = e < n ? op.value(e) : 0;
        }
        int s = 0;
    #pragma unroll
        for (int q = 0; q < SCAN_ITEMS; ++q) s += v[q];
        int tot;
        block_exscan(s, sh, &tot);
        if (threadIdx.x == 0) tile_sums[vbid] = tot;
    }
};
template <class Op>
__global__ void k_scan_reduce(Op op, int n, int* __restrict__ tile_sums) {
    k_scan_reduce_f<Op>::run(blockIdx.x, gridDim.x, op, n, tile_sums);
}

// One workgroup: exclusive scan of tile sums in place; total to *total.
struct k_scan_tiles_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int* __restrict__ tile_sums, int ntiles, int* __restrict__ total) {
        __shared__ int sh[4];
        int carry = 0;
        for (int base = 0; base < ntiles; base += WG) {
            const int i = base + threadIdx.x;
            const int v = i < ntiles ? tile_sums[i] : 0;
            int tot;
            const int ex = block_exscan(v, sh, &tot);
            if (i < ntiles) tile_sums[i] = carry + ex;
            carry += tot;
        }
        if (threadIdx.x == 0) *total = carry;
    }
};
__global__ void k_scan_tiles(int* __restrict__ tile_sums, int ntiles, int* __restrict__ total) {
    k_scan_tiles_f::run(blockIdx.x, gridDim.x, tile_sums, ntiles, total);
}

// The tile is loaded and emitted in coalesced stripes (element q*WG + t for thread t) and staged in
// LDS, where thread t scans its SCAN_ITEMS consecutive elements; the LDS rows are padded to
// SCAN_ITEMS + 1 words so that those reads do not conflict on the banks.
template <class Op>
struct k_scan_apply_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, Op op, int n, const int* __restrict__ tile_sums) {
        constexpr int RW = SCAN_ITEMS + 1;                 // padded row of one thread's elements
        __shared__ int sh[4];
        __shared__ int sv[WG * RW];
        const int tile0 = vbid * SCAN_TILE;
        auto at = [](int el) { return (el / SCAN_ITEMS) * RW + el % SCAN_ITEMS; };
        int lv[SCAN_ITEMS];
    #pragma unroll
        for (int q = 0; q < SCAN_ITEMS; ++q) {
            const int e = tile0 + q * WG + threadIdx.x;
            lv[q] = e < n ? op.value(e) : 0;
        }
    #pragma unroll
        for (int q = 0; q < SCAN_ITEMS; ++q) sv[at(q * WG + threadIdx.x)] = lv[q];
        __syncthreads();
        int v[SCAN_ITEMS];
        int s = 0;
    #pragma unroll
        for (int q = 0; q < SCAN_ITEMS; ++q) {
            v[q] = sv[threadIdx.x * RW + q];
            s += v[q];
        }
        int tot;
        int run = tile_sums[vbid] + block_exscan(s, sh, &tot);
        // exclusive prefixes of this thread's elements replace their values
    #pragma unroll
        for (int q = 0; q < SCAN_ITEMS; ++q) {
            sv[threadIdx.x * RW + q] = run;
            run += v[q];
        }
        __syncthreads();
    #pragma unroll
        for (int q = 0; q < SCAN_ITEMS; ++q) {
            const int el = q * WG + threadIdx.x, e = tile0 + el;
            if (e < n) op.emit(e, sv[at(el)], lv[q]);
        }
    }
};
template <class Op>
__global__ void k_scan_apply(Op op, int n, const int* __restrict__ tile_sums) {
    k_scan_apply_f<Op>::run(blockIdx.x, gridDim.x, op, n, tile_sums);
}

// Block compaction (SparseFrame.cu:175-196): slot[t] = rank or -1, coord[rank] = t.
struct OpBlocks {
    const int* flag;
    int* slot;
    int* coord;
    __device__ int value(int t) const { return flag[t] > 0 ? 1 : 0; }
    __device__ void emit(int t, int pre, int v) const {
        slot[t] = v ? pre : -1;
        if (v) coord[pre] = t;
    }
};

// Fluid enumeration (FluidProjection.cu:178-186): rank of every fluid cell in storage order.
struct OpFluid {
    const uint8_t* type;
    int* rank;
    __device__ int value(int c) const { return type[c] == kFluid ? 1 : 0; }
    __device__ void emit(int c, int pre, int v) const { rank[c] = v ? pre : -1; }
};

// Removal compaction, part 1: ranks of removed slots below newN; hole[rank] = slot.
struct OpHoles {
    const uint8_t* removed;
    int* hole;
    __device__ int value(int i) const { return removed[i] ? 1 : 0; }
    __device__ void emit(int i, int pre, int v) const {
        if (v) hole[pre] = i;
    }
};
// part 2: survivors at or above newN, in ascending order: src[rank] = newN + offset.
struct OpSurvivors {
    const uint8_t* removed;
    int base;
    int* src;
    __device__ int value(int o) const { return removed[base + o] ? 0 : 1; }
    __device__ void emit(int o, int pre, int v) const {
        if (v) src[pre] = base + o;
    }
};

// Neighbour block table: nbr[s*6+q] = slot of the neighbour block, -1 inactive, -2 outside.
struct k_block_nbr_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int nb, int fx, int fy, int fz, const int* __restrict__ coord, const int* __restrict__ slot,
                                int* __restrict__ nbr) {
        const int e = vbid * WG + threadIdx.x;
        if (e >= nb * 6) return;
        const int s = e / 6, q = e - 6 * (e / 6);
        const int t = coord[s];
        int bi = t % fx, bj = (t / fx) % fy, bk = t / (fx * fy);
        const int d = (q & 1) ? 1 : -1;
        if (q < 2) bi += d; else if (q < 4) bj += d; else bk += d;
        int r;
        if (bi < 0 || bi >= fx || bj < 0 || bj >= fy || bk < 0 || bk >= fz) r = -2;
        else r = slot[(bk * fy + bj) * fx + bi];
        nbr[e] = r;
    }
};
__global__ void k_block_nbr(int nb, int fx, int fy, int fz, const int* __restrict__ coord, const int* __restrict__ slot,
                            int* __restrict__ nbr) {
    k_block_nbr_f::run(blockIdx.x, gridDim.x, nb, fx, fy, fz, coord, slot, nbr);
}

// ------------------------------------------------------------------------------------------------
// P2G: Rasterization::PointsToVelField (Rasterization.cu:121-173) fused with PointsToCountField
// (Rasterization.cu:198-228).  One thread per particle, float atomics into the block-major
// fields.  Summation order is that of the atomics (the reference has the same non-determinism).
struct CellFields {
    float* mass;
    float* vx;
    float* vy;
    float* vz;
    float* cnt;
};

template <int B>
struct k_p2g_atomic_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
                                 VelPtr u, int np, FrameK f, CellFields cf) {
        const int n = vbid * WG + threadIdx.x;
        if (n >= np) return;
        const float x = px[n], y = py[n], z = pz[n];
        // Grid::inside (Grid.h:407-417)
        if (x < 0.f || x > (float)f.nx || y < 0.f || y > (float)f.ny || z < 0.f || z > (float)f.nz) return;
        const float vx = u.ld(0, n), vy = u.ld(1, n), vz = u.ld(2, n);
        const int ci = rmax((int)(x - 0.5f), 0), cj = rmax((int)(y - 0.5f), 0), ck = rmax((int)(z - 0.5f), 0);
        const int ie = rmin(ci + 1, f.nx - 1), je = rmin(cj + 1, f.ny - 1), ke = rmin(ck + 1, f.nz - 1);
        for (int k = ck; k <= ke; ++k)
            for (int j = cj; j <= je; ++j)
                for (int i = ci; i <= ie; ++i) {
                    const int c = cell_index<B>(f, i, j, k);
                    if (c < 0) continue;
                    const float dx = x - ((float)i + 0.5f), dy = y - ((float)j + 0.5f), dz = z - ((float)k + 0.5f);
                    const float xw = rmax(0.f, 1.f - rabs(dx));
                    const float yw = rmax(0.f, 1.f - rabs(dy));
                    const float zw = rmax(0.f, 1.f - rabs(dz));
                    const float w = xw * yw * zw + 1e-05f;
                    unsafeAtomicAdd(cf.mass + c, w);
                    unsafeAtomicAdd(cf.vx + c, w * vx);
                    unsafeAtomicAdd(cf.vy + c, w * vy);
                    unsafeAtomicAdd(cf.vz + c, w * vz);
                    unsafeAtomicAdd(cf.cnt + c, 1.f);
                }
    }
};
template <int B>
__global__ void k_p2g_atomic(const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
                             VelPtr u, int np, FrameK f, CellFields cf) {
    k_p2g_atomic_f<B>::run(blockIdx.x, gridDim.x, px, py, pz, u, np, f, cf);
}

// P2G normalise (Rasterization.cu:175-188) + cell classification and comp := prev
// (FLIPMethod.cu:242-280, curl disabled).
template <int B>
struct k_cell_finalize_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int nc, CellFields cf, float* __restrict__ cx, float* __restrict__ cy,
                                    float* __restrict__ cz, uint8_t* __restrict__ type, FrameK f, ColK col, float col_scale) {
        const int c = vbid * WG + threadIdx.x;
        if (c >= nc) return;
        const float m = cf.mass[c];
        float vx = 0.f, vy = 0.f, vz = 0.f;
        uint8_t t = kAir;
        if (m > 0.f) {
            vx = cf.vx[c] / m;
            vy = cf.vy[c] / m;
            vz = cf.vz[c] / m;
            t = kFluid;
        } else {
            cf.mass[c] = 0.f;
        }
        if (col.any()) {
            const int t0 = f.coord[c / Blk<B>::B3], l = c % Blk<B>::B3;
            const int i = (t0 % f.fx) * B + l % B, j = ((t0 / f.fx) % f.fy) * B + (l / B) % B,
                      k = (t0 / (f.fx * f.fy)) * B + l / (B * B);
            t = classify_cell<B>(col, f, col_scale, i, j, k, m, vx, vy, vz);
        }
        cf.vx[c] = vx; cf.vy[c] = vy; cf.vz[c] = vz;
        cx[c] = vx; cy[c] = vy; cz[c] = vz;
        type[c] = t;
    }
};
template <int B>
__global__ void k_cell_finalize(int nc, CellFields cf, float* __restrict__ cx, float* __restrict__ cy,
                                float* __restrict__ cz, uint8_t* __restrict__ type, FrameK f, ColK col, float col_scale) {
    k_cell_finalize_f<B>::run(blockIdx.x, gridDim.x, nc, cf, cx, cy, cz, type, f, col, col_scale);
}

// ------------------------------------------------------------------------------------------------
// Divergence + stencil descriptor: FluidProjection::buildLinearSystem K6 (FluidProjection.cu:203-267),
// matrix-free: the ELL row [diag, -1 per fluid neighbour] is encoded in 2 bytes.  Also initialises the
// CG vectors (x = 0, r = b - A*0 = b, p = r; LinearSystemSolver.cpp:17-19) and reduces |b|^2.
//
// Reductions: every reducing kernel runs a capped grid (grid-stride loop), each workgroup writes one
// FP64 partial, and a one-workgroup finalize kernel sums the partials in a fixed order (deterministic)
// and updates the device-resident solver scalars.  No atomics or device-scope fences are needed.
struct CgVecs {
    float* x;
    float* r;
    float* p;
    float* ap;
    float* b;
    float* z;         // preconditioned residual (multigrid); unused for none/Jacobi
    uint16_t* info;
};

struct CgState {
    double rr;        // r.r of the current residual (double accumulation)
    double rz;        // r.z (preconditioned)
    double bb;        // b.b
    double pap;
    float rsold;      // reference-mode scalars, kept in FP32 like LinearSystemSolver.cpp:21-41
    float alpha;
    float beta;
    float rzold;
    int iter;         // iterations executed
    int done;
    int count;        // return value of LinearSystemSolver::cg
    int converged;
    int mode;         // 0 reference, 1 relative
    int max_count;    // reference: params.projIteration; relative: max iterations
    int loop_limit;   // reference: b.size()+5
    float abs_tol;
    float rel_tol;
    int precond;      // 0 none, 1 Jacobi, 2 multigrid
    int nfluid;       // fluid cells (rows of the system)
    int ir_done;      // iterative refinement reached its tolerance
    double ir_bb;     // |b|^2 of the original system (refinement)
    double ir_rr;     // |b - A x64|^2 of the last refinement step
    double true_rr;   // |b - A x|^2 recomputed in FP64 from the final FP32 pressure
    double div_bb;    // |b(v_post)|^2: the divergence rule applied to the projected velocity
    int ir_rounds;    // refinement rounds run (k_fin_ir) since the solve's set-up
    int cont;         // the last round continues the FP32 solve (1) or restarts it (0)
    int last_count;   // count and iter of the FP32 solve the last round followed
    int last_iter;
};

// Cells visited by the solver's vector kernels: the cells of the listed blocks (those holding a
// fluid cell), virtual index v -> cell (blk[v >> sh3] << sh3) | (v & (2^sh3 - 1)).  Blocks without
// fluid keep x = r = p = 0 from k_divergence and are never read by the stencils.
struct CellList {
    const int* blk;
    int sh3;
    __device__ __forceinline__ int cell(int v) const { return (blk[v >> sh3] << sh3) | (v & ((1 << sh3) - 1)); }
};

// Exclusive scan of the per-block fluid flags -> list of blocks holding a fluid cell.
struct OpFluidBlocks {
    const int* flag;
    int* list;
    __device__ int value(int b) const { return flag[b] != 0 ? 1 : 0; }
    __device__ void emit(int b, int pre, int v) const { if (v) list[pre] = b; }
};

// NormOnly: only |b|^2 of the given velocity (the post-projection divergence report); the solver's
// vectors and block flags are left alone.
template <int B, bool NormOnly = false>
struct k_divergence_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int nc, FrameK f, const uint8_t* __restrict__ type, const float* __restrict__ cx,
                                 const float* __restrict__ cy, const float* __restrict__ cz, float wall_level, CgVecs v,
                                 double* __restrict__ partials, double* __restrict__ partials2, int* __restrict__ bfluid) {
        __shared__ double sh[4];
        double bb = 0.0, nfl = 0.0;
        // nc is a multiple of 64 and the loop stride a multiple of WG, so every wave stays whole and
        // its 64 cells lie in one block
        for (int c = vbid * WG + threadIdx.x; c < nc; c += vnbid * WG) {
            float d = 0.f;
            uint16_t info = 0;
            if (type[c] == kFluid) {
                int i, j, k;
                cell_coords<B>(f, c, i, j, k);
                const uint8_t wall = ((float)j + 0.5f <= wall_level + 1.f) ? kWall : kAir;
                int nid[6];
                uint8_t tn[6];
    #pragma unroll
                for (int q = 0; q < 6; ++q) {
                    nid[q] = cell_neighbor<B>(f, c, q);
                    tn[q] = nid[q] == -2 ? wall : (nid[q] < 0 ? (uint8_t)kAir : type[nid[q]]);
                }
                const float own[3] = {cx[c], cy[c], cz[c]};
                auto comp = [&](int q) -> float {       // neighbour velocity component along q's axis
                    const int id = nid[q];
                    const float* a = q < 2 ? cx : (q < 4 ? cy : cz);
                    return id >= 0 ? a[id] : 0.f;
                };
                // neighbour velocities with domain-boundary extrapolation (FluidProjection.cu:220-225)
                float nv[6];
    #pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const int lo = 2 * a, hi = 2 * a + 1;
                    nv[lo] = nid[lo] == -2 ? 2.f * own[a] - comp(hi) : comp(lo);
                    nv[hi] = nid[hi] == -2 ? 2.f * own[a] - comp(lo) : comp(hi);
                }
                // face velocities (FluidProjection.cu:227-232)
                float fv[6];
    #pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const int lo = 2 * a, hi = 2 * a + 1;
                    fv[lo] = tn[lo] == kAir ? 2.f * own[a] - nv[hi] : (tn[lo] == kWall ? rabs(own[a]) : nv[lo]);
                    fv[hi] = tn[hi] == kAir ? 2.f * own[a] - nv[lo] : (tn[hi] == kWall ? -rabs(own[a]) : nv[hi]);
                }
                d = -((fv[1] - fv[0]) * 0.5f + (fv[3] - fv[2]) * 0.5f + (fv[5] - fv[4]) * 0.5f);
                int diag = 6;
                uint16_t mask = 0;
    #pragma unroll
                for (int q = 0; q < 6; ++q) {
                    if (tn[q] == kAir) diag += 1;
                    else if (tn[q] == kSolid || tn[q] == kWall) diag -= 1;
                    else mask |= (uint16_t)(1u << q);
                }
                info = (uint16_t)(kInfoFluid | mask | (diag << 8));
                nfl += 1.0;
            }
            bb += (double)d * (double)d;
            if (NormOnly) continue;
            v.b[c] = d;
            v.info[c] = info;
            v.x[c] = 0.f;
            v.r[c] = d;
            v.p[c] = d;
            v.ap[c] = 0.f;
            if (__ballot(info != 0) && (threadIdx.x & 63) == 0) bfluid[c / Blk<B>::B3] = 1;
        }
        const double s = block_sum(bb, sh);
        if (NormOnly) {
            if (threadIdx.x == 0) partials[vbid] = s;
            return;
        }
        const double s2 = block_sum(nfl, sh);
        if (threadIdx.x == 0) { partials[vbid] = s; partials2[vbid] = s2; }
    }
};
template <int B, bool NormOnly = false>
__global__ void k_divergence(int nc, FrameK f, const uint8_t* __restrict__ type, const float* __restrict__ cx,
                             const float* __restrict__ cy, const float* __restrict__ cz, float wall_level, CgVecs v,
                             double* __restrict__ partials, double* __restrict__ partials2, int* __restrict__ bfluid) {
    k_divergence_f<B, NormOnly>::run(blockIdx.x, gridDim.x, nc, f, type, cx, cy, cz, wall_level, v, partials, partials2, bfluid);
}


// Deterministic sum of n partials by one 256-thread workgroup.  Each thread keeps 8 running sums over
// strided partials, so 8 loads are in flight per thread, and combines them in a fixed order.
// The finalize kernels run as one workgroup of kFinThreads threads: each thread sums its partials in
// rounds of 8 loads issued together (kFinThreads * 8 partials per round, one round for up to 8192
// producer workgroups), then the waves' sums are added in a fixed order.  The order of every addition
// is fixed, so the result is deterministic.
constexpr int kFinThreads = 1024;
constexpr int kFinWaves = kFinThreads / 64;
__device__ __forceinline__ double fin_sum(const double* __restrict__ partials, int n, double* sh /*[kFinWaves]*/) {
    constexpr int T = kFinThreads;
    double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int base = 0; base < n; base += 8 * T) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int q = base + j * T + threadIdx.x;
            a[j] += q < n ? partials[q] : 0.0;
        }
    }
    double v = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int q = 0; q < kFinWaves; ++q) r += sh[q];
    __syncthreads();
    return r;
}

struct k_fin_divergence_f {
    static constexpr int kLB = kFinThreads;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, CgState* __restrict__ st, const double* __restrict__ partials,
                                     const double* __restrict__ partials2, int n) {
        __shared__ double sh[kFinWaves];
        const double tot = fin_sum(partials, n, sh);
        const double tot2 = fin_sum(partials2, n, sh);
        if (threadIdx.x != 0) return;
        st->nfluid = (int)tot2;
        st->loop_limit = (int)tot2 + 5;     // LinearSystemSolver.cpp:25-27: i < b.size()+5
        st->bb = tot;
        st->ir_bb = tot;
        st->ir_rr = tot;
        st->ir_done = 0;
        st->true_rr = 0.0;
        st->rr = tot;
        st->iter = 0;
        st->count = 0;
        st->converged = 0;
        st->done = 0;
        const float rs = (float)tot;
        st->rsold = rs;
        st->rzold = rs;
        if (st->mode == 0) {
            if (sqrtf(rs) < 1e-30f) { st->done = 1; st->converged = 1; }    // LinearSystemSolver.cpp:23
        } else {
            if (tot == 0.0) { st->done = 1; st->converged = 1; }
        }
    }
};
__global__ void __launch_bounds__(kFinThreads) k_fin_divergence(CgState* __restrict__ st, const double* __restrict__ partials,
                                 const double* __restrict__ partials2, int n) {
    k_fin_divergence_f::run(blockIdx.x, gridDim.x, st, partials, partials2, n);
}

struct k_fin_divpost_f {
    static constexpr int kLB = kFinThreads;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, CgState* __restrict__ st, const double* __restrict__ partials,
                                                                   int n) {
        __shared__ double sh[kFinWaves];
        const double tot = fin_sum(partials, n, sh);
        if (threadIdx.x == 0) st->div_bb = tot;
    }
};
__global__ void __launch_bounds__(kFinThreads) k_fin_divpost(CgState* __restrict__ st, const double* __restrict__ partials,
                                                               int n) {
    k_fin_divpost_f::run(blockIdx.x, gridDim.x, st, partials, n);
}

// Jacobi preconditioner value for a cell (0 off the fluid set).
__device__ __forceinline__ float jacobi_inv(uint16_t info) {
    const int d = (info >> 8) & 0x1f;
    return ((info & kInfoFluid) && d > 0) ? 1.f / (float)d : 0.f;
}

// Relative mode set-up (after k_divergence): z = M^-1 r, p = z, partial r.z.  For multigrid the
// V-cycle has already written z.
struct k_cg_precond_init_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int nv, CellList cl, CgVecs v, double* __restrict__ partials,
                                      const CgState* __restrict__ st) {
        __shared__ double sh[4];
        double rz = 0.0;
        const int pc = st->precond;
        for (int vi = vbid * WG + threadIdx.x; vi < nv; vi += vnbid * WG) {
            const int c = cl.cell(vi);
            const float r = v.r[c];
            const float z = pc == 1 ? r * jacobi_inv(v.info[c]) : (pc == 2 ? v.z[c] : r);
            v.p[c] = z;
            rz += (double)r * (double)z;
        }
        const double s = block_sum(rz, sh);
        if (threadIdx.x == 0) partials[vbid] = s;
    }
};
__global__ void k_cg_precond_init(int nv, CellList cl, CgVecs v, double* __restrict__ partials,
                                  const CgState* __restrict__ st) {
    k_cg_precond_init_f::run(blockIdx.x, gridDim.x, nv, cl, v, partials, st);
}

struct k_fin_precond_init_f {
    static constexpr int kLB = kFinThreads;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, CgState* __restrict__ st, const double* __restrict__ partials, int n) {
        __shared__ double sh[kFinWaves];
        if (st->done) return;
        const double tot = fin_sum(partials, n, sh);
        if (threadIdx.x != 0) return;
        st->rz = tot;
        st->rzold = (float)tot;
    }
};
__global__ void __launch_bounds__(kFinThreads) k_fin_precond_init(CgState* __restrict__ st, const double* __restrict__ partials, int n) {
    k_fin_precond_init_f::run(blockIdx.x, gridDim.x, st, partials, n);
}

// alpha = rsold / p.Ap  (LinearSystemSolver.cpp:31); the partials are the products of k_spmv_blk.
struct k_fin_alpha_f {
    static constexpr int kLB = kFinThreads;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, CgState* __restrict__ st, const double* __restrict__ partials, int n) {
        __shared__ double sh[kFinWaves];
        if (st->done) return;
        const double tot = fin_sum(partials, n, sh);
        if (threadIdx.x != 0) return;
        st->pap = tot;
        st->alpha = (st->mode == 0 ? st->rsold : st->rzold) / (float)tot;
    }
};
__global__ void __launch_bounds__(kFinThreads) k_fin_alpha(CgState* __restrict__ st, const double* __restrict__ partials, int n) {
    k_fin_alpha_f::run(blockIdx.x, gridDim.x, st, partials, n);
}

// CG step 2: x += alpha p, r -= alpha Ap (LinearSystemSolver.cpp:33-35), partial r.r (and r.z for
// Jacobi).
struct k_cg_update_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int nv, CellList cl, CgVecs v, double* __restrict__ partials,
                                double* __restrict__ partials2, const CgState* __restrict__ st) {
        __shared__ double sh[4];
        if (st->done) return;
        const float alpha = st->alpha;
        const int precond = st->precond;
        double rr = 0.0, rz = 0.0;
        for (int vi = vbid * WG + threadIdx.x; vi < nv; vi += vnbid * WG) {
            const int c = cl.cell(vi);
            const float p = v.p[c];
            v.x[c] = v.x[c] + (alpha * p);
            const float r = v.r[c] - (alpha * v.ap[c]);
            v.r[c] = r;
            rr += (double)r * (double)r;
            if (precond == 1) rz += (double)r * (double)(r * jacobi_inv(v.info[c]));
        }
        const double s = block_sum(rr, sh);
        const double s2 = precond == 1 ? block_sum(rz, sh) : 0.0;
        if (threadIdx.x == 0) { partials[vbid] = s; partials2[vbid] = s2; }
    }
};
__global__ void k_cg_update(int nv, CellList cl, CgVecs v, double* __restrict__ partials,
                            double* __restrict__ partials2, const CgState* __restrict__ st) {
    k_cg_update_f::run(blockIdx.x, gridDim.x, nv, cl, v, partials, partials2, st);
}

// Convergence test and beta (LinearSystemSolver.cpp:37-46).  For multigrid, beta is computed after
// the V-cycle by k_fin_beta.
struct k_fin_update_f {
    static constexpr int kLB = kFinThreads;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, CgState* __restrict__ st, const double* __restrict__ partials,
                                 const double* __restrict__ partials2, int n) {
        __shared__ double sh[kFinWaves];
        if (st->done) return;
        const double tot = fin_sum(partials, n, sh);
        const double tot2 = st->precond == 1 ? fin_sum(partials2, n, sh) : 0.0;
        if (threadIdx.x != 0) return;
        st->rr = tot;
        const int i = st->iter;
        st->iter = i + 1;
        if (st->mode == 0) {
            const float rsnew = (float)tot;
            if (sqrtf(rsnew) < st->abs_tol) {
                st->done = 1; st->count = i; st->converged = 1;
            } else {
                st->beta = rsnew / st->rsold;
                st->rsold = rsnew;
                if (i > st->max_count) { st->done = 1; st->count = i; }
                else if (i + 1 >= st->loop_limit) { st->done = 1; st->count = i + 1; }
            }
        } else {
            if (sqrt(tot) <= (double)st->rel_tol * sqrt(st->bb)) {
                st->done = 1; st->count = i + 1; st->converged = 1;
            } else if (i + 1 >= st->max_count) {
                st->done = 1; st->count = i + 1;
            } else if (st->precond != 2) {
                const float rznew = st->precond == 1 ? (float)tot2 : (float)tot;
                st->beta = rznew / st->rzold;
                st->rzold = rznew;
            }
        }
    }
};
__global__ void __launch_bounds__(kFinThreads) k_fin_update(CgState* __restrict__ st, const double* __restrict__ partials,
                             const double* __restrict__ partials2, int n) {
    k_fin_update_f::run(blockIdx.x, gridDim.x, st, partials, partials2, n);
}

// Multigrid: beta from the r.z partials written by the V-cycle's last level-0 pass.
struct k_fin_beta_f {
    static constexpr int kLB = kFinThreads;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, CgState* __restrict__ st, const double* __restrict__ partials, int n) {
        __shared__ double sh[kFinWaves];
        if (st->done) return;
        const double tot = fin_sum(partials, n, sh);
        if (threadIdx.x != 0) return;
        const float rznew = (float)tot;
        st->beta = rznew / st->rzold;
        st->rzold = rznew;
    }
};
__global__ void __launch_bounds__(kFinThreads) k_fin_beta(CgState* __restrict__ st, const double* __restrict__ partials, int n) {
    k_fin_beta_f::run(blockIdx.x, gridDim.x, st, partials, n);
}

// ------------------------------------------------------------------------------------------------
// Pressure apply.  The solve runs in place on the cell layout, so FluidProjection.cu:297-306
// (scatter x to the field) has no counterpart.  Air extrapolation, FluidProjection.cu:314-335:
// each active air cell takes the mean pressure of its fluid neighbours (k_press_extrap_blk), then the
// gradient is subtracted at the fluid cells (k_press_grad_blk, FluidProjection.cu:342-378).  PT is the
// pressure type: float, or double for the FP64 pressure of iterative refinement (arithmetic in PT).
// ------------------------------------------------------------------------------------------------
// G2P + FLIP/PIC blend + advection + wall clamp + removal flag + gravity:
// FLIPMethod::updateParticles (FLIPMethod.cu:86-171) with SparseField::lerp/lerpMin
// (SparseField.h:150-196).  No collision source (CollisionSource.h:128-144 with empty fields:
// signedDistance = 5, so the collision branches are never taken).
struct G2PParams {
    float dt;
    float flip;
    float pic;
    float wall;
    float gx, gy, gz;
    int max_substeps;
    int tank;
    float band;
};

// Removal compaction, part 3: survivor src[S-1-k] moves into hole[k] (Array.h:995-1024 closed form:
// the k-th removed slot below newN, ascending, receives the k-th survivor at or above newN,
// descending).
struct PArrays {
    float* a[9];     // px,py,pz,vx,vy,vz,adj,vort,curv (vx..vz are FP16 when half != 0)
    int half;
    __device__ __forceinline__ VelPtr vel() const { return VelPtr{{a[3], a[4], a[5]}, half}; }
};
// Arrays moved together by one compaction: up to 8, each of 4- or 2-byte elements.
struct MoveSet {
    void* a[8];
    int esz[8];
    int n;
};
// The k-th hole (ascending) receives the k-th survivor from the end (descending): the closed form of
// Array::remove's swap-from-back loop (Array.h:995-1024).
struct k_remove_move_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int nmove, const int* __restrict__ hole, const int* __restrict__ src, MoveSet ms) {
        const int k = vbid * WG + threadIdx.x;
        if (k >= nmove) return;
        const int d = hole[k];
        const int s = src[nmove - 1 - k];
    #pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (q >= ms.n) break;
            if (ms.esz[q] == 2) ((_Float16*)ms.a[q])[d] = ((const _Float16*)ms.a[q])[s];
            else ((float*)ms.a[q])[d] = ((const float*)ms.a[q])[s];
        }
    }
};
__global__ void k_remove_move(int nmove, const int* __restrict__ hole, const int* __restrict__ src, MoveSet ms) {
    k_remove_move_f::run(blockIdx.x, gridDim.x, nmove, hole, src, ms);
}

// Logical removal mask from the physical one (G2P writes flags per physical particle).
struct k_removed_logical_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int np, const uint8_t* __restrict__ removed, const int* __restrict__ pid,
                                      uint8_t* __restrict__ removed_log) {
        const int q = vbid * WG + threadIdx.x;
        if (q >= np) return;
        if (removed[q]) removed_log[pid[q]] = 1;
    }
};
__global__ void k_removed_logical(int np, const uint8_t* __restrict__ removed, const int* __restrict__ pid,
                                  uint8_t* __restrict__ removed_log) {
    k_removed_logical_f::run(blockIdx.x, gridDim.x, np, removed, pid, removed_log);
}

// New logical index of every survivor at or above newN: the one the compaction moves into hole k.
struct k_remap_build_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int nmove, int newn, const int* __restrict__ hole, const int* __restrict__ src,
                                  int* __restrict__ remap) {
        const int k = vbid * WG + threadIdx.x;
        if (k >= nmove) return;
        remap[src[nmove - 1 - k] - newn] = hole[k];
    }
};
__global__ void k_remap_build(int nmove, int newn, const int* __restrict__ hole, const int* __restrict__ src,
                              int* __restrict__ remap) {
    k_remap_build_f::run(blockIdx.x, gridDim.x, nmove, newn, hole, src, remap);
}
struct k_pid_remap_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int n, int newn, const int* __restrict__ remap, int* __restrict__ pid) {
        const int q = vbid * WG + threadIdx.x;
        if (q >= n) return;
        const int l = pid[q];
        if (l >= newn) pid[q] = remap[l - newn];
    }
};
__global__ void k_pid_remap(int n, int newn, const int* __restrict__ remap, int* __restrict__ pid) {
    k_pid_remap_f::run(blockIdx.x, gridDim.x, n, newn, remap, pid);
}

struct k_iota_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int n, int base, int* __restrict__ out) {
        const int i = vbid * WG + threadIdx.x;
        if (i < n) out[i] = base + i;
    }
};
__global__ void k_iota(int n, int base, int* __restrict__ out) {
    k_iota_f::run(blockIdx.x, gridDim.x, n, base, out);
}

// ------------------------------------------------------------------------------------------------
// Voxel <-> world (Grid.h:431-457), particlesToVoxel / particlesToWorld (FLIPMethod.cu:480-514).
struct k_to_voxel_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int np, PArrays pa, float tx, float ty, float tz, float sx, float sy, float sz) {
        const int n = vbid * WG + threadIdx.x;
        if (n >= np) return;
        const VelPtr u = pa.vel();
        pa.a[0][n] = (pa.a[0][n] - tx) / sx;
        pa.a[1][n] = (pa.a[1][n] - ty) / sy;
        pa.a[2][n] = (pa.a[2][n] - tz) / sz;
        u.st(0, n, u.ld(0, n) / sx);
        u.st(1, n, u.ld(1, n) / sy);
        u.st(2, n, u.ld(2, n) / sz);
    }
};
__global__ void k_to_voxel(int np, PArrays pa, float tx, float ty, float tz, float sx, float sy, float sz) {
    k_to_voxel_f::run(blockIdx.x, gridDim.x, np, pa, tx, ty, tz, sx, sy, sz);
}
struct k_to_world_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int np, PArrays pa, float tx, float ty, float tz, float sx, float sy, float sz) {
        const int n = vbid * WG + threadIdx.x;
        if (n >= np) return;
        const VelPtr u = pa.vel();
        pa.a[0][n] = pa.a[0][n] * sx + tx;
        pa.a[1][n] = pa.a[1][n] * sy + ty;
        pa.a[2][n] = pa.a[2][n] * sz + tz;
        u.st(0, n, u.ld(0, n) * sx);
        u.st(1, n, u.ld(1, n) * sy);
        u.st(2, n, u.ld(2, n) * sz);
    }
};
__global__ void k_to_world(int np, PArrays pa, float tx, float ty, float tz, float sx, float sy, float sz) {
    k_to_world_f::run(blockIdx.x, gridDim.x, np, pa, tx, ty, tz, sx, sy, sz);
}

// Particle export: interleaved FP32 positions/velocities (voxel space, or world space when world
// != 0: Grid::worldPoint / worldVector, Grid.h:431-443).  Either output may be null.
// Particle export in the reference's order: physical particle q goes to slot pid[q] of the
// xyz-interleaved outputs (and of adj); optional world transform.
struct k_export_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int np, PArrays pa, const int* __restrict__ pid, int world, float tx, float ty, float tz,
                             float sx, float sy, float sz, float* __restrict__ pos, float* __restrict__ vel,
                             float* __restrict__ adj) {
        const int q = vbid * WG + threadIdx.x;
        if (q >= np) return;
        const int n = pid[q];
        const VelPtr u = pa.vel();
        float x = pa.a[0][q], y = pa.a[1][q], z = pa.a[2][q];
        float vx = u.ld(0, q), vy = u.ld(1, q), vz = u.ld(2, q);
        if (world) {
            x = x * sx + tx; y = y * sy + ty; z = z * sz + tz;
            vx = vx * sx; vy = vy * sy; vz = vz * sz;
        }
        if (pos) { pos[3 * n] = x; pos[3 * n + 1] = y; pos[3 * n + 2] = z; }
        if (vel) { vel[3 * n] = vx; vel[3 * n + 1] = vy; vel[3 * n + 2] = vz; }
        if (adj) adj[n] = pa.a[6][q];
    }
};
__global__ void k_export(int np, PArrays pa, const int* __restrict__ pid, int world, float tx, float ty, float tz,
                         float sx, float sy, float sz, float* __restrict__ pos, float* __restrict__ vel,
                         float* __restrict__ adj) {
    k_export_f::run(blockIdx.x, gridDim.x, np, pa, pid, world, tx, ty, tz, sx, sy, sz, pos, vel, adj);
}

// xyz-interleaved FP32 <-> SoA on the device (for device-pointer particle I/O and FP16 velocity
// storage): in/out are float[3n]; the SoA side is FP32 or FP16 (half).
struct k_deinterleave_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int n, const float* __restrict__ in, void* a, void* b, void* c, int half) {
        const int i = vbid * WG + threadIdx.x;
        if (i >= n) return;
        const VelPtr u{{a, b, c}, half};
        u.st(0, i, in[3 * i]); u.st(1, i, in[3 * i + 1]); u.st(2, i, in[3 * i + 2]);
    }
};
__global__ void k_deinterleave(int n, const float* __restrict__ in, void* a, void* b, void* c, int half) {
    k_deinterleave_f::run(blockIdx.x, gridDim.x, n, in, a, b, c, half);
}
struct k_interleave_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int n, const void* a, const void* b, const void* c, int half, float* __restrict__ out) {
        const int i = vbid * WG + threadIdx.x;
        if (i >= n) return;
        const VelPtr u{{(void*)a, (void*)b, (void*)c}, half};
        out[3 * i] = u.ld(0, i); out[3 * i + 1] = u.ld(1, i); out[3 * i + 2] = u.ld(2, i);
    }
};
__global__ void k_interleave(int n, const void* a, const void* b, const void* c, int half, float* __restrict__ out) {
    k_interleave_f::run(blockIdx.x, gridDim.x, n, a, b, c, half, out);
}

// Mark field in the reference encoding (uint64: 0 air, 1 solid, 4+fluid rank).
struct k_mark_u64_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int nc, const uint8_t* __restrict__ type, const int* __restrict__ rank,
                               unsigned long long* __restrict__ out) {
        const int c = vbid * WG + threadIdx.x;
        if (c >= nc) return;
        const uint8_t t = type[c];
        out[c] = t == kFluid ? (unsigned long long)(4 + rank[c]) : (unsigned long long)t;
    }
};
__global__ void k_mark_u64(int nc, const uint8_t* __restrict__ type, const int* __restrict__ rank,
                           unsigned long long* __restrict__ out) {
    k_mark_u64_f::run(blockIdx.x, gridDim.x, nc, type, rank, out);
}

}  // namespace vare

namespace vare {

// ------------------------------------------------------------------------------------------------
// Mixed-precision projection: iterative refinement.  The pressure accumulates in FP64 (x64); each
// round adds the FP32 solve's x to it (x64 += x, x then starts again at 0), recomputes the residual
// r = b - A x64 in FP64 (k_ir_residual_blk) and hands its FP32 rounding to the FP32 PCG, which either
// continues with it in place of its recursive residual or starts a correction solve A d = r afresh.
// The fused multigrid CG iteration leaves x += alpha p to the extra workgroups of its V-cycle's bottom
// launch, which does not run in the iteration that ends the solve; that last term is added here, after
// the solve: k = st->iter - 1 is the last iteration, alpha_k the state's alpha and p_k the direction
// buffer of its parity (p0 for even k).  p0 == nullptr: x is complete.
struct XTail {
    const float* p0;
    const float* p1;
    const CgState* st;
    __device__ __forceinline__ float apply(float x, int c) const {
        if (!p0) return x;
        const int it = st->iter;
        if (it <= 0) return x;
        const float* pk = (it & 1) ? p0 : p1;
        return x + (st->alpha * pk[c]);
    }
};

// x64 += d, with d the correction solve's x (its last CG term added, see XTail).  With done != nullptr
// it runs only once the solve is done (it is enqueued behind each chunk of iterations).
struct k_ir_accumulate_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int nv, CellList cl, double* __restrict__ x64, const float* __restrict__ d,
                                    float* __restrict__ x32, int last, XTail tail, const int* done) {
        if (done && !*done) return;
        for (int vi = vbid * WG + threadIdx.x; vi < nv; vi += vnbid * WG) {
            const int c = cl.cell(vi);
            const double v = x64[c] + (double)tail.apply(d[c], c);
            x64[c] = v;
            if (last) x32[c] = (float)v;
        }
    }
};
__global__ void k_ir_accumulate(int nv, CellList cl, double* __restrict__ x64, const float* __restrict__ d,
                                float* __restrict__ x32, int last, XTail tail, const int* done) {
    k_ir_accumulate_f::run(blockIdx.x, gridDim.x, nv, cl, x64, d, x32, last, tail, done);
}

// x = x + alpha_k p_k of the last iteration (XTail) when the solve is not refined.
struct k_cg_xflush_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int nv, CellList cl, float* __restrict__ x, XTail tail) {
        for (int vi = vbid * WG + threadIdx.x; vi < nv; vi += vnbid * WG) {
            const int c = cl.cell(vi);
            x[c] = tail.apply(x[c], c);
        }
    }
};
__global__ void k_cg_xflush(int nv, CellList cl, float* __restrict__ x, XTail tail) {
    k_cg_xflush_f::run(blockIdx.x, gridDim.x, nv, cl, x, tail);
}

struct k_ir_round_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int nc, const double* __restrict__ x64, float* __restrict__ x32) {
        for (int c = vbid * WG + threadIdx.x; c < nc; c += vnbid * WG) x32[c] = (float)x64[c];
    }
};
__global__ void k_ir_round(int nc, const double* __restrict__ x64, float* __restrict__ x32) {
    k_ir_round_f::run(blockIdx.x, gridDim.x, nc, x64, x32);
}

}  // namespace vare

namespace vare {

// Refinement bookkeeping after k_ir_residual_blk (it runs only once the FP32 solve is done): record the
// FP64 residual, stop at the target, or set up the next FP32 solve, whose relative tolerance is measured
// against |r| and only has to close the gap to the target: max(inner_tol, 0.8 * outer_tol * |b| / |r|),
// at most 0.5.  The solve continues (iteration count and the last r.z kept: a reliable update) when
// cont_allowed and at most max_reduction is left to gain, else it restarts.
struct k_fin_ir_f {
    static constexpr int kLB = kFinThreads;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, CgState* __restrict__ st, const double* __restrict__ partials, int n, float outer_tol,
                             float inner_tol, int cont_allowed, float max_reduction) {
        __shared__ double sh[kFinWaves];
        if (!st->done) return;                    // the solve is still running: this round comes later
        const double tot = fin_sum(partials, n, sh);
        if (threadIdx.x != 0) return;
        st->ir_rounds += 1;
        st->last_count = st->count;
        st->last_iter = st->iter;
        st->ir_rr = tot;
        st->true_rr = tot;            // |b - A x64|^2 of the current FP64 pressure, in FP64
        if (sqrt(tot) <= (double)outer_tol * sqrt(st->ir_bb) || tot == 0.0) {
            st->ir_done = 1;
            st->done = 1;
            return;
        }
        // continue the FP32 solve when at most max_reduction is left to gain, estimated from its last
        // recursive residual (the true one is within FP32 drift of it)
        const double rr_rec = st->rr;
        const int cont = cont_allowed && st->iter > 0 && rr_rec > 0.0 &&
                         (double)outer_tol * sqrt(st->ir_bb / rr_rec) * (double)max_reduction >= 1.0;
        st->cont = cont;
        // the next solve's tolerance: what the target still needs, with a margin for the FP32 recursive
        // residual running ahead of the true one (0.8: measured at 256^3, worst final residual 0.77x target)
        const double need = 0.8 * (double)outer_tol * sqrt(st->ir_bb / tot);
        st->rel_tol = (float)fmin(0.5, fmax((double)inner_tol, need));
        st->bb = tot;
        st->rr = tot;
        st->converged = 0;
        st->done = 0;
        if (cont) return;             // the solve goes on: iteration count and r.z of its last residual kept
        st->rsold = (float)tot;
        st->rzold = (float)tot;
        st->iter = 0;
        st->count = 0;
    }
};
__global__ void __launch_bounds__(kFinThreads) k_fin_ir(CgState* __restrict__ st, const double* __restrict__ partials, int n, float outer_tol,
                         float inner_tol, int cont_allowed, float max_reduction) {
    k_fin_ir_f::run(blockIdx.x, gridDim.x, st, partials, n, outer_tol, inner_tol, cont_allowed, max_reduction);
}

// |b - A x|^2 in FP64 from the pressure x (FP32, or FP64 when refining); reported with every solve.
template <int B, class PT>
struct k_true_residual_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int nc, FrameK f, const float* __restrict__ b, const PT* __restrict__ x,
                                    const uint16_t* __restrict__ info, double* __restrict__ partials) {
        __shared__ double sh[4];
        double rr = 0.0;
        for (int c = vbid * WG + threadIdx.x; c < nc; c += vnbid * WG) {
            const uint16_t in = info[c];
            if (!(in & kInfoFluid)) continue;
            double ax = (double)((in >> 8) & 0x1f) * (double)x[c];
    #pragma unroll
            for (int q = 0; q < 6; ++q)
                if (in & (1u << q)) ax -= (double)x[cell_neighbor<B>(f, c, q)];
            const double r = (double)b[c] - ax;
            rr += r * r;
        }
        const double s = block_sum(rr, sh);
        if (threadIdx.x == 0) partials[vbid] = s;
    }
};
template <int B, class PT>
__global__ void k_true_residual(int nc, FrameK f, const float* __restrict__ b, const PT* __restrict__ x,
                                const uint16_t* __restrict__ info, double* __restrict__ partials) {
    k_true_residual_f<B, PT>::run(blockIdx.x, gridDim.x, nc, f, b, x, info, partials);
}

struct k_fin_true_f {
    static constexpr int kLB = kFinThreads;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, CgState* __restrict__ st, const double* __restrict__ partials, int n) {
        __shared__ double sh[kFinWaves];
        const double tot = fin_sum(partials, n, sh);
        if (threadIdx.x == 0) st->true_rr = tot;
    }
};
__global__ void __launch_bounds__(kFinThreads) k_fin_true(CgState* __restrict__ st, const double* __restrict__ partials, int n) {
    k_fin_true_f::run(blockIdx.x, gridDim.x, st, partials, n);
}

}  // namespace vare
