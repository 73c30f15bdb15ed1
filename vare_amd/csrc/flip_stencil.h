// flip_stencil.h -- block-aligned 7-point stencil kernels on the fine level (edge B) and, for
// B = 8, on the first multigrid level (edge 4) over the same blocks.
//
// Each workgroup takes K consecutive entries of the solver's block list (the blocks holding a
// fluid cell; blocks without fluid hold no unknown and are skipped) with one thread per cell
// position (E^3 threads: 512 for E=8, 64 for E=4), so every thread serves K cells.  K is 2 or 4 per
// kernel: the passes that load many values per cell run shorter workgroups with K = 2, the lighter
// ones keep more bytes in flight with K = 4 (measured per kernel on the 256^3 dam break).
//
// Latency structure.  The slots of a workgroup's K blocks and their 6 face-neighbour slots are
// gathered once per solve into a table (k_wg_table, kTab ints per pair of blocks).  A kernel reads its
// entries with uniform loads (scalar registers) together with the solver's `done` flag, and then
// issues every value load of the pass at once: the K blocks' own values and the values across
// their faces.  A value across a face is loaded whenever the neighbour block is active and used
// only where the stencil descriptor marks that neighbour fluid, so it need not wait for the
// descriptor.  A pass is therefore one round trip for the table and one for the data.  Own values
// are staged in LDS for the neighbours inside a block.  Reductions are one FP64 partial per
// workgroup, summed in a fixed order by the finalize kernels.
#pragma once
#include "flip_kernels.h"
#include "flip_mg.h"

namespace vare {

template <int B> struct Fine {
    static constexpr int B3 = B * B * B;
    static constexpr int SH = (B == 4) ? 2 : 3;
    static constexpr int NW = B3 / 64;         // waves per workgroup
};

constexpr int kTabBlocks = 2;                  // blocks per table entry
constexpr int kTab = 16;                       // ints per table entry (kTabBlocks * 7 <= kTab)
static_assert(7 * kTabBlocks <= kTab, "workgroup table entry too small");

// Table entries needed for the solver's ns blocks: one per pair, plus one entry of -1 so that a
// 4-block workgroup reading two entries never passes the end.
inline int wg_table_entries(int ns) { return (ns + kTabBlocks - 1) / kTabBlocks + 1; }

// Table entry w: [0, 2) the slots of blocks 2w, 2w+1 of the list (-1 past its end), then the 6
// neighbour slots (-x,+x,-y,+y,-z,+z; -1 inactive, -2 outside the domain) of each.
struct k_wg_table_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, const int* __restrict__ blist, int nsb, const int* __restrict__ nbr, int nent,
                               int* __restrict__ tab) {
        constexpr int K = kTabBlocks;
        const int t = vbid * WG + threadIdx.x;
        if (t >= nent * kTab) return;
        const int w = t / kTab, e = t % kTab;
        int v = -1;
        if (e < K) {
            const int i = w * K + e;
            if (i < nsb) v = blist[i];
        } else if (e < 7 * K) {
            const int i = w * K + (e - K) / 6;
            if (i < nsb) v = nbr[blist[i] * 6 + (e - K) % 6];
        }
        tab[t] = v;
    }
};
__global__ void k_wg_table(const int* __restrict__ blist, int nsb, const int* __restrict__ nbr, int nent,
                           int* __restrict__ tab) {
    k_wg_table_f::run(blockIdx.x, gridDim.x, blist, nsb, nbr, nent, tab);
}

// The K blocks of this workgroup (K/2 consecutive table entries; uniform addresses, so they are
// loaded into scalar registers).  The empty asm statements make every value live at this point, so
// the loads are issued here, before the caller's early-exit test on the solver flag, and their
// latency overlaps the flag's.
template <int K>
__device__ __forceinline__ void wg_load(const unsigned vbid, const unsigned vnbid, const int* __restrict__ tab, int (&s)[K],
                                        int (&nb)[K][6]) {
    static_assert(K % kTabBlocks == 0, "K must be a multiple of the table entry");
    const int* t = tab + xcd_swizzle(vbid, vnbid) * (K / kTabBlocks) * kTab;
#pragma unroll
    for (int g = 0; g < K / kTabBlocks; ++g)
#pragma unroll
        for (int kb = 0; kb < kTabBlocks; ++kb) {
            s[g * kTabBlocks + kb] = t[g * kTab + kb];
#pragma unroll
            for (int q = 0; q < 6; ++q) nb[g * kTabBlocks + kb][q] = t[g * kTab + kTabBlocks + 6 * kb + q];
        }
#pragma unroll
    for (int kb = 0; kb < K; ++kb) {
        asm volatile("" ::"s"(s[kb]));
#pragma unroll
        for (int q = 0; q < 6; ++q) asm volatile("" ::"s"(nb[kb][q]));
    }
}

// Faces of the block that thread l lies on: per axis a, side[a] = 0 on the low face (direction
// 2a), 1 on the high face (2a+1), -1 for an interior position; fl[a] = the local index of the cell
// across that face in the neighbour block.
template <int B>
__device__ __forceinline__ void thread_faces(int l, int (&side)[3], int (&fl)[3]) {
    constexpr int SH = Fine<B>::SH;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int lc = (l >> (a * SH)) & (B - 1);
        const int stride = 1 << (a * SH);
        side[a] = lc == 0 ? 0 : (lc == B - 1 ? 1 : -1);
        fl[a] = lc == 0 ? l + (B - 1) * stride : l - (B - 1) * stride;
    }
}

// Slot of the neighbour block across this thread's face on axis a (negative: none).  `a` is a
// compile-time constant after unrolling, so both candidates are fixed scalar registers.
__device__ __forceinline__ int face_slot(const int (&nb)[6], int a, int side) {
    return side < 0 ? -1 : (side ? nb[2 * a + 1] : nb[2 * a]);
}

// Local index of the neighbour inside the block in direction q.
template <int B>
__device__ __forceinline__ int inner_nbr(int l, int q) {
    constexpr int SH = Fine<B>::SH;
    const int stride = 1 << ((q >> 1) * SH);
    return (q & 1) ? l + stride : l - stride;
}

template <int B>
__device__ __forceinline__ double fine_block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int q = 0; q < Fine<B>::NW; ++q) r += sh[q];
    return r;
}

// A applied to the staged field: diag * own - sum of the fluid neighbours (in the order -x,+x,
// -y,+y,-z,+z), the neighbour across a face from `halo`, inside the block from LDS.
template <int B>
__device__ __forceinline__ float apply_row(unsigned in, float own, const float* lds, int l, const int (&side)[3],
                                           const float (&halo)[3]) {
    float az = mg_diag(in) * own;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        if (!(in & (1u << q))) continue;
        const int a = q >> 1;
        az -= side[a] == (q & 1) ? halo[a] : lds[inner_nbr<B>(l, q)];
    }
    return az;
}

// Iterative refinement's FP64 residual r = b - A x64 over the solver's blocks (diag * x first, then
// the fluid neighbours subtracted in the order -x,+x,-y,+y,-z,+z), handing the FP32 rounding of r to
// the next FP32 solve (both r buffers, x = 0; the solve's set-up or its continuation forms the next
// direction), with one partial |r|^2 per workgroup.  It runs between solves: with done != nullptr only once the solve is done (it is enqueued
// behind each chunk of iterations).
template <int B, int K>
struct k_ir_residual_blk_f {
    static constexpr int kLB = Fine<B>::B3;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, const int* __restrict__ tab,
                                                                    const float* __restrict__ b,
                                                                    const double* __restrict__ x64,
                                                                    const uint16_t* __restrict__ info,
                                                                    float* __restrict__ r32, float* __restrict__ x,
                                                                    double* __restrict__ partials,
                                                                    float* __restrict__ r32b, const int* done) {
        constexpr int B3 = Fine<B>::B3;
        __shared__ double sx[K][B3];
        __shared__ double sh[Fine<B>::NW];
        const int dn = done ? *done : 1;
        int s[K], nb[K][6];
        wg_load<K>(vbid, vnbid, tab, s, nb);
        if (!dn) return;
        const int l = threadIdx.x;
        int side[3], fl[3];
        thread_faces<B>(l, side, fl);
        unsigned in[K];
        float bc[K];
        double xc[K], hx[K][3];
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) {
            in[kb] = 0;
            bc[kb] = 0.f;
            xc[kb] = 0.0;
            if (s[kb] >= 0) {
                const int c = s[kb] * B3 + l;
                in[kb] = info[c];
                bc[kb] = b[c];
                xc[kb] = x64[c];
            }
        }
    #pragma unroll
        for (int kb = 0; kb < K; ++kb)
    #pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int ns = face_slot(nb[kb], a, side[a]);
                hx[kb][a] = ns >= 0 ? x64[ns * B3 + fl[a]] : 0.0;
            }
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) sx[kb][l] = xc[kb];
        __syncthreads();
        double rr = 0.0;
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) {
            if (s[kb] < 0) continue;
            double r = 0.0;
            if (in[kb] & kInfoFluid) {
                double ax = (double)((in[kb] >> 8) & 0x1f) * xc[kb];
    #pragma unroll
                for (int q = 0; q < 6; ++q) {
                    if (!(in[kb] & (1u << q))) continue;
                    const int a = q >> 1;
                    ax -= side[a] == (q & 1) ? hx[kb][a] : sx[kb][inner_nbr<B>(l, q)];
                }
                r = (double)bc[kb] - ax;
            }
            const float rf = (float)r;
            const int c = s[kb] * B3 + l;
            r32[c] = rf;
            if (r32b) r32b[c] = rf;
            x[c] = 0.f;
            rr += r * r;
        }
        const double tot = fine_block_sum<B>(rr, sh);
        if (l == 0) partials[vbid] = tot;
    }
};
template <int B, int K>
__global__ void __launch_bounds__(Fine<B>::B3) k_ir_residual_blk(const int* __restrict__ tab,
                                                                const float* __restrict__ b,
                                                                const double* __restrict__ x64,
                                                                const uint16_t* __restrict__ info,
                                                                float* __restrict__ r32, float* __restrict__ x,
                                                                double* __restrict__ partials,
                                                                float* __restrict__ r32b, const int* done) {
    k_ir_residual_blk_f<B, K>::run(blockIdx.x, gridDim.x, tab, b, x64, info, r32, x, partials, r32b, done);
}

// Storage indices of the six neighbours of cell l of block s (cell_neighbor's values: -1 inactive, -2
// outside the domain), from the block's neighbour slots nb.
template <int B>
__device__ __forceinline__ void block_cell_nbrs(int s, int l, const int (&nb)[6], const int (&side)[3],
                                                const int (&fl)[3], int (&id)[6]) {
    constexpr int B3 = Fine<B>::B3;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const int a = q >> 1;
        if (side[a] == (q & 1)) id[q] = nb[q] < 0 ? nb[q] : nb[q] * B3 + fl[a];
        else id[q] = s * B3 + inner_nbr<B>(l, q);
    }
}

// Air extrapolation of the pressure (FluidProjection.cu:314-335: the mean of the fluid neighbours), one
// workgroup per active block: the block's neighbour slots are uniform loads, and an air cell loads its
// six neighbours' type and pressure together (the cell itself as a stand-in where a neighbour does not
// exist).  Only air cells are written, only fluid cells' values are used, so there is no race.
template <int B, class PT>
struct k_press_extrap_blk_f {
    static constexpr int kLB = Fine<B>::B3;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, FrameK f, const uint8_t* __restrict__ type,
                                                                     PT* __restrict__ pres) {
        constexpr int B3 = Fine<B>::B3;
        const int s = xcd_swizzle(vbid, vnbid);
        int nb[6];
    #pragma unroll
        for (int q = 0; q < 6; ++q) nb[q] = f.nbr[s * 6 + q];
        const int l = threadIdx.x, c = s * B3 + l;
        if (type[c] != kAir) return;
        int side[3], fl[3], id[6];
        thread_faces<B>(l, side, fl);
        block_cell_nbrs<B>(s, l, nb, side, fl, id);
        uint8_t tn[6];
        PT pv[6];
    #pragma unroll
        for (int q = 0; q < 6; ++q) {
            const int ids = id[q] >= 0 ? id[q] : c;
            tn[q] = type[ids];
            pv[q] = pres[ids];
        }
        PT press = PT(0);
        int cnt = 0;
    #pragma unroll
        for (int q = 0; q < 6; ++q)
            if (id[q] >= 0 && tn[q] == kFluid) { press += pv[q]; cnt++; }
        if (cnt > 0) pres[c] = press / (PT)cnt;
    }
};
template <int B, class PT>
__global__ void __launch_bounds__(Fine<B>::B3) k_press_extrap_blk(FrameK f, const uint8_t* __restrict__ type,
                                                                 PT* __restrict__ pres) {
    k_press_extrap_blk_f<B, PT>::run(blockIdx.x, gridDim.x, f, type, pres);
}

// Pressure gradient subtract (FluidProjection.cu:342-378, same operation order) over the solver's
// blocks (the only ones holding fluid cells) from the workgroup table: a fluid cell loads its
// pressure, velocity and its neighbours' type and pressure in one round.
template <int B, int K, class PT>
struct k_press_grad_blk_f {
    static constexpr int kLB = Fine<B>::B3;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, const int* __restrict__ tab, FrameK f,
                                                                   const uint8_t* __restrict__ type,
                                                                   const PT* __restrict__ pres, float wall_level,
                                                                   float* __restrict__ cx, float* __restrict__ cy,
                                                                   float* __restrict__ cz) {
        constexpr int B3 = Fine<B>::B3;
        constexpr int SH = Fine<B>::SH;
        int s[K], nb[K][6];
        wg_load<K>(vbid, vnbid, tab, s, nb);
        const int l = threadIdx.x;
        int side[3], fl[3];
        thread_faces<B>(l, side, fl);
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) {
            if (s[kb] < 0) continue;
            const int c = s[kb] * B3 + l;
            if (type[c] != kFluid) continue;
            const int t = f.coord[s[kb]];
            const int j = ((t / f.fx) % f.fy) * B + ((l >> SH) & (B - 1));
            const uint8_t wall = ((float)j + 0.5f <= wall_level + 1.f) ? kSolid : kAir;
            int nid[6];
            block_cell_nbrs<B>(s[kb], l, nb[kb], side, fl, nid);
            const PT p = pres[c];
            const float vx = cx[c], vy = cy[c], vz = cz[c];
            uint8_t tl[6];
            PT pl[6];
    #pragma unroll
            for (int q = 0; q < 6; ++q) {
                const int ids = nid[q] >= 0 ? nid[q] : c;
                tl[q] = type[ids];
                pl[q] = pres[ids];
            }
            uint8_t tn[6];
            PT pv[6];
    #pragma unroll
            for (int q = 0; q < 6; ++q) {
                tn[q] = nid[q] == -2 ? wall : (nid[q] < 0 ? (uint8_t)kAir : tl[q]);
                pv[q] = nid[q] >= 0 ? pl[q] : PT(0);
            }
            PT nbv[6];
    #pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int lo = 2 * a, hi = 2 * a + 1;
                nbv[lo] = nid[lo] == -2 ? PT(2) * p - pv[hi] : pv[lo];
                nbv[hi] = nid[hi] == -2 ? PT(2) * p - pv[lo] : pv[hi];
            }
            PT pf[6];
    #pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int lo = 2 * a, hi = 2 * a + 1;
                pf[lo] = tn[lo] == kSolid ? PT(2) * p - nbv[hi] : (tn[lo] == kAir ? -nbv[lo] : nbv[lo]);
                pf[hi] = tn[hi] == kSolid ? PT(2) * p - nbv[lo] : (tn[hi] == kAir ? -nbv[hi] : nbv[hi]);
            }
            cx[c] = (float)((PT)vx - (pf[1] - pf[0]) * PT(0.5));
            cy[c] = (float)((PT)vy - (pf[3] - pf[2]) * PT(0.5));
            cz[c] = (float)((PT)vz - (pf[5] - pf[4]) * PT(0.5));
        }
    }
};
template <int B, int K, class PT>
__global__ void __launch_bounds__(Fine<B>::B3) k_press_grad_blk(const int* __restrict__ tab, FrameK f,
                                                               const uint8_t* __restrict__ type,
                                                               const PT* __restrict__ pres, float wall_level,
                                                               float* __restrict__ cx, float* __restrict__ cy,
                                                               float* __restrict__ cz) {
    k_press_grad_blk_f<B, K, PT>::run(blockIdx.x, gridDim.x, tab, f, type, pres, wall_level, cx, cy, cz);
}

// Ap = A p and partial p.Ap (BLAS::mul, BLAS.cu:12-40).
template <int B, int K>
struct k_spmv_blk_f {
    static constexpr int kLB = Fine<B>::B3;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, const int* __restrict__ tab, CgVecs v,
                                                             double* __restrict__ partials,
                                                             const CgState* __restrict__ st) {
        constexpr int B3 = Fine<B>::B3;
        __shared__ float sp[K][B3];
        __shared__ double sh[Fine<B>::NW];
        const int dn = st->done;
        int s[K], nb[K][6];
        wg_load<K>(vbid, vnbid, tab, s, nb);
        if (dn) return;
        const int l = threadIdx.x;
        int side[3], fl[3];
        thread_faces<B>(l, side, fl);
        unsigned in[K];
        float pc[K], hv[K][3];
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) {
            in[kb] = 0;
            pc[kb] = 0.f;
            if (s[kb] >= 0) {
                in[kb] = v.info[s[kb] * B3 + l];
                pc[kb] = v.p[s[kb] * B3 + l];
            }
        }
    #pragma unroll
        for (int kb = 0; kb < K; ++kb)
    #pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int ns = face_slot(nb[kb], a, side[a]);
                hv[kb][a] = ns >= 0 ? v.p[ns * B3 + fl[a]] : 0.f;
            }
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) sp[kb][l] = pc[kb];
        __syncthreads();
        double pap = 0.0;
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) {
            if (s[kb] < 0) continue;
            float acc = 0.f;
            if (in[kb] & kInfoFluid) {
                acc = apply_row<B>(in[kb], pc[kb], sp[kb], l, side, hv[kb]);
                pap += (double)pc[kb] * (double)acc;
            }
            v.ap[s[kb] * B3 + l] = acc;
        }
        const double tot = fine_block_sum<B>(pap, sh);
        if (l == 0) partials[vbid] = tot;     // any fixed slot per workgroup keeps the sum order fixed
    }
};
template <int B, int K>
__global__ void __launch_bounds__(Fine<B>::B3) k_spmv_blk(const int* __restrict__ tab, CgVecs v,
                                                         double* __restrict__ partials,
                                                         const CgState* __restrict__ st) {
    k_spmv_blk_f<B, K>::run(blockIdx.x, gridDim.x, tab, v, partials, st);
}

// Search direction of one cell (BLAS::addmul, LinearSystemSolver.cpp:41): p = z + beta p, with z the
// preconditioned residual (PC 0: r, no preconditioner; 1: r / diag, Jacobi; 2: the V-cycle output,
// multigrid).
template <int PC>
__device__ __forceinline__ float pdir(float beta, unsigned in, float r, float z, float p) {
    const float zz = PC == 1 ? r * jacobi_inv((uint16_t)in) : (PC == 2 ? z : r);
    return zz + (beta * p);
}

// p = z + beta p (the direction update) fused with Ap = A p and partial p.Ap (k_spmv_blk) for every CG
// iteration after the first of a solve.  The new p across a face is recomputed from the neighbour's
// z (or r) and old p, exactly as the neighbour's own workgroup computes it, so no pass over p is
// needed between the two.  The old p is read from v.p and the new one written to pout, a second
// buffer: other workgroups still read the old values across their faces.
// PC is the preconditioner (the solve's st->precond), a template parameter so that each instance
// carries only its own form of z.
template <int B, int K, int PC>
struct k_pupdate_spmv_f {
    static constexpr int kLB = Fine<B>::B3;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, const int* __restrict__ tab, CgVecs v,
                                                                 float* __restrict__ pout,
                                                                 double* __restrict__ partials,
                                                                 const CgState* __restrict__ st) {
        constexpr int B3 = Fine<B>::B3;
        __shared__ float sp[K][B3];
        __shared__ double sh[Fine<B>::NW];
        const int dn = st->done;
        int s[K], nb[K][6];
        wg_load<K>(vbid, vnbid, tab, s, nb);
        if (dn) return;
        const float beta = st->beta;
        const float* __restrict__ zs = PC == 2 ? v.z : v.r;
        const int l = threadIdx.x;
        int side[3], fl[3];
        thread_faces<B>(l, side, fl);
        unsigned in[K];
        float po[K], zo[K], hp[K][3], hz[K][3];
        unsigned hi[K][3];
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) {
            in[kb] = 0;
            po[kb] = 0.f;
            zo[kb] = 0.f;
            if (s[kb] >= 0) {
                in[kb] = v.info[s[kb] * B3 + l];
                po[kb] = v.p[s[kb] * B3 + l];
                zo[kb] = zs[s[kb] * B3 + l];
            }
        }
    #pragma unroll
        for (int kb = 0; kb < K; ++kb)
    #pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int ns = face_slot(nb[kb], a, side[a]);
                hp[kb][a] = 0.f;
                hz[kb][a] = 0.f;
                hi[kb][a] = 0;
                if (ns >= 0) {
                    hp[kb][a] = v.p[ns * B3 + fl[a]];
                    hz[kb][a] = zs[ns * B3 + fl[a]];
                    if (PC == 1) hi[kb][a] = v.info[ns * B3 + fl[a]];
                }
            }
        float pn[K];
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) {
            pn[kb] = pdir<PC>(beta, in[kb], zo[kb], zo[kb], po[kb]);
            sp[kb][l] = pn[kb];
        }
        __syncthreads();
        double pap = 0.0;
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) {
            if (s[kb] < 0) continue;
            float acc = 0.f;
            if (in[kb] & kInfoFluid) {
                float h[3];
    #pragma unroll
                for (int a = 0; a < 3; ++a) h[a] = pdir<PC>(beta, hi[kb][a], hz[kb][a], hz[kb][a], hp[kb][a]);
                acc = apply_row<B>(in[kb], pn[kb], sp[kb], l, side, h);
                pap += (double)pn[kb] * (double)acc;
            }
            pout[s[kb] * B3 + l] = pn[kb];
            v.ap[s[kb] * B3 + l] = acc;
        }
        const double tot = fine_block_sum<B>(pap, sh);
        if (l == 0) partials[vbid] = tot;
    }
};
template <int B, int K, int PC>
__global__ void __launch_bounds__(Fine<B>::B3) k_pupdate_spmv(const int* __restrict__ tab, CgVecs v,
                                                             float* __restrict__ pout,
                                                             double* __restrict__ partials,
                                                             const CgState* __restrict__ st) {
    k_pupdate_spmv_f<B, K, PC>::run(blockIdx.x, gridDim.x, tab, v, pout, partials, st);
}

// ---- multigrid, level 0 (the CG residual r, stencil descriptor info) and level 1 for B = 8 ----
// First Jacobi sweep from zero at a cell: (omega / D) r on fluid cells, 0 elsewhere.
__device__ __forceinline__ float mg_first(unsigned in, float r, float omega) {
    return mg_fluid(in) ? r * mg_wd(in, omega) : 0.f;
}

// Two Jacobi sweeps from zero: zout = z1 + w (r - A z1)/D, z1 = w D^-1 r.
template <int B, int K>
struct k_mg0_smooth_first2_f {
    static constexpr int kLB = Fine<B>::B3;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, const int* __restrict__ tab, MgLevelK L,
                                                                      float omega, float omega2, float* __restrict__ zout,
                                                                      const int* done) {
        constexpr int B3 = Fine<B>::B3;
        __shared__ float sz[K][B3];
        const int dn = *done;
        int s[K], nb[K][6];
        wg_load<K>(vbid, vnbid, tab, s, nb);
        if (dn) return;
        const int l = threadIdx.x;
        int side[3], fl[3];
        thread_faces<B>(l, side, fl);
        unsigned in[K], hi[K][3];
        float rc[K], hr[K][3];
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) {
            in[kb] = 0;
            rc[kb] = 0.f;
            if (s[kb] >= 0) {
                in[kb] = L.info[s[kb] * B3 + l];
                rc[kb] = L.r[s[kb] * B3 + l];
            }
        }
    #pragma unroll
        for (int kb = 0; kb < K; ++kb)
    #pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int ns = face_slot(nb[kb], a, side[a]);
                hi[kb][a] = 0;
                hr[kb][a] = 0.f;
                if (ns >= 0) {
                    hi[kb][a] = L.info[ns * B3 + fl[a]];
                    hr[kb][a] = L.r[ns * B3 + fl[a]];
                }
            }
        float z1[K];
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) {
            z1[kb] = mg_first(in[kb], rc[kb], omega);
            sz[kb][l] = z1[kb];
        }
        __syncthreads();
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) {
            if (s[kb] < 0) continue;
            float out = 0.f;
            if (mg_fluid(in[kb])) {
                float hz[3];
    #pragma unroll
                for (int a = 0; a < 3; ++a) hz[a] = mg_first(hi[kb][a], hr[kb][a], omega);
                const float az = apply_row<B>(in[kb], z1[kb], sz[kb], l, side, hz);
                out = z1[kb] + (rc[kb] - az) * mg_wd(in[kb], omega2);
            }
            zout[s[kb] * B3 + l] = out;
        }
    }
};
template <int B, int K>
__global__ void __launch_bounds__(Fine<B>::B3) k_mg0_smooth_first2(const int* __restrict__ tab, MgLevelK L,
                                                                  float omega, float omega2, float* __restrict__ zout,
                                                                  const int* done) {
    k_mg0_smooth_first2_f<B, K>::run(blockIdx.x, gridDim.x, tab, L, omega, omega2, zout, done);
}

// r -= alpha Ap and partial r.r (k_cg_update, LinearSystemSolver.cpp:34-35) fused with the V-cycle's
// first level-0 pass (two Jacobi sweeps from zero on the new r, k_mg0_smooth_first2), for multigrid
// with two sweeps.  The new r across a face is recomputed from the neighbour's r and Ap, exactly as
// the neighbour's own workgroup computes it.  The old r is read from v.r and the new one written to
// rout, a second buffer, because other workgroups still read the old values across their faces.
// x += alpha p (LinearSystemSolver.cpp:33) is not done here: nothing reads x before the solve ends, so
// the extra workgroups of the same iteration's k_mg_bottom apply it while the bottom workgroup runs on
// an otherwise idle GPU, and the last iteration's term, whose V-cycle does not run once the solver is
// done, is applied after the solve (k_ir_accumulate, k_cg_xflush); the additions happen in the same
// order, so x is bit-identical.
template <int B, int K>
struct k_update_smooth2_f {
    static constexpr int kLB = Fine<B>::B3;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, const int* __restrict__ tab, CgVecs v,
                                                                   float* __restrict__ rout, float omega, float omega2,
                                                                   float* __restrict__ zout,
                                                                   double* __restrict__ partials,
                                                                   const CgState* __restrict__ st) {
        constexpr int B3 = Fine<B>::B3;
        __shared__ float sz[K][B3];
        __shared__ double sh[Fine<B>::NW];
        const int dn = st->done;
        int s[K], nb[K][6];
        wg_load<K>(vbid, vnbid, tab, s, nb);
        if (dn) return;
        const float alpha = st->alpha;
        const int l = threadIdx.x;
        int side[3], fl[3];
        thread_faces<B>(l, side, fl);
        unsigned in[K], hi[K][3];
        float ro[K], ao[K], hr[K][3], ha[K][3];
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) {
            in[kb] = 0;
            ro[kb] = ao[kb] = 0.f;
            if (s[kb] >= 0) {
                const int c = s[kb] * B3 + l;
                in[kb] = v.info[c];
                ro[kb] = v.r[c];
                ao[kb] = v.ap[c];
            }
        }
    #pragma unroll
        for (int kb = 0; kb < K; ++kb)
    #pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int ns = face_slot(nb[kb], a, side[a]);
                hi[kb][a] = 0;
                hr[kb][a] = 0.f;
                ha[kb][a] = 0.f;
                if (ns >= 0) {
                    hi[kb][a] = v.info[ns * B3 + fl[a]];
                    hr[kb][a] = v.r[ns * B3 + fl[a]];
                    ha[kb][a] = v.ap[ns * B3 + fl[a]];
                }
            }
        float rn[K], z1[K];
        double rr = 0.0;
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) {
            rn[kb] = ro[kb] - (alpha * ao[kb]);
            z1[kb] = mg_first(in[kb], rn[kb], omega);
            sz[kb][l] = z1[kb];
            if (s[kb] >= 0) rr += (double)rn[kb] * (double)rn[kb];
        }
        __syncthreads();
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) {
            if (s[kb] < 0) continue;
            float out = 0.f;
            if (mg_fluid(in[kb])) {
                float hz[3];
    #pragma unroll
                for (int a = 0; a < 3; ++a) hz[a] = mg_first(hi[kb][a], hr[kb][a] - (alpha * ha[kb][a]), omega);
                const float az = apply_row<B>(in[kb], z1[kb], sz[kb], l, side, hz);
                out = z1[kb] + (rn[kb] - az) * mg_wd(in[kb], omega2);
            }
            const int c = s[kb] * B3 + l;
            rout[c] = rn[kb];
            zout[c] = out;
        }
        const double tot = fine_block_sum<B>(rr, sh);
        if (l == 0) partials[vbid] = tot;
    }
};
template <int B, int K>
__global__ void __launch_bounds__(Fine<B>::B3) k_update_smooth2(const int* __restrict__ tab, CgVecs v,
                                                               float* __restrict__ rout, float omega, float omega2,
                                                               float* __restrict__ zout,
                                                               double* __restrict__ partials,
                                                               const CgState* __restrict__ st) {
    k_update_smooth2_f<B, K>::run(blockIdx.x, gridDim.x, tab, v, rout, omega, omega2, zout, partials, st);
}

// Residual and restriction to the next level inside the block (the 2^3 children of a coarse cell
// lie in the same block): rC = 0.5 * sum over fluid children of (r - A z), children summed in the
// fixed order of mg_child.  With first != 0, z is the first sweep from zero (w D^-1 r) and is
// stored to zf.
template <int B, int K>
struct k_mg0_restrict_f {
    static constexpr int kLB = Fine<B>::B3;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, const int* __restrict__ tab, MgLevelK L, MgLevelK C,
                                                                 float omega, int first, float* __restrict__ zf,
                                                                 const int* done) {
        constexpr int B3 = Fine<B>::B3;
        constexpr int SH = Fine<B>::SH;
        constexpr int BC = B / 2, CSH = SH - 1, C3 = BC * BC * BC;
        __shared__ float sz[K][B3];
        __shared__ float sres[K][B3];
        const int dn = *done;
        int s[K], nb[K][6];
        wg_load<K>(vbid, vnbid, tab, s, nb);
        if (dn) return;
        const int l = threadIdx.x;
        int side[3], fl[3];
        thread_faces<B>(l, side, fl);
        unsigned in[K];
        float rc[K], zc[K], hz[K][3];
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) {
            in[kb] = 0;
            rc[kb] = 0.f;
            zc[kb] = 0.f;
            if (s[kb] >= 0) {
                const int c = s[kb] * B3 + l;
                in[kb] = L.info[c];
                rc[kb] = L.r[c];
                if (!first) zc[kb] = zf[c];
            }
        }
        if (first) {
            unsigned hi[K][3];
            float hr[K][3];
    #pragma unroll
            for (int kb = 0; kb < K; ++kb)
    #pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const int ns = face_slot(nb[kb], a, side[a]);
                    hi[kb][a] = 0;
                    hr[kb][a] = 0.f;
                    if (ns >= 0) {
                        hi[kb][a] = L.info[ns * B3 + fl[a]];
                        hr[kb][a] = L.r[ns * B3 + fl[a]];
                    }
                }
    #pragma unroll
            for (int kb = 0; kb < K; ++kb) {
                zc[kb] = mg_first(in[kb], rc[kb], omega);
                if (s[kb] >= 0 && mg_fluid(in[kb])) zf[s[kb] * B3 + l] = zc[kb];
    #pragma unroll
                for (int a = 0; a < 3; ++a) hz[kb][a] = mg_first(hi[kb][a], hr[kb][a], omega);
            }
        } else {
    #pragma unroll
            for (int kb = 0; kb < K; ++kb)
    #pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const int ns = face_slot(nb[kb], a, side[a]);
                    hz[kb][a] = ns >= 0 ? zf[ns * B3 + fl[a]] : 0.f;
                }
        }
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) sz[kb][l] = zc[kb];
        __syncthreads();
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) {
            float res = 0.f;
            if (mg_fluid(in[kb])) res = rc[kb] - apply_row<B>(in[kb], zc[kb], sz[kb], l, side, hz[kb]);
            sres[kb][l] = res;
        }
        __syncthreads();
        if (l < C3) {
            const int li = l & (BC - 1), lj = (l >> CSH) & (BC - 1), lk = l >> (2 * CSH);
    #pragma unroll
            for (int kb = 0; kb < K; ++kb) {
                if (s[kb] < 0) continue;
                const int cc = s[kb] * C3 + l;
                float acc = 0.f;
                if (mg_fluid(C.info[cc])) {
    #pragma unroll
                    for (int a = 0; a < 8; ++a) {
                        const int fl8 = (((2 * lk + (a >> 2)) << SH) + (2 * lj + ((a >> 1) & 1))) * B + (2 * li + (a & 1));
                        acc += sres[kb][fl8];
                    }
                    acc *= 0.5f;
                }
                C.r[cc] = acc;
            }
        }
    }
};
template <int B, int K>
__global__ void __launch_bounds__(Fine<B>::B3) k_mg0_restrict(const int* __restrict__ tab, MgLevelK L, MgLevelK C,
                                                             float omega, int first, float* __restrict__ zf,
                                                             const int* done) {
    k_mg0_restrict_f<B, K>::run(blockIdx.x, gridDim.x, tab, L, C, omega, first, zf, done);
}

// Local index of the parent (edge B/2 level) of local cell ll of an edge-B block.
template <int B>
__device__ __forceinline__ int parent_local(int ll) {
    constexpr int SH = Fine<B>::SH, CSH = SH - 1;
    const int li = (ll & (B - 1)) >> 1, lj = ((ll >> SH) & (B - 1)) >> 1, lk = (ll >> (2 * SH)) >> 1;
    return (((lk << CSH) + lj) << CSH) + li;
}

// Prolongation (piecewise constant from the next level) fused with one Jacobi sweep.
template <int B, int K>
struct k_mg0_prolong_smooth_f {
    static constexpr int kLB = Fine<B>::B3;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, const int* __restrict__ tab, MgLevelK L,
                                                                       float omega, const float* __restrict__ z,
                                                                       const float* __restrict__ zc,
                                                                       float* __restrict__ zout, const int* done,
                                                                       double* __restrict__ rz_partials) {
        constexpr int B3 = Fine<B>::B3;
        constexpr int BC = B / 2, C3 = BC * BC * BC;
        __shared__ float szp[K][B3];
        __shared__ double sh[Fine<B>::NW];
        const int dn = *done;
        int s[K], nb[K][6];
        wg_load<K>(vbid, vnbid, tab, s, nb);
        if (dn) return;
        const int l = threadIdx.x;
        int side[3], fl[3];
        thread_faces<B>(l, side, fl);
        unsigned in[K];
        float zo[K], zq[K], rc[K], hz[K][3], hq[K][3];
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) {
            in[kb] = 0;
            zo[kb] = 0.f;
            zq[kb] = 0.f;
            rc[kb] = 0.f;
            if (s[kb] >= 0) {
                const int c = s[kb] * B3 + l;
                in[kb] = L.info[c];
                zo[kb] = z[c];
                zq[kb] = zc[s[kb] * C3 + parent_local<B>(l)];
                rc[kb] = L.r[c];
            }
        }
    #pragma unroll
        for (int kb = 0; kb < K; ++kb)
    #pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int ns = face_slot(nb[kb], a, side[a]);
                hz[kb][a] = 0.f;
                hq[kb][a] = 0.f;
                if (ns >= 0) {
                    hz[kb][a] = z[ns * B3 + fl[a]];
                    hq[kb][a] = zc[ns * C3 + parent_local<B>(fl[a])];
                }
            }
        float zp[K];
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) {
            zp[kb] = zo[kb] + zq[kb];
            szp[kb][l] = zp[kb];
        }
        __syncthreads();
        double rz = 0.0;             // r.z of the output (the PCG's r.z when this is the last level-0 pass)
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) {
            if (s[kb] < 0) continue;
            float out = 0.f;
            if (mg_fluid(in[kb])) {
                float h[3];
    #pragma unroll
                for (int a = 0; a < 3; ++a) h[a] = hz[kb][a] + hq[kb][a];
                const float az = apply_row<B>(in[kb], zp[kb], szp[kb], l, side, h);
                out = zp[kb] + (rc[kb] - az) * mg_wd(in[kb], omega);
            }
            zout[s[kb] * B3 + l] = out;
            rz += (double)rc[kb] * (double)out;
        }
        if (rz_partials) {
            const double tot = fine_block_sum<B>(rz, sh);
            if (l == 0) rz_partials[vbid] = tot;
        }
    }
};
template <int B, int K>
__global__ void __launch_bounds__(Fine<B>::B3) k_mg0_prolong_smooth(const int* __restrict__ tab, MgLevelK L,
                                                                   float omega, const float* __restrict__ z,
                                                                   const float* __restrict__ zc,
                                                                   float* __restrict__ zout, const int* done,
                                                                   double* __restrict__ rz_partials) {
    k_mg0_prolong_smooth_f<B, K>::run(blockIdx.x, gridDim.x, tab, L, omega, z, zc, zout, done, rz_partials);
}

// One Jacobi sweep: zout = z + w (r - A z)/D.
template <int B, int K>
struct k_mg0_smooth_f {
    static constexpr int kLB = Fine<B>::B3;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, const int* __restrict__ tab, MgLevelK L, float omega,
                                                               const float* __restrict__ z, float* __restrict__ zout,
                                                               const int* done, double* __restrict__ rz_partials) {
        constexpr int B3 = Fine<B>::B3;
        __shared__ float sz[K][B3];
        __shared__ double sh[Fine<B>::NW];
        const int dn = *done;
        int s[K], nb[K][6];
        wg_load<K>(vbid, vnbid, tab, s, nb);
        if (dn) return;
        const int l = threadIdx.x;
        int side[3], fl[3];
        thread_faces<B>(l, side, fl);
        unsigned in[K];
        float zc[K], rc[K], hz[K][3];
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) {
            in[kb] = 0;
            zc[kb] = 0.f;
            rc[kb] = 0.f;
            if (s[kb] >= 0) {
                const int c = s[kb] * B3 + l;
                in[kb] = L.info[c];
                zc[kb] = z[c];
                rc[kb] = L.r[c];
            }
        }
    #pragma unroll
        for (int kb = 0; kb < K; ++kb)
    #pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int ns = face_slot(nb[kb], a, side[a]);
                hz[kb][a] = ns >= 0 ? z[ns * B3 + fl[a]] : 0.f;
            }
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) sz[kb][l] = zc[kb];
        __syncthreads();
        double rz = 0.0;             // r.z of the output (the PCG's r.z when this is the last level-0 pass)
    #pragma unroll
        for (int kb = 0; kb < K; ++kb) {
            if (s[kb] < 0) continue;
            float out = 0.f;
            if (mg_fluid(in[kb])) {
                const float az = apply_row<B>(in[kb], zc[kb], sz[kb], l, side, hz[kb]);
                out = zc[kb] + (rc[kb] - az) * mg_wd(in[kb], omega);
            }
            zout[s[kb] * B3 + l] = out;
            rz += (double)rc[kb] * (double)out;
        }
        if (rz_partials) {
            const double tot = fine_block_sum<B>(rz, sh);
            if (l == 0) rz_partials[vbid] = tot;
        }
    }
};
template <int B, int K>
__global__ void __launch_bounds__(Fine<B>::B3) k_mg0_smooth(const int* __restrict__ tab, MgLevelK L, float omega,
                                                           const float* __restrict__ z, float* __restrict__ zout,
                                                           const int* done, double* __restrict__ rz_partials) {
    k_mg0_smooth_f<B, K>::run(blockIdx.x, gridDim.x, tab, L, omega, z, zout, done, rz_partials);
}

}  // namespace vare
