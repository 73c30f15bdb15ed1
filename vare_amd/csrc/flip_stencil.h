// flip_stencil.h -- block-aligned 7-point stencil kernels on the fine level (edge B).
//
// One workgroup per listed block (the blocks holding a fluid cell, `blist`), one thread per cell
// (B^3 threads: 512 for B=8, 64 for B=4).  Blocks without fluid hold no unknown and are skipped.
// The block's own values are staged in LDS; neighbours inside the block are read from LDS, the
// face neighbours of the 6 adjacent blocks from global memory through the block's neighbour slots.
// Reductions are one FP64 partial per block, summed in a fixed order by the finalize kernels.
// The arithmetic (term order included) is identical to the generic grid-stride kernels.
#pragma once
#include "flip_kernels.h"
#include "flip_mg.h"

namespace vare {

template <int B> struct Fine {
    static constexpr int B3 = B * B * B;
    static constexpr int SH = (B == 4) ? 2 : 3;
    static constexpr int NW = B3 / 64;         // waves per workgroup
};

// Local coordinates of thread l and whether the neighbour in direction q stays in the block;
// returns the local index inside the block (nl) or, across the face, the local index inside the
// neighbour block.
template <int B>
__device__ __forceinline__ bool fine_nbr(int l, int q, int& nl) {
    constexpr int SH = Fine<B>::SH;
    const int li = l & (B - 1), lj = (l >> SH) & (B - 1), lk = l >> (2 * SH);
    const int axis = q >> 1;
    const int dir = (q & 1) ? 1 : -1;
    const int lc = axis == 0 ? li : (axis == 1 ? lj : lk);
    const int stride = axis == 0 ? 1 : (axis == 1 ? B : B * B);
    const int t = lc + dir;
    if (t >= 0 && t < B) { nl = l + dir * stride; return true; }
    nl = l - dir * (B - 1) * stride;
    return false;
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

// Ap = A p and partial p.Ap (BLAS::mul, BLAS.cu:12-40).
template <int B>
__global__ void __launch_bounds__(Fine<B>::B3) k_spmv_blk(const int* __restrict__ blist, FrameK f, CgVecs v, double* __restrict__ partials,
                                                         const CgState* __restrict__ st) {
    constexpr int B3 = Fine<B>::B3;
    __shared__ float sp[B3];
    __shared__ int fnb[6];
    __shared__ double sh[Fine<B>::NW];
    if (st->done) return;
    const int s = blist[blockIdx.x], l = threadIdx.x, c = s * B3 + l;
    if (l < 6) fnb[l] = f.nbr[s * 6 + l];
    const uint16_t info = v.info[c];
    const float pc = v.p[c];
    sp[l] = pc;
    __syncthreads();
    float acc = 0.f;
    double pap = 0.0;
    if (info & kInfoFluid) {
        acc = (float)((info >> 8) & 0x1f) * pc;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            if (!(info & (1u << q))) continue;
            int nl;
            const float pv = fine_nbr<B>(l, q, nl) ? sp[nl] : v.p[fnb[q] * B3 + nl];
            acc = acc - pv;
        }
        pap = (double)pc * (double)acc;
    }
    v.ap[c] = acc;
    const double tot = fine_block_sum<B>(pap, sh);
    if (l == 0) partials[blockIdx.x] = tot;
}

// ---- multigrid, level 0 (the CG residual r, stencil descriptor info) ----
// Two Jacobi sweeps from zero: zout = z1 + w (r - A z1)/D, z1 = w D^-1 r.
template <int B>
__global__ void __launch_bounds__(Fine<B>::B3) k_mg0_smooth_first2(const int* __restrict__ blist, MgLevelK L, MgFrame F, float omega,
                                                                  float* __restrict__ zout, const int* done) {
    constexpr int B3 = Fine<B>::B3;
    __shared__ float sz[B3];
    __shared__ int fnb[6];
    if (*done) return;
    const int s = blist[blockIdx.x], l = threadIdx.x, c = s * B3 + l;
    if (l < 6) fnb[l] = F.nbr[s * 6 + l];
    const uint16_t in = L.info[c];
    const float rc = L.r[c];
    const float z1 = mg_fluid(in) ? omega * rc / mg_diag(in) : 0.f;
    sz[l] = z1;
    __syncthreads();
    float out = 0.f;
    if (mg_fluid(in)) {
        const float d = mg_diag(in);
        float az = d * z1;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            if (!(in & (1u << q))) continue;
            int nl;
            float zn;
            if (fine_nbr<B>(l, q, nl)) zn = sz[nl];
            else {
                const int n = fnb[q] * B3 + nl;
                const uint16_t ni = L.info[n];
                zn = mg_fluid(ni) ? omega * L.r[n] / mg_diag(ni) : 0.f;
            }
            az -= zn;
        }
        out = z1 + omega * (rc - az) / d;
    }
    zout[c] = out;
}

// Residual of level 0 and restriction to level 1 inside the block (the 2^3 children of a level-1
// cell lie in the same block): rC = 0.5 * sum over fluid children of (r - A z), children summed in
// the fixed order of mg_child.  With first != 0, z is the first sweep from zero (w D^-1 r) and is
// stored to zf.
template <int B>
__global__ void __launch_bounds__(Fine<B>::B3) k_mg0_restrict(const int* __restrict__ blist, MgLevelK L, MgLevelK C, MgFrame F, float omega,
                                                             int first, float* __restrict__ zf, const int* done) {
    constexpr int B3 = Fine<B>::B3;
    constexpr int SH = Fine<B>::SH;
    constexpr int BC = B / 2, CSH = SH - 1, C3 = BC * BC * BC;
    __shared__ float sz[B3];
    __shared__ float sres[B3];
    __shared__ int fnb[6];
    if (*done) return;
    const int s = blist[blockIdx.x], l = threadIdx.x, c = s * B3 + l;
    if (l < 6) fnb[l] = F.nbr[s * 6 + l];
    const uint16_t in = L.info[c];
    const float rc = L.r[c];
    float zc;
    if (first) {
        zc = mg_fluid(in) ? omega * rc / mg_diag(in) : 0.f;
        if (mg_fluid(in)) zf[c] = zc;
    } else {
        zc = zf[c];
    }
    sz[l] = zc;
    __syncthreads();
    float res = 0.f;
    if (mg_fluid(in)) {
        const float d = mg_diag(in);
        float az = d * zc;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            if (!(in & (1u << q))) continue;
            int nl;
            float zn;
            if (fine_nbr<B>(l, q, nl)) zn = sz[nl];
            else {
                const int n = fnb[q] * B3 + nl;
                if (first) {
                    const uint16_t ni = L.info[n];
                    zn = mg_fluid(ni) ? omega * L.r[n] / mg_diag(ni) : 0.f;
                } else {
                    zn = zf[n];
                }
            }
            az -= zn;
        }
        res = rc - az;
    }
    sres[l] = mg_fluid(in) ? res : 0.f;
    __syncthreads();
    if (l < C3) {
        const int cc = s * C3 + l;
        float acc = 0.f;
        if (mg_fluid(C.info[cc])) {
            const int li = l & (BC - 1), lj = (l >> CSH) & (BC - 1), lk = l >> (2 * CSH);
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                const int fl = (((2 * lk + (a >> 2)) << SH) + (2 * lj + ((a >> 1) & 1))) * B + (2 * li + (a & 1));
                acc += sres[fl];
            }
            acc *= 0.5f;
        }
        C.r[cc] = acc;
    }
}

// Prolongation (piecewise constant from level 1) fused with one Jacobi sweep on level 0.
template <int B>
__global__ void __launch_bounds__(Fine<B>::B3) k_mg0_prolong_smooth(const int* __restrict__ blist, MgLevelK L, MgLevelK C, MgFrame F, float omega,
                                                                   const float* __restrict__ z,
                                                                   const float* __restrict__ zc,
                                                                   float* __restrict__ zout, const int* done) {
    constexpr int B3 = Fine<B>::B3;
    constexpr int SH = Fine<B>::SH;
    constexpr int BC = B / 2, CSH = SH - 1, C3 = BC * BC * BC;
    __shared__ float szp[B3];
    __shared__ int fnb[6];
    if (*done) return;
    const int s = blist[blockIdx.x], l = threadIdx.x, c = s * B3 + l;
    if (l < 6) fnb[l] = F.nbr[s * 6 + l];
    auto parent_local = [](int ll) {
        const int li = (ll & (B - 1)) >> 1, lj = ((ll >> SH) & (B - 1)) >> 1, lk = (ll >> (2 * SH)) >> 1;
        return (((lk << CSH) + lj) << CSH) + li;
    };
    const uint16_t in = L.info[c];
    const float zp = z[c] + zc[s * C3 + parent_local(l)];
    szp[l] = zp;
    __syncthreads();
    float out = 0.f;
    if (mg_fluid(in)) {
        const float d = mg_diag(in);
        float az = d * zp;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            if (!(in & (1u << q))) continue;
            int nl;
            float zn;
            if (fine_nbr<B>(l, q, nl)) zn = szp[nl];
            else {
                const int ns = fnb[q];
                zn = z[ns * B3 + nl] + zc[ns * C3 + parent_local(nl)];
            }
            az -= zn;
        }
        out = zp + omega * (L.r[c] - az) / d;
    }
    zout[c] = out;
}

// One Jacobi sweep on level 0: zout = z + w (r - A z)/D.
template <int B>
__global__ void __launch_bounds__(Fine<B>::B3) k_mg0_smooth(const int* __restrict__ blist, MgLevelK L, MgFrame F, float omega,
                                                           const float* __restrict__ z, float* __restrict__ zout,
                                                           const int* done) {
    constexpr int B3 = Fine<B>::B3;
    __shared__ float sz[B3];
    __shared__ int fnb[6];
    if (*done) return;
    const int s = blist[blockIdx.x], l = threadIdx.x, c = s * B3 + l;
    if (l < 6) fnb[l] = F.nbr[s * 6 + l];
    const uint16_t in = L.info[c];
    const float zc = z[c];
    sz[l] = zc;
    __syncthreads();
    float out = 0.f;
    if (mg_fluid(in)) {
        const float d = mg_diag(in);
        float az = d * zc;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            if (!(in & (1u << q))) continue;
            int nl;
            az -= fine_nbr<B>(l, q, nl) ? sz[nl] : z[fnb[q] * B3 + nl];
        }
        out = zc + omega * (L.r[c] - az) / d;
    }
    zout[c] = out;
}

}  // namespace vare
