// flip_stencil.h -- block-aligned 7-point stencil kernels on the fine level (edge B).
//
// Each workgroup takes K consecutive entries of the solver's block list (`blist`: the blocks
// holding a fluid cell; blocks without fluid hold no unknown and are skipped) with one thread per
// cell position (B^3 threads: 512 for B=8, 64 for B=4), so every thread serves K cells.  A pass is
// two dependent memory round trips (own values, then the face neighbours in adjacent blocks), so
// the K blocks' loads are issued together in each round: the latency is paid once per K blocks.
// Own values are staged in LDS; neighbours inside a block are read from LDS, those across a face
// from global memory through the block's neighbour slots.  Reductions are one FP64 partial per
// workgroup, summed in a fixed order by the finalize kernels.
#pragma once
#include "flip_kernels.h"
#include "flip_mg.h"

namespace vare {

template <int B> struct Fine {
    static constexpr int B3 = B * B * B;
    static constexpr int SH = (B == 4) ? 2 : 3;
    static constexpr int NW = B3 / 64;         // waves per workgroup
    static constexpr int K = 4;                // blocks per workgroup
};

// Slots of this workgroup's K blocks (-1 past the end of the list) and, in LDS, their 6 neighbour
// slots each.  The caller synchronises before reading fnb.
template <int B>
__device__ __forceinline__ void wg_blocks(const int* __restrict__ blist, int nsb, const int* __restrict__ nbr,
                                          int (&s)[Fine<B>::K], int* fnb) {
    constexpr int K = Fine<B>::K;
    const int w = xcd_swizzle(blockIdx.x, gridDim.x);
#pragma unroll
    for (int kb = 0; kb < K; ++kb) {
        const int v = w * K + kb;
        s[kb] = v < nsb ? blist[v] : -1;
    }
    if (threadIdx.x < 6 * K) {
        const int v = w * K + threadIdx.x / 6;
        fnb[threadIdx.x] = v < nsb ? nbr[blist[v] * 6 + threadIdx.x % 6] : -1;
    }
}

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
__global__ void __launch_bounds__(Fine<B>::B3) k_spmv_blk(const int* __restrict__ blist, int nsb, FrameK f, CgVecs v,
                                                         double* __restrict__ partials,
                                                         const CgState* __restrict__ st) {
    constexpr int B3 = Fine<B>::B3, K = Fine<B>::K;
    __shared__ float sp[K][B3];
    __shared__ int fnb[6 * K];
    __shared__ double sh[Fine<B>::NW];
    if (st->done) return;
    const int l = threadIdx.x;
    int s[K];
    wg_blocks<B>(blist, nsb, f.nbr, s, fnb);
    uint16_t in[K];
    float pc[K];
#pragma unroll
    for (int kb = 0; kb < K; ++kb) {
        in[kb] = 0;
        pc[kb] = 0.f;
        if (s[kb] >= 0) {
            in[kb] = v.info[s[kb] * B3 + l];
            pc[kb] = v.p[s[kb] * B3 + l];
        }
        sp[kb][l] = pc[kb];
    }
    __syncthreads();
    float hv[K][6];
#pragma unroll
    for (int kb = 0; kb < K; ++kb)
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            int nl;
            hv[kb][q] = 0.f;
            if ((in[kb] & (1u << q)) && !fine_nbr<B>(l, q, nl)) hv[kb][q] = v.p[fnb[kb * 6 + q] * B3 + nl];
        }
    double pap = 0.0;
#pragma unroll
    for (int kb = 0; kb < K; ++kb) {
        if (s[kb] < 0) continue;
        float acc = 0.f;
        if (in[kb] & kInfoFluid) {
            acc = (float)((in[kb] >> 8) & 0x1f) * pc[kb];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                if (!(in[kb] & (1u << q))) continue;
                int nl;
                acc = acc - (fine_nbr<B>(l, q, nl) ? sp[kb][nl] : hv[kb][q]);
            }
            pap += (double)pc[kb] * (double)acc;
        }
        v.ap[s[kb] * B3 + l] = acc;
    }
    const double tot = fine_block_sum<B>(pap, sh);
    if (l == 0) partials[blockIdx.x] = tot;     // any fixed slot per workgroup keeps the sum order fixed
}

// ---- multigrid, level 0 (the CG residual r, stencil descriptor info) ----
// Two Jacobi sweeps from zero: zout = z1 + w (r - A z1)/D, z1 = w D^-1 r.
template <int B>
__global__ void __launch_bounds__(Fine<B>::B3) k_mg0_smooth_first2(const int* __restrict__ blist, int nsb, MgLevelK L,
                                                                  MgFrame F, float omega, float* __restrict__ zout,
                                                                  const int* done) {
    constexpr int B3 = Fine<B>::B3, K = Fine<B>::K;
    __shared__ float sz[K][B3];
    __shared__ int fnb[6 * K];
    if (*done) return;
    const int l = threadIdx.x;
    int s[K];
    wg_blocks<B>(blist, nsb, F.nbr, s, fnb);
    uint16_t in[K];
    float rc[K], z1[K];
#pragma unroll
    for (int kb = 0; kb < K; ++kb) {
        in[kb] = 0;
        rc[kb] = 0.f;
        if (s[kb] >= 0) {
            in[kb] = L.info[s[kb] * B3 + l];
            rc[kb] = L.r[s[kb] * B3 + l];
        }
        z1[kb] = mg_fluid(in[kb]) ? omega * rc[kb] / mg_diag(in[kb]) : 0.f;
        sz[kb][l] = z1[kb];
    }
    __syncthreads();
    uint16_t hi[K][6];
    float hr[K][6];
#pragma unroll
    for (int kb = 0; kb < K; ++kb)
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            int nl;
            hi[kb][q] = 0;
            hr[kb][q] = 0.f;
            if ((in[kb] & (1u << q)) && !fine_nbr<B>(l, q, nl)) {
                const int n = fnb[kb * 6 + q] * B3 + nl;
                hi[kb][q] = L.info[n];
                hr[kb][q] = L.r[n];
            }
        }
#pragma unroll
    for (int kb = 0; kb < K; ++kb) {
        if (s[kb] < 0) continue;
        float out = 0.f;
        if (mg_fluid(in[kb])) {
            const float d = mg_diag(in[kb]);
            float az = d * z1[kb];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                if (!(in[kb] & (1u << q))) continue;
                int nl;
                az -= fine_nbr<B>(l, q, nl) ? sz[kb][nl]
                                            : (mg_fluid(hi[kb][q]) ? omega * hr[kb][q] / mg_diag(hi[kb][q]) : 0.f);
            }
            out = z1[kb] + omega * (rc[kb] - az) / d;
        }
        zout[s[kb] * B3 + l] = out;
    }
}

// Residual of level 0 and restriction to level 1 inside the block (the 2^3 children of a level-1
// cell lie in the same block): rC = 0.5 * sum over fluid children of (r - A z), children summed in
// the fixed order of mg_child.  With first != 0, z is the first sweep from zero (w D^-1 r) and is
// stored to zf.
template <int B>
__global__ void __launch_bounds__(Fine<B>::B3) k_mg0_restrict(const int* __restrict__ blist, int nsb, MgLevelK L,
                                                             MgLevelK C, MgFrame F, float omega, int first,
                                                             float* __restrict__ zf, const int* done) {
    constexpr int B3 = Fine<B>::B3, K = Fine<B>::K;
    constexpr int SH = Fine<B>::SH;
    constexpr int BC = B / 2, CSH = SH - 1, C3 = BC * BC * BC;
    __shared__ float sz[K][B3];
    __shared__ float sres[K][B3];
    __shared__ int fnb[6 * K];
    if (*done) return;
    const int l = threadIdx.x;
    int s[K];
    wg_blocks<B>(blist, nsb, F.nbr, s, fnb);
    uint16_t in[K];
    float rc[K], zc[K];
#pragma unroll
    for (int kb = 0; kb < K; ++kb) {
        in[kb] = 0;
        rc[kb] = 0.f;
        zc[kb] = 0.f;
        if (s[kb] >= 0) {
            const int c = s[kb] * B3 + l;
            in[kb] = L.info[c];
            rc[kb] = L.r[c];
            if (first) {
                zc[kb] = mg_fluid(in[kb]) ? omega * rc[kb] / mg_diag(in[kb]) : 0.f;
                if (mg_fluid(in[kb])) zf[c] = zc[kb];
            } else {
                zc[kb] = zf[c];
            }
        }
        sz[kb][l] = zc[kb];
    }
    __syncthreads();
    float hz[K][6];
#pragma unroll
    for (int kb = 0; kb < K; ++kb)
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            int nl;
            hz[kb][q] = 0.f;
            if ((in[kb] & (1u << q)) && !fine_nbr<B>(l, q, nl)) {
                const int n = fnb[kb * 6 + q] * B3 + nl;
                if (first) {
                    const uint16_t ni = L.info[n];
                    hz[kb][q] = mg_fluid(ni) ? omega * L.r[n] / mg_diag(ni) : 0.f;
                } else {
                    hz[kb][q] = zf[n];
                }
            }
        }
#pragma unroll
    for (int kb = 0; kb < K; ++kb) {
        float res = 0.f;
        if (mg_fluid(in[kb])) {
            const float d = mg_diag(in[kb]);
            float az = d * zc[kb];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                if (!(in[kb] & (1u << q))) continue;
                int nl;
                az -= fine_nbr<B>(l, q, nl) ? sz[kb][nl] : hz[kb][q];
            }
            res = rc[kb] - az;
        }
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
                    const int fl = (((2 * lk + (a >> 2)) << SH) + (2 * lj + ((a >> 1) & 1))) * B + (2 * li + (a & 1));
                    acc += sres[kb][fl];
                }
                acc *= 0.5f;
            }
            C.r[cc] = acc;
        }
    }
}

// Prolongation (piecewise constant from level 1) fused with one Jacobi sweep on level 0.
template <int B>
__global__ void __launch_bounds__(Fine<B>::B3) k_mg0_prolong_smooth(const int* __restrict__ blist, int nsb,
                                                                   MgLevelK L, MgLevelK C, MgFrame F, float omega,
                                                                   const float* __restrict__ z,
                                                                   const float* __restrict__ zc,
                                                                   float* __restrict__ zout, const int* done,
                                                                   double* __restrict__ rz_partials) {
    constexpr int B3 = Fine<B>::B3, K = Fine<B>::K;
    constexpr int SH = Fine<B>::SH;
    constexpr int BC = B / 2, CSH = SH - 1, C3 = BC * BC * BC;
    __shared__ float szp[K][B3];
    __shared__ int fnb[6 * K];
    __shared__ double sh[Fine<B>::NW];
    if (*done) return;
    const int l = threadIdx.x;
    int s[K];
    wg_blocks<B>(blist, nsb, F.nbr, s, fnb);
    auto parent_local = [](int ll) {
        const int li = (ll & (B - 1)) >> 1, lj = ((ll >> SH) & (B - 1)) >> 1, lk = (ll >> (2 * SH)) >> 1;
        return (((lk << CSH) + lj) << CSH) + li;
    };
    uint16_t in[K];
    float zp[K], rc[K];
    double rz = 0.0;             // r.z of the output (the PCG's r.z when this is the last level-0 pass)
#pragma unroll
    for (int kb = 0; kb < K; ++kb) {
        in[kb] = 0;
        zp[kb] = 0.f;
        rc[kb] = 0.f;
        if (s[kb] >= 0) {
            const int c = s[kb] * B3 + l;
            in[kb] = L.info[c];
            zp[kb] = z[c] + zc[s[kb] * C3 + parent_local(l)];
            rc[kb] = L.r[c];
        }
        szp[kb][l] = zp[kb];
    }
    __syncthreads();
    float hz[K][6];
#pragma unroll
    for (int kb = 0; kb < K; ++kb)
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            int nl;
            hz[kb][q] = 0.f;
            if ((in[kb] & (1u << q)) && !fine_nbr<B>(l, q, nl)) {
                const int ns = fnb[kb * 6 + q];
                hz[kb][q] = z[ns * B3 + nl] + zc[ns * C3 + parent_local(nl)];
            }
        }
#pragma unroll
    for (int kb = 0; kb < K; ++kb) {
        if (s[kb] < 0) continue;
        float out = 0.f;
        if (mg_fluid(in[kb])) {
            const float d = mg_diag(in[kb]);
            float az = d * zp[kb];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                if (!(in[kb] & (1u << q))) continue;
                int nl;
                az -= fine_nbr<B>(l, q, nl) ? szp[kb][nl] : hz[kb][q];
            }
            out = zp[kb] + omega * (rc[kb] - az) / d;
        }
        zout[s[kb] * B3 + l] = out;
        rz += (double)rc[kb] * (double)out;
    }
    if (rz_partials) {
        const double tot = fine_block_sum<B>(rz, sh);
        if (l == 0) rz_partials[blockIdx.x] = tot;
    }
}

// One Jacobi sweep on level 0: zout = z + w (r - A z)/D.
template <int B>
__global__ void __launch_bounds__(Fine<B>::B3) k_mg0_smooth(const int* __restrict__ blist, int nsb, MgLevelK L,
                                                           MgFrame F, float omega, const float* __restrict__ z,
                                                           float* __restrict__ zout, const int* done,
                                                           double* __restrict__ rz_partials) {
    constexpr int B3 = Fine<B>::B3, K = Fine<B>::K;
    __shared__ float sz[K][B3];
    __shared__ int fnb[6 * K];
    __shared__ double sh[Fine<B>::NW];
    if (*done) return;
    const int l = threadIdx.x;
    int s[K];
    wg_blocks<B>(blist, nsb, F.nbr, s, fnb);
    uint16_t in[K];
    float zc[K], rc[K];
    double rz = 0.0;             // r.z of the output (the PCG's r.z when this is the last level-0 pass)
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
        sz[kb][l] = zc[kb];
    }
    __syncthreads();
    float hz[K][6];
#pragma unroll
    for (int kb = 0; kb < K; ++kb)
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            int nl;
            hz[kb][q] = 0.f;
            if ((in[kb] & (1u << q)) && !fine_nbr<B>(l, q, nl)) hz[kb][q] = z[fnb[kb * 6 + q] * B3 + nl];
        }
#pragma unroll
    for (int kb = 0; kb < K; ++kb) {
        if (s[kb] < 0) continue;
        float out = 0.f;
        if (mg_fluid(in[kb])) {
            const float d = mg_diag(in[kb]);
            float az = d * zc[kb];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                if (!(in[kb] & (1u << q))) continue;
                int nl;
                az -= fine_nbr<B>(l, q, nl) ? sz[kb][nl] : hz[kb][q];
            }
            out = zc[kb] + omega * (rc[kb] - az) / d;
        }
        zout[s[kb] * B3 + l] = out;
        rz += (double)rc[kb] * (double)out;
    }
    if (rz_partials) {
        const double tot = fine_block_sum<B>(rz, sh);
        if (l == 0) rz_partials[blockIdx.x] = tot;
    }
}

}  // namespace vare
