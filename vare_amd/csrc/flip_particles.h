// flip_particles.h -- cell-binned particle kernels: the binning (a stable counting sort of the
// particles by the storage index of their P2G base cell, ties in the reference's particle order),
// the reorder of the particle arrays into that order, P2G as a gather per block (deterministic: no
// float atomics), and G2P from an LDS tile per block.
//
// Storage: positions, velocities and adjacency live in physical order, which each step's binning
// sets to bin order so that a block's particles are contiguous; pid[q] is the reference (logical)
// index of physical particle q.  Vorticity and curvature stay in logical order.  A particle's P2G
// base cell is c = max((int)(p-0.5), 0) per axis (Rasterization.cu:139-147).  For a particle inside
// [0,n]^3 the block holding c is always active: it is the particle's block or the one below it,
// both inside the 3x3x3 neighbourhood that SparseFrame.cu:145-152 marks.  Particles outside [0,n]^3
// (skipped by P2G) go to an extra bin Nc, handled by G2P with global-memory lookups.
#pragma once
#include "flip_kernels.h"

namespace vare {

constexpr int kSortMax = 8192;      // longest bin sorted in LDS (64 KB of 64-bit keys)
constexpr int kSortReg = 16;        // bins up to this length are sorted in registers
constexpr unsigned long long kSortPad = ~0ull;

// Bin of every particle (storage index of its base cell, or nc) + its arrival rank inside the bin.
// A workgroup covers kKeySlices consecutive slices of WG particles, one particle per thread per slice
// (the slices' loads and slot lookups are independent, so their latencies overlap).  The counts are
// aggregated in LDS first: every key of the workgroup is entered in an open-addressing hash table
// (load factor at most 1/2) whose LDS counter hands out the rank among the workgroup's particles of
// that bin (one entry per run of equal keys in adjacent lanes), then one global atomic per distinct
// key reserves the workgroup's run in the bin.  The
// particles of a workgroup come from a few neighbouring cells, so this takes several times fewer
// global atomics than one per particle or per run of equal keys, and global atomic throughput is what
// bounds this kernel.  Ranks depend on the atomics' order; the bins are sorted by reference index
// afterwards, so the permutation does not.
constexpr int kBinSlices = 4;                    // k_pscatter
constexpr int kKeySlices = 8;                    // k_ckey
constexpr int kKeyHash = 2 * kKeySlices * WG;    // LDS hash slots of k_ckey (32 KB of keys and counts)

template <int B>
struct k_ckey_f {
    static constexpr int kLB = WG;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, const float* __restrict__ px, const float* __restrict__ py,
                                                 const float* __restrict__ pz, int np, FrameK f, int nc,
                                                 int* __restrict__ key, int* __restrict__ rank, int* __restrict__ count) {
        constexpr int B3 = B * B * B;
        constexpr int J = kKeySlices, H = kKeyHash;
        __shared__ int hk[H];                        // key of the slot, -1 when free
        __shared__ int hc[H];                        // count, then the global base of the slot's run
        for (int i = threadIdx.x; i < H; i += WG) { hk[i] = -1; hc[i] = 0; }
        int idx[J], k[J];
        bool valid[J];
    #pragma unroll
        for (int j = 0; j < J; ++j) {
            idx[j] = (vbid * J + j) * WG + threadIdx.x;
            valid[j] = idx[j] < np;
            k[j] = -1;
        }
        float x[J], y[J], z[J];
    #pragma unroll
        for (int j = 0; j < J; ++j) {
            const int i = valid[j] ? idx[j] : 0;
            x[j] = px[i]; y[j] = py[i]; z[j] = pz[i];
        }
    #pragma unroll
        for (int j = 0; j < J; ++j) {
            if (!valid[j]) continue;
            k[j] = nc;
            if (!(x[j] < 0.f || x[j] > (float)f.nx || y[j] < 0.f || y[j] > (float)f.ny || z[j] < 0.f ||
                  z[j] > (float)f.nz)) {
                const int ci = rmax((int)(x[j] - 0.5f), 0), cj = rmax((int)(y[j] - 0.5f), 0),
                          ck = rmax((int)(z[j] - 0.5f), 0);
                const int s = f.slot[((ck / B) * f.fy + cj / B) * f.fx + ci / B];
                if (s >= 0) k[j] = s * B3 + ((ck % B) * B + cj % B) * B + ci % B;
            }
            key[idx[j]] = k[j];
        }
        __syncthreads();
        // Runs of equal keys in adjacent lanes (particles arrive in the previous step's bin order) enter
        // the table once: the run's first lane finds the slot and adds the run length, the others take
        // that slot and base plus their distance from the first lane.
        const int lane = threadIdx.x & 63;
        const unsigned long long upto = lane == 63 ? ~0ull : (2ull << lane) - 1ull;     // lanes <= this one
        int hs[J], loc[J];
    #pragma unroll
        for (int j = 0; j < J; ++j) {
            const int kj = valid[j] ? k[j] : -1;
            const int pk = __shfl_up(kj, 1);
            const bool head = lane == 0 || kj != pk;
            const unsigned long long hm = __ballot(head);
            const int h = 63 - __clzll(hm & upto);                 // first lane of this lane's run
            const unsigned long long after = hm & ~upto;
            const int next = after ? __ffsll((long long)after) - 1 : 64;
            int q = 0, base = 0;
            if (head && kj >= 0) {
                // keys are >= 0 and at most J*WG are entered, so a free slot is always found
                q = (int)(((unsigned)kj * 2654435761u) >> 8) & (H - 1);
                while (true) {
                    const int old = atomicCAS(&hk[q], -1, kj);
                    if (old == -1 || old == kj) break;
                    q = (q + 1) & (H - 1);
                }
                base = atomicAdd(&hc[q], next - lane);
            }
            hs[j] = __shfl(q, h);
            loc[j] = __shfl(base, h) + (lane - h);
        }
        __syncthreads();
        {
            constexpr int U = H / WG;                // slots per thread; their atomics are in flight together
            int res[U];
    #pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = threadIdx.x + u * WG;
                const int c = hc[i];
                res[u] = c > 0 ? atomicAdd(count + hk[i], c) : 0;
            }
    #pragma unroll
            for (int u = 0; u < U; ++u) hc[threadIdx.x + u * WG] = res[u];
        }
        __syncthreads();
    #pragma unroll
        for (int j = 0; j < J; ++j)
            if (valid[j]) rank[idx[j]] = hc[hs[j]] + loc[j];
    }
};
template <int B>
__global__ void __launch_bounds__(WG) k_ckey(const float* __restrict__ px, const float* __restrict__ py,
                                             const float* __restrict__ pz, int np, FrameK f, int nc,
                                             int* __restrict__ key, int* __restrict__ rank, int* __restrict__ count) {
    k_ckey_f<B>::run(blockIdx.x, gridDim.x, px, py, pz, np, f, nc, key, rank, count);
}

// Exclusive scan of bin counts -> bin starts.
struct OpBuckets {
    const int* count;
    int* start;
    __device__ int value(int b) const { return count[b]; }
    __device__ void emit(int b, int pre, int v) const { start[b] = pre; }
};

// Sort key of physical particle q: its logical index above, q below (both < 2^31).  Same slices as
// k_ckey.
struct k_pscatter_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int np, const int* __restrict__ key, const int* __restrict__ rank,
                               const int* __restrict__ bstart, const int* __restrict__ pid,
                               unsigned long long* __restrict__ perm) {
        constexpr int J = kBinSlices;
        int idx[J], kk[J], rk[J], pd[J], st[J];
        const int wg = xcd_swizzle(vbid, vnbid);       // contiguous particle runs per XCD
    #pragma unroll
        for (int j = 0; j < J; ++j) {
            idx[j] = (wg * J + j) * WG + threadIdx.x;
            const int i = idx[j] < np ? idx[j] : 0;
            kk[j] = key[i]; rk[j] = rank[i]; pd[j] = pid[i];
        }
    #pragma unroll
        for (int j = 0; j < J; ++j) st[j] = bstart[kk[j]];
    #pragma unroll
        for (int j = 0; j < J; ++j)
            if (idx[j] < np) perm[st[j] + rk[j]] = ((unsigned long long)(unsigned)pd[j] << 32) | (unsigned)idx[j];
    }
};
__global__ void k_pscatter(int np, const int* __restrict__ key, const int* __restrict__ rank,
                           const int* __restrict__ bstart, const int* __restrict__ pid,
                           unsigned long long* __restrict__ perm) {
    k_pscatter_f::run(blockIdx.x, gridDim.x, np, key, rank, bstart, pid, perm);
}

// Sort of one bin of 2..kSortReg keys in place at p (LDS or global memory), in registers with a
// bitonic network; an already sorted bin is left as it is.
__device__ __forceinline__ void sort_bin_regs(unsigned long long* p, int n) {
    unsigned long long v[kSortReg];
#pragma unroll
    for (int q = 0; q < kSortReg; ++q) v[q] = q < n ? p[q] : kSortPad;
    bool sorted = true;
#pragma unroll
    for (int q = 0; q + 1 < kSortReg; ++q) sorted = sorted && v[q] < v[q + 1];
    if (sorted) return;
#pragma unroll
    for (int k = 2; k <= kSortReg; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int t = 0; t < kSortReg; ++t) {
                const int u = t ^ j;
                if (u > t) {
                    const unsigned long long lo = v[t] < v[u] ? v[t] : v[u], hi = v[t] < v[u] ? v[u] : v[t];
                    if ((t & k) == 0) { v[t] = lo; v[u] = hi; } else { v[t] = hi; v[u] = lo; }
                }
            }
#pragma unroll
    for (int q = 0; q < kSortReg; ++q)
        if (q < n) p[q] = v[q];
}

// Sort every bin's keys ascending, i.e. by logical index, which turns the permutation into a stable
// counting sort of the reference order that does not depend on the atomic arrival order.  One
// workgroup per block (its B^3 bins are consecutive, so its keys are one contiguous range) and one
// for the extra bin Nc: the range is staged through LDS with coalesced loads and stores, and each
// bin of up to kSortReg keys is sorted there by one thread in registers.  Bins of up to 64 keys are
// listed for k_sort_wave, longer ones for k_sort_long (both sort in global memory afterwards; the
// range written back holds their keys unchanged).  A block with more than kSortBlockCap keys is
// sorted bin by bin in global memory.
constexpr int kSortBlockCap = 6144;                // keys of one block staged in LDS (48 KB)
constexpr int kStageLoads = 8;
template <int B>
struct k_sort_block_f {
    static constexpr int kLB = WG;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, const int* __restrict__ bstart, int nb, int nc,
                                                       unsigned long long* __restrict__ perm, int* __restrict__ midbins,
                                                       int* __restrict__ nmid, int* __restrict__ longbins,
                                                       int* __restrict__ nlong) {
        constexpr int B3 = B * B * B;
        __shared__ unsigned long long sh[kSortBlockCap];
        __shared__ int cs[B3 + 1];
        const int s = vbid;
        const int b0 = s < nb ? s * B3 : nc;           // first bin of this workgroup
        const int nbin = s < nb ? B3 : 1;
        for (int l = threadIdx.x; l <= nbin; l += WG) cs[l] = bstart[b0 + l];
        __syncthreads();
        const int r0 = cs[0], n = cs[nbin] - r0;
        if (n <= 1) return;
        const bool staged = n <= kSortBlockCap;
        if (staged) {                                  // kStageLoads loads per thread in flight at a time
            constexpr int U = kStageLoads;
            for (int t0 = 0; t0 < n; t0 += U * WG) {
                unsigned long long v[U];
    #pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int t = t0 + u * WG + threadIdx.x;
                    v[u] = t < n ? perm[r0 + t] : 0ull;
                }
    #pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int t = t0 + u * WG + threadIdx.x;
                    if (t < n) sh[t] = v[u];
                }
            }
            __syncthreads();
        }
        for (int l = threadIdx.x; l < nbin; l += WG) {
            const int a = cs[l] - r0, m = cs[l + 1] - cs[l];
            if (m <= 1) continue;
            if (m > kSortReg) {
                if (m <= 64) midbins[atomicAdd(nmid, 1)] = b0 + l;
                else longbins[atomicAdd(nlong, 1)] = b0 + l;
                continue;
            }
            if (staged) sort_bin_regs(sh + a, m);
            else sort_bin_regs(perm + r0 + a, m);
        }
        if (staged) {
            __syncthreads();
            for (int t = threadIdx.x; t < n; t += WG) perm[r0 + t] = sh[t];
        }
    }
};
template <int B>
__global__ void __launch_bounds__(WG) k_sort_block(const int* __restrict__ bstart, int nb, int nc,
                                                   unsigned long long* __restrict__ perm, int* __restrict__ midbins,
                                                   int* __restrict__ nmid, int* __restrict__ longbins,
                                                   int* __restrict__ nlong) {
    k_sort_block_f<B>::run(blockIdx.x, gridDim.x, bstart, nb, nc, perm, midbins, nmid, longbins, nlong);
}

// Bins of kSortReg+1 .. 64 entries: one wave per bin, one key per lane, bitonic sort across the
// lanes with butterfly shuffles (no LDS, no barrier).  The grid is fixed and strides over the list.
__device__ __forceinline__ unsigned long long shfl_xor_u64(unsigned long long v, int m) {
    const unsigned lo = (unsigned)__shfl_xor((int)(unsigned)(v & 0xffffffffull), m, 64);
    const unsigned hi = (unsigned)__shfl_xor((int)(unsigned)(v >> 32), m, 64);
    return ((unsigned long long)hi << 32) | lo;
}
struct k_sort_wave_f {
    static constexpr int kLB = WG;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, const int* __restrict__ bstart, const int* __restrict__ bins,
                                                      const int* __restrict__ nbins, unsigned long long* __restrict__ perm) {
        const int lane = threadIdx.x & 63;
        const int nw = *nbins;
        for (int e = vbid * (WG / 64) + (threadIdx.x >> 6); e < nw; e += vnbid * (WG / 64)) {
            const int b = bins[e];
            const int s0 = bstart[b], n = bstart[b + 1] - s0;
            unsigned long long v = lane < n ? perm[s0 + lane] : kSortPad;
            for (int k = 2; k <= 64; k <<= 1)
                for (int j = k >> 1; j > 0; j >>= 1) {
                    const unsigned long long o = shfl_xor_u64(v, j);
                    const bool keep_min = ((lane & j) == 0) == ((lane & k) == 0);
                    v = keep_min ? (o < v ? o : v) : (o < v ? v : o);
                }
            if (lane < n) perm[s0 + lane] = v;
        }
    }
};
__global__ void __launch_bounds__(WG) k_sort_wave(const int* __restrict__ bstart, const int* __restrict__ bins,
                                                  const int* __restrict__ nbins, unsigned long long* __restrict__ perm) {
    k_sort_wave_f::run(blockIdx.x, gridDim.x, bstart, bins, nbins, perm);
}

// Bins longer than 64: one 1024-thread workgroup per listed bin, bitonic sort in LDS.  The
// grid is fixed and strides over the list, whose length is read on the device.  Bins longer than
// kSortMax keep their arrival order and are counted in *overflow.
struct k_sort_long_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, const int* __restrict__ bstart, const int* __restrict__ longbins,
                                                        const int* __restrict__ nlong,
                                                        unsigned long long* __restrict__ perm, int* __restrict__ overflow) {
        __shared__ unsigned long long sh[kSortMax];
        const int nl = *nlong;
        for (int e = vbid; e < nl; e += vnbid) {
            const int b = longbins[e];
            const int s0 = bstart[b], s1 = bstart[b + 1];
            const int n = s1 - s0;
            int unsorted = 0;
            for (int t = s0 + threadIdx.x; t + 1 < s1; t += blockDim.x) unsorted |= perm[t] > perm[t + 1];
            if (!__syncthreads_or(unsorted)) continue;
            if (n > kSortMax) {
                if (threadIdx.x == 0) atomicAdd(overflow, 1);
                continue;
            }
            int m = 1;
            while (m < n) m <<= 1;
            for (int t = threadIdx.x; t < m; t += blockDim.x) sh[t] = t < n ? perm[s0 + t] : kSortPad;
            __syncthreads();
            for (int k = 2; k <= m; k <<= 1) {
                for (int j = k >> 1; j > 0; j >>= 1) {
                    for (int t = threadIdx.x; t < m; t += blockDim.x) {
                        const int u = t ^ j;
                        if (u > t) {
                            const unsigned long long a = sh[t], c = sh[u];
                            const bool up = (t & k) == 0;
                            if ((a > c) == up) { sh[t] = c; sh[u] = a; }
                        }
                    }
                    __syncthreads();
                }
            }
            for (int t = threadIdx.x; t < n; t += blockDim.x) perm[s0 + t] = sh[t];
            __syncthreads();
        }
    }
};
__global__ void __launch_bounds__(1024) k_sort_long(const int* __restrict__ bstart, const int* __restrict__ longbins,
                                                    const int* __restrict__ nlong,
                                                    unsigned long long* __restrict__ perm, int* __restrict__ overflow) {
    k_sort_long_f::run(blockIdx.x, gridDim.x, bstart, longbins, nlong, perm, overflow);
}

// Physical reorder into bin order: new particle q takes the old particle in the low half of
// perm[q] and its logical index from the high half.  `vsz` is the byte size of the velocity
// elements (4, or 2 for FP16 storage).
struct ReorderArrays {
    const void* src[7];     // px, py, pz, vx, vy, vz, adj
    void* dst[7];
    int vsz;
    int n;                  // arrays moved: 7, or 6 when adj is rewritten before it is read again
};
struct k_reorder_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int np, const unsigned long long* __restrict__ perm, ReorderArrays ra,
                              int* __restrict__ pid_out) {
        const int q = vbid * WG + threadIdx.x;
        if (q >= np) return;
        const unsigned long long k = perm[q];
        const int o = (int)(unsigned)(k & 0xffffffffull);
        pid_out[q] = (int)(k >> 32);
    #pragma unroll
        for (int a = 0; a < 7; ++a) {
            if (a == 6 && ra.n < 7) break;
            if (a >= 3 && a < 6 && ra.vsz == 2) ((_Float16*)ra.dst[a])[q] = ((const _Float16*)ra.src[a])[o];
            else ((float*)ra.dst[a])[q] = ((const float*)ra.src[a])[o];
        }
    }
};
__global__ void k_reorder(int np, const unsigned long long* __restrict__ perm, ReorderArrays ra,
                          int* __restrict__ pid_out) {
    k_reorder_f::run(blockIdx.x, gridDim.x, np, perm, ra, pid_out);
}

// k_reorder over the particles from *first (read on the device) to np: the bin of the particles outside
// the domain, which no block's P2G workgroup covers when P2G moves the particles itself.  A fixed grid
// strides over the range (usually empty).
struct k_reorder_tail_f {
    static constexpr int kLB = WG;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, const int* __restrict__ first, int np,
                                               const unsigned long long* __restrict__ perm, ReorderArrays ra,
                                               int* __restrict__ pid_out) {
        for (int q = *first + (int)(vbid * WG + threadIdx.x); q < np; q += (int)(vnbid * WG)) {
            const unsigned long long k = perm[q];
            const int o = (int)(unsigned)(k & 0xffffffffull);
            pid_out[q] = (int)(k >> 32);
    #pragma unroll
            for (int a = 0; a < 7; ++a) {
                if (a == 6 && ra.n < 7) break;
                if (a >= 3 && a < 6 && ra.vsz == 2) ((_Float16*)ra.dst[a])[q] = ((const _Float16*)ra.src[a])[o];
                else ((float*)ra.dst[a])[q] = ((const float*)ra.src[a])[o];
            }
        }
    }
};
__global__ void __launch_bounds__(WG) k_reorder_tail(const int* __restrict__ first, int np,
                                                     const unsigned long long* __restrict__ perm, ReorderArrays ra,
                                                     int* __restrict__ pid_out) {
    k_reorder_tail_f::run(blockIdx.x, gridDim.x, first, np, perm, ra, pid_out);
}

// Slots of the 3x3x3 block neighbourhood of block slot s (index (dz+1)*9+(dy+1)*3+(dx+1)),
// -1 when inactive or outside the frame.
__device__ __forceinline__ void load_nbhd(const FrameK& f, int s, int* nsl) {
    if (threadIdx.x < 27) {
        const int t = f.coord[s];
        const int bi = t % f.fx, bj = (t / f.fx) % f.fy, bk = t / (f.fx * f.fy);
        const int dx = threadIdx.x % 3 - 1, dy = (threadIdx.x / 3) % 3 - 1, dz = threadIdx.x / 9 - 1;
        const int i = bi + dx, j = bj + dy, k = bk + dz;
        nsl[threadIdx.x] = (i < 0 || j < 0 || k < 0 || i >= f.fx || j >= f.fy || k >= f.fz)
                               ? -1 : f.slot[(k * f.fy + j) * f.fx + i];
    }
}

// Global storage index of tile cell (ti,tj,tk) in [0,B+2)^3 around block (origin = block*B - 1).
template <int B>
__device__ __forceinline__ int tile_cell(const int* nsl, int ti, int tj, int tk) {
    const int ai = ti - 1, aj = tj - 1, ak = tk - 1;     // cell offset from the block origin, -1..B
    const int dx = ai < 0 ? -1 : (ai >= B ? 1 : 0);
    const int dy = aj < 0 ? -1 : (aj >= B ? 1 : 0);
    const int dz = ak < 0 ? -1 : (ak >= B ? 1 : 0);
    const int s = nsl[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)];
    if (s < 0) return -1;
    const int li = ai - dx * B, lj = aj - dy * B, lk = ak - dz * B;
    return s * Blk<B>::B3 + (lk * B + lj) * B + li;
}

// ---- P2G in two passes: each block adds only its own particles ----
//
// Pass 1 (k_p2g_own, one workgroup per active block, one thread per own base cell): a thread sums the
// 8 corner terms of its base cell's particles in registers, and the block's cells plus the layer of
// cells one past its +x/+y/+z faces (the "+1 layer") receive them in 8 LDS passes, one per corner
// offset.  A wave owns 64 consecutive base cells, whose bins are one contiguous particle range: the
// wave stages that range through LDS in rounds of C particles with coalesced loads, and each lane
// then reads its bin's particles from LDS.  Cells whose 8 contributing base cells are all in the
// block (local x, y, z >= 1) are finished here; the block's face cells (a local coordinate 0) keep
// their partial sums in the output fields.  A +1-layer cell belongs to one of the 7 blocks above
// (+x, +y, +z and the edge and corner blocks): its partial goes to that block's "incoming" area, at
// the same +1-layer index (each entry has exactly one writer), and a block zeroes the entries of its
// own area whose writer is inactive or outside the frame.  A corner in an inactive block is dropped
// (isWritable).
// Pass 2 (k_p2g_faces, one thread per face cell) adds a face cell's incoming partials in the fixed
// order -x, -y, -x-y, -z, -x-z, -y-z, -x-y-z, then normalises and classifies.  Every sum is in a fixed
// order, so P2G is bit-reproducible; no particle is read twice.
template <int B> struct P2GOwn {
    static constexpr int B3 = B * B * B;
    static constexpr int THREADS = B3;               // one thread per own base cell
    static constexpr int NW = B3 / 64;               // waves per workgroup
    static constexpr int T = B + 1, T3 = T * T * T;  // the block's cells and its +1 layer
    static constexpr int NPART = T3 - B3;            // +1-layer cells: T^2 (x = B) + B*T (y = B) + B^2 (z = B)
    static constexpr int C = 256;                    // particles staged per wave and round
    static constexpr int NFACE = B3 - (B - 1) * (B - 1) * (B - 1);   // cells with a local coordinate 0
    static constexpr int FACE_THREADS = (NFACE + 63) / 64 * 64;
    static_assert(B3 % 64 == 0, "a wave owns 64 base cells");
};

// Index of a +1-layer cell (x, y, z) in [0,B]^3 with a coordinate equal to B, and back.
template <int B> __device__ __forceinline__ int p2g_layer_index(int x, int y, int z) {
    constexpr int T = B + 1;
    if (x == B) return y + T * z;
    if (y == B) return T * T + x + B * z;
    return T * T + B * T + x + B * y;
}
template <int B> __device__ __forceinline__ void p2g_layer_cell(int q, int& x, int& y, int& z) {
    constexpr int T = B + 1;
    if (q < T * T) { x = B; y = q % T; z = q / T; return; }
    q -= T * T;
    if (q < B * T) { y = B; x = q % B; z = q / B; return; }
    q -= B * T;
    z = B; x = q % B; y = q / B;
}

// kernel 2 of Rasterization::ParticlesToVelField and the classification (Rasterization.cu:171-190,
// FLIPMethod.cu:242-280): normalise a cell's sums and write every P2G output of cell c.
template <int B>
__device__ __forceinline__ void p2g_finish(int c, int i, int j, int k, float m, float sx, float sy, float sz, float cn,
                                           const CellFields& cf, float* __restrict__ cx, float* __restrict__ cy,
                                           float* __restrict__ cz, uint8_t* __restrict__ type, const ColK& col,
                                           const FrameK& f, float col_scale) {
    // The quotients are formed for every cell and selected (where m = 0 they are not finite and not
    // kept), so the sums' loads are not moved into a branch taken only by cells with mass.
    const bool has = m > 0.f;
    const float qx = sx / m, qy = sy / m, qz = sz / m;
    float vx = has ? qx : 0.f, vy = has ? qy : 0.f, vz = has ? qz : 0.f;
    uint8_t ty8 = has ? (uint8_t)kFluid : (uint8_t)kAir;
    if (col.any()) ty8 = classify_cell<B>(col, f, col_scale, i, j, k, m, vx, vy, vz);
    cf.mass[c] = m > 0.f ? m : 0.f;
    cf.vx[c] = vx; cf.vy[c] = vy; cf.vz[c] = vz;
    cf.cnt[c] = cn;
    cx[c] = vx; cy[c] = vy; cz[c] = vz;
    type[c] = ty8;
}

// The reorder into bin order done by P2G's first pass (RO = true): each staged slot q of the bin-order
// range takes the particle perm[q] names (its previous physical index below, its logical index above)
// from the previous arrays, which the pass reads, and writes it, with its logical index, at q of the
// next arrays.  This is k_reorder's work without its separate pass over the particles.
struct P2GReorder {
    const unsigned long long* perm;
    float* pos[3];
    VelPtr vel;
    int* pid;
};

template <int B, bool VH, bool RO = false>
struct k_p2g_own_f {
    static constexpr int kLB = P2GOwn<B>::THREADS;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const int vbid, const int vnbid,
            const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz, VelPtr uv,
            const int* __restrict__ cstart, FrameK f, CellFields cf,
            float* __restrict__ cx, float* __restrict__ cy, float* __restrict__ cz, uint8_t* __restrict__ type, ColK col,
            float col_scale, float* __restrict__ part, P2GReorder ro) {
        using P = P2GOwn<B>;
        constexpr int B3 = P::B3, T = P::T, T3 = P::T3, C = P::C, R = C / 64;
        const VelT<VH> u = vel_typed<VH>(uv);
        __shared__ float st[P::NW][6][C];
        __shared__ float tm[T3], tx[T3], ty[T3], tz[T3], tc[T3];
        __shared__ int nsl[27];
        const int s = xcd_swizzle(vbid, vnbid);
        const int t = threadIdx.x, lane = t & 63, w = t >> 6;
        load_nbhd(f, s, nsl);
        for (int q = t; q < T3; q += P::THREADS) { tm[q] = 0.f; tx[q] = 0.f; ty[q] = 0.f; tz[q] = 0.f; tc[q] = 0.f; }
        const int t0 = f.coord[s];
        const int ox = (t0 % f.fx) * B, oy = ((t0 / f.fx) % f.fy) * B, oz = (t0 / (f.fx * f.fy)) * B;
        const int li = t % B, lj = (t / B) % B, lk = t / (B * B);          // base cell - origin
        const int q0 = cstart[s * B3 + t], q1 = cstart[s * B3 + t + 1];
        const int w0 = __shfl(q0, 0, 64), w1 = __shfl(q1, 63, 64);        // the wave's particle range
        const int ci = ox + li, cj = oy + lj, ck = oz + lk;
        float am[8], ax[8], ay[8], az[8];
    #pragma unroll
        for (int c = 0; c < 8; ++c) { am[c] = 0.f; ax[c] = 0.f; ay[c] = 0.f; az[c] = 0.f; }
        for (int c0 = w0; c0 < w1; c0 += C) {
            float r[R][6];
            if constexpr (RO) {
                unsigned long long kk[R];
    #pragma unroll
                for (int k = 0; k < R; ++k) {
                    const int n = c0 + k * 64 + lane;
                    kk[k] = ro.perm[n < w1 ? n : w0];                     // lanes past the range re-read the first
                }
    #pragma unroll
                for (int k = 0; k < R; ++k) {
                    const unsigned m = (unsigned)(kk[k] & 0xffffffffull);  // previous physical index
                    r[k][0] = ld_u32(px, m); r[k][1] = ld_u32(py, m); r[k][2] = ld_u32(pz, m);
                    r[k][3] = u.ldu(0, m); r[k][4] = u.ldu(1, m); r[k][5] = u.ldu(2, m);
                }
                const VelT<VH> ud = vel_typed<VH>(ro.vel);
    #pragma unroll
                for (int k = 0; k < R; ++k) {
                    const int n = c0 + k * 64 + lane;
                    if (n < w1) {
                        const unsigned un = (unsigned)n;
                        st_u32(ro.pos[0], un, r[k][0]); st_u32(ro.pos[1], un, r[k][1]); st_u32(ro.pos[2], un, r[k][2]);
                        ud.stu(0, un, r[k][3]); ud.stu(1, un, r[k][4]); ud.stu(2, un, r[k][5]);
                        st_u32(ro.pid, un, (int)(kk[k] >> 32));
                    }
                }
            } else {
    #pragma unroll
                for (int k = 0; k < R; ++k) {
                    const int n = c0 + k * 64 + lane;
                    const unsigned m = (unsigned)(n < w1 ? n : w0);           // lanes past the range re-read the first
                    r[k][0] = ld_u32(px, m); r[k][1] = ld_u32(py, m); r[k][2] = ld_u32(pz, m);
                    r[k][3] = u.ldu(0, m); r[k][4] = u.ldu(1, m); r[k][5] = u.ldu(2, m);
                }
            }
    #pragma unroll
            for (int k = 0; k < R; ++k)
    #pragma unroll
                for (int a = 0; a < 6; ++a) st[w][a][k * 64 + lane] = r[k][a];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const int lo = q0 > c0 ? q0 : c0, hi = q1 < c0 + C ? q1 : c0 + C;
            for (int q = lo; q < hi; ++q) {
                const int e = q - c0;
                const float x = st[w][0][e], y = st[w][1][e], z = st[w][2][e];
                const float vx = st[w][3][e], vy = st[w][4][e], vz = st[w][5][e];
    #pragma unroll
                for (int c = 0; c < 8; ++c) {
                    const int i = ci + (c & 1), j = cj + ((c >> 1) & 1), k = ck + (c >> 2);
                    const float xw = tent(x - ((float)i + 0.5f));
                    const float yw = tent(y - ((float)j + 0.5f));
                    const float zw = tent(z - ((float)k + 0.5f));
                    // w = xw*yw*zw + 1e-5 and the momentum terms with fused multiply-adds
                    const float wt = __builtin_fmaf(xw * yw, zw, 1e-05f);
                    am[c] += wt;
                    ax[c] = __builtin_fmaf(wt, vx, ax[c]);
                    ay[c] = __builtin_fmaf(wt, vy, ay[c]);
                    az[c] = __builtin_fmaf(wt, vz, az[c]);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        const float ac = (float)(q1 - q0);              // count term: 1 per particle at every corner
        __syncthreads();
    #pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int ai = li + (c & 1), aj = lj + ((c >> 1) & 1), ak = lk + (c >> 2);
            // a corner past n-1 is skipped (the Min(c+1, n-1) bound)
            if (ox + ai <= f.nx - 1 && oy + aj <= f.ny - 1 && oz + ak <= f.nz - 1) {
                const int q = (ak * T + aj) * T + ai;
                tm[q] += am[c];
                tx[q] += ax[c];
                ty[q] += ay[c];
                tz[q] += az[c];
                tc[q] += ac;
            }
            __syncthreads();
        }
        constexpr int NP = P::NPART;
        for (int q = t; q < NP; q += P::THREADS) {       // own incoming entries whose writer does not exist
            int x, y, z;
            p2g_layer_cell<B>(q, x, y, z);
            const int ex = x == B, ey = y == B, ez = z == B;
            if (nsl[(1 - ez) * 9 + (1 - ey) * 3 + (1 - ex)] < 0) {
                float* o = part + (size_t)s * NP * 5 + q;
    #pragma unroll
                for (int v = 0; v < 5; ++v) o[v * NP] = 0.f;
            }
        }
        for (int q = t; q < T3; q += P::THREADS) {
            const int x = q % T, y = (q / T) % T, z = q / (T * T);
            if (x == B || y == B || z == B) {
                const int ex = x == B, ey = y == B, ez = z == B;
                const int ts = nsl[(1 + ez) * 9 + (1 + ey) * 3 + (1 + ex)];     // the block the cell lies in
                if (ts >= 0) {
                    float* o = part + (size_t)ts * NP * 5 + p2g_layer_index<B>(x, y, z);
                    o[0] = tm[q]; o[NP] = tx[q]; o[2 * NP] = ty[q]; o[3 * NP] = tz[q]; o[4 * NP] = tc[q];
                }
                continue;
            }
            const int c = s * B3 + (z * B + y) * B + x;
            if (x > 0 && y > 0 && z > 0) {
                p2g_finish<B>(c, ox + x, oy + y, oz + z, tm[q], tx[q], ty[q], tz[q], tc[q], cf, cx, cy, cz, type, col, f,
                              col_scale);
            } else {                                     // face cell: partial sums, finished by k_p2g_faces
                cf.mass[c] = tm[q]; cf.vx[c] = tx[q]; cf.vy[c] = ty[q]; cf.vz[c] = tz[q]; cf.cnt[c] = tc[q];
            }
        }
    }
};
template <int B, bool VH, bool RO = false>
__global__ void __launch_bounds__(P2GOwn<B>::THREADS) k_p2g_own(
        const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz, VelPtr uv,
        const int* __restrict__ cstart, FrameK f, CellFields cf,
        float* __restrict__ cx, float* __restrict__ cy, float* __restrict__ cz, uint8_t* __restrict__ type, ColK col,
        float col_scale, float* __restrict__ part, P2GReorder ro) {
    k_p2g_own_f<B, VH, RO>::run(blockIdx.x, gridDim.x, px, py, pz, uv, cstart, f, cf, cx, cy, cz, type, col, col_scale,
                                part, ro);
}

template <int B>
struct k_p2g_faces_f {
    static constexpr int kLB = P2GOwn<B>::FACE_THREADS;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, FrameK f, CellFields cf, float* __restrict__ cx,
                                                                           float* __restrict__ cy, float* __restrict__ cz,
                                                                           uint8_t* __restrict__ type, ColK col,
                                                                           float col_scale, const float* __restrict__ part) {
        using P = P2GOwn<B>;
        constexpr int NP = P::NPART;
        const int s = vbid;
        int fc = threadIdx.x;
        if (fc >= P::NFACE) return;
        int x, y, z;                                     // face cells: z = 0, then y = 0, then x = 0
        if (fc < B * B) { x = fc % B; y = fc / B; z = 0; }
        else if ((fc -= B * B) < B * (B - 1)) { x = fc % B; y = 0; z = 1 + fc / B; }
        else { fc -= B * (B - 1); x = 0; y = 1 + fc % (B - 1); z = 1 + fc / (B - 1); }
        const int c = s * P::B3 + (z * B + y) * B + x;
        const int t0 = f.coord[s];
        float m = cf.mass[c], sx = cf.vx[c], sy = cf.vy[c], sz = cf.vz[c], cn = cf.cnt[c];
        const float* in = part + (size_t)s * NP * 5;
        // The entries of the up to seven blocks below are loaded together, one round trip to memory (an
        // entry that does not apply to this cell reads entry 0), then added in the order e = 1..7 with
        // selects, so a skipped entry leaves the sums exactly as they are.
        float pv[7][5];
        bool use[7];
    #pragma unroll
        for (int e = 1; e < 8; ++e) {
            const int ex = e & 1, ey = (e >> 1) & 1, ez = e >> 2;
            use[e - 1] = !((ex && x) || (ey && y) || (ez && z));
            const int q = use[e - 1] ? p2g_layer_index<B>(x + ex * B, y + ey * B, z + ez * B) : 0;
    #pragma unroll
            for (int v = 0; v < 5; ++v) pv[e - 1][v] = in[v * NP + q];
        }
    #pragma unroll
        for (int e = 0; e < 7; ++e) {
            m = use[e] ? m + pv[e][0] : m;
            sx = use[e] ? sx + pv[e][1] : sx;
            sy = use[e] ? sy + pv[e][2] : sy;
            sz = use[e] ? sz + pv[e][3] : sz;
            cn = use[e] ? cn + pv[e][4] : cn;
        }
        const int ox = (t0 % f.fx) * B, oy = ((t0 / f.fx) % f.fy) * B, oz = (t0 / (f.fx * f.fy)) * B;
        p2g_finish<B>(c, ox + x, oy + y, oz + z, m, sx, sy, sz, cn, cf, cx, cy, cz, type, col, f, col_scale);
    }
};
template <int B>
__global__ void __launch_bounds__(P2GOwn<B>::FACE_THREADS) k_p2g_faces(FrameK f, CellFields cf, float* __restrict__ cx,
                                                                       float* __restrict__ cy, float* __restrict__ cz,
                                                                       uint8_t* __restrict__ type, ColK col,
                                                                       float col_scale, const float* __restrict__ part) {
    k_p2g_faces_f<B>::run(blockIdx.x, gridDim.x, f, cf, cx, cy, cz, type, col, col_scale, part);
}

// G2P for one particle given a cell-value fetcher (FLIPMethod.cu:86-171, SparseField.h:150-196).
// COL = false compiles the collision queries out (no collision source is active), which leaves the
// kernel with fewer registers and so more waves in flight.
template <int B, bool COL, class Vel, class Fetch>
__device__ __forceinline__ int g2p_particle(int n, float* __restrict__ px, float* __restrict__ py, float* __restrict__ pz,
                                            const Vel& u, float* __restrict__ adj, const FrameK& f, const G2PParams& gp,
                                            const ColK& col, uint8_t* __restrict__ removed, Fetch fetch) {
    const unsigned un = (unsigned)n;
    float x = ld_u32(px, un), y = ld_u32(py, un), z = ld_u32(pz, un);
    float vx = u.ldu(0, un), vy = u.ldu(1, un), vz = u.ldu(2, un);
    const int i = rclamp((int)(x - 0.5f), 0, f.nx - 2);
    const int j = rclamp((int)(y - 0.5f), 0, f.ny - 2);
    const int k = rclamp((int)(z - 0.5f), 0, f.nz - 2);
    // Clamp(d, 0, 1) as max/min (one clamped max): the same value, since d = p - (i + 0.5) is never
    // -0 and never NaN for a particle that has a bin
    const float fxw = __builtin_fminf(__builtin_fmaxf(x - ((float)i + 0.5f), 0.f), 1.f);
    const float fyw = __builtin_fminf(__builtin_fmaxf(y - ((float)j + 0.5f), 0.f), 1.f);
    const float fzw = __builtin_fminf(__builtin_fmaxf(z - ((float)k + 0.5f), 0.f), 1.f);
    // fields fi: 0 cnt, 1-3 cx, cy, cz, 4-6 vx0, vy0, vz0; corner q = dx + 2dy + 4dz.  One field is
    // fetched and interpolated at a time, which keeps 8 values live instead of 56.
    auto corner = [&](int fi, int q) -> float { return fetch(fi, i + (q & 1), j + ((q >> 1) & 1), k + (q >> 2)); };
    {
        // counts are sums of +1 (or the stand-in 10): never NaN or -0, so fminf is the reference's Min
        float a = __builtin_fminf(corner(0, 0), corner(0, 1));
        a = __builtin_fminf(a, __builtin_fminf(corner(0, 4), corner(0, 5)));
        a = __builtin_fminf(a, __builtin_fminf(corner(0, 2), corner(0, 3)));
        a = __builtin_fminf(a, __builtin_fminf(corner(0, 6), corner(0, 7)));
        st_u32(adj, un, a);
    }
    auto lerp = [&](int fi) -> float {
        float v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = corner(fi, q);
        const float c00 = v[0] * (1.f - fxw) + v[1] * fxw;
        const float c01 = v[4] * (1.f - fxw) + v[5] * fxw;
        const float c10 = v[2] * (1.f - fxw) + v[3] * fxw;
        const float c11 = v[6] * (1.f - fxw) + v[7] * fxw;
        const float c0 = c00 * (1.f - fyw) + c10 * fyw;
        const float c1 = c01 * (1.f - fyw) + c11 * fyw;
        return c0 * (1.f - fzw) + c1 * fzw;
    };
    const float gvx = lerp(1), gvy = lerp(2), gvz = lerp(3);
    const float fx_ = gvx - lerp(4), fy_ = gvy - lerp(5), fz_ = gvz - lerp(6);
    vx += fx_; vy += fy_; vz += fz_;
    vx = vx * gp.flip + gvx * gp.pic;
    vy = vy * gp.flip + gvy * gp.pic;
    vz = vz * gp.flip + gvz * gp.pic;
    if (gp.tank && y <= gp.wall) {
        const float lx = (float)f.nx + 0.5f, lz = (float)f.nz + 0.5f, band = gp.band;
        float w = 1.f;
        if (x < band) w = rmin(w, rclamp((x - 0.f) / band, 0.f, 1.f));
        if (x > lx - band) w = rmin(w, rclamp((lx - x) / band, 0.f, 1.f));
        if (z < band) w = rmin(w, rclamp((z - 0.f) / band, 0.f, 1.f));
        if (z > lz - band) w = rmin(w, rclamp((lz - z) / band, 0.f, 1.f));
        vx *= w; vy *= w; vz *= w;
    }
    const bool cols = COL && col.any();
    if (cols && col_phi<B>(col, f, x, y, z) < 0.f) {          // collision push (FLIPMethod.cu:116-119)
        float cv[3];
        col_vec<B>(col, f, 0, x, y, z, cv);
        x += cv[0] * gp.dt;
        y += cv[1] * gp.dt;
        z += cv[2] * gp.dt;
    }
    // advection substeps (FLIPMethod.cu:121-159)
    const float len = sqrtf(vx * vx + vy * vy + vz * vz);
    const int iter = 1 + rmin((int)(len * gp.dt * 1.5f), rmax(gp.max_substeps, 1));
    const float sub = gp.dt / (float)iter;
    const float fnx = (float)f.nx, fny = (float)f.ny, fnz = (float)f.nz;
    bool hit = false;
    for (int s = 0; s < iter; ++s) {
        x += vx * sub;
        y += vy * sub;
        z += vz * sub;
        if (y <= gp.wall) {
            if (x < 0.f) { x = 0.f; if (vx < 0.f) vx = 0.f; }
            if (y < 0.f) { y = 0.f; if (vy < 0.f) vy = 0.f; }
            if (z < 0.f) { z = 0.f; if (vz < 0.f) vz = 0.f; }
            if (x > fnx) { x = fnx; if (vx > 0.f) vx = 0.f; }
            if (y > fny) { y = fny; if (vy > 0.f) vy = 0.f; }
            if (z > fnz) { z = fnz; if (vz > 0.f) vz = 0.f; }
        }
        if (cols) {                                             // projection out of the source
            const float phi = col_phi<B>(col, f, x, y, z);
            if (phi < 0.f) {
                float nr[3];
                col_vec<B>(col, f, 1, x, y, z, nr);
                normalize3(nr);
                x -= nr[0] * phi;
                y -= nr[1] * phi;
                z -= nr[2] * phi;
                hit = true;
            }
        }
    }
    int rem = 0;
    if (hit) {                                                  // FLIPMethod.cu:161-164
        const float phi = col_phi<B>(col, f, x, y, z);
        if (phi > 3.f || phi < 0.f) rem = 1;
    }
    if (x < 0.f || x > fnx || y < 0.f || y > fny || z < 0.f || z > fnz) rem = 1;   // Grid::inside
    st_u32(removed, un, (uint8_t)rem);
    vx += gp.gx * gp.dt;
    vy += gp.gy * gp.dt;
    vz += gp.gz * gp.dt;
    st_u32(px, un, x); st_u32(py, un, y); st_u32(pz, un, z);
    u.stu(0, un, vx); u.stu(1, un, vy); u.stu(2, un, vz);
    return rem;
}

struct G2PFields {
    const float* cnt;
    const float* cx;
    const float* cy;
    const float* cz;
    const float* vx0;
    const float* vy0;
    const float* vz0;
};

// G2P for the particles whose base cell lies in one block (bins s*B^3 .. (s+1)*B^3-1), from an LDS
// tile of the 7 gathered fields over the block and a 1-cell halo, which holds every stencil
// i = clamp((int)(p-0.5), 0, n-2) .. i+1 of those particles.  Workgroup nb takes bin Nc (particles
// outside the domain) and reads the fields through the block table.
template <int B, bool COL, bool VH>
struct k_g2p_tile_f {
    static constexpr int kLB = WG;
    static constexpr int kWPE = COL ? 4 : 5;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, float* __restrict__ px, float* __restrict__ py, float* __restrict__ pz,
                                                    VelPtr uv, float* __restrict__ adj,
                                                    const int* __restrict__ cstart, int nb, FrameK f, G2PFields gf,
                                                    G2PParams gp, ColK col, uint8_t* __restrict__ removed,
                                                    int* __restrict__ nremoved) {
        constexpr int T = B + 2, T3 = T * T * T, B3 = B * B * B;
        const VelT<VH> u = vel_typed<VH>(uv);
        __shared__ float tile[7][T3];
        __shared__ int nsl[27];
        const int s = xcd_swizzle(vbid, vnbid);
        const int p0 = cstart[s * B3], p1 = s < nb ? cstart[(s + 1) * B3] : cstart[nb * B3 + 1];
        if (p0 == p1) return;
        const float* src[7] = {gf.cnt, gf.cx, gf.cy, gf.cz, gf.vx0, gf.vy0, gf.vz0};
        int rem_count = 0;
        if (s == nb) {
            // particles outside the domain: global lookups (cells of inactive blocks read background)
            auto fetch = [&](int fi, int i, int j, int k) -> float {
                const int c = cell_index<B>(f, i, j, k);
                return c >= 0 ? src[fi][c] : (fi == 0 ? 10.f : 0.f);
            };
            for (int q = p0 + threadIdx.x; q < p1; q += WG)
                rem_count += g2p_particle<B, COL>(q, px, py, pz, u, adj, f, gp, col, removed, fetch);
        } else {
            load_nbhd(f, s, nsl);
            __syncthreads();
            // the tile's T^3 cells, R per thread: every load of a thread is issued before any is
            // stored to LDS (stand-in cell 0 where the cell's block is inactive)
            constexpr int R = (T3 + WG - 1) / WG;
            float v[R][7];
            int cl[R];
    #pragma unroll
            for (int u = 0; u < R; ++u) {
                const int t = threadIdx.x + u * WG;
                cl[u] = t < T3 ? tile_cell<B>(nsl, t % T, (t / T) % T, t / (T * T)) : -1;
                const int cs = cl[u] >= 0 ? cl[u] : 0;
    #pragma unroll
                for (int fi = 0; fi < 7; ++fi) v[u][fi] = src[fi][cs];
            }
    #pragma unroll
            for (int u = 0; u < R; ++u) {
                const int t = threadIdx.x + u * WG;
                if (t < T3) {
    #pragma unroll
                    for (int fi = 0; fi < 7; ++fi) tile[fi][t] = cl[u] >= 0 ? v[u][fi] : (fi == 0 ? 10.f : 0.f);
                }
            }
            __syncthreads();
            const int t0 = f.coord[s];
            const int ox = (t0 % f.fx) * B - 1, oy = ((t0 / f.fx) % f.fy) * B - 1, oz = (t0 / (f.fx * f.fy)) * B - 1;
            // Every corner lies in the tile: a particle binned here has its P2G base cell c (per axis
            // max((int)(p-0.5), 0)) in this block, and its G2P stencil base clamp((int)(p-0.5), 0, n-2)
            // is c, or c-1 when c = n-1, so its corners lie in [c-1, c+1], inside the block +- 1 cell.
            // The clamp only keeps the LDS index in range.
            auto fetch = [&](int fi, int i, int j, int k) -> float {
                const int ti = rclamp(i - ox, 0, T - 1), tj = rclamp(j - oy, 0, T - 1), tk = rclamp(k - oz, 0, T - 1);
                return tile[fi][(tk * T + tj) * T + ti];
            };
            for (int q = p0 + threadIdx.x; q < p1; q += WG)
                rem_count += g2p_particle<B, COL>(q, px, py, pz, u, adj, f, gp, col, removed, fetch);
        }
        // wave-aggregated removal count
        const int w = wave_sum_i(rem_count);
        if ((threadIdx.x & 63) == 0 && w) atomicAdd(nremoved, w);
    }
};
template <int B, bool COL, bool VH>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(COL ? 4 : 5))) k_g2p_tile(float* __restrict__ px, float* __restrict__ py, float* __restrict__ pz,
                                                VelPtr uv, float* __restrict__ adj,
                                                const int* __restrict__ cstart, int nb, FrameK f, G2PFields gf,
                                                G2PParams gp, ColK col, uint8_t* __restrict__ removed,
                                                int* __restrict__ nremoved) {
    k_g2p_tile_f<B, COL, VH>::run(blockIdx.x, gridDim.x, px, py, pz, uv, adj, cstart, nb, f, gf, gp, col, removed, nremoved);
}

// G2P with one thread per particle in storage order and every field read through the block table
// (particle_path 1).
template <int B>
struct k_g2p_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, float* __restrict__ px, float* __restrict__ py, float* __restrict__ pz, VelPtr u,
                          float* __restrict__ adj, int np, FrameK f, G2PFields gf, G2PParams gp, ColK col,
                          uint8_t* __restrict__ removed, int* __restrict__ nremoved) {
        const int n = vbid * WG + threadIdx.x;
        const float* src[7] = {gf.cnt, gf.cx, gf.cy, gf.cz, gf.vx0, gf.vy0, gf.vz0};
        auto fetch = [&](int fi, int i, int j, int k) -> float {
            const int c = cell_index<B>(f, i, j, k);
            return c >= 0 ? src[fi][c] : (fi == 0 ? 10.f : 0.f);
        };
        const int rem = n < np ? g2p_particle<B, true>(n, px, py, pz, u, adj, f, gp, col, removed, fetch) : 0;
        const unsigned long long m = __ballot(rem);
        if ((threadIdx.x & 63) == 0 && m) atomicAdd(nremoved, (int)__popcll(m));
    }
};
template <int B>
__global__ void k_g2p(float* __restrict__ px, float* __restrict__ py, float* __restrict__ pz, VelPtr u,
                      float* __restrict__ adj, int np, FrameK f, G2PFields gf, G2PParams gp, ColK col,
                      uint8_t* __restrict__ removed, int* __restrict__ nremoved) {
    k_g2p_f<B>::run(blockIdx.x, gridDim.x, px, py, pz, u, adj, np, f, gf, gp, col, removed, nremoved);
}

}  // namespace vare
