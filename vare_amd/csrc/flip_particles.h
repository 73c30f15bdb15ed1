// flip_particles.h -- block-binned particle kernels: the binning permutation (a stable counting
// sort of particle indices by the active slot of the block that contains them), P2G accumulated in
// an LDS tile per block, and G2P gathering from an LDS tile per block.
//
// The particle arrays stay in the reference's order (Array::remove order); the permutation only
// says in which order each block's workgroup visits them.  A particle's block is
// c = clamp((int)p/B, 0, f-1) per axis (SparseFrame.cu:145-147 before the neighbourhood loop); it is
// always active for a particle inside the domain.  Particles whose clamped block is inactive (they
// lie outside the domain) go to an extra bucket Nb, handled with global-memory lookups.
#pragma once
#include "flip_kernels.h"

namespace vare {

constexpr int kSortMax = 16384;     // longest bucket sorted in LDS (64 KB)

__device__ __forceinline__ unsigned long long lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Bucket of every particle + its arrival rank inside the bucket (wave-aggregated atomics).
template <int B>
__global__ void k_pkey(const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz, int np,
                       FrameK f, int nb, int* __restrict__ key, int* __restrict__ rank, int* __restrict__ bcount) {
    const int i = blockIdx.x * WG + threadIdx.x;
    const bool valid = i < np;
    int k = -1;
    if (valid) {
        const int bi = rmin(rmax((int)(px[i]) / B, 0), f.fx - 1);
        const int bj = rmin(rmax((int)(py[i]) / B, 0), f.fy - 1);
        const int bk = rmin(rmax((int)(pz[i]) / B, 0), f.fz - 1);
        const int s = f.slot[(bk * f.fy + bj) * f.fx + bi];
        k = s >= 0 ? s : nb;
        key[i] = k;
    }
    unsigned long long pending = __ballot(valid);
    int myrank = 0;
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int lk = __shfl(k, leader, 64);
        const unsigned long long m = __ballot(valid && k == lk);
        int base = 0;
        if ((threadIdx.x & 63) == leader) base = atomicAdd(bcount + lk, (int)__popcll(m));
        base = __shfl(base, leader, 64);
        if (valid && k == lk) myrank = base + (int)__popcll(m & lanemask_lt());
        pending &= ~m;
    }
    if (valid) rank[i] = myrank;
}

// Exclusive scan of bucket counts -> bucket starts.
struct OpBuckets {
    const int* count;
    int* start;
    __device__ int value(int b) const { return count[b]; }
    __device__ void emit(int b, int pre, int v) const { start[b] = pre; }
};

__global__ void k_pscatter(int np, const int* __restrict__ key, const int* __restrict__ rank,
                           const int* __restrict__ bstart, int* __restrict__ perm) {
    const int i = blockIdx.x * WG + threadIdx.x;
    if (i >= np) return;
    perm[bstart[key[i]] + rank[i]] = i;
}

// Sort every bucket's indices ascending (makes the permutation a stable counting sort, independent
// of the atomic arrival order).  One 1024-thread workgroup per bucket, bitonic sort in LDS.
__global__ void __launch_bounds__(1024) k_psort(const int* __restrict__ bstart, int nbucket, int np,
                                                int* __restrict__ perm, int* __restrict__ overflow) {
    __shared__ int sh[kSortMax];
    const int b = blockIdx.x;
    if (b >= nbucket) return;
    const int s0 = bstart[b];
    const int s1 = b + 1 < nbucket ? bstart[b + 1] : np;
    const int n = s1 - s0;
    if (n <= 1) return;
    // arrival order is usually already ascending: skip buckets that need no sorting
    int unsorted = 0;
    for (int t = s0 + threadIdx.x; t + 1 < s1; t += blockDim.x) unsorted |= perm[t] > perm[t + 1];
    if (!__syncthreads_or(unsorted)) return;
    if (n > kSortMax) {
        if (threadIdx.x == 0) atomicAdd(overflow, 1);
        return;
    }
    int m = 1;
    while (m < n) m <<= 1;
    for (int t = threadIdx.x; t < m; t += blockDim.x) sh[t] = t < n ? perm[s0 + t] : 0x7fffffff;
    __syncthreads();
    for (int k = 2; k <= m; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = threadIdx.x; t < m; t += blockDim.x) {
                const int u = t ^ j;
                if (u > t) {
                    const int a = sh[t], c = sh[u];
                    const bool up = (t & k) == 0;
                    if ((a > c) == up) { sh[t] = c; sh[u] = a; }
                }
            }
            __syncthreads();
        }
    }
    for (int t = threadIdx.x; t < n; t += blockDim.x) perm[s0 + t] = sh[t];
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

// P2G of one block's particles into an LDS tile of (B+2)^3 cells (Rasterization.cu:121-173 and
// :198-228 per particle), then one global atomic add per touched tile cell and field.
template <int B>
__global__ void __launch_bounds__(WG) k_p2g_tile(const float* __restrict__ px, const float* __restrict__ py,
                                                const float* __restrict__ pz, VelPtr u,
                                                const int* __restrict__ perm, const int* __restrict__ bstart, FrameK f,
                                                CellFields cf) {
    constexpr int T = B + 2, T3 = T * T * T;
    __shared__ float tm[T3], tx[T3], ty[T3], tz[T3], tc[T3];
    __shared__ int nsl[27];
    const int s = blockIdx.x;
    const int p0 = bstart[s], p1 = bstart[s + 1];
    if (p0 == p1) return;
    load_nbhd(f, s, nsl);
    for (int t = threadIdx.x; t < T3; t += WG) { tm[t] = 0.f; tx[t] = 0.f; ty[t] = 0.f; tz[t] = 0.f; tc[t] = 0.f; }
    __syncthreads();
    const int t0 = f.coord[s];
    const int ox = (t0 % f.fx) * B - 1, oy = ((t0 / f.fx) % f.fy) * B - 1, oz = (t0 / (f.fx * f.fy)) * B - 1;
    // Each thread walks a contiguous run of the bucket.  Particles sharing the base cell
    // c = max((int)(p-0.5), 0) touch the same (up to) 8 cells, so their contributions are summed in
    // registers and flushed to the LDS tile only when the base cell changes.
    const int cnt = p1 - p0;
    const int chunk = (cnt + WG - 1) / WG;
    const int q0 = p0 + threadIdx.x * chunk;
    const int q1 = rmin(q0 + chunk, p1);
    float am[8], ax[8], ay[8], az[8], ac[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) { am[c] = 0.f; ax[c] = 0.f; ay[c] = 0.f; az[c] = 0.f; ac[c] = 0.f; }
    int cur = -1;                       // packed tile coordinates of the current base cell
    unsigned cur_mask = 0;              // corners that exist and are writable
    auto flush = [&]() {
        if (cur < 0) return;
        const int bi = cur & 1023, bj = (cur >> 10) & 1023, bk = cur >> 20;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if (cur_mask & (1u << c)) {
                const int t = ((bk + (c >> 2)) * T + (bj + ((c >> 1) & 1))) * T + (bi + (c & 1));
                atomicAdd(tm + t, am[c]);
                atomicAdd(tx + t, ax[c]);
                atomicAdd(ty + t, ay[c]);
                atomicAdd(tz + t, az[c]);
                atomicAdd(tc + t, ac[c]);
            }
            am[c] = 0.f; ax[c] = 0.f; ay[c] = 0.f; az[c] = 0.f; ac[c] = 0.f;    // every corner restarts
        }
    };
    for (int q = q0; q < q1; ++q) {
        const int n = perm[q];
        const float x = px[n], y = py[n], z = pz[n];
        if (x < 0.f || x > (float)f.nx || y < 0.f || y > (float)f.ny || z < 0.f || z > (float)f.nz) continue;
        const float vx = u.ld(0, n), vy = u.ld(1, n), vz = u.ld(2, n);
        const int ci = rmax((int)(x - 0.5f), 0), cj = rmax((int)(y - 0.5f), 0), ck = rmax((int)(z - 0.5f), 0);
        const int ti = ci - ox, tj = cj - oy, tk = ck - oz;
        const int key = (tk << 20) | (tj << 10) | ti;
        if (key != cur) {
            flush();
            cur = key;
            cur_mask = 0;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int i = ci + (c & 1), j = cj + ((c >> 1) & 1), k = ck + (c >> 2);
                if (i > f.nx - 1 || j > f.ny - 1 || k > f.nz - 1) continue;    // Min(c+1, n-1) bound
                const int ai = i - ox - 1, aj = j - oy - 1, ak = k - oz - 1;
                const int dx = ai < 0 ? -1 : (ai >= B ? 1 : 0);
                const int dy = aj < 0 ? -1 : (aj >= B ? 1 : 0);
                const int dz = ak < 0 ? -1 : (ak >= B ? 1 : 0);
                if (nsl[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)] >= 0) cur_mask |= 1u << c;   // isWritable
            }
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int i = ci + (c & 1), j = cj + ((c >> 1) & 1), k = ck + (c >> 2);
            const float ddx = x - ((float)i + 0.5f), ddy = y - ((float)j + 0.5f), ddz = z - ((float)k + 0.5f);
            const float xw = rmax(0.f, 1.f - rabs(ddx));
            const float yw = rmax(0.f, 1.f - rabs(ddy));
            const float zw = rmax(0.f, 1.f - rabs(ddz));
            const float w = xw * yw * zw + 1e-05f;
            am[c] += w;
            ax[c] += w * vx;
            ay[c] += w * vy;
            az[c] += w * vz;
            ac[c] += 1.f;
        }
    }
    flush();
    __syncthreads();
    for (int t = threadIdx.x; t < T3; t += WG) {
        if (tc[t] == 0.f) continue;
        const int ti = t % T, tj = (t / T) % T, tk = t / (T * T);
        const int c = tile_cell<B>(nsl, ti, tj, tk);
        if (c < 0) continue;
        unsafeAtomicAdd(cf.mass + c, tm[t]);
        unsafeAtomicAdd(cf.vx + c, tx[t]);
        unsafeAtomicAdd(cf.vy + c, ty[t]);
        unsafeAtomicAdd(cf.vz + c, tz[t]);
        unsafeAtomicAdd(cf.cnt + c, tc[t]);
    }
}

// G2P for one particle given a cell-value fetcher (FLIPMethod.cu:86-171, SparseField.h:150-196).
template <class Fetch>
__device__ __forceinline__ int g2p_particle(int n, float* __restrict__ px, float* __restrict__ py, float* __restrict__ pz,
                                            const VelPtr& u, float* __restrict__ adj, const FrameK& f, const G2PParams& gp,
                                            uint8_t* __restrict__ removed, Fetch fetch) {
    float x = px[n], y = py[n], z = pz[n];
    float vx = u.ld(0, n), vy = u.ld(1, n), vz = u.ld(2, n);
    const int i = rclamp((int)(x - 0.5f), 0, f.nx - 2);
    const int j = rclamp((int)(y - 0.5f), 0, f.ny - 2);
    const int k = rclamp((int)(z - 0.5f), 0, f.nz - 2);
    const float fxw = rclamp(x - ((float)i + 0.5f), 0.f, 1.f);
    const float fyw = rclamp(y - ((float)j + 0.5f), 0.f, 1.f);
    const float fzw = rclamp(z - ((float)k + 0.5f), 0.f, 1.f);
    float v[7][8];     // fields: cnt, cx, cy, cz, vx0, vy0, vz0 at the 8 corners (dx + 2dy + 4dz)
#pragma unroll
    for (int q = 0; q < 8; ++q) fetch(i + (q & 1), j + ((q >> 1) & 1), k + (q >> 2), q, v);
    {
        float a = rmin(v[0][0], v[0][1]);
        a = rmin(a, rmin(v[0][4], v[0][5]));
        a = rmin(a, rmin(v[0][2], v[0][3]));
        a = rmin(a, rmin(v[0][6], v[0][7]));
        adj[n] = a;
    }
    auto lerp = [&](int fi) -> float {
        const float c00 = v[fi][0] * (1.f - fxw) + v[fi][1] * fxw;
        const float c01 = v[fi][4] * (1.f - fxw) + v[fi][5] * fxw;
        const float c10 = v[fi][2] * (1.f - fxw) + v[fi][3] * fxw;
        const float c11 = v[fi][6] * (1.f - fxw) + v[fi][7] * fxw;
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
    const float len = sqrtf(vx * vx + vy * vy + vz * vz);
    const int iter = 1 + rmin((int)(len * gp.dt * 1.5f), rmax(gp.max_substeps, 1));
    const float sub = gp.dt / (float)iter;
    const float fnx = (float)f.nx, fny = (float)f.ny, fnz = (float)f.nz;
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
    }
    const int rem = (x < 0.f || x > fnx || y < 0.f || y > fny || z < 0.f || z > fnz) ? 1 : 0;
    removed[n] = (uint8_t)rem;
    vx += gp.gx * gp.dt;
    vy += gp.gy * gp.dt;
    vz += gp.gz * gp.dt;
    px[n] = x; py[n] = y; pz[n] = z;
    u.st(0, n, vx); u.st(1, n, vy); u.st(2, n, vz);
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

// G2P for one block's particles from an LDS tile of the 7 gathered fields; bucket Nb (particles
// outside the domain) reads the fields through the block table.
template <int B>
__global__ void __launch_bounds__(WG) k_g2p_tile(float* __restrict__ px, float* __restrict__ py, float* __restrict__ pz,
                                                VelPtr u, float* __restrict__ adj, const int* __restrict__ perm,
                                                const int* __restrict__ bstart, int nb, FrameK f, G2PFields gf,
                                                G2PParams gp, uint8_t* __restrict__ removed, int* __restrict__ nremoved) {
    constexpr int T = B + 2, T3 = T * T * T;
    __shared__ float tile[7][T3];
    __shared__ int nsl[27];
    const int s = blockIdx.x;
    const int p0 = bstart[s], p1 = bstart[s + 1];
    if (p0 == p1) return;
    const float* src[7] = {gf.cnt, gf.cx, gf.cy, gf.cz, gf.vx0, gf.vy0, gf.vz0};
    int rem_count = 0;
    if (s == nb) {
        // particles outside the domain: global lookups (cells of inactive blocks read background)
        auto fetch = [&](int i, int j, int k, int q, float (&v)[7][8]) {
            const int c = cell_index<B>(f, i, j, k);
#pragma unroll
            for (int fi = 0; fi < 7; ++fi) v[fi][q] = c >= 0 ? src[fi][c] : (fi == 0 ? 10.f : 0.f);
        };
        for (int q = p0 + threadIdx.x; q < p1; q += WG)
            rem_count += g2p_particle(perm[q], px, py, pz, u, adj, f, gp, removed, fetch);
    } else {
        load_nbhd(f, s, nsl);
        __syncthreads();
        for (int t = threadIdx.x; t < T3; t += WG) {
            const int c = tile_cell<B>(nsl, t % T, (t / T) % T, t / (T * T));
#pragma unroll
            for (int fi = 0; fi < 7; ++fi) tile[fi][t] = c >= 0 ? src[fi][c] : (fi == 0 ? 10.f : 0.f);
        }
        __syncthreads();
        const int t0 = f.coord[s];
        const int ox = (t0 % f.fx) * B - 1, oy = ((t0 / f.fx) % f.fy) * B - 1, oz = (t0 / (f.fx * f.fy)) * B - 1;
        auto fetch = [&](int i, int j, int k, int q, float (&v)[7][8]) {
            const int ti = i - ox, tj = j - oy, tk = k - oz;
            if (ti >= 0 && tj >= 0 && tk >= 0 && ti < T && tj < T && tk < T) {
                const int t = (tk * T + tj) * T + ti;
#pragma unroll
                for (int fi = 0; fi < 7; ++fi) v[fi][q] = tile[fi][t];
            } else {
                const int c = cell_index<B>(f, i, j, k);
#pragma unroll
                for (int fi = 0; fi < 7; ++fi) v[fi][q] = c >= 0 ? src[fi][c] : (fi == 0 ? 10.f : 0.f);
            }
        };
        for (int q = p0 + threadIdx.x; q < p1; q += WG)
            rem_count += g2p_particle(perm[q], px, py, pz, u, adj, f, gp, removed, fetch);
    }
    // wave-aggregated removal count
    const int w = wave_sum_i(rem_count);
    if ((threadIdx.x & 63) == 0 && w) atomicAdd(nremoved, w);
}

}  // namespace vare
