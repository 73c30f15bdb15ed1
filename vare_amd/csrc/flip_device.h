// flip_device.h -- device-side building blocks of the gfx950 FLIP step.
//
// Layout in HBM (one handle):
//   particles  : structure of arrays px,py,pz,vx,vy,vz,adj (float; velocities optionally FP16) in
//                physical order, which each step's binning sets to bin order; pid[q] is the
//                reference (logical) index of physical particle q.  vort,curv stay in logical order.
//   frame      : slot[fx*fy*fz] (int32: active slot or -1), coord[slot] (linear block index),
//                nbr[slot*6] (slot of the -x,+x,-y,+y,-z,+z neighbour block, -1 inactive,
//                -2 outside the domain).
//   cells      : Nb*B^3 cells, block-major, x fastest inside a block (SparseField.h:57-61), one
//                array per scalar component (mass, vel.xyz, comp.xyz, count, p, b, r, p, Ap) plus
//                a 1-byte cell type and a 2-byte stencil descriptor.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vare {

constexpr int WG = 256;        // threads per workgroup for every kernel (4 waves of 64)

// cell types (FluidProjection.h:32-39); fluid cells carry kFluid here, the reference's
// 4+rank numbering is produced on demand by a scan.
enum : uint8_t { kAir = 0, kSolid = 1, kWall = 2, kFluid = 4 };

// stencil descriptor (uint16) of a fluid cell: bit q (q=0..5 for -x,+x,-y,+y,-z,+z) = neighbour
// is fluid, bits 8..11 = diagonal 6 + nAir - nSolidOrWall (0..12), bit 15 = cell is fluid.
constexpr uint16_t kInfoFluid = 0x8000u;

__host__ __device__ inline float rand01(uint32_t seed) {        // Random.h:18-27
    uint32_t i = (seed ^ 12345391u) * 2654435769u;
    i ^= (i << 6) ^ (i >> 26);
    i *= 2654435769u;
    i += (i << 5) ^ (i >> 12);
    return (float)i / 4294967296.0f;                             // (float)UINT_MAX rounds to 2^32
}

template <class T> __device__ __forceinline__ T rmin(T a, T b) { return (a < b) ? a : b; }   // MathUtils.h:174-178
template <class T> __device__ __forceinline__ T rmax(T a, T b) { return (a > b) ? a : b; }   // MathUtils.h:195-199
__device__ __forceinline__ float rabs(float x) { return (x < 0.f) ? -x : x; }                // MathUtils.h:129-132
// Max(0, 1 - Abs(d)) (MathUtils.h:129-132, :195-199) as one v_max_f32 with an abs source modifier:
// the same value for every d that is not NaN (a particle with a NaN coordinate has no bin).
__device__ __forceinline__ float tent(float d) { return __builtin_fmaxf(0.f, 1.f - __builtin_fabsf(d)); }
template <class T> __device__ __forceinline__ T rclamp(T x, T lo, T hi) {                   // MathUtils.h:286-292
    if (x < lo) return lo;
    if (x > hi) return hi;
    return x;
}

// Particle velocity arrays, stored as FP32 or FP16 (half = 1).  The branch is uniform across a
// launch.  FP16 stores round to nearest even.
struct VelPtr {
    void* p[3];
    int half;
    __device__ __forceinline__ float ld(int a, int i) const {
        return half ? (float)((const _Float16*)p[a])[i] : ((const float*)p[a])[i];
    }
    __device__ __forceinline__ void st(int a, int i, float v) const {
        if (half) ((_Float16*)p[a])[i] = (_Float16)v;
        else ((float*)p[a])[i] = v;
    }
    __device__ __forceinline__ float ldu(int a, unsigned i) const { return ld(a, (int)i); }
    __device__ __forceinline__ void stu(int a, unsigned i, float v) const { st(a, (int)i, v); }
};

// The same arrays with the element type fixed at compile time (H: FP16), for the hot particle
// kernels: with no run-time branch per load, every load of a particle issues before any conversion.
// Element i of an array by a 32-bit byte offset from its base pointer, which lets a load take a
// uniform base in scalar registers and one vector offset.  Particle arrays hold at most 2^30
// elements (ensure_particles), so i * sizeof(T) < 2^32.
template <class T> __device__ __forceinline__ T ld_u32(const T* p, unsigned i) {
    return *(const T*)((const char*)p + i * (unsigned)sizeof(T));
}
template <class T> __device__ __forceinline__ void st_u32(T* p, unsigned i, T v) {
    *(T*)((char*)p + i * (unsigned)sizeof(T)) = v;
}

template <bool H> struct VelT {
    void* p[3];
    __device__ __forceinline__ float ld(int a, int i) const {
        if constexpr (H) return (float)((const _Float16*)p[a])[i];
        else return ((const float*)p[a])[i];
    }
    __device__ __forceinline__ float ldu(int a, unsigned i) const {
        if constexpr (H) return (float)ld_u32((const _Float16*)p[a], i);
        else return ld_u32((const float*)p[a], i);
    }
    __device__ __forceinline__ void stu(int a, unsigned i, float v) const {
        if constexpr (H) st_u32((_Float16*)p[a], i, (_Float16)v);
        else st_u32((float*)p[a], i, v);
    }
    __device__ __forceinline__ void st(int a, int i, float v) const {
        if constexpr (H) ((_Float16*)p[a])[i] = (_Float16)v;
        else ((float*)p[a])[i] = v;
    }
};
template <bool H> __device__ __forceinline__ VelT<H> vel_typed(const VelPtr& u) { return VelT<H>{{u.p[0], u.p[1], u.p[2]}}; }

// Frame seen by kernels.
struct FrameK {
    int fx, fy, fz;
    int nx, ny, nz;          // padded cell dims = f*B
    const int* slot;         // [fx*fy*fz]
    const int* coord;        // [Nb] linear block index
    const int* nbr;          // [Nb*6]
};

// MI355X deals workgroups to its 8 XCDs round-robin by id, and each XCD has its own L2.  This
// bijection of [0, g) gives each XCD a contiguous range of logical workgroup indices, so blocks that
// are neighbours in storage order (and share halos) run on the same XCD.
__device__ __forceinline__ int xcd_swizzle(int w, int g) {
    constexpr int X = 8;
    const int x = w % X, i = w / X, q = g / X, r = g % X;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

template <int B> struct Blk {
    static constexpr int B2 = B * B;
    static constexpr int B3 = B * B * B;
    static constexpr int SH = (B == 4) ? 2 : 3;
};

// Storage index of cell (i,j,k) (0 <= i < nx ...), or -1 when its block is inactive
// (SparseField::findIndex, SparseField.h:89-95).
template <int B>
__device__ __forceinline__ int cell_index(const FrameK& f, int i, int j, int k) {
    constexpr int SH = Blk<B>::SH;
    const int s = f.slot[((k >> SH) * f.fy + (j >> SH)) * f.fx + (i >> SH)];
    if (s < 0) return -1;
    return (s << (3 * SH)) + ((((k & (B - 1)) << SH) + (j & (B - 1))) << SH) + (i & (B - 1));
}

// Cell coordinates of storage index c (SparseField::findIndices, SparseField.h:103-112).
template <int B>
__device__ __forceinline__ void cell_coords(const FrameK& f, int c, int& i, int& j, int& k) {
    constexpr int SH = Blk<B>::SH;
    const int s = c >> (3 * SH);
    const int l = c & (Blk<B>::B3 - 1);
    const int t = f.coord[s];
    const int bi = t % f.fx, bj = (t / f.fx) % f.fy, bk = t / (f.fx * f.fy);
    i = (bi << SH) + (l & (B - 1));
    j = (bj << SH) + ((l >> SH) & (B - 1));
    k = (bk << SH) + (l >> (2 * SH));
}

// Neighbour of cell c in direction q (0:-x 1:+x 2:-y 3:+y 4:-z 5:+z).  Returns the storage
// index, -1 when the neighbour block is inactive, -2 when the neighbour is outside the domain.
template <int B>
__device__ __forceinline__ int cell_neighbor(const FrameK& f, int c, int q) {
    constexpr int SH = Blk<B>::SH;
    const int s = c >> (3 * SH);
    const int l = c & (Blk<B>::B3 - 1);
    const int li = l & (B - 1), lj = (l >> SH) & (B - 1), lk = l >> (2 * SH);
    const int axis = q >> 1;
    const int dir = (q & 1) ? 1 : -1;
    const int lc = axis == 0 ? li : (axis == 1 ? lj : lk);
    const int stride = axis == 0 ? 1 : (axis == 1 ? B : B * B);
    const int nl = lc + dir;
    if (nl >= 0 && nl < B) return c + dir * stride;
    const int ns = f.nbr[s * 6 + q];
    if (ns < 0) return ns;
    return (ns << (3 * SH)) + l - dir * (B - 1) * stride;
}

// ---- wave / workgroup reductions (wave64) ----
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over a 256-thread workgroup; result valid in every thread.
__device__ __forceinline__ double block_sum(double v, double* sh /*[4]*/) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    const double r = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return r;
}

// Exclusive scan over a 256-thread workgroup of one int per thread; total returned in *tot.
__device__ __forceinline__ int block_exscan(int v, int* sh /*[4]*/, int* tot) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[w] = x;
    __syncthreads();
    int off = 0;
    for (int q = 0; q < w; ++q) off += sh[q];
    *tot = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return off + x - v;
}

}  // namespace vare
