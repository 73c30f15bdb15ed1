// flip_batch.h -- one launch of a step kernel over the workgroups of many tanks.
//
// Every kernel of the step is written as a functor K_f with a static run(vbid, vnbid, args...): vbid
// is the workgroup's index within the launch it would have had on its own and vnbid that launch's
// workgroup count.  The kernel k_x(args) is run(blockIdx.x, gridDim.x, args).  k_batched<K_f, ...>
// runs the same body for a list of tanks in one launch: tank t owns the workgroups
// [starts[t], starts[t+1]) and its arguments are packs[t], so every tank's workgroups see exactly
// the indices, sizes and arguments of its own launch and compute bit for bit what it would alone.
#pragma once
#include <hip/hip_runtime.h>
#include <climits>
#include <cstddef>
#include <cstring>
#include <utility>

namespace vare {

constexpr int kBatchMaxTanks = 32;     // tanks per batched launch (at most 64: one lane of a wave each)

// The kernel arguments of one launch as a plain aggregate: leaf I holds argument I.  The host
// (x86-64) and the device (amdgcn) lay it out alike: both follow the Itanium C++ ABI, and every
// argument type is a scalar, a pointer or a struct of them with the same size and alignment on both.
template <size_t I, class T> struct PackLeaf { T v; };
template <class Seq, class... P> struct PackImpl;
template <size_t... I, class... P> struct PackImpl<std::index_sequence<I...>, P...> : PackLeaf<I, P>... {};
template <class... P> using Pack = PackImpl<std::index_sequence_for<P...>, P...>;
template <size_t I, class T> __host__ __device__ __forceinline__ T& pget(PackLeaf<I, T>& l) { return l.v; }
template <size_t I, class T> __host__ __device__ __forceinline__ const T& pget(const PackLeaf<I, T>& l) {
    return l.v;
}

template <class F, class PK, size_t... I>
__device__ __forceinline__ void call_packed(unsigned vbid, unsigned vnbid, const PK& pk, std::index_sequence<I...>) {
    F::run(vbid, vnbid, pget<I>(pk)...);
}

// Arguments of a merged launch, passed by value: like any kernel argument they live in the kernarg
// segment, so the compiler reads them with scalar loads and knows the pointers in them address
// global memory, exactly as in the tanks' own launches.  NT tanks at most (the segment holds a few KB).
constexpr size_t kBatchArgBytes = 3072;
template <class... P> constexpr int batch_tanks() {
    const size_t n = kBatchArgBytes / sizeof(Pack<P...>);
    return n < 1 ? 1 : (n > (size_t)kBatchMaxTanks ? kBatchMaxTanks : (int)n);
}
template <int NT, class... P> struct BatchArgs {
    int ntank;
    int starts[NT + 1];
    Pack<P...> packs[NT];
};

// The workgroup's tank is the last t with starts[t] <= blockIdx.x: lane t of each wave reads
// starts[t] and a ballot counts them.  LB and WPE repeat the functor's launch bounds and waves per
// SIMD, so the body compiles to the same register budget as its own kernel.
template <class F, int LB, int WPE, int NT, class... P>
__global__ void __launch_bounds__(LB) __attribute__((amdgpu_waves_per_eu(WPE)))
k_batched(const BatchArgs<NT, P...> a) {
    const int lane = (int)(threadIdx.x & 63u);
    const int b = (int)blockIdx.x;
    const int sv = lane < a.ntank ? a.starts[lane] : INT_MAX;
    const int t = __builtin_amdgcn_readfirstlane(__popcll(__ballot(sv <= b)) - 1);
    const int s0 = a.starts[t], s1 = a.starts[t + 1];
    call_packed<F>((unsigned)(b - s0), (unsigned)(s1 - s0), a.packs[t], std::index_sequence_for<P...>{});
}

// Launches k_batched<F, ...> for n <= NT tanks: grid[q] workgroups and the argument pack at packs[q]
// for tank q.  Returns the launch's error.
using IssueFn = hipError_t (*)(dim3 block, unsigned shm, hipStream_t s, int n, const unsigned* grid,
                               const unsigned char* const* packs);
template <class F, int LB, int WPE, class... P>
hipError_t batched_issue(dim3 block, unsigned shm, hipStream_t s, int n, const unsigned* grid,
                         const unsigned char* const* packs) {
    constexpr int NT = batch_tanks<P...>();
    BatchArgs<NT, P...> a;
    a.ntank = n;
    unsigned tot = 0;
    for (int q = 0; q < n; ++q) {
        a.starts[q] = (int)tot;
        tot += grid[q];
        std::memcpy((void*)&a.packs[q], packs[q], sizeof(Pack<P...>));
    }
    a.starts[n] = (int)tot;
    hipLaunchKernelGGL((k_batched<F, LB, WPE, NT, P...>), dim3(tot), block, shm, s, a);
    return hipGetLastError();
}

template <class PK, size_t... I, class... A>
inline void pack_set(PK& pk, std::index_sequence<I...>, A&&... a) {
    ((pget<I>(pk) = std::forward<A>(a)), ...);
}

// One recorded operation of a tank's step: a kernel launch (its batched launcher, workgroup count,
// block shape and argument pack), a device-to-host copy or an event record.
struct ROp {
    enum Kind : int { kKernel = 0, kCopy = 1, kEvent = 2 };
    int kind = kKernel;
    IssueFn issue = nullptr;
    unsigned grid = 0;
    dim3 block;
    unsigned shm = 0;
    unsigned pack_off = 0, pack_len = 0;
    int max_tanks = 1;                // tanks one merged launch of this kernel takes
    void* dst = nullptr;
    const void* src = nullptr;
    size_t bytes = 0;
    hipEvent_t ev = nullptr;
    bool same(const ROp& o) const {
        if (kind != o.kind) return false;
        if (kind != kKernel) return true;
        return issue == o.issue && pack_len == o.pack_len && block.x == o.block.x && block.y == o.block.y && block.z == o.block.z && shm == o.shm;
    }
};

}  // namespace vare
