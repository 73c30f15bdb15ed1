// flip_api.hip -- host runtime + C ABI (include/vare_flip.h) of the gfx950 FLIP step.
//
// One vare_flip handle owns its device state and one HIP stream (default priority).  The step is
// the reference's FLIPMethod::advanceOneStep (base/source/FLIPMethod.cu:208-298) split into six
// stages; each stage enqueues hand-written kernels from flip_kernels.h on the handle's stream.
#include <hip/hip_runtime.h>
#include <cstddef>
#include <hip/hip_ext.h>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include <thread>
#include <atomic>
#include <mutex>
#include <memory>
#include <chrono>
#include <condition_variable>
#include <unordered_map>
#include <array>

#include "../../include/vare_flip.h"
#include "flip_kernels.h"
#include "flip_mg.h"
#include "flip_particles.h"
#include "flip_stencil.h"
#include "flip_post.h"
#include "flip_batch.h"

using namespace vare;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(VARE_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),   \
                        __FILE__, __LINE__);                                                 \
    } while (0)

#define CHK(expr)                       \
    do {                                \
        int rc_ = (expr);               \
        if (rc_ != VARE_OK) return rc_; \
    } while (0)

inline unsigned nblocks(size_t n) { return (unsigned)((n + WG - 1) / WG); }

// Capped grid for the grid-stride reducing kernels: 8 workgroups of 256 threads per CU.
constexpr unsigned kReduceGrid = 2048;
inline unsigned rgrid(size_t n) { return std::max(1u, std::min(nblocks(n), kReduceGrid)); }

}  // namespace

struct vare_flip {
    // geometry (SparseFrame::initialize)
    int B = 4;
    int nx = 0, ny = 0, nz = 0;
    int fx = 0, fy = 0, fz = 0;
    int nblk = 0;
    float t[3] = {0, 0, 0}, s[3] = {1, 1, 1};
    vare_flip_params prm;
    float vox_force[3] = {0, 0, 0};
    float vox_wall = 0.f;
    hipStream_t stream = nullptr;
    int device = 0;               // the device current at create (its stream's device)
    int noalloc = 0;              // inside vare_flip_step_noalloc: scratch from the static array, no growth
    int post_done = 0;            // the cell fields hold the post-process results (until the next stage 0)

    // particles, SoA: px,py,pz,vx,vy,vz,adj in physical order (bin order after each binning),
    // vort,curv in the reference (logical) order; pid[q] = logical index of physical particle q
    size_t np = 0, pcap = 0;
    float* pa[9] = {};            // pa[3..5] hold FP16 values when vhalf != 0
    float* pb[7] = {};            // second buffers of pa[0..6] for the reorder
    int* pid = nullptr;
    int* pid2 = nullptr;
    int vhalf = 0;
    uint8_t* removed = nullptr;   // per physical particle (written by G2P)
    uint8_t* removed_log = nullptr;   // per logical index: the mask of the last step
    int* hole = nullptr;
    int* src = nullptr;
    int* remap = nullptr;         // new logical index of the survivors moved by the last compaction
    int* pkey = nullptr;          // bin (storage index of the particle's base cell, or Nc)
    int* prank = nullptr;         // arrival rank inside the bin
    unsigned long long* perm = nullptr;   // binning: (logical index << 32 | physical index) per bin slot
    int* longbins = nullptr;      // bins longer than 64, [pcap/(kSortReg+1)+2]
    int* midbins = nullptr;       // bins of kSortReg+1 .. 64 particles, same capacity
    int* ccount = nullptr;        // particles per bin, [bcap*B^3+2]
    int* cstart = nullptr;        // bin starts, [bcap*B^3+2]
    size_t last_np_before_remove = 0;
    int nremoved = 0;
    int last_iters[20] = {};      // iterations of each solve (first, then each refinement) of the previous step

    // frame
    int* flag = nullptr;
    int* occ = nullptr;           // [(fx+1)(fy+1)(fz+1)] step stamp of the blocks holding a particle (k_bin_occ)
    int occ_stamp = 0;
    int* slot = nullptr;
    int* coord = nullptr;     // capacity nblk (4 bytes per frame block)
    int* nbr = nullptr;       // capacity bcap*6
    size_t bcap = 0;
    int nb = 0;
    int* bfluid = nullptr;        // [bcap] 1 when the block holds a fluid cell (set by k_divergence)
    int* blist = nullptr;         // [bcap] slots of those blocks, ascending: the solver's blocks
    int ns = 0;                   // number of listed blocks
    int* wgtab = nullptr;         // [wg_table_entries(bcap)*kTab] block and neighbour slots per pair (k_wg_table)

    // cells, capacity bcap*B^3
    float *mass = nullptr, *vx = nullptr, *vy = nullptr, *vz = nullptr, *cnt = nullptr;
    float *cx = nullptr, *cy = nullptr, *cz = nullptr;
    float *x = nullptr, *r = nullptr, *p = nullptr, *ap = nullptr, *b = nullptr, *z = nullptr;
    // second buffers of p and r: the fused iteration kernels read the old vector across block faces
    // while writing the new one, so iteration k of a solve writes p and r to the buffer of parity k
    float *p2 = nullptr, *r2 = nullptr;
    double* x64 = nullptr;        // FP64 pressure of the iterative refinement
    uint16_t* info = nullptr;
    uint8_t* type = nullptr;
    int* rank = nullptr;
    double* partials = nullptr;
    double* partials2 = nullptr;
    float* p2gpart = nullptr;       // P2G: each block's incoming +1-layer partial sums (5 per cell)
    int in_step = 0;                // inside vare_flip_step: the binning leaves the adjacency in place
    int reorder_pending = 0;        // the binning left the move into bin order to P2G's first pass

    // multigrid preconditioner: level 0 aliases r/info/z of the CG; zt0 is its second buffer
    std::vector<MgLevelK> mg;
    std::vector<float*> mg_z, mg_zt;
    float* zt0 = nullptr;
    size_t mg_bcap = 0;

    // scan scratch
    int* tiles = nullptr;
    size_t tiles_cap = 0;
    int* dint = nullptr;          // [0] scan total, [1] removed count, [2] second scan total,
                                  // [3] bin total, [4] bins left unsorted, [5] long-bin count,
                                  // [6] solver block count, [7] mid-bin count
    CgState* st = nullptr;
    CgState* st_host = nullptr;   // pinned mirror
    int* dint_host = nullptr;     // pinned mirror of dint

    // collision source (CollisionSource): block table over the frame + fields of its active cells
    int* col_slot = nullptr;      // [nblk]
    float* col_phi = nullptr;
    float* col_v[3] = {};
    float* col_n[3] = {};
    size_t col_cells = 0;
    float col_bg = 5.f;           // CollisionSource.h:56, the value before the first update
    int col_active = 0;

    // stats + timing
    vare_step_stats stats;
    int stage_next = 0;
    unsigned timing = 0;          // bitmask of VARE_KERNEL_* classes bracketed by events
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> ev_kid;
    std::vector<int64_t> ev_units;
    std::vector<int> ev_iter;     // solver iteration of the timed launch (-1 outside the iterations)
    size_t ev_used = 0;
    int cur_iter = -1;            // iteration being enqueued (enqueue_iteration), -1 otherwise
    double kms[VARE_KERNEL_COUNT] = {};
    int64_t klaunch[VARE_KERNEL_COUNT] = {};
    int64_t kunits[VARE_KERNEL_COUNT] = {};
    hipEvent_t step_ev[4] = {};

    // device memory accounting: every allocation of the handle goes through dalloc/dfree
    std::unordered_map<void*, size_t> allocs;
    size_t dev_bytes = 0;         // live bytes
    size_t peak_bytes = 0;        // most live bytes at any time
    int dry = 0;                  // planning handle (vare_flip_plan_bytes): count, allocate nothing
    uintptr_t dry_next = 0;

    // batched stepping (vare_flip_step_batched): while `rec` is set, launches, readbacks and event
    // records are recorded here and issued, merged with the other tanks' operations, at the next
    // sync point; device frees wait for that flush (a recorded launch may still use the array)
    struct BatchRun* rec = nullptr;
    std::vector<ROp> rops;
    std::vector<unsigned char> rpack;
    std::vector<void*> rfree;
    int rec_rc = 0;

    size_t ncells() const { return (size_t)nb * (size_t)B * B * B; }
    size_t vsz() const { return vhalf ? 2 : 4; }
    void* vptr(int a, size_t off) const { return (char*)pa[3 + a] + off * vsz(); }
    VelPtr vel() const { return VelPtr{{pa[3], pa[4], pa[5]}, vhalf}; }
    PArrays parrays(size_t off = 0) const {
        PArrays r;
        for (int q = 0; q < 9; ++q) r.a[q] = (q >= 3 && q < 6) ? (float*)vptr(q - 3, off) : pa[q] + off;
        r.half = vhalf;
        return r;
    }
    FrameK frame() const { return FrameK{fx, fy, fz, nx, ny, nz, slot, coord, nbr}; }
    CellList clist() const { return CellList{blist, B == 8 ? 9 : 6}; }
    int nvirt() const { return ns * B * B * B; }          // cells of the listed blocks
    size_t nocc() const { return (size_t)(fx + 1) * (fy + 1) * (fz + 1); }
    ColK colk() const {
        return ColK{col_slot, col_phi, {col_v[0], col_v[1], col_v[2]}, {col_n[0], col_n[1], col_n[2]}, col_bg,
                    col_active};
    }
};

namespace {

// Debug hook (vare_debug_inject_alloc_failure): let `skip` more allocations through, then fail `fail`.
std::atomic<long> g_alloc_skip{-1};
std::atomic<long> g_alloc_fail{0};

bool injected_failure() {
    if (g_alloc_skip.load() < 0) return false;
    if (g_alloc_skip.load() > 0) { g_alloc_skip.fetch_sub(1); return false; }
    if (g_alloc_fail.load() <= 0) { g_alloc_skip.store(-1); return false; }
    g_alloc_fail.fetch_sub(1);
    return true;
}

// Every device array of a handle is allocated here, so the handle knows its live and peak bytes.  A
// planning handle (dry) hands out distinct fake addresses that are never dereferenced.
int dalloc_bytes(vare_flip* h, void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 1;
    if (injected_failure()) return fail(VARE_ENOMEM, "hipMalloc of %zu bytes failed: injected failure", bytes);
    if (h->dry) {
        h->dry_next += (bytes + 255) & ~(size_t)255;
        *p = (void*)h->dry_next;
    } else {
        hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            *p = nullptr;
            return fail(VARE_ENOMEM, "hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
        }
    }
    h->allocs[*p] = bytes;
    h->dev_bytes += bytes;
    h->peak_bytes = std::max(h->peak_bytes, h->dev_bytes);
    return VARE_OK;
}
void dfree_bytes(vare_flip* h, void* p) {
    if (!p) return;
    auto it = h->allocs.find(p);
    if (it != h->allocs.end()) {
        h->dev_bytes -= it->second;
        h->allocs.erase(it);
    }
    if (h->dry) return;
    if (h->rec) h->rfree.push_back(p);
    else (void)hipFree(p);
}
template <class T> int dalloc(vare_flip* h, T** p, size_t n) {
    void* v = nullptr;
    const int rc = dalloc_bytes(h, &v, (n ? n : 1) * sizeof(T));
    *p = (T*)v;
    return rc;
}
template <class T> void dfree(vare_flip* h, T*& p) {
    dfree_bytes(h, (void*)p);
    p = nullptr;
}

// ---- launches, readbacks and syncs of the step: direct on the handle's stream, or recorded while
// the handle steps in a batch ----
int batch_sync(vare_flip* h);
int batch_chunk_hint(const vare_flip* h, int k);

template <class F, class... P, class... A>
void record_kernel(vare_flip* h, dim3 grid, dim3 block, size_t shm, A&&... a) {
    if (grid.x == 0) return;
    Pack<P...> pk;
    pack_set(pk, std::index_sequence_for<P...>{}, std::forward<A>(a)...);
    ROp op;
    op.kind = ROp::kKernel;
    op.issue = &batched_issue<F, F::kLB, F::kWPE, P...>;
    op.grid = grid.x;
    op.block = block;
    op.shm = (unsigned)shm;
    op.pack_off = (unsigned)h->rpack.size();
    op.pack_len = (unsigned)sizeof pk;
    op.max_tanks = batch_tanks<P...>();
    const unsigned char* b = (const unsigned char*)&pk;
    h->rpack.insert(h->rpack.end(), b, b + sizeof pk);
    h->rops.push_back(op);
}

// Launch of kernel `kern` (functor F) with one-dimensional grid `grid`.
template <class F, class... P, class... A>
void vlaunch(vare_flip* h, void (*kern)(P...), dim3 grid, dim3 block, size_t shm, A&&... a) {
    if (!h->rec) {
        hipLaunchKernelGGL(kern, grid, block, shm, h->stream, std::forward<A>(a)...);
        return;
    }
    record_kernel<F, P...>(h, grid, block, shm, std::forward<A>(a)...);
}
// The same with start/stop events on the kernel's own dispatch (kernel timing; never set while
// the handle steps in a batch).
template <class F, class... P, class... A>
void vlaunch_ev(vare_flip* h, void (*kern)(P...), dim3 grid, dim3 block, size_t shm, hipEvent_t e0, hipEvent_t e1,
                A&&... a) {
    if (!h->rec) {
        hipExtLaunchKernelGGL(kern, grid, block, shm, h->stream, e0, e1, 0, std::forward<A>(a)...);
        return;
    }
    record_kernel<F, P...>(h, grid, block, shm, std::forward<A>(a)...);
}

hipError_t hevent(vare_flip* h, hipEvent_t e) {
    if (!h->rec) return hipEventRecord(e, h->stream);
    ROp op;
    op.kind = ROp::kEvent;
    op.ev = e;
    h->rops.push_back(op);
    return hipSuccess;
}

// Waits for everything the handle has enqueued.
int hsync(vare_flip* h) {
    if (h->rec) return batch_sync(h);
    HIPCHK(hipStreamSynchronize(h->stream));
    return VARE_OK;
}

// Device-to-host copy of a few scalars, then a sync.
int hread(vare_flip* h, void* dst, const void* src, size_t bytes) {
    if (!h->rec) {
        HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        return VARE_OK;
    }
    ROp op;
    op.kind = ROp::kCopy;
    op.dst = dst;
    op.src = src;
    op.bytes = bytes;
    h->rops.push_back(op);
    return batch_sync(h);
}

void free_mg(vare_flip* h) {
    for (size_t l = 1; l < h->mg.size(); ++l) {
        dfree(h, h->mg[l].r);
        dfree(h, h->mg[l].info);
        dfree(h, h->mg_z[l]);
        dfree(h, h->mg_zt[l]);
    }
    dfree(h, h->zt0);
    h->mg.clear();
    h->mg_z.clear();
    h->mg_zt.clear();
    h->mg_bcap = 0;
}

void free_collision(vare_flip* h) {
    dfree(h, h->col_slot);
    dfree(h, h->col_phi);
    for (int a = 0; a < 3; ++a) { dfree(h, h->col_v[a]); dfree(h, h->col_n[a]); }
    h->col_cells = 0;
    h->col_active = 0;
}

void free_cells(vare_flip* h) {
    free_mg(h);
    dfree(h, h->nbr);
    dfree(h, h->ccount); dfree(h, h->cstart);
    dfree(h, h->bfluid); dfree(h, h->blist); dfree(h, h->wgtab);
    dfree(h, h->mass); dfree(h, h->vx); dfree(h, h->vy); dfree(h, h->vz); dfree(h, h->cnt);
    dfree(h, h->cx); dfree(h, h->cy); dfree(h, h->cz);
    dfree(h, h->x); dfree(h, h->r); dfree(h, h->p); dfree(h, h->ap); dfree(h, h->b); dfree(h, h->z); dfree(h, h->x64);
    dfree(h, h->r2); dfree(h, h->p2);
    dfree(h, h->info); dfree(h, h->type); dfree(h, h->rank);
    dfree(h, h->partials); dfree(h, h->partials2); dfree(h, h->p2gpart);
    h->bcap = 0;
}

int alloc_cells(vare_flip* h, size_t cap);

// A failed allocation leaves no cell arrays (capacity 0), so the next call allocates again.
int ensure_blocks(vare_flip* h, size_t nb) {
    if (nb <= h->bcap && h->mass) return VARE_OK;
    if (h->noalloc)
        return fail(VARE_ENOMEM, "%zu active blocks exceed the block capacity %zu (the no-allocation step never grows it)",
                    nb, h->bcap);
    size_t cap = std::max(nb, h->bcap + h->bcap / 4);
    cap = std::min(cap, (size_t)h->nblk);
    cap = std::max(cap, nb);
    free_cells(h);
    const int rc = alloc_cells(h, cap);
    if (rc != VARE_OK) free_cells(h);
    return rc;
}

int alloc_cells(vare_flip* h, size_t cap) {
    const size_t nc = cap * (size_t)h->B * h->B * h->B;
    const size_t npart = std::max((size_t)kReduceGrid, cap) + 16;
    CHK(dalloc(h, &h->nbr, cap * 6));
    CHK(dalloc(h, &h->ccount, nc + 2)); CHK(dalloc(h, &h->cstart, nc + 2));
    CHK(dalloc(h, &h->bfluid, cap + 1)); CHK(dalloc(h, &h->blist, cap + 1));
    CHK(dalloc(h, &h->wgtab, (size_t)wg_table_entries((int)cap) * kTab));
    CHK(dalloc(h, &h->mass, nc)); CHK(dalloc(h, &h->vx, nc)); CHK(dalloc(h, &h->vy, nc)); CHK(dalloc(h, &h->vz, nc));
    CHK(dalloc(h, &h->cnt, nc));
    CHK(dalloc(h, &h->cx, nc)); CHK(dalloc(h, &h->cy, nc)); CHK(dalloc(h, &h->cz, nc));
    CHK(dalloc(h, &h->x, nc)); CHK(dalloc(h, &h->r, nc)); CHK(dalloc(h, &h->p, nc)); CHK(dalloc(h, &h->ap, nc));
    CHK(dalloc(h, &h->r2, nc)); CHK(dalloc(h, &h->p2, nc));
    CHK(dalloc(h, &h->b, nc)); CHK(dalloc(h, &h->z, nc)); CHK(dalloc(h, &h->x64, nc));
    CHK(dalloc(h, &h->info, nc)); CHK(dalloc(h, &h->type, nc)); CHK(dalloc(h, &h->rank, nc));
    CHK(dalloc(h, &h->partials, npart)); CHK(dalloc(h, &h->partials2, npart));
    const size_t b1 = (size_t)h->B + 1;
    CHK(dalloc(h, &h->p2gpart, cap * (b1 * b1 * b1 - (size_t)h->B * h->B * h->B) * 5));
    h->bcap = cap;
    return VARE_OK;
}

int ensure_particles(vare_flip* h, size_t n) {
    if (n <= h->pcap && h->pa[0]) return VARE_OK;
    // kernels address particle arrays by 32-bit byte offsets: at most 2^30 particles
    constexpr size_t kMaxParticles = (size_t)1 << 30;
    if (n > kMaxParticles) return fail(VARE_ENOMEM, "%zu particles exceed the limit of 2^30", n);
    size_t cap = std::max(n, h->pcap + h->pcap / 2);
    cap = std::min(std::max<size_t>(cap, 1024), kMaxParticles);
    // Every new array is allocated before an old one is released, so a failed allocation leaves the
    // handle with its old arrays, capacity and particles.
    struct Arr { void* member; size_t bytes; size_t keep; };     // keep: live bytes to copy over
    const size_t vs = h->vsz();
    std::vector<Arr> arrs;
    for (int q = 0; q < 9; ++q) {
        const size_t esz = (q >= 3 && q < 6) ? vs : sizeof(float);
        arrs.push_back({&h->pa[q], cap * esz, h->np * esz});
    }
    for (int q = 0; q < 7; ++q) arrs.push_back({&h->pb[q], cap * ((q >= 3 && q < 6) ? vs : sizeof(float)), 0});
    arrs.push_back({&h->pid, cap * sizeof(int), h->np * sizeof(int)});
    arrs.push_back({&h->pid2, cap * sizeof(int), 0});
    arrs.push_back({&h->removed, cap, 0});
    arrs.push_back({&h->removed_log, cap + 4, 0});
    for (int** m : {&h->hole, &h->src, &h->remap, &h->pkey, &h->prank}) arrs.push_back({m, cap * sizeof(int), 0});
    arrs.push_back({&h->perm, cap * sizeof(unsigned long long), 0});
    arrs.push_back({&h->longbins, (cap / (kSortReg + 1) + 2) * sizeof(int), 0});
    arrs.push_back({&h->midbins, (cap / (kSortReg + 1) + 2) * sizeof(int), 0});
    std::vector<void*> fresh(arrs.size(), nullptr);
    auto release = [&]() { for (void*& f : fresh) { dfree_bytes(h, f); f = nullptr; } };
    for (size_t i = 0; i < arrs.size(); ++i) {
        const int rc = dalloc_bytes(h, &fresh[i], arrs[i].bytes);
        if (rc != VARE_OK) { release(); return rc; }
    }
    if (!h->dry && h->np) {
        for (size_t i = 0; i < arrs.size(); ++i) {
            if (!arrs[i].keep) continue;
            void* old;
            std::memcpy(&old, arrs[i].member, sizeof old);
            if (hipMemcpyAsync(fresh[i], old, arrs[i].keep, hipMemcpyDeviceToDevice, h->stream) != hipSuccess) {
                (void)hipStreamSynchronize(h->stream);
                release();
                return fail(VARE_EHIP, "particle array copy failed");
            }
        }
        if (hipStreamSynchronize(h->stream) != hipSuccess) { release(); return fail(VARE_EHIP, "particle array copy failed"); }
    }
    for (size_t i = 0; i < arrs.size(); ++i) {
        void* old;
        std::memcpy(&old, arrs[i].member, sizeof old);
        dfree_bytes(h, old);
        std::memcpy(arrs[i].member, &fresh[i], sizeof(void*));
    }
    h->pcap = cap;
    return VARE_OK;
}

int ensure_tiles(vare_flip* h, size_t ntiles) {
    if (ntiles <= h->tiles_cap && h->tiles) return VARE_OK;
    if (h->noalloc) return fail(VARE_ENOMEM, "a scan of %zu tiles exceeds the no-allocation scratch", ntiles);
    dfree(h, h->tiles);
    CHK(dalloc(h, &h->tiles, ntiles + 16));
    h->tiles_cap = ntiles + 16;
    return VARE_OK;
}

int fill_u32(vare_flip* h, void* p, uint32_t v, size_t n) {
    if (n == 0) return VARE_OK;
    vlaunch<k_fill_u32_f>(h, k_fill_u32, dim3(nblocks(n)), dim3(WG), 0, (uint32_t*)p, v, n);
    HIPCHK(hipGetLastError());
    return VARE_OK;
}

// Three-phase scan; the total lands in dint[total_slot].
template <class Op> int run_scan(vare_flip* h, const Op& op, int n, int total_slot) {
    if (n <= 0) return fill_u32(h, h->dint + total_slot, 0, 1);
    const size_t ntiles = ((size_t)n + SCAN_TILE - 1) / SCAN_TILE;
    CHK(ensure_tiles(h, ntiles));
    vlaunch<k_scan_reduce_f<Op>>(h, k_scan_reduce<Op>, dim3((unsigned)ntiles), dim3(WG), 0, op, n, h->tiles);
    vlaunch<k_scan_tiles_f>(h, k_scan_tiles, dim3(1), dim3(WG), 0, h->tiles, (int)ntiles, h->dint + total_slot);
    vlaunch<k_scan_apply_f<Op>>(h, k_scan_apply<Op>, dim3((unsigned)ntiles), dim3(WG), 0, op, n, (const int*)h->tiles);
    HIPCHK(hipGetLastError());
    return VARE_OK;
}

// ---- kernel timing (HIP events around launches of one class on the handle's stream) ----
bool tgrow(vare_flip* h) {
    if (h->ev_used + 2 <= h->ev_pool.size()) return true;
    for (int q = 0; q < 64; ++q) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return false;
        h->ev_pool.push_back(e);
        h->ev_kid.push_back(-1);
        h->ev_units.push_back(0);
        h->ev_iter.push_back(-1);
    }
    return true;
}
void tbegin(vare_flip* h, int kid) {
    if (!(h->timing & (1u << kid))) return;
    if (!tgrow(h)) return;
    h->ev_kid[h->ev_used] = kid;
    h->ev_units[h->ev_used] = 0;
    h->ev_iter[h->ev_used] = -1;
    (void)hipEventRecord(h->ev_pool[h->ev_used], h->stream);
    h->ev_used++;
}
// Start/stop events for a launch through hipExtLaunchKernelGGL, which records them on the kernel's
// own dispatch (its execution time alone); `units` = work items of the launch (cells or
// particles).  Both null when the class is not timed.
void tpair(vare_flip* h, int kid, int64_t units, hipEvent_t& a, hipEvent_t& b) {
    a = b = nullptr;
    if (!(h->timing & (1u << kid)) || !tgrow(h)) return;
    h->ev_kid[h->ev_used] = kid;
    h->ev_units[h->ev_used] = units;
    h->ev_iter[h->ev_used] = h->cur_iter;
    a = h->ev_pool[h->ev_used++];
    h->ev_kid[h->ev_used] = -2;
    b = h->ev_pool[h->ev_used++];
}
void tend(vare_flip* h, int kid, int64_t units = 0) {
    if (!(h->timing & (1u << kid))) return;
    if (h->ev_used == 0 || h->ev_kid[h->ev_used - 1] != kid) return;
    h->ev_units[h->ev_used - 1] = units;
    h->ev_kid[h->ev_used] = -2;
    (void)hipEventRecord(h->ev_pool[h->ev_used], h->stream);
    h->ev_used++;
}
// After a solve of `iters` iterations: drop the timed launches of iterations enqueued past its end
// (they return at the device-side done flag).  The iteration that ends the solve runs its A*p and
// update kernels but not its V-cycle.  Events from index `from` on belong to this solve.
void tdrop_noop(vare_flip* h, size_t from, int iters) {
    for (size_t q = from; q + 1 < h->ev_used; q += 2) {
        const int k = h->ev_iter[q], kid = h->ev_kid[q];
        if (k < 0 || kid < 0) continue;
        const bool ran = kid == VARE_KERNEL_MG0 ? k < iters - 1 : k < iters;
        if (!ran) h->ev_kid[q] = -3;
    }
}

int tcollect(vare_flip* h) {
    if (h->ev_used == 0) return VARE_OK;
    HIPCHK(hipStreamSynchronize(h->stream));
    for (size_t q = 0; q + 1 < h->ev_used; q += 2) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, h->ev_pool[q], h->ev_pool[q + 1]));
        const int kid = h->ev_kid[q];
        if (kid >= 0 && kid < VARE_KERNEL_COUNT) { h->kms[kid] += ms; h->klaunch[kid]++; h->kunits[kid] += h->ev_units[q]; }
    }
    h->ev_used = 0;
    return VARE_OK;
}

bool finite3(const float* a, size_t n) {
    for (size_t i = 0; i < 3 * n; ++i)
        if (!std::isfinite(a[i])) return false;
    return true;
}

G2PParams g2p_params(const vare_flip* h, float dt) {
    G2PParams gp;
    gp.dt = dt;
    // FLIPMethod.cu:99 uses the literal pair 0.95f / 0.05f
    gp.flip = h->prm.flip_ratio;
    gp.pic = h->prm.flip_ratio == 0.95f ? 0.05f : 1.f - h->prm.flip_ratio;
    gp.wall = h->vox_wall;
    gp.gx = h->vox_force[0]; gp.gy = h->vox_force[1]; gp.gz = h->vox_force[2];
    gp.max_substeps = h->prm.max_substeps;
    gp.tank = h->prm.enable_water_tank;
    gp.band = h->prm.damping_band / h->prm.voxel_size;
    return gp;
}

struct k_cg_setup_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, CgState* st, int mode, int max_count, float abs_tol, float rel_tol, int precond) {
        if (threadIdx.x != 0) return;
        st->mode = mode;
        st->max_count = max_count;
        st->abs_tol = abs_tol;
        st->rel_tol = rel_tol;
        st->precond = precond;
        st->done = 0;
        st->iter = 0;
        st->count = 0;
        st->converged = 0;
        st->ir_rounds = 0;
        st->cont = 0;
    }
};
__global__ void k_cg_setup(CgState* st, int mode, int max_count, float abs_tol, float rel_tol, int precond) {
    k_cg_setup_f::run(blockIdx.x, gridDim.x, st, mode, max_count, abs_tol, rel_tol, precond);
}

// Binning: bin every particle by the storage index of its P2G base cell (bin Nc for particles
// outside the domain), counting sort, then sort each bin by logical index so the order equals a
// stable counting sort of the reference order, and reorder the physical arrays into it.
constexpr int kSortLongGrid = 512;   // workgroups of k_sort_long (strides over the long-bin list)

template <int B> int bin_particles(vare_flip* h) {
    const int nc = (int)h->ncells();
    h->reorder_pending = 0;
    CHK(fill_u32(h, h->ccount, 0, (size_t)nc + 2));
    CHK(fill_u32(h, h->dint + 4, 0, 2));           // [4] sort overflow, [5] long-bin count
    CHK(fill_u32(h, h->dint + 7, 0, 1));           // [7] mid-bin count
    if (h->np)
        vlaunch<k_ckey_f<B>>(h, k_ckey<B>, dim3(nblocks((h->np + kKeySlices - 1) / kKeySlices)), dim3(WG), 0,
                           (const float*)h->pa[0],
                           (const float*)h->pa[1], (const float*)h->pa[2], (int)h->np, h->frame(), nc, h->pkey, h->prank,
                           h->ccount);
    OpBuckets op{h->ccount, h->cstart};
    CHK(run_scan(h, op, nc + 2, 3));
    if (h->np) {
        vlaunch<k_pscatter_f>(h, k_pscatter, dim3(nblocks((h->np + kBinSlices - 1) / kBinSlices)), dim3(WG), 0,
                           (int)h->np, (const int*)h->pkey,
                           (const int*)h->prank, (const int*)h->cstart, (const int*)h->pid, h->perm);
        vlaunch<k_sort_block_f<B>>(h, k_sort_block<B>, dim3(h->nb + 1), dim3(WG), 0, (const int*)h->cstart, (int)h->nb,
                           nc, h->perm, h->midbins, h->dint + 7, h->longbins, h->dint + 5);
        vlaunch<k_sort_wave_f>(h, k_sort_wave, dim3(kSortLongGrid * 4), dim3(WG), 0, (const int*)h->cstart,
                           (const int*)h->midbins, (const int*)(h->dint + 7), h->perm);
        vlaunch<k_sort_long_f>(h, k_sort_long, dim3(kSortLongGrid), dim3(1024), 0, (const int*)h->cstart,
                           (const int*)h->longbins, (const int*)(h->dint + 5), h->perm, h->dint + 4);
        // move the particles into bin order (double-buffered) and keep their logical indices.
        // Inside vare_flip_step P2G's first pass does it as it stages each block's particles (and the
        // adjacency, which G2P rewrites for every particle before anything reads it, is not moved)
        if (h->in_step && h->prm.particle_path == 0) {
            h->reorder_pending = 1;
            HIPCHK(hipGetLastError());
            return VARE_OK;
        }
        ReorderArrays ra;
        for (int a = 0; a < 7; ++a) { ra.src[a] = h->pa[a]; ra.dst[a] = h->pb[a]; }
        ra.vsz = (int)h->vsz();
        ra.n = h->in_step ? 6 : 7;
        vlaunch<k_reorder_f>(h, k_reorder, dim3(nblocks(h->np)), dim3(WG), 0, (int)h->np,
                           (const unsigned long long*)h->perm, ra, h->pid2);
        for (int a = 0; a < ra.n; ++a) std::swap(h->pa[a], h->pb[a]);
        std::swap(h->pid, h->pid2);
    }
    HIPCHK(hipGetLastError());
    return VARE_OK;
}

// ---- stages ----
template <int B> int stage_prepare(vare_flip* h) {
    CHK(fill_u32(h, h->flag, 0, (size_t)h->nblk));
    tbegin(h, VARE_KERNEL_BIN);
    if (h->np) {
        h->occ_stamp = h->occ_stamp == INT_MAX ? 1 : h->occ_stamp + 1;
        if (h->occ_stamp == 1) CHK(fill_u32(h, h->occ, 0, h->nocc()));
        vlaunch<k_bin_occ_f<B>>(h, k_bin_occ<B>, dim3(nblocks((h->np + kOccSlices - 1) / kOccSlices)), dim3(WG), 0,
                           h->pa[0], h->pa[1], h->pa[2],
                           (int)h->np, h->fx, h->fy, h->fz, h->occ_stamp, h->occ);
        vlaunch<k_bin_dilate_f>(h, k_bin_dilate, dim3(nblocks(h->nocc())), dim3(WG), 0, h->fx, h->fy, h->fz,
                           h->occ_stamp, (const int*)h->occ, h->flag);
    }
    tend(h, VARE_KERNEL_BIN);
    if (h->prm.enable_water_tank) {
        const int tall = std::min(std::max((int)(h->vox_wall) / B, 0), h->fy);   // SparseFrame.cu:165
        vlaunch<k_bin_tank_f>(h, k_bin_tank, dim3(nblocks(h->nblk)), dim3(WG), 0, h->fx, h->fy, h->fz, tall, h->flag);
    }
    HIPCHK(hipGetLastError());
    OpBlocks op{h->flag, h->slot, h->coord};
    CHK(run_scan(h, op, h->nblk, 0));
    CHK(hread(h, h->dint_host, h->dint, sizeof(int)));
    const int nb = h->dint_host[0];
    CHK(ensure_blocks(h, (size_t)nb));
    h->nb = nb;
    if (nb)
        vlaunch<k_block_nbr_f>(h, k_block_nbr, dim3(nblocks((size_t)nb * 6)), dim3(WG), 0, nb, h->fx, h->fy, h->fz,
                           (const int*)h->coord, (const int*)h->slot, h->nbr);
    HIPCHK(hipGetLastError());
    if (h->prm.particle_path == 0) CHK(bin_particles<B>(h));
    return VARE_OK;
}

template <int B> int stage_p2g(vare_flip* h) {
    const size_t nc = h->ncells();
    CellFields cf{h->mass, h->vx, h->vy, h->vz, h->cnt};
    if (h->prm.particle_path == 0) {
        // two passes over the blocks: own particles into the block and its +1 layer, then the face
        // cells; every cell of every active block is written, normalised and classified
        const bool ro = h->reorder_pending != 0;
        P2GReorder rk{};
        if (ro) {
            rk.perm = (const unsigned long long*)h->perm;
            for (int a = 0; a < 3; ++a) rk.pos[a] = h->pb[a];
            rk.vel = VelPtr{{h->pb[3], h->pb[4], h->pb[5]}, h->vhalf};
            rk.pid = h->pid2;
        }
        tbegin(h, VARE_KERNEL_P2G);
        if (h->nb) {
            const FrameK fk = h->frame();
#define VARE_P2G_OWN(VH, RO)                                                                                    \
    vlaunch<k_p2g_own_f<B, VH, RO>>(h, k_p2g_own<B, VH, RO>, dim3(h->nb), dim3(P2GOwn<B>::THREADS), 0,            \
                                    (const float*)h->pa[0], (const float*)h->pa[1], (const float*)h->pa[2], h->vel(), \
                                    (const int*)h->cstart, fk, cf, h->cx, h->cy, h->cz, h->type, h->colk(),         \
                                    h->prm.col_vel_scale, h->p2gpart, rk)
            if (h->vhalf) { if (ro) VARE_P2G_OWN(true, true); else VARE_P2G_OWN(true, false); }
            else { if (ro) VARE_P2G_OWN(false, true); else VARE_P2G_OWN(false, false); }
#undef VARE_P2G_OWN
            vlaunch<k_p2g_faces_f<B>>(h, k_p2g_faces<B>, dim3(h->nb), dim3(P2GOwn<B>::FACE_THREADS), 0, fk, cf, h->cx, h->cy,
                               h->cz, h->type, h->colk(), h->prm.col_vel_scale, (const float*)h->p2gpart);
        }
        // algorithmic bytes of both passes: positions and velocities read once (24 B per particle, 18 with
        // FP16 velocities), mass, velocity, count, projected velocity and type written per cell (33 B);
        // moving the particles into bin order adds the permutation read (8 B) and the positions,
        // velocities and logical index written (28 B, 22 with FP16 velocities)
        const int64_t pb = (h->vhalf ? 18 : 24) + (ro ? (h->vhalf ? 30 : 36) : 0);
        tend(h, VARE_KERNEL_P2G, (int64_t)h->np * pb + (int64_t)nc * 33);
        if (ro) {
            // the particles outside the domain (the last bin) are moved by a small strided kernel
            ReorderArrays ra;
            for (int a = 0; a < 7; ++a) { ra.src[a] = h->pa[a]; ra.dst[a] = h->pb[a]; }
            ra.vsz = (int)h->vsz();
            ra.n = 6;
            vlaunch<k_reorder_tail_f>(h, k_reorder_tail, dim3(64), dim3(WG), 0, (const int*)h->cstart + nc, (int)h->np,
                               (const unsigned long long*)h->perm, ra, h->pid2);
            for (int a = 0; a < 6; ++a) std::swap(h->pa[a], h->pb[a]);
            std::swap(h->pid, h->pid2);
            h->reorder_pending = 0;
        }
        HIPCHK(hipGetLastError());
        return VARE_OK;
    }
    CHK(fill_u32(h, h->mass, 0, nc)); CHK(fill_u32(h, h->vx, 0, nc)); CHK(fill_u32(h, h->vy, 0, nc));
    CHK(fill_u32(h, h->vz, 0, nc)); CHK(fill_u32(h, h->cnt, 0, nc));
    tbegin(h, VARE_KERNEL_P2G);
    if (h->np)
        vlaunch<k_p2g_atomic_f<B>>(h, k_p2g_atomic<B>, dim3(nblocks(h->np)), dim3(WG), 0, h->pa[0], h->pa[1], h->pa[2],
                           h->vel(), (int)h->np, h->frame(), cf);
    tend(h, VARE_KERNEL_P2G);
    if (nc)
        vlaunch<k_cell_finalize_f<B>>(h, k_cell_finalize<B>, dim3(nblocks(nc)), dim3(WG), 0, (int)nc, cf, h->cx, h->cy,
                           h->cz, h->type, h->frame(), h->colk(), h->prm.col_vel_scale);
    HIPCHK(hipGetLastError());
    return VARE_OK;
}

CgVecs vecs(vare_flip* h) { return CgVecs{h->x, h->r, h->p, h->ap, h->b, h->z, h->info}; }

int mg_setup(vare_flip* h);
void enqueue_vcycle(vare_flip* h, double* rz = nullptr, float* r0 = nullptr, bool skip_first = false,
                    XTerm xt = XTerm{});

// z = M^-1 r (multigrid: already in z), p = z and r.z over the solver's blocks.
void enqueue_precond_init(vare_flip* h) {
    const int nv = h->nvirt();
    const unsigned gs = rgrid((size_t)nv);
    if (nv)
        vlaunch<k_cg_precond_init_f>(h, k_cg_precond_init, dim3(gs), dim3(WG), 0, nv, h->clist(), vecs(h), h->partials,
                           (const CgState*)h->st);
    vlaunch<k_fin_precond_init_f>(h, k_fin_precond_init, dim3(1), dim3(kFinThreads), 0, h->st, (const double*)h->partials,
                       nv ? (int)gs : 0);
}

template <int B> int stage_system(vare_flip* h) {
    const size_t nc = h->ncells();
    const int mode = h->prm.solver_mode == VARE_SOLVER_RELATIVE ? 1 : 0;
    const int precond = mode == 1 ? h->prm.preconditioner : 0;
    const int maxc = mode == 0 ? h->prm.proj_iteration : h->prm.max_iterations;
    vlaunch<k_cg_setup_f>(h, k_cg_setup, dim3(1), dim3(64), 0, h->st, mode, maxc, h->prm.cg_abs_tol,
                       h->prm.cg_rel_tol, precond);
    const unsigned g = rgrid(nc);
    CHK(fill_u32(h, h->bfluid, 0, (size_t)h->nb));
    tbegin(h, VARE_KERNEL_DIV);
    vlaunch<k_divergence_f<B>>(h, k_divergence<B>, dim3(g), dim3(WG), 0, (int)nc, h->frame(), (const uint8_t*)h->type,
                       (const float*)h->cx, (const float*)h->cy, (const float*)h->cz, h->vox_wall, vecs(h), h->partials,
                       h->partials2, h->bfluid);
    tend(h, VARE_KERNEL_DIV);
    vlaunch<k_fin_divergence_f>(h, k_fin_divergence, dim3(1), dim3(kFinThreads), 0, h->st, (const double*)h->partials,
                       (const double*)h->partials2, (int)g);
    // the solver's blocks: those holding a fluid cell
    OpFluidBlocks ofb{h->bfluid, h->blist};
    CHK(run_scan(h, ofb, h->nb, 6));
    CHK(hread(h, h->dint_host + 6, h->dint + 6, sizeof(int)));
    h->ns = h->dint_host[6];
    if (h->ns) {
        const int nent = wg_table_entries(h->ns);
        vlaunch<k_wg_table_f>(h, k_wg_table, dim3(nblocks((size_t)nent * kTab)), dim3(WG), 0,
                           (const int*)h->blist, h->ns, (const int*)h->nbr, nent, h->wgtab);
    }
    if (mode == 1) {
        if (precond == VARE_PRECOND_MG) {
            CHK(mg_setup(h));
            enqueue_vcycle(h);
        }
        enqueue_precond_init(h);
    }
    HIPCHK(hipGetLastError());
    return VARE_OK;
}

// ---- multigrid preconditioner (VARE_PRECOND_MG) ----
// Levels: sparse edges B, B/2, ..., 1 over the active blocks, then dense halvings of the block
// frame until a level has at most kMgCoarseMax cells.  Arrays are sized for the block capacity.
int alloc_mg(vare_flip* h);

// A failed allocation leaves no levels, so the next call allocates again.
int ensure_mg(vare_flip* h) {
    if (!h->mg.empty() && h->mg_bcap == h->bcap) return VARE_OK;
    if (h->noalloc) return fail(VARE_ENOMEM, "multigrid levels not allocated (the no-allocation step never allocates them)");
    free_mg(h);
    const int rc = alloc_mg(h);
    if (rc != VARE_OK) free_mg(h);
    return rc;
}

int alloc_mg(vare_flip* h) {
    const int sh0 = h->B == 8 ? 3 : 2;
    const size_t cap = h->bcap;
    MgLevelK L0{};
    L0.dense = 0; L0.sh = sh0; L0.dx = h->fx; L0.dy = h->fy; L0.dz = h->fz; L0.scale_sh = 0;
    h->mg.push_back(L0);
    h->mg_z.push_back(nullptr);
    h->mg_zt.push_back(nullptr);
    CHK(dalloc(h, &h->zt0, cap << (3 * sh0)));
    // a level joins the list before its arrays are allocated, so free_mg releases a partial level
    auto add_level = [&](const MgLevelK& L, size_t n) -> int {
        h->mg.push_back(L);
        h->mg_z.push_back(nullptr);
        h->mg_zt.push_back(nullptr);
        MgLevelK& M = h->mg.back();
        M.r = nullptr;
        M.info = nullptr;
        CHK(dalloc(h, &M.r, n)); CHK(dalloc(h, &M.info, n));
        CHK(dalloc(h, &h->mg_z.back(), n)); CHK(dalloc(h, &h->mg_zt.back(), n));
        return VARE_OK;
    };
    for (int sh = sh0 - 1; sh >= 0; --sh) {
        MgLevelK L{};
        L.dense = 0; L.sh = sh; L.dx = h->fx; L.dy = h->fy; L.dz = h->fz; L.scale_sh = sh0 - sh;
        CHK(add_level(L, cap << (3 * sh)));
    }
    int dx = h->fx, dy = h->fy, dz = h->fz, ssh = sh0;
    do {
        dx = (dx + 1) / 2; dy = (dy + 1) / 2; dz = (dz + 1) / 2; ++ssh;
        MgLevelK L{};
        L.dense = 1; L.sh = 0; L.dx = dx; L.dy = dy; L.dz = dz; L.scale_sh = ssh;
        L.ncell = dx * dy * dz;
        CHK(add_level(L, (size_t)L.ncell));
    } while ((size_t)dx * dy * dz > (size_t)kMgCoarseMax);
    h->mg_bcap = h->bcap;
    return VARE_OK;
}

MgFrame mg_frame(const vare_flip* h) { return MgFrame{h->slot, h->coord, h->nbr, h->vox_wall}; }
const int* done_ptr(const vare_flip* h) { return &h->st->done; }

// Per solve: level sizes for the current block count, coarse fluid flags and stencils.
int mg_setup(vare_flip* h) {
    CHK(ensure_mg(h));
    const MgFrame F = mg_frame(h);
    MgLevelK& L0 = h->mg[0];
    L0.ncell = (int)h->ncells();
    L0.r = h->r;
    L0.info = h->info;
    h->mg_z[0] = h->z;
    h->mg_zt[0] = h->zt0;
    for (size_t l = 1; l < h->mg.size(); ++l) {
        MgLevelK& L = h->mg[l];
        if (!L.dense) L.ncell = h->nb << (3 * L.sh);
        if (L.ncell)
            vlaunch<k_mg_flag_f>(h, k_mg_flag, dim3(nblocks(L.ncell)), dim3(WG), 0, h->mg[l - 1], L, F);
        if (L.ncell) vlaunch<k_mg_info_f>(h, k_mg_info, dim3(nblocks(L.ncell)), dim3(WG), 0, L, F);
    }
    HIPCHK(hipGetLastError());
    return VARE_OK;
}

// Blocks per workgroup (K) of the block-aligned kernels, chosen per kernel by measurement: the
// V-cycle passes at edge 8 load few values per cell and run best with K = 4; those at edge 4 and the
// fused CG passes (k_pupdate_spmv, k_update_smooth2) with K = 2.
template <int E> constexpr int mg_block_k() { return E == 8 ? 4 : 2; }
constexpr int kIterK = 2;
constexpr int kSpmvK = 4;
int nwg_k(const vare_flip* h, int K) { return (h->ns + K - 1) / K; }

// Block-aligned kernels for a sparse level of edge E (E^3 cells per slot, slot-major like the fine
// cells): level 0, and for B = 8 also level 1 (edge 4), run over the solver's blocks, mg_block_k<E>()
// per workgroup.  A level-l cell is fluid only if a level-0 cell of its block is, so the listed
// blocks hold every fluid cell of these levels.
// skip_first: the level-0 pass of two Jacobi sweeps from zero has already written mg_z[0]
// (k_update_smooth2), so only the restriction runs.
// Weights of the two sweeps of a smoothing pair: mg_omega_a / mg_omega_b when both are set (two
// sweeps), else mg_omega for both.
struct SweepW { float a, b; };
SweepW sweep_w(const vare_flip* h) {
    const vare_flip_params& p = h->prm;
    if (p.mg_sweeps == 2 && p.mg_omega_a > 0.f && p.mg_omega_b > 0.f) return SweepW{p.mg_omega_a, p.mg_omega_b};
    return SweepW{p.mg_omega, p.mg_omega};
}

template <int E> void enqueue_block_down(vare_flip* h, int l, int nu, bool skip_first = false) {
    const float w = h->prm.mg_omega;
    const SweepW sw = sweep_w(h);
    const int* done = done_ptr(h);
    const MgLevelK& L = h->mg[l];
    const MgLevelK& C = h->mg[l + 1];
    constexpr int K = mg_block_k<E>();
    const unsigned nwg = (unsigned)nwg_k(h, K);
    const int* tab = h->wgtab;
    constexpr unsigned E3 = (unsigned)(E * E * E);
    if (!nwg) return;
    if (nu == 1) {
        vlaunch<k_mg0_restrict_f<E, K>>(h, k_mg0_restrict<E, K>, dim3(nwg), dim3(E3), 0, tab, L, C, w, 1, h->mg_zt[l], done);
    } else {
        if (!skip_first)
            vlaunch<k_mg0_smooth_first2_f<E, K>>(h, k_mg0_smooth_first2<E, K>, dim3(nwg), dim3(E3), 0, tab, L, sw.a, sw.b,
                               h->mg_z[l], done);
        vlaunch<k_mg0_restrict_f<E, K>>(h, k_mg0_restrict<E, K>, dim3(nwg), dim3(E3), 0, tab, L, C, w, 0, h->mg_z[l], done);
    }
}

// rz != nullptr (level 0 only): the last pass also writes per-workgroup partials of r.z.
template <int E> void enqueue_block_up(vare_flip* h, int l, int nu, double* rz) {
    const float w = h->prm.mg_omega;
    const SweepW sw = sweep_w(h);
    const int* done = done_ptr(h);
    const MgLevelK& L = h->mg[l];
    constexpr int K = mg_block_k<E>();
    const unsigned nwg = (unsigned)nwg_k(h, K);
    const int* tab = h->wgtab;
    constexpr unsigned E3 = (unsigned)(E * E * E);
    if (!nwg) return;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (l == 0) tpair(h, VARE_KERNEL_MG0, (int64_t)h->ns * E3, e0, e1);
    if (nu == 1) {
        vlaunch_ev<k_mg0_prolong_smooth_f<E, K>>(h, k_mg0_prolong_smooth<E, K>, dim3(nwg), dim3(E3), 0, e0, e1, tab, L, w,
                              (const float*)h->mg_zt[l], (const float*)h->mg_z[l + 1], h->mg_z[l], done, rz);
    } else {
        vlaunch_ev<k_mg0_prolong_smooth_f<E, K>>(h, k_mg0_prolong_smooth<E, K>, dim3(nwg), dim3(E3), 0, e0, e1, tab, L, sw.a,
                              (const float*)h->mg_z[l], (const float*)h->mg_z[l + 1], h->mg_zt[l], done,
                              (double*)nullptr);
        vlaunch<k_mg0_smooth_f<E, K>>(h, k_mg0_smooth<E, K>, dim3(nwg), dim3(E3), 0, tab, L, sw.b, (const float*)h->mg_zt[l],
                           h->mg_z[l], done, rz);
    }
}

// Down and up passes of level l (l >= the block-kernel levels, above the bottom workgroup) with the
// generic kernels instantiated for the level's kind: two Jacobi sweeps from zero and the restriction
// (nu = 1: the restriction with the one sweep computed inside), then the prolongation fused with the
// first post-sweep and, for nu = 2, the second post-sweep.
template <int KIND> void enqueue_level_down(vare_flip* h, int l, int nu) {
    const MgFrame F = mg_frame(h);
    const float w = h->prm.mg_omega;
    const int* done = done_ptr(h);
    const MgLevelK& L = h->mg[l];
    const MgLevelK& C = h->mg[l + 1];
    const dim3 gr(nblocks((size_t)C.ncell * 8));
    if (nu == 1) {
        if (C.ncell)
            vlaunch<k_mg_restrict_f<KIND, true>>(h, k_mg_restrict<KIND, true>, gr, dim3(WG), 0, L, C, F, w, h->mg_zt[l], done);
        return;
    }
    const SweepW sw = sweep_w(h);
    if (L.ncell)
        vlaunch<k_mg_smooth_first2_f<KIND>>(h, k_mg_smooth_first2<KIND>, dim3(nblocks(L.ncell)), dim3(WG), 0, L, F, sw.a, sw.b,
                           h->mg_z[l], done);
    if (C.ncell)
        vlaunch<k_mg_restrict_f<KIND, false>>(h, k_mg_restrict<KIND, false>, gr, dim3(WG), 0, L, C, F, w, h->mg_z[l], done);
}

template <int KIND> void enqueue_level_up(vare_flip* h, int l, int nu) {
    const MgFrame F = mg_frame(h);
    const float w = h->prm.mg_omega;
    const int* done = done_ptr(h);
    const MgLevelK& L = h->mg[l];
    const MgLevelK& C = h->mg[l + 1];
    const dim3 gr(nblocks(L.ncell));
    if (nu == 1) {
        vlaunch<k_mg_prolong_smooth_f<KIND>>(h, k_mg_prolong_smooth<KIND>, gr, dim3(WG), 0, L, C, F, w,
                           (const float*)h->mg_zt[l], (const float*)h->mg_z[l + 1], h->mg_z[l], done);
        return;
    }
    const SweepW sw = sweep_w(h);
    vlaunch<k_mg_prolong_smooth_f<KIND>>(h, k_mg_prolong_smooth<KIND>, gr, dim3(WG), 0, L, C, F, sw.a, (const float*)h->mg_z[l],
                       (const float*)h->mg_z[l + 1], h->mg_zt[l], done);
    vlaunch<k_mg_smooth_f<KIND>>(h, k_mg_smooth<KIND>, gr, dim3(WG), 0, L, F, sw.b, (const float*)h->mg_zt[l], h->mg_z[l],
                       done);
}

// z = V-cycle(r) into the CG's z vector; with rz != nullptr the last level-0 pass also writes the
// per-workgroup partials of r.z (one per level-0 workgroup).
// r0 != nullptr: the level-0 right-hand side (the CG residual) is read from r0 instead of h->r;
// skip_first: the first level-0 pass is already done (see enqueue_block_down).
// xt.x != nullptr: the extra workgroups of k_mg_bottom apply the CG's x += alpha p (see XTerm).
template <int B> void enqueue_vcycle_t(vare_flip* h, double* rz, float* r0, bool skip_first, XTerm xt) {
    float* const r0_saved = h->mg[0].r;
    if (r0) h->mg[0].r = r0;
    const MgFrame F = mg_frame(h);
    const float w = h->prm.mg_omega;
    const int nu = h->prm.mg_sweeps;
    const int* done = done_ptr(h);
    const int m = (int)h->mg.size() - 1;
    const int nblk_levels = B == 8 ? 2 : 1;                    // levels run by the block kernels
    // levels lb..m run inside k_mg_bottom: every level below the block-kernel levels whose dense view
    // (a sparse level: the frame at its resolution) fits, together at most kMgBottomCells cells
    auto view_cells = [&](const MgLevelK& L) -> long long {
        return L.dense ? (long long)L.ncell : ((long long)h->fx * h->fy * h->fz) << (3 * L.sh);
    };
    int lb = m;
    long long tot = view_cells(h->mg[m]);
    while (lb - 1 >= nblk_levels && tot + view_cells(h->mg[lb - 1]) <= kMgBottomCells &&
           m - lb + 2 <= kMgBottomLevels) {
        --lb;
        tot += view_cells(h->mg[lb]);
    }
    for (int l = 0; l < lb; ++l) {
        const MgLevelK& L = h->mg[l];
        if (l < nblk_levels) {
            if (l == 0) enqueue_block_down<B>(h, 0, nu, skip_first);
            else enqueue_block_down<4>(h, 1, nu);
            continue;
        }
        switch (mg_kind(L)) {
            case kMgBlk: enqueue_level_down<kMgBlk>(h, l, nu); break;
            case kMgEdge1: enqueue_level_down<kMgEdge1>(h, l, nu); break;
            default: enqueue_level_down<kMgDense>(h, l, nu); break;
        }
    }
    MgBottom bt{};
    bt.nlev = m - lb + 1;
    bt.slot = h->slot;
    bt.fx = h->fx;
    bt.fy = h->fy;
    for (int q = 0, off = 0; q < bt.nlev; ++q) {
        const MgLevelK& L = h->mg[lb + q];
        bt.sh[q] = L.dense ? -1 : L.sh;
        bt.dx[q] = L.dense ? L.dx : h->fx << L.sh;
        bt.dy[q] = L.dense ? L.dy : h->fy << L.sh;
        bt.dz[q] = L.dense ? L.dz : h->fz << L.sh;
        bt.off[q] = off;
        bt.info[q] = L.info;
        off += bt.dx[q] * bt.dy[q] * bt.dz[q];
    }
    bt.r0 = h->mg[lb].r;
    bt.z0 = h->mg_z[lb];
    unsigned nxwg = 0;
    if (xt.x && xt.nblk > 0) {
        const int per = (kMgBottomThreads >> xt.sh3) * kXUpdPasses;     // blocks per x-update workgroup
        nxwg = (unsigned)((xt.nblk + per - 1) / per);
    }
    const SweepW sw = sweep_w(h);
    vlaunch<k_mg_bottom_f>(h, k_mg_bottom, dim3(1 + nxwg), dim3(kMgBottomThreads), 0, bt, w, sw.a, sw.b, nu,
                       h->prm.mg_coarse_sweeps, done, xt);
    for (int l = lb - 1; l >= 0; --l) {
        const MgLevelK& L = h->mg[l];
        if (!L.ncell) continue;
        if (l < nblk_levels) {
            if (l == 0) enqueue_block_up<B>(h, 0, nu, rz);
            else enqueue_block_up<4>(h, 1, nu, (double*)nullptr);
            continue;
        }
        switch (mg_kind(L)) {
            case kMgBlk: enqueue_level_up<kMgBlk>(h, l, nu); break;
            case kMgEdge1: enqueue_level_up<kMgEdge1>(h, l, nu); break;
            default: enqueue_level_up<kMgDense>(h, l, nu); break;
        }
    }
    h->mg[0].r = r0_saved;
}

void enqueue_vcycle(vare_flip* h, double* rz, float* r0, bool skip_first, XTerm xt) {
    if (h->B == 4) enqueue_vcycle_t<4>(h, rz, r0, skip_first, xt);
    else enqueue_vcycle_t<8>(h, rz, r0, skip_first, xt);
}

// One CG/PCG iteration (LinearSystemSolver.cpp:29-46), iteration k of the current solve, over the
// solver's blocks; every kernel returns at once when the device-side `done` flag is set, so the host
// can enqueue iterations ahead without syncing.
//   * Iteration k >= 1 starts with the previous iteration's direction update p = z + beta p fused
//     into A p (k_pupdate_spmv); iteration 0 uses the p of the solve's set-up.  p of iteration k
//     lives in the buffer of parity k (p, p2).
//   * With the multigrid preconditioner and two sweeps, the x, r update is fused with the V-cycle's
//     first level-0 pass (k_update_smooth2), and r alternates too: the residual before iteration k
//     is in the buffer of parity k (r, r2).  Otherwise r is updated in place by k_cg_update.
template <int B> void enqueue_iteration_k(vare_flip* h, int k);
template <int B> void enqueue_iteration(vare_flip* h, int k) {
    h->cur_iter = k;
    enqueue_iteration_k<B>(h, k);
    h->cur_iter = -1;
}
template <int B> void enqueue_iteration_k(vare_flip* h, int k) {
    const int nv = h->nvirt();
    if (!nv) return;
    const int nwg_it = nwg_k(h, kIterK);                   // fused CG passes
    const int nwg_sp = nwg_k(h, kSpmvK);                   // plain A p of a solve's first iteration
    const int nwg_v0 = nwg_k(h, mg_block_k<B>());         // the V-cycle's level-0 passes (r.z partials)
    const bool mg = h->prm.solver_mode == VARE_SOLVER_RELATIVE && h->prm.preconditioner == VARE_PRECOND_MG;
    const bool fuse = mg && h->prm.mg_sweeps == 2;
    float* const pbuf[2] = {h->p, h->p2};
    float* const rbuf[2] = {h->r, h->r2};
    float* const pk = pbuf[k & 1];
    float* const rin = fuse ? rbuf[k & 1] : h->r;
    float* const rout = fuse ? rbuf[(k + 1) & 1] : h->r;
    const int* tab = h->wgtab;
    CgVecs v = vecs(h);
    v.r = rin;
    if (k == 0) {
        v.p = pk;
        vlaunch<k_spmv_blk_f<B, kSpmvK>>(h, k_spmv_blk<B, kSpmvK>, dim3(nwg_sp), dim3(B * B * B), 0, tab, v, h->partials,
                           (const CgState*)h->st);
    } else {
        v.p = pbuf[(k + 1) & 1];
        hipEvent_t e0, e1;
        tpair(h, VARE_KERNEL_SPMV, (int64_t)nv, e0, e1);
        const int pc = h->prm.solver_mode == VARE_SOLVER_RELATIVE ? h->prm.preconditioner : VARE_PRECOND_NONE;
        if (pc == VARE_PRECOND_MG)
            vlaunch_ev<k_pupdate_spmv_f<B, kIterK, 2>>(h, k_pupdate_spmv<B, kIterK, 2>, dim3(nwg_it), dim3(B * B * B), 0, e0, e1, tab, v, pk, h->partials, (const CgState*)h->st);
        else if (pc == VARE_PRECOND_JACOBI)
            vlaunch_ev<k_pupdate_spmv_f<B, kIterK, 1>>(h, k_pupdate_spmv<B, kIterK, 1>, dim3(nwg_it), dim3(B * B * B), 0, e0, e1, tab, v, pk, h->partials, (const CgState*)h->st);
        else
            vlaunch_ev<k_pupdate_spmv_f<B, kIterK, 0>>(h, k_pupdate_spmv<B, kIterK, 0>, dim3(nwg_it), dim3(B * B * B), 0, e0, e1, tab, v, pk, h->partials, (const CgState*)h->st);
    }
    vlaunch<k_fin_alpha_f>(h, k_fin_alpha, dim3(1), dim3(kFinThreads), 0, h->st, (const double*)h->partials,
                       k == 0 ? nwg_sp : nwg_it);
    v.p = pk;
    if (fuse) {
        hipEvent_t u0, u1;
        tpair(h, VARE_KERNEL_UPDATE, (int64_t)nv, u0, u1);
        vlaunch_ev<k_update_smooth2_f<B, kIterK>>(h, k_update_smooth2<B, kIterK>, dim3(nwg_it), dim3(B * B * B), 0, u0, u1,
                              tab, v, rout, sweep_w(h).a, sweep_w(h).b, h->mg_z[0], h->partials, (const CgState*)h->st);
        vlaunch<k_fin_update_f>(h, k_fin_update, dim3(1), dim3(kFinThreads), 0, h->st, (const double*)h->partials,
                           (const double*)h->partials2, nwg_it);
        // the rest of the V-cycle on the new residual; its last level-0 pass reduces r.z as well
        // the V-cycle's bottom launch also applies this iteration's x += alpha p (p_k, alpha_k)
        const XTerm xt{h->x, pk, (const float*)((const char*)h->st + offsetof(CgState, alpha)), h->blist,
                       B == 8 ? 9 : 6, h->ns};
        enqueue_vcycle(h, h->partials, rout, true, xt);
        vlaunch<k_fin_beta_f>(h, k_fin_beta, dim3(1), dim3(kFinThreads), 0, h->st, (const double*)h->partials, nwg_v0);
        return;
    }
    const unsigned gs = rgrid((size_t)nv);
    tbegin(h, VARE_KERNEL_UPDATE);
    vlaunch<k_cg_update_f>(h, k_cg_update, dim3(gs), dim3(WG), 0, nv, h->clist(), v, h->partials, h->partials2,
                       (const CgState*)h->st);
    tend(h, VARE_KERNEL_UPDATE);
    vlaunch<k_fin_update_f>(h, k_fin_update, dim3(1), dim3(kFinThreads), 0, h->st, (const double*)h->partials,
                       (const double*)h->partials2, (int)gs);
    if (mg) {
        // the V-cycle's last level-0 pass reduces r.z as well (one partial per workgroup)
        enqueue_vcycle(h, h->partials);
        vlaunch<k_fin_beta_f>(h, k_fin_beta, dim3(1), dim3(kFinThreads), 0, h->st, (const double*)h->partials, nwg_v0);
    }
}

// Run CG/PCG iterations until the device-side `done` flag is set.  The iteration count of a solve
// changes slowly from step to step: the count of the same solve (the k-th of the step) in the
// previous step is enqueued before the first host check, then small chunks; iterations enqueued
// past convergence return at once.  k0 > 0 continues a solve after its first k0 iterations (their
// direction p and r.z kept on the device, the iteration count too).  With `round` (FP64 refinement)
// a refinement round is enqueued behind every chunk; it runs only once the solve is done, so the host
// reads the state once per round, after the round, and the solve's count is the one the round saved.
struct IrRound;
template <int B> void enqueue_ir_round(vare_flip* h, const IrRound& ir);
template <int B> int run_solver(vare_flip* h, int cap, int k, int& total_iters, int k0 = 0,
                                const IrRound* round = nullptr) {
    int& last = h->last_iters[std::min(k, 19)];
    const size_t ev_from = h->ev_used;
    int launched = 0;
    int chunk = last > 0 ? last : 8;
    // in a batch, every tank enqueues the largest count any of them needed for this solve last step,
    // so their iterations stay aligned and share launches (iterations past a tank's convergence
    // return at its done flag)
    if (h->rec) chunk = std::max(chunk, batch_chunk_hint(h, k));
    // the first chunk is enqueued without a check: a solve that is done at its set-up returns at
    // the device flag in every kernel
    const int rounds0 = h->st_host->ir_rounds;
    for (;;) {
        for (int q = 0; q < chunk; ++q) enqueue_iteration<B>(h, k0 + launched + q);
        if (round) enqueue_ir_round<B>(h, *round);
        HIPCHK(hipGetLastError());
        launched += chunk;
        chunk = last > 0 ? std::max(2, last / 8) : std::min(chunk * 2, 64);
        CHK(hread(h, h->st_host, h->st, sizeof(CgState)));
        if (round ? h->st_host->ir_rounds != rounds0 : h->st_host->done != 0) break;
        if (launched > cap + 1) break;                       // defensive: the device flag ends first
    }
    const int count = round ? h->st_host->last_count : h->st_host->count;
    last = count - k0;
    total_iters += count - k0;
    tdrop_noop(h, ev_from, round ? h->st_host->last_iter : h->st_host->iter);
    return VARE_OK;
}

// One FP64 refinement round, gated on the solve's done flag: x64 += x (with the last CG term), the FP64
// residual into both r buffers (x = 0), and k_fin_ir, which stops the refinement at its target or sets
// up the next FP32 solve: continued (a reliable update) or restarted.
struct IrRound {
    XTail tail;
    int cont_allowed;
};
template <int B> void enqueue_ir_round(vare_flip* h, const IrRound& ir) {
    const int nv = h->nvirt();
    const unsigned gv = rgrid((size_t)nv);
    const int* done = done_ptr(h);
    // over the solver's blocks only: the other blocks hold no fluid cell, and x and x64 stay 0 there
    vlaunch<k_ir_accumulate_f>(h, k_ir_accumulate, dim3(gv), dim3(WG), 0, nv, h->clist(), h->x64, (const float*)h->x,
                               h->x, 0, ir.tail, done);
    const int nwg_ir = nwg_k(h, kIterK);
    if (nwg_ir)
        vlaunch<k_ir_residual_blk_f<B, kIterK>>(h, k_ir_residual_blk<B, kIterK>, dim3(nwg_ir), dim3(B * B * B), 0,
                           (const int*)h->wgtab, (const float*)h->b, (const double*)h->x64,
                           (const uint16_t*)h->info, h->r, h->x, h->partials, h->r2, done);
    // the solve continues when at most kContMaxReduction of reduction is left (see stage_solve)
    constexpr float kContMaxReduction = 1e4f;
    vlaunch<k_fin_ir_f>(h, k_fin_ir, dim3(1), dim3(kFinThreads), 0, h->st, (const double*)h->partials, nwg_ir,
                        h->prm.refine_tol, h->prm.cg_rel_tol, ir.cont_allowed, kContMaxReduction);
}

template <int B> int stage_solve(vare_flip* h) {
    const size_t nc = h->ncells();
    const unsigned g = rgrid(nc);
    const int mode = h->prm.solver_mode == VARE_SOLVER_RELATIVE ? 1 : 0;
    const int cap = mode == 0 ? h->prm.proj_iteration + 2 : h->prm.max_iterations;
    HIPCHK(hevent(h, h->step_ev[2]));
    const bool refine = mode == 1 && h->prm.refine_tol > 0.f;
    int total_iters = 0, outer = 0;
    const int nv = h->nvirt();
    const unsigned gv = rgrid((size_t)nv);
    if (refine) CHK(fill_u32(h, h->x64, 0, 2 * nc));
    // the fused multigrid iteration leaves the last iteration's x += alpha p to after the solve (XTail)
    const bool fused = mode == 1 && h->prm.preconditioner == VARE_PRECOND_MG && h->prm.mg_sweeps == 2;
    const XTail tail{fused ? (const float*)h->p : nullptr, (const float*)h->p2, (const CgState*)h->st};
    const IrRound ir{tail, fused ? 1 : 0};
    h->st_host->ir_rounds = 0;               // k_cg_setup cleared the device's count; no round has run
    CHK(run_solver<B>(h, cap, 0, total_iters, 0, refine ? &ir : nullptr));
    if (!refine && fused && nv)
        vlaunch<k_cg_xflush_f>(h, k_cg_xflush, dim3(gv), dim3(WG), 0, nv, h->clist(), h->x, tail);
    if (refine) {
        // FP64 iterative refinement: x64 += d, r = b - A x64 in FP64, then the FP32 solve goes on with r.
        // With the fused multigrid iteration the solve may continue (a reliable update): the true
        // residual replaces the recursive one in the r buffers, the V-cycle gives its z and r.z,
        // beta = r.z / (r.z of the last recursive residual), and the next iteration's direction is
        // z + beta p with the last direction p kept, so the Krylov directions built so far are not
        // thrown away.  That pays when the next solve has a modest reduction left to make (at 256^3 with
        // a 1e-5 target: 12.8 -> 12.1 iterations per step); when it must reduce the residual by more
        // than 1e4 (the 512^3 pool at 1e-8: 14.4 -> 15.3), a fresh start (p = z) converges faster, so
        // each correction solve A d = r starts afresh.  k_fin_ir decides, from the solve's last
        // recursive residual (the true one is within FP32 drift of it).
        for (;;) {
            ++outer;                          // the round behind the last chunk has run
            if (h->st_host->ir_done || outer >= h->prm.refine_max) break;
            const int k0 = h->st_host->iter;  // continuing: the solve's iterations so far; restarting: 0
            if (h->st_host->cont) {
                enqueue_vcycle(h, h->partials, (k0 & 1) ? h->r2 : h->r);
                vlaunch<k_fin_beta_f>(h, k_fin_beta, dim3(1), dim3(kFinThreads), 0, h->st, (const double*)h->partials,
                                      nwg_k(h, mg_block_k<B>()));
            } else {
                if (h->prm.preconditioner == VARE_PRECOND_MG) enqueue_vcycle(h);
                enqueue_precond_init(h);
            }
            CHK(run_solver<B>(h, cap, outer, total_iters, k0, &ir));
        }
    }
    // FP64 recompute of the final residual from the pressure the gradient will use, reported with
    // every solve.  When refining, the last k_ir_residual_blk computed it from the final FP64 pressure
    // (k_fin_ir stored it); otherwise it is recomputed from the FP32 pressure.
    if (!refine) {
        vlaunch<k_true_residual_f<B, float>>(h, k_true_residual<B, float>, dim3(g), dim3(WG), 0, (int)nc, h->frame(),
                           (const float*)h->b, (const float*)h->x, (const uint16_t*)h->info, h->partials);
        vlaunch<k_fin_true_f>(h, k_fin_true, dim3(1), dim3(kFinThreads), 0, h->st, (const double*)h->partials, (int)g);
    }
    // (the refinement loop ends with a read of the state after its last kernel)
    if (!refine) CHK(hread(h, h->st_host, h->st, sizeof(CgState)));
    tbegin(h, VARE_KERNEL_APPLY);
    const bool apply64 = refine && !(h->prm.flags & VARE_FLAG_FP32_APPLY);
    if (nc && refine && !apply64)       // the refined pressure, rounded once, applied in FP32
        vlaunch<k_ir_round_f>(h, k_ir_round, dim3(g), dim3(WG), 0, (int)nc, (const double*)h->x64, h->x);
    if (nc && apply64) {
        // pressure apply in FP64 on the refined pressure; the FP32 field keeps its rounding
        vlaunch<k_press_extrap_blk_f<B, double>>(h, k_press_extrap_blk<B, double>, dim3(h->nb), dim3(B * B * B), 0, h->frame(),
                           (const uint8_t*)h->type, h->x64);
        if (h->ns)
            vlaunch<k_press_grad_blk_f<B, kIterK, double>>(h, k_press_grad_blk<B, kIterK, double>, dim3(nwg_k(h, kIterK)), dim3(B * B * B), 0, (const int*)h->wgtab, h->frame(), (const uint8_t*)h->type,
                               (const double*)h->x64, h->vox_wall, h->cx, h->cy, h->cz);
        vlaunch<k_ir_round_f>(h, k_ir_round, dim3(g), dim3(WG), 0, (int)nc, (const double*)h->x64, h->x);
    } else if (nc) {
        vlaunch<k_press_extrap_blk_f<B, float>>(h, k_press_extrap_blk<B, float>, dim3(h->nb), dim3(B * B * B), 0, h->frame(),
                           (const uint8_t*)h->type, h->x);
        if (h->ns)
            vlaunch<k_press_grad_blk_f<B, kIterK, float>>(h, k_press_grad_blk<B, kIterK, float>, dim3(nwg_k(h, kIterK)), dim3(B * B * B), 0, (const int*)h->wgtab, h->frame(), (const uint8_t*)h->type,
                               (const float*)h->x, h->vox_wall, h->cx, h->cy, h->cz);
    }
    tend(h, VARE_KERNEL_APPLY);
    HIPCHK(hipGetLastError());
    HIPCHK(hevent(h, h->step_ev[3]));
    h->stats.div_ratio = 0.0;
    if (nc && (h->prm.flags & VARE_FLAG_REPORT_DIVERGENCE)) {
        // the reference's divergence of the projected velocity on the same classification
        vlaunch<k_divergence_f<B, true>>(h, k_divergence<B, true>, dim3(g), dim3(WG), 0, (int)nc, h->frame(),
                           (const uint8_t*)h->type, (const float*)h->cx, (const float*)h->cy, (const float*)h->cz,
                           h->vox_wall, vecs(h), h->partials, h->partials2, h->bfluid);
        vlaunch<k_fin_divpost_f>(h, k_fin_divpost, dim3(1), dim3(kFinThreads), 0, h->st, (const double*)h->partials,
                           (int)g);
        HIPCHK(hipGetLastError());
        CgState* hs = h->st_host;
        CHK(hread(h, &hs->div_bb, &h->st->div_bb, sizeof(double)));
        h->stats.div_ratio = hs->ir_bb > 0.0 ? std::sqrt(std::max(hs->div_bb, 0.0) / hs->ir_bb) : 0.0;
    }
    const CgState& s = *h->st_host;
    h->stats.cg_iterations = refine ? total_iters : s.count;
    h->stats.converged = refine ? s.ir_done : s.converged;
    h->stats.residual_norm = std::sqrt(std::max(refine ? s.ir_rr : s.rr, 0.0));
    h->stats.b_norm = std::sqrt(std::max(s.ir_bb, 0.0));
    h->stats.num_fluid = s.nfluid;
    h->stats.refine_steps = outer;
    h->stats.true_rel_residual = s.ir_bb > 0.0 ? std::sqrt(std::max(s.true_rr, 0.0) / s.ir_bb) : 0.0;
    return VARE_OK;
}

template <int B> int stage_g2p(vare_flip* h, float dt) {
    CHK(fill_u32(h, h->dint + 1, 0, 1));
    CHK(fill_u32(h, h->removed_log, 0, (h->np + 3) / 4));
    tbegin(h, VARE_KERNEL_G2P);
    if (h->np && h->prm.particle_path == 0) {
        G2PFields gf{h->cnt, h->cx, h->cy, h->cz, h->vx, h->vy, h->vz};
        const bool col = h->colk().any();
        // collision queries and the velocity element type are compile-time choices of the kernel
        if (col && h->vhalf)
            vlaunch<k_g2p_tile_f<B, true, true>>(h, k_g2p_tile<B, true, true>, dim3(h->nb + 1), dim3(WG), 0, h->pa[0],
                               h->pa[1], h->pa[2], h->vel(), h->pa[6], (const int*)h->cstart, h->nb, h->frame(), gf,
                               g2p_params(h, dt), h->colk(), h->removed, h->dint + 1);
        else if (col)
            vlaunch<k_g2p_tile_f<B, true, false>>(h, k_g2p_tile<B, true, false>, dim3(h->nb + 1), dim3(WG), 0, h->pa[0],
                               h->pa[1], h->pa[2], h->vel(), h->pa[6], (const int*)h->cstart, h->nb, h->frame(), gf,
                               g2p_params(h, dt), h->colk(), h->removed, h->dint + 1);
        else if (h->vhalf)
            vlaunch<k_g2p_tile_f<B, false, true>>(h, k_g2p_tile<B, false, true>, dim3(h->nb + 1), dim3(WG), 0, h->pa[0],
                               h->pa[1], h->pa[2], h->vel(), h->pa[6], (const int*)h->cstart, h->nb, h->frame(), gf,
                               g2p_params(h, dt), h->colk(), h->removed, h->dint + 1);
        else
            vlaunch<k_g2p_tile_f<B, false, false>>(h, k_g2p_tile<B, false, false>, dim3(h->nb + 1), dim3(WG), 0, h->pa[0],
                               h->pa[1], h->pa[2], h->vel(), h->pa[6], (const int*)h->cstart, h->nb, h->frame(), gf,
                               g2p_params(h, dt), h->colk(), h->removed, h->dint + 1);
    } else if (h->np) {
        G2PFields gf{h->cnt, h->cx, h->cy, h->cz, h->vx, h->vy, h->vz};
        vlaunch<k_g2p_f<B>>(h, k_g2p<B>, dim3(nblocks(h->np)), dim3(WG), 0, h->pa[0], h->pa[1], h->pa[2],
                           h->vel(), h->pa[6], (int)h->np, h->frame(), gf, g2p_params(h, dt), h->colk(), h->removed,
                           h->dint + 1);
    }
    tend(h, VARE_KERNEL_G2P);
    HIPCHK(hipGetLastError());
    CHK(hread(h, h->dint_host + 1, h->dint + 1, sizeof(int)));
    h->nremoved = h->dint_host[1];
    h->last_np_before_remove = h->np;
    if (h->nremoved > 0)
        vlaunch<k_removed_logical_f>(h, k_removed_logical, dim3(nblocks(h->np)), dim3(WG), 0, (int)h->np,
                           (const uint8_t*)h->removed, (const int*)h->pid, h->removed_log);
    HIPCHK(hipGetLastError());
    return VARE_OK;
}

// Array::remove (Array.h:995-1024) on the reference order, in closed form: the k-th removed
// logical index below newN (ascending) receives the k-th surviving logical index at or above newN
// (descending).  Vorticity and curvature (logical storage) move accordingly; the physical arrays are
// compacted the same way on their own order, and every surviving particle above newN gets its new
// logical index through `remap`.
int compaction_scans(vare_flip* h, const uint8_t* mask, int n, int newn, int& nmove) {
    OpHoles oh{mask, h->hole};
    CHK(run_scan(h, oh, newn, 0));
    OpSurvivors os{mask, newn, h->src};
    CHK(run_scan(h, os, n - newn, 2));
    CHK(hread(h, h->dint_host, h->dint, 3 * sizeof(int)));
    nmove = h->dint_host[0];
    if (nmove != h->dint_host[2])
        return fail(VARE_EINVAL, "removal compaction mismatch (%d holes, %d survivors)", nmove, h->dint_host[2]);
    return VARE_OK;
}

int stage_remove(vare_flip* h) {
    const int nrem = h->nremoved;
    h->stats.num_removed = nrem;
    if (nrem <= 0) return VARE_OK;
    const int n = (int)h->np;
    const int newn = n - nrem;
    // logical order: vorticity and curvature, and the new index of each moved survivor
    int nmove = 0;
    CHK(compaction_scans(h, h->removed_log, n, newn, nmove));
    if (nmove) {
        MoveSet lg{};
        lg.a[0] = h->pa[7]; lg.a[1] = h->pa[8];
        lg.esz[0] = 4; lg.esz[1] = 4;
        lg.n = 2;
        vlaunch<k_remove_move_f>(h, k_remove_move, dim3(nblocks(nmove)), dim3(WG), 0, nmove, (const int*)h->hole,
                           (const int*)h->src, lg);
        vlaunch<k_remap_build_f>(h, k_remap_build, dim3(nblocks(nmove)), dim3(WG), 0, nmove, newn,
                           (const int*)h->hole, (const int*)h->src, h->remap);
    }
    // physical order: positions, velocities, adjacency and logical indices
    int pmove = 0;
    CHK(compaction_scans(h, h->removed, n, newn, pmove));
    if (pmove) {
        MoveSet ph{};
        for (int a = 0; a < 7; ++a) { ph.a[a] = h->pa[a]; ph.esz[a] = (a >= 3 && a < 6) ? (int)h->vsz() : 4; }
        ph.a[7] = h->pid;
        ph.esz[7] = 4;
        ph.n = 8;
        vlaunch<k_remove_move_f>(h, k_remove_move, dim3(nblocks(pmove)), dim3(WG), 0, pmove, (const int*)h->hole,
                           (const int*)h->src, ph);
    }
    if (nmove && newn)
        vlaunch<k_pid_remap_f>(h, k_pid_remap, dim3(nblocks(newn)), dim3(WG), 0, newn, newn,
                           (const int*)h->remap, h->pid);
    HIPCHK(hipGetLastError());
    h->np = (size_t)newn;
    return VARE_OK;
}

// FLIPMethod.cu:320-363 (water tank), on the host as in the reference: air cells of the surface field
// below the tank level that the flood fill from the level's layer does not reach become liquid (-5).
int tank_flood_fill(vare_flip* h) {
    const size_t nc = h->ncells();
    const int B = h->B, B3 = B * B * B, nx = h->nx, ny = h->ny, nz = h->nz;
    std::vector<float> surf(nc);
    std::vector<int> slot((size_t)h->nblk), coord((size_t)h->nb);
    HIPCHK(hipStreamSynchronize(h->stream));
    if (nc) HIPCHK(hipMemcpy(surf.data(), h->cnt, nc * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(slot.data(), h->slot, slot.size() * sizeof(int), hipMemcpyDeviceToHost));
    if (h->nb) HIPCHK(hipMemcpy(coord.data(), h->coord, coord.size() * sizeof(int), hipMemcpyDeviceToHost));
    auto index = [&](int i, int j, int k) -> long long {
        const int sl = slot[(size_t)(((k / B) * h->fy + j / B) * h->fx + i / B)];
        return sl < 0 ? -1 : (long long)sl * B3 + ((k % B) * B + (j % B)) * B + (i % B);
    };
    auto F = [&](int i, int j, int k) -> float {
        const long long c = index(i, j, k);
        return c < 0 ? kSurfBackground : surf[(size_t)c];
    };
    std::vector<char> scan((size_t)nx * ny * nz, 0);
    auto S = [&](int i, int j, int k) -> char& { return scan[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)]; };
    const int wall = (int)(size_t)h->vox_wall;                   // const size_t wall = _voxelWallLevel
    std::vector<std::array<int, 3>> cells;
    cells.push_back({nx - 1, ny - 1, nz - 1});
    if (wall >= 0 && wall < ny)
        for (int k = 0; k < nz; ++k)
            for (int i = 0; i < nx; ++i)
                if (F(i, wall, k) > 0.f) cells.push_back({i, wall, k});
    while (!cells.empty()) {
        const std::array<int, 3> c = cells.back();
        cells.pop_back();
        S(c[0], c[1], c[2]) = 1;
        if (c[1] > wall) continue;
        static const int d[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
        for (int q = 0; q < 6; ++q) {
            const int i = c[0] + d[q][0], j = c[1] + d[q][1], k = c[2] + d[q][2];
            if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) continue;
            if (F(i, j, k) > 0.f && S(i, j, k) == 0) cells.push_back({i, j, k});
        }
    }
    bool changed = false;
    for (size_t c = 0; c < nc; ++c) {
        const int t = coord[c / B3], l = (int)(c % B3);
        const int i = (t % h->fx) * B + l % B, j = ((t / h->fx) % h->fy) * B + (l / B) % B,
                  k = (t / (h->fx * h->fy)) * B + l / (B * B);
        if (j > wall) continue;
        if (surf[c] > 0.f && S(i, j, k) == 0) { surf[c] = -5.f; changed = true; }
    }
    if (changed) HIPCHK(hipMemcpy(h->cnt, surf.data(), nc * sizeof(float), hipMemcpyHostToDevice));
    return VARE_OK;
}

// FLIPMethod::postProcess (FLIPMethod.cu:300-405) on the voxel-space particles.  Afterwards the cell
// fields hold the level set (count field), the splatted velocity (P2G velocity) and the normals
// (projected velocity); pressure, mass and cell types are 0, as the reference's fields after its frame
// rebuild.
template <int B> int run_post_process(vare_flip* h) {
    h->stage_next = 0;
    CHK(stage_prepare<B>(h));                                   // prepareNextStep, :302
    if (h->prm.particle_path != 0) CHK(bin_particles<B>(h));    // the splat reads the bins
    const size_t nc = h->ncells();
    const unsigned g = rgrid(nc);
    const FrameK fk = h->frame();
    if (nc)
        vlaunch<k_splat_gather_f<B>>(h, k_splat_gather<B>, dim3(nblocks(nc)), dim3(WG), 0, (int)nc, (const float*)h->pa[0],
                           (const float*)h->pa[1], (const float*)h->pa[2], h->vel(), (const int*)h->cstart, fk, h->cnt,
                           h->vx, h->vy, h->vz);
    HIPCHK(hipGetLastError());
    const bool tank = h->prm.enable_water_tank != 0;
    if (tank) CHK(tank_flood_fill(h));                         // :320-363
    // Rasterization::ReinitializeSurface, the pressure field as the second buffer (:365)
    float* cur = h->cnt;
    float* nxt = h->x;
    for (int it = 0; it < h->prm.redist_iteration && nc; ++it) {
        if (tank)
            vlaunch<k_redistance_f<B, true>>(h, k_redistance<B, true>, dim3(g), dim3(WG), 0, (int)nc, fk, (const float*)cur,
                               nxt, h->vox_wall);
        else
            vlaunch<k_redistance_f<B, false>>(h, k_redistance<B, false>, dim3(g), dim3(WG), 0, (int)nc, fk,
                               (const float*)cur, nxt, h->vox_wall);
        std::swap(cur, nxt);
    }
    if (cur != h->cnt && nc)
        HIPCHK(hipMemcpyAsync(h->cnt, cur, nc * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
    if (nc) {
        const Vec3Field velo{h->vx, h->vy, h->vz}, comp{h->cx, h->cy, h->cz};
        vlaunch<k_curl_f<B>>(h, k_curl<B>, dim3(g), dim3(WG), 0, (int)nc, fk, velo, h->cx, h->cy, h->cz);   // :367
        const float dis = powf(1.f - h->prm.vort_pres_rate, h->prm.dt);                                       // :373
        if (h->np)
            vlaunch<k_particle_vorticity_f<B>>(h, k_particle_vorticity<B>, dim3(nblocks(h->np)), dim3(WG), 0, (int)h->np,
                               (const float*)h->pa[0], (const float*)h->pa[1], (const float*)h->pa[2],
                               (const int*)h->pid, fk, comp, h->s[0], h->s[1], h->s[2], dis, h->pa[7]);
        vlaunch<k_normal_gradient_f<B>>(h, k_normal_gradient<B>, dim3(g), dim3(WG), 0, (int)nc, fk, (const float*)h->cnt,
                           kSurfBackground, h->cx, h->cy, h->cz);                                            // :390
        vlaunch<k_divergence_vec_f<B>>(h, k_divergence_vec<B>, dim3(g), dim3(WG), 0, (int)nc, fk, comp, h->x);   // :391
        if (h->np)
            vlaunch<k_particle_curvature_f<B>>(h, k_particle_curvature<B>, dim3(nblocks(h->np)), dim3(WG), 0, (int)h->np,
                               (const float*)h->pa[0], (const float*)h->pa[1], (const float*)h->pa[2],
                               (const int*)h->pid, fk, (const float*)h->x, h->pa[8]);                      // :393-398
        CHK(fill_u32(h, h->x, 0, nc));                                                                       // :402-404
        CHK(fill_u32(h, h->mass, 0, nc));
        vlaunch<k_fill_u32_f>(h, k_fill_u32, dim3(nblocks((nc + 3) / 4)), dim3(WG), 0, (uint32_t*)h->type, 0u,
                           (nc + 3) / 4);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    h->post_done = 1;
    return VARE_OK;
}

template <int B> int run_stage(vare_flip* h, int stage, float dt) {
    switch (stage) {
        case 0: return stage_prepare<B>(h);
        case 1: return stage_p2g<B>(h);
        case 2: return stage_system<B>(h);
        case 3: return stage_solve<B>(h);
        case 4: return stage_g2p<B>(h, dt);
        case 5: return stage_remove(h);
    }
    return fail(VARE_EINVAL, "bad stage %d", stage);
}

int dispatch_stage(vare_flip* h, int stage, float dt) {
    if (stage != h->stage_next)
        return fail(VARE_EINVAL, "stage %d called out of order (expected %d)", stage, h->stage_next);
    if (stage == 0) h->post_done = 0;
    int rc = h->B == 4 ? run_stage<4>(h, stage, dt) : run_stage<8>(h, stage, dt);
    if (rc != VARE_OK) { h->stage_next = 0; return rc; }
    h->stage_next = (stage + 1) % 6;
    if (stage == 0) {
        h->stats.num_blocks = h->nb;
        h->stats.num_cells = (int64_t)h->ncells();
    }
    h->stats.num_particles = (int64_t)h->np;
    return VARE_OK;
}

int validate_params(const vare_flip_params* p) {
    if (p->enable_curl) return fail(VARE_ENOTIMPL, "enable_curl: curl noise is outside the built path");
    if (p->solver_mode != VARE_SOLVER_REFERENCE && p->solver_mode != VARE_SOLVER_RELATIVE)
        return fail(VARE_EINVAL, "bad solver_mode %d", p->solver_mode);
    if (p->preconditioner != VARE_PRECOND_NONE && p->preconditioner != VARE_PRECOND_JACOBI &&
        p->preconditioner != VARE_PRECOND_MG)
        return fail(VARE_EINVAL, "bad preconditioner %d", p->preconditioner);
    if (p->voxel_size <= 0.f) return fail(VARE_EINVAL, "voxel_size must be positive");
    if (p->max_iterations < 1 || p->proj_iteration < 0) return fail(VARE_EINVAL, "bad iteration cap");
    if (p->preconditioner == VARE_PRECOND_MG &&
        (p->mg_sweeps < 1 || p->mg_sweeps > 2 || !(p->mg_omega > 0.f && p->mg_omega < 1.f) || p->mg_coarse_sweeps < 1 ||
         !(p->mg_omega_a >= 0.f && p->mg_omega_a < 2.f) || !(p->mg_omega_b >= 0.f && p->mg_omega_b < 2.f)))
        return fail(VARE_EINVAL, "bad multigrid settings");
    return VARE_OK;
}

// SparseFrame::initialize (SparseFrame.cu:72-117) + Grid::initialize (Grid.h:123-138): pad to a block
// multiple, isotropic cell size h = min(dx, dy, dz), recentred bounding box.
int init_geometry(vare_flip* h, const vare_grid_desc* g) {
    const int B = g->block_size;
    h->B = B;
    h->fx = (int)(g->nx / B + ((g->nx % B) ? 1 : 0));
    h->fy = (int)(g->ny / B + ((g->ny % B) ? 1 : 0));
    h->fz = (int)(g->nz / B + ((g->nz % B) ? 1 : 0));
    const int64_t nblk = (int64_t)h->fx * h->fy * h->fz;
    const int64_t ncell = nblk * B * B * B;
    if (ncell > (int64_t)0x7fffffff) return fail(VARE_EINVAL, "frame of %lld cells exceeds 2^31-1", (long long)ncell);
    h->nblk = (int)nblk;
    h->nx = h->fx * B; h->ny = h->fy * B; h->nz = h->fz * B;
    float cen[3], d[3];
    for (int a = 0; a < 3; ++a) cen[a] = (g->min[a] + g->max[a]) * 0.5f;
    d[0] = (g->max[0] - g->min[0]) / (float)g->nx;
    d[1] = (g->max[1] - g->min[1]) / (float)g->ny;
    d[2] = (g->max[2] - g->min[2]) / (float)g->nz;
    const float hh = std::min(std::min(d[0], d[1]), d[2]);
    const int n3[3] = {h->nx, h->ny, h->nz};
    for (int a = 0; a < 3; ++a) {
        const float l = (float)n3[a] * hh;
        const float mn = cen[a] - l * 0.5f;
        const float mx = mn + l;
        float L = mx - mn;
        if (a == 2 && n3[2] == 1) L = 1.f;     // Grid.h:131
        h->t[a] = mn;
        h->s[a] = L / (float)n3[a];
    }
    return VARE_OK;
}

// Frame tables and solver scalars, allocated once per handle.
int alloc_frame(vare_flip* h) {
    CHK(dalloc(h, &h->flag, (size_t)h->nblk));
    CHK(dalloc(h, &h->occ, h->nocc()));
    CHK(dalloc(h, &h->slot, (size_t)h->nblk));
    CHK(dalloc(h, &h->coord, (size_t)h->nblk));
    CHK(dalloc(h, &h->dint, 8));
    CHK(dalloc(h, &h->st, 1));
    return VARE_OK;
}

// Block capacity reserved at create: the caller's, or the whole frame up to 64M cells.
size_t initial_block_capacity(const vare_flip* h, size_t bcap) {
    const size_t B3 = (size_t)h->B * h->B * h->B;
    const size_t bc = bcap ? std::min(bcap, (size_t)h->nblk) : std::min((size_t)h->nblk, (size_t)(64u << 20) / B3);
    return std::max<size_t>(bc, 1);
}

// Scan tiles the longest scan of a step needs: over the frame's blocks, the bins (cells + 2) or the
// particles (removal compaction).
size_t step_scan_tiles(const vare_flip* h, size_t np) {
    const size_t n = std::max(std::max((size_t)h->nblk, h->bcap * h->B * h->B * h->B + 2), np);
    return (n + SCAN_TILE - 1) / SCAN_TILE;
}

void apply_params(vare_flip* h, const vare_flip_params* p) {
    h->prm = *p;
    for (int a = 0; a < 3; ++a) h->vox_force[a] = p->ext_force[a] / h->s[a];   // FLIPMethod.cu:42
    h->vox_wall = (p->wall_level - h->t[1]) / h->s[1];                        // FLIPMethod.cu:44-46
}

}  // namespace

// Scratch of vare_flip_step_noalloc: the reductions' per-workgroup partials and the scans' tile sums.
// Host code never uses this name as an address; it uses the per-device address that
// hipGetSymbolAddress returns (noalloc_scratch).
static __device__ __align__(256) unsigned char scratch[VARE_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::once_flag g_scratch_once[kMaxDevices];
void* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

// Device address of `scratch` on device `dev`, looked up once per device with that device current.
int noalloc_scratch(int dev, void** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(VARE_EINVAL, "device %d out of range", dev);
    std::call_once(g_scratch_once[dev], [dev]() {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != dev) e = hipSetDevice(dev);
        if (e == hipSuccess) e = hipGetSymbolAddress(&g_scratch_ptr[dev], HIP_SYMBOL(scratch));
        if (cur >= 0 && cur != dev) (void)hipSetDevice(cur);
        g_scratch_err[dev] = e;
    });
    if (g_scratch_err[dev] != hipSuccess || !g_scratch_ptr[dev])
        return fail(VARE_EHIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                    hipGetErrorString(g_scratch_err[dev]));
    *out = g_scratch_ptr[dev];
    return VARE_OK;
}

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

const char* vare_last_error(void) { return g_err.c_str(); }

float vare_rand(uint32_t seed) { return rand01(seed); }

int vare_flip_params_default(vare_flip_params* p) {
    if (!p) return fail(VARE_EINVAL, "null params");
    std::memset(p, 0, sizeof *p);
    p->ext_force[0] = 0.f; p->ext_force[1] = -9.80665f; p->ext_force[2] = 0.f;   // FLIPMethod.h:53-66
    p->max_substeps = 1;
    p->dt = 1.f / 24.f;
    p->wall_level = -1e+30f;
    p->enable_water_tank = 0;
    p->enable_curl = 0;
    p->damping_band = 0.f;
    p->voxel_size = 0.1f;
    p->col_vel_scale = 1.f;
    p->pts_per_cell = 8;
    p->vort_pres_rate = 0.2f;
    p->adjacency_under = 2.f;
    p->proj_iteration = 200;
    p->redist_iteration = 80;
    p->flip_ratio = 0.95f;
    p->solver_mode = VARE_SOLVER_REFERENCE;
    p->cg_abs_tol = 1e-3f;
    p->cg_rel_tol = 1e-5f;
    p->max_iterations = 2000;
    p->preconditioner = VARE_PRECOND_NONE;
    p->mg_sweeps = 2;
    p->mg_omega = 0.8f;
    // the two roots of the degree-2 Chebyshev polynomial on [0.4, 2]: 1 / (1.2 -+ 0.8 / sqrt(2))
    p->mg_omega_a = 1.576514f;
    p->mg_omega_b = 0.566352f;
    p->mg_coarse_sweeps = 20;
    p->refine_tol = 0.f;
    p->refine_max = 8;
    return VARE_OK;
}

int vare_grid_from_voxel_size(float voxel, const float mn[3], const float mx[3], int block_size, vare_grid_desc* out) {
    if (!out || !mn || !mx || !(voxel > 0.f)) return fail(VARE_EINVAL, "bad grid arguments");
    out->nx = (int64_t)(size_t)((mx[0] - mn[0]) / voxel);     // Grid.h:145-147 (truncation)
    out->ny = (int64_t)(size_t)((mx[1] - mn[1]) / voxel);
    out->nz = (int64_t)(size_t)((mx[2] - mn[2]) / voxel);
    for (int a = 0; a < 3; ++a) { out->min[a] = mn[a]; out->max[a] = mx[a]; }
    out->block_size = block_size;
    return VARE_OK;
}

int vare_flip_create(const vare_grid_desc* g, const vare_flip_params* prm, size_t pcap, size_t bcap, vare_flip** out) {
    if (!g || !out) return fail(VARE_EINVAL, "null argument");
    *out = nullptr;
    if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0) return fail(VARE_EINVAL, "grid has zero cells");   // Grid.h:107
    if (g->block_size != 4 && g->block_size != 8) return fail(VARE_EINVAL, "block_size must be 4 or 8");
    vare_flip_params def;
    vare_flip_params_default(&def);
    const vare_flip_params* p = prm ? prm : &def;
    CHK(validate_params(p));
    vare_flip* h = new vare_flip();
    std::memset(&h->stats, 0, sizeof h->stats);
    if (int rc0 = init_geometry(h, g)) { delete h; return rc0; }
    apply_params(h, p);
    h->vhalf = p->velocity_fp16 ? 1 : 0;
    if (hipGetDevice(&h->device) != hipSuccess) { delete h; return fail(VARE_EHIP, "hipGetDevice failed"); }
    if (hipStreamCreate(&h->stream) != hipSuccess) { delete h; return fail(VARE_EHIP, "hipStreamCreate failed"); }
    int rc = alloc_frame(h);
    if (rc == VARE_OK && hipHostMalloc((void**)&h->st_host, sizeof(CgState), 0) != hipSuccess) rc = fail(VARE_EHIP, "hipHostMalloc failed");
    if (rc == VARE_OK && hipHostMalloc((void**)&h->dint_host, 8 * sizeof(int), 0) != hipSuccess) rc = fail(VARE_EHIP, "hipHostMalloc failed");
    for (int q = 0; q < 4 && rc == VARE_OK; ++q)
        if (hipEventCreate(&h->step_ev[q]) != hipSuccess) rc = fail(VARE_EHIP, "hipEventCreate failed");
    if (rc == VARE_OK && pcap) rc = ensure_particles(h, pcap);
    if (rc == VARE_OK) rc = ensure_blocks(h, initial_block_capacity(h, bcap));
    if (rc == VARE_OK && p->solver_mode == VARE_SOLVER_RELATIVE && p->preconditioner == VARE_PRECOND_MG)
        rc = ensure_mg(h);
    if (rc == VARE_OK) {
        rc = fill_u32(h, h->slot, 0xffffffffu, (size_t)h->nblk);
        if (rc == VARE_OK && hipStreamSynchronize(h->stream) != hipSuccess) rc = fail(VARE_EHIP, "init sync failed");
    }
    if (rc != VARE_OK) { vare_flip_destroy(h); return rc; }
    *out = h;
    return VARE_OK;
}

int vare_flip_plan_bytes(const vare_grid_desc* g, const vare_flip_params* prm, size_t pcap, size_t bcap,
                         size_t* bytes) {
    if (!g || !bytes) return fail(VARE_EINVAL, "null argument");
    *bytes = 0;
    if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0) return fail(VARE_EINVAL, "grid has zero cells");
    if (g->block_size != 4 && g->block_size != 8) return fail(VARE_EINVAL, "block_size must be 4 or 8");
    vare_flip_params def;
    vare_flip_params_default(&def);
    const vare_flip_params* p = prm ? prm : &def;
    CHK(validate_params(p));
    // the allocations of vare_flip_create and of the first steps, run on a planning handle that
    // counts bytes and allocates nothing
    vare_flip plan;
    plan.dry = 1;
    CHK(init_geometry(&plan, g));
    apply_params(&plan, p);
    plan.vhalf = p->velocity_fp16 ? 1 : 0;
    CHK(alloc_frame(&plan));
    if (pcap) CHK(ensure_particles(&plan, pcap));
    CHK(ensure_blocks(&plan, initial_block_capacity(&plan, bcap)));
    if (p->solver_mode == VARE_SOLVER_RELATIVE && p->preconditioner == VARE_PRECOND_MG) CHK(ensure_mg(&plan));
    CHK(ensure_tiles(&plan, step_scan_tiles(&plan, plan.pcap)));
    // staging of a host upload (3 floats per particle) or download (7 floats per particle)
    float* staging = nullptr;
    CHK(dalloc(&plan, &staging, 7 * std::max<size_t>(plan.pcap, 1)));
    dfree(&plan, staging);
    *bytes = plan.peak_bytes;
    return VARE_OK;
}

int vare_flip_get_device_bytes(const vare_flip* h, size_t* live, size_t* peak) {
    if (!h) return fail(VARE_EINVAL, "null handle");
    if (live) *live = h->dev_bytes;
    if (peak) *peak = h->peak_bytes;
    return VARE_OK;
}

int vare_debug_inject_alloc_failure(long skip, long count) {
    g_alloc_fail.store(count > 0 ? count : 0);
    g_alloc_skip.store(count > 0 ? std::max(skip, 0L) : -1);
    return VARE_OK;
}

int vare_flip_destroy(vare_flip* h) {
    if (!h) return VARE_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (int q = 0; q < 9; ++q) dfree(h, h->pa[q]);
    for (int q = 0; q < 7; ++q) dfree(h, h->pb[q]);
    dfree(h, h->pid); dfree(h, h->pid2); dfree(h, h->removed); dfree(h, h->removed_log); dfree(h, h->hole); dfree(h, h->src);
    dfree(h, h->remap); dfree(h, h->pkey); dfree(h, h->prank); dfree(h, h->perm); dfree(h, h->longbins); dfree(h, h->midbins);
    dfree(h, h->flag); dfree(h, h->occ); dfree(h, h->slot); dfree(h, h->coord);
    free_cells(h);
    free_collision(h);
    dfree(h, h->tiles); dfree(h, h->dint); dfree(h, h->st);
    if (h->st_host) (void)hipHostFree(h->st_host);
    if (h->dint_host) (void)hipHostFree(h->dint_host);
    for (auto e : h->ev_pool) (void)hipEventDestroy(e);
    for (int q = 0; q < 4; ++q) if (h->step_ev[q]) (void)hipEventDestroy(h->step_ev[q]);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return VARE_OK;
}

int vare_flip_set_params(vare_flip* h, const vare_flip_params* p) {
    if (!h || !p) return fail(VARE_EINVAL, "null argument");
    CHK(validate_params(p));
    if ((p->velocity_fp16 ? 1 : 0) != h->vhalf)
        return fail(VARE_EINVAL, "velocity_fp16 is fixed when the handle is created");
    apply_params(h, p);
    if (p->solver_mode == VARE_SOLVER_RELATIVE && p->preconditioner == VARE_PRECOND_MG) CHK(ensure_mg(h));
    return VARE_OK;
}

int vare_flip_get_frame(const vare_flip* h, vare_frame_info* o) {
    if (!h || !o) return fail(VARE_EINVAL, "null argument");
    o->nx = h->nx; o->ny = h->ny; o->nz = h->nz;
    o->fx = h->fx; o->fy = h->fy; o->fz = h->fz;
    o->block_size = h->B;
    for (int a = 0; a < 3; ++a) { o->translate[a] = h->t[a]; o->scale[a] = h->s[a]; o->voxel_ext_force[a] = h->vox_force[a]; }
    o->voxel_wall_level = h->vox_wall;
    return VARE_OK;
}

static int upload_particles(vare_flip* h, size_t off, const float* pos, const float* vel, const float* vort,
                            const float* curv, const float* adj, size_t n, int space, int mem) {
    if (n == 0) return VARE_OK;
    float* dv = nullptr;          // interleaved velocities on the device (converted to FP16 if stored so)
    if (mem == VARE_MEM_HOST) {
        if (!finite3(pos, n) || !finite3(vel, n)) return fail(VARE_ENONFINITE, "NaN/Inf in particle input");
        std::vector<float> tmp(n);
        for (int a = 0; a < 3; ++a) {
            for (size_t i = 0; i < n; ++i) tmp[i] = pos[3 * i + a];
            HIPCHK(hipMemcpy(h->pa[a] + off, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice));
        }
        CHK(dalloc(h, &dv, 3 * n));
        int rc = hipMemcpy(dv, vel, 3 * n * sizeof(float), hipMemcpyHostToDevice) == hipSuccess
                     ? VARE_OK : fail(VARE_EHIP, "velocity upload failed");
        if (rc != VARE_OK) { dfree(h, dv); return rc; }
        const float* sc[3] = {adj, vort, curv};
        for (int q = 0; q < 3; ++q) {
            if (sc[q]) HIPCHK(hipMemcpy(h->pa[6 + q] + off, sc[q], n * sizeof(float), hipMemcpyHostToDevice));
            else CHK(fill_u32(h, h->pa[6 + q] + off, 0, n));
        }
    } else {
        vlaunch<k_deinterleave_f>(h, k_deinterleave, dim3(nblocks(n)), dim3(WG), 0, (int)n, pos, (void*)(h->pa[0] + off),
                           (void*)(h->pa[1] + off), (void*)(h->pa[2] + off), 0);
        const float* sc[3] = {adj, vort, curv};
        for (int q = 0; q < 3; ++q) {
            if (sc[q]) HIPCHK(hipMemcpyAsync(h->pa[6 + q] + off, sc[q], n * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
            else CHK(fill_u32(h, h->pa[6 + q] + off, 0, n));
        }
    }
    vlaunch<k_deinterleave_f>(h, k_deinterleave, dim3(nblocks(n)), dim3(WG), 0, (int)n, mem == VARE_MEM_HOST ? dv : vel,
                       h->vptr(0, off), h->vptr(1, off), h->vptr(2, off), h->vhalf);
    // new particles are stored at the end in their reference order
    vlaunch<k_iota_f>(h, k_iota, dim3(nblocks(n)), dim3(WG), 0, (int)n, (int)off, h->pid + off);
    if (space == VARE_SPACE_WORLD)
        vlaunch<k_to_voxel_f>(h, k_to_voxel, dim3(nblocks(n)), dim3(WG), 0, (int)n, h->parrays(off), h->t[0],
                           h->t[1], h->t[2], h->s[0], h->s[1], h->s[2]);
    const hipError_t e = hipGetLastError();
    const hipError_t e2 = hipStreamSynchronize(h->stream);
    dfree(h, dv);
    if (e != hipSuccess || e2 != hipSuccess) return fail(VARE_EHIP, "particle upload failed");
    return VARE_OK;
}

int vare_flip_add_particles(vare_flip* h, const float* pos, const float* vel, size_t n, int space, int mem) {
    if (!h || (n && (!pos || !vel))) return fail(VARE_EINVAL, "null argument");
    CHK(ensure_particles(h, h->np + n));
    CHK(upload_particles(h, h->np, pos, vel, nullptr, nullptr, nullptr, n, space, mem));
    h->np += n;
    h->stage_next = 0;
    return VARE_OK;
}

int vare_flip_set_particles(vare_flip* h, const float* pos, const float* vel, const float* vort, const float* curv,
                            const float* adj, size_t n, int space, int mem) {
    if (!h || (n && (!pos || !vel))) return fail(VARE_EINVAL, "null argument");
    CHK(ensure_particles(h, n));
    CHK(upload_particles(h, 0, pos, vel, vort, curv, adj, n, space, mem));
    h->np = n;
    h->stage_next = 0;
    return VARE_OK;
}

int vare_flip_get_particles(vare_flip* h, float* pos, float* vel, float* adj, float* vort, float* curv, size_t cap,
                            size_t* n, int space, int mem) {
    if (!h) return fail(VARE_EINVAL, "null handle");
    if (n) *n = h->np;
    if (cap < h->np) return fail(VARE_EINVAL, "capacity %zu < %zu particles", cap, h->np);
    const size_t np = h->np;
    if (np == 0) return VARE_OK;
    const int world = space == VARE_SPACE_WORLD ? 1 : 0;
    float* lg[2] = {vort, curv};                  // stored in the reference order already
    if (mem == VARE_MEM_HOST) {
        float* d = nullptr;
        CHK(dalloc(h, &d, 7 * np));
        vlaunch<k_export_f>(h, k_export, dim3(nblocks(np)), dim3(WG), 0, (int)np, h->parrays(),
                           (const int*)h->pid, world, h->t[0], h->t[1], h->t[2], h->s[0], h->s[1], h->s[2],
                           pos ? d : nullptr, vel ? d + 3 * np : nullptr, adj ? d + 6 * np : nullptr);
        int rc = hipStreamSynchronize(h->stream) == hipSuccess ? VARE_OK : fail(VARE_EHIP, "particle export failed");
        if (rc == VARE_OK && pos && hipMemcpy(pos, d, 3 * np * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(VARE_EHIP, "particle download failed");
        if (rc == VARE_OK && vel && hipMemcpy(vel, d + 3 * np, 3 * np * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(VARE_EHIP, "particle download failed");
        if (rc == VARE_OK && adj && hipMemcpy(adj, d + 6 * np, np * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(VARE_EHIP, "particle download failed");
        for (int q = 0; q < 2 && rc == VARE_OK; ++q)
            if (lg[q] && hipMemcpy(lg[q], h->pa[7 + q], np * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
                rc = fail(VARE_EHIP, "particle download failed");
        dfree(h, d);
        return rc;
    }
    vlaunch<k_export_f>(h, k_export, dim3(nblocks(np)), dim3(WG), 0, (int)np, h->parrays(), (const int*)h->pid,
                       world, h->t[0], h->t[1], h->t[2], h->s[0], h->s[1], h->s[2], pos, vel, adj);
    for (int q = 0; q < 2; ++q)
        if (lg[q]) HIPCHK(hipMemcpyAsync(lg[q], h->pa[7 + q], np * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    return VARE_OK;
}

int vare_flip_num_particles(const vare_flip* h, size_t* n) {
    if (!h || !n) return fail(VARE_EINVAL, "null argument");
    *n = h->np;
    return VARE_OK;
}

static int convert_all(vare_flip* h, bool to_voxel) {
    if (!h) return fail(VARE_EINVAL, "null handle");
    if (!h->np) return VARE_OK;
    const PArrays pa = h->parrays();
    if (to_voxel)
        vlaunch<k_to_voxel_f>(h, k_to_voxel, dim3(nblocks(h->np)), dim3(WG), 0, (int)h->np, pa, h->t[0], h->t[1],
                           h->t[2], h->s[0], h->s[1], h->s[2]);
    else
        vlaunch<k_to_world_f>(h, k_to_world, dim3(nblocks(h->np)), dim3(WG), 0, (int)h->np, pa, h->t[0], h->t[1],
                           h->t[2], h->s[0], h->s[1], h->s[2]);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    return VARE_OK;
}
int vare_flip_particles_to_voxel(vare_flip* h) { return convert_all(h, true); }
int vare_flip_particles_to_world(vare_flip* h) { return convert_all(h, false); }

int vare_flip_stage(vare_flip* h, int stage, float dt) {
    if (!h) return fail(VARE_EINVAL, "null handle");
    if (stage < 0 || stage > 5) return fail(VARE_EINVAL, "bad stage %d", stage);
    if (stage == 0) h->stats.num_removed = 0;
    CHK(dispatch_stage(h, stage, dt));
    return tcollect(h);
}

int vare_flip_step(vare_flip* h, float dt) {
    if (!h) return fail(VARE_EINVAL, "null handle");
    h->stage_next = 0;
    h->stats.num_removed = 0;
    HIPCHK(hevent(h, h->step_ev[0]));
    h->in_step = 1;
    for (int s = 0; s < 6; ++s) {
        const int rc = dispatch_stage(h, s, dt);
        if (rc != VARE_OK) { h->in_step = 0; return rc; }
    }
    h->in_step = 0;
    HIPCHK(hevent(h, h->step_ev[1]));
    CHK(hsync(h));
    float ms = 0.f, sms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, h->step_ev[0], h->step_ev[1]));
    HIPCHK(hipEventElapsedTime(&sms, h->step_ev[2], h->step_ev[3]));
    h->stats.step_ms = ms;
    h->stats.solve_ms = sms;
    CHK(tcollect(h));
    if (h->prm.solver_mode == VARE_SOLVER_RELATIVE && !h->stats.converged)
        return fail(VARE_ENOCONV, "PCG stopped at %d iterations, |r|=%g |b|=%g", h->stats.cg_iterations,
                    h->stats.residual_norm, h->stats.b_norm);
    return VARE_OK;
}

int vare_flip_step_noalloc(vare_flip* h, float dt) {
    if (!h) return fail(VARE_EINVAL, "null handle");
    if (h->timing) return fail(VARE_EINVAL, "kernel timing creates events on demand: switch it off for the no-allocation step");
    if (!h->mass) return fail(VARE_ENOMEM, "the handle has no cell arrays");
    // every reduction of the step writes at most max(kReduceGrid, ceil(blocks / 2)) partials (the
    // block kernels take two or more blocks per workgroup); the longest scan covers the frame, the
    // bins or the particles
    const size_t npart = std::max((size_t)kReduceGrid, (h->bcap + 1) / 2) + 16;
    const size_t ntile = step_scan_tiles(h, h->np) + 16;
    const size_t off2 = align256(npart * sizeof(double));
    const size_t offt = off2 + align256(npart * sizeof(double));
    const size_t need = offt + ntile * sizeof(int);
    if (need > (size_t)VARE_NOALLOC_SCRATCH_BYTES)
        return fail(VARE_ENOMEM, "the step needs %zu bytes of scratch (block capacity %zu, %zu particles); the no-allocation "
                    "step holds %zu", need, h->bcap, h->np, (size_t)VARE_NOALLOC_SCRATCH_BYTES);
    void* base = nullptr;
    CHK(noalloc_scratch(h->device, &base));
    double* const own_p = h->partials;
    double* const own_p2 = h->partials2;
    int* const own_t = h->tiles;
    const size_t own_tcap = h->tiles_cap;
    h->partials = (double*)base;
    h->partials2 = (double*)((char*)base + off2);
    h->tiles = (int*)((char*)base + offt);
    h->tiles_cap = ntile;
    h->noalloc = 1;
    const int rc = vare_flip_step(h, dt);
    h->noalloc = 0;
    h->partials = own_p;
    h->partials2 = own_p2;
    h->tiles = own_t;
    h->tiles_cap = own_tcap;
    return rc;
}

int vare_flip_post_process(vare_flip* h) {
    if (!h) return fail(VARE_EINVAL, "null handle");
    const int rc = h->B == 4 ? run_post_process<4>(h) : run_post_process<8>(h);
    h->stage_next = 0;
    return rc;
}

int vare_flip_advance_frame(vare_flip* h) {
    if (!h) return fail(VARE_EINVAL, "null handle");
    if (h->np == 0) return VARE_OK;                         // FLIPMethod.cu:179
    CHK(convert_all(h, true));                              // :191
    int rc = vare_flip_step(h, h->prm.dt);                  // :193
    if (rc == VARE_OK || rc == VARE_ENOCONV) {
        const int rp = vare_flip_post_process(h);           // :198
        if (rp != VARE_OK) rc = rp;
    }
    CHK(convert_all(h, false));                             // :200
    return rc;
}

int vare_flip_export_surface(vare_flip* h, int32_t* ijk, float* phi, size_t cap, size_t* n, float* background,
                             float xform[4]) {
    if (!h) return fail(VARE_EINVAL, "null handle");
    if (!h->post_done) return fail(VARE_EINVAL, "the surface field is built by vare_flip_post_process");
    const int B = h->B, B3 = B * B * B;
    const size_t nc = h->ncells();
    const float dx = h->s[0];
    if (background) *background = kSurfBackground * dx;               // FloatGrid::create(background*dx)
    if (xform) { xform[0] = dx; for (int a = 0; a < 3; ++a) xform[1 + a] = h->t[a] + dx * 0.5f; }
    std::vector<float> surf(nc);
    std::vector<int> coord((size_t)h->nb);
    if (nc) HIPCHK(hipMemcpy(surf.data(), h->cnt, nc * sizeof(float), hipMemcpyDeviceToHost));
    if (h->nb) HIPCHK(hipMemcpy(coord.data(), h->coord, coord.size() * sizeof(int), hipMemcpyDeviceToHost));
    // convertSurfaceVDB (FieldCast.h:362-391): every active cell, phi = surf * dx
    std::vector<std::array<int, 3>> cells(nc);
    std::vector<float> vals(nc);
    std::unordered_map<long long, size_t> where;
    const long long ox = 1 << 20;                                     // key offset: shell cells reach -3
    auto key = [&](int i, int j, int k) { return (((long long)(k + 8) * ox) + (j + 8)) * ox + (i + 8); };
    for (size_t c = 0; c < nc; ++c) {
        const int t = coord[c / B3], l = (int)(c % B3);
        cells[c] = {(t % h->fx) * B + l % B, ((t / h->fx) % h->fy) * B + (l / B) % B, (t / (h->fx * h->fy)) * B + l / (B * B)};
        vals[c] = surf[c] * dx;
    }
    if (h->prm.enable_water_tank) {
        // FieldCast.h:392-440: a 3-cell shell outside the domain below the tank level holds
        // min(world cell centre y - level, value), the faces' cells included
        for (size_t c = 0; c < nc; ++c) where[key(cells[c][0], cells[c][1], cells[c][2])] = c;
        const int EI = h->nx - 1, EJ = h->ny - 1, EK = h->nz - 1, bnd = 3;
        (void)EJ;
        const float level = h->prm.wall_level;
        const int lv = (int)((level - h->t[1]) / dx + 1.f);
        auto put = [&](int i, int j, int k) {
            const float cy = ((float)j + 0.5f) * h->s[1] + h->t[1];       // Grid::worldPoint
            const long long kk = key(i, j, k);
            auto it = where.find(kk);
            if (it == where.end()) {
                where[kk] = cells.size();
                cells.push_back({i, j, k});
                vals.push_back(std::min(cy - level, kSurfBackground * dx));
            } else {
                vals[it->second] = std::min(cy - level, vals[it->second]);
            }
        };
        for (int k = -bnd; k <= EK + bnd; ++k) for (int j = -bnd; j <= 0; ++j) for (int i = -bnd; i <= EI + bnd; ++i) put(i, j, k);
        for (int k = -bnd; k <= EK + bnd; ++k) for (int j = -bnd; j <= lv + bnd; ++j) for (int i = -bnd; i <= 0; ++i) put(i, j, k);
        for (int k = -bnd; k <= EK + bnd; ++k) for (int j = -bnd; j <= lv + bnd; ++j) for (int i = EI; i <= EI + bnd; ++i) put(i, j, k);
        for (int k = -bnd; k <= 0; ++k) for (int j = -bnd; j <= lv + bnd; ++j) for (int i = -bnd; i <= EI + bnd; ++i) put(i, j, k);
        for (int k = EK; k <= EK + bnd; ++k) for (int j = -bnd; j <= lv + bnd; ++j) for (int i = -bnd; i <= EI + bnd; ++i) put(i, j, k);
    }
    if (n) *n = cells.size();
    if (!ijk && !phi) return VARE_OK;
    if (cap < cells.size()) return fail(VARE_EINVAL, "capacity %zu < %zu voxels", cap, cells.size());
    for (size_t q = 0; q < cells.size(); ++q) {
        if (ijk) for (int a = 0; a < 3; ++a) ijk[3 * q + a] = cells[q][a];
        if (phi) phi[q] = vals[q];
    }
    return VARE_OK;
}

int vare_flip_export_velocity(vare_flip* h, int32_t* ijk, float* vel, size_t cap, size_t* n, float xform[4]) {
    if (!h) return fail(VARE_EINVAL, "null handle");
    if (!h->post_done) return fail(VARE_EINVAL, "the splatted velocity is built by vare_flip_post_process");
    const int B = h->B, B3 = B * B * B;
    const size_t nc = h->ncells();
    const float dx = h->s[0];
    if (xform) { xform[0] = dx; for (int a = 0; a < 3; ++a) xform[1 + a] = h->t[a] + dx * 0.5f; }
    std::vector<float> v3[3];
    std::vector<int> coord((size_t)h->nb);
    const float* src[3] = {h->vx, h->vy, h->vz};
    for (int a = 0; a < 3; ++a) {
        v3[a].resize(nc);
        if (nc) HIPCHK(hipMemcpy(v3[a].data(), src[a], nc * sizeof(float), hipMemcpyDeviceToHost));
    }
    if (h->nb) HIPCHK(hipMemcpy(coord.data(), h->coord, coord.size() * sizeof(int), hipMemcpyDeviceToHost));
    // convertVelocityVDB (FieldCast.h:14-47): active cells whose world velocity is not below 1e-30
    size_t cnt = 0;
    for (size_t c = 0; c < nc; ++c) {
        const float w[3] = {v3[0][c] * h->s[0], v3[1][c] * h->s[1], v3[2][c] * h->s[2]};   // Grid::worldVector
        if (std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]) < 1e-30f) continue;
        if (ijk || vel) {
            if (cnt >= cap) return fail(VARE_EINVAL, "capacity %zu too small", cap);
            const int t = coord[c / B3], l = (int)(c % B3);
            if (ijk) {
                ijk[3 * cnt] = (t % h->fx) * B + l % B;
                ijk[3 * cnt + 1] = ((t / h->fx) % h->fy) * B + (l / B) % B;
                ijk[3 * cnt + 2] = (t / (h->fx * h->fy)) * B + l / (B * B);
            }
            if (vel) for (int a = 0; a < 3; ++a) vel[3 * cnt + a] = w[a];
        }
        cnt++;
    }
    if (n) *n = cnt;
    return VARE_OK;
}

int vare_flip_get_stats(const vare_flip* h, vare_step_stats* o) {
    if (!h || !o) return fail(VARE_EINVAL, "null argument");
    *o = h->stats;
    return VARE_OK;
}

int vare_flip_get_block_table(vare_flip* h, int64_t* pointers, int32_t* coords, size_t cap, size_t* nb) {
    if (!h) return fail(VARE_EINVAL, "null handle");
    if (nb) *nb = (size_t)h->nb;
    const int B3 = h->B * h->B * h->B;
    if (pointers) {
        std::vector<int> sl((size_t)h->nblk);
        HIPCHK(hipMemcpy(sl.data(), h->slot, sl.size() * sizeof(int), hipMemcpyDeviceToHost));
        for (size_t t = 0; t < sl.size(); ++t) pointers[t] = sl[t] < 0 ? -1 : (int64_t)sl[t] * B3;
    }
    if (coords) {
        if (cap < (size_t)h->nb) return fail(VARE_EINVAL, "coords capacity too small");
        std::vector<int> co((size_t)h->nb);
        if (h->nb) HIPCHK(hipMemcpy(co.data(), h->coord, co.size() * sizeof(int), hipMemcpyDeviceToHost));
        for (size_t q = 0; q < co.size(); ++q) {
            const int t = co[q];
            coords[3 * q] = t % h->fx;
            coords[3 * q + 1] = (t / h->fx) % h->fy;
            coords[3 * q + 2] = t / (h->fx * h->fy);
        }
    }
    return VARE_OK;
}

int vare_flip_get_cell_field(vare_flip* h, int which, void* out, size_t cap_bytes, size_t* n) {
    if (!h) return fail(VARE_EINVAL, "null handle");
    const size_t nc = h->ncells();
    if (n) *n = nc;
    if (!out) return VARE_OK;
    size_t need = nc * 4;
    if (which == VARE_FIELD_VELOCITY || which == VARE_FIELD_COMP) need = nc * 12;
    if (which == VARE_FIELD_MARK) need = nc * 8;
    if (cap_bytes < need) return fail(VARE_EINVAL, "buffer too small (%zu < %zu)", cap_bytes, need);
    if (nc == 0) return VARE_OK;
    const float* one = nullptr;
    switch (which) {
        case VARE_FIELD_MASS: one = h->mass; break;
        case VARE_FIELD_PRESSURE: one = h->x; break;
        case VARE_FIELD_COUNT: one = h->cnt; break;
        case VARE_FIELD_DIV: one = h->b; break;
        default: break;
    }
    if (one) {
        HIPCHK(hipMemcpy(out, one, nc * 4, hipMemcpyDeviceToHost));
        return VARE_OK;
    }
    if (which == VARE_FIELD_VELOCITY || which == VARE_FIELD_COMP) {
        const float* c3[3] = {which == VARE_FIELD_VELOCITY ? h->vx : h->cx, which == VARE_FIELD_VELOCITY ? h->vy : h->cy,
                              which == VARE_FIELD_VELOCITY ? h->vz : h->cz};
        std::vector<float> buf(nc);
        float* o = (float*)out;
        for (int a = 0; a < 3; ++a) {
            HIPCHK(hipMemcpy(buf.data(), c3[a], nc * 4, hipMemcpyDeviceToHost));
            for (size_t c = 0; c < nc; ++c) o[3 * c + a] = buf[c];
        }
        return VARE_OK;
    }
    if (which == VARE_FIELD_MARK) {
        OpFluid op{h->type, h->rank};
        CHK(run_scan(h, op, (int)nc, 2));
        unsigned long long* tmp = nullptr;
        CHK(dalloc(h, &tmp, nc));
        vlaunch<k_mark_u64_f>(h, k_mark_u64, dim3(nblocks(nc)), dim3(WG), 0, (int)nc, (const uint8_t*)h->type,
                           (const int*)h->rank, tmp);
        int rc = VARE_OK;
        if (hipStreamSynchronize(h->stream) != hipSuccess || hipMemcpy(out, tmp, nc * 8, hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(VARE_EHIP, "mark download failed");
        dfree(h, tmp);
        return rc;
    }
    return fail(VARE_EINVAL, "bad field id %d", which);
}

int vare_flip_get_binning(vare_flip* h, int32_t* perm, int32_t* bstart, size_t cap, size_t* n, size_t* nbucket) {
    if (!h) return fail(VARE_EINVAL, "null handle");
    if (n) *n = h->np;
    if (nbucket) *nbucket = h->ncells() + 1;
    if (h->prm.particle_path != 0) return fail(VARE_EINVAL, "binning is only built on particle_path 0");
    if (perm) {                                   // the physical order is the bin order: its logical indices
        if (cap < h->np) return fail(VARE_EINVAL, "capacity too small");
        if (h->np) HIPCHK(hipMemcpy(perm, h->pid, h->np * sizeof(int), hipMemcpyDeviceToHost));
    }
    if (bstart) HIPCHK(hipMemcpy(bstart, h->cstart, (h->ncells() + 2) * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(h->dint_host + 4, h->dint + 4, sizeof(int), hipMemcpyDeviceToHost));
    if (h->dint_host[4] != 0) return fail(VARE_ENOTIMPL, "%d bin(s) longer than %d were left unsorted", h->dint_host[4], kSortMax);
    return VARE_OK;
}

int vare_flip_get_removed_mask(vare_flip* h, uint8_t* out, size_t cap, size_t* n) {
    if (!h) return fail(VARE_EINVAL, "null handle");
    const size_t m = h->last_np_before_remove;
    if (n) *n = m;
    if (!out) return VARE_OK;
    if (cap < m) return fail(VARE_EINVAL, "capacity too small");
    if (m) HIPCHK(hipMemcpy(out, h->removed_log, m, hipMemcpyDeviceToHost));
    return VARE_OK;
}

}  // extern "C"

namespace {

// Host seeding shared by the seeders: for every cell whose centre satisfies `inside`, ppc points
// cell + (Rand(s+3q), Rand(s+3q+1), Rand(s+3q+2)), q = ppc*(i + nx*(j + ny*k)) + n, kept when the
// point itself satisfies `inside`.  Cells are visited in (k, j, i) order; z-slabs are counted and
// filled by parallel host threads, which keeps the order.
template <class Inside>
int seed_points(int64_t nx, int64_t ny, int64_t nz, int ppc, uint32_t seed, float* pos, size_t cap, size_t* n,
                Inside inside) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || ppc <= 0) return fail(VARE_EINVAL, "bad seeding arguments");
    auto slab = [&](int64_t k, float* out, size_t limit) {
        size_t cnt = 0;
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                if (!inside((float)i + 0.5f, (float)j + 0.5f, (float)k + 0.5f)) continue;
                const uint64_t cell = (uint64_t)i + (uint64_t)nx * ((uint64_t)j + (uint64_t)ny * (uint64_t)k);
                for (int q = 0; q < ppc; ++q) {
                    const uint32_t qq = (uint32_t)((uint64_t)ppc * cell + (uint64_t)q);
                    const float px = (float)i + rand01(seed + 3u * qq);
                    const float py = (float)j + rand01(seed + 3u * qq + 1u);
                    const float pz = (float)k + rand01(seed + 3u * qq + 2u);
                    if (!inside(px, py, pz)) continue;
                    if (out && cnt < limit) { out[3 * cnt] = px; out[3 * cnt + 1] = py; out[3 * cnt + 2] = pz; }
                    cnt++;
                }
            }
        return cnt;
    };
    const int nt = (int)std::max<unsigned>(1u, std::min<unsigned>(16u, std::thread::hardware_concurrency()));
    std::vector<size_t> counts((size_t)nz, 0);
    auto run = [&](auto body) {
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t)
            th.emplace_back([&, t]() { for (int64_t k = t; k < nz; k += nt) body(k); });
        for (auto& x : th) x.join();
    };
    run([&](int64_t k) { counts[(size_t)k] = slab(k, nullptr, 0); });
    size_t total = 0;
    std::vector<size_t> offs((size_t)nz, 0);
    for (int64_t k = 0; k < nz; ++k) { offs[(size_t)k] = total; total += counts[(size_t)k]; }
    if (n) *n = total;
    if (!pos) return VARE_OK;
    if (total > cap) return fail(VARE_EINVAL, "capacity %zu < %zu points", cap, total);
    run([&](int64_t k) { slab(k, pos + 3 * offs[(size_t)k], counts[(size_t)k]); });
    return VARE_OK;
}

// CollisionSource::update (CollisionSource.h:66-126), host-side as in the reference.  The OpenVDB
// grids are replaced by `seeds` (world points of the source's active voxels, nseed of them) and a
// sampler of signed distance, velocity and gradient (world units) at a world point:
//   * the source's block table marks the 3x3x3 block neighbourhood of every seed's block
//     (SparseFrame::buildFromPoints, SparseFrame.cu:145-157, with its index aliasing; negative
//     indices are dropped), active blocks numbered in ascending linear order;
//   * every active cell centre is sampled: phi = sdf / h (h = Lx / nx), velocity and gradient are
//     converted with voxelVector and the gradient is normalised (Vector3::normalized);
//   * inactive cells read background / h.
// `threads` > 1 samples blocks on several host threads (the sampler must then be reentrant).
template <class Sampler>
int build_collision(vare_flip* h, const float* seeds, size_t nseed, float bg_world, Sampler sample, int threads) {
    const int B = h->B, fx = h->fx, fy = h->fy, fz = h->fz;
    const int B3 = B * B * B;
    const float hv = h->s[0];
    std::vector<char> mark((size_t)h->nblk, 0);
    for (size_t q = 0; q < nseed; ++q) {
        float vp[3];
        for (int a = 0; a < 3; ++a) vp[a] = (seeds[3 * q + a] - h->t[a]) / h->s[a];      // Grid::voxelPoint
        const int ci = std::max((int)vp[0] / B, 0), cj = std::max((int)vp[1] / B, 0), ck = std::max((int)vp[2] / B, 0);
        for (int k = ck - 1; k <= std::min(ck + 1, fz - 1); ++k)
            for (int j = cj - 1; j <= std::min(cj + 1, fy - 1); ++j)
                for (int i = ci - 1; i <= std::min(ci + 1, fx - 1); ++i) {
                    const long long t = (long long)(k * fx * fy + j * fx + i);
                    if (t >= 0) mark[(size_t)t] = 1;
                }
    }
    std::vector<int> slot((size_t)h->nblk, -1), blocks;
    for (size_t t = 0; t < mark.size(); ++t)
        if (mark[t]) { slot[t] = (int)blocks.size(); blocks.push_back((int)t); }
    const size_t nbc = blocks.size(), nc = nbc * (size_t)B3;
    std::vector<float> phi(nc), vel(3 * nc), nrm(3 * nc);
    auto fill_block = [&](size_t b) {
        const int t = blocks[b];
        const int bi = t % fx, bj = (t / fx) % fy, bk = t / (fx * fy);
        for (int l = 0; l < B3; ++l) {
            const size_t c = b * (size_t)B3 + (size_t)l;
            const int i = bi * B + l % B, j = bj * B + (l / B) % B, k = bk * B + l / (B * B);
            const float ctr[3] = {(float)i + 0.5f, (float)j + 0.5f, (float)k + 0.5f};   // Grid::cellCenter
            float wp[3];
            for (int a = 0; a < 3; ++a) wp[a] = ctr[a] * h->s[a] + h->t[a];              // Grid::worldPoint
            float sdf = 0.f, v[3] = {0.f, 0.f, 0.f}, g[3] = {0.f, 0.f, 0.f};
            sample(wp, &sdf, v, g);
            phi[c] = sdf / hv;
            for (int a = 0; a < 3; ++a) { vel[3 * c + a] = v[a] / h->s[a]; g[a] = g[a] / h->s[a]; }
            const float d = 1.f / std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2] + 1e-30f);
            for (int a = 0; a < 3; ++a) nrm[3 * c + a] = g[a] * d;
        }
    };
    if (threads > 1 && nbc > 1) {
        std::vector<std::thread> th;
        for (int w = 0; w < threads; ++w)
            th.emplace_back([&, w]() { for (size_t b = (size_t)w; b < nbc; b += (size_t)threads) fill_block(b); });
        for (auto& x : th) x.join();
    } else {
        for (size_t b = 0; b < nbc; ++b) fill_block(b);
    }
    // upload: block table + structure-of-arrays fields
    HIPCHK(hipStreamSynchronize(h->stream));
    free_collision(h);
    h->col_bg = bg_world / hv;                                   // CollisionSource.h:90
    if (nbc == 0) return VARE_OK;                                // size()==0: every query returns background
    CHK(dalloc(h, &h->col_slot, (size_t)h->nblk));
    CHK(dalloc(h, &h->col_phi, nc));
    for (int a = 0; a < 3; ++a) { CHK(dalloc(h, &h->col_v[a], nc)); CHK(dalloc(h, &h->col_n[a], nc)); }
    HIPCHK(hipMemcpy(h->col_slot, slot.data(), slot.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->col_phi, phi.data(), nc * sizeof(float), hipMemcpyHostToDevice));
    std::vector<float> tmp(nc);
    for (int a = 0; a < 3; ++a) {
        for (size_t c = 0; c < nc; ++c) tmp[c] = vel[3 * c + a];
        HIPCHK(hipMemcpy(h->col_v[a], tmp.data(), nc * sizeof(float), hipMemcpyHostToDevice));
        for (size_t c = 0; c < nc; ++c) tmp[c] = nrm[3 * c + a];
        HIPCHK(hipMemcpy(h->col_n[a], tmp.data(), nc * sizeof(float), hipMemcpyHostToDevice));
    }
    h->col_cells = nc;
    h->col_active = 1;
    return VARE_OK;
}

// Axis-aligned box [lo,hi] (world): exact signed distance (negative inside) and gradient, in FP64,
// rounded once to FP32.  With q = |p - c| - e per axis (c centre, e half extent): outside,
// sdf = |max(q,0)| and grad = sign(p-c) max(q,0) / sdf; inside, sdf = max(q) and grad is the unit
// vector along the axis holding max(q) (x before y before z on ties), signed by p - c.
struct BoxSdf {
    double c[3], e[3];
    BoxSdf(const float lo[3], const float hi[3]) {
        for (int a = 0; a < 3; ++a) {
            c[a] = 0.5 * ((double)lo[a] + (double)hi[a]);
            e[a] = 0.5 * ((double)hi[a] - (double)lo[a]);
        }
    }
    void operator()(const float p[3], float* sdf, float grad[3]) const {
        double d[3], q[3], o[3];
        for (int a = 0; a < 3; ++a) {
            d[a] = (double)p[a] - c[a];
            q[a] = std::fabs(d[a]) - e[a];
            o[a] = q[a] > 0.0 ? q[a] : 0.0;
        }
        const double out = std::sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
        const double qmax = std::max(q[0], std::max(q[1], q[2]));
        *sdf = (float)(out + std::min(qmax, 0.0));
        if (out > 0.0) {
            for (int a = 0; a < 3; ++a) grad[a] = (float)((d[a] < 0.0 ? -o[a] : o[a]) / out);
            return;
        }
        const int ax = (q[0] >= q[1] && q[0] >= q[2]) ? 0 : (q[1] >= q[2] ? 1 : 2);
        for (int a = 0; a < 3; ++a) grad[a] = a != ax ? 0.f : (d[a] < 0.0 ? -1.f : 1.f);
    }
};

int host_threads() { return (int)std::max<unsigned>(1u, std::min<unsigned>(16u, std::thread::hardware_concurrency())); }

}  // namespace

extern "C" {

int vare_seed_box(int64_t nx, int64_t ny, int64_t nz, float wx, float wy, float wz, int ppc, uint32_t seed, float* pos,
                  size_t cap, size_t* n) {
    return seed_points(nx, ny, nz, ppc, seed, pos, cap, n,
                       [=](float x, float y, float z) { return x < wx && y < wy && z < wz; });
}

int vare_seed_pool_sphere(int64_t nx, int64_t ny, int64_t nz, float depth, float cx, float cy, float cz, float r, int ppc,
                          uint32_t seed, float* pos, size_t cap, size_t* n) {
    const float r2 = r * r;
    return seed_points(nx, ny, nz, ppc, seed, pos, cap, n, [=](float x, float y, float z) {
        const float dx = x - cx, dy = y - cy, dz = z - cz;
        return y < depth || dx * dx + dy * dy + dz * dz < r2;
    });
}

int vare_flip_set_collision(vare_flip* h, const float* seeds, size_t nseed, float background,
                            vare_collision_sampler fn, void* user) {
    if (!h || !fn || (nseed && !seeds)) return fail(VARE_EINVAL, "null argument");
    return build_collision(
        h, seeds, nseed, background,
        [&](const float* wp, float* sdf, float* v, float* g) { fn(user, wp, sdf, v, g); }, 1);
}

int vare_flip_set_collision_box(vare_flip* h, const float lo[3], const float hi[3], const float vel[3], float band) {
    if (!h || !lo || !hi) return fail(VARE_EINVAL, "null argument");
    if (!(band > 0.f) || !std::isfinite(band)) return fail(VARE_EINVAL, "band must be a positive number of voxels");
    for (int a = 0; a < 3; ++a)
        if (!(lo[a] <= hi[a]) || !std::isfinite(lo[a]) || !std::isfinite(hi[a]))
            return fail(VARE_EINVAL, "box needs finite lo <= hi");
    const BoxSdf box(lo, hi);
    const float v[3] = {vel ? vel[0] : 0.f, vel ? vel[1] : 0.f, vel ? vel[2] : 0.f};
    const float hv = h->s[0];
    // active voxels: cells whose centre lies within `band` voxels of the surface; only cells of the
    // box's bounding region grown by band+2 voxels can qualify
    int lo_i[3], hi_i[3];
    const int64_t n[3] = {h->nx, h->ny, h->nz};
    for (int a = 0; a < 3; ++a) {
        const double vl = ((double)lo[a] - h->t[a]) / h->s[a], vh = ((double)hi[a] - h->t[a]) / h->s[a];
        lo_i[a] = (int)std::max<double>(0.0, std::min<double>((double)n[a], std::floor(vl - band - 2.0)));
        hi_i[a] = (int)std::max<double>(0.0, std::min<double>((double)n[a], std::ceil(vh + band + 2.0)));
    }
    std::vector<float> seeds;
    for (int k = lo_i[2]; k < hi_i[2]; ++k)
        for (int j = lo_i[1]; j < hi_i[1]; ++j)
            for (int i = lo_i[0]; i < hi_i[0]; ++i) {
                const float ctr[3] = {(float)i + 0.5f, (float)j + 0.5f, (float)k + 0.5f};
                float wp[3], sdf, g[3];
                for (int a = 0; a < 3; ++a) wp[a] = ctr[a] * h->s[a] + h->t[a];
                box(wp, &sdf, g);
                if (std::fabs(sdf / hv) < band) seeds.insert(seeds.end(), wp, wp + 3);
            }
    return build_collision(
        h, seeds.data(), seeds.size() / 3, band * hv,
        [&](const float* wp, float* sdf, float* vo, float* g) {
            box(wp, sdf, g);
            for (int a = 0; a < 3; ++a) vo[a] = v[a];
        },
        host_threads());
}

int vare_flip_clear_collision(vare_flip* h) {
    if (!h) return fail(VARE_EINVAL, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    free_collision(h);
    h->col_bg = 5.f;
    return VARE_OK;
}

int vare_flip_get_collision(vare_flip* h, int64_t* pointers, float* phi, float* vel, float* nrm, size_t cap, size_t* nc,
                            float* background) {
    if (!h) return fail(VARE_EINVAL, "null handle");
    const size_t m = h->col_cells;
    if (nc) *nc = m;
    if (background) *background = h->col_bg;
    const int B3 = h->B * h->B * h->B;
    if (pointers) {
        std::vector<int> slot((size_t)h->nblk, -1);
        if (h->col_slot) HIPCHK(hipMemcpy(slot.data(), h->col_slot, slot.size() * sizeof(int), hipMemcpyDeviceToHost));
        for (size_t t = 0; t < slot.size(); ++t) pointers[t] = slot[t] < 0 ? -1 : (int64_t)slot[t] * B3;
    }
    if ((phi || vel || nrm) && cap < m) return fail(VARE_EINVAL, "capacity %zu < %zu cells", cap, m);
    if (m == 0) return VARE_OK;
    if (phi) HIPCHK(hipMemcpy(phi, h->col_phi, m * sizeof(float), hipMemcpyDeviceToHost));
    std::vector<float> tmp(m);
    for (int a = 0; a < 3; ++a) {
        if (vel) {
            HIPCHK(hipMemcpy(tmp.data(), h->col_v[a], m * sizeof(float), hipMemcpyDeviceToHost));
            for (size_t c = 0; c < m; ++c) vel[3 * c + a] = tmp[c];
        }
        if (nrm) {
            HIPCHK(hipMemcpy(tmp.data(), h->col_n[a], m * sizeof(float), hipMemcpyDeviceToHost));
            for (size_t c = 0; c < m; ++c) nrm[3 * c + a] = tmp[c];
        }
    }
    return VARE_OK;
}

int vare_device_synchronize(void) {
    HIPCHK(hipDeviceSynchronize());
    return VARE_OK;
}

int vare_device_name(char* buf, size_t cap) {
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, 0));
    if (buf && cap) {
        snprintf(buf, cap, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    }
    return VARE_OK;
}

int vare_flip_set_kernel_timing(vare_flip* h, int enable) {
    if (!h) return fail(VARE_EINVAL, "null handle");
    CHK(tcollect(h));
    h->timing = (unsigned)enable;
    return VARE_OK;
}

int vare_flip_get_kernel_timing(vare_flip* h, double ms[VARE_KERNEL_COUNT], int64_t launches[VARE_KERNEL_COUNT],
                                int64_t units[VARE_KERNEL_COUNT], int reset) {
    if (!h) return fail(VARE_EINVAL, "null handle");
    CHK(tcollect(h));
    for (int q = 0; q < VARE_KERNEL_COUNT; ++q) {
        if (ms) ms[q] = h->kms[q];
        if (launches) launches[q] = h->klaunch[q];
        if (units) units[q] = h->kunits[q];
        if (reset) { h->kms[q] = 0; h->klaunch[q] = 0; h->kunits[q] = 0; }
    }
    return VARE_OK;
}

}  // extern "C"

// ---- batched stepping of many tanks (vare_flip_step_batched) ----
//
// Each tank's step runs its ordinary host code on a thread of its own, with the handle in record
// mode: launches, readbacks and event records are appended to the handle's list instead of being
// issued.  A sync point (a readback of the block count, the solver state, the removal count; the
// end of the step) is a barrier over the tanks still stepping: when every one of them has reached
// one, the last to arrive merges all the lists and issues them on one stream.  Consecutive
// launches of the same kernel by different tanks become one launch of k_batched over all their
// workgroups; each tank's operations keep their order.  Tanks are independent (no array is shared),
// so only that order matters, and every tank computes exactly what it computes alone.
struct BatchRun {
    std::mutex call;                  // one batched step at a time per device
    std::mutex m;                     // barrier state below
    int active = 0;                   // tanks still inside their step
    int waiting = 0;                  // of those, waiting at a sync point
    std::atomic<uint64_t> gen{0};     // flushes so far (waiters spin on it outside the lock)
    std::vector<vare_flip*> members;
    int device = 0;
    hipStream_t stream = nullptr;
    // worker threads: one per tank of the largest batch so far
    std::mutex pm;
    std::condition_variable pcv, pdone;
    std::vector<std::thread> threads;
    uint64_t job = 0;
    int njob = 0, left = 0;
    vare_flip* const* jh = nullptr;
    const float* jdt = nullptr;
    int* jrc = nullptr;
    std::vector<std::string> jerr;
    std::chrono::steady_clock::time_point last_flush_end{};   // within the current batched step
    int chunk_hint[20] = {};          // per solve of the step: the most iterations a member took last step
    vare_batch_stats stats{};
};

namespace {

constexpr size_t kLookahead = 64;     // operations searched ahead when the tanks' next kernels differ

// A merged launch: the operations sel[first, first + n), all of one kernel, or one copy or event.
struct MergedItem {
    int kind;
    size_t first, n;
};

// Chooses, among the tanks' next operations, the one to issue now: the kernel that the most tanks
// are at, unless another tank meets it within kLookahead operations (it would then run without that
// tank); operations no other tank will reach soon go first.  Returns the index of a tank at it.
size_t pick_head(const std::vector<vare_flip*>& ms, const std::vector<size_t>& head) {
    size_t first = ms.size();
    bool uniform = true;
    for (size_t i = 0; i < ms.size(); ++i) {
        if (head[i] >= ms[i]->rops.size()) continue;
        if (first == ms.size()) first = i;
        else if (!ms[i]->rops[head[i]].same(ms[first]->rops[head[first]])) uniform = false;
    }
    if (uniform) return first;
    size_t best = ms.size();
    int best_count = -1;
    bool best_free = false;
    for (size_t i = 0; i < ms.size(); ++i) {
        if (head[i] >= ms[i]->rops.size()) continue;
        const ROp& c = ms[i]->rops[head[i]];
        bool seen = false;
        for (size_t j = 0; j < i && !seen; ++j)
            if (head[j] < ms[j]->rops.size() && ms[j]->rops[head[j]].same(c)) seen = true;
        if (seen) continue;
        int count = 0;
        bool free_ = true;
        for (size_t j = 0; j < ms.size(); ++j) {
            if (head[j] >= ms[j]->rops.size()) continue;
            if (ms[j]->rops[head[j]].same(c)) { ++count; continue; }
            const size_t end = std::min(ms[j]->rops.size(), head[j] + 1 + kLookahead);
            for (size_t q = head[j] + 1; q < end && free_; ++q)
                if (ms[j]->rops[q].same(c)) free_ = false;
        }
        if (best == ms.size() || (free_ && !best_free) || (free_ == best_free && count > best_count)) {
            best = i;
            best_count = count;
            best_free = free_;
        }
    }
    return best;
}

// Merges and issues the recorded operations of the tanks in ms, then waits for them.
// The issue order of the tanks' recorded operations: sel lists them, items groups consecutive
// entries of sel into merged launches (one kernel for several tanks) or single copies and events.
void plan_merged(const std::vector<vare_flip*>& ms, std::vector<std::pair<vare_flip*, const ROp*>>& sel,
                 std::vector<MergedItem>& items) {
    std::vector<size_t> head(ms.size(), 0);
    sel.clear();
    items.clear();
    for (;;) {
        const size_t i0 = pick_head(ms, head);
        if (i0 >= ms.size()) break;
        const ROp ref = ms[i0]->rops[head[i0]];
        if (ref.kind != ROp::kKernel) {
            items.push_back(MergedItem{ref.kind, sel.size(), 1});
            sel.emplace_back(ms[i0], &ms[i0]->rops[head[i0]]);
            ++head[i0];
            continue;
        }
        MergedItem it{ROp::kKernel, sel.size(), 0};
        for (size_t i = 0; i < ms.size(); ++i) {
            if (head[i] >= ms[i]->rops.size() || !ms[i]->rops[head[i]].same(ref)) continue;
            if (it.n == (size_t)ref.max_tanks) {
                items.push_back(it);
                it = MergedItem{ROp::kKernel, sel.size(), 0};
            }
            sel.emplace_back(ms[i], &ms[i]->rops[head[i]]);
            ++head[i];
            ++it.n;
        }
        items.push_back(it);
    }
}

hipError_t run_merged(BatchRun& R, const std::vector<vare_flip*>& ms) {
    std::vector<std::pair<vare_flip*, const ROp*>> sel;
    std::vector<MergedItem> items;
    plan_merged(ms, sel, items);
    std::vector<unsigned> grid;
    std::vector<const unsigned char*> packs;
    for (const MergedItem& it : items) {
        const ROp& op = *sel[it.first].second;
        hipError_t e;
        if (it.kind == ROp::kKernel) {
            grid.clear();
            packs.clear();
            for (size_t q = 0; q < it.n; ++q) {
                const auto& pr = sel[it.first + q];
                grid.push_back(pr.second->grid);
                packs.push_back(pr.first->rpack.data() + pr.second->pack_off);
            }
            e = op.issue(op.block, op.shm, R.stream, (int)it.n, grid.data(), packs.data());
            R.stats.launches++;
            R.stats.tank_launches += (int64_t)it.n;
        } else if (it.kind == ROp::kCopy) {
            e = hipMemcpyAsync(op.dst, op.src, op.bytes, hipMemcpyDeviceToHost, R.stream);
        } else {
            e = hipEventRecord(op.ev, R.stream);
        }
        if (e != hipSuccess) return e;
    }
    const auto t0 = std::chrono::steady_clock::now();
    const hipError_t es = hipStreamSynchronize(R.stream);
    R.stats.wait_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return es;
}

// With R.m held, when every stepping tank waits (or the last one left with work recorded).
void flush_locked(BatchRun& R) {
    std::vector<vare_flip*> ms;
    for (vare_flip* h : R.members)
        if (!h->rops.empty()) ms.push_back(h);
    hipError_t e = hipSuccess;
    const auto t0 = std::chrono::steady_clock::now();
    if (R.last_flush_end.time_since_epoch().count())
        R.stats.gap_ms += std::chrono::duration<double, std::milli>(t0 - R.last_flush_end).count();
    if (!ms.empty()) e = run_merged(R, ms);
    R.stats.flush_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    R.stats.flushes++;
    for (vare_flip* h : ms) {
        if (e != hipSuccess) h->rec_rc = VARE_EHIP;
        h->rops.clear();
        h->rpack.clear();
    }
    for (vare_flip* h : R.members) {
        for (void* p : h->rfree) (void)hipFree(p);
        h->rfree.clear();
    }
    if (e != hipSuccess) (void)hipGetLastError();
    R.waiting = 0;
    R.last_flush_end = std::chrono::steady_clock::now();
    R.gen.fetch_add(1, std::memory_order_release);
}

// Waits until the flush counter moves past g.  The waiters do not take the lock again on the way
// out (a condition variable would wake them one after the other through the mutex): they spin,
// then yield, then sleep briefly.
void spin_until_past(const std::atomic<uint64_t>& gen, uint64_t g) {
    for (int i = 0; gen.load(std::memory_order_acquire) == g; ++i) {
        if (i < 2048) std::this_thread::yield();
        else std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

int batch_sync(vare_flip* h) {
    BatchRun& R = *h->rec;
    uint64_t g = 0;
    bool flushed = false;
    {
        std::lock_guard<std::mutex> lk(R.m);
        R.waiting++;
        if (R.waiting == R.active) {
            flush_locked(R);
            flushed = true;
        } else {
            g = R.gen.load(std::memory_order_relaxed);
        }
    }
    if (!flushed) spin_until_past(R.gen, g);
    if (h->rec_rc != VARE_OK) return fail(h->rec_rc, "a merged launch of the batched step failed");
    return VARE_OK;
}

int batch_chunk_hint(const vare_flip* h, int k) { return h->rec->chunk_hint[std::min(k, 19)]; }

// The tank's step has returned: the others no longer wait for it.
void batch_leave(vare_flip* h) {
    BatchRun& R = *h->rec;
    std::unique_lock<std::mutex> lk(R.m);
    R.active--;
    bool pending = false;
    for (vare_flip* m : R.members) pending = pending || !m->rops.empty() || !m->rfree.empty();
    if (R.waiting == R.active && (R.waiting > 0 || pending)) flush_locked(R);
}

void batch_worker(BatchRun* R, int idx) {
    uint64_t seen = 0;
    for (;;) {
        std::unique_lock<std::mutex> lk(R->pm);
        R->pcv.wait(lk, [&] { return R->job != seen; });
        seen = R->job;
        if (idx >= R->njob) continue;
        vare_flip* h = R->jh[idx];
        const float dt = R->jdt[idx];
        lk.unlock();
        (void)hipSetDevice(h->device);
        const int rc = vare_flip_step(h, dt);
        const char* err = vare_last_error();
        batch_leave(h);
        lk.lock();
        R->jrc[idx] = rc;
        R->jerr[(size_t)idx] = rc == VARE_OK ? std::string() : std::string(err ? err : "");
        if (--R->left == 0) R->pdone.notify_all();
    }
}

std::mutex g_batch_mu;
BatchRun* g_batch[64] = {};

// The device's runner (created on first use and kept until the process ends; its threads wait).
int batch_runner(int dev, BatchRun** out) {
    if (dev < 0 || dev >= 64) return fail(VARE_EINVAL, "device %d", dev);
    std::lock_guard<std::mutex> lk(g_batch_mu);
    if (!g_batch[dev]) {
        BatchRun* R = new BatchRun();
        R->device = dev;
        int cur = -1;
        (void)hipGetDevice(&cur);
        hipError_t e = hipSetDevice(dev);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&R->stream, hipStreamNonBlocking);
        if (cur >= 0 && cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess) {
            delete R;
            return fail(VARE_EHIP, "batch stream: %s", hipGetErrorString(e));
        }
        g_batch[dev] = R;
    }
    *out = g_batch[dev];
    return VARE_OK;
}

}  // namespace

extern "C" {

int vare_flip_step_batched(vare_flip* const* hs, int n, const float* dt, int* status) {
    if (!hs || !dt || n < 1) return fail(VARE_EINVAL, "vare_flip_step_batched: no tanks");
    for (int i = 0; i < n; ++i) {
        if (!hs[i]) return fail(VARE_EINVAL, "vare_flip_step_batched: null handle %d", i);
        if (hs[i]->timing) return fail(VARE_EINVAL, "vare_flip_step_batched: kernel timing is on for tank %d", i);
        if (hs[i]->rec) return fail(VARE_EINVAL, "vare_flip_step_batched: tank %d is already stepping in a batch", i);
        if (hs[i]->device != hs[0]->device) return fail(VARE_EINVAL, "vare_flip_step_batched: tanks on different devices");
        for (int j = 0; j < i; ++j)
            if (hs[j] == hs[i]) return fail(VARE_EINVAL, "vare_flip_step_batched: tank %d given twice", i);
    }
    BatchRun* R = nullptr;
    CHK(batch_runner(hs[0]->device, &R));
    std::lock_guard<std::mutex> call(R->call);
    const auto t_call = std::chrono::steady_clock::now();
    R->last_flush_end = std::chrono::steady_clock::time_point{};
    // the handles' own streams may still hold work (uploads): it comes first
    for (int i = 0; i < n; ++i) HIPCHK(hipStreamSynchronize(hs[i]->stream));
    std::vector<int> rc((size_t)n, VARE_OK);
    {
        std::lock_guard<std::mutex> lk(R->m);
        R->members.assign(hs, hs + n);
        for (int k = 0; k < 20; ++k) {
            R->chunk_hint[k] = 0;
            for (int i = 0; i < n; ++i) R->chunk_hint[k] = std::max(R->chunk_hint[k], hs[i]->last_iters[k]);
        }
        R->active = n;
        R->waiting = 0;
        for (int i = 0; i < n; ++i) {
            hs[i]->rec = R;
            hs[i]->rec_rc = VARE_OK;
        }
    }
    {
        std::unique_lock<std::mutex> lk(R->pm);
        while ((int)R->threads.size() < n) {
            R->threads.emplace_back(batch_worker, R, (int)R->threads.size());
            R->threads.back().detach();
        }
        R->jh = hs;
        R->jdt = dt;
        R->jrc = rc.data();
        R->jerr.assign((size_t)n, std::string());
        R->njob = n;
        R->left = n;
        R->job++;
        R->pcv.notify_all();
        R->pdone.wait(lk, [&] { return R->left == 0; });
        R->njob = 0;
    }
    {
        std::lock_guard<std::mutex> lk(R->m);
        for (int i = 0; i < n; ++i) {
            hs[i]->rec = nullptr;
            for (void* p : hs[i]->rfree) (void)hipFree(p);     // none left after the last flush
            hs[i]->rfree.clear();
        }
        R->members.clear();
        R->stats.steps++;
        R->stats.tanks += n;
        R->stats.step_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_call).count();
    }
    int first = VARE_OK;
    for (int i = 0; i < n; ++i) {
        if (status) status[i] = rc[(size_t)i];
        if (first == VARE_OK && rc[(size_t)i] != VARE_OK) {
            first = rc[(size_t)i];
            g_err = R->jerr[(size_t)i];
        }
    }
    return first;
}

int vare_debug_batch_plan(const int32_t* keys, const int32_t* lens, int ntank, int32_t* out_tank, int32_t* out_pos,
                          int32_t* out_item, size_t cap, size_t* nout) {
    if (!lens || ntank < 1 || !nout) return fail(VARE_EINVAL, "vare_debug_batch_plan: no sequences");
    // stand-in tanks whose recorded operations carry only the keys (no device state is touched)
    std::vector<std::unique_ptr<vare_flip>> tanks;
    std::vector<vare_flip*> ms;
    size_t total = 0, off = 0;
    for (int t = 0; t < ntank; ++t) {
        if (lens[t] < 0) return fail(VARE_EINVAL, "vare_debug_batch_plan: negative length");
        tanks.emplace_back(new vare_flip());
        for (int q = 0; q < lens[t]; ++q) {
            const int32_t k = keys[off + (size_t)q];
            ROp op;
            if (k < 0) {
                op.kind = k == -1 ? ROp::kCopy : ROp::kEvent;
            } else {
                op.kind = ROp::kKernel;
                op.issue = reinterpret_cast<IssueFn>((uintptr_t)(k + 1) * 16);
                op.max_tanks = kBatchMaxTanks;
            }
            tanks.back()->rops.push_back(op);
        }
        off += (size_t)lens[t];
        total += (size_t)lens[t];
        ms.push_back(tanks.back().get());
    }
    *nout = total;
    if (!out_tank || !out_pos || !out_item) return VARE_OK;
    if (cap < total) return fail(VARE_EINVAL, "vare_debug_batch_plan: capacity %zu < %zu", cap, total);
    std::vector<std::pair<vare_flip*, const ROp*>> sel;
    std::vector<MergedItem> items;
    plan_merged(ms, sel, items);
    for (size_t it = 0; it < items.size(); ++it)
        for (size_t q = 0; q < items[it].n; ++q) {
            const auto& pr = sel[items[it].first + q];
            const size_t e = items[it].first + q;
            const int t = (int)(std::find(ms.begin(), ms.end(), pr.first) - ms.begin());
            out_tank[e] = t;
            out_pos[e] = (int32_t)(pr.second - pr.first->rops.data());
            out_item[e] = (int32_t)it;
        }
    return VARE_OK;
}

int vare_batch_get_stats(int device, vare_batch_stats* out, int reset) {
    if (!out) return fail(VARE_EINVAL, "null output");
    BatchRun* R = nullptr;
    CHK(batch_runner(device, &R));
    std::lock_guard<std::mutex> call(R->call);
    *out = R->stats;
    if (reset) std::memset(&R->stats, 0, sizeof R->stats);
    return VARE_OK;
}

}  // extern "C"
