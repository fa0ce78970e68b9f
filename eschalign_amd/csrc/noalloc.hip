// esch_score_device_noalloc: esch_score_batch_i32's scoring for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc). The caller passes the
// reads and the output in device memory and a stream; every other buffer
// the kernels need (the batch descriptor, the length-bucketing bins and
// lists, the stripe boundaries of the DP kernels) lives in one file-scope
// device array of fixed size, `scratch`, whose address is resolved once per
// device. Nothing here allocates: the call only launches kernels on the
// caller's stream, so it can itself be captured into a graph.
//
// Calls that overlap on one GPU share that GPU's array and race: a caller
// serialises them (one stream, or a wait between calls). Each call clears
// what must start at zero (k_zero: bins and counters) and rewrites the
// descriptor by a kernel, never by hipMemsetAsync or a copy.
#include "host_internal.h"

#include <algorithm>
#include <atomic>
#include <mutex>

namespace {

using namespace esch;

constexpr size_t kScratchBytes = ESCH_NOALLOC_SCRATCH_BYTES;
constexpr size_t kMaxRows = ESCH_NOALLOC_MAX_ROWS;
constexpr int kMaxDevices = 64;

// The only device memory this path uses besides the caller's. Each GPU the
// library runs this path on holds its own copy until the process ends.
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

// Per device, the address of that device's `scratch` (hipGetSymbolAddress
// with the device current): host code uses it, never the array's name,
// which on the host is only a stand-in for the symbol.
std::atomic<uint8_t *> g_scratch[kMaxDevices];
std::mutex g_init_mu;

esch_status scratch_on(int ordinal, uint8_t **out) {
    if (ordinal < 0 || ordinal >= kMaxDevices)
        return fail(ESCH_INVALID_ARG, "device %d outside the allocation-free path's 0..%d", ordinal, kMaxDevices - 1);
    uint8_t *p = g_scratch[ordinal].load(std::memory_order_acquire);
    if (!p) {
        std::lock_guard<std::mutex> g(g_init_mu);
        p = g_scratch[ordinal].load(std::memory_order_relaxed);
        if (!p) {
            void *addr = nullptr;
            const hipError_t e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !addr)
                return fail(ESCH_DRIVER, "hipGetSymbolAddress(scratch) on device %d: %s", ordinal,
                            hipGetErrorString(e));
            p = static_cast<uint8_t *>(addr);
            g_scratch[ordinal].store(p, std::memory_order_release);
        }
    }
    *out = p;
    return ESCH_OK;
}

// The run's descriptor, written on the device (no copy into memory a graph
// may replay): every field esch_batch_run sets for an int32 scoring run
// without selection; k_zero clears the counters after it.
struct Plan {
    int32_t c[kLenClasses];
};

__global__ void k_noalloc_desc(DevDesc *d, int32_t n, int32_t ratio, int32_t blocks16, Plan plan) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    d->n_queries = n;
    d->n_reads = n;
    d->both = 0;
    d->pairs = 0;
    d->dropin = 0;
    d->force32 = 0;
    d->min_score = 0;
    d->tie_cap = 1;
    d->max_ties = -1;
    d->tstrands = 1;
    d->diag_min_len = 0;
    d->diag_ratio = ratio;
    d->blocks16 = blocks16;
    for (int c = 0; c < kLenClasses; ++c) d->plan[c] = plan.c[c];
}

// scored[q] = 1 when query q was scored (class 1, 2 or 3), else 0.
__global__ void k_noalloc_scored(const uint8_t *cls, uint8_t *scored, int32_t n) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
        scored[q] = cls[q] ? 1 : 0;
}

constexpr size_t align256(size_t v) { return (v + 255) / 256 * 256; }

}  // namespace

extern "C" esch_status esch_score_device_noalloc(esch_scorer *s, const uint8_t *codes, const uint64_t *offs,
                                                 const int32_t *lens, size_t n, int32_t *scores, uint8_t *scored,
                                                 void *stream) {
    if (!s) return fail(ESCH_INVALID_ARG, "null scorer");
    if (n > kMaxRows)
        return fail(ESCH_CAPACITY, "%zu queries exceed the allocation-free path's %zu", n, kMaxRows);
    if (n && (!codes || !offs || !lens || !scores || !scored)) return fail(ESCH_INVALID_ARG, "null argument");
    int cur = -1;
    HIPC(hipGetDevice(&cur));
    if (cur != s->ordinal) HIPC(hipSetDevice(s->ordinal));
    uint8_t *base = nullptr;
    if (esch_status st = scratch_on(s->ordinal, &base)) return st;
    if (n == 0) return ESCH_OK;  // a call with n = 0 only sets the device up

    Launch geo;
    int ratio;
    Plan plan;
    {
        std::lock_guard<std::mutex> g(s->geo_mu);
        geo = s->geo;
        ratio = s->diag_ratio;
        for (int c = 0; c < kLenClasses; ++c) plan.c[c] = s->diag_plan[c];
    }
    const size_t L = s->max_len;
    // layout: descriptor, length bins, per-row class and lists, then one
    // region the DP kernels share in turn (k_dp16, then the anti-diagonal
    // kernel, then k_dp32, in stream order), sized to what is left
    const size_t off_hist = align256(sizeof(DevDesc));
    const size_t bins = (size_t)(kMaxLen16 + 1) * 4;
    const size_t off_start = off_hist + align256(bins);
    const size_t off_fill = off_start + align256(bins);
    const size_t off_cls = off_fill + align256(bins);
    const size_t off_sorted = off_cls + align256(kMaxRows);
    const size_t off_list32 = off_sorted + align256(kMaxRows * 4);
    const size_t off_listD = off_list32 + align256(kMaxRows * 4);
    const size_t off_dp = off_listD + align256(kMaxRows * 4);
    const size_t dp_bytes = kScratchBytes - off_dp;
    // persistent grids cut to the region: k_dp16 takes (L + kScrPad) x 64
    // boundary slots per wave, k_dp32 L per thread, k_diag16 L per row task
    const size_t per16 = (size_t)4 * (L + kScrPad) * 64 * sizeof(uint2);
    const size_t per32 = (size_t)geo.block32 * L * sizeof(int2);
    const size_t perD = geo.diag_lds ? 0 : (size_t)4 * L * sizeof(uint2);
    const int grid16 = (int)std::min<size_t>((size_t)s->grid16_v[geo.variant16], dp_bytes / per16);
    const int grid32 = (int)std::min<size_t>((size_t)geo.grid32, dp_bytes / per32);
    const int gridD = perD ? (int)std::min<size_t>((size_t)geo.grid_diag, dp_bytes / perD) : geo.grid_diag;
    if (grid16 < 1 || grid32 < 1 || gridD < 1)
        return fail(ESCH_CAPACITY,
                    "the panel's longest reference (%zu nt) needs more boundary scratch than the "
                    "allocation-free path's %zu bytes hold",
                    L, kScratchBytes);
    geo.grid16 = grid16;
    geo.grid32 = grid32;
    geo.grid_diag = gridD;

    KArgs a{};
    a.panel = s->panel;
    a.sc = s->sc;
    a.desc = reinterpret_cast<DevDesc *>(base);
    a.qcodes = codes;
    a.qoff = offs;
    a.qlen = lens;
    a.cls = base + off_cls;
    a.hist = reinterpret_cast<int32_t *>(base + off_hist);
    a.bin_start = reinterpret_cast<int32_t *>(base + off_start);
    a.bin_fill = reinterpret_cast<int32_t *>(base + off_fill);
    a.sorted16 = reinterpret_cast<int32_t *>(base + off_sorted);
    a.list32 = reinterpret_cast<int32_t *>(base + off_list32);
    a.listD = reinterpret_cast<int32_t *>(base + off_listD);
    a.out32 = scores;
    a.out16 = nullptr;
    a.scratch16 = reinterpret_cast<uint2 *>(base + off_dp);
    a.scratch32 = reinterpret_cast<int2 *>(base + off_dp);
    a.scratchD = reinterpret_cast<uint2 *>(base + off_dp);
    a.rc_codes = nullptr;  // desc->both = 0: k_revcomp writes nothing
    a.swap = 0;

    hipStream_t st = static_cast<hipStream_t>(stream);
    const bool local = s->sc.local != 0;
    hipLaunchKernelGGL(k_noalloc_desc, dim3(1), dim3(64), 0, st, a.desc, (int32_t)n, (int32_t)ratio,
                       (int32_t)grid16, plan);
    HIPC(hipGetLastError());
    HIPC(launch_prep(a, geo, st));
    HIPC(launch_dp16(a, geo, local, st));
    HIPC(launch_diag(a, geo, local, st));
    HIPC(launch_dp32(a, geo, local, 1, st));
    const int gs = (int)std::min<size_t>((n + 255) / 256, (size_t)geo.grid_aux);
    hipLaunchKernelGGL(k_noalloc_scored, dim3(gs), dim3(256), 0, st, a.cls, scored, (int32_t)n);
    HIPC(hipGetLastError());
    return ESCH_OK;
}
