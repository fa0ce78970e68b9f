// Host-side internals shared by the library's translation units: the
// scorer and batch objects behind the opaque C-ABI handles, grow-only
// device and pinned buffers, the error channel and the batch pool.
#pragma once

#include "../../include/eschalign_amd.h"
#include "esch_internal.h"

#include <hip/hip_runtime.h>

#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace esch {

// Sets the calling thread's esch_last_error() text and returns st.
esch_status fail(esch_status st, const char *fmt, ...);

#define HIPC(call)                                                                     \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ::esch::fail(ESCH_DRIVER, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

}  // namespace esch

// the types of the kernels' interface, as the structs below use them
using esch::DevDesc;
using esch::DevPanel;
using esch::Launch;
using esch::Scoring;
using esch::kDiagRatioDefault;
using esch::kDp16Variants;
using esch::kDp16F16;
using esch::fits_f16;

struct TraceCache;
struct MapCache;
void trace_cache_free(esch_scorer *s);
void map_cache_free(esch_scorer *s);

struct esch_scorer {
    int ordinal = 0;
    std::string device_name;
    int n_cu = 0;
    Scoring sc{};
    esch_mode mode = ESCH_LOCAL;
    size_t n_refs = 0, max_len = 0;
    int64_t total_len = 0;
    uint32_t *d_ref2 = nullptr, *d_refN = nullptr, *d_w2off = nullptr, *d_nmoff = nullptr;
    int32_t *d_rlen = nullptr, *d_rorder = nullptr;
    DevPanel panel{};
    Launch geo{};
    int grid16_v[kDp16Variants] = {};   // persistent grid of each k_dp16 variant
    int regs16_v[kDp16Variants] = {}, local16_v[kDp16Variants] = {};
    int max_grid16 = 0;                 // scratch is sized for the largest
    int regs16 = 0, local16 = 0;
    bool f16_ok = false;             // the panel fits k_dp16's fp16 variant (fits_f16)
    bool usable16[kDp16Variants] = {};  // variants exact on this panel (fp16 and shift bounds)
    // geo (its variant and grid), regs16 / local16, variant_fixed and the
    // routing ratio below are written by esch_set_variant, esch_autotune and
    // esch_set_diag_ratio, possibly while another thread runs batches of this
    // scorer (a server's scheduler): both sides take geo_mu, and a run works
    // from a snapshot taken under it.
    mutable std::mutex geo_mu;
    bool variant_fixed = false;         // set by the caller (esch_set_variant): no per-batch upgrade
    int diag_min_len = 0;               // reads this long take k_diag16 (0: never)
    int diag_ratio = kDiagRatioDefault; // padding-heavy length bins take k_diag16 (0: never)
    bool diag_ratio_tunable = true;     // false once the caller set the ratio or the plan
    int32_t diag_plan[esch::kLenClasses] = {kDiagRatioDefault, kDiagRatioDefault, kDiagRatioDefault,
                                            kDiagRatioDefault, kDiagRatioDefault, kDiagRatioDefault,
                                            kDiagRatioDefault, kDiagRatioDefault};  // per length class
    std::mutex pool_mu;
    std::vector<esch_batch *> pool;  // idle batches for the drop-in API
    std::vector<int32_t> h_rlen;     // reference lengths (traceback sizing)
    float trace_ms = 0.0f;           // device time of the last esch_align_pairs kernels
    size_t trace_diag = 0;           // pairs of the last call traced by the wavefront kernel
    TraceCache *trace = nullptr;     // traceback buffers kept between calls
    MapCache *map = nullptr;         // mapping buffers kept between calls
    size_t map_chunk = 0;            // reads per mapping chunk (0: by the rows' memory)
    // the references as queries, for runs with the roles swapped
    uint8_t *d_pq_codes = nullptr;
    uint64_t *d_pq_off = nullptr;
    int32_t *d_pq_len = nullptr;
    int32_t swap_rows = -1;          // swap roles up to this many rows (-1: by the rule)
};

// Tie capacity (tie indices per read) a batch allocates when it is created;
// a run asking for more reallocates, which waits for the device.
constexpr size_t kTieCapAlloc = 8;

struct SwapBufs;

struct esch_batch {
    esch_scorer *s = nullptr;
    size_t max_reads = 0, max_bases = 0;  // capacity before strand expansion
    size_t longest = 0;                   // the loaded reads' longest length
    int last_variant = -1;                // k_dp16 variant of the last run (-1: none, or swapped)
    size_t tie_cap_alloc = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
    // device
    uint8_t *d_codes = nullptr;
    uint64_t *d_qoff = nullptr;
    int32_t *d_qlen = nullptr;
    uint8_t *d_cls = nullptr;
    int32_t *d_hist = nullptr, *d_bin_start = nullptr, *d_bin_fill = nullptr;
    int32_t *d_sorted16 = nullptr, *d_list32 = nullptr;
    int32_t *d_out32 = nullptr;
    int16_t *d_out16 = nullptr;
    uint2 *d_scr16 = nullptr;
    int2 *d_scr32 = nullptr;
    int32_t *d_recs = nullptr, *d_ties = nullptr;
    int32_t *d_listD = nullptr;
    uint2 *d_scrD = nullptr;
    DevDesc *d_desc = nullptr;
    // pinned staging
    uint8_t *h_codes = nullptr;
    uint64_t *h_qoff = nullptr;
    int32_t *h_qlen = nullptr;
    DevDesc *h_desc = nullptr;
    // last load / run
    size_t n_reads = 0, n_rows = 0;
    int both = 0, pairs = 0;
    int last_what = 0;
    bool timed_dp = false;
    std::map<int, hipGraphExec_t> graphs;
    SwapBufs *swap = nullptr;   // the swapped path's buffers (allocated on first use)
    bool last_swapped = false;  // the last run scored with the roles swapped
    // a sliced run (esch::batch_run_sliced): k_dp16 in n_slices launches,
    // slices_out of them queued so far; the kernel variant and run flags
    // they launch with, and a start / end event per slice
    int n_slices = 0, slices_out = 0;
    esch::Launch slice_geo{};
    int slice_what = 0;
    hipEvent_t sev[2 * esch::kMaxSlices] = {};
    hipEvent_t pump_ev = nullptr;
};

namespace esch {

// Grow-only device buffer of the traceback cache.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t need(size_t bytes) {
        bytes = std::max<size_t>(bytes, 256);
        if (bytes <= cap) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        const hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) cap = bytes;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

// Grow-only pinned host buffer (staging of the traceback's largest copies).
struct HostBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t need(size_t bytes) {
        bytes = std::max<size_t>(bytes, 256);
        if (bytes <= cap) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        const hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
        if (e == hipSuccess) cap = bytes;
        return e;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

// A batch from the scorer's idle pool that holds n reads / bases (a new one
// when none does), and its return to the pool.
esch_status take_batch(esch_scorer *s, size_t n, size_t bases, esch_batch **out);
void give_batch(esch_scorer *s, esch_batch *b);

// The traceback of n (read, reference) pairs; read k is get(k) ->
// (codes, length). With dev_codes set, get returns device addresses inside
// dev_codes (reads already resident, e.g. a batch's codes; the caller has
// synchronised the stream that wrote them) and nothing is staged. With
// ops_vec set, the runs go there (resized to fit) instead of ops / ops_cap.
// With pump set, each wait for the device first queues pump's next slice
// (batch_pump) after the traceback's work.
esch_status align_core(esch_scorer *s, size_t n, const std::function<void(size_t, const uint8_t *&, size_t &)> &get,
                       const int32_t *refs, esch_alignment *out, uint32_t *ops, size_t ops_cap, size_t *ops_used,
                       const uint8_t *dev_codes = nullptr, std::vector<uint32_t> *ops_vec = nullptr,
                       esch_batch *pump = nullptr);

// Sliced runs. k_dp16 is persistent: once launched it holds every CU until
// its last work item, so a short kernel of another stream waits for the
// whole scoring. A sliced run cuts k_dp16's work items into n_slices
// launches (at most kMaxSlices): the run queues the bucketing and slice 0;
// batch_pump queues the next slice behind the work that stream `after` has
// queued so far (a mapping call's selection or tracebacks of the previous
// chunk), which so runs between two slices; the last slice brings the rest
// of the run (the anti-diagonal and int32 kernels). Every wait for the
// batch, read of its results, reload or new run first queues the slices
// still out (batch_drain). The swapped-roles path and the plain kernel run
// unsliced (the pump is then a no-op). Not captured in a graph.
esch_status batch_run_sliced(esch_batch *b, int what, int n_slices);
esch_status batch_pump(esch_batch *b, hipStream_t after);
esch_status batch_drain(esch_batch *b);

}  // namespace esch
