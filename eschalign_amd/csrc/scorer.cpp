// Host side of the C ABI (include/eschalign_amd.h): the panel resident on
// the device, device batches with their own stream, scratch and captured
// graphs, and the drop-in prepare / submit / finish of GpuScorer
// (lib/cuda/mod.rs:328-749).
#include "../../include/eschalign_amd.h"
#include "esch_internal.h"
#include "host_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

using namespace esch;

namespace {
thread_local std::string g_err;
}  // namespace

namespace esch {

esch_status fail(esch_status st, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return st;
}

}  // namespace esch

namespace {

int32_t iabs(int32_t v) { return v < 0 ? -v : v; }

// lib/scoring.rs:70-100
esch_status validate(int32_t ma, int32_t mi, int32_t o, int32_t e) {
    if (ma <= 0) return fail(ESCH_SCORING_INVALID, "invalid scoring: match must be positive, got %d", ma);
    if (mi >= ma)
        return fail(ESCH_SCORING_INVALID, "invalid scoring: mismatch (%d) must score below match (%d)", mi, ma);
    if (o > 0 || e > 0)
        return fail(ESCH_SCORING_INVALID,
                    "invalid scoring: gap scores are penalties and must be zero or negative, got open %d extend %d",
                    o, e);
    for (int32_t v : {ma, mi, o, e})
        if (iabs(v) > 1000)
            return fail(ESCH_SCORING_INVALID, "invalid scoring: |%d| exceeds the supported magnitude 1000", v);
    return ESCH_OK;
}

bool fits_i16(size_t n, size_t L, int32_t ma, int32_t mi, int32_t o, int32_t e) {
    const int64_t k = (int64_t)std::min(n, L);
    const int64_t M = std::max(iabs(ma), iabs(mi));
    return k * M + iabs(o) + iabs(e) < 15000 && L < 15000;
}

}  // namespace


struct esch_prepared {
    size_t n = 0;
    std::vector<uint8_t> codes;
    std::vector<uint64_t> off;
    std::vector<int32_t> len;
    std::vector<uint8_t> scored;
};

struct esch_pending {
    esch_batch *b = nullptr;
    size_t n = 0;
    std::vector<uint8_t> scored;
};

extern "C" {

const char *esch_last_error(void) { return g_err.c_str(); }

int esch_device_visible(void) {
    static std::once_flag once;
    static int visible = 0;
    std::call_once(once, [] {
        int n = 0;
        visible = (hipGetDeviceCount(&n) == hipSuccess && n > 0) ? 1 : 0;
    });
    return visible;
}

esch_status esch_scoring_validate(int32_t ma, int32_t mi, int32_t o, int32_t e) {
    return validate(ma, mi, o, e);
}

int esch_fits_i16(size_t n, size_t L, int32_t ma, int32_t mi, int32_t o, int32_t e) {
    return fits_i16(n, L, ma, mi, o, e) ? 1 : 0;
}

esch_status esch_scorer_new(int ordinal, const uint8_t *const *refs, const size_t *ref_lens,
                            size_t n_refs, int32_t ma, int32_t mi, int32_t o, int32_t e,
                            esch_mode mode, esch_scorer **out) {
    if (!out) return fail(ESCH_INVALID_ARG, "out is null");
    *out = nullptr;
    // the mode's integer, read as stored: a caller across the C ABI may pass
    // any int, and a value outside the enumeration is not one C++ may load
    int mode_i = 0;
    static_assert(sizeof mode_i == sizeof mode, "esch_mode is an int");
    memcpy(&mode_i, &mode, sizeof mode_i);
    if (mode_i != ESCH_LOCAL && mode_i != ESCH_SEMIGLOBAL) return fail(ESCH_INVALID_ARG, "unknown mode %d", mode_i);
    if (esch_status st = validate(ma, mi, o, e)) return st;
    if (n_refs == 0) return fail(ESCH_EMPTY_PANEL, "the reference panel is empty");
    if (!refs || !ref_lens) return fail(ESCH_INVALID_ARG, "refs or ref_lens is null");
    size_t max_len = 0;
    int64_t total = 0;
    for (size_t i = 0; i < n_refs; ++i) {
        if (ref_lens[i] == 0) return fail(ESCH_EMPTY_REFERENCE, "reference %zu has no bases", i);
        if (!refs[i]) return fail(ESCH_INVALID_ARG, "reference %zu is null", i);
        for (size_t j = 0; j < ref_lens[i]; ++j)
            if (refs[i][j] > 4) return fail(ESCH_INVALID_ARG, "reference %zu carries code %d at %zu (codes are 0..=4)", i, refs[i][j], j);
        max_len = std::max(max_len, ref_lens[i]);
        total += (int64_t)ref_lens[i];
    }
    if (max_len > ESCH_MAX_REF_LEN)
        return fail(ESCH_REFERENCE_TOO_LONG,
                    "the panel's longest reference (%zu nt) exceeds the GPU kernel's %d nt", max_len,
                    ESCH_MAX_REF_LEN);
    if (!esch_device_visible()) return fail(ESCH_UNAVAILABLE, "no HIP device is visible");
    int ndev = 0;
    HIPC(hipGetDeviceCount(&ndev));
    if (ordinal < 0 || ordinal >= ndev) return fail(ESCH_UNAVAILABLE, "no HIP device %d (%d visible)", ordinal, ndev);
    HIPC(hipSetDevice(ordinal));
    hipDeviceProp_t prop;
    HIPC(hipGetDeviceProperties(&prop, ordinal));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(ESCH_CODE_OBJECT, "device %d is %s; this library carries gfx950 code objects only", ordinal,
                    prop.gcnArchName);

    std::unique_ptr<esch_scorer> s(new esch_scorer);
    s->ordinal = ordinal;
    s->device_name = prop.name[0] ? prop.name : prop.gcnArchName;
    s->n_cu = prop.multiProcessorCount;
    s->sc = Scoring{ma, mi, o, e, mode == ESCH_LOCAL ? 1 : 0};
    s->mode = mode;
    s->n_refs = n_refs;
    s->max_len = max_len;
    s->total_len = total;

    // 2-bit codes, 16 per word (a wildcard stores 0 and sets its bit in the
    // wildcard bitmap, 32 per word); every reference starts on a new word.
    std::vector<uint32_t> ref2, refN, w2off(n_refs), nmoff(n_refs);
    std::vector<int32_t> rlen(n_refs);
    for (size_t i = 0; i < n_refs; ++i) {
        const size_t L = ref_lens[i];
        w2off[i] = (uint32_t)ref2.size();
        nmoff[i] = (uint32_t)refN.size();
        rlen[i] = (int32_t)L;
        ref2.resize(ref2.size() + (L + 15) / 16 + 1, 0u);
        refN.resize(refN.size() + (L + 31) / 32 + 1, 0u);
        for (size_t j = 0; j < L; ++j) {
            const uint8_t c = refs[i][j];
            if (c == 4) refN[nmoff[i] + j / 32] |= 1u << (j % 32);
            else ref2[w2off[i] + j / 16] |= (uint32_t)c << (2 * (j % 16));
        }
    }
    HIPC(hipMalloc(&s->d_ref2, ref2.size() * 4));
    HIPC(hipMalloc(&s->d_refN, refN.size() * 4));
    HIPC(hipMalloc(&s->d_w2off, n_refs * 4));
    HIPC(hipMalloc(&s->d_nmoff, n_refs * 4));
    HIPC(hipMalloc(&s->d_rlen, n_refs * 4));
    HIPC(hipMemcpy(s->d_ref2, ref2.data(), ref2.size() * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(s->d_refN, refN.data(), refN.size() * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(s->d_w2off, w2off.data(), n_refs * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(s->d_nmoff, nmoff.data(), n_refs * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(s->d_rlen, rlen.data(), n_refs * 4, hipMemcpyHostToDevice));
    s->h_rlen = rlen;
    // Longest first, stable (as lib/simd/mod.rs:299-300 orders its lane
    // groups): the four references of one workgroup then have nearly equal
    // lengths and finish their stripes together.
    std::vector<int32_t> order(n_refs);
    for (size_t i = 0; i < n_refs; ++i) order[i] = (int32_t)i;
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return rlen[x] > rlen[y]; });
    HIPC(hipMalloc(&s->d_rorder, n_refs * 4));
    HIPC(hipMemcpy(s->d_rorder, order.data(), n_refs * 4, hipMemcpyHostToDevice));
    s->panel = DevPanel{s->d_ref2, s->d_refN, s->d_w2off, s->d_nmoff, s->d_rlen, s->d_rorder,
                        (int32_t)n_refs, (int32_t)max_len, total};
    // The references as queries (byte codes, offsets, lengths), for batches
    // scored with the roles of reads and references swapped.
    {
        std::vector<uint8_t> codes((size_t)total);
        std::vector<uint64_t> off(n_refs);
        size_t at = 0;
        for (size_t i = 0; i < n_refs; ++i) {
            off[i] = at;
            memcpy(codes.data() + at, refs[i], ref_lens[i]);
            at += ref_lens[i];
        }
        HIPC(hipMalloc(&s->d_pq_codes, std::max<size_t>(at, 1)));
        HIPC(hipMalloc(&s->d_pq_off, n_refs * 8));
        HIPC(hipMalloc(&s->d_pq_len, n_refs * 4));
        HIPC(hipMemcpy(s->d_pq_codes, codes.data(), at, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(s->d_pq_off, off.data(), n_refs * 8, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(s->d_pq_len, rlen.data(), n_refs * 4, hipMemcpyHostToDevice));
    }

    // Launch geometry: persistent grids sized by occupancy.
    const bool local = mode == ESCH_LOCAL;
    for (int v = 0; v < kDp16Variants; ++v) {
        hipFuncAttributes attr;
        hipError_t ae = dp16_attributes(local, v, &attr);
        if (ae != hipSuccess)
            return fail(ESCH_CODE_OBJECT, "no gfx950 code object for the score kernel: %s", hipGetErrorString(ae));
        s->regs16_v[v] = attr.numRegs;
        s->local16_v[v] = (int)attr.localSizeBytes;
        int b16 = 0;
        HIPC(dp16_occupancy(local, v, &b16));
        s->grid16_v[v] = std::max(1, b16) * s->n_cu;
        s->max_grid16 = std::max(s->max_grid16, s->grid16_v[v]);
    }
    int b32 = 0;
    int block32 = (int)std::min<size_t>(256, ((n_refs + 63) / 64) * 64);
    HIPC(dp32_occupancy(local, block32, &b32));
    // the default variant (chosen below: the fewest ops per cell the panel
    // admits); esch_autotune re-decides on the caller's reads
    s->f16_ok = fits_f16((int64_t)max_len, s->sc.ma, s->sc.mi, s->sc.o, s->sc.e);
    {
        // The shifted variants: rows shifted in semiglobal mode, columns in
        // local mode, both in the diagonal shift (semiglobal mode only). The
        // column part restarts every 2^cshift columns where the whole
        // reference does not fit the arithmetic's bound (column_block_log2;
        // the diagonal shift's taken at 32 rows, which covers 24); -1: no
        // block fits and the variants are refused.
        const auto block = [&](int64_t lim) {
            return esch::column_block_log2(local, (int64_t)max_len, (int64_t)max_len, s->sc.ma, s->sc.mi, s->sc.o,
                                           s->sc.e, 32, lim);
        };
        s->geo.cshift_f16 = block(esch::kBoundF16);
        s->geo.cshift_i16 = block(esch::kBound16);
        for (int v = 0; v < kDp16Variants; ++v) {
            const bool f16 = esch::dp16_is_f16(v);
            const int64_t lim = f16 ? esch::kBoundF16 : esch::kBound16;
            const int cs = f16 ? s->geo.cshift_f16 : s->geo.cshift_i16;
            bool ok = !f16 || s->f16_ok;
            if (esch::dp16_is_diag(v))
                ok = ok && !local && cs >= 0;
            else if (esch::dp16_is_shifted(v))
                ok = ok && (local ? cs >= 0
                                  : esch::shifted_bound((int64_t)max_len, s->sc.ma, s->sc.mi, s->sc.o, s->sc.e,
                                                        esch::dp16_stripe(v)) <= lim);
            s->usable16[v] = ok;
        }
    }
    // semiglobal mode: the diagonal shift on 32-row stripes (fp16, else
    // int16) where its bound holds; then the shifted variants where theirs
    // does
    const int vd16 = esch::dp16_variant(12, 32), vd = esch::dp16_variant(14, 32);
    const int v0 = s->usable16[vd16] ? vd16
                   : s->usable16[vd] ? vd
                   : s->usable16[7]  ? 7
                   : s->f16_ok       ? kDp16F16
                   : s->usable16[6]  ? 6
                                     : kDp16Default;
    s->geo.variant16 = v0;
    s->geo.grid16 = s->grid16_v[v0];
    s->regs16 = s->regs16_v[v0];
    s->local16 = s->local16_v[v0];
    s->geo.block32 = block32;
    s->geo.grid32 = std::max(1, b32) * s->n_cu;
    s->geo.grid_aux = 4 * s->n_cu;
    int bd = 0;
    // the wave-wide anti-diagonal kernel keeps its pass boundary in LDS;
    // panels with longer references use the 16-lane one with global scratch
    s->geo.diag_lds = max_len <= (size_t)kDiag64MaxLen ? (max_len + kDiag64Pad) * sizeof(uint2) : 0;
    HIPC(diag_occupancy(local, s->geo.diag_lds, &bd));
    s->geo.grid_diag = std::max(1, bd) * s->n_cu;
    *out = s.release();
    return ESCH_OK;
}

void esch_batch_free(esch_batch *b);

void esch_scorer_free(esch_scorer *s) {
    if (!s) return;
    for (esch_batch *b : s->pool) esch_batch_free(b);
    trace_cache_free(s);
    map_cache_free(s);
    (void)hipFree(s->d_ref2);
    (void)hipFree(s->d_refN);
    (void)hipFree(s->d_w2off);
    (void)hipFree(s->d_nmoff);
    (void)hipFree(s->d_rlen);
    (void)hipFree(s->d_rorder);
    (void)hipFree(s->d_pq_codes);
    (void)hipFree(s->d_pq_off);
    (void)hipFree(s->d_pq_len);
    delete s;
}

int esch_scores_len(const esch_scorer *s, size_t len) {
    if (!s) return 0;
    return (len > 0 && len <= ESCH_MAX_READ_LEN &&
            fits_i16(len, s->max_len, s->sc.ma, s->sc.mi, s->sc.o, s->sc.e))
               ? 1
               : 0;
}

size_t esch_n_refs(const esch_scorer *s) { return s ? s->n_refs : 0; }

esch_status esch_device_name(const esch_scorer *s, char *buf, size_t cap) {
    if (!s || !buf || cap == 0) return fail(ESCH_INVALID_ARG, "null argument");
    snprintf(buf, cap, "%s", s->device_name.c_str());
    return ESCH_OK;
}

esch_status esch_valu_peak16(int32_t ordinal, double *lane_ops_per_s, int *n_cu, int *clock_khz) {
    if (!lane_ops_per_s) return fail(ESCH_INVALID_ARG, "null argument");
    if (!esch_device_visible()) return fail(ESCH_UNAVAILABLE, "no HIP device visible");
    HIPC(hipSetDevice(ordinal));
    hipDeviceProp_t prop;
    HIPC(hipGetDeviceProperties(&prop, ordinal));
    if (n_cu) *n_cu = prop.multiProcessorCount;
    if (clock_khz) *clock_khz = prop.clockRate;
    // 8 workgroups of 4 waves per CU: 8 waves per SIMD, every issue slot
    // covered; a ~15 ms warm-up, then ~110 ms timed, so the clock has ramped
    const int blocks = 8 * prop.multiProcessorCount, iters = 1 << 19;
    uint32_t *out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    esch_status rc = ESCH_OK;
    float ms = 0.0f;
    if (hipMalloc(&out, (size_t)blocks * 256 * 4) != hipSuccess || hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = fail(ESCH_DRIVER, "allocation for the peak probe failed");
    } else if (launch_peak16(out, blocks, 1 << 16, st) != hipSuccess || hipEventRecord(e0, st) != hipSuccess ||
               launch_peak16(out, blocks, iters, st) != hipSuccess || hipEventRecord(e1, st) != hipSuccess ||
               hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = fail(ESCH_DRIVER, "peak probe launch failed: %s", hipGetErrorString(hipGetLastError()));
    } else {
        // 16 packed ops x 64 lanes x 2 halves per iteration per wave
        const double waves = (double)blocks * 4.0;
        *lane_ops_per_s = waves * iters * 16.0 * 128.0 / (ms * 1e-3);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (out) (void)hipFree(out);
    return rc;
}

void esch_geometry(const esch_scorer *s, size_t *lanes, size_t *groups, size_t *blocks) {
    if (!s) return;
    if (lanes) *lanes = 256;
    if (groups) *groups = 4;
    if (blocks) {
        std::lock_guard<std::mutex> g(s->geo_mu);
        *blocks = (size_t)s->geo.grid16;
    }
}

int esch_registers(const esch_scorer *s) {
    if (!s) return 0;
    std::lock_guard<std::mutex> g(s->geo_mu);
    return s->regs16;
}
int esch_local_bytes(const esch_scorer *s) {
    if (!s) return 0;
    std::lock_guard<std::mutex> g(s->geo_mu);
    return s->local16;
}

// ------------------------------------------------------------------ batches

// k_diag16's per-row-task pass boundary in global memory (k_diag64 keeps it
// in LDS and needs none)
static size_t diag_scratch_bytes(const esch_scorer *s) {
    if (s->geo.diag_lds) return sizeof(uint2);
    return (size_t)s->geo.grid_diag * 4 * s->max_len * sizeof(uint2);
}

// The swapped path's buffers: the batch's rows packed as a panel (2-bit
// words, wildcard bitmap, offsets, lengths, longest-first order) and the
// scorer's references bucketed as queries, with a descriptor of its own.
// One device block and one pinned staging block share the layout of the
// rows' panel; both are sized for the batch's capacity and allocated on the
// first swapped run.
struct SwapBufs {
    size_t cap_rows = 0, cap_w2 = 0, cap_wn = 0;
    size_t off_nm = 0, off_rlen = 0, off_rord = 0, off_w2 = 0, off_wn = 0, off_desc = 0, bytes = 0;
    uint8_t *d_block = nullptr, *h_stage = nullptr;
    uint8_t *d_cls = nullptr;
    int32_t *d_hist = nullptr, *d_bin_start = nullptr, *d_bin_fill = nullptr;
    int32_t *d_sorted16 = nullptr, *d_list32 = nullptr, *d_listD = nullptr;
};

static void swap_release(SwapBufs *w) {
    (void)hipFree(w->d_block);
    (void)hipHostFree(w->h_stage);
    (void)hipFree(w->d_cls);
    (void)hipFree(w->d_hist);
    (void)hipFree(w->d_bin_start);
    (void)hipFree(w->d_bin_fill);
    (void)hipFree(w->d_sorted16);
    (void)hipFree(w->d_list32);
    (void)hipFree(w->d_listD);
}

static void swap_layout(size_t max_reads, size_t max_bases, SwapBufs &w) {
    w.cap_rows = 2 * std::max<size_t>(max_reads, 1);
    const size_t bases = 2 * std::max<size_t>(max_bases, 1);
    // a row starts a new word and carries a padding word
    w.cap_w2 = bases / 16 + 2 * w.cap_rows + 1;
    w.cap_wn = bases / 32 + 2 * w.cap_rows + 1;
    w.off_nm = w.cap_rows * 4;
    w.off_rlen = w.off_nm + w.cap_rows * 4;
    w.off_rord = w.off_rlen + w.cap_rows * 4;
    w.off_w2 = w.off_rord + w.cap_rows * 4;
    w.off_wn = w.off_w2 + w.cap_w2 * 4;
    w.off_desc = (w.off_wn + w.cap_wn * 4 + 15) / 16 * 16;
    w.bytes = w.off_desc + sizeof(DevDesc);
}

static size_t swap_bytes(const esch_scorer *s, size_t max_reads, size_t max_bases) {
    SwapBufs w;
    swap_layout(max_reads, max_bases, w);
    return w.bytes + s->n_refs * (1 + 3 * 4) + 3 * 4 * (kMaxLen16 + 1);
}

size_t esch_batch_footprint(const esch_scorer *s, size_t max_reads, size_t max_bases) {
    if (!s) return 0;
    const size_t rows = 2 * std::max<size_t>(max_reads, 1);
    const size_t bases = 2 * std::max<size_t>(max_bases, 1);
    size_t f = bases + rows * (8 + 4 + 1 + 4 + 4) + 3 * 4 * (kMaxLen16 + 1) + sizeof(DevDesc);
    f += rows * s->n_refs * (4 + 2);                                     // out32 + out16
    f += (size_t)s->max_grid16 * 4 * (s->max_len + kScrPad) * 64 * sizeof(uint2);  // scratch16
    f += (size_t)s->geo.grid32 * s->geo.block32 * s->max_len * sizeof(int2);  // scratch32
    f += rows * 5 * 4;                                                   // records
    f += rows * 4 + diag_scratch_bytes(s);                               // anti-diagonal
    f += std::max<size_t>(max_reads, 1) * kTieCapAlloc * 4;              // ties
    f += swap_bytes(s, max_reads, max_bases);                            // swapped path
    return f;
}

esch_status esch_batch_create(esch_scorer *s, size_t max_reads, size_t max_bases, esch_batch **out) {
    if (!s || !out) return fail(ESCH_INVALID_ARG, "null argument");
    *out = nullptr;
    HIPC(hipSetDevice(s->ordinal));
    std::unique_ptr<esch_batch> b(new esch_batch);
    b->s = s;
    b->max_reads = std::max<size_t>(max_reads, 1);
    b->max_bases = std::max<size_t>(max_bases, 1);
    const size_t rows = 2 * b->max_reads, bases = 2 * b->max_bases;
    HIPC(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    HIPC(hipEventCreate(&b->ev0));
    HIPC(hipEventCreate(&b->ev1));
    HIPC(hipEventCreate(&b->ev2));
    HIPC(hipEventCreate(&b->ev3));
    HIPC(hipMalloc(&b->d_codes, bases));
    HIPC(hipMalloc(&b->d_qoff, rows * 8));
    HIPC(hipMalloc(&b->d_qlen, rows * 4));
    HIPC(hipMalloc(&b->d_cls, rows));
    HIPC(hipMalloc(&b->d_hist, 4 * (kMaxLen16 + 1)));
    HIPC(hipMalloc(&b->d_bin_start, 4 * (kMaxLen16 + 1)));
    HIPC(hipMalloc(&b->d_bin_fill, 4 * (kMaxLen16 + 1)));
    HIPC(hipMalloc(&b->d_sorted16, rows * 4));
    HIPC(hipMalloc(&b->d_list32, rows * 4));
    HIPC(hipMalloc(&b->d_out32, rows * s->n_refs * 4));
    HIPC(hipMalloc(&b->d_out16, rows * s->n_refs * 2));
    HIPC(hipMalloc(&b->d_scr16, (size_t)s->max_grid16 * 4 * (s->max_len + kScrPad) * 64 * sizeof(uint2)));
    HIPC(hipMalloc(&b->d_scr32, (size_t)s->geo.grid32 * s->geo.block32 * s->max_len * sizeof(int2)));
    HIPC(hipMalloc(&b->d_recs, rows * 5 * 4));
    HIPC(hipMalloc(&b->d_listD, rows * 4));
    HIPC(hipMalloc(&b->d_scrD, diag_scratch_bytes(s)));
    HIPC(hipMalloc(&b->d_desc, sizeof(DevDesc)));
    HIPC(hipMalloc(&b->d_ties, b->max_reads * kTieCapAlloc * 4));
    b->tie_cap_alloc = kTieCapAlloc;
    HIPC(hipHostMalloc(&b->h_codes, bases, hipHostMallocDefault));
    HIPC(hipHostMalloc(&b->h_qoff, rows * 8, hipHostMallocDefault));
    HIPC(hipHostMalloc(&b->h_qlen, rows * 4, hipHostMallocDefault));
    HIPC(hipHostMalloc(&b->h_desc, sizeof(DevDesc), hipHostMallocDefault));
    memset(b->h_desc, 0, sizeof(DevDesc));
    *out = b.release();
    return ESCH_OK;
}

void esch_batch_free(esch_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->s->ordinal);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    (void)hipFree(b->d_codes);
    (void)hipFree(b->d_qoff);
    (void)hipFree(b->d_qlen);
    (void)hipFree(b->d_cls);
    (void)hipFree(b->d_hist);
    (void)hipFree(b->d_bin_start);
    (void)hipFree(b->d_bin_fill);
    (void)hipFree(b->d_sorted16);
    (void)hipFree(b->d_list32);
    (void)hipFree(b->d_out32);
    (void)hipFree(b->d_out16);
    (void)hipFree(b->d_scr16);
    (void)hipFree(b->d_scr32);
    (void)hipFree(b->d_recs);
    (void)hipFree(b->d_ties);
    (void)hipFree(b->d_listD);
    (void)hipFree(b->d_scrD);
    (void)hipFree(b->d_desc);
    (void)hipHostFree(b->h_codes);
    (void)hipHostFree(b->h_qoff);
    (void)hipHostFree(b->h_qlen);
    (void)hipHostFree(b->h_desc);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->ev2) (void)hipEventDestroy(b->ev2);
    if (b->ev3) (void)hipEventDestroy(b->ev3);
    for (hipEvent_t e : b->sev)
        if (e) (void)hipEventDestroy(e);
    if (b->pump_ev) (void)hipEventDestroy(b->pump_ev);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->swap) {
        swap_release(b->swap);
        delete b->swap;
    }
    delete b;
}

namespace {

// Stage reads (and, for both strands, their reverse-complement rows, whose
// codes the device writes) into the pinned buffers, then upload.
esch_status batch_stage(esch_batch *b, size_t n, int both, int pairs,
                        const std::function<void(size_t, const uint8_t *&, size_t &)> &get) {
    if (pairs && (n % 2)) return fail(ESCH_INVALID_ARG, "pairs need an even number of mates");
    size_t total = 0;
    for (size_t k = 0; k < n; ++k) {
        const uint8_t *p;
        size_t len;
        get(k, p, len);
        if (len > (size_t)INT32_MAX) return fail(ESCH_INVALID_ARG, "read %zu is too long", k);
        total += len;
    }
    if (n > b->max_reads || total > b->max_bases)
        return fail(ESCH_CAPACITY, "batch holds %zu reads / %zu bases; got %zu / %zu", b->max_reads,
                    b->max_bases, n, total);
    b->n_slices = b->slices_out = 0;        // a sliced run's slices still out: the new reads replace it
    HIPC(hipStreamSynchronize(b->stream));  // staging buffers may still feed a copy
    // offsets first, then the gather into pinned staging on up to 8 host
    // threads (1M short reads are a million small copies), then one
    // vectorisable pass for the code check; the offending read is looked
    // up only when the check fails
    size_t at = 0, longest = 0;
    for (size_t k = 0; k < n; ++k) {
        const uint8_t *p;
        size_t len;
        get(k, p, len);
        b->h_qoff[k] = at;
        b->h_qlen[k] = (int32_t)len;
        at += len;
        longest = std::max(longest, len);
    }
    b->longest = longest;
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = total < ((size_t)8 << 20) ? 1 : std::min<size_t>(8, hw);
    std::vector<uint8_t> bad(nt, 0);
    auto gather = [&](size_t t) {
        const size_t k0 = n * t / nt, k1 = n * (t + 1) / nt;
        uint8_t mx = 0;
        for (size_t k = k0; k < k1; ++k) {
            const uint8_t *p;
            size_t len;
            get(k, p, len);
            uint8_t *dst = b->h_codes + b->h_qoff[k];
            if (len) memcpy(dst, p, len);
        }
        if (k1 > k0) {
            const uint8_t *q = b->h_codes + b->h_qoff[k0];
            const size_t end = (k1 < n ? b->h_qoff[k1] : total) - b->h_qoff[k0];
            for (size_t j = 0; j < end; ++j) mx = q[j] > mx ? q[j] : mx;
        }
        bad[t] = mx > 4;
    };
    if (nt == 1) {
        gather(0);
    } else {
        std::vector<std::thread> pool;
        for (size_t t = 1; t < nt; ++t) pool.emplace_back(gather, t);
        gather(0);
        for (auto &th : pool) th.join();
    }
    for (size_t t = 0; t < nt; ++t) {
        if (!bad[t]) continue;
        for (size_t k = 0; k < n; ++k) {
            const uint8_t *q = b->h_codes + b->h_qoff[k];
            for (int32_t j = 0; j < b->h_qlen[k]; ++j)
                if (q[j] > 4) return fail(ESCH_INVALID_ARG, "read %zu carries code %d (codes are 0..=4)", k, q[j]);
        }
    }
    const size_t rows = both ? 2 * n : n;
    if (both) {
        for (size_t k = 0; k < n; ++k) {
            b->h_qoff[n + k] = total + b->h_qoff[k];
            b->h_qlen[n + k] = b->h_qlen[k];
        }
    }
    b->n_reads = n;
    b->n_rows = rows;
    b->both = both;
    b->pairs = pairs;
    if (total) HIPC(hipMemcpyAsync(b->d_codes, b->h_codes, total, hipMemcpyHostToDevice, b->stream));
    if (rows) {
        HIPC(hipMemcpyAsync(b->d_qoff, b->h_qoff, rows * 8, hipMemcpyHostToDevice, b->stream));
        HIPC(hipMemcpyAsync(b->d_qlen, b->h_qlen, rows * 4, hipMemcpyHostToDevice, b->stream));
    }
    return ESCH_OK;
}

KArgs make_args(esch_batch *b, int what) {
    esch_scorer *s = b->s;
    KArgs a{};
    a.panel = s->panel;
    a.sc = s->sc;
    a.desc = b->d_desc;
    a.qcodes = b->d_codes;
    a.qoff = b->d_qoff;
    a.qlen = b->d_qlen;
    a.cls = b->d_cls;
    a.hist = b->d_hist;
    a.bin_start = b->d_bin_start;
    a.bin_fill = b->d_bin_fill;
    a.sorted16 = b->d_sorted16;
    a.list32 = b->d_list32;
    a.out32 = (what & ESCH_RUN_SCORES16) ? nullptr : b->d_out32;
    a.out16 = (what & ESCH_RUN_SCORES16) ? b->d_out16 : nullptr;
    a.scratch16 = b->d_scr16;
    a.scratch32 = b->d_scr32;
    a.recs = b->d_recs;
    a.ties = b->d_ties;
    a.rc_codes = b->d_codes;
    a.listD = b->d_listD;
    a.scratchD = b->d_scrD;
    return a;
}

// What follows k_dp16 in a run: the anti-diagonal and int32 kernels, then
// the fused selection.
esch_status enqueue_tail(esch_batch *b, int what, bool timed, const Launch &g) {
    const KArgs a = make_args(b, what);
    const bool local = b->s->sc.local != 0;
    hipStream_t st = b->stream;
    HIPC(launch_diag(a, g, local, st));
    HIPC(launch_dp32(a, g, local, 1, st));
    if (timed) HIPC(hipEventRecord(b->ev2, st));
    if (what & ESCH_RUN_SELECT) HIPC(launch_select(a, g, st));
    if (what & ESCH_RUN_PAIRS) HIPC(launch_pairs(a, g, st));
    return ESCH_OK;
}

// Slice k of a sliced run, with its events; the last one brings the tail.
esch_status launch_slice(esch_batch *b, int k) {
    KArgs a = make_args(b, b->slice_what);
    a.slice = k;
    a.n_slices = b->n_slices;
    hipStream_t st = b->stream;
    HIPC(hipEventRecord(b->sev[2 * k], st));
    HIPC(launch_dp16(a, b->slice_geo, b->s->sc.local != 0, st));
    HIPC(hipEventRecord(b->sev[2 * k + 1], st));
    b->slices_out = k + 1;
    if (k + 1 == b->n_slices) {
        if (esch_status e = enqueue_tail(b, b->slice_what, true, b->slice_geo)) return e;
        HIPC(hipEventRecord(b->ev3, st));
    }
    return ESCH_OK;
}

// A run's launches: bucketing, then k_dp16 (a sliced run: its slice 0, the
// rest from batch_pump) and the tail, or the plain kernel.
esch_status enqueue(esch_batch *b, int what, bool timed, const Launch &g, int n_slices = 0) {
    esch_scorer *s = b->s;
    const KArgs a = make_args(b, what);
    const bool local = s->sc.local != 0;
    hipStream_t st = b->stream;
    HIPC(launch_prep(a, g, st));
    if (timed) HIPC(hipEventRecord(b->ev1, st));
    if (what & ESCH_RUN_PLAIN) {
        HIPC(launch_dp32(a, g, local, 0, st));
        HIPC(launch_dp32(a, g, local, 1, st));
        if (timed) HIPC(hipEventRecord(b->ev2, st));
        if (what & ESCH_RUN_SELECT) HIPC(launch_select(a, g, st));
        if (what & ESCH_RUN_PAIRS) HIPC(launch_pairs(a, g, st));
        return ESCH_OK;
    }
    if (n_slices > 1) {
        b->n_slices = n_slices;
        b->slice_geo = g;
        b->slice_what = what;
        return launch_slice(b, 0);
    }
    HIPC(launch_dp16(a, g, local, st));
    return enqueue_tail(b, what, timed, g);
}

// The k_dp16 variant a batch runs, and the column block of its fp16
// shifted-column form (*cs_f16: the panel's unless the batch upgrades):
// the scorer's variant, unless the caller fixed none (the default or the
// autotuner's pick) and the panel admits no fp16 variant while this batch's
// reads do. Every value of the fp16 recurrences is bounded through
// k = min(read, reference), the shifted columns' span aside, and a column
// shift restarted often enough spans a block whatever the references. So
// the batch's longest read caps k as the panel's longest reference does,
// and short reads on a panel of long references still get fp16:
//  - local mode: the column-shifted fp16 schedule (7 on 16-row stripes,
//    the windowed 10 on taller ones; 6.5 ops per cell) with the widest
//    block the batch admits, else the unshifted 4 (8.5 ops);
//  - semiglobal mode, an int16 diagonal-shift pick: the diagonal shift in
//    fp16 (schedule 12, 5 ops) with its block; where no block fits, the
//    pick stays (the row-shifted fp16 form is no fewer ops);
//  - semiglobal mode, another int16 pick: the row-shifted fp16 schedule
//    (7, or 10 for a windowed pick; 6 ops), else 4 (7 ops).
int batch_variant(const esch_batch *b, int v, bool fixed, int *cs_f16) {
    const esch_scorer *s = b->s;
    if (fixed || esch::dp16_is_f16(v)) return v;
    const Scoring &sc = s->sc;
    const int64_t L = (int64_t)s->max_len;
    const int64_t k = (int64_t)std::min<size_t>(s->max_len, b->longest);
    if (!fits_f16(k, sc.ma, sc.mi, sc.o, sc.e)) return v;
    const int st = esch::dp16_stripe(v);
    const bool local = sc.local != 0;
    if (local || esch::dp16_is_diag(v)) {
        const int kb = esch::column_block_log2(local, k, L, sc.ma, sc.mi, sc.o, sc.e, 32, esch::kBoundF16);
        if (kb >= 0) {
            *cs_f16 = kb;
            return local ? esch::dp16_variant(st == 16 ? 7 : 10, st) : esch::dp16_variant(12, st);
        }
        return local ? kDp16F16 : v;
    }
    const int up = esch::dp16_variant(esch::dp16_schedule(v) == 11 ? 10 : 7, st);
    return esch::shifted_bound(k, sc.ma, sc.mi, sc.o, sc.e, st) <= esch::kBoundF16 ? up : kDp16F16;
}

// Rows up to which a batch is scored with the roles swapped. With the
// roles swapped the references are the queries, bucketed into well-filled
// 128-row chunks, and every row is a reference of its exact length, so no
// row pads another; the work items are (chunk of references, 4 rows). On
// one MI355X that wins at every batch size measured (6 to 6,400 rows)
// against 10,000 references, and loses against 1,000 or 164, whose few
// reference chunks leave the device underfilled: the default swaps for
// panels of at least 32 chunks and batches of at most one row per
// reference.
size_t swap_limit(const esch_scorer *s) {
    if (s->swap_rows >= 0) return (size_t)s->swap_rows;
    return s->n_refs >= 32 * 128 ? s->n_refs : 0;
}

bool swap_eligible(const esch_batch *b, int what) {
    const esch_scorer *s = b->s;
    if ((what & (ESCH_RUN_SCORES16 | ESCH_RUN_PLAIN | ESCH_RUN_PAIRS)) || !(what & ESCH_RUN_SCORES32)) return false;
    if (s->diag_min_len > 0) return false;  // the caller routes by length explicitly
    if (b->n_rows == 0 || b->n_rows > swap_limit(s)) return false;
    // every row a non-empty reference no longer than the panel's (the
    // scratch is sized by the panel's longest reference)
    for (size_t r = 0; r < b->n_rows; ++r)
        if (b->h_qlen[r] <= 0 || (size_t)b->h_qlen[r] > s->max_len) return false;
    return true;
}

esch_status swap_ensure(esch_batch *b) {
    if (b->swap) return ESCH_OK;
    std::unique_ptr<SwapBufs> w(new SwapBufs);
    swap_layout(b->max_reads, b->max_bases, *w);
    const size_t nq = b->s->n_refs;
    esch_status st = ESCH_OK;
    auto dev = [&](auto **p, size_t n) {
        if (!st && hipMalloc(p, std::max<size_t>(n, 1)) != hipSuccess) st = fail(ESCH_DRIVER, "hipMalloc failed");
    };
    dev(&w->d_block, w->bytes);
    dev(&w->d_cls, nq);
    dev(&w->d_hist, 4 * (kMaxLen16 + 1));
    dev(&w->d_bin_start, 4 * (kMaxLen16 + 1));
    dev(&w->d_bin_fill, 4 * (kMaxLen16 + 1));
    dev(&w->d_sorted16, 4 * nq);
    dev(&w->d_list32, 4 * nq);
    dev(&w->d_listD, 4 * nq);
    if (!st && hipHostMalloc(&w->h_stage, w->bytes, hipHostMallocDefault) != hipSuccess)
        st = fail(ESCH_DRIVER, "hipHostMalloc failed");
    if (st) {  // nothing half-built stays behind: the next run tries again
        swap_release(w.get());
        return st;
    }
    b->swap = w.release();
    return ESCH_OK;
}

// Score the batch's rows against the panel with the roles swapped: the
// rows packed as a panel (reverse complements made here), the scorer's
// references as the queries, k_dp16 / k_dp32 writing every score where an
// ordinary run would (KArgs::swap), then the ordinary selection. The score
// of a read against a reference equals the reference's against the read:
// the substitution, the gaps, the zero borders and the best cell (anywhere
// in local mode; on the last row or column in semiglobal mode) are the same
// with the two sequences exchanged.
esch_status run_swapped(esch_batch *b, int what, const Launch &geo) {
    esch_scorer *s = b->s;
    if (esch_status st = swap_ensure(b)) return st;
    SwapBufs &w = *b->swap;
    const size_t rows = b->n_rows, n = b->n_reads;
    uint8_t *h = w.h_stage;
    uint32_t *w2off = reinterpret_cast<uint32_t *>(h);
    uint32_t *nmoff = reinterpret_cast<uint32_t *>(h + w.off_nm);
    int32_t *rlen = reinterpret_cast<int32_t *>(h + w.off_rlen);
    int32_t *rord = reinterpret_cast<int32_t *>(h + w.off_rord);
    uint32_t *ref2 = reinterpret_cast<uint32_t *>(h + w.off_w2);
    uint32_t *refN = reinterpret_cast<uint32_t *>(h + w.off_wn);
    DevDesc *desc = reinterpret_cast<DevDesc *>(h + w.off_desc);
    size_t n2 = 0, nn = 0;
    int32_t max_len = 0;
    int64_t total = 0;
    for (size_t r = 0; r < rows; ++r) {
        const int32_t L = b->h_qlen[r];
        const uint8_t *src = b->h_codes + b->h_qoff[r < n ? r : r - n];
        const bool rc = r >= n;
        w2off[r] = (uint32_t)n2;
        nmoff[r] = (uint32_t)nn;
        rlen[r] = L;
        const size_t k2 = (size_t)(L + 15) / 16 + 1, kn = (size_t)(L + 31) / 32 + 1;
        memset(ref2 + n2, 0, k2 * 4);
        memset(refN + nn, 0, kn * 4);
        for (int32_t j = 0; j < L; ++j) {
            uint8_t c = rc ? src[L - 1 - j] : src[j];
            if (rc && c < 4) c = (uint8_t)(3 - c);
            if (c == 4) refN[nn + (size_t)j / 32] |= 1u << (j % 32);
            else ref2[n2 + (size_t)j / 16] |= (uint32_t)c << (2 * (j % 16));
        }
        n2 += k2;
        nn += kn;
        max_len = std::max(max_len, L);
        total += L;
    }
    for (size_t r = 0; r < rows; ++r) rord[r] = (int32_t)r;
    std::stable_sort(rord, rord + rows, [&](int32_t x, int32_t y) { return rlen[x] > rlen[y]; });
    memset(desc, 0, sizeof(DevDesc));
    desc->n_queries = (int32_t)s->n_refs;
    desc->n_reads = (int32_t)s->n_refs;
    desc->tie_cap = 1;
    desc->max_ties = -1;
    desc->tstrands = 1;
    desc->blocks16 = (int32_t)geo.grid16;  // routing off: no anti-diagonal rows
    hipStream_t st = b->stream;
    uint8_t *d = w.d_block;
    HIPC(hipMemcpyAsync(d, h, rows * 4, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(d + w.off_nm, h + w.off_nm, rows * 4, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(d + w.off_rlen, h + w.off_rlen, rows * 4, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(d + w.off_rord, h + w.off_rord, rows * 4, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(d + w.off_w2, h + w.off_w2, n2 * 4, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(d + w.off_wn, h + w.off_wn, nn * 4, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(d + w.off_desc, h + w.off_desc, sizeof(DevDesc), hipMemcpyHostToDevice, st));

    const KArgs a0 = make_args(b, what);
    // the batch's own reverse-complement rows, as the ordinary path writes
    // them: whatever reads the batch's codes afterwards (a mapping call's
    // tracebacks) finds every row there
    if (b->both) HIPC(launch_revcomp(a0, geo, st));
    KArgs a = a0;
    a.panel = DevPanel{reinterpret_cast<uint32_t *>(d + w.off_w2), reinterpret_cast<uint32_t *>(d + w.off_wn),
                       reinterpret_cast<uint32_t *>(d), reinterpret_cast<uint32_t *>(d + w.off_nm),
                       reinterpret_cast<int32_t *>(d + w.off_rlen), reinterpret_cast<int32_t *>(d + w.off_rord),
                       (int32_t)rows, max_len, total};
    a.desc = reinterpret_cast<DevDesc *>(d + w.off_desc);
    a.qcodes = s->d_pq_codes;
    a.qoff = s->d_pq_off;
    a.qlen = s->d_pq_len;
    a.cls = w.d_cls;
    a.hist = w.d_hist;
    a.bin_start = w.d_bin_start;
    a.bin_fill = w.d_bin_fill;
    a.sorted16 = w.d_sorted16;
    a.list32 = w.d_list32;
    a.listD = w.d_listD;
    a.out16 = nullptr;
    a.swap = 1;
    const bool local = s->sc.local != 0;
    HIPC(launch_prep(a, geo, st));
    HIPC(hipEventRecord(b->ev1, st));
    HIPC(launch_dp16(a, geo, local, st));
    HIPC(launch_dp32(a, geo, local, 1, st));
    HIPC(hipEventRecord(b->ev2, st));
    HIPC(hipMemsetAsync(b->d_cls, 1, rows, st));  // every row scored (a run outside any graph)
    if (what & ESCH_RUN_SELECT) HIPC(launch_select(a0, geo, st));
    return ESCH_OK;
}

esch_status run_impl(esch_batch *b, int what, int32_t min_score, int64_t max_ties, size_t tie_cap, int n_slices) {
    esch_scorer *s = b->s;
    HIPC(hipSetDevice(s->ordinal));
    b->n_slices = b->slices_out = 0;  // an earlier sliced run's slices still out are dropped: this run replaces it
    if ((what & ESCH_RUN_PAIRS) && !b->pairs) return fail(ESCH_INVALID_ARG, "batch was not loaded as pairs");
    if ((what & (ESCH_RUN_SELECT | ESCH_RUN_PAIRS)) && (what & ESCH_RUN_SCORES16))
        return fail(ESCH_INVALID_ARG, "selection reads the int32 matrix");
    if (what & (ESCH_RUN_SELECT | ESCH_RUN_PAIRS)) {
        if (tie_cap == 0) tie_cap = 1;
        if (tie_cap > b->tie_cap_alloc) {
            HIPC(hipStreamSynchronize(b->stream));
            if (b->d_ties) (void)hipFree(b->d_ties);
            b->d_ties = nullptr;
            HIPC(hipMalloc(&b->d_ties, std::max<size_t>(b->max_reads, 1) * tie_cap * 4));
            b->tie_cap_alloc = tie_cap;
            for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
            b->graphs.clear();
        }
    }
    DevDesc &d = *b->h_desc;
    HIPC(hipStreamSynchronize(b->stream));  // the pinned descriptor may still feed a copy
    d.n_queries = (int32_t)b->n_rows;
    d.n_reads = (int32_t)b->n_reads;
    d.both = b->both;
    d.pairs = b->pairs;
    d.dropin = (what & ESCH_RUN_SCORES16) ? 1 : 0;
    d.force32 = (what & ESCH_RUN_PLAIN) ? 1 : 0;
    d.min_score = min_score;
    d.tie_cap = (int32_t)std::max<size_t>(tie_cap, 1);
    d.max_ties = max_ties;
    d.tstrands = (what & ESCH_RUN_PAIR_BOTH) ? 2 : 1;
    d.diag_min_len = (what & (ESCH_RUN_PLAIN | ESCH_RUN_SCORES16)) ? 0 : s->diag_min_len;
    Launch geo;
    int ratio;
    bool fixed;
    {
        std::lock_guard<std::mutex> g(s->geo_mu);
        geo = s->geo;
        ratio = s->diag_ratio;
        fixed = s->variant_fixed;
        for (int c = 0; c < esch::kLenClasses; ++c) d.plan[c] = s->diag_plan[c];
    }
    d.diag_ratio = (what & ESCH_RUN_PLAIN) ? 0 : ratio;
    const Launch geo_scorer = geo;  // the swapped path runs the scorer's own variant
    int cs_f16 = geo.cshift_f16;
    const int v16 = batch_variant(b, geo.variant16, fixed, &cs_f16);
    geo.variant16 = v16;
    geo.cshift_f16 = cs_f16;
    geo.grid16 = s->grid16_v[v16];
    d.blocks16 = (int32_t)geo.grid16;
    HIPC(hipMemcpyAsync(b->d_desc, b->h_desc, sizeof(DevDesc), hipMemcpyHostToDevice, b->stream));
    b->last_what = what;
    const bool graph = !(what & ESCH_RUN_NO_GRAPH);
    // the template-strand count is read from the descriptor, not baked in;
    // the kernel variant and its fp16 column block are baked into the
    // captured launches (the run flags take bits 0-6, the variant 16-20)
    const int key = (what & ~(ESCH_RUN_NO_GRAPH | ESCH_RUN_PAIR_BOTH)) | (v16 << 16) | (((cs_f16 + 1) & 31) << 24);
    HIPC(hipEventRecord(b->ev0, b->stream));
    b->last_swapped = swap_eligible(b, what);
    b->last_variant = b->last_swapped ? -1 : v16;
    if (b->last_swapped) {
        b->timed_dp = true;
        if (esch_status st = run_swapped(b, what, geo_scorer)) return st;
    } else if (n_slices > 1 && !(what & ESCH_RUN_PLAIN)) {
        b->timed_dp = true;
        n_slices = std::min(n_slices, esch::kMaxSlices);
        for (int k = 0; k < 2 * n_slices; ++k)
            if (!b->sev[k]) HIPC(hipEventCreate(&b->sev[k]));
        if (!b->pump_ev) HIPC(hipEventCreateWithFlags(&b->pump_ev, hipEventDisableTiming));
        // ev3 comes with the last slice
        return enqueue(b, key, true, geo, n_slices);
    } else if (!graph) {
        b->timed_dp = true;
        if (esch_status st = enqueue(b, key, true, geo)) return st;
    } else {
        b->timed_dp = false;
        auto it = b->graphs.find(key);
        if (it == b->graphs.end()) {
            hipGraph_t g;
            HIPC(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            esch_status st = enqueue(b, key, false, geo);
            hipError_t ce = hipStreamEndCapture(b->stream, &g);
            if (st) return st;
            if (ce != hipSuccess) return fail(ESCH_DRIVER, "hipStreamEndCapture: %s", hipGetErrorString(ce));
            hipGraphExec_t ge;
            hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) return fail(ESCH_DRIVER, "hipGraphInstantiate: %s", hipGetErrorString(ie));
            it = b->graphs.emplace(key, ge).first;
        }
        HIPC(hipGraphLaunch(it->second, b->stream));
    }
    HIPC(hipEventRecord(b->ev3, b->stream));
    return ESCH_OK;
}

}  // namespace

esch_status esch_batch_load(esch_batch *b, const uint8_t *const *reads, const size_t *lens, size_t n,
                            int both, int pairs) {
    if (!b || (n && (!reads || !lens))) return fail(ESCH_INVALID_ARG, "null argument");
    HIPC(hipSetDevice(b->s->ordinal));
    return batch_stage(b, n, both ? 1 : 0, pairs ? 1 : 0, [&](size_t k, const uint8_t *&p, size_t &len) {
        p = reads[k];
        len = lens[k];
    });
}

esch_status esch_batch_load_concat(esch_batch *b, const uint8_t *codes, const uint64_t *offs,
                                   const int32_t *lens, size_t n, int both, int pairs) {
    if (!b || (n && (!codes || !offs || !lens))) return fail(ESCH_INVALID_ARG, "null argument");
    HIPC(hipSetDevice(b->s->ordinal));
    return batch_stage(b, n, both ? 1 : 0, pairs ? 1 : 0, [&](size_t k, const uint8_t *&p, size_t &len) {
        p = codes + offs[k];
        len = (size_t)(lens[k] < 0 ? 0 : lens[k]);
    });
}

esch_status esch_batch_run(esch_batch *b, int what, int32_t min_score, int64_t max_ties, size_t tie_cap) {
    if (!b) return fail(ESCH_INVALID_ARG, "null batch");
    return run_impl(b, what, min_score, max_ties, tie_cap, 0);
}

esch_status esch_batch_wait(esch_batch *b) {
    if (!b) return fail(ESCH_INVALID_ARG, "null batch");
    HIPC(hipSetDevice(b->s->ordinal));
    if (esch_status e = esch::batch_drain(b)) return e;
    HIPC(hipStreamSynchronize(b->stream));
    return ESCH_OK;
}

int esch_batch_ready(esch_batch *b) {
    if (!b) return 0;
    if (esch::batch_drain(b)) return 0;
    return hipStreamQuery(b->stream) == hipSuccess ? 1 : 0;
}

esch_status esch_batch_timing(const esch_batch *b, float *dp_ms, float *total_ms) {
    if (!b) return fail(ESCH_INVALID_ARG, "null batch");
    if (b->slices_out < b->n_slices) return fail(ESCH_INVALID_ARG, "the run has slices still to launch");
    if (b->n_slices > 1) {
        // the slices alone (other streams' work ran between them), then the
        // tail; the bucketing before and the selection after for the total
        float dp = 0.0f, x = 0.0f;
        for (int k = 0; k < b->n_slices; ++k) {
            HIPC(hipEventElapsedTime(&x, b->sev[2 * k], b->sev[2 * k + 1]));
            dp += x;
        }
        HIPC(hipEventElapsedTime(&x, b->sev[2 * b->n_slices - 1], b->ev2));
        dp += x;
        if (dp_ms) *dp_ms = dp;
        if (total_ms) {
            float pre = 0.0f, post = 0.0f;
            HIPC(hipEventElapsedTime(&pre, b->ev0, b->ev1));
            HIPC(hipEventElapsedTime(&post, b->ev2, b->ev3));
            *total_ms = pre + dp + post;
        }
        return ESCH_OK;
    }
    if (dp_ms) {
        *dp_ms = -1.0f;
        if (b->timed_dp) HIPC(hipEventElapsedTime(dp_ms, b->ev1, b->ev2));
    }
    if (total_ms) HIPC(hipEventElapsedTime(total_ms, b->ev0, b->ev3));
    return ESCH_OK;
}

// The descriptor of the last run's scoring (the swapped path keeps its own).
static const DevDesc *scoring_desc(const esch_batch *b) {
    return b->last_swapped ? reinterpret_cast<const DevDesc *>(b->swap->d_block + b->swap->off_desc) : b->d_desc;
}

esch_status esch_batch_stats(const esch_batch *b, uint64_t *cells, uint64_t *n16, uint64_t *n32) {
    if (!b) return fail(ESCH_INVALID_ARG, "null batch");
    DevDesc d;
    HIPC(hipMemcpy(&d, scoring_desc(b), sizeof d, hipMemcpyDeviceToHost));
    if (cells) *cells = d.cells;
    if (n16) *n16 = (uint64_t)d.n16 + (uint64_t)d.nd;
    if (n32) *n32 = (uint64_t)d.n32;
    return ESCH_OK;
}

esch_status esch_batch_stats_diag(const esch_batch *b, uint64_t *n_diag) {
    if (!b || !n_diag) return fail(ESCH_INVALID_ARG, "null argument");
    DevDesc d;
    HIPC(hipMemcpy(&d, scoring_desc(b), sizeof d, hipMemcpyDeviceToHost));
    *n_diag = (uint64_t)d.nd;
    return ESCH_OK;
}

esch_status esch_batch_read_scores32(esch_batch *b, int32_t *scores, uint8_t *scored) {
    if (!b) return fail(ESCH_INVALID_ARG, "null batch");
    if (b->last_what & ESCH_RUN_SCORES16) return fail(ESCH_INVALID_ARG, "last run wrote the int16 matrix");
    HIPC(hipSetDevice(b->s->ordinal));
    if (esch_status e = esch::batch_drain(b)) return e;
    HIPC(hipStreamSynchronize(b->stream));
    const size_t rows = b->n_rows, nr = b->s->n_refs;
    if (scores && rows) HIPC(hipMemcpy(scores, b->d_out32, rows * nr * 4, hipMemcpyDeviceToHost));
    if (scored && rows) {
        HIPC(hipMemcpy(scored, b->d_cls, rows, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < rows; ++k) scored[k] = scored[k] ? 1 : 0;
    }
    return ESCH_OK;
}

esch_status esch_batch_read_scores16(esch_batch *b, int16_t *scores, uint8_t *scored) {
    if (!b) return fail(ESCH_INVALID_ARG, "null batch");
    if (!(b->last_what & ESCH_RUN_SCORES16)) return fail(ESCH_INVALID_ARG, "last run wrote the int32 matrix");
    HIPC(hipSetDevice(b->s->ordinal));
    if (esch_status e = esch::batch_drain(b)) return e;
    HIPC(hipStreamSynchronize(b->stream));
    const size_t rows = b->n_rows, nr = b->s->n_refs;
    if (scores && rows) HIPC(hipMemcpy(scores, b->d_out16, rows * nr * 2, hipMemcpyDeviceToHost));
    if (scored && rows) {
        HIPC(hipMemcpy(scored, b->d_cls, rows, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < rows; ++k) scored[k] = scored[k] ? 1 : 0;
    }
    return ESCH_OK;
}

esch_status esch_batch_read_select(esch_batch *b, esch_read_select *out, int32_t *ties) {
    if (!b) return fail(ESCH_INVALID_ARG, "null batch");
    if (!(b->last_what & ESCH_RUN_SELECT)) return fail(ESCH_INVALID_ARG, "last run did not select");
    HIPC(hipSetDevice(b->s->ordinal));
    HIPC(hipStreamSynchronize(b->stream));
    // copies ordered on the batch's own stream: they wait for nothing that
    // other streams have in flight
    const size_t n = b->n_reads;
    if (out && n)
        HIPC(hipMemcpyAsync(out, b->d_recs, n * sizeof(esch_read_select), hipMemcpyDeviceToHost, b->stream));
    if (ties && n)
        HIPC(hipMemcpyAsync(ties, b->d_ties, n * b->h_desc->tie_cap * 4, hipMemcpyDeviceToHost, b->stream));
    HIPC(hipStreamSynchronize(b->stream));
    return ESCH_OK;
}

esch_status esch_batch_read_pairs(esch_batch *b, esch_pair_select *out, int32_t *top) {
    if (!b) return fail(ESCH_INVALID_ARG, "null batch");
    if (!(b->last_what & ESCH_RUN_PAIRS)) return fail(ESCH_INVALID_ARG, "last run did not select pairs");
    HIPC(hipSetDevice(b->s->ordinal));
    HIPC(hipStreamSynchronize(b->stream));
    const size_t n = b->n_reads / 2;
    if (out && n)
        HIPC(hipMemcpyAsync(out, b->d_recs, n * sizeof(esch_pair_select), hipMemcpyDeviceToHost, b->stream));
    if (top && n)
        HIPC(hipMemcpyAsync(top, b->d_ties, n * b->h_desc->tie_cap * 4, hipMemcpyDeviceToHost, b->stream));
    HIPC(hipStreamSynchronize(b->stream));
    return ESCH_OK;
}

// ------------------------------------------------------------------ drop-in

esch_prepared *esch_prepare(const esch_scorer *s, const uint8_t *const *queries, const size_t *lens, size_t n) {
    if (!s || (n && (!queries || !lens))) {
        fail(ESCH_INVALID_ARG, "null argument");
        return nullptr;
    }
    std::unique_ptr<esch_prepared> p(new esch_prepared);
    p->n = n;
    p->off.resize(n);
    p->len.resize(n);
    p->scored.resize(n);
    size_t total = 0;
    for (size_t k = 0; k < n; ++k) {
        p->scored[k] = (uint8_t)esch_scores_len(s, lens[k]);
        // unscored queries travel as empty rows (lib/cuda/mod.rs:555-563)
        const size_t L = p->scored[k] ? lens[k] : 0;
        p->off[k] = total;
        p->len[k] = (int32_t)L;
        total += L;
    }
    p->codes.resize(std::max<size_t>(total, 1));
    for (size_t k = 0; k < n; ++k) {
        if (!p->len[k]) continue;
        for (size_t j = 0; j < lens[k]; ++j) {
            if (queries[k][j] > 4) {
                fail(ESCH_INVALID_ARG, "query %zu carries code %d (codes are 0..=4)", k, queries[k][j]);
                return nullptr;
            }
        }
        memcpy(p->codes.data() + p->off[k], queries[k], lens[k]);
    }
    return p.release();
}

void esch_prepared_free(esch_prepared *p) { delete p; }
size_t esch_prepared_len(const esch_prepared *p) { return p ? p->n : 0; }
size_t esch_prepared_n_scored(const esch_prepared *p) {
    if (!p) return 0;
    size_t c = 0;
    for (uint8_t v : p->scored) c += v;
    return c;
}

}  // extern "C"

namespace esch {

esch_status take_batch(esch_scorer *s, size_t n, size_t bases, esch_batch **out) {
    {
        std::lock_guard<std::mutex> g(s->pool_mu);
        for (size_t i = 0; i < s->pool.size(); ++i) {
            esch_batch *b = s->pool[i];
            if (b->max_reads >= n && b->max_bases >= bases) {
                s->pool.erase(s->pool.begin() + (long)i);
                *out = b;
                return ESCH_OK;
            }
        }
    }
    // grow geometrically so a stream of similar batches reuses one buffer set
    size_t cap_n = 1024, cap_b = 1 << 16;
    while (cap_n < n) cap_n *= 2;
    while (cap_b < bases) cap_b *= 2;
    return esch_batch_create(s, cap_n, cap_b, out);
}

void give_batch(esch_scorer *s, esch_batch *b) {
    std::lock_guard<std::mutex> g(s->pool_mu);
    s->pool.push_back(b);
    if (s->pool.size() > 4) {  // keep a few idle buffer sets, free the oldest
        esch_batch *old = s->pool.front();
        s->pool.erase(s->pool.begin());
        esch_batch_free(old);
    }
}

esch_status batch_run_sliced(esch_batch *b, int what, int n_slices) {
    if (!b) return fail(ESCH_INVALID_ARG, "null batch");
    if (what & (ESCH_RUN_SELECT | ESCH_RUN_PAIRS | ESCH_RUN_PLAIN))
        return fail(ESCH_INVALID_ARG, "a sliced run scores only");
    return run_impl(b, what | ESCH_RUN_NO_GRAPH, 0, -1, 1, n_slices);
}

esch_status batch_pump(esch_batch *b, hipStream_t after) {
    if (!b || b->slices_out >= b->n_slices) return ESCH_OK;
    HIPC(hipSetDevice(b->s->ordinal));
    if (after) {
        HIPC(hipEventRecord(b->pump_ev, after));
        HIPC(hipStreamWaitEvent(b->stream, b->pump_ev, 0));
    }
    return launch_slice(b, b->slices_out);
}

esch_status batch_drain(esch_batch *b) {
    while (b && b->slices_out < b->n_slices)
        if (esch_status e = launch_slice(b, b->slices_out)) return e;
    return ESCH_OK;
}

}  // namespace esch

extern "C" {

esch_status esch_submit(esch_scorer *s, esch_prepared *p, esch_pending **out) {
    std::unique_ptr<esch_prepared> prep(p);
    if (!s || !p || !out) return fail(ESCH_INVALID_ARG, "null argument");
    *out = nullptr;
    HIPC(hipSetDevice(s->ordinal));
    std::unique_ptr<esch_pending> pend(new esch_pending);
    pend->n = p->n;
    pend->scored = p->scored;
    size_t total = 0;
    for (int32_t L : p->len) total += (size_t)L;
    esch_batch *b = nullptr;
    if (esch_status st = take_batch(s, p->n, total, &b)) return st;
    pend->b = b;
    esch_status st = esch_batch_load_concat(b, p->codes.data(), p->off.data(), p->len.data(), p->n, 0, 0);
    if (!st) st = esch_batch_run(b, ESCH_RUN_SCORES16, 0, -1, 1);
    if (st) {
        give_batch(s, b);
        return st;
    }
    *out = pend.release();
    return ESCH_OK;
}

esch_status esch_finish(esch_scorer *s, esch_pending *p, int16_t *scores, uint8_t *scored) {
    std::unique_ptr<esch_pending> pend(p);
    if (!s || !p) return fail(ESCH_INVALID_ARG, "null argument");
    if (p->n && (!scores || !scored)) {
        give_batch(s, p->b);
        return fail(ESCH_INVALID_ARG, "null output");
    }
    esch_status st = ESCH_OK;
    if (p->n) st = esch_batch_read_scores16(p->b, scores, nullptr);
    else st = esch_batch_wait(p->b);
    give_batch(s, p->b);
    if (st) return st;
    if (p->n) memcpy(scored, p->scored.data(), p->n);
    return ESCH_OK;
}

esch_status esch_score_batch(esch_scorer *s, const uint8_t *const *queries, const size_t *lens, size_t n,
                             int16_t *scores, uint8_t *scored) {
    esch_prepared *p = esch_prepare(s, queries, lens, n);
    if (!p) return ESCH_INVALID_ARG;
    esch_pending *pend = nullptr;
    if (esch_status st = esch_submit(s, p, &pend)) return st;
    return esch_finish(s, pend, scores, scored);
}

}  // extern "C"

namespace {

// Select k_dp16 variant `variant` for the scorer's runs. fixed: the caller
// chose it (no per-batch fp16 upgrade, batch_variant); otherwise it is the
// autotuner's pick, which a batch of short enough reads still upgrades.
esch_status set_variant(esch_scorer *s, int variant, bool fixed) {
    if (!s || variant < 0 || variant >= kDp16Variants)
        return fail(ESCH_INVALID_ARG, "variant %d outside 0..%d", variant, kDp16Variants - 1);
    if (esch::dp16_is_f16(variant) && !s->f16_ok)
        return fail(ESCH_INVALID_ARG,
                    "variant %d (fp16) needs longest reference x max(|match|, |mismatch|) + |open| + "
                    "|extend| <= %lld",
                    variant, (long long)esch::kBoundF16);
    if (!s->usable16[variant] && esch::dp16_is_diag(variant))
        return fail(ESCH_INVALID_ARG,
                    s->sc.local ? "variant %d (diagonal shift) is a semiglobal form; local mode refuses it"
                                : "variant %d (diagonal shift) needs longest reference x max(|match|, |mismatch|) + "
                                  "8 |extend| + |open| + 34 |extend| + max(|match|, |mismatch|) <= %lld (its "
                                  "column shift restarted every 8 columns at the least)",
                    variant, (long long)(esch::dp16_is_f16(variant) ? esch::kBoundF16 : esch::kBound16));
    if (!s->usable16[variant])
        return fail(ESCH_INVALID_ARG,
                    s->sc.local ? "variant %d (column-shifted in local mode) needs longest reference x "
                                  "max(|match|, |mismatch|) + 8 |extend| + |open| + 3 |extend| + "
                                  "max(|match|, |mismatch|) <= %lld (its shift restarted every 8 columns at the "
                                  "least)"
                                : "variant %d (row-shifted) needs longest reference x max(|match|, |mismatch|) + "
                                  "|open| + (stripe rows + 1) |extend| <= %lld",
                    variant, (long long)(esch::dp16_is_f16(variant) ? esch::kBoundF16 : esch::kBound16));
    std::lock_guard<std::mutex> g(s->geo_mu);
    s->geo.variant16 = variant;
    s->variant_fixed = fixed;
    s->geo.grid16 = s->grid16_v[variant];
    s->regs16 = s->regs16_v[variant];
    s->local16 = s->local16_v[variant];
    return ESCH_OK;
}

// The routing ratio, and the plan of every length class with it.
void set_ratio(esch_scorer *s, int32_t ratio_x16) {
    std::lock_guard<std::mutex> g(s->geo_mu);
    s->diag_ratio = ratio_x16;
    for (int c = 0; c < esch::kLenClasses; ++c) s->diag_plan[c] = ratio_x16;
}

void set_plan_entry(esch_scorer *s, int c, int32_t v) {
    std::lock_guard<std::mutex> g(s->geo_mu);
    s->diag_plan[c] = v;
}

bool plan_value_ok(int32_t v) { return v == 0 || v == 1 || (v >= 16 && v <= 1024); }

}  // namespace

extern "C" {

esch_status esch_set_variant(esch_scorer *s, int variant) { return set_variant(s, variant, true); }

int esch_get_variant(const esch_scorer *s) {
    if (!s) return -1;
    std::lock_guard<std::mutex> g(s->geo_mu);
    return s->geo.variant16;
}

esch_status esch_set_swap_rows(esch_scorer *s, int32_t rows) {
    if (!s || rows < -1) return fail(ESCH_INVALID_ARG, "swap rows must be -1 (the rule), 0 (never) or a row count");
    s->swap_rows = rows;
    return ESCH_OK;
}

int32_t esch_get_swap_rows(const esch_scorer *s) { return s ? (int32_t)swap_limit(s) : -1; }

int esch_batch_swapped(const esch_batch *b) { return b && b->last_swapped ? 1 : 0; }

int esch_batch_variant(const esch_batch *b) { return b ? b->last_variant : -1; }

esch_status esch_set_diag_min_len(esch_scorer *s, int32_t min_len) {
    if (!s || min_len < 0) return fail(ESCH_INVALID_ARG, "bad anti-diagonal threshold");
    s->diag_min_len = min_len;
    return ESCH_OK;
}

int32_t esch_get_diag_min_len(const esch_scorer *s) { return s ? s->diag_min_len : -1; }

esch_status esch_set_diag_ratio(esch_scorer *s, int32_t ratio_x16) {
    if (!s) return fail(ESCH_INVALID_ARG, "null scorer");
    if (ratio_x16 != 0 && (ratio_x16 < 16 || ratio_x16 > 1024))
        return fail(ESCH_INVALID_ARG, "diag ratio must be 0 or 16..1024 (sixteenths)");
    std::lock_guard<std::mutex> g(s->geo_mu);
    s->diag_ratio = ratio_x16;
    for (int c = 0; c < esch::kLenClasses; ++c) s->diag_plan[c] = ratio_x16;
    s->diag_ratio_tunable = false;
    return ESCH_OK;
}

esch_status esch_set_diag_plan(esch_scorer *s, const int32_t *plan, size_t n) {
    if (!s || !plan || n != (size_t)esch::kLenClasses)
        return fail(ESCH_INVALID_ARG, "the plan holds %d length classes", esch::kLenClasses);
    for (size_t c = 0; c < n; ++c)
        if (!plan_value_ok(plan[c]))
            return fail(ESCH_INVALID_ARG, "plan entry %zu is %d: 0 (inter-sequence), 1 (anti-diagonal) or 16..1024",
                        c, plan[c]);
    std::lock_guard<std::mutex> g(s->geo_mu);
    for (size_t c = 0; c < n; ++c) s->diag_plan[c] = plan[c];
    s->diag_ratio_tunable = false;
    return ESCH_OK;
}

esch_status esch_get_diag_plan(const esch_scorer *s, int32_t *plan, size_t n) {
    if (!s || !plan || n != (size_t)esch::kLenClasses)
        return fail(ESCH_INVALID_ARG, "the plan holds %d length classes", esch::kLenClasses);
    std::lock_guard<std::mutex> g(s->geo_mu);
    for (size_t c = 0; c < n; ++c) plan[c] = s->diag_plan[c];
    return ESCH_OK;
}

int32_t esch_get_diag_ratio(const esch_scorer *s) {
    if (!s) return -1;
    std::lock_guard<std::mutex> g(s->geo_mu);
    return s->diag_ratio;
}
int esch_n_variants(void) { return kDp16Variants; }

esch_status esch_variant_info(int variant, int *schedule, int *stripe_rows, int *f16, int *shifted) {
    if (variant < 0 || variant >= kDp16Variants)
        return fail(ESCH_INVALID_ARG, "variant %d outside 0..%d", variant, kDp16Variants - 1);
    if (schedule) *schedule = esch::dp16_schedule(variant);
    if (stripe_rows) *stripe_rows = esch::dp16_stripe(variant);
    if (f16) *f16 = esch::dp16_is_f16(variant) ? 1 : 0;
    if (shifted) *shifted = esch::dp16_is_diag(variant) ? 2 : esch::dp16_is_shifted(variant) ? 1 : 0;
    return ESCH_OK;
}

int esch_variant_usable(const esch_scorer *s, int variant) {
    if (!s || variant < 0 || variant >= kDp16Variants) return 0;
    return s->usable16[variant] ? 1 : 0;
}

int esch_column_block(const esch_scorer *s, int variant) {
    if (!s || variant < 0 || variant >= kDp16Variants || !s->usable16[variant]) return -1;
    const bool cols = esch::dp16_is_diag(variant) || (s->sc.local && esch::dp16_is_shifted(variant));
    if (!cols) return 0;
    const int k = esch::dp16_is_f16(variant) ? s->geo.cshift_f16 : s->geo.cshift_i16;
    return k > 0 ? 1 << k : 0;
}

esch_status esch_autotune(esch_scorer *s, const uint8_t *const *reads, const size_t *lens, size_t n,
                          int both, float *ms, int *chosen) {
    if (!s || (n && (!reads || !lens))) return fail(ESCH_INVALID_ARG, "null argument");
    size_t total = 0;
    for (size_t k = 0; k < n; ++k) total += lens[k];
    esch_batch *b = nullptr;
    if (esch_status st = take_batch(s, n, total, &b)) return st;
    esch_status st = esch_batch_load(b, reads, lens, n, both, 0);
    // one untimed pass, then the timed one; DP time of the whole DP stage
    auto timed = [&](float *out) -> esch_status {
        for (int rep = 0; rep < 2; ++rep) {
            if (esch_status e = esch_batch_run(b, ESCH_RUN_SCORES32 | ESCH_RUN_NO_GRAPH, 0, -1, 1)) return e;
            if (esch_status e = esch_batch_wait(b)) return e;
        }
        return esch_batch_timing(b, out, nullptr);
    };
    // 1. the inter-sequence kernel's variants, routing as configured; the
    // scorer's current variant stays unless another is faster by more than
    // 1% (near-ties, such as schedules 12 and 13 on 32 rows, would
    // otherwise flip between runs)
    const int v_in = esch_get_variant(s);
    int best_v = v_in;
    float best_ms = 0.0f, in_ms = -1.0f;
    bool any = false;
    for (int v = 0; !st && v < kDp16Variants; ++v) {
        if (!s->usable16[v]) {
            if (ms) ms[v] = -1.0f;  // not applicable to this panel
            continue;
        }
        st = set_variant(s, v, true);  // exactly this variant while it is timed
        float dp = 0.0f;
        if (!st) st = timed(&dp);
        if (st) break;
        if (ms) ms[v] = dp;
        if (v == v_in) in_ms = dp;
        if (!any || dp < best_ms) {
            best_ms = dp;
            best_v = v;
            any = true;
        }
    }
    if (in_ms > 0.0f && best_ms >= 0.99f * in_ms) best_v = v_in;
    // the pick is the autotuner's, not the caller's: a batch whose reads
    // admit an fp16 variant the panel does not still takes it
    if (!st) st = set_variant(s, best_v, false);
    // 2. the per-length-bucket routing between the inter-sequence and the
    // anti-diagonal kernel: the ratio rho is the anti-diagonal kernel's
    // per-cell cost relative to the inter-sequence one, a property of the
    // two kernels on this panel rather than of the read distribution, so a
    // sample measures it. Another ratio replaces the default only when it
    // is faster by more than 2% (a sample without long tails times the same
    // under every ratio and keeps the default).
    bool tunable;
    {
        std::lock_guard<std::mutex> g(s->geo_mu);
        tunable = s->diag_ratio_tunable;
    }
    if (!st && tunable) {
        static const int32_t kRatios[] = {0, 24, 32, 48, 64};
        const int32_t base = kDiagRatioDefault;
        float base_ms = 0.0f;
        set_ratio(s, base);
        st = timed(&base_ms);
        int32_t best_r = base;
        float best_r_ms = base_ms;
        for (int32_t r : kRatios) {
            if (st || r == base) continue;
            set_ratio(s, r);
            float dp = 0.0f;
            st = timed(&dp);
            if (!st && dp < best_r_ms && dp < 0.98f * base_ms) {
                best_r_ms = dp;
                best_r = r;
            }
        }
        set_ratio(s, best_r);
        // 3. per length class (the per-bucket plan): for each class the
        // sample holds reads of (128 nt and over), keep the class on the
        // inter-sequence kernel, send it to the anti-diagonal kernel, or
        // route it by its own ratio, whichever is more than 2% faster than
        // the class's current choice
        int counts[esch::kLenClasses] = {};
        for (size_t k = 0; k < n; ++k) ++counts[esch::len_class((int)std::min<size_t>(lens[k], INT32_MAX))];
        if (!st) {
            float cur_ms = 0.0f;
            st = timed(&cur_ms);
            static const int32_t kOptions[] = {0, 1, 24, 32, 48, 64};
            for (int c = 1; !st && c < esch::kLenClasses; ++c) {
                if (!counts[c]) continue;
                const int32_t keep = best_r;
                int32_t best_o = keep;
                float best_o_ms = cur_ms;
                for (int32_t opt : kOptions) {
                    if (st || opt == keep) continue;
                    set_plan_entry(s, c, opt);
                    float dp = 0.0f;
                    st = timed(&dp);
                    if (!st && dp < best_o_ms && dp < 0.98f * cur_ms) {
                        best_o_ms = dp;
                        best_o = opt;
                    }
                }
                set_plan_entry(s, c, best_o);
                cur_ms = best_o_ms;
            }
        }
    }
    give_batch(s, b);
    if (st) return st;
    if (chosen) *chosen = best_v;
    return ESCH_OK;
}

esch_status esch_score_batch_i32(esch_scorer *s, const uint8_t *const *queries, const size_t *lens, size_t n,
                                 int32_t *scores, uint8_t *scored) {
    if (!s || (n && (!queries || !lens || !scores || !scored))) return fail(ESCH_INVALID_ARG, "null argument");
    size_t total = 0;
    for (size_t k = 0; k < n; ++k) total += lens[k];
    esch_batch *b = nullptr;
    if (esch_status st = take_batch(s, n, total, &b)) return st;
    esch_status st = esch_batch_load(b, queries, lens, n, 0, 0);
    if (!st) st = esch_batch_run(b, ESCH_RUN_SCORES32, 0, -1, 1);
    if (!st) st = esch_batch_read_scores32(b, scores, scored);
    give_batch(s, b);
    return st;
}

esch_status esch_select_batch(esch_scorer *s, const uint8_t *const *reads, const size_t *lens, size_t n,
                              int both, int32_t min_score, int64_t max_ties, size_t tie_cap,
                              esch_read_select *out, int32_t *ties) {
    if (!s || (n && (!reads || !lens || !out))) return fail(ESCH_INVALID_ARG, "null argument");
    size_t total = 0;
    for (size_t k = 0; k < n; ++k) total += lens[k];
    esch_batch *b = nullptr;
    if (esch_status st = take_batch(s, n, total, &b)) return st;
    esch_status st = esch_batch_load(b, reads, lens, n, both, 0);
    if (!st) st = esch_batch_run(b, ESCH_RUN_SCORES32 | ESCH_RUN_SELECT, min_score, max_ties, tie_cap);
    if (!st) st = esch_batch_read_select(b, out, ties);
    give_batch(s, b);
    return st;
}

esch_status esch_select_pairs(esch_scorer *s, const uint8_t *const *mates1, const size_t *lens1,
                              const uint8_t *const *mates2, const size_t *lens2, size_t n, int both_strands,
                              int32_t min_score, size_t top_cap, esch_pair_select *out, int32_t *top) {
    if (!s || (n && (!mates1 || !lens1 || !mates2 || !lens2 || !out)))
        return fail(ESCH_INVALID_ARG, "null argument");
    std::vector<const uint8_t *> ptrs(2 * n);
    std::vector<size_t> lens(2 * n);
    size_t total = 0;
    for (size_t k = 0; k < n; ++k) {
        ptrs[2 * k] = mates1[k];
        ptrs[2 * k + 1] = mates2[k];
        lens[2 * k] = lens1[k];
        lens[2 * k + 1] = lens2[k];
        total += lens1[k] + lens2[k];
    }
    esch_batch *b = nullptr;
    if (esch_status st = take_batch(s, 2 * n, total, &b)) return st;
    esch_status st = esch_batch_load(b, ptrs.data(), lens.data(), 2 * n, 1, 1);
    const int what = ESCH_RUN_SCORES32 | ESCH_RUN_PAIRS | (both_strands ? ESCH_RUN_PAIR_BOTH : 0);
    if (!st) st = esch_batch_run(b, what, min_score, -1, top_cap);
    if (!st) st = esch_batch_read_pairs(b, out, top);
    give_batch(s, b);
    return st;
}

}  // extern "C"

// ------------------------------------------------------------ winner traceback

namespace {

// Trace + boundary bytes per launch: a call whose regions do not fit is cut
// into launches run one after the other, each with its own tail, so the
// arena is sized for HBM (16 GiB holds the regions of a million winners).
constexpr size_t kTraceArenaMax = (size_t)16 << 30;
constexpr size_t kTraceArenaKeep = (size_t)4 << 30;     // cached across calls up to this
constexpr uint64_t kTraceMaxCells = (uint64_t)1 << 27;  // read x reference: a 4 GiB group region (one lane a pair)
constexpr uint64_t kTraceDiagMinCells = (uint64_t)1 << 15;  // below this a pair never takes the wavefront kernel

// Trace bytes of one pair of the anti-diagonal traceback (k_trace_diag).
uint64_t trace_diag_bytes(size_t n, int m) {
    const int R = trace_diag_rows((int)n);
    const uint64_t passes = (n + 64 * (size_t)R - 1) / (64 * (size_t)R);
    return passes * ((uint64_t)m + 63) * 256;
}

}  // namespace

// Device buffers, pinned staging, stream and events of esch_align_pairs,
// kept by the scorer between calls (one call at a time per scorer).
struct TraceCache {
    std::mutex mu;
    hipStream_t st = nullptr, st2 = nullptr;  // one-lane kernel; wavefront kernel beside it
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    DevBuf codes, qoff, qlen, ref, slot, gbase, gS, gM, dbase, arena, arena2, recs, ops_wc, ops_off, out_off, out;
    HostBuf h_codes, h_recs, h_off, h_ops, h_qoff, h_qlen, h_ref, h_slot, h_opsoff;
    ~TraceCache() {
        if (st) (void)hipStreamSynchronize(st);
        for (HostBuf *h : {&h_codes, &h_recs, &h_off, &h_ops, &h_qoff, &h_qlen, &h_ref, &h_slot, &h_opsoff})
            h->release();
        if (st2) (void)hipStreamSynchronize(st2);
        for (DevBuf *b : {&codes, &qoff, &qlen, &ref, &slot, &gbase, &gS, &gM, &dbase, &arena, &arena2, &recs,
                          &ops_wc, &ops_off, &out_off, &out})
            b->release();
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (e2) (void)hipEventDestroy(e2);
        if (st) (void)hipStreamDestroy(st);
        if (st2) (void)hipStreamDestroy(st2);
    }
};

void trace_cache_free(esch_scorer *s) {
    if (s->trace) {
        (void)hipSetDevice(s->ordinal);
        delete s->trace;
        s->trace = nullptr;
    }
}

// The traceback of n pairs whose reads come from get(k) -> (codes, length).
esch_status esch::align_core(esch_scorer *s, size_t n,
                             const std::function<void(size_t, const uint8_t *&, size_t &)> &get,
                             const int32_t *refs, esch_alignment *out, uint32_t *ops, size_t ops_cap,
                             size_t *ops_used, const uint8_t *dev_codes, std::vector<uint32_t> *ops_vec,
                             esch_batch *pump) {
    if (ops_used) *ops_used = 0;
    std::unique_ptr<size_t[]> lens(new size_t[std::max<size_t>(n, 1)]);
    std::unique_ptr<const uint8_t *[]> ptr(new const uint8_t *[std::max<size_t>(n, 1)]);
    // The caller's pairs, checked on up to 16 host threads (a mapping call
    // traces millions; from 2^16 pairs on, the check and the staging below
    // are threaded: 210k winner pairs traced from host buffers in 12.8-13.8
    // ms against 14.7-15.4 on one thread and 8); an error names the first
    // bad pair, raised here on the calling thread (the message is per
    // thread).
    struct Part {
        size_t np = 0, bases = 0, max_len = 0, bad = SIZE_MAX;
        uint64_t W = 0, max_cells = 0;
        int why = 0;
    };
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = n < ((size_t)1 << 16) ? 1 : std::min<size_t>(16, hw);
    std::vector<Part> part(nt);
    auto check = [&](size_t t) {
        Part &P = part[t];
        for (size_t k = n * t / nt; k < n * (t + 1) / nt; ++k) {
            get(k, ptr[k], lens[k]);
            const size_t len = lens[k];
            if (refs[k] < 0 || (size_t)refs[k] >= s->n_refs) {
                P.bad = k;
                P.why = 1;
                return;
            }
            // one lane a pair: a group region is 64 pairs x (read / 16
            // stripes) x reference x 8 B; a larger pair takes the wavefront
            // kernel, whose region (read x (reference + 63) / 2 B) must fit
            // the arena
            const int rl = s->h_rlen[refs[k]];
            const uint64_t cells = (uint64_t)len * (uint64_t)rl;
            if (cells > kTraceMaxCells && (rl > kTraceDiagMaxRef || trace_diag_bytes(len, rl) > kTraceArenaMax)) {
                P.bad = k;
                P.why = 2;
                return;
            }
            if (len && !ptr[k]) {
                P.bad = k;
                P.why = 3;
                return;
            }
            out[k] = esch_alignment{};  // lib/scalar.rs:185-195: empty in, empty out
            if (len) {
                ++P.np;
                P.bases += len;
                P.max_len = std::max(P.max_len, len);
                P.W += cells;
                P.max_cells = std::max(P.max_cells, cells);
            }
        }
    };
    {
        std::vector<std::thread> pool;
        for (size_t t = 1; t < nt; ++t) pool.emplace_back(check, t);
        check(0);
        for (auto &th : pool) th.join();
    }
    size_t np = 0, bases = 0, max_len = 0;
    uint64_t W = 0, max_cells = 0;
    for (const Part &P : part) {
        if (P.bad != SIZE_MAX) {
            const size_t k = P.bad;
            if (P.why == 1) return fail(ESCH_INVALID_ARG, "pair %zu names reference %d of %zu", k, refs[k], s->n_refs);
            if (P.why == 2)
                return fail(ESCH_INVALID_ARG, "pair %zu: %zu x %d cells is over the device traceback bound", k,
                            lens[k], s->h_rlen[refs[k]]);
            return fail(ESCH_INVALID_ARG, "pair %zu: null read", k);
        }
        np += P.np;
        bases += P.bases;
        max_len = std::max(max_len, P.max_len);
        W += P.W;
        max_cells = std::max(max_cells, P.max_cells);
    }
    if (np == 0) return ESCH_OK;
    // Groups of similar read and reference lengths (lib/simd/mod.rs:433
    // sorts ascending), longest first: a wave's time is its group's longest
    // read x longest reference, so the longest groups are dispatched first.
    // Order: read length descending, then reference length descending,
    // stable; one counting sort on the pair of lengths when its bins are
    // few, else two (by reference length, then by read length).
    std::vector<int32_t> order(np);
    {
        const size_t RM = (size_t)s->max_len + 1, QM = max_len + 1;
        if (RM * QM <= ((size_t)1 << 20)) {
            std::vector<uint32_t> cnt(RM * QM + 1, 0);
            auto key = [&](size_t k) { return (max_len - lens[k]) * RM + ((size_t)s->max_len - (size_t)s->h_rlen[refs[k]]); };
            for (size_t k = 0; k < n; ++k)
                if (lens[k]) ++cnt[key(k)];
            uint32_t run = 0;
            for (uint32_t &c : cnt) {
                const uint32_t x = c;
                c = run;
                run += x;
            }
            for (size_t k = 0; k < n; ++k)
                if (lens[k]) order[cnt[key(k)]++] = (int32_t)k;
        } else {
            std::vector<int32_t> tmp(np);
            std::vector<size_t> cnt((size_t)s->max_len + 2, 0);
            for (size_t k = 0; k < n; ++k)
                if (lens[k]) ++cnt[s->max_len - (size_t)s->h_rlen[refs[k]]];
            size_t run = 0;
            for (size_t &c : cnt) {
                const size_t x = c;
                c = run;
                run += x;
            }
            for (size_t k = 0; k < n; ++k)
                if (lens[k]) tmp[cnt[s->max_len - (size_t)s->h_rlen[refs[k]]]++] = (int32_t)k;
            std::vector<size_t> cl(max_len + 2, 0);
            for (size_t i = 0; i < np; ++i) ++cl[max_len - lens[tmp[i]]];
            run = 0;
            for (size_t &c : cl) {
                const size_t x = c;
                c = run;
                run += x;
            }
            for (size_t i = 0; i < np; ++i) order[cl[max_len - lens[tmp[i]]]++] = tmp[i];
        }
    }
    // Pairs for the wavefront kernel first: those a lane could not hold, and
    // those long enough that one lane would set the call's time: more cells
    // than 2^-18 of the call's (a device of 1,024 SIMDs traces about 2^16
    // lanes' worth at once, and a long pair's lane shares its SIMD with a
    // few other waves), and at least 2^15 (below that the wave per pair of
    // the wavefront kernel costs more than it saves). Measured on the 1M-
    // and 200k-read winner sets: 17.5 and 5.1 ms against 22.2 and 12.1 ms
    // with every pair on the one-lane kernel. No pair qualifies (the usual
    // case for short reads) when the largest is under both bounds.
    size_t n_diag = 0;
    const uint64_t thr = std::max<uint64_t>(kTraceDiagMinCells, W >> 18);
    if (max_cells >= thr || max_cells > kTraceMaxCells) {
        std::vector<int32_t> dl, rest;
        rest.reserve(np);
        for (size_t i = 0; i < np; ++i) {
            const int32_t k = order[i];
            const int rl = s->h_rlen[refs[k]];
            const uint64_t cells = (uint64_t)lens[k] * (uint64_t)rl;
            const bool diag = rl <= kTraceDiagMaxRef && (cells > kTraceMaxCells || (lens[k] >= 128 && cells >= thr));
            (diag ? dl : rest).push_back(k);
        }
        n_diag = dl.size();
        if (n_diag) {
            std::copy(dl.begin(), dl.end(), order.begin());
            std::copy(rest.begin(), rest.end(), order.begin() + (long)n_diag);
        }
    }
    HIPC(hipSetDevice(s->ordinal));
    if (!s->trace) s->trace = new TraceCache;
    TraceCache &T = *s->trace;
    std::lock_guard<std::mutex> lock(T.mu);
    if (T.st) HIPC(hipStreamSynchronize(T.st));  // the pinned staging may still feed a copy
    if (!dev_codes) HIPC(T.h_codes.need(bases));
    HIPC(T.h_recs.need(n * kTraceRec * 4));
    // per sorted pair (pinned, kept between calls): read offset, length,
    // reference, caller index; per caller pair: its worst-case runs' offset
    HIPC(T.h_qoff.need(np * 8));
    HIPC(T.h_qlen.need(np * 4));
    HIPC(T.h_ref.need(np * 4));
    HIPC(T.h_slot.need(np * 4));
    HIPC(T.h_opsoff.need(n * 8));
    uint8_t *const hcodes = T.h_codes.as<uint8_t>();
    uint64_t *const hqoff = T.h_qoff.as<uint64_t>();
    int32_t *const hqlen = T.h_qlen.as<int32_t>(), *const href = T.h_ref.as<int32_t>();
    int32_t *const hslot = T.h_slot.as<int32_t>();
    uint64_t *const ops_off = T.h_opsoff.as<uint64_t>();
    {
        // random reads of the caller's arrays, so spread over threads
        auto fill = [&](size_t i0, size_t i1) {
            for (size_t i = i0; i < i1; ++i) {
                const int32_t k = order[i];
                hqoff[i] = dev_codes ? (uint64_t)(ptr[k] - dev_codes) : 0;
                hqlen[i] = (int32_t)lens[k];
                href[i] = refs[k];
                hslot[i] = k;
            }
        };
        const size_t ft = np < ((size_t)1 << 16) ? 1 : std::min<size_t>(16, hw);
        std::vector<std::thread> pool;
        for (size_t t = 1; t < ft; ++t) pool.emplace_back(fill, np * t / ft, np * (t + 1) / ft);
        fill(0, np / ft);
        for (auto &th : pool) th.join();
        if (!dev_codes) {  // staged reads are packed in sorted order
            size_t at = 0;
            for (size_t i = 0; i < np; ++i) {
                hqoff[i] = at;
                at += (size_t)hqlen[i];
            }
        }
    }
    // the gather into staging and the code check on up to 16 host threads
    // (resident reads were checked when they were loaded)
    if (!dev_codes) {
        const size_t gt = bases < ((size_t)8 << 20) ? 1 : std::min<size_t>(16, hw);
        std::vector<uint8_t> bad(gt, 0);
        auto gather = [&](size_t t) {
            const size_t i0 = np * t / gt, i1 = np * (t + 1) / gt;
            uint8_t mx = 0;
            for (size_t i = i0; i < i1; ++i) {
                // the reads sit anywhere in the caller's memory: fetch ahead
                if (i + 16 < i1) __builtin_prefetch(ptr[order[i + 16]]);
                const uint8_t *q = ptr[order[i]];
                uint8_t *d = hcodes + hqoff[i];
                memcpy(d, q, (size_t)hqlen[i]);
                for (int32_t j = 0; j < hqlen[i]; ++j) mx = d[j] > mx ? d[j] : mx;
            }
            bad[t] = mx > 4;
        };
        if (gt == 1) {
            gather(0);
        } else {
            std::vector<std::thread> pool;
            for (size_t t = 1; t < gt; ++t) pool.emplace_back(gather, t);
            gather(0);
            for (auto &th : pool) th.join();
        }
        for (size_t t = 0; t < gt; ++t) {
            if (!bad[t]) continue;
            for (size_t k = 0; k < n; ++k)
                for (size_t j = 0; j < lens[k]; ++j)
                    if (ptr[k][j] > 4)
                        return fail(ESCH_INVALID_ARG, "read %zu carries code %d (codes are 0..=4)", k, ptr[k][j]);
        }
    }
    size_t wc = 0;  // worst-case runs: read + reference length per pair
    for (size_t k = 0; k < n; ++k) {
        ops_off[k] = wc;
        if (lens[k]) wc += lens[k] + (size_t)s->h_rlen[refs[k]];
    }
    s->trace_diag = n_diag;
    const size_t n_lane = np - n_diag;  // pairs of the one-lane kernel, after the wavefront ones
    const int ng = (int)((n_lane + 63) / 64);
    std::vector<int32_t> gS(ng), gM(ng);
    std::vector<uint64_t> gbytes(ng), gbase(ng);
    for (int g = 0; g < ng; ++g) {
        int32_t mx_n = 0, mx_m = 0;
        for (size_t i = n_diag + (size_t)g * 64; i < std::min(np, n_diag + (size_t)g * 64 + 64); ++i) {
            mx_n = std::max(mx_n, hqlen[i]);
            mx_m = std::max(mx_m, s->h_rlen[href[i]]);
        }
        gS[g] = (mx_n + 15) / 16;
        gM[g] = mx_m;
        gbytes[g] = ((uint64_t)gS[g] * gM[g] + gM[g]) * 512;
    }
    // launches of consecutive wavefront pairs, then of consecutive groups,
    // whose regions fit the arena together
    std::vector<uint64_t> dbase(std::max<size_t>(n_diag, 1));
    std::vector<std::pair<int, int>> dchunks, chunks;
    size_t arena = 0, arena2 = 0;  // the wavefront kernel runs beside the other: its own arena
    int diag_max_ref = 0;
    for (size_t d0 = 0; d0 < n_diag;) {
        uint64_t sum = 0;
        size_t d1 = d0;
        while (d1 < n_diag) {
            const uint64_t b = trace_diag_bytes((size_t)hqlen[d1], s->h_rlen[href[d1]]);
            if (d1 > d0 && sum + b > kTraceArenaMax) break;
            dbase[d1] = sum;
            sum += b;
            diag_max_ref = std::max(diag_max_ref, s->h_rlen[href[d1]]);
            ++d1;
        }
        dchunks.emplace_back((int)d0, (int)d1);
        arena2 = std::max<size_t>(arena2, sum);
        d0 = d1;
    }
    for (int g0 = 0; g0 < ng;) {
        uint64_t sum = 0;
        int g1 = g0;
        while (g1 < ng && (g1 == g0 || sum + gbytes[g1] <= kTraceArenaMax)) {
            gbase[g1] = sum;
            sum += gbytes[g1];
            ++g1;
        }
        chunks.emplace_back(g0, g1);
        arena = std::max<size_t>(arena, sum);
        g0 = g1;
    }

    if (!T.st) HIPC(hipStreamCreateWithFlags(&T.st, hipStreamNonBlocking));
    if (!T.e0) HIPC(hipEventCreate(&T.e0));
    if (!T.e1) HIPC(hipEventCreate(&T.e1));
    if (n_diag && !T.st2) HIPC(hipStreamCreateWithFlags(&T.st2, hipStreamNonBlocking));
    if (n_diag && !T.e2) HIPC(hipEventCreateWithFlags(&T.e2, hipEventDisableTiming));
    if (!dev_codes) HIPC(T.codes.need(bases));
    HIPC(T.qoff.need(np * 8));
    HIPC(T.qlen.need(np * 4));
    HIPC(T.ref.need(np * 4));
    HIPC(T.slot.need(np * 4));
    HIPC(T.gbase.need((size_t)ng * 8));
    HIPC(T.gS.need((size_t)ng * 4));
    HIPC(T.gM.need((size_t)ng * 4));
    HIPC(T.dbase.need(dbase.size() * 8));
    HIPC(T.arena.need(arena));
    if (n_diag) HIPC(T.arena2.need(arena2));
    HIPC(T.recs.need(n * kTraceRec * 4));
    HIPC(T.ops_wc.need(wc * 4));
    HIPC(T.ops_off.need(n * 8));
    hipStream_t st = T.st;
    if (!dev_codes) HIPC(hipMemcpyAsync(T.codes.p, hcodes, bases, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(T.qoff.p, hqoff, np * 8, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(T.qlen.p, hqlen, np * 4, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(T.ref.p, href, np * 4, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(T.slot.p, hslot, np * 4, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(T.gbase.p, gbase.data(), (size_t)ng * 8, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(T.gS.p, gS.data(), (size_t)ng * 4, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(T.gM.p, gM.data(), (size_t)ng * 4, hipMemcpyHostToDevice, st));
    if (n_diag) HIPC(hipMemcpyAsync(T.dbase.p, dbase.data(), n_diag * 8, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(T.ops_off.p, ops_off, n * 8, hipMemcpyHostToDevice, st));
    // empty pairs are never traced: their records must read as zero runs
    HIPC(hipMemsetAsync(T.recs.p, 0, n * kTraceRec * 4, st));
    TraceArgs ta{s->panel, s->sc, dev_codes ? dev_codes : T.codes.as<uint8_t>(), T.qoff.as<uint64_t>(), T.qlen.as<int32_t>(),
                 T.ref.as<int32_t>(), T.slot.as<int32_t>(), T.gbase.as<uint64_t>(), T.gS.as<int32_t>(),
                 T.gM.as<int32_t>(), T.arena.as<uint8_t>(), T.recs.as<int32_t>(), T.ops_wc.as<uint32_t>(),
                 T.ops_off.as<uint64_t>(), (int32_t)np, 0, (int32_t)n_diag, T.dbase.as<uint64_t>(),
                 (int32_t)n_diag, 0};
    const bool local = s->sc.local != 0;
    HIPC(hipEventRecord(T.e0, st));
    if (n_diag) {
        // the long pairs' waves run beside the one-lane groups (after the
        // copies above, which e0 follows on st)
        HIPC(hipStreamWaitEvent(T.st2, T.e0, 0));
        TraceArgs td = ta;
        td.arena = T.arena2.as<uint8_t>();
        for (const auto &c : dchunks) {
            td.d0 = c.first;
            HIPC(launch_trace_diag(td, c.second - c.first, diag_max_ref, local, T.st2));
        }
        HIPC(hipEventRecord(T.e2, T.st2));
    }
    for (const auto &c : chunks) {
        ta.g0 = c.first;
        HIPC(launch_trace(ta, c.second - c.first, local, st));
    }
    if (n_diag) HIPC(hipStreamWaitEvent(st, T.e2, 0));
    HIPC(hipEventRecord(T.e1, st));
    int32_t *const recs = T.h_recs.as<int32_t>();
    HIPC(hipMemcpyAsync(recs, T.recs.p, n * kTraceRec * 4, hipMemcpyDeviceToHost, st));
    if (esch_status e = esch::batch_pump(pump, st)) return e;
    HIPC(hipStreamSynchronize(st));
    HIPC(hipEventElapsedTime(&s->trace_ms, T.e0, T.e1));
    if (T.arena.cap > kTraceArenaKeep) T.arena.release();
    if (T.arena2.cap > kTraceArenaKeep) T.arena2.release();
    HIPC(T.h_off.need(n * 8));
    uint64_t *const out_off = T.h_off.as<uint64_t>();  // pinned: copied to the device
    size_t runs = 0;
    for (size_t k = 0; k < n; ++k) {
        out_off[k] = runs;
        if (lens[k]) runs += (size_t)recs[k * kTraceRec + 5];
    }
    if (ops_used) *ops_used = runs;
    if (ops_vec) ops_vec->clear();  // filled from the pinned staging below
    if (!ops_vec && (runs > ops_cap || (runs && !ops)))
        return fail(ESCH_CAPACITY, "the alignments have %zu operation runs; ops holds %zu", runs, ops_cap);
    if (runs) {
        HIPC(T.out_off.need(n * 8));
        HIPC(T.out.need(runs * 4));
        HIPC(T.h_ops.need(runs * 4));
        // untraced (empty) pairs have zero runs and copy nothing
        HIPC(hipMemcpyAsync(T.out_off.p, out_off, n * 8, hipMemcpyHostToDevice, st));
        HIPC(launch_trace_compact(T.ops_wc.as<uint32_t>(), T.ops_off.as<uint64_t>(), T.recs.as<int32_t>(),
                                  T.out_off.as<uint64_t>(), T.out.as<uint32_t>(), (int32_t)n, st));
        HIPC(hipMemcpyAsync(T.h_ops.p, T.out.p, runs * 4, hipMemcpyDeviceToHost, st));
        if (esch_status e = esch::batch_pump(pump, st)) return e;
        HIPC(hipStreamSynchronize(st));
        // pinned staging to the caller's memory
        const uint32_t *h = T.h_ops.as<uint32_t>();
        if (ops_vec) ops_vec->assign(h, h + runs);
        else memcpy(ops, h, runs * 4);
    }
    for (size_t k = 0; k < n; ++k) {
        if (!lens[k]) continue;
        const int32_t *r = recs + k * kTraceRec;
        out[k] = esch_alignment{r[0], r[1], r[2], r[3], r[4], r[5], (int64_t)out_off[k]};
    }
    return ESCH_OK;
}

esch_status esch_align_pairs(esch_scorer *s, const uint8_t *const *queries, const size_t *lens,
                             const int32_t *refs, size_t n, esch_alignment *out, uint32_t *ops,
                             size_t ops_cap, size_t *ops_used) {
    if (!s || (n && (!queries || !lens || !refs || !out))) return fail(ESCH_INVALID_ARG, "null argument");
    return align_core(
        s, n,
        [&](size_t k, const uint8_t *&p, size_t &len) {
            p = queries[k];
            len = lens[k];
        },
        refs, out, ops, ops_cap, ops_used);
}

esch_status esch_align_pairs_concat(esch_scorer *s, const uint8_t *codes, const uint64_t *offs,
                                    const int32_t *lens, const int32_t *refs, size_t n, esch_alignment *out,
                                    uint32_t *ops, size_t ops_cap, size_t *ops_used) {
    if (!s || (n && (!codes || !offs || !lens || !refs || !out))) return fail(ESCH_INVALID_ARG, "null argument");
    return align_core(
        s, n,
        [&](size_t k, const uint8_t *&p, size_t &len) {
            p = codes + offs[k];
            len = (size_t)(lens[k] < 0 ? 0 : lens[k]);
        },
        refs, out, ops, ops_cap, ops_used);
}

float esch_align_last_ms(const esch_scorer *s) { return s ? s->trace_ms : 0.0f; }
size_t esch_align_last_wavefront(const esch_scorer *s) { return s ? s->trace_diag : 0; }
