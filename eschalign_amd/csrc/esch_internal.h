// Internal layout shared by the gfx950 kernels and the host library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace esch {

// Rows of the read held in registers per DP stripe.
constexpr int kStripe = 16;
// Query code of a padding row placed above a read (substitution score 0,
// which reproduces the DP's zero top border exactly).
constexpr uint8_t kPreCode = 5;
// Reads per wave in the batched kernel: 64 lanes x 2 packed int16 halves.
constexpr int kChunk = 128;
// Longest read of the int16 kernel (length-bucketing bins 0 .. kMaxLen16);
// a longer read goes to the int32 kernel.
constexpr int kMaxLen16 = 65536;
// Largest |value| bound under which the packed int16 recurrence is exact:
// every real cell, E/F and intermediate sum lies in
// [-(k*M + |o| + |e|), k*M] with k = min(read, ref) and M = max(|ma|, |mi|),
// and -32768 (saturating) is the only -inf, so a bound below 32767 keeps
// every real value away from both clamps.
constexpr int64_t kBound16 = 32000;
// The same bound for k_dp16's fp16 variant: fp16 holds every integer of
// magnitude up to 2048 exactly (11 significant bits), and -inf exists, so a
// panel whose value bound, at k = its longest reference, is at most 2048
// scores every read exactly in fp16 arithmetic.
constexpr int64_t kBoundF16 = 2048;
__host__ __device__ inline bool fits_f16(int64_t max_ref_len, int32_t ma, int32_t mi, int32_t o, int32_t e) {
    const int64_t M = ma < 0 ? -(int64_t)ma : ma, Mi = mi < 0 ? -(int64_t)mi : mi;
    const int64_t O = o < 0 ? -(int64_t)o : o, Ee = e < 0 ? -(int64_t)e : e;
    return max_ref_len * (M > Mi ? M : Mi) + O + Ee <= kBoundF16;
}
// k_dp16's row-shifted semiglobal column (variants 6 and 7) holds a stripe
// row r's values shifted by |e| r (r = -1 .. ST - 1 in a stripe of ST rows)
// and offset by o or e: every value lies in [-(kM + |o| + 4|e|),
// kM + ST |e|], and the profile's stripe-top entries s - o + e h (h <= ST)
// within M + |o| + ST |e|. With (ST + 1) |e| for the shift, the unshifted
// bounds carry over.
__host__ __device__ inline int64_t shifted_bound(int64_t max_ref_len, int32_t ma, int32_t mi, int32_t o, int32_t e,
                                                 int stripe = kStripe) {
    const int64_t M = ma < 0 ? -(int64_t)ma : ma, Mi = mi < 0 ? -(int64_t)mi : mi;
    const int64_t O = o < 0 ? -(int64_t)o : o, Ee = e < 0 ? -(int64_t)e : e;
    return max_ref_len * (M > Mi ? M : Mi) + O + (stripe + 1) * Ee;
}
// In local mode those variants hold reference column j's values shifted by
// |e| j (j = -1 .. L - 1): every value lies within
// [-(|o| + 3|e| + M), kM + |e| (L + 2) + M], the floor and the carried best
// included, so the bound grows with the longest reference L. Restarted
// every `block` columns (a power of two, 0: never), the shift spans at most
// a block, which takes L's place in the |e| term.
// The two lengths apart: k bounds the read-reference overlap (min(read,
// reference) over a batch), span the columns the shift spans.
__host__ __device__ inline int64_t colshift_bound_ks(int64_t k, int64_t span, int32_t ma, int32_t mi, int32_t o,
                                                     int32_t e) {
    const int64_t M = ma < 0 ? -(int64_t)ma : ma, Mi = mi < 0 ? -(int64_t)mi : mi;
    const int64_t O = o < 0 ? -(int64_t)o : o, Ee = e < 0 ? -(int64_t)e : e;
    const int64_t MM = M > Mi ? M : Mi;
    return k * MM + span * Ee + O + 3 * Ee + MM;
}
__host__ __device__ inline int64_t colshift_bound(int64_t max_ref_len, int32_t ma, int32_t mi, int32_t o, int32_t e,
                                                  int64_t block = 0) {
    const int64_t span = block > 0 && block < max_ref_len ? block : max_ref_len;
    return colshift_bound_ks(max_ref_len, span, ma, mi, o, e);
}
// The diagonal-shift semiglobal schedules (12-14) hold a value of stripe
// row r in reference column j shifted by |e| (r + j) (r = -1 .. ST - 1,
// j = -1 .. L - 1) and offset by o or e: every value lies within
// [-(kM + |o| + M + (ST + 1)|e|), kM + |e| (ST + L - 2)], the profile's
// stripe-top entries s - o - e + e h included, so the bound grows with the
// longest reference L as the column shift's does, and with the stripe as
// the row shift's does; a column part restarted every `block` columns (0:
// never) spans at most a block, which takes L's place in the |e| term.
__host__ __device__ inline int64_t diagshift_bound_ks(int64_t k, int64_t span, int32_t ma, int32_t mi, int32_t o,
                                                      int32_t e, int stripe) {
    const int64_t M = ma < 0 ? -(int64_t)ma : ma, Mi = mi < 0 ? -(int64_t)mi : mi;
    const int64_t O = o < 0 ? -(int64_t)o : o, Ee = e < 0 ? -(int64_t)e : e;
    const int64_t MM = M > Mi ? M : Mi;
    return k * MM + span * Ee + O + (stripe + 2) * Ee + MM;
}
__host__ __device__ inline int64_t diagshift_bound(int64_t max_ref_len, int32_t ma, int32_t mi, int32_t o, int32_t e,
                                                   int stripe, int64_t block = 0) {
    const int64_t span = block > 0 && block < max_ref_len ? block : max_ref_len;
    return diagshift_bound_ks(max_ref_len, span, ma, mi, o, e, stripe);
}
// The column-shift block (log2) of the column-shifted (local) or
// diagonal-shift (semiglobal) variants under `lim`, for overlaps of at
// most k on references of at most L: 0 (never restarted) when the whole
// reference fits, else the widest power of two from 4096 down to 8 that
// does; -1 when none does (the variants are refused). Restarting costs two
// packed adds per stripe row each block (1/4 op per cell at 8 columns,
// 1/16 at 32).
__host__ __device__ inline int column_block_log2(bool local, int64_t k, int64_t L, int32_t ma, int32_t mi, int32_t o,
                                                 int32_t e, int stripe, int64_t lim) {
    auto bound = [&](int64_t span) {
        return local ? colshift_bound_ks(k, span, ma, mi, o, e) : diagshift_bound_ks(k, span, ma, mi, o, e, stripe);
    };
    if (bound(L) <= lim) return 0;
    for (int b = 12; b >= 3; --b)
        if (((int64_t)1 << b) < L && bound((int64_t)1 << b) <= lim) return b;
    return -1;
}
// -inf of the int32 kernel, the oracle's i32::MIN / 4 (lib/scalar.rs:42).
constexpr int32_t kNeg32 = INT32_MIN / 4;

struct Scoring {
    int32_t ma, mi, o, e;
    int32_t local;
    // k_dp16's column-shifted (local) and diagonal-shift (semiglobal)
    // variants restart the column part of their shift every 2^cshift
    // columns (0: never); set per launch from the variant's arithmetic
    int32_t cshift;
};

// A panel resident on the device: 2-bit codes (16 per word, A0 C1 G2 T3,
// wildcards stored as 0) plus a wildcard bitmap (32 per word), per
// reference offsets in words.
struct DevPanel {
    const uint32_t *ref2;
    const uint32_t *refN;
    const uint32_t *w2off;
    const uint32_t *nmoff;
    const int32_t *rlen;
    const int32_t *rorder;  // panel indices, longest first (stable)
    int32_t n_refs;
    int32_t max_len;
    int64_t total_len;
};

// Length classes of the per-length-bucket routing plan: a read of length L
// is in class len_class(L): 0 under 128 nt (never routed), then 128-255,
// 256-511, 512-1023, 1024-2047, 2048-4095, 4096-8191, and 8192 and over.
constexpr int kLenClasses = 8;
__host__ __device__ inline int len_class(int L) {
    if (L < 128) return 0;
    int c = 1;
    for (int b = 256; c < kLenClasses - 1 && L >= b; b *= 2) ++c;
    return c;
}

// Device-resident batch descriptor: every kernel reads its sizes from here,
// so one captured graph replays for any batch that fits the buffers.
// Most launches a sliced k_dp16 run is cut into.
constexpr int kMaxSlices = 16;

struct DevDesc {
    int32_t n_queries;   // rows scored (reads, then reverse complements)
    int32_t n_reads;     // reads as loaded (mates interleaved when pairs)
    int32_t both;        // reverse complements appended
    int32_t pairs;
    int32_t dropin;      // classify by scores_len (int16 rows only)
    int32_t force32;     // score everything with the int32 kernel
    int32_t min_score;
    int32_t tie_cap;
    int64_t max_ties;
    int32_t n16, n32;    // queries on the int16 and int32 kernels
    uint32_t next16, next32a, next32b;  // work-queue cursors
    uint32_t next_slice[kMaxSlices];     // k_dp16's cursors in a sliced run, one per slice
    int32_t tstrands;    // pairs: template strands considered (1 or 2)
    int32_t diag_min_len;  // reads this long go to the anti-diagonal kernel (0: none)
    int32_t nd;          // queries on the anti-diagonal kernel
    uint32_t nextD;      // its work-queue cursor
    int32_t diag_ratio;  // adaptive routing to the anti-diagonal kernel, x16 (0: off)
    int32_t blocks16;    // resident blocks of the inter-sequence kernel's grid
    int32_t max16;       // longest query on the int16 path (bins above it are empty)
    uint64_t cells;      // sum of len x panel bases over scored queries
    // per length class (len_class), the routing of its length bins while
    // diag_ratio > 0: 0 = they stay on the inter-sequence kernel, 1 = they
    // take the anti-diagonal kernel, 16..1024 = the ratio rule with rho =
    // plan / 16 (diag_ratio itself rules the small-batch cases)
    int32_t plan[kLenClasses];
};

struct KArgs {
    DevPanel panel;
    Scoring sc;
    DevDesc *desc;
    const uint8_t *qcodes;
    const uint64_t *qoff;
    const int32_t *qlen;
    uint8_t *cls;        // 0 unscored, 1 int16 kernel, 2 int32 kernel
    int32_t *hist;       // [kMaxLen16 + 1] reads per length
    int32_t *bin_start;  // [kMaxLen16 + 1]
    int32_t *bin_fill;   // [kMaxLen16 + 1]
    int32_t *sorted16;   // int16-kernel queries, longest first
    int32_t *list32;     // int32-kernel queries
    int32_t *out32;      // [n_queries * n_refs] or null
    int16_t *out16;      // [n_queries * n_refs] or null
    uint2 *scratch16;    // [grid16 * 4 waves * (max_len + kScrPad) * 64]
    int2 *scratch32;     // [max_len * grid32 threads]
    int32_t *recs;       // select / pair records, 5 int32 each
    int32_t *ties;       // [n * tie_cap]
    uint8_t *rc_codes;   // where reverse complements are written
    int32_t *listD;      // anti-diagonal-kernel queries
    uint2 *scratchD;     // [grid_diag * 4 rows * max_len]
    // 0: out[q * n_refs + r]. 1: the roles are swapped (the queries are a
    // panel's references and `panel` holds the reads), and the score of
    // query q against reference r is the read r's against panel reference
    // q: out[r * n_queries + q], the layout of an ordinary run.
    int32_t swap;
    // k_dp16 in a sliced run (n_slices > 1): this launch scores slice
    // `slice` of the run's work items, [n_items slice / n_slices,
    // n_items (slice + 1) / n_slices), from its own cursor; 0 or 1: all
    int32_t slice, n_slices;
};

// Where a kernel writes the score of query q against reference r.
__device__ __forceinline__ size_t out_index(const KArgs &a, int q, int r) {
    return a.swap ? (size_t)r * (size_t)a.desc->n_queries + (size_t)q : (size_t)q * (size_t)a.panel.n_refs + (size_t)r;
}
// Rows per lane of the anti-diagonal kernel (a pass covers 16 * kDiagRows).
constexpr int kDiagRows = 8;
constexpr int kDiag64MaxLen = 4096;  // longest reference k_diag64 keeps in LDS
constexpr int kDiag64Pad = 192;      // LDS columns past max_len (block reads ahead)
constexpr int kScrPad = 8;  // scratch columns past max_len per wave (prefetch overrun)
constexpr int kDiagRatioDefault = 32;  // 2.0 in sixteenths

// Host-side launchers (kernels.hip); every kernel runs on `st`.
struct Launch {
    int grid16;   // persistent workgroups of k_dp16 (4 waves each)
    int grid32;   // persistent blocks of k_dp32
    int block32;  // threads per k_dp32 block
    int grid_aux; // blocks of the bucketing / selection kernels
    int variant16;  // k_dp16 variant: 0 profile per column, 1 ping-pong, 2 pair schedule, 3 quad schedule,
                    // 4 quad schedule in fp16, 5 eight-column schedule in fp16 (the
                    // fp16 variants: panels within kBoundF16 only), 6, 7 and 8: 3, 5
                    // and 4 with the row-shifted semiglobal column (shifted_bound),
                    // in local mode the column-shifted one (colshift_bound)
    int grid_diag;  // persistent waves of the anti-diagonal kernel
    size_t diag_lds;  // k_diag64's LDS boundary bytes; 0: k_diag16 (global scratch)
    // the column-shift block (column_block_log2) of the fp16 and the int16
    // shifted-column variants on this panel; launch_dp16 passes the one of
    // the variant it launches in Scoring::cshift
    int cshift_f16, cshift_i16;
};
// Variants 0-8 are k_dp16's schedules (above) on 16-row stripes; 9-16 run
// some of them on taller stripes: 9, 10, 11 = schedules 7, 8, 6 on 24 rows;
// 12, 13, 14 = 7, 8, 6 on 32 rows; 15, 16 = 4, 5 on 24 rows. A stripe
// boundary per 24 or 32 rows instead of 16 means less HBM traffic and the
// per-stripe work spread over more rows, held in more registers (fewer
// resident waves where they pass 168). 17-22 are the windowed-profile
// schedules 9, 10, 11 (8, 7 and 6 with each column's plane read one quad
// ahead of its rows, not a column pair ahead), on 24 and 32 rows: the
// plane no longer holds a register per row, so 24-row stripes fit 4 waves
// per SIMD and 32-row ones 3. 23-28 are the diagonal-shift schedules 12,
// 13, 14 (9, 10, 11 with every value shifted by row and column, so neither
// gap chain adds its extension: 5 packed ops per cell in fp16, 6 in int16;
// semiglobal mode only, diagshift_bound) on 24 and 32 rows.
constexpr int kDp16Variants = 29;
constexpr int kDp16Default = 3;  // the int16 default; panels within kBoundF16 take kDp16F16
constexpr int kDp16F16 = 4;      // the first variant computing in fp16
constexpr int kDp16Sched[kDp16Variants] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 7, 8, 6, 7, 8, 6, 4, 5, 9, 9, 10, 10, 11, 11,
                                           12, 12, 13, 13, 14, 14};
constexpr int kDp16Rows[kDp16Variants] = {16, 16, 16, 16, 16, 16, 16, 16, 16, 24, 24, 24, 32, 32, 32,
                                          24, 24, 24, 32, 24, 32, 24, 32, 24, 32, 24, 32, 24, 32};
constexpr int dp16_schedule(int v) { return kDp16Sched[v]; }
constexpr int dp16_stripe(int v) { return kDp16Rows[v]; }
// variants computing in fp16, and those with a shifted column (row-shifted
// in semiglobal mode, column-shifted in local mode)
constexpr bool dp16_is_f16(int v) {
    return dp16_schedule(v) == 4 || dp16_schedule(v) == 5 || dp16_schedule(v) == 7 || dp16_schedule(v) == 8 ||
           dp16_schedule(v) == 9 || dp16_schedule(v) == 10 || dp16_schedule(v) == 12 || dp16_schedule(v) == 13;
}
constexpr bool dp16_is_shifted(int v) { return dp16_schedule(v) >= 6; }
// the diagonal-shift schedules (semiglobal mode only)
constexpr bool dp16_is_diag(int v) { return dp16_schedule(v) >= 12; }
// The variant with schedule `schedule` on stripes of `stripe` rows (-1: none).
constexpr int dp16_variant(int schedule, int stripe) {
    for (int v = 0; v < kDp16Variants; ++v)
        if (kDp16Sched[v] == schedule && kDp16Rows[v] == stripe) return v;
    return -1;
}
hipError_t launch_prep(const KArgs &a, const Launch &g, hipStream_t st);
// k_revcomp alone: the reverse-complement rows of a both-strand batch
// written into its codes (launch_prep does it for the ordinary path).
hipError_t launch_revcomp(const KArgs &a, const Launch &g, hipStream_t st);
hipError_t launch_dp16(const KArgs &a, const Launch &g, bool local, hipStream_t st);
hipError_t launch_dp32(const KArgs &a, const Launch &g, bool local, int which, hipStream_t st);
hipError_t launch_select(const KArgs &a, const Launch &g, hipStream_t st);
hipError_t launch_diag(const KArgs &a, const Launch &g, bool local, hipStream_t st);
hipError_t diag_occupancy(bool local, size_t lds, int *blocks_per_cu);
hipError_t launch_peak16(uint32_t *out, int blocks, int iters, hipStream_t st);

// ---- winner traceback (trace.hip) ----
constexpr int kTraceRec = 7;  // score, qs, qe, rs, re, runs, first run's slot in the worst-case region

struct TraceArgs {
    DevPanel panel;
    Scoring sc;
    const uint8_t *qcodes;   // read codes, concatenated
    const uint64_t *qoff;    // per sorted pair: offset of its read
    const int32_t *qlen;     // per sorted pair: read length
    const int32_t *ref;      // per sorted pair: panel index
    const int32_t *slot;     // per sorted pair: the caller's pair index
    const uint64_t *gbase;   // per group: byte offset of its arena region
    const int32_t *gS, *gM;  // per group: stripes of its longest read, its longest reference
    uint8_t *arena;          // trace nibbles + stripe boundary, per group region
    int32_t *recs;           // [caller index * kTraceRec]
    uint32_t *ops_wc;        // worst-case run regions (read + reference length each)
    const uint64_t *ops_off; // per caller index: its region's offset
    int32_t n_pairs;
    int32_t g0;              // first group of this launch
    int32_t pair0;           // sorted index of the first pair of the groups (the anti-diagonal ones come first)
    // anti-diagonal traceback (k_trace_diag): sorted pairs 0 .. n_diag - 1,
    // one wave each, pair k's trace region at arena + dbase[k]
    const uint64_t *dbase;
    int32_t n_diag;
    int32_t d0;              // first of them in this launch
};
// k_trace_diag keeps a pass's bottom row in LDS: references up to this long
constexpr int kTraceDiagMaxRef = 4096;
constexpr int kTraceDiagPad = 192;  // LDS columns past the reference (block reads ahead)
// Rows per lane of k_trace_diag for a read of n bases: R in {2, 4, 8}
// minimising passes x R (the steps of the task), ties to the larger R.
__host__ __device__ inline int trace_diag_rows(int n) {
    int best = 8, cost = ((n + 511) / 512) * 8;
    for (int r = 4; r >= 2; r /= 2) {
        const int c = ((n + 64 * r - 1) / (64 * r)) * r;
        if (c < cost) {
            cost = c;
            best = r;
        }
    }
    return best;
}

hipError_t launch_trace(const TraceArgs &a, int groups, bool local, hipStream_t st);
hipError_t launch_trace_diag(const TraceArgs &a, int pairs, int max_ref, bool local, hipStream_t st);
hipError_t launch_trace_compact(const uint32_t *ops_wc, const uint64_t *ops_off, const int32_t *recs,
                                const uint64_t *out_off, uint32_t *out, int32_t n_pairs, hipStream_t st);
hipError_t launch_pairs(const KArgs &a, const Launch &g, hipStream_t st);
// Selection over a batch's resident int32 rows for the mapping tier loop
// (resolve.hip). Rows: read x as sequenced is row x, its reverse complement
// row n_reads + x; pair p's mates are reads 2p and 2p + 1. One wave per
// listed entry. A list entry's candidates go to cands + cand_off[e] when
// cand_off is set (room for all of them), else to cands + e * cap (the first
// cap of them; the record still counts them all).
// With half set (forward template strand only), pair p holds two rows:
// 2p = mate 1 as sequenced, 2p + 1 = mate 2 reverse-complemented.
struct TierArgs {
    const int32_t *out32;
    int32_t n_reads, n_refs;
    int32_t half;
    int32_t strands;        // template strands: 1 forward, 2 both (not with half)
    int32_t min_score;
    const int32_t *list;    // pair indices
    const int32_t *below;   // per entry: only sums below this (INT32_MAX: all)
    int32_t n_list;
    int32_t cap;
    const int64_t *cand_off;
    int32_t *recs;          // per entry: n_tier, sum, found (kTierRec)
    int32_t *cands;         // (ref << 1) | template strand, (ref, strand) order
};
constexpr int kTierRec = 3;

struct SuboptArgs {
    const int32_t *out32;
    int32_t n_reads, n_refs, half, strands;
    const int32_t *list;     // pair indices
    const int64_t *set_off;  // per entry: its tie set in `set`
    const int32_t *set_n;
    const int32_t *set;      // (ref << 1) | template strand, ascending
    int32_t n_list;
    int32_t *out;            // per entry: suboptimal of mate 1, of mate 2, flags (kSuboptRec)
};
constexpr int kSuboptRec = 3;

struct MateSelArgs {
    const int32_t *out32;
    int32_t n_reads, n_refs;
    int32_t both;            // reverse-complement rows present and candidates
    int32_t min_score;
    int64_t max_ties;        // < 0: all
    const int32_t *list;     // read (row) indices
    int32_t n_list;
    int32_t cap;
    const int64_t *cand_off;
    int32_t *recs;           // per entry: best, n_tied, subopt, flags, n_take (kSelRec)
    int32_t *ties;           // (ref << 1) | reverse, tie order
};
constexpr int kSelRec = 5;

hipError_t launch_pair_tier(const TierArgs &a, int max_blocks, hipStream_t st);
hipError_t launch_pair_subopt(const SuboptArgs &a, int max_blocks, hipStream_t st);
hipError_t launch_mate_select(const MateSelArgs &a, int max_blocks, hipStream_t st);
hipError_t dp16_attributes(bool local, int variant, hipFuncAttributes *attr);
hipError_t dp32_attributes(bool local, hipFuncAttributes *attr);
hipError_t dp16_occupancy(bool local, int variant, int *blocks_per_cu);
hipError_t dp32_occupancy(bool local, int block, int *blocks_per_cu);

}  // namespace esch
