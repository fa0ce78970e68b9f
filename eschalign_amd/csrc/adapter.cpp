// 3' adapter read-through detection, the pre-step that defines a read
// pair's queries (lib/adapter.rs:68-110, called by the CLI's GPU thread at
// crates/eschalign-cli/src/align.rs:2844-2862 before scoring). Host code:
// one pass over a read's 3' end, far below the cost of scoring it.
#include "../../include/eschalign_amd.h"

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <thread>
#include <vector>

namespace {

inline uint8_t code_of(uint8_t b) {
    switch (b) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': case 'U': case 'u': return 3;
    default: return 4;
    }
}

// adapter::find on codes (read and adapter already encoded).
size_t find_codes(const uint8_t *read, size_t n, const uint8_t *adapter, size_t alen, size_t min_overlap) {
    if (min_overlap < 1) min_overlap = 1;
    if (n < min_overlap || alen == 0) return 0;
    for (size_t p = 0; p + min_overlap <= n; ++p) {
        const size_t overlap = alen < n - p ? alen : n - p;
        if (overlap < min_overlap) break;
        const size_t allowed = overlap / 10;
        size_t mism = 0;
        for (size_t k = 0; k < overlap && mism <= allowed; ++k) {
            const uint8_t a = read[p + k], b = adapter[k];
            if (a != b && a != 4 && b != 4) ++mism;
        }
        if (mism <= allowed) return n - p;
    }
    return 0;
}

}  // namespace

extern "C" {

// adapter::find (lib/adapter.rs:68-91): the leftmost position p whose
// stretch read[p..] matches the adapter's prefix over
// min(adapter length, read length - p) >= min_overlap bases with at most
// overlap / 10 mismatches (wildcards match); returns read length - p, or 0.
size_t esch_adapter_find(const uint8_t *read, size_t n, const uint8_t *adapter, size_t alen,
                         size_t min_overlap) {
    if (min_overlap < 1) min_overlap = 1;
    if (!read || !adapter || n < min_overlap || alen == 0) return 0;
    for (size_t p = 0; p + min_overlap <= n; ++p) {
        const size_t overlap = alen < n - p ? alen : n - p;
        if (overlap < min_overlap) break;
        const size_t allowed = overlap / 10;
        size_t mism = 0;
        for (size_t k = 0; k < overlap && mism <= allowed; ++k) {
            const uint8_t a = code_of(read[p + k]), b = code_of(adapter[k]);
            if (a != b && a != 4 && b != 4) ++mism;
        }
        if (mism <= allowed) return n - p;
    }
    return 0;
}

// adapter::detect_pair (lib/adapter.rs:95-110): both mates' adapters, and
// when both are found but disagree on where the insert ends, the weaker
// (shorter overlap) detection is dropped, both on a tie.
void esch_adapter_detect_pair(const uint8_t *r1, size_t l1, const uint8_t *r2, size_t l2,
                              const uint8_t *a1, size_t la1, const uint8_t *a2, size_t la2,
                              size_t min_overlap, size_t out[2]) {
    size_t d0 = esch_adapter_find(r1, l1, a1, la1, min_overlap);
    size_t d1 = esch_adapter_find(r2, l2, a2, la2, min_overlap);
    if (d0 > 0 && d1 > 0 && l1 - d0 != l2 - d1) {
        const size_t e1 = d0 < la1 ? d0 : la1, e2 = d1 < la2 ? d1 : la2;
        if (e1 > e2) d1 = 0;
        else if (e1 < e2) d0 = 0;
        else d0 = d1 = 0;
    }
    out[0] = d0;
    out[1] = d1;
}

// detect_pair over a batch of pairs held as codes, mates interleaved
// (mate m of pair p at codes[offs[2p + m] ..], lens[2p + m] bases): out[k]
// = lens[k] less mate k's adapter, the query the CLI aligns
// (crates/eschalign-cli/src/align.rs:1642-1662). Spread over up to 8 host
// threads.
void esch_adapter_trim_pairs(const uint8_t *codes, const uint64_t *offs, const int32_t *lens, size_t n_pairs,
                             const uint8_t *a1, size_t la1, const uint8_t *a2, size_t la2, size_t min_overlap,
                             int32_t *out) {
    if (!codes || !offs || !lens || !out) return;
    std::vector<uint8_t> c1(la1), c2(la2);
    for (size_t k = 0; k < la1; ++k) c1[k] = code_of(a1[k]);
    for (size_t k = 0; k < la2; ++k) c2[k] = code_of(a2[k]);
    auto work = [&](size_t p0, size_t p1) {
        for (size_t p = p0; p < p1; ++p) {
            const size_t l1 = (size_t)std::max(lens[2 * p], 0), l2 = (size_t)std::max(lens[2 * p + 1], 0);
            size_t d0 = find_codes(codes + offs[2 * p], l1, c1.data(), la1, min_overlap);
            size_t d1 = find_codes(codes + offs[2 * p + 1], l2, c2.data(), la2, min_overlap);
            // the reconciliation of lib/adapter.rs:101-108 (esch_adapter_detect_pair)
            if (d0 > 0 && d1 > 0 && l1 - d0 != l2 - d1) {
                const size_t e1 = d0 < la1 ? d0 : la1, e2 = d1 < la2 ? d1 : la2;
                if (e1 > e2) d1 = 0;
                else if (e1 < e2) d0 = 0;
                else d0 = d1 = 0;
            }
            out[2 * p] = (int32_t)(l1 - d0);
            out[2 * p + 1] = (int32_t)(l2 - d1);
        }
    };
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = n_pairs < 4096 ? 1 : std::min<size_t>(8, hw);
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nt; ++t) pool.emplace_back(work, n_pairs * t / nt, n_pairs * (t + 1) / nt);
    work(0, n_pairs / nt);
    for (auto &th : pool) th.join();
}

}  // extern "C"
