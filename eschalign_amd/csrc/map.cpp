// Read and pair mapping on the device: Aligner::map_reads_scored
// (lib/mapper.rs:400-550) and Aligner::map_pairs_scored with its tier loop
// and Separate fallback (lib/mapper/pair.rs:157-387).
//
// A chunk of reads (or pairs, mates interleaved) is loaded into a device
// batch with both orientations and scored into resident int32 rows; every
// selection then runs on those rows (resolve.hip) and every traceback reads
// its read straight from the batch's codes (k_trace), so the host moves only
// lists of indices, tie lists and the traced records.
//
// Pairs. Round r lists every pair still open with the sum of the tier that
// failed in round r - 1 as its bound (none in round 1); k_pair_tier returns
// each listed pair's next tier, whose candidates are traced, both mates
// each; a pair whose tier holds a candidate passing pair_ok
// (lib/mapper/pair.rs:392-410) is closed with that tier's passing
// candidates as its tie set, a pair whose tier fails stays open, and a pair
// with no tier left is Separate. Every round is batched over all open pairs,
// as the reference's loop is (lib/mapper/pair.rs:236-286). Then
// k_pair_subopt gives each proper pair's per-mate suboptimal, and the
// Separate pairs' mates are mapped like unpaired reads on both strands.
#include "host_internal.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

using namespace esch;

namespace {

// Candidates per pair (ties per read) returned inline per selection round;
// a wider tier or tie list takes a second pass with room for all of it.
// Most tiers hold one candidate, so a small inline width keeps the copy
// back to the host small.
constexpr int kTierCap = 4;
constexpr int kSelCap = 4;
constexpr int kResolveBlocks = 4096;
constexpr size_t kChunkRowBytes = (size_t)4 << 30;  // resident int32 rows per chunk

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

bool has_match(const uint32_t *ops, int32_t n) {
    for (int32_t k = 0; k < n; ++k)
        if ((ops[k] >> 30) == 0u) return true;
    return false;
}

// Traced (read, reference) jobs: records and their runs.
struct Traced {
    std::vector<esch_alignment> al;
    std::vector<uint32_t> ops;
    bool aligned(size_t k) const { return has_match(ops.data() + al[k].ops_offset, al[k].n_ops); }
};

// Hits of one read or pair, collected out of order and placed in input
// order at the end: hits[k].alignment.ops_offset indexes ops.
struct Pool {
    std::vector<esch_hit> hits;
    std::vector<uint32_t> ops;
    void add(const Traced &t, size_t k, int32_t ref, int32_t rev, int32_t trev, int32_t insert) {
        esch_hit h{ref, rev, trev, insert, t.al[k]};
        h.alignment.ops_offset = (int64_t)ops.size();
        const uint32_t *src = t.ops.data() + t.al[k].ops_offset;
        ops.insert(ops.end(), src, src + t.al[k].n_ops);
        hits.push_back(h);
    }
};

}  // namespace

// Buffers kept by the scorer between mapping calls; one call at a time per
// scorer (the mutex), as the traceback cache.
struct MapCache {
    std::mutex mu;
    DevBuf list, below, recs, cands, cand_off, set, set_off, set_n, out;
    std::vector<uint8_t> rc2;   // mate 2's reverse complements of a chunk (host)
    std::vector<size_t> rc2_at;
};

void map_cache_free(esch_scorer *s) {
    if (s->map) {
        (void)hipSetDevice(s->ordinal);
        delete s->map;
        s->map = nullptr;
    }
}

struct esch_mapped {
    std::vector<esch_read_mapping> reads;
    std::vector<esch_pair_mapping> pairs;
    std::vector<esch_hit> hits;
    std::vector<uint32_t> ops;
    bool is_pairs = false;
    esch_map_stats stats{};
};

namespace {

template <class T>
esch_status upload(DevBuf &d, const std::vector<T> &h, hipStream_t st) {
    HIPC(d.need(h.size() * sizeof(T)));
    if (!h.empty()) HIPC(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
    return ESCH_OK;
}

template <class T>
esch_status download(std::vector<T> &h, const DevBuf &d, size_t n, hipStream_t st) {
    h.resize(n);
    if (n) HIPC(hipMemcpyAsync(h.data(), d.p, n * sizeof(T), hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    return ESCH_OK;
}

// One mapping call's context: the scorer, the batch holding the chunk and
// the cache of small device buffers.
struct Ctx {
    esch_scorer *s;
    esch_batch *b;
    MapCache &C;
    esch_map_stats &stats;
    // the next chunk's sliced scoring (or null): each wait for the device
    // below first queues its next slice behind the work just queued, so
    // that work runs between two slices and the device stays busy while
    // the host waits and then uses the results
    esch_batch *next = nullptr;
    esch_status pump(hipStream_t st) const { return batch_pump(next, st); }
    // row x of the batch: its codes on the device and its length
    const uint8_t *row_ptr(int32_t x) const { return b->d_codes + b->h_qoff[x]; }
    size_t row_len(int32_t x) const { return (size_t)b->h_qlen[x]; }

    // Trace (row, ref) jobs from the batch's resident codes.
    esch_status trace(const std::vector<int32_t> &rows, const std::vector<int32_t> &refs, Traced &t) {
        t.al.assign(rows.size(), esch_alignment{});
        t.ops.clear();
        if (rows.empty()) return ESCH_OK;
        const auto tw = Clock::now();
        esch_status st = align_core(
            s, rows.size(),
            [&](size_t k, const uint8_t *&p, size_t &len) {
                p = row_ptr(rows[k]);
                len = row_len(rows[k]);
            },
            refs.data(), t.al.data(), nullptr, 0, nullptr, b->d_codes, &t.ops, next);
        stats.trace_ms += s->trace_ms;
        stats.trace_wall_ms += ms_since(tw);
        stats.traced += (int64_t)rows.size();
        return st;
    }

    // The device half of mapping listed rows: the fused selection (ties past
    // the inline capacity again, with room for all) and the tracebacks of
    // every taken tie (lib/mapper.rs:418-496). `build` then makes the
    // records and hits on the host, so a caller can launch the next chunk's
    // scoring between the two.
    struct Selected {
        std::vector<int32_t> recs, ties, ties_over;
        std::vector<int64_t> tie_at;
        std::vector<size_t> first;
        Traced t;
        int32_t tie(size_t e, int32_t j) const {
            return tie_at[e] >= 0 ? ties_over[(size_t)tie_at[e] + (size_t)j] : ties[e * kSelCap + (size_t)j];
        }
    };
    esch_status select_trace(const std::vector<int32_t> &rows, int both, int32_t min_score, int64_t max_ties,
                             Selected &S) {
        std::vector<int32_t> jrows, jrefs;
        if (esch_status e = select(rows, both, min_score, max_ties, S, jrows, jrefs)) return e;
        return trace(jrows, jrefs, S.t);
    }

    // The selection half of select_trace: the records and tie lists, and the
    // (row, reference) job of every taken tie for `trace`.
    esch_status select(const std::vector<int32_t> &rows, int both, int32_t min_score, int64_t max_ties,
                       Selected &S, std::vector<int32_t> &jrows, std::vector<int32_t> &jrefs) {
        const size_t n = rows.size();
        if (!n) return ESCH_OK;
        hipStream_t st = b->stream;
        const auto t0 = Clock::now();
        if (esch_status e = upload(C.list, rows, st)) return e;
        HIPC(C.recs.need(n * kSelRec * 4));
        HIPC(C.cands.need(n * kSelCap * 4));
        MateSelArgs a{b->d_out32, (int32_t)b->n_reads, (int32_t)s->n_refs, both, min_score, max_ties,
                      C.list.as<int32_t>(), (int32_t)n, kSelCap, nullptr, C.recs.as<int32_t>(), C.cands.as<int32_t>()};
        HIPC(launch_mate_select(a, kResolveBlocks, st));
        if (esch_status e = pump(st)) return e;
        if (esch_status e = download(S.recs, C.recs, n * kSelRec, st)) return e;
        if (esch_status e = download(S.ties, C.cands, n * kSelCap, st)) return e;
        const std::vector<int32_t> &recs = S.recs;
        std::vector<int32_t> over, over_rows;
        std::vector<int64_t> over_off;
        int64_t total = 0;
        for (size_t e = 0; e < n; ++e)
            if (recs[e * kSelRec + 4] > kSelCap) {
                over.push_back((int32_t)e);
                over_rows.push_back(rows[e]);
                over_off.push_back(total);
                total += recs[e * kSelRec + 4];
            }
        if (!over.empty()) {
            if (esch_status e = upload(C.list, over_rows, st)) return e;
            if (esch_status e = upload(C.cand_off, over_off, st)) return e;
            HIPC(C.out.need((size_t)total * 4));
            HIPC(C.set.need(over.size() * kSelRec * 4));
            MateSelArgs o = a;
            o.list = C.list.as<int32_t>();
            o.n_list = (int32_t)over.size();
            o.cand_off = C.cand_off.as<int64_t>();
            o.recs = C.set.as<int32_t>();
            o.ties = C.out.as<int32_t>();
            HIPC(launch_mate_select(o, kResolveBlocks, st));
            if (esch_status e = pump(st)) return e;
            if (esch_status e = download(S.ties_over, C.out, (size_t)total, st)) return e;
        }
        stats.select_ms += ms_since(t0);
        S.tie_at.assign(n, -1);  // into ties_over for the overflowed
        for (size_t i = 0; i < over.size(); ++i) S.tie_at[over[i]] = over_off[i];
        // every taken tie traced (lib/mapper.rs:486-496)
        jrows.clear();
        jrefs.clear();
        S.first.assign(n, 0);
        for (size_t e = 0; e < n; ++e) {
            S.first[e] = jrows.size();
            for (int32_t j = 0; j < recs[e * kSelRec + 4]; ++j) {
                const int32_t c = S.tie(e, j);
                jrows.push_back((c & 1) ? (int32_t)b->n_reads + rows[e] : rows[e]);
                jrefs.push_back(c >> 1);
            }
        }
        return ESCH_OK;
    }

    // The host half: records and hits (lib/mapper.rs:498-550); no device use.
    static void build(size_t n, const Selected &S, std::vector<esch_read_mapping> &out, Pool &pool) {
        out.assign(n, esch_read_mapping{});
        for (size_t e = 0; e < n; ++e) {
            const int32_t *r = &S.recs[e * kSelRec];
            esch_read_mapping &m = out[e];
            m.best = r[0];
            m.n_tied = r[1];
            m.subopt = r[2];
            m.flags = r[3];
            m.hit_offset = (int64_t)pool.hits.size();
            const int32_t take = r[4];
            // unmapped: nothing taken, or the primary aligns no base
            // (lib/mapper.rs:508-512)
            if (take == 0 || !S.t.aligned(S.first[e])) continue;
            for (int32_t j = 0; j < take; ++j) {
                const int32_t c = S.tie(e, j);
                pool.add(S.t, S.first[e] + (size_t)j, c >> 1, c & 1, 0, 0);
            }
            m.n_hits = take;
        }
    }

    // The unpaired selection and tracebacks of listed rows (lib/mapper.rs:
    // 418-550): out[e] for rows[e]; hits[e] into pool with hit_offset = the
    // pool index of the read's first hit.
    esch_status map_rows(const std::vector<int32_t> &rows, int both, int32_t min_score, int64_t max_ties,
                         std::vector<esch_read_mapping> &out, Pool &pool) {
        Selected S;
        if (esch_status e = select_trace(rows, both, min_score, max_ties, S)) return e;
        build(rows.size(), S, out, pool);
        return ESCH_OK;
    }
};

// Copy a read mapping's hits from the pool to the result, in order.
void place_read(esch_mapped *M, const Pool &pool, esch_read_mapping &m) {
    const int64_t h0 = m.hit_offset;
    m.hit_offset = (int64_t)M->hits.size();
    for (int32_t k = 0; k < m.n_hits; ++k) {
        esch_hit h = pool.hits[(size_t)(h0 + k)];
        const uint32_t *src = pool.ops.data() + h.alignment.ops_offset;
        h.alignment.ops_offset = (int64_t)M->ops.size();
        M->ops.insert(M->ops.end(), src, src + h.alignment.n_ops);
        M->hits.push_back(h);
    }
}

using Getter = std::function<void(size_t, const uint8_t *&, size_t &)>;

// Room for `extra` more elements, growing geometrically: an exact reserve
// per chunk would copy everything gathered so far at every chunk.
template <class V>
void reserve_more(V &v, size_t extra) {
    const size_t need = v.size() + extra;
    if (need > v.capacity()) v.reserve(std::max(need, 2 * v.capacity()));
}

size_t chunk_rows(const esch_scorer *s) {
    // resident int32 rows of a chunk: each read (mate) and its reverse complement
    if (s->map_chunk) return 2 * s->map_chunk;
    return std::max<size_t>(2048, kChunkRowBytes / (4 * std::max<size_t>(s->n_refs, 1)));
}

esch_status scorer_ready(esch_scorer *s) {
    HIPC(hipSetDevice(s->ordinal));
    static std::mutex create_mu;
    std::lock_guard<std::mutex> g(create_mu);
    if (!s->map) s->map = new MapCache;
    return ESCH_OK;
}

// Load reads [k0, k1) of get into a batch (both orientations when both):
// the gather into pinned staging and the upload, queued on the batch's
// stream. The scoring is launched by score_chunk.
esch_status load_chunk(esch_scorer *s, const Getter &get, size_t k0, size_t k1, int both, int pairs,
                       esch_batch **out, esch_map_stats &stats) {
    const auto t0 = Clock::now();
    const size_t n = k1 - k0;
    std::vector<const uint8_t *> ptrs(n);
    std::vector<size_t> lens(n);
    size_t bases = 0;
    for (size_t k = 0; k < n; ++k) {
        get(k0 + k, ptrs[k], lens[k]);
        bases += lens[k];
    }
    esch_batch *b = nullptr;
    if (esch_status st = take_batch(s, n, bases, &b)) return st;
    if (esch_status st = esch_batch_load(b, ptrs.data(), lens.data(), n, both, pairs)) {
        give_batch(s, b);
        return st;
    }
    stats.load_ms += ms_since(t0);
    *out = b;
    return ESCH_OK;
}

// Launch the scoring of a loaded batch into resident int32 rows.
esch_status score_chunk(esch_batch *b) { return esch_batch_run(b, ESCH_RUN_SCORES32, 0, -1, 1); }

// Chunk boundaries of a mapping call of n reads (or pairs) in chunks of
// `chunk`: whole chunks and, with taper, a first chunk of a quarter (its
// staging comes before any scoring) and the last chunk's worth cut in
// halves down to a quarter chunk (.., chunk / 2, chunk / 4, chunk / 4):
// the work after the last scoring (its selection, tracebacks and records)
// has no scoring left to run beside, so the last chunk is kept small.
// Measured on 1M reads: 439-442 ms in the library with a whole first
// chunk, 434-435 with a quarter (pairs: no change).
std::vector<size_t> chunk_cuts(size_t n, size_t chunk, bool taper) {
    std::vector<size_t> cuts{0};
    for (size_t at = 0; at < n;) {
        const size_t left = n - at;
        size_t take = std::min(left, chunk);
        if (taper && at == 0) take = std::max<size_t>(1, chunk / 4);  // the first scoring starts sooner
        else if (taper && left <= chunk && 4 * left > chunk) take = (left + 1) / 2;
        at += take;
        cuts.push_back(at);
    }
    return cuts;
}

// The same in slices (batch_run_sliced), for a chunk scored while the
// previous chunk's selection and tracebacks run: enough slices that those
// kernels wait at most a slice for the device, few enough that the slices'
// drains stay a small part of the scoring. Measured on 1M reads of the
// 164-reference panel and 2M pairs of 500 references (reads / pairs per
// s): 4 slices 2.281 M / 1.255 M, 6 2.300 M / 1.255 M, 8 2.293 M / 1.250 M,
// 12 2.270 M / 1.236 M.
constexpr int kMapSlices = 6;
esch_status score_chunk_sliced(esch_batch *b) { return batch_run_sliced(b, ESCH_RUN_SCORES32, kMapSlices); }

// Wait for a batch's scoring; count its device time and cells.
esch_status finish_chunk(esch_batch *b, esch_map_stats &stats) {
    if (esch_status st = esch_batch_wait(b)) return st;
    float total = 0.0f;
    uint64_t cells = 0;
    if (esch_status st = esch_batch_timing(b, nullptr, &total)) return st;
    if (esch_status st = esch_batch_stats(b, &cells, nullptr, nullptr)) return st;
    stats.score_ms += total;
    stats.cells += (int64_t)cells;
    return ESCH_OK;
}

// The three in a row: reads [k0, k1) of get scored into a batch's rows.
esch_status load_and_score(esch_scorer *s, const Getter &get, size_t k0, size_t k1, int both, int pairs,
                           esch_batch **out, esch_map_stats &stats) {
    esch_batch *b = nullptr;
    if (esch_status st = load_chunk(s, get, k0, k1, both, pairs, &b, stats)) return st;
    esch_status st = score_chunk(b);
    if (!st) st = finish_chunk(b, stats);
    if (st) {
        give_batch(s, b);
        return st;
    }
    *out = b;
    return ESCH_OK;
}

esch_status map_reads_core(esch_scorer *s, size_t n, const Getter &get, int both, int32_t min_score,
                           int64_t max_ties, esch_mapped **out) {
    const auto t0 = Clock::now();
    std::unique_ptr<esch_mapped> M(new esch_mapped);
    M->reads.resize(n);
    if (esch_status st = scorer_ready(s)) return st;
    std::lock_guard<std::mutex> call(s->map->mu);
    // reads per chunk: what the rows' memory allows, and by default about a
    // quarter of the call (at least 2^17 reads). Chunk ci, once scored:
    //  - the next chunk's scoring launched in slices (score_chunk_sliced);
    //  - its selection (resolve.hip) on its resident rows, between two of
    //    those slices;
    //  - the chunk after that staged (gathered and uploaded);
    //  - its tracebacks, between two more slices;
    //  - its records built on a worker thread, beside the next chunk.
    // So the device idles only between the last chunk's scoring and its
    // records. Measured on 1M reads of the 164-reference panel with the
    // tracebacks between the scorings instead: 1.64 M reads/s in one chunk,
    // 1.75 M in four, 1.67 M in sixteen.
    const size_t cap = std::max<size_t>(1, chunk_rows(s) / 2);
    const size_t chunk = s->map_chunk ? cap : std::min(cap, std::max<size_t>((size_t)1 << 17, (n + 3) / 4));
    const std::vector<size_t> cuts = chunk_cuts(n, chunk, !s->map_chunk && n > chunk);
    const size_t n_chunks = cuts.size() - 1;
    // chunks ci, ci + 1 and ci + 2, and the records' worker; joined and
    // handed back on every exit (before M goes)
    struct Held {
        esch_scorer *s;
        esch_batch *b[3] = {nullptr, nullptr, nullptr};
        std::thread builder;
        double build_ms = 0.0;  // the last builder's time, read after its join
        void join(esch_map_stats &stats) {
            if (!builder.joinable()) return;
            builder.join();
            stats.assemble_ms += build_ms;
        }
        ~Held() {
            if (builder.joinable()) builder.join();
            for (esch_batch *x : b)
                if (x) give_batch(s, x);
        }
    } held{s};
    auto range = [&](size_t ci) { return std::make_pair(cuts[ci], cuts[ci + 1]); };
    if (n_chunks) {
        if (esch_status st = load_chunk(s, get, 0, range(0).second, both, 0, &held.b[0], M->stats)) return st;
        if (esch_status st = score_chunk(held.b[0])) return st;
    }
    if (n_chunks > 1)
        if (esch_status st = load_chunk(s, get, range(1).first, range(1).second, both, 0, &held.b[1], M->stats))
            return st;
    for (size_t ci = 0; ci < n_chunks; ++ci) {
        const size_t k0 = range(ci).first, k1 = range(ci).second;
        esch_batch *b = held.b[0];
        if (esch_status st = finish_chunk(b, M->stats)) return st;
        if (held.b[1])
            if (esch_status st = score_chunk_sliced(held.b[1])) return st;
        Ctx cx{s, b, *s->map, M->stats, held.b[1]};
        std::vector<int32_t> rows;  // non-empty reads; empty ones have best None
        for (size_t k = k0; k < k1; ++k)
            if (b->h_qlen[k - k0]) rows.push_back((int32_t)(k - k0));
        std::unique_ptr<Ctx::Selected> S(new Ctx::Selected);
        std::vector<int32_t> jrows, jrefs;
        if (esch_status st = cx.select(rows, both, min_score, max_ties, *S, jrows, jrefs)) return st;
        if (ci + 2 < n_chunks)
            if (esch_status st = load_chunk(s, get, range(ci + 2).first, range(ci + 2).second, both, 0, &held.b[2],
                                            M->stats))
                return st;
        if (esch_status st = cx.trace(jrows, jrefs, S->t)) return st;
        // The chunk's records on the worker. Its hits go straight after the
        // earlier chunks' (the result's arrays serve as the pool: offsets
        // come out absolute), and a read without hits points where the next
        // read's hits start; one builder at a time keeps that order.
        held.join(M->stats);
        esch_mapped *m = M.get();
        held.builder = std::thread([m, k0, k1, rows = std::move(rows), S = std::move(S), &held]() {
            const auto ta = Clock::now();
            std::vector<esch_read_mapping> maps;
            Pool pool;
            pool.hits.swap(m->hits);
            pool.ops.swap(m->ops);
            Ctx::build(rows.size(), *S, maps, pool);
            m->hits.swap(pool.hits);
            m->ops.swap(pool.ops);
            for (size_t k = k0; k < k1; ++k) m->reads[k] = esch_read_mapping{};
            for (size_t e = 0; e < rows.size(); ++e) m->reads[k0 + (size_t)rows[e]] = maps[e];
            int64_t next = (int64_t)m->hits.size();  // reads without hits, from the back
            for (size_t k = k1; k-- > k0;) {
                esch_read_mapping &r = m->reads[k];
                if (r.n_hits) next = r.hit_offset;
                else r.hit_offset = next;
            }
            held.build_ms = ms_since(ta);
        });
        give_batch(s, held.b[0]);
        held.b[0] = held.b[1];
        held.b[1] = held.b[2];
        held.b[2] = nullptr;
    }
    held.join(M->stats);
    M->stats.total_ms = ms_since(t0);
    *out = M.release();
    return ESCH_OK;
}

// pair_ok (lib/mapper/pair.rs:392-410) on the traced mates of a candidate.
bool pair_ok(const esch_alignment &a0, bool al0, const esch_alignment &a1, bool al1, int t, int64_t max_insert) {
    if (!al0 || !al1) return false;
    // the mate read as sequenced on the template strand is the forward one
    const esch_alignment &fw = t ? a1 : a0, &rv = t ? a0 : a1;
    if (fw.ref_start >= rv.ref_end) return false;
    const int32_t lo = std::min(a0.ref_start, a1.ref_start), hi = std::max(a0.ref_end, a1.ref_end);
    return (int64_t)(hi - lo) <= max_insert;
}

// Reverse complements of mate 2 of pairs [p0, p1) into one host buffer
// (lib/alphabet.rs:58-64), over up to 8 threads; false when a mate carries a
// code outside 0..=4 (*bad = its pair).
bool rc_mates2(const Getter &mate, size_t p0, size_t p1, std::vector<uint8_t> &buf, std::vector<size_t> &at,
               size_t *bad) {
    const size_t c = p1 - p0;
    at.assign(c + 1, 0);
    for (size_t i = 0; i < c; ++i) {
        const uint8_t *p;
        size_t len;
        mate(2 * (p0 + i) + 1, p, len);
        at[i + 1] = at[i] + len;
    }
    if (buf.size() < at[c] + 1) buf.resize(at[c] + 1);
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = at[c] < ((size_t)8 << 20) ? 1 : std::min<size_t>(8, hw);
    std::vector<size_t> first_bad(nt, SIZE_MAX);
    auto work = [&](size_t t) {
        for (size_t i = c * t / nt; i < c * (t + 1) / nt; ++i) {
            const uint8_t *p;
            size_t len;
            mate(2 * (p0 + i) + 1, p, len);
            uint8_t *d = buf.data() + at[i];
            uint8_t mx = 0;
            for (size_t j = 0; j < len; ++j) {
                const uint8_t x = p[len - 1 - j];
                mx = x > mx ? x : mx;
                d[j] = x < 4 ? (uint8_t)(3 - x) : (uint8_t)4;
            }
            if (mx > 4 && first_bad[t] == SIZE_MAX) first_bad[t] = p0 + i;
        }
    };
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nt; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
    for (size_t t = 0; t < nt; ++t)
        if (first_bad[t] != SIZE_MAX) {
            *bad = first_bad[t];
            return false;
        }
    return true;
}

esch_status map_pairs_core(esch_scorer *s, size_t n_pairs, const Getter &mate, int both_strands, int32_t min_score,
                           int64_t max_ties, int64_t max_insert, esch_mapped **out) {
    const auto t0 = Clock::now();
    std::unique_ptr<esch_mapped> M(new esch_mapped);
    M->is_pairs = true;
    M->pairs.resize(n_pairs);
    if (esch_status st = scorer_ready(s)) return st;
    std::lock_guard<std::mutex> call(s->map->mu);
    const int strands = both_strands ? 2 : 1;
    // Forward template strand only: the tiers need mate 1 as sequenced and
    // mate 2 reverse-complemented, nothing else, so only those two rows per
    // pair are scored (half); the other orientations are scored later for
    // the mates of Separate pairs alone (the reference scores rows on
    // demand too, lib/mapper/pair.rs:195-207). Both strands: all four.
    const bool half = !both_strands;
    // pairs per chunk: what the rows' memory allows, and by default about a
    // sixteenth of the call (at least 2^16 pairs), so that staging, the tier
    // loop and the assembly of one chunk are a small part of the whole
    // (those of the first and last chunk are not hidden behind scoring).
    // Measured on 2M pairs of 500 references: 897 k pairs/s in two chunks,
    // 944 k in eight, 959 k in sixteen.
    const size_t cap = std::max<size_t>(1, chunk_rows(s) / (half ? 2 : 4));
    const size_t chunk = s->map_chunk ? cap : std::min(cap, std::max<size_t>((size_t)1 << 16, (n_pairs + 15) / 16));
    const std::vector<size_t> cuts = chunk_cuts(n_pairs, chunk, !s->map_chunk && n_pairs > chunk);
    const size_t n_chunks = cuts.size() - 1;
    std::vector<uint8_t> &rc2 = s->map->rc2;  // grow-only across calls
    std::vector<size_t> &rc2_at = s->map->rc2_at;
    // Stage chunk q into a batch: mate 2's reverse complements (half rows),
    // the gather into pinned staging and the upload. Its scoring is launched
    // separately, so the next chunk is staged while this one scores; once a
    // chunk is scored, the next one scores in slices, the chunk's tier loop
    // runs between them and its records are assembled beside the last ones.
    auto stage = [&](size_t q, esch_batch **out) -> esch_status {
        const size_t p0 = cuts[q], p1 = cuts[q + 1], c = p1 - p0;
        if (!half) return load_chunk(s, mate, 2 * p0, 2 * p1, 1, 1, out, M->stats);
        const auto t = Clock::now();
        size_t bad = 0;
        if (!rc_mates2(mate, p0, p1, rc2, rc2_at, &bad))
            return fail(ESCH_INVALID_ARG, "pair %zu: mate 2 carries a code outside 0..=4", bad);
        M->stats.load_ms += ms_since(t);
        const Getter rows = [&, p0](size_t k, const uint8_t *&p, size_t &len) {
            if (k & 1) {
                p = rc2.data() + rc2_at[k >> 1];
                len = rc2_at[(k >> 1) + 1] - rc2_at[k >> 1];
            } else {
                mate(2 * p0 + k, p, len);
            }
        };
        return load_chunk(s, rows, 0, 2 * c, 0, 0, out, M->stats);
    };
    // cur: the chunk being resolved (scored or scoring); next: staged
    struct Held {
        esch_scorer *s;
        esch_batch *cur = nullptr, *next = nullptr;
        ~Held() {
            if (cur) give_batch(s, cur);
            if (next) give_batch(s, next);
        }
    } held{s};
    if (n_chunks) {
        if (esch_status st = stage(0, &held.cur)) return st;
        if (esch_status st = score_chunk(held.cur)) return st;
    }
    for (size_t ci = 0; ci < n_chunks; ++ci) {
        const size_t p0 = cuts[ci], p1 = cuts[ci + 1], c = p1 - p0;
        if (ci + 1 < n_chunks)
            if (esch_status st = stage(ci + 1, &held.next)) return st;
        esch_batch *b = held.cur;
        if (esch_status st = finish_chunk(b, M->stats)) return st;
        // the next chunk scores in slices, this chunk's tier loop between them
        if (held.next)
            if (esch_status e = score_chunk_sliced(held.next)) return e;
        struct Guard {
            esch_scorer *s;
            esch_batch *b;
            ~Guard() {
                if (b) give_batch(s, b);
            }
        } guard_sep{s, nullptr};
        Ctx cx{s, b, *s->map, M->stats, held.next};
        MapCache &C = *s->map;
        hipStream_t st = b->stream;
        const int32_t nrd = (int32_t)b->n_reads;  // 2c
        // the batch rows of pair i's mates on template strand t
        auto row1 = [&](int32_t i, int32_t t) { return half ? 2 * i : (t ? nrd + 2 * i : 2 * i); };
        auto row2 = [&](int32_t i, int32_t t) { return half ? 2 * i + 1 : (t ? 2 * i + 1 : nrd + 2 * i + 1); };
        std::vector<esch_pair_mapping> res(c);
        Pool pool;
        std::vector<int32_t> set_codes;           // tie sets of the proper pairs
        std::vector<int64_t> set_off(c, -1);
        std::vector<int32_t> set_n(c, 0);
        std::vector<int32_t> list, below, sep;
        for (size_t i = 0; i < c; ++i) {
            if (b->h_qlen[2 * i] && b->h_qlen[2 * i + 1]) {
                list.push_back((int32_t)i);
                below.push_back(INT32_MAX);
            } else {
                sep.push_back((int32_t)i);  // an empty mate has no candidates
            }
        }
        while (!list.empty()) {
            ++M->stats.rounds;
            const size_t n = list.size();
            const auto ts = Clock::now();
            if (esch_status e = upload(C.list, list, st)) return e;
            if (esch_status e = upload(C.below, below, st)) return e;
            HIPC(C.recs.need(n * kTierRec * 4));
            HIPC(C.cands.need(n * kTierCap * 4));
            TierArgs a{b->d_out32, nrd, (int32_t)s->n_refs, half ? 1 : 0, strands, min_score, C.list.as<int32_t>(),
                       C.below.as<int32_t>(), (int32_t)n, kTierCap, nullptr, C.recs.as<int32_t>(),
                       C.cands.as<int32_t>()};
            HIPC(launch_pair_tier(a, kResolveBlocks, st));
            std::vector<int32_t> recs, cands;
            if (esch_status e = cx.pump(st)) return e;
            if (esch_status e = download(recs, C.recs, n * kTierRec, st)) return e;
            if (esch_status e = download(cands, C.cands, n * kTierCap, st)) return e;
            // tiers wider than the inline capacity: those pairs again
            std::vector<int32_t> ol, ob;
            std::vector<int64_t> oo, at(n, -1);
            int64_t total = 0;
            for (size_t e = 0; e < n; ++e)
                if (recs[e * kTierRec] > kTierCap) {
                    at[e] = total;
                    ol.push_back(list[e]);
                    ob.push_back(below[e]);
                    oo.push_back(total);
                    total += recs[e * kTierRec];
                }
            std::vector<int32_t> wide;
            if (!ol.empty()) {
                if (esch_status e = upload(C.list, ol, st)) return e;
                if (esch_status e = upload(C.below, ob, st)) return e;
                if (esch_status e = upload(C.cand_off, oo, st)) return e;
                HIPC(C.out.need((size_t)total * 4));
                HIPC(C.set.need(ol.size() * kTierRec * 4));
                TierArgs w = a;
                w.list = C.list.as<int32_t>();
                w.below = C.below.as<int32_t>();
                w.n_list = (int32_t)ol.size();
                w.cand_off = C.cand_off.as<int64_t>();
                w.recs = C.set.as<int32_t>();
                w.cands = C.out.as<int32_t>();
                HIPC(launch_pair_tier(w, kResolveBlocks, st));
                if (esch_status e = cx.pump(st)) return e;
                if (esch_status e = download(wide, C.out, (size_t)total, st)) return e;
            }
            M->stats.select_ms += ms_since(ts);
            auto cand = [&](size_t e, int32_t j) -> int32_t {
                return at[e] >= 0 ? wide[(size_t)at[e] + (size_t)j] : cands[e * kTierCap + (size_t)j];
            };
            // both mates of every candidate of every listed tier
            std::vector<int32_t> jrows, jrefs;
            std::vector<size_t> first(n, 0);
            jrows.reserve(2 * n);
            jrefs.reserve(2 * n);
            for (size_t e = 0; e < n; ++e) {
                first[e] = jrows.size();
                const int32_t i = list[e];
                for (int32_t j = 0; j < recs[e * kTierRec]; ++j) {
                    const int32_t cd = cand(e, j), t = cd & 1;
                    jrows.push_back(row1(i, t));
                    jrows.push_back(row2(i, t));
                    jrefs.push_back(cd >> 1);
                    jrefs.push_back(cd >> 1);
                }
            }
            Traced tr;
            if (esch_status e = cx.trace(jrows, jrefs, tr)) return e;
            std::vector<int32_t> nl, nb, ok;
            reserve_more(pool.hits, jrows.size());
            reserve_more(pool.ops, tr.ops.size());
            for (size_t e = 0; e < n; ++e) {
                const int32_t i = list[e];
                const int32_t nt = recs[e * kTierRec];
                if (!recs[e * kTierRec + 2]) {
                    sep.push_back(i);  // no tier left
                    continue;
                }
                ok.clear();
                for (int32_t j = 0; j < nt; ++j) {
                    const size_t k = first[e] + 2 * (size_t)j;
                    if (pair_ok(tr.al[k], tr.aligned(k), tr.al[k + 1], tr.aligned(k + 1), cand(e, j) & 1, max_insert))
                        ok.push_back(j);
                }
                if (ok.empty()) {
                    nl.push_back(i);
                    nb.push_back(recs[e * kTierRec + 1]);
                    continue;
                }
                // the tie set (lib/mapper/pair.rs:292-358): primary first, up
                // to max_ties + 1 kept
                esch_pair_mapping &m = res[(size_t)i];
                m.proper = 1;
                m.score = recs[e * kTierRec + 1];
                m.n_tied = (int32_t)ok.size();
                m.hit_offset = (int64_t)pool.hits.size();
                set_off[(size_t)i] = (int64_t)set_codes.size();
                set_n[(size_t)i] = (int32_t)ok.size();
                int64_t keep = (int64_t)ok.size();
                if (max_ties >= 0 && max_ties + 1 < keep) keep = max_ties + 1;
                for (size_t q = 0; q < ok.size(); ++q) {
                    const int32_t cd = cand(e, ok[q]);
                    set_codes.push_back(cd);
                    if ((int64_t)q >= keep) continue;
                    const size_t k = first[e] + 2 * (size_t)ok[q];
                    const esch_alignment &x = tr.al[k], &y = tr.al[k + 1];
                    const int32_t insert = std::max(x.ref_end, y.ref_end) - std::min(x.ref_start, y.ref_start);
                    const int32_t t = cd & 1;
                    pool.add(tr, k, cd >> 1, t, t, insert);       // mate 1: reversed iff t
                    pool.add(tr, k + 1, cd >> 1, !t, t, insert);  // mate 2: reversed iff !t
                }
                m.n_hits = (int32_t)keep;
            }
            list.swap(nl);
            below.swap(nb);
        }
        // per-mate suboptimal of the proper pairs
        std::vector<int32_t> pl, pn;
        std::vector<int64_t> po;
        for (size_t i = 0; i < c; ++i)
            if (res[i].proper) {
                pl.push_back((int32_t)i);
                po.push_back(set_off[i]);
                pn.push_back(set_n[i]);
            }
        if (!pl.empty()) {
            const auto ts = Clock::now();
            if (esch_status e = upload(C.list, pl, st)) return e;
            if (esch_status e = upload(C.cand_off, po, st)) return e;
            if (esch_status e = upload(C.set_n, pn, st)) return e;
            if (esch_status e = upload(C.set, set_codes, st)) return e;
            HIPC(C.out.need(pl.size() * kSuboptRec * 4));
            SuboptArgs a{b->d_out32, nrd, (int32_t)s->n_refs, half ? 1 : 0, strands, C.list.as<int32_t>(),
                         C.cand_off.as<int64_t>(), C.set_n.as<int32_t>(), C.set.as<int32_t>(), (int32_t)pl.size(),
                         C.out.as<int32_t>()};
            HIPC(launch_pair_subopt(a, kResolveBlocks, st));
            std::vector<int32_t> so;
            if (esch_status e = cx.pump(st)) return e;
            if (esch_status e = download(so, C.out, pl.size() * kSuboptRec, st)) return e;
            M->stats.select_ms += ms_since(ts);
            for (size_t q = 0; q < pl.size(); ++q) {
                esch_pair_mapping &m = res[(size_t)pl[q]];
                m.subopt[0] = so[q * kSuboptRec];
                m.subopt[1] = so[q * kSuboptRec + 1];
                m.flags = so[q * kSuboptRec + 2];
            }
        }
        // the Separate pairs: each non-empty mate as an unpaired read on both
        // strands (lib/mapper/pair.rs:360-387); with half rows, those mates
        // are scored again in both orientations in a batch of their own
        std::sort(sep.begin(), sep.end());
        std::vector<int32_t> mates_of;  // chunk-relative mate index 2i + m
        for (int32_t i : sep)
            for (int m = 0; m < 2; ++m)
                if (b->h_qlen[2 * i + m]) mates_of.push_back(2 * i + m);
        std::vector<esch_read_mapping> mm;
        if (half && !mates_of.empty()) {
            const Getter sm = [&](size_t k, const uint8_t *&p, size_t &len) { mate(2 * p0 + (size_t)mates_of[k], p, len); };
            esch_batch *bs = nullptr;
            if (esch_status e = load_and_score(s, sm, 0, mates_of.size(), 1, 0, &bs, M->stats)) return e;
            guard_sep.b = bs;
            Ctx cs{s, bs, *s->map, M->stats, held.next};
            std::vector<int32_t> rows(mates_of.size());
            for (size_t q = 0; q < rows.size(); ++q) rows[q] = (int32_t)q;
            if (esch_status e = cs.map_rows(rows, 1, min_score, max_ties, mm, pool)) return e;
        } else if (esch_status e = cx.map_rows(mates_of, 1, min_score, max_ties, mm, pool)) {
            return e;
        }
        for (size_t e = 0; e < mates_of.size(); ++e) {
            const int32_t x = mates_of[e];
            res[(size_t)(x / 2)].separate[x % 2] = mm[e];
        }
        // this chunk's device work is done: the next chunk's slices still
        // out run beside the assembly
        if (esch_status e = batch_drain(held.next)) return e;
        // into the result, in pair order
        const auto ta = Clock::now();
        reserve_more(M->hits, pool.hits.size());
        reserve_more(M->ops, pool.ops.size());
        for (size_t i = 0; i < c; ++i) {
            esch_pair_mapping &m = res[i];
            if (m.proper) {
                const int64_t h0 = m.hit_offset;
                m.hit_offset = (int64_t)M->hits.size();
                for (int32_t k = 0; k < 2 * m.n_hits; ++k) {
                    esch_hit h = pool.hits[(size_t)(h0 + k)];
                    const uint32_t *src = pool.ops.data() + h.alignment.ops_offset;
                    h.alignment.ops_offset = (int64_t)M->ops.size();
                    M->ops.insert(M->ops.end(), src, src + h.alignment.n_ops);
                    M->hits.push_back(h);
                }
            } else {
                m.hit_offset = (int64_t)M->hits.size();
                for (int k = 0; k < 2; ++k) {
                    if (m.separate[k].n_hits) place_read(M.get(), pool, m.separate[k]);
                    else m.separate[k].hit_offset = (int64_t)M->hits.size();
                }
            }
            M->pairs[p0 + i] = m;
        }
        M->stats.assemble_ms += ms_since(ta);
        give_batch(s, held.cur);
        held.cur = held.next;
        held.next = nullptr;
    }
    M->stats.total_ms = ms_since(t0);
    *out = M.release();
    return ESCH_OK;
}

}  // namespace

extern "C" {

esch_status esch_map_reads(esch_scorer *s, const uint8_t *const *reads, const size_t *lens, size_t n,
                           int both_strands, int32_t min_score, int64_t max_ties, esch_mapped **out) {
    if (!out) return fail(ESCH_INVALID_ARG, "out is null");
    *out = nullptr;
    if (!s || (n && (!reads || !lens))) return fail(ESCH_INVALID_ARG, "null argument");
    return map_reads_core(
        s, n,
        [&](size_t k, const uint8_t *&p, size_t &len) {
            p = reads[k];
            len = lens[k];
        },
        both_strands ? 1 : 0, min_score, max_ties, out);
}

esch_status esch_map_reads_concat(esch_scorer *s, const uint8_t *codes, const uint64_t *offs, const int32_t *lens,
                                  size_t n, int both_strands, int32_t min_score, int64_t max_ties,
                                  esch_mapped **out) {
    if (!out) return fail(ESCH_INVALID_ARG, "out is null");
    *out = nullptr;
    if (!s || (n && (!codes || !offs || !lens))) return fail(ESCH_INVALID_ARG, "null argument");
    return map_reads_core(
        s, n,
        [&](size_t k, const uint8_t *&p, size_t &len) {
            p = codes + offs[k];
            len = (size_t)(lens[k] < 0 ? 0 : lens[k]);
        },
        both_strands ? 1 : 0, min_score, max_ties, out);
}

esch_status esch_map_pairs(esch_scorer *s, const uint8_t *const *mates1, const size_t *lens1,
                           const uint8_t *const *mates2, const size_t *lens2, size_t n, int both_strands,
                           int32_t min_score, int64_t max_ties, int64_t max_insert, esch_mapped **out) {
    if (!out) return fail(ESCH_INVALID_ARG, "out is null");
    *out = nullptr;
    if (!s || (n && (!mates1 || !lens1 || !mates2 || !lens2))) return fail(ESCH_INVALID_ARG, "null argument");
    if (max_insert < 0) return fail(ESCH_INVALID_ARG, "max_insert is negative (%lld)", (long long)max_insert);
    return map_pairs_core(
        s, n,
        [&](size_t k, const uint8_t *&p, size_t &len) {
            p = (k & 1) ? mates2[k >> 1] : mates1[k >> 1];
            len = (k & 1) ? lens2[k >> 1] : lens1[k >> 1];
        },
        both_strands, min_score, max_ties, max_insert, out);
}

esch_status esch_map_pairs_concat(esch_scorer *s, const uint8_t *codes, const uint64_t *offs, const int32_t *lens,
                                  size_t n_pairs, int both_strands, int32_t min_score, int64_t max_ties,
                                  int64_t max_insert, esch_mapped **out) {
    if (!out) return fail(ESCH_INVALID_ARG, "out is null");
    *out = nullptr;
    if (!s || (n_pairs && (!codes || !offs || !lens))) return fail(ESCH_INVALID_ARG, "null argument");
    if (max_insert < 0) return fail(ESCH_INVALID_ARG, "max_insert is negative (%lld)", (long long)max_insert);
    return map_pairs_core(
        s, n_pairs,
        [&](size_t k, const uint8_t *&p, size_t &len) {
            p = codes + offs[k];
            len = (size_t)(lens[k] < 0 ? 0 : lens[k]);
        },
        both_strands, min_score, max_ties, max_insert, out);
}

esch_status esch_set_map_chunk(esch_scorer *s, size_t reads) {
    if (!s) return fail(ESCH_INVALID_ARG, "null scorer");
    if (reads == 1) return fail(ESCH_INVALID_ARG, "a chunk holds at least one pair (2 reads)");
    s->map_chunk = reads;
    return ESCH_OK;
}

size_t esch_mapped_len(const esch_mapped *m) {
    if (!m) return 0;
    return m->is_pairs ? m->pairs.size() : m->reads.size();
}
const esch_read_mapping *esch_mapped_reads(const esch_mapped *m) {
    return (m && !m->is_pairs) ? m->reads.data() : nullptr;
}
const esch_pair_mapping *esch_mapped_pairs(const esch_mapped *m) {
    return (m && m->is_pairs) ? m->pairs.data() : nullptr;
}
const esch_hit *esch_mapped_hits(const esch_mapped *m, size_t *n) {
    if (n) *n = m ? m->hits.size() : 0;
    return m ? m->hits.data() : nullptr;
}
const uint32_t *esch_mapped_ops(const esch_mapped *m, size_t *n) {
    if (n) *n = m ? m->ops.size() : 0;
    return m ? m->ops.data() : nullptr;
}
void esch_mapped_stats(const esch_mapped *m, esch_map_stats *out) {
    if (!out) return;
    *out = m ? m->stats : esch_map_stats{};
}
void esch_mapped_free(esch_mapped *m) { delete m; }

}  // extern "C"
