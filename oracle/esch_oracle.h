/*
 * CPU oracle for the panel-scoring path: TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of eschalign's score-only Gotoh DP and of the
 * selection that consumes its scores, used as the checker for the HIP
 * library. Nothing in the product links or calls this; only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg load it.
 *
 * Every function cites the reference file:line it restates (paths are
 * relative to the eschalign repository root; lib/ = crates/eschalign/src/).
 *
 * Pinning (what checks each function against something other than itself):
 *   eo_score          192 parasail 1.3.4 results (tests/golden/align_golden.json),
 *                     closed forms from lib/scalar.rs:378-430, brute-force
 *                     enumeration of every alignment for n, m <= 5.
 *   eo_align          the same 192 results' end and start coordinates,
 *                     the CIGAR span / first-last-op / rescore checks of
 *                     tests/parasail_golden.rs:82-110, the known answers of
 *                     lib/scalar.rs:378-430, and a full-matrix restatement of
 *                     the documented tie rules (lib/scalar.rs:28-37) on every
 *                     pair of tiny sequences.
 *   eo_fits_i16       the reference's own bounds test (lib/simd/mod.rs:1136-1140)
 *                     and the closed-form bound.
 *   eo_simd_*         equality with eo_score on the reference's own random
 *                     suites (xorshift64* inputs reproduced bit for bit).
 *   eo_select         lib/mapper.rs:571-589, 709-739 closed-form cases, and a
 *                     sort-based restatement in the tests.
 *   eo_pair_cands     lib/mapper/pair.rs:450-556 cases (score 120 paralog
 *                     choice, min-score, strands) and a brute-force sort.
 *   eo_map_reads      lib/mapper.rs:571-589 (ties in reference order, capped
 *                     hits), :709-739 (min-score unmaps, reverse strand).
 *   eo_map_pairs      every case of lib/mapper/pair.rs:449-573 (paralog
 *                     choice, insert and orientation limits, template strands,
 *                     min-score per mate, batched = one at a time) and the
 *                     outcomes crates/eschalign-cli/tests/align_paired.rs
 *                     states for the nine pairs of the paired fixture
 *                     (positions, inserts, proper or not, ties).
 *   eo_md_nm          the 21 samtools 1.23.1 calmd results of
 *                     tests/golden/calmd_golden.json (crates/eschalign/tests/
 *                     sam_fields.rs) and lib/sam.rs:150-156.
 *   eo_cigar_string   the same golden CIGARs round-tripped, lib/sam.rs:143-148.
 *   eo_adapter_*      lib/adapter.rs:116-197 known-answer tests.
 *   eo_encode etc.    lib/alphabet.rs:86-105 known answers.
 */
#ifndef ESCH_ORACLE_H
#define ESCH_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EO_WILDCARD 4
#define EO_NEG (INT32_MIN / 4)

/* lib/alphabet.rs:29-55 — letter -> code {A0,C1,G2,T/U3, everything else 4}. */
void eo_encode(const uint8_t *letters, size_t n, uint8_t *out);
/* lib/alphabet.rs:58-64 — reverse complement of codes, wildcard kept. */
void eo_revcomp_codes(const uint8_t *codes, size_t n, uint8_t *out);

/* lib/scoring.rs:70-100 — 0 when valid, else which rule failed (1..5). */
int eo_scoring_validate(int32_t ma, int32_t mi, int32_t o, int32_t e);
/* lib/scoring.rs:105-111 */
int32_t eo_substitution(int32_t ma, int32_t mi, uint8_t a, uint8_t b);

/* lib/scalar.rs:120-163 — the score-only DP (local=1: Smith-Waterman). */
int32_t eo_score(const uint8_t *q, size_t n, const uint8_t *r, size_t m,
                 int32_t ma, int32_t mi, int32_t o, int32_t e, int local);

/* lib/scalar.rs:164-364 — the traceback DP and walk. ops[] receives the
 * run-length operations (op << 30 | len; op 0 = M, 1 = I, 2 = D). Returns
 * the number of runs, or -1 if ops_cap is too small. */
typedef struct {
    int32_t score;
    int32_t query_start, query_end, ref_start, ref_end;
} eo_alignment;
int eo_align(const uint8_t *q, size_t n, const uint8_t *r, size_t m, int32_t ma,
             int32_t mi, int32_t o, int32_t e, int local, eo_alignment *out,
             uint32_t *ops, size_t ops_cap);

/* lib/scalar.rs:292-364 — the walk from end cell (bi, bj) through
 * traceback bytes tb[(j * stride + i) * lanes + lane]; shared by eo_align
 * (lanes = 1) and the SIMD pair kernel (lanes = 32). */
int eo_walk_trace(const uint8_t *tb, size_t stride, size_t lanes, size_t lane, int32_t best, size_t bi,
                  size_t bj, eo_alignment *out, uint32_t *ops, size_t ops_cap);

/* eo_align over pairs on `threads` workers; recs[k * 6] = score,
 * query_start, query_end, ref_start, ref_end, number of runs. */
void eo_align_batch(const uint8_t *qcodes, const uint64_t *qoff, const int32_t *qlen,
                    const uint8_t *rcodes, const uint64_t *roff, const int32_t *rlen,
                    const int32_t *refs, size_t n, int32_t ma, int32_t mi, int32_t o,
                    int32_t e, int local, int threads, int32_t *recs);

/* simd::align_pairs with the AVX-512BW pair kernel (lib/simd/mod.rs:384-499,
 * lib/simd/avx512.rs:120-247): the pairs (read k against reference
 * refs[k]) in chunks of 64 in input order, as the CLI's workers trace a
 * chunk's winners; within a chunk the SIMD-eligible pairs (non-empty, both
 * under 32,000, fits_i16) sorted by (read length, reference length) and cut
 * into groups of 32 lanes, one pair per lane, each lane walked by the
 * scalar walk; the others (and groups over 16 MiB of traceback) by
 * eo_align. Chunks are dealt to `threads` workers. recs as eo_align_batch;
 * ops (may be NULL) receives pair k's runs at ops + ops_off[k] (read +
 * reference length + 1 slots each). Without AVX-512BW every pair takes
 * eo_align. */
void eo_simd_align_pairs(const uint8_t *qcodes, const uint64_t *qoff, const int32_t *qlen,
                         const uint8_t *rcodes, const uint64_t *roff, const int32_t *rlen,
                         const int32_t *refs, size_t n, int32_t ma, int32_t mi, int32_t o, int32_t e,
                         int local, int threads, int32_t *recs, uint32_t *ops, const uint64_t *ops_off);

/* The same on the calling thread for one call's pairs (a chunk of reads'
 * taken ties, as map_reads_scored traces them, lib/mapper.rs:484-495): pair
 * k is q[k] (ql[k] codes) against r[k] (rl[k]); al[k], runs[k], and its
 * runs at ops + ops_off[k] (ql + rl + 1 slots). Buffers live in a scratch
 * kept by the caller between calls. */
typedef struct eo_pair_scratch eo_pair_scratch;
eo_pair_scratch *eo_pair_scratch_new(void);
void eo_pair_scratch_free(eo_pair_scratch *s);
void eo_simd_align_chunk(eo_pair_scratch *s, const uint8_t *const *q, const size_t *ql, const uint8_t *const *r,
                         const size_t *rl, size_t cnt, int32_t ma, int32_t mi, int32_t o, int32_t e, int local,
                         eo_alignment *al, int *runs, uint32_t *ops, const size_t *ops_off);

/* lib/simd/mod.rs:222-227 */
int eo_fits_i16(size_t n, size_t max_ref_len, int32_t ma, int32_t mi,
                int32_t o, int32_t e);

/* Every query against every reference with eo_score, row-major
 * out[q * n_refs + r], spread over `threads` POSIX threads. */
void eo_score_matrix(const uint8_t *qcodes, const uint64_t *qoff,
                     const int32_t *qlen, size_t nq, const uint8_t *rcodes,
                     const uint64_t *roff, const int32_t *rlen, size_t nr,
                     int32_t ma, int32_t mi, int32_t o, int32_t e, int local,
                     int threads, int32_t *out);

/* lib/mapper.rs:418-482 — one read's score-stage selection.
 * fwd: n_refs scores; rev: n_refs scores or NULL (forward strand only).
 * Candidates are taken in tie order (ref r forward, ref r reverse, r+1 ...).
 * max_ties < 0 means "all" (None). Writes up to `cap` taken ties as
 * (ref << 1) | reverse into ties. Returns the number taken. */
int32_t eo_select(const int32_t *fwd, const int32_t *rev, size_t n_refs,
                  int32_t min_score, int64_t max_ties, int32_t *best,
                  int32_t *n_tied, int32_t *has_subopt, int32_t *subopt,
                  int32_t *ties, size_t cap);

/* lib/mapper/pair.rs:171-231 — eligible candidates of one pair, sorted by
 * (-sum, ref, template strand). rows[m*2 + rv] are mate m's n_refs scores
 * as sequenced (rv=0) or reverse-complemented (rv=1); a NULL row marks an
 * empty mate (no candidates). Writes cand = (ref << 1) | template_reverse,
 * sums and per-mate scores; returns the count (<= 2 * n_refs). */
int32_t eo_pair_cands(const int32_t *const rows[4], size_t n_refs,
                      int both_strands, int32_t min_score, int32_t *cand,
                      int32_t *sum, int32_t *s0, int32_t *s1);

/* ---- read and pair mapping (esch_map.c) ----
 * Aligner::map_reads_scored (lib/mapper.rs:400-550) and map_pairs_scored
 * (lib/mapper/pair.rs:157-387) on caller-computed panel scores, with every
 * traceback by eo_align. The records have the layout of the library's
 * esch_hit / esch_read_mapping / esch_pair_mapping. */
typedef struct {
    int32_t reference;
    int32_t reverse;          /* the read's / mate's own strand */
    int32_t template_reverse; /* proper-pair hits: PairHit::reverse */
    int32_t insert;           /* proper-pair hits: PairHit::insert */
    int32_t score, query_start, query_end, ref_start, ref_end, n_ops;
    int64_t ops_offset;
} eo_hit;
typedef struct {
    int32_t best, n_tied, subopt;
    int32_t flags; /* bit 0 best present (non-empty read), bit 1 suboptimal present */
    int32_t n_hits, reserved;
    int64_t hit_offset;
} eo_read_mapping;
typedef struct {
    int32_t proper; /* 1 Proper, 0 Separate */
    int32_t score, n_tied;
    int32_t flags; /* proper: bit m = suboptimal[m] present */
    int32_t subopt[2];
    int32_t n_hits; /* proper: pair hits, 2 eo_hit each (mate 1, mate 2) */
    int32_t reserved;
    int64_t hit_offset;
    eo_read_mapping separate[2];
} eo_pair_mapping;
typedef struct eo_mapped eo_mapped;
/* fwd_rows[k * n_refs + r] (and rev_rows, NULL = forward strand only): read
 * k's scores as sequenced / reverse-complemented. max_ties < 0 = all. */
eo_mapped *eo_map_reads(const uint8_t *codes, const uint64_t *off, const int32_t *len, size_t n,
                        const int32_t *fwd_rows, const int32_t *rev_rows, const uint8_t *rcodes,
                        const uint64_t *roff, const int32_t *rlen, size_t n_refs, int32_t ma, int32_t mi,
                        int32_t o, int32_t e, int local, int32_t min_score, int64_t max_ties, int threads);
/* Mates interleaved: read 2p is pair p's mate 1, 2p + 1 its mate 2; both
 * row sets are required (the fallback maps each mate on both strands). */
eo_mapped *eo_map_pairs(const uint8_t *codes, const uint64_t *off, const int32_t *len, size_t n_pairs,
                        const int32_t *fwd_rows, const int32_t *rev_rows, const uint8_t *rcodes,
                        const uint64_t *roff, const int32_t *rlen, size_t n_refs, int32_t ma, int32_t mi,
                        int32_t o, int32_t e, int local, int32_t min_score, int64_t max_ties, int64_t max_insert,
                        int both_strands, int threads);
size_t eo_mapped_len(const eo_mapped *m);
const eo_read_mapping *eo_mapped_reads(const eo_mapped *m);
const eo_pair_mapping *eo_mapped_pairs(const eo_mapped *m);
const eo_hit *eo_mapped_hits(const eo_mapped *m, size_t *n);
const uint32_t *eo_mapped_ops(const eo_mapped *m, size_t *n);
void eo_mapped_free(eo_mapped *m);

/* sam::md_nm (lib/sam.rs:76-124) on letters: MD into md (NUL-terminated),
 * NM into *nm; returns MD's length, -1 if md_cap is short or the ops run
 * past a sequence. sam::cigar_string (lib/sam.rs:32-52): soft clips around
 * the ops; returns the length, -1 if cap is short. */
int eo_md_nm(const uint8_t *read, size_t n, const uint8_t *ref, size_t m, size_t qs, size_t rs,
             const uint32_t *ops, size_t n_ops, char *md, size_t md_cap, uint32_t *nm);
int eo_cigar_string(const uint32_t *ops, size_t n_ops, size_t query_len, size_t qs, size_t qe, char *out,
                    size_t cap);

/* lib/adapter.rs:68-91 — adapter bases at the 3' end of a read (letters). */
size_t eo_adapter_find(const uint8_t *read, size_t n, const uint8_t *adapter,
                       size_t alen, size_t min_overlap);
/* lib/adapter.rs:95-110 */
void eo_adapter_detect_pair(const uint8_t *r1, size_t l1, const uint8_t *r2,
                            size_t l2, const uint8_t *a1, size_t la1,
                            const uint8_t *a2, size_t la2, size_t min_overlap,
                            size_t out[2]);

/* ---- SIMD restatement (the CPU baseline; esch_simd.c) ---- */

/* Which kernels the host CPU supports: bit 0 AVX2, bit 1 AVX-512BW. */
int eo_simd_supported(void);

typedef struct eo_profile eo_profile;
/* lib/simd/mod.rs:286-331 (+ tprof :255-269): refs longest-first, stable,
 * cut into groups of `lanes` (16 = AVX2, 32 = AVX-512BW). */
eo_profile *eo_profile_build(const uint8_t *rcodes, const uint64_t *roff,
                             const int32_t *rlen, size_t nr, int32_t ma,
                             int32_t mi, int32_t o, int32_t e, int lanes);
void eo_profile_free(eo_profile *p);
/* lib/mapper.rs:216-222 + 242-252 score_all: row-major below
 * transposed_min (<= 0: never transposed), transposed at and above, scalar
 * when !fits_i16. kernel: -1 dispatch, 0 force row-major, 1 force transposed. */
void eo_simd_score_all(const eo_profile *p, const uint8_t *q, size_t n,
                       int local, int transposed_min, int kernel, int32_t *out);
/* Batch driver for the CPU baseline: reads dealt to `threads` workers in
 * chunks of 64 (crates/eschalign-cli/src/align.rs:1470), every query
 * scored with eo_simd_score_all; out row-major [q * n_refs + r]. */
void eo_simd_score_batch(const eo_profile *p, const uint8_t *qcodes,
                         const uint64_t *qoff, const int32_t *qlen, size_t nq,
                         int local, int threads, int32_t *out);

#ifdef __cplusplus
}
#endif
#endif
