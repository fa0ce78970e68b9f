/*
 * eschalign_amd — MI355X (gfx950) panel scoring behind a C ABI.
 *
 * Drop-in for eschalign's GPU score backend: every entry point below is what
 * the reference's `GpuScorer` / `ScoreMatrix` boundary would bind through a
 * thin FFI crate. Citations are relative to the eschalign repository root
 * (lib/ = crates/eschalign/src/).
 *
 * Conventions (the reference's, lib/cuda/mod.rs:516-519, lib/alphabet.rs):
 *   - sequences are alphabet CODES 0..=4 (A,C,G,T/U, wildcard), not letters;
 *   - scores row-major [query * n_refs + ref], references in panel order;
 *   - no exceptions cross the ABI: every call returns an esch_status, and
 *     esch_last_error() gives the text of the calling thread's last error.
 *
 * Threading (lib/cuda/mod.rs:328-341, crates/eschalign-cli/src/align.rs:2101):
 * esch_prepare is thread-safe and touches no device state; esch_submit /
 * esch_score_batch on one scorer must not run concurrently (one thread per
 * scorer, as the reference's &mut self). Device batches (esch_batch_*) are
 * independent of each other and may run on several threads at once.
 */
#ifndef ESCHALIGN_AMD_H
#define ESCHALIGN_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* GpuError (lib/cuda/mod.rs:144-153) plus the crate's AlignError variants
 * that the boundary can raise (lib/error.rs:7-28). */
typedef enum {
    ESCH_OK = 0,
    ESCH_UNAVAILABLE = 1,           /* no HIP runtime / no device       GpuError::Unavailable */
    ESCH_DRIVER = 2,                /* a HIP call failed                 GpuError::Driver */
    ESCH_CODE_OBJECT = 3,           /* no gfx950 code object for device  GpuError::Compile */
    ESCH_REFERENCE_TOO_LONG = 4,    /* panel reference > ESCH_MAX_REF_LEN GpuError::ReferenceTooLong */
    ESCH_INVALID_ARG = 5,           /* null pointer, bad size, bad mode */
    ESCH_SCORING_INVALID = 6,       /* AlignError::ScoringInvalid (lib/scoring.rs:70-100) */
    ESCH_EMPTY_PANEL = 7,           /* AlignError::EmptyPanel */
    ESCH_EMPTY_REFERENCE = 8,       /* AlignError::EmptyReference */
    ESCH_CAPACITY = 9,              /* a device batch is smaller than the input */
    ESCH_TIMEOUT = 10               /* a served request is not finished yet */
} esch_status;

typedef enum { ESCH_LOCAL = 0, ESCH_SEMIGLOBAL = 1 } esch_mode; /* lib/scoring.rs:151-165 */

/* Longest read the drop-in score matrix scores (lib/cuda/mod.rs:90-91,
 * the same scheduling bound as the reference). */
#define ESCH_MAX_READ_LEN 4096
/* Longest panel reference this library accepts. The reference's CUDA
 * kernel stops at 3,072 nt (shared memory, lib/cuda/mod.rs:94-95); here a
 * reference is streamed through scalar registers and its DP boundary lives
 * in device scratch sized by the panel, so the bound is the i16 profile
 * index of the reference's own SIMD layout (lens: Vec<i16>). */
#define ESCH_MAX_REF_LEN 16384

typedef struct esch_scorer esch_scorer;     /* GpuScorer        lib/cuda/mod.rs:328-341 */
typedef struct esch_prepared esch_prepared; /* PreparedBatch    lib/cuda/mod.rs:711-720 */
typedef struct esch_pending esch_pending;   /* PendingBatch     lib/cuda/mod.rs:730-736 */
typedef struct esch_batch esch_batch;       /* device-resident batch (no counterpart; see below) */

/* Text of the calling thread's last error ("" if none). */
const char *esch_last_error(void);

/* cuda::device_visible() — lib/cuda/mod.rs:130-140. Cached per process;
 * 0 when the HIP runtime sees no device. */
int esch_device_visible(void);

/* Validate a scoring scheme exactly as Scoring::validate
 * (lib/scoring.rs:70-100): ESCH_OK or ESCH_SCORING_INVALID. */
esch_status esch_scoring_validate(int32_t match, int32_t mismatch, int32_t gap_open,
                                  int32_t gap_extend);

/* simd::fits_i16 — lib/simd/mod.rs:222-227. */
int esch_fits_i16(size_t n, size_t max_ref_len, int32_t match, int32_t mismatch,
                  int32_t gap_open, int32_t gap_extend);

/* GpuScorer::new_on_device(ordinal, &Panel, Scoring, Mode) —
 * lib/cuda/mod.rs:350-443 (with Panel::new's refusals, lib/panel.rs:32-57).
 * Uploads the panel once (2-bit codes + wildcard bitmap) and sizes the
 * kernels for this device. refs[i] points at ref_lens[i] codes. */
esch_status esch_scorer_new(int ordinal, const uint8_t *const *refs, const size_t *ref_lens,
                            size_t n_refs, int32_t match, int32_t mismatch, int32_t gap_open,
                            int32_t gap_extend, esch_mode mode, esch_scorer **out);
/* Drop for GpuScorer. NULL is a no-op. */
void esch_scorer_free(esch_scorer *s);

/* GpuScorer::scores_len — lib/cuda/mod.rs:506-508:
 * 0 < len <= ESCH_MAX_READ_LEN && fits_i16(len, panel max_len). */
int esch_scores_len(const esch_scorer *s, size_t len);
size_t esch_n_refs(const esch_scorer *s);

/* GpuScorer::prepare — lib/cuda/mod.rs:533-576. Host only, thread-safe,
 * may overlap a running kernel. Copies the queries; the caller keeps
 * ownership of its buffers. Returns NULL (and sets the error) on bad args. */
esch_prepared *esch_prepare(const esch_scorer *s, const uint8_t *const *queries,
                            const size_t *lens, size_t n);
void esch_prepared_free(esch_prepared *p);
/* Rows of the prepared batch and how many of them will be scored. */
size_t esch_prepared_len(const esch_prepared *p);
size_t esch_prepared_n_scored(const esch_prepared *p);

/* GpuScorer::submit — lib/cuda/mod.rs:581-699. Consumes `p` (also on
 * error). Uploads and queues the kernels on the scorer's stream; returns
 * without waiting. */
esch_status esch_submit(esch_scorer *s, esch_prepared *p, esch_pending **out);

/* GpuScorer::finish + ScoreMatrix::from_parts — lib/cuda/mod.rs:702-708,
 * lib/mapper.rs:107-115. Waits, then writes scores[n * n_refs] (int16,
 * rows of unscored queries zero) and scored[n] (1 = row present, 0 = the
 * caller scores it on the CPU, ScoreMatrix::row -> None). Consumes `p`. */
esch_status esch_finish(esch_scorer *s, esch_pending *p, int16_t *scores, uint8_t *scored);

/* GpuScorer::score_batch — lib/cuda/mod.rs:524-528: prepare + submit +
 * finish. n == 0 is valid (an empty matrix). */
esch_status esch_score_batch(esch_scorer *s, const uint8_t *const *queries, const size_t *lens,
                             size_t n, int16_t *scores, uint8_t *scored);

/* Diagnostics — lib/cuda/mod.rs:446-488. geometry = (threads per block of
 * the batched kernel, references per work item, persistent blocks). */
esch_status esch_device_name(const esch_scorer *s, char *buf, size_t cap);
void esch_geometry(const esch_scorer *s, size_t *lanes, size_t *groups, size_t *blocks);
/* Measured packed-int16 VALU peak of device `ordinal` (lane-ops/s: a
 * saturating v_pk_add_i16 / v_pk_max_i16 microbenchmark over every SIMD),
 * with the CU count and the reported peak clock; no counterpart in the
 * reference (the roofline of the batched kernel is quoted against it). */
esch_status esch_valu_peak16(int32_t ordinal, double *lane_ops_per_s, int *n_cu, int *clock_khz);
int esch_registers(const esch_scorer *s);   /* VGPRs of the batched kernel */
int esch_local_bytes(const esch_scorer *s); /* private (spill) bytes of the batched kernel */

/* ---------------------------------------------------------------------
 * Extensions beyond the reference's i16 ScoreMatrix (no counterpart in the
 * reference; they serve the same callers, Aligner::map_reads_scored and
 * map_pairs_scored, lib/mapper.rs:400-405, lib/mapper/pair.rs:157-162).
 * --------------------------------------------------------------------- */

/* Every query scored, int32 out[n * n_refs]; a query outside the packed
 * int16 bound is rescored by the int32 kernel, so nothing is left for the
 * CPU except empty queries (row of zeros, scored[k] = 0). */
esch_status esch_score_batch_i32(esch_scorer *s, const uint8_t *const *queries,
                                 const size_t *lens, size_t n, int32_t *scores, uint8_t *scored);

/* Allocation-free scoring. No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). The scoring of
 * esch_score_batch_i32 (every non-empty query scored, the int32 rescue
 * outside the int16 bound; an empty query gives a zero row, scored = 0)
 * on queries already in DEVICE memory: query k is codes[offs[k] ..
 * offs[k] + lens[k]); scores[n * n_refs] (int32) and scored[n] are device
 * pointers too. The work is queued on `stream` (a hipStream_t; NULL = the
 * null stream) and the call returns without waiting. It calls no hipMalloc
 * or hipFree: every buffer of its kernels lives in one device array of
 * ESCH_NOALLOC_SCRATCH_BYTES per GPU, held from the first call on that GPU
 * until the process ends. So n is at most ESCH_NOALLOC_MAX_ROWS
 * (ESCH_CAPACITY beyond), and a panel whose longest reference needs more
 * DP boundary scratch than the array holds is refused with ESCH_CAPACITY
 * (on a panel of long references the persistent grids are cut to fit the
 * array). Calls that overlap on one GPU share its array and race: serialise
 * them (one stream, or a wait between calls). The first call on a device
 * resolves the array's address: make it, n = 0 does only that, before
 * capturing a graph. */
#define ESCH_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
#define ESCH_NOALLOC_MAX_ROWS ((size_t)1 << 20)
esch_status esch_score_device_noalloc(esch_scorer *s, const uint8_t *codes, const uint64_t *offs,
                                      const int32_t *lens, size_t n, int32_t *scores, uint8_t *scored,
                                      void *stream);

/* One read's score-stage selection, fused on device: the first half of
 * Aligner::map_reads_scored (lib/mapper.rs:444-472). */
typedef struct {
    int32_t best;      /* best score over every (ref, strand); 0 when empty */
    int32_t n_tied;    /* size of the tie set */
    int32_t subopt;    /* best score outside the tie set (valid when has_subopt) */
    int32_t flags;     /* bit 0: read non-empty (best present); bit 1: has_subopt */
    int32_t n_take;    /* ties to trace: min(n_tied, max_ties + 1), 0 below min_score */
} esch_read_select;

/* Reads (codes) scored forward and, with both_strands, reverse-complemented
 * on device; the tie set in reference order, forward before reverse.
 * max_ties < 0 = all (None). ties[k * tie_cap + i] = (ref << 1) | reverse
 * for i < min(n_take, tie_cap). */
esch_status esch_select_batch(esch_scorer *s, const uint8_t *const *reads, const size_t *lens,
                              size_t n, int both_strands, int32_t min_score, int64_t max_ties,
                              size_t tie_cap, esch_read_select *out, int32_t *ties);

/* One pair's candidates (lib/mapper/pair.rs:185-231), fused on device:
 * the top tier of eligible (ref, template strand) candidates by summed mate
 * score, both mates >= min_score. */
typedef struct {
    int32_t n_cand;     /* eligible candidates */
    int32_t top_sum;    /* best summed score (valid when n_cand > 0) */
    int32_t n_top;      /* candidates in the top tier */
    int32_t next_sum;   /* the next lower sum (valid when flags bit 1) */
    int32_t flags;      /* bit 0: both mates non-empty; bit 1: a second tier exists */
} esch_pair_select;

/* mates1[k], mates2[k]: codes of pair k's mates (adapters already
 * trimmed). top[k * top_cap + i] = (ref << 1) | template_reverse. */
esch_status esch_select_pairs(esch_scorer *s, const uint8_t *const *mates1, const size_t *lens1,
                              const uint8_t *const *mates2, const size_t *lens2, size_t n,
                              int both_strands, int32_t min_score, size_t top_cap,
                              esch_pair_select *out, int32_t *top);

/* Reads at least min_len long are scored by the anti-diagonal wavefront
 * kernel (one read against two references per 16-lane DPP row) instead of
 * the length-bucketed inter-sequence kernel; 0 (the default) sends none.
 * For batches whose long reads are too few per length to fill 128-read
 * chunks. Identical scores either way. */
esch_status esch_set_diag_min_len(esch_scorer *s, int32_t min_len);
int32_t esch_get_diag_min_len(const esch_scorer *s);
/* Adaptive routing per length bucket (on by default): after bucketing, the
 * reads of length L take the anti-diagonal kernel when L > rho * L' + 64,
 * L' the length of the read 127 ranks later (the chunk they would lead),
 * or when L >= 512 and the batch is too small to fill the device with
 * inter-sequence work items (128-read chunks x 4 references). Reads left
 * that do not fill one 128-read chunk all go when rho times their rows,
 * padded to the anti-diagonal kernel's passes (64 R rows; R = 2, 4, 8 for
 * lengths up to 128, 256, beyond), is less than 128 x the longest of them.
 * rho = ratio_x16 / 16, default 32 (2.0); 0 turns the routing off;
 * otherwise 16..1024. esch_autotune sweeps it. Identical scores either
 * way. */
esch_status esch_set_diag_ratio(esch_scorer *s, int32_t ratio_x16);
int32_t esch_get_diag_ratio(const esch_scorer *s);
/* The per-length-bucket plan of that routing: one entry per length class,
 * c = 0 under 128 nt (never routed), 1: 128-255, 2: 256-511, 3: 512-1023,
 * 4: 1024-2047, 5: 2048-4095, 6: 4096-8191, 7: 8192 and over. An entry is
 * 0 (the class's reads stay on the inter-sequence kernel), 1 (they take
 * the anti-diagonal kernel) or 16..1024 (the ratio rule above with rho =
 * entry / 16). esch_set_diag_ratio sets every entry to its ratio (so 0
 * turns all routing off); esch_autotune tunes them per class. The ratio
 * alone rules the small-batch cases (a batch that cannot fill the device,
 * kept rows that do not fill a chunk). Identical scores whatever the plan.
 * n must be ESCH_LEN_CLASSES. */
#define ESCH_LEN_CLASSES 8
esch_status esch_set_diag_plan(esch_scorer *s, const int32_t *plan, size_t n);
esch_status esch_get_diag_plan(const esch_scorer *s, int32_t *plan, size_t n);
/* A batch of up to `rows` rows (reads x strands; by default as many rows
 * as the panel has references, when it has at least 4,096, and none
 * otherwise) is scored with the roles swapped: the panel's references as
 * the queries, bucketed into well-filled 128-row chunks, and the batch's
 * rows as the references. The score of a read against a reference is the
 * reference's against the read (the same substitutions, gaps, zero borders
 * and best cell), so the results are identical; a few reads against a
 * large panel no longer pay a whole chunk per group of references, and no
 * read pads another. Only for non-empty rows no longer
 * than the panel's longest reference, int32 scores (the fused selection
 * included), and no length threshold set (esch_set_diag_min_len).
 * rows: -1 restores the rule, 0 turns it off. esch_get_swap_rows returns
 * the limit in force. */
esch_status esch_set_swap_rows(esch_scorer *s, int32_t rows);
int32_t esch_get_swap_rows(const esch_scorer *s);

/* adapter::find — lib/adapter.rs:68-91. Letters (not codes): how many bases
 * at the 3' end of `read` are the adapter's prefix (0 = none). The pre-step
 * that trims mates before scoring (crates/eschalign-cli/src/align.rs:2844-2862). */
size_t esch_adapter_find(const uint8_t *read, size_t n, const uint8_t *adapter, size_t alen,
                         size_t min_overlap);
/* adapter::detect_pair — lib/adapter.rs:95-110: out = [ad1, ad2]. */
void esch_adapter_detect_pair(const uint8_t *r1, size_t l1, const uint8_t *r2, size_t l2,
                              const uint8_t *a1, size_t la1, const uint8_t *a2, size_t la2,
                              size_t min_overlap, size_t out[2]);

/* The same pre-step over a batch of pairs held as codes, mates interleaved
 * (mate m of pair p at codes[offs[2p + m]], lens[2p + m] bases); adapters are
 * letters. out[k] = lens[k] less mate k's detected adapter: the queries the
 * CLI aligns (crates/eschalign-cli/src/align.rs:1642-1662). */
void esch_adapter_trim_pairs(const uint8_t *codes, const uint64_t *offs, const int32_t *lens, size_t n_pairs,
                             const uint8_t *a1, size_t la1, const uint8_t *a2, size_t la2, size_t min_overlap,
                             int32_t *out);

/* ---- winner tracebacks (lib/simd/mod.rs:405-499, lib/scalar.rs:164-364) ----
 * The alignment of read queries[k] (codes, lens[k] bases) against panel
 * reference refs[k], for every k: the score, the aligned spans (0-based,
 * half-open: query[query_start..query_end] against
 * reference[ref_start..ref_end]) and the run-length operations, written to
 * ops[ops_offset .. ops_offset + n_ops) as (op << 30 | length) with op 0 = M,
 * 1 = I, 2 = D. The optimum chosen among several is the reference's (first
 * maximum of a column-major scan; diagonal, then insertion, then deletion;
 * gaps extended before opened; local stops at the first zero). An empty read
 * gives an empty alignment (all zero); so does a local alignment that cannot
 * score above zero. *ops_used receives the number of runs; CAPACITY when
 * ops_cap is smaller (read + reference length per pair always suffices). */
typedef struct {
    int32_t score;
    int32_t query_start, query_end, ref_start, ref_end;
    int32_t n_ops;
    int64_t ops_offset;
} esch_alignment;
esch_status esch_align_pairs(esch_scorer *s, const uint8_t *const *queries, const size_t *lens,
                             const int32_t *refs, size_t n, esch_alignment *out, uint32_t *ops,
                             size_t ops_cap, size_t *ops_used);
/* The same for reads in one concatenated buffer: read k is
 * codes[offs[k] .. offs[k] + lens[k]). */
esch_status esch_align_pairs_concat(esch_scorer *s, const uint8_t *codes, const uint64_t *offs,
                                    const int32_t *lens, const int32_t *refs, size_t n, esch_alignment *out,
                                    uint32_t *ops, size_t ops_cap, size_t *ops_used);
/* Device time (ms, HIP events) of the traceback kernels of the scorer's
 * last esch_align_pairs call (diagnostics; no counterpart in the reference). */
float esch_align_last_ms(const esch_scorer *s);
/* How many pairs of that call took the anti-diagonal (one wave per pair)
 * traceback kernel rather than one lane each: pairs over 2^27 cells, and
 * pairs of at least 128 read bases whose cells reach 2^-18 of the call's
 * (and 2^15), on references up to 4,096 nt. Diagnostics. */
size_t esch_align_last_wavefront(const esch_scorer *s);

/* ---- read and pair mapping: scoring, selection, tracebacks and the pair
 * tier loop on the device (lib/mapper.rs:400-550, lib/mapper/pair.rs:157-387).
 * Reads and mates are codes (adapters already trimmed, as the CLI does,
 * crates/eschalign-cli/src/align.rs:1642-1662); MD / NM are left to the
 * caller (lib/sam.rs:76-140 needs only the letters and these records).
 *
 * esch_hit — Hit (lib/mapper.rs:43-56) without md / nm; for the hits of a
 * proper pair also PairHit::reverse and PairHit::insert
 * (lib/mapper/pair.rs:69-81). */
typedef struct {
    int32_t reference;
    int32_t reverse;          /* the read's (mate's) own strand */
    int32_t template_reverse; /* proper-pair hits: the template strand; else 0 */
    int32_t insert;           /* proper-pair hits: the span of both mates; else 0 */
    esch_alignment alignment; /* ops in esch_mapped_ops */
} esch_hit;

/* ReadMapping (lib/mapper.rs:58-72). */
typedef struct {
    int32_t best;     /* best_score (valid when flags bit 0) */
    int32_t n_tied;
    int32_t subopt;   /* suboptimal (valid when flags bit 1) */
    int32_t flags;    /* bit 0: best present (read non-empty); bit 1: suboptimal present */
    int32_t n_hits;   /* 0: unmapped */
    int32_t reserved;
    int64_t hit_offset; /* hits[hit_offset .. hit_offset + n_hits) */
} esch_read_mapping;

/* PairMapping (lib/mapper/pair.rs:90-107). */
typedef struct {
    int32_t proper;    /* 1: Proper; 0: Separate */
    int32_t score;     /* Proper: the summed score of the tie set */
    int32_t n_tied;    /* Proper: size of the tie set */
    int32_t flags;     /* Proper: bit m = suboptimal[m] present */
    int32_t subopt[2]; /* Proper: per mate, the best score outside the tie set */
    int32_t n_hits;    /* Proper: pair hits kept (primary first, up to max_ties + 1);
                          hit h is hits[hit_offset + 2h] (mate 1) and [.. + 2h + 1] (mate 2) */
    int32_t reserved;
    int64_t hit_offset;
    esch_read_mapping separate[2]; /* Separate: mate 1 and mate 2 as unpaired reads */
} esch_pair_mapping;

typedef struct esch_mapped esch_mapped;

/* Aligner::map_reads_scored (lib/mapper.rs:400-550) with the scores computed
 * on the device: selection over (ref r forward, ref r reverse, r + 1, ...),
 * the taken ties (max_ties + 1, all when max_ties < 0; none below
 * min_score) traced, a read unmapped when its primary traces to nothing. */
esch_status esch_map_reads(esch_scorer *s, const uint8_t *const *reads, const size_t *lens, size_t n,
                           int both_strands, int32_t min_score, int64_t max_ties, esch_mapped **out);
/* The same with reads in one buffer: read k is codes[offs[k] .. + lens[k]). */
esch_status esch_map_reads_concat(esch_scorer *s, const uint8_t *codes, const uint64_t *offs, const int32_t *lens,
                                  size_t n, int both_strands, int32_t min_score, int64_t max_ties,
                                  esch_mapped **out);
/* Aligner::map_pairs_scored (lib/mapper/pair.rs:157-387): candidates
 * (reference, template strand) with both mates >= min_score, taken tier by
 * tier in descending summed score; the first tier with a candidate passing
 * pair_ok (both mates aligned, no outie, insert <= max_insert) is the tie
 * set; pairs with none are mapped mate by mate on both strands. */
esch_status esch_map_pairs(esch_scorer *s, const uint8_t *const *mates1, const size_t *lens1,
                           const uint8_t *const *mates2, const size_t *lens2, size_t n, int both_strands,
                           int32_t min_score, int64_t max_ties, int64_t max_insert, esch_mapped **out);
/* The same with mates in one buffer, interleaved: mate m of pair p is
 * codes[offs[2p + m] .. + lens[2p + m]). */
esch_status esch_map_pairs_concat(esch_scorer *s, const uint8_t *codes, const uint64_t *offs, const int32_t *lens,
                                  size_t n_pairs, int both_strands, int32_t min_score, int64_t max_ties,
                                  int64_t max_insert, esch_mapped **out);
size_t esch_mapped_len(const esch_mapped *m);
const esch_read_mapping *esch_mapped_reads(const esch_mapped *m); /* NULL for pairs */
const esch_pair_mapping *esch_mapped_pairs(const esch_mapped *m); /* NULL for reads */
const esch_hit *esch_mapped_hits(const esch_mapped *m, size_t *n);
const uint32_t *esch_mapped_ops(const esch_mapped *m, size_t *n);
/* Where the time of the call went (diagnostics; no counterpart in the
 * reference), in ms. A chunk's selection and tracebacks run between two
 * launches of the next chunk's scoring, so select_ms and trace_ms include
 * the wait for the launch they run behind: */
typedef struct {
    double score_ms;      /* device scoring pipelines (HIP events; their launches alone) */
    double select_ms;     /* tier / suboptimal / selection kernels with their copies */
    double trace_ms;      /* device traceback kernels (HIP events) */
    double trace_wall_ms; /* the traceback calls, host staging and copies included */
    double load_ms;       /* staging reads into device batches */
    double assemble_ms;   /* records placed in input order */
    double total_ms;      /* the whole call */
    int32_t rounds;       /* tier rounds */
    int32_t reserved;
    int64_t traced;       /* (read, reference) tracebacks */
    int64_t cells;        /* DP cells scored (read length x panel bases) */
} esch_map_stats;
void esch_mapped_stats(const esch_mapped *m, esch_map_stats *out);
void esch_mapped_free(esch_mapped *m);
/* Reads (mates) per device chunk of the mapping calls; 0 (the default)
 * sizes chunks so their resident int32 rows take about 4 GiB. Results do
 * not depend on it (no counterpart in the reference). */
esch_status esch_set_map_chunk(esch_scorer *s, size_t reads);

/* ---- SAM fields (lib/sam.rs:32-124; host code) ----
 * sam::md_nm: MD and NM as samtools calmd writes them for an alignment of
 * read letters (in the orientation aligned) against reference letters (as
 * in the FASTA), ops beginning at query_start / ref_start. MD is written
 * NUL-terminated into md (truncated to md_cap - 1 characters); returns its
 * full length, or (size_t)-1 when the ops run past either sequence. */
size_t esch_md_nm(const uint8_t *read, size_t n, const uint8_t *reference, size_t m, size_t query_start,
                  size_t ref_start, const uint32_t *ops, size_t n_ops, char *md, size_t md_cap, uint32_t *nm);
/* sam::cigar_string: the CIGAR of an aligned span inside a read of
 * query_len bases, soft clips before query_start and after query_end;
 * returns the full length ((size_t)-1 on inconsistent spans). */
size_t esch_cigar_string(const uint32_t *ops, size_t n_ops, size_t query_len, size_t query_start, size_t query_end,
                         char *out, size_t cap);

/* ---- output records (crates/eschalign-cli/src/align.rs:1084-1196; host code) ----
 * build_records over the mappings of an unpaired run, each output record
 * laid out as a BAM alignment record (block_size, then the fixed fields,
 * name, CIGAR, sequence, qualities and data; SAM/BAM specification 4.2), as
 * the CLI writes them. `in` holds the n input records as an uncompressed BAM
 * record stream (each a u32 block_size and the record), in input order,
 * primary records only; maps[k] is read k's mapping (hits / ops as
 * esch_mapped holds them: esch_map_reads on the reads' codes, a read the
 * caller keeps from alignment, such as one over --max-read-len, as an
 * unmapped record with n_hits = 0). References are the panel's letters (as
 * in the FASTA) and names. Each record keeps its input's name, sequence,
 * qualities, QC-fail / duplicate flags and the data fields the run does not
 * own; NM, MD, AS, XS, XA are rewritten (removed, the record's last field
 * taking a removed one's place, then appended); a reverse-strand primary
 * carries the reverse-complemented sequence and reversed qualities; MAPQ is
 * 60 for a unique best and 0 for a tie; with secondary set, each other hit
 * follows as a 0x100 record (MAPQ 0, no sequence, NM / MD / AS / RG). The
 * records go to out (*out_len bytes, *n_records records); ESCH_CAPACITY
 * when out_cap is smaller (*out_len says how much is needed). */
esch_status esch_bam_build_records(const uint8_t *in, size_t in_bytes, size_t n, const esch_read_mapping *maps,
                                   const esch_hit *hits, size_t n_hits, const uint32_t *ops, size_t n_ops,
                                   const uint8_t *const *ref_letters, const size_t *ref_lens,
                                   const char *const *ref_names, size_t n_refs, int secondary, uint8_t *out,
                                   size_t out_cap, size_t *out_len, size_t *n_records);

/* Kernel variants of the batched DP, and the autotuner that picks one by
 * timing each on a sample of the caller's reads (no counterpart in the
 * reference; the scores are identical whichever variant runs):
 *   0  profile plane loaded in the column that uses it (fewest registers);
 *   1  profile planes ping-ponged, stripe boundary prefetched;
 *   2  every memory access of a column pair issued at its start, stores
 *      deferred by one pair, no data-dependent branch around a memory
 *      access;
 *   3  columns in quads: two pairs whose boundary registers swap roles;
 *   4  variant 3 in packed
 *      fp16 arithmetic, whose three-way maximum is one instruction: 7
 *      packed ops per cell pair in semiglobal mode instead of 8. Exact only
 *      while every value is an integer fp16 holds exactly, so only for
 *      panels with longest reference x max(|match|, |mismatch|) + |open| +
 *      |extend| <= 2048; esch_set_variant refuses it (ESCH_INVALID_ARG)
 *      elsewhere;
 *   5  the fp16 arithmetic of variant 4 over eight columns per iteration,
 *      the stripe boundary loaded four columns ahead; the same condition;
 *   6  variant 3 with shifted values, which take one gap chain's extension add out:
 *      semiglobal mode shifts every value of stripe row r by |e| r (7 packed
 *      ops instead of 8; needs longest reference x M + |open| + 17 |extend|
 *      <= 32000, M = max(|match|, |mismatch|)); local mode shifts reference
 *      column j by |e| j (9 instead of 10; needs longest reference x (M +
 *      |extend|) + |open| + 3 |extend| + M <= 32000);
 *   7  variant 5 with those shifts: 6 packed ops
 *      instead of 7 in semiglobal mode, 6.5 instead of 8.5 in local mode
 *      (where the zero floor rides on the shifted chain's maximum);
 *      the same sums at most 2048;
 *   8  variant 4 with those shifts (fewer registers than 7, one more
 *      resident wave, a shallower prefetch); the condition of 7.
 *      In the shifted variants the two columns of a pair run one row apart,
 *      so their serial chains interleave.
 *   9-16  the same schedules on taller stripes (the tile shape): 9, 10, 11
 *      = 7, 8, 6 with 24 rows per stripe; 12, 13, 14 = 7, 8, 6 with 32 rows;
 *      15, 16 = 4, 5 with 24 rows. A stripe hands its bottom row to the next
 *      through device memory, so taller stripes move fewer boundary bytes
 *      and spread each stripe's fixed work over more rows, in more
 *      registers. The row-shifted ones need (rows + 1) |extend| in place of
 *      17 |extend| in the condition of 6 / 7.
 *   17-22  schedules 8, 7 and 6 with a windowed profile: each column reads
 *      its plane a quad of rows ahead of use rather than a column pair
 *      ahead, so the plane holds a few registers instead of one per row;
 *      17, 18 = 8 on 24 / 32 rows; 19, 20 = 7 on 24 / 32; 21, 22 = 6 (int16)
 *      on 24 / 32. The conditions of the schedules they window.
 *   23-28  schedules 12, 13 and 14: the windowed 8, 7 and 6 with every value
 *      of stripe row r in reference column j shifted by |e| (r + j), so
 *      that neither gap chain adds its extension: 5 packed ops per cell in
 *      fp16 (12, 13), 6 in int16 (14). Semiglobal mode only (local mode
 *      refuses them); they need longest reference x (M + |extend|) +
 *      |open| + (stripe rows + 2) |extend| + M within 2048 (fp16) or 32000
 *      (int16). 23, 24 = 12 on 24 / 32 rows; 25, 26 = 13; 27, 28 = 14.
 *   The column-shifted (local) and diagonal-shift variants also run where
 *   the longest reference is too long for those sums: their column shift
 *   then restarts every K columns (esch_column_block), and K |extend|
 *   takes the place of longest reference x |extend| in the condition, K at
 *   least 8.
 * A new scorer runs the first of 24, 28, 7, 4, 6 and 3 its mode and panel
 * admit (local mode: 7, 4, 6, 3). */
int esch_n_variants(void);
/* A variant's shape: its schedule (variants 0-8 are schedules 0-8 on 16
 * rows; 9, 10 and 11 are the windowed 8, 7 and 6; 12, 13 and 14 those with
 * the diagonal shift), rows per stripe, whether it computes in fp16 (the
 * 2048 condition) and how its values are shifted: 0 not at all, 1 along
 * one axis (rows in semiglobal mode, columns in local mode), 2 along both
 * (the diagonal shift). Returns ESCH_INVALID_ARG for a variant outside
 * 0 .. esch_n_variants() - 1; any output pointer may be NULL. */
esch_status esch_variant_info(int variant, int *schedule, int *stripe_rows, int *f16, int *shifted);
/* 1 when the scorer admits `variant` (its mode, and the panel within the
 * variant's exactness condition: esch_set_variant accepts it and the
 * autotuner times it), else 0. */
int esch_variant_usable(const esch_scorer *s, int variant);
/* The column block of `variant` on this scorer: the column-shifted (local)
 * and diagonal-shift (semiglobal) variants restart the column part of
 * their shift every that many columns where the panel's longest reference
 * does not fit the arithmetic's bound whole (a power of two from 8 to
 * 4096; each restart costs two packed adds per stripe row). 0: never
 * restarted, or a variant without a column shift; -1: a variant the
 * scorer refuses. */
int esch_column_block(const esch_scorer *s, int variant);
int esch_get_variant(const esch_scorer *s);
esch_status esch_set_variant(esch_scorer *s, int variant);
/* ms[v] (may be NULL, esch_n_variants() entries) receives each variant's DP
 * time on the sample (-1 for a variant the panel does not admit); the
 * fastest is selected, unless the scorer's current variant is within 1% of
 * it (which then stays), and returned in *chosen.
 * Then, unless esch_set_diag_ratio or esch_set_diag_plan fixed it, the
 * routing ratio is swept over {off, 1.5, 2, 3, 4} on the same sample
 * (another one replaces the default 2.0 only when more than 2% faster), and
 * then, for each length class the sample holds reads of, the class's plan
 * entry over {inter-sequence, anti-diagonal, 1.5, 2, 3, 4}, kept when more
 * than 2% faster (esch_get_diag_ratio / esch_get_diag_plan read the
 * result). The variants include the stripe heights (16, 24, 32 rows), so
 * the choice of variant is the tile shape's. */
esch_status esch_autotune(esch_scorer *s, const uint8_t *const *reads, const size_t *lens, size_t n,
                          int both_strands, float *ms, int *chosen);

/* ---------------------------------------------------------------------
 * Device-resident batches: inputs uploaded once, every kernel reading its
 * sizes from device memory, the launch sequence captured in a hipGraph and
 * replayed. Used by the serving loop and the benchmark (no counterpart in
 * the reference, whose batches always come from host memory).
 * --------------------------------------------------------------------- */

enum {
    ESCH_RUN_SCORES16 = 1,   /* int16 matrix, drop-in rows only (scores_len) */
    ESCH_RUN_SCORES32 = 2,   /* int32 matrix, every non-empty query */
    ESCH_RUN_SELECT = 4,     /* fused unpaired selection (needs SCORES32 rows) */
    ESCH_RUN_PAIRS = 8,      /* fused pair selection */
    ESCH_RUN_PLAIN = 16,     /* score with the plain one-thread-per-pair int32 kernel */
    ESCH_RUN_NO_GRAPH = 32,  /* launch eagerly instead of replaying a captured graph */
    ESCH_RUN_PAIR_BOTH = 64  /* pairs: consider the reverse template strand too */
};

esch_status esch_batch_create(esch_scorer *s, size_t max_queries, size_t max_bases,
                              esch_batch **out);
void esch_batch_free(esch_batch *b);
/* Device bytes held by a batch of this capacity for this scorer. */
size_t esch_batch_footprint(const esch_scorer *s, size_t max_queries, size_t max_bases);
/* Upload reads (codes). both_strands appends their reverse complements on
 * device; pairs = 1 treats reads as mate1, mate2 interleaved. */
esch_status esch_batch_load(esch_batch *b, const uint8_t *const *reads, const size_t *lens,
                            size_t n, int both_strands, int pairs);
/* Same, from one concatenated code buffer and per-read offsets. */
esch_status esch_batch_load_concat(esch_batch *b, const uint8_t *codes, const uint64_t *offs,
                                   const int32_t *lens, size_t n, int both_strands, int pairs);
/* Queue the device pipeline (async on the batch's own stream). */
esch_status esch_batch_run(esch_batch *b, int what, int32_t min_score, int64_t max_ties,
                           size_t tie_cap);
esch_status esch_batch_wait(esch_batch *b);
/* 1 when the batch's queued work has finished (non-blocking), else 0. */
int esch_batch_ready(esch_batch *b);
/* Milliseconds of the last run: the DP kernel alone (HIP events on its
 * stream) and the whole device pipeline. */
esch_status esch_batch_timing(const esch_batch *b, float *dp_ms, float *total_ms);
/* Cells of the last run (sum over scored queries of len x panel bases) and
 * how many queries went to the int16 and int32 kernels. */
esch_status esch_batch_stats(const esch_batch *b, uint64_t *cells, uint64_t *n16, uint64_t *n32);
/* How many queries of the last run took the anti-diagonal kernel. */
esch_status esch_batch_stats_diag(const esch_batch *b, uint64_t *n_diag);
esch_status esch_batch_read_scores32(esch_batch *b, int32_t *scores, uint8_t *scored);
esch_status esch_batch_read_scores16(esch_batch *b, int16_t *scores, uint8_t *scored);
esch_status esch_batch_read_select(esch_batch *b, esch_read_select *out, int32_t *ties);
/* 1 when the batch's last run scored with the roles swapped (esch_set_swap_rows). */
int esch_batch_swapped(const esch_batch *b);
/* The k_dp16 variant the batch's last run used (-1 before any run, or when
 * it ran with the roles swapped). It is the scorer's variant, except when
 * the caller fixed none (esch_set_variant; the autotuner's pick does not
 * count), the scorer's variant computes in int16, and the batch's reads
 * admit fp16: every value is bounded through min(read, reference), so a
 * batch whose longest read n keeps min(n, longest reference) x
 * max(|match|, |mismatch|) + |open| + |extend| within 2048 upgrades to
 * an fp16 form (semiglobal: the diagonal shift for a diagonal-shift pick,
 * else the row shift; local: the column shift), its column shift restarted
 * as often as the batch needs (esch_column_block); where the shifted form
 * does not fit, to the unshifted fp16 schedule (a diagonal-shift pick
 * stays). No counterpart in the reference. */
int esch_batch_variant(const esch_batch *b);
esch_status esch_batch_read_pairs(esch_batch *b, esch_pair_select *out, int32_t *top);

/* ---------------------------------------------------------------------
 * Serving: device-resident panels and a continuous-batching request loop
 * (no counterpart in the reference). Requests are cut into units of at
 * most one batch (max_batch_reads, max_batch_bases, max_batch_cells). A
 * scheduler thread packs queued units that share a panel and options into
 * device batches, the units of the request with the least queued work
 * first, and keeps up to max_slots batches in flight on their own streams.
 * Idle slots are kept for reuse and released, least recently used first,
 * only when a new slot would not fit hbm_budget. Every result equals the
 * request run alone.
 * --------------------------------------------------------------------- */
typedef struct esch_server esch_server;
typedef struct {
    int ordinal;             /* device */
    size_t hbm_budget;       /* bytes of device memory for batch slots */
    size_t max_batch_reads;  /* reads per device batch (before strands) */
    size_t max_batch_bases;  /* bases per device batch */
    int max_slots;           /* batches in flight at once */
    size_t max_batch_cells;  /* DP cells per device batch (read bases x strands x panel
                                bases; 0: no limit): a larger request is cut into batches
                                of at most this much work, so smaller requests run between
                                them (a single read over the limit is a batch of its own) */
} esch_server_config;
typedef struct {
    uint64_t completed;      /* requests finished */
    uint64_t batches;        /* device batches launched */
    size_t queued_units;     /* request parts waiting for a slot */
    size_t hbm_in_use;       /* bytes held by slots now */
    size_t hbm_peak;         /* most bytes ever held by slots */
    size_t slots;            /* slots allocated */
    size_t busy_slots;       /* slots with a batch in flight */
    uint64_t swapped_batches; /* batches scored with the roles swapped (esch_set_swap_rows) */
} esch_server_stats_t;

esch_status esch_server_new(const esch_server_config *cfg, esch_server **out);
void esch_server_free(esch_server *s);
/* Make a panel resident (a scorer under one scoring and mode); returns its id. */
esch_status esch_server_add_panel(esch_server *s, const uint8_t *const *refs, const size_t *ref_lens,
                                  size_t n_refs, int32_t match, int32_t mismatch, int32_t gap_open,
                                  int32_t gap_extend, esch_mode mode, int *panel_id);
/* The scorer of a resident panel (owned by the server), e.g. to autotune it. */
esch_scorer *esch_server_scorer(esch_server *s, int panel_id);
/* Queue reads (codes, copied) for fused selection against a panel; returns at once. */
esch_status esch_server_submit(esch_server *s, int panel_id, const uint8_t *const *reads,
                               const size_t *lens, size_t n, int both_strands, int32_t min_score,
                               int64_t max_ties, size_t tie_cap, uint64_t *request_id);
/* Block until the request finishes (timeout_ms < 0: no limit); ESCH_TIMEOUT if not yet. */
esch_status esch_server_wait(esch_server *s, uint64_t request_id, int64_t timeout_ms);
/* Copy a finished request's records (n) and ties (n * tie_cap) and its latency
 * from submit to completion; releases the request. */
esch_status esch_server_result(esch_server *s, uint64_t request_id, esch_read_select *out, int32_t *ties,
                               double *latency_s);
esch_status esch_server_stats(esch_server *s, esch_server_stats_t *out);

#ifdef __cplusplus
}
#endif
#endif
