/*
 * CPU oracle, SIMD part: TEST INFRASTRUCTURE AND CPU BASELINE ONLY.
 *
 * A C restatement of eschalign's inter-sequence i16 score kernels — one
 * SIMD lane per reference, 16 lanes (AVX2) or 32 (AVX-512BW), wrapping i16
 * adds with NEG16 = -16384 as -inf — and of the dispatch around them. It is
 * what the reference's CPU path computes, timed on the host cores as the
 * CPU baseline, and pinned to eo_score by equality in the tests.
 *
 *   profile (groups, prof, tprof, ends)  lib/simd/mod.rs:232-331, 255-269
 *   row-major kernel                     lib/simd/avx512.rs:43-107, avx2.rs:41-106
 *   transposed kernel                    lib/simd/avx512.rs:278-354, avx2.rs:290-359
 *   dispatch (kernel_for, score_all)     lib/mapper.rs:216-222, 242-252
 *   score_scalar fallback                lib/simd/mod.rs:665-675
 *   winner tracebacks (align_pairs,      lib/simd/mod.rs:384-499,
 *     the AVX-512BW pair kernel)           lib/simd/avx512.rs:120-247
 *
 * The pair kernel is pinned to eo_align (itself pinned to the parasail
 * goldens) pair for pair, runs included, in tests/test_oracle_align.py.
 */
#include "esch_oracle.h"

#include <immintrin.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define NEG16 ((int16_t)-16384)
#define N_CODES 5

typedef struct {
    size_t len;       /* longest reference in the group: columns walked */
    int16_t *lens;    /* per lane (0 = empty lane) */
    int64_t *refs;    /* per lane panel index (-1 = empty lane) */
    int16_t *prof;    /* prof[(j * 5 + c) * lanes + lane] */
    int16_t *tprof;   /* tprof[(c * len + j) * lanes + lane] */
    uint32_t *ends;   /* ends[j]: lane bits whose last column is j */
} eo_group;

struct eo_profile {
    int lanes;
    size_t n_groups, n_refs, max_len;
    eo_group *groups;
    int32_t ma, mi, o, e;
    /* the panel itself, for the scalar fallback */
    const uint8_t **rptr;
    int32_t *rlen;
};

int eo_simd_supported(void) {
    __builtin_cpu_init();
    int s = 0;
    if (__builtin_cpu_supports("avx2")) s |= 1;
    if (__builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw")) s |= 2;
    return s;
}

typedef struct {
    size_t idx, len;
} ord_t;

static int ord_cmp(const void *a, const void *b) {
    const ord_t *x = (const ord_t *)a, *y = (const ord_t *)b;
    if (x->len != y->len) return x->len > y->len ? -1 : 1; /* longest first */
    return x->idx < y->idx ? -1 : (x->idx > y->idx ? 1 : 0); /* stable */
}

/* lib/simd/mod.rs:297-330 (from_codes) and :255-269 (tprof). */
eo_profile *eo_profile_build(const uint8_t *rcodes, const uint64_t *roff,
                             const int32_t *rlen, size_t nr, int32_t ma,
                             int32_t mi, int32_t o, int32_t e, int lanes) {
    eo_profile *p = (eo_profile *)calloc(1, sizeof(eo_profile));
    p->lanes = lanes;
    p->n_refs = nr;
    p->ma = ma; p->mi = mi; p->o = o; p->e = e;
    p->rptr = (const uint8_t **)malloc(sizeof(uint8_t *) * (nr ? nr : 1));
    p->rlen = (int32_t *)malloc(sizeof(int32_t) * (nr ? nr : 1));
    ord_t *order = (ord_t *)malloc(sizeof(ord_t) * (nr ? nr : 1));
    for (size_t i = 0; i < nr; i++) {
        p->rptr[i] = rcodes + roff[i];
        p->rlen[i] = rlen[i];
        if ((size_t)rlen[i] > p->max_len) p->max_len = (size_t)rlen[i];
        order[i].idx = i;
        order[i].len = (size_t)rlen[i];
    }
    qsort(order, nr, sizeof(ord_t), ord_cmp);
    p->n_groups = (nr + (size_t)lanes - 1) / (size_t)lanes;
    p->groups = (eo_group *)calloc(p->n_groups ? p->n_groups : 1, sizeof(eo_group));
    for (size_t g = 0; g < p->n_groups; g++) {
        eo_group *G = &p->groups[g];
        const size_t first = g * (size_t)lanes;
        size_t cnt = nr - first < (size_t)lanes ? nr - first : (size_t)lanes;
        size_t len = 0;
        for (size_t k = 0; k < cnt; k++)
            if (order[first + k].len > len) len = order[first + k].len;
        G->len = len;
        G->lens = (int16_t *)calloc((size_t)lanes, sizeof(int16_t));
        G->refs = (int64_t *)malloc(sizeof(int64_t) * (size_t)lanes);
        for (int l = 0; l < lanes; l++) G->refs[l] = -1;
        const size_t psz = len * N_CODES * (size_t)lanes;
        G->prof = (int16_t *)malloc(sizeof(int16_t) * (psz ? psz : 1));
        for (size_t k = 0; k < psz; k++) G->prof[k] = (int16_t)mi; /* padding = mismatch */
        G->ends = (uint32_t *)calloc(len ? len : 1, sizeof(uint32_t));
        for (size_t lane = 0; lane < cnt; lane++) {
            const size_t ri = order[first + lane].idx;
            const uint8_t *codes = rcodes + roff[ri];
            const size_t L = (size_t)rlen[ri];
            G->lens[lane] = (int16_t)L;
            G->refs[lane] = (int64_t)ri;
            if (L > 0) G->ends[L - 1] |= 1u << lane;
            for (size_t j = 0; j < L; j++)
                for (int c = 0; c < N_CODES; c++)
                    G->prof[(j * N_CODES + (size_t)c) * (size_t)lanes + lane] =
                        (int16_t)eo_substitution(ma, mi, (uint8_t)c, codes[j]);
        }
        const size_t tsz = N_CODES * len * (size_t)lanes;
        G->tprof = (int16_t *)malloc(sizeof(int16_t) * (tsz ? tsz : 1));
        for (size_t k = 0; k < tsz; k++) G->tprof[k] = NEG16;
        for (int lane = 0; lane < lanes; lane++) {
            for (size_t j = 0; j < (size_t)G->lens[lane]; j++)
                for (int c = 0; c < N_CODES; c++)
                    G->tprof[((size_t)c * len + j) * (size_t)lanes + (size_t)lane] =
                        G->prof[(j * N_CODES + (size_t)c) * (size_t)lanes + (size_t)lane];
        }
    }
    free(order);
    return p;
}

void eo_profile_free(eo_profile *p) {
    if (!p) return;
    for (size_t g = 0; g < p->n_groups; g++) {
        free(p->groups[g].lens);
        free(p->groups[g].refs);
        free(p->groups[g].prof);
        free(p->groups[g].tprof);
        free(p->groups[g].ends);
    }
    free(p->groups);
    free(p->rptr);
    free(p->rlen);
    free(p);
}

/* ---------------- AVX-512BW, 32 lanes ---------------- */

/* lib/simd/avx512.rs:43-107 (row-major). */
__attribute__((target("avx512f,avx512bw"))) static void
rm512(const uint8_t *query, size_t n, const eo_group *g, int16_t vo_, int16_t ve_,
      int local, int16_t *hbuf, int16_t *ebuf, int16_t *best) {
    const __m512i zero = _mm512_setzero_si512();
    const __m512i neg = _mm512_set1_epi16(NEG16);
    const __m512i vo = _mm512_set1_epi16(vo_);
    const __m512i ve = _mm512_set1_epi16(ve_);
    __m512i *hp = (__m512i *)hbuf, *ep = (__m512i *)ebuf;
    for (size_t i = 0; i < n; i++) {
        _mm512_storeu_si512(hp + i, zero);
        _mm512_storeu_si512(ep + i, neg);
    }
    const __m512i lens = _mm512_loadu_si512((const void *)g->lens);
    __m512i acc = local ? zero : neg;
    for (size_t j = 0; j < g->len; j++) {
        const int16_t *pj = g->prof + j * N_CODES * 32;
        __m512i hdiag = zero, hup = zero, f = neg, colmax = neg;
        for (size_t i = 0; i < n; i++) {
            const __m512i s = _mm512_loadu_si512((const void *)(pj + (size_t)query[i] * 32));
            const __m512i hleft = _mm512_loadu_si512(hp + i);
            const __m512i e = _mm512_max_epi16(_mm512_add_epi16(hleft, vo),
                                               _mm512_add_epi16(_mm512_loadu_si512(ep + i), ve));
            _mm512_storeu_si512(ep + i, e);
            __m512i h = _mm512_max_epi16(_mm512_add_epi16(hdiag, s), e);
            hdiag = hleft;
            f = _mm512_max_epi16(_mm512_add_epi16(hup, vo), _mm512_add_epi16(f, ve));
            h = _mm512_max_epi16(h, f);
            if (local) h = _mm512_max_epi16(h, zero);
            _mm512_storeu_si512(hp + i, h);
            hup = h;
            colmax = _mm512_max_epi16(colmax, h);
        }
        const __mmask32 valid = _mm512_cmpgt_epi16_mask(lens, _mm512_set1_epi16((int16_t)j));
        if (local) {
            acc = _mm512_mask_max_epi16(acc, valid, acc, colmax);
        } else {
            acc = _mm512_mask_max_epi16(acc, valid, acc, hup);
            const __mmask32 last = _mm512_cmpeq_epi16_mask(lens, _mm512_set1_epi16((int16_t)(j + 1)));
            acc = _mm512_mask_max_epi16(acc, last, acc, colmax);
        }
    }
    _mm512_storeu_si512((void *)best, acc);
}

/* lib/simd/avx512.rs:278-354 (transposed). */
__attribute__((target("avx512f,avx512bw"))) static void
tr512(const uint8_t *query, size_t n, const eo_group *g, int16_t vo_, int16_t ve_,
      int local, int16_t *hbuf, int16_t *fbuf, int16_t *best) {
    const size_t l = g->len;
    const __m512i zero = _mm512_setzero_si512();
    const __m512i neg = _mm512_set1_epi16(NEG16);
    const __m512i vo = _mm512_set1_epi16(vo_);
    const __m512i ve = _mm512_set1_epi16(ve_);
    __m512i *hp = (__m512i *)hbuf, *fp = (__m512i *)fbuf;
    for (size_t j = 0; j < l; j++) {
        _mm512_storeu_si512(hp + j, zero);
        _mm512_storeu_si512(fp + j, neg);
    }
    __m512i acc = local ? zero : neg;
    for (size_t i = 0; i < n; i++) {
        const int16_t *row = g->tprof + (size_t)query[i] * l * 32;
        __m512i hdiag = zero, hleft = zero, e = neg;
        for (size_t j = 0; j < l; j++) {
            const __m512i s = _mm512_loadu_si512((const void *)(row + j * 32));
            const __m512i hup = _mm512_loadu_si512(hp + j);
            const __m512i f = _mm512_max_epi16(_mm512_add_epi16(hup, vo),
                                               _mm512_add_epi16(_mm512_loadu_si512(fp + j), ve));
            _mm512_storeu_si512(fp + j, f);
            e = _mm512_max_epi16(_mm512_add_epi16(hleft, vo), _mm512_add_epi16(e, ve));
            __m512i h = _mm512_max_epi16(_mm512_add_epi16(hdiag, s), e);
            hdiag = hup;
            h = _mm512_max_epi16(h, f);
            if (local) h = _mm512_max_epi16(h, zero);
            _mm512_storeu_si512(hp + j, h);
            hleft = h;
            if (local) {
                acc = _mm512_max_epi16(acc, h);
            } else {
                const uint32_t end = g->ends[j];
                if (end) acc = _mm512_mask_max_epi16(acc, (__mmask32)end, acc, h);
            }
        }
    }
    if (!local && n > 0) {
        const __m512i lens = _mm512_loadu_si512((const void *)g->lens);
        for (size_t j = 0; j < l; j++) {
            const __mmask32 valid = _mm512_cmpgt_epi16_mask(lens, _mm512_set1_epi16((int16_t)j));
            acc = _mm512_mask_max_epi16(acc, valid, acc, _mm512_loadu_si512(hp + j));
        }
    }
    _mm512_storeu_si512((void *)best, acc);
}

/* ---------------- AVX2, 16 lanes ---------------- */

/* lib/simd/avx2.rs:41-106 (row-major); masks are blend vectors. */
__attribute__((target("avx2"))) static void
rm256(const uint8_t *query, size_t n, const eo_group *g, int16_t vo_, int16_t ve_,
      int local, int16_t *hbuf, int16_t *ebuf, int16_t *best) {
    const __m256i zero = _mm256_setzero_si256();
    const __m256i neg = _mm256_set1_epi16(NEG16);
    const __m256i vo = _mm256_set1_epi16(vo_);
    const __m256i ve = _mm256_set1_epi16(ve_);
    __m256i *hp = (__m256i *)hbuf, *ep = (__m256i *)ebuf;
    for (size_t i = 0; i < n; i++) {
        _mm256_storeu_si256(hp + i, zero);
        _mm256_storeu_si256(ep + i, neg);
    }
    const __m256i lens = _mm256_loadu_si256((const __m256i *)g->lens);
    __m256i acc = local ? zero : neg;
    for (size_t j = 0; j < g->len; j++) {
        const int16_t *pj = g->prof + j * N_CODES * 16;
        __m256i hdiag = zero, hup = zero, f = neg, colmax = neg;
        for (size_t i = 0; i < n; i++) {
            const __m256i s = _mm256_loadu_si256((const __m256i *)(pj + (size_t)query[i] * 16));
            const __m256i hleft = _mm256_loadu_si256(hp + i);
            const __m256i e = _mm256_max_epi16(_mm256_add_epi16(hleft, vo),
                                               _mm256_add_epi16(_mm256_loadu_si256(ep + i), ve));
            _mm256_storeu_si256(ep + i, e);
            __m256i h = _mm256_max_epi16(_mm256_add_epi16(hdiag, s), e);
            hdiag = hleft;
            f = _mm256_max_epi16(_mm256_add_epi16(hup, vo), _mm256_add_epi16(f, ve));
            h = _mm256_max_epi16(h, f);
            if (local) h = _mm256_max_epi16(h, zero);
            _mm256_storeu_si256(hp + i, h);
            hup = h;
            colmax = _mm256_max_epi16(colmax, h);
        }
        const __m256i valid = _mm256_cmpgt_epi16(lens, _mm256_set1_epi16((int16_t)j));
        if (local) {
            acc = _mm256_blendv_epi8(acc, _mm256_max_epi16(acc, colmax), valid);
        } else {
            acc = _mm256_blendv_epi8(acc, _mm256_max_epi16(acc, hup), valid);
            const __m256i last = _mm256_cmpeq_epi16(lens, _mm256_set1_epi16((int16_t)(j + 1)));
            acc = _mm256_blendv_epi8(acc, _mm256_max_epi16(acc, colmax), last);
        }
    }
    _mm256_storeu_si256((__m256i *)best, acc);
}

/* lib/simd/avx2.rs:290-359 (transposed). */
__attribute__((target("avx2"))) static void
tr256(const uint8_t *query, size_t n, const eo_group *g, int16_t vo_, int16_t ve_,
      int local, int16_t *hbuf, int16_t *fbuf, int16_t *best) {
    const size_t l = g->len;
    const __m256i zero = _mm256_setzero_si256();
    const __m256i neg = _mm256_set1_epi16(NEG16);
    const __m256i vo = _mm256_set1_epi16(vo_);
    const __m256i ve = _mm256_set1_epi16(ve_);
    __m256i *hp = (__m256i *)hbuf, *fp = (__m256i *)fbuf;
    for (size_t j = 0; j < l; j++) {
        _mm256_storeu_si256(hp + j, zero);
        _mm256_storeu_si256(fp + j, neg);
    }
    __m256i acc = local ? zero : neg;
    for (size_t i = 0; i < n; i++) {
        const int16_t *row = g->tprof + (size_t)query[i] * l * 16;
        __m256i hdiag = zero, hleft = zero, e = neg;
        for (size_t j = 0; j < l; j++) {
            const __m256i s = _mm256_loadu_si256((const __m256i *)(row + j * 16));
            const __m256i hup = _mm256_loadu_si256(hp + j);
            const __m256i f = _mm256_max_epi16(_mm256_add_epi16(hup, vo),
                                               _mm256_add_epi16(_mm256_loadu_si256(fp + j), ve));
            _mm256_storeu_si256(fp + j, f);
            e = _mm256_max_epi16(_mm256_add_epi16(hleft, vo), _mm256_add_epi16(e, ve));
            __m256i h = _mm256_max_epi16(_mm256_add_epi16(hdiag, s), e);
            hdiag = hup;
            h = _mm256_max_epi16(h, f);
            if (local) h = _mm256_max_epi16(h, zero);
            _mm256_storeu_si256(hp + j, h);
            hleft = h;
            if (local) {
                acc = _mm256_max_epi16(acc, h);
            } else {
                const uint32_t end = g->ends[j];
                if (end) {
                    /* lanes whose bit is set in `end` */
                    const __m256i bits = _mm256_setr_epi16(1, 2, 4, 8, 16, 32, 64, 128, 256, 512,
                                                           1024, 2048, 4096, 8192, 16384,
                                                           (short)0x8000);
                    const __m256i sel = _mm256_cmpeq_epi16(
                        _mm256_and_si256(_mm256_set1_epi16((int16_t)end), bits), bits);
                    acc = _mm256_blendv_epi8(acc, _mm256_max_epi16(acc, h), sel);
                }
            }
        }
    }
    if (!local && n > 0) {
        const __m256i lens = _mm256_loadu_si256((const __m256i *)g->lens);
        for (size_t j = 0; j < l; j++) {
            const __m256i valid = _mm256_cmpgt_epi16(lens, _mm256_set1_epi16((int16_t)j));
            acc = _mm256_blendv_epi8(acc, _mm256_max_epi16(acc, _mm256_loadu_si256(hp + j)), valid);
        }
    }
    _mm256_storeu_si256((__m256i *)best, acc);
}

/* lib/mapper.rs:242-252 score_all + :216-222 kernel_for + lib/simd/mod.rs:
 * 507-569 score_profile. Per-thread scratch is allocated per call here. */
void eo_simd_score_all(const eo_profile *p, const uint8_t *q, size_t n,
                       int local, int transposed_min, int kernel, int32_t *out) {
    for (size_t r = 0; r < p->n_refs; r++) out[r] = 0;
    if (n == 0) return;
    const int simd_ok = p->max_len < 32000 &&
                        eo_fits_i16(n, p->max_len, p->ma, p->mi, p->o, p->e);
    if (!simd_ok) {
        for (size_t r = 0; r < p->n_refs; r++)
            out[r] = eo_score(q, n, p->rptr[r], (size_t)p->rlen[r], p->ma, p->mi,
                              p->o, p->e, local);
        return;
    }
    int transposed;
    if (kernel >= 0) transposed = kernel;
    else transposed = transposed_min > 0 && n >= (size_t)transposed_min;
    const int lanes = p->lanes;
    size_t need = transposed ? p->max_len : n;
    need = (need ? need : 1) * (size_t)lanes;
    /* per-thread H/E rows reused across reads, lib/simd/mod.rs:340-343 */
    static __thread int16_t *hbuf = NULL, *ebuf = NULL;
    static __thread size_t cap = 0;
    if (cap < need) {
        free(hbuf);
        free(ebuf);
        cap = need;
        hbuf = (int16_t *)aligned_alloc(64, ((cap * 2 + 63) / 64) * 64);
        ebuf = (int16_t *)aligned_alloc(64, ((cap * 2 + 63) / 64) * 64);
    }
    int16_t best[32];
    for (size_t g = 0; g < p->n_groups; g++) {
        const eo_group *G = &p->groups[g];
        if (lanes == 32) {
            if (transposed) tr512(q, n, G, (int16_t)p->o, (int16_t)p->e, local, hbuf, ebuf, best);
            else rm512(q, n, G, (int16_t)p->o, (int16_t)p->e, local, hbuf, ebuf, best);
        } else {
            if (transposed) tr256(q, n, G, (int16_t)p->o, (int16_t)p->e, local, hbuf, ebuf, best);
            else rm256(q, n, G, (int16_t)p->o, (int16_t)p->e, local, hbuf, ebuf, best);
        }
        for (int lane = 0; lane < lanes; lane++)
            if (G->refs[lane] >= 0) out[G->refs[lane]] = best[lane];
    }
}

typedef struct {
    const eo_profile *p;
    const uint8_t *qcodes;
    const uint64_t *qoff;
    const int32_t *qlen;
    size_t nq;
    int local;
    int32_t *out;
    size_t *next; /* shared chunk cursor */
    pthread_mutex_t *mu;
} batch_job;

static void *batch_worker(void *arg) {
    batch_job *j = (batch_job *)arg;
    for (;;) {
        pthread_mutex_lock(j->mu);
        const size_t start = *j->next;
        *j->next += 64; /* chunks of 64 reads, crates/eschalign-cli/src/align.rs:1470 */
        pthread_mutex_unlock(j->mu);
        if (start >= j->nq) break;
        const size_t end = start + 64 < j->nq ? start + 64 : j->nq;
        for (size_t qi = start; qi < end; qi++)
            eo_simd_score_all(j->p, j->qcodes + j->qoff[qi], (size_t)j->qlen[qi], j->local,
                              256, -1, j->out + qi * j->p->n_refs);
    }
    return NULL;
}

void eo_simd_score_batch(const eo_profile *p, const uint8_t *qcodes,
                         const uint64_t *qoff, const int32_t *qlen, size_t nq,
                         int local, int threads, int32_t *out) {
    if (threads < 1) threads = 1;
    size_t next = 0;
    pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    batch_job *jobs = (batch_job *)malloc(sizeof(batch_job) * (size_t)threads);
    for (int t = 0; t < threads; t++) {
        batch_job j = {p, qcodes, qoff, qlen, nq, local, out, &next, &mu};
        jobs[t] = j;
        pthread_create(&th[t], NULL, batch_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th);
    free(jobs);
}

/* ------------------------------------------------------------------------
 * Winner tracebacks: simd::align_pairs (lib/simd/mod.rs:384-499) with the
 * AVX-512BW pair kernel (lib/simd/avx512.rs:120-247). One (read, reference)
 * pair per i16 lane, 32 lanes: the scalar fill's recurrence with wrapping
 * i16 adds, its traceback byte per cell and lane, and each lane's end cell
 * in the scalar path's scan order; every lane is then walked by
 * eo_walk_trace, the scalar walk.
 * ------------------------------------------------------------------------ */

#define PAIR_LANES 32
#define MAX_TRACE_BYTES ((size_t)16 << 20)

__attribute__((target("avx512f,avx512bw"))) static void
pairs512(int local, size_t n, size_t m, const int16_t *qlens, const int16_t *rlens, const int16_t *qt,
         const int16_t *rt, int16_t ma, int16_t mi, int16_t go, int16_t ge, int16_t *hbuf, int16_t *ebuf,
         uint8_t *tb, int16_t *best_out, int16_t *bi_out, int16_t *bj_out) {
    const size_t stride = n + 1;
    const __m512i zero = _mm512_setzero_si512(), neg = _mm512_set1_epi16(NEG16);
    const __m512i wild = _mm512_set1_epi16(4), vm = _mm512_set1_epi16(ma), vx = _mm512_set1_epi16(mi);
    const __m512i vo = _mm512_set1_epi16(go), ve = _mm512_set1_epi16(ge);
    const __m512i t_diag = _mm512_set1_epi16(1), t_ins = _mm512_set1_epi16(2), t_del = _mm512_set1_epi16(3);
    const __m512i t_eext = _mm512_set1_epi16(4), t_fext = _mm512_set1_epi16(8);
    for (size_t i = 0; i < n; i++) {
        _mm512_storeu_si512((void *)(hbuf + i * PAIR_LANES), zero);
        _mm512_storeu_si512((void *)(ebuf + i * PAIR_LANES), neg);
    }
    const __m512i ql = _mm512_loadu_si512((const void *)qlens), rl = _mm512_loadu_si512((const void *)rlens);
    __m512i best = local ? zero : neg, bi = zero, bj = zero;
    for (size_t j = 0; j < m; j++) {
        const __m512i rv = _mm512_loadu_si512((const void *)(rt + j * PAIR_LANES));
        const __mmask32 rwild = _mm512_cmpeq_epi16_mask(rv, wild);
        const __m512i jv = _mm512_set1_epi16((int16_t)(j + 1));
        const __mmask32 jvalid = _mm512_cmpgt_epi16_mask(rl, _mm512_set1_epi16((int16_t)j));
        const __mmask32 jlast = _mm512_cmpeq_epi16_mask(rl, jv);
        uint8_t *tcol = tb + (j + 1) * stride * PAIR_LANES;
        __m512i hdiag = zero, hup = zero, f = neg;
        for (size_t i = 0; i < n; i++) {
            const __m512i qv = _mm512_loadu_si512((const void *)(qt + i * PAIR_LANES));
            const __m512i i0 = _mm512_set1_epi16((int16_t)i), iv = _mm512_set1_epi16((int16_t)(i + 1));
            const __mmask32 same = _mm512_cmpeq_epi16_mask(qv, rv) | rwild | _mm512_cmpeq_epi16_mask(qv, wild);
            const __m512i s = _mm512_mask_blend_epi16(same, vx, vm);
            const __m512i hleft = _mm512_loadu_si512((const void *)(hbuf + i * PAIR_LANES));
            const __m512i e_ext = _mm512_add_epi16(_mm512_loadu_si512((const void *)(ebuf + i * PAIR_LANES)), ve);
            const __m512i ev = _mm512_max_epi16(_mm512_add_epi16(hleft, vo), e_ext);
            _mm512_storeu_si512((void *)(ebuf + i * PAIR_LANES), ev);
            const __m512i f_ext = _mm512_add_epi16(f, ve);
            f = _mm512_max_epi16(_mm512_add_epi16(hup, vo), f_ext);
            const __m512i hd = _mm512_add_epi16(hdiag, s);
            hdiag = hleft;
            __m512i h = _mm512_max_epi16(_mm512_max_epi16(hd, ev), f);
            if (local) h = _mm512_max_epi16(h, zero);
            /* the source, in the scalar path's order of preference */
            __m512i t = t_del;
            t = _mm512_mask_mov_epi16(t, _mm512_cmpeq_epi16_mask(h, f), t_ins);
            t = _mm512_mask_mov_epi16(t, _mm512_cmpeq_epi16_mask(h, hd), t_diag);
            if (local) t = _mm512_mask_mov_epi16(t, _mm512_cmpeq_epi16_mask(h, zero), zero);
            t = _mm512_mask_add_epi16(t, _mm512_cmpeq_epi16_mask(ev, e_ext), t, t_eext);
            t = _mm512_mask_add_epi16(t, _mm512_cmpeq_epi16_mask(f, f_ext), t, t_fext);
            _mm256_storeu_si256((__m256i *)(tcol + (i + 1) * PAIR_LANES), _mm512_cvtepi16_epi8(t));
            _mm512_storeu_si512((void *)(hbuf + i * PAIR_LANES), h);
            hup = h;
            /* candidate end cells, in the scalar scan order */
            const __mmask32 ivalid = _mm512_cmpgt_epi16_mask(ql, i0);
            const __mmask32 cand = local ? (ivalid & jvalid)
                                         : ((_mm512_cmpeq_epi16_mask(ql, iv) & jvalid) | (jlast & ivalid));
            const __mmask32 up = _mm512_cmpgt_epi16_mask(h, best) & cand;
            best = _mm512_mask_mov_epi16(best, up, h);
            bi = _mm512_mask_mov_epi16(bi, up, iv);
            bj = _mm512_mask_mov_epi16(bj, up, jv);
        }
    }
    _mm512_storeu_si512((void *)best_out, best);
    _mm512_storeu_si512((void *)bi_out, bi);
    _mm512_storeu_si512((void *)bj_out, bj);
}

typedef struct {
    size_t k, ql, rl;
} pair_ord;

static int pair_ord_cmp(const void *a, const void *b) {
    const pair_ord *x = (const pair_ord *)a, *y = (const pair_ord *)b;
    if (x->ql != y->ql) return x->ql < y->ql ? -1 : 1;
    if (x->rl != y->rl) return x->rl < y->rl ? -1 : 1;
    return x->k < y->k ? -1 : (x->k > y->k ? 1 : 0); /* stable */
}

/* A worker's buffers, grown as needed and kept between calls. */
struct eo_pair_scratch {
    int16_t *hbuf, *ebuf, *qt, *rt;
    uint8_t *tb;
    pair_ord *ord;
    size_t cap_h, cap_tb, cap_q, cap_r, cap_ord;
    int simd;
};

eo_pair_scratch *eo_pair_scratch_new(void) {
    eo_pair_scratch *s = (eo_pair_scratch *)calloc(1, sizeof(eo_pair_scratch));
    s->simd = (eo_simd_supported() & 2) != 0;
    return s;
}

void eo_pair_scratch_free(eo_pair_scratch *s) {
    if (!s) return;
    free(s->hbuf);
    free(s->ebuf);
    free(s->qt);
    free(s->rt);
    free(s->tb);
    free(s->ord);
    free(s);
}

static void *grow(void *p, size_t *cap, size_t need, size_t elem) {
    if (need <= *cap) return p;
    free(p);
    *cap = need;
    p = malloc(need * elem);
    if (!p) abort();
    return p;
}

/* simd::align_pairs (lib/simd/mod.rs:405-499) over one call's pairs, on the
 * calling thread: pair k is q[k] (ql[k] codes) against r[k] (rl[k]); its
 * alignment goes to al[k], its runs to ops + ops_off[k] (ql + rl + 1 slots)
 * and their count to runs[k]. The SIMD-eligible pairs (non-empty, both
 * under 32,000, fits_i16) are sorted by (read length, reference length) and
 * cut into groups of 32 lanes; the others, groups over 16 MiB of
 * traceback, and every pair without AVX-512BW take eo_align. */
void eo_simd_align_chunk(eo_pair_scratch *s, const uint8_t *const *q, const size_t *ql, const uint8_t *const *r,
                         const size_t *rl, size_t cnt, int32_t ma, int32_t mi, int32_t o, int32_t e, int local,
                         eo_alignment *al, int *runs, uint32_t *ops, const size_t *ops_off) {
    s->ord = (pair_ord *)grow(s->ord, &s->cap_ord, cnt ? cnt : 1, sizeof(pair_ord));
    size_t no = 0;
    for (size_t k = 0; k < cnt; k++) {
        const int ok = s->simd && ql[k] > 0 && rl[k] > 0 && ql[k] < 32000 && rl[k] < 32000 &&
                       eo_fits_i16(ql[k], rl[k], ma, mi, o, e);
        if (ok) {
            pair_ord x = {k, ql[k], rl[k]};
            s->ord[no++] = x;
        } else {
            runs[k] = eo_align(q[k], ql[k], r[k], rl[k], ma, mi, o, e, local, &al[k], ops + ops_off[k],
                               ql[k] + rl[k] + 1);
        }
    }
    qsort(s->ord, no, sizeof(pair_ord), pair_ord_cmp);
    for (size_t g0 = 0; g0 < no; g0 += PAIR_LANES) {
        const size_t gn = no - g0 < PAIR_LANES ? no - g0 : PAIR_LANES;
        const pair_ord *g = s->ord + g0;
        size_t rows = 0, cols = 0;
        for (size_t l = 0; l < gn; l++) {
            rows = g[l].ql > rows ? g[l].ql : rows;
            cols = g[l].rl > cols ? g[l].rl : cols;
        }
        const size_t need = (rows + 1) * (cols + 1) * PAIR_LANES;
        if (need > MAX_TRACE_BYTES) {
            for (size_t l = 0; l < gn; l++) {
                const size_t k = g[l].k;
                runs[k] = eo_align(q[k], ql[k], r[k], rl[k], ma, mi, o, e, local, &al[k], ops + ops_off[k],
                                   ql[k] + rl[k] + 1);
            }
            continue;
        }
        size_t ch = s->cap_h;
        s->hbuf = (int16_t *)grow(s->hbuf, &s->cap_h, rows * PAIR_LANES, sizeof(int16_t));
        if (s->cap_h != ch) {
            free(s->ebuf);
            s->ebuf = (int16_t *)malloc(s->cap_h * sizeof(int16_t));
        }
        s->tb = (uint8_t *)grow(s->tb, &s->cap_tb, need, 1);
        s->qt = (int16_t *)grow(s->qt, &s->cap_q, rows * PAIR_LANES, sizeof(int16_t));
        s->rt = (int16_t *)grow(s->rt, &s->cap_r, cols * PAIR_LANES, sizeof(int16_t));
        memset(s->qt, 0, rows * PAIR_LANES * sizeof(int16_t));
        memset(s->rt, 0, cols * PAIR_LANES * sizeof(int16_t));
        int16_t qlens[PAIR_LANES] = {0}, rlens[PAIR_LANES] = {0};
        for (size_t l = 0; l < gn; l++) {
            const size_t k = g[l].k;
            qlens[l] = (int16_t)ql[k];
            rlens[l] = (int16_t)rl[k];
            for (size_t i = 0; i < ql[k]; i++) s->qt[i * PAIR_LANES + l] = q[k][i];
            for (size_t c = 0; c < rl[k]; c++) s->rt[c * PAIR_LANES + l] = r[k][c];
        }
        int16_t best[PAIR_LANES], bi[PAIR_LANES], bj[PAIR_LANES];
        pairs512(local, rows, cols, qlens, rlens, s->qt, s->rt, (int16_t)ma, (int16_t)mi, (int16_t)o, (int16_t)e,
                 s->hbuf, s->ebuf, s->tb, best, bi, bj);
        for (size_t l = 0; l < gn; l++) {
            const size_t k = g[l].k;
            runs[k] = eo_walk_trace(s->tb, rows + 1, PAIR_LANES, l, best[l], (size_t)bi[l], (size_t)bj[l], &al[k],
                                    ops + ops_off[k], ql[k] + rl[k] + 1);
        }
    }
}

typedef struct {
    const uint8_t *qcodes, *rcodes;
    const uint64_t *qoff, *roff, *ops_off;
    const int32_t *qlen, *rlen, *refs;
    size_t n;
    int32_t ma, mi, o, e;
    int local;
    int32_t *recs;
    uint32_t *ops;
    size_t *next;
    pthread_mutex_t *mu;
} pair_job;

static void *pair_worker(void *arg) {
    const pair_job *j = (const pair_job *)arg;
    eo_pair_scratch *s = eo_pair_scratch_new();
    const uint8_t *q[64], *r[64];
    size_t ql[64], rl[64], off[64];
    eo_alignment al[64];
    int runs[64];
    uint32_t *ops = NULL;
    size_t cap_ops = 0;
    for (;;) {
        pthread_mutex_lock(j->mu);
        const size_t start = *j->next;
        *j->next += 64; /* a worker's chunk: 64 reads' winners (crates/eschalign-cli/src/align.rs:1470) */
        pthread_mutex_unlock(j->mu);
        if (start >= j->n) break;
        const size_t end = start + 64 < j->n ? start + 64 : j->n, cnt = end - start;
        size_t room = 0;
        for (size_t t = 0; t < cnt; t++) {
            const size_t k = start + t;
            q[t] = j->qcodes + j->qoff[k];
            ql[t] = (size_t)j->qlen[k];
            r[t] = j->rcodes + j->roff[j->refs[k]];
            rl[t] = (size_t)j->rlen[j->refs[k]];
            off[t] = room;
            room += ql[t] + rl[t] + 1;
        }
        ops = (uint32_t *)grow(ops, &cap_ops, room, sizeof(uint32_t));
        eo_simd_align_chunk(s, q, ql, r, rl, cnt, j->ma, j->mi, j->o, j->e, j->local, al, runs, ops, off);
        for (size_t t = 0; t < cnt; t++) {
            const size_t k = start + t;
            int32_t *o = j->recs + k * 6;
            o[0] = al[t].score;
            o[1] = al[t].query_start;
            o[2] = al[t].query_end;
            o[3] = al[t].ref_start;
            o[4] = al[t].ref_end;
            o[5] = runs[t];
            if (j->ops && runs[t] > 0) memcpy(j->ops + j->ops_off[k], ops + off[t], (size_t)runs[t] * sizeof(uint32_t));
        }
    }
    free(ops);
    eo_pair_scratch_free(s);
    return NULL;
}

void eo_simd_align_pairs(const uint8_t *qcodes, const uint64_t *qoff, const int32_t *qlen,
                         const uint8_t *rcodes, const uint64_t *roff, const int32_t *rlen,
                         const int32_t *refs, size_t n, int32_t ma, int32_t mi, int32_t o, int32_t e,
                         int local, int threads, int32_t *recs, uint32_t *ops, const uint64_t *ops_off) {
    if (threads < 1) threads = 1;
    size_t next = 0;
    pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    pair_job *jobs = (pair_job *)malloc(sizeof(pair_job) * (size_t)threads);
    for (int t = 0; t < threads; t++) {
        pair_job j = {qcodes, rcodes, qoff, roff, ops_off, qlen, rlen, refs, n, ma, mi, o, e, local,
                      recs, ops, &next, &mu};
        jobs[t] = j;
        pthread_create(&th[t], NULL, pair_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th);
    free(jobs);
}
