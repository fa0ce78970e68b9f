/*
 * CPU oracle, scalar part: TEST INFRASTRUCTURE ONLY (see esch_oracle.h for
 * what pins each function). Plain C, integer arithmetic throughout.
 */
#include "esch_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* lib/alphabet.rs:29-44: A/C/G/T and U fold to 0..3, case-insensitive;
 * every other byte is the wildcard. */
static uint8_t encode_base(uint8_t b) {
    switch (b) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': case 'U': case 'u': return 3;
    default: return EO_WILDCARD;
    }
}

void eo_encode(const uint8_t *letters, size_t n, uint8_t *out) {
    for (size_t i = 0; i < n; i++) out[i] = encode_base(letters[i]);
}

/* lib/alphabet.rs:58-64 */
void eo_revcomp_codes(const uint8_t *codes, size_t n, uint8_t *out) {
    for (size_t i = 0; i < n; i++) {
        uint8_t c = codes[n - 1 - i];
        out[i] = c == EO_WILDCARD ? EO_WILDCARD : (uint8_t)(3 - c);
    }
}

static int32_t iabs32(int32_t v) { return v < 0 ? -v : v; }

/* lib/scoring.rs:70-100, rules in the reference's order. */
int eo_scoring_validate(int32_t ma, int32_t mi, int32_t o, int32_t e) {
    if (ma <= 0) return 1;
    if (mi >= ma) return 2;
    if (o > 0 || e > 0) return 3;
    const int32_t v[4] = {ma, mi, o, e};
    for (int k = 0; k < 4; k++)
        if (iabs32(v[k]) > 1000) return 4;
    return 0;
}

/* lib/scoring.rs:105-111 */
int32_t eo_substitution(int32_t ma, int32_t mi, uint8_t a, uint8_t b) {
    return (a == b || a == EO_WILDCARD || b == EO_WILDCARD) ? ma : mi;
}

static int32_t max32(int32_t a, int32_t b) { return a > b ? a : b; }

/* lib/scalar.rs:120-163: reference column outer, read row inner; hcol/ecol
 * hold column j-1's H and E per row; zero borders, E/F start at -inf. */
int32_t eo_score(const uint8_t *q, size_t n, const uint8_t *r, size_t m,
                 int32_t ma, int32_t mi, int32_t o, int32_t e, int local) {
    if (n == 0 || m == 0) return 0;
    int32_t *hcol = (int32_t *)calloc(n + 1, sizeof(int32_t));
    int32_t *ecol = (int32_t *)malloc((n + 1) * sizeof(int32_t));
    for (size_t i = 0; i <= n; i++) ecol[i] = EO_NEG;
    int32_t best = local ? 0 : EO_NEG;
    for (size_t j = 0; j < m; j++) {
        const uint8_t rj = r[j];
        int32_t hdiag = 0, hup = 0, f = EO_NEG, colmax = EO_NEG;
        for (size_t i = 1; i <= n; i++) {
            const int32_t hleft = hcol[i];
            const int32_t ev = max32(hleft + o, ecol[i] + e);
            ecol[i] = ev;
            f = max32(hup + o, f + e);
            int32_t h = max32(max32(hdiag + eo_substitution(ma, mi, q[i - 1], rj), ev), f);
            if (local) h = max32(h, 0);
            hdiag = hleft;
            hcol[i] = h;
            hup = h;
            colmax = max32(colmax, h);
        }
        if (local) {
            best = max32(best, colmax);
        } else {
            best = max32(best, hup);
            if (j + 1 == m) best = max32(best, colmax);
        }
    }
    free(hcol);
    free(ecol);
    return best;
}

/* lib/scalar.rs:164-174: the traceback byte of a cell. */
enum { EO_SRC_STOP = 0, EO_SRC_DIAG = 1, EO_SRC_INS = 2, EO_SRC_DEL = 3 };
#define EO_E_EXTENDED 4u
#define EO_F_EXTENDED 8u

/* lib/scalar.rs:176-364: fill (column-major trace, trace[j * rows + i]),
 * the end cell as the first maximum of a column-major scan over the mode's
 * candidates, then the walk from it. */
int eo_align(const uint8_t *q, size_t n, const uint8_t *r, size_t m, int32_t ma,
             int32_t mi, int32_t o, int32_t e, int local, eo_alignment *out,
             uint32_t *ops, size_t ops_cap) {
    memset(out, 0, sizeof(*out));
    if (n == 0 || m == 0) return 0;
    const size_t rows = n + 1;
    uint8_t *trace = (uint8_t *)calloc(rows * (m + 1), 1);
    int32_t *hcol = (int32_t *)calloc(rows, sizeof(int32_t));
    int32_t *ecol = (int32_t *)malloc(rows * sizeof(int32_t));
    for (size_t i = 0; i < rows; i++) ecol[i] = EO_NEG;
    int32_t best = local ? 0 : EO_NEG;
    size_t bi = 0, bj = 0;
    for (size_t j = 1; j <= m; j++) {
        const uint8_t rj = r[j - 1];
        uint8_t *col = trace + j * rows;
        int32_t hdiag = 0, hup = 0, f = EO_NEG;
        const int last_col = j == m;
        for (size_t i = 1; i <= n; i++) {
            const int32_t hleft = hcol[i];
            const int32_t e_ext = ecol[i] + e;
            const int32_t ev = max32(hleft + o, e_ext);
            const int32_t f_ext = f + e;
            f = max32(hup + o, f_ext);
            const int32_t hd = hdiag + eo_substitution(ma, mi, q[i - 1], rj);
            int32_t h = max32(max32(hd, ev), f);
            if (local) h = max32(h, 0);
            uint8_t t = h == f ? EO_SRC_INS : EO_SRC_DEL;
            if (h == hd) t = EO_SRC_DIAG;
            if (local && h == 0) t = EO_SRC_STOP;
            if (ev == e_ext) t |= EO_E_EXTENDED;
            if (f == f_ext) t |= EO_F_EXTENDED;
            col[i] = t;
            ecol[i] = ev;
            hdiag = hleft;
            hcol[i] = h;
            hup = h;
            if ((local || last_col) && h > best) {
                best = h;
                bi = i;
                bj = j;
            }
        }
        if (!local && !last_col && hup > best) {
            best = hup;
            bi = n;
            bj = j;
        }
    }
    free(hcol);
    free(ecol);
    const int nops = eo_walk_trace(trace, rows, 1, 0, best, bi, bj, out, ops, ops_cap);
    free(trace);
    return nops;
}

/* lib/scalar.rs:292-364 (walk): from end cell (bi, bj) of score `best`, the
 * runs of the alignment through the traceback bytes, cell (i, j) at
 * tb[(j * stride + i) * lanes + lane] (the scalar fill's one column per
 * stride rows, lanes = 1; the SIMD pair kernel's 32 lanes per cell). Runs
 * are collected last-first, then reversed; -1 when ops_cap is too small. */
int eo_walk_trace(const uint8_t *tb, size_t stride, size_t lanes, size_t lane, int32_t best, size_t bi,
                  size_t bj, eo_alignment *out, uint32_t *ops, size_t ops_cap) {
    memset(out, 0, sizeof(*out));
    int nops = 0;
    if (bi != 0 && bj != 0) {
        size_t i = bi, j = bj;
        int state = 0; /* 0 H, 1 F (insertion), 2 E (deletion) */
        uint32_t cur = 3u, len = 0;
        while (i > 0 && j > 0) {
            const uint8_t t = tb[(j * stride + i) * lanes + lane];
            uint32_t op = 3u;
            if (state == 0) {
                const int src = t & 3;
                if (src == EO_SRC_STOP) break;
                if (src == EO_SRC_DIAG) {
                    op = 0;
                    i--;
                    j--;
                } else if (src == EO_SRC_INS) {
                    state = 1;
                } else {
                    state = 2;
                }
            } else if (state == 1) {
                op = 1;
                if (!(t & EO_F_EXTENDED)) state = 0;
                i--;
            } else {
                op = 2;
                if (!(t & EO_E_EXTENDED)) state = 0;
                j--;
            }
            if (op == 3u) continue;
            if (op == cur) {
                len++;
            } else {
                if (len) {
                    if ((size_t)nops >= ops_cap) return -1;
                    ops[nops++] = cur << 30 | len;
                }
                cur = op;
                len = 1;
            }
        }
        if (len) {
            if ((size_t)nops >= ops_cap) return -1;
            ops[nops++] = cur << 30 | len;
        }
        for (int a = 0, b = nops - 1; a < b; a++, b--) {
            const uint32_t x = ops[a];
            ops[a] = ops[b];
            ops[b] = x;
        }
        out->score = best;
        out->query_start = (int32_t)i;
        out->query_end = (int32_t)bi;
        out->ref_start = (int32_t)j;
        out->ref_end = (int32_t)bj;
    }
    return nops;
}

/* eo_align over pairs (query k against reference refs[k]) on `threads`
 * workers, pairs dealt in interleaved order: recs[k * 6] = score,
 * query_start, query_end, ref_start, ref_end, runs (the CPU side of the
 * traceback measurement and its equality check). */
typedef struct {
    const uint8_t *qcodes, *rcodes;
    const uint64_t *qoff, *roff;
    const int32_t *qlen, *rlen, *refs;
    size_t n;
    int32_t ma, mi, o, e;
    int local, t, threads;
    int32_t *recs;
} align_job;

static void *align_worker(void *arg) {
    const align_job *j = (const align_job *)arg;
    size_t cap = 0;
    uint32_t *ops = NULL;
    for (size_t k = (size_t)j->t; k < j->n; k += (size_t)j->threads) {
        const int32_t r = j->refs[k];
        const size_t need = (size_t)j->qlen[k] + (size_t)j->rlen[r] + 1;
        if (need > cap) {
            free(ops);
            cap = need;
            ops = (uint32_t *)malloc(cap * sizeof(uint32_t));
        }
        eo_alignment a;
        const int runs = eo_align(j->qcodes + j->qoff[k], (size_t)j->qlen[k], j->rcodes + j->roff[r],
                                  (size_t)j->rlen[r], j->ma, j->mi, j->o, j->e, j->local, &a, ops, cap);
        int32_t *o = j->recs + k * 6;
        o[0] = a.score;
        o[1] = a.query_start;
        o[2] = a.query_end;
        o[3] = a.ref_start;
        o[4] = a.ref_end;
        o[5] = runs;
    }
    free(ops);
    return NULL;
}

void eo_align_batch(const uint8_t *qcodes, const uint64_t *qoff, const int32_t *qlen,
                    const uint8_t *rcodes, const uint64_t *roff, const int32_t *rlen,
                    const int32_t *refs, size_t n, int32_t ma, int32_t mi, int32_t o,
                    int32_t e, int local, int threads, int32_t *recs) {
    if (threads < 1) threads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    align_job *jobs = (align_job *)malloc(sizeof(align_job) * (size_t)threads);
    for (int t = 0; t < threads; t++) {
        align_job j = {qcodes, rcodes, qoff, roff, qlen, rlen, refs, n, ma, mi, o, e, local, t, threads, recs};
        jobs[t] = j;
        pthread_create(&th[t], NULL, align_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th);
    free(jobs);
}

/* lib/simd/mod.rs:222-227 */
int eo_fits_i16(size_t n, size_t max_ref_len, int32_t ma, int32_t mi,
                int32_t o, int32_t e) {
    const int64_t k = (int64_t)(n < max_ref_len ? n : max_ref_len);
    const int64_t mm = iabs32(ma) > iabs32(mi) ? iabs32(ma) : iabs32(mi);
    const int64_t bound = k * mm + iabs32(o) + iabs32(e);
    return bound < 15000 && max_ref_len < 15000;
}

typedef struct {
    const uint8_t *qcodes, *rcodes;
    const uint64_t *qoff, *roff;
    const int32_t *qlen, *rlen;
    size_t nq, nr;
    int32_t ma, mi, o, e;
    int local, tid, threads;
    int32_t *out;
} matrix_job;

static void *matrix_worker(void *arg) {
    matrix_job *j = (matrix_job *)arg;
    for (size_t qi = (size_t)j->tid; qi < j->nq; qi += (size_t)j->threads) {
        for (size_t ri = 0; ri < j->nr; ri++) {
            j->out[qi * j->nr + ri] =
                eo_score(j->qcodes + j->qoff[qi], (size_t)j->qlen[qi],
                         j->rcodes + j->roff[ri], (size_t)j->rlen[ri], j->ma,
                         j->mi, j->o, j->e, j->local);
        }
    }
    return NULL;
}

void eo_score_matrix(const uint8_t *qcodes, const uint64_t *qoff,
                     const int32_t *qlen, size_t nq, const uint8_t *rcodes,
                     const uint64_t *roff, const int32_t *rlen, size_t nr,
                     int32_t ma, int32_t mi, int32_t o, int32_t e, int local,
                     int threads, int32_t *out) {
    if (threads < 1) threads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    matrix_job *jobs = (matrix_job *)malloc(sizeof(matrix_job) * (size_t)threads);
    for (int t = 0; t < threads; t++) {
        matrix_job j = {qcodes, rcodes, qoff, roff, qlen, rlen, nq, nr,
                        ma, mi, o, e, local, t, threads, out};
        jobs[t] = j;
        pthread_create(&th[t], NULL, matrix_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th);
    free(jobs);
}

/* lib/mapper.rs:444-472. Candidate k of the tie order is ref k/2 on strand
 * k%2 when both strands are scored, else ref k forward. */
int32_t eo_select(const int32_t *fwd, const int32_t *rev, size_t n_refs,
                  int32_t min_score, int64_t max_ties, int32_t *best,
                  int32_t *n_tied, int32_t *has_subopt, int32_t *subopt,
                  int32_t *ties, size_t cap) {
    const size_t per = rev ? 2 : 1;
    const size_t nc = n_refs * per;
#define CAND(k) ((rev && ((k) & 1)) ? rev[(k) >> 1] : (rev ? fwd[(k) >> 1] : fwd[(k)]))
    int32_t b = CAND(0);
    for (size_t k = 1; k < nc; k++) b = max32(b, CAND(k));
    int32_t nt = 0, hs = 0, so = 0;
    for (size_t k = 0; k < nc; k++) {
        const int32_t s = CAND(k);
        if (s == b) {
            nt++;
        } else {
            so = hs ? max32(so, s) : s;
            hs = 1;
        }
    }
    *best = b;
    *n_tied = nt;
    *has_subopt = hs;
    *subopt = hs ? so : 0;
    if (b < min_score) return 0;
    int64_t take = nt;
    if (max_ties >= 0) {
        const int64_t lim = max_ties == INT64_MAX ? INT64_MAX : max_ties + 1;
        if (lim < take) take = lim;
    }
    int32_t taken = 0;
    for (size_t k = 0; k < nc && taken < take; k++) {
        if (CAND(k) != b) continue;
        if ((size_t)taken < cap) {
            const int32_t r = (int32_t)(rev ? (k >> 1) : k);
            const int32_t rv = rev ? (int32_t)(k & 1) : 0;
            ties[taken] = (r << 1) | rv;
        }
        taken++;
    }
#undef CAND
    return taken;
}

/* lib/mapper/pair.rs:135-137: mate m on template strand t is reversed
 * iff t ^ (m == 1). */
static int mate_reverse(int m, int t) { return t ^ (m == 1); }

typedef struct {
    int32_t sum, ref, rv, a, b;
} pcand;

static int pcand_cmp(const void *x, const void *y) {
    const pcand *a = (const pcand *)x, *b = (const pcand *)y;
    if (a->sum != b->sum) return a->sum > b->sum ? -1 : 1;
    if (a->ref != b->ref) return a->ref < b->ref ? -1 : 1;
    return a->rv - b->rv;
}

/* lib/mapper/pair.rs:185-231 */
int32_t eo_pair_cands(const int32_t *const rows[4], size_t n_refs,
                      int both_strands, int32_t min_score, int32_t *cand,
                      int32_t *sum, int32_t *s0, int32_t *s1) {
    if (!rows[0] || !rows[2]) return 0; /* an empty mate: no candidates */
    pcand *c = (pcand *)malloc(sizeof(pcand) * 2 * (n_refs ? n_refs : 1));
    int32_t nc = 0;
    for (int t = 0; t < (both_strands ? 2 : 1); t++) {
        const int32_t *r0 = rows[0 * 2 + mate_reverse(0, t)];
        const int32_t *r1 = rows[1 * 2 + mate_reverse(1, t)];
        for (size_t r = 0; r < n_refs; r++) {
            const int32_t a = r0[r], b = r1[r];
            if (a >= min_score && b >= min_score) {
                pcand p = {a + b, (int32_t)r, t, a, b};
                c[nc++] = p;
            }
        }
    }
    qsort(c, (size_t)nc, sizeof(pcand), pcand_cmp);
    for (int32_t k = 0; k < nc; k++) {
        cand[k] = (c[k].ref << 1) | c[k].rv;
        sum[k] = c[k].sum;
        s0[k] = c[k].a;
        s1[k] = c[k].b;
    }
    free(c);
    return nc;
}

/* lib/adapter.rs:68-91: leftmost prefix match of at least min_overlap
 * bases with at most overlap/10 mismatches (wildcards match). */
size_t eo_adapter_find(const uint8_t *read, size_t n, const uint8_t *adapter,
                       size_t alen, size_t min_overlap) {
    if (min_overlap < 1) min_overlap = 1;
    if (n < min_overlap || alen == 0) return 0;
    for (size_t p = 0; p + min_overlap <= n; p++) {
        const size_t overlap = alen < n - p ? alen : n - p;
        if (overlap < min_overlap) break;
        size_t mism = 0;
        for (size_t k = 0; k < overlap; k++) {
            const uint8_t a = encode_base(read[p + k]), b = encode_base(adapter[k]);
            if (a != b && a != EO_WILDCARD && b != EO_WILDCARD) mism++;
        }
        if (mism <= overlap / 10) return n - p;
    }
    return 0;
}

/* lib/adapter.rs:95-110 */
void eo_adapter_detect_pair(const uint8_t *r1, size_t l1, const uint8_t *r2,
                            size_t l2, const uint8_t *a1, size_t la1,
                            const uint8_t *a2, size_t la2, size_t min_overlap,
                            size_t out[2]) {
    size_t ad0 = eo_adapter_find(r1, l1, a1, la1, min_overlap);
    size_t ad1 = eo_adapter_find(r2, l2, a2, la2, min_overlap);
    if (ad0 > 0 && ad1 > 0 && l1 - ad0 != l2 - ad1) {
        const size_t e1 = ad0 < la1 ? ad0 : la1;
        const size_t e2 = ad1 < la2 ? ad1 : la2;
        if (e1 > e2) ad1 = 0;
        else if (e1 < e2) ad0 = 0;
        else ad0 = ad1 = 0;
    }
    out[0] = ad0;
    out[1] = ad1;
}
