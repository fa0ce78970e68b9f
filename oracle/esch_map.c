/*
 * CPU oracle of read and pair mapping: TEST INFRASTRUCTURE ONLY (see
 * esch_oracle.h). Restates
 *   Aligner::map_reads_scored   lib/mapper.rs:400-550
 *   Aligner::map_pairs_scored   lib/mapper/pair.rs:157-387
 *   pair_ok / span              lib/mapper/pair.rs:392-420
 * on panel scores the caller computed (the `prescored` rows of both), with
 * eo_select / eo_pair_cands for the selection and eo_align for every
 * traceback. MD / NM strings are not part of the records; eo_md_nm and
 * eo_cigar_string (lib/sam.rs:32-124) at the end of this file give them.
 *
 * Pairs are independent of one another (lib/mapper/pair.rs:558-573 checks
 * that batching does not change a result), so each is mapped on its own,
 * spread over POSIX threads, and the results are concatenated in order.
 */
#include "esch_oracle.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    eo_hit *hits;
    size_t n_hits, cap_hits;
    uint32_t *ops;
    size_t n_ops, cap_ops;
} acc;

struct eo_mapped {
    size_t n;
    eo_read_mapping *reads; /* map_reads */
    eo_pair_mapping *pairs; /* map_pairs */
    eo_hit *hits;
    size_t n_hits;
    uint32_t *ops;
    size_t n_ops;
};

typedef struct {
    const uint8_t *rcodes;
    const uint64_t *roff;
    const int32_t *rlen;
    size_t n_refs;
    int32_t ma, mi, o, e;
    int local;
    int32_t min_score;
    int64_t max_ties;
} ctx;

static void *xrealloc(void *p, size_t bytes) {
    void *q = realloc(p, bytes ? bytes : 1);
    if (!q) abort();
    return q;
}

static eo_hit *acc_hit(acc *a) {
    if (a->n_hits == a->cap_hits) {
        a->cap_hits = a->cap_hits ? 2 * a->cap_hits : 8;
        a->hits = (eo_hit *)xrealloc(a->hits, a->cap_hits * sizeof(eo_hit));
    }
    eo_hit *h = &a->hits[a->n_hits++];
    memset(h, 0, sizeof *h);
    return h;
}

static void acc_ops(acc *a, const uint32_t *ops, size_t n) {
    if (a->n_ops + n > a->cap_ops) {
        while (a->n_ops + n > a->cap_ops) a->cap_ops = a->cap_ops ? 2 * a->cap_ops : 64;
        a->ops = (uint32_t *)xrealloc(a->ops, a->cap_ops * sizeof(uint32_t));
    }
    memcpy(a->ops + a->n_ops, ops, n * sizeof(uint32_t));
    a->n_ops += n;
}

/* One traced alignment with its runs (scalar::align, lib/scalar.rs:182-364). */
typedef struct {
    eo_alignment al;
    int n_ops;
    uint32_t *ops;
} traced;

static void trace(const ctx *c, const uint8_t *q, size_t n, int32_t ref, traced *t) {
    const uint8_t *r = c->rcodes + c->roff[ref];
    const size_t m = (size_t)c->rlen[ref];
    const size_t cap = n + m + 1;
    t->ops = (uint32_t *)malloc(cap * sizeof(uint32_t));
    t->n_ops = eo_align(q, n, r, m, c->ma, c->mi, c->o, c->e, c->local, &t->al, t->ops, cap);
    if (t->n_ops < 0) abort(); /* n + m runs always suffice */
}

/* Alignment::is_empty (lib/scalar.rs:83-85): no read base aligned to a
 * reference base, i.e. no M run. */
static int is_empty(const traced *t) {
    for (int k = 0; k < t->n_ops; k++)
        if ((t->ops[k] >> 30) == 0) return 0;
    return 1;
}

static void put_hit(acc *a, const traced *t, int32_t ref, int32_t reverse, int32_t template_reverse,
                    int32_t insert) {
    eo_hit *h = acc_hit(a);
    h->reference = ref;
    h->reverse = reverse;
    h->template_reverse = template_reverse;
    h->insert = insert;
    h->score = t->al.score;
    h->query_start = t->al.query_start;
    h->query_end = t->al.query_end;
    h->ref_start = t->al.ref_start;
    h->ref_end = t->al.ref_end;
    h->n_ops = t->n_ops;
    h->ops_offset = (int64_t)a->n_ops;
    acc_ops(a, t->ops, (size_t)t->n_ops);
}

/* One read of map_reads_scored (lib/mapper.rs:420-550): selection over
 * (ref r forward, ref r reverse, r + 1, ...), the taken ties traced, and
 * the read unmapped (no hits) when nothing is taken or the primary traces
 * to nothing. rev = NULL: forward strand only. */
static void map_read(const ctx *c, const uint8_t *q, size_t n, const int32_t *fwd, const int32_t *rev,
                     acc *a, eo_read_mapping *out) {
    memset(out, 0, sizeof *out);
    out->hit_offset = (int64_t)a->n_hits;
    if (n == 0) return; /* best_score None (lib/mapper.rs:425-434) */
    int32_t *ties = (int32_t *)malloc(2 * c->n_refs * sizeof(int32_t));
    int32_t best, nt, hs, so;
    const int32_t take = eo_select(fwd, rev, c->n_refs, c->min_score, c->max_ties, &best, &nt, &hs, &so,
                                   ties, 2 * c->n_refs);
    out->best = best;
    out->n_tied = nt;
    out->subopt = hs ? so : 0;
    out->flags = 1 | (hs ? 2 : 0);
    if (take > 0) {
        uint8_t *rc = (uint8_t *)malloc(n);
        eo_revcomp_codes(q, n, rc);
        traced *t = (traced *)calloc((size_t)take, sizeof(traced));
        for (int32_t k = 0; k < take; k++) trace(c, (ties[k] & 1) ? rc : q, n, ties[k] >> 1, &t[k]);
        if (!is_empty(&t[0])) {
            for (int32_t k = 0; k < take; k++) put_hit(a, &t[k], ties[k] >> 1, ties[k] & 1, 0, 0);
            out->n_hits = take;
        }
        for (int32_t k = 0; k < take; k++) free(t[k].ops);
        free(t);
        free(rc);
    }
    free(ties);
}

/* map_reads_scored over a chunk of reads (lib/mapper.rs:400-553), as the
 * CLI's workers run it on 64 reads at a time: every read's selection, then
 * every taken tie of the chunk traced in one simd::align_pairs call (the
 * AVX-512BW pair kernel, eo_simd_align_chunk), then the hits. Equal to
 * map_read per read: the pair kernel equals eo_align pair for pair. */
static void map_read_chunk(const ctx *c, eo_pair_scratch *ps, const uint8_t *const *qs, const size_t *ns,
                           const int32_t *const *fwd, const int32_t *const *rev, size_t cnt, acc *a,
                           eo_read_mapping *out) {
    const size_t nr = c->n_refs;
    int32_t *ties = (int32_t *)malloc(cnt * 2 * nr * sizeof(int32_t) + 1);
    int32_t *take = (int32_t *)calloc(cnt + 1, sizeof(int32_t));
    uint8_t **rcs = (uint8_t **)calloc(cnt + 1, sizeof(uint8_t *));
    size_t n_pairs = 0;
    for (size_t i = 0; i < cnt; i++) {
        memset(&out[i], 0, sizeof out[i]);
        if (ns[i] == 0) continue; /* best_score None (lib/mapper.rs:425-434) */
        int32_t best, nt, hs, so;
        take[i] = eo_select(fwd[i], rev[i], nr, c->min_score, c->max_ties, &best, &nt, &hs, &so, ties + i * 2 * nr,
                            2 * nr);
        out[i].best = best;
        out[i].n_tied = nt;
        out[i].subopt = hs ? so : 0;
        out[i].flags = 1 | (hs ? 2 : 0);
        if (take[i] > 0) {
            rcs[i] = (uint8_t *)malloc(ns[i]);
            eo_revcomp_codes(qs[i], ns[i], rcs[i]);
            n_pairs += (size_t)take[i];
        }
    }
    const uint8_t **pq = (const uint8_t **)malloc((n_pairs + 1) * sizeof(uint8_t *));
    const uint8_t **pr = (const uint8_t **)malloc((n_pairs + 1) * sizeof(uint8_t *));
    size_t *pql = (size_t *)malloc((n_pairs + 1) * sizeof(size_t)), *prl = (size_t *)malloc((n_pairs + 1) * sizeof(size_t));
    size_t *poff = (size_t *)malloc((n_pairs + 1) * sizeof(size_t));
    eo_alignment *al = (eo_alignment *)malloc((n_pairs + 1) * sizeof(eo_alignment));
    int *runs = (int *)malloc((n_pairs + 1) * sizeof(int));
    size_t k = 0, room = 0;
    for (size_t i = 0; i < cnt; i++)
        for (int32_t t = 0; t < take[i]; t++, k++) {
            const int32_t tie = ties[i * 2 * nr + (size_t)t], ref = tie >> 1;
            pq[k] = (tie & 1) ? rcs[i] : qs[i];
            pql[k] = ns[i];
            pr[k] = c->rcodes + c->roff[ref];
            prl[k] = (size_t)c->rlen[ref];
            poff[k] = room;
            room += pql[k] + prl[k] + 1;
        }
    uint32_t *ops = (uint32_t *)malloc((room + 1) * sizeof(uint32_t));
    eo_simd_align_chunk(ps, pq, pql, pr, prl, n_pairs, c->ma, c->mi, c->o, c->e, c->local, al, runs, ops, poff);
    k = 0;
    for (size_t i = 0; i < cnt; i++) {
        out[i].hit_offset = (int64_t)a->n_hits;
        if (take[i] <= 0) continue;
        traced t0 = {al[k], runs[k], ops + poff[k]};
        if (!is_empty(&t0)) {  /* the primary traces to something (lib/mapper.rs:508-512) */
            for (int32_t t = 0; t < take[i]; t++) {
                const int32_t tie = ties[i * 2 * nr + (size_t)t];
                traced tt = {al[k + (size_t)t], runs[k + (size_t)t], ops + poff[k + (size_t)t]};
                put_hit(a, &tt, tie >> 1, tie & 1, 0, 0);
            }
            out[i].n_hits = take[i];
        }
        k += (size_t)take[i];
    }
    for (size_t i = 0; i < cnt; i++) free(rcs[i]);
    free(rcs);
    free(take);
    free(ties);
    free(pq);
    free(pr);
    free(pql);
    free(prl);
    free(poff);
    free(al);
    free(runs);
    free(ops);
}

static int mate_reverse(int m, int t) { return t ^ (m == 1); } /* lib/mapper/pair.rs:135-137 */

/* pair_ok (lib/mapper/pair.rs:392-410) on two traced mates. */
static int pair_ok(const traced al[2], int template_reverse, int64_t max_insert) {
    if (is_empty(&al[0]) || is_empty(&al[1])) return 0;
    const eo_alignment *fw = template_reverse ? &al[1].al : &al[0].al;
    const eo_alignment *rv = template_reverse ? &al[0].al : &al[1].al;
    if (fw->ref_start >= rv->ref_end) return 0;
    const int32_t lo = al[0].al.ref_start < al[1].al.ref_start ? al[0].al.ref_start : al[1].al.ref_start;
    const int32_t hi = al[0].al.ref_end > al[1].al.ref_end ? al[0].al.ref_end : al[1].al.ref_end;
    return (int64_t)(hi - lo) <= max_insert;
}

/* span (lib/mapper/pair.rs:413-420) */
static int32_t span(const traced al[2]) {
    const int32_t lo = al[0].al.ref_start < al[1].al.ref_start ? al[0].al.ref_start : al[1].al.ref_start;
    const int32_t hi = al[0].al.ref_end > al[1].al.ref_end ? al[0].al.ref_end : al[1].al.ref_end;
    return hi - lo;
}

/* One pair of map_pairs_scored (lib/mapper/pair.rs:171-387). q[m], n[m]:
 * mate m's codes as sequenced; rows[m * 2 + rv]: mate m's scores as
 * sequenced (rv 0) or reverse-complemented (rv 1). */
static void map_pair(const ctx *c, const uint8_t *const q[2], const size_t n[2], const int32_t *const rows[4],
                     int both_strands, int64_t max_insert, acc *a, eo_pair_mapping *out) {
    memset(out, 0, sizeof *out);
    out->hit_offset = (int64_t)a->n_hits;
    const size_t nr = c->n_refs;
    int32_t *cand = (int32_t *)malloc(2 * nr * sizeof(int32_t));
    int32_t *sum = (int32_t *)malloc(2 * nr * sizeof(int32_t));
    int32_t *s0 = (int32_t *)malloc(2 * nr * sizeof(int32_t));
    int32_t *s1 = (int32_t *)malloc(2 * nr * sizeof(int32_t));
    int32_t nc = 0;
    if (n[0] && n[1]) {
        /* Phase 1 (lib/mapper/pair.rs:185-231): eligible candidates sorted
         * by (-sum, ref, template strand); tiers are runs of equal sum. */
        const int32_t *r4[4] = {rows[0], rows[1], rows[2], rows[3]};
        nc = eo_pair_cands(r4, nr, both_strands, c->min_score, cand, sum, s0, s1);
    }
    uint8_t *rc[2] = {NULL, NULL};
    for (int m = 0; m < 2; m++) {
        rc[m] = (uint8_t *)malloc(n[m] ? n[m] : 1);
        eo_revcomp_codes(q[m], n[m], rc[m]);
    }
    /* Phase 2 (lib/mapper/pair.rs:236-286): tier by tier, every candidate
     * of the tier traced (both mates), until some pass pair_ok. */
    traced *tr = NULL;
    int32_t f0 = 0, f1 = 0; /* the tier holding the tie set: cand[f0 .. f1) */
    unsigned char *ok = NULL;
    for (int32_t t0 = 0; t0 < nc;) {
        int32_t t1 = t0;
        while (t1 < nc && sum[t1] == sum[t0]) t1++;
        tr = (traced *)calloc((size_t)(2 * (t1 - t0)), sizeof(traced));
        ok = (unsigned char *)calloc((size_t)(t1 - t0), 1);
        int found = 0;
        for (int32_t k = t0; k < t1; k++) {
            const int ref = cand[k] >> 1, t = cand[k] & 1;
            traced *al = &tr[2 * (k - t0)];
            for (int m = 0; m < 2; m++) trace(c, mate_reverse(m, t) ? rc[m] : q[m], n[m], ref, &al[m]);
            ok[k - t0] = (unsigned char)pair_ok(al, t, max_insert);
            found += ok[k - t0];
        }
        if (found) {
            f0 = t0;
            f1 = t1;
            break;
        }
        for (int32_t k = 0; k < 2 * (t1 - t0); k++) free(tr[k].ops);
        free(tr);
        free(ok);
        tr = NULL;
        ok = NULL;
        t0 = t1;
    }
    if (tr) {
        /* Phase 3 (lib/mapper/pair.rs:292-358): the proper pair. */
        int32_t n_tied = 0;
        for (int32_t k = f0; k < f1; k++) n_tied += ok[k - f0];
        out->proper = 1;
        out->score = sum[f0];
        out->n_tied = n_tied;
        /* best score per mate outside the tie set, over the strands taken */
        unsigned char *in_set = (unsigned char *)calloc(2 * nr, 1);
        for (int32_t k = f0; k < f1; k++)
            if (ok[k - f0]) in_set[cand[k]] = 1; /* (ref << 1) | strand */
        for (int t = 0; t < (both_strands ? 2 : 1); t++)
            for (size_t r = 0; r < nr; r++) {
                if (in_set[2 * r + (size_t)t]) continue;
                for (int m = 0; m < 2; m++) {
                    const int32_t s = rows[m * 2 + mate_reverse(m, t)][r];
                    if (!(out->flags & (1 << m)) || s > out->subopt[m]) out->subopt[m] = s;
                    out->flags |= 1 << m;
                }
            }
        free(in_set);
        int64_t keep = n_tied;
        if (c->max_ties >= 0 && c->max_ties + 1 < keep) keep = c->max_ties + 1;
        int32_t kept = 0;
        for (int32_t k = f0; k < f1 && kept < keep; k++) {
            if (!ok[k - f0]) continue;
            const traced *al = &tr[2 * (k - f0)];
            const int32_t ins = span(al), ref = cand[k] >> 1, t = cand[k] & 1;
            for (int m = 0; m < 2; m++) put_hit(a, &al[m], ref, mate_reverse(m, t), t, ins);
            kept++;
        }
        out->n_hits = kept;
        for (int32_t k = 0; k < 2 * (f1 - f0); k++) free(tr[k].ops);
        free(tr);
        free(ok);
    } else {
        /* The fallback (lib/mapper/pair.rs:360-387): each mate mapped as an
         * unpaired read on both strands. */
        for (int m = 0; m < 2; m++) map_read(c, q[m], n[m], rows[m * 2], rows[m * 2 + 1], a, &out->separate[m]);
    }
    free(rc[0]);
    free(rc[1]);
    free(cand);
    free(sum);
    free(s0);
    free(s1);
}

typedef struct {
    const ctx *c;
    int pairs;
    const uint8_t *codes;
    const uint64_t *off;
    const int32_t *len;
    const int32_t *fwd, *rev;
    int both_strands;
    int64_t max_insert;
    size_t i0, i1;
    acc a;
    eo_read_mapping *reads;
    eo_pair_mapping *prs;
    eo_pair_scratch *ps;
} job;

static void *run_job(void *p) {
    job *j = (job *)p;
    const ctx *c = j->c;
    const size_t nr = c->n_refs;
    for (size_t i = j->i0; i < j->i1; i++) {
        if (j->pairs) {
            const uint8_t *q[2];
            size_t n[2];
            const int32_t *rows[4];
            for (int m = 0; m < 2; m++) {
                const size_t k = 2 * i + (size_t)m;
                q[m] = j->codes + j->off[k];
                n[m] = (size_t)j->len[k];
                rows[m * 2] = j->fwd + k * nr;
                rows[m * 2 + 1] = j->rev + k * nr;
            }
            map_pair(c, q, n, rows, j->both_strands, j->max_insert, &j->a, &j->prs[i]);
        } else {
            /* chunks of 64 reads, their winners traced together */
            const uint8_t *qs[64];
            size_t ns[64];
            const int32_t *fw[64], *rv[64];
            const size_t cnt = j->i1 - i < 64 ? j->i1 - i : 64;
            for (size_t t = 0; t < cnt; t++) {
                qs[t] = j->codes + j->off[i + t];
                ns[t] = (size_t)j->len[i + t];
                fw[t] = j->fwd + (i + t) * nr;
                rv[t] = j->rev ? j->rev + (i + t) * nr : NULL;
            }
            if (!j->ps) j->ps = eo_pair_scratch_new();
            map_read_chunk(c, j->ps, qs, ns, fw, rv, cnt, &j->a, &j->reads[i]);
            i += cnt - 1;
        }
    }
    eo_pair_scratch_free(j->ps);
    j->ps = NULL;
    return NULL;
}

static eo_mapped *run(const ctx *c, int pairs, const uint8_t *codes, const uint64_t *off, const int32_t *len,
                      size_t n, const int32_t *fwd, const int32_t *rev, int both_strands, int64_t max_insert,
                      int threads) {
    eo_mapped *m = (eo_mapped *)calloc(1, sizeof(eo_mapped));
    m->n = n;
    if (pairs) m->pairs = (eo_pair_mapping *)calloc(n ? n : 1, sizeof(eo_pair_mapping));
    else m->reads = (eo_read_mapping *)calloc(n ? n : 1, sizeof(eo_read_mapping));
    if (threads < 1) threads = 1;
    if ((size_t)threads > n) threads = n ? (int)n : 1;
    job *jobs = (job *)calloc((size_t)threads, sizeof(job));
    pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    for (int t = 0; t < threads; t++) {
        job *j = &jobs[t];
        j->c = c;
        j->pairs = pairs;
        j->codes = codes;
        j->off = off;
        j->len = len;
        j->fwd = fwd;
        j->rev = rev;
        j->both_strands = both_strands;
        j->max_insert = max_insert;
        j->i0 = n * (size_t)t / (size_t)threads;
        j->i1 = n * (size_t)(t + 1) / (size_t)threads;
        j->reads = m->reads;
        j->prs = m->pairs;
        if (t) pthread_create(&th[t], NULL, run_job, j);
    }
    run_job(&jobs[0]);
    for (int t = 1; t < threads; t++) pthread_join(th[t], NULL);
    /* concatenate the workers' hits and runs in input order */
    size_t nh = 0, no = 0;
    for (int t = 0; t < threads; t++) {
        nh += jobs[t].a.n_hits;
        no += jobs[t].a.n_ops;
    }
    m->hits = (eo_hit *)malloc((nh ? nh : 1) * sizeof(eo_hit));
    m->ops = (uint32_t *)malloc((no ? no : 1) * sizeof(uint32_t));
    size_t hb = 0, ob = 0;
    for (int t = 0; t < threads; t++) {
        job *j = &jobs[t];
        for (size_t h = 0; h < j->a.n_hits; h++) {
            m->hits[hb + h] = j->a.hits[h];
            m->hits[hb + h].ops_offset += (int64_t)ob;
        }
        if (j->a.n_ops) memcpy(m->ops + ob, j->a.ops, j->a.n_ops * sizeof(uint32_t));
        for (size_t i = j->i0; i < j->i1; i++) {
            if (pairs) {
                m->pairs[i].hit_offset += (int64_t)hb;
                m->pairs[i].separate[0].hit_offset += (int64_t)hb;
                m->pairs[i].separate[1].hit_offset += (int64_t)hb;
            } else {
                m->reads[i].hit_offset += (int64_t)hb;
            }
        }
        hb += j->a.n_hits;
        ob += j->a.n_ops;
        free(j->a.hits);
        free(j->a.ops);
    }
    m->n_hits = nh;
    m->n_ops = no;
    free(jobs);
    free(th);
    return m;
}

eo_mapped *eo_map_reads(const uint8_t *codes, const uint64_t *off, const int32_t *len, size_t n,
                        const int32_t *fwd_rows, const int32_t *rev_rows, const uint8_t *rcodes,
                        const uint64_t *roff, const int32_t *rlen, size_t n_refs, int32_t ma, int32_t mi,
                        int32_t o, int32_t e, int local, int32_t min_score, int64_t max_ties, int threads) {
    const ctx c = {rcodes, roff, rlen, n_refs, ma, mi, o, e, local, min_score, max_ties};
    return run(&c, 0, codes, off, len, n, fwd_rows, rev_rows, 0, 0, threads);
}

eo_mapped *eo_map_pairs(const uint8_t *codes, const uint64_t *off, const int32_t *len, size_t n_pairs,
                        const int32_t *fwd_rows, const int32_t *rev_rows, const uint8_t *rcodes,
                        const uint64_t *roff, const int32_t *rlen, size_t n_refs, int32_t ma, int32_t mi,
                        int32_t o, int32_t e, int local, int32_t min_score, int64_t max_ties, int64_t max_insert,
                        int both_strands, int threads) {
    const ctx c = {rcodes, roff, rlen, n_refs, ma, mi, o, e, local, min_score, max_ties};
    return run(&c, 1, codes, off, len, n_pairs, fwd_rows, rev_rows, both_strands, max_insert, threads);
}

size_t eo_mapped_len(const eo_mapped *m) { return m ? m->n : 0; }
const eo_read_mapping *eo_mapped_reads(const eo_mapped *m) { return m ? m->reads : NULL; }
const eo_pair_mapping *eo_mapped_pairs(const eo_mapped *m) { return m ? m->pairs : NULL; }
const eo_hit *eo_mapped_hits(const eo_mapped *m, size_t *n) {
    if (n) *n = m ? m->n_hits : 0;
    return m ? m->hits : NULL;
}
const uint32_t *eo_mapped_ops(const eo_mapped *m, size_t *n) {
    if (n) *n = m ? m->n_ops : 0;
    return m ? m->ops : NULL;
}
void eo_mapped_free(eo_mapped *m) {
    if (!m) return;
    free(m->reads);
    free(m->pairs);
    free(m->hits);
    free(m->ops);
    free(m);
}

/* ---- SAM fields: sam::md_nm and sam::cigar_string (lib/sam.rs:32-124) ---- */

static unsigned char eo_upper(unsigned char b) { return (b >= 'a' && b <= 'z') ? (unsigned char)(b - 32) : b; }

/* calmd's match (lib/sam.rs:57-69): both letters A/C/G/T, U as T, case
 * ignored, and equal. */
static int eo_calmd_match(unsigned char q, unsigned char r) {
    q = eo_upper(q);
    r = eo_upper(r);
    if (q == 'U') q = 'T';
    if (r == 'U') r = 'T';
    return q == r && (r == 'A' || r == 'C' || r == 'G' || r == 'T');
}

static void eo_put_num(char *md, size_t *k, unsigned v) { *k += (size_t)sprintf(md + *k, "%u", v); }

int eo_md_nm(const uint8_t *read, size_t n, const uint8_t *ref, size_t m, size_t qs, size_t rs,
             const uint32_t *ops, size_t n_ops, char *md, size_t md_cap, uint32_t *nm) {
    /* worst case: a number and a letter per aligned base, plus the deletions */
    size_t need = 16;
    for (size_t k = 0; k < n_ops; k++) need += 12 * (size_t)(ops[k] & 0x3FFFFFFFu) + 12;
    if (md_cap < need) return -1;
    size_t i = qs, j = rs, w = 0;
    unsigned run = 0, edits = 0;
    for (size_t k = 0; k < n_ops; k++) {
        const unsigned op = ops[k] >> 30, len = ops[k] & 0x3FFFFFFFu;
        if (op == 0) {
            for (unsigned t = 0; t < len; t++, i++, j++) {
                if (i >= n || j >= m) return -1;
                if (eo_calmd_match(read[i], ref[j])) {
                    run++;
                } else {
                    eo_put_num(md, &w, run);
                    md[w++] = (char)eo_upper(ref[j]);
                    run = 0;
                    edits++;
                }
            }
        } else if (op == 1) {
            i += len;
            edits += len;
        } else {
            eo_put_num(md, &w, run);
            md[w++] = '^';
            for (unsigned t = 0; t < len; t++) {
                if (j + t >= m) return -1;
                md[w++] = (char)eo_upper(ref[j + t]);
            }
            run = 0;
            j += len;
            edits += len;
        }
    }
    eo_put_num(md, &w, run);
    md[w] = 0;
    *nm = edits;
    return (int)w;
}

int eo_cigar_string(const uint32_t *ops, size_t n_ops, size_t query_len, size_t qs, size_t qe, char *out,
                    size_t cap) {
    if (cap < 24 * (n_ops + 2)) return -1;
    size_t w = 0;
    if (qs > 0) w += (size_t)sprintf(out + w, "%zuS", qs);
    for (size_t k = 0; k < n_ops; k++) w += (size_t)sprintf(out + w, "%u%c", ops[k] & 0x3FFFFFFFu, "MID"[ops[k] >> 30]);
    if (query_len > qe) w += (size_t)sprintf(out + w, "%zuS", query_len - qe);
    out[w] = 0;
    return (int)w;
}
