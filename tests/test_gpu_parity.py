"""The gfx950 kernels against the CPU oracle, by equality (runs on an MI355X).

A port of the reference's crates/eschalign/tests/gpu_parity.rs — the same
seeded panels and reads (xorshift64*, seeds 0x5eed_0101..0104), the same five
schemes (two of them with a positive mismatch or free gaps), every read
length 1..40 around the 16-row stripe, reverse complements in the same
batch, a 300-reference panel, the real tRNA fixture reads on both
47-reference panels, unscored rows and refusal — plus the parts of this
build that have no CUDA counterpart: the int32 rescue, the batched kernel
over many length bins, the plain kernel, fused selection and graph replay.
"""
import os

import numpy as np
import pytest

import eschalign_amd as ea
from eschalign_amd import synth
from oracle import pyoracle as O
import refgen as G

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def codes_of(reads):
    return [G.encode(r) for r in reads]


def oracle(queries, panel_codes, scoring, mode):
    return O.score_matrix(queries, panel_codes, scoring, mode == ea.Mode.LOCAL, THREADS)


def check(panel_letters, reads, scoring, mode, what):
    """Every query (and its reverse complement) against every reference in
    one batch, drop-in i16 matrix and int32 extension both equal to the
    oracle (gpu_parity.rs:128-186)."""
    panel = ea.Panel(panel_letters)
    gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), mode)
    queries = codes_of(reads)
    queries += [ea.reverse_complement_codes(q) for q in queries]
    want = oracle(queries, panel.codes, scoring, mode)
    m = gpu.score_batch(queries)
    assert len(m) == len(queries) and m.n_refs() == len(panel)
    for k, q in enumerate(queries):
        row = m.row(k)
        assert row is not None, f"{what}: query {k} ({len(q)} nt) left unscored"
        bad = np.nonzero(row.astype(np.int32) != want[k])[0]
        assert bad.size == 0, (f"{what} {ea.Mode.name(mode)} {scoring}: query {k} ({len(q)} nt) "
                               f"vs ref {bad[0]}: {row[bad[0]]} != {want[k][bad[0]]}")
    s32, ok = gpu.score_batch_i32(queries)
    assert ok.all()
    assert np.array_equal(s32, want), what
    return len(queries) * len(panel)


def test_gpu_scores_equal_scalar():
    """gpu_parity.rs:207-250."""
    total = 0
    rng = G.Rng(0x5EED_0101)
    small = G.random_panel(rng, 37, 20, 280)
    reads = G.reads_for(rng, small, 60)
    for scoring in G.SCHEMES_GPU:
        for mode in (ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL):
            total += check(small, reads, scoring, mode, "random 37-ref panel")
    wide = G.random_panel(rng, 300, 20, 600)
    reads = G.reads_for(rng, wide, 6)
    for mode in (ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL):
        total += check(wide, reads, (2, -1, -10, -1), mode, "random 300-ref panel")
    fixture = G.fixture_reads()
    assert len(fixture) >= 5
    for fasta in ("trna_reference.fa", "trna_reference_ambiguous.fa"):
        panel = G.fixture_panel(fasta)
        assert len(panel) == 47
        for scoring in G.SCHEMES_GPU:
            for mode in (ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL):
                total += check(panel, fixture, scoring, mode, fasta)
    print(f"[gpu_parity] {total} (query, reference) pairs equal to the oracle")


def test_gpu_leaves_what_it_cannot_score_to_the_cpu():
    """gpu_parity.rs:356-378."""
    rng = G.Rng(0x5EED_0102)
    panel = ea.Panel(G.random_panel(rng, 20, 50, 100))
    scoring = (2, -1, -10, -1)
    gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), ea.Mode.LOCAL)
    long = G.encode(G.random_seq(rng, ea.MAX_READ_LEN + 1, 0.0))
    ok = G.encode(G.random_seq(rng, 120, 0.0))
    edge = G.encode(G.random_seq(rng, ea.MAX_READ_LEN, 0.0))
    queries = [ok, np.zeros(0, np.uint8), long, edge, ok]
    m = gpu.score_batch(queries)
    assert m.row(1) is None, "empty read"
    assert m.row(2) is None, "read over MAX_READ_LEN"
    want = oracle(queries, panel.codes, scoring, ea.Mode.LOCAL)
    for k in (0, 3, 4):
        assert np.array_equal(m.row(k).astype(np.int32), want[k]), k
    assert m.scores[1].sum() == 0 and m.scores[2].sum() == 0
    assert gpu.score_batch([]).is_empty()
    # the int32 extension scores the long read too
    s32, sc = gpu.score_batch_i32(queries)
    assert list(sc) == [True, False, True, True, True]
    for k in (0, 2, 3, 4):
        assert np.array_equal(s32[k], want[k]), k


def test_gpu_refuses_a_reference_it_cannot_hold():
    """gpu_parity.rs:380-389, at this library's bound."""
    with pytest.raises(ea.GpuError) as e:
        ea.GpuScorer(ea.Panel([b"A" * (ea.MAX_REF_LEN + 1)]), ea.Scoring(), ea.Mode.LOCAL)
    assert e.value.kind == "ReferenceTooLong"


def test_batch_load_refuses_codes_outside_the_alphabet():
    """A read code above 4 is refused with INVALID_ARG naming the read, in a
    small batch (one staging thread) and in the middle of a 12 MB batch (the
    multi-threaded gather); a clean reload of the same batch then scores."""
    rng = np.random.default_rng(7)
    panel = ea.Panel([bytes(rng.choice(list(b"ACGT"), 150).astype(np.uint8)) for _ in range(8)])
    gpu = ea.GpuScorer(panel, ea.Scoring(), ea.Mode.LOCAL)
    for n, length in ((20, 50), (80_000, 150)):
        reads = [rng.integers(0, 5, length).astype(np.uint8) for _ in range(n)]
        bad_at = n // 2 + 3
        reads[bad_at] = reads[bad_at].copy()
        reads[bad_at][length // 2] = 9
        b = gpu.batch(n, n * length)
        with pytest.raises(ea.GpuError) as e:
            b.load(reads)
        assert e.value.kind == "InvalidArg" and f"read {bad_at} " in str(e.value)
        reads[bad_at][length // 2] = 2
        b.load(reads)
        b.run(ea.RUN_SCORES32)
        got, ok = b.scores32()
        assert ok.all()
        k = [0, bad_at, n - 1]
        want = oracle([reads[i] for i in k], panel.codes, (2, -1, -10, -1), ea.Mode.LOCAL)
        assert np.array_equal(got[k], want)
        b.close()


def test_prepare_submit_finish_overlap():
    """Two batches in flight at once (prepare k+1 while k runs,
    crates/eschalign-cli/src/align.rs:2879-2887) give the same rows."""
    rng = G.Rng(0x5EED_0201)
    panel_l = G.random_panel(rng, 23, 60, 120)
    panel = ea.Panel(panel_l)
    gpu = ea.GpuScorer(panel, ea.Scoring(), ea.Mode.SEMIGLOBAL)
    b1 = codes_of(G.reads_for(rng, panel_l, 50))
    b2 = codes_of(G.reads_for(rng, panel_l, 70))
    p1 = gpu.submit(gpu.prepare(b1))
    p2 = gpu.submit(gpu.prepare(b2))
    m2 = gpu.finish(p2)
    m1 = gpu.finish(p1)
    for batch, m in ((b1, m1), (b2, m2)):
        want = oracle(batch, panel.codes, (2, -1, -10, -1), ea.Mode.SEMIGLOBAL)
        assert np.array_equal(m.scores.astype(np.int32), want)


def test_int32_rescue_outside_the_int16_bound():
    """Queries whose DP values can reach the int16 clamps are scored by the
    int32 kernel and still equal the oracle; the drop-in matrix leaves them
    to the CPU exactly as scores_len says."""
    rng = G.Rng(0x5EED_0301)
    panel_l = G.random_panel(rng, 9, 200, 200)
    panel = ea.Panel(panel_l)
    scoring = (1000, -1000, -1000, -1000)
    reads = [G.random_seq(rng, 20, 0.0), panel_l[0][:150], G.mutate(rng, panel_l[3]),
             G.random_seq(rng, 300, 0.05)]
    queries = codes_of(reads)
    for mode in (ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL):
        gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), mode)
        want = oracle(queries, panel.codes, scoring, mode)
        s32, ok = gpu.score_batch_i32(queries)
        assert ok.all()
        assert np.array_equal(s32, want)
        m = gpu.score_batch(queries)
        for k, q in enumerate(queries):
            assert (m.row(k) is not None) == gpu.scores_len(len(q))
            assert gpu.scores_len(len(q)) == O.fits_i16(len(q), panel.max_len(), scoring)


def test_batched_kernel_many_bins_and_chunks():
    """Hundreds of reads of every length 1..1000 (chunks that straddle
    length bins, padding rows in several stripes), both modes, three schemes,
    a panel with wildcards and ragged reference lengths."""
    rng = G.Rng(0x5EED_0401)
    panel_l = G.random_panel(rng, 41, 30, 400)
    panel = ea.Panel(panel_l)
    reads = []
    for t in range(700):
        length = 1 + rng.below(1000) if t % 3 else 1 + rng.below(40)
        if t % 2:
            r = panel_l[rng.below(len(panel_l))]
            read = G.mutate(rng, (r * (1 + length // max(1, len(r))))[:length])
        else:
            read = G.random_seq(rng, length, 0.02)
        reads.append(read if read else b"A")
    queries = codes_of(reads)
    for scoring in ((2, -1, -10, -1), (1, -1, -2, -1), (3, 1, -5, -2)):
        for mode in (ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL):
            gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), mode)
            want = oracle(queries, panel.codes, scoring, mode)
            s32, ok = gpu.score_batch_i32(queries)
            assert ok.all()
            bad = np.argwhere(s32 != want)
            assert bad.size == 0, (scoring, mode, bad[:5], len(queries[bad[0][0]]))


def test_plain_kernel_equals_batched_kernel_and_oracle():
    rng = G.Rng(0x5EED_0501)
    panel_l = G.random_panel(rng, 70, 100, 100)
    panel = ea.Panel(panel_l)
    reads = codes_of(G.reads_for(rng, panel_l, 300))
    for mode in (ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL):
        gpu = ea.GpuScorer(panel, ea.Scoring(1, -1, -2, -1), mode)
        b = gpu.batch(len(reads), sum(len(r) for r in reads))
        b.load(reads, both_strands=True)
        b.run(ea.RUN_SCORES32 | ea.RUN_PLAIN)
        plain, _ = b.scores32()
        b.run(ea.RUN_SCORES32)
        batched, _ = b.scores32()
        assert np.array_equal(plain, batched)
        rows = reads + [ea.reverse_complement_codes(q) for q in reads]
        assert np.array_equal(batched, oracle(rows, panel.codes, (1, -1, -2, -1), mode))


def test_select_equals_oracle():
    """The fused reduction against lib/mapper.rs:444-472 (oracle eo_select),
    both strands, forward only, max_ties and min_score."""
    rng = G.Rng(0x5EED_0601)
    # paralog-rich panel so that ties are common
    seeds = G.random_panel(rng, 12, 60, 60)
    panel_l = []
    for s in seeds:
        panel_l += [s, G.mutate(rng, s), s]
    panel = ea.Panel(panel_l)
    reads = codes_of(G.reads_for(rng, panel_l, 200)) + [np.zeros(0, np.uint8)]
    scoring = (2, -1, -10, -1)
    for mode in (ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL):
        gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), mode)
        fwd = oracle(reads, panel.codes, scoring, mode)
        rev = oracle([ea.reverse_complement_codes(q) for q in reads], panel.codes, scoring, mode)
        for both, min_score, max_ties in ((True, 0, None), (False, 0, None), (True, 40, 1),
                                          (True, -5, 0)):
            got = gpu.select_batch(reads, both_strands=both, min_score=min_score,
                                   max_ties=max_ties)
            for k, q in enumerate(reads):
                if len(q) == 0:
                    assert got[k].best is None and got[k].take == []
                    continue
                want = O.select(fwd[k], rev[k] if both else None, min_score, max_ties)
                g = got[k]
                assert (g.best, g.n_tied, g.suboptimal, g.take) == want, (k, both, min_score)


def test_select_pairs_equals_oracle():
    """Top tier of summed mate scores (lib/mapper/pair.rs:185-231)."""
    rng = G.Rng(0x5EED_0701)
    base = G.random_panel(rng, 10, 70, 25)
    panel_l = []
    for s in base:
        t = bytearray(s)
        t[len(t) // 2] = ord("A") if t[len(t) // 2] != ord("A") else ord("C")
        panel_l += [s, bytes(t)]
    panel = ea.Panel(panel_l)
    scoring = (2, -1, -10, -1)
    m1, m2 = [], []
    for k in range(150):
        r = panel_l[rng.below(len(panel_l))]
        a = G.mutate(rng, r[: len(r) // 2 + 10])
        b = G.mutate(rng, r[len(r) // 2 - 10:])
        b = bytes(b"ACGTN"[c] for c in ea.reverse_complement_codes(G.encode(b)))
        if k % 5 == 4:
            a, b = b, a
        m1.append(G.encode(a))
        m2.append(G.encode(b) if k % 17 else np.zeros(0, np.uint8))
    for mode in (ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL):
        gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), mode)
        rows = {}
        for name, mates in (("m1", m1), ("m2", m2)):
            rows[name] = (oracle(mates, panel.codes, scoring, mode),
                          oracle([ea.reverse_complement_codes(q) for q in mates], panel.codes,
                                 scoring, mode))
        for both, min_score in ((False, 0), (True, 0), (True, 30)):
            got = gpu.select_pairs(m1, m2, both_strands=both, min_score=min_score)
            for k in range(len(m1)):
                g = got[k]
                if len(m1[k]) == 0 or len(m2[k]) == 0:
                    assert g.n_cand == 0 and g.top == []
                    continue
                r = [[rows["m1"][0][k], rows["m1"][1][k]], [rows["m2"][0][k], rows["m2"][1][k]]]
                cands = O.pair_cands(r, min_score, both)
                assert g.n_cand == len(cands), k
                if not cands:
                    assert g.top == []
                    continue
                top = cands[0][0]
                tier = [(c[1], c[2]) for c in cands if c[0] == top]
                below = [c[0] for c in cands if c[0] != top]
                assert g.top_sum == top and g.top == tier and g.n_top == len(tier), k
                assert g.next_sum == (below[0] if below else None), k


def test_graph_replay_matches_eager():
    """A captured pipeline replayed for batches of different sizes, with
    eager work between replays, equals the eager launch every time."""
    rng = G.Rng(0x5EED_0801)
    panel_l = G.random_panel(rng, 30, 80, 80)
    panel = ea.Panel(panel_l)
    gpu = ea.GpuScorer(panel, ea.Scoring(), ea.Mode.LOCAL)
    b = gpu.batch(400, 400 * 400)
    for rep in range(4):
        reads = codes_of(G.reads_for(rng, panel_l, 100 + 80 * rep))
        b.load(reads, both_strands=bool(rep % 2))
        b.run(ea.RUN_SCORES32 | ea.RUN_NO_GRAPH)
        eager, _ = b.scores32()
        b.run(ea.RUN_SCORES32)
        graph, _ = b.scores32()
        b.run(ea.RUN_SCORES32)
        graph2, _ = b.scores32()
        assert np.array_equal(eager, graph) and np.array_equal(graph, graph2), rep
        rows = reads + ([ea.reverse_complement_codes(q) for q in reads] if rep % 2 else [])
        assert np.array_equal(graph, oracle(rows, panel.codes, (2, -1, -10, -1), ea.Mode.LOCAL))


def test_result_is_the_same_alone_and_in_a_mix():
    """Batch composition changes chunking and bins, never a score."""
    rng = G.Rng(0x5EED_0901)
    panel_l = G.random_panel(rng, 17, 100, 60)
    panel = ea.Panel(panel_l)
    gpu = ea.GpuScorer(panel, ea.Scoring(1, -1, -2, -1), ea.Mode.SEMIGLOBAL)
    reads = codes_of(G.reads_for(rng, panel_l, 260))
    s_all, _ = gpu.score_batch_i32(reads)
    for k in (0, 7, 55, 199, len(reads) - 1):
        alone, _ = gpu.score_batch_i32([reads[k]])
        assert np.array_equal(alone[0], s_all[k])


def test_every_kernel_variant_equals_oracle_and_autotune_picks_one():
    """Every batched-kernel variant (the autotuner's choices) gives the
    oracle's scores: both modes, lengths 1..900 (first-stripe heights
    4/8/12/16, several stripes), a panel with wildcards."""
    rng = G.Rng(0x5EED_0A01)
    panel_l = G.random_panel(rng, 13, 40, 300)
    panel = ea.Panel(panel_l)
    reads = [G.random_seq(rng, 1 + rng.below(900), 0.03) for _ in range(300)]
    reads += [G.random_seq(rng, n, 0.0) for n in range(1, 34)]
    queries = codes_of(reads)
    for scoring in ((2, -1, -10, -1), (3, 1, -5, -2)):
        for mode in (ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL):
            gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), mode)
            want = oracle(queries, panel.codes, scoring, mode)
            usable = gpu.usable_variants()
            # the diagonal-shift variants are semiglobal forms
            diag = [v for v, x in ea.variants().items() if x.diag]
            assert set(diag) <= set(usable) if mode == ea.Mode.SEMIGLOBAL else not set(diag) & set(usable)
            for v in usable:
                gpu.variant = v
                s32, ok = gpu.score_batch_i32(queries)
                assert ok.all() and np.array_equal(s32, want), (scoring, mode, v)
            for v in set(range(ea.load_library().esch_n_variants())) - set(usable):
                with pytest.raises(ea.GpuError):
                    gpu.variant = v
            chosen, ms = gpu.autotune(queries[:120])
            assert gpu.variant == chosen and chosen in usable
            # the autotuner times what the scorer admits and marks the rest -1
            assert all((ms[v] > 0) == (v in usable) for v in range(len(ms)))
            assert all(ms[v] == -1.0 for v in range(len(ms)) if v not in usable)


def test_fp16_variant_at_its_bound():
    """Variant 4 computes in fp16, exact for integers of magnitude up to
    2048. Panels at that bound (longest reference x max(|match|, |mismatch|)
    + |open| + |extend| = 2047 and 2048), with reads that reach the extreme
    values (a reference read back whole: the highest score; runs of one
    base against a reference of another: cells down at the low end of the
    bound), equal the
    oracle and variant 3 in both modes. One base longer, the panel refuses
    the fp16 variants and defaults to an int16 one."""
    rng = G.Rng(0x5EED_0A03)
    for scoring, L in (((2, -1, -10, -1), 1018), ((1, -1, -2, -1), 2045)):
        refs = [G.random_seq(rng, L, 0.0), b"C" * L, G.random_seq(rng, L // 2, 0.02),
                G.random_seq(rng, 60, 0.0)]
        reads = [refs[0], b"A" * L, b"A" * (L // 3), b"G" * 5, refs[0][100:900],
                 G.random_seq(rng, 700, 0.03), G.random_seq(rng, L + 40, 0.0)]
        reads += [G.random_seq(rng, 1 + rng.below(400), 0.02) for _ in range(60)]
        queries = codes_of(reads)
        queries += [ea.reverse_complement_codes(q) for q in queries]
        panel = ea.Panel(refs)
        for mode in (ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL):
            gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), mode)
            # semiglobal mode defaults to the diagonal shift in int16 (its
            # 32000 bound holds, its fp16 one does not)
            assert gpu.variant == (4 if mode == ea.Mode.LOCAL else 28), (scoring, mode)
            want = oracle(queries, panel.codes, scoring, mode)
            s32, ok = gpu.score_batch_i32(queries)
            assert ok.all() and np.array_equal(s32, want), (scoring, mode)
            assert want.max() >= L - 20, want.max()
            for v in (4, 5, 3):
                gpu.variant = v
                sv, okv = gpu.score_batch_i32(queries)
                assert okv.all() and np.array_equal(sv, want), (scoring, mode, v)
        longer = ea.Panel(refs + [G.random_seq(rng, L + 1, 0.0)])
        gpu = ea.GpuScorer(longer, ea.Scoring(*scoring), ea.Mode.LOCAL)
        assert gpu.variant == 6  # the int16 column-shifted variant (its bound holds)
        for v in F16_VARIANTS:
            with pytest.raises(ea.GpuError):
                gpu.variant = v
        chosen, ms = gpu.autotune(queries[:40])
        assert chosen not in F16_VARIANTS and all(ms[v] == -1.0 for v in F16_VARIANTS)


F16_VARIANTS = (4, 5, 7, 8, 9, 10, 12, 13, 15, 16, 17, 18, 19, 20, 23, 24, 25, 26)


def test_row_shifted_variants_at_their_bound():
    """Variants 6 (int16), 7 and 8 (fp16) hold every value of a stripe row r
    shifted by |e| r: exact while longest reference x M + |o| + 17 |e| stays
    within the arithmetic's bound (2048 for fp16, 32000 for int16); the same
    schedules on 24- and 32-row stripes (9-14, and the windowed 17-22) need
    25 |e| and 33 |e|. Panels
    exactly at it, with reads that reach both ends of the value range (a
    reference read back whole, runs of one base against a reference of
    another, reads several stripes long), equal the oracle in semiglobal
    mode; one base longer, the panel refuses the variant (semiglobal only:
    in local mode the variants are the unshifted 3 and 5)."""
    rng = G.Rng(0x5EED_0A04)
    # (scoring, L, variants, stripe rows): L x M + |o| + (rows + 1) |e|
    # equals the bound; the tall-stripe variants 9-14 on 24 and 32 rows
    for scoring, L, others, st in (((1, -1, -2, -1), 2029, (7, 8), 16), ((2, -1, -11, -3), 993, (7, 8), 16),
                                   ((100, -100, -200, -100), 301, (6,), 16),
                                   ((1, -1, -2, -1), 2021, (9, 10, 17, 19), 24),
                                   ((100, -100, -200, -100), 293, (11, 21), 24),
                                   ((2, -1, -11, -3), 969, (12, 13, 18, 20), 32),
                                   ((100, -100, -200, -100), 285, (14, 22), 32)):
        M = max(abs(scoring[0]), abs(scoring[1]))
        f16 = ea.variant_info(others[0]).f16
        assert all(ea.variant_info(w).stripe_rows == st and ea.variant_info(w).f16 == f16 for w in others)
        assert L * M + abs(scoring[2]) + (st + 1) * abs(scoring[3]) == (2048 if f16 else 32000)
        refs = [G.random_seq(rng, L, 0.0), b"C" * L, G.random_seq(rng, L // 2, 0.02),
                G.random_seq(rng, 40, 0.0)]
        reads = [refs[0], b"A" * L, b"A" * (L // 3), b"G" * 5, refs[0][L // 9:L - L // 9],
                 G.random_seq(rng, L + 40, 0.03), refs[0] + refs[0][:300]]
        reads += [G.random_seq(rng, 1 + rng.below(L), 0.02) for _ in range(40)]
        queries = codes_of(reads)
        queries += [ea.reverse_complement_codes(q) for q in queries]
        panel = ea.Panel(refs)
        gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), ea.Mode.SEMIGLOBAL)
        gpu.diag_ratio = 0
        want = oracle(queries, panel.codes, scoring, ea.Mode.SEMIGLOBAL)
        for w in others:
            gpu.variant = w
            s32, ok = gpu.score_batch_i32(queries)
            assert ok.all()
            bad = np.argwhere(s32 != want)
            assert bad.size == 0, (scoring, w, bad[:4])
        assert want.max() >= L * scoring[0] - 20 * abs(scoring[0]), want.max()
        longer = ea.Panel(refs + [G.random_seq(rng, L + 1, 0.0)])
        gpu = ea.GpuScorer(longer, ea.Scoring(*scoring), ea.Mode.SEMIGLOBAL)
        for w in others:
            with pytest.raises(ea.GpuError):
                gpu.variant = w


def _bound_case(rng, L, extra=()):
    """Panel and reads that reach both ends of a shifted variant's value
    range: a reference read back whole (the top score), runs of one base
    against a reference of another (the low end), reads several stripes
    long, and reads longer than the references."""
    refs = [G.random_seq(rng, L, 0.0), b"C" * L, G.random_seq(rng, L // 2, 0.02),
            G.random_seq(rng, 40, 0.0)]
    reads = [refs[0], b"A" * L, b"A" * (L // 3), b"G" * 5, refs[0][L // 9:L - L // 9],
             G.random_seq(rng, L + 40, 0.03), refs[0] + refs[0][:300], b"C" * 3 + refs[0]]
    reads += [G.random_seq(rng, 1 + rng.below(L), 0.02) for _ in range(40)] + list(extra)
    queries = codes_of(reads)
    queries += [ea.reverse_complement_codes(q) for q in queries]
    return refs, queries


def _check_variants(refs, queries, scoring, mode, vs, block):
    """The variants `vs` on a panel of `refs` run with the column block
    `block` (0: the shift never restarts) and equal the oracle."""
    panel = ea.Panel(refs)
    gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), mode)
    gpu.diag_ratio = 0
    want = oracle(queries, panel.codes, scoring, mode)
    for w in vs:
        assert gpu.column_block(w) == block, (scoring, mode, w, gpu.column_block(w), block)
        gpu.variant = w
        s32, ok = gpu.score_batch_i32(queries)
        assert ok.all()
        bad = np.argwhere(s32 != want)
        assert bad.size == 0, (scoring, mode, w, block, bad[:4])
    return want


def test_diagonal_shifted_variants_at_their_bound():
    """Variants 23-28 hold every value of stripe row r in reference column j
    shifted by |e| (r + j) (semiglobal mode): exact while longest reference
    L x (M + |e|) + |o| + (rows + 2) |e| + M stays within 2048 (fp16: 23-26)
    or 32000 (int16: 27, 28), the block the scorer takes at 32 rows. Past
    it the column part of the shift restarts every K columns, K |e| in
    place of L |e|, down to K = 8. Panels exactly at the whole-reference
    bound (no restart), one base over (a restart), exactly at the K = 8
    bound (a restart every 8 columns), with reads that reach both ends of
    the value range, equal the oracle; one base past that the panel refuses
    the variants; local mode refuses them on any panel."""
    rng = G.Rng(0x5EED_0A07)
    # (scoring, L at the whole-reference bound (32 rows), L at the K = 8 bound, variants)
    for scoring, L0, L8, vs in (((1, -1, -3, -1), 1005, 2002, (23, 24, 25, 26)),
                                ((2, -1, -9, -3), 387, 955, (24, 26)),
                                ((100, -100, -500, -100), 140, 272, (27, 28))):
        M, O_, E = max(abs(scoring[0]), abs(scoring[1])), abs(scoring[2]), abs(scoring[3])
        lim = 2048 if ea.variant_info(vs[0]).f16 else 32000
        assert L0 * (M + E) + O_ + 34 * E + M == lim
        assert L8 * M + 8 * E + O_ + 34 * E + M in (lim, lim - 1) and (L8 + 1) * M + 42 * E + O_ + M > lim
        refs, queries = _bound_case(rng, L0)
        want = _check_variants(refs, queries, scoring, ea.Mode.SEMIGLOBAL, vs, 0)
        assert want.max() >= L0 * scoring[0] - 20 * abs(scoring[0]), want.max()
        # one base over: the widest power of two that fits
        K1 = next(K for K in (4096, 2048, 1024, 512, 256, 128, 64, 32, 16, 8)
                  if (L0 + 1) * M + K * E + O_ + 34 * E + M <= lim)
        refs1 = refs + [G.random_seq(rng, L0 + 1, 0.0)]
        _check_variants(refs1, queries, scoring, ea.Mode.SEMIGLOBAL, vs, K1)
        refs8, q8 = _bound_case(rng, L8)
        want8 = _check_variants(refs8, q8[:20] + q8[-20:], scoring, ea.Mode.SEMIGLOBAL, vs, 8)
        assert want8.max() >= L8 * scoring[0] - 20 * abs(scoring[0]), want8.max()
        longer = ea.Panel(refs8 + [G.random_seq(rng, L8 + 1, 0.0)])
        gpu = ea.GpuScorer(longer, ea.Scoring(*scoring), ea.Mode.SEMIGLOBAL)
        for w in vs:
            assert gpu.column_block(w) == -1
            with pytest.raises(ea.GpuError):
                gpu.variant = w
        gpu = ea.GpuScorer(ea.Panel(refs), ea.Scoring(*scoring), ea.Mode.LOCAL)
        for w in vs:
            with pytest.raises(ea.GpuError):
                gpu.variant = w


def test_column_shifted_local_variants_at_their_bound():
    """In local mode variants 6, 7 and 8 (and their tall and windowed forms)
    hold every value of reference column j shifted by |e| j: exact while
    longest reference L x (M + |e|) + |o| + 3 |e| + M stays within 2048
    (fp16) or 32000 (int16). Past it the shift restarts every K columns,
    K |e| in place of L |e|, down to K = 8. Panels exactly at the
    whole-reference bound (no restart), one base over (a restart) and
    exactly at the K = 8 bound (a restart every 8 columns), with reads
    reaching the top score (a reference read back whole) and the floor
    (runs of one base against a reference of another), several stripes
    long, equal the oracle; one base past that local mode refuses the
    variants while semiglobal mode (its row-shift bound) keeps them."""
    rng = G.Rng(0x5EED_0A05)
    f16, i16 = (7, 8, 9, 10, 12, 13, 17, 18, 19, 20), (6, 11, 14, 21, 22)
    # (scoring, L at the whole-reference bound, L at the K = 8 bound, variants)
    for scoring, L0, L8, vs in (((1, -1, -2, -1), 1021, 2034, f16), ((2, -1, -12, -3), 405, 1000, f16),
                                ((100, -100, -200, -100), 157, 306, i16)):
        M, O_, E = max(abs(scoring[0]), abs(scoring[1])), abs(scoring[2]), abs(scoring[3])
        lim = 32000 if vs == i16 else 2048
        assert L0 * (M + E) + O_ + 3 * E + M == lim
        assert L8 * M + 8 * E + O_ + 3 * E + M in (lim, lim - 1) and (L8 + 1) * M + 11 * E + O_ + M > lim
        refs, queries = _bound_case(rng, L0, [b"C" * 3 + G.random_seq(rng, L0, 0.0)])
        want = _check_variants(refs, queries, scoring, ea.Mode.LOCAL, vs, 0)
        assert want.max() == L0 * scoring[0], want.max()
        K1 = next(K for K in (4096, 2048, 1024, 512, 256, 128, 64, 32, 16, 8)
                  if (L0 + 1) * M + K * E + O_ + 3 * E + M <= lim)
        refs1 = refs + [G.random_seq(rng, L0 + 1, 0.0)]
        _check_variants(refs1, queries, scoring, ea.Mode.LOCAL, vs, K1)
        refs8, q8 = _bound_case(rng, L8)
        want8 = _check_variants(refs8, q8[:20] + q8[-20:], scoring, ea.Mode.LOCAL, vs[:4], 8)
        assert want8.max() == L8 * scoring[0], want8.max()
        longer = ea.Panel(refs8 + [G.random_seq(rng, L8 + 1, 0.0)])
        gpu = ea.GpuScorer(longer, ea.Scoring(*scoring), ea.Mode.LOCAL)
        for w in vs:
            with pytest.raises(ea.GpuError):
                gpu.variant = w
        gpu = ea.GpuScorer(longer, ea.Scoring(*scoring), ea.Mode.SEMIGLOBAL)
        for w in vs:
            if gpu.column_block(w) >= 0:  # the row shift's bound is its own
                gpu.variant = w


def test_every_kernel_variant_on_short_references():
    """References of every length 1..40 (column counts that leave 0-3
    columns after the paired / prefetching loops, the reference's padding
    word read by the profile prefetch) against reads 1..300, every variant,
    both modes, against the oracle."""
    rng = G.Rng(0x5EED_0A02)
    panel_l = [G.random_seq(rng, n, 0.05) for n in range(1, 41)]
    panel = ea.Panel(panel_l)
    reads = [G.random_seq(rng, 1 + rng.below(300), 0.03) for _ in range(200)]
    queries = codes_of(reads)
    for scoring in ((2, -1, -10, -1), (1, -1, -2, -1)):
        for mode in (ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL):
            gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), mode)
            gpu.diag_ratio = 0
            want = oracle(queries, panel.codes, scoring, mode)
            for v in gpu.usable_variants():
                gpu.variant = v
                s32, ok = gpu.score_batch_i32(queries)
                assert ok.all()
                bad = np.argwhere(s32 != want)
                assert bad.size == 0, (scoring, mode, v, bad[:4])


def test_anti_diagonal_kernel_equals_oracle():
    """The anti-diagonal wavefront kernel, forced for every read (threshold 1)
    and for long reads only: lengths 1..3000 (one to several 128-row passes),
    ragged reference pairs, wildcards in reads and references, five schemes,
    both modes, against the oracle; and the routing counts it."""
    rng = G.Rng(0x5EED_0B01)
    panel_l = G.random_panel(rng, 21, 30, 700)
    panel = ea.Panel(panel_l)
    reads = [G.random_seq(rng, 1 + rng.below(40), 0.05) for _ in range(40)]
    reads += [G.chimeric_read(rng, panel_l, 100 + rng.below(2900)) for _ in range(40)]
    queries = codes_of(reads)
    for scoring in G.SCHEMES_GPU:
        for mode in (ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL):
            gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), mode)
            gpu.diag_ratio = 0      # only the fixed threshold routes here
            want = oracle(queries, panel.codes, scoring, mode)
            for thr in (1, 1000):
                gpu.diag_min_len = thr
                b = gpu.batch(len(queries), sum(len(q) for q in queries))
                b.load(queries)
                b.run(ea.RUN_SCORES32)
                got, ok = b.scores32()
                assert ok.all()
                bad = np.argwhere(got != want)
                assert bad.size == 0, (scoring, mode, thr, bad[:4], len(queries[bad[0][0]]))
                assert b.n_diag() == sum(len(q) >= thr for q in queries)
                b.close()


def test_anti_diagonal_fallback_for_long_references():
    """A panel with a reference over 4096 nt keeps the anti-diagonal pass
    boundary in global memory (the 16-lane kernel) instead of LDS: forced
    for every read and for long ones, both modes, against the oracle."""
    rng = G.Rng(0x5EED_0B05)
    panel_l = [G.random_seq(rng, 4500, 0.01)] + G.random_panel(rng, 6, 50, 400)
    panel = ea.Panel(panel_l)
    reads = [G.chimeric_read(rng, panel_l[1:], 100 + rng.below(1500)) for _ in range(12)]
    reads += [G.random_seq(rng, 1 + rng.below(60), 0.02) for _ in range(8)]
    queries = codes_of(reads)
    scoring = (2, -1, -10, -1)
    for mode in (ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL):
        gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), mode)
        gpu.diag_ratio = 0
        want = oracle(queries, panel.codes, scoring, mode)
        for thr in (1, 500):
            gpu.diag_min_len = thr
            b = gpu.batch(len(queries), sum(len(q) for q in queries))
            b.load(queries)
            b.run(ea.RUN_SCORES32)
            got, ok = b.scores32()
            assert ok.all()
            bad = np.argwhere(got != want)
            assert bad.size == 0, (mode, thr, bad[:4])
            assert b.n_diag() == sum(len(q) >= thr for q in queries)
            b.close()


def _routed(gpu, queries, ratio):
    gpu.diag_ratio = ratio
    b = gpu.batch(len(queries), sum(len(q) for q in queries))
    b.load(queries)
    b.run(ea.RUN_SCORES32)
    got, ok = b.scores32()
    nd = b.n_diag()
    b.close()
    return got, ok, nd


def diag_padded(L):
    """Rows the anti-diagonal kernel computes for a read of length L: whole
    passes of 64 R rows, R = 2, 4 or 8 by length."""
    rp = 64 * (2 if L <= 128 else 4 if L <= 256 else 8)
    return -(-L // rp) * rp


def routed_count(lens, ratio, n_refs, blocks16):
    """How many reads the per-length-bucket rule routes: a length bin L
    (>= 128) goes when L > ratio * L' + 64, L' the length of the read 127
    ranks later (or the shortest read), or when L >= 512 and the batch's
    inter-sequence work items (128-read chunks x groups of 4 references)
    are fewer than that kernel's resident blocks. Then, when the kept reads
    do not fill one 128-read chunk and ratio x their padded rows is less
    than 128 x the longest kept length, the kept reads go too."""
    srt = sorted(lens, reverse=True)
    total = len(srt)
    starved = -(-total // 128) * -(-n_refs // 4) < blocks16
    r16 = round(ratio * 16)
    routed = set()
    for L in set(lens):
        if L < 128:
            continue
        r = sum(1 for x in srt if x > L)
        L2 = srt[min(r + 127, total - 1)]
        if 16 * L > r16 * L2 + 1024 or (starved and L >= 512):
            routed.add(L)
    kept = [L for L in lens if L not in routed]
    if r16 > 0 and 0 < len(kept) < 128 and r16 * sum(diag_padded(L) for L in kept) < 16 * 128 * max(kept):
        return len(lens)
    return len(lens) - len(kept)


def test_adaptive_routing_per_length_bucket():
    """The per-length-bucket routing between the inter-sequence and the
    anti-diagonal kernel: a long tail over many short reads (each long read
    is over twice the read 127 ranks later), a handful of long reads only
    (too few chunks to fill the device: all routed) and a small uniform
    600-1400 batch (starved: the >= 512 bins go). Every routing gives the
    oracle's scores, and the routed count is the rule's."""
    rng = G.Rng(0x5EED_0B02)
    panel_l = G.random_panel(rng, 24, 150, 400)
    panel = ea.Panel(panel_l)
    scoring = (1, -1, -2, -1)
    short = [G.chimeric_read(rng, panel_l, 50 + rng.below(100)) for _ in range(600)]
    tail = [G.chimeric_read(rng, panel_l, 1500 + rng.below(2500)) for _ in range(12)]
    few = [G.chimeric_read(rng, panel_l, 2000 + rng.below(3000)) for _ in range(10)]
    small = [G.chimeric_read(rng, panel_l, 300 + rng.below(1100)) for _ in range(200)]
    cases = [("tail", codes_of(short + tail)), ("few", codes_of(few)), ("small", codes_of(small))]
    for mode in (ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL):
        gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), mode)
        assert gpu.diag_ratio == 2.0
        blocks16 = gpu.geometry()[2]
        for name, queries in cases:
            lens = [len(q) for q in queries]
            want = oracle(queries, panel.codes, scoring, mode)
            off, ok0, nd0 = _routed(gpu, queries, 0)
            assert ok0.all() and nd0 == 0
            assert np.array_equal(off, want), (name, mode)
            for ratio in (2.0, 1.5, 4.0):
                got, ok, nd = _routed(gpu, queries, ratio)
                assert ok.all()
                bad = np.argwhere(got != want)
                assert bad.size == 0, (name, mode, ratio, bad[:4])
                assert nd == routed_count(lens, ratio, len(panel_l), blocks16), (name, ratio, nd)
        assert routed_count([len(q) for q in cases[0][1]], 2.0, len(panel_l), blocks16) == len(tail)
        assert routed_count([len(q) for q in cases[1][1]], 2.0, len(panel_l), blocks16) == len(few)


def test_adaptive_routing_dense_batch_keeps_the_chunks():
    """A batch large enough to fill the device with 128-read chunks of
    near-equal lengths keeps them on the inter-sequence kernel and routes
    only its 10 long outliers; at this size the routed scores equal the
    unrouted ones (each path pinned to the oracle above)."""
    panel = synth.panel(seed=0x5EED_0B04, n_refs=164, lo=140, hi=160)
    reads = synth.reads(panel, seed=0x5EED_0B05, n=20_000, length=("uniform", 600, 900))
    reads += synth.reads(panel, seed=0x5EED_0B06, n=10, length=("uniform", 2000, 3000))
    gpu = ea.GpuScorer(ea.Panel(panel), ea.Scoring(1, -1, -2, -1), ea.Mode.SEMIGLOBAL)
    blocks16 = gpu.geometry()[2]
    lens = [len(q) for q in reads]
    assert -(-len(reads) // 128) * -(-164 // 4) >= blocks16
    off, ok0, _ = _routed(gpu, reads, 0)
    for ratio in (2.0, 1.0):
        got, ok, nd = _routed(gpu, reads, ratio)
        assert ok.all() and ok0.all() and np.array_equal(got, off), ratio
        assert nd == routed_count(lens, ratio, 164, blocks16), (ratio, nd)
    assert routed_count(lens, 2.0, 164, blocks16) == 10


def test_small_batch_against_a_wide_panel_takes_the_anti_diagonal_kernel():
    """A batch too small to fill one 128-row chunk of the inter-sequence
    kernel (a few reads against 900 references) goes to the anti-diagonal
    kernel whole when its padded rows cost less there; reads so short that
    their passes would be mostly padding stay. Scores equal the oracle's
    either way, and the routed count is the rule's."""
    rng = G.Rng(0x5EED_0B04)
    panel_l = G.random_panel(rng, 900, 60, 300)
    panel = ea.Panel(panel_l)
    scoring = (1, -1, -2, -1)
    cases = {
        "six reads": [G.chimeric_read(rng, panel_l, 80 + rng.below(100)) for _ in range(6)],
        "fifty reads": [G.chimeric_read(rng, panel_l, 90 + rng.below(300)) for _ in range(50)],
        "short reads": [G.random_seq(rng, 20 + rng.below(20), 0.0) for _ in range(70)],
    }
    for mode in (ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL):
        gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), mode)
        gpu.swap_rows = 0  # the swapped path would take these batches first
        blocks16 = gpu.geometry()[2]
        for name, reads in cases.items():
            queries = codes_of(reads)
            lens = [len(q) for q in queries]
            want = oracle(queries, panel.codes, scoring, mode)
            got, ok, nd = _routed(gpu, queries, 2.0)
            assert ok.all() and np.array_equal(got, want), (name, mode)
            expect = routed_count(lens, 2.0, len(panel), blocks16)
            assert nd == expect, (name, mode, nd, expect)
            if name == "short reads":
                assert nd == 0
            elif name == "six reads":
                assert nd == len(queries)


def _scores(gpu, queries):
    b = gpu.batch(len(queries), sum(len(q) for q in queries))
    b.load(queries)
    b.run(ea.RUN_SCORES32)
    got, ok = b.scores32()
    swapped = b.swapped()
    cells = b.stats()[0]
    b.close()
    return got, ok, swapped, cells


def test_swapped_roles_equal_the_oracle():
    """A few reads against many references are scored with the roles
    swapped (the references as the queries, the reads as a panel): every
    scheme of the parity suite, both modes, wildcards in reads and panel,
    reads shorter and longer than most references, the int32 matrix and the
    fused selection on both strands, equal to the oracle and to the ordinary
    path; a high-scoring scheme whose long references exceed the int16
    bound takes the int32 rescue with the roles swapped too."""
    rng = G.Rng(0x5EED_0B06)
    panel_l = G.random_panel(rng, 300, 40, 460)
    reads = [G.random_seq(rng, 1 + rng.below(30), 0.05) for _ in range(6)]
    reads += [G.chimeric_read(rng, panel_l, 60 + rng.below(300)) for _ in range(14)]
    reads.append(panel_l[7])
    assert max(map(len, reads)) <= max(map(len, panel_l))
    queries = codes_of(reads)
    panel = ea.Panel(panel_l)
    for scoring in G.SCHEMES_GPU:
        for mode in (ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL):
            gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), mode)
            assert gpu.swap_rows == 0  # by the rule only from 4,096 references
            gpu.swap_rows = 127
            want = oracle(queries, panel.codes, scoring, mode)
            got, ok, swapped, cells = _scores(gpu, queries)
            assert swapped and ok.all(), (scoring, mode)
            bad = np.argwhere(got != want)
            assert bad.size == 0, (scoring, mode, bad[:4])
            gpu.swap_rows = 0
            got0, ok0, swapped0, cells0 = _scores(gpu, queries)
            assert not swapped0 and ok0.all() and np.array_equal(got0, want) and cells0 == cells
            sel0 = gpu.select_batch(queries, both_strands=True, min_score=3, max_ties=2)
            gpu.swap_rows = 127
            sel = gpu.select_batch(queries, both_strands=True, min_score=3, max_ties=2)
            assert sel == sel0, (scoring, mode)
            n = len(queries)
            rows = queries + [ea.reverse_complement_codes(q) for q in queries]
            sm = oracle(rows, panel.codes, scoring, mode)
            for k in range(n):
                b, nt, so, take = O.select(sm[k], sm[n + k], 3, 2)
                assert (sel[k].best, sel[k].n_tied, sel[k].suboptimal, sel[k].take) == (b, nt, so, take), k
    # the int32 rescue with the roles swapped: long references at a high
    # match score exceed the int16 bound
    big = [G.random_seq(rng, 1900 + rng.below(100), 0.01) for _ in range(3)] + G.random_panel(rng, 260, 40, 200)
    long_reads = [big[0][100:1800], big[2][:1500]] + [G.chimeric_read(rng, big[3:], 300) for _ in range(3)]
    lq = codes_of(long_reads)
    panel = ea.Panel(big)
    for mode in (ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL):
        gpu = ea.GpuScorer(panel, ea.Scoring(20, -1, -10, -1), mode)
        gpu.swap_rows = 64
        want = oracle(lq, panel.codes, (20, -1, -10, -1), mode)
        got, ok, swapped, _ = _scores(gpu, lq)
        assert swapped and ok.all() and np.array_equal(got, want), mode
        assert want.max() > 32000




def test_column_restart_through_the_swapped_path():
    """A panel with references up to 1,900 nt under 1,-1,-2,-1: the
    default variants shift columns with a restart (local mode: the column
    shift every 128 columns; semiglobal mode: the diagonal shift every 64),
    and the swapped-roles path runs them on the reads as a panel with the
    scorer's block. Ordinary and swapped scores equal the oracle in both
    modes, and so does the fused selection on both strands."""
    rng = G.Rng(0x5EED_0B08)
    panel_l = [G.random_seq(rng, n, 0.01) for n in (1900, 1700, 1500)] + G.random_panel(rng, 130, 40, 400)
    reads = [panel_l[0][200:1400], panel_l[1][:900], G.random_seq(rng, 1700, 0.02)]
    reads += [G.chimeric_read(rng, panel_l, 60 + rng.below(500)) for _ in range(12)]
    queries = codes_of(reads)
    panel = ea.Panel(panel_l)
    L = max(map(len, panel_l))
    for mode, block in ((ea.Mode.LOCAL, 128), (ea.Mode.SEMIGLOBAL, 64)):
        gpu = ea.GpuScorer(panel, ea.Scoring(1, -1, -2, -1), mode)
        v = gpu.variant
        assert ea.variant_info(v).f16 and gpu.column_block(v) == block, (mode, v, gpu.column_block(v), L)
        want = oracle(queries, panel.codes, (1, -1, -2, -1), mode)
        gpu.swap_rows = 0
        got0, ok0, swapped0, _ = _scores(gpu, queries)
        assert not swapped0 and ok0.all() and np.array_equal(got0, want), mode
        gpu.swap_rows = 127
        got, ok, swapped, _ = _scores(gpu, queries)
        assert swapped and ok.all(), mode
        bad = np.argwhere(got != want)
        assert bad.size == 0, (mode, bad[:4])
        sel = gpu.select_batch(queries, both_strands=True, max_ties=3)
        gpu.swap_rows = 0
        assert gpu.select_batch(queries, both_strands=True, max_ties=3) == sel, mode


def test_drop_in_matrix_on_a_panel_that_restarts_its_shift():
    """References up to 1,900 nt under 1,-1,-2,-1, where the default
    variants restart their column shift (local mode every 128 columns,
    semiglobal mode every 64): the drop-in int16 matrix and the int32
    extension, every query and its reverse complement, both modes, equal
    to the oracle; reads of 1-1,700 nt, then a batch of short reads only."""
    rng = G.Rng(0x5EED_0B09)
    panel_l = [G.random_seq(rng, n, 0.01) for n in (1900, 1700, 1500)] + G.random_panel(rng, 60, 40, 400)
    reads = [panel_l[0][200:1400], panel_l[1][:900], G.random_seq(rng, 1700, 0.02)]
    reads += G.reads_for(rng, panel_l, 40)
    short = G.reads_for(rng, panel_l, 1)[:40]  # lengths 1-40
    for mode in (ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL):
        gpu = ea.GpuScorer(ea.Panel(panel_l), ea.Scoring(1, -1, -2, -1), mode)
        assert gpu.column_block(gpu.variant) > 0, mode
        del gpu
        check(panel_l, reads, (1, -1, -2, -1), mode, "long-reference panel")
        check(panel_l, short, (1, -1, -2, -1), mode, "short reads on a long-reference panel")


def test_swapped_roles_by_the_rule_on_a_wide_panel():
    """By the default rule a panel of at least 4,096 references swaps roles
    for batches of up to one row per reference: 40 reads on both strands
    against 4,100 references, fused selection, equal to the ordinary path
    and, on a sample, to the oracle."""
    rng = G.Rng(0x5EED_0B07)
    panel_l = G.random_panel(rng, 4100, 60, 120)
    reads = [G.chimeric_read(rng, panel_l, 40 + rng.below(120)) for _ in range(40)]
    queries = codes_of(reads)
    panel = ea.Panel(panel_l)
    gpu = ea.GpuScorer(panel, ea.Scoring(2, -1, -10, -1), ea.Mode.LOCAL)
    assert gpu.swap_rows == 4100
    got, ok, swapped, _ = _scores(gpu, queries)
    assert swapped and ok.all()
    sel = gpu.select_batch(queries, both_strands=True, max_ties=3)
    gpu.swap_rows = 0
    got0, ok0, swapped0, _ = _scores(gpu, queries)
    assert not swapped0 and np.array_equal(got, got0)
    assert gpu.select_batch(queries, both_strands=True, max_ties=3) == sel
    want = oracle(queries[:6], panel.codes, (2, -1, -10, -1), ea.Mode.LOCAL)
    assert np.array_equal(got[:6], want)


def test_autotune_sweeps_the_routing_ratio():
    """The autotuner times the variants, then the routing ratios, on a
    sample; it leaves an explicitly set ratio alone."""
    rng = G.Rng(0x5EED_0B03)
    panel_l = G.random_panel(rng, 40, 150, 400)
    reads = [G.chimeric_read(rng, panel_l, 50 + rng.below(100)) for _ in range(3000)]
    reads += [G.chimeric_read(rng, panel_l, 2000 + rng.below(2000)) for _ in range(8)]
    queries = codes_of(reads)
    panel = ea.Panel(panel_l)
    gpu = ea.GpuScorer(panel, ea.Scoring(1, -1, -2, -1), ea.Mode.LOCAL)
    chosen, ms = gpu.autotune(queries, both_strands=False)
    usable = gpu.usable_variants()
    assert chosen in usable and all((ms[v] > 0) == (v in usable) for v in range(len(ms)))
    assert gpu.diag_ratio in (0.0, 1.5, 2.0, 3.0, 4.0)
    gpu.diag_ratio = 3.0
    gpu.autotune(queries, both_strands=False)
    assert gpu.diag_ratio == 3.0
    with pytest.raises(ea.GpuError):
        gpu.diag_ratio = 0.5
    got, ok, _ = _routed(gpu, queries, 3.0)
    assert ok.all() and np.array_equal(got, oracle(queries, panel.codes, (1, -1, -2, -1), ea.Mode.LOCAL))


def test_autotune_per_length_bucket_plan():
    """The routing plan per length class (0: stay on the inter-sequence
    kernel, 1: take the anti-diagonal kernel, 16..1024: the ratio rule):
    reads in six classes (50 nt to 4 kb); every plan tried gives the
    oracle's scores, a plan of 1s sends exactly the reads of 128 nt and
    over to the anti-diagonal kernel, the autotuner returns a valid plan
    (scores again equal), and a plan the caller set is left alone."""
    rng = G.Rng(0x5EED_0B05)
    panel_l = G.random_panel(rng, 30, 150, 400)
    spans = ((50, 127, 600), (128, 255, 300), (256, 511, 150), (512, 1023, 60), (1024, 2047, 20),
             (2048, 4095, 6))
    reads = [G.chimeric_read(rng, panel_l, lo + rng.below(hi - lo + 1)) for lo, hi, k in spans for _ in range(k)]
    queries = codes_of(reads)
    panel = ea.Panel(panel_l)
    scoring = (2, -1, -10, -1)
    want = oracle(queries, panel.codes, scoring, ea.Mode.SEMIGLOBAL)
    gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), ea.Mode.SEMIGLOBAL)
    long_reads = sum(len(q) >= 128 for q in queries)
    for plan in ([0] * 8, [1] * 8, [32, 1, 0, 24, 1, 64, 32, 32], [0, 0, 0, 1, 1, 1, 1, 1], [16, 1024, 48, 0, 16, 1, 0, 1]):
        gpu.diag_plan = plan
        assert gpu.diag_plan == plan
        b = gpu.batch(len(queries), sum(len(q) for q in queries))
        b.load(queries)
        b.run(ea.RUN_SCORES32)
        got, ok = b.scores32()
        nd = b.n_diag()
        b.close()
        assert ok.all() and np.array_equal(got, want), plan
        if plan == [1] * 8:
            assert nd == long_reads, (nd, long_reads)
    for bad in ([2] + [0] * 7, [0] * 7, [2000] * 8):
        with pytest.raises(ea.GpuError):
            gpu.diag_plan = bad
    tuned = ea.GpuScorer(panel, ea.Scoring(*scoring), ea.Mode.SEMIGLOBAL)
    tuned.autotune(queries, both_strands=True)
    plan = tuned.diag_plan
    assert len(plan) == 8 and all(v in (0, 1) or 16 <= v <= 1024 for v in plan), plan
    got, ok = tuned.score_batch_i32(queries)
    assert ok.all() and np.array_equal(got, want), plan
    fixed = [0, 1, 0, 1, 0, 1, 0, 1]
    tuned.diag_plan = fixed
    tuned.autotune(queries[:300], both_strands=True)
    assert tuned.diag_plan == fixed


@pytest.mark.gpu
def test_short_reads_take_fp16_on_a_panel_of_long_references():
    """Every value of the fp16 recurrences is bounded through min(read,
    reference), the shifted columns' span aside, and a column shift
    restarted every K columns spans K whatever the references. So on a
    panel of 2,100-nt references (no fp16 variant panel-wide: 2,100 +
    |o| + |e| > 2,048), a batch of reads up to 1,500 nt takes the fp16 form
    of the scorer's int16 variant: the diagonal shift in semiglobal mode
    (default 28 -> 24, with a 256-column block) and the column shift in
    local mode (default 6 -> 7, a 512-column block). With scoring
    1,-1,-20,-15 no block fits fp16 for those reads and the semiglobal
    scorer keeps the diagonal-shift int16 variant 28 (the row-shifted fp16
    form takes as many ops). A batch holding a 2,100-nt read (the top score,
    2,100, is beyond fp16's exact integers) stays on the scorer's int16
    variant, and so does every batch once the caller fixes a variant (an
    autotuned scorer still upgrades). Scores equal the oracle's in every
    case (the captured graphs are keyed by the variant and its block)."""
    rng = G.Rng(0x5EED_0A06)
    L = 2100
    refs = [G.random_seq(rng, L, 0.0), G.random_seq(rng, L - 300, 0.01), b"C" * 900]
    short = [refs[0][100:1600], refs[1][:700], b"A" * 400, G.random_seq(rng, 1500, 0.03)]
    short += [G.random_seq(rng, 1 + rng.below(1500), 0.02) for _ in range(30)]
    longer = short + [refs[0]]
    panel = ea.Panel(refs)
    for mode, scoring, default, fast in ((ea.Mode.SEMIGLOBAL, (1, -1, -2, -1), 28, 24),
                                         (ea.Mode.SEMIGLOBAL, (1, -1, -20, -15), 28, 28),
                                         (ea.Mode.LOCAL, (1, -1, -2, -1), 6, 7)):
        gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), mode)
        assert gpu.variant == default, (mode, scoring)
        for reads, want_v in ((short, fast), (longer, default), (short, fast)):
            q = codes_of(reads)
            b = gpu.batch(len(q), sum(len(x) for x in q))
            b.load(q)
            b.run(ea.RUN_SCORES32)
            b.wait()
            s32, ok = b.scores32()
            assert b.variant() == want_v, (mode, scoring, len(reads), b.variant())
            want = oracle(q, panel.codes, scoring, mode)
            assert ok.all() and np.array_equal(s32, want), (mode, scoring, want_v)
        # the autotuner's pick (among the int16 variants this panel admits)
        # is not the caller's: a batch of short reads still takes fp16
        chosen, _ = gpu.autotune(codes_of(short), both_strands=False)
        info = ea.variant_info(chosen)
        assert not info.f16  # an int16 variant
        rows = ea.variant_info(chosen).stripe_rows

        def on_rows(sched):
            return next(v for v, x in ea.variants().items() if x.schedule == sched and x.stripe_rows == rows)

        if mode == ea.Mode.LOCAL:
            up = on_rows(7 if rows == 16 else 10)
        elif info.diag:
            up = on_rows(12) if scoring == (1, -1, -2, -1) else chosen
        else:
            up = on_rows(10 if info.schedule == 11 else 7)
        q = codes_of(short)
        b = gpu.batch(len(q), sum(len(x) for x in q))
        b.load(q)
        b.run(ea.RUN_SCORES32)
        b.wait()
        assert b.variant() == up, (mode, scoring, chosen, b.variant())
        assert np.array_equal(b.scores32()[0], oracle(q, panel.codes, scoring, mode))
        gpu.variant = 6
        q = codes_of(short)
        b = gpu.batch(len(q), sum(len(x) for x in q))
        b.load(q)
        b.run(ea.RUN_SCORES32)
        b.wait()
        assert b.variant() == 6
        assert np.array_equal(b.scores32()[0], oracle(q, panel.codes, scoring, mode))
