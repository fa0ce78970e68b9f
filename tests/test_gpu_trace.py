"""Winner tracebacks on an MI355X (esch_align_pairs, the device counterpart
of simd::align_pairs, lib/simd/mod.rs:405-499) against the oracle's
traceback (eo_align, lib/scalar.rs:164-364), which is pinned to parasail's
coordinates and to a restatement of the tie rules in test_oracle_align.py.
Equality of score, coordinates and every run."""
import numpy as np
import pytest

import eschalign_amd as ea
from oracle import pyoracle as O
import refgen as G

pytestmark = pytest.mark.gpu


def oracle_aligns(queries, panel_codes, refs, scoring, local):
    out = []
    for q, r in zip(queries, refs):
        a = O.align(q, panel_codes[r], scoring, local)
        out.append(ea.Alignment(a[0], a[1], a[2], a[3], a[4], tuple(a[5])))
    return out


def check(panel_l, queries, refs, scoring, mode):
    panel = ea.Panel(panel_l)
    gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), mode)
    got = gpu.align_pairs(queries, refs)
    want = oracle_aligns(queries, panel.codes, refs, scoring, mode == ea.Mode.LOCAL)
    for k, (g, w) in enumerate(zip(got, want)):
        assert g == w, (k, len(queries[k]), refs[k], scoring, mode, g, w)
    return got


def test_trace_matches_parasail_golden_through_the_device():
    """The 192 golden results (score and coordinates, parasail 1.3.4), each
    pair traced on the device against a one-reference panel per pair."""
    g = G.golden()
    schemes = g["schemes"]
    by = {}
    for p in g["pairs"]:
        for res in p["results"]:
            by.setdefault((res["mode"], res["scheme"]), []).append((p, res))
    n = 0
    for (mode_s, scheme), items in by.items():
        scoring = tuple(schemes[scheme])
        mode = ea.Mode.LOCAL if mode_s == "local" else ea.Mode.SEMIGLOBAL
        panel_l = [p["reference"].encode() for p, _ in items]
        queries = [G.encode(p["query"].encode()) for p, _ in items]
        got = check(panel_l, queries, list(range(len(items))), scoring, mode)
        for a, (_, res) in zip(got, items):
            assert a.score == res["score"]
            assert (a.query_start, a.query_end, a.ref_start, a.ref_end) == (
                res["query_start"], res["query_end"], res["ref_start"], res["ref_end"])
            n += 1
    assert n == 192


def test_trace_random_pairs_every_scheme():
    """Reads of lengths 1-700 (one to 44 stripes, ragged within a group)
    against references of 1-500 nt with wildcards, chimeric and random
    reads, five schemes, both modes."""
    rng = G.Rng(0x5EED_0C01)
    panel_l = [G.random_seq(rng, 1 + rng.below(500), 0.02) for _ in range(60)]
    reads = [G.random_seq(rng, 1 + rng.below(60), 0.03) for _ in range(100)]
    reads += [G.chimeric_read(rng, panel_l, 20 + rng.below(680)) for _ in range(200)]
    queries = [G.encode(r) for r in reads]
    refs = [rng.below(len(panel_l)) for _ in queries]
    for scoring in G.SCHEMES_GPU:
        for mode in (ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL):
            check(panel_l, queries, refs, scoring, mode)


def test_trace_edges():
    """Empty reads (empty alignment), one-base reads and references, a local
    alignment that cannot score (empty), exact and reverse-complement
    matches, the same pair many times in one group, and many groups."""
    rng = G.Rng(0x5EED_0C02)
    panel_l = [b"A", b"ACGT", b"CCCC", G.random_seq(rng, 300, 0.0), G.random_seq(rng, 17, 0.0)]
    panel = ea.Panel(panel_l)
    queries = [np.zeros(0, np.uint8), G.encode(b"A"), G.encode(b"T"), G.encode(b"AAAA"),
               G.encode(panel_l[3]), ea.reverse_complement_codes(G.encode(panel_l[3])),
               G.encode(panel_l[4][3:16])]
    refs = [3, 0, 0, 2, 3, 3, 4]
    queries += [G.encode(panel_l[4])] * 130
    refs += [4] * 130
    for scoring in ((2, -1, -10, -1), (1, -1, -2, -1)):
        for mode in (ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL):
            got = check(panel_l, queries, refs, scoring, mode)
            assert got[0] == ea.Alignment(0, 0, 0, 0, 0, ())
            if mode == ea.Mode.LOCAL:
                assert got[3].is_empty() and got[3].score == 0
            assert got[4].cigar() == "300M" and got[4].score == 300 * scoring[0]
    gpu = ea.GpuScorer(panel, ea.Scoring(), ea.Mode.LOCAL)
    with pytest.raises(ea.GpuError):
        gpu.align_pairs([G.encode(b"ACGT")], [len(panel_l)])
    with pytest.raises(ea.GpuError):
        gpu.align_pairs([np.array([0, 9], np.uint8)], [0])


def test_trace_of_the_selected_winners():
    """The path's use: score reads against the panel on the device, select
    the winners (ties included), trace every winner on the device; equal to
    the oracle's traceback of each (read or reverse complement, reference)."""
    panel_l = G.fixture_panel("trna_reference.fa")
    panel = ea.Panel(panel_l)
    reads = [G.encode(r) for r in G.fixture_reads()]
    rng = G.Rng(0x5EED_0C03)
    reads += [G.encode(G.chimeric_read(rng, panel_l, 60 + rng.below(120))) for _ in range(200)]
    scoring = (2, -1, -10, -1)
    gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), ea.Mode.LOCAL)
    sel = gpu.select_batch(reads, both_strands=True, tie_cap=8)
    qs, rs = [], []
    for k, s in enumerate(sel):
        for ref, rev in s.take:
            qs.append(ea.reverse_complement_codes(reads[k]) if rev else reads[k])
            rs.append(ref)
    assert len(qs) >= len(reads) // 2
    got = gpu.align_pairs(qs, rs)
    want = oracle_aligns(qs, panel.codes, rs, scoring, True)
    assert got == want
    for a, q, r in zip(got, qs, rs):
        assert a.score == O.score(q, panel.codes[r], scoring, True)


def test_trace_concat_entry_and_repeated_calls():
    """esch_align_pairs_concat equals esch_align_pairs; the cached buffers
    give the same answers on a second, smaller call and on a larger one."""
    rng = G.Rng(0x5EED_0C04)
    panel_l = [G.random_seq(rng, 50 + rng.below(300), 0.01) for _ in range(20)]
    panel = ea.Panel(panel_l)
    gpu = ea.GpuScorer(panel, ea.Scoring(1, -1, -2, -1), ea.Mode.SEMIGLOBAL)
    for n in (500, 40, 3000):
        qs = [G.encode(G.chimeric_read(rng, panel_l, 1 + rng.below(400))) for _ in range(n)]
        rs = [rng.below(len(panel_l)) for _ in qs]
        a = gpu.align_pairs(qs, rs)
        codes = np.concatenate(qs)
        lens = np.array([len(q) for q in qs], np.int32)
        offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.uint64)
        recs, ops = gpu.align_pairs_concat(codes, offs, lens, rs)
        for k, al in enumerate(a):
            r = recs[k]
            assert (al.score, al.query_start, al.query_end, al.ref_start, al.ref_end) == tuple(int(x) for x in r[:5])
            seg = ops[int(r[6]):int(r[6]) + int(r[5])]
            assert al.ops == tuple(("MID"[int(x) >> 30], int(x) & 0x3FFFFFFF) for x in seg)
        want = oracle_aligns(qs[:200], panel.codes, rs[:200], (1, -1, -2, -1), False)
        assert a[:200] == want


def test_trace_long_reads_take_the_wavefront_kernel():
    """Reads of 300-5,000 nt against references of 100-1,200 nt (wildcards,
    chimeric reads), both modes, three schemes: the pairs that would set the
    one-lane kernel's time take the anti-diagonal kernel (R = 2, 4 and 8
    rows per lane, one to ten passes); equal to the oracle."""
    rng = G.Rng(0x5EED_0C05)
    panel_l = [G.random_seq(rng, 100 + rng.below(1100), 0.02) for _ in range(12)]
    lens = [300, 330, 400, 511, 512, 513, 700, 1024, 1500, 2600, 5000] + [300 + rng.below(2000) for _ in range(9)]
    queries = [G.encode(G.chimeric_read(rng, panel_l, n)) for n in lens]
    queries += [G.encode(panel_l[3][:200] + panel_l[5] + panel_l[7][50:])]
    refs = [rng.below(len(panel_l)) for _ in queries]
    for scoring in ((2, -1, -10, -1), (1, -1, -2, -1), (2, -1, 0, 0)):
        for mode in (ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL):
            panel = ea.Panel(panel_l)
            gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), mode)
            got = gpu.align_pairs(queries, refs)
            # the routing rule: at least 128 bases and max(2^15, cells / 2^18) cells
            cells = [len(q) * len(panel_l[r]) for q, r in zip(queries, refs)]
            thr = max(1 << 15, sum(cells) >> 18)
            assert gpu.align_last_wavefront() == sum(1 for q, c in zip(queries, cells) if len(q) >= 128 and c >= thr)
            assert gpu.align_last_wavefront() >= len(queries) - 3
            want = oracle_aligns(queries, panel.codes, refs, scoring, mode == ea.Mode.LOCAL)
            for k, (g, w) in enumerate(zip(got, want)):
                assert g == w, (k, len(queries[k]), refs[k], scoring, mode)


def test_trace_mixed_call_both_kernels():
    """One call with 3,000 short pairs and a few long ones: the long ones on
    the wavefront kernel, the rest on the one-lane kernel, every pair equal
    to the oracle and to the same pairs traced alone."""
    rng = G.Rng(0x5EED_0C06)
    panel_l = [G.random_seq(rng, 80 + rng.below(400), 0.01) for _ in range(30)]
    panel = ea.Panel(panel_l)
    queries = [G.encode(G.chimeric_read(rng, panel_l, 20 + rng.below(150))) for _ in range(3000)]
    queries += [G.encode(G.chimeric_read(rng, panel_l, 900 + rng.below(2000))) for _ in range(6)]
    refs = [rng.below(len(panel_l)) for _ in queries]
    for mode in (ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL):
        gpu = ea.GpuScorer(panel, ea.Scoring(2, -1, -10, -1), mode)
        got = gpu.align_pairs(queries, refs)
        assert 6 <= gpu.align_last_wavefront() < len(queries) // 2
        want = oracle_aligns(queries, panel.codes, refs, (2, -1, -10, -1), mode == ea.Mode.LOCAL)
        assert got == want
        short = gpu.align_pairs(queries[:3000], refs[:3000])
        assert short == got[:3000]


def test_trace_pair_over_the_one_lane_bound():
    """A 36,000-nt read against a 4,000-nt reference (1.44e8 cells, over the
    one-lane kernel's 2^27): traced by the wavefront kernel instead of being
    refused; equal to the oracle."""
    rng = G.Rng(0x5EED_0C07)
    ref = G.random_seq(rng, 4000, 0.0)
    read = G.random_seq(rng, 16000, 0.0) + G.mutate(rng, ref) + G.random_seq(rng, 16000, 0.0)
    read = read[:36000]
    panel = ea.Panel([ref])
    gpu = ea.GpuScorer(panel, ea.Scoring(2, -1, -10, -1), ea.Mode.LOCAL)
    got = gpu.align_pairs([G.encode(read)], [0])
    assert gpu.align_last_wavefront() == 1
    assert got == oracle_aligns([G.encode(read)], panel.codes, [0], (2, -1, -10, -1), True)
    assert got[0].score > 4000
