"""Read and pair mapping on an MI355X (esch_map_reads / esch_map_pairs: the
device counterparts of Aligner::map_reads_scored, lib/mapper.rs:400-550, and
map_pairs_scored with its tier loop and Separate fallback,
lib/mapper/pair.rs:157-387) against the oracle's eo_map_reads /
eo_map_pairs, which test_oracle_map.py pins to the reference's own cases and
to the paired fixture. Equality of every record: kind, scores, tie counts,
suboptimal scores, every hit's reference, strands, insert, coordinates and
runs."""
import os

import numpy as np
import pytest

import eschalign_amd as ea
from eschalign_amd import synth
from oracle import pyoracle as O
import refgen as G
from test_oracle_map import TRUSEQ_R1, TRUSEQ_R2, rc, xorshift_panel

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT = (2, -1, -10, -1)


def dev_reads(gpu, reads, **kw):
    recs, hits, ops, _ = gpu.map_reads_raw(reads, **kw)
    return [O.read_form(r, hits, ops) for r in recs]


def dev_pairs(gpu, pairs, **kw):
    recs, hits, ops, stats = gpu.map_pairs_raw([p[0] for p in pairs], [p[1] for p in pairs], **kw)
    return [O.pair_form(r, hits, ops) for r in recs], stats


def compare_pairs(gpu, panel_codes, pairs, scoring, local, **kw):
    got, stats = dev_pairs(gpu, pairs, **kw)
    want = O.map_pairs(pairs, panel_codes, scoring, local, **kw)
    for k, (g, w) in enumerate(zip(got, want)):
        assert g == w, (k, kw, g, w)
    assert len(got) == len(pairs)
    return got, stats


def test_map_reads_equals_oracle():
    """Random panels and the gpu_parity read mix (every length 1-40, reads
    from the panel, random reads), an empty and an all-wildcard read; both
    modes, one and both strands, min-score and max-ties."""
    rng = G.Rng(0x5EED0201)
    panel_l = G.random_panel(rng, 24, 40, 200)
    reads = [G.encode(r) for r in G.reads_for(rng, panel_l, 160)]
    reads += [np.zeros(0, np.uint8), G.encode(b"NNNNNN")]
    panel = ea.Panel(panel_l)
    for scoring, mode in ((DEFAULT, ea.Mode.LOCAL), ((1, -1, -2, -1), ea.Mode.SEMIGLOBAL)):
        gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), mode)
        local = mode == ea.Mode.LOCAL
        for both, min_score, max_ties in ((False, 0, None), (True, 0, None), (True, 25, 0), (True, 0, 2)):
            kw = dict(both_strands=both, min_score=min_score, max_ties=max_ties)
            got = dev_reads(gpu, reads, **kw)
            want = O.map_reads(reads, panel.codes, scoring, local, **kw)
            for k, (g, w) in enumerate(zip(got, want)):
                assert g == w, (k, kw, g, w)
        # the concatenated entry point gives the same records
        qc, qo, ql = O.concat(reads)
        recs, hits, ops, _ = gpu.map_reads_concat(qc, qo, ql, both_strands=True, min_score=0, max_ties=2)
        assert [O.read_form(r, hits, ops) for r in recs] == dev_reads(gpu, reads, both_strands=True, max_ties=2)
        # the object API reads the same records
        objs = gpu.map_reads(reads[:50], both_strands=True)
        forms = dev_reads(gpu, reads[:50], both_strands=True)
        for o, f in zip(objs, forms):
            assert (o.best_score, o.n_tied, o.suboptimal) == f[:3]
            assert tuple((h.reference, h.reverse, h.alignment.score, h.alignment.query_start,
                          h.alignment.query_end, h.alignment.ref_start, h.alignment.ref_end,
                          tuple(h.alignment.ops)) for h in o.hits) == f[3]


def test_map_pairs_reference_cases():
    """lib/mapper/pair.rs:449-573 on the device: the same outcomes, and
    records equal to the oracle's under every option."""
    a, b, c = xorshift_panel()
    panel = ea.Panel([a, b, c])
    gpu = ea.GpuScorer(panel, ea.Scoring(*DEFAULT), ea.Mode.SEMIGLOBAL)
    e = G.encode
    m = gpu.map_pairs([(e(a[:30]), e(rc(a[30:])))])[0]
    assert isinstance(m, ea.ProperPair) and m.n_tied == 1 and m.score == 120
    assert m.hits[0].reference == 0 and m.hits[0].insert == 60
    assert not m.hits[0].mates[0].reverse and m.hits[0].mates[1].reverse
    opts = dict(min_score=30)
    assert isinstance(gpu.map_pairs([(e(a[:20]), e(rc(a[40:])))], max_insert=50, **opts)[0], ea.SeparatePair)
    assert isinstance(gpu.map_pairs([(e(a[:20]), e(rc(a[40:])))], **opts)[0], ea.ProperPair)
    assert isinstance(gpu.map_pairs([(e(a[:20]), e(a[40:]))], **opts)[0], ea.SeparatePair)
    assert isinstance(gpu.map_pairs([(e(a[40:]), e(rc(a[:20])))], **opts)[0], ea.SeparatePair)
    assert isinstance(gpu.map_pairs([(e(rc(a[40:])), e(a[:20]))], **opts)[0], ea.SeparatePair)
    m = gpu.map_pairs([(e(rc(a[40:])), e(a[:20]))], both_strands=True, **opts)[0]
    assert isinstance(m, ea.ProperPair) and m.hits[0].reverse
    assert m.hits[0].mates[0].reverse and not m.hits[0].mates[1].reverse
    assert isinstance(gpu.map_pairs([(e(a[:30]), e(rc(a[30:])))], min_score=61)[0], ea.SeparatePair)
    pairs = [(e(a[:30]), e(rc(a[30:]))), (e(b[:30]), e(rc(b[30:]))), (e(a[:20]), e(a[40:])),
             (np.zeros(0, np.uint8), e(a[:20])), (e(rc(a[40:])), e(a[:20])), (e(a[40:]), e(rc(a[:20])))]
    for mode in (ea.Mode.SEMIGLOBAL, ea.Mode.LOCAL):
        g2 = ea.GpuScorer(panel, ea.Scoring(*DEFAULT), mode)
        for kw in (dict(), dict(both_strands=True), dict(min_score=30, max_insert=50),
                   dict(both_strands=True, min_score=30, max_ties=0)):
            compare_pairs(g2, panel.codes, pairs, DEFAULT, mode == ea.Mode.LOCAL, **kw)


def fixture_pairs():
    panel = G.read_fasta(os.path.join(HERE, "golden", "pairs", "pairs_panel.fa"))
    r1 = G.read_fastq(os.path.join(HERE, "golden", "pairs", "pairs_R1.fq"))
    r2 = G.read_fastq(os.path.join(HERE, "golden", "pairs", "pairs_R2.fq"))
    names, pairs = [], []
    for (n1, s1), (_, s2) in zip(r1, r2):
        ad = ea.adapter_detect_pair(s1, s2)
        names.append(n1)
        pairs.append((ea.encode(s1[:len(s1) - ad[0]]), ea.encode(s2[:len(s2) - ad[1]])))
    return [n for n, _ in panel], [s for _, s in panel], names, pairs


def test_map_pairs_paired_fixture():
    """crates/eschalign-cli/tests/align_paired.rs on the fixture (mates less
    their adapters, semi-global, --min-score 40): equal to the oracle under
    the default and --max-insert 120, and the outcomes the CLI tests state."""
    ref_names, refs, names, pairs = fixture_pairs()
    panel = ea.Panel(list(zip(ref_names, refs)))
    for mode in (ea.Mode.SEMIGLOBAL, ea.Mode.LOCAL):
        gpu = ea.GpuScorer(panel, ea.Scoring(*DEFAULT), mode)
        for kw in (dict(min_score=40), dict(min_score=40, max_insert=120), dict(min_score=40, both_strands=True)):
            compare_pairs(gpu, panel.codes, pairs, DEFAULT, mode == ea.Mode.LOCAL, **kw)
    gpu = ea.GpuScorer(panel, ea.Scoring(*DEFAULT), ea.Mode.SEMIGLOBAL)
    got = dict(zip(names, gpu.map_pairs(pairs, min_score=40)))
    ref = {n: i for i, n in enumerate(ref_names)}
    m = got["paralog"]
    assert isinstance(m, ea.ProperPair) and m.n_tied == 1 and m.hits[0].reference == ref["P1"]
    assert m.hits[0].insert == 130
    assert [x.alignment.ref_start for x in m.hits[0].mates] == [20, 90]
    assert m.suboptimal[0] == m.hits[0].mates[0].alignment.score
    m = got["pair_tie"]
    assert m.n_tied == 2 and [h.reference for h in m.hits] == [ref["P1"], ref["P2"]]
    m = got["short_insert"]
    assert m.hits[0].insert == 40 and m.hits[0].reference == ref["U1"]
    assert [x.alignment.cigar() for x in m.hits[0].mates] == ["40M", "40M"]
    m = got["one_mate_reads_through"]
    assert [x.alignment.cigar() for x in m.hits[0].mates] == ["40M", "35M"]
    assert isinstance(got["wrong_orientation"], ea.SeparatePair)
    m = got["one_unmapped"]
    assert isinstance(m, ea.SeparatePair) and m.mates[0].is_mapped() and not m.mates[1].is_mapped()
    sep = gpu.map_pairs([pairs[names.index("over_max_insert")]], min_score=40, max_insert=120)[0]
    assert isinstance(sep, ea.SeparatePair)
    assert [x.hits[0].reference for x in sep.mates] == [ref["U1"]] * 2


def rand_seq(rng, n):
    return bytes(np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, n)])


def random_pairs(seed, n_pairs):
    """Panels of paralog pairs (one substitution apart) and pairs of every
    kind: FR, RF, FF, outies, junk mates, empty mates, long gaps apart."""
    rng = np.random.default_rng(seed)
    base = [rand_seq(rng, int(rng.integers(60, 160))) for _ in range(10)]
    refs = []
    for s in base:
        refs.append(s)
        t = bytearray(s)
        t[int(rng.integers(len(t)))] = ord("ACGT"[int(rng.integers(4))])
        refs.append(bytes(t))
    pairs = []
    for _ in range(n_pairs):
        s = refs[int(rng.integers(len(refs)))]
        i = int(rng.integers(0, len(s) - 30))
        j = int(rng.integers(i + 15, len(s) + 1))
        x = s[i:i + int(rng.integers(12, 45))]
        y = s[max(0, j - int(rng.integers(12, 45))):j]
        kind = int(rng.integers(7))
        junk = rand_seq(rng, len(y))
        p = [(x, rc(y)), (rc(y), x), (x, y), (rc(x), y), (x, junk), (b"", x), (x, rc(y))][kind]
        pairs.append((ea.encode(p[0]), ea.encode(p[1])))
    return refs, pairs


def test_map_pairs_random_equal_oracle_and_chunking():
    """Random pairs of every kind on paralog panels under many options
    (fall-through to lower tiers, ties, both template strands, max-ties);
    results identical when the call is cut into small device chunks."""
    refs, pairs = random_pairs(0x9A1F, 300)
    panel = ea.Panel(refs)
    for scoring, mode in ((DEFAULT, ea.Mode.SEMIGLOBAL), ((1, -1, -2, -1), ea.Mode.LOCAL)):
        gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), mode)
        local = mode == ea.Mode.LOCAL
        for kw in (dict(), dict(both_strands=True, min_score=10), dict(min_score=15, max_insert=60, max_ties=1),
                   dict(both_strands=True, max_ties=0, max_insert=80)):
            whole, stats = compare_pairs(gpu, panel.codes, pairs, scoring, local, **kw)
            assert stats["rounds"] >= 2  # some pairs fall through their top tier
            gpu.set_map_chunk(38)  # 19 pairs per chunk
            cut, _ = dev_pairs(gpu, pairs, **kw)
            gpu.set_map_chunk(0)
            assert cut == whole
        reads = [p[k] for p in pairs for k in range(2)]
        full = dev_reads(gpu, reads, both_strands=True)
        gpu.set_map_chunk(25)
        assert dev_reads(gpu, reads, both_strands=True) == full
        gpu.set_map_chunk(0)


def test_map_pairs_wide_tiers_and_tie_lists():
    """40 identical references: a pair's tier (and a Separate mate's tie
    list) is wider than what one selection round returns inline, so the
    device takes its second pass for them; equal to the oracle with all ties
    and with max_ties."""
    rng = np.random.default_rng(7)
    s = rand_seq(rng, 120)
    other = rand_seq(rng, 120)
    refs = [s] * 40 + [other]
    panel = ea.Panel(refs)
    e = ea.encode
    pairs = [(e(s[:40]), e(rc(s[70:110]))),   # proper on all 40: a 40-wide tie set
             (e(s[:40]), e(s[70:110])),       # FF: Separate under a floor, 40 ties per mate
             (e(rc(s[70:110])), e(s[:40])),   # RF: proper only on the reverse template
             (e(other[:50]), e(rc(other[60:])))]
    for mode in (ea.Mode.SEMIGLOBAL, ea.Mode.LOCAL):
        gpu = ea.GpuScorer(panel, ea.Scoring(*DEFAULT), mode)
        for kw in (dict(), dict(both_strands=True), dict(min_score=30), dict(min_score=30, max_ties=3),
                   dict(both_strands=True, max_ties=20)):
            compare_pairs(gpu, panel.codes, pairs, DEFAULT, mode == ea.Mode.LOCAL, **kw)
        # (a floor makes the FF pair's reversed mate "not aligned": Separate)
        got, _ = dev_pairs(gpu, pairs, min_score=30)
        assert got[0][0] == "proper" and got[0][2] == 40 and len(got[0][4]) == 40
        assert got[1][0] == "separate" and got[1][1][1] == 40 and len(got[1][1][3]) == 40


def test_map_pairs_paired_workload_sample():
    """500 references of 70-95 nt, 1,500 pairs of 150-nt mates from inserts
    N(120, 30) that read into the TruSeq adapters, mates trimmed, semi-global
    2,-1,-10,-1, both template strands and forward only: equal to the oracle
    (its rows from the SIMD restatement, pinned to the scalar DP)."""
    panel_l = synth.panel(seed=0xC0F4, n_refs=500, lo=70, hi=95)
    codes, off, lens = synth.pair_arrays(panel_l, seed=0xC0F5, n_pairs=1500)
    letters = [bytes(b"ACGTN"[c] for c in codes[int(o):int(o) + int(n)]) for o, n in zip(off, lens)]
    pairs = []
    for k in range(0, len(letters), 2):
        a, b = letters[k], letters[k + 1]
        ad = ea.adapter_detect_pair(a, b)
        pairs.append((ea.encode(a[:len(a) - ad[0]]), ea.encode(b[:len(b) - ad[1]])))
    sup = O.simd_supported()
    prof = O.Profile(panel_l, DEFAULT, 32 if sup & 2 else 16)
    mates = [x for p in pairs for x in p]
    fwd = prof.score_batch(*O.concat(mates), False, 8)
    rev = prof.score_batch(*O.concat([O.revcomp(x) for x in mates]), False, 8)
    gpu = ea.GpuScorer(ea.Panel(panel_l), ea.Scoring(*DEFAULT), ea.Mode.SEMIGLOBAL)
    for kw in (dict(), dict(both_strands=True)):
        got, stats = dev_pairs(gpu, pairs, **kw)
        want = O.map_pairs(pairs, panel_l, DEFAULT, False, rows=(fwd, rev), **kw)
        for k, (g, w) in enumerate(zip(got, want)):
            assert g == w, (k, kw)
        kinds = [g[0] for g in got]
        assert kinds.count("proper") > len(pairs) // 2


def test_map_reads_longer_than_the_drop_in_bound():
    """Reads of 4,500-6,000 nt (over the reference's 4,096-nt GPU bound,
    cli/align.rs:1751-1774 keeps them on the CPU) mapped whole on the
    device: scored by the int16 path at that length, their winners traced
    by the anti-diagonal kernel; both modes, both strands, equal to the
    oracle. Under 40,-20,-60,-20 their scores pass the int16 bound and take
    the int32 kernel. Cut into chunks of 8 reads, the later ones scored in
    slices with those kernels after the last, and with every long read sent
    to the anti-diagonal kernel by length, the records are the same."""
    rng = G.Rng(0x5EED0207)
    panel_l = [G.random_seq(rng, 200 + rng.below(800), 0.01) for _ in range(8)]
    reads = [G.encode(G.chimeric_read(rng, panel_l, 4500 + rng.below(1500))) for _ in range(16)]
    reads += [G.encode(G.mutate(rng, panel_l[k % 8])) for k in range(40)]
    panel = ea.Panel(panel_l)
    for scoring, mode in ((DEFAULT, ea.Mode.LOCAL), ((1, -1, -2, -1), ea.Mode.SEMIGLOBAL),
                          ((40, -20, -60, -20), ea.Mode.SEMIGLOBAL)):
        gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), mode)
        got = dev_reads(gpu, reads, both_strands=True)
        want = O.map_reads(reads, panel.codes, scoring, mode == ea.Mode.LOCAL, both_strands=True)
        for k, (g, w) in enumerate(zip(got, want)):
            assert g == w, (k, len(reads[k]))
        assert all(g[3] for g in got[:16])  # every long read maps
        if scoring[0] == 40:  # every long read over the int16 bound
            b = gpu.batch(16, sum(len(r) for r in reads[:16]))
            b.load(reads[:16], both_strands=True)
            b.run(ea.RUN_SCORES32)
            b.wait()
            _, n16, n32 = b.stats()
            del b
            assert n32 == 32 and n16 == 0, (n16, n32)
        gpu.set_map_chunk(8)
        assert dev_reads(gpu, reads, both_strands=True) == got, scoring
        gpu.set_map_chunk(0)
        gpu.diag_min_len = 1000  # every long read on the anti-diagonal kernel by length
        assert dev_reads(gpu, reads, both_strands=True) == got, scoring
        gpu.diag_min_len = 0


def test_default_chunking_equals_one_chunk():
    """Calls over more than the default chunk (2^17 reads, 2^16 pairs): each
    chunk after the first scores in slices with the previous chunk's
    selection, rounds and tracebacks between them, and the last chunk's
    worth is cut in halves. Every record, hit and run equals the same call
    in one chunk (scored whole, unsliced), and the last reads' records equal
    the oracle's."""
    panel_l = synth.panel(seed=0x0123456789ABCDEF, n_refs=164, lo=140, hi=160)
    codes, off, lens = synth.read_arrays(panel_l, seed=0xC0F6, n=300_000)
    scoring = (1, -1, -2, -1)
    gpu = ea.GpuScorer(ea.Panel(panel_l), ea.Scoring(*scoring), ea.Mode.SEMIGLOBAL)
    cut = gpu.map_reads_concat(codes, off, lens, both_strands=True)[:3]
    gpu.set_map_chunk(len(lens))
    whole = gpu.map_reads_concat(codes, off, lens, both_strands=True)[:3]
    gpu.set_map_chunk(0)
    for a, b in zip(cut, whole):
        assert len(a) == len(b) and np.array_equal(a, b)
    tail = range(len(lens) - 400, len(lens))
    reads = [codes[int(off[k]):int(off[k]) + int(lens[k])] for k in tail]
    want = O.map_reads(reads, panel_l, scoring, False, both_strands=True)
    recs, hits, ops = cut
    for k, w in zip(tail, want):
        assert O.read_form(recs[k], hits, ops) == w, k

    ppanel = synth.panel(seed=0xC0F4, n_refs=500, lo=70, hi=95)
    pc, po, pl = synth.pair_arrays(ppanel, seed=0xC0F7, n_pairs=100_000)
    pg = ea.GpuScorer(ea.Panel(ppanel), ea.Scoring(*DEFAULT), ea.Mode.SEMIGLOBAL)
    for kw in (dict(), dict(both_strands=True)):
        cut = pg.map_pairs_concat(pc, po, pl, **kw)[:3]
        pg.set_map_chunk(2 * len(pl))
        whole = pg.map_pairs_concat(pc, po, pl, **kw)[:3]
        pg.set_map_chunk(0)
        for a, b in zip(cut, whole):
            assert len(a) == len(b) and np.array_equal(a, b), kw


def test_map_reads_with_swapped_chunks():
    """Chunks scored with the roles swapped (at most swap_rows rows each, so
    every chunk takes that path; a swapped run is never sliced): every
    record equals the oracle's and the same call unchunked and unswapped,
    the reverse strand's alignments included (the tracebacks read the
    reverse complements from the batch's codes); pairs on one and both
    template strands equal the oracle the same way."""
    rng = G.Rng(0x5EED_0C0A)
    panel_l = G.random_panel(rng, 200, 60, 180)
    L = max(map(len, panel_l))
    # swapped, a read is a panel reference: none longer than the panel's longest
    reads = [G.encode(r) for r in G.reads_for(rng, panel_l, 1000) if len(r) <= L]
    panel = ea.Panel(panel_l)
    scoring = (1, -1, -2, -1)
    for mode in (ea.Mode.SEMIGLOBAL, ea.Mode.LOCAL):
        gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), mode)
        plain = dev_reads(gpu, reads, both_strands=True, max_ties=3)
        gpu.swap_rows = 500
        gpu.set_map_chunk(200)  # 400 rows a chunk on both strands
        swapped = dev_reads(gpu, reads, both_strands=True, max_ties=3)
        gpu.set_map_chunk(0)
        gpu.swap_rows = 0
        assert swapped == plain, mode
        want = O.map_reads(reads, panel.codes, scoring, mode == ea.Mode.LOCAL, both_strands=True, max_ties=3)
        for k, (g, w) in enumerate(zip(swapped, want)):
            assert g == w, (mode, k)
        pairs = [(reads[2 * k], reads[2 * k + 1]) for k in range(len(reads) // 2)]
        gpu.swap_rows = 500
        gpu.set_map_chunk(100)  # 200 rows a chunk: 100 pairs on one strand, 50 on both
        for kw in (dict(), dict(both_strands=True)):
            compare_pairs(gpu, panel.codes, pairs, scoring, mode == ea.Mode.LOCAL, **kw)
        gpu.set_map_chunk(0)
        gpu.swap_rows = 0


def test_mapped_arrays_outlive_the_scorer_and_each_other():
    """The mirror's arrays view the library's result buffers: they stay whole
    after the scorer is freed and a collection runs, two calls share no
    memory, and writing one call's arrays changes nothing in the other's."""
    import gc

    rng = G.Rng(0x5EED0213)
    panel_l = G.random_panel(rng, 12, 40, 120)
    reads = [G.encode(r) for r in G.reads_for(rng, panel_l, 300)]
    panel = ea.Panel(panel_l)
    gpu = ea.GpuScorer(panel, ea.Scoring(*DEFAULT), ea.Mode.LOCAL)
    a = gpu.map_reads_raw(reads, both_strands=True)
    b = gpu.map_reads_raw(reads, both_strands=True)
    del gpu
    gc.collect()
    want = O.map_reads(reads, panel.codes, DEFAULT, True, both_strands=True)
    for recs, hits, ops, _ in (a, b):
        assert recs.flags.writeable and hits.flags.writeable
        got = [O.read_form(r, hits, ops) for r in recs]
        assert got == want
    for x, y in zip(a[:3], b[:3]):
        assert not np.shares_memory(x, y)
    recs_a, hits_a = a[0], a[1]
    del a
    gc.collect()
    assert [O.read_form(r, b[1], b[2]) for r in b[0]] == want
    recs_a[:] = np.zeros(1, recs_a.dtype)
    hits_a[:] = np.zeros(1, hits_a.dtype)
    assert [O.read_form(r, b[1], b[2]) for r in b[0]] == want


def test_mapping_calls_with_no_reads():
    """A call with no reads (no pairs) returns empty arrays and leaves the
    scorer ready for the next call."""
    rng = G.Rng(0x5EED0214)
    panel_l = G.random_panel(rng, 6, 40, 80)
    panel = ea.Panel(panel_l)
    gpu = ea.GpuScorer(panel, ea.Scoring(*DEFAULT), ea.Mode.SEMIGLOBAL)
    recs, hits, ops, st = gpu.map_reads_raw([], both_strands=True)
    assert len(recs) == 0 and len(hits) == 0 and int(st["traced"]) == 0
    precs, phits, pops, pst = gpu.map_pairs_raw([], [], both_strands=True)
    assert len(precs) == 0 and len(phits) == 0
    reads = [G.encode(r) for r in G.reads_for(rng, panel_l, 40)]
    assert dev_reads(gpu, reads) == O.map_reads(reads, panel.codes, DEFAULT, False)
