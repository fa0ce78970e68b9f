"""The benchmark configurations as parity cases on an MI355X: each one's
synthetic workload (at a size the CPU checker finishes in seconds) scored on
the device and compared with the CPU kernels, by equality.

The checker for the large panels is the oracle's AVX-512/AVX2 restatement of
the reference's CPU kernel, itself pinned to the scalar oracle in
tests/test_oracle.py; the scalar oracle checks a sample directly."""
import os

import numpy as np
import pytest

import eschalign_amd as ea
from eschalign_amd import synth
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def simd_rows(panel_codes, queries, scoring, local):
    sup = O.simd_supported()
    lanes = 32 if sup & 2 else 16
    prof = O.Profile(panel_codes, scoring, lanes)
    qc, qo, ql = O.concat(queries)
    return prof.score_batch(qc, qo, ql, local, THREADS)


def test_config1_plain_cpu_reference_path():
    """164 refs of 140-160 nt, 10k reads of 60-150 nt, local 2,-1,-10,-1,
    forward strand only: every score equals the CPU kernel; a sample equals
    the scalar oracle."""
    panel = synth.panel(seed=0x0123456789ABCDEF, n_refs=164, lo=140, hi=160)
    reads = synth.reads(panel, seed=0xC0F1, n=10_000, length=("uniform", 60, 150))
    scoring = (2, -1, -10, -1)
    gpu = ea.GpuScorer(ea.Panel(panel), ea.Scoring(*scoring), ea.Mode.LOCAL)
    got, ok = gpu.score_batch_i32(reads)
    assert ok.all()
    assert np.array_equal(got, simd_rows(panel, reads, scoring, True))
    assert np.array_equal(got[:200], O.score_matrix(reads[:200], panel, scoring, True, THREADS))
    m = gpu.score_batch(reads)          # the drop-in i16 matrix agrees
    assert m.n_scored() == len(reads) and np.array_equal(m.scores.astype(np.int32), got)


def test_config2_batched_uneven_lengths():
    """The benchmark's own workload at 24,000 reads: the 164-ref panel of
    140-160 nt (seed 0x0123456789ABCDEF), reads from synth.read_arrays with
    seed 0x5EED_2002 and lognormal(90, 0.6) lengths clipped 20-1000, both
    strands in one device batch, semiglobal 1,-1,-2,-1. The timed batch's
    own int32 rows (as bench.py reads them) equal the CPU kernel for the
    diagonal-shift variants 24, 25 (fp16, the default and the autotuner's
    picks here) and 28 (int16), the row-shifted fp16 variant 7, the int16
    variant 6 and the plain one-thread-per-pair kernel, eagerly and over
    three graph replays; the
    fused selection equals the oracle's on those rows; a sample equals the
    scalar oracle."""
    panel = synth.panel(seed=0x0123456789ABCDEF, n_refs=164, lo=140, hi=160)
    n = 24_000
    codes, off, lens = synth.read_arrays(panel, seed=0x5EED_2002, n=n,
                                         length=("lognormal", 90, 0.6, 20, 1000))
    reads = [codes[int(off[k]):int(off[k]) + int(lens[k])] for k in range(n)]
    rows = reads + [ea.reverse_complement_codes(q) for q in reads]
    scoring = (1, -1, -2, -1)
    want = simd_rows(panel, rows, scoring, False)
    assert np.array_equal(want[:150], O.score_matrix(rows[:150], panel, scoring, False, THREADS))
    assert np.array_equal(want[n:n + 150], O.score_matrix(rows[n:n + 150], panel, scoring, False, THREADS))
    sel_want = [O.select(want[k], want[n + k], 0, None) for k in range(n)]
    gpu = ea.GpuScorer(ea.Panel(panel), ea.Scoring(*scoring), ea.Mode.SEMIGLOBAL)
    batch = gpu.batch(n, int(lens.sum()))
    batch.load_concat(codes, off, lens, both_strands=True)
    tie_cap = 8

    def check(what, label):
        batch.run(what, tie_cap=tie_cap)
        batch.wait()
        got, ok = batch.scores32()
        assert ok.all(), label
        bad = np.nonzero((got != want).any(axis=1))[0]
        assert bad.size == 0, f"{label}: {bad.size} rows differ, first {bad[:5]}"
        sel = batch.select()
        for k in range(n):
            b, nt, so, take = sel_want[k]
            s = sel[k]
            assert (s.best, s.n_tied, s.suboptimal, s.take) == (b, nt, so, take[:tie_cap]), f"{label}: read {k}"

    base = ea.RUN_SCORES32 | ea.RUN_SELECT
    # the diagonal-shift variants the headline runs (fp16 24, 25 and their
    # int16 counterpart 28), and the row-shifted 7 (fp16) and 6 (int16)
    assert gpu.variant == 24  # the default on this panel
    for v in (24, 25, 28, 7, 6):
        gpu.variant = v
        check(base | ea.RUN_NO_GRAPH, f"variant {v} eager")
        assert batch.variant() == v
        for rep in range(3):
            check(base, f"variant {v} graph replay {rep}")
    check(base | ea.RUN_PLAIN | ea.RUN_NO_GRAPH, "plain kernel")


def test_config3_large_panel_near_ties():
    """A large panel of seeds plus 4 copies each at 2% substitutions (ties
    and near-ties common), refs 200-1000 nt, reads 100-1000 nt, local
    2,-1,-10,-1: scores equal the CPU kernel and the fused tie sets equal
    the oracle's selection on them."""
    panel = synth.panel_with_copies(seed=0xC0F3, n_seeds=120, copies=4, lo=200, hi=1000)
    reads = synth.reads(panel, seed=0xC0F3 + 1, n=300, length=("uniform", 100, 1000))
    scoring = (2, -1, -10, -1)
    gpu = ea.GpuScorer(ea.Panel(panel), ea.Scoring(*scoring), ea.Mode.LOCAL)
    got, ok = gpu.score_batch_i32(reads)
    want = simd_rows(panel, reads, scoring, True)
    assert ok.all() and np.array_equal(got, want)
    sel = gpu.select_batch(reads, both_strands=False, tie_cap=64)
    ties_seen = 0
    for k in range(len(reads)):
        b, nt, so, take = O.select(want[k], None, 0, None)
        ties_seen += nt > 1
        assert (sel[k].best, sel[k].n_tied, sel[k].suboptimal, sel[k].take) == (b, nt, so, take[:64])
    assert ties_seen > 0, "the paralog panel should produce ties"


def test_config3_at_scale_10k_references():
    """The large-panel configuration's own panel (10,000 references of
    200-1000 nt: 2,000 seeds and 4 copies of each at 2% substitutions) with
    400 of its reads (100-1000 nt), local 2,-1,-10,-1: 2,500 reference
    groups per chunk and the longest-first reference order at scale. The
    ordinary path (swapping off), the swapped-roles path a batch of at most
    one row per reference takes by default, and the autotuner's tuned tile
    shape and plan all give the CPU kernel's scores; the fused tie sets
    equal the oracle's selection."""
    panel = synth.panel_with_copies(seed=0xC0F3, n_seeds=2000, copies=4, lo=200, hi=1000)
    assert len(panel) == 10_000
    codes, off, lens = synth.read_arrays(panel, seed=0xC0F3 + 1, n=400, length=("uniform", 100, 1000))
    reads = [codes[int(off[k]):int(off[k]) + int(lens[k])] for k in range(400)]
    scoring = (2, -1, -10, -1)
    want = simd_rows(panel, reads, scoring, True)
    gpu = ea.GpuScorer(ea.Panel(panel), ea.Scoring(*scoring), ea.Mode.LOCAL)
    assert gpu.swap_rows == 10_000
    got, ok = gpu.score_batch_i32(reads)            # 400 rows: the roles swapped
    assert ok.all() and np.array_equal(got, want)
    gpu.swap_rows = 0
    got, ok = gpu.score_batch_i32(reads)            # the ordinary path
    assert ok.all() and np.array_equal(got, want)
    chosen, ms = gpu.autotune(reads[:200], both_strands=False)
    assert ms[chosen] > 0
    got, ok = gpu.score_batch_i32(reads)
    assert ok.all() and np.array_equal(got, want), (chosen, gpu.diag_plan)
    sel = gpu.select_batch(reads, both_strands=False, tie_cap=64)
    tied = 0
    for k in range(len(reads)):
        b, nt, so, take = O.select(want[k], None, 0, None)
        tied += nt > 1
        assert (sel[k].best, sel[k].n_tied, sel[k].suboptimal, sel[k].take) == (b, nt, so, take[:64]), k
    assert tied > 0


def test_config4_paired_with_adapter_read_through():
    """500 refs of 70-95 nt, mates of 150 nt from inserts N(120, 30) reading
    into the TruSeq adapters: mates trimmed by the library's adapter
    detection (checked against the oracle), scored on both orientations,
    and the device's top pair tier equal to the oracle's candidates."""
    panel = synth.panel(seed=0xC0F4, n_refs=500, lo=70, hi=95)
    codes, off, lens = synth.pair_arrays(panel, seed=0xC0F4 + 1, n_pairs=400)
    letters = [bytes(b"ACGTN"[c] for c in codes[int(o):int(o) + int(n)]) for o, n in zip(off, lens)]
    m1, m2 = [], []
    for k in range(0, len(letters), 2):
        a, b = letters[k], letters[k + 1]
        ad = ea.adapter_detect_pair(a, b)
        assert ad == O.adapter_detect_pair(a, b, ea.TRUSEQ_R1, ea.TRUSEQ_R2)
        m1.append(ea.encode(a[:len(a) - ad[0]]))
        m2.append(ea.encode(b[:len(b) - ad[1]]))
    scoring = (2, -1, -10, -1)
    gpu = ea.GpuScorer(ea.Panel(panel), ea.Scoring(*scoring), ea.Mode.SEMIGLOBAL)
    rows = {}
    for name, mates in (("m1", m1), ("m2", m2)):
        rc = [ea.reverse_complement_codes(q) for q in mates]
        rows[name] = (simd_rows(panel, mates, scoring, False), simd_rows(panel, rc, scoring, False))
    for both in (False, True):
        got = gpu.select_pairs(m1, m2, both_strands=both, min_score=0, top_cap=32)
        for k in range(len(m1)):
            r = [[rows["m1"][0][k], rows["m1"][1][k]], [rows["m2"][0][k], rows["m2"][1][k]]]
            cands = O.pair_cands(r, 0, both)
            g = got[k]
            assert g.n_cand == len(cands)
            if not cands:
                continue
            top = cands[0][0]
            tier = [(c[1], c[2]) for c in cands if c[0] == top]
            assert g.top_sum == top and g.n_top == len(tier) and g.top == tier[:32], k
