"""The oracle's traceback (eo_align, lib/scalar.rs:164-364) pinned against
what fixes it independently of its own code:

- the 192 parasail 1.3.4 results of align_golden.json: score AND the start
  and end coordinates (tests/parasail_golden.rs:48-77);
- the CIGAR checks of tests/parasail_golden.rs:82-110 (the ops re-score to
  the score, span exactly the coordinates, start and end with a match);
- the known answers of lib/scalar.rs:378-430;
- a full-matrix restatement of the documented choices (lib/scalar.rs:28-37:
  end cell = first maximum of a column-major scan over the mode's
  candidates; traceback prefers diagonal, then insertion, then deletion;
  inside a gap, extending over opening; local stops at the first H = 0),
  written from those rules with whole H/E/F matrices, on every pair of
  sequences over ACGTN up to length 3 and random pairs up to 7.
"""
import itertools
import random

import numpy as np
import pytest

from oracle import pyoracle as O
import refgen as G

NEG = -(2 ** 31) // 4


def sub(scoring, a, b):
    ma, mi, _, _ = scoring
    return ma if (a == b or a == 4 or b == 4) else mi


def rescore(q, r, scoring, qs, rs, ops):
    """Alignment::rescore (lib/scalar.rs:91-116)."""
    _, _, o, e = scoring
    i, j, total = qs, rs, 0
    for op, n in ops:
        if op == "M":
            for _ in range(n):
                total += sub(scoring, q[i], r[j])
                i += 1
                j += 1
        else:
            total += o + (n - 1) * e
            if op == "I":
                i += n
            else:
                j += n
    return total, i, j


def check_structure(q, r, scoring, a, ends_in_match=True):
    """The ops re-score to the score and span the coordinates; on the golden
    schemes they also start and end with a match (with a zero-cost gap, as
    2,-1,0,0 has, an optimal path may legitimately start with one)."""
    score, qs, qe, rs, re, ops = a
    if not ops:
        return
    got, i, j = rescore(q, r, scoring, qs, rs, ops)
    assert got == score
    assert (i, j) == (qe, re)
    if ends_in_match:
        assert ops[0][0] == "M" and ops[-1][0] == "M"


def test_align_known_answers():
    """lib/scalar.rs:378-430."""
    enc = O.encode
    d = (2, -1, -10, -1)
    a = O.align(enc(b"ACGTACGT"), enc(b"TTACGTACGTTT"), d, True)
    assert a == (16, 0, 8, 2, 10, [("M", 8)])
    s = (2, -4, -3, -1)
    q, r = enc(b"AAAACCCGGGGTTTT"), enc(b"AAAACCCCGGGGTTTT")
    a = O.align(q, r, s, True)
    assert a[0] == 15 * 2 - 3 and ("D", 1) in a[5]
    check_structure(q, r, s, a)
    a = O.align(enc(b"ACGTACGTCCCCC"), enc(b"GGGGGACGTACGT"), d, False)
    assert a[:5] == (16, 0, 8, 5, 13)
    a = O.align(enc(b"AAAA"), enc(b"CCCC"), d, True)
    assert a == (0, 0, 0, 0, 0, [])
    # empty inputs (lib/scalar.rs:185-195)
    assert O.align(enc(b""), enc(b"ACGT"), d, True)[:5] == (0, 0, 0, 0, 0)
    assert O.align(enc(b"ACGT"), enc(b""), d, False)[:5] == (0, 0, 0, 0, 0)


def test_align_matches_parasail_golden():
    """tests/parasail_golden.rs:48-110: every one of the 192 results."""
    g = G.golden()
    schemes = g["schemes"]
    n = 0
    for p in g["pairs"]:
        q, r = O.encode(p["query"].encode()), O.encode(p["reference"].encode())
        for res in p["results"]:
            scoring = tuple(schemes[res["scheme"]])
            a = O.align(q, r, scoring, res["mode"] == "local")
            label = (p["name"], res["mode"], res["scheme"])
            assert a[0] == res["score"], label
            assert a[1:5] == (res["query_start"], res["query_end"], res["ref_start"],
                              res["ref_end"]), label
            assert a[0] == O.score(q, r, scoring, res["mode"] == "local"), label
            check_structure(list(q), list(r), scoring, a)
            n += 1
    assert n == 192


def matrix_align(q, r, scoring, local, prefer_del=False, prefer_open=False):
    """The documented choices (lib/scalar.rs:28-37) on whole matrices.
    prefer_del / prefer_open plant the two plausible mistakes (deletion
    before insertion; opening a gap before extending it) for the check that
    the equality tests can see them."""
    _, _, o, e = scoring
    n, m = len(q), len(r)
    if n == 0 or m == 0:
        return (0, 0, 0, 0, 0, [])
    H = [[0] * (m + 1) for _ in range(n + 1)]
    E = [[NEG] * (m + 1) for _ in range(n + 1)]
    F = [[NEG] * (m + 1) for _ in range(n + 1)]
    for i in range(1, n + 1):
        for j in range(1, m + 1):
            E[i][j] = max(H[i][j - 1] + o, E[i][j - 1] + e)
            F[i][j] = max(H[i - 1][j] + o, F[i - 1][j] + e)
            h = max(H[i - 1][j - 1] + sub(scoring, q[i - 1], r[j - 1]), E[i][j], F[i][j])
            H[i][j] = max(h, 0) if local else h
    # end cell: first maximum, column-major, over the mode's candidates
    best, bi, bj = (0 if local else NEG), 0, 0
    for j in range(1, m + 1):
        for i in range(1, n + 1):
            cand = local or i == n or j == m
            if cand and H[i][j] > best:
                best, bi, bj = H[i][j], i, j
    if bi == 0:
        return (0, 0, 0, 0, 0, [])
    i, j, state, rev = bi, bj, "H", []
    while i > 0 and j > 0:
        if state == "H":
            h = H[i][j]
            if local and h == 0:
                break
            if h == H[i - 1][j - 1] + sub(scoring, q[i - 1], r[j - 1]):
                rev.append("M")
                i, j = i - 1, j - 1
            elif prefer_del and h == E[i][j]:
                state = "E"
            elif h == F[i][j]:
                state = "F"
            else:
                state = "E"
        elif state == "F":
            rev.append("I")
            if F[i][j] != F[i - 1][j] + e or (prefer_open and F[i][j] == H[i - 1][j] + o):
                state = "H"
            i -= 1
        else:
            rev.append("D")
            if E[i][j] != E[i][j - 1] + e or (prefer_open and E[i][j] == H[i][j - 1] + o):
                state = "H"
            j -= 1
    ops = []
    for op in reversed(rev):
        if ops and ops[-1][0] == op:
            ops[-1] = (op, ops[-1][1] + 1)
        else:
            ops.append((op, 1))
    return (best, i, bi, j, bj, ops)


SCHEMES = [(2, -1, -10, -1), (1, -1, -2, -1), (2, -4, -3, -1), (3, 1, -5, -2), (2, -1, 0, 0)]


def test_align_equals_matrix_restatement_exhaustive():
    """Every pair of sequences over ACGTN of lengths 1-3 (and 0), 5 schemes,
    both modes: identical alignment, coordinates and ops."""
    seqs = [()]
    for k in (1, 2, 3):
        seqs += list(itertools.product(range(5), repeat=k))
    rng = random.Random(7)
    pairs = [(a, b) for a in seqs for b in seqs if len(a) + len(b) <= 5]
    pairs += [(rng.choice(seqs), rng.choice(seqs)) for _ in range(3000)]
    for scoring in SCHEMES:
        for local in (True, False):
            for a, b in pairs:
                want = matrix_align(list(a), list(b), scoring, local)
                got = O.align(np.array(a, np.uint8), np.array(b, np.uint8), scoring, local)
                assert got == want, (a, b, scoring, local, got, want)


def test_align_equals_matrix_restatement_random():
    rng = random.Random(11)
    for _ in range(4000):
        n, m = rng.randint(1, 7), rng.randint(1, 7)
        q = [rng.choice([0, 1, 2, 3, 0, 1, 2, 3, 4]) for _ in range(n)]
        r = [rng.choice([0, 1, 2, 3, 0, 1, 2, 3, 4]) for _ in range(m)]
        scoring = rng.choice(SCHEMES)
        local = rng.random() < 0.5
        want = matrix_align(q, r, scoring, local)
        got = O.align(np.array(q, np.uint8), np.array(r, np.uint8), scoring, local)
        assert got == want, (q, r, scoring, local)
        check_structure(q, r, scoring, got, ends_in_match=False)


def test_matrix_restatement_catches_planted_mistakes():
    """The restatement's tie choices are not vacuous: planting either
    plausible traceback mistake (deletion before insertion, or opening before
    extending) in it changes the answer on inputs the exhaustive test covers,
    so an oracle making that mistake would fail the equality."""
    found = {"del": False, "open": False}
    seqs = [t for k in (2, 3, 4) for t in itertools.product(range(4), repeat=k)]
    rng = random.Random(5)
    for scoring in [(2, -1, -1, -1), (1, -1, -2, -1), (2, -1, -10, -1)]:
        for _ in range(20000):
            q, r = list(rng.choice(seqs)), list(rng.choice(seqs))
            local = rng.random() < 0.5
            base = matrix_align(q, r, scoring, local)
            if not found["del"] and matrix_align(q, r, scoring, local, prefer_del=True) != base:
                found["del"] = True
            if not found["open"] and matrix_align(q, r, scoring, local, prefer_open=True) != base:
                found["open"] = True
            if all(found.values()):
                break
    assert found == {"del": True, "open": True}


def _simd_vs_scalar(queries, panel, refs, scoring, local):
    qc, qo, ql = O.concat(queries)
    recs, ops, off = O.simd_align_batch_concat(qc, qo, ql, panel, np.array(refs, np.int32), scoring, local, 4,
                                               with_ops=True)
    for k, (q, r) in enumerate(zip(queries, refs)):
        want = O.align(q, panel[r], scoring, local)
        runs = [(O.OPS[int(x) >> 30], int(x) & 0x3FFFFFFF) for x in ops[int(off[k]):int(off[k]) + recs[k][5]]]
        assert tuple(int(v) for v in recs[k][:5]) + (runs,) == tuple(want[:5]) + (want[5],), (k, scoring, local)


def test_simd_pair_traceback_equals_scalar_align():
    """The AVX-512BW pair traceback (simd::align_pairs, lib/simd/avx512.rs:
    120-247; the CPU baseline of the device tracebacks) equals eo_align
    pair for pair, score, coordinates and every run, as the reference's own
    lib/simd/mod.rs:869-912 checks: random pairs over five schemes and both
    modes (wildcards, reads 1-700 nt, references 1-500 nt, chunks that mix
    lane groups), empty reads (the scalar path), and a pair outside the
    i16 bound (the scalar path)."""
    if not O.simd_supported() & 2:
        pytest.skip("no AVX-512BW on this host: every pair takes the scalar path")
    rng = G.Rng(0x5EED_0C01)
    panel = [G.encode(G.random_seq(rng, 1 + rng.below(500), 0.03)) for _ in range(30)]
    queries = [G.encode(G.random_seq(rng, 1 + rng.below(700), 0.03)) for _ in range(300)]
    queries += [np.zeros(0, np.uint8), G.encode(G.random_seq(rng, 5, 0.0))]
    refs = [rng.below(len(panel)) for _ in queries]
    for scoring in G.SCHEMES_GPU:
        for local in (True, False):
            _simd_vs_scalar(queries, panel, refs, scoring, local)
    # outside fits_i16 at 1000,-1000,-1000,-1000: every pair on the scalar path
    _simd_vs_scalar(queries[:20], panel, refs[:20], (1000, -1000, -1000, -1000), True)


def test_simd_pair_traceback_on_the_golden_pairs():
    """The 192 parasail results through the pair kernel (both schemes, both
    modes): equal to eo_align, which they pin."""
    if not O.simd_supported() & 2:
        pytest.skip("no AVX-512BW on this host")
    g = G.golden()
    for scoring in ((2, -1, -10, -1), (1, -1, -2, -1)):
        qs = [G.encode(p["query"].encode()) for p in g["pairs"]]
        rs = [G.encode(p["reference"].encode()) for p in g["pairs"]]
        for local in (True, False):
            _simd_vs_scalar(qs, rs, list(range(len(rs))), scoring, local)
