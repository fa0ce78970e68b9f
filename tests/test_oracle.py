"""Pin the CPU oracle before trusting it (CPU only).

* eo_score against the reference's 192 parasail 1.3.4 golden results, the
  closed forms of lib/scalar.rs:378-430, and exhaustive enumeration of every
  alignment on tiny inputs (a dropped term, a wrong sign, an off-by-one in
  the borders or a mis-ordered max each fail one of these);
* the SIMD restatement (CPU baseline) against eo_score on the reference's
  own random suites, inputs regenerated bit for bit;
* selection, pair candidates and adapter detection against the reference's
  closed-form tests.
"""
import itertools

import numpy as np
import pytest

from oracle import pyoracle as O
import refgen as G


def sc_name(s):
    return ",".join(map(str, s))


# ---------------------------------------------------------------- alphabet

def test_alphabet_known_answers():
    # lib/alphabet.rs:86-105
    assert list(O.encode(b"ACGTUacgtu")) == [0, 1, 2, 3, 3, 0, 1, 2, 3, 3]
    for b in b"NnRYKMSWBDHV-*.X":
        assert list(O.encode(bytes([b]))) == [4]
    assert list(O.revcomp(O.encode(b"ACGTN"))) == list(O.encode(b"NACGT"))
    # the test helper's table encoder agrees with the oracle's on every byte
    every = bytes(range(256))
    assert np.array_equal(G.encode(every), O.encode(every))


def test_scoring_validate_matches_reference_rules():
    # lib/scoring.rs:192-213 refuses_nonsense, and the rule order of :70-100
    assert O.validate(2, -1, -10, -1) == 0
    assert O.validate(2, -1, 10, -1) != 0
    assert O.validate(0, -1, -10, -1) != 0
    assert O.validate(2, 3, -10, -1) != 0
    assert O.validate(1001, -1, -10, -1) != 0
    assert O.validate(2, -1, 0, 0) == 0
    assert O.validate(3, 1, -5, -2) == 0


# ---------------------------------------------------------------- golden

def golden_cases():
    g = G.golden()
    schemes = g["schemes"]
    for p in g["pairs"]:
        for res in p["results"]:
            yield (p["name"], p["query"].encode(), p["reference"].encode(),
                   tuple(schemes[res["scheme"]]), res["mode"], res["score"])


def test_score_matches_parasail_golden():
    """tests/parasail_golden.rs:48-77: every one of the 192 results."""
    cases = list(golden_cases())
    assert len(cases) == 192
    for name, q, r, scoring, mode, want in cases:
        got = O.score(O.encode(q), O.encode(r), scoring, mode == "local")
        assert got == want, f"{name} {mode} {scoring}"


def test_closed_forms():
    # lib/scalar.rs:378-430
    d = (2, -1, -10, -1)
    assert O.score(O.encode(b"ACGTACGT"), O.encode(b"TTACGTACGTTT"), d, True) == 16
    assert O.score(O.encode(b"ACGTA"), O.encode(b"ACNTA"), d, True) == 10
    assert O.score(O.encode(b"AAAACCCGGGGTTTT"), O.encode(b"AAAACCCCGGGGTTTT"),
                   (2, -4, -3, -1), True) == 15 * 2 - 3
    assert O.score(O.encode(b"ACGTACGTCCCCC"), O.encode(b"GGGGGACGTACGT"), d, False) == 16
    assert O.score(O.encode(b"AAAA"), O.encode(b"CCCC"), d, True) == 0
    # empty inputs score 0 in both modes (lib/scalar.rs:123-125)
    assert O.score(b"", O.encode(b"ACGT"), d, False) == 0
    assert O.score(O.encode(b"ACGT"), b"", d, True) == 0
    # a one-base semiglobal mismatch has a negative best
    assert O.score(O.encode(b"A"), O.encode(b"C"), d, False) == -1


def _global_best(A, B, scoring):
    """Max over every alignment path of A against B (exhaustive)."""
    ma, mi, o, e = scoring
    best = [None]

    def rec(i, j, last, acc):
        if i == len(A) and j == len(B):
            if best[0] is None or acc > best[0]:
                best[0] = acc
            return
        if i < len(A) and j < len(B):
            a, b = A[i], B[j]
            s = ma if (a == b or a == 4 or b == 4) else mi
            rec(i + 1, j + 1, "M", acc + s)
        if i < len(A):
            rec(i + 1, j, "I", acc + (e if last == "I" else o))
        if j < len(B):
            rec(i, j + 1, "D", acc + (e if last == "D" else o))

    rec(0, 0, None, 0)
    return best[0]


def brute_local(q, r, scoring):
    cache = {}
    best = 0
    for i0 in range(len(q) + 1):
        for i1 in range(i0, len(q) + 1):
            for j0 in range(len(r) + 1):
                for j1 in range(j0, len(r) + 1):
                    key = (tuple(q[i0:i1]), tuple(r[j0:j1]))
                    if key not in cache:
                        cache[key] = _global_best(*key, scoring)
                    best = max(best, cache[key])
    return best


def brute_semiglobal(q, r, scoring):
    n, m = len(q), len(r)
    starts = [(i, 0) for i in range(n + 1)] + [(0, j) for j in range(1, m + 1)]
    ends = set([(n, j) for j in range(1, m + 1)] + [(i, m) for i in range(1, n + 1)])
    best = None
    for i0, j0 in starts:
        for i1, j1 in ends:
            if i1 < i0 or j1 < j0 or (i1, j1) == (i0, j0):
                continue
            v = _global_best(tuple(q[i0:i1]), tuple(r[j0:j1]), scoring)
            best = v if best is None else max(best, v)
    return best


def test_score_equals_exhaustive_enumeration():
    rng = G.Rng(0xB0B0_0001)
    schemes = G.SCHEMES_GPU + [(2, -4, -3, -1), (5, -3, -1, 0)]
    count = 0
    for scoring in schemes:
        for _ in range(30):
            n, m = 1 + rng.below(4), 1 + rng.below(4)
            alpha = b"ACGTN" if rng.chance(0.3) else b"AC"
            q = [b"ACGTN".index(alpha[rng.below(len(alpha))]) for _ in range(n)]
            r = [b"ACGTN".index(alpha[rng.below(len(alpha))]) for _ in range(m)]
            qa, ra = np.array(q, np.uint8), np.array(r, np.uint8)
            assert O.score(qa, ra, scoring, True) == brute_local(q, r, scoring), (q, r, scoring)
            assert O.score(qa, ra, scoring, False) == brute_semiglobal(q, r, scoring), (q, r, scoring)
            count += 1
    assert count == 30 * len(schemes)


def test_exhaustive_checker_catches_planted_mistakes():
    """The enumeration distinguishes plausible oracle mistakes: scoring a gap
    as k*open, or flooring semiglobal cells at zero, both change answers on
    tiny inputs that the enumeration covers."""
    q, r = [0, 1, 1, 0], [0, 0]
    s = (2, -1, -3, -1)
    right = brute_semiglobal(q, r, s)
    assert right == O.score(np.array(q, np.uint8), np.array(r, np.uint8), s, False)
    # a planted wrong gap model (every gap base charged `open`) differs
    wrong = _global_best(tuple(q), tuple(r), (2, -1, -3, -3))
    assert wrong != _global_best(tuple(q), tuple(r), s)
    # a semiglobal best that is negative (no floor) is reported as such
    assert O.score(np.array([0], np.uint8), np.array([1], np.uint8), s, False) == -1


def test_fits_i16_bounds():
    # lib/simd/mod.rs:1136-1140
    d = (2, -1, -10, -1)
    assert O.fits_i16(5000, 200, d)
    assert not O.fits_i16(20_000, 20_000, d)
    # closed form: k*M + |o| + |e| < 15000 and L < 15000
    assert O.fits_i16(7494, 8000, d)          # 14988 + 11 = 14999
    assert not O.fits_i16(7495, 8000, d)      # 15001
    assert not O.fits_i16(10, 15000, d)


# ---------------------------------------------------------------- SIMD restatement

def _simd_backends():
    s = O.simd_supported()
    out = []
    if s & 1:
        out.append(16)
    if s & 2:
        out.append(32)
    return out


@pytest.mark.parametrize("lanes,seed", [(16, 0x5EED_0001), (32, 0x5EED_0002)])
def test_simd_matches_scalar_random(lanes, seed):
    """simd/mod.rs:752-810: 37-ref random panel, 3 schemes x 2 modes."""
    if lanes not in _simd_backends():
        pytest.skip("CPU lacks this SIMD width")
    rng = G.Rng(seed)
    panel = G.random_panel_simd(rng, 37)
    pcodes = [G.encode(r) for r in panel]
    for scoring in G.SCHEMES_SIMD:
        for local in (True, False):
            prof = O.Profile(pcodes, scoring, lanes)
            reads = G.simd_random_reads(rng, panel, 40)
            for read in reads:
                q = G.encode(read)
                got = prof.score_all(q, local)
                want = O.score_matrix([q], pcodes, scoring, local)[0]
                assert np.array_equal(got, want), (lanes, scoring, local, read)


def test_simd_matches_scalar_on_fixture_reads():
    """simd/mod.rs:816-862 on both 47-reference panels."""
    reads = [G.encode(r) for r in G.fixture_reads()]
    assert len(reads) >= 5
    for fasta in ("trna_reference.fa", "trna_reference_ambiguous.fa"):
        pcodes = [G.encode(r) for r in G.fixture_panel(fasta)]
        assert len(pcodes) == 47
        for scoring in G.SCHEMES_SIMD:
            for local in (True, False):
                want = O.score_matrix(reads, pcodes, scoring, local)
                for lanes in _simd_backends():
                    prof = O.Profile(pcodes, scoring, lanes)
                    for k, q in enumerate(reads):
                        assert np.array_equal(prof.score_all(q, local), want[k])


def test_transposed_matches_scalar_random():
    """simd/mod.rs:940-1002 (long reads shortened to 2-4 kb to keep the
    CPU suite quick; the recurrence is the same at any length)."""
    backends = _simd_backends()
    if not backends:
        pytest.skip("no SIMD backend")
    rng = G.Rng(0x5EED_0004)
    panel = G.random_panel_simd(rng, 37)
    pcodes = [G.encode(r) for r in panel]
    combo = 0
    for scoring in G.SCHEMES_TRANSPOSED:
        for local in (True, False):
            reads = []
            for k in range(12):
                length = 1 + rng.below(320)
                reads.append(G.random_seq(rng, length, 0.02) if k % 3 == 2
                             else G.chimeric_read(rng, panel, length))
            for length in (255, 257):
                reads.append(G.chimeric_read(rng, panel, length))
            reads.append(G.chimeric_read(rng, panel, 2000 + 200 * combo))
            combo += 1
            qs = [G.encode(r) for r in reads]
            want = O.score_matrix(qs, pcodes, scoring, local)
            for lanes in backends:
                prof = O.Profile(pcodes, scoring, lanes)
                for k, q in enumerate(qs):
                    got = prof.score_all(q, local, kernel=1)
                    assert np.array_equal(got, want[k]), (lanes, scoring, local, len(q))


def test_simd_batch_driver_equals_per_read():
    backends = _simd_backends()
    if not backends:
        pytest.skip("no SIMD backend")
    rng = G.Rng(0x5EED_0005)
    panel = G.random_panel_simd(rng, 37)
    pcodes = [G.encode(r) for r in panel]
    reads = [G.encode(r) for r in G.simd_random_reads(rng, panel, 150)] + [np.zeros(0, np.uint8)]
    qc, qo, ql = O.concat(reads)
    for lanes in backends:
        prof = O.Profile(pcodes, (2, -1, -10, -1), lanes)
        batch = prof.score_batch(qc, qo, ql, True, 4)
        want = O.score_matrix(reads, pcodes, (2, -1, -10, -1), True)
        assert np.array_equal(batch, want)


# ---------------------------------------------------------------- selection

def _panel3():
    return [b"GGGGACGTACGTACGTTTTT", b"CCCCACGTACGTACGTAAAA", b"TTTTTTTTTTTTTTTTTTTT"]


def test_select_ties_in_reference_order():
    """lib/mapper.rs:571-589."""
    refs = [G.encode(r) for r in _panel3()]
    q = G.encode(b"ACGTACGTACGT")
    row = O.score_matrix([q], refs, (2, -1, -10, -1), True)[0]
    best, nt, sub, take = O.select(row, None)
    assert best == 24 and nt == 2
    assert take == [(0, False), (1, False)]
    assert sub is not None and sub < 24
    _, nt2, _, take2 = O.select(row, None, max_ties=0)
    assert nt2 == 2 and take2 == [(0, False)]


def test_select_min_score_and_reverse():
    """lib/mapper.rs:709-739."""
    refs = [G.encode(r) for r in _panel3()]
    q = G.encode(b"ACGTACGTACGT")
    row = O.score_matrix([q], refs, (2, -1, -10, -1), True)[0]
    best, _, _, take = O.select(row, None, min_score=1000)
    assert best == 24 and take == []
    # reverse complement of the unique tail of "b"
    read = G.revcomp_codes(G.encode(b"CGTACGTAAAA"))
    fwd = O.score_matrix([read], refs, (2, -1, -10, -1), True)[0]
    rev = O.score_matrix([G.revcomp_codes(read)], refs, (2, -1, -10, -1), True)[0]
    b1, _, _, _ = O.select(fwd, None)
    b2, _, _, take = O.select(fwd, rev)
    assert b2 > b1 and take[0] == (1, True)


def test_select_equals_sort_restatement():
    """Tie order, n_tied and suboptimal against an independent restatement
    by sorting the (ref, strand) candidates."""
    rng = np.random.default_rng(7)
    for trial in range(300):
        n = int(rng.integers(1, 40))
        fwd = rng.integers(-3, 4, n).astype(np.int32)
        rev = rng.integers(-3, 4, n).astype(np.int32) if trial % 2 else None
        mt = None if trial % 3 == 0 else int(rng.integers(0, 5))
        ms = int(rng.integers(-2, 3))
        cands = []
        for r in range(n):
            cands.append((r, False, int(fwd[r])))
            if rev is not None:
                cands.append((r, True, int(rev[r])))
        best = max(c[2] for c in cands)
        ties = [(r, rv) for r, rv, s in cands if s == best]
        others = [s for _, _, s in cands if s != best]
        sub = max(others) if others else None
        take = [] if best < ms else ties[: (len(ties) if mt is None else min(len(ties), mt + 1))]
        assert O.select(fwd, rev, ms, mt) == (best, len(ties), sub, take)


# ---------------------------------------------------------------- pairs

def _pair_panel():
    """lib/mapper/pair.rs:427-443: paralogs differing at base 50."""
    x = 0x2545F4914F6CDD1D
    M = (1 << 64) - 1

    def base():
        nonlocal x
        x ^= (x << 13) & M
        x ^= x >> 7
        x ^= (x << 17) & M
        return b"ACGT"[x % 4]

    a = bytes(base() for _ in range(60))
    b = bytearray(a)
    b[50] = ord("C") if a[50] == ord("A") else ord("A")
    c = bytes(base() for _ in range(60))
    return [a, bytes(b), c]


def _rows(mates, refs, scoring, local):
    rows = []
    for m in mates:
        q = G.encode(m)
        rows.append([O.score_matrix([q], refs, scoring, local)[0],
                     O.score_matrix([G.revcomp_codes(q)], refs, scoring, local)[0]])
    return rows


def test_pair_paralog_choice_score_120():
    """lib/mapper/pair.rs:450-474: the top tier is (ref a, forward), sum 120."""
    a, b, c = _pair_panel()
    refs = [G.encode(s) for s in (a, b, c)]
    r1 = a[:30]
    r2 = G.revcomp_codes(G.encode(a[30:]))
    r2_letters = bytes(b"ACGT"[v] for v in r2)
    rows = _rows([r1, r2_letters], refs, (2, -1, -10, -1), False)
    cands = O.pair_cands(rows)
    assert cands[0][:3] == (120, 0, False)
    assert cands[1][0] < 120  # the paralog b loses by the base at 50


def test_pair_both_mates_must_clear_min_score():
    """lib/mapper/pair.rs:543-556."""
    a, b, c = _pair_panel()
    refs = [G.encode(s) for s in (a, b, c)]
    r2 = bytes(b"ACGT"[v] for v in G.revcomp_codes(G.encode(a[30:])))
    rows = _rows([a[:30], r2], refs, (2, -1, -10, -1), False)
    assert all(c[3] >= 61 and c[4] >= 61 for c in O.pair_cands(rows, min_score=61))
    assert O.pair_cands(rows, min_score=61) == []


def test_pair_cands_equal_sort_restatement():
    rng = np.random.default_rng(11)
    for trial in range(200):
        n = int(rng.integers(1, 30))
        rows = [[rng.integers(-5, 6, n).astype(np.int32) for _ in range(2)] for _ in range(2)]
        both = bool(trial % 2)
        ms = int(rng.integers(-3, 3))
        want = []
        for t in ([0, 1] if both else [0]):
            ra = rows[0][t ^ 0]
            rb = rows[1][t ^ 1]
            for r in range(n):
                if ra[r] >= ms and rb[r] >= ms:
                    want.append((int(ra[r] + rb[r]), r, bool(t), int(ra[r]), int(rb[r])))
        want.sort(key=lambda c: (-c[0], c[1], c[2]))
        assert O.pair_cands(rows, ms, both) == want


# ---------------------------------------------------------------- adapter

R1 = b"AGATCGGAAGAGCACACGTCTGAACTCCAGTCA"
R2 = b"AGATCGGAAGAGCGTCGTGTAGGGAAAGAGTGT"


def test_adapter_known_answers():
    """lib/adapter.rs:116-197."""
    f = O.adapter_find
    assert f(b"ACGTACGTAC" + R1[:20], R1, 5) == 20
    assert f(b"ACGTACGTAC" + R1[:6], R1, 5) == 6
    assert f(b"ACGTACGTAC" + R1 + b"TTTT", R1, 5) == len(R1) + 4
    a = bytearray(R1[:20])
    a[5] = ord("T")
    a[12] = ord("T")
    assert f(b"G" * 12 + bytes(a), R1, 5) == 20
    a[15] = ord("C")
    assert f(b"G" * 12 + bytes(a), R1, 5) == 0
    assert f(b"G" * 12 + b"AGATCGGAT", R1, 5) == 0
    assert f(b"G" * 12 + b"AGAT", R1, 5) == 0
    assert f(b"G" * 12 + b"AGAT", R1, 4) == 4
    assert f(b"AG", R1, 5) == 0
    assert f(b"", R1, 5) == 0
    assert f(b"G" * 12 + b"AGATNGGAAGAG", R1, 5) == 12


def test_adapter_pairs_reconcile():
    insert = b"GGGGCCCCGGGGCCCCGGGG"
    mk = lambda ad, extra: insert + ad[:extra]
    shifted = lambda ad, extra: insert + b"CC" + ad[:extra]
    d = lambda x, y: O.adapter_detect_pair(x, y, R1, R2)
    a, b = mk(R1, 12), mk(R2, 12)
    assert d(a, b) == [12, 12]
    assert d(a, shifted(R2, 9)) == [12, 0]
    assert d(shifted(R1, 7), b) == [0, 12]
    assert d(mk(R1, 8), shifted(R2, 8)) == [0, 0]
    assert d(a, insert) == [12, 0]
