"""The oracle's read and pair mapping (eo_map_reads / eo_map_pairs,
lib/mapper.rs:400-550, lib/mapper/pair.rs:157-387) pinned against what fixes
it independently of its own code:

- every case of lib/mapper/pair.rs:449-573 on the same xorshift panel
  (paralog choice with score 120 and insert 60, max-insert and orientation
  limits, the reverse template strand only with both strands, min-score per
  mate, batched equal to one at a time);
- the map_reads cases of lib/mapper.rs:569-739 (ties in reference order,
  max_ties caps the hits not the count, min-score unmaps, the reverse strand
  wins when asked);
- the outcomes crates/eschalign-cli/tests/align_paired.rs states for the
  paired fixture (tests/golden/pairs): reference, position, insert, CIGAR
  core, proper or not, tie sets, the unmapped mate;
- properties the semantics fix (lib/mapper/pair.rs:5-39): a proper pair's
  hits all pass the pair constraints at the pair score, and no eligible
  candidate with a higher sum passes them (checked by tracing every one
  with eo_align); a Separate pair is exactly map_reads of each mate on both
  strands.
"""
import os
import random

import numpy as np

from oracle import pyoracle as O
import refgen as G

DEFAULT = (2, -1, -10, -1)
TRUSEQ_R1 = b"AGATCGGAAGAGCACACGTCTGAACTCCAGTCA"  # lib/adapter.rs:40
TRUSEQ_R2 = b"AGATCGGAAGAGCGTCGTGTAGGGAAAGAGTGT"  # lib/adapter.rs:42
HERE = os.path.dirname(os.path.abspath(__file__))


def xorshift_panel():
    """lib/mapper/pair.rs:426-442: paralogs a and b (one base apart at 50)
    and an unrelated c, 60 nt each, from xorshift64 (13, 7, 17)."""
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
    return a, bytes(b), c


def rc(s: bytes) -> bytes:
    return s[::-1].translate(bytes.maketrans(b"ACGTN", b"TGCAN"))


def enc(s: bytes):
    return O.encode(s)


def mp(pairs, refs, local=False, **kw):
    return O.map_pairs([(enc(p), enc(q)) for p, q in pairs], [enc(r) for r in refs], DEFAULT, local, **kw)


def test_pair_unit_cases():
    """lib/mapper/pair.rs:449-556."""
    a, b, c = xorshift_panel()
    refs = [a, b, c]
    # the paralog its mates support (:449-474)
    r1, r2 = a[:30], rc(a[30:])
    alone = O.map_reads([enc(r1)], [enc(r) for r in refs], DEFAULT, False)[0]
    assert alone[1] >= 2, "mate 1 alone is a tie"
    m = mp([(r1, r2)], refs)[0]
    assert m[0] == "proper" and m[2] == 1 and m[1] == 120
    ref, trev, insert, h1, h2 = m[4][0]
    assert ref == 0 and insert == 60 and not h1[1] and h2[1] and not trev
    # over max insert and wrong orientation (:476-513)
    r1, r2 = a[:20], rc(a[40:])
    assert mp([(r1, r2)], refs, max_insert=50, min_score=30)[0][0] == "separate"
    assert mp([(r1, r2)], refs, min_score=30)[0][0] == "proper"
    assert mp([(r1, a[40:])], refs, min_score=30)[0][0] == "separate"  # FF
    assert mp([(a[40:], rc(a[:20]))], refs, min_score=30)[0][0] == "separate"  # outie
    # the reverse template strand only with both strands (:515-540)
    r1, r2 = rc(a[40:]), a[:20]
    assert mp([(r1, r2)], refs, min_score=30)[0][0] == "separate"
    m = mp([(r1, r2)], refs, min_score=30, both_strands=True)[0]
    assert m[0] == "proper"
    _, trev, _, h1, h2 = m[4][0]
    assert trev and h1[1] and not h2[1]
    # both mates must clear min_score (:542-556): each scores 60, sum 120
    assert mp([(a[:30], rc(a[30:]))], refs, min_score=61)[0][0] == "separate"


def test_pairs_batched_equal_one_at_a_time():
    """lib/mapper/pair.rs:558-573 (an empty mate included)."""
    a, b, c = xorshift_panel()
    refs = [a, b, c]
    pairs = [(a[:30], rc(a[30:])), (b[:30], rc(b[30:])), (a[:20], a[40:]), (b"", a[:20])]
    for local in (False, True):
        for both in (False, True):
            all_ = mp(pairs, refs, local=local, both_strands=both)
            for p, m in zip(pairs, all_):
                assert mp([p], refs, local=local, both_strands=both)[0] == m
    m = mp([(b"", a[:20])], refs)[0]
    assert m[0] == "separate" and m[1][0] is None and m[1][1] == 0 and m[2][3]


def test_map_reads_unit_cases():
    """lib/mapper.rs:569-589 and 709-739."""
    refs = [enc(s) for s in (b"GGGGACGTACGTACGTTTTT", b"CCCCACGTACGTACGTAAAA", b"TTTTTTTTTTTTTTTTTTTT")]
    m = O.map_reads([enc(b"ACGTACGTACGT")], refs, DEFAULT, True)[0]
    assert m[0] == 24 and m[1] == 2 and len(m[3]) == 2
    assert m[3][0][0] == 0 and m[3][1][0] == 1 and m[2] < 24
    capped = O.map_reads([enc(b"ACGTACGTACGT")], refs, DEFAULT, True, max_ties=0)[0]
    assert capped[1] == 2 and len(capped[3]) == 1
    m = O.map_reads([enc(b"ACGTACGTACGT")], refs, DEFAULT, True, min_score=1000)[0]
    assert m[3] == () and m[0] == 24
    read = enc(rc(b"CGTACGTAAAA"))
    fwd = O.map_reads([read], refs, DEFAULT, True)[0]
    both = O.map_reads([read], refs, DEFAULT, True, both_strands=True)[0]
    assert both[0] > fwd[0]
    assert both[3][0][0] == 1 and both[3][0][1]
    # md "11": all 11 bases match, one run of 11 aligned bases
    assert both[3][0][7] == (("M", 11),)
    # batched = one at a time (lib/mapper.rs:592-614), an empty read and
    # an all-wildcard read included
    reads = [b"ACGTACGTACGT", b"CCCCACGTACGTACGTAAAA", b"TTTTTTTTTTTT", b"", b"GGGGACGTACGTACGTTTTT", b"NNNN"]
    for local in (True, False):
        batched = O.map_reads([enc(r) for r in reads], refs, DEFAULT, local, both_strands=True)
        for r, m in zip(reads, batched):
            assert O.map_reads([enc(r)], refs, DEFAULT, local, both_strands=True)[0] == m
        assert batched[3] == (None, 0, None, ())


def fixture():
    panel = G.read_fasta(os.path.join(HERE, "golden", "pairs", "pairs_panel.fa"))
    r1 = G.read_fastq(os.path.join(HERE, "golden", "pairs", "pairs_R1.fq"))
    r2 = G.read_fastq(os.path.join(HERE, "golden", "pairs", "pairs_R2.fq"))
    names = [n for n, _ in panel]
    refs = [s for _, s in panel]
    pairs, ads = {}, {}
    for (n1, s1), (n2, s2) in zip(r1, r2):
        assert n1 == n2
        ad = O.adapter_detect_pair(s1, s2, TRUSEQ_R1, TRUSEQ_R2, 5)
        ads[n1] = ad
        pairs[n1] = (s1[:len(s1) - ad[0]], s2[:len(s2) - ad[1]])
    return names, refs, pairs, ads


def test_paired_fixture_outcomes():
    """crates/eschalign-cli/tests/align_paired.rs:193-403 on the fixture:
    semi-global 2,-1,-10,-1, --min-score 40, mates less their adapters."""
    names, refs, pairs, ads = fixture()
    order = list(pairs)
    got = dict(zip(order, mp([pairs[k] for k in order], refs, min_score=40)))
    ref = {n: i for i, n in enumerate(names)}
    # paralog: proper on P1, unique; POS 21 / 91, TLEN 130, mate 2 60M;
    # mate 1 reports its would-be tie as XS = AS
    m = got["paralog"]
    assert m[0] == "proper" and m[2] == 1
    r, trev, ins, h1, h2 = m[4][0]
    assert r == ref["P1"] and not trev and ins == 130
    assert (h1[5], h2[5]) == (20, 90) and h2[7] == (("M", 60),)
    assert not h1[1] and h2[1]
    assert m[3][0] == h1[2]
    # pair_tie: a tie between P1 and P2, P1 first
    m = got["pair_tie"]
    assert m[0] == "proper" and m[2] == 2
    assert [ph[0] for ph in m[4]] == [ref["P1"], ref["P2"]]
    # short_insert: 40-nt insert, each mate's 20 adapter bases trimmed
    assert ads["short_insert"] == [20, 20]
    m = got["short_insert"]
    r, _, ins, h1, h2 = m[4][0]
    assert m[0] == "proper" and r == ref["U1"] and ins == 40
    assert h1[7] == (("M", 40),) and h2[7] == (("M", 40),)
    assert (h1[5], h2[5]) == (30, 30)
    # one_mate_reads_through: 40M (+15 clipped) and 35M, insert 40
    assert ads["one_mate_reads_through"] == [15, 0]
    m = got["one_mate_reads_through"]
    _, _, ins, h1, h2 = m[4][0]
    assert h1[7] == (("M", 40),) and h2[7] == (("M", 35),) and ins == 40
    # terminal_mismatch: semi-global keeps both end mismatches (60M)
    m = got["terminal_mismatch"]
    _, _, _, h1, h2 = m[4][0]
    assert h1[7] == (("M", 60),) and h2[7] == (("M", 60),)
    loc = mp([pairs["terminal_mismatch"]], refs, local=True, min_score=40)[0]
    assert loc[4][0][3][7] == (("M", 59),) and loc[4][0][3][4] == 59  # 59M1S
    # over_max_insert: separate under --max-insert 120, both at U1 (mate 2
    # reverse), proper under the default 500
    sep = mp([pairs["over_max_insert"]], refs, min_score=40, max_insert=120)[0]
    assert sep[0] == "separate"
    assert sep[1][3][0][0] == ref["U1"] and sep[2][3][0][0] == ref["U1"]
    assert not sep[1][3][0][1] and sep[2][3][0][1]
    assert got["over_max_insert"][0] == "proper"
    # wrong_orientation: separate, both forward on U2
    m = got["wrong_orientation"]
    assert m[0] == "separate"
    assert [(x[3][0][0], x[3][0][1]) for x in m[1:]] == [(ref["U2"], False)] * 2
    # one_unmapped: mate 1 mapped, mate 2 (junk) below --min-score
    m = got["one_unmapped"]
    assert m[0] == "separate" and m[1][3] and m[2][3] == ()
    assert m[2][0] < 40


def traced(q, r, local):
    a = O.align(q, r, DEFAULT, local)
    return a


def pair_ok(al, trev, max_insert):
    """The constraints of lib/mapper/pair.rs:20-30 on two traced mates."""
    if not any(op == "M" for op, _ in al[0][5]) or not any(op == "M" for op, _ in al[1][5]):
        return False
    fw, rv = (al[1], al[0]) if trev else (al[0], al[1])
    if fw[3] >= rv[4]:
        return False
    return max(al[0][4], al[1][4]) - min(al[0][3], al[1][3]) <= max_insert


def test_pair_properties_random():
    """Random panels with paralog copies and random pairs (FR, RF, FF,
    outies, junk): the semantics of lib/mapper/pair.rs:5-39 checked by
    tracing every candidate above the chosen score."""
    rng = random.Random(0x9A1E)
    seen = {"fell_through": 0, "tie": 0, "separate": 0, "reverse_template": 0}
    for trial in range(6):
        base = [bytes(rng.choice(b"ACGT") for _ in range(rng.randint(50, 90))) for _ in range(4)]
        refs = []
        for s in base:
            refs.append(s)
            t = bytearray(s)
            t[rng.randrange(len(t))] = rng.choice(b"ACGT")
            refs.append(bytes(t))
        pairs = []
        for _ in range(40):
            s = rng.choice(refs)
            i = rng.randrange(0, len(s) - 20)
            j = rng.randrange(i + 10, len(s) + 1)
            x, y = s[i:i + rng.randint(12, 30)], s[max(0, j - rng.randint(12, 30)):j]
            kind = rng.randrange(5)
            if kind == 0:
                p = (x, rc(y))
            elif kind == 1:
                p = (rc(y), x)
            elif kind == 2:
                p = (x, y)
            elif kind == 3:
                p = (rc(x), y)
            else:
                p = (x, bytes(rng.choice(b"ACGT") for _ in range(len(y))))
            pairs.append(p)
        local = trial % 2 == 1
        both = trial % 3 != 0
        min_score = rng.choice([0, 10, 20])
        max_insert = rng.choice([500, 40])
        max_ties = rng.choice([None, 0, 1])
        got = mp(pairs, refs, local=local, both_strands=both, min_score=min_score,
                 max_insert=max_insert, max_ties=max_ties)
        R = [enc(r) for r in refs]
        for (p1, p2), m in zip(pairs, got):
            q = [enc(p1), enc(p2)]
            qr = [O.revcomp(q[0]), O.revcomp(q[1])]
            sc = {(m_, rv): [O.score((qr if rv else q)[m_], r, DEFAULT, local) for r in R]
                  for m_ in range(2) for rv in range(2)}
            cands = []
            if len(p1) and len(p2):
                for t in ((0, 1) if both else (0,)):
                    for ri in range(len(R)):
                        a_, b_ = sc[(0, t)][ri], sc[(1, 1 - t)][ri]
                        if a_ >= min_score and b_ >= min_score:
                            cands.append((a_ + b_, ri, t))

            def trace_cand(ri, t):
                return [O.align((qr if (t ^ (mm == 1)) else q)[mm], R[ri], DEFAULT, local) for mm in range(2)]

            if m[0] == "proper":
                score, n_tied, sub, hits = m[1], m[2], m[3], m[4]
                seen["fell_through"] += score < max(c_[0] for c_ in cands)
                seen["tie"] += n_tied > 1
                seen["reverse_template"] += hits[0][1]
                # no higher-sum eligible candidate passes the constraints
                for s_, ri, t in cands:
                    if s_ > score:
                        assert not pair_ok(trace_cand(ri, t), t, max_insert)
                # the tie set: every eligible candidate at the score that passes
                ties = [(ri, t) for s_, ri, t in sorted(cands, key=lambda c: (-c[0], c[1], c[2]))
                        if s_ == score and pair_ok(trace_cand(ri, t), t, max_insert)]
                assert n_tied == len(ties)
                keep = len(ties) if max_ties is None else min(len(ties), max_ties + 1)
                assert [(h[0], int(h[1])) for h in hits] == ties[:keep]
                for ri, trev, ins, h1, h2 in hits:
                    al = trace_cand(ri, int(trev))
                    assert (h1[2], h1[3], h1[4], h1[5], h1[6]) == al[0][:5]
                    assert (h2[2], h2[3], h2[4], h2[5], h2[6]) == al[1][:5]
                    assert h1[2] + h2[2] == score
                    assert ins == max(h1[6], h2[6]) - min(h1[5], h2[5])
                # XS per mate: best outside the tie set over the strands taken
                for mm in range(2):
                    outside = [sc[(mm, t ^ (mm == 1))][ri] for t in ((0, 1) if both else (0,))
                               for ri in range(len(R)) if (ri, t) not in ties]
                    assert sub[mm] == (max(outside) if outside else None)
            else:
                seen["separate"] += 1
                for s_, ri, t in cands:
                    assert not pair_ok(trace_cand(ri, t), t, max_insert)
                want = O.map_reads([q[0], q[1]], R, DEFAULT, local, min_score=min_score, max_ties=max_ties,
                                   both_strands=True)
                assert (m[1], m[2]) == (want[0], want[1])
    # the cases the semantics distinguish all occur
    assert all(v > 0 for v in seen.values()), seen
