"""SAM fields: the oracle's md_nm / cigar_string (lib/sam.rs:32-124) and the
library's esch_md_nm / esch_cigar_string, both pinned to the 21 samtools
1.23.1 calmd results of tests/golden/calmd_golden.json (the data of
crates/eschalign/tests/sam_fields.rs, reference N / IUPAC / lower case / U
positions included) and to lib/sam.rs:143-156; then equal to each other on
random alignments from the oracle's traceback over IUPAC-laden letters."""
import json
import os
import random

import eschalign_amd as ea
from oracle import pyoracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def parse_cigar(s):
    """crates/eschalign/tests/sam_fields.rs:11-32: (lead clip, ops, tail clip)."""
    lead = tail = 0
    ops, num = [], 0
    for ch in s:
        if ch.isdigit():
            num = num * 10 + int(ch)
            continue
        if ch in "MID":
            ops.append((ch, num))
        elif ch == "S" and not ops:
            lead = num
        elif ch == "S":
            tail = num
        else:
            raise ValueError(ch)
        num = 0
    return lead, ops, tail


def golden():
    with open(os.path.join(HERE, "golden", "calmd_golden.json")) as f:
        return json.load(f)


def test_md_nm_match_samtools_calmd():
    g = golden()
    ref = g["reference"]["seq"].encode()
    names = {r["name"] for r in g["records"]}
    assert {"n_under_read_base", "n_under_read_n", "n_inside_deletion"} <= names
    for r in g["records"]:
        seq = r["seq"].encode()
        lead, ops, tail = parse_cigar(r["cigar"])
        for f in (O.md_nm, ea.md_nm):
            assert f(seq, ref, lead, r["pos"] - 1, ops) == (r["md"], r["nm"]), (f, r["name"])
        for f in (O.cigar_string, ea.cigar_string):
            assert f(ops, len(seq), lead, len(seq) - tail) == r["cigar"], r["name"]
    assert len(g["records"]) == 21


def test_sam_known_answers():
    """lib/sam.rs:143-156."""
    ops = [("M", 5), ("I", 1), ("M", 3)]
    for f in (O.cigar_string, ea.cigar_string):
        assert f(ops, 12, 2, 11) == "2S5M1I3M1S"
        assert f(ops, 9, 0, 9) == "5M1I3M"
    for f in (O.md_nm, ea.md_nm):
        assert f(b"ACGTAACGTA", b"ACGTANCGTA", 0, 0, [("M", 10)]) == ("5N4", 1)


def test_md_nm_library_equals_oracle_random():
    rng = random.Random(3)
    alpha = b"ACGTACGTACGTNnacgtURYU"
    for _ in range(3000):
        r = bytes(rng.choice(alpha) for _ in range(rng.randint(1, 40)))
        q = bytes(rng.choice(alpha) for _ in range(rng.randint(1, 40)))
        a = O.align(O.encode(q), O.encode(r), (2, -1, -3, -1), rng.random() < 0.5)
        if not a[5]:
            continue
        assert ea.md_nm(q, r, a[1], a[3], a[5]) == O.md_nm(q, r, a[1], a[3], a[5])
        assert ea.cigar_string(a[5], len(q), a[1], a[2]) == O.cigar_string(a[5], len(q), a[1], a[2])


def test_hit_md_nm_on_reverse_strand():
    """lib/mapper.rs:709-738: the reverse-strand hit of b's tail reads MD
    "11" (the read is reverse-complemented before MD is taken)."""
    ref = b"CCCCACGTACGTACGTAAAA"
    read = ea.reverse_complement(b"CGTACGTAAAA")
    # lib/alphabet.rs:69-80: case folded, U as T, anything else N
    assert ea.reverse_complement(b"ACGTUNacgtx") == b"NACGTNAACGT"
    a = O.align(O.encode(ea.reverse_complement(read)), O.encode(ref), (2, -1, -10, -1), True)
    hit = ea.Hit(1, True, ea.Alignment(a[0], a[1], a[2], a[3], a[4], tuple(a[5])))
    assert hit.md_nm(read, ref) == ("11", 0)
