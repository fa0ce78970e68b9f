"""The unpaired run's output records (esch_bam_build_records, host code)
against the reference's own golden digests.

crates/eschalign-cli/tests/align_unpaired.rs runs the CLI on the 60-read
fixture BAM against the 47-tRNA panel under six flag sets and keeps, per
set, the record count and the FNV-1a 64 digest of the uncompressed BAM
record stream, header skipped (fixtures/unpaired_golden.tsv, copied to
tests/golden/). Here the same input goes through this build's pipeline:
reads from the BAM, reads over --max-read-len kept out of alignment, the
mapping (the oracle's map_reads_scored restatement on the CPU here; the
device in tests/test_gpu_bam.py), the records built and laid out by the
library, the coordinate sort when asked. Equal digests pin every byte of
every record: flags, positions, MAPQ, bins, CIGARs, sequences and
qualities, the kept input fields in their order, NM / MD / AS / XS / XA and
the secondary records."""
import os
import struct

import numpy as np
import pytest

import eschalign_amd as ea
from eschalign_amd import bam as B
from oracle import pyoracle as O
import refgen as G

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

# align_unpaired.rs:21-36
CONFIGS = [
    ("default", B.RunOptions()),
    ("semiglobal_both_secondary", B.RunOptions(mode="semiglobal", both_strands=True, secondary=True)),
    ("sort_coordinate", B.RunOptions(sort_coordinate=True)),
    ("min_score_max_ties", B.RunOptions(min_score=60, max_ties=1)),
    ("scoring_bwa", B.RunOptions(scoring=(1, -1, -2, -1), both_strands=True)),
    ("max_read_len", B.RunOptions(max_read_len=60)),
]


def golden():
    out = {}
    with open(os.path.join(GOLDEN, "unpaired_golden.tsv")) as f:
        for line in f:
            label, n, h = line.rstrip("\n").split("\t")
            out[label] = (int(n), h)
    return out


def panel():
    return ea.Panel(B.read_fasta(os.path.join(GOLDEN, "trna_reference.fa")))


def oracle_mapper(p):
    def run(codes, opts):
        return O.map_reads(codes, p.codes, opts.scoring, opts.mode == "local", min_score=opts.min_score,
                           max_ties=opts.max_ties, both_strands=opts.both_strands, raw=True)
    return run


@pytest.mark.parametrize("label,opts", CONFIGS, ids=[c[0] for c in CONFIGS])
def test_record_stream_digest_equals_the_references(label, opts):
    p = panel()
    bam = B.read_bam(os.path.join(GOLDEN, "trna_mappings_padded.bam"))
    stream, n = B.run_unpaired(bam, p, opts, oracle_mapper(p))
    assert (n, f"{B.fnv1a64(stream):016x}") == golden()[label], label


def test_fixture_inputs():
    """The inputs as the reference reads them: 47 references named by the
    header's first word, 60 primary forward-strand records whose header
    lists the same references; FNV-1a's published test vectors."""
    p = panel()
    assert len(p) == 47 and p.names[0] == "tRNA-Ala-GGC-1-1"
    assert [s for s in p.seqs] == G.fixture_panel("trna_reference.fa")
    bam = B.read_bam(os.path.join(GOLDEN, "trna_mappings_padded.bam"))
    assert bam.ref_names == p.names and len(bam.records) == 60
    assert len(B.primary_records(bam.records)) == 60
    assert B.fnv1a64(b"") == 0xCBF29CE484222325
    assert B.fnv1a64(b"a") == 0xAF63DC4C8601EC8C
    assert B.fnv1a64(b"foobar") == 0x85944171F73967E8


def test_owned_fields_move_and_records_read_back():
    """Every output record parses back; a mapped primary's owned fields come
    last (NM, MD, AS, then XS / XA) and its other fields keep their values
    (the last input field taking the first removed one's place); an
    unmapped record has no position, MAPQ 0 and bin 4680; a reverse primary
    carries the reverse-complemented sequence."""
    p = panel()
    bam = B.read_bam(os.path.join(GOLDEN, "trna_mappings_padded.bam"))
    opts = B.RunOptions(scoring=(1, -1, -2, -1), both_strands=True, max_read_len=150)
    stream, n = B.run_unpaired(bam, p, opts, oracle_mapper(p))
    recs = B.split_stream(stream)
    assert len(recs) == n == 60
    seen_rev = seen_unmapped = 0
    for k, (r_in, r_out) in enumerate(zip(bam.records, recs)):
        r = r_out[4:]
        tid, pos, l_name, mapq, bin_ = struct.unpack_from("<iiBBH", r, 0)
        flag = B.record_flag(r)
        assert B.record_name(r) == B.record_name(r_in)
        tags = _tags(r)
        assert not {"NM", "MD", "AS", "XS", "XA"} & set(tags[:len(tags) - 5]) or flag & B.FLAG_UNMAPPED
        if flag & B.FLAG_UNMAPPED:
            seen_unmapped += 1
            assert (tid, pos, mapq, bin_) == (-1, -1, 0, 4680)
            assert not {"NM", "MD", "AS", "XS", "XA"} & set(tags)
            continue
        t_in = _tags(r_in)
        assert tags[0] == t_in[-1], (k, tags[:3], t_in[-3:])  # PT moved into NM's place
        assert tags[-len([t for t in tags if t in ("NM", "MD", "AS", "XS", "XA")]):][:3] == ["NM", "MD", "AS"]
        if flag & B.FLAG_REVERSE:
            seen_rev += 1
            letters = B.record_letters(r_in)
            comp = bytes({ord("A"): ord("T"), ord("C"): ord("G"), ord("G"): ord("C"), ord("T"): ord("A")}.get(c, ord("N"))
                         for c in reversed(letters))
            assert B.record_letters(r) == comp
    assert seen_unmapped > 0  # the reads over 150 nt


def _tags(r):
    """The data fields' tags of a record, in order."""
    l_name, n_cigar, l_seq = r[8], struct.unpack_from("<H", r, 12)[0], struct.unpack_from("<I", r, 16)[0]
    p = 32 + l_name + 4 * n_cigar + (l_seq + 1) // 2 + l_seq
    tags = []
    while p < len(r):
        t, ty = r[p:p + 2].decode(), chr(r[p + 2])
        p += 3
        if ty in "AcC":
            p += 1
        elif ty in "sS":
            p += 2
        elif ty in "iIf":
            p += 4
        elif ty in "ZH":
            p = r.index(b"\0", p) + 1
        elif ty == "B":
            sub, cnt = chr(r[p]), struct.unpack_from("<I", r, p + 1)[0]
            p += 5 + cnt * {"c": 1, "C": 1, "s": 2, "S": 2, "i": 4, "I": 4, "f": 4}[sub]
        tags.append(t)
    return tags


def test_refusals():
    """A reverse-strand aligned input record is refused (align.rs:631-640)
    and secondary / supplementary ones skipped; a malformed record stream or
    a hit outside the panel is INVALID_ARG, never a crash."""
    bam = B.read_bam(os.path.join(GOLDEN, "trna_mappings_padded.bam"))
    r = bytearray(bam.records[0])
    struct.pack_into("<H", r, 14, 0x10)
    with pytest.raises(ValueError):
        B.primary_records([bytes(r)])
    struct.pack_into("<H", r, 14, 0x100)
    assert B.primary_records([bytes(r), bam.records[1]]) == [bam.records[1]]
    p = panel()
    maps = np.zeros(1, ea.READ_MAPPING_DTYPE)
    with pytest.raises(ea.GpuError):
        B.build_records([bam.records[0][:40]], maps, np.zeros(0, ea.HIT_DTYPE), np.zeros(0, np.uint32), p)
    maps["n_hits"] = 1
    hits = np.zeros(1, ea.HIT_DTYPE)
    hits["reference"] = 99
    with pytest.raises(ea.GpuError):
        B.build_records([bam.records[0]], maps, hits, np.zeros(0, np.uint32), p)
