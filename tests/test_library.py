"""The C-ABI library without a GPU: it loads, exports every symbol the
header declares, and its host-side logic (validation, refusals, the
scores_len bound) matches the reference's rules. No compute calls."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import eschalign_amd as ea
from oracle import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "eschalign_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(esch_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = ea.load_library()
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), f"{n} declared in the header but not exported"
    assert sorted(ea.EXPORTS) == names


def test_scoring_validation_matches_reference_rules():
    for s in [(2, -1, -10, -1), (1, -1, -2, -1), (3, 1, -5, -2), (2, -1, 0, 0), (1000, -1000, -1000, -1000)]:
        assert ea.load_library().esch_scoring_validate(*s) == 0
        assert O.validate(*s) == 0
    for s in [(0, -1, -10, -1), (2, 3, -10, -1), (2, -1, 10, -1), (2, -1, -10, 1), (1001, -1, -1, -1)]:
        assert ea.load_library().esch_scoring_validate(*s) == ea.SCORING_INVALID
        assert O.validate(*s) != 0
        with pytest.raises(ValueError):
            ea.Scoring(*s)
    assert str(ea.Scoring.parse("2,-1,-10,-1")) == "2,-1,-10,-1"
    for bad in ("2,-1,-10", "2,-1,-10,-x"):
        with pytest.raises(ValueError):
            ea.Scoring.parse(bad)
    assert ea.Scoring().gap(1) == -10 and ea.Scoring().gap(3) == -12


def test_fits_i16_matches_oracle():
    rng = np.random.default_rng(3)
    for _ in range(500):
        n = int(rng.integers(0, 20000))
        L = int(rng.integers(1, 20000))
        s = (int(rng.integers(1, 30)), int(rng.integers(-30, 1)), int(rng.integers(-50, 1)),
             int(rng.integers(-50, 1)))
        assert ea.fits_i16(n, L, ea.Scoring(*s)) == O.fits_i16(n, L, s)


def test_encode_matches_oracle():
    every = bytes(range(256))
    assert np.array_equal(ea.encode(every), O.encode(every))
    c = ea.encode(b"ACGTNacgtu")
    assert np.array_equal(ea.reverse_complement_codes(c), O.revcomp(c))


def test_refusals_before_any_device_work():
    """Panel::new and GpuScorer::new refusals are host-side: empty panel,
    empty reference, invalid scoring, a reference over the bound, and codes
    outside 0..=4 — the same answer with or without a device."""
    L = ea.load_library()
    out = C.c_void_p()

    def new(refs, scoring=(2, -1, -10, -1), mode=0):
        arrs = [np.ascontiguousarray(r, np.uint8) for r in refs]
        ptrs = (C.c_void_p * max(1, len(arrs)))(*[a.ctypes.data for a in arrs])
        lens = (C.c_size_t * max(1, len(arrs)))(*[a.size for a in arrs])
        return L.esch_scorer_new(0, ptrs, lens, len(arrs), *scoring, mode, C.byref(out))

    assert new([]) == ea.EMPTY_PANEL
    assert new([np.zeros(0, np.uint8)]) == ea.EMPTY_REFERENCE
    assert new([np.zeros(10, np.uint8)], (2, 3, -1, -1)) == ea.SCORING_INVALID
    assert new([np.zeros(ea.MAX_REF_LEN + 1, np.uint8)]) == ea.REFERENCE_TOO_LONG
    assert new([np.full(10, 7, np.uint8)]) == ea.INVALID_ARG
    assert new([np.zeros(10, np.uint8)], mode=9) == ea.INVALID_ARG
    assert b"reference" in L.esch_last_error() or b"mode" in L.esch_last_error()


def test_variant_shapes():
    """esch_variant_info (host-only): 0-8 are the schedules on 16 rows, 9-16
    some of them on 24 / 32 rows, 17-22 the windowed schedules 9-11 and
    23-28 the diagonal-shift schedules 12-14 on 24 and 32 rows; fp16 and
    the shift follow the schedule; out of range refused."""
    vs = ea.variants()
    assert len(vs) == 29
    assert [vs[v].schedule for v in range(9)] == list(range(9))
    assert all(vs[v].stripe_rows == 16 for v in range(9))
    assert [(vs[v].schedule, vs[v].stripe_rows) for v in range(17, 29)] == [
        (9, 24), (9, 32), (10, 24), (10, 32), (11, 24), (11, 32),
        (12, 24), (12, 32), (13, 24), (13, 32), (14, 24), (14, 32)]
    for v, x in vs.items():
        assert x.f16 == (x.schedule in (4, 5, 7, 8, 9, 10, 12, 13)), v
        assert x.shifted == (2 if x.schedule >= 12 else 1 if x.schedule >= 6 else 0), v
        assert x.diag == (x.schedule >= 12), v
    # each (schedule, stripe height) once
    assert len({(x.schedule, x.stripe_rows) for x in vs.values()}) == len(vs)
    for bad in (-1, 29):
        with pytest.raises(ea.GpuError):
            ea.variant_info(bad)


def test_no_device_is_an_error_not_a_crash():
    """lib/cuda/mod.rs:356-366: a missing device is GpuError::Unavailable."""
    if ea.device_visible():
        pytest.skip("a device is visible here")
    with pytest.raises(ea.GpuError) as e:
        ea.GpuScorer(ea.Panel([b"ACGT" * 10]), ea.Scoring(), ea.Mode.LOCAL)
    assert e.value.kind == "Unavailable"


def test_score_matrix_semantics():
    """lib/mapper.rs:94-139: rows of unscored queries are None."""
    m = ea.ScoreMatrix(3, np.arange(9, dtype=np.int16), np.array([1, 0, 1], bool))
    assert len(m) == 3 and m.n_refs() == 3 and m.n_scored() == 2
    assert list(m.row(0)) == [0, 1, 2] and m.row(1) is None
    with pytest.raises(ValueError):
        ea.ScoreMatrix(4, np.zeros(9, np.int16), np.zeros(3, bool))


R1 = ea.TRUSEQ_R1
R2 = ea.TRUSEQ_R2


def test_adapter_find_known_answers():
    """lib/adapter.rs:116-164 through the C ABI."""
    f = ea.adapter_find
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
    assert f(b"AG", R1, 5) == 0 and f(b"", R1, 5) == 0
    assert f(b"G" * 12 + b"AGATNGGAAGAG", R1, 5) == 12


def test_adapter_detect_pair_known_answers():
    """lib/adapter.rs:166-197."""
    insert = b"GGGGCCCCGGGGCCCCGGGG"
    a, b = insert + R1[:12], insert + R2[:12]
    shifted = lambda ad, extra: insert + b"CC" + ad[:extra]
    d = ea.adapter_detect_pair
    assert d(a, b) == [12, 12]
    assert d(a, shifted(R2, 9)) == [12, 0]
    assert d(shifted(R1, 7), b) == [0, 12]
    assert d(insert + R1[:8], shifted(R2, 8)) == [0, 0]
    assert d(a, insert) == [12, 0]


def test_adapter_equals_oracle_random():
    rng = np.random.default_rng(5)
    for _ in range(2000):
        n = int(rng.integers(0, 60))
        read = bytes(rng.choice(list(b"ACGTN"), n, p=[0.24, 0.24, 0.24, 0.24, 0.04]).astype(np.uint8))
        if n and rng.random() < 0.7:
            k = int(rng.integers(0, min(n, len(R1)) + 1))
            tail = bytearray(R1[:k])
            for i in range(len(tail)):
                if rng.random() < 0.08:
                    tail[i] = ord("ACGT"[int(rng.integers(0, 4))])
            read = read[: n - k] + bytes(tail)
        mo = int(rng.integers(1, 8))
        assert ea.adapter_find(read, R1, mo) == O.adapter_find(read, R1, mo)


def test_alignment_record_layout_matches_the_header(tmp_path):
    """The Python mirror reads esch_alignment as 32-byte records with the
    int64 ops offset at byte 24; the header's struct, compiled, agrees."""
    import subprocess
    src = tmp_path / "layout.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "eschalign_amd.h"\n'
                   'int main(void) { printf("%zu %zu", sizeof(esch_alignment), '
                   'offsetof(esch_alignment, ops_offset)); return 0; }\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    assert out.split() == ["32", "24"]


def test_mapping_record_layouts_match_the_header(tmp_path):
    """esch_hit / esch_read_mapping / esch_pair_mapping, compiled from the
    header, have the size and field offsets of the Python dtypes (and of the
    oracle's records, which tests compare with them)."""
    import subprocess
    from oracle import pyoracle as O
    fields = {
        "esch_hit": ["reference", "reverse", "template_reverse", "insert", "alignment"],
        "esch_read_mapping": ["best", "n_tied", "subopt", "flags", "n_hits", "hit_offset"],
        "esch_pair_mapping": ["proper", "score", "n_tied", "flags", "subopt", "n_hits", "hit_offset",
                              "separate"],
    }
    body = "".join(
        f'printf("{t} %zu", sizeof({t}));' + "".join(f'printf(" %zu", offsetof({t}, {f}));' for f in fs)
        + 'printf("\\n");' for t, fs in fields.items())
    src = tmp_path / "maplayout.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "eschalign_amd.h"\n'
                   f'int main(void) {{ {body} return 0; }}\n')
    exe = tmp_path / "maplayout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {ln.split()[0]: [int(x) for x in ln.split()[1:]]
           for ln in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines()}
    hit = ea.HIT_DTYPE
    assert got["esch_hit"] == [hit.itemsize] + [hit.fields[f][1] for f in
                                                ["reference", "reverse", "template_reverse", "insert", "score"]]
    rd = ea.READ_MAPPING_DTYPE
    assert got["esch_read_mapping"] == [rd.itemsize] + [rd.fields[f][1] for f in fields["esch_read_mapping"]]
    pd = ea.PAIR_MAPPING_DTYPE
    assert got["esch_pair_mapping"] == [pd.itemsize] + [pd.fields[f][1] for f in fields["esch_pair_mapping"]]
    assert (O.HIT_DTYPE, O.READ_DTYPE, O.PAIR_DTYPE) == (hit, rd, pd)


def test_adapter_trim_pairs_equals_detect_pair():
    """The batched pre-step on codes gives, for every pair of the paired
    workload's generator (inserts N(120, 30) reading into the adapters) and of
    the paired fixture, the oracle's detect_pair (lib/adapter.rs:95-110)."""
    from eschalign_amd import synth
    panel = synth.panel(seed=0xC0F4, n_refs=50, lo=70, hi=95)
    codes, off, lens = synth.pair_arrays(panel, seed=0xAD, n_pairs=6000)
    got = ea.adapter_trim_pairs(codes, off, lens)
    letters = np.frombuffer(b"ACGTN", np.uint8)[codes].tobytes()
    trimmed = 0
    for p in range(len(lens) // 2):
        a = letters[int(off[2 * p]):int(off[2 * p]) + int(lens[2 * p])]
        b = letters[int(off[2 * p + 1]):int(off[2 * p + 1]) + int(lens[2 * p + 1])]
        d = O.adapter_detect_pair(a, b, R1, R2, 5)
        assert [int(got[2 * p]), int(got[2 * p + 1])] == [len(a) - d[0], len(b) - d[1]], p
        trimmed += d[0] > 0
    assert trimmed > 1000


def test_mapping_and_sam_entry_points_refuse_bad_arguments():
    """The mapping calls and the SAM field writers refuse null or
    inconsistent arguments with a status, never a crash, before any device
    work."""
    L = ea.load_library()
    out = C.c_void_p()
    assert L.esch_map_reads(None, None, None, 0, 0, 0, -1, C.byref(out)) == ea.INVALID_ARG
    assert L.esch_map_reads(None, None, None, 0, 0, 0, -1, None) == ea.INVALID_ARG
    assert L.esch_map_pairs(None, None, None, None, None, 0, 0, 0, -1, 500, C.byref(out)) == ea.INVALID_ARG
    assert L.esch_map_pairs_concat(None, None, None, None, 0, 0, 0, -1, 500, C.byref(out)) == ea.INVALID_ARG
    assert L.esch_map_reads_concat(None, None, None, None, 0, 0, 0, -1, C.byref(out)) == ea.INVALID_ARG
    assert L.esch_set_map_chunk(None, 0) == ea.INVALID_ARG
    assert L.esch_mapped_len(None) == 0
    assert L.esch_mapped_pairs(None) is None and L.esch_mapped_reads(None) is None
    n = C.c_size_t(7)
    assert L.esch_mapped_hits(None, C.byref(n)) is None and n.value == 0
    L.esch_mapped_free(None)
    # ops running past the read, or spans out of order
    with pytest.raises(ValueError):
        ea.md_nm(b"ACGT", b"ACGT", 0, 0, [("M", 5)])
    with pytest.raises(ValueError):
        ea.md_nm(b"ACGT", b"ACGT", 0, 2, [("D", 3)])
    with pytest.raises(ValueError):
        ea.cigar_string([("M", 2)], 4, 3, 2)
