"""eschalign_amd: MI355X-native panel scoring for eschalign.

The host-side mirror of eschalign's score-backend interface, over the C ABI
of libeschalign_amd.so (include/eschalign_amd.h). Names and semantics follow
the reference so that code written against it reads the same:

    Scoring, Mode          lib/scoring.rs:27-165
    encode, reverse_complement_codes   lib/alphabet.rs:29-64
    Panel                  lib/panel.rs:10-82
    ScoreMatrix            lib/mapper.rs:94-139
    GpuScorer, GpuError, device_visible, MAX_READ_LEN, MAX_REF_LEN
                           lib/cuda/mod.rs:86-749

There is no CPU fallback in this package: every score comes from the gfx950
kernels, and a missing library raises at import of the scorer.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

__all__ = [
    "Scoring", "Mode", "Panel", "ScoreMatrix", "GpuScorer", "GpuError", "DeviceBatch",
    "ReadSelect", "PairSelect", "encode", "reverse_complement_codes", "device_visible",
    "fits_i16", "MAX_READ_LEN", "MAX_REF_LEN", "lib_path", "load_library",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
lib_path = os.path.join(_HERE, "libeschalign_amd.so")

MAX_READ_LEN = 4096
NOALLOC_MAX_ROWS = 1 << 20
# length classes of the per-bucket routing plan: [lo, hi) nt
LEN_CLASSES = 8
LEN_CLASS_BOUNDS = [(0, 128), (128, 256), (256, 512), (512, 1024), (1024, 2048), (2048, 4096), (4096, 8192),
                    (8192, None)]
NOALLOC_SCRATCH_BYTES = 256 << 20
MAX_REF_LEN = 16384

# esch_status
OK, UNAVAILABLE, DRIVER, CODE_OBJECT, REFERENCE_TOO_LONG, INVALID_ARG = 0, 1, 2, 3, 4, 5
SCORING_INVALID, EMPTY_PANEL, EMPTY_REFERENCE, CAPACITY, TIMEOUT = 6, 7, 8, 9, 10
_STATUS = {
    UNAVAILABLE: "Unavailable", DRIVER: "Driver", CODE_OBJECT: "Compile",
    REFERENCE_TOO_LONG: "ReferenceTooLong", INVALID_ARG: "InvalidArg",
    SCORING_INVALID: "ScoringInvalid", EMPTY_PANEL: "EmptyPanel",
    EMPTY_REFERENCE: "EmptyReference", CAPACITY: "Capacity", TIMEOUT: "Timeout",
}

RUN_SCORES16, RUN_SCORES32, RUN_SELECT, RUN_PAIRS, RUN_PLAIN = 1, 2, 4, 8, 16
RUN_NO_GRAPH, RUN_PAIR_BOTH = 32, 64

# Every symbol the header declares; tests check the library exports them.
EXPORTS = [
    "esch_last_error", "esch_device_visible", "esch_scoring_validate", "esch_fits_i16",
    "esch_scorer_new", "esch_scorer_free", "esch_scores_len", "esch_n_refs", "esch_prepare",
    "esch_prepared_free", "esch_prepared_len", "esch_prepared_n_scored", "esch_submit",
    "esch_finish", "esch_score_batch", "esch_device_name", "esch_geometry", "esch_registers",
    "esch_local_bytes", "esch_score_batch_i32", "esch_select_batch", "esch_select_pairs",
    "esch_batch_create", "esch_batch_free", "esch_batch_footprint", "esch_batch_load",
    "esch_batch_load_concat", "esch_batch_run", "esch_batch_wait", "esch_batch_timing",
    "esch_batch_stats", "esch_batch_read_scores32", "esch_batch_read_scores16",
    "esch_batch_read_select", "esch_batch_read_pairs", "esch_n_variants", "esch_variant_info", "esch_variant_usable", "esch_column_block", "esch_get_variant",
    "esch_set_variant", "esch_autotune", "esch_adapter_find", "esch_adapter_detect_pair",
    "esch_batch_ready", "esch_server_new", "esch_server_free", "esch_server_add_panel",
    "esch_server_scorer", "esch_server_submit", "esch_server_wait", "esch_server_result",
    "esch_server_stats", "esch_set_diag_min_len", "esch_get_diag_min_len",
    "esch_set_swap_rows", "esch_get_swap_rows", "esch_batch_swapped", "esch_batch_variant",
    "esch_batch_stats_diag", "esch_set_diag_ratio", "esch_get_diag_ratio", "esch_valu_peak16",
    "esch_align_pairs", "esch_align_last_ms", "esch_align_pairs_concat", "esch_map_reads",
    "esch_map_pairs", "esch_map_pairs_concat", "esch_mapped_len", "esch_mapped_reads",
    "esch_mapped_pairs", "esch_mapped_hits", "esch_mapped_ops", "esch_mapped_stats", "esch_mapped_free",
    "esch_set_map_chunk", "esch_adapter_trim_pairs", "esch_md_nm", "esch_cigar_string",
    "esch_map_reads_concat", "esch_align_last_wavefront", "esch_score_device_noalloc",
    "esch_set_diag_plan", "esch_get_diag_plan", "esch_bam_build_records",
]


class ServerConfig(C.Structure):
    _fields_ = [("ordinal", C.c_int), ("hbm_budget", C.c_size_t),
                ("max_batch_reads", C.c_size_t), ("max_batch_bases", C.c_size_t),
                ("max_slots", C.c_int), ("max_batch_cells", C.c_size_t)]


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_uint64), ("batches", C.c_uint64),
                ("queued_units", C.c_size_t), ("hbm_in_use", C.c_size_t),
                ("hbm_peak", C.c_size_t), ("slots", C.c_size_t), ("busy_slots", C.c_size_t),
                ("swapped_batches", C.c_uint64)]


class MapStats(C.Structure):
    """esch_map_stats."""
    _fields_ = [("score_ms", C.c_double), ("select_ms", C.c_double), ("trace_ms", C.c_double),
                ("trace_wall_ms", C.c_double), ("load_ms", C.c_double), ("assemble_ms", C.c_double),
                ("total_ms", C.c_double), ("rounds", C.c_int32), ("reserved", C.c_int32),
                ("traced", C.c_int64), ("cells", C.c_int64)]


# Records of esch_map_reads / esch_map_pairs (esch_hit, esch_read_mapping,
# esch_pair_mapping in include/eschalign_amd.h).
HIT_DTYPE = np.dtype([("reference", "<i4"), ("reverse", "<i4"), ("template_reverse", "<i4"),
                      ("insert", "<i4"), ("score", "<i4"), ("query_start", "<i4"),
                      ("query_end", "<i4"), ("ref_start", "<i4"), ("ref_end", "<i4"),
                      ("n_ops", "<i4"), ("ops_offset", "<i8")])
READ_MAPPING_DTYPE = np.dtype([("best", "<i4"), ("n_tied", "<i4"), ("subopt", "<i4"), ("flags", "<i4"),
                               ("n_hits", "<i4"), ("reserved", "<i4"), ("hit_offset", "<i8")])
PAIR_MAPPING_DTYPE = np.dtype([("proper", "<i4"), ("score", "<i4"), ("n_tied", "<i4"), ("flags", "<i4"),
                               ("subopt", "<i4", (2,)), ("n_hits", "<i4"), ("reserved", "<i4"),
                               ("hit_offset", "<i8"), ("separate", READ_MAPPING_DTYPE, (2,))])


class GpuError(Exception):
    """lib/cuda/mod.rs:144-182: why the scorer could not be built or run."""

    def __init__(self, status: int, message: str):
        self.status = status
        self.kind = _STATUS.get(status, f"status {status}")
        super().__init__(f"{self.kind}: {message}")


_lib = None


def load_library():
    """Load libeschalign_amd.so; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(lib_path):
        raise ImportError(
            f"{lib_path} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(lib_path)
    vp, sz, i32, i64 = C.c_void_p, C.c_size_t, C.c_int32, C.c_int64
    L.esch_last_error.restype = C.c_char_p
    L.esch_device_visible.restype = C.c_int
    L.esch_scoring_validate.argtypes = [i32] * 4
    L.esch_fits_i16.argtypes = [sz, sz, i32, i32, i32, i32]
    L.esch_scorer_new.argtypes = [C.c_int, vp, vp, sz, i32, i32, i32, i32, C.c_int, C.POINTER(vp)]
    L.esch_scorer_free.argtypes = [vp]
    L.esch_scorer_free.restype = None
    L.esch_scores_len.argtypes = [vp, sz]
    L.esch_n_refs.argtypes = [vp]
    L.esch_n_refs.restype = sz
    L.esch_prepare.argtypes = [vp, vp, vp, sz]
    L.esch_prepare.restype = vp
    L.esch_prepared_free.argtypes = [vp]
    L.esch_prepared_free.restype = None
    L.esch_prepared_len.argtypes = [vp]
    L.esch_prepared_len.restype = sz
    L.esch_prepared_n_scored.argtypes = [vp]
    L.esch_prepared_n_scored.restype = sz
    L.esch_submit.argtypes = [vp, vp, C.POINTER(vp)]
    L.esch_finish.argtypes = [vp, vp, vp, vp]
    L.esch_score_batch.argtypes = [vp, vp, vp, sz, vp, vp]
    L.esch_device_name.argtypes = [vp, C.c_char_p, sz]
    L.esch_geometry.argtypes = [vp, C.POINTER(sz), C.POINTER(sz), C.POINTER(sz)]
    L.esch_geometry.restype = None
    L.esch_registers.argtypes = [vp]
    L.esch_local_bytes.argtypes = [vp]
    L.esch_score_batch_i32.argtypes = [vp, vp, vp, sz, vp, vp]
    L.esch_score_device_noalloc.argtypes = [vp, vp, vp, vp, sz, vp, vp, vp]
    L.esch_select_batch.argtypes = [vp, vp, vp, sz, C.c_int, i32, i64, sz, vp, vp]
    L.esch_select_pairs.argtypes = [vp, vp, vp, vp, vp, sz, C.c_int, i32, sz, vp, vp]
    L.esch_batch_create.argtypes = [vp, sz, sz, C.POINTER(vp)]
    L.esch_batch_free.argtypes = [vp]
    L.esch_batch_free.restype = None
    L.esch_batch_footprint.argtypes = [vp, sz, sz]
    L.esch_batch_footprint.restype = sz
    L.esch_batch_load.argtypes = [vp, vp, vp, sz, C.c_int, C.c_int]
    L.esch_batch_load_concat.argtypes = [vp, vp, vp, vp, sz, C.c_int, C.c_int]
    L.esch_batch_run.argtypes = [vp, C.c_int, i32, i64, sz]
    L.esch_batch_wait.argtypes = [vp]
    L.esch_batch_timing.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.esch_batch_stats.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_uint64)]
    L.esch_batch_read_scores32.argtypes = [vp, vp, vp]
    L.esch_batch_read_scores16.argtypes = [vp, vp, vp]
    L.esch_batch_read_select.argtypes = [vp, vp, vp]
    L.esch_batch_read_pairs.argtypes = [vp, vp, vp]
    L.esch_n_variants.restype = C.c_int
    L.esch_variant_info.argtypes = [C.c_int] + [C.POINTER(C.c_int)] * 4
    L.esch_variant_usable.argtypes = [vp, C.c_int]
    L.esch_column_block.argtypes = [vp, C.c_int]
    L.esch_get_variant.argtypes = [vp]
    L.esch_set_variant.argtypes = [vp, C.c_int]
    L.esch_autotune.argtypes = [vp, vp, vp, sz, C.c_int, vp, C.POINTER(C.c_int)]
    L.esch_adapter_find.argtypes = [C.c_char_p, sz, C.c_char_p, sz, sz]
    L.esch_adapter_find.restype = sz
    L.esch_adapter_detect_pair.argtypes = [C.c_char_p, sz, C.c_char_p, sz, C.c_char_p, sz,
                                           C.c_char_p, sz, sz, C.POINTER(sz * 2)]
    L.esch_adapter_detect_pair.restype = None
    L.esch_batch_ready.argtypes = [vp]
    L.esch_server_new.argtypes = [C.POINTER(ServerConfig), C.POINTER(vp)]
    L.esch_server_free.argtypes = [vp]
    L.esch_server_free.restype = None
    L.esch_server_add_panel.argtypes = [vp, vp, vp, sz, i32, i32, i32, i32, C.c_int,
                                        C.POINTER(C.c_int)]
    L.esch_server_scorer.argtypes = [vp, C.c_int]
    L.esch_server_scorer.restype = vp
    L.esch_server_submit.argtypes = [vp, C.c_int, vp, vp, sz, C.c_int, i32, i64, sz,
                                     C.POINTER(C.c_uint64)]
    L.esch_server_wait.argtypes = [vp, C.c_uint64, i64]
    L.esch_server_result.argtypes = [vp, C.c_uint64, vp, vp, C.POINTER(C.c_double)]
    L.esch_server_stats.argtypes = [vp, C.POINTER(ServerStats)]
    L.esch_set_diag_min_len.argtypes = [vp, i32]
    L.esch_get_diag_min_len.argtypes = [vp]
    L.esch_get_diag_min_len.restype = i32
    L.esch_set_swap_rows.argtypes = [vp, i32]
    L.esch_get_swap_rows.argtypes = [vp]
    L.esch_get_swap_rows.restype = i32
    L.esch_batch_swapped.argtypes = [vp]
    L.esch_batch_swapped.restype = C.c_int
    L.esch_batch_variant.argtypes = [vp]
    L.esch_batch_variant.restype = C.c_int
    L.esch_batch_stats_diag.argtypes = [vp, C.POINTER(C.c_uint64)]
    L.esch_set_diag_ratio.argtypes = [vp, C.c_int32]
    L.esch_set_diag_plan.argtypes = [vp, vp, sz]
    L.esch_get_diag_plan.argtypes = [vp, vp, sz]
    L.esch_align_pairs.argtypes = [vp, vp, vp, vp, sz, vp, vp, sz, C.POINTER(sz)]
    L.esch_align_last_ms.argtypes = [vp]
    L.esch_align_pairs_concat.argtypes = [vp, vp, vp, vp, vp, sz, vp, vp, sz, C.POINTER(sz)]
    L.esch_align_last_ms.restype = C.c_float
    L.esch_align_last_wavefront.argtypes = [vp]
    L.esch_align_last_wavefront.restype = sz
    L.esch_valu_peak16.argtypes = [C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.esch_get_diag_ratio.argtypes = [vp]
    L.esch_map_reads.argtypes = [vp, vp, vp, sz, C.c_int, i32, i64, C.POINTER(vp)]
    L.esch_map_pairs.argtypes = [vp, vp, vp, vp, vp, sz, C.c_int, i32, i64, i64, C.POINTER(vp)]
    L.esch_map_reads_concat.argtypes = [vp, vp, vp, vp, sz, C.c_int, i32, i64, C.POINTER(vp)]
    L.esch_map_pairs_concat.argtypes = [vp, vp, vp, vp, sz, C.c_int, i32, i64, i64, C.POINTER(vp)]
    L.esch_mapped_len.argtypes = [vp]
    L.esch_mapped_len.restype = sz
    for f in ("esch_mapped_reads", "esch_mapped_pairs"):
        getattr(L, f).argtypes = [vp]
        getattr(L, f).restype = vp
    for f in ("esch_mapped_hits", "esch_mapped_ops"):
        getattr(L, f).argtypes = [vp, C.POINTER(sz)]
        getattr(L, f).restype = vp
    L.esch_mapped_stats.argtypes = [vp, C.POINTER(MapStats)]
    L.esch_mapped_stats.restype = None
    L.esch_mapped_free.argtypes = [vp]
    L.esch_mapped_free.restype = None
    L.esch_set_map_chunk.argtypes = [vp, sz]
    L.esch_adapter_trim_pairs.argtypes = [vp, vp, vp, sz, C.c_char_p, sz, C.c_char_p, sz, sz, vp]
    L.esch_adapter_trim_pairs.restype = None
    L.esch_md_nm.argtypes = [C.c_char_p, sz, C.c_char_p, sz, sz, sz, vp, sz, C.c_char_p, sz, C.POINTER(C.c_uint32)]
    L.esch_md_nm.restype = sz
    L.esch_cigar_string.argtypes = [vp, sz, sz, sz, sz, C.c_char_p, sz]
    L.esch_bam_build_records.argtypes = [C.c_char_p, sz, sz, vp, vp, sz, vp, sz, vp, vp, vp, sz, C.c_int, vp, sz,
                                         C.POINTER(sz), C.POINTER(sz)]
    L.esch_cigar_string.restype = sz
    _lib = L
    return L


def _check(st: int):
    if st != OK:
        raise GpuError(st, load_library().esch_last_error().decode(errors="replace"))


@dataclass(frozen=True)
class Alignment:
    """lib/scalar.rs:63-79: a scored alignment with its traceback; spans are
    0-based half-open, ops run-length (op, length) with op in "MID"."""
    score: int
    query_start: int
    query_end: int
    ref_start: int
    ref_end: int
    ops: tuple

    def is_empty(self) -> bool:
        """lib/scalar.rs:84-88: no read base aligned to a reference base."""
        return not any(op == "M" for op, _ in self.ops)

    def cigar(self) -> str:
        return "".join(f"{n}{op}" for op, n in self.ops)


def reverse_complement(letters: bytes) -> bytes:
    """alphabet::reverse_complement on letters (lib/alphabet.rs:69-80):
    reversed, A/C/G/T(U) complemented upper-case, anything else N."""
    return bytes(reversed(letters.upper())).translate(_RC_LETTERS)


@dataclass(frozen=True)
class Hit:
    """lib/mapper.rs:43-56: one placement of a read. MD / NM come from the
    letters (md_nm below), as the reference computes them when it builds
    the record (lib/mapper.rs:520-532)."""
    reference: int
    reverse: bool
    alignment: Alignment

    def md_nm(self, read: bytes, reference: bytes):
        """(MD, NM) of this hit for the read's letters as sequenced and the
        reference's letters as in the FASTA."""
        letters = reverse_complement(read) if self.reverse else read
        a = self.alignment
        return md_nm(letters, reference, a.query_start, a.ref_start, a.ops)


@dataclass(frozen=True)
class ReadMapping:
    """lib/mapper.rs:58-72."""
    best_score: Optional[int]
    n_tied: int
    suboptimal: Optional[int]
    hits: tuple

    def is_mapped(self) -> bool:
        return bool(self.hits)


@dataclass(frozen=True)
class PairHit:
    """lib/mapper/pair.rs:69-88: a placement of both mates on one reference."""
    reference: int
    reverse: bool
    mates: tuple
    insert: int

    @property
    def score(self) -> int:
        return self.mates[0].alignment.score + self.mates[1].alignment.score


@dataclass(frozen=True)
class ProperPair:
    """PairMapping::Proper (lib/mapper/pair.rs:94-103)."""
    score: int
    n_tied: int
    suboptimal: tuple
    hits: tuple


@dataclass(frozen=True)
class SeparatePair:
    """PairMapping::Separate (lib/mapper/pair.rs:104-106): each mate on its
    own, as an unpaired read on both strands."""
    mates: tuple


def _hit(h, ops) -> Hit:
    o0, k = int(h["ops_offset"]), int(h["n_ops"])
    runs = tuple(("MID"[int(x) >> 30], int(x) & 0x3FFFFFFF) for x in ops[o0:o0 + k])
    return Hit(int(h["reference"]), bool(h["reverse"]),
               Alignment(int(h["score"]), int(h["query_start"]), int(h["query_end"]),
                         int(h["ref_start"]), int(h["ref_end"]), runs))


def _read_mapping(r, hits, ops) -> ReadMapping:
    flags, h0 = int(r["flags"]), int(r["hit_offset"])
    return ReadMapping(int(r["best"]) if flags & 1 else None, int(r["n_tied"]),
                       int(r["subopt"]) if flags & 2 else None,
                       tuple(_hit(hits[h0 + i], ops) for i in range(int(r["n_hits"]))))


def _pair_mapping(p, hits, ops):
    if not int(p["proper"]):
        return SeparatePair((_read_mapping(p["separate"][0], hits, ops),
                             _read_mapping(p["separate"][1], hits, ops)))
    flags, h0 = int(p["flags"]), int(p["hit_offset"])
    phs = []
    for i in range(int(p["n_hits"])):
        a, b = hits[h0 + 2 * i], hits[h0 + 2 * i + 1]
        phs.append(PairHit(int(a["reference"]), bool(a["template_reverse"]),
                           (_hit(a, ops), _hit(b, ops)), int(a["insert"])))
    return ProperPair(int(p["score"]), int(p["n_tied"]),
                      tuple(int(p["subopt"][m]) if flags & (1 << m) else None for m in range(2)),
                      tuple(phs))


class _MappedOwner:
    """Owns an esch_mapped whose buffers arrays view in place; frees it when
    the last such array is gone."""

    __slots__ = ("m",)

    def __init__(self, m):
        self.m = m

    def __del__(self):
        m, self.m = self.m, None
        if m is not None and m.value and _lib is not None:
            _lib.esch_mapped_free(m)


def _view(owner: _MappedOwner, ptr, n: int, dt: np.dtype) -> np.ndarray:
    """n records of dtype dt at a pointer into owner's esch_mapped, viewed
    without a copy (writable); the view keeps owner, and so the buffer,
    alive."""
    if not n:
        return np.empty(0, dt)
    buf = (C.c_char * (n * dt.itemsize)).from_address(ptr)
    buf._owner = owner
    return np.frombuffer(buf, dt)


def _take_mapped(m, pairs: bool):
    """An esch_mapped's records, hits and operation runs as arrays viewing
    its buffers (freed with the last of them), and its stats: (records,
    hits, ops, stats dict)."""
    L = _lib
    owner = _MappedOwner(m)
    n = int(L.esch_mapped_len(m))
    dt = PAIR_MAPPING_DTYPE if pairs else READ_MAPPING_DTYPE
    rp = L.esch_mapped_pairs(m) if pairs else L.esch_mapped_reads(m)
    nh, no = C.c_size_t(), C.c_size_t()
    hp = L.esch_mapped_hits(m, C.byref(nh))
    op = L.esch_mapped_ops(m, C.byref(no))
    recs = _view(owner, rp, n, dt)
    hits = _view(owner, hp, nh.value, HIT_DTYPE)
    ops = _view(owner, op, no.value, np.dtype(np.uint32))
    st = MapStats()
    L.esch_mapped_stats(m, C.byref(st))
    stats = {f: getattr(st, f) for f, _ in MapStats._fields_ if f != "reserved"}
    return recs, hits, ops, stats


def _runs(ops) -> np.ndarray:
    return np.array([("MID".index(o) << 30) | int(n) for o, n in ops] or [0], np.uint32)


def md_nm(read: bytes, reference: bytes, query_start: int, ref_start: int, ops):
    """sam::md_nm (lib/sam.rs:76-124): (MD, NM) as samtools calmd writes them
    for `ops` ([(op, len)], op in "MID") aligning read letters (as aligned)
    against reference letters from query_start / ref_start."""
    L = load_library()
    r = _runs(ops)
    nm = C.c_uint32()
    need = L.esch_md_nm(read, len(read), reference, len(reference), query_start, ref_start, r.ctypes.data,
                        len(ops), None, 0, C.byref(nm))
    if need == C.c_size_t(-1).value:
        raise ValueError("the operations run past the read or the reference")
    buf = C.create_string_buffer(need + 1)
    L.esch_md_nm(read, len(read), reference, len(reference), query_start, ref_start, r.ctypes.data, len(ops),
                 buf, need + 1, C.byref(nm))
    return buf.value.decode(), nm.value


def cigar_string(ops, query_len: int, query_start: int, query_end: int) -> str:
    """sam::cigar_string (lib/sam.rs:32-52): soft clips around the ops."""
    L = load_library()
    r = _runs(ops)
    need = L.esch_cigar_string(r.ctypes.data, len(ops), query_len, query_start, query_end, None, 0)
    if need == C.c_size_t(-1).value:
        raise ValueError("inconsistent span")
    buf = C.create_string_buffer(need + 1)
    L.esch_cigar_string(r.ctypes.data, len(ops), query_len, query_start, query_end, buf, need + 1)
    return buf.value.decode()


def valu_peak16(ordinal: int = 0) -> dict:
    """The measured packed-int16 VALU peak (lane-ops/s) with the CU count and
    the reported peak clock of device `ordinal`."""
    v, cu, clk = C.c_double(), C.c_int(), C.c_int()
    _check(load_library().esch_valu_peak16(ordinal, C.byref(v), C.byref(cu), C.byref(clk)))
    return {"lane_ops_per_s": v.value, "n_cu": cu.value, "clock_khz": clk.value}


@dataclass(frozen=True)
class VariantInfo:
    schedule: int
    stripe_rows: int
    f16: bool
    shifted: int  # 0 none, 1 one axis (rows / columns), 2 both (the diagonal shift)

    @property
    def diag(self) -> bool:
        """The diagonal shift (semiglobal mode only)."""
        return self.shifted == 2


def variant_info(variant: int) -> VariantInfo:
    """esch_variant_info: a batched-kernel variant's schedule, rows per
    stripe, fp16 arithmetic (the 2048 condition) and shifted values."""
    out = [C.c_int() for _ in range(4)]
    _check(load_library().esch_variant_info(variant, *[C.byref(x) for x in out]))
    return VariantInfo(out[0].value, out[1].value, bool(out[2].value), out[3].value)


def variants() -> dict:
    """Every batched-kernel variant's VariantInfo, by number."""
    return {v: variant_info(v) for v in range(load_library().esch_n_variants())}


def device_visible() -> bool:
    """cuda::device_visible (lib/cuda/mod.rs:130-140)."""
    return bool(load_library().esch_device_visible())


# ---------------------------------------------------------------- alphabet

_RC_LETTERS = bytes(
    {ord("A"): ord("T"), ord("C"): ord("G"), ord("G"): ord("C"), ord("T"): ord("A"), ord("U"): ord("A")}.get(b, ord("N"))
    for b in range(256))
_TABLE = np.full(256, 4, dtype=np.uint8)
for _ch, _c in ((b"A", 0), (b"C", 1), (b"G", 2), (b"T", 3), (b"U", 3)):
    _TABLE[_ch[0]] = _c
    _TABLE[_ch.lower()[0]] = _c
WILDCARD = 4


def encode(letters) -> np.ndarray:
    """Letters -> codes (lib/alphabet.rs:29-55): A/C/G/T(U) 0..3, else 4."""
    if isinstance(letters, str):
        letters = letters.encode()
    if len(letters) == 0:
        return np.zeros(0, np.uint8)
    return _TABLE[np.frombuffer(bytes(letters), dtype=np.uint8)]


def reverse_complement_codes(codes) -> np.ndarray:
    """lib/alphabet.rs:58-64: reversed, 3 - c, wildcard kept."""
    c = np.asarray(codes, dtype=np.uint8)[::-1].copy()
    m = c != WILDCARD
    c[m] = 3 - c[m]
    return c


# ---------------------------------------------------------------- scoring

@dataclass(frozen=True)
class Scoring:
    """lib/scoring.rs:27-137. Default 2,-1,-10,-1."""

    match_score: int = 2
    mismatch: int = -1
    gap_open: int = -10
    gap_extend: int = -1

    def __post_init__(self):
        st = load_library().esch_scoring_validate(self.match_score, self.mismatch,
                                                  self.gap_open, self.gap_extend)
        if st != OK:
            raise ValueError(load_library().esch_last_error().decode())

    @classmethod
    def parse(cls, s: str) -> "Scoring":
        parts = [p.strip() for p in s.split(",")]
        if len(parts) != 4:
            raise ValueError(f"scoring {s!r} is not MATCH,MISMATCH,GAP_OPEN,GAP_EXTEND")
        try:
            v = [int(p) for p in parts]
        except ValueError:
            raise ValueError(f"scoring {s!r} is not MATCH,MISMATCH,GAP_OPEN,GAP_EXTEND") from None
        return cls(*v)

    def substitution(self, a: int, b: int) -> int:
        return self.match_score if (a == b or a == WILDCARD or b == WILDCARD) else self.mismatch

    def gap(self, length: int) -> int:
        return self.gap_open + (length - 1) * self.gap_extend

    def tuple(self):
        return (self.match_score, self.mismatch, self.gap_open, self.gap_extend)

    def __str__(self):
        return ",".join(map(str, self.tuple()))


class Mode:
    """lib/scoring.rs:151-190."""

    LOCAL = 0
    SEMIGLOBAL = 1

    @staticmethod
    def parse(s: str) -> int:
        if s == "local":
            return Mode.LOCAL
        if s in ("semiglobal", "semi-global", "overlap"):
            return Mode.SEMIGLOBAL
        raise ValueError(f"unknown mode {s!r}, expected local|semiglobal")

    @staticmethod
    def name(m: int) -> str:
        return "local" if m == Mode.LOCAL else "semiglobal"


def fits_i16(n: int, max_ref_len: int, scoring: Scoring) -> bool:
    """lib/simd/mod.rs:222-227."""
    return bool(load_library().esch_fits_i16(n, max_ref_len, *scoring.tuple()))


# ---------------------------------------------------------------- panel

class Panel:
    """lib/panel.rs:10-82: references in the order given (= tie order)."""

    def __init__(self, refs):
        self.names: List[str] = []
        self.seqs: List[bytes] = []
        self.codes: List[np.ndarray] = []
        for item in refs:
            if isinstance(item, (tuple, list)) and len(item) == 2 and isinstance(item[0], str):
                name, seq = item
            else:
                name, seq = f"r{len(self.names)}", item
            if isinstance(seq, np.ndarray):
                codes = np.ascontiguousarray(seq, dtype=np.uint8)
                seq = bytes(b"ACGTN"[c] for c in codes)
            else:
                seq = seq.encode() if isinstance(seq, str) else bytes(seq)
                codes = encode(seq)
            if len(seq) == 0:
                raise ValueError(f"reference {name!r} has no bases")
            self.names.append(name)
            self.seqs.append(seq)
            self.codes.append(codes)
        if not self.names:
            raise ValueError("the reference panel is empty")

    def __len__(self):
        return len(self.names)

    def max_len(self) -> int:
        return max(len(c) for c in self.codes)

    def total_len(self) -> int:
        return int(sum(len(c) for c in self.codes))


# ---------------------------------------------------------------- adapters

TRUSEQ_R1 = b"AGATCGGAAGAGCACACGTCTGAACTCCAGTCA"  # lib/adapter.rs:40
TRUSEQ_R2 = b"AGATCGGAAGAGCGTCGTGTAGGGAAAGAGTGT"  # lib/adapter.rs:42
DEFAULT_MIN_OVERLAP = 5                            # lib/adapter.rs:44


def adapter_find(read: bytes, adapter: bytes = TRUSEQ_R1, min_overlap: int = DEFAULT_MIN_OVERLAP) -> int:
    """adapter::find (lib/adapter.rs:68-91), letters in."""
    return int(load_library().esch_adapter_find(bytes(read), len(read), bytes(adapter),
                                                 len(adapter), min_overlap))


def adapter_trim_pairs(codes: np.ndarray, offs: np.ndarray, lens: np.ndarray, a1: bytes = TRUSEQ_R1,
                       a2: bytes = TRUSEQ_R2, min_overlap: int = DEFAULT_MIN_OVERLAP) -> np.ndarray:
    """esch_adapter_trim_pairs: mates interleaved as codes; each mate's
    length less its detected 3' adapter (int32, one per mate)."""
    L = load_library()
    codes = np.ascontiguousarray(codes, np.uint8)
    offs = np.ascontiguousarray(offs, np.uint64)
    lens = np.ascontiguousarray(lens, np.int32)
    if len(lens) % 2:
        raise ValueError("mates come in pairs")
    out = np.zeros(len(lens), np.int32)
    if len(lens):
        L.esch_adapter_trim_pairs(codes.ctypes.data, offs.ctypes.data, lens.ctypes.data, len(lens) // 2,
                                  a1, len(a1), a2, len(a2), min_overlap, out.ctypes.data)
    return out


def adapter_detect_pair(r1: bytes, r2: bytes, a1: bytes = TRUSEQ_R1, a2: bytes = TRUSEQ_R2,
                        min_overlap: int = DEFAULT_MIN_OVERLAP):
    """adapter::detect_pair (lib/adapter.rs:95-110): [ad1, ad2]."""
    out = (C.c_size_t * 2)()
    load_library().esch_adapter_detect_pair(bytes(r1), len(r1), bytes(r2), len(r2), bytes(a1),
                                            len(a1), bytes(a2), len(a2), min_overlap,
                                            C.byref(out))
    return [int(out[0]), int(out[1])]


# ---------------------------------------------------------------- score matrix

class ScoreMatrix:
    """lib/mapper.rs:94-139: row-major i16 scores + which rows are present."""

    def __init__(self, n_refs: int, scores: np.ndarray, scored: np.ndarray):
        scores = np.asarray(scores)
        scored = np.asarray(scored, dtype=bool)
        if scores.size != scored.size * n_refs:
            raise ValueError("row-major matrix")
        self._n_refs = n_refs
        self.scores = scores.reshape(scored.size, n_refs) if n_refs else scores.reshape(scored.size, 0)
        self.scored = scored

    from_parts = classmethod(lambda cls, n_refs, scores, scored: cls(n_refs, scores, scored))

    def __len__(self):
        return int(self.scored.size)

    def is_empty(self):
        return self.scored.size == 0

    def n_refs(self):
        return self._n_refs

    def n_scored(self):
        return int(self.scored.sum())

    def row(self, k: int) -> Optional[np.ndarray]:
        return self.scores[k] if self.scored[k] else None


# ---------------------------------------------------------------- selection records

@dataclass
class ReadSelect:
    """Score stage of Aligner::map_reads_scored (lib/mapper.rs:444-472)."""

    best: Optional[int]
    n_tied: int
    suboptimal: Optional[int]
    take: List[tuple]  # (ref, reverse), tie order, truncated to max_ties + 1


@dataclass
class PairSelect:
    """Top tier of lib/mapper/pair.rs:185-231."""

    n_cand: int
    top_sum: Optional[int]
    top: List[tuple]  # (ref, template_reverse)
    n_top: int
    next_sum: Optional[int]


def _ptr_array(seqs):
    """(keepalive, char** array, size_t* lens) for a list of code arrays."""
    arrs = [np.ascontiguousarray(s, dtype=np.uint8) for s in seqs]
    n = len(arrs)
    ptrs = (C.c_void_p * max(n, 1))()
    lens = (C.c_size_t * max(n, 1))()
    for i, a in enumerate(arrs):
        ptrs[i] = a.ctypes.data if a.size else None
        lens[i] = a.size
    return arrs, ptrs, lens


# ---------------------------------------------------------------- scorer

class PreparedBatch:
    def __init__(self, ptr, n):
        self._ptr, self.n = ptr, n

    def __del__(self):
        if getattr(self, "_ptr", None) and _lib is not None:
            _lib.esch_prepared_free(self._ptr)


class PendingBatch:
    def __init__(self, ptr, n):
        self._ptr, self.n = ptr, n


class GpuScorer:
    """lib/cuda/mod.rs:328-709: a panel resident on the device under one
    scoring and one mode."""

    def __init__(self, panel: Panel, scoring: Optional[Scoring] = None, mode: int = Mode.LOCAL,
                 ordinal: int = 0):
        L = load_library()
        scoring = Scoring() if scoring is None else scoring
        self.panel = panel
        self.scoring = scoring
        self.mode = mode
        keep, ptrs, lens = _ptr_array(panel.codes)
        out = C.c_void_p()
        _check(L.esch_scorer_new(ordinal, ptrs, lens, len(panel.codes), *scoring.tuple(), mode,
                                 C.byref(out)))
        self._h = out
        self.n_refs = len(panel)
        self._rlen = [len(c) for c in panel.codes]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.esch_scorer_free(h)
            self._h = None

    close = __del__

    # diagnostics (lib/cuda/mod.rs:446-488)
    def device_name(self) -> str:
        buf = C.create_string_buffer(256)
        _check(_lib.esch_device_name(self._h, buf, 256))
        return buf.value.decode()

    def geometry(self):
        a, b, c = C.c_size_t(), C.c_size_t(), C.c_size_t()
        _lib.esch_geometry(self._h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def registers(self) -> int:
        return _lib.esch_registers(self._h)

    def local_bytes(self) -> int:
        return _lib.esch_local_bytes(self._h)

    def scores_len(self, n: int) -> bool:
        return bool(_lib.esch_scores_len(self._h, n))

    # prepare / submit / finish (lib/cuda/mod.rs:524-708)
    def prepare(self, queries: Sequence[np.ndarray]) -> PreparedBatch:
        keep, ptrs, lens = _ptr_array(queries)
        p = _lib.esch_prepare(self._h, ptrs, lens, len(keep))
        if not p:
            raise GpuError(INVALID_ARG, _lib.esch_last_error().decode())
        return PreparedBatch(p, len(keep))

    def submit(self, prepared: PreparedBatch) -> PendingBatch:
        ptr = prepared._ptr
        prepared._ptr = None  # consumed
        out = C.c_void_p()
        _check(_lib.esch_submit(self._h, ptr, C.byref(out)))
        return PendingBatch(out, prepared.n)

    def finish(self, pending: PendingBatch) -> ScoreMatrix:
        n = pending.n
        scores = np.zeros(max(n, 1) * max(self.n_refs, 1), np.int16)
        scored = np.zeros(max(n, 1), np.uint8)
        ptr = pending._ptr
        pending._ptr = None
        _check(_lib.esch_finish(self._h, ptr, scores.ctypes.data, scored.ctypes.data))
        return ScoreMatrix(self.n_refs, scores[: n * self.n_refs], scored[:n].astype(bool))

    def score_batch(self, queries: Sequence[np.ndarray]) -> ScoreMatrix:
        return self.finish(self.submit(self.prepare(queries)))

    # extensions
    def score_batch_i32(self, queries: Sequence[np.ndarray]):
        keep, ptrs, lens = _ptr_array(queries)
        n = len(keep)
        scores = np.zeros(max(n, 1) * self.n_refs, np.int32)
        scored = np.zeros(max(n, 1), np.uint8)
        _check(_lib.esch_score_batch_i32(self._h, ptrs, lens, n, scores.ctypes.data,
                                         scored.ctypes.data))
        return scores[: n * self.n_refs].reshape(n, self.n_refs), scored[:n].astype(bool)

    def score_device_noalloc(self, codes: int, offs: int, lens: int, n: int, scores: int, scored: int,
                             stream: int = 0):
        """esch_score_device_noalloc: the scoring of score_batch_i32 on
        queries already in device memory (addresses as ints: codes u8,
        offs u64, lens i32; out scores i32[n * n_refs], scored u8[n]),
        queued on `stream` without any device allocation."""
        _check(_lib.esch_score_device_noalloc(self._h, codes, offs, lens, n, scores, scored, stream or None))

    def select_batch(self, reads: Sequence[np.ndarray], both_strands=False, min_score=0,
                     max_ties=None, tie_cap=None) -> List[ReadSelect]:
        keep, ptrs, lens = _ptr_array(reads)
        n = len(keep)
        cap = tie_cap or 2 * self.n_refs
        recs = np.zeros((max(n, 1), 5), np.int32)
        ties = np.zeros(max(n, 1) * cap, np.int32)
        mt = -1 if max_ties is None else int(max_ties)
        _check(_lib.esch_select_batch(self._h, ptrs, lens, n, int(both_strands), min_score, mt,
                                      cap, recs.ctypes.data, ties.ctypes.data))
        return [_read_select(recs[k], ties[k * cap:(k + 1) * cap]) for k in range(n)]

    def select_pairs(self, mates1, mates2, both_strands=False, min_score=0, top_cap=None):
        k1, p1, l1 = _ptr_array(mates1)
        k2, p2, l2 = _ptr_array(mates2)
        n = len(k1)
        if len(k2) != n:
            raise ValueError("mates1 and mates2 differ in length")
        cap = top_cap or 2 * self.n_refs
        recs = np.zeros((max(n, 1), 5), np.int32)
        top = np.zeros(max(n, 1) * cap, np.int32)
        _check(_lib.esch_select_pairs(self._h, p1, l1, p2, l2, n, int(both_strands), min_score,
                                      cap, recs.ctypes.data, top.ctypes.data))
        return [_pair_select(recs[k], top[k * cap:(k + 1) * cap]) for k in range(n)]

    # kernel variants and the autotuner
    @property
    def variant(self) -> int:
        return _lib.esch_get_variant(self._h)

    @variant.setter
    def variant(self, v: int):
        _check(_lib.esch_set_variant(self._h, int(v)))

    def column_block(self, variant: int) -> int:
        """esch_column_block: the columns after which `variant`'s column
        shift restarts on this panel (0: never; -1: refused)."""
        return int(_lib.esch_column_block(self._h, int(variant)))

    def usable_variants(self) -> list:
        """The batched-kernel variants this scorer admits (its mode, and the
        panel within each variant's exactness condition)."""
        return [v for v in range(_lib.esch_n_variants()) if _lib.esch_variant_usable(self._h, v)]

    @property
    def diag_min_len(self) -> int:
        """Reads this long take the anti-diagonal kernel (0: none)."""
        return _lib.esch_get_diag_min_len(self._h)

    @diag_min_len.setter
    def diag_min_len(self, v: int):
        _check(_lib.esch_set_diag_min_len(self._h, int(v)))

    @property
    def swap_rows(self) -> int:
        """Batches of up to this many rows are scored with the roles of
        reads and references swapped (identical scores); set -1 for the
        default rule, 0 to turn it off."""
        return _lib.esch_get_swap_rows(self._h)

    @swap_rows.setter
    def swap_rows(self, v: int):
        _check(_lib.esch_set_swap_rows(self._h, int(v)))

    def align_pairs_raw(self, queries: Sequence[np.ndarray], refs: Sequence[int]):
        """esch_align_pairs: (records int64 [n, 7] = score, query_start,
        query_end, ref_start, ref_end, n_ops, ops_offset; ops uint32, each
        op << 30 | length with op 0 = M, 1 = I, 2 = D)."""
        if len(queries) != len(refs):
            raise ValueError("one reference per query")
        keep, ptrs, lens = _ptr_array(queries)
        n = len(keep)
        refs_a = np.ascontiguousarray(refs, np.int32)
        recs = np.zeros((max(n, 1), 4), np.int64)  # esch_alignment: 6 x int32, then int64 at byte 24
        nr = len(self._rlen)
        rl = np.asarray(self._rlen, np.int64)
        ok = (refs_a >= 0) & (refs_a < nr)
        cap = int(sum(len(q) for q in keep)) + int(rl[refs_a[ok]].sum()) + 1
        ops = np.zeros(cap, np.uint32)
        used = C.c_size_t()
        _check(_lib.esch_align_pairs(self._h, ptrs, lens, refs_a.ctypes.data, n, recs.ctypes.data,
                                     ops.ctypes.data, cap, C.byref(used)))
        r32 = recs.view(np.int32).reshape(-1, 8)[:n]
        out = np.zeros((n, 7), np.int64)
        out[:, :6] = r32[:, :6]
        out[:, 6] = recs[:n, 3]
        return out, ops[:used.value]

    def align_pairs_concat(self, codes: np.ndarray, offs: np.ndarray, lens: np.ndarray, refs):
        """esch_align_pairs_concat: read k is codes[offs[k]:offs[k] + lens[k]];
        returns what align_pairs_raw does."""
        codes = np.ascontiguousarray(codes, np.uint8)
        offs = np.ascontiguousarray(offs, np.uint64)
        lens = np.ascontiguousarray(lens, np.int32)
        refs_a = np.ascontiguousarray(refs, np.int32)
        n = len(lens)
        if len(refs_a) != n:
            raise ValueError("one reference per query")
        rl = np.asarray(self._rlen, np.int64)
        ok = (refs_a >= 0) & (refs_a < len(rl))
        cap = int(lens.astype(np.int64).sum()) + int(rl[refs_a[ok]].sum()) + 1
        recs = np.zeros((max(n, 1), 4), np.int64)
        ops = np.zeros(cap, np.uint32)
        used = C.c_size_t()
        _check(_lib.esch_align_pairs_concat(self._h, codes.ctypes.data, offs.ctypes.data, lens.ctypes.data,
                                            refs_a.ctypes.data, n, recs.ctypes.data, ops.ctypes.data, cap,
                                            C.byref(used)))
        r32 = recs.view(np.int32).reshape(-1, 8)[:n]
        out = np.zeros((n, 7), np.int64)
        out[:, :6] = r32[:, :6]
        out[:, 6] = recs[:n, 3]
        return out, ops[:used.value]

    def align_pairs(self, queries: Sequence[np.ndarray], refs: Sequence[int]) -> list:
        """simd::align_pairs (lib/simd/mod.rs:405-499) on the device: the
        alignment of queries[k] (codes) against panel reference refs[k], in
        this scorer's scoring and mode."""
        recs, ops = self.align_pairs_raw(queries, refs)
        out = []
        for r in recs:
            seg = ops[int(r[6]):int(r[6]) + int(r[5])]
            out.append(Alignment(int(r[0]), int(r[1]), int(r[2]), int(r[3]), int(r[4]),
                                 tuple(("MID"[int(x) >> 30], int(x) & 0x3FFFFFFF) for x in seg)))
        return out

    def set_map_chunk(self, reads: int):
        """Reads (mates) per device chunk of the mapping calls (0: by memory)."""
        _check(_lib.esch_set_map_chunk(self._h, int(reads)))

    def map_reads_raw(self, reads: Sequence[np.ndarray], both_strands=False, min_score=0, max_ties=None):
        """esch_map_reads: (records READ_MAPPING_DTYPE, hits HIT_DTYPE, ops,
        stats)."""
        keep, ptrs, lens = _ptr_array(reads)
        m = C.c_void_p()
        mt = -1 if max_ties is None else int(max_ties)
        _check(_lib.esch_map_reads(self._h, ptrs, lens, len(keep), int(both_strands), int(min_score), mt,
                                   C.byref(m)))
        return _take_mapped(m, False)

    def map_reads_concat(self, codes: np.ndarray, offs: np.ndarray, lens: np.ndarray, both_strands=False,
                         min_score=0, max_ties=None):
        """esch_map_reads_concat: read k is codes[offs[k]:offs[k] + lens[k]];
        returns what map_reads_raw does."""
        codes = np.ascontiguousarray(codes, np.uint8)
        offs = np.ascontiguousarray(offs, np.uint64)
        lens = np.ascontiguousarray(lens, np.int32)
        m = C.c_void_p()
        mt = -1 if max_ties is None else int(max_ties)
        _check(_lib.esch_map_reads_concat(self._h, codes.ctypes.data, offs.ctypes.data, lens.ctypes.data,
                                          len(lens), int(both_strands), int(min_score), mt, C.byref(m)))
        return _take_mapped(m, False)

    def map_reads(self, reads: Sequence[np.ndarray], both_strands=False, min_score=0, max_ties=None) -> list:
        """Aligner::map_reads_scored (lib/mapper.rs:400-550) on the device:
        one ReadMapping per read (codes)."""
        recs, hits, ops, _ = self.map_reads_raw(reads, both_strands, min_score, max_ties)
        return [_read_mapping(r, hits, ops) for r in recs]

    def map_pairs_raw(self, mates1: Sequence[np.ndarray], mates2: Sequence[np.ndarray], both_strands=False,
                      min_score=0, max_ties=None, max_insert=500):
        """esch_map_pairs: (records PAIR_MAPPING_DTYPE, hits HIT_DTYPE, ops,
        stats)."""
        if len(mates1) != len(mates2):
            raise ValueError("mates1 and mates2 differ in length")
        k1, p1, l1 = _ptr_array(mates1)
        k2, p2, l2 = _ptr_array(mates2)
        m = C.c_void_p()
        mt = -1 if max_ties is None else int(max_ties)
        _check(_lib.esch_map_pairs(self._h, p1, l1, p2, l2, len(k1), int(both_strands), int(min_score), mt,
                                   int(max_insert), C.byref(m)))
        return _take_mapped(m, True)

    def map_pairs_concat(self, codes: np.ndarray, offs: np.ndarray, lens: np.ndarray, both_strands=False,
                         min_score=0, max_ties=None, max_insert=500):
        """esch_map_pairs_concat: mates interleaved (2p mate 1, 2p + 1 mate 2)
        in one buffer; returns what map_pairs_raw does."""
        codes = np.ascontiguousarray(codes, np.uint8)
        offs = np.ascontiguousarray(offs, np.uint64)
        lens = np.ascontiguousarray(lens, np.int32)
        if len(lens) % 2:
            raise ValueError("mates come in pairs")
        m = C.c_void_p()
        mt = -1 if max_ties is None else int(max_ties)
        _check(_lib.esch_map_pairs_concat(self._h, codes.ctypes.data, offs.ctypes.data, lens.ctypes.data,
                                          len(lens) // 2, int(both_strands), int(min_score), mt, int(max_insert),
                                          C.byref(m)))
        return _take_mapped(m, True)

    def map_pairs(self, pairs, both_strands=False, min_score=0, max_ties=None, max_insert=500) -> list:
        """Aligner::map_pairs_scored (lib/mapper/pair.rs:157-387) on the
        device: pairs = [(mate 1, mate 2), ...] as codes (adapters trimmed);
        one ProperPair or SeparatePair each."""
        recs, hits, ops, _ = self.map_pairs_raw([p[0] for p in pairs], [p[1] for p in pairs], both_strands,
                                                min_score, max_ties, max_insert)
        return [_pair_mapping(r, hits, ops) for r in recs]

    def align_last_ms(self) -> float:
        """Device time of the last align_pairs call's traceback kernels."""
        return float(_lib.esch_align_last_ms(self._h))

    def align_last_wavefront(self) -> int:
        """Pairs of the last align_pairs call traced by the anti-diagonal
        (one wave per pair) kernel."""
        return int(_lib.esch_align_last_wavefront(self._h))

    @property
    def diag_ratio(self) -> float:
        """Adaptive routing per length bucket: reads of length L take the
        anti-diagonal kernel when L > ratio * L' + 64 (L' the read 127 ranks
        later) or when they lead a mostly empty chunk; 0.0 turns it off."""
        return _lib.esch_get_diag_ratio(self._h) / 16.0

    @diag_ratio.setter
    def diag_ratio(self, ratio: float):
        _check(_lib.esch_set_diag_ratio(self._h, int(round(ratio * 16))))

    @property
    def diag_plan(self) -> List[int]:
        """Per length class (LEN_CLASS_BOUNDS), the routing to the
        anti-diagonal kernel: 0 never, 1 always, 16..1024 the ratio rule
        (rho x 16)."""
        plan = np.zeros(LEN_CLASSES, np.int32)
        _check(_lib.esch_get_diag_plan(self._h, plan.ctypes.data, LEN_CLASSES))
        return [int(x) for x in plan]

    @diag_plan.setter
    def diag_plan(self, plan):
        plan = np.ascontiguousarray(plan, np.int32)
        _check(_lib.esch_set_diag_plan(self._h, plan.ctypes.data, len(plan)))

    def autotune(self, reads: Sequence[np.ndarray], both_strands=True):
        """Time every batched-kernel variant on `reads`, keep the fastest,
        then sweep the per-length-bucket routing ratio (unless it was set);
        returns (chosen, [ms per variant]); `diag_ratio` reads the ratio."""
        keep, ptrs, lens = _ptr_array(reads)
        nv = _lib.esch_n_variants()
        ms = np.zeros(nv, np.float32)
        chosen = C.c_int()
        _check(_lib.esch_autotune(self._h, ptrs, lens, len(keep), int(both_strands),
                                  ms.ctypes.data, C.byref(chosen)))
        return chosen.value, [float(x) for x in ms]

    def batch(self, max_reads: int, max_bases: int) -> "DeviceBatch":
        return DeviceBatch(self, max_reads, max_bases)

    def batch_footprint(self, max_reads: int, max_bases: int) -> int:
        return int(_lib.esch_batch_footprint(self._h, max_reads, max_bases))


def _read_select(rec, ties) -> ReadSelect:
    best, nt, so, flags, ntake = (int(v) for v in rec)
    if not flags & 1:
        return ReadSelect(None, 0, None, [])
    take = [(int(t) >> 1, bool(t & 1)) for t in ties[:min(ntake, len(ties))]]
    return ReadSelect(best, nt, so if flags & 2 else None, take)


def _pair_select(rec, top) -> PairSelect:
    ncand, tsum, ntop, nsum, flags = (int(v) for v in rec)
    members = [(int(t) >> 1, bool(t & 1)) for t in top[:min(ntop, len(top))]]
    return PairSelect(ncand, tsum if ncand else None, members, ntop,
                      nsum if flags & 2 else None)


class DeviceBatch:
    """A device-resident batch: upload once, run the captured pipeline, read
    back what is needed (include/eschalign_amd.h, esch_batch_*)."""

    def __init__(self, scorer: GpuScorer, max_reads: int, max_bases: int):
        self.scorer = scorer
        out = C.c_void_p()
        _check(_lib.esch_batch_create(scorer._h, max_reads, max_bases, C.byref(out)))
        self._h = out
        self.n_reads = 0
        self.n_rows = 0
        self.pairs = False

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.esch_batch_free(h)
            self._h = None

    close = __del__

    def load(self, reads, both_strands=False, pairs=False):
        keep, ptrs, lens = _ptr_array(reads)
        _check(_lib.esch_batch_load(self._h, ptrs, lens, len(keep), int(both_strands), int(pairs)))
        self._set(len(keep), both_strands, pairs)

    def load_concat(self, codes: np.ndarray, offs: np.ndarray, lens: np.ndarray,
                    both_strands=False, pairs=False):
        codes = np.ascontiguousarray(codes, np.uint8)
        offs = np.ascontiguousarray(offs, np.uint64)
        lens = np.ascontiguousarray(lens, np.int32)
        _check(_lib.esch_batch_load_concat(self._h, codes.ctypes.data, offs.ctypes.data,
                                           lens.ctypes.data, len(lens), int(both_strands),
                                           int(pairs)))
        self._set(len(lens), both_strands, pairs)

    def _set(self, n, both, pairs):
        self.n_reads = n
        self.n_rows = 2 * n if both else n
        self.pairs = bool(pairs)

    def run(self, what: int, min_score=0, max_ties=None, tie_cap=1):
        mt = -1 if max_ties is None else int(max_ties)
        self.tie_cap = max(1, int(tie_cap))
        _check(_lib.esch_batch_run(self._h, what, min_score, mt, self.tie_cap))
        self.what = what

    def wait(self):
        _check(_lib.esch_batch_wait(self._h))

    def timing(self):
        dp, tot = C.c_float(), C.c_float()
        _check(_lib.esch_batch_timing(self._h, C.byref(dp), C.byref(tot)))
        return dp.value, tot.value

    def stats(self):
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(_lib.esch_batch_stats(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def n_diag(self) -> int:
        v = C.c_uint64()
        _check(_lib.esch_batch_stats_diag(self._h, C.byref(v)))
        return v.value

    def swapped(self) -> bool:
        """The last run scored with the roles of reads and references swapped."""
        return bool(_lib.esch_batch_swapped(self._h))

    def variant(self) -> int:
        """The batched kernel variant of the last run (-1: none, or swapped)."""
        return int(_lib.esch_batch_variant(self._h))

    def scores32(self):
        nr = self.scorer.n_refs
        s = np.zeros(max(self.n_rows, 1) * nr, np.int32)
        ok = np.zeros(max(self.n_rows, 1), np.uint8)
        _check(_lib.esch_batch_read_scores32(self._h, s.ctypes.data, ok.ctypes.data))
        return s[: self.n_rows * nr].reshape(self.n_rows, nr), ok[: self.n_rows].astype(bool)

    def scores16(self):
        nr = self.scorer.n_refs
        s = np.zeros(max(self.n_rows, 1) * nr, np.int16)
        ok = np.zeros(max(self.n_rows, 1), np.uint8)
        _check(_lib.esch_batch_read_scores16(self._h, s.ctypes.data, ok.ctypes.data))
        return s[: self.n_rows * nr].reshape(self.n_rows, nr), ok[: self.n_rows].astype(bool)

    def select(self) -> List[ReadSelect]:
        n, cap = self.n_reads, self.tie_cap
        recs = np.zeros((max(n, 1), 5), np.int32)
        ties = np.zeros(max(n, 1) * cap, np.int32)
        _check(_lib.esch_batch_read_select(self._h, recs.ctypes.data, ties.ctypes.data))
        return [_read_select(recs[k], ties[k * cap:(k + 1) * cap]) for k in range(n)]

    def select_raw(self):
        n, cap = self.n_reads, self.tie_cap
        recs = np.zeros((max(n, 1), 5), np.int32)
        ties = np.zeros(max(n, 1) * cap, np.int32)
        _check(_lib.esch_batch_read_select(self._h, recs.ctypes.data, ties.ctypes.data))
        return recs[:n], ties[: n * cap].reshape(n, cap)

    def pairs_select(self) -> List[PairSelect]:
        n, cap = self.n_reads // 2, self.tie_cap
        recs = np.zeros((max(n, 1), 5), np.int32)
        top = np.zeros(max(n, 1) * cap, np.int32)
        _check(_lib.esch_batch_read_pairs(self._h, recs.ctypes.data, top.ctypes.data))
        return [_pair_select(recs[k], top[k * cap:(k + 1) * cap]) for k in range(n)]


class Server:
    """Continuous-batching request loop over device-resident panels
    (include/eschalign_amd.h, esch_server_*)."""

    def __init__(self, hbm_budget=32 << 30, max_batch_reads=50_000, max_batch_bases=8_000_000,
                 max_slots=2, ordinal=0, max_batch_cells=1 << 37):
        L = load_library()
        cfg = ServerConfig(ordinal, hbm_budget, max_batch_reads, max_batch_bases, max_slots,
                           max_batch_cells)
        out = C.c_void_p()
        _check(L.esch_server_new(C.byref(cfg), C.byref(out)))
        self._h = out
        self._caps = {}

    def close(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.esch_server_free(h)
            self._h = None

    __del__ = close

    def add_panel(self, panel: Panel, scoring: Optional[Scoring] = None, mode: int = Mode.LOCAL) -> int:
        scoring = Scoring() if scoring is None else scoring
        keep, ptrs, lens = _ptr_array(panel.codes)
        pid = C.c_int()
        _check(_lib.esch_server_add_panel(self._h, ptrs, lens, len(keep), *scoring.tuple(), mode,
                                          C.byref(pid)))
        self._caps.setdefault("n_refs", {})[pid.value] = len(panel)
        return pid.value

    def autotune(self, panel_id: int, reads, both_strands=True):
        h = _lib.esch_server_scorer(self._h, panel_id)
        keep, ptrs, lens = _ptr_array(reads)
        nv = _lib.esch_n_variants()
        ms = np.zeros(nv, np.float32)
        chosen = C.c_int()
        _check(_lib.esch_autotune(h, ptrs, lens, len(keep), int(both_strands), ms.ctypes.data,
                                  C.byref(chosen)))
        return chosen.value, [float(x) for x in ms]

    def submit(self, panel_id: int, reads, both_strands=True, min_score=0, max_ties=None,
               tie_cap=8) -> int:
        keep, ptrs, lens = _ptr_array(reads)
        rid = C.c_uint64()
        mt = -1 if max_ties is None else int(max_ties)
        _check(_lib.esch_server_submit(self._h, panel_id, ptrs, lens, len(keep), int(both_strands),
                                       min_score, mt, tie_cap, C.byref(rid)))
        self._caps[rid.value] = (len(keep), tie_cap)
        return rid.value

    def wait(self, rid: int, timeout_ms: int = -1) -> bool:
        st = _lib.esch_server_wait(self._h, rid, timeout_ms)
        if st == TIMEOUT:
            return False
        _check(st)
        return True

    def result(self, rid: int):
        """(list of ReadSelect, latency in seconds) of a finished request."""
        n, cap = self._caps.pop(rid)
        recs = np.zeros((max(n, 1), 5), np.int32)
        ties = np.zeros(max(n, 1) * cap, np.int32)
        lat = C.c_double()
        _check(_lib.esch_server_result(self._h, rid, recs.ctypes.data, ties.ctypes.data,
                                       C.byref(lat)))
        return [_read_select(recs[k], ties[k * cap:(k + 1) * cap]) for k in range(n)], lat.value

    def stats(self) -> dict:
        st = ServerStats()
        _check(_lib.esch_server_stats(self._h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}
