"""ctypes front-end to the CPU oracle (liboracle.so): TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module. It is the checker, never the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
U64P = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
U8P = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build() -> None:
    """Compile liboracle.so from oracle/*.c (plain gcc, seconds)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.eo_encode.argtypes = [U8P, C.c_size_t, U8P]
        L.eo_revcomp_codes.argtypes = [U8P, C.c_size_t, U8P]
        L.eo_scoring_validate.argtypes = [C.c_int32] * 4
        L.eo_scoring_validate.restype = C.c_int
        L.eo_score.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t,
                               C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int]
        L.eo_score.restype = C.c_int32
        L.eo_fits_i16.argtypes = [C.c_size_t, C.c_size_t] + [C.c_int32] * 4
        L.eo_align.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t] + [C.c_int32] * 4 + \
            [C.c_int, C.c_void_p, C.c_void_p, C.c_size_t]
        L.eo_align.restype = C.c_int
        L.eo_align_batch.argtypes = [U8P, U64P, I32P, U8P, U64P, I32P, I32P, C.c_size_t] + \
            [C.c_int32] * 4 + [C.c_int, C.c_int, I32P]
        L.eo_simd_align_pairs.argtypes = [U8P, U64P, I32P, U8P, U64P, I32P, I32P, C.c_size_t] + \
            [C.c_int32] * 4 + [C.c_int, C.c_int, I32P, C.c_void_p, C.c_void_p]
        L.eo_simd_align_pairs.restype = None
        L.eo_fits_i16.restype = C.c_int
        L.eo_score_matrix.argtypes = [U8P, U64P, I32P, C.c_size_t, U8P, U64P, I32P,
                                      C.c_size_t, C.c_int32, C.c_int32, C.c_int32,
                                      C.c_int32, C.c_int, C.c_int, I32P]
        L.eo_select.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int32, C.c_int64,
                                C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                C.c_void_p, C.c_size_t]
        L.eo_select.restype = C.c_int32
        L.eo_pair_cands.argtypes = [C.c_void_p * 4, C.c_size_t, C.c_int, C.c_int32,
                                    I32P, I32P, I32P, I32P]
        L.eo_pair_cands.restype = C.c_int32
        L.eo_map_reads.argtypes = [U8P, U64P, I32P, C.c_size_t, I32P, C.c_void_p, U8P, U64P, I32P,
                                   C.c_size_t] + [C.c_int32] * 4 + [C.c_int, C.c_int32, C.c_int64, C.c_int]
        L.eo_map_reads.restype = C.c_void_p
        L.eo_map_pairs.argtypes = [U8P, U64P, I32P, C.c_size_t, I32P, I32P, U8P, U64P, I32P,
                                   C.c_size_t] + [C.c_int32] * 4 + [C.c_int, C.c_int32, C.c_int64,
                                                                    C.c_int64, C.c_int, C.c_int]
        L.eo_map_pairs.restype = C.c_void_p
        L.eo_mapped_len.argtypes = [C.c_void_p]
        L.eo_mapped_len.restype = C.c_size_t
        for f in ("eo_mapped_reads", "eo_mapped_pairs"):
            getattr(L, f).argtypes = [C.c_void_p]
            getattr(L, f).restype = C.c_void_p
        for f in ("eo_mapped_hits", "eo_mapped_ops"):
            getattr(L, f).argtypes = [C.c_void_p, C.POINTER(C.c_size_t)]
            getattr(L, f).restype = C.c_void_p
        L.eo_mapped_free.argtypes = [C.c_void_p]
        L.eo_md_nm.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t, C.c_size_t, C.c_size_t,
                               C.c_void_p, C.c_size_t, C.c_char_p, C.c_size_t, C.POINTER(C.c_uint32)]
        L.eo_md_nm.restype = C.c_int
        L.eo_cigar_string.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, C.c_char_p,
                                      C.c_size_t]
        L.eo_cigar_string.restype = C.c_int
        L.eo_adapter_find.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t,
                                      C.c_size_t]
        L.eo_adapter_find.restype = C.c_size_t
        L.eo_adapter_detect_pair.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t,
                                             C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t,
                                             C.c_size_t, C.POINTER(C.c_size_t * 2)]
        L.eo_simd_supported.restype = C.c_int
        L.eo_profile_build.argtypes = [U8P, U64P, I32P, C.c_size_t, C.c_int32, C.c_int32,
                                       C.c_int32, C.c_int32, C.c_int]
        L.eo_profile_build.restype = C.c_void_p
        L.eo_profile_free.argtypes = [C.c_void_p]
        L.eo_simd_score_all.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_int,
                                        C.c_int, I32P]
        L.eo_simd_score_batch.argtypes = [C.c_void_p, U8P, U64P, I32P, C.c_size_t, C.c_int,
                                          C.c_int, I32P]
        _lib = L
    return _lib


def _codes(x) -> np.ndarray:
    return np.ascontiguousarray(np.frombuffer(bytes(x), dtype=np.uint8)) if not isinstance(
        x, np.ndarray) else np.ascontiguousarray(x, dtype=np.uint8)


def encode(letters: bytes) -> np.ndarray:
    a = _codes(letters)
    out = np.empty(len(a), dtype=np.uint8)
    lib().eo_encode(a if len(a) else np.zeros(1, np.uint8), len(a), out if len(a) else np.zeros(1, np.uint8))
    return out


def revcomp(codes) -> np.ndarray:
    a = _codes(codes)
    out = np.empty(len(a), dtype=np.uint8)
    if len(a):
        lib().eo_revcomp_codes(a, len(a), out)
    return out


def validate(ma, mi, o, e) -> int:
    return lib().eo_scoring_validate(ma, mi, o, e)


def score(q, r, scoring, local: bool) -> int:
    q = _codes(q)
    r = _codes(r)
    ma, mi, o, e = scoring
    return lib().eo_score(q.ctypes.data, len(q), r.ctypes.data, len(r), ma, mi, o, e, int(local))


OPS = "MID"


def align(q, r, scoring, local: bool):
    """eo_align (lib/scalar.rs:164-364): (score, query_start, query_end,
    ref_start, ref_end, [(op char, len), ...])."""
    q = _codes(q)
    r = _codes(r)
    rec = (C.c_int32 * 5)()
    cap = len(q) + len(r) + 1
    ops = np.zeros(cap, np.uint32)
    k = lib().eo_align(q.ctypes.data, len(q), r.ctypes.data, len(r), *scoring, int(local),
                       C.byref(rec), ops.ctypes.data, cap)
    assert k >= 0
    return (rec[0], rec[1], rec[2], rec[3], rec[4],
            [(OPS[int(x) >> 30], int(x) & 0x3FFFFFFF) for x in ops[:k]])


def align_batch(queries, refs_codes, ref_idx, scoring, local: bool, threads: int | None = None):
    """eo_align for query k against refs_codes[ref_idx[k]] on threads:
    int32 [n, 6] = score, query_start, query_end, ref_start, ref_end, runs."""
    return align_batch_concat(*concat(queries), refs_codes, ref_idx, scoring, local, threads)


def align_batch_concat(qc, qo, ql, refs_codes, ref_idx, scoring, local: bool, threads: int | None = None):
    """align_batch for queries qc[qo[k]:qo[k] + ql[k]]."""
    threads = threads or min(8, os.cpu_count() or 1)
    qc = np.ascontiguousarray(qc, np.uint8)
    qo = np.ascontiguousarray(qo, np.uint64)
    ql = np.ascontiguousarray(ql, np.int32)
    n = len(ql)
    rc, ro, rl = concat(refs_codes)
    idx = np.ascontiguousarray(ref_idx, np.int32)
    out = np.zeros((max(n, 1), 6), np.int32)
    if n:
        lib().eo_align_batch(qc, qo, ql, rc, ro, rl, idx, n, *scoring, int(local), threads,
                             out.reshape(-1))
    return out[:n]


def simd_align_batch_concat(qc, qo, ql, refs_codes, ref_idx, scoring, local: bool, threads: int | None = None,
                            with_ops=False):
    """simd::align_pairs with the AVX-512BW pair kernel (the reference's CPU
    traceback of winners) for queries qc[qo[k]:qo[k] + ql[k]] against
    refs_codes[ref_idx[k]]: int32 [n, 6] as align_batch_concat, and with
    with_ops the runs of pair k at ops[off[k]:off[k] + runs] as (recs, ops,
    off)."""
    threads = threads or min(8, os.cpu_count() or 1)
    qc = np.ascontiguousarray(qc, np.uint8)
    qo = np.ascontiguousarray(qo, np.uint64)
    ql = np.ascontiguousarray(ql, np.int32)
    n = len(ql)
    rc, ro, rl = concat(refs_codes)
    idx = np.ascontiguousarray(ref_idx, np.int32)
    out = np.zeros((max(n, 1), 6), np.int32)
    ops = off = None
    if with_ops:
        room = ql.astype(np.int64) + rl[idx].astype(np.int64) + 1 if n else np.zeros(0, np.int64)
        off = np.zeros(max(n, 1), np.uint64)
        if n > 1:
            off[1:n] = np.cumsum(room[:-1]).astype(np.uint64)
        ops = np.zeros(max(int(room.sum()) if n else 1, 1), np.uint32)
    if n:
        lib().eo_simd_align_pairs(qc, qo, ql, rc, ro, rl, idx, n, *scoring, int(local), threads, out.reshape(-1),
                                  None if ops is None else ops.ctypes.data, None if off is None else off.ctypes.data)
    if with_ops:
        return out[:n], ops, off[:n]
    return out[:n]


def fits_i16(n, max_ref_len, scoring) -> bool:
    return bool(lib().eo_fits_i16(n, max_ref_len, *scoring))


def concat(seqs):
    """Concatenate code arrays: (codes, offsets u64, lengths i32)."""
    lens = np.array([len(s) for s in seqs], dtype=np.int32)
    off = np.zeros(len(seqs), dtype=np.uint64)
    if len(seqs):
        off[1:] = np.cumsum(lens[:-1], dtype=np.uint64)
    codes = np.concatenate([_codes(s) for s in seqs]) if len(seqs) and lens.sum() else np.zeros(1, np.uint8)
    if len(codes) == 0:
        codes = np.zeros(1, np.uint8)
    return np.ascontiguousarray(codes), off, lens


def score_matrix(queries, refs, scoring, local: bool, threads: int | None = None) -> np.ndarray:
    """eo_score for every (query, ref): int32 [len(queries), len(refs)]."""
    threads = threads or min(8, os.cpu_count() or 1)
    qc, qo, ql = concat(queries)
    rc, ro, rl = concat(refs)
    out = np.zeros((max(len(queries), 1), max(len(refs), 1)), dtype=np.int32)
    if len(queries) and len(refs):
        lib().eo_score_matrix(qc, qo, ql, len(queries), rc, ro, rl, len(refs), *scoring,
                              int(local), threads, out.reshape(-1))
    return out[:len(queries), :len(refs)]


def select(fwd, rev, min_score=0, max_ties=None, cap=4096):
    """Score-stage selection of one read (lib/mapper.rs:444-472).

    Returns (best, n_tied, suboptimal or None, take list of (ref, reverse))."""
    fwd = np.ascontiguousarray(fwd, dtype=np.int32)
    revp = None
    if rev is not None:
        rev = np.ascontiguousarray(rev, dtype=np.int32)
        revp = rev.ctypes.data
    b, nt, hs, so = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    ties = np.zeros(cap, dtype=np.int32)
    mt = -1 if max_ties is None else int(max_ties)
    k = lib().eo_select(fwd.ctypes.data, revp, len(fwd), min_score, mt, C.byref(b), C.byref(nt),
                        C.byref(hs), C.byref(so), ties.ctypes.data, cap)
    take = [(int(t) >> 1, bool(t & 1)) for t in ties[:min(k, cap)]]
    return b.value, nt.value, (so.value if hs.value else None), take


def pair_cands(rows, min_score=0, both_strands=False):
    """Eligible candidates of one pair sorted by (-sum, ref, strand)
    (lib/mapper/pair.rs:185-231). rows[m][rv] or None for an empty mate."""
    n_refs = None
    arrs = []
    ptrs = (C.c_void_p * 4)()
    for m in range(2):
        for rv in range(2):
            row = rows[m][rv] if rows[m] is not None else None
            if row is None:
                ptrs[m * 2 + rv] = None
                arrs.append(None)
                continue
            a = np.ascontiguousarray(row, dtype=np.int32)
            arrs.append(a)
            n_refs = len(a)
            ptrs[m * 2 + rv] = a.ctypes.data
    if n_refs is None:
        return []
    cap = 2 * n_refs
    cand = np.zeros(cap, np.int32)
    sm = np.zeros(cap, np.int32)
    s0 = np.zeros(cap, np.int32)
    s1 = np.zeros(cap, np.int32)
    k = lib().eo_pair_cands(ptrs, n_refs, int(both_strands), min_score, cand, sm, s0, s1)
    return [(int(sm[i]), int(cand[i]) >> 1, bool(cand[i] & 1), int(s0[i]), int(s1[i]))
            for i in range(k)]


# Records of eo_map_reads / eo_map_pairs (the layout of the library's
# esch_hit, esch_read_mapping and esch_pair_mapping).
HIT_DTYPE = np.dtype([("reference", "<i4"), ("reverse", "<i4"), ("template_reverse", "<i4"),
                      ("insert", "<i4"), ("score", "<i4"), ("query_start", "<i4"),
                      ("query_end", "<i4"), ("ref_start", "<i4"), ("ref_end", "<i4"),
                      ("n_ops", "<i4"), ("ops_offset", "<i8")])
READ_DTYPE = np.dtype([("best", "<i4"), ("n_tied", "<i4"), ("subopt", "<i4"), ("flags", "<i4"),
                       ("n_hits", "<i4"), ("reserved", "<i4"), ("hit_offset", "<i8")])
PAIR_DTYPE = np.dtype([("proper", "<i4"), ("score", "<i4"), ("n_tied", "<i4"), ("flags", "<i4"),
                       ("subopt", "<i4", (2,)), ("n_hits", "<i4"), ("reserved", "<i4"),
                       ("hit_offset", "<i8"), ("separate", READ_DTYPE, (2,))])


def _mapped_arrays(L, m, pairs: bool):
    n = int(L.eo_mapped_len(m))
    nh, no = C.c_size_t(), C.c_size_t()
    hp = L.eo_mapped_hits(m, C.byref(nh))
    op = L.eo_mapped_ops(m, C.byref(no))
    dt = PAIR_DTYPE if pairs else READ_DTYPE
    rp = L.eo_mapped_pairs(m) if pairs else L.eo_mapped_reads(m)
    recs = np.frombuffer(C.string_at(rp, n * dt.itemsize), dtype=dt).copy() if n else np.zeros(0, dt)
    hits = (np.frombuffer(C.string_at(hp, nh.value * HIT_DTYPE.itemsize), dtype=HIT_DTYPE).copy()
            if nh.value else np.zeros(0, HIT_DTYPE))
    ops = (np.frombuffer(C.string_at(op, no.value * 4), dtype=np.uint32).copy()
           if no.value else np.zeros(0, np.uint32))
    return recs, hits, ops


def hit_form(h, ops):
    """A hit as (reference, reverse, score, query_start, query_end,
    ref_start, ref_end, ((op, len), ...))."""
    o0, k = int(h["ops_offset"]), int(h["n_ops"])
    runs = tuple((OPS[int(x) >> 30], int(x) & 0x3FFFFFFF) for x in ops[o0:o0 + k])
    return (int(h["reference"]), bool(h["reverse"]), int(h["score"]), int(h["query_start"]),
            int(h["query_end"]), int(h["ref_start"]), int(h["ref_end"]), runs)


def read_form(r, hits, ops):
    """ReadMapping (lib/mapper.rs:58-72) as (best_score or None, n_tied,
    suboptimal or None, (hit, ...))."""
    flags = int(r["flags"])
    h0, nh = int(r["hit_offset"]), int(r["n_hits"])
    return (int(r["best"]) if flags & 1 else None, int(r["n_tied"]),
            int(r["subopt"]) if flags & 2 else None,
            tuple(hit_form(hits[h0 + i], ops) for i in range(nh)))


def pair_form(p, hits, ops):
    """PairMapping (lib/mapper/pair.rs:90-107): ("proper", score, n_tied,
    (suboptimal per mate), ((reference, template reverse, insert, mate 1 hit,
    mate 2 hit), ...)) or ("separate", mate 1 ReadMapping, mate 2 ReadMapping)."""
    if int(p["proper"]):
        flags = int(p["flags"])
        sub = tuple(int(p["subopt"][m]) if flags & (1 << m) else None for m in range(2))
        h0 = int(p["hit_offset"])
        phs = []
        for i in range(int(p["n_hits"])):
            a, b = hits[h0 + 2 * i], hits[h0 + 2 * i + 1]
            phs.append((int(a["reference"]), bool(a["template_reverse"]), int(a["insert"]),
                        hit_form(a, ops), hit_form(b, ops)))
        return ("proper", int(p["score"]), int(p["n_tied"]), sub, tuple(phs))
    return ("separate", read_form(p["separate"][0], hits, ops), read_form(p["separate"][1], hits, ops))


def map_reads(reads, refs, scoring, local: bool, min_score=0, max_ties=None, both_strands=False,
              threads: int | None = None, rows=None, raw=False):
    """Aligner::map_reads_scored (lib/mapper.rs:400-550) on codes: one
    read_form per read. rows = (fwd, rev) panel scores if already computed.
    raw: the (records, hits, ops) arrays instead, in the library's layout."""
    threads = threads or min(8, os.cpu_count() or 1)
    L = lib()
    reads = [_codes(r) for r in reads]
    if rows is None:
        fwd = score_matrix(reads, refs, scoring, local, threads)
        rev = score_matrix([revcomp(r) for r in reads], refs, scoring, local, threads) if both_strands else None
    else:
        fwd, rev = rows
    fwd = np.ascontiguousarray(fwd, np.int32).reshape(-1)
    rev = None if rev is None else np.ascontiguousarray(rev, np.int32).reshape(-1)
    qc, qo, ql = concat(reads)
    rc, ro, rl = concat(refs)
    mt = -1 if max_ties is None else int(max_ties)
    m = L.eo_map_reads(qc, qo, ql, len(reads), fwd if len(fwd) else np.zeros(1, np.int32),
                       None if rev is None else rev.ctypes.data, rc, ro, rl, len(refs), *scoring,
                       int(local), min_score, mt, threads)
    try:
        recs, hits, ops = _mapped_arrays(L, m, False)
    finally:
        L.eo_mapped_free(m)
    if raw:
        return recs, hits, ops
    return [read_form(r, hits, ops) for r in recs]


def map_pairs(pairs, refs, scoring, local: bool, min_score=0, max_ties=None, max_insert=500,
              both_strands=False, threads: int | None = None, rows=None, raw=False):
    """Aligner::map_pairs_scored (lib/mapper/pair.rs:157-387) on codes:
    pairs = [(mate 1, mate 2), ...]; one pair_form per pair (raw: the
    record arrays). rows = (fwd, rev) scores of the interleaved mates."""
    threads = threads or min(8, os.cpu_count() or 1)
    L = lib()
    mates = [_codes(x) for p in pairs for x in p]
    if rows is None:
        fwd = score_matrix(mates, refs, scoring, local, threads)
        rev = score_matrix([revcomp(x) for x in mates], refs, scoring, local, threads)
    else:
        fwd, rev = rows
    fwd = np.ascontiguousarray(fwd, np.int32).reshape(-1)
    rev = np.ascontiguousarray(rev, np.int32).reshape(-1)
    if not len(fwd):
        fwd = rev = np.zeros(1, np.int32)
    qc, qo, ql = concat(mates)
    rc, ro, rl = concat(refs)
    mt = -1 if max_ties is None else int(max_ties)
    m = L.eo_map_pairs(qc, qo, ql, len(pairs), fwd, rev, rc, ro, rl, len(refs), *scoring, int(local),
                       min_score, mt, int(max_insert), int(both_strands), threads)
    try:
        recs, hits, ops = _mapped_arrays(L, m, True)
    finally:
        L.eo_mapped_free(m)
    if raw:
        return recs, hits, ops
    return [pair_form(p, hits, ops) for p in recs]


def _runs(ops):
    """[(op char, len), ...] -> uint32 runs (op << 30 | len)."""
    return np.array([(OPS.index(o) << 30) | n for o, n in ops] or [0], np.uint32), len(ops)


def md_nm(read: bytes, ref: bytes, query_start: int, ref_start: int, ops):
    """sam::md_nm (lib/sam.rs:76-124) on letters: (MD, NM)."""
    r, k = _runs(ops)
    cap = 16 + sum(12 * n + 12 for _, n in ops)
    buf = C.create_string_buffer(cap)
    nm = C.c_uint32()
    w = lib().eo_md_nm(read, len(read), ref, len(ref), query_start, ref_start, r.ctypes.data, k, buf, cap,
                       C.byref(nm))
    assert w >= 0
    return buf.value.decode(), nm.value


def cigar_string(ops, query_len: int, query_start: int, query_end: int) -> str:
    """sam::cigar_string (lib/sam.rs:32-52)."""
    r, k = _runs(ops)
    cap = 24 * (k + 2)
    buf = C.create_string_buffer(cap)
    assert lib().eo_cigar_string(r.ctypes.data, k, query_len, query_start, query_end, buf, cap) >= 0
    return buf.value.decode()


def adapter_find(read: bytes, adapter: bytes, min_overlap: int) -> int:
    return int(lib().eo_adapter_find(read, len(read), adapter, len(adapter), min_overlap))


def adapter_detect_pair(r1: bytes, r2: bytes, a1: bytes, a2: bytes, min_overlap=5):
    out = (C.c_size_t * 2)()
    lib().eo_adapter_detect_pair(r1, len(r1), r2, len(r2), a1, len(a1), a2, len(a2),
                                 min_overlap, C.byref(out))
    return [int(out[0]), int(out[1])]


def simd_supported() -> int:
    return lib().eo_simd_supported()


class Profile:
    """The SIMD restatement's panel layout (lib/simd/mod.rs:286-331)."""

    def __init__(self, refs, scoring, lanes: int):
        rc, ro, rl = concat(refs)
        self._keep = (rc, ro, rl)
        self.n_refs = len(refs)
        self.p = lib().eo_profile_build(rc, ro, rl, len(refs), *scoring, lanes)

    def score_all(self, q, local: bool, transposed_min=256, kernel=-1) -> np.ndarray:
        q = _codes(q)
        out = np.zeros(self.n_refs, np.int32)
        buf = q if len(q) else np.zeros(1, np.uint8)
        lib().eo_simd_score_all(self.p, buf.ctypes.data, len(q), int(local), transposed_min,
                                kernel, out)
        return out

    def score_batch(self, qcodes, qoff, qlen, local: bool, threads: int) -> np.ndarray:
        out = np.zeros(len(qlen) * self.n_refs, np.int32)
        lib().eo_simd_score_batch(self.p, qcodes, qoff, qlen, len(qlen), int(local), threads, out)
        return out.reshape(len(qlen), self.n_refs)

    def __del__(self):
        try:
            lib().eo_profile_free(self.p)
        except Exception:
            pass
