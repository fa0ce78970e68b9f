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
