"""Seeded synthetic workloads (panels, reads, read pairs) for the benchmark
and the tests, generated by libesch_synth.so (csrc/synth.cpp): xorshift64*
and the error model 5% substitutions, 3% insertions, 3% deletions with
Geometric(0.7) gap lengths unless stated otherwise."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "libesch_synth.so")
_lib = None

TRUSEQ_R1 = b"AGATCGGAAGAGCACACGTCTGAACTCCAGTCA"  # lib/adapter.rs:40
TRUSEQ_R2 = b"AGATCGGAAGAGCGTCGTGTAGGGAAAGAGTGT"  # lib/adapter.rs:42


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            raise ImportError(f"{_PATH} is missing: run __graft_entry__.build()")
        L = C.CDLL(_PATH)
        vp = C.c_void_p
        L.esch_synth_panel.argtypes = [C.c_uint64, C.c_int32, C.c_int32, C.c_int32, vp, C.c_uint64,
                                       vp, vp]
        L.esch_synth_panel.restype = C.c_int64
        L.esch_synth_reads.argtypes = [vp, vp, vp, C.c_int32, C.c_uint64, C.c_int64, C.c_int32,
                                       C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_double,
                                       C.c_double, C.c_double, C.c_double, C.c_double, vp,
                                       C.c_uint64, vp, vp]
        L.esch_synth_reads.restype = C.c_int64
        L.esch_synth_substitute.argtypes = [C.c_uint64, vp, vp, C.c_uint64, C.c_double]
        L.esch_synth_substitute.restype = None
        L.esch_synth_pairs.argtypes = [vp, vp, vp, C.c_int32, C.c_uint64, C.c_int64, C.c_int32,
                                       C.c_double, C.c_double, vp, C.c_int32, vp, C.c_int32,
                                       C.c_double, C.c_double, C.c_double, vp, C.c_uint64, vp, vp]
        L.esch_synth_pairs.restype = C.c_int64
        _lib = L
    return _lib


def _concat(seqs):
    lens = np.array([len(s) for s in seqs], np.int32)
    off = np.zeros(len(seqs), np.uint64)
    if len(seqs):
        off[1:] = np.cumsum(lens[:-1], dtype=np.uint64)
    codes = np.concatenate(seqs).astype(np.uint8) if len(seqs) else np.zeros(1, np.uint8)
    return np.ascontiguousarray(codes), off, lens


def _split(codes, off, lens):
    return [codes[int(o):int(o) + int(n)] for o, n in zip(off, lens)]


def panel_arrays(seed, n_refs, lo, hi):
    cap = n_refs * hi
    codes = np.zeros(max(cap, 1), np.uint8)
    off = np.zeros(n_refs, np.uint64)
    lens = np.zeros(n_refs, np.int32)
    tot = lib().esch_synth_panel(seed, n_refs, lo, hi, codes.ctypes.data, cap, off.ctypes.data,
                                 lens.ctypes.data)
    assert tot >= 0
    return codes[:tot], off, lens


def panel(seed, n_refs, lo, hi):
    """n_refs uniform-ACGT references, lengths uniform in [lo, hi]."""
    return _split(*panel_arrays(seed, n_refs, lo, hi))


def panel_with_copies(seed, n_seeds, copies, lo, hi, p_sub=0.02):
    """n_seeds random references, each followed by `copies` copies carrying
    independent substitutions at p_sub (the large-panel workload)."""
    seeds = panel(seed, n_seeds, lo, hi)
    out = []
    for k, s in enumerate(seeds):
        out.append(s)
        for c in range(copies):
            d = np.empty_like(s)
            lib().esch_synth_substitute(seed ^ (0x9E37 + k * 131 + c), s.ctypes.data, d.ctypes.data,
                                        s.size, p_sub)
            out.append(d)
    return out


def read_arrays(panel_codes, seed, n, length=("lognormal", 90, 0.6, 20, 1000), frac_ref=0.9,
                p_sub=0.05, p_ins=0.03, p_del=0.03, geo=0.7):
    """Concatenated reads (codes, offsets, lengths); see reads()."""
    pc, po, pl = _concat(panel_codes)
    kind, a, b, lo, hi = length if length[0] == "lognormal" else ("uniform", 0, 0, length[1],
                                                                   length[2])
    k = 1 if kind == "lognormal" else 0
    # generous capacity: insertions can lengthen a read
    cap = int(n * (hi * 1.6 + 16)) if k == 0 else int(n * (max(a * np.exp(2 * b), 1) * 1.6 + hi * 0.01 + 64))
    cap = max(cap, 1024)
    while True:
        codes = np.zeros(cap, np.uint8)
        off = np.zeros(max(n, 1), np.uint64)
        lens = np.zeros(max(n, 1), np.int32)
        tot = lib().esch_synth_reads(pc.ctypes.data, po.ctypes.data, pl.ctypes.data,
                                     len(panel_codes), seed, n, k, a, b, lo, hi, frac_ref, p_sub,
                                     p_ins, p_del, geo, codes.ctypes.data, cap, off.ctypes.data,
                                     lens.ctypes.data)
        if tot >= 0:
            return codes[:max(tot, 1)], off[:n], lens[:n]
        cap *= 2


def reads(panel_codes, seed, n, length=("lognormal", 90, 0.6, 20, 1000), **kw):
    """n reads: 90% from a random reference on a random strand through the
    error model, 10% uniform random. length = ("lognormal", median, sigma,
    lo, hi) or ("uniform", lo, hi)."""
    return _split(*read_arrays(panel_codes, seed, n, length, **kw))


def pair_arrays(panel_codes, seed, n_pairs, mate_len=150, mu=120.0, sd=30.0, p_sub=0.005,
                p_indel=0.0005, geo=0.7):
    pc, po, pl = _concat(panel_codes)
    a1 = np.frombuffer(TRUSEQ_R1, np.uint8).copy()
    a2 = np.frombuffer(TRUSEQ_R2, np.uint8).copy()
    t = np.full(256, 4, np.uint8)
    t[ord("A")], t[ord("C")], t[ord("G")], t[ord("T")] = 0, 1, 2, 3
    a1, a2 = t[a1], t[a2]
    cap = int(2 * n_pairs * (mate_len + 8))
    codes = np.zeros(cap, np.uint8)
    off = np.zeros(2 * n_pairs, np.uint64)
    lens = np.zeros(2 * n_pairs, np.int32)
    tot = lib().esch_synth_pairs(pc.ctypes.data, po.ctypes.data, pl.ctypes.data, len(panel_codes),
                                 seed, n_pairs, mate_len, mu, sd, a1.ctypes.data, len(a1),
                                 a2.ctypes.data, len(a2), p_sub, p_indel, geo, codes.ctypes.data,
                                 cap, off.ctypes.data, lens.ctypes.data)
    assert tot >= 0
    return codes[:tot], off, lens
