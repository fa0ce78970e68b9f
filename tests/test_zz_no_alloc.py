"""The allocation-free entry point, esch_score_device_noalloc, through the C
ABI on an MI355X: reads and outputs in device memory (allocated here by
the test with hipMalloc), every other buffer in the library's fixed device
array. Its scores and scored flags equal esch_score_batch_i32's on the same
queries, by equality, over two calls in a row, the second with a smaller
input (the array keeps the first call's contents); a sample also equals the
scalar oracle. Sizes beyond the array's limits are refused on the host."""
import ctypes as C
import os

import numpy as np
import pytest

import eschalign_amd as ea
from eschalign_amd import synth
from oracle import pyoracle as O
import refgen as G

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DevArray:
    """A device copy of a numpy array (or an uninitialised buffer)."""

    def __init__(self, arr=None, nbytes=0):
        self.nbytes = max(int(arr.nbytes if arr is not None else nbytes), 1)
        p = C.c_void_p()
        assert hip().hipMalloc(C.byref(p), self.nbytes) == 0
        self.ptr = p.value
        if arr is not None and arr.nbytes:
            arr = np.ascontiguousarray(arr)
            assert hip().hipMemcpy(self.ptr, arr.ctypes.data, arr.nbytes, 1) == 0  # host to device

    def get(self, dtype, count):
        out = np.zeros(max(count, 1), dtype)
        if count:
            assert hip().hipMemcpy(out.ctypes.data, self.ptr, out.itemsize * count, 2) == 0  # device to host
        return out[:count]

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None


def noalloc_scores(gpu, queries):
    """Scores and scored flags of esch_score_device_noalloc for `queries`."""
    codes, off, lens = O.concat(queries)
    n = len(queries)
    d_codes, d_off, d_lens = DevArray(codes), DevArray(off), DevArray(lens)
    d_scores = DevArray(nbytes=max(n, 1) * gpu.n_refs * 4)
    d_scored = DevArray(nbytes=max(n, 1))
    try:
        gpu.score_device_noalloc(d_codes.ptr, d_off.ptr, d_lens.ptr, n, d_scores.ptr, d_scored.ptr)
        assert hip().hipDeviceSynchronize() == 0
        s = d_scores.get(np.int32, n * gpu.n_refs).reshape(n, gpu.n_refs)
        ok = d_scored.get(np.uint8, n).astype(bool)
    finally:
        for d in (d_codes, d_off, d_lens, d_scores, d_scored):
            d.free()
    return s, ok


def check_equal(gpu, queries, label):
    got, ok = noalloc_scores(gpu, queries)
    want, want_ok = gpu.score_batch_i32(queries)
    assert np.array_equal(ok, want_ok), label
    bad = np.nonzero((got != want).any(axis=1))[0]
    assert bad.size == 0, f"{label}: {bad.size} rows differ, first {bad[:5]}"
    return got


def test_noalloc_equals_score_batch_i32_two_calls():
    """The benchmark's panel and reads (lognormal lengths) plus empty
    queries and a tail of 1.5-3 kb reads (the anti-diagonal kernel's
    routing), semiglobal and local; call 1 with 6,000 queries, call 2 with
    700 other ones right after."""
    panel = synth.panel(seed=0x0123456789ABCDEF, n_refs=164, lo=140, hi=160)
    big = synth.reads(panel, seed=0x5EED_7001, n=6000)
    rng = G.Rng(0x5EED_7002)
    big[100] = np.zeros(0, np.uint8)
    big[4000] = np.zeros(0, np.uint8)
    for k in range(5000, 5012):
        big[k] = G.encode(G.random_seq(rng, 1500 + rng.below(1500), 0.01))
    small = synth.reads(panel, seed=0x5EED_7003, n=700, length=("uniform", 20, 400))
    for scoring, mode in (((1, -1, -2, -1), ea.Mode.SEMIGLOBAL), ((2, -1, -10, -1), ea.Mode.LOCAL)):
        gpu = ea.GpuScorer(ea.Panel(panel), ea.Scoring(*scoring), mode)
        got = check_equal(gpu, big, f"{ea.Mode.name(mode)} call 1")
        assert not got[100].any() and not got[4000].any()
        got2 = check_equal(gpu, small, f"{ea.Mode.name(mode)} call 2 (smaller)")
        local = mode == ea.Mode.LOCAL
        assert np.array_equal(got2[:100], O.score_matrix(small[:100], panel, scoring, local, THREADS))


def test_noalloc_with_a_restarted_column_shift():
    """References up to 1,900 nt under 1,-1,-2,-1: the default variants
    restart their column shift (local every 128 columns, semiglobal every
    64); the allocation-free entry point runs them with the scorer's block
    and equals esch_score_batch_i32 and, on a sample, the scalar oracle, over
    two calls, the second smaller."""
    rng = G.Rng(0x5EED_7006)
    panel_l = [G.random_seq(rng, n, 0.01) for n in (1900, 1700, 1500)] + G.random_panel(rng, 60, 40, 400)
    reads = [panel_l[0][200:1400], G.random_seq(rng, 1700, 0.02)]
    reads += [G.random_seq(rng, 20 + rng.below(600), 0.02) for _ in range(40)]
    q = [G.encode(r) for r in reads]
    for mode, block in ((ea.Mode.LOCAL, 128), (ea.Mode.SEMIGLOBAL, 64)):
        gpu = ea.GpuScorer(ea.Panel(panel_l), ea.Scoring(1, -1, -2, -1), mode)
        assert gpu.column_block(gpu.variant) == block, (mode, gpu.variant)
        got = check_equal(gpu, q, f"{ea.Mode.name(mode)} restart call 1")
        check_equal(gpu, q[:7], f"{ea.Mode.name(mode)} restart call 2")
        want = O.score_matrix(q[:6], ea.Panel(panel_l).codes, (1, -1, -2, -1), mode == ea.Mode.LOCAL, THREADS)
        assert np.array_equal(got[:6], want), mode


def test_noalloc_int32_rescue_and_long_references():
    """Every query outside the int16 bound (1000,-1000,-1000,-1000) takes
    the int32 kernel; a panel of 5,000-nt references (the 16-lane
    anti-diagonal kernel's global boundary, the cut grids) with reads up to
    6 kb; two calls each, the second smaller."""
    rng = G.Rng(0x5EED_7004)
    panel_l = G.random_panel(rng, 9, 200, 200)
    reads = [G.random_seq(rng, 20 + rng.below(400), 0.02) for _ in range(40)] + [panel_l[0][:150]]
    gpu = ea.GpuScorer(ea.Panel(panel_l), ea.Scoring(1000, -1000, -1000, -1000), ea.Mode.LOCAL)
    q = [G.encode(r) for r in reads]
    check_equal(gpu, q, "int32 rescue call 1")
    check_equal(gpu, q[:9], "int32 rescue call 2")
    long_panel = [G.random_seq(rng, 4000 + rng.below(1000), 0.0) for _ in range(6)]
    reads = [long_panel[k % 6][rng.below(2000):][:500 + rng.below(5500)] for k in range(30)]
    reads += [G.random_seq(rng, 100 + rng.below(3000), 0.01) for _ in range(30)]
    q = [G.encode(r) for r in reads]
    for mode in (ea.Mode.SEMIGLOBAL, ea.Mode.LOCAL):
        gpu = ea.GpuScorer(ea.Panel(long_panel), ea.Scoring(2, -1, -10, -1), mode)
        check_equal(gpu, q, f"long references {ea.Mode.name(mode)} call 1")
        check_equal(gpu, q[:11], f"long references {ea.Mode.name(mode)} call 2")


def test_noalloc_refuses_sizes_beyond_its_array():
    """More queries than ESCH_NOALLOC_MAX_ROWS are refused with CAPACITY on
    the host, before any launch; n = 0 only sets the device up."""
    panel = synth.panel(seed=0x0123456789ABCDEF, n_refs=16, lo=140, hi=160)
    gpu = ea.GpuScorer(ea.Panel(panel), ea.Scoring(), ea.Mode.LOCAL)
    gpu.score_device_noalloc(0, 0, 0, 0, 0, 0)
    with pytest.raises(ea.GpuError) as e:
        gpu.score_device_noalloc(8, 8, 8, ea.NOALLOC_MAX_ROWS + 1, 8, 8)
    assert e.value.kind == "Capacity"
