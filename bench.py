#!/usr/bin/env python3
"""Benchmark of the panel-scoring hot path on one MI355X.

Workload (the batched uneven-length configuration): the 164-reference panel
(uniform ACGT, lengths 140-160), 1M reads with lognormal lengths (median 90,
sigma 0.6, clipped 20-1000), 90% drawn from the panel on a random strand
through 5% substitution / 3% insertion / 3% deletion errors, 10% random;
both strands, semiglobal mode, scoring 1,-1,-2,-1. Synthetic data.

A step is one pass of the device pipeline over the batch, already resident
in HBM: length bucketing, the batched packed 16-bit DP (in fp16 on panels
whose value bound fits its exact integers, as this one does, with every
value shifted by row and column so that a cell takes 5 packed ops; int16
and int32 rescue where needed) and the fused best / tie selection. value =
reads per second.

Also reported: the DP kernel's time from HIP events on its stream, GCUPS and
the fraction of the packed 16-bit VALU lane-op roofline at the kernel's own
minimum op count per cell; the plain one-thread-per-
pair int32 kernel on the same batch; the reference's CPU kernel (the
oracle's AVX-512/AVX2 restatement) on a bounded sample on the host cores;
the winner tracebacks of the step's selection; map_reads_scored whole on
the same reads (scores to records); the paired workload (2M
pairs of 150-nt mates on a 500-reference panel) mapped end to end with the
pair tier loop on the device, against the oracle on a sample; and serving.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

# MI355X: 256 CUs at 2.4 GHz; one packed int16 VALU op does 2 lane-ops in
# each of 64 lanes, 4 SIMDs x 16 lanes per clock per CU -> 128 lane-ops/clk/CU.
N_CU, CLOCK_HZ = 256, 2.4e9
PEAK_I16 = N_CU * CLOCK_HZ * 128           # lane-ops/s
PEAK_I32 = N_CU * CLOCK_HZ * 64
# Minimum packed 16-bit lane-ops per DP cell of the recurrence: 4 adds and
# 4 two-input maxima in semiglobal mode, plus the zero floor and the running
# best in local mode; with the fp16 variants' three-input maximum H takes
# one max instead of two, and the column best a three-way tree. A shift
# along one axis (shifted 1: rows in semiglobal mode, columns in local
# mode) drops one gap chain's extension add, and in local mode lets the
# zero floor ride on the horizontal chain's maximum (free in fp16, whose
# three-input maximum is one op); the diagonal shift (shifted 2,
# semiglobal mode only) drops both adds. Keyed (mode, fp16, shifted).
OPS_PER_CELL = {("semiglobal", False, 0): 8, ("local", False, 0): 10,
                ("semiglobal", True, 0): 7, ("local", True, 0): 8.5,
                ("semiglobal", False, 1): 7, ("local", False, 1): 9,
                ("semiglobal", True, 1): 6, ("local", True, 1): 6.5,
                ("semiglobal", False, 2): 6, ("semiglobal", True, 2): 5}


def vinfo(v):
    """A k_dp16 variant's schedule, rows per stripe, fp16 arithmetic and
    shifted column, as the library reports them."""
    import eschalign_amd as ea
    return ea.variant_info(v)


def int16_counterpart(v):
    """The int16 variant with v's schedule family and stripe height: the
    diagonal-shift int16 schedule 14 for a diagonal-shift fp16 one, the
    windowed int16 schedule 11 for a windowed fp16 one, the row-shifted int16
    schedule 6 for the other shifted ones, 3 for the unshifted."""
    import eschalign_amd as ea
    info = vinfo(v)
    if not info.shifted:
        return 3
    sched = 14 if info.diag else 11 if info.schedule >= 9 else 6
    return next(u for u, x in ea.variants().items() if x.schedule == sched and x.stripe_rows == info.stripe_rows)


# SQ_INSTS_VALU x 128 / cells of one k_dp16 launch of the default workload,
# from the counter summary in profiles/ (mode, variant) -> lane-ops per cell
VALU_OPS_PER_CELL_ISSUED = {("semiglobal", 3): round(344_021_737_036 * 128 / 5_303_272_453_076, 3),
                            ("semiglobal", 4): round(304_396_230_832 * 128 / 5_303_272_453_076, 3),
                            ("semiglobal", 5): round(301_256_340_206 * 128 / 5_303_272_453_076, 3),
                            ("semiglobal", 6): round(306_294_990_254 * 128 / 5_303_272_453_076, 3),
                            ("semiglobal", 7): round(263_781_725_920 * 128 / 5_303_272_453_076, 3),
                            ("semiglobal", 10): round(262_088_845_368 * 128 / 5_303_272_453_076, 3),
                            ("semiglobal", 11): round(303_721_597_544 * 128 / 5_303_272_453_076, 3),
                            ("semiglobal", 24): round(220_033_994_246 * 128 / 5_303_272_453_076, 3),
                            ("semiglobal", 26): round(219_282_137_274 * 128 / 5_303_272_453_076, 3),
                            ("semiglobal", 28): round(261_701_617_397 * 128 / 5_303_272_453_076, 3)}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(panel_codes, codes, off, lens, scoring, local, gpu_check):
    """The oracle's SIMD restatement of the reference CPU kernel, both
    strands, threads = host cores (at most 16), on a bounded sample sized
    to about 10-20 s of CPU work; its scores are checked against the GPU's
    on the same sample."""
    from oracle import pyoracle as O
    import eschalign_amd as ea

    sup = O.simd_supported()
    lanes = 32 if sup & 2 else (16 if sup & 1 else 0)
    threads = min(16, os.cpu_count() or 1)
    if not lanes:
        return None
    prof = O.Profile(panel_codes, scoring, lanes)

    def run(n):
        qs = [codes[int(off[k]):int(off[k]) + int(lens[k])] for k in range(n)]
        qs += [ea.reverse_complement_codes(q) for q in qs]
        qc, qo, ql = O.concat(qs)
        t = time.perf_counter()
        out = prof.score_batch(qc, qo, ql, local, threads)
        return time.perf_counter() - t, qs, out

    dt, _, _ = run(2000)
    n = int(min(len(lens), max(2000, 2000 * 12.0 / max(dt, 1e-3))))
    dt, qs, out = run(n)
    cells = sum(len(q) for q in qs) * sum(len(p) for p in panel_codes)
    gpu_check(qs, out)
    avx2 = None
    if lanes == 32 and sup & 1:
        # the AVX2 16-lane kernel on a quarter of the sample, reported apart
        prof = O.Profile(panel_codes, scoring, 16)
        n2 = max(2000, n // 4)
        dt2, qs2, out2 = run(n2)
        gpu_check(qs2, out2)
        avx2 = {"reads_per_s": round(n2 / dt2, 1), "reads": n2,
                "cells_per_s": round(sum(len(q) for q in qs2) * sum(len(p) for p in panel_codes) / dt2, 1)}
    return {
        "value": round(n / dt, 1),
        "unit": "reads/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {n} reads of the workload, both strands ({2 * n} queries x "
                  f"{len(panel_codes)} refs), {'AVX-512BW 32' if lanes == 32 else 'AVX2 16'}-lane "
                  f"inter-sequence i16 kernel restated from the reference, {threads} threads",
        "cells_per_s": round(cells / dt, 1),
        "seconds": round(dt, 2),
        "host": host_cpu(),
        "avx2": avx2,
    }


def host_cpu():
    """Model, logical CPUs and SMT of the host, from /proc/cpuinfo."""
    model, siblings, cores = None, None, None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "siblings" and siblings is None:
                    siblings = int(v)
                elif k == "cpu cores" and cores is None:
                    cores = int(v)
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(),
            "smt": None if not (siblings and cores) else siblings > cores}


def trace_leg(gpu, batch, panel_codes, codes, off, lens, scoring, local, n_reads, cpu):
    """The winner tracebacks of the first n_reads reads: every selected
    (read or reverse complement, reference) of the step's fused selection,
    traced on the device (esch_align_pairs_concat); on a bounded sample,
    the oracle's restatement of the reference's AVX-512BW pair traceback
    (simd::align_pairs) on the host cores, checked equal (score,
    coordinates, runs)."""
    recs, ties = batch.select_raw()
    nr = min(n_reads, len(recs))
    take = np.where((recs[:nr, 3] & 1) != 0, np.minimum(recs[:nr, 4], ties.shape[1]), 0)
    mask = np.arange(ties.shape[1])[None, :] < take[:, None]
    k_idx = np.nonzero(mask)[0]          # read order, then tie order
    ent = ties[:nr][mask].astype(np.int64)
    refs, rev = (ent >> 1).astype(np.int32), (ent & 1) != 0
    # reverse complements as slices of the whole code array reverse-complemented
    total = len(codes)
    big = np.concatenate([codes, np.where(codes < 4, 3 - codes, 4).astype(np.uint8)[::-1]])
    o, ln = off[k_idx].astype(np.int64), lens[k_idx].astype(np.int64)
    qoff = np.where(rev, total + (total - o - ln), o).astype(np.uint64)
    qlen = ln.astype(np.int32)
    rlen = np.array([len(p) for p in panel_codes], np.int64)
    cells = int((ln * rlen[refs]).sum())
    gpu.align_pairs_concat(big, qoff, qlen, refs)  # warm: code objects, the call's buffers (steady state)
    t = time.perf_counter()
    al, _ = gpu.align_pairs_concat(big, qoff, qlen, refs)
    py_wall = time.perf_counter() - t
    kms = gpu.align_last_ms()
    # the C ABI call alone, host buffers in and the caller's records and runs
    # out, those allocated once as a caller reusing them would (the Python
    # mirror's call above adds its allocations and the conversion of the
    # records to an int64 table)
    import ctypes as C
    import eschalign_amd as ea
    lib = ea.load_library()
    n = len(qlen)
    cap = int(qlen.astype(np.int64).sum()) + int(rlen[refs].sum()) + 1
    recs_buf = np.empty((n, 4), np.int64)
    ops_buf = np.empty(cap, np.uint32)
    used = C.c_size_t()
    qoff_c, qlen_c, refs_c = (np.ascontiguousarray(qoff, np.uint64), np.ascontiguousarray(qlen, np.int32),
                              np.ascontiguousarray(refs, np.int32))

    def abi_call():
        st = lib.esch_align_pairs_concat(gpu._h, big.ctypes.data, qoff_c.ctypes.data, qlen_c.ctypes.data,
                                         refs_c.ctypes.data, n, recs_buf.ctypes.data, ops_buf.ctypes.data, cap,
                                         C.byref(used))
        assert st == 0, "esch_align_pairs_concat failed"
    abi_call()  # warm: the outputs' pages
    walls = []
    for _ in range(3):
        t = time.perf_counter()
        abi_call()
        walls.append(time.perf_counter() - t)
    wall = float(np.median(walls))
    assert np.array_equal(recs_buf.view(np.int32).reshape(-1, 8)[:n, :6], al[:, :6])
    out = {"reads": int(nr), "pairs": int(len(refs)), "cells": cells,
           "pairs_per_s": round(len(refs) / wall, 1), "wall_ms": round(wall * 1e3, 2),
           "scope": "one esch_align_pairs_concat call from host buffers to the caller's records and "
                    "runs (median of 3); python_* is the Python mirror's call with its own allocations",
           "python_wall_ms": round(py_wall * 1e3, 2), "python_pairs_per_s": round(len(refs) / py_wall, 1),
           "kernel_ms": round(kms, 3), "kernel_gcups": round(cells / (kms / 1e3) / 1e9, 1),
           "wall_gcups": round(cells / wall / 1e9, 1)}
    if cpu:
        from oracle import pyoracle as O
        threads = min(16, os.cpu_count() or 1)
        m = min(len(refs), 2000)
        t = time.perf_counter()
        O.simd_align_batch_concat(big, qoff[:m], qlen[:m], panel_codes, refs[:m], scoring, local, threads)
        dt = time.perf_counter() - t
        m = int(min(len(refs), max(m, m * 8.0 / max(dt, 1e-3))))  # about 8 s of CPU work
        t = time.perf_counter()
        ref = O.simd_align_batch_concat(big, qoff[:m], qlen[:m], panel_codes, refs[:m], scoring, local, threads)
        dt = time.perf_counter() - t
        kernel = "AVX-512BW pair kernel" if O.simd_supported() & 2 else "scalar (no AVX-512BW on this host)"
        assert np.array_equal(ref, al[:m, :6].astype(np.int32)), "device traceback != CPU traceback"
        cpu_cells = int((ln[:m] * rlen[refs[:m]]).sum())
        out["cpu_baseline"] = {"pairs_per_s": round(m / dt, 1), "cells_per_s": round(cpu_cells / dt, 1),
                               "threads": threads, "kind": "port",
                               "sample": f"first {m} traced pairs, simd::align_pairs with the "
                                         f"{kernel} restated from lib/simd/mod.rs:405-499 and "
                                         f"lib/simd/avx512.rs:120-247 (chunks of 64 pairs, 32 lanes, "
                                         f"the scalar walk), equal on every pair"}
    return out


def pairs_leg(n_pairs, cpu, both_strands=False):
    """The paired workload: 500 references of 70-95 nt; n_pairs pairs of
    150-nt mates from inserts N(120, 30) that read into the TruSeq adapters
    (0.5% substitutions, 0.05% indels); each mate less its adapter
    (esch_adapter_trim_pairs); semi-global 2,-1,-10,-1, --max-insert 500.
    One esch_map_pairs_concat call end to end from host buffers: staging and
    upload, the orientations the tiers read scored (forward template strand:
    mate 1 as sequenced and mate 2 reverse-complemented; both strands: all
    four), the tier loop, the tracebacks, the per-mate suboptimal scores, the
    Separate fallback (its mates' other orientations scored then) and the
    records back. On a bounded sample, the oracle on the host cores (panel
    rows by the SIMD restatement of the reference's CPU kernel, then the
    restated map_pairs_scored with its scalar tracebacks), equal on every
    pair."""
    import eschalign_amd as ea
    from eschalign_amd import synth
    from oracle import pyoracle as O

    scoring = (2, -1, -10, -1)
    t0 = time.perf_counter()
    panel = synth.panel(seed=0xC0F4, n_refs=500, lo=70, hi=95)
    codes, off, lens = synth.pair_arrays(panel, seed=0xC0F5, n_pairs=n_pairs)
    tl = ea.adapter_trim_pairs(codes, off, lens)
    log(f"[bench] paired workload: {n_pairs} pairs, mates {lens.mean():.1f} nt, "
        f"{tl.mean():.1f} nt after adapter trimming ({time.perf_counter() - t0:.1f} s to generate)")
    gpu = ea.GpuScorer(ea.Panel(panel), ea.Scoring(*scoring), ea.Mode.SEMIGLOBAL)
    opts = dict(both_strands=both_strands, min_score=0, max_ties=None, max_insert=500)
    gpu.map_pairs_concat(codes, off, tl, **opts)  # warm: code objects, the call's buffers (steady state)
    t = time.perf_counter()
    recs, hits, ops, st = gpu.map_pairs_concat(codes, off, tl, **opts)
    wall = time.perf_counter() - t
    cells = int(st["cells"])  # what was scored: two orientations per pair forward-only, four with both
    proper = int(recs["proper"].sum())
    out = {"pairs": n_pairs, "pairs_per_s": round(n_pairs / wall, 1), "wall_ms": round(wall * 1e3, 1),
           "lib_ms": round(st["total_ms"], 1), "load_ms": round(st["load_ms"], 1),
           "score_ms": round(st["score_ms"], 1), "select_ms": round(st["select_ms"], 1),
           "trace_ms": round(st["trace_ms"], 1), "trace_wall_ms": round(st["trace_wall_ms"], 1),
           "assemble_ms": round(st["assemble_ms"], 1), "rounds": int(st["rounds"]),
           "traced": int(st["traced"]), "proper": proper, "separate": n_pairs - proper, "cells": cells,
           "score_gcups": round(cells / (st["score_ms"] / 1e3) / 1e9, 1),
           "workload": "500 refs 70-95 nt, 150-nt mates, inserts N(120, 30) into TruSeq adapters, "
                       f"trimmed, semiglobal 2,-1,-10,-1, max_insert 500, "
                       f"{'both template strands' if both_strands else 'forward template strand'}"}
    if cpu:
        sup = O.simd_supported()
        lanes = 32 if sup & 2 else 16
        threads = min(16, os.cpu_count() or 1)
        prof = O.Profile(panel, scoring, lanes)

        def run(m):
            # rows on demand, as the reference's CPU path scores them
            # (lib/mapper/pair.rs:195-207): forward template strand only, mate
            # 1 as sequenced and mate 2 reverse-complemented; both orientations
            # of a mate only when its pair falls back to Separate
            mates = [codes[int(off[k]):int(off[k]) + int(tl[k])] for k in range(2 * m)]
            pairs = [(mates[2 * k], mates[2 * k + 1]) for k in range(m)]
            t = time.perf_counter()
            nr = len(panel)
            fwd = np.zeros((2 * m, nr), np.int32)
            rev = np.zeros((2 * m, nr), np.int32)
            if both_strands:
                fwd[:] = prof.score_batch(*O.concat(mates), False, threads)
                rev[:] = prof.score_batch(*O.concat([O.revcomp(x) for x in mates]), False, threads)
            else:
                fwd[0::2] = prof.score_batch(*O.concat(mates[0::2]), False, threads)
                rev[1::2] = prof.score_batch(*O.concat([O.revcomp(x) for x in mates[1::2]]), False, threads)
            res = O.map_pairs(pairs, panel, scoring, False, rows=(fwd, rev), threads=threads, **opts)
            sep = [k for k in range(m) if res[k][0] == "separate"]
            if sep and not both_strands:
                sf = prof.score_batch(*O.concat([mates[2 * k + 1] for k in sep]), False, threads)
                sr = prof.score_batch(*O.concat([O.revcomp(mates[2 * k]) for k in sep]), False, threads)
                f2 = np.stack([r for q, k in enumerate(sep) for r in (fwd[2 * k], sf[q])])
                r2 = np.stack([r for q, k in enumerate(sep) for r in (sr[q], rev[2 * k + 1])])
                again = O.map_pairs([pairs[k] for k in sep], panel, scoring, False, rows=(f2, r2),
                                    threads=threads, **opts)
                for k, r in zip(sep, again):
                    res[k] = r
            return time.perf_counter() - t, res

        dt, _ = run(1000)
        m = int(min(n_pairs, max(1000, 1000 * 10.0 / max(dt, 1e-3))))
        dt, res = run(m)
        for k in range(m):
            assert O.pair_form(recs[k], hits, ops) == res[k], f"device pair mapping != CPU on pair {k}"
        out["cpu_baseline"] = {"pairs_per_s": round(m / dt, 1), "threads": threads, "kind": "port",
                               "sample": f"first {m} pairs: the orientations the tiers need scored by the "
                                         f"{'AVX-512BW' if lanes == 32 else 'AVX2'} restatement of the "
                                         f"reference CPU kernel (the others only for Separate pairs), then "
                                         f"map_pairs_scored restated with scalar tracebacks "
                                         f"(lib/mapper/pair.rs:157-387), equal on every pair"}
    return out


def map_reads_leg(gpu, panel_codes, codes, off, lens, scoring, local, n_reads, cpu):
    """Aligner::map_reads_scored whole on the device (esch_map_reads): the
    first n_reads reads of the workload from host buffers, both strands,
    scored, selected, every taken tie traced, records out. On a bounded
    sample, the oracle on the host cores (rows by the SIMD restatement of
    the reference CPU kernel, then the restated map_reads_scored with scalar
    tracebacks), equal on every read."""
    from oracle import pyoracle as O

    gpu.map_reads_concat(codes, off[:n_reads], lens[:n_reads], both_strands=True)  # warm: the call's buffers
    calls = []
    for _ in range(3):
        t = time.perf_counter()
        res = gpu.map_reads_concat(codes, off[:n_reads], lens[:n_reads], both_strands=True)
        calls.append((time.perf_counter() - t, res))
        del res
    calls.sort(key=lambda c: c[0])
    wall, (recs, hits, ops, st) = calls[1]  # the median of three calls
    del calls
    out = {"reads": n_reads, "calls": "median of 3", "reads_per_s": round(n_reads / wall, 1), "wall_ms": round(wall * 1e3, 1),
           "lib_ms": round(st["total_ms"], 1), "load_ms": round(st["load_ms"], 1),
           "score_ms": round(st["score_ms"], 1), "select_ms": round(st["select_ms"], 1),
           "trace_ms": round(st["trace_ms"], 1), "trace_wall_ms": round(st["trace_wall_ms"], 1),
           "assemble_ms": round(st["assemble_ms"], 1), "traced": int(st["traced"]),
           "mapped": int((recs["n_hits"] > 0).sum())}
    if cpu:
        sup = O.simd_supported()
        lanes = 32 if sup & 2 else 16
        threads = min(16, os.cpu_count() or 1)
        prof = O.Profile(panel_codes, scoring, lanes)

        def run(m):
            qs = [codes[int(off[k]):int(off[k]) + int(lens[k])] for k in range(m)]
            t = time.perf_counter()
            fwd = prof.score_batch(*O.concat(qs), local, threads)
            rev = prof.score_batch(*O.concat([O.revcomp(q) for q in qs]), local, threads)
            res = O.map_reads(qs, panel_codes, scoring, local, both_strands=True, threads=threads, rows=(fwd, rev))
            return time.perf_counter() - t, res

        dt, _ = run(2000)
        m = int(min(n_reads, max(2000, 2000 * 10.0 / max(dt, 1e-3))))
        dt, res = run(m)
        for k in range(m):
            assert O.read_form(recs[k], hits, ops) == res[k], f"device read mapping != CPU on read {k}"
        out["cpu_baseline"] = {"reads_per_s": round(m / dt, 1), "threads": threads, "kind": "port",
                               "sample": f"first {m} reads: both strands scored by the "
                                         f"{'AVX-512BW' if lanes == 32 else 'AVX2'} restatement of the "
                                         f"reference CPU kernel, then map_reads_scored restated with the "
                                         f"winners of each 64 reads traced together by the AVX-512BW pair "
                                         f"kernel (lib/mapper.rs:400-550), equal on every read"}
    return out


def _timed_steps(batch, what, steps, tie_cap=1):
    """Warm once, then `steps` eager runs: (wall s per step, mean DP ms from
    the HIP events on the batch's stream)."""
    batch.run(what, tie_cap=tie_cap)
    batch.wait()
    dps = []
    t = time.perf_counter()
    for _ in range(steps):
        batch.run(what, tie_cap=tie_cap)
        batch.wait()
        dps.append(batch.timing()[0])
    return (time.perf_counter() - t) / steps, float(np.mean(dps))


def _roof(cells, dp_ms, mode, v):
    info = vinfo(v)
    ops = OPS_PER_CELL[(mode, info.f16, info.shifted)]
    achieved = cells * ops / (dp_ms / 1e3)
    return {"bound": "valu", "achieved": round(achieved / 1e12, 3), "peak": round(PEAK_I16 / 1e12, 3),
            "unit": "Tlane-op/s (packed 16-bit)", "frac": round(achieved / PEAK_I16, 4), "ops_per_cell": ops}


def config1_leg(steps, cpu):
    """The plain CPU reference path's workload, on the device and on the
    host: the 164-reference panel (140-160 nt), 10k reads (90% substrings
    of 60-150 nt through the error model, 10% random), local 2,-1,-10,-1,
    forward strand only, int32 scores. Device: the autotuned batched kernel
    and the plain one-thread-per-pair kernel, from HBM; host: the oracle's
    AVX-512 restatement of the reference's CPU kernel over all 10k reads on
    the host cores, equal to the device's scores."""
    import eschalign_amd as ea
    from eschalign_amd import synth

    panel = synth.panel(seed=0x0123456789ABCDEF, n_refs=164, lo=140, hi=160)
    n = 10_000
    codes, off, lens = synth.read_arrays(panel, seed=0xC0F1, n=n, length=("uniform", 60, 150))
    reads = [codes[int(off[k]):int(off[k]) + int(lens[k])] for k in range(n)]
    scoring = (2, -1, -10, -1)
    gpu = ea.GpuScorer(ea.Panel(panel), ea.Scoring(*scoring), ea.Mode.LOCAL)
    chosen, _ = gpu.autotune(reads, both_strands=False)
    batch = gpu.batch(n, int(lens.sum()))
    batch.load_concat(codes, off, lens, both_strands=False)
    wall, dp = _timed_steps(batch, ea.RUN_SCORES32 | ea.RUN_NO_GRAPH, steps)
    got, ok = batch.scores32()
    cells, _, _ = batch.stats()
    pwall, pdp = _timed_steps(batch, ea.RUN_SCORES32 | ea.RUN_PLAIN | ea.RUN_NO_GRAPH, 1)
    gotp, _ = batch.scores32()
    assert ok.all() and np.array_equal(gotp, got), "plain kernel != batched kernel"
    out = {"workload": "164-ref panel (140-160 nt), 10k reads 60-150 nt (90% substrings with errors, "
                       "10% random), local 2,-1,-10,-1, forward strand",
           "reads": n, "cells": int(cells), "kernel_variant": chosen,
           "reads_per_s": round(n / wall, 1), "ms_per_step": round(wall * 1e3, 3), "dp_ms": round(dp, 3),
           "gcups": round(cells / (dp / 1e3) / 1e9, 1), "roofline": _roof(cells, dp, "local", chosen),
           "plain_kernel": {"reads_per_s": round(n / pwall, 1), "dp_ms": round(pdp, 3),
                            "gcups": round(cells / (pdp / 1e3) / 1e9, 1)}}
    if cpu:
        from oracle import pyoracle as O
        sup = O.simd_supported()
        lanes = 32 if sup & 2 else 16
        threads = min(16, os.cpu_count() or 1)
        prof = O.Profile(panel, scoring, lanes)
        prof.score_batch(codes, off, lens, True, threads)  # warm
        t = time.perf_counter()
        want = prof.score_batch(codes, off, lens, True, threads)
        dt = time.perf_counter() - t
        assert np.array_equal(want, got), "device != CPU kernel on the plain CPU path's workload"
        out["cpu_baseline"] = {"reads_per_s": round(n / dt, 1), "cells_per_s": round(cells / dt, 1),
                               "cores": threads, "kind": "port",
                               "sample": f"all {n} reads, {'AVX-512BW 32' if lanes == 32 else 'AVX2 16'}-lane "
                                         f"restatement of the reference's CPU kernel, equal on every score"}
    return out


def config3_leg(n_reads, cpu, cpu_reads=2000, variant=-1):
    """The large-panel workload: 10k references of 200-1000 nt (2k seeds
    plus 4 copies of each at 2% substitutions, so ties and near-ties are
    common), n_reads reads of 100-1000 nt, local 2,-1,-10,-1, forward
    strand. The autotuner sweeps the batched kernel's variants (schedules
    and stripe heights) and the per-length-bucket routing to the
    anti-diagonal kernel on the first 500 reads; one timed pass of the whole batch with the fused selection. The
    CPU baseline is the oracle's AVX-512 restatement of the reference's CPU
    kernel on the first cpu_reads reads (a seeded subsample), equal to the
    device's rows and tie sets."""
    import eschalign_amd as ea
    from eschalign_amd import synth

    t0 = time.perf_counter()
    panel = synth.panel_with_copies(seed=0xC0F3, n_seeds=2000, copies=4, lo=200, hi=1000)
    codes, off, lens = synth.read_arrays(panel, seed=0xC0F3 + 1, n=n_reads, length=("uniform", 100, 1000))
    gen_s = time.perf_counter() - t0
    scoring = (2, -1, -10, -1)
    gpu = ea.GpuScorer(ea.Panel(panel), ea.Scoring(*scoring), ea.Mode.LOCAL)
    sample = [codes[int(off[k]):int(off[k]) + int(lens[k])] for k in range(min(n_reads, 500))]
    t0 = time.perf_counter()
    if variant >= 0:  # fixed (profiling): no autotuning dispatches
        gpu.variant = variant
        chosen, tune_ms = variant, []
    else:
        chosen, tune_ms = gpu.autotune(sample, both_strands=False)
    tune_s = time.perf_counter() - t0
    batch = gpu.batch(n_reads, int(lens.sum()))
    batch.load_concat(codes, off, lens, both_strands=False)
    batch.wait()
    tie_cap = 64
    t = time.perf_counter()
    batch.run(ea.RUN_SCORES32 | ea.RUN_SELECT | ea.RUN_NO_GRAPH, tie_cap=tie_cap)
    batch.wait()
    wall = time.perf_counter() - t
    dp, _ = batch.timing()
    cells, n16, n32 = batch.stats()
    n_diag = batch.n_diag()
    recs, ties = batch.select_raw()
    tied = int((recs[:, 1] > 1).sum())
    out = {"workload": f"10k refs 200-1000 nt (2k seeds + 4 copies at 2% substitutions), {n_reads} reads "
                       f"100-1000 nt, local 2,-1,-10,-1, forward strand",
           "reads": n_reads, "n_refs": len(panel), "panel_bases": int(sum(len(p) for p in panel)),
           "cells": int(cells), "kernel_variant": chosen, "stripe_rows": vinfo(chosen).stripe_rows,
           "diag_ratio": gpu.diag_ratio, "diag_plan_x16": gpu.diag_plan,
           "autotune_ms": [round(x, 3) for x in tune_ms], "autotune_s": round(tune_s, 2),
           "reads_per_s": round(n_reads / wall, 1), "wall_ms": round(wall * 1e3, 1), "dp_ms": round(dp, 3),
           "gcups": round(cells / (dp / 1e3) / 1e9, 1), "roofline": _roof(cells, dp, "local", chosen),
           "queries_int16": int(n16), "queries_anti_diagonal": int(n_diag), "queries_int32": int(n32),
           "reads_with_ties": tied, "generate_s": round(gen_s, 1)}
    if cpu:
        from oracle import pyoracle as O
        sup = O.simd_supported()
        lanes = 32 if sup & 2 else 16
        threads = min(16, os.cpu_count() or 1)
        prof = O.Profile(panel, scoring, lanes)
        m = min(cpu_reads, n_reads)
        t = time.perf_counter()
        want = prof.score_batch(codes, off[:m], lens[:m], True, threads)
        dt = time.perf_counter() - t
        got, ok = batch.scores32()
        assert ok[:m].all() and np.array_equal(got[:m], want), "device != CPU kernel on the large panel"
        del got
        for k in range(m):
            b, nt, so, take = O.select(want[k], None, 0, None)
            r = recs[k]
            assert (r[0], r[1], r[4]) == (b, nt, len(take)), f"large panel: selection of read {k}"
            assert list(ties[k, :min(len(take), tie_cap)]) == [(rr << 1) | int(rv) for rr, rv in take[:tie_cap]], k
        cpu_cells = int(lens[:m].astype(np.int64).sum()) * out["panel_bases"]
        out["cpu_baseline"] = {"reads_per_s": round(m / dt, 1), "cells_per_s": round(cpu_cells / dt, 1),
                               "cores": threads, "kind": "port",
                               "sample": f"the first {m} reads (seeded), "
                                         f"{'AVX-512BW 32' if lanes == 32 else 'AVX2 16'}-lane restatement of the "
                                         f"reference's CPU kernel (transposed at 256 nt and over), equal on "
                                         f"every score and tie set"}
    batch.close()
    return out


def parse_lengths(text):
    """'lognormal:MEDIAN:SIGMA:LO:HI' or 'uniform:LO:HI' -> synth length spec."""
    kind, *vals = text.split(":")
    if kind == "lognormal" and len(vals) == 4:
        return ("lognormal", float(vals[0]), float(vals[1]), int(vals[2]), int(vals[3]))
    if kind == "uniform" and len(vals) == 2:
        return ("uniform", int(vals[0]), int(vals[1]))
    raise SystemExit(f"bad --lengths {text!r}")


def describe_lengths(length):
    if length[0] == "lognormal":
        return f"lognormal(median {length[1]:g}, sigma {length[2]:g}) clipped {length[3]}-{length[4]}"
    return f"uniform {length[1]}-{length[2]}"


def recorded_traffic(mode, reads, variant, lengths):
    """HBM bytes per k_dp16 launch recorded from rocprofv3 counter runs
    (FETCH_SIZE + WRITE_SIZE, each scaled by what the counter reads for a
    device copy of known size) in the committed profile summaries, for the
    same workload (mode, read count, kernel variant, length distribution);
    (None, None) when no summary matches. Counters cannot be read from
    inside this process."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(path) as f:
            text = f.read()
        for m in re.finditer(r"<!-- bench-traffic (\{.*?\}) -->", text):
            d = json.loads(m.group(1))
            key = (d.get("mode"), d.get("reads"), d.get("variant"), d.get("lengths", "lognormal:90:0.6:20:1000"))
            if key == (mode, reads, variant, lengths):
                return int(d["fetch_bytes"]) + int(d["write_bytes"]), os.path.relpath(path, ROOT)
    return None, None


def calibrate_copy(nbytes):
    """One hipMemcpy device-to-device of nbytes (the runtime's copy kernel),
    the known-size reference for rocprof's HBM byte counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    print(json.dumps({"calibrate_copy_bytes": nbytes}), flush=True)


def serving(rate, duration, seed=0x5E4E):
    """The serving configuration: three resident panels (164, 1k and 10k
    random references of 140-1000 nt), Poisson arrivals at `rate` requests/s
    for `duration` s, each request 1-50k reads (log-uniform) of the batched
    workload's lengths against a uniformly chosen panel, both strands,
    semiglobal 1,-1,-2,-1, fused selection. Requests are cut into device
    batches of at most 2^37 DP cells, two batches in flight. Reports latency
    from each request's scheduled arrival to its result (the server's
    submit-to-result time plus any lag of the submitting thread) and reads/s
    over the run."""
    import threading
    import eschalign_amd as ea
    from eschalign_amd import synth

    rng = np.random.default_rng(seed)
    panels = [synth.panel(seed=seed + k, n_refs=n, lo=140, hi=1000)
              for k, n in enumerate((164, 1000, 10000))]
    scoring = ea.Scoring(1, -1, -2, -1)
    server = ea.Server(hbm_budget=96 << 30, max_batch_reads=50_000, max_batch_bases=8_000_000,
                       max_slots=2, max_batch_cells=1 << 37)
    ids = [server.add_panel(ea.Panel(p), scoring, ea.Mode.SEMIGLOBAL) for p in panels]
    for pid, p in zip(ids, panels):
        server.autotune(pid, synth.reads(p, seed=seed + 7, n=5000))
    n_req = max(1, int(rng.poisson(rate * duration)))
    arrivals = np.sort(rng.uniform(0, duration, n_req))
    reqs = []
    for r in range(n_req):
        n = int(round(np.exp(rng.uniform(0, np.log(50_000)))))
        pid = int(rng.integers(0, 3))
        reqs.append((pid, synth.reads(panels[pid], seed=seed + 100 + r, n=n)))
    done = [None] * n_req
    lag = [0.0] * n_req
    kept = {}
    t0 = time.perf_counter()

    def feeder():
        for r, (pid, reads) in enumerate(reqs):
            dt = arrivals[r] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            lag[r] = max(0.0, time.perf_counter() - t0 - arrivals[r])
            done[r] = server.submit(ids[pid], reads, both_strands=True, tie_cap=8)

    th = threading.Thread(target=feeder)
    th.start()
    lat = []
    for r in range(n_req):
        while done[r] is None:
            time.sleep(0.0005)
        server.wait(done[r])
        res, l = server.result(done[r])
        lat.append(l + lag[r])
        if len(reqs[r][1]) <= 2_000:
            kept[r] = res
    th.join()
    wall = time.perf_counter() - t0
    st = server.stats()
    server.close()
    # a served result equals the request run alone: the two smallest
    # requests per panel, again through a scorer of their own
    checked = 0
    for k, p in enumerate(panels):
        mine = sorted((len(reqs[r][1]), r) for r in kept if reqs[r][0] == k)[:2]
        if not mine:
            continue
        alone = ea.GpuScorer(ea.Panel(p), scoring, ea.Mode.SEMIGLOBAL)
        for _, r in mine:
            want = alone.select_batch(reqs[r][1], both_strands=True, tie_cap=8)
            assert want == kept[r], f"served request {r} != the same request alone"
            checked += 1
        del alone
    total = sum(len(x[1]) for x in reqs)
    lat = np.array(lat)

    def pct(v):
        return {"requests": int(len(v)),
                "p50_ms": round(float(np.percentile(v, 50)) * 1e3, 2),
                "p90_ms": round(float(np.percentile(v, 90)) * 1e3, 2),
                "p99_ms": round(float(np.percentile(v, 99)) * 1e3, 2),
                "mean_ms": round(float(np.mean(v)) * 1e3, 2)} if len(v) else {"requests": 0}

    pids = np.array([x[0] for x in reqs])
    per_panel = {f"{len(p)}_refs": pct(lat[pids == k]) for k, p in enumerate(panels)}
    for k, p in enumerate(panels):
        per_panel[f"{len(p)}_refs"]["reads"] = int(sum(len(x[1]) for x in reqs if x[0] == k))
    return {"requests": n_req, "rate_per_s": rate, "window_s": duration, "reads": total,
            "reads_per_s": round(total / wall, 1), "wall_s": round(wall, 2),
            **{k: v for k, v in pct(lat).items() if k != "requests"},
            "per_panel": per_panel, "checked_alone": checked,
            "batches": int(st["batches"]), "hbm_peak_gib": round(st["hbm_peak"] / 2**30, 2),
            "panels": "164 / 1,000 / 10,000 refs of 140-1000 nt"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--reads", type=int, default=1_000_000)
    ap.add_argument("--mode", default="semiglobal", choices=["local", "semiglobal"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-kernel comparison")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving run")
    ap.add_argument("--no-trace", action="store_true", help="skip the winner-traceback leg")
    ap.add_argument("--no-pairs", action="store_true", help="skip the paired-mapping leg")
    ap.add_argument("--no-map", action="store_true", help="skip the read-mapping leg")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the plain-CPU-path (10k reads) and large-panel (10k refs) legs")
    ap.add_argument("--config3-reads", type=int, default=100_000, help="reads of the large-panel leg")
    ap.add_argument("--only", choices=["config1", "config3", "serving"], default=None,
                    help="only run that leg and print its record (profiling)")
    ap.add_argument("--map-reads", type=int, default=1_000_000, help="reads of the read-mapping leg")
    ap.add_argument("--pairs", type=int, default=2_000_000, help="pairs of the paired-mapping leg")
    ap.add_argument("--pairs-only", action="store_true",
                    help="only run the paired-mapping leg and print its record (profiling)")
    ap.add_argument("--pairs-both", action="store_true",
                    help="paired-mapping leg: consider both template strands")
    ap.add_argument("--trace-reads", type=int, default=200_000,
                    help="reads whose selected winners the traceback leg traces")
    ap.add_argument("--lengths", default="lognormal:90:0.6:20:1000",
                    help="read lengths: lognormal:MEDIAN:SIGMA:LO:HI or uniform:LO:HI")
    ap.add_argument("--diag-ratio", type=float, default=None,
                    help="per-length-bucket routing ratio to the anti-diagonal kernel "
                         "(0 = off; default: the library's, or the autotuner's)")
    ap.add_argument("--variant", type=int, default=-1,
                    help="batched-kernel variant to run (default: autotune on a sample)")
    ap.add_argument("--serve-rate", type=float, default=3.0, help="serving arrivals per second")
    ap.add_argument("--serve-seconds", type=float, default=70.0, help="serving arrival window")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES (scales rocprof's "
                         "FETCH_SIZE / WRITE_SIZE counters to bytes)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(f"bench.py measures one GPU only (the scoring path shards by reads with no "
              f"collective); --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    if args.only == "config1":
        print(json.dumps(config1_leg(args.steps, not args.no_cpu)), flush=True)
        return
    if args.only == "config3":
        print(json.dumps(config3_leg(args.config3_reads, not args.no_cpu, variant=args.variant)), flush=True)
        return
    if args.only == "serving":
        print(json.dumps(serving(args.serve_rate, args.serve_seconds)), flush=True)
        return
    if args.pairs_only:
        print(json.dumps(pairs_leg(args.pairs, not args.no_cpu, args.pairs_both)), flush=True)
        return

    import eschalign_amd as ea
    from eschalign_amd import synth

    scoring = (1, -1, -2, -1)
    mode = ea.Mode.parse(args.mode)
    local = mode == ea.Mode.LOCAL
    t0 = time.perf_counter()
    panel_codes = synth.panel(seed=0x0123456789ABCDEF, n_refs=164, lo=140, hi=160)
    length = parse_lengths(args.lengths)
    codes, off, lens = synth.read_arrays(panel_codes, seed=0x5EED_2002, n=args.reads, length=length)
    log(f"[bench] workload: {args.reads} reads, mean {lens.mean():.1f} nt, panel "
        f"{len(panel_codes)} refs / {sum(len(p) for p in panel_codes)} bases "
        f"({time.perf_counter() - t0:.1f} s to generate)")

    panel = ea.Panel(panel_codes)
    gpu = ea.GpuScorer(panel, ea.Scoring(*scoring), mode)
    if args.diag_ratio is not None:
        gpu.diag_ratio = args.diag_ratio
    if args.variant >= 0:
        gpu.variant = args.variant
        chosen, tune_ms = args.variant, []
    else:
        # autotune the batched kernel on a sample of the workload's own reads
        n_tune = min(args.reads, 100_000)
        tune_reads = [codes[int(off[k]):int(off[k]) + int(lens[k])] for k in range(n_tune)]
        chosen, tune_ms = gpu.autotune(tune_reads, both_strands=True)
        log(f"[bench] autotune on {n_tune} reads: variant DP ms {tune_ms} -> variant {chosen} "
            f"({gpu.registers()} VGPRs), routing ratio {gpu.diag_ratio}")
    batch = gpu.batch(args.reads, int(lens.sum()))
    batch.load_concat(codes, off, lens, both_strands=True)
    batch.wait()
    eager = ea.RUN_SCORES32 | ea.RUN_SELECT | ea.RUN_NO_GRAPH
    graph = ea.RUN_SCORES32 | ea.RUN_SELECT
    tie_cap = 8

    for _ in range(args.warmup):
        for what in (eager, graph):  # the first graph run captures the graph
            batch.run(what, tie_cap=tie_cap)
            batch.wait()
    # 1. eager launches: HIP events on the batch's stream around the DP
    # kernels give the kernel interval the roofline is priced on
    dp_ms = []
    t = time.perf_counter()
    for _ in range(args.steps):
        batch.run(eager, tie_cap=tie_cap)
        batch.wait()  # every step's events are read before the next is queued
        dp_ms.append(batch.timing()[0])
    eager_wall = time.perf_counter() - t
    eager_s32, _ = batch.scores32()
    # 2. the captured graph replayed: the headline (the library's default
    # run mode); its output is what the parity checks below read
    t = time.perf_counter()
    for _ in range(args.steps):
        batch.run(graph, tie_cap=tie_cap)
        batch.wait()
    wall = time.perf_counter() - t
    timed_s32, timed_ok = batch.scores32()
    timed_recs, timed_ties = batch.select_raw()
    assert timed_ok.all(), "a row of the timed batch was left unscored"
    assert np.array_equal(timed_s32, eager_s32), "graph replay != eager launches"
    del eager_s32
    cells, n16, n32 = batch.stats()
    n_diag = batch.n_diag()
    # the boundary hands over host buffers: time loading the same batch again
    # (gather into pinned staging + host-to-device copy); not part of `value`
    t = time.perf_counter()
    batch.load_concat(codes, off, lens, both_strands=True)
    batch.wait()
    load_s = time.perf_counter() - t
    # 3. whole steps from host buffers to host records: load (H2D), the
    # graph, the selection records and tie lists back (D2H)
    n_xfer = max(1, min(args.steps, 5))
    t = time.perf_counter()
    for _ in range(n_xfer):
        batch.load_concat(codes, off, lens, both_strands=True)
        batch.run(graph, tie_cap=tie_cap)
        xr, xt = batch.select_raw()
    xfer = (time.perf_counter() - t) / n_xfer
    assert np.array_equal(xr, timed_recs) and np.array_equal(xt, timed_ties)
    # 4. the same with two batches, as a caller keeps one batch scoring
    # while it packs the next: step k + 1 gathered and uploaded while step k
    # scores, launched as step k ends (queued earlier, its persistent DP
    # kernel would hold the CUs that step k's selection waits for), and step
    # k's records copied back while step k + 1 scores
    other = gpu.batch(args.reads, int(lens.sum()))
    other.load_concat(codes, off, lens, both_strands=True)
    other.run(graph, tie_cap=tie_cap)  # its own graph captured
    other.wait()
    two = (batch, other)
    n_pipe = 6  # the first load and the last copy back amortised over six steps
    t = time.perf_counter()
    two[0].load_concat(codes, off, lens, both_strands=True)
    two[0].run(graph, tie_cap=tie_cap)
    for k in range(1, n_pipe + 1):
        if k < n_pipe:
            two[k % 2].load_concat(codes, off, lens, both_strands=True)
        two[(k - 1) % 2].wait()
        if k < n_pipe:
            two[k % 2].run(graph, tie_cap=tie_cap)
        pr, pt = two[(k - 1) % 2].select_raw()
        assert np.array_equal(pr, timed_recs) and np.array_equal(pt, timed_ties)
    pipe = (time.perf_counter() - t) / n_pipe
    del other, two, pr, pt
    ms_step = wall / args.steps * 1e3
    value = args.reads / (wall / args.steps)
    dp = float(np.mean(dp_ms)) / 1e3
    gcups = cells / dp / 1e9
    f16 = vinfo(chosen).f16
    shifted = vinfo(chosen).shifted
    ops = OPS_PER_CELL[(args.mode, f16, shifted)]
    achieved = cells * ops / dp
    log(f"[bench] {value:,.0f} reads/s, {ms_step:.1f} ms/step (graph; eager "
        f"{eager_wall / args.steps * 1e3:.1f}), DP {dp * 1e3:.1f} ms ({gcups:,.0f} GCUPS), {n16} int16 / "
        f"{n_diag} anti-diagonal / {n32} int32 queries; from host buffers {xfer * 1e3:.1f} ms/step")

    # With an fp16 variant chosen, the same steps again with its int16
    # counterpart on the same stripes, so both arithmetics' rates are on
    # record.
    int16 = None
    v16 = int16_counterpart(chosen)
    if f16:
        gpu.variant = v16
        batch.run(eager, tie_cap=tie_cap)
        batch.wait()
        d16 = []
        t = time.perf_counter()
        for _ in range(args.steps):
            batch.run(eager, tie_cap=tie_cap)
            batch.wait()
            d16.append(batch.timing()[0])
        w16 = time.perf_counter() - t
        got16, _ = batch.scores32()
        assert np.array_equal(got16, timed_s32), f"int16 variant {v16} != variant {chosen}"
        del got16
        gpu.variant = chosen
        dp16 = float(np.mean(d16)) / 1e3
        ops16 = OPS_PER_CELL[(args.mode, False, shifted)]
        int16 = {"variant": v16, "reads_per_s": round(args.reads / (w16 / args.steps), 1),
                 "ms_per_step": round(w16 / args.steps * 1e3, 3), "dp_ms": round(dp16 * 1e3, 3),
                 "gcups": round(cells / dp16 / 1e9, 1), "ops_per_cell": ops16,
                 "roofline_frac": round(cells * ops16 / dp16 / PEAK_I16, 4)}
        log(f"[bench] int16 variant {v16} on the same batch: {int16}")

    plain = None
    if not args.no_plain:
        pl = ea.RUN_SCORES32 | ea.RUN_SELECT | ea.RUN_PLAIN | ea.RUN_NO_GRAPH
        batch.run(pl, tie_cap=tie_cap)
        batch.wait()
        t = time.perf_counter()
        batch.run(pl, tie_cap=tie_cap)
        batch.wait()
        pw = time.perf_counter() - t
        pdp, _ = batch.timing()
        gotp, _ = batch.scores32()
        assert np.array_equal(gotp, timed_s32), "plain kernel != batched kernel"
        del gotp
        plain_rate = args.reads / pw
        plain = {"reads_per_s": round(plain_rate, 1),
                 "dp_ms": round(pdp, 3),
                 "gcups": round(cells / (pdp / 1e3) / 1e9, 1),
                 "speedup": round(value / plain_rate, 2),
                 "speedup_int16": round(int16["reads_per_s"] / plain_rate, 2) if int16 else None}
        log(f"[bench] plain int32 kernel: {plain}")

    cpu = None
    if not args.no_cpu:
        n_all = args.reads

        def gpu_check(qs, cpu_scores):
            # the timed batch's own rows: reads 0..n-1, then their reverse
            # complements at n_all + 0..n-1 (the CPU sample's order)
            n = len(qs) // 2
            got = np.concatenate([timed_s32[:n], timed_s32[n_all:n_all + n]])
            assert np.array_equal(got, cpu_scores), "timed GPU batch != CPU kernel on the sample"
            # and the fused selection of the timed batch, on the first 20k
            for k in range(min(n, 20_000)):
                b, nt, so, take = O.select(cpu_scores[k], cpu_scores[n + k], 0, None)
                r = timed_recs[k]
                want = [(rr << 1) | int(rv) for rr, rv in take[:tie_cap]]
                assert (r[0], r[1], r[4]) == (b, nt, len(take)), f"timed selection != CPU on read {k}"
                assert (so is None) == (not (r[3] & 2)) and (so is None or r[2] == so), k
                assert list(timed_ties[k, :min(len(take), tie_cap)]) == want, k
        from oracle import pyoracle as O
        cpu = cpu_baseline(panel_codes, codes, off, lens, scoring, local, gpu_check)
        log(f"[bench] CPU baseline: {cpu}")
    del timed_s32

    trace = None
    if not args.no_trace:
        batch.run(graph, tie_cap=tie_cap)  # the selection the winners come from
        batch.wait()
        trace = trace_leg(gpu, batch, panel_codes, codes, off, lens, scoring, local, args.trace_reads,
                          not args.no_cpu)
        log(f"[bench] winner tracebacks: {trace}")
    batch.close()

    mapped = None
    if not args.no_map:
        mapped = map_reads_leg(gpu, panel_codes, codes, off, lens, scoring, local,
                               min(args.map_reads, args.reads), not args.no_cpu)
        log(f"[bench] read mapping: {mapped}")
    gpu_ratio, gpu_plan = gpu.diag_ratio, gpu.diag_plan
    del gpu

    configs = {}
    if not args.no_configs:
        configs["plain_cpu_path_10k"] = config1_leg(args.steps, not args.no_cpu)
        log(f"[bench] config plain CPU path: {configs['plain_cpu_path_10k']}")
        configs["large_panel_10k_refs"] = config3_leg(args.config3_reads, not args.no_cpu)
        log(f"[bench] config large panel: {configs['large_panel_10k_refs']}")

    pairs = None
    if not args.no_pairs:
        pairs = pairs_leg(args.pairs, not args.no_cpu, args.pairs_both)
        log(f"[bench] paired mapping: {pairs}")
        if not args.pairs_both:
            # both template strands: every orientation of every mate scored
            # (the reference's GPU thread scores all four for a paired run)
            pairs["both_strands"] = pairs_leg(args.pairs, False, True)
            log(f"[bench] paired mapping, both strands: {pairs['both_strands']}")

    serve = None
    if not args.no_serving:
        serve = serving(args.serve_rate, args.serve_seconds)
        log(f"[bench] serving: {serve}")

    traffic, traffic_src = recorded_traffic(args.mode, args.reads, chosen, args.lengths)
    peak = ea.valu_peak16(0)
    log(f"[bench] measured int16 VALU peak {peak['lane_ops_per_s'] / 1e12:.2f} T lane-op/s "
        f"({peak['n_cu']} CUs, reported clock {peak['clock_khz'] / 1e3:.0f} MHz)")
    roof = {"bound": "valu", "achieved": round(achieved / 1e12, 3),
            "peak": round(PEAK_I16 / 1e12, 3), "unit": "Tlane-op/s (packed 16-bit)",
            "frac": round(achieved / PEAK_I16, 4), "traffic": traffic,
            "peak_measured": round(peak["lane_ops_per_s"] / 1e12, 3),
            "frac_of_measured": round(achieved / peak["lane_ops_per_s"], 4),
            "ops_per_cell": ops,
            "ops_basis": "the minimum packed lane-ops per cell of the formulation the kernel runs "
                         "(5: fp16 with the diagonal shift; 6: fp16 with the row shift or int16 with the "
                         "diagonal shift; 7: int16 row-shifted or fp16 unshifted; 8: int16)",
            "valu_lane_ops_per_cell_issued": VALU_OPS_PER_CELL_ISSUED.get((args.mode, chosen)),
            "traffic_unit": "B per k_dp16 launch (HBM, rocprofv3 FETCH_SIZE + WRITE_SIZE)",
            "traffic_source": traffic_src}
    fracs = [roof["frac"], roof["frac_of_measured"]]
    if int16:
        fracs.append(int16["roofline_frac"])
    for c in configs.values():
        fracs.append(c["roofline"]["frac"])
    assert all(0 < f <= 1 for f in fracs), f"a roofline fraction outside (0, 1]: {fracs}"
    line = {
        "metric": "reads/s + GCUPS, bit-exact scores vs CPU; % VALU roofline; p50/p99 latency",
        "value": round(value, 1),
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16x2 (integer-exact; int16 / int32 rescue)" if f16 else "i16x2 (int32 rescue)",
        "data": "synthetic",
        "config": {"workload": "batched uneven lengths: 164-ref panel (140-160 nt), "
                               f"{args.reads} reads {describe_lengths(length)}, "
                               f"both strands, {args.mode}, scoring 1,-1,-2,-1",
                   "reads": args.reads, "queries": 2 * args.reads, "n_refs": len(panel),
                   "queries_int16": int(n16), "queries_anti_diagonal": int(n_diag),
                   "queries_int32": int(n32),
                   "parallelism": "single GPU"},
        "value_scope": "device pipeline (bucketing, DP, fused selection) replayed from its captured "
                       "hipGraph, reads resident in HBM; host transfers in transfer_inclusive",
        "eager_ms_per_step": round(eager_wall / args.steps * 1e3, 3),
        "graph_ms_per_step": round(ms_step, 3),
        "gcups": round(gcups, 1),
        "dp_ms": round(dp * 1e3, 3),
        "cells": int(cells),
        "roofline": roof,
        "transfer_inclusive": {"reads_per_s": round(args.reads / xfer, 1), "ms_per_step": round(xfer * 1e3, 3),
                               "h2d_bytes": int(lens.sum()) + 12 * args.reads,
                               "d2h_bytes": int(xr.nbytes + xt.nbytes), "load_ms": round(load_s * 1e3, 3),
                               "scope": "reads from host buffers (gather into pinned staging, copy), the "
                                        "graph, selection records and tie lists copied back",
                               "two_batches_reads_per_s": round(args.reads / pipe, 1),
                               "two_batches_ms_per_step": round(pipe * 1e3, 3),
                               "two_batches_scope": "the same steps with two batches alternating: step k + 1 "
                                                    "loaded while step k scores, launched as it ends, step k's "
                                                    "records copied back while step k + 1 scores"},
        "kernel_variant": chosen,
        "stripe_rows": vinfo(chosen).stripe_rows,
        "int16_variant": int16,
        "diag_ratio": gpu_ratio,
        "diag_plan": {"classes_nt": ["<128", "128-255", "256-511", "512-1023", "1024-2047", "2048-4095",
                                     "4096-8191", ">=8192"],
                      "per_class_x16": gpu_plan,
                      "meaning": "0 inter-sequence, 1 anti-diagonal, 16..1024 ratio rule (rho x 16)"},
        "autotune_ms": [round(x, 3) for x in tune_ms],
        "plain_kernel": plain,
        "cpu_baseline": cpu,
        "configs": configs,
        "trace": trace,
        "map_reads": mapped,
        "pairs": pairs,
        "serving": serve,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
