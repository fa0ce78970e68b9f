"""The serving loop on an MI355X: continuous batching over three resident
panels, admission by HBM footprint, requests arriving while others run.
Every result must equal the same request run alone (and, on a sample, the
CPU oracle)."""
import os
import random
import time

import numpy as np
import pytest

import eschalign_amd as ea
from eschalign_amd import synth
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cells", [None, 300_000_000])
def test_served_results_equal_alone_results(cells):
    # cells: a DP-cell limit per device batch small enough to cut most
    # requests into many units of length-sorted reads
    scoring = ea.Scoring(1, -1, -2, -1)
    mode = ea.Mode.SEMIGLOBAL
    panels = [synth.panel(seed=11 + k, n_refs=n, lo=140, hi=400) for k, n in enumerate((40, 90, 160))]
    # a budget of about two slots of the largest panel: admission has to wait
    probe = ea.GpuScorer(ea.Panel(panels[2]), scoring, mode)
    slot = probe.batch_footprint(4000, 600_000)
    del probe
    budget = int(2.5 * slot)
    kw = {} if cells is None else {"max_batch_cells": cells}
    server = ea.Server(hbm_budget=budget, max_batch_reads=4000, max_batch_bases=600_000, max_slots=3, **kw)
    ids = [server.add_panel(ea.Panel(p), scoring, mode) for p in panels]
    alone = [ea.GpuScorer(ea.Panel(p), scoring, mode) for p in panels]

    rng = random.Random(1234)
    requests = []
    for r in range(48):
        pid = rng.randrange(3)
        n = int(round(np.exp(rng.uniform(0, np.log(3000)))))
        reads = synth.reads(panels[pid], seed=1000 + r, n=n, length=("lognormal", 90, 0.6, 20, 500))
        opts = rng.choice([dict(min_score=0, max_ties=None), dict(min_score=20, max_ties=2),
                           dict(min_score=-5, max_ties=0)])
        requests.append((pid, reads, opts))

    t0 = time.perf_counter()
    submitted = []
    for pid, reads, opts in requests:
        time.sleep(rng.expovariate(1 / 0.004))   # Poisson arrivals, mean gap 4 ms
        rid = server.submit(ids[pid], reads, both_strands=True, tie_cap=8, **opts)
        submitted.append(rid)
    lat = []
    results = []
    for rid in submitted:
        assert server.wait(rid, timeout_ms=120_000)
        res, l = server.result(rid)
        results.append(res)
        lat.append(l)
    wall = time.perf_counter() - t0
    st = server.stats()
    assert st["completed"] == len(requests)
    assert st["hbm_peak"] <= budget
    assert st["batches"] >= 1
    if cells is not None:
        assert st["batches"] > 2 * len(requests)
    print(f"[serving] {len(requests)} requests in {wall:.2f} s, {st['batches']} batches, "
          f"p50 {np.percentile(lat, 50) * 1e3:.1f} ms, p99 {np.percentile(lat, 99) * 1e3:.1f} ms, "
          f"peak HBM {st['hbm_peak'] / 2**30:.2f} GiB of {budget / 2**30:.2f}")

    for (pid, reads, opts), res in zip(requests, results):
        want = alone[pid].select_batch(reads, both_strands=True, tie_cap=8, **opts)
        assert res == want
    # a sample against the CPU oracle
    pid, reads, opts = requests[0]
    rows = reads[:50] + [ea.reverse_complement_codes(q) for q in reads[:50]]
    sm = O.score_matrix(rows, panels[pid], scoring.tuple(), False, min(16, os.cpu_count() or 1))
    n = len(reads[:50])
    for k in range(n):
        b, nt, so, take = O.select(sm[k], sm[n + k], opts["min_score"], opts["max_ties"])
        r = results[0][k]
        assert (r.best, r.n_tied, r.suboptimal, r.take) == (b, nt, so, take[:8])
    server.close()


def test_small_request_overtakes_large_one():
    """A request with little work, submitted after a large one, finishes
    first: units of the request with the least queued work go first."""
    scoring = ea.Scoring(1, -1, -2, -1)
    mode = ea.Mode.SEMIGLOBAL
    panel = synth.panel(seed=23, n_refs=300, lo=140, hi=400)
    server = ea.Server(hbm_budget=16 << 30, max_batch_reads=4000, max_batch_bases=600_000, max_slots=2,
                       max_batch_cells=500_000_000)
    pid = server.add_panel(ea.Panel(panel), scoring, mode)
    big = synth.reads(panel, seed=5, n=6000, length=("lognormal", 90, 0.6, 20, 500))
    small = synth.reads(panel, seed=6, n=5, length=("lognormal", 90, 0.6, 20, 500))
    t_big = time.perf_counter()
    r_big = server.submit(pid, big, both_strands=True, tie_cap=4)
    t_small = time.perf_counter()
    r_small = server.submit(pid, small, both_strands=True, tie_cap=4)
    assert server.wait(r_small, timeout_ms=120_000)
    assert server.wait(r_big, timeout_ms=120_000)
    res_small, l_small = server.result(r_small)
    res_big, l_big = server.result(r_big)
    st = server.stats()
    print(f"[serving] large {l_big * 1e3:.1f} ms, small {l_small * 1e3:.1f} ms, {st['batches']} batches")
    assert st["batches"] > 50
    assert t_small + l_small < t_big + l_big
    alone = ea.GpuScorer(ea.Panel(panel), scoring, mode)
    assert res_small == alone.select_batch(small, both_strands=True, tie_cap=4)
    assert res_big == alone.select_batch(big, both_strands=True, tie_cap=4)
    server.close()


def test_config5_real_panels_and_swapped_roles():
    """The serving configuration's own panels: 164, 1,000 and 10,000
    references of 140-1000 nt, semiglobal 1,-1,-2,-1, reads lognormal(90,
    0.6) clipped 20-1000, Poisson arrivals, request sizes log-uniform. On
    the 10,000-reference panel a device batch of at most one row per
    reference is scored with the roles swapped (the panel's references as
    the queries); the server takes that path for the small requests, and
    every served result equals the same request run alone on a scorer with
    swapping off (the ordinary path), and a sample equals the CPU kernel."""
    scoring = ea.Scoring(1, -1, -2, -1)
    mode = ea.Mode.SEMIGLOBAL
    panels = [synth.panel(seed=0x5E4E + k, n_refs=n, lo=140, hi=1000) for k, n in enumerate((164, 1000, 10000))]
    server = ea.Server(hbm_budget=64 << 30, max_batch_reads=50_000, max_batch_bases=8_000_000, max_slots=2,
                       max_batch_cells=1 << 38)
    ids = [server.add_panel(ea.Panel(p), scoring, mode) for p in panels]
    alone = [ea.GpuScorer(ea.Panel(p), scoring, mode) for p in panels]
    for a in alone:
        a.swap_rows = 0
    rng = random.Random(0x5E4E)
    requests = []
    for r in range(36):
        pid = r % 3 if r < 6 else rng.randrange(3)
        n = int(round(np.exp(rng.uniform(0, np.log(4000)))))
        reads = synth.reads(panels[pid], seed=0x5E4E + 100 + r, n=n)
        requests.append((pid, reads))
    submitted = []
    for pid, reads in requests:
        time.sleep(rng.expovariate(1 / 0.01))
        submitted.append(server.submit(ids[pid], reads, both_strands=True, tie_cap=8))
    results = []
    for rid in submitted:
        assert server.wait(rid, timeout_ms=180_000)
        results.append(server.result(rid)[0])
    st = server.stats()
    assert st["completed"] == len(requests)
    assert st["swapped_batches"] > 0, st
    print(f"[serving] config-5 panels: {st['batches']} batches, {st['swapped_batches']} with the roles swapped")
    for k, ((pid, reads), res) in enumerate(zip(requests, results)):
        assert res == alone[pid].select_batch(reads, both_strands=True, tie_cap=8), (k, pid, len(reads))
    # a sample of the 10k-panel requests against the CPU kernel
    sup = O.simd_supported()
    for k, (pid, reads) in enumerate(requests):
        if pid != 2:
            continue
        q = reads[:12]
        rows = q + [ea.reverse_complement_codes(x) for x in q]
        prof = O.Profile(panels[2], scoring.tuple(), 32 if sup & 2 else 16)
        sm = prof.score_batch(*O.concat(rows), False, min(16, os.cpu_count() or 1))
        n = len(q)
        for i in range(n):
            b, nt, so, take = O.select(sm[i], sm[n + i], 0, None)
            r = results[k][i]
            assert (r.best, r.n_tied, r.suboptimal, r.take) == (b, nt, so, take[:8]), (k, i)
        break
    server.close()


def test_serving_local_mode_and_panels_that_restart_their_shift():
    """Two resident panels of references up to 1,900 nt, one in local mode
    and one in semiglobal mode, under 1,-1,-2,-1: their default variants
    restart the column shift (local every 128 columns, semiglobal every
    64), and a batch of short reads may take another block. Requests of
    1-2,000 reads arriving while others run: every result equals the same
    request alone, and a sample of each panel's equals the CPU oracle."""
    scoring = ea.Scoring(1, -1, -2, -1)
    rng = random.Random(4321)
    base = synth.panel(seed=77, n_refs=60, lo=140, hi=1000)
    long_refs = synth.panel(seed=78, n_refs=3, lo=1500, hi=1900)
    panels = [long_refs + base, long_refs[::-1] + base[::-1]]
    modes = [ea.Mode.LOCAL, ea.Mode.SEMIGLOBAL]
    server = ea.Server(hbm_budget=16 << 30, max_batch_reads=4000, max_batch_bases=2_000_000, max_slots=3,
                       max_batch_cells=2_000_000_000)
    ids = [server.add_panel(ea.Panel(p), scoring, m) for p, m in zip(panels, modes)]
    alone = [ea.GpuScorer(ea.Panel(p), scoring, m) for p, m in zip(panels, modes)]
    for g in alone:
        assert g.column_block(g.variant) > 0
    requests = []
    for r in range(24):
        pid = r % 2
        n = int(round(np.exp(rng.uniform(0, np.log(2000)))))
        reads = synth.reads(panels[pid], seed=2000 + r, n=n, length=("lognormal", 90, 0.8, 20, 1500))
        requests.append((pid, reads))
    submitted = []
    for pid, reads in requests:
        time.sleep(rng.expovariate(1 / 0.002))
        submitted.append(server.submit(ids[pid], reads, both_strands=True, tie_cap=8))
    results = []
    for rid in submitted:
        assert server.wait(rid, timeout_ms=120_000)
        results.append(server.result(rid)[0])
    assert server.stats()["completed"] == len(requests)
    server.close()
    for (pid, reads), res in zip(requests, results):
        assert res == alone[pid].select_batch(reads, both_strands=True, tie_cap=8)
    for pid in (0, 1):
        k0 = next(k for k, (p, reads) in enumerate(requests) if p == pid and len(reads) >= 30)
        reads = requests[k0][1][:30]
        rows = reads + [ea.reverse_complement_codes(q) for q in reads]
        sm = O.score_matrix(rows, panels[pid], scoring.tuple(), modes[pid] == ea.Mode.LOCAL,
                            min(16, os.cpu_count() or 1))
        n = len(reads)
        for k in range(n):
            b, nt, so, take = O.select(sm[k], sm[n + k], 0, None)
            r = results[k0][k]
            assert (r.best, r.n_tied, r.suboptimal, r.take) == (b, nt, so, take[:8]), (pid, k)
