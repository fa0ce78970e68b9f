"""Random scoring schemes, panels and reads, every k_dp16 variant a panel
admits (and the default), both modes: the int32 scores equal the oracle's
score matrix (lib/scalar.rs:120-163). Schemes are drawn from what
Scoring::validate accepts (lib/scoring.rs:70-100), including zero gaps,
positive mismatches and large matches. Those sit near or beyond the fp16
and shifted bounds, so the refusals, the int16 / int32 fallbacks and the
per-batch fp16 choice are all crossed. Seeded and bounded (40 panels)."""
import random

import numpy as np
import pytest

import eschalign_amd as ea
from oracle import pyoracle as O
import refgen as G

THREADS = 16


@pytest.mark.gpu
def test_random_schemes_every_variant_equals_oracle():
    rng = random.Random(0x5EED_0C01)
    lib = ea.load_library()
    panels = checked = 0
    while panels < 40:
        ma = rng.choice([1, 2, 3, 5, 10, 40])
        mi = rng.choice([-ma, -1, -2 * ma, 0, -3]) if rng.random() < 0.9 else 1
        o = -rng.choice([0, 1, 2, 5, 10, 30])
        e = -rng.choice([0, 1, 2, 3, 7]) if o < 0 else 0
        e = max(e, o)
        if O.validate(ma, mi, o, e) != 0:
            continue
        panels += 1
        refs = [G.random_seq(G.Rng(rng.randrange(1 << 30)), rng.randint(1, rng.choice([20, 80, 200, 600, 1500])), 0.03)
                for _ in range(rng.randint(1, 40))]
        reads = [G.random_seq(G.Rng(rng.randrange(1 << 30)), rng.randint(1, rng.choice([50, 300, 1200])), 0.03)
                 for _ in range(rng.randint(1, 120))]
        reads += [r[rng.randrange(max(1, len(r) // 2)):] for r in refs[:10]]  # high scores
        q = [G.encode(r) for r in reads]
        panel = ea.Panel(refs)
        for mode in (ea.Mode.SEMIGLOBAL, ea.Mode.LOCAL):
            want = O.score_matrix(q, panel.codes, (ma, mi, o, e), mode == ea.Mode.LOCAL, THREADS)
            gpu = ea.GpuScorer(panel, ea.Scoring(ma, mi, o, e), mode)
            for v in [None] + list(range(lib.esch_n_variants())):
                if v is not None:
                    try:
                        gpu.variant = v
                    except ea.GpuError:
                        continue  # not exact on this panel: refused
                s32, ok = gpu.score_batch_i32(q)
                checked += 1
                assert ok.all() and np.array_equal(s32, want), ((ma, mi, o, e), mode, v, gpu.variant)
    assert checked > 400
