"""The shifted value spaces of k_dp16's shifted variants (eschalign_amd/csrc/
dp16.h: column16s / column16s_pair with the stripe rebase in semiglobal
mode, column16c / column16c_pair in local mode, and column16s' diagonal
shift, by row and column at once, in semiglobal mode), modelled step for
step in Python on the CPU.

Each model walks a read the way the kernel does. The read is padded at the
top to its chunk's row plan (a first stripe of RR rows, then stripes of 16),
columns go in pairs with a single tail column, each stripe's boundary is
handed on in that stripe's own space, and the best is kept as the kernel
keeps it. The models are compared with the oracle's score
(lib/scalar.rs:120-163, restated in oracle/eschalign_oracle.c and pinned to
the reference's vectors), so a wrong shift, rebase, floor or best
conversion fails.

They also record every value they form, against the ranges the exactness
bounds in esch_internal.h claim (shifted_bound, colshift_bound,
diagshift_bound). A value
outside them would let the fp16 or int16 arithmetic round or clamp.
"""
import random

import numpy as np
import pytest

from oracle import pyoracle as O
import refgen as G

NEG = None  # -inf: an identity for max, absorbing for +


def mx(*v):
    w = [x for x in v if x is not None]
    return max(w) if w else None


def add(a, b):
    return None if a is None or b is None else a + b


class Range:
    def __init__(self):
        self.lo, self.hi = 0, 0

    def see(self, *v):
        for x in v:
            if x is not None:
                self.lo, self.hi = min(self.lo, x), max(self.hi, x)
        return v[0] if len(v) == 1 else v


def plan(nmax, st=16):
    """k_dp16's row plan for a chunk whose longest read is nmax, on stripes
    of st rows (16; 24 or 32 for variants 9-14): S stripes, the first of RR
    rows (an even count), T rows in all."""
    S = (nmax + st - 1) // st
    RR = ((nmax - st * (S - 1) + 1) // 2) * 2
    return S, RR, RR + st * (S - 1)


def subst(x, c, ma, mi):
    if x is None:
        return 0  # a top padding row
    return ma if (x == c or x == 4 or c == 4) else mi


def walk_semiglobal_rows(q, r, scoring, nmax, rng_vals, st=16):
    """Variant 7's semiglobal walk: row r of a stripe holds P = H - e r + o,
    R = E - e r + e, Q = F - e r + e; the row above arrives in the previous
    stripe's space, lower by c = e h, added to Q at row 0 and to the
    profile's row 0; bests in H + o space (the last column's rows through
    + e r, the last stripe's bottom row through + e (RR - 1))."""
    ma, mi, o, e = scoring
    S, RR, T = plan(nmax, st)
    rows = [None] * (T - len(q)) + list(q)
    pad = T - len(q)
    m = len(r)
    best = None
    bottom = None
    row0 = 0
    for s in range(S):
        h = RR if s == 0 else st
        c = 0 if s == 0 else e * (RR if s == 1 else st)
        Pv = [o - e * k for k in range(h)]
        R = [NEG] * h
        hdt = o + e - c
        bestL = NEG
        newb = []

        def column(j, pu, qq, pd):
            for k in range(h):
                sc = subst(rows[row0 + k], r[j], ma, mi) - o + (c if k == 0 else 0)
                R[k] = rng_vals.see(add(mx(Pv[k], R[k]), e))
                qq = rng_vals.see(mx(pu, qq))
                if k == 0 and s > 0:
                    qq = rng_vals.see(add(qq, c))
                hh = rng_vals.see(mx(rng_vals.see(pd + sc), R[k], qq))
                pd = Pv[k]
                Pv[k] = rng_vals.see(hh + (o - e))
                pu = Pv[k]
            return pu, qq

        j = 0
        while j < m:
            for jj in ((j, j + 1) if j + 1 < m else (j,)):
                pu, qq = (o + e, NEG) if s == 0 else bottom[jj]
                pd = hdt
                hdt = pu
                pu, qq = column(jj, pu, qq, pd)
                newb.append((pu, qq))
                bestL = mx(bestL, pu)
            j += 2
        for k in range(h):
            if row0 + k >= pad:
                best = mx(best, rng_vals.see(Pv[k] + e * k))
        if s == S - 1:
            best = mx(best, rng_vals.see(add(bestL, e * (h - 1))))
        bottom = newb
        row0 += h
    return best - o


def walk_semiglobal_diag(q, r, scoring, nmax, rng_vals, st=16, advance_top=True, rebase=0):
    """The diagonal shift's semiglobal walk (variants 23-28): row r of a
    stripe in reference column j holds P = H - e (r + j) + o,
    R = E - e (r + j) + e, Q = F - e (r + j) + e, so neither gap chain adds
    its extension and the profile carries s - o - e. The row above arrives
    in the previous stripe's space, lower by c = e h (the column part is the
    same in every stripe); the first stripe's top border P = o + e - e j is
    advanced per column; the last stripe's bottom-row best is carried in
    column j's space (a pair: best - 2e against the pair's A - e and B) and
    the last column's rows come back to H + o through + e (r + m - 1).
    advance_top=False leaves the top border at column 0's value (a planted
    mistake). rebase = K (even) restarts the column part of the shift every
    K columns, as walk_local_columns does: the rows' P and R, the corner,
    the top border and the bottom-row best move by + e K at each block
    start, and the last column comes back through + e ((m - 1) mod K)."""
    ma, mi, o, e = scoring
    S, RR, T = plan(nmax, st)
    rows = [None] * (T - len(q)) + list(q)
    pad = T - len(q)
    m = len(r)
    best = None
    bottom = None
    row0 = 0
    for s in range(S):
        h = RR if s == 0 else st
        c = 0 if s == 0 else e * (RR if s == 1 else st)
        Pv = [o + e - e * k for k in range(h)]  # the left border, column -1
        R = [NEG] * h
        hdt = o + 2 * e - c
        bestL = NEG
        vtop = o + e
        newb = []

        def column(j, pu, qq, pd):
            for k in range(h):
                sc = subst(rows[row0 + k], r[j], ma, mi) - o - e + (c if k == 0 else 0)
                R[k] = rng_vals.see(mx(Pv[k], R[k]))
                qq = rng_vals.see(mx(pu, qq))
                if k == 0 and s > 0:
                    qq = rng_vals.see(add(qq, c))
                hh = rng_vals.see(mx(rng_vals.see(pd + sc), R[k], qq))
                pd = Pv[k]
                Pv[k] = rng_vals.see(hh + (o - e))
                pu = Pv[k]
            return pu, qq

        j = 0
        while j < m:
            if rebase and j and j % rebase == 0:
                for k in range(h):
                    Pv[k] += e * rebase
                    R[k] = add(R[k], e * rebase)
                hdt += e * rebase
                vtop += e * rebase
                bestL = add(bestL, e * rebase)
            cols = (j, j + 1) if j + 1 < m else (j,)
            for jj in cols:
                if s == 0:
                    pu, qq = rng_vals.see(vtop), NEG
                    if advance_top:
                        vtop -= e
                else:
                    pu, qq = bottom[jj]
                pd = hdt
                hdt = pu
                pu, qq = column(jj, pu, qq, pd)
                newb.append((pu, qq))
            if len(cols) == 2:
                bestL = rng_vals.see(mx(add(bestL, -2 * e), rng_vals.see(newb[-2][0] - e), newb[-1][0]))
            else:
                bestL = rng_vals.see(mx(add(bestL, -e), newb[-1][0]))
            j += 2
        jl = (m - 1) % rebase if rebase else m - 1  # column m - 1's place in its block
        for k in range(h):
            if row0 + k >= pad:
                best = mx(best, rng_vals.see(Pv[k] + e * k + e * jl))
        if s == S - 1:
            best = mx(best, rng_vals.see(add(bestL, e * (h - 1) + e * jl)))
        bottom = newb
        row0 += h
    return best - o


def walk_local_columns(q, r, scoring, nmax, rng_vals, st=16, floor_on_chain=True, floor_shift=0, rebase=0):
    """Variant 7's local walk: column j holds P = H - e j + o,
    R = E - e j + e, Q = F - e j + e, the zero floor is z_j = e (1 - j), and
    the best is carried in the space of the last column done (two columns a
    pair: best - 2e against the pair's tA - e and tB), back to H + o at the
    stripe's end. The floor rides on the horizontal chain, R = max(P, R, z_j)
    (E' = max(E, 0), which changes no H), so the cell's own maximum needs no
    fourth operand; floor_on_chain=False applies it to the cell instead (the
    earlier form, equal). floor_shift moves the floor by that many extension
    steps (a planted mistake when not 0). rebase = K (even) restarts the
    shift every K columns: at column j = K, 2K, ... every value carried in
    from column j - 1 (the rows' P and R, the corner, the top border, the
    floor and the best) moves into the new block's space by + e K, and the
    best returns to H + o through + e ((m - 1) mod K)."""
    ma, mi, o, e = scoring
    S, RR, T = plan(nmax, st)
    rows = [None] * (T - len(q)) + list(q)
    m = len(r)
    best = o  # H = 0
    bottom = None
    row0 = 0
    for s in range(S):
        h = RR if s == 0 else st
        Pv = [o + e] * h
        R = [NEG] * h
        hdt = o + e
        vtop = o
        z = e
        b = best + e
        newb = []

        def column(j, pu, qq, pd, zf):
            for k in range(h):
                sc = subst(rows[row0 + k], r[j], ma, mi) - o
                if floor_on_chain:
                    R[k] = rng_vals.see(mx(Pv[k], R[k], zf + floor_shift * e))
                    qq = rng_vals.see(add(mx(pu, qq), e))
                    hh = rng_vals.see(mx(rng_vals.see(pd + sc), R[k], qq))
                else:
                    R[k] = rng_vals.see(mx(Pv[k], R[k]))
                    qq = rng_vals.see(add(mx(pu, qq), e))
                    hh = rng_vals.see(mx(mx(rng_vals.see(pd + sc), R[k], qq), zf))
                pd = Pv[k]
                Pv[k] = rng_vals.see(hh + (o - e))
                pu = Pv[k]
            return pu, qq, max(Pv)

        j = 0
        while j < m:
            if rebase and j and j % rebase == 0:
                for k in range(h):
                    Pv[k] += e * rebase
                    R[k] = add(R[k], e * rebase)
                hdt += e * rebase
                vtop += e * rebase
                z += e * rebase
                b += e * rebase
            cols = (j, j + 1) if j + 1 < m else (j,)
            ts = []
            for jj in cols:
                if s == 0:
                    pu, qq = rng_vals.see(vtop), NEG
                    vtop -= e
                else:
                    pu, qq = bottom[jj]
                pd = hdt
                hdt = pu
                pu, qq, t = column(jj, pu, qq, pd, rng_vals.see(z))
                z -= e
                ts.append(t)
                newb.append((pu, qq))
            if len(cols) == 2:
                b = rng_vals.see(max(rng_vals.see(b - 2 * e), rng_vals.see(ts[0] - e), ts[1]))
            else:
                b = rng_vals.see(max(rng_vals.see(b - e), ts[0]))
            j += 2
        best = rng_vals.see(b + e * ((m - 1) % rebase if rebase else m - 1))
        bottom = newb
        row0 += h
    return best - o


def cases(seed, n_cases, n_max=60):
    rng = random.Random(seed)
    schemes = G.SCHEMES_GPU + [(1, -3, -1, -1), (2, -3, -5, -2)]
    for _ in range(n_cases):
        sc = schemes[rng.randrange(len(schemes))]
        n, m = rng.randint(1, n_max), rng.randint(1, 24)
        q = [rng.randrange(5) if rng.random() < 0.1 else rng.randrange(4) for _ in range(n)]
        ref = [rng.randrange(5) if rng.random() < 0.1 else rng.randrange(4) for _ in range(m)]
        if rng.random() < 0.3:  # a read that matches the reference: the top of the range
            q = (ref * (n // m + 1))[:n]
        yield sc, q, ref, n + rng.randint(0, 20)


def bounds(scoring):
    """M = max(|match|, |mismatch|), |open|, |extend|."""
    ma, mi, o, e = scoring
    return max(abs(ma), abs(mi)), abs(o), abs(e)


@pytest.mark.parametrize("seed,st", [(1, 16), (2, 16), (6, 24), (7, 32)])
def test_row_shifted_semiglobal_walk_equals_oracle(seed, st):
    """The row shift (semiglobal), on stripes of 16 rows and of the tall
    variants' 24 and 32: scores equal the oracle's, and every value lies in
    [-(kM + |o| + 4|e|), kM + st |e|] with k the read-reference overlap
    bound min(n, m), the range shifted_bound(..., st) rests on."""
    for sc, q, ref, nmax in cases(seed, 150, n_max=4 * st):
        vals = Range()
        got = walk_semiglobal_rows(q, ref, sc, nmax, vals, st)
        want = O.score(np.array(q, np.uint8), np.array(ref, np.uint8), sc, False)
        assert got == want, (sc, q, ref, nmax, st)
        M, O_, E = bounds(sc)
        k = min(len(q), len(ref))
        assert -(k * M + O_ + 4 * E) <= vals.lo and vals.hi <= k * M + st * E, (sc, vals.lo, vals.hi, k, st)


@pytest.mark.parametrize("seed,st", [(3, 16), (4, 16), (8, 24), (9, 32)])
def test_column_shifted_local_walk_equals_oracle(seed, st):
    """The column shift (local), on 16-, 24- and 32-row stripes, with the
    zero floor on the horizontal chain (the kernel's form) and on the cell:
    scores equal the oracle's, and every value lies in [-(|o| + 3|e| + M),
    kM + |e| (m + 2) + M], m the reference length, the range colshift_bound
    rests on (whatever the stripe height)."""
    for sc, q, ref, nmax in cases(seed, 150, n_max=4 * st):
        vals = Range()
        got = walk_local_columns(q, ref, sc, nmax, vals, st)
        want = O.score(np.array(q, np.uint8), np.array(ref, np.uint8), sc, True)
        assert got == want, (sc, q, ref, nmax)
        assert walk_local_columns(q, ref, sc, nmax, vals, st, floor_on_chain=False) == want
        M, O_, E = bounds(sc)
        k = min(len(q), len(ref))
        lo, hi = -(O_ + 3 * E + M), k * M + E * (len(ref) + 2) + M
        assert lo <= vals.lo and vals.hi <= hi, (sc, vals.lo, vals.hi, lo, hi)


@pytest.mark.parametrize("seed,st", [(11, 16), (12, 24), (13, 32), (14, 32)])
def test_diagonal_shifted_semiglobal_walk_equals_oracle(seed, st):
    """The diagonal shift (semiglobal, rows and columns at once), on 16-,
    24- and 32-row stripes: scores equal the oracle's, and every value lies
    in [-(kM + |o| + M + (st + 1)|e|), kM + |e| (st + m - 2)], m the
    reference length: the range diagshift_bound(..., st) rests on."""
    for sc, q, ref, nmax in cases(seed, 200, n_max=4 * st):
        vals = Range()
        got = walk_semiglobal_diag(q, ref, sc, nmax, vals, st)
        want = O.score(np.array(q, np.uint8), np.array(ref, np.uint8), sc, False)
        assert got == want, (sc, q, ref, nmax, st)
        M, O_, E = bounds(sc)
        k = min(len(q), len(ref))
        lo, hi = -(k * M + O_ + M + (st + 1) * E), k * M + E * (st + len(ref) - 2)
        assert lo <= vals.lo and vals.hi <= hi, (sc, vals.lo, vals.hi, lo, hi, st)
        # diagshift_bound covers both ends
        assert max(-lo, hi) <= len(ref) * (M + E) + O_ + (st + 2) * E + M


@pytest.mark.parametrize("seed,st,K", [(31, 16, 8), (32, 24, 16), (33, 32, 8), (34, 32, 2)])
def test_diagonal_shift_rebased_every_k_columns_equals_oracle(seed, st, K):
    """The diagonal shift with its column part restarted every K columns:
    scores equal the oracle's on references several blocks long, and every
    value lies in [-(kM + |o| + M + (st + 1)|e|), kM + |e| (st + K - 2)]:
    the range diagshift_bound rests on with the block width in place of the
    reference length."""
    for sc, q, ref, nmax in cases(seed, 150, n_max=4 * st):
        ref = (ref * 3)[: len(ref) * 3]
        vals = Range()
        got = walk_semiglobal_diag(q, ref, sc, nmax, vals, st, rebase=K)
        want = O.score(np.array(q, np.uint8), np.array(ref, np.uint8), sc, False)
        assert got == want, (sc, q, ref, nmax, st, K)
        M, O_, E = bounds(sc)
        k = min(len(q), len(ref))
        lo, hi = -(k * M + O_ + M + (st + 1) * E), k * M + E * (st + K - 2)
        assert lo <= vals.lo and vals.hi <= hi, (sc, vals.lo, vals.hi, lo, hi, st, K)


def test_a_wrong_diagonal_top_border_is_caught():
    """The diagonal-shift model with the first stripe's top border left at
    column 0's value (not advanced by -e per column) disagrees with the
    oracle, so the equality above sees that mistake."""
    wrong = tried = 0
    for sc, q, ref, nmax in cases(15, 120):
        if sc[3] == 0 or len(ref) < 3:
            continue
        tried += 1
        got = walk_semiglobal_diag(q, ref, sc, nmax, Range(), 16, advance_top=False)
        want = O.score(np.array(q, np.uint8), np.array(ref, np.uint8), sc, False)
        wrong += got != want
    assert tried and wrong > 0


@pytest.mark.parametrize("seed,st,K", [(21, 16, 8), (22, 24, 16), (23, 32, 8), (24, 16, 2)])
def test_column_shift_rebased_every_k_columns_equals_oracle(seed, st, K):
    """The column shift restarted every K columns (local mode on panels
    whose references are too long for the shift's range): scores equal the
    oracle's on references several blocks long, and every value lies in
    [-(|o| + 3|e| + M), kM + |e| (K + 2) + M]: the range colshift_bound
    rests on with the block width in place of the reference length."""
    for sc, q, ref, nmax in cases(seed, 150, n_max=4 * st):
        ref = (ref * 3)[: len(ref) * 3]  # 3 to 72 columns: several blocks
        vals = Range()
        got = walk_local_columns(q, ref, sc, nmax, vals, st, rebase=K)
        want = O.score(np.array(q, np.uint8), np.array(ref, np.uint8), sc, True)
        assert got == want, (sc, q, ref, nmax, K)
        M, O_, E = bounds(sc)
        k = min(len(q), len(ref))
        lo, hi = -(O_ + 3 * E + M), k * M + E * (K + 2) + M
        assert lo <= vals.lo and vals.hi <= hi, (sc, vals.lo, vals.hi, lo, hi, K)


def test_a_missing_block_rebase_is_caught():
    """The rebased column shift with the best left in the old block's space
    at a block boundary (its + e K dropped) disagrees with the oracle."""
    wrong = tried = 0
    for sc, q, ref, nmax in cases(25, 120):
        if sc[3] == 0:
            continue
        ref = (ref * 3)[: len(ref) * 3]
        tried += 1
        want = O.score(np.array(q, np.uint8), np.array(ref, np.uint8), sc, True)
        assert walk_local_columns(q, ref, sc, nmax, Range(), 16, rebase=4) == want
        wrong += walk_local_columns_rebase_without_best(q, ref, sc, nmax, 4) != want
    assert tried and wrong > 0


def walk_local_columns_rebase_without_best(q, r, scoring, nmax, K):
    """walk_local_columns(rebase=K) with the best's block conversion left
    out (the mistake)."""
    ma, mi, o, e = scoring
    S, RR, T = plan(nmax, 16)
    rows = [None] * (T - len(q)) + list(q)
    m = len(r)
    best = o
    bottom = None
    row0 = 0
    for s in range(S):
        h = RR if s == 0 else 16
        Pv = [o + e] * h
        R = [NEG] * h
        hdt, vtop, z, b = o + e, o, e, best + e
        newb = []
        for j in range(m):
            if j and j % K == 0:
                Pv = [x + e * K for x in Pv]
                R = [add(x, e * K) for x in R]
                hdt += e * K
                vtop += e * K
                z += e * K
            if s == 0:
                pu, qq = vtop, NEG
                vtop -= e
            else:
                pu, qq = bottom[j]
            pd = hdt
            hdt = pu
            for k in range(h):
                sc = subst(rows[row0 + k], r[j], ma, mi) - o
                R[k] = mx(Pv[k], R[k], z)
                qq = add(mx(pu, qq), e)
                hh = mx(pd + sc, R[k], qq)
                pd = Pv[k]
                Pv[k] = hh + (o - e)
                pu = Pv[k]
            z -= e
            b = max(b - e, max(Pv))
            newb.append((pu, qq))
        best = b + e * ((m - 1) % K)
        bottom = newb
        row0 += h
    return best - o


def test_a_wrong_chain_floor_is_caught():
    """The local model with the chain's floor one extension low (H held at
    or above e instead of 0) disagrees with the oracle, so the equality
    above sees a floor in the wrong column's space."""
    wrong = tried = 0
    for sc, q, ref, nmax in cases(16, 120):
        if sc[3] == 0:
            continue
        tried += 1
        got = walk_local_columns(q, ref, sc, nmax, Range(), 16, floor_shift=1)
        want = O.score(np.array(q, np.uint8), np.array(ref, np.uint8), sc, True)
        wrong += got != want
    assert tried and wrong > 0


def test_a_wrong_rebase_is_caught():
    """The semiglobal model with the stripe rebase left out (c = 0: the
    shift restarted per stripe without moving the boundary into the new
    stripe's space) disagrees with the oracle on reads of several stripes,
    so the equality above sees the mistake it guards against."""
    wrong = tried = 0
    for sc, q, ref, nmax in cases(5, 80):
        if nmax <= 16 or sc[3] == 0:
            continue
        tried += 1
        got = walk_semiglobal_rows_without_rebase(q, ref, sc, nmax)
        want = O.score(np.array(q, np.uint8), np.array(ref, np.uint8), sc, False)
        wrong += got != want
    assert tried and wrong > 0


def walk_semiglobal_rows_without_rebase(q, r, scoring, nmax):
    """walk_semiglobal_rows with c forced to 0 (the mistake)."""
    ma, mi, o, e = scoring
    S, RR, T = plan(nmax)
    rows = [None] * (T - len(q)) + list(q)
    pad = T - len(q)
    m = len(r)
    best = None
    bottom = None
    row0 = 0
    for s in range(S):
        h = RR if s == 0 else 16
        Pv = [o - e * k for k in range(h)]
        R = [NEG] * h
        hdt = o + e
        bestL = NEG
        newb = []
        for j in range(m):
            pu, qq = (o + e, NEG) if s == 0 else bottom[j]
            pd = hdt
            hdt = pu
            for k in range(h):
                sc = subst(rows[row0 + k], r[j], ma, mi) - o
                R[k] = add(mx(Pv[k], R[k]), e)
                qq = mx(pu, qq)
                hh = mx(pd + sc, R[k], qq)
                pd = Pv[k]
                Pv[k] = hh + (o - e)
                pu = Pv[k]
            newb.append((pu, qq))
            bestL = mx(bestL, pu)
        for k in range(h):
            if row0 + k >= pad:
                best = mx(best, Pv[k] + e * k)
        if s == S - 1:
            best = mx(best, add(bestL, e * (h - 1)))
        bottom = newb
        row0 += h
    return best - o
