// gfx950 kernels of the winner traceback: the alignment (score, start and
// end coordinates, run-length operations) of given (read, panel reference)
// pairs, as the reference's simd::align_pairs / scalar::align compute it
// (lib/simd/mod.rs:405-499, lib/scalar.rs:164-364).
//
//   k_trace          the traceback DP and the walk, one pair per lane
//   k_trace_compact  the walked operations packed into the caller's order
//
// Layout. Pairs are sorted by (read length, reference length), longest
// first, on the host and cut into groups of 64; a 64-lane wave takes one
// group, one pair per lane. Each lane runs its own exact int32 Gotoh DP (the reference's
// recurrence and sentinel, lib/scalar.rs:120-163) over its read in stripes
// of 16 rows held in registers, walking its reference's columns (2-bit
// codes and wildcard bitmap, per-lane loads); the stripe's bottom row (H,
// F) per column goes to the group's boundary buffer for the next stripe.
//
// Trace. Every cell gets the reference's traceback code (lib/scalar.rs:
// 166-174: source diag/insertion/deletion/stop in 2 bits, E-extended and
// F-extended flags), 4 bits, written as 16 nibbles (8 bytes) per lane per
// column per stripe into the group's arena region:
//   trace[(s * M_g + j) * 64 + lane]   (s = stripe, j = column, M_g = the
//   group's longest reference), so a wave's stores are coalesced.
//
// End cell. The reference takes the first maximum of a column-major scan
// over the mode's candidate cells (local: every cell; semiglobal: the last
// row and the last column). The kernel visits cells stripe by stripe, so a
// cell replaces the current best when it is greater, or equal and in an
// earlier column: within one stripe and column cells come in row order, and
// an earlier column of a later stripe precedes in column-major order.
//
// Walk. The lane then follows its trace from the end cell exactly as
// scalar::walk does (lib/scalar.rs:292-364), writing run-length operations
// (op << 30 | length; 0 = M, 1 = I, 2 = D) backwards into the pair's
// worst-case region (read length + reference length entries), so the
// region's tail holds them in forward order.
#include "esch_internal.h"

namespace esch {

namespace {

constexpr int32_t kNeg = INT32_MIN / 4;  // lib/scalar.rs:41

__device__ __forceinline__ int32_t max3i(int32_t a, int32_t b, int32_t c) { return max(max(a, b), c); }

template <bool LOCAL>
__global__ __launch_bounds__(64) void k_trace(TraceArgs a) {
    const int lane = threadIdx.x;
    const int g = a.g0 + blockIdx.x;
    const int p = a.pair0 + g * 64 + lane;
    const bool active = p < a.n_pairs;
    const DevPanel &P = a.panel;
    const int n = active ? a.qlen[p] : 0;
    const int rf = active ? a.ref[p] : 0;
    const int m = active ? P.rlen[rf] : 0;
    const int S_g = a.gS[g], M_g = a.gM[g];
    uint8_t *region = a.arena + a.gbase[g];
    uint2 *trace = reinterpret_cast<uint2 *>(region);
    int2 *bnd = reinterpret_cast<int2 *>(region + (size_t)S_g * M_g * 512);
    const uint8_t *qc = a.qcodes + (active ? a.qoff[p] : 0);
    const uint32_t *w2 = P.ref2 + P.w2off[rf];
    const uint32_t *wn = P.refN + P.nmoff[rf];
    const int32_t mi = a.sc.mi, o = a.sc.o, e = a.sc.e, dmm = a.sc.ma - a.sc.mi;
    const int S = (n + 15) / 16;

    int32_t best = LOCAL ? 0 : kNeg;
    int bi = 0, bj = 0;
    for (int s = 0; s < S_g; ++s) {
        const bool srow = s < S;
        // per row, bit c set when reference code c scores a match against the
        // read base (lib/scoring.rs:105-111: equal, or either a wildcard);
        // rows past the read get 0 (their cells are never read back)
        uint32_t qm[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int i = 16 * s + r;
            const int q = (srow && i < n) ? (int)qc[i] : -1;
            qm[r] = q < 0 ? 0u : (q == 4 ? 31u : ((1u << q) | 16u));
        }
        int32_t H[16], E[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            H[r] = 0;  // H[i][0] = 0
            E[r] = kNeg;
        }
        int32_t top_prev = 0;  // H[16 s][j - 1]; the left border is 0
        uint32_t w = 0, wnw = 0;
        for (int cj = 0; cj < M_g; ++cj) {
            const bool col = srow && cj < m;
            if (col && (cj & 15) == 0) w = w2[cj >> 4];
            if (col && (cj & 31) == 0) wnw = wn[cj >> 5];
            const int c = ((wnw >> (cj & 31)) & 1u) ? 4 : (int)((w >> (2 * (cj & 15))) & 3u);
            int32_t hup = 0, f = kNeg;  // the row above the stripe: the border, or the previous stripe
            if (col && s > 0) {
                const int2 b = bnd[(size_t)cj * 64 + lane];
                hup = b.x;
                f = b.y;
            }
            int32_t hdiag = top_prev;
            top_prev = hup;
            const int j = cj + 1;
            uint32_t lo = 0, hi = 0;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int32_t hleft = H[r];
                const int32_t e_ext = E[r] + e;
                const int32_t ev = max(hleft + o, e_ext);
                const int32_t f_ext = f + e;
                f = max(hup + o, f_ext);
                const int32_t hd = hdiag + mi + (int32_t)((qm[r] >> c) & 1u) * dmm;
                int32_t h = max3i(hd, ev, f);
                if (LOCAL) h = max(h, 0);
                uint32_t t = h == f ? 2u : 3u;       // insertion : deletion
                t = h == hd ? 1u : t;                // diagonal first
                if (LOCAL) t = h == 0 ? 0u : t;      // stop
                t |= ev == e_ext ? 4u : 0u;          // E extended
                t |= f == f_ext ? 8u : 0u;           // F extended
                if (r < 8) lo |= t << (4 * r);
                else hi |= t << (4 * (r - 8));
                E[r] = ev;
                hdiag = hleft;
                H[r] = h;
                hup = h;
            }
            // Unconditional stores: a lane past its read or reference writes
            // only its own slots of the group's region, which its walk and
            // its next stripe never read (the walk reads i <= n, j <= m; the
            // next stripe reads the boundary only where col holds). Under a
            // branch, the compiler sinks every row's trace code into it and
            // holds the whole column's operands live; the stores also come
            // before the candidate branch below, for the same reason.
            trace[((size_t)s * M_g + cj) * 64 + lane] = make_uint2(lo, hi);
            bnd[(size_t)cj * 64 + lane] = make_int2(hup, f);
            // the column's candidates, in column-major order (rows of this
            // stripe in order; see the header for why ties take j < bj)
            const int nr = min(16, n - 16 * s);  // this stripe's real rows
            if (LOCAL) {
                int32_t cmax = kNeg;
#pragma unroll
                for (int r = 0; r < 16; ++r) cmax = max(cmax, r < nr ? H[r] : kNeg);
                if (col && (cmax > best || (cmax == best && j < bj))) {
                    int ci = 0;
#pragma unroll
                    for (int r = 15; r >= 0; --r) ci = (r < nr && H[r] == cmax) ? r : ci;
                    best = cmax;
                    bi = 16 * s + ci + 1;
                    bj = j;
                }
            } else if (col) {
                if (j == m) {
                    // the last column: every real row, in order
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        if (r < nr && H[r] > best) {
                            best = H[r];
                            bi = 16 * s + r + 1;
                            bj = j;
                        }
                    }
                } else if (s + 1 == S) {
                    // the last row (row n - 1 - 16 s of this stripe)
                    int32_t hn = kNeg;
#pragma unroll
                    for (int r = 0; r < 16; ++r) hn = r == nr - 1 ? H[r] : hn;
                    if (hn > best || (hn == best && j < bj)) {
                        best = hn;
                        bi = n;
                        bj = j;
                    }
                }
            }
        }
    }

    if (!active) return;
    int32_t *rec = a.recs + (size_t)a.slot[p] * kTraceRec;
    const int cap = n + m;
    uint32_t *ops = a.ops_wc + a.ops_off[a.slot[p]];
    int i = bi, j = bj, k = 0, wp = cap;
    if (bi != 0 && bj != 0) {
        int state = 0;  // 0 H, 1 F (insertion), 2 E (deletion)
        uint32_t cur = 3u, len = 0;
        while (i > 0 && j > 0) {
            const int s = (i - 1) >> 4, rr = (i - 1) & 15;
            const uint2 wv = trace[((size_t)s * M_g + (j - 1)) * 64 + lane];
            const uint32_t t = ((rr < 8 ? wv.x : wv.y) >> (4 * (rr & 7))) & 15u;
            uint32_t op = 3u;
            if (state == 0) {
                const uint32_t src = t & 3u;
                if (src == 0u) break;
                if (src == 1u) {
                    op = 0u;
                    --i;
                    --j;
                } else {
                    state = src == 2u ? 1 : 2;
                }
            } else if (state == 1) {
                op = 1u;
                if (!(t & 8u)) state = 0;
                --i;
            } else {
                op = 2u;
                if (!(t & 4u)) state = 0;
                --j;
            }
            if (op == 3u) continue;
            if (op == cur) {
                ++len;
            } else {
                if (len) {
                    ops[--wp] = cur << 30 | len;
                    ++k;
                }
                cur = op;
                len = 1;
            }
        }
        if (len) {
            ops[--wp] = cur << 30 | len;
            ++k;
        }
        rec[0] = best;
        rec[1] = i;
        rec[2] = bi;
        rec[3] = j;
        rec[4] = bj;
    } else {
        rec[0] = rec[1] = rec[2] = rec[3] = rec[4] = 0;
    }
    rec[5] = k;
    rec[6] = wp;
}

// ---------------------------------------------------------------------------
// k_trace_diag — the traceback of one long pair per wave, as an
// anti-diagonal wavefront (k_trace gives a pair one lane, so a group's time
// is its longest pair's, and a few long reads would set the whole call's).
//
// The read is cut into passes of 64 R rows (R = 2, 4 or 8,
// trace_diag_rows), top-padded to whole passes with rows of substitution 0
// (which reproduce the zero top border exactly for every real cell). Lane l
// holds rows l R .. l R + R - 1 of the pass in registers and processes
// column j at step t = j + l: the row above (H, F of lane l - 1's bottom row
// at column j, one step earlier) comes by a `wave_shr:1` DPP move; lane 0
// takes it from the previous pass's bottom row, which lane 63 kept in LDS
// and which the wave reads back one 64-column block ahead (as k_diag64).
// Every step each lane stores its R trace codes (the nibbles of k_trace) as
// one dword at trace[(pass (m + 63) + t) 64 + l], so a step's stores are one
// coalesced 256-byte line.
//
// End cell: each lane keeps the best of its candidate cells by (score
// descending, column ascending, row ascending), the order of the
// reference's first maximum in a column-major scan; a wave reduction in the
// same order picks the cell. Lane 0 then walks the codes exactly as
// k_trace's lanes do.
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool cell_before(int32_t h, int j, int i, int32_t bh, int bj, int bi) {
    return h > bh || (h == bh && (j < bj || (j == bj && i < bi)));
}

template <int R, bool LOCAL>
__device__ void trace_diag_task(const TraceArgs &a, const int k, uint2 *__restrict__ lb, const int lane) {
    const DevPanel &P = a.panel;
    const int n = a.qlen[k];
    const int rf = a.ref[k];
    const int m = P.rlen[rf];
    const uint8_t *qc = a.qcodes + a.qoff[k];
    uint32_t *trace = reinterpret_cast<uint32_t *>(a.arena + a.dbase[k]);
    const uint32_t *w2 = P.ref2 + P.w2off[rf];
    const uint32_t *wn = P.refN + P.nmoff[rf];
    const int32_t mi = a.sc.mi, o = a.sc.o, e = a.sc.e, dmm = a.sc.ma - a.sc.mi;
    constexpr int RP = 64 * R;
    const int Pn = (n + RP - 1) / RP;
    const int pre = Pn * RP - n;   // padding rows above the read
    const size_t steps = (size_t)m + 63;

    int32_t bh = LOCAL ? 0 : kNeg;
    int bi = 0, bj = 0;            // 1-based real coordinates of the lane's best (0: none)
    for (int p = 0; p < Pn; ++p) {
        // the lane's rows: padded index p RP + lane R + r; real row (1-based)
        // ip - pre + 1 when ip >= pre
        uint32_t qm[R];
        uint32_t padbits = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int ip = p * RP + lane * R + r;
            const bool pad = ip < pre;
            padbits |= (pad ? 1u : 0u) << r;
            const int q = pad ? 0 : (int)qc[ip - pre];
            qm[r] = pad ? 0u : (q == 4 ? 31u : ((1u << q) | 16u));
        }
        const int row1 = p * RP + lane * R - pre + 1;  // real row of register 0
        int32_t H[R], E[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            H[r] = 0;    // left border
            E[r] = kNeg;
        }
        const bool first = p == 0, lastpass = p + 1 == Pn;
        int32_t bot_h = 0, bot_f = kNeg;  // this lane's bottom row, last step
        int32_t in_prev = 0;              // H above-left (column -1: border)
        uint32_t w = 0, wnw = 0;
        uint2 cur = make_uint2(0u, (uint32_t)kNeg), nxt = cur;
        if (!first) {
            cur = lb[lane];
            nxt = lb[64 + lane];
        }
        uint32_t *tp = trace + (size_t)p * steps * 64 + lane;
        for (int t = 0; t < m + 63; ++t) {
            if ((t & 63) == 0 && t > 0 && !first) {
                cur = nxt;
                nxt = lb[t + 64 + lane];  // still the previous pass's (this pass writes t - 63)
            }
            int32_t h0 = 0, f0 = kNeg;    // lane 0's row above: border or the previous pass
            if (!first && t < m) {
                h0 = __builtin_amdgcn_readlane((int)cur.x, t & 63);
                f0 = __builtin_amdgcn_readlane((int)cur.y, t & 63);
            }
            const int32_t hin = __builtin_amdgcn_update_dpp(h0, bot_h, 0x138, 0xF, 0xF, false);
            const int32_t fin = __builtin_amdgcn_update_dpp(f0, bot_f, 0x138, 0xF, 0xF, false);
            const int32_t hdiag0 = in_prev;
            in_prev = hin;
            const int j = t - lane;   // 0-based column
            const bool col = j >= 0 && j < m;
            if (col && (j & 15) == 0) w = w2[j >> 4];
            if (col && (j & 31) == 0) wnw = wn[j >> 5];
            const int c = ((wnw >> (j & 31)) & 1u) ? 4 : (int)((w >> (2 * (j & 15))) & 3u);
            int32_t hup = hin, f = fin, hdiag = hdiag0;
            uint32_t word = 0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int32_t hleft = H[r];
                const int32_t e_ext = E[r] + e;
                const int32_t ev = max(hleft + o, e_ext);
                const int32_t f_ext = f + e;
                f = max(hup + o, f_ext);
                // padding rows score 0 against every base (their H stays 0)
                const int32_t sub = ((padbits >> r) & 1u) ? 0 : mi + (int32_t)((qm[r] >> c) & 1u) * dmm;
                const int32_t hd = hdiag + sub;
                int32_t h = max(max(hd, ev), f);
                if (LOCAL) h = max(h, 0);
                uint32_t tc = h == f ? 2u : 3u;
                tc = h == hd ? 1u : tc;
                if (LOCAL) tc = h == 0 ? 0u : tc;
                tc |= ev == e_ext ? 4u : 0u;
                tc |= f == f_ext ? 8u : 0u;
                word |= tc << (4 * r);
                if (col) {
                    E[r] = ev;
                    H[r] = h;
                }
                hdiag = hleft;
                hup = h;
            }
            tp[(size_t)t * 64] = word;  // unconditional: a step's line is whole (slots off the matrix are never read)
            if (col) {
                bot_h = hup;
                bot_f = f;
                if (lane == 63 && !lastpass) lb[j] = make_uint2((uint32_t)bot_h, (uint32_t)bot_f);
                // candidate cells, rows in order (their order within a column)
                if (LOCAL) {
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (!((padbits >> r) & 1u) && cell_before(H[r], j + 1, row1 + r, bh, bj, bi)) {
                            bh = H[r];
                            bj = j + 1;
                            bi = row1 + r;
                        }
                } else {
                    if (j == m - 1) {
#pragma unroll
                        for (int r = 0; r < R; ++r)
                            if (!((padbits >> r) & 1u) && cell_before(H[r], j + 1, row1 + r, bh, bj, bi)) {
                                bh = H[r];
                                bj = j + 1;
                                bi = row1 + r;
                            }
                    } else if (lastpass && lane == 63 && cell_before(H[R - 1], j + 1, n, bh, bj, bi)) {
                        bh = H[R - 1];  // the read's last row
                        bj = j + 1;
                        bi = n;
                    }
                }
            }
        }
    }
    // the wave's first maximum in column-major order
    for (int off = 32; off > 0; off >>= 1) {
        const int32_t oh = __shfl_xor(bh, off, 64);
        const int oj = __shfl_xor(bj, off, 64), oi = __shfl_xor(bi, off, 64);
        if (oj != 0 && (bj == 0 ? (LOCAL ? oh > bh : true) : cell_before(oh, oj, oi, bh, bj, bi))) {
            bh = oh;
            bj = oj;
            bi = oi;
        }
    }
    // every lane's codes stored before lane 0 reads them
    __threadfence_block();
    __syncthreads();
    if (lane != 0) return;
    int32_t *rec = a.recs + (size_t)a.slot[k] * kTraceRec;
    uint32_t *ops = a.ops_wc + a.ops_off[a.slot[k]];
    const int cap = n + m;
    int i = bi, j = bj, nk = 0, wp = cap;
    if (bi != 0 && bj != 0) {
        int state = 0;  // 0 H, 1 F (insertion), 2 E (deletion)
        uint32_t cur = 3u, len = 0;
        while (i > 0 && j > 0) {
            const int ip = i - 1 + pre;
            const int pp = ip / RP, l = (ip - pp * RP) / R, r = ip % R;
            const uint32_t word = trace[((size_t)pp * steps + (size_t)(j - 1 + l)) * 64 + l];
            const uint32_t tc = (word >> (4 * r)) & 15u;
            uint32_t op = 3u;
            if (state == 0) {
                const uint32_t src = tc & 3u;
                if (src == 0u) break;
                if (src == 1u) {
                    op = 0u;
                    --i;
                    --j;
                } else {
                    state = src == 2u ? 1 : 2;
                }
            } else if (state == 1) {
                op = 1u;
                if (!(tc & 8u)) state = 0;
                --i;
            } else {
                op = 2u;
                if (!(tc & 4u)) state = 0;
                --j;
            }
            if (op == 3u) continue;
            if (op == cur) {
                ++len;
            } else {
                if (len) {
                    ops[--wp] = cur << 30 | len;
                    ++nk;
                }
                cur = op;
                len = 1;
            }
        }
        if (len) {
            ops[--wp] = cur << 30 | len;
            ++nk;
        }
        rec[0] = bh;
        rec[1] = i;
        rec[2] = bi;
        rec[3] = j;
        rec[4] = bj;
    } else {
        rec[0] = rec[1] = rec[2] = rec[3] = rec[4] = 0;
    }
    rec[5] = nk;
    rec[6] = wp;
}

// One block (one wave) per long pair; dynamic LDS: (longest reference of
// the launch + kTraceDiagPad) columns of 8 B.
template <bool LOCAL>
__global__ __launch_bounds__(64) void k_trace_diag(TraceArgs a) {
    extern __shared__ uint2 lb[];
    const int k = a.d0 + blockIdx.x;
    const int R = trace_diag_rows(a.qlen[k]);
    if (R == 2) trace_diag_task<2, LOCAL>(a, k, lb, threadIdx.x);
    else if (R == 4) trace_diag_task<4, LOCAL>(a, k, lb, threadIdx.x);
    else trace_diag_task<8, LOCAL>(a, k, lb, threadIdx.x);
}

// One thread per pair: its runs (the tail of its worst-case region) to the
// caller's packed offset.
__global__ void k_trace_compact(const uint32_t *ops_wc, const uint64_t *ops_off, const int32_t *recs,
                                const uint64_t *out_off, uint32_t *out, int32_t n_pairs) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pairs) return;
    const int32_t *rec = recs + (size_t)p * kTraceRec;
    const uint32_t *src = ops_wc + ops_off[p] + rec[6];
    uint32_t *dst = out + out_off[p];
    for (int k = 0; k < rec[5]; ++k) dst[k] = src[k];
}

}  // namespace

hipError_t launch_trace(const TraceArgs &a, int groups, bool local, hipStream_t st) {
    if (local) hipLaunchKernelGGL(k_trace<true>, dim3(groups), dim3(64), 0, st, a);
    else hipLaunchKernelGGL(k_trace<false>, dim3(groups), dim3(64), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_trace_diag(const TraceArgs &a, int pairs, int max_ref, bool local, hipStream_t st) {
    if (pairs <= 0) return hipSuccess;
    const size_t lds = (size_t)(max_ref + kTraceDiagPad) * sizeof(uint2);
    if (local) hipLaunchKernelGGL(k_trace_diag<true>, dim3(pairs), dim3(64), lds, st, a);
    else hipLaunchKernelGGL(k_trace_diag<false>, dim3(pairs), dim3(64), lds, st, a);
    return hipGetLastError();
}

hipError_t launch_trace_compact(const uint32_t *ops_wc, const uint64_t *ops_off, const int32_t *recs,
                                const uint64_t *out_off, uint32_t *out, int32_t n_pairs, hipStream_t st) {
    if (n_pairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_trace_compact, dim3((n_pairs + 255) / 256), dim3(256), 0, st, ops_wc, ops_off, recs,
                       out_off, out, n_pairs);
    return hipGetLastError();
}

}  // namespace esch
