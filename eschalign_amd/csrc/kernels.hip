// gfx950 kernels of the panel-scoring path.
//
//   k_dp16   batched packed-int16 score DP (the hot path)
//   k_dp32   plain one-thread-per-pair int32 score DP (baseline + rescue)
//   k_zero, k_revcomp, k_classify, k_scan, k_scatter   length bucketing
//   k_select, k_pairs   fused best / tie reductions
//
// The recurrence is eschalign's score-only Gotoh DP (lib/scalar.rs:120-163):
//   E[i][j] = max(H[i][j-1] + o, E[i][j-1] + e)
//   F[i][j] = max(H[i-1][j] + o, F[i-1][j] + e)
//   H[i][j] = max(H[i-1][j-1] + s(q_i, r_j), E, F [, 0 local])
// with zero borders and E/F = -inf on them; local mode takes the best cell
// anywhere, semiglobal the best of the last row and the last column.
#include "esch_internal.h"
#include "dp16.h"

namespace esch {

// ---------------------------------------------------------------------------
// k_diag16 — anti-diagonal (intra-task) packed-int16 DP for long reads.
//
// For reads too few per length to fill a 128-read chunk. A row of 16 lanes
// (one DPP row) scores ONE read against TWO references packed in the int16
// halves; a 64-lane wave runs four such tasks over 8 length-adjacent
// references of the same read. Lane l of a row holds R consecutive rows of
// the read in registers (a pass covers 16 R rows; longer reads take several
// passes, the last row of a pass handed to the next through scratch) and
// processes column j at step j + l: the wavefront sweeps the anti-diagonal.
// At every step each lane takes from lane l - 1, with a `row_shr:1` DPP
// move, the bottom row's (Ho, F) of the column it now computes and that
// column's reference-code selector; lane 0 takes the previous pass's
// boundary (or the border) and the next column's codes as the DPP `old`
// operand. The substitution is one v_perm_b32 per row: the row's profile
// s(q_r, c) - o for c = 0..3 sits in two registers as four int16, and the
// selector picks code cA's int16 into the low half and cB's into the high.
// A wildcard reference base (rare) takes a slow path. The recurrence and
// the Ho = H + o bookkeeping are k_dp16's: 9 packed VALU ops per row per
// two cells.
// ---------------------------------------------------------------------------

constexpr int kDiagRefs = 8;  // references per work item: 4 rows x 2 halves

__device__ __forceinline__ uint32_t half_sel(int c) {
    // v_perm selector bytes of int16 entry c (0..3) of {tab1:tab0}; 0x0C0C = zero
    return c < 4 ? (uint32_t)(2 * c) | ((uint32_t)(2 * c + 1) << 8) : 0x0C0Cu;
}

template <int R, bool LOCAL>
__global__ __launch_bounds__(64) void k_diag16(KArgs a) {
    const int lane = threadIdx.x;
    const int grp = lane >> 4;
    const int l = lane & 15;
    const DevPanel &P = a.panel;
    const int n_refs = P.n_refs;
    const int n_rg = (n_refs + kDiagRefs - 1) / kDiagRefs;
    const int nd = uni(a.desc->nd);
    const uint32_t n_items = (uint32_t)nd * (uint32_t)n_rg;
    if (n_items == 0) return;  // nothing routed here (uniform: before any barrier)
    uint2 *scr = a.scratchD + ((size_t)blockIdx.x * 4 + grp) * (size_t)P.max_len;
    const v2s vo = splat(a.sc.o), ve = splat(a.sc.e), NEG = splat(-32768);
    const int nval_real = a.sc.ma - a.sc.o, nval_pre = -a.sc.o;
    constexpr int RP = 16 * R;

    __shared__ uint32_t s_item;
    for (;;) {
        // the item index through LDS between barriers (see k_diag64)
        __syncthreads();
        if (lane == 0) s_item = atomicAdd(&a.desc->nextD, 1u);
        __syncthreads();
        const uint32_t t0 = (uint32_t)uni((int)s_item);
        if (t0 >= n_items) break;
        const int qi = (int)(t0 / (uint32_t)n_rg);
        const int g = (int)(t0 - (uint32_t)qi * (uint32_t)n_rg);
        const int q = uni(a.listD[qi]);
        const int n = uni(a.qlen[q]);
        const uint8_t *qc = a.qcodes + a.qoff[q];
        const int slotA = g * kDiagRefs + 2 * grp, slotB = slotA + 1;
        const int refA = slotA < n_refs ? P.rorder[slotA] : -1;
        const int refB = slotB < n_refs ? P.rorder[slotB] : -1;
        const int mA = refA >= 0 ? P.rlen[refA] : 0;
        const int mB = refB >= 0 ? P.rlen[refB] : 0;
        // longest first: the item's first reference is its longest
        const int M = uni(P.rlen[P.rorder[g * kDiagRefs]]);
        const uint32_t *w2A = P.ref2 + (refA >= 0 ? P.w2off[refA] : 0);
        const uint32_t *wnA = P.refN + (refA >= 0 ? P.nmoff[refA] : 0);
        const uint32_t *w2B = P.ref2 + (refB >= 0 ? P.w2off[refB] : 0);
        const uint32_t *wnB = P.refN + (refB >= 0 ? P.nmoff[refB] : 0);
        const int Pn = (n + RP - 1) / RP;
        const int pre = Pn * RP - n;
        v2s best = LOCAL ? vo : NEG;  // Ho space

        for (int p = 0; p < Pn; ++p) {
            // this pass's rows of lane l: p RP + l R + r
            uint32_t tab0[R], tab1[R];
            uint32_t prebits = 0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = p * RP + l * R + r;
                const int x = i < pre ? (int)kPreCode : (int)qc[i - pre];
                prebits |= (i < pre ? 1u : 0u) << r;
                const int v0 = prof_val(x, 0, a.sc), v1 = prof_val(x, 1, a.sc);
                const int v2 = prof_val(x, 2, a.sc), v3 = prof_val(x, 3, a.sc);
                tab0[r] = pack2(v0, v1);
                tab1[r] = pack2(v2, v3);
            }
            v2s Ho[R], E[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                Ho[r] = vo;
                E[r] = NEG;
            }
            const bool lastpass = p + 1 == Pn;
            uint32_t bot_h = as_u32(vo), bot_f = as_u32(NEG), sel_prev = 0, nb_prev = 0;
            uint32_t in_prev = as_u32(vo);  // Ho above-left (column -1: border)
            uint32_t wA = 0, wB = 0, nA = 0, nB = 0;
            for (int t = 0; t < M + 15; ++t) {
                // lane 0's inputs for column t: codes of both references
                // each reference's words are read only inside its own length
                if ((t & 15) == 0) {
                    if (t < mA) wA = w2A[t >> 4];
                    if (t < mB) wB = w2B[t >> 4];
                }
                if ((t & 31) == 0) {
                    if (t < mA) nA = wnA[t >> 5];
                    if (t < mB) nB = wnB[t >> 5];
                }
                uint32_t sel0 = 0x0C0C0C0Cu, nb0 = 0;
                if (t < M) {
                    const int cA = t < mA ? (int)((wA >> (2 * (t & 15))) & 3u) : 4;
                    const int cB = t < mB ? (int)((wB >> (2 * (t & 15))) & 3u) : 4;
                    const bool NA = t < mA && ((nA >> (t & 31)) & 1u);
                    const bool NB = t < mB && ((nB >> (t & 31)) & 1u);
                    sel0 = half_sel(NA ? 4 : cA) | (half_sel(NB ? 4 : cB) << 16);
                    nb0 = (NA ? 1u : 0u) | (NB ? 2u : 0u);
                }
                uint2 b0 = make_uint2(as_u32(vo), as_u32(NEG));
                if (p > 0 && t < M) b0 = scr[t];
                // shift the wavefront one lane down the row
                const uint32_t hin = (uint32_t)__builtin_amdgcn_update_dpp((int)b0.x, (int)bot_h, 0x111, 0xF, 0xF, false);
                const uint32_t fin = (uint32_t)__builtin_amdgcn_update_dpp((int)b0.y, (int)bot_f, 0x111, 0xF, 0xF, false);
                const uint32_t sel = (uint32_t)__builtin_amdgcn_update_dpp((int)sel0, (int)sel_prev, 0x111, 0xF, 0xF, false);
                const uint32_t nb = (uint32_t)__builtin_amdgcn_update_dpp((int)nb0, (int)nb_prev, 0x111, 0xF, 0xF, false);
                const uint32_t hdiag = in_prev;
                in_prev = hin;
                sel_prev = sel;
                nb_prev = nb;
                const int j = t - l;
                if (j >= 0 && j < M) {
                    v2s sc[R];
#pragma unroll
                    for (int r = 0; r < R; ++r) sc[r] = as_v2s(__builtin_amdgcn_perm(tab1[r], tab0[r], sel));
                    if (nb) {
                        // wildcard reference base: a match for every real row
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            const short nv = (short)(((prebits >> r) & 1u) ? nval_pre : nval_real);
                            if (nb & 1u) sc[r].x = nv;
                            if (nb & 2u) sc[r].y = nv;
                        }
                    }
                    v2s hou = as_v2s(hin), f = as_v2s(fin);
                    column16<R, LOCAL>(sc, Ho, E, hou, f, as_v2s(hdiag), vo, ve, best);
                    bot_h = as_u32(hou);
                    bot_f = as_u32(f);
                    if (l == 15 && !lastpass) scr[j] = make_uint2(bot_h, bot_f);
                    if (!LOCAL) {
                        if (lastpass && l == 15) {
                            // the read's last row (register R - 1 of lane 15, last pass)
                            v2s v = hou;
                            if (j >= mA) v.x = (short)-32768;
                            if (j >= mB) v.y = (short)-32768;
                            best = vmax(best, v);
                        }
                        if (j == mA - 1 || j == mB - 1) {
                            // a reference's last column: every real row
#pragma unroll
                            for (int r = 0; r < R; ++r) {
                                v2s v = Ho[r];
                                const bool real = !((prebits >> r) & 1u);
                                if (!(real && j == mA - 1)) v.x = (short)-32768;
                                if (!(real && j == mB - 1)) v.y = (short)-32768;
                                best = vmax(best, v);
                            }
                        }
                    }
                }
            }
        }
        // the row's best over its 16 lanes
        int bx = best.x, by = best.y;
        for (int off = 8; off > 0; off >>= 1) {
            bx = max(bx, __shfl_xor(bx, off, 16));
            by = max(by, __shfl_xor(by, off, 16));
        }
        if (l == 0) {
            if (refA >= 0) {
                const size_t o = out_index(a, q, refA);
                if (a.out32) a.out32[o] = bx - a.sc.o;
                if (a.out16) a.out16[o] = (int16_t)(bx - a.sc.o);
            }
            if (refB >= 0) {
                const size_t o = out_index(a, q, refB);
                if (a.out32) a.out32[o] = by - a.sc.o;
                if (a.out16) a.out16[o] = (int16_t)(by - a.sc.o);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// k_diag64 — the anti-diagonal DP with the whole wave as one row.
//
// One wave scores one read against TWO references (int16 halves). The 64
// lanes form a single anti-diagonal: lane l holds R consecutive rows of the
// pass (64 R rows: 128, 256 or 512 by the read's length) and processes
// column j at step j + l, taking the bottom row (Ho, F) of lane l - 1 and
// the column's reference selector with a `wave_shr:1` DPP move. Compared
// with k_diag16 (four 16-lane rows per wave, one task each) a task needs a
// quarter of the passes, which is what sets the latency of a few long
// reads. The pass boundary (the pass's bottom row per column) stays in LDS:
// lane 63 writes column j at step j + 63, and lane 0 reads it in the next
// pass from a 64-column block that the whole wave loads one block ahead
// (lane c holds column 64 b + c; v_readlane picks the step's column and
// v_writelane puts it into lane 0). Columns are written and read in
// program order by the one wave, so the single buffer is updated in place.
// ---------------------------------------------------------------------------

template <int R, bool LOCAL>
__device__ __forceinline__ void diag64_task(const KArgs &a, const int q, const int n, const int refA,
                                            const int refB, uint2 *__restrict__ lb, const int lane) {
    const DevPanel &P = a.panel;
    const v2s vo = splat(a.sc.o), ve = splat(a.sc.e), NEG = splat(-32768);
    const int nval_real = a.sc.ma - a.sc.o, nval_pre = -a.sc.o;
    constexpr int RP = 64 * R;
    const int mA = refA >= 0 ? uni(P.rlen[refA]) : 0;
    const int mB = refB >= 0 ? uni(P.rlen[refB]) : 0;
    const int M = max(mA, mB);
    const uint8_t *qc = a.qcodes + a.qoff[q];
    const cwords w2A = (cwords)(P.ref2 + (refA >= 0 ? uni((int)P.w2off[refA]) : 0));
    const cwords wnA = (cwords)(P.refN + (refA >= 0 ? uni((int)P.nmoff[refA]) : 0));
    const cwords w2B = (cwords)(P.ref2 + (refB >= 0 ? uni((int)P.w2off[refB]) : 0));
    const cwords wnB = (cwords)(P.refN + (refB >= 0 ? uni((int)P.nmoff[refB]) : 0));
    const int Pn = (n + RP - 1) / RP;
    const int pre = Pn * RP - n;
    v2s best = LOCAL ? vo : NEG;  // Ho space
    const uint32_t bh0 = as_u32(vo), bf0 = as_u32(NEG);

    for (int p = 0; p < Pn; ++p) {
        uint32_t tab0[R], tab1[R];
        uint32_t prebits = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = p * RP + lane * R + r;
            const int x = i < pre ? (int)kPreCode : (int)qc[i - pre];
            prebits |= (i < pre ? 1u : 0u) << r;
            tab0[r] = pack2(prof_val(x, 0, a.sc), prof_val(x, 1, a.sc));
            tab1[r] = pack2(prof_val(x, 2, a.sc), prof_val(x, 3, a.sc));
        }
        v2s Ho[R], E[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            Ho[r] = vo;
            E[r] = NEG;
        }
        const bool first = p == 0, lastpass = p + 1 == Pn;
        uint32_t bot_h = bh0, bot_f = bf0, sel_prev = 0, nb_prev = 0;
        uint32_t in_prev = bh0;  // Ho above-left (column -1: border)
        uint32_t wA = 0, wB = 0, nA = 0, nB = 0;
        // the previous pass's boundary, columns 64 b + lane (b = current block)
        uint2 cur = make_uint2(bh0, bf0), nxt = make_uint2(bh0, bf0);
        if (!first) {
            cur = lb[lane];
            nxt = lb[64 + lane];
        }
        for (int t = 0; t < M + 63; ++t) {
            if ((t & 63) == 0 && t > 0 && !first) {
                cur = nxt;
                nxt = lb[t + 64 + lane];  // written by the previous pass only (this pass is at t - 63)
            }
            if ((t & 15) == 0) {
                if (t < mA) wA = w2A[t >> 4];
                if (t < mB) wB = w2B[t >> 4];
            }
            if ((t & 31) == 0) {
                if (t < mA) nA = wnA[t >> 5];
                if (t < mB) nB = wnB[t >> 5];
            }
            uint32_t sel0 = 0x0C0C0C0Cu, nb0 = 0;
            if (t < M) {
                const int cA = t < mA ? (int)((wA >> (2 * (t & 15))) & 3u) : 4;
                const int cB = t < mB ? (int)((wB >> (2 * (t & 15))) & 3u) : 4;
                const bool NA = t < mA && ((nA >> (t & 31)) & 1u);
                const bool NB = t < mB && ((nB >> (t & 31)) & 1u);
                sel0 = half_sel(NA ? 4 : cA) | (half_sel(NB ? 4 : cB) << 16);
                nb0 = (NA ? 1u : 0u) | (NB ? 2u : 0u);
            }
            // lane 0's boundary for column t: the block's lane t & 63
            uint32_t h0 = bh0, f0 = bf0;
            if (!first && t < M) {
                h0 = (uint32_t)__builtin_amdgcn_readlane((int)cur.x, t & 63);
                f0 = (uint32_t)__builtin_amdgcn_readlane((int)cur.y, t & 63);
            }
            // shift the wavefront one lane down the wave; lane 0 keeps `old`
            const uint32_t hin = (uint32_t)__builtin_amdgcn_update_dpp((int)h0, (int)bot_h, 0x138, 0xF, 0xF, false);
            const uint32_t fin = (uint32_t)__builtin_amdgcn_update_dpp((int)f0, (int)bot_f, 0x138, 0xF, 0xF, false);
            const uint32_t sel = (uint32_t)__builtin_amdgcn_update_dpp((int)sel0, (int)sel_prev, 0x138, 0xF, 0xF, false);
            const uint32_t nb = (uint32_t)__builtin_amdgcn_update_dpp((int)nb0, (int)nb_prev, 0x138, 0xF, 0xF, false);
            const uint32_t hdiag = in_prev;
            in_prev = hin;
            sel_prev = sel;
            nb_prev = nb;
            const int j = t - lane;
            if (j >= 0 && j < M) {
                v2s sc[R];
#pragma unroll
                for (int r = 0; r < R; ++r) sc[r] = as_v2s(__builtin_amdgcn_perm(tab1[r], tab0[r], sel));
                if (nb) {
                    // wildcard reference base: a match for every real row
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const short nv = (short)(((prebits >> r) & 1u) ? nval_pre : nval_real);
                        if (nb & 1u) sc[r].x = nv;
                        if (nb & 2u) sc[r].y = nv;
                    }
                }
                v2s hou = as_v2s(hin), f = as_v2s(fin);
                column16<R, LOCAL>(sc, Ho, E, hou, f, as_v2s(hdiag), vo, ve, best);
                bot_h = as_u32(hou);
                bot_f = as_u32(f);
                if (lane == 63 && !lastpass) lb[j] = make_uint2(bot_h, bot_f);
                if (!LOCAL) {
                    if (lastpass && lane == 63) {
                        // the read's last row (register R - 1 of lane 63, last pass)
                        v2s v = hou;
                        if (j >= mA) v.x = (short)-32768;
                        if (j >= mB) v.y = (short)-32768;
                        best = vmax(best, v);
                    }
                    if (j == mA - 1 || j == mB - 1) {
                        // a reference's last column: every real row
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            v2s v = Ho[r];
                            const bool real = !((prebits >> r) & 1u);
                            if (!(real && j == mA - 1)) v.x = (short)-32768;
                            if (!(real && j == mB - 1)) v.y = (short)-32768;
                            best = vmax(best, v);
                        }
                    }
                }
            }
        }
    }
    // the task's best over the wave
    int bx = best.x, by = best.y;
    for (int off = 32; off > 0; off >>= 1) {
        bx = max(bx, __shfl_xor(bx, off, 64));
        by = max(by, __shfl_xor(by, off, 64));
    }
    if (lane == 0) {
        if (refA >= 0) {
            const size_t o = out_index(a, q, refA);
            if (a.out32) a.out32[o] = bx - a.sc.o;
            if (a.out16) a.out16[o] = (int16_t)(bx - a.sc.o);
        }
        if (refB >= 0) {
            const size_t o = out_index(a, q, refB);
            if (a.out32) a.out32[o] = by - a.sc.o;
            if (a.out16) a.out16[o] = (int16_t)(by - a.sc.o);
        }
    }
}

// Work item t = (read qi of listD, pair g of length-adjacent references);
// dynamic LDS: (max_len + kDiag64Pad) columns of 8 B.
template <bool LOCAL>
__global__ __launch_bounds__(64) void k_diag64(KArgs a) {
    extern __shared__ uint2 lb[];
    __shared__ uint32_t s_item;
    const int lane = threadIdx.x;
    const DevPanel &P = a.panel;
    const int n_refs = P.n_refs;
    const int n_pairs = (n_refs + 1) / 2;
    const int nd = uni(a.desc->nd);
    const uint32_t n_items = (uint32_t)nd * (uint32_t)n_pairs;
    if (n_items == 0) return;  // nothing routed here (uniform: before any barrier)
    for (;;) {
        // the item index goes through LDS between two barriers (as in
        // k_dp16): every lane reads it after lane 0 fetched it, whatever
        // the compiler does with the lane-0 branches around the loop edge
        __syncthreads();
        if (lane == 0) s_item = atomicAdd(&a.desc->nextD, 1u);
        __syncthreads();
        const uint32_t t0 = (uint32_t)uni((int)s_item);
        if (t0 >= n_items) break;
        const int qi = (int)(t0 / (uint32_t)n_pairs);
        const int g = (int)(t0 - (uint32_t)qi * (uint32_t)n_pairs);
        const int q = uni(a.listD[qi]);
        const int n = uni(a.qlen[q]);
        const int refA = uni(P.rorder[2 * g]);
        const int refB = 2 * g + 1 < n_refs ? uni(P.rorder[2 * g + 1]) : -1;
        if (n <= 128) diag64_task<2, LOCAL>(a, q, n, refA, refB, lb, lane);
        else if (n <= 256) diag64_task<4, LOCAL>(a, q, n, refA, refB, lb, lane);
        else diag64_task<8, LOCAL>(a, q, n, refA, refB, lb, lane);
    }
}

// ---------------------------------------------------------------------------
// k_dp32 — plain one-thread-per-pair int32 DP (the baseline kernel, and the
// exact rescue for queries outside the int16 bound).
//
// A block scores one read against up to blockDim.x references, one per
// thread; the read is uniform across the block (bottom-aligned in 16-row
// register stripes exactly as above), each thread walks its own reference
// to its own length. The substitution is the 4-bit mask intersection
// (A=1, C=2, G=4, T=8, wildcard 0xF) with padding rows scoring 0. The
// stripe boundary (H, F) lives in scratch[j * total_threads + thread].
// ---------------------------------------------------------------------------

template <bool LOCAL>
__global__ __launch_bounds__(256) void k_dp32(KArgs a, int which) {
    const DevPanel &P = a.panel;
    const int n_refs = P.n_refs;
    const int G = blockDim.x;
    const int groups = (n_refs + G - 1) / G;
    const int count = uni(which == 0 ? a.desc->n16 : a.desc->n32);
    const int *list = which == 0 ? a.sorted16 : a.list32;
    uint32_t *cursor = which == 0 ? &a.desc->next32a : &a.desc->next32b;
    const uint32_t n_items = (uint32_t)count * (uint32_t)groups;
    if (n_items == 0) return;  // an empty list (uniform: before any barrier)
    const size_t total = (size_t)gridDim.x * G;
    const size_t gtid = (size_t)blockIdx.x * G + threadIdx.x;
    int2 *scr = a.scratch32 + gtid;
    const int32_t ma = a.sc.ma, mi = a.sc.mi, o = a.sc.o, e = a.sc.e;
    __shared__ uint32_t s_item;

    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_item = atomicAdd(cursor, 1u);
        __syncthreads();
        const uint32_t t = s_item;
        if (t >= n_items) break;
        const int qi = (int)(t / (uint32_t)groups);
        const int g = (int)(t - (uint32_t)qi * (uint32_t)groups);
        const int q = list[qi];
        const int n = a.qlen[q];
        const int S = (n + kStripe - 1) / kStripe;
        const int p = kStripe * S - n;
        const uint8_t *qc = a.qcodes + a.qoff[q];
        const int ref = g * G + threadIdx.x;
        if (ref >= n_refs) continue;
        const int m = P.rlen[ref];
        const uint32_t *w2 = P.ref2 + P.w2off[ref];
        const uint32_t *wn = P.refN + P.nmoff[ref];
        int32_t best = LOCAL ? 0 : kNeg32;

        for (int s = 0; s < S; ++s) {
            uint32_t qm[kStripe];
            int32_t sma[kStripe], smi[kStripe];
#pragma unroll
            for (int r = 0; r < kStripe; ++r) {
                const int i = kStripe * s + r;
                if (i < p) {
                    qm[r] = 0;
                    sma[r] = 0;
                    smi[r] = 0;
                } else {
                    const int x = qc[i - p];
                    qm[r] = x == 4 ? 0xFu : (1u << x);
                    sma[r] = ma;
                    smi[r] = mi;
                }
            }
            int32_t H[kStripe], E[kStripe];
#pragma unroll
            for (int r = 0; r < kStripe; ++r) {
                H[r] = 0;
                E[r] = kNeg32;
            }
            int32_t hdt = 0;
            const bool first = s == 0, last = s + 1 == S;
            for (int j = 0; j < m; ++j) {
                const uint32_t c2 = (w2[j >> 4] >> (2 * (j & 15))) & 3u;
                const uint32_t rm = ((wn[j >> 5] >> (j & 31)) & 1u) ? 0xFu : (1u << c2);
                int32_t hup, f;
                if (first) {
                    hup = 0;
                    f = kNeg32;
                } else {
                    const int2 b = scr[(size_t)j * total];
                    hup = b.x;
                    f = b.y;
                }
                int32_t hdiag = hdt;
                hdt = hup;
#pragma unroll
                for (int r = 0; r < kStripe; ++r) {
                    const int32_t hleft = H[r];
                    const int32_t ev = max(hleft + o, E[r] + e);
                    E[r] = ev;
                    f = max(hup + o, f + e);
                    const int32_t sv = (qm[r] & rm) ? sma[r] : smi[r];
                    int32_t h = max(max(hdiag + sv, ev), f);
                    if (LOCAL) {
                        h = max(h, 0);
                        best = max(best, h);
                    }
                    hdiag = hleft;
                    H[r] = h;
                    hup = h;
                }
                if (!last) scr[(size_t)j * total] = make_int2(hup, f);
                if (!LOCAL && last) best = max(best, hup);
            }
            if (!LOCAL) {
#pragma unroll
                for (int r = 0; r < kStripe; ++r)
                    if (kStripe * s + r >= p) best = max(best, H[r]);
            }
        }
        const size_t oidx = out_index(a, q, ref);
        if (a.out32) a.out32[oidx] = best;
        if (a.out16) a.out16[oidx] = (int16_t)best;
    }
}

// ---------------------------------------------------------------------------
// Length bucketing: classify -> scan -> scatter. Reads go to the int16
// kernel in stripe-count bins, longest first; chunk k of 128 consecutive
// reads then shares one stripe count (its first read's).
// ---------------------------------------------------------------------------

__global__ void k_zero(KArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= kMaxLen16) {
        a.hist[i] = 0;
        a.bin_fill[i] = 0;
    }
    if (i == 0) {
        a.desc->n16 = 0;
        a.desc->n32 = 0;
        a.desc->nd = 0;
        a.desc->nextD = 0;
        a.desc->next16 = 0;
        a.desc->next32a = 0;
        a.desc->next32b = 0;
        a.desc->max16 = 0;
        a.desc->cells = 0;
        for (int k = 0; k < kMaxSlices; ++k) a.desc->next_slice[k] = 0;
    }
}

// Reverse complements of the loaded reads (lib/alphabet.rs:58-64) into the
// rows n_reads .. 2 * n_reads; one wave per read.
__global__ void k_revcomp(KArgs a) {
    const int nr = a.desc->n_reads;
    if (!a.desc->both) return;
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int waves = (gridDim.x * blockDim.x) >> 6;
    for (int q = wave; q < nr; q += waves) {
        const int n = a.qlen[q];
        const uint8_t *src = a.qcodes + a.qoff[q];
        uint8_t *dst = a.rc_codes + a.qoff[nr + q];
        for (int i = lane; i < n; i += 64) {
            const uint8_t c = src[n - 1 - i];
            dst[i] = c == 4 ? 4 : (uint8_t)(3 - c);
        }
    }
}

__device__ __forceinline__ bool fits_i16_dev(int64_t n, int64_t L, const Scoring &sc) {
    // lib/simd/mod.rs:222-227
    const int64_t k = n < L ? n : L;
    const int64_t M = max(abs(sc.ma), abs(sc.mi));
    return k * M + abs(sc.o) + abs(sc.e) < 15000 && L < 15000;
}

// A list append with one atomic per wave: each lane that wants a slot gets
// the list's old length plus its rank among those lanes (slots come in no
// particular order; nothing downstream depends on it). Every active lane of
// the wave must call it.
__device__ __forceinline__ int wave_append(int32_t *counter, bool want) {
    const uint64_t m = __ballot(want);
    if (!m) return 0;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, leader, 64);
    return base + __popcll(m & ((1ull << lane) - 1ull));
}

// Length bins shorter than this are counted per block in LDS and added to
// the global bins once per block: reads of one length are common (mates
// trimmed to similar lengths), and same-address global atomics serialise.
constexpr int kLdsBins = 4096;

// A block's contiguous share of n items (coalesced, and the same for every
// pass of a kernel over them).
__device__ __forceinline__ void block_range(int n, int &q0, int &q1) {
    const int per = (n + (int)gridDim.x - 1) / (int)gridDim.x;
    q0 = min(n, (int)blockIdx.x * per);
    q1 = min(n, q0 + per);
}

__global__ __launch_bounds__(256) void k_classify(KArgs a) {
    __shared__ int32_t hist[kLdsBins];
    for (int t = threadIdx.x; t < kLdsBins; t += blockDim.x) hist[t] = 0;
    __syncthreads();
    const int nq = a.desc->n_queries;
    const int dropin = a.desc->dropin, force32 = a.desc->force32;
    const int64_t L = a.panel.max_len;
    const int64_t M = max(abs(a.sc.ma), abs(a.sc.mi));
    uint64_t cells = 0;
    int max16 = 0;
    int q0, q1;
    block_range(nq, q0, q1);
    for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
        const int n = a.qlen[q];
        uint8_t c = 0;
        if (n > 0) {
            if (dropin) {
                // GpuScorer::scores_len (lib/cuda/mod.rs:506-508).
                c = (n <= 4096 && fits_i16_dev(n, L, a.sc)) ? 1 : 0;
            } else {
                const int64_t k = n < L ? n : L;
                const bool b16 = k * M + abs(a.sc.o) + abs(a.sc.e) < kBound16 && n <= kMaxLen16;
                c = (b16 && !force32) ? 1 : 2;
            }
        }
        if (c == 1 && a.desc->diag_min_len > 0 && n >= a.desc->diag_min_len) c = 3;
        a.cls[q] = c;
        const int posD = wave_append(&a.desc->nd, c == 3);
        const int pos32 = wave_append(&a.desc->n32, c == 2);
        if (c == 3) {
            a.listD[posD] = q;
        } else if (c == 1) {
            if (n < kLdsBins) atomicAdd(&hist[n], 1);
            else atomicAdd(&a.hist[n], 1);
            max16 = max(max16, n);
        } else if (c == 2) {
            a.list32[pos32] = q;
        } else if (a.out16 || a.out32) {
            // Unscored rows read back as zeros (ScoreMatrix::from_parts).
            const int nr = a.panel.n_refs;
            for (int r = 0; r < nr; ++r) {
                if (a.out16) a.out16[out_index(a, q, r)] = 0;
                if (a.out32) a.out32[out_index(a, q, r)] = 0;
            }
        }
        if (c) cells += (uint64_t)n;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kLdsBins; t += blockDim.x)
        if (hist[t]) atomicAdd(&a.hist[t], hist[t]);
    // one atomic per wave for the cell count and the longest int16 query
    for (int off = 32; off > 0; off >>= 1) {
        cells += __shfl_down(cells, off, 64);
        max16 = max(max16, __shfl_down(max16, off, 64));
    }
    if ((threadIdx.x & 63) == 0 && cells) atomicAdd((unsigned long long *)&a.desc->cells,
                                                    (unsigned long long)(cells * (uint64_t)a.panel.total_len));
    if ((threadIdx.x & 63) == 0 && max16) atomicMax(&a.desc->max16, max16);
}

// Exclusive prefix over the length bins in DESCENDING length (up to the
// longest int16 query), one block of 1024 threads, with the adaptive
// routing to the anti-diagonal kernel.
//
// rank(L) = reads longer than L. A read of length L would share its chunk
// with the reads down to rank(L) + 127, of length L'. When L > rho L' + 64
// (rho = diag_ratio / 16, the per-cell cost of the anti-diagonal kernel
// relative to the inter-sequence one, 2 by default) the chunk would be
// mostly padding, so the whole length bin goes to k_diag16 instead
// (bin_start = -1). And when the whole batch makes fewer inter-sequence
// work items (chunks x groups of 4 references) than that kernel keeps
// resident blocks, it cannot fill the device, so every bin of length >= 512
// goes to k_diag16 too, whose items are (read, 8 references). A batch whose
// kept rows do not fill one chunk goes there whole when that is cheaper.
// The remaining bins are compacted into sorted16 order.
__device__ int32_t block_exclusive_scan(int32_t v, int32_t *part, int32_t *total) {
    const int t = threadIdx.x;
    part[t] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int32_t x = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    const int32_t incl = part[t];
    *total = part[1023];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(1024) void k_scan(KArgs a) {
    __shared__ int32_t part[1024];
    // bins 0 .. top (longer ones are empty); thread t owns PER consecutive
    // bins top - t PER, top - t PER - 1, ... (one bin each up to 1024)
    const int top = a.desc->max16;
    const int PER = (top + 1 + 1023) / 1024;
    const int t = threadIdx.x;
    const int b0 = top - t * PER;
    const int nb = max(0, min(PER, b0 + 1));  // this thread's bins
    int32_t sum = 0;
    for (int k = 0; k < nb; ++k) sum += a.hist[b0 - k];
    int32_t total;
    int32_t run = block_exclusive_scan(sum, part, &total);
    // rank(L) = reads longer than L (the routing test reads other threads'
    // bins)
    {
        int32_t r = run;
        for (int k = 0; k < nb; ++k) {
            a.bin_start[b0 - k] = r;
            r += a.hist[b0 - k];
        }
    }
    __syncthreads();
    // the small-batch rules (starved, all) follow the ratio; each length
    // class's bins follow its plan entry (0 for every class: no routing)
    const int ratio = a.desc->diag_ratio;
    bool planned = false;
#pragma unroll
    for (int c = 0; c < kLenClasses; ++c) planned |= a.desc->plan[c] != 0;
    const int64_t items16 = (int64_t)((total + kChunk - 1) / kChunk) * ((a.panel.n_refs + 3) / 4);
    const bool starved = ratio > 0 && items16 < (int64_t)a.desc->blocks16;
    // routing flags into bin_fill (k_zero cleared it; cleared again below)
    int32_t kept = sum;
    if (ratio > 0 || planned) {
        for (int k = 0; k < nb; ++k) {
            const int L = b0 - k;
            const int cnt = a.hist[L];
            if (cnt == 0 || L < 128) continue;
            // the read at rank x (the chunk's last) has length L' and
            // 16 L > ratio L' + 1024  <=>  L' <= Lt = (16 L - 1025) / ratio
            //                         <=>  reads longer than Lt <= x
            const int r = a.bin_start[L];
            const int x = min(r + kChunk - 1, total - 1);
            const int num = 16 * L - 1025;
            // the length class's own rule (the autotuner's per-bucket plan)
            const int rc = a.desc->plan[len_class(L)];
            const bool thin = rc == 1 || (rc >= 16 && num >= 0 && a.bin_start[num / rc] <= x);  // num / rc < L
            if (thin || (starved && L >= 512)) {
                a.bin_fill[L] = 1;
                kept -= cnt;
            }
        }
    }
    __syncthreads();
    // compacted prefix over the kept bins; routed bins get -1
    int32_t n16;
    run = block_exclusive_scan(kept, part, &n16);
    // Kept rows that do not fill one chunk still cost k_dp16 a whole chunk:
    // 128 rows at the longest kept length, however few they are (a small
    // request against a large panel). The anti-diagonal kernel costs rho
    // times each row padded to its passes of 64 R rows (R = 2, 4, 8 by
    // length, as diag64_task); when that is less, every kept bin goes.
    __shared__ int32_t s_top;
    bool all = false;
    if (ratio > 0 && n16 > 0 && n16 < kChunk) {
        if (t == 0) s_top = 0;
        __syncthreads();
        int32_t pad = 0, top = 0;
        for (int k = 0; k < nb; ++k) {
            const int L = b0 - k;
            const int cnt = a.hist[L];
            if (cnt == 0 || a.bin_fill[L]) continue;
            const int rp = 64 * (L <= 128 ? 2 : L <= 256 ? 4 : 8);
            pad += cnt * ((L + rp - 1) / rp) * rp;
            top = max(top, L);
        }
        if (top) atomicMax(&s_top, top);
        int32_t pad_total;
        (void)block_exclusive_scan(pad, part, &pad_total);
        all = (int64_t)ratio * pad_total < (int64_t)16 * kChunk * s_top;
    }
    for (int k = 0; k < nb; ++k) {
        const int b = b0 - k;
        if (a.bin_fill[b] || all) {
            a.bin_start[b] = -1;
            a.bin_fill[b] = 0;
        } else {
            a.bin_start[b] = run;
            run += a.hist[b];
        }
    }
    if (t == 0) a.desc->n16 = all ? 0 : n16;
}

// Each int16 query's slot in sorted16: its bin's start plus its rank in the
// bin. A block first counts its queries per short bin in LDS, reserves a
// range of each bin with one global atomic, then hands out the slots from
// LDS; longer bins take global atomics directly.
__global__ __launch_bounds__(256) void k_scatter(KArgs a) {
    __shared__ int32_t cnt[kLdsBins], base[kLdsBins];
    for (int t = threadIdx.x; t < kLdsBins; t += blockDim.x) cnt[t] = 0;
    __syncthreads();
    const int nq = a.desc->n_queries;
    int q0, q1;
    block_range(nq, q0, q1);
    for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
        if (a.cls[q] != 1) continue;
        const int b = a.qlen[q];
        if (b < kLdsBins && a.bin_start[b] >= 0) atomicAdd(&cnt[b], 1);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kLdsBins; t += blockDim.x)
        if (cnt[t]) {
            base[t] = atomicAdd(&a.bin_fill[t], cnt[t]);
            cnt[t] = 0;
        }
    __syncthreads();
    for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
        const bool is16 = a.cls[q] == 1;
        const int b = is16 ? a.qlen[q] : 0;
        const int start = is16 ? a.bin_start[b] : 0;
        const bool diag = is16 && start < 0;  // the bin was routed to the anti-diagonal kernel
        const int posD = wave_append(&a.desc->nd, diag);
        if (diag) {
            a.cls[q] = 3;
            a.listD[posD] = q;
        } else if (is16) {
            const int pos = b < kLdsBins ? base[b] + atomicAdd(&cnt[b], 1) : atomicAdd(&a.bin_fill[b], 1);
            a.sorted16[start + pos] = q;
        }
    }
}

// ---------------------------------------------------------------------------
// Fused selection (lib/mapper.rs:444-472): one wave per read. Candidate k of
// the tie order is reference k/2 on strand k%2 with both strands, else
// reference k forward. best = max; ties = every candidate equal to it, in
// order; suboptimal = the best candidate that is not a tie.
// ---------------------------------------------------------------------------

__device__ __forceinline__ int wave_max(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ int wave_sum(int v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Candidates a lane of k_select holds in registers: a read's whole
// candidate list (both strands) up to 64 x this many is loaded once.
constexpr int kSelectRegs = 8;

__global__ __launch_bounds__(256) void k_select(KArgs a) {
    const DevDesc &d = *a.desc;
    const int n_reads = d.n_reads, both = d.both, nr = a.panel.n_refs;
    const int cap = d.tie_cap;
    const int64_t max_ties = d.max_ties;
    const int min_score = d.min_score;
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int waves = (gridDim.x * blockDim.x) >> 6;
    const int nc = nr * (both ? 2 : 1);
    for (int rd = wave; rd < n_reads; rd += waves) {
        int32_t *rec = a.recs + (size_t)rd * 5;
        if (a.qlen[rd] == 0) {
            if (lane < 5) rec[lane] = 0;
            continue;
        }
        const int32_t *fwd = a.out32 + (size_t)rd * nr;
        const int32_t *rev = a.out32 + (size_t)(n_reads + rd) * nr;
        auto cand = [&](int k) -> int {
            if (!both) return fwd[k];
            return (k & 1) ? rev[k >> 1] : fwd[k >> 1];
        };
        const int64_t take_lim = max_ties < 0 ? INT64_MAX : max_ties + 1;
        const int lim = (int)min((int64_t)cap, take_lim);
        int nt = 0, so = INT32_MIN, hs = 0, base = 0;
        // the tie pass over candidates c0 .. c0 + 63, lane k holding s
        auto tie_pass = [&](int c0, int s, int b) {
            const int k = c0 + lane;
            bool tie = false;
            if (k < nc) {
                tie = s == b;
                if (!tie) {
                    so = max(so, s);
                    hs = 1;
                }
            }
            const uint64_t mask = __ballot(tie);
            const int rank = __popcll(mask & ((1ull << lane) - 1ull));
            if (tie && base + rank < lim) {
                const int r = both ? (k >> 1) : k;
                const int rv = both ? (k & 1) : 0;
                a.ties[(size_t)rd * cap + base + rank] = (r << 1) | rv;
            }
            base += __popcll(mask);
        };
        int b = INT32_MIN;
        if (nc <= 64 * kSelectRegs) {
            // every candidate loaded once, into registers, all loads in
            // flight together; the tie pass reads the registers
            int sv[kSelectRegs];
#pragma unroll
            for (int i = 0; i < kSelectRegs; ++i) {
                const int k = 64 * i + lane;
                sv[i] = k < nc ? cand(k) : INT32_MIN;
                b = max(b, sv[i]);
            }
            b = wave_max(b);
#pragma unroll
            for (int i = 0; i < kSelectRegs; ++i)
                if (64 * i < nc) tie_pass(64 * i, sv[i], b);
        } else {
            for (int k = lane; k < nc; k += 64) b = max(b, cand(k));
            b = wave_max(b);
            for (int c0 = 0; c0 < nc; c0 += 64) tie_pass(c0, c0 + lane < nc ? cand(c0 + lane) : 0, b);
        }
        nt = base;
        so = wave_max(so);
        hs = wave_max(hs);
        if (lane == 0) {
            const int64_t take = b < min_score ? 0 : min((int64_t)nt, take_lim);
            rec[0] = b;
            rec[1] = nt;
            rec[2] = hs ? so : 0;
            rec[3] = 1 | (hs ? 2 : 0);
            rec[4] = (int32_t)take;
        }
    }
}

// Pair candidates (lib/mapper/pair.rs:185-231): one wave per pair. Mates are
// interleaved (row 2p = mate 1, 2p + 1 = mate 2), their reverse complements
// in the second half. On template strand t mate 1 is reversed iff t, mate 2
// iff !t; a candidate is eligible when both mates reach min_score, and the
// top tier is every eligible candidate with the best sum, in (ref, strand)
// order.
__global__ __launch_bounds__(256) void k_pairs(KArgs a) {
    const DevDesc &d = *a.desc;
    const int n_pairs = d.n_reads / 2;
    const int strands = d.tstrands;
    const int nr = a.panel.n_refs, cap = d.tie_cap, min_score = d.min_score;
    const int n_reads = d.n_reads;
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int waves = (gridDim.x * blockDim.x) >> 6;
    const int nc = nr * strands;
    for (int p = wave; p < n_pairs; p += waves) {
        int32_t *rec = a.recs + (size_t)p * 5;
        if (a.qlen[2 * p] == 0 || a.qlen[2 * p + 1] == 0) {
            if (lane < 5) rec[lane] = 0;
            continue;
        }
        const int32_t *m1f = a.out32 + (size_t)(2 * p) * nr;
        const int32_t *m2f = a.out32 + (size_t)(2 * p + 1) * nr;
        const int32_t *m1r = a.out32 + (size_t)(n_reads + 2 * p) * nr;
        const int32_t *m2r = a.out32 + (size_t)(n_reads + 2 * p + 1) * nr;
        auto cand = [&](int k, bool &ok) -> int {
            const int r = strands == 2 ? (k >> 1) : k;
            const int t = strands == 2 ? (k & 1) : 0;
            const int s1 = t ? m1r[r] : m1f[r];
            const int s2 = t ? m2f[r] : m2r[r];
            ok = s1 >= min_score && s2 >= min_score;
            return s1 + s2;
        };
        int b = INT32_MIN, cnt = 0;
        for (int k = lane; k < nc; k += 64) {
            bool ok;
            const int s = cand(k, ok);
            if (ok) {
                b = max(b, s);
                cnt++;
            }
        }
        b = wave_max(b);
        cnt = wave_sum(cnt);
        int nxt = INT32_MIN, hn = 0, base = 0;
        for (int c0 = 0; c0 < nc; c0 += 64) {
            const int k = c0 + lane;
            bool top = false;
            if (k < nc) {
                bool ok;
                const int s = cand(k, ok);
                if (ok) {
                    top = s == b;
                    if (!top) {
                        nxt = max(nxt, s);
                        hn = 1;
                    }
                }
            }
            const uint64_t mask = __ballot(top);
            const int rank = __popcll(mask & ((1ull << lane) - 1ull));
            if (top && base + rank < cap) {
                const int r = strands == 2 ? (k >> 1) : k;
                const int t = strands == 2 ? (k & 1) : 0;
                a.ties[(size_t)p * cap + base + rank] = (r << 1) | t;
            }
            base += __popcll(mask);
        }
        nxt = wave_max(nxt);
        hn = wave_max(hn);
        if (lane == 0) {
            rec[0] = cnt;
            rec[1] = cnt ? b : 0;
            rec[2] = base;
            rec[3] = hn ? nxt : 0;
            rec[4] = 1 | (hn ? 2 : 0);
        }
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers.
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// k_peak16 — the packed-int16 VALU peak, measured: every lane runs 8
// independent chains of the k_dp16 column's two operations (saturating
// v_pk_add_i16 and v_pk_max_i16), `iters` x 16 of them (unrolled to 128 per
// loop branch), over a grid that fills every SIMD. The result is stored so
// nothing is dead.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_peak16(uint32_t *out, int iters, uint32_t seed) {
    v2s x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = as_v2s(seed * (threadIdx.x + 1 + k) ^ blockIdx.x);
    const v2s a = as_v2s(seed | 1u), b = as_v2s(seed >> 3);
#pragma unroll 8
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            x[k] = adds(x[k], a);
            x[k] = vmax(x[k], b);
        }
    }
    v2s r = x[0];
#pragma unroll
    for (int k = 1; k < 8; ++k) r = vmax(r, x[k]);
    out[blockIdx.x * blockDim.x + threadIdx.x] = as_u32(r);
}

hipError_t launch_peak16(uint32_t *out, int blocks, int iters, hipStream_t st) {
    hipLaunchKernelGGL(k_peak16, dim3(blocks), dim3(256), 0, st, out, iters, 0x9E3779B9u);
    return hipGetLastError();
}

hipError_t launch_prep(const KArgs &a, const Launch &g, hipStream_t st) {
    hipLaunchKernelGGL(k_zero, dim3((kMaxLen16 + 1 + 255) / 256), dim3(256), 0, st, a);
    // one wave per read, latency-bound: twice the auxiliary grid's waves
    hipLaunchKernelGGL(k_revcomp, dim3(2 * g.grid_aux), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_classify, dim3(g.grid_aux), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, a);
    hipLaunchKernelGGL(k_scatter, dim3(g.grid_aux), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_revcomp(const KArgs &a, const Launch &g, hipStream_t st) {
    hipLaunchKernelGGL(k_revcomp, dim3(2 * g.grid_aux), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_dp16(const KArgs &a0, const Launch &g, bool local, hipStream_t st) {
    // the column-shift block of the launched variant's arithmetic
    KArgs a = a0;
    a.sc.cshift = dp16_is_f16(g.variant16) ? g.cshift_f16 : g.cshift_i16;
    if (a.sc.cshift < 0) a.sc.cshift = 0;  // a variant the scorer refuses never runs shifted columns
    switch (dp16_stripe(g.variant16)) {
        case 24: return local ? launch_dp16_lo24(a, g, st) : launch_dp16_sg24(a, g, st);
        case 32: return local ? launch_dp16_lo32(a, g, st) : launch_dp16_sg32(a, g, st);
        default: return local ? launch_dp16_lo16(a, g, st) : launch_dp16_sg16(a, g, st);
    }
}

static const void *dp16_fn(bool local, int variant) {
    switch (dp16_stripe(variant)) {
        case 24: return local ? dp16_fn_lo24(variant) : dp16_fn_sg24(variant);
        case 32: return local ? dp16_fn_lo32(variant) : dp16_fn_sg32(variant);
        default: return local ? dp16_fn_lo16(variant) : dp16_fn_sg16(variant);
    }
}

hipError_t launch_dp32(const KArgs &a, const Launch &g, bool local, int which, hipStream_t st) {
    if (local) hipLaunchKernelGGL(k_dp32<true>, dim3(g.grid32), dim3(g.block32), 0, st, a, which);
    else hipLaunchKernelGGL(k_dp32<false>, dim3(g.grid32), dim3(g.block32), 0, st, a, which);
    return hipGetLastError();
}

hipError_t launch_diag(const KArgs &a, const Launch &g, bool local, hipStream_t st) {
    if (g.diag_lds) {
        if (local) hipLaunchKernelGGL((k_diag64<true>), dim3(g.grid_diag), dim3(64), g.diag_lds, st, a);
        else hipLaunchKernelGGL((k_diag64<false>), dim3(g.grid_diag), dim3(64), g.diag_lds, st, a);
    } else {
        if (local) hipLaunchKernelGGL((k_diag16<kDiagRows, true>), dim3(g.grid_diag), dim3(64), 0, st, a);
        else hipLaunchKernelGGL((k_diag16<kDiagRows, false>), dim3(g.grid_diag), dim3(64), 0, st, a);
    }
    return hipGetLastError();
}

hipError_t diag_occupancy(bool local, size_t lds, int *blocks) {
    if (lds)
        return local ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k_diag64<true>, 64, lds)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k_diag64<false>, 64, lds);
    return local ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k_diag16<kDiagRows, true>, 64, 0)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k_diag16<kDiagRows, false>, 64, 0);
}

hipError_t launch_select(const KArgs &a, const Launch &g, hipStream_t st) {
    // two waves per read slot of the auxiliary grid: a wave's reads are
    // latency-bound loads, so more resident waves hide more of it
    hipLaunchKernelGGL(k_select, dim3(2 * g.grid_aux), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_pairs(const KArgs &a, const Launch &g, hipStream_t st) {
    hipLaunchKernelGGL(k_pairs, dim3(g.grid_aux), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t dp16_attributes(bool local, int variant, hipFuncAttributes *attr) {
    return hipFuncGetAttributes(attr, dp16_fn(local, variant));
}

hipError_t dp32_attributes(bool local, hipFuncAttributes *attr) {
    return local ? hipFuncGetAttributes(attr, reinterpret_cast<const void *>(&k_dp32<true>))
                 : hipFuncGetAttributes(attr, reinterpret_cast<const void *>(&k_dp32<false>));
}

hipError_t dp16_occupancy(bool local, int variant, int *blocks) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        blocks, dp16_fn(local, variant), 64 * kWaves16, 0);
}

hipError_t dp32_occupancy(bool local, int block, int *blocks) {
    return local ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k_dp32<true>, block, 0)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k_dp32<false>, block, 0);
}

}  // namespace esch
