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

namespace esch {

typedef short v2s __attribute__((ext_vector_type(2)));
// Panel words are read-only for a kernel's lifetime: read them through the
// constant address space so uniform indices become scalar loads.
typedef const __attribute__((address_space(4))) uint32_t *cwords;

__device__ __forceinline__ v2s adds(v2s a, v2s b) { return __builtin_elementwise_add_sat(a, b); }
__device__ __forceinline__ v2s vmax(v2s a, v2s b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ v2s splat(int v) { return v2s{(short)v, (short)v}; }
__device__ __forceinline__ uint32_t as_u32(v2s v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ v2s as_v2s(uint32_t v) { return __builtin_bit_cast(v2s, v); }
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

typedef _Float16 v2h __attribute__((ext_vector_type(2)));

// Packed arithmetic of k_dp16: two int16 lanes with saturating adds (-32768
// is -inf), or, for a panel whose value bound fits the integers fp16 holds
// exactly (kBoundF16), two fp16 lanes (-inf is -inf, and the max of three
// is one v_pk_maximum3_f16 where int16 needs two v_pk_max_i16).
template <bool F16>
struct Pk {
    using T = v2s;
    static __device__ __forceinline__ T add(T a, T b) { return __builtin_elementwise_add_sat(a, b); }
    static __device__ __forceinline__ T max(T a, T b) { return __builtin_elementwise_max(a, b); }
    static __device__ __forceinline__ T max3(T a, T b, T c) { return max(max(a, b), c); }
    static __device__ __forceinline__ T splat(int v) { return T{(short)v, (short)v}; }
    static __device__ __forceinline__ T neg() { return splat(-32768); }
    static __device__ __forceinline__ T from(uint32_t v) { return __builtin_bit_cast(T, v); }
    static __device__ __forceinline__ uint32_t bits(T v) { return __builtin_bit_cast(uint32_t, v); }
    static __device__ __forceinline__ uint32_t enc(int v) { return (uint32_t)(uint16_t)(int16_t)v; }
    static __device__ __forceinline__ void neg_lo(T &v) { v.x = (short)-32768; }
    static __device__ __forceinline__ void neg_hi(T &v) { v.y = (short)-32768; }
    static __device__ __forceinline__ int lo(T v) { return v.x; }
    static __device__ __forceinline__ int hi(T v) { return v.y; }
};

template <>
struct Pk<true> {
    using T = v2h;
    static __device__ __forceinline__ T add(T a, T b) { return a + b; }
    // IEEE maximum: no NaN can arise, so no quieting ops are needed
    static __device__ __forceinline__ T max(T a, T b) { return __builtin_elementwise_maximum(a, b); }
    static __device__ __forceinline__ T max3(T a, T b, T c) {
        return __builtin_elementwise_maximum(__builtin_elementwise_maximum(a, b), c);
    }
    static __device__ __forceinline__ T splat(int v) { return T{(_Float16)v, (_Float16)v}; }
    static __device__ __forceinline__ T neg() { return from(0xFC00FC00u); }
    static __device__ __forceinline__ T from(uint32_t v) { return __builtin_bit_cast(T, v); }
    static __device__ __forceinline__ uint32_t bits(T v) { return __builtin_bit_cast(uint32_t, v); }
    static __device__ __forceinline__ uint32_t enc(int v) {
        return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)v);
    }
    static __device__ __forceinline__ void neg_lo(T &v) { v = from((bits(v) & 0xFFFF0000u) | 0xFC00u); }
    static __device__ __forceinline__ void neg_hi(T &v) { v = from((bits(v) & 0x0000FFFFu) | 0xFC000000u); }
    static __device__ __forceinline__ int lo(T v) { return (int)(float)v.x; }
    static __device__ __forceinline__ int hi(T v) { return (int)(float)v.y; }
};

// ---------------------------------------------------------------------------
// k_dp16 — batched packed-int16 DP.
//
// Layout. A workgroup of four waves scores one work item: a chunk of 128
// reads of the length-bucketed order (two reads per lane, packed as the low
// and high int16 halves of every register) against four panel references
// of neighbouring length, one reference per wave. Within a wave the
// reference is uniform, so its code per column is a scalar; the read side
// of the substitution comes from a query profile in LDS, shared by the four
// waves: plane c holds s(q_r, c) - o for the 16 rows of the current stripe
// (plane 4, a wildcard reference base, the best of the four). Selecting a
// plane is a uniform LDS offset, so a column is four ds_read_b128 and a
// branch-free body of 8 packed VALU ops per row (4 saturating adds, 4 max)
// in semiglobal mode and 10 in local mode, for two cells.
//
// Stripes. The read is walked in stripes of 16 rows held in registers
// (Ho = H + o and E per row); the stripe's bottom row (Ho, F) per column is
// the only state that leaves registers, to a per-wave scratch the next
// stripe reads back (coalesced: scratch[j * 64 + lane]).
//
// Bottom alignment. Every read of a chunk is padded at the TOP to the
// chunk's stripe count with rows whose substitution score is 0. Starting
// from the zero border such rows keep H = 0 and E = F = o, which is exactly
// what the first real row would see from the border (H = 0, F = -inf:
// max(0 + o, -inf + e) = o), so padding changes no real cell, and the
// read's last row is always register 15 of the last stripe.
//
// Saturation. Adds clamp to int16; a query is sent here only when its
// value bound (kBound16) keeps every real value off both clamps, so the
// result equals the int32 recurrence exactly. Others go to k_dp32.
// ---------------------------------------------------------------------------

constexpr int kWaves16 = 4;  // waves (references) per k_dp16 workgroup

__device__ __forceinline__ int prof_val(int x, int c, const Scoring &sc) {
    // s(x, c) - o for read code x against reference code c; padding rows 0 - o.
    if (x == kPreCode) return -sc.o;
    return ((x == c || x == 4) ? sc.ma : sc.mi) - sc.o;
}

__device__ __forceinline__ uint32_t pack2(int lo, int hi) {
    return (uint32_t)(uint16_t)(int16_t)lo | ((uint32_t)(uint16_t)(int16_t)hi << 16);
}

// The maximum of W packed values by three-way maxima, a level at a time;
// a level's leftovers (W mod 3) carry up: W / 2 ops.
template <bool F16, int W>
__device__ __forceinline__ typename Pk<F16>::T max3_all(const typename Pk<F16>::T (&t)[W]) {
    using P = Pk<F16>;
    if constexpr (W == 1) {
        return t[0];
    } else if constexpr (W == 2) {
        return P::max(t[0], t[1]);
    } else {
        constexpr int NW = W / 3 + W % 3;
        typename P::T u[NW];
#pragma unroll
        for (int i = 0; i < W / 3; ++i) u[i] = P::max3(t[3 * i], t[3 * i + 1], t[3 * i + 2]);
#pragma unroll
        for (int i = 0; i < W % 3; ++i) u[W / 3 + i] = t[3 * (W / 3) + i];
        return max3_all<F16, NW>(u);
    }
}

template <int RR, bool LOCAL, bool F16 = false>
__device__ __forceinline__ void column16(const typename Pk<F16>::T (&sc)[RR], typename Pk<F16>::T (&Ho)[RR],
                                         typename Pk<F16>::T (&E)[RR], typename Pk<F16>::T &hou,
                                         typename Pk<F16>::T &f, typename Pk<F16>::T hod,
                                         typename Pk<F16>::T vo, typename Pk<F16>::T ve,
                                         typename Pk<F16>::T &best) {
    using P = Pk<F16>;
    using T = typename P::T;
#pragma unroll
    for (int r = 0; r < RR; ++r) {
        E[r] = P::max(Ho[r], P::add(E[r], ve));
        f = P::max(hou, P::add(f, ve));
        T h = P::max3(P::add(hod, sc[r]), E[r], f);
        if constexpr (LOCAL) h = P::max(h, P::splat(0));
        hod = Ho[r];
        Ho[r] = P::add(h, vo);
        hou = Ho[r];
    }
    if constexpr (LOCAL && F16) {
        // The column's best cell folded into `best` by a tree of three-way
        // maxima (RR / 2 ops for RR rows).
        T t[RR + 1];
#pragma unroll
        for (int r = 0; r < RR; ++r) t[r] = Ho[r];
        t[RR] = best;
        best = max3_all<F16, RR + 1>(t);
    } else if constexpr (LOCAL) {
        // The column's best cell, in Ho = H + o space, as a depth-4 tree
        // (same 16 max ops as a running max, without a serial chain).
        v2s t[RR / 2];
#pragma unroll
        for (int r = 0; r < RR / 2; ++r) t[r] = vmax(Ho[2 * r], Ho[2 * r + 1]);
        int w = RR / 2;
#pragma unroll
        while (w > 1) {
#pragma unroll
            for (int r = 0; r < w / 2; ++r) t[r] = vmax(t[2 * r], t[2 * r + 1]);
            if (w & 1) t[0] = vmax(t[0], t[w - 1]);
            w /= 2;
        }
        best = vmax(best, t[0]);
    }
}

// The semiglobal column with every value of stripe row r shifted by -e r
// (r = 0 .. RR-1 within the stripe; -1 for the row above), which takes the
// extension out of the vertical gap chain. In this space
//   P = H - e r + o,  R = E - e r + e,  Q = F - e r + e
// and the recurrence of the header becomes
//   R[j]    = max(P[j-1], R[j-1]) + e
//   Q[r]    = max(P[r-1], Q[r-1])
//   P[r][j] = max(P[r-1][j-1] + (s - o), R, Q) + (o - e)
// (expand each side: the shifts cancel), 6 packed ops per row where the
// unshifted column takes 7 (F's extension), and the profile is unchanged
// (s - o). The row above arrives from the previous stripe in that stripe's
// space, whose bottom row r' = h - 1 (h its height) is this stripe's row -1:
// values there are lower by c = e h here. ADJ adds c to Q at row 0, and
// the profile's row 0 carries + c for the diagonal term.
template <int RR, bool ADJ, bool F16>
__device__ __forceinline__ void column16s(const typename Pk<F16>::T (&sc)[RR], typename Pk<F16>::T (&Pv)[RR],
                                          typename Pk<F16>::T (&R)[RR], typename Pk<F16>::T &pu,
                                          typename Pk<F16>::T &q, typename Pk<F16>::T pd,
                                          typename Pk<F16>::T voe, typename Pk<F16>::T ve,
                                          typename Pk<F16>::T vc) {
    using P = Pk<F16>;
#pragma unroll
    for (int r = 0; r < RR; ++r) {
        R[r] = P::add(P::max(Pv[r], R[r]), ve);
        q = P::max(pu, q);
        if constexpr (ADJ) {
            if (r == 0) q = P::add(q, vc);
        }
        const typename P::T h = P::max3(P::add(pd, sc[r]), R[r], q);
        pd = Pv[r];
        Pv[r] = P::add(h, voe);
        pu = Pv[r];
    }
}

// Rows q*4 .. q*4+3 of a profile plane (q < RR / 4), or its last two rows
// (q = RR / 4, RR % 4 == 2), from the plane at pc (this lane's column).
template <int RR, bool F16>
__device__ __forceinline__ void plane_part(const uint4 *pc, int q, typename Pk<F16>::T (&sc)[RR]) {
    using P = Pk<F16>;
    if (4 * q + 4 <= RR) {
        const uint4 v = pc[q * 64];
        sc[4 * q + 0] = P::from(v.x);
        sc[4 * q + 1] = P::from(v.y);
        sc[4 * q + 2] = P::from(v.z);
        sc[4 * q + 3] = P::from(v.w);
    } else if (4 * q + 2 <= RR) {
        const uint2 v = *reinterpret_cast<const uint2 *>(pc + q * 64);
        sc[4 * q + 0] = P::from(v.x);
        sc[4 * q + 1] = P::from(v.y);
    }
}

// Two consecutive columns of column16s, A then B, run one row apart: B's
// row r - 1 needs only A's rows up to r - 1, so the two serial chains (Q,
// then H, then P, row after row) interleave instead of running back to
// back; each cell's arithmetic is column16s'. A plane's quad of rows is
// reloaded, from nA / nB (the next pair's columns), once its column has
// used it, so the next pair's profile is in flight during this one.
template <int RR, bool ADJ, bool F16>
__device__ __forceinline__ void column16s_pair(typename Pk<F16>::T (&scA)[RR], typename Pk<F16>::T (&scB)[RR],
                                               typename Pk<F16>::T (&Pv)[RR], typename Pk<F16>::T (&R)[RR],
                                               typename Pk<F16>::T &puA, typename Pk<F16>::T &qA,
                                               typename Pk<F16>::T pdA, typename Pk<F16>::T &puB,
                                               typename Pk<F16>::T &qB, typename Pk<F16>::T pdB,
                                               typename Pk<F16>::T voe, typename Pk<F16>::T ve,
                                               typename Pk<F16>::T vc, const uint4 *nA, const uint4 *nB) {
    using P = Pk<F16>;
    using T = typename P::T;
    constexpr int NP = (RR + 3) / 4;  // quads, the last one possibly half
    auto row = [&](int r, const T (&sc)[RR], T &pu, T &q, T &pd) {
        R[r] = P::add(P::max(Pv[r], R[r]), ve);
        q = P::max(pu, q);
        if constexpr (ADJ) {
            if (r == 0) q = P::add(q, vc);
        }
        const T h = P::max3(P::add(pd, sc[r]), R[r], q);
        pd = Pv[r];
        Pv[r] = P::add(h, voe);
        pu = Pv[r];
    };
    row(0, scA, puA, qA, pdA);
#pragma unroll
    for (int r = 1; r < RR; ++r) {
        row(r, scA, puA, qA, pdA);
        row(r - 1, scB, puB, qB, pdB);
        if ((r & 3) == 3) plane_part<RR, F16>(nA, r >> 2, scA);
        if (((r - 1) & 3) == 3) plane_part<RR, F16>(nB, (r - 1) >> 2, scB);
    }
    row(RR - 1, scB, puB, qB, pdB);
    // the quads not yet reloaded: A's from row (RR / 4) * 4 on, B's from
    // ((RR - 1) / 4) * 4 on
#pragma unroll
    for (int q = RR / 4; q < NP; ++q) plane_part<RR, F16>(nA, q, scA);
#pragma unroll
    for (int q = (RR - 1) / 4; q < NP; ++q) {
        if (4 * q + 3 > RR - 2 || 4 * q + 3 >= RR) plane_part<RR, F16>(nB, q, scB);
    }
}

// The local column with every value of reference column j shifted by -e j,
// which takes the extension out of the horizontal gap chain instead:
//   P = H - e j + o,  R = E - e j + e,  Q = F - e j + e
//   R[j]    = max(P[j-1], R[j-1])
//   Q[r]    = max(P[r-1], Q[r-1]) + e
//   P[r][j] = max(P[r-1][j-1] + (s - o), R, Q, z_j) + (o - e)
// where z_j = e (1 - j) is the zero floor H >= 0 in this space: one value
// per column, shared by every row (a row shift would need one per row, and a
// best per row). 7 packed ops per row where the unshifted local column takes
// 8, and a column's rows share one shift, so their best folds by a tree of
// three-way maxima as before. Returns the column's best P (column j space).
template <int RR, bool F16>
__device__ __forceinline__ typename Pk<F16>::T column16c(const typename Pk<F16>::T (&sc)[RR],
                                                         typename Pk<F16>::T (&Pv)[RR],
                                                         typename Pk<F16>::T (&R)[RR], typename Pk<F16>::T &pu,
                                                         typename Pk<F16>::T &q, typename Pk<F16>::T pd,
                                                         typename Pk<F16>::T voe, typename Pk<F16>::T ve,
                                                         typename Pk<F16>::T z) {
    using P = Pk<F16>;
#pragma unroll
    for (int r = 0; r < RR; ++r) {
        R[r] = P::max(Pv[r], R[r]);
        q = P::add(P::max(pu, q), ve);
        const typename P::T h = P::max(P::max3(P::add(pd, sc[r]), R[r], q), z);
        pd = Pv[r];
        Pv[r] = P::add(h, voe);
        pu = Pv[r];
    }
    return max3_all<F16, RR>(Pv);
}

// Two consecutive columns of column16c run one row apart (as column16s_pair),
// the next pair's planes reloaded quad by quad; tA / tB receive the two
// columns' best P, each folded two rows at a time while its values are live.
template <int RR, bool F16>
__device__ __forceinline__ void column16c_pair(typename Pk<F16>::T (&scA)[RR], typename Pk<F16>::T (&scB)[RR],
                                               typename Pk<F16>::T (&Pv)[RR], typename Pk<F16>::T (&R)[RR],
                                               typename Pk<F16>::T &puA, typename Pk<F16>::T &qA,
                                               typename Pk<F16>::T pdA, typename Pk<F16>::T &puB,
                                               typename Pk<F16>::T &qB, typename Pk<F16>::T pdB,
                                               typename Pk<F16>::T voe, typename Pk<F16>::T ve,
                                               typename Pk<F16>::T zA, typename Pk<F16>::T zB,
                                               typename Pk<F16>::T &tA, typename Pk<F16>::T &tB,
                                               const uint4 *nA, const uint4 *nB) {
    using P = Pk<F16>;
    using T = typename P::T;
    constexpr int NP = (RR + 3) / 4;
    auto row = [&](int r, const T (&sc)[RR], T &pu, T &q, T &pd, T z) {
        R[r] = P::max(Pv[r], R[r]);
        q = P::add(P::max(pu, q), ve);
        const T h = P::max(P::max3(P::add(pd, sc[r]), R[r], q), z);
        pd = Pv[r];
        Pv[r] = P::add(h, voe);
        pu = Pv[r];
    };
    // rows r - 1 and r (r odd) into a column's running best
    auto fold = [&](int r, T &t) { t = r == 1 ? P::max(Pv[0], Pv[1]) : P::max3(t, Pv[r - 1], Pv[r]); };
    row(0, scA, puA, qA, pdA, zA);
#pragma unroll
    for (int r = 1; r < RR; ++r) {
        row(r, scA, puA, qA, pdA, zA);
        if (r & 1) fold(r, tA);  // before B overwrites row r - 1
        row(r - 1, scB, puB, qB, pdB, zB);
        if (!(r & 1) && r >= 2) fold(r - 1, tB);
        if ((r & 3) == 3) plane_part<RR, F16>(nA, r >> 2, scA);
        if (((r - 1) & 3) == 3) plane_part<RR, F16>(nB, (r - 1) >> 2, scB);
    }
    row(RR - 1, scB, puB, qB, pdB, zB);
    fold(RR - 1, tB);
#pragma unroll
    for (int q = RR / 4; q < NP; ++q) plane_part<RR, F16>(nA, q, scA);
#pragma unroll
    for (int q = (RR - 1) / 4; q < NP; ++q) plane_part<RR, F16>(nB, q, scB);
}

// One stripe of RR rows (RR = 16, or 4/8/12 for a chunk's first stripe)
// against the wave's reference, columns 0 .. m-1. row0 is the stripe's first
// row in the chunk's padded row numbering, pA / pB the padding rows of the
// lane's two reads. The loop is unrolled by two with ping-pong registers so
// that column j + 1's profile plane (LDS) and column j + 2's boundary (HBM)
// are in flight while column j computes.
template <int RR, bool LOCAL, bool PF>
__device__ __forceinline__ void stripe16(const uint4 *__restrict__ prof, uint2 *__restrict__ scr,
                                         const int lane, const cwords w2, const cwords wn,
                                         const int m, const bool first, const bool last,
                                         const int row0, const int pA, const int pB, const v2s vo,
                                         const v2s ve, v2s &best) {
    constexpr int NQ = RR / 4;
    const v2s NEG = splat(-32768);
    v2s Ho[RR], E[RR];
#pragma unroll
    for (int r = 0; r < RR; ++r) {
        Ho[r] = vo;  // H = 0 on the left border
        E[r] = NEG;
    }
    v2s hdt = vo;  // Ho of the row above at column j - 1 (border: H = 0)
    auto code = [&](int j) -> int {
        const uint32_t c2 = (w2[j >> 4] >> (2 * (j & 15))) & 3u;
        return ((wn[j >> 5] >> (j & 31)) & 1u) ? 4 : (int)c2;
    };
    auto plane = [&](int c, v2s (&sc)[RR]) {
        const uint4 *pc = prof + c * 4 * 64 + lane;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const uint4 v = pc[q * 64];
            sc[4 * q + 0] = as_v2s(v.x);
            sc[4 * q + 1] = as_v2s(v.y);
            sc[4 * q + 2] = as_v2s(v.z);
            sc[4 * q + 3] = as_v2s(v.w);
        }
    };
    const uint2 border = make_uint2(as_u32(vo), as_u32(NEG));
    auto bnd = [&](int j) -> uint2 { return (first || j >= m) ? border : scr[(size_t)j * 64 + lane]; };
    auto col = [&](int j, const v2s (&sc)[RR], uint2 b) {
        v2s hou = as_v2s(b.x), f = as_v2s(b.y);
        const v2s hod = hdt;
        hdt = hou;
        column16<RR, LOCAL>(sc, Ho, E, hou, f, hod, vo, ve, best);
        if (!last) scr[(size_t)j * 64 + lane] = make_uint2(as_u32(hou), as_u32(f));
        if (!LOCAL && last) best = vmax(best, hou);  // last row, Ho space
    };

    if constexpr (PF) {
        v2s sa[RR], sb[RR];
        plane(code(0), sa);
        uint2 ba = bnd(0), bb = bnd(1);
        int j = 0;
        for (; j + 1 < m; j += 2) {
            plane(code(j + 1), sb);
            const uint2 bc = bnd(j + 2);
            col(j, sa, ba);
            if (j + 2 < m) plane(code(j + 2), sa);
            const uint2 bd = bnd(j + 3);
            col(j + 1, sb, bb);
            ba = bc;
            bb = bd;
        }
        if (j < m) col(j, sa, ba);
    } else {
        // Profile plane loaded in the column that uses it (fewer registers,
        // one more resident wave); the boundary still one column ahead.
        uint2 bn = bnd(0);
        for (int j = 0; j < m; ++j) {
            const uint2 b = bn;
            bn = bnd(j + 1);
            v2s sc[RR];
            plane(code(j), sc);
            col(j, sc, b);
        }
    }

    if (!LOCAL) {
        // Last column: every real row of this stripe (padding rows out).
#pragma unroll
        for (int r = 0; r < RR; ++r) {
            v2s v = Ho[r];
            if (row0 + r < pA) v.x = (short)-32768;
            if (row0 + r < pB) v.y = (short)-32768;
            best = vmax(best, v);
        }
    }
}

// The same stripe with every memory access of a column pair issued at the
// start of the pair: the boundary loads of the next pair and the deferred
// stores of the previous one, then a scheduling barrier, then the two
// columns. gfx9-class hardware counts loads and stores on one counter, so
// with both in flight any wait for a loaded value drains the counter; here
// that wait falls at the end of the pair, two columns after the requests.
// No data-dependent branch surrounds a memory access: the stores of the
// pair before column 0 go to two padding columns in front of the wave's
// scratch, loads past the last column read padding behind it (columns up
// to m + 1), the profile prefetch reads at most the reference's padding
// word (column m), and whether the stripe is the first (border instead of
// scratch) or the last (no stores) is a template argument. SH (semiglobal
// only) runs the row-shifted column (column16s), c being the shift offset
// of the row above (0 for a first stripe, whose border is given in this
// stripe's space); CS (local only) the column-shifted one (column16c), with
// `best` carried in column j's space through the stripe.
template <int RR, bool LOCAL, bool FIRST, bool LAST, bool QUAD, bool F16, bool OCT = false, bool SH = false,
          bool CS = false>
__device__ __forceinline__ void stripe16d(const uint4 *__restrict__ prof, uint2 *__restrict__ scr,
                                          const int lane, const cwords w2, const cwords wn,
                                          const int m, const int row0, const int pA, const int pB,
                                          const typename Pk<F16>::T vo, const typename Pk<F16>::T ve,
                                          typename Pk<F16>::T &best, const int c = 0) {
    static_assert(RR % 2 == 0 && RR >= 2 && RR <= 16, "an even number of rows");
    static_assert(!(SH && LOCAL), "the row shift is a semiglobal form");
    static_assert(!(CS && !LOCAL), "the column shift is a local form");
    constexpr bool XS = SH || CS;  // a shifted column: pairs run interleaved
    using P = Pk<F16>;
    using v2s = typename P::T;  // the packed value type of this instance
    constexpr int NQ = RR / 4;          // whole quads of rows
    constexpr bool HALF = RR % 4 == 2;  // and the first two rows of one more
    const v2s NEG = P::neg();
    // SH: o and e as ints again (vo, ve are their splats), the shifted
    // space's constants, and the last stripe's bottom-row best kept apart
    // (it is in row RR - 1's space)
    const int so = XS ? P::lo(vo) : 0, se = XS ? P::lo(ve) : 0;
    // wave-uniform constants through readfirstlane: scalar registers, each
    // an operand of a packed op, where a vector copy would cost a VGPR
    auto us = [](v2s v) { return P::from((uint32_t)uni((int)P::bits(v))); };
    const v2s voe = us(P::splat(so - se)), vc = us(P::splat(c));
    const v2s vne = us(P::splat(-se)), vse = XS ? us(ve) : ve;
    v2s bestL = NEG;
    // CS: the top border's P = o - e j (first stripe), the floor z_j =
    // e (1 - j), both advanced by -e a column, and best moved from H + o
    // (column 0's space) to column -1's
    v2s vtop = P::splat(so), z = P::splat(se);
    const v2s vn2e = us(P::splat(-2 * se));
    if constexpr (CS) best = P::add(best, us(P::splat(se)));
    v2s Ho[RR], E[RR];
    // SH: the left border's o - e r as a chain of adds from a base the
    // compiler cannot see through, so the RR constants are not hoisted out
    // of the stripe loop into registers held for the whole kernel
    v2s t = vo;
    if constexpr (SH) asm volatile("" : "+v"(t));
#pragma unroll
    for (int r = 0; r < RR; ++r) {
        Ho[r] = CS ? us(P::splat(so + se)) : t;  // H = 0 on the left border
        if constexpr (SH) t = P::add(t, vne);
        E[r] = NEG;
    }
    v2s hdt = SH ? P::splat(so + se - c) : CS ? us(P::splat(so + se)) : vo;
    auto code = [&](int j) -> int {
        const uint32_t c2 = (w2[j >> 4] >> (2 * (j & 15))) & 3u;
        return ((wn[j >> 5] >> (j & 31)) & 1u) ? 4 : (int)c2;
    };
    auto plane = [&](int c, v2s (&sc)[RR]) {
        const uint4 *pc = prof + c * 4 * 64 + lane;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const uint4 v = pc[q * 64];
            sc[4 * q + 0] = P::from(v.x);
            sc[4 * q + 1] = P::from(v.y);
            sc[4 * q + 2] = P::from(v.z);
            sc[4 * q + 3] = P::from(v.w);
        }
        if constexpr (HALF) {
            const uint2 v = *reinterpret_cast<const uint2 *>(pc + NQ * 64);
            sc[4 * NQ + 0] = P::from(v.x);
            sc[4 * NQ + 1] = P::from(v.y);
        }
    };
    uint2 *const s0 = scr + 2 * 64 + lane;  // column 0 (two padding columns in front)
    // the row above a first stripe: H = 0, F = -inf (SH: P = o + e)
    const uint2 border = make_uint2(P::bits(SH ? P::splat(so + se) : vo), P::bits(NEG));
    auto bnd = [&](int j) -> uint2 {
        if constexpr (FIRST) return border;
        else return s0[(size_t)j * 64];
    };
    auto col = [&](const v2s (&sc)[RR], uint2 b) -> uint2 {
        v2s hou = P::from(b.x), f = P::from(b.y);
        if constexpr (CS && FIRST) {
            hou = vtop;
            vtop = P::add(vtop, vne);
        }
        const v2s hod = hdt;
        hdt = hou;
        if constexpr (CS) {
            const v2s tcol = column16c<RR, F16>(sc, Ho, E, hou, f, hod, voe, vse, z);
            z = P::add(z, vne);
            best = P::max(P::add(best, vne), tcol);
        } else if constexpr (SH) {
            column16s<RR, !FIRST, F16>(sc, Ho, E, hou, f, hod, voe, vse, vc);
            if constexpr (LAST) bestL = P::max(bestL, hou);
        } else {
            column16<RR, LOCAL, F16>(sc, Ho, E, hou, f, hod, vo, ve, best);
            if constexpr (!LOCAL && LAST) best = P::max(best, hou);
        }
        return make_uint2(P::bits(hou), P::bits(f));
    };

    v2s sa[RR], sb[RR];
    plane(code(0), sa);
    if constexpr (XS) plane(code(1 < m ? 1 : m), sb);
    uint2 o0 = border, o1 = border;
    int j = 0;
    // Columns jj (plane sa) and jj + 1 into o0 and o1. SH: both planes are
    // loaded on entry, and the pair loads those of columns jj + 2 and jj + 3
    // (at most the padding column m) for the next one as it goes.
    auto pair = [&](int jj, uint2 x0, uint2 x1) {
        if constexpr (CS) {
            v2s puA = P::from(x0.x), qA = P::from(x0.y), puB = P::from(x1.x), qB = P::from(x1.y);
            if constexpr (FIRST) {
                puA = vtop;
                puB = P::add(vtop, vne);
                vtop = P::add(puB, vne);
            }
            const v2s pdA = hdt, pdB = puA;
            hdt = puB;
            const v2s zA = z, zB = P::add(z, vne);
            z = P::add(zB, vne);
            const int ja = jj + 2 < m ? jj + 2 : m, jb = jj + 3 < m ? jj + 3 : m;
            v2s tA, tB;
            column16c_pair<RR, F16>(sa, sb, Ho, E, puA, qA, pdA, puB, qB, pdB, voe, vse, zA, zB, tA, tB,
                                    prof + code(ja) * 4 * 64 + lane, prof + code(jb) * 4 * 64 + lane);
            // best from column jj - 1's space to jj + 1's
            best = P::max3(P::add(best, vn2e), P::add(tA, vne), tB);
            o0 = make_uint2(P::bits(puA), P::bits(qA));
            o1 = make_uint2(P::bits(puB), P::bits(qB));
        } else if constexpr (SH) {
            v2s puA = P::from(x0.x), qA = P::from(x0.y), puB = P::from(x1.x), qB = P::from(x1.y);
            const v2s pdA = hdt, pdB = puA;
            hdt = puB;
            const int ja = jj + 2 < m ? jj + 2 : m, jb = jj + 3 < m ? jj + 3 : m;
            column16s_pair<RR, !FIRST, F16>(sa, sb, Ho, E, puA, qA, pdA, puB, qB, pdB, voe, vse, vc,
                                            prof + code(ja) * 4 * 64 + lane, prof + code(jb) * 4 * 64 + lane);
            if constexpr (LAST) bestL = P::max3(bestL, puA, puB);
            o0 = make_uint2(P::bits(puA), P::bits(qA));
            o1 = make_uint2(P::bits(puB), P::bits(qB));
        } else {
            plane(code(jj + 1), sb);
            o0 = col(sa, x0);
            plane(code(jj + 2), sa);  // jj + 2 <= m: at most the padding word
            o1 = col(sb, x1);
        }
    };
    if constexpr (OCT) {
    uint2 a0 = bnd(0), a1 = bnd(1), b0 = bnd(2), b1 = bnd(3);
    // Eight columns per iteration as four pairs over four sets of boundary
    // registers (a, b, c, d), each loaded two pairs before its use: the
    // boundary loads run four columns ahead (the farthest, j + 11 <= m + 3,
    // reads the scratch padding) and no loaded register is copied.
    for (; j + 7 < m; j += 8) {
        const uint2 c0 = bnd(j + 4), c1 = bnd(j + 5);
        if constexpr (!LAST) {
            s0[(ptrdiff_t)(j - 2) * 64] = o0;
            s0[(ptrdiff_t)(j - 1) * 64] = o1;
        }
        __builtin_amdgcn_sched_barrier(0);
        pair(j, a0, a1);
        __builtin_amdgcn_sched_barrier(0);
        const uint2 d0 = bnd(j + 6), d1 = bnd(j + 7);
        if constexpr (!LAST) {
            s0[(ptrdiff_t)j * 64] = o0;
            s0[(ptrdiff_t)(j + 1) * 64] = o1;
        }
        __builtin_amdgcn_sched_barrier(0);
        pair(j + 2, b0, b1);
        __builtin_amdgcn_sched_barrier(0);
        a0 = bnd(j + 8);
        a1 = bnd(j + 9);
        if constexpr (!LAST) {
            s0[(ptrdiff_t)(j + 2) * 64] = o0;
            s0[(ptrdiff_t)(j + 3) * 64] = o1;
        }
        __builtin_amdgcn_sched_barrier(0);
        pair(j + 4, c0, c1);
        __builtin_amdgcn_sched_barrier(0);
        b0 = bnd(j + 10);
        b1 = bnd(j + 11);
        if constexpr (!LAST) {
            s0[(ptrdiff_t)(j + 4) * 64] = o0;
            s0[(ptrdiff_t)(j + 5) * 64] = o1;
        }
        __builtin_amdgcn_sched_barrier(0);
        pair(j + 6, d0, d1);
        __builtin_amdgcn_sched_barrier(0);
    }
    // 0-7 columns left: pairs with the sets moved down (copies, once per
    // stripe), then a single column
    while (j + 1 < m) {
        const uint2 c0 = bnd(j + 4), c1 = bnd(j + 5);  // j + 5 <= m + 3
        if constexpr (!LAST) {
            s0[(ptrdiff_t)(j - 2) * 64] = o0;
            s0[(ptrdiff_t)(j - 1) * 64] = o1;
        }
        pair(j, a0, a1);
        a0 = b0;
        a1 = b1;
        b0 = c0;
        b1 = c1;
        j += 2;
    }
    if constexpr (!LAST) {
        s0[(ptrdiff_t)(j - 2) * 64] = o0;
        s0[(ptrdiff_t)(j - 1) * 64] = o1;
    }
    if (j < m) {
        o0 = col(sa, a0);
        if constexpr (!LAST) s0[(size_t)j * 64] = o0;
    }

    } else if constexpr (!QUAD) {
    uint2 r0 = bnd(0), r1 = bnd(1);
    for (; j + 1 < m; j += 2) {
        const uint2 q0 = bnd(j + 2), q1 = bnd(j + 3);
        if constexpr (!LAST) {
            s0[(ptrdiff_t)(j - 2) * 64] = o0;
            s0[(ptrdiff_t)(j - 1) * 64] = o1;
        }
        __builtin_amdgcn_sched_barrier(0);
        pair(j, r0, r1);
        __builtin_amdgcn_sched_barrier(0);  // the copies (and their wait) stay at the end
        r0 = q0;
        r1 = q1;
    }
    if constexpr (!LAST) {
        s0[(ptrdiff_t)(j - 2) * 64] = o0;
        s0[(ptrdiff_t)(j - 1) * 64] = o1;
    }
    if (j < m) {
        o0 = col(sa, r0);
        if constexpr (!LAST) s0[(size_t)j * 64] = o0;
    }

    } else {
    uint2 r0 = bnd(0), r1 = bnd(1);
    // Four columns per iteration as two pairs whose boundary registers swap
    // roles (r for the first pair, q for the second), so no loaded register
    // is copied: a copy would need its load complete where it is made.
    for (; j + 3 < m; j += 4) {
        const uint2 q0 = bnd(j + 2), q1 = bnd(j + 3);
        if constexpr (!LAST) {
            s0[(ptrdiff_t)(j - 2) * 64] = o0;
            s0[(ptrdiff_t)(j - 1) * 64] = o1;
        }
        __builtin_amdgcn_sched_barrier(0);
        pair(j, r0, r1);
        __builtin_amdgcn_sched_barrier(0);
        r0 = bnd(j + 4);
        r1 = bnd(j + 5);
        if constexpr (!LAST) {
            s0[(ptrdiff_t)j * 64] = o0;
            s0[(ptrdiff_t)(j + 1) * 64] = o1;
        }
        __builtin_amdgcn_sched_barrier(0);
        pair(j + 2, q0, q1);
        __builtin_amdgcn_sched_barrier(0);
    }
    // 0-3 columns left: one more pair, then a single column
    if (j + 1 < m) {
        const uint2 q0 = bnd(j + 2), q1 = bnd(j + 3);
        if constexpr (!LAST) {
            s0[(ptrdiff_t)(j - 2) * 64] = o0;
            s0[(ptrdiff_t)(j - 1) * 64] = o1;
        }
        pair(j, r0, r1);
        r0 = q0;
        r1 = q1;
        j += 2;
    }
    if constexpr (!LAST) {
        s0[(ptrdiff_t)(j - 2) * 64] = o0;
        s0[(ptrdiff_t)(j - 1) * 64] = o1;
    }
    if (j < m) {
        o0 = col(sa, r0);
        if constexpr (!LAST) s0[(size_t)j * 64] = o0;
    }

    }

    if constexpr (CS) best = P::add(best, us(P::splat(se * (m - 1))));  // column m - 1's space to H + o
    if constexpr (!LOCAL) {
        // the last column's real rows, in Ho = H + o space (SH: P + e r,
        // the e r again a chain from an opaque zero)
        v2s k = P::splat(0);
        if constexpr (SH) asm volatile("" : "+v"(k));
#pragma unroll
        for (int r = 0; r < RR; ++r) {
            v2s v = SH ? P::add(Ho[r], k) : Ho[r];
            if constexpr (SH) {
                if (r + 1 < RR) k = P::add(k, vse);
            }
            if (row0 + r < pA) P::neg_lo(v);
            if (row0 + r < pB) P::neg_hi(v);
            best = P::max(best, v);
        }
        if constexpr (SH && LAST) best = P::max(best, P::add(bestL, k));  // k = e (RR - 1)
    }
}

template <int RR, bool LOCAL, bool FIRST, bool QUAD, bool F16, bool OCT, bool SH = false, bool CS = false>
__device__ __forceinline__ void stripe16d_any(const uint4 *prof, uint2 *scr, const int lane,
                                              const cwords w2, const cwords wn, const int m,
                                              const bool last, const int row0, const int pA,
                                              const int pB, const typename Pk<F16>::T vo,
                                              const typename Pk<F16>::T ve, typename Pk<F16>::T &best,
                                              const int c = 0) {
    if (last) stripe16d<RR, LOCAL, FIRST, true, QUAD, F16, OCT, SH, CS>(prof, scr, lane, w2, wn, m, row0, pA, pB, vo, ve, best, c);
    else stripe16d<RR, LOCAL, FIRST, false, QUAD, F16, OCT, SH, CS>(prof, scr, lane, w2, wn, m, row0, pA, pB, vo, ve, best, c);
}

template <bool LOCAL, int PF>
__global__ __launch_bounds__(64 * kWaves16) void k_dp16(KArgs a) {
    // prof[(c * 4 + quad) * 64 + lane] = rows 4 quad .. 4 quad + 3 of plane c
    __shared__ uint4 prof[5 * 4 * 64];
    __shared__ uint32_t s_item;
    const int lane = threadIdx.x & 63;
    const int wv = uni(threadIdx.x >> 6);
    const DevPanel &P = a.panel;
    const int n_refs = P.n_refs;
    const int n_groups = (n_refs + kWaves16 - 1) / kWaves16;
    const int n16 = a.desc->n16;
    const int n_chunks = (n16 + kChunk - 1) / kChunk;
    const uint32_t n_items = (uint32_t)n_chunks * (uint32_t)n_groups;
    uint2 *scr = a.scratch16 + ((size_t)blockIdx.x * kWaves16 + wv) * (size_t)(P.max_len + kScrPad) * 64;
    // variants 4 and 5: variant 3's schedule, and eight columns per
    // iteration with the boundary four columns ahead, in fp16 arithmetic;
    // 6, 7 and 8: variants 3, 5 and 4 with the row-shifted semiglobal
    // column (column16s), in local mode the column-shifted one (column16c)
    constexpr bool F16 = PF == 4 || PF == 5 || PF == 7 || PF == 8;
    constexpr bool OCT = PF == 5 || PF == 7;
    constexpr bool SH = !LOCAL && PF >= 6;
    constexpr bool CS = LOCAL && PF >= 6;  // local mode: the column shift (column16c)
    using PK = Pk<F16>;
    using V = typename PK::T;
    const V vo = PK::splat(a.sc.o), ve = PK::splat(a.sc.e), NEG = PK::neg();
    // The profile's values, encoded once: s - o for a match and for a
    // mismatch, and -o for the top padding rows (the values prof_val gives).
    const uint32_t eMa = PK::enc(a.sc.ma - a.sc.o), eMi = PK::enc(a.sc.mi - a.sc.o);
    const uint32_t ePre = PK::enc(-a.sc.o);

    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_item = atomicAdd(&a.desc->next16, 1u);
        __syncthreads();
        const uint32_t t = s_item;
        if (t >= n_items) break;
        const int k = (int)(t / (uint32_t)n_groups);
        const int g = (int)(t - (uint32_t)k * (uint32_t)n_groups);
        const int slot = g * kWaves16 + wv;
        const bool active = slot < n_refs;
        const int ref = active ? uni(P.rorder[slot]) : 0;

        const int base = k * kChunk + 2 * lane;
        const int qa = base < n16 ? a.sorted16[base] : -1;
        const int qb = base + 1 < n16 ? a.sorted16[base + 1] : -1;
        const int nA = qa >= 0 ? a.qlen[qa] : 0;
        const int nB = qb >= 0 ? a.qlen[qb] : 0;
        // Longest first: the chunk's first read sets its row plan, S stripes
        // of which the first has RR rows (a multiple of 4) and the rest 16.
        const int nmax = uni(a.qlen[a.sorted16[k * kChunk]]);
        const int S = (nmax + kStripe - 1) / kStripe;
        // variants 2 and 3 take first stripes of any even height, the
        // others multiples of 4 (less top padding: 0.5 rows a read on
        // average against 1.5)
        constexpr int G = PF >= 2 ? 2 : 4;
        const int RR = ((nmax - kStripe * (S - 1) + G - 1) / G) * G;
        const int T = RR + kStripe * (S - 1);
        const int pA = T - nA, pB = T - nB;
        const uint8_t *qA = a.qcodes + (qa >= 0 ? a.qoff[qa] : 0);
        const uint8_t *qB = a.qcodes + (qb >= 0 ? a.qoff[qb] : 0);

        const int m = active ? uni(P.rlen[ref]) : 0;
        const cwords w2 = (cwords)(P.ref2 + uni((int)P.w2off[ref]));
        const cwords wn = (cwords)(P.refN + uni((int)P.nmoff[ref]));

        V best = LOCAL ? vo : NEG;  // Ho = H + o space (local starts at H = 0)

        for (int s = 0; s < S; ++s) {
            const int row0 = s == 0 ? 0 : RR + kStripe * (s - 1);
            const int height = s == 0 ? RR : kStripe;
            // SH: the row above, the previous stripe's bottom row, is lower
            // by e x (the previous stripe's height) in this stripe's space
            const int cadj = (SH && s > 0) ? a.sc.e * (s == 1 ? RR : kStripe) : 0;
            // This stripe's profile: wave wv builds rows row0 + 4 wv .. + 3.
            __syncthreads();
            if (4 * wv < height) {
                uint32_t pl[5][4];
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    const int i = row0 + 4 * wv + rr;
                    // rows past the chunk's plan (a first stripe of 2, 6, 10 or
                    // 14 rows builds whole quads) are never read: padding codes
                    const int xa = (i < pA || i >= T) ? kPreCode : qA[i - pA];
                    const int xb = (i < pB || i >= T) ? kPreCode : qB[i - pB];
                    // the stripe's row 0 carries the shift offset (SH)
                    const bool top = SH && cadj != 0 && i == row0;
                    const uint32_t rMa = top ? PK::enc(a.sc.ma - a.sc.o + cadj) : eMa;
                    const uint32_t rMi = top ? PK::enc(a.sc.mi - a.sc.o + cadj) : eMi;
                    const uint32_t rPre = top ? PK::enc(-a.sc.o + cadj) : ePre;
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const uint32_t ea = xa == kPreCode ? rPre : (xa == c || xa == 4) ? rMa : rMi;
                        const uint32_t eb = xb == kPreCode ? rPre : (xb == c || xb == 4) ? rMa : rMi;
                        pl[c][rr] = ea | (eb << 16);
                    }
                    // plane 4, a wildcard reference base: a match for every
                    // read base (also the best of the four, mismatch < match)
                    pl[4][rr] = (xa == kPreCode ? rPre : rMa) | ((xb == kPreCode ? rPre : rMa) << 16);
                }
#pragma unroll
                for (int c = 0; c < 5; ++c)
                    prof[(c * 4 + wv) * 64 + lane] = make_uint4(pl[c][0], pl[c][1], pl[c][2], pl[c][3]);
            }
            __syncthreads();
            if (!active) continue;
            const bool first = s == 0, last = s + 1 == S;
            if constexpr (PF >= 2) {
                constexpr bool Q = PF == 3 || PF == 4 || PF == 6 || PF == 8;
                if (s > 0) {
                    stripe16d_any<16, LOCAL, false, Q, F16, OCT, SH, CS>(prof, scr, lane, w2, wn, m, last, row0, pA, pB, vo, ve, best, cadj);
                } else {
                    switch (RR) {
#define ESCH_FIRST(H) \
    case H: stripe16d_any<H, LOCAL, true, Q, F16, OCT, SH, CS>(prof, scr, lane, w2, wn, m, last, row0, pA, pB, vo, ve, best); break;
                        ESCH_FIRST(16) ESCH_FIRST(14) ESCH_FIRST(12) ESCH_FIRST(10)
                        ESCH_FIRST(8) ESCH_FIRST(6) ESCH_FIRST(4)
                        default: stripe16d_any<2, LOCAL, true, Q, F16, OCT, SH, CS>(prof, scr, lane, w2, wn, m, last, row0, pA, pB, vo, ve, best);
#undef ESCH_FIRST
                    }
                }
            } else {
                constexpr bool PP = PF == 1;
                if (s > 0) {
                    stripe16<16, LOCAL, PP>(prof, scr, lane, w2, wn, m, false, last, row0, pA, pB, vo, ve, best);
                } else if (RR == 16) {
                    stripe16<16, LOCAL, PP>(prof, scr, lane, w2, wn, m, true, last, row0, pA, pB, vo, ve, best);
                } else if (RR == 12) {
                    stripe16<12, LOCAL, PP>(prof, scr, lane, w2, wn, m, true, last, row0, pA, pB, vo, ve, best);
                } else if (RR == 8) {
                    stripe16<8, LOCAL, PP>(prof, scr, lane, w2, wn, m, true, last, row0, pA, pB, vo, ve, best);
                } else {
                    stripe16<4, LOCAL, PP>(prof, scr, lane, w2, wn, m, true, last, row0, pA, pB, vo, ve, best);
                }
            }
            (void)first;
        }
        if (!active) continue;
        const int bA = PK::lo(best) - a.sc.o, bB = PK::hi(best) - a.sc.o;
        if (qa >= 0) {
            const size_t o = out_index(a, qa, ref);
            if (a.out32) a.out32[o] = bA;
            if (a.out16) a.out16[o] = (int16_t)bA;
        }
        if (qb >= 0) {
            const size_t o = out_index(a, qb, ref);
            if (a.out32) a.out32[o] = bB;
            if (a.out16) a.out16[o] = (int16_t)bB;
        }
    }
}

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
    const int ratio = a.desc->diag_ratio;
    const int64_t items16 = (int64_t)((total + kChunk - 1) / kChunk) * ((a.panel.n_refs + 3) / 4);
    const bool starved = items16 < (int64_t)a.desc->blocks16;
    // routing flags into bin_fill (k_zero cleared it; cleared again below)
    int32_t kept = sum;
    if (ratio > 0) {
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
            const bool thin = num >= 0 && a.bin_start[num / ratio] <= x;  // num / ratio < L
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
        if (ratio > 0 && (a.bin_fill[b] || all)) {
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
        int b = INT32_MIN;
        for (int k = lane; k < nc; k += 64) b = max(b, cand(k));
        b = wave_max(b);
        const int64_t take_lim = max_ties < 0 ? INT64_MAX : max_ties + 1;
        const int lim = (int)min((int64_t)cap, take_lim);
        int nt = 0, so = INT32_MIN, hs = 0, base = 0;
        for (int c0 = 0; c0 < nc; c0 += 64) {
            const int k = c0 + lane;
            int s = 0;
            bool tie = false;
            if (k < nc) {
                s = cand(k);
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
    hipLaunchKernelGGL(k_revcomp, dim3(g.grid_aux), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_classify, dim3(g.grid_aux), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, a);
    hipLaunchKernelGGL(k_scatter, dim3(g.grid_aux), dim3(256), 0, st, a);
    return hipGetLastError();
}

// Kernel variants of k_dp16, chosen per scorer (set by the autotuner).
template <bool LOCAL>
static const void *dp16_fn(int variant) {
    if (variant == 8) return reinterpret_cast<const void *>(&k_dp16<LOCAL, 8>);
    if (variant == 7) return reinterpret_cast<const void *>(&k_dp16<LOCAL, 7>);
    if (variant == 6) return reinterpret_cast<const void *>(&k_dp16<LOCAL, 6>);
    if (variant == 5) return reinterpret_cast<const void *>(&k_dp16<LOCAL, 5>);
    if (variant == 4) return reinterpret_cast<const void *>(&k_dp16<LOCAL, 4>);
    if (variant == 3) return reinterpret_cast<const void *>(&k_dp16<LOCAL, 3>);
    if (variant == 2) return reinterpret_cast<const void *>(&k_dp16<LOCAL, 2>);
    if (variant == 1) return reinterpret_cast<const void *>(&k_dp16<LOCAL, 1>);
    return reinterpret_cast<const void *>(&k_dp16<LOCAL, 0>);
}

hipError_t launch_dp16(const KArgs &a, const Launch &g, bool local, hipStream_t st) {
    const int v = g.variant16;
    const dim3 grid(g.grid16), block(64 * kWaves16);
    if (local) {
        if (v == 8) hipLaunchKernelGGL((k_dp16<true, 8>), grid, block, 0, st, a);
        else if (v == 7) hipLaunchKernelGGL((k_dp16<true, 7>), grid, block, 0, st, a);
        else if (v == 6) hipLaunchKernelGGL((k_dp16<true, 6>), grid, block, 0, st, a);
        else if (v == 5) hipLaunchKernelGGL((k_dp16<true, 5>), grid, block, 0, st, a);
        else if (v == 4) hipLaunchKernelGGL((k_dp16<true, 4>), grid, block, 0, st, a);
        else if (v == 3) hipLaunchKernelGGL((k_dp16<true, 3>), grid, block, 0, st, a);
        else if (v == 2) hipLaunchKernelGGL((k_dp16<true, 2>), grid, block, 0, st, a);
        else if (v == 1) hipLaunchKernelGGL((k_dp16<true, 1>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((k_dp16<true, 0>), grid, block, 0, st, a);
    } else {
        if (v == 8) hipLaunchKernelGGL((k_dp16<false, 8>), grid, block, 0, st, a);
        else if (v == 7) hipLaunchKernelGGL((k_dp16<false, 7>), grid, block, 0, st, a);
        else if (v == 6) hipLaunchKernelGGL((k_dp16<false, 6>), grid, block, 0, st, a);
        else if (v == 5) hipLaunchKernelGGL((k_dp16<false, 5>), grid, block, 0, st, a);
        else if (v == 4) hipLaunchKernelGGL((k_dp16<false, 4>), grid, block, 0, st, a);
        else if (v == 3) hipLaunchKernelGGL((k_dp16<false, 3>), grid, block, 0, st, a);
        else if (v == 2) hipLaunchKernelGGL((k_dp16<false, 2>), grid, block, 0, st, a);
        else if (v == 1) hipLaunchKernelGGL((k_dp16<false, 1>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((k_dp16<false, 0>), grid, block, 0, st, a);
    }
    return hipGetLastError();
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
    hipLaunchKernelGGL(k_select, dim3(g.grid_aux), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_pairs(const KArgs &a, const Launch &g, hipStream_t st) {
    hipLaunchKernelGGL(k_pairs, dim3(g.grid_aux), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t dp16_attributes(bool local, int variant, hipFuncAttributes *attr) {
    return hipFuncGetAttributes(attr, local ? dp16_fn<true>(variant) : dp16_fn<false>(variant));
}

hipError_t dp32_attributes(bool local, hipFuncAttributes *attr) {
    return local ? hipFuncGetAttributes(attr, reinterpret_cast<const void *>(&k_dp32<true>))
                 : hipFuncGetAttributes(attr, reinterpret_cast<const void *>(&k_dp32<false>));
}

hipError_t dp16_occupancy(bool local, int variant, int *blocks) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        blocks, local ? dp16_fn<true>(variant) : dp16_fn<false>(variant), 64 * kWaves16, 0);
}

hipError_t dp32_occupancy(bool local, int block, int *blocks) {
    return local ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k_dp32<true>, block, 0)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k_dp32<false>, block, 0);
}

}  // namespace esch
