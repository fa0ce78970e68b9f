// k_dp16, the batched packed 16-bit score DP, and the device helpers the
// other score kernels share. Included by the per-mode translation units
// (kernels_sg.hip, kernels_lo.hip), which instantiate its variants, and by
// kernels.hip for the helpers.
#pragma once
#include <type_traits>
#include <utility>

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
//
// DS (the diagonal shift): every value of stripe row r in reference column
// j is shifted by -e (r + j) instead, which takes the extension out of both
// gap chains:
//   P = H - e (r + j) + o,  R = E - e (r + j) + e,  Q = F - e (r + j) + e
//   R[j]    = max(P[j-1], R[j-1])
//   Q[r]    = max(P[r-1], Q[r-1])
//   P[r][j] = max(P[r-1][j-1] + (s - o - e), R, Q) + (o - e)
// 5 packed ops per row in fp16 (6 in int16). The column part of the shift
// is the same in every stripe, so the rebase by c is the row shift's; the
// profile carries s - o - e.
template <int RR, bool ADJ, bool F16, bool DS = false>
__device__ __forceinline__ void column16s(const typename Pk<F16>::T (&sc)[RR], typename Pk<F16>::T (&Pv)[RR],
                                          typename Pk<F16>::T (&R)[RR], typename Pk<F16>::T &pu,
                                          typename Pk<F16>::T &q, typename Pk<F16>::T pd,
                                          typename Pk<F16>::T voe, typename Pk<F16>::T ve,
                                          typename Pk<F16>::T vc) {
    using P = Pk<F16>;
#pragma unroll
    for (int r = 0; r < RR; ++r) {
        R[r] = DS ? P::max(Pv[r], R[r]) : P::add(P::max(Pv[r], R[r]), ve);
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
//   R[j]    = max(P[j-1], R[j-1], z_j)
//   Q[r]    = max(P[r-1], Q[r-1]) + e
//   P[r][j] = max(P[r-1][j-1] + (s - o), R, Q) + (o - e)
// where z_j = e (1 - j) is the zero floor H >= 0 in this space: one value
// per column, shared by every row (a row shift would need one per row, and a
// best per row). The floor rides on the horizontal chain: carrying
// E' = max(E, 0) changes no H (by induction E'[j] = max(H[j-1] + o,
// E'[j-1] + e, 0) = max(E[j], 0), e <= 0), and the cell's maximum then
// includes it. 6 packed ops per row in fp16 (R's three-input maximum is one
// op) and 8 in int16; a column's rows share one shift, so their best folds
// by a tree of three-way maxima as before (1/2 op a row in fp16, 1 in
// int16): 6.5 and 9 per cell where the unshifted local column takes 8.5 and
// 10. Returns the column's best P (column j space).
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
        R[r] = P::max3(Pv[r], R[r], z);
        q = P::add(P::max(pu, q), ve);
        const typename P::T h = P::max3(P::add(pd, sc[r]), R[r], q);
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
        R[r] = P::max3(Pv[r], R[r], z);  // the floor on the horizontal chain
        q = P::add(P::max(pu, q), ve);
        const T h = P::max3(P::add(pd, sc[r]), R[r], q);
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

// Calls f(std::integral_constant<int, I>) for I = B .. E - 1 in order: a
// loop unrolled by construction, whatever its body's size, so that row
// indices stay compile-time constants (a loop the unroller declines would
// index the register arrays at run time).
template <int B, typename F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, B + I>{}), ...);
}
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (E > B) static_for_impl<B>(f, std::make_integer_sequence<int, E - B>{});
}

// The profile reloads of a windowed column pair (column16s_pairw,
// column16c_pairw), issued once a column has computed row r: its own quad
// k + LA once quad k is done (LA quads ahead), and the next pair's first LA
// quads as its last LA quads finish, into the registers those first quads
// left. A column's plane then occupies about LA + 1 quads of registers
// rather than all of its rows, which leaves room for taller stripes or one
// more resident wave.
template <int RR, int LA, bool F16, int r>
__device__ __forceinline__ void window_loads(const uint4 *cur, const uint4 *nxt, typename Pk<F16>::T (&sc)[RR]) {
    constexpr int NP = (RR + 3) / 4;
    constexpr int SKIP = NP > LA ? NP - LA : 0;
    constexpr int k = r == RR - 1 ? NP - 1 : r / 4;  // the quad row r ends, if it ends one
    if constexpr (r == RR - 1 || r % 4 == 3) {
        if constexpr (k + LA < NP) plane_part<RR, F16>(cur, k + LA, sc);
        if constexpr (k >= SKIP && k - SKIP < LA) plane_part<RR, F16>(nxt, k - SKIP, sc);
    }
}

// column16s_pair with the windowed profile: on entry quads 0 .. LA - 1 of
// both columns' planes (cA, cB) are loaded; on exit those of the next pair's
// (nA, nB). Each cell's arithmetic is column16s' (DS: the diagonal shift).
template <int RR, bool ADJ, bool F16, int LA, bool DS = false>
__device__ __forceinline__ void column16s_pairw(typename Pk<F16>::T (&scA)[RR], typename Pk<F16>::T (&scB)[RR],
                                                typename Pk<F16>::T (&Pv)[RR], typename Pk<F16>::T (&R)[RR],
                                                typename Pk<F16>::T &puA, typename Pk<F16>::T &qA,
                                                typename Pk<F16>::T pdA, typename Pk<F16>::T &puB,
                                                typename Pk<F16>::T &qB, typename Pk<F16>::T pdB,
                                                typename Pk<F16>::T voe, typename Pk<F16>::T ve,
                                                typename Pk<F16>::T vc, const uint4 *cA, const uint4 *cB,
                                                const uint4 *nA, const uint4 *nB) {
    using P = Pk<F16>;
    using T = typename P::T;
    auto row = [&](auto rc, const T (&sc)[RR], T &pu, T &q, T &pd) {
        constexpr int r = decltype(rc)::value;
        if constexpr (DS) R[r] = P::max(Pv[r], R[r]);
        else R[r] = P::add(P::max(Pv[r], R[r]), ve);
        q = P::max(pu, q);
        if constexpr (ADJ && r == 0) q = P::add(q, vc);
        const T h = P::max3(P::add(pd, sc[r]), R[r], q);
        pd = Pv[r];
        Pv[r] = P::add(h, voe);
        pu = Pv[r];
    };
    // step r: A's row r, then B's row r - 1 (one row behind)
    static_for<0, RR + 1>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        if constexpr (r < RR) {
            row(std::integral_constant<int, r>{}, scA, puA, qA, pdA);
            window_loads<RR, LA, F16, r>(cA, nA, scA);
        }
        if constexpr (r >= 1) {
            row(std::integral_constant<int, r - 1>{}, scB, puB, qB, pdB);
            window_loads<RR, LA, F16, r - 1>(cB, nB, scB);
        }
    });
}

// column16c_pair with the windowed profile (as column16s_pairw).
template <int RR, bool F16, int LA>
__device__ __forceinline__ void column16c_pairw(typename Pk<F16>::T (&scA)[RR], typename Pk<F16>::T (&scB)[RR],
                                                typename Pk<F16>::T (&Pv)[RR], typename Pk<F16>::T (&R)[RR],
                                                typename Pk<F16>::T &puA, typename Pk<F16>::T &qA,
                                                typename Pk<F16>::T pdA, typename Pk<F16>::T &puB,
                                                typename Pk<F16>::T &qB, typename Pk<F16>::T pdB,
                                                typename Pk<F16>::T voe, typename Pk<F16>::T ve,
                                                typename Pk<F16>::T zA, typename Pk<F16>::T zB,
                                                typename Pk<F16>::T &tA, typename Pk<F16>::T &tB,
                                                const uint4 *cA, const uint4 *cB, const uint4 *nA,
                                                const uint4 *nB) {
    using P = Pk<F16>;
    using T = typename P::T;
    auto row = [&](auto rc, const T (&sc)[RR], T &pu, T &q, T &pd, T z) {
        constexpr int r = decltype(rc)::value;
        R[r] = P::max3(Pv[r], R[r], z);  // the floor on the horizontal chain
        q = P::add(P::max(pu, q), ve);
        const T h = P::max3(P::add(pd, sc[r]), R[r], q);
        pd = Pv[r];
        Pv[r] = P::add(h, voe);
        pu = Pv[r];
    };
    // rows r - 1 and r (r odd) into a column's running best
    auto fold = [&](auto rc, T &t) {
        constexpr int r = decltype(rc)::value;
        if constexpr (r == 1) t = P::max(Pv[0], Pv[1]);
        else t = P::max3(t, Pv[r - 1], Pv[r]);
    };
    static_for<0, RR + 1>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        if constexpr (r < RR) {
            row(std::integral_constant<int, r>{}, scA, puA, qA, pdA, zA);
            window_loads<RR, LA, F16, r>(cA, nA, scA);
            if constexpr (r & 1) fold(std::integral_constant<int, r>{}, tA);  // before B overwrites row r - 1
        }
        if constexpr (r >= 1) {
            row(std::integral_constant<int, r - 1>{}, scB, puB, qB, pdB, zB);
            window_loads<RR, LA, F16, r - 1>(cB, nB, scB);
            if constexpr ((r - 1) & 1) fold(std::integral_constant<int, r - 1>{}, tB);
        }
    });
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
// `best` carried in column j's space through the stripe. DS (with SH and
// the windowed profile) shifts by row and column at once (column16s' DS):
// the first stripe's top border P = o + e - e j is advanced per column as
// CS's, the last stripe's bottom-row best is carried in column j's space,
// and the last column's rows come back to H + o through + e (r + m - 1).
// CS and DS restart the column part of the shift every cm + 1 columns (cm
// = -1: never): at a column j with j & cm == 0 every value carried in from
// column j - 1 (the rows' P and R, the corner, the top border, the floor
// and the carried bests) moves by + e (cm + 1) into the new block's space,
// and the stripe's bests return to H + o through column m - 1's place in
// its block. Every stripe restarts at the same columns, so the boundary a
// stripe leaves for column j is in the space the next one reads it in.
template <int RR, bool LOCAL, bool FIRST, bool LAST, bool QUAD, bool F16, bool OCT = false, bool SH = false,
          bool CS = false, int PQ = 4, int LA = 0, bool DS = false>
__device__ __forceinline__ void stripe16d(const uint4 *__restrict__ prof, uint2 *__restrict__ scr,
                                          const int lane, const cwords w2, const cwords wn,
                                          const int m, const int row0, const int pA, const int pB,
                                          const typename Pk<F16>::T vo, const typename Pk<F16>::T ve,
                                          typename Pk<F16>::T &best, const int c = 0, const int cm = -1) {
    static_assert(RR % 2 == 0 && RR >= 2 && RR <= 4 * PQ, "an even number of rows, at most the plane's");
    static_assert(!(SH && LOCAL), "the row shift is a semiglobal form");
    static_assert(!(CS && !LOCAL), "the column shift is a local form");
    static_assert(!DS || (SH && LA > 0), "the diagonal shift extends the windowed row-shifted schedules");
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
    // (column 0's space) to column -1's. DS: the top border's P = o + e - e j.
    v2s vtop = P::splat(DS ? so + se : so), z = P::splat(se);
    const v2s vn2e = us(P::splat(-2 * se));
    if constexpr (CS) best = P::add(best, us(P::splat(se)));
    v2s Ho[RR], E[RR];
    // SH: the left border's o - e r (DS: o + e - e r, column -1) as a chain
    // of adds from a base the compiler cannot see through, so the RR
    // constants are not hoisted out of the stripe loop into registers held
    // for the whole kernel
    v2s t = DS ? P::splat(so + se) : vo;
    if constexpr (SH) asm volatile("" : "+v"(t));
#pragma unroll
    for (int r = 0; r < RR; ++r) {
        Ho[r] = CS ? us(P::splat(so + se)) : t;  // H = 0 on the left border
        if constexpr (SH) t = P::add(t, vne);
        E[r] = NEG;
    }
    // the corner above column 0's row 0, in the space of the row above
    // (lower by c after a first stripe)
    v2s hdt = DS ? P::splat(so + 2 * se - c) : SH ? P::splat(so + se - c) : CS ? us(P::splat(so + se)) : vo;
    // CS / DS: the start of a column block (see above); wave-uniform
    auto rebase = [&](int j) {
        if constexpr (CS || DS) {
            if (j == 0 || (j & cm) != 0) return;
            const v2s vk = us(P::splat(se * (cm + 1)));
#pragma unroll
            for (int r = 0; r < RR; ++r) {
                Ho[r] = P::add(Ho[r], vk);
                E[r] = P::add(E[r], vk);
            }
            hdt = P::add(hdt, vk);
            if constexpr (FIRST) vtop = P::add(vtop, vk);
            if constexpr (CS) {
                z = P::add(z, vk);
                best = P::add(best, vk);
            }
            if constexpr (DS && LAST) bestL = P::add(bestL, vk);
        }
    };
    auto code = [&](int j) -> int {
        const uint32_t c2 = (w2[j >> 4] >> (2 * (j & 15))) & 3u;
        return ((wn[j >> 5] >> (j & 31)) & 1u) ? 4 : (int)c2;
    };
    auto plane = [&](int c, v2s (&sc)[RR]) {
        const uint4 *pc = prof + c * PQ * 64 + lane;
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
        if constexpr ((CS || DS) && FIRST) {
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
            column16s<RR, !FIRST, F16, DS>(sc, Ho, E, hou, f, hod, voe, vse, vc);
            if constexpr (LAST) bestL = DS ? P::max(P::add(bestL, vne), hou) : P::max(bestL, hou);
        } else {
            column16<RR, LOCAL, F16>(sc, Ho, E, hou, f, hod, vo, ve, best);
            if constexpr (!LOCAL && LAST) best = P::max(best, hou);
        }
        return make_uint2(P::bits(hou), P::bits(f));
    };

    v2s sa[RR], sb[RR];
    // LA > 0 (shifted columns): the windowed profile, each column's plane
    // read LA quads ahead of its rows (column16s_pairw / column16c_pairw);
    // pa / pb are the planes of the pair about to run
    static_assert(LA == 0 || XS, "the windowed profile runs the shifted column pairs");
    const uint4 *pa = prof + code(0) * PQ * 64 + lane;
    const uint4 *pb = prof + code(1 < m ? 1 : m) * PQ * 64 + lane;
    if constexpr (LA > 0) {
#pragma unroll
        for (int q = 0; q < LA && q < (RR + 3) / 4; ++q) {
            plane_part<RR, F16>(pa, q, sa);
            plane_part<RR, F16>(pb, q, sb);
        }
    } else {
        plane(code(0), sa);
        if constexpr (XS) plane(code(1 < m ? 1 : m), sb);
    }
    uint2 o0 = border, o1 = border;
    int j = 0;
    // Columns jj (plane sa) and jj + 1 into o0 and o1. SH: both planes are
    // loaded on entry, and the pair loads those of columns jj + 2 and jj + 3
    // (at most the padding column m) for the next one as it goes.
    auto pair = [&](int jj, uint2 x0, uint2 x1) {
        rebase(jj);
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
            const uint4 *nA = prof + code(ja) * PQ * 64 + lane, *nB = prof + code(jb) * PQ * 64 + lane;
            v2s tA, tB;
            if constexpr (LA > 0) {
                column16c_pairw<RR, F16, LA>(sa, sb, Ho, E, puA, qA, pdA, puB, qB, pdB, voe, vse, zA, zB, tA,
                                             tB, pa, pb, nA, nB);
                pa = nA;
                pb = nB;
            } else {
                column16c_pair<RR, F16>(sa, sb, Ho, E, puA, qA, pdA, puB, qB, pdB, voe, vse, zA, zB, tA, tB, nA,
                                        nB);
            }
            // best from column jj - 1's space to jj + 1's
            best = P::max3(P::add(best, vn2e), P::add(tA, vne), tB);
            o0 = make_uint2(P::bits(puA), P::bits(qA));
            o1 = make_uint2(P::bits(puB), P::bits(qB));
        } else if constexpr (SH) {
            v2s puA = P::from(x0.x), qA = P::from(x0.y), puB = P::from(x1.x), qB = P::from(x1.y);
            if constexpr (DS && FIRST) {
                puA = vtop;
                puB = P::add(vtop, vne);
                vtop = P::add(puB, vne);
            }
            const v2s pdA = hdt, pdB = puA;
            hdt = puB;
            const int ja = jj + 2 < m ? jj + 2 : m, jb = jj + 3 < m ? jj + 3 : m;
            const uint4 *nA = prof + code(ja) * PQ * 64 + lane, *nB = prof + code(jb) * PQ * 64 + lane;
            if constexpr (LA > 0) {
                column16s_pairw<RR, !FIRST, F16, LA, DS>(sa, sb, Ho, E, puA, qA, pdA, puB, qB, pdB, voe, vse, vc,
                                                         pa, pb, nA, nB);
                pa = nA;
                pb = nB;
            } else {
                column16s_pair<RR, !FIRST, F16>(sa, sb, Ho, E, puA, qA, pdA, puB, qB, pdB, voe, vse, vc, nA, nB);
            }
            // DS: bestL from column jj - 1's space to jj + 1's
            if constexpr (LAST && DS) bestL = P::max3(P::add(bestL, vn2e), P::add(puA, vne), puB);
            else if constexpr (LAST) bestL = P::max3(bestL, puA, puB);
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
        if constexpr (LA > 0) plane(code(j), sa);  // the window held its first LA quads only
        rebase(j);
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
        if constexpr (LA > 0) plane(code(j), sa);
        rebase(j);
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
        if constexpr (LA > 0) plane(code(j), sa);
        rebase(j);
        o0 = col(sa, r0);
        if constexpr (!LAST) s0[(size_t)j * 64] = o0;
    }

    }

    // column m - 1's space (its place in its block) to H + o
    if constexpr (CS) best = P::add(best, us(P::splat(se * ((m - 1) & cm))));
    if constexpr (!LOCAL) {
        // the last column's real rows, in Ho = H + o space (SH: P + e r,
        // the e r again a chain from an opaque zero; DS: from e (m - 1))
        v2s k = P::splat(DS ? se * ((m - 1) & cm) : 0);
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
        // k = e (RR - 1) (DS: e (RR - 1 + m - 1), bestL being in column m - 1's space)
        if constexpr (SH && LAST) best = P::max(best, P::add(bestL, k));
    }
}

template <int RR, bool LOCAL, bool FIRST, bool QUAD, bool F16, bool OCT, bool SH = false, bool CS = false, int PQ = 4,
          int LA = 0, bool DS = false>
__device__ __forceinline__ void stripe16d_any(const uint4 *prof, uint2 *scr, const int lane,
                                              const cwords w2, const cwords wn, const int m,
                                              const bool last, const int row0, const int pA,
                                              const int pB, const typename Pk<F16>::T vo,
                                              const typename Pk<F16>::T ve, typename Pk<F16>::T &best,
                                              const int c = 0, const int cm = -1) {
    if (last) stripe16d<RR, LOCAL, FIRST, true, QUAD, F16, OCT, SH, CS, PQ, LA, DS>(prof, scr, lane, w2, wn, m, row0, pA, pB, vo, ve, best, c, cm);
    else stripe16d<RR, LOCAL, FIRST, false, QUAD, F16, OCT, SH, CS, PQ, LA, DS>(prof, scr, lane, w2, wn, m, row0, pA, pB, vo, ve, best, c, cm);
}

template <bool LOCAL, int PF, int ST = kStripe>
__global__ __launch_bounds__(64 * kWaves16) void k_dp16(KArgs a) {
    static_assert(ST == 16 || ST == 24 || ST == 32, "16-, 24- or 32-row stripes");
    static_assert(PF >= 2 || ST == 16, "variants 0 and 1 take 16-row stripes");
    constexpr int PQ = ST / 4;  // quads of rows per profile plane
    // prof[(c * PQ + quad) * 64 + lane] = rows 4 quad .. 4 quad + 3 of plane c
    __shared__ uint4 prof[5 * PQ * 64];
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
    // 9, 10 and 11: schedules 8, 7 and 6 with the windowed profile (each
    // column's plane read one quad ahead of its rows instead of a pair
    // ahead: fewer registers, so taller stripes or one more wave)
    // 12, 13 and 14 (semiglobal only): 9, 10 and 11 with the diagonal shift
    // (column16s' DS: both gap chains without their extension add)
    static_assert(!(LOCAL && PF >= 12), "the diagonal shift is a semiglobal form");
    constexpr bool F16 = PF == 4 || PF == 5 || PF == 7 || PF == 8 || PF == 9 || PF == 10 || PF == 12 || PF == 13;
    constexpr bool OCT = PF == 5 || PF == 7 || PF == 10 || PF == 13;
    constexpr int LA = PF >= 9 ? 1 : 0;
    constexpr bool SH = !LOCAL && PF >= 6;
    constexpr bool CS = LOCAL && PF >= 6;  // local mode: the column shift (column16c)
    constexpr bool DS = !LOCAL && PF >= 12;
    using PK = Pk<F16>;
    using V = typename PK::T;
    const V vo = PK::splat(a.sc.o), ve = PK::splat(a.sc.e), NEG = PK::neg();
    // The profile's values, encoded once: s - o for a match and for a
    // mismatch, and -o for the top padding rows (the values prof_val gives);
    // DS: each less e.
    const int dse = DS ? a.sc.e : 0;
    // the column-shift block mask (CS, DS): restart every 2^cshift columns, -1: never
    const int cm = a.sc.cshift > 0 ? (1 << a.sc.cshift) - 1 : -1;
    const uint32_t eMa = PK::enc(a.sc.ma - a.sc.o - dse), eMi = PK::enc(a.sc.mi - a.sc.o - dse);
    const uint32_t ePre = PK::enc(-a.sc.o - dse);

    // the items this launch takes: all of them, or one slice of a sliced
    // run (items [lo, hi) from the slice's own cursor)
    uint32_t lo = 0, hi = n_items;
    uint32_t *cursor = &a.desc->next16;
    if (a.n_slices > 1) {
        lo = (uint32_t)((uint64_t)n_items * (uint32_t)a.slice / (uint32_t)a.n_slices);
        hi = (uint32_t)((uint64_t)n_items * (uint32_t)(a.slice + 1) / (uint32_t)a.n_slices);
        cursor = &a.desc->next_slice[a.slice];
    }
    if (lo >= hi) return;  // no work items (uniform: before any barrier)
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_item = lo + atomicAdd(cursor, 1u);
        __syncthreads();
        const uint32_t t = s_item;
        if (t >= hi) break;
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
        const int S = (nmax + ST - 1) / ST;
        // variants 2 and 3 take first stripes of any even height, the
        // others multiples of 4 (less top padding: 0.5 rows a read on
        // average against 1.5)
        constexpr int G = PF >= 2 ? 2 : 4;
        const int RR = ((nmax - ST * (S - 1) + G - 1) / G) * G;
        const int T = RR + ST * (S - 1);
        const int pA = T - nA, pB = T - nB;
        const uint8_t *qA = a.qcodes + (qa >= 0 ? a.qoff[qa] : 0);
        const uint8_t *qB = a.qcodes + (qb >= 0 ? a.qoff[qb] : 0);

        const int m = active ? uni(P.rlen[ref]) : 0;
        const cwords w2 = (cwords)(P.ref2 + uni((int)P.w2off[ref]));
        const cwords wn = (cwords)(P.refN + uni((int)P.nmoff[ref]));

        V best = LOCAL ? vo : NEG;  // Ho = H + o space (local starts at H = 0)

        for (int s = 0; s < S; ++s) {
            const int row0 = s == 0 ? 0 : RR + ST * (s - 1);
            const int height = s == 0 ? RR : ST;
            // SH: the row above, the previous stripe's bottom row, is lower
            // by e x (the previous stripe's height) in this stripe's space
            const int cadj = (SH && s > 0) ? a.sc.e * (s == 1 ? RR : ST) : 0;
            // This stripe's profile: wave wv builds the quads wv, wv + 4, ...
            // (rows row0 + 4 quad .. + 3).
            __syncthreads();
#pragma unroll
            for (int qd0 = 0; qd0 < PQ; qd0 += kWaves16) {
            const int qd = qd0 + wv;
            if (4 * qd < height) {
                uint32_t pl[5][4];
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    const int i = row0 + 4 * qd + rr;
                    // rows past the chunk's plan (a first stripe of 2, 6, 10 or
                    // 14 rows builds whole quads) are never read: padding codes
                    const int xa = (i < pA || i >= T) ? kPreCode : qA[i - pA];
                    const int xb = (i < pB || i >= T) ? kPreCode : qB[i - pB];
                    // the stripe's row 0 carries the shift offset (SH)
                    const bool top = SH && cadj != 0 && i == row0;
                    const uint32_t rMa = top ? PK::enc(a.sc.ma - a.sc.o - dse + cadj) : eMa;
                    const uint32_t rMi = top ? PK::enc(a.sc.mi - a.sc.o - dse + cadj) : eMi;
                    const uint32_t rPre = top ? PK::enc(-a.sc.o - dse + cadj) : ePre;
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
                    prof[(c * PQ + qd) * 64 + lane] = make_uint4(pl[c][0], pl[c][1], pl[c][2], pl[c][3]);
            }
            }
            __syncthreads();
            if (!active) continue;
            const bool first = s == 0, last = s + 1 == S;
            if constexpr (PF >= 2) {
                constexpr bool Q = PF == 3 || PF == 4 || PF == 6 || PF == 8 || PF == 9 || PF == 11 || PF == 12 || PF == 14;
                if (s > 0) {
                    stripe16d_any<ST, LOCAL, false, Q, F16, OCT, SH, CS, PQ, LA, DS>(prof, scr, lane, w2, wn, m, last, row0, pA, pB, vo, ve, best, cadj, cm);
                } else {
                    switch (RR) {
#define ESCH_FIRST(H) \
    case H: stripe16d_any<(H <= ST ? H : 2), LOCAL, true, Q, F16, OCT, SH, CS, PQ, LA, DS>(prof, scr, lane, w2, wn, m, last, row0, pA, pB, vo, ve, best, 0, cm); break;
                        ESCH_FIRST(16) ESCH_FIRST(14) ESCH_FIRST(12) ESCH_FIRST(10)
                        ESCH_FIRST(8) ESCH_FIRST(6) ESCH_FIRST(4)
#undef ESCH_FIRST
#define ESCH_FIRST(H) \
    case H: if constexpr (ST >= H) stripe16d_any<(H <= ST ? H : 2), LOCAL, true, Q, F16, OCT, SH, CS, PQ, LA, DS>(prof, scr, lane, w2, wn, m, last, row0, pA, pB, vo, ve, best, 0, cm); break;
                        ESCH_FIRST(32) ESCH_FIRST(30) ESCH_FIRST(28) ESCH_FIRST(26) ESCH_FIRST(24)
                        ESCH_FIRST(22) ESCH_FIRST(20) ESCH_FIRST(18)
#undef ESCH_FIRST
                        default: stripe16d_any<2, LOCAL, true, Q, F16, OCT, SH, CS, PQ, LA, DS>(prof, scr, lane, w2, wn, m, last, row0, pA, pB, vo, ve, best, 0, cm);
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

// The variants of one mode and stripe height: the kernel's host handle and
// its launch. Each (mode, stripe height) has a translation unit of its own
// (kernels_{sg,lo}{,24,32}.hip), so they compile in parallel; the 24-row
// units instantiate the schedules 4-14 (local mode 4-11), the 32-row ones
// 6-14 (local 6-11). The diagonal-shift schedules 12-14 are semiglobal
// forms: in local mode, where the scorer refuses them, their variants name
// the schedules they extend (9-11), so every variant has a code object.
template <bool LOCAL>
constexpr int dp16_mode_schedule(int variant) {
    const int sc = dp16_schedule(variant);
    return LOCAL && sc >= 12 ? sc - 3 : sc;
}

template <bool LOCAL, int ST, int PF>
inline const void *dp16_handle() {
    return reinterpret_cast<const void *>(&k_dp16<LOCAL, PF, ST>);
}

// Calls f(std::integral_constant<int, PF>) for the schedule PF that
// `variant` runs in this mode on ST-row stripes; false when this unit has
// no such instance.
template <bool LOCAL, int ST, typename F>
inline bool dp16_dispatch(int variant, F &&f) {
    const int sc = ST == 16 ? variant : dp16_mode_schedule<LOCAL>(variant);
    if constexpr (ST >= 24 && !LOCAL) {
        switch (sc) {
            case 14: f(std::integral_constant<int, 14>{}); return true;
            case 13: f(std::integral_constant<int, 13>{}); return true;
            case 12: f(std::integral_constant<int, 12>{}); return true;
            default: break;
        }
    }
    if constexpr (ST >= 24) {
        switch (sc) {
            case 11: f(std::integral_constant<int, 11>{}); return true;
            case 10: f(std::integral_constant<int, 10>{}); return true;
            case 9: f(std::integral_constant<int, 9>{}); return true;
            default: break;
        }
    }
    switch (sc) {
        case 8: f(std::integral_constant<int, 8>{}); return true;
        case 7: f(std::integral_constant<int, 7>{}); return true;
        case 6: f(std::integral_constant<int, 6>{}); return true;
        default: break;
    }
    if constexpr (ST <= 24) {
        switch (sc) {
            case 5: f(std::integral_constant<int, 5>{}); return true;
            case 4: f(std::integral_constant<int, 4>{}); return true;
            default: break;
        }
    }
    if constexpr (ST == 16) {
        switch (sc) {
            case 3: f(std::integral_constant<int, 3>{}); return true;
            case 2: f(std::integral_constant<int, 2>{}); return true;
            case 1: f(std::integral_constant<int, 1>{}); return true;
            default: f(std::integral_constant<int, 0>{}); return true;
        }
    }
    return false;
}

template <bool LOCAL, int ST>
inline const void *dp16_fn_impl(int variant) {
    const void *fn = nullptr;
    dp16_dispatch<LOCAL, ST>(variant, [&](auto pf) { fn = dp16_handle<LOCAL, ST, decltype(pf)::value>(); });
    return fn;
}

template <bool LOCAL, int ST>
inline hipError_t launch_dp16_impl(const KArgs &a, const Launch &g, hipStream_t st) {
    const dim3 grid(g.grid16), block(64 * kWaves16);
    const bool ok = dp16_dispatch<LOCAL, ST>(g.variant16, [&](auto pf) {
        hipLaunchKernelGGL((k_dp16<LOCAL, decltype(pf)::value, ST>), grid, block, 0, st, a);
    });
    return ok ? hipGetLastError() : hipErrorInvalidValue;
}

// Defined in kernels_{sg,lo}{,24,32}.hip.
#define ESCH_DP16_UNIT(M, ST)                            \
    const void *dp16_fn_##M##ST(int variant);            \
    hipError_t launch_dp16_##M##ST(const KArgs &a, const Launch &g, hipStream_t st);
ESCH_DP16_UNIT(sg, 16)
ESCH_DP16_UNIT(sg, 24)
ESCH_DP16_UNIT(sg, 32)
ESCH_DP16_UNIT(lo, 16)
ESCH_DP16_UNIT(lo, 24)
ESCH_DP16_UNIT(lo, 32)
#undef ESCH_DP16_UNIT

}  // namespace esch
