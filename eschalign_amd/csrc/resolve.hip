// gfx950 kernels of the mapping tier loop: selections over a batch's
// resident int32 score rows, one 64-lane wave per listed read or pair.
//
//   k_pair_tier     the next tier of a pair's candidates (lib/mapper/pair.rs:
//                   185-231): every eligible (reference, template strand)
//                   whose summed mate score is the largest below a bound
//   k_pair_subopt   per mate, the best score outside a proper pair's tie set
//                   (lib/mapper/pair.rs:300-316)
//   k_mate_select   the unpaired selection of listed reads (lib/mapper.rs:
//                   444-472), for map_reads and the Separate fallback
//
// A wave strides its lanes over the candidates in tie order, reduces with
// cross-lane shuffles, and writes the tie list in order with a ballot and a
// prefix popcount per 64 candidates, so no atomics are involved and a list
// is the same whatever the launch shape. The host loops tier by tier: each
// round lists the pairs still open with the sum of the tier that failed as
// their bound.
#include "esch_internal.h"

namespace esch {

namespace {

__device__ __forceinline__ int wmax(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

struct PairRows {
    const int32_t *m1f, *m2f, *m1r, *m2r;
    __device__ PairRows(const int32_t *out32, int n_reads, int nr, int p, int half) {
        if (half) {  // the forward template strand's two rows only
            m1f = out32 + (size_t)(2 * p) * nr;
            m2r = out32 + (size_t)(2 * p + 1) * nr;
            m1r = m2f = nullptr;
            return;
        }
        m1f = out32 + (size_t)(2 * p) * nr;
        m2f = out32 + (size_t)(2 * p + 1) * nr;
        m1r = out32 + (size_t)(n_reads + 2 * p) * nr;
        m2r = out32 + (size_t)(n_reads + 2 * p + 1) * nr;
    }
    // lib/mapper/pair.rs:135-137: on template strand t mate 1 is
    // reverse-complemented iff t, mate 2 iff !t
    __device__ int32_t mate1(int r, int t) const { return t ? m1r[r] : m1f[r]; }
    __device__ int32_t mate2(int r, int t) const { return t ? m2f[r] : m2r[r]; }
};

__global__ __launch_bounds__(256) void k_pair_tier(TierArgs a) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int waves = (gridDim.x * blockDim.x) >> 6;
    const int nr = a.n_refs, strands = a.strands, nc = nr * strands;
    for (int e = wave; e < a.n_list; e += waves) {
        const PairRows R(a.out32, a.n_reads, nr, a.list[e], a.half);
        const int32_t lim = a.below[e];
        // candidate k: reference k / strands on template strand k % strands
        auto sum_of = [&](int k, bool &ok) -> int32_t {
            const int r = strands == 2 ? (k >> 1) : k;
            const int t = strands == 2 ? (k & 1) : 0;
            const int32_t s1 = R.mate1(r, t), s2 = R.mate2(r, t);
            ok = s1 >= a.min_score && s2 >= a.min_score && s1 + s2 < lim;
            return s1 + s2;
        };
        int32_t b = INT32_MIN;
        int any = 0;
        for (int k = lane; k < nc; k += 64) {
            bool ok;
            const int32_t s = sum_of(k, ok);
            if (ok) {
                b = max(b, s);
                any = 1;
            }
        }
        b = wmax(b);
        any = wmax(any);
        int32_t *rec = a.recs + (size_t)e * kTierRec;
        if (!any) {
            if (lane < kTierRec) rec[lane] = 0;
            continue;
        }
        int32_t *dst = a.cand_off ? a.cands + a.cand_off[e] : a.cands + (size_t)e * a.cap;
        const int cap = a.cand_off ? INT32_MAX : a.cap;
        int base = 0;
        for (int c0 = 0; c0 < nc; c0 += 64) {
            const int k = c0 + lane;
            bool top = false;
            if (k < nc) {
                bool ok;
                const int32_t s = sum_of(k, ok);
                top = ok && s == b;
            }
            const uint64_t mask = __ballot(top);
            const int rank = __popcll(mask & ((1ull << lane) - 1ull));
            if (top && base + rank < cap) {
                const int r = strands == 2 ? (k >> 1) : k;
                const int t = strands == 2 ? (k & 1) : 0;
                dst[base + rank] = (r << 1) | t;
            }
            base += __popcll(mask);
        }
        if (lane == 0) {
            rec[0] = base;
            rec[1] = b;
            rec[2] = 1;
        }
    }
}

__global__ __launch_bounds__(256) void k_pair_subopt(SuboptArgs a) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int waves = (gridDim.x * blockDim.x) >> 6;
    const int nr = a.n_refs, strands = a.strands, nc = nr * strands;
    for (int e = wave; e < a.n_list; e += waves) {
        const PairRows R(a.out32, a.n_reads, nr, a.list[e], a.half);
        const int32_t *set = a.set + a.set_off[e];
        const int ns = a.set_n[e];
        int32_t s1 = INT32_MIN, s2 = INT32_MIN;
        int has = 0;
        for (int k = lane; k < nc; k += 64) {
            const int r = strands == 2 ? (k >> 1) : k;
            const int t = strands == 2 ? (k & 1) : 0;
            const int32_t code = (r << 1) | t;
            int lo = 0, hi = ns;  // the set is ascending: binary search
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (set[mid] < code) lo = mid + 1;
                else hi = mid;
            }
            if (lo < ns && set[lo] == code) continue;
            s1 = max(s1, R.mate1(r, t));
            s2 = max(s2, R.mate2(r, t));
            has = 1;
        }
        s1 = wmax(s1);
        s2 = wmax(s2);
        has = wmax(has);
        if (lane == 0) {
            int32_t *o = a.out + (size_t)e * kSuboptRec;
            o[0] = has ? s1 : 0;
            o[1] = has ? s2 : 0;
            o[2] = has ? 3 : 0;
        }
    }
}

__global__ __launch_bounds__(256) void k_mate_select(MateSelArgs a) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int waves = (gridDim.x * blockDim.x) >> 6;
    const int nr = a.n_refs, both = a.both, nc = nr * (both ? 2 : 1);
    const int64_t take_lim = a.max_ties < 0 ? INT64_MAX : a.max_ties + 1;
    for (int e = wave; e < a.n_list; e += waves) {
        const int x = a.list[e];
        const int32_t *fwd = a.out32 + (size_t)x * nr;
        const int32_t *rev = a.out32 + (size_t)(a.n_reads + x) * nr;
        // tie order: reference r forward, reference r reverse, r + 1, ...
        auto cand = [&](int k) -> int32_t {
            if (!both) return fwd[k];
            return (k & 1) ? rev[k >> 1] : fwd[k >> 1];
        };
        int32_t b = INT32_MIN;
        for (int k = lane; k < nc; k += 64) b = max(b, cand(k));
        b = wmax(b);
        const int64_t take_all = b < a.min_score ? 0 : take_lim;
        int32_t *dst = a.cand_off ? a.ties + a.cand_off[e] : a.ties + (size_t)e * a.cap;
        const int64_t lim = a.cand_off ? take_all : min((int64_t)a.cap, take_all);
        int32_t so = INT32_MIN;
        int hs = 0, base = 0;
        for (int c0 = 0; c0 < nc; c0 += 64) {
            const int k = c0 + lane;
            bool tie = false;
            if (k < nc) {
                const int32_t s = cand(k);
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
                dst[base + rank] = (r << 1) | rv;
            }
            base += __popcll(mask);
        }
        so = wmax(so);
        hs = wmax(hs);
        if (lane == 0) {
            int32_t *rec = a.recs + (size_t)e * kSelRec;
            rec[0] = b;
            rec[1] = base;
            rec[2] = hs ? so : 0;
            rec[3] = 1 | (hs ? 2 : 0);
            rec[4] = (int32_t)min((int64_t)base, take_all);
        }
    }
}

int grid_for(int n_list, int max_blocks) {
    const int blocks = (n_list + 3) / 4;  // 4 waves per 256-thread block
    return blocks < 1 ? 1 : (blocks < max_blocks ? blocks : max_blocks);
}

}  // namespace

hipError_t launch_pair_tier(const TierArgs &a, int max_blocks, hipStream_t st) {
    if (a.n_list <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pair_tier, dim3(grid_for(a.n_list, max_blocks)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_pair_subopt(const SuboptArgs &a, int max_blocks, hipStream_t st) {
    if (a.n_list <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pair_subopt, dim3(grid_for(a.n_list, max_blocks)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_mate_select(const MateSelArgs &a, int max_blocks, hipStream_t st) {
    if (a.n_list <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mate_select, dim3(grid_for(a.n_list, max_blocks)), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace esch
