// Seeded synthetic panels and reads for the benchmark workloads.
//
// RNG: xorshift64*, the generator of the reference's own tests
// (crates/eschalign/src/simd/mod.rs:685-699). Error model per template base:
// with p_del delete a run of Geometric(geo) bases (P(k) = geo (1-geo)^(k-1));
// otherwise with p_sub substitute one of the other three bases uniformly;
// after the base, with p_ins insert Geometric(geo) uniform random bases.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed ? seed : 0x9E3779B97F4A7C15ull) {}
    uint64_t next() {
        s ^= s >> 12;
        s ^= s << 25;
        s ^= s >> 27;
        return s * 0x2545F4914F6CDD1Dull;
    }
    uint64_t below(uint64_t n) { return next() % n; }
    double unit() { return (double)(next() >> 11) / (double)(1ull << 53); }
    bool chance(double p) { return unit() < p; }
    int geometric(double p) {
        int k = 1;
        while (!chance(p)) ++k;
        return k;
    }
    double normal() {
        double u1 = unit(), u2 = unit();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

void mutate(Rng &rng, const uint8_t *t, size_t n, double p_sub, double p_ins, double p_del, double geo,
            std::vector<uint8_t> &out) {
    out.clear();
    size_t i = 0;
    while (i < n) {
        if (rng.chance(p_del)) {
            i += (size_t)rng.geometric(geo);
            continue;
        }
        uint8_t b = t[i];
        if (rng.chance(p_sub)) b = (uint8_t)((b + 1 + rng.below(3)) & 3u);
        out.push_back(b);
        if (rng.chance(p_ins)) {
            const int k = rng.geometric(geo);
            for (int x = 0; x < k; ++x) out.push_back((uint8_t)rng.below(4));
        }
        ++i;
    }
}

int draw_len(Rng &rng, int kind, double a, double b, int lo, int hi) {
    if (kind == 0) return lo + (int)rng.below((uint64_t)(hi - lo + 1));
    // lognormal with median a and sigma b, clipped to [lo, hi]
    double v = a * std::exp(b * rng.normal());
    int L = (int)std::lround(v);
    if (L < lo) L = lo;
    if (L > hi) L = hi;
    return L;
}

// template of `L` bases from reference `r` on strand `rev`: a uniform
// substring, or the whole reference with random flanks when L exceeds it
void make_template(Rng &rng, const uint8_t *r, size_t m, int L, bool rev, std::vector<uint8_t> &t) {
    t.clear();
    if ((size_t)L <= m) {
        const size_t start = rng.below(m - (size_t)L + 1);
        t.assign(r + start, r + start + L);
    } else {
        const size_t extra = (size_t)L - m;
        const size_t left = rng.below(extra + 1);
        for (size_t k = 0; k < left; ++k) t.push_back((uint8_t)rng.below(4));
        t.insert(t.end(), r, r + m);
        while (t.size() < (size_t)L) t.push_back((uint8_t)rng.below(4));
    }
    if (rev) {
        std::vector<uint8_t> rc(t.rbegin(), t.rend());
        for (auto &c : rc) c = c == 4 ? 4 : (uint8_t)(3 - c);
        t.swap(rc);
    }
}

}  // namespace

extern "C" {

// n_refs references of uniform ACGT, lengths uniform in [lo, hi].
// Returns total bases written (codes), or -1 if cap is too small.
int64_t esch_synth_panel(uint64_t seed, int32_t n_refs, int32_t lo, int32_t hi, uint8_t *codes,
                         uint64_t cap, uint64_t *off, int32_t *len) {
    Rng rng(seed);
    uint64_t at = 0;
    for (int32_t k = 0; k < n_refs; ++k) {
        const int L = lo + (int)rng.below((uint64_t)(hi - lo + 1));
        if (at + (uint64_t)L > cap) return -1;
        off[k] = at;
        len[k] = L;
        for (int j = 0; j < L; ++j) codes[at + j] = (uint8_t)rng.below(4);
        at += (uint64_t)L;
    }
    return (int64_t)at;
}

// Reads: a fraction frac_ref sampled from a uniformly chosen reference on a
// random strand and passed through the error model, the rest uniform random
// sequence; lengths uniform [lo, hi] (kind 0) or lognormal(median a,
// sigma b) clipped to [lo, hi] (kind 1). Returns total bases or -1.
int64_t esch_synth_reads(const uint8_t *pcodes, const uint64_t *poff, const int32_t *plen, int32_t n_refs,
                         uint64_t seed, int64_t n, int32_t kind, double a, double b, int32_t lo, int32_t hi,
                         double frac_ref, double p_sub, double p_ins, double p_del, double geo,
                         uint8_t *codes, uint64_t cap, uint64_t *off, int32_t *len) {
    Rng rng(seed);
    std::vector<uint8_t> t, m;
    uint64_t at = 0;
    for (int64_t k = 0; k < n; ++k) {
        const int L = draw_len(rng, kind, a, b, lo, hi);
        if (rng.chance(frac_ref)) {
            const int32_t ri = (int32_t)rng.below((uint64_t)n_refs);
            const bool rev = rng.chance(0.5);
            make_template(rng, pcodes + poff[ri], (size_t)plen[ri], L, rev, t);
            mutate(rng, t.data(), t.size(), p_sub, p_ins, p_del, geo, m);
            if (m.empty()) m.push_back((uint8_t)rng.below(4));
        } else {
            m.resize((size_t)L);
            for (int j = 0; j < L; ++j) m[j] = (uint8_t)rng.below(4);
        }
        if (at + m.size() > cap) return -1;
        off[k] = at;
        len[k] = (int32_t)m.size();
        memcpy(codes + at, m.data(), m.size());
        at += m.size();
    }
    return (int64_t)at;
}

// Copies of seed references with independent substitutions at rate p_sub.
void esch_synth_substitute(uint64_t seed, const uint8_t *src, uint8_t *dst, uint64_t n, double p_sub) {
    Rng rng(seed);
    for (uint64_t i = 0; i < n; ++i) {
        uint8_t c = src[i];
        if (c < 4 && rng.chance(p_sub)) c = (uint8_t)((c + 1 + rng.below(3)) & 3u);
        dst[i] = c;
    }
}

// Read pairs of mate length `mate_len` from inserts of length ~N(mu, sd)
// (at least 1): an insert is a substring of a uniformly chosen reference, or
// the whole reference with random flanks when longer; mate 1 reads the
// insert forward, mate 2 its reverse complement, each continuing into its
// adapter (ad1 / ad2, codes) and then random sequence when the insert is
// shorter than the mate; each mate then passes the error model.
// Mates are written interleaved: read 2k = mate 1, 2k + 1 = mate 2.
int64_t esch_synth_pairs(const uint8_t *pcodes, const uint64_t *poff, const int32_t *plen, int32_t n_refs,
                         uint64_t seed, int64_t n_pairs, int32_t mate_len, double mu, double sd,
                         const uint8_t *ad1, int32_t la1, const uint8_t *ad2, int32_t la2, double p_sub,
                         double p_indel, double geo, uint8_t *codes, uint64_t cap, uint64_t *off, int32_t *len) {
    Rng rng(seed);
    std::vector<uint8_t> ins, t, m;
    uint64_t at = 0;
    for (int64_t k = 0; k < n_pairs; ++k) {
        int L = (int)std::lround(mu + sd * rng.normal());
        if (L < 1) L = 1;
        const int32_t ri = (int32_t)rng.below((uint64_t)n_refs);
        make_template(rng, pcodes + poff[ri], (size_t)plen[ri], L, false, ins);
        for (int mate = 0; mate < 2; ++mate) {
            t.clear();
            if (mate == 0) {
                t = ins;
                t.insert(t.end(), ad1, ad1 + la1);
            } else {
                t.assign(ins.rbegin(), ins.rend());
                for (auto &c : t) c = c == 4 ? 4 : (uint8_t)(3 - c);
                t.insert(t.end(), ad2, ad2 + la2);
            }
            while (t.size() < (size_t)mate_len) t.push_back((uint8_t)rng.below(4));
            t.resize((size_t)mate_len);
            mutate(rng, t.data(), t.size(), p_sub, p_indel, p_indel, geo, m);
            if (m.size() > (size_t)mate_len) m.resize((size_t)mate_len);
            if (m.empty()) m.push_back((uint8_t)rng.below(4));
            if (at + m.size() > cap) return -1;
            off[2 * k + mate] = at;
            len[2 * k + mate] = (int32_t)m.size();
            memcpy(codes + at, m.data(), m.size());
            at += m.size();
        }
    }
    return (int64_t)at;
}

}  // extern "C"
