// SAM fields of an alignment (lib/sam.rs:32-124): the CIGAR string with the
// soft clips around the aligned span, and MD / NM as `samtools calmd`
// writes them for that CIGAR. Host code over letters; an alignment's runs
// are the library's (op << 30 | length; op 0 = M, 1 = I, 2 = D).
//
// calmd's rules (lib/sam.rs:5-27): a position matches only when both
// letters are A, C, G or T (U counting as T, case ignored) and equal, so a
// reference N or other IUPAC letter is a mismatch against every read base;
// mismatched and deleted reference letters are written upper-cased but
// otherwise as they are; NM = mismatches + inserted + deleted bases; MD puts
// a 0 between adjacent mismatches, after a deletion followed by a mismatch,
// and at the end after a final mismatch.
#include "../../include/eschalign_amd.h"

#include <cstdio>
#include <cstring>
#include <string>

namespace {

inline uint8_t upper(uint8_t b) { return (b >= 'a' && b <= 'z') ? (uint8_t)(b - 32) : b; }
inline uint8_t norm(uint8_t b) {
    const uint8_t u = upper(b);
    return u == 'U' ? (uint8_t)'T' : u;
}
// sam::calmd_match (lib/sam.rs:65-69)
inline bool calmd_match(uint8_t read_base, uint8_t ref_base) {
    const uint8_t q = norm(read_base), r = norm(ref_base);
    return q == r && (r == 'A' || r == 'C' || r == 'G' || r == 'T');
}

size_t put(const std::string &s, char *out, size_t cap) {
    if (out && cap) {
        const size_t k = s.size() < cap - 1 ? s.size() : cap - 1;
        memcpy(out, s.data(), k);
        out[k] = '\0';
    }
    return s.size();
}

}  // namespace

extern "C" {

size_t esch_md_nm(const uint8_t *read, size_t n, const uint8_t *reference, size_t m, size_t query_start,
                  size_t ref_start, const uint32_t *ops, size_t n_ops, char *md, size_t md_cap, uint32_t *nm) {
    std::string s;
    uint32_t edits = 0, run = 0;
    size_t i = query_start, j = ref_start;
    for (size_t k = 0; k < n_ops; ++k) {
        const uint32_t op = ops[k] >> 30, len = ops[k] & 0x3FFFFFFFu;
        if (op == 0u) {
            if (i + len > n || j + len > m) return (size_t)-1;
            for (uint32_t t = 0; t < len; ++t, ++i, ++j) {
                if (calmd_match(read[i], reference[j])) {
                    ++run;
                } else {
                    s += std::to_string(run);
                    s += (char)upper(reference[j]);
                    run = 0;
                    ++edits;
                }
            }
        } else if (op == 1u) {
            if (i + len > n) return (size_t)-1;
            i += len;
            edits += len;
        } else if (op == 2u) {
            if (j + len > m) return (size_t)-1;
            s += std::to_string(run);
            s += '^';
            for (uint32_t t = 0; t < len; ++t) s += (char)upper(reference[j + t]);
            run = 0;
            j += len;
            edits += len;
        } else {
            return (size_t)-1;
        }
    }
    s += std::to_string(run);
    if (nm) *nm = edits;
    return put(s, md, md_cap);
}

size_t esch_cigar_string(const uint32_t *ops, size_t n_ops, size_t query_len, size_t query_start, size_t query_end,
                         char *out, size_t cap) {
    if (query_start > query_end || query_end > query_len) return (size_t)-1;
    static const char kOp[3] = {'M', 'I', 'D'};
    std::string s;
    if (query_start > 0) s += std::to_string(query_start) + "S";
    for (size_t k = 0; k < n_ops; ++k) {
        const uint32_t op = ops[k] >> 30;
        if (op > 2u) return (size_t)-1;
        s += std::to_string(ops[k] & 0x3FFFFFFFu);
        s += kOp[op];
    }
    if (query_len > query_end) s += std::to_string(query_len - query_end) + "S";
    return put(s, out, cap);
}

}  // extern "C"
