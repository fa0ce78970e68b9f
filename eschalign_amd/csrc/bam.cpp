// The output records of an unpaired run: build_records
// (crates/eschalign-cli/src/align.rs:1084-1196) over the mappings, each
// record then laid out as a BAM alignment record (SAM/BAM format
// specification, section 4.2) the way the CLI's writer lays it out
// (noodles-bam 0.95 / noodles-sam 0.90, pinned in the reference's
// Cargo.lock). Host code: the input records come in as the uncompressed
// BAM record stream, the output goes out as one.
//
// What a record keeps from its input: the name, the sequence and qualities
// (reverse-complemented / reversed for a reverse-strand primary), the
// QC-fail and duplicate flags, and every data field the run does not own.
// The owned fields (NM, MD, AS, XS, XA) are removed first; a removal moves
// the record's last field into the removed one's place (the writer's data
// buffer removes by swapping). The mate fields are cleared. An unmapped
// read keeps everything else; a mapped one gets its primary hit's
// reference, position, MAPQ (60 unique, 0 tied), CIGAR with soft clips,
// then NM, MD, AS, XS (when a suboptimal score exists) and XA (the other
// hits, bwa's `ref,+-pos,CIGAR,NM;`) appended. With secondary set, each
// other hit also becomes a secondary record (flag 0x100, MAPQ 0, no
// sequence) carrying NM, MD, AS and the read group.
#include "../../include/eschalign_amd.h"
#include "host_internal.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

namespace {

using esch::fail;

constexpr uint16_t kPaired = 0x1, kUnmapped = 0x4, kReverse = 0x10, kSecondary = 0x100, kQcFail = 0x200,
                   kDuplicate = 0x400, kSupplementary = 0x800;

struct Field {
    char tag[2];
    char type;
    std::string value;  // the value's bytes as stored (Z / H with their NUL)
};

struct InRecord {
    std::string name;  // without the NUL; "*" when absent
    uint16_t flag = 0;
    std::string seq;   // letters, =ACMGRSVTWYHKDBN
    std::string qual;  // raw bytes, l_seq of them (0xFF when absent)
    std::vector<Field> data;
};

uint32_t rd32(const uint8_t *p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }
uint16_t rd16(const uint8_t *p) { return (uint16_t)(p[0] | p[1] << 8); }

void put32(std::string &o, uint32_t v) {
    const char b[4] = {(char)(v & 0xFF), (char)(v >> 8 & 0xFF), (char)(v >> 16 & 0xFF), (char)(v >> 24 & 0xFF)};
    o.append(b, 4);
}
void put16(std::string &o, uint16_t v) {
    const char b[2] = {(char)(v & 0xFF), (char)(v >> 8)};
    o.append(b, 2);
}

const char kSeqLetters[] = "=ACMGRSVTWYHKDBN";

int seq_code(uint8_t b) {
    for (int c = 0; c < 16; ++c)
        if ((uint8_t)kSeqLetters[c] == b) return c;
    return 15;  // N
}

// The value bytes of a field of type t at p (end e): its length, or 0 when
// the field runs past the record.
size_t value_len(char t, const uint8_t *p, const uint8_t *e) {
    switch (t) {
        case 'A': case 'c': case 'C': return p + 1 <= e ? 1 : 0;
        case 's': case 'S': return p + 2 <= e ? 2 : 0;
        case 'i': case 'I': case 'f': return p + 4 <= e ? 4 : 0;
        case 'Z': case 'H': {
            const uint8_t *z = p;
            while (z < e && *z) ++z;
            return z < e ? (size_t)(z - p) + 1 : 0;
        }
        case 'B': {
            if (p + 5 > e) return 0;
            size_t w = 0;
            switch ((char)p[0]) {
                case 'c': case 'C': w = 1; break;
                case 's': case 'S': w = 2; break;
                case 'i': case 'I': case 'f': w = 4; break;
                default: return 0;
            }
            const size_t len = 5 + w * (size_t)rd32(p + 1);
            return (size_t)(e - p) >= len ? len : 0;
        }
        default: return 0;
    }
}

bool parse_record(const uint8_t *r, size_t len, InRecord &out) {
    if (len < 32) return false;
    const uint8_t l_name = r[8];
    const uint16_t n_cigar = rd16(r + 12);
    out.flag = rd16(r + 14);
    const uint32_t l_seq = rd32(r + 16);
    size_t p = 32;
    if (l_name == 0 || p + l_name > len) return false;
    out.name.assign((const char *)r + p, l_name - 1);
    p += l_name;
    p += 4 * (size_t)n_cigar;
    const size_t sb = ((size_t)l_seq + 1) / 2;
    if (p + sb + l_seq > len) return false;
    out.seq.resize(l_seq);
    for (uint32_t i = 0; i < l_seq; ++i) {
        const uint8_t b = r[p + i / 2];
        out.seq[i] = kSeqLetters[(i & 1) ? (b & 0xF) : (b >> 4)];
    }
    p += sb;
    out.qual.assign((const char *)r + p, l_seq);
    p += l_seq;
    out.data.clear();
    const uint8_t *e = r + len;
    while (p < len) {
        if (p + 3 > len) return false;
        Field f;
        f.tag[0] = (char)r[p];
        f.tag[1] = (char)r[p + 1];
        f.type = (char)r[p + 2];
        p += 3;
        const size_t vl = value_len(f.type, r + p, e);
        if (!vl) return false;
        f.value.assign((const char *)r + p, vl);
        p += vl;
        out.data.push_back(std::move(f));
    }
    return true;
}

bool is_tag(const Field &f, const char *t) { return f.tag[0] == t[0] && f.tag[1] == t[1]; }

// Remove field `t`: the last field takes its place.
void remove_field(std::vector<Field> &d, const char *t) {
    for (size_t i = 0; i < d.size(); ++i) {
        if (is_tag(d[i], t)) {
            if (i + 1 != d.size()) d[i] = std::move(d.back());
            d.pop_back();
            return;
        }
    }
}

// Insert (replace in place, or append) field t.
void insert_field(std::vector<Field> &d, const char *t, char type, std::string value) {
    for (Field &f : d) {
        if (is_tag(f, t)) {
            f.type = type;
            f.value = std::move(value);
            return;
        }
    }
    Field f;
    f.tag[0] = t[0];
    f.tag[1] = t[1];
    f.type = type;
    f.value = std::move(value);
    d.push_back(std::move(f));
}

// An integer field in its smallest type, unsigned first (the CLI's int_value,
// crates/eschalign-cli/src/align.rs:1001-1015).
void insert_int(std::vector<Field> &d, const char *t, int64_t v) {
    std::string b;
    char type;
    if (v >= 0 && v <= 255) {
        type = 'C';
        b.push_back((char)v);
    } else if (v >= -128 && v < 0) {
        type = 'c';
        b.push_back((char)(int8_t)v);
    } else if (v >= 0 && v <= 65535) {
        type = 'S';
        put16(b, (uint16_t)v);
    } else if (v >= -32768 && v < 0) {
        type = 's';
        put16(b, (uint16_t)(int16_t)v);
    } else if (v >= 0 && v <= 4294967295LL) {
        type = 'I';
        put32(b, (uint32_t)v);
    } else {
        type = 'i';
        put32(b, (uint32_t)(int32_t)v);
    }
    insert_field(d, t, type, b);
}

void insert_string(std::vector<Field> &d, const char *t, const std::string &s) {
    std::string v = s;
    v.push_back('\0');
    insert_field(d, t, 'Z', v);
}

// alphabet::reverse_complement (lib/alphabet.rs:69-80): upper-cased,
// A<->T, C<->G, U -> A, anything else N.
std::string revcomp_letters(const std::string &s) {
    std::string o(s.size(), 'N');
    for (size_t i = 0; i < s.size(); ++i) {
        uint8_t b = (uint8_t)s[s.size() - 1 - i];
        if (b >= 'a' && b <= 'z') b = (uint8_t)(b - 32);
        o[i] = b == 'A' ? 'T' : b == 'C' ? 'G' : b == 'G' ? 'C' : (b == 'T' || b == 'U') ? 'A' : 'N';
    }
    return o;
}

// reg2bin (SAM/BAM specification 5.3) of the 0-based half-open [beg, end).
uint16_t reg2bin(int64_t beg, int64_t end) {
    --end;
    if (beg >> 14 == end >> 14) return (uint16_t)(((1 << 15) - 1) / 7 + (beg >> 14));
    if (beg >> 17 == end >> 17) return (uint16_t)(((1 << 12) - 1) / 7 + (beg >> 17));
    if (beg >> 20 == end >> 20) return (uint16_t)(((1 << 9) - 1) / 7 + (beg >> 20));
    if (beg >> 23 == end >> 23) return (uint16_t)(((1 << 6) - 1) / 7 + (beg >> 23));
    if (beg >> 26 == end >> 26) return (uint16_t)(((1 << 3) - 1) / 7 + (beg >> 26));
    return 0;
}

struct OutRecord {
    std::string name;
    uint16_t flag = 0;
    int32_t ref = -1, pos = -1;  // 0-based position
    uint8_t mapq = 0;
    std::vector<uint32_t> cigar;  // len << 4 | op (M 0, I 1, D 2, S 4)
    std::string seq, qual;
    std::vector<Field> data;
};

void encode(std::string &o, const OutRecord &r) {
    std::string b;
    int64_t span = 0;
    for (uint32_t c : r.cigar) {
        const uint32_t op = c & 0xF;
        if (op == 0 || op == 2 || op == 3 || op == 7 || op == 8) span += c >> 4;
    }
    const uint16_t bin = (r.pos >= 0 && span > 0) ? reg2bin(r.pos, r.pos + span) : 4680;
    put32(b, (uint32_t)r.ref);
    put32(b, (uint32_t)r.pos);
    b.push_back((char)(uint8_t)(r.name.size() + 1));
    b.push_back((char)r.mapq);
    put16(b, bin);
    put16(b, (uint16_t)r.cigar.size());
    put16(b, r.flag);
    put32(b, (uint32_t)r.seq.size());
    put32(b, 0xFFFFFFFFu);  // mate reference: none
    put32(b, 0xFFFFFFFFu);  // mate position: none
    put32(b, 0);            // template length
    b += r.name;
    b.push_back('\0');
    for (uint32_t c : r.cigar) put32(b, c);
    for (size_t i = 0; i < r.seq.size(); i += 2) {
        const int hi = seq_code((uint8_t)r.seq[i]);
        const int lo = i + 1 < r.seq.size() ? seq_code((uint8_t)r.seq[i + 1]) : 0;
        b.push_back((char)(uint8_t)(hi << 4 | lo));
    }
    if (r.qual.size() == r.seq.size()) b += r.qual;
    else b.append(r.seq.size(), (char)0xFF);
    for (const Field &f : r.data) {
        b.push_back(f.tag[0]);
        b.push_back(f.tag[1]);
        b.push_back(f.type);
        b += f.value;
    }
    put32(o, (uint32_t)b.size());
    o += b;
}

// The hit's CIGAR as BAM operations: soft clips around the aligned span
// (crates/eschalign-cli/src/align.rs:1017-1051).
std::vector<uint32_t> bam_cigar(const esch_hit &h, const uint32_t *ops, size_t read_len) {
    std::vector<uint32_t> c;
    const esch_alignment &a = h.alignment;
    if (a.query_start > 0) c.push_back((uint32_t)a.query_start << 4 | 4u);
    for (int32_t k = 0; k < a.n_ops; ++k) {
        const uint32_t v = ops[a.ops_offset + k];
        c.push_back((v & 0x3FFFFFFFu) << 4 | (v >> 30));  // M 0, I 1, D 2 in both encodings
    }
    if (read_len > (size_t)a.query_end) c.push_back((uint32_t)(read_len - (size_t)a.query_end) << 4 | 4u);
    return c;
}

std::string sam_cigar(const esch_hit &h, const uint32_t *ops, size_t read_len) {
    std::string s;
    s.resize(64 + 12 * (size_t)h.alignment.n_ops);
    const size_t n = esch_cigar_string(ops + h.alignment.ops_offset, (size_t)h.alignment.n_ops, read_len,
                                       (size_t)h.alignment.query_start, (size_t)h.alignment.query_end, &s[0],
                                       s.size());
    s.resize(n == (size_t)-1 ? 0 : std::min(n, s.size() - 1));
    return s;
}

}  // namespace

extern "C" esch_status esch_bam_build_records(const uint8_t *in, size_t in_bytes, size_t n,
                                              const esch_read_mapping *maps, const esch_hit *hits, size_t n_hits,
                                              const uint32_t *ops, size_t n_ops, const uint8_t *const *ref_letters,
                                              const size_t *ref_lens, const char *const *ref_names, size_t n_refs,
                                              int secondary, uint8_t *out, size_t out_cap, size_t *out_len,
                                              size_t *n_records) {
    if ((n && (!in || !maps)) || !out_len || (n_refs && (!ref_letters || !ref_lens || !ref_names)))
        return fail(ESCH_INVALID_ARG, "null argument");
    std::string o;
    size_t count = 0, at = 0;
    InRecord rec;
    std::string md;
    for (size_t k = 0; k < n; ++k) {
        if (at + 4 > in_bytes) return fail(ESCH_INVALID_ARG, "input record %zu runs past the stream", k);
        const size_t len = rd32(in + at);
        if (at + 4 + len > in_bytes || !parse_record(in + at + 4, len, rec))
            return fail(ESCH_INVALID_ARG, "input record %zu is not a BAM record", k);
        at += 4 + len;
        if (rec.flag & (kSecondary | kSupplementary))
            return fail(ESCH_INVALID_ARG, "input record %zu is secondary or supplementary (the reader skips those)", k);
        if (!(rec.flag & kUnmapped) && (rec.flag & kReverse))
            return fail(ESCH_INVALID_ARG, "input record %zu is aligned to the reverse strand", k);
        const esch_read_mapping &m = maps[k];
        const uint16_t kept = rec.flag & (kQcFail | kDuplicate);
        for (const char *t : {"NM", "MD", "AS", "XS", "XA"}) remove_field(rec.data, t);
        OutRecord r;
        r.name = rec.name;
        r.data = std::move(rec.data);
        const size_t read_len = rec.seq.size();
        if (m.n_hits <= 0) {
            r.flag = kept | kUnmapped;
            r.seq = std::move(rec.seq);
            r.qual = std::move(rec.qual);
            encode(o, r);
            ++count;
            continue;
        }
        if (m.hit_offset < 0 || (size_t)m.hit_offset + (size_t)m.n_hits > n_hits)
            return fail(ESCH_INVALID_ARG, "mapping %zu names hits outside the array", k);
        const esch_hit *hk = hits + m.hit_offset;
        for (int32_t h = 0; h < m.n_hits; ++h) {
            const esch_alignment &a = hk[h].alignment;
            if (hk[h].reference < 0 || (size_t)hk[h].reference >= n_refs || a.n_ops < 0 || a.ops_offset < 0 ||
                (size_t)a.ops_offset + (size_t)a.n_ops > n_ops || a.query_end > (int32_t)read_len)
                return fail(ESCH_INVALID_ARG, "mapping %zu: hit %d does not fit the read or the panel", k, h);
        }
        const std::string rev_letters =
            [&] {
                for (int32_t h = 0; h < m.n_hits; ++h)
                    if (hk[h].reverse) return revcomp_letters(rec.seq);
                return std::string();
            }();
        // MD / NM of each hit (lib/mapper.rs:520-536), on the read as aligned
        std::vector<std::string> mds((size_t)m.n_hits);
        std::vector<uint32_t> nms((size_t)m.n_hits);
        for (int32_t h = 0; h < m.n_hits; ++h) {
            const esch_hit &x = hk[h];
            const std::string &letters = x.reverse ? rev_letters : rec.seq;
            const uint8_t *refl = ref_letters[x.reference];
            md.resize(64 + 4 * (size_t)x.alignment.n_ops + read_len);
            const size_t ml = esch_md_nm((const uint8_t *)letters.data(), letters.size(), refl,
                                         ref_lens[x.reference], (size_t)x.alignment.query_start,
                                         (size_t)x.alignment.ref_start, ops + x.alignment.ops_offset,
                                         (size_t)x.alignment.n_ops, &md[0], md.size(), &nms[h]);
            if (ml == (size_t)-1) return fail(ESCH_INVALID_ARG, "mapping %zu: hit %d runs past its sequences", k, h);
            if (ml >= md.size()) {
                md.resize(ml + 1);
                (void)esch_md_nm((const uint8_t *)letters.data(), letters.size(), refl, ref_lens[x.reference],
                                 (size_t)x.alignment.query_start, (size_t)x.alignment.ref_start,
                                 ops + x.alignment.ops_offset, (size_t)x.alignment.n_ops, &md[0], md.size(),
                                 &nms[h]);
            }
            mds[h].assign(md.data(), ml);
        }
        const esch_hit &p = hk[0];
        uint16_t flag = kept;
        if (p.reverse) {
            flag |= kReverse;
            r.seq = rev_letters;
            r.qual.assign(rec.qual.rbegin(), rec.qual.rend());
        } else {
            r.seq = std::move(rec.seq);
            r.qual = std::move(rec.qual);
        }
        r.flag = flag;
        r.ref = p.reference;
        r.pos = p.alignment.ref_start;
        r.mapq = m.n_tied == 1 ? 60 : 0;
        r.cigar = bam_cigar(p, ops, read_len);
        insert_int(r.data, "NM", nms[0]);
        insert_string(r.data, "MD", mds[0]);
        insert_int(r.data, "AS", m.best);
        if (m.flags & 2) insert_int(r.data, "XS", m.subopt);
        if (m.n_hits > 1) {
            std::string xa;
            for (int32_t h = 1; h < m.n_hits; ++h) {
                const esch_hit &x = hk[h];
                xa += ref_names[x.reference];
                xa += ',';
                xa += x.reverse ? '-' : '+';
                xa += std::to_string(x.alignment.ref_start + 1);
                xa += ',';
                xa += sam_cigar(x, ops, read_len);
                xa += ',';
                xa += std::to_string(nms[h]);
                xa += ';';
            }
            insert_string(r.data, "XA", xa);
        }
        encode(o, r);
        ++count;
        if (secondary) {
            const Field *rg = nullptr;
            for (const Field &f : r.data)
                if (is_tag(f, "RG")) rg = &f;
            for (int32_t h = 1; h < m.n_hits; ++h) {
                const esch_hit &x = hk[h];
                OutRecord s;
                s.name = r.name;
                s.flag = (uint16_t)(kSecondary | kept | (x.reverse ? kReverse : 0));
                s.ref = x.reference;
                s.pos = x.alignment.ref_start;
                s.mapq = 0;
                s.cigar = bam_cigar(x, ops, read_len);
                insert_int(s.data, "NM", nms[h]);
                insert_string(s.data, "MD", mds[h]);
                insert_int(s.data, "AS", m.best);
                if (rg) insert_field(s.data, "RG", rg->type, rg->value);
                encode(o, s);
                ++count;
            }
        }
    }
    (void)kPaired;
    *out_len = o.size();
    if (n_records) *n_records = count;
    if (o.size() > out_cap || (!out && !o.empty()))
        return fail(ESCH_CAPACITY, "the records take %zu bytes; the buffer holds %zu", o.size(), out_cap);
    if (!o.empty()) memcpy(out, o.data(), o.size());
    return ESCH_OK;
}
