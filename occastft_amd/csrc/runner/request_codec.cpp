// request_codec.cpp -- see request_codec.hpp.
#include "request_codec.hpp"

#include <cstring>

namespace ostft_wire {
namespace {

constexpr uint32_t kDataWords = 6;
constexpr uint32_t kPointers = 4;
constexpr uint64_t kMaxListElems = (1ull << 29) - 1;

struct Segments {
    std::vector<const uint8_t*> base;
    std::vector<uint64_t> words;
};

uint64_t load64(const uint8_t* p)
{
    uint64_t v;
    std::memcpy(&v, p, 8);
    return v;
}

bool word_at(const Segments& m, uint32_t seg, uint64_t idx, uint64_t& v)
{
    if (seg >= m.base.size() || idx >= m.words[seg])
        return false;
    v = load64(m.base[seg] + 8 * idx);
    return true;
}

int64_t signed_offset(uint64_t p)
{
    int64_t off = (int64_t)((p >> 2) & 0x3FFFFFFFull);
    if (off & (1ll << 29))
        off -= (1ll << 30);
    return off;
}

// A resolved (non-far) pointer: where its content starts and what it is.
struct Target {
    bool null = true;
    uint32_t kind = 0;  // 0 struct, 1 list
    uint32_t seg = 0;
    uint64_t start = 0;  // word index of the content
    uint32_t dsize = 0, psize = 0;  // struct
    uint32_t esize = 0;             // list element size tag
    uint64_t count = 0;             // list elements
};

bool resolve(const Segments& m, uint32_t seg, uint64_t at, Target& t, std::string& err)
{
    uint64_t p;
    if (!word_at(m, seg, at, p)) {
        err = "pointer outside its segment";
        return false;
    }
    t = Target();
    if (p == 0)
        return true;
    uint32_t kind = (uint32_t)(p & 3);
    uint64_t tag = p;       // the pointer that describes the content
    uint32_t cseg = seg;    // segment of the content
    int64_t cstart = 0;     // word index of the content
    if (kind == 2) {
        const bool dbl = (p >> 2) & 1;
        const uint64_t pad = (p >> 3) & 0x1FFFFFFFull;
        const uint32_t pseg = (uint32_t)(p >> 32);
        uint64_t land;
        if (!word_at(m, pseg, pad, land)) {
            err = "far pointer landing pad out of range";
            return false;
        }
        if (!dbl) {
            if ((land & 3) == 2) {
                err = "far pointer lands on a far pointer";
                return false;
            }
            tag = land;
            cseg = pseg;
            cstart = (int64_t)pad + 1 + signed_offset(land);
        } else {
            uint64_t tagw;
            if ((land & 3) != 2 || ((land >> 2) & 1) || !word_at(m, pseg, pad + 1, tagw)) {
                err = "malformed double-far landing pad";
                return false;
            }
            tag = tagw;
            cseg = (uint32_t)(land >> 32);
            cstart = (int64_t)((land >> 3) & 0x1FFFFFFFull);
        }
        kind = (uint32_t)(tag & 3);
    } else {
        cstart = (int64_t)at + 1 + signed_offset(p);
    }
    if (kind != 0 && kind != 1) {
        err = "unsupported pointer kind";
        return false;
    }
    if (cseg >= m.base.size() || cstart < 0) {
        err = "pointer target out of range";
        return false;
    }
    t.null = false;
    t.kind = kind;
    t.seg = cseg;
    t.start = (uint64_t)cstart;
    uint64_t words = 0;
    if (kind == 0) {
        t.dsize = (uint32_t)((tag >> 32) & 0xFFFF);
        t.psize = (uint32_t)(tag >> 48);
        words = (uint64_t)t.dsize + t.psize;
    } else {
        t.esize = (uint32_t)((tag >> 32) & 7);
        t.count = tag >> 35;
        static const uint32_t bits[8] = {0, 1, 8, 16, 32, 64, 64, 0};
        if (t.esize == 7) {
            words = t.count;  // composite: count is the word count
        } else {
            words = (t.count * bits[t.esize] + 63) / 64;
        }
    }
    if (t.start + words > m.words[cseg] || t.start + words < t.start) {
        err = "pointer content exceeds its segment";
        return false;
    }
    return true;
}

bool read_text(const Segments& m, const Target& t, std::string& s, std::string& err)
{
    s.clear();
    if (t.null)
        return true;
    if (t.kind != 1 || t.esize != 2) {
        err = "text field is not a byte list";
        return false;
    }
    const char* p = reinterpret_cast<const char*>(m.base[t.seg] + 8 * t.start);
    uint64_t n = t.count;
    if (n > 0 && p[n - 1] == '\0')
        --n;
    s.assign(p, (size_t)n);
    const size_t z = s.find('\0');
    if (z != std::string::npos)
        s.resize(z);
    return true;
}

uint64_t data_u64(const Segments& m, const Target& st, uint32_t byte_off, uint32_t bytes)
{
    if (byte_off + bytes > 8ull * st.dsize)
        return 0;  // field beyond the sender's data section: default (0)
    const uint8_t* p = m.base[st.seg] + 8 * st.start + byte_off;
    uint64_t v = 0;
    std::memcpy(&v, p, bytes);
    return v;
}

}  // namespace

bool decode_request(const uint8_t* buf, size_t len, Request& out, std::string& err)
{
    out = Request();
    if (len < 8 || (len & 7)) {
        err = "message is not a whole number of words";
        return false;
    }
    uint32_t nseg_m1;
    std::memcpy(&nseg_m1, buf, 4);
    const uint64_t nseg = (uint64_t)nseg_m1 + 1;
    if (nseg > 512) {
        err = "too many segments";
        return false;
    }
    const uint64_t table = ((4 + 4 * nseg) + 7) & ~7ull;
    if (table > len) {
        err = "truncated segment table";
        return false;
    }
    Segments m;
    uint64_t off = table;
    for (uint64_t i = 0; i < nseg; ++i) {
        uint32_t w;
        std::memcpy(&w, buf + 4 + 4 * i, 4);
        if (off + 8ull * w > len) {
            err = "truncated segment";
            return false;
        }
        m.base.push_back(buf + off);
        m.words.push_back(w);
        off += 8ull * w;
    }
    Target root;
    if (!resolve(m, 0, 0, root, err))
        return false;
    if (root.null || root.kind != 0) {
        err = "root is not a struct";
        return false;
    }
    out.memPTR = data_u64(m, root, 0, 8);
    out.posixFileDes = (int64_t)data_u64(m, root, 8, 8);
    out.windowsHandlePTR = data_u64(m, root, 16, 8);
    out.windowRadix = (uint32_t)data_u64(m, root, 24, 4);
    const uint32_t ov_bits = (uint32_t)data_u64(m, root, 28, 4);
    std::memcpy(&out.overlapRatio, &ov_bits, 4);
    out.dataLength = data_u64(m, root, 32, 8);
    out.overlapdataLength = data_u64(m, root, 40, 8);

    Target ptr[kPointers];
    for (uint32_t i = 0; i < kPointers; ++i) {
        if (i < root.psize) {
            if (!resolve(m, root.seg, root.start + root.dsize + i, ptr[i], err))
                return false;
        }
    }
    if (!read_text(m, ptr[0], out.sharedMemory, err) || !read_text(m, ptr[2], out.mappedID, err)
        || !read_text(m, ptr[3], out.options, err))
        return false;
    const Target& d = ptr[1];
    if (!d.null) {
        if (d.kind != 1 || d.esize != 4) {
            err = "data is not a List(Float32)";
            return false;
        }
        out.data.resize((size_t)d.count);
        if (d.count)  // an empty list: nothing to copy (and no buffer to copy to)
            std::memcpy(out.data.data(), m.base[d.seg] + 8 * d.start, 4 * (size_t)d.count);
    }
    return true;
}

namespace {

uint64_t list_ptr(uint64_t at, uint64_t content, uint32_t esize, uint64_t count)
{
    const int64_t offset = (int64_t)content - (int64_t)at - 1;
    return 1ull | (((uint64_t)offset & 0x3FFFFFFFull) << 2) | ((uint64_t)esize << 32) | (count << 35);
}

uint64_t text_words(const std::string& s) { return (s.size() + 1 + 7) / 8; }

}  // namespace

bool encode_request(const Request& r, const float* data, uint64_t n, std::string& out)
{
    if (n > kMaxListElems)
        return false;
    const uint64_t struct_at = 1;
    const uint64_t ptr_at = struct_at + kDataWords;
    uint64_t cursor = ptr_at + kPointers;
    const uint64_t sm_at = cursor;
    cursor += text_words(r.sharedMemory);
    const uint64_t id_at = cursor;
    cursor += text_words(r.mappedID);
    const uint64_t op_at = cursor;
    cursor += text_words(r.options);
    const uint64_t data_at = cursor;
    cursor += (n * 4 + 7) / 8;
    const uint64_t words = cursor;
    if (words > 0xFFFFFFFFull)
        return false;

    out.assign(8 + 8 * words, '\0');
    uint8_t* b = reinterpret_cast<uint8_t*>(&out[0]);
    const uint32_t nseg_m1 = 0, w32 = (uint32_t)words;
    std::memcpy(b, &nseg_m1, 4);
    std::memcpy(b + 4, &w32, 4);
    uint8_t* seg = b + 8;
    auto put64 = [&](uint64_t word, uint64_t v) { std::memcpy(seg + 8 * word, &v, 8); };
    // root struct pointer: offset 0, 6 data words, 4 pointers
    put64(0, 0ull | (0ull << 2) | ((uint64_t)kDataWords << 32) | ((uint64_t)kPointers << 48));
    uint8_t* ds = seg + 8 * struct_at;
    std::memcpy(ds + 0, &r.memPTR, 8);
    std::memcpy(ds + 8, &r.posixFileDes, 8);
    std::memcpy(ds + 16, &r.windowsHandlePTR, 8);
    std::memcpy(ds + 24, &r.windowRadix, 4);
    std::memcpy(ds + 28, &r.overlapRatio, 4);
    std::memcpy(ds + 32, &r.dataLength, 8);
    std::memcpy(ds + 40, &r.overlapdataLength, 8);
    auto put_text = [&](uint64_t pslot, uint64_t at, const std::string& s) {
        put64(pslot, list_ptr(pslot, at, 2, s.size() + 1));
        std::memcpy(seg + 8 * at, s.data(), s.size());
    };
    put_text(ptr_at + 0, sm_at, r.sharedMemory);
    put_text(ptr_at + 2, id_at, r.mappedID);
    put_text(ptr_at + 3, op_at, r.options);
    if (n > 0) {
        put64(ptr_at + 1, list_ptr(ptr_at + 1, data_at, 4, n));
        std::memcpy(seg + 8 * data_at, data, 4 * (size_t)n);
    }
    return true;
}

}  // namespace ostft_wire
