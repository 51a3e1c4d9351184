// PNG (ISO/IEC 15948, with DEFLATE per RFC 1951 inside zlib per RFC 1950)
// and binary PNM decoding into RGBA8 for texture maps.
#include "host_image.h"

#include <cstdio>
#include <cstring>

namespace lfh {
namespace {

struct Bits {
    const uint8_t* p;
    size_t n, pos = 0;
    uint32_t buf = 0;
    int cnt = 0;
    bool bad = false;
    uint32_t get(int k) {
        while (cnt < k) {
            if (pos >= n) {
                bad = true;
                return 0;
            }
            buf |= (uint32_t)p[pos++] << cnt;
            cnt += 8;
        }
        const uint32_t v = buf & ((1u << k) - 1u);
        buf >>= k;
        cnt -= k;
        return v;
    }
    void align() {  // drop the rest of the current byte
        buf = 0;
        cnt = 0;
    }
};

// Canonical Huffman code from code lengths (RFC 1951, 3.2.2).
struct Huff {
    uint16_t count[16] = {0};
    std::vector<uint16_t> sym;
    bool build(const uint8_t* len, int n) {
        std::memset(count, 0, sizeof count);
        for (int i = 0; i < n; ++i) count[len[i]]++;
        count[0] = 0;
        int left = 1;
        for (int l = 1; l < 16; ++l) {
            left <<= 1;
            left -= count[l];
            if (left < 0) return false;  // over-subscribed
        }
        uint16_t offs[16];
        offs[1] = 0;
        for (int l = 1; l < 15; ++l) offs[l + 1] = (uint16_t)(offs[l] + count[l]);
        sym.assign(n, 0);
        for (int i = 0; i < n; ++i)
            if (len[i]) sym[offs[len[i]]++] = (uint16_t)i;
        return true;
    }
    int decode(Bits& b) const {
        int code = 0, first = 0, index = 0;
        for (int l = 1; l < 16; ++l) {
            code |= (int)b.get(1);
            if (b.bad) return -1;
            const int c = count[l];
            if (code - first < c) return sym[index + (code - first)];
            index += c;
            first += c;
            first <<= 1;
            code <<= 1;
        }
        return -1;
    }
};

// RFC 1951, 3.2.5: length and distance bases and extra bits.
const uint16_t kLenBase[29] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27,
                               31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
const uint8_t kLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
const uint16_t kDistBase[30] = {1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129,
                                193, 257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
const uint8_t kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6,
                                6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
const uint8_t kClOrder[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

bool inflate_block(Bits& b, const Huff& lit, const Huff& dist, std::vector<uint8_t>* out) {
    for (;;) {
        const int s = lit.decode(b);
        if (s < 0) return false;
        if (s < 256) {
            out->push_back((uint8_t)s);
        } else if (s == 256) {
            return true;
        } else {
            const int li = s - 257;
            if (li >= 29) return false;
            const int len = kLenBase[li] + (int)b.get(kLenExtra[li]);
            const int ds = dist.decode(b);
            if (ds < 0 || ds >= 30) return false;
            const size_t d = kDistBase[ds] + b.get(kDistExtra[ds]);
            if (b.bad || d > out->size()) return false;
            const size_t from = out->size() - d;
            for (int k = 0; k < len; ++k) out->push_back((*out)[from + k]);
        }
    }
}

bool inflate_raw(Bits& b, std::vector<uint8_t>* out, std::string* err) {
    for (;;) {
        const uint32_t last = b.get(1), type = b.get(2);
        if (b.bad) break;
        if (type == 0) {  // stored
            b.align();
            if (b.pos + 4 > b.n) break;
            const uint32_t len = b.p[b.pos] | (b.p[b.pos + 1] << 8), nlen = b.p[b.pos + 2] | (b.p[b.pos + 3] << 8);
            b.pos += 4;
            if ((len ^ 0xFFFFu) != nlen || b.pos + len > b.n) break;
            out->insert(out->end(), b.p + b.pos, b.p + b.pos + len);
            b.pos += len;
        } else if (type == 1) {  // fixed Huffman codes (3.2.6)
            uint8_t l[288];
            for (int i = 0; i < 144; ++i) l[i] = 8;
            for (int i = 144; i < 256; ++i) l[i] = 9;
            for (int i = 256; i < 280; ++i) l[i] = 7;
            for (int i = 280; i < 288; ++i) l[i] = 8;
            uint8_t dl[30];
            for (int i = 0; i < 30; ++i) dl[i] = 5;
            Huff lit, dist;
            lit.build(l, 288);
            dist.build(dl, 30);
            if (!inflate_block(b, lit, dist, out)) break;
        } else if (type == 2) {  // dynamic Huffman codes (3.2.7)
            const int hlit = (int)b.get(5) + 257, hdist = (int)b.get(5) + 1, hclen = (int)b.get(4) + 4;
            uint8_t cl[19] = {0};
            for (int i = 0; i < hclen; ++i) cl[kClOrder[i]] = (uint8_t)b.get(3);
            Huff clh;
            if (b.bad || !clh.build(cl, 19)) break;
            uint8_t lens[320] = {0};
            int i = 0;
            bool ok = true;
            while (i < hlit + hdist) {
                const int s = clh.decode(b);
                if (s < 0) { ok = false; break; }
                if (s < 16) {
                    lens[i++] = (uint8_t)s;
                } else {
                    int rep = 0;
                    uint8_t v = 0;
                    if (s == 16) {
                        if (i == 0) { ok = false; break; }
                        v = lens[i - 1];
                        rep = 3 + (int)b.get(2);
                    } else if (s == 17) {
                        rep = 3 + (int)b.get(3);
                    } else {
                        rep = 11 + (int)b.get(7);
                    }
                    if (i + rep > hlit + hdist) { ok = false; break; }
                    for (int k = 0; k < rep; ++k) lens[i++] = v;
                }
            }
            Huff lit, dist;
            if (!ok || b.bad || !lit.build(lens, hlit) || !dist.build(lens + hlit, hdist)) break;
            if (!inflate_block(b, lit, dist, out)) break;
        } else {
            break;
        }
        if (last) return true;
    }
    if (err) *err = "corrupt DEFLATE stream";
    return false;
}

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

bool decode_png(const uint8_t* d, size_t n, Image* out, std::string* err) {
    auto fail = [&](const char* m) {
        if (err) *err = std::string("PNG: ") + m;
        return false;
    };
    size_t pos = 8;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte, trns;
    bool seen_end = false;
    while (pos + 12 <= n && !seen_end) {
        const uint32_t len = be32(d + pos);
        const uint8_t* type = d + pos + 4;
        const uint8_t* body = d + pos + 8;
        if (pos + 12 + (size_t)len > n) return fail("truncated chunk");
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len < 13) return fail("bad IHDR");
            w = be32(body);
            h = be32(body + 4);
            depth = body[8];
            ctype = body[9];
            interlace = body[12];
        } else if (!std::memcmp(type, "PLTE", 4)) {
            plte.assign(body, body + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            trns.assign(body, body + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), body, body + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            seen_end = true;
        }
        pos += 12 + len;
    }
    if (w == 0 || h == 0 || w > (1u << 15) || h > (1u << 15)) return fail("bad size");
    if (interlace) return fail("interlaced images are not supported");
    int chans = 0;
    switch (ctype) {
    case 0: chans = 1; break;
    case 2: chans = 3; break;
    case 3: chans = 1; break;
    case 4: chans = 2; break;
    case 6: chans = 4; break;
    default: return fail("bad color type");
    }
    const bool depth_ok = (ctype == 0 || ctype == 3) ? (depth == 1 || depth == 2 || depth == 4 || depth == 8 ||
                                                         (ctype == 0 && depth == 16))
                                                      : (depth == 8 || depth == 16);
    if (!depth_ok) return fail("unsupported bit depth");
    if (ctype == 3 && plte.size() < 3) return fail("palette missing");
    std::vector<uint8_t> raw;
    if (!zlib_inflate(idat.data(), idat.size(), &raw, err)) return false;
    const size_t bits_pp = (size_t)chans * depth;
    const size_t stride = ((size_t)w * bits_pp + 7) / 8;
    const size_t bpp = (bits_pp + 7) / 8;   // filter byte distance
    if (raw.size() < (stride + 1) * h) return fail("image data too short");
    std::vector<uint8_t> img(stride * h), prev(stride, 0);
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t f = raw[y * (stride + 1)];
        const uint8_t* src = &raw[y * (stride + 1) + 1];
        uint8_t* cur = &img[y * stride];
        for (size_t x = 0; x < stride; ++x) {
            const int a = x >= bpp ? cur[x - bpp] : 0, b = prev[x], c = x >= bpp ? prev[x - bpp] : 0;
            int v = src[x];
            switch (f) {
            case 0: break;
            case 1: v += a; break;
            case 2: v += b; break;
            case 3: v += (a + b) >> 1; break;
            case 4: v += paeth(a, b, c); break;
            default: return fail("bad filter type");
            }
            cur[x] = (uint8_t)v;
        }
        std::memcpy(prev.data(), cur, stride);
    }
    out->width = (int)w;
    out->height = (int)h;
    out->rgba.assign((size_t)w * h * 4, 255);
    auto sample = [&](const uint8_t* row, uint32_t x, int c) -> int {   // channel c of pixel x, as 8 bits
        if (depth == 8) return row[x * chans + c];
        if (depth == 16) return row[(x * chans + c) * 2];   // high byte
        const uint32_t bit = x * depth;
        const int v = (row[bit >> 3] >> (8 - depth - (bit & 7))) & ((1 << depth) - 1);
        return v;
    };
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t* row = &img[y * stride];
        uint8_t* o = &out->rgba[(size_t)y * w * 4];
        for (uint32_t x = 0; x < w; ++x, o += 4) {
            if (ctype == 3) {
                const int i = sample(row, x, 0);
                if ((size_t)i * 3 + 2 >= plte.size()) return fail("palette index out of range");
                o[0] = plte[i * 3];
                o[1] = plte[i * 3 + 1];
                o[2] = plte[i * 3 + 2];
                o[3] = (size_t)i < trns.size() ? trns[i] : 255;
            } else if (ctype == 0 || ctype == 4) {
                int g = sample(row, x, 0);
                if (depth < 8) g = g * 255 / ((1 << depth) - 1);
                o[0] = o[1] = o[2] = (uint8_t)g;
                if (ctype == 4) o[3] = (uint8_t)sample(row, x, 1);
            } else {
                o[0] = (uint8_t)sample(row, x, 0);
                o[1] = (uint8_t)sample(row, x, 1);
                o[2] = (uint8_t)sample(row, x, 2);
                if (ctype == 6) o[3] = (uint8_t)sample(row, x, 3);
            }
        }
    }
    return true;
}

// Binary PNM: P5 (gray) and P6 (RGB), maxval <= 255.
bool decode_pnm(const uint8_t* d, size_t n, Image* out, std::string* err) {
    size_t pos = 2;
    auto skip = [&]() {
        for (;;) {
            while (pos < n && (d[pos] == ' ' || d[pos] == '\t' || d[pos] == '\r' || d[pos] == '\n')) ++pos;
            if (pos < n && d[pos] == '#') {
                while (pos < n && d[pos] != '\n') ++pos;
                continue;
            }
            return;
        }
    };
    auto num = [&]() -> long {
        skip();
        long v = 0;
        bool any = false;
        while (pos < n && d[pos] >= '0' && d[pos] <= '9') {
            v = v * 10 + (d[pos++] - '0');
            any = true;
            if (v > (1 << 20)) return -1;
        }
        return any ? v : -1;
    };
    const bool rgb = d[1] == '6';
    const long w = num(), h = num(), maxv = num();
    if (w <= 0 || h <= 0 || maxv <= 0 || maxv > 255 || w > (1 << 15) || h > (1 << 15)) {
        if (err) *err = "PNM: bad header";
        return false;
    }
    ++pos;   // the single whitespace after maxval
    const size_t c = rgb ? 3 : 1;
    if (pos + (size_t)w * h * c > n) {
        if (err) *err = "PNM: data too short";
        return false;
    }
    out->width = (int)w;
    out->height = (int)h;
    out->rgba.assign((size_t)w * h * 4, 255);
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        const uint8_t* s = d + pos + i * c;
        auto sc = [&](uint8_t v) { return (uint8_t)(maxv == 255 ? v : (v * 255 + maxv / 2) / maxv); };
        out->rgba[i * 4 + 0] = sc(s[0]);
        out->rgba[i * 4 + 1] = sc(rgb ? s[1] : s[0]);
        out->rgba[i * 4 + 2] = sc(rgb ? s[2] : s[0]);
    }
    return true;
}

}  // namespace

bool zlib_inflate(const uint8_t* data, size_t size, std::vector<uint8_t>* out, std::string* err) {
    if (size < 2 || (data[0] & 0x0F) != 8 || ((data[0] << 8) | data[1]) % 31 != 0 || (data[1] & 0x20)) {
        if (err) *err = "not a zlib stream";
        return false;
    }
    Bits b{data, size};
    b.pos = 2;
    out->clear();
    return inflate_raw(b, out, err);
}

bool decode_image(const uint8_t* data, size_t size, Image* out, std::string* err) {
    static const uint8_t kPng[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    if (size >= 8 && !std::memcmp(data, kPng, 8)) return decode_png(data, size, out, err);
    if (size >= 3 && data[0] == 'P' && (data[1] == '5' || data[1] == '6')) return decode_pnm(data, size, out, err);
    if (err) *err = "unsupported image format (PNG and binary PPM/PGM are read)";
    return false;
}

bool load_image(const std::string& path, Image* out, std::string* err) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        if (err) *err = "cannot open " + path;
        return false;
    }
    std::vector<uint8_t> buf;
    uint8_t tmp[1 << 16];
    size_t k;
    while ((k = std::fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + k);
    std::fclose(f);
    std::string e;
    if (!decode_image(buf.data(), buf.size(), out, &e)) {
        if (err) *err = path + ": " + e;
        return false;
    }
    return true;
}

}  // namespace lfh
