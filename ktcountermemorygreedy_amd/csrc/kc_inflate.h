// Raw DEFLATE decoding (RFC 1951) from any bit position, for inflating one
// gzip member on several threads.  A decoder that starts inside the stream
// does not know the 32 KB of output before its start; a back-reference that
// reaches there is stored as a marker, 256 + (32768 + offset) for offset
// -32768..-1 relative to the decoder's first output byte, and literals as
// their byte value, so the output is an array of 16-bit symbols.  Once the
// output before the start is known, every marker is replaced by the byte it
// names (resolve).  The tables are RFC 1951's (length and distance bases,
// code-length order, fixed codes).
#pragma once
#include <stdint.h>

#include <algorithm>
#include <cstring>
#include <vector>

namespace kcz {

// Reads bits LSB-first from a byte array; the position may run past the end
// (reads there see zero bytes, and callers check the position).
struct BitIn {
    const uint8_t *base = nullptr;
    uint64_t nbytes = 0;
    uint64_t pos = 0;  // next bit
    uint64_t peek(int n) const
    {
        const uint64_t b = pos >> 3;
        uint64_t v = 0;
        if (b + 8 <= nbytes) {
            std::memcpy(&v, base + b, 8);
        } else {
            for (uint64_t i = 0; b + i < nbytes && i < 8; ++i) v |= (uint64_t)base[b + i] << (8 * i);
        }
        v >>= (pos & 7);
        return n >= 57 ? v : v & ((1ull << n) - 1);
    }
    uint64_t bits(int n)
    {
        const uint64_t v = n ? peek(n) : 0;
        pos += (uint64_t)n;
        return v;
    }
    bool past_end() const { return pos > nbytes * 8; }
};

// Canonical Huffman code: a table of the first FB bits for short codes,
// codes longer than FB (and invalid bit patterns) by counting per length.
struct Huff {
    static constexpr int FB = 10;
    uint16_t fast[1 << FB];  // (length << 12) | symbol, 0 = not in the table
    uint16_t count[16];
    uint16_t sym[320];
    // false for an over-subscribed set, or an incomplete one unless allowed
    bool build(const uint8_t *len, int n, bool allow_incomplete)
    {
        std::memset(count, 0, sizeof(count));
        for (int s = 0; s < n; ++s) count[len[s]]++;
        count[0] = 0;
        int left = 1;
        for (int l = 1; l < 16; ++l) {
            left <<= 1;
            left -= count[l];
            if (left < 0) return false;
        }
        if (left > 0 && !allow_incomplete) return false;
        uint16_t offs[16];
        offs[1] = 0;
        for (int l = 1; l < 15; ++l) offs[l + 1] = (uint16_t)(offs[l] + count[l]);
        for (int s = 0; s < n; ++s)
            if (len[s]) sym[offs[len[s]]++] = (uint16_t)s;
        std::memset(fast, 0, sizeof(fast));
        uint32_t code = 0;
        int k = 0;
        for (int l = 1; l < 16; ++l) {
            for (int c = 0; c < count[l]; ++c, ++code) {
                const uint16_t s = sym[k++];
                if (l > FB) continue;
                uint32_t r = 0;  // the code's bits in stream order
                for (int b = 0; b < l; ++b) r |= ((code >> b) & 1u) << (l - 1 - b);
                for (uint32_t x = r; x < (1u << FB); x += 1u << l) fast[x] = (uint16_t)((l << 12) | s);
            }
            code <<= 1;
        }
        return true;
    }
    // the next symbol, or -1 for a bit pattern that is no code
    int decode(BitIn &in) const
    {
        const uint64_t v = in.peek(15);
        const uint16_t e = fast[v & ((1u << FB) - 1)];
        if (e) {
            in.pos += e >> 12;
            return e & 0xFFF;
        }
        int code = 0, first = 0, index = 0;
        for (int l = 1; l < 16; ++l) {
            code |= (int)((v >> (l - 1)) & 1);
            const int cnt = count[l];
            if (code - cnt < first) {
                in.pos += (uint64_t)l;
                return sym[index + (code - first)];
            }
            index += cnt;
            first += cnt;
            first <<= 1;
            code <<= 1;
        }
        return -1;
    }
};

constexpr uint16_t kLenBase[29] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27,
                                   31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
constexpr uint8_t kLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
constexpr uint16_t kDistBase[30] = {1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129, 193,
                                    257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
constexpr uint8_t kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
constexpr uint8_t kClOrder[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
constexpr uint32_t kWindow = 32768;
constexpr uint16_t kMarker = 256;  // marker of offset o (-32768..-1): kMarker + kWindow + o

// Output symbols (see the file comment), grown in large steps.
struct Out16 {
    std::vector<uint16_t> v;
    uint64_t n = 0;
    void reserve_more(uint64_t k)
    {
        if (n + k > v.size()) v.resize(std::max<uint64_t>(n + k, v.size() * 2 + (1u << 20)));
    }
};

enum BlockResult { kBlockOk = 0, kBlockFinal = 1, kBlockError = -1 };

struct Decoder {
    BitIn in;
    Huff lit, dist;
    bool unknown_window = false;  // back-references before the first output become markers
    bool text_only = false;       // literals must be text bytes (9..126): a plausibility check for a guessed start
    uint64_t max_out = ~0ull;     // more output symbols than this fail the block (a guessed start gone wrong)

    bool read_dynamic_tables()
    {
        const int hlit = (int)in.bits(5) + 257, hdist = (int)in.bits(5) + 1, hclen = (int)in.bits(4) + 4;
        if (hlit > 286 || hdist > 30) return false;
        uint8_t cl[19] = {};
        for (int i = 0; i < hclen; ++i) cl[kClOrder[i]] = (uint8_t)in.bits(3);
        Huff clh;
        if (!clh.build(cl, 19, false)) return false;
        uint8_t lens[286 + 30];
        int i = 0;
        while (i < hlit + hdist) {
            const int s = clh.decode(in);
            if (s < 0) return false;
            if (s < 16) {
                lens[i++] = (uint8_t)s;
                continue;
            }
            int rep;
            uint8_t v = 0;
            if (s == 16) {
                if (i == 0) return false;
                v = lens[i - 1];
                rep = 3 + (int)in.bits(2);
            } else if (s == 17) {
                rep = 3 + (int)in.bits(3);
            } else {
                rep = 11 + (int)in.bits(7);
            }
            if (i + rep > hlit + hdist) return false;
            while (rep--) lens[i++] = v;
        }
        if (in.past_end() || lens[256] == 0) return false;  // no end-of-block code
        auto max_len = [](const uint8_t *l, int n) {
            int m = 0;
            for (int k = 0; k < n; ++k) m = l[k] > m ? l[k] : m;
            return m;
        };
        // incomplete codes only as zlib takes them: a single code of length 1
        return lit.build(lens, hlit, max_len(lens, hlit) == 1) &&
               dist.build(lens + hlit, hdist, max_len(lens + hlit, hdist) <= 1);
    }

    void fixed_tables()
    {
        uint8_t l[288];
        for (int s = 0; s < 144; ++s) l[s] = 8;
        for (int s = 144; s < 256; ++s) l[s] = 9;
        for (int s = 256; s < 280; ++s) l[s] = 7;
        for (int s = 280; s < 288; ++s) l[s] = 8;
        lit.build(l, 288, false);
        uint8_t d[30];
        for (int s = 0; s < 30; ++s) d[s] = 5;
        dist.build(d, 30, true);
    }

    // The block at in.pos, appended to out.
    BlockResult block(Out16 &out)
    {
        const bool final_block = in.bits(1) != 0;
        const int type = (int)in.bits(2);
        if (type == 0) {
            in.pos = (in.pos + 7) & ~7ull;
            const uint64_t b = in.pos >> 3;
            if (b + 4 > in.nbytes) return kBlockError;
            const uint16_t len = (uint16_t)(in.base[b] | (in.base[b + 1] << 8));
            const uint16_t nlen = (uint16_t)(in.base[b + 2] | (in.base[b + 3] << 8));
            if ((uint16_t)~len != nlen || b + 4 + len > in.nbytes || out.n + len > max_out) return kBlockError;
            out.reserve_more(len);
            for (uint32_t k = 0; k < len; ++k) {
                const uint8_t c = in.base[b + 4 + k];
                if (text_only && (c < 9 || c > 126)) return kBlockError;
                out.v[out.n++] = c;
            }
            in.pos += (4ull + len) * 8;
            return final_block ? kBlockFinal : kBlockOk;
        }
        if (type == 1) fixed_tables();
        else if (type == 2) {
            if (!read_dynamic_tables()) return kBlockError;
        } else {
            return kBlockError;
        }
        if (!codes(out)) return kBlockError;
        return final_block ? kBlockFinal : kBlockOk;
    }

    // The symbols of a Huffman-coded block up to its end-of-block code, with
    // a 64-bit bit buffer refilled once per symbol: a literal/length code,
    // its extra bits, a distance code and its extra bits take at most 48
    // bits, and a refill leaves at least 56.
    bool codes(Out16 &out)
    {
        const uint8_t *base = in.base;
        const uint64_t nbytes = in.nbytes;
        uint64_t bp = in.pos >> 3;  // next byte to load (past the end: zeros)
        uint64_t buf = 0;
        int cnt = 0;
        auto refill = [&]() {
            if (bp + 8 <= nbytes) {
                uint64_t v;
                std::memcpy(&v, base + bp, 8);
                buf |= v << cnt;
                bp += (uint64_t)((63 - cnt) >> 3);
                cnt |= 56;
            } else {
                while (cnt <= 56) {
                    if (bp < nbytes) buf |= (uint64_t)base[bp] << cnt;
                    ++bp;
                    cnt += 8;
                }
            }
        };
        refill();
        buf >>= (in.pos & 7);
        cnt -= (int)(in.pos & 7);
        auto take = [&](int n) {
            const uint64_t v = buf & ((1ull << n) - 1);
            buf >>= n;
            cnt -= n;
            return (uint32_t)v;
        };
        // the next symbol of h from the buffer, or -1 for no code
        auto sym_of = [&](const Huff &h) -> int {
            const uint16_t e = h.fast[buf & ((1u << Huff::FB) - 1)];
            if (e) {
                const int l = e >> 12;
                buf >>= l;
                cnt -= l;
                return e & 0xFFF;
            }
            int code = 0, first = 0, index = 0;
            for (int l = 1; l < 16; ++l) {
                code |= (int)((buf >> (l - 1)) & 1);
                const int c = h.count[l];
                if (code - c < first) {
                    buf >>= l;
                    cnt -= l;
                    return h.sym[index + (code - first)];
                }
                index += c;
                first += c;
                first <<= 1;
                code <<= 1;
            }
            return -1;
        };
        uint64_t n = out.n;
        uint16_t *o = out.v.data();
        uint64_t cap = out.v.size();
        const uint64_t end_bits = nbytes * 8;
        bool ok = false;
        for (;;) {
            refill();
            if (bp * 8 - (uint64_t)cnt > end_bits) break;  // read past the input
            if (n > max_out) break;
            if (n + 258 > cap) {
                out.n = n;
                out.reserve_more(258);
                o = out.v.data();
                cap = out.v.size();
            }
            const int s = sym_of(lit);
            if (s < 0) break;
            if (s < 256) {
                if (text_only && (s < 9 || s > 126)) break;
                o[n++] = (uint16_t)s;
                continue;
            }
            if (s == 256) {
                ok = true;
                break;
            }
            const int li = s - 257;
            if (li >= 29) break;
            const uint32_t len = kLenBase[li] + take(kLenExtra[li]);
            const int ds = sym_of(dist);
            if (ds < 0 || ds >= 30) break;
            const uint32_t d = kDistBase[ds] + take(kDistExtra[ds]);
            const int64_t src = (int64_t)n - (int64_t)d;
            if (src >= 0) {
                if (d >= len) std::memcpy(o + n, o + src, (size_t)len * 2);
                else
                    for (uint32_t k = 0; k < len; ++k) o[n + k] = o[src + k];  // overlapping copies repeat
            } else {
                if (!unknown_window || src < -(int64_t)kWindow) break;
                for (uint32_t k = 0; k < len; ++k) {
                    const int64_t q = src + (int64_t)k;
                    o[n + k] = q < 0 ? (uint16_t)(kMarker + kWindow + q) : o[q];
                }
            }
            n += len;
        }
        out.n = n;
        in.pos = bp * 8 - (uint64_t)cnt;
        return ok && !in.past_end();
    }
};

// The first bit position in [from, to) where a dynamic block starts that
// decodes to its end as text with an unknown window, or `to` if none.
inline uint64_t find_block_start(const uint8_t *base, uint64_t nbytes, uint64_t from, uint64_t to)
{
    Decoder dec;
    dec.in.base = base;
    dec.in.nbytes = nbytes;
    dec.unknown_window = true;
    dec.text_only = true;
    Out16 scratch;
    for (uint64_t p = from; p < to; ++p) {
        // BFINAL = 0 and BTYPE = 2 (dynamic): bits 0b100 from p on
        dec.in.pos = p;
        if ((dec.in.peek(3) & 7) != 4) continue;
        dec.in.pos = p;
        scratch.n = 0;
        if (dec.block(scratch) == kBlockOk) return p;
    }
    return to;
}

// Replaces markers by the bytes of window (the wn <= 32768 bytes before the
// symbols' start; a marker before them is an error): symbols -> bytes, in
// place over out[0, n) (uint8 view of the same storage).  Returns false on a
// marker the window cannot answer.
inline bool resolve(const uint16_t *sym, uint64_t n, const uint8_t *window, uint32_t wn, uint8_t *out)
{
    for (uint64_t i = 0; i < n; ++i) {
        const uint16_t v = sym[i];
        if (v < kMarker) {
            out[i] = (uint8_t)v;
            continue;
        }
        const int64_t back = (int64_t)kWindow - (int64_t)(v - kMarker);  // 1..32768 bytes before the start
        if (back > (int64_t)wn) return false;
        out[i] = window[wn - back];
    }
    return true;
}

}  // namespace kcz
