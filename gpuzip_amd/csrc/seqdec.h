// seqdec.h - a one-thread Deflate decoder with the result contract of the reference decoder's
// run() and probe() (gpuinflate/inflate.cuh:418-648): for every start, damaged input included,
// the same status, end and end-of-block positions, resume state, landing, block and symbol counts
// and last block. The rescue decides on those fields, so this decoder, not the lockstep one of
// deflate_core.h (whose table cache and early header checks change where a failure is reported),
// decodes its pieces: one per GPU lane for the pieces of a file, and on the host for the rescue's
// targeted re-decodes. Tables are built per block from the header, as the format defines them.
#pragma once
#include <stdint.h>

#include "deflate_core.h"

namespace seqd {

using gz::Start;

// The reference Result's fields (inflate.cuh:55-68)
struct Outcome {
    uint64_t end_bit;
    uint64_t eob_bit;         // after the first end-of-block code (UINT64_MAX: none)
    Start resume;             // ERR_OVERFLOW, STOPPED: the state to go on from
    uint32_t eob_out;         // output by that end-of-block code
    uint32_t out_len;
    int32_t status;           // gz::Status (never NEW_TABLE)
    int32_t landed;           // index in the targets (-1: none)
    int32_t end_final;        // BFINAL of the block it ended in (-1 unknown)
    uint32_t blocks;
    uint64_t last_block_bit;  // header of the last block it started (UINT64_MAX: none)
    uint32_t last_block_out;  // where that block's output begins
    uint32_t symbols;         // literals and matches
};

// LSB-first bits of in[0..size) from 32-bit word loads; words past the input's last read as zero,
// so reads run on past the end and pos() tells how far
class Reader {
  public:
    GZ_INL void at(const uint8_t* in, uint64_t size, uint64_t bit) {
        in_ = in;
        words_ = (size + 3) / 4;
        size_bits = size * 8;
        next_ = bit >> 5;
        acc_ = 0;
        have_ = 0;
        fill();
        const int skip = (int)(bit & 31);
        acc_ >>= skip;
        have_ -= skip;
    }
    GZ_INL void fill() {
        while (have_ <= 32) {
            const uint64_t w = next_ < words_ ? (uint64_t)load(next_) : 0;
            acc_ |= w << have_;
            next_++;
            have_ += 32;
        }
    }
    GZ_INL uint32_t look(int n) const { return (uint32_t)(acc_ & ((1ull << n) - 1)); }
    GZ_INL void skip(int n) {
        acc_ >>= n;
        have_ -= n;
    }
    GZ_INL uint32_t take(int n) {
        if (have_ < n) fill();
        const uint32_t v = look(n);
        skip(n);
        return v;
    }
    GZ_INL uint64_t pos() const { return next_ * 32 - (uint64_t)have_; }
    GZ_INL bool overrun() const { return pos() > size_bits; }
    GZ_INL void to_byte() {
        const int r = (int)(pos() & 7);
        if (r) take(8 - r);
    }
    GZ_INL uint64_t window() {  // the next 33+ bits
        if (have_ < 33) fill();
        return acc_;
    }
    uint64_t size_bits = 0;

  private:
    GZ_INL uint32_t load(uint64_t i) const {
        uint32_t w;
        memcpy(&w, in_ + 4 * i, 4);
        return w;
    }
    const uint8_t* in_ = nullptr;
    uint64_t words_ = 0, next_ = 0, acc_ = 0;
    int have_ = 0;
};

// A canonical code: codes of up to ROOT bits by one lookup ((length << 9) | symbol; 0: longer,
// or no code), longer ones by walking the count/symbol form (gz::canon_decode)
template <int ROOT, int NSYM>
struct Canon {
    uint16_t fast[1 << ROOT];
    uint16_t count[16];
    uint16_t sym[NSYM];
    int nonzero;

    // 0 complete, 1 incomplete, -1 over-subscribed (then unusable)
    GZ_INL int build(const uint8_t* lens, int n) {
        const int k = gz::canon_build<NSYM>(lens, n, count, sym, nonzero);
        if (k < 0) return k;
        for (int i = 0; i < (1 << ROOT); i++) fast[i] = 0;
        uint16_t next[16];
        int code = 0;
        for (int b = 1; b < 16; b++) {
            code = (code + (b > 1 ? count[b - 1] : 0)) << 1;
            next[b] = (uint16_t)code;
        }
        for (int s = 0; s < n; s++) {
            const int len = lens[s];
            if (!len) continue;
            const uint32_t c = next[len]++;
            if (len > ROOT) continue;
            const uint32_t rev = gz::reverse_bits(c, len);
            for (uint32_t x = rev; x < (1u << ROOT); x += 1u << len) fast[x] = (uint16_t)((len << 9) | s);
        }
        return k;
    }
    GZ_INL bool one_code() const { return nonzero == 1 && count[1] == 1; }
    // the next symbol, or -1 (nothing consumed) for bits that are no code
    GZ_INL int decode(Reader& br) const {
        const uint64_t w = br.window();
        const uint16_t e = fast[w & ((1u << ROOT) - 1)];
        if (e) {
            br.skip(e >> 9);
            return e & 511;
        }
        int len = 0;
        const int s = gz::canon_decode(count, sym, (uint32_t)w, &len);
        if (s >= 0) br.skip(len);
        return s;
    }
};

struct Inflater {
    Canon<10, 288> lit;
    Canon<8, 30> dist;
    Canon<7, 19> clc;
    uint8_t lens[320];
    bool fixed_ready = false;

    // a dynamic block's codes from its header (br just past BFINAL and BTYPE), by zlib's
    // acceptance rules: 0, or -1 with br where the header was found wrong
    GZ_INL int dynamic_codes(Reader& br) {
        static constexpr uint8_t order[19] = GZ_ORDER;
        fixed_ready = false;
        const int hlit = (int)br.take(5) + 257, hdist = (int)br.take(5) + 1, hclen = (int)br.take(4) + 4;
        if (hlit > 286 || hdist > 30) return -1;
        uint8_t cl[19];
        for (int i = 0; i < 19; i++) cl[i] = 0;
        for (int i = 0; i < hclen; i++) cl[order[i]] = (uint8_t)br.take(3);
        if (clc.build(cl, 19) != 0) return -1;  // the code-length code must be complete
        const int total = hlit + hdist;
        for (int i = 0; i < total;) {
            const int s = clc.decode(br);
            if (s < 0) return -1;
            if (s < 16) {
                lens[i++] = (uint8_t)s;
                continue;
            }
            uint8_t val = 0;
            int rep;
            if (s == 16) {
                if (i == 0) return -1;
                val = lens[i - 1];
                rep = 3 + (int)br.take(2);
            } else {
                rep = s == 17 ? 3 + (int)br.take(3) : 11 + (int)br.take(7);
            }
            if (i + rep > total) return -1;
            while (rep--) lens[i++] = val;
        }
        if (lens[256] == 0) return -1;
        int k = lit.build(lens, hlit);
        if (k < 0 || (k > 0 && !lit.one_code())) return -1;
        k = dist.build(lens + hlit, hdist);
        if (k < 0 || (k > 0 && !(dist.nonzero == 0 || dist.one_code()))) return -1;
        return 0;
    }
    GZ_INL void fixed_codes() {
        if (fixed_ready) return;
        fixed_ready = true;
        for (int i = 0; i < 288; i++) lens[i] = i < 144 ? 8 : i < 256 ? 9 : i < 280 ? 7 : 8;
        lit.build(lens, 288);
        for (int i = 0; i < 30; i++) lens[i] = 5;
        dist.build(lens, 30);
    }

    // One block's symbols up to its end-of-block code (0); a landing between symbols when LAND (1,
    // o.landed); a full buffer, the stop bit, bad data or the input's end (-1, o.status set)
    template <bool LAND>
    GZ_INL int block_symbols(Reader& br, uint16_t* out, uint32_t& op, uint32_t cap, bool stream_start, const Start* tg,
                             int nt, int& t, bool fixed, uint64_t header, int final, uint64_t limit, Outcome& o) {
        static constexpr uint16_t lbase[29] = GZ_LBASE;
        static constexpr uint8_t lext[29] = GZ_LEXT;
        static constexpr uint16_t dbase[30] = GZ_DBASE;
        static constexpr uint8_t dext[30] = GZ_DEXT;
        const uint8_t mode = fixed ? gz::IN_FIXED : gz::IN_DYNAMIC;
        for (;;) {
            if (LAND) {
                const uint64_t p = br.pos();
                while (t < nt && tg[t].bit < p) t++;
                if (t < nt && tg[t].bit == p && tg[t].mode == mode && (fixed || tg[t].header == header)) {
                    o.landed = t;
                    return 1;
                }
            }
            if (op + gz::MAX_SYMBOL_OUT > cap || br.pos() >= limit) {
                const uint64_t p = br.pos();
                if (p > br.size_bits) {
                    o.status = gz::ERR_INPUT;
                    return -1;
                }
                o.status = p >= limit ? gz::STOPPED : gz::ERR_OVERFLOW;
                o.resume = make_start(p, header, mode, final);
                return -1;
            }
            int s = lit.decode(br);
            if (s < 256) {
                if (s < 0) {
                    o.status = gz::ERR_DATA;
                    return -1;
                }
                out[op++] = (uint16_t)s;
                o.symbols++;
                continue;
            }
            if (s == 256) return 0;
            o.symbols++;
            s -= 257;
            if (s >= 29) {
                o.status = gz::ERR_DATA;
                return -1;
            }
            const uint32_t len = lbase[s] + br.take(lext[s]);
            const int ds = dist.decode(br);
            if (ds < 0 || ds >= 30) {
                o.status = gz::ERR_DATA;
                return -1;
            }
            const uint32_t d = dbase[ds] + br.take(dext[ds]);
            if (d > op && (stream_start || d - op > gz::WINDOW)) {
                o.status = gz::ERR_DATA;
                return -1;
            }
            const uint16_t v = (uint16_t)(gz::REF + d);
            for (uint32_t k = 0; k < len; k++) out[op + k] = v;
            op += len;
        }
    }

    // a stored block's copy listed for later (true), or to be made now
    GZ_INL static bool defer_stored(gz::StoredList* stored, uint64_t byte, uint64_t dst, uint32_t len) {
#if defined(__HIP_DEVICE_COMPILE__)
        if (!stored || !len) return false;
        const unsigned i = atomicAdd(stored->count, 1u);
        if (i >= stored->cap) return false;
        stored->list[i] = {byte, dst, len, 0};
        return true;
#else
        (void)stored, (void)byte, (void)dst, (void)len;
        return false;
#endif
    }

    GZ_INL static Start make_start(uint64_t bit, uint64_t header, uint8_t mode, int final) {
        Start s;
        s.bit = bit;
        s.header = header;
        s.mode = mode;
        s.final = (int8_t)final;
        for (int i = 0; i < 6; i++) s.pad[i] = 0;
        return s;
    }

    // Decodes from s until it lands on one of tg[0..nt) (sorted by bit), a full buffer, the stop
    // bit, the end of the last block, or an error (inflate.cuh:467-581's contract). On the GPU, with
    // `stored` given, a stored block's bytes are not copied but listed (input byte, out_index + its
    // first symbol, length) for a copy kernel, while the list has room.
    GZ_INL Outcome run(const uint8_t* in, uint64_t size, Start s, const Start* tg, int nt, uint16_t* out, uint32_t cap,
                       bool stream_start, uint64_t stop = UINT64_MAX, gz::StoredList* stored = nullptr,
                       uint64_t out_index = 0) {
        Outcome o;
        o.end_bit = 0;
        o.eob_bit = UINT64_MAX;
        o.resume = make_start(0, 0, gz::AT_BLOCK, -1);
        o.eob_out = 0;
        o.out_len = 0;
        o.status = gz::OK;
        o.landed = -1;
        o.end_final = s.mode == gz::AT_BLOCK ? -1 : s.final;
        o.blocks = 0;
        o.last_block_bit = UINT64_MAX;
        o.last_block_out = 0;
        o.symbols = 0;
        Reader br;
        uint32_t op = 0;
        int t = 0;
        bool in_block = false, fixed = false;
        uint64_t header = s.header;
        if (s.mode == gz::IN_FIXED) {
            fixed_codes();
            in_block = fixed = true;
        } else if (s.mode == gz::IN_DYNAMIC) {
            br.at(in, size, s.header + 3);
            if (dynamic_codes(br) != 0) {
                o.status = gz::ERR_DATA;
                o.end_bit = s.bit;
                return o;
            }
            in_block = true;
        }
        br.at(in, size, s.bit);
        const uint64_t limit = stop < br.size_bits + 1 ? stop : br.size_bits + 1;
        for (;;) {
            if (!in_block) {
                const uint64_t p = br.pos();
                while (t < nt && tg[t].bit < p) t++;
                if (t < nt && tg[t].bit == p && tg[t].mode == gz::AT_BLOCK) {
                    o.landed = t;
                    break;
                }
                if (p >= stop || op + gz::MAX_SYMBOL_OUT > cap || o.blocks >= gz::MAX_BLOCKS) {
                    o.status = p >= stop ? gz::STOPPED : gz::ERR_OVERFLOW;
                    o.resume = make_start(p, p, gz::AT_BLOCK, -1);
                    break;
                }
                if (p >= br.size_bits) {
                    o.status = gz::ERR_INPUT;
                    break;
                }
                header = p;
                o.last_block_bit = p;
                o.last_block_out = op;
                const int final = (int)br.take(1);
                const int type = (int)br.take(2);
                o.blocks++;
                o.end_final = final;
                if (type == 0) {
                    br.to_byte();
                    const uint32_t len = br.take(16), nlen = br.take(16);
                    if (len != (~nlen & 0xffffu)) {
                        o.status = gz::ERR_DATA;
                        break;
                    }
                    if (op + len > cap) {
                        o.status = gz::ERR_OVERFLOW;
                        o.resume = make_start(p, p, gz::AT_BLOCK, -1);
                        break;
                    }
                    const uint64_t byte = br.pos() >> 3;
                    if (byte + len > size) {
                        o.status = gz::ERR_INPUT;
                        break;
                    }
                    if (!defer_stored(stored, byte, out_index + op, len))
                        for (uint32_t k = 0; k < len; k++) out[op + k] = in[byte + k];
                    op += len;
                    br.at(in, size, (byte + len) * 8);
                    if (final) break;
                    continue;
                }
                if (type == 3) {
                    o.status = gz::ERR_DATA;
                    break;
                }
                if (type == 1) {
                    fixed_codes();
                    fixed = true;
                } else {
                    if (dynamic_codes(br) != 0) {
                        o.status = gz::ERR_DATA;
                        break;
                    }
                    fixed = false;
                }
                in_block = true;
            }
            // landing between symbols: in every fixed block; in a dynamic one when the next
            // target names its header
            const bool land = fixed || (t < nt && tg[t].mode == gz::IN_DYNAMIC && tg[t].header == header);
            const int how = land ? block_symbols<true>(br, out, op, cap, stream_start, tg, nt, t, fixed, header,
                                                       o.end_final, limit, o)
                                 : block_symbols<false>(br, out, op, cap, stream_start, tg, nt, t, fixed, header,
                                                        o.end_final, limit, o);
            if (how != 0) break;
            in_block = false;
            if (br.overrun()) {
                o.status = gz::ERR_INPUT;
                break;
            }
            if (o.eob_bit == UINT64_MAX) {
                o.eob_bit = br.pos();
                o.eob_out = op;
            }
            if (o.end_final == 1) break;
        }
        o.end_bit = br.pos();
        o.out_len = op;
        return o;
    }

    // The first symbol boundary at or after at_bit, decoding (without output) from from_bit with
    // the codes of the block at `header` (mode AT_BLOCK), the fixed code, or the dynamic header
    // at `header`, one bit later after each failure, 64 attempts at most (inflate.cuh:591-648's
    // contract). bit UINT64_MAX: none.
    GZ_INL Start probe(const uint8_t* in, uint64_t size, uint8_t mode, uint64_t header, uint64_t from_bit,
                       uint64_t at_bit) {
        static constexpr uint8_t lext[29] = GZ_LEXT;
        static constexpr uint8_t dext[30] = GZ_DEXT;
        const Start none = make_start(UINT64_MAX, UINT64_MAX, gz::AT_BLOCK, -1);
        bool fixed0 = mode == gz::IN_FIXED;
        if (mode == gz::AT_BLOCK) {
            Reader hb;
            hb.at(in, size, header);
            hb.take(1);
            const int type0 = (int)hb.take(2);
            if (type0 == 0 || type0 == 3) return none;
            fixed0 = type0 == 1;
        }
        uint64_t from = from_bit;
        for (int attempt = 0; attempt < 64 && from < at_bit; attempt++) {
            bool fixed = fixed0;
            if (fixed) {
                fixed_codes();
            } else {
                Reader h2;
                h2.at(in, size, header + 3);
                if (dynamic_codes(h2) != 0) return none;
            }
            uint64_t cur = header, p;
            Reader br;
            br.at(in, size, from);
            for (;;) {
                p = br.pos();
                if (p >= at_bit) return make_start(p, cur, fixed ? gz::IN_FIXED : gz::IN_DYNAMIC, -1);
                int s = lit.decode(br);
                if (s < 0 || s > 285) break;
                if (s < 256) continue;
                if (s == 256) {
                    const uint64_t h = br.pos();
                    if (h >= at_bit) return make_start(h, h, gz::AT_BLOCK, -1);
                    br.take(1);
                    const int ty = (int)br.take(2);
                    if (ty == 1) {
                        fixed_codes();
                        fixed = true;
                    } else if (ty == 2 && dynamic_codes(br) == 0) {
                        fixed = false;
                    } else {
                        break;  // stored, or bad: try again a bit later
                    }
                    cur = h;
                    continue;
                }
                s -= 257;
                br.take(lext[s]);
                const int ds = dist.decode(br);
                if (ds < 0 || ds >= 30) break;
                br.take(dext[ds]);
            }
            from = p + 1;
        }
        return none;
    }
};

}  // namespace seqd
