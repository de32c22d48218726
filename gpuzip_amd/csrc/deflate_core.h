// deflate_core.h - the Deflate (RFC 1951) decoding core of the MI355X decoder. One source for
// the gfx950 kernels and for a host build that unit-tests the same logic without a GPU.
//
// Design:
//  * Huffman tables live in a table cache in device memory, one entry per block header
//    (`Table`), built once by a whole workgroup (k_tables) and shared by every decoder that
//    decodes inside that block. First-level lookup entries are 32-bit and carry the code
//    length, the number of extra bits and the base value, so a length or distance costs one
//    lookup. Codes longer than the lookup bits take a canonical walk over count/symbol arrays.
//  * A decoder writes 16-bit symbols: a byte (<= 255), or REF + d for "the byte d back". Matches
//    are never copied while decoding; references are resolved afterwards, all at once.
//  * A decoder lands exactly on a sorted list of targets (block starts and probed symbol
//    boundaries), stops at a stop bit, or fills its buffer, and then reports the exact state to
//    continue from. Crossing into a dynamic block whose table is not in the cache, it stops
//    at that block's header (NEW_TABLE) and the next round continues from there.
// Semantics follow the reference decoder, gpuinflate/inflate.cuh:418-581 (run) and
// inflate.cuh:285-332 (block-start predicates), restated for this design.
#pragma once
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GZ_HD __host__ __device__
#define GZ_INL __host__ __device__ __forceinline__
#define GZ_NOINL __host__ __device__ __noinline__
#else
#define GZ_HD
#define GZ_INL inline
#define GZ_NOINL inline
#endif

namespace gz {

constexpr uint32_t WINDOW = 32768;
constexpr uint16_t REF = 255;            // 16-bit symbols above this: the byte (v - REF) back
constexpr uint32_t MAX_SYMBOL_OUT = 258; // the most one symbol writes
constexpr uint32_t MAX_BLOCKS = 1u << 16;
constexpr int LIT_BITS = 10;             // first-level lookup bits, literal/length code
constexpr int DIST_BITS = 8;             // first-level lookup bits, distance code

enum Status : int32_t { OK = 0, ERR_DATA = 1, ERR_OVERFLOW = 2, ERR_INPUT = 3, STOPPED = 4, NEW_TABLE = 5 };
enum Mode : uint8_t { AT_BLOCK = 0, IN_FIXED = 1, IN_DYNAMIC = 2 };

// A decoder's starting state, or a target it may land on (inflate.cuh:43-53)
struct Start {
    uint64_t bit;
    uint64_t header;  // IN_DYNAMIC: the block's header; AT_BLOCK: the bit itself
    uint8_t mode;
    int8_t final;     // BFINAL of the block it starts in; -1 unknown
    uint8_t pad[6];
};
static_assert(sizeof(Start) == 24, "Start layout");

struct Result {
    uint64_t end_bit;    // where decoding stopped
    uint64_t eob_bit;    // after the first end-of-block code; UINT64_MAX none
    Start resume;        // ERR_OVERFLOW, STOPPED, NEW_TABLE: where to continue
    uint32_t eob_out;    // symbols written by the first end-of-block code
    uint32_t out_len;    // symbols written
    int32_t status;
    int32_t landed;      // index in the job's target list; -1: none
    int32_t end_final;   // BFINAL of the block it stopped in; -1 unknown
    uint32_t blocks;
    uint32_t reach;      // max (d - symbols before it) over matches reaching before the piece
    uint32_t reach_eob;  // the same over matches before the first end-of-block code
};
static_assert(sizeof(Result) == 72, "Result layout");

// lookup entries (32-bit)
//  literal/length: [3:0] code length (0: not in the first level), [7:4] extra bits,
//                  [9:8] kind (0 literal, 1 length, 2 end of block, 3 invalid symbol),
//                  [31:16] literal byte or base length
//  distance:       [3:0] code length, [7:4] extra bits, [8] invalid symbol, [31:16] base
enum : uint32_t { K_LIT = 0, K_LEN = 1, K_EOB = 2, K_BAD = 3 };

// A block's code, as the decoders use it (the table cache holds one per block header)
struct Table {
    uint32_t lit[1 << LIT_BITS];
    uint32_t dist[1 << DIST_BITS];
    uint16_t lcount[16];
    uint16_t dcount[16];
    uint16_t lsym[288];
    uint16_t dsym[32];
    int32_t status;    // TAB_DYNAMIC: a valid dynamic code; else what the header holds
    uint32_t hdr_bits; // dynamic: bits from the block header to the first symbol
    uint32_t pad[2];
};
enum : int32_t { TAB_DYNAMIC = 0, TAB_FIXED = 1, TAB_STORED = 2, TAB_BAD = -1 };
static_assert(sizeof(Table) % 16 == 0, "Table alignment");

// RFC 1951 3.2.5 / 3.2.7 (the format fixes these)
#define GZ_LBASE {3,4,5,6,7,8,9,10,11,13,15,17,19,23,27,31,35,43,51,59,67,83,99,115,131,163,195,227,258}
#define GZ_LEXT {0,0,0,0,0,0,0,0,1,1,1,1,2,2,2,2,3,3,3,3,4,4,4,4,5,5,5,5,0}
#define GZ_DBASE {1,2,3,4,5,7,9,13,17,25,33,49,65,97,129,193,257,385,513,769,1025,1537,2049,3073,4097,6145,8193,12289,16385,24577}
#define GZ_DEXT {0,0,0,0,1,1,2,2,3,3,4,4,5,5,6,6,7,7,8,8,9,9,10,10,11,11,12,12,13,13}
#define GZ_ORDER {16,17,18,0,8,7,9,6,10,5,11,4,12,3,13,2,14,1,15}

GZ_INL uint32_t lit_entry(int sym, int len) {
    static constexpr uint16_t lb[29] = GZ_LBASE;
    static constexpr uint8_t le[29] = GZ_LEXT;
    if (sym < 256) return (uint32_t)len | (K_LIT << 8) | ((uint32_t)sym << 16);
    if (sym == 256) return (uint32_t)len | (K_EOB << 8);
    if (sym > 285) return (uint32_t)len | (K_BAD << 8);
    int i = sym - 257;
    return (uint32_t)len | ((uint32_t)le[i] << 4) | (K_LEN << 8) | ((uint32_t)lb[i] << 16);
}

GZ_INL uint32_t dist_entry(int sym, int len) {
    static constexpr uint16_t db[30] = GZ_DBASE;
    static constexpr uint8_t de[30] = GZ_DEXT;
    if (sym >= 30) return (uint32_t)len | 0x100u;
    return (uint32_t)len | ((uint32_t)de[sym] << 4) | ((uint32_t)db[sym] << 16);
}

// ------------------------------------------------------------------ bit reader
// LSB-first, 32-bit aligned loads into a 64-bit buffer. Words at or past `nwords` read as
// zero, and the last word is masked to the input's bytes, so nothing past the input counts.
struct Bits {
    const uint32_t* w;
    uint64_t nfull;     // words wholly inside the input
    uint32_t tail_mask; // the partial last word's valid bytes (0: none)
    uint64_t next;      // next word to load
    uint64_t buf;
    uint32_t cnt;       // valid bits in buf

    GZ_INL void setup(const uint32_t* words, uint64_t size_bytes) {
        w = words;
        nfull = size_bytes >> 2;
        uint32_t r = (uint32_t)(size_bytes & 3);
        tail_mask = r ? (uint32_t)((1ull << (8 * r)) - 1) : 0;
    }
    GZ_INL uint32_t word(uint64_t i) const {
        if (i < nfull) return w[i];
        if (i == nfull && tail_mask) return w[i] & tail_mask;
        return 0;
    }
    GZ_INL void refill() {
        while (cnt <= 32) {
            buf |= (uint64_t)word(next) << cnt;
            next++;
            cnt += 32;
        }
    }
    GZ_INL void seek(uint64_t bit) {
        next = bit >> 5;
        buf = 0;
        cnt = 0;
        refill();
        uint32_t s = (uint32_t)(bit & 31);
        buf >>= s;
        cnt -= s;
    }
    GZ_INL uint64_t pos() const { return next * 32 - cnt; }
    GZ_INL uint32_t peek(int n) const { return (uint32_t)(buf & ((1ull << n) - 1)); }
    GZ_INL void drop(int n) {
        buf >>= n;
        cnt -= (uint32_t)n;
    }
    GZ_INL uint32_t get(int n) {
        if (cnt < (uint32_t)n) refill();
        uint32_t v = peek(n);
        drop(n);
        return v;
    }
};

// The decoders' reader: the same LSB-first bit buffer as Bits, fed from 16-byte chunks with the
// next chunk already loading while the current one is consumed, so a decoder seldom waits on
// device memory. Chunks wholly inside the input are read with one 16-byte load (when the input
// is 16-byte aligned); the rest word by word with Bits' masking.
struct alignas(16) Q4 {
    uint32_t x, y, z, w;
};

struct FBits {
    const uint32_t* w;
    uint64_t nfull;
    uint32_t tail_mask;
    bool aligned;
    uint64_t next;       // next word to consume
    uint64_t buf;
    uint32_t cnt;
    Q4 cur;              // words not consumed yet of the current chunk, in cur.x..
    uint32_t cur_n;      // how many
    Q4 nxt;              // the following chunk (loading)
    uint64_t nxt_chunk;  // index of the chunk after nxt

    GZ_INL void setup(const uint32_t* words, uint64_t size_bytes) {
        w = words;
        nfull = size_bytes >> 2;
        uint32_t r = (uint32_t)(size_bytes & 3);
        tail_mask = r ? (uint32_t)((1ull << (8 * r)) - 1) : 0;
        aligned = ((uintptr_t)words & 15) == 0;
    }
    GZ_INL uint32_t word(uint64_t i) const {
        if (i < nfull) return w[i];
        if (i == nfull && tail_mask) return w[i] & tail_mask;
        return 0;
    }
    GZ_INL Q4 chunk(uint64_t c) const {
        if (aligned && (c + 1) * 4 <= nfull) return ((const Q4*)w)[c];
        Q4 q;
        q.x = word(4 * c);
        q.y = word(4 * c + 1);
        q.z = word(4 * c + 2);
        q.w = word(4 * c + 3);
        return q;
    }
    GZ_INL uint32_t take() {
        if (!cur_n) {
            cur = nxt;
            cur_n = 4;
            nxt = chunk(nxt_chunk++);
        }
        uint32_t v = cur.x;
        cur.x = cur.y;
        cur.y = cur.z;
        cur.z = cur.w;
        cur_n--;
        next++;
        return v;
    }
    GZ_INL void refill() {
        while (cnt <= 32) {
            buf |= (uint64_t)take() << cnt;
            cnt += 32;
        }
    }
    GZ_INL void seek(uint64_t bit) {
        next = bit >> 5;
        const uint64_t c = next >> 2;
        cur = chunk(c);
        uint32_t skip = (uint32_t)(next & 3);
        for (uint32_t k = 0; k < skip; k++) {
            cur.x = cur.y;
            cur.y = cur.z;
            cur.z = cur.w;
        }
        cur_n = 4 - skip;
        nxt = chunk(c + 1);
        nxt_chunk = c + 2;
        buf = 0;
        cnt = 0;
        refill();
        uint32_t s = (uint32_t)(bit & 31);
        buf >>= s;
        cnt -= s;
    }
    GZ_INL uint64_t pos() const { return next * 32 - cnt; }
    GZ_INL uint32_t peek(int n) const { return (uint32_t)(buf & ((1ull << n) - 1)); }
    GZ_INL void drop(int n) {
        buf >>= n;
        cnt -= (uint32_t)n;
    }
    GZ_INL uint32_t get(int n) {
        if (cnt < (uint32_t)n) refill();
        uint32_t v = peek(n);
        drop(n);
        return v;
    }
};

// 64 bits of the stream at `bit` (zeros past the input)
GZ_INL uint64_t bits64_at(const Bits& proto, uint64_t bit) {
    uint64_t i = bit >> 5;
    uint32_t s = (uint32_t)(bit & 31);
    uint64_t a = proto.word(i), b = proto.word(i + 1), c = proto.word(i + 2);
    uint64_t lo = a | (b << 32);
    return s ? (lo >> s) | (c << (64 - s)) : lo;
}

// ------------------------------------------------------------------ canonical codes

// count/symbol arrays of a canonical code (puff's form). 0 complete, 1 incomplete, -1 over.
template <int NSYM>
GZ_INL int canon_build(const uint8_t* lens, int n, uint16_t* count, uint16_t* sym, int& nonzero) {
    for (int i = 0; i < 16; i++) count[i] = 0;
    for (int s = 0; s < n; s++) count[lens[s]]++;
    nonzero = n - count[0];
    count[0] = 0;
    int left = 1;
    for (int len = 1; len < 16; len++) {
        left = 2 * left - count[len];
        if (left < 0) return -1;
    }
    uint16_t offs[16];
    offs[1] = 0;
    for (int len = 1; len < 15; len++) offs[len + 1] = (uint16_t)(offs[len] + count[len]);
    for (int s = 0; s < n; s++)
        if (lens[s]) sym[offs[lens[s]]++] = (uint16_t)s;
    return left > 0 ? 1 : 0;
}

// decode by walking code lengths bit by bit from `v` (the next 15+ stream bits): symbol and
// length in *len, or -1
GZ_INL int canon_decode(const uint16_t* count, const uint16_t* sym, uint32_t v, int* len_out) {
    int code = 0, first = 0, index = 0;
    for (int len = 1; len < 16; len++) {
        code |= (int)(v & 1);
        v >>= 1;
        int c = count[len];
        if (code - c < first) {
            *len_out = len;
            return sym[index + (code - first)];
        }
        index += c;
        first = (first + c) << 1;
        code <<= 1;
    }
    return -1;
}

// canonical code values (MSB-first) of lens[0..n) into codes[], for lookup filling
GZ_INL void canon_codes(const uint8_t* lens, int n, uint16_t* codes) {
    uint16_t count[16];
    for (int i = 0; i < 16; i++) count[i] = 0;
    for (int s = 0; s < n; s++) count[lens[s]]++;
    count[0] = 0;
    uint16_t next[16];
    int code = 0;
    next[0] = 0;
    for (int b = 1; b < 16; b++) {
        code = (code + count[b - 1]) << 1;
        next[b] = (uint16_t)code;
    }
    for (int s = 0; s < n; s++) codes[s] = lens[s] ? next[lens[s]]++ : 0;
}

GZ_INL uint32_t reverse_bits(uint32_t c, int len) {
    uint32_t r = 0;
    for (int b = 0; b < len; b++) r |= ((c >> b) & 1u) << (len - 1 - b);
    return r;
}

// ------------------------------------------------------------------ dynamic header parse
// The code lengths of a dynamic block, zlib's acceptance rules (inflate.cuh:231-265, 393-402).
// br is just past BFINAL/BTYPE. Returns 0 and hlit/hdist, or -1.
struct HeaderParse {
    uint8_t lens[320];
    int hlit, hdist;
};

GZ_INL int parse_code_lengths(Bits& br, HeaderParse& h) {
    static constexpr uint8_t order[19] = GZ_ORDER;
    h.hlit = (int)br.get(5) + 257;
    h.hdist = (int)br.get(5) + 1;
    int hclen = (int)br.get(4) + 4;
    if (h.hlit > 286 || h.hdist > 30) return -1;
    uint8_t cl[19];
    for (int i = 0; i < 19; i++) cl[i] = 0;
    for (int i = 0; i < hclen; i++) cl[order[i]] = (uint8_t)br.get(3);
    uint16_t ccount[16], csym[19];
    int nz;
    if (canon_build<19>(cl, 19, ccount, csym, nz) != 0) return -1;  // must be complete
    int total = h.hlit + h.hdist;
    // Kraft sums (in units of 2^-15) of the literal/length and the distance lengths read so
    // far: past 2^15 a code is over-subscribed, which both users reject, so stop at once
    uint32_t ksum[2] = {0, 0};
    for (int i = 0; i < total;) {
        if (br.cnt < 32) br.refill();
        int len;
        int s = canon_decode(ccount, csym, (uint32_t)br.buf, &len);
        if (s < 0) return -1;
        br.drop(len);
        if (s < 16) {
            if (s) ksum[i >= h.hlit] += 1u << (15 - s);
            h.lens[i++] = (uint8_t)s;
        } else {
            int rep;
            uint8_t val = 0;
            if (s == 16) {
                if (i == 0) return -1;
                val = h.lens[i - 1];
                rep = 3 + (int)br.get(2);
            } else if (s == 17) {
                rep = 3 + (int)br.get(3);
            } else {
                rep = 11 + (int)br.get(7);
            }
            if (i + rep > total) return -1;
            while (rep--) {
                if (val) ksum[i >= h.hlit] += 1u << (15 - val);
                h.lens[i++] = val;
            }
        }
        if (ksum[0] > 32768u || ksum[1] > 32768u) return -1;
    }
    if (h.lens[256] == 0) return -1;
    return 0;
}

// Kraft state only: 0 complete, 1 incomplete, -1 over-subscribed
GZ_INL int kraft_state(const uint8_t* lens, int n, int& nonzero, int& ones) {
    int count[16];
    for (int i = 0; i < 16; i++) count[i] = 0;
    for (int s = 0; s < n; s++) count[lens[s]]++;
    nonzero = n - count[0];
    ones = count[1];
    int left = 1;
    for (int len = 1; len < 16; len++) {
        left = 2 * left - count[len];
        if (left < 0) return -1;
    }
    return left > 0 ? 1 : 0;
}

// zlib's rules on the parsed lengths: literal code complete or one code of length 1; distance
// code complete, empty, or one code of length 1
GZ_INL bool lengths_acceptable(const HeaderParse& h) {
    int nz, ones;
    int r = kraft_state(h.lens, h.hlit, nz, ones);
    if (r < 0 || (r > 0 && !(nz == 1 && ones == 1))) return false;
    r = kraft_state(h.lens + h.hlit, h.hdist, nz, ones);
    if (r < 0 || (r > 0 && !(nz == 0 || (nz == 1 && ones == 1)))) return false;
    return true;
}

// ------------------------------------------------------------------ block-start predicates
// inflate.cuh:285-332

GZ_INL bool header_fields_ok(uint64_t v) {
    return ((v >> 1) & 3) == 2 && ((v >> 3) & 31) <= 29 && ((v >> 8) & 31) <= 29;
}

// the code-length code over the 128 bits at a candidate (v1: bits 0-63, v2: 64-127) is complete
GZ_INL bool cl_code_complete(uint64_t v1, uint64_t v2) {
    int hclen = (int)((v1 >> 13) & 15) + 4;
    uint64_t f = (v1 >> 17) | (v2 << 47);  // 19 fields of 3 bits from bit 17
    int sum = 0;
    for (int i = 0; i < hclen; i++) {
        int l = (int)((f >> (3 * i)) & 7);
        sum += l ? 128 >> l : 0;
    }
    return sum == 128;
}

GZ_INL bool header_prefilter128(uint64_t v1, uint64_t v2) { return header_fields_ok(v1) && cl_code_complete(v1, v2); }

// Kraft sum (x 128) of four 3-bit code-length fields packed in 12 bits: the finder's lookup
GZ_INL uint32_t kraft4(uint32_t f12) {
    uint32_t sum = 0;
    for (int k = 0; k < 4; k++) {
        uint32_t l = (f12 >> (3 * k)) & 7;
        sum += l ? 128u >> l : 0u;
    }
    return sum;
}

// cl_code_complete with the sum taken from a 4096-entry kraft4 table
GZ_INL bool cl_code_complete_lut(uint64_t v1, uint64_t v2, const uint16_t* k4) {
    int hclen = (int)((v1 >> 13) & 15) + 4;
    uint64_t f = (v1 >> 17) | (v2 << 47);
    uint32_t sum = 0;
#pragma unroll
    for (int g = 0; g < 5; g++) {
        int nf = hclen - 4 * g;
        nf = nf < 0 ? 0 : nf > 4 ? 4 : nf;
        sum += k4[(uint32_t)(f >> (12 * g)) & ((1u << (3 * nf)) - 1)];
    }
    return sum == 128;
}

// The candidate block starts among the 32 bit positions of stream word `word` that lie in
// (bit_lo, bit_hi), as a bit mask (bit o: position 32*word + o). a..d are words word..word+3.
// The header fields (BTYPE 2, HLIT <= 29, HDIST <= 29) are tested for all 32 positions at once
// on bit planes (plane k, bit o = stream bit 32*word + o + k); the survivors get the Kraft sum
// of their code-length code, which must be complete (inflate.cuh:300-311).
GZ_INL uint32_t find_word_hits(uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint64_t word, uint64_t bit_lo,
                               uint64_t bit_hi) {
    uint64_t lo = (uint64_t)a | ((uint64_t)b << 32), hi = (uint64_t)c | ((uint64_t)d << 32);
#define GZ_PLANE(k) ((uint32_t)(lo >> (k)))
    uint32_t m = ~GZ_PLANE(1) & GZ_PLANE(2);                                 // BTYPE = 2
    m &= ~(GZ_PLANE(4) & GZ_PLANE(5) & GZ_PLANE(6) & GZ_PLANE(7));          // HLIT <= 29
    m &= ~(GZ_PLANE(9) & GZ_PLANE(10) & GZ_PLANE(11) & GZ_PLANE(12));       // HDIST <= 29
#undef GZ_PLANE
    uint64_t base = word * 32;
    if (base <= bit_lo) m &= bit_lo - base >= 31 ? 0u : ~0u << (uint32_t)(bit_lo - base + 1);
    if (base + 32 > bit_hi) m &= bit_hi <= base ? 0u : (uint32_t)((1ull << (bit_hi - base)) - 1);
    uint32_t hits = 0;
    while (m) {
        int o = __builtin_ctz(m);
        m &= m - 1;
        uint64_t v1 = o ? (lo >> o) | (hi << (64 - o)) : lo;
        uint64_t v2 = hi >> o;
        if (cl_code_complete(v1, v2)) hits |= 1u << o;
    }
    return hits;
}

// a dynamic block header that zlib would accept, not reading past the input
GZ_INL bool plausible_dynamic_header(const Bits& proto, uint64_t size_bits, uint64_t bit) {
    Bits br = proto;
    br.seek(bit);
    if (!header_fields_ok(br.peek(17))) return false;
    br.drop(3);
    HeaderParse h;
    if (parse_code_lengths(br, h) != 0) return false;
    if (!lengths_acceptable(h)) return false;
    return br.pos() <= size_bits;
}

// plausible_dynamic_header without the 320-entry length array: running Kraft sums, nonzero and
// length-1 counts per code, and the end-of-block length are all the predicate needs. `work` is
// 64 bytes of per-lane scratch (LDS on the GPU) for the code-length code. Same result.
GZ_INL bool plausible_dynamic_header_lean(const Bits& proto, uint64_t size_bits, uint64_t bit, uint8_t* work) {
    static constexpr uint8_t order[19] = GZ_ORDER;
    uint8_t* cl = work;         // 19: code-length code lengths by symbol
    uint8_t* cnt = work + 20;   // 8: codes per length
    uint8_t* offs = work + 28;  // 8
    uint8_t* sym = work + 36;   // 19: symbols in canonical order
    Bits br = proto;
    br.seek(bit);
    if (!header_fields_ok(br.peek(17))) return false;
    br.drop(3);
    const int hlit = (int)br.get(5) + 257;
    const int hdist = (int)br.get(5) + 1;
    const int hclen = (int)br.get(4) + 4;
    for (int k = 0; k < 19; k++) cl[k] = 0;
    for (int k = 0; k < hclen; k++) cl[order[k]] = (uint8_t)br.get(3);
    for (int l = 0; l < 8; l++) cnt[l] = 0;
    for (int k = 0; k < 19; k++) cnt[cl[k]]++;
    int left = 1;
    for (int l = 1; l < 8; l++) {
        left = 2 * left - cnt[l];
        if (left < 0) return false;
    }
    if (left != 0) return false;  // the code-length code must be complete
    offs[1] = 0;
    for (int l = 1; l < 7; l++) offs[l + 1] = (uint8_t)(offs[l] + cnt[l]);
    for (int k = 0; k < 19; k++)
        if (cl[k]) sym[offs[cl[k]]++] = (uint8_t)k;
    const int total = hlit + hdist;
    uint32_t ks0 = 0, ks1 = 0;  // Kraft sums (units of 2^-15): literal/length code, distance code
    int nz0 = 0, nz1 = 0, ones0 = 0, ones1 = 0, eob = 0, prev = 0;
    for (int i = 0; i < total;) {
        if (br.cnt < 32) br.refill();
        uint32_t v = (uint32_t)br.buf;
        int code = 0, first = 0, index = 0, s = -1, len = 0;
        for (int l = 1; l < 8; l++) {
            code |= (int)(v & 1);
            v >>= 1;
            int c = cnt[l];
            if (code - c < first) {
                s = sym[index + (code - first)];
                len = l;
                break;
            }
            index += c;
            first = (first + c) << 1;
            code <<= 1;
        }
        if (s < 0) return false;
        br.drop(len);
        int val, rep;
        if (s < 16) {
            val = s;
            rep = 1;
        } else if (s == 16) {
            if (i == 0) return false;
            val = prev;
            rep = 3 + (int)br.get(2);
        } else if (s == 17) {
            val = 0;
            rep = 3 + (int)br.get(3);
        } else {
            val = 0;
            rep = 11 + (int)br.get(7);
        }
        if (i + rep > total) return false;
        if (val == 0) {
            if (i <= 256 && 256 < i + rep) eob = 0;
            i += rep;
        } else {
            for (; rep; rep--, i++) {
                if (i < hlit) {
                    ks0 += 1u << (15 - val);
                    nz0++;
                    ones0 += val == 1;
                } else {
                    ks1 += 1u << (15 - val);
                    nz1++;
                    ones1 += val == 1;
                }
                if (i == 256) eob = val;
            }
            if (ks0 > 32768u || ks1 > 32768u) return false;
        }
        prev = val;
    }
    if (eob == 0) return false;
    if (ks0 != 32768u && !(nz0 == 1 && ones0 == 1)) return false;
    if (ks1 != 32768u && !(nz1 == 0 || (nz1 == 1 && ones1 == 1))) return false;
    return br.pos() <= size_bits;
}

// ------------------------------------------------------------------ table build (serial)
// Everything but the first-level lookup fill, which the kernel spreads over a workgroup.
// Returns the table status; lens/codes receive what the fill needs.
GZ_INL int32_t table_prepare(const Bits& proto, uint64_t header_bit, Table& t, uint8_t* lens, uint16_t* codes,
                             int& hlit, int& hdist) {
    Bits br = proto;
    br.seek(header_bit);
    uint32_t h3 = br.peek(3);
    uint32_t type = h3 >> 1;
    if (type == 1) return TAB_FIXED;
    if (type == 0) return TAB_STORED;
    if (type == 3) return TAB_BAD;
    br.drop(3);
    HeaderParse hp;
    if (parse_code_lengths(br, hp) != 0) return TAB_BAD;
    int nz;
    int r = canon_build<288>(hp.lens, hp.hlit, t.lcount, t.lsym, nz);
    if (r < 0 || (r > 0 && !(nz == 1 && t.lcount[1] == 1))) return TAB_BAD;
    r = canon_build<32>(hp.lens + hp.hlit, hp.hdist, t.dcount, t.dsym, nz);
    if (r < 0 || (r > 0 && !(nz == 0 || (nz == 1 && t.dcount[1] == 1)))) return TAB_BAD;
    t.hdr_bits = (uint32_t)(br.pos() - header_bit);
    hlit = hp.hlit;
    hdist = hp.hdist;
    for (int i = 0; i < hlit + hdist; i++) lens[i] = hp.lens[i];
    canon_codes(hp.lens, hlit, codes);
    canon_codes(hp.lens + hlit, hdist, codes + hlit);
    return TAB_DYNAMIC;
}

// the fixed code (RFC 1951 3.2.6) as a Table; lens/codes for the fill
GZ_INL void table_prepare_fixed(Table& t, uint8_t* lens, uint16_t* codes) {
    for (int i = 0; i < 288; i++) lens[i] = i < 144 ? 8 : i < 256 ? 9 : i < 280 ? 7 : 8;
    for (int i = 0; i < 30; i++) lens[288 + i] = 5;
    int nz;
    canon_build<288>(lens, 288, t.lcount, t.lsym, nz);
    canon_build<32>(lens + 288, 30, t.dcount, t.dsym, nz);
    canon_codes(lens, 288, codes);
    canon_codes(lens + 288, 30, codes + 288);
    t.hdr_bits = 3;
}

// first-level entries of one symbol (all k with k's low `len` bits = the reversed code)
GZ_INL void fill_lit_symbol(Table& t, int s, int len, uint16_t code) {
    if (!len || len > LIT_BITS) return;
    uint32_t e = lit_entry(s, len);
    for (uint32_t k = reverse_bits(code, len); k < (1u << LIT_BITS); k += 1u << len) t.lit[k] = e;
}
GZ_INL void fill_dist_symbol(Table& t, int s, int len, uint16_t code) {
    if (!len || len > DIST_BITS) return;
    uint32_t e = dist_entry(s, len);
    for (uint32_t k = reverse_bits(code, len); k < (1u << DIST_BITS); k += 1u << len) t.dist[k] = e;
}

// ------------------------------------------------------------------ decode

// one literal/length symbol's entry from the buffer (refilled to >= 32 bits); 0 = invalid code
template <class B>
GZ_INL uint32_t lit_lookup(const Table* T, const B& br) {
    uint32_t e = T->lit[br.buf & ((1u << LIT_BITS) - 1)];
    if (e & 15u) return e;
    int len;
    int s = canon_decode(T->lcount, T->lsym, (uint32_t)br.buf, &len);
    return s < 0 ? 0u : lit_entry(s, len);
}
template <class B>
GZ_INL uint32_t dist_lookup(const Table* T, const B& br) {
    uint32_t e = T->dist[br.buf & ((1u << DIST_BITS) - 1)];
    if (e & 15u) return e;
    int len;
    int s = canon_decode(T->dcount, T->dsym, (uint32_t)br.buf, &len);
    return s < 0 ? 0u : dist_entry(s, len);
}

// deferred stored-block copies (device); the host build always copies inline
struct StoredCopy {
    uint64_t src;  // input byte
    uint64_t dst;  // 16-bit pool index
    uint32_t len;
    uint32_t pad;
};
struct StoredList {
    StoredCopy* list;
    unsigned* count;
    unsigned cap;
};

// the table of the block whose header is at `key`, from the cache's sorted keys; null: none
GZ_INL const Table* cache_lookup(const Table* tabs, const uint64_t* keys, const uint32_t* slots, int nkeys,
                                 uint64_t key) {
    int lo = 0, hi = nkeys;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (keys[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < nkeys && keys[lo] == key ? tabs + slots[lo] : nullptr;
}

struct Job {
    const uint32_t* words;  // the stream
    uint64_t size;          // its bytes
    Start s;
    const Table* tabs;      // the cache; tabs[0] holds the fixed code
    const uint64_t* keys;   // the cache's sorted header keys and their slots
    const uint32_t* slots;
    int nkeys;
    const Table* own;       // the table of s (IN_DYNAMIC, or AT_BLOCK at a dynamic header); may be null
    const Start* targets;   // sorted by bit, all after s
    int ntargets;
    uint16_t* out;          // the job's region of the 16-bit pool
    uint64_t out_index;     // where that region starts in the pool (for deferred copies)
    uint32_t cap;
    bool stream_start;
    uint64_t stop;
};

// p[0..n) = v, with 16-byte stores where the alignment allows: a long match is a few stores
// (inflate.cuh:335-349 does the same)
GZ_INL void fill16(uint16_t* p, uint32_t n, uint16_t v) {
    while (n && ((uintptr_t)p & 15)) {
        *p++ = v;
        n--;
    }
    const uint32_t w = (uint32_t)v | ((uint32_t)v << 16);
#if defined(__HIP_DEVICE_COMPILE__)
    const uint4 q = make_uint4(w, w, w, w);
    for (; n >= 8; n -= 8, p += 8) *(uint4*)p = q;
#else
    const uint64_t q = (uint64_t)w | ((uint64_t)w << 32);
    for (; n >= 4; n -= 4, p += 4) memcpy(p, &q, 8);
#endif
    while (n--) *p++ = v;
}

GZ_INL void set_resume(Result& r, uint64_t bit, uint64_t header, uint8_t mode, int final) {
    r.resume.bit = bit;
    r.resume.header = header;
    r.resume.mode = mode;
    r.resume.final = (int8_t)final;
    for (int i = 0; i < 6; i++) r.resume.pad[i] = 0;
}

// Decoder::run for this design (inflate.cuh:474-581): decode from j.s until landing on a
// target, the end of the stream, the stop bit, a full buffer, or a dynamic block header whose
// table is not j.own (NEW_TABLE; resume there).
GZ_INL Result run(const Job& j, StoredList* stored) {
    Result r;
    r.end_bit = 0;
    r.eob_bit = UINT64_MAX;
    set_resume(r, 0, 0, AT_BLOCK, -1);
    r.eob_out = 0;
    r.out_len = 0;
    r.status = OK;
    r.landed = -1;
    r.end_final = j.s.mode == AT_BLOCK ? -1 : j.s.final;
    r.blocks = 0;
    r.reach = 0;
    r.reach_eob = 0;
    FBits br;
    br.setup(j.words, j.size);
    const uint64_t size_bits = j.size * 8;
    const Table* T = nullptr;
    uint32_t op = 0;
    int t = 0;
    bool in_block = false, fixed = false;
    uint64_t header = j.s.header;
    if (j.s.mode == IN_FIXED) {
        T = j.tabs;
        in_block = fixed = true;
    } else if (j.s.mode == IN_DYNAMIC) {
        if (!j.own || j.own->status != TAB_DYNAMIC) {
            r.status = ERR_DATA;
            r.end_bit = j.s.bit;
            return r;
        }
        T = j.own;
        in_block = true;
    }
    br.seek(j.s.bit);
    const uint64_t limit = j.stop < size_bits + 1 ? j.stop : size_bits + 1;
    for (;;) {
        if (!in_block) {
            uint64_t p = br.pos();
            while (t < j.ntargets && j.targets[t].bit < p) t++;
            if (t < j.ntargets && j.targets[t].bit == p && j.targets[t].mode == AT_BLOCK) {
                r.landed = t;
                break;
            }
            if (p >= j.stop || op + MAX_SYMBOL_OUT > j.cap || r.blocks >= MAX_BLOCKS) {
                r.status = p >= j.stop ? STOPPED : ERR_OVERFLOW;
                set_resume(r, p, p, AT_BLOCK, -1);
                break;
            }
            if (p >= size_bits) {
                r.status = ERR_INPUT;
                break;
            }
            if (br.cnt < 32) br.refill();
            uint32_t type = (br.peek(3) >> 1) & 3;
            const Table* dyn = nullptr;
            if (type == 2) {
                dyn = p == j.s.bit && j.s.mode == AT_BLOCK && j.own ? j.own
                                                                    : cache_lookup(j.tabs, j.keys, j.slots, j.nkeys, p);
                if (!dyn) {
                    // a dynamic block whose header has no table: it is built in the next round
                    r.status = NEW_TABLE;
                    set_resume(r, p, p, AT_BLOCK, -1);
                    break;
                }
            }
            header = p;
            int fin = (int)br.get(1);
            br.get(2);
            r.blocks++;
            r.end_final = fin;
            if (type == 0) {
                uint32_t rem = (uint32_t)(br.pos() & 7);
                if (rem) br.get((int)(8 - rem));
                uint32_t len = br.get(16), nlen = br.get(16);
                if (len != (~nlen & 0xffffu)) {
                    r.status = ERR_DATA;
                    break;
                }
                if ((uint64_t)op + len > j.cap) {
                    r.status = ERR_OVERFLOW;
                    set_resume(r, p, p, AT_BLOCK, -1);
                    break;
                }
                uint64_t byte = br.pos() >> 3;
                if (byte + len > j.size) {
                    r.status = ERR_INPUT;
                    break;
                }
                bool deferred = false;
#if defined(__HIP_DEVICE_COMPILE__)
                if (stored && len) {
                    unsigned i = atomicAdd(stored->count, 1u);
                    if (i < stored->cap) {
                        stored->list[i] = {byte, j.out_index + op, len, 0};
                        deferred = true;
                    }
                }
#else
                (void)stored;
#endif
                if (!deferred) {
                    const uint8_t* src = (const uint8_t*)j.words;
                    for (uint32_t k = 0; k < len; k++) j.out[op + k] = src[byte + k];
                }
                op += len;
                br.seek((byte + len) * 8);
                if (fin) break;
                continue;
            }
            if (type == 3) {
                r.status = ERR_DATA;
                break;
            }
            if (type == 1) {
                T = j.tabs;
                fixed = true;
            } else {
                if (dyn->status != TAB_DYNAMIC) {
                    r.status = ERR_DATA;
                    break;
                }
                T = dyn;
                br.seek(p + T->hdr_bits);
                fixed = false;
            }
            in_block = true;
        }
        // landing between symbols: every fixed block; a dynamic block when the next target
        // names its header
        bool land = fixed || (t < j.ntargets && j.targets[t].mode == IN_DYNAMIC && j.targets[t].header == header);
        const uint8_t mode = fixed ? IN_FIXED : IN_DYNAMIC;
        int how = 0;  // 0 end of block, 1 landed, -1 stopped (r.status)
        // the next target's bit, kept in a register: the list is read only once the decoder reaches it
        uint64_t tbit = land && t < j.ntargets ? j.targets[t].bit : UINT64_MAX;
        for (;;) {
            if (br.pos() >= tbit) {
                uint64_t p = br.pos();
                while (t < j.ntargets && j.targets[t].bit < p) t++;
                if (t < j.ntargets && j.targets[t].bit == p && j.targets[t].mode == mode &&
                    (fixed || j.targets[t].header == header)) {
                    r.landed = t;
                    how = 1;
                    break;
                }
                tbit = t < j.ntargets ? j.targets[t].bit : UINT64_MAX;
            }
            if (op + MAX_SYMBOL_OUT > j.cap || br.pos() >= limit) {
                uint64_t p = br.pos();
                if (p > size_bits) {
                    r.status = ERR_INPUT;
                    how = -1;
                    break;
                }
                r.status = p >= limit ? STOPPED : ERR_OVERFLOW;
                set_resume(r, p, header, mode, r.end_final);
                how = -1;
                break;
            }
            if (br.cnt < 32) br.refill();
            uint32_t e = lit_lookup(T, br);
            uint32_t kind = (e >> 8) & 3;
            if (!e || kind == K_BAD) {
                if (e) br.drop((int)(e & 15));
                r.status = ERR_DATA;
                how = -1;
                break;
            }
            br.drop((int)(e & 15));
            if (kind == K_LIT) {
                j.out[op++] = (uint16_t)(e >> 16);
                continue;
            }
            if (kind == K_EOB) break;
            uint32_t len = (e >> 16) + br.peek((int)((e >> 4) & 15));
            br.drop((int)((e >> 4) & 15));
            if (br.cnt < 32) br.refill();
            uint32_t de = dist_lookup(T, br);
            if (!de || (de & 0x100u)) {
                r.status = ERR_DATA;
                how = -1;
                break;
            }
            br.drop((int)(de & 15));
            uint32_t d = (de >> 16) + br.peek((int)((de >> 4) & 15));
            br.drop((int)((de >> 4) & 15));
            if (d > op) {
                if (j.stream_start) {
                    r.status = ERR_DATA;
                    how = -1;
                    break;
                }
                uint32_t reach = d - op;
                if (reach > r.reach) r.reach = reach;
                if (r.eob_bit == UINT64_MAX && reach > r.reach_eob) r.reach_eob = reach;
            }
            fill16(j.out + op, len, (uint16_t)(REF + d));
            op += len;
        }
        if (how != 0) break;
        in_block = false;
        if (br.pos() > size_bits) {
            r.status = ERR_INPUT;
            break;
        }
        if (r.eob_bit == UINT64_MAX) {
            r.eob_bit = br.pos();
            r.eob_out = op;
        }
        if (r.end_final == 1) break;
    }
    r.end_bit = br.pos();
    r.out_len = op;
    return r;
}

// Resynchronise inside a block (after inflate.cuh:591-648): with the code of the block at
// `header` (own: its table, or the fixed code for IN_FIXED), decode without writing from
// from_bit, restarting one bit later on data that makes no sense, and return the first
// symbol boundary at or after at_bit (UINT64_MAX: none). Past an end-of-block code it follows a
// fixed block; another dynamic or a stored block ends the attempt.
GZ_INL Start probe(const uint32_t* words, uint64_t size, const Table* tabs, const Table* own, uint8_t mode,
                   uint64_t header, uint64_t from_bit, uint64_t at_bit) {
    Start none;
    none.bit = UINT64_MAX;
    none.header = UINT64_MAX;
    none.mode = AT_BLOCK;
    none.final = -1;
    for (int i = 0; i < 6; i++) none.pad[i] = 0;
    FBits br;
    br.setup(words, size);
    bool fixed0 = mode == IN_FIXED;
    if (mode == AT_BLOCK) {
        br.seek(header);
        uint32_t type = (br.peek(3) >> 1) & 3;
        if (type == 0 || type == 3) return none;
        fixed0 = type == 1;
    }
    const Table* T0 = fixed0 ? tabs : own;
    if (!T0 || (!fixed0 && T0->status != TAB_DYNAMIC)) return none;
    uint64_t start = from_bit;
    for (int attempt = 0; attempt < 64 && start < at_bit; attempt++) {
        const Table* T = T0;
        bool fixed = fixed0;
        uint64_t cur = header;
        br.seek(start);
        uint64_t p;
        for (;;) {
            p = br.pos();
            if (p >= at_bit) {
                Start s = none;
                s.bit = p;
                s.header = cur;  // the block's header (a fixed block's too: it tells which block)
                s.mode = fixed ? IN_FIXED : IN_DYNAMIC;
                return s;
            }
            if (br.cnt < 32) br.refill();
            uint32_t e = lit_lookup(T, br);
            uint32_t kind = (e >> 8) & 3;
            if (!e || kind == K_BAD) break;
            br.drop((int)(e & 15));
            if (kind == K_LIT) continue;
            if (kind == K_EOB) {
                uint64_t h = br.pos();
                if (h >= at_bit) {
                    Start s = none;
                    s.bit = h;
                    s.header = h;
                    return s;
                }
                if (br.cnt < 32) br.refill();
                uint32_t ty = (br.peek(3) >> 1) & 3;
                if (ty != 1) break;
                br.drop(3);
                T = tabs;
                fixed = true;
                cur = h;
                continue;
            }
            br.drop((int)((e >> 4) & 15));
            if (br.cnt < 32) br.refill();
            uint32_t de = dist_lookup(T, br);
            if (!de || (de & 0x100u)) break;
            br.drop((int)(de & 15));
            br.drop((int)((de >> 4) & 15));
        }
        start = p + 1;
    }
    return none;
}

// ------------------------------------------------------------------ groups
// A group decodes one interval of the stream that starts at an exact state with GROUP_LANES
// lanes: lane 0 from the start, lane i from the first symbol boundary at or after an even split
// point (found by a probe with a short lead, in the interval's own code), each lane landing on
// a later lane's start or on a target after the interval; the group then walks its lanes'
// chain itself (the same rules the host applies to jobs), and reports it as one piece.
constexpr int GROUP_LANES = 32;
constexpr int GROUP_TARGETS = 16;  // targets after the interval a lane may land on

struct GroupDesc {
    Start s;             // the interval's start (exact)
    uint64_t end;        // where it ends: the next target
    uint64_t pool;       // lane l writes at pool + l * lane_cap
    uint64_t split_end;  // lanes split [first symbol, split_end): stored data after it is the last lane's
    uint32_t lane_cap;
    uint32_t last_extra; // room the last lane has beyond lane_cap
    int32_t tfirst;      // the first target after s
    int32_t job;
};
static_assert(sizeof(GroupDesc) == 64, "GroupDesc layout");

struct GroupResult {
    Result r;                  // the chain as one piece: out_len, exit status/landed/resume/end_bit,
                               // the chain's first end-of-block (eob_bit, eob_out), end_final, reach
    uint32_t chain;            // lanes in the chain (bit l)
    uint32_t eob_lane;         // lane of the chain's first end-of-block code (GROUP_LANES: none)
    uint32_t eob_len;          // that lane's symbols up to it
    uint32_t pad;
    uint32_t len[GROUP_LANES]; // symbols each chain lane contributes
};

// lane i's split point in [first, end)
GZ_INL uint64_t group_point(uint64_t first, uint64_t end, int i) {
    return first + (end - first) * (uint64_t)i / GROUP_LANES;
}

// the chain of lane results from lane 0. lane_of[t] for a target index t < nv is the lane that
// starts there; landed indices are relative to each lane's own target list, which starts at
// `first_t[lane]` of the combined list (valid lane starts, then the targets after the interval).
GZ_INL void group_walk(const Start* st, const Result* res, const int* lane_of, const int* first_t, int nv,
                       GroupResult& g) {
    g.r.end_bit = 0;
    g.r.eob_bit = UINT64_MAX;
    g.r.eob_out = 0;
    g.r.out_len = 0;
    g.r.status = OK;
    g.r.landed = -1;
    g.r.blocks = 0;
    g.r.reach = 0;
    g.r.reach_eob = 0;
    set_resume(g.r, 0, 0, AT_BLOCK, -1);
    g.chain = 0;
    g.eob_lane = GROUP_LANES;
    g.eob_len = 0;
    g.pad = 0;
    for (int l = 0; l < GROUP_LANES; l++) g.len[l] = 0;
    int final = st[0].final;  // BFINAL of the block the next lane starts in (-1 unknown here)
    uint32_t total = 0;
    int k = 0;
    for (int steps = 0; steps < GROUP_LANES; steps++) {
        const Result& r = res[k];
        const Start& s = st[k];
        const bool last = s.mode != AT_BLOCK && s.final < 0 && final == 1 && r.eob_bit != UINT64_MAX;
        const uint32_t len = last ? r.eob_out : r.out_len;
        const uint32_t reach = last ? r.reach_eob : r.reach;
        if (reach > total && reach - total > g.r.reach) g.r.reach = reach - total;
        if (g.eob_lane == GROUP_LANES) {
            const uint32_t re = r.eob_bit != UINT64_MAX ? r.reach_eob : reach;
            if (re > total && re - total > g.r.reach_eob) g.r.reach_eob = re - total;
            if (r.eob_bit != UINT64_MAX) {
                g.eob_lane = (uint32_t)k;
                g.eob_len = r.eob_out;
                g.r.eob_bit = r.eob_bit;
                g.r.eob_out = total + r.eob_out;
            }
        }
        g.chain |= 1u << k;
        g.len[k] = len;
        total += len;
        g.r.blocks += r.blocks;
        if (r.end_final >= 0) final = r.end_final;
        if (last) {
            g.r.end_bit = r.eob_bit;
            break;
        }
        const int t = r.landed >= 0 ? first_t[k] + r.landed : -1;
        if (r.status != OK || t < 0 || t >= nv) {
            // the chain leaves the group here
            g.r.status = r.status;
            g.r.landed = t < 0 ? -1 : t - nv;
            g.r.resume = r.resume;
            g.r.end_bit = r.end_bit;
            break;
        }
        k = lane_of[t];
    }
    g.r.out_len = total;
    g.r.end_final = final;
}

}  // namespace gz
