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
// A code's decoding table. Codes up to LIT_BITS / DIST_BITS bits: one first-level entry.
// Longer codes: their length is found by comparing the next 15 bits, read MSB-first, with the
// left-justified end of each longer length's code range (`llim`/`dlim`: lengths ROOT+1.. at
// [0..], each `end | (offset - first code) << 16`), and their entry is then read from the
// entries in canonical order (`lent`/`dent`) -- no loop, so lanes of a wavefront that meet long
// codes at different times stay together.
struct Table {
    uint16_t lit[1 << LIT_BITS];
    uint32_t dist[1 << DIST_BITS];
    uint32_t llim[8];
    uint32_t dlim[8];
    uint16_t lent[288];
    uint32_t dent[32];
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

// A literal/length entry (16 bits): bits 0-3 the code's length (0: no code of at most LIT_BITS
// bits there, or none at all). A length symbol has bit 15 set, its extra bits in 4-6 and its base
// length - 3 in 7-14; any other symbol has bit 15 clear, its kind in 4-5 (K_LIT, K_EOB, K_BAD)
// and a literal's byte in 6-13.
GZ_INL uint32_t lit_entry(int sym, int len) {
    static constexpr uint16_t lb[29] = GZ_LBASE;
    static constexpr uint8_t le[29] = GZ_LEXT;
    if (sym < 256) return (uint32_t)len | (K_LIT << 4) | ((uint32_t)sym << 6);
    if (sym == 256) return (uint32_t)len | (K_EOB << 4);
    if (sym > 285) return (uint32_t)len | (K_BAD << 4);
    int i = sym - 257;
    return (uint32_t)len | ((uint32_t)le[i] << 4) | ((uint32_t)(lb[i] - 3) << 7) | 0x8000u;
}
GZ_INL uint32_t lit_kind(uint32_t e) { return (e & 0x8000u) ? (uint32_t)K_LEN : (e >> 4) & 3; }

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
    GZ_INL void fill32() {
        if (cnt < 32) refill();
    }
    GZ_INL uint64_t bits() const { return buf; }
};

struct alignas(16) Q4 {
    uint32_t x, y, z, w;
};

// The decoders' reader: the same LSB-first bit buffer as Bits, fed from words staged in fast
// memory (LDS on the device; `stg[k * stride]` holds word sbase + k, k < SW). A decoder then
// reads no device memory between stagings, so on the GPU its loop never waits on the stores it
// has in flight (loads and stores share one counter there). Staging is done by all decoders of a
// wavefront together (see run()), one wait for all of them.
constexpr int SW = 24;

struct SBits {
    const uint32_t* w;
    uint64_t nfull;
    uint32_t tail_mask;
    bool aligned;
    uint32_t* stg;
    int stride;
    uint64_t sbase;    // the first staged word
    uint64_t next;     // the next word to take into buf
    uint64_t buf;
    uint32_t cnt;      // valid bits in buf
    uint32_t nw;       // the staged word at `next`, read ahead
    bool pend;         // a seek waits for staging
    uint64_t pend_bit;

    GZ_INL void setup(const uint32_t* words, uint64_t size_bytes, uint32_t* stage_words, int stage_stride) {
        w = words;
        nfull = size_bytes >> 2;
        uint32_t r = (uint32_t)(size_bytes & 3);
        tail_mask = r ? (uint32_t)((1ull << (8 * r)) - 1) : 0;
        aligned = ((uintptr_t)words & 15) == 0;
        stg = stage_words;
        stride = stage_stride;
        sbase = 1ull << 62;  // nothing staged
        next = 0;
        buf = 0;
        cnt = 0;
        nw = 0;
        pend = false;
        pend_bit = 0;
    }
    GZ_INL uint32_t word(uint64_t i) const {
        if (i < nfull) return w[i];
        if (i == nfull && tail_mask) return w[i] & tail_mask;
        return 0;
    }
    GZ_INL uint32_t avail() const { return next >= sbase && next <= sbase + SW ? (uint32_t)(sbase + SW - next) : 0; }
    GZ_INL uint32_t staged(uint64_t i) const {
        uint64_t k = i - sbase;
        return stg[(int)(k < (uint64_t)SW ? k : SW - 1) * stride];
    }
    GZ_INL void pull() {
        buf |= (uint64_t)nw << cnt;
        cnt += 32;
        next++;
        nw = staged(next);
    }
    GZ_INL void finish_seek() {
        next = pend_bit >> 5;
        buf = 0;
        cnt = 0;
        nw = staged(next);
        pull();
        pull();
        uint32_t s = (uint32_t)(pend_bit & 31);
        buf >>= s;
        cnt -= s;
        pend = false;
    }
    // SW words from the word in use (a pending seek's) rounded down to 16 bytes
    GZ_INL void stage() {
        const uint64_t b = (pend ? pend_bit >> 5 : next) & ~3ull;
        sbase = b;
        if (aligned && b + SW <= nfull) {
            const Q4* q = (const Q4*)w + (b >> 2);
#pragma unroll
            for (int c = 0; c < SW / 4; c++) {
                const Q4 v = q[c];
                stg[(4 * c) * stride] = v.x;
                stg[(4 * c + 1) * stride] = v.y;
                stg[(4 * c + 2) * stride] = v.z;
                stg[(4 * c + 3) * stride] = v.w;
            }
        } else {
            for (int k = 0; k < SW; k++) stg[k * stride] = word(b + k);
        }
        if (pend)
            finish_seek();
        else
            nw = staged(next);
    }
    // to `bit`: at once when it is staged (with room for a step after it), else at the next stage()
    GZ_INL void seek(uint64_t bit) {
        pend = true;
        pend_bit = bit;
        const uint64_t wi = bit >> 5;
        if (wi >= sbase && wi + 4 <= sbase + SW) finish_seek();
    }
    GZ_INL uint64_t pos() const { return pend ? pend_bit : next * 32 - cnt; }
    GZ_INL void need32() {
        if (cnt < 32) pull();
    }
    // need32 without a branch: the staged word after the next one is read either way (for the
    // symbol loop, whose lanes would otherwise split on it at almost every step)
    GZ_INL void need32_flat() {
        const bool p = cnt < 32;
        const uint64_t nx = next + (p ? 1 : 0);
        const uint32_t ahead = staged(nx);
        buf = p ? buf | ((uint64_t)nw << (cnt & 31)) : buf;
        cnt = p ? cnt + 32 : cnt;
        next = nx;
        nw = p ? ahead : nw;
    }
    GZ_INL uint32_t peek(int n) const { return (uint32_t)(buf & ((1ull << n) - 1)); }
    GZ_INL void drop(int n) {
        buf >>= n;
        cnt -= (uint32_t)n;
    }
    GZ_INL uint32_t get(int n) {
        if (cnt < (uint32_t)n) pull();
        uint32_t v = peek(n);
        drop(n);
        return v;
    }
};

// SBits for one lane on its own (the header checks): stages its next words whenever the staged
// ones run out
struct VBits {
    SBits s;
    GZ_INL void seek(uint64_t bit) {
        s.seek(bit);
        if (s.pend) s.stage();
    }
    GZ_INL void more() {
        if (s.avail() < 1) s.stage();
        s.pull();
    }
    GZ_INL void fill32() {
        if (s.cnt < 32) more();
    }
    GZ_INL uint32_t peek(int n) const { return s.peek(n); }
    GZ_INL void drop(int n) { s.drop(n); }
    GZ_INL uint32_t get(int n) {
        if (s.cnt < (uint32_t)n) more();
        uint32_t v = s.peek(n);
        s.drop(n);
        return v;
    }
    GZ_INL uint64_t pos() const { return s.pos(); }
    GZ_INL uint64_t bits() const { return s.buf; }
};

// whether p holds for any running lane of the wavefront (the lane itself on the host)
GZ_INL bool wave_any(bool p) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ballot_w64(p) != 0;
#else
    return p;
#endif
}

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

// cl_code_complete from g, the 64 stream bits from HCLEN (a candidate's bit 13): the lengths past
// HCLEN + 4 masked off, the Kraft sum taken from a 4096-entry kraft4 table, four lengths a lookup
// (the table may hold its sums saturated at 255: any group over 128 fails either way)
template <class K>
GZ_INL bool cl_code_complete_g(uint64_t g, const K* k4) {
    const uint32_t nf = 3 * ((uint32_t)(g & 15) + 4);
    const uint64_t f = (g >> 4) & ((1ull << nf) - 1);
    const uint32_t sum = (uint32_t)k4[f & 0xFFF] + k4[(f >> 12) & 0xFFF] + k4[(f >> 24) & 0xFFF] +
                         k4[(f >> 36) & 0xFFF] + k4[(f >> 48) & 0x1FF];
    return sum == 128;
}

// the same over the 128 bits at a candidate (v1: bits 0-63, v2: 64-127)
GZ_INL bool cl_code_complete_lut(uint64_t v1, uint64_t v2, const uint16_t* k4) {
    return cl_code_complete_g((v1 >> 13) | (v2 << 51), k4);
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
// 64 bytes of per-lane scratch (LDS on the GPU) for the code-length code. Same result. With LENS
// it also writes the code lengths to lens[0 .. hlit + hdist) and hlit, hdist (a table build);
// br is then just past the header.
template <class B, bool LENS = false>
GZ_INL bool plausible_dynamic_header_lean(B& br, uint64_t size_bits, uint64_t bit, uint8_t* work,
                                          uint8_t* lens = nullptr, int* hl_out = nullptr, int* hd_out = nullptr) {
    static constexpr uint8_t order[19] = GZ_ORDER;
    uint8_t* cl = work;         // 19: code-length code lengths by symbol
    uint8_t* cnt = work + 20;   // 8: codes per length
    uint8_t* offs = work + 28;  // 8
    uint8_t* sym = work + 36;   // 19: symbols in canonical order
    br.seek(bit);
    if (!header_fields_ok(br.peek(17))) return false;
    br.drop(3);
    const int hlit = (int)br.get(5) + 257;
    const int hdist = (int)br.get(5) + 1;
    const int hclen = (int)br.get(4) + 4;
    if (LENS) {
        *hl_out = hlit;
        *hd_out = hdist;
    }
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
    // decoding without a loop: the left-justified (7-bit) end of each length's codes, and the
    // offset that turns a code into its index among the symbols in canonical order
    uint32_t lim[8];
    int base[8];
    {
        uint32_t first = 0;
        for (int l = 1; l < 8; l++) {
            lim[l] = (first + cnt[l]) << (7 - l);
            base[l] = (int)offs[l] - (int)first;
            first = (first + cnt[l]) << 1;
        }
    }
    for (int k = 0; k < 19; k++)
        if (cl[k]) sym[offs[cl[k]]++] = (uint8_t)k;
    const int total = hlit + hdist;
    uint32_t ks0 = 0, ks1 = 0;  // Kraft sums (units of 2^-15): literal/length code, distance code
    int nz0 = 0, nz1 = 0, ones0 = 0, ones1 = 0, eob = 0, prev = 0;
    for (int i = 0; i < total;) {
        br.fill32();
        const uint32_t v = (uint32_t)br.bits();
#if defined(__HIP_DEVICE_COMPILE__)
        const uint32_t rev = __builtin_bitreverse32(v) >> 25;
#else
        const uint32_t rev = reverse_bits(v & 127, 7);
#endif
        int len = 1;
#pragma unroll
        for (int l = 1; l < 7; l++) len += rev >= lim[l] ? 1 : 0;
        // the code is complete, so every 7-bit pattern starts a code
        const int s = sym[base[len] + (int)(rev >> (7 - len))];
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
            if (LENS)
                for (int r = 0; r < rep; r++) lens[i + r] = 0;
            i += rep;
        } else {
            for (; rep; rep--, i++) {
                if (LENS) lens[i] = (uint8_t)val;
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

// The longer-code levels of a Table (see there) for code lengths lens[0..n) and first-level
// ROOT bits: ends and offsets of lengths ROOT+1..15, and every symbol's entry in canonical order.
template <int ROOT, bool LIT, class E>
GZ_INL void long_levels(const uint8_t* lens, int n, uint32_t* lim, E* ent) {
    uint16_t count[16], offs[16], first[16];
    for (int i = 0; i < 16; i++) count[i] = 0;
    for (int s = 0; s < n; s++) count[lens[s]]++;
    count[0] = 0;
    offs[0] = offs[1] = 0;
    for (int len = 1; len < 15; len++) offs[len + 1] = (uint16_t)(offs[len] + count[len]);
    uint32_t code = 0;
    first[0] = 0;
    for (int b = 1; b < 16; b++) {
        code = (code + count[b - 1]) << 1;
        first[b] = (uint16_t)code;
    }
    uint16_t at[16];
    for (int i = 0; i < 16; i++) at[i] = offs[i];
    for (int s = 0; s < n; s++)
        if (lens[s]) ent[at[lens[s]]++] = (E)(LIT ? lit_entry(s, lens[s]) : dist_entry(s, lens[s]));
    for (int k = 0; k < 8; k++) {
        const int len = ROOT + 1 + k;
        if (len > 15) {
            lim[k] = 0;
            continue;
        }
        uint32_t end = ((uint32_t)first[len] + count[len]) << (15 - len);
        if (end > 32768u) end = 32768u;  // an over-full code never gets here (rejected earlier)
        int32_t base = (int32_t)offs[len] - (int32_t)first[len];
        lim[k] = end | ((uint32_t)(uint16_t)(int16_t)base << 16);
    }
}

// The entry of a code longer than ROOT bits at the front of buf (>= 15 valid bits), or 0
template <int ROOT, class PL, class PE>
GZ_INL uint32_t long_code(PL lim, PE ent, uint64_t buf) {
    constexpr int NL = 15 - ROOT;
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t rev = __builtin_bitreverse32((uint32_t)buf) >> 17;
#else
    const uint32_t rev = reverse_bits((uint32_t)buf & 0x7FFF, 15);
#endif
    uint32_t cur = lim[0];
    int k = 0;
#pragma unroll
    for (int i = 1; i < NL; i++) {
        const uint32_t nx = lim[i];
        const bool past = rev >= (cur & 0xFFFFu);
        cur = past ? nx : cur;
        k += past ? 1 : 0;
    }
    if (rev >= (cur & 0xFFFFu)) return 0;
    const int len = ROOT + 1 + k;
    return ent[(int32_t)(int16_t)(cur >> 16) + (int32_t)(rev >> (15 - len))];
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
    if (!lengths_acceptable(hp)) return TAB_BAD;
    long_levels<LIT_BITS, true>(hp.lens, hp.hlit, t.llim, t.lent);
    long_levels<DIST_BITS, false>(hp.lens + hp.hlit, hp.hdist, t.dlim, t.dent);
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
    long_levels<LIT_BITS, true>(lens, 288, t.llim, t.lent);
    long_levels<DIST_BITS, false>(lens + 288, 30, t.dlim, t.dent);
    canon_codes(lens, 288, codes);
    canon_codes(lens + 288, 30, codes + 288);
    t.hdr_bits = 3;
}

// first-level entries of one symbol (all k with k's low `len` bits = the reversed code)
GZ_INL void fill_lit_symbol(Table& t, int s, int len, uint16_t code) {
    if (!len || len > LIT_BITS) return;
    uint32_t e = lit_entry(s, len);
    for (uint32_t k = reverse_bits(code, len); k < (1u << LIT_BITS); k += 1u << len) t.lit[k] = (uint16_t)e;
}
GZ_INL void fill_dist_symbol(Table& t, int s, int len, uint16_t code) {
    if (!len || len > DIST_BITS) return;
    uint32_t e = dist_entry(s, len);
    for (uint32_t k = reverse_bits(code, len); k < (1u << DIST_BITS); k += 1u << len) t.dist[k] = e;
}

// ------------------------------------------------------------------ decode

// one literal/length symbol's entry from the buffer (refilled to >= 32 bits); length 0 = invalid
// TP: a Table pointer, generic or (on the device) into LDS
template <class TP, class B>
GZ_INL uint32_t lit_lookup(TP T, const B& br) {
    uint32_t e = T->lit[br.buf & ((1u << LIT_BITS) - 1)];
    if (e & 15u) return e;
    return long_code<LIT_BITS>(&T->llim[0], &T->lent[0], br.buf);
}
template <class TP, class B>
GZ_INL uint32_t dist_lookup(TP T, const B& br) {
    uint32_t e = T->dist[br.buf & ((1u << DIST_BITS) - 1)];
    if (e & 15u) return e;
    return long_code<DIST_BITS>(&T->dlim[0], &T->dent[0], br.buf);
}

// Tables in LDS are read through an LDS pointer on the device: a generic pointer would make
// every lookup wait for the decoder's stores in flight as well
#if defined(__HIP_DEVICE_COMPILE__)
#define GZ_LDS __attribute__((address_space(3)))
GZ_INL bool in_lds(const void* p) { return __builtin_amdgcn_is_shared(p); }
#else
#define GZ_LDS
GZ_INL bool in_lds(const void*) { return false; }
#endif
typedef const GZ_LDS Table* LdsTable;

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

// A decoder's landing targets, sorted by bit: na entries of a (through the index ai when given:
// a group's later lane starts, kept where the group probed them), then nb entries of b (the
// targets after the interval, where the planner put them)
struct TargetList {
    const Start* a = nullptr;
    const int* ai = nullptr;
    int na = 0;
    const Start* b = nullptr;
    int nb = 0;
    GZ_INL const Start& operator[](int t) const { return t < na ? a[ai ? ai[t] : t] : b[t - na]; }
};
GZ_INL TargetList target_list(const Start* t, int n) {
    TargetList l;
    l.a = t;
    l.na = n;
    return l;
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
    TargetList targets;     // sorted by bit, all after s
    int ntargets;           // targets.na + targets.nb
    uint16_t* out;          // the job's region of the 16-bit pool
    uint64_t out_index;     // where that region starts in the pool (for deferred copies)
    uint32_t cap;
    bool stream_start;
    uint64_t stop;
    uint64_t hard_end;      // a stored block's LEN bit: Huffman data reaching it stops the decoder
                            // (ERR_OVERFLOW, resumable), until the decoder has passed a stored block.
                            // A correct decode crosses that bit only inside the stored block, so a
                            // lane that reaches it otherwise started wrong (and stops early) or met a
                            // false stored candidate (and its continuation goes on without the bound)
    const Table* fixed;     // the fixed code's table (a copy in LDS, or tabs)
    uint32_t* stg;          // this decoder's staging words (stride apart)
    int stride;
};

// the furthest one symbol's stores reach past its start: 258 bytes in 8-symbol stores
constexpr uint32_t SYMBOL_SPAN = 264;  // (a decode that writes nothing reserves MAX_SYMBOL_OUT)

// p[0..8) = v: one 16-byte store (symbols past the one being written are written again later,
// by the same lane, in order); gfx950 runs with unaligned memory access enabled
GZ_INL void put8(uint16_t* p, uint16_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
    typedef uint4 __attribute__((aligned(2))) U4;
    const uint32_t x = (uint32_t)v | ((uint32_t)v << 16);
    *(U4*)p = make_uint4(x, x, x, x);
#else
    for (int i = 0; i < 8; i++) p[i] = v;
#endif
}

GZ_INL void set_resume(Result& r, uint64_t bit, uint64_t header, uint8_t mode, int final) {
    r.resume.bit = bit;
    r.resume.header = header;
    r.resume.mode = mode;
    r.resume.final = (int8_t)final;
    for (int i = 0; i < 6; i++) r.resume.pad[i] = 0;
}

// events that end a wavefront's run of symbols
enum : uint32_t { EV_NONE = 0, EV_STAGE = 1, EV_LAND = 2, EV_CAP = 3, EV_EOB = 4, EV_BAD = 5 };

template <uint32_t SPAN = SYMBOL_SPAN>
GZ_INL uint32_t sym_check(const SBits& br, uint32_t op, uint32_t cap, uint64_t tbit, uint64_t limit) {
    const uint64_t p = br.pos();
    return p >= tbit                         ? EV_LAND
           : (op + SPAN > cap || p >= limit) ? EV_CAP
           : br.avail() < 2                         ? EV_STAGE
                                                    : EV_NONE;
}

// run()'s symbols in lockstep: every lane of the wavefront decodes one literal or match per
// step (a match's 16-bit symbols written with 16-byte stores) until any lane has an event
// OUT false: nothing is written (a decode that only counts, as a header's verification needs)
//
// A step is one straight sequence for every lane: the literal/length code, then (for a length; a
// literal's lanes read but consume nothing) the distance code, the bit buffer refilled without a
// branch before each; a lane's outcome (a symbol, the end of its block, a bad code) is chosen by
// selects, so the lanes of a wavefront part only for codes longer than the first lookup level
// and for a match's stores past its first 8 symbols.
template <bool OUT, class TP>
GZ_INL uint32_t symbols(TP T, SBits& br, const Job& j, Result& r, uint32_t& op, uint64_t tbit, uint64_t limit,
                        uint32_t ev) {
    while (!wave_any(ev != EV_NONE)) {
        br.need32_flat();
        const uint32_t e = lit_lookup(T, br);
        const uint32_t kind = lit_kind(e);
        const uint32_t cl = e & 15;
        br.drop((int)cl);
        const bool is_len = kind == K_LEN;
        // a length's extra bits (none for a literal)
        const int lx = is_len ? (int)((e >> 4) & 7) : 0;
        const uint32_t v = ((e >> 7) & 255) + 3 + br.peek(lx);
        br.drop(lx);
        br.need32_flat();
        uint32_t de = T->dist[br.buf & ((1u << DIST_BITS) - 1)];
        if (!(de & 15u) && is_len) de = long_code<DIST_BITS>(&T->dlim[0], &T->dent[0], br.buf);
        const bool dist_ok = de && !(de & 0x100u);
        const bool take_d = is_len && dist_ok;
        const int dl = take_d ? (int)(de & 15) : 0;
        br.drop(dl);
        const int dx = take_d ? (int)((de >> 4) & 15) : 0;
        const uint32_t d = (de >> 16) + br.peek(dx);
        br.drop(dx);
        const bool far = take_d && d > op;
        const bool bad = !cl || kind == K_BAD || (is_len && !dist_ok) || (far && j.stream_start);
        if (far && !j.stream_start) {
            const uint32_t reach = d - op;
            if (reach > r.reach) r.reach = reach;
            if (r.eob_bit == UINT64_MAX && reach > r.reach_eob) r.reach_eob = reach;
        }
        if (bad) {
            ev = EV_BAD;
            continue;
        }
        if (kind == K_EOB) {
            ev = EV_EOB;
            continue;
        }
        const uint32_t len = is_len ? v : 1u;
        const uint16_t val = is_len ? (uint16_t)(REF + d) : (uint16_t)((e >> 6) & 255);
        if (OUT) {
            // a literal is one 2-byte store, a match's symbols go 8 at a time: the stores are the
            // compute unit's memory pipeline's work, shared by its wavefronts (gathering literals
            // into 16-byte stores was measured slower: the selects cost more than they saved)
            if (is_len) {
                put8(j.out + op, val);
                for (uint32_t k = 8; k < len; k += 8) put8(j.out + op + k, val);
            } else {
                j.out[op] = val;
            }
        }
        op += len;
        ev = sym_check<OUT ? SYMBOL_SPAN : MAX_SYMBOL_OUT>(br, op, j.cap, tbit, limit);
    }
    return ev;
}

// Speculative decode from j.s (after inflate.cuh:474-581): until landing exactly on a target
// (AT_BLOCK at a block start; IN_FIXED / IN_DYNAMIC with the same header between symbols), a
// full buffer (ERR_OVERFLOW) or the stop bit (STOPPED) with the state to resume from, a
// dynamic header that has no table (NEW_TABLE), the final block's end (OK), or an error. Output:
// one 16-bit symbol per byte, v < 256 a literal, v > REF the byte (v - REF) back. Stored blocks
// are copied later (deferred) when a list is given.
//
// The decoders of a wavefront run in rounds: each round stages input where needed (all lanes
// together), handles block starts, then decodes symbols in lockstep until any lane meets an
// event (landing check, buffer end, end of block, bad code, staging running out); each lane
// then handles its own event. The common step has no loop and no branch that the lanes of a
// wavefront take apart for long.
template <bool OUT = true>
GZ_INL Result run(const Job& j, StoredList* stored) {
    constexpr uint32_t SPAN = OUT ? SYMBOL_SPAN : MAX_SYMBOL_OUT;
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
    const uint64_t size_bits = j.size * 8;
    const Table* T = nullptr;
    uint32_t op = 0;
    int t = 0;
    bool in_block = false, fixed = false;
    uint64_t header = j.s.header;
    if (j.s.mode == IN_FIXED) {
        T = j.fixed;
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
    bool entered = in_block;
    SBits br;
    br.setup(j.words, j.size, j.stg, j.stride);
    br.seek(j.s.bit);
    const uint64_t stop_limit = j.stop < size_bits + 1 ? j.stop : size_bits + 1;
    uint64_t hard = j.hard_end;
    uint64_t limit = stop_limit < hard ? stop_limit : hard;
    uint64_t tbit = UINT64_MAX;
    for (;;) {
        // staging: when one decoder needs it, every decoder of the wavefront reloads
        if (wave_any(br.pend || br.avail() < 3)) br.stage();
        if (!in_block) {
            // a block's start
            const uint64_t p = br.pos();
            while (t < j.ntargets && j.targets[t].bit < p) t++;
            if (t < j.ntargets && j.targets[t].bit == p && j.targets[t].mode == AT_BLOCK) {
                r.landed = t;
                break;
            }
            if (p >= j.stop || op + SPAN > j.cap || r.blocks >= MAX_BLOCKS) {
                r.status = p >= j.stop ? STOPPED : ERR_OVERFLOW;
                set_resume(r, p, p, AT_BLOCK, -1);
                break;
            }
            if (p >= hard) {
                // resumable: the stored block there may be a false candidate
                r.status = ERR_OVERFLOW;
                set_resume(r, p, p, AT_BLOCK, -1);
                break;
            }
            if (p >= size_bits) {
                r.status = ERR_INPUT;
                break;
            }
            br.need32();
            const uint32_t type = (br.peek(3) >> 1) & 3;
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
            const int fin = (int)br.get(1);
            br.get(2);
            r.blocks++;
            r.end_final = fin;
            if (type == 0) {
                const uint32_t rem = (uint32_t)(br.pos() & 7);
                if (rem) br.get((int)(8 - rem));
                const uint32_t len = br.get(16), nlen = br.get(16);
                if (len != (~nlen & 0xffffu)) {
                    r.status = ERR_DATA;
                    break;
                }
                if ((uint64_t)op + len > j.cap) {
                    r.status = ERR_OVERFLOW;
                    set_resume(r, p, p, AT_BLOCK, -1);
                    break;
                }
                const uint64_t byte = br.pos() >> 3;
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
                if (!deferred && OUT) {
                    const uint8_t* src = (const uint8_t*)j.words;
                    for (uint32_t k = 0; k < len; k++) j.out[op + k] = src[byte + k];
                }
                op += len;
                br.seek((byte + len) * 8);
                hard = UINT64_MAX;
                limit = stop_limit;
                if (fin) break;
                continue;
            }
            if (type == 3) {
                r.status = ERR_DATA;
                break;
            }
            if (type == 1) {
                T = j.fixed;
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
            entered = true;
        }
        if (entered && !br.pend) {
            // landing between symbols: every fixed block; a dynamic block when the next target
            // names its header (the next target's bit is kept in a register)
            const bool land =
                fixed || (t < j.ntargets && j.targets[t].mode == IN_DYNAMIC && j.targets[t].header == header);
            tbit = land && t < j.ntargets ? j.targets[t].bit : UINT64_MAX;
            entered = false;
        }
        // symbols, while no decoder of the wavefront meets an event
        uint32_t ev = EV_STAGE;
        if (!br.pend && !entered) ev = sym_check<SPAN>(br, op, j.cap, tbit, limit);
        if (in_lds(T))
            ev = symbols<OUT>((LdsTable)T, br, j, r, op, tbit, limit, ev);
        else
            ev = symbols<OUT>(T, br, j, r, op, tbit, limit, ev);
        if (ev == EV_LAND) {
            const uint64_t p = br.pos();
            while (t < j.ntargets && j.targets[t].bit < p) t++;
            const uint8_t mode = fixed ? IN_FIXED : IN_DYNAMIC;
            if (t < j.ntargets && j.targets[t].bit == p && j.targets[t].mode == mode &&
                (fixed || j.targets[t].header == header)) {
                r.landed = t;
                break;
            }
            // a target here of another kind (a block start, another header) is passed over, as
            // the reference passes it (inflate.cuh:422-428): the next check is at a later target.
            // Left as the next target, it would end every step at once and the decoder (and its
            // wavefront) would never move again.
            while (t < j.ntargets && j.targets[t].bit <= p) t++;
            tbit = t < j.ntargets ? j.targets[t].bit : UINT64_MAX;
        } else if (ev == EV_CAP) {
            const uint64_t p = br.pos();
            if (p > size_bits) {
                r.status = ERR_INPUT;
                break;
            }
            // a full buffer or `hard` (resumable: the stored block there may be a false candidate)
            r.status = p >= stop_limit ? STOPPED : ERR_OVERFLOW;
            set_resume(r, p, header, fixed ? IN_FIXED : IN_DYNAMIC, r.end_final);
            break;
        } else if (ev == EV_EOB) {
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
        } else if (ev == EV_BAD) {
            r.status = ERR_DATA;
            break;
        }
    }
    r.end_bit = br.pos();
    r.out_len = op;
    return r;
}

// probe()'s symbols in lockstep, without output; p0 follows the current symbol's first bit
template <class TP>
GZ_INL uint32_t probe_symbols(TP T, SBits& br, uint64_t at_bit, uint64_t& p0, uint32_t ev) {
    while (!wave_any(ev != EV_NONE)) {
        br.need32();
        const uint32_t e = lit_lookup(T, br);
        const uint32_t kind = lit_kind(e);
        if (!(e & 15) || kind == K_BAD) {
            ev = EV_BAD;
            continue;
        }
        br.drop((int)(e & 15));
        if (kind == K_EOB) {
            ev = EV_EOB;
            continue;
        }
        if (kind == K_LEN) {
            br.drop((int)((e >> 4) & 7));
            br.need32();
            const uint32_t de = dist_lookup(T, br);
            if (!de || (de & 0x100u)) {
                ev = EV_BAD;
                continue;
            }
            br.drop((int)(de & 15));
            br.drop((int)((de >> 4) & 15));
        }
        p0 = br.pos();
        ev = p0 >= at_bit ? EV_LAND : br.avail() < 2 ? EV_STAGE : EV_NONE;
    }
    return ev;
}

// Resynchronise inside a block (after inflate.cuh:591-648): with the code of the block at
// `header` (own: its table; fixed: the fixed code's), decode without writing from from_bit,
// restarting one bit later on data that makes no sense (`tries` attempts; the reference makes 64),
// and return the first symbol
// boundary at or after at_bit (bit UINT64_MAX: none). Past an end-of-block code it follows a
// fixed block; another dynamic or a stored block ends the attempt. Runs in wavefront rounds as
// run() does; stg/stride: this lane's staging words.
GZ_INL Start probe(const uint32_t* words, uint64_t size, const Table* fixed_tab, const Table* own, uint8_t mode,
                   uint64_t header, uint64_t from_bit, uint64_t at_bit, uint32_t* stg, int stride, int tries = 64) {
    Start none;
    none.bit = UINT64_MAX;
    none.header = UINT64_MAX;
    none.mode = AT_BLOCK;
    none.final = -1;
    for (int i = 0; i < 6; i++) none.pad[i] = 0;
    SBits br;
    br.setup(words, size, stg, stride);
    bool fixed0 = mode == IN_FIXED;
    if (mode == AT_BLOCK) {
        const uint64_t v = br.word(header >> 5) | (uint64_t)br.word((header >> 5) + 1) << 32;
        const uint32_t type = (uint32_t)(v >> (header & 31) >> 1) & 3;
        if (type == 0 || type == 3) return none;
        fixed0 = type == 1;
    }
    const Table* T0 = fixed0 ? fixed_tab : own;
    if (!T0 || (!fixed0 && T0->status != TAB_DYNAMIC)) return none;
    if (from_bit >= at_bit) return none;
    const Table* T = T0;
    bool fixed = fixed0;
    uint64_t cur = header;
    uint64_t start = from_bit;
    int attempt = 0;
    br.seek(start);
    Start res = none;
    for (;;) {
        if (wave_any(br.pend || br.avail() < 3)) br.stage();
        uint32_t ev = EV_STAGE;
        uint64_t p0 = br.pos();  // the current symbol's first bit (a restart goes one past it)
        if (!br.pend) ev = p0 >= at_bit ? EV_LAND : br.avail() < 2 ? EV_STAGE : EV_NONE;
        if (in_lds(T))
            ev = probe_symbols((LdsTable)T, br, at_bit, p0, ev);
        else
            ev = probe_symbols(T, br, at_bit, p0, ev);
        if (ev == EV_LAND) {
            res.bit = p0;
            res.header = cur;  // the block's header (a fixed block's too: it tells which block)
            res.mode = fixed ? IN_FIXED : IN_DYNAMIC;
            break;
        }
        if (ev == EV_EOB) {
            const uint64_t h = br.pos();
            if (h >= at_bit) {
                res.bit = h;
                res.header = h;
                break;
            }
            br.need32();
            if (((br.peek(3) >> 1) & 3) == 1) {
                br.drop(3);
                T = fixed_tab;
                fixed = true;
                cur = h;
                continue;
            }
            ev = EV_BAD;
        }
        if (ev == EV_BAD) {
            start = p0 + 1;
            if (++attempt >= tries || start >= at_bit) break;
            T = T0;
            fixed = fixed0;
            cur = header;
            br.seek(start);
        }
    }
    return res;
}

// ------------------------------------------------------------------ groups
// A group decodes one interval of the stream that starts at an exact state with L lanes (16, 32
// or 64: a quarter, a half or all of a wavefront; a schedule field, GROUP_LANES_DEFAULT unless
// tuned): lane 0 from the start, lane i from the first symbol boundary at or after an even split
// point (found by a probe with a short lead, in the interval's own code), each lane landing on
// a later lane's start or on a target after the interval; the group then walks its lanes'
// chain itself (the same rules the host applies to jobs), and reports it as one piece.
constexpr int GROUP_LANES_MAX = 64;      // a wavefront
constexpr int GROUP_LANES_DEFAULT = 64;  // a group a wavefront (measured 4% faster than two on 4 GiB)
GZ_INL constexpr bool group_lanes_ok(int l) { return l == 16 || l == 32 || l == 64; }
constexpr int GROUP_TARGETS = 16;  // targets after the interval a lane may land on

struct GroupDesc {
    Start s;             // the interval's start (exact)
    uint64_t end;        // where it ends: the next target
    uint64_t pool;       // lane l writes at pool + l * lane_cap
    uint64_t split_end;  // lanes split [first symbol, split_end): stored data after it is the last lane's
    uint64_t hard;       // that stored block's LEN bit (0: none): Huffman data reaching it stops a lane
    uint32_t lane_cap;
    uint32_t last_extra; // room the last lane has beyond lane_cap
    int32_t tfirst;      // the first target after s
    int32_t job;
    uint64_t lim;        // the end (byte) of the input its stream owns (0: all of it)
};
static_assert(sizeof(GroupDesc) == 80, "GroupDesc layout");

struct GroupResult {
    Result r;                  // the chain as one piece: out_len, exit status/landed/resume/end_bit,
                               // the chain's first end-of-block (eob_bit, eob_out), end_final, reach
    uint64_t chain;            // lanes in the chain (bit l)
    uint32_t eob_lane;         // lane of the chain's first end-of-block code (GROUP_LANES_MAX: none)
    uint32_t eob_len;          // that lane's symbols up to it
    uint32_t len[GROUP_LANES_MAX]; // symbols each chain lane contributes
};
// what the host's chain walk reads of a GroupResult (its lane lengths stay on the device)
struct GroupHead {
    Result r;
    uint64_t chain;
    uint32_t eob_lane, eob_len;
};
static_assert(sizeof(GroupHead) == 88 && offsetof(GroupResult, len) == sizeof(GroupHead), "GroupHead is GroupResult's head");

// lane i's split point in [first, end), of a group of `lanes`
GZ_INL uint64_t group_point(uint64_t first, uint64_t end, int i, int lanes) {
    return first + (end - first) * (uint64_t)i / (uint64_t)lanes;
}

// the chain of lane results from lane 0. lane_of[t] for a target index t < nv is the lane that
// starts there; landed indices are relative to each lane's own target list, which starts at
// `first_t[lane]` of the combined list (valid lane starts, then the targets after the interval).
GZ_INL void group_walk(const Start* st, const Result* res, const int* lane_of, const int* first_t, int nv, int lanes,
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
    g.eob_lane = GROUP_LANES_MAX;
    g.eob_len = 0;
    for (int l = 0; l < lanes; l++) g.len[l] = 0;
    int final = st[0].final;  // BFINAL of the block the next lane starts in (-1 unknown here)
    uint32_t total = 0;
    int k = 0;
    for (int steps = 0; steps < lanes; steps++) {
        const Result& r = res[k];
        const Start& s = st[k];
        const bool last = s.mode != AT_BLOCK && s.final < 0 && final == 1 && r.eob_bit != UINT64_MAX;
        const uint32_t len = last ? r.eob_out : r.out_len;
        const uint32_t reach = last ? r.reach_eob : r.reach;
        if (reach > total && reach - total > g.r.reach) g.r.reach = reach - total;
        if (g.eob_lane == GROUP_LANES_MAX) {
            const uint32_t re = r.eob_bit != UINT64_MAX ? r.reach_eob : reach;
            if (re > total && re - total > g.r.reach_eob) g.r.reach_eob = re - total;
            if (r.eob_bit != UINT64_MAX) {
                g.eob_lane = (uint32_t)k;
                g.eob_len = r.eob_out;
                g.r.eob_bit = r.eob_bit;
                g.r.eob_out = total + r.eob_out;
            }
        }
        g.chain |= 1ull << k;
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
