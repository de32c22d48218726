/* inflate_oracle.c - CPU restatement of the reference Deflate decoder's semantics.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP decoder. Loaded by tests/, smoke() and
 * bench.py's cpu_baseline leg; never by the product path.
 *
 * Pinned by (tests/test_oracle.py): system zlib's inflate (bytes and consumed input) over the
 * eight-kind stream corpus and the config-0/1 generators; the CRC-32 known answer
 * "123456789" -> 0xCBF43926 and zlib.crc32; brute-force enumeration of every header the
 * predicates accept on tiny inputs against an independent bit-by-bit parse; the reference's
 * own decoder (gpuinflate/inflate.cuh compiled from its sources into oracle/_ref) on fuzzed
 * streams: the same status, bytes and in_used.
 */
#include "inflate_oracle.h"

#include <stdlib.h>
#include <string.h>

/* RFC 1951 3.2.5 length and distance codes, 3.2.7 code-length order (the format fixes them) */
static const uint16_t LEN_BASE[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                                      31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
static const uint8_t LEN_EXTRA[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
static const uint16_t DIST_BASE[30] = {1,   2,   3,   4,   5,   7,    9,    13,   17,   25,   33,   49,   65,    97,    129,
                                       193, 257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
static const uint8_t DIST_EXTRA[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
static const uint8_t CL_ORDER[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

/* ------------------------------------------------------------------ CRC-32 */

static uint32_t crc_tab[256];
static int crc_ready;

static void crc_init(void) {
    for (uint32_t n = 0; n < 256; n++) {
        uint32_t c = n;
        for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
        crc_tab[n] = c;
    }
    crc_ready = 1;
}

uint32_t or_crc32(uint32_t crc, const uint8_t* p, size_t n) {
    if (!crc_ready) crc_init();
    crc = ~crc;
    for (size_t i = 0; i < n; i++) crc = crc_tab[(crc ^ p[i]) & 0xff] ^ (crc >> 8);
    return ~crc;
}

/* ------------------------------------------------------------------ bits
 * Position-based LSB-first reader; bytes past the input read as zero. `pos` counts bits from
 * the start of the input, exactly the reference's Bits::pos() (inflate.cuh:144). */

typedef struct {
    const uint8_t* in;
    uint64_t size;
    uint64_t pos;
} bits_t;

uint64_t or_bits64(const uint8_t* in, uint64_t size, uint64_t bit) {
    uint64_t byte = bit >> 3;
    uint8_t b[16];
    for (int i = 0; i < 16; i++) b[i] = byte + i < size ? in[byte + i] : 0;
    uint64_t lo = 0, hi = 0;
    for (int i = 7; i >= 0; i--) lo = lo << 8 | b[i];
    for (int i = 15; i >= 8; i--) hi = hi << 8 | b[i];
    int s = (int)(bit & 7);
    return s ? (lo >> s) | (hi << (64 - s)) : lo;
}

static uint32_t peek(const bits_t* b, int n) {
    if (n == 0) return 0;
    uint64_t v;
    uint64_t byte = b->pos >> 3;
    if (byte + 8 <= b->size) {
        memcpy(&v, b->in + byte, 8);
        v >>= (b->pos & 7);
    } else {
        v = or_bits64(b->in, b->size, b->pos);
    }
    return (uint32_t)(v & ((1ull << n) - 1));
}

static uint32_t getb(bits_t* b, int n) {
    uint32_t v = peek(b, n);
    b->pos += (uint64_t)n;
    return v;
}

/* ------------------------------------------------------------------ Huffman
 * Canonical code kept as puff.c does: how many codes of each length, and the symbols in
 * canonical order. Decoding walks the lengths one bit at a time. Acceptance rules are zlib's
 * (inflate.cuh:161-199, 231-265, 393-411). */

typedef struct {
    uint16_t count[16];
    uint16_t symbol[288];
    int nonzero;
} huff_t;

/* 0 complete, 1 incomplete, -1 over-subscribed */
static int huff_build(huff_t* h, const uint8_t* lens, int n) {
    for (int i = 0; i < 16; i++) h->count[i] = 0;
    for (int s = 0; s < n; s++) h->count[lens[s]]++;
    h->nonzero = n - h->count[0];
    h->count[0] = 0;
    int left = 1;
    for (int len = 1; len < 16; len++) {
        left = left * 2 - h->count[len];
        if (left < 0) return -1;
    }
    uint16_t offs[16];
    offs[1] = 0;
    for (int len = 1; len < 15; len++) offs[len + 1] = (uint16_t)(offs[len] + h->count[len]);
    for (int s = 0; s < n; s++)
        if (lens[s]) h->symbol[offs[lens[s]]++] = (uint16_t)s;
    return left > 0 ? 1 : 0;
}

static int huff_single(const huff_t* h) { return h->nonzero == 1 && h->count[1] == 1; }

static int huff_decode(const huff_t* h, bits_t* b) {
    uint32_t v = peek(b, 15);
    int code = 0, first = 0, index = 0;
    for (int len = 1; len < 16; len++) {
        code |= (int)(v & 1);
        v >>= 1;
        int c = h->count[len];
        if (code - c < first) {
            b->pos += (uint64_t)len;
            return h->symbol[index + (code - first)];
        }
        index += c;
        first = (first + c) << 1;
        code <<= 1;
    }
    return -1;
}

/* after BFINAL and BTYPE: the code lengths of a dynamic block. 0 = valid */
static int code_lengths(bits_t* b, uint8_t* lens, int* hlit_out, int* hdist_out) {
    int hlit = (int)getb(b, 5) + 257;
    int hdist = (int)getb(b, 5) + 1;
    int hclen = (int)getb(b, 4) + 4;
    if (hlit > 286 || hdist > 30) return -1;
    uint8_t cl[19];
    memset(cl, 0, sizeof cl);
    for (int i = 0; i < hclen; i++) cl[CL_ORDER[i]] = (uint8_t)getb(b, 3);
    huff_t clh;
    if (huff_build(&clh, cl, 19) != 0) return -1; /* zlib: the code-length code must be complete */
    int total = hlit + hdist;
    int i = 0;
    while (i < total) {
        int s = huff_decode(&clh, b);
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
            rep = 3 + (int)getb(b, 2);
        } else if (s == 17) {
            rep = 3 + (int)getb(b, 3);
        } else {
            rep = 11 + (int)getb(b, 7);
        }
        if (i + rep > total) return -1;
        while (rep--) lens[i++] = val;
    }
    if (lens[256] == 0) return -1; /* no end-of-block code */
    *hlit_out = hlit;
    *hdist_out = hdist;
    return 0;
}

typedef struct {
    huff_t lit, dist;
} tables_t;

static int dynamic_tables(tables_t* t, bits_t* b) {
    uint8_t lens[320];
    int hlit, hdist;
    if (code_lengths(b, lens, &hlit, &hdist) != 0) return -1;
    int r = huff_build(&t->lit, lens, hlit);
    if (r < 0 || (r > 0 && !huff_single(&t->lit))) return -1;
    r = huff_build(&t->dist, lens + hlit, hdist);
    if (r < 0 || (r > 0 && !(t->dist.nonzero == 0 || huff_single(&t->dist)))) return -1;
    return 0;
}

static void fixed_tables(tables_t* t) {
    uint8_t lens[288];
    for (int i = 0; i < 288; i++) lens[i] = i < 144 ? 8 : i < 256 ? 9 : i < 280 ? 7 : 8;
    huff_build(&t->lit, lens, 288);
    for (int i = 0; i < 30; i++) lens[i] = 5;
    huff_build(&t->dist, lens, 30);
}

/* ------------------------------------------------------------------ predicates
 * inflate.cuh:285-332 */

static int hdr_fields_ok(uint64_t v) {
    return ((v >> 1) & 3) == 2 && ((v >> 3) & 31) <= 29 && ((v >> 8) & 31) <= 29;
}

int or_header_prefilter(uint64_t v) {
    if (!hdr_fields_ok(v)) return 0;
    int hclen = (int)((v >> 13) & 15) + 4;
    int n = hclen < 15 ? hclen : 15; /* 17 + 3*15 = 62: what 64 bits hold */
    int sum = 0;
    for (int i = 0; i < n; i++) {
        int l = (int)((v >> (17 + 3 * i)) & 7);
        if (l) sum += 128 >> l;
    }
    return sum <= 128 && (hclen > 15 || sum == 128);
}

int or_header_prefilter128(uint64_t v1, uint64_t v2) {
    if (!hdr_fields_ok(v1)) return 0;
    int hclen = (int)((v1 >> 13) & 15) + 4;
    int sum = 0;
    for (int i = 0; i < hclen; i++) {
        int p = 17 + 3 * i; /* the 3-bit field at bits p..p+2 of the 128 */
        uint32_t l;
        if (p + 3 <= 64)
            l = (uint32_t)(v1 >> p);
        else if (p >= 64)
            l = (uint32_t)(v2 >> (p - 64));
        else
            l = (uint32_t)((v1 >> p) | (v2 << (64 - p)));
        l &= 7;
        if (l) sum += 128 >> l;
    }
    return sum == 128;
}

/* Kraft sum of lens[0..n): 0 complete, 1 incomplete, -1 over-subscribed */
static int kraft(const uint8_t* lens, int n, int* nonzero, int* ones) {
    int count[16] = {0};
    for (int s = 0; s < n; s++) count[lens[s]]++;
    *nonzero = n - count[0];
    *ones = count[1];
    int left = 1;
    for (int len = 1; len < 16; len++) {
        left = left * 2 - count[len];
        if (left < 0) return -1;
    }
    return left > 0 ? 1 : 0;
}

int or_plausible_dynamic_header(const uint8_t* in, uint64_t size, uint64_t bit) {
    bits_t b = {in, size, bit};
    if (!hdr_fields_ok(peek(&b, 17))) return 0;
    b.pos += 3;
    uint8_t lens[320];
    int hlit, hdist;
    if (code_lengths(&b, lens, &hlit, &hdist) != 0) return 0;
    int nz, ones;
    int r = kraft(lens, hlit, &nz, &ones);
    if (r < 0 || (r > 0 && !(nz == 1 && ones == 1))) return 0;
    r = kraft(lens + hlit, hdist, &nz, &ones);
    if (r < 0 || (r > 0 && !(nz == 0 || (nz == 1 && ones == 1)))) return 0;
    return b.pos <= size * 8; /* not past the input */
}

uint64_t or_find_headers(const uint8_t* in, uint64_t size, uint64_t lo, uint64_t hi, uint64_t* out, uint64_t cap) {
    uint64_t n = 0;
    for (uint64_t b = lo; b < hi; b++) {
        uint64_t v1 = or_bits64(in, size, b);
        if (!or_header_prefilter(v1)) continue;
        if (!or_header_prefilter128(v1, or_bits64(in, size, b + 64))) continue;
        if (!or_plausible_dynamic_header(in, size, b)) continue;
        if (n < cap) out[n] = b;
        n++;
    }
    return n;
}

/* ------------------------------------------------------------------ run
 * Decoder::run (inflate.cuh:418-581), restated as one loop with an explicit in-block flag. */

static void fill_ref(uint16_t* p, uint32_t n, uint16_t v) {
    for (uint32_t k = 0; k < n; k++) p[k] = v;
}

or_result or_run(const uint8_t* in, uint64_t size, or_start s, const or_start* targets, int ntargets, uint16_t* out,
                 uint32_t cap, int stream_start, uint64_t stop) {
    or_result r;
    memset(&r, 0, sizeof r);
    r.eob_bit = UINT64_MAX;
    r.status = OR_OK;
    r.landed = -1;
    r.end_final = s.mode == OR_AT_BLOCK ? -1 : s.final;
    r.last_block_bit = UINT64_MAX;
    const uint64_t size_bits = size * 8;
    tables_t tab;
    bits_t b = {in, size, 0};
    uint32_t op = 0;
    int t = 0;
    int in_block = 0, fixed = 0;
    uint64_t header = s.header;
    if (s.mode == OR_IN_FIXED) {
        fixed_tables(&tab);
        in_block = fixed = 1;
    } else if (s.mode == OR_IN_DYNAMIC) {
        b.pos = s.header + 3;
        if (dynamic_tables(&tab, &b) != 0) {
            r.status = OR_ERR_DATA;
            r.end_bit = s.bit;
            return r;
        }
        in_block = 1;
    }
    b.pos = s.bit;
    const uint64_t limit = stop < size_bits + 1 ? stop : size_bits + 1;
    for (;;) {
        if (!in_block) {
            uint64_t p = b.pos;
            while (t < ntargets && targets[t].bit < p) t++;
            if (t < ntargets && targets[t].bit == p && targets[t].mode == OR_AT_BLOCK) {
                r.landed = t;
                break;
            }
            if (p >= stop || op + OR_MAX_SYMBOL_OUT > cap || r.blocks >= OR_MAX_BLOCKS) {
                r.status = p >= stop ? OR_STOPPED : OR_ERR_OVERFLOW;
                r.resume.bit = p;
                r.resume.header = p;
                r.resume.mode = OR_AT_BLOCK;
                r.resume.final = -1;
                break;
            }
            if (p >= size_bits) {
                r.status = OR_ERR_INPUT;
                break;
            }
            header = p;
            r.last_block_bit = p;
            r.last_block_out = op;
            int fin = (int)getb(&b, 1);
            int type = (int)getb(&b, 2);
            r.blocks++;
            r.end_final = fin;
            if (type == 0) {
                if (b.pos & 7) b.pos += 8 - (b.pos & 7);
                uint32_t len = getb(&b, 16), nlen = getb(&b, 16);
                if (len != (~nlen & 0xffffu)) {
                    r.status = OR_ERR_DATA;
                    break;
                }
                if ((uint64_t)op + len > cap) {
                    r.status = OR_ERR_OVERFLOW;
                    r.resume.bit = p;
                    r.resume.header = p;
                    r.resume.mode = OR_AT_BLOCK;
                    r.resume.final = -1;
                    break;
                }
                uint64_t byte = b.pos >> 3;
                if (byte + len > size) {
                    r.status = OR_ERR_INPUT;
                    break;
                }
                for (uint32_t k = 0; k < len; k++) out[op + k] = in[byte + k];
                op += len;
                b.pos = (byte + len) * 8;
                if (fin) break;
                continue;
            }
            if (type == 3) {
                r.status = OR_ERR_DATA;
                break;
            }
            if (type == 1) {
                fixed_tables(&tab);
                fixed = 1;
            } else {
                if (dynamic_tables(&tab, &b) != 0) {
                    r.status = OR_ERR_DATA;
                    break;
                }
                fixed = 0;
            }
            in_block = 1;
        }
        /* landing between symbols is tried only where the next target could be one for this
         * block: every fixed block, or a dynamic block whose header the next target names */
        int land = fixed || (t < ntargets && targets[t].mode == OR_IN_DYNAMIC && targets[t].header == header);
        uint8_t mode = fixed ? OR_IN_FIXED : OR_IN_DYNAMIC;
        int how = 0; /* 0 end of block, 1 landed, -1 stopped with r.status */
        for (;;) {
            if (land) {
                uint64_t p = b.pos;
                while (t < ntargets && targets[t].bit < p) t++;
                if (t < ntargets && targets[t].bit == p && targets[t].mode == mode && (fixed || targets[t].header == header)) {
                    r.landed = t;
                    how = 1;
                    break;
                }
            }
            if (op + OR_MAX_SYMBOL_OUT > cap || b.pos >= limit) {
                uint64_t p = b.pos;
                if (p > size_bits) {
                    r.status = OR_ERR_INPUT;
                    how = -1;
                    break;
                }
                r.status = p >= limit ? OR_STOPPED : OR_ERR_OVERFLOW;
                r.resume.bit = p;
                r.resume.header = header;
                r.resume.mode = mode;
                r.resume.final = (int8_t)r.end_final;
                how = -1;
                break;
            }
            int sym = huff_decode(&tab.lit, &b);
            if (sym < 256) {
                if (sym < 0) {
                    r.status = OR_ERR_DATA;
                    how = -1;
                    break;
                }
                out[op++] = (uint16_t)sym;
                r.symbols++;
                continue;
            }
            if (sym == 256) break;
            r.symbols++;
            sym -= 257;
            if (sym >= 29) {
                r.status = OR_ERR_DATA;
                how = -1;
                break;
            }
            uint32_t len = LEN_BASE[sym] + getb(&b, LEN_EXTRA[sym]);
            int ds = huff_decode(&tab.dist, &b);
            if (ds < 0 || ds >= 30) {
                r.status = OR_ERR_DATA;
                how = -1;
                break;
            }
            uint32_t d = DIST_BASE[ds] + getb(&b, DIST_EXTRA[ds]);
            if (d > op && (stream_start || d - op > OR_WINDOW)) {
                r.status = OR_ERR_DATA;
                how = -1;
                break;
            }
            fill_ref(out + op, len, (uint16_t)(OR_REF + d));
            op += len;
        }
        if (how != 0) break;
        in_block = 0;
        if (b.pos > size_bits) {
            r.status = OR_ERR_INPUT;
            break;
        }
        if (r.eob_bit == UINT64_MAX) {
            r.eob_bit = b.pos;
            r.eob_out = op;
        }
        if (r.end_final == 1) break;
    }
    r.end_bit = b.pos;
    r.out_len = op;
    return r;
}

/* ------------------------------------------------------------------ whole stream
 * gi_test.cpp:47-91: decode from bit 0, continuing after every full buffer, then resolve the
 * back-references in order. */

int or_inflate(const uint8_t* in, uint64_t size, uint8_t* out, uint64_t max_out, uint64_t* out_len, uint64_t* in_used,
               int* too_big) {
    *out_len = 0;
    *in_used = 0;
    *too_big = 0;
    uint64_t n16 = max_out + 2 * OR_MAX_SYMBOL_OUT;
    uint16_t* o16 = (uint16_t*)malloc(n16 * sizeof(uint16_t));
    if (!o16) return -1;
    or_start s = {0, 0, OR_AT_BLOCK, -1};
    uint64_t op = 0;
    int status = OR_GI_OK;
    for (;;) {
        uint64_t room = n16 - op;
        uint32_t cap = room > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)room;
        or_result r = or_run(in, size, s, NULL, 0, o16 + op, cap, op == 0, UINT64_MAX);
        op += r.out_len;
        if (r.status == OR_ERR_OVERFLOW) {
            if (n16 - op < 64 * 1024) {
                *too_big = 1;
                free(o16);
                return OR_GI_OK;
            }
            s = r.resume;
            continue;
        }
        if (r.status != OR_OK) {
            status = r.status == OR_ERR_INPUT ? OR_GI_TRUNCATED : OR_GI_DATA;
            free(o16);
            return status;
        }
        *in_used = (r.end_bit + 7) / 8;
        break;
    }
    if (op > max_out) {
        *too_big = 1;
        free(o16);
        return OR_GI_OK;
    }
    for (uint64_t i = 0; i < op; i++) {
        uint16_t v = o16[i];
        if (v <= OR_REF) {
            out[i] = (uint8_t)v;
        } else if ((uint64_t)(v - OR_REF) > i) {
            free(o16);
            *in_used = 0;
            return OR_GI_DATA; /* a match before the stream's start */
        } else {
            out[i] = out[i - (v - OR_REF)];
        }
    }
    *out_len = op;
    free(o16);
    return status;
}
