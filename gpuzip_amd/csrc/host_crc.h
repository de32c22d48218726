// host_crc.h - gzip's CRC-32 on the host (reflected polynomial 0xEDB88320, register preset and
// final complement), for checking bytes after they reach host memory.
//
// Long buffers are folded 64 bytes a step with carry-less multiplication (PCLMULQDQ), the method
// of Gopal et al., "Fast CRC Computation for Generic Polynomials Using PCLMULQDQ Instruction"
// (Intel, 2009): four 128-bit lanes, each folded 512 bits forward by multiplying its two halves by
// x^(512+32) and x^(512-32) mod P; then the lanes folded into one by x^(128+32) and x^(128-32);
// the last 16 bytes and the tail go through a byte table from a zero register (the folded value
// is congruent, modulo P, to the bytes it replaces). The constants are computed from P when first
// needed. Without PCLMULQDQ, slice-by-8 tables.
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#if defined(__x86_64__)
#include <immintrin.h>
#endif

namespace hostcrc {

struct Tables {
    uint32_t t[8][256];
    uint64_t k512[2], k128[2];  // folding constants: (low half, high half)
    Tables() {
        for (uint32_t n = 0; n < 256; n++) {
            uint32_t c = n;
            for (int b = 0; b < 8; b++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            t[0][n] = c;
        }
        for (int s = 1; s < 8; s++)
            for (uint32_t n = 0; n < 256; n++) t[s][n] = (t[s - 1][n] >> 8) ^ t[0][t[s - 1][n] & 255];
        k512[0] = fold_constant(512 + 32), k512[1] = fold_constant(512 - 32);
        k128[0] = fold_constant(128 + 32), k128[1] = fold_constant(128 - 32);
    }
    // x^e mod P, bit-reflected into 32 bits and shifted up one: a 33-bit factor whose product with a
    // reflected 64-bit half lands aligned with the 128-bit lane it is folded into
    static uint64_t fold_constant(int e) {
        uint64_t v = 1;  // normal bit order: bit i is the coefficient of x^i; P = 0x104C11DB7
        for (int i = 0; i < e; i++) {
            v <<= 1;
            if (v >> 32) v ^= 0x104C11DB7ull;
        }
        uint32_t r = 0;
        for (int i = 0; i < 32; i++) r |= (uint32_t)((v >> i) & 1) << (31 - i);
        return (uint64_t)r << 1;
    }
};

inline const Tables& tables() {
    static const Tables T;
    return T;
}

// the register (not complemented) after n bytes
inline uint32_t table_update(uint32_t reg, const uint8_t* p, size_t n) {
    const Tables& T = tables();
    for (; n >= 8; n -= 8, p += 8) {
        uint32_t a, b;
        memcpy(&a, p, 4);
        memcpy(&b, p + 4, 4);
        a ^= reg;
        reg = T.t[7][a & 255] ^ T.t[6][a >> 8 & 255] ^ T.t[5][a >> 16 & 255] ^ T.t[4][a >> 24] ^ T.t[3][b & 255] ^
              T.t[2][b >> 8 & 255] ^ T.t[1][b >> 16 & 255] ^ T.t[0][b >> 24];
    }
    while (n--) reg = T.t[0][(reg ^ *p++) & 255] ^ (reg >> 8);
    return reg;
}

#if defined(__x86_64__)
__attribute__((target("pclmul,sse4.1"))) inline __m128i fold(__m128i x, __m128i k) {
    return _mm_xor_si128(_mm_clmulepi64_si128(x, k, 0x00), _mm_clmulepi64_si128(x, k, 0x11));
}

// n >= 64: the register after n bytes from register reg
__attribute__((target("pclmul,sse4.1"))) inline uint32_t clmul_update(uint32_t reg, const uint8_t* p, size_t n) {
    const Tables& T = tables();
    const __m128i k512 = _mm_set_epi64x((long long)T.k512[1], (long long)T.k512[0]);
    const __m128i k128 = _mm_set_epi64x((long long)T.k128[1], (long long)T.k128[0]);
    auto load = [](const uint8_t* q) { return _mm_loadu_si128((const __m128i*)q); };
    // the preset register folds into the first four bytes
    __m128i x0 = _mm_xor_si128(load(p), _mm_cvtsi32_si128((int)reg));
    __m128i x1 = load(p + 16), x2 = load(p + 32), x3 = load(p + 48);
    p += 64;
    n -= 64;
    for (; n >= 64; n -= 64, p += 64) {
        x0 = _mm_xor_si128(fold(x0, k512), load(p));
        x1 = _mm_xor_si128(fold(x1, k512), load(p + 16));
        x2 = _mm_xor_si128(fold(x2, k512), load(p + 32));
        x3 = _mm_xor_si128(fold(x3, k512), load(p + 48));
    }
    __m128i x = _mm_xor_si128(fold(x0, k128), x1);
    x = _mm_xor_si128(fold(x, k128), x2);
    x = _mm_xor_si128(fold(x, k128), x3);
    for (; n >= 16; n -= 16, p += 16) x = _mm_xor_si128(fold(x, k128), load(p));
    uint8_t last[16];
    _mm_storeu_si128((__m128i*)last, x);
    return table_update(table_update(0, last, 16), p, n);
}

inline bool has_clmul() {
    static const bool yes = __builtin_cpu_supports("pclmul") && __builtin_cpu_supports("sse4.1");
    return yes;
}
#endif

// Combining CRCs (the method of zlib's crc32_combine, crc32.c): with A and B byte strings and
// |B| = n, crc(A B) = x^(8n) crc(A) + crc(B) modulo P, in the reflected bit order where bit 31 is
// the coefficient of x^0. multmodp is a carry-less multiply modulo P; x^(8n) comes from the powers
// x^(2^k) by square and multiply.
inline uint32_t multmodp(uint32_t a, uint32_t b) {
    uint32_t p = 0;
    for (uint32_t m = 1u << 31; m; m >>= 1) {
        if (a & m) p ^= b;
        b = (b & 1) ? (b >> 1) ^ 0xEDB88320u : b >> 1;
    }
    return p;
}
// x^(8n) mod P
inline uint32_t x8nmodp(uint64_t n) {
    static const struct Powers {
        uint32_t x2k[64];  // x^(2^k)
        Powers() {
            uint32_t p = 1u << 30;  // x^1
            for (int k = 0; k < 64; k++) {
                x2k[k] = p;
                p = multmodp(p, p);
            }
        }
    } pw;
    uint32_t p = 1u << 31;  // x^0
    for (int k = 3; n; n >>= 1, k++)
        if (n & 1) p = multmodp(pw.x2k[k], p);
    return p;
}
// the CRC-32 of A B from crc(A), crc(B) and |B|
inline uint32_t combine(uint32_t crc_a, uint32_t crc_b, uint64_t len_b) { return multmodp(x8nmodp(len_b), crc_a) ^ crc_b; }

// crc32(crc, p, n): gzip's CRC-32 of the bytes after those whose CRC-32 is crc (0 to start)
inline uint32_t crc32(uint32_t crc, const void* data, size_t n) {
    const uint8_t* p = (const uint8_t*)data;
    uint32_t reg = ~crc;
#if defined(__x86_64__)
    if (n >= 64 && has_clmul()) return ~clmul_update(reg, p, n);
#endif
    return ~table_update(reg, p, n);
}

}  // namespace hostcrc
