/* inflate_oracle.h - CPU restatement of the reference Deflate decoder's semantics.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP decoder in gpuzip_amd/:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and there
 * only to check or to time a CPU baseline. The product path never calls it.
 *
 * It restates, in plain C and in its own structure, what gpuinflate/inflate.cuh and
 * gpuinflate/gi_test.cpp (reference) compute:
 *   - the LSB-first bit reader whose reads past the input return zeros (inflate.cuh:108-150);
 *   - canonical Huffman codes with zlib's acceptance rules (inflate.cuh:154-281, 393-411);
 *   - Decoder::run: speculative decode with landing on targets, stop bit, output cap and the
 *     exact resume state (inflate.cuh:418-581);
 *   - the block-start predicates header_prefilter/header_prefilter128/plausible_dynamic_header
 *     (inflate.cuh:285-332);
 *   - whole-stream decode with back-references resolved afterwards (gi_test.cpp:47-91);
 *   - CRC-32 as gzip defines it (reflected 0xEDB88320, init and final xor 0xFFFFFFFF).
 */
#ifndef INFLATE_ORACLE_H
#define INFLATE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_WINDOW 32768
#define OR_REF 255 /* 16-bit symbols above this are back-references: the byte (v - 255) back */
#define OR_MAX_SYMBOL_OUT 258
#define OR_MAX_BLOCKS 65536

enum { OR_OK = 0, OR_ERR_DATA = 1, OR_ERR_OVERFLOW = 2, OR_ERR_INPUT = 3, OR_STOPPED = 4 };
enum { OR_AT_BLOCK = 0, OR_IN_FIXED = 1, OR_IN_DYNAMIC = 2 };
/* whole-stream result codes, the gpuinflate.h values */
enum { OR_GI_OK = 0, OR_GI_DATA = 1, OR_GI_TRUNCATED = 2 };

typedef struct {
    uint64_t bit;
    uint64_t header; /* IN_DYNAMIC: the block's header; IN_FIXED: header or UINT64_MAX */
    uint8_t mode;
    int8_t final;    /* BFINAL of the block it starts in; -1 unknown */
} or_start;

typedef struct {
    uint64_t end_bit;
    uint64_t eob_bit;        /* after the first end-of-block code; UINT64_MAX none */
    or_start resume;         /* OVERFLOW / STOPPED: where to continue */
    uint32_t eob_out;
    uint32_t out_len;
    int32_t status;
    int32_t landed;          /* target index, -1: none */
    int32_t end_final;
    uint32_t blocks;
    uint64_t last_block_bit;
    uint32_t last_block_out;
    uint32_t symbols;
} or_result;

/* CRC-32 (gzip). crc is the running value (0 to begin). */
uint32_t or_crc32(uint32_t crc, const uint8_t* p, size_t n);

int or_header_prefilter(uint64_t v);
int or_header_prefilter128(uint64_t v1, uint64_t v2);
int or_plausible_dynamic_header(const uint8_t* in, uint64_t size, uint64_t bit);
/* 64 bits of the stream starting at `bit` (zeros past the input) */
uint64_t or_bits64(const uint8_t* in, uint64_t size, uint64_t bit);

/* Decoder::run. out receives cap 16-bit symbols at most. */
or_result or_run(const uint8_t* in, uint64_t size, or_start s, const or_start* targets, int ntargets, uint16_t* out,
                 uint32_t cap, int stream_start, uint64_t stop);

/* Whole stream from bit 0 (gi_test reference()): returns OR_GI_*; *too_big set when the
 * output would exceed max_out. out must hold max_out bytes. */
int or_inflate(const uint8_t* in, uint64_t size, uint8_t* out, uint64_t max_out, uint64_t* out_len, uint64_t* in_used,
               int* too_big);

/* Every bit b in [lo, hi) where plausible_dynamic_header holds, ascending; returns the count
 * (writes at most cap). Brute force: used to pin the GPU block finder. */
uint64_t or_find_headers(const uint8_t* in, uint64_t size, uint64_t lo, uint64_t hi, uint64_t* out, uint64_t cap);

#ifdef __cplusplus
}
#endif

#endif
