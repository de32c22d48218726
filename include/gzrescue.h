/* gzrescue.h - the rescue interface (ABI version 4) that the reference's host loads as
 * libgzrescue.so (src/rescue.rs:154-175; reference gpuinflate/rescue.h:38-107): the same names,
 * struct layouts, codes and mask values, so the reference's Rust binding loads this library in
 * place of its CPU one. What it does: the bytes a damaged gzip file (or raw Deflate stream) still
 * holds, in order, with a mask byte per output byte saying whether it is exact, suspect (inside a
 * damaged block), inferred (copied from data lost with a damaged block, its value inferred) or
 * unknown (written as the fill byte). Here the header search, the decode of every piece and the
 * provenance of every output byte run on the GPU; the decisions run on the host. */
#ifndef GZRESCUE_H
#define GZRESCUE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(GR_BUILD) && defined(__GNUC__)
#define GR_API __attribute__((visibility("default")))
#else
#define GR_API
#endif

/* rescue.h:38 */
#define GR_ABI_VERSION 4

/* rescue.h:40-43: return codes; mask values */
enum { GR_OK = 0, GR_ERR_NO_DEFLATE = 1, GR_ERR_ABORTED = 2, GR_ERR_ARG = 3, GR_ERR_MEMORY = 4, GR_ERR_TOO_BIG = 5, GR_ERR_INTERNAL = 6 };
enum { GR_UNKNOWN = 0, GR_EXACT = 1, GR_SUSPECT = 2, GR_INFERRED = 3 };

/* rescue.h:45-56: versioned by size (fields beyond it take their defaults) */
typedef struct {
    uint32_t size;          /* sizeof(gr_options) */
    uint8_t fill;           /* the byte written where one is unknown (size >= 5) */
    uint8_t no_infer;       /* 1: leave bytes copied from lost data unknown (size >= 6) */
    uint8_t cr_lost;        /* CR LF to LF transfer: 0 judge from the data, 1 it happened, 2 never (size >= 7) */
    uint8_t crs_expected;   /* raw stream with cr_lost 1: CRs it lost, 1-8; 255 none; 0 unknown (size >= 8) */
    uint64_t max_output;    /* GR_ERR_TOO_BIG past this many output bytes; 0 no limit (size >= 16) */
    uint64_t cr_budget;     /* input bytes the search for lost CRs may decode; 0: 1 GiB (size >= 24) */
} gr_options;

/* rescue.h:58-66 */
typedef struct {
    uint64_t out_offset, length;  /* in the output */
    uint64_t decoded, known;      /* bytes the decode produced; of them, not unknown */
    uint64_t in_byte_start, in_byte_end;
    int32_t suspect;              /* 1: damaged */
    int32_t status;               /* how its decode ended */
    uint32_t blocks, symbols;
    const char* why;              /* suspect: how the damage showed; else "" */
} gr_piece;

/* rescue.h:68-85 */
typedef struct {
    uint64_t input_size, output_size;
    uint64_t exact_bytes, suspect_bytes, unknown_bytes;
    uint32_t headers, pieces, damaged, members;
    int32_t crc_check;      /* 1 ok, 0 mismatch, -1 not checkable */
    int64_t length_off;     /* output length minus the trailer's, every block in step; else 0 */
    int32_t truncated;
    int32_t header_damaged;
    uint64_t input_end;     /* byte after the first member's last decoded block */
    uint64_t inferred_bytes;
    int32_t ascii_undone;   /* an LF to CR LF transfer was undone */
    uint32_t cr_restored;   /* CRs put back */
    uint64_t cr_decoded;    /* input bytes the search for them decoded */
    int32_t hidden_damage;  /* damage of a kind that can stay unseen; exact bytes are not vouched for */
} gr_stats;

/* rescue.h:88: the output in order, with its mask; nonzero stops the rescue */
typedef int (*gr_sink)(void* user, const uint8_t* data, const uint8_t* mask, size_t len);

typedef struct gr_result gr_result;

/* rescue.h:92 */
GR_API uint32_t gr_abi_version(void);
/* rescue.h:96: a gzip file; *result holds statistics, pieces and notes (free with gr_free) */
GR_API int gr_rescue(const uint8_t* in, uint64_t size, const gr_options* opt, gr_sink sink, void* user, gr_result** result);
/* rescue.h:102-103: a raw Deflate stream at in[0] (in may run on past it); expected_size
 * (UINT64_MAX unknown) and expected_crc (when has_crc) stand in for a gzip trailer */
GR_API int gr_rescue_raw(const uint8_t* in, uint64_t size, uint64_t expected_size, int has_crc, uint32_t expected_crc,
                         const gr_options* opt, gr_sink sink, void* user, gr_result** result);
/* rescue.h:104-107 */
GR_API const gr_stats* gr_result_stats(const gr_result* r);
GR_API uint32_t gr_result_pieces(const gr_result* r, const gr_piece** pieces);
GR_API uint32_t gr_result_notes(const gr_result* r, const char* const** notes);
GR_API void gr_free(gr_result* r);

/* ---------------------------------------------------------------- Additions (no counterpart
 * in the reference's interface; the reference host never calls them) */

/* The HIP device the rescues after this call run on (0 unless set); GR_ERR_ARG if negative */
GR_API int gzr_set_device(int device);
/* The wall seconds each step of the rescue took, in order, up to cap of them: names[i] (valid
 * until gr_free) and seconds[i] (either may be NULL); returns how many steps there are */
GR_API uint32_t gzr_result_phases(const gr_result* r, const char** names, double* seconds, uint32_t cap);
/* GPU time per kernel of the rescue, from HIP events recorded around every launch on its
 * stream: up to max entries (names are static strings); returns how many kernels ran.
 * *stored_bytes (may be NULL): the output bytes copied from stored blocks by their own kernel */
typedef struct {
    const char* name;
    uint64_t launches;
    double seconds;
} gzr_kernel_time;
GR_API int gzr_result_kernels(const gr_result* r, gzr_kernel_time* out, int max, uint64_t* stored_bytes);
/* Why this thread's last rescue returned GR_ERR_INTERNAL or GR_ERR_MEMORY (a GPU failure among
 * them); "" after one that did not */
GR_API const char* gzr_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
