/* gzhost - the gzip member driver above gpuinflate: every member of a .gz file decoded on the
 * GPU through gi_inflate, its CRC-32 and size checked against the member's trailer.
 *
 * It has the role of the reference host's member loop, gz::decode (src/gz.rs:154-264, reference)
 * and its header parser parse_header (src/gz.rs:62-110): the same failure classes and messages.
 *   - Every member goes to the GPU. The reference sends members under 8 MiB (auto) or 1 MiB
 *     (--mode gpu) to zlib-rs on the CPU (src/gz.rs:15-19, 177-181); that choice stays with the
 *     caller, which can look at gz_header's member size first.
 *   - Two or more BGZF members in a row (each one's size in its header's BC subfield, as bgzip
 *     writes them) are decoded together by gi_inflate_many, a few sets of launches for thousands
 *     of members; a member whose batched result does not match its trailer is decoded again by
 *     itself, which reports it as the one-member path does.
 *   - The trailer's CRC-32 must equal both the CRC the GPU computed as it wrote the output
 *     (gi_last_crc32) and a host CRC of the bytes as the sink received them (as src/gz.rs:205-213
 *     checks the sink's bytes), so a fault on the way to the sink is a GZ_RETRY as well.
 * Error classes follow gz::Fail: a damaged file is GZ_CORRUPT; a GPU failure, or GPU output that
 * fails the trailer check, is GZ_RETRY (the caller decodes on the CPU); a sink that returns
 * nonzero is GZ_STOPPED. */
#ifndef GZHOST_H
#define GZHOST_H

#include "gpuinflate.h"

#ifdef __cplusplus
extern "C" {
#endif

#if defined(GZ_BUILD) && defined(__GNUC__)
#define GZ_API __attribute__((visibility("default")))
#else
#define GZ_API
#endif

enum { GZ_OK = 0, GZ_CORRUPT = 1, GZ_RETRY = 2, GZ_STOPPED = 3 };

typedef struct {
    uint32_t members;   /* members decoded */
    uint32_t pad;
    uint64_t packed;    /* bytes of the file */
    uint64_t out;       /* bytes decoded */
    gi_stats stats;     /* gi_inflate's, summed over the members (seconds: the calls' sum) */
    char info[256];     /* zero bytes after the last member, if any ("" none) */
    char note[256];     /* other bytes after the last member, ignored ("" none) */
    char error[256];    /* why the call failed ("" if it did not) */
} gz_report;

/* A member header (RFC 1952) at d[0..n): 0 and its length (*len) and, for a BGZF member (an
 * FEXTRA "BC" subfield), the whole member's size (*member; else 0); or GZ_CORRUPT with a message
 * in err (src/gz.rs:62-110). No GPU needed. */
GZ_API int gz_header(const uint8_t* d, uint64_t n, uint64_t* len, uint64_t* member, char* err, size_t err_len);

/* Decodes every member of the gzip file data[0..size) on ctx's GPU, passing the output to sink
 * in order, never two calls at once: from gi_inflate's thread for a member decoded by itself,
 * from the calling thread for a run of BGZF members. rep may be NULL. */
GZ_API int gz_decode(gi_ctx* ctx, const uint8_t* data, uint64_t size, gi_sink sink, void* user, gz_report* rep);

#ifdef __cplusplus
}
#endif

#endif
