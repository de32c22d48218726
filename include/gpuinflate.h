/* gpuinflate - decompress one raw Deflate stream (RFC 1951) in parallel on an AMD Instinct
 * MI355X (gfx950) GPU.
 *
 * The C interface is the reference library's, version 1, unchanged: gpuinflate/gpuinflate.h:31-84
 * (reference), so a host built against the reference loads this libgpuinflate.so in its place.
 * The declarations below the line "Additions" are new symbols; GI_ABI_VERSION stays 1 because
 * the reference host checks it for equality (src/inflate.rs:124).
 *
 * The stream is decoded speculatively by many thousands of GPU lanes at once, in segments of
 * output; the output reaches the caller in order. Corrupt or malicious input makes the calls
 * return an error; they never read or write outside their buffers. A context is not
 * thread-safe: use one per thread. */
#ifndef GPUINFLATE_H
#define GPUINFLATE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(GI_BUILD) && defined(__GNUC__)
#define GI_API __attribute__((visibility("default"))) /* the only symbols exported with -fvisibility=hidden */
#else
#define GI_API
#endif

#define GI_ABI_VERSION 1

/* gpuinflate.h:33-40 */
enum {
    GI_OK = 0,
    GI_ERR_DATA = 1,      /* the stream is not valid Deflate */
    GI_ERR_TRUNCATED = 2, /* the input ends before the stream's last block does */
    GI_ERR_ABORTED = 3,   /* the sink returned nonzero */
    GI_ERR_GPU = 4,       /* no usable GPU, out of GPU memory, or a HIP error: decode on the CPU */
    GI_ERR_ARG = 5        /* a bad argument */
};

typedef struct gi_ctx gi_ctx;

/* gpuinflate.h:44-49; 16 bytes */
typedef struct {
    uint32_t size;           /* sizeof(gi_config) */
    int32_t device;          /* HIP device ordinal */
    uint32_t segment_out_mb; /* most output per segment, MiB (0: from the GPU's free memory, at most 4096) */
    uint32_t segment_in_kb;  /* compressed input per segment, KiB (0: from the stream's ratio) */
} gi_config;

/* gpuinflate.h:51-59; 80 bytes */
typedef struct {
    uint64_t in_used;  /* bytes of input the stream took, through its last block (byte-aligned) */
    uint64_t out_size; /* bytes of output */
    uint32_t segments;
    uint32_t rounds;   /* decode rounds, all segments */
    uint64_t jobs;     /* decoders started, all segments */
    double seconds;    /* the whole call */
    double upload_s, find_s, decode_s, resolve_s, download_s; /* GPU time by stage */
} gi_stats;

/* gpuinflate.h:61-63: receives the output in order, from a thread of the library's own, never
 * two calls at once. Return nonzero to stop. */
typedef int (*gi_sink)(void* user, const uint8_t* data, size_t len);

/* gpuinflate.h:65 */
GI_API uint32_t gi_abi_version(void);

/* gpuinflate.h:67-69: allocates the GPU and pinned host buffers. On failure returns an error
 * code and describes it in err (when given). */
GI_API int gi_create(const gi_config* cfg, gi_ctx** ctx, char* err, size_t err_len);

/* gpuinflate.h:71-76: decompresses the raw Deflate stream at the start of in[0..in_size). Bytes
 * after the stream's last block are allowed and not read past what the stream needs;
 * stats->in_used says where the stream ended. out_hint: the expected output size, or 0 (it
 * only tunes the first segment). stats may be NULL. */
GI_API int gi_inflate(gi_ctx* ctx, const uint8_t* in, uint64_t in_size, uint64_t out_hint, gi_sink sink, void* user,
                      gi_stats* stats);

/* gpuinflate.h:78-79: what went wrong in the last call that failed */
GI_API const char* gi_last_error(const gi_ctx* ctx);

/* gpuinflate.h:81-82: the GPU's name, for reports */
GI_API const char* gi_device_name(const gi_ctx* ctx);

/* gpuinflate.h:84 */
GI_API void gi_destroy(gi_ctx* ctx);

/* ---------------------------------------------------------------- Additions (no counterpart
 * in the reference's interface; the reference host never calls them) */

/* A gi_config with the decoder's schedule after it, for gi_create and gi_batch_create: set
 * base.size = sizeof(gi_config_tuned). The reference's size field versions gi_config
 * (gpuinflate.cu:888-892): a config of 16 bytes keeps every default. 0 in a field keeps that
 * field's default; a value outside its range is GI_ERR_ARG. Results do not depend on the
 * schedule, only the speed does; gpuzip_amd/autotune.py chooses one by measurement. */
typedef struct {
    gi_config base;
    uint32_t chunk_kib;          /* search stride: one job per this many KiB of input at most (4..64; 4) */
    uint32_t probe_spacing_kbit; /* stretches longer than this many Kibit get probes (64..1024; 1024) */
    uint32_t probe_lead_bits;    /* a probe decodes this many bits before its point (1024..8192; 4096) */
    uint32_t segment_in_max_mb;  /* compressed bytes per segment at most (64..1024; 1024) */
    uint32_t lanes_per_group;    /* lanes that decode one interval together: 16, 32 or 64, a quarter,
                                    half or whole wavefront (64) */
    uint32_t persistent_per_cu;  /* the group decoder as one persistent launch of this many workgroups
                                    a compute unit, taking work from a device queue (1..16; 0: a
                                    workgroup per queue item). A config of 32 bytes keeps both
                                    defaults. */
} gi_config_tuned;

/* The CRC-32 (gzip's) of the output of the last gi_inflate or gi_inflate_device call that
 * returned GI_OK, computed on the GPU as the output was made. */
GI_API uint32_t gi_last_crc32(const gi_ctx* ctx);

/* Decodes a raw Deflate stream that is already in device memory into device memory: no host
 * copies. d_in: in_size bytes, readable up to the next multiple of 4; d_out: out_cap bytes
 * (GI_ERR_ARG when the output is larger). *out_size and *crc32 (both may be NULL) receive the
 * output's size and CRC-32. Work is ordered on the context's own stream and complete on return. */
GI_API int gi_inflate_device(gi_ctx* ctx, const void* d_in, uint64_t in_size, void* d_out, uint64_t out_cap,
                             uint64_t out_hint, uint64_t* out_size, uint32_t* crc32, gi_stats* stats);

/* GPU time per kernel of the last gi_inflate or gi_inflate_device call, from HIP events
 * recorded around every launch on the context's stream: up to max entries, one per kernel that
 * ran; returns how many. Names are static strings. The block finder's four kernels run as one
 * HIP graph a segment and are timed together as "find_graph" (with GZ_NO_GRAPH set when the
 * context is created, they are launched and timed one by one). */
typedef struct {
    const char* name;
    uint64_t launches;
    double seconds;
} gi_kernel_time;
GI_API int gi_kernel_times(const gi_ctx* ctx, gi_kernel_time* out, int max);

/* The rescue's header search (no counterpart in gpuinflate.h: it replaces Rescuer::find_headers,
 * gpuinflate/rescue.cpp:597-626, the first step of gr_rescue): every dynamic-block header in
 * [from, 8 * size - 128) of the host buffer in[0..size) that zlib would accept and from which a
 * decode goes on cleanly for 256 Ki symbols, or to the stream's end, or across two block starts
 * (verify_at, rescue.cpp:549-554); in bit order, the first `cap` of them in out, their number in
 * *count. from >= 1 (the reference passes the gzip header's end, or 1 for a raw stream); size at
 * most one segment's search (1 GiB by default). */
GI_API int gi_rescue_headers(gi_ctx* ctx, const uint8_t* in, uint64_t size, uint64_t from, uint64_t* out, uint64_t cap,
                             uint64_t* count);

/* gi_rescue_headers without allocating: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The same search as gi_rescue_headers, over d_in[0..size) already in device memory on
 * `device`, with every device buffer it needs taken from one static array of the library's
 * (96 MiB at most, held by each device the call runs on until the process ends): no hipMalloc,
 * hipFree or other allocating call. size is at most 16 MiB (GI_ERR_ARG above), and an input with
 * more than 524,288 header candidates or 16,384 plausible headers is GI_ERR_ARG too; err (may be
 * NULL) receives the reason. The call runs on the null stream, is complete on return, and holds a
 * process-wide lock: calls from one process never overlap; another process's calls on the same
 * GPU have their own array. Headers go to out and *count as for gi_rescue_headers. */
GI_API int gi_rescue_headers_noalloc(int device, const void* d_in, uint64_t size, uint64_t from, uint64_t* out,
                                     uint64_t cap, uint64_t* count, char* err, size_t err_len);

/* Batches (no counterpart in the reference: its batched path is nvCOMP's, src/gpu.rs:297-408).
 * Many streams already in device memory, decoded by worker threads, each with a decoder context
 * of its own (its own HIP streams and buffers) on the device of cfg. A worker takes every queued
 * job that fits one batch segment (their inputs packed together, at most 16384 jobs) and decodes
 * them with one set of launches; a job too large to share a segment is decoded alone. Each job's
 * results are those of decoding it alone. A job's results are set when it is done; its buffers
 * must stay valid until then. */
typedef struct gi_batch gi_batch;
typedef struct {
    const void* d_in;  /* the raw Deflate stream, in device memory, readable to the next multiple of 4 bytes */
    uint64_t in_size;
    void* d_out;       /* room for its output, in device memory */
    uint64_t out_cap;
    uint64_t out_hint; /* its output size if known (0: unknown) */
    void* user;        /* the caller's */
    /* results */
    uint64_t out_size;
    uint64_t in_used;
    uint32_t crc32;    /* gzip's CRC-32 of the output, computed on the GPU */
    int32_t status;    /* GI_OK or gi_inflate's error codes */
    double queued_s;   /* from submission to the start of its decode */
    double seconds;    /* its decode */
} gi_job;
/* workers: 1..4 decoder contexts, each made with cfg (its segment_out_mb bounds a batch
 * segment's output; 0 sizes it to the free memory, at most 4 GiB). Two keep the GPU busy while a
 * worker plans on the host. */
GI_API int gi_batch_create(const gi_config* cfg, int workers, gi_batch** out, char* err, size_t err_len);
/* Decodes jobs[0..n), largest first; returns when all are done: GI_OK, or the first failed job's status */
GI_API int gi_batch_run(gi_batch* b, gi_job* jobs, uint64_t n);
/* Queues one job behind those already queued (arrival order), for a serving loop */
GI_API int gi_batch_submit(gi_batch* b, gi_job* job);
/* The next job submitted with gi_batch_submit that is done, or NULL after timeout_s seconds */
GI_API gi_job* gi_batch_wait(gi_batch* b, double timeout_s);
GI_API int gi_batch_workers(const gi_batch* b);
/* The batch's GPU work since the last reset (or its creation): per kernel (names as
 * gi_kernel_times'), launches and HIP-event time summed over the workers, up to max entries,
 * returning how many; *segments: batch segments decoded (each a find graph, one group-decoder
 * launch per decode round and a resolve graph, whatever its number of streams); *alone: streams
 * decoded by themselves. reset: the counts start again from zero. */
GI_API int gi_batch_kernel_times(gi_batch* b, gi_kernel_time* out, int max, uint64_t* segments, uint64_t* alone,
                                 int reset);
GI_API void gi_batch_destroy(gi_batch* b);

/* Serving (no counterpart in the reference): decode requests that arrive while others are on the
 * GPU. A request is a gzip member or a raw Deflate stream in host memory and a host buffer for
 * its output. The server admits requests in arrival order when their device memory (the
 * compressed input and the output room; the decoders' marker pools belong to the workers'
 * contexts) fits beside what is already admitted within its HBM budget, and keeps them there
 * while they decode. The batch workers fold admitted requests into their next batch segment
 * with whatever else is on the device, so requests that arrive while others decode overlap them.
 * Each request's results are those of decoding it alone. */
typedef struct gi_server gi_server;
typedef struct {
    const uint8_t* in;   /* a gzip member (gzip = 1: header read, trailer checked) or raw Deflate */
    uint64_t in_size;
    uint8_t* out;        /* host memory for the output */
    uint64_t out_cap;    /* its size; a gzip member's output room is its trailer's ISIZE if smaller */
    int32_t gzip;
    int32_t pad;
    void* user;          /* the caller's */
    /* results */
    uint64_t out_size;
    uint32_t crc32;      /* gzip's CRC-32 of the output, computed on the GPU */
    int32_t status;      /* GI_OK or gi_inflate's codes; GI_ERR_DATA for a bad gzip header or
                            trailer; GI_ERR_ARG for a request larger than the whole budget */
    double admitted_s;   /* from submission to admission */
    double done_s;       /* from submission to the output in host memory */
    char error[96];      /* what went wrong, when status is not GI_OK */
} gi_request;
/* workers: as gi_batch_create; hbm_budget: device bytes for admitted requests (one allocation,
 * made here) */
GI_API int gi_serve_create(const gi_config* cfg, int workers, uint64_t hbm_budget, gi_server** out, char* err,
                           size_t err_len);
/* Queues a request behind those not yet admitted; its buffers must stay valid until it is done */
GI_API int gi_serve_submit(gi_server* s, gi_request* r);
/* The next request that is done (completion order), or NULL after timeout_s seconds */
GI_API gi_request* gi_serve_wait(gi_server* s, double timeout_s);
/* Device bytes held by admitted requests now, and the most ever held (may be NULL) */
GI_API uint64_t gi_serve_in_use(const gi_server* s, uint64_t* peak);
/* Waits for every admitted request to finish; requests never admitted are dropped */
GI_API void gi_serve_destroy(gi_server* s);

/* Many raw Deflate streams in one host buffer, decoded on one context as batch segments (every
 * stream that fits a segment's share decoded with one set of launches; a larger one alone), the
 * outputs written to host memory: the many members of a BGZF file, for one. No counterpart in the
 * reference (its host sends each member to the GPU by itself, src/gz.rs:215). Each stream's
 * results are those of gi_inflate on it alone. */
typedef struct {
    uint64_t offset;     /* the stream: in[offset, offset + size) */
    uint64_t size;
    uint64_t out_offset; /* its output: out[out_offset, out_offset + out_room) */
    uint64_t out_room;   /* an output larger than this is GI_ERR_ARG for the stream */
    /* results */
    uint64_t out_size;
    uint64_t in_used;
    uint32_t crc32;      /* gzip's CRC-32 of the output, computed on the GPU */
    int32_t status;      /* GI_OK or gi_inflate's error codes */
} gi_stream;
/* Returns GI_OK when every stream decoded, else the status of the first one (in the array's
 * order) that did not; GI_ERR_ARG before any decode when a stream's output lies outside
 * out[0, out_cap). stats (may be NULL) sums the streams. */
GI_API int gi_inflate_many(gi_ctx* ctx, const uint8_t* in, gi_stream* streams, uint64_t n, uint8_t* out, uint64_t out_cap,
                           gi_stats* stats);

#ifdef __cplusplus
}
#endif

#endif
