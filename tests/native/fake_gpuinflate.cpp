// A stand-in for the gi_* calls libgzhost makes (gi_inflate, gi_inflate_many, gi_last_crc32,
// gi_last_error), decoding on the CPU with zlib, so the gzip member driver (gpuzip_amd/csrc/gzhost.cpp) can be
// tested without a GPU: built with it into one library by tests/test_gzhost_cpu.py.
// FAKE_GI_FAULT picks a failure to act out:
//   "gpu:N"   the N-th call (from 1) returns GI_ERR_GPU, as a device fault would
//   "flip:N"  the N-th call hands the sink a buffer with one byte changed after its CRC was taken
//             (the GPU's CRC is right, the delivered bytes are not)
//   "crc:N"   the N-th call reports a wrong GPU CRC
//   "many:N"  the N-th gi_inflate_many call returns GI_ERR_GPU
//   "mcrc:K"  gi_inflate_many reports a wrong GPU CRC for its stream K (from 0)
// Test infrastructure only.
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#include <string>
#include <vector>

#include "../../include/gpuinflate.h"

struct gi_ctx {
    std::string err;
    uint32_t crc = 0;
    int calls = 0, many_calls = 0;
};

namespace {
bool fault(const char* kind, int call) {
    const char* f = getenv("FAKE_GI_FAULT");
    if (!f) return false;
    const size_t n = strlen(kind);
    return strncmp(f, kind, n) == 0 && f[n] == ':' && atoi(f + n + 1) == call;
}
}  // namespace

extern "C" {

__attribute__((visibility("default"))) gi_ctx* fake_gi_create() { return new gi_ctx; }
__attribute__((visibility("default"))) void fake_gi_destroy(gi_ctx* g) { delete g; }
// how many gi_inflate and gi_inflate_many calls the context took
__attribute__((visibility("default"))) int fake_gi_calls(const gi_ctx* g, int many) { return many ? g->many_calls : g->calls; }

__attribute__((visibility("default"))) int gi_inflate(gi_ctx* g, const uint8_t* in, uint64_t in_size, uint64_t,
                                                      gi_sink sink, void* user, gi_stats* stats) {
    g->err.clear();
    const int call = ++g->calls;
    if (fault("gpu", call)) {
        g->err = "injected device fault";
        return GI_ERR_GPU;
    }
    z_stream z;
    memset(&z, 0, sizeof z);
    if (inflateInit2(&z, -15) != Z_OK) return GI_ERR_GPU;
    std::vector<uint8_t> out(1 << 16);
    z.next_in = (Bytef*)in;
    z.avail_in = (uInt)in_size;
    uint32_t crc = 0;
    bool flipped = false;
    int rc;
    for (;;) {
        z.next_out = out.data();
        z.avail_out = (uInt)out.size();
        rc = inflate(&z, Z_NO_FLUSH);
        const size_t n = out.size() - z.avail_out;
        if (n) {
            crc = (uint32_t)crc32(crc, out.data(), (uInt)n);
            if (fault("flip", call) && !flipped) {
                out[n / 2] ^= 0x40;
                flipped = true;
            }
            if (sink(user, out.data(), n)) {
                inflateEnd(&z);
                g->err = "stopped by the caller";
                return GI_ERR_ABORTED;
            }
        }
        if (rc == Z_STREAM_END) break;
        if (rc != Z_OK || (n == 0 && z.avail_in == 0)) {
            inflateEnd(&z);
            g->err = rc == Z_OK || rc == Z_BUF_ERROR ? "the input ends inside the stream" : "invalid Deflate data";
            return rc == Z_OK || rc == Z_BUF_ERROR ? GI_ERR_TRUNCATED : GI_ERR_DATA;
        }
    }
    g->crc = fault("crc", call) ? ~crc : crc;
    if (stats) {
        memset(stats, 0, sizeof *stats);
        stats->in_used = in_size - z.avail_in;
        stats->out_size = z.total_out;
        stats->segments = 1;
    }
    inflateEnd(&z);
    return GI_OK;
}

__attribute__((visibility("default"))) int gi_inflate_many(gi_ctx* g, const uint8_t* in, gi_stream* st, uint64_t n,
                                                           uint8_t* out, uint64_t out_cap, gi_stats* stats) {
    g->err.clear();
    const int call = ++g->many_calls;
    if (stats) memset(stats, 0, sizeof *stats);
    if (fault("many", call)) {
        g->err = "injected device fault";
        return GI_ERR_GPU;
    }
    int first = GI_OK;
    for (uint64_t i = 0; i < n; i++) {
        gi_stream& s = st[i];
        if (s.out_offset + s.out_room > out_cap) return GI_ERR_ARG;
        z_stream z;
        memset(&z, 0, sizeof z);
        inflateInit2(&z, -15);
        std::vector<uint8_t> buf(s.out_room + 1);
        z.next_in = (Bytef*)(in + s.offset);
        z.avail_in = (uInt)s.size;
        z.next_out = buf.data();
        z.avail_out = (uInt)buf.size();
        const int rc = inflate(&z, Z_FINISH);
        s.status = rc == Z_STREAM_END ? (z.total_out <= s.out_room ? GI_OK : GI_ERR_ARG)
                   : rc == Z_BUF_ERROR && z.avail_out ? GI_ERR_TRUNCATED : GI_ERR_DATA;
        s.out_size = s.status == GI_OK ? z.total_out : 0;
        s.in_used = s.status == GI_OK ? s.size - z.avail_in : 0;
        s.crc32 = s.status == GI_OK ? (uint32_t)crc32(0, buf.data(), (uInt)z.total_out) : 0;
        if (fault("mcrc", (int)i)) s.crc32 = ~s.crc32;
        if (s.status == GI_OK) memcpy(out + s.out_offset, buf.data(), s.out_size);
        if (first == GI_OK) first = s.status;
        inflateEnd(&z);
        if (stats) stats->out_size += s.out_size;
    }
    if (stats) stats->segments = 1;
    return first;
}

__attribute__((visibility("default"))) uint32_t gi_last_crc32(const gi_ctx* g) { return g->crc; }
__attribute__((visibility("default"))) const char* gi_last_error(const gi_ctx* g) { return g->err.c_str(); }

}  // extern "C"
