// A stand-in for the four gi_* calls libgzhost makes (gi_inflate, gi_last_crc32, gi_last_error),
// decoding on the CPU with zlib, so the gzip member driver (gpuzip_amd/csrc/gzhost.cpp) can be
// tested without a GPU: built with it into one library by tests/test_gzhost_cpu.py.
// FAKE_GI_FAULT picks a failure to act out:
//   "gpu:N"   the N-th call (from 1) returns GI_ERR_GPU, as a device fault would
//   "flip:N"  the N-th call hands the sink a buffer with one byte changed after its CRC was taken
//             (the GPU's CRC is right, the delivered bytes are not)
//   "crc:N"   the N-th call reports a wrong GPU CRC
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
    int calls = 0;
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

__attribute__((visibility("default"))) uint32_t gi_last_crc32(const gi_ctx* g) { return g->crc; }
__attribute__((visibility("default"))) const char* gi_last_error(const gi_ctx* g) { return g->err.c_str(); }

}  // extern "C"
