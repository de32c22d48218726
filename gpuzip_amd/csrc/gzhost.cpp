// gzhost.cpp - the gzip member driver (include/gzhost.h): the reference host's gz::decode and
// parse_header (src/gz.rs:62-110, 154-264) over this library's gi_inflate, for callers that
// hold whole .gz files.
#define GZ_BUILD
#include "gzhost.h"

#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>

namespace {

// CRC-32 (gzip's), for the header's FHCRC only: the members' output CRC comes from the GPU
uint32_t crc32_bytes(const uint8_t* p, size_t n) {
    static uint32_t tab[256];
    static bool ready = false;
    if (!ready) {
        for (uint32_t k = 0; k < 256; k++) {
            uint32_t c = k;
            for (int b = 0; b < 8; b++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            tab[k] = c;
        }
        ready = true;
    }
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; i++) c = tab[(c ^ p[i]) & 255] ^ (c >> 8);
    return ~c;
}

void put(char* dst, size_t cap, const std::string& s) {
    if (!dst || !cap) return;
    size_t n = s.size() < cap - 1 ? s.size() : cap - 1;
    memcpy(dst, s.data(), n);
    dst[n] = 0;
}

uint32_t le32(const uint8_t* p) { return p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }

bool is_gzip(const uint8_t* d, uint64_t n) { return n >= 3 && d[0] == 0x1f && d[1] == 0x8b && d[2] == 8; }

// src/gz.rs:62-110
bool parse_header(const uint8_t* d, uint64_t n, uint64_t& len, uint64_t& member, std::string& err) {
    member = 0;
    if (n < 10 || d[0] != 0x1f || d[1] != 0x8b) {
        err = "not a gzip member";
        return false;
    }
    if (d[2] != 8) {
        err = "compression method " + std::to_string(d[2]) + " is not Deflate";
        return false;
    }
    const uint8_t flg = d[3];
    if (flg & 0xe0) {
        err = "reserved header flags are set";
        return false;
    }
    const std::string shortmsg = "the file ends inside a member header";
    uint64_t p = 10;
    if (flg & 4) {
        // FEXTRA: subfields of 2 id bytes, a 2-byte length, the data; BGZF's "BC" holds the
        // member's size less one
        if (p + 2 > n) {
            err = shortmsg;
            return false;
        }
        const uint64_t xn = d[p] | (uint64_t)d[p + 1] << 8;
        if (p + 2 + xn > n) {
            err = shortmsg;
            return false;
        }
        const uint8_t* extra = d + p + 2;
        uint64_t x = 0;
        while (x + 4 <= xn) {
            const uint64_t sl = extra[x + 2] | (uint64_t)extra[x + 3] << 8;
            if (extra[x] == 'B' && extra[x + 1] == 'C' && sl == 2 && x + 6 <= xn)
                member = (extra[x + 4] | (uint64_t)extra[x + 5] << 8) + 1;
            x += 4 + sl;
        }
        p += 2 + xn;
    }
    for (uint8_t flag : {(uint8_t)8, (uint8_t)16}) {
        // FNAME, FCOMMENT: zero-terminated
        if (flg & flag) {
            const void* z = p < n ? memchr(d + p, 0, n - p) : nullptr;
            if (!z) {
                err = shortmsg;
                return false;
            }
            p = (uint64_t)((const uint8_t*)z - d) + 1;
        }
    }
    if (flg & 2) {
        // FHCRC: the low 16 bits of the header's CRC-32
        if (p + 2 > n) {
            err = shortmsg;
            return false;
        }
        if ((crc32_bytes(d, p) & 0xFFFF) != (uint32_t)(d[p] | d[p + 1] << 8)) {
            err = "header CRC mismatch";
            return false;
        }
        p += 2;
    }
    if (p > n) {
        err = shortmsg;
        return false;
    }
    len = p;
    return true;
}

// the caller's sink, and the member's output size (its CRC comes from the GPU)
struct Counted {
    gi_sink sink;
    void* user;
    uint64_t size;
};

int counted_sink(void* u, const uint8_t* data, size_t len) {
    Counted* c = (Counted*)u;
    c->size += len;
    return c->sink ? c->sink(c->user, data, len) : 0;
}

void add_stats(gi_stats& a, const gi_stats& b) {
    a.in_used += b.in_used;
    a.out_size += b.out_size;
    a.segments += b.segments;
    a.rounds += b.rounds;
    a.jobs += b.jobs;
    a.seconds += b.seconds;
    a.upload_s += b.upload_s;
    a.find_s += b.find_s;
    a.decode_s += b.decode_s;
    a.resolve_s += b.resolve_s;
    a.download_s += b.download_s;
}

}  // namespace

extern "C" {

GZ_API int gz_header(const uint8_t* d, uint64_t n, uint64_t* len, uint64_t* member, char* err, size_t err_len) {
    uint64_t l = 0, m = 0;
    std::string e;
    if (!d || !parse_header(d, n, l, m, e)) {
        put(err, err_len, d ? e : "no data");
        return GZ_CORRUPT;
    }
    if (len) *len = l;
    if (member) *member = m;
    put(err, err_len, "");
    return GZ_OK;
}

// src/gz.rs:154-264, every member on the GPU
GZ_API int gz_decode(gi_ctx* ctx, const uint8_t* data, uint64_t size, gi_sink sink, void* user, gz_report* rep) {
    gz_report local;
    gz_report& r = rep ? *rep : local;
    memset(&r, 0, sizeof r);
    auto fail = [&](int code, const std::string& msg) {
        put(r.error, sizeof r.error, msg);
        return code;
    };
    if (!ctx || (!data && size)) return fail(GZ_CORRUPT, "no data");
    uint64_t pos = 0;
    while (pos < size) {
        const uint8_t* rest = data + pos;
        const uint64_t nrest = size - pos;
        if (pos > 0 && !is_gzip(rest, nrest)) {
            bool zeros = true;
            for (uint64_t i = 0; i < nrest && zeros; i++) zeros = rest[i] == 0;
            if (zeros)
                put(r.info, sizeof r.info, std::to_string(nrest) + " zero bytes after the last member");
            else
                put(r.note, sizeof r.note,
                    std::to_string(nrest) + " bytes after the last member are not gzip data; ignored");
            break;
        }
        uint64_t hlen = 0, member = 0;
        std::string err;
        if (!parse_header(rest, nrest, hlen, member, err)) return fail(GZ_CORRUPT, err);
        (void)member;  // the reference routes small members to the CPU by it; here every member is the GPU's
        // the last 4 bytes are the size of a single member's output (mod 4 GiB): a hint
        const uint64_t hint = pos == 0 && size >= 18 ? le32(data + size - 4) : 0;
        Counted c{sink, user, 0};
        gi_stats st;
        memset(&st, 0, sizeof st);
        const int rc = gi_inflate(ctx, rest + hlen, nrest - hlen, hint, counted_sink, &c, &st);
        if (rc == GI_ERR_DATA || rc == GI_ERR_TRUNCATED) return fail(GZ_CORRUPT, gi_last_error(ctx));
        if (rc == GI_ERR_ABORTED) return fail(GZ_STOPPED, "stopped by the caller");
        if (rc != GI_OK) return fail(GZ_RETRY, std::string("GPU error: ") + gi_last_error(ctx));
        add_stats(r.stats, st);
        const uint64_t t = hlen + st.in_used;
        if (t + 8 > nrest) return fail(GZ_CORRUPT, "the file ends before a member's trailer");
        const uint32_t want_crc = le32(rest + t), want_size = le32(rest + t + 4);
        if (gi_last_crc32(ctx) != want_crc || (uint32_t)c.size != want_size)
            return fail(GZ_RETRY, std::string("GPU output failed the ") + ((uint32_t)c.size != want_size ? "size" : "CRC") +
                                      " check");
        r.members++;
        r.out += c.size;
        pos += t + 8;
    }
    if (r.members == 0) return fail(GZ_CORRUPT, "no gzip member");
    r.packed = size;
    return GZ_OK;
}

}  // extern "C"
