// gzip_format.h - the gzip member format (RFC 1952) as the host reads it: a member's header, its
// optional fields and BGZF size, with the reference host's acceptance rules and messages
// (src/gz.rs:62-110). Used by the member driver (gzhost.cpp) and the serving layer of
// gpuinflate.hip.
#pragma once
#include <stdint.h>
#include <string.h>

#include <string>

#include "host_crc.h"

namespace gzf {

// Reads a member header front to back. Every read past the end of the data marks the reader
// short (and returns zeros), so a header is parsed straight through and checked once at the end.
class HeaderCursor {
  public:
    HeaderCursor(const uint8_t* d, uint64_t n) : d_(d), n_(n) {}
    uint8_t byte() { return has(1) ? d_[p_++] : 0; }
    uint32_t le16() {
        const uint32_t lo = byte();
        return lo | (uint32_t)byte() << 8;
    }
    const uint8_t* take(uint64_t k) {
        if (!has(k)) return nullptr;
        const uint8_t* at = d_ + p_;
        p_ += k;
        return at;
    }
    // past the next zero byte (a zero-terminated field)
    void skip_cstring() {
        const void* z = p_ < n_ ? memchr(d_ + p_, 0, (size_t)(n_ - p_)) : nullptr;
        if (!z) {
            short_ = true;
            p_ = n_;
            return;
        }
        p_ = (uint64_t)((const uint8_t*)z - d_) + 1;
    }
    uint64_t pos() const { return p_; }
    bool is_short() const { return short_; }

  private:
    bool has(uint64_t k) {
        if (k <= n_ - p_) return true;
        short_ = true;
        p_ = n_;
        return false;
    }
    const uint8_t* d_;
    uint64_t n_, p_ = 0;
    bool short_ = false;
};

enum : uint8_t { FTEXT = 1, FHCRC = 2, FEXTRA = 4, FNAME = 8, FCOMMENT = 16, FRESERVED = 0xE0 };

struct MemberHeader {
    uint64_t len = 0;         // header bytes
    uint64_t bgzf_total = 0;  // BGZF: the whole member's bytes, from the BC subfield (0: not given)
};

// the BGZF "BC" subfield of an FEXTRA field (SAM/BAM specification: BSIZE, the member's size
// less one); 0 when absent
inline uint64_t bgzf_size(const uint8_t* x, uint64_t n) {
    uint64_t found = 0;
    for (uint64_t at = 0; at + 4 <= n;) {
        const uint64_t sublen = x[at + 2] | (uint64_t)x[at + 3] << 8;
        if (x[at] == 'B' && x[at + 1] == 'C' && sublen == 2 && at + 6 <= n) found = (x[at + 4] | (uint64_t)x[at + 5] << 8) + 1;
        at += 4 + sublen;
    }
    return found;
}

// RFC 1952 section 2.3: ID1 ID2 CM FLG MTIME(4) XFL OS, then the optional fields FLG names
inline bool read_header(const uint8_t* d, uint64_t n, MemberHeader& h, std::string& why) {
    const char* const truncated = "the file ends inside a member header";
    if (n < 10 || d[0] != 0x1f || d[1] != 0x8b) {
        why = "not a gzip member";
        return false;
    }
    if (d[2] != 8) {
        why = "compression method " + std::to_string(d[2]) + " is not Deflate";
        return false;
    }
    const uint8_t flags = d[3];
    if (flags & FRESERVED) {
        why = "reserved header flags are set";
        return false;
    }
    HeaderCursor c(d, n);
    c.take(10);
    if (flags & FEXTRA) {
        const uint64_t xlen = c.le16();
        const uint8_t* x = c.take(xlen);
        if (x) h.bgzf_total = bgzf_size(x, xlen);
    }
    if (flags & FNAME) c.skip_cstring();
    if (flags & FCOMMENT) c.skip_cstring();
    if (flags & FHCRC) {
        const uint64_t covered = c.pos();
        const uint32_t want = c.le16();
        if (!c.is_short() && (hostcrc::crc32(0, d, (size_t)covered) & 0xFFFF) != want) {
            why = "header CRC mismatch";
            return false;
        }
    }
    if (c.is_short()) {
        why = truncated;
        return false;
    }
    h.len = c.pos();
    return true;
}

inline uint32_t le32(const uint8_t* p) { return p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }

}  // namespace gzf
