// gzhost.cpp - the gzip member driver (include/gzhost.h) above gi_inflate: a .gz file held in
// memory is walked member by member (RFC 1952), each member's Deflate data decoded on the GPU,
// and each member's trailer checked against what reached the caller. It has the role of the
// reference host's member loop (src/gz.rs:154-264) and header parser (src/gz.rs:62-110): the
// same failure classes and messages, the same treatment of bytes after the last member.
//
// The trailer's CRC-32 is checked twice: against the CRC the GPU computed as it wrote the output
// (gi_last_crc32), and against a CRC of the bytes as the sink received them, computed here on the
// host (host_crc.h, about 12 GB/s a core), so that a fault between the GPU and the sink (a copy,
// a pinned buffer) is a GZ_RETRY too, as the reference's host CRC makes it (src/gz.rs:205-213).
#define GZ_BUILD
#include "gzhost.h"

#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "gzip_format.h"
#include "host_crc.h"

namespace {

using gzf::le32;
using gzf::MemberHeader;
using gzf::read_header;

void set_text(char* dst, size_t cap, const std::string& s) {
    if (dst && cap) snprintf(dst, cap, "%s", s.c_str());
}

// Between gi_inflate and the caller's sink: the bytes that pass, counted and CRC'd
struct Delivery {
    gi_sink sink;
    void* user;
    uint64_t bytes = 0;
    uint32_t crc = 0;
    static int pass(void* self, const uint8_t* data, size_t len) {
        Delivery& d = *(Delivery*)self;
        d.crc = hostcrc::crc32(d.crc, data, len);
        d.bytes += len;
        return d.sink ? d.sink(d.user, data, len) : 0;
    }
};

void accumulate(gi_stats& sum, const gi_stats& s) {
    sum.in_used += s.in_used;
    sum.out_size += s.out_size;
    sum.segments += s.segments;
    sum.rounds += s.rounds;
    sum.jobs += s.jobs;
    sum.seconds += s.seconds;
    sum.upload_s += s.upload_s;
    sum.find_s += s.find_s;
    sum.decode_s += s.decode_s;
    sum.resolve_s += s.resolve_s;
    sum.download_s += s.download_s;
}

// One call of gz_decode: the members from the start of the file, one at a time
class MemberLoop {
  public:
    MemberLoop(gi_ctx* ctx, const uint8_t* data, uint64_t size, gi_sink sink, void* user, gz_report& rep)
        : ctx_(ctx), data_(data), size_(size), sink_(sink), user_(user), rep_(rep) {}

    int run() {
        uint64_t at = 0;
        while (at < size_) {
            if (at > 0 && !starts_member(at)) {
                describe_trailing(at);
                break;
            }
            // a run of BGZF members from here; the member it stops at (if it stopped at one that
            // did not check) and any other member go by themselves
            uint64_t next = at;
            bool stopped = false;
            int rc = bgzf_run(at, next, stopped);
            if (rc != GZ_OK) return rc;
            if (next > at && !stopped) {
                at = next;
                continue;
            }
            at = next;
            rc = member(at, next);
            if (rc != GZ_OK) return rc;
            at = next;
        }
        if (rep_.members == 0) return fail(GZ_CORRUPT, "no gzip member");
        rep_.packed = size_;
        return GZ_OK;
    }

  private:
    bool starts_member(uint64_t at) const {
        return size_ - at >= 3 && data_[at] == 0x1f && data_[at + 1] == 0x8b && data_[at + 2] == 8;
    }
    // what follows the last member: zeros (padding some tools write) or other data, ignored
    void describe_trailing(uint64_t at) {
        const uint64_t n = size_ - at;
        uint64_t z = 0;
        while (z < n && data_[at + z] == 0) z++;
        if (z == n)
            set_text(rep_.info, sizeof rep_.info, std::to_string(n) + " zero bytes after the last member");
        else
            set_text(rep_.note, sizeof rep_.note, std::to_string(n) + " bytes after the last member are not gzip data; ignored");
    }
    int fail(int code, const std::string& why) {
        set_text(rep_.error, sizeof rep_.error, why);
        return code;
    }
    // BGZF members (each one's size in its header's BC subfield, SAM/BAM specification) from byte
    // `at` on, two or more, decoded together by gi_inflate_many: a file of many small members is
    // then a few sets of launches instead of a call a member. The members whose results match
    // their trailers reach the sink in order; the first that does not is left to member(), which
    // decodes it again and reports it as it would have. next: the byte after the last one taken
    // (at: none taken); stopped: the run ended at a member that did not check.
    int bgzf_run(uint64_t at, uint64_t& next, bool& stopped) {
        constexpr size_t RUN_MEMBERS = 16384;
        constexpr uint64_t RUN_OUT = 256ull << 20;
        struct Member {
            uint64_t at, hlen, total;
            uint32_t crc, isize;
        };
        std::vector<Member> run;
        uint64_t p = at, out = 0;
        while (p < size_ && run.size() < RUN_MEMBERS && out < RUN_OUT && starts_member(p)) {
            MemberHeader h;
            std::string why;
            if (!read_header(data_ + p, size_ - p, h, why) || !h.bgzf_total) break;
            const uint64_t total = h.bgzf_total;
            if (total < h.len + 8 || total > size_ - p) break;
            run.push_back(Member{p, h.len, total, le32(data_ + p + total - 8), le32(data_ + p + total - 4)});
            out += run.back().isize;
            p += total;
        }
        next = at;
        if (run.size() < 2) return GZ_OK;
        std::vector<gi_stream> st(run.size());
        uint64_t room = 0;
        for (size_t k = 0; k < run.size(); k++) {
            const Member& m = run[k];
            memset(&st[k], 0, sizeof st[k]);
            st[k].offset = m.at + m.hlen;
            st[k].size = m.total - m.hlen - 8;
            st[k].out_offset = room;
            st[k].out_room = m.isize;
            room += m.isize;
        }
        buf_.resize(room);
        gi_stats stats;
        const int rc = gi_inflate_many(ctx_, data_, st.data(), st.size(), buf_.data(), room, &stats);
        if (rc == GI_ERR_GPU) return fail(GZ_RETRY, std::string("GPU error: ") + gi_last_error(ctx_));
        for (size_t k = 0; k < run.size(); k++) {
            const Member& m = run[k];
            const gi_stream& s = st[k];
            if (s.status != GI_OK || s.in_used != s.size || s.out_size != m.isize || s.crc32 != m.crc) {
                stopped = true;
                break;
            }
            Delivery got{sink_, user_};
            if (s.out_size && Delivery::pass(&got, buf_.data() + s.out_offset, (size_t)s.out_size))
                return fail(GZ_STOPPED, "stopped by the caller");
            if (got.crc != m.crc) return fail(GZ_RETRY, "GPU output failed the CRC check");
            rep_.members++;
            rep_.out += got.bytes;
            rep_.stats.in_used += s.in_used;
            rep_.stats.out_size += s.out_size;
            next = m.at + m.total;
        }
        rep_.stats.segments += stats.segments;
        rep_.stats.jobs += stats.jobs;
        rep_.stats.seconds += stats.seconds;
        return GZ_OK;
    }

    // the member at byte `at`; *next: the byte after its trailer
    int member(uint64_t at, uint64_t& next) {
        const uint8_t* m = data_ + at;
        const uint64_t avail = size_ - at;
        MemberHeader h;
        std::string why;
        if (!read_header(m, avail, h, why)) return fail(GZ_CORRUPT, why);
        // a file of one member ends with its output size mod 2^32: the first member's hint
        const uint64_t hint = at == 0 && size_ >= 18 ? le32(data_ + size_ - 4) : 0;
        Delivery got{sink_, user_};
        gi_stats st;
        memset(&st, 0, sizeof st);
        const int rc = gi_inflate(ctx_, m + h.len, avail - h.len, hint, &Delivery::pass, &got, &st);
        switch (rc) {
            case GI_OK:
                break;
            case GI_ERR_DATA:
            case GI_ERR_TRUNCATED:
                return fail(GZ_CORRUPT, gi_last_error(ctx_));
            case GI_ERR_ABORTED:
                return fail(GZ_STOPPED, "stopped by the caller");
            default:
                return fail(GZ_RETRY, std::string("GPU error: ") + gi_last_error(ctx_));
        }
        accumulate(rep_.stats, st);
        const uint64_t trailer = h.len + st.in_used;
        if (trailer + 8 > avail) return fail(GZ_CORRUPT, "the file ends before a member's trailer");
        const uint32_t crc = le32(m + trailer), isize = le32(m + trailer + 4);
        if ((uint32_t)got.bytes != isize) return fail(GZ_RETRY, "GPU output failed the size check");
        if (gi_last_crc32(ctx_) != crc || got.crc != crc) return fail(GZ_RETRY, "GPU output failed the CRC check");
        rep_.members++;
        rep_.out += got.bytes;
        next = at + trailer + 8;
        return GZ_OK;
    }

    gi_ctx* const ctx_;
    const uint8_t* const data_;
    const uint64_t size_;
    const gi_sink sink_;
    void* const user_;
    gz_report& rep_;
    std::vector<uint8_t> buf_;  // a BGZF run's output
};

}  // namespace

extern "C" {

GZ_API int gz_header(const uint8_t* d, uint64_t n, uint64_t* len, uint64_t* member, char* err, size_t err_len) {
    if (d == nullptr) {
        set_text(err, err_len, "no data");
        return GZ_CORRUPT;
    }
    MemberHeader h;
    std::string why;
    if (!read_header(d, n, h, why)) {
        set_text(err, err_len, why);
        return GZ_CORRUPT;
    }
    if (len) *len = h.len;
    if (member) *member = h.bgzf_total;
    set_text(err, err_len, "");
    return GZ_OK;
}

GZ_API int gz_decode(gi_ctx* ctx, const uint8_t* data, uint64_t size, gi_sink sink, void* user, gz_report* rep) {
    gz_report scratch;
    gz_report& r = rep ? *rep : scratch;
    memset(&r, 0, sizeof r);
    if (ctx == nullptr || (data == nullptr && size)) {
        set_text(r.error, sizeof r.error, "no data");
        return GZ_CORRUPT;
    }
    return MemberLoop(ctx, data, size, sink, user, r).run();
}

}  // extern "C"
