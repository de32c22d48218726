// rescuecheck.cpp - the rescue (rescue_core.h) over a CPU backend, and the sequential decoder
// (seqdec.h) alone: librescuecheck.so, for tests only. The GPU library (libgzrescue.so) runs the
// same decisions over its GPU backend; this build lets every decision be checked against the
// reference's rescue on the CPU, and the decoder against the reference decoder field by field.
// Nothing in the product loads it.
#include <stdint.h>
#include <string.h>

#include <atomic>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "rescue_core.h"

#define RC_API extern "C" __attribute__((visibility("default")))

using gz::Start;

namespace {

// the host's threads over [0, n)
template <class F>
void each_index(size_t n, F&& fn) {
    const unsigned nt = (unsigned)std::min<size_t>(std::max(1u, std::thread::hardware_concurrency()), n);
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; t++)
        pool.emplace_back([&] {
            for (size_t i; (i = next.fetch_add(1)) < n;) fn(i);
        });
    for (auto& th : pool) th.join();
}

struct CpuRescue {
    const uint8_t* in = nullptr;
    uint64_t size = 0;
    std::vector<std::vector<uint16_t>> outs;  // each decode's 16-bit output, by id
    std::mutex mu;

    void attach(const uint8_t* p, uint64_t n) {
        in = p;
        size = n;
    }
    // every dynamic header zlib accepts in [from, 8 size - 128) from which a decode goes on for
    // VERIFY_OUT symbols, two blocks, or to the end (rescue.cpp:548-626)
    void find(uint64_t from, std::vector<Start>& out) {
        out.clear();
        const uint64_t nbits = size * 8, end = nbits > 128 ? nbits - 128 : 0;
        if (from >= end) return;
        std::vector<uint32_t> words((size + 3) / 4 + 8, 0);
        memcpy(words.data(), in, size);
        gz::Bits proto;
        proto.setup(words.data(), size);
        const uint64_t w0 = (from - 1) / 32, w1 = (end + 31) / 32;
        constexpr uint64_t PART = 1 << 16;
        const size_t nparts = (size_t)((w1 - w0 + PART - 1) / PART);
        std::vector<std::vector<Start>> found(nparts);
        each_index(nparts, [&](size_t i) {
            seqd::Inflater dec;
            std::vector<uint16_t> scratch(rsc::VERIFY_OUT + 2 * gz::MAX_SYMBOL_OUT);
            for (uint64_t w = w0 + i * PART; w < std::min(w1, w0 + (i + 1) * PART); w++) {
                uint32_t m = gz::find_word_hits(proto.word(w), proto.word(w + 1), proto.word(w + 2), proto.word(w + 3), w,
                                                from - 1, end);
                while (m) {
                    const int o = __builtin_ctz(m);
                    m &= m - 1;
                    const uint64_t bit = w * 32 + (uint64_t)o;
                    if (!gz::plausible_dynamic_header(proto, nbits, bit)) continue;
                    const Start s = rsc::at_block(bit);
                    const seqd::Outcome r = dec.run(in, size, s, nullptr, 0, scratch.data(), (uint32_t)rsc::VERIFY_OUT, false);
                    const bool ok = r.status == gz::OK ||
                                    (r.status == gz::ERR_OVERFLOW && (r.blocks >= rsc::VERIFY_BLOCKS ||
                                                                      r.out_len >= rsc::VERIFY_OUT - gz::MAX_SYMBOL_OUT));
                    if (ok) found[i].push_back(s);
                }
            }
        });
        for (auto& f : found) out.insert(out.end(), f.begin(), f.end());
    }
    // a piece decoded to its end, resumed past full buffers (rescue.cpp:661-750's loop)
    void decode(const Start& s0, bool stream_start, size_t t0, const std::vector<Start>& h, rsc::Decoded& d,
                seqd::Inflater& dec, std::vector<uint16_t>& scratch, std::vector<uint16_t>& out) {
        Start s = s0;
        const Start* tg = h.data() + std::min(t0, h.size());
        const int nt = (int)(h.size() - std::min(t0, h.size()));
        d = rsc::Decoded{};
        for (;;) {
            const seqd::Outcome r = dec.run(in, size, s, tg, nt, scratch.data(), (uint32_t)scratch.size(), stream_start && out.empty());
            const uint64_t base = out.size();
            out.insert(out.end(), scratch.begin(), scratch.begin() + r.out_len);
            d.blocks += r.blocks;
            d.symbols += r.symbols;
            if (r.last_block_bit != UINT64_MAX) {
                d.last_block_bit = r.last_block_bit;
                d.last_block_at = base + r.last_block_out;
            }
            if (r.status == gz::ERR_OVERFLOW) {
                s = r.resume;
                stream_start = false;
                continue;
            }
            d.status = r.status;
            d.end_bit = r.end_bit;
            d.landed = r.landed;
            break;
        }
        d.len = out.size();
    }
    void decode_headers(const std::vector<Start>& h, std::vector<rsc::Decoded>& res) {
        res.assign(h.size(), rsc::Decoded{});
        const size_t first = outs.size();
        outs.resize(first + h.size());
        each_index(h.size(), [&](size_t i) {
            static thread_local seqd::Inflater dec;
            static thread_local std::vector<uint16_t> scratch(rsc::SCRATCH + 2 * gz::MAX_SYMBOL_OUT);
            decode(h[i], false, i + 1, h, res[i], dec, scratch, outs[first + i]);
            res[i].id = (int)(first + i);
        });
    }
    void decode_other(const Start& s, bool stream_start, size_t t0, const std::vector<Start>& h, rsc::Decoded& d) {
        seqd::Inflater dec;
        std::vector<uint16_t> scratch(rsc::SCRATCH + 2 * gz::MAX_SYMBOL_OUT), out;
        decode(s, stream_start, t0, h, d, dec, scratch, out);
        d.id = (int)outs.size();
        outs.push_back(std::move(out));
    }
    // a pass's bytes in output order: each reference copies what it reaches, or the lost
    // window's variable, or nothing known
    void resolve(const std::vector<rsc::Span>& spans, uint64_t total, rsc::Provenance& prov) {
        rsc::HostProv& pv = rsc::HostProv::make(prov, total);
        for (const rsc::Span& sp : spans) {
            const uint16_t* m = outs[sp.id].data();
            for (uint64_t i = 0; i < sp.len; i++) {
                const uint64_t x = sp.out + i;
                const uint16_t v = m[i];
                if (v <= gz::REF) {
                    pv.byte[x] = (uint8_t)v;
                    pv.tag[x] = rsc::Provenance::EXACT;
                    continue;
                }
                const uint64_t d = v - gz::REF;
                if (d <= i || (sp.has_window && d - i <= sp.window)) {
                    pv.byte[x] = pv.byte[x - d];
                    pv.tag[x] = pv.tag[x - d];
                } else {
                    pv.byte[x] = 0;
                    pv.tag[x] = sp.var_base && d - i <= gz::WINDOW ? sp.var_base + (uint32_t)(d - i) : 0;
                }
            }
            if (sp.blank_tail)
                for (uint64_t i = sp.len > gz::MAX_SYMBOL_OUT ? sp.len - gz::MAX_SYMBOL_OUT : 0; i < sp.len; i++) pv.tag[sp.out + i] = 0;
        }
    }
    void summarize(const rsc::Provenance& pv, const std::vector<rsc::PieceRange>& pieces, std::vector<rsc::PieceFacts>& facts,
                   const std::vector<std::pair<uint64_t, uint64_t>>& crc_ranges, std::vector<uint32_t>& crcs) {
        rsc::HostProv::summarize(pv, pieces, facts, crc_ranges, crcs);
    }
    std::vector<rsc::Place> all_places;  // of the last var_tags call
    void var_tags(const rsc::Provenance& pv, const std::vector<std::pair<uint64_t, uint64_t>>& stretches,
                  std::vector<uint32_t>& tags) {
        rsc::HostProv::var_places(pv, stretches, all_places);
        tags.resize(all_places.size());
        for (size_t i = 0; i < all_places.size(); i++) tags[i] = all_places[i].tag;
    }
    void places(const std::vector<uint64_t>& picked, std::vector<rsc::Place>& out) {
        out.resize(picked.size());
        for (size_t j = 0; j < picked.size(); j++) out[j] = all_places.at(picked[j]);
    }
    void fetch(const rsc::Provenance& pv, uint64_t at, uint64_t len, uint8_t* byte, uint32_t* tag) {
        rsc::HostProv::fetch(pv, at, len, byte, tag);
    }
    int emit(const rsc::Provenance& pv, const std::vector<rsc::EmitPiece>& pieces, const rsc::Inferred& value, uint8_t fill,
             gr_sink sink, void* user, std::vector<rsc::EmitCounts>& counts, uint64_t& emitted) {
        return rsc::HostProv::emit(pv, pieces, value, fill, sink, user, counts, emitted);
    }
};

}  // namespace

struct gr_result : rsc::RescueResult {};

namespace {
int rescue_entry(const uint8_t* in, uint64_t size, bool gzip, uint64_t expected_size, bool has_crc, uint32_t expected_crc,
                 const gr_options* opt, gr_sink sink, void* user, gr_result** result) {
    if (result) *result = nullptr;
    if (!in && size) return GR_ERR_ARG;
    try {
        CpuRescue be;
        std::unique_ptr<gr_result> res(new gr_result);
        const int rc = rsc::run_rescue(be, in, size, gzip, expected_size, has_crc, expected_crc, rsc::read_options(opt), sink,
                                       user, *res);
        if (rc != GR_OK && rc != GR_ERR_ABORTED) return rc;
        if (result) *result = res.release();
        return rc;
    } catch (const std::bad_alloc&) {
        return GR_ERR_MEMORY;
    } catch (const std::runtime_error&) {
        return GR_ERR_INTERNAL;
    }
}
}  // namespace

RC_API int rc_gr_rescue(const uint8_t* in, uint64_t size, const gr_options* opt, gr_sink sink, void* user, gr_result** result) {
    return rescue_entry(in, size, true, UINT64_MAX, false, 0, opt, sink, user, result);
}
RC_API int rc_gr_rescue_raw(const uint8_t* in, uint64_t size, uint64_t expected_size, int has_crc, uint32_t expected_crc,
                            const gr_options* opt, gr_sink sink, void* user, gr_result** result) {
    return rescue_entry(in, size, false, expected_size, has_crc != 0, expected_crc, opt, sink, user, result);
}
RC_API const gr_stats* rc_gr_result_stats(const gr_result* r) { return r ? &r->stats : nullptr; }
RC_API uint32_t rc_gr_result_pieces(const gr_result* r, const gr_piece** pieces) {
    if (!r) return 0;
    if (pieces) *pieces = r->pieces.data();
    return (uint32_t)r->pieces.size();
}
RC_API uint32_t rc_gr_result_notes(const gr_result* r, const char* const** notes) {
    if (!r) return 0;
    if (notes) *notes = r->note_ptrs.data();
    return (uint32_t)r->note_ptrs.size();
}
RC_API void rc_gr_free(gr_result* r) { delete r; }
// the host's CRC combine (host_crc.h), for its test against direct CRCs
RC_API uint32_t rc_crc32_combine(uint32_t crc_a, uint32_t crc_b, uint64_t len_b) { return hostcrc::combine(crc_a, crc_b, len_b); }
RC_API uint32_t rc_gzr_result_phases(const gr_result* r, const char** names, double* seconds, uint32_t cap) {
    if (!r) return 0;
    for (uint32_t i = 0; i < cap && i < r->phases.size(); i++) {
        if (names) names[i] = r->phases[i].first.c_str();
        if (seconds) seconds[i] = r->phases[i].second;
    }
    return (uint32_t)r->phases.size();
}

// the sequential decoder alone: the reference Result's layout (oracle_lib.OrResult)
struct RcResult {
    uint64_t end_bit, eob_bit;
    struct {
        uint64_t bit, header;
        uint8_t mode;
        int8_t final;
    } resume;
    uint32_t eob_out, out_len;
    int32_t status, landed, end_final;
    uint32_t blocks;
    uint64_t last_block_bit;
    uint32_t last_block_out, symbols;
};
struct RcStart {
    uint64_t bit, header;
    uint8_t mode;
    int8_t final;
};

RC_API void rc_run(const uint8_t* in, uint64_t size, const RcStart* s, const RcStart* targets, int ntargets, uint16_t* out,
                   uint32_t cap, int stream_start, uint64_t stop, RcResult* res) {
    std::vector<uint8_t> pad((size + 3) / 4 * 4 + 64, 0);
    memcpy(pad.data(), in, size);
    std::vector<Start> tg(ntargets);
    for (int i = 0; i < ntargets; i++) {
        tg[i] = rsc::at_block(targets[i].bit);
        tg[i].header = targets[i].header;
        tg[i].mode = targets[i].mode;
        tg[i].final = targets[i].final;
    }
    Start st = rsc::at_block(s->bit);
    st.header = s->header;
    st.mode = s->mode;
    st.final = s->final;
    static thread_local seqd::Inflater dec;
    const seqd::Outcome r = dec.run(pad.data(), size, st, tg.data(), ntargets, out, cap, stream_start != 0, stop);
    res->end_bit = r.end_bit;
    res->eob_bit = r.eob_bit;
    res->resume.bit = r.resume.bit;
    res->resume.header = r.resume.header;
    res->resume.mode = r.resume.mode;
    res->resume.final = r.resume.final;
    res->eob_out = r.eob_out;
    res->out_len = r.out_len;
    res->status = r.status;
    res->landed = r.landed;
    res->end_final = r.end_final;
    res->blocks = r.blocks;
    res->last_block_bit = r.last_block_bit;
    res->last_block_out = r.last_block_out;
    res->symbols = r.symbols;
}

RC_API void rc_probe(const uint8_t* in, uint64_t size, uint8_t mode, uint64_t header, uint64_t from_bit, uint64_t at_bit,
                     RcStart* out) {
    std::vector<uint8_t> pad((size + 3) / 4 * 4 + 64, 0);
    memcpy(pad.data(), in, size);
    static thread_local seqd::Inflater dec;
    const Start s = dec.probe(pad.data(), size, mode, header, from_bit, at_bit);
    out->bit = s.bit;
    out->header = s.header;
    out->mode = s.mode;
    out->final = s.final;
}
