// rescue_core.h - the decisions of a rescue (include/gzrescue.h, the reference's gr_* interface
// v4, gpuinflate/rescue.h:38-107): what a damaged gzip or raw Deflate stream still holds, with a
// per-byte mask. Host code, over a Backend that does the data-parallel work:
//
//   Backend B provides
//     const uint8_t* in; uint64_t size;                      the prepared input (zero padded)
//     void find(uint64_t from, std::vector<Start>& out)       every confirmed dynamic header in
//                                                             [from, 8 size - 128), in bit order
//     void decode_headers(const std::vector<Start>& h, std::vector<Decoded>& out)
//                                                             the piece decoded from each header,
//                                                             landing on the headers after it
//     int  decode_other(Start s, bool stream_start, size_t t0, const std::vector<Start>& h, Decoded& out)
//     void resolve(const std::vector<Span>& spans, uint64_t total, Provenance& prov)
//                                                             each output byte of a pass: its value,
//                                                             exact or copied from a lost window
//                                                             (kept by the backend, in prov.store)
//     and what the decisions read off a pass's provenance:
//     void summarize(prov, pieces, facts, crc_ranges, crcs)  each piece's exactness, first variable
//                                                             at a multiple of 64, kept ranges with
//                                                             variables; CRC-32 of byte ranges
//     void var_tags(prov, stretches, tags)                    the tag of every byte that copies a
//                                                             variable, in order
//     void places(picked, places)                             of those, the picked ones with their
//                                                             neighbours
//     void fetch(prov, at, len, byte, tag)                    a few bytes and tags (model training)
//     int  emit(prov, pieces, inferred, fill, sink, user, counts, emitted)
//                                                             the output and its mask to the sink
//   (the GPU backend in gzrescue.hip, provenance in HBM; a CPU one, HostProv below, for the tests).
//
// The steps and their outcomes follow the reference rescue (gpuinflate/rescue.cpp); the
// structure is this design's: a pass is first walked on the pieces' decode results alone (which
// start where, which is damaged, which window each has), and the provenance of every output byte
// of the pass is then resolved in one step over the whole output, instead of byte by byte while
// decoding (rescue.cpp:661-750). Whatever the walk or a later check needs to decode again is
// decoded on the host by seqdec.h, a decoder with the reference's result contract.
#pragma once
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gzrescue.h"
#include "host_crc.h"
#include "seqdec.h"

namespace rsc {

using gz::Start;
using seqd::Outcome;

// the reference's thresholds (rescue.cpp:67-80): the outcomes depend on them
constexpr uint32_t SCRATCH = 32u << 20;          // symbols one host decode writes before it resumes
constexpr uint64_t VERIFY_OUT = 256u << 10;
constexpr uint32_t VERIFY_BLOCKS = 2;
constexpr uint64_t FIXED_SEARCH = 8ull << 23;     // bits past a failure searched for a fixed-code resync
constexpr uint64_t RUN = 16;                      // zero or 0xFF bytes in a row that mark where damage begins
constexpr uint64_t MAX_SYMBOL_BITS = 48;          // the most bits one symbol reads
constexpr size_t MAX_OCC = 64;                    // places per lost window byte inference reads
constexpr uint64_t TRAIN = 1u << 20;              // exact bytes the inference model learns from
constexpr uint64_t FLIP_BITS = 16384;             // how far back from a failure a flipped bit is sought
constexpr int FLIP_PIECES = 8;
constexpr uint32_t NO_VAR_LEFT = 0xF0000000u;     // variables are numbered below this

inline Start at_block(uint64_t bit) {
    Start s{};
    s.bit = bit;
    s.header = bit;
    s.mode = gz::AT_BLOCK;
    s.final = -1;
    return s;
}
inline Start no_start() { return at_block(UINT64_MAX); }

// One decode of a pass: from a start, landing on the headers after it, resumed past full
// buffers until it ends (rescue.cpp:661-750's loop over run())
struct Decoded {
    int id = -1;              // the backend's handle on its 16-bit output
    uint64_t len = 0;         // output bytes
    int32_t status = gz::OK;  // of its last run
    uint64_t end_bit = 0;
    int32_t landed = -1;
    uint32_t blocks = 0, symbols = 0;
    uint64_t last_block_bit = UINT64_MAX;  // header of the last block it started
    uint64_t last_block_at = 0;            // where that block's output begins
};

// How a stretch of a pass's output is made: decode `id`'s first len bytes; references before
// them reach the window of the output before it (the `window` bytes before `out`, when it has
// one) or else a lost window, whose byte k back is variable var_base + k (0: unknown)
struct Span {
    int id;
    uint64_t out, len;
    bool has_window;
    uint64_t window;    // bytes of it that references may reach
    uint32_t var_base;
    bool blank_tail;    // the last MAX_SYMBOL_OUT bytes may hold bits past the file's end: unknown
};

// Each output byte of a pass: its value (0 where unknown) and a tag: EXACT, or the variable it
// copies (0: nothing known). The backend keeps them (host arrays, or device memory) and answers
// what the decisions ask of them.
struct Provenance {
    static constexpr uint32_t EXACT = 0xFFFFFFFFu;
    uint64_t total = 0;
    std::shared_ptr<void> store;
};

struct Piece {
    uint64_t in_bit_start = 0, in_bit_end = 0;
    Start start = no_start();
    uint64_t len = 0;
    size_t src = 0;        // the pass decode it is (part of)
    uint64_t src_off = 0;  // its first byte in that decode
    bool all_exact = true;
    uint32_t var_base = 0;
    bool suspect = false;
    std::string why;
    uint64_t true_len = 0;
    bool stream_start = false, landed = false, anchored = false, by_crc = false, length_exact = false;
    int status = gz::OK;
    uint32_t blocks = 0, symbols = 0;
    uint64_t last_block_bit = UINT64_MAX;
    uint64_t last_block_at = 0;
};

struct Source {
    Start start = no_start();
    bool stream_start = false;
    size_t t0 = 0;
    uint32_t var_base = 0;
    uint64_t out = 0, len = 0;
    bool window_after = false, blank_tail = false;
    int id = -1;
    uint64_t lost64 = UINT64_MAX;  // the first offset, a multiple of 64, of a byte that copies a variable
};

struct Member {
    size_t first_piece = 0, last_piece = 0;
    bool trailer = false, has_crc = false, size_known = false;
    uint32_t crc = 0, isize = 0;
    uint32_t out_crc = 0;
    int crc_check = -1;
};

// One place a variable appears: the bytes around it in its decode, 5 before and 6 after: a byte
// (< VAR), another variable (VAR + its number) or nothing known (NONE) (rescue.cpp:139-144)
struct Occ {
    static constexpr uint32_t NONE = 0xFFFFFFFFu, VAR = 0x100;
    static constexpr int BEFORE = 5, AFTER = 6, CELLS = 11;
    uint64_t pos;
    uint32_t cell[CELLS];  // i < BEFORE: i + 1 bytes before; else i - BEFORE + 1 after
};

// What a pass's summaries ask of a piece: its stretch of the provenance (from its decode's first
// byte on, len bytes), and whether its ranges with variables are kept
struct PieceRange {
    uint64_t at, len;
    bool keep;
};
// ... and what they learn: no byte other than EXACT; the first offset, a multiple of 64, of a byte
// that copies a variable (len: none); the ranges [a, b) with variables (kept pieces only)
struct PieceFacts {
    bool all_exact = true;
    uint64_t var64 = 0;
    std::vector<std::pair<uint64_t, uint64_t>> runs;
};
// A byte that copies a variable (tag), with the cells around it in its stretch
struct Place {
    uint64_t pos;
    uint32_t tag;
    uint32_t cell[Occ::CELLS];
};
// A piece as emitted: the provenance's bytes [at, at + min(decoded, length)), then the fill byte
// up to length
struct EmitPiece {
    uint64_t at, decoded, length;
    bool suspect;
};
struct EmitCounts {
    uint64_t known = 0, exact = 0, suspect = 0, unknown = 0;
};

// The variables a pass met, in the order met: how many bytes copy each, and its first MAX_OCC
// places (those kept; with drop_alike, a place alike a kept one is only counted)
struct Vars {
    struct Var {
        uint32_t var = 0;
        uint16_t met = 0;
        uint64_t count = 0;
        std::vector<Occ> places;
    };
    std::vector<Var> list;
    std::vector<uint32_t> slots;  // open addressing: index + 1
    static uint32_t slot_of(uint32_t v, uint32_t mask) { return (uint32_t)(((uint64_t)v * 0x9E3779B97F4A7C15ull) >> 32) & mask; }
    uint32_t find(uint32_t v) const {
        if (slots.empty()) return UINT32_MAX;
        const uint32_t mask = (uint32_t)slots.size() - 1;
        for (uint32_t i = slot_of(v, mask);; i = (i + 1) & mask) {
            const uint32_t k = slots[i];
            if (!k) return UINT32_MAX;
            if (list[k - 1].var == v) return k - 1;
        }
    }
    uint32_t get(uint32_t v) {
        if (2 * (list.size() + 1) > slots.size()) {
            std::vector<uint32_t> grown(slots.empty() ? 1024 : 2 * slots.size(), 0);
            const uint32_t mask = (uint32_t)grown.size() - 1;
            for (size_t k = 0; k < list.size(); k++) {
                uint32_t i = slot_of(list[k].var, mask);
                while (grown[i]) i = (i + 1) & mask;
                grown[i] = (uint32_t)k + 1;
            }
            slots.swap(grown);
        }
        const uint32_t mask = (uint32_t)slots.size() - 1;
        for (uint32_t i = slot_of(v, mask);; i = (i + 1) & mask) {
            const uint32_t k = slots[i];
            if (!k) {
                list.push_back({});
                list.back().var = v;
                slots[i] = (uint32_t)list.size();
                return (uint32_t)list.size() - 1;
            }
            if (list[k - 1].var == v) return k - 1;
        }
    }
};

struct PassResult {
    std::vector<Piece> pieces;
    std::vector<Member> members;
    std::vector<std::string> notes;
    std::vector<Source> sources;
    bool truncated = false, too_big = false;
    uint64_t input_end = 0;
    uint64_t total = 0;
    // resolved by finish_pass()
    Provenance prov;
    std::map<size_t, std::vector<std::pair<uint64_t, uint64_t>>> runs;  // ranges with variables, by source (damaged and last)
    Vars vars;
};

struct RescueResult {
    gr_stats stats{};
    std::vector<gr_piece> pieces;
    std::vector<std::string> why, notes;
    std::vector<const char*> note_ptrs;
    // wall seconds of each step, in order (name, seconds): where a rescue's time goes
    std::vector<std::pair<std::string, double>> phases;
};

// Charges the time since the last mark to a named step of a rescue
class PhaseClock {
  public:
    explicit PhaseClock(RescueResult& r) : r_(r), t_(std::chrono::steady_clock::now()) {}
    void mark(const char* name) {
        const auto now = std::chrono::steady_clock::now();
        const double s = std::chrono::duration<double>(now - t_).count();
        t_ = now;
        for (auto& ph : r_.phases)
            if (ph.first == name) {
                ph.second += s;
                return;
            }
        r_.phases.emplace_back(name, s);
    }

  private:
    RescueResult& r_;
    std::chrono::steady_clock::time_point t_;
};

// fn(i, worker) for i in [0, n) on the host's cores (at most 64 threads)
template <class F>
void on_cores(size_t n, F&& fn) {
    const unsigned nt = (unsigned)std::min<size_t>(std::max(1u, std::min(64u, std::thread::hardware_concurrency())), n);
    if (nt <= 1) {
        for (size_t i = 0; i < n; i++) fn(i, 0u);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nt; w++)
        pool.emplace_back([&, w] {
            for (size_t i; (i = next.fetch_add(1)) < n;) fn(i, w);
        });
    for (auto& t : pool) t.join();
}
inline unsigned cores() { return std::max(1u, std::min(64u, std::thread::hardware_concurrency())); }

// the length of the run of EXACT tags at the start of tg[0, n)
inline uint64_t exact_run(const uint32_t* tg, uint64_t n, uint32_t exact) {
    uint64_t i = 0;
    for (; i + 64 <= n; i += 64) {
        uint32_t any = 0;
        for (int j = 0; j < 64; j++) any |= tg[i + j] ^ exact;
        if (any) break;
    }
    while (i < n && tg[i] == exact) i++;
    return i;
}

// the first x in [0, n), a multiple of 64, whose tag is a variable (neither EXACT nor 0); n if none
inline uint64_t first_var64(const uint32_t* tg, uint64_t n, uint32_t exact) {
    for (uint64_t x = 0; x < n; x += 64)
        if (tg[x] != exact && tg[x] != 0) return x;
    return n;
}

// the ranges [a, b) of tg[0, n) whose tags are variables, in order
inline void var_ranges(const uint32_t* tg, uint64_t n, uint32_t exact, std::vector<std::pair<uint64_t, uint64_t>>& rr) {
    rr.clear();
    for (uint64_t i = 0; i < n;) {
        if (i + 64 <= n) {
            uint32_t any = 0;
            for (int j = 0; j < 64; j++) any |= (uint32_t)(tg[i + j] != exact && tg[i + j] != 0);
            if (!any) {
                i += 64;
                continue;
            }
        }
        const uint64_t e = std::min(n, i + 64);
        for (; i < e; i++) {
            const uint32_t v = tg[i];
            if (v == exact || v == 0) continue;
            if (!rr.empty() && rr.back().second == i)
                rr.back().second++;
            else
                rr.push_back({i, i + 1});
        }
    }
}


// A gzip member header at byte h: its length, or 0 when none is there (rescue.cpp:530-546)
inline uint64_t member_header(const uint8_t* in, uint64_t size, uint64_t h) {
    if (h + 10 > size || in[h] != 0x1f || in[h + 1] != 0x8b || in[h + 2] != 8) return 0;
    const uint8_t flg = in[h + 3];
    uint64_t p = h + 10;
    if (flg & 4) {
        if (p + 2 > size) return 0;
        p += 2 + (in[p] | (uint64_t)in[p + 1] << 8);
    }
    if (flg & 8) {
        while (p < size && in[p]) p++;
        p++;
    }
    if (flg & 16) {
        while (p < size && in[p]) p++;
        p++;
    }
    if (flg & 2) p += 2;
    return p <= size ? p - h : 0;
}

// the 64 input bits from `bit`
inline uint64_t bits_at(const uint8_t* in, uint64_t bit) {
    uint64_t lo, hi;
    memcpy(&lo, in + (bit >> 3), 8);
    memcpy(&hi, in + (bit >> 3) + 8, 8);
    const int off = (int)(bit & 7);
    return off ? (lo >> off) | (hi << (64 - off)) : lo;
}

// ------------------------------------------------------------------ transfer damage

// A file ending in 64 or more zero bytes ends where they begin (Deflate never writes them);
// returns the bytes cut (rescue.cpp:2106-2113)
inline uint64_t trim_zero_tail(uint64_t& size, const uint8_t* in) {
    uint64_t z = size;
    while (z > 0 && in[z - 1] == 0) z--;
    if (size - z < 64) return 0;
    const uint64_t cut = size - z;
    size = z;
    return cut;
}

// The LF bytes, CR bytes and CR LF pairs of in[0, size), counted in parts on the host's cores
struct LineEnds {
    uint64_t lf = 0, cr = 0, crlf = 0;
};
inline LineEnds line_ends(const uint8_t* in, uint64_t size) {
    constexpr uint64_t PART = 16u << 20;
    const size_t np = (size_t)((size + PART - 1) / PART);
    std::vector<LineEnds> part(np);
    on_cores(np, [&](size_t k, unsigned) {
        const uint64_t a = k * PART, e = std::min(size, a + PART);
        uint64_t lf = 0, cr = 0, crlf = 0;
        uint8_t prev = a ? in[a - 1] : 0;
        for (uint64_t i = a; i < e; i++) {
            const uint8_t b = in[i];
            lf += b == '\n';
            cr += b == '\r';
            crlf += (b == '\n') & (prev == '\r');
            prev = b;
        }
        part[k] = LineEnds{lf, cr, crlf};
    });
    LineEnds t;
    for (const LineEnds& x : part) t.lf += x.lf, t.cr += x.cr, t.crlf += x.crlf;
    return t;
}

// Nearly every LF preceded by CR: an LF to CR LF transfer, undone in place (rescue.cpp:2117-2133);
// the bytes freed at the end become zeros (in holds size + pad bytes)
inline bool undo_crlf(uint8_t* in, uint64_t& size, uint64_t pad, const LineEnds& le) {
    if (le.lf < 16 || le.crlf < 0.98 * le.lf) return false;
    uint64_t o = 0;
    for (uint64_t i = 0; i < size; i++) {
        if (in[i] == '\r' && i + 1 < size && in[i + 1] == '\n') continue;
        in[o++] = in[i];
    }
    memset(in + o, 0, size + pad - o);
    size = o;
    return true;
}

// Almost no CR LF where chance would put dozens: a CR LF to LF transfer (rescue.cpp:2141-2153);
// size: the input's, less a zero tail (zeros change none of the counts)
inline bool lost_crs_likely(const LineEnds& le, uint64_t size) {
    const double chance = size ? (double)le.lf * (double)le.cr / (double)size : 0;
    return chance >= 8.0 && 16.0 * (double)le.crlf <= chance;
}

// ------------------------------------------------------------------ encoder habits

// How the stream's encoder ends blocks (rescue.cpp:85-97)
struct Habit {
    enum Kind { NONE, GNU, ZLIB, ONE } kind = NONE;
    uint32_t regular = 0;
    bool fits(uint32_t symbols, uint32_t blocks) const {
        switch (kind) {
            case GNU: return symbols == 32767 || (symbols < 32768 && symbols % 4096 == 0);
            case ZLIB: return symbols == regular || blocks > 1;
            case ONE: return symbols == regular;
            default: return true;
        }
    }
};

// ------------------------------------------------------------------ a rescue session

constexpr uint64_t DELAY = gz::MAX_SYMBOL_OUT + Occ::AFTER;  // bytes of a piece settled only at its end

template <class B>
struct Session {
    B& be;
    uint8_t* in;  // the prepared input (a bit is flipped in it and put back by one_flip_mends)
    uint64_t size, nbits;
    uint64_t max_output = 0;
    bool drop_alike = false;
    std::vector<Start> headers;
    std::vector<Decoded> from_header;  // the piece decoded from each header, landing on the later ones
    std::map<std::tuple<uint64_t, uint64_t, int, int, int, size_t>, Decoded> others;
    seqd::Inflater dec;
    std::vector<uint16_t> scratch;

    Session(B& backend, uint8_t* input, uint64_t n)
        : be(backend), in(input), size(n), nbits(n * 8), scratch(SCRATCH + 2 * gz::MAX_SYMBOL_OUT) {}

    // one run of the host decoder (at most `cap` symbols)
    Outcome host_run(const Start& s, const Start* tg, int nt, bool first, uint64_t stop, uint32_t cap = 0) {
        return dec.run(in, size, s, tg, nt, scratch.data(), cap ? cap : (uint32_t)scratch.size(), first, stop);
    }

    // The pieces' decodes: every header's in bulk (the backend's), any other start when first met
    void decode_all_headers() { be.decode_headers(headers, from_header); }
    const Decoded& decode_at(const Start& s, bool stream_start, size_t t0) {
        if (!stream_start && s.mode == gz::AT_BLOCK && t0 > 0 && t0 - 1 < headers.size() && headers[t0 - 1].bit == s.bit)
            return from_header[t0 - 1];
        auto key = std::make_tuple(s.bit, s.header, (int)s.mode, (int)s.final, stream_start ? 1 : 0, t0);
        auto it = others.find(key);
        if (it != others.end()) return it->second;
        Decoded d;
        be.decode_other(s, stream_start, t0, headers, d);
        return others.emplace(key, d).first->second;
    }

    // Inside a fixed-Huffman block the codes resynchronise: a symbol boundary after e from which
    // the decode is believed (rescue.cpp:630-653)
    Start fixed_restart(uint64_t e, uint64_t limit, bool gzip) {
        limit = std::min(limit, std::min(nbits, e + FIXED_SEARCH));
        for (uint64_t from = e + 1; from + 4096 * 8 < limit; from += 2048 * 8) {
            const Start p = dec.probe(in, size, gz::IN_FIXED, UINT64_MAX, from, from + 2048 * 8);
            if (p.bit != UINT64_MAX && p.bit < limit && fixed_believed(p, gzip)) return p;
        }
        return no_start();
    }
    bool fixed_believed(const Start& p, bool gzip) {
        const Outcome r = host_run(p, nullptr, 0, false, UINT64_MAX, (uint32_t)VERIFY_OUT);
        if (r.status == gz::ERR_OVERFLOW) return r.blocks >= VERIFY_BLOCKS || r.out_len >= VERIFY_OUT - gz::MAX_SYMBOL_OUT;
        if (r.status != gz::OK) return false;
        const uint64_t t = (r.end_bit + 7) / 8, next = gzip ? t + 8 : t;
        return next == size || (gzip && next < size && member_header(in, size, next));
    }

    // The stored block of piece p's decode whose data holds `bit`: its data's input bytes
    // [first, second), or {0, 0}; with_len: its length fields count too (rescue.cpp:879-900)
    std::pair<uint64_t, uint64_t> stored_around(const Piece& p, uint64_t bit, bool with_len = false) {
        if (p.start.bit == UINT64_MAX || bit <= p.start.bit) return {0, 0};
        Start st = p.start;
        bool first = p.stream_start;
        for (;;) {
            const Outcome q = host_run(st, nullptr, 0, first, bit);
            first = false;
            if (q.status == gz::ERR_OVERFLOW) {
                st = q.resume;
                continue;
            }
            const uint64_t h = q.last_block_bit;
            if (q.status != gz::STOPPED || h == UINT64_MAX || ((bits_at(in, h) >> 1) & 3) != 0) return {0, 0};
            const uint64_t at = (h + 3 + 7) / 8, from = at + 4, to = from + (in[at] | (uint64_t)in[at + 1] << 8);
            if (bit < (with_len ? h + 1 : from * 8) || bit >= to * 8 || q.end_bit != to * 8) return {0, 0};
            return {from, to};
        }
    }

    // Does a block of piece p's decode begin at bit? (rescue.cpp:902-916)
    bool block_start_of(const Piece& p, uint64_t bit) {
        if (p.start.bit == UINT64_MAX || bit <= p.start.bit) return bit == p.start.bit;
        Start st = p.start;
        bool first = p.stream_start;
        for (;;) {
            const Outcome q = host_run(st, nullptr, 0, first, bit);
            first = false;
            if (q.status == gz::ERR_OVERFLOW) {
                st = q.resume;
                continue;
            }
            return q.status == gz::STOPPED && q.resume.mode == gz::AT_BLOCK && q.resume.bit == bit;
        }
    }

    // From headers[s], landing header to header: 1 the chain reaches headers[j]; 0 it ends
    // cleanly before (an embedded stream's headers), or what it reaches merges with p's own
    // stored blocks; -1 it breaks (damage); -2 it breaks inside the stored block of p's decode
    // where it started (rescue.cpp:918-963)
    int chain_outcome(size_t s, size_t j, const Piece& p) {
        auto skip_to = [&](uint64_t bit) {
            bit = std::max(bit, headers[s].bit + 1);
            while (s < j && headers[s].bit < bit) s++;
        };
        while (s < j) {
            Start st = headers[s];
            for (;;) {
                const Outcome q = host_run(st, headers.data() + s + 1, (int)(j - s), false, UINT64_MAX);
                if (q.status == gz::ERR_OVERFLOW) {
                    st = q.resume;
                    continue;
                }
                if (q.status == gz::OK && q.landed >= 0) {
                    const size_t began = s;
                    s = s + 1 + (size_t)q.landed;
                    if (s < j) break;
                    const bool merged = stored_around(p, headers[began].bit).second && q.last_block_bit != UINT64_MAX &&
                                        block_start_of(p, q.last_block_bit);
                    return merged ? 0 : 1;
                }
                if (q.status == gz::OK) {
                    skip_to(q.end_bit);
                    break;
                }
                const auto st_data = stored_around(p, headers[s].bit);
                if (st_data.second == 0) return -1;
                if (q.end_bit + MAX_SYMBOL_BITS < st_data.second * 8) return -2;
                skip_to(std::max(q.end_bit, st_data.second * 8));
                break;
            }
        }
        return 0;
    }

    // literals and matches of the block at bit, to its end-of-block code or failure
    uint64_t block_symbols(uint64_t bit) {
        Start st = at_block(bit);
        uint64_t n = 0;
        for (;;) {
            const Outcome q = host_run(st, nullptr, 0, false, UINT64_MAX);
            n += q.symbols;
            if (q.status != gz::ERR_OVERFLOW || q.eob_bit != UINT64_MAX) return n;
            st = q.resume;
        }
    }

    // the first run of RUN or more 0x00 or 0xFF bytes that starts in [from, to)
    uint64_t first_run(uint64_t from, uint64_t to, uint64_t& len, uint8_t& value) const {
        to = std::min(to, size);
        for (uint64_t i = from; i < to;) {
            const uint8_t v = in[i];
            uint64_t j = i + 1;
            if (v == 0 || v == 0xff) {
                while (j < size && in[j] == v) j++;
                if (j - i >= RUN) {
                    len = j - i;
                    value = v;
                    return i;
                }
            }
            i = j;
        }
        return UINT64_MAX;
    }

    // what p's decode produced from bits before `damage`, every symbol read wholly before it:
    // {output bytes, symbols} (rescue.cpp:1056-1080)
    std::pair<uint64_t, uint64_t> output_before(const Piece& p, uint64_t damage) {
        if (p.start.bit == UINT64_MAX || damage < p.start.bit + MAX_SYMBOL_BITS) return {0, 0};
        const uint64_t stop = damage - MAX_SYMBOL_BITS;
        Start st = p.start;
        bool first = p.stream_start;
        uint64_t out = 0, syms = 0;
        for (;;) {
            const Outcome q = host_run(st, nullptr, 0, first, stop);
            first = false;
            if (q.status == gz::ERR_OVERFLOW) {
                out += q.out_len;
                syms += q.symbols;
                st = q.resume;
                continue;
            }
            if (q.status != gz::STOPPED) return {0, 0};
            // a stored block is copied whole: one begun before the stop may reach past the damage
            if (q.end_bit > damage && q.last_block_bit != UINT64_MAX && q.out_len > q.last_block_out)
                return {out + q.last_block_out, syms + q.symbols};
            return {out + q.out_len, syms + q.symbols};
        }
    }

    // piece t takes p's bytes from offset at on
    static void split_bytes(Piece& p, Piece& t, uint64_t at) {
        t.src = p.src;
        t.src_off = p.src_off + at;
        t.len = p.len - at;
        t.all_exact = p.all_exact;
        p.len = at;
    }

    // ------------------------------------------------------------------ one pass

    // A pass over the file (rescue.cpp:1266-1463), on the decodes' results alone: the pieces,
    // which are damaged, what window each decodes against. `forced`: pieces (by start bit) known
    // damaged from an earlier pass. finish_pass() then resolves its bytes.
    PassResult walk(bool gzip, uint64_t hlen, const std::map<uint64_t, std::string>& forced, uint64_t expected_size,
                    bool has_crc, uint32_t expected_crc, std::vector<Span>& spans) {
        PassResult o;
        spans.clear();
        bool win_valid = false;
        uint64_t win_size = 0;
        uint32_t next_var = 1;
        bool stop_pending = false;  // the output passed max_output as the last piece settled
        bool stream_start = !gzip || hlen != 0;
        bool in_step = stream_start;
        if (gzip && !hlen) o.notes.push_back("the gzip header is damaged: decoding starts at the first block header found");
        Start cur = !gzip ? at_block(0) : hlen ? at_block(hlen * 8) : headers[0];
        o.members.push_back({});
        if (!gzip) {
            Member& m = o.members.back();
            m.trailer = expected_size != UINT64_MAX;
            m.isize = (uint32_t)expected_size;
            m.crc = expected_crc;
            m.has_crc = has_crc;
            m.size_known = expected_size != UINT64_MAX;
        }
        // the piece is settled: it joins the pass, and the window after it holds its last bytes or nothing
        auto commit = [&](Piece& p, Span sp, bool blank, bool window_after) {
            Source& s = o.sources.back();
            s.len = p.len;
            s.blank_tail = blank;
            s.window_after = window_after;
            sp.blank_tail = blank;
            spans.push_back(sp);
            if (max_output && o.total + p.len > max_output) stop_pending = true;
            o.total += p.len;
            o.pieces.push_back(std::move(p));
            if (window_after) {
                win_size = std::min<uint64_t>(gz::WINDOW, (win_valid ? win_size : 0) + o.pieces.back().len);
                win_valid = true;
            } else {
                win_valid = false;
                win_size = 0;
            }
        };
        size_t t0 = 0;
        while (cur.bit < nbits) {
            while (t0 < headers.size() && headers[t0].bit <= cur.bit) t0++;
            if (stop_pending) {  // the next decode finds the pass over its limit at once
                o.too_big = true;
                break;
            }
            const Decoded r = decode_at(cur, stream_start, t0);
            Piece p;
            p.in_bit_start = cur.bit;
            p.start = cur;
            p.stream_start = stream_start;
            p.src = o.sources.size();
            uint32_t var_base = 0;
            if (!win_valid && !stream_start && next_var < NO_VAR_LEFT) {
                var_base = next_var;
                next_var += gz::WINDOW + 1;
            }
            p.var_base = var_base;
            Source src;
            src.start = cur;
            src.stream_start = stream_start;
            src.t0 = t0;
            src.var_base = var_base;
            src.out = o.total;
            src.id = r.id;
            o.sources.push_back(src);
            if (max_output && r.len > DELAY && o.total + (r.len - DELAY) > max_output) {
                o.too_big = true;  // passed while decoding
                break;
            }
            const Span sp{r.id, o.total, r.len, win_valid, win_size, var_base, false};
            p.len = r.len;
            p.in_bit_end = r.end_bit;
            p.status = r.status;
            p.landed = r.status == gz::OK && r.landed >= 0;
            p.blocks = r.blocks;
            p.symbols = r.symbols;
            p.last_block_bit = r.last_block_bit;
            p.last_block_at = r.last_block_at;
            stream_start = false;
            p.anchored = in_step || (r.status == gz::OK && r.landed == 0);
            in_step = false;
            o.members.back().last_piece = o.pieces.size();
            const auto fb = forced.find(p.in_bit_start);
            const bool forced_bad = fb != forced.end();
            if (forced_bad) {
                p.suspect = true;
                p.why = fb->second;
            }
            const uint64_t t = (r.end_bit + 7) / 8, next = gzip ? t + 8 : t;
            const uint64_t nh = gzip && next < size ? member_header(in, size, next) : 0;
            // a final block is believed when the file ends after its trailer, another member
            // follows, no block header is left, or (raw) the output has the size expected
            const bool real_end = next >= size || nh || t0 >= headers.size() ||
                                  (!gzip && expected_size != UINT64_MAX && o.total + p.len == expected_size);
            const bool ok = r.status == gz::OK;
            if (ok && r.landed == 0 && !forced_bad) {  // in step: landed on the next header
                commit(p, sp, false, true);
                cur = headers[t0];
                in_step = true;
                continue;
            }
            if (ok && r.landed > 0 && !forced_bad) {
                // passed over headers: real ones missed through damage, or an embedded stream's
                const size_t j = t0 + (size_t)r.landed;
                int chain = chain_outcome(t0, j, p);
                if (chain == 1) {
                    bool all_stored = true;  // headers inside stored blocks the decode copied are the data's
                    for (size_t q = t0; q < j && all_stored; q++) all_stored = stored_around(p, headers[q].bit, true).second != 0;
                    if (all_stored) chain = 0;
                }
                if (chain == 0) {
                    commit(p, sp, false, true);
                    cur = headers[j];
                    t0 = j;
                    in_step = true;
                    continue;
                }
                if (chain == -2) {
                    p.suspect = p.length_exact = true;
                    p.why = "a Deflate stream embedded in its stored data breaks inside it: the stored bytes are damaged";
                    commit(p, sp, false, false);
                    cur = headers[j];
                    t0 = j;
                    in_step = true;
                    continue;
                }
                p.why = "its decode passed over the next block header without landing on it";
            }
            if (ok && r.landed < 0 && real_end) {  // the member's last block
                Member& m = o.members.back();
                commit(p, sp, false, !forced_bad && !nh);
                if (gzip && t + 8 <= size) {
                    m.trailer = m.has_crc = m.size_known = true;
                    memcpy(&m.crc, in + t, 4);
                    memcpy(&m.isize, in + t + 4, 4);
                }
                if (o.input_end == 0) o.input_end = t;
                if (!gzip) break;
                if (nh) {
                    Member nm;
                    nm.first_piece = nm.last_piece = o.pieces.size();
                    o.members.push_back(nm);
                    cur = at_block((next + nh) * 8);
                    stream_start = in_step = true;
                    continue;
                }
                if (next < size) o.notes.push_back(std::to_string(size - next) + " bytes after the last member were not gzip data");
                break;
            }
            if (r.status == gz::ERR_INPUT && t0 >= headers.size()) {
                commit(p, sp, true, false);  // its last symbol may have read past the end
                o.notes.push_back("the file ends inside the stream (truncated)");
                o.truncated = true;
                break;
            }
            if (p.why.empty())
                p.why = ok && r.landed < 0              ? "a final block was read where the data goes on"
                        : r.status == gz::ERR_INPUT ? "the decode ran off the end of the file"
                                                    : "the decode hit invalid data";
            // damaged: restart at the next header, but at none inside a stored block this decode
            // copied before it broke (those are the data's)
            if (!forced_bad || !(ok && r.landed > 0)) {
                uint64_t stored_to = 0;
                while (t0 < headers.size() && headers[t0].bit < r.end_bit) {
                    if (headers[t0].bit < stored_to) {
                        t0++;
                        continue;
                    }
                    const uint64_t to = stored_around(p, headers[t0].bit, true).second;
                    if (to == 0) break;
                    stored_to = to * 8;
                    t0++;
                }
            }
            p.suspect = true;
            commit(p, sp, false, false);
            if (forced_bad && ok && r.landed > 0) t0 += (size_t)r.landed;
            if (t0 < headers.size()) {
                cur = headers[t0];
                continue;
            }
            const Start fx = fixed_restart(r.end_bit, nbits, gzip);
            if (fx.bit == UINT64_MAX) {
                o.notes.push_back("no block header after byte " + std::to_string(r.end_bit / 8) + ": the rest is lost");
                break;
            }
            o.notes.push_back("resynchronized inside a fixed-Huffman block at byte " + std::to_string(fx.bit / 8));
            cur = fx;
        }
        return o;
    }

    // The pass's bytes resolved (the backend), and what the decisions read off them: each
    // piece's exactness, each member's CRC, the ranges of bytes with variables
    void finish_pass(PassResult& o, const std::vector<Span>& spans, PhaseClock& clock) {
        be.resolve(spans, o.total, o.prov);
        clock.mark("resolve");
        // per piece, over its decode's first len bytes: all exact; the first variable at a multiple
        // of 64; the suspect and the last pieces' ranges with variables. Per member, the CRC-32
        // of its output.
        const size_t np = o.pieces.size();
        const size_t last_k = np ? np - 1 : SIZE_MAX;
        std::vector<PieceRange> ranges(np);
        for (size_t k = 0; k < np; k++)
            ranges[k] = PieceRange{o.sources[o.pieces[k].src].out, o.pieces[k].len, o.pieces[k].suspect || k == last_k};
        std::vector<std::pair<uint64_t, uint64_t>> crc_ranges;
        std::vector<Member*> crc_members;
        for (Member& m : o.members) {
            if (m.first_piece >= np) continue;
            const size_t last = std::min(m.last_piece, np - 1);
            crc_ranges.push_back({o.sources[o.pieces[m.first_piece].src].out,
                                  o.sources[o.pieces[last].src].out + o.pieces[last].len});
            crc_members.push_back(&m);
        }
        std::vector<PieceFacts> facts;
        std::vector<uint32_t> crcs;
        be.summarize(o.prov, ranges, facts, crc_ranges, crcs);
        // a decode's lost64 comes from the first of its pieces that has one; its ranges from the last kept
        for (size_t k = 0; k < np; k++) {
            Piece& p = o.pieces[k];
            Source& s = o.sources[p.src];
            p.all_exact = facts[k].all_exact;
            if (s.lost64 == UINT64_MAX && facts[k].var64 < p.len) s.lost64 = facts[k].var64;
            if (ranges[k].keep) o.runs[p.src] = std::move(facts[k].runs);
        }
        for (size_t i = 0; i < crc_members.size(); i++) crc_members[i]->out_crc = crcs[i];
        clock.mark("pass summaries");
    }

    // ------------------------------------------------------------------ after a pass

    // A member decoded as one piece, in step, whose CRC or length fails was damaged where the
    // decoder fell back in step: flagged from the first run of zeros or 0xFF bytes in its data,
    // else whole (rescue.cpp:1082-1118). Returns the pieces to flag, by start bit, and why.
    std::map<uint64_t, std::string> crc_misses(const PassResult& o, std::vector<std::string>& notes) {
        std::map<uint64_t, std::string> out;
        for (const Member& m : o.members) {
            if (!m.trailer || m.first_piece != m.last_piece || m.last_piece >= o.pieces.size() ||
                o.pieces[m.last_piece].status != gz::OK)
                continue;
            const Piece& p = o.pieces[m.last_piece];
            if (p.suspect || !p.all_exact) continue;
            if (!((m.has_crc && m.out_crc != m.crc) || (m.size_known && (uint32_t)p.len != m.isize))) continue;
            const std::string what =
                std::string("the stream's ") + (m.has_crc ? "CRC" : "length") + " does not match what was decoded";
            uint64_t n = 0;
            uint8_t v = 0;
            const uint64_t at = first_run(p.in_bit_start / 8, (p.in_bit_end + 7) / 8, n, v);
            if (at != UINT64_MAX) {
                const std::string run = "a run of " + std::to_string(n) + (v ? " 0xFF" : " zero") + " bytes at input byte " +
                                        std::to_string(at);
                out[p.in_bit_start] = what + ", and " + run + " lies in this block: damage that decodes as valid codes";
                notes.push_back(what + ", though it decoded in step: the damage is placed where the " +
                                (v ? "0xFF bytes" : "zeros") +
                                " begin, and the output from there on is flagged. The bytes before them decoded cleanly; with the "
                                "CRC failing, nothing more confirms them");
            } else {
                out[p.in_bit_start] = what + ", and nothing shows where in this, the stream's only piece, the damage lies";
                notes.push_back(what + ", though it decoded in step, and nothing shows where the damage lies: decoded in one piece (" +
                                std::to_string(p.blocks) + " block" + (p.blocks == 1 ? "" : "s") + "), all of it is flagged");
            }
        }
        return out;
    }

    // Does one flipped bit in the FLIP_BITS before where piece p broke mend its block: a decode
    // that lands on the next header with a count the stream's habit allows? (rescue.cpp:1120-1183)
    bool one_flip_mends(const Piece& p, const Habit& habit) {
        if (p.start.mode != gz::AT_BLOCK || p.start.bit == UINT64_MAX || p.status != gz::ERR_DATA) return false;
        const size_t t = (size_t)(std::partition_point(headers.begin(), headers.end(),
                                                       [&](const Start& h) { return h.bit <= p.start.bit; }) -
                                  headers.begin());
        if (t == headers.size()) return false;
        const Start* targets = headers.data() + t;
        const int ntargets = (int)(headers.size() - t);
        const uint64_t e = std::min(p.in_bit_end, nbits), lo = e > p.start.bit + FLIP_BITS ? e - FLIP_BITS : p.start.bit;
        // symbol boundaries from lo on, about every 64 bits, with the symbols and blocks before
        // each: a trial decodes again from the last one before its flipped bit
        struct Snap {
            Start at;
            uint32_t symbols, blocks;
        };
        std::vector<Snap> snaps{{p.start, 0, 0}};
        Start st = p.start;
        uint32_t syms = 0, blocks = 0;
        bool first = p.stream_start;
        for (uint64_t stop = lo; stop < e;) {
            const Outcome q = host_run(st, nullptr, 0, first, stop);
            first = false;
            syms += q.symbols;
            blocks += q.blocks;
            if (q.status == gz::ERR_OVERFLOW) {
                st = q.resume;
                continue;
            }
            if (q.status != gz::STOPPED) break;
            if (q.resume.bit > snaps.back().at.bit) snaps.push_back({q.resume, syms, blocks});
            st = q.resume;
            stop = q.resume.bit + 64;
        }
        for (uint64_t f = e; f-- > lo;) {
            size_t k = snaps.size() - 1;
            while (k > 0 && snaps[k].at.bit > f) k--;
            in[f >> 3] ^= (uint8_t)(1u << (f & 7));
            Start s2 = snaps[k].at;
            uint32_t sy = snaps[k].symbols, bl = snaps[k].blocks;
            bool from_start = k == 0 && p.stream_start, mended = false;
            for (;;) {
                const Outcome q = host_run(s2, targets, ntargets, from_start, UINT64_MAX);
                from_start = false;
                sy += q.symbols;
                bl += q.blocks;
                if (q.status == gz::ERR_OVERFLOW) {
                    s2 = q.resume;
                    continue;
                }
                mended = q.status == gz::OK && q.landed == 0 && habit.fits(sy, bl);
                break;
            }
            in[f >> 3] ^= (uint8_t)(1u << (f & 7));
            if (mended) return true;
        }
        return false;
    }

    // A truncated decode whose last block holds more symbols than the stream's blocks do ran
    // through damage to the end: that block is suspect (rescue.cpp:998-1034)
    void suspect_overfull_tail(PassResult& o, uint32_t regular) {
        if (!regular || !o.truncated || o.pieces.empty()) return;
        Piece& p = o.pieces.back();
        if (p.suspect || p.status != gz::ERR_INPUT || p.last_block_bit == UINT64_MAX || p.last_block_at > p.len) return;
        const uint64_t n = block_symbols(p.last_block_bit);
        if (n <= regular) return;
        Piece t;
        split_bytes(p, t, p.last_block_at);
        t.in_bit_start = p.last_block_bit;
        t.start = at_block(p.last_block_bit);
        t.anchored = p.anchored;
        t.in_bit_end = p.in_bit_end;
        t.status = p.status;
        t.blocks = 1;
        t.symbols = (uint32_t)std::min<uint64_t>(n, UINT32_MAX);
        t.suspect = true;
        t.why = "its block holds " + std::to_string(n) + " symbols, more than the stream's regular " + std::to_string(regular) +
                ": the decode ran through damaged data to the end of the file";
        p.in_bit_end = p.last_block_bit;
        p.status = gz::OK;
        p.blocks = p.blocks ? p.blocks - 1 : 0;
        p.symbols = p.symbols > t.symbols ? p.symbols - t.symbols : 0;
        if (p.len == 0) {
            t.stream_start = p.stream_start;
            p = std::move(t);
        } else {
            o.pieces.push_back(std::move(t));
        }
        o.members.back().last_piece = o.pieces.size() - 1;
        for (std::string& note : o.notes)
            if (note == "the file ends inside the stream (truncated)")
                note = "the decode ran to the end of the file through damaged data: its last block holds more symbols than the "
                       "stream's blocks do";
    }

    // A damaged piece decoded from a start in step up to a run of zeros or 0xFF bytes, breaking
    // only in or after it, is exact up to the run and suspect from it on: split there; a truncated
    // decode that ran through such a run is suspect from it on (rescue.cpp:1185-1260)
    void localize(PassResult& o) {
        std::vector<Piece> out;
        std::vector<size_t> head_at(o.pieces.size()), tail_at(o.pieces.size());
        for (size_t k = 0; k < o.pieces.size(); k++) {
            Piece& p = o.pieces[k];
            head_at[k] = out.size();
            const bool runs_out = !p.suspect && o.truncated && k + 1 == o.pieces.size() && p.status == gz::ERR_INPUT;
            uint64_t n = 0, at = UINT64_MAX;
            uint8_t v = 0;
            if (p.suspect && p.anchored && p.start.mode == gz::AT_BLOCK) {
                // the damage lies before where the decode broke (its last symbol read up to
                // MAX_SYMBOL_BITS on) or landed, and before any header it passed over that is not
                // inside one of its stored blocks
                uint64_t limit = p.in_bit_end + (p.status == gz::OK ? 0 : MAX_SYMBOL_BITS);
                size_t h = (size_t)(std::partition_point(headers.begin(), headers.end(),
                                                         [&](const Start& x) { return x.bit <= p.in_bit_start; }) -
                                    headers.begin());
                while (h < headers.size() && headers[h].bit < p.in_bit_end && stored_around(p, headers[h].bit).second) h++;
                if (h < headers.size() && headers[h].bit < p.in_bit_end) limit = std::min(limit, headers[h].bit);
                at = first_run(p.in_bit_start / 8, (limit + 7) / 8, n, v);
            } else if (runs_out) {
                at = first_run(p.in_bit_start / 8, size, n, v);
            }
            std::pair<uint64_t, uint64_t> before = at == UINT64_MAX ? std::pair<uint64_t, uint64_t>{0, 0} : output_before(p, at * 8);
            const uint64_t nb = std::min<uint64_t>(before.first, p.len), ns = before.second;
            if (at == UINT64_MAX || (nb == 0 && !runs_out)) {
                out.push_back(std::move(p));
                tail_at[k] = out.size() - 1;
                continue;
            }
            const std::string run =
                "a run of " + std::to_string(n) + (v ? " 0xFF" : " zero") + " bytes at input byte " + std::to_string(at);
            Piece t;
            split_bytes(p, t, nb);
            t.in_bit_start = at * 8;
            t.in_bit_end = p.in_bit_end;
            t.suspect = true;
            t.by_crc = p.by_crc;
            t.length_exact = p.length_exact;
            t.status = p.status;
            t.landed = p.landed;
            t.blocks = 1;
            t.symbols = (uint32_t)(p.symbols > ns ? p.symbols - ns : 0);
            if (runs_out)
                t.why = "the decode ran through " + run +
                        " and on to the end of the file: zeros where compressed data should be decode as valid codes, so nothing "
                        "from them on is trusted";
            else if (p.by_crc)
                t.why = p.why;
            else
                t.why = p.why + "; the damage begins at " + run + ": the " + std::to_string(nb) +
                        " bytes decoded before it are exact";
            if (nb > 0) {
                p.in_bit_end = at * 8;
                p.suspect = p.by_crc = p.landed = false;
                p.why.clear();
                p.status = gz::STOPPED;
                p.symbols = (uint32_t)ns;
                out.push_back(std::move(p));
            } else {
                t.stream_start = p.stream_start;
            }
            out.push_back(std::move(t));
            tail_at[k] = out.size() - 1;
            if (runs_out)
                for (std::string& note : o.notes)
                    if (note == "the file ends inside the stream (truncated)")
                        note = "the decode ran through " + run + " to the end of the file: from there on the output is flagged";
        }
        for (Member& m : o.members) {
            if (m.first_piece < head_at.size()) m.first_piece = head_at[m.first_piece];
            if (m.last_piece < tail_at.size()) m.last_piece = tail_at[m.last_piece];
        }
        o.pieces.swap(out);
    }

    // a kept place of the variable that agrees with oc on the 3 cells either side (rescue.cpp:361-375)
    static bool alike(const Vars::Var& v, const Occ& o) {
        for (const Occ& e : v.places)
            if (e.cell[0] == o.cell[0] && e.cell[1] == o.cell[1] && e.cell[2] == o.cell[2] && e.cell[5] == o.cell[5] &&
                e.cell[6] == o.cell[6] && e.cell[7] == o.cell[7])
                return true;
        return false;
    }
};

// A member's CRC and length over what a pass decoded: 1 right, 0 wrong, -1 not checkable
// (rescue.cpp:1466-1479)
inline int member_crc(const PassResult& o, const Member& m) {
    if (!m.trailer || m.last_piece >= o.pieces.size() || o.pieces[m.last_piece].status != gz::OK) return -1;
    uint64_t len = 0;
    for (size_t k = m.first_piece; k <= m.last_piece; k++) {
        const Piece& p = o.pieces[k];
        if (p.suspect || !p.all_exact) return -1;
        len += p.len;
    }
    const bool size_ok = !m.size_known || (uint32_t)len == m.isize;
    if (m.has_crc) return m.out_crc == m.crc && size_ok;
    return m.size_known && !size_ok ? 0 : -1;
}

// Pieces of one block, in step, whose symbol count does not fit the stream's encoder habit were
// damaged in a way landing cannot see (rescue.cpp:1481-1587). `whole`: pieces of members whose
// CRC checked, never flagged. Returns their start bits; `habit` the one found.
inline std::set<uint64_t> count_anomalies(const PassResult& o, const std::vector<char>& whole, bool cr_stripped,
                                          std::string& note, Habit& habit) {
    habit = Habit{};
    const size_t n = o.pieces.size();
    std::vector<char> last(n, 0), chunk_end(n, 0);
    std::vector<uint64_t> end(n, 0);  // where each piece's output ends, in its member's
    for (const Member& m : o.members) {
        if (m.last_piece < n) last[m.last_piece] = 1;
        uint64_t at = 0;
        for (size_t k = m.first_piece; k <= m.last_piece && k < n; k++) end[k] = at += o.pieces[k].len;
    }
    auto eligible = [&](size_t k) {
        const Piece& p = o.pieces[k];
        return p.blocks == 1 && p.landed && !last[k] && !p.suspect;
    };
    std::map<uint32_t, int> freq;
    int considered = 0;
    for (size_t k = 0; k < n; k++)
        if (eligible(k)) {
            freq[o.pieces[k].symbols]++;
            considered++;
        }
    std::set<uint64_t> out;
    if (considered < 10) return out;
    // pigz's chunk: the commonest output distance between marked chunk ends (pieces of several
    // blocks), when a multiple of 1 KiB that half the distances are
    std::map<uint64_t, int> gaps;
    int ngaps = 0;
    for (const Member& m : o.members) {
        uint64_t prev = 0;
        for (size_t k = m.first_piece; k <= m.last_piece && k < n; k++) {
            if (o.pieces[k].suspect) prev = UINT64_MAX;
            if (o.pieces[k].suspect || o.pieces[k].blocks < 2 || last[k]) continue;
            if (prev != UINT64_MAX) {
                gaps[end[k] - prev]++;
                ngaps++;
            }
            prev = end[k];
        }
    }
    uint64_t chunk = 0;
    if (!gaps.empty()) {
        auto g = std::max_element(gaps.begin(), gaps.end(), [](const auto& a, const auto& b) { return a.second < b.second; });
        if (g->first >= 32768 && g->first % 1024 == 0 && 2 * g->second >= ngaps) chunk = g->first;
    }
    if (chunk)
        for (const Member& m : o.members) {
            uint64_t base = 0;
            for (size_t k = m.first_piece; k <= m.last_piece && k < n; k++) {
                const Piece& p = o.pieces[k];
                if (p.suspect)
                    base = UINT64_MAX;
                else if (base != UINT64_MAX && (end[k] - base) % chunk == 0)
                    chunk_end[k] = 1;
                if (!p.suspect && p.blocks >= 2) base = end[k];
            }
        }
    auto gnu = [](uint32_t x) { return x == 32767 || (x < 32768 && x % 4096 == 0); };
    int gnu_fit = 0, zlib_fit = 0;
    for (const auto& [x, c] : freq) gnu_fit += gnu(x) ? c : 0;
    uint32_t zlib = 0;
    for (int b = 7; b <= 15; b++) {
        auto f = freq.find((1u << b) - 1);
        if (f != freq.end() && f->second > zlib_fit) {
            zlib = f->first;
            zlib_fit = f->second;
        }
    }
    int ends = 0;
    for (size_t k = 0; k < n; k++) ends += eligible(k) && o.pieces[k].symbols != zlib && chunk_end[k];
    if (zlib) zlib_fit += ends;
    auto best = std::max_element(freq.begin(), freq.end(), [](const auto& a, const auto& b) { return a.second < b.second; });
    const int need = cr_stripped ? (considered + 3) / 4 : (considered + 1) / 2;
    if (freq.count(32767) && gnu_fit >= need && gnu_fit >= zlib_fit)
        habit = {Habit::GNU, 32767};
    else if (zlib && zlib_fit >= need)
        habit = {Habit::ZLIB, zlib};
    else if (best->second >= considered * 7 / 10)
        habit = {Habit::ONE, best->first};
    if (habit.kind == Habit::NONE) return out;
    const int fit = habit.kind == Habit::GNU ? gnu_fit : habit.kind == Habit::ZLIB ? zlib_fit : best->second;
    auto fits = [&](size_t k) { return habit.fits(o.pieces[k].symbols, 1) || (habit.kind == Habit::ZLIB && chunk_end[k]); };
    for (size_t k = 0; k < n; k++)
        if (eligible(k) && !whole[k] && !fits(k)) out.insert(o.pieces[k].in_bit_start);
    if (!out.empty())
        note = std::to_string(out.size()) + " block(s) with a symbol count " +
               (habit.kind == Habit::GNU
                    ? std::string("GNU gzip never writes (it ends a block at 32767 symbols or a multiple of 4096: ")
                    : "other than the stream's regular " + std::to_string(habit.regular) +
                          (habit.kind == Habit::ZLIB && chunk
                               ? " that do not end one of pigz's " + std::to_string(chunk >> 10) + " KiB chunks"
                               : std::string()) +
                          " (") +
               std::to_string(fit) + " of " + std::to_string(considered) + " blocks fit that)";
    return out;
}

// ------------------------------------------------------------------ the lost windows' bytes

// fn(i) for i in [0, n) on the host's cores

// A model of the file's exact bytes: for contexts of 1 to 5 bytes, how often each byte
// followed, blended from the longest context seen down (rescue.cpp:1592-1733). Its counts are
// kept in the reference's table (the same keys, slots and 64-probe limit), since a count a full
// probe sequence drops changes what is inferred.
struct ByteModel {
    static constexpr int ORDERS = 5;
    static constexpr uint32_t BITS = 23;
    std::vector<uint64_t> keys;
    std::vector<uint32_t> vals;
    uint32_t uni[256] = {};
    uint64_t n = 0;

    static uint64_t key(uint64_t ctx, int k, int c) {
        uint64_t x = ctx & ((1ull << (8 * k)) - 1);
        x = x * 0x9E3779B97F4A7C15ull + ((uint64_t)k << 52) + (uint64_t)(c + 1) * 0xC2B2AE3D27D4EB4Full;
        x ^= x >> 31;
        x *= 0xBF58476D1CE4E5B9ull;
        x ^= x >> 29;
        return x | 1;  // 0 marks an empty slot
    }
    void add(uint64_t k) {
        const uint32_t mask = (1u << BITS) - 1;
        uint32_t i = (uint32_t)(k >> 20) & mask;
        for (int probe = 0; probe < 64; probe++, i = (i + 1) & mask) {
            if (keys[i] == k) {
                vals[i]++;
                return;
            }
            if (!keys[i]) {
                keys[i] = k;
                vals[i] = 1;
                return;
            }
        }
    }
    uint32_t get(uint64_t k) const {
        const uint32_t mask = (1u << BITS) - 1;
        uint32_t i = (uint32_t)(k >> 20) & mask;
        for (int probe = 0; probe < 64; probe++, i = (i + 1) & mask) {
            if (keys[i] == k) return vals[i];
            if (!keys[i]) return 0;
        }
        return 0;
    }
    // the exact bytes (tag EXACT) of one piece, up to the budget, in stretches: ctx and run carry
    // the context from one stretch of the piece to the next (both 0 at the piece's start)
    void learn(const uint8_t* b, const uint32_t* tag, uint64_t len, uint64_t& budget, uint64_t& ctx, int& run) {
        for (uint64_t i = 0; i < len && budget; i++) {
            if (tag[i] != Provenance::EXACT) {
                run = 0;
                ctx = 0;
                continue;
            }
            const uint8_t x = b[i];
            uni[x]++;
            n++;
            budget--;
            for (int k = 1; k <= ORDERS && k <= run; k++) {
                add(key(ctx, k, x));
                add(key(ctx, k, -1));
            }
            ctx = ctx << 8 | x;
            run++;
        }
    }
    // log P(c | the bytes before), ctx newest lowest, `known` of them known
    float lp(uint64_t ctx, int known, int c) const {
        double pr = (uni[c] + 1.0) / (n + 256.0);
        for (int k = 1; k <= ORDERS && k <= known; k++) {
            const double t = get(key(ctx, k, -1));
            if (t == 0) break;  // a context never seen: no longer one was
            const double e = get(key(ctx, k, c));
            pr = (e + 1.5 * pr) / (t + 1.5);
        }
        return (float)std::log(pr);
    }
};

// The lost window bytes' values: for each variable, -1 (none) or its byte
struct Inferred {
    const Vars* vars = nullptr;
    std::vector<int16_t> value;  // by index in vars->list
    int get(uint32_t var) const {
        if (!vars) return -1;
        const uint32_t d = vars->find(var);
        return d == UINT32_MAX ? -1 : value[d];
    }
};

// Infers the bytes of lost windows (rescue.cpp:1735-2100): each variable's places give contexts
// of known bytes around it; a value is accepted when the contexts, scored by the model, favour it
// by MARGIN over the next, one of them by SOLO on its own, and none favours another by AGAINST.
// Accepted values become context for the next round. `done`: the output bytes they give a value.
// fetch(at, len, byte, tag): the provenance's bytes and tags of [at, at + len)
using Fetch = std::function<void(uint64_t, uint64_t, uint8_t*, uint32_t*)>;

inline Inferred infer_lost(PassResult& o, const Fetch& fetch, uint64_t& done, std::string& note) {
    constexpr float MARGIN = 14.0f, SOLO = 6.0f, AGAINST = 2.0f;
    constexpr int ROUNDS = 60, AHEAD = Occ::AFTER;
    done = 0;
    Inferred out;
    Vars& vt = o.vars;
    const uint32_t nv = (uint32_t)vt.list.size();
    if (!nv) return out;
    std::vector<Piece>& pieces = o.pieces;
    const size_t np = pieces.size();
    std::vector<uint64_t> start(np);
    for (size_t k = 0; k < np; k++) start[k] = o.sources[pieces[k].src].out + pieces[k].src_off;
    // a piece has lost bytes when one at a multiple of 64 from its start copies a variable
    std::vector<uint8_t> lost(np, 0);
    for (size_t k = 0; k < np; k++) {
        const Piece& p = pieces[k];
        const Source& s = o.sources[p.src];
        if (p.src_off == 0) {
            lost[k] = s.lost64 < p.len;
            continue;
        }
        auto it = o.runs.find(p.src);
        if (it == o.runs.end()) throw std::runtime_error("the ranges with variables of a split decode were not kept");
        const uint64_t a = p.src_off, e = a + p.len;
        for (const auto& [r0, r1] : it->second) {
            if (r1 <= a || r0 >= e) continue;
            const uint64_t x = a + (std::max(r0, a) - a + 63) / 64 * 64;
            if (x < std::min(r1, e)) {
                lost[k] = 1;
                break;
            }
        }
    }
    // cells of a place that lie outside its piece (pieces split since) are not known
    on_cores((nv + 4095) / 4096, [&](size_t b, unsigned) {
        for (uint32_t d = (uint32_t)(b * 4096); d < nv && d < (b + 1) * 4096; d++)
            for (Occ& oc : vt.list[d].places) {
                const uint64_t pos = oc.pos;
                const size_t k = (size_t)(std::upper_bound(start.begin(), start.end(), pos) - start.begin()) - 1;
                const uint64_t a = start[k], e = a + pieces[k].len;
                for (int q = 1; q <= Occ::BEFORE; q++)
                    if (pos < a + q) oc.cell[q - 1] = Occ::NONE;
                for (int q = 1; q <= Occ::AFTER; q++)
                    if (pos + q >= e) oc.cell[Occ::BEFORE + q - 1] = Occ::NONE;
            }
    });
    uint64_t unknown = 0;
    for (const Vars::Var& x : vt.list) unknown += x.count;
    // the model learns from the pieces nearest those with lost bytes, exact ones only
    std::vector<int64_t> dist(np, INT64_MAX);
    for (int64_t j = 0, last = -1; j < (int64_t)np; j++) {
        if (lost[j]) last = j;
        if (last >= 0) dist[j] = j - last;
    }
    for (int64_t j = (int64_t)np - 1, nxt = -1; j >= 0; j--) {
        if (lost[j]) nxt = j;
        if (nxt >= 0) dist[j] = std::min<int64_t>(dist[j], nxt - j);
    }
    std::vector<size_t> ord(np);
    for (size_t k = 0; k < np; k++) ord[k] = k;
    std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return dist[a] < dist[b]; });
    ByteModel m;
    m.keys.assign(1u << ByteModel::BITS, 0);
    m.vals.assign(1u << ByteModel::BITS, 0);
    uint64_t budget = TRAIN;
    {
        // a piece's bytes a stretch at a time: the budget is usually spent within its first one
        constexpr uint64_t STRETCH = 1u << 20;
        std::vector<uint8_t> b(STRETCH);
        std::vector<uint32_t> tg(STRETCH);
        for (size_t k : ord) {
            if (!budget) break;
            if (pieces[k].suspect) continue;
            uint64_t ctx = 0;
            int run = 0;
            for (uint64_t a = 0; a < pieces[k].len && budget; a += STRETCH) {
                const uint64_t n = std::min(STRETCH, pieces[k].len - a);
                fetch(start[k] + a, n, b.data(), tg.data());
                m.learn(b.data(), tg.data(), n, budget, ctx, run);
            }
        }
    }
    if (m.n < 4096) return out;
    out.vars = &vt;
    out.value.assign(nv, -1);
    // a cell as a context byte for variable self: -1 when not known (itself included)
    auto val = [&](uint32_t c, uint32_t self) -> int {
        if (c < Occ::VAR) return (int)c;
        if (c == Occ::NONE || c - Occ::VAR == self) return -1;
        const uint32_t d = vt.find(c - Occ::VAR);
        return d == UINT32_MAX ? -1 : out.value[d];
    };
    std::vector<uint32_t> todo(nv);
    for (uint32_t d = 0; d < nv; d++) todo[d] = d;
    std::sort(todo.begin(), todo.end(), [&](uint32_t a, uint32_t b) { return vt.list[a].var < vt.list[b].var; });
    // the variables each one lies beside other than its window neighbours v +- q (which are scored
    // again by number): wlist[wstart[u] .. wstart[u + 1]) are scored again when u is resolved
    auto nearby = [&](uint32_t d, std::vector<uint32_t>& near) {
        near.clear();
        const uint32_t v = vt.list[d].var;
        for (const Occ& oc : vt.list[d].places)
            for (int j = 0; j < Occ::CELLS; j++) {
                const uint32_t c = oc.cell[j];
                if (c == Occ::NONE || c < Occ::VAR || c - Occ::VAR == v) continue;
                const uint32_t q = j < Occ::BEFORE ? (uint32_t)j + 1 : (uint32_t)(Occ::BEFORE - 1 - j);
                if (c - Occ::VAR == v + q) continue;
                const uint32_t u = vt.find(c - Occ::VAR);
                if (u != UINT32_MAX) near.push_back(u);
            }
        std::sort(near.begin(), near.end());
        near.erase(std::unique(near.begin(), near.end()), near.end());
    };
    std::vector<uint32_t> wstart(nv + 1, 0), wlist;
    {
        std::vector<std::vector<uint32_t>> nears(nv);
        on_cores((nv + 4095) / 4096, [&](size_t b, unsigned) {
            for (uint32_t d = (uint32_t)(b * 4096); d < nv && d < (b + 1) * 4096; d++) nearby(d, nears[d]);
        });
        for (uint32_t d = 0; d < nv; d++)
            for (uint32_t u : nears[d]) wstart[u + 1]++;
        for (uint32_t d = 0; d < nv; d++) wstart[d + 1] += wstart[d];
        wlist.resize(wstart[nv]);
        std::vector<uint32_t> at(wstart.begin(), wstart.end() - 1);
        for (uint32_t d = 0; d < nv; d++)
            for (uint32_t u : nears[d]) wlist[at[u]++] = d;
    }
    std::vector<uint8_t> dirty(nv, 1);
    struct Ctx {
        uint64_t before;  // newest lowest
        int nbefore;
        int after[AHEAD];
        int nafter;
    };
    struct Work {
        std::vector<Ctx> ctxs;
        std::vector<float> score, per;
    };
    std::vector<Work> work(cores());
    auto judge = [&](uint32_t d, Work& w) -> int {
        const uint32_t v = vt.list[d].var;
        // the distinct contexts the byte appears in, with known bytes on both sides; those alike
        // on the 3 bytes either side are one piece of evidence
        std::vector<Ctx>& ctxs = w.ctxs;
        ctxs.clear();
        for (const Occ& oc : vt.list[d].places) {
            Ctx c{0, 0, {}, 0};
            for (int q = 1; q <= ByteModel::ORDERS; q++) {
                const int b = val(oc.cell[q - 1], v);
                if (b < 0) break;
                c.before |= (uint64_t)b << (8 * (q - 1));
                c.nbefore = q;
            }
            for (int q = 1; q <= AHEAD; q++) {
                const int b = val(oc.cell[Occ::BEFORE + q - 1], v);
                if (b < 0) break;
                c.after[q - 1] = b;
                c.nafter = q;
            }
            if (c.nbefore < 2 || c.nafter < 2) continue;
            bool dup = false;
            for (const Ctx& e : ctxs) {
                const int nb = std::min(3, std::min(e.nbefore, c.nbefore)), na = std::min(3, std::min(e.nafter, c.nafter));
                if (std::min(3, e.nbefore) != std::min(3, c.nbefore) || std::min(3, e.nafter) != std::min(3, c.nafter)) continue;
                if (((e.before ^ c.before) & ((1ull << (8 * nb)) - 1)) == 0 && std::equal(e.after, e.after + na, c.after)) {
                    dup = true;
                    break;
                }
            }
            if (!dup) ctxs.push_back(c);
        }
        if (ctxs.empty()) return -1;
        std::vector<float>& score = w.score;
        std::vector<float>& per = w.per;
        score.assign(256, 0.0f);
        per.assign(ctxs.size() * 256, 0.0f);
        for (int x = 0; x < 256; x++) {
            float sc = 0;
            for (size_t ci = 0; ci < ctxs.size(); ci++) {
                const Ctx& c = ctxs[ci];
                uint64_t ctx = c.before;
                int kn = c.nbefore;
                float s1 = m.lp(ctx, kn, x);
                ctx = ctx << 8 | (uint64_t)x;
                kn++;
                for (int q = 0; q < c.nafter; q++) {
                    s1 += m.lp(ctx, kn, c.after[q]);
                    ctx = ctx << 8 | (uint64_t)c.after[q];
                    kn++;
                }
                per[ci * 256 + x] = s1;
                sc += s1;
            }
            score[x] = sc;
        }
        float best = -1e30f, second = -1e30f;
        int bc = 0;
        for (int x = 0; x < 256; x++) {
            if (score[x] > best) {
                second = best;
                best = score[x];
                bc = x;
            } else if (score[x] > second) {
                second = score[x];
            }
        }
        bool favoured = false, against = false;
        for (size_t ci = 0; ci < ctxs.size(); ci++) {
            const float* row = &per[ci * 256];
            float other = -1e30f;
            for (int x = 0; x < 256; x++)
                if (x != bc) other = std::max(other, row[x]);
            favoured |= row[bc] - other >= SOLO;
            against |= other - row[bc] >= AGAINST;
        }
        return favoured && !against && best - second >= MARGIN ? bc : -1;
    };
    for (int round = 0; round < ROUNDS && !todo.empty(); round++) {
        std::vector<uint32_t> scored;
        for (uint32_t d : todo)
            if (dirty[d]) scored.push_back(d);
        std::vector<int16_t> verdict(scored.size());
        on_cores(scored.size(), [&](size_t i, unsigned w) { verdict[i] = (int16_t)judge(scored[i], work[w]); });
        // accepted in the variables' order, as one pass over them would
        std::vector<std::pair<uint32_t, int>> accepted;
        std::vector<uint32_t> left;
        size_t si = 0;
        for (uint32_t d : todo) {
            int bc = -1;
            if (si < scored.size() && scored[si] == d) bc = verdict[si++];
            dirty[d] = 0;
            if (bc >= 0)
                accepted.push_back({d, bc});
            else
                left.push_back(d);
        }
        if (accepted.empty()) break;
        for (auto& [d, c] : accepted) {
            out.value[d] = (int16_t)c;
            for (uint32_t i = wstart[d]; i < wstart[d + 1]; i++) dirty[wlist[i]] = 1;
            const uint32_t u = vt.list[d].var;
            for (uint32_t w = u > (uint32_t)Occ::BEFORE ? u - Occ::BEFORE : 1; w <= u + Occ::AFTER; w++) {
                const uint32_t e = w == u ? UINT32_MAX : vt.find(w);
                if (e != UINT32_MAX) dirty[e] = 1;
            }
        }
        todo.swap(left);
    }
    uint64_t resolved = 0;
    for (uint32_t d = 0; d < nv; d++)
        if (out.value[d] >= 0) {
            done += vt.list[d].count;
            resolved++;
        }
    if (done)
        note = std::to_string(done) + " of " + std::to_string(unknown) +
               " bytes that copied from data lost with a damaged block were inferred (" + std::to_string(resolved) +
               " lost window bytes, from where their copies appear and the file's own statistics)";
    return out;
}

// ------------------------------------------------------------------ the result and the output


// The output in order, a megabyte at a time to the sink, from host provenance: every piece's
// bytes (those copied from lost windows given their inferred values), cut or padded with the fill
// byte to its emitted length (rescue.cpp:2392-2480). A sink that returns nonzero is called no
// more; the counts are still made.
class Emitter {
  public:
    static constexpr size_t CHUNK = 1u << 20;
    Emitter(gr_sink s, void* u, uint8_t f) : sink_(s), user_(u), fill_(f), buf_(CHUNK), mask_(CHUNK) {}
    int rc() const { return rc_; }
    uint64_t pos() const { return pos_; }
    void flush() {
        if (n_ && rc_ == GR_OK && sink_ && sink_(user_, buf_.data(), mask_.data(), n_)) rc_ = GR_ERR_ABORTED;
        n_ = 0;
    }
    EmitCounts piece(const EmitPiece& p, const uint8_t* b, const uint32_t* tag, const Inferred& value) {
        EmitCounts c;
        const uint64_t e = std::min(p.decoded, p.length);
        for (uint64_t i = 0; i < e; i++) {
            if (tag[i] == Provenance::EXACT) {
                const uint64_t r = exact_run(tag + i, e - i, Provenance::EXACT);
                c.known += r;
                (p.suspect ? c.suspect : c.exact) += r;
                put_run(b + i, p.suspect ? GR_SUSPECT : GR_EXACT, r);
                i += r - 1;
                continue;
            }
            const int got = tag[i] ? value.get(tag[i]) : -1;
            if (got < 0) {
                c.unknown++;
                put((uint8_t)fill_, GR_UNKNOWN);
                continue;
            }
            c.known++;
            if (p.suspect) {
                c.suspect++;
                put((uint8_t)got, GR_SUSPECT);
            } else {
                put((uint8_t)got, GR_INFERRED);
            }
        }
        if (p.length > p.decoded) {
            c.unknown += p.length - p.decoded;
            fills(GR_UNKNOWN, p.length - p.decoded);
        }
        return c;
    }

  private:
    void put(uint8_t x, uint8_t m) {
        buf_[n_] = x;
        mask_[n_] = m;
        n_++;
        pos_++;
        if (n_ == CHUNK) flush();
    }
    // len bytes from x, each with mask m
    void put_run(const uint8_t* x, uint8_t m, uint64_t len) {
        while (len) {
            const size_t k = (size_t)std::min<uint64_t>(len, CHUNK - n_);
            memcpy(buf_.data() + n_, x, k);
            memset(mask_.data() + n_, m, k);
            n_ += k;
            pos_ += k;
            x += k;
            len -= k;
            if (n_ == CHUNK) flush();
        }
    }
    // len fill bytes with mask m
    void fills(uint8_t m, uint64_t len) {
        while (len) {
            const size_t k = (size_t)std::min<uint64_t>(len, CHUNK - n_);
            memset(buf_.data() + n_, fill_, k);
            memset(mask_.data() + n_, m, k);
            n_ += k;
            pos_ += k;
            len -= k;
            if (n_ == CHUNK) flush();
        }
    }
    gr_sink sink_;
    void* user_;
    uint8_t fill_;
    std::vector<uint8_t> buf_, mask_;
    size_t n_ = 0;
    uint64_t pos_ = 0;
    int rc_ = GR_OK;
};

// Provenance in host memory, and the answers a backend gives from it: the CPU backend's (the
// tests' host build); every answer is the one the GPU backend computes on the device
struct HostProv {
    std::vector<uint8_t> byte;
    std::vector<uint32_t> tag;

    static HostProv& make(Provenance& pv, uint64_t total) {
        auto h = std::make_shared<HostProv>();
        h->byte.assign(total, 0);
        h->tag.assign(total, 0);
        pv.total = total;
        pv.store = h;
        return *h;
    }
    static const HostProv& of(const Provenance& pv) { return *static_cast<const HostProv*>(pv.store.get()); }

    static void summarize(const Provenance& pv, const std::vector<PieceRange>& pieces, std::vector<PieceFacts>& facts,
                          const std::vector<std::pair<uint64_t, uint64_t>>& crc_ranges, std::vector<uint32_t>& crcs) {
        const HostProv& h = of(pv);
        const uint32_t EXACT = Provenance::EXACT;
        facts.assign(pieces.size(), PieceFacts{});
        on_cores(pieces.size(), [&](size_t k, unsigned) {
            const PieceRange& r = pieces[k];
            const uint32_t* tg = h.tag.data() + r.at;
            PieceFacts& f = facts[k];
            f.all_exact = exact_run(tg, r.len, EXACT) == r.len;
            f.var64 = f.all_exact ? r.len : first_var64(tg, r.len, EXACT);
            if (r.keep) var_ranges(tg, r.len, EXACT, f.runs);
        });
        crcs.resize(crc_ranges.size());
        for (size_t i = 0; i < crc_ranges.size(); i++)
            crcs[i] = hostcrc::crc32(0, h.byte.data() + crc_ranges[i].first,
                                     (size_t)(crc_ranges[i].second - crc_ranges[i].first));
    }
    // every byte of the stretches [out, out + len) whose tag is a variable, stretch by stretch, with
    // the cells around it
    static void var_places(const Provenance& pv, const std::vector<std::pair<uint64_t, uint64_t>>& stretches,
                           std::vector<Place>& places) {
        const HostProv& h = of(pv);
        const uint32_t EXACT = Provenance::EXACT;
        places.clear();
        for (const auto& [out, len] : stretches) {
            const uint8_t* b = h.byte.data() + out;
            const uint32_t* tg = h.tag.data() + out;
            auto cell = [&](uint64_t j) -> uint32_t {
                if (tg[j] == EXACT) return b[j];
                return tg[j] ? Occ::VAR + tg[j] : Occ::NONE;
            };
            for (uint64_t x = 0; x < len; x++) {
                if (tg[x] == EXACT || tg[x] == 0) continue;
                Place pl;
                pl.pos = out + x;
                pl.tag = tg[x];
                for (int q = 1; q <= Occ::BEFORE; q++) pl.cell[q - 1] = x >= (uint64_t)q ? cell(x - q) : Occ::NONE;
                for (int q = 1; q <= Occ::AFTER; q++) pl.cell[Occ::BEFORE + q - 1] = x + q < len ? cell(x + q) : Occ::NONE;
                places.push_back(pl);
            }
        }
    }
    static void fetch(const Provenance& pv, uint64_t at, uint64_t len, uint8_t* byte, uint32_t* tag) {
        const HostProv& h = of(pv);
        memcpy(byte, h.byte.data() + at, len);
        memcpy(tag, h.tag.data() + at, len * 4);
    }
    static int emit(const Provenance& pv, const std::vector<EmitPiece>& pieces, const Inferred& value, uint8_t fill,
                    gr_sink sink, void* user, std::vector<EmitCounts>& counts, uint64_t& emitted) {
        const HostProv& h = of(pv);
        Emitter em(sink, user, fill);
        counts.resize(pieces.size());
        for (size_t k = 0; k < pieces.size(); k++)
            counts[k] = em.piece(pieces[k], h.byte.data() + pieces[k].at, h.tag.data() + pieces[k].at, value);
        em.flush();
        emitted = em.pos();
        return em.rc();
    }
};

struct Options {
    uint8_t fill = 0;
    bool infer = true;
    uint8_t cr_lost = 0, crs_expected = 0;
    uint64_t max_output = 0, cr_budget = 0;
};

// gr_options read by its size field (rescue.cpp:2492-2506)
inline Options read_options(const gr_options* opt) {
    Options o;
    if (!opt) return o;
    if (opt->size >= 5) o.fill = opt->fill;
    if (opt->size >= 6) o.infer = !opt->no_infer;
    if (opt->size >= 7) o.cr_lost = opt->cr_lost;
    if (opt->size >= 8) o.crs_expected = opt->crs_expected;
    if (opt->size >= 16) o.max_output = opt->max_output;
    if (opt->size >= 24) o.cr_budget = opt->cr_budget;
    return o;
}

// The variables of a pass's output and their places (rescue.cpp:785-834's gather, over every decode
// of the pass in output order): cells within the decode's own bytes
template <class B>
void gather_places(Session<B>& ss, PassResult& o) {
    std::vector<std::pair<uint64_t, uint64_t>> stretches;
    for (const Source& s : o.sources) stretches.push_back({s.out, s.len});
    // every byte that copies a variable counts; the first MAX_OCC of a variable are its places
    std::vector<uint32_t> tags;
    ss.be.var_tags(o.prov, stretches, tags);
    std::vector<uint64_t> picked;
    std::vector<uint32_t> picked_var;
    for (uint64_t i = 0; i < tags.size(); i++) {
        const uint32_t d = o.vars.get(tags[i]);
        Vars::Var& vr = o.vars.list[d];
        vr.count++;
        if (vr.met >= MAX_OCC) continue;
        vr.met++;
        picked.push_back(i);
        picked_var.push_back(d);
    }
    std::vector<Place> places;
    ss.be.places(picked, places);
    // with drop_alike, a place alike one its variable keeps is only counted
    for (size_t j = 0; j < picked.size(); j++) {
        Vars::Var& vr = o.vars.list[picked_var[j]];
        Occ oc;
        oc.pos = places[j].pos;
        memcpy(oc.cell, places[j].cell, sizeof oc.cell);
        if (ss.drop_alike && Session<B>::alike(vr, oc)) continue;
        vr.places.push_back(oc);
    }
}

// A whole rescue (rescue.cpp:2488-2765's steps): input prepared, headers found, the passes, the
// checks after them, the lost windows inferred, the output and result written. The backend is
// given the prepared input by attach(in, size).
template <class B>
int run_rescue(B& be, const uint8_t* in_, uint64_t size, bool gzip, uint64_t expected_size, bool has_crc,
               uint32_t expected_crc, const Options& opt, gr_sink sink, void* user, RescueResult& res) {
    PhaseClock clock(res);
    // the input copied in parts on the host's cores, with zero padding: the readers look ahead
    constexpr uint64_t PAD = 64, PART = 64u << 20;
    std::unique_ptr<uint8_t[]> buf(new uint8_t[size + PAD]);
    uint8_t* const in = buf.get();
    on_cores((size_t)((size + PART - 1) / PART), [&](size_t k, unsigned) {
        memcpy(in + k * PART, in_ + k * PART, (size_t)std::min(PART, size - k * PART));
    });
    memset(in + size, 0, PAD);
    const LineEnds le = line_ends(in, size);
    const bool ascii_undone = undo_crlf(in, size, PAD, le);
    const uint64_t zero_tail = trim_zero_tail(size, in);
    const bool cr_stripped = !ascii_undone && (opt.cr_lost == 1 || (opt.cr_lost == 0 && lost_crs_likely(le, size)));
    const uint64_t crs_restored = 0, cr_decoded = 0;  // putting lost CRs back in a raw stream is not done here
    clock.mark("prepare input");
    be.attach(in, size);
    clock.mark("input");
    Session<B> ss(be, in, size);
    ss.drop_alike = cr_stripped && !crs_restored && gzip;
    const uint64_t hlen = gzip ? member_header(in, size, 0) : 0;
    be.find(gzip ? (hlen ? hlen * 8 : 80) : 1, ss.headers);
    clock.mark("find headers");
    if (gzip && ss.headers.empty() && !hlen) return GR_ERR_NO_DEFLATE;
    if (!gzip && size == 0) return GR_ERR_NO_DEFLATE;
    ss.max_output = opt.max_output;
    ss.decode_all_headers();
    clock.mark("decode from headers");
    std::vector<Span> spans;
    auto pass = [&](const std::map<uint64_t, std::string>& forced) {
        PassResult o = ss.walk(gzip, hlen, forced, expected_size, has_crc, expected_crc, spans);
        clock.mark("walk");
        if (!o.too_big) ss.finish_pass(o, spans, clock);
        return o;
    };
    PassResult o = pass({});
    std::string anomaly_note;
    Habit habit;
    std::map<uint64_t, std::string> forced;
    const bool crs_lost = cr_stripped && !crs_restored;
    std::vector<char> whole(o.pieces.size(), 0);
    std::set<uint64_t> odd = count_anomalies(o, whole, crs_lost, anomaly_note, habit);
    // a member whose CRC checks is whole: counts that stand out in it are its encoder's
    bool crc_failed = false;
    if (!odd.empty()) {
        bool any = false;
        for (const Member& m : o.members) {
            const int c = member_crc(o, m);
            crc_failed |= c == 0 && m.last_piece > m.first_piece;
            if (c == 1) {
                for (size_t k = m.first_piece; k <= m.last_piece && k < whole.size(); k++) whole[k] = 1;
                any = true;
            }
        }
        if (any) {
            anomaly_note.clear();
            odd = count_anomalies(o, whole, crs_lost, anomaly_note, habit);
        }
    }
    std::string hidden;  // what showed damage of a kind that can keep the decoder in step
    if (crs_lost && gzip)
        hidden = "CRs removed by an ASCII-mode transfer";
    else if (crc_failed)
        hidden = "the CRC fails on the stream as decoded, every block in step";
    auto no_run = [&](const Piece& p) {
        uint64_t n = 0;
        uint8_t v = 0;
        return ss.first_run(p.in_bit_start / 8, (p.in_bit_end + 7) / 8 + 8, n, v) == UINT64_MAX;
    };
    for (const Piece& p : o.pieces)
        if (hidden.empty() && odd.count(p.in_bit_start) && no_run(p))
            hidden = "a block decoded in step with a symbol count its encoder never writes, and no run of zeros in it";
    for (uint64_t b : odd) forced[b] = "its symbol count differs from the stream's regular one";
    if (!forced.empty() && !o.too_big) o = pass(forced);
    std::vector<std::string> crc_notes;
    std::map<uint64_t, std::string> placed;
    if (!o.too_big) placed = ss.crc_misses(o, crc_notes);
    if (!placed.empty() && !o.too_big) {
        forced.insert(placed.begin(), placed.end());
        o = pass(forced);
        for (Piece& p : o.pieces) p.by_crc = placed.count(p.in_bit_start) > 0;
    }
    if (o.too_big) return GR_ERR_TOO_BIG;
    clock.mark("checks");
    if (!anomaly_note.empty()) o.notes.push_back(anomaly_note);
    o.notes.insert(o.notes.end(), crc_notes.begin(), crc_notes.end());
    // bit rot that broke a decode: one flipped bit mends the block
    int looked = 0;
    for (const Piece& p : o.pieces) {
        if (!hidden.empty() || looked >= FLIP_PIECES) break;
        if (!p.suspect || p.by_crc || p.status != gz::ERR_DATA || !no_run(p)) continue;
        looked++;
        if (ss.one_flip_mends(p, habit)) hidden = "a damaged block that one flipped bit mends";
    }
    ss.suspect_overfull_tail(o, habit.regular);
    ss.localize(o);
    clock.mark("checks");
    if (ascii_undone) o.notes.push_back("the file had gone through an ASCII-mode transfer (every LF became CR LF): undone");
    if (crs_restored)
        o.notes.push_back(std::to_string(crs_restored) + " CR" + (crs_restored == 1 ? "" : "s") +
                          " put back that an ASCII-mode transfer had removed; the CRC confirms it");
    else if (cr_stripped && gzip)
        o.notes.push_back("the file went through an ASCII-mode transfer that removed the CR of every CR LF (it has almost no CR LF "
                          "pair where chance would put dozens): the blocks it broke are flagged; the CRs cannot be put back "
                          "here, with one CRC for the whole file to judge them. A copy transferred in binary mode is whole");
    if (zero_tail)
        o.notes.push_back("the file ends in " + std::to_string(zero_tail) +
                          " zero bytes (written by a crash or an unfinished download, not by gzip): the data ends where they begin");
    std::vector<Piece>& pieces = o.pieces;
    uint64_t inferred = 0;
    Inferred value;
    if (opt.infer) {
        gather_places(ss, o);
        std::string note;
        clock.mark("gather places");
        const Fetch fetch = [&](uint64_t at, uint64_t len, uint8_t* b, uint32_t* t) { be.fetch(o.prov, at, len, b, t); };
        value = infer_lost(o, fetch, inferred, note);
        if (!note.empty()) o.notes.push_back(note);
        clock.mark("infer");
    }
    gr_stats& st = res.stats;
    st.input_size = size;
    st.headers = (uint32_t)ss.headers.size();
    st.members = (uint32_t)o.members.size();
    st.truncated = o.truncated;
    st.header_damaged = gzip && hlen == 0;
    st.crc_check = -1;
    st.input_end = o.input_end;
    st.inferred_bytes = inferred;
    st.ascii_undone = ascii_undone;
    st.cr_restored = (uint32_t)crs_restored;
    st.cr_decoded = cr_decoded;
    // each suspect piece's true length: the trailer's size less what is certain, spread over the
    // suspect pieces by their compressed size (rescue.cpp:2629-2679); and each member's CRC
    for (Member& m : o.members) {
        uint64_t good = 0, suspect_bits = 0, sus_decoded = 0, sized = 0;
        int nsus = 0, nspread = 0;
        for (size_t k = m.first_piece; k <= m.last_piece && k < pieces.size(); k++) {
            const Piece& p = pieces[k];
            if (p.suspect) nsus++;
            if (p.suspect && !p.length_exact) {
                nspread++;
                sus_decoded += p.len;
                suspect_bits += p.in_bit_end > p.in_bit_start ? p.in_bit_end - p.in_bit_start : 0;
            } else if (p.suspect) {
                sized += p.len;
            } else {
                good += p.len;
            }
        }
        const uint32_t missing = m.isize - (uint32_t)(good + sized);
        if (m.trailer && nspread > 0 && missing <= 4 * sus_decoded + (64u << 20)) {
            uint64_t given = 0;
            int seen = 0;
            for (size_t k = m.first_piece; k <= m.last_piece && k < pieces.size(); k++) {
                if (!pieces[k].suspect || pieces[k].length_exact) continue;
                seen++;
                const uint64_t bits =
                    pieces[k].in_bit_end > pieces[k].in_bit_start ? pieces[k].in_bit_end - pieces[k].in_bit_start : 0;
                pieces[k].true_len =
                    seen == nspread ? missing - given : (suspect_bits ? (uint64_t)missing * bits / suspect_bits : 0);
                given += pieces[k].true_len;
            }
        }
        if (m.trailer && nsus == 0 && !pieces.empty() && pieces[m.last_piece].status == gz::OK) {
            bool all_known = true;
            for (size_t k = m.first_piece; k <= m.last_piece; k++) all_known &= pieces[k].all_exact;
            if (all_known) {
                const bool size_ok = !m.size_known || (uint32_t)good == m.isize;
                if (m.has_crc)
                    m.crc_check = m.out_crc == m.crc && size_ok;
                else if (m.size_known)
                    m.crc_check = size_ok ? -1 : 0;
                if (&m == &o.members.front()) st.crc_check = m.crc_check;
                if (m.size_known && (uint32_t)good != m.isize) {
                    st.length_off = (int64_t)(int32_t)((uint32_t)good - m.isize);
                    o.notes.push_back("the output is " + std::to_string(-st.length_off) +
                                      " bytes off the length the trailer gives: a block was damaged in a way that went "
                                      "unnoticed (the decoder fell back into step inside it); the CRC says where it is not");
                }
            }
        }
    }
    if (st.crc_check == -1) {
        bool damaged = false;
        for (const Piece& p : pieces) damaged |= p.suspect && !p.by_crc;
        if (damaged && !hidden.empty())
            o.notes.push_back("with damaged blocks in it, the stream's CRC cannot be checked, and its damage is of a kind that can "
                              "keep the decoder in step (" + hidden + "): damage like that elsewhere would not show, so the bytes "
                              "marked exact outside the flagged blocks are not vouched for");
        else if (damaged)
            o.notes.push_back("with damaged blocks in it, the stream's CRC cannot be checked: bytes marked exact are exact as far as "
                              "decoding can tell, and damage that kept the decoder in step (a flipped bit that changed a literal "
                              "or a match length) would not show");
    }
    st.hidden_damage = !hidden.empty() && st.crc_check != 1;
    uint64_t out_size = 0;
    for (const Piece& p : pieces) out_size += p.suspect && p.true_len ? p.true_len : p.len;
    if (opt.max_output && out_size > opt.max_output) return GR_ERR_TOO_BIG;
    std::vector<EmitPiece> emit(pieces.size());
    for (size_t k = 0; k < pieces.size(); k++) {
        const Piece& p = pieces[k];
        emit[k] = EmitPiece{o.sources[p.src].out + p.src_off, p.len, p.suspect && p.true_len ? p.true_len : p.len, p.suspect};
    }
    std::vector<EmitCounts> counts;
    uint64_t emitted = 0;
    const int erc = be.emit(o.prov, emit, value, opt.fill, sink, user, counts, emitted);
    clock.mark("emit");
    uint64_t at = 0;
    for (size_t k = 0; k < pieces.size(); k++) {
        const Piece& p = pieces[k];
        const EmitCounts& c = counts[k];
        st.exact_bytes += c.exact;
        st.suspect_bytes += c.suspect;
        st.unknown_bytes += c.unknown;
        res.why.push_back(p.why);
        res.pieces.push_back(gr_piece{at, emit[k].length, p.len, c.known, p.in_bit_start / 8, p.in_bit_end / 8,
                                      p.suspect ? 1 : 0, p.status, p.blocks, p.symbols, nullptr});
        if (p.suspect) st.damaged++;
        at += emit[k].length;
    }
    for (size_t k = 0; k < res.pieces.size(); k++) res.pieces[k].why = res.why[k].c_str();
    st.output_size = emitted;
    st.pieces = (uint32_t)res.pieces.size();
    res.notes = o.notes;
    for (const std::string& nn : res.notes) res.note_ptrs.push_back(nn.c_str());
    return erc;
}

}  // namespace rsc
