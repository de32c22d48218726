// plan.h - the host side of one stream's parallel decode: segments, block starts, probes,
// decoding jobs and their continuations, stitching the pieces, and the segment loop. It runs
// against a Backend that does the data-parallel work (the HIP backend in gpuinflate.hip; a
// CPU backend over the same core for tests), so every planning decision is testable without a
// GPU.
//
// Semantics follow the reference driver (gpuinflate/gpuinflate.cu:87-285 planning, 638-873
// the segment loop), in this design's terms:
//  * the whole input is in device memory, and every bit position is a position in the stream
//    (no per-segment header copy: a segment that starts inside a dynamic block names its
//    header, wherever it lies);
//  * every verified dynamic header in a segment gets a Huffman table in the cache, so
//    decoders cross block boundaries without rebuilding tables; one header per chunk starts
//    a job;
//  * a match reaching before the stream's start is found while stitching, in stream order
//    (Result::reach), so the first error in the stream decides the error class.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <future>
#include <numeric>
#include <string>
#include <vector>

#include "deflate_core.h"

namespace gz {

constexpr uint64_t SLACK = 8192;              // output symbols every job gets beyond its share
constexpr int MAX_ROUNDS = 64;                // decode rounds per segment before cutting it short
constexpr uint64_t CHUNK_BITS = 4096 * 8;     // one job per chunk at most, from the block finder
constexpr uint64_t PROBE_SPACING = 1ull << 20;  // global probes split intervals longer than this
constexpr uint64_t PROBE_LEAD = 4096;         // a probe decodes this many bits before its point
constexpr uint32_t JOB_CAP_MAX = 4u << 20;    // symbols one decoder writes per round
constexpr uint32_t JOB_CAP_MIN = 4096;
constexpr uint64_t SEG_IN_MAX = 256ull << 20; // compressed bytes per segment at most, by default
constexpr uint64_t SEG_IN_MIN = 256 << 10;

struct Probe {
    uint64_t header, point;
    uint8_t mode;
    uint8_t pad[7];
};
static_assert(sizeof(Probe) == 24, "Probe layout");

// What the block finder reports for a segment: verified dynamic headers (bit positions), and
// byte positions where a stored block's LEN and NLEN could be (LEN == ~NLEN), packed as
// byte << 16 | LEN
struct Found {
    std::vector<uint64_t> dyn;
    std::vector<uint32_t> est;  // per dyn[i] (when given): its code's estimated output bytes per input byte, x256
    std::vector<uint64_t> stored;
    uint32_t slot_base = 1;  // the segment's table slots start here
    bool prebuilt = false;   // the table of dyn[i] is built already, in slot slot_base + i
};

// A stored block is confirmed when what follows its data is itself a block start the finder
// saw: another stored block (LEN/NLEN one byte after the header byte) or a verified dynamic
// header. Returns the confirmed blocks as (LEN byte, end byte), sorted.
// Membership in a sorted vector by buckets of its value range (about one value a bucket): a
// lookup scans one bucket
struct SortedSet {
    const std::vector<uint64_t>* v = nullptr;
    uint64_t lo = 0, hi = 0;
    int shift = 0;
    std::vector<uint32_t> first;  // first[b]: index of the first value in bucket b or after it
    explicit SortedSet(const std::vector<uint64_t>& sorted) : v(&sorted) {
        if (sorted.empty()) return;
        lo = sorted.front(), hi = sorted.back();
        while (((hi - lo) >> shift) > sorted.size()) shift++;
        const size_t nb = (size_t)((hi - lo) >> shift) + 2;
        first.assign(nb, (uint32_t)sorted.size());
        for (size_t i = sorted.size(); i-- > 0;) first[(size_t)((sorted[i] - lo) >> shift)] = (uint32_t)i;
        for (size_t b = nb - 1; b-- > 0;) first[b] = std::min(first[b], first[b + 1]);
    }
    bool has(uint64_t x) const {
        if (first.empty() || x < lo || x > hi) return false;
        const size_t b = (size_t)((x - lo) >> shift);
        for (uint32_t i = first[b]; i < first[b + 1]; i++)
            if ((*v)[i] >= x) return (*v)[i] == x;
        return false;
    }
};

inline std::vector<std::pair<uint64_t, uint64_t>> confirm_stored(const std::vector<uint64_t>& sorted_dyn,
                                                                 std::vector<uint64_t> stored,
                                                                 std::vector<uint64_t>& bytes) {
    std::sort(stored.begin(), stored.end());
    bytes.resize(stored.size());
    for (size_t i = 0; i < stored.size(); i++) bytes[i] = stored[i] >> 16;
    const SortedSet in_bytes(bytes), in_dyn(sorted_dyn);
    std::vector<std::pair<uint64_t, uint64_t>> out;
    for (uint64_t c : stored) {
        uint64_t at = c >> 16, q = at + 4 + (c & 0xFFFF);
        if (in_bytes.has(q + 1) || in_dyn.has(q * 8)) out.push_back({at, q});
    }
    return out;
}

// A segment's finder report made ready for planning: the verified headers sorted (raw_dyn keeps
// the report's order, which the prebuilt tables' slots follow), their ratio estimates sorted by
// bit, and the stored candidates confirmed. Needs nothing of the segment before, so it can run
// beside that segment's decode.
struct Prepared {
    Found f;                                       // f.dyn sorted
    std::vector<uint64_t> raw_dyn;                 // as reported
    std::vector<uint32_t> dyn_order;               // indices of raw_dyn in bit order
    std::vector<std::pair<uint64_t, uint32_t>> est;  // (header bit, estimate), sorted
    std::vector<uint64_t> cand_bytes;
    std::vector<std::pair<uint64_t, uint64_t>> stored;
    size_t nstored_cand = 0;
    double confirm_s = 0;
};
inline Prepared prepare_found(Found f) {
    Prepared p;
    p.raw_dyn = f.dyn;
    p.dyn_order.resize(f.dyn.size());
    std::iota(p.dyn_order.begin(), p.dyn_order.end(), 0u);
    std::sort(p.dyn_order.begin(), p.dyn_order.end(), [&](uint32_t a, uint32_t b) { return p.raw_dyn[a] < p.raw_dyn[b]; });
    for (size_t i = 0; i < f.dyn.size(); i++) f.dyn[i] = p.raw_dyn[p.dyn_order[i]];
    if (f.est.size() == f.dyn.size()) {
        p.est.resize(f.dyn.size());
        for (size_t i = 0; i < f.dyn.size(); i++) p.est[i] = {f.dyn[i], f.est[p.dyn_order[i]]};
    }
    p.nstored_cand = f.stored.size();
    const auto t = std::chrono::steady_clock::now();
    p.stored = confirm_stored(f.dyn, std::move(f.stored), p.cand_bytes);
    p.confirm_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
    p.f = std::move(f);
    return p;
}

// Job starts: the segment's start; the verified dynamic headers outside confirmed stored data,
// one per chunk at most; the block after every confirmed stored block (marked `hint` 1 when it
// is itself stored). `hint` travels in Start::pad[0].
inline std::vector<Start> block_starts(const std::vector<uint64_t>& sorted_dyn,
                                      const std::vector<std::pair<uint64_t, uint64_t>>& stored,
                                      const std::vector<uint64_t>& cand_bytes, const Start& s0, uint64_t end,
                                      uint64_t chunk_bits = CHUNK_BITS) {
    std::vector<Start> v;
    auto push = [&](uint64_t b, uint8_t hint) {
        Start s{};
        s.bit = b;
        s.header = b;
        s.mode = AT_BLOCK;
        s.final = -1;
        s.pad[0] = hint;
        v.push_back(s);
    };
    // stored data regions [8*LEN byte, 8*end byte), sorted
    size_t r = 0;
    uint64_t last_chunk = UINT64_MAX;
    for (uint64_t b : sorted_dyn) {
        if (b <= s0.bit || b >= end) continue;
        while (r < stored.size() && stored[r].second * 8 <= b) r++;
        if (r < stored.size() && stored[r].first * 8 <= b) continue;  // inside stored data
        uint64_t c = (b - s0.bit) / chunk_bits;
        if (c == last_chunk) continue;
        last_chunk = c;
        push(b, 0);
    }
    const size_t nd = v.size();  // the headers' starts, in bit order; then the stored blocks' ends
    for (auto& st : stored) {
        uint64_t b = st.second * 8;
        if (b <= s0.bit || b >= end) continue;
        bool next_stored = std::binary_search(cand_bytes.begin(), cand_bytes.end(), st.second + 1);
        push(b, next_stored ? 1 : 0);
    }
    // stored blocks are sorted by LEN byte; their ends are too unless blocks overlap (false
    // candidates can), so sort that run only when it is out of order, then merge the two runs
    auto by_bit = [](const Start& a, const Start& b) { return a.bit < b.bit; };
    if (!std::is_sorted(v.begin() + nd, v.end(), by_bit)) std::sort(v.begin() + nd, v.end(), by_bit);
    std::inplace_merge(v.begin(), v.begin() + nd, v.end(), by_bit);
    std::vector<Start> out{s0};
    for (auto& x : v)
        if (x.bit > out.back().bit) out.push_back(x);
    return out;
}

// Probe points every `spacing` bits inside every stretch between starts, with the code the
// stretch begins with; none in stretches that begin with a stored block, and none inside
// confirmed stored data (gpuinflate.cu:94-103 probes only stretches longer than 2 * spacing;
// here every stretch is split, since a short lead resynchronises)
inline std::vector<Probe> probe_points(const std::vector<Start>& s, uint64_t end, uint64_t spacing,
                                       const std::vector<std::pair<uint64_t, uint64_t>>& stored) {
    std::vector<Probe> pts;
    size_t r = 0;
    for (size_t k = 0; k < s.size(); k++) {
        uint64_t e = k + 1 < s.size() ? s[k + 1].bit : end;
        if (s[k].pad[0] == 1 || e <= s[k].bit) continue;
        const uint64_t header = s[k].mode == AT_BLOCK ? s[k].bit : s[k].header;
        // points at bit + i * spacing, i >= 1, each at least half a spacing before the stretch ends
        const uint64_t half = spacing / 2;
        const uint64_t npts = e > s[k].bit + half ? (e - 1 - half - s[k].bit) / spacing : 0;
        for (uint64_t i = 1; i <= npts; i++) {
            const uint64_t b = s[k].bit + i * spacing;
            while (r < stored.size() && stored[r].second * 8 <= b) r++;
            if (r < stored.size() && stored[r].first * 8 <= b + 16) continue;  // in stored data or its header
            Probe p{};
            p.header = header;
            p.point = b;
            p.mode = s[k].mode;
            pts.push_back(p);
        }
    }
    return pts;
}

// The block starts with the probes' findings merged in: both runs sorted by bit, a probed start
// outside (lo, end) dropped, and where two starts share a bit the block start is kept (among
// probed ones, the first probe's). `starts` is sorted with distinct bits.
inline std::vector<Start> merge_probed(const std::vector<Start>& starts, const std::vector<Start>& probed, uint64_t lo,
                                       uint64_t end) {
    std::vector<Start> extra;
    extra.reserve(probed.size());
    for (const Start& p : probed)
        if (p.bit > lo && p.bit < end) extra.push_back(p);
    std::stable_sort(extra.begin(), extra.end(), [](const Start& a, const Start& b) { return a.bit < b.bit; });
    std::vector<Start> out;
    out.reserve(starts.size() + extra.size());
    size_t a = 0, b = 0;
    while (a < starts.size() || b < extra.size()) {
        // the block start first on a tie
        const bool from_starts = b == extra.size() || (a < starts.size() && starts[a].bit <= extra[b].bit);
        const Start& x = from_starts ? starts[a++] : extra[b++];
        if (out.empty() || out.back().bit != x.bit) out.push_back(x);
    }
    return out;
}

// Decoding jobs: first one per target, then continuations (gpuinflate.cu:116-174). A first-round
// job is a group (`lanes` lanes over the interval up to the next target, with cap symbols per
// lane) unless its interval starts with a stored block; continuations are single lanes.
enum : uint8_t { JOB_LANE = 0, JOB_GROUP = 1 };
constexpr uint32_t LANE_SLACK = 2048;  // output symbols a group lane gets beyond its share
constexpr uint64_t GROUP_MIN_BITS = 16384;  // a continuation over a shorter stretch is one lane

struct Jobs {
    std::vector<Start> targets;  // job k < targets.size() starts at targets[k]
    std::vector<Start> start;
    std::vector<int> tfirst;     // a job lands only on targets[tfirst..]
    std::vector<uint64_t> off;   // its region of the 16-bit pool
    std::vector<uint32_t> cap;   // symbols (a group: per lane)
    std::vector<uint8_t> kind;
    std::vector<uint64_t> end;   // a group's interval end
    std::vector<uint64_t> split_end;
    std::vector<uint64_t> hard;  // a group's first stored block's LEN bit: Huffman data stops there (0: none)
    std::vector<uint32_t> extra; // a group's last lane: room beyond cap
    std::vector<Result> res;
    std::vector<GroupHead> ghead;  // a group's result, but for its lanes' lengths (the backend keeps those)
    std::vector<int> cont;       // the job continuing it, or -1
    std::vector<uint64_t> lim;   // the end (byte) of the input the job's stream owns (0: the whole input)
    uint64_t used = 0, pool = 0;
    size_t max_jobs = 0;
    int lanes = GROUP_LANES_DEFAULT;  // a group's lanes
    uint64_t seg_end = 0;        // the segment's last bit (a job past every target runs to it)
    std::vector<std::pair<uint64_t, uint64_t>> stored;  // confirmed stored blocks (LEN byte, end byte)
    // a batch: the streams' (first byte, end byte) in the packed input, sorted (empty: one stream)
    std::vector<std::pair<uint64_t, uint64_t>> streams;

    // the headers' ratio estimates (bit, output bytes per input byte x256), sorted by bit
    std::vector<std::pair<uint64_t, uint32_t>> est;
    // the ratio to size a job at bit b by: the estimate of the last header at or before b, else dflt
    double ratio_at(uint64_t b, double dflt) const {
        auto it = std::upper_bound(est.begin(), est.end(), b,
                                   [](uint64_t v, const std::pair<uint64_t, uint32_t>& x) { return v < x.first; });
        return ratio_of((size_t)(it - est.begin()), dflt);
    }
    // the same for bits that do not decrease from one call to the next: *cur (start at 0) is the
    // number of estimates at or before the last bit asked for
    double ratio_at(uint64_t b, double dflt, size_t* cur) const {
        if (*cur && est[*cur - 1].first > b) return ratio_at(b, dflt);
        while (*cur < est.size() && est[*cur].first <= b) ++*cur;
        return ratio_of(*cur, dflt);
    }
    double ratio_of(size_t n_before, double dflt) const {
        if (!n_before || !est[n_before - 1].second) return dflt;
        return std::clamp((double)est[n_before - 1].second / 256.0, 1.0, 1032.0);
    }
    // the end byte of the stream that holds bit b (0: one stream)
    uint64_t lim_of(uint64_t b) const {
        if (streams.empty()) return 0;
        size_t i = (size_t)(std::upper_bound(streams.begin(), streams.end(), b / 8,
                                             [](uint64_t v, const std::pair<uint64_t, uint64_t>& x) { return v < x.first; }) -
                            streams.begin());
        return i ? streams[i - 1].second : streams.front().second;
    }

    // takes the vectors (emptied, their memory kept) of a Jobs done with: a segment's jobs then
    // reuse the pages of the segment before's instead of faulting in fresh ones
    void take_memory(Jobs& old) {
        auto take = [](auto& mine, auto& theirs) {
            mine.swap(theirs);
            mine.clear();
        };
        take(start, old.start), take(tfirst, old.tfirst), take(off, old.off), take(cap, old.cap);
        take(kind, old.kind), take(end, old.end), take(split_end, old.split_end), take(hard, old.hard);
        take(extra, old.extra), take(lim, old.lim), take(res, old.res), take(ghead, old.ghead), take(cont, old.cont);
    }
    void reserve(size_t n) {
        start.reserve(n);
        tfirst.reserve(n);
        off.reserve(n);
        cap.reserve(n);
        kind.reserve(n);
        end.reserve(n);
        split_end.reserve(n);
        hard.reserve(n);
        extra.reserve(n);
        lim.reserve(n);
        res.reserve(n);
        ghead.reserve(n);
        cont.reserve(n);
    }
    // One job appended: its region of the pool follows the last job's
    struct Spec {
        Start start;
        int tfirst;                 // the first target it may land on
        uint32_t cap;               // symbols (a group: per lane)
        uint8_t kind = JOB_LANE;
        uint64_t end = 0;           // a group's interval end
        uint64_t split_end = 0;     // where its lanes' split ends (0: end)
        uint32_t extra = 0;         // its last lane's room beyond cap
        uint64_t hard = 0;          // its first stored block's LEN bit (0: none)
    };
    uint64_t room_of(const Spec& x) const { return x.kind == JOB_GROUP ? (uint64_t)x.cap * lanes + x.extra : x.cap; }
    void append(const Spec& x) {
        off.push_back(used);
        used += room_of(x);
        start.push_back(x.start);
        tfirst.push_back(x.tfirst);
        cap.push_back(x.cap);
        kind.push_back(x.kind);
        end.push_back(x.end);
        split_end.push_back(x.split_end ? x.split_end : x.end);
        hard.push_back(x.hard);
        extra.push_back(x.extra);
        lim.push_back(lim_of(x.start.bit));
        res.emplace_back();
        ghead.emplace_back();
        cont.push_back(-1);
    }
};

// The shape of a group of `lanes` over bits [b, e): where its lanes' split ends (the first stored
// block in it, whose data only the last lane crosses; 0: none) and the room that lane needs for
// the stored bytes; returns the output symbols one lane wants.
// cur (optional, start at 0): for b that do not decrease between calls, a cursor that replaces
// the search for the first stored block ending after b
inline double group_shape(uint64_t b, uint64_t e, double ratio, const std::vector<std::pair<uint64_t, uint64_t>>& stored,
                          int lanes, uint64_t& split, uint64_t& extra, uint64_t& hard, size_t* cur = nullptr) {
    split = 0;
    extra = 0;
    hard = 0;
    size_t r;
    if (cur) {
        while (*cur < stored.size() && stored[*cur].second <= b / 8) ++*cur;
        r = *cur;
    } else {
        r = (size_t)(std::upper_bound(stored.begin(), stored.end(), b / 8,
                                      [](uint64_t v, const std::pair<uint64_t, uint64_t>& a) { return v < a.second; }) -
                     stored.begin());
    }
    while (r > 0 && stored[r - 1].second * 8 > b) r--;
    for (size_t q = r; q < stored.size() && stored[q].first * 8 < e; q++) {
        // a stored block whose LEN bit is at or before b is a false candidate: a start is never
        // inside real stored data (a stored block's decode resumes at its header)
        if (stored[q].first * 8 <= b) continue;
        if (!split) {
            split = stored[q].first * 8 > b + 16 ? stored[q].first * 8 - 16 : b + 1;
            hard = stored[q].first * 8;
        }
        extra += (stored[q].second - stored[q].first) + SLACK;
    }
    const uint64_t span = e > b ? e - b : 0;
    const uint64_t sspan = split ? split - b : span;
    return std::min<double>(2 * ratio * (double)(sspan / 8) / lanes + LANE_SLACK, JOB_CAP_MAX);
}

// Output room a lane job asks for over `bytes` of input at `ratio` output bytes per input byte:
// HEADROOM times the expected output, plus slack, at most JOB_CAP_MAX
constexpr double HEADROOM = 2.0;
inline uint64_t lane_room(uint64_t bytes, double ratio) {
    return (uint64_t)std::min<double>(HEADROOM * ratio * (double)bytes + SLACK, JOB_CAP_MAX);
}

// the index of the first target strictly after bit
inline int first_target_after(const std::vector<Start>& targets, uint64_t bit) {
    return (int)(std::partition_point(targets.begin(), targets.end(), [&](const Start& t) { return t.bit <= bit; }) -
                 targets.begin());
}

// The first round: one job per target. A target that starts with a stored block is a lane job;
// the others are groups, shaped by group_shape (a group whose interval holds stored data splits
// only the part before it, and its last lane gets room for the stored bytes). The first round
// takes at most half the pool's free part (the rest is for continuations): when the jobs ask for
// more, every request is scaled down by the same factor, to a floor per job.
// stored: the confirmed stored blocks (LEN byte, end byte), sorted. A batch passes its streams'
// byte ranges, and the pool's first `pool_base` symbols stay free. Groups have `lanes` lanes.
inline Jobs make_jobs(std::vector<Start> targets, uint64_t end, double ratio, uint64_t pool, size_t max_jobs, int lanes,
                      const std::vector<std::pair<uint64_t, uint64_t>>& stored = {},
                      const std::vector<std::pair<uint64_t, uint64_t>>& streams = {}, uint64_t pool_base = 0,
                      std::vector<std::pair<uint64_t, uint32_t>> est = {}, Jobs* spare = nullptr) {
    Jobs j;
    if (spare) j.take_memory(*spare);
    j.lanes = lanes;
    j.pool = pool;
    j.used = pool_base;
    j.streams = streams;
    j.est = std::move(est);
    j.max_jobs = max_jobs;
    j.targets = std::move(targets);
    j.seg_end = end;
    j.stored = stored;
    const size_t n = j.targets.size();
    j.reserve(n + n / 8 + 64);  // room for the continuations of a round or two
    std::vector<Jobs::Spec> spec(n);
    std::vector<double> asks(n);  // symbols asked for: per lane for a group
    double asked = 0;
    size_t scur = 0, ecur = 0;  // the targets are in bit order: cursors into stored and est
    for (size_t k = 0; k < n; k++) {
        const Start& t = j.targets[k];
        Jobs::Spec& x = spec[k];
        x.start = t;
        x.tfirst = (int)k + 1;
        x.end = k + 1 < n ? j.targets[k + 1].bit : end;
        if (t.pad[0] == 1) {  // a stored block first: one lane
            x.kind = JOB_LANE;
            asks[k] = (double)lane_room(x.end > t.bit ? (x.end - t.bit) / 8 : 0, ratio);
            asked += asks[k];
            continue;
        }
        x.kind = JOB_GROUP;
        uint64_t split, extra, hard;
        asks[k] = group_shape(t.bit, x.end, j.ratio_at(t.mode == IN_DYNAMIC ? t.header : t.bit, ratio, &ecur), stored,
                              lanes, split, extra, hard, &scur);
        x.split_end = split;
        x.hard = hard;
        x.extra = (uint32_t)std::min<uint64_t>(extra, JOB_CAP_MAX);
        asked += asks[k] * lanes + (double)extra;
    }
    const uint64_t first_round = (pool - pool_base) / 2;
    const double scale = asked > (double)first_round ? (double)first_round / asked : 1.0;
    for (size_t k = 0; k < n; k++) {
        const double floor = spec[k].kind == JOB_GROUP ? JOB_CAP_MIN / 4 : JOB_CAP_MIN;
        spec[k].cap = (uint32_t)std::max(floor, asks[k] * scale);
        j.append(spec[k]);
    }
    return j;
}

// A continuation of job k from the state where it stopped. Jobs other than the chain's own leave
// `reserve` symbols of the pool free. Over a long stretch to the next target it is a group (the
// stretch's targets whose own first-round decode failed are passed over: they are not block
// starts); else one lane. A lane that continues an overflow is sized by what job k measured: the
// output it made per input byte it read, over the input left to the next target, at least four
// times the room job k had, so a run of overflows grows geometrically.
inline bool add_continuation(Jobs& j, size_t k, uint64_t reserve, double ratio) {
    if (j.cont[k] >= 0 || j.start.size() >= j.max_jobs) return false;
    const Start r = j.res[k].resume;
    const int tf = first_target_after(j.targets, r.bit);  // one at r.bit may be where job k stopped
    const uint64_t left = j.used < j.pool ? j.pool - j.used : 0;
    if (left < reserve + JOB_CAP_MIN) return false;
    {
        size_t te = (size_t)tf;
        while (te < j.targets.size() && te < j.res.size() &&
               (j.res[te].status == ERR_DATA || j.res[te].status == ERR_INPUT))
            te++;
        const uint64_t e = te < j.targets.size() ? j.targets[te].bit : j.seg_end;
        uint64_t split = 0, extra = 0, hard = 0;
        const double rr = j.ratio_at(r.mode == IN_DYNAMIC ? r.header : r.bit, ratio);
        double want = e > r.bit ? group_shape(r.bit, e, rr, j.stored, j.lanes, split, extra, hard) : 0;
        // not a group: a short stretch, one that starts with stored data, or no progress since job k
        const bool group = e > r.bit && e - r.bit >= GROUP_MIN_BITS && !(split && split <= r.bit + 1) &&
                           !(j.kind[k] == JOB_GROUP && r.bit == j.start[k].bit);
        if (group) {
            uint64_t c = (uint64_t)std::max<double>(want, JOB_CAP_MIN / 4);
            extra = std::min<uint64_t>(extra, JOB_CAP_MAX);
            if (c * j.lanes + extra + reserve > left) {
                if (left < reserve + extra + (uint64_t)j.lanes * (JOB_CAP_MIN / 4)) return false;
                c = (left - reserve - extra) / j.lanes;
            }
            j.cont[k] = (int)j.start.size();
            Jobs::Spec x;
            x.start = r;
            x.tfirst = tf;
            x.cap = (uint32_t)c;
            x.kind = JOB_GROUP;
            x.end = e;
            x.split_end = split;
            x.extra = (uint32_t)extra;
            x.hard = hard;
            j.append(x);
            return true;
        }
    }
    const uint64_t next = (size_t)tf < j.targets.size() ? j.targets[tf].bit : 0;
    uint64_t want;
    if (j.res[k].status == NEW_TABLE) {
        // stopped at a block boundary, not for want of room: sized like a first-round job
        const uint64_t e = next ? next : r.bit + 32768ull * 8;
        want = std::max<uint64_t>(lane_room((e - r.bit) / 8, ratio), JOB_CAP_MIN);
    } else {
        const uint64_t had = j.kind[k] == JOB_GROUP ? (uint64_t)j.cap[k] * j.lanes : j.cap[k];
        want = 4 * had;
        const uint64_t read_bits = r.bit > j.start[k].bit ? r.bit - j.start[k].bit : 0;
        const uint64_t to = next ? next : j.seg_end;  // the input left to the next target
        const uint64_t ahead = to > r.bit ? (to - r.bit) / 8 : 0;
        if (j.kind[k] == JOB_LANE && read_bits >= 64) {
            const double measured = (double)j.res[k].out_len / ((double)read_bits / 8);
            want = std::max<uint64_t>(want, lane_room(ahead, measured));
        }
        want = std::min<uint64_t>(want, JOB_CAP_MAX);
    }
    j.cont[k] = (int)j.start.size();
    Jobs::Spec x;
    x.start = r;
    x.tfirst = tf;
    x.cap = (uint32_t)std::min<uint64_t>(want, left - reserve);
    j.append(x);
    return true;
}

// One job's share of a segment's output, in stream order. A lane job gives one piece; a group
// gives its chain lanes in order, one piece each (lane l at off + l * cap), the lane trim_lane
// cut to trim_len (the output ends in it).
struct ChainEntry {
    uint64_t off;        // the job's region of the pool
    uint64_t dst;        // where its first symbol goes in the segment's output
    uint32_t cap;        // a group's lane stride; 0: a lane job
    uint32_t len;        // symbols the job gives
    uint32_t piece0;     // index of its first piece
    int32_t job;
    uint64_t chain;      // a group: the lanes that give pieces
    uint32_t trim_lane;  // GROUP_LANES_MAX: none
    uint32_t trim_len;
};
static_assert(sizeof(ChainEntry) == 48, "ChainEntry layout");

struct Pieces {
    std::vector<ChainEntry> jobs;
    uint32_t npieces = 0;
    uint64_t total = 0;
};

// the pieces of the entries (the device builds them with k_pieces): pool index, length, output index
inline void expand_pieces(const Pieces& p, const std::vector<GroupResult>& gres, std::vector<uint64_t>& off,
                          std::vector<uint32_t>& len, std::vector<uint64_t>& dst) {
    off.resize(p.npieces);
    len.resize(p.npieces);
    dst.resize(p.npieces);
    for (const ChainEntry& e : p.jobs) {
        uint32_t q = e.piece0;
        uint64_t d = e.dst;
        if (!e.cap) {
            off[q] = e.off, len[q] = e.len, dst[q] = d;
            continue;
        }
        const GroupResult& g = gres[e.job];
        for (uint64_t c = e.chain; c; c &= c - 1) {
            const uint32_t l = (uint32_t)__builtin_ctzll(c);
            const uint32_t ll = l == e.trim_lane ? e.trim_len : g.len[l];
            off[q] = e.off + (uint64_t)l * e.cap, len[q] = ll, dst[q] = d;
            d += ll;
            q++;
        }
    }
}

// ------------------------------------------------------------------ the output chain
// A segment's output is the chain of pieces that follows from its first job: a job that landed
// hands on to the job at the target it landed on; one that stopped for room, or at a block whose
// table it lacked, hands on to its continuation. How the chain passes through a job depends on
// one bit it carries along: whether the block the job starts in is the stream's last (BFINAL).
// A job that starts inside a block without knowing that block's BFINAL, when the block is the
// last, holds the stream's end: its piece ends at its first end-of-block code.

enum class ChainEnd : uint8_t {
    Finished,  // the stream's last block ended in the chain (end_bit)
    Pending,   // the chain reached a job that stopped, and its continuation is not decoded yet
    Cut,       // the output limit or the segment's stop: the next segment goes on at resume_at
    Broken,    // invalid or truncated data, or pieces that do not join
};

struct Chain {
    ChainEnd end = ChainEnd::Broken;
    Pieces out;            // the chain's pieces, in stream order
    Start resume_at{};     // Cut, Pending: the exact decoder state where the chain stops
    uint64_t end_bit = 0;  // Finished: the bit after the stream's last block
    int pending = -1;      // Pending: the job whose continuation is missing
    int status = 0;        // Broken: the decoder's status (-1: the pieces do not join)
    uint64_t error_bit = 0;
    int error_job = -1;    // Broken: the job whose decode failed (-1: none)
    // Pending: the BFINAL state after the pending job and the jobs visited, for going on from
    // there once its continuation is decoded (follow_chain's `from`)
    int bfinal = 0;
    size_t visits = 0;
};

inline bool needs_continuation(int32_t status) { return status == ERR_OVERFLOW || status == NEW_TABLE; }

// What job k adds to the chain, given the BFINAL state carried into it
struct Link {
    enum Next : uint8_t { TO_TARGET, TO_CONTINUATION, STOP, STREAM_END, ERROR } next;
    bool at_eob;     // STREAM_END: the stream ends at the job's first end-of-block code
    uint32_t len;    // its symbols
    uint32_t reach;  // how far before its first symbol its matches reach (0: they do not)
    int to;          // TO_TARGET: the job that starts at the target it landed on
    uint64_t bit;    // STREAM_END: the stream's end; ERROR: where the decode failed
    int bfinal;      // BFINAL after it (-1: unchanged)
};

inline Link link_of(const Jobs& j, int k, int bfinal) {
    const Result& r = j.res[k];
    const Start& st = j.start[k];
    Link l{};
    l.bfinal = r.end_final;
    if (st.mode != AT_BLOCK && st.final < 0 && bfinal == 1 && r.eob_bit != UINT64_MAX) {
        l.next = Link::STREAM_END;
        l.at_eob = true;
        l.len = r.eob_out;
        l.reach = r.reach_eob;
        l.bit = r.eob_bit;
        return l;
    }
    l.len = r.out_len;
    l.reach = r.reach;
    if (r.status == OK) {
        if (r.landed < 0) {
            l.next = Link::STREAM_END;
            l.bit = r.end_bit;
        } else {
            l.next = Link::TO_TARGET;
            l.to = j.tfirst[k] + r.landed;
        }
    } else if (needs_continuation(r.status)) {
        l.next = Link::TO_CONTINUATION;
    } else if (r.status == STOPPED) {
        l.next = Link::STOP;
    } else {
        l.next = Link::ERROR;
        l.bit = r.end_bit;
    }
    return l;
}

// the chain entry of job k (a group gives its chain lanes in order, trimmed at the chain's first
// end-of-block code when the stream ends there)
inline void append_entry(Pieces& p, const Jobs& j, int k, uint32_t len, bool at_eob) {
    ChainEntry e{};
    e.off = j.off[k];
    e.dst = p.total;
    e.len = len;
    e.piece0 = p.npieces;
    e.job = k;
    e.trim_lane = GROUP_LANES_MAX;
    uint32_t pieces = 1;
    if (j.kind[k] == JOB_GROUP) {
        const GroupHead& g = j.ghead[k];
        e.cap = j.cap[k];
        e.chain = g.chain;
        if (at_eob && g.eob_lane < (uint32_t)GROUP_LANES_MAX) {
            e.chain &= g.eob_lane == 63 ? ~0ull : (2ull << g.eob_lane) - 1;
            e.trim_lane = g.eob_lane;
            e.trim_len = g.eob_len;
        }
        pieces = (uint32_t)__builtin_popcountll(e.chain);
    }
    p.jobs.push_back(e);
    p.npieces += pieces;
    p.total += len;
}

// The chain from job k0 (a batch follows each stream from its first job), at most out_limit
// symbols. `out_before`: the stream's output before this segment, which matches may reach into.
// reuse (optional): an emptied vector whose memory the chain's entries take.
// from (optional): a Pending chain of an earlier round, whose pending job's continuation is
// decoded now: the walk goes on from there (its entries taken over) instead of from k0.
inline Chain follow_chain(const Jobs& j, uint64_t out_limit, uint64_t out_before, int k0 = 0,
                          std::vector<ChainEntry>* reuse = nullptr, Chain* from = nullptr) {
    Chain c;
    int bfinal = j.start[k0].final >= 0 ? j.start[k0].final : 0;
    int k = k0;
    size_t visits = 0;
    if (from && from->end == ChainEnd::Pending && j.cont[from->pending] >= 0) {
        c.out = std::move(from->out);
        c.resume_at = from->resume_at;
        bfinal = from->bfinal;
        visits = from->visits;
        k = j.cont[from->pending];
    } else if (reuse) {
        c.out.jobs.swap(*reuse);
        c.out.jobs.clear();
    }
    // a state inside a block takes the carried BFINAL, unless it records its own
    auto carrying = [&](Start s) {
        if (s.mode != AT_BLOCK && s.final < 0) s.final = (int8_t)bfinal;
        return s;
    };
    auto broken = [&](int status, uint64_t bit, int job) {
        c.end = ChainEnd::Broken;
        c.status = status;
        c.error_bit = bit;
        c.error_job = job;
        return c;
    };
    // every job is visited once at most: more visits mean the links loop
    for (; visits <= j.start.size(); visits++) {
        if (j.res[k].status < 0) return broken(-1, 0, -1);  // not decoded
        const Link l = link_of(j, k, bfinal);
        if (l.reach && (uint64_t)l.reach > out_before + c.out.total) return broken(ERR_DATA, j.start[k].bit, -1);
        if (l.next == Link::ERROR) return broken(j.res[k].status, l.bit, k);
        if (c.out.total + l.len > out_limit) {
            c.end = ChainEnd::Cut;
            c.resume_at = carrying(j.start[k]);
            return c;
        }
        append_entry(c.out, j, k, l.len, l.at_eob);
        if (l.bfinal >= 0) bfinal = l.bfinal;
        switch (l.next) {
            case Link::STREAM_END:
                c.end = ChainEnd::Finished;
                c.end_bit = l.bit;
                return c;
            case Link::STOP:
                c.end = ChainEnd::Cut;
                c.resume_at = carrying(j.res[k].resume);
                return c;
            case Link::TO_CONTINUATION: {
                c.resume_at = carrying(j.res[k].resume);
                const int nx = j.cont[k];
                if (nx < 0 || j.res[nx].status < 0) {
                    c.end = ChainEnd::Pending;
                    c.pending = k;
                    c.bfinal = bfinal;
                    c.visits = visits + 1;
                    return c;
                }
                k = nx;
                break;
            }
            default:  // TO_TARGET
                k = l.to;
        }
    }
    return broken(-1, 0, -1);
}

struct Stats {
    uint64_t in_used = 0, out_size = 0;
    uint32_t segments = 0, rounds = 0;
    uint64_t jobs = 0;
    double seconds = 0, upload_s = 0, find_s = 0, decode_s = 0, resolve_s = 0, download_s = 0;
};

struct Params {
    uint64_t seg_out_max = 256ull << 20;  // output per segment at most
    uint64_t forced_seg_in = 0;           // bytes, 0: from the ratio
    uint64_t seg_in_max = SEG_IN_MAX;     // compressed bytes per segment at most
    uint64_t pool = 0;                    // 16-bit pool symbols
    size_t max_jobs = 0;
    size_t max_probes = 0;
    uint64_t probe_spacing = PROBE_SPACING;  // bits between probe points
    uint64_t chunk_bits = CHUNK_BITS;        // one job per chunk of this many bits at most (>= CHUNK_BITS)
    int group_lanes = GROUP_LANES_DEFAULT;   // lanes of a group job: 16, 32 or 64
    bool debug = false;
};

// Table keys: the sorted header bits with a table in the cache (slot 0 holds the fixed code)
struct TableKeys {
    std::vector<uint64_t> keys;   // sorted
    std::vector<uint32_t> slots;  // slots[i]: cache slot of keys[i]
    uint32_t next_slot = 1;
    uint32_t base = 1;  // the first slot of this segment's tables
    void clear(uint32_t first = 1) {
        keys.clear();
        slots.clear();
        base = next_slot = first;
    }
    // keys whose tables are already in slots base, base + 1, ... (in this order)
    void prebuilt(const std::vector<uint64_t>& ks, uint32_t base) {
        std::vector<uint32_t> order(ks.size());
        std::iota(order.begin(), order.end(), 0u);
        std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return ks[a] < ks[b]; });
        prebuilt_sorted(ks, order, base);
    }
    // the same, with order[] listing the indices of ks by key
    void prebuilt_sorted(const std::vector<uint64_t>& ks, const std::vector<uint32_t>& order, uint32_t base) {
        keys.resize(ks.size());
        slots.resize(ks.size());
        for (size_t i = 0; i < ks.size(); i++) keys[i] = ks[order[i]], slots[i] = base + order[i];
        next_slot = base + (uint32_t)ks.size();
    }
    bool has(uint64_t k) const { return std::binary_search(keys.begin(), keys.end(), k); }
    // adds the keys not yet present; returns (key, slot) of the new ones
    std::vector<std::pair<uint64_t, uint32_t>> add(std::vector<uint64_t> ks) {
        std::sort(ks.begin(), ks.end());
        ks.erase(std::unique(ks.begin(), ks.end()), ks.end());
        std::vector<std::pair<uint64_t, uint32_t>> fresh;
        for (uint64_t k : ks)
            if (k != UINT64_MAX && !has(k)) fresh.push_back({k, next_slot++});
        if (fresh.empty()) return fresh;
        // merge the fresh keys (in key order) into the sorted ones, from the back
        const size_t n0 = keys.size();
        keys.resize(n0 + fresh.size());
        slots.resize(n0 + fresh.size());
        size_t a = n0, b = fresh.size(), o = keys.size();
        while (b) {
            if (a && keys[a - 1] > fresh[b - 1].first) {
                --a, --o;
                keys[o] = keys[a], slots[o] = slots[a];
            } else {
                --b, --o;
                keys[o] = fresh[b].first, slots[o] = fresh[b].second;
            }
        }
        return fresh;
    }
};

// the header whose table a job (or probe) starting at s needs, or UINT64_MAX
inline uint64_t table_key(const Start& s) {
    if (s.mode == IN_DYNAMIC) return s.header;
    if (s.mode == AT_BLOCK) return s.bit;
    return UINT64_MAX;
}

// wall time of the segment loop's phases (reported with Params::debug)
struct PhaseClock {
    enum { FIND, PREP, TABLES, PROBE, JOBS, DECODE, STITCH, RESOLVE, DELIVER, N };
    double t[N] = {};
    std::chrono::steady_clock::time_point at = std::chrono::steady_clock::now();
    void lap(int k) {
        auto now = std::chrono::steady_clock::now();
        t[k] += std::chrono::duration<double>(now - at).count();
        at = now;
    }
    void print() const {
        static const char* names[N] = {"find", "prep", "tables", "probe", "jobs", "decode", "stitch", "resolve", "deliver"};
        fprintf(stderr, "    host phases (ms):");
        for (int k = 0; k < N; k++) fprintf(stderr, " %s %.1f", names[k], t[k] * 1e3);
        fprintf(stderr, "\n");
    }
};

// Error codes of the C interface
enum : int { RC_OK = 0, RC_DATA = 1, RC_TRUNCATED = 2, RC_ABORTED = 3, RC_GPU = 4, RC_ARG = 5 };

// One stream, segment by segment (gpuinflate.cu:638-873). Backend B provides:
//   uint64_t nbits()                                           stream bits
//   void find_launch(uint64_t lo, uint64_t hi)                 starts looking for block-start evidence in
//                                                              (lo, hi); it may run beside other work
//   Found find_wait()                                          that evidence, any order
//   void tables(const std::vector<std::pair<uint64_t,uint32_t>>& fresh, const TableKeys& all)
//   std::vector<Start> probe(const std::vector<Probe>& pts, const TableKeys& keys)
//   void decode(Jobs& j, size_t a, size_t b, const TableKeys& keys, bool stream_start, uint64_t stop, bool first_round)
//   int resolve(const Jobs& j, const Pieces& p, uint32_t W, uint64_t out_before, std::string& err)
//                                                              builds the output
//   bool deliver(uint64_t total, uint32_t W)                   false: the sink stopped
//   void begin_segment(), double stage timers through mark(i)
template <class B>
int inflate_stream(B& be, const Params& P, uint64_t hint, uint64_t in_size, Stats& stats, std::string& err) {
    const uint64_t nbits = in_size * 8;
    Start cur{};
    cur.bit = 0;
    cur.header = 0;
    cur.mode = AT_BLOCK;
    cur.final = -1;
    uint32_t W = 0;          // output bytes before the segment that its references may reach
    uint64_t out_before = 0; // the stream's output before the segment
    double ratio = hint ? std::clamp((double)hint / (double)in_size, 1.0, 1032.0) : 4.0;
    TableKeys keys;
    auto seg_bits = [&]() {
        return 8 * (P.forced_seg_in ? std::min(P.forced_seg_in, P.seg_in_max)
                                    : std::clamp<uint64_t>((uint64_t)(0.6 * (double)P.seg_out_max / ratio), SEG_IN_MIN,
                                                           P.seg_in_max));
    };
    // the block finder runs one segment ahead: the next segment's range (lo, hi) is known as soon
    // as this one's stop bit is (the next segment starts at or just after it)
    uint64_t ahead_lo = UINT64_MAX, ahead_hi = 0;
    // a backend whose finder may be waited for from another thread has the look-ahead report
    // prepared there (B::async_find)
    std::future<Prepared> ahead;
    Jobs spare_jobs;                      // the segment before's jobs, their memory reused
    std::vector<ChainEntry> spare_ents;   // the same for the chain entries
    auto ahead_wait = [&]() { return ahead.valid() ? ahead.get() : prepare_found(be.find_wait()); };
    for (;;) {
        if (be.aborted()) break;
        be.begin_segment();
        Start s0 = cur;
        if (s0.mode == AT_BLOCK) s0.header = s0.bit;
        uint64_t stop_bit;
        PhaseClock pc;
        Prepared pr;
        if (ahead_hi && s0.bit >= ahead_lo && s0.bit < ahead_hi) {
            stop_bit = ahead_hi;
            pr = ahead_wait();
        } else {
            if (ahead_hi) ahead_wait();  // run ahead for a range this segment does not start in
            // a call's first segment is a third of the others: nothing overlaps its search, and
            // the next segment's search runs beside its decode
            const uint64_t sb = stats.segments == 0 && !P.forced_seg_in
                                    ? std::max<uint64_t>(seg_bits() / 3, 8 * SEG_IN_MIN)
                                    : seg_bits();
            stop_bit = std::min(s0.bit + sb, nbits);
            be.find_launch(s0.bit, stop_bit);
            pr = prepare_found(be.find_wait());
        }
        ahead_hi = 0;
        bool last = stop_bit >= nbits;
        uint64_t stop = last ? UINT64_MAX : stop_bit;
        uint64_t end = last ? nbits : stop_bit;
        if (!last) {
            // The next segment at most doubles, so that its search (about 0.4 of a decode and
            // resolve of the same input) ends before this segment's decode and resolve do. Near
            // the end the segments halve: a segment takes half of what is left when that is less
            // than two of them, so the last segments, whose decode rounds and resolve nothing
            // overlaps, are the small ones (measured on a 4 GiB stream: 215, 430, 754, 377 and
            // 377 MB of input, 6% faster than five segments growing by half to an 846 MB last).
            uint64_t next = P.forced_seg_in ? seg_bits() : std::min(seg_bits(), (stop_bit - s0.bit) * 2);
            const uint64_t rest = nbits - stop_bit;
            if (!P.forced_seg_in && rest < 2 * next && rest > 2 * 8 * SEG_IN_MIN) next = rest / 2;
            ahead_lo = stop_bit;
            ahead_hi = std::min(stop_bit + next, nbits);
            if (nbits - ahead_hi < seg_bits() / 4) ahead_hi = nbits;  // no short segment at the end
            be.find_launch(ahead_lo, ahead_hi);
            if (B::async_find) ahead = std::async(std::launch::async, [&be] { return prepare_found(be.find_wait()); });
        }
        pc.lap(PhaseClock::FIND);

        // 1. block starts: every verified dynamic header gets a table; one per chunk a job; the
        //    block after every confirmed stored block starts one too
        Found& f = pr.f;
        keys.clear(f.slot_base);
        if (f.prebuilt) keys.prebuilt_sorted(pr.raw_dyn, pr.dyn_order, f.slot_base);
        std::vector<std::pair<uint64_t, uint32_t>> est = std::move(pr.est);
        const std::vector<uint64_t>& cand_bytes = pr.cand_bytes;
        const size_t nstored_cand = pr.nstored_cand;
        const std::vector<std::pair<uint64_t, uint64_t>>& stored = pr.stored;
        const double t_confirm = pr.confirm_s;
        const auto t_s = std::chrono::steady_clock::now();
        std::vector<Start> starts = block_starts(f.dyn, stored, cand_bytes, s0, end, P.chunk_bits);
        const double t_starts = t_confirm + std::chrono::duration<double>(std::chrono::steady_clock::now() - t_s).count();
        std::vector<uint64_t> want;
        if (!f.prebuilt) want = f.dyn;
        want.push_back(table_key(s0));
        if (P.debug)
            fprintf(stderr, "    prep: %zu stored candidates, %zu confirmed; confirm %.1f ms, starts %.1f ms\n", nstored_cand,
                    stored.size(), t_confirm * 1e3, (t_starts - t_confirm) * 1e3);
        pc.lap(PhaseClock::PREP);
        be.tables(keys.add(want), keys);
        pc.lap(PhaseClock::TABLES);

        // 2. probes
        std::vector<Probe> pts = probe_points(starts, end, P.probe_spacing, stored);
        if (pts.size() > P.max_probes) pts.resize(P.max_probes);
        if (!pts.empty()) {
            std::vector<uint64_t> pk;
            for (auto& p : pts) pk.push_back(p.mode == IN_FIXED ? UINT64_MAX : p.header);
            be.tables(keys.add(pk), keys);
            std::vector<Start> probed = be.probe(pts, keys);
            starts = merge_probed(starts, probed, s0.bit, end);
        }
        be.mark(2);
        pc.lap(PhaseClock::PROBE);

        // 3-4. decode and stitch
        Jobs j = make_jobs(starts, end, ratio, P.pool, P.max_jobs, P.group_lanes, stored, {}, 0, std::move(est),
                           &spare_jobs);
        {
            // probed starts name headers that have tables; a block start that is not a verified
            // dynamic header needs none unless its block turns out dynamic (then the decoder
            // stops there, NEW_TABLE, and its continuation brings the key)
            std::vector<uint64_t> jk;
            for (auto& s : j.start)
                if (s.mode == IN_DYNAMIC) jk.push_back(s.header);
            be.tables(keys.add(jk), keys);
        }
        pc.lap(PhaseClock::JOBS);
        int rounds = 0;
        const size_t first_cont = j.start.size();
        Chain c;
        size_t a = 0;
        bool first_round = true;
        const uint64_t out_limit = P.seg_out_max;
        for (int round = 1;; round++) {
            if (round > MAX_ROUNDS) {
                c.end = ChainEnd::Cut;
                break;
            }
            for (size_t k = a; k < j.start.size(); k++) j.res[k].status = -1;
            be.decode(j, a, j.start.size(), keys, out_before == 0, stop, first_round);
            pc.lap(PhaseClock::DECODE);
            first_round = false;
            rounds++;
            if (c.end == ChainEnd::Pending) {
                c = follow_chain(j, out_limit, out_before, 0, nullptr, &c);  // on from the pending job
            } else {
                spare_ents.swap(c.out.jobs);
                c = follow_chain(j, out_limit, out_before, 0, &spare_ents);
            }
            pc.lap(PhaseClock::STITCH);
            if (c.end != ChainEnd::Pending) break;
            size_t first = j.start.size();
            if (!add_continuation(j, (size_t)c.pending, 0, ratio)) {
                c.end = ChainEnd::Cut;
                break;
            }
            for (size_t k = a; k < first; k++)
                if (needs_continuation(j.res[k].status) && j.cont[k] < 0) add_continuation(j, k, 2ull * JOB_CAP_MAX, ratio);
            std::vector<uint64_t> nk;
            for (size_t k = first; k < j.start.size(); k++) nk.push_back(table_key(j.start[k]));
            be.tables(keys.add(nk), keys);
            pc.lap(PhaseClock::TABLES);
            a = first;
        }
        be.mark(3);
        stats.rounds += (uint32_t)rounds;
        stats.jobs += j.start.size();
        if (P.debug) {
            uint64_t mx = 0, sum = 0, n_over = 0, n_land = 0, n_err = 0;
            for (auto& r : j.res) {
                mx = std::max<uint64_t>(mx, r.out_len);
                sum += r.out_len;
                n_over += needs_continuation(r.status);
                n_land += r.landed >= 0;
                n_err += r.status == ERR_DATA || r.status == ERR_INPUT;
            }
            for (size_t k = 0; k < j.res.size(); k++)
                if ((needs_continuation(j.res[k].status) || k >= first_cont) && getenv("GZ_DEBUG_JOBS"))
                    fprintf(stderr, "      job %zu kind %d at %llu..%llu cap %u: status %d out %u resume %llu end %llu landed %d chain %llx cont %d\n", k,
                            j.kind[k], (unsigned long long)j.start[k].bit, (unsigned long long)j.end[k], j.cap[k],
                            j.res[k].status, j.res[k].out_len, (unsigned long long)j.res[k].resume.bit,
                            (unsigned long long)j.res[k].end_bit, j.res[k].landed, (unsigned long long)j.ghead[k].chain,
                            j.cont[k]);
            fprintf(stderr, "    jobs: output max %llu mean %.0f symbols (pool use %.2fx the chain); %llu landed, %llu continued, %llu errors\n",
                    (unsigned long long)mx, j.res.empty() ? 0.0 : (double)sum / (double)j.res.size(),
                    c.out.total ? (double)sum / (double)c.out.total : 0.0, (unsigned long long)n_land,
                    (unsigned long long)n_over, (unsigned long long)n_err);
        }
        if (P.debug)
            fprintf(stderr, "  segment %u at bit %llu: %zu hits, %zu starts (%zu probes), %zu jobs, %d rounds, %zu pieces, %llu bytes, walk %d\n",
                    stats.segments, (unsigned long long)cur.bit, f.dyn.size(), starts.size(), pts.size(), j.start.size(),
                    rounds, (size_t)c.out.npieces, (unsigned long long)c.out.total, (int)c.end);
        if (c.end == ChainEnd::Broken && P.debug && c.error_job >= 0) {
            const int k = c.error_job;
            const Start& st = j.start[k];
            fprintf(stderr, "    chain fails at job %d (kind %d) from bit %llu mode %d header %llu final %d to %llu: status %d, out %u, "
                            "landed %d, tfirst %d; group chain %llx eob_lane %u\n",
                    k, j.kind[k], (unsigned long long)st.bit, st.mode, (unsigned long long)st.header, st.final,
                    (unsigned long long)j.end[k], j.res[k].status, j.res[k].out_len, j.res[k].landed, j.tfirst[k],
                    (unsigned long long)j.ghead[k].chain, j.ghead[k].eob_lane);
            fprintf(stderr, "      split_end %llu hard %llu cap %u extra %u\n", (unsigned long long)j.split_end[k],
                    (unsigned long long)j.hard[k], j.cap[k], j.extra[k]);
        }
        if (c.end == ChainEnd::Broken) {
            err = c.status == ERR_INPUT ? "the input ends inside the stream"
                  : c.status < 0        ? "the pieces did not chain (internal error)"
                                        : "invalid Deflate data";
            err += " (near byte " + std::to_string(c.error_bit / 8) + ")";
            return c.status == ERR_INPUT ? RC_TRUNCATED : RC_DATA;
        }
        if (c.end == ChainEnd::Cut && c.out.total == 0 && c.resume_at.bit <= cur.bit) {
            err = "no progress (internal error)";
            return RC_DATA;
        }

        // 5. resolve, 6. deliver
        uint64_t total = c.out.total;
        pc.lap(PhaseClock::STITCH);
        int rc = be.resolve(j, c.out, W, out_before, err);
        if (rc) return rc;
        be.mark(4);
        pc.lap(PhaseClock::RESOLVE);
        if (!be.deliver(total, W)) break;
        pc.lap(PhaseClock::DELIVER);
        if (P.debug) pc.print();
        be.mark(5);
        stats.out_size += total;
        stats.segments++;
        be.add_stage_times(stats);
        out_before += total;
        uint32_t newW = (uint32_t)std::min<uint64_t>(WINDOW, out_before);
        if (c.end == ChainEnd::Finished) {
            stats.in_used = (c.end_bit + 7) / 8;
            if (ahead_hi) ahead_wait();
            break;
        }
        Start nx = c.resume_at;
        if (nx.mode == AT_BLOCK) nx.header = nx.bit;
        if (nx.mode == IN_FIXED) nx.header = UINT64_MAX;
        if (nx.bit > cur.bit) ratio = std::clamp((double)total / ((double)(nx.bit - cur.bit) / 8), 1.0, 1032.0);
        W = newW;
        cur = nx;
        spare_jobs = std::move(j);  // the device has its copies of both (the rounds ended with a sync)
        spare_ents.swap(c.out.jobs);
    }
    if (be.finish_delivery()) {
        err = "stopped by the caller";
        return RC_ABORTED;
    }
    return RC_OK;
}

// ------------------------------------------------------------------ batches
// Many whole streams decoded as one segment: packed one after another in one input, searched,
// planned, decoded and resolved by the same launches, each walked from its own exact start. A
// job reads only its own stream's bytes (Jobs::lim), so each stream decodes as it would alone.
// In the batch's output space every stream ends on a BATCH_ALIGN boundary, with zeros before it
// up to the previous one's end: the resolve's CRC tiles then lie within one stream each (leading
// zeros leave a CRC register at zero unchanged).
constexpr uint64_t BATCH_ALIGN = 16384;
constexpr uint64_t BATCH_ZERO = BATCH_ALIGN;  // pool symbols kept at zero: the gaps' source

struct BatchStream {
    uint64_t base = 0;     // its first byte in the packed input
    uint64_t size = 0;     // its bytes
    uint64_t out_cap = 0;  // room for its output
    // results
    int rc = -1;           // RC_*; -1: not finished in the batch (decode it alone)
    std::string err;
    uint64_t out_size = 0, in_used = 0;
    uint64_t out_off = 0;  // its first byte in the batch's output space
    uint32_t crc = 0;      // CRC-32 of its output
};

// Backend B as for inflate_stream, plus
//   void resolve_batch(const Jobs& j, const Pieces& p, std::vector<BatchStream>& bs)
//        builds the output of the finished streams (out_off, out_size) and their CRC-32s
// bs: sorted by base. Streams left with rc -1 ran out of rounds or room here.
template <class B>
void inflate_batch(B& be, const Params& P, std::vector<BatchStream>& bs, uint64_t in_size, Stats& stats) {
    if (bs.empty()) return;
    const uint64_t nbits = in_size * 8;
    be.begin_segment();
    be.find_launch(0, nbits);
    Found f = be.find_wait();
    TableKeys keys;
    keys.clear(f.slot_base);
    if (f.prebuilt) keys.prebuilt(f.dyn, f.slot_base);
    std::vector<std::pair<uint64_t, uint32_t>> est;
    if (f.est.size() == f.dyn.size()) {
        est.resize(f.dyn.size());
        for (size_t i = 0; i < f.dyn.size(); i++) est[i] = {f.dyn[i], f.est[i]};
        std::sort(est.begin(), est.end());
    }
    std::sort(f.dyn.begin(), f.dyn.end());
    std::vector<uint64_t> cand_bytes;
    std::vector<std::pair<uint64_t, uint64_t>> stored = confirm_stored(f.dyn, std::move(f.stored), cand_bytes);
    std::vector<Start> heads(bs.size());
    std::vector<std::pair<uint64_t, uint64_t>> ranges(bs.size());
    uint64_t cap_sum = 0;
    for (size_t i = 0; i < bs.size(); i++) {
        Start h{};
        h.bit = bs[i].base * 8;
        h.header = h.bit;
        h.mode = AT_BLOCK;
        h.final = -1;
        h.pad[1] = 1;  // a stream's start: no match may reach before it
        heads[i] = h;
        ranges[i] = {bs[i].base, bs[i].base + bs[i].size};
        cap_sum += bs[i].out_cap;
    }
    // the streams' heads, then the evidence's starts; a head wins a tie
    std::vector<Start> starts = heads;
    {
        std::vector<Start> ev = block_starts(f.dyn, stored, cand_bytes, heads[0], nbits, P.chunk_bits);
        starts.insert(starts.end(), ev.begin(), ev.end());
        std::stable_sort(starts.begin(), starts.end(), [](const Start& a, const Start& b) { return a.bit < b.bit; });
        std::vector<Start> u;
        for (auto& x : starts)
            if (u.empty() || x.bit != u.back().bit) u.push_back(x);
        starts.swap(u);
    }
    std::vector<uint64_t> want;
    if (!f.prebuilt) want = f.dyn;
    for (auto& h : heads) want.push_back(table_key(h));
    be.tables(keys.add(want), keys);
    std::vector<Probe> pts = probe_points(starts, nbits, P.probe_spacing, stored);
    if (pts.size() > P.max_probes) pts.resize(P.max_probes);
    if (!pts.empty()) {
        std::vector<uint64_t> pk;
        for (auto& p : pts) pk.push_back(p.mode == IN_FIXED ? UINT64_MAX : p.header);
        be.tables(keys.add(pk), keys);
        std::vector<Start> probed = be.probe(pts, keys);
        starts = merge_probed(starts, probed, 0, nbits);
    }
    be.mark(2);
    const double ratio = std::clamp((double)cap_sum / (double)std::max<uint64_t>(in_size, 1), 1.0, 1032.0);
    Jobs j = make_jobs(starts, nbits, ratio, P.pool, P.max_jobs, P.group_lanes, stored, ranges, BATCH_ZERO,
                       std::move(est));
    {
        std::vector<uint64_t> jk;
        for (auto& st : j.start)
            if (st.mode == IN_DYNAMIC) jk.push_back(st.header);
        be.tables(keys.add(jk), keys);
    }
    // each stream's first job: the target at its head
    std::vector<int> first_job(bs.size());
    for (size_t i = 0; i < bs.size(); i++) {
        auto it = std::lower_bound(j.targets.begin(), j.targets.end(), heads[i].bit,
                                   [](const Start& a, uint64_t b) { return a.bit < b; });
        first_job[i] = (int)(it - j.targets.begin());
    }
    std::vector<Chain> chains(bs.size()), pend(bs.size());
    std::vector<char> open(bs.size(), 1);
    size_t a = 0;
    bool first_round = true;
    int rounds = 0;
    for (int round = 1; round <= MAX_ROUNDS; round++) {
        for (size_t k = a; k < j.start.size(); k++) j.res[k].status = -1;
        be.decode(j, a, j.start.size(), keys, false, UINT64_MAX, first_round);
        first_round = false;
        rounds++;
        const size_t first = j.start.size();
        bool more = false;
        for (size_t i = 0; i < bs.size(); i++) {
            if (!open[i]) continue;
            Chain c = follow_chain(j, bs[i].out_cap, 0, first_job[i], nullptr, &pend[i]);
            if (c.end == ChainEnd::Pending) {
                if (j.cont[c.pending] >= 0 || add_continuation(j, (size_t)c.pending, 0, ratio)) {
                    more = true;
                    pend[i] = std::move(c);  // the next round goes on from its pending job
                    continue;
                }
                open[i] = 0;  // no room left here: decoded alone
                continue;
            }
            open[i] = 0;
            BatchStream& b = bs[i];
            const uint64_t b0 = b.base * 8;
            if (c.end == ChainEnd::Finished) {
                b.rc = RC_OK;
                b.out_size = c.out.total;
                b.in_used = (c.end_bit - b0 + 7) / 8;
                chains[i] = std::move(c);
            } else if (c.end == ChainEnd::Broken) {
                b.rc = c.status == ERR_INPUT ? RC_TRUNCATED : RC_DATA;
                b.err = c.status == ERR_INPUT ? "the input ends inside the stream"
                        : c.status < 0        ? "the pieces did not chain (internal error)"
                                              : "invalid Deflate data";
                b.err += " (near byte " + std::to_string(c.error_bit >= b0 ? (c.error_bit - b0) / 8 : 0) + ")";
            } else {
                b.rc = RC_ARG;
                b.err = "the output is larger than the buffer given";
            }
        }
        if (!more) break;
        for (size_t k = a; k < first; k++)
            if (needs_continuation(j.res[k].status) && j.cont[k] < 0) add_continuation(j, k, 2ull * JOB_CAP_MAX, ratio);
        std::vector<uint64_t> nk;
        for (size_t k = first; k < j.start.size(); k++) nk.push_back(table_key(j.start[k]));
        be.tables(keys.add(nk), keys);
        a = first;
    }
    be.mark(3);
    // the output space: each finished stream ends on a BATCH_ALIGN boundary, zeros before it
    Pieces all;
    uint64_t cur = 0;
    for (size_t i = 0; i < bs.size(); i++) {
        BatchStream& b = bs[i];
        if (b.rc != RC_OK) continue;
        const uint64_t n = b.out_size;
        const uint64_t o = n ? (cur + n + BATCH_ALIGN - 1) / BATCH_ALIGN * BATCH_ALIGN - n : cur;
        if (o > cur) {
            ChainEntry z{};
            z.off = 0;
            z.dst = cur;
            z.len = (uint32_t)(o - cur);
            z.piece0 = all.npieces++;
            z.job = -1;
            z.trim_lane = GROUP_LANES_MAX;
            all.jobs.push_back(z);
        }
        for (ChainEntry e : chains[i].out.jobs) {
            e.dst += o;
            e.piece0 += all.npieces;
            all.jobs.push_back(e);
        }
        all.npieces += chains[i].out.npieces;
        b.out_off = o;
        cur = o + n;
        stats.out_size += n;
    }
    all.total = cur;
    be.resolve_batch(j, all, bs);
    be.mark(4);
    be.mark(5);
    stats.rounds += (uint32_t)rounds;
    stats.jobs += j.start.size();
    stats.segments++;
    be.add_stage_times(stats);
}

}  // namespace gz
