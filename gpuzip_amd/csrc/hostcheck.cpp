// hostcheck.cpp - the decoding core (deflate_core.h) and planner (plan.h) built for the host,
// with a CPU backend that does the data-parallel steps as loops: libgzhostcheck.so.
//
// For tests only. It lets the planning, stitching and decoding logic that the GPU library runs
// be checked in a container with no GPU, against the oracle. It is not a decoder for users:
// nothing in the product loads it, and the library the product loads (libgpuinflate.so) has no
// CPU path.
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "deflate_core.h"
#include "plan.h"

#define HC_API extern "C" __attribute__((visibility("default")))

using namespace gz;

namespace {

// fn(i) for i in [0, n) on up to `threads` threads, indices handed out one at a time
template <class F>
void parallel_for(size_t n, unsigned threads, F&& fn) {
    const unsigned nt = (unsigned)std::min<size_t>(std::max(1u, threads), n);
    if (nt <= 1) {
        for (size_t i = 0; i < n; i++) fn(i);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; t++)
        pool.emplace_back([&] {
            for (size_t i; (i = next.fetch_add(1)) < n;) fn(i);
        });
    for (auto& th : pool) th.join();
}

void build_table(const Bits& proto, uint64_t key, Table& t) {
    memset(&t, 0, sizeof t);
    uint8_t lens[320];
    uint16_t codes[320];
    int hl = 0, hd = 0;
    if (key == UINT64_MAX) {
        table_prepare_fixed(t, lens, codes);
        t.status = TAB_FIXED;
        hl = 288;
        hd = 30;
    } else {
        t.status = table_prepare(proto, key, t, lens, codes, hl, hd);
        if (t.status != TAB_DYNAMIC) return;
    }
    for (int s = 0; s < hl; s++) fill_lit_symbol(t, s, lens[s], codes[s]);
    for (int s = 0; s < hd; s++) fill_dist_symbol(t, s, lens[hl + s], codes[hl + s]);
}

struct CpuBackend {
    static constexpr bool async_find = false;  // its finder runs inside find_wait
    unsigned threads = 1;  // the data-parallel steps run on this many host threads
    std::vector<uint32_t> words;  // the input, padded
    uint64_t size;
    std::vector<Table> tabs;
    TableKeys keys;
    // the decoders' 16-bit output, left uninitialised (a decoder writes what it gives) but for
    // the first BATCH_ZERO symbols, a batch's zero source
    struct Pool {
        std::unique_ptr<uint16_t[]> mem;
        size_t n = 0;
        uint16_t* data() { return mem.get(); }
        size_t size() const { return n; }
        uint16_t operator[](size_t i) const { return mem[i]; }
        void resize(size_t m) {
            if (m <= n) return;
            std::unique_ptr<uint16_t[]> grown(new uint16_t[m]);
            if (n) memcpy(grown.get(), mem.get(), n * sizeof(uint16_t));
            for (size_t i = n; i < std::min<size_t>(m, BATCH_ZERO); i++) grown[i] = 0;
            mem = std::move(grown);
            n = m;
        }
    } pool;
    std::vector<uint8_t> out;     // everything delivered
    std::vector<uint8_t> seg;     // [window | segment] bytes of the current segment
    std::vector<uint8_t> window;
    uint64_t limit;
    uint64_t probe_lead = PROBE_LEAD;  // bits a probe decodes before its point

    Bits proto() const {
        Bits b;
        b.setup(words.data(), size);
        return b;
    }
    const Table* lookup(uint64_t key) const {
        auto it = std::lower_bound(keys.keys.begin(), keys.keys.end(), key);
        if (it == keys.keys.end() || *it != key) return nullptr;
        return &tabs[keys.slots[it - keys.keys.begin()]];
    }
    bool aborted() { return false; }
    bool finish_delivery() { return false; }
    void begin_segment() {}
    void mark(int) {}
    void add_stage_times(Stats&) {}

    uint64_t find_lo = 0, find_hi = 0;
    void find_launch(uint64_t lo, uint64_t hi) {
        find_lo = lo;
        find_hi = hi;
    }
    Found find_wait() { return find(find_lo, find_hi); }
    Found find(uint64_t lo, uint64_t hi) {
        Found f;
        if (hi <= lo + 1) return f;
        // the range in pieces of 2^20 words, each searched by one thread; the reports in order
        constexpr uint64_t PART = 1ull << 20;
        const uint64_t w0 = lo / 32, w1 = (hi + 31) / 32;
        const size_t nparts = (size_t)((w1 - w0 + PART - 1) / PART);
        std::vector<Found> parts(nparts);
        const uint8_t* in = (const uint8_t*)words.data();
        const uint64_t byte_hi = std::min<uint64_t>(size, hi / 8 + 8);
        parallel_for(nparts, threads, [&](size_t i) {
            Bits b = proto();
            Found& pf = parts[i];
            const uint64_t a = w0 + i * PART, e = std::min(w1, a + PART);
            for (uint64_t w = a; w < e; w++) {
                uint32_t m = find_word_hits(b.word(w), b.word(w + 1), b.word(w + 2), b.word(w + 3), w, lo, hi);
                while (m) {
                    int o = __builtin_ctz(m);
                    m &= m - 1;
                    uint64_t bit = w * 32 + (uint64_t)o;
                    if (plausible_dynamic_header(b, size * 8, bit)) pf.dyn.push_back(bit);
                }
            }
            // the stored-block candidates at bytes [4a, 4e), clipped to the search's bytes
            const uint64_t blo = std::max<uint64_t>(lo / 8, a * 4), bhi = std::min<uint64_t>(byte_hi, e * 4);
            for (uint64_t at = blo; at < bhi && at + 4 <= byte_hi; at++) {
                uint32_t len = in[at] | ((uint32_t)in[at + 1] << 8), nlen = in[at + 2] | ((uint32_t)in[at + 3] << 8);
                if (len == (~nlen & 0xFFFFu)) pf.stored.push_back((at << 16) | len);
            }
        });
        for (auto& pf : parts) {
            f.dyn.insert(f.dyn.end(), pf.dyn.begin(), pf.dyn.end());
            f.stored.insert(f.stored.end(), pf.stored.begin(), pf.stored.end());
        }
        // bytes past the last word's part (the look-ahead bytes of a block starting near hi)
        for (uint64_t at = std::max<uint64_t>(lo / 8, w1 * 4); at + 4 <= byte_hi; at++) {
            uint32_t len = in[at] | ((uint32_t)in[at + 1] << 8), nlen = in[at + 2] | ((uint32_t)in[at + 3] << 8);
            if (len == (~nlen & 0xFFFFu)) f.stored.push_back((at << 16) | len);
        }
        return f;
    }
    void tables(const std::vector<std::pair<uint64_t, uint32_t>>& fresh, const TableKeys& all) {
        keys = all;
        uint32_t need = all.next_slot;
        if (tabs.size() < need) tabs.resize(need);
        parallel_for(fresh.size(), threads, [&](size_t i) { build_table(proto(), fresh[i].first, tabs[fresh[i].second]); });
    }
    std::vector<Start> probe(const std::vector<Probe>& pts, const TableKeys&) {
        std::vector<Start> r(pts.size());
        parallel_for(pts.size(), threads, [&](size_t i) {
            const Probe& p = pts[i];
            const Table* own = p.mode == IN_FIXED ? nullptr : lookup(p.header);
            uint64_t from = p.point > probe_lead ? p.point - probe_lead : 0;
            uint32_t stg[SW];
            r[i] = gz::probe(words.data(), size, tabs.data(), own, p.mode, p.header, from, p.point, stg, 1);
        });
        return r;
    }
    Job base_job() const {
        Job jb;
        jb.words = words.data();
        jb.size = size;
        jb.tabs = tabs.data();
        jb.keys = keys.keys.data();
        jb.slots = keys.slots.data();
        jb.nkeys = (int)keys.keys.size();
        jb.own = nullptr;
        jb.targets = TargetList{};
        jb.ntargets = 0;
        jb.out = nullptr;
        jb.out_index = 0;
        jb.cap = 0;
        jb.stream_start = false;
        jb.stop = UINT64_MAX;
        jb.fixed = tabs.data();
        jb.hard_end = UINT64_MAX;
        jb.stg = nullptr;
        jb.stride = 1;
        return jb;
    }
    const Table* own_of(const Start& st) const {
        uint64_t key = table_key(st);
        return key == UINT64_MAX ? nullptr : lookup(key);
    }
    // k_groups, one lane after another
    GroupResult run_group(const Jobs& j, size_t k, bool stream_start, uint64_t stop) {
        const Start s = j.start[k];
        const uint64_t end = j.end[k];
        const uint64_t lim = j.lim[k] ? j.lim[k] : size;  // its stream's bytes
        Bits in;
        in.setup(words.data(), lim);
        int kind = 0;
        uint64_t first = s.bit;
        const Table* T = nullptr;
        if (s.mode == IN_FIXED) {
            kind = 1;
        } else if (s.mode == IN_DYNAMIC) {
            T = lookup(s.header);
            kind = T && T->status == TAB_DYNAMIC ? 2 : 0;
        } else {
            in.seek(s.bit);
            uint32_t h3 = in.peek(3);
            if ((h3 >> 1) == 1) {
                kind = 1;
                first = s.bit + 3;
            } else if ((h3 >> 1) == 2) {
                T = lookup(s.bit);
                if (T && T->status == TAB_DYNAMIC) {
                    kind = 2;
                    first = s.bit + T->hdr_bits;
                }
            }
        }
        const uint64_t header0 = s.mode == AT_BLOCK ? s.bit : s.header;
        const int L = j.lanes;
        Start lst[GROUP_LANES_MAX];
        lst[0] = s;
        for (int l = 1; l < L; l++) {
            Start st{};
            st.bit = UINT64_MAX;
            st.header = UINT64_MAX;
            st.mode = AT_BLOCK;
            st.final = -1;
            if (kind && j.split_end[k] > first) {
                uint64_t point = group_point(first, j.split_end[k], l, L);
                uint64_t from = point > first + probe_lead ? point - probe_lead : first;
                uint32_t stg[SW];
                st = gz::probe(words.data(), lim, tabs.data(), kind == 2 ? T : nullptr, s.mode, header0, from, point,
                               stg, 1, 4);  // the GPU groups' attempts
                // BFINAL stays unknown: a probe may have synchronised in a later block with the same
                // code (any fixed block); the chain walk learns it from the lane before
            }
            lst[l] = st;
        }
        Start comb[GROUP_LANES_MAX + GROUP_TARGETS];
        int lane_of[GROUP_LANES_MAX], first_t[GROUP_LANES_MAX];
        int nv = 0;
        uint64_t prev = s.bit;
        first_t[0] = 0;
        for (int l = 1; l < L; l++) {
            if (lst[l].bit != UINT64_MAX && lst[l].bit > prev && lst[l].bit < end) {
                comb[nv] = lst[l];
                lane_of[nv] = l;
                nv++;
                first_t[l] = nv;
                prev = lst[l].bit;
            } else {
                first_t[l] = -1;
            }
        }
        int ngt = std::min<int>(GROUP_TARGETS, std::max<int>(0, (int)j.targets.size() - j.tfirst[k]));
        for (int t = 0; t < ngt; t++) comb[nv + t] = j.targets[j.tfirst[k] + t];
        Result lres[GROUP_LANES_MAX];
        for (int l = 0; l < L; l++) {
            if (first_t[l] < 0) continue;
            Job jb = base_job();
            jb.size = lim;
            jb.s = lst[l];
            uint64_t key = table_key(lst[l]);
            jb.own = kind == 2 && key == header0 ? T : own_of(lst[l]);
            jb.targets = target_list(comb + first_t[l], nv + ngt - first_t[l]);
            jb.ntargets = nv + ngt - first_t[l];
            jb.out_index = j.off[k] + (uint64_t)l * j.cap[k];
            jb.out = pool.data() + jb.out_index;
            jb.cap = !(kind && j.split_end[k] > first) && l == 0 ? j.cap[k] * L + j.extra[k]
                                                                  : j.cap[k] + (l == L - 1 ? j.extra[k] : 0);
            jb.hard_end = j.hard[k] ? j.hard[k] : UINT64_MAX;
            jb.stream_start = l == 0 && ((stream_start && k == 0) || s.pad[1] == 1);
            jb.stop = stop;
            uint32_t stg[SW];
            jb.stg = stg;
            lres[l] = run(jb, nullptr);
        }
        GroupResult g;
        group_walk(lst, lres, lane_of, first_t, nv, L, g);
        return g;
    }
    std::vector<GroupResult> gres;  // the groups' whole results, by job (their lanes' lengths included)
    void decode(Jobs& j, size_t a, size_t b, const TableKeys&, bool stream_start, uint64_t stop, bool) {
        if (pool.size() < std::max(j.pool, j.used)) pool.resize(std::max(j.pool, j.used));
        if (gres.size() < b) gres.resize(b);
        // jobs write their own pool regions and results: any order, any thread
        parallel_for(b - a, threads, [&](size_t i) {
            const size_t k = a + i;
            if (j.kind[k] == JOB_GROUP) {
                const GroupResult g = run_group(j, k, stream_start, stop);
                memcpy((void*)&gres[k], (const void*)&g, sizeof g);
                memcpy((void*)&j.ghead[k], (const void*)&g, sizeof(GroupHead));
                j.res[k] = g.r;
                return;
            }
            Job jb = base_job();
            if (j.lim[k]) jb.size = j.lim[k];
            jb.s = j.start[k];
            jb.own = own_of(jb.s);
            jb.targets = target_list(j.targets.data() + j.tfirst[k], (int)j.targets.size() - j.tfirst[k]);
            jb.ntargets = (int)j.targets.size() - j.tfirst[k];
            jb.out = pool.data() + j.off[k];
            jb.out_index = j.off[k];
            jb.cap = j.cap[k];
            jb.stream_start = (stream_start && k == 0) || jb.s.pad[1] == 1;
            jb.stop = stop;
            uint32_t stg[SW];
            jb.stg = stg;
            j.res[k] = run(jb, nullptr);
        });
    }
    // Markers to bytes in tiles of RTILE output bytes, three passes: (1) every tile at once
    // resolves what refers inside it and notes, for each byte that refers before it, how far
    // before the tile its source lies; (2) tile by tile in order, those of its last WINDOW bytes
    // (all that the next tile can reach) take their bytes; (3) every tile at once, the rest.
    static constexpr uint64_t RTILE = 1ull << 20;
    int resolve(const Jobs& j, const Pieces& ps, uint32_t W, uint64_t, std::string& err) {
        seg.assign(window.end() - W, window.end());
        seg.resize(W + ps.total);
        std::vector<uint64_t> poff, pdst;
        std::vector<uint32_t> plen;
        expand_pieces(ps, gres, poff, plen, pdst);
        const uint64_t total = ps.total;
        const size_t ntiles = (size_t)((total + RTILE - 1) / RTILE);
        std::vector<std::vector<std::pair<uint32_t, uint32_t>>> pending(ntiles);  // (offset in tile, distance before it)
        std::atomic<bool> reach_error{false};
        parallel_for(ntiles, threads, [&](size_t t) {
            const uint64_t t0 = (uint64_t)t * RTILE, t1 = std::min(total, t0 + RTILE);
            std::vector<uint32_t> before(t1 - t0, 0);  // 0: resolved; else how far before the tile its source is
            size_t q = (size_t)(std::upper_bound(pdst.begin(), pdst.end(), t0) - pdst.begin()) - 1;
            for (uint64_t x = t0; x < t1; x++) {
                while (x >= pdst[q] + plen[q]) q++;
                const uint16_t v = pool[poff[q] + (x - pdst[q])];
                const uint64_t pos = W + x, i = x - t0;
                if (v <= REF) {
                    seg[pos] = (uint8_t)v;
                    continue;
                }
                const uint64_t d = v - REF;
                if (d > pos) {
                    reach_error = true;
                    continue;
                }
                if (d <= i) {
                    if (before[i - d] == 0)
                        seg[pos] = seg[pos - d];
                    else
                        before[i] = before[i - d];
                } else {
                    before[i] = (uint32_t)(d - i);
                }
                if (before[i]) pending[t].push_back({(uint32_t)i, before[i]});
            }
        });
        if (reach_error) {
            err = "invalid Deflate data (a match reaches before the window)";
            return RC_DATA;
        }
        for (size_t t = 0; t < ntiles; t++) {
            const uint64_t base = W + (uint64_t)t * RTILE, len = std::min(total - (uint64_t)t * RTILE, RTILE);
            for (auto& [i, k] : pending[t])
                if (i + WINDOW >= len) seg[base + i] = seg[base - k];
        }
        parallel_for(ntiles, threads, [&](size_t t) {
            const uint64_t base = W + (uint64_t)t * RTILE, len = std::min(total - (uint64_t)t * RTILE, RTILE);
            for (auto& [i, k] : pending[t])
                if (i + WINDOW < len) seg[base + i] = seg[base - k];
        });
        return RC_OK;
    }
    // a batch: the pieces into one output space, then each finished stream's bytes and CRC-32
    std::vector<std::vector<uint8_t>> batch_out;
    void resolve_batch(const Jobs& j, const Pieces& ps, std::vector<BatchStream>& bs) {
        std::vector<uint8_t> space(ps.total);
        std::vector<uint64_t> poff, pdst;
        std::vector<uint32_t> plen;
        expand_pieces(ps, gres, poff, plen, pdst);
        std::vector<uint8_t> cover(ps.total, 0);
        for (size_t q = 0; q < poff.size(); q++)
            for (uint32_t i = 0; i < plen[q]; i++) {
                uint16_t v = pool[poff[q] + i];
                uint64_t pos = pdst[q] + i;
                space[pos] = v <= REF ? (uint8_t)v : space[pos - (v - REF)];
                cover[pos] = 1;
            }
        for (uint64_t p = 0; p < ps.total; p++)
            if (!cover[p]) throw std::runtime_error("batch output position not covered");
        batch_out.assign(bs.size(), {});
        for (size_t i = 0; i < bs.size(); i++) {
            if (bs[i].rc != RC_OK) continue;
            batch_out[i].assign(space.begin() + bs[i].out_off, space.begin() + bs[i].out_off + bs[i].out_size);
            for (uint64_t p = (bs[i].out_off / BATCH_ALIGN) * BATCH_ALIGN; p < bs[i].out_off; p++)
                if (space[p]) throw std::runtime_error("batch gap not zero");
            if ((bs[i].out_off + bs[i].out_size) % BATCH_ALIGN) throw std::runtime_error("batch stream not aligned");
            uint32_t c = 0xFFFFFFFFu;
            for (uint8_t b : batch_out[i]) {
                c ^= b;
                for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            }
            bs[i].crc = ~c;
        }
    }

    bool overflow = false;
    bool deliver(uint64_t total, uint32_t W) {
        if (out.size() + total > limit) {
            overflow = true;
            return false;
        }
        if (out.capacity() < out.size() + total) out.reserve(std::max<size_t>(2 * out.capacity(), out.size() + total));
        out.insert(out.end(), seg.begin() + W, seg.begin() + W + total);
        size_t keep = std::min<size_t>(WINDOW, seg.size());
        window.assign(seg.end() - keep, seg.end());
        return true;
    }
};

}  // namespace

// the same segment loop as the GPU library, on the CPU. Returns the RC_* code.
// The decoder's schedule for the calls after this one (the fields of gi_config_tuned that the
// planner and probes use; 0 keeps a default): results must not depend on it
static uint64_t g_chunk_bits = CHUNK_BITS, g_probe_spacing = PROBE_SPACING, g_probe_lead = PROBE_LEAD;
static int g_group_lanes = GROUP_LANES_DEFAULT;
static unsigned g_threads = 1;
// host threads for the data-parallel steps of the calls after this one (results do not depend on it)
HC_API void hc_set_threads(unsigned n) { g_threads = std::max(1u, n); }
HC_API void hc_set_schedule(uint32_t chunk_kib, uint32_t probe_spacing_kbit, uint32_t probe_lead_bits,
                            uint32_t lanes_per_group) {
    g_group_lanes = group_lanes_ok((int)lanes_per_group) ? (int)lanes_per_group : GROUP_LANES_DEFAULT;
    g_chunk_bits = chunk_kib ? (uint64_t)chunk_kib * 8192 : CHUNK_BITS;
    g_probe_spacing = probe_spacing_kbit ? (uint64_t)probe_spacing_kbit << 10 : PROBE_SPACING;
    g_probe_lead = probe_lead_bits ? probe_lead_bits : PROBE_LEAD;
}

HC_API int hc_inflate(const uint8_t* in, uint64_t size, uint64_t hint, uint64_t seg_out, uint64_t seg_in_kb,
                      uint8_t* out, uint64_t out_cap, uint64_t* out_len, uint64_t* in_used, uint32_t* segments,
                      uint32_t* rounds, char* err, size_t err_len) {
    CpuBackend be;
    be.threads = g_threads;
    be.size = size;
    be.words.assign((size + 3) / 4 + 8, 0);
    memcpy(be.words.data(), in, size);
    be.limit = out_cap;
    Table fixed;
    build_table(be.proto(), UINT64_MAX, fixed);
    be.tabs.push_back(fixed);
    Params P;
    P.seg_out_max = seg_out;
    P.forced_seg_in = seg_in_kb << 10;
    P.debug = getenv("GI_DEBUG") != nullptr;
    const char* sim = getenv("GZ_SEG_IN_MB");
    P.seg_in_max = sim && *sim ? strtoull(sim, nullptr, 10) << 20 : SEG_IN_MAX;
    uint64_t nchunks_max = P.seg_in_max * 8 / CHUNK_BITS + 2;
    P.max_jobs = (size_t)(4 * nchunks_max + 8192);
    P.max_probes = P.max_jobs / 2;
    P.chunk_bits = g_chunk_bits;
    P.probe_spacing = g_probe_spacing;
    P.group_lanes = g_group_lanes;
    be.probe_lead = g_probe_lead;
    P.pool = std::max<uint64_t>(seg_out / 2 * 5, 64ull * JOB_CAP_MIN);
    Stats st;
    std::string e;
    int rc;
    if (size == 0) {
        rc = RC_TRUNCATED;
        e = "empty input";
    } else {
        rc = inflate_stream(be, P, hint, size, st, e);
    }
    if (rc == RC_OK && be.overflow) {
        rc = RC_ARG;
        e = "the output is larger than the buffer given";
    }
    *out_len = be.out.size();
    if (rc == RC_OK) memcpy(out, be.out.data(), be.out.size());
    *in_used = st.in_used;
    if (segments) *segments = st.segments;
    if (rounds) *rounds = st.rounds;
    if (err && err_len) snprintf(err, err_len, "%s", e.c_str());
    return rc;
}

// A batch segment (plan.h inflate_batch) on the CPU: streams[i] = in[offs[i] .. offs[i] + sizes[i]),
// packed as the GPU library packs them (16-byte aligned, at least 16 zero bytes apart). Per
// stream: rcs[i] (RC_*, -1 not finished in the batch), out_sizes, in_used, crcs; the finished
// streams' bytes go to out one after another in stream order. Returns 0, or -1 with err set.
HC_API int hc_inflate_batch(const uint8_t* in, const uint64_t* offs, const uint64_t* sizes, const uint64_t* caps, int n,
                            uint64_t seg_out, int* rcs, uint64_t* out_sizes, uint64_t* in_used, uint32_t* crcs,
                            uint8_t* out, uint64_t out_cap, char* err, size_t err_len) {
    try {
        std::vector<BatchStream> bs(n);
        uint64_t at = 0;
        for (int i = 0; i < n; i++) {
            bs[i].base = at;
            bs[i].size = sizes[i];
            bs[i].out_cap = caps[i];
            at = (at + sizes[i] + 15) / 16 * 16 + 16;
        }
        CpuBackend be;
        be.size = at;
        be.words.assign((at + 3) / 4 + 8, 0);
        for (int i = 0; i < n; i++) memcpy((uint8_t*)be.words.data() + bs[i].base, in + offs[i], sizes[i]);
        Table fixed;
        build_table(be.proto(), UINT64_MAX, fixed);
        be.tabs.push_back(fixed);
        Params P;
        P.seg_out_max = seg_out;
        P.seg_in_max = SEG_IN_MAX;
        uint64_t nchunks_max = P.seg_in_max * 8 / CHUNK_BITS + 2;
        P.max_jobs = (size_t)(4 * nchunks_max + 8192);
        P.max_probes = P.max_jobs / 2;
        P.chunk_bits = g_chunk_bits;
        P.probe_spacing = g_probe_spacing;
        P.group_lanes = g_group_lanes;
        be.probe_lead = g_probe_lead;
        P.pool = std::max<uint64_t>(seg_out / 2 * 5, 64ull * JOB_CAP_MIN);
        Stats st;
        inflate_batch(be, P, bs, at, st);
        uint64_t o = 0;
        for (int i = 0; i < n; i++) {
            rcs[i] = bs[i].rc;
            out_sizes[i] = bs[i].rc == RC_OK ? bs[i].out_size : 0;
            in_used[i] = bs[i].rc == RC_OK ? bs[i].in_used : 0;
            crcs[i] = bs[i].rc == RC_OK ? bs[i].crc : 0;
            if (bs[i].rc != RC_OK) continue;
            if (o + bs[i].out_size > out_cap) throw std::runtime_error("output buffer too small");
            memcpy(out + o, be.batch_out[i].data(), bs[i].out_size);
            o += bs[i].out_size;
        }
        if (err && err_len) err[0] = 0;
        return 0;
    } catch (const std::exception& e) {
        if (err && err_len) snprintf(err, err_len, "%s", e.what());
        return -1;
    }
}

// The rescue's header search (gi_rescue_headers' steps) on the CPU: every plausible header from
// `from` on gets a table, then each one before 8 * size - 128 is verified by a decode that writes
// nothing. Returns how many headers are confirmed; the first `cap` go to out.
HC_API uint64_t hc_rescue_headers(const uint8_t* in, uint64_t size, uint64_t from, uint64_t* out, uint64_t cap) {
    CpuBackend be;
    be.size = size;
    be.words.assign((size + 3) / 4 + 8, 0);
    memcpy(be.words.data(), in, size);
    Table fixed;
    build_table(be.proto(), UINT64_MAX, fixed);
    be.tabs.push_back(fixed);
    const uint64_t nbits = size * 8, end = nbits > 128 ? nbits - 128 : 0;
    if (from == 0 || from >= end) return 0;
    Found f = be.find(from - 1, nbits);
    TableKeys keys;
    keys.clear(1);
    be.tables(keys.add(f.dyn), keys);
    std::sort(f.dyn.begin(), f.dyn.end());
    uint64_t k = 0;
    for (uint64_t b : f.dyn) {
        if (b < from || b >= end) continue;
        Job jb = be.base_job();
        jb.s.bit = b;
        jb.s.header = b;
        jb.s.mode = AT_BLOCK;
        jb.s.final = -1;
        jb.own = be.lookup(b);
        jb.cap = 256u << 10;
        uint32_t stg[SW];
        jb.stg = stg;
        Result r = run<false>(jb, nullptr);
        const bool ok = r.status == OK ||
                        (r.status == ERR_OVERFLOW && (r.blocks >= 2 || r.out_len >= (256u << 10) - MAX_SYMBOL_OUT));
        if (ok) {
            if (k < cap) out[k] = b;
            k++;
        }
    }
    return k;
}

// the block finder's hits in (lo, hi): the word-sliced prefilter, then zlib's header checks
HC_API uint64_t hc_find(const uint8_t* in, uint64_t size, uint64_t lo, uint64_t hi, uint64_t* out, uint64_t cap) {
    CpuBackend be;
    be.size = size;
    be.words.assign((size + 3) / 4 + 8, 0);
    memcpy(be.words.data(), in, size);
    std::vector<uint64_t> h = be.find(lo, hi).dyn;
    for (size_t i = 0; i < h.size() && i < cap; i++) out[i] = h[i];
    return h.size();
}

// the prefilter alone over (lo, hi)
HC_API uint64_t hc_prefilter(const uint8_t* in, uint64_t size, uint64_t lo, uint64_t hi, uint64_t* out, uint64_t cap) {
    std::vector<uint32_t> words((size + 3) / 4 + 8, 0);
    memcpy(words.data(), in, size);
    Bits b;
    b.setup(words.data(), size);
    uint64_t n = 0;
    for (uint64_t w = lo / 32; w < (hi + 31) / 32; w++) {
        uint32_t m = find_word_hits(b.word(w), b.word(w + 1), b.word(w + 2), b.word(w + 3), w, lo, hi);
        while (m) {
            int o = __builtin_ctz(m);
            m &= m - 1;
            if (n < cap) out[n] = w * 32 + (uint64_t)o;
            n++;
        }
    }
    return n;
}

HC_API uint32_t hc_sizeof(int what) {
    switch (what) {
        case 0: return sizeof(Start);
        case 1: return sizeof(Result);
        case 2: return sizeof(Table);
        default: return 0;
    }
}

// For each point: the first symbol boundary at or after it that a decoder from the block
// header reaches (truth), and what a probe started `lead` bits before the point returns.
HC_API void hc_probe_check(const uint8_t* in, uint64_t size, uint64_t header, const uint64_t* points, int n,
                           uint64_t lead, uint64_t* truth, uint64_t* probed) {
    std::vector<uint32_t> words((size + 3) / 4 + 8, 0);
    memcpy(words.data(), in, size);
    Bits b;
    b.setup(words.data(), size);
    std::vector<Table> tabs(2);
    build_table(b, UINT64_MAX, tabs[0]);
    build_table(b, header, tabs[1]);
    std::vector<uint16_t> out(1u << 24);
    for (int i = 0; i < n; i++) {
        Job jb;
        jb.words = words.data();
        jb.size = size;
        jb.s = Start{};
        jb.s.bit = header;
        jb.s.header = header;
        jb.s.mode = AT_BLOCK;
        jb.s.final = -1;
        jb.tabs = tabs.data();
        jb.keys = nullptr;
        jb.slots = nullptr;
        jb.nkeys = 0;
        jb.own = &tabs[1];
        jb.targets = TargetList{};
        jb.ntargets = 0;
        jb.out = out.data();
        jb.out_index = 0;
        jb.cap = (uint32_t)out.size();
        jb.stream_start = false;
        jb.stop = points[i];
        jb.fixed = tabs.data();
        jb.hard_end = UINT64_MAX;
        uint32_t stg[SW];
        jb.stg = stg;
        jb.stride = 1;
        Result r = run(jb, nullptr);
        truth[i] = r.status == STOPPED ? r.resume.bit : UINT64_MAX;
        uint64_t from = points[i] > lead ? points[i] - lead : 0;
        probed[i] = probe(words.data(), size, tabs.data(), &tabs[1], AT_BLOCK, header, from, points[i], stg, 1).bit;
    }
}

// the planner's bucketed membership test over a sorted vector: out[i] = queries[i] is in values
HC_API void hc_sorted_set_has(const uint64_t* values, uint64_t n, const uint64_t* queries, uint64_t nq, uint8_t* out) {
    const std::vector<uint64_t> v(values, values + n);
    const SortedSet set(v);
    for (uint64_t i = 0; i < nq; i++) out[i] = set.has(queries[i]) ? 1 : 0;
}

// the finder's forms of the code-length completeness test, for a check that they agree (use_lut
// 0: direct sum, 1: 16-bit table, 2: the GPU's 64-bit form with 8-bit saturated sums)
HC_API int hc_cl_complete(uint64_t v1, uint64_t v2, int use_lut) {
    static uint16_t k4[4096];
    static uint8_t k4s[4096];  // the GPU finder's form: sums saturated at 255
    static bool ready = false;
    if (!ready) {
        for (uint32_t k = 0; k < 4096; k++) {
            k4[k] = (uint16_t)kraft4(k);
            k4s[k] = (uint8_t)std::min<uint32_t>(kraft4(k), 255u);
        }
        ready = true;
    }
    if (use_lut == 2) return cl_code_complete_g((v1 >> 13) | (v2 << 51), k4s);
    return use_lut ? cl_code_complete_lut(v1, v2, k4) : cl_code_complete(v1, v2);
}

// both forms of the dynamic-header predicate at one bit
HC_API int hc_plausible(const uint8_t* in, uint64_t size, uint64_t bit, int lean) {
    std::vector<uint32_t> words((size + 3) / 4 + 8, 0);
    memcpy(words.data(), in, size);
    Bits b;
    b.setup(words.data(), size);
    uint8_t work[64];
    if (lean == 2) {
        // the staged reader the device uses
        uint32_t stg[SW];
        VBits v;
        v.s.setup(words.data(), size, stg, 1);
        return plausible_dynamic_header_lean(v, size * 8, bit, work);
    }
    return lean ? plausible_dynamic_header_lean(b, size * 8, bit, work) : plausible_dynamic_header(b, size * 8, bit);
}
