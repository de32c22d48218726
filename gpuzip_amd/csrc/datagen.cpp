// datagen.cpp - seeded synthetic inputs for the decode workloads, and a parallel zlib
// compressor, as a small C library (libgzgen.so) for tests and the benchmark.
//
// Content (the same bytes on every machine for a seed):
//   log text   lines of U{40..160} bytes ending in LF; space-separated tokens, each a word drawn
//              Zipf(s = 1.1) by rank from a seeded 4096-word vocabulary (lowercase ASCII,
//              lengths U{3..10}), or with probability 0.2 a decimal number U[0, 10^6);
//   mixed      log text interleaved with uniform random byte runs of U[64 KiB, 4 MiB]; each run
//              follows a text stretch of 4 * run * U[0.5, 1.5] bytes, so about 20% of the bytes
//              are random (deflate stores those: stored blocks among dynamic and fixed ones).
// Text is made in independent pages of 1 MiB (seeded by page index), so generation runs on
// all cores and any prefix of a stream is the same whatever its total length.
//
// Compression: zlib deflate, raw (windowBits -15) with a 32 KiB window at the given level. With
// chunk > 0 the input is cut into chunks compressed in parallel, each primed with the 32 KiB
// before it (deflateSetDictionary) and ended with a sync flush (an empty stored block), as
// pigz does: one valid Deflate stream, byte-identical output for the same chunk size.
#include <stdint.h>
#include <string.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#define GG_API extern "C" __attribute__((visibility("default")))

namespace {

struct Rng {  // splitmix64
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

uint64_t mix(uint64_t a, uint64_t b) {
    Rng r{a * 0x2545F4914F6CDD1Dull ^ (b + 0x632BE59BD9B4E019ull)};
    r.next();
    return r.next();
}

constexpr int VOCAB = 4096;
constexpr double ZIPF_S = 1.1;
constexpr double NUM_P = 0.2;
constexpr uint64_t PAGE = 1u << 20;

struct Vocab {
    std::vector<std::string> words;
    std::vector<double> cdf;
    explicit Vocab(uint64_t seed) {
        Rng r{mix(seed, 0xA0CAB)};
        words.resize(VOCAB);
        for (auto& w : words) {
            int len = 3 + (int)r.below(8);
            w.resize(len);
            for (auto& c : w) c = (char)('a' + r.below(26));
        }
        cdf.resize(VOCAB);
        double sum = 0;
        for (int k = 0; k < VOCAB; k++) cdf[k] = (sum += 1.0 / std::pow((double)(k + 1), ZIPF_S));
        for (auto& c : cdf) c /= sum;
    }
    const std::string& draw(Rng& r) const {
        double u = r.unit();
        int k = (int)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        return words[std::min(k, VOCAB - 1)];
    }
};

// one page of log text: whole lines, the last one cut at the page end
void text_page(const Vocab& v, uint64_t seed, uint64_t page, uint8_t* out, uint64_t n) {
    Rng r{mix(seed, page + 1)};
    uint64_t at = 0;
    char num[16];
    while (at < n) {
        uint64_t L = 40 + r.below(121);  // line length including LF
        uint64_t used = 0;
        uint64_t body = L - 1;
        while (used < body && at < n) {
            const char* tok;
            size_t tl;
            if (r.unit() < NUM_P) {
                int k = snprintf(num, sizeof num, "%llu", (unsigned long long)r.below(1000000));
                tok = num;
                tl = (size_t)k;
            } else {
                const std::string& w = v.draw(r);
                tok = w.data();
                tl = w.size();
            }
            if (used) {
                out[at++] = ' ';
                used++;
                if (used >= body || at >= n) break;
            }
            size_t take = std::min<uint64_t>({tl, body - used, n - at});
            memcpy(out + at, tok, take);
            at += take;
            used += take;
        }
        if (at < n) out[at++] = '\n';
    }
}

template <class F>
void parallel(uint64_t n, int threads, F&& f) {
    threads = std::max(1, std::min<int>(threads, (int)std::min<uint64_t>(n, 256)));
    std::atomic<uint64_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++)
        pool.emplace_back([&] {
            for (uint64_t i; (i = next.fetch_add(1)) < n;) f(i);
        });
    for (auto& th : pool) th.join();
}

int default_threads() {
    unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(h, 16u));
}

}  // namespace

// log text of n bytes
GG_API void gg_text(uint64_t seed, uint64_t n, uint8_t* out, int threads) {
    if (threads <= 0) threads = default_threads();
    Vocab v(seed);
    uint64_t pages = (n + PAGE - 1) / PAGE;
    parallel(pages, threads, [&](uint64_t p) { text_page(v, seed, p, out + p * PAGE, std::min(PAGE, n - p * PAGE)); });
}

// log text interleaved with random runs (random_frac of the bytes on average, runs of
// [run_min, run_max] bytes); 0 random_frac gives plain text
GG_API void gg_mixed(uint64_t seed, uint64_t n, double random_frac, uint64_t run_min, uint64_t run_max, uint8_t* out,
                     int threads) {
    if (threads <= 0) threads = default_threads();
    Vocab v(seed);
    struct Piece {
        uint64_t at, len;
        bool random;
        uint64_t key;
    };
    std::vector<Piece> plan;
    Rng r{mix(seed, 0xB1A9)};
    double text_per_random = random_frac > 0 ? (1.0 - random_frac) / random_frac : 0;
    uint64_t at = 0, k = 0;
    while (at < n) {
        uint64_t run = random_frac > 0 ? run_min + r.below(run_max - run_min + 1) : n;
        uint64_t text = random_frac > 0 ? (uint64_t)(text_per_random * (double)run * (0.5 + r.unit())) : n;
        uint64_t tl = std::min(text, n - at);
        // text pieces are cut into pages so they generate in parallel
        for (uint64_t o = 0; o < tl; o += PAGE) plan.push_back({at + o, std::min(PAGE, tl - o), false, k++});
        at += tl;
        if (at >= n || random_frac <= 0) break;
        uint64_t rl = std::min(run, n - at);
        for (uint64_t o = 0; o < rl; o += PAGE) plan.push_back({at + o, std::min(PAGE, rl - o), true, k++});
        at += rl;
    }
    parallel(plan.size(), threads, [&](uint64_t i) {
        const Piece& p = plan[i];
        if (p.random) {
            Rng q{mix(seed ^ 0x5EED, p.key)};
            uint64_t j = 0;
            for (; j + 8 <= p.len; j += 8) {
                uint64_t x = q.next();
                memcpy(out + p.at + j, &x, 8);
            }
            for (uint64_t x = q.next(); j < p.len; j++, x >>= 8) out[p.at + j] = (uint8_t)x;
        } else {
            text_page(v, seed, p.key, out + p.at, p.len);
        }
    });
}

// Raw Deflate of in[0..n) at `level` (strategy: zlib's), windowBits -15, memLevel 8. chunk 0:
// one deflate call sequence; else pigz-style parallel chunks. Returns the compressed size, or
// 0 when out_cap is too small or zlib fails.
GG_API uint64_t gg_deflate(const uint8_t* in, uint64_t n, int level, int strategy, uint64_t chunk, int threads,
                           uint8_t* out, uint64_t out_cap) {
    if (threads <= 0) threads = default_threads();
    if (chunk == 0 || chunk >= n) chunk = n ? n : 1;
    uint64_t nch = n ? (n + chunk - 1) / chunk : 1;
    std::vector<std::vector<uint8_t>> parts(nch);
    std::atomic<bool> fail{false};
    parallel(nch, nch == 1 ? 1 : threads, [&](uint64_t c) {
        uint64_t a = c * chunk, len = std::min(chunk, n - a);
        z_stream z{};
        if (deflateInit2(&z, level, Z_DEFLATED, -15, 8, strategy) != Z_OK) {
            fail = true;
            return;
        }
        if (a > 0) {
            uint64_t d = std::min<uint64_t>(a, 32768);
            deflateSetDictionary(&z, in + a - d, (uInt)d);
        }
        std::vector<uint8_t>& o = parts[c];
        o.resize(deflateBound(&z, (uLong)std::min<uint64_t>(len, 1ull << 30)) + 64 + len / 1000);
        uint64_t done_in = 0;
        size_t used = 0;
        int flush_end = c + 1 == nch ? Z_FINISH : Z_SYNC_FLUSH;
        for (;;) {
            uint64_t step = std::min<uint64_t>(len - done_in, 1ull << 30);
            z.next_in = (Bytef*)(in + a + done_in);
            z.avail_in = (uInt)step;
            bool last = done_in + step == len;
            int flush = last ? flush_end : Z_NO_FLUSH;
            for (;;) {
                if (o.size() - used < 65536) o.resize(o.size() * 2);
                z.next_out = o.data() + used;
                z.avail_out = (uInt)std::min<size_t>(o.size() - used, 1u << 30);
                int rc = deflate(&z, flush);
                used = (size_t)(z.next_out - o.data());
                if (rc == Z_STREAM_ERROR) {
                    fail = true;
                    break;
                }
                if (flush == Z_FINISH ? rc == Z_STREAM_END : (z.avail_in == 0 && z.avail_out != 0)) break;
            }
            done_in += step;
            if (last || fail) break;
        }
        deflateEnd(&z);
        o.resize(used);
    });
    if (fail) return 0;
    uint64_t total = 0;
    for (auto& p : parts) total += p.size();
    if (total > out_cap) return 0;
    uint64_t at = 0;
    for (auto& p : parts) {
        memcpy(out + at, p.data(), p.size());
        at += p.size();
    }
    return total;
}

// zlib's CRC-32 (for host-side checks of generated data), chunked over threads and combined
GG_API uint32_t gg_crc32(uint32_t crc, const uint8_t* p, uint64_t n, int threads) {
    if (threads <= 0) threads = default_threads();
    const uint64_t piece = 64ull << 20;
    uint64_t np = (n + piece - 1) / piece;
    if (np <= 1) {
        for (uint64_t a = 0; a < n; a += 1u << 30) crc = crc32(crc, p + a, (uInt)std::min<uint64_t>(n - a, 1u << 30));
        return crc;
    }
    std::vector<uint32_t> c(np);
    parallel(np, threads, [&](uint64_t i) {
        uint64_t a = i * piece, len = std::min(piece, n - a);
        c[i] = crc32(0, p + a, (uInt)len);
    });
    for (uint64_t i = 0; i < np; i++) crc = crc32_combine(crc, c[i], (z_off_t)std::min(piece, n - i * piece));
    return crc;
}
