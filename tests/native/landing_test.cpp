// landing_test.cpp - the decoding core's landing rule on the CPU (host build of deflate_core.h).
//
// A decoder checks its landing list at symbol boundaries only (inflate.cuh:422-428): at a
// boundary p it lands on a target at p whose mode (and, in a dynamic block, header) is its own;
// a target at p of another kind is passed over and decoding goes on. A decoder out of step with
// the stream (a false probe synchronisation) reaches such targets at its boundaries, and so does
// one whose list names a start that is not a block boundary. Each case below puts a target of
// the wrong kind exactly on a symbol boundary of the decode; the decode must then run to the
// stream's end as if the target were not there. Each case runs in a child process with an alarm,
// so a decode that never returns fails the case instead of hanging the test.
//
// Built and run by tests/test_native.py; prints "N failures".
#include <signal.h>
#include <stdio.h>
#include <string.h>
#include <sys/wait.h>
#include <unistd.h>
#include <zlib.h>

#include <string>
#include <vector>

#include "deflate_core.h"

using namespace gz;

namespace {

std::vector<uint8_t> text(size_t n) {
    static const char* words[] = {"alpha", "bravo", "charlie", "delta", "echo", "foxtrot", "golf", "hotel",
                                  "india", "juliet", "kilo", "lima", "mike", "november", "oscar", "papa"};
    std::vector<uint8_t> out;
    uint64_t x = 0x9E3779B97F4A7C15ull;
    while (out.size() < n) {
        x ^= x << 13, x ^= x >> 7, x ^= x << 17;
        const char* w = words[x % 16];
        out.insert(out.end(), w, w + strlen(w));
        out.push_back((x >> 8) % 11 == 0 ? '\n' : ' ');
        if ((x >> 16) % 7 == 0) {
            char num[16];
            int k = snprintf(num, sizeof num, "%u ", (unsigned)(x >> 40) % 100000);
            out.insert(out.end(), num, num + k);
        }
    }
    out.resize(n);
    return out;
}

std::vector<uint8_t> deflate_raw(const std::vector<uint8_t>& d, int level, int strategy) {
    z_stream z;
    memset(&z, 0, sizeof z);
    deflateInit2(&z, level, Z_DEFLATED, -15, 8, strategy);
    std::vector<uint8_t> out(deflateBound(&z, d.size()) + 64);
    z.next_in = (Bytef*)d.data();
    z.avail_in = (uInt)d.size();
    z.next_out = out.data();
    z.avail_out = (uInt)out.size();
    deflate(&z, Z_FINISH);
    out.resize(z.total_out);
    deflateEnd(&z);
    return out;
}

void build(const Bits& proto, uint64_t key, Table& t) {
    memset(&t, 0, sizeof t);
    uint8_t lens[320];
    uint16_t codes[320];
    int hl = 288, hd = 30;
    if (key == UINT64_MAX) {
        table_prepare_fixed(t, lens, codes);
        t.status = TAB_FIXED;
    } else {
        t.status = table_prepare(proto, key, t, lens, codes, hl, hd);
    }
    for (int s = 0; s < hl; s++) fill_lit_symbol(t, s, lens[s], codes[s]);
    for (int s = 0; s < hd; s++) fill_dist_symbol(t, s, lens[hl + s], codes[hl + s]);
}

struct Stream {
    std::vector<uint32_t> words;  // the stream, zero padded
    uint64_t size = 0;
    std::vector<Table> tabs;      // tabs[0] the fixed code, tabs[1] the first block's (if dynamic)
    uint64_t keys[1] = {0};
    uint32_t slots[1] = {1};
    int nkeys = 0;
    uint32_t stg[SW];
    std::vector<uint16_t> out;

    explicit Stream(const std::vector<uint8_t>& s) : size(s.size()), tabs(2) {
        words.assign(s.size() / 4 + 16, 0);
        memcpy(words.data(), s.data(), s.size());
        Bits b;
        b.setup(words.data(), size);
        build(b, UINT64_MAX, tabs[0]);
        if (((s[0] >> 1) & 3) == 2) {  // the first block is dynamic: its table is the cache's one key
            build(b, 0, tabs[1]);
            nkeys = 1;
        }
    }
    Result run_from_start(const std::vector<Start>& targets) {
        out.assign(1 << 22, 0);
        Job j;
        memset(&j, 0, sizeof j);
        j.words = words.data();
        j.size = size;
        j.s = Start{0, 0, AT_BLOCK, -1, {0}};
        j.tabs = tabs.data();
        j.keys = keys;
        j.slots = slots;
        j.nkeys = nkeys;
        j.own = nkeys ? &tabs[1] : nullptr;
        j.targets = target_list(targets.data(), (int)targets.size());
        j.ntargets = (int)targets.size();
        j.out = out.data();
        j.out_index = 0;
        j.cap = (uint32_t)out.size();
        j.stream_start = true;
        j.stop = UINT64_MAX;
        j.fixed = tabs.data();
        j.hard_end = UINT64_MAX;
        j.stg = stg;
        j.stride = 1;
        return run(j, nullptr);
    }
};

Start target(uint64_t bit, uint8_t mode, uint64_t header) { return Start{bit, header, mode, -1, {0}}; }

// the bits in [lo, hi) where a decode from the start lands on a target of its own kind there:
// its symbol boundaries in the first block
std::vector<uint64_t> boundaries(Stream& s, uint64_t lo, uint64_t hi, uint8_t mode, uint64_t header) {
    std::vector<uint64_t> b;
    for (uint64_t q = lo; q < hi; q++) {
        Result r = s.run_from_start({target(q, mode, header)});
        if (r.landed == 0) b.push_back(q);
    }
    return b;
}

int failures = 0;

// runs fn in a child with an alarm; false when it did not return within `seconds` or failed
template <class F>
bool in_child(const char* what, unsigned seconds, F&& fn) {
    fflush(stdout);
    pid_t pid = fork();
    if (pid == 0) {
        alarm(seconds);
        _exit(fn() ? 0 : 1);
    }
    int st = 0;
    waitpid(pid, &st, 0);
    const bool ok = WIFEXITED(st) && WEXITSTATUS(st) == 0;
    if (!ok) {
        failures++;
        printf("FAIL %s: %s\n", what, WIFSIGNALED(st) && WTERMSIG(st) == SIGALRM ? "the decode never returned" : "wrong result");
    }
    return ok;
}

void check_case(const char* name, const std::vector<uint8_t>& data, int level, int strategy, uint8_t mode) {
    Stream s(deflate_raw(data, level, strategy));
    const Result whole = s.run_from_start({});
    if (whole.status != OK || whole.out_len != data.size()) {
        failures++;
        printf("FAIL %s: the plain decode (status %d, %u symbols)\n", name, whole.status, whole.out_len);
        return;
    }
    std::vector<uint16_t> want(s.out.begin(), s.out.begin() + whole.out_len);
    const uint64_t header = mode == IN_FIXED ? UINT64_MAX : 0;
    const std::vector<uint64_t> b = boundaries(s, 200, 1400, mode, header);
    if (b.size() < 20) {
        failures++;
        printf("FAIL %s: only %zu boundaries found\n", name, b.size());
        return;
    }
    int cases = 0;
    for (size_t i = 0; i < b.size() && cases < 12; i += b.size() / 12 + 1, cases++) {
        const uint64_t q = b[i];
        // (1) a block-start target on the boundary, the first in the list (in a fixed block every
        //     target is checked; in a dynamic one a list that starts with another header's target
        //     is not checked at all)
        // (2) a target of the decoder's own kind one bit before the boundary (inside the symbol that
        //     ends there: passed over), then a block-start target on the boundary
        // (3) as (2) with a target of another dynamic header on the boundary
        std::vector<std::vector<Start>> lists = {
            {target(q, AT_BLOCK, q)},
            {target(q - 1, mode, header), target(q, AT_BLOCK, q)},
            {target(q - 1, mode, header), target(q, IN_DYNAMIC, q + 12345)},
        };
        for (size_t k = 0; k < lists.size(); k++) {
            char what[160];
            snprintf(what, sizeof what, "%s, boundary %llu, list %zu", name, (unsigned long long)q, k + 1);
            in_child(what, 2, [&] {
                const Result r = s.run_from_start(lists[k]);
                return r.status == OK && r.landed == -1 && r.out_len == want.size() &&
                       memcmp(s.out.data(), want.data(), want.size() * 2) == 0;
            });
        }
    }
    printf("%s: %zu boundaries, %d checked\n", name, b.size(), cases);
}

}  // namespace

int main() {
    const std::vector<uint8_t> d = text(40000);
    check_case("fixed Huffman", d, 6, Z_FIXED, IN_FIXED);
    check_case("dynamic Huffman", d, 6, Z_DEFAULT_STRATEGY, IN_DYNAMIC);
    printf("%d failures\n", failures);
    return failures ? 1 : 0;
}
