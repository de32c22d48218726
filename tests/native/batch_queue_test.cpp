// The gi_batch job queue (gpuzip_amd/csrc/batch_queue.h) driven by host threads only: workers
// taking batch segments and finishing them, callers running whole batches, a serving loop
// submitting and waiting, all at once. Built with -fsanitize=thread by tests/test_native.py; a
// data race or a job taken twice, lost, or handed back wrongly fails it.
#include <atomic>
#include <cstdio>
#include <random>
#include <thread>
#include <vector>

#include "../../gpuzip_amd/csrc/batch_queue.h"

struct FakeJob {
    uint64_t in_size = 0, out_cap = 0;
    std::atomic<int> taken{0};
    int worker = -1;
};

struct Sizes {
    static uint64_t in_bytes(const FakeJob* j) { return j->in_size + 16; }
    static uint64_t out_bytes(const FakeJob* j) { return j->out_cap; }
    static bool alone(const FakeJob* j, const gz::SegmentRoom& r) { return j->in_size > r.in_bytes / 4; }
};

int main() {
    using Q = gz::BatchQueue<FakeJob, Sizes>;
    Q q;
    const gz::SegmentRoom room{64, 1 << 20, 4 << 20};
    std::atomic<int> failures{0};
    auto fail = [&](const char* what) {
        if (failures++ < 10) fprintf(stderr, "FAIL: %s\n", what);
    };
    std::vector<std::thread> workers;
    for (int w = 0; w < 3; w++)
        workers.emplace_back([&, w] {
            std::mt19937 rng(w);
            std::vector<Q::Item> items;
            while (q.take(room, items)) {
                if (items.empty()) fail("empty batch");
                uint64_t in = 0;
                for (auto& it : items) {
                    if (it.job->taken.fetch_add(1) != 0) fail("a job taken twice");
                    it.job->worker = w;
                    in += Sizes::in_bytes(it.job);
                }
                if (items.size() > 1 && in > room.in_bytes) fail("a batch over its room");
                std::this_thread::sleep_for(std::chrono::microseconds(rng() % 200));
                q.finish(items);
            }
        });
    // two callers with whole batches, one serving loop
    auto make = [](std::vector<FakeJob>& v, unsigned seed) {
        std::mt19937 rng(seed);
        for (auto& j : v) j.in_size = 1 + rng() % (rng() % 8 == 0 ? 600000 : 20000), j.out_cap = 4 * j.in_size;
    };
    std::vector<FakeJob> a(500), b(700), c(400);
    make(a, 1), make(b, 2), make(c, 3);
    std::thread ra([&] { q.run(a.data(), a.size(), [](const FakeJob* j) { return j->in_size; }); });
    std::thread rb([&] { q.run(b.data(), b.size(), [](const FakeJob* j) { return j->in_size; }); });
    std::thread serve([&] {
        size_t back = 0;
        for (auto& j : c) q.submit(&j);
        while (back < c.size()) {
            FakeJob* j = q.wait_done(10.0);
            if (!j) {
                fail("a submitted job never came back");
                break;
            }
            if (j < c.data() || j >= c.data() + c.size()) fail("a job handed back that was not submitted");
            back++;
        }
    });
    ra.join();
    rb.join();
    serve.join();
    for (auto* v : {&a, &b, &c})
        for (auto& j : *v)
            if (j.taken.load() != 1) fail("a job not taken exactly once");
    if (q.wait_done(0.01) != nullptr) fail("more handed back than submitted");
    q.stop();
    for (auto& t : workers) t.join();
    printf("batch queue: %d failures\n", failures.load());
    return failures.load() ? 1 : 0;
}
