// batch_queue.h - the queue between gi_batch's callers and its worker threads, host code only
// (the HIP library instantiates it over gi_job; tests/native/batch_queue_test.cpp runs it over a
// stand-in job under ThreadSanitizer).
//
// Jobs arrive two ways: a whole batch (run: queued largest first, the caller waits for all of
// them) or one at a time (submit: arrival order, handed back by wait_done as they finish).
// A worker takes, in queue order, every queued job that fits one batch segment of its context,
// or a single job too large to share one. Every field shared between threads is read and
// written under the one mutex.
#pragma once
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <mutex>
#include <vector>

namespace gz {

// What one worker's batch segment holds at most
struct SegmentRoom {
    uint64_t jobs = 0;       // streams
    uint64_t in_bytes = 0;   // packed compressed input
    uint64_t out_bytes = 0;  // summed output room
};

// Policy P gives, for a job: its packed input bytes, its output room, and whether it is too large
// to share a segment:
//   static uint64_t in_bytes(const Job*), out_bytes(const Job*);  static bool alone(const Job*, const SegmentRoom&)
template <class Job, class P>
class BatchQueue {
  public:
    using Clock = std::chrono::steady_clock;
    struct Item {
        Job* job;
        Clock::time_point queued;
        bool submitted;  // came through submit: handed back by wait_done
    };

    // the caller's whole batch; returns when every one of its jobs is done
    void run(Job* jobs, uint64_t n, uint64_t (*size_of)(const Job*)) {
        std::vector<Job*> order(n);
        for (uint64_t i = 0; i < n; i++) order[i] = jobs + i;
        std::stable_sort(order.begin(), order.end(), [&](Job* a, Job* b) { return size_of(a) > size_of(b); });
        const Clock::time_point now = Clock::now();
        std::unique_lock<std::mutex> lk(mu_);
        for (Job* j : order) queue_.push_back({j, now, false});
        running_ += n;
        work_.notify_all();
        done_.wait(lk, [&] { return running_ == 0; });
    }
    void submit(Job* j) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            queue_.push_back({j, Clock::now(), true});
        }
        work_.notify_one();
    }
    // the next submitted job that is done, or null once timeout_s seconds pass
    Job* wait_done(double timeout_s) {
        std::unique_lock<std::mutex> lk(mu_);
        const auto limit = std::chrono::duration<double>(timeout_s > 0 ? timeout_s : 0);
        if (!done_.wait_for(lk, limit, [&] { return !finished_.empty(); })) return nullptr;
        Job* j = finished_.front();
        finished_.pop_front();
        return j;
    }

    // A worker's next batch segment, in queue order: false when the queue is stopped and empty.
    bool take(const SegmentRoom& room, std::vector<Item>& out) {
        out.clear();
        bool more;
        {
            std::unique_lock<std::mutex> lk(mu_);
            work_.wait(lk, [&] { return stop_ || !queue_.empty(); });
            if (queue_.empty()) return false;  // stopping
            uint64_t in_sum = 0, out_sum = 0;
            while (!queue_.empty()) {
                const Job* j = queue_.front().job;
                if (P::alone(j, room)) {
                    if (out.empty()) {  // a large job is a segment of its own
                        out.push_back(queue_.front());
                        queue_.pop_front();
                    }
                    break;
                }
                const uint64_t in_next = in_sum + P::in_bytes(j), out_next = out_sum + P::out_bytes(j);
                if (!out.empty() && (out.size() >= room.jobs || in_next > room.in_bytes || out_next > room.out_bytes)) break;
                out.push_back(queue_.front());
                queue_.pop_front();
                in_sum = in_next;
                out_sum = out_next;
            }
            more = !queue_.empty();
        }
        if (more) work_.notify_one();  // what is left goes to another worker
        return true;
    }
    // a worker finished the items it took
    void finish(const std::vector<Item>& items) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (const Item& it : items) {
                if (it.submitted)
                    finished_.push_back(it.job);
                else
                    running_--;
            }
        }
        done_.notify_all();
    }
    // workers return from take once the queue is empty
    void stop() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        work_.notify_all();
    }

  private:
    std::mutex mu_;
    std::condition_variable work_, done_;
    std::deque<Item> queue_;
    std::deque<Job*> finished_;
    uint64_t running_ = 0;  // run()'s jobs not yet finished
    bool stop_ = false;
};

}  // namespace gz
