// ORACLE -- test infrastructure only. Nothing in the product may include, link or run this.
//
// A host thread pool for the CPU baseline's parallel mode: the stages the reference hands to its
// job system (world.h:1462-1481 for_range_: tight bounds 1489-1499, pair discovery 1545-1577 with
// per-worker pair buckets, narrowphase 1598-1645) run as contiguous chunks on `workers` threads,
// the calling thread included; everything else stays serial, as in the reference. Chunks write
// only their own indices or their own bucket, so the results equal the serial run's.
#pragma once
#include <condition_variable>
#include <cstddef>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace pxo {

class Pool {
public:
    explicit Pool(unsigned workers) : n_(workers < 1 ? 1 : workers) {
        for (unsigned w = 1; w < n_; ++w) threads_.emplace_back([this, w] { loop(w); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> lock(mu_);
            stop_ = true;
            ++gen_;
        }
        cv_.notify_all();
        for (auto& t : threads_) t.join();
    }
    Pool(const Pool&) = delete;
    Pool& operator=(const Pool&) = delete;
    unsigned workers() const { return n_; }

    // fn(begin, end, worker) over [0, count) in n_ contiguous chunks; returns when all are done
    void run(size_t count, const std::function<void(size_t, size_t, unsigned)>& fn) {
        if (n_ == 1 || count < 2 * n_) {
            fn(0, count, 0);
            return;
        }
        {
            std::lock_guard<std::mutex> lock(mu_);
            fn_ = &fn;
            count_ = count;
            pending_ = n_ - 1;
            ++gen_;
        }
        cv_.notify_all();
        fn(0, count / n_, 0);
        std::unique_lock<std::mutex> lock(mu_);
        done_.wait(lock, [this] { return pending_ == 0; });
        fn_ = nullptr;
    }

private:
    void loop(unsigned w) {
        unsigned long seen = 0;
        for (;;) {
            const std::function<void(size_t, size_t, unsigned)>* fn;
            size_t count;
            {
                std::unique_lock<std::mutex> lock(mu_);
                cv_.wait(lock, [&] { return gen_ != seen; });
                seen = gen_;
                if (stop_) return;
                fn = fn_;
                count = count_;
            }
            (*fn)(count * w / n_, count * (w + 1) / n_, w);
            {
                std::lock_guard<std::mutex> lock(mu_);
                if (--pending_ == 0) done_.notify_one();
            }
        }
    }
    unsigned n_;
    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    const std::function<void(size_t, size_t, unsigned)>* fn_ = nullptr;
    size_t count_ = 0;
    unsigned pending_ = 0;
    unsigned long gen_ = 0;
    bool stop_ = false;
};

}  // namespace pxo
