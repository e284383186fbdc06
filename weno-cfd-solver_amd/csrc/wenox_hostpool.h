// wenox_hostpool.h -- persistent host worker threads for the drop-in calls'
// host-side packing and copies (included by wenox_abi.cpp; header-only so a
// host test can build it alone).
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace wenox {

// Host threads for the drop-in calls' packing and copies (memory-bound; one
// thread moves a few GB/s): workers started on first use and kept for the
// life of the process (starting threads per call cost about as much as the
// copies).  A job runs fn over T row ranges; start() hands it to the workers
// and returns, so the caller can issue device work meanwhile; wait() runs the
// ranges no worker has taken on the calling thread (so a job always
// completes, also in a forked child that has no workers) and returns when all
// are done.  One job at a time.  The library keeps one pool for the life of
// the process and never destroys it: its idle workers end with the process.
class HostPool {
public:
    static constexpr int MAX_THREADS = 16;
    int threads()
    {
        const unsigned hw = std::thread::hardware_concurrency();
        return (int)std::min<unsigned>(hw ? hw : 1, MAX_THREADS);
    }
    void start(int64_t M, int T, std::function<void(int64_t, int64_t)> fn)
    {
        std::unique_lock<std::mutex> lk(mu_);
        idle_.wait(lk, [this] { return !active_; });
        while ((int)th_.size() < std::min(T, MAX_THREADS - 1)) th_.emplace_back([this] { work(); });
        active_ = true;
        fn_ = std::move(fn);
        m_ = M;
        parts_ = T;
        next_ = 0;
        left_ = T;
        ++gen_;
        cv_.notify_all();
    }
    void wait()
    {
        take();
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return left_ == 0; });
        fn_ = nullptr;
        active_ = false;
        idle_.notify_all();
    }
    void run(int64_t M, int T, std::function<void(int64_t, int64_t)> fn)
    {
        start(M, T, std::move(fn));
        wait();
    }

private:
    // run the job's ranges no thread has taken yet
    void take()
    {
        std::unique_lock<std::mutex> lk(mu_);
        while (active_ && next_ < parts_) {
            const int t = next_++;
            const int64_t M = m_, T = parts_;
            lk.unlock();
            fn_(t * M / T, (t + 1) * M / T);   // fn_ stays until every range is done
            lk.lock();
            if (--left_ == 0) done_.notify_all();
        }
    }
    void work()
    {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
            }
            take();
        }
    }
    std::mutex mu_;
    std::condition_variable cv_, done_, idle_;
    std::vector<std::thread> th_;
    std::function<void(int64_t, int64_t)> fn_;
    bool active_ = false;
    int64_t m_ = 0;
    int parts_ = 0, next_ = 0, left_ = 0;
    uint64_t gen_ = 0;
};

}  // namespace wenox
