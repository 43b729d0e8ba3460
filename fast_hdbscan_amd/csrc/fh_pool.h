// Host worker threads for the label stage: created once per process and
// kept, so a call's parallel passes cost a wake-up each instead of a thread
// creation per task.  One job at a time; a caller that finds the pool busy
// (another thread's label stage) runs its tasks on threads of its own.
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace fh {

class HostPool {
  public:
    // threads a job may use: the host's cores, at most 16
    static int width() {
        static const int w = std::max(1, std::min(16, static_cast<int>(std::thread::hardware_concurrency())));
        return w;
    }

    // fn(k) for k in [0, tasks), on the pool's workers and the calling
    // thread; returns when every task has finished
    static void run(int tasks, const std::function<void(int)> &fn) {
        if (tasks <= 0) return;
        if (tasks == 1) {
            fn(0);
            return;
        }
        HostPool &p = instance();
        std::unique_lock<std::mutex> busy(p.job_mu_, std::try_to_lock);
        if (!busy.owns_lock()) {
            std::vector<std::thread> th;
            for (int k = 1; k < tasks; ++k) th.emplace_back(fn, k);
            fn(0);
            for (auto &t : th) t.join();
            return;
        }
        {
            std::lock_guard<std::mutex> g(p.mu_);
            p.fn_ = &fn;
            p.tasks_ = tasks;
            p.next_ = 0;
            p.pending_ = tasks;
            ++p.gen_;
        }
        p.cv_.notify_all();
        p.work();
        std::unique_lock<std::mutex> g(p.mu_);
        p.done_cv_.wait(g, [&] { return p.pending_ == 0; });
        p.fn_ = nullptr;
    }

  private:
    static HostPool &instance() {
        // never destroyed: the workers wait on it until the process ends
        static HostPool *p = new HostPool(width() - 1);
        return *p;
    }

    explicit HostPool(int workers) {
        for (int i = 0; i < workers; ++i) std::thread([this] { loop(); }).detach();
    }

    // claims tasks of the current job until none are left; a task is only
    // claimed under the lock for the job whose function this thread holds,
    // and a job ends only when its claimed tasks have finished
    void work() {
        const std::function<void(int)> *fn;
        uint64_t gen;
        {
            std::lock_guard<std::mutex> g(mu_);
            fn = fn_;
            gen = gen_;
        }
        if (!fn) return;
        for (;;) {
            int k;
            {
                std::lock_guard<std::mutex> g(mu_);
                if (gen_ != gen || next_ >= tasks_) return;
                k = next_++;
            }
            (*fn)(k);
            std::lock_guard<std::mutex> g(mu_);
            if (--pending_ == 0) done_cv_.notify_all();
        }
    }

    void loop() {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> g(mu_);
                cv_.wait(g, [&] { return gen_ != seen; });
                seen = gen_;
            }
            work();
        }
    }

    std::mutex job_mu_;  // one job at a time
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(int)> *fn_ = nullptr;
    int tasks_ = 0;
    int next_ = 0;
    int pending_ = 0;
    uint64_t gen_ = 0;
};

}  // namespace fh
