// rk_gang.hpp -- a fixed set of host threads for large host copies.
// Host-only (no HIP): the sanitizer builds of the CPU code include it.
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace rk {

// A fixed set of host threads that run one job at a time: run(f) calls
// f(0) .. f(size-1), f(0) on the caller's thread, and returns when every
// part is done.  Kept by a context for its large host copies, so a call
// starts no threads.
class Gang {
public:
    explicit Gang(unsigned size) : size_(std::max(1u, size))
    {
        for (unsigned k = 1; k < size_; ++k) th_.emplace_back([this, k] { loop(k); });
    }
    Gang(const Gang&) = delete;
    Gang& operator=(const Gang&) = delete;
    ~Gang()
    {
        {
            std::lock_guard<std::mutex> l(m_);
            stop_ = true;
        }
        go_.notify_all();
        for (std::thread& t : th_) t.join();
    }
    unsigned size() const { return size_; }
    void run(const std::function<void(unsigned)>& f)
    {
        {
            std::lock_guard<std::mutex> l(m_);
            job_ = &f;
            pending_ = size_ - 1;
            ++gen_;
        }
        go_.notify_all();
        f(0);
        std::unique_lock<std::mutex> l(m_);
        done_.wait(l, [&] { return pending_ == 0; });
        job_ = nullptr;
    }

private:
    void loop(unsigned k)
    {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> l(m_);
        for (;;) {
            go_.wait(l, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            const std::function<void(unsigned)>* f = job_;
            l.unlock();
            (*f)(k);
            l.lock();
            if (--pending_ == 0) done_.notify_one();
        }
    }
    unsigned size_;
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable go_, done_;
    const std::function<void(unsigned)>* job_ = nullptr;
    unsigned pending_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

} // namespace rk
