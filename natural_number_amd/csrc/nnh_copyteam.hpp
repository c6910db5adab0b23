// nnh_copyteam.hpp — a fixed team of host threads that splits one memcpy
// with the calling thread (the host side of the pinned staging ring in
// nnh_hostio.cpp).  Plain C++: no HIP, so it is tested on the CPU under
// ThreadSanitizer.
#pragma once
#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace nnh {

template <int NHELP>
class CopyTeam {
public:
    ~CopyTeam() { stop(); }
    void copy(void* dst, const void* src, uint64_t n) {
        if (n < (2ull << 20)) {
            std::memcpy(dst, src, n);
            return;
        }
        start();
        const uint64_t parts = NHELP + 1, per = ((n + parts - 1) / parts + 63) & ~63ull;  // parts * per >= n
        {
            std::lock_guard<std::mutex> g(mu_);
            for (int i = 0; i < NHELP; i++) {
                const uint64_t off = std::min<uint64_t>(n, per * (i + 1)), end = std::min<uint64_t>(n, off + per);
                job_[i] = Job{(char*)dst + off, (const char*)src + off, end - off};
            }
            left_ = NHELP;
            gen_++;
        }
        go_.notify_all();
        std::memcpy(dst, src, std::min<uint64_t>(n, per));
        std::unique_lock<std::mutex> l(mu_);
        done_.wait(l, [&] { return left_ == 0; });
    }
    void stop() {
        {
            std::lock_guard<std::mutex> g(mu_);
            if (th_.empty()) return;
            quit_ = true;
        }
        go_.notify_all();
        for (auto& t : th_) t.join();
        th_.clear();
        quit_ = false;
    }

private:
    struct Job {
        char* dst;
        const char* src;
        uint64_t n;
    };
    // (Threads started after a stop() must not take the last job of the
    // previous team for a new one: each starts from the current generation.)
    void start() {
        if (!th_.empty()) return;
        uint64_t g;
        {
            std::lock_guard<std::mutex> l(mu_);
            g = gen_;
        }
        for (int i = 0; i < NHELP; i++) th_.emplace_back([this, i, g] { run(i, g); });
    }
    void run(int i, uint64_t seen) {
        for (;;) {
            std::unique_lock<std::mutex> l(mu_);
            go_.wait(l, [&] { return quit_ || gen_ != seen; });
            if (quit_) return;
            seen = gen_;
            const Job j = job_[i];
            l.unlock();
            if (j.n) std::memcpy(j.dst, j.src, j.n);
            l.lock();
            if (--left_ == 0) done_.notify_one();
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable go_, done_;
    Job job_[NHELP]{};
    int left_ = 0;
    uint64_t gen_ = 0;
    bool quit_ = false;
};


}  // namespace nnh
