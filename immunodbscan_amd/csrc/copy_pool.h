// Process-wide pool of host threads for the copies between caller arrays
// and the pinned staging buffer (host code only, no HIP).
#pragma once
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <vector>

// Host copies between caller arrays and pinned staging, split over a few
// threads: one core's memcpy moves about 60 GB/s, and a pinned transfer runs
// at about 45 GB/s up and down against 26 and 19 GB/s for the runtime's own
// staging of pageable copies (MI355X host, 8 MB).  One process-wide pool;
// calls into it are serialised.
class CopyPool {
public:
    struct Job {
        void *dst;
        const void *src;
        size_t bytes;
    };
    static CopyPool &get()
    {
        static CopyPool *p = new CopyPool();  // never destroyed: its threads live until the process ends
        return *p;
    }
    void run(const Job *jobs, size_t n)
    {
        std::lock_guard<std::mutex> call(call_m_);
        split(jobs, nullptr, n, 1);
        start();
        work();
        finish();
    }
    // Copies like run(); job k belongs to part part_of[k] (non-decreasing,
    // below kMaxParts).  The calling thread calls issue(p) for p = 0, 1, ...
    // in order, each as soon as every byte of part p is copied, while the
    // workers go on with the later parts: one wake-up of the pool per call.
    static constexpr uint32_t kMaxParts = 8;
    template <class F>
    void run_parts(const Job *jobs, const uint32_t *part_of, size_t n, uint32_t parts, F &&issue)
    {
        std::lock_guard<std::mutex> call(call_m_);
        if (parts > kMaxParts) throw std::length_error("copy pool: more parts than it tracks");
        for (size_t k = 0; k < n; ++k)
            if (part_of[k] >= parts || (k > 0 && part_of[k] < part_of[k - 1]))
                throw std::invalid_argument("copy pool: job parts must be non-decreasing and below the part count");
        split(jobs, part_of, n, parts);
        start();
        uint32_t issued = 0;
        auto flush = [&] {
            while (issued < parts && left_[issued].load(std::memory_order_acquire) == 0) issue(issued++);
        };
        try {
            for (size_t i; (i = next_.fetch_add(1)) < pieces_.size();) {
                copy(i);
                flush();
            }
            while (issued < parts) {
                flush();
                if (issued < parts) std::this_thread::yield();
            }
        } catch (...) {
            // a failed issue: the workers stop claiming pieces and are
            // waited for before the pieces they read go away
            next_.store(pieces_.size());
            finish();
            throw;
        }
        finish();
    }

private:
    static constexpr size_t kPiece = 256 << 10;
    CopyPool()
    {
        const unsigned hw = std::thread::hardware_concurrency();
        workers_ = std::min(3u, hw > 1 ? hw - 1 : 0u);
        for (unsigned k = 0; k < workers_; ++k) std::thread([this] { loop(); }).detach();
    }
    void loop()
    {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return gen_ != seen; });
                seen = gen_;
            }
            work();
            std::lock_guard<std::mutex> g(m_);
            if (--active_ == 0) done_.notify_one();
        }
    }
    // pieces of at most kPiece bytes, in job order, with their parts
    void split(const Job *jobs, const uint32_t *part_of, size_t n, uint32_t parts)
    {
        pieces_.clear();
        piece_part_.clear();
        total_ = 0;
        size_t count[kMaxParts] = {};
        for (size_t k = 0; k < n; ++k) {
            const uint32_t part = part_of ? part_of[k] : 0u;
            for (size_t o = 0; o < jobs[k].bytes; o += kPiece) {
                pieces_.push_back({(uint8_t *)jobs[k].dst + o, (const uint8_t *)jobs[k].src + o,
                                   std::min(kPiece, jobs[k].bytes - o)});
                piece_part_.push_back(part);
                ++count[part];
            }
            total_ += jobs[k].bytes;
        }
        for (uint32_t q = 0; q < kMaxParts; ++q) left_[q].store(q < parts ? count[q] : 0, std::memory_order_relaxed);
        next_.store(0);
    }
    // wakes the workers when the copy is worth sharing
    void start()
    {
        woke_ = workers_ > 0 && total_ >= 2 * kPiece;
        if (!woke_) return;
        {
            std::lock_guard<std::mutex> g(m_);
            ++gen_;
            active_ = workers_;
        }
        cv_.notify_all();
    }
    void finish()
    {
        if (!woke_) return;
        std::unique_lock<std::mutex> g(m_);
        done_.wait(g, [&] { return active_ == 0; });
    }
    void copy(size_t i)
    {
        std::memcpy(pieces_[i].dst, pieces_[i].src, pieces_[i].bytes);
        left_[piece_part_[i]].fetch_sub(1, std::memory_order_release);
    }
    void work()
    {
        for (size_t i; (i = next_.fetch_add(1)) < pieces_.size();) copy(i);
    }
    std::mutex call_m_, m_;
    std::condition_variable cv_, done_;
    std::vector<Job> pieces_;
    std::vector<uint32_t> piece_part_;
    std::atomic<size_t> left_[kMaxParts];  // pieces of each part not yet copied
    std::atomic<size_t> next_{0};
    size_t total_ = 0;
    bool woke_ = false;
    uint64_t gen_ = 0;
    unsigned active_ = 0, workers_ = 0;
};
