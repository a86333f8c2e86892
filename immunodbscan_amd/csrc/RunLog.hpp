// RunLog.hpp — the run log of the reference's host program, kept by the
// drop-in InterlinkClass.
//
// The reference collects messages and stage timers in a process-wide object
// and writes them as a two-column TSV (cpp/BenchmarkSingleton.hpp:33-115):
// header "Timer/Message<TAB>Value"; a message row is ("At <s> sec", text)
// with the seconds since logging began; a timer row is (name, "<s> sec").
// Numbers use std::to_string's six decimals.  RunLog keeps rows in that
// format; InterlinkClass appends one clustering run's rows in the order the
// reference's DBSCAN writes them (cuda/PSDBSCAN.hcu:986-997, 1113, 1198,
// 1279-1292, 1300-1354), with the library's stage times.
#pragma once
#include <chrono>
#include <cstdio>
#include <fstream>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

namespace idb {

class RunLog {
  public:
    static RunLog &instance()
    {
        static RunLog log;
        return log;
    }

    void message(const std::string &text)
    {
        std::lock_guard<std::mutex> lock(mu_);
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
        rows_.emplace_back("At " + std::to_string(s) + " sec", text);
        if (verbose_) std::printf("\n%s", text.c_str());
    }

    void timer(const std::string &name, double seconds)
    {
        std::lock_guard<std::mutex> lock(mu_);
        rows_.emplace_back(name, std::to_string(seconds) + " sec");
        if (verbose_) std::printf("\nTimer %s: %f sec", name.c_str(), seconds);
    }

    // Named timers: start (or restart) one, stop it to add its row; a timer
    // never started adds nothing.
    void start(const std::string &name)
    {
        std::lock_guard<std::mutex> lock(mu_);
        starts_[name] = std::chrono::steady_clock::now();
    }

    void stop(const std::string &name)
    {
        std::chrono::steady_clock::time_point t0;
        {
            std::lock_guard<std::mutex> lock(mu_);
            const auto it = starts_.find(name);
            if (it == starts_.end()) return;
            t0 = it->second;
        }
        timer(name, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }

    // Messages and timers are also printed as they arrive.
    void be_verbose() { verbose_ = true; }

    std::vector<std::pair<std::string, std::string>> rows() const
    {
        std::lock_guard<std::mutex> lock(mu_);
        return rows_;
    }

    // The TSV file, as the reference's write_map produces it.
    bool write(const std::string &path) const
    {
        std::ofstream out(path);
        if (!out.is_open()) return false;
        out << "Timer/Message\tValue\n";
        for (const auto &r : rows()) out << r.first << "\t" << r.second << "\n";
        return static_cast<bool>(out);
    }

  private:
    RunLog() : t0_(std::chrono::steady_clock::now()) {}
    mutable std::mutex mu_;
    std::chrono::steady_clock::time_point t0_;
    std::vector<std::pair<std::string, std::string>> rows_;
    std::map<std::string, std::chrono::steady_clock::time_point> starts_;
    bool verbose_ = false;
};

}  // namespace idb
