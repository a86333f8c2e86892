// Host-side check of the copy pool behind the staged transfers: every byte
// copied, the parts handed over in order and only once complete, a failing
// hand-over leaves the pool usable.
#include <cstdio>
#include <random>
#include <stdexcept>
#include "../../immunodbscan_amd/csrc/copy_pool.h"

int main()
{
    std::mt19937 rng(1);
    for (int trial = 0; trial < 120; ++trial) {
        const size_t total = 1 + rng() % (12u << 20);
        std::vector<uint8_t> src(total), dst(total, 0);
        for (size_t k = 0; k < total; k += 97) src[k] = (uint8_t)(k * 7 + trial);
        const uint32_t parts = 1 + rng() % 4;
        const size_t step = (total + parts - 1) / parts;
        std::vector<CopyPool::Job> jobs;
        std::vector<uint32_t> part_of;
        uint32_t np = 0;
        for (size_t a = 0; a < total; a += step, ++np) {
            const size_t b = std::min(total, a + step), m = a + (b - a) / 3;  // two jobs per part
            if (m > a) {
                jobs.push_back({dst.data() + a, src.data() + a, m - a});
                part_of.push_back(np);
            }
            jobs.push_back({dst.data() + m, src.data() + m, b - m});
            part_of.push_back(np);
        }
        std::vector<uint32_t> order;
        bool early = false;
        CopyPool::get().run_parts(jobs.data(), part_of.data(), jobs.size(), np, [&](uint32_t q) {
            order.push_back(q);
            const size_t a = q * step, b = std::min(total, a + step);
            if (std::memcmp(dst.data() + a, src.data() + a, b - a) != 0) early = true;
        });
        if (early || order.size() != np) {
            printf("FAIL part handed over before it was complete (trial %d)\n", trial);
            return 1;
        }
        for (uint32_t q = 0; q < np; ++q)
            if (order[q] != q) {
                printf("FAIL parts out of order (trial %d)\n", trial);
                return 1;
            }
        if (std::memcmp(dst.data(), src.data(), total)) {
            printf("FAIL bytes (trial %d)\n", trial);
            return 1;
        }
        if (trial % 3 == 0) {
            std::vector<uint8_t> d2(total);
            CopyPool::Job j{d2.data(), src.data(), total};
            CopyPool::get().run(&j, 1);
            if (std::memcmp(d2.data(), src.data(), total)) {
                printf("FAIL plain run (trial %d)\n", trial);
                return 1;
            }
        }
        if (trial == 60) {
            bool thrown = false;
            try {
                CopyPool::get().run_parts(jobs.data(), part_of.data(), jobs.size(), np, [&](uint32_t q) {
                    if (q == 0) throw std::runtime_error("hand-over failed");
                });
            } catch (const std::runtime_error &) {
                thrown = true;
            }
            if (!thrown) {
                printf("FAIL exception lost\n");
                return 1;
            }
        }
    }
    // part lists the pool cannot track are refused before any copy
    {
        std::vector<uint8_t> a(64), b(64);
        CopyPool::Job j[2] = {{a.data(), b.data(), 32}, {a.data() + 32, b.data() + 32, 32}};
        const uint32_t bad_order[2] = {1, 0}, too_many[2] = {0, 9};
        int refused = 0;
        try {
            CopyPool::get().run_parts(j, bad_order, 2, 2, [](uint32_t) {});
        } catch (const std::invalid_argument &) {
            ++refused;
        }
        try {
            CopyPool::get().run_parts(j, too_many, 2, 10, [](uint32_t) {});
        } catch (const std::length_error &) {
            ++refused;
        }
        if (refused != 2) {
            printf("FAIL bad part lists accepted\n");
            return 1;
        }
    }
    printf("OK copy pool\n");
    return 0;
}
