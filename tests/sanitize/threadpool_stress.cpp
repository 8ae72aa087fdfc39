// The drop-in ThreadPool (host/src/ThreadPool.cpp) under ThreadSanitizer: back-to-back
// parallel-fors of every shape (fewer items than threads, one chunk, uneven chunks, many small
// ones), several pools at once, pools destroyed right after their last call, and a job that
// itself reads what the previous job wrote. Exits 0 when every element was visited exactly once
// per call and TSan reported nothing (it exits non-zero on a report with halt_on_error=1).
#include <atomic>
#include <cstdio>
#include <thread>
#include <vector>

#include "u3d/ThreadPool.h"

namespace {

bool run(u3d::ThreadPool &pool, int n, int grain, std::vector<int> &hits, int round)
{
    hits.assign(size_t(n > 0 ? n : 1), 0);
    std::atomic<long> sum{0};
    pool.parallelFor(
        n,
        [&](int b, int e, int t) {
            long local = 0;
            for (int i = b; i < e; ++i) {
                hits[size_t(i)] += 1; // disjoint chunks: no two threads write one element
                local += i;
            }
            if (t < 0 || t >= pool.size())
                std::abort();
            sum.fetch_add(local, std::memory_order_relaxed);
        },
        grain);
    for (int i = 0; i < n; ++i)
        if (hits[size_t(i)] != 1) {
            std::fprintf(stderr, "round %d n %d grain %d: element %d visited %d times\n", round, n,
                         grain, i, hits[size_t(i)]);
            return false;
        }
    return sum.load() == long(n) * long(n - 1) / 2 || n <= 0;
}

} // namespace

int main()
{
    const int shapes[][2] = {{0, 1}, {1, 1}, {3, 1}, {7, 2}, {64, 64}, {1000, 1}, {1000, 7}, {4097, 64}, {50000, 333}};
    int round = 0;
    for (int threads : {1, 2, 3, 8}) {
        u3d::ThreadPool pool(threads);
        std::vector<int> hits;
        for (int rep = 0; rep < 40; ++rep)
            for (const auto &s : shapes)
                if (!run(pool, s[0], s[1], hits, round++))
                    return 1;
    }
    // several pools driven from several threads at once, each destroyed after its last call
    std::vector<std::thread> drivers;
    std::atomic<int> failures{0};
    for (int d = 0; d < 4; ++d)
        drivers.emplace_back([&, d] {
            for (int k = 0; k < 10; ++k) {
                u3d::ThreadPool pool(2 + d);
                std::vector<int> hits;
                for (const auto &s : shapes)
                    if (!run(pool, s[0], s[1], hits, k))
                        failures.fetch_add(1);
            }
        });
    for (std::thread &t : drivers)
        t.join();
    // a job reading the previous job's output: the call returns only after every chunk is done
    u3d::ThreadPool pool(4);
    std::vector<float> a(10000, 1.0f), b(10000, 0.0f);
    for (int k = 0; k < 50; ++k) {
        pool.parallelFor(int(a.size()), [&](int s, int e, int) { for (int i = s; i < e; ++i) b[size_t(i)] = a[size_t(i)] + 1.0f; }, 97);
        pool.parallelFor(int(a.size()), [&](int s, int e, int) { for (int i = s; i < e; ++i) a[size_t(i)] = b[size_t(i)]; }, 101);
    }
    for (float v : a)
        if (v != 51.0f)
            return 1;
    std::printf("threadpool_stress: %d rounds, %d failures\n", round, failures.load());
    return failures.load() == 0 ? 0 : 1;
}
