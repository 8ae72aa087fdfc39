// Fixed-size pool with a blocking chunked parallel-for (drop-in for engine/src/ThreadPool.cpp).
#include "u3d/ThreadPool.h"

#include <algorithm>

namespace u3d {

ThreadPool::ThreadPool(int threads)
{
    if (threads <= 0)
        threads = int(std::max(1u, std::thread::hardware_concurrency()));
    for (int i = 1; i < threads; ++i)
        m_threads.emplace_back([this, i] { worker(i); });
}

ThreadPool::~ThreadPool()
{
    {
        std::lock_guard<std::mutex> g(m_lock);
        m_quit = true;
    }
    m_start.notify_all();
    for (std::thread &t : m_threads)
        t.join();
}

void ThreadPool::worker(int index)
{
    unsigned seen = 0;
    for (;;) {
        const std::function<void(int, int, int)> *job;
        {
            std::unique_lock<std::mutex> g(m_lock);
            m_start.wait(g, [&] { return m_quit || m_epoch != seen; });
            if (m_quit)
                return;
            seen = m_epoch;
            job = m_job;
        }
        for (;;) {
            const int b = m_cursor.fetch_add(m_chunk);
            if (b >= m_total)
                break;
            (*job)(b, std::min(m_total, b + m_chunk), index);
        }
        std::lock_guard<std::mutex> g(m_lock);
        if (--m_busy == 0)
            m_finish.notify_one();
    }
}

void ThreadPool::parallelFor(int n, const std::function<void(int, int, int)> &fn, int grain)
{
    if (n <= 0)
        return;
    grain = std::max(1, grain);
    if (m_threads.empty() || n <= grain) {
        for (int b = 0; b < n; b += grain)
            fn(b, std::min(n, b + grain), 0);
        return;
    }
    {
        std::lock_guard<std::mutex> g(m_lock);
        m_job = &fn;
        m_total = n;
        m_chunk = grain;
        m_cursor.store(0);
        m_busy = int(m_threads.size());
        ++m_epoch;
    }
    m_start.notify_all();
    for (;;) {
        const int b = m_cursor.fetch_add(grain);
        if (b >= n)
            break;
        fn(b, std::min(n, b + grain), 0);
    }
    std::unique_lock<std::mutex> g(m_lock);
    m_finish.wait(g, [&] { return m_busy == 0; });
    m_job = nullptr;
}

} // namespace u3d
