// Drop-in declaration of the reference's worker pool (engine/include/u3d/ThreadPool.h). The
// GPU trainer does not use it for compute; it stays in the signatures so callers that construct
// one and pass it to the trainer compile unchanged, and remains a working parallel-for.
#pragma once

#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace u3d {

class ThreadPool
{
public:
    explicit ThreadPool(int threads = 0); // <= 0: all hardware threads
    ~ThreadPool();
    ThreadPool(const ThreadPool &) = delete;
    ThreadPool &operator=(const ThreadPool &) = delete;

    int size() const { return int(m_threads.size()) + 1; }

    // fn(begin, end, threadIndex) over chunks of [0, n) of at most `grain` items; the caller
    // takes part as thread 0 and the call returns when every chunk is done.
    void parallelFor(int n, const std::function<void(int, int, int)> &fn, int grain = 1);

private:
    void worker(int index);

    std::vector<std::thread> m_threads;
    std::mutex m_lock;
    std::condition_variable m_start, m_finish;
    const std::function<void(int, int, int)> *m_job = nullptr;
    int m_total = 0, m_chunk = 1;
    std::atomic<int> m_cursor{0};
    int m_busy = 0;
    unsigned m_epoch = 0;
    bool m_quit = false;
};

} // namespace u3d
