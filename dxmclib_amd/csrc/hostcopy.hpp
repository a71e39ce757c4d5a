// Host-side copies of the staged transfers, free of HIP so that they build
// and are tested on any host.
//
// CopyHelpers: three persistent helper threads that copy parts of one block
// while the calling thread copies the first part. One core copies 27-30 GB/s
// out of page-locked memory on the MI355X host, four 80-100 GB/s, while the
// device-to-host DMA of a staging chunk runs at 53-57 GB/s. A context starts
// its helpers with its staging chunks and stops them with them; its lock
// keeps one copy at a time. When the threads cannot start, every copy runs
// on the calling thread.
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <thread>

namespace dxmc_backend {

class CopyHelpers {
public:
    static constexpr int PARTS = 4;
    // blocks below this are copied by the calling thread alone
    static constexpr std::size_t PART_MIN = std::size_t { 1 } << 22;

    CopyHelpers() = default;
    CopyHelpers(const CopyHelpers&) = delete;
    CopyHelpers& operator=(const CopyHelpers&) = delete;
    ~CopyHelpers() { stop(); }

    // starts the helper threads; false (and no helpers) if one cannot start
    bool start()
    {
        if (m_running)
            return true;
        m_quit = false;
        for (int k = 1; k < PARTS; ++k) {
            try {
                m_threads[k - 1] = std::thread([this, k] { work(k); });
            } catch (...) {
                stop();
                return false;
            }
        }
        m_running = true;
        return true;
    }

    void stop()
    {
        {
            std::lock_guard lock(m_mutex);
            m_quit = true;
        }
        m_wake.notify_all();
        for (auto& t : m_threads)
            if (t.joinable())
                t.join();
        m_running = false;
    }

    bool running() const { return m_running; }

    // dst[0, len) = src[0, len); with the helpers running and len >=
    // PART_MIN, in PARTS parts on PARTS threads
    void copy(void* dst, const void* src, std::size_t len)
    {
        if (!m_running || len < PART_MIN) {
            std::memcpy(dst, src, len);
            return;
        }
        {
            std::lock_guard lock(m_mutex);
            m_dst = static_cast<char*>(dst);
            m_src = static_cast<const char*>(src);
            m_len = len;
            m_per = (len + PARTS - 1) / PARTS;
            m_pending = PARTS - 1;
            ++m_generation;
        }
        m_wake.notify_all();
        part(0);
        std::unique_lock lock(m_mutex);
        m_finished.wait(lock, [this] { return m_pending == 0; });
    }

private:
    void part(int k) const
    {
        const std::size_t b = std::min(m_len, k * m_per), e = std::min(m_len, b + m_per);
        if (b < e)
            std::memcpy(m_dst + b, m_src + b, e - b);
    }

    void work(int k)
    {
        unsigned seen = 0;
        for (;;) {
            {
                std::unique_lock lock(m_mutex);
                m_wake.wait(lock, [&] { return m_quit || m_generation != seen; });
                if (m_quit)
                    return;
                seen = m_generation;
            }
            part(k);
            bool last = false;
            {
                std::lock_guard lock(m_mutex);
                last = --m_pending == 0;
            }
            if (last)
                m_finished.notify_one();
        }
    }

    std::thread m_threads[PARTS - 1];
    bool m_running = false;
    std::mutex m_mutex;
    std::condition_variable m_wake, m_finished;
    bool m_quit = false;
    unsigned m_generation = 0;
    int m_pending = 0;
    char* m_dst = nullptr;
    const char* m_src = nullptr;
    std::size_t m_len = 0, m_per = 0;
};

}
