// Host arrays kept for reuse by the request server.
//
// A fresh large allocation costs one page fault per 4 KiB page at its first
// touch: tens of milliseconds for the arrays of a 256^3 request, as much as
// its transport. An array handed back to the pool keeps its pages, so the
// next request that fits fills memory that is already mapped. Arrays beyond
// the pool's byte limit are freed as usual.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <sys/mman.h>
#include <vector>

namespace dxmc_host_detail {

template <typename T>
class VectorPool {
public:
    using Handle = std::shared_ptr<std::vector<T>>;

    explicit VectorPool(std::size_t maxBytes)
        : m_max(maxBytes)
    {
    }
    VectorPool(const VectorPool&) = delete;
    VectorPool& operator=(const VectorPool&) = delete;

    // A vector with capacity for n elements and unspecified contents; it
    // returns to the pool when the last handle goes. A pooled vector is
    // taken only if it is at most 8 times larger than asked for.
    Handle take(std::size_t n)
    {
        std::unique_ptr<std::vector<T>> v;
        {
            std::lock_guard lock(m_mutex);
            auto it = m_free.lower_bound(n);
            if (it != m_free.end() && it->first <= 8 * n + 4096) {
                v = std::move(it->second);
                m_held -= it->first * sizeof(T);
                m_free.erase(it);
            }
        }
        if (!v) {
            v = std::make_unique<std::vector<T>>();
            v->reserve(n);
        }
        return Handle(v.release(), [this](std::vector<T>* p) { give(p); });
    }

    std::size_t heldBytes()
    {
        std::lock_guard lock(m_mutex);
        return m_held;
    }

private:
    void give(std::vector<T>* p)
    {
        std::unique_ptr<std::vector<T>> v(p);
        const std::size_t bytes = v->capacity() * sizeof(T);
        std::lock_guard lock(m_mutex);
        if (bytes == 0 || m_held + bytes > m_max)
            return;
        m_held += bytes;
        m_free.emplace(v->capacity(), std::move(v));
    }

    std::mutex m_mutex;
    std::multimap<std::size_t, std::unique_ptr<std::vector<T>>> m_free;
    std::size_t m_held = 0;
    const std::size_t m_max;
};

// Raw blocks for arrays that are written whole before they are read (the
// result tallies): never initialised, 2 MiB aligned and advised for
// transparent huge pages, so a first touch faults once per 2 MiB instead of
// once per 4 KiB; reused like the vectors above. They stay pageable: the
// device would fill page-locked blocks faster (2 ms instead of 12 ms per
// request on the serving stream), but the host then reads them about five
// times slower, and every result is read once by its collector.
class BlockPool {
public:
    static constexpr std::size_t ALIGN = std::size_t { 2 } << 20;
    using Handle = std::shared_ptr<std::byte>;

    explicit BlockPool(std::size_t maxBytes)
        : m_max(maxBytes)
    {
    }
    BlockPool(const BlockPool&) = delete;
    BlockPool& operator=(const BlockPool&) = delete;
    ~BlockPool()
    {
        for (auto& [bytes, b] : m_free)
            release(b);
    }

    Handle take(std::size_t bytes)
    {
        Block b {};
        {
            std::lock_guard lock(m_mutex);
            auto it = m_free.lower_bound(bytes);
            if (it != m_free.end() && it->first <= 8 * bytes + ALIGN) {
                b = it->second;
                m_held -= b.bytes;
                m_free.erase(it);
            }
        }
        if (!b.p) {
            b.bytes = std::max<std::size_t>(ALIGN, (bytes + ALIGN - 1) / ALIGN * ALIGN);
            b.p = std::aligned_alloc(ALIGN, b.bytes);
            if (!b.p)
                throw std::bad_alloc();
            madvise(b.p, b.bytes, MADV_HUGEPAGE);
        }
        return Handle(static_cast<std::byte*>(b.p), [this, b](std::byte*) { give(b); });
    }

private:
    struct Block {
        void* p = nullptr;
        std::size_t bytes = 0;
    };

    static void release(const Block& b) { std::free(b.p); }

    void give(const Block& b)
    {
        {
            std::lock_guard lock(m_mutex);
            if (m_held + b.bytes <= m_max) {
                m_held += b.bytes;
                m_free.emplace(b.bytes, b);
                return;
            }
        }
        release(b);
    }

    std::mutex m_mutex;
    std::multimap<std::size_t, Block> m_free;
    std::size_t m_held = 0;
    const std::size_t m_max;
};

}
