// Progress, ETA, cancellation and dose preview of a running transport
// (reference progressbar.hpp:43-235). The GPU run loop writes the number of
// finished exposures into exposureCounter() and polls cancelFlag() between
// kernel launches, so another thread can watch getETA() and call
// setCancel(true) exactly as with the reference's CPU workers. The dose
// preview (computeDoseProgressImage) is the maximum intensity projection of
// the tally so far: the run loop computes it on the device every
// previewInterval() seconds and hands it over through previewCallback().
#pragma once

#include "dxmc/floating.hpp"

#include <array>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <iomanip>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

namespace dxmc {

template <Floating T>
struct DoseProgressImageData {
    std::array<std::size_t, 2> dimensions = { 0, 0 };
    std::array<T, 2> spacing = { 0, 0 };
    std::vector<std::uint8_t> image;
};

template <Floating T>
class ProgressBar {
public:
    enum class Axis { X,
        Y,
        Z };
    ProgressBar() { }
    ProgressBar(std::uint64_t total) { setTotalExposures(total); }
    void setTotalExposures(std::uint64_t total)
    {
        m_total = total;
        m_done = 0;
        m_start = std::chrono::steady_clock::now();
    }
    std::uint64_t totalExposures() const { return m_total; }
    std::uint64_t completedExposures() const { return m_done; }
    void setPrefixMessage(const std::string& msg) { m_message = msg; }
    const std::string& prefixMessage() const { return m_message; }
    void exposureCompleted() { ++m_done; }
    std::string getETA() const
    {
        const std::uint64_t done = m_done;
        if (done == 0)
            return m_message + "ETA: estimating... [0%]";
        const std::uint64_t total = m_total;
        const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - m_start).count();
        const double remaining = elapsed / done * (total - done);
        char buf[128];
        std::snprintf(buf, sizeof(buf), "ETA: %.0f s [%.0f%%]", remaining, 100.0 * done / (total ? total : 1));
        return m_message + buf;
    }
    void setCancel(bool c) { m_cancel = c ? 1 : 0; }
    bool cancel(void) const { return m_cancel != 0; }
    void setPlaneNormal(Axis a) { m_axis = a; }
    Axis planeNormal() const { return m_axis; }

    // the device run loop writes / polls these through the C ABI, with atomic
    // release stores and acquire loads of the same width as these atomics
    volatile std::uint64_t* exposureCounter() { return reinterpret_cast<volatile std::uint64_t*>(&m_done); }
    const volatile std::int32_t* cancelFlag() const { return reinterpret_cast<const volatile std::int32_t*>(&m_cancel); }

    // seconds between two device previews during a run (0: after every launch)
    void setPreviewInterval(double seconds) { m_previewInterval = seconds < 0 ? 0 : seconds; }
    double previewInterval() const { return m_previewInterval; }
    // receives a device preview (dxmc_preview_fn): the projection of the dose
    // along the plane normal, width x height, first index fastest
    static void previewCallback(void* self, const float* image, std::uint64_t width, std::uint64_t height)
    {
        auto* pb = static_cast<ProgressBar*>(self);
        std::scoped_lock lock(pb->m_mutex);
        pb->m_preview.assign(image, image + width * height);
        pb->m_previewSize = { width, height };
    }

    void setDoseData(const T* dose, const std::array<std::size_t, 3>& dims, const std::array<T, 3>& spacing)
    {
        std::scoped_lock lock(m_mutex);
        m_dose = dose;
        m_dims = dims;
        m_spacing = spacing;
    }
    void clearDoseData()
    {
        std::scoped_lock lock(m_mutex);
        m_dose = nullptr;
        m_preview.clear();
    }
    // 8-bit maximum intensity projection of the dose along the plane normal,
    // scaled to 255 at the maximum: from the latest device preview while a
    // run goes on, else from the attached dose array
    std::shared_ptr<DoseProgressImageData<T>> computeDoseProgressImage()
    {
        std::scoped_lock lock(m_mutex);
        if (!m_dose && m_preview.empty())
            return nullptr;
        const int a = m_axis == Axis::X ? 0 : (m_axis == Axis::Y ? 1 : 2);
        const int u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
        auto img = std::make_shared<DoseProgressImageData<T>>();
        img->dimensions = { m_dims[u], m_dims[v] };
        img->spacing = { m_spacing[u], m_spacing[v] };
        std::vector<T> mip;
        if (!m_preview.empty() && m_previewSize[0] == m_dims[u] && m_previewSize[1] == m_dims[v]) {
            mip.assign(m_preview.cbegin(), m_preview.cend());
        } else if (m_dose) {
            mip.assign(m_dims[u] * m_dims[v], T { 0 });
            for (std::size_t z = 0; z < m_dims[2]; ++z)
                for (std::size_t y = 0; y < m_dims[1]; ++y)
                    for (std::size_t x = 0; x < m_dims[0]; ++x) {
                        const std::size_t idx[3] = { x, y, z };
                        const T d = m_dose[(z * m_dims[1] + y) * m_dims[0] + x];
                        T& m = mip[idx[v] * m_dims[u] + idx[u]];
                        if (d > m)
                            m = d;
                    }
        } else {
            return nullptr;
        }
        T mx = 0;
        for (const T m : mip)
            mx = std::max(mx, m);
        const T scale = T { 255 } / mx;
        img->image.resize(mip.size());
        for (std::size_t i = 0; i < mip.size(); ++i)
            img->image[i] = mx > 0 ? static_cast<std::uint8_t>(mip[i] * scale) : 0;
        return img;
    }

private:
    static_assert(sizeof(std::atomic<std::uint64_t>) == sizeof(std::uint64_t) && std::atomic<std::uint64_t>::is_always_lock_free);
    static_assert(sizeof(std::atomic<std::int32_t>) == sizeof(std::int32_t) && std::atomic<std::int32_t>::is_always_lock_free);
    std::atomic<std::uint64_t> m_total { 0 };
    std::atomic<std::uint64_t> m_done { 0 };
    std::atomic<std::int32_t> m_cancel { 0 };
    std::chrono::steady_clock::time_point m_start = std::chrono::steady_clock::now();
    std::string m_message;
    Axis m_axis = Axis::Z;
    std::mutex m_mutex;
    const T* m_dose = nullptr;
    std::array<std::size_t, 3> m_dims = { 0, 0, 0 };
    std::array<T, 3> m_spacing = { 0, 0, 0 };
    double m_previewInterval = 1.0;
    std::vector<T> m_preview;
    std::array<std::size_t, 2> m_previewSize = { 0, 0 };
};
}
