// Random number generation and the host side of the samplers.
//
// RandomState        PCG32 XSH-RR (reference dxmcrandom.hpp:37-167); the same
//                    generator runs on the device (csrc/device/rng.hpp).
// RandomDistribution alias-method table (reference :173-279)
// SpecterDistribution energy spectrum sampler (reference :284-330)
// RITA               adaptive rational inverse-CDF table (reference :333-501)
//
// The table builders below produce the arrays the GPU kernels sample from.
// Their arithmetic follows the reference step by step (same expressions in
// the same order) because stream-level parity needs identical tables: any
// alias table samples the same distribution, but only the identical one maps
// the same random draws to the same bin.
#pragma once

#include "dxmc/floating.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <concepts>
#include <cstdint>
#include <random>
#include <type_traits>
#include <vector>

namespace dxmc {

// splitmix64 finaliser used to derive one PCG stream per history
constexpr std::uint64_t splitmix64(std::uint64_t x) noexcept
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

class RandomState {
public:
    // seeded from the machine's random device, as the reference's default
    RandomState()
    {
        std::random_device d;
        std::uniform_int_distribution<std::uint64_t> dist(0);
        m_state[0] = dist(d);
        m_state[1] = dist(d);
    }
    RandomState(std::uint64_t state[2])
    {
        m_state[0] = state[0];
        m_state[1] = state[1];
    }
    RandomState(std::uint64_t s0, std::uint64_t s1)
    {
        m_state[0] = s0;
        m_state[1] = s1;
    }
    // stream of history h under a run seed: state1 = 2h+1 (odd increment,
    // one distinct PCG stream per history), state0 = splitmix64(seed ^ splitmix64(h))
    static RandomState forHistory(std::uint64_t seed, std::uint64_t h)
    {
        return RandomState(splitmix64(seed ^ splitmix64(h)), 2 * h + 1);
    }
    RandomState(const RandomState&) = delete;
    RandomState& operator=(const RandomState&) = delete;

    // u32 * 2^-32; for float the converted integer rounds, so 1.0f is possible
    template <typename T>
    inline T randomUniform() noexcept
    {
        static_assert(std::is_floating_point<T>::value);
        constexpr T scale = T { 2.32830643653869628906e-010 };
        return pcg32() * scale;
    }

    template <typename T>
    inline T randomUniform(const T max) noexcept
    {
        if constexpr (std::is_floating_point<T>::value) {
            return randomUniform<T>() * max;
        } else {
            // rejection threshold computed in the width of T, then truncated to 32
            // bits, exactly as the reference does (it is 2^64 mod max, not 2^32 mod max)
            const std::uint32_t threshold = static_cast<std::uint32_t>(static_cast<T>(-max) % max);
            for (;;) {
                const std::uint32_t r = pcg32();
                if (r >= threshold)
                    return static_cast<T>(r % static_cast<std::uint32_t>(max));
            }
        }
    }

    template <typename T>
    inline T randomUniform(const T min, const T max) noexcept
    {
        if constexpr (std::is_floating_point<T>::value) {
            const T r = randomUniform<T>();
            return min + r * (max - min);
        } else {
            return min + randomUniform<T>(max - min);
        }
    }

    // unsigned integer in [0, max) by rejection below (2^w - max) mod max, w
    // the width of T: randomInteger for types up to 32 bits, and
    // randomInteger32BitCapped for wider types whose values fit 32 bits
    // (reference dxmcrandom.hpp:125-147)
    template <std::unsigned_integral T>
    inline T randomInteger(const T max) noexcept
    {
        static_assert(sizeof(T) <= 4, "randomInteger draws up to 32 bits; use randomInteger32BitCapped for wider types");
        return rejectionInteger(max);
    }
    template <std::unsigned_integral T>
    inline T randomInteger32BitCapped(const T max) noexcept
    {
        static_assert(sizeof(T) > 4, "randomInteger32BitCapped is for types wider than 32 bits; use randomInteger");
        return rejectionInteger(max);
    }

    inline std::uint32_t pcg32() noexcept
    {
        const std::uint64_t old = m_state[0];
        m_state[0] = old * 6364136223846793005ULL + (m_state[1] | 1ULL);
        const std::uint32_t xs = static_cast<std::uint32_t>(((old >> 18u) ^ old) >> 27u);
        const std::uint32_t rot = static_cast<std::uint32_t>(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    std::uint64_t m_state[2];

private:
    template <typename T>
    inline T rejectionInteger(const T max) noexcept
    {
        const T threshold = static_cast<T>(-max) % max;
        for (;;) {
            const T r = pcg32();
            if (r >= threshold)
                return r % max;
        }
    }
};

template <Floating T = double>
class RandomDistribution {
public:
    RandomDistribution(const std::vector<T>& weights) { buildAliasTable(weights); }
    RandomDistribution(const RandomDistribution& o)
        : m_state(o.m_state.m_state[0], o.m_state.m_state[1])
        , m_alias(o.m_alias)
        , m_probs(o.m_probs)
    {
    }
    RandomDistribution& operator=(const RandomDistribution& o)
    {
        m_alias = o.m_alias;
        m_probs = o.m_probs;
        return *this;
    }

    std::size_t sampleIndex() { return sampleIndex(m_state); }
    // draw order: the uniform first, then the bin (reference :211-216)
    std::size_t sampleIndex(RandomState& state) const
    {
        const T r = state.randomUniform<T>();
        const std::size_t k = state.randomUniform<std::size_t>(size());
        return r < m_probs[k] ? k : static_cast<std::size_t>(m_alias[k]);
    }
    std::size_t size() const { return m_probs.size(); }
    const std::vector<std::uint64_t>& aliasingData() const { return m_alias; }
    const std::vector<T>& probabilityData() const { return m_probs; }

protected:
    RandomState m_state { 0x853c49e6748fea9bULL, 0xda3e39cb94b95bdbULL };

    // Vose/Walker "square histogram": bins scaled to mean 1, pushed to the
    // small/large work lists from the last bin to the first, paired from the
    // back of both lists; leftovers get probability 1.
    void buildAliasTable(const std::vector<T>& weights)
    {
        const std::size_t n = weights.size();
        m_probs.assign(n, T { 0 });
        m_alias.assign(n, 0);
        T sum = T { 0 };
        for (const T w : weights)
            sum = sum + w;
        const T scale = n / sum;
        std::vector<T> p(n);
        for (std::size_t i = 0; i < n; ++i)
            p[i] = weights[i] * scale;
        std::vector<std::int64_t> small, large;
        small.reserve(n);
        large.reserve(n);
        for (std::int64_t i = static_cast<std::int64_t>(n) - 1; i >= 0; --i)
            (p[i] < T { 1 } ? small : large).push_back(i);
        while (!small.empty() && !large.empty()) {
            const auto s = small.back();
            small.pop_back();
            const auto l = large.back();
            large.pop_back();
            m_probs[s] = p[s];
            m_alias[s] = static_cast<std::uint64_t>(l);
            p[l] = p[l] + p[s] - 1;
            (p[l] < 1 ? small : large).push_back(l);
        }
        for (const auto l : large)
            m_probs[l] = 1;
        for (const auto s : small)
            m_probs[s] = 1;
    }

private:
    std::vector<std::uint64_t> m_alias;
    std::vector<T> m_probs;
};

template <Floating T = double>
class SpecterDistribution : public RandomDistribution<T> {
public:
    SpecterDistribution(const std::vector<T>& weights, const std::vector<T>& energies)
        : RandomDistribution<T>(weights)
        , m_energies(energies)
    {
    }
    SpecterDistribution()
        : RandomDistribution<T>(std::vector<T> { 1 })
        , m_energies { 60 }
    {
    }
    T sampleValue() { return sampleValue(this->m_state); }
    // uniform inside the sampled bin [E_k, E_k+1); the last bin returns E_last
    T sampleValue(RandomState& state) const
    {
        const std::size_t k = this->sampleIndex(state);
        return k < m_energies.size() - 1 ? state.randomUniform(m_energies[k], m_energies[k + 1]) : m_energies[k];
    }
    const std::vector<T>& energies() const { return m_energies; }

private:
    std::vector<T> m_energies;
};

// Adaptive rational inverse transform sampling (Penelope's RITA, as used by
// the reference for the Rayleigh form factor). N nodes are placed by starting
// from 10 equidistant nodes and repeatedly bisecting the interval with the
// largest interpolation error.
template <Floating T, int N = 20>
class RITA {
public:
    template <typename F>
    RITA(const T min, const T max, F pdf)
    {
        struct Node {
            T x, e, a, b, error;
        };
        std::size_t n = 10;
        std::vector<Node> v(n);
        v.reserve(N);
        for (std::size_t i = 0; i < n; ++i) {
            v[i].x = min + i * (max - min) / (n - 1);
            v[i].e = 0;
            v[i].a = 0;
            v[i].b = 0;
            v[i].error = -1;
        }
        auto cumulate = [&]() {
            for (std::size_t j = 1; j < n; ++j)
                v[j].e = v[j - 1].e + simpson(v[j - 1].x, v[j].x, pdf);
        };
        auto normalise = [&]() {
            for (std::size_t j = 1; j < n; ++j)
                v[j].e = v[j].e / v[n - 1].e;
        };
        cumulate();
        const T total = v[n - 1].e; // integral of the un-normalised pdf
        normalise();
        while (n != N) {
            for (std::size_t i = 0; i + 1 < n; ++i) {
                if (!(v[i].error < 0)) // fitted already (a NaN error is never refitted)
                    continue;
                const T dx = v[i + 1].x - v[i].x;
                const T slope = (v[i + 1].e - v[i].e) / dx;
                const T p0 = pdf(v[i].x) / total;
                const T p1 = pdf(v[i + 1].x) / total;
                v[i].b = (p0 > 0 && p1 > 0) ? 1 - slope * slope / (p0 * p1) : 0;
                v[i].a = p0 > 0 ? slope / p0 - v[i].b - 1 : 0;
                v[i].error = 0;
                for (std::size_t j = 1; j < 50; ++j) {
                    const T x = v[i].x + j * (v[i + 1].x - v[i].x) / 50;
                    const T t = (x - v[i].x) / (v[i + 1].x - v[i].x);
                    const T a = v[i].a;
                    const T b = v[i].b;
                    const T f = (1 + a + b - a * t);
                    const T nn = f * (1 - std::sqrt(1 - 4 * b * t * t / (f * f))) / (2 * b * t);
                    const T p1n = (1 + a * nn + b * nn * nn);
                    const T p = p1n * p1n * (v[i + 1].e - v[i].e) / ((1 + a + b) * (1 - b * nn * nn) * (v[i + 1].x - v[i].x));
                    v[i].error += std::abs(p - pdf(x)) * (v[i + 1].x - v[i].x) / 50;
                }
            }
            // first node with the largest error (the comparison std::max_element makes)
            std::size_t worst = 0;
            for (std::size_t i = 1; i < v.size(); ++i)
                if (v[worst].error < v[i].error)
                    worst = i;
            const T x0 = v[worst].x;
            v[worst].error = -1;
            const T x1 = v[worst + 1].x;
            v.insert(v.begin() + static_cast<std::ptrdiff_t>(worst + 1), Node { x0 + (x1 - x0) / 2, 0, 0, 0, -1 });
            ++n;
            cumulate();
            normalise();
        }
        for (std::size_t i = 0; i < static_cast<std::size_t>(N); ++i) {
            m_x[i] = v[i].x;
            m_e[i] = v[i].e;
            m_a[i] = v[i].a;
            m_b[i] = v[i].b;
        }
    }

    // sample below maxValue: draws are confined to the CDF of the node that
    // brackets maxValue from above and rejected while the result exceeds it
    T operator()(RandomState& state, const T maxValue) const
    {
        const std::size_t im = std::upper_bound(m_x.cbegin(), m_x.cend(), maxValue) - m_x.cbegin();
        const T cdfMax = im != static_cast<std::size_t>(N) ? m_e[im] : T { 1 };
        T res;
        do {
            const T r = state.randomUniform<T>(cdfMax);
            std::size_t i = std::upper_bound(m_e.cbegin(), m_e.cend(), r) - m_e.cbegin() - 1;
            i = std::min<std::size_t>(i, N - 2); // r == 1 would index past the table
            res = interval(i, r);
        } while (res > maxValue);
        return res;
    }
    T operator()(RandomState& state) const
    {
        const T r = state.randomUniform<T>();
        const std::size_t i = std::upper_bound(m_e.cbegin(), m_e.cend(), r) - m_e.cbegin();
        if (i == 0)
            return m_x[0];
        return interval(std::min<std::size_t>(i - 1, N - 2), r);
    }

    const std::array<T, N>& x() const { return m_x; }
    const std::array<T, N>& e() const { return m_e; }
    const std::array<T, N>& a() const { return m_a; }
    const std::array<T, N>& b() const { return m_b; }

protected:
    template <typename F>
    static T simpson(const T start, const T stop, F pdf)
    {
        const T h = (stop - start) / 50;
        T s = pdf(start) + pdf(stop);
        for (std::size_t i = 1; i < 50; ++i) {
            const T w = i % 2 == 0 ? 2 : 4;
            s += w * pdf(start + h * i);
        }
        return h * s / 3;
    }
    T interval(std::size_t i, T r) const
    {
        const T v = r - m_e[i];
        const T d = m_e[i + 1] - m_e[i];
        return m_x[i] + (1 + m_a[i] + m_b[i]) * d * v / (d * d + m_a[i] * d * v + m_b[i] * v * v) * (m_x[i + 1] - m_x[i]);
    }

private:
    std::array<T, N> m_x {};
    std::array<T, N> m_e {};
    std::array<T, N> m_a {};
    std::array<T, N> m_b {};
};
}
