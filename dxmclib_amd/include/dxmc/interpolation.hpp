// Interpolation helpers (reference interpolation.hpp:33-177).
//
// CubicSplineInterpolator<T,N> is a natural cubic spline through N nodes on
// a uniform grid whose step is (stop-start)/(N-2), so the last node lies one
// step beyond 'stop' (reference quirk, kept for parity). Each interval stores
// the cubic's coefficients in absolute x: y = c0 + c1 x + c2 x^2 + c3 x^3.
// The same coefficients are evaluated on the device.
#pragma once

#include "dxmc/floating.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <vector>

namespace dxmc {

// Sum in the order libstdc++ evaluates std::reduce(first, last, init) for
// random access ranges: blocks of four as ((a+b)+(c+d)) folded into init,
// then the remainder one by one. Reproducing it keeps host tables that the
// reference builds with std::reduce bit-identical.
template <typename It, typename U>
inline U stlReduce(It first, It last, U init)
{
    while (last - first >= 4) {
        const U v1 = first[0] + first[1];
        const U v2 = first[2] + first[3];
        init = init + (v1 + v2);
        first += 4;
    }
    for (; first != last; ++first)
        init = init + *first;
    return init;
}

template <Floating T>
inline T interp(T x0, T x1, T y0, T y1, T x)
{
    return y0 + (y1 - y0) * (x - x0) / (x1 - x0);
}

template <Floating T>
inline T interpolate(const std::vector<T>& xs, const std::vector<T>& ys, T x)
{
    const auto up = std::upper_bound(xs.cbegin(), xs.cend(), x);
    if (up == xs.cbegin())
        return ys.front();
    if (up == xs.cend())
        return ys.back();
    const auto i = static_cast<std::size_t>(up - xs.cbegin());
    return interp(xs[i - 1], xs[i], ys[i - 1], ys[i], x);
}

template <Floating T, int N = 30>
class CubicSplineInterpolator {
public:
    template <typename F>
    CubicSplineInterpolator(const T start, const T stop, F function)
    {
        std::vector<T> y(N);
        m_start = start;
        m_step = (stop - start) / (N - 2);
        for (std::size_t i = 0; i < N; ++i) {
            m_x[i] = m_start + m_step * i;
            y[i] = function(m_x[i]);
        }
        m_stop = m_x[N - 1];

        // second derivatives from the tridiagonal system with natural ends
        std::vector<T> h(N, T { 0 }), slope(N, T { 1 }), diag(N), rhs(N);
        for (std::size_t i = 0; i + 1 < N; ++i)
            h[i] = m_x[i + 1] - m_x[i];
        for (std::size_t i = 0; i + 1 < N; ++i)
            slope[i] = (y[i + 1] - y[i]) / h[i];
        for (std::size_t i = 1; i < N; ++i)
            diag[i] = 2 * (h[i - 1] + h[i]);
        diag[0] = diag[1];
        for (std::size_t i = 1; i < N; ++i)
            rhs[i] = 6 * (slope[i] - slope[i - 1]);
        rhs[N - 1] = 0;
        rhs[0] = 0;
        const auto s = solveTridiagonal(h, diag, rhs);

        for (std::size_t i = 0; i + 1 < N; ++i) {
            const T x0 = m_x[i], x1 = m_x[i + 1];
            T* c = &m_coefficients[i * 4];
            c[0] = (s[i] * x1 * x1 * x1 - s[i + 1] * x0 * x0 * x0 + 6 * (y[i] * x1 - y[i + 1] * x0)) / (6 * h[i]);
            c[0] += h[i] * (s[i + 1] * x0 - s[i] * x1) / 6;
            c[1] = (s[i + 1] * x0 * x0 - s[i] * x1 * x1 + 2 * (y[i + 1] - y[i])) / (2 * h[i]) + h[i] * (s[i] - s[i + 1]) / 6;
            c[2] = (s[i] * x1 - s[i + 1] * x0) / (2 * h[i]);
            c[3] = (s[i + 1] - s[i]) / (6 * h[i]);
        }
    }

    T operator()(const T xv) const
    {
        const T x = std::clamp(xv, m_start, m_stop);
        const std::size_t k = x > m_start ? static_cast<std::size_t>((x - m_start) / m_step) : 0;
        const std::size_t o = k < N - 1 ? k * 4 : (N - 2) * 4;
        return m_coefficients[o] + m_coefficients[o + 1] * x + m_coefficients[o + 2] * x * x + m_coefficients[o + 3] * x * x * x;
    }

    const std::array<T, (N - 1) * 4>& coefficients() const { return m_coefficients; }
    T start() const { return m_start; }
    T step() const { return m_step; }
    T stop() const { return m_stop; }

protected:
    // Thomas elimination with off diagonals h (sub == super) and diagonal d
    static std::vector<T> solveTridiagonal(const std::vector<T>& h, std::vector<T> d, std::vector<T> r)
    {
        const std::size_t n = r.size();
        for (std::size_t i = 1; i < n; ++i) {
            const T w = h[i - 1] / d[i - 1];
            d[i] -= w * h[i - 1];
            r[i] -= w * r[i - 1];
        }
        std::vector<T> x(n);
        x[n - 1] = r[n - 1] / d[n - 1];
        for (std::size_t i = n - 1; i-- > 0;)
            x[i] = (r[i] - h[i] * x[i + 1]) / d[i];
        return x;
    }

private:
    std::array<T, (N - 1) * 4> m_coefficients {};
    std::array<T, N> m_x {};
    T m_step = 0;
    T m_start = 0;
    T m_stop = 0;
};
}
