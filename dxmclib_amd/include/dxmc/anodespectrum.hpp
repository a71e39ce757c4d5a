// Thick-target bremsstrahlung of a tungsten anode and its K lines: the
// spectrum model of the reference (betheHeitlerCrossSection.hpp:174-407,
// used by tube.hpp:189-258).
//
// The intensity at photon energy hv leaving the anode at take-off angle a is
//   I(hv) = sum_x [ sum_u BH(hv, T0 u) rho(u, x) du ] exp(-mu_W(hv) x / sin a) dx
// with BH the semi-relativistic Bethe-Heitler cross section, rho the density
// of electrons at relative energy u and mass depth x (tabulated 100 keV data
// rescaled by the Thomson-Whiddington range, anodedata.hpp) and the sums
// taken as the reference takes them: x = 0, 0.1, ... <= 14 mg/cm^2 and
// u = 0.005, 0.010, ... <= 1, each formed by repeated addition, and every
// partial sum accumulated in the same order.
//
// The bracketed inner sum depends on (T0, hv) only, and rho on T0 only, so
// both are computed once and kept: a spectrum at another take-off angle
// (the heel filter asks for several) or another filtration costs one
// exponential per depth. The values are the ones the nested loop gives,
// bit for bit, since every term is formed by the same operations.
#pragma once

#include "dxmc/anodedata.hpp"
#include "dxmc/constants.hpp"
#include "dxmc/material.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <utility>
#include <vector>

namespace dxmc::anode {

inline constexpr int TUNGSTEN = 74;
inline constexpr double TABLE_ENERGY = 100; // keV, energy of the tabulated electron densities
inline constexpr double X_STEP = 0.1, X_MAX = 14; // mg/cm^2
inline constexpr double U_STEP = 0.005, U_MAX = 1;

// Thomson-Whiddington range [mg/cm^2] of an electron of energy T0 [keV] in
// tungsten (valid for 50-150 keV)
inline double twRange(double T0) { return 0.0119 * std::pow(T0, 1.513); }

// First of the two grid points used at v: the first point not below v,
// moved back when fewer than two points follow it. A value between two
// points therefore uses the segment that starts at the upper one (the
// reference's choice, kept: it sets the interpolated values).
inline int segment(const double* grid, int n, double v)
{
    const int i = static_cast<int>(std::lower_bound(grid, grid + n, v) - grid);
    const int left = n - i;
    return left < 2 ? n - 3 + left : i;
}

// Thomson-Whiddington law: (T/T0)^2 of an electron of initial energy T0 at
// depth x, clamped at 0
inline double twLaw(double x, double T0)
{
    const double* v = anodedata::twVoltage;
    const double* c = anodedata::twConstant;
    const int i = segment(v, anodedata::N_TW, T0);
    const double C = c[i] + ((c[i + 1] - c[i]) / (v[i + 1] - v[i])) * (T0 - v[i]);
    const double f = (T0 * T0 - C * x) / (T0 * T0);
    return f < 0 ? 0 : f;
}

// number fractions of the two electron populations at depth x
inline double fractionF(double x, double T0) { return std::pow(twLaw(x, T0), 1.753); }
inline double fractionM(double x, double T0)
{
    constexpr double K = 18.0, Bd = 0.584, B0 = 0.5;
    const double e = 1 - std::exp(-K * x / twRange(T0));
    const double F = Bd * e;
    const double B = B0 + (Bd - B0) * e;
    return fractionF(x, T0) * B * (F + 1) / (1 - B * F);
}

// Semi-relativistic Bethe-Heitler cross section of tungsten for emitting hv
// from an electron of kinetic energy Ti, with the Elwert Coulomb factor
inline double betheHeitler(double hv, double Ti)
{
    constexpr double m = ELECTRON_REST_MASS<double>();
    constexpr double alpha = 7.29735308E-03;
    constexpr double re = 2.81794092E-15; // m
    constexpr double phiBar = (TUNGSTEN * TUNGSTEN) * re * re * alpha;
    constexpr double k = (phiBar * 2) / 3;
    const double Ei = m + Ti;
    const double Ef = Ei - hv;
    const double pi2 = Ei * Ei - m * m;
    const double pi = std::sqrt(pi2);
    const double pf2 = Ef * Ef - m * m;
    if (pf2 <= 0)
        return 0;
    const double pf = std::sqrt(pf2);
    const double L = 2 * std::log((Ei * Ef + pi * pf - m * m) / (m * hv));
    const double coulomb = pi / pf;
    return k * (4 * Ei * Ef * L - 7 * pi * pf) / (hv * pi * pi) * coulomb;
}

// The integration nodes and the electron density at every node for one
// tube voltage: rho[ix * u.size() + iu].
struct DepthGrid {
    double T0 = 0;
    std::vector<double> x, u, rho;
};

inline DepthGrid makeDepthGrid(double T0)
{
    DepthGrid g;
    g.T0 = T0;
    for (double x = 0; x <= X_MAX; x = x + X_STEP)
        g.x.push_back(x);
    for (double u = U_STEP; u <= U_MAX; u = u + U_STEP)
        g.u.push_back(u);
    const double scale = twRange(TABLE_ENERGY) / twRange(T0);
    const std::size_t nu = g.u.size();
    g.rho.resize(g.x.size() * nu);
    // bilinear interpolation of a [u][depth] table at (u, x), both clamped
    // to the table's grid
    auto table = [](const double (&t)[anodedata::N_U][anodedata::N_X], const double* depth, double uv, double xv) {
        constexpr int NX = anodedata::N_X, NU = anodedata::N_U;
        const double* ug = anodedata::u;
        const double x = std::clamp(xv, depth[0], depth[NX - 1]);
        const double u = std::clamp(uv, ug[0], ug[NU - 1]);
        const int ix = segment(depth, NX, x), iu = segment(ug, NU, u);
        const double x1 = depth[ix], x2 = depth[ix + 1], u1 = ug[iu], u2 = ug[iu + 1];
        const double a1 = (x2 - x) / (x2 - x1);
        const double a2 = (x - x1) / (x2 - x1);
        const double r1 = a1 * t[iu][ix] + a2 * t[iu][ix + 1];
        const double r2 = a1 * t[iu + 1][ix] + a2 * t[iu + 1][ix + 1];
        return ((u2 - u) / (u2 - u1)) * r1 + ((u - u1) / (u2 - u1)) * r2;
    };
    for (std::size_t ix = 0; ix < g.x.size(); ++ix) {
        const double x = g.x[ix];
        const double fF = fractionF(x, T0), fM = fractionM(x, T0);
        for (std::size_t iu = 0; iu < nu; ++iu) {
            const double u = g.u[iu];
            g.rho[ix * nu + iu] = fF * table(anodedata::densityF, anodedata::depthF, u, x * scale)
                + fM * table(anodedata::densityM, anodedata::depthM, u, x * scale);
        }
    }
    return g;
}

// Per-process cache of the energy-independent grids and of the inner sums
// S(x) = sum_u BH(hv, T0 u) rho(u, x) du, keyed by the exact bits of T0 and
// hv. A bounded number of tube voltages is kept.
class SpectrumCache {
public:
    static SpectrumCache& instance()
    {
        static SpectrumCache c;
        return c;
    }

    // bremsstrahlung intensity at hv for tube voltage T0 and take-off angle
    double bremsstrahlung(double T0, double hv, double takeoff)
    {
        if (hv <= 0)
            return 0;
        const auto sums = innerSums(T0, hv);
        const auto& x = sums->first;
        const auto& s = sums->second;
        const double mu = Material::getTotalAttenuation(TUNGSTEN, hv);
        const double sinA = std::sin(takeoff);
        double I = 0;
        for (std::size_t ix = 0; ix < s.size(); ++ix)
            I = I + s[ix] * std::exp(-mu * x[ix] * 0.001 / sinA) * X_STEP;
        return I;
    }

private:
    using Sums = std::pair<std::vector<double>, std::vector<double>>; // depths, S(x)
    static constexpr std::size_t MAX_VOLTAGES = 64;

    struct Voltage {
        std::shared_ptr<const DepthGrid> grid;
        std::map<std::uint64_t, std::shared_ptr<const Sums>> sums;
    };

    static std::uint64_t bits(double v)
    {
        std::uint64_t b;
        std::memcpy(&b, &v, sizeof b);
        return b;
    }

    std::shared_ptr<const Sums> innerSums(double T0, double hv)
    {
        std::shared_ptr<const DepthGrid> grid;
        {
            std::lock_guard lock(m_mutex);
            auto& v = voltage(T0);
            auto it = v.sums.find(bits(hv));
            if (it != v.sums.end())
                return it->second;
            grid = v.grid;
        }
        if (!grid) {
            auto g = std::make_shared<const DepthGrid>(makeDepthGrid(T0));
            std::lock_guard lock(m_mutex);
            auto& v = voltage(T0);
            if (!v.grid)
                v.grid = g;
            grid = v.grid;
        }
        const std::size_t nu = grid->u.size();
        std::vector<double> bh(nu);
        for (std::size_t iu = 0; iu < nu; ++iu)
            bh[iu] = betheHeitler(hv, grid->T0 * grid->u[iu]);
        auto out = std::make_shared<Sums>();
        out->first = grid->x;
        out->second.resize(grid->x.size());
        for (std::size_t ix = 0; ix < grid->x.size(); ++ix) {
            const double* r = &grid->rho[ix * nu];
            double step = 0;
            for (std::size_t iu = 0; iu < nu; ++iu)
                step = step + bh[iu] * r[iu] * U_STEP;
            out->second[ix] = step;
        }
        std::lock_guard lock(m_mutex);
        return voltage(T0).sums.emplace(bits(hv), std::move(out)).first->second;
    }

    // the entry of T0 (caller holds the mutex); the cache is emptied when
    // it would exceed MAX_VOLTAGES tube voltages
    Voltage& voltage(double T0)
    {
        const auto key = bits(T0);
        auto it = m_voltages.find(key);
        if (it == m_voltages.end()) {
            if (m_voltages.size() >= MAX_VOLTAGES)
                m_voltages.clear();
            it = m_voltages.emplace(key, Voltage {}).first;
        }
        return it->second;
    }

    std::mutex m_mutex;
    std::map<std::uint64_t, Voltage> m_voltages;
};

// the tungsten K lines: energy [keV] and relative intensity
inline constexpr std::array<std::pair<double, double>, 4> K_LINES { { { 59.3, 0.505 }, { 58.0, 0.291 }, { 67.2, 0.162 }, { 69.1, 0.042 } } };

// intensity of a K line: (1 + P) f r_K omega_K times the bremsstrahlung at
// the line energy (P = 0.33, r_K = 4.4, omega_K = 0.94)
inline double kLineIntensity(double T0, double energy, double fraction, double takeoff)
{
    constexpr double P = 0.33, omegaK = 0.94, rK = 4.4;
    return (1 + P) * fraction * rK * omegaK * SpectrumCache::instance().bremsstrahlung(T0, energy, takeoff);
}

}
