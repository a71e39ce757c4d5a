// ORACLE - test infrastructure, not product code.
//
// CPU restatement of DXMClib's photon transport hot path, used by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker of
// the HIP backend, never as the thing measured or shipped. Written separately
// from the product (dxmclib_amd/): the only shared code is the physics data
// provider (dxmclib_amd/src/material.cpp), which stands in for xraylib and is
// an input to both sides, not an algorithm under test.
//
// Every function cites the reference file:line it restates (paths under the
// reference's include/dxmc/). Floating point expressions keep the reference's
// operation order; the library is built with -ffp-contract=off.
//
// How each part is pinned (tests/test_oracle_*.py):
//   PCG32, uniforms, integer draws         reference build (oracle/_ref), PCG
//                                          definition, hand values
//   alias tables + spectrum sampling       reference build; P(i) invariant
//   RITA build + sampling                  reference build; analytic x^2 pdf
//   cubic spline                           reference build; node interpolation
//   rotate / peturb / world entry          reference build; geometry identities
//   Klein-Nishina, Thomson (L=0)           reference build; chi-square vs pdf
//   attenuation LUT + majorant             exactness at nodes, log-log
//                                          linearity, majorant >= rho*mu,
//                                          lagged segment extrapolation
//   Livermore Compton, RITA Rayleigh (L=1) chi-square against KN x S(q) and
//                                          (1+cos^2) F^2(q); energy conservation
//   IA Compton (L=2)                       pz recovered from (E, E', cos) against
//                                          the analytic profile (KS), shell
//                                          choice, Q9, Klein-Nishina limit,
//                                          angle distribution
//   L=2 fluorescence                       line probabilities in closed form
//   bowtie / heel weights                  np.interp of |angle| (the reference's
//                                          bowtie does not compile with
//                                          libstdc++); heel index rules in
//                                          closed form; product heel table =
//                                          normalised tube spectrum
//   Woodcock tracking                      uncollided transmission exp(-mu L),
//                                          majorant independence, energy balance
//   normalisation                          closed form
//   tube spectrum (Bethe-Heitler anode)     literal nested-loop integral on
//                                          the tables of a fixture extracted
//                                          from the reference; TG-195 RQR-8/9
#include "dxmc/material.hpp"
#include "dxmc_hip.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <numbers>
#include <string>
#include <thread>
#include <vector>

namespace oracle {

// loop bounds shared with the device code: a rejection loop keeps its last
// trial after MAX_TRIALS, a photon is dropped after MAX_STEPS Woodcock steps
constexpr unsigned MAX_TRIALS = 1000;
constexpr unsigned MAX_STEPS = 1u << 20;

// ------------------------------------------------------------------ RNG
// RandomState::pcg32 (dxmcrandom.hpp:156-165), uniforms (:66-118)
struct Rng {
    std::uint64_t state;
    std::uint64_t increment;
    std::uint32_t next()
    {
        const std::uint64_t s = state;
        state = s * 6364136223846793005ULL + (increment | 1ULL);
        const std::uint32_t x = static_cast<std::uint32_t>(((s >> 18u) ^ s) >> 27u);
        const std::uint32_t r = static_cast<std::uint32_t>(s >> 59u);
        return (x >> r) | (x << ((0u - r) & 31u));
    }
    template <typename T>
    T uni()
    {
        const std::uint32_t u = next();
        return u * T { 2.32830643653869628906e-010 };
    }
    template <typename T>
    T uni(T max) { return uni<T>() * max; }
    template <typename T>
    T uni(T lo, T hi)
    {
        const T r = uni<T>();
        return lo + r * (hi - lo);
    }
    // randomUniform<std::size_t>(n) (dxmcrandom.hpp:87-92)
    std::uint64_t index(std::uint64_t n)
    {
        const std::uint32_t thr = static_cast<std::uint32_t>((0ULL - n) % n);
        while (true) {
            const std::uint32_t r = next();
            if (r >= thr)
                return r % static_cast<std::uint32_t>(n);
        }
    }
};

std::uint64_t mix64(std::uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// per history stream: increment 2h+1, state mix64(seed ^ mix64(h))
Rng historyRng(std::uint64_t seed, std::uint64_t h) { return Rng { mix64(seed ^ mix64(h)), 2 * h + 1 }; }

template <typename T>
std::size_t upperIndex(const T* a, std::size_t n, T v)
{
    return static_cast<std::size_t>(std::upper_bound(a, a + n, v) - a);
}

// --------------------------------------------------------- alias table
// RandomDistribution::generateTable (dxmcrandom.hpp:229-274)
void aliasTable(const double* w, std::size_t n, double* prob, std::uint32_t* alias)
{
    double sum = 0.0;
    for (std::size_t i = 0; i < n; ++i)
        sum = sum + w[i];
    const double scale = static_cast<double>(n) / sum;
    std::vector<double> p(n);
    for (std::size_t i = 0; i < n; ++i) {
        p[i] = w[i] * scale;
        prob[i] = 0;
        alias[i] = 0;
    }
    std::vector<std::int64_t> sm(n), lg(n);
    std::int64_t ns = 0, nl = 0;
    for (std::int64_t i = static_cast<std::int64_t>(n) - 1; i >= 0; --i) {
        if (p[i] < 1.0)
            sm[ns++] = i;
        else
            lg[nl++] = i;
    }
    while (ns && nl) {
        const auto s = sm[--ns];
        const auto l = lg[--nl];
        prob[s] = p[s];
        alias[s] = static_cast<std::uint32_t>(l);
        p[l] = p[l] + p[s] - 1;
        if (p[l] < 1)
            sm[ns++] = l;
        else
            lg[nl++] = l;
    }
    while (nl)
        prob[lg[--nl]] = 1;
    while (ns)
        prob[sm[--ns]] = 1;
}

// ------------------------------------------------------------------ RITA
// RITA<T,N> constructor (dxmcrandom.hpp:384-468), tabulated into x,e,a,b
template <typename F>
double simpson50(double a, double b, F f)
{
    const double h = (b - a) / 50;
    double s = f(a) + f(b);
    for (std::size_t i = 1; i < 50; ++i) {
        const double k = i % 2 == 0 ? 2 : 4;
        s += k * f(a + h * i);
    }
    return h * s / 3;
}

template <typename F>
void ritaBuild(double lo, double hi, F pdf, int N, double* ox, double* oe, double* oa, double* ob)
{
    std::vector<double> x, e, a, b, err;
    std::size_t n = 10;
    for (std::size_t i = 0; i < n; ++i) {
        x.push_back(lo + i * (hi - lo) / (n - 1));
        e.push_back(0);
        a.push_back(0);
        b.push_back(0);
        err.push_back(-1);
    }
    auto integrate = [&]() {
        for (std::size_t j = 1; j < n; ++j)
            e[j] = e[j - 1] + simpson50(x[j - 1], x[j], pdf);
    };
    auto scale = [&]() {
        for (std::size_t j = 1; j < n; ++j)
            e[j] = e[j] / e[n - 1];
    };
    integrate();
    const double norm = e[n - 1];
    scale();
    while (n != static_cast<std::size_t>(N)) {
        for (std::size_t i = 0; i + 1 < n; ++i) {
            if (!(err[i] < 0))
                continue;
            const double slope = (e[i + 1] - e[i]) / (x[i + 1] - x[i]);
            const double f0 = pdf(x[i]) / norm;
            const double f1 = pdf(x[i + 1]) / norm;
            b[i] = (f0 > 0 && f1 > 0) ? 1 - slope * slope / (f0 * f1) : 0;
            a[i] = f0 > 0 ? slope / f0 - b[i] - 1 : 0;
            err[i] = 0;
            for (std::size_t j = 1; j < 50; ++j) {
                const double xx = x[i] + j * (x[i + 1] - x[i]) / 50;
                const double t = (xx - x[i]) / (x[i + 1] - x[i]);
                const double ai = a[i], bi = b[i];
                const double f = (1 + ai + bi - ai * t);
                const double nu = f * (1 - std::sqrt(1 - 4 * bi * t * t / (f * f))) / (2 * bi * t);
                const double q = (1 + ai * nu + bi * nu * nu);
                const double approx = q * q * (e[i + 1] - e[i]) / ((1 + ai + bi) * (1 - bi * nu * nu) * (x[i + 1] - x[i]));
                err[i] += std::abs(approx - pdf(xx)) * (x[i + 1] - x[i]) / 50;
            }
        }
        std::size_t m = 0;
        for (std::size_t i = 1; i < n; ++i)
            if (err[m] < err[i])
                m = i;
        const double x0 = x[m];
        err[m] = -1;
        const double x1 = x[m + 1];
        x.insert(x.begin() + m + 1, x0 + (x1 - x0) / 2);
        e.insert(e.begin() + m + 1, 0.0);
        a.insert(a.begin() + m + 1, 0.0);
        b.insert(b.begin() + m + 1, 0.0);
        err.insert(err.begin() + m + 1, -1.0);
        ++n;
        integrate();
        scale();
    }
    for (int i = 0; i < N; ++i) {
        ox[i] = x[i];
        oe[i] = e[i];
        oa[i] = a[i];
        ob[i] = b[i];
    }
}

// RITA::operator()(state, max) (dxmcrandom.hpp:483-500); the index is kept
// inside the table when a draw equals the last CDF node
template <typename T>
T ritaSample(const T* x, const T* e, const T* a, const T* b, int N, Rng& rng, T maxValue)
{
    const std::size_t im = upperIndex(x, N, maxValue);
    const T cdfMax = im != static_cast<std::size_t>(N) ? e[im] : T { 1 };
    T res;
    unsigned trials = 0;
    do {
        const T r = rng.uni<T>(cdfMax);
        std::size_t i = upperIndex(e, N, r);
        i = i == 0 ? 0 : i - 1;
        i = std::min<std::size_t>(i, N - 2);
        const T v = r - e[i];
        const T d = e[i + 1] - e[i];
        res = x[i] + (1 + a[i] + b[i]) * d * v / (d * d + a[i] * d * v + b[i] * v * v) * (x[i + 1] - x[i]);
    } while (res > maxValue && ++trials < MAX_TRIALS);
    return res;
}

// --------------------------------------------------------------- spline
// CubicSplineInterpolator<T,16> (interpolation.hpp:123-176): 60 coefficients,
// start, step, stop
template <typename F>
void splineBuild(double start, double stop, F f, int N, double* out)
{
    std::vector<double> xs(N), y(N), h(N, 0.0), dl(N, 1.0), H(N, 0.0), d(N, 0.0);
    const double step = (stop - start) / (N - 2);
    for (int i = 0; i < N; ++i) {
        xs[i] = start + step * i;
        y[i] = f(xs[i]);
    }
    for (int i = 0; i + 1 < N; ++i)
        h[i] = xs[i + 1] - xs[i];
    for (int i = 0; i + 1 < N; ++i)
        dl[i] = (y[i + 1] - y[i]) / h[i];
    for (int i = 1; i < N; ++i)
        H[i] = 2 * (h[i - 1] + h[i]);
    H[0] = H[1];
    for (int i = 1; i < N; ++i)
        d[i] = 6 * (dl[i] - dl[i - 1]);
    d[N - 1] = 0;
    d[0] = 0;
    for (int i = 1; i < N; ++i) {
        const double w = h[i - 1] / H[i - 1];
        H[i] -= w * h[i - 1];
        d[i] -= w * d[i - 1];
    }
    std::vector<double> s(N);
    s[N - 1] = d[N - 1] / H[N - 1];
    for (int i = N - 2; i >= 0; --i)
        s[i] = (d[i] - h[i] * s[i + 1]) / H[i];
    for (int i = 0; i + 1 < N; ++i) {
        const double x0 = xs[i], x1 = xs[i + 1];
        double c0 = (s[i] * x1 * x1 * x1 - s[i + 1] * x0 * x0 * x0 + 6 * (y[i] * x1 - y[i + 1] * x0)) / (6 * h[i]);
        c0 += h[i] * (s[i + 1] * x0 - s[i] * x1) / 6;
        out[4 * i] = c0;
        out[4 * i + 1] = (s[i + 1] * x0 * x0 - s[i] * x1 * x1 + 2 * (y[i + 1] - y[i])) / (2 * h[i]) + h[i] * (s[i] - s[i + 1]) / 6;
        out[4 * i + 2] = (s[i] * x1 - s[i + 1] * x0) / (2 * h[i]);
        out[4 * i + 3] = (s[i + 1] - s[i]) / (6 * h[i]);
    }
    out[4 * (N - 1)] = start;
    out[4 * (N - 1) + 1] = step;
    out[4 * (N - 1) + 2] = xs[N - 1];
}

template <typename T>
T splineEval(const T* c, T q)
{
    const T start = c[60], step = c[61], stop = c[62];
    const T x = std::clamp(q, start, stop);
    const std::size_t k = x > start ? static_cast<std::size_t>((x - start) / step) : 0;
    const std::size_t o = k < 15 ? k * 4 : 14 * 4;
    return c[o] + c[o + 1] * x + c[o + 2] * x * x + c[o + 3] * x * x * x;
}

// ------------------------------------------------------------- tables
struct Shell {
    double binding, electrons, j0, photoProb, yield, lineProb[3], lineEnergy[3];
};

template <typename T>
struct TablesT {
    std::size_t nMat = 0, nodes = 0, linearIndex = 0;
    T linearStep = 0, linearEnergy = 0;
    std::vector<T> grid, coef, maj, rita, spline;
    std::vector<Shell> shells;
};

struct Tables {
    TablesT<double> d;
    TablesT<float> f;
    double maxEnergy = 0, minEnergy = 0;
};

// AttenuationLutInterpolator::generate + generateMaxAttenuationLut
// (attenuationinterpolator.hpp:48-84, 119-205) for the world tables
// (maxEnergy >= 50, minEnergy >= 0.1)
void buildLut(const std::vector<dxmc::Material>& mats, const std::vector<double>& maxDens, double maxE, double minE, TablesT<double>& t)
{
    minE = std::max(0.1, minE);
    maxE = std::max(maxE, 50.0);
    std::size_t res = std::max(static_cast<std::size_t>(maxE - minE), std::size_t { 10 });
    const double lmin = std::log10(minE), lmax = std::log10(maxE);
    std::vector<double> edges;
    for (const auto& m : mats)
        for (const double be : m.getBindingEnergies(minE)) {
            edges.push_back(std::log10(be - 0.001));
            edges.push_back(std::log10(be));
        }
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    std::vector<double> x;
    for (std::size_t i = 0; i <= res; ++i)
        x.push_back(lmin + (i * (lmax - lmin)) / (res));
    for (const double e : edges)
        x.push_back(e);
    std::sort(x.begin(), x.end());
    x.erase(std::unique(x.begin(), x.end()), x.end());
    const std::size_t nseg = x.size() - 1;
    t.nMat = mats.size();
    t.coef.assign(mats.size() * nseg * 6, 0.0);
    for (std::size_t m = 0; m < mats.size(); ++m)
        for (int k = 0; k < 3; ++k) {
            std::vector<double> y(x.size());
            for (std::size_t i = 0; i < x.size(); ++i) {
                const double E = std::pow(10.0, x[i]);
                const double mu = k == 0 ? mats[m].getPhotoelectricAttenuation(E) : (k == 1 ? mats[m].getComptonAttenuation(E) : mats[m].getRayleightAttenuation(E));
                y[i] = std::log10(mu);
            }
            for (std::size_t i = 0; i < nseg; ++i) {
                const double slope = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
                t.coef[m * nseg * 6 + k * 2 + i * 6] = y[i] - x[i] * slope;
                t.coef[m * nseg * 6 + k * 2 + i * 6 + 1] = slope;
            }
        }
    const double dropped = x.front();
    t.grid.assign(x.begin() + 1, x.end());
    if (!edges.empty()) {
        const std::size_t u = upperIndex(t.grid.data(), t.grid.size(), edges.back());
        if (u == t.grid.size()) {
            t.linearIndex = t.grid.size();
            t.linearStep = t.grid[t.grid.size() - 1] - t.grid[t.grid.size() - 2];
            t.linearEnergy = t.grid.back();
        } else {
            t.linearIndex = u;
            t.linearStep = t.grid[u + 1] - t.grid[u];
            t.linearEnergy = t.grid[u];
        }
    } else {
        t.linearIndex = 0;
        t.linearStep = t.grid[1] - t.grid[0];
        t.linearEnergy = t.grid.front();
    }
    t.nodes = t.grid.size();
    std::vector<double> xm = t.grid;
    xm.insert(xm.begin(), dropped);
    std::vector<double> ym(xm.size());
    for (std::size_t i = 0; i < xm.size(); ++i) {
        double top = 0;
        for (std::size_t m = 0; m < maxDens.size(); ++m) {
            const double E = std::pow(10.0, xm[i]);
            const double le = std::log10(E);
            std::size_t k;
            if (le > t.linearEnergy)
                k = std::min(static_cast<std::size_t>((le - t.linearEnergy) / t.linearStep) + t.linearIndex, t.nodes - 1);
            else {
                const std::size_t u = upperIndex(t.grid.data(), t.nodes, le);
                k = u != t.nodes ? u : t.nodes - 1;
            }
            const double* c = &t.coef[m * t.nodes * 6 + k * 6];
            const double v = maxDens[m] * ((std::pow(10.0, c[0] + c[1] * le) + std::pow(10.0, c[2] + c[3] * le)) + std::pow(10.0, c[4] + c[5] * le));
            top = std::max(top, v);
        }
        ym[i] = std::log10(1 / top);
    }
    t.maj.assign(t.nodes * 2, 0.0);
    for (std::size_t i = 0; i + 1 < xm.size(); ++i) {
        const double slope = (ym[i + 1] - ym[i]) / (xm[i + 1] - xm[i]);
        t.maj[2 * i] = ym[i] - xm[i] * slope;
        t.maj[2 * i + 1] = slope;
    }
}

Tables* buildTables(const std::vector<dxmc::Material>& mats, const std::vector<double>& maxDens, double maxEnergy, double minEnergy)
{
    auto* T = new Tables();
    // AttenuationLut::generate (attenuationlut.hpp:64-96)
    const double lo = std::max(0.5, std::min(maxEnergy, minEnergy));
    const double hi = std::min(2 * 510.9989461, std::max(maxEnergy, minEnergy));
    T->maxEnergy = hi;
    T->minEnergy = lo;
    auto& t = T->d;
    const std::size_t n = mats.size();
    t.rita.assign(n * 4 * 56, 0.0);
    t.spline.assign(n * 63, 0.0);
    const double qmax = hi * (1 / 12.398520);
    for (std::size_t m = 0; m < n; ++m) {
        // form factor extent (attenuationlut.hpp:216-233)
        double qa = 1;
        double ff = mats[m].getRayleightFormFactorSquared(qa);
        while (qa < qmax * qmax && ff > 0.001) {
            qa += ff > 0.5 ? 0.5 : 0.1;
            ff = mats[m].getRayleightFormFactorSquared(qa);
        }
        double* r = &t.rita[m * 4 * 56];
        ritaBuild(
            0.0, qa, [&](double q2) { return mats[m].getRayleightFormFactorSquared(std::sqrt(q2)); }, 56, r, r + 56, r + 112, r + 168);
        // scatter factor extent (attenuationlut.hpp:248-263)
        double qs = 0.5;
        double sf = mats[m].getComptonNormalizedScatterFactor(qs);
        while (sf < 0.999 && qs < qmax) {
            sf = mats[m].getComptonNormalizedScatterFactor(qs);
            qs += sf < 0.5 ? 0.5 : 0.1;
        }
        splineBuild(
            0.0, qs, [&](double q) { return mats[m].getComptonNormalizedScatterFactor(q); }, 16, &t.spline[m * 63]);
        const auto conf = mats[m].getElectronConfiguration();
        for (const auto& c : conf)
            t.shells.push_back(Shell { c.bindingEnergy, c.numberElectrons, c.hartreeFockOrbital_0, c.photoIonizationProbability, c.fluorescenceYield,
                { c.fluorLineProbabilities[0], c.fluorLineProbabilities[1], c.fluorLineProbabilities[2] },
                { c.fluorLineEnergies[0], c.fluorLineEnergies[1], c.fluorLineEnergies[2] } });
    }
    buildLut(mats, maxDens, hi, lo, t);
    // float copy in the reference's base-10 form
    auto& f = T->f;
    f.nMat = t.nMat;
    f.nodes = t.nodes;
    f.linearIndex = t.linearIndex;
    f.linearStep = static_cast<float>(t.linearStep);
    f.linearEnergy = static_cast<float>(t.linearEnergy);
    auto cast = [](const std::vector<double>& v) { return std::vector<float>(v.cbegin(), v.cend()); };
    f.grid = cast(t.grid);
    f.coef = cast(t.coef);
    f.maj = cast(t.maj);
    f.rita = cast(t.rita);
    f.spline = cast(t.spline);
    f.shells = t.shells;
    return T;
}

// ------------------------------------------------------------ lookups
// AttenuationLutInterpolator::operator() / maxAttenuationInverse
// (attenuationinterpolator.hpp:207-248)
template <typename T>
std::size_t segment(const TablesT<T>& t, T le)
{
    if (le > t.linearEnergy)
        return std::min(static_cast<std::size_t>((le - t.linearEnergy) / t.linearStep) + t.linearIndex, t.nodes - 1);
    const std::size_t u = upperIndex(t.grid.data(), t.nodes, le);
    return u != t.nodes ? u : t.nodes - 1;
}
template <typename T>
void attenuation(const TablesT<T>& t, std::size_t mat, T E, T out[3])
{
    const T le = std::log10(E);
    const T* c = &t.coef[mat * t.nodes * 6 + segment(t, le) * 6];
    for (int k = 0; k < 3; ++k)
        out[k] = std::pow(T { 10 }, c[2 * k] + c[2 * k + 1] * le);
}
template <typename T>
T majorantInv(const TablesT<T>& t, T E)
{
    const T le = std::log10(E);
    const std::size_t k = segment(t, le);
    return std::pow(T { 10 }, t.maj[2 * k] + t.maj[2 * k + 1] * le);
}

// ------------------------------------------------------------ geometry
template <typename T>
void rotateVec(T v[3], const T ax[3], T angle)
{
    const T s = std::sin(angle), c = std::cos(angle);
    const T m = (T { 1 } - c) * (v[0] * ax[0] + v[1] * ax[1] + v[2] * ax[2]);
    const T a = c * v[0] + m * ax[0] + s * (ax[1] * v[2] - ax[2] * v[1]);
    const T b = c * v[1] + m * ax[1] + s * (-ax[0] * v[2] + ax[2] * v[0]);
    const T d = c * v[2] + m * ax[2] + s * (ax[0] * v[1] - ax[1] * v[0]);
    v[0] = a;
    v[1] = b;
    v[2] = d;
}
// vectormath::peturb (vectormath.hpp:197-221)
template <typename T>
void deflect(T d[3], T theta, T phi)
{
    const T x = std::abs(d[0]), y = std::abs(d[1]), z = std::abs(d[2]);
    const int k = x <= y ? (x <= z ? 0 : 2) : (y <= z ? 1 : 2);
    T e[3] = { 0, 0, 0 };
    e[k] = 1;
    T o[3] = { d[1] * e[2] - d[2] * e[1], d[2] * e[0] - d[0] * e[2], d[0] * e[1] - d[1] * e[0] };
    rotateVec(o, d, phi);
    const T st = std::sin(theta), ct = std::cos(theta);
    for (int i = 0; i < 3; ++i)
        d[i] = d[i] * ct + o[i] * st;
}

template <typename T>
struct Photon {
    T pos[3], dir[3], energy, weight;
};

struct WorldView {
    std::uint64_t dims[3];
    double spacing[3], extent[6];
    const std::uint8_t* material;
    const double* density;
    const std::uint8_t* measurement;
};

template <typename T>
bool insideWorld(const WorldView& w, const T p[3])
{
    return (p[0] > T(w.extent[0]) && p[0] < T(w.extent[1])) && (p[1] > T(w.extent[2]) && p[1] < T(w.extent[3])) && (p[2] > T(w.extent[4]) && p[2] < T(w.extent[5]));
}

// transportParticleToWorld (transport.hpp:702-728)
template <typename T>
bool enterWorld(const WorldView& w, Photon<T>& p)
{
    if (insideWorld(w, p.pos))
        return true;
    T lo = std::numeric_limits<T>::lowest(), hi = std::numeric_limits<T>::max();
    for (int i = 0; i < 3; ++i) {
        if (std::abs(p.dir[i]) > T { 1.0e-9 }) {
            const T a0 = (T(w.extent[2 * i]) - p.pos[i]) / p.dir[i];
            const T a1 = (T(w.extent[2 * i + 1]) - p.pos[i]) / p.dir[i];
            lo = std::max(lo, std::min(a0, a1));
            hi = std::min(hi, std::max(a0, a1));
        }
    }
    if (lo < hi && lo > 0) {
        for (int i = 0; i < 3; ++i)
            p.pos[i] += lo * p.dir[i];
        return true;
    }
    return false;
}

// indexFromPosition (transport.hpp:496-521), kept inside the grid
template <typename T>
std::uint64_t voxel(const WorldView& w, const T p[3])
{
    std::uint64_t i[3];
    for (int k = 0; k < 3; ++k) {
        i[k] = static_cast<std::uint64_t>((p[k] - T(w.extent[2 * k])) / T(w.spacing[k]));
        i[k] = std::min<std::uint64_t>(i[k], w.dims[k] - 1);
    }
    return i[2] * w.dims[0] * w.dims[1] + i[1] * w.dims[0] + i[0];
}

// ------------------------------------------------------------- sources
struct SourceTables {
    std::vector<std::uint32_t> specOffset { 0 };
    std::vector<double> specProb, specEnergy;
    std::vector<std::uint32_t> specAlias;
    std::vector<dxmc_heel> heel;
    std::vector<double> heelWeight;
    std::vector<std::uint32_t> bowOffset { 0 };
    std::vector<double> bowAngle, bowWeight;
};

// SpecterDistribution::sampleValue (dxmcrandom.hpp:211-216, 322-326)
template <typename T>
T sampleSpectrum(const SourceTables& s, int id, Rng& rng)
{
    const std::uint32_t off = s.specOffset[id];
    const std::uint64_t n = s.specOffset[id + 1] - off;
    const T r = rng.uni<T>();
    const std::uint64_t k = rng.index(n);
    const std::uint64_t bin = r < T(s.specProb[off + k]) ? k : s.specAlias[off + k];
    return bin < n - 1 ? rng.uni<T>(T(s.specEnergy[off + bin]), T(s.specEnergy[off + bin + 1])) : T(s.specEnergy[off + bin]);
}

// BowTieFilter::sampleIntensityWeight (beamfilters.hpp:150-190)
template <typename T>
T bowtie(const SourceTables& s, int id, T signedAngle)
{
    const T a = std::abs(signedAngle);
    const std::uint32_t off = s.bowOffset[id];
    const std::size_t n = s.bowOffset[id + 1] - off;
    const double* xs = &s.bowAngle[off];
    const double* ys = &s.bowWeight[off];
    std::size_t first = 0, last = n - 1, it = (last - first) / 2;
    while (it != first) {
        if (a < T(xs[it])) {
            last = it;
            it = first;
        } else
            first = it;
        it += (last - first) / 2;
    }
    if (a < T(xs[first]))
        return T(ys[0]);
    if (a > T(xs[last]))
        return T(ys[n - 1]);
    return T(ys[first]) + (a - T(xs[first])) * (T(ys[last]) - T(ys[first])) / (T(xs[last]) - T(xs[first]));
}

// HeelFilter::sampleIntensityWeight (beamfilters.hpp:514-538)
template <typename T>
T heel(const SourceTables& s, int id, T angle, T energy)
{
    const dxmc_heel& h = s.heel[id];
    std::uint64_t ei = energy < T(h.energy_start) ? 0 : static_cast<std::uint64_t>((energy - T(h.energy_start) + T { 0.5 } * T(h.energy_step)) / T(h.energy_step));
    ei = std::min<std::uint64_t>(ei, h.energy_size - 1);
    std::uint64_t ai = angle < T(h.angle_start) ? 0 : static_cast<std::uint64_t>((angle - T(h.angle_start)) / T(h.angle_step));
    ai = std::min<std::uint64_t>(ai, h.angle_size - 1);
    const double* w = &s.heelWeight[h.weight_offset + ei * h.angle_size + ai];
    if (ai < h.angle_size - 1) {
        const T a0 = T(h.angle_start) + T(h.angle_step) * ai;
        const T a1 = T(h.angle_start) + T(h.angle_step) * (ai + 1);
        return T(w[0]) + (T(w[1]) - T(w[0])) * (angle - a0) / (a1 - a0);
    }
    return T(w[0]);
}

// Exposure::sampleParticle (exposure.hpp:280-304)
template <typename T>
Photon<T> emit(const dxmc_exposure& x, const SourceTables& s, Rng& rng)
{
    const T theta = rng.uni<T>(T(x.collimation_angles[0]), T(x.collimation_angles[1]));
    const T phi = rng.uni<T>(T(x.collimation_angles[2]), T(x.collimation_angles[3]));
    Photon<T> p;
    T cy[3], cx[3];
    for (int i = 0; i < 3; ++i) {
        p.pos[i] = T(x.position[i]);
        p.dir[i] = T(x.beam_direction[i]);
        cx[i] = T(x.direction_cosines[i]);
        cy[i] = T(x.direction_cosines[3 + i]);
    }
    p.weight = T(x.weight);
    rotateVec(p.dir, cy, theta);
    rotateVec(p.dir, cx, phi);
    p.energy = x.spectrum >= 0 ? sampleSpectrum<T>(s, x.spectrum, rng) : T(x.mono_energy);
    if (x.bowtie >= 0)
        p.weight *= bowtie<T>(s, x.bowtie, theta);
    if (x.heel >= 0)
        p.weight *= heel<T>(s, x.heel, phi, p.energy);
    return p;
}

// -------------------------------------------------------------- physics
constexpr double MEC2 = 510.9989461;
constexpr double KEV_A = 12.398520;

struct Tally {
    double* dose;
    std::uint32_t* events;
    double* variance;
    std::uint64_t* nEvents; // this thread's scoring event count
    // Transport::safeValueAdd (transport.hpp:208-214)
    void add(std::uint64_t i, double e)
    {
        std::atomic_ref<double>(dose[i]).fetch_add(e);
        std::atomic_ref<std::uint32_t>(events[i]).fetch_add(1u);
        std::atomic_ref<double>(variance[i]).fetch_add(e * e);
        ++*nEvents;
    }
};

// photoAbsorption (transport.hpp:216-263)
template <typename T, int L>
T photo(const TablesT<T>& t, Photon<T>& p, std::size_t mat, Rng& rng)
{
    const T E = p.energy;
    p.energy = 0;
    if (L < 2)
        return E;
    const Shell* sh = &t.shells[mat * 12];
    std::array<T, 12> cum;
    for (int i = 0; i < 12; ++i)
        cum[i] = E > T(sh[i].binding) ? T(sh[i].photoProb) : T { 0 };
    for (int i = 1; i < 12; ++i)
        cum[i] = cum[i - 1] + cum[i];
    std::size_t k = 0;
    const T u = cum[11] * rng.uni<T>();
    while (cum[k] < u && k < 11)
        ++k;
    if (T(sh[k].binding) > E || T(sh[k].binding) < T { 1 })
        return E;
    if (rng.uni<T>() <= T(sh[k].yield)) {
        std::size_t line = 0;
        T r = rng.uni<T>() - T(sh[k].lineProb[0]);
        while (r > 0 && line < 2) {
            ++line;
            r -= T(sh[k].lineProb[line]);
        }
        p.energy = T(sh[k].lineEnergy[line]);
        const T theta = rng.uni<T>(std::numbers::pi_v<T>);
        const T phi = rng.uni<T>(std::numbers::pi_v<T> + std::numbers::pi_v<T>);
        deflect(p.dir, theta, phi);
        return E - p.energy;
    }
    return E;
}

// rayleightScatter (transport.hpp:264-298)
template <typename T, int L>
void rayleigh(const TablesT<T>& t, Photon<T>& p, std::size_t mat, Rng& rng)
{
    if (L == 0) {
        const T top = (4 * std::numbers::sqrt2_v<T>) / (3 * std::numbers::sqrt3_v<T>);
        T theta;
        bool again = true;
        unsigned trials = 0;
        while (again && trials++ < MAX_TRIALS) {
            const T r = rng.uni<T>(T { 0 }, top);
            theta = rng.uni<T>(T { 0 }, std::numbers::pi_v<T>);
            const T s = std::sin(theta);
            again = r > ((2 - s * s) * s);
        }
        deflect(p.dir, theta, rng.uni<T>(std::numbers::pi_v<T> + std::numbers::pi_v<T>));
        return;
    }
    const T k = 1 / T(KEV_A);
    const T qm = p.energy * k;
    const T qm2 = qm * qm;
    const T* r = &t.rita[mat * 4 * 56];
    T mu;
    unsigned trials = 0;
    do {
        const T q2 = ritaSample(r, r + 56, r + 112, r + 168, 56, rng, qm2);
        const T inv = T(KEV_A) / p.energy;
        mu = 1 - 2 * q2 * inv * inv;
    } while ((1 + mu * mu) * T { 0.5 } < rng.uni<T>() && ++trials < MAX_TRIALS);
    const T theta = std::acos(mu);
    deflect(p.dir, theta, rng.uni<T>(std::numbers::pi_v<T> + std::numbers::pi_v<T>));
}

// comptonScatterNRC (transport.hpp:342-483)
template <typename T>
T comptonIA(const TablesT<T>& t, Photon<T>& p, std::size_t mat, Rng& rng)
{
    const Shell* sh = &t.shells[mat * 12];
    std::array<T, 12> cum;
    for (int i = 0; i < 12; ++i)
        cum[i] = (T(sh[i].binding) < p.energy && T(sh[i].j0) > 0) ? T(sh[i].electrons) : T { 0 };
    for (int i = 1; i < 12; ++i)
        cum[i] = cum[i - 1] + cum[i];
    int k = 0;
    const T u = cum[11] * rng.uni<T>();
    while (cum[k] < u && k < 11)
        ++k;
    const T U = T(sh[k].binding) / T(MEC2);
    const T pb = std::sqrt(2 * U + U * U);
    const T J0 = T(sh[k].j0);
    const T kk = p.energy / T(MEC2);
    if (U > kk)
        return 0;
    const T emin = 1 / (1 + 2 * kk);
    const T gmi = 1 / (1 / emin + emin);
    T e, mu;
    bool rej;
    unsigned trials = 0;
    do {
        const T r1 = rng.uni<T>();
        e = r1 + (1 - r1) * emin;
        const T tt = (1 - e) / (kk * e);
        const T s2 = tt * (2 - tt);
        mu = 1 - tt;
        const T g = (1 / e + e - s2) * gmi;
        rej = rng.uni<T>() > g;
        if (rej)
            continue;
        const T pi = (kk * (kk - U) * (1 - mu) - U) / std::sqrt(2 * kk * (kk - U) * (1 - mu) + U * U);
        const T kc = kk * e;
        const T qc = std::sqrt(kk * kk + kc * kc - 2 * kk * kc * mu);
        const T alpha = qc * (1 + kc * (kc - kk * mu) / (qc * qc)) / kk;
        const T bp = 1 + 2 * J0 * std::abs(pi);
        const T b = (bp * bp + 1) / 2;
        const T eb = std::exp(-b);
        T S;
        if (pi <= -pb)
            S = (1 - alpha * pb) * eb / 2;
        else if (pi < pb) {
            const T sp2 = 1 / (std::numbers::inv_sqrtpi_v<T> / std::numbers::sqrt2_v<T>);
            const T part1 = alpha * sp2 / (4 * J0);
            const T se = std::sqrt(std::numbers::e_v<T>);
            const T tp = 1 / (1 + T { 0.332673 } * (1 + 2 * J0 * std::abs(pb)));
            const T tq = 1 / (1 + T { 0.332673 } * (1 + 2 * J0 * std::abs(pi)));
            const T fp = se - eb * tp * (T { 0.34802 } + T { -0.0958798 } * tp + T { 0.7478556 } * tp * tp);
            const T fq = se - eb * tq * (T { 0.34802 } + T { -0.0958798 } * tq + T { 0.7478556 } * tq * tq);
            S = pi <= 0 ? (1 - alpha * pi) * eb / 2 - part1 * (fp - fq) : 1 - (1 - alpha * pi) * eb / 2 - part1 * (fp - fq);
        } else
            S = 1 - (1 - alpha * pb) * eb / 2;
        rej = rng.uni<T>() > S;
        if (rej)
            continue;
        const T Fmax = pi <= -pb ? 1 - alpha * pb : (pi >= pb ? 1 + alpha * pb : 1 + alpha * pi);
        const T r4 = rng.uni<T>();
        const T rb = 2 * r4 * eb;
        T pz;
        if (rb < 1)
            pz = (1 - std::sqrt(1 - 2 * std::log(rb))) / (2 * J0) / T(MEC2);
        else
            pz = (std::sqrt(1 - 2 * std::log(2 - rb)) - 1) / (2 * J0) / T(MEC2);
        const T Fpz = pz <= -pb ? 1 - alpha * pb : (pz >= pb ? 1 + alpha * pb : 1 + alpha * pz);
        rej = rng.uni<T>() > Fpz / Fmax;
        if (!rej) {
            const T part = std::sqrt(1 - 2 * e * mu + e * e * (1 - pz * pz * s2));
            const T kbar = kc / (1 - pz * pz * e * e) * (1 - pz * pz * e * mu + pz * part);
            e = kbar / kk;
        }
    } while (rej && ++trials < MAX_TRIALS);
    const T theta = std::acos(mu);
    deflect(p.dir, theta, rng.uni<T>(std::numbers::pi_v<T> + std::numbers::pi_v<T>));
    const T E = p.energy;
    p.energy *= e;
    return E - p.energy;
}

// comptonScatter (transport.hpp:300-340)
template <typename T, int L>
T compton(const TablesT<T>& t, Photon<T>& p, std::size_t mat, Rng& rng)
{
    if (L == 2)
        return comptonIA(t, p, mat, rng);
    const T E = p.energy;
    const T kk = E / T(MEC2);
    const T emin = 1 / (1 + 2 * kk);
    const T gmi = 1 / (1 / emin + emin);
    T e, mu;
    bool rej;
    unsigned trials = 0;
    do {
        const T r1 = rng.uni<T>();
        e = r1 + (1 - r1) * emin;
        const T tt = (1 - e) / (kk * e);
        const T s2 = tt * (2 - tt);
        mu = 1 - tt;
        const T g = (1 / e + e - s2) * gmi;
        const T r2 = rng.uni<T>();
        if (L == 1) {
            const T q = E * (1 / T(KEV_A)) * std::sqrt(T { 0.5 } - mu * T { 0.5 });
            rej = r2 > g * splineEval(&t.spline[mat * 63], q);
        } else
            rej = r2 > g;
    } while (rej && ++trials < MAX_TRIALS);
    const T theta = std::acos(mu);
    deflect(p.dir, theta, rng.uni<T>(std::numbers::pi_v<T> + std::numbers::pi_v<T>));
    p.energy *= e;
    return E - p.energy;
}

template <typename T>
bool scoreAndCutoff(Tally& tl, std::uint64_t idx, Photon<T>& p, T e, bool& refresh)
{
    if (p.energy < T { 1 }) {
        tl.add(idx, static_cast<double>((e + p.energy) * p.weight));
        p.energy = 0;
        return false;
    }
    tl.add(idx, static_cast<double>(e * p.weight));
    refresh = true;
    return true;
}

// computeInteractions (transport.hpp:583-638)
template <typename T, int L>
bool interactNaive(const TablesT<T>& t, const T a[3], Photon<T>& p, std::size_t mat, Tally& tl, std::uint64_t idx, Rng& rng, bool& refresh)
{
    const T tot = (a[0] + a[1]) + a[2];
    const T r = rng.uni<T>(tot);
    if (r < a[0])
        return scoreAndCutoff(tl, idx, p, photo<T, L>(t, p, mat, rng), refresh);
    if (r < (a[0] + a[1]))
        return scoreAndCutoff(tl, idx, p, compton<T, L>(t, p, mat, rng), refresh);
    rayleigh<T, L>(t, p, mat, rng);
    return true;
}

// computeInteractionsForced (transport.hpp:523-581)
template <typename T, int L>
bool interactForced(const TablesT<T>& t, T prob, const T a[3], Photon<T>& p, std::size_t mat, Tally& tl, std::uint64_t idx, Rng& rng, bool& refresh)
{
    const T tot = (a[0] + a[1]) + a[2];
    const T pp = a[0] / tot;
    const T wc = prob * pp;
    {
        Photon<T> v = p;
        const T ev = photo<T, L>(t, v, mat, rng);
        tl.add(idx, static_cast<double>(v.energy < T { 1 } ? (ev + v.energy) * v.weight * wc : ev * v.weight * wc));
    }
    if (rng.uni<T>() < prob * (1 - pp)) {
        const T r2 = rng.uni<T>(a[1] + a[2]);
        if (r2 < a[1]) {
            if (!scoreAndCutoff(tl, idx, p, compton<T, L>(t, p, mat, rng), refresh))
                return false;
        } else
            rayleigh<T, L>(t, p, mat, rng);
    }
    p.weight *= (1 - wc);
    return true;
}

// woodcockParticleTracking (transport.hpp:640-700)
template <typename T, int L>
void track(const WorldView& w, const TablesT<T>& t, Photon<T>& p, Rng& rng, Tally& tl, std::uint64_t& steps)
{
    T inv = 0;
    bool refresh = true;
    bool alive = true;
    for (unsigned n = 0; alive && n < MAX_STEPS; ++n) {
        if (refresh) {
            inv = majorantInv(t, p.energy);
            refresh = false;
        }
        const T r1 = rng.uni<T>();
        const T len = -std::log(r1) * inv * T { 10 };
        for (int i = 0; i < 3; ++i)
            p.pos[i] += p.dir[i] * len;
        if (!insideWorld(w, p.pos))
            break;
        ++steps;
        const std::uint64_t idx = voxel(w, p.pos);
        const std::size_t mat = w.material[idx];
        const T dens = T(w.density[idx]);
        const std::uint8_t meas = w.measurement ? w.measurement[idx] : 0;
        T a[3];
        attenuation(t, mat, p.energy, a);
        const T prob = ((a[0] + a[1]) + a[2]) * dens * inv;
        if (meas == 0) {
            if (rng.uni<T>() < prob)
                alive = interactNaive<T, L>(t, a, p, mat, tl, idx, rng, refresh);
        } else
            alive = interactForced<T, L>(t, prob, a, p, mat, tl, idx, rng, refresh);
        if (alive && p.energy * p.weight < T { 5 }) {
            if (rng.uni<T>() < T { 0.8 })
                alive = false;
            else
                p.weight *= T { 1 } / (T { 1 } - T { 0.8 });
        }
    }
}

template <typename T, int L>
void runHistories(const WorldView& w, const TablesT<T>& t, const SourceTables& s, const dxmc_exposure* x, const std::vector<std::uint64_t>& start,
    std::uint64_t seed, std::uint64_t first, std::uint64_t hb, std::uint64_t he, Tally tl, std::uint64_t& steps)
{
    std::size_t e = upperIndex(start.data(), start.size(), hb) - 1;
    for (std::uint64_t h = hb; h < he; ++h) {
        while (start[e + 1] <= h)
            ++e;
        Rng rng = historyRng(seed, first + h);
        Photon<T> p = emit<T>(x[e], s, rng);
        if (enterWorld(w, p))
            track<T, L>(w, t, p, rng, tl, steps);
    }
}

struct Context {
    std::unique_ptr<Tables> tables;
    SourceTables sources;
};
}

using namespace oracle;

extern "C" {

// ---------------------------------------------------------------- RNG
void oracle_rng_streams(std::uint64_t seed, std::uint64_t first, std::uint64_t n, std::uint32_t draws, std::uint32_t* out)
{
    for (std::uint64_t i = 0; i < n; ++i) {
        Rng r = historyRng(seed, first + i);
        for (std::uint32_t k = 0; k < draws; ++k)
            out[i * draws + k] = r.next();
    }
}
void oracle_pcg32(std::uint64_t s0, std::uint64_t s1, std::uint64_t n, std::uint32_t* out)
{
    Rng r { s0, s1 };
    for (std::uint64_t i = 0; i < n; ++i)
        out[i] = r.next();
}
void oracle_uniform(std::uint64_t s0, std::uint64_t s1, std::uint64_t n, int f32, double* out)
{
    Rng r { s0, s1 };
    for (std::uint64_t i = 0; i < n; ++i)
        out[i] = f32 ? static_cast<double>(r.uni<float>()) : r.uni<double>();
}
void oracle_random_index(std::uint64_t s0, std::uint64_t s1, std::uint64_t max, std::uint64_t n, std::uint64_t* out)
{
    Rng r { s0, s1 };
    for (std::uint64_t i = 0; i < n; ++i)
        out[i] = r.index(max);
}
std::uint64_t oracle_splitmix64(std::uint64_t x) { return mix64(x); }

// --------------------------------------------------------- samplers
void oracle_alias_table(const double* w, std::uint64_t n, double* prob, std::uint32_t* alias) { aliasTable(w, n, prob, alias); }

void oracle_rita_build(double lo, double hi, double (*pdf)(double), int N, double* x, double* e, double* a, double* b)
{
    ritaBuild(lo, hi, pdf, N, x, e, a, b);
}
void oracle_rita_sample(const double* x, const double* e, const double* a, const double* b, int N, std::uint64_t s0, std::uint64_t s1, double maxValue,
    std::uint64_t n, double* out)
{
    Rng r { s0, s1 };
    for (std::uint64_t i = 0; i < n; ++i)
        out[i] = ritaSample(x, e, a, b, N, r, maxValue);
}
void oracle_spline_build(double start, double stop, double (*f)(double), double* out63) { splineBuild(start, stop, f, 16, out63); }
double oracle_spline_eval(const double* c63, double q) { return splineEval(c63, q); }
void oracle_rotate(double* v, const double* axis, double angle) { rotateVec(v, axis, angle); }
void oracle_peturb(double* d, double theta, double phi) { deflect(d, theta, phi); }

// ------------------------------------------------------------ tables
void* oracle_create(std::uint32_t n, const char* const* names, const double* maxDens, double maxEnergy, double minEnergy)
{
    std::vector<dxmc::Material> mats;
    for (std::uint32_t i = 0; i < n; ++i) {
        dxmc::Material m(names[i]);
        if (!m.isValid())
            return nullptr;
        mats.push_back(m);
    }
    auto* c = new Context();
    c->tables.reset(buildTables(mats, std::vector<double>(maxDens, maxDens + n), maxEnergy, minEnergy));
    return c;
}
void oracle_destroy(void* h) { delete static_cast<Context*>(h); }

// table views in the layout of dxmc_tables_desc
void oracle_table_sizes(void* h, std::uint64_t* sizes /* nMat, nodes, linearIndex */, double* linear /* step, energy */)
{
    const auto& t = static_cast<Context*>(h)->tables->d;
    sizes[0] = t.nMat;
    sizes[1] = t.nodes;
    sizes[2] = t.linearIndex;
    linear[0] = t.linearStep;
    linear[1] = t.linearEnergy;
}
void oracle_table_arrays(void* h, double* grid, double* coef, double* maj, double* rita, double* spline, double* shells /* [nMat*12][11] */)
{
    const auto& t = static_cast<Context*>(h)->tables->d;
    std::copy(t.grid.cbegin(), t.grid.cend(), grid);
    std::copy(t.coef.cbegin(), t.coef.cend(), coef);
    std::copy(t.maj.cbegin(), t.maj.cend(), maj);
    std::copy(t.rita.cbegin(), t.rita.cend(), rita);
    std::copy(t.spline.cbegin(), t.spline.cend(), spline);
    for (std::size_t i = 0; i < t.shells.size(); ++i) {
        const Shell& s = t.shells[i];
        const double v[11] = { s.binding, s.electrons, s.j0, s.photoProb, s.yield, s.lineProb[0], s.lineProb[1], s.lineProb[2], s.lineEnergy[0],
            s.lineEnergy[1], s.lineEnergy[2] };
        std::copy(v, v + 11, shells + i * 11);
    }
}

// source tables: spectra from raw weights (alias built here), heel and bowtie as given
int oracle_add_spectrum(void* h, const double* weights, const double* energies, std::uint32_t n)
{
    auto& s = static_cast<Context*>(h)->sources;
    std::vector<double> prob(n);
    std::vector<std::uint32_t> alias(n);
    aliasTable(weights, n, prob.data(), alias.data());
    s.specProb.insert(s.specProb.end(), prob.cbegin(), prob.cend());
    s.specAlias.insert(s.specAlias.end(), alias.cbegin(), alias.cend());
    s.specEnergy.insert(s.specEnergy.end(), energies, energies + n);
    s.specOffset.push_back(static_cast<std::uint32_t>(s.specProb.size()));
    return static_cast<int>(s.specOffset.size() - 2);
}
int oracle_add_heel(void* h, const dxmc_heel* heel, const double* weights)
{
    auto& s = static_cast<Context*>(h)->sources;
    dxmc_heel d = *heel;
    d.weight_offset = static_cast<std::uint32_t>(s.heelWeight.size());
    s.heelWeight.insert(s.heelWeight.end(), weights, weights + std::size_t(heel->energy_size) * heel->angle_size);
    s.heel.push_back(d);
    return static_cast<int>(s.heel.size() - 1);
}
int oracle_add_bowtie(void* h, const double* angles, const double* weights, std::uint32_t n)
{
    auto& s = static_cast<Context*>(h)->sources;
    s.bowAngle.insert(s.bowAngle.end(), angles, angles + n);
    s.bowWeight.insert(s.bowWeight.end(), weights, weights + n);
    s.bowOffset.push_back(static_cast<std::uint32_t>(s.bowAngle.size()));
    return static_cast<int>(s.bowOffset.size() - 2);
}

// bowtie weight of table id at nq angles; heel weight of table id at n
// (angle, energy) pairs
void oracle_bowtie_weight(void* h, int id, std::uint64_t nq, const double* q, double* out)
{
    const auto& s = static_cast<Context*>(h)->sources;
    for (std::uint64_t i = 0; i < nq; ++i)
        out[i] = bowtie<double>(s, id, q[i]);
}
void oracle_heel_weight(void* h, int id, std::uint64_t n, const double* angle, const double* energy, double* out)
{
    const auto& s = static_cast<Context*>(h)->sources;
    for (std::uint64_t i = 0; i < n; ++i)
        out[i] = heel<double>(s, id, angle[i], energy[i]);
}

// ----------------------------------------------------------- lookups
void oracle_lut_lookup(void* h, int f32, std::uint64_t n, const std::uint32_t* mat, const double* E, double* att, double* maj)
{
    const auto& T = *static_cast<Context*>(h)->tables;
    for (std::uint64_t i = 0; i < n; ++i) {
        if (f32) {
            float a[3];
            attenuation(T.f, mat[i], static_cast<float>(E[i]), a);
            for (int k = 0; k < 3; ++k)
                att[i * 3 + k] = a[k];
            maj[i] = majorantInv(T.f, static_cast<float>(E[i]));
        } else {
            double a[3];
            attenuation(T.d, mat[i], E[i], a);
            for (int k = 0; k < 3; ++k)
                att[i * 3 + k] = a[k];
            maj[i] = majorantInv(T.d, E[i]);
        }
    }
}

}

// single interactions, semantics of dxmc_hip_interact
template <int L>
void interactL(const TablesT<double>& t, int kind, std::uint64_t n, const std::uint32_t* mat, double fp, std::uint64_t seed, std::uint64_t first,
    double* part, double* eout, int* aout)
{
    for (std::uint64_t i = 0; i < n; ++i) {
        Rng rng = historyRng(seed, first + i);
        Photon<double> p;
        double* q = part + 8 * i;
        std::copy(q, q + 3, p.pos);
        std::copy(q + 3, q + 6, p.dir);
        p.energy = q[6];
        p.weight = q[7];
        double dose = 0, var = 0;
        std::uint32_t ev = 0;
        std::uint64_t cnt = 0;
        Tally tl { &dose, &ev, &var, &cnt };
        bool refresh = false;
        int alive = 1;
        double e = 0;
        if (kind == 0)
            e = photo<double, L>(t, p, mat[i], rng);
        else if (kind == 1)
            e = compton<double, L>(t, p, mat[i], rng);
        else if (kind == 2)
            rayleigh<double, L>(t, p, mat[i], rng);
        else {
            double a[3];
            attenuation(t, mat[i], p.energy, a);
            alive = kind == 3 ? interactNaive<double, L>(t, a, p, mat[i], tl, 0, rng, refresh) : interactForced<double, L>(t, fp, a, p, mat[i], tl, 0, rng, refresh);
            e = dose;
        }
        std::copy(p.pos, p.pos + 3, q);
        std::copy(p.dir, p.dir + 3, q + 3);
        q[6] = p.energy;
        q[7] = p.weight;
        eout[i] = e;
        aout[i] = alive;
    }
}
extern "C" {
void oracle_interact(void* h, int L, int kind, std::uint64_t n, const std::uint32_t* mat, double fp, std::uint64_t seed, std::uint64_t first, double* part,
    double* eout, int* aout)
{
    const auto& t = static_cast<Context*>(h)->tables->d;
    if (L == 0)
        interactL<0>(t, kind, n, mat, fp, seed, first, part, eout, aout);
    else if (L == 1)
        interactL<1>(t, kind, n, mat, fp, seed, first, part, eout, aout);
    else
        interactL<2>(t, kind, n, mat, fp, seed, first, part, eout, aout);
}

// comptonScatterNRC (transport.hpp:342-483) on an explicit shell table, the
// harness of the impulse-approximation tests: shells3 [12][3] = binding
// energy (keV), electrons, J0; photon n starts at `energy` along +z with
// history stream first + n; out [n][5] = returned energy, E', direction.
void oracle_compton_ia(const double* shells3, double energy, std::uint64_t seed, std::uint64_t first, std::uint64_t n, double* out)
{
    TablesT<double> t;
    t.nMat = 1;
    t.shells.assign(12, Shell {});
    for (int i = 0; i < 12; ++i) {
        t.shells[i].binding = shells3[3 * i];
        t.shells[i].electrons = shells3[3 * i + 1];
        t.shells[i].j0 = shells3[3 * i + 2];
    }
    for (std::uint64_t i = 0; i < n; ++i) {
        Rng rng = historyRng(seed, first + i);
        Photon<double> p {};
        p.dir[2] = 1;
        p.energy = energy;
        p.weight = 1;
        double* q = out + 5 * i;
        q[0] = comptonIA(t, p, 0, rng);
        q[1] = p.energy;
        std::copy(p.dir, p.dir + 3, q + 2);
    }
}

// transportParticleToWorld + indexFromPosition for given particles [n][8]
void oracle_enter_world(const WorldView* w, std::uint64_t n, double* part, int* inside, std::uint64_t* index)
{
    for (std::uint64_t i = 0; i < n; ++i) {
        double* q = part + 8 * i;
        Photon<double> p;
        std::copy(q, q + 3, p.pos);
        std::copy(q + 3, q + 6, p.dir);
        p.energy = q[6];
        p.weight = q[7];
        inside[i] = enterWorld(*w, p) ? 1 : 0;
        index[i] = inside[i] ? voxel(*w, p.pos) : 0;
        std::copy(p.pos, p.pos + 3, q);
    }
}

void oracle_sample_particles(void* h, const WorldView* w, const dxmc_exposure* x, std::uint64_t seed, std::uint64_t first, std::uint64_t n, double* out,
    int* inside)
{
    const auto& s = static_cast<Context*>(h)->sources;
    for (std::uint64_t i = 0; i < n; ++i) {
        Rng rng = historyRng(seed, first + i);
        Photon<double> p = emit<double>(*x, s, rng);
        inside[i] = enterWorld(*w, p) ? 1 : 0;
        std::copy(p.pos, p.pos + 3, out + 8 * i);
        std::copy(p.dir, p.dir + 3, out + 8 * i + 3);
        out[8 * i + 6] = p.energy;
        out[8 * i + 7] = p.weight;
    }
}

// ---------------------------------------------------------- transport
// Transport::parallellRun (transport.hpp:745-778) with per-history streams:
// threads take blocks of histories from a shared counter; scoring uses
// atomic adds on the shared tallies (tallies are accumulated, not reset).
int oracle_transport(void* h, const WorldView* w, const dxmc_exposure* x, std::uint64_t nExp, int L, int f32, std::uint64_t seed, std::uint64_t first,
    std::uint32_t nThreads, double* dose, std::uint32_t* events, double* variance, std::uint64_t* counters /* histories, steps, events */)
{
    auto* c = static_cast<Context*>(h);
    std::vector<std::uint64_t> start(nExp + 1, 0);
    for (std::uint64_t i = 0; i < nExp; ++i)
        start[i + 1] = start[i] + x[i].n_histories;
    const std::uint64_t total = start[nExp];
    nThreads = std::max<std::uint32_t>(1, nThreads);
    std::atomic<std::uint64_t> next { 0 };
    std::vector<std::uint64_t> steps(nThreads, 0), evs(nThreads, 0);
    constexpr std::uint64_t block = 4096;
    auto worker = [&](std::uint32_t id) {
        Tally tl { dose, events, variance, &evs[id] };
        for (;;) {
            const std::uint64_t hb = next.fetch_add(block);
            if (hb >= total)
                break;
            const std::uint64_t he = std::min(total, hb + block);
            if (f32) {
                if (L == 0)
                    runHistories<float, 0>(*w, c->tables->f, c->sources, x, start, seed, first, hb, he, tl, steps[id]);
                else if (L == 1)
                    runHistories<float, 1>(*w, c->tables->f, c->sources, x, start, seed, first, hb, he, tl, steps[id]);
                else
                    runHistories<float, 2>(*w, c->tables->f, c->sources, x, start, seed, first, hb, he, tl, steps[id]);
            } else {
                if (L == 0)
                    runHistories<double, 0>(*w, c->tables->d, c->sources, x, start, seed, first, hb, he, tl, steps[id]);
                else if (L == 1)
                    runHistories<double, 1>(*w, c->tables->d, c->sources, x, start, seed, first, hb, he, tl, steps[id]);
                else
                    runHistories<double, 2>(*w, c->tables->d, c->sources, x, start, seed, first, hb, he, tl, steps[id]);
            }
        }
    };
    std::vector<std::thread> pool;
    for (std::uint32_t i = 1; i < nThreads; ++i)
        pool.emplace_back(worker, i);
    worker(0);
    for (auto& t : pool)
        t.join();
    if (counters) {
        counters[0] = total;
        counters[1] = 0;
        counters[2] = 0;
        for (std::uint32_t i = 0; i < nThreads; ++i) {
            counters[1] += steps[i];
            counters[2] += evs[i];
        }
    }
    return 0;
}

// normalizeScoring / energyImpartedToDose (transport.hpp:780-816)
void oracle_normalize(std::uint64_t n, double* dose, double* variance, int mode, std::uint64_t nHist, const double* density, const double* spacing,
    double calibration)
{
    if (mode == DXMC_OUTPUT_EV_PER_HISTORY) {
        const double hi = 1.0 / (nHist - 1);
        const double hd = 1e3 / nHist;
        const double hv = 1e6 / nHist;
        for (std::uint64_t i = 0; i < n; ++i) {
            dose[i] = dose[i] * hd;
            variance[i] = (variance[i] * hv - dose[i] * dose[i]) * hi;
        }
        return;
    }
    const double vol = spacing[0] * spacing[1] * spacing[2] / 1000;
    for (std::uint64_t i = 0; i < n; ++i) {
        const double mass = density[i] * vol * 0.001;
        const double f = calibration / mass;
        dose[i] = density[i] > 0 ? dose[i] * f : 0;
        variance[i] = density[i] > 0 ? variance[i] * f * f : 0;
    }
}
}

// ---------------------------------------------------------------- tube spectrum
// Restatement of the reference's tube spectrum as the reference evaluates it:
// betheHeitlerCrossSection.hpp:174-391 (nested x, u loops per photon energy,
// electron density re-interpolated at every node), tube.hpp:189-258 (K lines
// at the first bin >= the line within 2 keV, filtration, normalisation).
// The electron-penetration tables come in as arrays (from a fixture), so the
// product's own copy of them is checked too.
namespace {
struct AnodeTables {
    const double* u;
    std::int32_t nu;
    const double* depthF;
    const double* depthM;
    std::int32_t nx;
    const double* densityF; // [nu][nx]
    const double* densityM;
    const double* twVoltage;
    const double* twConstant;
    std::int32_t ntw;
    const double* kEnergy;
    const double* kFraction;
    std::int32_t nk;
};

// betheHeitlerCrossSection.hpp:182-206 and the segment rule of 240-252
std::int64_t refSegment(const double* g, std::int64_t n, double v)
{
    std::int64_t i = std::lower_bound(g, g + n, v) - g;
    const std::int64_t rest = n - i;
    if (rest < 2)
        i = n + (-3 + rest);
    return i;
}

double refTwLaw(const AnodeTables& t, double x, double T0)
{
    const std::int64_t i = refSegment(t.twVoltage, t.ntw, T0);
    const double C = t.twConstant[i] + ((t.twConstant[i + 1] - t.twConstant[i]) / (t.twVoltage[i + 1] - t.twVoltage[i])) * (T0 - t.twVoltage[i]);
    const double twl = (T0 * T0 - C * x) / (T0 * T0);
    return twl < 0 ? 0 : twl;
}
double refRange(double T0) { return 0.0119 * std::pow(T0, 1.513); } // :175-179
double refFracF(const AnodeTables& t, double x, double T0) { return std::pow(refTwLaw(t, x, T0), 1.753); } // :208-213
double refFracM(const AnodeTables& t, double x, double T0) // :214-224
{
    const double K = 18.0, Bd = 0.584, B0 = 0.5;
    const double ex = 1 - std::exp(-K * x / refRange(T0));
    const double F = Bd * ex;
    const double B = B0 + (Bd - B0) * ex;
    return refFracF(t, x, T0) * B * (F + 1) / (1 - B * F);
}
// :226-311 (bilinear, clamped, row u / column depth)
double refDensity(const AnodeTables& t, const double* depth, const double* table, double uval, double xval)
{
    const double x = std::clamp(xval, depth[0], depth[t.nx - 1]);
    const double u = std::clamp(uval, t.u[0], t.u[t.nu - 1]);
    const std::int64_t ix = refSegment(depth, t.nx, x), iu = refSegment(t.u, t.nu, u);
    const double q11 = table[iu * t.nx + ix], q21 = table[iu * t.nx + ix + 1];
    const double q22 = table[(iu + 1) * t.nx + ix + 1], q12 = table[(iu + 1) * t.nx + ix];
    const double x1 = depth[ix], x2 = depth[ix + 1], y1 = t.u[iu], y2 = t.u[iu + 1];
    const double xf1 = ((x2 - x) / (x2 - x1));
    const double xf2 = ((x - x1) / (x2 - x1));
    const double r1 = xf1 * q11 + xf2 * q21;
    const double r2 = xf1 * q12 + xf2 * q22;
    return ((y2 - u) / (y2 - y1)) * r1 + ((u - y1) / (y2 - y1)) * r2;
}
double refElectronDensity(const AnodeTables& t, double u, double x, double T0) // :313-318
{
    const double f = refRange(100.0) / refRange(T0);
    return refFracF(t, x, T0) * refDensity(t, t.depthF, t.densityF, u, x * f) + refFracM(t, x, T0) * refDensity(t, t.depthM, t.densityM, u, x * f);
}
double refBetheHeitler(double hv, double Ti) // :349-366
{
    const double m = 510.9989461;
    const double phiBar = (74 * 74) * 2.81794092E-15 * 2.81794092E-15 * 7.29735308E-03;
    const double k = (phiBar * 2) / 3;
    const double Ei = m + Ti;
    const double Ef = Ei - hv;
    const double pic_sqr = Ei * Ei - m * m;
    const double pic = std::sqrt(pic_sqr);
    const double pfc_sqr = Ef * Ef - m * m;
    if (pfc_sqr <= 0)
        return 0;
    const double pfc = std::sqrt(pfc_sqr);
    const double L = 2 * std::log((Ei * Ef + pic * pfc - m * m) / (m * hv));
    const double cc = pic / pfc;
    return k * (4 * Ei * Ef * L - 7 * pic * pfc) / (hv * pic * pic) * cc;
}
double refBremsstrahlung(const AnodeTables& t, double T0, double hv, double takeoff) // :368-392
{
    const double att = dxmc::Material::getTotalAttenuation(74, hv);
    double x = 0, I = 0;
    if (hv <= 0)
        return 0;
    while (x <= 14.0) {
        double step = 0;
        double u = 0.005;
        while (u <= 1.0) {
            step = step + refBetheHeitler(hv, T0 * u) * refElectronDensity(t, u, x, T0) * 0.005;
            u = u + 0.005;
        }
        I = I + step * std::exp(-att * x * 0.001 / std::sin(takeoff)) * 0.1;
        x = x + 0.1;
    }
    return I;
}
}

extern "C" {
// tube.hpp:189-208 with the filters applied in order; out[n]
void oracle_tube_spectrum(const AnodeTables* t, double T0, double takeoff, const double* energies, std::uint64_t n, const std::int32_t* filterZ,
    const double* filterMm, std::int32_t nFilters, std::int32_t normalize, double* out)
{
    for (std::uint64_t i = 0; i < n; ++i)
        out[i] = refBremsstrahlung(*t, T0, energies[i], takeoff);
    // tube.hpp:216-236, betheHeitlerCrossSection.hpp:394-407
    for (std::int32_t k = 0; k < t->nk; ++k) {
        const double e = t->kEnergy[k];
        const double intensity = (1 + 0.33) * t->kFraction[k] * 4.4 * 0.94 * refBremsstrahlung(*t, T0, e, takeoff);
        const double* it = std::lower_bound(energies, energies + n, e);
        if (it != energies + n && std::abs(e - *it) <= 2.0)
            out[it - energies] = out[it - energies] + intensity;
    }
    for (std::int32_t f = 0; f < nFilters; ++f) { // :237-244
        const dxmc::Material mat(filterZ[f]);
        const double cm = filterMm[f] * 0.1;
        for (std::uint64_t i = 0; i < n; ++i)
            out[i] = out[i] * std::exp(-mat.getTotalAttenuation(energies[i]) * mat.standardDensity() * cm);
    }
    if (normalize) { // :245-249, summed in order
        double sum = 0;
        for (std::uint64_t i = 0; i < n; ++i)
            sum = sum + out[i];
        for (std::uint64_t i = 0; i < n; ++i)
            out[i] = out[i] / sum;
    }
}
}
