// Built-in photon interaction model behind dxmc::Material.
//
// Replaces the xraylib calls of the reference (src/material.cpp:31-377) with a
// self-contained model. What is modelled, per element Z:
//  * subshell occupation by Madelung filling in xraylib's shell numbering
//    (K=0, L1..L3=1..3, M1..M5=4..8, N1..N7=9..15, O1..O7=16..22, P1..P5=23..27, Q1=28);
//  * binding energies: tabulated K edges, L/M edges from smooth fits of
//    their ratio to the K / L3 edge, Slater-screened hydrogenic values for the
//    remaining shells (never above 95 % of any inner shell's energy);
//  * form factor F(q) = sum_s n_s f_s(q), f_s = (1 + (2 pi q a_s)^2)^-2 with
//    the hydrogenic radius a_s = a0 / sqrt(E_s / Ry) (q = sin(theta/2)/lambda);
//  * incoherent scatter function S(q) = Z - sum_s n_s f_s(q)^2 (Waller-Hartree
//    without exchange);
//  * coherent / incoherent cross sections by Gauss-Legendre integration of
//    Thomson*F^2 and Klein-Nishina*S over momentum transfer;
//  * photoelectric cross section C Z^4.2 E^-3 shared among shells with
//    empirical jump ratios (K: 125/Z + 3.5), so the total drops by the jump
//    ratio below each edge and partial (shell) cross sections exist;
//  * fluorescence yields Z^4/(Z^4+a), three lines per K/L shell with energies
//    from the edge differences.
#include "dxmc/material.hpp"
#include "elementdata.hpp"

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>

namespace dxmc {
namespace {

    constexpr double RYDBERG_KEV = 0.0136057;
    constexpr double BOHR_ANGSTROM = 0.529177;
    constexpr double RE2_BARN = 0.0794078; // classical electron radius squared [barn]
    constexpr double AVOGADRO = 0.60221408; // 1e24/mol, so barn*N_A/A -> cm2/g
    constexpr double MEC2 = 510.9989461;
    constexpr double KEV_ANGSTROM = 12.398520;
    constexpr double PHOTO_CONSTANT = 21.8; // barn keV^3 / Z^4.2
    constexpr double PHOTO_Z_EXPONENT = 4.2;
    constexpr int N_SHELLS = 29;

    // capacity, principal quantum number, azimuthal quantum number per xraylib shell index
    constexpr int shellCapacity[N_SHELLS] = { 2, 2, 2, 4, 2, 2, 4, 4, 6, 2, 2, 4, 4, 6, 6, 8, 2, 2, 4, 4, 6, 6, 8, 2, 2, 4, 4, 6, 2 };
    constexpr int shellN[N_SHELLS] = { 1, 2, 2, 2, 3, 3, 3, 3, 3, 4, 4, 4, 4, 4, 4, 4, 5, 5, 5, 5, 5, 5, 5, 6, 6, 6, 6, 6, 7 };
    constexpr int shellL[N_SHELLS] = { 0, 0, 1, 1, 0, 1, 1, 2, 2, 0, 1, 1, 2, 2, 3, 3, 0, 1, 1, 2, 2, 3, 3, 0, 1, 1, 2, 2, 0 };
    // Madelung filling order expressed in shell indices
    constexpr int fillOrder[N_SHELLS] = { 0, 1, 2, 3, 4, 5, 6, 9, 7, 8, 10, 11, 16, 12, 13, 17, 18, 23, 14, 15, 19, 20, 24, 25, 28, 21, 22, 26, 27 };

    double piecewise(const double* x, const double* y, int n, double v)
    {
        if (v <= x[0])
            return y[0];
        if (v >= x[n - 1])
            return y[n - 1];
        int i = 1;
        while (x[i] < v)
            ++i;
        return y[i - 1] + (y[i] - y[i - 1]) * (v - x[i - 1]) / (x[i] - x[i - 1]);
    }

    struct ShellModel {
        int Z = 0;
        int nShells = 0; // contiguous shells starting at K that hold electrons
        std::array<double, N_SHELLS> occupancy {};
        std::array<double, N_SHELLS> binding {};
        std::array<double, N_SHELLS> radius {}; // Angstrom
        std::array<double, N_SHELLS> photoWeight {};
    };

    ShellModel buildShellModel(int Z)
    {
        ShellModel m;
        m.Z = Z;
        int left = Z;
        for (int k = 0; k < N_SHELLS && left > 0; ++k) {
            const int s = fillOrder[k];
            const int e = std::min(left, shellCapacity[s]);
            m.occupancy[s] = e;
            left -= e;
        }
        // xraylib enumerates shells until the first missing one
        m.nShells = 0;
        while (m.nShells < N_SHELLS && m.occupancy[m.nShells] > 0)
            ++m.nShells;

        // Slater screened hydrogenic estimate for every occupied shell
        for (int s = 0; s < N_SHELLS; ++s) {
            if (m.occupancy[s] <= 0)
                continue;
            const int n = shellN[s];
            const int l = shellL[s];
            double screen = 0;
            for (int t = 0; t < N_SHELLS; ++t) {
                if (m.occupancy[t] <= 0)
                    continue;
                const int nt = shellN[t];
                const int lt = shellL[t];
                const bool sameGroup = (nt == n) && ((l <= 1 && lt <= 1) || (lt == l));
                double e = m.occupancy[t];
                if (t == s)
                    e -= 1;
                if (sameGroup)
                    screen += e * (n == 1 ? 0.30 : 0.35);
                else if (l <= 1) {
                    if (nt == n - 1)
                        screen += e * 0.85;
                    else if (nt < n - 1)
                        screen += e;
                } else if (nt < n || (nt == n && lt < l))
                    screen += e;
            }
            const double zeff = std::max(Z - screen, 0.7);
            const double neff = n <= 3 ? n : (n == 4 ? 3.7 : (n == 5 ? 4.0 : 4.2));
            m.binding[s] = RYDBERG_KEV * zeff * zeff / (neff * neff);
        }
        m.binding[0] = elementdata::table[Z].kEdge;
        if (Z >= 11) {
            static const double zl[] = { 11, 13, 16, 20, 26, 29, 47, 53, 74, 82, 92 };
            static const double r3[] = { 0.0286, 0.0466, 0.0657, 0.0858, 0.0996, 0.1037, 0.131, 0.137, 0.1468, 0.148, 0.1485 };
            static const double r2[] = { 1.0, 1.0014, 1.005, 1.0104, 1.018, 1.021, 1.051, 1.065, 1.131, 1.166, 1.209 };
            static const double r1[] = { 2.1, 1.619, 1.40, 1.264, 1.195, 1.177, 1.143, 1.138, 1.185, 1.217, 1.263 };
            const double L3 = m.binding[0] * piecewise(zl, r3, 11, Z);
            m.binding[3] = L3;
            m.binding[2] = L3 * piecewise(zl, r2, 11, Z);
            m.binding[1] = L3 * piecewise(zl, r1, 11, Z);
            if (Z >= 30) {
                static const double zm[] = { 30, 47, 50, 74, 82, 92 };
                static const double rm[] = { 0.135, 0.2146, 0.225, 0.276, 0.295, 0.324 };
                const double M1 = L3 * piecewise(zm, rm, 6, Z);
                const double mf[5] = { 1.0, 0.92, 0.80, 0.67, 0.645 };
                for (int s = 4; s <= 8; ++s)
                    if (m.occupancy[s] > 0)
                        m.binding[s] = M1 * mf[s - 4];
            }
        }
        // outer shells never bind stronger than inner ones
        for (int s = 1; s < N_SHELLS; ++s) {
            if (m.occupancy[s] <= 0)
                continue;
            for (int t = 0; t < N_SHELLS; ++t)
                if (m.occupancy[t] > 0 && shellN[t] < shellN[s])
                    m.binding[s] = std::min(m.binding[s], 0.95 * m.binding[t]);
        }
        for (int s = 0; s < N_SHELLS; ++s)
            if (m.occupancy[s] > 0)
                m.radius[s] = BOHR_ANGSTROM / std::sqrt(m.binding[s] / RYDBERG_KEV);

        // photoelectric share of each shell: K gets 1 - 1/rK, then the L group
        // takes (1 - 1/rL) of the rest, M the same with rM; remaining weight goes
        // to the outermost shells in proportion to their electrons
        const double rK = Z > 1 ? 125.0 / Z + 3.5 : 1e30;
        const double rL = 3.0, rM = 3.0;
        double rest = 1.0;
        m.photoWeight.fill(0);
        if (Z == 1) {
            m.photoWeight[0] = 1;
            return m;
        }
        m.photoWeight[0] = rest * (1 - 1 / rK);
        rest /= rK;
        auto group = [&](int s0, int s1, double r, bool last) {
            double e = 0;
            for (int s = s0; s <= s1; ++s)
                e += m.occupancy[s];
            if (e <= 0)
                return;
            const double share = last ? rest : rest * (1 - 1 / r);
            for (int s = s0; s <= s1; ++s)
                m.photoWeight[s] += share * m.occupancy[s] / e;
            rest -= share;
        };
        const bool hasM = m.occupancy[4] > 0;
        const bool hasN = m.occupancy[9] > 0;
        group(1, 3, rL, !hasM);
        if (hasM)
            group(4, 8, rM, !hasN);
        if (hasN)
            group(9, N_SHELLS - 1, 1.0, true);
        return m;
    }

    const ShellModel& shellModel(int Z)
    {
        static std::array<ShellModel, elementdata::MAX_Z + 1> cache;
        static std::once_flag flag;
        std::call_once(flag, [] {
            for (int z = 1; z <= elementdata::MAX_Z; ++z)
                cache[z] = buildShellModel(z);
        });
        if (Z < 1 || Z > elementdata::MAX_Z)
            throw std::out_of_range("atomic number out of range");
        return cache[Z];
    }

    double modelFormFactor(int Z, double q)
    {
        const auto& m = shellModel(Z);
        double F = 0;
        for (int s = 0; s < N_SHELLS; ++s) {
            if (m.occupancy[s] <= 0)
                continue;
            const double x = 2 * 3.14159265358979323846 * q * m.radius[s];
            const double d = 1 + x * x;
            F += m.occupancy[s] / (d * d);
        }
        return F;
    }

    double modelScatterFunction(int Z, double q)
    {
        const auto& m = shellModel(Z);
        double S = Z;
        for (int s = 0; s < N_SHELLS; ++s) {
            if (m.occupancy[s] <= 0)
                continue;
            const double x = 2 * 3.14159265358979323846 * q * m.radius[s];
            const double d = 1 + x * x;
            const double f = 1 / (d * d);
            S -= m.occupancy[s] * f * f;
        }
        return std::max(S, 0.0);
    }

    // 24-point Gauss-Legendre nodes/weights on [-1, 1]
    constexpr double GL_X[12] = { 0.0640568928626056, 0.1911188674736163, 0.3150426796961634, 0.4337935076260451,
        0.5454214713888396, 0.6480936519369755, 0.7401241915785544, 0.8200019859739029,
        0.8864155270044011, 0.9382745520027328, 0.9747285559713095, 0.9951872199970214 };
    constexpr double GL_W[12] = { 0.1279381953467522, 0.1258374563468283, 0.1216704729278034, 0.1155056680537256,
        0.1074442701159656, 0.0976186521041139, 0.0861901615319533, 0.0733464814110803,
        0.0592985849154368, 0.0442774388174198, 0.0285313886289337, 0.0123412297999872 };

    // Momentum-transfer panels of the scattering integrals: [0, 1e-4], then
    // each panel four times longer than the one before, the last one cut at
    // t_max = (E / hc)^2. Only that last panel depends on the energy, so the
    // form factor and scatter function at the Gauss nodes of the full panels
    // are evaluated once per element and reused; every node and every sum is
    // formed by the same operations as with fresh evaluations, so the
    // integrals are bit-identical. 16 panels reach t = 1e-4 * 4^16 ~ 4e5,
    // beyond (2 m_e c^2 / hc)^2 ~ 6.8e3.
    constexpr int FIXED_PANELS = 16;
    struct PanelNodes {
        double value[FIXED_PANELS][24];
    };

    template <typename V>
    const PanelNodes& panelNodes(int kind, int Z, V value)
    {
        static std::map<std::pair<int, int>, std::unique_ptr<PanelNodes>> cache;
        static std::mutex mutex;
        std::lock_guard lock(mutex);
        auto& nodes = cache[{ kind, Z }];
        if (!nodes) {
            nodes = std::make_unique<PanelNodes>();
            double a = 0, b = 1e-4;
            for (int j = 0; j < FIXED_PANELS; ++j) {
                const double c = 0.5 * (a + b), h = 0.5 * (b - a);
                for (int i = 0; i < 12; ++i) {
                    nodes->value[j][2 * i] = value(c - h * GL_X[i]);
                    nodes->value[j][2 * i + 1] = value(c + h * GL_X[i]);
                }
                a = b;
                b *= 4;
            }
        }
        return *nodes;
    }

    // integrate g(t, v(t)) over t = q^2 in [0, tmax] with 24-point
    // Gauss-Legendre on each panel; v comes from the panel cache where the
    // panel is whole
    template <typename V, typename G>
    double integrateMomentum(double tmax, const PanelNodes& nodes, V value, G g)
    {
        double sum = 0;
        double a = 0, b = 1e-4;
        for (int j = 0; a < tmax; ++j) {
            const bool whole = b <= tmax && j < FIXED_PANELS;
            b = std::min(b, tmax);
            const double c = 0.5 * (a + b), h = 0.5 * (b - a);
            double s = 0;
            for (int i = 0; i < 12; ++i) {
                const double t0 = c - h * GL_X[i], t1 = c + h * GL_X[i];
                const double v0 = whole ? nodes.value[j][2 * i] : value(t0);
                const double v1 = whole ? nodes.value[j][2 * i + 1] : value(t1);
                s += GL_W[i] * (g(t0, v0) + g(t1, v1));
            }
            sum += s * h;
            a = b;
            b *= 4;
        }
        return sum;
    }

    double modelCoherent(int Z, double E)
    {
        const double k = E / KEV_ANGSTROM;
        const double k2 = k * k;
        const auto formFactor = [Z](double t) { return modelFormFactor(Z, std::sqrt(t)); };
        const double integral = integrateMomentum(k2, panelNodes(0, Z, formFactor), formFactor, [&](double t, double F) {
            const double mu = 1 - 2 * t / k2;
            return (1 + mu * mu) * F * F;
        });
        return 3.14159265358979323846 * RE2_BARN * 2 / k2 * integral;
    }

    double kleinNishinaFactor(double mu, double kappa)
    {
        const double r = 1 / (1 + kappa * (1 - mu));
        return r * r * (r + 1 / r - 1 + mu * mu);
    }

    double modelIncoherent(int Z, double E, bool energyTransfer)
    {
        const double k = E / KEV_ANGSTROM;
        const double k2 = k * k;
        const double kappa = E / MEC2;
        const auto scatterFunction = [Z](double t) { return modelScatterFunction(Z, std::sqrt(t)); };
        const double integral = integrateMomentum(k2, panelNodes(1, Z, scatterFunction), scatterFunction, [&](double t, double S) {
            const double mu = 1 - 2 * t / k2;
            const double kn = kleinNishinaFactor(mu, kappa);
            if (energyTransfer) {
                const double r = 1 / (1 + kappa * (1 - mu));
                return kn * S * (1 - r);
            }
            return kn * S;
        });
        return 3.14159265358979323846 * RE2_BARN * 2 / k2 * integral;
    }

    double modelPhotoPartial(int Z, int shell, double E)
    {
        const auto& m = shellModel(Z);
        // a shell contributes from its edge energy on (the LUT places a node at
        // the edge and one 1 eV below it, so the jump lies between the two)
        if (shell < 0 || shell >= N_SHELLS || m.occupancy[shell] <= 0 || E < m.binding[shell])
            return 0;
        return m.photoWeight[shell] * PHOTO_CONSTANT * std::pow(static_cast<double>(Z), PHOTO_Z_EXPONENT) / (E * E * E);
    }

    double modelPhoto(int Z, double E)
    {
        double s = 0;
        for (int sh = 0; sh < N_SHELLS; ++sh)
            s += modelPhotoPartial(Z, sh, E);
        return s;
    }

    double fluorYieldK(int Z)
    {
        const double z4 = std::pow(static_cast<double>(Z), 4);
        return z4 / (z4 + 9.1e5);
    }
    double fluorYieldL(int Z)
    {
        const double z4 = std::pow(static_cast<double>(Z), 4);
        return z4 / (z4 + 8.5e7);
    }

    // three principal fluorescence lines (target shell of the transition and
    // relative rate) for the K and L subshells
    struct Line {
        int from;
        double rate;
    };
    constexpr Line kLines[3] = { { 3, 0.58 }, { 2, 0.30 }, { 6, 0.12 } };
    constexpr Line l1Lines[3] = { { 6, 0.50 }, { 5, 0.35 }, { 11, 0.15 } };
    constexpr Line l2Lines[3] = { { 7, 0.75 }, { 12, 0.15 }, { 4, 0.10 } };
    constexpr Line l3Lines[3] = { { 8, 0.75 }, { 7, 0.09 }, { 13, 0.16 } };

    int modelShell(int Z, int shell, double* binding, double* electrons, double* J0, double* yield, double lineProb[3], double lineEnergy[3])
    {
        const auto& m = shellModel(Z);
        if (shell < 0 || shell >= m.nShells)
            return 1;
        *binding = m.binding[shell];
        *electrons = m.occupancy[shell];
        // hydrogenic Compton profile at pz=0: 8/(3 pi Zeff/n) per electron (atomic units)
        *J0 = m.occupancy[shell] * 8.0 / (3.0 * 3.14159265358979323846 * std::sqrt(m.binding[shell] / RYDBERG_KEV));
        *yield = 0;
        const Line* lines = nullptr;
        if (shell == 0) {
            *yield = fluorYieldK(Z);
            lines = kLines;
        } else if (shell <= 3) {
            *yield = fluorYieldL(Z);
            lines = shell == 1 ? l1Lines : (shell == 2 ? l2Lines : l3Lines);
        }
        double sum = 0;
        for (int i = 0; i < 3; ++i) {
            lineProb[i] = 0;
            lineEnergy[i] = 0;
            if (!lines)
                continue;
            const int from = lines[i].from;
            if (m.occupancy[from] > 0 && m.binding[from] < m.binding[shell]) {
                lineProb[i] = lines[i].rate;
                lineEnergy[i] = m.binding[shell] - m.binding[from];
                sum += lineProb[i];
            }
        }
        if (sum > 0)
            for (int i = 0; i < 3; ++i)
                lineProb[i] /= sum;
        else
            *yield = 0;
        return 0;
    }

    double modelEnergyAbsorption(int Z, double E)
    {
        // photoelectric energy minus the mean fluorescence energy escaping
        const auto& m = shellModel(Z);
        double photo = 0;
        for (int sh = 0; sh < m.nShells; ++sh) {
            const double p = modelPhotoPartial(Z, sh, E);
            if (p <= 0)
                continue;
            double b, n, j, w, lp[3], le[3];
            modelShell(Z, sh, &b, &n, &j, &w, lp, le);
            double meanLine = 0;
            for (int i = 0; i < 3; ++i)
                meanLine += lp[i] * le[i];
            photo += p * (1 - w * meanLine / E);
        }
        return photo + modelIncoherent(Z, E, true);
    }

    Material::ElementData& provider()
    {
        static Material::ElementData data {
            modelPhoto,
            modelPhotoPartial,
            [](int Z, double E) { return modelIncoherent(Z, E, false); },
            modelCoherent,
            modelEnergyAbsorption,
            modelFormFactor,
            modelScatterFunction,
            modelShell
        };
        return data;
    }

    // per element cross sections are cached per (Z, E) because table builders
    // ask for the same energies for every material
    double cachedCrossSection(int type, int Z, double E)
    {
        static std::map<std::tuple<int, int, double>, double> cache;
        static std::mutex mutex;
        const auto key = std::make_tuple(type, Z, E);
        {
            std::lock_guard lock(mutex);
            auto it = cache.find(key);
            if (it != cache.end())
                return it->second;
        }
        const auto& p = provider();
        double v = 0;
        if (type == 0)
            v = p.photo(Z, E);
        else if (type == 1)
            v = p.incoherent(Z, E);
        else if (type == 2)
            v = p.coherent(Z, E);
        else
            v = p.energyAbsorption(Z, E);
        std::lock_guard lock(mutex);
        cache.emplace(key, v);
        return v;
    }

    // parse a chemical formula into (Z, atom count)
    bool parseFormula(const std::string& s, std::size_t& pos, std::map<int, double>& out)
    {
        while (pos < s.size()) {
            const char c = s[pos];
            if (c == '(') {
                ++pos;
                std::map<int, double> inner;
                if (!parseFormula(s, pos, inner))
                    return false;
                if (pos >= s.size() || s[pos] != ')')
                    return false;
                ++pos;
                double mult = 0;
                std::size_t start = pos;
                while (pos < s.size() && (std::isdigit(static_cast<unsigned char>(s[pos])) || s[pos] == '.'))
                    ++pos;
                mult = pos > start ? std::stod(s.substr(start, pos - start)) : 1.0;
                for (const auto& [z, n] : inner)
                    out[z] += n * mult;
            } else if (c == ')') {
                return true;
            } else if (std::isupper(static_cast<unsigned char>(c))) {
                std::string sym(1, c);
                ++pos;
                while (pos < s.size() && std::islower(static_cast<unsigned char>(s[pos])))
                    sym.push_back(s[pos++]);
                const int Z = Material::getAtomicNumberFromSymbol(sym);
                if (Z <= 0)
                    return false;
                std::size_t start = pos;
                while (pos < s.size() && (std::isdigit(static_cast<unsigned char>(s[pos])) || s[pos] == '.'))
                    ++pos;
                const double n = pos > start ? std::stod(s.substr(start, pos - start)) : 1.0;
                out[Z] += n;
            } else {
                return false;
            }
        }
        return true;
    }
}

void Material::setDataProvider(const ElementData& data) { provider() = data; }
const Material::ElementData& Material::dataProvider() { return provider(); }

double Material::getAtomicWeight(int Z)
{
    if (Z < 1 || Z > elementdata::MAX_Z)
        return 0;
    return elementdata::table[Z].atomicWeight;
}
std::string Material::getAtomicNumberToSymbol(int Z) { return getSymbolFromAtomicNumber(Z); }
std::string Material::getSymbolFromAtomicNumber(int Z)
{
    if (Z < 1 || Z > elementdata::MAX_Z)
        return "";
    return elementdata::table[Z].symbol;
}
int Material::getAtomicNumberFromSymbol(const std::string& symbol)
{
    for (int Z = 1; Z <= elementdata::MAX_Z; ++Z)
        if (symbol == elementdata::table[Z].symbol)
            return Z;
    return 0;
}
std::vector<std::string> Material::getNISTCompoundNames(void)
{
    std::vector<std::string> names;
    for (const auto& m : elementdata::mixtures)
        names.emplace_back(m.name);
    return names;
}
double Material::getTotalAttenuation(int Z, double E)
{
    const double s = cachedCrossSection(0, Z, E) + cachedCrossSection(1, Z, E) + cachedCrossSection(2, Z, E);
    return s * AVOGADRO / getAtomicWeight(Z);
}

Material::Material(int atomicNumber) { setByAtomicNumber(atomicNumber); }

Material::Material(const std::string& name, const std::string& prettyName, const double density)
{
    setByMaterialName(name);
    if (!m_valid)
        setByCompoundName(name);
    if (!m_valid && !name.empty()) {
        const int Z = getAtomicNumberFromSymbol(name);
        if (Z > 0) {
            setByAtomicNumber(Z);
        }
    }
    m_prettyName = prettyName.empty() ? m_name : prettyName;
    if (density > 0)
        setStandardDensity(density);
}

Material::Material(const std::vector<int>& elements, const std::vector<double>& massFractions, const std::string& name, double density)
{
    if (elements.size() == massFractions.size() && !elements.empty()) {
        setComposition(elements, massFractions);
        m_name = name;
        m_prettyName = name;
        m_valid = true;
        setStandardDensity(density);
    }
}

const std::string& Material::prettyName(void) const { return m_prettyName.empty() ? m_name : m_prettyName; }

std::string Material::compositionKey(void) const
{
    std::string k;
    char buf[48];
    for (std::size_t i = 0; i < m_elements.size(); ++i) {
        std::uint64_t bits;
        std::memcpy(&bits, &m_elementNumberFraction[i], sizeof(bits));
        std::snprintf(buf, sizeof(buf), "%d:%016llx;", m_elements[i], static_cast<unsigned long long>(bits));
        k += buf;
    }
    return k;
}

void Material::setStandardDensity(double density)
{
    if (density > 0) {
        m_density = density;
        m_hasDensity = true;
    }
}

void Material::setComposition(const std::vector<int>& elements, const std::vector<double>& massFractions)
{
    m_elements = elements;
    m_elementNumberFraction.resize(elements.size());
    for (std::size_t i = 0; i < elements.size(); ++i)
        m_elementNumberFraction[i] = massFractions[i] / getAtomicWeight(elements[i]);
    const double sum = std::accumulate(m_elementNumberFraction.cbegin(), m_elementNumberFraction.cend(), 0.0);
    for (auto& f : m_elementNumberFraction)
        f /= sum;
}

void Material::setByAtomicNumber(int Z)
{
    if (Z < 1 || Z > elementdata::MAX_Z)
        return;
    m_name = elementdata::table[Z].symbol;
    m_density = elementdata::table[Z].density;
    m_hasDensity = true;
    m_valid = true;
    m_elements = { Z };
    m_elementNumberFraction = { 1.0 };
}

void Material::setByMaterialName(const std::string& name)
{
    for (const auto& mix : elementdata::mixtures) {
        if (name == mix.name) {
            std::vector<int> el(mix.Z, mix.Z + mix.n);
            std::vector<double> w(mix.w, mix.w + mix.n);
            setComposition(el, w);
            m_name = mix.name;
            m_density = mix.density;
            m_hasDensity = true;
            m_valid = true;
            return;
        }
    }
}

void Material::setByCompoundName(const std::string& name)
{
    if (name.empty() || !std::isupper(static_cast<unsigned char>(name[0])))
        return;
    std::map<int, double> atoms;
    std::size_t pos = 0;
    if (!parseFormula(name, pos, atoms) || pos != name.size() || atoms.empty())
        return;
    if (atoms.size() == 1 && atoms.begin()->second == 1.0) {
        setByAtomicNumber(atoms.begin()->first);
        return;
    }
    std::vector<int> el;
    std::vector<double> w;
    for (const auto& [z, n] : atoms) {
        el.push_back(z);
        w.push_back(n * getAtomicWeight(z));
    }
    const double sum = std::accumulate(w.cbegin(), w.cend(), 0.0);
    for (auto& x : w)
        x /= sum;
    setComposition(el, w);
    m_name = name;
    m_valid = true;
    m_hasDensity = false;
}

namespace {
    // mass attenuation of a mixture from per atom cross sections:
    // sum_i n_i sigma_i * N_A / sum_i n_i A_i
    double mixtureMass(const std::vector<int>& el, const std::vector<double>& nf, int type, double E)
    {
        double s = 0, a = 0;
        for (std::size_t i = 0; i < el.size(); ++i) {
            s += nf[i] * cachedCrossSection(type, el[i], E);
            a += nf[i] * Material::getAtomicWeight(el[i]);
        }
        return a > 0 ? s * AVOGADRO / a : 0;
    }
}

double Material::getPhotoelectricAttenuation(double E) const { return mixtureMass(m_elements, m_elementNumberFraction, 0, E); }
double Material::getComptonAttenuation(double E) const { return mixtureMass(m_elements, m_elementNumberFraction, 1, E); }
double Material::getRayleightAttenuation(double E) const { return mixtureMass(m_elements, m_elementNumberFraction, 2, E); }
double Material::getMassEnergyAbsorbtion(double E) const { return mixtureMass(m_elements, m_elementNumberFraction, 3, E); }
double Material::getTotalAttenuation(double E) const
{
    return getPhotoelectricAttenuation(E) + getComptonAttenuation(E) + getRayleightAttenuation(E);
}

double Material::getRayleightFormFactorSquared(const double q) const
{
    double s = 0;
    for (std::size_t i = 0; i < m_elements.size(); ++i) {
        const double f = provider().formFactor(m_elements[i], q);
        s += m_elementNumberFraction[i] * f * f;
    }
    return s;
}

double Material::getComptonNormalizedScatterFactor(const double q) const
{
    double s = 0;
    for (std::size_t i = 0; i < m_elements.size(); ++i)
        s += m_elementNumberFraction[i] * provider().scatterFunction(m_elements[i], q) / m_elements[i];
    return s;
}

std::vector<double> Material::getBindingEnergies(const double minValue) const
{
    std::vector<double> edges;
    for (const int Z : m_elements) {
        for (int sh = 0;; ++sh) {
            double b, n, j, w, lp[3], le[3];
            if (provider().shell(Z, sh, &b, &n, &j, &w, lp, le) != 0)
                break;
            if (b > minValue)
                edges.push_back(b);
        }
    }
    std::sort(edges.begin(), edges.end(), std::greater<double>());
    return edges;
}

// Grouped shell configuration, reference material.cpp:143-273: every shell of
// every element weighted by its number fraction, sorted by binding energy,
// truncated to 12, electron counts normalised to one, and the photoionisation
// probability taken as electrons * sum of the shell's partial photoelectric
// cross section at 0.5, 1.5, ..., 399.5 keV, normalised over the 12 entries.
std::array<ElectronShellConfiguration<double>, 12> Material::getElectronConfiguration() const
{
    std::vector<ElectronShellConfiguration<double>> all;
    for (std::size_t i = 0; i < m_elements.size(); ++i) {
        const int Z = m_elements[i];
        for (int sh = 0;; ++sh) {
            double b, n, j, w, lp[3], le[3];
            if (provider().shell(Z, sh, &b, &n, &j, &w, lp, le) != 0)
                break;
            ElectronShellConfiguration<double> c;
            c.bindingEnergy = b;
            c.numberElectrons = m_elementNumberFraction[i] * n;
            c.hartreeFockOrbital_0 = j;
            c.photoIonizationProbability = 0.0;
            c.fluorescenceYield = w;
            c.fluorLineProbabilities = { lp[0], lp[1], lp[2] };
            c.fluorLineEnergies = { le[0], le[1], le[2] };
            c.Z = Z;
            c.shell = sh;
            all.push_back(c);
        }
    }
    std::stable_sort(all.begin(), all.end(), [](const auto& a, const auto& b) { return a.bindingEnergy > b.bindingEnergy; });
    std::array<ElectronShellConfiguration<double>, 12> conf;
    for (std::size_t i = 0; i < std::min(all.size(), conf.size()); ++i)
        conf[i] = all[i];
    double electrons = 0;
    for (const auto& c : conf)
        electrons += c.numberElectrons;
    for (auto& c : conf)
        c.numberElectrons /= electrons;
    for (auto& c : conf) {
        if (c.Z > 0) {
            double s = 0;
            for (int k = 0; k < 400; ++k)
                s += provider().photoPartial(c.Z, c.shell, 0.5 + k);
            c.photoIonizationProbability = c.numberElectrons * s;
        }
    }
    double total = 0;
    for (const auto& c : conf)
        total += c.photoIonizationProbability;
    if (total > 0)
        for (auto& c : conf)
            c.photoIonizationProbability /= total;
    return conf;
}
}
