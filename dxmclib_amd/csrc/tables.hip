// Host side of the table upload: validation of dxmc_tables_desc and the
// construction of the FP64 and FP32 device copies.
//
// FP32 copy: the attenuation LUT is re-based from log10 to log2 and anchored
// at each segment's upper grid node, so the kernels evaluate
// mu = exp2(y_k + a (log2 E - x_k)) with y_k = log2 mu(x_k) (no cancellation
// for steep segments). The segment search runs on the FP64 grid, corrected
// by per-segment energy bounds that a small kernel derives from the FP64
// copy, so both paths pick the same segment. All other tables are rounded
// to float. The conversion is done in double on the host.
#include "internal.hpp"

#include <cmath>
#include <map>
#include <mutex>

namespace dxmc_backend {

namespace {
    // errors are kept per calling thread: a message read by one thread is
    // never rewritten by another thread's call
    thread_local std::string gErr;
    thread_local std::map<const dxmc_ctx*, std::string> gCtxErr;
    constexpr double LOG2_10 = 3.321928094887362347870319429489390175864831393;
}

void setGlobalError(const std::string& s) { gErr = s; }
const char* globalError() { return gErr.c_str(); }
std::map<const dxmc_ctx*, std::string>& threadErrors() { return gCtxErr; }

// page-locked blocks from dxmc_hip_host_alloc: start -> size
namespace {
    std::mutex gPinnedMutex;
    std::map<uintptr_t, size_t> gPinned;
}
void registerPinned(const void* p, size_t bytes)
{
    std::lock_guard lock(gPinnedMutex);
    gPinned[reinterpret_cast<uintptr_t>(p)] = bytes;
}
bool unregisterPinned(const void* p)
{
    std::lock_guard lock(gPinnedMutex);
    return gPinned.erase(reinterpret_cast<uintptr_t>(p)) != 0;
}
bool pinnedHost(const void* p, size_t bytes)
{
    const auto a = reinterpret_cast<uintptr_t>(p);
    std::lock_guard lock(gPinnedMutex);
    auto it = gPinned.upper_bound(a);
    if (it == gPinned.begin())
        return false;
    --it;
    return a >= it->first && a + bytes <= it->first + it->second;
}

bool validateTables(const dxmc_tables_desc& d, std::string& why)
{
    if (d.n_materials == 0 || d.n_materials > 256)
        return why = "need 1..256 materials", false;
    if (d.lut_nodes < 2 || !d.lut_grid || !d.lut_coefficients || !d.lut_majorant)
        return why = "attenuation LUT missing", false;
    // LUT rows are addressed with 24-bit products and 32-bit offsets
    if (d.lut_nodes >= (1u << 20))
        return why = "attenuation LUT larger than 2^20 nodes", false;
    if (d.lut_linear_index > d.lut_nodes || !(d.lut_linear_step > 0))
        return why = "LUT linear index parameters invalid", false;
    for (uint32_t i = 1; i < d.lut_nodes; ++i)
        if (!(d.lut_grid[i - 1] < d.lut_grid[i]))
            return why = "LUT grid not increasing", false;
    if (!d.rita || !d.spline || !d.shells)
        return why = "RITA, spline or shell tables missing", false;
    for (uint32_t m = 0; m < d.n_materials; ++m) {
        const double* s = d.spline + m * 63;
        if (!(s[61] > 0))
            return why = "spline step must be positive", false;
    }
    if (d.n_spectra) {
        if (!d.spectrum_offset || !d.spectrum_probability || !d.spectrum_alias || !d.spectrum_energy)
            return why = "spectrum arrays missing", false;
        if (d.spectrum_offset[0] != 0)
            return why = "spectrum offsets must start at 0", false;
        for (uint32_t s = 0; s < d.n_spectra; ++s) {
            const uint32_t a = d.spectrum_offset[s], b = d.spectrum_offset[s + 1];
            if (b <= a)
                return why = "empty spectrum", false;
            for (uint32_t i = a; i < b; ++i)
                if (d.spectrum_alias[i] >= b - a)
                    return why = "spectrum alias index out of range", false;
        }
    }
    if (d.n_heel) {
        if (!d.heel || !d.heel_weight)
            return why = "heel arrays missing", false;
        for (uint32_t h = 0; h < d.n_heel; ++h)
            if (d.heel[h].energy_size < 1 || d.heel[h].angle_size < 1 || !(d.heel[h].energy_step > 0) || !(d.heel[h].angle_step > 0))
                return why = "heel grid invalid", false;
    }
    if (d.n_bowtie) {
        if (!d.bowtie_offset || !d.bowtie_angle || !d.bowtie_weight || d.bowtie_offset[0] != 0)
            return why = "bowtie arrays missing", false;
        for (uint32_t b = 0; b < d.n_bowtie; ++b)
            if (d.bowtie_offset[b + 1] <= d.bowtie_offset[b])
                return why = "empty bowtie filter", false;
    }
    return true;
}

namespace {
    // lowest FP64 energy that the FP64 path places in segment k or above
    // (bisection over the ordered bit patterns of positive doubles, with the
    // FP64 path's own log10 and segment search)
    __global__ void segmentBounds(dxmc_dev::Tables<double> t, double* bound)
    {
        const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
        if (k >= t.lutNodes)
            return;
        if (k == 0) {
            bound[0] = 0.0;
            return;
        }
        auto segment = [&](uint64_t bits) { return dxmc_dev::lutInterval(t, log10(__longlong_as_double(static_cast<long long>(bits)))); };
        uint64_t lo = static_cast<uint64_t>(__double_as_longlong(1e-6));
        uint64_t hi = static_cast<uint64_t>(__double_as_longlong(1e6));
        if (segment(hi) < k) {
            bound[k] = HUGE_VAL;
            return;
        }
        if (segment(lo) >= k) {
            bound[k] = 1e-6;
            return;
        }
        while (hi - lo > 1) {
            const uint64_t mid = lo + (hi - lo) / 2;
            if (segment(mid) >= k)
                hi = mid;
            else
                lo = mid;
        }
        bound[k] = __longlong_as_double(static_cast<long long>(hi));
    }

    // Bounds of the FP32 total attenuation (photo + Compton + Rayleigh) of
    // each material over each LUT segment: the energies the FP64 segment
    // search sends to segment k are [bound[k], bound[k+1]); each component
    // 2^(c0 + c1 (log2 E - anchor_k)) is monotone there, so its extremes sit
    // at the range's ends (infinite ends give 0 or infinity). The sums are
    // widened by a relative margin of 2e-5 + 4e-6 sum|c1|, ten times the
    // error of the device's evaluation (hardware log2 / exp2 within a few
    // ulp, float coefficients), then rounded outward to float. A segment no
    // energy reaches gets an empty interval (lo > hi).
    std::vector<float2> totalAttenuationBounds(const std::vector<float>& coef, const std::vector<float>& anchor, const std::vector<double>& bound,
        uint32_t nMat, uint32_t nodes)
    {
        auto below = [](double v) {
            float f = static_cast<float>(v);
            return static_cast<double>(f) > v ? std::nextafter(f, -HUGE_VALF) : f;
        };
        auto above = [](double v) {
            float f = static_cast<float>(v);
            return static_cast<double>(f) < v ? std::nextafter(f, HUGE_VALF) : f;
        };
        auto component = [](double c0, double c1, double anchorK, double E) {
            if (E <= 0)
                return c1 > 0 ? 0.0 : (c1 < 0 ? HUGE_VAL : std::exp2(c0));
            if (std::isinf(E))
                return c1 < 0 ? 0.0 : (c1 > 0 ? HUGE_VAL : std::exp2(c0));
            return std::exp2(c0 + c1 * (std::log2(E) - anchorK));
        };
        std::vector<float2> out(size_t(nMat) * nodes);
        for (uint32_t m = 0; m < nMat; ++m)
            for (uint32_t k = 0; k < nodes; ++k) {
                const double eLo = k == 0 ? 0.0 : bound[k];
                const double eHi = k + 1 < nodes ? bound[k + 1] : HUGE_VAL;
                float2& r = out[size_t(m) * nodes + k];
                if (!(eLo < eHi)) {
                    r = make_float2(HUGE_VALF, 0.0f);
                    continue;
                }
                const float* c = &coef[(size_t(m) * nodes + k) * 6];
                double lo = 0, hi = 0, slope = 0;
                for (int i = 0; i < 3; ++i) {
                    const double a = component(c[2 * i], c[2 * i + 1], anchor[k], eLo);
                    const double b = component(c[2 * i], c[2 * i + 1], anchor[k], eHi);
                    lo += std::min(a, b);
                    hi += std::max(a, b);
                    slope += std::fabs(c[2 * i + 1]);
                }
                const double margin = 2e-5 + 4e-6 * slope;
                r = make_float2(below(lo * (1 - margin)), above(hi * (1 + margin)));
            }
        return out;
    }

    template <typename T, typename S>
    int upload(dxmc_ctx* c, dxmc_tables& t, const std::vector<S>& host, const T** dst)
    {
        T* p = nullptr;
        if (int st = c->alloc(&p, host.size() * sizeof(S)))
            return st;
        t.allocations.push_back(p);
        *dst = p;
        return c->copy(p, host.data(), host.size() * sizeof(S));
    }
    template <typename T>
    std::vector<T> cast(const double* a, size_t n, double scale = 1.0)
    {
        std::vector<T> r(n);
        for (size_t i = 0; i < n; ++i)
            r[i] = static_cast<T>(a[i] * scale);
        return r;
    }

    template <typename T>
    int build(dxmc_ctx* c, const dxmc_tables_desc& d, dxmc_tables& t, dxmc_dev::Tables<T>& out, bool base2)
    {
        const double ls = base2 ? LOG2_10 : 1.0;
        out.nMaterials = d.n_materials;
        out.lutNodes = d.lut_nodes;
        out.lutLinearIndex = d.lut_linear_index;
        out.lutLinearStep = d.lut_linear_step;
        out.lutLinearEnergy = d.lut_linear_energy;
        out.lutLinearEnergy2 = static_cast<float>(d.lut_linear_energy * LOG2_10);
        out.lutLinearStepInv2 = static_cast<float>(1.0 / (d.lut_linear_step * LOG2_10));
        out.nSpectra = d.n_spectra;
        int st;
        if ((st = upload(c, t, cast<double>(d.lut_grid, d.lut_nodes), &out.lutGrid)))
            return st;
        out.lutBound = nullptr;
        out.lutBoundUp = nullptr;
        out.lutAnchor = nullptr;
        if (base2) {
            // segment bounds from the FP64 copy (built first), float anchors
            double* bound = nullptr;
            if ((st = c->alloc(&bound, d.lut_nodes * sizeof(double))))
                return st;
            t.allocations.push_back(bound);
            hipLaunchKernelGGL(segmentBounds, dim3((d.lut_nodes + 63) / 64), dim3(64), 0, c->stream, t.f64, bound);
            if ((st = c->check(hipGetLastError())) || (st = c->check(hipStreamSynchronize(c->stream))))
                return st;
            out.lutBound = bound;
            if ((st = upload(c, t, cast<float>(d.lut_grid, d.lut_nodes, LOG2_10), &out.lutAnchor)))
                return st;
        }
        // FP64: (b, a) as built; FP32: (log2 mu at the upper node x_k, a)
        auto pair = [&](const double* ba, size_t node, T* out2) {
            if (base2) {
                out2[0] = static_cast<T>((ba[0] + ba[1] * d.lut_grid[node]) * ls);
                out2[1] = static_cast<T>(ba[1]);
            } else {
                out2[0] = static_cast<T>(ba[0]);
                out2[1] = static_cast<T>(ba[1]);
            }
        };
        const size_t nc = size_t(d.n_materials) * d.lut_nodes * 6;
        std::vector<T> coef(nc);
        for (size_t i = 0; i < nc; i += 2)
            pair(d.lut_coefficients + i, (i / 6) % d.lut_nodes, &coef[i]);
        if ((st = upload(c, t, coef, &out.lutCoef)))
            return st;
        out.lutSumBound = nullptr;
        if constexpr (std::is_same_v<T, float>) {
            std::vector<double> bounds(d.lut_nodes);
            if ((st = c->download(bounds.data(), out.lutBound, d.lut_nodes * sizeof(double))))
                return st;
            const auto anchors = cast<float>(d.lut_grid, d.lut_nodes, LOG2_10);
            const auto sums = totalAttenuationBounds(coef, anchors, bounds, d.n_materials, d.lut_nodes);
            if ((st = upload(c, t, sums, &out.lutSumBound)))
                return st;
            // segment bounds rounded up to float: a float energy compares
            // with them (>=, <) exactly as with the FP64 values
            std::vector<float> up(d.lut_nodes);
            for (size_t k = 0; k < up.size(); ++k) {
                float f = static_cast<float>(bounds[k]);
                if (static_cast<double>(f) < bounds[k])
                    f = std::nextafter(f, HUGE_VALF);
                up[k] = f;
            }
            if ((st = upload(c, t, up, &out.lutBoundUp)))
                return st;
        }
        std::vector<T> maj(size_t(d.lut_nodes) * 2);
        for (size_t i = 0; i < maj.size(); i += 2)
            pair(d.lut_majorant + i, i / 2, &maj[i]);
        if ((st = upload(c, t, maj, &out.lutMajorant)))
            return st;
        if ((st = upload(c, t, cast<T>(d.rita, size_t(d.n_materials) * 4 * 56), &out.rita)))
            return st;
        if ((st = upload(c, t, cast<T>(d.spline, size_t(d.n_materials) * 63), &out.spline)))
            return st;
        std::vector<dxmc_dev::Shell<T>> sh(size_t(d.n_materials) * 12);
        for (size_t i = 0; i < sh.size(); ++i) {
            const dxmc_shell& s = d.shells[i];
            sh[i].binding = static_cast<T>(s.binding_energy);
            sh[i].electrons = static_cast<T>(s.number_electrons);
            sh[i].j0 = static_cast<T>(s.hartree_fock_j0);
            sh[i].photoProb = static_cast<T>(s.photo_ionization_probability);
            sh[i].yield = static_cast<T>(s.fluorescence_yield);
            for (int k = 0; k < 3; ++k) {
                sh[i].lineProb[k] = static_cast<T>(s.line_probability[k]);
                sh[i].lineEnergy[k] = static_cast<T>(s.line_energy[k]);
            }
        }
        if ((st = upload(c, t, sh, &out.shells)))
            return st;
        // spectra (a dummy entry keeps the pointers valid when there are none)
        const uint32_t ns = d.n_spectra;
        std::vector<uint32_t> off(ns + 1, 0), thr(ns ? ns : 1, 0), alias;
        std::vector<T> prob, energy;
        for (uint32_t s = 0; s < ns; ++s) {
            off[s + 1] = d.spectrum_offset[s + 1];
            const uint64_t n = d.spectrum_offset[s + 1] - d.spectrum_offset[s];
            thr[s] = static_cast<uint32_t>((0ULL - n) % n);
        }
        const uint32_t total = ns ? d.spectrum_offset[ns] : 1;
        out.specTotal = total;
        prob = ns ? cast<T>(d.spectrum_probability, total) : std::vector<T>(1, T(1));
        energy = ns ? cast<T>(d.spectrum_energy, total) : std::vector<T>(1, T(0));
        alias = ns ? std::vector<uint32_t>(d.spectrum_alias, d.spectrum_alias + total) : std::vector<uint32_t>(1, 0);
        if ((st = upload(c, t, off, &out.specOffset)) || (st = upload(c, t, thr, &out.specThreshold)) || (st = upload(c, t, prob, &out.specProb))
            || (st = upload(c, t, alias, &out.specAlias)) || (st = upload(c, t, energy, &out.specEnergy)))
            return st;
        std::vector<dxmc_dev::Heel<T>> heel(d.n_heel ? d.n_heel : 1);
        size_t nhw = 1;
        for (uint32_t h = 0; h < d.n_heel; ++h) {
            const dxmc_heel& s = d.heel[h];
            heel[h] = { static_cast<T>(s.energy_start), static_cast<T>(s.energy_step), static_cast<T>(s.angle_start), static_cast<T>(s.angle_step),
                s.energy_size, s.angle_size, s.weight_offset, 0 };
            nhw = std::max<size_t>(nhw, size_t(s.weight_offset) + size_t(s.energy_size) * s.angle_size);
        }
        const std::vector<T> hw = d.n_heel ? cast<T>(d.heel_weight, nhw) : std::vector<T>(1, T(1));
        out.nHeel = static_cast<uint32_t>(heel.size());
        out.heelWeightTotal = static_cast<uint32_t>(nhw);
        if ((st = upload(c, t, heel, &out.heel)) || (st = upload(c, t, hw, &out.heelWeight)))
            return st;
        const uint32_t nb = d.n_bowtie;
        std::vector<uint32_t> boff(nb + 1, 0);
        for (uint32_t b = 0; b <= nb && nb; ++b)
            boff[b] = d.bowtie_offset[b];
        const uint32_t btotal = nb ? d.bowtie_offset[nb] : 1;
        out.nBowtie = nb;
        out.bowtieTotal = btotal;
        const std::vector<T> ba = nb ? cast<T>(d.bowtie_angle, btotal) : std::vector<T>(1, T(0));
        const std::vector<T> bw = nb ? cast<T>(d.bowtie_weight, btotal) : std::vector<T>(1, T(1));
        if ((st = upload(c, t, boff, &out.bowtieOffset)) || (st = upload(c, t, ba, &out.bowtieAngle)) || (st = upload(c, t, bw, &out.bowtieWeight)))
            return st;
        return DXMC_OK;
    }
}

int buildDeviceTables(dxmc_ctx* c, const dxmc_tables_desc& d, dxmc_tables& t)
{
    t.nMaterials = d.n_materials;
    t.nSpectra = d.n_spectra;
    t.nHeel = d.n_heel;
    t.nBowtie = d.n_bowtie;
    if (int st = build<double>(c, d, t, t.f64, false))
        return st;
    return build<float>(c, d, t, t.f32, true);
}
}
