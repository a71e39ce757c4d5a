// Flattening of host tables and exposures into the plain structures of the
// C ABI (include/dxmc_hip.h): TableSet owns the arrays a dxmc_tables_desc
// points into; ExposureTable collects exposures and assigns table indices
// to the spectra, heel and bowtie filters they use.
#pragma once

#include "dxmc/attenuationlut.hpp"
#include "dxmc/beamfilters.hpp"
#include "dxmc/exposure.hpp"
#include "dxmc_hip.h"

#include <map>
#include <stdexcept>
#include <vector>

namespace dxmc {

struct TableSet {
    std::vector<double> grid, coefficients, majorant, rita, spline;
    std::vector<dxmc_shell> shells;
    std::vector<std::uint32_t> spectrumOffset { 0 }, spectrumAlias;
    std::vector<double> spectrumProbability, spectrumEnergy;
    std::vector<dxmc_heel> heel;
    std::vector<double> heelWeight;
    std::vector<std::uint32_t> bowtieOffset { 0 };
    std::vector<double> bowtieAngle, bowtieWeight;
    std::uint32_t nMaterials = 0;
    std::uint32_t linearIndex = 0;
    double linearStep = 0, linearEnergy = 0;

    void setMaterialTables(const AttenuationLut<double>& lut)
    {
        const auto& ip = lut.interpolator();
        nMaterials = static_cast<std::uint32_t>(lut.formFactorTables().size());
        grid = ip.grid();
        coefficients = ip.coefficients();
        majorant = ip.maxCoefficients();
        linearIndex = static_cast<std::uint32_t>(ip.linearIndex());
        linearStep = ip.linearStep();
        linearEnergy = ip.linearEnergy();
        constexpr int R = AttenuationLut<double>::RITA_NODES;
        rita.assign(std::size_t(nMaterials) * 4 * R, 0.0);
        for (std::uint32_t m = 0; m < nMaterials; ++m) {
            const auto& r = lut.formFactorTables()[m];
            for (int i = 0; i < R; ++i) {
                rita[(m * 4 + 0) * R + i] = r.x()[i];
                rita[(m * 4 + 1) * R + i] = r.e()[i];
                rita[(m * 4 + 2) * R + i] = r.a()[i];
                rita[(m * 4 + 3) * R + i] = r.b()[i];
            }
        }
        spline.assign(std::size_t(nMaterials) * 63, 0.0);
        for (std::uint32_t m = 0; m < nMaterials; ++m) {
            const auto& s = lut.scatterFactorTables()[m];
            for (int i = 0; i < 60; ++i)
                spline[m * 63 + i] = s.coefficients()[i];
            spline[m * 63 + 60] = s.start();
            spline[m * 63 + 61] = s.step();
            spline[m * 63 + 62] = s.stop();
        }
        shells.assign(std::size_t(nMaterials) * 12, dxmc_shell {});
        for (std::uint32_t m = 0; m < nMaterials; ++m)
            for (int i = 0; i < 12; ++i) {
                const auto& c = lut.shellConfigurations()[m][i];
                dxmc_shell& s = shells[m * 12 + i];
                s.binding_energy = c.bindingEnergy;
                s.number_electrons = c.numberElectrons;
                s.hartree_fock_j0 = c.hartreeFockOrbital_0;
                s.photo_ionization_probability = c.photoIonizationProbability;
                s.fluorescence_yield = c.fluorescenceYield;
                for (int k = 0; k < 3; ++k) {
                    s.line_probability[k] = c.fluorLineProbabilities[k];
                    s.line_energy[k] = c.fluorLineEnergies[k];
                }
            }
    }

    template <Floating T>
    std::int32_t addSpectrum(const SpecterDistribution<T>& s)
    {
        for (std::size_t i = 0; i < s.size(); ++i) {
            spectrumProbability.push_back(static_cast<double>(s.probabilityData()[i]));
            spectrumAlias.push_back(static_cast<std::uint32_t>(s.aliasingData()[i]));
            spectrumEnergy.push_back(static_cast<double>(s.energies()[i]));
        }
        spectrumOffset.push_back(static_cast<std::uint32_t>(spectrumProbability.size()));
        return static_cast<std::int32_t>(spectrumOffset.size() - 2);
    }
    template <Floating T>
    std::int32_t addHeel(const HeelFilter<T>& h)
    {
        dxmc_heel d {};
        d.energy_start = h.energyStart();
        d.energy_step = h.energyStep();
        d.angle_start = h.angleStart();
        d.angle_step = h.angleStep();
        d.energy_size = static_cast<std::uint32_t>(h.energySize());
        d.angle_size = static_cast<std::uint32_t>(h.angleSize());
        d.weight_offset = static_cast<std::uint32_t>(heelWeight.size());
        for (const auto w : h.weights())
            heelWeight.push_back(static_cast<double>(w));
        heel.push_back(d);
        return static_cast<std::int32_t>(heel.size() - 1);
    }
    template <Floating T>
    std::int32_t addBowtie(const BowTieFilter<T>& b)
    {
        for (const auto& [a, w] : b.data()) {
            bowtieAngle.push_back(static_cast<double>(a));
            bowtieWeight.push_back(static_cast<double>(w));
        }
        bowtieOffset.push_back(static_cast<std::uint32_t>(bowtieAngle.size()));
        return static_cast<std::int32_t>(bowtieOffset.size() - 2);
    }

    dxmc_tables_desc desc() const
    {
        dxmc_tables_desc d {};
        d.n_materials = nMaterials;
        d.lut_nodes = static_cast<std::uint32_t>(grid.size());
        d.lut_grid = grid.data();
        d.lut_coefficients = coefficients.data();
        d.lut_majorant = majorant.data();
        d.lut_linear_index = linearIndex;
        d.lut_linear_step = linearStep;
        d.lut_linear_energy = linearEnergy;
        d.rita = rita.data();
        d.spline = spline.data();
        d.shells = shells.data();
        d.n_spectra = static_cast<std::uint32_t>(spectrumOffset.size() - 1);
        d.spectrum_offset = spectrumOffset.data();
        d.spectrum_probability = spectrumProbability.data();
        d.spectrum_alias = spectrumAlias.data();
        d.spectrum_energy = spectrumEnergy.data();
        d.n_heel = static_cast<std::uint32_t>(heel.size());
        d.heel = heel.data();
        d.heel_weight = heelWeight.data();
        d.n_bowtie = static_cast<std::uint32_t>(bowtieOffset.size() - 1);
        d.bowtie_offset = bowtieOffset.data();
        d.bowtie_angle = bowtieAngle.data();
        d.bowtie_weight = bowtieWeight.data();
        return d;
    }
};

// Collects aligned exposures and registers their sampler tables once each.
template <Floating T>
class ExposureTable {
public:
    explicit ExposureTable(TableSet& tables)
        : m_tables(tables)
    {
    }
    void add(const Exposure<T>& e)
    {
        dxmc_exposure r {};
        for (int i = 0; i < 3; ++i) {
            r.position[i] = e.position()[i];
            r.beam_direction[i] = e.beamDirection()[i];
        }
        for (int i = 0; i < 6; ++i)
            r.direction_cosines[i] = e.directionCosines()[i];
        for (int i = 0; i < 4; ++i)
            r.collimation_angles[i] = e.collimationAngles()[i];
        r.weight = e.beamIntensityWeight();
        r.mono_energy = e.monoenergeticPhotonEnergy();
        r.n_histories = e.numberOfHistories();
        r.spectrum = -1;
        r.heel = -1;
        r.bowtie = -1;
        if (const auto* s = e.specterDistribution()) {
            auto it = m_spectra.find(s);
            r.spectrum = it != m_spectra.end() ? it->second : (m_spectra[s] = m_tables.addSpectrum(*s));
        }
        if (const auto* h = e.heelFilter()) {
            auto it = m_heels.find(h);
            r.heel = it != m_heels.end() ? it->second : (m_heels[h] = m_tables.addHeel(*h));
        }
        if (const auto* f = e.beamFilter()) {
            const auto* b = dynamic_cast<const BowTieFilter<T>*>(f);
            if (!b)
                throw std::invalid_argument("only bowtie filters can act per photon");
            auto it = m_bowties.find(b);
            r.bowtie = it != m_bowties.end() ? it->second : (m_bowties[b] = m_tables.addBowtie(*b));
        }
        m_records.push_back(r);
    }
    const std::vector<dxmc_exposure>& records() const { return m_records; }

private:
    TableSet& m_tables;
    std::vector<dxmc_exposure> m_records;
    std::map<const void*, std::int32_t> m_spectra, m_heels, m_bowties;
};
}
