// Material: photon interaction data for elements, compounds and NIST-style
// mixtures. The public interface is the reference's (material.hpp:60-139) so
// World/Tube/Source code written against DXMClib keeps compiling.
//
// The reference fills this class from xraylib. xraylib is not available to
// this project, so the data come from a compact physics model of our own
// (src/material.cpp): tabulated edges/atomic weights plus screened
// hydrogenic form factors, Waller-Hartree incoherent scatter functions,
// numerically integrated Klein-Nishina and Thomson cross sections and a
// jump-ratio photoelectric model. Absolute values therefore differ from
// xraylib by percents; every consumer (host tables, GPU kernels, oracle) sees
// the same numbers, so GPU/CPU parity is unaffected. A tabulated data set can
// replace the model through Material::setDataProvider().
#pragma once

#include "dxmc/floating.hpp"

#include <array>
#include <functional>
#include <string>
#include <vector>

namespace dxmc {

// One (grouped) electron shell of a material: reference material.hpp:29-58.
template <Floating T>
struct ElectronShellConfiguration {
    T bindingEnergy = 0; // keV
    T numberElectrons = 0; // fraction of all electrons of the material in this shell
    T hartreeFockOrbital_0 = 0; // Compton profile J(pz=0)
    T photoIonizationProbability = 1;
    T fluorescenceYield = 0;
    std::array<T, 3> fluorLineProbabilities = { 1, 1, 1 };
    std::array<T, 3> fluorLineEnergies = { 0, 0, 0 };
    int Z = 0;
    int shell = 0;

    template <Floating U>
    ElectronShellConfiguration<U> cast() const
    {
        ElectronShellConfiguration<U> c;
        c.bindingEnergy = static_cast<U>(bindingEnergy);
        c.numberElectrons = static_cast<U>(numberElectrons);
        c.hartreeFockOrbital_0 = static_cast<U>(hartreeFockOrbital_0);
        c.fluorescenceYield = static_cast<U>(fluorescenceYield);
        c.photoIonizationProbability = static_cast<U>(photoIonizationProbability);
        for (std::size_t i = 0; i < 3; ++i) {
            c.fluorLineProbabilities[i] = static_cast<U>(fluorLineProbabilities[i]);
            c.fluorLineEnergies[i] = static_cast<U>(fluorLineEnergies[i]);
        }
        c.Z = Z;
        c.shell = shell;
        return c;
    }
};

class Material {
public:
    // name: a NIST-style mixture name ("Water, Liquid"), a chemical formula
    // ("H2O", "C5H8O2", "Ca10(PO4)6(OH)2") or an element symbol ("Al").
    Material(const std::string& materialNameOrCompound = "", const std::string& prettyName = "", const double density = -1.0);
    Material(int atomicNumber);
    // explicit composition by mass fractions
    Material(const std::vector<int>& elements, const std::vector<double>& massFractions, const std::string& name, double density);

    bool isValid(void) const { return m_valid && m_hasDensity; }
    const std::string& name(void) const { return m_name; }
    // the composition as text (atomic numbers and the bits of the number
    // fractions): equal keys mean equal physics data
    std::string compositionKey(void) const;
    const std::string& prettyName(void) const;

    bool hasStandardDensity(void) const { return m_hasDensity; }
    double standardDensity(void) const { return m_density; }
    void setStandardDensity(double density); // g/cm3

    // sum over elements of number fraction * F(q,Z)^2, q in 1/Angstrom
    double getRayleightFormFactorSquared(const double momentumTransfer) const;
    // sum over elements of number fraction * S(q,Z)/Z
    double getComptonNormalizedScatterFactor(const double momentumTransfer) const;

    template <Floating T>
    T getRayleightFormFactorSquared(const T q) const { return static_cast<T>(getRayleightFormFactorSquared(static_cast<double>(q))); }
    template <Floating T>
    T getComptonNormalizedScatterFactor(const T q) const { return static_cast<T>(getComptonNormalizedScatterFactor(static_cast<double>(q))); }

    // mass attenuation coefficients [cm2/g], energy in keV
    double getPhotoelectricAttenuation(double energy) const;
    double getRayleightAttenuation(double energy) const;
    double getComptonAttenuation(double energy) const;
    double getTotalAttenuation(double energy) const;
    double getMassEnergyAbsorbtion(double energy) const;

    // all shell binding energies above minValue [keV], descending
    std::vector<double> getBindingEnergies(const double minValue = 1) const;
    template <Floating T>
    std::vector<T> getBindingEnergies(const T minValue = 1) const
    {
        const auto v = getBindingEnergies(static_cast<double>(minValue));
        return std::vector<T>(v.cbegin(), v.cend());
    }

    std::array<ElectronShellConfiguration<double>, 12> getElectronConfiguration() const;
    template <Floating T>
    std::array<ElectronShellConfiguration<T>, 12> getElectronConfiguration() const
    {
        const auto d = getElectronConfiguration();
        std::array<ElectronShellConfiguration<T>, 12> r;
        for (std::size_t i = 0; i < r.size(); ++i)
            r[i] = d[i].template cast<T>();
        return r;
    }

    const std::vector<int>& elements() const { return m_elements; }
    const std::vector<double>& elementNumberFractions() const { return m_elementNumberFraction; }

    static double getAtomicWeight(int Z);
    static std::string getAtomicNumberToSymbol(int Z);
    static std::vector<std::string> getNISTCompoundNames(void);
    static int getAtomicNumberFromSymbol(const std::string& symbol);
    static std::string getSymbolFromAtomicNumber(int Z);
    static double getTotalAttenuation(int atomicNumber, double energy);

    // Element-level physics data. The default provider is the built-in model;
    // a tabulated provider (e.g. exported xraylib tables) can be installed
    // before any Material is used.
    struct ElementData {
        std::function<double(int Z, double E)> photo; // barn/atom
        std::function<double(int Z, int shell, double E)> photoPartial; // barn/atom
        std::function<double(int Z, double E)> incoherent; // barn/atom
        std::function<double(int Z, double E)> coherent; // barn/atom
        std::function<double(int Z, double E)> energyAbsorption; // barn/atom
        std::function<double(int Z, double q)> formFactor; // F(q,Z)
        std::function<double(int Z, double q)> scatterFunction; // S(q,Z)
        std::function<int(int Z, int shell, double* binding, double* electrons, double* J0, double* yield,
            double lineProb[3], double lineEnergy[3])>
            shell; // returns 0 if the shell exists
    };
    static void setDataProvider(const ElementData& data);
    static const ElementData& dataProvider();

protected:
    void setByCompoundName(const std::string& name);
    void setByMaterialName(const std::string& name);
    void setByAtomicNumber(int index);
    void setComposition(const std::vector<int>& elements, const std::vector<double>& massFractions);

private:
    std::string m_name;
    std::string m_prettyName;
    std::vector<int> m_elements;
    std::vector<double> m_elementNumberFraction;
    double m_density = -1.0;
    bool m_valid = false;
    bool m_hasDensity = false;
};
}
