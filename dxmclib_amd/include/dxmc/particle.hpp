// Photon state (reference particle.hpp:30-47): position [mm], direction,
// energy [keV] and statistical weight.
#pragma once
#include "dxmc/floating.hpp"
#include <array>

namespace dxmc {
template <Floating T = double>
struct Particle {
    std::array<T, 3> pos;
    std::array<T, 3> dir;
    T energy;
    T weight;
};
}
