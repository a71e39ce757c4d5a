// dxmc host API: floating point concept used by every template.
// Mirrors the reference's concept (floating.hpp:25) so that user code written
// against dxmc::World<T>, dxmc::Transport<T> ... compiles unchanged.
#pragma once
#include <type_traits>

namespace dxmc {
template <typename T>
concept Floating = std::is_floating_point_v<T>;
}
