// Low-energy correction models (reference lowenergycorrectionmodel.hpp:22-26).
// NONE: free-electron Klein-Nishina / Thomson; LIVERMORE: scatter-factor and
// form-factor corrections; IA: impulse approximation with shell binding.
#pragma once

namespace dxmc {
enum class LOWENERGYCORRECTION {
    NONE,
    LIVERMORE,
    IA
};
}
