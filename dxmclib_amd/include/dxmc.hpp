// Umbrella header of the dxmc host API (reference include/dxmc.hpp:18-25).
#pragma once
#include "dxmc/attenuationlut.hpp"
#include "dxmc/beamfilters.hpp"
#include "dxmc/dxmcrandom.hpp"
#include "dxmc/exposure.hpp"
#include "dxmc/material.hpp"
#include "dxmc/progressbar.hpp"
#include "dxmc/source.hpp"
#include "dxmc/transport.hpp"
#include "dxmc/tube.hpp"
#include "dxmc/world.hpp"
