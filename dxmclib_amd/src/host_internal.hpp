// Helpers shared by the host C API (host_capi.cpp) and the request server
// (server.cpp).
#pragma once

#include "dxmc/source.hpp"
#include "dxmc/world.hpp"
#include "dxmc_host.h"

#include <memory>
#include <string>

namespace dxmc_host_detail {
// records the message for dxmc_host_last_error() and returns status
int fail(int status, const std::string& msg);
// the Source subclass a spec describes, configured; nullptr for an unknown type
std::unique_ptr<dxmc::Source<double>> makeSource(const dxmc_source_spec& s);
// a validated World from plain arrays; false with the reason otherwise
bool buildWorld(const dxmc_host_world& w, dxmc::World<double>& world, std::string& why);
}
