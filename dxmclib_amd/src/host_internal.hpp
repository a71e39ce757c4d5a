// Helpers shared by the host C API (host_capi.cpp) and the request server
// (server.cpp).
#pragma once

#include "dxmc/source.hpp"
#include "dxmc/world.hpp"
#include "dxmc_host.h"

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

namespace dxmc_host_detail {
// records the message for dxmc_host_last_error() and returns status
int fail(int status, const std::string& msg);
// the Source subclass a spec describes, configured; nullptr for an unknown type
std::unique_ptr<dxmc::Source<double>> makeSource(const dxmc_source_spec& s);
// Where buildWorld() takes the World's arrays from: each call returns an
// empty vector with capacity for n elements (the request server reuses them).
struct WorldArrays {
    std::function<std::shared_ptr<std::vector<double>>(std::size_t)> density;
    std::function<std::shared_ptr<std::vector<std::uint8_t>>(std::size_t)> bytes;
};
// a validated World from plain arrays; false with the reason otherwise. With
// 'arrays', an absent measurement map is supplied as zeros from it too.
bool buildWorld(const dxmc_host_world& w, dxmc::World<double>& world, std::string& why, const WorldArrays* arrays = nullptr);
}
