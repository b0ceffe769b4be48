// Population synthesis on the device (init.hip): the agent draws, the public-transport
// cap, the starting-infection selection, the start cells and the occupancy grid, written
// straight into the SoA and grid of one simulation.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <string>
#include <vector>

#include "config.hpp"
#include "population.hpp"

namespace epi {

// Device destinations of one simulation (slices of the batch's concatenated arrays).
struct InitTargets {
  uint32_t* pos;
  uint32_t* st;
  uint32_t* tm;
  uint32_t* wsa;
  uint8_t* occ;  // the simulation's padded grid
  uint32_t pitch;
  uint64_t cells;
};

// Fills `t` with the same agents synthesize_population builds on the host, on `stream`.
// `cfg` must have passed prepare_population. Returns hipSuccess or the first HIP error;
// `err` names the failing step.
hipError_t synthesize_on_device(const Config& cfg, uint64_t seed, const Geometry& geo, const InitTargets& t,
                                hipStream_t stream, std::string& err);

// The office layout's slot maps of one simulation (model.hpp) from its home-layout state
// words `st` (n agents): office_agent[slot] = id and office_slot[id] = slot, with the
// number of Normal and Essential workers of each office and in all, on `stream`.
hipError_t build_office_maps_on_device(const uint32_t* st, uint32_t n, uint32_t n_offices, uint32_t* office_agent,
                                      uint32_t* office_slot, std::vector<uint32_t>& workers_per_office,
                                      uint32_t& n_workers, hipStream_t stream, std::string& err);

}  // namespace epi
