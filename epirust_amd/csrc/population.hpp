// Geometry and population synthesis for one simulation, producing the packed SoA
// agent state that is uploaded to HBM.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "config.hpp"
#include "model.hpp"

namespace epi {

struct Geometry {
  uint32_t g = 0, stride = 0, height = 0;
  Rect housing{}, transport{}, work{}, hosp0{}, hosp1{};
  int32_t house_sx = 0, house_sy = 0, house_nx = 0, house_ny = 0;
  int32_t office_sx = 0, office_sy = 0, office_nx = 0, office_ny = 0;
  uint32_t n_houses = 0, n_offices = 0, n_transport_points = 0;
};

struct Population {
  Geometry geo;
  std::vector<uint32_t> pos, st, tm, wsa;
  uint32_t counts0[7] = {0, 0, 0, 0, 0, 0, 0};  // hour 0, S E I H R D (utils/util.rs:45-51)
};

// rand 0.8.5 Bernoulli threshold: floor(p * 2^64), p == 1 -> always.
uint64_t bernoulli_threshold(double p);

// Geography (geography/mod.rs:33-70, grid.rs:240-261) for `agents` on grid `g`.
bool build_geometry(const Config& cfg, Geometry& geo, std::string& err);

// Geometry, the reference's panic checks (house capacity, starting infections > N,
// Population::Csv home points) and the hour-0 counts; returns false with the panic text.
bool prepare_population(const Config& cfg, Population& pop, std::string& err);

// Deterministic-mode population synthesis on the host: prepare_population, then the
// agents' SoA words (the device path is synthesize_on_device in init.hpp).
bool synthesize_population(const Config& cfg, uint64_t seed, Population& pop, std::string& err);

// Bit helpers shared with the runtime's read-back.
inline uint32_t st_state(uint32_t w) { return (w >> B_STATE) & M_STATE; }
inline uint32_t st_sev(uint32_t w) { return (w >> B_SEV) & M_SEV; }
inline uint32_t st_ws(uint32_t w) { return (w >> B_WS) & M_WS; }
inline uint32_t st_area(uint32_t w) { return (w >> B_AREA) & M_AREA; }
inline int32_t st_imm(uint32_t w) { return static_cast<int32_t>((w >> B_IMM) & M_IMM) - 2; }
inline uint32_t st_day(uint32_t w) { return (w >> B_DAY) & M_DAY; }

}  // namespace epi
