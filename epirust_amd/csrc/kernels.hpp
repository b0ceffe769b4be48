// Launch interface of the hour-step kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>

#include "model.hpp"

namespace epi {

// Device pointers handed to every kernel by value.
struct DevView {
  // Agent arrays: pos and st in one layout, home (slot = id) or office (office != 0);
  // tm and wsa always in the home layout (indexed by id).
  uint32_t* pos;
  uint32_t* st;
  uint32_t* tm;
  uint32_t* wsa;
  uint32_t* tgt;
  uint8_t* occ;
  uint32_t* claim;  // per cell: lowest claim key of the day's latest claim hour, ~0u = none
  const SimDesc* desc;
  SimState* state;
  const AgentBlock* agent_blocks;  // per agent block; the office view has its own table
  const uint32_t* office_agent;    // office view: slot -> agent id (per simulation from agent_off)
  const uint32_t* office_slot;     // office view: agent id -> slot
  const BlockRef* hosp_blocks;   // per hospital-scan block: (sim, first cell)
  Counts7* rows;
  Event* events;
  const uint32_t* vacc_hour;
  const unsigned long long* vacc_thr;
  uint32_t* block_counts;  // 6 per agent block: the hour's partial S E I H R D
  // Host-visible (pinned) copies that every hour's epilogue writes: its row (indexed like
  // rows) and the simulation's state after it. A step call reads them after one drain.
  Counts7* host_rows;
  SimState* host_state;
  uint32_t office;         // 1: the agent arrays above are in the office layout
};

// One tile of a simulation's layout move: agents k * n_offices + j for k in k0..k0+31 and
// j in j0..j0+31.
struct TileRef {
  uint32_t sim, k0, j0;
};

struct LaunchShape {
  uint32_t block = 256;          // threads per agent block (home layout, and the night launch)
  uint32_t office_block = 256;   // threads per agent block of the office layout (hours of day 8..15)
  uint32_t agent_blocks = 0;     // total agent blocks over the batch (home layout)
  uint32_t office_agent_blocks = 0;  // the same for the office layout's table
  uint32_t hosp_block = 256;     // threads per hospital-scan block
  uint32_t hosp_cells_per_thread = 8;
  uint32_t hosp_blocks = 0;
  uint32_t n_sims = 0;
};

// One hour: [hour of day 0] scan, [hours of day 0 and 7..23] step, then the commit launch:
// the epilogue in its first blocks and the hour's agent pass in the rest (the commit at
// hours of day 0 and 7..23, the pending sweep / area reset / claim reset at 1..6).
// `hour` only needs the right parity, so a captured day of hours 1..24 replays for any day.
hipError_t launch_hospital_scan(const DevView& v, const LaunchShape& s, hipStream_t stream);
hipError_t launch_step(const DevView& v, const LaunchShape& s, hipStream_t stream);
hipError_t launch_commit(const DevView& v, const LaunchShape& s, uint32_t hour, hipStream_t stream);
// Copies st and pos between the home and office layouts (to_office: home -> office), one
// 32 x 32 tile of (k, j) per block.
hipError_t launch_relayout(const DevView& home, const DevView& office, const TileRef* tiles, uint32_t n_tiles,
                           bool to_office, hipStream_t stream);
// The epilogues of n_hours consecutive hours of day 3..6 in one launch, with no agent pass
// (only valid when no sweep can be pending at those hours: no vaccination at 2..5).
hipError_t launch_quiet_hours(const DevView& v, const LaunchShape& s, uint32_t n_hours, hipStream_t stream);
// Hours of day 1..6 in one launch (home layout): their epilogues, and the night's agent work
// (hour of day 0's sweep, areas to home, claim reset). Only valid when no vaccination falls
// at hours of day 1..5.
hipError_t launch_night(const DevView& v, const LaunchShape& s, hipStream_t stream);
// Re-applies the last decided sweep of every simulation (idempotent), before agent read-back.
hipError_t launch_apply_sweep(const DevView& v, const LaunchShape& s, hipStream_t stream);
inline bool hour_has_step(uint32_t hour) { return hour % 24u == 0 || hour % 24u >= 7; }
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t value, hipStream_t stream);
// Every simulation's active flag to `active` (device address of pinned host memory), for
// the host's daily poll inside a multi-day step call.
hipError_t launch_export_active(const DevView& v, uint32_t n_sims, uint32_t* active, hipStream_t stream);

}  // namespace epi
