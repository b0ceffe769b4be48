// Launch interface of the citizen-state formatter (stream.hip).
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>

#include "model.hpp"

namespace epi {

// Longest record: comma + fixed text (52 bytes) + a 10-digit id + two 5-digit coordinates.
constexpr uint32_t kMaxRecord = 73;

struct StreamArgs {
  const uint32_t* pos;  // the simulation's slice of the SoA
  const uint32_t* st;
  uint32_t n;
  const SimState* state;  // records are written only if state->hour == hour
  uint32_t hour;
  unsigned long long* block_off;  // stream_blocks(n) entries of scratch
  unsigned long long* total;      // bytes written this hour
  char* text;                     // >= n * max_record_bytes(...) bytes
};

// Agents the no-allocation formatter handles: its per-block offsets live in a static
// device array of kNoAllocMaxBlocks entries (8 bytes each).
constexpr uint32_t kNoAllocMaxBlocks = 65536;
constexpr uint32_t kNoAllocMaxAgents = kNoAllocMaxBlocks * 256u;

// The current device's copy of the no-allocation scratch array (resolved once per device
// with hipGetSymbolAddress; call with the device current).
hipError_t noalloc_scratch(unsigned long long** out);

// Upper bound of one record's length for `n` agents on a grid of the given extent.
uint32_t max_record_bytes(uint32_t n, uint32_t stride, uint32_t height);
uint32_t stream_blocks(uint32_t n);

// Formats the hour's records (the message body between '[' and ']') into a.text.
hipError_t launch_citizen_states(const StreamArgs& a, hipStream_t stream);

}  // namespace epi
