// Host runtime of the C ABI (include/epirust_gpu.h): one handle = a batch of one or more
// independent simulations stepped together on one HIP stream.
#include <hip/hip_runtime_api.h>

#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "epirust_gpu.h"
#include "config.hpp"
#include "disease.hpp"
#include "init.hpp"
#include "kernels.hpp"
#include "model.hpp"
#include "population.hpp"
#include "stream.hpp"

using namespace epi;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                                  \
  do {                                                                                                 \
    hipError_t e_ = (expr);                                                                            \
    if (e_ != hipSuccess) return fail(EPI_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

bool valid_p(double p) { return p >= 0.0 && p <= 1.0; }

Rect building_rect(int32_t sx, int32_t sy, int32_t nx, int32_t size, uint32_t k) {
  int32_t x0 = sx + static_cast<int32_t>(k % static_cast<uint32_t>(nx)) * size;
  int32_t y0 = sy + static_cast<int32_t>(k / static_cast<uint32_t>(nx)) * size;
  return Rect{x0, y0, x0 + size - 1, y0 + size - 1};
}

bool rect_inside(const Rect& r, uint32_t stride, uint32_t height) {
  return r.x0 >= 0 && r.y0 >= 0 && r.x1 < static_cast<int32_t>(stride) && r.y1 < static_cast<int32_t>(height) &&
         r.x0 <= r.x1 && r.y0 <= r.y1;
}

struct DeviceBuffers {
  std::vector<void*> owned;
  template <class T>
  hipError_t alloc(T** p, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) owned.push_back(q);
    *p = static_cast<T*>(q);
    return e;
  }
  ~DeviceBuffers() {
    for (void* p : owned) (void)hipFree(p);
  }
};

// File sink of the citizen-state stream: one JSON line per message of the reference's
// citizen_states_updated topic (listeners/events_kafka_producer.rs:62-67, 90-121). The
// hour's body is formatted on the device (stream.hip) and copied into one of two pinned
// buffers, so the file write of hour h overlaps the step of hour h + 1.
struct StateSink {
  FILE* f = nullptr;
  std::string path;
  uint64_t cap = 0;  // text bytes of one hour, upper bound
  char* d_text = nullptr;
  unsigned long long* d_block_off = nullptr;
  unsigned long long* d_total = nullptr;
  char* h_text[2] = {nullptr, nullptr};
  unsigned long long* h_total = nullptr;  // pinned
  hipEvent_t ev_total = nullptr;
  hipEvent_t ev_text[2] = {nullptr, nullptr};
  ~StateSink() {
    if (f) std::fclose(f);
    for (hipEvent_t e : {ev_total, ev_text[0], ev_text[1]})
      if (e) (void)hipEventDestroy(e);
    for (char* p : h_text)
      if (p) (void)hipHostFree(p);
    if (h_total) (void)hipHostFree(h_total);
    for (void* p : {static_cast<void*>(d_text), static_cast<void*>(d_block_off), static_cast<void*>(d_total)})
      if (p) (void)hipFree(p);
  }
};

}  // namespace

struct epi_sim {
  std::vector<Config> cfg;
  std::vector<uint64_t> seed;
  std::vector<Geometry> geo;
  std::vector<SimDesc> desc;
  std::vector<uint32_t> counts0;  // 7 per sim
  uint32_t n_sims = 0;
  uint64_t n_agents = 0, n_cells = 0, n_rows = 0, n_events = 0;
  uint32_t max_hour = 0;  // largest hour any simulation can reach
  bool any_empty = false;
  uint32_t flags = 0;
  LaunchShape shape;
  std::vector<AgentBlock> agent_blocks, office_blocks;
  std::vector<BlockRef> hosp_blocks;
  // per simulation: the number of Normal and Essential workers of each office, in office
  // order (their slots come first in the office layout, office by office)
  std::vector<std::vector<uint32_t>> office_workers;
  std::vector<uint32_t> vacc_hour;
  std::vector<unsigned long long> vacc_thr;
  DeviceBuffers mem;
  DevView v{};         // agent arrays in the home layout
  DevView v_office{};  // the same view over the office-layout copies (hours of day 8..15)
  AgentBlock* d_agent_blocks = nullptr;
  AgentBlock* d_office_blocks = nullptr;
  uint32_t* d_office_agent = nullptr;  // office layout: slot -> agent id
  uint32_t* d_office_slot = nullptr;   // office layout: agent id -> slot
  TileRef* d_tiles = nullptr;          // layout move tiles of every simulation
  uint32_t n_tiles = 0;
  uint64_t cap_blocks = 0, cap_office_blocks = 0;  // entries of the two device tables
  // Day graphs run hours of day 3..6 as one epilogue-only launch: no simulation has a
  // vaccination at an hour of day 2..5, so no sweep is pending at 3..6.
  bool quiet_hours = false;
  // ... and hours of day 1..6 as one launch when none has one at 1..5 either, and no
  // simulation's last hour falls at 1..6 (a batch's day graph may hold a shorter run's
  // last hour): then no simulation stops inside the launch, whose agent blocks read the
  // stop flag while its epilogues run.
  bool night = false;
  SimDesc* d_desc = nullptr;
  hipStream_t stream = nullptr;
  hipGraphExec_t day_graph = nullptr;
  hipEvent_t poll_event[2] = {nullptr, nullptr};
  hipEvent_t loop_event[2] = {nullptr, nullptr};  // around the hour loop of one step call
  double last_loop_ms = 0.0;                      // device time of the last step call's hours
  uint32_t last_loop_hours = 0;
  bool loop_timed = false;  // loop_event brackets the last advance's launches
  uint32_t* pinned_active = nullptr;  // 2 x n_sims
  SimState* pinned_state = nullptr;   // n_sims: the states a step call reads back
  Counts7* pinned_rows = nullptr;     // n_rows: staging of the rows a step call returns
  // the same three buffers' device addresses, which the export kernels write
  uint32_t* d_pinned_active = nullptr;
  SimState* d_pinned_state = nullptr;
  Counts7* d_pinned_rows = nullptr;
  // The home-layout agent arrays hold the current state with the last decided sweep
  // applied; cleared by every step call, restored by settle() before the host reads agents.
  bool settled = true;
  uint32_t host_hour = 0;             // hours enqueued so far (every active simulation is at this hour)
  bool all_stopped = false;           // every simulation reported inactive
  std::vector<uint32_t> rows_returned;
  std::vector<SimState> state_h;
  std::unique_ptr<StateSink> sink;  // citizen-state stream, single-simulation handles only

  ~epi_sim() {
    if (day_graph) (void)hipGraphExecDestroy(day_graph);
    for (auto& e : poll_event)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : loop_event)
      if (e) (void)hipEventDestroy(e);
    if (pinned_active) (void)hipHostFree(pinned_active);
    if (pinned_state) (void)hipHostFree(pinned_state);
    if (pinned_rows) (void)hipHostFree(pinned_rows);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

int check_device() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(EPI_ERR_DEVICE, "no HIP device available (a gfx950 GPU is required)");
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(EPI_ERR_DEVICE, std::string("device is ") + prop.gcnArchName + ", this build targets gfx950");
  return EPI_OK;
}

int validate(const Config& c) {
  const std::pair<const char*, double> ps[] = {
      {"public_transport_percentage", c.public_transport_percentage},
      {"working_percentage", c.working_percentage},
      {"regular_transmission_rate", c.regular_transmission_rate},
      {"high_transmission_rate", c.high_transmission_rate},
      {"death_rate", c.death_rate},
      {"percentage_asymptomatic_population", c.percentage_asymptomatic_population},
      {"percentage_severe_infected_population", c.percentage_severe_infected_population},
      {"hospital_beds_percentage", c.hospital_beds_percentage},
      {"essential_workers_population", c.essential_workers_population}};
  for (auto& p : ps)
    if (!valid_p(p.second))
      return fail(EPI_ERR_PANIC, std::string("called `Result::unwrap()` on an `Err` value: InvalidProbability (") +
                                     p.first + "=" + std::to_string(p.second) + ")");
  for (auto& v : c.vaccinate)
    if (!valid_p(v.percent))
      return fail(EPI_ERR_PANIC, "called `Result::unwrap()` on an `Err` value: InvalidProbability (Vaccinate.percent=" +
                                     std::to_string(v.percent) + ")");
  if (c.hours > 24u * kMaxDay) return fail(EPI_ERR_CONFIG, "hours above the supported maximum of 98280");
  if (c.number_of_agents > kMaxAgents) return fail(EPI_ERR_CONFIG, "number_of_agents above 2^27 - 1");
  return EPI_OK;
}

// XCDs of an MI355X. The dispatcher hands a launch's workgroups to them round-robin
// (workgroup w to XCD w % 8), in increasing order within each, and each XCD has its own L2.
constexpr uint32_t kXcds = 8;

int build_launch_tables(epi_sim* s) {
  const uint32_t B = s->shape.block;
  s->agent_blocks.clear();
  for (uint32_t i = 0; i < s->n_sims; ++i) {
    const uint32_t n = s->desc[i].n_slots;
    const uint32_t nb = (n + B - 1) / B;
    s->desc[i].n_blocks = nb;
    s->desc[i].block_begin = static_cast<uint32_t>(s->agent_blocks.size());
    // Home-layout blocks in dispatch order. Housemates (ids congruent mod n_houses) sit
    // in different blocks, one per round of ids, and read the same windows and each
    // other's target words. Sorting the blocks by the house of their first slot makes a
    // house's blocks neighbours; spreading the sorted list so that each XCD gets one
    // contiguous run (table entry j runs on XCD j % 8) puts them on one XCD, close in
    // time, sharing its L2, with that XCD's band of house rows. The order changes no
    // result: claims are resolved by key and counts are summed.
    const uint32_t H = std::max(s->desc[i].n_houses, 1u);
    std::vector<uint32_t> sorted(nb);
    for (uint32_t b = 0; b < nb; ++b) sorted[b] = b;
    std::stable_sort(sorted.begin(), sorted.end(),
                     [&](uint32_t a, uint32_t b) { return (a * B) % H < (b * B) % H; });
    uint32_t start[kXcds];
    for (uint32_t x = 0, acc = 0; x < kXcds; ++x) {
      start[x] = acc;
      acc += nb / kXcds + (x < nb % kXcds ? 1u : 0u);
    }
    for (uint32_t j = 0; j < nb; ++j) {
      const uint32_t first = sorted[start[j % kXcds] + j / kXcds] * B;
      s->agent_blocks.push_back(AgentBlock{i, first, std::min(B, n - first), 1, 0, 1, 0, 0});
    }
  }
  s->shape.agent_blocks = static_cast<uint32_t>(s->agent_blocks.size());
  // Office layout, in blocks of Bo threads (shape.office_block, tuned apart from B): the
  // workers' blocks hold whole offices, as many consecutive ones as fit Bo slots and the
  // claim filter's kMaxFilterOffices; an office of more than Bo workers
  // spans blocks of its own. Then blocks of Bo slots over the other agents. Idle lanes cost
  // latency hiding, and offices are packed in order, so a block is more than half full
  // unless its next office alone would not fit beside it.
  // The two kinds of blocks are interleaved evenly in dispatch order: the other agents'
  // blocks make more global claims and fewer LDS ones, and run beside the workers' instead
  // of all at the end of the launch.
  s->office_blocks.clear();
  const uint32_t Bo = s->shape.office_block;
  std::vector<AgentBlock> wb, rb;
  for (uint32_t i = 0; i < s->n_sims && !s->office_workers.empty(); ++i) {
    SimDesc& D = s->desc[i];
    const std::vector<uint32_t>& cnt = s->office_workers[i];
    D.office_block_begin = static_cast<uint32_t>(s->office_blocks.size());
    wb.clear();
    rb.clear();
    uint32_t slot = 0, j = 0;
    const uint32_t O = static_cast<uint32_t>(cnt.size());
    while (j < O) {
      if (cnt[j] > Bo) {
        for (uint32_t k = 0; k < cnt[j]; k += Bo) wb.push_back(AgentBlock{i, slot + k, std::min(Bo, cnt[j] - k), j, j, 1, 0, 0});
        slot += cnt[j];
        ++j;
        continue;
      }
      const uint32_t first = slot, j_lo = j;
      uint32_t count = 0;
      while (j < O && j - j_lo < kMaxFilterOffices && count + cnt[j] <= Bo) count += cnt[j++];
      if (count) wb.push_back(AgentBlock{i, first, count, j_lo, j - 1, j_lo, j - 1, 0});
      slot += count;
    }
    for (uint32_t first = slot; first < D.n; first += Bo)
      rb.push_back(AgentBlock{i, first, std::min(Bo, D.n - first), 1, 0, 1, 0, 0});
    const size_t total = wb.size() + rb.size();
    std::vector<AgentBlock> merged;
    merged.reserve(total);
    for (size_t t = 0, w = 0, r = 0; t < total; ++t) {
      // block t is the r-th other-agent block when r * total <= t * |rb| (an even spread)
      if (r < rb.size() && (w == wb.size() || r * total <= t * rb.size())) merged.push_back(rb[r++]);
      else merged.push_back(wb[w++]);
    }
    // Each XCD runs one contiguous run of the merged list (table entry t runs on XCD
    // t % 8), so that neighbouring offices, and neighbouring houses, share its L2.
    const uint32_t nb = static_cast<uint32_t>(total);
    uint32_t start[kXcds];
    for (uint32_t x = 0, acc = 0; x < kXcds; ++x) {
      start[x] = acc;
      acc += nb / kXcds + (x < nb % kXcds ? 1u : 0u);
    }
    for (uint32_t t = 0; t < nb; ++t) s->office_blocks.push_back(merged[start[t % kXcds] + t / kXcds]);
    D.office_n_blocks = static_cast<uint32_t>(s->office_blocks.size()) - D.office_block_begin;
  }
  s->shape.office_agent_blocks = static_cast<uint32_t>(s->office_blocks.size());
  if (s->agent_blocks.size() > s->cap_blocks || s->office_blocks.size() > s->cap_office_blocks)
    return fail(EPI_ERR_DEVICE, "agent block tables exceed their allocation");
  return EPI_OK;
}

// The office layout's slot order (model.hpp): per simulation, the Normal and Essential
// workers office by office (office = id % n_offices), each office's in id order, then the
// other agents in id order. Work statuses are fixed at creation, so the slot maps are
// built once, on the device from the synthesized state words, with the layout move's tiles.
int build_office_order(epi_sim* s) {
  s->office_workers.assign(s->n_sims, {});
  std::vector<TileRef> tiles;
  for (uint32_t i = 0; i < s->n_sims; ++i) {
    SimDesc& D = s->desc[i];
    const uint32_t n = D.n, O = D.n_offices;
    std::string err;
    const hipError_t e = build_office_maps_on_device(s->v.st + D.agent_off, n, O, s->d_office_agent + D.agent_off,
                                                     s->d_office_slot + D.agent_off, s->office_workers[i],
                                                     D.n_workers, s->stream, err);
    if (e != hipSuccess) return fail(EPI_ERR_DEVICE, "office layout (" + err + "): " + hipGetErrorString(e));
    // Tiles column by column (32 offices, every k), so that consecutive tiles share the
    // cache lines of their office slots: an office's workers, split over a column's k
    // tiles, and the other agents of neighbouring columns, whose slots follow id order.
    const uint32_t K = (n + O - 1u) / O;  // rows k of ids k * O + j
    for (uint32_t j0 = 0; j0 < O; j0 += 32)
      for (uint32_t k0 = 0; k0 < K; k0 += 32) tiles.push_back(TileRef{i, k0, j0});
  }
  // Each XCD runs one contiguous run of that order (launch entry t runs on XCD t % 8), so
  // that the partial lines of neighbouring tiles meet in one L2 before they are written back.
  const uint32_t nt = static_cast<uint32_t>(tiles.size());
  uint32_t start[kXcds];
  for (uint32_t x = 0, acc = 0; x < kXcds; ++x) {
    start[x] = acc;
    acc += nt / kXcds + (x < nt % kXcds ? 1u : 0u);
  }
  std::vector<TileRef> spread(nt);
  for (uint32_t t = 0; t < nt; ++t) spread[t] = tiles[start[t % kXcds] + t / kXcds];
  s->n_tiles = nt;
  HIP_TRY(s->mem.alloc(&s->d_tiles, spread.size()));
  if (!spread.empty())
    HIP_TRY(hipMemcpy(s->d_tiles, spread.data(), spread.size() * sizeof(TileRef), hipMemcpyHostToDevice));
  return EPI_OK;
}

// The device table is allocated for 64-thread blocks (the most blocks any shape needs).
int upload_tables(epi_sim* s) {
  if (s->shape.agent_blocks)
    HIP_TRY(hipMemcpyAsync(s->d_agent_blocks, s->agent_blocks.data(), s->agent_blocks.size() * sizeof(AgentBlock),
                           hipMemcpyHostToDevice, s->stream));
  if (s->shape.office_agent_blocks)
    HIP_TRY(hipMemcpyAsync(s->d_office_blocks, s->office_blocks.data(), s->office_blocks.size() * sizeof(AgentBlock),
                           hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(s->d_desc, s->desc.data(), s->desc.size() * sizeof(SimDesc), hipMemcpyHostToDevice,
                         s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return EPI_OK;
}

// Copies the agent arrays into the other layout.
int relayout(epi_sim* s, bool to_office) {
  HIP_TRY(launch_relayout(s->v, s->v_office, s->d_tiles, s->n_tiles, to_office, s->stream));
  return EPI_OK;
}

// Before the host reads agents: the home layout holds the current state (it is copied back
// from the office layout after an hour of day 8..15) and includes the last decided sweep.
// Step calls leave this to the first agent read after them; returns with the stream drained.
int settle(epi_sim* s) {
  if (!s->settled) {
    if (s->host_hour > 0 && office_hour(s->host_hour)) {
      const int rc = relayout(s, false);
      if (rc != EPI_OK) return rc;
    }
    HIP_TRY(launch_apply_sweep(s->v, s->shape, s->stream));
    s->settled = true;
  }
  HIP_TRY(hipStreamSynchronize(s->stream));
  return EPI_OK;
}

int enqueue_hour(epi_sim* s, uint32_t next_hour) {
  const uint32_t hd = next_hour % 24u;
  if (hd == 0) HIP_TRY(launch_hospital_scan(s->v, s->shape, s->stream));
  // the office layout holds the agent arrays from hour of day 8's step to 15's commit
  if (hd == 8 || hd == 16) {
    const int rc = relayout(s, hd == 8);
    if (rc != EPI_OK) return rc;
  }
  const DevView& v = office_hour(next_hour) ? s->v_office : s->v;
  if (hour_has_step(next_hour)) HIP_TRY(launch_step(v, s->shape, s->stream));
  HIP_TRY(launch_commit(v, s->shape, next_hour, s->stream));
  return EPI_OK;
}

// Hours first..last of one day (24d+1 .. 24d+24). Hours of day 1..6 are one launch when
// the handle allows the night launch and the range holds all six, and 3..6 one
// epilogue-only launch when it allows quiet hours and the range holds all four.
int enqueue_hours(epi_sim* s, uint32_t first, uint32_t last) {
  for (uint32_t h = first; h <= last;) {
    const uint32_t hd = h % 24u;
    if (s->night && hd == 1 && last - h >= 5u) {
      HIP_TRY(launch_night(s->v, s->shape, s->stream));
      h += 6;
      continue;
    }
    if (s->quiet_hours && hd == 3 && last - h >= 3u) {
      HIP_TRY(launch_quiet_hours(s->v, s->shape, 4, s->stream));
      h += 4;
      continue;
    }
    const int rc = enqueue_hour(s, h);
    if (rc != EPI_OK) return rc;
    ++h;
  }
  return EPI_OK;
}

// The same hours as separate launches each (EPI_FLAG_NO_GRAPH).
int enqueue_hours_plain(epi_sim* s, uint32_t first, uint32_t last) {
  for (uint32_t h = first; h <= last; ++h) {
    const int rc = enqueue_hour(s, h);
    if (rc != EPI_OK) return rc;
  }
  return EPI_OK;
}

int capture_day_graph(epi_sim* s) {
  if (s->day_graph) return EPI_OK;
  hipGraph_t graph = nullptr;
  HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_hours(s, 1, 24);  // the hours of day of 24d+1 .. 24d+24
  hipError_t e = hipStreamEndCapture(s->stream, &graph);
  if (rc != EPI_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return fail(EPI_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&s->day_graph, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    s->day_graph = nullptr;
    return fail(EPI_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  return EPI_OK;
}

int fetch_states(epi_sim* s) {
  HIP_TRY(hipMemcpyAsync(s->state_h.data(), s->v.state, s->n_sims * sizeof(SimState), hipMemcpyDeviceToHost,
                         s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return EPI_OK;
}

bool all_inactive(const epi_sim* s, const uint32_t* flags) {
  for (uint32_t i = 0; i < s->n_sims; ++i)
    if (flags[i]) return false;
  return true;
}

int write_state_line(StateSink& k, uint32_t hour, uint64_t bytes, const char* text) {
  // CitizenStatesAtHr serialized by serde_json: {"hr":h,"citizen_states":[...]}
  if (std::fprintf(k.f, "{\"hr\":%u,\"citizen_states\":[", hour) < 0 || std::fwrite(text, 1, bytes, k.f) != bytes ||
      std::fputs("]}\n", k.f) < 0)
    return fail(EPI_ERR_IO, "Failed to write citizen states to " + k.path);
  return EPI_OK;
}

// The hour loop while the citizen-state stream is open: hour by hour, each hour's records
// formatted on the device right after its commit. The file write of one hour runs while
// the GPU steps the next. A zero-length body means the simulation stopped earlier.
int advance_streaming(epi_sim* s, uint32_t target) {
  StateSink& k = *s->sink;
  const SimDesc& D = s->desc[0];
  StreamArgs a{s->v.pos + D.agent_off, s->v.st + D.agent_off, D.n, s->v.state, 0u,
               k.d_block_off, k.d_total, k.d_text};
  uint32_t slot = 0, prev_hour = 0;
  uint64_t prev_bytes = 0;
  bool have_prev = false;
  int rc = EPI_OK;
  while (s->host_hour < target) {
    const uint32_t h = s->host_hour + 1;
    rc = enqueue_hour(s, h);
    if (rc != EPI_OK) return rc;
    if (office_hour(h)) {  // the records are formatted from the home layout
      rc = relayout(s, false);
      if (rc != EPI_OK) return rc;
    }
    a.hour = h;
    HIP_TRY(launch_citizen_states(a, s->stream));
    HIP_TRY(hipMemcpyAsync(k.h_total + slot, k.d_total, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           s->stream));
    HIP_TRY(hipEventRecord(k.ev_total, s->stream));
    ++s->host_hour;
    if (have_prev) {  // overlaps the hour just enqueued
      HIP_TRY(hipEventSynchronize(k.ev_text[slot ^ 1u]));
      rc = write_state_line(k, prev_hour, prev_bytes, k.h_text[slot ^ 1u]);
      if (rc != EPI_OK) return rc;
      have_prev = false;
    }
    HIP_TRY(hipEventSynchronize(k.ev_total));
    const uint64_t bytes = k.h_total[slot];
    if (bytes == 0) break;
    if (bytes > k.cap) return fail(EPI_ERR_DEVICE, "citizen-state text exceeds its bound");
    HIP_TRY(hipMemcpyAsync(k.h_text[slot], k.d_text, bytes, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipEventRecord(k.ev_text[slot], s->stream));
    prev_hour = h;
    prev_bytes = bytes;
    have_prev = true;
    slot ^= 1u;
  }
  if (have_prev) {
    HIP_TRY(hipEventSynchronize(k.ev_text[slot ^ 1u]));
    rc = write_state_line(k, prev_hour, prev_bytes, k.h_text[slot ^ 1u]);
    if (rc != EPI_OK) return rc;
  }
  if (std::fflush(k.f) != 0) return fail(EPI_ERR_IO, "Failed to write citizen states to " + k.path);
  HIP_TRY(hipStreamSynchronize(s->stream));
  return EPI_OK;
}

// Enqueue hours until `target` (or until every simulation has stopped). Whole days are
// replayed from the captured graph; stop flags are polled one day behind the GPU.
int advance(epi_sim* s, uint32_t target) {
  s->last_loop_ms = 0.0;
  s->last_loop_hours = 0;
  s->loop_timed = false;
  if (s->any_empty && target > s->host_hour) return fail(EPI_ERR_PANIC, "No citizens!");
  if (s->all_stopped) return EPI_OK;
  if (s->sink) {
    const uint32_t first = s->host_hour;
    const int rc = advance_streaming(s, target);
    if (s->host_hour != first) s->settled = false;
    return rc;
  }
  const bool graph_ok = !(s->flags & EPI_FLAG_NO_GRAPH);
  const uint32_t first_hour = s->host_hour;
  HIP_TRY(hipEventRecord(s->loop_event[0], s->stream));
  uint32_t day_index = 0;
  bool pending[2] = {false, false};
  while (s->host_hour < target) {
    // poll the flags copied two days ago
    const uint32_t slot = day_index & 1u;
    if (pending[slot]) {
      HIP_TRY(hipEventSynchronize(s->poll_event[slot]));
      pending[slot] = false;
      if (all_inactive(s, s->pinned_active + slot * s->n_sims)) break;
    }
    if (graph_ok && s->host_hour % 24u == 0 && target - s->host_hour >= 24u) {
      int rc = capture_day_graph(s);
      if (rc != EPI_OK) return rc;
      HIP_TRY(hipGraphLaunch(s->day_graph, s->stream));
      s->host_hour += 24;
    } else {
      const uint32_t end = std::min(target, (s->host_hour / 24u + 1u) * 24u);
      const int rc = graph_ok ? enqueue_hours(s, s->host_hour + 1, end) : enqueue_hours_plain(s, s->host_hour + 1, end);
      if (rc != EPI_OK) return rc;
      s->host_hour = end;
    }
    // the active flags of every simulation into pinned memory, polled two chunks later
    // (not after the call's last chunk: the call's read-back has the states)
    if (s->host_hour < target) {
      HIP_TRY(launch_export_active(s->v, s->n_sims, s->d_pinned_active + slot * s->n_sims, s->stream));
      HIP_TRY(hipEventRecord(s->poll_event[slot], s->stream));
      pending[slot] = true;
    }
    ++day_index;
  }
  HIP_TRY(hipEventRecord(s->loop_event[1], s->stream));
  s->loop_timed = true;
  s->last_loop_hours = s->host_hour - first_hour;
  if (s->last_loop_hours) s->settled = false;
  return EPI_OK;
}

// After the stream has drained: device time of the last advance's hour loop.
int record_loop_time(epi_sim* s) {
  if (!s->loop_timed) return EPI_OK;
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, s->loop_event[0], s->loop_event[1]));
  s->last_loop_ms = ms;
  return EPI_OK;
}


// The hours are enqueued and the stream is drained once. Each hour's epilogue has written
// its row and its simulation's state into pinned memory, where they are read.
int step_impl(epi_sim* s, uint32_t n_hours, epi_counts* rows, uint32_t row_stride, uint32_t* n_written,
              int* stopped) {
  const uint64_t want = static_cast<uint64_t>(s->host_hour) + n_hours;
  const uint32_t target = static_cast<uint32_t>(std::min<uint64_t>(want, s->max_hour));
  const uint32_t start = s->host_hour;
  int rc = advance(s, target);
  if (rc != EPI_OK) return rc;
  const uint32_t ran = s->host_hour - start;
  HIP_TRY(hipStreamSynchronize(s->stream));
  rc = record_loop_time(s);
  if (rc != EPI_OK) return rc;
  std::memcpy(s->state_h.data(), s->pinned_state, s->n_sims * sizeof(SimState));
  s->all_stopped = true;
  for (uint32_t i = 0; i < s->n_sims; ++i) s->all_stopped = s->all_stopped && !s->state_h[i].active;
  for (uint32_t i = 0; i < s->n_sims; ++i) {
    const SimState& st = s->state_h[i];
    if (st.error) return fail(EPI_ERR_PANIC, "device reported an error state");
    const uint32_t from = s->rows_returned[i], to = st.rows_written;
    const uint32_t k = to - from;
    if (k > n_hours || k > ran) return fail(EPI_ERR_DEVICE, "row accounting mismatch");
    if (rows && k)
      std::memcpy(rows + static_cast<size_t>(i) * row_stride, s->pinned_rows + s->desc[i].row_off + from,
                  k * sizeof(Counts7));
    s->rows_returned[i] = to;
    if (n_written) n_written[i] = k;
    if (stopped) stopped[i] = st.active ? 0 : 1;
  }
  return EPI_OK;
}

int create_impl(const std::vector<std::string>& texts, const std::vector<uint64_t>& seeds, uint32_t flags,
                epi_sim** out) {
  if (!out) return fail(EPI_ERR_ARG, "out is null");
  *out = nullptr;
  if (texts.empty()) return fail(EPI_ERR_ARG, "no simulations");
  std::unique_ptr<epi_sim> s(new epi_sim());
  s->n_sims = static_cast<uint32_t>(texts.size());
  s->flags = flags;
  s->cfg.resize(s->n_sims);
  s->seed = seeds;
  for (uint32_t i = 0; i < s->n_sims; ++i) {
    std::string err;
    if (!parse_config(texts[i], s->cfg[i], err)) return fail(EPI_ERR_CONFIG, err);
    if (s->cfg[i].csv_population && !read_population_csv(s->cfg[i], err)) return fail(EPI_ERR_PANIC, err);
    int rc = validate(s->cfg[i]);
    if (rc != EPI_OK) return rc;
  }
  int rc = check_device();
  if (rc != EPI_OK) return rc;
  {  // set-up of the no-allocation entry point: this device's copy of its scratch array
    unsigned long long* scratch = nullptr;
    HIP_TRY(noalloc_scratch(&scratch));
  }

  // geometry and the reference's panic checks on the host; the agents themselves are
  // synthesized on the device unless EPI_FLAG_HOST_INIT asks for the host restatement
  const bool host_init = (flags & EPI_FLAG_HOST_INIT) != 0;
  std::vector<Population> pops(s->n_sims);
  std::vector<std::string> errs(s->n_sims);
  std::vector<char> ok(s->n_sims, 0);
  {
    const unsigned nt = host_init ? std::max(1u, std::min<unsigned>(s->n_sims, 8u)) : 1u;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
      th.emplace_back([&, t] {
        for (uint32_t i = t; i < s->n_sims; i += nt)
          ok[i] = (host_init ? synthesize_population(s->cfg[i], s->seed[i], pops[i], errs[i])
                             : prepare_population(s->cfg[i], pops[i], errs[i]))
                      ? 1
                      : 0;
      });
    for (auto& x : th) x.join();
  }
  for (uint32_t i = 0; i < s->n_sims; ++i)
    if (!ok[i]) return fail(EPI_ERR_PANIC, errs[i]);

  // descriptors and offsets
  s->desc.resize(s->n_sims);
  s->geo.resize(s->n_sims);
  s->counts0.resize(7 * s->n_sims);
  s->rows_returned.assign(s->n_sims, 0);
  s->state_h.resize(s->n_sims);
  uint64_t agent_off = 0, cell_off = 0, row_off = 0, event_off = 0;
  for (uint32_t i = 0; i < s->n_sims; ++i) {
    const Config& c = s->cfg[i];
    const Geometry& G = pops[i].geo;
    s->geo[i] = G;
    SimDesc& D = s->desc[i];
    std::memset(&D, 0, sizeof D);
    D.n = c.number_of_agents;
    D.g = G.g;
    D.stride = G.stride;
    D.pitch = padded_pitch(G.stride);
    D.height = G.height;
    D.hours = c.hours;
    D.agent_off = agent_off;
    D.cell_off = cell_off;
    D.row_off = row_off;
    D.event_off = event_off;
    D.event_cap = 4u + static_cast<uint32_t>(c.vaccinate.size()) + 2u * (c.hours / 24u + 1u);
    D.seed_lo = static_cast<uint32_t>(s->seed[i]);
    D.seed_hi = static_cast<uint32_t>(s->seed[i] >> 32);
    D.housing = G.housing;
    D.transport = G.transport;
    D.work = G.work;
    D.hosp[0] = G.hosp0;
    D.hosp[1] = G.hosp1;
    D.house_sx = G.house_sx;
    D.house_sy = G.house_sy;
    D.house_nx = std::max(G.house_nx, 1);
    D.n_houses = std::max(G.n_houses, 1u);
    D.office_sx = G.office_sx;
    D.office_sy = G.office_sy;
    D.office_nx = std::max(G.office_nx, 1);
    D.n_offices = std::max(G.n_offices, 1u);
    D.fd_houses = make_fastdiv(D.n_houses);
    D.fd_house_nx = make_fastdiv(static_cast<uint32_t>(D.house_nx));
    D.fd_offices = make_fastdiv(D.n_offices);
    D.fd_office_nx = make_fastdiv(static_cast<uint32_t>(D.office_nx));
    D.n_slots = D.n;
    D.dz = Disease::params(c);
    D.has_lockdown = c.has_lockdown;
    D.lockdown_at = c.lockdown_at_number_of_infections;
    D.has_hospital = c.has_build_new_hospital;
    D.spread_rate_threshold = c.spread_rate_threshold;
    // vaccination schedule: HashMap<hour, percent> semantics (vaccination.rs:37-50)
    std::vector<std::pair<uint32_t, double>> sched;
    for (auto& v : c.vaccinate) {
      auto it = std::find_if(sched.begin(), sched.end(), [&](auto& e) { return e.first == v.at_hour; });
      if (it != sched.end()) it->second = v.percent; else sched.emplace_back(v.at_hour, v.percent);
    }
    std::sort(sched.begin(), sched.end());
    D.vacc_off = static_cast<uint32_t>(s->vacc_hour.size());
    D.n_vacc = static_cast<uint32_t>(sched.size());
    for (auto& e : sched) {
      s->vacc_hour.push_back(e.first);
      s->vacc_thr.push_back(bernoulli_threshold(e.second));
    }
    for (const Rect* r : {&G.housing, &G.transport, &G.work, &G.hosp0, &G.hosp1})
      if (!rect_inside(*r, G.stride, G.height))
        return fail(EPI_ERR_CONFIG, "grid_size too small for the engine's area layout");
    if (c.number_of_agents == 0 && c.hours > 1) s->any_empty = true;
    agent_off += D.n_slots;
    cell_off += padded_cells(G.stride, G.height);
    row_off += c.hours > 1 ? c.hours - 1 : 0;
    event_off += D.event_cap;
    s->max_hour = std::max(s->max_hour, c.hours > 1 ? c.hours - 1 : 0u);
    std::memcpy(&s->counts0[7 * i], pops[i].counts0, sizeof(pops[i].counts0));
    const auto cells = [](const Rect& r) {
      return static_cast<uint32_t>(r.x1 - r.x0 + 1) * static_cast<uint32_t>(r.y1 - r.y0 + 1);
    };
    const uint32_t hosp_cells = std::max(cells(G.hosp0), cells(G.hosp1));
    const uint32_t per_block = s->shape.hosp_block * s->shape.hosp_cells_per_thread;
    for (uint32_t first = 0; first < hosp_cells; first += per_block)
      s->hosp_blocks.push_back(BlockRef{i, first});
  }
  s->n_agents = agent_off;
  s->n_cells = cell_off;
  s->n_rows = row_off;
  s->n_events = event_off;
  s->shape.hosp_blocks = static_cast<uint32_t>(s->hosp_blocks.size());
  s->shape.n_sims = s->n_sims;
  // Default threads per agent block, as measured on MI355X: 512 for one simulation of at
  // least 2^19 agents (1M agents: 31 against 33 µs per hour at 256; 10M: 231 against 250),
  // 256 for batches and smaller runs, the same at office hours. epi_set_block_size
  // overrides both, epi_set_office_block_size then the office layout's alone.
  s->shape.block = (s->n_sims == 1 && s->desc[0].n >= (1u << 19)) ? 512u : 256u;
  s->shape.office_block = s->shape.block;
  s->quiet_hours = true;
  s->night = true;
  for (uint32_t h : s->vacc_hour) {
    if (h % 24u >= 2u && h % 24u <= 5u) s->quiet_hours = false;
    if (h % 24u >= 1u && h % 24u <= 5u) s->night = false;
  }
  for (uint32_t i = 0; i < s->n_sims; ++i) {
    const uint32_t last = s->desc[i].hours > 0 ? s->desc[i].hours - 1u : 0u;
    if (last % 24u >= 1u && last % 24u <= 6u) s->night = false;
  }

  // device memory
  HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  DeviceBuffers& M = s->mem;
  DevView& v = s->v;
  HIP_TRY(M.alloc(&v.pos, s->n_agents));
  HIP_TRY(M.alloc(&v.st, s->n_agents));
  HIP_TRY(M.alloc(&v.tm, s->n_agents));
  HIP_TRY(M.alloc(&v.wsa, s->n_agents));
  HIP_TRY(M.alloc(&v.tgt, s->n_agents));
  DevView& vo = s->v_office;  // completed from v below
  HIP_TRY(M.alloc(&vo.pos, s->n_agents));
  HIP_TRY(M.alloc(&vo.st, s->n_agents));
  HIP_TRY(M.alloc(&v.occ, s->n_cells));
  HIP_TRY(M.alloc(&v.claim, s->n_cells));
  HIP_TRY(M.alloc(&s->d_desc, s->n_sims));
  v.desc = s->d_desc;
  HIP_TRY(M.alloc(&v.state, s->n_sims));
  // tables sized for the smallest legal block (64 threads) so block-size changes never
  // reallocate: the office layout has at most one packed block per office, one split
  // block per office and per 64 workers, and one block per 64 other agents
  uint64_t max_blocks = 0, max_office_blocks = 0;
  for (uint32_t i = 0; i < s->n_sims; ++i) {
    max_blocks += (s->desc[i].n_slots + 63) / 64;
    max_office_blocks += 2ull * s->desc[i].n_offices + 2ull * ((s->desc[i].n_slots + 63) / 64) + 2;
  }
  HIP_TRY(M.alloc(&s->d_agent_blocks, max_blocks));
  HIP_TRY(M.alloc(&s->d_office_blocks, max_office_blocks));
  s->cap_blocks = max_blocks;
  s->cap_office_blocks = max_office_blocks;
  HIP_TRY(M.alloc(&s->d_office_agent, s->n_agents));
  HIP_TRY(M.alloc(&s->d_office_slot, s->n_agents));
  HIP_TRY(M.alloc(&v.block_counts, 6 * std::max(max_blocks, max_office_blocks)));
  v.agent_blocks = s->d_agent_blocks;
  BlockRef* d_hosp_blocks = nullptr;
  HIP_TRY(M.alloc(&d_hosp_blocks, s->hosp_blocks.size()));
  v.hosp_blocks = d_hosp_blocks;
  HIP_TRY(M.alloc(&v.rows, s->n_rows));
  HIP_TRY(M.alloc(&v.events, s->n_events));
  uint32_t* d_vacc_hour = nullptr;
  unsigned long long* d_vacc_thr = nullptr;
  HIP_TRY(M.alloc(&d_vacc_hour, s->vacc_hour.size()));
  HIP_TRY(M.alloc(&d_vacc_thr, s->vacc_thr.size()));
  v.vacc_hour = d_vacc_hour;
  v.vacc_thr = d_vacc_thr;

  std::vector<uint8_t> grid;
  for (uint32_t i = 0; i < s->n_sims; ++i) {
    const SimDesc& D = s->desc[i];
    const Population& P = pops[i];
    const size_t nb = static_cast<size_t>(D.n) * sizeof(uint32_t);
    SimState st;
    std::memset(&st, 0, sizeof st);
    st.active = (s->cfg[i].hours > 1 && D.n > 0) ? 1u : 0u;
    st.commit_hour[1] = st.active;  // the agent pass of hour 1 (odd)
    for (int k = 0; k < 6; ++k) st.last_counts[k] = pops[i].counts0[k + 1];  // hour 1 repeats them
    st.first_vacant = ~0u;
    s->state_h[i] = st;
    if (!host_init) {
      const InitTargets t{v.pos + D.agent_off, v.st + D.agent_off, v.tm + D.agent_off, v.wsa + D.agent_off,
                          v.occ + D.cell_off, D.pitch, padded_cells(P.geo.stride, P.geo.height)};
      std::string err;
      const hipError_t e = synthesize_on_device(s->cfg[i], s->seed[i], P.geo, t, s->stream, err);
      if (e != hipSuccess) return fail(EPI_ERR_DEVICE, "device population synthesis (" + err + "): " + hipGetErrorString(e));
      continue;
    }
    if (D.n) {
      HIP_TRY(hipMemcpy(v.pos + D.agent_off, P.pos.data(), nb, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(v.st + D.agent_off, P.st.data(), nb, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(v.tm + D.agent_off, P.tm.data(), nb, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(v.wsa + D.agent_off, P.wsa.data(), nb, hipMemcpyHostToDevice));
    }
    const size_t cells = padded_cells(P.geo.stride, P.geo.height);
    grid.assign(cells, 0);
    for (uint32_t a = 0; a < D.n; ++a) {
      const uint32_t p = P.pos[a];
      grid[static_cast<size_t>((p >> 16) + kPadY) * D.pitch + (p & 0xFFFFu) + kPadX] =
          occupant_byte(D.dz, P.st[a]);
    }
    HIP_TRY(hipMemcpy(v.occ + D.cell_off, grid.data(), cells, hipMemcpyHostToDevice));
  }
  HIP_TRY(hipMemcpy(v.state, s->state_h.data(), s->n_sims * sizeof(SimState), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d_hosp_blocks, s->hosp_blocks.data(), s->hosp_blocks.size() * sizeof(BlockRef),
                    hipMemcpyHostToDevice));
  if (!s->vacc_hour.empty()) {
    HIP_TRY(hipMemcpy(d_vacc_hour, s->vacc_hour.data(), s->vacc_hour.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_vacc_thr, s->vacc_thr.data(), s->vacc_thr.size() * sizeof(unsigned long long),
                      hipMemcpyHostToDevice));
  }
  HIP_TRY(launch_fill_u32(v.claim, s->n_cells, ~0u, s->stream));
  {  // the office-layout view shares everything with v but pos and st
    DevView o = v;
    o.pos = vo.pos;
    o.st = vo.st;
    o.agent_blocks = s->d_office_blocks;
    o.office_agent = s->d_office_agent;
    o.office_slot = s->d_office_slot;
    o.office = 1;
    v.office = 0;
    v.office_agent = nullptr;
    v.office_slot = nullptr;
    vo = o;
  }
  rc = build_office_order(s.get());
  if (rc != EPI_OK) return rc;
  rc = build_launch_tables(s.get());
  if (rc != EPI_OK) return rc;
  rc = upload_tables(s.get());
  if (rc != EPI_OK) return rc;
  HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&s->pinned_active), 2 * s->n_sims * sizeof(uint32_t),
                        hipHostMallocDefault));
  HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&s->pinned_state), s->n_sims * sizeof(SimState),
                        hipHostMallocDefault));
  HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&s->pinned_rows), std::max<uint64_t>(s->n_rows, 1) * sizeof(Counts7),
                        hipHostMallocDefault));
  HIP_TRY(hipHostGetDevicePointer(reinterpret_cast<void**>(&s->d_pinned_active), s->pinned_active, 0));
  HIP_TRY(hipHostGetDevicePointer(reinterpret_cast<void**>(&s->d_pinned_state), s->pinned_state, 0));
  HIP_TRY(hipHostGetDevicePointer(reinterpret_cast<void**>(&s->d_pinned_rows), s->pinned_rows, 0));
  // every hour's epilogue writes its row and the state after it here (both views)
  std::memcpy(s->pinned_state, s->state_h.data(), s->n_sims * sizeof(SimState));
  s->v.host_rows = s->v_office.host_rows = s->d_pinned_rows;
  s->v.host_state = s->v_office.host_state = s->d_pinned_state;
  HIP_TRY(hipEventCreateWithFlags(&s->poll_event[0], hipEventDisableTiming));
  HIP_TRY(hipEventCreate(&s->loop_event[0]));
  HIP_TRY(hipEventCreate(&s->loop_event[1]));
  HIP_TRY(hipEventCreateWithFlags(&s->poll_event[1], hipEventDisableTiming));
  HIP_TRY(hipStreamSynchronize(s->stream));
  *out = s.release();
  return EPI_OK;
}

std::string utc_stamp() {
  std::time_t t = std::time(nullptr);
  std::tm tm{};
  gmtime_r(&t, &tm);
  char buf[32];
  std::strftime(buf, sizeof buf, "%Y-%m-%dT%H:%M:%S", &tm);
  return buf;
}

// EPI_OUTPUT_DIR prefixing of a relative listener path (engine/src/utils/environment.rs:33-38)
std::string listener_path(const std::string& name) {
  if (!name.empty() && name[0] == '/') return name;
  const char* env = std::getenv("EPI_OUTPUT_DIR");
  std::string base = env ? env : ".";
  if (!base.empty() && base.back() != '/') base += '/';
  return base + name;
}

const char* event_name(uint32_t kind) {
  return kind == EV_VACCINATION ? "vaccination" : (kind == EV_HOSPITAL ? "build_new_hospital" : "lockdown");
}
const char* event_data(uint32_t kind) {
  if (kind == EV_LOCKED) return "{\"status\":\"locked_down\"}";
  if (kind == EV_UNLOCKED) return "{\"status\":\"lockdown_revoked\"}";
  return "{}";
}

std::string area_json(const Rect& r) {
  // Area {location_id, start_offset, end_offset} (geography/area.rs:30-35); location_id is
  // the standalone engine id "0" (engine_app.rs:30)
  char buf[160];
  std::snprintf(buf, sizeof buf,
                "{\"location_id\":\"0\",\"start_offset\":{\"x\":%d,\"y\":%d},\"end_offset\":{\"x\":%d,\"y\":%d}}",
                r.x0, r.y0, r.x1, r.y1);
  return buf;
}

// The Grid message published once before the first hour (grid_updated,
// events_kafka_producer.rs:101-115): serde field order of Grid (geography/grid.rs:35-47),
// houses and offices in area_factory order (geography/area.rs:95-117).
int write_grid_line(StateSink& k, const Geometry& G, uint32_t hosp_version) {
  std::string out;
  out.reserve(96u * (G.n_houses + G.n_offices) + 1024u);
  out += "{\"grid_size\":" + std::to_string(G.g);
  out += ",\"housing_area\":" + area_json(G.housing);
  out += ",\"work_area\":" + area_json(G.work);
  out += ",\"transport_area\":" + area_json(G.transport);
  out += ",\"hospital_area\":" + area_json(hosp_version ? G.hosp1 : G.hosp0);
  out += ",\"houses\":[";
  for (uint32_t h = 0; h < G.n_houses; ++h) {
    if (h) out += ',';
    out += area_json(building_rect(G.house_sx, G.house_sy, G.house_nx, 2, h));
  }
  out += "],\"offices\":[";
  for (uint32_t o = 0; o < G.n_offices; ++o) {
    if (o) out += ',';
    out += area_json(building_rect(G.office_sx, G.office_sy, G.office_nx, 10, o));
  }
  out += "]}\n";
  if (std::fwrite(out.data(), 1, out.size(), k.f) != out.size())
    return fail(EPI_ERR_IO, "Failed to write citizen states to " + k.path);
  return EPI_OK;
}

int open_sink(epi_sim* s, const std::string& path) {
  if (s->n_sims != 1) return fail(EPI_ERR_ARG, "the citizen-state stream needs a single-simulation handle");
  int rc = fetch_states(s);
  if (rc != EPI_OK) return rc;
  std::unique_ptr<StateSink> k(new StateSink());
  k->path = path;
  const SimDesc& D = s->desc[0];
  k->cap = static_cast<uint64_t>(D.n) * max_record_bytes(D.n, D.stride, D.height);
  HIP_TRY(hipMalloc(reinterpret_cast<void**>(&k->d_text), std::max<uint64_t>(k->cap, 1)));
  HIP_TRY(hipMalloc(reinterpret_cast<void**>(&k->d_block_off),
                    std::max<uint64_t>(stream_blocks(D.n), 1) * sizeof(unsigned long long)));
  HIP_TRY(hipMalloc(reinterpret_cast<void**>(&k->d_total), sizeof(unsigned long long)));
  for (char*& p : k->h_text) HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&p), std::max<uint64_t>(k->cap, 1)));
  HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&k->h_total), 2 * sizeof(unsigned long long)));
  HIP_TRY(hipEventCreateWithFlags(&k->ev_total, hipEventDisableTiming));
  for (hipEvent_t& e : k->ev_text) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  k->f = std::fopen(path.c_str(), "w");
  if (!k->f) return fail(EPI_ERR_IO, "Failed to create citizen state file: " + path);
  rc = write_grid_line(*k, s->geo[0], s->state_h[0].hosp_version);
  if (rc != EPI_OK) return rc;
  s->sink = std::move(k);
  return EPI_OK;
}

// Closes the stream with the reference's end message (events_kafka_producer.rs:75-87).
int close_sink(epi_sim* s) {
  if (!s->sink) return EPI_OK;
  StateSink& k = *s->sink;
  const bool ok = std::fputs("{\"simulation_ended\": true}\n", k.f) >= 0 && std::fclose(k.f) == 0;
  k.f = nullptr;
  const std::string path = k.path;
  s->sink.reset();
  if (!ok) return fail(EPI_ERR_IO, "Failed to write citizen states to " + path);
  return EPI_OK;
}

}  // namespace

extern "C" {

int epi_stream_citizen_states(epi_sim* sim, const char* path) {
  if (!sim) return fail(EPI_ERR_ARG, "sim is null");
  int rc = close_sink(sim);
  if (rc != EPI_OK || !path) return rc;
  return open_sink(sim, path);
}

int epi_citizen_states_text_noalloc(epi_sim* sim, char* d_text, uint64_t cap, uint64_t* d_len) {
  if (!sim || !d_text || !d_len) return fail(EPI_ERR_ARG, "sim, d_text or d_len is null");
  if (sim->n_sims != 1) return fail(EPI_ERR_ARG, "the citizen-state text needs a single-simulation handle");
  const SimDesc& D = sim->desc[0];
  if (D.n > kNoAllocMaxAgents)
    return fail(EPI_ERR_ARG, "epi_citizen_states_text_noalloc handles at most " + std::to_string(kNoAllocMaxAgents) +
                                 " agents (its scratch array is fixed); this simulation has " + std::to_string(D.n));
  const uint64_t need = static_cast<uint64_t>(D.n) * max_record_bytes(D.n, D.stride, D.height);
  if (cap < need)
    return fail(EPI_ERR_ARG, "d_text holds " + std::to_string(cap) + " bytes; the records may need " +
                                 std::to_string(need));
  unsigned long long* scratch = nullptr;
  HIP_TRY(noalloc_scratch(&scratch));
  int rc = settle(sim);  // the home layout holds the current state
  if (rc != EPI_OK) return rc;
  rc = fetch_states(sim);
  if (rc != EPI_OK) return rc;
  unsigned long long* len = reinterpret_cast<unsigned long long*>(d_len);
  if (D.n == 0) {
    HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t*>(len), 2, 0u, sim->stream));
  } else {
    const StreamArgs a{sim->v.pos + D.agent_off, sim->v.st + D.agent_off, D.n, sim->v.state,
                       sim->state_h[0].hour, scratch, len, d_text};
    HIP_TRY(launch_citizen_states(a, sim->stream));
  }
  HIP_TRY(hipStreamSynchronize(sim->stream));
  return EPI_OK;
}

int epi_citizen_states(epi_sim* sim, uint32_t index, epi_citizen_state* out, uint32_t cap, uint32_t* n,
                       uint32_t* hour) {
  if (!sim || index >= sim->n_sims) return fail(EPI_ERR_ARG, "bad handle or index");
  int rc = settle(sim);
  if (rc != EPI_OK) return rc;
  rc = fetch_states(sim);
  if (rc != EPI_OK) return rc;
  const SimDesc& D = sim->desc[index];
  if (n) *n = D.n;
  if (hour) *hour = sim->state_h[index].hour;
  const uint32_t k = std::min(cap, D.n);
  if (!out || k == 0) return EPI_OK;
  std::vector<uint32_t> pos(k), st(k);
  HIP_TRY(hipMemcpy(pos.data(), sim->v.pos + D.agent_off, k * 4ull, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(st.data(), sim->v.st + D.agent_off, k * 4ull, hipMemcpyDeviceToHost));
  static const char letters[5] = {'s', 'e', 'i', 'r', 'd'};
  for (uint32_t i = 0; i < k; ++i) {
    out[i].citizen_id = i;
    out[i].x = static_cast<int32_t>(pos[i] & 0xFFFFu);
    out[i].y = static_cast<int32_t>(pos[i] >> 16);
    out[i].state = letters[std::min<uint32_t>(st_state(st[i]), 4u)];
  }
  return EPI_OK;
}

const char* epi_last_error(void) { return g_last_error.c_str(); }

int epi_create(const char* config_json, uint64_t seed, uint32_t flags, epi_sim** out) {
  if (!config_json) return fail(EPI_ERR_ARG, "config_json is null");
  g_last_error.clear();
  return create_impl({std::string(config_json)}, {seed}, flags, out);
}

int epi_batch_create(const char* const* configs, const uint64_t* seeds, uint32_t n, uint32_t flags, epi_sim** out) {
  if (!configs || !seeds || n == 0) return fail(EPI_ERR_ARG, "configs/seeds null or n == 0");
  g_last_error.clear();
  std::vector<std::string> texts;
  for (uint32_t i = 0; i < n; ++i) {
    if (!configs[i]) return fail(EPI_ERR_ARG, "configs[" + std::to_string(i) + "] is null");
    texts.emplace_back(configs[i]);
  }
  return create_impl(texts, std::vector<uint64_t>(seeds, seeds + n), flags, out);
}

void epi_destroy(epi_sim* sim) { delete sim; }

int epi_step(epi_sim* sim, uint32_t n_hours, epi_counts* rows, uint32_t* n_written, int* stopped) {
  if (!sim) return fail(EPI_ERR_ARG, "sim is null");
  if (sim->n_sims != 1) return fail(EPI_ERR_ARG, "batch handle: use epi_batch_step");
  return step_impl(sim, n_hours, rows, n_hours, n_written, stopped);
}

int epi_batch_step(epi_sim* sim, uint32_t n_hours, epi_counts* rows, uint32_t* n_written, int* stopped) {
  if (!sim) return fail(EPI_ERR_ARG, "sim is null");
  return step_impl(sim, n_hours, rows, n_hours, n_written, stopped);
}

int epi_info(epi_sim* sim, uint32_t index, uint32_t* n_sims, uint32_t* n_agents, uint32_t* grid_size,
             uint32_t* hours) {
  if (!sim || index >= sim->n_sims) return fail(EPI_ERR_ARG, "bad handle or index");
  if (n_sims) *n_sims = sim->n_sims;
  if (n_agents) *n_agents = sim->desc[index].n;
  if (grid_size) *grid_size = sim->desc[index].g;
  if (hours) *hours = sim->desc[index].hours;
  return EPI_OK;
}

int epi_geometry(epi_sim* sim, uint32_t index, int32_t rects[20], uint32_t* hospital_version) {
  if (!sim || index >= sim->n_sims || !rects) return fail(EPI_ERR_ARG, "bad handle, index or rects");
  const Geometry& G = sim->geo[index];
  const Rect rs[5] = {G.housing, G.transport, G.work, G.hosp0, G.hosp1};
  for (int k = 0; k < 5; ++k) {
    rects[4 * k] = rs[k].x0;
    rects[4 * k + 1] = rs[k].y0;
    rects[4 * k + 2] = rs[k].x1;
    rects[4 * k + 3] = rs[k].y1;
  }
  if (hospital_version) {
    int rc = fetch_states(sim);
    if (rc != EPI_OK) return rc;
    *hospital_version = sim->state_h[index].hosp_version;
  }
  return EPI_OK;
}

int epi_current_counts(epi_sim* sim, uint32_t index, epi_counts* out) {
  if (!sim || index >= sim->n_sims || !out) return fail(EPI_ERR_ARG, "bad handle, index or out");
  int rc = fetch_states(sim);
  if (rc != EPI_OK) return rc;
  const SimState& st = sim->state_h[index];
  if (st.rows_written == 0) {
    const uint32_t* c = &sim->counts0[7 * index];
    *out = epi_counts{c[0], c[1], c[2], c[3], c[4], c[5], c[6]};
    return EPI_OK;
  }
  Counts7 row;
  HIP_TRY(hipMemcpy(&row, sim->v.rows + sim->desc[index].row_off + st.rows_written - 1, sizeof row,
                    hipMemcpyDeviceToHost));
  *out = epi_counts{row.hour, row.s, row.e, row.i, row.h, row.r, row.d};
  return EPI_OK;
}

int epi_events(epi_sim* sim, uint32_t index, epi_event* out, uint32_t cap, uint32_t* n) {
  if (!sim || index >= sim->n_sims) return fail(EPI_ERR_ARG, "bad handle or index");
  int rc = fetch_states(sim);
  if (rc != EPI_OK) return rc;
  const SimDesc& D = sim->desc[index];
  const uint32_t total = sim->state_h[index].n_events;
  const uint32_t avail = std::min(total, D.event_cap);
  if (total > D.event_cap) return fail(EPI_ERR_DEVICE, "event log overflow");
  if (n) *n = total;
  const uint32_t k = std::min(avail, cap);
  if (out && k) HIP_TRY(hipMemcpy(out, sim->v.events + D.event_off, k * sizeof(Event), hipMemcpyDeviceToHost));
  return EPI_OK;
}

int epi_read_agents(epi_sim* sim, uint32_t index, epi_agents_soa* dst) {
  if (!sim || index >= sim->n_sims || !dst) return fail(EPI_ERR_ARG, "bad handle, index or dst");
  const SimDesc& D = sim->desc[index];
  const Geometry& G = sim->geo[index];
  const uint32_t n = D.n;
  int rc = settle(sim);
  if (rc != EPI_OK) return rc;
  std::vector<uint32_t> pos(n), st(n), tm(n), wsa(n);
  if (n) {
    HIP_TRY(hipMemcpy(pos.data(), sim->v.pos + D.agent_off, n * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(st.data(), sim->v.st + D.agent_off, n * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(tm.data(), sim->v.tm + D.agent_off, n * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(wsa.data(), sim->v.wsa + D.agent_off, n * 4, hipMemcpyDeviceToHost));
  }
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t w = st[i];
    const uint32_t state = st_state(w), sev = st_sev(w), ws = st_ws(w);
    if (dst->x) dst->x[i] = static_cast<int32_t>(pos[i] & 0xFFFFu);
    if (dst->y) dst->y[i] = static_cast<int32_t>(pos[i] >> 16);
    if (dst->state) dst->state[i] = static_cast<uint8_t>(state);
    if (dst->severity) dst->severity[i] = static_cast<uint8_t>(state == ST_I ? sev : 0);
    if (dst->infection_day) dst->infection_day[i] = state == ST_I ? st_day(w) : 0;
    if (dst->at_hour) dst->at_hour[i] = Disease::timed(w) ? tm[i] : 0;
    if (dst->flags)
      dst->flags[i] = static_cast<uint8_t>(((w & F_HOSP) ? 1 : 0) | ((w & F_VACC) ? 2 : 0) | ((w & F_ISO) ? 4 : 0) |
                                           ((w & F_WQ) ? 8 : 0) | ((w & F_PT) ? 16 : 0));
    if (dst->work_status) dst->work_status[i] = static_cast<uint8_t>(ws);
    if (dst->work_start_at) dst->work_start_at[i] = ws == WS_STAFF ? wsa[i] : 0;
    if (dst->immunity) dst->immunity[i] = static_cast<int8_t>(st_imm(w));
    if (dst->current_area) {
      const Rect home = building_rect(G.house_sx, G.house_sy, std::max(G.house_nx, 1), 2, i % std::max(G.n_houses, 1u));
      Rect r{};
      switch (st_area(w)) {
        case AREA_HOME: r = home; break;
        case AREA_WORK:
          r = ws == WS_NA ? home
                          : building_rect(G.office_sx, G.office_sy, std::max(G.office_nx, 1), 10,
                                          i % std::max(G.n_offices, 1u));
          break;
        case AREA_TRANSPORT: r = G.transport; break;
        case AREA_HOUSING: r = G.housing; break;
        case AREA_HOSP0: r = G.hosp0; break;
        default: r = G.hosp1; break;
      }
      dst->current_area[4 * i] = r.x0;
      dst->current_area[4 * i + 1] = r.y0;
      dst->current_area[4 * i + 2] = r.x1;
      dst->current_area[4 * i + 3] = r.y1;
    }
  }
  return EPI_OK;
}

int epi_run(epi_sim* sim, const char* output_dir, char* final_path, size_t cap) {
  if (!sim || !output_dir) return fail(EPI_ERR_ARG, "sim or output_dir is null");
  if (sim->n_sims != 1) return fail(EPI_ERR_ARG, "epi_run drives a single simulation");
  // create_out_dir_if_not_present (engine/src/utils/util.rs:28-35): errors ignored
  std::string dir(output_dir);
  std::string out_dir = dir + (dir.empty() || dir.back() == '/' ? "" : "/") + "output";
  (void)::mkdir(out_dir.c_str(), 0777);
  const std::string base = out_dir + "/simulation_0_" + utc_stamp();
  // citizen-state messages go to a JSON-lines file in place of the Kafka topic
  if (sim->cfg[0].enable_citizen_state_messages) {
    int rc = epi_stream_citizen_states(sim, listener_path(base + "_citizen_states.jsonl").c_str());
    if (rc != EPI_OK) return rc;
  }
  std::vector<epi_counts> rows;
  const uint32_t chunk = 24u * 50u;
  std::vector<epi_counts> buf(chunk);
  for (;;) {
    uint32_t nw = 0;
    int stopped = 0;
    int rc = step_impl(sim, chunk, buf.data(), chunk, &nw, &stopped);
    if (rc != EPI_OK) return rc;
    rows.insert(rows.end(), buf.begin(), buf.begin() + nw);
    if (stopped || nw == 0) break;
  }
  if (sim->sink) {
    int rc = close_sink(sim);
    if (rc != EPI_OK) return rc;
  }
  // CSV (listeners/csv_service.rs:44-71): serde field order of Counts, one row per hour
  const std::string csv_path = listener_path(base + ".csv");
  {
    std::ofstream f(csv_path);
    if (!f) return fail(EPI_ERR_IO, "Failed to write to file: " + csv_path);
    f << "hour,susceptible,exposed,infected,hospitalized,recovered,deceased\n";
    for (const auto& r : rows)
      f << r.hour << ',' << r.susceptible << ',' << r.exposed << ',' << r.infected << ',' << r.hospitalized << ','
        << r.recovered << ',' << r.deceased << '\n';
    if (!f) return fail(EPI_ERR_IO, "Failed to write to file: " + csv_path);
  }
  // interventions report (listeners/intervention_reporter.rs:53-58): compact serde_json
  uint32_t n_ev = 0;
  int rc = epi_events(sim, 0, nullptr, 0, &n_ev);
  if (rc != EPI_OK) return rc;
  std::vector<epi_event> ev(n_ev);
  if (n_ev) {
    rc = epi_events(sim, 0, ev.data(), n_ev, &n_ev);
    if (rc != EPI_OK) return rc;
  }
  const std::string json_path = listener_path(base + "_interventions.json");
  {
    std::ofstream f(json_path);
    if (!f) return fail(EPI_ERR_IO, "Failed to create intervention report file: " + json_path);
    f << '[';
    for (size_t k = 0; k < ev.size(); ++k) {
      if (k) f << ',';
      f << "{\"hour\":" << ev[k].hour << ",\"intervention\":\"" << event_name(ev[k].kind) << "\",\"data\":"
        << event_data(ev[k].kind) << '}';
    }
    f << ']';
  }
  if (final_path && cap) std::snprintf(final_path, cap, "%s", csv_path.c_str());
  return EPI_OK;
}

int epi_profile(epi_sim* sim, uint32_t n_hours, double ms_mean[3], uint32_t launches[3], uint32_t* hours_run) {
  if (!sim || !ms_mean || !launches) return fail(EPI_ERR_ARG, "null argument");
  if (sim->any_empty) return fail(EPI_ERR_PANIC, "No citizens!");
  const uint32_t target = std::min<uint64_t>(static_cast<uint64_t>(sim->host_hour) + n_hours, sim->max_hour);
  double sum[3] = {0, 0, 0};
  uint32_t cnt[3] = {0, 0, 0};
  const uint32_t per_batch = 24;
  std::vector<hipEvent_t> ev(4 * per_batch, nullptr);
  for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
  uint32_t ran = 0;
  int rc = EPI_OK;
  while (sim->host_hour < target && rc == EPI_OK) {
    const uint32_t nb = std::min(per_batch, target - sim->host_hour);
    std::vector<int> had_scan(nb, 0), had_step(nb, 0);
    for (uint32_t k = 0; k < nb && rc == EPI_OK; ++k) {
      const uint32_t h = sim->host_hour + 1;
      hipEvent_t* e = &ev[4 * k];
      (void)hipEventRecord(e[0], sim->stream);
      if (h % 24u == 0) {
        had_scan[k] = 1;
        if (launch_hospital_scan(sim->v, sim->shape, sim->stream) != hipSuccess) rc = fail(EPI_ERR_DEVICE, "launch");
      }
      (void)hipEventRecord(e[1], sim->stream);
      had_step[k] = hour_has_step(h) ? 1 : 0;
      // the layout moves of hours of day 8 and 16 are timed with the step
      if ((h % 24u == 8 || h % 24u == 16) && rc == EPI_OK) rc = relayout(sim, h % 24u == 8);
      const DevView& v = office_hour(h) ? sim->v_office : sim->v;
      if (had_step[k] && launch_step(v, sim->shape, sim->stream) != hipSuccess) rc = fail(EPI_ERR_DEVICE, "launch");
      (void)hipEventRecord(e[2], sim->stream);
      if (launch_commit(v, sim->shape, h, sim->stream) != hipSuccess) rc = fail(EPI_ERR_DEVICE, "launch");
      (void)hipEventRecord(e[3], sim->stream);
      ++sim->host_hour;
    }
    if (hipStreamSynchronize(sim->stream) != hipSuccess) rc = fail(EPI_ERR_DEVICE, "hipStreamSynchronize");
    for (uint32_t k = 0; k < nb && rc == EPI_OK; ++k) {
      float t[3] = {0, 0, 0};
      hipEvent_t* e = &ev[4 * k];
      (void)hipEventElapsedTime(&t[0], e[0], e[1]);
      (void)hipEventElapsedTime(&t[1], e[1], e[2]);
      (void)hipEventElapsedTime(&t[2], e[2], e[3]);
      if (had_scan[k]) { sum[0] += t[0]; ++cnt[0]; }
      if (had_step[k]) { sum[1] += t[1]; ++cnt[1]; }
      sum[2] += t[2]; ++cnt[2];
    }
    ran += nb;
  }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (rc != EPI_OK) return rc;
  for (int k = 0; k < 3; ++k) {
    ms_mean[k] = cnt[k] ? sum[k] / cnt[k] : 0.0;
    launches[k] = cnt[k];
  }
  if (hours_run) *hours_run = ran;
  if (ran) sim->settled = false;
  rc = settle(sim);
  if (rc != EPI_OK) return rc;
  // keep row accounting consistent for later epi_step calls
  rc = fetch_states(sim);
  if (rc != EPI_OK) return rc;
  for (uint32_t i = 0; i < sim->n_sims; ++i) sim->rows_returned[i] = sim->state_h[i].rows_written;
  return EPI_OK;
}

int epi_last_run_ms(epi_sim* sim, double* device_ms, uint32_t* hours) {
  if (!sim || !device_ms || !hours) return fail(EPI_ERR_ARG, "null argument");
  *device_ms = sim->last_loop_ms;
  *hours = sim->last_loop_hours;
  return EPI_OK;
}

static int set_launch_blocks(epi_sim* sim, uint32_t home, uint32_t office) {
  if (!sim) return fail(EPI_ERR_ARG, "sim is null");
  if (office < 64 || office > 1024 || office % 64 || home < 64 || home > 1024 || home % 64)
    return fail(EPI_ERR_ARG, "block size must be 64..1024, x64");
  HIP_TRY(hipStreamSynchronize(sim->stream));
  sim->shape.block = home;
  sim->shape.office_block = office;
  int rc = build_launch_tables(sim);
  if (rc != EPI_OK) return rc;
  rc = upload_tables(sim);
  if (rc != EPI_OK) return rc;
  if (sim->day_graph) {
    (void)hipGraphExecDestroy(sim->day_graph);
    sim->day_graph = nullptr;
  }
  return EPI_OK;
}

int epi_set_block_size(epi_sim* sim, uint32_t threads) { return set_launch_blocks(sim, threads, threads); }

int epi_set_office_block_size(epi_sim* sim, uint32_t threads) {
  if (!sim) return fail(EPI_ERR_ARG, "sim is null");
  return set_launch_blocks(sim, sim->shape.block, threads);
}

}  // extern "C"
