// Agent state layout in HBM and the per-simulation descriptors shared by the host
// runtime and the gfx950 kernels.
//
// Agent state is struct-of-arrays, one 32-bit word per field so a wavefront's 64
// lanes read 256 contiguous bytes per field:
//   pos[i]  x | y << 16                       (grid_size <= 32000)
//   st[i]   packed state word (bit layout below)
//   tm[i]   Exposed{at_hour} / Pre{at_hour}
//   wsa[i]  HospitalStaff{work_start_at}      (read only for the 0.2% staff)
//   tgt[i]  the cell chosen this hour          (step -> commit hand-off)
// The grid is a dense byte array: 0 vacant, 1 occupied by an agent that cannot infect,
// 2 / 3 occupied by an infectious agent in the regular / high transmission window. Cell
// (x, y) lives at (y + kPadY) * pitch + x + kPadX: a zero border of kPadX columns / kPadY
// rows (and pitch >= stride + 9, a multiple of 4) lets every kernel read the 5x5
// neighbourhood of any cell as ten aligned dwords without bounds checks. Claims are a
// parallel u32 array with the same indexing: a claim key is the rank of the claim hour
// within the day (hours of day 7..23 then 0, later hours smaller) in the top kClaimHourBits
// bits over the agent id, and the array is reset at hour of day 2, before the day's first
// claim hour, so a stale claim never ranks with the current hour's.
#pragma once
#include <cstdint>

#include "fastdiv.hpp"

namespace epi {

// Hours of day whose step and commit run on the office layout of the agent arrays.
inline constexpr bool office_hour(uint32_t hour) { return hour % 24u >= 8u && hour % 24u <= 15u; }

// ---- state word
enum : uint32_t {
  ST_S = 0, ST_E = 1, ST_I = 2, ST_R = 3, ST_D = 4,
  SEV_PRE = 0, SEV_ASYM = 1, SEV_MILD = 2, SEV_SEVERE = 3,
  WS_NORMAL = 0, WS_ESSENTIAL = 1, WS_STAFF = 2, WS_NA = 3,
  AREA_HOME = 0, AREA_WORK = 1, AREA_TRANSPORT = 2, AREA_HOUSING = 3, AREA_HOSP0 = 4, AREA_HOSP1 = 5,
};

constexpr uint32_t B_STATE = 0, M_STATE = 0x7u;
constexpr uint32_t B_SEV = 3, M_SEV = 0x3u;
constexpr uint32_t F_HOSP = 1u << 5;
constexpr uint32_t F_VACC = 1u << 6;
constexpr uint32_t F_ISO = 1u << 7;
constexpr uint32_t F_WQ = 1u << 8;
constexpr uint32_t B_WS = 9, M_WS = 0x3u;
constexpr uint32_t B_AREA = 11, M_AREA = 0x7u;
constexpr uint32_t F_PT = 1u << 14;
// bit 15 is unused
constexpr uint32_t F_DIRTY = 1u << 16;  // transient: the agent's grid byte changes this hour
constexpr uint32_t B_IMM = 17, M_IMM = 0x7u;  // immunity + 2
constexpr uint32_t B_DAY = 20, M_DAY = 0xFFFu;
constexpr uint32_t kMaxDay = M_DAY;

// ---- Philox sites (counter word 2); the hour word is 0xFFFFFFFF during population synthesis
constexpr uint32_t kInitHour = 0xFFFFFFFFu;
enum : uint32_t {
  SITE_MOVE = 1,       // w0 x, w1 y of the hour's one random area point; w2 neighbour choice
  SITE_DISEASE = 2,    // w0 exposed random factor; (w2,w3) the hour's disease Bernoulli
  SITE_INFECT = 3,     // blocks 0..3: Moore neighbour k uses block k>>1, words 2(k&1), 2(k&1)+1
  SITE_VACCINATE = 4,  // (w0,w1)
  SITE_INIT_A = 16,    // (w0,w1) working, (w2,w3) public transport
  SITE_INIT_B = 17,    // (w0,w1) hospital staff, w2 immunity
  SITE_INIT_C = 18,    // (w0,w1) essential worker, (w2,w3) starting-infection key
  SITE_INIT_HOUSE = 19 // counter id = house index; w0,w1 x permutation, w2,w3 y permutation
};

// ---- sweeps decided after an hour's counts, applied by the next hour's agent pass
enum : uint32_t { SWEEP_VACC = 1, SWEEP_LOCK = 2, SWEEP_UNLOCK = 4 };
// ---- intervention event kinds (intervention_reporter.rs names)
enum : uint32_t { EV_VACCINATION = 0, EV_LOCKED = 1, EV_UNLOCKED = 2, EV_HOSPITAL = 3 };

struct Rect {
  int32_t x0, y0, x1, y1;
};

constexpr uint32_t kPadX = 4, kPadY = 2;
constexpr uint32_t kClaimHourBits = 5, kClaimIdBits = 32 - kClaimHourBits;
constexpr uint32_t kMaxAgents = (1u << kClaimIdBits) - 1u;  // agent ids fit the claim key
inline uint32_t padded_pitch(uint32_t stride) { return (stride + 9u + 3u) & ~3u; }
inline uint64_t padded_cells(uint32_t stride, uint32_t height) {
  return static_cast<uint64_t>(padded_pitch(stride)) * (height + 2u * kPadY);
}

// The constants of the reference's Disease (common/src/disease/mod.rs:26-44) in the form the
// step uses: window bounds in days, durations in hours, rates as Bernoulli thresholds, and
// bit c of hosp_window set when is_to_be_hospitalized holds in transmission window c. A
// disease handler (disease.hpp) fills them from the config and is their only reader.
struct DiseaseParams {
  uint32_t regular_start, high_start, last_day;
  uint32_t exposed_duration, pre_symptomatic_duration;
  uint64_t thr_regular, thr_high, thr_death, thr_symptoms, thr_severe;
  uint32_t hosp_window;
};

// Read-only per-simulation descriptor (device copy built once at create).
struct SimDesc {
  uint32_t n;          // agents
  uint32_t g;          // grid_size
  uint32_t stride;     // logical grid width (>= g+1)
  uint32_t pitch;      // padded row pitch of the grid / claim arrays
  uint32_t height;     // logical grid height (g+1)
  uint32_t hours;      // config hours: the loop runs 1..hours-1
  uint64_t agent_off;  // first agent in the concatenated SoA
  uint64_t cell_off;   // first cell in the concatenated grid / claim arrays
  uint64_t row_off;    // first row in the concatenated per-hour counts
  uint64_t event_off;  // first slot in the concatenated event log
  uint32_t event_cap;
  uint32_t seed_lo, seed_hi;
  Rect housing, transport, work, hosp[2];
  int32_t house_sx, house_sy, house_nx;
  uint32_t n_houses;
  int32_t office_sx, office_sy, office_nx;
  uint32_t n_offices;
  FastDiv fd_houses, fd_house_nx, fd_offices, fd_office_nx;  // id -> building index -> (column, row)
  DiseaseParams dz;  // the compiled-in disease handler's constants (disease.hpp)
  // interventions
  uint32_t has_lockdown, lockdown_at;
  uint32_t has_hospital, spread_rate_threshold;
  uint32_t vacc_off, n_vacc;  // sorted, last-entry-wins schedule in the shared vacc arrays
  // Agent slots: n_slots = n from agent_off in both layouts of the agent arrays. Home
  // layout: slot = id. Office layout (hours of day 8..15): first the workers (Normal and
  // Essential) of office 0, 1, ... (an agent's office is id % n_offices), each office's in
  // id order, then every other agent (not working, hospital staff) in id order; the
  // office_agent array maps slot -> id. Workers spend these hours in their office, the
  // others near their house, so both groups keep neighbours in neighbouring slots.
  uint32_t n_slots;
  uint32_t n_workers;         // office layout: slots 0 .. n_workers-1 hold the office workers
  uint32_t n_blocks;          // agent blocks of this simulation (step / commit), over n_slots
  uint32_t block_begin;       // index of its first agent block
  uint32_t office_n_blocks, office_block_begin;  // the same in the office layout's table
};

// Mutable per-simulation scalars, owned by the device.
struct SimState {
  uint32_t hour;            // last hour stepped
  uint32_t active;          // 0 once stopped (E=I=H=0) or out of hours
  uint32_t commit_hour[2];  // [hour & 1]: the hour the agent pass of that parity runs, 0 = none
  // [hour & 1]: the sweep decided at that hour (SWEEP_*), applied by the next hour's agent
  // pass (idempotent), with the hour and threshold of its vaccination draw
  uint32_t sweep[2];
  uint32_t sweep_vacc_hour[2];
  uint64_t sweep_vacc_thr[2];
  uint32_t first_vacant;    // row-major index in the hospital, or ~0u
  uint32_t hosp_version;    // 0 after resize_hospital, 1 after increase_hospital_size
  uint32_t locked, zero_infection_hour;
  uint32_t hosp_new_infections, hosp_applied;
  uint32_t vacc_next;       // first schedule entry not yet passed
  uint32_t n_events;
  uint32_t last_counts[6];  // S E I H R D of the last row (the population's before hour 1)
  uint32_t rows_written;
  uint32_t error;
  // The sweep decided at the last hour of day 0, which no other hour overwrites: the night
  // launch (hours of day 1..6 together) applies it while its epilogues refill the slots.
  uint32_t night_sweep, night_vacc_hour;
  uint64_t night_vacc_thr;
  uint32_t pad[2];
};

// Launch-table entry: (simulation, first hospital cell).
struct BlockRef {
  uint32_t sim, first;
};

// Most offices one office-layout block holds, so that its claim filter (100 cells per
// 10 x 10 office) fits its LDS array.
constexpr uint32_t kMaxFilterOffices = 20;

// Agent-block table entry: slots first .. first + count - 1 of simulation sim. In the
// office layout a block holds either workers of offices j_lo .. j_hi (every worker of
// offices full_lo .. full_hi among them) or only agents that are not office workers
// (j_lo > j_hi); in the home layout j_lo > j_hi too.
struct AgentBlock {
  uint32_t sim, first, count;
  uint32_t j_lo, j_hi, full_lo, full_hi;
  uint32_t pad;
};

struct Counts7 {
  uint32_t hour, s, e, i, h, r, d;
};

struct Event {
  uint32_t hour, kind;
};

}  // namespace epi
