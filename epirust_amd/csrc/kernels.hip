// Hour-step kernels for gfx950.
//
// One hour of EpiRust's standalone engine (engine/src/allocation_map.rs:67-134 and
// :306-337, engine/src/citizen/mod.rs:227-454, default_disease_handler.rs:31-103) is
//   [hour-of-day 0 only] k_hospital_scan : first vacant hospital cell of the snapshot
//   k_step   : [hours of day 0, 7..23] per agent, the previous hour's vaccination /
//              lockdown sweep, then routine + disease update against the start-of-hour
//              grid; state written in place (it is read by no other agent) when it
//              changes, the chosen cell written to the target word and claimed with a
//              32-bit atomicMin keyed (claim hour, agent id) unless the office filter or
//              housemate arbitration settles it; S/E/I/H/R/D counted with wavefront
//              ballots into per-block partials.
//   k_commit : every hour. Its first blocks, one per simulation, run the hour's epilogue:
//              reduce the partials, write the row, run the intervention triggers and the
//              stop rule on device; the row and the state go to device memory and to the
//              host's pinned copies, which a step call reads after its one drain. The other blocks run the hour's agent pass, which
//              reads nothing the same hour's epilogue writes (its hour and its sweep come
//              from parity slots the previous epilogue filled), so the two share a launch.
//              At hours of day 0 and 7..23 the agent pass commits the moves (the claim
//              winner moves, grid bytes are updated). Hours of day 1..6 move nobody and
//              change no state but flags and areas (citizen/mod.rs:244-248): the epilogue
//              repeats the last counts and the agent pass applies the pending sweep, sets
//              non-staff areas to home at hour of day 1 and resets the claim array at 2.
//   k_relayout: [hours of day 8 and 16] moves st and pos between the home layout
//              (slot = agent id) and the office layout that hours of day 8..15 run on:
//              each office's workers in consecutive slots, packed whole into blocks,
//              then everyone else in id order. There k_step resolves the claims of each
//              office's cells among the block's workers in LDS, and the winners on
//              offices that lie wholly in one block (all of them when an office fits a
//              block) claim nothing globally.
//   k_night / k_quiet_hours: in day graphs, hours of day 1..6 (or 3..6) as one launch
//              when no sweep can be pending inside them (no vaccination there).
// A sweep decided at hour h (allocation_map.rs:306-337) is applied by the agent pass of
// hour h+1 (before the routine, in k_step), which is what the reference's next hour sees;
// sweeps are idempotent, so k_apply_sweep re-applies the last one before the host reads
// agents.
// Every kernel is batched: blocks map to (simulation, agent range) through a table, so
// many independent replicates share one launch. Nothing here uses MFMA: the step is
// integer work, priced against HBM bandwidth (its per-agent bytes are few; at 1M agents
// the grid stays in the caches and the step is bound by VALU issue and memory latency).
#include <hip/hip_runtime.h>

#include <cstddef>

#include "disease.hpp"
#include "kernels.hpp"
#include "philox.hpp"

namespace epi {
namespace {

constexpr int kWave = 64;
// Office-hour block claim filter: 100 cells per 10 x 10 office, for the at most
// kMaxFilterOffices offices of a block.
constexpr uint32_t kOfficeCells = 100u, kFilterCells = kMaxFilterOffices * kOfficeCells, kNoCell = ~0u;
// Target word flags (x and y never exceed 32000, so bits 15 and 31 are free):
//   kTgtWon   the move has already won (office-hour block claim filter);
//   kTgtHouse a home-layout move into a cell of the agent's own house, arbitrated in the
//             commit among the house's agents instead of by a global claim.
// Every agent writes its target word at every step hour, movers or not, so the commit
// tells movers by it (no flag in the state word, which then changes, and is written, only
// when the agent's state does) and the words housemates read are this hour's.
constexpr uint32_t kTgtWon = 1u << 31;
constexpr uint32_t kTgtHouse = 1u << 15;
constexpr uint32_t kTgtCell = ~(kTgtWon | kTgtHouse);
// A non-mover: its cell part (x = y = 32767) is no cell, so it equals no move's word.
constexpr uint32_t kTgtNone = ~0u;

__device__ __forceinline__ uint32_t first_lane(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Agent id of lane `lane` of a block: the slot itself in the home layout, the office
// layout's slot -> id map in the office layout; ~0u past the block's slots.
__device__ __forceinline__ uint32_t agent_of_lane(const DevView& v, const SimDesc& D, const AgentBlock& bm,
                                                  uint32_t lane) {
  const uint32_t slot = first_lane(bm.first) + lane;
  if (lane >= first_lane(bm.count)) return ~0u;
  return v.office ? v.office_agent[D.agent_off + slot] : slot;
}


// Makes the values pass through registers at this point: the loads producing them are
// issued before it and waited for once, together. Without it the compiler sinks each load
// to its first use, behind the wait for another, and a lane pays one memory round trip per
// word instead of one for all of them. The asm has outputs and no side effects, so it is
// no memory clobber: descriptor reads after it still come from the scalar cache.
__device__ __forceinline__ void loaded_together(uint32_t& a, uint32_t& b, uint32_t& c) {
  asm("" : "+v"(a), "+v"(b), "+v"(c));
}
// The same for two loaded words and a draw block computed while they are in flight: the
// draw's words pass through the barrier too, so its rounds are scheduled before the wait.
__device__ __forceinline__ void loaded_together(uint32_t& a, uint32_t& b, u32x4& d) {
  asm("" : "+v"(a), "+v"(b), "+v"(d.w[0]), "+v"(d.w[1]), "+v"(d.w[2]), "+v"(d.w[3]));
}

__device__ __forceinline__ bool contains(const Rect& a, int32_t x, int32_t y) {
  return a.x0 <= x && a.x1 >= x && a.y0 <= y && a.y1 >= y;
}
__device__ __forceinline__ bool same_rect(const Rect& a, const Rect& b) {
  return a.x0 == b.x0 && a.y0 == b.y0 && a.x1 == b.x1 && a.y1 == b.y1;
}

// Moore neighbour order (point.rs:57-61): (-1,-1) (0,-1) (1,-1) (-1,0) (1,0) (-1,1) (0,1) (1,1)
__device__ constexpr int moore_dx(int k) { return k < 3 ? k - 1 : (k == 3 ? -1 : (k == 4 ? 1 : k - 6)); }
__device__ constexpr int moore_dy(int k) { return k < 3 ? -1 : (k < 5 ? 0 : 1); }

// Compile-time loop over the eight Moore neighbours: f(Idx<k>) for k = 0..7.
template <int V>
struct Idx {
  static constexpr int value = V;
};
template <int K = 0, class F>
__device__ __forceinline__ void for_moore(F&& f) {
  if constexpr (K < 8) {
    f(Idx<K>{});
    for_moore<K + 1>(f);
  }
}

__device__ __forceinline__ uint32_t fld(uint32_t w, uint32_t b, uint32_t m) { return (w >> b) & m; }
__device__ __forceinline__ uint32_t set_fld(uint32_t w, uint32_t b, uint32_t m, uint32_t v) {
  return (w & ~(m << b)) | ((v & m) << b);
}

// Index of cell (x, y) in the padded grid / claim arrays of one simulation. A simulation
// has fewer than 2^32 padded cells (grid_size <= 32,000), so the index is 32-bit and a
// load adds it to the simulation's base as an unsigned offset.
__device__ __forceinline__ uint32_t cell_index(const SimDesc& D, int32_t x, int32_t y) {
  return (static_cast<uint32_t>(y) + kPadY) * D.pitch + static_cast<uint32_t>(x) + kPadX;
}

// 3 x 3 neighbourhood masks: bit 3 * (dy + 1) + (dx + 1) stands for the cell at offset
// (dx, dy), row-major, so bit 4 is the centre. moore8 drops the centre, which leaves the
// Moore order of point.rs:57-61: (-1,-1) (0,-1) (1,-1) (-1,0) (1,0) (-1,1) (0,1) (1,1).
// Index of the n-th (0-based) set bit of an 8-bit mask m, n < popcount(m): two halving
// steps on popcounts and a last compare, the same work for every n (a loop clearing n bits
// runs as long as the wavefront's largest n).
__device__ __forceinline__ uint32_t nth_bit8(uint32_t m, uint32_t n) {
  uint32_t base = 0;
  const uint32_t c4 = __builtin_popcount(m & 0xFu);
  if (n >= c4) {
    n -= c4;
    m >>= 4;
    base = 4;
  }
  const uint32_t c2 = __builtin_popcount(m & 0x3u);
  if (n >= c2) {
    n -= c2;
    m >>= 2;
    base += 2;
  }
  return base + (n >= (m & 1u) ? 1u : 0u);
}

__device__ __forceinline__ uint32_t moore8(uint32_t m9) { return (m9 & 0xFu) | ((m9 >> 1) & 0xF0u); }

// Bit j (j = 0..2) set iff lo <= base + j <= hi.
__device__ __forceinline__ uint32_t span3(int32_t base, int32_t lo, int32_t hi) {
  const int32_t from = min(max(lo - base, 0), 3), past = min(max(base + 2 - hi, 0), 3);
  return (0x7u << from) & (0x7u >> past) & 0x7u;
}

// Three grid bytes as 24 bits (columns px-1, px, px+1 of a window row).
__device__ __forceinline__ uint32_t three(uint64_t row, int32_t byte_off) {
  return static_cast<uint32_t>(row >> (8 * byte_off)) & 0xFFFFFFu;
}
// Per-byte tests of three bytes, each to bits 0..2. Exact: adding 0x7F to a byte's low
// seven bits never carries into the next byte.
__device__ __forceinline__ uint32_t pack3(uint32_t top_bits) {  // bits 7, 15, 23 -> 0, 1, 2
  return ((top_bits >> 7) & 1u) | ((top_bits >> 14) & 2u) | ((top_bits >> 21) & 4u);
}
__device__ __forceinline__ uint32_t nonzero3(uint32_t v) {
  return pack3((((v & 0x7F7F7Fu) + 0x7F7F7Fu) | v) & 0x808080u);
}
__device__ __forceinline__ uint32_t zero3(uint32_t v) { return nonzero3(v) ^ 0x7u; }

// Claim key of agent `id` at `hour`: hours of day 7..23, 0 (the hours that move anyone)
// rank 0..17 since the reset at hour of day 2; a later hour takes a smaller top field, so
// its claims undercut any left from earlier hours of the day, and within an hour the
// lowest id wins.
__device__ __forceinline__ uint32_t claim_key(uint32_t hour, uint32_t id) {
  const uint32_t rank = (hour % 24u + 17u) % 24u;
  return ((31u - rank) << kClaimIdBits) | id;
}

// Hour of day 2: block b of a simulation's n_blocks resets its share of the claim array
// (all padded cells, in 16-byte stores; the array of each simulation is a multiple of 4
// cells long and starts at a multiple of 4). The share is computed before the block's
// first store, while the descriptor still comes from the scalar cache.
struct ClaimShare {
  uint4* q;
  uint32_t begin, end;
};
__device__ __forceinline__ ClaimShare claim_share(const DevView& v, const SimDesc& D, uint32_t b) {
  const uint32_t quads = D.pitch / 4u * (D.height + 2u * kPadY);
  const uint32_t per = (quads + D.n_blocks - 1u) / D.n_blocks;
  return ClaimShare{reinterpret_cast<uint4*>(v.claim + D.cell_off), b * per, min(quads, (b + 1u) * per)};
}
__device__ __forceinline__ void reset_claims(const ClaimShare& c) {
  for (uint32_t k = c.begin + threadIdx.x; k < c.end; k += blockDim.x) c.q[k] = make_uint4(~0u, ~0u, ~0u, ~0u);
}

// Per-thread state of one agent's hour.
struct Agent {
  const SimDesc& D;
  const SimState& S;
  const uint8_t* __restrict__ occ;  // this simulation's grid
  // Transmission thresholds, read once up front: a per-lane choice between two descriptor
  // fields would otherwise become a per-lane load of the chosen field.
  const uint64_t thr_regular, thr_high;
  uint32_t id, h, w, tm, wsa;
  uint32_t tm0 = 0;  // tm as stored: loaded for the states that carry an at_hour, else 0
  const uint32_t* tm_arr = nullptr;  // this agent's stored tm word
  int32_t x, y;                     // current cell
  int32_t tx, ty;                   // chosen cell
  int32_t hx = 0, hy = 0;           // home house origin (house id % H)
  int32_t ox = 0, oy = 0;           // office origin (office id % O)
  // The hour's movement and disease draw blocks, computed once up front (one inlined
  // Philox each) for the hours / states that can use them.
  u32x4 mv{{0, 0, 0, 0}}, dz{{0, 0, 0, 0}};
  // 5x5 grid window around (wx, wy), one 64-bit row per grid row: bytes of columns
  // c0 .. c0+7 where c0 = (x + kPadX - 2) & ~3 and woff = x + kPadX - c0 (2..5).
  bool have_win = false;
  int32_t wx = 0, wy = 0, woff = 0;
  uint64_t r0 = 0, r1 = 0, r2 = 0, r3 = 0, r4 = 0;  // named, so no variable-index access exists

  __device__ Agent(const SimDesc& d, const SimState& s, const uint8_t* o)
      : D(d), S(s), occ(o), thr_regular(Disease::transmission_threshold(d.dz, 1u)),
        thr_high(Disease::transmission_threshold(d.dz, 2u)) {}

  __device__ __forceinline__ u32x4 dr(uint32_t site, uint32_t block) const {
    return philox4x32_10(id, h, site, block, D.seed_lo, D.seed_hi);
  }
  __device__ __forceinline__ const u32x4& move_words() const { return mv; }
  __device__ __forceinline__ const u32x4& disease_words() const { return dz; }

  __device__ __forceinline__ uint32_t state() const { return fld(w, B_STATE, M_STATE); }
  __device__ __forceinline__ uint32_t sev() const { return fld(w, B_SEV, M_SEV); }
  __device__ __forceinline__ uint32_t ws() const { return fld(w, B_WS, M_WS); }
  __device__ __forceinline__ bool symptomatic() const {
    return state() == ST_I && (sev() == SEV_MILD || sev() == SEV_SEVERE);
  }
  __device__ __forceinline__ bool can_move() const {  // citizen/mod.rs:452-454
    return !(symptomatic() || (w & F_HOSP) || state() == ST_D || (w & F_ISO));
  }
  // Home house id % H and office id % O, tiled row-major (area.rs:95-117), once per hour
  // with the division constants of the tiling. The office is needed only at hours of day
  // 8..16: a worker's area is its office from 8 to 16 (goto home resets it at 16, the
  // sleep hours reset it to home), the office is the target at 8, and the symptomatic-at-
  // work override of goto home is tested at 16; at other hours ox, oy are left unset and
  // every expression that reads them is masked by a condition that is false then.
  __device__ __forceinline__ void locate_buildings(bool office) {
    const uint32_t hk = fast_mod(id, D.fd_houses), hrow = fast_div(hk, D.fd_house_nx);
    hx = D.house_sx + static_cast<int32_t>(hk - hrow * D.fd_house_nx.d) * 2;
    hy = D.house_sy + static_cast<int32_t>(hrow) * 2;
    if (office) {
      const uint32_t ok = fast_mod(id, D.fd_offices), orow = fast_div(ok, D.fd_office_nx);
      ox = D.office_sx + static_cast<int32_t>(ok - orow * D.fd_office_nx.d) * 10;
      oy = D.office_sy + static_cast<int32_t>(orow) * 10;
    }
  }
  __device__ __forceinline__ Rect home() const { return Rect{hx, hy, hx + 1, hy + 1}; }
  __device__ __forceinline__ bool in_house(int32_t cx, int32_t cy) const {
    return cx >= hx && cx <= hx + 1 && cy >= hy && cy <= hy + 1;
  }
  __device__ __forceinline__ Rect work() const {  // non-workers' work location is home
    return ws() == WS_NA ? home() : Rect{ox, oy, ox + 9, oy + 9};
  }
  __device__ __forceinline__ Rect area_rect(uint32_t tag) const {
    switch (tag) {
      case AREA_HOME: return home();
      case AREA_WORK: return work();
      case AREA_TRANSPORT: return D.transport;
      case AREA_HOUSING: return D.housing;
      case AREA_HOSP0: return D.hosp[0];
      default: return D.hosp[1];
    }
  }
  __device__ __forceinline__ Rect current_area() const { return area_rect(fld(w, B_AREA, M_AREA)); }
  __device__ __forceinline__ void set_area(uint32_t tag) { w = set_fld(w, B_AREA, M_AREA, tag); }

  __device__ __forceinline__ bool vacant(int32_t cx, int32_t cy) const {
    return occ[cell_index(D, cx, cy)] == 0;
  }
  // 3 x 3 mask of the cells around (cx, cy) that lie in rect a and in the grid
  // (area.rs:83-88 and allocation_map.rs:156-159 for each neighbour).
  __device__ __forceinline__ uint32_t allowed9(const Rect& a, int32_t cx, int32_t cy) const {
    const int32_t last = static_cast<int32_t>(D.g) - 1;
    const uint32_t cols = span3(cx - 1, max(a.x0, 0), min(a.x1, last));
    const uint32_t rows = span3(cy - 1, max(a.y0, 0), min(a.y1, last));
    return cols * ((rows & 1u) | ((rows & 2u) << 2) | ((rows & 4u) << 4));  // cols copied to each row
  }
  __device__ __forceinline__ bool in_grid(int32_t cx, int32_t cy) const {
    return cx >= 0 && cy >= 0 && cx < static_cast<int32_t>(D.g) && cy < static_cast<int32_t>(D.g);
  }
  // Ten independent aligned dword loads: rows cy-2..cy+2, 8 bytes each.
  __device__ __forceinline__ void load_window(int32_t cx, int32_t cy) {
    const uint32_t px = static_cast<uint32_t>(cx) + kPadX;
    const uint32_t c0 = (px - 2u) & ~3u;
    const uint32_t* p = reinterpret_cast<const uint32_t*>(occ + ((static_cast<uint32_t>(cy) + kPadY - 2u) * D.pitch + c0));
    const uint32_t q = D.pitch >> 2;
    uint32_t lo[5], hi[5];
#pragma unroll
    for (int r = 0; r < 5; ++r) {
      lo[r] = p[r * q];
      hi[r] = p[r * q + 1];
    }
    r0 = join64(hi[0], lo[0]);
    r1 = join64(hi[1], lo[1]);
    r2 = join64(hi[2], lo[2]);
    r3 = join64(hi[3], lo[3]);
    r4 = join64(hi[4], lo[4]);
    wx = cx;
    wy = cy;
    woff = static_cast<int32_t>(px - c0);
    have_win = true;
  }
  template <int R>
  __device__ __forceinline__ uint64_t row() const {
    static_assert(R >= 0 && R < 5, "window row");
    if constexpr (R == 0) return r0;
    else if constexpr (R == 1) return r1;
    else if constexpr (R == 2) return r2;
    else if constexpr (R == 3) return r3;
    else return r4;
  }
  // Byte of cell (wx + dx, wy + DY), |dx| <= 2, |DY| <= 2.
  template <int DY>
  __device__ __forceinline__ uint32_t wbyte(int32_t dx) const {
    return static_cast<uint32_t>(row<DY + 2>() >> (8 * (woff + dx))) & 0xFFu;
  }
  // Make the window cover the 3x3 neighbourhood of (cx, cy).
  __device__ __forceinline__ void cover(int32_t cx, int32_t cy) {
    if (!have_win || cx - wx > 1 || wx - cx > 1 || cy - wy > 1 || wy - cy > 1) load_window(cx, cy);
  }
  __device__ __forceinline__ void random_point(const Rect& a, int32_t& px, int32_t& py) {  // area.rs:76-81
    const u32x4& m = move_words();
    px = a.x0 + static_cast<int32_t>(below(m.w[0], static_cast<uint32_t>(a.x1 - a.x0 + 1)));
    py = a.y0 + static_cast<int32_t>(below(m.w[1], static_cast<uint32_t>(a.y1 - a.y0 + 1)));
  }
  // move_agent (allocation_map.rs:136-142): the target if vacant in the snapshot, else stay
  __device__ __forceinline__ void move_to(int32_t px, int32_t py) {
    if (vacant(px, py)) { tx = px; ty = py; }
  }

  // The hour's movement primitive. Every path of perform_movements ends in at most one of
  //   ACT_LOCAL: move_agent_from (citizen/mod.rs:415-432): from the current cell, or from a
  //              random point of the current area when outside it, to a random vacant
  //              Moore neighbour inside the area;
  //   ACT_TELE : goto_area for a working citizen (citizen/mod.rs:367-395): a random point
  //              of the target area if vacant (a non-worker's goto_area is ACT_LOCAL).
  // Lanes choose their primitive first and one shared copy runs, so a wavefront mixing work
  // statuses pays for one window load and one neighbour scan.
  enum : uint32_t { ACT_NONE = 0, ACT_LOCAL = 1, ACT_TELE = 2 };

  // goto_area's gate for a working citizen: can_move, or the symptomatic-at-work override.
  __device__ __forceinline__ uint32_t tele_act(const Rect& target) const {
    const uint32_t s = ws();
    const bool override_movement = (s == WS_NORMAL || s == WS_ESSENTIAL) && symptomatic() &&
                                   same_rect(target, home()) && contains(work(), x, y);
    return (can_move() || override_movement) ? ACT_TELE : ACT_NONE;
  }

  __device__ __forceinline__ void move(uint32_t act, const Rect& target) {
    if (act == ACT_NONE) return;
    const bool tele = act == ACT_TELE;
    if (!tele && !can_move()) return;
    const Rect a = tele ? target : current_area();
    int32_t lx = x, ly = y;
    if (tele || !contains(a, x, y)) random_point(a, lx, ly);  // teleport origin, vacancy not checked
    load_window(lx, ly);
    if (tele) {  // vacancy of the point; the window also serves the infection scan around it
      if (wbyte<0>(0) == 0) { tx = lx; ty = ly; }
      return;
    }
    // vacant Moore neighbours of (lx, ly) in the area and the grid: the window rows
    // ly-1 .. ly+1, bytes woff-1 .. woff+1, tested as three bytes at a time
    const uint32_t vac9 = zero3(three(r1, woff - 1)) | (zero3(three(r2, woff - 1)) << 3) |
                          (zero3(three(r3, woff - 1)) << 6);
    uint32_t cand = moore8(vac9 & allowed9(a, lx, ly));
    const uint32_t n = __builtin_popcount(cand);
    if (n == 0) return;
    const int k = static_cast<int>(nth_bit8(cand, below(move_words().w[2], n)));
    tx = lx + moore_dx(k);
    ty = ly + moore_dy(k);
  }

  // default_disease_handler.rs:64-86 as an OR over fixed per-neighbour draws: first the
  // mask of eligible infectious neighbours of (tx, ty) from the window (rows chosen with
  // bit masks, so no variable indexing), then one Philox block per pair of neighbours,
  // drawn in Moore order only for eligible ones.
  __device__ __forceinline__ bool on_susceptible() {
    if (!Disease::can_catch(w)) return false;
    const Rect a = current_area();
    cover(tx, ty);
    // No infectious cell (byte >= 2) anywhere in the 5 x 5 window, a superset of the
    // neighbourhood: nothing to test, and no draw is made. Most agents, and whole waves,
    // take this exit while infections are rare.
    if (((r0 | r1 | r2 | r3 | r4) & 0xFEFEFEFEFEFEFEFEull) == 0) return false;
    const int32_t ox = tx - wx, oy = ty - wy;  // both in [-1, 1]
    const uint64_t mm = oy < 0 ? ~0ull : 0ull, m0 = oy == 0 ? ~0ull : 0ull, mp = oy > 0 ? ~0ull : 0ull;
    const uint64_t top = (r0 & mm) | (r1 & m0) | (r2 & mp);
    const uint64_t mid = (r1 & mm) | (r2 & m0) | (r3 & mp);
    const uint64_t bot = (r2 & mm) | (r3 & m0) | (r4 & mp);
    // infectious neighbours (grid byte >= 2, i.e. a byte with a bit above bit 0) and the
    // high-window ones (byte 3) among the cells of the area and the grid
    const int32_t c = woff + ox - 1;
    const uint32_t bt = three(top, c), bm = three(mid, c), bb = three(bot, c);
    const uint32_t ok = allowed9(a, tx, ty);
    uint32_t hot = moore8((nonzero3(bt & 0xFEFEFEu) | (nonzero3(bm & 0xFEFEFEu) << 3) |
                           (nonzero3(bb & 0xFEFEFEu) << 6)) & ok);
    const uint32_t high = moore8((nonzero3(bt & (bt >> 1) & 0x010101u) | (nonzero3(bm & (bm >> 1) & 0x010101u) << 3) |
                                  (nonzero3(bb & (bb >> 1) & 0x010101u) << 6)) & ok);
    int blk = -1;
    uint32_t q0 = 0, q1 = 0, q2 = 0, q3 = 0;
    while (hot) {
      const int k = __builtin_ctz(hot);
      hot &= hot - 1;
      if ((k >> 1) != blk) {
        blk = k >> 1;
        const u32x4 q = dr(SITE_INFECT, static_cast<uint32_t>(blk));
        q0 = q.w[0]; q1 = q.w[1]; q2 = q.w[2]; q3 = q.w[3];
      }
      const uint64_t v = (k & 1) ? join64(q2, q3) : join64(q0, q1);
      const uint64_t thr = ((high >> k) & 1u) ? thr_high : thr_regular;
      if (bernoulli(thr, v)) return true;
    }
    return false;
  }

  // The stored at_hour of a state that carries one (Exposed, Pre), read where it is needed.
  __device__ __forceinline__ void load_tm() {
    if (Disease::timed(w)) tm = tm0 = *tm_arr;
  }

  // DiseaseStateMachine::next (disease_state_machine.rs:53-70): the handler's transition of
  // the current state; on_susceptible is the neighbour scan above.
  __device__ __forceinline__ void update_infection_dynamics() {
    const uint32_t s = state();
    if (s == ST_S) {
      if (on_susceptible()) Disease::on_caught(h, w, tm);
    } else if (s == ST_E) {
      load_tm();
      Disease::on_exposed(D.dz, h, disease_words(), w, tm);
    } else if (s == ST_I) {
      load_tm();
      Disease::on_infected(D.dz, h, disease_words(), w, tm);
    }
  }

  // citizen/mod.rs:257-349: choose the primitive and the area it leads to, run it, then the
  // disease update (skipped by the hospital-staff quarantine transitions, which return early).
  __device__ __forceinline__ void perform_movements(uint32_t hd) {
    const uint32_t s = ws();
    uint32_t act = ACT_NONE, new_area = ~0u;
    Rect target = D.housing;
    bool disease = true;
    if (s == WS_NORMAL || s == WS_ESSENTIAL) {
      if (hd == 7 || hd == 17) {
        if (w & F_PT) {
          target = D.transport;
          act = tele_act(target);
          new_area = AREA_TRANSPORT;
        } else {
          act = ACT_LOCAL;
        }
      } else if (hd == 8) {
        target = work();
        act = tele_act(target);
        new_area = AREA_WORK;
      } else if (hd == 16) {
        target = home();
        act = tele_act(target);
        new_area = AREA_HOME;
      } else {
        act = ACT_LOCAL;
      }
    } else if (s == WS_STAFF) {
      const uint32_t since = h > wsa ? h - wsa : 0u;  // saturating_sub
      if (since == 24u * 14u) {
        w |= F_WQ;
        disease = false;
      } else if (since == 24u * 14u * 2u) {
        target = home();
        act = tele_act(target);
        new_area = AREA_HOME;
        wsa = h + 24u * 14u;
        disease = false;
      } else if (hd == 8) {
        const uint32_t ver = S.hosp_version;
        if (!same_rect(current_area(), D.hosp[ver]) && wsa <= h) {
          target = D.hosp[ver];
          act = tele_act(target);
          new_area = AREA_HOSP0 + ver;
          wsa = h;
        }
        w &= ~F_WQ;
      } else if (hd == 16) {
        w |= F_WQ;
      } else if (!(w & F_WQ)) {
        act = ACT_LOCAL;
      }
    } else {  // NA: goto_area of a non-worker is move_agent_from in the current area
      act = ACT_LOCAL;
      if (hd == 8) new_area = AREA_HOUSING;
      else if (hd == 12) new_area = AREA_HOME;
    }
    move(act, target);
    if (new_area != ~0u) set_area(new_area);
    if (disease) update_infection_dynamics();
  }

  // citizen/mod.rs:351-365 with allocation_map.rs:144-150
  __device__ __forceinline__ void hospitalize() {
    if ((w & F_HOSP) || !Disease::is_to_be_hospitalize(D.dz, w)) return;
    const uint32_t fv = S.first_vacant;
    if (fv != ~0u) {
      const Rect hr = D.hosp[S.hosp_version];
      const uint32_t wd = static_cast<uint32_t>(hr.x1 - hr.x0 + 1);
      tx = hr.x0 + static_cast<int32_t>(fv % wd);
      ty = hr.y0 + static_cast<int32_t>(fv / wd);
      w |= F_HOSP;
    } else {
      mv = dr(SITE_MOVE, 0);  // hour of day 0 draws it only here
      int32_t px, py;
      random_point(home(), px, py);
      move_to(px, py);
    }
  }

  // citizen/mod.rs:397-413: the handler's on_routine_end, then the recovered go home and the
  // recovered or deceased leave hospital.
  __device__ __forceinline__ void deceased() {
    load_tm();
    Disease::on_routine_end(D.dz, disease_words(), w, tm);
    if (state() == ST_R) {
      mv = dr(SITE_MOVE, 0);  // hour of day 23 draws it only here
      int32_t px, py;
      random_point(home(), px, py);
      move_to(px, py);
    }
    if ((state() == ST_R || state() == ST_D) && (w & F_HOSP)) w &= ~F_HOSP;
  }
};

// The descriptor fields the epilogue reads.
struct EpiDesc {
  uint64_t row_off, event_off;
  uint32_t event_cap, hours, n_vacc, vacc_off;
  uint32_t has_lockdown, lockdown_at, has_hospital, spread_rate_threshold;
};

__device__ __forceinline__ void add_event(const DevView& v, const EpiDesc& D, SimState& st, uint32_t hour,
                                          uint32_t kind) {
  const uint32_t n = st.n_events;
  if (n < D.event_cap) v.events[D.event_off + n] = Event{hour, kind};
  st.n_events = n + 1;
}

// What the epilogue reads besides the counts, loaded by its thread 0 before the counts are
// reduced so the loads overlap: the simulation's descriptor and state, and the next entry
// of the vaccination schedule.
struct EpilogueIn {
  EpiDesc D;
  SimState st;
  uint32_t vacc_h;           // hour of schedule entry st.vacc_next, ~0u past the end
  unsigned long long vacc_t;
};

__device__ __forceinline__ void load_epilogue_in(const DevView& v, uint32_t sim, EpilogueIn& in) {
  const SimDesc& d = v.desc[sim];
  in.D = EpiDesc{d.row_off, d.event_off, d.event_cap, d.hours, d.n_vacc, d.vacc_off,
                 d.has_lockdown, d.lockdown_at, d.has_hospital, d.spread_rate_threshold};
  in.st = v.state[sim];
  const bool more = in.st.vacc_next < in.D.n_vacc;
  in.vacc_h = more ? v.vacc_hour[in.D.vacc_off + in.st.vacc_next] : ~0u;
  in.vacc_t = more ? v.vacc_thr[in.D.vacc_off + in.st.vacc_next] : 0ull;
}

// Hour epilogue for one simulation (after k_step):
// the CSV row (counts.rs:126-140), interventions (allocation_map.rs:306-337,
// lockdown.rs:55-73, hospital.rs:51-74, vaccination.rs:52-54) and the standalone stop
// rule (epidemiology_simulation.rs:564-575). Works on the registers of `in` and stores
// the state once at the end.
__device__ void hour_epilogue(const DevView& v, EpilogueIn& in, SimState* S, uint32_t h, const uint32_t* c) {
  const EpiDesc& D = in.D;
  SimState& st = in.st;
  const Counts7 row{h, c[0], c[1], c[2], c[3], c[4], c[5]};
  v.rows[D.row_off + (h - 1)] = row;
  v.host_rows[D.row_off + (h - 1)] = row;  // what a step call returns, without a copy kernel
  const uint32_t infected = c[2];
  uint32_t sweep = 0;
  unsigned long long vacc_thr = 0;
  if (D.n_vacc) {
    // schedule sorted by hour, one entry per hour (the last config entry wins); hours
    // advance by one per call, so a cursor replaces a search
    uint32_t k = st.vacc_next, vh = in.vacc_h;
    unsigned long long vt = in.vacc_t;
    while (k < D.n_vacc && vh < h) {
      ++k;
      if (k < D.n_vacc) {
        vh = v.vacc_hour[D.vacc_off + k];
        vt = v.vacc_thr[D.vacc_off + k];
      }
    }
    if (k < D.n_vacc && vh == h) {
      sweep |= SWEEP_VACC;
      vacc_thr = vt;
      add_event(v, D, st, h, EV_VACCINATION);
      ++k;
    }
    st.vacc_next = k;
  }
  if (D.has_lockdown && !st.locked && h % 24u == 0 && infected > D.lockdown_at) {
    st.locked = 1;
    sweep |= SWEEP_LOCK;
    add_event(v, D, st, h, EV_LOCKED);
  }
  if (st.locked && h == st.zero_infection_hour + 21u * 24u) {
    sweep |= SWEEP_UNLOCK;
    st.locked = 0;
    st.zero_infection_hour = 0;
    add_event(v, D, st, h, EV_UNLOCKED);
  }
  if (h % 24u == 0) st.hosp_new_infections = infected > st.hosp_new_infections ? infected - st.hosp_new_infections : 0u;
  if (!st.hosp_applied && h % 24u == 0 && D.has_hospital && st.hosp_new_infections >= D.spread_rate_threshold) {
    st.hosp_version = 1;
    st.hosp_applied = 1;
    add_event(v, D, st, h, EV_HOSPITAL);
  }
  if ((c[1] == 0 && c[2] == 0 && c[3] == 0) || h + 1 >= D.hours) st.active = 0;
  st.hour = h;
  st.rows_written = h;
  st.first_vacant = ~0u;  // this hour's scan is consumed; the next one starts fresh
#pragma unroll
  for (int k = 0; k < 6; ++k) st.last_counts[k] = c[k];
  // Slots of this hour's parity: the sweep (read by the next hour's agent pass) and the
  // next hour's agent pass (0 once stopped). Constant indices keep st in registers.
  const uint32_t next = st.active ? h + 1u : 0u;
  if (h % 24u == 0) {
    st.night_sweep = sweep;
    st.night_vacc_hour = h;
    st.night_vacc_thr = vacc_thr;
  }
  if (h & 1u) {
    st.sweep[1] = sweep;
    st.sweep_vacc_hour[1] = h;
    st.sweep_vacc_thr[1] = vacc_thr;
    st.commit_hour[0] = next;
  } else {
    st.sweep[0] = sweep;
    st.sweep_vacc_hour[0] = h;
    st.sweep_vacc_thr[0] = vacc_thr;
    st.commit_hour[1] = next;
  }
  *S = st;
  v.host_state[S - v.state] = st;
}

// A decided sweep (allocation_map.rs:349-387): vaccination of susceptible agents (a
// Bernoulli keyed by the vaccination hour), lockdown isolation of everyone but essential
// workers, unlock. Applying it twice gives the same word.
struct Sweep {
  uint32_t mask, vacc_hour;
  unsigned long long vacc_thr;
};
// The sweep in the slot of parity p (filled by the epilogue of an hour of that parity).
__device__ __forceinline__ Sweep sweep_slot(const SimState& S, uint32_t p) {
  return Sweep{S.sweep[p], S.sweep_vacc_hour[p], S.sweep_vacc_thr[p]};
}
__device__ __forceinline__ uint32_t apply_sweep(uint32_t w, const Sweep& sw, uint32_t local, const SimDesc& D) {
  const uint32_t sweep = sw.mask;
  if ((sweep & SWEEP_VACC) && fld(w, B_STATE, M_STATE) == ST_S) {
    const u32x4 d = philox4x32_10(local, sw.vacc_hour, SITE_VACCINATE, 0, D.seed_lo, D.seed_hi);
    if (bernoulli(sw.vacc_thr, join64(d.w[0], d.w[1]))) w |= F_VACC;
  }
  if ((sweep & SWEEP_LOCK) && fld(w, B_WS, M_WS) != WS_ESSENTIAL) w |= F_ISO;
  if (sweep & SWEEP_UNLOCK) w &= ~F_ISO;
  return w;
}

// Block-wide hour epilogue of one simulation: reduce the hour's per-block partial counts
// (hours of day 1..6 repeat the last counts: nobody's state changes), then hour_epilogue
// on thread 0, whose other inputs are loaded first so they arrive with the partials.
__device__ void run_epilogue(const DevView& v, uint32_t sim, const SimDesc& D, SimState* S, uint32_t h) {
  __shared__ uint32_t s_sum[6];
  EpilogueIn in;
  if (h % 24u >= 1 && h % 24u <= 6) {
    if (threadIdx.x == 0) {
      load_epilogue_in(v, sim, in);
      uint32_t c[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) c[k] = in.st.last_counts[k];
      hour_epilogue(v, in, S, h, c);
    }
    return;
  }
  if (threadIdx.x == 0) load_epilogue_in(v, sim, in);
  if (threadIdx.x < 6) s_sum[threadIdx.x] = 0;
  __syncthreads();
  uint32_t acc[6] = {0, 0, 0, 0, 0, 0};
  // the partials of the block table the hour ran on
  const uint32_t n_blocks = v.office ? D.office_n_blocks : D.n_blocks;
  const uint32_t begin = v.office ? D.office_block_begin : D.block_begin;
  for (uint32_t b = threadIdx.x; b < n_blocks; b += blockDim.x) {
    const uint32_t* bc = v.block_counts + 6u * (begin + b);
#pragma unroll
    for (int k = 0; k < 6; ++k) acc[k] += bc[k];
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    uint32_t x = acc[k];
    for (int off = kWave / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0 && x) atomicAdd(&s_sum[k], x);
  }
  __syncthreads();
  if (threadIdx.x == 0) hour_epilogue(v, in, S, h, s_sum);
}

__global__ void k_step(DevView v) {
  const AgentBlock bm = v.agent_blocks[blockIdx.x];
  const uint32_t sim = first_lane(bm.sim);
  const SimDesc& D = v.desc[sim];
  SimState* S = v.state + sim;
  if (!S->active) return;  // uniform over the block
  const uint32_t h = S->hour + 1;  // an hour of day 0 or 7..23
  const uint32_t hd = h % 24u;
  const Sweep sweep = sweep_slot(*S, (h - 1u) & 1u);  // decided at hour h - 1
  const uint32_t slot = first_lane(bm.first) + threadIdx.x;
  const uint32_t local = agent_of_lane(v, D, bm, threadIdx.x);  // the agent id
  // Every slot below the block's count holds an agent (the office slot map is a
  // permutation of the ids), so validity does not wait for the slot -> id read, and the
  // agent words below are loaded alongside it.
  const bool valid = threadIdx.x < first_lane(bm.count);
  const uint8_t* occ = v.occ + D.cell_off;

  uint32_t cat = 6;
  // This lane's claim and state word, carried past the block's claim filter to the stores.
  bool moves = false;
  uint32_t key = 0, w_out = 0, w_in = 0, tgt_word = 0, office_cell = kNoCell;
  bool whole_office = false;  // the claimed cell's office has all its slots in this block
  bool house_move = false;    // a home-layout move into the agent's own house
  const uint64_t gi = D.agent_off + slot;
  uint64_t ci = 0;
  // Office hours: the offices of this block's workers (j_lo .. j_hi, none in a block of
  // other agents and in the home layout), each with 100 cells in the block's claim filter.
  const uint32_t j_lo = first_lane(bm.j_lo), j_hi = first_lane(bm.j_hi);
  const uint32_t full_lo = first_lane(bm.full_lo), full_hi = first_lane(bm.full_hi);
  const bool filter = v.office && j_lo <= j_hi && (j_hi - j_lo + 1u) <= kMaxFilterOffices;
  if (valid) {
    // st and pos are loaded together, then the work that needs neither (the buildings and
    // the hour's movement draw, which hours of day 0 and 23 do not use: hospitalize draws it
    // for its rare fallback, deceased for the recovered, who go home), then one wait (wsa
    // and tm are read by few agents and loaded where they are needed).
    uint32_t w_stored = v.st[gi];
    uint32_t p = v.pos[gi];
    Agent a(D, *S, occ);
    a.id = local;
    a.h = h;
    a.locate_buildings(hd >= 8 && hd <= 16);
    if (hd != 0 && hd != 23) a.mv = a.dr(SITE_MOVE, 0);
    loaded_together(w_stored, p, a.mv);
    a.w = sweep.mask ? apply_sweep(w_stored, sweep, local, D) : w_stored;
    a.x = static_cast<int32_t>(p & 0xFFFFu);
    a.y = static_cast<int32_t>(p >> 16);
    a.tx = a.x;
    a.ty = a.y;
    a.tm = 0;
    const uint64_t gid = D.agent_off + local;  // tm and wsa stay in the home layout (by id)
    a.tm_arr = v.tm + gid;
    const uint32_t w0 = w_stored;
    const bool staff = a.ws() == WS_STAFF;
    a.wsa = staff ? v.wsa[gid] : 0u;
    const uint32_t wsa0 = a.wsa;

    if (Disease::draws_needed(a.w, hd)) a.dz = a.dr(SITE_DISEASE, 0);
    if (hd == 0) {
      if (a.state() == ST_I) {  // increment_infection_day (state.rs:69-73)
        const uint32_t day = fld(a.w, B_DAY, M_DAY);
        a.w = set_fld(a.w, B_DAY, M_DAY, day < kMaxDay ? day + 1 : day);
      }
      a.hospitalize();
    } else if (hd == 23) {
      a.deceased();
    } else {
      a.perform_movements(hd);
    }
    // Everything the stores need is computed first: a descriptor field read after a store
    // can no longer come from the scalar cache and costs a vector round trip.
    key = claim_key(h, local);
    moves = a.tx != a.x || a.ty != a.y;
    // the grid byte depends only on the state, the hospitalized flag, the infection day and
    // the immunity, so it is recomputed only when one of those changed this hour
    constexpr uint32_t kByteFields = (M_STATE << B_STATE) | F_HOSP | (M_DAY << B_DAY) | (M_IMM << B_IMM);
    if (((a.w ^ w0) & kByteFields) && occupant_byte(D.dz, a.w) != occupant_byte(D.dz, w0)) a.w |= F_DIRTY;
    ci = D.cell_off + cell_index(D, a.tx, a.ty);
    const uint32_t s = a.state();
    cat = s == ST_S ? 0u : s == ST_E ? 1u : s == ST_I ? ((a.w & F_HOSP) ? 3u : 2u) : s == ST_R ? 4u : 5u;
    w_out = a.w;
    w_in = w0;
    tgt_word = static_cast<uint32_t>(a.tx) | (static_cast<uint32_t>(a.ty) << 16);
    // a worker's claim of a cell of its own office (the only agents that claim office
    // cells at these hours) enters the filter
    const uint32_t j = fast_mod(local, D.fd_offices);  // the agent's office
    if (filter && moves && a.tx >= a.ox && a.tx <= a.ox + 9 && a.ty >= a.oy && a.ty <= a.oy + 9 && j >= j_lo &&
        j <= j_hi) {
      office_cell = (j - j_lo) * kOfficeCells + static_cast<uint32_t>(a.ty - a.oy) * 10u +
                    static_cast<uint32_t>(a.tx - a.ox);
      whole_office = j >= full_lo && j <= full_hi;
    }
    // Home-layout hours (16..23, 0..7): only a house's own agents (ids congruent mod
    // n_houses) choose cells of that house. Local moves stay in the current area, which is
    // home for everyone but public-transport users (transport area) and hospital staff;
    // the teleports into a house (going home at 16, recovered going home at 23, the
    // hospital fallback at 0) are its own agents'; no one has the housing area then. So
    // the lowest-id claimant of such a cell is found among the agent's lower housemates,
    // whose target words the commit reads: no global claim is made.
    house_move = !v.office && moves && a.in_house(a.tx, a.ty);
    if (a.tm != a.tm0) v.tm[gid] = a.tm;
    if (staff && a.wsa != wsa0) v.wsa[gid] = a.wsa;
  }
  // Block claim filter (office hours). The office layout puts each office's workers in one
  // or two blocks, and they are the only claimants of its cells, so most same-cell rivals
  // meet here: an LDS atomicMin per claimed cell leaves only the block's lowest key, and a
  // lane that is not it has lost to a lower id and makes no global claim. At most one
  // global atomic per cell and block remains, instead of one per rival. On an office whose
  // workers all lie in this block the filter has seen every claimant, so its survivor has
  // won: it makes no global claim and flags its target word for the commit.
  if (filter) {  // uniform
    __shared__ uint32_t s_cell[kFilterCells];
    const uint32_t n_cells = (j_hi - j_lo + 1u) * kOfficeCells;
    for (uint32_t i = threadIdx.x; i < n_cells; i += blockDim.x) s_cell[i] = ~0u;
    __syncthreads();
    if (office_cell != kNoCell) atomicMin(&s_cell[office_cell], key);
    __syncthreads();
    if (office_cell != kNoCell && s_cell[office_cell] != key) moves = false;
  }
  if (valid) {
    if (moves) {
      if (whole_office) tgt_word |= kTgtWon;
      else if (house_move) tgt_word |= kTgtHouse;
      else atomicMin(&v.claim[ci], key);
    }
    v.tgt[gi] = moves ? tgt_word : kTgtNone;
    if (w_out != w_in) v.st[gi] = w_out;
  }

  // S/E/I/H/R/D via wavefront ballots into LDS, then one plain store per block and
  // category; the epilogue reduces the block partials.
  __shared__ uint32_t s_counts[6];
  if (threadIdx.x < 6) s_counts[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (uint32_t k = 0; k < 6; ++k) {
    const unsigned long long m = __ballot(cat == k);
    if (lane == 0 && m) atomicAdd(&s_counts[k], static_cast<uint32_t>(__popcll(m)));
  }
  __syncthreads();
  if (threadIdx.x < 6) v.block_counts[6u * blockIdx.x + threadIdx.x] = s_counts[threadIdx.x];
}

// The hour's epilogue of simulation `sim` by one block. A stopped simulation only clears
// the commit slot of the parity it last stepped (the slot of the next commit but one).
__device__ __forceinline__ void epilogue_block(const DevView& v, uint32_t sim) {
  SimState* S = v.state + sim;
  if (!S->active) {  // uniform
    if (threadIdx.x == 0) S->commit_hour[S->hour & 1u] = 0u;
    return;
  }
  run_epilogue(v, sim, v.desc[sim], S, S->hour + 1);
}

// Epilogues of hours of day 3..6 in one launch (one block per simulation, thread 0 walks
// the hours), for day graphs of simulations with no vaccination at hours of day 2..5. Then
// no sweep is pending at any of those hours (lockdown, unlock and hospital decisions fall
// on hours of day 0 only), so their agent passes would change nothing. Each hour does
// what epilogue_block does for it.
__device__ void quiet_epilogues(const DevView& v, uint32_t sim, uint32_t n_hours) {
  SimState* S = v.state + sim;
  for (uint32_t k = 0; k < n_hours; ++k) {
    if (!S->active) {
      S->commit_hour[S->hour & 1u] = 0u;
      continue;
    }
    EpilogueIn in;
    load_epilogue_in(v, sim, in);
    uint32_t c[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) c[q] = in.st.last_counts[q];
    hour_epilogue(v, in, S, in.st.hour + 1u, c);
  }
}
__global__ void k_quiet_hours(DevView v, uint32_t n_hours) {
  if (threadIdx.x == 0) quiet_epilogues(v, blockIdx.x, n_hours);
}

// Hours of day 1..6 in one launch, for day graphs of simulations with no vaccination at
// hours of day 1..5 (so the only sweep pending in the night is hour of day 0's). Blocks
// 0 .. n_epi-1 walk the six epilogues of simulation blockIdx.x; the agent blocks (home
// layout) do the night's agent work: hour of day 0's sweep, every non-staff current area
// to home (hour of day 1), and their share of the daily claim reset (hour of day 2). The
// epilogues rewrite the parity slots meanwhile, so the agent blocks read hour of day 0's
// sweep from its own slot, and whether the simulation runs from `active`, which these
// hours cannot change (counts repeat, and a day graph never holds the last hour).
__global__ void k_night(DevView v, uint32_t n_epi) {
  if (blockIdx.x < n_epi) {  // uniform
    if (threadIdx.x == 0) quiet_epilogues(v, blockIdx.x, 6);
    return;
  }
  const uint32_t b = blockIdx.x - n_epi;
  const AgentBlock bm = v.agent_blocks[b];
  const uint32_t sim = first_lane(bm.sim);
  const SimDesc& D = v.desc[sim];
  const SimState* S = v.state + sim;
  if (!S->active) return;  // uniform
  const Sweep sweep{S->night_sweep, S->night_vacc_hour, S->night_vacc_thr};
  const ClaimShare share = claim_share(v, D, b - D.block_begin);
  const uint32_t local = first_lane(bm.first) + threadIdx.x;
  if (local < D.n) {
    const uint64_t gi = D.agent_off + local;
    const uint32_t w = v.st[gi];
    uint32_t nw = sweep.mask ? apply_sweep(w, sweep, local, D) : w;
    if (fld(nw, B_WS, M_WS) != WS_STAFF) nw = set_fld(nw, B_AREA, M_AREA, AREA_HOME);
    if (nw != w) v.st[gi] = nw;
  }
  reset_claims(share);
}

// Agent pass of an hour of day 1..6 (block b of the simulation's agent blocks): nobody
// moves and no disease update runs while asleep (citizen/mod.rs:244-248), so only the
// sweep decided at hour h - 1 is applied, hour of day 1 sets every non-staff current area
// to home, and hour of day 2 resets the claim array before the day's first claim hour.
__device__ __forceinline__ void rest_pass(const DevView& v, const SimDesc& D, const SimState& S, uint32_t h,
                                          uint32_t local, uint32_t b) {
  const uint32_t hd = h % 24u;
  const Sweep sweep = sweep_slot(S, (h - 1u) & 1u);
  ClaimShare share{nullptr, 0, 0};
  if (hd == 2) share = claim_share(v, D, b);
  if (local < D.n && (sweep.mask || hd == 1)) {
    const uint64_t gi = D.agent_off + local;
    const uint32_t w = v.st[gi];
    uint32_t nw = sweep.mask ? apply_sweep(w, sweep, local, D) : w;
    if (hd == 1 && fld(nw, B_WS, M_WS) != WS_STAFF) nw = set_fld(nw, B_AREA, M_AREA, AREA_HOME);
    if (nw != w) v.st[gi] = nw;
  }
  reset_claims(share);
}

// `parity` is the hour's parity (an hour of day 24k + j has the parity of j). Blocks
// 0 .. n_epi-1 run the epilogue of simulation blockIdx.x; the rest run the agent pass:
// the commit at hours of day 0 and 7..23, rest_pass at 1..6.
__global__ void k_commit(DevView v, uint32_t parity, uint32_t n_epi) {
  if (blockIdx.x < n_epi) {  // uniform
    epilogue_block(v, blockIdx.x);
    return;
  }
  const AgentBlock bm = v.agent_blocks[blockIdx.x - n_epi];
  const uint32_t sim = first_lane(bm.sim);
  const SimDesc& D = v.desc[sim];
  const SimState* S = v.state + sim;
  const uint32_t h = S->commit_hour[parity];
  if (h == 0) return;  // uniform: the simulation did not step this hour
  const uint32_t slot = first_lane(bm.first) + threadIdx.x;
  if (h % 24u >= 1 && h % 24u <= 6) {  // uniform; home layout
    rest_pass(v, D, *S, h, slot, blockIdx.x - n_epi - D.block_begin);
    return;
  }
  const uint32_t local = agent_of_lane(v, D, bm, threadIdx.x);  // the agent id
  if (threadIdx.x < first_lane(bm.count)) {  // a valid slot, known without the slot -> id read
    const uint64_t gi = D.agent_off + slot;
    // st, pos and tgt are loaded together, and at home-layout hours with them the target
    // words of the agent's lower housemates (ids local - k * n_houses, k = 1..3: a house
    // holds at most 4 agents, grid.rs:133-142), consecutive for consecutive ids. A house
    // move then pays one round trip; a claimed move two (these words, then its claim).
    uint32_t w = v.st[gi];
    uint32_t p = v.pos[gi];
    uint32_t t = v.tgt[gi];
    uint32_t m1 = kTgtNone, m2 = kTgtNone, m3 = kTgtNone;
    if (!v.office) {  // uniform
      const uint32_t lower = fast_div(local, D.fd_houses);  // lower housemates
      const uint32_t H = D.fd_houses.d;
      if (lower >= 1) m1 = v.tgt[gi - H];
      if (lower >= 2) m2 = v.tgt[gi - 2ull * H];
      if (lower >= 3) m3 = v.tgt[gi - 3ull * H];
      loaded_together(m1, m2, m3);
    }
    loaded_together(w, p, t);
    const bool mover = t != kTgtNone;
    if (mover || (w & F_DIRTY)) {
      uint8_t* occ = v.occ + D.cell_off;
      const uint32_t old_ci = cell_index(D, static_cast<int32_t>(p & 0xFFFFu), static_cast<int32_t>(p >> 16));
      const uint8_t byte = occupant_byte(D.dz, w);
      bool moved = false;
      if (mover) {
        // won: decided by the office-hour claim filter; a house move wins unless a lower
        // housemate chose the same cell (the same word, kTgtHouse included); any other
        // move reads back its claim
        const bool house = (t & kTgtHouse) != 0u;
        const bool won = house ? (m1 != t && m2 != t && m3 != t) : (t & kTgtWon) != 0u;
        t &= kTgtCell;
        const uint32_t ci = cell_index(D, static_cast<int32_t>(t & 0xFFFFu), static_cast<int32_t>(t >> 16));
        if (won || (!house && (v.claim + D.cell_off)[ci] == claim_key(h, local))) {
          occ[old_ci] = 0;
          occ[ci] = byte;
          v.pos[gi] = t;
          moved = true;
        }
      }
      if (!moved && (w & F_DIRTY)) occ[old_ci] = byte;
      if (w & F_DIRTY) v.st[gi] = w & ~F_DIRTY;
    }
  }
}

// Re-applies each simulation's last decided sweep (idempotent) so that agent state read
// by the host includes it; the next hour's agent pass applies it again to the same words.
__global__ void k_apply_sweep(DevView v) {
  const AgentBlock bm = v.agent_blocks[blockIdx.x];
  const uint32_t sim = first_lane(bm.sim);
  const SimDesc& D = v.desc[sim];
  const SimState* S = v.state + sim;
  const Sweep sweep = sweep_slot(*S, S->hour & 1u);
  const uint32_t local = first_lane(bm.first) + threadIdx.x;
  if (sweep.mask == 0 || local >= D.n) return;
  const uint64_t gi = D.agent_off + local;
  const uint32_t w = v.st[gi];
  const uint32_t nw = apply_sweep(w, sweep, local, D);
  if (nw != w) v.st[gi] = nw;
}

// First vacant cell of the current hospital area in row-major order (area.rs:119-146 +
// allocation_map.rs:145), reduced with atomicMin over the snapshot grid.
__global__ void k_hospital_scan(DevView v, uint32_t cells_per_thread) {
  const BlockRef bm = v.hosp_blocks[blockIdx.x];
  const uint32_t sim = first_lane(bm.sim);
  const SimDesc& D = v.desc[sim];
  SimState* S = v.state + sim;
  if (!S->active || (S->hour + 1) % 24u != 0) return;
  const uint32_t first = first_lane(bm.first);
  if (__hip_atomic_load(&S->first_vacant, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < first)
    return;  // an earlier block already found one
  const Rect hr = D.hosp[S->hosp_version];
  const uint32_t wd = static_cast<uint32_t>(hr.x1 - hr.x0 + 1);
  const uint32_t ncell = wd * static_cast<uint32_t>(hr.y1 - hr.y0 + 1);
  const uint8_t* occ = v.occ + D.cell_off;
  uint32_t best = ~0u;
  for (uint32_t k = 0; k < cells_per_thread; ++k) {
    const uint32_t c = first + k * blockDim.x + threadIdx.x;
    if (c >= ncell) break;
    const uint32_t cx = static_cast<uint32_t>(hr.x0) + c % wd, cy = static_cast<uint32_t>(hr.y0) + c / wd;
    if (occ[cell_index(D, static_cast<int32_t>(cx), static_cast<int32_t>(cy))] == 0) { best = c; break; }
  }
  // wavefront min then one atomic per wave
  for (int off = kWave / 2; off > 0; off >>= 1) {
    const uint32_t o = __shfl_xor(best, off, kWave);
    best = o < best ? o : best;
  }
  if ((threadIdx.x & (kWave - 1)) == 0 && best != ~0u) atomicMin(&S->first_vacant, best);
}

// Layout move of st and pos (tm and wsa, which few agents read, stay in the home layout and
// are indexed by id in both), one tile of 32 x 32 agents id = k * n_offices + j per block.
// Along j the ids are consecutive: the home side and the slot map are read (or written)
// 128 B at a time, and so are the office slots of the tile's other agents, which keep id
// order. Along k lie the workers of one office, which hold consecutive office slots: the
// workers go through LDS rows of 33 words so that their office side is also accessed a
// column of the tile at a time. Each thread owns four rows of the tile and issues the loads
// of all four before any store, so that a wave keeps four times the bytes in flight.
__global__ void __launch_bounds__(256) k_relayout(DevView home, DevView office, const TileRef* tiles,
                                                  uint32_t to_office) {
  constexpr uint32_t kRows = 4;           // rows r = ly + 8 * i of the tile per thread
  __shared__ uint32_t tile[2][32][33];  // [array][k - k0][j - j0]: workers' words
  __shared__ uint32_t tslot[32][33];    // office slot of (k, j); ~0u: no agent or not a worker
  const TileRef tr = tiles[blockIdx.x];
  const SimDesc& D = home.desc[first_lane(tr.sim)];
  const uint32_t O = D.n_offices, W = D.n_workers, n = D.n;
  const uint64_t base = D.agent_off;
  const uint32_t k0 = first_lane(tr.k0), j0 = first_lane(tr.j0);
  const uint32_t lx = threadIdx.x & 31u, ly = threadIdx.x >> 5;
  const uint32_t j = j0 + lx;
  // along j: the slot map, the home words, and the non-workers' office words
  uint64_t id[kRows];
  uint32_t slot[kRows], w[kRows], p[kRows];
#pragma unroll
  for (uint32_t i = 0; i < kRows; ++i) {
    id[i] = static_cast<uint64_t>(k0 + ly + 8u * i) * O + j;
    const bool valid = j < O && id[i] < n;
    slot[i] = valid ? office.office_slot[base + id[i]] : ~0u;  // ~0u >= W: never a worker
    if (to_office && valid) {
      w[i] = home.st[base + id[i]];
      p[i] = home.pos[base + id[i]];
    }
  }
  if (!to_office) {  // uniform: the non-workers' office words, by slot
#pragma unroll
    for (uint32_t i = 0; i < kRows; ++i)
      if (slot[i] >= W && slot[i] != ~0u) {
        w[i] = office.st[base + slot[i]];
        p[i] = office.pos[base + slot[i]];
      }
  }
#pragma unroll
  for (uint32_t i = 0; i < kRows; ++i) {
    const uint32_t r = ly + 8u * i;
    if (slot[i] < W) {
      if (to_office) {
        tile[0][r][lx] = w[i];
        tile[1][r][lx] = p[i];
      }
    } else if (slot[i] != ~0u) {
      if (to_office) {
        office.st[base + slot[i]] = w[i];
        office.pos[base + slot[i]] = p[i];
      } else {
        home.st[base + id[i]] = w[i];
        home.pos[base + id[i]] = p[i];
      }
    }
    tslot[r][lx] = slot[i] < W ? slot[i] : ~0u;
  }
  __syncthreads();
  // along k: the workers of office j0 + r, in consecutive office slots
  uint32_t ks[kRows];
#pragma unroll
  for (uint32_t i = 0; i < kRows; ++i) ks[i] = tslot[lx][ly + 8u * i];
  if (to_office) {  // uniform
#pragma unroll
    for (uint32_t i = 0; i < kRows; ++i)
      if (ks[i] != ~0u) {
        office.st[base + ks[i]] = tile[0][lx][ly + 8u * i];
        office.pos[base + ks[i]] = tile[1][lx][ly + 8u * i];
      }
    return;
  }
#pragma unroll
  for (uint32_t i = 0; i < kRows; ++i)
    if (ks[i] != ~0u) {
      w[i] = office.st[base + ks[i]];
      p[i] = office.pos[base + ks[i]];
    }
#pragma unroll
  for (uint32_t i = 0; i < kRows; ++i)
    if (ks[i] != ~0u) {
      tile[0][lx][ly + 8u * i] = w[i];
      tile[1][lx][ly + 8u * i] = p[i];
    }
  __syncthreads();
  // along j: the workers' home words
#pragma unroll
  for (uint32_t i = 0; i < kRows; ++i) {
    const uint32_t r = ly + 8u * i;
    if (tslot[r][lx] == ~0u) continue;
    home.st[base + id[i]] = tile[0][r][lx];
    home.pos[base + id[i]] = tile[1][r][lx];
  }
}

__global__ void k_export_active(DevView v, uint32_t n_sims, uint32_t* active) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_sims) active[i] = v.state[i].active;
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t value) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t k = i; k < n; k += stride) p[k] = value;
}

}  // namespace

hipError_t launch_hospital_scan(const DevView& v, const LaunchShape& s, hipStream_t stream) {
  if (s.hosp_blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_hospital_scan, dim3(s.hosp_blocks), dim3(s.hosp_block), 0, stream, v, s.hosp_cells_per_thread);
  return hipGetLastError();
}

hipError_t launch_step(const DevView& v, const LaunchShape& s, hipStream_t stream) {
  const uint32_t blocks = v.office ? s.office_agent_blocks : s.agent_blocks;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_step, dim3(blocks), dim3(v.office ? s.office_block : s.block), 0, stream, v);
  return hipGetLastError();
}

hipError_t launch_commit(const DevView& v, const LaunchShape& s, uint32_t hour, hipStream_t stream) {
  const uint32_t blocks = v.office ? s.office_agent_blocks : s.agent_blocks;
  if (blocks + s.n_sims == 0) return hipSuccess;
  hipLaunchKernelGGL(k_commit, dim3(s.n_sims + blocks), dim3(v.office ? s.office_block : s.block), 0, stream, v,
                     hour & 1u, s.n_sims);
  return hipGetLastError();
}

hipError_t launch_night(const DevView& v, const LaunchShape& s, hipStream_t stream) {
  if (s.agent_blocks + s.n_sims == 0) return hipSuccess;
  hipLaunchKernelGGL(k_night, dim3(s.n_sims + s.agent_blocks), dim3(s.block), 0, stream, v, s.n_sims);
  return hipGetLastError();
}

hipError_t launch_quiet_hours(const DevView& v, const LaunchShape& s, uint32_t n_hours, hipStream_t stream) {
  if (s.n_sims == 0) return hipSuccess;
  hipLaunchKernelGGL(k_quiet_hours, dim3(s.n_sims), dim3(64), 0, stream, v, n_hours);
  return hipGetLastError();
}

hipError_t launch_apply_sweep(const DevView& v, const LaunchShape& s, hipStream_t stream) {
  if (s.agent_blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_apply_sweep, dim3(s.agent_blocks), dim3(s.block), 0, stream, v);
  return hipGetLastError();
}

hipError_t launch_relayout(const DevView& home, const DevView& office, const TileRef* tiles, uint32_t n_tiles,
                           bool to_office, hipStream_t stream) {
  if (n_tiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_relayout, dim3(n_tiles), dim3(256), 0, stream, home, office, tiles, to_office ? 1u : 0u);
  return hipGetLastError();
}

hipError_t launch_export_active(const DevView& v, uint32_t n_sims, uint32_t* active, hipStream_t stream) {
  if (n_sims == 0) return hipSuccess;
  hipLaunchKernelGGL(k_export_active, dim3((n_sims + 255) / 256), dim3(256), 0, stream, v, n_sims, active);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t value, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const uint64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_fill_u32, dim3(static_cast<uint32_t>(blocks < 65536 ? blocks : 65536)), dim3(256), 0, stream, p,
                     n, value);
  return hipGetLastError();
}

}  // namespace epi
