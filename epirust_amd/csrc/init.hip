// Population synthesis on gfx950.
//
// The reference builds the population on the CPU (engine/src/geography/grid.rs:83-155,
// engine/src/citizen/citizen_factory.rs:31-134, epidemiology_simulation.rs:75-192); in
// deterministic mode every draw is a Philox block at hour 0xFFFFFFFF, so each step is a
// data-parallel kernel over agents or houses:
//   k_init_agents  : per agent, working / public transport / work status / immunity
//                    (sites INIT_A, INIT_B, INIT_C); the public-transport draw is kept in
//                    F_PT and counted per block;
//   k_scan_exclusive + k_init_pt_cap : the transport cap in id order
//                    (citizen_factory.rs:72-83) is a rank test against an exclusive scan
//                    of the per-block counts;
//   k_select_hist  : radix selection over the 96-bit composite (INIT_C key, id), 12 bits
//                    a pass, finds the agent of a given rank without sorting; the starting
//                    infections (the `total` smallest composites, taking exposed /
//                    asymptomatic / mild / severe in composite order) need at most four
//                    such ranks;
//   k_init_infect  : per agent, its group from the four rank bounds;
//   k_init_cells   : per house, random_points(k) of the 2x2 house for its agents;
//   k_init_grid    : per agent, its byte of the padded occupancy grid.
// All of it is integer work over a few bytes per agent and runs once per simulation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "disease.hpp"
#include "init.hpp"
#include "philox.hpp"

namespace epi {
namespace {

constexpr uint32_t kInitBlock = 256;
constexpr uint32_t kSelectBits = 12;
constexpr uint32_t kSelectBins = 1u << kSelectBits;
constexpr uint32_t kSelectTopShift = 96 - kSelectBits;  // 84

struct InitArgs {
  uint32_t n;
  uint32_t seed_lo, seed_hi;
  uint64_t thr_work, thr_pt, thr_staff, thr_ess;
  const uint8_t* csv_working;  // Population::Csv columns, or null for Auto
  const uint8_t* csv_pt;
  uint32_t pt_cap;
  uint32_t* pos;
  uint32_t* st;
  uint32_t* tm;
  uint32_t* wsa;
  uint32_t* block_pt;  // per init block: public-transport draws, then their exclusive prefix
};

// A point of the 96-bit order (key, id): key in bits 95..32, id in bits 31..0.
struct Composite {
  unsigned long long key;
  uint32_t id;
};

__device__ __forceinline__ uint64_t seed_of(const InitArgs& a) { return join64(a.seed_hi, a.seed_lo); }

__device__ __forceinline__ uint64_t infection_key(const InitArgs& a, uint32_t i) {
  const u32x4 c = draw(seed_of(a), i, kInitHour, SITE_INIT_C, 0);
  return join64(c.w[2], c.w[3]);
}

// Bits [shift, shift + 12) of the composite.
__device__ __forceinline__ uint32_t digit(uint64_t key, uint32_t id, uint32_t shift) {
  if (shift >= 32) return static_cast<uint32_t>(key >> (shift - 32)) & (kSelectBins - 1);
  const uint64_t low = (key << 32) | id;
  return static_cast<uint32_t>(low >> shift) & (kSelectBins - 1);
}

__device__ __forceinline__ bool greater(uint64_t key, uint32_t id, const Composite& c) {
  return key > c.key || (key == c.key && id > c.id);
}

__global__ __launch_bounds__(kInitBlock) void k_init_agents(InitArgs a) {
  const uint32_t i = blockIdx.x * kInitBlock + threadIdx.x;
  bool pt = false;
  if (i < a.n) {
    const uint64_t seed = seed_of(a);
    const u32x4 A = draw(seed, i, kInitHour, SITE_INIT_A, 0);
    const u32x4 B = draw(seed, i, kInitHour, SITE_INIT_B, 0);
    const u32x4 C = draw(seed, i, kInitHour, SITE_INIT_C, 0);
    bool working;
    if (a.csv_working) {  // Citizen::from_record (citizen/mod.rs:155-180)
      working = a.csv_working[i] != 0;
      pt = a.csv_pt[i] != 0;
    } else {
      working = bernoulli(a.thr_work, join64(A.w[0], A.w[1]));
      pt = working && bernoulli(a.thr_pt, join64(A.w[2], A.w[3]));
    }
    uint32_t ws;
    if (!working) ws = WS_NA;
    else if (bernoulli(a.thr_staff, join64(B.w[0], B.w[1]))) ws = WS_STAFF;
    else ws = bernoulli(a.thr_ess, join64(C.w[0], C.w[1])) ? WS_ESSENTIAL : WS_NORMAL;
    const uint32_t imm = below(B.w[2], 5);  // immunity -2..2 stored as immunity + 2
    a.st[i] = (ST_S << B_STATE) | (ws << B_WS) | (AREA_HOME << B_AREA) | (imm << B_IMM) | (pt ? F_PT : 0u);
    a.wsa[i] = ws == WS_STAFF ? 8u : 0u;  // HospitalStaff { work_start_at: ROUTINE_WORK_TIME }
    a.tm[i] = 0;
  }
  const int count = __syncthreads_count(pt);
  if (threadIdx.x == 0) a.block_pt[blockIdx.x] = static_cast<uint32_t>(count);
}

// Exclusive prefix sum of v[0..n) in place, one block of 1024 threads: each thread sums a
// contiguous chunk, the chunk sums are scanned in LDS, then each chunk is rewritten.
__global__ __launch_bounds__(1024) void k_scan_exclusive(uint32_t* v, uint32_t n) {
  __shared__ uint32_t part[1024];
  const uint32_t per = (n + 1023u) / 1024u;
  const uint32_t b = threadIdx.x * per, e = min(n, b + per);
  uint32_t sum = 0;
  for (uint32_t k = b; k < e; ++k) sum += v[k];
  part[threadIdx.x] = sum;
  __syncthreads();
  for (uint32_t off = 1; off < 1024u; off <<= 1) {
    const uint32_t t = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
    __syncthreads();
    part[threadIdx.x] += t;
    __syncthreads();
  }
  uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
  for (uint32_t k = b; k < e; ++k) {
    const uint32_t x = v[k];
    v[k] = run;
    run += x;
  }
}

// Clears F_PT on every public-transport draw whose rank in id order is >= the cap.
__global__ __launch_bounds__(kInitBlock) void k_init_pt_cap(InitArgs a) {
  __shared__ uint32_t wave_count[kInitBlock / 64];
  const uint32_t i = blockIdx.x * kInitBlock + threadIdx.x;
  const bool pt = i < a.n && (a.st[i] & F_PT) != 0;
  const uint64_t mask = __ballot(pt);
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  if (lane == 0) wave_count[wave] = static_cast<uint32_t>(__popcll(mask));
  __syncthreads();
  if (!pt) return;
  uint32_t rank = a.block_pt[blockIdx.x] + static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
  for (uint32_t w = 0; w < wave; ++w) rank += wave_count[w];
  if (rank >= a.pt_cap) a.st[i] &= ~F_PT;
}

// Histogram of digit(shift) over the agents whose composite matches `prefix` on `mask`.
__global__ __launch_bounds__(kInitBlock) void k_select_hist(InitArgs a, Composite prefix, Composite mask,
                                                            uint32_t shift, uint32_t* hist) {
  __shared__ uint32_t h[kSelectBins];
  for (uint32_t k = threadIdx.x; k < kSelectBins; k += kInitBlock) h[k] = 0;
  __syncthreads();
  const uint32_t step = gridDim.x * kInitBlock;
  for (uint32_t i = blockIdx.x * kInitBlock + threadIdx.x; i < a.n; i += step) {
    const uint64_t key = infection_key(a, i);
    if (((key ^ prefix.key) & mask.key) == 0 && ((i ^ prefix.id) & mask.id) == 0)
      atomicAdd(&h[digit(key, i, shift)], 1u);
    if (i > 0xFFFFFFFFu - step) break;
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < kSelectBins; k += kInitBlock)
    if (h[k]) atomicAdd(&hist[k], h[k]);
}

// bound[j] is the composite of rank c_j - 1 for the cumulative group sizes
// c = (exposed, +asymptomatic, +mild, +severe); bit j of `zero` marks c_j == 0. An agent
// of rank r belongs to group #{j : r >= c_j}; group 4 means not selected.
struct Bounds {
  Composite bound[4];
  uint32_t zero;
};

__global__ __launch_bounds__(kInitBlock) void k_init_infect(InitArgs a, Bounds b) {
  const uint32_t i = blockIdx.x * kInitBlock + threadIdx.x;
  if (i >= a.n) return;
  const uint64_t key = infection_key(a, i);
  uint32_t group = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (((b.zero >> j) & 1u) || greater(key, i, b.bound[j])) ++group;
  if (group >= 4) return;
  // starting_infections.rs: exposed {at_hour 0}; infected asymptomatic / mild / severe on day 1
  const uint32_t state = group == 0 ? ST_E : ST_I;
  const uint32_t sev = group == 0 ? 0u : (group == 1 ? SEV_ASYM : (group == 2 ? SEV_MILD : SEV_SEVERE));
  const uint32_t day = group == 0 ? 0u : 1u;
  uint32_t w = a.st[i];
  w = (w & ~((M_STATE << B_STATE) | (M_SEV << B_SEV) | (M_DAY << B_DAY))) | (state << B_STATE) | (sev << B_SEV) |
      (day << B_DAY);
  a.st[i] = w;
  a.tm[i] = 0;
}

// Partial Fisher-Yates over {lo, lo + 1} from fixed words: min(m, 2) values.
__device__ __forceinline__ uint32_t choose2(int32_t lo, uint32_t m, uint32_t w0, uint32_t w1, int32_t* out) {
  int32_t v0 = lo, v1 = lo + 1;
  if (m > 2) m = 2;
  if (m >= 1 && below(w0, 2) == 1) {
    const int32_t t = v0;
    v0 = v1;
    v1 = t;
  }
  // the second step draws from a one-element range and swaps nothing (below(w1, 1) == 0)
  (void)w1;
  out[0] = v0;
  out[1] = v1;
  return m;
}

struct HouseLayout {
  int32_t sx, sy, nx;
  uint32_t n_houses;
};

// House h holds agents h, h + H, ...; they take random_points(k) of the house in id order.
__global__ __launch_bounds__(kInitBlock) void k_init_cells(InitArgs a, HouseLayout L) {
  const uint32_t h = blockIdx.x * kInitBlock + threadIdx.x;
  const uint32_t H = L.n_houses;
  if (h >= min(H, a.n)) return;
  const uint32_t k = (a.n - h + H - 1) / H;
  const int32_t x0 = L.sx + static_cast<int32_t>(h % static_cast<uint32_t>(L.nx)) * 2;
  const int32_t y0 = L.sy + static_cast<int32_t>(h / static_cast<uint32_t>(L.nx)) * 2;
  const u32x4 w = draw(seed_of(a), h, kInitHour, SITE_INIT_HOUSE, 0);
  uint32_t side = 1;  // ceil(sqrt(k)) in integers
  while (side * side < k) ++side;
  int32_t xs[2], ys[2];
  const uint32_t mx = choose2(x0, side, w.w[0], w.w[1], xs);
  const uint32_t my = choose2(y0, side, w.w[2], w.w[3], ys);
  uint32_t j = 0;
  for (uint32_t p = 0; p < mx && j < k; ++p)
    for (uint32_t q = 0; q < my && j < k; ++q, ++j)
      a.pos[h + j * H] = static_cast<uint32_t>(xs[p]) | (static_cast<uint32_t>(ys[q]) << 16);
}

struct GridArgs {
  uint8_t* occ;
  uint32_t pitch;
  DiseaseParams dz;
};

// The agent's grid byte: 1 occupied, 2 / 3 infectious in the regular / high window.
__global__ __launch_bounds__(kInitBlock) void k_init_grid(InitArgs a, GridArgs g) {
  const uint32_t i = blockIdx.x * kInitBlock + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t p = a.pos[i], w = a.st[i];
  g.occ[static_cast<size_t>((p >> 16) + kPadY) * g.pitch + (p & 0xFFFFu) + kPadX] = occupant_byte(g.dz, w);
}

void set_digit(Composite& c, uint32_t shift, uint32_t d) {
  if (shift >= 32) {
    c.key |= static_cast<unsigned long long>(d) << (shift - 32);
  } else {
    const uint64_t v = static_cast<uint64_t>(d) << shift;
    c.id |= static_cast<uint32_t>(v);
    c.key |= v >> 32;
  }
}

struct DeviceTemp {
  std::vector<void*> owned;
  template <class T>
  hipError_t alloc(T** p, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) owned.push_back(q);
    *p = static_cast<T*>(q);
    return e;
  }
  ~DeviceTemp() {
    for (void* p : owned) (void)hipFree(p);
  }
};

#define INIT_TRY(step, expr)      \
  do {                            \
    hipError_t e_ = (expr);       \
    if (e_ != hipSuccess) {       \
      err = step;                 \
      return e_;                  \
    }                             \
  } while (0)

// The composite of rank r (0-based) among all agents: eight 12-bit passes, each narrowing
// the prefix to the bin that holds rank r.
hipError_t select_rank(const InitArgs& a, uint32_t r, uint32_t* d_hist, uint32_t grid, hipStream_t stream,
                       Composite& out, std::string& err) {
  Composite prefix{0, 0}, mask{0, 0};
  std::vector<uint32_t> h(kSelectBins);
  for (int shift = static_cast<int>(kSelectTopShift); shift >= 0; shift -= static_cast<int>(kSelectBits)) {
    INIT_TRY("select: clear histogram", hipMemsetAsync(d_hist, 0, kSelectBins * sizeof(uint32_t), stream));
    hipLaunchKernelGGL(k_select_hist, dim3(grid), dim3(kInitBlock), 0, stream, a, prefix, mask,
                       static_cast<uint32_t>(shift), d_hist);
    INIT_TRY("select: k_select_hist", hipGetLastError());
    INIT_TRY("select: read histogram",
             hipMemcpyAsync(h.data(), d_hist, kSelectBins * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    INIT_TRY("select: sync", hipStreamSynchronize(stream));
    uint32_t b = 0;
    while (b < kSelectBins && r >= h[b]) r -= h[b++];
    if (b == kSelectBins) {
      err = "select: rank beyond the population";
      return hipErrorInvalidValue;
    }
    set_digit(prefix, static_cast<uint32_t>(shift), b);
    set_digit(mask, static_cast<uint32_t>(shift), kSelectBins - 1);
  }
  out = prefix;
  return hipSuccess;
}

// ---- office layout slot maps (model.hpp): Normal and Essential workers office by office
// (office j = id % n_offices), each office's in id order, then the other agents in id order.
__device__ __forceinline__ bool office_worker(uint32_t w) {
  const uint32_t ws = (w >> B_WS) & M_WS;
  return ws == WS_NORMAL || ws == WS_ESSENTIAL;
}

// One thread per office: its workers (ids j, j + O, ... in order, coalesced over offices).
__global__ __launch_bounds__(kInitBlock) void k_office_count(const uint32_t* st, uint32_t n, uint32_t O,
                                                             uint32_t* cnt) {
  const uint32_t j = blockIdx.x * kInitBlock + threadIdx.x;
  if (j >= O) return;
  uint32_t c = 0;
  for (uint64_t id = j; id < n; id += O) c += office_worker(st[id]) ? 1u : 0u;
  cnt[j] = c;
}

__global__ __launch_bounds__(kInitBlock) void k_office_place(const uint32_t* st, uint32_t n, uint32_t O,
                                                             const uint32_t* base, uint32_t* office_agent,
                                                             uint32_t* office_slot) {
  const uint32_t j = blockIdx.x * kInitBlock + threadIdx.x;
  if (j >= O) return;
  uint32_t slot = base[j];
  for (uint64_t id = j; id < n; id += O)
    if (office_worker(st[id])) {
      office_agent[slot] = static_cast<uint32_t>(id);
      office_slot[id] = slot;
      ++slot;
    }
}

constexpr uint32_t kRestChunk = 1024;  // ids per block of the other agents' passes

// Other agents per chunk of kRestChunk ids.
__global__ __launch_bounds__(kRestChunk) void k_rest_count(const uint32_t* st, uint32_t n, uint32_t* cnt) {
  const uint32_t id = blockIdx.x * kRestChunk + threadIdx.x;
  const int rest = id < n && !office_worker(st[id]);
  const int c = __syncthreads_count(rest);
  if (threadIdx.x == 0) cnt[blockIdx.x] = static_cast<uint32_t>(c);
}

// Slot of each other agent: n_workers + the chunk's base + its rank in the chunk (wave
// ballots, then the waves' totals in LDS).
__global__ __launch_bounds__(kRestChunk) void k_rest_place(const uint32_t* st, uint32_t n, uint32_t n_workers,
                                                          const uint32_t* base, uint32_t* office_agent,
                                                          uint32_t* office_slot) {
  __shared__ uint32_t wave_total[kRestChunk / 64];
  const uint32_t id = blockIdx.x * kRestChunk + threadIdx.x;
  const bool rest = id < n && !office_worker(st[id]);
  const unsigned long long m = __ballot(rest);
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t below = static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)));
  if (lane == 0) wave_total[wave] = static_cast<uint32_t>(__popcll(m));
  __syncthreads();
  uint32_t before = 0;
  for (uint32_t w = 0; w < wave; ++w) before += wave_total[w];
  if (rest) {
    const uint32_t slot = n_workers + base[blockIdx.x] + before + below;
    office_agent[slot] = id;
    office_slot[id] = slot;
  }
}

}  // namespace

hipError_t synthesize_on_device(const Config& cfg, uint64_t seed, const Geometry& geo, const InitTargets& t,
                                hipStream_t stream, std::string& err) {
  const uint32_t N = cfg.number_of_agents;
  INIT_TRY("clear grid", hipMemsetAsync(t.occ, 0, t.cells, stream));
  if (N == 0) return hipSuccess;
  DeviceTemp tmp;
  const uint32_t blocks = (N + kInitBlock - 1) / kInitBlock;
  InitArgs a{};
  a.n = N;
  a.seed_lo = static_cast<uint32_t>(seed);
  a.seed_hi = static_cast<uint32_t>(seed >> 32);
  a.thr_work = bernoulli_threshold(cfg.working_percentage);
  a.thr_pt = bernoulli_threshold(cfg.public_transport_percentage);
  a.thr_staff = bernoulli_threshold(0.002);  // HOSPITAL_STAFF_PERCENTAGE
  a.thr_ess = bernoulli_threshold(cfg.has_lockdown ? cfg.essential_workers_population : 0.0);
  a.pos = t.pos;
  a.st = t.st;
  a.tm = t.tm;
  a.wsa = t.wsa;
  INIT_TRY("alloc block counts", tmp.alloc(&a.block_pt, blocks));
  if (cfg.csv_population) {
    uint8_t* cols = nullptr;
    INIT_TRY("alloc csv columns", tmp.alloc(&cols, 2 * static_cast<size_t>(N)));
    INIT_TRY("upload csv working",
             hipMemcpyAsync(cols, cfg.csv_working.data(), N, hipMemcpyHostToDevice, stream));
    INIT_TRY("upload csv pub_transport",
             hipMemcpyAsync(cols + N, cfg.csv_pub_transport.data(), N, hipMemcpyHostToDevice, stream));
    a.csv_working = cols;
    a.csv_pt = cols + N;
  }
  hipLaunchKernelGGL(k_init_agents, dim3(blocks), dim3(kInitBlock), 0, stream, a);
  INIT_TRY("k_init_agents", hipGetLastError());
  // public transport users are capped in id order; Population::Csv records keep their flag
  if (!cfg.csv_population) {
    a.pt_cap = geo.n_transport_points;
    hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, stream, a.block_pt, blocks);
    INIT_TRY("k_scan_exclusive", hipGetLastError());
    hipLaunchKernelGGL(k_init_pt_cap, dim3(blocks), dim3(kInitBlock), 0, stream, a);
    INIT_TRY("k_init_pt_cap", hipGetLastError());
  }
  // starting infections
  const uint32_t c[4] = {cfg.si_exposed, cfg.si_exposed + cfg.si_asymptomatic,
                         cfg.si_exposed + cfg.si_asymptomatic + cfg.si_mild,
                         cfg.si_exposed + cfg.si_asymptomatic + cfg.si_mild + cfg.si_severe};
  if (c[3] > 0) {
    uint32_t* d_hist = nullptr;
    INIT_TRY("alloc histogram", tmp.alloc(&d_hist, kSelectBins));
    const uint32_t grid = std::min(blocks, 512u);
    Bounds b{};
    for (int j = 0; j < 4; ++j) {
      if (c[j] == 0) {
        b.zero |= 1u << j;
      } else if (j > 0 && c[j] == c[j - 1]) {
        b.bound[j] = b.bound[j - 1];
      } else {
        hipError_t e = select_rank(a, c[j] - 1, d_hist, grid, stream, b.bound[j], err);
        if (e != hipSuccess) return e;
      }
    }
    hipLaunchKernelGGL(k_init_infect, dim3(blocks), dim3(kInitBlock), 0, stream, a, b);
    INIT_TRY("k_init_infect", hipGetLastError());
  }
  // start cells, then the grid
  const HouseLayout L{geo.house_sx, geo.house_sy, std::max(geo.house_nx, 1), geo.n_houses};
  const uint32_t houses = std::min(geo.n_houses, N);
  if (houses) {
    hipLaunchKernelGGL(k_init_cells, dim3((houses + kInitBlock - 1) / kInitBlock), dim3(kInitBlock), 0, stream, a,
                       L);
    INIT_TRY("k_init_cells", hipGetLastError());
  }
  const GridArgs g{t.occ, t.pitch, Disease::params(cfg)};
  hipLaunchKernelGGL(k_init_grid, dim3(blocks), dim3(kInitBlock), 0, stream, a, g);
  INIT_TRY("k_init_grid", hipGetLastError());
  // the temporaries are freed on return: finish with them first
  INIT_TRY("init sync", hipStreamSynchronize(stream));
  return hipSuccess;
}

hipError_t build_office_maps_on_device(const uint32_t* st, uint32_t n, uint32_t n_offices, uint32_t* office_agent,
                                      uint32_t* office_slot, std::vector<uint32_t>& workers_per_office,
                                      uint32_t& n_workers, hipStream_t stream, std::string& err) {
  const uint32_t O = std::max(n_offices, 1u);
  workers_per_office.assign(O, 0);
  n_workers = 0;
  if (n == 0) return hipSuccess;
  const uint32_t chunks = (n + kRestChunk - 1) / kRestChunk;
  DeviceTemp tmp;
  uint32_t *d_cnt = nullptr, *d_base = nullptr, *d_rcnt = nullptr, *d_rbase = nullptr;
  INIT_TRY("office maps: alloc", tmp.alloc(&d_cnt, O));
  INIT_TRY("office maps: alloc", tmp.alloc(&d_base, O));
  INIT_TRY("office maps: alloc", tmp.alloc(&d_rcnt, chunks));
  INIT_TRY("office maps: alloc", tmp.alloc(&d_rbase, chunks));
  const uint32_t ob = (O + kInitBlock - 1) / kInitBlock;
  hipLaunchKernelGGL(k_office_count, dim3(ob), dim3(kInitBlock), 0, stream, st, n, O, d_cnt);
  INIT_TRY("k_office_count", hipGetLastError());
  hipLaunchKernelGGL(k_rest_count, dim3(chunks), dim3(kRestChunk), 0, stream, st, n, d_rcnt);
  INIT_TRY("k_rest_count", hipGetLastError());
  std::vector<uint32_t> rest(chunks), base(O), rbase(chunks);
  INIT_TRY("office maps: read counts", hipMemcpyAsync(workers_per_office.data(), d_cnt, O * sizeof(uint32_t),
                                                       hipMemcpyDeviceToHost, stream));
  INIT_TRY("office maps: read counts",
           hipMemcpyAsync(rest.data(), d_rcnt, chunks * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  INIT_TRY("office maps: sync", hipStreamSynchronize(stream));
  for (uint32_t j = 0; j < O; ++j) {
    base[j] = n_workers;
    n_workers += workers_per_office[j];
  }
  for (uint32_t c = 0, acc = 0; c < chunks; ++c) {
    rbase[c] = acc;
    acc += rest[c];
  }
  INIT_TRY("office maps: write bases",
           hipMemcpyAsync(d_base, base.data(), O * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
  INIT_TRY("office maps: write bases",
           hipMemcpyAsync(d_rbase, rbase.data(), chunks * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL(k_office_place, dim3(ob), dim3(kInitBlock), 0, stream, st, n, O, d_base, office_agent,
                     office_slot);
  INIT_TRY("k_office_place", hipGetLastError());
  hipLaunchKernelGGL(k_rest_place, dim3(chunks), dim3(kRestChunk), 0, stream, st, n, n_workers, d_rbase,
                     office_agent, office_slot);
  INIT_TRY("k_rest_place", hipGetLastError());
  // the temporaries and the host bases are freed on return: finish with them first
  INIT_TRY("office maps: sync", hipStreamSynchronize(stream));
  return hipSuccess;
}

}  // namespace epi
