// Population synthesis in deterministic mode.
//
// Reference flow (engine/src/geography/grid.rs:83-155, engine/src/citizen/citizen_factory.rs:31-134,
// engine/src/epidemiology_simulation.rs:75-192), with every rand draw replaced by a
// fixed Philox counter at hour 0xFFFFFFFF:
//   agent i: working ~ Bernoulli(wp); home = houses[i % H]; office = offices[i % O];
//            public transport ~ Bernoulli(pt) && working && users < transport points;
//            working -> HospitalStaff{8} ~ Bernoulli(0.002) else Normal; else NA;
//            immunity ~ uniform{-2..2}; later Normal -> Essential ~ Bernoulli(essential%).
//   starting infections: the `total` agents with the smallest 64-bit Philox keys, in key
//            order, take exposed / asymptomatic / mild / severe (replaces choose_multiple).
//   start cells: house h's agents h, h+H, ... take random_points(k) of the 2x2 house in
//            id order (x and y choices are partial Fisher-Yates draws keyed by house).
#include "population.hpp"

#include <algorithm>
#include <cmath>
#include <thread>

#include "philox.hpp"

namespace epi {

uint64_t bernoulli_threshold(double p) {
  if (p == 1.0) return kAlways;
  return static_cast<uint64_t>(p * 18446744073709551616.0);
}

static Rect building(int32_t sx, int32_t sy, int32_t nx, int32_t size, uint32_t k) {
  int32_t x0 = sx + static_cast<int32_t>(k % static_cast<uint32_t>(nx)) * size;
  int32_t y0 = sy + static_cast<int32_t>(k / static_cast<uint32_t>(nx)) * size;
  return Rect{x0, y0, x0 + size - 1, y0 + size - 1};
}

bool build_geometry(const Config& cfg, Geometry& geo, std::string& err) {
  const uint32_t g = cfg.grid_size;
  if (g == 0 || g > 32000) {
    err = "grid_size must be in 1..=32000";
    return false;
  }
  geo.g = g;
  // geography/mod.rs:33-48 (f64 ceil of the relative strip widths)
  const int32_t home_w = static_cast<int32_t>(std::ceil(static_cast<double>(g) * 0.4));
  const int32_t t_end = home_w + static_cast<int32_t>(std::ceil(static_cast<double>(g) * 0.2));
  const int32_t w_end = t_end + static_cast<int32_t>(std::ceil(static_cast<double>(g) * 0.2));
  const int32_t h_end = w_end + static_cast<int32_t>(std::ceil(static_cast<double>(g) * 0.1));
  const int32_t gy = static_cast<int32_t>(g);
  geo.housing = Rect{0, 0, home_w - 1, gy};
  geo.transport = Rect{home_w, 0, t_end - 1, gy};
  geo.work = Rect{t_end, 0, w_end - 1, gy};
  geo.hosp0 = Rect{w_end, 0, h_end - 1, gy};
  // area_factory (area.rs:95-117): 2x2 houses, 10x10 offices, row-major
  geo.house_sx = geo.housing.x0;
  geo.house_sy = geo.housing.y0;
  geo.house_nx = std::max(0, (geo.housing.x1 - geo.housing.x0 + 1) / 2);
  geo.house_ny = std::max(0, (geo.housing.y1 - geo.housing.y0 + 1) / 2);
  geo.office_sx = geo.work.x0;
  geo.office_sy = geo.work.y0;
  geo.office_nx = std::max(0, (geo.work.x1 - geo.work.x0 + 1) / 10);
  geo.office_ny = std::max(0, (geo.work.y1 - geo.work.y0 + 1) / 10);
  geo.n_houses = static_cast<uint32_t>(geo.house_nx) * static_cast<uint32_t>(geo.house_ny);
  geo.n_offices = static_cast<uint32_t>(geo.office_nx) * static_cast<uint32_t>(geo.office_ny);
  // resize_hospital(N, 0.002, beds%) (grid.rs:240-261); cells use (x1-x0)*(y1-y0) (area.rs:90-92)
  const double n = static_cast<double>(static_cast<int32_t>(cfg.number_of_agents));
  const double beds_f = std::ceil(n * cfg.hospital_beds_percentage + n * 0.002);
  const uint32_t beds = beds_f <= 0.0 ? 0u : (beds_f >= 4294967295.0 ? 0xFFFFFFFFu : static_cast<uint32_t>(beds_f));
  const uint32_t cells = static_cast<uint32_t>((geo.hosp0.x1 - geo.hosp0.x0) * (geo.hosp0.y1 - geo.hosp0.y0));
  if (beds <= cells) {
    const uint32_t w = static_cast<uint32_t>(geo.hosp0.x1 - geo.hosp0.x0);
    if (w == 0) {
      err = "attempt to divide by zero";
      return false;
    }
    geo.hosp0.y1 = static_cast<int32_t>(beds / w);
  }
  // increase_hospital_size (grid.rs:233-238), applied later by BuildNewHospital
  geo.hosp1 = Rect{geo.hosp0.x0, geo.hosp0.y0, gy, gy};
  const int32_t max_x = std::max({gy, geo.hosp0.x1, geo.hosp1.x1});
  geo.stride = static_cast<uint32_t>(max_x) + 1;
  geo.height = g + 1;
  // transport points: ceil(N (pt+0.1)(wp+0.1)) requested; random_points draws ceil(sqrt)
  // distinct xs and ys (f32 sqrt, area.rs:64-74) so the count is capped by the strip.
  const double req_f = static_cast<double>(cfg.number_of_agents) * (cfg.public_transport_percentage + 0.1) *
                       (cfg.working_percentage + 0.1);
  const uint64_t req = static_cast<uint64_t>(std::ceil(req_f));
  const uint64_t side = static_cast<uint64_t>(std::ceil(std::sqrt(static_cast<float>(req))));
  const uint64_t lx = static_cast<uint64_t>(geo.transport.x1 - geo.transport.x0 + 1);
  const uint64_t ly = static_cast<uint64_t>(geo.transport.y1 - geo.transport.y0 + 1);
  const uint64_t avail = std::min(side, lx) * std::min(side, ly);
  geo.n_transport_points = static_cast<uint32_t>(std::min(req, avail));
  return true;
}

namespace {

struct SelKey {
  uint64_t key;
  uint32_t id;
  bool operator<(const SelKey& o) const { return key != o.key ? key < o.key : id < o.id; }
};

template <class F>
void parallel_for(uint32_t n, F&& f) {
  const uint32_t min_chunk = 1u << 16;
  unsigned t = std::max(1u, std::min(std::thread::hardware_concurrency(), 16u));
  if (n < 2 * min_chunk || t == 1) {
    f(0u, n);
    return;
  }
  t = std::min<unsigned>(t, (n + min_chunk - 1) / min_chunk);
  std::vector<std::thread> th;
  const uint32_t chunk = (n + t - 1) / t;
  for (unsigned k = 0; k < t; ++k) {
    uint32_t b = k * chunk, e = std::min(n, b + chunk);
    if (b < e) th.emplace_back([&f, b, e] { f(b, e); });
  }
  for (auto& x : th) x.join();
}

inline uint32_t pack_pos(int32_t x, int32_t y) {
  return static_cast<uint32_t>(x) | (static_cast<uint32_t>(y) << 16);
}

// Partial Fisher-Yates over {lo, lo+1} driven by fixed words (rand's choose_multiple on a
// two-element range, restated): returns min(m, 2) values.
uint32_t choose2(int32_t lo, uint32_t m, const uint32_t* w, int32_t* out) {
  int32_t a[2] = {lo, lo + 1};
  if (m > 2) m = 2;
  for (uint32_t t = 0; t < m; ++t) {
    uint32_t r = t + below(w[t], 2 - t);
    std::swap(a[t], a[r]);
  }
  for (uint32_t t = 0; t < m; ++t) out[t] = a[t];
  return m;
}

}  // namespace

bool prepare_population(const Config& cfg, Population& pop, std::string& err) {
  if (!build_geometry(cfg, pop.geo, err)) return false;
  const Geometry& G = pop.geo;
  const uint32_t N = cfg.number_of_agents;
  const uint32_t H = G.n_houses, O = G.n_offices;
  if (N > 0 && (H == 0 || O == 0)) {
    err = "attempt to calculate the remainder with a divisor of zero";
    return false;
  }
  const uint32_t total_si = cfg.si_exposed + cfg.si_asymptomatic + cfg.si_mild + cfg.si_severe;
  if (total_si > N) {
    err = "There are " + std::to_string(total_si) + " people set to infect, but only " + std::to_string(N) +
          " agents available";
    return false;
  }
  if (cfg.csv_population && N > 4u * H) {  // grid.rs:212-219
    err = "Cannot accommodate citizens into homes! There are " + std::to_string(N) + " citizens, but " +
          std::to_string(4u * H) + " home points";
    return false;
  }
  if (H > 0 && N > 4u * H) {
    err = "There are " + std::to_string((N + H - 1) / H) +
          " agents assigned to a house, but house capacity is 4";
    return false;
  }
  pop.counts0[0] = 0;
  pop.counts0[1] = N - total_si;
  pop.counts0[2] = cfg.si_exposed;
  pop.counts0[3] = total_si - cfg.si_exposed;
  return true;
}

bool synthesize_population(const Config& cfg, uint64_t seed, Population& pop, std::string& err) {
  if (!prepare_population(cfg, pop, err)) return false;
  const Geometry& G = pop.geo;
  const uint32_t N = cfg.number_of_agents;
  const uint32_t H = G.n_houses;
  const uint32_t total_si = cfg.si_exposed + cfg.si_asymptomatic + cfg.si_mild + cfg.si_severe;
  pop.pos.assign(N, 0);
  pop.st.assign(N, 0);
  pop.tm.assign(N, 0);
  pop.wsa.assign(N, 0);
  const uint64_t thr_work = bernoulli_threshold(cfg.working_percentage);
  const uint64_t thr_pt = bernoulli_threshold(cfg.public_transport_percentage);
  const uint64_t thr_staff = bernoulli_threshold(0.002);  // HOSPITAL_STAFF_PERCENTAGE
  const uint64_t thr_ess = bernoulli_threshold(cfg.has_lockdown ? cfg.essential_workers_population : 0.0);
  std::vector<uint8_t> pt_draw(N);
  std::vector<SelKey> keys(total_si ? N : 0);

  parallel_for(N, [&](uint32_t b, uint32_t e) {
    for (uint32_t i = b; i < e; ++i) {
      const u32x4 a = draw(seed, i, kInitHour, SITE_INIT_A, 0);
      const u32x4 s = draw(seed, i, kInitHour, SITE_INIT_B, 0);
      const u32x4 c = draw(seed, i, kInitHour, SITE_INIT_C, 0);
      // Population::Csv takes working / public transport from the record
      // (Citizen::from_record, citizen/mod.rs:155-180); Auto draws them.
      const bool working = cfg.csv_population ? cfg.csv_working[i] != 0 : bernoulli(thr_work, join64(a.w[0], a.w[1]));
      pt_draw[i] = cfg.csv_population ? cfg.csv_pub_transport[i] != 0
                                      : working && bernoulli(thr_pt, join64(a.w[2], a.w[3]));
      uint32_t ws;
      if (!working) ws = WS_NA;
      else if (bernoulli(thr_staff, join64(s.w[0], s.w[1]))) ws = WS_STAFF;
      else ws = bernoulli(thr_ess, join64(c.w[0], c.w[1])) ? WS_ESSENTIAL : WS_NORMAL;
      const uint32_t imm = below(s.w[2], 5);  // IMMUNITY_RANGE index; stored as immunity + 2
      pop.st[i] = (ST_S << B_STATE) | (ws << B_WS) | (AREA_HOME << B_AREA) | (imm << B_IMM);
      pop.wsa[i] = ws == WS_STAFF ? 8u : 0u;  // HospitalStaff { work_start_at: ROUTINE_WORK_TIME }
      if (total_si) keys[i] = SelKey{join64(c.w[2], c.w[3]), i};
    }
  });
  // public transport users are capped in id order (citizen_factory.rs:72-83); records of a
  // Population::Csv keep their flag uncapped
  uint32_t users = 0;
  for (uint32_t i = 0; i < N; ++i)
    if (pt_draw[i] && (cfg.csv_population || users < G.n_transport_points)) {
      pop.st[i] |= F_PT;
      ++users;
    }
  // starting infections
  if (total_si) {
    std::partial_sort(keys.begin(), keys.begin() + total_si, keys.end());
    uint32_t k = 0;
    auto set = [&](uint32_t count, uint32_t state, uint32_t sev, uint32_t day) {
      for (uint32_t j = 0; j < count; ++j, ++k) {
        uint32_t& w = pop.st[keys[k].id];
        w = (w & ~((M_STATE << B_STATE) | (M_SEV << B_SEV) | (M_DAY << B_DAY))) | (state << B_STATE) |
            (sev << B_SEV) | (day << B_DAY);
        pop.tm[keys[k].id] = 0;  // Exposed { at_hour: 0 }
      }
    };
    set(cfg.si_exposed, ST_E, 0, 0);
    set(cfg.si_asymptomatic, ST_I, SEV_ASYM, 1);
    set(cfg.si_mild, ST_I, SEV_MILD, 1);
    set(cfg.si_severe, ST_I, SEV_SEVERE, 1);
  }
  // start cells: random_points(k) of each house, handed out in id order
  const uint32_t n_houses_used = std::min(H, N);
  parallel_for(n_houses_used, [&](uint32_t b, uint32_t e) {
    for (uint32_t h = b; h < e; ++h) {
      const uint32_t k = (N - h + H - 1) / H;
      const Rect hr = building(G.house_sx, G.house_sy, G.house_nx, 2, h);
      const u32x4 w = draw(seed, h, kInitHour, SITE_INIT_HOUSE, 0);
      const uint32_t side = static_cast<uint32_t>(std::ceil(std::sqrt(static_cast<float>(k))));
      int32_t xs[2], ys[2];
      const uint32_t mx = choose2(hr.x0, side, &w.w[0], xs);
      const uint32_t my = choose2(hr.y0, side, &w.w[2], ys);
      uint32_t j = 0;
      for (uint32_t a = 0; a < mx && j < k; ++a)
        for (uint32_t c = 0; c < my && j < k; ++c, ++j) pop.pos[h + j * H] = pack_pos(xs[a], ys[c]);
    }
  });
  return true;
}

}  // namespace epi
