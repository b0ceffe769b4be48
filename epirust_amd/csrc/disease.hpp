// The disease model of the hour step, as a handler type fixed when the library is built.
//
// The reference makes the disease model a plug-in: the DiseaseHandler trait
// (engine/src/state_machine/disease_handler.rs:27-44) is the generic parameter D of
// Epidemiology (epidemiology_simulation.rs:59-72) and of EngineApp::start_standalone
// (engine_app.rs:89-125), resolved by the Rust compiler, and its one implementation is
// Disease (state_machine/default_disease_handler.rs:31-103). Here the handler is likewise a
// type resolved by the compiler: the step's kernels call its static functions, which inline
// into them, so a handler costs nothing at run time. The handler compiled in is
// EPI_DISEASE_HANDLER from the header EPI_DISEASE_HEADER when both are defined on the
// compiler's command line (epirust_amd.build's --disease-handler option does that), else
// DefaultDisease below.
//
// A handler is a struct with these static members, the trait's methods restated for the
// packed state word (model.hpp: state, severity, infection day, immunity + 2, hospitalized
// and vaccinated / work-quarantined flags) and the Exposed / Pre at_hour word `tm`:
//
//   static DiseaseParams params(const Config& c);                       host: per-simulation constants
//   EPI_HD static uint32_t transmission_class(const DiseaseParams&, uint32_t w);
//       0, 1 or 2 for an infected, not hospitalized occupant: the transmission window of its
//       rate (Citizen::get_infection_transmission_rate, citizen/mod.rs:182-185); it is kept in
//       the cell's grid byte, which the neighbour scan of on_susceptible reads
//   EPI_HD static uint64_t transmission_threshold(const DiseaseParams&, uint32_t cls);
//       Bernoulli threshold of one neighbour of class 1 or 2 (philox.hpp: bernoulli)
//   EPI_HD static bool can_catch(uint32_t w);                            on_susceptible's gate
//   EPI_HD static void on_caught(uint32_t hour, uint32_t& w, uint32_t& tm);  on_susceptible's Some(state)
//   EPI_HD static bool timed(uint32_t w);      the state carries an at_hour in tm (tm is 0 otherwise)
//   EPI_HD static bool draws_needed(uint32_t w, uint32_t hour_of_day);   the update reads `draws`
//   EPI_HD static void on_exposed(const DiseaseParams&, uint32_t hour, const u32x4& draws, uint32_t& w, uint32_t& tm);
//   EPI_HD static void on_infected(const DiseaseParams&, uint32_t hour, const u32x4& draws, uint32_t& w, uint32_t& tm);
//   EPI_HD static void on_routine_end(const DiseaseParams&, const u32x4& draws, uint32_t& w, uint32_t& tm);
//   EPI_HD static bool is_to_be_hospitalize(const DiseaseParams&, uint32_t w);
//
// on_susceptible itself is the step's: a scan of the eight Moore neighbours of the chosen cell
// in the current area over the hour's snapshot, one Bernoulli per infectious neighbour with
// transmission_threshold of its class, first success infects
// (default_disease_handler.rs:64-86). `draws` is the agent's Philox block of the hour for the
// disease (model.hpp: SITE_DISEASE), computed only where draws_needed says so; the RandomWrapper
// of the trait's methods becomes these four words, so a handler's outcome depends only on
// (seed, agent, hour). A class-0 neighbour never transmits. tm is read from memory only for
// timed states and written back only when a call changed it, so a handler that leaves a
// timed state sets tm to 0. on_exposed and
// on_infected run at hours of day 7..22 for agents that can update (citizen/mod.rs:199-208),
// on_routine_end and is_to_be_hospitalize at the hours the routine calls them
// (citizen/mod.rs:351-365, 397-413).
#pragma once
#include <cstdint>

#include "config.hpp"
#include "model.hpp"
#include "philox.hpp"

namespace epi {

// rand 0.8.5 Bernoulli threshold: floor(p * 2^64), p == 1 -> always (population.cpp).
uint64_t bernoulli_threshold(double p);

EPI_HD uint32_t dz_state(uint32_t w) { return (w >> B_STATE) & M_STATE; }
EPI_HD uint32_t dz_sev(uint32_t w) { return (w >> B_SEV) & M_SEV; }
EPI_HD uint32_t dz_day(uint32_t w) { return (w >> B_DAY) & M_DAY; }
EPI_HD uint32_t dz_set(uint32_t w, uint32_t b, uint32_t m, uint32_t v) { return (w & ~(m << b)) | ((v & m) << b); }

// Disease as the reference's DiseaseHandler (state_machine/default_disease_handler.rs:31-103).
struct DefaultDisease {
  static DiseaseParams params(const Config& c) {
    DiseaseParams P{};
    P.regular_start = c.regular_transmission_start_day;
    P.high_start = c.high_transmission_start_day;
    P.last_day = c.last_day;
    P.exposed_duration = c.exposed_duration;
    P.pre_symptomatic_duration = c.pre_symptomatic_duration;
    P.thr_regular = bernoulli_threshold(c.regular_transmission_rate);
    P.thr_high = bernoulli_threshold(c.high_transmission_rate);
    P.thr_death = bernoulli_threshold(c.death_rate);
    P.thr_symptoms = bernoulli_threshold(1.0 - c.percentage_asymptomatic_population);
    P.thr_severe = bernoulli_threshold(c.percentage_severe_infected_population);
    // is_to_be_hospitalized: rate(day) >= high_transmission_rate (disease/mod.rs:97-99),
    // one bit per window, compared in f64 as the reference does
    if (0.0 >= c.high_transmission_rate) P.hosp_window |= 1u;
    if (c.regular_transmission_rate >= c.high_transmission_rate) P.hosp_window |= 2u;
    if (c.high_transmission_rate >= c.high_transmission_rate) P.hosp_window |= 4u;
    return P;
  }

  // Window of (infection_day + immunity) as u32 (citizen/mod.rs:182-185, disease/mod.rs:88-95):
  // 1 regular, 2 high, 0 rate 0.0. Not infected: day 0.
  EPI_HD static uint32_t transmission_class(const DiseaseParams& P, uint32_t w) {
    const uint32_t day = dz_state(w) == ST_I ? dz_day(w) : 0u;
    const uint32_t d = static_cast<uint32_t>(static_cast<int32_t>(day) + static_cast<int32_t>((w >> B_IMM) & M_IMM) - 2);
    if (P.regular_start < d && d <= P.high_start) return 1u;
    if (P.high_start < d && d <= P.last_day) return 2u;
    return 0u;
  }

  EPI_HD static uint64_t transmission_threshold(const DiseaseParams& P, uint32_t cls) {
    return cls == 2u ? P.thr_high : P.thr_regular;
  }

  // !work_quarantined && !is_vaccinated (default_disease_handler.rs:72)
  EPI_HD static bool can_catch(uint32_t w) { return !(w & (F_WQ | F_VACC)); }

  // The state on_susceptible returns when a neighbour transmits: Exposed{at_hour: hour}
  // (default_disease_handler.rs:81-83).
  EPI_HD static void on_caught(uint32_t hour, uint32_t& w, uint32_t& tm) {
    w = dz_set(w, B_STATE, M_STATE, ST_E);
    tm = hour;
  }

  // The states that carry an at_hour in tm: Exposed and Infected{Pre} (state.rs:22-74).
  EPI_HD static bool timed(uint32_t w) {
    return dz_state(w) == ST_E || (dz_state(w) == ST_I && dz_sev(w) == SEV_PRE);
  }

  // Whether the update at hour of day hd reads `draws`: on_exposed and on_infected of Pre
  // at 7..22, on_routine_end of Severe at 23.
  EPI_HD static bool draws_needed(uint32_t w, uint32_t hd) {
    const uint32_t s = dz_state(w), v = dz_sev(w);
    return hd >= 7 && (s == ST_E || (s == ST_I && (v == SEV_PRE || (hd == 23 && v == SEV_SEVERE))));
  }

  // default_disease_handler.rs:52-62: random factor from draws.w[0] over {-1, 0, 1}, the
  // symptoms Bernoulli from (w[2], w[3]).
  EPI_HD static void on_exposed(const DiseaseParams& P, uint32_t hour, const u32x4& d, uint32_t& w, uint32_t& tm) {
    const int32_t rf = static_cast<int32_t>(below(d.w[0], 3)) - 1;  // RANGE_FOR_EXPOSED
    if (hour - tm >= static_cast<uint32_t>(static_cast<int32_t>(P.exposed_duration) + rf)) {
      const bool symptoms = bernoulli(P.thr_symptoms, join64(d.w[2], d.w[3]));
      w = dz_set(w, B_STATE, M_STATE, ST_I);
      w = dz_set(w, B_DAY, M_DAY, 0);
      w = dz_set(w, B_SEV, M_SEV, symptoms ? SEV_PRE : SEV_ASYM);
      tm = symptoms ? hour : 0u;
    }
  }

  // default_disease_handler.rs:41-50: Pre{at_hour} becomes Severe or Mild after the
  // pre-symptomatic duration; other severities have no transition.
  EPI_HD static void on_infected(const DiseaseParams& P, uint32_t hour, const u32x4& d, uint32_t& w, uint32_t& tm) {
    if (dz_sev(w) != SEV_PRE) return;
    if (hour - tm >= P.pre_symptomatic_duration) {
      w = dz_set(w, B_SEV, M_SEV, bernoulli(P.thr_severe, join64(d.w[2], d.w[3])) ? SEV_SEVERE : SEV_MILD);
      tm = 0;
    }
  }

  // default_disease_handler.rs:88-102 with the constants ASYMPTOMATIC_LAST_DAY = 9 and
  // MILD_INFECTED_LAST_DAY = 12 (engine/src/constants.rs); the death draw from (w[2], w[3]).
  EPI_HD static void on_routine_end(const DiseaseParams& P, const u32x4& d, uint32_t& w, uint32_t& tm) {
    if (dz_state(w) != ST_I) return;
    const uint32_t day = dz_day(w), sv = dz_sev(w);
    uint32_t ns = ST_I;
    if (sv == SEV_ASYM && day == 9u) ns = ST_R;
    else if (sv == SEV_MILD && day == 12u) ns = ST_R;
    else if (sv == SEV_SEVERE && day == P.last_day) ns = bernoulli(P.thr_death, join64(d.w[2], d.w[3])) ? ST_D : ST_R;
    if (ns != ST_I) {
      w = dz_set(w, B_STATE, M_STATE, ns);
      w = dz_set(w, B_SEV, M_SEV, 0);
      w = dz_set(w, B_DAY, M_DAY, 0);
      tm = 0;
    }
  }

  // default_disease_handler.rs:32-39: Infected and Severe, and the rate of
  // (infection_day + immunity) at least the high rate.
  EPI_HD static bool is_to_be_hospitalize(const DiseaseParams& P, uint32_t w) {
    return dz_state(w) == ST_I && dz_sev(w) == SEV_SEVERE && ((P.hosp_window >> transmission_class(P, w)) & 1u);
  }
};

}  // namespace epi

#if defined(EPI_DISEASE_HEADER) && defined(EPI_DISEASE_HANDLER)
#include EPI_DISEASE_HEADER
namespace epi {
using Disease = EPI_DISEASE_HANDLER;
}
#else
namespace epi {
using Disease = DefaultDisease;
}
#endif

namespace epi {
// Grid byte of an occupied cell: 1, or 1 + the class of an infected, not hospitalized
// occupant (model.hpp).
EPI_HD uint8_t occupant_byte(const DiseaseParams& P, uint32_t w) {
  const bool infectious = dz_state(w) == ST_I && !(w & F_HOSP);
  return static_cast<uint8_t>(1u + (infectious ? Disease::transmission_class(P, w) : 0u));
}
}  // namespace epi
