/*
 * epicpu -- CPU oracle for the per-hour agent step of EpiRust, deterministic mode.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP product is
 * compared against. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it. The product library (libepigpu) never links it.
 *
 * It restates, in plain C, the reference engine's data flow for one standalone
 * engine (reference paths, relative to the reference repository root):
 *   engine/src/epidemiology_simulation.rs:211-274   hour loop, stop rule
 *   engine/src/allocation_map.rs:67-163, 306-387     step, merge, interventions
 *   engine/src/citizen/mod.rs:215-454                routine / movement
 *   engine/src/state_machine/default_disease_handler.rs:31-103  disease handler
 *   common/src/disease/mod.rs:88-107                 transmission windows
 *   engine/src/geography/{mod,area,grid,point}.rs    geography
 *   engine/src/citizen/citizen_factory.rs:31-134     population synthesis
 * with the reference's OS-seeded thread_rng replaced by Philox4x32-10 streams
 * keyed by (seed, agent id, hour, site) and collisions resolved lowest-agent-id-wins.
 */
#ifndef EPICPU_H
#define EPICPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EPICPU_MAX_VACC 4096

/* Flat restatement of the reference Config (common/src/config/mod.rs:44-58). */
typedef struct {
    /* population.Auto (common/src/config/population.rs:34-42) */
    uint32_t number_of_agents;
    double public_transport_percentage;
    double working_percentage;
    /* disease (common/src/disease/mod.rs:26-44) */
    uint32_t regular_transmission_start_day;
    uint32_t high_transmission_start_day;
    uint32_t last_day;
    uint32_t asymptomatic_last_day;  /* parsed, ignored by the engine (constants.rs:48-50) */
    uint32_t mild_infected_last_day; /* parsed, ignored by the engine */
    double regular_transmission_rate;
    double high_transmission_rate;
    double death_rate;
    double percentage_asymptomatic_population;
    double percentage_severe_infected_population;
    uint32_t exposed_duration;
    uint32_t pre_symptomatic_duration;
    /* geography_parameters */
    uint32_t grid_size;
    double hospital_beds_percentage;
    uint32_t hours;
    /* starting_infections (common/src/config/starting_infections.rs:22-28) */
    uint32_t si_infected_mild_asymptomatic;
    uint32_t si_infected_mild_symptomatic;
    uint32_t si_infected_severe;
    uint32_t si_exposed;
    /* interventions (common/src/config/intervention_config.rs:23-48) */
    uint32_t has_lockdown;
    uint32_t lockdown_at_number_of_infections;
    double lockdown_essential_workers_population;
    uint32_t has_build_new_hospital;
    uint32_t spread_rate_threshold;
    uint32_t n_vaccinate;
    uint32_t vaccinate_at_hour[EPICPU_MAX_VACC];
    double vaccinate_percent[EPICPU_MAX_VACC];
    uint64_t seed;
    /* Population::Csv (grid.rs:194-231): when csv_n > 0 the population is csv_n records
       with these working / pub_transport flags (number_of_agents is ignored). */
    uint32_t csv_n;
    const uint8_t* csv_working;
    const uint8_t* csv_pub_transport;
} epicpu_config;

/* Canonical per-agent dump, identical in layout to the product's epi_agents_soa. */
typedef struct {
    int32_t* x;
    int32_t* y;
    uint8_t* state;        /* 0 S, 1 E, 2 I, 3 R, 4 D */
    uint8_t* severity;     /* 0 Pre, 1 Asymptomatic, 2 Mild, 3 Severe (0 unless I) */
    uint32_t* infection_day;
    uint32_t* at_hour;     /* Exposed{at_hour} or Pre{at_hour}, else 0 */
    uint8_t* flags;        /* 1 hospitalized, 2 vaccinated, 4 isolated, 8 work_quarantined, 16 uses_public_transport */
    uint8_t* work_status;  /* 0 Normal, 1 Essential, 2 HospitalStaff, 3 NA */
    uint32_t* work_start_at;
    int8_t* immunity;
    int32_t* current_area; /* 4 per agent: x0 y0 x1 y1 */
} epicpu_agents;

typedef struct epicpu_sim epicpu_sim;

/* Returns NULL on error and writes the message (mirroring the reference panic text). */
epicpu_sim* epicpu_create(const epicpu_config* cfg, char* err, size_t errlen);
void epicpu_destroy(epicpu_sim* s);
/* Reference-order mode (on = 1), for distributional checks only: each agent-hour consumes
   one sequential random stream in the reference's call order (rand 0.8.5 gen_range with
   rejection, IteratorRandom::choose reservoir sampling, short-circuit infection find with a
   gen_bool per infectious neighbour) and the merge visits agents in a random order per
   hour (standing in for hash order) instead of id order. Not bit-comparable with the GPU. */
int epicpu_set_reference_order(epicpu_sim* s, int on);
/* Deterministic-mode draws with the merge of reference-order mode: collisions go to the
   agent visited first in a random order per hour instead of the lowest id. For
   distributional checks of the collision rule alone; not bit-comparable with the GPU. */
int epicpu_set_random_merge(epicpu_sim* s, int on);
/* Threads for the per-agent map (the merge stays sequential); results do not depend on it. */
void epicpu_set_threads(epicpu_sim* s, uint32_t threads);
/* Runs up to n_hours hours; rows gets 7 u32 per hour (hour,S,E,I,H,R,D). */
int epicpu_step(epicpu_sim* s, uint32_t n_hours, uint32_t* rows, uint32_t* n_written, int* stopped);
void epicpu_read_agents(const epicpu_sim* s, epicpu_agents* out);
/* Hour-0 counts (utils/util.rs:45-51) and the current hour. */
void epicpu_counts(const epicpu_sim* s, uint32_t out7[7]);
/* Intervention events: kind 0 vaccination, 1 locked_down, 2 lockdown_revoked, 3 build_new_hospital. */
uint32_t epicpu_events(const epicpu_sim* s, uint32_t* hours, uint32_t* kinds, uint32_t cap);
/* Geometry of the simulation: 5 rects (housing, transport, work, hospital, hospital-now) x0 y0 x1 y1,
   plus n_houses, n_offices and the transport-point count. */
void epicpu_geometry(const epicpu_sim* s, int32_t rects[20], uint32_t counts3[3]);

/* Known-answer test hook: one explicit agent. Codes and flag bits are those of
   epicpu_agents; rects are x0 y0 x1 y1. */
typedef struct {
    int32_t x, y;
    int32_t home[4], work[4], current_area[4];
    uint32_t state, severity, work_status, flags;
    uint32_t infection_day, at_hour, work_start_at;
    int32_t immunity;
} epicpu_agent_init;

/* A simulation with the config's geography, disease and interventions whose population is
   exactly `agents` (ids 0..n-1, no synthesis), having run hour `last_hour`: the next
   epicpu_step simulates last_hour + 1. `hospital` (x0 y0 x1 y1) replaces the resized
   hospital area when not NULL. Returns NULL on an invalid scenario (a position outside
   the (g+1)^2 grid, two agents on one cell). */
epicpu_sim* epicpu_create_scenario(const epicpu_config* cfg, const epicpu_agent_init* agents, uint32_t n,
                                   uint32_t last_hour, const int32_t* hospital, char* err, size_t errlen);

/* Building blocks exposed for known-answer tests. */
void epicpu_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void epicpu_define_geography(uint32_t grid_size, int32_t rects[16]);
uint32_t epicpu_area_factory(int32_t sx, int32_t sy, int32_t ex, int32_t ey, uint32_t size, int32_t* out, uint32_t cap);
void epicpu_resize_hospital(int32_t rect[4], int32_t number_of_agents, double staff_pct, double beds_pct);
void epicpu_increase_hospital_size(int32_t rect[4], uint32_t grid_size);
double epicpu_transmission_rate(const epicpu_config* cfg, uint32_t day);
int epicpu_is_to_be_hospitalized(const epicpu_config* cfg, uint32_t day);
uint32_t epicpu_neighbors_in_area(int32_t px, int32_t py, const int32_t area[4], int32_t* out);
uint32_t epicpu_area_iter(const int32_t area[4], int32_t* out, uint32_t cap);
uint64_t epicpu_bernoulli_threshold(double p);
/* Intervention trigger rules in isolation (lockdown.rs:55-73, hospital.rs:51-74). */
int epicpu_lockdown_should_apply(int locked, uint32_t at_number, uint32_t hour, uint32_t infected);
int epicpu_lockdown_should_unlock(int locked, uint32_t zero_infection_hour, uint32_t hour);
uint32_t epicpu_hospital_counts_updated(uint32_t new_infections, uint32_t hour, uint32_t infected);
int epicpu_hospital_should_apply(int applied, uint32_t has, uint32_t threshold, uint32_t new_infections, uint32_t hour);
/* first vacant hospital cell, given an occupancy predicate array for an area (row-major) */
int epicpu_first_vacant(const int32_t area[4], const int32_t* occupied_xy, uint32_t n_occupied, int32_t out_xy[2]);

#ifdef __cplusplus
}
#endif
#endif
