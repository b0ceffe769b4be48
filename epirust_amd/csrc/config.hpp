// Simulation config: the reference engine's standalone JSON schema
// (common/src/config/mod.rs:44-58 and its sub-structs), parsed without serde.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace epi {

struct VaccinateConfig {
  uint32_t at_hour;
  double percent;
};

struct Config {
  // population: Auto (common/src/config/population.rs:34-42) or Csv (:26-31)
  uint32_t number_of_agents = 0;
  double public_transport_percentage = 0, working_percentage = 0;
  bool csv_population = false;
  std::string csv_file;
  std::vector<uint8_t> csv_working, csv_pub_transport;  // per record, filled by read_population_csv
  // disease (common/src/disease/mod.rs:26-44)
  uint32_t regular_transmission_start_day = 0, high_transmission_start_day = 0, last_day = 0;
  uint32_t asymptomatic_last_day = 0, mild_infected_last_day = 0;
  double regular_transmission_rate = 0, high_transmission_rate = 0, death_rate = 0;
  double percentage_asymptomatic_population = 0, percentage_severe_infected_population = 0;
  uint32_t exposed_duration = 0, pre_symptomatic_duration = 0;
  // geography_parameters (common/src/config/geography_parameters.rs:23-28)
  uint32_t grid_size = 0;
  double hospital_beds_percentage = 0;
  uint32_t hours = 0;
  // starting_infections, default exposed: 1 (starting_infections.rs:65-69)
  uint32_t si_asymptomatic = 0, si_mild = 0, si_severe = 0, si_exposed = 1;
  // interventions (intervention_config.rs:23-48): the first Lockdown / BuildNewHospital win
  bool has_lockdown = false;
  uint32_t lockdown_at_number_of_infections = 0;
  double essential_workers_population = 0;
  bool has_build_new_hospital = false;
  uint32_t spread_rate_threshold = 0;
  std::vector<VaccinateConfig> vaccinate;
  bool enable_citizen_state_messages = false;
  // project extension: Philox key for deterministic mode
  uint64_t seed = 0;
  bool has_seed = false;
};

// Parses `text`; on failure returns false and writes a serde-style message to `err`.
bool parse_config(const std::string& text, Config& out, std::string& err);

// Reads the Population::Csv file named by the config (grid.rs:194-231,
// citizen/population_record.rs:23-43): records with ind (u32), age, working and
// pub_transport ("True" / "False"). Sets number_of_agents to the record count.
bool read_population_csv(Config& cfg, std::string& err);

}  // namespace epi
