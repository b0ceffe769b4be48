// A disease handler for the plug-in tests: the reference's Disease with two of its trait
// methods replaced. on_exposed waits one day longer before symptoms can start, and
// on_routine_end lets every Severe case recover at its last day. The library built with it
// must therefore run exactly like the default library on the same config with
// exposed_duration + 24 and death_rate 0, which the oracle checks.
#pragma once

namespace epi {

struct RecoveringDisease : DefaultDisease {
  EPI_HD static void on_exposed(const DiseaseParams& P, uint32_t hour, const u32x4& d, uint32_t& w, uint32_t& tm) {
    DiseaseParams later = P;
    later.exposed_duration += 24u;
    DefaultDisease::on_exposed(later, hour, d, w, tm);
  }
  EPI_HD static void on_routine_end(const DiseaseParams& P, const u32x4& d, uint32_t& w, uint32_t& tm) {
    DiseaseParams no_death = P;
    no_death.thr_death = 0;  // bernoulli(0, v) is false for every v
    DefaultDisease::on_routine_end(no_death, d, w, tm);
  }
};

}  // namespace epi
