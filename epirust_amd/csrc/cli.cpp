// `epirust-gpu`: the engine-app command line (engine-app/src/main.rs:60-87, 99-177) for
// the standalone mode, driving the C ABI.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "config.hpp"
#include "epirust_gpu.h"

namespace {

void usage() {
  std::fprintf(stderr,
               "usage: epirust-gpu [-c|--config FILE] [-m|--mode standalone] [-i|--id ID] [-t|--threads N] "
               "[-o|--output-dir DIR]\n");
}

bool take(int argc, char** argv, int& i, const char* s, const char* l, std::string& out) {
  const std::string a = argv[i];
  if (a == s || a == l) {
    if (i + 1 >= argc) return false;
    out = argv[++i];
    return true;
  }
  const std::string pre = std::string(l) + "=";
  if (a.compare(0, pre.size(), pre) == 0) {
    out = a.substr(pre.size());
    return true;
  }
  return false;
}

}  // namespace

extern "C" int epi_cli_main(int argc, char** argv) {
  std::string config = "config/default.json", mode = "standalone", id = "default_engine", threads = "4",
              out_dir = "/tmp";
  for (int i = 1; i < argc; ++i) {
    if (take(argc, argv, i, "-c", "--config", config) || take(argc, argv, i, "-m", "--mode", mode) ||
        take(argc, argv, i, "-i", "--id", id) || take(argc, argv, i, "-t", "--threads", threads) ||
        take(argc, argv, i, "-o", "--output-dir", out_dir))
      continue;
    if (!std::strcmp(argv[i], "-h") || !std::strcmp(argv[i], "--help")) {
      usage();
      return 0;
    }
    std::fprintf(stderr, "error: unexpected argument '%s'\n", argv[i]);
    usage();
    return 2;
  }
  // clap's ValueEnum names of Mode (main.rs:41-46), matched case-sensitively
  if (mode == "kafka" || mode == "mpi") {
    std::fprintf(stderr, "error: mode '%s' is not provided by this engine (standalone only)\n", mode.c_str());
    return 2;
  }
  if (mode != "standalone") {
    std::fprintf(stderr,
                 "error: invalid value '%s' for '--mode <MODE>'\n  [possible values: kafka, mpi, standalone]\n",
                 mode.c_str());
    return 2;
  }
  std::printf("Standalone\n");
  std::ifstream f(config);
  if (!f) {
    std::fprintf(stderr, "Failed to read config file: %s\n", config.c_str());
    return 1;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string text = ss.str();
  epi::Config parsed;
  std::string err;
  if (!epi::parse_config(text, parsed, err)) {
    std::fprintf(stderr, "Failed to read config file: %s\n", err.c_str());
    return 1;
  }
  const auto t0 = std::chrono::steady_clock::now();
  epi_sim* sim = nullptr;
  if (epi_create(text.c_str(), parsed.has_seed ? parsed.seed : 0, EPI_FLAG_DETERMINISTIC, &sim) != EPI_OK) {
    std::fprintf(stderr, "error: %s\n", epi_last_error());
    return 1;
  }
  const auto t1 = std::chrono::steady_clock::now();
  std::printf("Initialization completed in %.3f seconds\n", std::chrono::duration<double>(t1 - t0).count());
  char path[4096] = {0};
  const int rc = epi_run(sim, out_dir.c_str(), path, sizeof path);
  const auto t2 = std::chrono::steady_clock::now();
  if (rc != EPI_OK) {
    std::fprintf(stderr, "error: %s\n", epi_last_error());
    epi_destroy(sim);
    return 1;
  }
  epi_counts last{};
  epi_current_counts(sim, 0, &last);
  const double secs = std::chrono::duration<double>(t2 - t1).count();
  std::printf("Number of iterations: %u, Total Time taken %.6f seconds\n", last.hour, secs);
  std::printf("Iterations/sec: %.3f\n", secs > 0 ? last.hour / secs : 0.0);
  std::printf("Output: %s\n", path);
  epi_destroy(sim);
  return 0;
}
