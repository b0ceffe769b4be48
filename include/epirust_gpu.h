/*
 * epirust_gpu.h -- C ABI of the MI355X (gfx950) per-hour agent step for one EpiRust
 * engine region.
 *
 * The library owns all device memory; callers own every host buffer they pass. A handle
 * is driven by one caller thread at a time (calls on one handle are serialized on its
 * own HIP stream); separate handles may be used concurrently. No entry point aborts:
 * conditions on which the reference engine panics return a nonzero status and
 * epi_last_error() returns the reference's panic text.
 *
 * The disease model is fixed when the library is built, as the reference's DiseaseHandler
 * type parameter (engine/src/state_machine/disease_handler.rs:27-44) is fixed when the
 * engine is compiled: libepigpu.so carries the reference's Disease, and a library built
 * with another handler type exports these same entry points.
 *
 * Reference paths below are relative to the reference repository root.
 */
#ifndef EPIRUST_GPU_H
#define EPIRUST_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define EPI_OK 0
#define EPI_ERR_CONFIG 1   /* config JSON rejected (serde would fail)                 */
#define EPI_ERR_PANIC 2    /* a condition on which the reference engine panics        */
#define EPI_ERR_DEVICE 3   /* HIP runtime failure or no gfx950 device                 */
#define EPI_ERR_ARG 4      /* invalid argument (null pointer, index out of range)     */
#define EPI_ERR_IO 5       /* output file could not be written                        */

/* flags for epi_create / epi_batch_create */
#define EPI_FLAG_DETERMINISTIC 1u   /* Philox streams keyed (seed, agent id, hour); the only mode */
#define EPI_FLAG_NO_GRAPH 2u        /* launch hour kernels directly instead of replaying day graphs */
#define EPI_FLAG_HOST_INIT 4u       /* synthesize the population with the host restatement, not on the device */

/* One CSV row of the epicurve; mirrors Counts (engine/src/models/events/counts.rs:25-34). */
typedef struct {
  uint32_t hour, susceptible, exposed, infected, hospitalized, recovered, deceased;
} epi_counts;

/* One entry of the interventions report (engine/src/listeners/intervention_reporter.rs:28-63).
 * kind: 0 vaccination, 1 lockdown {"status": "locked_down"}, 2 lockdown {"status":
 * "lockdown_revoked"}, 3 build_new_hospital. */
typedef struct {
  uint32_t hour, kind;
} epi_event;

/* Per-agent dump for parity checks; caller allocates n_agents entries per array
 * (current_area: 4 per agent, x0 y0 x1 y1). Field meaning follows Citizen
 * (engine/src/citizen/mod.rs:48-63) and State (engine/src/state_machine/state.rs:22-37). */
typedef struct {
  int32_t* x;
  int32_t* y;
  uint8_t* state;         /* 0 Susceptible, 1 Exposed, 2 Infected, 3 Recovered, 4 Deceased */
  uint8_t* severity;      /* Infected only: 0 Pre, 1 Asymptomatic, 2 Mild, 3 Severe */
  uint32_t* infection_day;
  uint32_t* at_hour;      /* Exposed{at_hour} / Pre{at_hour}, else 0 */
  uint8_t* flags;         /* 1 hospitalized, 2 vaccinated, 4 isolated, 8 work_quarantined, 16 uses_public_transport */
  uint8_t* work_status;   /* 0 Normal, 1 Essential, 2 HospitalStaff, 3 NA */
  uint32_t* work_start_at;
  int8_t* immunity;
  int32_t* current_area;
} epi_agents_soa;

typedef struct epi_sim epi_sim;

/* Replaces Epidemiology::new (engine/src/epidemiology_simulation.rs:75-135), called from
 * EngineApp::start_standalone (engine/src/engine_app.rs:89-125) with the config read by
 * Config::read (common/src/config/mod.rs:119-123). config_json is the reference's
 * standalone config schema; `seed` keys the Philox streams. A simulation holds at most
 * 2^27 - 1 agents (agent ids share a 32-bit claim word with the claim hour): larger
 * populations return EPI_ERR_CONFIG. */
int epi_create(const char* config_json, uint64_t seed, uint32_t flags, epi_sim** out);

/* Batched replicates: n independent simulations sharing every launch. Same role as the
 * reference's serial `launch-batch` of engine-app runs (engine/scripts/epirust:58-99). */
int epi_batch_create(const char* const* configs, const uint64_t* seeds, uint32_t n, uint32_t flags, epi_sim** out);

/* Runs up to n_hours of the hour loop (epidemiology_simulation.rs:223-258): the step
 * (CitizenLocationMap::simulate, engine/src/allocation_map.rs:67-129), the counts row,
 * process_interventions (allocation_map.rs:306-337) and the standalone stop rule
 * (epidemiology_simulation.rs:564-575). rows receives one epi_counts per hour run. */
int epi_step(epi_sim* sim, uint32_t n_hours, epi_counts* rows, uint32_t* n_written, int* stopped);

/* Batch form: rows is n x n_hours (row-major by simulation); n_written and stopped have
 * one entry per simulation. */
int epi_batch_step(epi_sim* sim, uint32_t n_hours, epi_counts* rows, uint32_t* n_written, int* stopped);

/* The whole run plus its output files: replaces Epidemiology::run / run_single_engine
 * (epidemiology_simulation.rs:194-274) with the CSV listener (listeners/csv_service.rs:44-71)
 * and interventions report (listeners/intervention_reporter.rs:28-63), written to
 * <output_dir>/output/simulation_0_<UTC time>.csv and ..._interventions.json
 * (engine/src/utils/util.rs:28-43). final_path (optional, cap bytes) gets the CSV path. */
int epi_run(epi_sim* sim, const char* output_dir, char* final_path, size_t cap);

/* Current per-agent state of simulation `index` of the handle (0 for epi_create). */
int epi_read_agents(epi_sim* sim, uint32_t index, epi_agents_soa* dst);

/* One citizen-state record: CitizenState {citizen_id, state, location}
 * (engine/src/models/events/citizen_state.rs:26-31); state is 's', 'e', 'i', 'r' or 'd'
 * (:35-43) and citizen_id the dense deterministic id in place of the UUID. */
typedef struct {
  uint32_t citizen_id;
  int32_t x, y;
  char state;
} epi_citizen_state;

/* The records of simulation `index` after its last stepped hour, in id order: the
 * CitizenStatesAtHr buffer (citizen_state.rs:46-68) filled by citizen_state_updated
 * (engine/src/allocation_map.rs:122-124). *n = agents, *hour = last stepped hour. */
int epi_citizen_states(epi_sim* sim, uint32_t index, epi_citizen_state* out, uint32_t cap, uint32_t* n,
                       uint32_t* hour);

/* Opens a file sink for the citizen-state stream in place of the Kafka topic
 * citizen_states_updated (engine/src/listeners/events_kafka_producer.rs:62-121): one JSON
 * line per message, the Grid first, then {"hr":h,"citizen_states":[...]} for every hour
 * epi_step runs (formatted on the device). path NULL closes the sink after the
 * {"simulation_ended": true} message. Single-simulation handles only. epi_run opens and
 * closes it itself when the config sets enable_citizen_state_messages. */
int epi_stream_citizen_states(epi_sim* sim, const char* path);

/* The citizen-state message body of the handle's simulation after its last stepped hour,
 * formatted on the device into caller-owned device memory: d_text receives the records
 * joined by commas, exactly the text between '[' and ']' of the {"hr":h,"citizen_states":
 * [...]} line that epi_stream_citizen_states writes for that hour, and d_len (8 device
 * bytes) its length. No counterpart in the reference: a variant of the stream's device
 * formatter for callers that cannot allocate (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). It calls no allocating routine: the formatter's
 * per-block offsets live in one static device array per GPU (512 KiB, held until the
 * process ends), so it handles at most 16,777,216 agents, and calls on one GPU that
 * overlap race on that array. cap must hold the records' upper bound, agents x the
 * longest record (EPI_ERR_ARG otherwise, before anything is launched). Single-simulation
 * handles only; returns once the text is written. */
int epi_citizen_states_text_noalloc(epi_sim* sim, char* d_text, uint64_t cap, uint64_t* d_len);

/* Counts of the last completed hour (hour 0: counts_at_start, engine/src/utils/util.rs:45-51). */
int epi_current_counts(epi_sim* sim, uint32_t index, epi_counts* out);

/* Intervention events recorded so far for simulation `index`. */
int epi_events(epi_sim* sim, uint32_t index, epi_event* out, uint32_t cap, uint32_t* n);

/* Sizes of a handle: simulations, and agents of simulation `index`. */
int epi_info(epi_sim* sim, uint32_t index, uint32_t* n_sims, uint32_t* n_agents, uint32_t* grid_size, uint32_t* hours);

/* Geometry of simulation `index`: housing, transport, work, initial hospital and the
 * hospital after BuildNewHospital (5 x {x0,y0,x1,y1}) and the current hospital version. */
int epi_geometry(epi_sim* sim, uint32_t index, int32_t rects[20], uint32_t* hospital_version);

/* Profiling: runs n_hours with a HIP event pair around every kernel on the handle's stream
 * and returns the mean duration in ms of each of an hour's kernels (hospital scan; step;
 * the commit launch with its folded epilogue), each mean over the hours that launched it
 * (scan: hour of day 0; step: hours of day 0 and 7..23; commit: every hour), the number of
 * launches of each, and how many hours ran. */
int epi_profile(epi_sim* sim, uint32_t n_hours, double ms_mean[3], uint32_t launches[3], uint32_t* hours_run);

/* Device time of the hour loop of the last epi_step / epi_batch_step / epi_run call: HIP
 * events recorded on the handle's stream before its first hour's launches and after its
 * last hour's, and the number of hours enqueued between them (0 and 0 when no hour ran or
 * the citizen-state stream was open). Measurement only; no reference counterpart. */
int epi_last_run_ms(epi_sim* sim, double* device_ms, uint32_t* hours);

/* Launch shape (no reference counterpart): threads per agent block (64..1024, multiple of
 * 64), for the home and the office layout alike. Takes effect on the next step; discards
 * the captured day graph. Default: 512 for a single simulation of at least 2^19 agents,
 * else 256. */
int epi_set_block_size(epi_sim* sim, uint32_t threads);

/* Launch shape (no reference counterpart): threads per agent block of the office layout
 * alone (hours of day 8..15, whose worker blocks pack whole offices into that many slots),
 * 64..1024, multiple of 64. epi_set_block_size resets it to the home value. */
int epi_set_office_block_size(epi_sim* sim, uint32_t threads);

/* Message of the last failing call on this thread ("" if none). */
const char* epi_last_error(void);

void epi_destroy(epi_sim* sim);

/* ---- Learned surrogate (no reference counterpart: the reference only collates ensemble
 * mean/std, engine/plot/models/EpiCurves.py:30-40). An MLP widths[0] -> ... -> widths[n-1]
 * with GELU after every hidden layer and an identity output, bf16 weights/activations on
 * MFMA with FP32 accumulation, MSE loss, Adam. Parameters are laid out W0 (K x N row-major),
 * b0, W1, b1, ... as fp32. */
typedef struct epi_mlp epi_mlp;
int epi_mlp_create(const uint32_t* widths, uint32_t n_widths, uint64_t seed, uint32_t max_batch, epi_mlp** out);
uint64_t epi_mlp_param_count(epi_mlp* mlp);
int epi_mlp_get_params(epi_mlp* mlp, float* params, uint64_t n);
int epi_mlp_set_params(epi_mlp* mlp, const float* params, uint64_t n);
/* y (batch x widths[n-1]) = network(x (batch x widths[0])). */
int epi_mlp_forward(epi_mlp* mlp, const float* x, uint32_t batch, float* y);
/* One Adam step on the MSE of network(x) against t; loss (optional) gets the batch MSE. */
int epi_mlp_train_step(epi_mlp* mlp, const float* x, const float* t, uint32_t batch, float lr, float* loss);
/* Parameter gradients of the batch MSE (same layout as the parameters), for checks. */
int epi_mlp_gradients(epi_mlp* mlp, const float* x, const float* t, uint32_t batch, float* grads, float* loss);
const char* epi_mlp_last_error(void);
void epi_mlp_destroy(epi_mlp* mlp);
/* C (M x N, fp32) = A * B with bf16 inputs on the MFMA kernel; A is M x K row-major, or
 * K x M when trans_a; B is K x N, or N x K when trans_b. Host buffers. */
int epi_gemm_bf16(const uint16_t* A, const uint16_t* B, float* C, int M, int N, int K, int trans_a, int trans_b);

/* Command-line entry equivalent to engine-app (engine-app/src/main.rs:60-87, 167-173):
 * -c/--config, -m/--mode (standalone only), -i/--id, -t/--threads (accepted, unused),
 * -o/--output-dir (default /tmp). Returns a process exit code. */
int epi_cli_main(int argc, char** argv);

#ifdef __cplusplus
}
#endif
#endif
