//! extern "C" binding of libepigpu.so, the MI355X (gfx950) per-hour agent step of EpiRust,
//! for the reference's Rust workspace. Declarations mirror include/epirust_gpu.h one for
//! one; `Simulation` is a thin owning wrapper; `start_standalone` is what
//! EngineApp::start_standalone (engine/src/engine_app.rs:89-125) becomes when it drives
//! the GPU step instead of Epidemiology::new(...).run(threads).
//!
//! Error contract: no entry point aborts. The conditions on which the reference engine
//! panics return EPI_ERR_PANIC with the reference's message from `epi_last_error`, so a
//! caller that wants the reference's behaviour panics with that text.

use std::ffi::{CStr, CString};
use std::os::raw::{c_char, c_int};

pub const EPI_OK: c_int = 0;
pub const EPI_ERR_CONFIG: c_int = 1;
pub const EPI_ERR_PANIC: c_int = 2;
pub const EPI_ERR_DEVICE: c_int = 3;
pub const EPI_ERR_ARG: c_int = 4;
pub const EPI_ERR_IO: c_int = 5;

pub const EPI_FLAG_DETERMINISTIC: u32 = 1;
pub const EPI_FLAG_NO_GRAPH: u32 = 2;
pub const EPI_FLAG_HOST_INIT: u32 = 4;

/// One CSV row, in the serde field order of Counts (engine/src/models/events/counts.rs:25-34).
#[repr(C)]
#[derive(Clone, Copy, Default, Debug, PartialEq, Eq)]
pub struct EpiCounts {
    pub hour: u32,
    pub susceptible: u32,
    pub exposed: u32,
    pub infected: u32,
    pub hospitalized: u32,
    pub recovered: u32,
    pub deceased: u32,
}

/// kind: 0 vaccination, 1 lockdown locked_down, 2 lockdown lockdown_revoked, 3 build_new_hospital.
#[repr(C)]
#[derive(Clone, Copy, Default, Debug, PartialEq, Eq)]
pub struct EpiEvent {
    pub hour: u32,
    pub kind: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct EpiCitizenState {
    pub citizen_id: u32,
    pub x: i32,
    pub y: i32,
    pub state: u8,
}

#[repr(C)]
pub struct EpiAgentsSoa {
    pub x: *mut i32,
    pub y: *mut i32,
    pub state: *mut u8,
    pub severity: *mut u8,
    pub infection_day: *mut u32,
    pub at_hour: *mut u32,
    pub flags: *mut u8,
    pub work_status: *mut u8,
    pub work_start_at: *mut u32,
    pub immunity: *mut i8,
    pub current_area: *mut i32,
}

#[repr(C)]
pub struct EpiSim {
    _private: [u8; 0],
}

#[repr(C)]
pub struct EpiMlp {
    _private: [u8; 0],
}

extern "C" {
    pub fn epi_create(config_json: *const c_char, seed: u64, flags: u32, out: *mut *mut EpiSim) -> c_int;
    pub fn epi_batch_create(configs: *const *const c_char, seeds: *const u64, n: u32, flags: u32,
                            out: *mut *mut EpiSim) -> c_int;
    pub fn epi_step(sim: *mut EpiSim, n_hours: u32, rows: *mut EpiCounts, n_written: *mut u32,
                    stopped: *mut c_int) -> c_int;
    pub fn epi_batch_step(sim: *mut EpiSim, n_hours: u32, rows: *mut EpiCounts, n_written: *mut u32,
                          stopped: *mut c_int) -> c_int;
    pub fn epi_run(sim: *mut EpiSim, output_dir: *const c_char, final_path: *mut c_char, cap: usize) -> c_int;
    pub fn epi_read_agents(sim: *mut EpiSim, index: u32, dst: *mut EpiAgentsSoa) -> c_int;
    pub fn epi_citizen_states(sim: *mut EpiSim, index: u32, out: *mut EpiCitizenState, cap: u32, n: *mut u32,
                              hour: *mut u32) -> c_int;
    pub fn epi_stream_citizen_states(sim: *mut EpiSim, path: *const c_char) -> c_int;
    pub fn epi_citizen_states_text_noalloc(sim: *mut EpiSim, d_text: *mut c_char, cap: u64, d_len: *mut u64) -> c_int;
    pub fn epi_current_counts(sim: *mut EpiSim, index: u32, out: *mut EpiCounts) -> c_int;
    pub fn epi_events(sim: *mut EpiSim, index: u32, out: *mut EpiEvent, cap: u32, n: *mut u32) -> c_int;
    pub fn epi_info(sim: *mut EpiSim, index: u32, n_sims: *mut u32, n_agents: *mut u32, grid_size: *mut u32,
                    hours: *mut u32) -> c_int;
    pub fn epi_geometry(sim: *mut EpiSim, index: u32, rects: *mut i32, hospital_version: *mut u32) -> c_int;
    pub fn epi_profile(sim: *mut EpiSim, n_hours: u32, ms_mean: *mut f64, launches: *mut u32,
                       hours_run: *mut u32) -> c_int;
    pub fn epi_last_run_ms(sim: *mut EpiSim, device_ms: *mut f64, hours: *mut u32) -> c_int;
    pub fn epi_set_block_size(sim: *mut EpiSim, threads: u32) -> c_int;
    pub fn epi_set_office_block_size(sim: *mut EpiSim, threads: u32) -> c_int;
    pub fn epi_last_error() -> *const c_char;
    pub fn epi_destroy(sim: *mut EpiSim);

    pub fn epi_mlp_create(widths: *const u32, n_widths: u32, seed: u64, max_batch: u32, out: *mut *mut EpiMlp) -> c_int;
    pub fn epi_mlp_param_count(mlp: *mut EpiMlp) -> u64;
    pub fn epi_mlp_get_params(mlp: *mut EpiMlp, params: *mut f32, n: u64) -> c_int;
    pub fn epi_mlp_set_params(mlp: *mut EpiMlp, params: *const f32, n: u64) -> c_int;
    pub fn epi_mlp_forward(mlp: *mut EpiMlp, x: *const f32, batch: u32, y: *mut f32) -> c_int;
    pub fn epi_mlp_train_step(mlp: *mut EpiMlp, x: *const f32, t: *const f32, batch: u32, lr: f32,
                              loss: *mut f32) -> c_int;
    pub fn epi_mlp_gradients(mlp: *mut EpiMlp, x: *const f32, t: *const f32, batch: u32, grads: *mut f32,
                             loss: *mut f32) -> c_int;
    pub fn epi_mlp_last_error() -> *const c_char;
    pub fn epi_mlp_destroy(mlp: *mut EpiMlp);
    pub fn epi_gemm_bf16(a: *const u16, b: *const u16, c: *mut f32, m: c_int, n: c_int, k: c_int, trans_a: c_int,
                         trans_b: c_int) -> c_int;

    pub fn epi_cli_main(argc: c_int, argv: *mut *mut c_char) -> c_int;
}

/// The library's message for the last failing call on this thread.
pub fn last_error() -> String {
    unsafe { CStr::from_ptr(epi_last_error()).to_string_lossy().into_owned() }
}

#[derive(Debug)]
pub struct EpiError {
    pub code: c_int,
    pub message: String,
}

fn check(code: c_int) -> Result<(), EpiError> {
    if code == EPI_OK {
        Ok(())
    } else {
        Err(EpiError { code, message: last_error() })
    }
}

/// One engine region on the GPU (Epidemiology, engine/src/epidemiology_simulation.rs:59-135).
pub struct Simulation {
    raw: *mut EpiSim,
}

impl Simulation {
    /// config_json: the reference's standalone config; `"seed"` in it, or `seed`, keys the
    /// deterministic mode.
    pub fn new(config_json: &str, seed: u64) -> Result<Simulation, EpiError> {
        let text = CString::new(config_json).expect("config without NUL bytes");
        let mut raw = std::ptr::null_mut();
        check(unsafe { epi_create(text.as_ptr(), seed, EPI_FLAG_DETERMINISTIC, &mut raw) })?;
        Ok(Simulation { raw })
    }

    /// Up to n_hours of the hour loop (epidemiology_simulation.rs:223-258); the rows of the
    /// hours run and whether the run has stopped (epidemiology_simulation.rs:564-575).
    pub fn step(&mut self, n_hours: u32) -> Result<(Vec<EpiCounts>, bool), EpiError> {
        let mut rows = vec![EpiCounts::default(); n_hours.max(1) as usize];
        let (mut n, mut stopped) = (0u32, 0 as c_int);
        check(unsafe { epi_step(self.raw, n_hours, rows.as_mut_ptr(), &mut n, &mut stopped) })?;
        rows.truncate(n as usize);
        Ok((rows, stopped != 0))
    }

    /// Intervention events so far, in the order InterventionReporter receives them.
    pub fn events(&mut self) -> Result<Vec<EpiEvent>, EpiError> {
        let mut n = 0u32;
        check(unsafe { epi_events(self.raw, 0, std::ptr::null_mut(), 0, &mut n) })?;
        let mut out = vec![EpiEvent::default(); n as usize];
        check(unsafe { epi_events(self.raw, 0, out.as_mut_ptr(), n, &mut n) })?;
        Ok(out)
    }

    /// The whole run with the reference's CSV and interventions JSON under
    /// <output_dir>/output/; returns the CSV path.
    pub fn run(&mut self, output_dir: &str) -> Result<String, EpiError> {
        let dir = CString::new(output_dir).expect("path without NUL bytes");
        let mut path = vec![0 as c_char; 4096];
        check(unsafe { epi_run(self.raw, dir.as_ptr(), path.as_mut_ptr(), path.len()) })?;
        Ok(unsafe { CStr::from_ptr(path.as_ptr()) }.to_string_lossy().into_owned())
    }

    pub fn as_raw(&self) -> *mut EpiSim {
        self.raw
    }
}

impl Drop for Simulation {
    fn drop(&mut self) {
        unsafe { epi_destroy(self.raw) }
    }
}

/// EngineApp::start_standalone (engine/src/engine_app.rs:89-125) on the GPU step: the
/// reference panics where the library reports EPI_ERR_PANIC, with the same text.
pub fn start_standalone(config_json: &str, seed: u64, output_dir: &str) -> String {
    let mut sim = Simulation::new(config_json, seed).unwrap_or_else(|e| panic!("{}", e.message));
    sim.run(output_dir).unwrap_or_else(|e| panic!("{}", e.message))
}
