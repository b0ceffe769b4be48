"""MI355X (gfx950) per-hour agent step of EpiRust, behind the engine's config JSON, run
entry point and CSV/JSON outputs.

The compute path is libepigpu.so (HIP kernels + C ABI, include/epirust_gpu.h). This module
is a thin ctypes mirror of the reference's standalone entry points:

    EngineApp.start_standalone(config, output_dir)   engine/src/engine_app.rs:89-125
    Epidemiology(config).run()                       engine/src/epidemiology_simulation.rs:194-274

There is no CPU fallback: if the library is missing or no gfx950 device is present the
calls raise.

The disease model is compiled into the library (the reference's DiseaseHandler type
parameter, engine/src/state_machine/disease_handler.rs:27-44): EPIRUST_GPU_LIB names a
library built with another handler (python -m epirust_amd.build --disease-handler
HEADER:TYPE --out PATH) to load instead of libepigpu.so.
"""
import ctypes
import json
import os

import numpy as np

__all__ = ["Simulation", "Batch", "EngineApp", "EpiError", "lib", "LIB_PATH"]

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("EPIRUST_GPU_LIB") or os.path.join(PKG, "libepigpu.so")
EXE_PATH = os.path.join(PKG, "epirust-gpu")

EPI_FLAG_DETERMINISTIC = 1
EPI_FLAG_NO_GRAPH = 2
EPI_FLAG_HOST_INIT = 4
CITIZEN_STATE_DTYPE = np.dtype({"names": ["citizen_id", "x", "y", "state"],
                                 "formats": [np.uint32, np.int32, np.int32, "S1"],
                                 "offsets": [0, 4, 8, 12], "itemsize": 16})
EVENT_NAMES = {0: ("vaccination", {}), 1: ("lockdown", {"status": "locked_down"}),
               2: ("lockdown", {"status": "lockdown_revoked"}), 3: ("build_new_hospital", {})}

AGENT_FIELDS = [
    ("x", np.int32, 1), ("y", np.int32, 1), ("state", np.uint8, 1), ("severity", np.uint8, 1),
    ("infection_day", np.uint32, 1), ("at_hour", np.uint32, 1), ("flags", np.uint8, 1),
    ("work_status", np.uint8, 1), ("work_start_at", np.uint32, 1), ("immunity", np.int8, 1),
    ("current_area", np.int32, 4),
]

ROW_DTYPE = np.dtype([(k, np.uint32) for k in
                      ("hour", "susceptible", "exposed", "infected", "hospitalized", "recovered", "deceased")])


class EpiError(RuntimeError):
    def __init__(self, code, message):
        super().__init__("epirust_gpu error %d: %s" % (code, message))
        self.code = code
        self.message = message


class _Counts(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint32) for k in ROW_DTYPE.names]


class _Event(ctypes.Structure):
    _fields_ = [("hour", ctypes.c_uint32), ("kind", ctypes.c_uint32)]


class _Agents(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name, _, _ in AGENT_FIELDS]


_P = ctypes.POINTER
_lib = None

SYMBOLS = {
    "epi_create": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint32, _P(ctypes.c_void_p)]),
    "epi_batch_create": (ctypes.c_int, [_P(ctypes.c_char_p), _P(ctypes.c_uint64), ctypes.c_uint32, ctypes.c_uint32,
                                        _P(ctypes.c_void_p)]),
    "epi_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, _P(ctypes.c_uint32),
                                _P(ctypes.c_int)]),
    "epi_batch_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p]),
    "epi_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t]),
    "epi_read_agents": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, _P(_Agents)]),
    "epi_current_counts": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, _P(_Counts)]),
    "epi_events": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32,
                                  _P(ctypes.c_uint32)]),
    "epi_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32] + [_P(ctypes.c_uint32)] * 4),
    "epi_geometry": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, _P(ctypes.c_int32), _P(ctypes.c_uint32)]),
    "epi_profile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, _P(ctypes.c_double), _P(ctypes.c_uint32),
                                   _P(ctypes.c_uint32)]),
    "epi_set_block_size": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "epi_set_office_block_size": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "epi_last_run_ms": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_double), _P(ctypes.c_uint32)]),
    "epi_citizen_states": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32,
                                          _P(ctypes.c_uint32), _P(ctypes.c_uint32)]),
    "epi_stream_citizen_states": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "epi_citizen_states_text_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                                       ctypes.c_void_p]),
    "epi_last_error": (ctypes.c_char_p, []),
    "epi_destroy": (None, [ctypes.c_void_p]),
    "epi_cli_main": (ctypes.c_int, [ctypes.c_int, _P(ctypes.c_char_p)]),
    "epi_mlp_create": (ctypes.c_int, [_P(ctypes.c_uint32), ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32,
                                      _P(ctypes.c_void_p)]),
    "epi_mlp_param_count": (ctypes.c_uint64, [ctypes.c_void_p]),
    "epi_mlp_get_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]),
    "epi_mlp_set_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]),
    "epi_mlp_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]),
    "epi_mlp_train_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                          ctypes.c_float, _P(ctypes.c_float)]),
    "epi_mlp_gradients": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                         ctypes.c_void_p, _P(ctypes.c_float)]),
    "epi_mlp_last_error": (ctypes.c_char_p, []),
    "epi_mlp_destroy": (None, [ctypes.c_void_p]),
    "epi_gemm_bf16": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int]),
}


def lib():
    """Load libepigpu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EpiError(-1, "%s is not built (python -m epirust_amd.build)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SYMBOLS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise EpiError(rc, lib().epi_last_error().decode(errors="replace"))


def _config_text(config):
    if isinstance(config, dict):
        return json.dumps(config)
    if isinstance(config, bytes):
        return config.decode()
    if isinstance(config, str) and os.path.exists(config) and not config.lstrip().startswith("{"):
        with open(config) as f:
            return f.read()
    return config


def _seed_of(text, seed):
    if seed is not None:
        return int(seed) & 0xFFFFFFFFFFFFFFFF
    try:
        return int(json.loads(text).get("seed", 0)) & 0xFFFFFFFFFFFFFFFF
    except (ValueError, AttributeError):
        return 0


class _Handle:
    def __init__(self):
        self._h = ctypes.c_void_p()

    def close(self):
        if self._h:
            lib().epi_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self, index=0):
        vals = [ctypes.c_uint32() for _ in range(4)]
        _check(lib().epi_info(self._h, index, *[ctypes.byref(v) for v in vals]))
        return dict(zip(("n_sims", "n_agents", "grid_size", "hours"), [v.value for v in vals]))

    def agents(self, index=0):
        """Per-agent state (deterministic-mode parity dump) as a dict of numpy arrays."""
        n = self.info(index)["n_agents"]
        arrays = {name: np.zeros(max(n, 1) * w, dtype=dt) for name, dt, w in AGENT_FIELDS}
        s = _Agents(**{name: a.ctypes.data for name, a in arrays.items()})
        _check(lib().epi_read_agents(self._h, index, ctypes.byref(s)))
        out = {k: v[: n * dict((f, w) for f, _, w in AGENT_FIELDS)[k]] for k, v in arrays.items()}
        out["current_area"] = out["current_area"].reshape(-1, 4)
        return out

    def counts(self, index=0):
        c = _Counts()
        _check(lib().epi_current_counts(self._h, index, ctypes.byref(c)))
        return [getattr(c, k) for k in ROW_DTYPE.names]

    def events(self, index=0):
        n = ctypes.c_uint32()
        _check(lib().epi_events(self._h, index, None, 0, ctypes.byref(n)))
        buf = (_Event * max(n.value, 1))()
        _check(lib().epi_events(self._h, index, buf, n.value, ctypes.byref(n)))
        return [(buf[i].hour, buf[i].kind) for i in range(n.value)]

    def geometry(self, index=0):
        r = (ctypes.c_int32 * 20)()
        ver = ctypes.c_uint32()
        _check(lib().epi_geometry(self._h, index, r, ctypes.byref(ver)))
        return [tuple(r[4 * k:4 * k + 4]) for k in range(5)], ver.value

    def profile(self, n_hours):
        """Mean kernel durations (ms) measured with HIP events on the handle's stream."""
        ms = (ctypes.c_double * 3)()
        n = (ctypes.c_uint32 * 3)()
        hours = ctypes.c_uint32()
        _check(lib().epi_profile(self._h, n_hours, ms, n, ctypes.byref(hours)))
        return {"hospital_scan": (ms[0], n[0]), "step": (ms[1], n[1]), "commit": (ms[2], n[2]),
                "hours": hours.value}

    def set_block_size(self, threads):
        """Threads per agent block, home and office layout alike."""
        _check(lib().epi_set_block_size(self._h, threads))

    def set_office_block_size(self, threads):
        """Threads per agent block of the office layout alone (hours of day 8..15)."""
        _check(lib().epi_set_office_block_size(self._h, threads))

    def last_run_ms(self):
        """(device ms, hours) of the last step call's hour loop, from HIP events on the
        handle's stream."""
        ms, hours = ctypes.c_double(), ctypes.c_uint32()
        _check(lib().epi_last_run_ms(self._h, ctypes.byref(ms), ctypes.byref(hours)))
        return ms.value, hours.value

    def citizen_states(self, index=0):
        """(hour, records) after the last stepped hour: a structured array of
        (citizen_id, x, y, state) in id order (CitizenStatesAtHr, citizen_state.rs:46-68)."""
        n, hour = ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().epi_citizen_states(self._h, index, None, 0, ctypes.byref(n), ctypes.byref(hour)))
        out = np.zeros(max(n.value, 1), dtype=CITIZEN_STATE_DTYPE)
        _check(lib().epi_citizen_states(self._h, index, out.ctypes.data, n.value, ctypes.byref(n),
                                        ctypes.byref(hour)))
        return hour.value, out[: n.value]

    def stream_citizen_states(self, path):
        """Open (path) or close (None) the JSON-lines citizen-state sink."""
        _check(lib().epi_stream_citizen_states(self._h, None if path is None else os.fsencode(path)))


class Simulation(_Handle):
    """One engine region in deterministic mode (Epidemiology, epidemiology_simulation.rs:59-135)."""

    def __init__(self, config, seed=None, flags=EPI_FLAG_DETERMINISTIC):
        super().__init__()
        text = _config_text(config)
        _check(lib().epi_create(text.encode(), _seed_of(text, seed), flags, ctypes.byref(self._h)))

    def step(self, n_hours):
        """Run up to n_hours; returns (rows as a (k, 7) uint32 array, stopped)."""
        rows = np.zeros((max(n_hours, 1), 7), dtype=np.uint32)
        nw = ctypes.c_uint32()
        st = ctypes.c_int()
        _check(lib().epi_step(self._h, n_hours, rows.ctypes.data, ctypes.byref(nw), ctypes.byref(st)))
        return rows[: nw.value], bool(st.value)

    def run(self, output_dir):
        """Whole run + the reference's CSV and interventions JSON; returns the CSV path."""
        buf = ctypes.create_string_buffer(4096)
        _check(lib().epi_run(self._h, os.fsencode(output_dir), buf, 4096))
        return buf.value.decode()


class Batch(_Handle):
    """Many independent simulations sharing every launch (replicate ensembles)."""

    def __init__(self, configs, seeds=None, flags=EPI_FLAG_DETERMINISTIC):
        super().__init__()
        texts = [_config_text(c) for c in configs]
        seeds = [_seed_of(t, None if seeds is None else seeds[i]) for i, t in enumerate(texts)]
        n = len(texts)
        arr = (ctypes.c_char_p * n)(*[t.encode() for t in texts])
        sd = (ctypes.c_uint64 * n)(*seeds)
        _check(lib().epi_batch_create(arr, sd, n, flags, ctypes.byref(self._h)))
        self.n = n

    def step(self, n_hours):
        """Returns (list of (k_i, 7) row arrays, list of stopped flags)."""
        rows = np.zeros((self.n, max(n_hours, 1), 7), dtype=np.uint32)
        nw = np.zeros(self.n, dtype=np.uint32)
        st = np.zeros(self.n, dtype=np.int32)
        _check(lib().epi_batch_step(self._h, n_hours, rows.ctypes.data, nw.ctypes.data, st.ctypes.data))
        return [rows[i, : nw[i]] for i in range(self.n)], [bool(x) for x in st]


class EngineApp:
    """Mirror of engine/src/engine_app.rs:32-126 for the standalone mode."""

    STANDALONE_SIM_ID = "0"

    @staticmethod
    def start_standalone(config, output_dir, seed=None):
        with Simulation(config, seed=seed) as sim:
            return sim.run(output_dir)


def cli(argv):
    """engine-app compatible command line (epi_cli_main)."""
    args = [b"epirust-gpu"] + [os.fsencode(a) for a in argv]
    arr = (ctypes.c_char_p * len(args))(*args)
    return lib().epi_cli_main(len(args), arr)
