"""ctypes binding of the CPU oracle (oracle/epicpu.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker. The product package
(epirust_amd) never imports this module.

`config_from_json` restates the reference's serde schema for the standalone
Config (common/src/config/mod.rs:44-58): unknown keys are ignored, `disease_overrides`
and `output_file` are accepted and unused, `starting_infections` defaults to
`exposed: 1` (common/src/config/starting_infections.rs:65-69), and the project's
extra key `seed` selects the Philox key.
"""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libepicpu.so")
REF_LIB_PATH = os.path.join(HERE, "librefcpu.so")
MAX_VACC = 4096


class EpicpuConfig(ctypes.Structure):
    _fields_ = [
        ("number_of_agents", ctypes.c_uint32),
        ("public_transport_percentage", ctypes.c_double),
        ("working_percentage", ctypes.c_double),
        ("regular_transmission_start_day", ctypes.c_uint32),
        ("high_transmission_start_day", ctypes.c_uint32),
        ("last_day", ctypes.c_uint32),
        ("asymptomatic_last_day", ctypes.c_uint32),
        ("mild_infected_last_day", ctypes.c_uint32),
        ("regular_transmission_rate", ctypes.c_double),
        ("high_transmission_rate", ctypes.c_double),
        ("death_rate", ctypes.c_double),
        ("percentage_asymptomatic_population", ctypes.c_double),
        ("percentage_severe_infected_population", ctypes.c_double),
        ("exposed_duration", ctypes.c_uint32),
        ("pre_symptomatic_duration", ctypes.c_uint32),
        ("grid_size", ctypes.c_uint32),
        ("hospital_beds_percentage", ctypes.c_double),
        ("hours", ctypes.c_uint32),
        ("si_infected_mild_asymptomatic", ctypes.c_uint32),
        ("si_infected_mild_symptomatic", ctypes.c_uint32),
        ("si_infected_severe", ctypes.c_uint32),
        ("si_exposed", ctypes.c_uint32),
        ("has_lockdown", ctypes.c_uint32),
        ("lockdown_at_number_of_infections", ctypes.c_uint32),
        ("lockdown_essential_workers_population", ctypes.c_double),
        ("has_build_new_hospital", ctypes.c_uint32),
        ("spread_rate_threshold", ctypes.c_uint32),
        ("n_vaccinate", ctypes.c_uint32),
        ("vaccinate_at_hour", ctypes.c_uint32 * MAX_VACC),
        ("vaccinate_percent", ctypes.c_double * MAX_VACC),
        ("seed", ctypes.c_uint64),
        ("csv_n", ctypes.c_uint32),
        ("csv_working", ctypes.c_void_p),
        ("csv_pub_transport", ctypes.c_void_p),
    ]


_P = ctypes.POINTER
AGENT_FIELDS = [
    ("x", np.int32, 1),
    ("y", np.int32, 1),
    ("state", np.uint8, 1),
    ("severity", np.uint8, 1),
    ("infection_day", np.uint32, 1),
    ("at_hour", np.uint32, 1),
    ("flags", np.uint8, 1),
    ("work_status", np.uint8, 1),
    ("work_start_at", np.uint32, 1),
    ("immunity", np.int8, 1),
    ("current_area", np.int32, 4),
]


class EpicpuAgents(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name, _, _ in AGENT_FIELDS]


class EpicpuAgentInit(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int32), ("y", ctypes.c_int32), ("home", ctypes.c_int32 * 4),
                ("work", ctypes.c_int32 * 4), ("current_area", ctypes.c_int32 * 4), ("state", ctypes.c_uint32),
                ("severity", ctypes.c_uint32), ("work_status", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("infection_day", ctypes.c_uint32), ("at_hour", ctypes.c_uint32), ("work_start_at", ctypes.c_uint32),
                ("immunity", ctypes.c_int32)]


_libs = {}


def lib(path=LIB_PATH):
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError("oracle library missing: build it with `make -C oracle`")
        L = ctypes.CDLL(path)
        L.epicpu_create.restype = ctypes.c_void_p
        L.epicpu_create.argtypes = [_P(EpicpuConfig), ctypes.c_char_p, ctypes.c_size_t]
        L.epicpu_destroy.argtypes = [ctypes.c_void_p]
        L.epicpu_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        L.epicpu_set_reference_order.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.epicpu_set_random_merge.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.epicpu_step.argtypes = [ctypes.c_void_p, ctypes.c_uint32, _P(ctypes.c_uint32), _P(ctypes.c_uint32),
                                  _P(ctypes.c_int)]
        L.epicpu_read_agents.argtypes = [ctypes.c_void_p, _P(EpicpuAgents)]
        L.epicpu_counts.argtypes = [ctypes.c_void_p, _P(ctypes.c_uint32)]
        L.epicpu_events.restype = ctypes.c_uint32
        L.epicpu_events.argtypes = [ctypes.c_void_p, _P(ctypes.c_uint32), _P(ctypes.c_uint32), ctypes.c_uint32]
        L.epicpu_geometry.argtypes = [ctypes.c_void_p, _P(ctypes.c_int32), _P(ctypes.c_uint32)]
        L.epicpu_philox4x32_10.argtypes = [_P(ctypes.c_uint32), _P(ctypes.c_uint32), _P(ctypes.c_uint32)]
        L.epicpu_define_geography.argtypes = [ctypes.c_uint32, _P(ctypes.c_int32)]
        L.epicpu_area_factory.restype = ctypes.c_uint32
        L.epicpu_area_factory.argtypes = [ctypes.c_int32] * 4 + [ctypes.c_uint32, _P(ctypes.c_int32), ctypes.c_uint32]
        L.epicpu_resize_hospital.argtypes = [_P(ctypes.c_int32), ctypes.c_int32, ctypes.c_double, ctypes.c_double]
        L.epicpu_increase_hospital_size.argtypes = [_P(ctypes.c_int32), ctypes.c_uint32]
        L.epicpu_transmission_rate.restype = ctypes.c_double
        L.epicpu_transmission_rate.argtypes = [_P(EpicpuConfig), ctypes.c_uint32]
        L.epicpu_is_to_be_hospitalized.argtypes = [_P(EpicpuConfig), ctypes.c_uint32]
        L.epicpu_neighbors_in_area.restype = ctypes.c_uint32
        L.epicpu_neighbors_in_area.argtypes = [ctypes.c_int32, ctypes.c_int32, _P(ctypes.c_int32), _P(ctypes.c_int32)]
        L.epicpu_area_iter.restype = ctypes.c_uint32
        L.epicpu_area_iter.argtypes = [_P(ctypes.c_int32), _P(ctypes.c_int32), ctypes.c_uint32]
        L.epicpu_bernoulli_threshold.restype = ctypes.c_uint64
        L.epicpu_bernoulli_threshold.argtypes = [ctypes.c_double]
        L.epicpu_lockdown_should_apply.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        L.epicpu_lockdown_should_unlock.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32]
        L.epicpu_hospital_counts_updated.restype = ctypes.c_uint32
        L.epicpu_hospital_counts_updated.argtypes = [ctypes.c_uint32] * 3
        L.epicpu_hospital_should_apply.argtypes = [ctypes.c_int] + [ctypes.c_uint32] * 4
        L.epicpu_first_vacant.argtypes = [_P(ctypes.c_int32), _P(ctypes.c_int32), ctypes.c_uint32, _P(ctypes.c_int32)]
        L.epicpu_create_scenario.restype = ctypes.c_void_p
        L.epicpu_create_scenario.argtypes = [_P(EpicpuConfig), ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                             _P(ctypes.c_int32), ctypes.c_char_p, ctypes.c_size_t]
        _libs[path] = L
    return _libs[path]


def _arr(ctype, values):
    return (ctype * len(values))(*values)


def config_from_json(cfg, seed=None):
    """Build an EpicpuConfig from a reference-schema config (dict or JSON text)."""
    if isinstance(cfg, (str, bytes)):
        cfg = json.loads(cfg)
    c = EpicpuConfig()
    pop = cfg["population"]
    if "Csv" in pop:
        working, pub = read_population_csv(pop["Csv"]["file"])
        c._csv = (working, pub)  # keep the arrays alive with the struct
        c.csv_n = len(working)
        c.csv_working = working.ctypes.data
        c.csv_pub_transport = pub.ctypes.data
        c.number_of_agents = len(working)
    else:
        auto = pop["Auto"]
        c.number_of_agents = auto["number_of_agents"]
        c.public_transport_percentage = auto["public_transport_percentage"]
        c.working_percentage = auto["working_percentage"]
    d = cfg["disease"]
    for k in ("regular_transmission_start_day", "high_transmission_start_day", "last_day", "asymptomatic_last_day",
              "mild_infected_last_day", "regular_transmission_rate", "high_transmission_rate", "death_rate",
              "percentage_asymptomatic_population", "percentage_severe_infected_population", "exposed_duration",
              "pre_symptomatic_duration"):
        setattr(c, k, d[k])
    geo = cfg["geography_parameters"]
    c.grid_size = geo["grid_size"]
    c.hospital_beds_percentage = geo["hospital_beds_percentage"]
    c.hours = cfg["hours"]
    si = cfg.get("starting_infections", {"infected_mild_asymptomatic": 0, "infected_mild_symptomatic": 0,
                                          "infected_severe": 0, "exposed": 1})
    c.si_infected_mild_asymptomatic = si["infected_mild_asymptomatic"]
    c.si_infected_mild_symptomatic = si["infected_mild_symptomatic"]
    c.si_infected_severe = si["infected_severe"]
    c.si_exposed = si["exposed"]
    nv = 0
    for iv in cfg.get("interventions", []):
        (kind, body), = iv.items()
        if kind == "Lockdown" and not c.has_lockdown:
            c.has_lockdown = 1
            c.lockdown_at_number_of_infections = body["at_number_of_infections"]
            c.lockdown_essential_workers_population = body["essential_workers_population"]
        elif kind == "BuildNewHospital" and not c.has_build_new_hospital:
            c.has_build_new_hospital = 1
            c.spread_rate_threshold = body["spread_rate_threshold"]
        elif kind == "Vaccinate":
            c.vaccinate_at_hour[nv] = body["at_hour"]
            c.vaccinate_percent[nv] = body["percent"]
            nv += 1
    c.n_vaccinate = nv
    c.seed = int(seed if seed is not None else cfg.get("seed", 0)) & 0xFFFFFFFFFFFFFFFF
    return c


def read_population_csv(path):
    """Population::Csv records (engine/src/citizen/population_record.rs:23-43) read with
    Python's csv module: columns ind, age, working, pub_transport; flags "True"/"False"."""
    import csv
    working, pub = [], []
    with open(path, newline="") as f:
        for rec in csv.DictReader(f):
            for k in ("ind", "age", "working", "pub_transport"):
                if k not in rec:
                    raise ValueError("missing field `%s`" % k)
            int(rec["ind"])
            flags = []
            for k in ("working", "pub_transport"):
                if rec[k] not in ("True", "False"):
                    raise ValueError("expected True or False, got %r" % rec[k])
                flags.append(rec[k] == "True")
            working.append(flags[0])
            pub.append(flags[1])
    return np.array(working, dtype=np.uint8), np.array(pub, dtype=np.uint8)


def alloc_agents(n):
    arrays = {name: np.zeros(n * width, dtype=dt) for name, dt, width in AGENT_FIELDS}
    s = EpicpuAgents(**{name: a.ctypes.data for name, a in arrays.items()})
    return arrays, s


class OracleSim:
    """One deterministic-mode simulation on the CPU oracle."""

    LIB = LIB_PATH

    def __init__(self, cfg, seed=None):
        self.L = lib(self.LIB)
        self.cfg = config_from_json(cfg, seed)
        err = ctypes.create_string_buffer(512)
        self.h = self.L.epicpu_create(ctypes.byref(self.cfg), err, 512)
        if not self.h:
            raise RuntimeError(err.value.decode())
        self.n = self.cfg.number_of_agents

    @classmethod
    def scenario(cls, cfg, agents, last_hour, hospital=None, seed=None):
        """A simulation whose population is exactly `agents` (dicts of EpicpuAgentInit
        fields; ids in list order) that has run hour `last_hour` (known-answer tests)."""
        self = cls.__new__(cls)
        self.L = lib(cls.LIB)
        self.cfg = config_from_json(cfg, seed)
        arr = (EpicpuAgentInit * max(len(agents), 1))()
        for i, a in enumerate(agents):
            for k, v in a.items():
                if k in ("home", "work", "current_area"):
                    getattr(arr[i], k)[:] = list(v)
                else:
                    setattr(arr[i], k, v)
        hosp = _arr(ctypes.c_int32, hospital) if hospital is not None else None
        err = ctypes.create_string_buffer(512)
        self.h = self.L.epicpu_create_scenario(ctypes.byref(self.cfg), ctypes.cast(arr, ctypes.c_void_p), len(agents),
                                              last_hour, hosp, err, 512)
        if not self.h:
            raise RuntimeError(err.value.decode())
        self.n = len(agents)
        return self

    def set_threads(self, threads):
        self.L.epicpu_set_threads(self.h, threads)

    def set_reference_order(self, on=True):
        """Sequential per-agent-hour streams in the reference's call order and a random
        merge order (distributional checks; not bit-comparable with deterministic mode)."""
        if self.L.epicpu_set_reference_order(self.h, 1 if on else 0) != 0:
            raise MemoryError("epicpu_set_reference_order")

    def set_random_merge(self, on=True):
        """Deterministic-mode draws, collisions won by the first agent of a random order per
        hour instead of the lowest id (the collision rule alone, for distributional checks)."""
        if self.L.epicpu_set_random_merge(self.h, 1 if on else 0) != 0:
            raise MemoryError("epicpu_set_random_merge")

    def close(self):
        if self.h:
            self.L.epicpu_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def step(self, n_hours):
        rows = np.zeros((max(n_hours, 1), 7), dtype=np.uint32)
        nw = ctypes.c_uint32(0)
        st = ctypes.c_int(0)
        rc = self.L.epicpu_step(self.h, n_hours, rows.ctypes.data_as(_P(ctypes.c_uint32)), ctypes.byref(nw),
                               ctypes.byref(st))
        if rc != 0:
            raise RuntimeError("epicpu_step failed: %d" % rc)
        return rows[: nw.value], bool(st.value)

    def agents(self):
        arrays, s = alloc_agents(self.n)
        self.L.epicpu_read_agents(self.h, ctypes.byref(s))
        arrays["current_area"] = arrays["current_area"].reshape(-1, 4)
        return arrays

    def counts(self):
        out = (ctypes.c_uint32 * 7)()
        self.L.epicpu_counts(self.h, out)
        return list(out)

    def events(self):
        cap = 65536
        hours = (ctypes.c_uint32 * cap)()
        kinds = (ctypes.c_uint32 * cap)()
        n = self.L.epicpu_events(self.h, hours, kinds, cap)
        return [(hours[i], kinds[i]) for i in range(min(n, cap))]

    def geometry(self):
        r = (ctypes.c_int32 * 20)()
        c = (ctypes.c_uint32 * 3)()
        self.L.epicpu_geometry(self.h, r, c)
        rects = [tuple(r[4 * i:4 * i + 4]) for i in range(5)]
        return rects, list(c)


class RefCpuSim(OracleSim):
    """The same routine on the reference's store (oracle/librefcpu.so): citizens by value
    in two FNV hash maps keyed by cell, swapped every hour (allocation_map.rs:44-65), a
    threaded map over the buckets, a single-thread merge in bucket order so collision
    winners follow hash order, and draws in the reference's call order. The timed CPU
    baseline and the stand-in for the unmodified engine; not bit-comparable with
    deterministic mode."""

    LIB = REF_LIB_PATH

    def set_reference_order(self, on=True):
        if not on:
            raise ValueError("refcpu always draws in the reference's order")

    def set_random_merge(self, on=True):
        if on:
            raise ValueError("refcpu merges in bucket order")


def philox(ctr, key):
    out = (ctypes.c_uint32 * 4)()
    lib().epicpu_philox4x32_10(_arr(ctypes.c_uint32, ctr), _arr(ctypes.c_uint32, key), out)
    return list(out)
