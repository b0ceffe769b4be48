"""epi_citizen_states_text_noalloc: the allocation-free variant of the citizen-state
stream's device formatter, checked through the C ABI against the stream that
epi_stream_citizen_states writes for the same hour, on two calls in a row, the second on a
smaller simulation (the scratch array keeps the first call's contents)."""
import copy
import ctypes

import pytest

import epirust_amd as G
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu


def small(cfg1, n, g):
    c = copy.deepcopy(cfg1)
    c["population"]["Auto"]["number_of_agents"] = n
    c["geography_parameters"]["grid_size"] = g
    c["disease"]["regular_transmission_rate"] = 0.3
    c["disease"]["high_transmission_rate"] = 0.6
    c["starting_infections"]["exposed"] = 20
    return c


def streamed_body(cfg, hours, path):
    sim = G.Simulation(cfg)
    sim.stream_citizen_states(path)
    rows, _ = sim.step(hours)
    sim.stream_citizen_states(None)
    sim.close()
    assert len(rows) == hours
    line = open(path).read().split("\n")[hours]
    head = '{"hr":%d,"citizen_states":[' % hours
    assert line.startswith(head) and line.endswith("]}")
    return line[len(head):-2].encode()


class DeviceBuffer:
    """Caller-owned device memory, from the HIP runtime the library itself uses."""

    def __init__(self, nbytes):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.n = nbytes
        self.p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(nbytes)) == 0

    def write(self, data):
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        assert self.hip.hipMemcpy(self.p, buf, ctypes.c_size_t(len(data)), 1) == 0  # host to device

    def read(self, nbytes):
        buf = ctypes.create_string_buffer(nbytes)
        assert self.hip.hipMemcpy(buf, self.p, ctypes.c_size_t(nbytes), 2) == 0  # device to host
        return buf.raw

    def free(self):
        assert self.hip.hipFree(self.p) == 0


def test_noalloc_text_equals_the_stream_on_two_calls(cfg1, tmp_path):
    L = G.lib()
    text = DeviceBuffer(3000 * 80)
    n_len = DeviceBuffer(8)
    for k, (n, g, hours) in enumerate(((3000, 120, 34), (1100, 70, 17))):
        cfg = small(cfg1, n, g)
        say("[noalloc] call %d: %d agents, hour %d" % (k + 1, n, hours))
        want = streamed_body(cfg, hours, str(tmp_path / ("s%d.jsonl" % k)))
        sim = G.Simulation(cfg)
        sim.step(hours)
        rc = L.epi_citizen_states_text_noalloc(sim._h, text.p, text.n, n_len.p)
        assert rc == 0, L.epi_last_error()
        got_len = int.from_bytes(n_len.read(8), "little")
        assert got_len == len(want)
        assert text.read(got_len) == want
        sim.close()
    text.free()
    n_len.free()


def test_noalloc_rejects_a_short_buffer_before_launching(cfg1):
    L = G.lib()
    sim = G.Simulation(small(cfg1, 500, 50))
    sim.step(3)
    text = DeviceBuffer(100)
    n_len = DeviceBuffer(8)
    n_len.write((7).to_bytes(8, "little"))
    rc = L.epi_citizen_states_text_noalloc(sim._h, text.p, text.n, n_len.p)
    assert rc == 4 and b"may need" in L.epi_last_error()
    assert int.from_bytes(n_len.read(8), "little") == 7  # nothing was launched
    sim.close()
    text.free()
    n_len.free()
