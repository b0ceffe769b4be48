import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    if not os.path.exists(os.path.join(ROOT, "oracle", "libepicpu.so")):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def load_config(name):
    import json
    with open(os.path.join(ROOT, "tests", "golden", "configs", name)) as f:
        return json.load(f)


@pytest.fixture
def cfg1():
    return load_config("cfg1_10k.json")
