"""Writes oracle fixtures for full-size parity checks that the oracle cannot rerun within a
GPU test's time budget.

    python tests/golden/make_golden.py

cfg2_1m_oracle.npz: the 2,399 epicurve rows of the 1,000,000-agent / 2000x2000 run
(configs/cfg2_1m.json, seed 42, deterministic mode) and a SHA-256 digest of every agent's
final state, both computed by the CPU oracle (oracle/epicpu.c).
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
FIELDS = ["x", "y", "state", "severity", "infection_day", "at_hour", "flags", "work_status", "work_start_at",
          "immunity", "current_area"]


def state_digest(agents):
    h = hashlib.sha256()
    for f in FIELDS:
        h.update(f.encode())
        h.update(np.ascontiguousarray(agents[f]).tobytes())
    return h.hexdigest()


def main():
    sys.path.insert(0, ROOT)
    from oracle import epicpu as O
    cfg = json.load(open(os.path.join(HERE, "configs", "cfg2_1m.json")))
    t = time.time()
    sim = O.OracleSim(cfg)
    rows, stopped = sim.step(2399)
    digest = state_digest(sim.agents())
    np.savez_compressed(os.path.join(HERE, "cfg2_1m_oracle.npz"), rows=rows, digest=np.array(digest))
    print("cfg2: %d rows, stopped=%s, digest %s, %.1fs" % (len(rows), stopped, digest, time.time() - t))


if __name__ == "__main__":
    main()
