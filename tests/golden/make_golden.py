"""Writes oracle fixtures for full-size parity checks that the oracle cannot rerun within a
GPU test's time budget.

    python tests/golden/make_golden.py [cfg2] [cfg3] [cfg4] [--threads T] [--workers W]

cfg2_1m_oracle.npz: the 2,399 epicurve rows of the 1,000,000-agent / 2000x2000 run
(configs/cfg2_1m.json, seed 42, deterministic mode) and a SHA-256 digest of every agent's
final state, both computed by the CPU oracle (oracle/epicpu.c).

cfg3_10m_oracle.npz: the 10,000,000-agent / 6000x6000 run with lockdown, hospital growth
and daily vaccination (configs/cfg3_10m_interventions.json, 4,799 hours): every row, the
intervention events (hour, kind), and SHA-256 state digests after the hour that follows
each lockdown / unlock / hospital-growth event, after hours 504, 505 and 721 (the
standalone unlock hour and the first vaccination), and after the last hour.

cfg4_replicates_oracle.npz: every row of each of the 64 mixed replicates
(configs/cfg4_replicates.json, each run alone on the oracle for 2,399 hours or until it
stops), its row count and its final-state digest.
"""
import argparse
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
CFG3_FIXED_DIGEST_HOURS = (504, 505, 721)


def state_digest(agents):
    h = hashlib.sha256()
    for f in FIELDS:
        h.update(f.encode())
        h.update(np.ascontiguousarray(agents[f]).tobytes())
    return h.hexdigest()


def cfg3_digest_hours(events, last_hour):
    """Hours after which the cfg3 fixture holds a state digest: the hour after each lockdown,
    unlock or hospital-growth event (its sweep or area change has then acted on one hour),
    the fixed hours, and the last hour."""
    hours = {h + 1 for h, kind in events if kind in (1, 2, 3)}
    hours.update(CFG3_FIXED_DIGEST_HOURS)
    return sorted(x for x in hours if x < last_hour) + [last_hour]


def _oracle():
    sys.path.insert(0, ROOT)
    from oracle import epicpu as O
    return O


def make_cfg2(threads):
    O = _oracle()
    cfg = json.load(open(os.path.join(HERE, "configs", "cfg2_1m.json")))
    t = time.time()
    sim = O.OracleSim(cfg)
    sim.set_threads(threads)
    rows, stopped = sim.step(2399)
    digest = state_digest(sim.agents())
    np.savez_compressed(os.path.join(HERE, "cfg2_1m_oracle.npz"), rows=rows, digest=np.array(digest))
    print("cfg2: %d rows, stopped=%s, digest %s, %.1fs" % (len(rows), stopped, digest, time.time() - t))


def make_cfg3(threads):
    O = _oracle()
    cfg = json.load(open(os.path.join(HERE, "configs", "cfg3_10m_interventions.json")))
    t = time.time()
    sim = O.OracleSim(cfg)
    sim.set_threads(threads)
    rows, digests = [], {}
    h, stopped, seen = 0, False, 0
    pending = set(CFG3_FIXED_DIGEST_HOURS)
    while not stopped:
        r, stopped = sim.step(1)
        if len(r) == 0:
            break
        rows.append(r[0])
        h = int(r[0][0])
        ev = sim.events()
        for eh, kind in ev[seen:]:
            if kind in (1, 2, 3):
                pending.add(eh + 1)
        seen = len(ev)
        if h in pending or stopped:
            digests[h] = state_digest(sim.agents())
            print("cfg3: digest after hour %d (%.0fs)" % (h, time.time() - t), flush=True)
        if h % 240 == 0:
            print("cfg3: hour %d, row %s (%.0fs)" % (h, r[0].tolist(), time.time() - t), flush=True)
    events = np.array(sim.events(), dtype=np.uint32).reshape(-1, 2)
    want = cfg3_digest_hours([tuple(e) for e in events.tolist()], h)
    assert sorted(digests) == want, (sorted(digests), want)
    np.savez_compressed(os.path.join(HERE, "cfg3_10m_oracle.npz"), rows=np.array(rows, dtype=np.uint32),
                        events=events, digest_hours=np.array(want, dtype=np.uint32),
                        digests=np.array([digests[x] for x in want]))
    print("cfg3: %d rows, %d events, digests at %s, %.1fs" % (len(rows), len(events), want, time.time() - t))


def _run_replicate(args):
    i, cfg, threads = args
    O = _oracle()
    sim = O.OracleSim(cfg)
    sim.set_threads(threads)
    rows, _ = sim.step(2399)
    digest = state_digest(sim.agents())
    sim.close()
    return i, rows, digest


def make_cfg4(threads, workers):
    import concurrent.futures as cf
    cfgs = json.load(open(os.path.join(HERE, "configs", "cfg4_replicates.json")))
    t = time.time()
    out = {}
    with cf.ProcessPoolExecutor(max_workers=workers) as ex:
        for i, rows, digest in ex.map(_run_replicate, [(i, c, threads) for i, c in enumerate(cfgs)]):
            out[i] = (rows, digest)
            print("cfg4: replicate %d, %d rows, digest %s (%.0fs)" % (i, len(rows), digest[:16], time.time() - t),
                  flush=True)
    n_rows = np.array([len(out[i][0]) for i in range(len(cfgs))], dtype=np.uint32)
    rows = np.zeros((len(cfgs), 2399, 7), dtype=np.uint32)
    for i in range(len(cfgs)):
        rows[i, :n_rows[i]] = out[i][0]
    np.savez_compressed(os.path.join(HERE, "cfg4_replicates_oracle.npz"), rows=rows, n_rows=n_rows,
                        digests=np.array([out[i][1] for i in range(len(cfgs))]))
    print("cfg4: %d replicates, %.1fs" % (len(cfgs), time.time() - t))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("which", nargs="*", default=["cfg2", "cfg3", "cfg4"])
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--workers", type=int, default=1)
    a = ap.parse_args()
    for w in a.which:
        if w == "cfg2":
            make_cfg2(a.threads)
        elif w == "cfg3":
            make_cfg3(a.threads)
        elif w == "cfg4":
            make_cfg4(a.threads, a.workers)
        else:
            raise SystemExit("unknown fixture %s" % w)


if __name__ == "__main__":
    main()
