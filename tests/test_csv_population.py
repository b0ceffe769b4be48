"""Population::Csv (engine/src/geography/grid.rs:194-231, citizen/mod.rs:155-180): the
records' working / pub_transport flags are used as given (no transport-point cap), ids are
file order, and the reference's capacity panic is an error."""
import copy
import os

import numpy as np
import pytest

from oracle import epicpu as O


def write_population(path, n, seed=4):
    rng = np.random.default_rng(seed)
    with open(path, "w") as f:
        f.write("ind,age,sex,working,pub_transport\n")
        for i in range(n):
            w = rng.random() < 0.6
            pt = rng.random() < 0.3
            f.write("%d,%d-%d,%s,%s,%s\n" % (i, 20 + i % 50, 24 + i % 50, "F" if i % 2 else "M",
                                             "True" if w else "False", "True" if pt else "False"))
    return path


def csv_config(cfg1, path):
    c = copy.deepcopy(cfg1)
    c["population"] = {"Csv": {"file": path, "cols": ["age", "sex", "working", "pub_transport"]}}
    c["starting_infections"]["exposed"] = 3
    return c


def test_oracle_uses_record_flags(cfg1, tmp_path):
    path = write_population(str(tmp_path / "pop.csv"), 5000)
    working, pub = O.read_population_csv(path)
    s = O.OracleSim(csv_config(cfg1, path))
    a = s.agents()
    assert len(a["x"]) == 5000
    assert np.array_equal(a["work_status"] != 3, working.astype(bool))
    assert np.array_equal((a["flags"] & 16) != 0, pub.astype(bool))  # uncapped, even for non-workers
    rows, _ = s.step(200)
    assert (rows[:, 1:].sum(axis=1) == 5000).all()


def test_capacity_panic_text(cfg1, tmp_path):
    path = write_population(str(tmp_path / "big.csv"), 6250 * 4 + 1)
    with pytest.raises(RuntimeError, match="Cannot accommodate citizens into homes! There are 25001 citizens, "
                                           "but 25000 home points"):
        O.OracleSim(csv_config(cfg1, path))


def test_bad_boolean_is_rejected(tmp_path):
    p = tmp_path / "bad.csv"
    p.write_text("ind,age,working,pub_transport\n0,30,yes,False\n")
    with pytest.raises(ValueError):
        O.read_population_csv(str(p))
