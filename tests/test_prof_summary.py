"""The rocprof summary's grouping of dispatches into simulated hours
(epirust_amd/prof_summary.py), on a synthetic kernel trace: one hour per k_commit, four
per k_quiet_hours (hours of day 3..6), six per k_night (1..6)."""
import csv

from epirust_amd import prof_summary as P

NAME = "epi::(anonymous namespace)::%s(epi::DevView)"


def write_trace(path, kernels):
    """kernels: (name, duration ns) dispatched back to back, 1 us apart."""
    t = 0
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        for k, d in kernels:
            w.writerow({"Kernel_Name": NAME % k, "Start_Timestamp": t, "End_Timestamp": t + d})
            t += d + 1000


def day(quiet):
    if quiet == "night":
        ks = [("k_night", 11000)]  # hours of day 1..6
    else:
        ks = [("k_commit", 6000), ("k_commit", 9000)]  # hours of day 1, 2
        ks += [("k_quiet_hours", 8000)] if quiet else [("k_commit", 4000)] * 4
    for hd in range(7, 25):
        if hd % 24 == 0:
            ks.append(("k_hospital_scan", 4000))
        if hd in (8, 16):
            ks.append(("k_relayout", 5000))
        ks += [("k_step", 20000), ("k_commit", 10000)]
    return ks


def rows(text):
    out = {}
    for line in text.splitlines():
        cells = [c.strip() for c in line.strip("|").split("|")]
        if cells and cells[0].isdigit():
            out[int(cells[0])] = [float(c) for c in cells[1:]]
    return out


def test_hours_are_counted_by_commits_and_quiet_dispatches(tmp_path):
    for quiet in (False, True, "night"):
        p = tmp_path / ("q.csv" if quiet else "c.csv")
        write_trace(p, day(quiet) + day(quiet))
        r = rows(P.by_hour_of_day(str(p), 48))
        cols = list(P.HOUR_KERNELS)
        assert r[0][cols.index("k_hospital_scan")] == 4.0 and r[0][cols.index("k_step")] == 20.0
        assert r[8][cols.index("k_relayout")] == 5.0 and r[9][cols.index("k_relayout")] == 0.0
        assert r[7][cols.index("k_step")] == 20.0 and r[7][cols.index("k_commit")] == 10.0
        if quiet == "night":
            assert r[1][cols.index("k_night")] == 11.0
            assert all(sum(r[hd][:len(cols)]) == 0.0 for hd in (2, 3, 4, 5, 6))
            assert all(r[hd][-1] == 2.0 for hd in (1, 2, 3, 4, 5, 6))  # (11 + 1) / 6
            continue
        assert r[1][cols.index("k_commit")] == 6.0 and r[2][cols.index("k_commit")] == 9.0
        if quiet:
            assert r[3][cols.index("k_quiet_hours")] == 8.0
            assert all(sum(r[hd][:len(cols)]) == 0.0 for hd in (4, 5, 6))
            # the quiet dispatch's 9 us (8 running + 1 gap) split over its four hours
            assert all(r[hd][-1] == 2.2 for hd in (3, 4, 5, 6))
        else:
            assert all(r[hd][cols.index("k_commit")] == 4.0 for hd in (3, 4, 5, 6))
