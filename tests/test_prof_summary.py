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


def test_hourly_counters_follow_the_commit_count(tmp_path):
    """Two days of per-hour dispatches (--profile-hours: hours 1..6 as commits alone): each
    k_step / k_commit counter value lands on its hour of day, averaged over the days, and
    the L2 hit rate is TCC_HIT / (TCC_HIT + TCC_MISS)."""
    path = tmp_path / "counters.csv"
    fields = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        d = 0
        for day_i in range(2):
            for h in range(1, 25):
                hd = h % 24
                ks = ["k_commit"] if 1 <= hd <= 6 else ["k_step", "k_commit"]
                for k in ks:
                    d += 1
                    hit = 100 * hd + (10 if k == "k_commit" else 0) + day_i  # day 0 and 1 differ by 1
                    w.writerow({"Dispatch_Id": d, "Kernel_Name": NAME % k, "Counter_Name": "TCC_HIT_sum",
                                "Counter_Value": hit})
                    w.writerow({"Dispatch_Id": d, "Kernel_Name": NAME % k, "Counter_Name": "TCC_MISS_sum",
                                "Counter_Value": 3 * hit})
    text = P.hourly([str(path)], "cmd")
    table = {}
    for line in text.splitlines():
        cells = [c.strip() for c in line.strip("|").split("|")]
        if len(cells) == 5 and cells[0].isdigit():
            table[(int(cells[0]), cells[1])] = (float(cells[2]), float(cells[3]), cells[4])
    # four significant digits are printed
    assert table[(9, "k_step")] == (900.5, 2702.0, "25%")
    assert table[(9, "k_commit")] == (910.5, 2732.0, "25%")
    assert table[(0, "k_step")] == (0.5, 1.5, "25%")
    assert (3, "k_commit") not in table  # only hours of day 7..23 and 0 are listed
