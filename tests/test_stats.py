"""Production statistics against the reference's own printed report
(tests/golden/stats_ref.json, made by make_stats_ref.py): the oracle replays
each reference run from its printed seeds, driven by the library's main-loop
schedule (fastmc_amd.gcmc.run_main_loop: equilibration, production start at
a check, the end after the check that finds the production count reached,
the final window); its averaging windows (gcmc.f:1050-1130, with the
surface_check tally of mc_moves.f:10-44) go through the library's host
statistics (fmc_stats_from_windows: calc_Qst / calc_Cv, window-weighted
means and deviations, avunion / stdunion).  Every value of the node report
(f20.9) and of the final report (f15.6, errors f12.3 / f15.6) must match as
printed, NaN included (the 'sampling' case closes an empty window, as the
reference does).  Two runs count equilibration and / or production in
cycles (negative step counts), whose ends and windows follow the guest
count at each check."""
import json
import math
import os
import shutil
import tempfile

import pytest

from _helpers import GOLDEN
from fastmc_amd import gcmc
from fastmc_amd import inputs as I
from fastmc_amd.lib import stats_from_windows
from fastmc_amd.system import read_system
from oracle.oracle import Oracle

REF = json.load(open(os.path.join(GOLDEN, "stats_ref.json")))


def stats_system(case):
    d = tempfile.mkdtemp(prefix="fmc_stats_")
    try:
        if case["name"].startswith("calf20"):
            for f in ("FIELD", "CONFIG"):
                shutil.copy(os.path.join(GOLDEN, "calf20", f), os.path.join(d, f))
        else:
            I.config0(d, pressure_bar=1.0)
        with open(os.path.join(d, "CONTROL"), "w") as fh:
            fh.write(case["control"])
        return read_system(d)
    finally:
        shutil.rmtree(d, ignore_errors=True)


def printed(ours, text, decimals):
    """`ours` rounds to the printed `text` (Fortran F format, NaN as NaN)."""
    if text == "NaN":
        return math.isnan(ours)
    v = float(text)
    # half a unit in the last printed place, plus a few ulps of the value
    return abs(ours - v) <= 0.5 * 10.0 ** (-decimals) * 1.0001 + 4e-16 * abs(v)


class OracleEngine:
    def __init__(self, o):
        self.o = o

    def run(self, n, flags):
        self.o.run(n, stats=flags)

    def prodcount(self):
        return self.o.windows()[2]

    def guests(self):
        return self.o.guest_total()


def replay(case):
    sysm = stats_system(case)
    o = Oracle(sysm, seed=tuple(case["seeds"]))
    nwind, nwindsteps = gcmc.window_steps(sysm.steps, sysm.nwind)
    assert nwindsteps == case["nwindsteps"]
    # in cycles the windows close by cycle count (run_main_loop's CLOSE flags)
    o.set_sampling(0 if sysm.steps < 0 else nwindsteps, sysm.accsample)
    eng = OracleEngine(o)
    minchk = gcmc.check_interval(sysm.eqsteps, sysm.nnumg)
    r = gcmc.run_main_loop(eng, sysm.eqsteps, sysm.steps, minchk, sysm.accsample, eng.prodcount, nwindsteps)
    win, nw, pc = o.windows()
    o.close()
    return sysm, r, win, nw, pc


@pytest.mark.parametrize("case", REF, ids=[c["name"] for c in REF])
def test_oracle_statistics_match_reference_report(case):
    sysm, r, win, nw, pc = replay(case)
    assert (r.start, r.steps) == (case["start"], case["total"])
    assert pc == r.prodcount == int(case["node"][0]["Multiplier"])
    stats = stats_from_windows(win, nw, len(sysm.guests), sysm.temperature)
    node_keys = [("<N>", "avg", "<N>"), ("<E>", "avg", "<E>"), ("<E*N>", "avg", "<E*N>"), ("<N*N>", "avg", "<N*N>"),
                 ("<E*E>", "avg", "<E*E>"), ("Qst", "avg", "Qst"), ("Qst_err", "std", "Qst"), ("Cv", "avg", "Cv"),
                 ("Cv_err", "std", "Cv"), ("<NF>", "avg", "<NF>")]
    for g, (ref_node, ref_final) in enumerate(zip(case["node"], case["final"])):
        st = stats[g]
        for key, part, name in node_keys:
            assert printed(st[part][name], ref_node[key], 9), (g, key, st[part][name], ref_node[key])
        for name in ("<N>", "<E>", "<E*N>", "<N*N>", "<E*E>"):
            assert printed(st["avg_final"][name], ref_final[name], 6), (g, name)
            assert printed(st["std_final"][name], ref_final[name + "_err"], 3), (g, name + "_err")
        finals = [("avg_guests", "avg_final", "<N>"), ("avg_guests_err", "std_final", "<N>"),
                  ("Qst", "avg_final", "Qst"), ("Qst_err", "std_final", "Qst"), ("surface", "avg_final", "<NF>"),
                  ("surface_err", "std_final", "<NF>"), ("Cv", "avg_final", "Cv"), ("Cv_err", "std_final", "Cv")]
        for key, part, name in finals:
            assert printed(st[part][name], ref_final[key], 6), (g, key, st[part][name], ref_final[key])
        assert int(ref_final["counted"]) == pc


def test_schedule_closed_form():
    # gcmc.f:704-765 with checks every minchk steps
    assert gcmc.schedule(0, 3000, 400) == (1, 3200)
    assert gcmc.schedule(1000, 4000, 400) == (1200, 5200)
    assert gcmc.schedule(750000, 500000, 400) == (750000, 1250000)
    assert gcmc.check_interval(750000) == 400 and gcmc.check_interval(2000) == 2 and gcmc.check_interval(600) == 1
    assert gcmc.window_steps(500000) == (5, 100000) and gcmc.window_steps(1000) == (1, 1000)


def test_report_parses_like_the_reference_test_harness(tmp_path):
    """The final report's lines parse as run_test.py:71-76 reads OUTPUT."""
    case = REF[0]
    sysm, r, win, nw, pc = replay(case)
    stats = stats_from_windows(win, nw, len(sysm.guests), sysm.temperature)
    names = [gcmc.species_name(s) for s in sysm.guests]
    lines = gcmc.final_report(names, stats, pc, sysm.surftol)
    avg, std = [], []
    for line in lines:
        if "average number of guests" in line:
            avg.append(float(line.strip().split()[4]))
        if "standard error" in line:
            std.append(float(line.strip().split()[2]))
    assert avg == [float(f["avg_guests"]) for f in case["final"]]
    # two 'standard error' lines per guest (N, then surface), as the reference prints
    assert [std[0], std[2]] == [float(f["avg_guests_err"]) for f in case["final"]]
