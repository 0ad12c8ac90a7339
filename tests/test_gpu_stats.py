"""Production statistics of the device chains (fmc_set_sampling, the window
records of k_run, fmc_production_stats) against the reference's own printed
report and the oracle: each reference run of tests/golden/stats_ref.json is
replayed in FP64 on the device from the reference's printed seeds through the
library's main-loop driver.  The windows must hold the oracle's step counts,
N and surface sums exactly and its energy sums to 1e-10 relative; every
printed statistic must match the reference's to its printed digits (or
1e-12 relative where that is looser).  The surface_check flags and the
surface count of every step's trace match the oracle's.  The run-directory
driver writes an OUTPUT that the reference's test harness parses."""
import math
import os
import shutil
import sys

import numpy as np
import pytest

from test_stats import REF, OracleEngine, stats_system
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def printed_dev(ours, text, decimals):
    if text == "NaN":
        return math.isnan(ours)
    v = float(text)
    return abs(ours - v) <= max(0.5 * 10.0 ** (-decimals) * 1.0001, 1e-12 * abs(v))


class DeviceEngine:
    def __init__(self, h):
        self.h = h

    def run(self, n, flags):
        self.h.run(n, count_stats=flags)

    def prodcount(self):
        return int(self.h.stats(0).prodcount)

    def guests(self):
        st = self.h.stats(0)
        return int(sum(st.nmols[g] for g in range(len(self.h.sysm.guests))))


@pytest.mark.parametrize("case", REF, ids=[c["name"] for c in REF])
def test_gpu_statistics_match_reference_and_oracle(case):
    from fastmc_amd import gcmc
    from fastmc_amd.lib import FMC_FP64, Handle
    say(case["name"])
    sysm = stats_system(case)
    ng = len(sysm.guests)
    nwind, nwindsteps = gcmc.window_steps(sysm.steps, sysm.nwind)
    minchk = gcmc.check_interval(sysm.eqsteps, sysm.nnumg)
    h = Handle(sysm, precision=FMC_FP64, seeds=[tuple(case["seeds"])])
    try:
        h.set_sampling(0 if sysm.steps < 0 else nwindsteps, sysm.accsample, 64)
        eng = DeviceEngine(h)
        r = gcmc.run_main_loop(eng, sysm.eqsteps, sysm.steps, minchk, sysm.accsample, eng.prodcount, nwindsteps)
        win, nw = h.windows(0)
        st = h.stats(0)
        stats = h.production_stats(0)
    finally:
        h.close()
    o = Oracle(sysm, seed=tuple(case["seeds"]))
    o.set_sampling(0 if sysm.steps < 0 else nwindsteps, sysm.accsample)
    oe = OracleEngine(o)
    ro = gcmc.run_main_loop(oe, sysm.eqsteps, sysm.steps, minchk, sysm.accsample, oe.prodcount, nwindsteps)
    owin, onw, opc = o.windows()
    osurf = o.surface()
    o.close()
    assert (r.start, r.steps) == (ro.start, ro.steps) == (case["start"], case["total"])
    assert nw == onw and int(st.prodcount) == opc == int(case["node"][0]["Multiplier"])
    assert st.overflow == 0
    assert list(st.surface[:ng]) == list(osurf[:ng])
    for j in range(nw):
        assert win[j].count == owin[j].count
        for g in range(ng):
            a, b = np.array(win[j].sums[g][:]), np.array(owin[j].sums[g][:])
            assert a[0] == b[0] and a[3] == b[3] and a[5] == b[5], (j, g, a, b)
            for q in (1, 2, 4):
                assert abs(a[q] - b[q]) <= 1e-10 * max(abs(b[q]), 1.0), (j, g, q, a[q], b[q])
    node_keys = [("<N>", "avg", "<N>"), ("<E>", "avg", "<E>"), ("<E*N>", "avg", "<E*N>"), ("<N*N>", "avg", "<N*N>"),
                 ("<E*E>", "avg", "<E*E>"), ("Qst", "avg", "Qst"), ("Qst_err", "std", "Qst"), ("Cv", "avg", "Cv"),
                 ("Cv_err", "std", "Cv"), ("<NF>", "avg", "<NF>")]
    for g in range(ng):
        for key, part, name in node_keys:
            assert printed_dev(stats[g][part][name], case["node"][g][key], 9), (g, key, stats[g][part][name])
        fin = case["final"][g]
        for key, part, name in [("avg_guests", "avg_final", "<N>"), ("avg_guests_err", "std_final", "<N>"),
                                ("Qst", "avg_final", "Qst"), ("Qst_err", "std_final", "Qst"),
                                ("surface", "avg_final", "<NF>"), ("surface_err", "std_final", "<NF>"),
                                ("Cv", "avg_final", "Cv"), ("Cv_err", "std_final", "Cv")]:
            assert printed_dev(stats[g][part][name], fin[key], 6), (g, key, stats[g][part][name], fin[key])


def test_gpu_trace_surface_flags_match_oracle():
    from fastmc_amd.lib import FMC_FP64, Handle
    case = REF[0]   # CALF-20, insert / delete / displace / jump / swap / switch
    sysm = stats_system(case)
    n = 3000
    h = Handle(sysm, precision=FMC_FP64, seeds=[tuple(case["seeds"])])
    try:
        td = h.run(n, trace=True)
    finally:
        h.close()
    o = Oracle(sysm, seed=tuple(case["seeds"]))
    to = o.run(n)
    o.close()
    assert np.array_equal(td["move"], to["move"]) and np.array_equal(td["accepted"], to["accepted"])
    clean = to["overlap"] == 0   # an overlapping trial's flags stop at the overlap in the reference
    say("surface flags set on", int((to["surface"][clean] != 0).sum()), "of", int(clean.sum()), "clean steps")
    assert np.array_equal(td["surface"][clean], to["surface"][clean])
    assert np.array_equal(td["nsurface"], to["nsurface"])


def test_gpu_run_directory_writes_a_parsable_report(tmp_path):
    from fastmc_amd import gcmc
    from fastmc_amd import inputs as I
    I.config0(str(tmp_path), pressure_bar=1.0)
    p = tmp_path / "CONTROL"
    ctl = p.read_text().replace("steps    100000", "steps    2000").replace("equilibration    100000",
                                                                              "equilibration    1000")
    p.write_text(ctl)
    r, stats = gcmc.run_directory(str(tmp_path), seed=(12, 34, 56, 78))
    assert (r.start, r.steps, r.prodcount) == (1200, 3200, 2001)
    avg, std = [], []
    for line in (tmp_path / "OUTPUT").read_text().splitlines():
        if "average number of guests" in line:
            avg.append(float(line.strip().split()[4]))
        if "standard error" in line:
            std.append(float(line.strip().split()[2]))
    assert len(avg) == 1 and len(std) == 2
    assert abs(avg[0] - stats[0]["avg_final"]["<N>"]) <= 5e-7 and avg[0] > 0
