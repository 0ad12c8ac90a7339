"""Drop-in boundary: the reference's own main program (gcmc.f and the
reference modules, compiled from the reference sources by `make -C oracle
dropin`) running over the GPU module mc_moves (fastmc_amd/fortran/
mc_moves_fmc.f90 -> libfastmc_hip.so).

Every trial the Fortran chain sent to the GPU is logged by the module
(FMC_SHIM_TRACE) and replayed through the oracle's one-trial seam with the
same accept/reject decisions: same overlap flags, dU within the FP64 bound
(mixed: within the mixed bound).  With the duni seeds the run printed, the
oracle's own chain must make the same sequence of moves and decisions, i.e.
gcmc.f over the GPU module is the reference chain, and the report it prints
(production statistics with the device's surface flags) equals the oracle
chain's statistics as printed."""
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from _helpers import GOLDEN, KB_KCAL, du_tol
from fastmc_amd import inputs as I
from fastmc_amd.system import read_system
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
EXE = os.path.join(ROOT, "oracle", "_ref", "dropin", "gcmc_fmc.x")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(tmp, steps, precision, calf20=False):
    if calf20:
        from make_moves_ref import control
        for f in ("FIELD", "CONFIG"):
            shutil.copy(os.path.join(GOLDEN, "calf20", f), os.path.join(tmp, f))
        open(os.path.join(tmp, "CONTROL"), "w").write(control(steps))
    else:
        I.config0(tmp, pressure_bar=10.0, steps=steps, equilibration=0)
    trace = os.path.join(tmp, "shim.trace")
    env = dict(os.environ, FMC_SHIM_TRACE=trace, FMC_SHIM_PRECISION=precision)
    r = subprocess.run([EXE], cwd=tmp, env=env, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    out = open(os.path.join(tmp, "OUTPUT"), errors="replace").read()
    _run.report = out
    m = re.search(r"i,j,k,l values for duni\(\)\s+(\d+)\s+(\d+)\s+(\d+)\s+(\d+)", out)
    seeds = tuple(int(x) for x in m.groups())
    trials = []
    _run.events = []
    for ln in open(trace):
        f = ln.split()
        _run.events.append(f[0])
        if f[0] == "T":
            kind, g, choice, natms = (int(x) for x in f[1:5])
            est = float(f[5].replace("D", "E"))
            ovl = int(f[6])
            xyz = np.array([float(x) for x in f[7:7 + 3 * natms]]).reshape(natms, 3)
            trials.append({"kind": kind, "g": g, "choice": choice, "estep": est, "overlap": ovl,
                           "xyz": xyz, "accepted": None})
        elif f[0] == "D":
            trials[-1]["accepted"] = int(f[1])
    assert all(t["accepted"] is not None for t in trials)
    return seeds, trials, read_system(tmp)


def _replay(sysm, seeds, trials, tol_scale):
    o = Oracle(sysm, seed=seeds)
    o.run(0)
    T = sysm.temperature
    worst = 0.0
    for i, t in enumerate(trials):
        if t["kind"] == 1:
            a = o.trial_insert(t["g"], t["xyz"])
        elif t["kind"] == 2:
            a = o.trial_delete(t["g"], t["choice"])
        else:
            a = o.trial_displace(t["g"], t["choice"], t["xyz"])
        assert a.overlap == t["overlap"], (i, t["kind"])
        if not a.overlap or t["kind"] == 2:
            err = abs(a.estep - t["estep"]) / max(abs(a.estep), KB_KCAL * T)
            worst = max(worst, err)
            assert abs(a.estep - t["estep"]) <= du_tol(a.estep, T) * tol_scale, (i, t["kind"], a.estep, t["estep"])
        if t["accepted"]:
            o.accept()
        else:
            o.reject()
    return o, worst


@pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in build absent (make -C oracle dropin)")
def test_dropin_gcmc_main_program_fp64():
    tmp = tempfile.mkdtemp(prefix="fmc_dropin_")
    try:
        say("drop-in: reference gcmc.f over the GPU mc_moves module, FP64")
        seeds, trials, sysm = _run(tmp, 1000, "fp64")
        say("drop-in:", len(trials), "trials, seeds", seeds)
        assert len(trials) >= 1000
        o, worst = _replay(sysm, seeds, trials, 1.0)
        say("drop-in: worst scaled dU difference", worst)
        # the oracle's own chain from the printed seeds makes the same moves and decisions
        ref = Oracle(sysm, seed=seeds).run(len(trials))
        for i, t in enumerate(trials):
            r = ref[i]
            assert min(int(r["move"]), 3) == t["kind"], i
            assert int(r["accepted"]) == t["accepted"], i
            if t["kind"] != 1:
                assert int(r["choice"]) == t["choice"], i
        # its printed statistics are the oracle chain's (surface counts included)
        from fastmc_amd import gcmc
        from fastmc_amd.lib import stats_from_windows
        from test_stats import OracleEngine, printed
        oc = Oracle(sysm, seed=seeds)
        _, nwindsteps = gcmc.window_steps(sysm.steps, sysm.nwind)
        oc.set_sampling(nwindsteps)
        eng = OracleEngine(oc)
        gcmc.run_main_loop(eng, sysm.eqsteps, sysm.steps, gcmc.check_interval(sysm.eqsteps, sysm.nnumg), False,
                           eng.prodcount)
        win, nw, _ = oc.windows()
        st = stats_from_windows(win, nw, 1, sysm.temperature)[0]
        final = _run.report[_run.report.index("final stats"):]
        got = dict(re.findall(r"\n\s*(average number of guests|Isosteric heat of adsorption \(kcal/mol\)|"
                              r"average surface adsorption):\s+([-0-9.]+)", final))
        say("drop-in report:", got)
        assert printed(st["avg_final"]["<N>"], got["average number of guests"], 6)
        assert printed(st["avg_final"]["Qst"], got["Isosteric heat of adsorption (kcal/mol)"], 6)
        assert printed(st["avg_final"]["<NF>"], got["average surface adsorption"], 6)
        assert st["avg_final"]["<NF>"] > 0
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


@pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in build absent (make -C oracle dropin)")
def test_dropin_gcmc_main_program_mixed():
    tmp = tempfile.mkdtemp(prefix="fmc_dropin_")
    try:
        say("drop-in: reference gcmc.f over the GPU mc_moves module, mixed")
        seeds, trials, sysm = _run(tmp, 600, "mixed")
        o, worst = _replay(sysm, seeds, trials, 1e4)   # 1e-10 * 1e4 = the mixed bound 1e-6
        say("drop-in mixed: worst scaled dU difference", worst, "over", len(trials), "trials")
        assert worst <= 1e-6
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


@pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in build absent (make -C oracle dropin)")
def test_dropin_calf20_six_move_set_fp64():
    """gcmc.f over the GPU module with the reference test's six-move set on
    CALF-20 (two guest kinds, triclinic cell): the oracle's chain from the
    printed seeds makes every move and decision the drop-in chain made, each
    trial's dU agrees within the FP64 bound (a swap: its deletion and
    insertion together), rejected swaps were reloaded on the device, and the
    printed statistics of both guests are the oracle chain's."""
    from fastmc_amd.lib import FMC_MOVE_SWAP, FMC_MOVE_SWITCH
    tmp = tempfile.mkdtemp(prefix="fmc_dropin6_")
    try:
        say("drop-in: CALF-20, six-move set, FP64")
        seeds, trials, sysm = _run(tmp, 1500, "fp64", calf20=True)
        ev = _run.events
        say("drop-in:", len(trials), "trials,", ev.count("R"), "rejected swaps,", ev.count("L"), "reloads; seeds",
            seeds)
        # gcmc.f ends production at a check interval past `steps`: run the
        # oracle chain long enough and walk it until every trial is matched
        ref = Oracle(sysm, seed=seeds).run(sysm.steps + 1000)
        T = sysm.temperature
        i = 0
        nswap = swaps_acc = 0
        for n, r in enumerate(ref):
            if i == len(trials):
                break
            mv = int(r["move"])
            if mv == FMC_MOVE_SWITCH:
                continue  # two guest kinds: mc_switch returns before any trial
            if mv == FMC_MOVE_SWAP:
                d, ins = trials[i], trials[i + 1]
                assert (d["kind"], ins["kind"]) == (2, 1), (n, d["kind"], ins["kind"])
                assert d["choice"] == int(r["choice"]) and d["g"] == int(r["iguest"]), n
                assert ins["overlap"] == int(r["overlap"]), n
                if not ins["overlap"]:
                    du = ins["estep"] - d["estep"]
                    assert abs(du - r["estep"]) <= du_tol(r["estep"], T) * 10, (n, du, r["estep"])
                i += 2
                nswap += 1
                swaps_acc += int(r["accepted"])
                continue
            t = trials[i]
            assert t["kind"] == min(mv, 3), (n, mv, t["kind"])  # translate, rotate, jump: displacements
            assert t["g"] == int(r["iguest"]), n
            assert t["overlap"] == int(r["overlap"]), n
            assert t["accepted"] == int(r["accepted"]), n
            if t["kind"] != 1:
                assert t["choice"] == int(r["choice"]), n
            if not t["overlap"]:
                assert abs(t["estep"] - r["estep"]) <= du_tol(r["estep"], T), (n, t["estep"], r["estep"])
            i += 1
        assert i == len(trials)
        assert nswap > 0 and ev.count("R") > 0 and ev.count("L") > 0
        assert ev.count("R") == nswap - swaps_acc
        # the printed final statistics of both guests are the oracle chain's
        from fastmc_amd import gcmc
        from fastmc_amd.lib import stats_from_windows
        from test_stats import OracleEngine, printed
        oc = Oracle(sysm, seed=seeds)
        _, nwindsteps = gcmc.window_steps(sysm.steps, sysm.nwind)
        oc.set_sampling(nwindsteps)
        eng = OracleEngine(oc)
        gcmc.run_main_loop(eng, sysm.eqsteps, sysm.steps, gcmc.check_interval(sysm.eqsteps, sysm.nnumg), False,
                           eng.prodcount)
        win, nw, _ = oc.windows()
        st = stats_from_windows(win, nw, 2, sysm.temperature)
        final = _run.report[_run.report.index("final stats"):]
        got = re.findall(r"\n\s*average number of guests:\s+([-0-9.]+)", final)
        say("drop-in report <N>:", got, "oracle", [s_["avg_final"]["<N>"] for s_ in st])
        assert len(got) == 2
        for g in range(2):
            assert printed(st[g]["avg_final"]["<N>"], got[g], 6)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
