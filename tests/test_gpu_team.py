"""Teams: a walker's step split over several workgroups (each takes a share of
the pair atoms and of the k-vectors; the members exchange their sums once per
step and run the proposal and decision alike).  The chains must be the
reference's as with one workgroup per walker:

* the golden reference traces replayed with teams of 3 to 16 on one walker,
  in the all-in-LDS variant, the global-memory variants (the supercell of
  config 1, whose members keep private guest coordinates) and the
  three-species CALF-20 system;
* swap, jump and switch (the reference's CALF-20 move set) against the oracle
  with a team of 4;
* the 64 state points of the batched isotherm with teams of 2 and 4, walkers
  at each temperature following their own oracle chains through hipGraph
  sweeps;
* Wang-Landau windows on the config-4 supercell with a team of 2 against the
  oracle."""
import math
import sys

import numpy as np
import pytest

from _helpers import case_system, compare_trace, fixture
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name,team,env", [("cfg0_moves5", 4, ""), ("cfg0_p10", 16, ""),
                                           ("calf20_co2n2", 3, ""), ("cfg1_p10", 8, ""),
                                           ("cfg0_moves5", 4, "FMC_NO_LDS_ALL"),
                                           ("cfg0_p10", 2, "FMC_NO_LDS_ALL+FMC_NO_LJ_LDS")])
def test_gpu_team_replays_reference_trace(name, team, env, monkeypatch):
    from fastmc_amd.lib import FMC_FP64, Handle
    for e in filter(None, env.split("+")):
        monkeypatch.setenv(e, "1")
    fx = fixture(name)
    say(f"replay {name} with a team of {team} {env}")
    h = Handle(case_system(name), precision=FMC_FP64, seeds=[fx["seeds"]])
    h.set_team(team)
    assert h.get_team() == team
    tr = h.run(len(fx["trace"]["step"]), trace=True)
    bad, worst = compare_trace(tr, fx)
    assert not bad, f"{name}: first mismatching steps {bad[:5]}"
    assert worst <= 1e-10
    assert h.stats(0).overflow == 0
    last = fx["sk_checks"][-1]
    sk = h.sk()
    assert abs(sk[:, 0].sum() - last[1]) < 1e-8 * max(1.0, abs(last[1]))
    assert abs(sk[:, 1].sum() - last[2]) < 1e-8 * max(1.0, abs(last[2]))
    np.testing.assert_allclose(tr[-1]["energy"][: h.maxmls], fx["trace"]["energy"][-1], rtol=1e-9, atol=1e-9)
    h.close()


def test_gpu_team_moves_match_oracle():
    from fastmc_amd.lib import FMC_FP64, Handle
    from test_gpu_moves import first_difference
    from test_moves import REF, moves_system
    case = REF[0]
    sysm = moves_system(case)
    seeds = tuple(case["seeds"])
    n = min(case["steps"], 3000)
    say(f"{case['name']}: {n} steps with swap / jump / switch, team of 4")
    h = Handle(sysm, precision=FMC_FP64, seeds=[seeds])
    h.set_team(4)
    tr = h.run(n, trace=True)
    o = Oracle_run(sysm, seeds, n)
    d = first_difference(tr, o["trace"], n, sysm.temperature)
    assert d is None, d
    for g in (1, 2):
        np.testing.assert_allclose(h.guests(g), o["guests"][g], rtol=0, atol=1e-9)
    np.testing.assert_allclose(h.sk(), o["sk"], rtol=0, atol=1e-8)
    h.close()


def Oracle_run(sysm, seeds, n):
    o = O.Oracle(sysm, seed=seeds)
    out = {"trace": o.run(n), "guests": {g: o.guests(g) for g in (1, 2)}, "sk": o.sk()}
    o.close()
    return out


@pytest.mark.parametrize("team", [2, 4])
def test_gpu_team_batched_isotherm_follows_oracle(team):
    from fastmc_amd import inputs as I
    from fastmc_amd.lib import FMC_FP64
    from fastmc_amd.system import read_system
    from test_gpu_config2 import bench_seed, chosen, first_difference, make_handle
    import shutil
    import tempfile
    tmp = tempfile.mkdtemp(prefix="fmc_team_")
    try:
        systems = [read_system(d) for d in I.config2(tmp, n=64)]
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    pick = chosen(systems)
    h = make_handle(systems, FMC_FP64)
    h.set_team(team)
    ntr, nsw, nsteps = 400, 400, 3
    say(f"64 state points, team of {team}: {len(pick)} traced walkers, then {nsteps} graph sweeps of {nsw}")
    traces = []
    for w in pick:
        h.set_trace_walker(w)
        traces.append(h.run(ntr, trace=True))
    h.run_graph(nsteps, nsw)
    total = ntr * len(pick) + nsw * nsteps
    for k, w in enumerate(pick):
        s = systems[w]
        o = O.Oracle(s, seed=bench_seed(w))
        otr = o.run(total)
        d = first_difference(traces[k], otr[k * ntr:(k + 1) * ntr], s.temperature)
        assert d is None, (w, d)
        st = h.stats(w)
        assert st.steps == total and st.overflow == 0
        np.testing.assert_allclose(h.guests(1, w=w), o.guests(1), rtol=0, atol=1e-9)
        np.testing.assert_allclose(h.sk(w), o.sk(), rtol=0, atol=1e-8)
        say(f"  walker {w}: T={s.temperature:g} K, {total} steps identical, N={len(o.guests(1)) // 3}")
        o.close()
    h.close()


def test_gpu_team_wang_landau_windows_follow_oracle():
    from fastmc_amd.lib import FMC_FP64, Handle
    from test_gpu_config4 import config4_system, seeds
    sysm = config4_system()
    nw, nwin, maxn, nsteps = 128, 8, 400, 600
    sd = seeds(nw, 3)
    kw = dict(visittol=10000, flatcoeff=0.8, prectol=1e-6)
    say(f"config 4 with a team of 2: {nw} walkers, {nwin} windows, {nsteps} steps")
    h = Handle(sysm, precision=FMC_FP64, seeds=sd, n=nw)
    try:
        h.wl_setup(maxn=maxn, nwindows=nwin, **kw)
        h.set_team(2)
        h.run(nsteps)
        for k in (0, 3, 7):
            w = k * nw // nwin
            mn, mx, _ = O.wl_window(k, nwin, maxn, 0)
            o = O.Oracle(sysm, seed=sd[w])
            o.wl_setup(mn, mx, maxn, wlprec=math.e, flat=kw["flatcoeff"], visittol=kw["visittol"],
                       prectol=kw["prectol"])
            o.wl_run(nsteps)
            lng, hist, st = h.wl_get(w)
            olng, ohist, ost = o.wl_get()
            np.testing.assert_array_equal(hist, ohist)
            np.testing.assert_array_equal(lng, olng)
            for g in (1, 2):
                np.testing.assert_allclose(h.guests(g, w=w), o.guests(g), rtol=0, atol=1e-9)
            say(f"  window {k} walker {w}: identical")
            o.close()
    finally:
        h.close()


def test_gpu_team_size_changes_follow_oracle():
    """One chain continued through runs of different team sizes on one handle
    (4, 2, 4, 3, 1, 4): every change clears the exchange records, whose row
    values carry their exchange in two mantissa bits, so no record of an
    earlier team can pass for a current one.  The chain is the oracle's."""
    from fastmc_amd import inputs as I
    from fastmc_amd.lib import FMC_FP64, Handle
    from fastmc_amd.system import read_system
    from test_gpu_config2 import first_difference
    import shutil
    import tempfile
    d = tempfile.mkdtemp(prefix="fmc_teamchg_")
    try:
        I.config0(d, pressure_bar=10.0)
        sysm = read_system(d)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    seeds = (31, 47, 59, 71)
    teams, n = (4, 2, 4, 3, 1, 4), 150
    say(f"config 0 at 10 bar, FP64: {n} steps per run with teams {teams}")
    h = Handle(sysm, precision=FMC_FP64, seeds=[seeds])
    parts = []
    for t in teams:
        h.set_team(t)
        assert h.get_team() == t
        parts.append(h.run(n, trace=True))
    tr = np.concatenate(parts)
    o = O.Oracle(sysm, seed=seeds)
    otr = o.run(n * len(teams))
    diff = first_difference(tr, otr, sysm.temperature)
    assert diff is None, diff
    assert h.stats(0).overflow == 0
    np.testing.assert_allclose(h.guests(1), o.guests(1), rtol=0, atol=1e-9)
    np.testing.assert_allclose(h.sk(), o.sk(), rtol=0, atol=1e-8)
    say(f"  {n * len(teams)} steps identical, N={len(o.guests(1)) // 3}")
    o.close()
    h.close()
