"""GPU Wang-Landau walkers (wang_landau.f) against the oracle's restatement:
the same seeds give the same chain step by step (moves, molecule choices,
overlaps, decisions, dU within the FP64 bound), the same visit histogram and
bit-identical ln g(N); windows, preloading and a two-guest system included.
The merged / stitched result, the isotherm and the window split through the
C-ABI equal the oracle's numpy restatements, and the ideal gas converges to
its exact density of states."""
import math
import sys

import numpy as np
import pytest

from _helpers import KB_KCAL, case_system, du_tol, fixture
from oracle import oracle as O
from test_wang_landau import exact_lng, ideal_gas_system

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _lib():
    from fastmc_amd import lib
    return lib


@pytest.mark.parametrize("env", [None, "FMC_NO_LDS_ALL"])
def test_gpu_wl_chain_matches_oracle(env, monkeypatch):
    lib = _lib()
    if env:
        monkeypatch.setenv(env, "1")
    name = "cfg0_p10"
    sysm = case_system(name)
    seed = tuple(fixture(name)["seeds"])
    say("wl chain", name, env)
    kw = dict(visittol=40, flatcoeff=0.6, prectol=1e-3)
    h = lib.Handle(sysm, precision=lib.FMC_FP64, seeds=[seed])
    h.wl_setup(maxn=8, nwindows=1, **kw)
    tr = h.run(6000, trace=True)
    o = O.Oracle(sysm, seed=seed)
    o.wl_setup(0, 8, 8, wlprec=math.e, flat=kw["flatcoeff"], visittol=kw["visittol"], prectol=kw["prectol"])
    otr = o.wl_run(6000)
    n = int((otr["step"] > 0).sum())
    assert n > 1000
    T = sysm.temperature
    for i in range(n):
        a, b = tr[i], otr[i]
        for f in ("move", "choice", "nmols", "overlap", "accepted"):
            assert a[f] == b[f], (i, f, a[f], b[f])
        if b["move"] and not b["overlap"]:
            assert abs(a["estep"] - b["estep"]) <= du_tol(b["estep"], T), (i, a["estep"], b["estep"])
    lng, hist, st = h.wl_get(0)
    olng, ohist, ost = o.wl_get()
    np.testing.assert_array_equal(hist, ohist)
    np.testing.assert_array_equal(lng, olng)
    assert (st.sweeps, st.done, st.minmol, st.maxmol) == (ost["sweeps"], ost["done"], ost["minmol"], ost["maxmol"])
    assert st.f == ost["f"] and st.lnf == ost["lnf"] and st.lambda3 == ost["lambda3"]
    say("wl chain", n, "steps, sweeps", st.sweeps, "ln f", st.lnf)
    h.close()
    o.close()


def test_gpu_wl_windows_preload_two_guests():
    """Four walkers over two windows of N(CO2) in the CO2/N2 system: window 1
    first loads to its lower edge; N2 keeps grand-canonical moves."""
    lib = _lib()
    name = "calf20_co2n2"
    sysm = case_system(name)
    seeds = [(11, 22, 33, 44), (55, 66, 77, 88), (99, 100, 101, 102), (150, 151, 152, 153)]
    kw = dict(visittol=30, flatcoeff=0.6, prectol=1e-4)
    say("wl windows", name)
    h = lib.Handle(sysm, precision=lib.FMC_FP64, seeds=seeds, n=4)
    h.wl_setup(maxn=10, nwindows=2, **kw)
    h.run(3000)
    for w in range(4):
        mn, mx, _ = O.wl_window(w * 2 // 4, 2, 10, 0)
        o = O.Oracle(sysm, seed=seeds[w])
        o.wl_setup(mn, mx, 10, wlprec=math.e, flat=kw["flatcoeff"], visittol=kw["visittol"],
                   prectol=kw["prectol"])
        o.wl_run(3000)
        lng, hist, st = h.wl_get(w)
        olng, ohist, ost = o.wl_get()
        assert (st.minmol, st.maxmol) == (mn, mx)
        np.testing.assert_array_equal(hist, ohist)
        np.testing.assert_array_equal(lng, olng)
        assert st.sweeps == ost["sweeps"] and st.done == ost["done"]
        for g in (1, 2):
            np.testing.assert_allclose(h.guests(g, w=w), o.guests(g), rtol=0, atol=1e-9)
        say("walker", w, "window", (mn, mx), "sweeps", st.sweeps, "N", len(h.guests(1, w=w)) // 3)
        o.close()
    h.close()


def test_gpu_wl_ideal_gas_converges_and_merges():
    lib = _lib()
    s = ideal_gas_system()
    maxn, nwin, nw = 24, 2, 8
    say("wl ideal gas", nw, "walkers", nwin, "windows")
    h = lib.Handle(s, precision=lib.FMC_FP64, seeds=[(7 + i, 13 + i, 29 + i, 31 + i) for i in range(nw)], n=nw)
    h.wl_setup(maxn=maxn, nwindows=nwin, visittol=10 ** 9, flatcoeff=0.8, prectol=1e-3)
    for _ in range(40):
        h.run(20000)
        if all(h.wl_get(w)[2].done for w in range(nw)):
            break
    sts = [h.wl_get(w) for w in range(nw)]
    assert all(st.done and st.error == 0 for _, _, st in sts)
    res = h.wl_result()
    # merge + stitch restated: per window, walkers aligned at the first bin and averaged
    wins = np.zeros((nwin, maxn + 1))
    for k in range(nwin):
        mn, mx, _ = O.wl_window(k, nwin, maxn, 0)
        ws = [w for w in range(nw) if w * nwin // nw == k]
        for w in ws:
            g = sts[w][0]
            wins[k, mn: mx + 1] += (g[mn: mx + 1] - g[mn]) / len(ws)
    np.testing.assert_allclose(res, O.wl_stitch(wins, maxn, 0), rtol=0, atol=1e-9)
    l3 = sts[0][2].lambda3
    err = res - exact_lng(s.volume(), l3, maxn)
    say("ideal gas: max |ln g - exact|", np.abs(err).max())
    assert np.abs(err).max() < 0.6
    p = O.reference_pressures()[:5] * 1e-4
    np.testing.assert_allclose(lib.wl_isotherm(res, maxn, 0, s.temperature, l3, p, 0),
                               O.wl_isotherm(res, maxn, 0, s.temperature, l3, p, 0), rtol=1e-12)
    h.close()


def test_gpu_wl_library_host_functions():
    lib = _lib()
    for nwin in (1, 2, 3, 8):
        for maxn in (20, 100, 400):
            for k in range(nwin):
                assert lib.wl_window(k, nwin, maxn, 0) == O.wl_window(k, nwin, maxn, 0)
    rng = np.random.default_rng(9)
    for nwin in (1, 2, 5):
        wins = rng.normal(0, 30, size=(nwin, 121))
        np.testing.assert_allclose(lib.wl_stitch(wins, 120, 0), O.wl_stitch(wins, 120, 0), rtol=0, atol=1e-12)
    g = np.cumsum(rng.normal(2.0, 1.0, size=61))
    p = rng.uniform(1e-3, 1e6, size=7)
    for units in (0, 1):
        np.testing.assert_allclose(lib.wl_isotherm(g, 60, 0, 293.0, 0.0036, p, units),
                                   O.wl_isotherm(g, 60, 0, 293.0, 0.0036, p, units), rtol=1e-12)
