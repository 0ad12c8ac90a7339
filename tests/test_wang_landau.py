"""Wang-Landau in N (wang_landau.f): the oracle's walker, window split,
stitching and isotherm against closed forms.

The reference's Wang-Landau build does not compile (argument mismatch), so
the oracle restates wang_landau.f and is pinned here by the ideal gas: with
no charges, no LJ and no overlap test every trial has dU = 0, and the density
of states is exactly ln g(N) - ln g(0) = N ln(V / Lambda^3) - ln N!; its
grand-canonical isotherm is the Poisson mean."""
import math
import os
import shutil
import tempfile

import numpy as np

from fastmc_amd import inputs as I
from fastmc_amd.system import read_system
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def ideal_gas_system(pressure_bar=1.0):
    d = tempfile.mkdtemp(prefix="fmc_ig_")
    try:
        I.config0(d, pressure_bar=pressure_bar)
        s = read_system(d)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    for sp in s.species:
        sp.charge = np.zeros_like(sp.charge)
    s.pairs = [(i, j, k, 0.0, sg) for (i, j, k, e, sg) in s.pairs]
    s.overlap = 0.0
    s.sumchg = 0.0
    return s


def exact_lng(V, lambda3, maxn):
    n = np.arange(maxn + 1)
    return n * math.log(V / lambda3) - np.array([math.lgamma(k + 1) for k in n])


def test_wl_thermal_wavelength_closed_form():
    """Lambda = h / sqrt(2 pi m kB T) with m the guest's molar mass (wang_landau.f:121-123)."""
    s = ideal_gas_system()
    o = O.Oracle(s, seed=(12, 34, 56, 78))
    o.wl_setup(0, 4, 4)
    _, _, st = o.wl_get()
    m = sum(s.guests[0].mass) / 1000.0 / 6.02214076e23          # kg per molecule (CODATA 2018)
    lam = 6.62607015e-34 / math.sqrt(2 * math.pi * m * 1.380649e-23 * s.temperature) * 1e10
    # the reference's constants are older (kB 1.3806503e-23, h 6.62607004e-34): 1.4e-6 apart
    assert abs(st["lambda3"] / lam ** 3 - 1.0) < 5e-6
    o.close()


def test_wl_oracle_ideal_gas_converges_to_exact_dos():
    s = ideal_gas_system()
    maxn = 12
    o = O.Oracle(s, seed=(12, 34, 56, 78))
    o.wl_setup(0, maxn, maxn, wlprec=math.e, flat=0.8, visittol=10 ** 9, prectol=1e-4)
    n = 0
    while True:
        tr = o.wl_run(20000)
        n += 20000
        lng, hist, st = o.wl_get()
        if st["done"] or n > 600000:
            break
    assert st["done"] and st["sweeps"] == 14           # ln f: 1, 1/2, ... < 1e-4
    assert abs(st["lnf"] - math.log(math.e ** (0.5 ** 14))) < 1e-15
    # every trial of the ideal gas has dU = 0
    assert np.all(tr["estep"][tr["move"] > 0] == 0.0)
    err = (lng - lng[0]) - exact_lng(s.volume(), st["lambda3"], maxn)
    assert np.abs(err).max() < 0.5, err
    o.close()


def test_wl_oracle_window_edges_and_preload():
    """A window [3, 7]: insert_guests first loads 3 molecules; then N stays inside,
    and only in-window bins are visited."""
    s = ideal_gas_system()
    o = O.Oracle(s, seed=(3, 5, 7, 11))
    o.wl_setup(3, 7, 10, wlprec=math.e, flat=0.8, visittol=10 ** 9, prectol=1e-2)
    tr = o.wl_run(4000)
    tr = tr[tr["step"] > 0]          # a walker stops once ln f < prectol
    lng, hist, st = o.wl_get()
    assert st["done"] and st["sweeps"] == 7
    pre = np.nonzero(tr["nmols"] >= 3)[0][0]
    assert pre == 3 and np.all(tr["accepted"][:3] == 1)            # three preload insertions
    n_after = tr["nmols"][pre:] + np.where(tr["accepted"][pre:] == 1,
                                           np.where(tr["move"][pre:] == 1, 1,
                                                    np.where(tr["move"][pre:] == 2, -1, 0)), 0)
    assert n_after.min() >= 3 and n_after.max() <= 7
    assert hist[:3].sum() == 0 and hist[8:].sum() == 0
    assert st["minmol"] == 3 and st["maxmol"] == 7
    o.close()


def test_wl_windows_divide_jobs():
    """divide_jobs for 8 ranks over N = 0..400 (tail 3): contiguous, overlapping windows."""
    w = [O.wl_window(i, 8, 400, 0) for i in range(8)]
    assert w[0][:2] == (0, 153) and w[-1][:2] == (384, 400)
    for a, b in zip(w, w[1:]):
        assert b[0] < a[1]                       # windows overlap by the tails
    assert O.wl_window(0, 1, 40, 0) == (0, 40, 0)


def test_wl_stitch_recovers_exact_dos():
    """Windows cut from one exact ln g, each with its own additive constant,
    stitch back to ln g - ln g(0)."""
    maxn = 60
    g = exact_lng(17173.512, 0.0036, maxn)
    rng = np.random.default_rng(5)
    for nwin in (1, 2, 4):
        wins = []
        for k in range(nwin):
            mn, mx, _ = O.wl_window(k, nwin, maxn, 0)
            a = np.zeros(maxn + 1)
            a[mn: mx + 1] = g[mn: mx + 1] + rng.uniform(-50, 50)
            wins.append(a)
        st = O.wl_stitch(wins, maxn, 0)
        ref = g - g[0]
        if nwin == 1:
            # stitch_branches shifts only varchunk = maxn - minn bins of window 0
            np.testing.assert_allclose(st[:-1], ref[:-1], atol=1e-9)
        else:
            np.testing.assert_allclose(st, ref, atol=1e-9)


def test_wl_isotherm_ideal_gas_poisson_mean():
    """<N> from the exact ideal-gas ln g: the reference's chemical potential
    (beta mu = ln(p Lambda^3 beta)) gives the Poisson mean p beta V; the SI
    form gives p V / (kB T)."""
    V, l3, T = 17173.512, 0.0036339, 293.0
    maxn = 80
    g = exact_lng(V, l3, maxn)
    beta = 1.0 / (O.KBOLTZ * T)
    p0 = np.array([1e-4, 5e-4, 1e-3])   # means 3 .. 30, far below maxn
    n0 = O.wl_isotherm(g, maxn, 0, T, l3, p0, units=0)
    np.testing.assert_allclose(n0, p0 * beta * V, rtol=1e-9)
    p1 = np.array([1e5, 5e5, 2e6])
    n1 = O.wl_isotherm(g, maxn, 0, T, l3, p1, units=1)
    np.testing.assert_allclose(n1, p1 * V * 1e-30 / (O.BOLTZ * T), rtol=1e-9)
    p = O.reference_pressures()
    assert len(p) == 51 and abs(p[0] - 0.1) < 1e-12 and abs(p[-1] - 500000.1) < 1e-6
