"""Isotherm loadings of the device chains against the reference's own runs:
the config-0 system (424-atom framework, CO2) at T = 293 K and 0.1, 1 and
10 bar, and at 273 K and 323 K and 1 bar, with the reference CONTROL's 1e5
equilibration and 1e5 production steps.  The reference side
(tests/golden/isotherm_ref_cfg0.json, made by make_isotherm_ref.py from the
reference's serial program) holds 16 runs per state point: each run's duni
seeds and its printed <N> +/- block-average error (gcmc.f:1458-1557).

* FP64: 80 walkers, one per reference run, started from that run's printed
  seeds and driven through the reference's main-loop schedule, print every
  run's <N> to the reference's 6 decimals -- the whole 2e5-step chains are
  the reference's.
* Mixed precision: 128 walkers per state point with random seeds, each run
  in mixed precision and in FP64 from the same seeds.  The FP64 chains are the
  reference's chains (above), so the paired difference mixed - FP64 isolates
  what the precision does: |mean difference| <= 3 standard errors of the
  paired differences (most mixed chains never take a different decision
  from their FP64 chain, and then the difference is exactly zero).  The
  loading criterion for mixed precision, <N>_mixed within one reference block-average error
  of the reference mean, is checked too; z against the reference mean is
  printed for both precisions.  The reference mean of 16 runs is itself a
  sample: at 10 bar it lies 2.2 molecules (about 3.5 of its standard errors)
  above the mean of 512 FP64 chains, seeded at random or by the reference's
  own clock rule alike (190.8 +/- 0.15)."""
import json
import os
import shutil
import sys
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
NPER = 128


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def walker_seeds(n, salt):
    """Distinct duni seeds (i, j, k in 1..178, l in 0..168, utility_pack.f:2801-2804)
    drawn at random, so that the chains share no structure."""
    rng = np.random.default_rng(20261204 + salt)
    seen, out = set(), []
    while len(out) < n:
        s = (int(rng.integers(1, 179)), int(rng.integers(1, 179)), int(rng.integers(1, 179)),
             int(rng.integers(0, 169)))
        if s not in seen:
            seen.add(s)
            out.append(s)
    return out


def reference():
    return json.load(open(os.path.join(HERE, "golden", "isotherm_ref_cfg0.json")))


def config0_system():
    from fastmc_amd import inputs as I
    from fastmc_amd.system import read_system
    d = tempfile.mkdtemp(prefix="fmc_iso_")
    try:
        I.config0(d, pressure_bar=1.0)
        return read_system(d)
    finally:
        shutil.rmtree(d, ignore_errors=True)


class Engine:
    """All walkers of a handle through the reference's main-loop schedule."""

    def __init__(self, h):
        self.h = h

    def run(self, n, flags):
        self.h.run(n, count_stats=flags)

    def prodcount(self):
        return int(self.h.stats(0).prodcount)


def test_gpu_isotherm_fp64_reproduces_reference_runs():
    from fastmc_amd import gcmc
    from fastmc_amd.lib import FMC_FP64, Handle
    from fastmc_amd.system import make_walker
    ref = reference()
    sysm = config0_system()
    walkers, keep, want = [], [], []
    for pt in ref["points"]:
        for run in pt["runs"]:
            walkers.append(make_walker(sysm, seed=tuple(run["seeds"]), temperature=pt["temperature"],
                                       pressures=[pt["pressure_bar"] * 1e5], keep=keep))
            want.append((pt["temperature"], pt["pressure_bar"], run["mean_n"]))
    eq, prod = ref["equilibration"], ref["production"]
    say(f"fp64: {len(walkers)} walkers from the reference runs' seeds, {eq} + {prod} steps")
    nwind, nwindsteps = gcmc.window_steps(prod)
    h = Handle(sysm, walkers=walkers, precision=FMC_FP64)
    try:
        h.set_sampling(nwindsteps, False, 16)
        eng = Engine(h)
        gcmc.run_main_loop(eng, eq, prod, gcmc.check_interval(eq, 1000), False, eng.prodcount, nwindsteps)
        got = [h.production_stats(w)[0]["avg_final"]["<N>"] for w in range(len(walkers))]
        for w in range(len(walkers)):
            assert h.stats(w).overflow == 0
    finally:
        h.close()
    bad = [(t, p, n, round(g, 6)) for (t, p, n), g in zip(want, got) if abs(g - n) > 0.5e-6 * 1.0001]
    say(f"fp64: {len(want) - len(bad)} of {len(want)} reference runs reproduced to the printed digits")
    assert not bad, bad[:5]


def run_walkers(sysm, pts, seeds, precision, eq, prod):
    from fastmc_amd.lib import Handle
    from fastmc_amd.system import make_walker
    walkers, keep = [], []
    for ip, pt in enumerate(pts):
        for r in range(NPER):
            walkers.append(make_walker(sysm, seed=seeds[ip * NPER + r], temperature=pt["temperature"],
                                       pressures=[pt["pressure_bar"] * 1e5], keep=keep))
    h = Handle(sysm, walkers=walkers, precision=precision)
    try:
        h.run(eq)
        h.reset_stats()
        h.run(prod, count_stats=True)
        means = []
        for w in range(len(walkers)):
            st = h.stats(w)
            assert st.counted == prod and st.overflow == 0
            means.append(st.sum_n[0] / st.counted)
    finally:
        h.close()
    return np.array(means).reshape(len(pts), NPER)


def test_gpu_isotherm_mixed_matches_fp64_and_reference():
    from fastmc_amd.lib import FMC_FP64, FMC_MIXED
    ref = reference()
    sysm = config0_system()
    pts = ref["points"]
    eq, prod = ref["equilibration"], ref["production"]
    seeds = walker_seeds(NPER * len(pts), 0)
    say(f"{len(seeds)} walkers x (mixed, fp64), {eq} + {prod} steps")
    mixed = run_walkers(sysm, pts, seeds, FMC_MIXED, eq, prod)
    fp64 = run_walkers(sysm, pts, seeds, FMC_FP64, eq, prod)
    bad = []
    for ip, pt in enumerate(pts):
        d = mixed[ip] - fp64[ip]
        ndiff = int(np.count_nonzero(d))  # chains whose decisions ever differed
        se_d = d.std(ddof=1) / np.sqrt(NPER)
        z_pair = d.mean() / se_d if ndiff > 1 else 0.0
        se_ref = pt["sd_between_runs"] / np.sqrt(len(pt["runs"]))
        line = f"T = {pt['temperature']:g} K, P = {pt['pressure_bar']} bar: "
        for name, m in (("mixed", mixed[ip]), ("fp64", fp64[ip])):
            se = m.std(ddof=1) / np.sqrt(NPER)
            line += (f"{name} {m.mean():.3f} +/- {se:.3f} (z vs reference "
                     f"{(m.mean() - pt['mean_n']) / np.hypot(se, se_ref):+.2f}); ")
        say(line + f"reference {pt['mean_n']:.3f} +/- {se_ref:.3f}; mixed - fp64 = {d.mean():+.3f} +/- {se_d:.3f} "
            f"(z {z_pair:+.2f}; {NPER - ndiff} of {NPER} mixed chains identical to their FP64 chain)")
        if abs(z_pair) > 3.0:
            bad.append(("paired", pt["temperature"], pt["pressure_bar"], round(float(z_pair), 2)))
        if abs(mixed[ip].mean() - pt["mean_n"]) > pt["mean_stderr_n"]:
            bad.append(("block error", pt["temperature"], pt["pressure_bar"]))
    assert not bad, bad
