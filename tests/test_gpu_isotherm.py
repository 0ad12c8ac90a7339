"""Isotherm loadings of the device chains against the reference's own: the
config-0 system (424-atom framework, CO2) at T = 293 K and 0.1, 1 and 10 bar,
and at 273 K and 323 K and 1 bar, 64 independent walkers per state point, each
run for the reference CONTROL's 1e5 equilibration and 1e5 production steps.
The reference side (tests/golden/isotherm_ref_cfg0.json, made by
make_isotherm_ref.py from the reference's serial program) holds 16 runs per
state point with each run's <N> (gcmc.f:1458-1557).

Both sides are means of independent chains of the same Markov process, so
their difference is tested against the standard error of the difference of
the two means: |<N>_dev - <N>_ref| <= 3 sqrt(SE_dev^2 + SE_ref^2), with SE the
standard deviation between chains over the square root of their number.  The
mixed-precision run is the north-star check; the FP64 run (whose chains are
the reference's step for step, as the trace replays show) is the control
that separates a precision effect from sampling noise.  z-scores are printed."""
import json
import os
import shutil
import sys
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
NPER = 64


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


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_gpu_isotherm_loadings_match_reference(precision):
    from fastmc_amd import inputs as I
    from fastmc_amd.lib import FMC_FP64, FMC_MIXED, Handle
    from fastmc_amd.system import make_walker, read_system
    ref = json.load(open(os.path.join(HERE, "golden", "isotherm_ref_cfg0.json")))
    d = tempfile.mkdtemp(prefix="fmc_iso_")
    try:
        I.config0(d, pressure_bar=1.0)
        sysm = read_system(d)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    pts = ref["points"]
    seeds = walker_seeds(NPER * len(pts), 0 if precision == "mixed" else 1)
    walkers, keep = [], []
    for ip, pt in enumerate(pts):
        for r in range(NPER):
            walkers.append(make_walker(sysm, seed=seeds[ip * NPER + r], temperature=pt["temperature"],
                                       pressures=[pt["pressure_bar"] * 1e5], keep=keep))
    say(f"{precision}: {len(walkers)} walkers, {ref['equilibration']} + {ref['production']} steps")
    h = Handle(sysm, walkers=walkers, precision=FMC_MIXED if precision == "mixed" else FMC_FP64)
    try:
        h.run(ref["equilibration"])
        h.reset_stats()
        h.run(ref["production"], count_stats=True)
        means = []
        for w in range(len(walkers)):
            st = h.stats(w)
            assert st.counted == ref["production"]
            means.append(st.sum_n[0] / st.counted)
    finally:
        h.close()
    means = np.array(means).reshape(len(pts), NPER)
    bad = []
    for ip, pt in enumerate(pts):
        dev = means[ip].mean()
        se_dev = means[ip].std(ddof=1) / np.sqrt(NPER)
        se_ref = pt["sd_between_runs"] / np.sqrt(len(pt["runs"]))
        z = (dev - pt["mean_n"]) / np.hypot(se_dev, se_ref)
        say(f"{precision} T = {pt['temperature']:g} K, P = {pt['pressure_bar']} bar: device <N> = {dev:.3f} "
            f"+/- {se_dev:.3f} ({NPER} chains); reference {pt['mean_n']:.3f} +/- {se_ref:.3f} "
            f"({len(pt['runs'])} runs); z = {z:+.2f}")
        if abs(z) > 3.0:
            bad.append((pt["temperature"], pt["pressure_bar"], round(float(z), 2)))
    assert not bad, bad
