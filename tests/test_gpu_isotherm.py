"""Isotherm loadings of the mixed-precision device chains against the
reference's own: the config-0 system (424-atom framework, CO2, T = 293 K) at
0.1, 1 and 10 bar, 64 independent walkers per pressure, each run for the
reference CONTROL's 1e5 equilibration and 1e5 production steps.  The
reference side (tests/golden/isotherm_ref_cfg0.json, made by
make_isotherm_ref.py from the reference's serial program) holds 16 runs per
pressure with the reference's own <N> +/- block-average standard error.  The
device loading (mean over walkers of each walker's production average) must
lie within one reference standard error of the reference mean."""
import json
import os
import shutil
import sys
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_gpu_isotherm_loadings_within_reference_error():
    from fastmc_amd import inputs as I
    from fastmc_amd.lib import FMC_MIXED, Handle
    from fastmc_amd.system import make_walker, read_system
    ref = json.load(open(os.path.join(HERE, "golden", "isotherm_ref_cfg0.json")))
    d = tempfile.mkdtemp(prefix="fmc_iso_")
    try:
        I.config0(d, pressure_bar=1.0)
        sysm = read_system(d)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    nper = 64
    walkers, keep = [], []
    for ip, pt in enumerate(ref["points"]):
        for r in range(nper):
            k = ip * nper + r
            seed = (1 + k % 100, 2 + k // 100, 3 + (k % 7) * 20, (5 * k) % 168)  # distinct for k < 10000
            walkers.append(make_walker(sysm, seed=seed, pressures=[pt["pressure_bar"] * 1e5], keep=keep))
    say(f"{len(walkers)} walkers, {ref['equilibration']} + {ref['production']} steps")
    h = Handle(sysm, walkers=walkers, precision=FMC_MIXED)
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
    means = np.array(means).reshape(len(ref["points"]), nper)
    for ip, pt in enumerate(ref["points"]):
        dev = means[ip].mean()
        dev_err = means[ip].std(ddof=1) / np.sqrt(nper)
        say(f"P = {pt['pressure_bar']} bar: device <N> = {dev:.3f} +/- {dev_err:.3f} ({nper} walkers); "
            f"reference {pt['mean_n']:.3f} ({len(pt['runs'])} runs, sd {pt['sd_between_runs']:.3f}), "
            f"reference block-average error {pt['mean_stderr_n']:.3f}")
        assert abs(dev - pt["mean_n"]) <= pt["mean_stderr_n"], pt["pressure_bar"]
