"""The reference's own isotherm test (run_test.py, 'long'), run on the device
chain: the CALF-20 CO2/N2 system (its FIELD / CONFIG, tests/golden/calf20) at
its ten state points (test/T293.0P*/CONTROL: 750000 + 500000 steps, insert /
delete / displace / swap / switch / jump at 1/6 each), four independent FP64
walkers per point in one launch.  The reference passes a point when its
expected average numbers of CO2 and N2 (run_test.py:27-38) lie in
[(N - s) * 0.8, (N + s) * 1.2] with N the run's average and s its standard
error (run_test.py:88-113); here N is the mean over the four walkers and s
the standard error of that mean."""
import json
import os
import shutil
import sys
import tempfile

import numpy as np
import pytest

from _helpers import GOLDEN

pytestmark = pytest.mark.gpu

REF = json.load(open(os.path.join(GOLDEN, "calf20_isotherm_ref.json")))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_gpu_reference_isotherm_test_passes():
    from fastmc_amd.lib import FMC_FP64, Handle
    from fastmc_amd.system import make_walker, read_system
    pts = REF["points"]
    d = tempfile.mkdtemp(prefix="fmc_calf20_iso_")
    try:
        for f in ("FIELD", "CONFIG"):
            shutil.copy(os.path.join(GOLDEN, "calf20", f), os.path.join(d, f))
        with open(os.path.join(d, "CONTROL"), "w") as fh:
            fh.write(pts[0]["control"])
        sysm = read_system(d)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    nrep = 4
    walkers, keep = [], []
    for ip, pt in enumerate(pts):
        for r in range(nrep):
            k = ip * nrep + r
            walkers.append(make_walker(sysm, seed=(1 + k, 3 + 2 * k, 7 + 3 * k, (11 * k) % 168),
                                       pressures=[p * 1e5 for p in pt["pressures_bar"]], keep=keep))
    eq, prod = pts[0]["equilibration"], pts[0]["steps"]
    say(len(walkers), "walkers,", eq, "+", prod, "steps")
    h = Handle(sysm, walkers=walkers, precision=FMC_FP64)
    try:
        h.run(eq)
        h.reset_stats()
        h.run(prod, count_stats=True)
        means = np.array([[h.stats(w).sum_n[g] / h.stats(w).counted for g in range(2)] for w in range(len(walkers))])
    finally:
        h.close()
    means = means.reshape(len(pts), nrep, 2)
    failed = []
    for ip, pt in enumerate(pts):
        n = means[ip].mean(0)
        s = means[ip].std(0, ddof=1) / np.sqrt(nrep)
        ref = pt["expected_co2_n2"]
        ok = all((n[g] - s[g]) * 0.8 <= ref[g] <= (n[g] + s[g]) * 1.2 for g in range(2))
        say(f"P = {pt['total_bar']:.6f} bar: CO2 {n[0]:.2f} +/- {s[0]:.2f} (ref {ref[0]}), "
            f"N2 {n[1]:.3f} +/- {s[1]:.3f} (ref {ref[1]})  {'PASS' if ok else 'FAIL'}")
        if not ok:
            failed.append(pt["total_bar"])
    assert not failed, failed
