"""Widom grid on the device (fmc_widom) against the reference's own output
and the C oracle: Henry's constant of every guest to the reference's printed
digits, the non-overlapping attempt count, and every attempt's estep and
overlap flag against the oracle (FP64 within the chains' bound; mixed
precision, FP32 high-k phase products only, within the mixed bound), on
config 0 and the CALF-20 CO2/N2 system
(triclinic cell, two guests; guest 2 continues the walker's duni sequence).
Also: a traced chain run after an energy-gradient call leaves the gradient
workspace intact."""
import json
import math
import os
import sys

import numpy as np
import pytest

from _helpers import GOLDEN
from test_widom import printed_tolerance, widom_system
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

REF = json.load(open(os.path.join(GOLDEN, "widom_ref.json")))
KB = 1.9872041e-3


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("case", REF, ids=[c["name"] for c in REF])
def test_gpu_widom_matches_reference_and_oracle(case):
    from fastmc_amd.lib import FMC_FP64, FMC_MIXED, Handle
    sysm = widom_system(case["name"])
    seeds = tuple(case["seeds"])
    T = sysm.temperature
    for prec in (FMC_FP64, FMC_MIXED):
        h = Handle(sysm, precision=prec, seeds=[seeds])
        o = Oracle(sysm, seed=seeds)
        accepted = 0
        for g, henry in enumerate(case["henry"], start=1):
            say(case["name"], "precision", prec, "guest", g)
            r = h.widom(g, case["insertions_per_point"], case["grid_spacing"], per_attempt=True)
            ro = o.widom(g, case["insertions_per_point"], case["grid_spacing"], max_attempts=r["attempts"])
            assert r["attempts"] == ro["attempts"] and tuple(r["ngrid"]) == tuple(case["grid"])
            np.testing.assert_array_equal(r["overlap"], ro["overlap"])
            ok = ro["overlap"] == 0
            du = np.abs(r["estep"][ok] - ro["estep"][ok])
            scale = np.maximum(np.abs(ro["estep"][ok]), KB * T)
            rel = du / scale
            worst = int(np.argmax(rel))
            say("  max |dU| / max(|dU|, kT):", float(rel[worst]), "at estep", float(ro["estep"][ok][worst]))
            assert (rel <= (1e-10 if prec == FMC_FP64 else 1e-6)).all()
            accepted += r["accepted"]
            if prec == FMC_FP64:
                assert abs(r["henry"] - henry) <= printed_tolerance(henry), (r["henry"], henry)
            else:
                assert abs(r["henry"] - ro["henry"]) <= 1e-4 * abs(ro["henry"])
        assert all(c == accepted for c in case["counted"])
        h.close()
        o.close()


def test_gpu_traced_run_keeps_gradient_workspace():
    from fastmc_amd import adjoint as A
    from fastmc_amd.lib import FMC_FP64, Handle
    sysm = widom_system("cfg0_p1")
    nmols, xyz = A.sample_configurations(sysm, 2, every=200, warmup=2000)
    h = Handle(sysm, precision=FMC_FP64, seeds=[(3, 5, 7, 9)])
    q, eps, sig = A.theta0(sysm)
    first = h.energy_grad(q, eps, sig, nmols, xyz, FMC_FP64)
    h.run(10, trace=True)
    h.run(50, trace=True)   # grows the trace buffer
    again = h.energy_grad(q, eps, sig, nmols, xyz, FMC_FP64)
    # equal up to the order of the kernels' FP64 atomic additions
    for a, b in zip(first, again):
        np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-15)
    h.close()


@pytest.mark.parametrize("case", REF, ids=[c["name"] for c in REF])
def test_gpu_run_directory_widom_mode(case, tmp_path):
    """CONTROL with the reference's 'henry' directives through the run driver:
    the OUTPUT it writes carries each guest's Henry's constant as the
    reference printed it (final report, E15.6) and its counted steps."""
    import re
    import shutil
    from fastmc_amd import gcmc
    from fastmc_amd import inputs as I
    from _helpers import GOLDEN
    d = str(tmp_path)
    if case["name"] == "calf20_co2n2":
        for f in ("FIELD", "CONFIG"):
            shutil.copy(os.path.join(GOLDEN, "calf20", f), os.path.join(d, f))
        ctl = json.load(open(os.path.join(GOLDEN, "moves_ref.json")))[0]["control"]   # CALF-20 at 293 K
    else:
        I.config0(d, pressure_bar=1.0)
        ctl = open(os.path.join(d, "CONTROL")).read()
    ctl = ctl.replace("finish", f"henry\nwidom insertions {case['insertions_per_point']}\n"
                                f"grid spacing {case['grid_spacing']}\nfinish")
    open(os.path.join(d, "CONTROL"), "w").write(ctl)
    res = gcmc.run_directory(d, seed=tuple(case["seeds"]))
    out = open(os.path.join(d, "OUTPUT")).read()
    final = out[out.index("final stats"):]
    printed = [float(x) for x in re.findall(r"Henry's Constant \(mol /kg /bar\):\s+([-0-9.E+]+)", final)]
    counted = [int(x) for x in re.findall(r"Total steps counted:\s+(\d+)", final)]
    assert len(printed) == len(case["henry"])
    for ours, ref in zip(printed, case["henry"]):
        assert abs(ours - ref) <= 0.5 * 10.0 ** (math.floor(math.log10(abs(ref))) - 5) * 1.001, (ours, ref)
    assert all(c == case["counted"][0] for c in counted)
    assert [r["henry"] for r in res] == pytest.approx(case["henry"], rel=1e-8)
