"""Guests with fewer sites than the three trial sites k_run keeps per pair
(a one-site argon-like guest, and a two-site diatomic, beside CO2 on the
config-0 framework): the padded site slots must contribute no energy and no
overlap or surface flag.  A five-site methane beside CO2 takes the general
trial-site path instead.  The device FP64 chain runs against the oracle from
the same seeds: every move, molecule choice, decision and molecule count
identical, and dU within 1e-10 * max(|dU|, kB T).  No pure rotation moves:
rotating a one-site molecule changes nothing, its dU is rounding noise of
either sign, and the reference draws the Metropolis uniform only when that
noise is >= 0 (gcmc.f:1879-1882), which no other order of summation
reproduces; displacements (a translation and a rotation) are covered."""
import os
import shutil
import sys
import tempfile

import numpy as np
import pytest

from fastmc_amd import inputs as I
from fastmc_amd.system import read_system
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

KB_KCAL = 1.9872041e-3


def argon():
    return I.GuestSpec(name="Ar", sites=[("Arx", 39.948, 0.0, 0.0, 0.0, 0.0)],
                       lj={"Arx": (119.8 * KB_KCAL, 3.40)})


def diatomic():
    return I.GuestSpec(name="X2", sites=[("Xa", 20.0, 0.3, -0.6, 0.0, 0.0), ("Xb", 20.0, -0.3, 0.6, 0.0, 0.0)],
                       lj={"Xa": (60.0 * KB_KCAL, 3.1), "Xb": (45.0 * KB_KCAL, 3.2)})


def methane():
    """Five sites (C and four H at 1.09 A, tetrahedral): more sites than the
    register slots, so k_run takes its general trial-site path."""
    h = 1.09 / 3 ** 0.5
    return I.GuestSpec(name="CH4", sites=[("Cm", 12.0107, -0.24, 0.0, 0.0, 0.0)] +
                       [("Hm", 1.00794, 0.06, sx * h, sy * h, sz * h)
                        for sx, sy, sz in ((1, 1, 1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1))],
                       lj={"Cm": (33.2 * KB_KCAL, 3.31), "Hm": (7.6 * KB_KCAL, 2.81)})


MOVES = {"insert": 0.3, "delete": 0.3, "displace": 0.3, "jump": 0.0, "swap": 0.1, "translate": 0.0,
         "rotate": 0.0, "switch": 0.0}

SYSTEMS = {
    "co2_ar": ([I.co2_guest(), argon()], [2.0, 5.0]),
    "co2_ar_x2": ([I.co2_guest(), argon(), diatomic()], [2.0, 5.0, 3.0]),
    "co2_ch4": ([I.co2_guest(), methane()], [2.0, 5.0]),
}


def system(name):
    guests, pressures = SYSTEMS[name]
    d = tempfile.mkdtemp(prefix=f"fmc_ragged_{name}_")
    try:
        I.write_system(d, I.make_framework(), guests, 293.0, pressures, moves=MOVES)
        return read_system(d)
    finally:
        shutil.rmtree(d, ignore_errors=True)


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_gpu_padded_sites_match_oracle(name):
    from fastmc_amd.lib import FMC_FP64, Handle
    print(name, flush=True)
    sys.stdout.flush()
    sysm = system(name)
    seeds = (21, 43, 65, 87)
    n = 3000
    h = Handle(sysm, precision=FMC_FP64, seeds=[seeds])
    try:
        td = h.run(n, trace=True)
    finally:
        h.close()
    o = Oracle(sysm, seed=seeds)
    to = o.run(n)
    o.close()
    for f in ("iguest", "move", "choice", "nmols", "overlap", "accepted"):
        assert np.array_equal(td[f], to[f]), (name, f, int(np.argmax(td[f] != to[f])))
    ok = to["overlap"] == 0
    scale = np.maximum(np.abs(to["estep"][ok]), KB_KCAL * 293.0)
    err = np.abs(td["estep"][ok] - to["estep"][ok]) / scale
    kinds = np.unique(td["iguest"][td["accepted"] == 1])
    print(name, "accepted moves per guest kind:", {int(k): int(((td["iguest"] == k) & (td["accepted"] == 1)).sum())
                                                  for k in kinds}, "max scaled dU error", float(err.max()), flush=True)
    assert len(kinds) == len(SYSTEMS[name][0])   # every guest kind moved
    assert err.max() <= 1e-10


NOSWAP = dict(MOVES, insert=1 / 3, delete=1 / 3, displace=1 / 3, swap=0.0)


@pytest.mark.parametrize("name", ["co2_ar", "co2_ch4"])
def test_gpu_mixed_precision_du_error_ragged(name):
    """Mixed mode on the same guests (the general trial-site path and the
    FP32 phase rows included): the FP64 device chain's trials replayed
    through the one-trial seam on a mixed and an FP64 handle, |dU_mixed -
    dU_fp64| <= 1e-6 * max(|dU_fp64|, kB T) on every trial without overlap."""
    from fastmc_amd.lib import FMC_FP64, FMC_MIXED, Handle
    print(name, "mixed", flush=True)
    guests, pressures = SYSTEMS[name]
    d = tempfile.mkdtemp(prefix=f"fmc_ragged_mixed_{name}_")
    try:
        I.write_system(d, I.make_framework(), guests, 293.0, pressures, moves=NOSWAP)
        sysm = read_system(d)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    seeds = (21, 43, 65, 87)
    natoms = [len(g.sites) for g in guests]
    hd = Handle(sysm, precision=FMC_FP64, seeds=[seeds])
    try:
        tr = hd.run(1200, trace=True)
    finally:
        hd.close()
    a_h = Handle(sysm, precision=FMC_FP64, seeds=[seeds])
    b_h = Handle(sysm, precision=FMC_MIXED, seeds=[seeds])
    errs = []
    try:
        for r in tr:
            g, mv, c = int(r["iguest"]), int(r["move"]), int(r["choice"])
            xyz = r["xyz"][:3 * natoms[g - 1]].reshape(natoms[g - 1], 3)
            if mv == 1:
                a, b = a_h.trial_insert(g, xyz), b_h.trial_insert(g, xyz)
            elif mv == 2:
                a, b = a_h.trial_delete(g, c), b_h.trial_delete(g, c)
            else:
                a, b = a_h.trial_displace(g, c, xyz), b_h.trial_displace(g, c, xyz)
            assert a.overlap == b.overlap
            if not a.overlap:
                errs.append(abs(a.estep - b.estep) / max(abs(a.estep), KB_KCAL * 293.0))
            for h in (a_h, b_h):
                h.accept() if r["accepted"] else h.reject()
    finally:
        a_h.close()
        b_h.close()
    errs = np.array(errs)
    print(name, "mixed dU scaled error: max", errs.max(), "n", len(errs), flush=True)
    assert len(errs) > 500 and errs.max() <= 1e-6
