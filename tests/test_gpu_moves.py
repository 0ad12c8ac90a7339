"""jump, swap and switch on the device chain: the reference's CALF-20 CO2/N2
test system with its test move set (insert, delete, displace, swap, switch,
jump at 1/6 each), from the seeds the reference printed.  FP64: every step's
guest, move, molecule choice, N, overlap, decision and draw equal the
oracle's (whose own decisions reproduce the reference's), estep within the
FP64 bound, and the walker's move counters give the reference's printed
acceptance ratios and accepted total.  Mixed precision: the same decisions
over the first 2000 steps."""
import json
import os
import sys

import numpy as np
import pytest

from _helpers import GOLDEN, du_tol
from test_moves import NAMES, REF, moves_system
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

FIELDS = ("iguest", "move", "choice", "nmols", "overlap", "accepted", "drew_rande")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def first_difference(tr, otr, n, temp):
    for i in range(n):
        a, b = tr[i], otr[i]
        for f in FIELDS:
            if a[f] != b[f]:
                return i, f, a[f], b[f]
        if not b["overlap"] and abs(a["estep"] - b["estep"]) > du_tol(b["estep"], temp):
            return i, "estep", a["estep"], b["estep"]
    return None


@pytest.mark.parametrize("case", REF, ids=[c["name"] for c in REF])
def test_gpu_moves_match_oracle_and_reference(case):
    from fastmc_amd.lib import FMC_FP64, Handle
    sysm = moves_system(case)
    seeds = tuple(case["seeds"])
    n = case["steps"]
    say(case["name"], n, "steps")
    h = Handle(sysm, precision=FMC_FP64, seeds=[seeds])
    tr = h.run(n, trace=True)
    st = h.stats(0)
    h.close()
    o = Oracle(sysm, seed=seeds)
    otr = o.run(n)
    o.close()
    d = first_difference(tr, otr, n, sysm.temperature)
    assert d is None, d
    moves = [int(x) for x in tr["move"][:n]]
    say("  moves:", {NAMES.get(k, k): moves.count(k) for k in sorted(set(moves))})
    att, acc = np.array(st.attempted), np.array(st.accepted)
    for k, name in NAMES.items():
        if att[k]:
            assert abs(acc[k] / att[k] - case["ratios"][name]) <= 5.0001e-10, (name, acc[k], att[k])
        else:
            assert name not in case["ratios"]
    assert int(acc[1:8].sum()) == case["total_accepted"]


def test_gpu_moves_mixed_precision_decisions():
    from fastmc_amd.lib import FMC_MIXED, Handle
    case = REF[0]
    sysm = moves_system(case)
    seeds = tuple(case["seeds"])
    h = Handle(sysm, precision=FMC_MIXED, seeds=[seeds])
    tr = h.run(2000, trace=True)
    h.close()
    o = Oracle(sysm, seed=seeds)
    otr = o.run(2000)
    o.close()
    for i in range(2000):
        for f in FIELDS:
            assert tr[i][f] == otr[i][f], (i, f)
        if not otr[i]["overlap"]:
            assert abs(tr[i]["estep"] - otr[i]["estep"]) <= du_tol(otr[i]["estep"], sysm.temperature, 1e-6)
