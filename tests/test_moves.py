"""jump, swap and switch (gcmc.f:1937-2452) in the C oracle against the
reference's own run of its CALF-20 CO2/N2 test system with the test's move
set (insert, delete, displace, swap, switch, jump at 1/6 each): from the seeds
the reference printed, the restatement must make every decision the
reference made, which the printed acceptance ratio of every move kind and
the total of accepted steps pin (tests/golden/moves_ref.json, made by
make_moves_ref.py).  Swaps are never accepted at this state point (the
reference's swap test multiplies by the volume in A^3), switches need three
guest kinds and do nothing here; both still draw and restore state, which
the later decisions check."""
import json
import os
import shutil
import tempfile

import pytest

from _helpers import GOLDEN
from fastmc_amd.system import read_system
from oracle.oracle import Oracle

REF = json.load(open(os.path.join(GOLDEN, "moves_ref.json")))
NAMES = {1: "insertion", 2: "deletion", 3: "displacement", 4: "translation", 5: "rotation", 6: "jump", 7: "swap",
         8: "switch"}


def moves_system(case):
    d = tempfile.mkdtemp(prefix="fmc_moves_")
    try:
        for f in ("FIELD", "CONFIG"):
            shutil.copy(os.path.join(GOLDEN, "calf20", f), os.path.join(d, f))
        with open(os.path.join(d, "CONTROL"), "w") as fh:
            fh.write(case["control"])
        return read_system(d)
    finally:
        shutil.rmtree(d, ignore_errors=True)


@pytest.mark.parametrize("case", REF, ids=[c["name"] for c in REF])
def test_oracle_moves_match_reference(case):
    sysm = moves_system(case)
    o = Oracle(sysm, seed=tuple(case["seeds"]))
    o.run(case["steps"])
    att, acc = o.move_counts()
    o.close()
    assert att.sum() <= case["steps"]
    for k, name in NAMES.items():
        if att[k] == 0:
            assert name not in case["ratios"], name
            continue
        # f15.9, rounded half away from zero: an exact tie sits 5e-10 away
        assert abs(acc[k] / att[k] - case["ratios"][name]) <= 5.0001e-10, (name, acc[k], att[k], case["ratios"][name])
    # the reference's total leaves switches out (gcmc.f:1464-1466)
    assert int(acc[1:8].sum()) == case["total_accepted"]
