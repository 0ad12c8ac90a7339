"""numguests.out (gcmc.f:1138-1163) in the C oracle against the reference's
own files (tests/golden/numguests_ref.json, made by make_numguests_ref.py):
replaying each reference run from its printed seeds, the oracle's records
every nnumg steps, through the library's writer (fastmc_amd.gcmc.
numguests_text), must reproduce every file line for line: step, molecule
count, energy and delE (f20.6) and the move flags of the step (insert,
delete, displace / translate / rotate as 1 / 2 / 3, jump, swap, switch).
The CALF-20 run has swaps and switches (a switch leaves delE as the step
before left it)."""
import json
import os
import shutil
import tempfile

import pytest

from _helpers import GOLDEN
from fastmc_amd import gcmc
from fastmc_amd import inputs as I
from fastmc_amd.system import read_system
from oracle.oracle import Oracle

REF = json.load(open(os.path.join(GOLDEN, "numguests_ref.json")))


def numg_system(case):
    d = tempfile.mkdtemp(prefix="fmc_numg_")
    try:
        if case["name"].startswith("calf20"):
            for f in ("FIELD", "CONFIG"):
                shutil.copy(os.path.join(GOLDEN, "calf20", f), os.path.join(d, f))
        else:
            I.config0(d, pressure_bar=1.0)
        with open(os.path.join(d, "CONTROL"), "w") as fh:
            fh.write(case["control"])
        return read_system(d)
    finally:
        shutil.rmtree(d, ignore_errors=True)


@pytest.mark.parametrize("case", REF, ids=[c["name"] for c in REF])
def test_oracle_numguests_match_reference_files(case):
    sysm = numg_system(case)
    o = Oracle(sysm, seed=tuple(case["seeds"]))
    o.set_numguests(sysm.nnumg)
    o.run(case["total"])
    recs, n = o.numguests()
    o.close()
    names = gcmc.numguests_names(len(sysm.guests))
    assert sorted(names) == sorted(case["files"])
    for g, name in enumerate(names, start=1):
        assert gcmc.numguests_text(recs, n, g) == case["files"][name], name
