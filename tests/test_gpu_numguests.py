"""numguests.out records of the device chain (fmc_set_numguests, written by
k_run every nnumg steps) against the reference's own files: each run of
tests/golden/numguests_ref.json replayed in FP64 on the device from the
reference's printed seeds gives every line's step, molecule count and move
flags exactly and its energy and delE to the sixth printed decimal (one
unit, since values within the chain's FP64 bound of 1e-10 relative can round
either way)."""
import sys

import numpy as np
import pytest

from test_numguests import REF, numg_system

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("case", REF, ids=[c["name"] for c in REF])
def test_gpu_numguests_match_reference_files(case):
    from fastmc_amd import gcmc
    from fastmc_amd.lib import FMC_FP64, Handle
    say(case["name"])
    sysm = numg_system(case)
    h = Handle(sysm, precision=FMC_FP64, seeds=[tuple(case["seeds"])])
    try:
        h.set_numguests(sysm.nnumg, case["total"] // sysm.nnumg + 1)
        h.run(case["total"])
        recs, n = h.numguests(0)
        assert h.stats(0).overflow == 0
    finally:
        h.close()
    for g, name in enumerate(gcmc.numguests_names(len(sysm.guests)), start=1):
        ours = gcmc.numguests_text(recs, n, g).splitlines()
        ref = case["files"][name].splitlines()
        assert len(ours) == len(ref)
        for a, b in zip(ours, ref):
            fa, fb = a.split(), b.split()
            assert fa[:2] == fb[:2] and fa[4:] == fb[4:], (a, b)
            for x, y in zip(fa[2:4], fb[2:4]):
                x, y = float(x), float(y)
                # a device energy within 1e-10 of the reference's may print one unit off in the 6th decimal
                assert abs(x - y) <= 1.0001e-6 + 1e-10 * abs(y), (a, b)
