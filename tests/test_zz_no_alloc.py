"""fmc_widom_noalloc, the variant of fmc_widom that allocates no device
memory (its buffers are one fixed device array per GPU), against fmc_widom
through the C-ABI: two handles with the same generator seeds make the same
calls in the same order (Widom attempts draw from the walker's generator),
and every call of the fixed-scratch variant gives fmc_widom's per-attempt
energies and overlap flags exactly, its non-overlapping count exactly, and
its Boltzmann average, Henry's constant and energy maps to 1e-12 (sums in a
different grouping).  Two calls in a row, the second smaller, so the second
runs on what the first left in the shared array; a map larger than the
array's limit is refused."""
import sys

import numpy as np
import pytest

from _helpers import case_system

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_gpu_widom_noalloc_matches_widom():
    from fastmc_amd.lib import FMC_MIXED, FastMCError, Handle
    sysm = case_system("cfg0_p10")
    seeds = [(12, 34, 56, 78)]
    a = Handle(sysm, precision=FMC_MIXED, seeds=seeds)
    b = Handle(sysm, precision=FMC_MIXED, seeds=seeds)
    try:
        a.run(400)
        b.run(400)
        # 1: 1.0 A grid (26^3 points) x 6 orientations = 105,456 attempts, two chunks of the fixed array
        # 2: 2.0 A grid (13^3) x 3 orientations, smaller, on what call 1 left in the array
        for griddim, nstep in ((1.0, 6), (2.0, 3)):
            say("widom noalloc", griddim, nstep)
            ra = a.widom(1, nstep, griddim, per_attempt=True, energy_map=True)
            rb = b.widom(1, nstep, griddim, per_attempt=True, energy_map=True, noalloc=True)
            assert ra["attempts"] == rb["attempts"] and ra["ngrid"] == rb["ngrid"]
            assert ra["accepted"] == rb["accepted"]
            np.testing.assert_array_equal(ra["estep"], rb["estep"])
            np.testing.assert_array_equal(ra["overlap"], rb["overlap"])
            assert abs(ra["avg_boltz"] - rb["avg_boltz"]) <= 1e-12 * abs(ra["avg_boltz"])
            assert abs(ra["henry"] - rb["henry"]) <= 1e-12 * abs(ra["henry"])
            np.testing.assert_allclose(rb["emap"], ra["emap"], rtol=1e-12, atol=1e-12 * np.abs(ra["emap"]).max())
            np.testing.assert_allclose(rb["cmap"], ra["cmap"], rtol=1e-12, atol=0)
            say("  attempts", ra["attempts"], "non-overlapping", ra["accepted"], "henry", ra["henry"])
        # without the per-attempt outputs and maps: the same averages
        ra = a.widom(1, 2, 1.5)
        rb = b.widom(1, 2, 1.5, noalloc=True)
        assert ra["accepted"] == rb["accepted"]
        assert abs(ra["avg_boltz"] - rb["avg_boltz"]) <= 1e-12 * abs(ra["avg_boltz"])
        # a map beyond the fixed array is refused (0.4 A grid: 65^3 = 274,625 points)
        with pytest.raises(FastMCError) as e:
            b.widom(1, 1, 0.4, energy_map=True, noalloc=True)
        assert e.value.code == 9005
    finally:
        a.close()
        b.close()
