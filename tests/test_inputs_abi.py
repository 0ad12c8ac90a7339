"""Input reader against the reference's parsed values; C-ABI export check."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from _helpers import CASES, case_system, fixture
from fastmc_amd import inputs as I
from fastmc_amd.system import dblstr, intstr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("name", CASES)
def test_reader_matches_reference_globals(name):
    fx = fixture(name)
    s = case_system(name)
    assert s.imcon == fx["sizes"]["imcon"]
    assert tuple(s.kmax) == (fx["sizes"]["kmax1"], fx["sizes"]["kmax2"], fx["sizes"]["kmax3"])
    assert s.alpha == fx["params"]["alpha"]
    assert np.array_equal(s.cell, np.array(fx["cell"]))
    assert abs(s.volume() - fx["params"]["volm"]) <= 1e-12 * fx["params"]["volm"]
    assert s.engunit == fx["params"]["engunit"]
    assert s.sumchg == fx["params"]["sumchg"]
    assert s.temperature == fx["params"]["temp"]
    assert s.overlap ** 2 == fx["params"]["overlap2"]
    assert np.array_equal(s.move_cum(0).astype(np.float64), np.array(fx["move_cum"]))
    for g, (p, fug) in enumerate(zip(s.pressures, fx["fugacity"])):
        assert p == fug
    # guest templates and charges bit-identical to the reference's module arrays
    for row in fx["guest_sites"]:
        g, site, x, y, z, q, m = row
        sp = s.guests[g - 1]
        assert tuple(sp.template[site - 1]) == (x, y, z)
        assert sp.charge[site - 1] == q and sp.mass[site - 1] == m
    # LJ parameters in internal units
    for k, kind, eps, sig, _ in fx["pairs"]:
        assert s.pairs[k - 1][2] == kind and s.pairs[k - 1][3] == eps and s.pairs[k - 1][4] == sig


def test_dblstr_grammar():
    assert dblstr("0.700000")[0] == 0.7000000000000001  # digit-by-digit accumulation
    assert dblstr("ewald precision  1d-6")[0] == 1e-6
    assert dblstr("  pressure (bar) 0.10000000")[0] == 0.1
    assert dblstr("-0.350000 1.16")[0] == -0.35000000000000003
    assert intstr("  0   3  1584")[0] == 0


def test_generator_is_deterministic_and_sized(tmp_path):
    fw = I.make_framework()
    assert len(fw.names) == 424 and abs(fw.charges.sum()) < 1e-9
    assert len(I.supercell(fw).names) == 3392
    pts = I.config2_state_points(64)
    assert len(pts) == 64 and sorted({t for t, _ in pts}) == [273.0, 293.0, 323.0]
    a = I.make_framework(jitter_seed=5).xyz
    b = I.make_framework(jitter_seed=5).xyz
    assert np.array_equal(a, b)


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "fastmc.h")).read()
    declared = set(re.findall(r"^int (fmc_\w+)\(", header, flags=re.M))
    lib = os.path.join(ROOT, "fastmc_amd", "libfastmc_hip.so")
    assert os.path.exists(lib), "build the library first (__graft_entry__.build())"
    L = C.CDLL(lib)
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing and len(declared) >= 19
