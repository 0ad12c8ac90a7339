"""Probability plots (storeprob / equitable_bin, utility_pack.f:622-687,
2407-2519; the cube files of writeprob, 773-871; the energy plot of
storeenprob / writeenprob, 550-620, 689-771) in the C oracle against the
reference's own cube files (tests/golden/prob_ref.json.gz, made by
make_prob_ref.py): the oracle replays each reference run from its printed
seeds through the library's main-loop driver, bins every production step,
and the library's cube writer (fastmc_amd.gcmc.cube_text) must reproduce
each file as printed (E15.6 values, header, framework atoms).  The config-0
framework's synthetic masses match no element of atmnumber's table, for
which the reference returns an undefined value: there the atom lines are
compared without their first field.  The energy plot of a GCMC run is all
zeros (the reference bins a variable that stays 0); that of a Widom run
('henry') holds the mean insertion energy per grid point, from the
oracle's widom_grid with the map.  Also the grid size (gcmc.f:258-281)."""
import gzip
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
from test_stats import OracleEngine

REF = json.load(gzip.open(os.path.join(GOLDEN, "prob_ref.json.gz"), "rt"))


def prob_system(case):
    d = tempfile.mkdtemp(prefix="fmc_prob_")
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


def plot_list(sysm):
    """[(iguest, sites)] in the reference's grid order, with cube file names."""
    plots, names = [], []
    for g, pl in enumerate(sysm.prob_plots, start=1):
        for k, sites in enumerate(pl, start=1):
            plots.append((g, sites))
            names.append(f"prob_guest{g:02d}_prob_{k:02d}.cube")
    return plots, names


def replay_oracle(case):
    sysm = prob_system(case)
    ngrid = gcmc.prob_grid_size(sysm.cell, sysm.griddim, sysm.gridfactor)
    plots, names = plot_list(sysm)
    o = Oracle(sysm, seed=tuple(case["seeds"]))
    _, nwindsteps = gcmc.window_steps(sysm.steps, sysm.nwind)
    o.set_sampling(nwindsteps)
    o.set_prob(ngrid, sysm.gridfactor, plots)
    eng = OracleEngine(o)
    r = gcmc.run_main_loop(eng, sysm.eqsteps, sysm.steps, gcmc.check_interval(sysm.eqsteps, sysm.nnumg), False,
                           eng.prodcount)
    grids = [o.prob(p) for p in range(len(plots))]
    o.close()
    return sysm, ngrid, names, grids, r


def same_cube(ours, ref, atom_z=True):
    a, b = ours.split("\n"), ref.split("\n")
    assert len(a) == len(b)
    natoms = int(b[2].split()[0])
    for i, (x, y) in enumerate(zip(a, b)):
        if not atom_z and 6 <= i < 6 + natoms:
            assert x[6:] == y[6:], (i, x, y)
        else:
            assert x == y, (i, x, y)


GCMC = [c for c in REF if not c["name"].endswith("widom_map")]
WIDOM = [c for c in REF if c["name"].endswith("widom_map")]


@pytest.mark.parametrize("case", GCMC, ids=[c["name"] for c in GCMC])
def test_oracle_probability_plots_match_reference_cubes(case):
    sysm, ngrid, names, grids, r = replay_oracle(case)
    assert list(ngrid) == case["grid"]
    assert (r.start, r.steps) == (case["start"], case["total"])
    enames = [f"Ehist_guest{g:02d}.cube" for g, on in enumerate(sysm.prob_energy, start=1) if on]
    assert sorted(names + enames) == sorted(case["cubes"])
    for name, grid in zip(names, grids):
        assert grid.sum() > 0
        text = gcmc.cube_text(sysm, grid, ngrid, sysm.gridfactor, r.prodcount)
        same_cube(text, case["cubes"][name], atom_z=case["name"].startswith("calf20"))
    zero = [0.0] * (ngrid[0] * ngrid[1] * ngrid[2])
    for name in enames:
        same_cube(gcmc.enprob_text(sysm, zero, zero, ngrid, sysm.gridfactor), case["cubes"][name],
                  atom_z=case["name"].startswith("calf20"))


@pytest.mark.parametrize("case", WIDOM, ids=[c["name"] for c in WIDOM])
def test_oracle_widom_energy_map_matches_reference_cube(case):
    sysm = prob_system(case)
    assert sysm.widom and sysm.prob_energy == [True]
    ngrid = gcmc.prob_grid_size(sysm.cell, sysm.griddim, sysm.gridfactor)
    assert list(ngrid) == case["grid"]
    o = Oracle(sysm, seed=tuple(case["seeds"]))
    r = o.widom(1, sysm.nwidstep, sysm.griddim, sysm.gridfactor, energy_map=True)
    o.close()
    assert r["attempts"] == case["total"]
    assert (r["cmap"] > 0).all()
    same_cube(gcmc.enprob_text(sysm, r["emap"], r["cmap"], ngrid, sysm.gridfactor), case["cubes"]["Ehist_guest01.cube"],
              atom_z=False)


def test_grid_size_and_atomic_numbers():
    # CALF-20 (triclinic) at 1.5 A and config 0 at 2.0 A, as the reference printed
    assert gcmc.prob_grid_size(prob_system(REF[1]).cell, 1.5) == tuple(REF[1]["grid"])
    assert gcmc.prob_grid_size(prob_system(REF[0]).cell, 2.0) == tuple(REF[0]["grid"])
    assert [gcmc.atmnumber(m) for m in (1.00794, 12.0107, 14.0067, 15.9994, 65.38, 39.948, 39.0983)] == \
        [1, 6, 7, 8, 30, 18, 19]
    assert gcmc.atmnumber(500.0) == 0
