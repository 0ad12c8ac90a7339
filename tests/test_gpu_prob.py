"""Probability plots on the device chain (fmc_set_prob: the binning of every
production step inside k_run, hardware FP64 atomics) against the oracle and
the reference's cube files: each run of tests/golden/prob_ref.json.gz is
replayed in FP64 on the device from the reference's printed seeds; every
grid value must equal the oracle's to 1e-12 relative (the atomics add in no
fixed order) and every printed cube value the reference's to its six
digits.  The run-directory driver writes the cube files."""
import math
import sys

import numpy as np
import pytest

from test_prob import GCMC, WIDOM, plot_list, prob_system, replay_oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class DeviceEngine:
    def __init__(self, h):
        self.h = h

    def run(self, n, flags):
        self.h.run(n, count_stats=flags)

    def prodcount(self):
        return int(self.h.stats(0).prodcount)


def cube_values(text, natoms):
    return np.array([float(v) for v in " ".join(text.split("\n")[6 + natoms:]).split()])


@pytest.mark.parametrize("case", GCMC, ids=[c["name"] for c in GCMC])
def test_gpu_probability_plots_match_oracle_and_reference(case):
    from fastmc_amd import gcmc
    from fastmc_amd.lib import FMC_FP64, Handle
    say(case["name"])
    sysm = prob_system(case)
    plots, names = plot_list(sysm)
    h = Handle(sysm, precision=FMC_FP64, seeds=[tuple(case["seeds"])])
    try:
        _, nwindsteps = gcmc.window_steps(sysm.steps, sysm.nwind)
        h.set_sampling(nwindsteps)
        ngrid = h.set_prob(sysm.griddim, sysm.gridfactor, plots)
        eng = DeviceEngine(h)
        r = gcmc.run_main_loop(eng, sysm.eqsteps, sysm.steps, gcmc.check_interval(sysm.eqsteps, sysm.nnumg), False,
                               eng.prodcount)
        grids = [h.prob(p) for p in range(len(plots))]
    finally:
        h.close()
    assert list(ngrid) == case["grid"]
    _, _, _, ogrids, ro = replay_oracle(case)
    assert (r.start, r.steps, r.prodcount) == (ro.start, ro.steps, ro.prodcount)
    natoms = sum(sp.nmols * len(sp.labels) for sp in sysm.framework_species)
    for name, g, og in zip(names, grids, ogrids):
        assert np.allclose(g, og, rtol=1e-12, atol=1e-12 * og.max()), (name, np.abs(g - og).max())
        ours = cube_values(gcmc.cube_text(sysm, g, ngrid, sysm.gridfactor, r.prodcount), natoms)
        ref = cube_values(case["cubes"][name], natoms)
        assert ours.shape == ref.shape
        assert np.all(np.abs(ours - ref) <= 0.5e-6 * 1.0001 * np.abs(ref) + 1e-300), name


@pytest.mark.parametrize("case", WIDOM, ids=[c["name"] for c in WIDOM])
def test_gpu_widom_energy_map_totals_match_oracle(case, tmp_path):
    """The 'e' plot of a Widom run, binned on the device by k_widom.  Every
    attempt of widom_grid sits exactly on a grid point, where equitable_bin
    splits the weight between the planes on either side according to the
    last bit of the centre of mass; the reference's map therefore depends
    on the rounding of its own arithmetic (libm sin / cos included), and the
    device's can differ from it voxel by voxel.  What the rounding cannot
    move is checked: the total count (non-overlapping attempts) and the
    total energy, against the oracle's map, and the run driver's cube file
    (header and framework as the reference's, every value finite)."""
    from fastmc_amd import gcmc
    from fastmc_amd import inputs as I
    from fastmc_amd.lib import FMC_FP64, Handle
    from oracle.oracle import Oracle
    sysm = prob_system(case)
    h = Handle(sysm, precision=FMC_FP64, seeds=[tuple(case["seeds"])])
    try:
        r = h.widom(1, sysm.nwidstep, sysm.griddim, sysm.gridfactor, energy_map=True, per_attempt=True)
    finally:
        h.close()
    o = Oracle(sysm, seed=tuple(case["seeds"]))
    ro = o.widom(1, sysm.nwidstep, sysm.griddim, sysm.gridfactor, energy_map=True)
    o.close()
    ok = r["overlap"] == 0
    assert abs(r["cmap"].sum() - ro["cmap"].sum()) <= 1e-9 and r["cmap"].sum() == pytest.approx(ok.sum())
    assert r["emap"].sum() == pytest.approx(ro["emap"].sum(), rel=1e-9)
    assert r["emap"].sum() == pytest.approx(r["estep"][ok].sum(), rel=1e-9)
    I.config0(str(tmp_path), pressure_bar=1.0)
    (tmp_path / "CONTROL").write_text(case["control"])
    gcmc.run_directory(str(tmp_path), seed=tuple(case["seeds"]))
    text = (tmp_path / "Ehist_guest01.cube").read_text()
    ref = case["cubes"]["Ehist_guest01.cube"]
    natoms = int(ref.split("\n")[2].split()[0])
    assert text.split("\n")[:6] == ref.split("\n")[:6]
    assert np.isfinite(cube_values(text, natoms)).all() and cube_values(text, natoms).size == cube_values(ref, natoms).size


def test_gpu_run_directory_writes_cube_files(tmp_path):
    from fastmc_amd import gcmc
    case = GCMC[0]
    sysm = prob_system(case)
    from fastmc_amd import inputs as I
    I.config0(str(tmp_path), pressure_bar=1.0)
    (tmp_path / "CONTROL").write_text(case["control"])
    r, _ = gcmc.run_directory(str(tmp_path), seed=tuple(case["seeds"]))
    assert r.prodcount == case["total"] - case["start"] + 1
    for name in case["cubes"]:
        text = (tmp_path / name).read_text()
        natoms = int(text.split("\n")[2].split()[0])
        ours, ref = cube_values(text, natoms), cube_values(case["cubes"][name], natoms)
        assert np.all(np.abs(ours - ref) <= 0.5e-6 * 1.0001 * np.abs(ref) + 1e-300), name
