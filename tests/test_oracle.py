"""The oracle against the reference: golden traces recorded from the compiled
reference, the reference's own known answer, and closed-form cases."""
import math

import numpy as np
import pytest

from _helpers import CASES, case_system, compare_trace, fixture
from oracle.oracle import Oracle, duni_seq


@pytest.mark.parametrize("name", CASES)
def test_oracle_replays_reference_trace(name):
    fx = fixture(name)
    o = Oracle(case_system(name), seed=fx["seeds"])
    assert o.newld == fx["sizes"]["newld"]
    tr = o.run(len(fx["trace"]["step"]))
    bad, worst = compare_trace(tr, fx)
    assert not bad, f"{name}: first mismatching steps {bad[:5]}"
    assert worst < 1e-10
    # per-species energy() after the last step
    np.testing.assert_allclose(tr[-1]["energy"][: o.maxmls], fx["trace"]["energy"][-1], rtol=1e-9, atol=1e-9)
    # S(k) checksum and guest count at the end
    last = fx["sk_checks"][-1]
    sk = o.sk()
    assert abs(sk[:, 0].sum() - last[1]) < 1e-8 * max(1.0, abs(last[1]))
    assert abs(sk[:, 1].sum() - last[2]) < 1e-8 * max(1.0, abs(last[2]))


def test_calf20_known_answer():
    """run_test.py:40: initial electrostatic energy of the reference's CALF-20 test."""
    o = Oracle(case_system("calf20_co2n2"), seed=(1, 2, 3, 4))
    assert f"{o.single_point()[0]:.6f}" == "-11190.606165"


@pytest.mark.parametrize("name", ["cfg0_p0.1", "calf20_co2n2"])
def test_oracle_tables_and_sk_match_reference(name):
    fx = fixture(name)
    o = Oracle(case_system(name), seed=fx["seeds"])
    erc = o.table(0)
    for i, v in fx["erc"]:
        assert abs(erc[i - 1] - v) <= 1e-14 * abs(v)
    for k, i, v in fx["vvv"]:
        col = o.table(1, k)
        assert abs(col[i - 1] - v) <= 1e-13 * max(abs(v), 1e-30)
    # lrcorrect's steadd (vdw_terms.f:587-588,639-645), indexed by the type
    # pair's triangular index; the fixture lists steadd(1..ntpvdw) beside the FIELD rows
    stead = o.table(2, n=len(fx["lstvdw"]))
    assert len(fx["pairs"]) > 0
    for ll, kind, eps, sig, st in fx["pairs"]:
        assert abs(stead[ll - 1] - st) <= 1e-14 * abs(st), (ll, stead[ll - 1], st)
    # total S(k) right after single_point at the sampled reference k indices (ewald1)
    sk = o.sk()
    assert len(fx["sk0"]) > 10
    for k, c, s, c2, s2 in fx["sk0"]:
        assert abs(sk[k - 1, 0] - c) <= 1e-12 * max(1.0, abs(c)), (k, sk[k - 1, 0], c)
        assert abs(sk[k - 1, 1] - s) <= 1e-12 * max(1.0, abs(s)), (k, sk[k - 1, 1], s)
    assert o.newld == fx["sizes"]["newld"] and o.mxegrd == fx["sizes"]["mxegrd"]


def test_duni_published_method_properties():
    """MZT universal generator: outputs in [0,1) on the 2^-24 grid, full-period-ish spread."""
    x = duni_seq((12, 34, 56, 78), 200000)
    assert x.min() >= 0.0 and x.max() < 1.0
    assert np.all(np.abs(x * 2 ** 24 - np.round(x * 2 ** 24)) == 0.0)
    assert abs(x.mean() - 0.5) < 5e-3 and abs(x.var() - 1 / 12) < 5e-3
    # the reference's fixed DEBUG seeds (utility_pack.f:2806-2809) give a repeatable stream
    a, b = duni_seq((25, 89, 100, 150), 10), duni_seq((25, 89, 100, 150), 10)
    assert np.array_equal(a, b)


def test_insert_delete_are_inverse():
    """Inserting x then deleting it: the deletion estep (the molecule's own energy, which
    mc_delete passes to energy_eval negated, gcmc.f:1795) equals the insertion estep, term by
    term (tail and self terms included)."""
    sysm = case_system("cfg0_p10")
    o = Oracle(sysm, seed=fixture("cfg0_p10")["seeds"])
    o.run(600)
    rng = np.random.default_rng(7)
    tmpl = sysm.guests[0].template
    for _ in range(5):
        com = rng.uniform(0, 25.8, 3)
        xyz = tmpl + com
        ti = o.trial_insert(1, xyz)
        if ti.overlap:
            o.reject()
            continue
        assert o.accept() == 0
        n = len(o.guests(1)) // len(tmpl)
        td = o.trial_delete(1, n)
        assert abs(ti.estep - td.estep) < 1e-9 * max(1.0, abs(ti.estep))
        for a in ("recip", "real", "vdw", "tail"):
            assert abs(getattr(ti, a) - getattr(td, a)) < 1e-9 * max(1.0, abs(getattr(ti, a)))
        o.reject()


def test_zero_and_lattice_displacement_give_zero_du():
    sysm = case_system("cfg0_p10")
    o = Oracle(sysm, seed=fixture("cfg0_p10")["seeds"])
    o.run(400)
    g = o.guests(1)
    for choice in (1, 2):
        xyz = g[3 * (choice - 1): 3 * choice]
        t0 = o.trial_displace(1, choice, xyz)
        assert abs(t0.estep) < 1e-10
        o.reject()
        t1 = o.trial_displace(1, choice, xyz + np.array([25.8, 0.0, -25.8]))
        assert abs(t1.estep) < 1e-8
        o.reject()


def test_sk_incremental_equals_resum():
    """S(k) after many incremental updates equals a fresh FP64 re-sum (ewald1)."""
    o = Oracle(case_system("cfg0_p10"), seed=fixture("cfg0_p10")["seeds"])
    o.run(1500)
    sk = o.sk().copy()
    drift = o.full_resum()
    assert drift < 1e-11 * max(1.0, np.abs(sk).max())


def test_tables_against_closed_forms():
    """erfc table = A&S 7.1.26 erfc(alpha r)/r within the formula's 1.5e-7 bound (erfcgen,
    ewald.f:10-25); LJ table = 4 eps ((s/r)^12 - (s/r)^6) at every grid point (forgen,
    vdw_terms.f:229-233); k-vector count and a_k = exp(-k^2/4a^2)/k^2."""
    sysm = case_system("cfg0_p0.1")
    o = Oracle(sysm, seed=(1, 2, 3, 4))
    a = sysm.alpha
    erc = o.table(0)
    r = 0.01 * np.arange(1, o.mxegrd + 1)
    ex = np.array([math.erfc(a * x) for x in r])
    assert np.max(np.abs(erc * r - ex)) < 1.5e-7
    for k, (i, j, kind, eps, sig) in enumerate(sysm.pairs, start=1):
        col = o.table(1, k)
        lj = 4 * eps * ((sig / r) ** 12 - (sig / r) ** 6)
        assert np.allclose(col, lj, rtol=1e-12, atol=0)
    lmn, akk = o.kvecs()
    L = 25.8
    k2 = ((2 * np.pi / L) ** 2) * (lmn ** 2).sum(1)
    assert np.allclose(akk, np.exp(-k2 / (4 * a * a)) / k2, rtol=1e-13)
    # half space: no k and -k pair both present, and the 1.05 sphere cut
    keys = {tuple(v) for v in lmn}
    assert not any(tuple(-np.array(v)) in keys for v in lmn)
    assert np.sqrt(k2).max() <= 7 * 2 * np.pi / L * 1.05 + 1e-12
