"""GPU parity: the HIP path (through the C-ABI) against the reference's golden
traces and the oracle.  FP64 bound: |dU - dU_ref| <= 1e-10 * max(|dU_ref|, kB T)."""
import sys

import numpy as np
import pytest

from _helpers import CASES, KB_KCAL, case_system, compare_trace, du_tol, fixture
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _lib():
    from fastmc_amd.lib import FMC_FP64, FMC_MIXED, Handle
    return Handle, FMC_FP64, FMC_MIXED


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", CASES)
def test_gpu_replays_reference_trace_fp64(name):
    _replay(name)


@pytest.mark.parametrize("env", ["FMC_NO_LDS_ALL", "FMC_NO_LDS_ALL+FMC_NO_LJ_LDS", "FMC_NO_NSITE"])
@pytest.mark.parametrize("name", ["cfg0_moves5", "calf20_co2n2"])
def test_gpu_replay_kernel_variants(name, env, monkeypatch):
    """The kernel variant with only the LJ columns in LDS, the all-global
    variant and the general site-count variant replay the same reference
    traces (the default plan takes the all-in-LDS, at-most-3-sites variant
    for these systems)."""
    for e in env.split("+"):
        monkeypatch.setenv(e, "1")
    _replay(name)


def _replay(name):
    Handle, FP64, _ = _lib()
    fx = fixture(name)
    say("replay", name)
    h = Handle(case_system(name), precision=FP64, seeds=[fx["seeds"]])
    assert h.newld == fx["sizes"]["newld"]
    tr = h.run(len(fx["trace"]["step"]), trace=True)
    bad, worst = compare_trace(tr, fx)
    assert not bad, f"{name}: first mismatching steps {bad[:5]}"
    assert worst <= 1e-10
    np.testing.assert_allclose(tr[-1]["energy"][: h.maxmls], fx["trace"]["energy"][-1], rtol=1e-9, atol=1e-9)
    last = fx["sk_checks"][-1]
    sk = h.sk()
    assert abs(sk[:, 0].sum() - last[1]) < 1e-8 * max(1.0, abs(last[1]))
    assert abs(sk[:, 1].sum() - last[2]) < 1e-8 * max(1.0, abs(last[2]))
    h.close()


def test_gpu_calf20_known_answer():
    """run_test.py:40: -11190.606165 kcal/mol, the reference test's initial Ewald energy."""
    Handle, FP64, _ = _lib()
    h = Handle(case_system("calf20_co2n2"), precision=FP64)
    sp = h.single_point()
    assert f"{sp[0]:.6f}" == "-11190.606165"
    h.close()


def _random_trials(sysm, n, seed):
    rng = np.random.default_rng(seed)
    a = sysm.cell[0]
    tmpl = [g.template for g in sysm.guests]
    for _ in range(n):
        yield rng.integers(0, 3), rng.integers(1, len(tmpl) + 1), rng.uniform(0, a, 3), rng.uniform(0, 1)


def _rot(rng_u):
    th = 2 * np.pi * rng_u
    c, s = np.cos(th), np.sin(th)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])


@pytest.mark.parametrize("name", ["cfg0_p10", "calf20_co2n2"])
def test_gpu_seam_matches_oracle(name):
    """insertion / deletion / displacement through the one-trial C-ABI, accept and reject."""
    Handle, FP64, _ = _lib()
    fx = fixture(name)
    sysm = case_system(name)
    say("seam", name)
    o = Oracle(sysm, seed=fx["seeds"])
    h = Handle(sysm, precision=FP64, seeds=[fx["seeds"]])
    o.run(0)
    T = sysm.temperature
    nchk = 0
    for kind, g, com, u in _random_trials(sysm, 160, 11):
        tmpl = sysm.guests[g - 1].template
        og = o.guests(g)
        nm = len(og) // len(tmpl)
        if kind == 0 or nm == 0:
            xyz = tmpl @ _rot(u).T + com
            a, b = o.trial_insert(g, xyz), h.trial_insert(g, xyz)
        elif kind == 1:
            c = 1 + int(u * nm)
            a, b = o.trial_delete(g, c), h.trial_delete(g, c)
        else:
            c = 1 + int(u * nm)
            mol = og[len(tmpl) * (c - 1): len(tmpl) * c]
            xyz = (mol - mol.mean(0)) @ _rot(u).T + mol.mean(0) + (com / sysm.cell[0] - 0.5)
            a, b = o.trial_displace(g, c, xyz), h.trial_displace(g, c, xyz)
        assert a.overlap == b.overlap
        if not a.overlap or kind == 1:
            assert abs(a.estep - b.estep) <= du_tol(a.estep, T), (kind, a.estep, b.estep)
            for i in range(sysm and len(sysm.species)):
                assert abs(a.delE[i] - b.delE[i]) <= du_tol(a.delE[i], T) * 10
            nchk += 1
        take = (not a.overlap) and (a.estep < 0 or u < 0.3)
        if take:
            assert o.accept() == 0
            h.accept()
        else:
            o.reject()
            h.reject()
    assert nchk > 50
    np.testing.assert_allclose(h.sk(), o.sk(), rtol=0, atol=1e-9)
    for g in range(1, len(sysm.guests) + 1):
        np.testing.assert_allclose(h.guests(g), o.guests(g), rtol=0, atol=1e-12)
    h.close()


def test_gpu_sk_resum_and_batched_walkers():
    """Batched chains: each walker follows its own reference-faithful chain; S(k) after the
    run equals a fresh FP64 re-sum."""
    Handle, FP64, _ = _lib()
    sysm = case_system("cfg0_p10")
    seeds = [(11, 22, 33, 44), (55, 66, 77, 88), (99, 100, 101, 102), (150, 151, 152, 153)]
    press = [[1.0e5], [3.0e5], [1.0e6], [2.0e6]]
    say("batched 4 walkers")
    h = Handle(sysm, precision=FP64, seeds=seeds, pressures=press, n=4)
    h.run(800)
    drift = h.full_resum()
    assert drift < 1e-10
    for w in range(4):
        o = Oracle(sysm, seed=seeds[w], pressures=press[w])
        o.run(800)
        st = h.stats(w)
        np.testing.assert_allclose(h.guests(1, w=w), o.guests(1), rtol=0, atol=1e-9)
        np.testing.assert_allclose(np.array(st.energy[:2]), o.energy(), rtol=1e-9, atol=1e-8)
    h.close()


def test_gpu_inkernel_fp64_resum():
    """The periodic in-kernel FP64 re-sum of S(k): in FP64 mode it changes nothing
    beyond round-off (the reference trace still replays with re-sums every 10
    accepted moves); in mixed mode it bounds the drift of the FP32-updated S(k)."""
    Handle, FP64, MIXED = _lib()
    name = "cfg0_moves5"
    fx = fixture(name)
    sysm = case_system(name)
    say("in-kernel resum, FP64 replay", name)
    h = Handle(sysm, precision=FP64, seeds=[fx["seeds"]])
    h.set_resum(10)
    tr = h.run(len(fx["trace"]["step"]), trace=True)
    bad, worst = compare_trace(tr, fx)
    assert not bad, f"first mismatching steps {bad[:5]}"
    assert worst <= 1e-10
    st = h.stats(0)
    assert st.resums == sum(fx["trace"]["accepted"]) // 10
    assert st.resum_drift < 1e-10
    h.close()
    say("in-kernel resum, mixed drift")
    hm = Handle(sysm, precision=MIXED, seeds=[fx["seeds"]])
    hm.set_resum(25)
    hm.run(3000)
    st = hm.stats(0)
    assert st.resums >= 10
    assert 0.0 < st.resum_drift < 1e-4
    assert hm.full_resum(0) < 1e-4
    hm.close()


def test_gpu_graph_replays_match_eager():
    """hipGraph-captured sweeps give the same chains as eager launches (replayed 3x)."""
    Handle, FP64, _ = _lib()
    sysm = case_system("cfg0_p0.1")
    seeds = [(3, 5, 7, 9), (13, 15, 17, 19)]
    a = Handle(sysm, precision=FP64, seeds=seeds, n=2)
    b = Handle(sysm, precision=FP64, seeds=seeds, n=2)
    for _ in range(3):
        a.run_graph(4, 50)
        a.run(10)
        for _ in range(4):
            b.run(50)
        b.run(10)
        for w in range(2):
            np.testing.assert_array_equal(a.guests(1, w=w), b.guests(1, w=w))
            np.testing.assert_array_equal(a.sk(w), b.sk(w))
    a.close()
    b.close()


@pytest.mark.parametrize("name", ["cfg0_p0.1", "cfg0_p10", "cfg1_p10"])
def test_gpu_mixed_precision_du_error(name):
    """Mixed mode (FP32 pair/phase arithmetic) against FP64 on the same trials.
    Bound: |dU_mixed - dU_fp64| <= 1e-6 * max(|dU_fp64|, kB T)."""
    Handle, FP64, MIXED = _lib()
    fx = fixture(name)
    sysm = case_system(name)
    say("mixed", name)
    hd = Handle(sysm, precision=FP64, seeds=[fx["seeds"]])
    hm = Handle(sysm, precision=MIXED, seeds=[fx["seeds"]])
    tr = hd.run(len(fx["trace"]["step"]), trace=True)
    hm2 = Handle(sysm, precision=MIXED, seeds=[fx["seeds"]])
    # replay the FP64 chain through the seam on a mixed handle and an FP64 handle
    hd2 = Handle(sysm, precision=FP64, seeds=[fx["seeds"]])
    T = sysm.temperature
    errs = []
    for r in tr[:1500]:
        g, mv, c = int(r["iguest"]), int(r["move"]), int(r["choice"])
        xyz = r["xyz"][:9].reshape(3, 3)
        if mv == 1:
            a, b = hd2.trial_insert(g, xyz), hm2.trial_insert(g, xyz)
        elif mv == 2:
            a, b = hd2.trial_delete(g, c), hm2.trial_delete(g, c)
        else:
            a, b = hd2.trial_displace(g, c, xyz), hm2.trial_displace(g, c, xyz)
        if not a.overlap:
            errs.append(abs(a.estep - b.estep) / max(abs(a.estep), KB_KCAL * T))
        if r["accepted"]:
            hd2.accept()
            hm2.accept()
        else:
            hd2.reject()
            hm2.reject()
    errs = np.array(errs)
    say(name, "mixed dU scaled error: max", errs.max(), "p99", np.quantile(errs, 0.99), "n", len(errs))
    assert errs.max() <= 1e-6
    for hh in (hd, hm, hm2, hd2):
        hh.close()


def test_gpu_seam_interleaved_walkers():
    """Trials of two walkers pending at once: each accept applies its own
    walker's trial (the accept data are kept per walker)."""
    Handle, FP64, _ = _lib()
    fx = fixture("cfg0_p10")
    sysm = case_system("cfg0_p10")
    seeds = [tuple(fx["seeds"]), (7, 11, 13, 17)]
    say("seam, two walkers interleaved")
    orc = [Oracle(sysm, seed=s) for s in seeds]
    for o in orc:
        o.run(300)
    h = Handle(sysm, precision=FP64, seeds=seeds, n=2)
    h.run(300)
    for w in range(2):
        np.testing.assert_allclose(h.guests(1, w=w), orc[w].guests(1), rtol=0, atol=1e-9)
    T = sysm.temperature
    tmpl = sysm.guests[0].template
    rng = np.random.default_rng(5)
    naccept = 0
    for it in range(60):
        trials = []
        for w in range(2):
            kind = int(rng.integers(0, 3))
            og = orc[w].guests(1)
            nm = len(og) // 3
            com = rng.uniform(0, sysm.cell[0], 3)
            if kind == 0 or nm == 0:
                xyz = tmpl @ _rot(rng.uniform()).T + com
                a, b = orc[w].trial_insert(1, xyz), h.trial_insert(1, xyz, w=w)
            elif kind == 1:
                c = 1 + int(rng.integers(0, nm))
                a, b = orc[w].trial_delete(1, c), h.trial_delete(1, c, w=w)
            else:
                c = 1 + int(rng.integers(0, nm))
                mol = og[3 * (c - 1): 3 * c]
                xyz = mol + rng.uniform(-0.5, 0.5, 3)
                a, b = orc[w].trial_displace(1, c, xyz), h.trial_displace(1, c, xyz, w=w)
            assert a.overlap == b.overlap
            if not a.overlap or kind == 1:
                assert abs(a.estep - b.estep) <= du_tol(a.estep, T), (it, w, kind, a.estep, b.estep)
            trials.append(a.overlap)
        # decide both after both trials ran: walker 1's trial must not leak into walker 0's accept
        for w in range(2):
            if not trials[w] and rng.uniform() < 0.6:
                assert orc[w].accept() == 0
                h.accept(w=w)
                naccept += 1
            else:
                orc[w].reject()
                h.reject(w=w)
    assert naccept > 20
    for w in range(2):
        np.testing.assert_allclose(h.guests(1, w=w), orc[w].guests(1), rtol=0, atol=1e-9)
        np.testing.assert_allclose(np.array(h.stats(w).energy[:2]), orc[w].energy(), rtol=1e-9, atol=1e-8)
        np.testing.assert_allclose(h.sk(w), orc[w].sk(), rtol=0, atol=1e-9)
    h.close()
