"""The batched isotherm workload that bench.py times, checked against the
oracle: 64 state points (T 273/293/323 K split 22/21/21, P log-spaced
1e-3..10 bar), each on its own seed-jittered copy of the 424-atom framework
and at its own temperature, all in one handle -- the per-walker framework
copies and per-walker temperatures of the kernel (beta per walker in
energy_eval, utility_pack.f:3637-3693; the move energies of mc_moves.f:150-484
over each walker's own framework).

FP64: walkers at each temperature follow the oracle chain built from that
state point's own FIELD / CONFIG / CONTROL (read independently of the batch),
step by step through eager traced runs and to the end state through hipGraph
sweeps (the bench's launch path): every decision identical, dU within
1e-10 * max(|dU|, kB T), final coordinates, energies and S(k) equal.

Mixed: the oracle's chains replayed trial by trial through the one-trial seam
on walkers of a mixed and an FP64 handle of the same 64 state points:
|dU_mixed - dU_fp64| <= 1e-6 * max(|dU_fp64|, kB T); the maximum and 99th
percentile per temperature are printed."""
import os
import shutil
import sys
import tempfile

import numpy as np
import pytest

from _helpers import KB_KCAL, du_tol

pytestmark = pytest.mark.gpu

FIELDS = ("iguest", "move", "choice", "nmols", "overlap", "accepted", "drew_rande")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bench_seed(i):
    # the seeds bench.py gives state point i
    return (1 + (7 * i) % 170, 1 + (13 * i) % 170, 1 + (29 * i) % 170, (37 * i) % 168)


@pytest.fixture(scope="module")
def batch():
    from fastmc_amd import inputs as I
    from fastmc_amd.system import read_system
    tmp = tempfile.mkdtemp(prefix="fmc_cfg2_")
    try:
        dirs = I.config2(tmp, n=64)
        systems = [read_system(d) for d in dirs]
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return systems


def make_handle(systems, precision):
    from fastmc_amd.lib import Handle
    from fastmc_amd.system import make_walker
    base, keep = systems[0], []
    walkers = [make_walker(base, seed=bench_seed(i), temperature=s.temperature, pressures=s.pressures,
                           framework_xyz=s.framework_xyz(), keep=keep) for i, s in enumerate(systems)]
    h = Handle(base, walkers=walkers, precision=precision)
    h._keep.extend(keep)
    return h


def chosen(systems):
    """One state point per temperature at its highest pressure (the most
    guests), and the 293 K point nearest 0.1 bar."""
    out = []
    for t in (273.0, 293.0, 323.0):
        idx = [i for i, s in enumerate(systems) if s.temperature == t]
        out.append(max(idx, key=lambda i: systems[i].pressures[0]))
    idx = [i for i, s in enumerate(systems) if s.temperature == 293.0]
    out.append(min(idx, key=lambda i: abs(np.log10(systems[i].pressures[0] / 1e4))))
    return out


def first_difference(tr, otr, temp):
    for i in range(len(tr)):
        a, b = tr[i], otr[i]
        for f in FIELDS:
            if a[f] != b[f]:
                return i, f, a[f], b[f]
        if not b["overlap"] and abs(a["estep"] - b["estep"]) > du_tol(b["estep"], temp):
            return i, "estep", a["estep"], b["estep"]
    return None


def test_gpu_config2_fp64_walkers_follow_their_oracle_chains(batch):
    from fastmc_amd.lib import FMC_FP64
    from oracle.oracle import Oracle
    systems = batch
    pick = chosen(systems)
    say("config 2 FP64: 64 state points; checked", [(i, systems[i].temperature, systems[i].pressures[0])
                                                      for i in pick])
    h = make_handle(systems, FMC_FP64)
    ntr, nsw, nsteps = 500, 500, 3
    traces = []
    for w in pick:
        h.set_trace_walker(w)
        traces.append(h.run(ntr, trace=True))
    h.run_graph(nsteps, nsw)
    total = ntr * len(pick) + nsw * nsteps
    for k, w in enumerate(pick):
        s = systems[w]
        o = Oracle(s, seed=bench_seed(w))
        otr = o.run(total)
        d = first_difference(traces[k], otr[k * ntr:(k + 1) * ntr], s.temperature)
        assert d is None, (w, d)
        st = h.stats(w)
        assert st.steps == total
        g_dev, g_orc = h.guests(1, w=w), o.guests(1)
        assert g_dev.shape == g_orc.shape, (w, g_dev.shape, g_orc.shape)
        np.testing.assert_allclose(g_dev, g_orc, rtol=0, atol=1e-9)
        np.testing.assert_allclose(np.array(st.energy[:2]), o.energy(), rtol=1e-9, atol=1e-8)
        np.testing.assert_allclose(h.sk(w), o.sk(), rtol=0, atol=1e-8)
        say(f"  walker {w}: T={s.temperature:g} K P={s.pressures[0] / 1e5:.4g} bar, {total} steps identical, "
            f"N={len(g_dev) // 3}, accepted {int(sum(otr['accepted']))}")
        o.close()
    h.close()


def test_gpu_config2_mixed_du_error(batch):
    from fastmc_amd.lib import FMC_FP64, FMC_MIXED
    from oracle.oracle import Oracle
    systems = batch
    pick = chosen(systems)[:3]
    hd = make_handle(systems, FMC_FP64)
    hm = make_handle(systems, FMC_MIXED)
    n = 1500
    worst = {}
    for w in pick:
        s = systems[w]
        T = s.temperature
        o = Oracle(s, seed=bench_seed(w))
        otr = o.run(n)
        o.close()
        errs = []
        for r in otr:
            g, mv, c = int(r["iguest"]), int(r["move"]), int(r["choice"])
            xyz = r["xyz"][:9].reshape(3, 3)
            if mv == 1:
                a, b = hd.trial_insert(g, xyz, w=w), hm.trial_insert(g, xyz, w=w)
            elif mv == 2:
                a, b = hd.trial_delete(g, c, w=w), hm.trial_delete(g, c, w=w)
            else:
                a, b = hd.trial_displace(g, c, xyz, w=w), hm.trial_displace(g, c, xyz, w=w)
            assert a.overlap == b.overlap == int(r["overlap"])
            if not a.overlap:
                # the FP64 seam is the oracle's trial energy
                assert abs(a.estep - r["estep"]) <= du_tol(r["estep"], T), (w, a.estep, r["estep"])
                errs.append(abs(a.estep - b.estep) / max(abs(a.estep), KB_KCAL * T))
            if r["accepted"]:
                hd.accept(w=w)
                hm.accept(w=w)
            else:
                hd.reject(w=w)
                hm.reject(w=w)
        errs = np.array(errs)
        worst[w] = (errs.max(), np.quantile(errs, 0.99), len(errs))
        say(f"config 2 mixed, walker {w} (T={T:g} K, P={s.pressures[0] / 1e5:.4g} bar): scaled dU error "
            f"max {errs.max():.3g}, p99 {np.quantile(errs, 0.99):.3g} over {len(errs)} trials")
    hd.close()
    hm.close()
    for w, (mx, p99, cnt) in worst.items():
        assert cnt > 500
        assert mx <= 1e-6, (w, mx)
