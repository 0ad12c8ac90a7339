"""The FP64 numpy oracle of the energy gradient (oracle/adjoint_ref.py),
pinned on the CPU: its energy equals the reference-faithful C oracle's
single_point energy of the same configuration, and its analytic gradient
equals central finite differences of that C-oracle energy (theta moved in the
FIELD / CONFIG the oracle reads) and of its own energy."""
import copy
import shutil
import tempfile

import numpy as np
import pytest

from fastmc_amd import inputs as I
from fastmc_amd.system import read_system
from oracle.adjoint_ref import EnergyModel
from oracle.oracle import Oracle


def _system(pressure_bar=1.0):
    d = tempfile.mkdtemp(prefix="fmc_adj_")
    try:
        I.config0(d, pressure_bar=pressure_bar)
        return read_system(d)
    finally:
        shutil.rmtree(d, ignore_errors=True)


def theta0(sysm):
    q = np.concatenate([np.tile(sp.charge, sp.nmols) for sp in sysm.framework_species])
    nt = len(sysm.labels)
    eps, sig = np.zeros(nt), np.ones(nt)
    for (i, j, kind, e, s) in sysm.pairs:
        if i == j and kind == 2:
            eps[i - 1], sig[i - 1] = e, s
    return q, eps, sig


def configuration(sysm, steps=3000, seed=(12, 34, 56, 78)):
    o = Oracle(sysm, seed=seed)
    o.run(steps)
    xyz = o.guests(1)
    o.close()
    return np.array([len(xyz) // sysm.guests[0].natoms]), xyz


def with_theta(sysm, q, eps, sig, nmols, xyz):
    """The same system with theta written into it (LB pairs for framework
    types) and the configuration's guests, for the C oracle's single_point."""
    s = copy.deepcopy(sysm)
    o = 0
    for sp in s.framework_species:
        n = sp.natoms * sp.nmols
        sp.charge = np.asarray(q[o: o + n].reshape(sp.nmols, sp.natoms)[0], dtype=float)
        assert np.allclose(np.tile(sp.charge, sp.nmols), q[o: o + n])
        o += n
    fwt = set(np.concatenate([sp.types for sp in s.framework_species]).tolist())
    pairs = []
    for (i, j, kind, e, sg) in s.pairs:
        if kind == 2 and (i in fwt or j in fwt):
            e, sg = np.sqrt(eps[i - 1] * eps[j - 1]), 0.5 * (sig[i - 1] + sig[j - 1])
        pairs.append((i, j, kind, e, sg))
    s.pairs = pairs
    g = s.guests[0]
    g.nmols = int(nmols[0])
    g.xyz = np.asarray(xyz, dtype=float).reshape(-1, 3)
    return s


def oracle_energy(sysm, q, eps, sig, nmols, xyz):
    o = Oracle(with_theta(sysm, q, eps, sig, nmols, xyz))
    sp = o.single_point()
    o.close()
    return sp[0] + sp[3]


@pytest.fixture(scope="module")
def case():
    sysm = _system()
    nmols, xyz = configuration(sysm)
    assert nmols[0] >= 3
    return sysm, EnergyModel(sysm), nmols, xyz


def test_numpy_energy_equals_oracle_single_point(case):
    sysm, model, nmols, xyz = case
    q, eps, sig = theta0(sysm)
    U = model.energy_grad(q, eps, sig, nmols, xyz)[0]
    Uo = oracle_energy(sysm, q, eps, sig, nmols, xyz)
    assert abs(U - Uo) <= 1e-9 * max(1.0, abs(Uo)), (U, Uo)


def test_numpy_gradient_equals_oracle_finite_differences(case):
    sysm, model, nmols, xyz = case
    q, eps, sig = theta0(sysm)
    U, dq, de, ds = model.energy_grad(q, eps, sig, nmols, xyz)
    rng = np.random.default_rng(3)
    h = 1e-5
    for i in rng.choice(len(q), 4, replace=False):
        qp, qm = q.copy(), q.copy()
        qp[i] += h
        qm[i] -= h
        fd = (oracle_energy(sysm, qp, eps, sig, nmols, xyz) - oracle_energy(sysm, qm, eps, sig, nmols, xyz)) / (2 * h)
        assert abs(fd - dq[i]) <= 1e-5 * max(1.0, abs(dq[i])), (i, fd, dq[i])
    fw = np.unique(np.concatenate([sp.types for sp in sysm.framework_species]))
    for t in fw[:2]:
        for which, grad in (("eps", de), ("sig", ds)):
            p = (eps if which == "eps" else sig).copy()
            step = h * max(1.0, abs(p[t - 1]))
            pp, pm = p.copy(), p.copy()
            pp[t - 1] += step
            pm[t - 1] -= step
            args_p = (pp, sig) if which == "eps" else (eps, pp)
            args_m = (pm, sig) if which == "eps" else (eps, pm)
            fd = (oracle_energy(sysm, q, *args_p, nmols, xyz) - oracle_energy(sysm, q, *args_m, nmols, xyz)) / (2 * step)
            assert abs(fd - grad[t - 1]) <= 1e-5 * max(1.0, abs(grad[t - 1])), (which, t, fd, grad[t - 1])


def test_numpy_gradient_equals_own_finite_differences(case):
    sysm, model, nmols, xyz = case
    q, eps, sig = theta0(sysm)
    U, dq, de, ds = model.energy_grad(q, eps, sig, nmols, xyz)
    h = 1e-5
    for i in (0, 17, 211, len(q) - 1):
        qp, qm = q.copy(), q.copy()
        qp[i] += h
        qm[i] -= h
        fd = (model.energy_grad(qp, eps, sig, nmols, xyz)[0] - model.energy_grad(qm, eps, sig, nmols, xyz)[0]) / (2 * h)
        assert abs(fd - dq[i]) <= 1e-6 * max(1.0, abs(dq[i]))
    fw = np.unique(np.concatenate([sp.types for sp in sysm.framework_species]))
    for t in fw:
        for p, grad, isig in ((eps, de, False), (sig, ds, True)):
            step = h * max(1.0, abs(p[t - 1]))
            pp, pm = p.copy(), p.copy()
            pp[t - 1] += step
            pm[t - 1] -= step
            up = model.energy_grad(q, eps if isig else pp, pp if isig else sig, nmols, xyz)[0]
            um = model.energy_grad(q, eps if isig else pm, pm if isig else sig, nmols, xyz)[0]
            fd = (up - um) / (2 * step)
            assert abs(fd - grad[t - 1]) <= 1e-6 * max(1.0, abs(grad[t - 1])), (t, isig, fd, grad[t - 1])
    # guest types carry no theta derivative
    gt = [t for t in range(1, len(eps) + 1) if t not in fw]
    assert np.all(de[np.array(gt) - 1] == 0.0)
