"""GPU energy gradient with respect to theta (framework charges, LJ type
parameters): fmc_energy_grad against the FP64 numpy oracle (itself pinned
to the reference-faithful C oracle on the CPU), central finite differences
(h = 1e-5) of its own FP64 energy, the mixed (FP32-product) mode, the
torch.autograd.Function and the fluctuation estimator of d<N>/dtheta on
configurations sampled at 1 bar."""
import sys

import numpy as np
import pytest
import torch

from oracle.adjoint_ref import EnergyModel
from test_adjoint_ref import _system

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup():
    from fastmc_amd import adjoint as A
    from fastmc_amd.lib import FMC_FP64, Handle
    sysm = _system(1.0)
    say("sampling 24 configurations at 1 bar on the device")
    nmols, xyz = A.sample_configurations(sysm, 24, every=300, warmup=20000)
    h = Handle(sysm, precision=FMC_FP64)
    yield sysm, h, nmols, xyz
    h.close()


def test_gpu_energy_grad_matches_numpy_fp64(setup):
    from fastmc_amd import adjoint as A
    from fastmc_amd.lib import FMC_FP64
    sysm, h, nmols, xyz = setup
    q, eps, sig = A.theta0(sysm)
    U, dq, de, ds = h.energy_grad(q, eps, sig, nmols, xyz, FMC_FP64)
    model = EnergyModel(sysm)
    for c in (0, 7, 23):
        Ur, dqr, der, dsr = model.energy_grad(q, eps, sig, nmols[c], xyz[c])
        assert abs(U[c] - Ur) <= 1e-10 * max(1.0, abs(Ur)), (c, U[c], Ur)
        scale = np.abs(dqr).max()
        assert np.abs(dq[c] - dqr).max() <= 1e-10 * scale
        np.testing.assert_allclose(de[c], der, rtol=1e-10, atol=1e-12 * np.abs(der).max())
        np.testing.assert_allclose(ds[c], dsr, rtol=1e-10, atol=1e-12 * np.abs(dsr).max())
    say("N per configuration", nmols[:, 0].tolist())


def test_gpu_energy_grad_central_differences(setup):
    """dU/dtheta against central differences of the device's own FP64 U, h = 1e-5
    (relative for the LJ parameters), for every theta entry of one configuration."""
    from fastmc_amd import adjoint as A
    from fastmc_amd.lib import FMC_FP64
    sysm, h, nmols, xyz = setup
    q, eps, sig = A.theta0(sysm)
    c = int(np.argmax(nmols[:, 0]))
    n1, x1 = nmols[c: c + 1], xyz[c: c + 1]
    U, dq, de, ds = h.energy_grad(q, eps, sig, n1, x1, FMC_FP64)
    f = lambda qq, ee, ss: h.energy_grad(qq, ee, ss, n1, x1, FMC_FP64, want_grad=False)[0][0]
    hq = 1e-5
    fd = np.zeros(len(q))
    for i in range(len(q)):
        qp, qm = q.copy(), q.copy()
        qp[i] += hq
        qm[i] -= hq
        fd[i] = (f(qp, eps, sig) - f(qm, eps, sig)) / (2 * hq)
    err = np.abs(fd - dq[0]).max() / np.abs(dq[0]).max()
    say("charges: max |FD - adjoint| / max |grad| =", err)
    assert err <= 1e-6
    for t in A.framework_types(sysm):
        for which in ("eps", "sig"):
            p = (eps if which == "eps" else sig).copy()
            st = 1e-5 * abs(p[t - 1])
            pp, pm = p.copy(), p.copy()
            pp[t - 1] += st
            pm[t - 1] -= st
            up = f(q, pp, sig) if which == "eps" else f(q, eps, pp)
            um = f(q, pm, sig) if which == "eps" else f(q, eps, pm)
            g = (de if which == "eps" else ds)[0][t - 1]
            assert abs((up - um) / (2 * st) - g) <= 1e-6 * max(1.0, abs(g)), (which, t)


def test_gpu_energy_grad_mixed_precision(setup):
    from fastmc_amd import adjoint as A
    from fastmc_amd.lib import FMC_FP64, FMC_MIXED
    sysm, h, nmols, xyz = setup
    q, eps, sig = A.theta0(sysm)
    U, dq, de, ds = h.energy_grad(q, eps, sig, nmols, xyz, FMC_FP64)
    Um, dqm, dem, dsm = h.energy_grad(q, eps, sig, nmols, xyz, FMC_MIXED)
    # the energy, LJ and tail terms are FP64 in both modes (equal up to the
    # order of the atomic additions across blocks)
    np.testing.assert_allclose(U, Um, rtol=1e-14)
    rel = np.abs(dqm - dq).max() / np.abs(dq).max()
    say("mixed: max |dU/dq error| / max |dU/dq| =", rel)
    assert rel <= 1e-5
    np.testing.assert_allclose(de, dem, rtol=1e-13, atol=1e-13 * np.abs(de).max())
    np.testing.assert_allclose(ds, dsm, rtol=1e-13, atol=1e-13 * np.abs(ds).max())


def test_gpu_autograd_function(setup):
    from fastmc_amd import adjoint as A
    from fastmc_amd.lib import FMC_FP64
    sysm, h, nmols, xyz = setup
    q0, e0, s0 = A.theta0(sysm)
    q = torch.tensor(q0, dtype=torch.float64, requires_grad=True)
    e = torch.tensor(e0, dtype=torch.float64, requires_grad=True)
    s = torch.tensor(s0, dtype=torch.float64, requires_grad=True)
    U = A.framework_energy(q, e, s, h, nmols[:6], xyz[:6], FMC_FP64)
    w = torch.linspace(0.5, 1.5, 6, dtype=torch.float64)
    (w * U).sum().backward()
    _, dq, de, ds = h.energy_grad(q0, e0, s0, nmols[:6], xyz[:6], FMC_FP64)
    np.testing.assert_allclose(q.grad.numpy(), w.numpy() @ dq, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(e.grad.numpy(), w.numpy() @ de, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(s.grad.numpy(), w.numpy() @ ds, rtol=1e-12, atol=1e-12)
    # torch's own finite-difference check on a few charges and one epsilon
    idx = torch.tensor([3, 50, 200])
    fw = int(A.framework_types(sysm)[0]) - 1

    def f(qs, es):
        qq = q.detach().clone()
        qq[idx] = qs
        ee = e.detach().clone()
        ee[fw] = es
        return A.framework_energy(qq, ee, s.detach(), h, nmols[:2], xyz[:2], FMC_FP64)

    qs = q.detach()[idx].clone().requires_grad_(True)
    es = e.detach()[fw].clone().requires_grad_(True)
    assert torch.autograd.gradcheck(f, (qs, es), eps=1e-6, atol=1e-5, rtol=1e-6)


def test_gpu_fluctuation_estimator(setup):
    """d<N>/dtheta = -beta (<N dU/dtheta> - <N><dU/dtheta>) over the sampled
    configurations, FP64 against mixed, and against the formula in numpy."""
    from fastmc_amd import adjoint as A
    from fastmc_amd.lib import FMC_FP64, FMC_MIXED
    sysm, h, nmols, xyz = setup
    q, eps, sig = A.theta0(sysm)
    _, dq, de, ds = h.energy_grad(q, eps, sig, nmols, xyz, FMC_FP64)
    _, dqm, _, _ = h.energy_grad(q, eps, sig, nmols, xyz, FMC_MIXED)
    N = nmols[:, 0].astype(float)
    g = A.dN_dtheta(N, np.concatenate([dq, de, ds], axis=1), sysm.temperature)
    beta = 1.0 / (1.9872041e-3 * sysm.temperature)
    G = np.concatenate([dq, de, ds], axis=1)
    ref = -beta * (np.mean(N[:, None] * G, axis=0) - np.mean(N) * np.mean(G, axis=0))
    np.testing.assert_allclose(g, ref, rtol=1e-12, atol=1e-15)
    gm = A.dN_dtheta(N, dqm, sysm.temperature)
    rel = np.abs(gm - g[: len(q)]).max() / max(np.abs(g[: len(q)]).max(), 1e-300)
    say("d<N>/dq: mixed vs FP64 max rel", rel, "| <N> =", N.mean())
    assert rel <= 1e-4


def test_gpu_energy_grad_config3_batch():
    """The config-3 batch the bench times: 256 configurations sampled every 200
    steps after 20 000 at 1 bar, one fmc_energy_grad call.  Six of them spread
    over the batch against the FP64 numpy oracle; the batch call equals
    one-configuration calls (the blocked contraction and the k-range partial
    sums do not mix configurations); mixed mode over all 256 within the
    bounds of the 24-configuration test."""
    from fastmc_amd import adjoint as A
    from fastmc_amd.lib import FMC_FP64, FMC_MIXED, Handle
    sysm = _system(1.0)
    say("sampling 256 configurations at 1 bar on the device")
    nmols, xyz = A.sample_configurations(sysm, 256, every=200, warmup=20000)
    say("N range", int(nmols[:, 0].min()), int(nmols[:, 0].max()))
    h = Handle(sysm, precision=FMC_FP64)
    try:
        q, eps, sig = A.theta0(sysm)
        U, dq, de, ds = h.energy_grad(q, eps, sig, nmols, xyz, FMC_FP64)
        model = EnergyModel(sysm)
        for c in (0, 51, 102, 153, 204, 255):
            Ur, dqr, der, dsr = model.energy_grad(q, eps, sig, nmols[c], xyz[c])
            assert abs(U[c] - Ur) <= 1e-10 * max(1.0, abs(Ur)), (c, U[c], Ur)
            assert np.abs(dq[c] - dqr).max() <= 1e-10 * np.abs(dqr).max(), c
            np.testing.assert_allclose(de[c], der, rtol=1e-10, atol=1e-12 * np.abs(der).max())
            np.testing.assert_allclose(ds[c], dsr, rtol=1e-10, atol=1e-12 * np.abs(dsr).max())
        for c in (3, 130, 254):
            U1, dq1, de1, ds1 = h.energy_grad(q, eps, sig, nmols[c: c + 1], xyz[c: c + 1], FMC_FP64)
            assert abs(U1[0] - U[c]) <= 1e-12 * max(1.0, abs(U[c])), c
            np.testing.assert_allclose(dq1[0], dq[c], rtol=0, atol=1e-12 * np.abs(dq[c]).max())
            np.testing.assert_allclose(de1[0], de[c], rtol=1e-12, atol=1e-14 * np.abs(de[c]).max())
            np.testing.assert_allclose(ds1[0], ds[c], rtol=1e-12, atol=1e-14 * np.abs(ds[c]).max())
        Um, dqm, dem, dsm = h.energy_grad(q, eps, sig, nmols, xyz, FMC_MIXED)
        np.testing.assert_allclose(U, Um, rtol=1e-14)
        rel = np.abs(dqm - dq).max() / np.abs(dq).max()
        say("256 configurations, mixed: max |dU/dq error| / max |dU/dq| =", rel)
        assert rel <= 1e-5
        np.testing.assert_allclose(de, dem, rtol=1e-13, atol=1e-13 * np.abs(de).max())
        np.testing.assert_allclose(ds, dsm, rtol=1e-13, atol=1e-13 * np.abs(ds).max())
    finally:
        h.close()
