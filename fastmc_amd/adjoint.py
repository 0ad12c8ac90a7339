"""Differentiable framework energy (torch) over the HIP adjoint kernels.

theta = (framework site charges, LJ epsilon and sigma of the atom types);
every LJ pair with a framework type is mixed from the type parameters
(Lorentz-Berthelot).  ``FrameworkEnergy`` returns the single_point total
energy (gcmc.f:2549-2661) of each guest configuration of a batch, in FIELD
units, and its backward pass is the hand-written adjoint of
``fmc_energy_grad`` (include/fastmc.h).  ``dN_dtheta`` is the fluctuation
estimator of the grand-canonical loading:

    d<N>/dtheta = -beta (<N dU/dtheta> - <N> <dU/dtheta>).

torch is used for autograd plumbing only; the arithmetic runs in
libfastmc_hip.so, and there is no CPU fallback.
"""
from __future__ import annotations

import numpy as np
import torch

from .lib import FMC_FP64, FMC_MIXED, Handle
from .system import System, make_walker

KBOLTZ = 1.9872041e-3  # kcal/mol/K, utility_pack.f:121


def theta0(sysm: System):
    """theta at the FIELD values: framework charges in FIELD/CONFIG order and,
    per atom type, the epsilon (internal units) and sigma of its self pair."""
    q = np.concatenate([np.tile(sp.charge, sp.nmols) for sp in sysm.framework_species]).astype(np.float64)
    nt = len(sysm.labels)
    eps, sig = np.zeros(nt), np.ones(nt)
    for (i, j, kind, e, s) in sysm.pairs:
        if i == j and kind == 2:
            eps[i - 1], sig[i - 1] = e, s
    return q, eps, sig


def framework_types(sysm: System) -> np.ndarray:
    """1-based atom types that occur in the framework."""
    return np.unique(np.concatenate([sp.types for sp in sysm.framework_species]))


class FrameworkEnergy(torch.autograd.Function):
    """U[c] = E(theta; configuration c); backward returns dE/dq, dE/deps, dE/dsig."""

    @staticmethod
    def forward(ctx, q, eps, sig, handle: Handle, nmols, xyz, precision):
        U, dq, de, ds = handle.energy_grad(q.detach().cpu().double().numpy(), eps.detach().cpu().double().numpy(),
                                           sig.detach().cpu().double().numpy(), nmols, xyz, precision)
        dev, dt = q.device, torch.float64
        ctx.save_for_backward(torch.as_tensor(dq, dtype=dt, device=dev), torch.as_tensor(de, dtype=dt, device=dev),
                              torch.as_tensor(ds, dtype=dt, device=dev))
        return torch.as_tensor(U, dtype=dt, device=dev)

    @staticmethod
    def backward(ctx, gU):
        dq, de, ds = ctx.saved_tensors
        g = gU.to(torch.float64)
        return g @ dq, g @ de, g @ ds, None, None, None, None


def framework_energy(q, eps, sig, handle: Handle, nmols, xyz, precision: int = FMC_MIXED):
    return FrameworkEnergy.apply(q, eps, sig, handle, nmols, xyz, precision)


def sample_configurations(sysm: System, n: int, every: int = 200, warmup: int = 20000,
                          seed=(21, 43, 65, 87), pressure_pa=None, precision: int = FMC_FP64):
    """n guest configurations from a device GCMC chain at the system's state
    point (or pressure_pa per guest), one every `every` steps after `warmup`.

    Returns (nmols [n][ntpguest] int32, xyz [n][gstride][3]) with gstride the
    largest number of guest sites seen."""
    press = None if pressure_pa is None else list(pressure_pa)
    h = Handle(sysm, walkers=[make_walker(sysm, seed=seed, pressures=press, keep=[])], precision=precision)
    try:
        h.run(warmup)
        ng = len(sysm.guests)
        snaps = []
        for _ in range(n):
            h.run(every)
            snaps.append([h.guests(g) for g in range(1, ng + 1)])
    finally:
        h.close()
    gstride = max(1, max(sum(len(a) for a in s) for s in snaps))
    nmols = np.zeros((n, ng), dtype=np.int32)
    xyz = np.zeros((n, gstride, 3))
    for c, s in enumerate(snaps):
        o = 0
        for g, a in enumerate(s):
            nmols[c, g] = len(a) // sysm.guests[g].natoms
            xyz[c, o: o + len(a)] = a
            o += len(a)
    return nmols, xyz


def dN_dtheta(N: np.ndarray, dU: np.ndarray, temperature: float) -> np.ndarray:
    """Fluctuation estimator -beta (<N dU/dtheta> - <N><dU/dtheta>) over samples
    (N [n], dU [n][p] in kcal/mol)."""
    beta = 1.0 / (KBOLTZ * temperature)
    N = np.asarray(N, dtype=np.float64)
    dU = np.asarray(dU, dtype=np.float64)
    return -beta * ((N[:, None] * dU).mean(0) - N.mean() * dU.mean(0))
