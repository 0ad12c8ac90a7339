"""Deterministic synthetic CONFIG / FIELD / CONTROL writers for the GCMC workloads.

The files follow the reference input grammar exactly, so the same directory
can be fed to the reference executable and to this package:

* FIELD  -- ``readinputs.f:11-307`` (``UNITS kcal``, guests first as
  ``&guest`` blocks with COM-centred template coordinates, framework last,
  ``VDW n`` pair table, ``close``).  Pairs are written explicitly with
  Lorentz-Berthelot mixing (``vdw_terms.f:35-142`` reads them verbatim).
* CONFIG -- ``readinputs.f:309-513`` (title, ``levcfg imcon natoms``, three
  cell rows, then per atom a name line and a ``3f20`` coordinate line).
* CONTROL -- ``readinputs.f:515-944`` (keyword lines; per-guest ``&guest``
  blocks carrying the pressure and the move frequencies).

Framework rule (the synthetic net is described as 424 atoms on a simple-cubic
strut net with nodes every 12.9 A and atoms every 1.4 A along the edges; a
single-strand strut with 1.4 A spacing gives only 224 atoms, so the rule used
here is the one that gives exactly 424):

* the a = 25.8 A cell holds 8 nodes on a 12.9 A simple-cubic lattice;
* each node is a 5-atom cluster: the node point plus 4 atoms 1.0 A away along
  the tetrahedral directions (+1,+1,+1), (+1,-1,-1), (-1,+1,-1), (-1,-1,+1);
* each of the 24 edges (+x, +y, +z from every node) is a two-strand ladder:
  8 atoms per strand spaced 1.4 A along the edge and centred on the edge
  midpoint (t = 6.45 + 1.4 (k - 4.5), k = 1..8), the strands offset by
  +-0.7 A along the next axis (x-edge -> y, y-edge -> z, z-edge -> x);
* 8 x 5 + 24 x 16 = 424 atoms; every coordinate receives N(0, 0.05 A) jitter.

Types (4) are drawn uniformly per atom, epsilon ~ U(0.03, 0.2) kcal/mol and
sigma ~ U(2.5, 3.5) A per type, charges ~ U(-0.8, 0.8) shifted to net zero.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

KBOLTZ_KCAL = 1.9872041e-3  # kcal/mol/K, utility_pack.f:121

DEFAULT_SEED = 20261203


@dataclass
class GuestSpec:
    name: str
    sites: list  # (label, mass, charge, x, y, z)
    lj: dict  # label -> (eps kcal/mol, sigma A)


def co2_guest() -> GuestSpec:
    """Rigid 3-site CO2: C-O 1.16 A, q_C=+0.70, q_O=-0.35, C 27 K/2.80 A, O 79 K/3.05 A."""
    return GuestSpec(
        name="CO2",
        sites=[("Cx", 12.0107, 0.70, 0.0, 0.0, 0.0),
               ("Ox", 15.9994, -0.35, 1.16, 0.0, 0.0),
               ("Ox", 15.9994, -0.35, -1.16, 0.0, 0.0)],
        lj={"Cx": (27.0 * KBOLTZ_KCAL, 2.80), "Ox": (79.0 * KBOLTZ_KCAL, 3.05)},
    )


def n2_guest() -> GuestSpec:
    """Three-site TraPPE-like N2: N-N 1.10 A, q_N=-0.482, q_COM=+0.964, N 36 K/3.31 A, COM eps=0."""
    return GuestSpec(
        name="N2",
        sites=[("Nx", 14.0067, -0.482, -0.55, 0.0, 0.0),
               ("Nx", 14.0067, -0.482, 0.55, 0.0, 0.0),
               ("COM", 0.0, 0.964, 0.0, 0.0, 0.0)],
        lj={"Nx": (36.0 * KBOLTZ_KCAL, 3.31), "COM": (0.0, 0.0)},
    )


@dataclass
class Framework:
    cell: np.ndarray            # (9,) row vectors a, b, c (CONFIG order)
    names: list                 # per atom type label
    xyz: np.ndarray             # (n, 3)
    charges: np.ndarray         # (n,)
    type_lj: dict = field(default_factory=dict)  # label -> (eps, sigma)
    mass: float = 50.0


FW_LABELS = ["Fa", "Fb", "Fc", "Fd"]


def strut_net_points(a: float = 25.8, node: float = 12.9) -> np.ndarray:
    """Undisplaced 424-point strut net of one 25.8 A cell (see module docstring)."""
    pts = []
    nn = int(round(a / node))
    tet = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], dtype=np.float64) / math.sqrt(3.0)
    for i in range(nn):
        for j in range(nn):
            for k in range(nn):
                c = np.array([i, j, k], dtype=np.float64) * node
                pts.append(c.copy())
                for d in tet:
                    pts.append(c + 1.0 * d)
                for ax in range(3):
                    off_ax = (ax + 1) % 3
                    for s in (-0.7, 0.7):
                        for kk in range(1, 9):
                            t = 0.5 * node + 1.4 * (kk - 4.5)
                            p = c.copy()
                            p[ax] += t
                            p[off_ax] += s
                            pts.append(p)
    return np.asarray(pts)


def make_framework(seed: int = DEFAULT_SEED, jitter_seed: int | None = None,
                   a: float = 25.8) -> Framework:
    """Config-0 framework: 424 jittered strut-net atoms, 4 LJ types, net-neutral charges.

    ``seed`` fixes the chemistry (types, charges, per-type LJ); ``jitter_seed``
    (default: ``seed``) fixes the Gaussian position jitter, so batched state
    points can share chemistry while each gets its own jittered copy.
    """
    rng = np.random.default_rng(seed)
    base = strut_net_points(a=a)
    n = base.shape[0]
    types = rng.integers(0, 4, size=n)
    eps = rng.uniform(0.03, 0.2, size=4)
    sig = rng.uniform(2.5, 3.5, size=4)
    q = rng.uniform(-0.8, 0.8, size=n)
    q = q - q.mean()
    # round to the printed precision first, then re-neutralise on the last atom
    q = np.round(q, 10)
    q[-1] = -np.sum(q[:-1])
    q = np.round(q, 10)
    jrng = np.random.default_rng(seed if jitter_seed is None else jitter_seed + 7919)
    xyz = base + jrng.normal(0.0, 0.05, size=base.shape)
    xyz = np.mod(xyz, a)
    cell = np.array([a, 0, 0, 0, a, 0, 0, 0, a], dtype=np.float64)
    names = [FW_LABELS[t] for t in types]
    return Framework(cell=cell, names=names, xyz=xyz, charges=q,
                     type_lj={FW_LABELS[i]: (float(eps[i]), float(sig[i])) for i in range(4)})


def supercell(fw: Framework, reps=(2, 2, 2)) -> Framework:
    """Replicate a framework (cubic/orthorhombic cells only)."""
    a, b, c = fw.cell[0], fw.cell[4], fw.cell[8]
    xs, names, qs = [], [], []
    for i in range(reps[0]):
        for j in range(reps[1]):
            for k in range(reps[2]):
                xs.append(fw.xyz + np.array([i * a, j * b, k * c]))
                names.extend(fw.names)
                qs.append(fw.charges)
    cell = np.array([a * reps[0], 0, 0, 0, b * reps[1], 0, 0, 0, c * reps[2]], dtype=np.float64)
    return Framework(cell=cell, names=names, xyz=np.concatenate(xs), charges=np.concatenate(qs),
                     type_lj=dict(fw.type_lj), mass=fw.mass)


def _fmt(x: float) -> str:
    return f"{x:.10f}"


def write_field(path: str, fw: Framework, guests: list, title: str = "synthetic strut net") -> None:
    """FIELD in the reference grammar: guests first (readinputs.f:127-131), framework last."""
    lines = [title, "UNITS   kcal", f"molecular types {len(guests) + 1}"]
    for g in guests:
        lines.append(f"&guest {g.name}")
        lines.append("NUMMOLS 0")
        lines.append(f"ATOMS {len(g.sites)}")
        for (lab, m, qq, x, y, z) in g.sites:
            lines.append(f"{lab:<8s}{m:14.6f}{qq:12.6f}{x:12.6f}{y:12.6f}{z:12.6f}")
        lines.append("rigid 1")
        lines.append(f"{len(g.sites)} " + " ".join(str(i + 1) for i in range(len(g.sites))))
        lines.append("finish")
    lines.append("Framework")
    lines.append("NUMMOLS 1")
    lines.append(f"ATOMS {len(fw.names)}")
    for lab, qq in zip(fw.names, fw.charges):
        lines.append(f"{lab:<8s}{fw.mass:14.6f}{qq:20.10f}      1      1")
    lines.append("finish")
    # unique labels in first-appearance order (guest sites first), as readfield builds unqatm
    labels = []
    lj = {}
    for g in guests:
        for s in g.sites:
            if s[0] not in labels:
                labels.append(s[0])
        lj.update(g.lj)
    for lab in fw.names:
        if lab not in labels:
            labels.append(lab)
    lj.update(fw.type_lj)
    pairs = []
    for i, li in enumerate(labels):
        for lj_ in labels[i:]:
            e1, s1 = lj[li]
            e2, s2 = lj[lj_]
            eps = math.sqrt(e1 * e2)
            sig = 0.5 * (s1 + s2) if (e1 > 0 and e2 > 0) else 0.0
            pairs.append(f"{li:<8s}{lj_:<8s}lj {eps:.8f} {sig:.8f}")
    lines.append(f"VDW {len(pairs)}")
    lines.extend(pairs)
    lines.append("close")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_config(path: str, fw: Framework, title: str = "synthetic strut net", imcon: int = 1) -> None:
    """CONFIG in DL_POLY levcfg 0 format (readinputs.f:331-415; coordinates read as 3f20.0)."""
    c = fw.cell
    lines = [title, f"{0:10d}{imcon:10d}{len(fw.names):10d}"]
    for r in range(3):
        lines.append(f"{c[3*r]:20.12f}{c[3*r+1]:20.12f}{c[3*r+2]:20.12f}")
    for i, (lab, p) in enumerate(zip(fw.names, fw.xyz)):
        lines.append(f"{lab:<8s}{i+1:10d}")
        lines.append(f"{p[0]:20.12f}{p[1]:20.12f}{p[2]:20.12f}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_control(path: str, temperature: float, pressures_bar: list, moves: dict,
                  steps: int = 100000, equilibration: int = 100000, cutoff: float = 12.5,
                  delr: float = 1.0, precision: str = "1d-6", numguests: int = 1000,
                  extra: list | None = None) -> None:
    """CONTROL with one ``&guest`` block per guest type (readinputs.f:615-706)."""
    lines = ["GCMC Run",
             f"temperature  {temperature:.6f}",
             f"steps    {steps}",
             f"equilibration    {equilibration}",
             f"cutoff          {cutoff:.6f} angstrom",
             f"delr            {delr:.1f} angstrom",
             f"ewald precision  {precision}",
             f"numguests {numguests}",
             "history 0"]
    for ig, p in enumerate(pressures_bar):
        lines.append(f"&guest {ig + 1}")
        lines.append(f"  pressure (bar) {p:.8f}")
        for mv, fr in moves.items():
            lines.append(f"  move {mv} {fr:.6f}")
        lines.append("&end")
    if extra:
        lines.extend(extra)
    lines.append("finish")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


# every move written explicitly: readcontrol leaves the frequencies of moves a
# CONTROL file omits to uninitialised locals (readinputs.f:539-543,830-885)
GCMC_MOVES = {"insert": 0.25, "delete": 0.25, "displace": 0.25, "jump": 0.0,
              "swap": 0.0, "translate": 0.0, "rotate": 0.25, "switch": 0.0}


def write_system(dirpath: str, fw: Framework, guests: list, temperature: float,
                 pressures_bar: list, moves: dict | None = None, **ctl) -> str:
    os.makedirs(dirpath, exist_ok=True)
    write_field(os.path.join(dirpath, "FIELD"), fw, guests)
    write_config(os.path.join(dirpath, "CONFIG"), fw)
    write_control(os.path.join(dirpath, "CONTROL"), temperature, pressures_bar,
                  GCMC_MOVES if moves is None else moves, **ctl)
    return dirpath


def config0(dirpath: str, pressure_bar: float = 0.1, temperature: float = 293.0,
            seed: int = DEFAULT_SEED, **ctl) -> str:
    """Config 0: 424-atom cubic framework, CO2, T=293 K, P=0.1 bar, ins/del/dis/rot 1/4 each."""
    return write_system(dirpath, make_framework(seed), [co2_guest()], temperature, [pressure_bar], **ctl)


def config1(dirpath: str, pressure_bar: float = 10.0, temperature: float = 293.0,
            seed: int = DEFAULT_SEED, **ctl) -> str:
    """Config 1: 2x2x2 supercell of config 0 (3392 atoms, a = 51.6 A), CO2 at 10 bar."""
    return write_system(dirpath, supercell(make_framework(seed)), [co2_guest()], temperature,
                        [pressure_bar], **ctl)


def config2_state_points(n: int = 64):
    """Config 2 grid: T in {273, 293, 323} K split 22/21/21, P log-spaced 1e-3..10 bar per T."""
    temps = [273.0] * 22 + [293.0] * 21 + [323.0] * 21
    temps = temps[:n] if n <= 64 else (temps * ((n + 63) // 64))[:n]
    pts = []
    counts = {}
    for t in temps:
        counts[t] = counts.get(t, 0) + 1
    idx = {t: 0 for t in counts}
    for t in temps:
        m = counts[t]
        k = idx[t]
        idx[t] += 1
        p = 10.0 ** (-3.0 + 4.0 * (k / max(m - 1, 1)))
        pts.append((t, p))
    return pts


def config2(rootdir: str, n: int = 64, seed: int = DEFAULT_SEED, **ctl) -> list:
    """Config 2: n state points, each with its own seed-jittered copy of the config-0 framework."""
    dirs = []
    for i, (t, p) in enumerate(config2_state_points(n)):
        fw = make_framework(seed, jitter_seed=seed + 1000 + i)
        d = os.path.join(rootdir, f"sp{i:03d}")
        write_system(d, fw, [co2_guest()], t, [p], **ctl)
        dirs.append(d)
    return dirs


def config4(dirpath: str, temperature: float = 293.0, seed: int = DEFAULT_SEED,
            pressures_bar=(0.2, 0.8), **ctl) -> str:
    """Config 4 chemistry: config-1 framework with CO2 + N2 (binary mixture)."""
    return write_system(dirpath, supercell(make_framework(seed)), [co2_guest(), n2_guest()],
                        temperature, list(pressures_bar), **ctl)
