"""Reader for the reference FIELD / CONFIG / CONTROL inputs and the ctypes
mirror of ``include/fastmc.h``.

The reader restates the reference parsers so that every number reaches the
library bit-identical to what the reference holds in its module globals:

* ``dblstr`` / ``intstr`` -- parse_module.f:81-256 (digit-by-digit
  accumulation, ``d``/``e`` exponents, sign taken from the character before
  the first digit);
* ``initscan`` -- utility_pack.f:207-508 (cutoff, delr, Ewald alpha / kmax from
  ``ewald precision``, vdw cutoff, averaging windows);
* ``readfield`` -- readinputs.f:11-307 (units, molecule types, guest templates
  centred with the reference's doubled-mass COM quirk, unique atom labels in
  first-appearance order, VDW table);
* ``readconfig`` -- readinputs.f:309-513 (cell rows, ``3f20.0`` coordinates);
* ``readcontrol`` -- readinputs.f:515-944 (temperature, per-guest pressure,
  move frequencies, delr, overlap) and the REAL(4) move-frequency
  normalisation of gcmc.f:341-366.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field

import numpy as np

PI = 3.141592653589793  # utility_pack.f:95
BOLTZ = 1.3806503e-23   # utility_pack.f:124
KBOLTZ = 1.9872041e-3   # utility_pack.f:121

FMC_MAX_SPECIES = 8
FMC_MAX_SITES = 8
FMC_MAX_TYPES = 24
FMC_NMOVES = 8

MOVE_KEYS = ("ins", "del", "dis", "jum", "swa", "swi", "tra", "rot")
# storage order of the cumulative table (gcmc.f:359-366): ins del dis jmp swp tra rot swi
MOVE_ORDER = ("ins", "del", "dis", "jum", "swa", "tra", "rot", "swi")


# ------------------------------------------------------------- number grammar
def intstr(word: str):
    """intstr (parse_module.f:81-146): first integer in ``word``; returns (value, rest)."""
    n = len(word)
    lst = 0
    isn = 1
    ksn = "+"
    val = 0
    count = False
    final = False
    while lst < n and not final:
        ch = word[lst]
        lst += 1
        flag = False
        if ch.isdigit():
            val = 10 * val + int(ch)
            count = True
            flag = True
        if count and not flag:
            final = True
        if flag and ksn == "-":
            isn = -1
        ksn = ch
    return isn * val, word[lst - 1:] if final else word[lst:]


def dblstr(word: str):
    """dblstr (parse_module.f:148-256): first real in ``word``; returns (value, rest).

    The value is built exactly as the reference builds it: ``v = ten*v + one*d``
    with ``ten`` dropping to 1 and ``one`` divided by 10 per character after the
    decimal point, then multiplied by ``10.d0**iexp``.
    """
    n = len(word)
    lst = 0
    sn = 1.0
    ksn = "+"
    ten = 10.0
    one = 1.0
    val = 0.0
    iexp = 0
    start = False
    ldot = False
    final = False
    exp_rest = None
    while lst < n and not final:
        ch = word[lst]
        lst += 1
        flag = False
        if ch.isdigit():
            val = ten * val + one * float(int(ch))
            flag = True
            start = True
        if ch == ".":
            flag = True
            ten = 1.0
            ldot = True
            start = True
        if flag and ksn == "-":
            sn = -1.0
        if ldot:
            one = one / 10.0
        ksn = ch
        if ksn == "D":
            ksn = "d"
        if ksn == "E":
            ksn = "e"
        if start:
            if ksn in ("d", "e"):
                e, exp_rest = intstr(word[lst:])
                iexp = e
                final = True
            if not flag:
                final = True
    value = sn * val * _pow10(iexp)
    if exp_rest is not None:
        return value, exp_rest
    return value, word[lst - 1:] if final else ""


def _pow10(k: int) -> float:
    """10.d0**k as Fortran evaluates a real base to an integer power."""
    if k == 0:
        return 1.0
    if k > 0:
        return _ipow(10.0, k)
    return 1.0 / _ipow(10.0, -k)


def _ipow(x: float, k: int) -> float:
    r = 1.0
    b = x
    while k:
        if k & 1:
            r *= b
        b *= b
        k >>= 1
    return r


def _lc(s: str) -> str:
    return s.lower()


# --------------------------------------------------------------- data model
@dataclass
class Species:
    name: str
    is_guest: bool
    nmols: int
    labels: list
    mass: np.ndarray
    charge: np.ndarray
    frozen: np.ndarray
    types: np.ndarray = None      # 1-based atom-type index per site
    template: np.ndarray = None   # guests: (natoms, 3) COM-centred template
    xyz: np.ndarray = None        # (nmols*natoms, 3) CONFIG coordinates
    molmass: float = 0.0

    @property
    def natoms(self) -> int:
        return len(self.labels)


@dataclass
class System:
    title: str
    engunit: float
    species: list
    labels: list                 # unique atom labels (unqatm)
    pairs: list                  # (i, j, kind, eps_internal, sigma)
    imcon: int
    cell: np.ndarray
    levcfg: int
    sumchg: float
    # CONTROL / initscan
    temperature: float = 0.0
    rcut: float = 0.0
    rvdw: float = 0.0
    delr: float = 1.0
    alpha: float = 0.0
    kmax: tuple = (0, 0, 0)
    eps_precision: float = 0.0
    overlap: float = 1.5
    surftol: float = 0.0
    steps: int = 1
    eqsteps: int = 0
    nwind: int = 0                  # 'averaging window' (0: not given, initscan default)
    accsample: bool = False         # 'sampling': count accepted steps only
    nnumg: int = 0                  # 'numguests' interval (0: not given, readcontrol default)
    griddim: float = 0.0            # 'grid spacing' (A) of probability plots / Widom grids
    gridfactor: tuple = (1, 1, 1)   # 'grid factor': unit cells per supercell axis
    prob_plots: list = field(default_factory=list)  # per guest: site lists (0 = COM) of 'probability'
    prob_energy: list = field(default_factory=list) # per guest: an 'e' (energy) plot was requested
    widom: bool = False             # 'henry': Widom insertions on the grid instead of GCMC
    nwidstep: int = 0               # 'widom insertions': orientations per grid point
    pressures: list = field(default_factory=list)   # Pa per guest
    moves: list = field(default_factory=list)       # per guest dict key -> fraction
    guest_delr: list = field(default_factory=list)

    @property
    def guests(self):
        return [s for s in self.species if s.is_guest]

    @property
    def framework_species(self):
        return [s for s in self.species if not s.is_guest]

    def framework_xyz(self) -> np.ndarray:
        return np.concatenate([s.xyz for s in self.framework_species], axis=0)

    def n_framework(self) -> int:
        return sum(s.natoms * s.nmols for s in self.framework_species)

    def volume(self) -> float:
        return dcell(self.cell)[9]

    def move_cum(self, ig: int) -> np.ndarray:
        """REAL(4) cumulative move table of guest ig (gcmc.f:341-366)."""
        f = np.float32
        mv = self.moves[ig]
        vals = [f(mv.get(k, 0.0)) for k in MOVE_ORDER]
        norm = f(0.0)
        for v in vals:
            norm = f(norm + v)
        if norm == f(0.0):
            norm = f(1.0)
            vals = [f(1.0 / 3.0), f(1.0 / 3.0), f(1.0 / 3.0)] + [f(0.0)] * 5
        out = np.zeros(FMC_NMOVES, dtype=np.float32)
        acc = f(vals[0] / norm)
        out[0] = acc
        for i in range(1, FMC_NMOVES):
            acc = f(acc + f(vals[i] / norm))
            out[i] = acc
        return out


def dcell(a) -> np.ndarray:
    """dcell (utility_pack.f:1904-1960), 0-based output bbb[0..9]."""
    a = [0.0] + list(map(float, a))
    b = [0.0] * 11
    b[1] = math.sqrt(a[1] * a[1] + a[2] * a[2] + a[3] * a[3])
    b[2] = math.sqrt(a[4] * a[4] + a[5] * a[5] + a[6] * a[6])
    b[3] = math.sqrt(a[7] * a[7] + a[8] * a[8] + a[9] * a[9])
    b[4] = (a[1] * a[4] + a[2] * a[5] + a[3] * a[6]) / (b[1] * b[2])
    b[5] = (a[1] * a[7] + a[2] * a[8] + a[3] * a[9]) / (b[1] * b[3])
    b[6] = (a[4] * a[7] + a[5] * a[8] + a[6] * a[9]) / (b[2] * b[3])
    axb1 = a[2] * a[6] - a[3] * a[5]
    axb2 = a[3] * a[4] - a[1] * a[6]
    axb3 = a[1] * a[5] - a[2] * a[4]
    bxc1 = a[5] * a[9] - a[6] * a[8]
    bxc2 = a[6] * a[7] - a[4] * a[9]
    bxc3 = a[4] * a[8] - a[5] * a[7]
    cxa1 = a[8] * a[3] - a[2] * a[9]
    cxa2 = a[1] * a[9] - a[3] * a[7]
    cxa3 = a[2] * a[7] - a[1] * a[8]
    b[10] = abs(a[1] * bxc1 + a[2] * bxc2 + a[3] * bxc3)
    b[7] = b[10] / math.sqrt(bxc1 * bxc1 + bxc2 * bxc2 + bxc3 * bxc3)
    b[8] = b[10] / math.sqrt(cxa1 * cxa1 + cxa2 * cxa2 + cxa3 * cxa3)
    b[9] = b[10] / math.sqrt(axb1 * axb1 + axb2 * axb2 + axb3 * axb3)
    return np.array(b[1:], dtype=np.float64)


class InputError(RuntimeError):
    """Raised where the reference would call error(idnode, code)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"error {code}: {msg}")
        self.code = code


# ------------------------------------------------------------------- FIELD
def read_field(path: str):
    with open(path) as f:
        lines = f.read().split("\n")
    it = iter(lines[1:])
    title = lines[0].strip()
    engunit = 1.0
    species = []
    labels = []
    pairs = []
    for raw in it:
        rec = _lc(raw).strip()
        if "units" in rec:
            rest = rec[rec.index("units") + 5:].strip()
            if rest == "":
                engunit = 1.0
            elif "ev" in rec:
                engunit = 9648.530821
            elif "kcal" in rec:
                engunit = 418.4
            elif "kj" in rec:
                engunit = 1.0e2
            elif "k" in rec:
                engunit = BOLTZ  # readinputs.f:82-85 (sic)
        elif "vdw" in rec:
            npair, _ = intstr(rec)
            for _ in range(npair):
                line = next(it)
                words = line.split()
                a1, a2 = words[0][:8], words[1][:8]
                rest = " ".join(words[2:])
                key = _lc(rest)[:4].strip()
                kind = {"12-6": 1, "lj": 2, "nm": 3, "buck": 4, "bhm": 5, "hbnd": 6, "snm": 7,
                        "hcnm": 7, "mors": 8, "wca": 9, "tab": 0}.get(key.split()[0] if key else "", -1)
                tail = rest.split(None, 1)[1] if len(rest.split(None, 1)) > 1 else ""
                p1, tail = dblstr(tail)
                p2, tail = dblstr(tail)
                i = labels.index(a1) + 1 if a1 in labels else 0
                j = labels.index(a2) + 1 if a2 in labels else 0
                p1 = p1 * engunit
                if kind in (1, 6):
                    p2 = p2 * engunit
                pairs.append((i, j, kind, p1, p2))
        elif "molecu" in rec:
            nmt, _ = intstr(rec)
            for _ in range(nmt):
                name_line = next(it)
                is_guest = "&guest" in _lc(name_line)
                nummols = 0
                sp = None
                while True:
                    r = _lc(next(it)).strip()
                    if "nummol" in r:
                        nummols, _ = intstr(r)
                    elif "atoms" in r:
                        nsit, _ = intstr(r)
                        labs, mass, chg, frz, tmpl = [], [], [], [], []
                        ksite = 0
                        while ksite < nsit:
                            line = next(it)
                            words = line.split()
                            lab = words[0][:8]
                            rest = line.split(None, 1)[1]
                            w, rest = dblstr(rest)
                            q, rest = dblstr(rest)
                            if is_guest:
                                x, rest = dblstr(rest)
                                y, rest = dblstr(rest)
                                z, rest = dblstr(rest)
                                nrept, ifrz = 1, 0
                                tmpl.append((x, y, z))
                            else:
                                nrept, rest = intstr(rest)
                                ifrz, rest = intstr(rest)
                            if nrept == 0:
                                nrept = 1
                            for _r in range(nrept):
                                ksite += 1
                                labs.append(lab)
                                mass.append(w)
                                chg.append(q)
                                frz.append(ifrz)
                            if lab not in labels:
                                labels.append(lab)
                        sp = Species(name=name_line.strip(), is_guest=is_guest, nmols=nummols,
                                     labels=labs, mass=np.array(mass), charge=np.array(chg),
                                     frozen=np.array(frz, dtype=np.int32))
                        if is_guest:
                            sp.template = np.array(tmpl, dtype=np.float64)
                    elif "finish" in r:
                        break
                sp.nmols = nummols
                sp.types = np.array([labels.index(l) + 1 for l in sp.labels], dtype=np.int32)
                mass_sum = float(sum(sp.mass))
                if sp.is_guest:
                    # readinputs.f:243-265: the site loop already summed the
                    # masses, the COM loop adds them again, so the template is
                    # shifted by half its true COM.
                    sp.molmass = mass_sum
                    tot = 0.0
                    for m in sp.mass:
                        tot += m
                    tot2 = mass_sum
                    for m in sp.mass:
                        tot2 += m
                    com = np.zeros(3)
                    cx = cy = cz = 0.0
                    for (x, y, z), m in zip(sp.template, sp.mass):
                        cx += x * m
                        cy += y * m
                        cz += z * m
                    com = np.array([cx / tot2, cy / tot2, cz / tot2])
                    sp.template = sp.template - com
                else:
                    sp.molmass = mass_sum * nummols
                species.append(sp)
        elif "close" in rec:
            break
    sumchg = 0.0
    for sp in species:
        for q in sp.charge:
            sumchg = sumchg + float(sp.nmols * q)
    return title, engunit, species, labels, pairs, sumchg


# ------------------------------------------------------------------ CONFIG
def read_config(path: str, species: list):
    with open(path) as f:
        lines = f.read().split("\n")
    levcfg, rest = intstr(lines[1])
    imcon, _ = intstr(rest)
    cell = []
    for r in range(3):
        rest = lines[2 + r]
        for _ in range(3):
            v, rest = dblstr(rest)
            cell.append(v)
    pos = 5
    for sp in species:
        n = sp.nmols * sp.natoms
        xyz = np.zeros((n, 3))
        for a in range(n):
            name = lines[pos][:8].strip()
            coord = lines[pos + 1]
            m = a % sp.natoms
            if name != sp.labels[m]:
                raise InputError(25, f"unidentified atom label {name!r} at atom {a + 1}")
            xyz[a] = [float(coord[0:20].replace("d", "e").replace("D", "e")),
                      float(coord[20:40].replace("d", "e").replace("D", "e")),
                      float(coord[40:60].replace("d", "e").replace("D", "e"))]
            pos += 2 + levcfg
        sp.xyz = xyz
    return imcon, levcfg, np.array(cell, dtype=np.float64)


# ----------------------------------------------------------------- CONTROL
def read_control(path: str, nguests: int):
    with open(path) as f:
        lines = f.read().split("\n")
    out = {"temperature": None, "rcut": None, "delr": None, "eps": None, "alpha": None,
           "kmax": None, "rvdw": None, "overlap": 1.5, "surftol": 0.0, "steps": 1, "eqsteps": 0,
           "nwind": 0, "accsample": False, "nnumg": 0, "griddim": 0.0, "gridfactor": (1, 1, 1),
           "prob_plots": [[] for _ in range(nguests)], "prob_energy": [False] * nguests, "widom": False,
           "nwidstep": 0,
           "pressures": [-1.0] * nguests, "moves": [dict() for _ in range(nguests)],
           "guest_delr": [None] * nguests, "topmoves": {}}
    i = 1
    while i < len(lines):
        rec = _lc(lines[i]).strip()
        i += 1
        if rec == "" or rec.startswith("#"):
            continue
        if "&guest" in rec:
            ig, _ = intstr(rec)
            while i < len(lines):
                r = _lc(lines[i]).strip()
                i += 1
                if r == "" or r.startswith("#"):
                    continue
                if "probability" in r:
                    # readinputs.f:642-672: N lines, each 'n s1 .. sn' (site
                    # 0: centre of mass) or starting with 'e' (energy plot)
                    nplot, _ = intstr(r)
                    got = 0
                    while got < nplot and i < len(lines):
                        pr = _lc(lines[i]).strip()
                        i += 1
                        if pr == "" or pr.startswith("#"):
                            continue
                        got += 1
                        if pr.startswith("e"):
                            out["prob_energy"][ig - 1] = True
                            continue
                        ns, rest = intstr(pr)
                        sites = []
                        for _ in range(ns):
                            v, rest = intstr(rest)
                            sites.append(v)
                        out["prob_plots"][ig - 1].append(sites)
                elif "press" in r and "move" not in r:
                    v, _ = dblstr(r)
                    out["pressures"][ig - 1] = v * 1.0e5
                elif "move" in r:
                    for k in MOVE_KEYS:
                        if k in r:
                            v, _ = dblstr(r)
                            out["moves"][ig - 1][k] = v
                            break
                elif "delr" in r:
                    v, _ = dblstr(r)
                    out["guest_delr"][ig - 1] = v
                elif "&end" in r:
                    break
            continue
        if "temp" in rec:
            out["temperature"], _ = dblstr(rec)
        elif "grid" in rec and ("spa" in rec or "fac" in rec):
            # utility_pack.f:436-442
            if "spa" in rec:
                out["griddim"], _ = dblstr(rec)
            else:
                a, rest = intstr(rec)
                b, rest = intstr(rest)
                c, rest = intstr(rest)
                out["gridfactor"] = (a, b, c)
        elif "averaging window" in rec:
            out["nwind"], _ = intstr(rec)           # utility_pack.f:394-396
        elif "henry" in rec:
            out["widom"] = True                     # utility_pack.f:389-390
        elif "widom ins" in rec:
            out["nwidstep"], _ = intstr(rec)        # readinputs.f:598-599
        elif "sampling" in rec:
            out["accsample"] = True                 # readinputs.f:608-613
        elif "numg" in rec:
            n, _ = intstr(rec)                      # readinputs.f:713-717
            out["nnumg"] = n if n > 0 else 1000
        elif "steps" in rec:
            out["steps"], _ = intstr(rec)
        elif "equilibr" in rec:
            out["eqsteps"], _ = intstr(rec)
        elif "cut" in rec:
            out["rcut"], _ = dblstr(rec)
        elif "overlap" in rec:
            out["overlap"], _ = dblstr(rec)
        elif "surface" in rec:
            out["surftol"], _ = dblstr(rec)
        elif "delr" in rec:
            out["delr"], _ = dblstr(rec)
        elif "rvdw" in rec:
            out["rvdw"], _ = dblstr(rec)
        elif "move" in rec:
            for k in MOVE_KEYS:
                if k in rec:
                    v, _ = dblstr(rec)
                    out["topmoves"][k] = v
                    break
        elif "ewald" in rec:
            if "precision" in rec:
                out["eps"], _ = dblstr(rec)
            else:
                a, rest = dblstr(rec)
                k1, rest = intstr(rest)
                k2, rest = intstr(rest)
                k3, rest = intstr(rest)
                out["alpha"] = a
                out["kmax"] = (k1, k2, k3)
        elif "finish" in rec:
            break
    return out


def ewald_from_precision(eps: float, rcut: float, cell, imcon: int):
    """initscan Ewald block, utility_pack.f:411-424 (Fortran nint = round half away)."""
    celprp = dcell(cell)
    eps = min(abs(eps), 0.5)
    tol = math.sqrt(abs(math.log(eps * rcut)))
    alpha = math.sqrt(abs(math.log(eps * rcut * tol))) / rcut
    tol1 = math.sqrt(-math.log(eps * rcut * (2.0 * tol * alpha) ** 2))
    fac = 1.0
    if imcon in (4, 5, 7):
        fac = 2.0 ** (1.0 / 3.0)

    def nint(x):
        return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))

    kmax = tuple(nint(0.25 + fac * celprp[i] * alpha * tol1 / PI) for i in range(3))
    return alpha, kmax


def read_system(dirpath: str) -> System:
    title, engunit, species, labels, pairs, sumchg = read_field(os.path.join(dirpath, "FIELD"))
    imcon, levcfg, cell = read_config(os.path.join(dirpath, "CONFIG"), species)
    nguests = sum(1 for s in species if s.is_guest)
    ctl = read_control(os.path.join(dirpath, "CONTROL"), nguests)
    if ctl["rcut"] is None:
        raise InputError(433, "cutoff missing")
    if ctl["delr"] is None:
        raise InputError(433, "delr missing")
    if ctl["temperature"] is None:
        raise InputError(380, "temperature missing")
    rcut = ctl["rcut"]
    if ctl["eps"] is not None:
        alpha, kmax = ewald_from_precision(ctl["eps"], rcut, cell, imcon)
    elif ctl["alpha"] is not None:
        alpha, kmax = ctl["alpha"], ctl["kmax"]
    else:
        raise InputError(2311, "no ewald directive")
    celprp = dcell(cell)
    width = min(celprp[6], celprp[7], celprp[8]) / 2.0
    if rcut > width:
        raise InputError(95, "cutoff exceeds half the cell width")
    moves = []
    for ig in range(nguests):
        mv = dict(ctl["moves"][ig])
        for k in MOVE_KEYS:
            if k not in mv:
                mv[k] = ctl["topmoves"].get(k, 0.0)
        moves.append(mv)
    gdelr = [d if d is not None else ctl["delr"] for d in ctl["guest_delr"]]
    return System(title=title, engunit=engunit, species=species, labels=labels, pairs=pairs,
                  imcon=imcon, cell=cell, levcfg=levcfg, sumchg=sumchg,
                  temperature=ctl["temperature"], rcut=rcut,
                  rvdw=ctl["rvdw"] if ctl["rvdw"] else rcut, delr=ctl["delr"], alpha=alpha,
                  kmax=kmax, eps_precision=ctl["eps"] or 0.0, overlap=ctl["overlap"],
                  surftol=ctl["surftol"], steps=ctl["steps"], eqsteps=ctl["eqsteps"],
                  nwind=ctl["nwind"], accsample=ctl["accsample"], nnumg=ctl["nnumg"],
                  griddim=ctl["griddim"], gridfactor=ctl["gridfactor"], prob_plots=ctl["prob_plots"],
                  prob_energy=ctl["prob_energy"], widom=ctl["widom"], nwidstep=ctl["nwidstep"],
                  pressures=ctl["pressures"], moves=moves, guest_delr=gdelr)


# ------------------------------------------------------- ctypes structures
class FmcSpecies(C.Structure):
    _fields_ = [("is_guest", C.c_int), ("natoms", C.c_int), ("nmols", C.c_int),
                ("site_charge", C.POINTER(C.c_double)), ("site_mass", C.POINTER(C.c_double)),
                ("site_type", C.POINTER(C.c_int)), ("site_frozen", C.POINTER(C.c_int)),
                ("template_xyz", C.POINTER(C.c_double)), ("xyz", C.POINTER(C.c_double))]


class FmcSystem(C.Structure):
    _fields_ = [("imcon", C.c_int), ("cell", C.c_double * 9), ("nspecies", C.c_int),
                ("species", C.POINTER(FmcSpecies)), ("ntypes", C.c_int), ("npairs", C.c_int),
                ("pair_i", C.POINTER(C.c_int)), ("pair_j", C.POINTER(C.c_int)),
                ("pair_kind", C.POINTER(C.c_int)), ("pair_eps", C.POINTER(C.c_double)),
                ("pair_sig", C.POINTER(C.c_double)), ("rcut", C.c_double), ("rvdw", C.c_double),
                ("delr", C.c_double), ("alpha", C.c_double), ("kmax", C.c_int * 3),
                ("engunit", C.c_double), ("epsq", C.c_double), ("overlap", C.c_double),
                ("surftol", C.c_double), ("sumchg", C.c_double)]


class FmcWalker(C.Structure):
    _fields_ = [("temperature", C.c_double), ("fugacity", C.c_double * FMC_MAX_SPECIES),
                ("move_cum", (C.c_float * FMC_NMOVES) * FMC_MAX_SPECIES),
                ("delr", C.c_double * FMC_MAX_SPECIES), ("seed", C.c_int * 4),
                ("framework_xyz", C.POINTER(C.c_double))]


class FmcTrial(C.Structure):
    _fields_ = [("estep", C.c_double), ("delE", C.c_double * FMC_MAX_SPECIES),
                ("recip", C.c_double), ("real", C.c_double), ("vdw", C.c_double),
                ("tail", C.c_double), ("overlap", C.c_int), ("surface", C.c_int)]


class FmcTraceRec(C.Structure):
    _fields_ = [("step", C.c_int), ("iguest", C.c_int), ("move", C.c_int), ("choice", C.c_int),
                ("nmols", C.c_int), ("overlap", C.c_int), ("accepted", C.c_int),
                ("drew_rande", C.c_int), ("estep", C.c_double), ("rande", C.c_double),
                ("xyz", C.c_double * (3 * FMC_MAX_SITES)),
                ("energy", C.c_double * FMC_MAX_SPECIES), ("delE", C.c_double * FMC_MAX_SPECIES),
                ("delrdisp", C.c_double), ("rotangle", C.c_double),
                ("rota_rotangle", C.c_double), ("surface", C.c_int), ("nsurface", C.c_int)]


class FmcWalkerStats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("accepted", C.c_int64 * 9), ("attempted", C.c_int64 * 9),
                ("nmols", C.c_int * FMC_MAX_SPECIES), ("overflow", C.c_int),
                ("energy", C.c_double * FMC_MAX_SPECIES), ("sum_n", C.c_double * FMC_MAX_SPECIES),
                ("sum_n2", C.c_double * FMC_MAX_SPECIES), ("sum_e", C.c_double),
                ("sum_e2", C.c_double), ("counted", C.c_int64), ("resums", C.c_int64),
                ("resum_drift", C.c_double), ("surface", C.c_int * FMC_MAX_SPECIES),
                ("prodcount", C.c_int64), ("windows", C.c_int64)]


FMC_WIN_STATS = 6


class FmcWindow(C.Structure):
    _fields_ = [("count", C.c_double), ("sums", (C.c_double * FMC_WIN_STATS) * FMC_MAX_SPECIES)]


FMC_MAX_PLOTS = 8


class FmcNumguestsRec(C.Structure):
    _fields_ = [("step", C.c_int64), ("nmols", C.c_int * FMC_MAX_SPECIES),
                ("moves", (C.c_int * 6) * FMC_MAX_SPECIES), ("energy", C.c_double * FMC_MAX_SPECIES),
                ("delE", C.c_double * FMC_MAX_SPECIES)]


class FmcProbPlot(C.Structure):
    _fields_ = [("iguest", C.c_int), ("nsites", C.c_int), ("sites", C.c_int * (FMC_MAX_SITES + 1))]


def prob_plots(plots):
    """(iguest, site list) pairs -> an FmcProbPlot array."""
    arr = (FmcProbPlot * max(len(plots), 1))()
    for i, (g, sites) in enumerate(plots):
        arr[i].iguest, arr[i].nsites = g, len(sites)
        for e, v in enumerate(sites):
            arr[i].sites[e] = v
    return arr


class FmcGuestStats(C.Structure):
    _fields_ = [("avg", C.c_double * 8), ("std", C.c_double * 8), ("avg_final", C.c_double * 8),
                ("std_final", C.c_double * 8)]


class FmcWidomResult(C.Structure):
    _fields_ = [("avg_boltz", C.c_double), ("henry", C.c_double), ("attempts", C.c_int64),
                ("accepted", C.c_int64), ("ngrid", C.c_int * 3), ("wngrid", C.c_int * 3)]


class FmcTheta(C.Structure):
    _fields_ = [("fw_charge", C.POINTER(C.c_double)), ("type_eps", C.POINTER(C.c_double)),
                ("type_sig", C.POINTER(C.c_double))]


class FmcWlParams(C.Structure):
    _fields_ = [("maxn", C.c_int), ("minn", C.c_int), ("nwindows", C.c_int), ("visittol", C.c_int),
                ("flatcoeff", C.c_double), ("wlprec", C.c_double), ("prectol", C.c_double),
                ("create_molecules", C.c_int), ("pad", C.c_int)]


class FmcWlStatus(C.Structure):
    _fields_ = [("minmol", C.c_int), ("maxmol", C.c_int), ("done", C.c_int), ("error", C.c_int),
                ("sweeps", C.c_int64), ("steps", C.c_int64), ("f", C.c_double), ("lnf", C.c_double),
                ("lambda3", C.c_double)]


TRACE_DTYPE = np.dtype([("step", "i4"), ("iguest", "i4"), ("move", "i4"), ("choice", "i4"),
                        ("nmols", "i4"), ("overlap", "i4"), ("accepted", "i4"),
                        ("drew_rande", "i4"), ("estep", "f8"), ("rande", "f8"),
                        ("xyz", "f8", (3 * FMC_MAX_SITES,)), ("energy", "f8", (FMC_MAX_SPECIES,)),
                        ("delE", "f8", (FMC_MAX_SPECIES,)), ("delrdisp", "f8"),
                        ("rotangle", "f8"), ("rota_rotangle", "f8"), ("surface", "i4"),
                        ("nsurface", "i4")])
assert TRACE_DTYPE.itemsize == C.sizeof(FmcTraceRec)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class CSystem:
    """Owns the numpy buffers behind an ``fmc_system`` struct."""

    def __init__(self, sysm: System):
        self.sysm = sysm
        self._keep = []
        sp_arr = (FmcSpecies * len(sysm.species))()
        for i, sp in enumerate(sysm.species):
            q = np.ascontiguousarray(sp.charge, dtype=np.float64)
            m = np.ascontiguousarray(sp.mass, dtype=np.float64)
            t = np.ascontiguousarray(sp.types, dtype=np.int32)
            fz = np.ascontiguousarray(sp.frozen, dtype=np.int32)
            tm = np.ascontiguousarray(sp.template if sp.template is not None
                                      else np.zeros((sp.natoms, 3)), dtype=np.float64)
            xyz = np.ascontiguousarray(sp.xyz if sp.xyz is not None and len(sp.xyz)
                                       else np.zeros((1, 3)), dtype=np.float64)
            self._keep += [q, m, t, fz, tm, xyz]
            sp_arr[i] = FmcSpecies(int(sp.is_guest), sp.natoms, sp.nmols, _dp(q), _dp(m), _ip(t),
                                   _ip(fz), _dp(tm), _dp(xyz))
        self._keep.append(sp_arr)
        pi = np.array([p[0] for p in sysm.pairs], dtype=np.int32)
        pj = np.array([p[1] for p in sysm.pairs], dtype=np.int32)
        pk = np.array([p[2] for p in sysm.pairs], dtype=np.int32)
        pe = np.array([p[3] for p in sysm.pairs], dtype=np.float64)
        ps = np.array([p[4] for p in sysm.pairs], dtype=np.float64)
        self._keep += [pi, pj, pk, pe, ps]
        s = FmcSystem()
        s.imcon = sysm.imcon
        for k in range(9):
            s.cell[k] = float(sysm.cell[k])
        s.nspecies = len(sysm.species)
        s.species = sp_arr
        s.ntypes = len(sysm.labels)
        s.npairs = len(sysm.pairs)
        s.pair_i, s.pair_j, s.pair_kind = _ip(pi), _ip(pj), _ip(pk)
        s.pair_eps, s.pair_sig = _dp(pe), _dp(ps)
        s.rcut, s.rvdw, s.delr, s.alpha = sysm.rcut, sysm.rvdw, sysm.delr, sysm.alpha
        for k in range(3):
            s.kmax[k] = int(sysm.kmax[k])
        s.engunit, s.epsq = sysm.engunit, 1.0
        s.overlap, s.surftol, s.sumchg = sysm.overlap, sysm.surftol, sysm.sumchg
        self.struct = s


def make_walker(sysm: System, seed=(25, 89, 100, 150), temperature=None, pressures=None,
                framework_xyz=None, keep=None) -> FmcWalker:
    """Walker settings from a System (temperature, fugacities, REAL(4) move table)."""
    w = FmcWalker()
    w.temperature = float(sysm.temperature if temperature is None else temperature)
    pres = sysm.pressures if pressures is None else pressures
    for g in range(len(sysm.guests)):
        w.fugacity[g] = float(pres[g])
        cum = sysm.move_cum(g)
        for k in range(FMC_NMOVES):
            w.move_cum[g][k] = float(cum[k])
        w.delr[g] = float(sysm.guest_delr[g])
    for k in range(4):
        w.seed[k] = int(seed[k])
    if framework_xyz is not None:
        arr = np.ascontiguousarray(framework_xyz, dtype=np.float64)
        if keep is not None:
            keep.append(arr)
        w.framework_xyz = _dp(arr)
    else:
        w.framework_xyz = C.POINTER(C.c_double)()
    return w


def duni_seeds_from_clock(seconds: int, millis: int):
    """The seed rule of utility_pack.f:2799-2804 for a given clock reading."""
    i = (millis * 3) % 177 + 1
    j = (seconds * 23) % 177 + 1
    k = (millis * 5) % 177 + 1
    l = (seconds * 3) % 168
    return (i, j, k, l)
