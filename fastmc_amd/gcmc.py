"""Host driver of a GCMC run on the device chains: the control flow of the
reference main program (gcmc.f:704-765: equilibration, production and its
end), its production statistics (gcmc.f:1050-1258) and its OUTPUT report
(gcmc.f:1272-1557), so that a run reads the same FIELD / CONFIG / CONTROL
and writes a report that the reference's own tooling parses
(run_test.py:71-76).

    python -m fastmc_amd.gcmc <run directory> [--precision fp64|mixed] [--seed i j k l]

The chains themselves run in libfastmc_hip.so; this module only sequences
launches and formats numbers.
"""
from __future__ import annotations

import argparse
import math
import os
import sys
import time

from . import lib as L
from .system import System, read_system

STAT_NAMES = L.STAT_NAMES


# ------------------------------------------------------------- schedule --
def window_steps(mcsteps: int, nwind: int = 0):
    """(nwind, nwindsteps) of initscan (utility_pack.f:453-470).  Without an
    'averaging window' directive: 5 windows above 1000 production steps,
    else 1.  With one, the reference reads its window count over an
    uninitialised step count; here the production step count is used."""
    cfactor = abs(mcsteps)
    if not nwind:
        nwind = 5 if cfactor > 1000 else 1
    nwindsteps = math.ceil(cfactor / nwind)
    return cfactor // nwindsteps, nwindsteps


def check_interval(eqsteps: int, nnumg: int = 0) -> int:
    """minchk = min(400, nnumg) (gcmc.f:658); nnumg defaults to eqsteps/1000
    above 1000 equilibration steps, else 1 (readinputs.f:910-916)."""
    if not nnumg:
        nnumg = eqsteps // 1000 if eqsteps > 1000 else 1
    return min(400, nnumg)


def schedule(eqsteps: int, mcsteps: int, minchk: int):
    """(first production step, last step) of the main loop when every
    production step counts: production starts at the first check
    (gcmccount a multiple of minchk) with gcmccount >= eqsteps, or at step 1
    without equilibration (gcmc.f:598,752-757); the run ends after the step
    of the first later check that finds mcsteps production steps done
    (gcmc.f:707-752), and that step still counts."""
    if eqsteps < 0 or mcsteps <= 0:
        raise ValueError("cycle counts (negative steps) have no closed-form schedule: run_main_loop drives them")
    s0 = 1 if eqsteps == 0 else minchk * math.ceil(eqsteps / minchk)
    return s0, minchk * math.ceil((s0 + mcsteps) / minchk)


class Run:
    """Steps done and production state of one run of the main loop."""

    def __init__(self):
        self.steps = 0
        self.start = 0          # first production step (0: not started)
        self.prodcount = 0


def run_main_loop(engine, eqsteps: int, mcsteps: int, minchk: int, accepted_only: bool = False,
                  prodcount=None, nwindsteps: int = 0) -> Run:
    """Drive `engine.run(nsteps, flags)` (flags as fmc_run's count_stats)
    through the reference's main-loop schedule.  With every production step
    counted the schedule is known up front (three launches at most); with
    accepted_only the run ends on the production count, which `prodcount()`
    reads at every check (gcmc.f:707-752).  Negative step counts are cycles
    (gcmc.f:700-800, run_cycles); the engine then also needs
    `guests()`, the total guest count, and `nwindsteps` the window length
    in cycles."""
    if eqsteps < 0 or mcsteps < 0:
        return run_cycles(engine, eqsteps, mcsteps, minchk, accepted_only, prodcount, nwindsteps)
    r = Run()
    COUNT, CLOSE = L.FMC_STATS_COUNT, L.FMC_STATS_CLOSE
    if not accepted_only:
        s0, last = schedule(eqsteps, mcsteps, minchk)
        if s0 > 1:
            engine.run(s0 - 1, 0)
        if last - s0 > 0:
            engine.run(last - s0, COUNT)
        engine.run(1, COUNT | CLOSE)
        r.steps, r.start, r.prodcount = last, s0, last - s0 + 1
        return r
    production = eqsteps == 0
    if production:
        r.start = 1
    g = minchk - 1
    if g > 0:
        engine.run(g, COUNT if production else 0)
    while True:
        # the check at the top of step g + 1 (a multiple of minchk)
        if production:
            if prodcount() >= mcsteps:
                engine.run(1, COUNT | CLOSE)
                g += 1
                break
        elif g + 1 >= eqsteps:
            production = True
            r.start = g + 1
        engine.run(minchk, COUNT if production else 0)
        g += minchk
    r.steps = g
    r.prodcount = prodcount()
    return r


CYCLE_MULT = 20   # gcmc.f:141


def run_cycles(engine, eqsteps: int, mcsteps: int, minchk: int, accepted_only: bool, prodcount,
               nwindsteps: int) -> Run:
    """The main loop with equilibration and / or production counted in
    cycles (gcmc.f:700-800): every step advances cycle_step; at each check
    (every minchk steps) a cycle ends once cycle_step reaches maxcycle, which
    then becomes max(guests, 20) (initially max(20 x initial guests, 1000));
    equilibration ends after -eqsteps cycles, production after -mcsteps.
    With production in cycles, windows close when the cycle count is a
    multiple of nwindsteps at the check step that ended a cycle
    (gcmc.f:1085-1092), and on the last step; the device's own windows by
    production count must be off (nwindsteps 0 in fmc_set_sampling)."""
    COUNT, CLOSE = L.FMC_STATS_COUNT, L.FMC_STATS_CLOSE
    r = Run()
    production = eqsteps == 0
    if production:
        r.start = 1
    maxcycle = max(engine.guests() * CYCLE_MULT, 1000)
    cycle_count, cycle_step, g = 1, 0, 0
    counted = 0
    while True:
        # steps up to the next check, then the check at the top of its step
        seg = minchk - 1 - (g % minchk)
        if seg > 0:
            engine.run(seg, COUNT if production else 0)
            counted += seg if production else 0
            g += seg
            cycle_step += seg
        g += 1
        cycle_step += 1
        last = False
        if production:
            if mcsteps < 0:
                if cycle_step >= maxcycle:
                    cycle_step = 0
                    maxcycle = max(engine.guests(), CYCLE_MULT)
                    cycle_count += 1
                last = cycle_count >= -mcsteps
            else:
                last = (prodcount() if accepted_only else counted) >= mcsteps
        elif eqsteps < 0:
            if cycle_step >= maxcycle:
                cycle_step = 0
                maxcycle = max(engine.guests(), CYCLE_MULT)
                cycle_count += 1
            if cycle_count >= -eqsteps:
                cycle_count, cycle_step = 1, 0
                maxcycle = max(engine.guests(), CYCLE_MULT)
                production = True
                r.start = g
        elif g >= eqsteps:
            production = True
            r.start = g
        flags = COUNT if production else 0
        if production and (last or (mcsteps < 0 and nwindsteps > 0 and cycle_count % nwindsteps == 0
                                    and cycle_step == 0)):
            flags |= CLOSE
        engine.run(1, flags)
        counted += 1 if production else 0
        if last:
            break
    r.steps = g
    r.prodcount = prodcount() if prodcount else counted
    return r


# --------------------------------------------------------------- report --
def _f(x: float, w: int, d: int) -> str:
    """Fortran Fw.d (NaN / Infinity as gfortran-style text)."""
    if math.isnan(x):
        return "NaN".rjust(w)
    if math.isinf(x):
        return ("Infinity" if x > 0 else "-Infinity").rjust(w)
    s = f"{x:{w}.{d}f}"
    return s if len(s) <= w else "*" * w


def _e(x: float, w: int, d: int) -> str:
    """Fortran Ew.d: [-]0.ddd...E+xx, d significant digits."""
    if not math.isfinite(x):
        return _f(x, w, d)
    if x == 0.0:
        return f"0.{'0' * d}E+00".rjust(w)
    m, e = f"{abs(x):.{d - 1}e}".split("e")
    digits = m.replace(".", "")
    ex = int(e) + 1
    return f"{'-' if x < 0 else ''}0.{digits}E{'+' if ex >= 0 else '-'}{abs(ex):02d}".rjust(w)


# ---------------------------------------------------- probability plots --
ANGS2BOHR = 1.889725989                     # utility_pack.f:100

# atmnumber (utility_pack.f:2521-2766): atomic number from a FIELD mass, the
# first (low, high, Z) range holding it; bounds are single-precision literals
# in the reference, so they are compared as float32 values.  A mass in no
# range gives 0 (the reference returns an undefined value).
_ATM_RANGES = (
    (0.0, 1.5, 1), (3.9, 4.5, 2), (6.5, 7.1, 3), (8.9, 9.5, 4), (10.5, 11.1, 5), (11.9, 12.5, 6), (13.9, 14.5, 7),
    (15.5, 16.1, 8), (18.5, 19.1, 9), (19.9, 20.5, 10), (22.5, 23.1, 11), (23.9, 24.5, 12), (26.5, 27.1, 13),
    (27.9, 28.5, 14), (30.5, 31.1, 15), (31.9, 32.5, 16), (34.9, 36.1, 17), (39.5, 39.9999, 18), (38.9, 39.4, 19),
    (40.0, 40.5, 20), (44.5, 45.1, 21), (47.5, 48.1, 22), (50.5, 51.1, 23), (51.5, 52.1, 24), (54.5, 55.1, 25),
    (55.5, 56.1, 26), (58.76, 59.1, 27), (58.5, 59.75, 28), (62.9, 64.1, 29), (64.9, 66.1, 30), (69.5, 70.1, 31),
    (72.5, 73.1, 32), (74.5, 75.1, 33), (78.5, 79.1, 34), (79.5, 80.1, 35), (83.5, 84.1, 36), (84.9, 86.1, 37),
    (87.5, 88.1, 38), (88.5, 89.1, 39), (90.9, 91.5, 40), (92.5, 93.1, 41), (95.5, 96.1, 42), (97.9, 98.1, 43),
    (109.9, 101.5, 44), (102.5, 103.1, 45), (105.9, 106.5, 46), (107.5, 108.1, 47), (111.9, 112.5, 48),
    (114.5, 115.1, 49), (118.5, 119.1, 50), (121.5, 122.1, 51), (127.5, 128.1, 52), (126.5, 127.1, 53),
    (130.9, 131.5, 54), (132.5, 133.1, 55), (136.9, 137.5, 56), (138.5, 139.1, 57), (139.9, 140.5, 58),
    (140.6, 141.1, 59), (144.0, 144.5, 60), (144.9, 145.1, 61), (150.0, 150.6, 62), (151.5, 152.1, 63),
    (156.9, 157.5, 64), (158.5, 159.1, 65), (162.0, 163.1, 66), (164.5, 165.1, 67), (166.5, 167.9, 68),
    (168.0, 169.1, 69), (172.9, 173.5, 70), (174.0, 175.1, 71), (178.0, 179.1, 72), (180.0, 181.1, 73),
    (183.0, 184.1, 74), (185.9, 186.5, 75), (189.9, 190.5, 76), (191.9, 192.5, 77), (194.9, 195.5, 78),
    (196.5, 197.1, 79), (200.0, 201.1, 80), (203.9, 204.6, 81), (206.9, 207.6, 82), (208.5, 209.1, 83),
    (209.9, 210.1, 85), (221.9, 222.1, 86), (222.9, 223.1, 87), (225.9, 226.1, 88), (226.9, 227.1, 89),
    (231.9, 232.1, 90), (230.9, 231.1, 91), (237.9, 238.1, 92), (243.9, 244.1, 94), (242.9, 243.1, 95),
    (246.9, 247.1, 96), (250.9, 251.1, 98), (251.9, 252.1, 99), (256.9, 257.1, 100), (257.9, 258.1, 101),
    (258.9, 259.1, 102), (261.9, 262.1, 103), (266.9, 267.1, 104), (267.9, 268.1, 105), (270.9, 271.1, 106),
    (269.9, 270.1, 107), (268.9, 269.1, 108), (277.9, 278.1, 109), (280.9, 281.1, 110), (284.9, 285.1, 112),
    (285.9, 286.1, 113), (288.9, 289.1, 114), (292.9, 293.1, 116), (293.9, 294.1, 117))


def atmnumber(mass: float) -> int:
    import numpy as np
    for lo, hi, z in _ATM_RANGES:
        if float(np.float32(lo)) <= mass <= float(np.float32(hi)):
            return z
    return 0


def prob_grid_size(cell, griddim: float, gridfactor=(1, 1, 1)):
    """Grid points per axis of the probability plots and Widom grid
    (gcmc.f:258-281)."""
    from .system import dcell
    celprp = dcell(cell)
    out = []
    for a in range(3):
        gd = celprp[a] / math.ceil(celprp[a] / gridfactor[a] / griddim) / gridfactor[a]
        out.append(math.ceil(celprp[a] / (gd * gridfactor[a])))
    return tuple(out)


def numguests_text(records, n: int, g: int) -> str:
    """numguests.out lines of guest type g (1-based) from n trace records
    (format (i9,i7,2f20.6,7i4), gcmc.f:1150-1152)."""
    out = []
    for r in records[:n]:
        out.append(f"{r.step:9d}{r.nmols[g - 1]:7d}{_f(r.energy[g - 1], 20, 6)}{_f(r.delE[g - 1], 20, 6)}"
                   + "".join(f"{r.moves[g - 1][k]:4d}" for k in range(6)))
    return "".join(line + "\n" for line in out)


def numguests_names(ntpguest: int):
    """branch01/numguests.out for one guest type, numguestsGG.out for more
    (gcmc.f:478-493)."""
    if ntpguest == 1:
        return ["numguests.out"]
    return [f"numguests{g:02d}.out" for g in range(1, ntpguest + 1)]


def cube_text(sysm, grid, ngrid, gridfactor, steps: int) -> str:
    """writeprob (utility_pack.f:773-871): a Gaussian cube file of one plot,
    the grid divided by the voxel volume (bohr^3) and the production step
    count; the framework atoms of one unit cell as the molecule list."""
    return _cube(sysm, ngrid, gridfactor, ("probability cube file", "outer loop a, middle loop     b, inner loop c"),
                 lambda vol, j: float(grid[j]) / vol / float(steps))


def enprob_text(sysm, emap, cmap, ngrid, gridfactor) -> str:
    """writeenprob (utility_pack.f:689-771): the energy plot, binned energy
    over binned count per grid point (0 where nothing was binned)."""
    return _cube(sysm, ngrid, gridfactor, ("Energy histogram cube file",
                                           "outer loop a,            middle loop b, inner loop c"),
                 lambda vol, j: float(emap[j]) / float(cmap[j]) if cmap[j] > 0.0 else 0.0)


def _cube(sysm, ngrid, gridfactor, title, value) -> str:
    na, nb, nc = ngrid
    g = [0.0] + [float(sysm.cell.ravel()[i]) / float(gridfactor[i // 3]) * ANGS2BOHR for i in range(9)]
    vol = ((g[2] * g[6] - g[3] * g[5]) / (na * nb) * g[7] / nc + (g[3] * g[4] - g[1] * g[6]) / (na * nb) * g[8] / nc
           + (g[1] * g[5] - g[2] * g[4]) / (na * nb) * g[9] / nc)
    scell = gridfactor[0] * gridfactor[1] * gridfactor[2]
    fw = sysm.framework_species
    nfram = 0
    for sp in fw:
        nfram = int(nfram + sp.nmols / float(scell) * len(sp.labels))
    lines = [title[0], title[1], f"{nfram:6d}" + "".join(_f(0.0, 12, 6) for _ in range(3))]
    for n, off in ((na, 1), (nb, 4), (nc, 7)):
        lines.append(f"{n:6d}" + "".join(_f(g[off + k] / n, 12, 6) for k in range(3)))
    for sp in fw:
        nmol = int(sp.nmols / float(scell))
        natms = len(sp.labels)
        for a in range(nmol * natms):
            x, y, z = sp.xyz[a]
            lines.append(f"{atmnumber(float(sp.mass[a % natms])):6d}" + _f(0.0, 12, 6) + _f(x * ANGS2BOHR, 12, 6)
                         + _f(y * ANGS2BOHR, 12, 6) + _f(z * ANGS2BOHR, 12, 6))
    row, ip = [], 0
    for j in range(1, na * nb * nc + 1):
        ip += 1
        row.append(_e(value(vol, j - 1), 15, 6))
        if j % nc == 0 or ip == 6:
            lines.append("".join(row))
            row, ip = [], 0
    return "\n".join(lines) + "\n"


def species_name(sp) -> str:
    """molnam: the FIELD record lower-cased, '&guest' dropped, 40 characters
    (readinputs.f:121-140)."""
    rec = sp.name.lower().strip()
    if sp.is_guest and rec.startswith("&guest"):
        rec = rec[len("&guest"):].strip()
    return rec[:40].ljust(40)


def node_report(names, stats, prodcount: int) -> list:
    """gcmc.f:1276-1311: the node's window-weighted averages."""
    out = ["", "*" * 100, "   Data reported from node   0", "*" * 100, ""]
    for i, (nm, st) in enumerate(zip(names, stats), start=1):
        a, s = st["avg"], st["std"]
        out += [f"     guest {i:2d}: {nm}", ""]
        rows = [("<N>: ", a["<N>"]), ("<E>: ", a["<E>"]), ("<E*N>: ", a["<E*N>"]), ("<N*N>: ", a["<N*N>"]),
                ("<E*E>: ", a["<E*E>"])]
        out += ["     " + k.rjust(60) + _f(v, 20, 9) for k, v in rows]
        out += ["     " + "Multiplier: ".rjust(60) + f"{prodcount:20d}"]
        rows = [("Isosteric heat of adsorption (kcal/mol): ", a["Qst"]), ("Isosteric heat error: ", s["Qst"]),
                ("Heat capacity, Cv (kcal/mol/K): ", a["Cv"]), ("Heat capacity error: ", s["Cv"]),
                ("<surface adsorbed N>:", a["<NF>"])]
        out += ["     " + k.rjust(60) + _f(v, 20, 9) for k, v in rows]
        out += [""]
    return out


def final_report(names, stats, counted: int, surftol: float) -> list:
    """gcmc.f:1506-1556: per guest the union over the one node."""
    out = []
    for i, (nm, st) in enumerate(zip(names, stats), start=1):
        a, s = st["avg_final"], st["std_final"]
        out += ["", "=" * 100, f"final stats for guest {i:2d}   {nm}", "=" * 100, "", ""]
        for k in ("<N>", "<E>", "<E*N>", "<N*N>", "<E*E>"):
            out.append((k + ": ").rjust(36) + _f(a[k], 15, 6) + " +/- " + _f(s[k], 12, 3))
        out += ["", ""]
        out.append("average number of guests: ".rjust(60) + _f(a["<N>"], 15, 6))
        out.append("standard error: ".rjust(60) + _f(s["<N>"], 15, 6))
        out.append("Isosteric heat of adsorption (kcal/mol): ".rjust(60) + _f(a["Qst"], 15, 6))
        out.append("Isosteric heat error: ".rjust(60) + _f(s["Qst"], 15, 6))
        if surftol >= 0.0:
            out.append("average surface adsorption: ".rjust(60) + _f(a["<NF>"], 15, 6))
            out.append("standard error: ".rjust(60) + _f(s["<NF>"], 15, 6))
        out.append("Heat capacity, Cv (kcal/mol/K): ".rjust(60) + _f(a["Cv"], 15, 6))
        out.append("Heat capacity error: ".rjust(60) + _f(s["Cv"], 15, 6))
        out.append("Total steps counted: ".rjust(60) + f"{counted:15d}")
        out.append("")
    return out


def move_report(steps: int, seconds: float, att, acc) -> list:
    """gcmc.f:1458-1491: elapsed time, accepted total (switches left out) and
    the acceptance ratio of every move kind that was tried."""
    out = ["", f"Time elapsed for {steps:9d} gcmc steps : {seconds:13.3f} seconds"]
    total = sum(int(acc[k]) for k in range(1, 8))
    out += ["", "Total accepted steps : ".rjust(30) + f"{total:9d}"]
    names = [(1, "insertion ratio: "), (2, "deletion ratio: "), (3, "displacement ratio: "), (6, "jump ratio: "),
             (7, "swap ratio: "), (8, "switch ratio: "), (4, "translation ratio: "), (5, "rotation ratio: ")]
    for k, nm in names:
        if att[k] > 0:
            out += ["", "   " + nm.rjust(21) + _f(acc[k] / att[k], 15, 9)]
    return out


def energy_report(sp_energies, names, e_vdw, e_coul, seeds) -> list:
    """Initial energies (gcmc.f:578-587) and the duni seeds (utility_pack.f:2814)."""
    out = []
    for i, e in enumerate(sp_energies, start=1):
        out.append("Initial guest ".rjust(23) + f"{i:3d}" + " energy :".rjust(9) + _f(e, 22, 6))
    out.append("van der Waals energy :".rjust(35) + _f(e_vdw, 22, 6))
    out.append("Electrostatic energy :".rjust(35) + _f(e_coul, 22, 6))
    out += ["", "i,j,k,l values for duni()" + "".join(f"{s:5d}" for s in seeds), ""]
    return out


# ---------------------------------------------------------------- driver --
class _Engine:
    def __init__(self, handle):
        self.h = handle

    def run(self, n, flags):
        self.h.run(n, count_stats=flags)

    def prodcount(self):
        return int(self.h.stats(0).prodcount)


def widom_report(names, henry, counted: int, attempts: int, seconds: float) -> list:
    """The Widom run's report (gcmc.f:1313-1321, 1458-1466, 1545-1550):
    Henry's constant per guest (E20.9 in the node report, E15.6 in the
    final one), the attempt count as the step count, the non-overlapping
    attempts as accepted and counted steps."""
    out = ["", "*" * 100, "   Data reported from node   0", "*" * 100, ""]
    for i, (nm, kh) in enumerate(zip(names, henry), start=1):
        out += [f"     guest {i:2d}: {nm}", "", "     " + "Henry's Constant (mol /kg /bar): ".rjust(60) + _e(kh, 20, 9),
                ""]
    out += ["", f"Time elapsed for {attempts:9d} gcmc steps : {seconds:13.3f} seconds"]
    out += ["", "Total accepted steps : ".rjust(30) + f"{counted:9d}"]
    for i, (nm, kh) in enumerate(zip(names, henry), start=1):
        out += ["", "=" * 100, f"final stats for guest {i:2d}   {nm}", "=" * 100, "", ""]
        out += ["Henry's Constant (mol /kg /bar): ".rjust(60) + _e(kh, 15, 6),
                "Total steps counted: ".rjust(60) + f"{counted:15d}", ""]
    return out


def run_widom(path: str, sysm: System, h, seed, out_name: str):
    """The 'henry' run: widom_grid for every guest type (gcmc.f:605-635,
    2454-2548) on the device, then its report."""
    names = [species_name(s) for s in sysm.guests]
    t0 = time.perf_counter()
    res = [h.widom(g, sysm.nwidstep, sysm.griddim, sysm.gridfactor, energy_map=sysm.prob_energy[g - 1])
           for g in range(1, len(sysm.guests) + 1)]
    seconds = time.perf_counter() - t0
    # the energy plots of the Widom run (writeenprob after widom_grid)
    for g, r in enumerate(res, start=1):
        if sysm.prob_energy[g - 1] and sum(int(x["accepted"]) for x in res) > 0:
            ng = prob_grid_size(sysm.cell, sysm.griddim, sysm.gridfactor)
            with open(os.path.join(path, f"Ehist_guest{g:02d}.cube"), "w") as fh:
                fh.write(enprob_text(sysm, r["emap"], r["cmap"], ng, sysm.gridfactor))
    counted = sum(int(r["accepted"]) for r in res)
    lines = energy_report([0.0] * len(names), names, *h.single_point()[[3, 0]], seed)
    lines += widom_report(names, [r["henry"] for r in res], counted, int(res[-1]["attempts"]), seconds)
    with open(os.path.join(path, out_name), "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return res


def run_directory(path: str, precision: int = L.FMC_FP64, seed=(12, 34, 56, 78), out_name: str = "OUTPUT",
                  max_windows: int = 256):
    """One GCMC run (or, with 'henry' in CONTROL, one Widom grid run) of the
    FIELD / CONFIG / CONTROL in `path` on the device; writes the report to
    path/out_name and returns (Run, statistics), or the Widom results."""
    sysm: System = read_system(path)
    h = L.Handle(sysm, precision=precision, seeds=[tuple(seed)])
    try:
        if sysm.widom:
            return run_widom(path, sysm, h, seed, out_name)
        sp = h.single_point()
        nwind, nwindsteps = window_steps(sysm.steps, sysm.nwind)
        minchk = check_interval(sysm.eqsteps, sysm.nnumg)
        cycles = sysm.steps < 0 or sysm.eqsteps < 0
        h.set_sampling(0 if sysm.steps < 0 else nwindsteps, sysm.accsample, max_windows)
        plots, cubes = [], []
        for g, pl in enumerate(sysm.prob_plots, start=1):
            for k, sites in enumerate(pl, start=1):
                plots.append((g, sites))
                cubes.append(f"prob_guest{g:02d}_prob_{k:02d}.cube")
        ngrid = h.set_prob(sysm.griddim, sysm.gridfactor, plots) if plots else None
        # numguests.out every nnumg steps (readinputs.f:713-717,910-916)
        nnumg = sysm.nnumg or (sysm.eqsteps // 1000 if sysm.eqsteps > 1000 else 1)
        last = 2000000 if cycles else schedule(sysm.eqsteps, sysm.steps, minchk)[1] * (4 if sysm.accsample else 1)
        h.set_numguests(nnumg, last // nnumg + 1)
        eng = _Engine(h)
        t0 = time.perf_counter()
        r = run_main_loop(eng, sysm.eqsteps, sysm.steps, minchk, sysm.accsample, eng.prodcount, nwindsteps)
        seconds = time.perf_counter() - t0
        st = h.stats(0)
        stats = h.production_stats(0)
        names = [species_name(s) for s in sysm.guests]
        lines = [f"Averages will be computed from {nwind:6d} averaging windows. "
                 f"Each window will average over {nwindsteps:10d} steps (or cycles)."]
        ens = [float(st.energy[i]) for i, s in enumerate(sysm.species) if s.is_guest]
        lines += energy_report(ens, names, sp[3], sp[0], seed)
        if sysm.eqsteps < 0:
            lines += [f"Completed at least {-sysm.eqsteps:10d} equilibration cycles for each guest",
                      f"Starting production at{r.start:10d} over all nodes"]
        elif r.start > 1:
            lines += [f"Completed at least {sysm.eqsteps:10d} equlibration steps",
                      f"Starting production at {r.start:10d}"]
        if sysm.steps < 0:
            lines += [f"Completed at least {-sysm.steps:10d} production cycles for each guest"]
        lines += node_report(names, stats, int(st.prodcount))
        lines += move_report(r.steps, seconds, list(st.attempted), list(st.accepted))
        lines += final_report(names, stats, int(st.prodcount), sysm.surftol)
        with open(os.path.join(path, out_name), "w") as fh:
            fh.write("\n".join(lines) + "\n")
        recs, nrec = h.numguests(0)
        os.makedirs(os.path.join(path, "branch01"), exist_ok=True)
        for g, name in enumerate(numguests_names(len(sysm.guests)), start=1):
            with open(os.path.join(path, "branch01", name), "w") as fh:
                fh.write(numguests_text(recs, nrec, g))
        if st.overflow & 4:
            sys.stderr.write("fastmc: numguests records truncated\n")
        if (plots or any(sysm.prob_energy)) and st.prodcount > 0:
            for p, name in enumerate(cubes):
                with open(os.path.join(path, name), "w") as fh:
                    fh.write(cube_text(sysm, h.prob(p), ngrid, sysm.gridfactor, int(st.prodcount)))
            # The energy plot of a GCMC run: the reference bins the main
            # program's guest_toten, which stays 0 (mc_displace sets a local
            # of the same name, gcmc.f:1840,1904), so every value is 0.
            eg = prob_grid_size(sysm.cell, sysm.griddim, sysm.gridfactor)
            zero = [0.0] * (eg[0] * eg[1] * eg[2])
            for g, on in enumerate(sysm.prob_energy, start=1):
                if on:
                    with open(os.path.join(path, f"Ehist_guest{g:02d}.cube"), "w") as fh:
                        fh.write(enprob_text(sysm, zero, zero, eg, sysm.gridfactor))
        return r, stats
    finally:
        h.close()


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("path")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--seed", type=int, nargs=4, default=(12, 34, 56, 78))
    a = ap.parse_args(argv)
    prec = L.FMC_FP64 if a.precision == "fp64" else L.FMC_MIXED
    r, _ = run_directory(a.path, prec, tuple(a.seed))
    print(f"{r.steps} steps, production from step {r.start}, {r.prodcount} counted; report in "
          f"{os.path.join(a.path, 'OUTPUT')}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
