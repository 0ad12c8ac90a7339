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
        raise NotImplementedError("cycle-based equilibration or production (negative steps) is not supported")
    s0 = 1 if eqsteps == 0 else minchk * math.ceil(eqsteps / minchk)
    return s0, minchk * math.ceil((s0 + mcsteps) / minchk)


class Run:
    """Steps done and production state of one run of the main loop."""

    def __init__(self):
        self.steps = 0
        self.start = 0          # first production step (0: not started)
        self.prodcount = 0


def run_main_loop(engine, eqsteps: int, mcsteps: int, minchk: int, accepted_only: bool = False,
                  prodcount=None) -> Run:
    """Drive `engine.run(nsteps, flags)` (flags as fmc_run's count_stats)
    through the reference's main-loop schedule.  With every production step
    counted the schedule is known up front (three launches at most); with
    accepted_only the run ends on the production count, which `prodcount()`
    reads at every check (gcmc.f:707-752)."""
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
    if eqsteps < 0 or mcsteps <= 0:
        raise NotImplementedError("cycle-based equilibration or production (negative steps) is not supported")
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
    """Fortran Ew.d: 0.dddE+xx."""
    if x == 0.0 or not math.isfinite(x):
        return _f(x, w, d) if not math.isfinite(x) else f"0.{'0' * d}E+00".rjust(w)
    e = math.floor(math.log10(abs(x))) + 1
    m = abs(x) / 10.0 ** e
    if round(m, d) >= 1.0:
        m /= 10.0
        e += 1
    s = f"{'-' if x < 0 else ''}0.{int(round(m * 10 ** d)):0{d}d}E{'+' if e >= 0 else '-'}{abs(e):02d}"
    return s.rjust(w)


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


def run_directory(path: str, precision: int = L.FMC_FP64, seed=(12, 34, 56, 78), out_name: str = "OUTPUT",
                  max_windows: int = 256):
    """One GCMC run of the FIELD / CONFIG / CONTROL in `path` on the device;
    writes the report to path/out_name and returns (Run, statistics)."""
    sysm: System = read_system(path)
    h = L.Handle(sysm, precision=precision, seeds=[tuple(seed)])
    try:
        sp = h.single_point()
        nwind, nwindsteps = window_steps(sysm.steps, sysm.nwind)
        minchk = check_interval(sysm.eqsteps, sysm.nnumg)
        h.set_sampling(nwindsteps, sysm.accsample, max_windows)
        eng = _Engine(h)
        t0 = time.perf_counter()
        r = run_main_loop(eng, sysm.eqsteps, sysm.steps, minchk, sysm.accsample, eng.prodcount)
        seconds = time.perf_counter() - t0
        st = h.stats(0)
        stats = h.production_stats(0)
        names = [species_name(s) for s in sysm.guests]
        lines = [f"Averages will be computed from {nwind:6d} averaging windows. "
                 f"Each window will average over {nwindsteps:10d} steps (or cycles)."]
        ens = [float(st.energy[i]) for i, s in enumerate(sysm.species) if s.is_guest]
        lines += energy_report(ens, names, sp[3], sp[0], seed)
        if r.start > 1:
            lines += [f"Completed at least {sysm.eqsteps:10d} equlibration steps",
                      f"Starting production at {r.start:10d}"]
        lines += node_report(names, stats, int(st.prodcount))
        lines += move_report(r.steps, seconds, list(st.attempted), list(st.accepted))
        lines += final_report(names, stats, int(st.prodcount), sysm.surftol)
        with open(os.path.join(path, out_name), "w") as fh:
            fh.write("\n".join(lines) + "\n")
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
