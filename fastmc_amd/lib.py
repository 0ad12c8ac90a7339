"""ctypes binding of libfastmc_hip.so (the C-ABI in include/fastmc.h).

The library is the only compute path: if it is missing or fails to load,
every entry point raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .system import (CSystem, FmcGuestStats, FmcNumguestsRec, FmcTheta, FmcTrial, FmcWidomResult, FmcTraceRec,
                     FmcWalker, FmcWalkerStats, FmcWindow, FmcWlParams, FmcWlStatus, TRACE_DTYPE, System, make_walker,
                     prob_plots)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfastmc_hip.so")

FMC_FP64 = 0
FMC_MIXED = 1
FMC_MOVE_INSERT, FMC_MOVE_DELETE, FMC_MOVE_DISPLACE, FMC_MOVE_TRANSLATE = 1, 2, 3, 4
FMC_MOVE_ROTATE, FMC_MOVE_JUMP, FMC_MOVE_SWAP, FMC_MOVE_SWITCH = 5, 6, 7, 8
FMC_STATS_COUNT = 1
FMC_STATS_CLOSE = 2
STAT_NAMES = ("<N>", "<E>", "<E*N>", "<N*N>", "<E*E>", "<NF>", "Qst", "Cv")

EXPORTS = ("fmc_create", "fmc_destroy", "fmc_info", "fmc_single_point", "fmc_full_resum",
           "fmc_trial_insert", "fmc_trial_delete", "fmc_trial_displace", "fmc_accept",
           "fmc_reject", "fmc_get_guests", "fmc_get_sk", "fmc_get_stats", "fmc_reset_stats",
           "fmc_run", "fmc_run_graph", "fmc_last_timing", "fmc_set_block", "fmc_get_block",
           "fmc_set_team", "fmc_get_team",
           "fmc_set_resum", "fmc_wl_setup", "fmc_wl_get", "fmc_wl_window", "fmc_wl_stitch",
           "fmc_wl_result", "fmc_wl_isotherm", "fmc_energy_grad", "fmc_widom", "fmc_set_sampling",
           "fmc_get_windows", "fmc_stats_from_windows", "fmc_production_stats", "fmc_set_prob", "fmc_get_prob",
           "fmc_set_numguests", "fmc_get_numguests", "fmc_load_guests",
           "fmc_set_trace_walker", "fmc_widom_noalloc", "fmc_phase_clock")

_lib = None


class FastMCError(RuntimeError):
    def __init__(self, fn: str, code: int):
        super().__init__(f"{fn} returned error code {code}")
        self.code = code


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libfastmc_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m fastmc_amd._build` "
                           "(or __graft_entry__.build()); there is no fallback path")
    L = C.CDLL(path)
    for name in EXPORTS:
        getattr(L, name).restype = C.c_int
    L.fmc_create.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]
    L.fmc_wl_isotherm.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_int,
                                  C.c_int, C.c_void_p]
    _lib = L
    return L


def _chk(fn, rc):
    if rc != 0:
        raise FastMCError(fn, rc)


def _guest_stats(arr, ntpguest):
    return [{"avg": dict(zip(STAT_NAMES, arr[g].avg)), "std": dict(zip(STAT_NAMES, arr[g].std)),
             "avg_final": dict(zip(STAT_NAMES, arr[g].avg_final)),
             "std_final": dict(zip(STAT_NAMES, arr[g].std_final))} for g in range(ntpguest)]


def stats_from_windows(windows, nwin: int, ntpguest: int, temperature: float):
    """The reference's end-of-run statistics (gcmc.f:1148-1258, avunion /
    stdunion gcmc.f:1591-1664) from closed averaging windows (an FmcWindow
    array): per guest type a dict of avg / std (node report) and avg_final /
    std_final (final report) keyed by STAT_NAMES.  Host arithmetic only."""
    L = load_library()
    out = (FmcGuestStats * ntpguest)()
    _chk("fmc_stats_from_windows", L.fmc_stats_from_windows(windows, C.c_int(nwin), C.c_int(ntpguest),
                                                             C.c_double(temperature), out))
    return _guest_stats(out, ntpguest)


PHASE_NAMES = ("statistics and records at the step's end", "proposal (or commit of a speculative one)",
               "barrier after the proposal", "pair terms (thread 0's share)", "wait for the phase tables",
               "reciprocal terms (thread 0's share)", "wave sums", "barrier after the wave sums", "cross-wave sums",
               "team exchange", "until the decision", "decision", "counters", "barrier after the decision",
               "accept")


def phase_clock(clear: bool = False):
    """k_run's step clock (library built with FMC_PHASE_CLOCK): per phase the
    mean microseconds per step of walker 0, the step count and the steps that
    took a speculative proposal."""
    L = load_library()
    t = (C.c_ulonglong * 18)()
    _chk("fmc_phase_clock", L.fmc_phase_clock(t, 18, 1 if clear else 0))
    steps = max(int(t[16]), 1)
    return {"us_per_step": {PHASE_NAMES[i]: t[i] / steps * 1e-2 for i in range(15)},
            "steps": int(t[16]), "speculative": int(t[17])}


class Handle:
    """A set of walkers (independent GCMC chains) on the GPU sharing one chemistry."""

    def __init__(self, sysm: System, walkers=None, precision: int = FMC_FP64,
                 seeds=None, temperatures=None, pressures=None, framework_xyz=None, n: int = 1):
        L = load_library()
        self.sysm = sysm
        self._keep = []
        self.csys = CSystem(sysm)
        if walkers is None:
            walkers = []
            for i in range(n):
                sd = seeds[i] if seeds is not None else (25 + i % 150, 89, 100, 150)
                t = temperatures[i] if temperatures is not None else None
                p = pressures[i] if pressures is not None else None
                fx = framework_xyz[i] if framework_xyz is not None else None
                walkers.append(make_walker(sysm, sd, t, p, fx, self._keep))
        self.nwalkers = len(walkers)
        arr = (FmcWalker * self.nwalkers)(*walkers)
        self._keep.append(arr)
        h = C.c_void_p()
        _chk("fmc_create", L.fmc_create(C.byref(self.csys.struct), self.nwalkers, arr, precision,
                                        C.byref(h)))
        self.h = h
        self.L = L
        nk, mx, nfw, mm = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _chk("fmc_info", L.fmc_info(self.h, C.byref(nk), C.byref(mx), C.byref(nfw), C.byref(mm)))
        self.newld, self.mxegrd, self.nframework, self.maxmls = nk.value, mx.value, nfw.value, mm.value

    def close(self):
        if getattr(self, "h", None):
            self.L.fmc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- whole-system
    def single_point(self, w: int = 0) -> np.ndarray:
        out = np.zeros(5)
        _chk("fmc_single_point", self.L.fmc_single_point(self.h, w, out.ctypes.data_as(C.c_void_p)))
        return out

    def load_guests(self, iguest: int, xyz, w: int = 0, sumchg=None):
        """Replace walker w's molecules of guest type iguest by the sites xyz
        ([nmols * natoms][3], molecule order) and rebuild counts, sums, S(k)
        and energies from the coordinates (fmc_load_guests)."""
        nat = self.sysm.guests[iguest - 1].natoms
        x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        if len(x) % nat:
            raise ValueError("xyz is not a whole number of molecules")
        q = C.byref(C.c_double(sumchg)) if sumchg is not None else None
        _chk("fmc_load_guests", self.L.fmc_load_guests(self.h, C.c_int(w), C.c_int(iguest),
                                                       x.ctypes.data_as(C.c_void_p), C.c_int(len(x) // nat), q))

    def full_resum(self, w: int = -1) -> float:
        d = C.c_double()
        _chk("fmc_full_resum", self.L.fmc_full_resum(self.h, w, C.byref(d)))
        return d.value

    # ---- seam
    def _trial(self, name, *args):
        t = FmcTrial()
        _chk(name, getattr(self.L, name)(self.h, *args, C.byref(t)))
        return t

    def trial_insert(self, iguest, xyz, w=0):
        x = np.ascontiguousarray(xyz, dtype=np.float64).ravel()
        return self._trial("fmc_trial_insert", C.c_int(w), C.c_int(iguest), x.ctypes.data_as(C.c_void_p))

    def trial_delete(self, iguest, choice, w=0):
        return self._trial("fmc_trial_delete", C.c_int(w), C.c_int(iguest), C.c_int(choice))

    def trial_displace(self, iguest, choice, xyz, w=0):
        x = np.ascontiguousarray(xyz, dtype=np.float64).ravel()
        return self._trial("fmc_trial_displace", C.c_int(w), C.c_int(iguest), C.c_int(choice),
                           x.ctypes.data_as(C.c_void_p))

    def accept(self, w=0):
        _chk("fmc_accept", self.L.fmc_accept(self.h, C.c_int(w)))

    def reject(self, w=0):
        _chk("fmc_reject", self.L.fmc_reject(self.h, C.c_int(w)))

    # ---- device chains
    def run(self, nsteps: int, trace: bool = False, count_stats=False):
        """nsteps steps of every walker; count_stats: False / True (production
        steps) or fmc_run's flags (FMC_STATS_COUNT | FMC_STATS_CLOSE)."""
        tr = np.zeros(nsteps, dtype=TRACE_DTYPE) if trace else None
        ptr = tr.ctypes.data_as(C.c_void_p) if trace else None
        _chk("fmc_run", self.L.fmc_run(self.h, C.c_int(nsteps), ptr, C.c_int(int(count_stats))))
        return tr

    def set_trace_walker(self, w: int):
        """Walker whose steps run(trace=True) records."""
        _chk("fmc_set_trace_walker", self.L.fmc_set_trace_walker(self.h, C.c_int(w)))

    def run_graph(self, nsweeps: int, nsteps_per_sweep: int, count_stats: bool = False):
        _chk("fmc_run_graph", self.L.fmc_run_graph(self.h, C.c_int(nsweeps), C.c_int(nsteps_per_sweep),
                                                   C.c_int(int(count_stats))))

    def last_timing(self):
        ms, n = C.c_double(), C.c_int()
        _chk("fmc_last_timing", self.L.fmc_last_timing(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def set_block(self, threads: int):
        _chk("fmc_set_block", self.L.fmc_set_block(self.h, C.c_int(threads)))

    def set_resum(self, every: int):
        _chk("fmc_set_resum", self.L.fmc_set_resum(self.h, C.c_int(every)))

    def get_block(self) -> int:
        t = C.c_int()
        _chk("fmc_get_block", self.L.fmc_get_block(self.h, C.byref(t)))
        return t.value

    def set_team(self, team: int):
        """Workgroups per walker (1..16); 0 lets set_block(0) try every team size."""
        _chk("fmc_set_team", self.L.fmc_set_team(self.h, C.c_int(team)))

    def get_team(self) -> int:
        t = C.c_int()
        _chk("fmc_get_team", self.L.fmc_get_team(self.h, C.byref(t)))
        return t.value

    # ---- state readback
    def guests(self, iguest: int = 1, w: int = 0, cap: int = 1 << 16) -> np.ndarray:
        nat = self.sysm.guests[iguest - 1].natoms
        buf = np.zeros(3 * nat * cap)
        n = C.c_int()
        _chk("fmc_get_guests", self.L.fmc_get_guests(self.h, C.c_int(w), C.c_int(iguest),
                                                     buf.ctypes.data_as(C.c_void_p), C.c_int(cap),
                                                     C.byref(n)))
        return buf[: 3 * nat * min(n.value, cap)].reshape(-1, 3)

    def sk(self, w: int = 0) -> np.ndarray:
        out = np.zeros(2 * self.newld)
        _chk("fmc_get_sk", self.L.fmc_get_sk(self.h, C.c_int(w), out.ctypes.data_as(C.c_void_p)))
        return out.reshape(-1, 2)

    def stats(self, w: int = 0) -> FmcWalkerStats:
        s = FmcWalkerStats()
        _chk("fmc_get_stats", self.L.fmc_get_stats(self.h, C.c_int(w), C.byref(s)))
        return s

    def reset_stats(self):
        _chk("fmc_reset_stats", self.L.fmc_reset_stats(self.h))

    # ---- production statistics
    def set_sampling(self, nwindsteps: int, accepted_only: bool = False, max_windows: int = 128):
        _chk("fmc_set_sampling", self.L.fmc_set_sampling(self.h, C.c_int(nwindsteps), C.c_int(int(accepted_only)),
                                                         C.c_int(max_windows)))

    def windows(self, w: int = 0):
        """(closed windows of walker w as an FmcWindow array, their count)."""
        n = C.c_int()
        _chk("fmc_get_windows", self.L.fmc_get_windows(self.h, C.c_int(w), None, C.c_int(0), C.byref(n)))
        buf = (FmcWindow * max(n.value, 1))()
        _chk("fmc_get_windows", self.L.fmc_get_windows(self.h, C.c_int(w), buf, C.c_int(n.value), C.byref(n)))
        return buf, n.value

    # ---- numguests.out trace
    def set_numguests(self, nnumg: int, cap: int):
        _chk("fmc_set_numguests", self.L.fmc_set_numguests(self.h, C.c_int(nnumg), C.c_int(cap)))

    def numguests(self, w: int = 0):
        """(records of walker w as an FmcNumguestsRec array, their count)."""
        n = C.c_int()
        _chk("fmc_get_numguests", self.L.fmc_get_numguests(self.h, C.c_int(w), None, C.c_int(0), C.byref(n)))
        buf = (FmcNumguestsRec * max(n.value, 1))()
        _chk("fmc_get_numguests", self.L.fmc_get_numguests(self.h, C.c_int(w), buf, C.c_int(n.value), C.byref(n)))
        return buf, n.value

    # ---- probability plots
    def set_prob(self, griddim: float, gridfactor=(1, 1, 1), plots=()):
        """Probability plots [(iguest, [sites])] (site 0: centre of mass) on a
        grid of spacing griddim; returns the grid size (na, nb, nc)."""
        arr = prob_plots(plots)
        ng = (C.c_int * 3)()
        _chk("fmc_set_prob", self.L.fmc_set_prob(self.h, C.c_double(griddim), (C.c_int * 3)(*gridfactor),
                                                 C.c_int(len(plots)), arr, ng))
        self._ngrid = tuple(ng)
        return self._ngrid

    def prob(self, plot: int, w: int = 0):
        out = np.zeros(int(np.prod(self._ngrid)))
        _chk("fmc_get_prob", self.L.fmc_get_prob(self.h, C.c_int(w), C.c_int(plot),
                                                 out.ctypes.data_as(C.c_void_p)))
        return out

    def production_stats(self, w: int = 0):
        ng = len(self.sysm.guests)
        out = (FmcGuestStats * ng)()
        _chk("fmc_production_stats", self.L.fmc_production_stats(self.h, C.c_int(w), out))
        return _guest_stats(out, ng)

    # ---- energy gradient with respect to framework parameters
    def energy_grad(self, fw_charge, type_eps, type_sig, nmols, xyz, precision: int = FMC_FP64,
                    want_grad: bool = True):
        """U[nconf] and dU/dq[nconf][nfw], dU/deps, dU/dsig [nconf][ntypes] (FIELD units).

        nmols: [nconf][ntpguest] molecule counts; xyz: [nconf][gstride][3] guest sites."""
        q = np.ascontiguousarray(fw_charge, dtype=np.float64)
        e = np.ascontiguousarray(type_eps, dtype=np.float64)
        sg = np.ascontiguousarray(type_sig, dtype=np.float64)
        nm = np.ascontiguousarray(nmols, dtype=np.int32)
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        nconf, gstride = x.shape[0], x.shape[1]
        nt = len(e)
        th = FmcTheta(q.ctypes.data_as(C.POINTER(C.c_double)), e.ctypes.data_as(C.POINTER(C.c_double)),
                      sg.ctypes.data_as(C.POINTER(C.c_double)))
        U = np.zeros(nconf)
        dq = np.zeros((nconf, len(q))) if want_grad else None
        de = np.zeros((nconf, nt)) if want_grad else None
        ds = np.zeros((nconf, nt)) if want_grad else None
        ptr = lambda a: a.ctypes.data_as(C.c_void_p) if a is not None else None
        _chk("fmc_energy_grad", self.L.fmc_energy_grad(self.h, C.byref(th), C.c_int(nconf),
                                                       nm.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p),
                                                       C.c_int(gstride), C.c_int(precision), ptr(U), ptr(dq),
                                                       ptr(de), ptr(ds)))
        return U, dq, de, ds

    # ---- Wang-Landau (wang_landau.f)
    def widom(self, iguest: int, nstep: int, griddim: float, gridfactor=(1, 1, 1), w: int = 0,
              per_attempt: bool = False, energy_map: bool = False, noalloc: bool = False):
        """Widom insertions on a grid (fmc_widom): dict of <exp(-beta estep)>, attempts, non-overlapping
        attempts, Henry's constant as the reference prints it, grid sizes; with per_attempt also the
        estep and overlap arrays in the reference's loop order; with energy_map the binned estep and
        count grids ("emap", "cmap") of the 'e' probability plot.  noalloc: the variant that allocates
        no device memory (fmc_widom_noalloc)."""
        gf = (C.c_int * 3)(*gridfactor)
        res = FmcWidomResult()
        est = ovl = emap = cmap = None
        if per_attempt:
            # size from the same grid rule, so the arrays can be handed over
            n = self._widom_attempts(iguest, nstep, griddim, gridfactor)
            est = np.zeros(n)
            ovl = np.zeros(n, dtype=np.int32)
        if energy_map:
            from .gcmc import prob_grid_size
            npt = int(np.prod(prob_grid_size(self.sysm.cell, griddim, gridfactor)))
            emap, cmap = np.zeros(npt), np.zeros(npt)

        def ptr(a):
            return a.ctypes.data_as(C.c_void_p) if a is not None else None
        fn = "fmc_widom_noalloc" if noalloc else "fmc_widom"
        _chk(fn, getattr(self.L, fn)(self.h, C.c_int(w), C.c_int(iguest), C.c_int(nstep), C.c_double(griddim),
                                     gf, ptr(est), ptr(ovl), ptr(emap), ptr(cmap), C.byref(res)))
        out = {"avg_boltz": res.avg_boltz, "attempts": res.attempts, "accepted": res.accepted, "henry": res.henry,
               "ngrid": tuple(res.ngrid), "wngrid": tuple(res.wngrid)}
        if per_attempt:
            out["estep"], out["overlap"] = est, ovl
        if energy_map:
            out["emap"], out["cmap"] = emap, cmap
        return out

    def _widom_attempts(self, iguest, nstep, griddim, gridfactor):
        import math
        cell = np.asarray(self.sysm.cell, dtype=np.float64).reshape(3, 3)
        n = nstep
        for a in range(3):
            length = float(np.sqrt((cell[a] ** 2).sum()))
            gd = length / math.ceil(length / gridfactor[a] / griddim) / gridfactor[a]
            n *= math.ceil(length / (gd * gridfactor[a]))
        return n

    def wl_setup(self, maxn: int, minn: int = 0, nwindows: int = 1, visittol: int = 10000,
                 flatcoeff: float = 0.7, wlprec: float = float(np.e), prectol: float = 1e-8,
                 create_molecules: int = 0):
        p = FmcWlParams(maxn, minn, nwindows, visittol, flatcoeff, wlprec, prectol, create_molecules, 0)
        _chk("fmc_wl_setup", self.L.fmc_wl_setup(self.h, C.byref(p)))
        self.wl_maxn, self.wl_minn, self.wl_nwin = maxn, minn, nwindows

    def wl_get(self, w: int = 0):
        lng = np.zeros(self.wl_maxn + 1)
        hist = np.zeros(self.wl_maxn + 1, dtype=np.int32)
        st = FmcWlStatus()
        _chk("fmc_wl_get", self.L.fmc_wl_get(self.h, C.c_int(w), lng.ctypes.data_as(C.c_void_p),
                                             hist.ctypes.data_as(C.c_void_p), C.byref(st)))
        return lng, hist, st

    def wl_result(self) -> np.ndarray:
        out = np.zeros(self.wl_maxn + 1)
        _chk("fmc_wl_result", self.L.fmc_wl_result(self.h, out.ctypes.data_as(C.c_void_p)))
        return out


def wl_window(win: int, nwindows: int, maxn: int, minn: int = 0):
    """Window of N for window `win` of `nwindows` (divide_jobs)."""
    L = load_library()
    mn, mx, t = C.c_int(), C.c_int(), C.c_int()
    _chk("fmc_wl_window", L.fmc_wl_window(win, nwindows, maxn, minn, C.byref(mn), C.byref(mx), C.byref(t)))
    return mn.value, mx.value, t.value


def wl_stitch(windows, maxn: int, minn: int = 0) -> np.ndarray:
    """Stitch per-window ln g arrays ([nwin][maxn+1]) into one (stitch_branches)."""
    L = load_library()
    a = np.ascontiguousarray(windows, dtype=np.float64)
    out = np.zeros(maxn + 1)
    _chk("fmc_wl_stitch", L.fmc_wl_stitch(a.ctypes.data_as(C.c_void_p), C.c_int(a.shape[0]), C.c_int(maxn),
                                          C.c_int(minn), out.ctypes.data_as(C.c_void_p)))
    return out


def wl_isotherm(lng, maxn: int, minn: int, temperature: float, lambda3: float, p_pa, units: int = 0):
    """<N>(p) from ln g(N) (compute_isotherm); p in Pa."""
    L = load_library()
    g = np.ascontiguousarray(lng, dtype=np.float64)
    p = np.ascontiguousarray(np.atleast_1d(p_pa), dtype=np.float64)
    out = np.zeros(len(p))
    _chk("fmc_wl_isotherm", L.fmc_wl_isotherm(g.ctypes.data_as(C.c_void_p), maxn, minn, temperature, lambda3,
                                              p.ctypes.data_as(C.c_void_p), len(p), units,
                                              out.ctypes.data_as(C.c_void_p)))
    return out
