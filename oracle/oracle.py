"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the C oracle (liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only to compare results with.  The product package never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from fastmc_amd.system import (CSystem, FmcNumguestsRec, FmcTrial, FmcTraceRec, FmcWindow, TRACE_DTYPE, make_walker,
                               prob_plots)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE], stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_destroy.argtypes = [C.c_void_p]
        for name in ("orc_info", "orc_single_point", "orc_trial_insert", "orc_trial_delete",
                     "orc_trial_displace", "orc_accept", "orc_reject", "orc_run", "orc_get_guests",
                     "orc_get_sk", "orc_get_table", "orc_get_kvec", "orc_get_energy",
                     "orc_full_resum", "orc_wl_setup", "orc_wl_run", "orc_wl_get", "orc_run_stats",
                     "orc_set_sampling", "orc_get_windows", "orc_get_surface", "orc_set_prob", "orc_get_prob", "orc_set_numguests",
                     "orc_get_numguests"):
            getattr(L, name).restype = C.c_int
        L.orc_duni_seq.restype = None
        L.orc_wl_setup.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                   C.c_int, C.c_double]
        _lib = L
    return _lib


def duni_seq(seed, n):
    out = np.zeros(n)
    s = (C.c_int * 4)(*seed)
    lib().orc_duni_seq(s, C.c_int(n), out.ctypes.data_as(C.c_void_p))
    return out


class Oracle:
    """One reference-faithful FP64 chain on the CPU."""

    def __init__(self, sysm, seed=(25, 89, 100, 150), temperature=None, pressures=None,
                 framework_xyz=None):
        self.sysm = sysm
        self._keep = []
        self.csys = CSystem(sysm)
        self.walker = make_walker(sysm, seed, temperature, pressures, framework_xyz, self._keep)
        self.h = lib().orc_create(C.byref(self.csys.struct), C.byref(self.walker))
        if not self.h:
            raise RuntimeError("orc_create failed")
        newld, mx, tot, mm = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        lib().orc_info(C.c_void_p(self.h), C.byref(newld), C.byref(mx), C.byref(tot), C.byref(mm))
        self.newld, self.mxegrd, self.totatm, self.maxmls = newld.value, mx.value, tot.value, mm.value

    def close(self):
        if self.h:
            lib().orc_destroy(C.c_void_p(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def widom(self, iguest, nstep, griddim, gridfactor=(1, 1, 1), max_attempts=0, energy_map=False):
        """widom_grid (gcmc.f:2454-2548): dict of <exp(-beta estep)>, attempts,
        non-overlapping attempts, Henry's constant as printed, grid sizes; with
        max_attempts > 0 (at least the attempt count) also per-attempt estep and
        overlap arrays."""
        out = np.zeros(10)
        gf = (C.c_int * 3)(*gridfactor)
        est = np.zeros(max(max_attempts, 1))
        ovl = np.zeros(max(max_attempts, 1), dtype=np.int32)
        emap = cmap = None
        if energy_map:
            from fastmc_amd.gcmc import prob_grid_size
            npt = int(np.prod(prob_grid_size(self.sysm.cell, griddim, gridfactor)))
            emap, cmap = np.zeros(npt), np.zeros(npt)
        rc = lib().orc_widom(C.c_void_p(self.h), C.c_int(iguest), C.c_int(nstep), C.c_double(griddim), gf,
                             out.ctypes.data_as(C.c_void_p), est.ctypes.data_as(C.c_void_p) if max_attempts else None,
                             ovl.ctypes.data_as(C.c_void_p) if max_attempts else None,
                             emap.ctypes.data_as(C.c_void_p) if energy_map else None,
                             cmap.ctypes.data_as(C.c_void_p) if energy_map else None)
        if rc:
            raise RuntimeError(f"orc_widom: {rc}")
        r = {"avg_boltz": out[0], "attempts": int(out[1]), "accepted": int(out[2]), "henry": out[3],
             "ngrid": tuple(int(x) for x in out[4:7]), "wngrid": tuple(int(x) for x in out[7:10])}
        if max_attempts:
            assert max_attempts >= r["attempts"]
            r["estep"], r["overlap"] = est[: r["attempts"]], ovl[: r["attempts"]]
        if energy_map:
            r["emap"], r["cmap"] = emap, cmap
        return r

    def move_counts(self):
        """(attempts, acceptances) by move kind FMC_MOVE_* (index 1..8), summed over guests."""
        att = np.zeros(9, dtype=np.int64)
        acc = np.zeros(9, dtype=np.int64)
        lib().orc_move_counts(C.c_void_p(self.h), att.ctypes.data_as(C.c_void_p), acc.ctypes.data_as(C.c_void_p))
        return att, acc

    def single_point(self):
        out = np.zeros(5)
        lib().orc_single_point(C.c_void_p(self.h), out.ctypes.data_as(C.c_void_p))
        return out

    def run(self, nsteps, stats=0):
        """nsteps steps of the chain; stats as fmc_run's count_stats flags
        (1: production steps, 2: close the open window after the last)."""
        tr = np.zeros(nsteps, dtype=TRACE_DTYPE)
        rc = lib().orc_run_stats(C.c_void_p(self.h), C.c_int(nsteps), tr.ctypes.data_as(C.c_void_p),
                                 C.c_int(stats))
        if rc:
            raise RuntimeError(f"orc_run rc={rc}")
        return tr

    def set_sampling(self, nwindsteps, accepted_only=False):
        rc = lib().orc_set_sampling(C.c_void_p(self.h), C.c_int(nwindsteps), C.c_int(int(accepted_only)))
        if rc:
            raise RuntimeError(f"orc_set_sampling rc={rc}")

    def windows(self):
        """(closed windows as an FmcWindow array, production count)."""
        n, pc = C.c_int(), C.c_longlong()
        lib().orc_get_windows(C.c_void_p(self.h), None, C.c_int(0), C.byref(n), C.byref(pc))
        buf = (FmcWindow * max(n.value, 1))()
        lib().orc_get_windows(C.c_void_p(self.h), buf, C.c_int(n.value), C.byref(n), C.byref(pc))
        return buf, n.value, pc.value

    def set_prob(self, ngrid, gridfactor, plots):
        """Probability plots: plots = [(iguest, [sites])] (site 0: centre of mass)."""
        arr = prob_plots(plots)
        rc = lib().orc_set_prob(C.c_void_p(self.h), (C.c_int * 3)(*ngrid), (C.c_int * 3)(*gridfactor),
                                C.c_int(len(plots)), arr)
        if rc:
            raise RuntimeError(f"orc_set_prob rc={rc}")
        self._ngrid = tuple(ngrid)

    def prob(self, plot):
        out = np.zeros(int(np.prod(self._ngrid)))
        lib().orc_get_prob(C.c_void_p(self.h), C.c_int(plot), out.ctypes.data_as(C.c_void_p))
        return out

    def set_numguests(self, nnumg):
        if lib().orc_set_numguests(C.c_void_p(self.h), C.c_int(nnumg)):
            raise RuntimeError("orc_set_numguests")

    def numguests(self):
        n = C.c_int()
        lib().orc_get_numguests(C.c_void_p(self.h), None, C.c_int(0), C.byref(n))
        buf = (FmcNumguestsRec * max(n.value, 1))()
        lib().orc_get_numguests(C.c_void_p(self.h), buf, C.c_int(n.value), C.byref(n))
        return buf, n.value

    def surface(self):
        out = np.zeros(8, dtype=np.int32)
        lib().orc_get_surface(C.c_void_p(self.h), out.ctypes.data_as(C.c_void_p))
        return out

    def _trial(self, fn, *args):
        t = FmcTrial()
        rc = fn(C.c_void_p(self.h), *args, C.byref(t))
        if rc:
            raise RuntimeError(f"oracle trial rc={rc}")
        return t

    def trial_insert(self, iguest, xyz):
        x = np.ascontiguousarray(xyz, dtype=np.float64).ravel()
        return self._trial(lib().orc_trial_insert, C.c_int(iguest), x.ctypes.data_as(C.c_void_p))

    def trial_delete(self, iguest, choice):
        return self._trial(lib().orc_trial_delete, C.c_int(iguest), C.c_int(choice))

    def trial_displace(self, iguest, choice, xyz):
        x = np.ascontiguousarray(xyz, dtype=np.float64).ravel()
        return self._trial(lib().orc_trial_displace, C.c_int(iguest), C.c_int(choice),
                           x.ctypes.data_as(C.c_void_p))

    def accept(self):
        return lib().orc_accept(C.c_void_p(self.h))

    def reject(self):
        return lib().orc_reject(C.c_void_p(self.h))

    def guests(self, iguest=1, cap=100000):
        nat = self.sysm.guests[iguest - 1].natoms
        buf = np.zeros(3 * nat * 4096)
        n = C.c_int()
        lib().orc_get_guests(C.c_void_p(self.h), C.c_int(iguest), buf.ctypes.data_as(C.c_void_p),
                             C.c_int(4096), C.byref(n))
        return buf[: 3 * nat * n.value].reshape(-1, 3)

    def guest_total(self):
        """Molecules of every guest type together."""
        n, tot = C.c_int(), 0
        buf = np.zeros(3)
        for g in range(1, len(self.sysm.guests) + 1):
            lib().orc_get_guests(C.c_void_p(self.h), C.c_int(g), buf.ctypes.data_as(C.c_void_p), C.c_int(0),
                                 C.byref(n))
            tot += n.value
        return tot

    def sk(self):
        out = np.zeros(2 * self.newld)
        lib().orc_get_sk(C.c_void_p(self.h), out.ctypes.data_as(C.c_void_p))
        return out.reshape(-1, 2)

    def energy(self):
        out = np.zeros(self.maxmls)
        lib().orc_get_energy(C.c_void_p(self.h), out.ctypes.data_as(C.c_void_p))
        return out

    def table(self, which, k=0, n=None):
        n = n or self.mxegrd
        out = np.zeros(n)
        rc = lib().orc_get_table(C.c_void_p(self.h), C.c_int(which), C.c_int(k),
                                 out.ctypes.data_as(C.c_void_p), C.c_int(n))
        if rc:
            raise RuntimeError("bad table")
        return out

    def kvecs(self):
        lmn = np.zeros(3 * self.newld, dtype=np.int32)
        akk = np.zeros(self.newld)
        lib().orc_get_kvec(C.c_void_p(self.h), lmn.ctypes.data_as(C.c_void_p),
                           akk.ctypes.data_as(C.c_void_p))
        return lmn.reshape(-1, 3), akk

    def full_resum(self):
        d = C.c_double()
        lib().orc_full_resum(C.c_void_p(self.h), C.byref(d))
        return d.value


# ---------------------------------------------------------------- Wang-Landau
# wang_landau.f restated for the checks: the walker itself is in the C oracle
# (orc_wl_setup / orc_wl_run); the window split, stitching and isotherm are
# restated here in numpy.

def _wl_methods():
    def wl_setup(self, minmol, maxmol, maxn, preload=0, wlprec=np.e, flat=0.7, visittol=10000, prectol=1e-8):
        rc = lib().orc_wl_setup(C.c_void_p(self.h), minmol, maxmol, maxn, preload, wlprec, flat, visittol,
                                prectol)
        if rc:
            raise RuntimeError(f"orc_wl_setup rc={rc}")
        self.wl_maxn = maxn

    def wl_run(self, nsteps):
        tr = np.zeros(nsteps, dtype=TRACE_DTYPE)
        rc = lib().orc_wl_run(C.c_void_p(self.h), C.c_int(nsteps), tr.ctypes.data_as(C.c_void_p))
        if rc:
            raise RuntimeError(f"orc_wl_run rc={rc}")
        return tr

    def wl_get(self):
        nb = self.wl_maxn + 2
        lng, hist = np.zeros(nb), np.zeros(nb, dtype=np.int32)
        ist, dst = np.zeros(5, dtype=np.int64), np.zeros(3)
        lib().orc_wl_get(C.c_void_p(self.h), lng.ctypes.data_as(C.c_void_p), hist.ctypes.data_as(C.c_void_p),
                         ist.ctypes.data_as(C.c_void_p), dst.ctypes.data_as(C.c_void_p))
        st = {"done": int(ist[0]), "error": int(ist[1]), "sweeps": int(ist[2]), "minmol": int(ist[3]),
              "maxmol": int(ist[4]), "f": dst[0], "lnf": dst[1], "lambda3": dst[2]}
        return lng[: self.wl_maxn + 1], hist[: self.wl_maxn + 1], st

    Oracle.wl_setup, Oracle.wl_run, Oracle.wl_get = wl_setup, wl_run, wl_get


_wl_methods()


def wl_window(win, nwindows, maxn, minn):
    """divide_jobs, wang_landau.f:1742-1781 (integer arithmetic)."""
    tail = 0 if nwindows == 1 else 3
    nodex = win + 1
    mx = ((maxn + 1) * nodex ** 2 + (maxn + 1) // nwindows * nodex ** 2) // (nodex ** 2 + nodex + 1) + tail
    mn = ((maxn + 1) * win ** 2 + (maxn + 1) // nwindows * win ** 2) // (win ** 2 + win + 1) - tail
    if nodex == 1:
        mn = minn
    if nodex == nwindows:
        mx = maxn
    return max(mn, 0), min(mx, maxn), tail


def wl_stitch(windows, maxn, minn):
    """stitch_branches, wang_landau.f:1548-1621, over per-window ln g arrays [nwin][maxn+1]."""
    nwin = len(windows)
    dos = np.array(windows[0], dtype=np.float64).copy()
    mn0, mx0, tail = wl_window(0, nwin, maxn, minn)
    shift = dos[0]
    dos[: mx0 - mn0] -= shift
    for w in range(1, nwin):
        buf = np.asarray(windows[w], dtype=np.float64)
        mn, mx, tail = wl_window(w, nwin, maxn, minn)
        const = dos[mn + 1] - buf[mn + 1]
        for ik, j in enumerate(range(mn, mx + 1), start=2):
            if 2 <= ik <= 2 * tail - 1:
                dos[j] = (dos[j] + buf[j] + const) / 2.0
            else:
                dos[j] = buf[j] + const
    return dos


KBOLTZ = 1.9872041e-3
BOLTZ = 1.3806503e-23


def wl_isotherm(lng, maxn, minn, temperature, lambda3, p_pa, units=0):
    """compute_isotherm + grand_canonical_partition, wang_landau.f:1623-1740."""
    beta = 1.0 / (KBOLTZ * temperature)
    out = []
    for p in np.atleast_1d(p_pa):
        if units == 0:
            mu = np.log(p * lambda3 * beta) / beta
            bmu = beta * mu
        else:
            bmu = np.log(p * lambda3 * 1e-30 / (BOLTZ * temperature))
        i = np.arange(0, maxn - minn + 1)
        zt = lng[i] + bmu * (i - minn)
        k = np.arange(1, maxn)
        nt = np.log(k) + lng[k] + bmu * k
        lz = zt.max() + np.log(np.exp(zt - zt.max()).sum())
        ln = nt.max() + np.log(np.exp(nt - nt.max()).sum())
        out.append(np.exp(ln - lz))
    return np.array(out)


def reference_pressures():
    """compute_isotherm's pressure grid: (ip - 1 + 1e-5) * 0.1 bar, ip = 1..51, in Pa."""
    return np.array([((ip - 1) + 1e-5) * 1e-1 * 1e5 for ip in range(1, 52)])
