"""TEST INFRASTRUCTURE ONLY: FP64 numpy restatement of the single_point energy
(gcmc.f:2549-2661) of one guest configuration as a function of theta (the
framework charges and the LJ type parameters, Lorentz-Berthelot for every
pair with a framework type), with its analytic gradient: the oracle that
fmc_energy_grad is compared with.  Only tests/ import it.

Terms (FIELD units after / engunit):
  recip   C sum_k a_k |S_k|^2 + self + charged-system term   (ewald1, ewald.f:354-661)
  real    sum_pairs q_i q_j r4pie0/eps erc(r)                  (ewald2, ewald.f:663-733)
  LJ      sum_pairs V_ij(r) from the tables                    (srfrce, vdw_terms.f:382-466)
  tail    2 pi / V sum_{i<=j} (n_i n_j - f_i f_j) eadd_ij      (lrcorrect, vdw_terms.f:581-645)
  - ewald3 of each guest molecule                              (ewald.f:735-790)
Pairs: framework-guest and guest-guest of different molecules (frozen
framework pairs are left out, nlist_builders.f:120).  Cubic / orthorhombic
cells only (the configurations used here).
"""
from __future__ import annotations

import math

import numpy as np

from oracle.oracle import Oracle

R4PIE0 = 138935.4835
PI = 3.141592653589793
SQRPI = 1.7724538509055159
A1, A2, A3, A4, A5, PP = 0.254829592, -0.284496736, 1.421413741, -1.453152027, 1.061405429, 0.3275911


def _interp(tab, r, rstep):
    """3-point interpolation, ewald.f:703-715 (tab indexed from 0, entry 0 unused)."""
    x = r * rstep
    l = x.astype(np.int64)
    p = x - l
    v0, v1, v2 = tab[..., l], tab[..., l + 1], tab[..., l + 2]
    t1 = v0 + (v1 - v0) * p
    t2 = v1 + (v2 - v1) * (p - 1.0)
    return t1 + (t2 - t1) * p * 0.5


class EnergyModel:
    def __init__(self, sysm):
        assert sysm.imcon in (1, 2)
        o = Oracle(sysm)
        lmn, self.akk = o.kvecs()
        o.close()
        self.l, self.m, self.n = (lmn[:, 0].astype(float), lmn[:, 1].astype(float), lmn[:, 2].astype(float))
        self.box = np.array([sysm.cell[0], sysm.cell[4], sysm.cell[8]])
        self.volm = float(np.prod(self.box))
        self.alpha, self.rcut, self.epsq, self.engunit = sysm.alpha, sysm.rcut, 1.0, sysm.engunit
        rvdw = sysm.rvdw if sysm.rvdw > 0 else sysm.rcut
        self.mxegrd = max(1000, int(rvdw / 0.01 + 0.5) + 4)
        self.drewd = self.rcut / (self.mxegrd - 4)
        self.dlrpot = rvdw / (self.mxegrd - 4)
        i = np.arange(1, self.mxegrd + 1, dtype=float)
        rr = i * self.drewd
        tt = 1.0 / (1.0 + PP * self.alpha * rr)
        self.erc = np.zeros(self.mxegrd + 3)
        self.erc[1: self.mxegrd + 1] = tt * (A1 + tt * (A2 + tt * (A3 + tt * (A4 + tt * A5)))) * np.exp(
            -(self.alpha * rr) ** 2) / rr
        self.rgrid = np.concatenate([[1.0], i * self.dlrpot, [1.0, 1.0]])  # entry 0 and the tail unused
        fw = sysm.framework_species
        self.fxyz = sysm.framework_xyz()
        self.ftyp = np.concatenate([np.tile(sp.types, sp.nmols) for sp in fw])
        self.guests = sysm.guests
        self.ntypes = len(sysm.labels)
        self.pairs = {}
        for (a, b, kind, e, s) in sysm.pairs:
            if kind == 2:
                self.pairs[(max(a, b), min(a, b))] = (e, s)
        self.isfw = np.zeros(self.ntypes + 1, dtype=bool)
        self.isfw[np.unique(self.ftyp)] = True
        self.e3 = [self._ewald3(g) for g in self.guests]

    def _ewald3(self, g):
        """ewald3 of one molecule from its template (ewald.f:735-790)."""
        x, q, a = g.template, g.charge, self.alpha
        tot = 0.0
        for i in range(len(q)):
            for j in range(i + 1, len(q)):
                r = float(np.linalg.norm(x[i] - x[j]))
                chg = q[i] / self.epsq * R4PIE0 * q[j]
                ar = r * a
                if ar < 1e-2:
                    erfr = 2 * chg * (a / SQRPI) * (1 + ar * ar * (-1 / 3 + ar * ar * (1 / 10 + ar * ar * (
                        -1 / 42 + ar * ar / 216))))
                else:
                    t = 1.0 / (1.0 + PP * a * r)
                    erfr = (1.0 - t * (A1 + t * (A2 + t * (A3 + t * (A4 + t * A5)))) * math.exp(-ar * ar)) * chg / r
                tot += erfr
        return tot

    def _lj(self, e, s):
        """tables of V, dV/deps, dV/dsigma at the grid points (forgen, vdw_terms.f:222-234)."""
        x6 = (s / self.rgrid) ** 6
        v = 4.0 * e * x6 * (x6 - 1.0)
        de = 4.0 * x6 * (x6 - 1.0)
        ds = 4.0 * e * (12.0 * x6 * x6 - 6.0 * x6) / s
        for t in (v, de, ds):
            t[0] = 0.0
            t[-2:] = 0.0
        return v, de, ds

    def _pair_params(self, ti, tj, teps, tsig):
        if self.isfw[ti] or self.isfw[tj]:
            return math.sqrt(teps[ti - 1] * teps[tj - 1]), 0.5 * (tsig[ti - 1] + tsig[tj - 1]), True
        e, s = self.pairs.get((max(ti, tj), min(ti, tj)), (0.0, 1.0))
        return e, s, False

    def _image(self, d):
        return d - self.box * np.rint(d / self.box)

    def energy_grad(self, q, teps, tsig, nmols, xyz):
        """U (FIELD units), dU/dq [nfw], dU/deps [ntypes], dU/dsig [ntypes]."""
        kq = R4PIE0 / self.epsq
        gx, gq, gt, gm = [], [], [], []
        o = 0
        for gi, g in enumerate(self.guests):
            na = g.natoms
            for mcount in range(nmols[gi]):
                for s in range(na):
                    gx.append(xyz[o]), gq.append(g.charge[s]), gt.append(g.types[s]), gm.append((gi, mcount))
                    o += 1
        gx = np.array(gx).reshape(-1, 3)
        gq, gt = np.array(gq, dtype=float), np.array(gt, dtype=int)
        nfw = len(self.fxyz)
        dq = np.zeros(nfw)
        de, ds = np.zeros(self.ntypes + 1), np.zeros(self.ntypes + 1)
        e2 = ev = 0.0

        def add_lj_grad(ti, tj, sde, sds, teps, e):
            if ti == tj:
                de[ti] += sde
                ds[ti] += sds
                return
            for t in (ti, tj):
                if teps[t - 1] > 0:
                    de[t] += sde * e / (2.0 * teps[t - 1])
                ds[t] += 0.5 * sds

        # framework - guest pairs
        if len(gx):
            d = self._image(gx[None, :, :] - self.fxyz[:, None, :])
            rsq = (d * d).sum(-1)
            inr = rsq < self.rcut ** 2
            r = np.sqrt(np.where(inr, rsq, 1.0))
            er = _interp(self.erc, r, 1.0 / self.drewd)
            chg = (q[:, None] / self.epsq * R4PIE0) * gq[None, :]
            use = inr & (np.abs(chg) > 1e-10)
            e2 += float((chg * er)[use].sum())
            dq += np.where(use, kq * gq[None, :] * er, 0.0).sum(1)
            for ti in np.unique(self.ftyp):
                for tj in np.unique(gt):
                    e, s, _ = self._pair_params(ti, tj, teps, tsig)
                    v, dve, dvs = self._lj(e, s)
                    if abs(v[1]) <= 1e-10:
                        continue
                    sel = inr & (self.ftyp[:, None] == ti) & (gt[None, :] == tj)
                    if not sel.any():
                        continue
                    rs = r[sel]
                    ev += float(_interp(v, rs, 1.0 / self.dlrpot).sum())
                    add_lj_grad(ti, tj, float(_interp(dve, rs, 1.0 / self.dlrpot).sum()),
                                float(_interp(dvs, rs, 1.0 / self.dlrpot).sum()), teps, e)
        # guest - guest pairs (different molecules, each pair once)
        for a in range(len(gx)):
            for b in range(a + 1, len(gx)):
                if gm[a] == gm[b]:
                    continue
                dd = self._image(gx[a] - gx[b])
                rr = float(np.sqrt(dd @ dd))
                if rr * rr >= self.rcut ** 2:
                    continue
                chg = gq[a] / self.epsq * R4PIE0 * gq[b]
                if abs(chg) > 1e-10:
                    e2 += chg * float(_interp(self.erc, np.array([rr]), 1.0 / self.drewd)[0])
                e, s, _ = self._pair_params(gt[a], gt[b], teps, tsig)
                v, _, _ = self._lj(e, s)
                if abs(v[1]) > 1e-10:
                    ev += float(_interp(v, np.array([rr]), 1.0 / self.dlrpot)[0])
        # reciprocal space
        allx = np.concatenate([self.fxyz, gx]) if len(gx) else self.fxyz
        allq = np.concatenate([q, gq])
        sfr = allx / self.box
        ph = 2.0 * PI * (np.outer(sfr[:, 0], self.l) + np.outer(sfr[:, 1], self.m) + np.outer(sfr[:, 2], self.n))
        c, s_ = np.cos(ph), np.sin(ph)
        Sx, Sy = allq @ c, allq @ s_
        C = 2.0 * (2.0 * PI / self.volm) * R4PIE0 / self.epsq
        urec = C * float((self.akk * (Sx * Sx + Sy * Sy)).sum())
        dq += 2.0 * C * (c[:nfw] @ (self.akk * Sx) + s_[:nfw] @ (self.akk * Sy))
        q2 = float((q * q).sum() + (gq * gq).sum())
        qs = float(q.sum() + gq.sum())
        engsic = -kq * self.alpha * q2 / SQRPI
        qfix = -(0.5 * PI * kq) * ((qs / self.alpha) ** 2 / self.volm)
        dq += -2.0 * kq * self.alpha * q / SQRPI - (PI * kq) * qs / (self.alpha ** 2 * self.volm)
        # tail
        nt = self.ntypes
        ncount, fcount = np.zeros(nt + 1), np.zeros(nt + 1)
        for t in self.ftyp:
            ncount[t] += 1
            fcount[t] += 1
        for t in gt:
            ncount[t] += 1
        rc = self.rcut
        elrc = 0.0
        for i in range(1, nt + 1):
            for j in range(1, i + 1):
                w = 2.0 * PI * (ncount[i] * ncount[j] - fcount[i] * fcount[j]) / self.volm
                two = 2.0 if i != j else 1.0
                e, s, th = self._pair_params(i, j, teps, tsig)
                if (i, j) not in self.pairs and not th:
                    continue
                shape = 4.0 * (s ** 12 / (9.0 * rc ** 9) - s ** 6 / (3.0 * rc ** 3))
                elrc += w * two * e * shape
                if th:
                    add_lj_grad(i, j, w * two * shape,
                                w * two * 4.0 * e * (12.0 * s ** 11 / (9.0 * rc ** 9) - 6.0 * s ** 5 / (3.0 * rc ** 3)),
                                teps, e)
        e3 = sum(nmols[gi] * self.e3[gi] for gi in range(len(self.guests)))
        U = (urec + engsic + qfix + e2 - e3 + ev + elrc) / self.engunit
        dmask = self.isfw[1:]
        return U, dq / self.engunit, np.where(dmask, de[1:], 0.0) / self.engunit, np.where(dmask, ds[1:], 0.0) / self.engunit
