// Kernels of the GCMC trial-move path.  Template parameters: NS = number of
// molecule species (maxmls, compile-time so the per-species accumulators live
// in registers) and T = arithmetic type of the pair and phase products
// (double: FP64 parity mode; float: mixed mode with FP64 S(k), FP64
// coordinate differences and FP64 accumulation).
#pragma once

#include "fmc_device.hpp"

namespace fmc {

template <int NS>
struct Acc {
  static constexpr int NACC = 6 * NS + 2;
  // layout: [0,NS) real_new, [NS,2NS) vdw_new, [2NS,3NS) real_old, [3NS,4NS) vdw_old,
  //         4NS rec_new, 4NS+1 rec_old, [4NS+2,5NS+2) mix_new, [5NS+2,6NS+2) mix_old
};

template <int NS>
struct Smem {
  WState ws;
  int kind, g, mol, natms, choice, nmols, has_new, has_old, excl_lo, excl_hi, nfw;
  int accepted, drew, ovl_new, ovl_old;
  double xn[KSITE][3], xo[KSITE][3];
  double2 ph[2][KSITE][3][KCAP];
  double red[KWAVES][Acc<NS>::NACC];
  double tot[Acc<NS>::NACC];
  int ovlw[KWAVES][2];
  double estep, rande, delrc;
  double delE[KMAXS + 1], delrc_mol[KMAXC + 1];
  double comp[4];
};

// ------------------------------------------------------------ tail term --
// gstlrcorrect, vdw_terms.f:686-776, in its linear form (coefficients built on
// the host from the same expressions): total delrc and the loc2(mol, jt)
// entries of delrc_mol.
template <int NS>
__device__ double gstlrcorrect(const DevParams &P, const WState &ws, int g, int lexisting,
                               double *delrc_mol) {
  int mol = P.locguest[g];
  double delrc = P.tl_c0[g][lexisting];
  for (int gg = 1; gg <= P.ntpguest; ++gg) delrc += P.tl_c1[g][lexisting][gg] * (double)ws.nmols[P.locguest[gg]];
  for (int t = 0; t <= KMAXC; ++t) delrc_mol[t] = 0.0;
#pragma unroll
  for (int jt = 1; jt <= NS; ++jt)
    delrc_mol[loc2(mol, jt)] = P.tm_c0[g][lexisting][jt] + P.tm_c1[g][lexisting][jt] * (double)ws.nmols[jt];
  return delrc;
}

// ------------------------------------------------------------- geometry --
// com, utility_pack.f:3565-3596
__device__ inline void com_of(const DevParams &P, int mol, int natms, const double (*x)[3],
                              double *c) {
  double wsum = 0.0, ax = 0.0, ay = 0.0, az = 0.0;
  for (int i = 0; i < natms; ++i) {
    double m = P.site_m[mol][i];
    wsum += m;
    ax += x[i][0] * m;
    ay += x[i][1] * m;
    az += x[i][2] * m;
  }
  c[0] = ax / wsum;
  c[1] = ay / wsum;
  c[2] = az / wsum;
}

// rotation, utility_pack.f:3480-3524
__device__ inline void rotate_about(double (*x)[3], int natms, const double *c, const double *q) {
  double Q[9];
  double q1 = q[0], q2 = q[1], q3 = q[2], q4 = q[3];
  Q[0] = 2.0 * q2 * q2 + 2.0 * q1 * q1 - 1.0;
  Q[1] = 2.0 * q2 * q3 - 2.0 * q1 * q4;
  Q[2] = 2.0 * q2 * q4 + 2.0 * q1 * q3;
  Q[3] = 2.0 * q2 * q3 + 2.0 * q1 * q4;
  Q[4] = 2.0 * q1 * q1 + 2.0 * q3 * q3 - 1.0;
  Q[5] = 2.0 * q3 * q4 - 2.0 * q1 * q2;
  Q[6] = 2.0 * q2 * q4 - 2.0 * q1 * q3;
  Q[7] = 2.0 * q3 * q4 + 2.0 * q1 * q2;
  Q[8] = 2.0 * q4 * q4 + 2.0 * q1 * q1 - 1.0;
  for (int i = 0; i < natms; ++i) {
    double a = x[i][0] - c[0], b = x[i][1] - c[1], d = x[i][2] - c[2];
    double tx = Q[0] * a + Q[1] * b + Q[2] * d;
    double ty = Q[3] * a + Q[4] * b + Q[5] * d;
    double tz = Q[6] * a + Q[7] * b + Q[8] * d;
    x[i][0] = tx + c[0];
    x[i][1] = ty + c[1];
    x[i][2] = tz + c[2];
  }
}

// random_ins / random_rot quaternion, utility_pack.f:3291-3310,3544-3560
__device__ inline void quat_from(double r1, double r2, double r3, double r4, double scale, double *q) {
  double theta = r1 * scale;
  double b1 = 2.0 * r2 - 1.0, b2 = 2.0 * r3 - 1.0, b3 = 2.0 * r4 - 1.0;
  double st = sin(theta / 2.0);
  double norm = sqrt(b1 * b1 + b2 * b2 + b3 * b3);
  q[0] = cos(theta / 2.0);
  q[1] = st * (b1 / norm);
  q[2] = st * (b2 / norm);
  q[3] = st * (b3 / norm);
}

template <int NS>
__device__ inline void load_molecule(const DevParams &P, const DevWalkers &W, int w, int g, int choice,
                                     double (*x)[3], int natms) {
  size_t base = (size_t)w * P.gatoms_total + P.gbase[g] + (size_t)(choice - 1) * natms;
  for (int i = 0; i < natms; ++i) {
    x[i][0] = W.gx[base + i];
    x[i][1] = W.gy[base + i];
    x[i][2] = W.gz[base + i];
  }
}

// ------------------------------------------------------------- proposal --
// Main-loop move choice and step-size adaptation (gcmc.f:805-911), then the
// trial geometry of mc_insert / mc_delete / mc_displace (gcmc.f:1693-1935)
// with random_ins, random_disp, random_rot, translate and rotation
// (utility_pack.f:3161-3563).  Thread 0 only.
template <int NS>
__device__ void propose(const DevParams &P, const DevWalkers &W, int w, Smem<NS> &sh) {
  WState &ws = sh.ws;
  int g = 1;
  if (P.ntpguest > 1) g = (int)floor(duni(ws) * P.ntpguest) + 1;
  int mol = P.locguest[g], natms = P.natoms[mol], nmols = ws.nmols[mol];
  double randmov = duni(ws);
  const float *f = ws.mcf[g];
  int kind;
  if (randmov < (double)f[0]) kind = FMC_MOVE_INSERT;
  else if (randmov < (double)f[1]) kind = FMC_MOVE_DELETE;
  else if (randmov < (double)f[2]) kind = FMC_MOVE_DISPLACE;
  else if (randmov < (double)f[3]) kind = FMC_MOVE_DISPLACE;  // jump: not on this path, never configured
  else if (randmov < (double)f[4]) kind = FMC_MOVE_DISPLACE;  // swap: idem
  else if (randmov < (double)f[5]) kind = FMC_MOVE_TRANSLATE;
  else if (randmov < (double)f[6]) kind = FMC_MOVE_ROTATE;
  else kind = FMC_MOVE_DISPLACE;
  if (nmols >= 1 && ws.disp_count[g] >= 1) {
    if (ws.disp_count[g] % 100 == 0) {
      if ((ws.accept_disp[g] / (double)ws.disp_count[g]) > 0.5) ws.delrdisp[g] *= 1.05;
      else if (ws.delrdisp[g] > DELRMIN) ws.delrdisp[g] *= 0.95;
    } else if (ws.disp_count[g] % 50 == 0) {
      if ((ws.accept_disp[g] / (double)ws.disp_count[g]) > 0.5) ws.rotangle *= 1.05;
      else if (ws.rotangle > MINANGLE) ws.rotangle *= 0.95;
    }
  }
  if (nmols >= 1 && ws.tran_count[g] >= 1 && ws.tran_count[g] % 100 == 0) {
    if ((ws.accept_tran[g] / (double)ws.tran_count[g]) > 0.5) ws.tran_delr[g] *= 1.05;
    else if (ws.tran_delr[g] > DELRMIN) ws.tran_delr[g] *= 0.95;
  }
  if (nmols >= 1 && ws.rota_count[g] >= 1 && ws.rota_count[g] % 100 == 0) {
    if ((ws.accept_rota[g] / (double)ws.rota_count[g]) > 0.5) ws.rota_rotangle *= 1.05;
    else if (ws.rota_rotangle > MINANGLE) ws.rota_rotangle *= 0.95;
  }
  if (nmols == 0) kind = FMC_MOVE_INSERT;

  sh.kind = kind;
  sh.g = g;
  sh.mol = mol;
  sh.natms = natms;
  sh.nmols = nmols;
  sh.choice = 0;
  sh.excl_lo = sh.excl_hi = 0;
  if (kind == FMC_MOVE_INSERT) {
    double ra = duni(ws), rb = duni(ws), rc = duni(ws);
    double r1 = duni(ws), r2 = duni(ws), r3 = duni(ws), r4 = duni(ws);
    const double *cl = P.cell;
    double xc = ra * cl[0] + rb * cl[3] + rc * cl[6];
    double yc = ra * cl[1] + rb * cl[4] + rc * cl[7];
    double zc = ra * cl[2] + rb * cl[5] + rc * cl[8];
    for (int i = 0; i < natms; ++i) {
      sh.xn[i][0] = P.tmpl[g][i][0];
      sh.xn[i][1] = P.tmpl[g][i][1];
      sh.xn[i][2] = P.tmpl[g][i][2];
    }
    double q[4], c0[3] = {0.0, 0.0, 0.0};
    quat_from(r1, r2, r3, r4, 2.0 * PI, q);
    rotate_about(sh.xn, natms, c0, q);
    for (int i = 0; i < natms; ++i) {
      sh.xn[i][0] += xc;
      sh.xn[i][1] += yc;
      sh.xn[i][2] += zc;
    }
    sh.has_new = 1;
    sh.has_old = 0;
  } else if (kind == FMC_MOVE_DELETE) {
    int choice = (int)floor(duni(ws) * nmols) + 1;
    sh.choice = choice;
    load_molecule<NS>(P, W, w, g, choice, sh.xo, natms);
    sh.has_new = 0;
    sh.has_old = 1;
    sh.excl_lo = (choice - 1) * natms;
    sh.excl_hi = choice * natms;
  } else {
    double a = 0.0, b = 0.0, c = 0.0, q[4] = {0.0, 0.0, 0.0, 0.0};
    if (kind == FMC_MOVE_TRANSLATE) {
      ws.tran_count[g] += 1;
      double h = ws.tran_delr[g] * 0.5;
      a = (2.0 * duni(ws) - 1.0) * h;
      b = (2.0 * duni(ws) - 1.0) * h;
      c = (2.0 * duni(ws) - 1.0) * h;
    } else if (kind == FMC_MOVE_ROTATE) {
      ws.rota_count[g] += 1;
      double r1 = duni(ws), r2 = duni(ws), r3 = duni(ws), r4 = duni(ws);
      quat_from(r1, r2, r3, r4, ws.rota_rotangle, q);
    } else {
      ws.disp_count[g] += 1;
      double h = ws.delrdisp[g] * 0.5;
      double da = duni(ws), db = duni(ws), dc = duni(ws);
      a = (2.0 * da - 1.0) * h;
      b = (2.0 * db - 1.0) * h;
      c = (2.0 * dc - 1.0) * h;
      double r1 = duni(ws), r2 = duni(ws), r3 = duni(ws), r4 = duni(ws);
      quat_from(r1, r2, r3, r4, ws.rotangle, q);
    }
    int choice = (int)floor(duni(ws) * nmols) + 1;
    sh.choice = choice;
    load_molecule<NS>(P, W, w, g, choice, sh.xo, natms);
    // translate, utility_pack.f:3161-3218
    for (int i = 0; i < natms; ++i) {
      sh.xn[i][0] = a + sh.xo[i][0];
      sh.xn[i][1] = b + sh.xo[i][1];
      sh.xn[i][2] = c + sh.xo[i][2];
    }
    double cm[3];
    com_of(P, mol, natms, sh.xn, cm);
    const double *rc = P.rcell, *cl = P.cell;
    double ssx = rc[0] * cm[0] + rc[3] * cm[1] + rc[6] * cm[2];
    double ssy = rc[1] * cm[0] + rc[4] * cm[1] + rc[7] * cm[2];
    double ssz = rc[2] * cm[0] + rc[5] * cm[1] + rc[8] * cm[2];
    double xss = ssx - floor(ssx), yss = ssy - floor(ssy), zss = ssz - floor(ssz);
    double x = cl[0] * xss + cl[3] * yss + cl[6] * zss;
    double y = cl[1] * xss + cl[4] * yss + cl[7] * zss;
    double z = cl[2] * xss + cl[5] * yss + cl[8] * zss;
    double sx = x - cm[0], sy = y - cm[1], sz = z - cm[2];
    cm[0] = x;
    cm[1] = y;
    cm[2] = z;
    for (int i = 0; i < natms; ++i) {
      sh.xn[i][0] += sx;
      sh.xn[i][1] += sy;
      sh.xn[i][2] += sz;
    }
    rotate_about(sh.xn, natms, cm, q);
    sh.has_new = 1;
    sh.has_old = 1;
    sh.excl_lo = (choice - 1) * natms;
    sh.excl_hi = choice * natms;
  }
}

// ----------------------------------------------------------- evaluation --
template <int NS, typename T>
__device__ inline void pair_site(const DevParams &P, const double *erc, const double *xs, double chgea, int ts, double xj,
                                 double yj, double zj, double qj, int tj, int spj, double *acc_real,
                                 double *acc_vdw, int &ovl) {
  // distance, cutoff and overlap tests and the real-space Coulomb term stay
  // in FP64 in both modes (pair Coulomb terms are large and cancel); T only
  // sets the precision of the LJ table interpolation, whose fractional
  // position is formed in FP64 first
  double dx = xs[0] - xj, dy = xs[1] - yj, dz = xs[2] - zj;
  min_image(P, dx, dy, dz);
  double rsq = dx * dx + dy * dy + dz * dz;
  if (rsq < P.overlap2) ovl = 1;
  if (rsq < P.rcut2) {
    double r = sqrt(rsq);
    double chgprd = chgea * qj;
    double e2 = 0.0, ev = 0.0;
    if (fabs(chgprd) > 1e-10) e2 = interp3<double>(erc, r, P.rdrewd) * chgprd;
    int k = P.pairidx[ts][tj];
    if (k >= 0) {
      double x = r * P.rdlrpot;
      int l = (int)x;
      T pp = (T)(x - (double)l);
      const T *vv;
      if constexpr (sizeof(T) == 8) vv = (const T *)P.vvv; else vv = (const T *)P.vvv32;
      vv += (size_t)k * P.vstride + l;
      T v0 = vv[0], v1 = vv[1], v2 = vv[2];
      T t1 = v0 + (v1 - v0) * pp;
      T t2 = v1 + (v2 - v1) * (pp - (T)1);
      ev = (double)(t1 + (t2 - t1) * pp * (T)0.5);
    }
#pragma unroll
    for (int i = 0; i < NS; ++i)
      if (spj == i + 1) {
        acc_real[i] += e2;
        acc_vdw[i] += ev;
      }
  }
}

// One reciprocal-space phase product for site s of configuration cfg at
// (ll, mm, nn): the reference's clm/slm then ckc/cks (ewald.f:191-251).
template <typename T>
__device__ inline void site_phase(const double2 (*ph)[KCAP], int ll, int mm, int nn, T q, T &ckc, T &cks) {
  int m = mm < 0 ? -mm : mm, n = nn < 0 ? -nn : nn;
  T lc = (T)ph[0][ll].x, ls = (T)ph[0][ll].y;
  T mc = (T)ph[1][m].x, ms = (T)ph[1][m].y;
  T nc = (T)ph[2][n].x, ns = (T)ph[2][n].y;
  T clm, slm;
  if (mm >= 0) {
    clm = lc * mc - ls * ms;
    slm = ls * mc + ms * lc;
  } else {
    clm = lc * mc + ls * ms;
    slm = ls * mc - ms * lc;
  }
  if (nn >= 0) {
    ckc = q * (clm * nc - slm * ns);
    cks = q * (slm * nc + clm * ns);
  } else {
    ckc = q * (clm * nc + slm * ns);
    cks = q * (slm * nc - clm * ns);
  }
}

// Trial energy terms of insertion / deletion / guest_energy (mc_moves.f:150-484)
// with ewald2 (ewald.f:663-733), srfrce (vdw_terms.f:382-466) and ewald1_guest
// (ewald.f:32-352).  Every thread of the workgroup takes part; the reduced
// sums land in sh.tot.
// Copies the read-only per-launch data into LDS: the erfc table always, the
// walker's framework and the k-vector list when the plan made room for them.
__device__ inline void stage_tables(const DevParams &P, const DevWalkers &W, int w, unsigned char *lds) {
  const int tid = threadIdx.x, nt = blockDim.x;
  double *erc = (double *)(lds + P.off_erc);
  for (int i = tid; i < P.mxegrd + 3; i += nt) erc[i] = P.erc[i];
  if (P.lds_fw) {
    const double *gx = P.fw_per_walker ? W.fwx + (size_t)w * P.nfw : P.fwx;
    const double *gy = P.fw_per_walker ? W.fwy + (size_t)w * P.nfw : P.fwy;
    const double *gz = P.fw_per_walker ? W.fwz + (size_t)w * P.nfw : P.fwz;
    double *fx = (double *)(lds + P.off_fw), *fy = fx + P.nfw, *fz = fy + P.nfw, *fq = fz + P.nfw;
    int *fi = (int *)(fq + P.nfw);
    for (int i = tid; i < P.nfw; i += nt) {
      fx[i] = gx[i];
      fy[i] = gy[i];
      fz[i] = gz[i];
      fq[i] = P.fwq[i];
      fi[i] = P.fwinfo[i];
    }
  }
  if (P.lds_k) {
    double *ak = (double *)(lds + P.off_k);
    int *kl = (int *)(ak + P.nk);
    for (int i = tid; i < P.nk; i += nt) {
      ak[i] = P.akk[i];
      kl[i] = P.klmn[i];
    }
  }
}

template <int NS, typename T>
__device__ void evaluate(const DevParams &P, const DevWalkers &W, int w, Smem<NS> &sh, unsigned char *lds) {
  using A = Acc<NS>;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int natms = sh.natms, mol = sh.mol, g = sh.g;
  // per-site 1-D phase tables, built by the reference recursion (ewald.f:108-147)
  for (int t = tid; t < 2 * KSITE * 3; t += nt) {
    int cfg = t / (KSITE * 3), rem = t % (KSITE * 3), s = rem / 3, ax = rem % 3;
    if (s >= natms || (cfg == 0 && !sh.has_new) || (cfg == 1 && !sh.has_old)) continue;
    const double *x = cfg == 0 ? sh.xn[s] : sh.xo[s];
    double fr = P.rcell[ax] * x[0] + P.rcell[3 + ax] * x[1] + P.rcell[6 + ax] * x[2];
    int kmax = ax == 0 ? P.kmax1 : (ax == 1 ? P.kmax2 : P.kmax3);
    double c1 = cos(2.0 * PI * fr), s1 = sin(2.0 * PI * fr);
    double2 *row = sh.ph[cfg][s][ax];
    row[0] = make_double2(1.0, 0.0);
    row[1] = make_double2(c1, s1);
    double c = c1, sn = s1;
    for (int l = 2; l <= kmax; ++l) {
      double cn = c * c1 - sn * s1;
      double snn = sn * c1 + c * s1;
      c = cn;
      sn = snn;
      row[l] = make_double2(c, sn);
    }
  }
  double acc[A::NACC];
#pragma unroll
  for (int i = 0; i < A::NACC; ++i) acc[i] = 0.0;
  int ovl_new = 0, ovl_old = 0;
  const double *erc = (const double *)(lds + P.off_erc);
  __syncthreads();
  // ---- pair terms: framework atoms (LDS when staged)
  auto fw_loop = [&](const double *fx, const double *fy, const double *fz, const double *fq, const int *fi) {
    for (int j = tid; j < P.nfw; j += nt) {
      double xj = fx[j], yj = fy[j], zj = fz[j], qj = fq[j];
      int inf = fi[j], tj = inf & 255, spj = (inf >> 8) & 255;
      for (int s = 0; s < natms; ++s) {
        if (sh.has_new)
          pair_site<NS, T>(P, erc, sh.xn[s], P.site_chgea[mol][s], P.site_t[mol][s], xj, yj, zj, qj, tj, spj,
                           acc, acc + NS, ovl_new);
        if (sh.has_old)
          pair_site<NS, T>(P, erc, sh.xo[s], P.site_chgea[mol][s], P.site_t[mol][s], xj, yj, zj, qj, tj, spj,
                           acc + 2 * NS, acc + 3 * NS, ovl_old);
      }
    }
  };
  if (P.lds_fw) {
    const double *fx = (const double *)(lds + P.off_fw), *fy = fx + P.nfw, *fz = fy + P.nfw, *fq = fz + P.nfw;
    fw_loop(fx, fy, fz, fq, (const int *)(fq + P.nfw));
  } else {
    const double *fx = P.fw_per_walker ? W.fwx + (size_t)w * P.nfw : P.fwx;
    const double *fy = P.fw_per_walker ? W.fwy + (size_t)w * P.nfw : P.fwy;
    const double *fz = P.fw_per_walker ? W.fwz + (size_t)w * P.nfw : P.fwz;
    fw_loop(fx, fy, fz, P.fwq, P.fwinfo);
  }
  // ---- pair terms: guest atoms (own molecule excluded)
  for (int gg = 1; gg <= P.ntpguest; ++gg) {
    int molg = P.locguest[gg], na = P.natoms[molg];
    int nact = sh.ws.nmols[molg] * na;
    size_t base = (size_t)w * P.gatoms_total + P.gbase[gg];
    for (int j = tid; j < nact; j += nt) {
      if (gg == g && j >= sh.excl_lo && j < sh.excl_hi) continue;
      int site = j % na;
      double xj = W.gx[base + j], yj = W.gy[base + j], zj = W.gz[base + j];
      double qj = P.site_q[molg][site];
      int tj = P.site_t[molg][site];
      for (int s = 0; s < natms; ++s) {
        if (sh.has_new)
          pair_site<NS, T>(P, erc, sh.xn[s], P.site_chgea[mol][s], P.site_t[mol][s], xj, yj, zj, qj, tj, molg,
                           acc, acc + NS, ovl_new);
        if (sh.has_old)
          pair_site<NS, T>(P, erc, sh.xo[s], P.site_chgea[mol][s], P.site_t[mol][s], xj, yj, zj, qj, tj, molg,
                           acc + 2 * NS, acc + 3 * NS, ovl_old);
      }
    }
  }
  // ---- reciprocal terms: one k-vector per thread
  const double2 *skw = W.sk + (size_t)w * (NS + 1) * P.nk;
  double2 *scr = W.sscr + (size_t)w * 2 * P.nk;
  auto recip_k = [&](int k, int lmn, double a, auto tag) {
    using U = decltype(tag);
    int ll = lmn & 255, mm = ((lmn >> 8) & 255) - 128, nn = ((lmn >> 16) & 255) - 128;
    double2 St = skw[(size_t)NS * P.nk + k];
    double2 Si[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) Si[i] = skw[(size_t)i * P.nk + k];
    double soc = 0.0, sos = 0.0;
    if (sh.has_old) {
      U c = (U)0, s_ = (U)0;
      for (int s = 0; s < natms; ++s) {
        U ckc, cks;
        site_phase<U>(sh.ph[1][s], ll, mm, nn, (U)P.site_q[mol][s], ckc, cks);
        c += ckc;
        s_ += cks;
      }
      soc = (double)c;
      sos = (double)s_;
      double mix = 2.0 * soc * St.x + 2.0 * sos * St.y;
      acc[4 * NS + 1] += a * (-mix + (soc * soc + sos * sos));
#pragma unroll
      for (int i = 0; i < NS; ++i) {
        double fkk = (i + 1 == mol) ? a : 2.0 * a;
        acc[5 * NS + 2 + i] += fkk * (soc * Si[i].x + sos * Si[i].y);
      }
      scr[P.nk + k] = make_double2(soc, sos);
    }
    if (sh.has_new) {
      U c = (U)0, s_ = (U)0;
      for (int s = 0; s < natms; ++s) {
        U ckc, cks;
        site_phase<U>(sh.ph[0][s], ll, mm, nn, (U)P.site_q[mol][s], ckc, cks);
        c += ckc;
        s_ += cks;
      }
      double snc = (double)c, sns = (double)s_;
      double Stx = St.x, Sty = St.y;
      if (sh.has_old) {
        Stx = Stx - soc;
        Sty = Sty - sos;
      }
      double mix = 2.0 * snc * Stx + 2.0 * sns * Sty;
      acc[4 * NS] += a * (mix + (snc * snc + sns * sns));
#pragma unroll
      for (int i = 0; i < NS; ++i) {
        double fkk = (i + 1 == mol) ? a : 2.0 * a;
        double sx = Si[i].x, sy = Si[i].y;
        if (sh.has_old && i + 1 == mol) {
          sx = sx - soc;
          sy = sy - sos;
        }
        acc[4 * NS + 2 + i] += fkk * (snc * sx + sns * sy);
      }
      scr[k] = make_double2(snc, sns);
    }
  };
  // the k list is sorted by a_k: the leading (large-weight) vectors always use
  // FP64 phase products, the tail uses T
  auto k_loops = [&](const int *kl, const double *ak) {
    for (int k = tid; k < P.nk64; k += nt) recip_k(k, kl[k], ak[k], 0.0);
    for (int k = P.nk64 + tid; k < P.nk; k += nt) recip_k(k, kl[k], ak[k], (T)0);
  };
  if (P.lds_k) {
    const double *ak = (const double *)(lds + P.off_k);
    k_loops((const int *)(ak + P.nk), ak);
  } else {
    k_loops(P.klmn, P.akk);
  }
  // ---- workgroup reduction: DPP within each wave, then across waves; only
  // the accumulators of the configurations present are reduced
  const int lane = tid & 63, wave = tid >> 6, nwaves = (nt + 63) >> 6;
#pragma unroll
  for (int i = 0; i < A::NACC; ++i) {
    bool is_new = (i < 2 * NS) || (i == 4 * NS) || (i >= 4 * NS + 2 && i < 5 * NS + 2);
    if (is_new ? sh.has_new : sh.has_old) {
      double v = wave_sum_dpp(acc[i]);
      if (lane == 63) sh.red[wave][i] = v;
    }
  }
  unsigned long long bn = __ballot(ovl_new != 0), bo = __ballot(ovl_old != 0);
  if (lane == 0) {
    sh.ovlw[wave][0] = bn != 0ULL;
    sh.ovlw[wave][1] = bo != 0ULL;
  }
  __syncthreads();
  if (tid < A::NACC) {
    double v = 0.0;
    for (int wv = 0; wv < nwaves; ++wv) v += sh.red[wv][tid];
    sh.tot[tid] = v;
  } else if (tid == A::NACC) {
    int a = 0, b = 0;
    for (int wv = 0; wv < nwaves; ++wv) {
      a |= sh.ovlw[wv][0];
      b |= sh.ovlw[wv][1];
    }
    sh.ovl_new = a;
    sh.ovl_old = b;
  }
  __syncthreads();
}

// ------------------------------------------------------------- decision --
// Combines the reduced sums into estep and delE exactly as insertion,
// deletion and displace_guest do, then applies energy_eval
// (utility_pack.f:3637-3693) with the mc_insert / mc_delete / mc_displace
// acceptance rules.  Thread 0.  If draw == 0 the Metropolis test is skipped
// (one-trial-at-a-time seam: the caller decides).
template <int NS>
__device__ void decide(const DevParams &P, Smem<NS> &sh, int draw) {
  using A = Acc<NS>;
  WState &ws = sh.ws;
  const int g = sh.g, mol = sh.mol, kind = sh.kind;
  const double *t = sh.tot;
  const double C = 2.0 * P.rvolm * R4PIE0 / P.epsq;
  const double alpha = P.alpha, volm = P.volm, engunit = P.engunit;
  const double chg = P.chgtmp[g];
  const double engsic0 = -R4PIE0 / P.epsq * alpha * P.selfq2[g] / SQRPI;
  const double qfix0 = -(0.5 * PI * R4PIE0 / P.epsq) * ((chg * ws.sumchg) / alpha + (chg / alpha) * (chg / alpha)) / volm;
  double e1new[KMAXS + 1], e1old[KMAXS + 1];
  for (int i = 1; i <= NS; ++i) {
    double fkk = (i == mol) ? 0.5 : 1.0;
    double sic = (i == mol) ? engsic0 : 0.0;
    double qfm = -(fkk * PI * R4PIE0 / P.epsq) * (ws.chgsum_mol[i] * chg) / (alpha * alpha * volm);
    e1new[i] = C * t[4 * NS + 2 + i - 1] + sic + qfm;
    e1old[i] = C * t[5 * NS + 2 + i - 1] + (-sic) + (-qfm);
  }
  double real_n = 0.0, vdw_n = 0.0, real_o = 0.0, vdw_o = 0.0;
  for (int i = 0; i < NS; ++i) {
    real_n += t[i];
    vdw_n += t[NS + i];
    real_o += t[2 * NS + i];
    vdw_o += t[3 * NS + i];
  }
  double ewld3 = P.ewald3en[mol];
  double estep = 0.0;
  int overlap = 0;
  for (int i = 0; i <= KMAXS; ++i) sh.delE[i] = 0.0;
  sh.delrc = 0.0;
  if (kind == FMC_MOVE_INSERT) {
    double delrc = gstlrcorrect<NS>(P, ws, g, 0, sh.delrc_mol);
    sh.delrc = delrc;
    overlap = sh.ovl_new;
    if (!overlap) {
      double ewld1eng = C * t[4 * NS] + engsic0 + qfix0;
      estep = (ewld1eng + real_n - ewld3 + vdw_n + delrc) / engunit;
      for (int i = 1; i <= NS; ++i) {
        int ik = loc2(mol, i);
        if (i != mol) sh.delE[i] += (e1new[i] + t[i - 1] + t[NS + i - 1] + sh.delrc_mol[ik]) / engunit;
      }
      sh.delE[mol] += estep;
      sh.comp[0] = (ewld1eng - ewld3) / engunit;
      sh.comp[1] = real_n / engunit;
      sh.comp[2] = vdw_n / engunit;
      sh.comp[3] = delrc / engunit;
    }
  } else if (kind == FMC_MOVE_DELETE) {
    double delrc = gstlrcorrect<NS>(P, ws, g, 1, sh.delrc_mol);
    sh.delrc = delrc;
    overlap = sh.ovl_old;
    double ewld1eng = C * t[4 * NS + 1] + (-engsic0) + (-qfix0);
    estep = (-ewld1eng + real_o + vdw_o - ewld3) / engunit;
    estep = estep - delrc / engunit;
    for (int i = 1; i <= NS; ++i) {
      int ik = loc2(mol, i);
      if (i != mol) sh.delE[i] -= (e1old[i] + t[2 * NS + i - 1] + t[3 * NS + i - 1] + sh.delrc_mol[ik]) / engunit;
    }
    sh.delE[mol] -= estep;
    sh.comp[0] = (-ewld1eng - ewld3) / engunit;
    sh.comp[1] = real_o / engunit;
    sh.comp[2] = vdw_o / engunit;
    sh.comp[3] = -delrc / engunit;
  } else {
    overlap = sh.ovl_new;
    double ewld1_old = C * t[4 * NS + 1] + (-engsic0) + (-qfix0);
    double est_old = sh.ovl_old ? 0.0 : (-ewld1_old + real_o + vdw_o - ewld3) / engunit;
    double ewld1_new = C * t[4 * NS] + engsic0 + qfix0;
    double est_new = overlap ? 0.0 : (ewld1_new + real_n + vdw_n - ewld3) / engunit;
    estep = est_new - est_old;
    for (int i = 1; i <= NS; ++i)
      if (i != mol)
        sh.delE[i] += (e1new[i] - e1old[i] + t[i - 1] - t[2 * NS + i - 1] + t[NS + i - 1] - t[3 * NS + i - 1]) / engunit;
    sh.delE[mol] += estep;
    sh.comp[0] = ((ewld1_new - ewld3) - (-ewld1_old - ewld3)) / engunit;
    sh.comp[1] = (real_n - real_o) / engunit;
    sh.comp[2] = (vdw_n - vdw_o) / engunit;
    sh.comp[3] = 0.0;
  }
  sh.estep = estep;
  sh.ovl_new = overlap;
  sh.rande = -1.0;
  sh.drew = 0;
  int accepted = 0;
  if (draw) {
    double beta = ws.beta;
    int ingsts = ws.nmols[mol];
    double ipress = ws.fug[g];
    if (kind == FMC_MOVE_INSERT) {
      if (!overlap) {
        sh.rande = duni(ws);
        sh.drew = 1;
        double e = estep, test;
        if (-1 * beta * e > 700.0) test = exp(700.0);
        else if (beta * e > 700.0) test = exp(-700.0);
        else test = P.statvolm * ipress / (((double)ingsts + 1.0) * BOLTZ * ws.temp) * exp(-1.0 * beta * e);
        accepted = sh.rande < test;
      }
    } else if (kind == FMC_MOVE_DELETE) {
      sh.rande = duni(ws);
      sh.drew = 1;
      double e = -estep, test;
      if (-1 * beta * e > 700.0) test = exp(700.0);
      else if (beta * e > 700.0) test = exp(-700.0);
      else if (fabs(ipress) < 1e-7) test = 1.0;
      else test = (double)ingsts * BOLTZ * ws.temp / (P.statvolm * ipress) * exp(-1.0 * beta * e);
      accepted = sh.rande < test;
    } else {
      if (!overlap) {
        if (estep < 0.0) {
          accepted = 1;
        } else {
          sh.rande = duni(ws);
          sh.drew = 1;
          double e = estep, test;
          if (-1 * beta * e > 700.0) test = exp(700.0);
          else if (beta * e > 700.0) test = exp(-700.0);
          else test = exp(-1.0 * beta * e);
          accepted = sh.rande < test;
        }
      }
    }
    if (accepted && kind == FMC_MOVE_INSERT && ws.nmols[mol] >= P.gcap[g]) {
      accepted = 0;
      ws.overflow = 1;
    }
  }
  sh.accepted = accepted;
}

// -------------------------------------------------------------- accept --
// accept_move, mc_moves.f:485-621 (plus the S(k) bookkeeping that
// mc_displace / displace_guest do around it, gcmc.f:1863-1930,2862-2871).
template <int NS>
__device__ void apply(const DevParams &P, const DevWalkers &W, int w, Smem<NS> &sh) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int g = sh.g, mol = sh.mol, natms = sh.natms, kind = sh.kind;
  double2 *skw = W.sk + (size_t)w * (NS + 1) * P.nk;
  const double2 *scr = W.sscr + (size_t)w * 2 * P.nk;
  for (int k = tid; k < P.nk; k += nt) {
    double2 sm = skw[(size_t)(mol - 1) * P.nk + k], st = skw[(size_t)NS * P.nk + k];
    if (kind == FMC_MOVE_INSERT) {
      double2 sn = scr[k];
      sm.x += sn.x; sm.y += sn.y;
      st.x += sn.x; st.y += sn.y;
    } else if (kind == FMC_MOVE_DELETE) {
      double2 so = scr[P.nk + k];
      sm.x -= so.x; sm.y -= so.y;
      st.x -= so.x; st.y -= so.y;
    } else {
      double2 sn = scr[k], so = scr[P.nk + k];
      sm.x = (sm.x - so.x) + sn.x; sm.y = (sm.y - so.y) + sn.y;
      st.x = (st.x - so.x) + sn.x; st.y = (st.y - so.y) + sn.y;
    }
    skw[(size_t)(mol - 1) * P.nk + k] = sm;
    skw[(size_t)NS * P.nk + k] = st;
  }
  size_t base = (size_t)w * P.gatoms_total + P.gbase[g];
  int nm = sh.ws.nmols[mol];
  if (kind == FMC_MOVE_INSERT) {
    if (tid < natms) {
      size_t a = base + (size_t)nm * natms + tid;
      W.gx[a] = sh.xn[tid][0];
      W.gy[a] = sh.xn[tid][1];
      W.gz[a] = sh.xn[tid][2];
    }
  } else if (kind == FMC_MOVE_DELETE) {
    // shift the molecules after `choice` down by one, keeping their order
    int lo = (sh.choice - 1) * natms, hi = (nm - 1) * natms;
    for (int b0 = lo; b0 < hi; b0 += nt) {
      int j = b0 + tid;
      double x = 0, y = 0, z = 0;
      if (j < hi) {
        x = W.gx[base + j + natms];
        y = W.gy[base + j + natms];
        z = W.gz[base + j + natms];
      }
      __syncthreads();
      if (j < hi) {
        W.gx[base + j] = x;
        W.gy[base + j] = y;
        W.gz[base + j] = z;
      }
      __syncthreads();
    }
  } else {
    if (tid < natms) {
      size_t a = base + (size_t)(sh.choice - 1) * natms + tid;
      W.gx[a] = sh.xn[tid][0];
      W.gy[a] = sh.xn[tid][1];
      W.gz[a] = sh.xn[tid][2];
    }
  }
  if (tid == 0) {
    WState &ws = sh.ws;
    int kk = loc2(mol, mol);
    for (int i = 1; i <= NS; ++i) ws.energy[i] += sh.delE[i];
    if (kind == FMC_MOVE_INSERT || kind == FMC_MOVE_DELETE) {
      double sg = kind == FMC_MOVE_INSERT ? 1.0 : -1.0;
      double engsic0 = -R4PIE0 / P.epsq * P.alpha * P.selfq2[g] / SQRPI;
      double engsictmp = kind == FMC_MOVE_INSERT ? engsic0 : -engsic0;  // ewald.f:324-326
      ws.engsic[kk] += sg * engsictmp;
      ws.chgsum_mol[mol] += sg * P.chgtmp[g];
      ws.sumchg += sg * P.chgtmp[g];
      int dn = kind == FMC_MOVE_INSERT ? 1 : -1;
      for (int i = 0; i < natms; ++i) {
        int ka = P.site_t[mol][i];
        ws.numtyp[ka] += dn;
        ws.numtyp_mol[mol][ka] += dn;
        if (P.site_f[mol][i] != 0) {
          ws.numfrz[ka] += dn;
          ws.numfrz_mol[mol][ka] += dn;
        }
      }
      ws.elrc += sg * sh.delrc;
      for (int t = 1; t <= KMAXC; ++t) ws.elrc_mol[t] += sg * sh.delrc_mol[t];
      ws.nmols[mol] += dn;
    }
  }
  __syncthreads();
}

template <int NS>
__device__ inline void stage_in(const DevWalkers &W, int w, Smem<NS> &sh) {
  const int *src = (const int *)(W.ws + w);
  int *dst = (int *)&sh.ws;
  for (int i = threadIdx.x; i < (int)(sizeof(WState) / 4); i += blockDim.x) dst[i] = src[i];
  __syncthreads();
}

template <int NS>
__device__ inline void stage_out(const DevWalkers &W, int w, Smem<NS> &sh) {
  __syncthreads();
  int *dst = (int *)(W.ws + w);
  const int *src = (const int *)&sh.ws;
  for (int i = threadIdx.x; i < (int)(sizeof(WState) / 4); i += blockDim.x) dst[i] = src[i];
}

template <int NS>
__device__ void record(const DevParams &P, const DevWalkers &W, Smem<NS> &sh, long long step) {
  fmc_trace_rec *r = W.trace + step;
  r->step = (int)sh.ws.steps;
  r->iguest = sh.g;
  r->move = sh.kind;
  r->choice = sh.choice;
  r->nmols = sh.nmols;
  r->overlap = sh.ovl_new;
  r->accepted = sh.accepted;
  r->drew_rande = sh.drew;
  r->estep = sh.estep;
  r->rande = sh.rande;
  for (int i = 0; i < 3 * FMC_MAX_SITES; ++i) r->xyz[i] = 0.0;
  for (int i = 0; i < sh.natms && i < FMC_MAX_SITES; ++i) {
    const double *x = sh.kind == FMC_MOVE_DELETE ? sh.xo[i] : sh.xn[i];
    r->xyz[3 * i] = x[0];
    r->xyz[3 * i + 1] = x[1];
    r->xyz[3 * i + 2] = x[2];
  }
  for (int i = 0; i < FMC_MAX_SPECIES; ++i) {
    r->energy[i] = (i < NS) ? sh.ws.energy[i + 1] : 0.0;
    r->delE[i] = (i < NS && sh.accepted) ? sh.delE[i + 1] : 0.0;
  }
  r->delrdisp = sh.ws.delrdisp[sh.g];
  r->rotangle = sh.ws.rotangle;
  r->rota_rotangle = sh.ws.rota_rotangle;
  r->pad = 0.0;
}

// ------------------------------------------------------------- kernels --
// Device-resident chains: one workgroup per walker, nsteps GCMC steps.
template <int NS, typename T>
__global__ void __launch_bounds__(512) k_run(const DevParams *__restrict__ Pp, DevWalkers W, int nsteps,
                                             int count_stats) {
  extern __shared__ __align__(16) unsigned char lds[];
  Smem<NS> &sh = *reinterpret_cast<Smem<NS> *>(lds);
  const DevParams &P = *Pp;
  const int w = blockIdx.x;
  stage_in<NS>(W, w, sh);
  stage_tables(P, W, w, lds);
  __syncthreads();
  for (int step = 0; step < nsteps; ++step) {
    if (threadIdx.x < 64) {
      if (sh.ws.mz_made - sh.ws.mz_used < 33) {
        mzt_refill(sh.ws, threadIdx.x);
        if (threadIdx.x == 0) sh.ws.mz_made += 33;
      }
      if (threadIdx.x == 0) propose<NS>(P, W, w, sh);
    }
    __syncthreads();
    evaluate<NS, T>(P, W, w, sh, lds);
    if (threadIdx.x == 0) {
      decide<NS>(P, sh, 1);
      WState &ws = sh.ws;
      int kind = sh.kind;
      ws.att[kind] += 1;
      if (sh.accepted) {
        ws.acc[kind] += 1;
        if (kind == FMC_MOVE_TRANSLATE) ws.accept_tran[sh.g] += 1;
        else if (kind == FMC_MOVE_ROTATE) ws.accept_rota[sh.g] += 1;
        else if (kind == FMC_MOVE_DISPLACE) ws.accept_disp[sh.g] += 1;
      }
      ws.steps += 1;
    }
    __syncthreads();
    if (sh.accepted) apply<NS>(P, W, w, sh);
    if (threadIdx.x == 0 && W.trace && w == 0) record<NS>(P, W, sh, step);
    if (threadIdx.x == 0 && count_stats) {
      WState &ws = sh.ws;
      double e = 0.0;
      for (int gg = 1; gg <= P.ntpguest; ++gg) {
        int m = P.locguest[gg];
        double n = ws.nmols[m];
        ws.sum_n[gg] += n;
        ws.sum_n2[gg] += n * n;
        e += ws.energy[m];
      }
      ws.sum_e += e;
      ws.sum_e2 += e * e;
      ws.counted += 1;
    }
  }
  stage_out<NS>(W, w, sh);
}

// One trial of the mc_moves.f seam: the proposal comes from the pending
// fields of walker w (written by the host); results go to W.trial_out.
template <int NS, typename T>
__global__ void __launch_bounds__(512) k_trial(const DevParams *__restrict__ Pp, DevWalkers W, int w) {
  extern __shared__ __align__(16) unsigned char lds[];
  Smem<NS> &sh = *reinterpret_cast<Smem<NS> *>(lds);
  const DevParams &P = *Pp;
  stage_in<NS>(W, w, sh);
  stage_tables(P, W, w, lds);
  if (threadIdx.x == 0) {
    WState &ws = sh.ws;
    int g = ws.pend_iguest, mol = P.locguest[g], natms = P.natoms[mol];
    sh.kind = ws.pend_kind;
    sh.g = g;
    sh.mol = mol;
    sh.natms = natms;
    sh.nmols = ws.nmols[mol];
    sh.choice = ws.pend_choice;
    sh.excl_lo = sh.excl_hi = 0;
    sh.has_new = sh.kind != FMC_MOVE_DELETE;
    sh.has_old = sh.kind != FMC_MOVE_INSERT;
    for (int i = 0; i < natms; ++i)
      for (int d = 0; d < 3; ++d) sh.xn[i][d] = ws.pend_xyz[3 * i + d];
    if (sh.has_old) {
      load_molecule<NS>(P, W, w, g, sh.choice, sh.xo, natms);
      sh.excl_lo = (sh.choice - 1) * natms;
      sh.excl_hi = sh.choice * natms;
    }
  }
  __syncthreads();
  evaluate<NS, T>(P, W, w, sh, lds);
  if (threadIdx.x == 0) {
    decide<NS>(P, sh, 0);
    fmc_trial *o = W.trial_out;
    o->estep = sh.estep;
    for (int i = 0; i < FMC_MAX_SPECIES; ++i) o->delE[i] = (i < NS) ? sh.delE[i + 1] : 0.0;
    o->recip = sh.comp[0];
    o->real = sh.comp[1];
    o->vdw = sh.comp[2];
    o->tail = sh.comp[3];
    o->overlap = sh.ovl_new;
    o->pad = 0;
    // keep what accept needs
    WState &ws = sh.ws;
    ws.pend_valid = 1;
    double *keep = (double *)(W.trial_out + 1);
    keep[0] = sh.delrc;
    for (int i = 0; i <= KMAXS; ++i) keep[1 + i] = sh.delE[i];
    for (int i = 0; i <= KMAXC; ++i) keep[2 + KMAXS + i] = sh.delrc_mol[i];
  }
  stage_out<NS>(W, w, sh);
}

template <int NS>
__global__ void __launch_bounds__(512) k_accept(const DevParams *__restrict__ Pp, DevWalkers W, int w) {
  extern __shared__ __align__(16) unsigned char lds[];
  Smem<NS> &sh = *reinterpret_cast<Smem<NS> *>(lds);
  const DevParams &P = *Pp;
  stage_in<NS>(W, w, sh);
  if (threadIdx.x == 0) {
    WState &ws = sh.ws;
    int g = ws.pend_iguest, mol = P.locguest[g];
    sh.kind = ws.pend_kind;
    sh.g = g;
    sh.mol = mol;
    sh.natms = P.natoms[mol];
    sh.choice = ws.pend_choice;
    for (int i = 0; i < sh.natms; ++i)
      for (int d = 0; d < 3; ++d) sh.xn[i][d] = ws.pend_xyz[3 * i + d];
    const double *keep = (const double *)(W.trial_out + 1);
    sh.delrc = keep[0];
    for (int i = 0; i <= KMAXS; ++i) sh.delE[i] = keep[1 + i];
    for (int i = 0; i <= KMAXC; ++i) sh.delrc_mol[i] = keep[2 + KMAXS + i];
    if (sh.kind == FMC_MOVE_TRANSLATE || sh.kind == FMC_MOVE_ROTATE) sh.kind = FMC_MOVE_DISPLACE;
    ws.pend_valid = 0;
    ws.pend_kind = 0;
  }
  __syncthreads();
  apply<NS>(P, W, w, sh);
  stage_out<NS>(W, w, sh);
}

// Full FP64 structure factor of every walker (ewald1, ewald.f:354-661): one
// thread per (walker, k); species rows plus the total row.
template <int NS>
__global__ void __launch_bounds__(256) k_resum(const DevParams *__restrict__ Pp, DevWalkers W, double *drift) {
  const DevParams &P = *Pp;
  const int w = blockIdx.y, k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P.nk) return;
  int lmn = P.klmn[k];
  double l = (double)(lmn & 255), m = (double)(((lmn >> 8) & 255) - 128), n = (double)(((lmn >> 16) & 255) - 128);
  double sr[NS + 1], si[NS + 1];
  for (int i = 0; i <= NS; ++i) sr[i] = si[i] = 0.0;
  const double *fx = P.fw_per_walker ? W.fwx + (size_t)w * P.nfw : P.fwx;
  const double *fy = P.fw_per_walker ? W.fwy + (size_t)w * P.nfw : P.fwy;
  const double *fz = P.fw_per_walker ? W.fwz + (size_t)w * P.nfw : P.fwz;
  const double *r = P.rcell;
  for (int j = 0; j < P.nfw; ++j) {
    double x = fx[j], y = fy[j], z = fz[j];
    double sx = r[0] * x + r[3] * y + r[6] * z, sy = r[1] * x + r[4] * y + r[7] * z, sz = r[2] * x + r[5] * y + r[8] * z;
    double sn, cs;
    sincospi(2.0 * (l * sx + m * sy + n * sz), &sn, &cs);
    double q = P.fwq[j];
    int sp = (P.fwinfo[j] >> 8) & 255;
    for (int i = 0; i < NS; ++i)
      if (sp == i + 1) {
        sr[i] += q * cs;
        si[i] += q * sn;
      }
    sr[NS] += q * cs;
    si[NS] += q * sn;
  }
  const WState &ws = W.ws[w];
  for (int gg = 1; gg <= P.ntpguest; ++gg) {
    int molg = P.locguest[gg], na = P.natoms[molg], nact = ws.nmols[molg] * na;
    size_t base = (size_t)w * P.gatoms_total + P.gbase[gg];
    for (int j = 0; j < nact; ++j) {
      double x = W.gx[base + j], y = W.gy[base + j], z = W.gz[base + j];
      double sx = r[0] * x + r[3] * y + r[6] * z, sy = r[1] * x + r[4] * y + r[7] * z, sz = r[2] * x + r[5] * y + r[8] * z;
      double sn, cs;
      sincospi(2.0 * (l * sx + m * sy + n * sz), &sn, &cs);
      double q = P.site_q[molg][j % na];
      sr[molg - 1] += q * cs;
      si[molg - 1] += q * sn;
      sr[NS] += q * cs;
      si[NS] += q * sn;
    }
  }
  double2 *skw = W.sk + (size_t)w * (NS + 1) * P.nk;
  if (drift) {
    double2 o = skw[(size_t)NS * P.nk + k];
    double d = fmax(fabs(o.x - sr[NS]), fabs(o.y - si[NS]));
    // per-walker maximum through an integer atomic on the non-negative bit pattern
    atomicMax((unsigned long long *)(drift + w), (unsigned long long)__double_as_longlong(d));
  }
  for (int i = 0; i <= NS; ++i) skw[(size_t)i * P.nk + k] = make_double2(sr[i], si[i]);
}

// Species-pair reciprocal sums of ewald1 (ewald.f:597-611) for walker w:
// out[w][0] = sum akk |S_tot|^2, out[w][kmol] = sum fkk akk Re(S_i S_j*).
template <int NS>
__global__ void __launch_bounds__(256) k_recip_sums(const DevParams *__restrict__ Pp, DevWalkers W, double *out) {
  const DevParams &P = *Pp;
  const int w = blockIdx.x;
  const double2 *skw = W.sk + (size_t)w * (NS + 1) * P.nk;
  constexpr int NC = NS * (NS + 1) / 2;
  double acc[NC + 1];
  for (int i = 0; i <= NC; ++i) acc[i] = 0.0;
  for (int k = threadIdx.x; k < P.nk; k += blockDim.x) {
    double a = P.akk[k];
    double2 st = skw[(size_t)NS * P.nk + k];
    acc[0] += a * (st.x * st.x + st.y * st.y);
    int kmol = 0;
    for (int i = 1; i <= NS; ++i) {
      double fkk = 2.0 * a;
      double2 si = skw[(size_t)(i - 1) * P.nk + k];
      for (int j = 1; j <= i; ++j) {
        ++kmol;
        if (i == j) fkk = a;
        double2 sj = skw[(size_t)(j - 1) * P.nk + k];
        acc[kmol] += fkk * (si.x * sj.x + si.y * sj.y);
      }
    }
  }
  __shared__ double red[KWAVES][KMAXC + 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  for (int i = 0; i <= NC; ++i) {
    double v = wave_sum(acc[i]);
    if (lane == 0) red[wave][i] = v;
  }
  __syncthreads();
  if (threadIdx.x <= NC) {
    double v = 0.0;
    for (int q = 0; q < nw; ++q) v += red[q][threadIdx.x];
    out[(size_t)w * (KMAXC + 1) + threadIdx.x] = v;
  }
}

// Real-space and LJ sums of single_point (gcmc.f:2604-2652) over every pair
// once (the parlst half list, nlist_builders.f:1-185): frozen-frozen pairs and
// pairs inside one guest molecule are skipped.  Accumulated per species pair.
template <int NS>
__global__ void __launch_bounds__(256) k_pair_sums(const DevParams *__restrict__ Pp, DevWalkers W, int ntot,
                                                   double *out) {
  const DevParams &P = *Pp;
  const int w = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double e2[KMAXC + 1], ev[KMAXC + 1];
  for (int t = 0; t <= KMAXC; ++t) e2[t] = ev[t] = 0.0;
  const double *fx = P.fw_per_walker ? W.fwx + (size_t)w * P.nfw : P.fwx;
  const double *fy = P.fw_per_walker ? W.fwy + (size_t)w * P.nfw : P.fwy;
  const double *fz = P.fw_per_walker ? W.fwz + (size_t)w * P.nfw : P.fwz;
  const WState &ws = W.ws[w];
  auto atom = [&](int a, double &x, double &y, double &z, double &q, int &t, int &sp, int &frz, int &molid) {
    if (a < P.nfw) {
      x = fx[a]; y = fy[a]; z = fz[a]; q = P.fwq[a];
      int inf = P.fwinfo[a];
      t = inf & 255; sp = (inf >> 8) & 255; frz = (inf >> 16) & 1; molid = -1;
      return;
    }
    a -= P.nfw;
    for (int gg = 1; gg <= P.ntpguest; ++gg) {
      int molg = P.locguest[gg], na = P.natoms[molg], nact = ws.nmols[molg] * na;
      if (a < nact) {
        size_t b = (size_t)w * P.gatoms_total + P.gbase[gg] + a;
        x = W.gx[b]; y = W.gy[b]; z = W.gz[b];
        q = P.site_q[molg][a % na]; t = P.site_t[molg][a % na]; sp = molg;
        frz = P.site_f[molg][a % na] != 0; molid = gg * 1000000 + a / na;
        return;
      }
      a -= nact;
    }
  };
  if (i < ntot) {
    double xi, yi, zi, qi; int ti, spi, fzi, mi;
    atom(i, xi, yi, zi, qi, ti, spi, fzi, mi);
    double chgea = qi / P.epsq * R4PIE0;
    for (int j = i + 1; j < ntot; ++j) {
      double xj, yj, zj, qj; int tj, spj, fzj, mj;
      atom(j, xj, yj, zj, qj, tj, spj, fzj, mj);
      if (fzi && fzj) continue;
      if (mi >= 0 && mi == mj) continue;
      double dx = xi - xj, dy = yi - yj, dz = zi - zj;
      min_image(P, dx, dy, dz);
      double rsq = dx * dx + dy * dy + dz * dz;
      if (rsq < P.rcut2) {
        double r = sqrt(rsq);
        int kk = loc2(spi, spj);
        double chgprd = chgea * qj;
        if (fabs(chgprd) > 1e-10) e2[kk] += interp3<double>(P.erc, r, P.rdrewd) * chgprd;
        int k = P.pairidx[ti][tj];
        if (k >= 0) ev[kk] += interp3<double>(P.vvv + (size_t)k * P.vstride, r, P.rdlrpot);
      }
    }
  }
  for (int t = 1; t <= KMAXC; ++t) {
    double a = wave_sum(e2[t]), b = wave_sum(ev[t]);
    if ((threadIdx.x & 63) == 0) {
      if (a != 0.0) atomicAdd(out + (size_t)w * 2 * (KMAXC + 1) + t, a);
      if (b != 0.0) atomicAdd(out + (size_t)w * 2 * (KMAXC + 1) + KMAXC + 1 + t, b);
    }
  }
}

}  // namespace fmc
