/*
 * TEST INFRASTRUCTURE ONLY -- the oracle.  A plain-C FP64 restatement of the
 * reference FastMC GCMC trial-move energy path (FastMC v1.4.7, Fortran).  It is
 * what the HIP library's results are compared with: tests/, smoke() and
 * bench.py's cpu_baseline leg may call it; the product never does.
 *
 * Every routine names the reference routine it restates (file:line).  Arrays
 * that the reference indexes from 1 are kept 1-based here (index 0 unused) so
 * each loop can be read side by side with the Fortran.  Parity of this file is
 * pinned by tests/test_oracle.py against the reference's own known answer
 * (run_test.py:40), against traces recorded from the compiled reference
 * (JSON fixtures under tests/golden, made by tests/golden/make_golden.py through
 * oracle/ref_driver.f90) and against closed-form cases.
 */
#include "fastmc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define R4PIE0 138935.4835          /* utility_pack.f:113 */
#define KBOLTZ 1.9872041e-3         /* utility_pack.f:121 */
#define BOLTZ 1.3806503e-23         /* utility_pack.f:124 */
#define PI 3.141592653589793        /* utility_pack.f:95 */
#define SQRPI 1.7724538509055159    /* utility_pack.f:96 */
#define DELRMIN 0.1                 /* utility_pack.f:107 */
#define MINANGLE (PI / 18.0)        /* utility_pack.f:104 */
#define MAXGUEST 20000              /* utility_pack.f:89 */
#define HPLANCK 6.62607004e-34      /* utility_pack.f:144 */
#define AVO 6.022140857e23          /* utility_pack.f:127 */

/* ------------------------------------------------------------------ RNG -- */
/* duni, utility_pack.f:2768-2857: Marsaglia-Zaman-Tsang universal generator in
 * REAL(4).  Every operation is exact in binary32 (values are multiples of
 * 2^-24 in [0,1)). */
typedef struct {
  float u[98];
  float c, cd, cm;
  int ir, jr;
} mzt_t;

static void mzt_seed(mzt_t *g, int i, int j, int k, int l) {
  for (int ii = 1; ii <= 97; ++ii) {
    float s = 0.0f, t = 0.5f;
    for (int jj = 1; jj <= 24; ++jj) {
      int m = ((i * j) % 179 * k) % 179;
      i = j;
      j = k;
      k = m;
      l = (53 * l + 1) % 169;
      if ((l * m) % 64 >= 32) s = s + t;
      t = 0.5f * t;
    }
    g->u[ii] = s;
  }
  g->c = 362436.0f / 16777216.0f;
  g->cd = 7654321.0f / 16777216.0f;
  g->cm = 16777213.0f / 16777216.0f;
  g->ir = 97;
  g->jr = 33;
}

static double mzt_next(mzt_t *g) {
  float uni = g->u[g->ir] - g->u[g->jr];
  if (uni < 0.0f) uni = uni + 1.0f;
  g->u[g->ir] = uni;
  g->ir -= 1;
  if (g->ir == 0) g->ir = 97;
  g->jr -= 1;
  if (g->jr == 0) g->jr = 97;
  g->c = g->c - g->cd;
  if (g->c < 0.0f) g->c = g->c + g->cm;
  uni = uni - g->c;
  if (uni < 0.0f) uni = uni + 1.0f;
  return (double)uni;
}

void orc_duni_seq(const int seed[4], int n, double *out) {
  mzt_t g;
  mzt_seed(&g, seed[0], seed[1], seed[2], seed[3]);
  for (int i = 0; i < n; ++i) out[i] = mzt_next(&g);
}

/* ---------------------------------------------------------------- state -- */
typedef struct {
  int ll, mm, nn;
  double rksq, akk;
} kvec_t;

struct orc_state {
  /* system */
  int imcon, maxmls, ntpguest, ntpfram, ntpatm, ntpvdw, maxvdw, mxcmls;
  double cell[10], rcell[10], volm, statvolm;
  double rcut, rvdw, ddelr, alpha, epsq, engunit, overlap, temp, beta, sumchg, surftol;
  int kmax1, kmax2, kmax3, newld, mxegrd;
  double drewd, dlrpot;
  double *erc;                 /* [mxegrd+1] */
  double *vvv;                 /* [(maxvdw+2)*(mxegrd+1)] column k at k*(mxegrd+1) */
  int *lstvdw, *ltpvdw;        /* [maxvdw+2] */
  double *prm_eps, *prm_sig, *steadd;
  kvec_t *kv;                  /* [newld+1] */
  /* species (1-based, FIELD order) */
  int locguest[FMC_MAX_SPECIES + 1], locfram[FMC_MAX_SPECIES + 1];
  int isguest[FMC_MAX_SPECIES + 1];
  int numatoms[FMC_MAX_SPECIES + 1], nummols[FMC_MAX_SPECIES + 1];
  double atmchg[FMC_MAX_SPECIES + 1][FMC_MAX_SITES + 1];
  double atmwght[FMC_MAX_SPECIES + 1][FMC_MAX_SITES + 1];
  int ltpsit[FMC_MAX_SPECIES + 1][FMC_MAX_SITES + 1];
  int lfzsite[FMC_MAX_SPECIES + 1][FMC_MAX_SITES + 1];
  double guestx[FMC_MAX_SPECIES + 1][FMC_MAX_SITES + 1];
  double guesty[FMC_MAX_SPECIES + 1][FMC_MAX_SITES + 1];
  double guestz[FMC_MAX_SPECIES + 1][FMC_MAX_SITES + 1];
  double *fwchg[FMC_MAX_SPECIES + 1];  /* framework species: per-site arrays */
  int *fwtyp[FMC_MAX_SPECIES + 1], *fwfrz[FMC_MAX_SPECIES + 1];
  double *molx[FMC_MAX_SPECIES + 1], *moly[FMC_MAX_SPECIES + 1], *molz[FMC_MAX_SPECIES + 1];
  int molcap[FMC_MAX_SPECIES + 1];
  /* condensed atoms (condense, utility_pack.f:3951-4050) */
  int totatm, maxalloc;
  double *xxx, *yyy, *zzz, *atmcharge;
  int *ltype, *moltype, *lfreezesite;
  /* S(k) per species + total: [(maxmls+2)*(newld+1)] */
  double *ckcsum, *ckssum, *ckcsnew, *ckssnew, *ckcsorig, *ckssorig;
  double ewald1en[64], ewald2en[64], vdwen[64], ewald3en[FMC_MAX_SPECIES + 1];
  double ewald1entmp[64], ewald2entmp[64], vdwentmp[64];
  double engsic[64], engsicorig[64], chgsum_mol[FMC_MAX_SPECIES + 1];
  double delE[64], energy[FMC_MAX_SPECIES + 2];
  double elrc, elrc_mol[64], delrc_mol[64];
  int numtyp[FMC_MAX_TYPES + 1], numfrz[FMC_MAX_TYPES + 1];
  int numtyp_mol[FMC_MAX_SPECIES + 2][FMC_MAX_TYPES + 1];
  int numfrz_mol[FMC_MAX_SPECIES + 2][FMC_MAX_TYPES + 1];
  int numtyp_gstmol[FMC_MAX_SPECIES + 1][FMC_MAX_TYPES + 1];
  int numfrz_gstmol[FMC_MAX_SPECIES + 1][FMC_MAX_TYPES + 1];
  /* trial scratch */
  double newx[FMC_MAX_SITES + 1], newy[FMC_MAX_SITES + 1], newz[FMC_MAX_SITES + 1];
  int ind[FMC_MAX_SITES + 1];
  int *gstlentry_dummy;
  int *gstlist[FMC_MAX_SITES + 1];
  int gstlentry[FMC_MAX_SITES + 1];
  double *xdf, *ydf, *zdf, *rsqdf;
  int *ilist, *moldf;
  double *elc0, *els0, *emc, *ems, *enc, *ens, *elc1, *els1, *clm, *slm;
  /* surface flags of the last trial (lnewsurf / lsurf, linitsurf) and
   * surfacemols by species */
  int lnewsurf, linitsurf;
  int surfacemols[FMC_MAX_SPECIES + 1];
  /* production statistics (gcmc.f:1050-1130): windows of nwindsteps counted
   * steps; win holds closed windows as fmc_window records */
  int st_nwind, st_accsample, nwin, capwin;
  long long prodcount, rollcount;
  double chain[FMC_MAX_SPECIES + 1][FMC_WIN_STATS];
  fmc_window *win;
  /* probability plots (storeprob): grid, unit-cell reciprocal, plots, grids
   * [nplots][gridsize + 1] with the reference's 1-based bins */
  int nplots, ngrida, ngridb, ngridc, gridsize;
  double rucell[10];
  fmc_prob_plot plots[FMC_MAX_PLOTS];
  double *grid;
  /* numguests.out records (gcmc.f:1138-1163) every nnumg steps */
  long long gcmccount;
  int nnumg, numg_n, numg_cap, swap_jguest;
  fmc_numguests_rec *numg;
  /* pending trial (for accept/reject) */
  int pend_kind, pend_iguest, pend_choice, pend_valid;
  double pend_delrc, pend_engsictmp, pend_chgtmp;
  /* MC driver state (gcmc.f main loop) */
  mzt_t rng;
  float mcf[FMC_MAX_SPECIES + 1][FMC_NMOVES];
  double gstfuga[FMC_MAX_SPECIES + 1], delr[FMC_MAX_SPECIES + 1];
  double delrdisp[FMC_MAX_SPECIES + 1], tran_delr[FMC_MAX_SPECIES + 1];
  long long mv_att[FMC_NMOVES + 1], mv_acc[FMC_NMOVES + 1];  /* by move kind, summed over guests */
  double rotangle, rota_rotangle;
  int disp_count[FMC_MAX_SPECIES + 1], accept_disp[FMC_MAX_SPECIES + 1];
  int tran_count[FMC_MAX_SPECIES + 1], accept_tran[FMC_MAX_SPECIES + 1];
  int rota_count[FMC_MAX_SPECIES + 1], accept_rota[FMC_MAX_SPECIES + 1];
  double disp_ratio, tran_ratio, rota_ratio;
  /* Wang-Landau walker (wang_landau.f): window, preload target, ln g and visits by N */
  int wl_on, wl_done, wl_min, wl_max, wl_preload, wl_error, wl_visittol, wl_nbins;
  long long wl_iter, wl_sweeps;
  double wl_f, wl_logf, wl_flat, wl_prectol, wl_lambda3;
  double *wl_lng;
  int *wl_hist;
};

static int loc2(int i, int j) {  /* utility_pack.f:2859-2872 */
  int a = i > j ? i : j, b = i > j ? j : i;
  return (a * (a - 1)) / 2 + b;
}

#define SK(arr, m, k) (s->arr[(size_t)(k) * (size_t)(s->maxmls + 2) + (size_t)(m)])
#define VVV(i, k) (s->vvv[(size_t)(k) * (size_t)(s->mxegrd + 1) + (size_t)(i)])

static double fnint(double x) { return round(x); } /* Fortran nint: half away from zero */

/* dcell, utility_pack.f:1904-1960 (bbb 1-based) */
static void dcell(const double *aaa, double *bbb) {
  bbb[1] = sqrt(aaa[1] * aaa[1] + aaa[2] * aaa[2] + aaa[3] * aaa[3]);
  bbb[2] = sqrt(aaa[4] * aaa[4] + aaa[5] * aaa[5] + aaa[6] * aaa[6]);
  bbb[3] = sqrt(aaa[7] * aaa[7] + aaa[8] * aaa[8] + aaa[9] * aaa[9]);
  bbb[4] = (aaa[1] * aaa[4] + aaa[2] * aaa[5] + aaa[3] * aaa[6]) / (bbb[1] * bbb[2]);
  bbb[5] = (aaa[1] * aaa[7] + aaa[2] * aaa[8] + aaa[3] * aaa[9]) / (bbb[1] * bbb[3]);
  bbb[6] = (aaa[4] * aaa[7] + aaa[5] * aaa[8] + aaa[6] * aaa[9]) / (bbb[2] * bbb[3]);
  double axb1 = aaa[2] * aaa[6] - aaa[3] * aaa[5];
  double axb2 = aaa[3] * aaa[4] - aaa[1] * aaa[6];
  double axb3 = aaa[1] * aaa[5] - aaa[2] * aaa[4];
  double bxc1 = aaa[5] * aaa[9] - aaa[6] * aaa[8];
  double bxc2 = aaa[6] * aaa[7] - aaa[4] * aaa[9];
  double bxc3 = aaa[4] * aaa[8] - aaa[5] * aaa[7];
  double cxa1 = aaa[8] * aaa[3] - aaa[2] * aaa[9];
  double cxa2 = aaa[1] * aaa[9] - aaa[3] * aaa[7];
  double cxa3 = aaa[2] * aaa[7] - aaa[1] * aaa[8];
  bbb[10] = fabs(aaa[1] * bxc1 + aaa[2] * bxc2 + aaa[3] * bxc3);
  bbb[7] = bbb[10] / sqrt(bxc1 * bxc1 + bxc2 * bxc2 + bxc3 * bxc3);
  bbb[8] = bbb[10] / sqrt(cxa1 * cxa1 + cxa2 * cxa2 + cxa3 * cxa3);
  bbb[9] = bbb[10] / sqrt(axb1 * axb1 + axb2 * axb2 + axb3 * axb3);
}

/* invert, utility_pack.f:1962-2004 */
static void invert(const double *a, double *b) {
  b[1] = a[5] * a[9] - a[6] * a[8];
  b[2] = a[3] * a[8] - a[2] * a[9];
  b[3] = a[2] * a[6] - a[3] * a[5];
  b[4] = a[6] * a[7] - a[4] * a[9];
  b[5] = a[1] * a[9] - a[3] * a[7];
  b[6] = a[3] * a[4] - a[1] * a[6];
  b[7] = a[4] * a[8] - a[5] * a[7];
  b[8] = a[2] * a[7] - a[1] * a[8];
  b[9] = a[1] * a[5] - a[2] * a[4];
  double d = a[1] * b[1] + a[4] * b[2] + a[7] * b[3];
  double r = 0.0;
  if (fabs(d) > 0.0) r = 1.0 / d;
  for (int i = 1; i <= 9; ++i) b[i] = r * b[i];
}

/* images, utility_pack.f:1697-1902 (imcon 1, 2, 3) */
static void images(const orc_state *s, int n, double *x, double *y, double *z) {
  const double *cell = s->cell;
  if (s->imcon == 1) {
    double aaa = 1.0 / cell[1];
    for (int i = 1; i <= n; ++i) {
      x[i] = x[i] - cell[1] * fnint(aaa * x[i]);
      y[i] = y[i] - cell[1] * fnint(aaa * y[i]);
      z[i] = z[i] - cell[1] * fnint(aaa * z[i]);
    }
  } else if (s->imcon == 2) {
    double aaa = 1.0 / cell[1], bbb = 1.0 / cell[5], ccc = 1.0 / cell[9];
    for (int i = 1; i <= n; ++i) {
      x[i] = x[i] - cell[1] * fnint(aaa * x[i]);
      y[i] = y[i] - cell[5] * fnint(bbb * y[i]);
      z[i] = z[i] - cell[9] * fnint(ccc * z[i]);
    }
  } else {
    double rc[10];
    invert(cell, rc);
    for (int i = 1; i <= n; ++i) {
      double ssx = (rc[1] * x[i] + rc[4] * y[i] + rc[7] * z[i]);
      double ssy = (rc[2] * x[i] + rc[5] * y[i] + rc[8] * z[i]);
      double ssz = (rc[3] * x[i] + rc[6] * y[i] + rc[9] * z[i]);
      double xss = ssx - fnint(ssx), yss = ssy - fnint(ssy), zss = ssz - fnint(ssz);
      x[i] = (cell[1] * xss + cell[4] * yss + cell[7] * zss);
      y[i] = (cell[2] * xss + cell[5] * yss + cell[8] * zss);
      z[i] = (cell[3] * xss + cell[6] * yss + cell[9] * zss);
    }
  }
}

/* ------------------------------------------------------------- tables -- */
/* erfcgen, ewald.f:1-30 */
static void erfcgen(orc_state *s) {
  const double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741;
  const double a4 = -1.453152027, a5 = 1.061405429, pp = 0.3275911;
  s->drewd = s->rcut / (double)(s->mxegrd - 4);
  for (int i = 1; i <= s->mxegrd; ++i) {
    double rrr = (double)i * s->drewd;
    double tt = 1.0 / (1.0 + pp * s->alpha * rrr);
    double ar = s->alpha * rrr;
    double exp1 = exp(-(ar * ar));
    s->erc[i] = tt * (a1 + tt * (a2 + tt * (a3 + tt * (a4 + tt * a5)))) * exp1 / rrr;
  }
}

static double ipow6(double x) {
  double x2 = x * x;
  return x2 * x2 * x2;
}

/* forgen LJ branch, vdw_terms.f:181-234, plus the zero column of
 * define_van_der_waals, vdw_terms.f:150-174 */
static void forgen(orc_state *s) {
  s->dlrpot = s->rvdw / (double)(s->mxegrd - 4);
  for (int k = 1; k <= s->ntpvdw; ++k) {
    double eps = s->prm_eps[k], sig = s->prm_sig[k];
    for (int i = 1; i <= s->mxegrd; ++i) {
      double rrr = (double)i * s->dlrpot;
      double sr6 = ipow6(sig / rrr);
      VVV(i, k) = 4.0 * eps * sr6 * (sr6 - 1.0);
    }
  }
  for (int i = 1; i <= s->mxegrd; ++i) VVV(i, s->ntpvdw + 1) = 0.0;
}

/* k-vector enumeration of ewald1 / ewald1_guest, ewald.f:87-94,151-222 */
static void build_kvecs(orc_state *s) {
  double buffer[11];
  double twopi = 2.0 * PI;
  dcell(s->rcell, buffer);
  double rcpcut = fmin(fmin((double)s->kmax1 * buffer[7], (double)s->kmax2 * buffer[8]),
                       (double)s->kmax3 * buffer[9]);
  rcpcut = rcpcut * 1.05 * twopi;
  double rcpct2 = rcpcut * rcpcut;
  double ralph = -0.25 / (s->alpha * s->alpha);
  int cap = (2 * s->kmax1 + 1) * (2 * s->kmax2 + 1) * (2 * s->kmax3 + 1);
  s->kv = (kvec_t *)calloc((size_t)cap + 1, sizeof(kvec_t));
  int kkk = 0, mmin = 0, nmin = 1;
  const double *rc = s->rcell;
  for (int ll = 0; ll <= s->kmax1; ++ll) {
    double tmp = twopi * (double)ll;
    double rkx1 = tmp * rc[1], rky1 = tmp * rc[4], rkz1 = tmp * rc[7];
    for (int mm = mmin; mm <= s->kmax2; ++mm) {
      tmp = twopi * (double)mm;
      double rkx2 = rkx1 + tmp * rc[2], rky2 = rky1 + tmp * rc[5], rkz2 = rkz1 + tmp * rc[8];
      for (int nn = nmin; nn <= s->kmax3; ++nn) {
        tmp = twopi * (double)nn;
        double rkx3 = rkx2 + tmp * rc[3], rky3 = rky2 + tmp * rc[6], rkz3 = rkz2 + tmp * rc[9];
        double rksq = rkx3 * rkx3 + rky3 * rky3 + rkz3 * rkz3;
        if (rksq <= rcpct2) {
          ++kkk;
          s->kv[kkk].ll = ll;
          s->kv[kkk].mm = mm;
          s->kv[kkk].nn = nn;
          s->kv[kkk].rksq = rksq;
          double rrksq = 1.0 / rksq;
          s->kv[kkk].akk = exp(ralph * rksq) * rrksq;
        }
      }
      nmin = -s->kmax3;
    }
    mmin = -s->kmax2;
  }
  s->newld = kkk;
}

/* ---------------------------------------------------------- condense -- */
static void condense(orc_state *s) {  /* utility_pack.f:3951-4050 (coordinates part) */
  int jj = 0;
  for (int g = 1; g <= s->ntpguest; ++g) {
    int mol = s->locguest[g], natms = s->numatoms[mol], nmol = s->nummols[mol];
    int ii = 0;
    for (int im = 1; im <= nmol; ++im)
      for (int is = 1; is <= natms; ++is) {
        ++ii;
        ++jj;
        s->xxx[jj] = s->molx[mol][ii];
        s->yyy[jj] = s->moly[mol][ii];
        s->zzz[jj] = s->molz[mol][ii];
        s->atmcharge[jj] = s->atmchg[mol][is];
        s->lfreezesite[jj] = s->lfzsite[mol][is];
        s->ltype[jj] = s->ltpsit[mol][is];
        s->moltype[jj] = mol;
      }
  }
  for (int f = 1; f <= s->ntpfram; ++f) {
    int mol = s->locfram[f], natms = s->numatoms[mol], nmol = s->nummols[mol];
    int ii = 0;
    for (int im = 1; im <= nmol; ++im)
      for (int k = 1; k <= natms; ++k) {
        ++jj;
        ++ii;
        s->xxx[jj] = s->molx[mol][ii];
        s->yyy[jj] = s->moly[mol][ii];
        s->zzz[jj] = s->molz[mol][ii];
        s->atmcharge[jj] = s->fwchg[mol][k];
        s->lfreezesite[jj] = s->fwfrz[mol][k];
        s->ltype[jj] = s->fwtyp[mol][k];
        s->moltype[jj] = mol;
      }
  }
  s->totatm = jj;
}

static double site_chg(const orc_state *s, int mol, int k) {
  return s->isguest[mol] ? s->atmchg[mol][k] : s->fwchg[mol][k];
}
static int site_typ(const orc_state *s, int mol, int k) {
  return s->isguest[mol] ? s->ltpsit[mol][k] : s->fwtyp[mol][k];
}
static int site_frz(const orc_state *s, int mol, int k) {
  return s->isguest[mol] ? s->lfzsite[mol][k] : s->fwfrz[mol][k];
}

/* lrcorrect LJ branch, vdw_terms.f:468-684 */
static void lrcorrect(orc_state *s) {
  double twopi = 2.0 * PI;
  double volm = s->volm;
  double elrc_mol0[64];
  memset(s->numtyp, 0, sizeof s->numtyp);
  memset(s->numfrz, 0, sizeof s->numfrz);
  memset(s->numtyp_mol, 0, sizeof s->numtyp_mol);
  memset(s->numfrz_mol, 0, sizeof s->numfrz_mol);
  memset(s->numtyp_gstmol, 0, sizeof s->numtyp_gstmol);
  memset(s->numfrz_gstmol, 0, sizeof s->numfrz_gstmol);
  for (int i = 0; i < 64; ++i) s->elrc_mol[i] = 0.0;
  for (int g = 1; g <= s->ntpguest; ++g) {
    int mol = s->locguest[g];
    for (int j = 1; j <= s->numatoms[mol]; ++j) {
      int ka = s->ltpsit[mol][j];
      s->numtyp_gstmol[g][ka] += 1;
      if (s->lfzsite[mol][j] != 0) s->numfrz_gstmol[g][ka] += 1;
    }
  }
  for (int i = 1; i <= s->totatm; ++i) {
    int ka = s->ltype[i], mol = s->moltype[i];
    s->numtyp[ka] += 1;
    s->numtyp_mol[mol][ka] += 1;
    if (s->lfreezesite[i] != 0) {
      s->numfrz[ka] += 1;
      s->numfrz_mol[mol][ka] += 1;
    }
  }
  s->elrc = 0.0;
  int ivdw = 0;
  for (int i = 1; i <= s->ntpatm; ++i)
    for (int j = 1; j <= i; ++j) {
      double eadd = 0.0;
      ++ivdw;
      int k = s->lstvdw[ivdw];
      if (s->ltpvdw[k] == 2) {
        double eps = s->prm_eps[k], sig = s->prm_sig[k];
        double sig6 = ipow6(sig), rc3 = s->rcut * s->rcut * s->rcut;
        double rc9 = rc3 * rc3 * rc3;
        eadd = 4.0 * eps * (sig6 * sig6 / (9.0 * rc9) - sig6 / (3.0 * rc3));
      }
      if (i != j) eadd = eadd * 2.0;
      s->steadd[ivdw] = eadd;
      double denprd = twopi * ((double)s->numtyp[i] * (double)s->numtyp[j] -
                               (double)s->numfrz[i] * (double)s->numfrz[j]) / (volm * volm);
      s->elrc = s->elrc + volm * denprd * eadd;
      for (int t = 0; t < 64; ++t) elrc_mol0[t] = 0.0;
      for (int it = 1; it <= s->maxmls; ++it)
        for (int jt = 1; jt <= s->maxmls; ++jt) {
          int kt = loc2(it, jt);
          double natyp = s->numtyp_mol[it][i], nbtyp = s->numtyp_mol[jt][j];
          double nafrz = s->numfrz_mol[it][i], nbfrz = s->numfrz_mol[jt][j];
          elrc_mol0[kt] += twopi * (natyp * nbtyp - nafrz * nbfrz) / (volm * volm);
        }
      for (int t = 1; t <= s->mxcmls; ++t) s->elrc_mol[t] += volm * eadd * elrc_mol0[t];
    }
}

/* gstlrcorrect, vdw_terms.f:686-776 */
static double gstlrcorrect(orc_state *s, int iguest, int lexisting) {
  double twopi = 2.0 * PI, volm = s->volm, delrc = 0.0;
  int mol = s->locguest[iguest];
  int ivdw = 0;
  for (int i = 1; i <= s->ntpatm; ++i)
    for (int j = 1; j <= i; ++j) {
      ++ivdw;
      double eadd = s->steadd[ivdw];
      double natyp = s->numtyp_gstmol[iguest][i], nbtyp = s->numtyp_gstmol[iguest][j];
      double nafrz = s->numfrz_gstmol[iguest][i], nbfrz = s->numfrz_gstmol[iguest][j];
      double denprd = (natyp * (double)s->numtyp[j] + nbtyp * (double)s->numtyp[i] -
                       nafrz * (double)s->numfrz[j] - nbfrz * (double)s->numfrz[i]);
      if (lexisting) denprd = -1.0 * denprd;
      denprd = denprd + (natyp * nbtyp - nafrz * nbfrz);
      delrc = delrc + volm * twopi * (denprd) / (volm * volm) * eadd;
      double delrc_mol0[64];
      for (int t = 0; t < 64; ++t) delrc_mol0[t] = 0.0;
      for (int jt = 1; jt <= s->maxmls; ++jt) {
        int kt = loc2(mol, jt);
        double nctyp = s->numtyp_mol[jt][j], ndtyp = s->numtyp_mol[jt][i];
        double ncfrz = s->numfrz_mol[jt][j], ndfrz = s->numfrz_mol[jt][i];
        double sic = (natyp * nctyp + nbtyp * ndtyp - nafrz * ncfrz - nbfrz * ndfrz);
        double self = 0.0;
        if (jt == mol) {
          self = (natyp * nbtyp - nafrz * nbfrz);
          if (lexisting) sic = -1.0 * sic;
        }
        delrc_mol0[kt] += twopi * (sic + self) / (volm * volm);
      }
      for (int t = 1; t <= s->mxcmls; ++t) s->delrc_mol[t] += volm * eadd * delrc_mol0[t];
    }
  return delrc;
}

/* ------------------------------------------------------- trial pieces -- */
/* guestlistgen, utility_pack.f:3711-3774 */
static void guestlistgen(orc_state *s, int natms) {
  double rclim = (s->rcut + s->ddelr) * (s->rcut + s->ddelr);
  for (int i = 1; i <= natms; ++i) {
    s->gstlentry[i] = 0;
    int ii = 0;
    for (int j = 1; j <= s->totatm; ++j) {
      int chk = 1;
      for (int it = 1; it <= natms; ++it)
        if (j == s->ind[it]) chk = 0;
      if (chk) {
        ++ii;
        s->xdf[ii] = s->newx[i] - s->xxx[j];
        s->ydf[ii] = s->newy[i] - s->yyy[j];
        s->zdf[ii] = s->newz[i] - s->zzz[j];
      }
    }
    images(s, ii, s->xdf, s->ydf, s->zdf);
    ii = 0;
    for (int j = 1; j <= s->totatm; ++j) {
      int chk = 1;
      for (int it = 1; it <= natms; ++it)
        if (j == s->ind[it]) chk = 0;
      if (chk) {
        ++ii;
        double rsq = s->xdf[ii] * s->xdf[ii] + s->ydf[ii] * s->ydf[ii] + s->zdf[ii] * s->zdf[ii];
        if (rsq <= rclim) {
          s->gstlentry[i] += 1;
          s->gstlist[i][s->gstlentry[i]] = j;
        }
      }
    }
  }
}

/* ewald2, ewald.f:663-733 */
static double ewald2(orc_state *s, double chg, int natm, int mol) {
  double rcsq = s->rcut * s->rcut, rdrewd = 1.0 / s->drewd, engcpe = 0.0;
  double chgea = chg / s->epsq * R4PIE0;
  for (int i = 1; i <= natm; ++i) {
    int jatm = s->ilist[i], jmol = s->moldf[i];
    double chgprd = chgea * s->atmcharge[jatm];
    if (fabs(chgprd) > 1e-10) {
      double rsq = s->rsqdf[i];
      if (rcsq > rsq) {
        double rrr = sqrt(rsq);
        int ll = (int)(rrr * rdrewd);
        double ppp = rrr * rdrewd - (double)ll;
        double vk0 = s->erc[ll], vk1 = s->erc[ll + 1], vk2 = s->erc[ll + 2];
        double t1 = vk0 + (vk1 - vk0) * ppp;
        double t2 = vk1 + (vk2 - vk1) * (ppp - 1.0);
        double erfcr = (t1 + (t2 - t1) * ppp * 0.5) * chgprd;
        engcpe = engcpe + erfcr;
        s->ewald2en[loc2(mol, jmol)] += erfcr;
      }
    }
  }
  return engcpe;
}

/* srfrce, vdw_terms.f:382-466 */
static double srfrce(orc_state *s, int sittyp, int ik, int mol) {
  double rcsq = s->rcut * s->rcut, rdr = 1.0 / s->dlrpot, engsrp = 0.0;
  double ai = (double)sittyp;
  for (int m = 1; m <= ik; ++m) {
    int jatm = s->ilist[m], jmol = s->moldf[m];
    double aj = (double)s->ltype[jatm], ab;
    if (ai > aj)
      ab = ai * (ai - 1.0) * 0.5 + aj + 0.5;
    else
      ab = aj * (aj - 1.0) * 0.5 + ai + 0.5;
    int k = s->lstvdw[(int)ab];
    if (s->ltpvdw[k] < 100 && fabs(VVV(1, k)) > 1e-10) {
      double rsq = s->rsqdf[m];
      if (rcsq > rsq) {
        double rrr = sqrt(rsq);
        int l = (int)(rrr * rdr);
        double ppp = rrr * rdr - (double)l;
        double vk0 = VVV(l, k), vk1 = VVV(l + 1, k), vk2 = VVV(l + 2, k);
        double t1 = vk0 + (vk1 - vk0) * ppp;
        double t2 = vk1 + (vk2 - vk1) * (ppp - 1.0);
        double omega = t1 + (t2 - t1) * ppp * 0.5;
        engsrp = omega + engsrp;
        s->vdwen[loc2(jmol, mol)] += omega;
      }
    }
  }
  return engsrp;
}

/* ewald1_guest, ewald.f:32-352 */
static double ewald1_guest(orc_state *s, int natms, int iguest, double *chgtmp, double *engsictmp,
                           int lexisting) {
  double twopi = 2.0 * PI;
  int mol = s->locguest[iguest], maxmls = s->maxmls;
  double engcpe = 0.0;
  for (int ik = 1; ik <= s->newld; ++ik) {
    SK(ckcsnew, mol, ik) = 0.0;
    SK(ckssnew, mol, ik) = 0.0;
    SK(ckcsnew, maxmls + 1, ik) = 0.0;
    SK(ckssnew, maxmls + 1, ik) = 0.0;
  }
  double rvolm = twopi / s->volm;
  const double *rc = s->rcell;
  *chgtmp = 0.0;
  *engsictmp = 0.0;
  for (int n = 1; n <= natms; ++n) {
    double chge = s->atmchg[mol][n];
    *chgtmp += chge;
    *engsictmp += chge * chge;
  }
  *engsictmp = -R4PIE0 / s->epsq * s->alpha * (*engsictmp) / SQRPI;
  int K2 = s->kmax2 + 1, K3 = s->kmax3 + 1;
  for (int i = 1; i <= natms; ++i) {
    double ssx = rc[1] * s->newx[i] + rc[4] * s->newy[i] + rc[7] * s->newz[i];
    double ssy = rc[2] * s->newx[i] + rc[5] * s->newy[i] + rc[8] * s->newz[i];
    double ssz = rc[3] * s->newx[i] + rc[6] * s->newy[i] + rc[9] * s->newz[i];
    s->elc0[i] = 1.0;
    s->els0[i] = 0.0;
    s->emc[i * K2 + 0] = 1.0;
    s->ems[i * K2 + 0] = 0.0;
    s->enc[i * K3 + 0] = 1.0;
    s->ens[i * K3 + 0] = 0.0;
    s->elc1[i] = cos(twopi * ssx);
    s->emc[i * K2 + 1] = cos(twopi * ssy);
    s->enc[i * K3 + 1] = cos(twopi * ssz);
    s->els1[i] = sin(twopi * ssx);
    s->ems[i * K2 + 1] = sin(twopi * ssy);
    s->ens[i * K3 + 1] = sin(twopi * ssz);
  }
  for (int l = 2; l <= s->kmax2; ++l)
    for (int i = 1; i <= natms; ++i) {
      s->emc[i * K2 + l] = s->emc[i * K2 + l - 1] * s->emc[i * K2 + 1] - s->ems[i * K2 + l - 1] * s->ems[i * K2 + 1];
      s->ems[i * K2 + l] = s->ems[i * K2 + l - 1] * s->emc[i * K2 + 1] + s->emc[i * K2 + l - 1] * s->ems[i * K2 + 1];
    }
  for (int l = 2; l <= s->kmax3; ++l)
    for (int i = 1; i <= natms; ++i) {
      s->enc[i * K3 + l] = s->enc[i * K3 + l - 1] * s->enc[i * K3 + 1] - s->ens[i * K3 + l - 1] * s->ens[i * K3 + 1];
      s->ens[i * K3 + l] = s->ens[i * K3 + l - 1] * s->enc[i * K3 + 1] + s->enc[i * K3 + l - 1] * s->ens[i * K3 + 1];
    }
  int curl = 0;
  for (int kkk = 1; kkk <= s->newld; ++kkk) {
    const kvec_t *kv = &s->kv[kkk];
    while (curl < kv->ll) {  /* advance elc(i,0) to power ll (ewald.f:164-181) */
      ++curl;
      for (int i = 1; i <= natms; ++i) {
        if (curl == 1) {
          s->elc0[i] = s->elc1[i];
          s->els0[i] = s->els1[i];
        } else {
          double cs = s->elc0[i];
          s->elc0[i] = cs * s->elc1[i] - s->els0[i] * s->els1[i];
          s->els0[i] = s->els0[i] * s->elc1[i] + cs * s->els1[i];
        }
      }
    }
    int m = abs(kv->mm), n = abs(kv->nn);
    double ckcs = 0.0, ckss = 0.0;
    for (int i = 1; i <= natms; ++i) {
      double clm, slm;
      if (kv->mm >= 0) {
        clm = s->elc0[i] * s->emc[i * K2 + m] - s->els0[i] * s->ems[i * K2 + m];
        slm = s->els0[i] * s->emc[i * K2 + m] + s->ems[i * K2 + m] * s->elc0[i];
      } else {
        clm = s->elc0[i] * s->emc[i * K2 + m] + s->els0[i] * s->ems[i * K2 + m];
        slm = s->els0[i] * s->emc[i * K2 + m] - s->ems[i * K2 + m] * s->elc0[i];
      }
      double tchge = s->atmchg[mol][i];
      double tenc = s->enc[i * K3 + n], tens = s->ens[i * K3 + n];
      double ckc, cks;
      if (kv->nn >= 0) {
        ckc = tchge * (clm * tenc - slm * tens);
        cks = tchge * (slm * tenc + clm * tens);
      } else {
        ckc = tchge * (clm * tenc + slm * tens);
        cks = tchge * (slm * tenc - clm * tens);
      }
      ckcs = ckcs + ckc;
      ckss = ckss + cks;
    }
    double ckcold = SK(ckcsum, maxmls + 1, kkk), cksold = SK(ckssum, maxmls + 1, kkk);
    SK(ckcsnew, mol, kkk) += ckcs;
    SK(ckssnew, mol, kkk) += ckss;
    SK(ckcsnew, maxmls + 1, kkk) += ckcs;
    SK(ckssnew, maxmls + 1, kkk) += ckss;
    double akk = kv->akk;
    double sumsqd = ckcs * ckcs + ckss * ckss;
    double mixterm = 2.0 * ckcs * ckcold + 2.0 * ckss * cksold;
    if (lexisting) mixterm = -1.0 * mixterm;
    engcpe = engcpe + akk * (mixterm + sumsqd);
    for (int i = 1; i <= maxmls; ++i) {
      double fkk = 2.0 * akk;
      if (i == mol) fkk = akk;
      int kmol = loc2(i, mol);
      s->ewald1en[kmol] += fkk * (SK(ckcsnew, mol, kkk) * SK(ckcsum, i, kkk) +
                                  SK(ckssnew, mol, kkk) * SK(ckssum, i, kkk));
    }
  }
  double chg = *chgtmp, alpha = s->alpha, volm = s->volm;
  double qfix = -(0.5 * PI * R4PIE0 / s->epsq) * ((chg * s->sumchg) / alpha + (chg / alpha) * (chg / alpha)) / volm;
  if (lexisting) {
    *engsictmp = -*engsictmp;
    qfix = -qfix;
  }
  engcpe = 2.0 * rvolm * R4PIE0 * engcpe / s->epsq + *engsictmp + qfix;
  for (int i = 1; i <= maxmls; ++i) {
    double fkk = 1.0, sic = 0.0;
    if (i == mol) {
      fkk = 0.5;
      sic = *engsictmp;
    }
    double qfixmol = -(fkk * PI * R4PIE0 / s->epsq) * (s->chgsum_mol[i] * chg) / (alpha * alpha * volm);
    if (lexisting) {
      sic = -sic;
      qfixmol = -qfixmol;
    }
    int kmol = loc2(i, mol);
    s->ewald1en[kmol] = 2.0 * rvolm * R4PIE0 * s->ewald1en[kmol] / s->epsq + sic + qfixmol;
  }
  return engcpe;
}

static void zero_components(orc_state *s) {
  for (int ik = 1; ik <= s->mxcmls; ++ik) {
    s->delE[ik] = 0.0;
    s->ewald1en[ik] = 0.0;
    s->ewald2en[ik] = 0.0;
    s->vdwen[ik] = 0.0;
    s->delrc_mol[ik] = 0.0;
  }
}

/* per-site pair loop shared by insertion / guest_energy / deletion:
 * gathers the neighbour list, applies images, then surface_check
 * (mc_moves.f:10-44: *surf if a frozen neighbour lies closer than
 * surftol + sigma of the pair, prmvdw(ivdw,2); the overlap test) and
 * ewald2 + srfrce. */
static int site_pairs(orc_state *s, int i, int mol, double chg, int sittyp, int stop_on_overlap,
                      double *e2, double *ev, int *overlap, int *surf) {
  int ik = 0;
  for (int j = 1; j <= s->gstlentry[i]; ++j) {
    ++ik;
    int jatm = s->gstlist[i][j];
    s->ilist[j] = jatm;
    s->moldf[j] = s->moltype[jatm];
    s->xdf[ik] = s->newx[i] - s->xxx[jatm];
    s->ydf[ik] = s->newy[i] - s->yyy[jatm];
    s->zdf[ik] = s->newz[i] - s->zzz[jatm];
  }
  images(s, ik, s->xdf, s->ydf, s->zdf);
  for (int l = 1; l <= s->gstlentry[i]; ++l) {
    s->rsqdf[l] = s->xdf[l] * s->xdf[l] + s->ydf[l] * s->ydf[l] + s->zdf[l] * s->zdf[l];
    {
      int jatm = s->ilist[l];
      double aa = (double)sittyp, ab = (double)s->ltype[jatm], ak;
      if (aa > ab)
        ak = aa * (aa - 1.0) * 0.5 + ab + 0.5;
      else
        ak = ab * (ab - 1.0) * 0.5 + aa + 0.5;
      double req = s->prm_sig[s->lstvdw[(int)ak]];
      double tol = (s->surftol + req) * (s->surftol + req);
      if (s->rsqdf[l] < tol && s->lfreezesite[jatm] == 1) *surf = 1;
    }
    if (s->rsqdf[l] < s->overlap) *overlap = 1;
    if (*overlap && stop_on_overlap) return 1;
  }
  *e2 = ewald2(s, chg, s->gstlentry[i], mol);
  *ev = srfrce(s, sittyp, s->gstlentry[i], mol);
  return 0;
}

/* get_guest, utility_pack.f:3776-3807 */
static void get_guest(orc_state *s, int iguest, int choice) {
  int mol = s->locguest[iguest], natms = s->numatoms[mol];
  int atmadd = 0;
  for (int i = 1; i <= iguest - 1; ++i) {
    int imol = s->locguest[i];
    atmadd += s->numatoms[imol] * s->nummols[imol];
  }
  int at = (choice - 1) * natms + 1, iatm = 0;
  for (int i = at; i <= at - 1 + natms; ++i) {
    ++iatm;
    s->ind[iatm] = atmadd + i;
    s->newx[iatm] = s->molx[mol][i];
    s->newy[iatm] = s->moly[mol][i];
    s->newz[iatm] = s->molz[mol][i];
  }
}

/* insertion, mc_moves.f:150-276 */
static int insertion(orc_state *s, int iguest, fmc_trial *out) {
  int mol = s->locguest[iguest], natms = s->numatoms[mol];
  int loverlap = 0, lnewsurf = 0;
  double ewld2sum = 0.0, vdwsum = 0.0, estep = 0.0;
  for (int i = 0; i <= FMC_MAX_SITES; ++i) s->ind[i] = 0;
  zero_components(s);
  guestlistgen(s, natms);
  double delrc = gstlrcorrect(s, iguest, 0);
  for (int i = 1; i <= natms; ++i) {
    double e2 = 0.0, ev = 0.0;
    if (site_pairs(s, i, mol, s->atmchg[mol][i], s->ltpsit[mol][i], 1, &e2, &ev, &loverlap, &lnewsurf)) {
      memset(out, 0, sizeof *out);
      out->overlap = 1;
      out->surface = lnewsurf;
      s->lnewsurf = lnewsurf;
      s->pend_delrc = delrc;
      return 1;
    }
    ewld2sum += e2;
    vdwsum += ev;
  }
  double chgtmp, engsictmp;
  double ewld1eng = ewald1_guest(s, natms, iguest, &chgtmp, &engsictmp, 0);
  double ewld3sum = s->ewald3en[mol];
  estep = estep + (ewld1eng + ewld2sum - ewld3sum + vdwsum + delrc) / s->engunit;
  for (int i = 1; i <= s->maxmls; ++i) {
    int ik = loc2(mol, i);
    if (i != mol) s->delE[i] += (s->ewald1en[ik] + s->ewald2en[ik] + s->vdwen[ik] + s->delrc_mol[ik]) / s->engunit;
  }
  s->delE[mol] += estep;
  memset(out, 0, sizeof *out);
  out->estep = estep;
  out->surface = lnewsurf;
  s->lnewsurf = lnewsurf;
  out->recip = (ewld1eng - ewld3sum) / s->engunit;
  out->real = ewld2sum / s->engunit;
  out->vdw = vdwsum / s->engunit;
  out->tail = delrc / s->engunit;
  for (int i = 1; i <= s->maxmls && i <= FMC_MAX_SPECIES; ++i) out->delE[i - 1] = s->delE[i];
  s->pend_delrc = delrc;
  s->pend_chgtmp = chgtmp;
  s->pend_engsictmp = engsictmp;
  return 0;
}

/* guest_energy, mc_moves.f:277-374: returns estep, sets *ov on overlap */
static double guest_energy(orc_state *s, int iguest, int lexisting, int *ov, double comps[3], int *lsurf) {
  int mol = s->locguest[iguest], natms = s->numatoms[mol];
  double ewld2sum = 0.0, vdwsum = 0.0;
  *ov = 0;
  *lsurf = 0;
  zero_components(s);
  guestlistgen(s, natms);
  double chgtmp, engsictmp;
  double ewld1eng = ewald1_guest(s, natms, iguest, &chgtmp, &engsictmp, lexisting);
  for (int i = 1; i <= natms; ++i) {
    int itatm = s->ind[i];
    double e2 = 0.0, ev = 0.0;
    if (site_pairs(s, i, mol, s->atmcharge[itatm], s->ltype[itatm], 1, &e2, &ev, ov, lsurf)) return 0.0;
    ewld2sum += e2;
    vdwsum += ev;
  }
  double ewld3sum = s->ewald3en[mol];
  comps[0] = lexisting ? (-ewld1eng - ewld3sum) : (ewld1eng - ewld3sum);
  comps[1] = ewld2sum;
  comps[2] = vdwsum;
  if (lexisting) return (-ewld1eng + ewld2sum + vdwsum - ewld3sum) / s->engunit;
  return (ewld1eng + ewld2sum + vdwsum - ewld3sum) / s->engunit;
}

/* deletion, mc_moves.f:375-484 */
static int deletion(orc_state *s, int iguest, int choice, fmc_trial *out) {
  double ewld2sum = 0.0, vdwsum = 0.0;
  get_guest(s, iguest, choice);
  int mol = s->locguest[iguest], natms = s->numatoms[mol];
  int loverlap = 0, linitsurf = 0;
  zero_components(s);
  double delrc = gstlrcorrect(s, iguest, 1);
  guestlistgen(s, natms);
  double chgtmp, engsictmp;
  double ewld1eng = ewald1_guest(s, natms, iguest, &chgtmp, &engsictmp, 1);
  for (int i = 1; i <= natms; ++i) {
    int itatm = s->ind[i];
    double e2 = 0.0, ev = 0.0;
    site_pairs(s, i, mol, s->atmcharge[itatm], s->ltype[itatm], 0, &e2, &ev, &loverlap, &linitsurf);
    ewld2sum += e2;
    vdwsum += ev;
  }
  double ewld3sum = s->ewald3en[mol];
  double estep = (-ewld1eng + ewld2sum + vdwsum - ewld3sum) / s->engunit;
  estep = estep - delrc / s->engunit;
  for (int i = 1; i <= s->maxmls; ++i) {
    int ik = loc2(mol, i);
    if (i != mol) s->delE[i] -= (s->ewald1en[ik] + s->ewald2en[ik] + s->vdwen[ik] + s->delrc_mol[ik]) / s->engunit;
  }
  s->delE[mol] -= estep;
  memset(out, 0, sizeof *out);
  out->estep = estep;
  out->overlap = loverlap;
  out->surface = linitsurf << 1;
  s->linitsurf = linitsurf;
  out->recip = (-ewld1eng - ewld3sum) / s->engunit;
  out->real = ewld2sum / s->engunit;
  out->vdw = vdwsum / s->engunit;
  out->tail = -delrc / s->engunit;
  for (int i = 1; i <= s->maxmls && i <= FMC_MAX_SPECIES; ++i) out->delE[i - 1] = s->delE[i];
  s->pend_delrc = delrc;
  s->pend_chgtmp = chgtmp;
  s->pend_engsictmp = engsictmp;
  return 0;
}

/* com, utility_pack.f:3565-3596 */
static void com(const orc_state *s, int natms, int mol, const double *x, const double *y, const double *z,
                double *cx, double *cy, double *cz) {
  double w = 0.0, ax = 0.0, ay = 0.0, az = 0.0;
  for (int i = 1; i <= natms; ++i) {
    double m = s->atmwght[mol][i];
    w += m;
    ax += x[i] * m;
    ay += y[i] * m;
    az += z[i] * m;
  }
  *cx = ax / w;
  *cy = ay / w;
  *cz = az / w;
}

/* translate, utility_pack.f:3161-3218 */
static void translate(orc_state *s, int natms, int mol, double a, double b, double c,
                      double *comx, double *comy, double *comz) {
  for (int i = 1; i <= natms; ++i) {
    s->newx[i] = a + s->newx[i];
    s->newy[i] = b + s->newy[i];
    s->newz[i] = c + s->newz[i];
  }
  com(s, natms, mol, s->newx, s->newy, s->newz, comx, comy, comz);
  double x = *comx, y = *comy, z = *comz;
  const double *rc = s->rcell, *cl = s->cell;
  double ssx = (rc[1] * x + rc[4] * y + rc[7] * z);
  double ssy = (rc[2] * x + rc[5] * y + rc[8] * z);
  double ssz = (rc[3] * x + rc[6] * y + rc[9] * z);
  double xss = ssx - floor(ssx), yss = ssy - floor(ssy), zss = ssz - floor(ssz);
  x = (cl[1] * xss + cl[4] * yss + cl[7] * zss);
  y = (cl[2] * xss + cl[5] * yss + cl[8] * zss);
  z = (cl[3] * xss + cl[6] * yss + cl[9] * zss);
  double sx = x - *comx, sy = y - *comy, sz = z - *comz;
  *comx = x;
  *comy = y;
  *comz = z;
  for (int i = 1; i <= natms; ++i) {
    s->newx[i] += sx;
    s->newy[i] += sy;
    s->newz[i] += sz;
  }
}

/* rotation, utility_pack.f:3480-3524 */
static void rotation(double *xr, double *yr, double *zr, double comx, double comy, double comz, int natms,
                     double q1, double q2, double q3, double q4) {
  double Q[10], tx[FMC_MAX_SITES + 1], ty[FMC_MAX_SITES + 1], tz[FMC_MAX_SITES + 1];
  for (int i = 1; i <= natms; ++i) {
    xr[i] -= comx;
    yr[i] -= comy;
    zr[i] -= comz;
  }
  Q[1] = 2.0 * q2 * q2 + 2.0 * q1 * q1 - 1.0;
  Q[2] = 2.0 * q2 * q3 - 2.0 * q1 * q4;
  Q[3] = 2.0 * q2 * q4 + 2.0 * q1 * q3;
  Q[4] = 2.0 * q2 * q3 + 2.0 * q1 * q4;
  Q[5] = 2.0 * q1 * q1 + 2.0 * q3 * q3 - 1.0;
  Q[6] = 2.0 * q3 * q4 - 2.0 * q1 * q2;
  Q[7] = 2.0 * q2 * q4 - 2.0 * q1 * q3;
  Q[8] = 2.0 * q3 * q4 + 2.0 * q1 * q2;
  Q[9] = 2.0 * q4 * q4 + 2.0 * q1 * q1 - 1.0;
  for (int i = 1; i <= natms; ++i) {
    tx[i] = Q[1] * xr[i] + Q[2] * yr[i] + Q[3] * zr[i];
    ty[i] = Q[4] * xr[i] + Q[5] * yr[i] + Q[6] * zr[i];
    tz[i] = Q[7] * xr[i] + Q[8] * yr[i] + Q[9] * zr[i];
  }
  for (int i = 1; i <= natms; ++i) {
    xr[i] = tx[i] + comx;
    yr[i] = ty[i] + comy;
    zr[i] = tz[i] + comz;
  }
}

/* quaternion from four uniforms: random_ins / random_rot, utility_pack.f:3291-3310,3544-3560 */
static void quat_from(double rand1, double rand2, double rand3, double rand4, double angle_scale,
                      double *q) {
  double theta = rand1 * angle_scale;
  double beta1 = 2.0 * rand2 - 1.0, beta2 = 2.0 * rand3 - 1.0, beta3 = 2.0 * rand4 - 1.0;
  double sintheta = sin(theta / 2.0);
  double norm = sqrt(beta1 * beta1 + beta2 * beta2 + beta3 * beta3);
  q[0] = cos(theta / 2.0);
  q[1] = sintheta * (beta1 / norm);
  q[2] = sintheta * (beta2 / norm);
  q[3] = sintheta * (beta3 / norm);
}

/* displace_guest, gcmc.f:2816-2901 (coordinates supplied by the caller) */
static int displace_eval(orc_state *s, int iguest, int choice, fmc_trial *out, int have_coords,
                         const double *xyz_new, double a, double b, double c, const double *q) {
  int mol = s->locguest[iguest], natms = s->numatoms[mol];
  int maxmls = s->maxmls;
  get_guest(s, iguest, choice);
  int ov = 0, linitsurf = 0, lnewsurf = 0;
  double c0[3] = {0, 0, 0}, c1[3] = {0, 0, 0};
  double enginit = guest_energy(s, iguest, 1, &ov, c0, &linitsurf);
  for (int ik = 1; ik <= s->mxcmls; ++ik) {
    s->ewald1entmp[ik] = s->ewald1en[ik];
    s->ewald2entmp[ik] = s->ewald2en[ik];
    s->vdwentmp[ik] = s->vdwen[ik];
  }
  for (int ik = 1; ik <= s->newld; ++ik) {
    SK(ckcsorig, mol, ik) = SK(ckcsum, mol, ik);
    SK(ckssorig, mol, ik) = SK(ckssum, mol, ik);
    SK(ckcsorig, maxmls + 1, ik) = SK(ckcsum, maxmls + 1, ik);
    SK(ckssorig, maxmls + 1, ik) = SK(ckssum, maxmls + 1, ik);
    SK(ckcsum, mol, ik) -= SK(ckcsnew, mol, ik);
    SK(ckssum, mol, ik) -= SK(ckssnew, mol, ik);
    SK(ckcsum, maxmls + 1, ik) -= SK(ckcsnew, maxmls + 1, ik);
    SK(ckssum, maxmls + 1, ik) -= SK(ckssnew, maxmls + 1, ik);
  }
  if (have_coords) {
    for (int i = 1; i <= natms; ++i) {
      s->newx[i] = xyz_new[3 * (i - 1) + 0];
      s->newy[i] = xyz_new[3 * (i - 1) + 1];
      s->newz[i] = xyz_new[3 * (i - 1) + 2];
    }
  } else {
    double comx, comy, comz;
    translate(s, natms, mol, a, b, c, &comx, &comy, &comz);
    rotation(s->newx, s->newy, s->newz, comx, comy, comz, natms, q[0], q[1], q[2], q[3]);
  }
  int ov2 = 0;
  double estep = guest_energy(s, iguest, 0, &ov2, c1, &lnewsurf);
  estep = estep - enginit;
  s->linitsurf = linitsurf;
  s->lnewsurf = lnewsurf;
  for (int i = 1; i <= maxmls; ++i) {
    int ik = loc2(mol, i);
    if (i != mol)
      s->delE[i] += (s->ewald1en[ik] - s->ewald1entmp[ik] + s->ewald2en[ik] - s->ewald2entmp[ik] +
                     s->vdwen[ik] - s->vdwentmp[ik]) / s->engunit;
  }
  s->delE[mol] += estep;
  memset(out, 0, sizeof *out);
  out->estep = estep;
  out->overlap = ov2;
  out->surface = lnewsurf | (linitsurf << 1);
  out->recip = (c1[0] - c0[0]) / s->engunit;
  out->real = (c1[1] - c0[1]) / s->engunit;
  out->vdw = (c1[2] - c0[2]) / s->engunit;
  for (int i = 1; i <= s->maxmls && i <= FMC_MAX_SPECIES; ++i) out->delE[i - 1] = s->delE[i];
  return 0;
}

static void ensure_cap(orc_state *s, int mol, int natoms_needed) {
  if (natoms_needed + 1 <= s->molcap[mol]) return;
  int cap = s->molcap[mol] * 2 + natoms_needed + 16;
  s->molx[mol] = (double *)realloc(s->molx[mol], sizeof(double) * (size_t)cap);
  s->moly[mol] = (double *)realloc(s->moly[mol], sizeof(double) * (size_t)cap);
  s->molz[mol] = (double *)realloc(s->molz[mol], sizeof(double) * (size_t)cap);
  s->molcap[mol] = cap;
}

/* accept_move, mc_moves.f:485-621 */
static void accept_move(orc_state *s, int kind, int iguest, int choice) {
  int mol = s->locguest[iguest], kk = loc2(mol, mol), natms = s->numatoms[mol], maxmls = s->maxmls;
  int at = (choice - 1) * natms + 1;
  for (int i = 1; i <= maxmls; ++i) s->energy[i] += s->delE[i];
  if (kind == FMC_MOVE_INSERT) {
    for (int k = 1; k <= s->newld; ++k) {
      SK(ckcsum, mol, k) += SK(ckcsnew, mol, k);
      SK(ckssum, mol, k) += SK(ckssnew, mol, k);
      SK(ckcsum, maxmls + 1, k) += SK(ckcsnew, maxmls + 1, k);
      SK(ckssum, maxmls + 1, k) += SK(ckssnew, maxmls + 1, k);
    }
    int mm = natms * s->nummols[mol];
    s->engsic[kk] += s->pend_engsictmp;
    s->chgsum_mol[mol] += s->pend_chgtmp;
    s->sumchg += s->pend_chgtmp;
    ensure_cap(s, mol, mm + natms);
    for (int i = 1; i <= natms; ++i) {
      s->molx[mol][mm + i] = s->newx[i];
      s->moly[mol][mm + i] = s->newy[i];
      s->molz[mol][mm + i] = s->newz[i];
      int ka = s->ltpsit[mol][i];
      s->numtyp[ka] += 1;
      s->numtyp_mol[mol][ka] += 1;
      if (s->lfzsite[mol][i] != 0) {
        s->numfrz[ka] += 1;
        s->numfrz_mol[mol][ka] += 1;
      }
    }
    s->elrc += s->pend_delrc;
    for (int t = 1; t <= s->mxcmls; ++t) s->elrc_mol[t] += s->delrc_mol[t];
    s->nummols[mol] += 1;
    condense(s);
  } else if (kind == FMC_MOVE_DELETE) {
    for (int k = 1; k <= s->newld; ++k) {
      SK(ckcsum, mol, k) -= SK(ckcsnew, mol, k);
      SK(ckssum, mol, k) -= SK(ckssnew, mol, k);
      SK(ckcsum, maxmls + 1, k) -= SK(ckcsnew, maxmls + 1, k);
      SK(ckssum, maxmls + 1, k) -= SK(ckssnew, maxmls + 1, k);
    }
    s->engsic[kk] -= s->pend_engsictmp;
    s->chgsum_mol[mol] -= s->pend_chgtmp;
    s->sumchg -= s->pend_chgtmp;
    int mm = natms * s->nummols[mol];
    int iatm = 0;
    ensure_cap(s, mol, mm + natms);
    for (int i = at; i <= mm; ++i) {
      ++iatm;
      s->molx[mol][i] = (i + natms <= mm) ? s->molx[mol][i + natms] : 0.0;
      s->moly[mol][i] = (i + natms <= mm) ? s->moly[mol][i + natms] : 0.0;
      s->molz[mol][i] = (i + natms <= mm) ? s->molz[mol][i + natms] : 0.0;
      if (iatm <= natms) {
        int ka = s->ltpsit[mol][iatm];
        s->numtyp[ka] -= 1;
        s->numtyp_mol[mol][ka] -= 1;
        if (s->lfzsite[mol][iatm] != 0) {
          s->numfrz[ka] -= 1;
          s->numfrz_mol[mol][ka] -= 1;
        }
      }
    }
    s->elrc -= s->pend_delrc;
    for (int t = 1; t <= s->mxcmls; ++t) s->elrc_mol[t] -= s->delrc_mol[t];
    s->nummols[mol] -= 1;
    condense(s);
  } else {
    for (int k = 1; k <= s->newld; ++k) {
      SK(ckcsum, mol, k) += SK(ckcsnew, mol, k);
      SK(ckssum, mol, k) += SK(ckssnew, mol, k);
      SK(ckcsum, maxmls + 1, k) += SK(ckcsnew, maxmls + 1, k);
      SK(ckssum, maxmls + 1, k) += SK(ckssnew, maxmls + 1, k);
    }
    int mm = 0;
    for (int i = at; i <= at - 1 + natms; ++i) {
      ++mm;
      s->molx[mol][i] = s->newx[mm];
      s->moly[mol][i] = s->newy[mm];
      s->molz[mol][i] = s->newz[mm];
    }
    condense(s);
  }
}

static void restore_sk(orc_state *s, int mol) {  /* gcmc.f:1921-1930 */
  int maxmls = s->maxmls;
  for (int ik = 1; ik <= s->newld; ++ik) {
    SK(ckcsum, mol, ik) = SK(ckcsorig, mol, ik);
    SK(ckssum, mol, ik) = SK(ckssorig, mol, ik);
    SK(ckcsum, maxmls + 1, ik) = SK(ckcsorig, maxmls + 1, ik);
    SK(ckssum, maxmls + 1, ik) = SK(ckssorig, maxmls + 1, ik);
  }
}

/* energy_eval, utility_pack.f:3637-3693 */
static int energy_eval(const orc_state *s, double eng, double rande, int iguest, int kind) {
  double test = 0.0;
  double ipress = s->gstfuga[iguest];
  int ingsts = s->nummols[s->locguest[iguest]];
  double beta = s->beta;
  if (-1 * beta * eng > 700.0)
    test = exp(700.0);
  else if (beta * eng > 700.0)
    test = exp(-700.0);
  else if (kind == FMC_MOVE_DISPLACE)
    test = exp(-1.0 * beta * eng);
  else if (kind == FMC_MOVE_INSERT)
    test = s->statvolm * ipress / (((double)ingsts + 1.0) * BOLTZ * s->temp) * exp(-1.0 * beta * eng);
  else if (kind == FMC_MOVE_DELETE) {
    if (fabs(ipress) < 1e-7)
      test = 1.0;
    else
      test = (double)ingsts * BOLTZ * s->temp / (s->statvolm * ipress) * exp(-1.0 * beta * eng);
  }
  return rande < test;
}

/* ------------------------------------------------------- single point -- */
/* single_point, gcmc.f:2549-2661, with the Brode-Ahlrichs half list of parlst
 * (nlist_builders.f:1-185: frozen-frozen and intramolecular pairs skipped). */
static void ewald1_full(orc_state *s, double *engcpe_out) {  /* ewald.f:354-661 */
  double twopi = 2.0 * PI, rvolm = twopi / s->volm;
  int maxmls = s->maxmls, mxatm = s->totatm;
  double engsicold = 0.0;
  size_t nsk = (size_t)(s->newld + 1) * (size_t)(maxmls + 2);
  for (size_t t = 0; t < nsk; ++t) {
    s->ckcsum[t] = 0.0;
    s->ckssum[t] = 0.0;
  }
  for (int i = 0; i < 64; ++i) {
    s->engsic[i] = 0.0;
    s->ewald1en[i] = 0.0;
  }
  for (int i = 0; i <= FMC_MAX_SPECIES; ++i) s->chgsum_mol[i] = 0.0;
  for (int n = 1; n <= mxatm; ++n) {
    int mol = s->moltype[n];
    double tchge = s->atmcharge[n] * s->atmcharge[n];
    int ik = loc2(mol, mol);
    engsicold += tchge;
    s->engsic[ik] += tchge;
    s->chgsum_mol[mol] += s->atmcharge[n];
  }
  engsicold = -R4PIE0 / s->epsq * s->alpha * engsicold / SQRPI;
  const double *rc = s->rcell;
  double *cx = (double *)malloc(sizeof(double) * (size_t)(mxatm + 1));
  double *sx = (double *)malloc(sizeof(double) * (size_t)(mxatm + 1));
  double *c1 = (double *)malloc(sizeof(double) * (size_t)(mxatm + 1));
  double *s1 = (double *)malloc(sizeof(double) * (size_t)(mxatm + 1));
  int K2 = s->kmax2 + 1, K3 = s->kmax3 + 1;
  double *emc = (double *)malloc(sizeof(double) * (size_t)(mxatm + 1) * K2);
  double *ems = (double *)malloc(sizeof(double) * (size_t)(mxatm + 1) * K2);
  double *enc = (double *)malloc(sizeof(double) * (size_t)(mxatm + 1) * K3);
  double *ens = (double *)malloc(sizeof(double) * (size_t)(mxatm + 1) * K3);
  for (int i = 1; i <= mxatm; ++i) {
    double ssx = rc[1] * s->xxx[i] + rc[4] * s->yyy[i] + rc[7] * s->zzz[i];
    double ssy = rc[2] * s->xxx[i] + rc[5] * s->yyy[i] + rc[8] * s->zzz[i];
    double ssz = rc[3] * s->xxx[i] + rc[6] * s->yyy[i] + rc[9] * s->zzz[i];
    cx[i] = 1.0;
    sx[i] = 0.0;
    emc[i * K2] = 1.0;
    ems[i * K2] = 0.0;
    enc[i * K3] = 1.0;
    ens[i * K3] = 0.0;
    c1[i] = cos(twopi * ssx);
    s1[i] = sin(twopi * ssx);
    if (s->kmax2 >= 1) {
      emc[i * K2 + 1] = cos(twopi * ssy);
      ems[i * K2 + 1] = sin(twopi * ssy);
    }
    if (s->kmax3 >= 1) {
      enc[i * K3 + 1] = cos(twopi * ssz);
      ens[i * K3 + 1] = sin(twopi * ssz);
    }
  }
  for (int l = 2; l <= s->kmax2; ++l)
    for (int i = 1; i <= mxatm; ++i) {
      emc[i * K2 + l] = emc[i * K2 + l - 1] * emc[i * K2 + 1] - ems[i * K2 + l - 1] * ems[i * K2 + 1];
      ems[i * K2 + l] = ems[i * K2 + l - 1] * emc[i * K2 + 1] + emc[i * K2 + l - 1] * ems[i * K2 + 1];
    }
  for (int l = 2; l <= s->kmax3; ++l)
    for (int i = 1; i <= mxatm; ++i) {
      enc[i * K3 + l] = enc[i * K3 + l - 1] * enc[i * K3 + 1] - ens[i * K3 + l - 1] * ens[i * K3 + 1];
      ens[i * K3 + l] = ens[i * K3 + l - 1] * enc[i * K3 + 1] + enc[i * K3 + l - 1] * ens[i * K3 + 1];
    }
  double engcpe = 0.0;
  int curl = 0;
  for (int kkk = 1; kkk <= s->newld; ++kkk) {
    const kvec_t *kv = &s->kv[kkk];
    while (curl < kv->ll) {
      ++curl;
      for (int i = 1; i <= mxatm; ++i) {
        if (curl == 1) {
          cx[i] = c1[i];
          sx[i] = s1[i];
        } else {
          double cs = cx[i];
          cx[i] = cs * c1[i] - sx[i] * s1[i];
          sx[i] = sx[i] * c1[i] + cs * s1[i];
        }
      }
    }
    int m = abs(kv->mm), n = abs(kv->nn);
    double ckcs = 0.0, ckss = 0.0;
    for (int i = 1; i <= mxatm; ++i) {
      double clm, slm;
      if (kv->mm >= 0) {
        clm = cx[i] * emc[i * K2 + m] - sx[i] * ems[i * K2 + m];
        slm = sx[i] * emc[i * K2 + m] + ems[i * K2 + m] * cx[i];
      } else {
        clm = cx[i] * emc[i * K2 + m] + sx[i] * ems[i * K2 + m];
        slm = sx[i] * emc[i * K2 + m] - ems[i * K2 + m] * cx[i];
      }
      double tchge = s->atmcharge[i], tenc = enc[i * K3 + n], tens = ens[i * K3 + n];
      double ckc, cks;
      if (kv->nn >= 0) {
        ckc = tchge * (clm * tenc - slm * tens);
        cks = tchge * (slm * tenc + clm * tens);
      } else {
        ckc = tchge * (clm * tenc + slm * tens);
        cks = tchge * (slm * tenc - clm * tens);
      }
      ckcs += ckc;
      ckss += cks;
      int mol = s->moltype[i];
      SK(ckcsum, mol, kkk) += ckc;
      SK(ckssum, mol, kkk) += cks;
      SK(ckcsum, maxmls + 1, kkk) += ckc;
      SK(ckssum, maxmls + 1, kkk) += cks;
    }
    double akk = kv->akk;
    engcpe = engcpe + akk * (ckcs * ckcs + ckss * ckss);
    int kmol = 0;
    for (int i = 1; i <= maxmls; ++i) {
      double fkk = 2.0 * akk;
      for (int j = 1; j <= i; ++j) {
        ++kmol;
        if (i == j) fkk = akk;
        s->ewald1en[kmol] += fkk * (SK(ckcsum, i, kkk) * SK(ckcsum, j, kkk) + SK(ckssum, i, kkk) * SK(ckssum, j, kkk));
      }
    }
  }
  double qfix = -(0.5 * PI * R4PIE0 / s->epsq) * ((s->sumchg / s->alpha) * (s->sumchg / s->alpha) / s->volm);
  engcpe = 2.0 * rvolm * R4PIE0 * engcpe / s->epsq + engsicold + qfix;
  for (int i = 1; i <= maxmls; ++i) {
    double fkk = 1.0;
    for (int j = 1; j <= i; ++j) {
      if (i == j) fkk = 0.5;
      int ik = loc2(i, j);
      double sic = -R4PIE0 / s->epsq * s->alpha * s->engsic[ik] / SQRPI;
      sic = sic - ((fkk * PI * R4PIE0 / s->epsq) * s->chgsum_mol[i] * s->chgsum_mol[j] / (s->alpha * s->alpha * s->volm));
      s->ewald1en[ik] = 2.0 * rvolm * R4PIE0 * s->ewald1en[ik] / s->epsq + sic;
    }
  }
  free(cx); free(sx); free(c1); free(s1); free(emc); free(ems); free(enc); free(ens);
  *engcpe_out = engcpe;
}

/* ewald3, ewald.f:735-790 */
static double ewald3(const orc_state *s, double chg, int mol, int natms, const int *jl,
                     const double *xd, const double *yd, const double *zd, int fw) {
  const double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741;
  const double a4 = -1.453152027, a5 = 1.061405429, pp = 0.3275911;
  const double rr3 = 0.333333333333, r10 = 0.1, r42 = 0.02380952381, r216 = 4.62962962963e-3;
  double engcpe = 0.0, chgea = chg / s->epsq * R4PIE0, alpha = s->alpha;
  for (int m = 1; m <= natms; ++m) {
    int jatm = jl[m];
    double chgprd = chgea * (fw ? s->fwchg[mol][(jatm - 1) % s->numatoms[mol] + 1] : s->atmchg[mol][jatm]);
    double rsq = xd[m] * xd[m] + yd[m] * yd[m] + zd[m] * zd[m];
    double rrr = sqrt(rsq), alpr = rrr * alpha, alpr2 = alpr * alpr, erfr;
    if (alpr < 1e-2) {
      erfr = 2.0 * chgprd * (alpha / SQRPI) * (1.0 + alpr2 * (-rr3 + alpr2 * (r10 + alpr2 * (-r42 + alpr2 * r216))));
    } else {
      double tt = 1.0 / (1.0 + pp * alpha * rrr);
      double exp1 = exp(-(alpha * rrr) * (alpha * rrr));
      erfr = (1.0 - tt * (a1 + tt * (a2 + tt * (a3 + tt * (a4 + tt * a5)))) * exp1) * chgprd / rrr;
    }
    engcpe += erfr;
  }
  return engcpe;
}

/* excluding_ewald_charges guest branch, gcmc.f:2662-2716 (the framework branch
 * only feeds ewald3en(framework), which no trial reads) */
static double excluding_ewald_charges(orc_state *s) {
  double ewald3sum = 0.0;
  for (int i = 0; i <= FMC_MAX_SPECIES; ++i) s->ewald3en[i] = 0.0;
  int jl[FMC_MAX_SITES + 2];
  double xd[FMC_MAX_SITES + 2], yd[FMC_MAX_SITES + 2], zd[FMC_MAX_SITES + 2];
  for (int g = 1; g <= s->ntpguest; ++g) {
    int mm = s->locguest[g], na = s->numatoms[mm];
    double ewald3mol = 0.0;
    for (int iatm = 1; iatm <= na - 1; ++iatm) {
      int ikk = 0;
      for (int jatm = iatm + 1; jatm <= na; ++jatm) {
        ++ikk;
        jl[ikk] = jatm;
        xd[ikk] = s->guestx[g][iatm] - s->guestx[g][jatm];
        yd[ikk] = s->guesty[g][iatm] - s->guesty[g][jatm];
        zd[ikk] = s->guestz[g][iatm] - s->guestz[g][jatm];
      }
      images(s, ikk, xd, yd, zd);
      double e = ewald3(s, s->atmchg[mm][iatm], mm, ikk, jl, xd, yd, zd, 0);
      ewald3mol += e;
      s->ewald3en[mm] += e;
    }
    ewald3sum += (double)s->nummols[mm] * ewald3mol;
  }
  return ewald3sum;
}

int orc_single_point(orc_state *s, double out[5]) {
  for (int ik = 1; ik <= s->maxmls; ++ik) {
    s->vdwen[ik] = 0.0;
    s->ewald2en[ik] = 0.0;
    s->ewald1en[ik] = 0.0;
    s->delrc_mol[ik] = 0.0;
  }
  lrcorrect(s);
  condense(s);
  double ewald1sum;
  ewald1_full(s, &ewald1sum);
  double ewald3eng = excluding_ewald_charges(s);
  /* real-space + LJ over the half list (parlst + single_point loop) */
  double ewald2sum = 0.0, vdwsum = 0.0, rclim = (s->rcut + s->ddelr) * (s->rcut + s->ddelr);
  int natms = s->totatm, mpm2 = natms / 2, npm2 = (natms - 1) / 2;
  int *lentry = (int *)calloc((size_t)natms + 1, sizeof(int));
  int cap = 64;
  int **list = (int **)calloc((size_t)natms + 1, sizeof(int *));
  int *lcap = (int *)calloc((size_t)natms + 1, sizeof(int));
  int *nexatm = (int *)calloc((size_t)natms + 1, sizeof(int));
  int *noxatm = (int *)calloc((size_t)natms + 1, sizeof(int));
  int (*lexatm)[FMC_MAX_SITES + 1] = calloc((size_t)natms + 1, sizeof *lexatm);
  /* exclusion lists of condense (utility_pack.f:3993-4019): guest sites of one molecule */
  {
    int jj = 0, lsite = 0;
    for (int g = 1; g <= s->ntpguest; ++g) {
      int mol = s->locguest[g], na = s->numatoms[mol];
      for (int im = 1; im <= s->nummols[mol]; ++im) {
        for (int is = 1; is <= na; ++is) {
          ++jj;
          int kk = 0;
          for (int k = 1; k <= na; ++k) {
            if (k == is) continue;
            int newatm = k + lsite;
            if (((newatm > jj) && (newatm - jj <= s->totatm / 2)) ||
                ((newatm < jj) && (newatm + s->totatm - jj <= (s->totatm - 1) / 2))) {
              ++kk;
              lexatm[jj][kk] = newatm;
              for (int j = kk; j >= 2; --j)
                if (lexatm[jj][j] < lexatm[jj][j - 1]) {
                  int t = lexatm[jj][j];
                  lexatm[jj][j] = lexatm[jj][j - 1];
                  lexatm[jj][j - 1] = t;
                }
            }
          }
          nexatm[jj] = kk;
        }
        lsite += na;
      }
    }
  }
  for (int i = 1; i <= natms; ++i) noxatm[i] = 1;
  int last = natms;
  for (int m = 1; m <= mpm2; ++m) {
    if (m > npm2) last = mpm2;
    int ii = 0;
    for (int i = 1; i <= last; ++i) {
      int j = i + m;
      if (j > natms) j -= natms;
      ++ii;
      s->xdf[ii] = s->xxx[i] - s->xxx[j];
      s->ydf[ii] = s->yyy[i] - s->yyy[j];
      s->zdf[ii] = s->zzz[i] - s->zzz[j];
    }
    images(s, ii, s->xdf, s->ydf, s->zdf);
    ii = 0;
    for (int i = 1; i <= last; ++i) {
      int lfrzi = s->lfreezesite[i] != 0;
      int j = i + m;
      if (j > natms) j -= natms;
      ++ii;
      if (nexatm[ii] > 0 && lexatm[ii][noxatm[ii]] == j) {
        noxatm[ii] = noxatm[ii] + 1 < nexatm[ii] ? noxatm[ii] + 1 : nexatm[ii];
      } else {
        int ldo = 1;
        if (lfrzi) ldo = (s->lfreezesite[j] == 0);
        if (ldo) {
          double rsq = s->xdf[ii] * s->xdf[ii] + s->ydf[ii] * s->ydf[ii] + s->zdf[ii] * s->zdf[ii];
          if (rsq < rclim && rsq >= 0.0) {
            if (lentry[ii] + 1 >= lcap[ii]) {
              lcap[ii] = lcap[ii] ? 2 * lcap[ii] : cap;
              list[ii] = (int *)realloc(list[ii], sizeof(int) * (size_t)lcap[ii]);
            }
            lentry[ii] += 1;
            list[ii][lentry[ii]] = j;
          }
        }
      }
    }
  }
  int ii = 0;
  for (int mol = 1; mol <= s->maxmls; ++mol) {
    for (int im = 1; im <= s->nummols[mol]; ++im)
      for (int ia = 1; ia <= s->numatoms[mol]; ++ia) {
        ++ii;
        if (lentry[ii] > 0) {
          int ikk = 0;
          for (int j = 1; j <= lentry[ii]; ++j) {
            ++ikk;
            int jatm = list[ii][j];
            s->ilist[j] = jatm;
            s->moldf[j] = s->moltype[jatm];
            s->xdf[ikk] = s->xxx[ii] - s->xxx[jatm];
            s->ydf[ikk] = s->yyy[ii] - s->yyy[jatm];
            s->zdf[ikk] = s->zzz[ii] - s->zzz[jatm];
          }
          images(s, ikk, s->xdf, s->ydf, s->zdf);
          for (int j = 1; j <= lentry[ii]; ++j)
            s->rsqdf[j] = s->xdf[j] * s->xdf[j] + s->ydf[j] * s->ydf[j] + s->zdf[j] * s->zdf[j];
          ewald2sum += ewald2(s, s->atmcharge[ii], lentry[ii], mol);
          vdwsum += srfrce(s, s->ltype[ii], lentry[ii], mol);
        }
      }
  }
  for (int i = 0; i <= natms; ++i) free(list[i]);
  free(list); free(lcap); free(lentry); free(nexatm); free(noxatm); free(lexatm);
  out[0] = (ewald2sum + ewald1sum - ewald3eng) / s->engunit;
  out[1] = ewald1sum / s->engunit;
  out[2] = ewald2sum / s->engunit;
  out[3] = (vdwsum + s->elrc) / s->engunit;
  out[4] = ewald3eng / s->engunit;
  /* initial per-guest energies, gcmc.f:554-582 */
  for (int g = 1; g <= s->ntpguest; ++g) {
    int mol = s->locguest[g];
    double sp = 0.0;
    for (int j = 1; j <= s->maxmls; ++j) {
      int ik = loc2(mol, j);
      sp += (s->ewald1en[ik] + s->ewald2en[ik] + s->elrc_mol[ik] + s->vdwen[ik]) / s->engunit;
    }
    if (s->nummols[mol] > 0) sp -= s->nummols[mol] * s->ewald3en[mol] / s->engunit;
    s->energy[mol] = sp;
  }
  return 0;
}

/* --------------------------------------------------------------- create -- */
orc_state *orc_create(const fmc_system *sys, const fmc_walker *wk) {
  if (sys->imcon < 1 || sys->imcon > 3) return NULL;
  if (sys->nspecies > FMC_MAX_SPECIES || sys->ntypes > FMC_MAX_TYPES) return NULL;
  orc_state *s = (orc_state *)calloc(1, sizeof(orc_state));
  s->imcon = sys->imcon;
  for (int i = 0; i < 9; ++i) s->cell[i + 1] = sys->cell[i];
  invert(s->cell, s->rcell);
  double celprp[11];
  dcell(s->cell, celprp);
  s->volm = celprp[10];
  s->statvolm = s->volm * 1e-30;
  s->rcut = sys->rcut;
  s->rvdw = sys->rvdw > 0 ? sys->rvdw : sys->rcut;
  s->ddelr = sys->delr;
  s->alpha = sys->alpha;
  s->epsq = sys->epsq;
  s->engunit = sys->engunit;
  s->overlap = sys->overlap * sys->overlap;
  s->surftol = sys->surftol;
  s->sumchg = sys->sumchg;
  s->kmax1 = sys->kmax[0];
  s->kmax2 = sys->kmax[1];
  s->kmax3 = sys->kmax[2];
  s->maxmls = sys->nspecies;
  s->mxcmls = s->maxmls * (s->maxmls - 1) / 2 + s->maxmls;
  s->ntpatm = sys->ntypes;
  s->ntpvdw = sys->npairs;
  s->maxvdw = s->ntpvdw > (s->ntpatm * (s->ntpatm + 1)) / 2 ? s->ntpvdw : (s->ntpatm * (s->ntpatm + 1)) / 2;
  s->mxegrd = 1000;
  {
    int g = (int)(s->rvdw / 0.01 + 0.5) + 4;
    if (g > s->mxegrd) s->mxegrd = g;
  }
  s->erc = (double *)calloc((size_t)s->mxegrd + 1, sizeof(double));
  s->vvv = (double *)calloc((size_t)(s->maxvdw + 2) * (size_t)(s->mxegrd + 1), sizeof(double));
  s->lstvdw = (int *)calloc((size_t)s->maxvdw + 2, sizeof(int));
  s->ltpvdw = (int *)calloc((size_t)s->maxvdw + 2, sizeof(int));
  s->prm_eps = (double *)calloc((size_t)s->maxvdw + 2, sizeof(double));
  s->prm_sig = (double *)calloc((size_t)s->maxvdw + 2, sizeof(double));
  s->steadd = (double *)calloc((size_t)s->maxvdw + 2, sizeof(double));
  /* define_van_der_waals, vdw_terms.f:28-174 */
  for (int i = 1; i <= s->maxvdw + 1; ++i) {
    s->lstvdw[i] = 0;
    s->ltpvdw[i] = -1;
  }
  for (int p = 1; p <= s->ntpvdw; ++p) {
    int key = loc2(sys->pair_i[p - 1], sys->pair_j[p - 1]);
    s->lstvdw[key] = p;
    s->ltpvdw[p] = sys->pair_kind[p - 1];
    s->prm_eps[p] = sys->pair_eps[p - 1];
    s->prm_sig[p] = sys->pair_sig[p - 1];
  }
  forgen(s);
  int ntab = (s->ntpatm * (s->ntpatm + 1)) / 2;
  if (s->ntpvdw < ntab)
    for (int i = 1; i <= ntab; ++i)
      if (s->lstvdw[i] == 0) s->lstvdw[i] = s->ntpvdw + 1;
  erfcgen(s);
  build_kvecs(s);
  /* species */
  int nfw_atoms = 0;
  for (int m = 1; m <= s->maxmls; ++m) {
    const fmc_species *sp = &sys->species[m - 1];
    s->isguest[m] = sp->is_guest;
    s->numatoms[m] = sp->natoms;
    s->nummols[m] = sp->nmols;
    if (sp->is_guest) {
      int g = ++s->ntpguest;
      s->locguest[g] = m;
      if (sp->natoms > FMC_MAX_SITES) { orc_destroy(s); return NULL; }
      for (int k = 1; k <= sp->natoms; ++k) {
        s->atmchg[m][k] = sp->site_charge[k - 1];
        s->atmwght[m][k] = sp->site_mass[k - 1];
        s->ltpsit[m][k] = sp->site_type[k - 1];
        s->lfzsite[m][k] = sp->site_frozen[k - 1];
        s->guestx[g][k] = sp->template_xyz[3 * (k - 1) + 0];
        s->guesty[g][k] = sp->template_xyz[3 * (k - 1) + 1];
        s->guestz[g][k] = sp->template_xyz[3 * (k - 1) + 2];
      }
      s->gstfuga[g] = wk->fugacity[g - 1];
      s->delr[g] = wk->delr[g - 1];
      for (int mv = 0; mv < FMC_NMOVES; ++mv) s->mcf[g][mv] = wk->move_cum[g - 1][mv];
    } else {
      int f = ++s->ntpfram;
      s->locfram[f] = m;
      s->fwchg[m] = (double *)calloc((size_t)sp->natoms + 1, sizeof(double));
      s->fwtyp[m] = (int *)calloc((size_t)sp->natoms + 1, sizeof(int));
      s->fwfrz[m] = (int *)calloc((size_t)sp->natoms + 1, sizeof(int));
      for (int k = 1; k <= sp->natoms; ++k) {
        s->fwchg[m][k] = sp->site_charge[k - 1];
        s->fwtyp[m][k] = sp->site_type[k - 1];
        s->fwfrz[m][k] = sp->site_frozen[k - 1];
        if (k <= FMC_MAX_SITES) s->atmwght[m][k] = sp->site_mass[k - 1];
      }
      nfw_atoms += sp->natoms * sp->nmols;
    }
    int n = sp->natoms * sp->nmols;
    s->molcap[m] = n + 16;
    s->molx[m] = (double *)calloc((size_t)s->molcap[m] + 1, sizeof(double));
    s->moly[m] = (double *)calloc((size_t)s->molcap[m] + 1, sizeof(double));
    s->molz[m] = (double *)calloc((size_t)s->molcap[m] + 1, sizeof(double));
    const double *src = sp->xyz;
    for (int a = 1; a <= n && src; ++a) {
      s->molx[m][a] = src[3 * (a - 1) + 0];
      s->moly[m][a] = src[3 * (a - 1) + 1];
      s->molz[m][a] = src[3 * (a - 1) + 2];
    }
  }
  if (wk->framework_xyz) {
    int off = 0;
    for (int f = 1; f <= s->ntpfram; ++f) {
      int m = s->locfram[f];
      for (int a = 1; a <= s->numatoms[m] * s->nummols[m]; ++a, ++off) {
        s->molx[m][a] = wk->framework_xyz[3 * off + 0];
        s->moly[m][a] = wk->framework_xyz[3 * off + 1];
        s->molz[m][a] = wk->framework_xyz[3 * off + 2];
      }
    }
  }
  s->maxalloc = nfw_atoms + MAXGUEST * 3 + 64;
  size_t ma = (size_t)s->maxalloc + 1;
  s->xxx = (double *)calloc(ma, sizeof(double));
  s->yyy = (double *)calloc(ma, sizeof(double));
  s->zzz = (double *)calloc(ma, sizeof(double));
  s->atmcharge = (double *)calloc(ma, sizeof(double));
  s->ltype = (int *)calloc(ma, sizeof(int));
  s->moltype = (int *)calloc(ma, sizeof(int));
  s->lfreezesite = (int *)calloc(ma, sizeof(int));
  s->xdf = (double *)calloc(ma, sizeof(double));
  s->ydf = (double *)calloc(ma, sizeof(double));
  s->zdf = (double *)calloc(ma, sizeof(double));
  s->rsqdf = (double *)calloc(ma, sizeof(double));
  s->ilist = (int *)calloc(ma, sizeof(int));
  s->moldf = (int *)calloc(ma, sizeof(int));
  for (int i = 0; i <= FMC_MAX_SITES; ++i) s->gstlist[i] = (int *)calloc(ma, sizeof(int));
  size_t nsk = (size_t)(s->newld + 1) * (size_t)(s->maxmls + 2);
  s->ckcsum = (double *)calloc(nsk, sizeof(double));
  s->ckssum = (double *)calloc(nsk, sizeof(double));
  s->ckcsnew = (double *)calloc(nsk, sizeof(double));
  s->ckssnew = (double *)calloc(nsk, sizeof(double));
  s->ckcsorig = (double *)calloc(nsk, sizeof(double));
  s->ckssorig = (double *)calloc(nsk, sizeof(double));
  int K2 = s->kmax2 + 1, K3 = s->kmax3 + 1;
  size_t ns = FMC_MAX_SITES + 1;
  s->elc0 = (double *)calloc(ns, sizeof(double));
  s->els0 = (double *)calloc(ns, sizeof(double));
  s->elc1 = (double *)calloc(ns, sizeof(double));
  s->els1 = (double *)calloc(ns, sizeof(double));
  s->emc = (double *)calloc(ns * K2, sizeof(double));
  s->ems = (double *)calloc(ns * K2, sizeof(double));
  s->enc = (double *)calloc(ns * K3, sizeof(double));
  s->ens = (double *)calloc(ns * K3, sizeof(double));
  s->temp = wk->temperature;
  s->beta = 1.0 / (KBOLTZ * s->temp);
  mzt_seed(&s->rng, wk->seed[0], wk->seed[1], wk->seed[2], wk->seed[3]);
  s->rotangle = PI / 3.0;
  s->rota_rotangle = s->rotangle;
  for (int g = 1; g <= s->ntpguest; ++g) {
    s->delrdisp[g] = s->delr[g];
    s->tran_delr[g] = s->delr[g];
  }
  s->disp_ratio = s->tran_ratio = s->rota_ratio = 0.5;
  condense(s);
  double sp[5];
  orc_single_point(s, sp);
  return s;
}

void orc_destroy(orc_state *s) {
  if (!s) return;
  free(s->win);
  free(s->grid);
  free(s->numg);
  free(s->erc); free(s->vvv); free(s->lstvdw); free(s->ltpvdw); free(s->prm_eps); free(s->prm_sig);
  free(s->steadd); free(s->kv);
  for (int m = 0; m <= FMC_MAX_SPECIES; ++m) {
    free(s->fwchg[m]); free(s->fwtyp[m]); free(s->fwfrz[m]);
    free(s->molx[m]); free(s->moly[m]); free(s->molz[m]);
  }
  free(s->xxx); free(s->yyy); free(s->zzz); free(s->atmcharge); free(s->ltype); free(s->moltype);
  free(s->lfreezesite); free(s->xdf); free(s->ydf); free(s->zdf); free(s->rsqdf); free(s->ilist);
  free(s->moldf);
  for (int i = 0; i <= FMC_MAX_SITES; ++i) free(s->gstlist[i]);
  free(s->ckcsum); free(s->ckssum); free(s->ckcsnew); free(s->ckssnew); free(s->ckcsorig); free(s->ckssorig);
  free(s->elc0); free(s->els0); free(s->elc1); free(s->els1); free(s->emc); free(s->ems); free(s->enc); free(s->ens);
  free(s->wl_lng); free(s->wl_hist);
  free(s);
}

int orc_info(const orc_state *s, int *newld, int *mxegrd, int *totatm, int *maxmls) {
  if (newld) *newld = s->newld;
  if (mxegrd) *mxegrd = s->mxegrd;
  if (totatm) *totatm = s->totatm;
  if (maxmls) *maxmls = s->maxmls;
  return 0;
}

/* ----------------------------------------------------- public trials -- */
/* surfacemols after an accepted move of species mol: accept_move counts an
 * inserted molecule at the surface up and a deleted one down
 * (mc_moves.f:526,568); a displacement or jump (chain = 1) moves the count
 * by the change of the molecule's own flag, never below zero
 * (gcmc.f:1906-1912,2021-2027). */
static void surf_tally(orc_state *s, int kind, int mol, int chain) {
  if (kind == FMC_MOVE_INSERT) {
    if (s->lnewsurf) s->surfacemols[mol] += 1;
  } else if (kind == FMC_MOVE_DELETE) {
    if (s->linitsurf) s->surfacemols[mol] -= 1;
  } else if (chain) {
    if (s->linitsurf && !s->lnewsurf) {
      s->surfacemols[mol] -= 1;
      if (s->surfacemols[mol] < 0) s->surfacemols[mol] = 0;
    } else if (!s->linitsurf && s->lnewsurf) {
      s->surfacemols[mol] += 1;
    }
  }
}

int orc_trial_insert(orc_state *s, int iguest, const double *xyz, fmc_trial *out) {
  if (iguest < 1 || iguest > s->ntpguest) return FMC_ERR_ARG;
  int natms = s->numatoms[s->locguest[iguest]];
  for (int i = 1; i <= natms; ++i) {
    s->newx[i] = xyz[3 * (i - 1) + 0];
    s->newy[i] = xyz[3 * (i - 1) + 1];
    s->newz[i] = xyz[3 * (i - 1) + 2];
  }
  for (int i = 0; i < 64; ++i) s->engsicorig[i] = s->engsic[i];
  insertion(s, iguest, out);
  s->pend_kind = FMC_MOVE_INSERT;
  s->pend_iguest = iguest;
  s->pend_choice = 0;
  s->pend_valid = !out->overlap;
  return 0;
}

int orc_trial_delete(orc_state *s, int iguest, int choice, fmc_trial *out) {
  if (iguest < 1 || iguest > s->ntpguest) return FMC_ERR_ARG;
  if (choice < 1 || choice > s->nummols[s->locguest[iguest]]) return FMC_ERR_ARG;
  for (int i = 0; i < 64; ++i) s->engsicorig[i] = s->engsic[i];
  deletion(s, iguest, choice, out);
  s->pend_kind = FMC_MOVE_DELETE;
  s->pend_iguest = iguest;
  s->pend_choice = choice;
  s->pend_valid = 1;
  return 0;
}

int orc_trial_displace(orc_state *s, int iguest, int choice, const double *xyz_new, fmc_trial *out) {
  if (iguest < 1 || iguest > s->ntpguest) return FMC_ERR_ARG;
  if (choice < 1 || choice > s->nummols[s->locguest[iguest]]) return FMC_ERR_ARG;
  displace_eval(s, iguest, choice, out, 1, xyz_new, 0, 0, 0, NULL);
  s->pend_kind = FMC_MOVE_DISPLACE;
  s->pend_iguest = iguest;
  s->pend_choice = choice;
  s->pend_valid = !out->overlap;
  return 0;
}

int orc_accept(orc_state *s) {
  if (!s->pend_kind || !s->pend_valid) return FMC_ERR_STATE;
  surf_tally(s, s->pend_kind, s->locguest[s->pend_iguest], 0);
  accept_move(s, s->pend_kind, s->pend_iguest, s->pend_choice);
  s->pend_kind = 0;
  return 0;
}

int orc_reject(orc_state *s) {
  if (!s->pend_kind) return FMC_ERR_STATE;
  if (s->pend_kind == FMC_MOVE_DISPLACE) restore_sk(s, s->locguest[s->pend_iguest]);
  for (int i = 0; i < 64; ++i) s->delE[i] = 0.0;
  s->pend_kind = 0;
  return 0;
}

/* ---------------------------------------------------------- MC driver -- */
/* gcmc.f:704-1045 restricted to insert / delete / displace / translate /
 * rotate, with mc_insert (1693-1752), mc_delete (1754-1810) and mc_displace
 * (1812-1935). */
/* mc_swap, gcmc.f:2269-2452: molecule ichoice of iguest is deleted (accepted
 * provisionally) and a molecule of another kind jguest inserted at its COM
 * with the template rotated by random_rot(rotangle) (the displacement's
 * current angle, as the reference passes it), also accepted provisionally;
 * then two tests with the reference's own factors (volm in A^3,
 * gcmc.f:2400-2413).  On rejection the coordinates, S(k), per-species
 * charges and energies, elrc and nummols are restored (reject_move restores
 * engsic, elrc_mol and the type counts); sumchg is not, as in the reference.
 * The provisional insertion runs accept_move even when it overlaps (with the
 * deletion's structure factor, charge and self-energy still in place), as
 * the reference does. */
static int swap_move(orc_state *s, int iguest, fmc_trial *tr, int *choice_out, double *rande_out, int *drew_out,
                     int *accepted_out, double *rec) {
  const int maxmls = s->maxmls, imol = s->locguest[iguest], natms = s->numatoms[imol], nmols = s->nummols[imol];
  if (s->ntpguest < 2) return FMC_ERR_ARG;
  s->mv_att[FMC_MOVE_SWAP] += 1;
  double chgsum_orig[FMC_MAX_SPECIES + 1], energy_orig[FMC_MAX_SPECIES + 2], elrc_mol_orig[64];
  for (int i = 0; i <= FMC_MAX_SPECIES; ++i) chgsum_orig[i] = s->chgsum_mol[i];
  for (int i = 0; i < FMC_MAX_SPECIES + 2; ++i) energy_orig[i] = s->energy[i];
  for (int i = 0; i < 64; ++i) {
    s->engsicorig[i] = s->engsic[i];
    elrc_mol_orig[i] = s->elrc_mol[i];
  }
  const double elrc_orig = s->elrc;
  int chosen[FMC_MAX_SPECIES + 1], n = 0;
  for (int j = 1; j <= s->ntpguest; ++j)
    if (j != iguest) chosen[++n] = j;
  const int ichoice = (int)floor(mzt_next(&s->rng) * nmols + 1);
  get_guest(s, iguest, ichoice);
  double comx, comy, comz;
  com(s, natms, imol, s->newx, s->newy, s->newz, &comx, &comy, &comz);
  const int jguest = chosen[(int)floor(mzt_next(&s->rng) * (s->ntpguest - 1) + 1)];
  s->swap_jguest = jguest;
  const int jmol = s->locguest[jguest], jnatms = s->numatoms[jmol];
  /* back-ups of both species' coordinates and S(k) rows */
  const int ni = natms * s->nummols[imol], nj = jnatms * s->nummols[jmol];
  double *bx = (double *)malloc(sizeof(double) * 3 * (size_t)(ni + nj + 2));
  if (!bx) return FMC_ERR_ALLOC;
  for (int a = 1; a <= ni; ++a) {
    bx[3 * a - 3] = s->molx[imol][a];
    bx[3 * a - 2] = s->moly[imol][a];
    bx[3 * a - 1] = s->molz[imol][a];
  }
  for (int a = 1; a <= nj; ++a) {
    bx[3 * (ni + a) - 3] = s->molx[jmol][a];
    bx[3 * (ni + a) - 2] = s->moly[jmol][a];
    bx[3 * (ni + a) - 1] = s->molz[jmol][a];
  }
  const int rows[3] = {imol, jmol, maxmls + 1};
  for (int r = 0; r < 3; ++r)
    for (int ik = 1; ik <= s->newld; ++ik) {
      SK(ckcsorig, rows[r], ik) = SK(ckcsum, rows[r], ik);
      SK(ckssorig, rows[r], ik) = SK(ckssum, rows[r], ik);
    }
  /* delete molecule ichoice of iguest */
  int surf_orig[FMC_MAX_SPECIES + 1];
  for (int i = 0; i <= FMC_MAX_SPECIES; ++i) surf_orig[i] = s->surfacemols[i];
  fmc_trial td;
  deletion(s, iguest, ichoice, &td);
  surf_tally(s, FMC_MOVE_DELETE, imol, 1);
  accept_move(s, FMC_MOVE_DELETE, iguest, ichoice);
  const double estepi = -td.estep;
  /* insert a jguest molecule at its COM */
  for (int a = 1; a <= jnatms; ++a) {
    s->newx[a] = s->guestx[jguest][a] + comx;
    s->newy[a] = s->guesty[jguest][a] + comy;
    s->newz[a] = s->guestz[jguest][a] + comz;
  }
  double r1 = mzt_next(&s->rng), r2 = mzt_next(&s->rng), r3 = mzt_next(&s->rng), r4 = mzt_next(&s->rng);
  double q[4];
  quat_from(r1, r2, r3, r4, s->rotangle, q);
  rotation(s->newx, s->newy, s->newz, comx, comy, comz, jnatms, q[0], q[1], q[2], q[3]);
  for (int a = 1; a <= jnatms; ++a) {
    rec[3 * (a - 1) + 0] = s->newx[a];
    rec[3 * (a - 1) + 1] = s->newy[a];
    rec[3 * (a - 1) + 2] = s->newz[a];
  }
  fmc_trial ti;
  const int ov = insertion(s, jguest, &ti);
  surf_tally(s, FMC_MOVE_INSERT, jmol, 1);
  accept_move(s, FMC_MOVE_INSERT, jguest, 0);
  const double estepj = ov ? 0.0 : ti.estep;
  int accepted = 0, drew = 0;
  double rande = -1.0;
  if (!ov) {
    rande = mzt_next(&s->rng);
    drew = 1;
    const double jpress = s->gstfuga[jguest], ipress = s->gstfuga[iguest];
    const int jnm = s->nummols[jmol], inm = s->nummols[imol] + 1;
    const double testi = (double)inm * BOLTZ * s->temp / (s->volm * ipress) * exp(-1.0 * s->beta * estepi);
    const double testj = s->volm * jpress / ((double)jnm * BOLTZ * s->temp) * exp(-1.0 * s->beta * estepj);
    if (rande < testi) {
      rande = mzt_next(&s->rng);
      if (rande < testj) accepted = 1;
    }
  }
  if (accepted) {
    s->mv_acc[FMC_MOVE_SWAP] += 1;
    condense(s);
  } else {
    /* reject_move(swap), mc_moves.f:623-673, then mc_swap's restores */
    for (int i = 0; i < 64; ++i) {
      s->delE[i] = 0.0;
      s->engsic[i] = s->engsicorig[i];
      s->elrc_mol[i] = elrc_mol_orig[i];
    }
    for (int a = 1; a <= natms; ++a) {
      int ka = s->ltpsit[imol][a];
      s->numtyp[ka] += 1;
      s->numtyp_mol[imol][ka] += 1;
      if (s->lfzsite[imol][a] != 0) {
        s->numfrz[ka] += 1;
        s->numfrz_mol[imol][ka] += 1;
      }
    }
    for (int a = 1; a <= jnatms; ++a) {
      int ka = s->ltpsit[jmol][a];
      s->numtyp[ka] -= 1;
      s->numtyp_mol[jmol][ka] -= 1;
      if (s->lfzsite[jmol][a] != 0) {
        s->numfrz[ka] -= 1;
        s->numfrz_mol[jmol][ka] -= 1;
      }
    }
    s->nummols[imol] += 1;
    s->nummols[jmol] -= 1;
    ensure_cap(s, imol, ni + natms);
    for (int a = 1; a <= ni; ++a) {
      s->molx[imol][a] = bx[3 * a - 3];
      s->moly[imol][a] = bx[3 * a - 2];
      s->molz[imol][a] = bx[3 * a - 1];
    }
    for (int a = 1; a <= nj; ++a) {
      s->molx[jmol][a] = bx[3 * (ni + a) - 3];
      s->moly[jmol][a] = bx[3 * (ni + a) - 2];
      s->molz[jmol][a] = bx[3 * (ni + a) - 1];
    }
    for (int r = 0; r < 3; ++r)
      for (int ik = 1; ik <= s->newld; ++ik) {
        SK(ckcsum, rows[r], ik) = SK(ckcsorig, rows[r], ik);
        SK(ckssum, rows[r], ik) = SK(ckssorig, rows[r], ik);
      }
    s->chgsum_mol[imol] = chgsum_orig[imol];
    s->chgsum_mol[jmol] = chgsum_orig[jmol];
    s->surfacemols[imol] = surf_orig[imol];
    s->surfacemols[jmol] = surf_orig[jmol];
    s->elrc = elrc_orig;
    condense(s);
    for (int i = 0; i < FMC_MAX_SPECIES + 2; ++i) s->energy[i] = energy_orig[i];
  }
  free(bx);
  memset(tr, 0, sizeof *tr);
  tr->estep = estepi + estepj;
  tr->overlap = ov;
  tr->surface = s->lnewsurf | (s->linitsurf << 1);
  *choice_out = ichoice;
  *rande_out = rande;
  *drew_out = drew;
  *accepted_out = accepted;
  return 0;
}

int orc_move_counts(const orc_state *s, long long *att, long long *acc) {
  for (int k = 0; k <= FMC_NMOVES; ++k) {
    att[k] = s->mv_att[k];
    acc[k] = s->mv_acc[k];
  }
  return 0;
}

/* The production block of the main loop (gcmc.f:1050-1130) after a step:
 * count it (every step, or only accepted ones with the 'sampling'
 * directive), add N, E, E*N, N*N, E*E and the surface count of every guest
 * type to the open window, and close the window when the production count
 * reaches a multiple of nwindsteps or on the run's final step (lgchk false). */
static void stats_step(orc_state *s, int accepted, int final_step) {
  if (!s->st_accsample || accepted) {
    s->prodcount += 1;
    s->rollcount += 1;
    for (int g = 1; g <= s->ntpguest; ++g) {
      int mol = s->locguest[g];
      double dm = (double)(float)s->nummols[mol];
      double m2 = dm * dm, e = s->energy[mol], e2 = e * e, sf = (double)(float)s->surfacemols[mol];
      double *c = s->chain[g];
      c[0] += dm;
      c[1] += e;
      c[2] += e * dm;
      c[3] += m2;
      c[4] += e2;
      c[5] += sf;
    }
  }
  if ((s->st_nwind > 0 && s->prodcount % s->st_nwind == 0) || final_step) {
    if (s->nwin == s->capwin) {
      s->capwin = s->capwin * 2 + 8;
      s->win = (fmc_window *)realloc(s->win, sizeof(fmc_window) * (size_t)s->capwin);
    }
    fmc_window *w = &s->win[s->nwin++];
    memset(w, 0, sizeof *w);
    w->count = (double)s->rollcount;
    for (int g = 1; g <= s->ntpguest && g <= FMC_MAX_SPECIES; ++g)
      for (int q = 0; q < FMC_WIN_STATS; ++q) {
        w->sums[g - 1][q] = s->chain[g][q];
        s->chain[g][q] = 0.0;
      }
    s->rollcount = 0;
  }
}

/* equitable_bin, utility_pack.f:2407-2519, into grid[1..ngrida*ngridb*ngridc] */
static void ebin(int ngrida, int ngridb, int ngridc, double *grid, double fposa, double fposb, double fposc,
                 double val) {
  const int gridsize = ngrida * ngridb * ngridc;
  int alidx, aridx, blidx, bridx, clidx, cridx;
  double al, ar, bl, br, cl, cr;
  double gpta = ngrida * (fposa);
  int aidx = (int)ceil(gpta);
  double apart = gpta - floor(gpta / 1.0) * 1.0;
  if (apart > 0.5) {
    if (aidx >= ngrida) {
      alidx = (ngrida - 1) * ngridb * ngridc;
      aridx = 0;
    } else {
      alidx = (aidx - 1) * ngridb * ngridc;
      aridx = (aidx)*ngridb * ngridc;
    }
    al = 1.5 - apart;
    ar = apart - 0.5;
  } else {
    if (aidx <= 1) {
      alidx = (ngrida - 1) * ngridb * ngridc;
      aridx = 0;
    } else {
      alidx = (aidx - 2) * ngridb * ngridc;
      aridx = (aidx - 1) * ngridb * ngridc;
    }
    al = 0.5 - apart;
    ar = 0.5 + apart;
  }
  double gptb = ngridb * (fposb);
  int bidx = (int)ceil(gptb);
  double bpart = gptb - floor(gptb / 1.0) * 1.0;
  if (bpart > 0.5) {
    if (bidx >= ngridb) {
      blidx = (ngridb - 1) * ngridc;
      bridx = 1;
    } else {
      blidx = (bidx - 1) * ngridc;
      bridx = (bidx)*ngridc;
    }
    bl = 1.5 - bpart;
    br = bpart - 0.5;
  } else {
    if (bidx <= 1) {
      blidx = (ngridb - 1) * ngridc;
      bridx = 0;
    } else {
      blidx = (bidx - 2) * ngridc;
      bridx = (bidx - 1) * ngridc;
    }
    bl = 0.5 - bpart;
    br = 0.5 + bpart;
  }
  double gptc = ngridc * (fposc);
  int cidx = (int)ceil(gptc);
  double cpart = gptc - floor(gptc / 1.0) * 1.0;
  if (cpart > 0.5) {
    if (cidx >= ngridc) {
      clidx = ngridc;
      cridx = 1;
    } else {
      clidx = cidx;
      cridx = cidx + 1;
    }
    cl = 1.5 - cpart;
    cr = cpart - 0.5;
  } else {
    if (cidx <= 1) {
      clidx = ngridc;
      cridx = 1;
    } else {
      clidx = cidx - 1;
      cridx = cidx;
    }
    cl = 0.5 - cpart;
    cr = 0.5 + cpart;
  }
  const int bins[8] = {alidx + blidx + clidx, aridx + blidx + clidx, alidx + bridx + clidx, alidx + blidx + cridx,
                       aridx + bridx + clidx, aridx + blidx + cridx, alidx + bridx + cridx, aridx + bridx + cridx};
  const double wts[8] = {(al * bl * cl) * val, (ar * bl * cl) * val, (al * br * cl) * val, (al * bl * cr) * val,
                         (ar * br * cl) * val, (ar * bl * cr) * val, (al * br * cr) * val, (ar * br * cr) * val};
  for (int i = 0; i < 8; ++i)
    if (bins[i] >= 1 && bins[i] <= gridsize) grid[bins[i]] = grid[bins[i]] + wts[i];
}

static void equitable_bin(orc_state *s, double fposa, double fposb, double fposc, int subgrid, double val) {
  ebin(s->ngrida, s->ngridb, s->ngridc, s->grid + (size_t)subgrid * (size_t)(s->gridsize + 1), fposa, fposb, fposc,
       val);
}

/* frac2 / frac, utility_pack.f:873-921: fractional coordinates in the unit
 * cell, folded by modulo(x, 1.0) */
static void frac1(const orc_state *s, double *x, double *y, double *z) {
  const double *r = s->rucell;
  double ssx = *x * r[1] + *y * r[4] + *z * r[7];
  double ssy = *x * r[2] + *y * r[5] + *z * r[8];
  double ssz = *x * r[3] + *y * r[6] + *z * r[9];
  *x = ssx - floor(ssx);
  *y = ssy - floor(ssy);
  *z = ssz - floor(ssz);
}

/* storeprob, utility_pack.f:622-687: per guest type and molecule, the centre
 * of mass and the sites in unit-cell fractional coordinates, binned into
 * the type's plots (site 0: the centre of mass) */
static void storeprob(orc_state *s) {
  for (int iguest = 1; iguest <= s->ntpguest; ++iguest) {
    int mol = s->locguest[iguest], nmols = s->nummols[mol], natms = s->numatoms[mol];
    int jj = 0;
    for (int itmols = 1; itmols <= nmols; ++itmols) {
      for (int itatm = 1; itatm <= natms; ++itatm) {
        ++jj;
        s->newx[itatm] = s->molx[mol][jj];
        s->newy[itatm] = s->moly[mol][jj];
        s->newz[itatm] = s->molz[mol][jj];
      }
      double comx, comy, comz;
      com(s, natms, mol, s->newx, s->newy, s->newz, &comx, &comy, &comz);
      frac1(s, &comx, &comy, &comz);
      for (int i = 1; i <= natms; ++i) frac1(s, &s->newx[i], &s->newy[i], &s->newz[i]);
      for (int p = 0; p < s->nplots; ++p) {
        const fmc_prob_plot *pl = &s->plots[p];
        if (pl->iguest != iguest) continue;
        for (int e = 0; e < pl->nsites; ++e) {
          int jatm = pl->sites[e];
          if (jatm == 0)
            equitable_bin(s, comx, comy, comz, p, 1.0);
          else if (jatm <= natms)
            equitable_bin(s, s->newx[jatm], s->newy[jatm], s->newz[jatm], p, 1.0);
        }
      }
    }
  }
}

int orc_set_prob(orc_state *s, const int ngrid[3], const int gridfactor[3], int nplots, const fmc_prob_plot *plots) {
  if (nplots < 0 || nplots > FMC_MAX_PLOTS) return FMC_ERR_ARG;
  s->ngrida = ngrid[0];
  s->ngridb = ngrid[1];
  s->ngridc = ngrid[2];
  s->gridsize = ngrid[0] * ngrid[1] * ngrid[2];
  double ucell[10];
  for (int i = 1; i <= 9; ++i) ucell[i] = s->cell[i] / (double)gridfactor[(i - 1) / 3];
  invert(ucell, s->rucell);
  s->nplots = nplots;
  for (int p = 0; p < nplots; ++p) s->plots[p] = plots[p];
  free(s->grid);
  s->grid = (double *)calloc((size_t)(nplots > 0 ? nplots : 1) * (size_t)(s->gridsize + 1), sizeof(double));
  return s->grid ? 0 : FMC_ERR_ALLOC;
}

int orc_set_numguests(orc_state *s, int nnumg) {
  if (nnumg < 0) return FMC_ERR_ARG;
  s->nnumg = nnumg;
  s->numg_n = 0;
  return 0;
}

int orc_get_numguests(const orc_state *s, fmc_numguests_rec *out, int cap, int *n) {
  for (int i = 0; i < s->numg_n && i < cap; ++i) out[i] = s->numg[i];
  if (n) *n = s->numg_n;
  return 0;
}

int orc_get_prob(const orc_state *s, int plot, double *out) {
  if (plot < 0 || plot >= s->nplots) return FMC_ERR_ARG;
  const double *g = s->grid + (size_t)plot * (size_t)(s->gridsize + 1);
  for (int b = 1; b <= s->gridsize; ++b) out[b - 1] = g[b];
  return 0;
}

int orc_set_sampling(orc_state *s, int nwindsteps, int accepted_only) {
  if (nwindsteps < 0) return FMC_ERR_ARG;
  s->st_nwind = nwindsteps;
  s->st_accsample = accepted_only != 0;
  s->prodcount = s->rollcount = 0;
  s->nwin = 0;
  memset(s->chain, 0, sizeof s->chain);
  return 0;
}

int orc_get_windows(const orc_state *s, fmc_window *out, int cap, int *nwin, long long *prodcount) {
  for (int i = 0; i < s->nwin && i < cap; ++i) out[i] = s->win[i];
  if (nwin) *nwin = s->nwin;
  if (prodcount) *prodcount = s->prodcount;
  return 0;
}

int orc_get_surface(const orc_state *s, int *surf) {
  for (int g = 1; g <= s->ntpguest; ++g) surf[g - 1] = s->surfacemols[s->locguest[g]];
  return 0;
}

static int run_chain(orc_state *s, int nsteps, fmc_trace_rec *trace, int flags);

int orc_run(orc_state *s, int nsteps, fmc_trace_rec *trace) { return run_chain(s, nsteps, trace, 0); }

int orc_run_stats(orc_state *s, int nsteps, fmc_trace_rec *trace, int flags) {
  return run_chain(s, nsteps, trace, flags);
}

static int run_chain(orc_state *s, int nsteps, fmc_trace_rec *trace, int flags) {
  for (int step = 1; step <= nsteps; ++step) {
    int iguest = 1;
    if (s->ntpguest > 1) iguest = (int)floor(mzt_next(&s->rng) * s->ntpguest) + 1;
    int mol = s->locguest[iguest], natms = s->numatoms[mol], nmols = s->nummols[mol];
    double randmov = mzt_next(&s->rng);
    const float *f = s->mcf[iguest];
    int kind;
    if (randmov < (double)f[0]) kind = FMC_MOVE_INSERT;
    else if (randmov < (double)f[1]) kind = FMC_MOVE_DELETE;
    else if (randmov < (double)f[2]) kind = FMC_MOVE_DISPLACE;
    else if (randmov < (double)f[3]) kind = FMC_MOVE_JUMP;
    else if (randmov < (double)f[4]) kind = FMC_MOVE_SWAP;
    else if (randmov < (double)f[5]) kind = FMC_MOVE_TRANSLATE;
    else if (randmov < (double)f[6]) kind = FMC_MOVE_ROTATE;
    else if (randmov < (double)f[7]) kind = FMC_MOVE_SWITCH;
    else kind = FMC_MOVE_DISPLACE;
    /* step-size adaptation, gcmc.f:859-900 */
    if (nmols >= 1 && s->disp_count[iguest] >= 1) {
      if (s->disp_count[iguest] % 100 == 0) {
        if ((s->accept_disp[iguest] / (double)s->disp_count[iguest]) > s->disp_ratio)
          s->delrdisp[iguest] *= 1.05;
        else if (s->delrdisp[iguest] > DELRMIN)
          s->delrdisp[iguest] *= 0.95;
      } else if (s->disp_count[iguest] % 50 == 0) {
        if ((s->accept_disp[iguest] / (double)s->disp_count[iguest]) > s->disp_ratio)
          s->rotangle *= 1.05;
        else if (s->rotangle > MINANGLE)
          s->rotangle *= 0.95;
      }
    }
    if (nmols >= 1 && s->tran_count[iguest] >= 1 && s->tran_count[iguest] % 100 == 0) {
      if ((s->accept_tran[iguest] / (double)s->tran_count[iguest]) > s->tran_ratio)
        s->tran_delr[iguest] *= 1.05;
      else if (s->tran_delr[iguest] > DELRMIN)
        s->tran_delr[iguest] *= 0.95;
    }
    if (nmols >= 1 && s->rota_count[iguest] >= 1 && s->rota_count[iguest] % 100 == 0) {
      if ((s->accept_rota[iguest] / (double)s->rota_count[iguest]) > s->rota_ratio)
        s->rota_rotangle *= 1.05;
      else if (s->rota_rotangle > MINANGLE)
        s->rota_rotangle *= 0.95;
    }
    if (nmols == 0) kind = FMC_MOVE_INSERT;

    fmc_trial tr;
    memset(&tr, 0, sizeof tr);
    double rande = -1.0;
    int drew = 0, accepted = 0, choice = 0;
    double rec[3 * FMC_MAX_SITES];
    if (kind == FMC_MOVE_INSERT) {
      for (int i = 0; i < 64; ++i) s->engsicorig[i] = s->engsic[i];
      /* random_ins, utility_pack.f:3220-3331 */
      double ra = mzt_next(&s->rng), rb = mzt_next(&s->rng), rcc = mzt_next(&s->rng);
      double r1 = mzt_next(&s->rng), r2 = mzt_next(&s->rng), r3 = mzt_next(&s->rng), r4 = mzt_next(&s->rng);
      const double *cl = s->cell;
      double xc = ra * cl[1] + rb * cl[4] + rcc * cl[7];
      double yc = ra * cl[2] + rb * cl[5] + rcc * cl[8];
      double zc = ra * cl[3] + rb * cl[6] + rcc * cl[9];
      for (int i = 1; i <= natms; ++i) {
        s->newx[i] = s->guestx[iguest][i];
        s->newy[i] = s->guesty[iguest][i];
        s->newz[i] = s->guestz[iguest][i];
      }
      double q[4];
      quat_from(r1, r2, r3, r4, 2.0 * PI, q);
      rotation(s->newx, s->newy, s->newz, 0.0, 0.0, 0.0, natms, q[0], q[1], q[2], q[3]);
      for (int i = 1; i <= natms; ++i) {
        s->newx[i] += xc;
        s->newy[i] += yc;
        s->newz[i] += zc;
      }
      for (int i = 1; i <= natms; ++i) {
        rec[3 * (i - 1) + 0] = s->newx[i];
        rec[3 * (i - 1) + 1] = s->newy[i];
        rec[3 * (i - 1) + 2] = s->newz[i];
      }
      insertion(s, iguest, &tr);
      s->mv_att[FMC_MOVE_INSERT] += 1;
      if (!tr.overlap) {
        rande = mzt_next(&s->rng);
        drew = 1;
        accepted = energy_eval(s, tr.estep, rande, iguest, FMC_MOVE_INSERT);
      }
      if (accepted) {
        s->mv_acc[FMC_MOVE_INSERT] += 1;
        surf_tally(s, FMC_MOVE_INSERT, mol, 1);
        accept_move(s, FMC_MOVE_INSERT, iguest, 0);
      } else {
        for (int i = 0; i < 64; ++i) s->engsic[i] = s->engsicorig[i];
        for (int i = 0; i < 64; ++i) s->delE[i] = 0.0;
      }
    } else if (kind == FMC_MOVE_DELETE) {
      for (int i = 0; i < 64; ++i) s->engsicorig[i] = s->engsic[i];
      choice = (int)floor(mzt_next(&s->rng) * nmols) + 1;
      deletion(s, iguest, choice, &tr);
      for (int i = 1; i <= natms; ++i) {
        rec[3 * (i - 1) + 0] = s->newx[i];
        rec[3 * (i - 1) + 1] = s->newy[i];
        rec[3 * (i - 1) + 2] = s->newz[i];
      }
      rande = mzt_next(&s->rng);
      drew = 1;
      accepted = energy_eval(s, -tr.estep, rande, iguest, FMC_MOVE_DELETE);
      s->mv_att[FMC_MOVE_DELETE] += 1;
      if (accepted) {
        s->mv_acc[FMC_MOVE_DELETE] += 1;
        surf_tally(s, FMC_MOVE_DELETE, mol, 1);
        accept_move(s, FMC_MOVE_DELETE, iguest, choice);
      } else {
        for (int i = 0; i < 64; ++i) s->engsic[i] = s->engsicorig[i];
        for (int i = 0; i < 64; ++i) s->delE[i] = 0.0;
      }
    } else if (kind == FMC_MOVE_JUMP) {
      /* mc_jump, gcmc.f:1937-2039: a displacement to a random point of the
       * cell (random_jump, utility_pack.f:3359-3380) with a random rotation */
      s->mv_att[FMC_MOVE_JUMP] += 1;
      choice = (int)floor(mzt_next(&s->rng) * nmols) + 1;
      get_guest(s, iguest, choice);
      double comx, comy, comz;
      com(s, natms, mol, s->newx, s->newy, s->newz, &comx, &comy, &comz);
      double ra = mzt_next(&s->rng), rb = mzt_next(&s->rng), rcc = mzt_next(&s->rng);
      const double *cl = s->cell;
      double xc = ra * cl[1] + rb * cl[4] + rcc * cl[7];
      double yc = ra * cl[2] + rb * cl[5] + rcc * cl[8];
      double zc = ra * cl[3] + rb * cl[6] + rcc * cl[9];
      double r1 = mzt_next(&s->rng), r2 = mzt_next(&s->rng), r3 = mzt_next(&s->rng), r4 = mzt_next(&s->rng);
      double q[4];
      quat_from(r1, r2, r3, r4, 8.0 * atan(1.0), q);
      displace_eval(s, iguest, choice, &tr, 0, NULL, xc - comx, yc - comy, zc - comz, q);
      for (int i = 1; i <= natms; ++i) {
        rec[3 * (i - 1) + 0] = s->newx[i];
        rec[3 * (i - 1) + 1] = s->newy[i];
        rec[3 * (i - 1) + 2] = s->newz[i];
      }
      if (!tr.overlap) {
        if (tr.estep < 0.0) {
          accepted = 1;
        } else {
          rande = mzt_next(&s->rng);
          drew = 1;
          accepted = energy_eval(s, tr.estep, rande, iguest, FMC_MOVE_DISPLACE);
        }
      }
      if (accepted) {
        s->mv_acc[FMC_MOVE_JUMP] += 1;
        accept_move(s, FMC_MOVE_DISPLACE, iguest, choice);
        surf_tally(s, FMC_MOVE_DISPLACE, mol, 1);
      } else {
        restore_sk(s, mol);
        for (int i = 0; i < 64; ++i) s->delE[i] = 0.0;
      }
    } else if (kind == FMC_MOVE_SWAP) {
      int rc = swap_move(s, iguest, &tr, &choice, &rande, &drew, &accepted, rec);
      if (rc) return rc;
    } else if (kind == FMC_MOVE_SWITCH) {
      /* mc_switch, gcmc.f:2041-2267: needs two other guest kinds with
       * molecules present (nswitchgst >= 2, line 2098); otherwise the step
       * does nothing */
      int nsw = 0;
      for (int g = 1; g <= s->ntpguest; ++g)
        if (g != iguest && s->nummols[s->locguest[g]] > 0) ++nsw;
      if (nsw >= 2) return FMC_ERR_ARG;  /* the three-kind switch is not restated */
    } else {
      double a = 0, b = 0, c = 0, q[4] = {0, 0, 0, 0};
      if (kind == FMC_MOVE_TRANSLATE) {
        s->tran_count[iguest] += 1;
        double h = s->tran_delr[iguest] * 0.5;
        a = (2.0 * mzt_next(&s->rng) - 1.0) * h;
        b = (2.0 * mzt_next(&s->rng) - 1.0) * h;
        c = (2.0 * mzt_next(&s->rng) - 1.0) * h;
      } else if (kind == FMC_MOVE_ROTATE) {
        s->rota_count[iguest] += 1;
        double r1 = mzt_next(&s->rng), r2 = mzt_next(&s->rng), r3 = mzt_next(&s->rng), r4 = mzt_next(&s->rng);
        quat_from(r1, r2, r3, r4, s->rota_rotangle, q);
      } else {
        s->disp_count[iguest] += 1;
        double h = s->delrdisp[iguest] * 0.5;
        double da = mzt_next(&s->rng), db = mzt_next(&s->rng), dc = mzt_next(&s->rng);
        a = (2.0 * da - 1.0) * h;
        b = (2.0 * db - 1.0) * h;
        c = (2.0 * dc - 1.0) * h;
        double r1 = mzt_next(&s->rng), r2 = mzt_next(&s->rng), r3 = mzt_next(&s->rng), r4 = mzt_next(&s->rng);
        quat_from(r1, r2, r3, r4, s->rotangle, q);
      }
      choice = (int)floor(mzt_next(&s->rng) * nmols) + 1;
      displace_eval(s, iguest, choice, &tr, 0, NULL, a, b, c, q);
      for (int i = 1; i <= natms; ++i) {
        rec[3 * (i - 1) + 0] = s->newx[i];
        rec[3 * (i - 1) + 1] = s->newy[i];
        rec[3 * (i - 1) + 2] = s->newz[i];
      }
      if (!tr.overlap) {
        if (tr.estep < 0.0) {
          accepted = 1;
        } else {
          rande = mzt_next(&s->rng);
          drew = 1;
          accepted = energy_eval(s, tr.estep, rande, iguest, FMC_MOVE_DISPLACE);
        }
      }
      s->mv_att[kind] += 1;
      if (accepted) {
        s->mv_acc[kind] += 1;
        accept_move(s, FMC_MOVE_DISPLACE, iguest, choice);
        surf_tally(s, FMC_MOVE_DISPLACE, mol, 1);
        if (kind == FMC_MOVE_TRANSLATE) s->accept_tran[iguest] += 1;
        else if (kind == FMC_MOVE_ROTATE) s->accept_rota[iguest] += 1;
        else s->accept_disp[iguest] += 1;
      } else {
        restore_sk(s, mol);
        for (int i = 0; i < 64; ++i) s->delE[i] = 0.0;
      }
    }
    if (trace) {
      fmc_trace_rec *r = &trace[step - 1];
      memset(r, 0, sizeof *r);
      r->step = step;
      r->iguest = iguest;
      r->move = kind;
      r->choice = choice;
      r->nmols = nmols;
      r->overlap = tr.overlap;
      r->accepted = accepted;
      r->drew_rande = drew;
      r->estep = tr.estep;
      r->rande = rande;
      for (int i = 0; i < 3 * natms && i < 3 * FMC_MAX_SITES; ++i) r->xyz[i] = rec[i];
      for (int i = 1; i <= s->maxmls && i <= FMC_MAX_SPECIES; ++i) {
        r->energy[i - 1] = s->energy[i];
        r->delE[i - 1] = s->delE[i];
      }
      r->delrdisp = s->delrdisp[iguest];
      r->rotangle = s->rotangle;
      r->rota_rotangle = s->rota_rotangle;
      r->surface = tr.surface;
      r->nsurface = s->surfacemols[mol];
    }
    if (flags & FMC_STATS_COUNT) {
      stats_step(s, accepted, (flags & FMC_STATS_CLOSE) && step == nsteps);
      if (s->nplots > 0) storeprob(s);
    }
    /* numguests.out: nummols, energy, delE and the step's move flags
     * (ins, del, dis 1/2/3, jmp, swp, swi) per guest, gcmc.f:1138-1163 */
    s->gcmccount += 1;
    if (s->nnumg > 0 && s->gcmccount % s->nnumg == 0) {
      if (s->numg_n == s->numg_cap) {
        s->numg_cap = s->numg_cap * 2 + 16;
        s->numg = (fmc_numguests_rec *)realloc(s->numg, sizeof(fmc_numguests_rec) * (size_t)s->numg_cap);
      }
      fmc_numguests_rec *q = &s->numg[s->numg_n++];
      memset(q, 0, sizeof *q);
      q->step = s->gcmccount;
      for (int g = 1; g <= s->ntpguest && g <= FMC_MAX_SPECIES; ++g) {
        int m = s->locguest[g];
        q->nmols[g - 1] = s->nummols[m];
        q->energy[g - 1] = s->energy[m];
        q->delE[g - 1] = s->delE[m];
        if (g == iguest) {
          q->moves[g - 1][0] = kind == FMC_MOVE_INSERT;
          q->moves[g - 1][1] = kind == FMC_MOVE_DELETE;
          q->moves[g - 1][2] = kind == FMC_MOVE_DISPLACE ? 1 : kind == FMC_MOVE_TRANSLATE ? 2
                               : kind == FMC_MOVE_ROTATE ? 3 : 0;
          q->moves[g - 1][3] = kind == FMC_MOVE_JUMP;
        }
        q->moves[g - 1][4] = kind == FMC_MOVE_SWAP && (g == iguest || g == s->swap_jguest);
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------- Wang-Landau -- */
/* The flat-histogram loop of wang_landau_sim (wang_landau.f:9-411) with
 * wl_insert / wl_delete / wl_displace (590-895), eval_wl_move (540-588),
 * convergence_check (413-459), adjust_factor (527-538) and reset_visit_hist
 * (513-525); before it, insert_guests (mc_moves.f:84-149) up to the preload
 * target, one attempt per step.  Bin N of ln g / visits is dos_hist(N+1).
 * The reference stops for more than one guest type (error 2318); here guest
 * types other than 1 keep grand-canonical moves.  A displacement with no
 * molecule present (undefined in the reference) is an empty step. */
int orc_wl_setup(orc_state *s, int minmol, int maxmol, int maxn, int preload, double wlprec, double flat,
                 int visittol, double prectol) {
  if (minmol < 0 || maxmol < minmol || maxn < maxmol) return FMC_ERR_ARG;
  free(s->wl_lng);
  free(s->wl_hist);
  s->wl_nbins = maxn + 2;
  s->wl_lng = (double *)calloc((size_t)s->wl_nbins, sizeof(double));
  s->wl_hist = (int *)calloc((size_t)s->wl_nbins, sizeof(int));
  s->wl_on = 1;
  s->wl_done = 0;
  s->wl_error = 0;
  s->wl_min = minmol;
  s->wl_max = maxmol;
  s->wl_preload = preload > minmol ? preload : minmol;
  s->wl_iter = 0;
  s->wl_sweeps = 0;
  s->wl_f = wlprec;
  s->wl_logf = log(wlprec);
  s->wl_flat = flat;
  s->wl_visittol = visittol;
  s->wl_prectol = prectol;
  /* step sizes of the Wang-Landau loop, wang_landau.f:105-110 */
  s->rotangle = PI / 3.0;
  s->rota_rotangle = s->rotangle;
  for (int g = 1; g <= s->ntpguest; ++g) {
    s->delrdisp[g] = s->delr[g];
    s->tran_delr[g] = s->delr[g];
  }
  int mol = s->locguest[1];
  double molmass = 0.0;
  for (int i = 1; i <= s->numatoms[mol]; ++i) molmass += s->atmwght[mol][i];
  double lambda = HPLANCK / sqrt(2.0 * PI * molmass / 1000.0 * BOLTZ * s->temp / AVO) / 1.0e-10;
  s->wl_lambda3 = lambda * lambda * lambda;
  return 0;
}

static void wl_visit(orc_state *s, int n) {
  s->wl_hist[n] += 1;
  s->wl_lng[n] += s->wl_logf;
}

static int wl_flat(const orc_state *s) {
  double ave = 0.0, minvar = 1.0e99;
  int ncount = 0;
  for (int b = s->wl_min; b <= s->wl_max; ++b) {
    double v = (double)s->wl_hist[b];
    ++ncount;
    if (v < minvar) minvar = v;
    ave += v;
  }
  ave /= (double)ncount;
  if (minvar <= 0.5) return 0;
  return (minvar >= s->wl_flat * ave) || ((int)minvar >= s->wl_visittol);
}

static int eval_wl_move(orc_state *s, int iguest, int insert, double estep) {
  int nmol = s->nummols[s->locguest[iguest]];
  double contrib;
  if (insert) {
    double unbiased = exp(-1.0 * s->beta * estep) * s->volm / (double)(1 + nmol);
    contrib = unbiased / s->wl_lambda3 * exp(s->wl_lng[nmol] - s->wl_lng[nmol + 1]);
  } else {
    double unbiased = exp(-1.0 * s->beta * estep) * (double)nmol / s->volm;
    contrib = unbiased * s->wl_lambda3 * exp(s->wl_lng[nmol] - s->wl_lng[nmol - 1]);
  }
  double randn = mzt_next(&s->rng);
  return randn < contrib;
}

static int wl_choose(const float *f, double randmov, int skip_ins, int skip_del) {
  if (!skip_ins && randmov < (double)f[0]) return FMC_MOVE_INSERT;
  if (skip_ins && randmov < (double)f[1]) return FMC_MOVE_DELETE;
  if (!skip_ins && !skip_del && randmov < (double)f[1]) return FMC_MOVE_DELETE;
  if (randmov < (double)f[2]) return FMC_MOVE_DISPLACE;
  if (randmov < (double)f[5]) return randmov < (double)f[4] ? FMC_MOVE_DISPLACE : FMC_MOVE_TRANSLATE;
  if (randmov < (double)f[6]) return FMC_MOVE_ROTATE;
  return FMC_MOVE_DISPLACE;
}

static void random_ins_xyz(orc_state *s, int iguest, int natms) {  /* utility_pack.f:3220-3331 */
  double ra = mzt_next(&s->rng), rb = mzt_next(&s->rng), rcc = mzt_next(&s->rng);
  double r1 = mzt_next(&s->rng), r2 = mzt_next(&s->rng), r3 = mzt_next(&s->rng), r4 = mzt_next(&s->rng);
  const double *cl = s->cell;
  double xc = ra * cl[1] + rb * cl[4] + rcc * cl[7];
  double yc = ra * cl[2] + rb * cl[5] + rcc * cl[8];
  double zc = ra * cl[3] + rb * cl[6] + rcc * cl[9];
  for (int i = 1; i <= natms; ++i) {
    s->newx[i] = s->guestx[iguest][i];
    s->newy[i] = s->guesty[iguest][i];
    s->newz[i] = s->guestz[iguest][i];
  }
  double q[4];
  quat_from(r1, r2, r3, r4, 2.0 * PI, q);
  rotation(s->newx, s->newy, s->newz, 0.0, 0.0, 0.0, natms, q[0], q[1], q[2], q[3]);
  for (int i = 1; i <= natms; ++i) {
    s->newx[i] += xc;
    s->newy[i] += yc;
    s->newz[i] += zc;
  }
}

int orc_wl_run(orc_state *s, int nsteps, fmc_trace_rec *trace) {
  if (!s->wl_on) return FMC_ERR_STATE;
  const int mol1 = s->locguest[1];
  for (int step = 1; step <= nsteps; ++step) {
    if (s->wl_done) break;
    fmc_trial tr;
    memset(&tr, 0, sizeof tr);
    double rande = -1.0;
    int drew = 0, accepted = 0, choice = 0, iguest = 1, kind, preload = 0;
    if (s->nummols[mol1] < s->wl_preload) {
      preload = 1;
      kind = FMC_MOVE_INSERT;
      s->wl_iter += 1;
    } else {
      iguest = (int)floor(mzt_next(&s->rng) * s->ntpguest) + 1;
      double randmov = mzt_next(&s->rng);
      const float *f = s->mcf[iguest];
      kind = wl_choose(f, randmov, 0, 0);
      int nmol = s->nummols[s->locguest[iguest]];
      if (iguest == 1) {
        if (nmol < s->wl_min) {
          kind = FMC_MOVE_INSERT;
        } else if (nmol <= s->wl_min && kind == FMC_MOVE_DELETE) {
          wl_visit(s, nmol);
          kind = wl_choose(f, randmov, 0, 1);
        }
        if (nmol > s->wl_max) {
          kind = FMC_MOVE_DELETE;
        } else if (nmol >= s->wl_max && kind == FMC_MOVE_INSERT) {
          wl_visit(s, nmol);
          kind = wl_choose(f, randmov, 1, 0);
        }
        if (nmol == 0 && kind != FMC_MOVE_INSERT) kind = 0;
      } else if (nmol == 0) {
        kind = FMC_MOVE_INSERT;
      }
    }
    int mol = s->locguest[iguest], natms = s->numatoms[mol], nmols = s->nummols[mol];
    int wlg = !preload && iguest == 1;
    double rec[3 * FMC_MAX_SITES];
    memset(rec, 0, sizeof rec);
    if (kind == FMC_MOVE_INSERT) {
      for (int i = 0; i < 64; ++i) s->engsicorig[i] = s->engsic[i];
      random_ins_xyz(s, iguest, natms);
      for (int i = 1; i <= natms; ++i) {
        rec[3 * (i - 1) + 0] = s->newx[i];
        rec[3 * (i - 1) + 1] = s->newy[i];
        rec[3 * (i - 1) + 2] = s->newz[i];
      }
      insertion(s, iguest, &tr);
      if (preload) {
        accepted = !tr.overlap;
      } else if (!tr.overlap) {
        if (wlg) {
          accepted = eval_wl_move(s, iguest, 1, tr.estep);
        } else {
          rande = mzt_next(&s->rng);
          accepted = energy_eval(s, tr.estep, rande, iguest, FMC_MOVE_INSERT);
        }
        drew = 1;
      }
      if (wlg) wl_visit(s, accepted ? nmols + 1 : nmols);
      if (accepted) {
        accept_move(s, FMC_MOVE_INSERT, iguest, 0);
      } else {
        for (int i = 0; i < 64; ++i) s->engsic[i] = s->engsicorig[i];
        for (int i = 0; i < 64; ++i) s->delE[i] = 0.0;
      }
      if (preload && s->wl_iter >= 1000LL * s->wl_preload && s->nummols[mol1] < s->wl_preload) {
        s->wl_error = FMC_ERR_INSERT_ITER;  /* mc_moves.f:125 */
        s->wl_done = 1;
      }
    } else if (kind == FMC_MOVE_DELETE) {
      for (int i = 0; i < 64; ++i) s->engsicorig[i] = s->engsic[i];
      choice = (int)floor(mzt_next(&s->rng) * nmols) + 1;
      deletion(s, iguest, choice, &tr);
      for (int i = 1; i <= natms; ++i) {
        rec[3 * (i - 1) + 0] = s->newx[i];
        rec[3 * (i - 1) + 1] = s->newy[i];
        rec[3 * (i - 1) + 2] = s->newz[i];
      }
      if (wlg) {
        accepted = eval_wl_move(s, iguest, 0, -tr.estep);
      } else {
        rande = mzt_next(&s->rng);
        accepted = energy_eval(s, -tr.estep, rande, iguest, FMC_MOVE_DELETE);
      }
      drew = 1;
      if (wlg) wl_visit(s, accepted ? nmols - 1 : nmols);
      if (accepted) {
        accept_move(s, FMC_MOVE_DELETE, iguest, choice);
      } else {
        for (int i = 0; i < 64; ++i) s->engsic[i] = s->engsicorig[i];
        for (int i = 0; i < 64; ++i) s->delE[i] = 0.0;
      }
    } else if (kind != 0) {
      /* wl_displace: fixed delr(iguest), rotangle and rota_rotangle (no adaptation) */
      double a = 0, b = 0, c = 0, q[4] = {0, 0, 0, 0};
      double h = s->delr[iguest] * 0.5;
      if (kind == FMC_MOVE_TRANSLATE) {
        s->tran_count[iguest] += 1;
        a = (2.0 * mzt_next(&s->rng) - 1.0) * h;
        b = (2.0 * mzt_next(&s->rng) - 1.0) * h;
        c = (2.0 * mzt_next(&s->rng) - 1.0) * h;
      } else if (kind == FMC_MOVE_ROTATE) {
        s->rota_count[iguest] += 1;
        double r1 = mzt_next(&s->rng), r2 = mzt_next(&s->rng), r3 = mzt_next(&s->rng), r4 = mzt_next(&s->rng);
        quat_from(r1, r2, r3, r4, s->rota_rotangle, q);
      } else {
        s->disp_count[iguest] += 1;
        double da = mzt_next(&s->rng), db = mzt_next(&s->rng), dc = mzt_next(&s->rng);
        a = (2.0 * da - 1.0) * h;
        b = (2.0 * db - 1.0) * h;
        c = (2.0 * dc - 1.0) * h;
        double r1 = mzt_next(&s->rng), r2 = mzt_next(&s->rng), r3 = mzt_next(&s->rng), r4 = mzt_next(&s->rng);
        quat_from(r1, r2, r3, r4, s->rotangle, q);
      }
      choice = (int)floor(mzt_next(&s->rng) * nmols) + 1;
      displace_eval(s, iguest, choice, &tr, 0, NULL, a, b, c, q);
      for (int i = 1; i <= natms; ++i) {
        rec[3 * (i - 1) + 0] = s->newx[i];
        rec[3 * (i - 1) + 1] = s->newy[i];
        rec[3 * (i - 1) + 2] = s->newz[i];
      }
      if (!tr.overlap) {
        if (tr.estep < 0.0) {
          accepted = 1;
        } else {
          rande = mzt_next(&s->rng);
          drew = 1;
          accepted = energy_eval(s, tr.estep, rande, iguest, FMC_MOVE_DISPLACE);
        }
      }
      if (accepted) {
        accept_move(s, FMC_MOVE_DISPLACE, iguest, choice);
        if (kind == FMC_MOVE_TRANSLATE) s->accept_tran[iguest] += 1;
        else if (kind == FMC_MOVE_ROTATE) s->accept_rota[iguest] += 1;
        else s->accept_disp[iguest] += 1;
      } else {
        restore_sk(s, mol);
        for (int i = 0; i < 64; ++i) s->delE[i] = 0.0;
      }
    }
    if (!preload && wl_flat(s)) {
      s->wl_sweeps += 1;
      s->wl_f = sqrt(s->wl_f);
      s->wl_logf = log(s->wl_f);
      if (s->wl_logf < s->wl_prectol) {
        s->wl_done = 1;
      } else {
        for (int bb = s->wl_min; bb <= s->wl_max; ++bb) s->wl_hist[bb] = 0;
      }
    }
    if (trace) {
      fmc_trace_rec *r = &trace[step - 1];
      memset(r, 0, sizeof *r);
      r->step = step;
      r->iguest = iguest;
      r->move = kind;
      r->choice = choice;
      r->nmols = nmols;
      r->overlap = tr.overlap;
      r->accepted = accepted;
      r->drew_rande = drew;
      r->estep = tr.estep;
      r->rande = rande;
      for (int i = 0; i < 3 * natms && i < 3 * FMC_MAX_SITES; ++i) r->xyz[i] = kind ? rec[i] : 0.0;
      for (int i = 1; i <= s->maxmls && i <= FMC_MAX_SPECIES; ++i) {
        r->energy[i - 1] = s->energy[i];
        r->delE[i - 1] = accepted ? s->delE[i] : 0.0;
      }
      r->delrdisp = s->delrdisp[iguest];
      r->rotangle = s->rotangle;
      r->rota_rotangle = s->rota_rotangle;
    }
  }
  return 0;
}

/* ln g, visits [nbins] and (done, error, sweeps) ; f, ln f, lambda^3 */
int orc_wl_get(orc_state *s, double *lng, int *hist, long long *ist, double *dst) {
  if (!s->wl_on) return FMC_ERR_STATE;
  for (int b = 0; b < s->wl_nbins; ++b) {
    if (lng) lng[b] = s->wl_lng[b];
    if (hist) hist[b] = s->wl_hist[b];
  }
  if (ist) {
    ist[0] = s->wl_done;
    ist[1] = s->wl_error;
    ist[2] = s->wl_sweeps;
    ist[3] = s->wl_min;
    ist[4] = s->wl_max;
  }
  if (dst) {
    dst[0] = s->wl_f;
    dst[1] = s->wl_logf;
    dst[2] = s->wl_lambda3;
  }
  return 0;
}

/* ----------------------------------------------------------- getters -- */
/* --------------------------------------------------------- Widom grid -- */
/* widom_grid, gcmc.f:2454-2548, with the grid set-up of gcmc.f:258-281,
 * 605-625: grid spacing griddim, factors gf; points (i/wngrida, j/wngridb,
 * k/wngridc) for i = 1..ngrida etc.; at each point nstep orientations, each a
 * random_rot quaternion (4 draws, angle 2 pi) applied to the COM-centred
 * template about the origin, then moved to the point; insertion energy; sum of
 * exp(-estep / (kboltz T)) over non-overlapping attempts, averaged over all
 * attempts.  out: [0] <exp(-beta estep)>, [1] attempts, [2] non-overlapping
 * attempts, [3] Henry's constant as printed (gcmc.f:1319-1321), [4..6]
 * ngrida, ngridb, ngridc, [7..9] wngrida, wngridb, wngridc.  estep / ovl (may
 * be NULL): per attempt in loop order, estep 0 on overlap. */
int orc_widom(orc_state *s, int iguest, int nstep, double griddim, const int gf[3], double out[10], double *estep,
              int *ovl, double *emap, double *cmap) {
  if (iguest < 1 || iguest > s->ntpguest || nstep < 1 || griddim <= 0.0) return FMC_ERR_ARG;
  double celprp[11];
  dcell(s->cell, celprp);
  double gd[3];
  int ng[3], wng[3];
  for (int a = 0; a < 3; ++a) {
    gd[a] = celprp[1 + a] / ceil(celprp[1 + a] / gf[a] / griddim) / gf[a];
    ng[a] = (int)ceil(celprp[1 + a] / (gd[a] * gf[a]));
    wng[a] = (int)ceil(celprp[1 + a] / gd[a]);
  }
  const int mol = s->locguest[iguest], natms = s->numatoms[mol];
  double hsum = 0.0;
  long long widcount = 0, prodcount = 0;
  /* energy map: storeenprob (utility_pack.f:550-620) into 1-based grids */
  const int npt = ng[0] * ng[1] * ng[2];
  double *eg = NULL, *cg = NULL, ucell[10], rucell[10];
  if (emap) {
    eg = (double *)calloc((size_t)npt + 1, sizeof(double));
    cg = (double *)calloc((size_t)npt + 1, sizeof(double));
    for (int i = 1; i <= 9; ++i) ucell[i] = s->cell[i] / (double)gf[(i - 1) / 3];
    invert(ucell, rucell);
  }
  fmc_trial tr;
  const double *cl = s->cell;
  for (int i = 1; i <= ng[0]; ++i)
    for (int j = 1; j <= ng[1]; ++j)
      for (int k = 1; k <= ng[2]; ++k) {
        double afr = (double)i / (double)wng[0], bfr = (double)j / (double)wng[1], cfr = (double)k / (double)wng[2];
        /* cartesian, utility_pack.f:3694-3709 */
        double cx = afr * cl[1] + bfr * cl[4] + cfr * cl[7];
        double cy = afr * cl[2] + bfr * cl[5] + cfr * cl[8];
        double cz = afr * cl[3] + bfr * cl[6] + cfr * cl[9];
        for (int st = 0; st < nstep; ++st) {
          for (int a = 1; a <= natms; ++a) {
            s->newx[a] = s->guestx[iguest][a];
            s->newy[a] = s->guesty[iguest][a];
            s->newz[a] = s->guestz[iguest][a];
          }
          double r1 = mzt_next(&s->rng), r2 = mzt_next(&s->rng), r3 = mzt_next(&s->rng), r4 = mzt_next(&s->rng);
          double q[4];
          quat_from(r1, r2, r3, r4, 2.0 * PI, q);
          rotation(s->newx, s->newy, s->newz, 0.0, 0.0, 0.0, natms, q[0], q[1], q[2], q[3]);
          for (int a = 1; a <= natms; ++a) {
            s->newx[a] += cx;
            s->newy[a] += cy;
            s->newz[a] += cz;
          }
          int ov = insertion(s, iguest, &tr);
          for (int m = 0; m < 64; ++m) s->delE[m] = 0.0;
          if (estep) estep[widcount] = ov ? 0.0 : tr.estep;
          if (ovl) ovl[widcount] = ov;
          widcount += 1;
          if (!ov && eg) {
            double comx, comy, comz;
            com(s, natms, mol, s->newx, s->newy, s->newz, &comx, &comy, &comz);
            double sx = comx * rucell[1] + comy * rucell[4] + comz * rucell[7];
            double sy = comx * rucell[2] + comy * rucell[5] + comz * rucell[8];
            double sz = comx * rucell[3] + comy * rucell[6] + comz * rucell[9];
            sx = sx - floor(sx);
            sy = sy - floor(sy);
            sz = sz - floor(sz);
            ebin(ng[0], ng[1], ng[2], eg, sx, sy, sz, tr.estep);
            ebin(ng[0], ng[1], ng[2], cg, sx, sy, sz, 1.0);
          }
          if (!ov) {
            prodcount += 1;
            hsum += exp(-1.0 * tr.estep / KBOLTZ / s->temp);
          }
        }
      }
  if (eg) {
    for (int b = 1; b <= npt; ++b) {
      emap[b - 1] = eg[b];
      cmap[b - 1] = cg[b];
    }
    free(eg);
    free(cg);
  }
  const double avg = hsum / (double)widcount;
  out[0] = avg;
  out[1] = (double)widcount;
  out[2] = (double)prodcount;
  out[3] = avg / AVO / BOLTZ / s->temp * 1.0e5;
  for (int a = 0; a < 3; ++a) {
    out[4 + a] = ng[a];
    out[7 + a] = wng[a];
  }
  return 0;
}

int orc_get_guests(orc_state *s, int iguest, double *xyz, int cap_mols, int *nmols) {
  if (iguest < 1 || iguest > s->ntpguest) return FMC_ERR_ARG;
  int mol = s->locguest[iguest], na = s->numatoms[mol], nm = s->nummols[mol];
  *nmols = nm;
  int n = (nm < cap_mols ? nm : cap_mols) * na;
  for (int a = 1; a <= n; ++a) {
    xyz[3 * (a - 1) + 0] = s->molx[mol][a];
    xyz[3 * (a - 1) + 1] = s->moly[mol][a];
    xyz[3 * (a - 1) + 2] = s->molz[mol][a];
  }
  return 0;
}

int orc_get_sk(orc_state *s, double *sk) {
  for (int k = 1; k <= s->newld; ++k) {
    sk[2 * (k - 1) + 0] = SK(ckcsum, s->maxmls + 1, k);
    sk[2 * (k - 1) + 1] = SK(ckssum, s->maxmls + 1, k);
  }
  return 0;
}

int orc_get_energy(orc_state *s, double *energy) {
  for (int i = 1; i <= s->maxmls; ++i) energy[i - 1] = s->energy[i];
  return 0;
}

/* which: 0 erc (k ignored), 1 vvv column k, 2 steadd, 3 lstvdw (as double) */
int orc_get_table(orc_state *s, int which, int k, double *out, int n) {
  if (which == 0) {
    for (int i = 1; i <= n && i <= s->mxegrd; ++i) out[i - 1] = s->erc[i];
  } else if (which == 1) {
    if (k < 1 || k > s->ntpvdw + 1) return FMC_ERR_ARG;
    for (int i = 1; i <= n && i <= s->mxegrd; ++i) out[i - 1] = VVV(i, k);
  } else if (which == 2) {
    for (int i = 1; i <= n && i <= s->maxvdw; ++i) out[i - 1] = s->steadd[i];
  } else if (which == 3) {
    for (int i = 1; i <= n && i <= s->maxvdw; ++i) out[i - 1] = (double)s->lstvdw[i];
  } else {
    return FMC_ERR_ARG;
  }
  return 0;
}

int orc_get_kvec(orc_state *s, int *lmn, double *akk) {
  for (int k = 1; k <= s->newld; ++k) {
    lmn[3 * (k - 1) + 0] = s->kv[k].ll;
    lmn[3 * (k - 1) + 1] = s->kv[k].mm;
    lmn[3 * (k - 1) + 2] = s->kv[k].nn;
    akk[k - 1] = s->kv[k].akk;
  }
  return 0;
}

/* FP64 re-sum of S(k) from the current coordinates; reports the largest
 * component difference against the incrementally updated totals. */
int orc_full_resum(orc_state *s, double *max_drift) {
  size_t nsk = (size_t)(s->newld + 1) * (size_t)(s->maxmls + 2);
  double *oc = (double *)malloc(nsk * sizeof(double)), *os = (double *)malloc(nsk * sizeof(double));
  memcpy(oc, s->ckcsum, nsk * sizeof(double));
  memcpy(os, s->ckssum, nsk * sizeof(double));
  double saved_e1[64];
  memcpy(saved_e1, s->ewald1en, sizeof saved_e1);
  double e;
  condense(s);
  ewald1_full(s, &e);
  double md = 0.0;
  for (int k = 1; k <= s->newld; ++k) {
    double d1 = fabs(oc[(size_t)k * (s->maxmls + 2) + s->maxmls + 1] - SK(ckcsum, s->maxmls + 1, k));
    double d2 = fabs(os[(size_t)k * (s->maxmls + 2) + s->maxmls + 1] - SK(ckssum, s->maxmls + 1, k));
    if (d1 > md) md = d1;
    if (d2 > md) md = d2;
  }
  memcpy(s->ewald1en, saved_e1, sizeof saved_e1);
  free(oc);
  free(os);
  *max_drift = md;
  return 0;
}
