// Widom insertions on a grid (widom_grid, gcmc.f:2454-2548): reject-only
// insertion energies of a rigid guest at every grid point in random
// orientations, on the current state of one walker (framework, guests and
// S(k) unchanged).  Throughput kernels: one wave per attempt, many attempts
// in flight per CU; the grid of attempts needs no ordering, only the
// orientations follow the reference's duni sequence, which k_draws writes
// out first.
#pragma once
#include "fmc_kernels.hpp"

namespace fmc {

struct WidomArgs {
  int w, g, nstep;
  int ng[3], wng[3];      // ngrida.. and wngrida.. of gcmc.f:258-281,605-610
  long long a0, count;    // attempts [a0, a0 + count) of this launch
  const float *uni;       // 4 draws per attempt, attempt a0 first
  double *estep;          // [count] or null: estep of each attempt (0 on overlap)
  int *overlap;           // [count] or null
  double *bsum;           // [gridDim] sum of exp(-estep / (kboltz T)) over non-overlapping attempts
  long long *bcnt;        // [gridDim] non-overlapping attempts
  double kboltz;         // the reference's exp(-1 * estep / kboltz / temp), temp the walker's
  // energy map (storeenprob, utility_pack.f:550-620): estep and 1 of every
  // non-overlapping attempt binned at the molecule's centre of mass, in the
  // unit cell's fractional coordinates, on the ng grid; null: no map
  double *emap, *cmap;
  double rucell[9];
};

// n consecutive duni draws of walker w into out, advancing the walker's
// generator exactly as n serial calls would (the ring first, then rounds of
// 33 draws made at once).  One wave.
__global__ void __launch_bounds__(64) k_draws(DevWalkers W, int w, long long n, float *out) {
  __shared__ WState ws;
  const int lane = threadIdx.x;
  {
    const int *src = (const int *)(W.ws + w);
    int *dst = (int *)&ws;
    for (int i = lane; i < (int)(sizeof(WState) / 4); i += 64) dst[i] = src[i];
  }
  __syncthreads();
  long long done = 0;
  while (done < n) {
    if (ws.mz_made - ws.mz_used < 33) {
      mzt_refill(ws, lane);
      __syncthreads();
      if (lane == 0) mzt_advance(ws);
      __syncthreads();
    }
    const long long avail = ws.mz_made - ws.mz_used;
    const int take = (int)(n - done < avail ? n - done : avail);
    if (lane < take) out[done + lane] = ws.rq[(int)((ws.mz_used + lane) & 127)];
    __syncthreads();
    if (lane == 0) ws.mz_used += take;
    done += take;
    __syncthreads();
  }
  {
    int *dst = (int *)(W.ws + w);
    const int *src = (const int *)&ws;
    for (int i = lane; i < (int)(sizeof(WState) / 4); i += 64) dst[i] = src[i];
  }
}

// LDS layout of k_widom (byte offsets, 16-byte aligned): erfc table, site
// coordinates and phase tables of each wave, LJ columns of the guest's sites,
// the launch constants.  Shared by the kernel and the host's launch.
constexpr int WIDOM_WPB = 4;  // waves per block
struct WidomLds {
  size_t erc, xyz, ph, col, cst, total;
};
__host__ __device__ inline size_t widom_align(size_t x) { return (x + 15) & ~(size_t)15; }
__host__ __device__ inline WidomLds widom_lds(const DevParams &P) {
  WidomLds L;
  L.erc = 0;
  L.xyz = widom_align(L.erc + sizeof(double) * (size_t)(P.mxegrd + 3));
  L.ph = widom_align(L.xyz + sizeof(double) * WIDOM_WPB * KSITE * 3);
  L.col = widom_align(L.ph + sizeof(double2) * (size_t)WIDOM_WPB * P.nsmax * 3 * P.kcp);
  L.cst = widom_align(L.col + sizeof(int) * KSITE * (KTYPE + 1));
  L.total = widom_align(L.cst + 64);
  return L;
}

// Per-launch constants of the insertion energy (the trial-independent part
// of prepare and decide for an insertion of guest g on the walker's state).
struct WidomConst {
  double engsic0, qfix0, delrc, ewld3, C, reu;
};

// One attempt per wave.  NSITE > 0: the guest has at most NSITE sites (kept
// in registers); 0: up to KSITE.  T: precision of the phase products of
// k >= nk64 (as in the chains); geometry, real-space and LJ terms are FP64.
template <typename T, int NSITE>
__global__ void __launch_bounds__(256) k_widom(const DevParams *__restrict__ Pp, DevWalkers W, WidomArgs A) {
  extern __shared__ __align__(16) unsigned char lds[];
  const DevParams &P = *Pp;
  constexpr int NSA = NSITE > 0 ? NSITE : KSITE;
  constexpr int WPB = WIDOM_WPB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = A.g, mol = P.locguest[g], natms = P.natoms[mol];
  const WState &ws = W.ws[A.w];
  const WidomLds L = widom_lds(P);
  double *erc = (double *)(lds + L.erc);
  double *sxyz = (double *)(lds + L.xyz);   // [WPB][KSITE][3]
  double2 *ph = (double2 *)(lds + L.ph);    // [WPB][nsmax][3][kcp]
  int *col = (int *)(lds + L.col);          // [KSITE][KTYPE + 1]
  WidomConst *K = (WidomConst *)(lds + L.cst);
  for (int i = tid; i < P.mxegrd + 3; i += blockDim.x) erc[i] = P.erc[i];
  for (int i = tid; i < KSITE * (KTYPE + 1); i += blockDim.x) {
    const int s = i / (KTYPE + 1), tj = i % (KTYPE + 1);
    col[i] = s < natms ? (int)P.pairidx[P.site_t[mol][s]][tj] : -1;
  }
  if (tid == 0) {
    // prepare (qfix0), gstlrcorrect with lexisting = 0, decide's insertion constants
    const double chg = P.chgtmp[g], alpha = P.alpha;
    WidomConst c;
    c.qfix0 = -(0.5 * PI * R4PIE0 / P.epsq) * ((chg * ws.sumchg) / alpha + (chg / alpha) * (chg / alpha)) / P.volm;
    double delrc_mol[KMAXC + 1];
    c.delrc = gstlrcorrect<KMAXS>(P, ws, g, 0, delrc_mol);
    c.engsic0 = P.dc_esic[g];
    c.ewld3 = P.ewald3en[mol];
    c.C = P.dc_C;
    c.reu = P.rengunit;
    *K = c;
  }
  __syncthreads();
  const WidomConst c = *K;
  // LJ from the FP64 tables in both modes: a grid visits every close contact,
  // where FP32 table entries alone move estep by ~1e-6 kcal/mol
  const LJTab vvv = lj_global(P, false);
  const double *fx = P.fw_per_walker ? W.fwx + (size_t)A.w * P.nfw : P.fwx;
  const double *fy = P.fw_per_walker ? W.fwy + (size_t)A.w * P.nfw : P.fwy;
  const double *fz = P.fw_per_walker ? W.fwz + (size_t)A.w * P.nfw : P.fwz;
  const double2 *Stot = W.sk + ((size_t)A.w * (P.maxmls + 1) + P.maxmls) * P.nk;
  const double *gxw = W.gx + (size_t)A.w * P.gatoms_total;
  const double *gyw = W.gy + (size_t)A.w * P.gatoms_total;
  const double *gzw = W.gz + (size_t)A.w * P.gatoms_total;
  // guests present: kind, first slot and atom count of each kind (uniform)
  int gn[KMAXS + 1], gb[KMAXS + 1], gm[KMAXS + 1];
  int ntot = P.nfw;
  for (int gg = 1; gg <= KMAXS; ++gg) {
    const bool real = gg <= P.ntpguest;
    const int m = real ? P.locguest[gg] : 1;
    gm[gg] = m;
    gb[gg] = real ? P.gbase[gg] : 0;
    gn[gg] = real ? ws.nmols[m] * P.natoms[m] : 0;
    ntot += gn[gg];
  }
  double *my_xyz = sxyz + wave * KSITE * 3;
  double2 *my_ph = ph + (size_t)wave * P.nsmax * 3 * P.kcp;
  double hsum = 0.0;
  long long hcnt = 0;
  const long long stride = (long long)gridDim.x * WPB;
  for (long long a = (long long)blockIdx.x * WPB + wave; a < A.count; a += stride) {
    const long long ga = A.a0 + a;
    // grid point of attempt ga: loops i (outer), j, k, then the orientations
    const long long pnt = ga / A.nstep;
    const int k3 = (int)(pnt % A.ng[2]) + 1;
    const int j3 = (int)((pnt / A.ng[2]) % A.ng[1]) + 1;
    const int i3 = (int)(pnt / ((long long)A.ng[2] * A.ng[1])) + 1;
    const double afr = (double)i3 / (double)A.wng[0], bfr = (double)j3 / (double)A.wng[1],
                 cfr = (double)k3 / (double)A.wng[2];
    const double *cl = P.cell;
    const double cx = afr * cl[0] + bfr * cl[3] + cfr * cl[6];
    const double cy = afr * cl[1] + bfr * cl[4] + cfr * cl[7];
    const double cz = afr * cl[2] + bfr * cl[5] + cfr * cl[8];
    // random_rot (utility_pack.f:3526-3563) and rotation about the origin
    const float *u = A.uni + 4 * a;
    double q[4], Q[9];
    quat_from((double)u[0], (double)u[1], (double)u[2], (double)u[3], 2.0 * PI, q);
    rot_matrix(q, Q);
    if (lane < natms) {
      const double tx = P.tmpl[g][lane][0], ty = P.tmpl[g][lane][1], tz = P.tmpl[g][lane][2];
      my_xyz[lane * 3 + 0] = (Q[0] * tx + Q[1] * ty + Q[2] * tz) + cx;
      my_xyz[lane * 3 + 1] = (Q[3] * tx + Q[4] * ty + Q[5] * tz) + cy;
      my_xyz[lane * 3 + 2] = (Q[6] * tx + Q[7] * ty + Q[8] * tz) + cz;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // per-site 1-D phase tables by the reference recursion (ewald.f:108-147)
    if (lane < 3 * natms) {
      const int s = lane / 3, ax = lane - 3 * s;
      const double *x = my_xyz + s * 3;
      const double fr = P.rcell[ax] * x[0] + P.rcell[3 + ax] * x[1] + P.rcell[6 + ax] * x[2];
      const int kmax = ax == 0 ? P.kmax1 : (ax == 1 ? P.kmax2 : P.kmax3);
      double s1, c1;
      sincospi(2.0 * fr, &s1, &c1);
      double2 *row = my_ph + (size_t)(s * 3 + ax) * P.kcp;
      row[0] = make_double2(1.0, 0.0);
      row[1] = make_double2(c1, s1);
      double cc = c1, sn = s1;
      for (int l = 2; l <= kmax; ++l) {
        const double cn = cc * c1 - sn * s1;
        const double snn = sn * c1 + cc * s1;
        cc = cn;
        sn = snn;
        row[l] = make_double2(cc, sn);
      }
    }
    double xs[NSA][3], cq[NSA], ov2[NSA], qs[NSA];
#pragma unroll
    for (int s = 0; s < NSA; ++s) {
      const bool real = s < natms;
      cq[s] = real ? P.site_chgea[mol][s] : 0.0;
      ov2[s] = real ? P.overlap2 : -1.0;
      qs[s] = real ? P.site_q[mol][s] : 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) xs[s][d] = real ? my_xyz[s * 3 + d] : 0.0;
    }
    // pair terms: framework atoms, then every guest present (guestlistgen)
    double er = 0.0, ev = 0.0;
    int ov = 0;
    for (int j = lane; j < ntot; j += 64) {
      double xj, yj, zj, qj;
      int tj;
      if (j < P.nfw) {
        xj = fx[j];
        yj = fy[j];
        zj = fz[j];
        qj = P.fwq[j];
        tj = P.fwinfo[j] & 255;
      } else {
        // the guest kind holding index j, without lane-indexed private arrays
        int jj = j - P.nfw, m = gm[1], bb = gb[1];
        bool found = jj < gn[1];
#pragma unroll
        for (int gg = 2; gg <= KMAXS; ++gg) {
          const bool here = !found && jj - gn[gg - 1] < gn[gg];
          jj -= found ? 0 : gn[gg - 1];
          m = here ? gm[gg] : m;
          bb = here ? gb[gg] : bb;
          found = found || here;
        }
        const int na = P.natoms[m], site = jj % na, at = bb + jj;
        xj = gxw[at];
        yj = gyw[at];
        zj = gzw[at];
        qj = P.site_q[m][site];
        tj = P.site_t[m][site];
      }
#pragma unroll
      for (int s = 0; s < NSA; ++s) {
        double e2, e6;
        pair_eval<double>(P, erc, vvv, xs[s], cq[s], col[s * (KTYPE + 1) + tj], ov2[s], -1.0, xj, yj, zj, qj, e2, e6,
                          ov);
        er += e2;
        ev += e6;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // reciprocal terms of the insertion against the walker's total S(k)
    double rec = 0.0;
    for (int k = lane; k < P.nk; k += 64) {
      const int lmn = P.klmn[k];
      const int ll = lmn & 255, mm = ((lmn >> 8) & 255) - 128, nn = ((lmn >> 16) & 255) - 128;
      double snc, sns;
      if (k < P.nk64) {
        double c0 = 0.0, s0 = 0.0;
#pragma unroll
        for (int s = 0; s < NSA; ++s) {
          if (s >= natms) break;
          double ckc, cks;
          site_phase<double>(my_ph + (size_t)s * 3 * P.kcp, P.kcp, ll, mm, nn, qs[s], ckc, cks);
          c0 += ckc;
          s0 += cks;
        }
        snc = c0;
        sns = s0;
      } else {
        T c0 = (T)0, s0 = (T)0;
#pragma unroll
        for (int s = 0; s < NSA; ++s) {
          if (s >= natms) break;
          T ckc, cks;
          site_phase<T>(my_ph + (size_t)s * 3 * P.kcp, P.kcp, ll, mm, nn, (T)qs[s], ckc, cks);
          c0 += ckc;
          s0 += cks;
        }
        snc = (double)c0;
        sns = (double)s0;
      }
      const double2 St = Stot[k];
      const double mix = 2.0 * snc * St.x + 2.0 * sns * St.y;
      rec += P.akk[k] * (mix + (snc * snc + sns * sns));
    }
    const unsigned long long bov = __ballot(ov != 0);
    er = wave_sum_dpp(er);
    ev = wave_sum_dpp(ev);
    rec = wave_sum_dpp(rec);
    if (lane == 63) {
      const int overlap = bov != 0ULL;
      double estep = 0.0;
      if (!overlap) {
        // decide's insertion branch (mc_moves.f:251-252)
        const double ewld1eng = c.C * rec + c.engsic0 + c.qfix0;
        estep = (ewld1eng + er - c.ewld3 + ev + c.delrc) * c.reu;
        hsum += exp(-1.0 * estep / A.kboltz / ws.temp);
        hcnt += 1;
        if (A.emap) {
          double wsum = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
          for (int s = 0; s < natms; ++s) {
            const double m = P.site_m[mol][s];
            wsum += m;
            mx += my_xyz[s * 3 + 0] * m;
            my += my_xyz[s * 3 + 1] * m;
            mz += my_xyz[s * 3 + 2] * m;
          }
          mx /= wsum;
          my /= wsum;
          mz /= wsum;
          const double *r = A.rucell;
          const double sx = mx * r[0] + my * r[3] + mz * r[6], sy = mx * r[1] + my * r[4] + mz * r[7],
                       sz = mx * r[2] + my * r[5] + mz * r[8];
          const double fa = sx - floor(sx), fb = sy - floor(sy), fc = sz - floor(sz);
          equitable_bin(A.ng[0], A.ng[1], A.ng[2], A.emap, fa, fb, fc, estep);
          equitable_bin(A.ng[0], A.ng[1], A.ng[2], A.cmap, fa, fb, fc, 1.0);
        }
      }
      if (A.estep) A.estep[a] = estep;
      if (A.overlap) A.overlap[a] = overlap;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  // block partials in wave order
  __shared__ double wsum[WPB];
  __shared__ long long wcnt[WPB];
  if (lane == 63) {
    wsum[wave] = hsum;
    wcnt[wave] = hcnt;
  }
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    long long n = 0;
    for (int i = 0; i < WPB; ++i) {
      s += wsum[i];
      n += wcnt[i];
    }
    A.bsum[blockIdx.x] = s;
    A.bcnt[blockIdx.x] = n;
  }
}

}  // namespace fmc
