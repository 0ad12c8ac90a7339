// Energy of a batch of guest configurations and its gradient with respect
// to the framework parameters theta: the framework site charges and the LJ
// epsilon / sigma of the atom types (Lorentz-Berthelot mixing for every pair
// with a framework type).  The energy is the single_point total
// (gcmc.f:2549-2661: ewald1 + ewald2 - ewald3 + srfrce + tail) of each
// configuration; its derivatives are taken term by term on the same tables:
//   reciprocal  d/dq_i  C sum_k a_k |S_k|^2 = 2 C sum_k a_k Re(e^{i k.r_i} S_k*)
//   real space  d/dq_i  sum_j q_i q_j r4pie0/eps erc(r_ij) = sum_j q_j r4pie0/eps erc(r_ij)
//   LJ          d/d(eps, sigma) of the interpolated table: the same 3-point
//               interpolation of the tables of dV/deps and dV/dsigma
// (the self term, the charged-system term and the tail are closed forms done
// on the host).  The reciprocal contraction over (framework atom, k) runs in
// T (float: FP32 products with FP64 accumulation; double: all FP64); the pair
// terms are FP64.
#pragma once

#include "fmc_kernels.hpp"

namespace fmc {

struct GradArgs {
  int nconf, nk, nfw, gstride, ntpguest;
  const int *nmols;          // [nconf][ntpguest]
  const double *gx;          // [nconf][gstride][3] guest sites, guest types in order, molecules in order
  const double *fwq;         // [nfw] theta charges
  const double *fx, *fy, *fz;
  const int *fwt;            // [nfw] framework atom type
  float2 *efw32;             // [nk][nfw] e^{i k.r} of the framework (FP32)
  double2 *efw64;            // [nk][nfw] (FP64)
  double2 *sfw;              // [nk] framework structure factor
  double2 *sc;               // [nconf][nk] total structure factor of each configuration
  double *urec;              // [nconf] sum_k a_k |S_k|^2
  double *dq;                // [nconf][nfw] d/dq_i (guest recip + real), before the host's closed forms
  double *dqfw;              // [nfw] d/dq_i of the framework-framework reciprocal part
  double *e2, *ev;           // [nconf] real-space and LJ pair sums
  double *dpe, *dps;         // [nconf][KTYPE+1][KTYPE+1] dV/deps_ij, dV/dsigma_ij summed over pairs
  const double *vt;          // [npt][vstride] LJ tables of the theta pairs: V, dV/deps, dV/dsigma
  const int *ptab;           // [KTYPE+1][KTYPE+1] -> theta pair table, -2 theta pair without LJ, -1 FIELD row
  int vstride, npt;
};

// e^{i k.r_j} of the framework atoms, k-major (one thread per (k, j))
__global__ void k_grad_phases(const DevParams *__restrict__ Pp, GradArgs A) {
  const DevParams &P = *Pp;
  const int j = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
  if (j >= A.nfw) return;
  int lmn = P.klmn[k];
  double l = (double)(lmn & 255), m = (double)(((lmn >> 8) & 255) - 128), n = (double)(((lmn >> 16) & 255) - 128);
  const double *r = P.rcell;
  double x = A.fx[j], y = A.fy[j], z = A.fz[j];
  double sx = r[0] * x + r[3] * y + r[6] * z, sy = r[1] * x + r[4] * y + r[7] * z, sz = r[2] * x + r[5] * y + r[8] * z;
  double sn, cs;
  sincospi(2.0 * (l * sx + m * sy + n * sz), &sn, &cs);
  A.efw64[(size_t)k * A.nfw + j] = make_double2(cs, sn);
  A.efw32[(size_t)k * A.nfw + j] = make_float2((float)cs, (float)sn);
}

// framework S(k) for the theta charges: one block of 256 threads per k
__global__ void __launch_bounds__(256) k_grad_sfw(GradArgs A) {
  const int k = blockIdx.x;
  double re = 0.0, im = 0.0;
  for (int j = threadIdx.x; j < A.nfw; j += blockDim.x) {
    double2 e = A.efw64[(size_t)k * A.nfw + j];
    re += A.fwq[j] * e.x;
    im += A.fwq[j] * e.y;
  }
  re = wave_sum(re);
  im = wave_sum(im);
  __shared__ double red[4][2];
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6][0] = re;
    red[threadIdx.x >> 6][1] = im;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      a += red[w][0];
      b += red[w][1];
    }
    A.sfw[k] = make_double2(a, b);
  }
}

// total S(k) of each configuration and sum_k a_k |S_k|^2: thread per k,
// block row per configuration
__global__ void __launch_bounds__(256) k_grad_sconf(const DevParams *__restrict__ Pp, GradArgs A) {
  const DevParams &P = *Pp;
  const int c = blockIdx.y, k = blockIdx.x * blockDim.x + threadIdx.x;
  double u = 0.0;
  if (k < A.nk) {
    int lmn = P.klmn[k];
    double l = (double)(lmn & 255), m = (double)(((lmn >> 8) & 255) - 128), n = (double)(((lmn >> 16) & 255) - 128);
    double2 s = A.sfw[k];
    const double *r = P.rcell;
    const double *g = A.gx + (size_t)c * A.gstride * 3;
    int a = 0;
    for (int gg = 1; gg <= A.ntpguest; ++gg) {
      int mol = P.locguest[gg], na = P.natoms[mol], nat = A.nmols[c * A.ntpguest + gg - 1] * na;
      for (int j = 0; j < nat; ++j, ++a) {
        double x = g[3 * a], y = g[3 * a + 1], z = g[3 * a + 2];
        double sx = r[0] * x + r[3] * y + r[6] * z, sy = r[1] * x + r[4] * y + r[7] * z,
               sz = r[2] * x + r[5] * y + r[8] * z;
        double sn, cs;
        sincospi(2.0 * (l * sx + m * sy + n * sz), &sn, &cs);
        double q = P.site_q[mol][j % na];
        s.x += q * cs;
        s.y += q * sn;
      }
    }
    A.sc[(size_t)c * A.nk + k] = s;
    u = P.akk[k] * (s.x * s.x + s.y * s.y);
  }
  u = wave_sum(u);
  if ((threadIdx.x & 63) == 0) atomicAdd(A.urec + c, u);
}

// d/dq_i of C sum_k a_k |S_k|^2 = 2C sum_k a_k Re(e_ik S_k*), split as
// S = S_framework + S_guests: fw = 1 computes the framework part once (into
// dqfw, FP64 by the caller's choice of T); fw = 0 the part of each
// configuration's guests (block row c), whose FP32 rounding is then relative
// to the guest term only.  One thread per framework atom; a_k S_k staged in
// LDS in chunks.
template <typename T>
__global__ void __launch_bounds__(256) k_grad_dq_recip(const DevParams *__restrict__ Pp, GradArgs A, int fw) {
  const DevParams &P = *Pp;
  const int c = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ double2 w[512];
  double acc = 0.0;
  for (int k0 = 0; k0 < A.nk; k0 += 512) {
    const int nkc = min(512, A.nk - k0);
    __syncthreads();
    for (int t = threadIdx.x; t < nkc; t += blockDim.x) {
      double2 sf = A.sfw[k0 + t];
      double2 s = sf;
      if (!fw) {
        double2 st = A.sc[(size_t)c * A.nk + k0 + t];
        s = make_double2(st.x - sf.x, st.y - sf.y);
      }
      double ak = P.akk[k0 + t];
      w[t] = make_double2(ak * s.x, ak * s.y);
    }
    __syncthreads();
    if (i < A.nfw) {
      if constexpr (sizeof(T) == 4) {
        const float2 *e = A.efw32 + (size_t)k0 * A.nfw + i;
        for (int t = 0; t < nkc; ++t) {
          float2 ek = e[(size_t)t * A.nfw];
          double2 wk = w[t];
          float p = ek.x * (float)wk.x + ek.y * (float)wk.y;
          acc += (double)p;
        }
      } else {
        const double2 *e = A.efw64 + (size_t)k0 * A.nfw + i;
        for (int t = 0; t < nkc; ++t) {
          double2 ek = e[(size_t)t * A.nfw];
          double2 wk = w[t];
          acc += ek.x * wk.x + ek.y * wk.y;
        }
      }
    }
  }
  if (i < A.nfw) {
    if (fw) A.dqfw[i] = 2.0 * P.dc_C * acc;
    else A.dq[(size_t)c * A.nfw + i] = 2.0 * P.dc_C * acc;
  }
}

// interpolated table value (ewald.f:703-715 / vdw_terms.f:441-453)
__device__ inline double interp_tab(const double *tab, double r, double rstep) {
  return interp3<double>(tab, r, rstep);
}

// Real-space and LJ pair terms (FP64) of every framework atom with the guest
// sites of configuration c (blockIdx.y), and of guest pairs in different
// molecules: energies, d/dq_i (real space), dV/deps_ij and dV/dsigma_ij per
// type pair.  Framework-framework pairs are frozen-frozen and skipped, as in
// parlst (nlist_builders.f:120).
__global__ void __launch_bounds__(256) k_grad_pairs(const DevParams *__restrict__ Pp, GradArgs A) {
  const DevParams &P = *Pp;
  const int c = blockIdx.y;
  const int tid = threadIdx.x;
  __shared__ double sdpe[(KTYPE + 1) * (KTYPE + 1)], sdps[(KTYPE + 1) * (KTYPE + 1)];
  __shared__ double red[2][4];
  for (int t = tid; t < (KTYPE + 1) * (KTYPE + 1); t += blockDim.x) sdpe[t] = sdps[t] = 0.0;
  __syncthreads();
  const double *g = A.gx + (size_t)c * A.gstride * 3;
  int ng = 0;
  for (int gg = 1; gg <= A.ntpguest; ++gg) ng += A.nmols[c * A.ntpguest + gg - 1] * P.natoms[P.locguest[gg]];
  const double kq = R4PIE0 / P.epsq;
  double e2 = 0.0, ev = 0.0;
  // atom index space: framework atoms, then guest atoms (for guest-guest pairs)
  for (int a = blockIdx.x * blockDim.x + tid; a < A.nfw + ng; a += gridDim.x * blockDim.x) {
    double xa, ya, za, qa;
    int ta, mola = -1, ia = -1;
    if (a < A.nfw) {
      xa = A.fx[a];
      ya = A.fy[a];
      za = A.fz[a];
      qa = A.fwq[a];
      ta = A.fwt[a];
    } else {
      int j = a - A.nfw, gg = 1, base = 0;
      while (true) {
        int nat = A.nmols[c * A.ntpguest + gg - 1] * P.natoms[P.locguest[gg]];
        if (j < base + nat || gg == A.ntpguest) break;
        base += nat;
        ++gg;
      }
      int mol = P.locguest[gg], na = P.natoms[mol], jj = j - base;
      xa = g[3 * j];
      ya = g[3 * j + 1];
      za = g[3 * j + 2];
      qa = P.site_q[mol][jj % na];
      ta = P.site_t[mol][jj % na];
      mola = gg * 1000000 + jj / na;
      ia = j;
    }
    double dqa = 0.0;
    int b = 0;
    for (int gg = 1; gg <= A.ntpguest; ++gg) {
      int mol = P.locguest[gg], na = P.natoms[mol], nat = A.nmols[c * A.ntpguest + gg - 1] * na;
      for (int jj = 0; jj < nat; ++jj, ++b) {
        if (ia >= 0 && (b <= ia || gg * 1000000 + jj / na == mola)) continue;  // each guest pair once, no intramolecular
        double dx = xa - g[3 * b], dy = ya - g[3 * b + 1], dz = za - g[3 * b + 2];
        min_image(P, dx, dy, dz);
        double rsq = dx * dx + dy * dy + dz * dz;
        if (rsq >= P.rcut2) continue;
        double r = sqrt(rsq);
        double qb = P.site_q[mol][jj % na];
        int tb = P.site_t[mol][jj % na];
        double chgprd = qa / P.epsq * R4PIE0 * qb;
        if (fabs(chgprd) > 1e-10) {
          double er = interp_tab(P.erc, r, P.rdrewd);
          e2 += er * chgprd;
          if (ia < 0) dqa += kq * qb * er;
        }
        int pt = A.ptab[ta * (KTYPE + 1) + tb];
        if (pt >= 0) {
          const double *tab = A.vt + (size_t)pt * 3 * A.vstride;
          ev += interp_tab(tab, r, P.rdlrpot);
          if (ia < 0) {
            int lo = min(ta, tb), hi = max(ta, tb);
            atomicAdd(&sdpe[hi * (KTYPE + 1) + lo], interp_tab(tab + A.vstride, r, P.rdlrpot));
            atomicAdd(&sdps[hi * (KTYPE + 1) + lo], interp_tab(tab + 2 * A.vstride, r, P.rdlrpot));
          }
        } else if (pt == -1) {  // no framework type: the FIELD row
          int k = P.pairidx[ta][tb];
          if (k >= 0) ev += interp_tab(P.vvv + (size_t)k * P.vstride, r, P.rdlrpot);
        }
      }
    }
    if (ia < 0) A.dq[(size_t)c * A.nfw + a] += dqa;
  }
  e2 = wave_sum(e2);
  ev = wave_sum(ev);
  if ((tid & 63) == 0) {
    red[0][tid >> 6] = e2;
    red[1][tid >> 6] = ev;
  }
  __syncthreads();
  if (tid == 0) {
    double a = 0.0, b = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      a += red[0][w];
      b += red[1][w];
    }
    atomicAdd(A.e2 + c, a);
    atomicAdd(A.ev + c, b);
  }
  for (int t = tid; t < (KTYPE + 1) * (KTYPE + 1); t += blockDim.x) {
    if (sdpe[t] != 0.0) atomicAdd(A.dpe + (size_t)c * (KTYPE + 1) * (KTYPE + 1) + t, sdpe[t]);
    if (sdps[t] != 0.0) atomicAdd(A.dps + (size_t)c * (KTYPE + 1) * (KTYPE + 1) + t, sdps[t]);
  }
}

}  // namespace fmc
