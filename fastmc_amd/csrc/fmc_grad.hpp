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

#include <type_traits>

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
  double *dqfw;              // [ksfw][nfw] d/dq_i of the framework-framework reciprocal part by k range
  int ntt;                   // type-pair stride of dpe / dps: ntypes + 1
  double *e2, *ev;           // [nconf] real-space and LJ pair sums
  double *dpe, *dps;         // [nconf][ntt][ntt] dV/deps_ij, dV/dsigma_ij summed over pairs
  const double *vt;          // [npt][vstride] LJ tables of the theta pairs: V, dV/deps, dV/dsigma
  const double *rpe;         // [npt] 1 / eps of each theta pair
  double *dqpart;            // [ksplit][nconf][nfw] reciprocal d/dq_i of the guests by k range
  int ksplit;
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

// total S(k) of each configuration (the framework's plus its guest sites')
// and sum_k a_k |S_k|^2: one block per configuration.  The guest sites go in
// batches of nb: each batch's axis phase rows e^{i 2 pi l s_d}, l = 0..kmax_d,
// by the recursion of ewald.f:108-147 (one row per site and axis) and its
// charges in LDS; every thread then adds the batch to its KG k-vectors with
// site_phase's products (sites in order, as ewald1 sums them).
constexpr int KG = 4;
__global__ void __launch_bounds__(256) k_grad_sconf(const DevParams *__restrict__ Pp, GradArgs A, int nb) {
  extern __shared__ __align__(16) unsigned char glds[];
  const DevParams &P = *Pp;
  const int c = blockIdx.x, tid = threadIdx.x, nt = blockDim.x, kcp = P.kcp;
  double2 *rows = (double2 *)glds;
  double *rq = (double *)(rows + (size_t)nb * 3 * kcp);
  const double *r = P.rcell;
  const double *g = A.gx + (size_t)c * A.gstride * 3;
  int gcut[KMAXS + 1], nsite = 0;
  gcut[0] = 0;
#pragma unroll
  for (int gg = 1; gg <= KMAXS; ++gg) {
    if (gg <= A.ntpguest) nsite += A.nmols[c * A.ntpguest + gg - 1] * P.natoms[P.locguest[gg]];
    gcut[gg] = nsite;
  }
  double usum = 0.0;
  for (int k0 = 0; k0 < A.nk; k0 += KG * nt) {
    double2 acc[KG];
    int kk[KG], ll[KG], mm[KG], nn[KG];
#pragma unroll
    for (int m = 0; m < KG; ++m) {
      kk[m] = k0 + m * nt + tid;
      const int lmn = kk[m] < A.nk ? P.klmn[kk[m]] : 0;
      ll[m] = lmn & 255;
      mm[m] = ((lmn >> 8) & 255) - 128;
      nn[m] = ((lmn >> 16) & 255) - 128;
      acc[m] = kk[m] < A.nk ? A.sfw[kk[m]] : make_double2(0.0, 0.0);
    }
    for (int a0 = 0; a0 < nsite; a0 += nb) {
      const int cnt = min(nb, nsite - a0);
      __syncthreads();  // the previous batch is consumed
      for (int t = tid; t < 3 * cnt; t += nt) {
        const int ia = t / 3, d = t - 3 * ia, a = a0 + ia;
        const double x = g[3 * a], y = g[3 * a + 1], z = g[3 * a + 2];
        const double fr = r[d] * x + r[3 + d] * y + r[6 + d] * z;
        double s1, c1;
        sincospi(2.0 * fr, &s1, &c1);
        double2 *row = rows + (size_t)(ia * 3 + d) * kcp;
        const int kmax = d == 0 ? P.kmax1 : (d == 1 ? P.kmax2 : P.kmax3);
        row[0] = make_double2(1.0, 0.0);
        double cs = c1, sn = s1;
        for (int l = 1; l <= kmax; ++l) {
          row[l] = make_double2(cs, sn);
          const double cn = cs * c1 - sn * s1, snn = sn * c1 + cs * s1;
          cs = cn;
          sn = snn;
        }
        if (d == 0) {
          int gg = 1, base = 0;  // the guest kind holding site a (selects: gcut stays in registers)
#pragma unroll
          for (int q = 1; q < KMAXS; ++q)
            if (a >= gcut[q]) {
              gg = q + 1;
              base = gcut[q];
            }
          const int mol = P.locguest[gg], na = P.natoms[mol];
          rq[ia] = P.site_q[mol][(a - base) % na];
        }
      }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < KG; ++m) {
        if (kk[m] >= A.nk) continue;
        double sx = acc[m].x, sy = acc[m].y;
        for (int ia = 0; ia < cnt; ++ia) {
          double ckc, cks;
          site_phase<double>(rows + (size_t)ia * 3 * kcp, kcp, ll[m], mm[m], nn[m], rq[ia], ckc, cks);
          sx += ckc;
          sy += cks;
        }
        acc[m] = make_double2(sx, sy);
      }
    }
#pragma unroll
    for (int m = 0; m < KG; ++m)
      if (kk[m] < A.nk) {
        A.sc[(size_t)c * A.nk + kk[m]] = acc[m];
        usum += P.akk[kk[m]] * (acc[m].x * acc[m].x + acc[m].y * acc[m].y);
      }
  }
  usum = wave_sum(usum);
  __shared__ double ured[4];
  if ((tid & 63) == 0) ured[tid >> 6] = usum;
  __syncthreads();
  if (tid == 0) {
    double u = 0.0;
    for (int wv = 0; wv < (nt >> 6); ++wv) u += ured[wv];
    A.urec[c] = u;
  }
}

// d/dq_i of C sum_k a_k |S_k|^2 = 2C sum_k a_k Re(e_ik S_k*), split as
// S = S_framework + S_guests.  The framework part (FP64, the same for every
// configuration): block (x, z) takes the framework atoms [64 x, 64 x + 64)
// and the z-th of ks equal ranges of k, writing its partial sum to
// dqfw[z][i] (the host adds the ranges in order).
__global__ void __launch_bounds__(64) k_grad_dq_fw(const DevParams *__restrict__ Pp, GradArgs A, int ks) {
  const DevParams &P = *Pp;
  const int i = blockIdx.x * blockDim.x + threadIdx.x, z = blockIdx.y;
  const int kb = (int)((long long)A.nk * z / ks), ke = (int)((long long)A.nk * (z + 1) / ks);
  if (i >= A.nfw) return;
  double acc = 0.0;
  for (int k = kb; k < ke; ++k) {
    const double2 ek = A.efw64[(size_t)k * A.nfw + i], sf = A.sfw[k];
    const double ak = P.akk[k];
    acc += ek.x * (ak * sf.x) + ek.y * (ak * sf.y);
  }
  A.dqfw[(size_t)z * A.nfw + i] = 2.0 * P.dc_C * acc;
}

// The guest part of the same contraction for all configurations at once, as
// one blocked product: block (x, y, z) takes the framework atoms
// [256 x, 256 x + 256), the configurations [CT y, CT y + CT) and the z-th of
// A.ksplit equal ranges of k; the weights a_k (S_k - S_framework,k) of its CT
// configurations are staged in LDS by chunks of k, and each framework phase
// e_ik read from memory serves CT configurations (CT accumulators per
// thread).  Each block writes its partial sums to dqpart[z]; k_grad_pairs
// adds the ranges in order.  T float: FP32 products of FP32 phases and
// weights, FP64 accumulation; double: all FP64.
constexpr int GCT = 16, GKC = 128;
template <typename T>
__global__ void __launch_bounds__(256) k_grad_dq_block(const DevParams *__restrict__ Pp, GradArgs A) {
  using T2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
  const DevParams &P = *Pp;
  const int c0 = blockIdx.y * GCT, i = blockIdx.x * blockDim.x + threadIdx.x;
  const int kb = (int)((long long)A.nk * blockIdx.z / A.ksplit), ke = (int)((long long)A.nk * (blockIdx.z + 1) / A.ksplit);
  __shared__ T2 w[GCT][GKC];
  double acc[GCT];
#pragma unroll
  for (int cc = 0; cc < GCT; ++cc) acc[cc] = 0.0;
  for (int k0 = kb; k0 < ke; k0 += GKC) {
    const int nkc = min(GKC, ke - k0);
    __syncthreads();
    for (int t = threadIdx.x; t < GCT * GKC; t += blockDim.x) {
      const int cc = t / GKC, kk = t - cc * GKC, c = c0 + cc;
      double2 v = make_double2(0.0, 0.0);
      if (c < A.nconf && kk < nkc) {
        const double2 sf = A.sfw[k0 + kk], st = A.sc[(size_t)c * A.nk + k0 + kk];
        const double ak = P.akk[k0 + kk];
        v = make_double2(ak * (st.x - sf.x), ak * (st.y - sf.y));
      }
      T2 wv;
      wv.x = (T)v.x;
      wv.y = (T)v.y;
      w[cc][kk] = wv;
    }
    __syncthreads();
    if (i < A.nfw) {
      for (int kk = 0; kk < nkc; ++kk) {
        T2 ek;
        if constexpr (sizeof(T) == 4) ek = A.efw32[(size_t)(k0 + kk) * A.nfw + i];
        else ek = A.efw64[(size_t)(k0 + kk) * A.nfw + i];
#pragma unroll
        for (int cc = 0; cc < GCT; ++cc) {
          const T2 wk = w[cc][kk];
          const T p = ek.x * wk.x + ek.y * wk.y;
          acc[cc] += (double)p;
        }
      }
    }
  }
  if (i < A.nfw) {
    double *part = A.dqpart + (size_t)blockIdx.z * A.nconf * A.nfw;
#pragma unroll
    for (int cc = 0; cc < GCT; ++cc)
      if (c0 + cc < A.nconf) part[(size_t)(c0 + cc) * A.nfw + i] = 2.0 * P.dc_C * acc[cc];
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
// parlst (nlist_builders.f:120).  The erfc table is staged in LDS; dV/deps
// of a pair is its interpolated V times 1/eps of the pair (the table of
// dV/deps is V/eps at every grid point), so two LJ tables are read per pair.
constexpr int GPS = 8;  // lanes per atom in k_grad_pairs
__global__ void __launch_bounds__(256) k_grad_pairs(const DevParams *__restrict__ Pp, GradArgs A) {
  extern __shared__ __align__(16) unsigned char glds[];
  const DevParams &P = *Pp;
  const int c = blockIdx.y;
  const int tid = threadIdx.x;
  __shared__ double sdpe[(KTYPE + 1) * (KTYPE + 1)], sdps[(KTYPE + 1) * (KTYPE + 1)];
  __shared__ double red[2][4];
  double *serc = (double *)glds;  // erfc table [mxegrd + 3]
  for (int t = tid; t < P.mxegrd + 3; t += blockDim.x) serc[t] = P.erc[t];
  for (int t = tid; t < (KTYPE + 1) * (KTYPE + 1); t += blockDim.x) sdpe[t] = sdps[t] = 0.0;
  __syncthreads();
  const double *g = A.gx + (size_t)c * A.gstride * 3;
  int ng = 0;
  for (int gg = 1; gg <= A.ntpguest; ++gg) ng += A.nmols[c * A.ntpguest + gg - 1] * P.natoms[P.locguest[gg]];
  const double kq = R4PIE0 / P.epsq;
  double e2 = 0.0, ev = 0.0;
  // atom index space: framework atoms, then guest atoms (for guest-guest
  // pairs); GPS consecutive lanes share an atom, lane `sub` of the group
  // taking the guest molecules m = sub, sub + GPS, ...
  for (int t = blockIdx.x * blockDim.x + tid; t < (A.nfw + ng) * GPS; t += gridDim.x * blockDim.x) {
    const int a = t / GPS, sub = t - a * GPS;
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
    int base = 0;
    for (int gg = 1; gg <= A.ntpguest; ++gg) {
      const int mol = P.locguest[gg], na = P.natoms[mol], nm = A.nmols[c * A.ntpguest + gg - 1];
      for (int m = sub; m < nm; m += GPS) {
        if (ia >= 0 && gg * 1000000 + m == mola) continue;  // no intramolecular pairs
        for (int s2 = 0; s2 < na; ++s2) {
          const int b = base + m * na + s2;
          if (ia >= 0 && b <= ia) continue;  // each guest pair once
          double dx = xa - g[3 * b], dy = ya - g[3 * b + 1], dz = za - g[3 * b + 2];
          min_image(P, dx, dy, dz);
          double rsq = dx * dx + dy * dy + dz * dz;
          if (rsq >= P.rcut2) continue;
          double r = sqrt(rsq);
          double qb = P.site_q[mol][s2];
          int tb = P.site_t[mol][s2];
          double chgprd = qa / P.epsq * R4PIE0 * qb;
          if (fabs(chgprd) > 1e-10) {
            double er = interp_tab(serc, r, P.rdrewd);
            e2 += er * chgprd;
            if (ia < 0) dqa += kq * qb * er;
          }
          int pt = A.ptab[ta * (KTYPE + 1) + tb];
          if (pt >= 0) {
            const double *tab = A.vt + (size_t)pt * 3 * A.vstride;
            const double v = interp_tab(tab, r, P.rdlrpot);
            ev += v;
            if (ia < 0) {
              int lo = min(ta, tb), hi = max(ta, tb);
              atomicAdd(&sdpe[hi * (KTYPE + 1) + lo], v * A.rpe[pt]);
              atomicAdd(&sdps[hi * (KTYPE + 1) + lo], interp_tab(tab + 2 * A.vstride, r, P.rdlrpot));
            }
          } else if (pt == -1) {  // no framework type: the FIELD row
            int k = P.pairidx[ta][tb];
            if (k >= 0) ev += interp_tab(P.vvv + (size_t)k * P.vstride, r, P.rdlrpot);
          }
        }
      }
      base += nm * na;
    }
    // the group's lanes add their molecules' terms of the atom (fixed order)
#pragma unroll
    for (int off = 1; off < GPS; off <<= 1) dqa += __shfl_xor(dqa, off, 64);
    if (ia < 0 && sub == 0) {
      // the reciprocal part (k ranges in order) and the real-space part
      double dr = 0.0;
      for (int z = 0; z < A.ksplit; ++z) dr += A.dqpart[((size_t)z * A.nconf + c) * A.nfw + a];
      A.dq[(size_t)c * A.nfw + a] = dr + dqa;
    }
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
    const int hi = t / (KTYPE + 1), lo = t - hi * (KTYPE + 1);
    if (hi >= A.ntt || lo >= A.ntt) continue;  // types beyond ntypes: no pairs
    const size_t o = (size_t)c * A.ntt * A.ntt + (size_t)hi * A.ntt + lo;
    if (sdpe[t] != 0.0) atomicAdd(A.dpe + o, sdpe[t]);
    if (sdps[t] != 0.0) atomicAdd(A.dps + o, sdps[t]);
  }
}

}  // namespace fmc
