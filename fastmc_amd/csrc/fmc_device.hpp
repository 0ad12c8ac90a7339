// Device-side data layout and building blocks of the GCMC trial-move path
// for CDNA4 (gfx950).  One workgroup owns one walker (an independent Markov
// chain at one state point); its mutable state is staged in LDS for the whole
// launch, the trial energy is split over the workgroup (atoms for the pair
// terms, k-vectors for the reciprocal terms) and reduced with wavefront
// shuffles.  Reference routines restated here are cited per function.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fastmc.h"

namespace fmc {

constexpr int KMAXS = 4;          // species handled by the kernels (framework + 3 guests)
constexpr int KMAXC = KMAXS * (KMAXS + 1) / 2;
constexpr int KSITE = 8;          // sites per guest molecule
constexpr int KTYPE = FMC_MAX_TYPES;
constexpr int KCAP = 20;          // kmax + 1 cap for the per-site phase tables
constexpr int KTRI = KTYPE * (KTYPE + 1) / 2;
constexpr int KWAVES = 16;        // up to 1024 threads per workgroup
// Teams: a walker's step split over up to KTEAM workgroups (pair atoms and
// k-vectors by member), their partial sums exchanged once per step through
// XREC-double records in global memory, each value tagged with its exchange
// in its low mantissa bits.  A wait that sees nothing for XSPIN polls gives
// up and flags the walker (FMC_OVF_TEAM), so that no wait is unbounded.
constexpr int KTEAM = 16;
constexpr int XREC = 32;
constexpr long long XSPIN = 1LL << 22;

constexpr double R4PIE0 = 138935.4835;      // utility_pack.f:113
constexpr double KBOLTZ = 1.9872041e-3;     // utility_pack.f:121
constexpr double BOLTZ = 1.3806503e-23;     // utility_pack.f:124
constexpr double PI = 3.141592653589793;    // utility_pack.f:95
constexpr double SQRPI = 1.7724538509055159;
constexpr double DELRMIN = 0.1;             // utility_pack.f:107
constexpr double MINANGLE = PI / 18.0;      // utility_pack.f:104
constexpr double HPLANCK = 6.62607004e-34;  // utility_pack.f:144
constexpr double AVO = 6.022140857e23;      // utility_pack.f:127

// Read-only system description shared by all walkers of a handle.
struct DevParams {
  int imcon, maxmls, ntpguest, ntypes, nk, mxegrd, nfw, fw_per_walker;
  int kmax1, kmax2, kmax3, ntri;
  int nk64;                       // k-vectors [0, nk64) use FP64 phase products (list sorted by a_k)
  // LDS plan (bytes from the start of the dynamic LDS block): Smem, erfc
  // table, per-site phase tables [2][nsmax][3][kcp] double2, constants block;
  // with lds_all also the framework (x, y, z, q, info), the k list (a_k,
  // packed l m n), the guest coordinates (x, y, z over all guest slots) and,
  // in FP32 mode, the LJ columns of the guest-site types
  int off_erc, off_ph, off_cst, lds_bytes;
  int off_ph32;                   // FP32 copy of the phase tables (mixed mode's low-weight k)
  int nsmax, kcp;
  int lds_all;
  int off_fw, off_k, off_g, off_vvv;
  int nvcol;                      // staged LJ columns
  // mixed-mode LJ table: entries [lj_lo, lj_hi) of a column, the repulsive
  // wall where |V| is large, are interpolated from FP64 values (entries up to
  // lj_hi + 1), all others from FP32 values.  LDS copy: off_vvv64 holds
  // [nvcol][lj_n64] doubles (entries lj_lo ..), off_vvv [nvcol][vstride] floats
  int lj_lo, lj_hi, lj_n64, off_vvv64;
  int off_rs, rs_batch;           // LDS rows of the in-kernel re-sum: rs_batch atoms x [3][kcp]
                                  // double2, then their charges and species (0: no room)
  int lj_lds;                     // without lds_all: the LJ columns alone in LDS (FP64 mode:
                                  // [nvcol][vstride] doubles at off_vvv64)
  int resum_every;                // full FP64 S(k) re-sum after this many accepted moves (0: never)
  int spec_prop;                  // k_run builds the next step's proposal speculatively (1) or not (0)
  // Wang-Landau sampling in N of guest 1 (wang_landau.f); wl_nbins = 0: off
  int wl_nbins;                   // bins per walker: N = 0 .. wl_nbins-1 (maxn + 2)
  int wl_off;                     // LDS offset of the walker's ln g (double) then visits (int)
  int wl_visittol;                // wl_visit_tolerance
  double wl_flat;                 // wl_flat_coeff
  double wl_prectol;              // stop once ln f < wl_prec_tolerance
  int vcol_src[64];               // staged column -> table column
  int8_t lpairidx[KTYPE + 1][KTYPE + 1];  // type pair -> staged column or -1
  // constants of insertion / deletion bookkeeping (ewald.f:96-105,316-345)
  double dc_C;                    // 2 rvolm r4pie0 / epsq
  double dc_esic[KMAXS + 1];      // engsictmp of inserting guest g: -r4pie0/epsq alpha sum q^2 / sqrt(pi)
  // gstlrcorrect as a linear function of the guest counts (exact algebra of
  // vdw_terms.f:717-765): delrc(g, lexisting) = tl_c0 + sum_g' tl_c1[g'] N_g',
  // delrc_mol(loc2(mol, jt)) = tm_c0[jt] + tm_c1[jt] N_jt
  double tl_c0[KMAXS + 1][2], tl_c1[KMAXS + 1][2][KMAXS + 1];
  double tm_c0[KMAXS + 1][2][KMAXS + 1], tm_c1[KMAXS + 1][2][KMAXS + 1];
  int locguest[KMAXS + 1];        // guest g (1-based) -> species index
  int gofspecies[KMAXS + 1];      // species -> guest g, 0 for framework
  int natoms[KMAXS + 1];          // by species
  int gcap[KMAXS + 1];            // guest g: molecule capacity
  int gbase[KMAXS + 1];           // guest g: first atom slot in the walker's guest arrays
  int gatoms_total;               // guest atom slots per walker
  double cell[9], rcell[9], inv_cell[3];  // inv_cell: 1/cell(1), 1/cell(5), 1/cell(9) (images)
  double volm, statvolm, rcut2, rclim2, overlap2, drewd, rdrewd, dlrpot, rdlrpot;
  double alpha, epsq, engunit, rvolm, rengunit;  // rengunit = 1 / engunit
  double site_q[KMAXS + 1][KSITE], site_m[KMAXS + 1][KSITE];
  double site_chgea[KMAXS + 1][KSITE];          // q / epsq * r4pie0 (ewald.f:686)
  double tmpl[KMAXS + 1][KSITE][3];            // guest g template (COM-centred)
  int site_t[KMAXS + 1][KSITE], site_f[KMAXS + 1][KSITE];
  double ewald3en[KMAXS + 1];
  double selfq2[KMAXS + 1], chgtmp[KMAXS + 1];  // guest g: sum q^2, sum q
  double steadd[KTRI + 1];
  int numtyp_g[KMAXS + 1][KTYPE + 1], numfrz_g[KMAXS + 1][KTYPE + 1];  // numtyp_gstmol
  int8_t pairidx[KTYPE + 1][KTYPE + 1];       // LJ table column or -1 (srfrce skip rule)
  // surface_check (mc_moves.f:10-44): (surftol + sigma of the type pair)^2,
  // sigma = prmvdw(ivdw,2) of the pair's VDW row (0 for a pair without one)
  double surf2[KTYPE + 1][KTYPE + 1];
  // production statistics (gcmc.f:1050-1130): window length in counted steps
  // (0: windows close only at the end), accepted steps only, window capacity
  int nwindsteps, accsample, maxwin;
  // probability plots (storeprob, utility_pack.f:622-687): grid, its size,
  // reciprocal of the unit cell used for folding, and per plot the guest
  // kind and site list (0: centre of mass)
  int nplots, pgrid[3], pgridsize;
  // guest-count trace every nnumg steps (0: off), numg_cap records per walker
  int nnumg, numg_cap;
  double prucell[9];
  int plot_g[FMC_MAX_PLOTS], plot_ns[FMC_MAX_PLOTS], plot_site[FMC_MAX_PLOTS][FMC_MAX_SITES + 1];
  // tables
  const double *erc;      // [mxegrd+1]
  const double *vvv;      // [npt][mxegrd+1]
  const float *erc32;
  const float *vvv32;
  int vstride;
  // k-vectors
  const int *klmn;        // packed (ll | (mm+128)<<8 | (nn+128)<<16)
  const double *akk;
  // framework (shared or per walker)
  const double *fwx, *fwy, *fwz;
  const double *fwq;
  const int *fwinfo;      // type | species<<8 | frozen<<16
};

// Per-walker mutable state (global memory; staged into LDS by each launch).
struct WState {
  // duni (utility_pack.f:2768-2857) as a lag-97/33 subtractive generator with
  // a pre-generated output ring: rq[n & 127] holds call n+1's value
  float u[98];
  float rq[128];
  long long mz_made, mz_used;
  int mz_ir, mz_jr, mz_c, mz_pad;  // ir, jr of the next call and c (units of 2^-24) after mz_made calls
  double temp, beta;
  double fug[KMAXS + 1];
  float mcf[KMAXS + 1][FMC_NMOVES];
  double delr[KMAXS + 1], delrdisp[KMAXS + 1], tran_delr[KMAXS + 1];
  double rotangle, rota_rotangle;
  int disp_count[KMAXS + 1], accept_disp[KMAXS + 1];
  int tran_count[KMAXS + 1], accept_tran[KMAXS + 1];
  int rota_count[KMAXS + 1], accept_rota[KMAXS + 1];
  int nmols[KMAXS + 1];                     // by species
  double energy[KMAXS + 2];
  double sumchg, chgsum_mol[KMAXS + 1], engsic[KMAXC + 1];
  double elrc, elrc_mol[KMAXC + 1];
  int numtyp[KTYPE + 1], numfrz[KTYPE + 1];
  int numtyp_mol[KMAXS + 2][KTYPE + 1], numfrz_mol[KMAXS + 2][KTYPE + 1];
  // pending trial of the one-trial-at-a-time seam, and what its accept needs
  // (kept per walker, so trials of several walkers may be pending at once)
  int pend_kind, pend_iguest, pend_choice, pend_valid;
  double pend_xyz[3 * KSITE];
  double pend_delrc, pend_delE[KMAXS + 1], pend_delrc_mol[KMAXC + 1];
  int pend_surface, pend_pad;
  // statistics
  long long steps, acc[9], att[9], counted;
  double sum_n[KMAXS + 1], sum_n2[KMAXS + 1], sum_e, sum_e2;
  // surfacemols by species; production count, steps in the open window,
  // closed windows, and the open window's sums per guest g of N, E, E*N,
  // N*N, E*E and the surface count (chainstats, gcmc.f:1062-1075)
  int surfmols[KMAXS + 1];
  long long prodcount, rollcount, nwin, nnumg_n;
  double delE_last[KMAXS + 1];    // delE after the last step that moved (a switch leaves it)
  double chain[KMAXS + 1][FMC_WIN_STATS];
  // Wang-Landau walker: window [wl_min, wl_max] of N, preload target,
  // incremental minimum of the window's visit histogram, f and ln f
  int wl_on, wl_done, wl_min, wl_max;
  int wl_preload, wl_hmin, wl_hmin_cnt, wl_error;
  long long wl_hsum, wl_sweeps, wl_iter;
  double wl_f, wl_logf;
  double wl_lambda3;              // thermal wavelength cubed of guest 1 at this T (A^3), wang_landau.f:121-123
  int overflow, acc_since_resum;
  long long nresum;               // in-kernel FP64 re-sums done
  double resum_drift;             // largest |S_total(incremental) - S_total(re-sum)| seen at a re-sum
};

// Per-walker device arrays.
struct DevWalkers {
  WState *ws;
  double *gx, *gy, *gz;      // [w][gatoms_total]
  double2 *sk;               // [w][maxmls+1][nk]: species rows then the total
  double2 *sscr;             // [w][2][nk]: s_new, s_old of the current trial
  double2 *sbak;             // [w][maxmls+1][nk]: S(k) rows kept across a swap
  double *fwx, *fwy, *fwz;   // per-walker framework copies (if fw_per_walker)
  fmc_trace_rec *trace;      // trace of walker trace_w, or null
  int trace_w;
  fmc_trial *trial_out;      // seam output
  double *wl_lng;            // [w][wl_nbins] ln g(N) (dos_hist)
  int *wl_hist;              // [w][wl_nbins] visits (visit_hist)
  fmc_window *win;           // [w][maxwin] closed averaging windows
  double *prob;              // [w][nplots][pgridsize] probability plots
  fmc_numguests_rec *numg;   // [w][numg_cap] guest-count trace
  // teams: exchange records [w][2 (exchange parity)][KTEAM][XREC] (row
  // values tagged with their exchange in the low mantissa bits; zero: no
  // value is current), the number of exchanges each walker has made, and
  // the private guest coordinates of members 1.. when they live in global
  // memory [w][team_alloc - 1][3][gatoms_total]
  double *xch;
  long long *xcnt;
  double *gxt;
  int team_alloc;
  int *xerr;                 // set when a team wait gave up (read and cleared by the host)
};

__host__ __device__ inline int loc2(int i, int j) {  // utility_pack.f:2859-2872
  int a = i > j ? i : j, b = i > j ? j : i;
  return (a * (a - 1)) / 2 + b;
}

// Next duni value from the ring (thread 0).  The ring is refilled by
// mzt_refill before each step, so at least 33 values are always available.
__device__ inline double duni(WState &w) {
  float v = w.rq[w.mz_used & 127];
  w.mz_used += 1;
  return (double)v;
}

// duni, utility_pack.f:2768-2857, 33 calls at once on lanes 0..32 of one
// wave.  Call m reads u(ir) - u(jr) where ir and jr step down by one per call
// (wrapping 1 -> 97) and jr equals the ir of call m-33, so the 33 calls of a
// round read only values made before the round; c steps by -7654321 modulo
// 16777213 in units of 2^-24.  All operations are exact in binary32.  The
// caller advances mz_ir, mz_jr, mz_c and mz_made after the round.
__device__ inline void mzt_refill(WState &w, int lane) {
  if (lane < 33) {
    int ir = ((w.mz_ir - 1 - lane) % 97 + 97) % 97 + 1;
    int jr = ((w.mz_jr - 1 - lane) % 97 + 97) % 97 + 1;
    float a = w.u[ir], b = w.u[jr];
    float uni = a - b;
    if (uni < 0.0f) uni = uni + 1.0f;
    int cc = ((w.mz_c - (lane + 1) * 7654321) % 16777213 + 16777213) % 16777213;
    float c = (float)cc * (1.0f / 16777216.0f);
    float out = uni - c;
    if (out < 0.0f) out = out + 1.0f;
    __builtin_amdgcn_wave_barrier();
    w.u[ir] = uni;
    w.rq[(w.mz_made + lane) & 127] = out;
  }
}

__device__ inline void mzt_advance(WState &w) {
  w.mz_ir = ((w.mz_ir - 1 - 33) % 97 + 97) % 97 + 1;
  w.mz_jr = ((w.mz_jr - 1 - 33) % 97 + 97) % 97 + 1;
  w.mz_c = ((w.mz_c - 33 * 7654321) % 16777213 + 16777213) % 16777213;
  w.mz_made += 33;
}

// A value every lane of the wave holds (read from LDS, so the compiler cannot
// tell): taken from the first active lane, so that what it indexes is read
// through the scalar unit instead of one vector load per lane.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Lane k's value of v (k uniform).
__device__ __forceinline__ double rdlane(double v, int k) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), k);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), k);
  return __hiloint2double(hi, lo);
}

// images, utility_pack.f:1697-1902 for imcon 1, 2, 3 (nint = round half
// away).  In the orthogonal cells the nearest integer is taken with rint
// (one instruction instead of six): the two differ only when a * dx is an
// exact half-integer, and then both images lie at distance cell / 2, so the
// squared distance is the same.  The triclinic cell keeps nint.
__device__ inline void min_image(const DevParams &P, double &dx, double &dy, double &dz) {
  if (P.imcon == 1) {
    double a = P.inv_cell[0];
    dx = dx - P.cell[0] * rint(a * dx);
    dy = dy - P.cell[0] * rint(a * dy);
    dz = dz - P.cell[0] * rint(a * dz);
  } else if (P.imcon == 2) {
    dx = dx - P.cell[0] * rint(P.inv_cell[0] * dx);
    dy = dy - P.cell[4] * rint(P.inv_cell[1] * dy);
    dz = dz - P.cell[8] * rint(P.inv_cell[2] * dz);
  } else {
    const double *r = P.rcell, *c = P.cell;
    double sx = r[0] * dx + r[3] * dy + r[6] * dz;
    double sy = r[1] * dx + r[4] * dy + r[7] * dz;
    double sz = r[2] * dx + r[5] * dy + r[8] * dz;
    sx -= round(sx);
    sy -= round(sy);
    sz -= round(sz);
    dx = c[0] * sx + c[3] * sy + c[6] * sz;
    dy = c[1] * sx + c[4] * sy + c[7] * sz;
    dz = c[2] * sx + c[5] * sy + c[8] * sz;
  }
}

// Square root of a squared pair distance: the correctly rounded FP64 sequence
// the compiler emits (reciprocal square root, one Goldschmidt step, two
// residual corrections) without its rescaling of tiny or huge arguments and
// its zero / infinity fix-up, which pair distances never need (x is clamped
// to 1e-200; an exactly coincident pair is an overlap, never accepted).
__device__ inline double pair_sqrt(double x) {
  x = x > 1e-200 ? x : 1e-200;
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double e = fma(-h, g, 0.5);
  g = fma(g, e, g);
  h = fma(h, e, h);
  double d = fma(-g, g, x);
  g = fma(d, h, g);
  d = fma(-g, g, x);
  return fma(d, h, g);
}

// 3-point interpolation of ewald.f:703-715 / vdw_terms.f:441-453
template <typename T>
__device__ inline T interp3(const T *tab, T r, T rinv_step) {
  T x = r * rinv_step;
  int l = (int)x;
  T p = x - (T)l;
  T v0 = tab[l], v1 = tab[l + 1], v2 = tab[l + 2];
  T t1 = v0 + (v1 - v0) * p;
  T t2 = v1 + (v2 - v1) * (p - (T)1);
  return t1 + (t2 - t1) * p * (T)0.5;
}

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int CTRL, int ROW, int BANK>
__device__ inline double dpp_f64(double x) {
  int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, ROW, BANK, true);
  int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, ROW, BANK, true);
  return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a wave with DPP row shifts and row broadcasts
// (gfx9 cross-lane pattern); the total lands in lane 63.
__device__ inline double wave_sum_dpp(double v) {
  double s = v + dpp_f64<0x111, 0xf, 0xf>(v);   // row_shr:1
  s += dpp_f64<0x112, 0xf, 0xf>(v);             // row_shr:2
  s += dpp_f64<0x113, 0xf, 0xf>(v);             // row_shr:3
  s += dpp_f64<0x114, 0xf, 0xe>(s);             // row_shr:4, banks 1-3
  s += dpp_f64<0x118, 0xf, 0xc>(s);             // row_shr:8, banks 2-3
  s += dpp_f64<0x142, 0xa, 0xf>(s);             // row_bcast:15 into rows 1, 3
  s += dpp_f64<0x143, 0xc, 0xf>(s);             // row_bcast:31 into rows 2, 3
  return s;
}

// One halving step of wave_sum_rows: lanes exchange half of the C values
// they carry with their DPP partner (CTRL, an involution within each row of
// 16 lanes that flips lane bit `up`), keep the other half and add.
template <int C, int CTRL>
__device__ inline void wave_halve(double *v, bool up) {
  constexpr int H = C / 2;
#pragma unroll
  for (int i = 0; i < H; ++i) {
    const double send = up ? v[i] : v[i + H];
    const double keep = up ? v[i + H] : v[i];
    v[i] = keep + dpp_f64<CTRL, 0xf, 0xf>(send);
  }
}

// Sums of N per-lane values over the 64 lanes of a wave, all N at once by
// recursive halving: row mirror, half-row mirror, quad xor 2 and quad xor 1
// each halve the values a lane carries (K = N rounded up to 16 or 32, J =
// K / 16 left per lane), so that lane L holds its row's partial sums of
// values J * (L & 15) + j; two exchanges across rows (xor 16, xor 32) finish
// them.  Lanes 0..15 write the totals to out[0..N).  About K adds and K
// exchanges per lane instead of 6 N of each for N separate reductions.
template <int N>
__device__ inline void wave_sum_rows(const double *acc, int lane, double *out) {
  static_assert(N >= 1 && N <= 32, "wave_sum_rows: 1..32 values");
  constexpr int K = N <= 16 ? 16 : 32;
  constexpr int J = K / 16;
  double v[K];
#pragma unroll
  for (int i = 0; i < K; ++i) v[i] = i < N ? acc[i] : 0.0;
  wave_halve<K, 0x140>(v, (lane >> 3) & 1);      // row_mirror: r <-> 15 - r
  wave_halve<K / 2, 0x141>(v, (lane >> 2) & 1);  // row_half_mirror: r <-> 7 - r within 8
  wave_halve<K / 4, 0x4E>(v, (lane >> 1) & 1);   // quad_perm [2,3,0,1]
  wave_halve<K / 8, 0xB1>(v, lane & 1);          // quad_perm [1,0,3,2]
#pragma unroll
  for (int j = 0; j < J; ++j) {
    v[j] += __shfl_xor(v[j], 16, 64);
    v[j] += __shfl_xor(v[j], 32, 64);
  }
  if (lane < 16) {
#pragma unroll
    for (int j = 0; j < J; ++j)
      if (J * lane + j < N) out[J * lane + j] = v[j];
  }
}

}  // namespace fmc
