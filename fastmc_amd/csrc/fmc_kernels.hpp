// Kernels of the GCMC trial-move path.  Template parameters: NS = number of
// molecule species (maxmls, compile-time so the per-species accumulators live
// in registers) and T = arithmetic type of the pair and phase products
// (double: FP64 parity mode; float: mixed mode with FP64 S(k), FP64
// coordinate differences and FP64 accumulation).
#pragma once

#include "fmc_device.hpp"

namespace fmc {

// Step clock (built with -DFMC_PHASE_CLOCK only): thread 0 of workgroup 0
// (walker 0, team member 0) adds the time of each phase of k_run's steps,
// read from the 100 MHz real-time counter, to an LDS row, and the row is
// added to fmc_phase_ticks at the end of the launch (fmc_phase_clock reads
// it).  Slot FMC_NPHASE holds the last stamp, FMC_NPHASE + 1 the steps and
// FMC_NPHASE + 2 the steps that took a speculative proposal.  Without the
// macro every stamp compiles to nothing.
constexpr int FMC_NPHASE = 15;
#ifdef FMC_PHASE_CLOCK
__device__ unsigned long long fmc_phase_ticks[FMC_NPHASE + 3];
__shared__ unsigned long long fmc_phase_row[FMC_NPHASE + 3];
__device__ __forceinline__ void phase_stamp(int i) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const unsigned long long t = __builtin_amdgcn_s_memrealtime();
    fmc_phase_row[i] += t - fmc_phase_row[FMC_NPHASE];
    fmc_phase_row[FMC_NPHASE] = t;
  }
}
#define FMC_PHASE(i) phase_stamp(i)
#define FMC_PHASE_COUNT(i) \
  do {                                                        \
    if (threadIdx.x == 0 && blockIdx.x == 0) fmc_phase_row[i] += 1; \
  } while (0)
#else
#define FMC_PHASE(i) ((void)0)
#define FMC_PHASE_COUNT(i) ((void)0)
#endif

template <int NS>
struct Acc {
  static constexpr int NACC = 6 * NS + 2;
  // layout: [0,NS) real_new, [NS,2NS) vdw_new, [2NS,3NS) real_old, [3NS,4NS) vdw_old,
  //         4NS rec_new, 4NS+1 rec_old, [4NS+2,5NS+2) mix_new, [5NS+2,6NS+2) mix_old
  // reduced with two more rows, the flag counts of the new and the existing
  // configuration (overlaps + 1024 x surface contacts, exact in FP64)
  static constexpr int NRED = NACC + 2;
};

// One proposal: the move and its trial sites as propose makes them, and the
// walker-state updates that come with it (ring position, the move's counter,
// adapted step sizes).  commit makes a record the step's live proposal; the
// two speculative records are built during a step for the next one.
struct Prop {
  int ok;  // a speculative record that commit may take (not a swap)
  int wlmode, mkind, swap_phase, kind, g, mol, natms, nmols, choice, has_new, has_old, excl_lo, excl_hi;
  int cnt, cval;  // counter of the move (1 disp_count, 2 tran_count, 3 rota_count of guest g) and its new value
  int chg;        // step sizes adapted: bit 0 delrdisp[g], 1 rotangle, 2 tran_delr[g], 3 rota_rotangle
  long long used0, used1;  // ring position before and after the proposal's draws
  double delrdisp, rotangle, tran_delr, rota_rotangle;
  double xn[KSITE][3], xo[KSITE][3];
};

template <int NS>
struct Smem {
  WState ws;
  int kind, g, mol, natms, choice, nmols, has_new, has_old, excl_lo, excl_hi, nfw;
  int wlmode;  // 0 grand-canonical move, 1 Wang-Landau move, 2 insert_guests preload
  // move kind of the step (FMC_MOVE_*; jump, swap evaluate as displacement /
  // deletion + insertion) and the swap's state: phase 1 the provisional
  // deletion of molecule swap_ichoice of guest swap_ig, phase 2 the
  // insertion of a guest swap_jg molecule at its COM
  int mkind, swap_phase, swap_ig, swap_jg, swap_ichoice, swap_ovl, swap_n0, swap_srf;
  double swap_com[3], swap_estepi, swap_rande;
  double bk_energy[KMAXS + 2], bk_chg[KMAXS + 1], bk_elrc, bk_elrc_mol[KMAXC + 1], bk_engsic[KMAXC + 1];
  int bk_surf[KMAXS + 1];
  int accepted, drew, ovl_new, ovl_old;
  int srf_new, srf_old;  // surface_check flags of the new / existing position
  double xn[KSITE][3], xo[KSITE][3];
  double red[KWAVES][Acc<NS>::NRED];
  double tot[Acc<NS>::NRED];
  double estep, rande, delrc;
  double delE[KMAXS + 1], delrc_mol[KMAXC + 1];
  double comp[4];
  // reciprocal cell and kmax per axis for lane-indexed access (from DevParams)
  double rc9[9];
  int kmx[3];
  // per species: site charges (as ewald2 uses them and bare); per guest kind:
  // species, sites, first guest slot.  Copies of DevParams entries, so that a
  // step reads them with independent LDS loads instead of chained scalar loads.
  double scq[KMAXS + 1][KSITE], sq[KMAXS + 1][KSITE];
  double ovl2;
  // cell, site masses by species and guest templates (COM-centred) by guest kind
  double cl9[9], sm[KMAXS + 1][KSITE], tmpl[KMAXS + 1][KSITE][3];
  int kmol[KMAXS + 1], kna[KMAXS + 1], kslot[KMAXS + 1];
  // first index of each guest kind in the pair loop's atom space (framework
  // atoms first) and the end of the space; kept current by apply
  int gstart[KMAXS + 1], ntot, a0, a1;
  // constants of the current trial, computed while the energy terms run
  double qfix0, pref;
  double qfm[KMAXS + 1];
  // team of the walker: size, this workgroup's member index and its
  // k-vectors [k0, k1); exchanges made so far; set when a wait gave up
  int team, member, k0, k1;
  long long xseq;
  int team_fail;
  // number of the last evaluation whose phase tables are written (evaluate)
  int tab_ready;
  // proposals: [0], [1] built speculatively for the next step assuming this
  // one is rejected after 0 / 1 draws of its decision, [2] the live one;
  // spec_on: [0] and [1] were built this step; mz_after: the ring position
  // after the live proposal's draws
  Prop prop[3];
  int spec_on;
  long long mz_after;
  // the move of the last completed step, for the work after it (numguests
  // records on wave 1) while wave 0 already commits the next proposal
  int done_mkind, done_g, done_swap_ig, done_swap_jg;
};

constexpr int KPT = 2;  // k-vectors per thread held in registers for a whole launch

// Register-resident share of the walker's structure factor: slot m of a
// thread is k = kslot(k0, m) of the workgroup's k-vectors [k0, k1).
template <int NS>
struct KSlot {
  double2 S[NS + 1];  // species rows, then the total
  double2 sn, so;     // s_new, s_old of the current trial
};

// k of register slot m (reverse thread order: the low threads hold the extra
// atoms of the pair loop, the high threads the extra k-vectors)
__device__ inline int kslot(int k0, int m) {
  return k0 + (int)(blockDim.x - 1 - threadIdx.x) + m * (int)blockDim.x;
}

template <int NS>
__device__ __forceinline__ void load_slots(const DevParams &P, const DevWalkers &W, int w, int k0, int k1,
                                           KSlot<NS> *ks) {
  const double2 *skw = W.sk + (size_t)w * (NS + 1) * P.nk;
#pragma unroll
  for (int m = 0; m < KPT; ++m) {
    int k = kslot(k0, m);
#pragma unroll
    for (int i = 0; i <= NS; ++i) ks[m].S[i] = k < k1 ? skw[(size_t)i * P.nk + k] : make_double2(0.0, 0.0);
    ks[m].sn = ks[m].so = make_double2(0.0, 0.0);
  }
}

template <int NS>
__device__ __forceinline__ void store_slots(const DevParams &P, const DevWalkers &W, int w, int k0, int k1,
                                            const KSlot<NS> *ks) {
  double2 *skw = W.sk + (size_t)w * (NS + 1) * P.nk;
#pragma unroll
  for (int m = 0; m < KPT; ++m) {
    int k = kslot(k0, m);
    if (k < k1) {
#pragma unroll
      for (int i = 0; i <= NS; ++i) skw[(size_t)i * P.nk + k] = ks[m].S[i];
    }
  }
}

// ------------------------------------------------------------ tail term --
// gstlrcorrect, vdw_terms.f:686-776, in its linear form (coefficients built on
// the host from the same expressions): total delrc and the loc2(mol, jt)
// entries of delrc_mol.
template <int NS>
__device__ double gstlrcorrect(const DevParams &P, const WState &ws, int g, int lexisting,
                               double *delrc_mol) {
  g = uni(g);
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
// Rotation matrix of quaternion q, rotation (utility_pack.f:3480-3524).
__device__ inline void rot_matrix(const double *q, double *Q) {
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
}

// random_ins / random_rot quaternion, utility_pack.f:3291-3310,3544-3560
__device__ inline void quat_from(double r1, double r2, double r3, double r4, double scale, double *q) {
  double theta = r1 * scale;
  double b1 = 2.0 * r2 - 1.0, b2 = 2.0 * r3 - 1.0, b3 = 2.0 * r4 - 1.0;
  double st, ct;
  sincos(theta / 2.0, &st, &ct);
  double norm = sqrt(b1 * b1 + b2 * b2 + b3 * b3);
  q[0] = ct;
  q[1] = st * (b1 / norm);
  q[2] = st * (b2 / norm);
  q[3] = st * (b3 / norm);
}

// Guest coordinates of the walker: read from the global arrays, or from
// their LDS copy when LF is set; gx, gy, gz are the global arrays this
// workgroup keeps current.  Team member 0 (and a walker without a team)
// owns the walker's arrays; a member m > 0 reads its LDS copy (LF) and
// writes no global array, or (without LF) keeps a private global copy.
template <bool LF>
struct GuestXYZ {
  const double *x, *y, *z;  // slot 0 of guest 1
  double *gx, *gy, *gz;
  __device__ GuestXYZ(const DevParams &P, const DevWalkers &W, int w, unsigned char *lds, int member = 0) {
    const size_t n = P.gatoms_total;
    if (member == 0) {
      gx = W.gx + (size_t)w * n;
      gy = W.gy + (size_t)w * n;
      gz = W.gz + (size_t)w * n;
    } else if (!LF) {
      gx = W.gxt + ((size_t)w * (W.team_alloc - 1) + member - 1) * 3 * n;
      gy = gx + n;
      gz = gy + n;
    } else {
      gx = gy = gz = nullptr;
    }
    if constexpr (LF) {
      x = (const double *)(lds + P.off_g);
      y = x + n;
      z = y + n;
    } else {
      x = gx;
      y = gy;
      z = gz;
    }
  }
};

template <bool LF>
__device__ inline void load_molecule(const DevParams &P, const GuestXYZ<LF> &G, int g, int choice,
                                     double (*x)[3], int natms) {
  int base = P.gbase[g] + (choice - 1) * natms;
  for (int i = 0; i < natms; ++i) {
    x[i][0] = G.x[base + i];
    x[i][1] = G.y[base + i];
    x[i][2] = G.z[base + i];
  }
}

// ---------------------------------------------------------- Wang-Landau --
// The walker's ln g(N) and visit histogram (dos_hist / visit_hist of
// wang_landau_module.f), staged in LDS for the launch; bin b holds N = b.
struct WLView {
  double *lng;
  int *hist;
  int nbins;
};

__device__ inline WLView wl_view(const DevParams &P, unsigned char *lds) {
  double *l = (double *)(lds + P.wl_off);
  return WLView{l, (int *)(l + P.wl_nbins), P.wl_nbins};
}

// minimum of the window's visits and how many bins hold it
__device__ inline void wl_rescan(WState &ws, const WLView &v) {
  int m = 0x7fffffff, c = 0;
  for (int b = ws.wl_min; b <= ws.wl_max; ++b) {
    int h = v.hist[b];
    if (h < m) {
      m = h;
      c = 1;
    } else if (h == m) {
      ++c;
    }
  }
  ws.wl_hmin = m;
  ws.wl_hmin_cnt = c;
}

// One visit of N = b: visit_hist += 1, dos_hist += ln f (wang_landau.f:148-150,
// 189-191, 668-677, 757-766), keeping the window minimum incrementally.
__device__ inline void wl_visit(WState &ws, const WLView &v, int b) {
  if (b < 0 || b >= v.nbins) return;  // outside the arrays (set-up keeps N <= maxn)
  v.lng[b] += ws.wl_logf;
  const bool inwin = b >= ws.wl_min && b <= ws.wl_max;
  if (inwin && v.hist[b] == ws.wl_hmin) ws.wl_hmin_cnt -= 1;
  v.hist[b] += 1;
  if (inwin) {
    ws.wl_hsum += 1;
    if (ws.wl_hmin_cnt == 0) wl_rescan(ws, v);
  }
}

// convergence_check (wang_landau.f:413-459) after every flat-histogram step:
// flat when every window bin was visited and min >= flatcoeff * mean or
// min >= visittol; then f -> sqrt(f) (adjust_factor, 527-538), ln f = log f,
// stop once ln f < prectol, else clear the window's visits (513-525).
__device__ inline void wl_check(const DevParams &P, WState &ws, const WLView &v) {
  if (ws.wl_hmin < 1) return;
  const int nb = ws.wl_max - ws.wl_min + 1;
  const double ave = (double)ws.wl_hsum / (double)nb;
  if (!((double)ws.wl_hmin >= P.wl_flat * ave || ws.wl_hmin >= P.wl_visittol)) return;
  ws.wl_sweeps += 1;
  ws.wl_f = sqrt(ws.wl_f);
  ws.wl_logf = log(ws.wl_f);
  if (ws.wl_logf < P.wl_prectol) {
    ws.wl_done = 1;
    return;
  }
  for (int b = ws.wl_min; b <= ws.wl_max; ++b) v.hist[b] = 0;
  ws.wl_hsum = 0;
  ws.wl_hmin = 0;
  ws.wl_hmin_cnt = nb;
}

// ------------------------------------------------------------- proposal --
// Main-loop move choice and step-size adaptation (gcmc.f:805-911), then the
// trial geometry of mc_insert / mc_delete / mc_displace (gcmc.f:1693-1935)
// with random_ins, random_disp, random_rot, translate and rotation
// (utility_pack.f:3161-3563).  Thread 0 only.
// Move from the cumulative fractions f (gcmc.f:821-844): insert, delete,
// displace, jump, swap, translate, rotate, switch.  gc: the grand-canonical
// loop; the Wang-Landau loop has no jump, swap or switch and displaces in
// their bands.  skip_ins / skip_del re-choose without that band, as the
// Wang-Landau loop does at the window edges (wang_landau.f:161-235).
__device__ inline int choose_move(const float *f, double randmov, bool skip_ins, bool skip_del, bool gc) {
  if (!skip_ins && randmov < (double)f[0]) return FMC_MOVE_INSERT;
  if (skip_ins && randmov < (double)f[1]) return FMC_MOVE_DELETE;
  if (!skip_ins && !skip_del && randmov < (double)f[1]) return FMC_MOVE_DELETE;
  if (randmov < (double)f[2]) return FMC_MOVE_DISPLACE;
  if (randmov < (double)f[3]) return gc ? FMC_MOVE_JUMP : FMC_MOVE_DISPLACE;
  if (randmov < (double)f[4]) return gc ? FMC_MOVE_SWAP : FMC_MOVE_DISPLACE;
  if (randmov < (double)f[5]) return FMC_MOVE_TRANSLATE;
  if (randmov < (double)f[6]) return FMC_MOVE_ROTATE;
  if (randmov < (double)f[7]) return gc ? FMC_MOVE_SWITCH : FMC_MOVE_DISPLACE;
  return FMC_MOVE_DISPLACE;
}

// Wave-parallel proposal: the 64 lanes of one wave run it with wave-uniform
// control flow.  The uniforms of the step are read once from the duni ring
// from position used0 (lane i holds call used0+i+1; the n-th draw is a
// readlane), lane i builds site i of the trial, and values every lane
// computes alike stay uniform.  The proposal goes to the record o (lane 0,
// and lane i for site i); the walker state is only read, except by the
// Wang-Landau bookkeeping (preload count, visits at the window edges: all
// lanes with identical values), which runs for live proposals only.
// spec: a speculative proposal for the next step (it leaves a swap's set-up
// undone and marks the record not to be taken).
// Kept out of line: one copy serves the live and the speculative calls, which
// keeps k_run's code within the instruction cache (the guest-coordinate view
// is passed by value for the same reason: it is two pointers).
template <int NS, bool LF>
__device__ __attribute__((noinline)) void propose(const DevParams &P, const GuestXYZ<LF> G, Smem<NS> &sh,
                                                  unsigned char *lds, const long long used0, const int slot,
                                                  const bool spec) {
  Prop &o = sh.prop[slot];
  // guest coordinates: with LF their LDS copy, addressed from the LDS base
  // here so that the reads are LDS reads
  const double *Gx = G.x, *Gy = G.y, *Gz = G.z;
  if constexpr (LF) {
    Gx = (const double *)(lds + P.off_g);
    Gy = Gx + P.gatoms_total;
    Gz = Gy + P.gatoms_total;
  }
  WState &ws = sh.ws;
  const int lane = threadIdx.x & 63;
  const int rbits = __float_as_int(ws.rq[(int)((used0 + lane) & 127)]);
  int nu = 0;
  auto draw = [&]() -> double {
    const float v = __int_as_float(__builtin_amdgcn_readlane(rbits, nu));
    ++nu;
    return (double)v;
  };
  const int wl_on = uni(ws.wl_on);
  int g = 1, kind, wlmode = 0;
  if (__builtin_expect(wl_on && uni(ws.nmols[P.locguest[1]]) < uni(ws.wl_preload), 0)) {
    // insert_guests (mc_moves.f:84-149) up to the window's lower edge
    wlmode = 2;
    kind = FMC_MOVE_INSERT;
    const long long it = ws.wl_iter;
    if (lane == 0) ws.wl_iter = it + 1;
  } else {
    // the grand-canonical loop draws the guest type only when there is a
    // choice (gcmc.f:806-808); the Wang-Landau loop always draws (wang_landau.f:137)
    if (P.ntpguest > 1 || wl_on) g = uni((int)floor(draw() * P.ntpguest) + 1);
    const int molg = P.locguest[g];
    const double randmov = draw();
    const float *f = ws.mcf[g];
    kind = uni(choose_move(f, randmov, false, false, !wl_on));
    if (__builtin_expect(wl_on && g == 1, 0)) {
      // window edges of the flat-histogram loop, wang_landau.f:161-235 (the
      // visit bookkeeping is written by every lane with the same values)
      wlmode = 1;
      const WLView v = wl_view(P, lds);
      const int nmol = uni(ws.nmols[molg]);
      const int wmin = uni(ws.wl_min), wmax = uni(ws.wl_max);
      if (nmol < wmin) {
        kind = FMC_MOVE_INSERT;
      } else if (nmol <= wmin && kind == FMC_MOVE_DELETE) {
        wl_visit(ws, v, nmol);
        kind = uni(choose_move(f, randmov, false, true, false));
      }
      if (nmol > wmax) {
        kind = FMC_MOVE_DELETE;
      } else if (nmol >= wmax && kind == FMC_MOVE_INSERT) {
        wl_visit(ws, v, nmol);
        kind = uni(choose_move(f, randmov, true, false, false));
      }
      if (nmol == 0 && kind != FMC_MOVE_INSERT) kind = 0;  // nothing to move: an empty step
    }
  }
  const int mol = P.locguest[g], natms = P.natoms[mol], nmols = uni(ws.nmols[mol]);
  if (kind == 0) {
    if (lane == 0) {
      o.ok = 1;
      o.wlmode = wlmode;
      o.mkind = 0;
      o.swap_phase = 0;
      o.kind = 0;
      o.g = g;
      o.mol = mol;
      o.natms = natms;
      o.nmols = nmols;
      o.choice = 0;
      o.has_new = o.has_old = 0;
      o.excl_lo = o.excl_hi = 0;
      o.cnt = o.cval = o.chg = 0;
      o.used0 = used0;
      o.used1 = used0 + nu;
    }
    return;
  }
  // step-size adaptation of the grand-canonical loop (the Wang-Landau loop has
  // none), gcmc.f:859-900: all counters read first, then the new sizes
  double delrdisp = ws.delrdisp[g], rotangle = ws.rotangle;
  double tran_delr = ws.tran_delr[g], rota_rotangle = ws.rota_rotangle;
  const int dcnt = uni(ws.disp_count[g]), tcnt = uni(ws.tran_count[g]), rcnt = uni(ws.rota_count[g]);
  const int dacc = uni(ws.accept_disp[g]), tacc = uni(ws.accept_tran[g]), racc = uni(ws.accept_rota[g]);
  int chg = 0, cnt = 0, cval = 0;
  if (!wl_on && nmols >= 1) {
    bool chd = false, chr = false, cht = false, chq = false;
    if (__builtin_expect(dcnt >= 1 && (dcnt % 50 == 0), 0)) {
      if (dcnt % 100 == 0) {
        if ((dacc / (double)dcnt) > 0.5) delrdisp *= 1.05, chd = true;
        else if (delrdisp > DELRMIN) delrdisp *= 0.95, chd = true;
      } else if (dcnt % 50 == 0) {
        if ((dacc / (double)dcnt) > 0.5) rotangle *= 1.05, chr = true;
        else if (rotangle > MINANGLE) rotangle *= 0.95, chr = true;
      }
    }
    if (__builtin_expect(tcnt >= 1 && tcnt % 100 == 0, 0)) {
      if ((tacc / (double)tcnt) > 0.5) tran_delr *= 1.05, cht = true;
      else if (tran_delr > DELRMIN) tran_delr *= 0.95, cht = true;
    }
    if (__builtin_expect(rcnt >= 1 && rcnt % 100 == 0, 0)) {
      if ((racc / (double)rcnt) > 0.5) rota_rotangle *= 1.05, chq = true;
      else if (rota_rotangle > MINANGLE) rota_rotangle *= 0.95, chq = true;
    }
    chg = (int)chd | ((int)chr << 1) | ((int)cht << 2) | ((int)chq << 3);
  }
  if (wlmode == 0 && nmols == 0) kind = FMC_MOVE_INSERT;
  const int mkind = kind;
  if (kind == FMC_MOVE_SWITCH) {
    // mc_switch (gcmc.f:2041-2267) needs two other guest kinds with molecules
    // present (nswitchgst >= 2, gcmc.f:2098); set-up refuses such systems,
    // so a switch is a step without a move (and without a count)
    kind = 0;
  }
  if (kind == FMC_MOVE_SWAP) kind = FMC_MOVE_DELETE;  // phase 1: provisional deletion

  const bool site = lane < natms;
  const int i = site ? lane : 0;
  int choice = 0;
  if (kind == 0) {
    // nothing to evaluate
  } else if (kind == FMC_MOVE_INSERT) {
    // random_ins + rotation of the COM-centred template (utility_pack.f:3220-3331, 3480-3524)
    const double ra = draw(), rb = draw(), rc = draw();
    const double r1 = draw(), r2 = draw(), r3 = draw(), r4 = draw();
    const double *cl = sh.cl9;
    const double xc = ra * cl[0] + rb * cl[3] + rc * cl[6];
    const double yc = ra * cl[1] + rb * cl[4] + rc * cl[7];
    const double zc = ra * cl[2] + rb * cl[5] + rc * cl[8];
    double q[4], Q[9];
    quat_from(r1, r2, r3, r4, 2.0 * PI, q);
    rot_matrix(q, Q);
    const double a = sh.tmpl[g][i][0], b = sh.tmpl[g][i][1], d = sh.tmpl[g][i][2];
    const double tx = Q[0] * a + Q[1] * b + Q[2] * d;
    const double ty = Q[3] * a + Q[4] * b + Q[5] * d;
    const double tz = Q[6] * a + Q[7] * b + Q[8] * d;
    if (site) {
      o.xn[i][0] = tx + xc;
      o.xn[i][1] = ty + yc;
      o.xn[i][2] = tz + zc;
    }
  } else if (kind == FMC_MOVE_DELETE) {
    // mc_delete; or mc_swap's choice of molecule (floor(duni*n + 1)) and of
    // the other guest kind (gcmc.f:2311-2322)
    choice = uni(mkind == FMC_MOVE_SWAP ? (int)floor(draw() * nmols + 1) : (int)floor(draw() * nmols) + 1);
    const int a = P.gbase[g] + (choice - 1) * natms + i;
    const double x = Gx[a], y = Gy[a], z = Gz[a];
    if (site) {
      o.xo[i][0] = x;
      o.xo[i][1] = y;
      o.xo[i][2] = z;
    }
    if (__builtin_expect(mkind == FMC_MOVE_SWAP && !spec, 0)) {
      double wsum = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
      for (int k = 0; k < natms; ++k) {
        const double m = sh.sm[mol][k];
        wsum += m;
        cx += rdlane(x, k) * m;
        cy += rdlane(y, k) * m;
        cz += rdlane(z, k) * m;
      }
      int jg = uni((int)floor(draw() * (P.ntpguest - 1) + 1));
      jg = jg >= g ? jg + 1 : jg;  // switch_chosen_guest: the guest kinds other than g, in order
      if (lane == 0) {
        sh.swap_com[0] = cx / wsum;
        sh.swap_com[1] = cy / wsum;
        sh.swap_com[2] = cz / wsum;
        sh.swap_ig = g;
        sh.swap_jg = jg;
        sh.swap_ichoice = choice;
        sh.swap_n0 = nmols;
      }
    }
  } else {
    // displacement, translation, rotation, jump.  A translation keeps q = 0,
    // which rotation turns into the inversion through the COM, as the
    // reference's mc_displace does (gcmc.f:1845,2875)
    double a = 0.0, b = 0.0, c = 0.0, q[4] = {0.0, 0.0, 0.0, 0.0};
    double ox, oy, oz;
    if (kind == FMC_MOVE_JUMP) {
      // mc_jump (gcmc.f:1937-2039): molecule first, then random_jump (a point
      // of the cell) and random_rot(2 pi); the displacement that takes the
      // old COM there
      choice = uni((int)floor(draw() * nmols) + 1);
      const int at = P.gbase[g] + (choice - 1) * natms + i;
      ox = Gx[at];
      oy = Gy[at];
      oz = Gz[at];
      double wsum = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
      for (int k = 0; k < natms; ++k) {
        const double m = sh.sm[mol][k];
        wsum += m;
        cx += rdlane(ox, k) * m;
        cy += rdlane(oy, k) * m;
        cz += rdlane(oz, k) * m;
      }
      cx = cx / wsum;
      cy = cy / wsum;
      cz = cz / wsum;
      const double ra = draw(), rb = draw(), rc = draw();
      const double *cl = sh.cl9;
      const double xc = ra * cl[0] + rb * cl[3] + rc * cl[6];
      const double yc = ra * cl[1] + rb * cl[4] + rc * cl[7];
      const double zc = ra * cl[2] + rb * cl[5] + rc * cl[8];
      const double r1 = draw(), r2 = draw(), r3 = draw(), r4 = draw();
      quat_from(r1, r2, r3, r4, 2.0 * PI, q);
      a = xc - cx;
      b = yc - cy;
      c = zc - cz;
    } else {
      if (kind == FMC_MOVE_TRANSLATE) {
        const double h = tran_delr * 0.5;
        a = (2.0 * draw() - 1.0) * h;
        b = (2.0 * draw() - 1.0) * h;
        c = (2.0 * draw() - 1.0) * h;
        cnt = 2;
        cval = tcnt + 1;
      } else if (kind == FMC_MOVE_ROTATE) {
        const double r1 = draw(), r2 = draw(), r3 = draw(), r4 = draw();
        quat_from(r1, r2, r3, r4, rota_rotangle, q);
        cnt = 3;
        cval = rcnt + 1;
      } else {
        const double h = delrdisp * 0.5;
        const double da = draw(), db = draw(), dc = draw();
        a = (2.0 * da - 1.0) * h;
        b = (2.0 * db - 1.0) * h;
        c = (2.0 * dc - 1.0) * h;
        const double r1 = draw(), r2 = draw(), r3 = draw(), r4 = draw();
        quat_from(r1, r2, r3, r4, rotangle, q);
        cnt = 1;
        cval = dcnt + 1;
      }
      choice = uni((int)floor(draw() * nmols) + 1);
      const int at = P.gbase[g] + (choice - 1) * natms + i;
      ox = Gx[at];
      oy = Gy[at];
      oz = Gz[at];
    }
    // translate, utility_pack.f:3161-3218: move, then wrap the COM into the cell
    double nx = a + ox, ny = b + oy, nz = c + oz;
    double wsum = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    for (int k = 0; k < natms; ++k) {
      const double m = sh.sm[mol][k];
      wsum += m;
      cx += rdlane(nx, k) * m;
      cy += rdlane(ny, k) * m;
      cz += rdlane(nz, k) * m;
    }
    cx = cx / wsum;
    cy = cy / wsum;
    cz = cz / wsum;
    const double *rcl = sh.rc9, *cl = sh.cl9;
    const double ssx = rcl[0] * cx + rcl[3] * cy + rcl[6] * cz;
    const double ssy = rcl[1] * cx + rcl[4] * cy + rcl[7] * cz;
    const double ssz = rcl[2] * cx + rcl[5] * cy + rcl[8] * cz;
    const double xss = ssx - floor(ssx), yss = ssy - floor(ssy), zss = ssz - floor(ssz);
    const double wx = cl[0] * xss + cl[3] * yss + cl[6] * zss;
    const double wy = cl[1] * xss + cl[4] * yss + cl[7] * zss;
    const double wz = cl[2] * xss + cl[5] * yss + cl[8] * zss;
    nx += wx - cx;
    ny += wy - cy;
    nz += wz - cz;
    // rotation about the wrapped COM, utility_pack.f:3480-3524
    double Q[9];
    rot_matrix(q, Q);
    const double dx = nx - wx, dy = ny - wy, dz = nz - wz;
    const double tx = Q[0] * dx + Q[1] * dy + Q[2] * dz;
    const double ty = Q[3] * dx + Q[4] * dy + Q[5] * dz;
    const double tz = Q[6] * dx + Q[7] * dy + Q[8] * dz;
    if (site) {
      o.xo[i][0] = ox;
      o.xo[i][1] = oy;
      o.xo[i][2] = oz;
      o.xn[i][0] = tx + wx;
      o.xn[i][1] = ty + wy;
      o.xn[i][2] = tz + wz;
    }
  }
  if (lane == 0) {
    o.ok = mkind != FMC_MOVE_SWAP;
    o.wlmode = wlmode;
    o.mkind = mkind;
    o.swap_phase = mkind == FMC_MOVE_SWAP ? 1 : 0;
    o.kind = kind;
    o.g = g;
    o.mol = mol;
    o.natms = natms;
    o.nmols = nmols;
    o.choice = choice;
    o.has_new = kind != 0 && kind != FMC_MOVE_DELETE;
    o.has_old = kind != 0 && kind != FMC_MOVE_INSERT;
    o.excl_lo = kind == 0 || kind == FMC_MOVE_INSERT ? 0 : (choice - 1) * natms;
    o.excl_hi = kind == 0 || kind == FMC_MOVE_INSERT ? 0 : choice * natms;
    o.cnt = cnt;
    o.cval = cval;
    o.chg = chg;
    o.delrdisp = delrdisp;
    o.rotangle = rotangle;
    o.tran_delr = tran_delr;
    o.rota_rotangle = rota_rotangle;
    o.used0 = used0;
    o.used1 = used0 + nu;
  }
}

// Makes record o the step's live proposal (wave 0): the trial sites (all
// slots, one per lane), the move, and the walker-state updates that come
// with it.
template <int NS>
__device__ inline void commit(Smem<NS> &sh, const Prop &o) {
  const int lane = threadIdx.x & 63;
  static_assert(2 * KSITE * 3 <= 64, "trial sites exceed one wave");
  if (lane < 2 * KSITE * 3) {
    const double *src = lane < KSITE * 3 ? &o.xn[0][0] + lane : &o.xo[0][0] + (lane - KSITE * 3);
    double *dst = lane < KSITE * 3 ? &sh.xn[0][0] + lane : &sh.xo[0][0] + (lane - KSITE * 3);
    *dst = *src;
  }
  if (lane == 0) {
    WState &ws = sh.ws;
    const int g = o.g;
    sh.wlmode = o.wlmode;
    sh.mkind = o.mkind;
    sh.swap_phase = o.swap_phase;
    sh.kind = o.kind;
    sh.g = g;
    sh.mol = o.mol;
    sh.natms = o.natms;
    sh.nmols = o.nmols;
    sh.choice = o.choice;
    sh.has_new = o.has_new;
    sh.has_old = o.has_old;
    sh.excl_lo = o.excl_lo;
    sh.excl_hi = o.excl_hi;
    if (o.cnt == 1) ws.disp_count[g] = o.cval;
    else if (o.cnt == 2) ws.tran_count[g] = o.cval;
    else if (o.cnt == 3) ws.rota_count[g] = o.cval;
    const int chg = o.chg;
    if (chg & 1) ws.delrdisp[g] = o.delrdisp;
    if (chg & 2) ws.rotangle = o.rotangle;
    if (chg & 4) ws.tran_delr[g] = o.tran_delr;
    if (chg & 8) ws.rota_rotangle = o.rota_rotangle;
    ws.mz_used = o.used1;
    sh.mz_after = o.used1;
  }
}

// mc_swap's insertion (gcmc.f:2360-2375): a guest swap_jg molecule, its
// template placed at the deleted molecule's COM and rotated about it by
// random_rot with the displacement's current angle (rotangle), as the
// reference passes it.  Wave 0, after the provisional deletion.
template <int NS>
__device__ void swap_insert_setup(const DevParams &P, Smem<NS> &sh) {
  WState &ws = sh.ws;
  const int lane = threadIdx.x & 63;
  const int jg = uni(sh.swap_jg), jmol = P.locguest[jg], jnatms = P.natoms[jmol];
  const long long used0 = ws.mz_used;
  const int rbits = __float_as_int(ws.rq[(int)((used0 + lane) & 127)]);
  double r[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) r[k] = (double)__int_as_float(__builtin_amdgcn_readlane(rbits, k));
  double q[4], Q[9];
  quat_from(r[0], r[1], r[2], r[3], ws.rotangle, q);
  rot_matrix(q, Q);
  const double cx = sh.swap_com[0], cy = sh.swap_com[1], cz = sh.swap_com[2];
  const int i = lane < jnatms ? lane : 0;
  // newx = guestx + com, then rotation subtracts the COM again
  const double dx = (sh.tmpl[jg][i][0] + cx) - cx, dy = (sh.tmpl[jg][i][1] + cy) - cy,
               dz = (sh.tmpl[jg][i][2] + cz) - cz;
  const double tx = Q[0] * dx + Q[1] * dy + Q[2] * dz;
  const double ty = Q[3] * dx + Q[4] * dy + Q[5] * dz;
  const double tz = Q[6] * dx + Q[7] * dy + Q[8] * dz;
  if (lane < jnatms) {
    sh.xn[lane][0] = tx + cx;
    sh.xn[lane][1] = ty + cy;
    sh.xn[lane][2] = tz + cz;
  }
  if (lane == 0) {
    ws.mz_used = used0 + 4;
    sh.swap_phase = 2;
    sh.kind = FMC_MOVE_INSERT;
    sh.g = jg;
    sh.mol = jmol;
    sh.natms = jnatms;
    sh.nmols = ws.nmols[jmol];
    sh.choice = 0;
    sh.has_new = 1;
    sh.has_old = 0;
    sh.excl_lo = sh.excl_hi = 0;
  }
}

// ----------------------------------------------------------- evaluation --
// Where a pair's LJ table value comes from: FP64 mode reads v64 (column
// stride s64) everywhere; mixed mode reads FP32 values from v32 (stride s32,
// whole columns) and, for the wall entries [lo, hi), FP64 values from v64
// (table index b64 at v64[0]).
struct LJTab {
  const double *v64;
  const float *v32;
  int s64, s32, b64, lo, hi;
};

__device__ inline LJTab lj_global(const DevParams &P, bool mixed) {
  LJTab t;
  t.v64 = P.vvv;
  t.v32 = P.vvv32;
  t.s64 = t.s32 = P.vstride;
  t.b64 = 0;
  t.lo = mixed ? P.lj_lo : 0;
  t.hi = mixed ? P.lj_hi : 0x7fffffff;
  return t;
}

// The LDS copy (stage_tables): FP64 mode the whole columns at off_vvv64;
// mixed mode the wall entries at off_vvv64 and the whole FP32 columns at off_vvv.
template <typename T>
__device__ inline LJTab lj_lds(const DevParams &P, const unsigned char *lds) {
  LJTab t;
  t.v64 = (const double *)(lds + P.off_vvv64);
  t.v32 = (const float *)(lds + P.off_vvv);
  t.s32 = P.vstride;
  if constexpr (sizeof(T) == 8) {
    t.s64 = P.vstride;
    t.b64 = 0;
  } else {
    t.s64 = P.lj_n64;
    t.b64 = P.lj_lo;
  }
  t.lo = P.lj_lo;
  t.hi = P.lj_hi;
  return t;
}

// One guest-site / atom pair of insertion / guest_energy / deletion: minimum
// image, surface and overlap tests (surface_check, mc_moves.f:10-44: bit 0
// of ovl rsq < overlap^2, bit 1 rsq < sg2 = (surftol + sigma)^2 for a frozen
// atom, sg2 < 0 otherwise), ewald2 and srfrce terms (ewald.f:689-726,
// vdw_terms.f:418-460).  Branch-free, so that the pairs of one atom with all
// trial sites form one straight-line block: table indices are clamped and
// out-of-range terms selected away.  Geometry and the real-space Coulomb term
// are FP64 in both modes (pair Coulomb terms are large and cancel).  T is the
// precision of the LJ interpolation away from the repulsive wall, whose
// fractional grid position is formed in FP64 first; on the wall (mixed mode's
// LJTab segment [lo, hi)) the LJ term is interpolated in FP64 from FP64
// values.  col < 0: no LJ term; ovl2 < 0: a padding site (never overlaps,
// chgea 0).
template <typename T>
__device__ inline void pair_eval(const DevParams &P, const double *erc, const LJTab &lj, const double *xs, double chgea,
                                 int col, double ovl2, double sg2, double xj, double yj, double zj, double qj,
                                 double &e2, double &ev, int &ovl) {
  double dx = xs[0] - xj, dy = xs[1] - yj, dz = xs[2] - zj;
  min_image(P, dx, dy, dz);
  double rsq = dx * dx + dy * dy + dz * dz;
  ovl |= (int)(rsq < ovl2) | ((int)(rsq < sg2) << 1);
  bool in = rsq < P.rcut2;
  double r = pair_sqrt(rsq);
  const int lmax = P.mxegrd;
  double x = r * P.rdrewd;
  int l = min((int)x, lmax);
  double p = x - (double)l;
  double v0 = erc[l], v1 = erc[l + 1], v2 = erc[l + 2];
  double t1 = v0 + (v1 - v0) * p;
  double t2 = v1 + (v2 - v1) * (p - 1.0);
  double chgprd = chgea * qj;
  double ec = (t1 + (t2 - t1) * p * 0.5) * chgprd;
  e2 = (in && fabs(chgprd) > 1e-10) ? ec : 0.0;
  double xv = r * P.rdlrpot;
  int lv = min((int)xv, lmax);
  const int c = col < 0 ? 0 : col;
  double el;
  if constexpr (sizeof(T) == 8) {
    const double pp = xv - (double)lv;
    const double *vv = lj.v64 + (size_t)c * lj.s64 + lv;
    const double w0 = vv[0], w1 = vv[1], w2 = vv[2];
    const double u1 = w0 + (w1 - w0) * pp;
    const double u2 = w1 + (w2 - w1) * (pp - 1.0);
    el = u1 + (u2 - u1) * pp * 0.5;
  } else {
    const float pp = (float)(xv - (double)lv);
    const float *vv = lj.v32 + (size_t)c * lj.s32 + lv;
    const float w0 = vv[0], w1 = vv[1], w2 = vv[2];
    const float u1 = w0 + (w1 - w0) * pp;
    const float u2 = w1 + (w2 - w1) * (pp - 1.0f);
    el = (double)(u1 + (u2 - u1) * pp * 0.5f);
    // the repulsive wall in FP64, read only by a wave with a lane on it
    // (close contacts are rare, so most waves skip it)
    const bool wall = lv >= lj.lo && lv < lj.hi;
    if (__any(wall)) {
      const double qq = xv - (double)lv;
      const double *a = lj.v64 + (size_t)c * lj.s64 + (min(max(lv, lj.lo), lj.hi - 1) - lj.b64);
      const double a0 = a[0], a1 = a[1], a2 = a[2];
      const double t1 = a0 + (a1 - a0) * qq;
      const double t2 = a1 + (a2 - a1) * (qq - 1.0);
      const double e64 = t1 + (t2 - t1) * qq * 0.5;
      el = wall ? e64 : el;
    }
  }
  ev = (in && col >= 0) ? el : 0.0;
}

// LDS block of per-launch constants (offsets from P.off_cst):
//   ltab [KMAXS+1][KSITE][KTYPE+1] int8: guest g, site s, atom type -> LJ column or -1
//   lsq  [KMAXS+1][KSITE] double: site charge by species
//   lst  [KMAXS+1][KSITE] int: site type by species
//   srf  [KMAXS+1][KSITE][KTYPE+1] double: guest g, site s, atom type ->
//        surface_check threshold (surftol + sigma)^2, -1 for padding sites
constexpr int CST_LTAB = 0;
constexpr int CST_LSQ = 1008;
constexpr int CST_LST = CST_LSQ + 8 * (KMAXS + 1) * KSITE;
constexpr int CST_SRF = CST_LST + 4 * (KMAXS + 1) * KSITE;
constexpr int CST_BYTES = CST_SRF + 8 * (KMAXS + 1) * KSITE * (KTYPE + 1);
static_assert(CST_SRF % 8 == 0, "srf alignment");
static_assert((KMAXS + 1) * KSITE * (KTYPE + 1) <= CST_LSQ, "ltab overlaps lsq");

// The same rows rounded to FP32 once when built (mixed mode's low-weight
// k-vectors: the products are formed in FP32 from these, the same values
// the FP64 rows convert to, without a conversion per use).
__device__ inline float2 *ph_site32(unsigned char *lds, const DevParams &P, int cfg, int s) {
  return (float2 *)(lds + P.off_ph32) + (size_t)((cfg * P.nsmax + s) * 3) * P.kcp;
}

__device__ inline double2 *ph_site(unsigned char *lds, const DevParams &P, int cfg, int s) {
  return (double2 *)(lds + P.off_ph) + (size_t)((cfg * P.nsmax + s) * 3) * P.kcp;
}

// One reciprocal-space phase product for one site at (ll, mm, nn): the
// reference's clm/slm then ckc/cks (ewald.f:191-251).  ph holds the site's
// three axis rows at stride kcp.
template <typename T>
__device__ inline void site_phase(const double2 *ph, int kcp, int ll, int mm, int nn, T q, T &ckc, T &cks) {
  int m = mm < 0 ? -mm : mm, n = nn < 0 ? -nn : nn;
  T lc = (T)ph[ll].x, ls = (T)ph[ll].y;
  T mc = (T)ph[kcp + m].x, ms = (T)ph[kcp + m].y;
  T nc = (T)ph[2 * kcp + n].x, ns = (T)ph[2 * kcp + n].y;
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

// The same product from FP32 rows (mixed mode's low-weight k-vectors).
__device__ inline void site_phase(const float2 *ph, int kcp, int ll, int mm, int nn, float q, float &ckc,
                                  float &cks) {
  const int m = mm < 0 ? -mm : mm, n = nn < 0 ? -nn : nn;
  const float2 L = ph[ll], M = ph[kcp + m], N = ph[2 * kcp + n];
  float clm, slm;
  if (mm >= 0) {
    clm = L.x * M.x - L.y * M.y;
    slm = L.y * M.x + M.y * L.x;
  } else {
    clm = L.x * M.x + L.y * M.y;
    slm = L.y * M.x - M.y * L.x;
  }
  if (nn >= 0) {
    ckc = q * (clm * N.x - slm * N.y);
    cks = q * (slm * N.x + clm * N.y);
  } else {
    ckc = q * (clm * N.x + slm * N.y);
    cks = q * (slm * N.x - clm * N.y);
  }
}

// Copies the read-only per-launch data into LDS: the erfc table and the
// constants block always; with LF (the plan made room for everything) also
// the walker's framework, the k list, the LJ columns of the guest sites (in
// the precision T of the mode) and the guest coordinates.
// LF: framework, k list and guest coordinates into LDS; LJL: the LJ columns
// (either precision).
template <typename T, bool LF, bool LJL>
__device__ inline void stage_tables(const DevParams &P, const DevWalkers &W, int w, unsigned char *lds) {
  constexpr bool LV = LJL;
  const int tid = threadIdx.x, nt = blockDim.x;
  double *erc = (double *)(lds + P.off_erc);
  for (int i = tid; i < P.mxegrd + 3; i += nt) erc[i] = P.erc[i];
  {
    int8_t *lt = (int8_t *)(lds + P.off_cst + CST_LTAB);
    double *lsq = (double *)(lds + P.off_cst + CST_LSQ);
    int *lst = (int *)(lds + P.off_cst + CST_LST);
    double *srf = (double *)(lds + P.off_cst + CST_SRF);
    const int per = KSITE * (KTYPE + 1);
    for (int i = tid; i < (KMAXS + 1) * per; i += nt) {
      int g = i / per, s = (i % per) / (KTYPE + 1), tj = i % (KTYPE + 1);
      int col = -1;
      double sg2 = -1.0;
      if (g >= 1 && g <= P.ntpguest) {
        int mol = P.locguest[g];
        if (s < P.natoms[mol]) {
          int ts = P.site_t[mol][s];
          col = LV ? P.lpairidx[ts][tj] : P.pairidx[ts][tj];
          sg2 = P.surf2[ts][tj];
        }
      }
      lt[i] = (int8_t)col;
      srf[i] = sg2;
    }
    for (int i = tid; i < (KMAXS + 1) * KSITE; i += nt) {
      lsq[i] = P.site_q[i / KSITE][i % KSITE];
      lst[i] = P.site_t[i / KSITE][i % KSITE];
    }
  }
  if constexpr (LV) {
    if constexpr (sizeof(T) == 4) {
      // LJ columns: whole in FP32, and the wall entries again in FP64 (LJTab)
      double *d64 = (double *)(lds + P.off_vvv64);
      const int n64 = P.nvcol * P.lj_n64;
      for (int i = tid; i < n64; i += nt) {
        int c = i / P.lj_n64, r = i - c * P.lj_n64;
        d64[i] = P.vvv[(size_t)P.vcol_src[c] * P.vstride + P.lj_lo + r];
      }
      float *d32 = (float *)(lds + P.off_vvv);
      const int n32 = P.nvcol * P.vstride;
      for (int i = tid; i < n32; i += nt) {
        int c = i / P.vstride, r = i - c * P.vstride;
        d32[i] = P.vvv32[(size_t)P.vcol_src[c] * P.vstride + r];
      }
    } else {
      double *d64 = (double *)(lds + P.off_vvv64);
      const int n64 = P.nvcol * P.vstride;
      for (int i = tid; i < n64; i += nt) {
        int c = i / P.vstride, r = i - c * P.vstride;
        d64[i] = P.vvv[(size_t)P.vcol_src[c] * P.vstride + r];
      }
    }
  }
  if constexpr (LF) {
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
    double *ak = (double *)(lds + P.off_k);
    int *kl = (int *)(ak + P.nk);
    for (int i = tid; i < P.nk; i += nt) {
      ak[i] = P.akk[i];
      kl[i] = P.klmn[i];
    }
    double *lx = (double *)(lds + P.off_g), *ly = lx + P.gatoms_total, *lz = ly + P.gatoms_total;
    const size_t b = (size_t)w * P.gatoms_total;
    for (int gg = 1; gg <= P.ntpguest; ++gg) {
      int mol = P.locguest[gg];
      int n = W.ws[w].nmols[mol] * P.natoms[mol], o = P.gbase[gg];
      for (int i = tid; i < n; i += nt) {
        lx[o + i] = W.gx[b + o + i];
        ly[o + i] = W.gy[b + o + i];
        lz[o + i] = W.gz[b + o + i];
      }
    }
  }
}

// Trial constants that depend only on the proposal: the tail correction
// (gstlrcorrect), the charged-system terms of ewald1_guest (ewald.f:96-105)
// and the acceptance prefactor of energy_eval (utility_pack.f:3655-3676).
// One thread, overlapped with the energy terms.
template <int NS>
__device__ void prepare(const DevParams &P, Smem<NS> &sh) {
  const WState &ws = sh.ws;
  const int g = uni(sh.g), mol = uni(sh.mol), kind = uni(sh.kind);
  const double alpha = P.alpha, volm = P.volm;
  const double chg = P.chgtmp[g];
  sh.qfix0 = -(0.5 * PI * R4PIE0 / P.epsq) * ((chg * ws.sumchg) / alpha + (chg / alpha) * (chg / alpha)) / volm;
  for (int i = 1; i <= NS; ++i) {
    double fkk = (i == mol) ? 0.5 : 1.0;
    sh.qfm[i] = -(fkk * PI * R4PIE0 / P.epsq) * (ws.chgsum_mol[i] * chg) / (alpha * alpha * volm);
  }
  sh.delrc = 0.0;
  if (kind == FMC_MOVE_INSERT) sh.delrc = gstlrcorrect<NS>(P, ws, g, 0, sh.delrc_mol);
  else if (kind == FMC_MOVE_DELETE) sh.delrc = gstlrcorrect<NS>(P, ws, g, 1, sh.delrc_mol);
  const int ingsts = ws.nmols[mol];
  const double ipress = ws.fug[g];
  if (kind == FMC_MOVE_INSERT) sh.pref = P.statvolm * ipress / (((double)ingsts + 1.0) * BOLTZ * ws.temp);
  else if (kind == FMC_MOVE_DELETE) sh.pref = (double)ingsts * BOLTZ * ws.temp / (P.statvolm * ipress);
  else sh.pref = 1.0;
}

// The team's sums of one step (lanes < NRED of wave 0, v = this member's
// sum of row `lane`): each member publishes its row sums in its record of
// the exchange's parity and gathers the other members' rows; the rows are
// then added in member order, so that every member holds the same bits.
//
// Each value carries its exchange in its two lowest mantissa bits (bit 1
// set, bit 0 = bit 1 of the exchange number x; the record of x's parity
// last held x - 2, whose bit differs, and a cleared record has bit 1 clear),
// and those bits are cleared before the value is used, by every member for
// every row, its own included (a relative change below 2^-50).  A value is
// then complete in itself: members store their rows with relaxed
// agent-scope 64-bit atomics, readers poll each row until its tag is the
// current one (bounded: XSPIN polls, then the walker is flagged and the
// members leave their loops), and no fence, store completion wait or
// separate flag orders the two.  The host clears the records whenever the
// team changes, so that a member's record holds only this team's exchanges.
// The gathered rows are staged in sh.red, which wave 0 has finished reading.
template <int NS>
__device__ __forceinline__ double team_exchange(const DevWalkers &W, int w, Smem<NS> &sh, double v, int lane, int team,
                                                int member) {
  constexpr int NR = Acc<NS>::NRED;
  static_assert(NR <= XREC, "exchange record too short");
  static_assert(KTEAM <= KWAVES, "exchange rows staged in sh.red");
  const long long x = sh.xseq + 1;
  const long long tag = 2 | ((x >> 1) & 1);
  long long *recs = (long long *)W.xch + ((size_t)w * 2 + (size_t)(x & 1)) * KTEAM * XREC;
  if (lane < NR) {
    const long long b = (__double_as_longlong(v) & ~3LL) | tag;
    __hip_atomic_store(recs + (size_t)member * XREC + lane, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    sh.red[member][lane] = __longlong_as_double(b & ~3LL);
  }
  int fail = 0;
  if (team * NR <= 64) {
    // one row per lane: lane q * NR + row polls member q's row
    const int q = lane / NR, row = lane - q * NR;
    if (lane < team * NR && q != member) {
      const long long *src = recs + (size_t)q * XREC + row;
      long long b, spin = 0;
      while (((b = __hip_atomic_load(src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) & 3) != tag) {
        if (__builtin_expect(++spin > XSPIN, 0)) {
          fail = 1;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      sh.red[q][row] = __longlong_as_double(b & ~3LL);
    }
    fail = __any(fail);
  } else {
    // lane takes the rows lane, lane + 64, ... of the other members (RMAX at
    // most); each poll reads every row it still lacks at once, so that a
    // large team costs one round trip per poll, not one per row of a lane
    constexpr int RMAX = (KTEAM * NR + 63) / 64;
    static_assert(RMAX <= 31, "row mask");
    int need = 0;
#pragma unroll
    for (int r = 0; r < RMAX; ++r) {
      const int i = lane + 64 * r, q = i / NR;
      if (i < team * NR && q != member) need |= 1 << r;
    }
    long long spin = 0;
    while (true) {
      long long got[RMAX];
#pragma unroll
      for (int r = 0; r < RMAX; ++r)
        if ((need >> r) & 1) {
          const int i = lane + 64 * r, q = i / NR, row = i - q * NR;
          got[r] = __hip_atomic_load(recs + (size_t)q * XREC + row, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
#pragma unroll
      for (int r = 0; r < RMAX; ++r)
        if (((need >> r) & 1) && (got[r] & 3) == tag) {
          const int i = lane + 64 * r, q = i / NR, row = i - q * NR;
          sh.red[q][row] = __longlong_as_double(got[r] & ~3LL);
          need &= ~(1 << r);
        }
      if (!__any(need)) break;
      if (__builtin_expect(++spin > XSPIN, 0)) {
        fail = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __builtin_amdgcn_wave_barrier();  // the rows other lanes staged are read below
  double t = 0.0;
  if (lane < NR)
    for (int q = 0; q < team; ++q) t += sh.red[q][lane];
  if (lane == 0) {
    sh.xseq = x;
    if (__builtin_expect(fail, 0)) {
      sh.team_fail = 1;
      sh.ws.overflow |= FMC_OVF_TEAM;
      *W.xerr = 1;
    }
  }
  return t;
}

// Trial energy terms of insertion / deletion / guest_energy (mc_moves.f:150-484)
// with ewald2 (ewald.f:663-733), srfrce (vdw_terms.f:382-466) and ewald1_guest
// (ewald.f:32-352).  Every thread of the workgroup takes part: atoms of the
// pair terms from the bottom thread up, k-vectors from the top down (register
// slots ks, then any k beyond KPT per thread through global memory); the
// reduced sums land in sh.tot.  LF: framework, k list and guests are read
// from LDS; LJL: the LJ columns are.  NSITE > 0: guests have at most NSITE
// sites (padding sites masked) and the site loops are straight-line code.
// seam: also write s_new / s_old of the register slots to the scratch rows.
// evn: the number of this evaluation in the launch (every thread counts them
// alike); the top wave publishes the phase tables under it.
template <int NS, typename T, bool LF, int NSITE, bool LJL, bool LJ32 = LJL>
__device__ __forceinline__ void evaluate(const DevParams &P, const DevWalkers &W, int w, Smem<NS> &sh, unsigned char *lds,
                         KSlot<NS> *ks, bool seam, const int evn) {
  using A = Acc<NS>;
  // LJ columns in LDS (stage_tables), in the precision T of the mode; read
  // from global memory (L2) they are FP64 in both modes: on gfx950 an FP32
  // interpolation issues at the FP64 rate and adds the conversions and the
  // wall test, so FP32 values pay only where they halve the LDS they take.
  // LJ32: the interpolation in T from the global FP32 columns (the seam
  // trial of a handle whose chains stage FP32 columns, same arithmetic)
  constexpr bool LV = LJL;
  using TP = typename std::conditional<LJ32, T, double>::type;
  constexpr int NSR = NSITE > 0 ? NSITE : KSITE;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int natms = uni(sh.natms), mol = uni(sh.mol), g = uni(sh.g);
  const int has_new = uni(sh.has_new), has_old = uni(sh.has_old);
  // The trial constants (prepare) are needed only by the decision and the
  // phase tables only by the reciprocal terms: both are built while the pair
  // terms run, and the barrier that publishes the tables sits before the
  // reciprocal loop.  The table rows are spread over the waves, a few lanes
  // each, so that no wave carries them all.
  const int lane = tid & 63, wave = tid >> 6, nwaves = (nt + 63) >> 6;
  // the top wave has the fewest pair atoms (they are dealt from thread 0 up),
  // so it builds the phase tables (at most 2 * KSITE * 3 = 48 rows, one lane
  // each) and the wave below it the trial constants
  if (tid == (nwaves >= 2 ? nt - 65 : nt - 1)) prepare<NS>(P, sh);
  if (wave == nwaves - 1) {
    constexpr int NSC = NSITE > 0 ? NSITE : 0;
    const int nsr = NSITE > 0 ? NSC : P.nsmax;
    const int ntask = 2 * nsr * 3;
    static_assert(2 * KSITE * 3 <= 64, "phase-table rows exceed one wave");
    const int t = lane;
    if (t < ntask) {
      // per-site 1-D phase tables, built by the reference recursion (ewald.f:108-147)
      const int cfg = t / (nsr * 3), rem = t - cfg * (nsr * 3), s = rem / 3, ax = rem - 3 * s;
      if (s < natms && (cfg == 0 ? has_new : has_old)) {
        const double *x = cfg == 0 ? sh.xn[s] : sh.xo[s];
        double fr = sh.rc9[ax] * x[0] + sh.rc9[3 + ax] * x[1] + sh.rc9[6 + ax] * x[2];
        int kmax = sh.kmx[ax];
        double s1, c1;
        sincospi(2.0 * fr, &s1, &c1);
        double2 *row = ph_site(lds, P, cfg, s) + ax * P.kcp;
        float2 *row32 = ph_site32(lds, P, cfg, s) + ax * P.kcp;
        row[0] = make_double2(1.0, 0.0);
        row[1] = make_double2(c1, s1);
        row32[0] = make_float2(1.0f, 0.0f);
        row32[1] = make_float2((float)c1, (float)s1);
        double c = c1, sn = s1;
        for (int l = 2; l <= kmax; ++l) {
          double cn = c * c1 - sn * s1;
          double snn = sn * c1 + c * s1;
          c = cn;
          sn = snn;
          row[l] = make_double2(c, sn);
          row32[l] = make_float2((float)c, (float)sn);
        }
      }
    }
    // the rows of every lane are written (this wave's LDS writes complete in
    // order): publish them under this evaluation's number
    if (lane == 0) __hip_atomic_store(&sh.tab_ready, evn, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  double acc[A::NACC];
#pragma unroll
  for (int i = 0; i < A::NACC; ++i) acc[i] = 0.0;
  int ovl_new = 0, ovl_old = 0;
  const double *erc = (const double *)(lds + P.off_erc);
  const int8_t *lt = (const int8_t *)(lds + P.off_cst + CST_LTAB) + g * KSITE * (KTYPE + 1);
  const double *srf = (const double *)(lds + P.off_cst + CST_SRF) + g * KSITE * (KTYPE + 1);
  const double *lsq = (const double *)(lds + P.off_cst + CST_LSQ);
  const int *lst = (const int *)(lds + P.off_cst + CST_LST);
  LJTab vvv;
  if constexpr (LV) vvv = lj_lds<T>(P, lds);
  else vvv = lj_global(P, LJ32 && sizeof(T) == 4);
  // NSITE > 0: site charges and overlap thresholds in registers; padding
  // sites (s >= natms) carry charge 0, no LJ column, no overlap or surface
  // test, so whatever their LDS coordinates hold is selected away.  The site
  // coordinates are read from LDS at each pair (same address on every lane,
  // a broadcast): held in registers they spill the scalar file and cost more.
  // NSITE == 0: everything from LDS.
  constexpr int NSA = NSITE > 0 ? NSITE : 1;
  double cq[NSA], ov2[NSA], qs[NSA];
  if constexpr (NSITE > 0) {
    const double ovl2 = sh.ovl2;
#pragma unroll
    for (int s = 0; s < NSITE; ++s) {
      const bool real = s < natms;
      const double c = sh.scq[mol][s], q = sh.sq[mol][s];  // loaded unconditionally, then selected
      cq[s] = real ? c : 0.0;
      ov2[s] = real ? ovl2 : -1.0;
      qs[s] = real ? q : 0.0;
    }
  }
  // atom index space: framework atoms, then the guests of each kind (per
  // guest kind: first index, species, sites, first slot; all uniform)
  int gstart[KMAXS + 1], gmol[KMAXS + 1], gna[KMAXS + 1], gslot[KMAXS + 1];
#pragma unroll
  for (int gg = 1; gg <= KMAXS; ++gg) {
    gstart[gg] = uni(sh.gstart[gg]);
    gmol[gg] = uni(sh.kmol[gg]);
    gna[gg] = uni(sh.kna[gg]);
    gslot[gg] = uni(sh.kslot[gg]);
  }
  const int ntot = uni(sh.ntot);
  const int excl_lo = uni(sh.excl_lo), excl_hi = uni(sh.excl_hi);
  const int team = uni(sh.team), member = uni(sh.member);
  const int k0 = uni(sh.k0), k1 = uni(sh.k1);
  // a team member takes the atoms [a0, a1) of the pair loop (gstart_from)
  const int a0 = uni(sh.a0), a1 = uni(sh.a1);
  const GuestXYZ<LF> G(P, W, w, lds, member);
  const double *fx, *fy, *fz, *fq;
  const int *fi;
  if constexpr (LF) {
    fx = (const double *)(lds + P.off_fw);
    fy = fx + P.nfw;
    fz = fy + P.nfw;
    fq = fz + P.nfw;
    fi = (const int *)(fq + P.nfw);
  } else {
    fx = P.fw_per_walker ? W.fwx + (size_t)w * P.nfw : P.fwx;
    fy = P.fw_per_walker ? W.fwy + (size_t)w * P.nfw : P.fwy;
    fz = P.fw_per_walker ? W.fwz + (size_t)w * P.nfw : P.fwz;
    fq = P.fwq;
    fi = P.fwinfo;
  }
  // ---- pair terms
  for (int j = a0 + tid; j < a1; j += nt) {
    double xj, yj, zj, qj;
    int tj, spj;
    bool skip = false, frz = false;
    if (j < P.nfw) {
      xj = fx[j];
      yj = fy[j];
      zj = fz[j];
      qj = fq[j];
      int inf = fi[j];
      tj = inf & 255;
      spj = (inf >> 8) & 255;
      frz = (inf >> 16) & 1;
    } else {
      // the guest kind holding index j, selected without lane-indexed loads
      int gg = 1, start = gstart[1], molg = gmol[1], na = gna[1], slot = gslot[1];
#pragma unroll
      for (int k = 2; k <= KMAXS; ++k)
        if (j >= gstart[k]) {
          gg = k;
          start = gstart[k];
          molg = gmol[k];
          na = gna[k];
          slot = gslot[k];
        }
      int jj = j - start;
      int site = jj % na;
      int a = slot + jj;
      xj = G.x[a];
      yj = G.y[a];
      zj = G.z[a];
      qj = lsq[molg * KSITE + site];
      tj = lst[molg * KSITE + site];
      spj = molg;
      skip = gg == g && jj >= excl_lo && jj < excl_hi;
    }
    if (skip) continue;
    int col[NSR];
    double sg2[NSR];
#pragma unroll
    for (int s = 0; s < NSR; ++s) {
      if (NSITE == 0 && s >= natms) break;
      col[s] = lt[s * (KTYPE + 1) + tj];
      sg2[s] = frz ? srf[s * (KTYPE + 1) + tj] : -1.0;
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      if (c == 0 ? !has_new : !has_old) continue;
      double er = 0.0, evd = 0.0;
      int ov = 0;
#pragma unroll
      for (int s = 0; s < NSR; ++s) {
        if (NSITE == 0 && s >= natms) break;
        double e2, ev;
        if constexpr (NSITE > 0)
          pair_eval<TP>(P, erc, vvv, c == 0 ? sh.xn[s] : sh.xo[s], cq[s], col[s], ov2[s], sg2[s], xj, yj, zj, qj, e2,
                       ev, ov);
        else
          pair_eval<TP>(P, erc, vvv, c == 0 ? sh.xn[s] : sh.xo[s], sh.scq[mol][s], col[s], sh.ovl2, sg2[s], xj, yj,
                       zj, qj, e2, ev, ov);
        er += e2;
        evd += ev;
      }
#pragma unroll
      for (int i = 0; i < NS; ++i) {
        acc[2 * c * NS + i] += (spj == i + 1) ? er : 0.0;
        acc[(2 * c + 1) * NS + i] += (spj == i + 1) ? evd : 0.0;
      }
      if (c == 0) ovl_new |= ov;
      else ovl_old |= ov;
    }
  }
  FMC_PHASE(3);
  // the reciprocal terms need the phase tables only: a wave waits for the
  // top wave's number rather than for every wave's pair terms, so that the
  // waves holding k-vectors overlap them with the pair terms of the others.
  // The top wave publishes unconditionally; the wait is bounded all the same
  // (XSPIN polls, then the walker is flagged as a team wait that gave up)
  if (wave != nwaves - 1) {
    long long spin = 0;
    while (__hip_atomic_load(&sh.tab_ready, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != evn) {
      if (__builtin_expect(++spin > XSPIN, 0)) {
        if (lane == 0) {
          atomicOr(&sh.ws.overflow, FMC_OVF_TEAM);
          sh.team_fail = 1;
          if (team > 1) *W.xerr = 1;
        }
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  FMC_PHASE(4);
  // ---- reciprocal terms
  const int *kl;
  const double *ak;
  if constexpr (LF) {
    ak = (const double *)(lds + P.off_k);
    kl = (const int *)(ak + P.nk);
  } else {
    ak = P.akk;
    kl = P.klmn;
  }
  double2 *scr = W.sscr + (size_t)w * 2 * P.nk;
  auto recip_k = [&](int k, const double2 *S, double2 &sn_out, double2 &so_out, auto tag) {
    using U = decltype(tag);
    const int lmn = kl[k];
    const double a = ak[k];
    const int ll = lmn & 255, mm = ((lmn >> 8) & 255) - 128, nn = ((lmn >> 16) & 255) - 128;
    const double2 St = S[NS];
    double soc = 0.0, sos = 0.0;
    if (has_old) {
      U c = (U)0, s_ = (U)0;
#pragma unroll
      for (int s = 0; s < NSR; ++s) {
        if (s >= natms) break;
        U ckc, cks;
        if constexpr (sizeof(U) == 4)
          site_phase(ph_site32(lds, P, 1, s), P.kcp, ll, mm, nn,
                        (U)(NSITE > 0 ? qs[s < NSA ? s : 0] : lsq[mol * KSITE + s]), ckc, cks);
        else
          site_phase<U>(ph_site(lds, P, 1, s), P.kcp, ll, mm, nn,
                        (U)(NSITE > 0 ? qs[s < NSA ? s : 0] : lsq[mol * KSITE + s]), ckc, cks);
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
        acc[5 * NS + 2 + i] += fkk * (soc * S[i].x + sos * S[i].y);
      }
      so_out = make_double2(soc, sos);
    }
    if (has_new) {
      U c = (U)0, s_ = (U)0;
#pragma unroll
      for (int s = 0; s < NSR; ++s) {
        if (s >= natms) break;
        U ckc, cks;
        if constexpr (sizeof(U) == 4)
          site_phase(ph_site32(lds, P, 0, s), P.kcp, ll, mm, nn,
                        (U)(NSITE > 0 ? qs[s < NSA ? s : 0] : lsq[mol * KSITE + s]), ckc, cks);
        else
          site_phase<U>(ph_site(lds, P, 0, s), P.kcp, ll, mm, nn,
                        (U)(NSITE > 0 ? qs[s < NSA ? s : 0] : lsq[mol * KSITE + s]), ckc, cks);
        c += ckc;
        s_ += cks;
      }
      double snc = (double)c, sns = (double)s_;
      double Stx = St.x, Sty = St.y;
      if (has_old) {
        Stx = Stx - soc;
        Sty = Sty - sos;
      }
      double mix = 2.0 * snc * Stx + 2.0 * sns * Sty;
      acc[4 * NS] += a * (mix + (snc * snc + sns * sns));
#pragma unroll
      for (int i = 0; i < NS; ++i) {
        double fkk = (i + 1 == mol) ? a : 2.0 * a;
        double sx = S[i].x, sy = S[i].y;
        if (has_old && i + 1 == mol) {
          sx = sx - soc;
          sy = sy - sos;
        }
        acc[4 * NS + 2 + i] += fkk * (snc * sx + sns * sy);
      }
      sn_out = make_double2(snc, sns);
    }
  };
#pragma unroll
  for (int m = 0; m < KPT; ++m) {
    const int k = kslot(k0, m);
    if (k < k1) {
      if (k < P.nk64) recip_k(k, ks[m].S, ks[m].sn, ks[m].so, 0.0);
      else recip_k(k, ks[m].S, ks[m].sn, ks[m].so, (T)0);
      if (seam) {
        scr[k] = ks[m].sn;
        scr[P.nk + k] = ks[m].so;
      }
    }
  }
  if (__builtin_expect(k1 - k0 > KPT * nt, 0)) {
    const double2 *skw = W.sk + (size_t)w * (NS + 1) * P.nk;
    for (int k = k0 + KPT * nt + tid; k < k1; k += nt) {
      double2 S[NS + 1], sn = make_double2(0.0, 0.0), so = make_double2(0.0, 0.0);
#pragma unroll
      for (int i = 0; i <= NS; ++i) S[i] = skw[(size_t)i * P.nk + k];
      if (k < P.nk64) recip_k(k, S, sn, so, 0.0);
      else recip_k(k, S, sn, so, (T)0);
      scr[k] = sn;
      scr[P.nk + k] = so;
    }
  }
  FMC_PHASE(5);
  // ---- workgroup reduction: within each wave by recursive halving (the
  // accumulators of an absent configuration hold zeros), then across waves
  {
    double ax[A::NRED];
#pragma unroll
    for (int i = 0; i < A::NACC; ++i) ax[i] = acc[i];
    // overlap (bit 0) and surface (bit 1) flags counted with the sums
    ax[A::NACC] = (double)((ovl_new & 1) + 1024 * ((ovl_new >> 1) & 1));
    ax[A::NACC + 1] = (double)((ovl_old & 1) + 1024 * ((ovl_old >> 1) & 1));
    wave_sum_rows<A::NRED>(ax, lane, sh.red[wave]);
  }
  FMC_PHASE(6);
  __syncthreads();
  FMC_PHASE(7);
  // wave 0 finishes the sums; the decision that reads them runs on wave 0
  // too, so no further barrier is needed before it (callers that read
  // sh.tot or the overlap flags from other waves synchronise first)
  if (wave == 0) {
    // pairwise over the waves (missing waves count as zero); the kernels that
    // evaluate trials run at most 512 threads, 8 waves
    constexpr int KWR = 8;
    static_assert(KWR <= KWAVES, "reduction rows");
    double v[KWR];
    if (lane < A::NRED) {
#pragma unroll
      for (int wv = 0; wv < KWR; ++wv) v[wv] = wv < nwaves ? sh.red[wv][lane] : 0.0;
#pragma unroll
      for (int st = KWR / 2; st >= 1; st >>= 1)
#pragma unroll
        for (int wv = 0; wv < st; ++wv) v[wv] += v[wv + st];
    } else {
      v[0] = 0.0;
    }
    FMC_PHASE(8);
    if (team > 1) v[0] = team_exchange<NS>(W, w, sh, v[0], lane, team, member);
    FMC_PHASE(9);
    if (lane < A::NRED) sh.tot[lane] = v[0];
    if (lane >= A::NACC && lane < A::NRED) {
      const long long n = (long long)v[0];
      int &ovl = lane == A::NACC ? sh.ovl_new : sh.ovl_old;
      int &srf = lane == A::NACC ? sh.srf_new : sh.srf_old;
      ovl = (n % 1024) != 0;
      srf = n >= 1024;
    }
  }
}

// ------------------------------------------------------------- decision --
// Combines the reduced sums into estep and delE exactly as insertion,
// deletion and displace_guest do, then applies energy_eval
// (utility_pack.f:3637-3693) with the mc_insert / mc_delete / mc_displace
// acceptance rules.  Thread 0.  If draw == 0 the Metropolis test is skipped
// (one-trial-at-a-time seam: the caller decides).
template <int NS>
__device__ void decide(const DevParams &P, Smem<NS> &sh, int draw, unsigned char *lds, bool want_comp) {
  WState &ws = sh.ws;
  if (sh.kind == 0) {
    // an empty Wang-Landau step (no molecule to move): only the flatness test
    sh.estep = 0.0;
    sh.ovl_new = 0;
    sh.srf_new = sh.srf_old = 0;
    sh.rande = -1.0;
    sh.drew = 0;
    sh.accepted = 0;
    for (int i = 0; i <= KMAXS; ++i) sh.delE[i] = 0.0;
    if (draw && ws.wl_on) wl_check(P, ws, wl_view(P, lds));
    return;
  }
  const int g = uni(sh.g), mol = uni(sh.mol), kind = uni(sh.kind);
  const double *t = sh.tot;
  const double C = P.dc_C, reu = P.rengunit;
  const double engsic0 = P.dc_esic[g], qfix0 = sh.qfix0;
  double e1new[KMAXS + 1], e1old[KMAXS + 1];
  for (int i = 1; i <= NS; ++i) {
    double sic = (i == mol) ? engsic0 : 0.0;
    double qfm = sh.qfm[i];
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
  const int swap_phase = uni(sh.swap_phase);
  // a swap's insertion adds to the deletion's delE, as the reference's do
  if (swap_phase != 2)
    for (int i = 0; i <= KMAXS; ++i) sh.delE[i] = 0.0;
  if (kind == FMC_MOVE_INSERT) {
    double delrc = sh.delrc;
    overlap = sh.ovl_new;
    if (!overlap) {
      double ewld1eng = C * t[4 * NS] + engsic0 + qfix0;
      estep = (ewld1eng + real_n - ewld3 + vdw_n + delrc) * reu;
      for (int i = 1; i <= NS; ++i) {
        int ik = loc2(mol, i);
        if (i != mol) sh.delE[i] += (e1new[i] + t[i - 1] + t[NS + i - 1] + sh.delrc_mol[ik]) * reu;
      }
      sh.delE[mol] += estep;
      if (want_comp) {
        sh.comp[0] = (ewld1eng - ewld3) * reu;
        sh.comp[1] = real_n * reu;
        sh.comp[2] = vdw_n * reu;
        sh.comp[3] = delrc * reu;
      }
    }
  } else if (kind == FMC_MOVE_DELETE) {
    double delrc = sh.delrc;
    overlap = sh.ovl_old;
    double ewld1eng = C * t[4 * NS + 1] + (-engsic0) + (-qfix0);
    estep = (-ewld1eng + real_o + vdw_o - ewld3) * reu;
    estep = estep - delrc * reu;
    for (int i = 1; i <= NS; ++i) {
      int ik = loc2(mol, i);
      if (i != mol) sh.delE[i] -= (e1old[i] + t[2 * NS + i - 1] + t[3 * NS + i - 1] + sh.delrc_mol[ik]) * reu;
    }
    sh.delE[mol] -= estep;
    if (want_comp) {
      sh.comp[0] = (-ewld1eng - ewld3) * reu;
      sh.comp[1] = real_o * reu;
      sh.comp[2] = vdw_o * reu;
      sh.comp[3] = -delrc * reu;
    }
  } else {
    overlap = sh.ovl_new;
    double ewld1_old = C * t[4 * NS + 1] + (-engsic0) + (-qfix0);
    double est_old = sh.ovl_old ? 0.0 : (-ewld1_old + real_o + vdw_o - ewld3) * reu;
    double ewld1_new = C * t[4 * NS] + engsic0 + qfix0;
    double est_new = overlap ? 0.0 : (ewld1_new + real_n + vdw_n - ewld3) * reu;
    estep = est_new - est_old;
    for (int i = 1; i <= NS; ++i)
      if (i != mol)
        sh.delE[i] += (e1new[i] - e1old[i] + t[i - 1] - t[2 * NS + i - 1] + t[NS + i - 1] - t[3 * NS + i - 1]) * reu;
    sh.delE[mol] += estep;
    if (want_comp) {
      sh.comp[0] = ((ewld1_new - ewld3) - (-ewld1_old - ewld3)) * reu;
      sh.comp[1] = (real_n - real_o) * reu;
      sh.comp[2] = (vdw_n - vdw_o) * reu;
      sh.comp[3] = 0.0;
    }
  }
  sh.estep = estep;
  sh.ovl_new = overlap;
  sh.rande = -1.0;
  sh.drew = 0;
  if (__builtin_expect(swap_phase == 1, 0)) {
    // mc_swap's deletion, accepted provisionally (gcmc.f:2345-2353)
    sh.swap_estepi = -estep;
    sh.accepted = 1;
    return;
  }
  if (__builtin_expect(swap_phase == 2, 0)) {
    // mc_swap's test (gcmc.f:2390-2413): the deletion and the insertion each
    // against the reference's own factors (volume in A^3), the second draw
    // only when the first passes; jnmols counts the inserted molecule
    const double estepi = sh.swap_estepi, estepj = overlap ? 0.0 : estep;
    int acc = 0;
    if (draw && !overlap) {
      const int ig = uni(sh.swap_ig), imol = P.locguest[ig];
      double rande = duni(ws);
      sh.drew = 1;
      const double inm = (double)(ws.nmols[imol] + 1), jnm = (double)(ws.nmols[mol] + 1);
      const double testi = inm * BOLTZ * ws.temp / (P.volm * ws.fug[ig]) * exp(-1.0 * ws.beta * estepi);
      const double testj = P.volm * ws.fug[g] / (jnm * BOLTZ * ws.temp) * exp(-1.0 * ws.beta * estepj);
      if (rande < testi) {
        rande = duni(ws);
        acc = rande < testj;
      }
      sh.rande = rande;
    }
    if (acc && ws.nmols[mol] >= P.gcap[g]) {
      acc = 0;  // guest slots full
      ws.overflow = 1;
    }
    sh.estep = estepi + estepj;
    sh.swap_ovl = overlap;
    sh.accepted = acc;
    return;
  }
  int accepted = 0;
  if (__builtin_expect(draw && sh.wlmode == 2, 0)) {
    // insert_guests: kept whenever it does not overlap (mc_moves.f:126-142)
    accepted = !overlap;
  } else if (__builtin_expect(draw && sh.wlmode == 1 && (kind == FMC_MOVE_INSERT || kind == FMC_MOVE_DELETE), 0)) {
    // eval_wl_move, wang_landau.f:540-588: the grand-canonical weight without
    // the fugacity, biased by exp(ln g(N) - ln g(N +- 1))
    const double beta = ws.beta;
    const double *lng = (const double *)(lds + P.wl_off);
    const int nmol = ws.nmols[mol];
    if (kind == FMC_MOVE_INSERT) {
      if (!overlap) {
        double unbiased = exp(-1.0 * beta * estep) * P.volm / (double)(1 + nmol);
        double contrib = unbiased / ws.wl_lambda3 * exp(lng[nmol] - lng[nmol + 1]);
        sh.rande = duni(ws);
        sh.drew = 1;
        accepted = sh.rande < contrib;
      }
    } else {
      double e = -estep;
      double unbiased = exp(-1.0 * beta * e) * (double)nmol / P.volm;
      double contrib = unbiased * ws.wl_lambda3 * exp(lng[nmol] - lng[nmol - 1]);
      sh.rande = duni(ws);
      sh.drew = 1;
      accepted = sh.rande < contrib;
    }
  } else if (draw) {
    // energy_eval, utility_pack.f:3637-3693; sh.pref holds the prefactor of
    // the insertion / deletion test
    const double beta = ws.beta;
    if (kind == FMC_MOVE_INSERT) {
      if (!overlap) {
        sh.rande = duni(ws);
        sh.drew = 1;
        double e = estep, test;
        if (-1 * beta * e > 700.0) test = exp(700.0);
        else if (beta * e > 700.0) test = exp(-700.0);
        else test = sh.pref * exp(-1.0 * beta * e);
        accepted = sh.rande < test;
      }
    } else if (kind == FMC_MOVE_DELETE) {
      sh.rande = duni(ws);
      sh.drew = 1;
      double e = -estep, test;
      if (-1 * beta * e > 700.0) test = exp(700.0);
      else if (beta * e > 700.0) test = exp(-700.0);
      else if (fabs(ws.fug[g]) < 1e-7) test = 1.0;
      else test = sh.pref * exp(-1.0 * beta * e);
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
  }
  if (__builtin_expect(draw && accepted && kind == FMC_MOVE_INSERT && ws.nmols[mol] >= P.gcap[g], 0)) {
    accepted = 0;  // guest slots full: the walker stops growing and says so
    ws.overflow = 1;
  }
  sh.accepted = accepted;
  if (__builtin_expect(draw && ws.wl_on, 0)) {
    const WLView v = wl_view(P, lds);
    if (sh.wlmode == 1) {
      // the bin of the resulting N (wang_landau.f:668-677, 757-766)
      const int nmol = ws.nmols[mol];
      if (kind == FMC_MOVE_INSERT) wl_visit(ws, v, accepted ? nmol + 1 : nmol);
      else if (kind == FMC_MOVE_DELETE) wl_visit(ws, v, accepted ? nmol - 1 : nmol);
    }
    if (sh.wlmode == 2) {
      if (ws.wl_iter >= 1000LL * ws.wl_preload && !(accepted && ws.nmols[mol] + 1 >= ws.wl_preload)) {
        ws.wl_error = FMC_ERR_INSERT_ITER;  // mc_moves.f:125
        ws.wl_done = 1;
      }
    } else {
      wl_check(P, ws, v);
    }
  }
}

// -------------------------------------------------------------- accept --
// accept_move, mc_moves.f:485-621 (plus the S(k) bookkeeping that
// mc_displace / displace_guest do around it, gcmc.f:1863-1930,2862-2871).
// S(k) update of one k-vector for the accepted move.
template <int NS>
__device__ inline void sk_update(int kind, int mol, double2 &sm, double2 &st, double2 sn, double2 so) {
  if (kind == FMC_MOVE_INSERT) {
    sm.x += sn.x; sm.y += sn.y;
    st.x += sn.x; st.y += sn.y;
  } else if (kind == FMC_MOVE_DELETE) {
    sm.x -= so.x; sm.y -= so.y;
    st.x -= so.x; st.y -= so.y;
  } else {
    sm.x = (sm.x - so.x) + sn.x; sm.y = (sm.y - so.y) + sn.y;
    st.x = (st.x - so.x) + sn.x; st.y = (st.y - so.y) + sn.y;
  }
}

// accept_move, mc_moves.f:485-621 (plus the S(k) bookkeeping that
// mc_displace / displace_guest do around it, gcmc.f:1863-1930,2862-2871).
// REG: the register slots ks hold S(k) and this trial's s_new / s_old for
// their k-vectors; every other k of the workgroup's range goes through
// global memory.  LF: the LDS copy of the guest coordinates is updated with
// the global arrays the workgroup keeps (GuestXYZ).
template <int NS, bool LF, bool REG>
__device__ __forceinline__ void apply(const DevParams &P, const DevWalkers &W, int w, Smem<NS> &sh, unsigned char *lds,
                      KSlot<NS> *ks) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int g = uni(sh.g), mol = uni(sh.mol), natms = uni(sh.natms), kind = uni(sh.kind);
  const int k0 = uni(sh.k0), k1 = uni(sh.k1);
  double2 *skw = W.sk + (size_t)w * (NS + 1) * P.nk;
  const double2 *scr = W.sscr + (size_t)w * 2 * P.nk;
  int kfirst = k0;
  if constexpr (REG) {
#pragma unroll
    for (int m = 0; m < KPT; ++m) {
      // mol - 1 selects the row at run time; keep the slot in registers
#pragma unroll
      for (int i = 0; i < NS; ++i)
        if (i == mol - 1) sk_update<NS>(kind, mol, ks[m].S[i], ks[m].S[NS], ks[m].sn, ks[m].so);
    }
    kfirst = k0 + KPT * nt;
  }
  for (int k = kfirst + tid; k < k1; k += nt) {
    double2 sm = skw[(size_t)(mol - 1) * P.nk + k], st = skw[(size_t)NS * P.nk + k];
    sk_update<NS>(kind, mol, sm, st, scr[k], scr[P.nk + k]);
    skw[(size_t)(mol - 1) * P.nk + k] = sm;
    skw[(size_t)NS * P.nk + k] = st;
  }
  const GuestXYZ<LF> G(P, W, w, lds, uni(sh.member));
  double *gxw = G.gx ? G.gx + P.gbase[g] : nullptr;
  double *gyw = G.gy ? G.gy + P.gbase[g] : nullptr;
  double *gzw = G.gz ? G.gz + P.gbase[g] : nullptr;
  double *lx = nullptr, *ly = nullptr, *lz = nullptr;
  if constexpr (LF) {
    lx = (double *)(lds + P.off_g) + P.gbase[g];
    ly = lx + P.gatoms_total;
    lz = ly + P.gatoms_total;
  }
  const int nm = uni(sh.ws.nmols[mol]), choice = uni(sh.choice);
  if (kind != FMC_MOVE_DELETE) {
    if (tid < natms) {
      int a = (kind == FMC_MOVE_INSERT ? nm : choice - 1) * natms + tid;
      double x = sh.xn[tid][0], y = sh.xn[tid][1], z = sh.xn[tid][2];
      if (gxw) {
        gxw[a] = x;
        gyw[a] = y;
        gzw[a] = z;
      }
      if constexpr (LF) {
        lx[a] = x;
        ly[a] = y;
        lz[a] = z;
      }
    }
  } else {
    // shift the molecules after `choice` down by one, keeping their order
    int lo = (choice - 1) * natms, hi = (nm - 1) * natms;
    for (int b0 = lo; b0 < hi; b0 += nt) {
      int j = b0 + tid;
      double x = 0, y = 0, z = 0;
      if (j < hi) {
        if constexpr (LF) {
          x = lx[j + natms];
          y = ly[j + natms];
          z = lz[j + natms];
        } else {
          x = gxw[j + natms];
          y = gyw[j + natms];
          z = gzw[j + natms];
        }
      }
      __syncthreads();
      if (j < hi) {
        if (gxw) {
          gxw[j] = x;
          gyw[j] = y;
          gzw[j] = z;
        }
        if constexpr (LF) {
          lx[j] = x;
          ly[j] = y;
          lz[j] = z;
        }
      }
      __syncthreads();
    }
  }
  if (tid == 0) {
    WState &ws = sh.ws;
    int kk = loc2(mol, mol);
    for (int i = 1; i <= NS; ++i) ws.energy[i] += sh.delE[i];
    if (kind == FMC_MOVE_INSERT || kind == FMC_MOVE_DELETE) {
      double sg = kind == FMC_MOVE_INSERT ? 1.0 : -1.0;
      double engsic0 = P.dc_esic[g];
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
      // accept_move's surface tally (mc_moves.f:526,568)
      if (kind == FMC_MOVE_INSERT ? sh.srf_new : sh.srf_old) ws.surfmols[mol] += dn;
      ws.nmols[mol] += dn;
      gstart_from(P, ws.nmols, sh);
    }
  }
  __syncthreads();
}

// mc_swap's back-ups before its provisional deletion (gcmc.f:2297-2343): the
// S(k) rows (register slots and global rows, each k by the thread that owns
// it) and the walker totals that a rejected swap restores.
template <int NS>
__device__ __forceinline__ void swap_backup(const DevParams &P, const DevWalkers &W, int w, Smem<NS> &sh, KSlot<NS> *ks) {
  const int tid = threadIdx.x, nt = blockDim.x;
  // every row, so that no row is selected by a run-time index (which would
  // move the register slots to scratch memory)
  double2 *bak = W.sbak + (size_t)w * (NS + 1) * P.nk;
  const double2 *skw = W.sk + (size_t)w * (NS + 1) * P.nk;
  const int k0 = uni(sh.k0), k1 = uni(sh.k1);
#pragma unroll
  for (int m = 0; m < KPT; ++m) {
    const int k = kslot(k0, m);
    if (k < k1) {
#pragma unroll
      for (int i = 0; i <= NS; ++i) bak[(size_t)i * P.nk + k] = ks[m].S[i];
    }
  }
  for (int k = k0 + KPT * nt + tid; k < k1; k += nt)
    for (int i = 0; i <= NS; ++i) bak[(size_t)i * P.nk + k] = skw[(size_t)i * P.nk + k];
  if (tid == 0) {
    const WState &ws = sh.ws;
    for (int i = 0; i < KMAXS + 2; ++i) sh.bk_energy[i] = ws.energy[i];
    for (int i = 0; i <= KMAXS; ++i) sh.bk_chg[i] = ws.chgsum_mol[i];
    for (int i = 0; i <= KMAXC; ++i) {
      sh.bk_elrc_mol[i] = ws.elrc_mol[i];
      sh.bk_engsic[i] = ws.engsic[i];
    }
    sh.bk_elrc = ws.elrc;
    for (int i = 0; i <= KMAXS; ++i) sh.bk_surf[i] = ws.surfmols[i];
    sh.swap_srf = sh.srf_old;
  }
}

// A rejected swap (gcmc.f:2418-2450 and reject_move, mc_moves.f:623-673):
// S(k) rows and walker totals back, the deleted molecule back in its place.
// sumchg keeps the provisional insertion's charge as in the reference (the
// deletion's, if the insertion overlapped: the reference then re-adds the
// charge its deletion left behind).
template <int NS, bool LF>
__device__ __forceinline__ void swap_restore(const DevParams &P, const DevWalkers &W, int w, Smem<NS> &sh, unsigned char *lds,
                             KSlot<NS> *ks) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int ig = uni(sh.swap_ig), imol = P.locguest[ig], natms = P.natoms[imol];
  const double2 *bak = W.sbak + (size_t)w * (NS + 1) * P.nk;
  double2 *skw = W.sk + (size_t)w * (NS + 1) * P.nk;
  const int k0 = uni(sh.k0), k1 = uni(sh.k1);
#pragma unroll
  for (int m = 0; m < KPT; ++m) {
    const int k = kslot(k0, m);
    if (k < k1) {
#pragma unroll
      for (int i = 0; i <= NS; ++i) ks[m].S[i] = bak[(size_t)i * P.nk + k];
    }
  }
  for (int k = k0 + KPT * nt + tid; k < k1; k += nt)
    for (int i = 0; i <= NS; ++i) skw[(size_t)i * P.nk + k] = bak[(size_t)i * P.nk + k];
  // molecule swap_ichoice back in place: the ones after it move up one slot
  const GuestXYZ<LF> G(P, W, w, lds, uni(sh.member));
  double *gxw = G.gx ? G.gx + P.gbase[ig] : nullptr;
  double *gyw = G.gy ? G.gy + P.gbase[ig] : nullptr;
  double *gzw = G.gz ? G.gz + P.gbase[ig] : nullptr;
  double *lx = nullptr, *ly = nullptr, *lz = nullptr;
  if constexpr (LF) {
    lx = (double *)(lds + P.off_g) + P.gbase[ig];
    ly = lx + P.gatoms_total;
    lz = ly + P.gatoms_total;
  }
  const int nm = uni(sh.ws.nmols[imol]), choice = uni(sh.swap_ichoice);
  const int lo = (choice - 1) * natms, hi = nm * natms;  // atoms [lo, hi) move to [lo + natms, hi + natms)
  for (int top = hi; top > lo; top -= nt) {
    const int j = top - 1 - tid;
    double x = 0, y = 0, z = 0;
    if (j >= lo) {
      if constexpr (LF) {
        x = lx[j];
        y = ly[j];
        z = lz[j];
      } else {
        x = gxw[j];
        y = gyw[j];
        z = gzw[j];
      }
    }
    __syncthreads();
    if (j >= lo) {
      if (gxw) {
        gxw[j + natms] = x;
        gyw[j + natms] = y;
        gzw[j + natms] = z;
      }
      if constexpr (LF) {
        lx[j + natms] = x;
        ly[j + natms] = y;
        lz[j + natms] = z;
      }
    }
    __syncthreads();
  }
  if (tid < natms) {
    const double x = sh.xo[tid][0], y = sh.xo[tid][1], z = sh.xo[tid][2];
    if (gxw) {
      gxw[lo + tid] = x;
      gyw[lo + tid] = y;
      gzw[lo + tid] = z;
    }
    if constexpr (LF) {
      lx[lo + tid] = x;
      ly[lo + tid] = y;
      lz[lo + tid] = z;
    }
  }
  if (tid == 0) {
    WState &ws = sh.ws;
    const int jg = sh.swap_jg;
    for (int i = 0; i < KMAXS + 2; ++i) ws.energy[i] = sh.bk_energy[i];
    for (int i = 0; i <= KMAXC; ++i) {
      ws.elrc_mol[i] = sh.bk_elrc_mol[i];
      ws.engsic[i] = sh.bk_engsic[i];
    }
    ws.elrc = sh.bk_elrc;
    for (int i = 0; i <= KMAXS; ++i) ws.surfmols[i] = sh.bk_surf[i];
    ws.chgsum_mol[imol] = sh.bk_chg[imol];
    ws.chgsum_mol[P.locguest[jg]] = sh.bk_chg[P.locguest[jg]];
    ws.sumchg += sh.swap_ovl ? P.chgtmp[ig] : P.chgtmp[jg];
    for (int i = 0; i < natms; ++i) {
      const int ka = P.site_t[imol][i];
      ws.numtyp[ka] += 1;
      ws.numtyp_mol[imol][ka] += 1;
      if (P.site_f[imol][i] != 0) {
        ws.numfrz[ka] += 1;
        ws.numfrz_mol[imol][ka] += 1;
      }
    }
    ws.nmols[imol] += 1;
    gstart_from(P, ws.nmols, sh);
    for (int i = 0; i <= KMAXS; ++i) sh.delE[i] = 0.0;
  }
  __syncthreads();
}

// Full FP64 structure factor of this walker (ewald1, ewald.f:354-661) for
// the workgroup's k-vectors: register slots and any k beyond them in global
// memory.  Same arithmetic as k_resum.  Records the largest change of the
// total row in ws.resum_drift.  Called by every thread.
template <int NS, bool LF>
__device__ __forceinline__ void resum_walker(const DevParams &P, const DevWalkers &W, int w, Smem<NS> &sh, unsigned char *lds,
                             KSlot<NS> *ks) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const GuestXYZ<LF> G(P, W, w, lds, uni(sh.member));
  const int k0 = uni(sh.k0), k1 = uni(sh.k1);
  const double *fx, *fy, *fz, *fq;
  const int *fi;
  const int *kl;
  if constexpr (LF) {
    fx = (const double *)(lds + P.off_fw);
    fy = fx + P.nfw;
    fz = fy + P.nfw;
    fq = fz + P.nfw;
    fi = (const int *)(fq + P.nfw);
    kl = (const int *)((const double *)(lds + P.off_k) + P.nk);
  } else {
    fx = P.fw_per_walker ? W.fwx + (size_t)w * P.nfw : P.fwx;
    fy = P.fw_per_walker ? W.fwy + (size_t)w * P.nfw : P.fwy;
    fz = P.fw_per_walker ? W.fwz + (size_t)w * P.nfw : P.fwz;
    fq = P.fwq;
    fi = P.fwinfo;
    kl = P.klmn;
  }
  const double *lsq = (const double *)(lds + P.off_cst + CST_LSQ);
  const double *r = P.rcell;
  double2 *skw = W.sk + (size_t)w * (NS + 1) * P.nk;
  double dmax = 0.0;
  if (P.rs_batch > 0) {
    // Atoms in batches of rs_batch: each batch's axis phase rows
    // e^{i 2 pi l s_d}, l = 0..kmax_d, by the recursion of ewald.f:108-147
    // (one row per atom and axis), its charges and species in LDS; every
    // thread then adds the batch to its k-vectors with the products of
    // site_phase.  The k-vectors go in groups of KPT per thread: the
    // register slots first, then the ones in global memory.
    const int kcp = P.kcp, nb = P.rs_batch;
    double2 *rows = (double2 *)(lds + P.off_rs);
    double *rq = (double *)(rows + (size_t)nb * 3 * kcp);
    int *rsp = (int *)(rq + nb);
    int natot = P.nfw, gcut[KMAXS + 2];
    gcut[0] = P.nfw;
    for (int gg = 1; gg <= P.ntpguest; ++gg) {
      natot += sh.ws.nmols[P.locguest[gg]] * P.natoms[P.locguest[gg]];
      gcut[gg] = natot;
    }
    const int nglob = k1 - k0 > KPT * nt ? k1 - k0 - KPT * nt : 0;
    const int ngroups = 1 + (nglob + KPT * nt - 1) / (KPT * nt);
    for (int grp = 0; grp < ngroups; ++grp) {
      double2 acc[KPT][NS + 1];
      int kk[KPT];
#pragma unroll
      for (int m = 0; m < KPT; ++m) {
        kk[m] = grp == 0 ? kslot(k0, m) : k0 + KPT * nt + (grp - 1) * KPT * nt + m * nt + tid;
#pragma unroll
        for (int i = 0; i <= NS; ++i) acc[m][i] = make_double2(0.0, 0.0);
      }
      for (int a0 = 0; a0 < natot; a0 += nb) {
        const int cnt = min(nb, natot - a0);
        __syncthreads();  // the previous batch is consumed
        for (int t = tid; t < 3 * cnt; t += nt) {
          const int ia = t / 3, d = t - 3 * ia, a = a0 + ia;
          double x, y, z, q;
          int sp;
          if (a < P.nfw) {
            x = fx[a];
            y = fy[a];
            z = fz[a];
            q = fq[a];
            sp = (fi[a] >> 8) & 255;
          } else {
            int gg = 1;
            while (a >= gcut[gg]) ++gg;
            const int molg = P.locguest[gg], j = a - gcut[gg - 1], s2 = P.gbase[gg] + j;
            x = G.x[s2];
            y = G.y[s2];
            z = G.z[s2];
            q = lsq[molg * KSITE + j % P.natoms[molg]];
            sp = molg;
          }
          const double fr = r[3 * 0 + d] * x + r[3 * 1 + d] * y + r[3 * 2 + d] * z;
          double s1, c1;
          sincospi(2.0 * fr, &s1, &c1);
          double2 *row = rows + (size_t)(ia * 3 + d) * kcp;
          const int kmax = d == 0 ? P.kmax1 : (d == 1 ? P.kmax2 : P.kmax3);
          row[0] = make_double2(1.0, 0.0);
          double c = c1, sn = s1;
          for (int l = 1; l <= kmax; ++l) {
            row[l] = make_double2(c, sn);
            const double cn = c * c1 - sn * s1, snn = sn * c1 + c * s1;
            c = cn;
            sn = snn;
          }
          if (d == 0) {
            rq[ia] = q;
            rsp[ia] = sp;
          }
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < KPT; ++m) {
          if (kk[m] >= k1) continue;
          const int lmn = kl[kk[m]];
          const int ll = lmn & 255, mm = ((lmn >> 8) & 255) - 128, nn = ((lmn >> 16) & 255) - 128;
          for (int ia = 0; ia < cnt; ++ia) {
            double ckc, cks;
            site_phase<double>(rows + (size_t)ia * 3 * kcp, kcp, ll, mm, nn, rq[ia], ckc, cks);
            const int sp = rsp[ia];
#pragma unroll
            for (int i = 0; i < NS; ++i)
              if (sp == i + 1) {
                acc[m][i].x += ckc;
                acc[m][i].y += cks;
              }
            acc[m][NS].x += ckc;
            acc[m][NS].y += cks;
          }
        }
      }
#pragma unroll
      for (int m = 0; m < KPT; ++m) {
        const int k = kk[m];
        if (k >= k1) continue;
        if (grp == 0) {
          dmax = fmax(dmax, fmax(fabs(ks[m].S[NS].x - acc[m][NS].x), fabs(ks[m].S[NS].y - acc[m][NS].y)));
#pragma unroll
          for (int i = 0; i <= NS; ++i) ks[m].S[i] = acc[m][i];
        } else {
          const double2 o = skw[(size_t)NS * P.nk + k];
          dmax = fmax(dmax, fmax(fabs(o.x - acc[m][NS].x), fabs(o.y - acc[m][NS].y)));
#pragma unroll
          for (int i = 0; i <= NS; ++i) skw[(size_t)i * P.nk + k] = acc[m][i];
        }
      }
    }
    __syncthreads();
  } else {
    // no LDS room for the rows: one phase per (atom, k) directly
    auto one_k = [&](int k, double2 *S) {
      int lmn = kl[k];
      double l = (double)(lmn & 255), m = (double)(((lmn >> 8) & 255) - 128), n = (double)(((lmn >> 16) & 255) - 128);
      double sr[NS + 1], si[NS + 1];
#pragma unroll
      for (int i = 0; i <= NS; ++i) sr[i] = si[i] = 0.0;
      for (int j = 0; j < P.nfw; ++j) {
        double x = fx[j], y = fy[j], z = fz[j];
        double sx = r[0] * x + r[3] * y + r[6] * z, sy = r[1] * x + r[4] * y + r[7] * z,
               sz = r[2] * x + r[5] * y + r[8] * z;
        double sn, cs;
        sincospi(2.0 * (l * sx + m * sy + n * sz), &sn, &cs);
        double q = fq[j];
        int sp = (fi[j] >> 8) & 255;
#pragma unroll
        for (int i = 0; i < NS; ++i)
          if (sp == i + 1) {
            sr[i] += q * cs;
            si[i] += q * sn;
          }
        sr[NS] += q * cs;
        si[NS] += q * sn;
      }
      for (int gg = 1; gg <= P.ntpguest; ++gg) {
        int molg = P.locguest[gg], na = P.natoms[molg], nact = sh.ws.nmols[molg] * na;
        for (int j = 0; j < nact; ++j) {
          int a = P.gbase[gg] + j;
          double x = G.x[a], y = G.y[a], z = G.z[a];
          double sx = r[0] * x + r[3] * y + r[6] * z, sy = r[1] * x + r[4] * y + r[7] * z,
                 sz = r[2] * x + r[5] * y + r[8] * z;
          double sn, cs;
          sincospi(2.0 * (l * sx + m * sy + n * sz), &sn, &cs);
          double q = lsq[molg * KSITE + j % na];
#pragma unroll
          for (int i = 0; i < NS; ++i)
            if (molg == i + 1) {
              sr[i] += q * cs;
              si[i] += q * sn;
            }
          sr[NS] += q * cs;
          si[NS] += q * sn;
        }
      }
      double d = fmax(fabs(S[NS].x - sr[NS]), fabs(S[NS].y - si[NS]));
#pragma unroll
      for (int i = 0; i <= NS; ++i) S[i] = make_double2(sr[i], si[i]);
      return d;
    };
#pragma unroll
    for (int m = 0; m < KPT; ++m) {
      int k = kslot(k0, m);
      if (k < k1) dmax = fmax(dmax, one_k(k, ks[m].S));
    }
    for (int k = k0 + KPT * nt + tid; k < k1; k += nt) {
      double2 S[NS + 1];
#pragma unroll
      for (int i = 0; i <= NS; ++i) S[i] = skw[(size_t)i * P.nk + k];
      dmax = fmax(dmax, one_k(k, S));
#pragma unroll
      for (int i = 0; i <= NS; ++i) skw[(size_t)i * P.nk + k] = S[i];
    }
  }
  // workgroup maximum: per wave, then across waves through LDS
  for (int off = 32; off >= 1; off >>= 1) dmax = fmax(dmax, __shfl_xor(dmax, off, 64));
  const int lane = tid & 63, wave = tid >> 6, nwaves = (nt + 63) >> 6;
  if (lane == 0) sh.red[wave][0] = dmax;
  __syncthreads();
  if (tid == 0) {
    double v = 0.0;
    for (int q = 0; q < nwaves; ++q) v = fmax(v, sh.red[q][0]);
    sh.ws.resum_drift = fmax(sh.ws.resum_drift, v);
    sh.ws.nresum += 1;
    sh.ws.acc_since_resum = 0;
  }
  __syncthreads();
}

// Guest kind gg's first index in the atom space of the pair loop, from the
// molecule counts by species (one thread).
template <int NS>
__device__ inline void gstart_from(const DevParams &P, const int *nmols, Smem<NS> &sh) {
  int n = P.nfw;
  for (int gg = 1; gg <= KMAXS; ++gg) {
    if (gg <= P.ntpguest) {
      const int m = P.locguest[gg];
      sh.gstart[gg] = n;
      n += nmols[m] * P.natoms[m];
    } else {
      sh.gstart[gg] = 0x7fffffff;
    }
  }
  sh.ntot = n;
  // this team member's atoms [a0, a1) of the pair loop (sh.team and
  // sh.member are set before the first call)
  sh.a0 = n * sh.member / sh.team;
  sh.a1 = n * (sh.member + 1) / sh.team;
}

// The walker's state into LDS; team / member: the workgroup's place in the
// walker's team (1 / 0 for a walker without one), which sets its k-vectors.
template <int NS>
__device__ inline void stage_in(const DevParams &P, const DevWalkers &W, int w, Smem<NS> &sh, int team = 1,
                                int member = 0) {
  const int *src = (const int *)(W.ws + w);
  int *dst = (int *)&sh.ws;
  for (int i = threadIdx.x; i < (int)(sizeof(WState) / 4); i += blockDim.x) dst[i] = src[i];
  const int t = threadIdx.x;
  if (t < 9) {
    sh.rc9[t] = P.rcell[t];
    sh.cl9[t] = P.cell[t];
  }
  for (int i = t; i < (KMAXS + 1) * KSITE * 3; i += blockDim.x) {
    const int gs = i / 3;
    sh.tmpl[gs / KSITE][gs % KSITE][i % 3] = P.tmpl[gs / KSITE][gs % KSITE][i % 3];
  }
  if (t == 9) {
    sh.kmx[0] = P.kmax1;
    sh.kmx[1] = P.kmax2;
    sh.kmx[2] = P.kmax3;
    sh.ovl2 = P.overlap2;
  }
  for (int i = t; i < (KMAXS + 1) * KSITE; i += blockDim.x) {
    sh.scq[i / KSITE][i % KSITE] = P.site_chgea[i / KSITE][i % KSITE];
    sh.sq[i / KSITE][i % KSITE] = P.site_q[i / KSITE][i % KSITE];
    sh.sm[i / KSITE][i % KSITE] = P.site_m[i / KSITE][i % KSITE];
  }
  if (t == 0) {
    sh.team = team;
    sh.member = member;
    gstart_from(P, W.ws[w].nmols, sh);
    sh.mkind = 0;
    sh.swap_phase = 0;
    sh.k0 = (int)((long long)P.nk * member / team);
    sh.k1 = (int)((long long)P.nk * (member + 1) / team);
    sh.xseq = team > 1 ? W.xcnt[w] : 0;
    sh.team_fail = 0;
    sh.tab_ready = 0;
    sh.spec_on = 0;
    sh.accepted = 0;
  }
  for (int i = t; i < (int)(sizeof(sh.prop) / 4); i += blockDim.x) ((int *)sh.prop)[i] = 0;
  if (t <= KMAXS) {
    // guest kinds beyond ntpguest: species 1 (never read), no sites
    const bool real = t >= 1 && t <= P.ntpguest;
    const int m = real ? P.locguest[t] : 1;
    sh.kmol[t] = m;
    sh.kna[t] = real ? P.natoms[m] : 1;
    sh.kslot[t] = real ? P.gbase[t] : 0;
  }
  __syncthreads();
}

template <int NS>
__device__ inline void stage_out(const DevWalkers &W, int w, Smem<NS> &sh) {
  __syncthreads();
  if (sh.member != 0) return;  // member 0 holds the walker's state
  if (threadIdx.x == 0 && sh.team > 1) W.xcnt[w] = sh.xseq;
  int *dst = (int *)(W.ws + w);
  const int *src = (const int *)&sh.ws;
  for (int i = threadIdx.x; i < (int)(sizeof(WState) / 4); i += blockDim.x) dst[i] = src[i];
}

template <int NS>
__device__ void record(const DevParams &P, const DevWalkers &W, Smem<NS> &sh, long long step) {
  fmc_trace_rec *r = W.trace + step;
  const bool swap = sh.mkind == FMC_MOVE_SWAP;
  r->step = (int)sh.ws.steps;
  r->iguest = swap ? sh.swap_ig : sh.g;
  r->move = sh.mkind;
  r->choice = swap ? sh.swap_ichoice : sh.choice;
  r->nmols = swap ? sh.swap_n0 : sh.nmols;
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
  r->surface = sh.srf_new | ((swap ? sh.swap_srf : sh.srf_old) << 1);
  r->nsurface = sh.ws.surfmols[P.locguest[swap ? sh.swap_ig : sh.g]];
}

// mc_displace / mc_jump's surface tally after an accepted move of an existing
// molecule (gcmc.f:1906-1912,2021-2027).  Thread 0.
template <int NS>
__device__ inline void surf_displaced(Smem<NS> &sh) {
  int &n = sh.ws.surfmols[sh.mol];
  if (sh.srf_old && !sh.srf_new) {
    n -= 1;
    if (n < 0) n = 0;
  } else if (!sh.srf_old && sh.srf_new) {
    n += 1;
  }
}

// equitable_bin (utility_pack.f:2407-2519): spread weight 1 over the 8 grid
// points around fractional position (fa, fb, fc) of a grid of na x nb x nc
// points (c fastest; the reference's 1-based bin b is grid[b - 1]), each
// axis split by the position's offset from the nearest point.  The b axis
// keeps the reference's offset of 1 (not 0) for its upper neighbour past the
// last plane.  Bins outside the grid (only reachable with nb = 1) are
// dropped.  Each point receives its weight times val.  Hardware FP64
// atomics: the order of the additions is free.
__device__ inline void equitable_bin(int na, int nb, int nc, double *grid, double fa, double fb, double fc,
                                     double val) {
  const int gsize = na * nb * nc;
  const double gpta = na * fa, gptb = nb * fb, gptc = nc * fc;
  const int aidx = (int)ceil(gpta), bidx = (int)ceil(gptb), cidx = (int)ceil(gptc);
  const double apart = gpta - floor(gpta), bpart = gptb - floor(gptb), cpart = gptc - floor(gptc);
  int alidx, aridx, blidx, bridx, clidx, cridx;
  double al, ar, bl, br, cl, cr;
  if (apart > 0.5) {
    alidx = aidx >= na ? (na - 1) * nb * nc : (aidx - 1) * nb * nc;
    aridx = aidx >= na ? 0 : aidx * nb * nc;
    al = 1.5 - apart;
    ar = apart - 0.5;
  } else {
    alidx = aidx <= 1 ? (na - 1) * nb * nc : (aidx - 2) * nb * nc;
    aridx = aidx <= 1 ? 0 : (aidx - 1) * nb * nc;
    al = 0.5 - apart;
    ar = 0.5 + apart;
  }
  if (bpart > 0.5) {
    blidx = bidx >= nb ? (nb - 1) * nc : (bidx - 1) * nc;
    bridx = bidx >= nb ? 1 : bidx * nc;
    bl = 1.5 - bpart;
    br = bpart - 0.5;
  } else {
    blidx = bidx <= 1 ? (nb - 1) * nc : (bidx - 2) * nc;
    bridx = bidx <= 1 ? 0 : (bidx - 1) * nc;
    bl = 0.5 - bpart;
    br = 0.5 + bpart;
  }
  if (cpart > 0.5) {
    clidx = cidx >= nc ? nc : cidx;
    cridx = cidx >= nc ? 1 : cidx + 1;
    cl = 1.5 - cpart;
    cr = cpart - 0.5;
  } else {
    clidx = cidx <= 1 ? nc : cidx - 1;
    cridx = cidx <= 1 ? 1 : cidx;
    cl = 0.5 - cpart;
    cr = 0.5 + cpart;
  }
  const int bins[8] = {alidx + blidx + clidx, aridx + blidx + clidx, alidx + bridx + clidx, alidx + blidx + cridx,
                       aridx + bridx + clidx, aridx + blidx + cridx, alidx + bridx + cridx, aridx + bridx + cridx};
  const double wts[8] = {al * bl * cl, ar * bl * cl, al * br * cl, al * bl * cr,
                         ar * br * cl, ar * bl * cr, al * br * cr, ar * br * cr};
#pragma unroll
  for (int i = 0; i < 8; ++i)
    if (bins[i] >= 1 && bins[i] <= gsize) unsafeAtomicAdd(grid + bins[i] - 1, wts[i] * val);
}

// storeprob (utility_pack.f:622-687) for one production step: every guest
// molecule of a plotted kind, its centre of mass (com, utility_pack.f:
// 3565-3596) and sites in fractional coordinates of the unit cell folded
// into [0, 1) (frac / frac2, utility_pack.f:873-921), binned into each of
// its kind's plots.  Threads t0, t0 + nthr, ... take one molecule each.
template <bool LF>
__device__ inline void prob_step(const DevParams &P, const DevWalkers &W, int w, const int *nmols,
                                 const GuestXYZ<LF> &G, int t0, int nthr) {
  const double *r = P.prucell;
  auto frac = [&](double x, double y, double z, double &a, double &b, double &c) {
    const double sx = x * r[0] + y * r[3] + z * r[6], sy = x * r[1] + y * r[4] + z * r[7],
                 sz = x * r[2] + y * r[5] + z * r[8];
    a = sx - floor(sx);
    b = sy - floor(sy);
    c = sz - floor(sz);
  };
  for (int g = 1; g <= P.ntpguest; ++g) {
    bool any = false;
    for (int p = 0; p < P.nplots; ++p) any |= P.plot_g[p] == g;
    if (!any) continue;
    const int mol = P.locguest[g], na = P.natoms[mol], nm = nmols[mol];
    for (int m = t0; m < nm; m += nthr) {
      const int base = P.gbase[g] + m * na;
      double wsum = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
      for (int s = 0; s < na; ++s) {
        const double ms = P.site_m[mol][s];
        wsum += ms;
        cx += G.x[base + s] * ms;
        cy += G.y[base + s] * ms;
        cz += G.z[base + s] * ms;
      }
      double ca, cb, cc;
      frac(cx / wsum, cy / wsum, cz / wsum, ca, cb, cc);
      for (int p = 0; p < P.nplots; ++p) {
        if (P.plot_g[p] != g) continue;
        double *grid = W.prob + ((size_t)w * P.nplots + p) * P.pgridsize;
        for (int e = 0; e < P.plot_ns[p]; ++e) {
          const int site = P.plot_site[p][e];
          if (site == 0) {
            equitable_bin(P.pgrid[0], P.pgrid[1], P.pgrid[2], grid, ca, cb, cc, 1.0);
          } else if (site <= na) {
            double a, b, c;
            frac(G.x[base + site - 1], G.y[base + site - 1], G.z[base + site - 1], a, b, c);
            equitable_bin(P.pgrid[0], P.pgrid[1], P.pgrid[2], grid, a, b, c, 1.0);
          }
        }
      }
    }
  }
}

// The numguests.out record of a step (gcmc.f:1138-1163), lane g - 1 of one
// wave for guest kind g: counts, energy, delE and the move flags of the
// step.  delE is the reference's module array: zero after a rejection, the
// accepted move's change after an acceptance, unchanged by a switch that
// does not move; it is kept for every step and written every nnumg steps.
template <int NS>
__device__ inline void numg_step(const DevParams &P, const DevWalkers &W, int w, Smem<NS> &sh) {
  WState &ws = sh.ws;
  const int lane = threadIdx.x & 63;
  if (lane < P.ntpguest && sh.done_mkind != FMC_MOVE_SWITCH) {
    const int m = P.locguest[lane + 1];
    ws.delE_last[m] = sh.accepted ? sh.delE[m] : 0.0;
  }
  if (ws.steps % P.nnumg != 0) return;
  const long long n = ws.nnumg_n;
  const bool room = n < P.numg_cap;
  if (room && lane < P.ntpguest) {
    fmc_numguests_rec *r = W.numg + (size_t)w * P.numg_cap + n;
    const int g = lane + 1, m = P.locguest[g], mk = sh.done_mkind;
    const bool swap = mk == FMC_MOVE_SWAP;
    const int gm = swap ? sh.done_swap_ig : sh.done_g;
    if (lane == 0) r->step = ws.steps;
    r->nmols[lane] = ws.nmols[m];
    r->energy[lane] = ws.energy[m];
    r->delE[lane] = ws.delE_last[m];
    r->moves[lane][0] = mk == FMC_MOVE_INSERT && gm == g;
    r->moves[lane][1] = mk == FMC_MOVE_DELETE && gm == g;
    r->moves[lane][2] = gm != g                       ? 0
                        : mk == FMC_MOVE_DISPLACE  ? 1
                        : mk == FMC_MOVE_TRANSLATE ? 2
                        : mk == FMC_MOVE_ROTATE    ? 3
                                                   : 0;
    r->moves[lane][3] = mk == FMC_MOVE_JUMP && gm == g;
    r->moves[lane][4] = swap && (sh.done_swap_ig == g || sh.done_swap_jg == g);
    r->moves[lane][5] = 0;  // a switch does not move with fewer than three guest kinds
  }
  if (lane == 0) {
    if (room) ws.nnumg_n = n + 1;
    else ws.overflow |= 4;
  }
}

// The production block after a step (gcmc.f:1050-1130), on one wave: count the
// step (only if accepted with the 'sampling' directive), add N, E, E*N, N*N,
// E*E and the surface count of every guest kind to the open window (lane
// 6(g-1)+q holds sum q of guest g), and close the window into W.win when the
// production count reaches a multiple of nwindsteps or, with close, after
// the run's last step.
template <int NS>
__device__ inline void stats_step(const DevParams &P, const DevWalkers &W, int w, Smem<NS> &sh, bool close) {
  WState &ws = sh.ws;
  const int lane = threadIdx.x & 63;
  const int ng = P.ntpguest;
  const bool take = !P.accsample || sh.accepted;
  const long long pc = ws.prodcount + (take ? 1 : 0);
  const long long rc = ws.rollcount + (take ? 1 : 0);
  const bool active = lane < FMC_WIN_STATS * ng;
  const int g = lane / FMC_WIN_STATS + 1, q = lane % FMC_WIN_STATS;
  double sum = 0.0;
  if (active) {
    const int m = P.locguest[g];
    const double n = (double)ws.nmols[m], e = ws.energy[m];
    double v = n;
    if (q == 1) v = e;
    else if (q == 2) v = e * n;
    else if (q == 3) v = n * n;
    else if (q == 4) v = e * e;
    else if (q == 5) v = (double)ws.surfmols[m];
    sum = ws.chain[g][q] + (take ? v : 0.0);
  }
  const bool shut = (P.nwindsteps > 0 && pc % P.nwindsteps == 0) || close;
  const long long nw = ws.nwin;
  const bool room = nw < P.maxwin && W.win != nullptr;
  if (shut && room) {
    fmc_window *rec = W.win + (size_t)w * P.maxwin + nw;
    if (lane == 0) rec->count = (double)rc;
    if (active) rec->sums[g - 1][q] = sum;
  }
  if (active) ws.chain[g][q] = shut ? 0.0 : sum;
  if (lane == 0) {
    ws.prodcount = pc;
    ws.rollcount = shut ? 0 : rc;
    if (shut) {
      if (room) ws.nwin = nw + 1;
      else ws.overflow |= 2;
    }
  }
}

// ------------------------------------------------------------- kernels --
// Device-resident chains, nsteps GCMC steps per walker.  LF / NSITE select
// the all-in-LDS variant and the site count of its straight-line pair code
// (see evaluate).  team == 1: one workgroup per walker.  team > 1: `team`
// workgroups per walker, each with the whole walker state in LDS, running
// the same proposal and decision; each evaluates its share of the pair atoms
// and of the k-vectors (whose S(k) it alone keeps), and the shares are
// added through team_exchange once per step.  The grid is then
// 8 * ceil(nw / 8) * team workgroups, which the host keeps within what the
// device holds at once (a member waits on the others).
template <int NS, typename T, bool LF, int NSITE, bool LJL>
__global__ void __launch_bounds__(512) k_run(const DevParams *__restrict__ Pp, DevWalkers W, int nsteps,
                                             int count_stats, int team, int nw) {
  extern __shared__ __align__(16) unsigned char lds[];
  Smem<NS> &sh = *reinterpret_cast<Smem<NS> *>(lds);
  const DevParams &P = *Pp;
  int w = blockIdx.x, member = 0;
  if (team > 1) {
    // workgroups are dealt to the 8 XCDs in turn, so blocks b, b + 8,
    // b + 16, ... share an XCD and its L2: a walker's members are taken
    // from one such sequence
    const int b = blockIdx.x, j = b >> 3;
    w = (b & 7) + 8 * (j / team);
    member = j - (j / team) * team;
    if (w >= nw) return;  // padding of the last XCD round
  }
  KSlot<NS> ks[KPT];
  load_slots<NS>(P, W, w, (int)((long long)P.nk * member / team), (int)((long long)P.nk * (member + 1) / team), ks);
  stage_in<NS>(P, W, w, sh, team, member);
  stage_tables<T, LF, LJL>(P, W, w, lds);
  if (!LF && member > 0) {
    // a private copy of the guest coordinates (the walker's own arrays are
    // member 0's; it writes them only after the first exchange)
    const size_t n = P.gatoms_total, src = (size_t)w * n;
    double *dst = W.gxt + ((size_t)w * (W.team_alloc - 1) + member - 1) * 3 * n;
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) {
      dst[i] = W.gx[src + i];
      dst[n + i] = W.gy[src + i];
      dst[2 * n + i] = W.gz[src + i];
    }
  }
  if (P.wl_nbins > 0) {
    const WLView v = wl_view(P, lds);
    for (int b = threadIdx.x; b < P.wl_nbins; b += blockDim.x) {
      v.lng[b] = W.wl_lng[(size_t)w * P.wl_nbins + b];
      v.hist[b] = W.wl_hist[(size_t)w * P.wl_nbins + b];
    }
  }
  const GuestXYZ<LF> G(P, W, w, lds, member);
#ifdef FMC_PHASE_CLOCK
  if (threadIdx.x < FMC_NPHASE + 3) fmc_phase_row[threadIdx.x] = 0;
#endif
  __syncthreads();
#ifdef FMC_PHASE_CLOCK
  if (threadIdx.x == 0) fmc_phase_row[FMC_NPHASE] = __builtin_amdgcn_s_memrealtime();
#endif
  // A swap takes two passes of the loop: its deletion (applied provisionally)
  // and the insertion of the other kind, accepted or undone as a whole; only
  // the second pass completes a step.
  int evn = 0;  // evaluations made (every thread counts them alike)
  for (int step = 0; step < nsteps;) {
    if (sh.ws.wl_done) break;  // converged Wang-Landau walker (uniform: written before the last barrier)
    if (sh.team_fail) break;   // a team wait gave up (the walker is flagged)
    if (threadIdx.x < 64) {
      if (__builtin_expect(sh.swap_phase == 1, 0)) {
        swap_insert_setup<NS>(P, sh);
      } else {
        // after a rejected step, the proposal built for it speculatively
        // from the ring position its decision left (0 or 1 draws on)
        int take = -1;
        if (uni(sh.spec_on) && !uni(sh.accepted)) {
          const long long u = sh.ws.mz_used;
          if (sh.prop[0].ok && sh.prop[0].used0 == u) take = 0;
          else if (sh.prop[1].ok && sh.prop[1].used0 == u) take = 1;
        }
        take = uni(take);
        if (take >= 0) {
          commit<NS>(sh, sh.prop[take]);
          FMC_PHASE_COUNT(FMC_NPHASE + 2);
        }
        // at least 33 draws ahead of the step: enough for this step and the
        // next one's speculative proposals (at most 12 per step)
        if (__builtin_expect(sh.ws.mz_made - sh.ws.mz_used < 33, 0)) {
          mzt_refill(sh.ws, threadIdx.x);
          if (threadIdx.x == 0) mzt_advance(sh.ws);
        }
        if (take < 0) {
          propose<NS, LF>(P, G, sh, lds, sh.ws.mz_used, 2, false);
          commit<NS>(sh, sh.prop[2]);
        }
      }
    }
    FMC_PHASE(1);
    __syncthreads();
    FMC_PHASE(2);
    if (sh.kind != 0) evaluate<NS, T, LF, NSITE, LJL>(P, W, w, sh, lds, ks, false, ++evn);
    {
      // the next step's proposal, assuming this step is rejected, on waves 1
      // and 2 (for 0 and 1 draws of the decision) while wave 0 finishes the
      // sums, waits for the team and decides; a swap or a Wang-Landau walker
      // (whose proposals update its visit histogram) proposes live.  Wave 1
      // records whether the proposals were made (read after the barrier).
      const int wv = threadIdx.x >> 6;
      if (wv == 1 || wv == 2) {
        const bool spec = P.spec_prop && uni(sh.swap_phase) == 0 && !uni(sh.ws.wl_on) && blockDim.x >= 192;
        if (spec) propose<NS, LF>(P, G, sh, lds, sh.mz_after + (wv - 1), wv - 1, true);
        if (threadIdx.x == 64) sh.spec_on = spec;
      }
    }
    if (threadIdx.x == 0) {
      FMC_PHASE(10);
      decide<NS>(P, sh, 1, lds, W.trace != nullptr && w == W.trace_w && member == 0);
      FMC_PHASE(11);
      if (sh.swap_phase != 1) {
        WState &ws = sh.ws;
        const int mk = sh.mkind;
        sh.done_mkind = mk;
        sh.done_g = sh.g;
        sh.done_swap_ig = sh.swap_ig;
        sh.done_swap_jg = sh.swap_jg;
        if (mk != FMC_MOVE_SWITCH) ws.att[mk] += 1;
        if (sh.accepted) {
          ws.acc[mk] += 1;
          if (mk == FMC_MOVE_TRANSLATE) ws.accept_tran[sh.g] += 1;
          else if (mk == FMC_MOVE_ROTATE) ws.accept_rota[sh.g] += 1;
          else if (mk == FMC_MOVE_DISPLACE) ws.accept_disp[sh.g] += 1;
        }
        ws.steps += 1;
      }
    }
    FMC_PHASE(12);
    __syncthreads();
    FMC_PHASE(13);
    const int phase = sh.swap_phase;
    if (__builtin_expect(phase == 1, 0)) swap_backup<NS>(P, W, w, sh, ks);
    if (sh.accepted) {
      apply<NS, LF, true>(P, W, w, sh, lds, ks);
      if (threadIdx.x == 0 && sh.kind != FMC_MOVE_INSERT && sh.kind != FMC_MOVE_DELETE) surf_displaced<NS>(sh);
      if (phase != 1 && P.resum_every > 0) {
        // apply ends with a barrier, so every thread sees the same count
        int due = sh.ws.acc_since_resum + 1 >= P.resum_every;
        __syncthreads();
        if (threadIdx.x == 0) sh.ws.acc_since_resum += 1;
        if (__builtin_expect(due, 0)) resum_walker<NS, LF>(P, W, w, sh, lds, ks);
      }
    } else if (__builtin_expect(phase == 2, 0)) {
      swap_restore<NS, LF>(P, W, w, sh, lds, ks);
    }
    FMC_PHASE(14);
    if (phase == 1) continue;
    if (member > 0) {
      // the statistics, records and plots are member 0's
      ++step;
      continue;
    }
    if (__builtin_expect(threadIdx.x == 0 && W.trace && w == W.trace_w, 0)) record<NS>(P, W, sh, step);
    // the step's statistics on wave 1, idle while wave 0 proposes the next
    // step (the fields written here are read nowhere before the next barrier)
    const int stat_wave = blockDim.x > 64 ? 1 : 0;
    if ((int)(threadIdx.x >> 6) == stat_wave && (count_stats & FMC_STATS_COUNT))
      stats_step<NS>(P, W, w, sh, (count_stats & FMC_STATS_CLOSE) && step == nsteps - 1);
    if (__builtin_expect((int)(threadIdx.x >> 6) == stat_wave && P.nnumg > 0, 0)) numg_step<NS>(P, W, w, sh);
    // probability plots on waves 2 and up (every thread of a smaller block)
    if (__builtin_expect(P.nplots > 0 && (count_stats & FMC_STATS_COUNT), 0)) {
      const int p0 = blockDim.x > 128 ? 128 : 0;
      if ((int)threadIdx.x >= p0) prob_step<LF>(P, W, w, sh.ws.nmols, G, threadIdx.x - p0, blockDim.x - p0);
    }
    if ((int)threadIdx.x == 64 * stat_wave && count_stats) {
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
    FMC_PHASE(0);
    FMC_PHASE_COUNT(FMC_NPHASE + 1);
    ++step;
  }
#ifdef FMC_PHASE_CLOCK
  if (threadIdx.x < FMC_NPHASE + 3 && threadIdx.x != FMC_NPHASE && blockIdx.x == 0)
    atomicAdd(&fmc_phase_ticks[threadIdx.x], fmc_phase_row[threadIdx.x]);
#endif
  store_slots<NS>(P, W, w, uni(sh.k0), uni(sh.k1), ks);
  if (P.wl_nbins > 0 && member == 0) {
    __syncthreads();
    const WLView v = wl_view(P, lds);
    for (int b = threadIdx.x; b < P.wl_nbins; b += blockDim.x) {
      W.wl_lng[(size_t)w * P.wl_nbins + b] = v.lng[b];
      W.wl_hist[(size_t)w * P.wl_nbins + b] = v.hist[b];
    }
  }
  stage_out<NS>(W, w, sh);
}

// One trial of the mc_moves.f seam: the proposal comes from the pending
// fields of walker w (written by the host); results go to W.trial_out and
// s_new / s_old of every k-vector to the walker's scratch rows.
template <int NS, typename T, bool LJ32>
__global__ void __launch_bounds__(512) k_trial(const DevParams *__restrict__ Pp, DevWalkers W, int w) {
  extern __shared__ __align__(16) unsigned char lds[];
  Smem<NS> &sh = *reinterpret_cast<Smem<NS> *>(lds);
  const DevParams &P = *Pp;
  KSlot<NS> ks[KPT];
  load_slots<NS>(P, W, w, 0, P.nk, ks);
  stage_in<NS>(P, W, w, sh);
  stage_tables<T, false, false>(P, W, w, lds);
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
      const GuestXYZ<false> G(P, W, w, lds);
      load_molecule<false>(P, G, g, sh.choice, sh.xo, natms);
      sh.excl_lo = (sh.choice - 1) * natms;
      sh.excl_hi = sh.choice * natms;
    }
  }
  __syncthreads();
  evaluate<NS, T, false, 0, false, LJ32>(P, W, w, sh, lds, ks, true, 1);
  if (threadIdx.x == 0) {
    decide<NS>(P, sh, 0, lds, true);
    fmc_trial *o = W.trial_out;
    o->estep = sh.estep;
    for (int i = 0; i < FMC_MAX_SPECIES; ++i) o->delE[i] = (i < NS) ? sh.delE[i + 1] : 0.0;
    o->recip = sh.comp[0];
    o->real = sh.comp[1];
    o->vdw = sh.comp[2];
    o->tail = sh.comp[3];
    o->overlap = sh.ovl_new;
    o->surface = sh.srf_new | (sh.srf_old << 1);
    // keep what accept needs in the walker's own state
    WState &ws = sh.ws;
    ws.pend_valid = 1;
    ws.pend_delrc = sh.delrc;
    for (int i = 0; i <= KMAXS; ++i) ws.pend_delE[i] = sh.delE[i];
    for (int i = 0; i <= KMAXC; ++i) ws.pend_delrc_mol[i] = sh.delrc_mol[i];
    ws.pend_surface = sh.srf_new | (sh.srf_old << 1);
  }
  stage_out<NS>(W, w, sh);
}

template <int NS>
__global__ void __launch_bounds__(512) k_accept(const DevParams *__restrict__ Pp, DevWalkers W, int w) {
  extern __shared__ __align__(16) unsigned char lds[];
  Smem<NS> &sh = *reinterpret_cast<Smem<NS> *>(lds);
  const DevParams &P = *Pp;
  stage_in<NS>(P, W, w, sh);
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
    sh.delrc = ws.pend_delrc;
    for (int i = 0; i <= KMAXS; ++i) sh.delE[i] = ws.pend_delE[i];
    for (int i = 0; i <= KMAXC; ++i) sh.delrc_mol[i] = ws.pend_delrc_mol[i];
    const int sf = ws.pend_surface;
    sh.srf_new = sf & 1;
    sh.srf_old = sf >> 1;
    if (sh.kind == FMC_MOVE_TRANSLATE || sh.kind == FMC_MOVE_ROTATE) sh.kind = FMC_MOVE_DISPLACE;
    ws.pend_valid = 0;
    ws.pend_kind = 0;
  }
  __syncthreads();
  apply<NS, false, false>(P, W, w, sh, lds, nullptr);
  stage_out<NS>(W, w, sh);
}

// Full FP64 structure factor of walkers w0 .. w0 + gridDim.y - 1 (ewald1,
// ewald.f:354-661): one thread per (walker, k); species rows plus the total row.
template <int NS>
__global__ void __launch_bounds__(256) k_resum(const DevParams *__restrict__ Pp, DevWalkers W, double *drift,
                                               int w0) {
  const DevParams &P = *Pp;
  const int w = w0 + (int)blockIdx.y, k = blockIdx.x * blockDim.x + threadIdx.x;
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
