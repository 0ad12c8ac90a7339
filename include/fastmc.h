/*
 * fastmc.h -- C-ABI of the MI355X GCMC trial-move energy library
 * (libfastmc_hip.so).  Plain C types only: no torch, no HIP types.
 *
 * The reference (FastMC, Fortran 77/90) has no plugin API; its seam is the
 * Fortran procedure interface of module mc_moves plus the module globals it
 * reads and writes.  Each entry point below names the reference procedure it
 * stands in for.  A Fortran caller binds them with iso_c_binding interfaces;
 * every pointer is caller-owned and every index that the reference counts
 * from 1 (guest type, molecule choice) is also passed 1-based here.
 *
 * Error convention: every function returns 0 on success or a positive code.
 * Where the reference would have called error(idnode, code) (error.f) the
 * same code is returned instead of terminating the process; codes >= 9000
 * are this library's own (HIP runtime failures, bad arguments).
 */
#ifndef FASTMC_H
#define FASTMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FMC_MAX_SPECIES 8  /* molecule types (maxmls), framework + guests */
#define FMC_MAX_SITES 8    /* sites of one guest molecule (reference cap mxguestsite=50) */
#define FMC_MAX_TYPES 24   /* unique atom labels (ntpatm) */
#define FMC_NMOVES 8       /* ins del dis jmp swp tra rot swi, gcmc.f:341-366 */

/* error codes (error.f numbering where the reference has one) */
#define FMC_OK 0
#define FMC_ERR_CUTOFF_WIDTH 95      /* rcut exceeds half the cell width, readinputs.f:504 */
#define FMC_ERR_CELL 410             /* inconsistent cell vectors, readinputs.f:425-461 */
#define FMC_ERR_ALLOC 1000           /* array allocation, utility_pack.f:1131 / ewald_module.f:83 */
#define FMC_ERR_NO_EWALD 2311        /* no ewald directive, utility_pack.f:499-501 */
#define FMC_ERR_INSERT_ITER 2320     /* insert_guests exceeded its attempts, mc_moves.f:125 */
#define FMC_ERR_ARG 9001             /* invalid argument / size beyond a compiled cap */
#define FMC_ERR_STATE 9002           /* call out of sequence (e.g. accept without a trial) */
#define FMC_ERR_HIP 9003             /* HIP runtime failure */
#define FMC_ERR_IMCON 9004           /* boundary condition other than imcon 1, 2, 3 */
#define FMC_ERR_SCRATCH 9005         /* request larger than a fixed scratch buffer */
#define FMC_ERR_TEAM 9006            /* a team member's sums did not arrive (a bounded wait expired) */
#define FMC_OVF_TEAM 8               /* fmc_walker_stats.overflow bit of FMC_ERR_TEAM */

/* precision modes */
#define FMC_FP64 0   /* every term in FP64 (parity mode) */
#define FMC_MIXED 1  /* pair and phase arithmetic in FP32; S(k), energies, ln g in FP64 */

/* move kinds as recorded in traces (gcmc.f:821-844) */
#define FMC_MOVE_INSERT 1
#define FMC_MOVE_DELETE 2
#define FMC_MOVE_DISPLACE 3   /* translate + rotate (mc_displace default branch) */
#define FMC_MOVE_TRANSLATE 4
#define FMC_MOVE_ROTATE 5
#define FMC_MOVE_JUMP 6       /* mc_jump, gcmc.f:1937-2039 */
#define FMC_MOVE_SWAP 7       /* mc_swap, gcmc.f:2269-2452 */
#define FMC_MOVE_SWITCH 8     /* mc_switch, gcmc.f:2041-2267 */

/* One molecule type of the FIELD file (readinputs.f:114-281), in FIELD order. */
typedef struct fmc_species {
  int is_guest;               /* 1 for an &guest block, 0 for framework */
  int natoms;                 /* ATOMS (numatoms) */
  int nmols;                  /* NUMMOLS (nummols) */
  const double *site_charge;  /* [natoms] atmchg */
  const double *site_mass;    /* [natoms] atmwght */
  const int *site_type;       /* [natoms] ltpsit, 1-based index into the atom-label list */
  const int *site_frozen;     /* [natoms] lfzsite */
  const double *template_xyz; /* guests: [3*natoms] COM-centred template (guestx/y/z) */
  const double *xyz;          /* [3*natoms*nmols] CONFIG coordinates (molxxx/y/z) or NULL */
} fmc_species;

/* The parsed FIELD / CONFIG / CONTROL content that the trial path needs. */
typedef struct fmc_system {
  int imcon;                  /* 1 cubic, 2 orthorhombic, 3 parallelepiped */
  double cell[9];             /* cell vectors a, b, c (CONFIG rows) */
  int nspecies;               /* maxmls */
  const fmc_species *species; /* [nspecies] */
  int ntypes;                 /* ntpatm */
  int npairs;                 /* ntpvdw: VDW table rows in FIELD order */
  const int *pair_i;          /* [npairs] 1-based atom types */
  const int *pair_j;
  const int *pair_kind;       /* [npairs] ltpvdw: 2 = lj (only kind on this path) */
  const double *pair_eps;     /* [npairs] eps already in internal units (x engunit) */
  const double *pair_sig;     /* [npairs] sigma in A */
  double rcut;                /* cutoff (A) */
  double rvdw;                /* vdw cutoff for the table grid (A), defaults to rcut */
  double delr;                /* neighbour skin (A), guestlistgen */
  double alpha;               /* Ewald convergence parameter (1/A) */
  int kmax[3];                /* Ewald kmax1..3 */
  double engunit;             /* energy unit in internal units (418.4 for kcal) */
  double epsq;                /* relative dielectric constant (1.0) */
  double overlap;             /* overlap distance (A, not squared) */
  double surftol;             /* surface tolerance (statistics only) */
  double sumchg;              /* total system charge after readfield */
} fmc_system;

/* Per-walker (state point) settings: temperature, fugacities, RNG seeds, moves. */
typedef struct fmc_walker {
  double temperature;                 /* K */
  double fugacity[FMC_MAX_SPECIES];   /* Pa, per guest type (gstfuga) */
  float move_cum[FMC_MAX_SPECIES][FMC_NMOVES]; /* REAL(4) cumulative move fractions, gcmc.f:359-366 */
  double delr[FMC_MAX_SPECIES];       /* initial displacement step per guest (delr) */
  int seed[4];                        /* duni seeds i, j, k, l (utility_pack.f:2801-2804) */
  const double *framework_xyz;        /* NULL: the system's framework; else [3*nframework] copy */
} fmc_walker;

/* One trial, as the reference trial routines report it. */
typedef struct fmc_trial {
  double estep;                     /* energy change in FIELD units (mc_moves.f:251-252 etc.) */
  double delE[FMC_MAX_SPECIES];     /* per-species energy change (delE) */
  double recip, real, vdw, tail;    /* components in FIELD units, for diagnostics */
  int overlap;                      /* loverlap */
  int surface;                      /* surface_check (mc_moves.f:10-44): bit 0 the new / inserted
                                       position (lnewsurf, lsurf), bit 1 the existing one (linitsurf) */
} fmc_trial;

/* One GCMC step of a device-driven chain (the fields a trace replay compares). */
typedef struct fmc_trace_rec {
  int step, iguest, move, choice, nmols, overlap, accepted, drew_rande;
  double estep, rande;
  double xyz[3 * FMC_MAX_SITES];
  double energy[FMC_MAX_SPECIES];
  double delE[FMC_MAX_SPECIES];
  double delrdisp, rotangle, rota_rotangle;
  int surface;                       /* fmc_trial.surface of the step's trial */
  int nsurface;                      /* surfacemols of the moved guest's species after the step */
} fmc_trace_rec;

/* Per-walker running totals read back after a run. */
typedef struct fmc_walker_stats {
  int64_t steps, accepted[9], attempted[9];  /* index by move kind FMC_MOVE_* (1..8) */
  int nmols[FMC_MAX_SPECIES];
  int overflow;                              /* bit 0: guest slots full; bit 1: window buffer full;
                                                bit 2: numguests records full; bit 3
                                                (FMC_OVF_TEAM): a team wait gave up */
  double energy[FMC_MAX_SPECIES];            /* cumulative per-species energy (FP64) */
  double sum_n[FMC_MAX_SPECIES];             /* sum of N over counted steps */
  double sum_n2[FMC_MAX_SPECIES];
  double sum_e, sum_e2;                      /* total guest energy moments */
  int64_t counted;                           /* steps that entered the sums */
  int64_t resums;                            /* in-kernel FP64 S(k) re-sums done */
  double resum_drift;                        /* largest |S_total| correction a re-sum made */
  int surface[FMC_MAX_SPECIES];              /* surfacemols by guest type (mc_moves.f:526,568; gcmc.f:1907-1912) */
  int64_t prodcount;                         /* production steps counted (gcmc.f:1052) */
  int64_t windows;                           /* averaging windows closed (avcount, gcmc.f:1095) */
} fmc_walker_stats;

typedef struct fmc_handle fmc_handle;

/* ---- lifecycle -------------------------------------------------------- */

/* Build device state for nwalkers independent chains sharing one chemistry.
 * Replaces the start-up of gcmc.f:229-548 (initscan .. erfcgen, single_point):
 * k-vector enumeration (ewald.f:151-222), erfc table (erfcgen, ewald.f:1-30),
 * LJ table (forgen, vdw_terms.f:181-234), tail terms (lrcorrect,
 * vdw_terms.f:468-684), full S(k) (ewald1, ewald.f:354-661) and intramolecular
 * exclusions (ewald3, ewald.f:735-790; gcmc.f:2662-2761). */
int fmc_create(const fmc_system *sys, int nwalkers, const fmc_walker *walkers,
               int precision, fmc_handle **out);
int fmc_destroy(fmc_handle *h);

/* Sizes the handle derived: newld (number of k-vectors), mxegrd, totatm. */
int fmc_info(const fmc_handle *h, int *newld, int *mxegrd, int *nframework, int *maxmls);

/* Full-system energy of walker w (single_point, gcmc.f:2549-2661):
 * out[0] electrostatic (ecoul), out[1] reciprocal (ewald1), out[2] real-space
 * pairs, out[3] vdw incl. tail, out[4] ewald3 exclusions; FIELD units. */
int fmc_single_point(fmc_handle *h, int w, double out[5]);

/* Re-sum S(k) of walker w in FP64 from its coordinates (ewald1, ewald.f:354-661)
 * and return the largest |S_incremental - S_resum| found before overwriting. */
int fmc_full_resum(fmc_handle *h, int w, double *max_drift);

/* Replace the molecules of guest type iguest in walker w by nmols molecules
 * whose sites are xyz [3*natoms*nmols] (molecule order), the way gcmc.f
 * writes molxxx/y/z and nummols back after a rejected swap
 * (gcmc.f:2416-2437) or switch (gcmc.f:2238-2262), then rebuild what follows
 * from the coordinates as the start-up does (gcmc.f:542-582): molecule and
 * atom-type counts, the guest kind's charge and self-interaction sums, the
 * tail sums (lrcorrect, vdw_terms.f:468-684), every species row and the total
 * of S(k) in FP64 (ewald1, ewald.f:354-661) and the per-species energies
 * (single_point).  sumchg: the system charge to hold from now on, or NULL to
 * keep the walker's (mc_swap does not restore it).  Surface counts are kept.
 * A pending seam trial of walker w is discarded. */
int fmc_load_guests(fmc_handle *h, int w, int iguest, const double *xyz, int nmols, const double *sumchg);

/* ---- one trial at a time (the mc_moves.f seam) ------------------------- */

/* insertion (mc_moves.f:150-276): xyz = [3*natoms] trial coordinates (newx/y/z). */
int fmc_trial_insert(fmc_handle *h, int w, int iguest, const double *xyz, fmc_trial *out);
/* deletion (mc_moves.f:375-484): molecule `choice` (1-based) of guest type iguest. */
int fmc_trial_delete(fmc_handle *h, int w, int iguest, int choice, fmc_trial *out);
/* displace_guest (gcmc.f:2816-2901) with guest_energy (mc_moves.f:277-374)
 * before and after: xyz_new = [3*natoms] displaced coordinates. */
int fmc_trial_displace(fmc_handle *h, int w, int iguest, int choice, const double *xyz_new,
                       fmc_trial *out);
/* accept_move / reject_move (mc_moves.f:485-673) for the last trial of walker w. */
int fmc_accept(fmc_handle *h, int w);
int fmc_reject(fmc_handle *h, int w);

/* Read guest coordinates of walker w: [3*natoms*nmols] of guest type iguest. */
int fmc_get_guests(fmc_handle *h, int w, int iguest, double *xyz, int cap_mols, int *nmols);
/* Read the FP64 structure factor S(k) (total row) of walker w: [2*newld] (re, im). */
int fmc_get_sk(fmc_handle *h, int w, double *sk);
/* Per-walker statistics (counts, N and E moments). */
int fmc_get_stats(fmc_handle *h, int w, fmc_walker_stats *out);
int fmc_reset_stats(fmc_handle *h);

/* ---- device-driven chains (gcmc.f:704-1045 main loop) ------------------ */

/* Advance every walker by nsteps GCMC steps on the device: move choice,
 * trial energy, Metropolis test (energy_eval, utility_pack.f:3637-3693) and
 * accept/reject, with the duni generator (utility_pack.f:2768-2857) per
 * walker.  If trace is non-NULL, one walker (walker 0 unless
 * fmc_set_trace_walker chose another) writes one record per step
 * (trace[nsteps]).  count_stats: bit 0 (FMC_STATS_COUNT) makes every step a
 * production step (gcmc.f:1050-1130: the N/E moments and the averaging
 * windows set by fmc_set_sampling); bit 1 (FMC_STATS_CLOSE) closes the open
 * window after the last step, as the reference does on its final step
 * (lgchk false, gcmc.f:1090-1092). */
#define FMC_STATS_COUNT 1
#define FMC_STATS_CLOSE 2
int fmc_run(fmc_handle *h, int nsteps, fmc_trace_rec *trace, int count_stats);

/* Same, captured once into a HIP graph of `sweeps_per_graph` launches and
 * replayed; nsweeps launches of nsteps_per_sweep steps in total. */
int fmc_run_graph(fmc_handle *h, int nsweeps, int nsteps_per_sweep, int count_stats);

/* The walker whose steps a traced fmc_run records (no counterpart in the
 * reference, which runs one chain per process). */
int fmc_set_trace_walker(fmc_handle *h, int w);

/* Time the last fmc_run / fmc_run_graph dispatches with HIP events on the
 * handle's stream: total milliseconds and launch count. */
int fmc_last_timing(fmc_handle *h, double *ms, int *launches);

/* Launch-shape control: threads per walker block (64..512, multiple of 64);
 * 0 = autotune on a snapshot of the walkers, over the block size and the
 * team size (below, up to the limit fmc_set_team last set). */
int fmc_set_block(fmc_handle *h, int threads);
int fmc_get_block(const fmc_handle *h, int *threads);

/* Workgroups per walker in fmc_run / fmc_run_graph (1..16): the step's pair
 * terms (the atoms of the neighbour loop, ewald.f:689-726 and
 * vdw_terms.f:418-460) and reciprocal terms (the k-vectors of ewald1_guest,
 * ewald.f:155-315) are split over the team, whose members exchange their
 * sums once per step; proposal and decision run in every member alike.  All
 * team workgroups must be resident together: FMC_ERR_ARG if the launch would
 * exceed what the device holds at once with the current block and LDS plan.
 * A run whose exchange wait gave up returns FMC_ERR_TEAM and flags the
 * walker (overflow bit FMC_OVF_TEAM).  Also the autotuner's upper limit;
 * 0 restores the autotuner's full range.  No reference counterpart (the
 * reference runs one chain per MPI rank). */
int fmc_set_team(fmc_handle *h, int team);
int fmc_get_team(const fmc_handle *h, int *team);

/* Full FP64 re-sum of each walker's S(k) (ewald1, ewald.f:354-661) inside the
 * device chain after every `every` accepted moves of that walker; 0 = never.
 * Defaults: 1000 in FMC_MIXED (re-syncs the FP32-updated phases), 0 in
 * FMC_FP64 (the reference never re-sums). */
int fmc_set_resum(fmc_handle *h, int every);

/* ---- Wang-Landau flat-histogram sampling in N (wang_landau.f) ---------- */

/* Parameters of the CONTROL directives wl_max_n, wl_min_n, wl_flat_coeff,
 * wl_visit_tolerance, wl_precision, wl_prec_tolerance (readinputs.f:725-745;
 * defaults gcmc.f:166-176) plus the window count (the reference's MPI ranks,
 * divide_jobs wang_landau.f:1742-1781) and "create molecules" of guest 1. */
typedef struct fmc_wl_params {
  int maxn, minn;          /* N range sampled (default 200, 0) */
  int nwindows;            /* N-windows; walkers are split over them in order, evenly */
  int visittol;            /* a window is flat once every bin has this many visits (10000) */
  double flatcoeff;        /* ... or once min >= flatcoeff * mean (0.7) */
  double wlprec;           /* initial f, ln g += ln f per visit (e) */
  double prectol;          /* a walker stops once ln f < prectol (1e-8) */
  int create_molecules;    /* preload to max(create_molecules, window minimum) (insert_guests) */
  int pad;
} fmc_wl_params;

/* Per-walker Wang-Landau status. */
typedef struct fmc_wl_status {
  int minmol, maxmol;      /* the walker's window of N */
  int done, error;         /* converged (ln f < prectol); error code (2320: preload failed) */
  int64_t sweeps;          /* flat histograms reached (f -> sqrt f each) */
  int64_t steps;           /* steps taken */
  double f, lnf, lambda3;  /* current f, ln f; thermal wavelength cubed (A^3) */
} fmc_wl_status;

/* Switch every walker of the handle to Wang-Landau sampling of the number of
 * guest-1 molecules (wang_landau_sim, wang_landau.f:9-411): moves of guest 1
 * are accepted with eval_wl_move (540-588), ln g(N) and the visits are kept
 * in FP64 / int per walker, windows from divide_jobs.  The reference stops
 * with error 2318 for more than one guest type; here other guest types keep
 * their grand-canonical moves at their fugacities.  fmc_run / fmc_run_graph
 * then advance the walkers; a converged walker stops. */
int fmc_wl_setup(fmc_handle *h, const fmc_wl_params *p);
/* Read walker w: ln g [maxn+1], visits [maxn+1] (either may be NULL). */
int fmc_wl_get(fmc_handle *h, int w, double *lng, int *visits, fmc_wl_status *st);
/* Window range of window `win` of `nwindows` (divide_jobs, wang_landau.f:1742-1781). */
int fmc_wl_window(int win, int nwindows, int maxn, int minn, int *minmol, int *maxmol, int *tail);
/* Stitch per-window ln g (each [maxn+1], window order) into one ln g(N)
 * (stitch_branches, wang_landau.f:1548-1621). */
int fmc_wl_stitch(const double *lng_windows, int nwindows, int maxn, int minn, double *lng_out);
/* Merge the handle's walkers: per window, the walkers' ln g aligned at the
 * window's first bin and averaged; then stitched. lng_out [maxn+1]. */
int fmc_wl_result(fmc_handle *h, double *lng_out);
/* <N>(p) from ln g (compute_isotherm / grand_canonical_partition,
 * wang_landau.f:1623-1740) with log-sum-exp in FP64 for the reference's
 * multiprecision sums.  p in Pa.  units 0: the reference's chemical potential
 * beta*mu = ln(p * lambda3 * beta) with lambda3 in A^3 and beta in mol/kcal;
 * units 1: beta*mu = ln(p * lambda3 * 1e-30 / (kB T)) (SI ideal gas). */
int fmc_wl_isotherm(const double *lng, int maxn, int minn, double temperature, double lambda3,
                    const double *p_pa, int np, int units, double *n_out);

/* ---- energy gradient with respect to framework parameters -------------- */

/* theta: the framework site charges and the LJ parameters of the atom types.
 * Every LJ pair with a framework type is mixed from the type parameters
 * (Lorentz-Berthelot: eps_ij = sqrt(eps_i eps_j), sig_ij = (sig_i + sig_j)/2);
 * guest-guest pairs keep their FIELD rows. */
typedef struct fmc_theta {
  const double *fw_charge;   /* [nframework] framework site charges, FIELD/CONFIG order */
  const double *type_eps;    /* [ntypes] LJ epsilon per atom type, internal units (x engunit) */
  const double *type_sig;    /* [ntypes] LJ sigma per atom type (A) */
} fmc_theta;

/* Total energy (single_point, gcmc.f:2549-2661) of nconf guest configurations
 * in the handle's (rigid, frozen) framework, and its gradient with respect to
 * theta, from one set of device kernels (no counterpart in the reference).
 * Configuration c: nmols[c*ntpguest + g-1] molecules of guest g; site
 * coordinates at xyz + 3*gstride*c, guest types in order, molecules in order.
 * Outputs in FIELD units, any may be NULL: U[nconf], dU_dq[nconf][nframework],
 * dU_deps[nconf][ntypes], dU_dsig[nconf][ntypes] (guest-type entries are 0).
 * precision FMC_FP64, or FMC_MIXED: the reciprocal contraction over (atom, k)
 * in FP32 with FP64 accumulation. */
int fmc_energy_grad(fmc_handle *h, const fmc_theta *theta, int nconf, const int *nmols, const double *xyz,
                    int gstride, int precision, double *U, double *dU_dq, double *dU_deps, double *dU_dsig);

/* Widom insertions on a grid: widom_grid, gcmc.f:2454-2548, with the grid
 * set-up of gcmc.f:258-281,605-610 (CONTROL "henry", "widom insertions N",
 * "grid spacing D", "grid factor a b c"; utility_pack.f:390,436-442,
 * readinputs.f:598-599).  At every grid point (i/wngrida, j/wngridb,
 * k/wngridc), i = 1..ngrida etc., nstep random orientations (random_rot,
 * 4 duni draws each, from walker w's generator) of guest iguest's template
 * are inserted into walker w's current state and rejected.  Outputs:
 * avg_boltz = sum over non-overlapping attempts of exp(-estep / (kboltz T)),
 * divided by all attempts (chainstats(istat+7)); henry = the printed Henry's
 * constant avg_boltz / avo / boltz / T * 1e5 (gcmc.f:1319-1321); per attempt,
 * in the reference's loop order, estep (0 on overlap) and the overlap flag if
 * the arrays are given (attempts = ngrida*ngridb*ngridc*nstep entries).
 * emap_out / cmap_out (both or neither, [ngrida*ngridb*ngridc], c fastest):
 * the energy map of the 'e' probability plot (storeenprob,
 * utility_pack.f:550-620): every non-overlapping attempt's estep and 1
 * spread by equitable_bin at the molecule's centre of mass in the unit
 * cell (cell / gridfactor); writeenprob prints emap / cmap. */
typedef struct fmc_widom_result {
  double avg_boltz, henry;
  long long attempts, accepted;  /* accepted: attempts without overlap */
  int ngrid[3], wngrid[3];
} fmc_widom_result;

int fmc_widom(fmc_handle *h, int w, int iguest, int nstep, double griddim, const int *gridfactor,
              double *estep_out, int *overlap_out, double *emap_out, double *cmap_out, fmc_widom_result *res);

/* fmc_widom that allocates no device memory: no counterpart in the
 * reference; it is for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same arguments
 * and results.  Its draws, per-attempt staging, block sums and maps live in
 * one fixed device array of FMC_WIDOM_SCRATCH_BYTES on each GPU the library
 * runs on (held until the process ends; fmc_create resolves its address);
 * attempts go FMC_WIDOM_SCRATCH_CHUNK per launch, and an energy map of more
 * than FMC_WIDOM_SCRATCH_MAP grid points returns FMC_ERR_SCRATCH.  Every call
 * on one GPU uses the same array: calls that overlap on one GPU, from any
 * handles or threads, race. */
#define FMC_WIDOM_SCRATCH_BYTES (4 << 20)
#define FMC_WIDOM_SCRATCH_CHUNK 65536
#define FMC_WIDOM_SCRATCH_MAP 143360
int fmc_widom_noalloc(fmc_handle *h, int w, int iguest, int nstep, double griddim, const int *gridfactor,
                      double *estep_out, int *overlap_out, double *emap_out, double *cmap_out,
                      fmc_widom_result *res);

/* ---- production statistics (gcmc.f:1050-1130, 1148-1258, 1458-1557) ---- */

/* Averaging windows of the production statistics: a window closes when the
 * production count reaches a multiple of nwindsteps (initscan's
 * nwindsteps = ceiling(steps / nwind), utility_pack.f:453-470); with
 * accepted_only only accepted steps are counted ('sampling' directive,
 * readinputs.f:608-613).  Each walker keeps up to max_windows closed windows
 * (the reference allocates nwind*20, utility_pack.f:1091).  Resets the
 * production counters and the windows of every walker. */
int fmc_set_sampling(fmc_handle *h, int nwindsteps, int accepted_only, int max_windows);

/* One closed window as the kernel records it: the number of steps it counted
 * (rollcount) and, per guest type, the sums of N, E, E*N, N*N, E*E and the
 * surface count over those steps (chainstats, gcmc.f:1062-1075). */
#define FMC_WIN_STATS 6
typedef struct fmc_window {
  double count;
  double sums[FMC_MAX_SPECIES][FMC_WIN_STATS];
} fmc_window;

/* Copy walker w's closed windows (at most cap) into out; *nwin = count. */
int fmc_get_windows(fmc_handle *h, int w, fmc_window *out, int cap, int *nwin);

/* The reference's final statistics of one guest type on one node: window
 * averages with Q_st and C_v per window (calc_Qst / calc_Cv,
 * utility_pack.f:2916-2956, Rgas = 8.314472/4184), their count-weighted
 * means and standard deviations (avg, std: the node report, gcmc.f:1156-1236,
 * 1276-1311) and the final report's union over the one node (avg_final,
 * std_final: avunion / stdunion, gcmc.f:1591-1664, 1512-1535).  Index order:
 * <N>, <E>, <E*N>, <N*N>, <E*E>, <surface N>, Q_st, C_v. */
typedef struct fmc_guest_stats {
  double avg[8];
  double std[8];
  double avg_final[8];
  double std_final[8];
} fmc_guest_stats;

/* Host arithmetic only (no device access): the statistics above from nwin
 * windows of ntpguest guest types at temperature temp (K). */
int fmc_stats_from_windows(const fmc_window *win, int nwin, int ntpguest, double temp, fmc_guest_stats *out);

/* fmc_get_windows + fmc_stats_from_windows at walker w's temperature;
 * out[ntpguest]. */
int fmc_production_stats(fmc_handle *h, int w, fmc_guest_stats *out);

/* ---- numguests.out (gcmc.f:1138-1163) ---------------------------------- */

/* One record of the guest-count trace, written after every step whose total
 * step count (gcmccount) is a multiple of nnumg: per guest type nummols,
 * energy, delE (zero after a rejected move) and the step's move flags ins,
 * del, dis (1 displace, 2 translate, 3 rotate), jmp, swp, swi as the
 * reference writes them to numguests.out (format (i9,i7,2f20.6,7i4)). */
typedef struct fmc_numguests_rec {
  int64_t step;
  int nmols[FMC_MAX_SPECIES];
  int moves[FMC_MAX_SPECIES][6];
  double energy[FMC_MAX_SPECIES];
  double delE[FMC_MAX_SPECIES];
} fmc_numguests_rec;

/* Record every nnumg steps (0: off), up to cap records per walker (more set
 * fmc_walker_stats.overflow bit 2).  Clears the records of every walker. */
int fmc_set_numguests(fmc_handle *h, int nnumg, int cap);
int fmc_get_numguests(fmc_handle *h, int w, fmc_numguests_rec *out, int cap, int *n);

/* ---- probability plots of guest positions (storeprob, utility_pack.f:622-687) */

/* One plot: the guest type and its site list (site 0: the molecule's centre
 * of mass; else a 1-based site), the CONTROL 'probability' lines
 * (readinputs.f:642-672). */
#define FMC_MAX_PLOTS 8
typedef struct fmc_prob_plot {
  int iguest;
  int nsites;
  int sites[FMC_MAX_SITES + 1];
} fmc_prob_plot;

/* Turn on the plots: every production step (fmc_run with FMC_STATS_COUNT)
 * bins the centre of mass or sites of every guest molecule of each plot's
 * type with equitable_bin's 8-corner weights (utility_pack.f:2407-2519) into
 * that walker's grid.  Grid: spacing griddim (A) and 'grid factor' unit cells
 * per axis (gcmc.f:258-281), positions folded into the unit cell
 * (cell / gridfactor, gcmc.f:299-306).  ngrid_out receives the grid size.
 * Clears every walker's grids.  The energy plot ('e') is not provided. */
int fmc_set_prob(fmc_handle *h, double griddim, const int gridfactor[3], int nplots, const fmc_prob_plot *plots,
                 int ngrid_out[3]);

/* Walker w's grid of plot p (0-based): [ngrid a * ngrid b * ngrid c], c
 * fastest, the reference's 1-based bin b at out[b-1]; weights summed over the
 * production steps (the cube files divide by voxel volume and step count). */
int fmc_get_prob(fmc_handle *h, int w, int plot, double *out);

/* Step clock (a profiling build: libfastmc_hip_phases.so, compiled with
 * FMC_PHASE_CLOCK).  No counterpart in the reference, whose timer covers the
 * whole loop (gcmc.f:1461-1462).  The run kernel of that build adds, for
 * walker 0 (team member 0), the time of each phase of every GCMC step in
 * ticks of the GPU's 100 MHz real-time counter: ticks[0] statistics and
 * records at the step's end, [1] proposal (or commit of a speculative one),
 * [2] barrier, [3] pair terms (thread 0's share), [4] wait for the phase
 * tables, [5] reciprocal
 * terms (thread 0's share), [6] wave sums, [7] barrier, [8] cross-wave sums,
 * [9] team exchange, [10] until the decision starts, [11] decision, [12]
 * counters, [13] barrier, [14] accept; ticks[16] counts the steps and
 * ticks[17] those that took a speculative proposal (ticks[15] is unused).
 * n >= 18; clear != 0 zeroes the sums.  A library built without the clock
 * returns FMC_ERR_ARG. */
int fmc_phase_clock(unsigned long long *ticks, int n, int clear);

#ifdef __cplusplus
}
#endif
#endif /* FASTMC_H */
