/*
 * TEST INFRASTRUCTURE ONLY.  The oracle is a plain-C FP64 restatement of the
 * reference FastMC trial-move energy path, used by tests/, smoke() and the
 * cpu_baseline leg of bench.py to compare results with.  The product
 * (libfastmc_hip.so) never links, loads or calls it.
 */
#ifndef FASTMC_ORACLE_H
#define FASTMC_ORACLE_H

#include "../include/fastmc.h"

typedef struct orc_state orc_state;

orc_state *orc_create(const fmc_system *sys, const fmc_walker *walker);
void orc_destroy(orc_state *s);

int orc_info(const orc_state *s, int *newld, int *mxegrd, int *totatm, int *maxmls);
int orc_single_point(orc_state *s, double out[5]);

int orc_trial_insert(orc_state *s, int iguest, const double *xyz, fmc_trial *out);
int orc_trial_delete(orc_state *s, int iguest, int choice, fmc_trial *out);
int orc_trial_displace(orc_state *s, int iguest, int choice, const double *xyz_new, fmc_trial *out);
int orc_accept(orc_state *s);
int orc_reject(orc_state *s);

int orc_run(orc_state *s, int nsteps, fmc_trace_rec *trace);
/* orc_run with the production block (gcmc.f:1050-1130); flags as fmc_run's
 * count_stats (FMC_STATS_COUNT, FMC_STATS_CLOSE). */
int orc_run_stats(orc_state *s, int nsteps, fmc_trace_rec *trace, int flags);
int orc_set_sampling(orc_state *s, int nwindsteps, int accepted_only);
int orc_get_windows(const orc_state *s, fmc_window *out, int cap, int *nwin, long long *prodcount);
/* probability plots (storeprob, utility_pack.f:622-687) on production steps;
 * orc_get_prob writes plot p's grid with bin b at out[b-1] */
int orc_set_prob(orc_state *s, const int ngrid[3], const int gridfactor[3], int nplots, const fmc_prob_plot *plots);
int orc_get_prob(const orc_state *s, int plot, double *out);
/* numguests.out records every nnumg steps of the chain (gcmc.f:1138-1163) */
int orc_set_numguests(orc_state *s, int nnumg);
int orc_get_numguests(const orc_state *s, fmc_numguests_rec *out, int cap, int *n);
/* surfacemols by guest type */
int orc_get_surface(const orc_state *s, int *surf);
/* attempts / acceptances by move kind (index FMC_MOVE_*), summed over guests */
int orc_move_counts(const orc_state *s, long long *att, long long *acc);

/* Wang-Landau walker over N of guest 1 (wang_landau.f). */
int orc_wl_setup(orc_state *s, int minmol, int maxmol, int maxn, int preload, double wlprec, double flat,
                 int visittol, double prectol);
int orc_wl_run(orc_state *s, int nsteps, fmc_trace_rec *trace);
int orc_wl_get(orc_state *s, double *lng, int *hist, long long *ist, double *dst);

/* Widom insertions on a grid (gcmc.f:2454-2548); out[10] as documented in the source. */
int orc_widom(orc_state *s, int iguest, int nstep, double griddim, const int gf[3], double out[10], double *estep,
              int *ovl, double *emap, double *cmap);

int orc_get_guests(orc_state *s, int iguest, double *xyz, int cap_mols, int *nmols);
int orc_get_sk(orc_state *s, double *sk);
int orc_get_table(orc_state *s, int which, int k, double *out, int n);
int orc_get_kvec(orc_state *s, int *lmn, double *akk);
int orc_get_energy(orc_state *s, double *energy);
int orc_full_resum(orc_state *s, double *max_drift);

/* duni (utility_pack.f:2768-2857) draws after seeding with (i, j, k, l). */
void orc_duni_seq(const int seed[4], int n, double *out);

#endif
