// Production statistics of the host side: the reference's end-of-run
// arithmetic over the averaging windows that the device chains record.
// Host code only (no device access); included by fastmc_hip.hip.
#pragma once

#include <cmath>

#include "../../include/fastmc.h"

namespace fmc_stats {

constexpr double KBOLTZ = 1.9872041e-3;          // kcal/mol/K, utility_pack.f:121
constexpr double RGAS = 8.314472 / 4184.0;       // kcal/mol/K, utility_pack.f:130

// calc_Qst, utility_pack.f:2916-2928
inline double calc_qst(double E, double N, double N2, double EN, double temp) {
  double q = 0.0;
  const double denom = N2 - N * N;
  if (std::fabs(denom) > 0.0) q = -1.0 * (EN - E * N) / denom + RGAS * temp;
  return q;
}

// calc_Cv, utility_pack.f:2943-2956
inline double calc_cv(double E2, double E, double N, double N2, double EN, double temp) {
  double c = 0.0;
  const double denom = N2 - N * N;
  if (std::fabs(denom) > 0.0 && N > 0.0) {
    const double d = EN - E * N;
    c = (3.0 / 2.0 * RGAS) + (E2 - (E * E) - d * d / denom) / (N * KBOLTZ * temp * temp);
  }
  return c;
}

}  // namespace fmc_stats

extern "C" int fmc_stats_from_windows(const fmc_window *win, int nwin, int ntpguest, double temp,
                                      fmc_guest_stats *out) {
  using namespace fmc_stats;
  if ((!win && nwin > 0) || nwin < 0 || !out || ntpguest < 1 || ntpguest > FMC_MAX_SPECIES) return FMC_ERR_ARG;
  // the node's weight: chainstats(1) = prodcount, every counted step of the
  // closed windows (gcmc.f:1054,1155)
  double weight = 0.0;
  for (int j = 0; j < nwin; ++j) weight += win[j].count;
  for (int g = 0; g < ntpguest; ++g) {
    // window averages with Q_st and C_v of each window (gcmc.f:1096-1118)
    // and their count-weighted means and deviations (gcmc.f:1156-1236)
    double avg[8] = {0, 0, 0, 0, 0, 0, 0, 0}, sd[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double sumweight = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
      sumweight = 0.0;
      for (int j = 0; j < nwin; ++j) {
        const double rc = win[j].count;
        const double *c = win[j].sums[g];
        double v[8];
        v[0] = c[0] / rc;  // <N>
        v[1] = c[1] / rc;  // <E>
        v[2] = c[2] / rc;  // <E*N>
        v[3] = c[3] / rc;  // <N*N>
        v[4] = c[4] / rc;  // <E*E>
        v[5] = c[5] / rc;  // <surface N>
        v[6] = calc_qst(v[1], v[0], v[3], v[2], temp);
        v[7] = calc_cv(v[4], v[1], v[0], v[3], v[2], temp);
        const double sw = rc;
        sumweight = sumweight + sw;
        for (int q = 0; q < 8; ++q) {
          if (pass == 0) {
            avg[q] = avg[q] + sw * v[q];
          } else {
            const double d = avg[q] - v[q];
            sd[q] = sd[q] + sw * (d * d);
          }
        }
      }
      for (int q = 0; q < 8; ++q) {
        if (pass == 0) avg[q] = avg[q] / sumweight;
        else sd[q] = std::sqrt(sd[q] / sumweight);
      }
    }
    // the final report over the one node: avunion / stdunion
    // (gcmc.f:1591-1664) with the node weight
    fmc_guest_stats &o = out[g];
    for (int q = 0; q < 8; ++q) {
      o.avg[q] = avg[q];
      o.std[q] = sd[q];
      o.avg_final[q] = (0.0 + weight * avg[q]) / weight;
    }
    for (int q = 0; q < 8; ++q) {
      const double t = 0.0 + weight * (sd[q] * sd[q] + avg[q] * avg[q]);
      o.std_final[q] = std::sqrt((t / weight) - o.avg_final[q] * o.avg_final[q]);
    }
  }
  return FMC_OK;
}
