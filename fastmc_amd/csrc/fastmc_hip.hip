// libfastmc_hip.so: C-ABI of the MI355X GCMC trial-move path (include/fastmc.h).
//
// Host side: builds the read-only tables once per handle (the erfc and LJ
// grids, the k-vector list, the tail constants), allocates the per-walker
// device state, and launches the kernels of fmc_kernels.hpp on one HIP stream
// per handle.  No computation of the trial path happens on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "fmc_grad.hpp"
#include "fmc_widom.hpp"
#include "fmc_stats.hpp"

using namespace fmc;

namespace {

#define HIPCHK(x)                                                                      \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) {                                                            \
      fprintf(stderr, "fastmc_hip: %s failed: %s\n", #x, hipGetErrorString(e_));      \
      return FMC_ERR_HIP;                                                              \
    }                                                                                  \
  } while (0)

__global__ void k_fill_zero(double *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

struct GraphKey {
  int nsweeps, nsteps, count;
  bool operator<(const GraphKey &o) const {
    if (nsweeps != o.nsweeps) return nsweeps < o.nsweeps;
    if (nsteps != o.nsteps) return nsteps < o.nsteps;
    return count < o.count;
  }
};

}  // namespace

// Fixed device scratch of fmc_widom_noalloc: one copy on every GPU the
// library runs on, held until the process ends.  On the host the name is only
// a stand-in for the device symbol: host code never passes it to a kernel or a
// copy, it uses the address that scratch_address resolves per device.
static __device__ __align__(256) unsigned char scratch[FMC_WIDOM_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
void *g_scratch_addr[64] = {};

// Device address of `scratch` on the current device, resolved once per device.
int scratch_address(void **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return FMC_ERR_HIP;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "fastmc_hip: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
      return FMC_ERR_HIP;
    }
    g_scratch_addr[dev] = p;
  }
  *out = g_scratch_addr[dev];
  return FMC_OK;
}

}  // namespace

struct fmc_handle {
  int nwalkers = 0, precision = FMC_FP64, ns = 0, block = 512;
  int team = 1;                  // workgroups per walker in fmc_run (k_run)
  int xch_team = 0;              // team whose exchanges the records hold (0: cleared)
  int variant_ovr = -1;          // k_run variant chosen by the autotuner (-1: the LDS plan's)
  int tune_team_max = KTEAM;     // largest team the block autotuner tries
  void *scratch = nullptr;       // this device's address of `scratch` (fmc_widom_noalloc)
  int trace_w = 0;               // walker whose steps fmc_run records
  int nsite = 0;                 // 3: every guest has at most 3 sites (straight-line pair code)
  DevParams hp;                  // host copy
  DevParams *dp = nullptr;       // device copy
  DevWalkers dw{};
  std::vector<void *> allocs;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  int last_launches = 0;
  std::vector<WState> ws_init;
  std::map<GraphKey, hipGraphExec_t> graphs;
  // pending seam trial (host shadow)
  int pend_kind[1] = {0};
  std::vector<int> pend_state;   // per walker: 0 none, 1 valid trial, 2 overlapped trial
  fmc_trace_rec *dtrace = nullptr;
  int dtrace_cap = 0;
  double *dscratch = nullptr;    // small device scratch for reductions (allocated at create)
  int nspecies_total = 0;
  // host mirrors for single_point
  std::vector<double> fw_q;
  double selfq2_all_fw = 0.0;
  std::vector<int> kperm;        // internal k -> reference enumeration index
  size_t lds_base_end = 0;       // LDS bytes of the always-staged part of the plan
  // the plan with the LJ columns alone after the always-staged part
  int ljo_off64 = 0, ljo_off32 = 0;
  size_t ljo_end = 0;
  bool ljo_fits = false;
  size_t plan_end = 0;           // end of the plan before the Wang-Landau arrays and the re-sum rows
  int wl_maxn = 0, wl_minn = 0, wl_nwin = 0;  // Wang-Landau set-up (0: off)
  // energy-gradient entry: framework types and frozen flags, FIELD LJ rows
  std::vector<int> fw_type, fw_frz, fw_spec;
  std::vector<double> pair_eps, pair_sig;   // per FIELD row
  std::vector<int> pair_row;                // [KTYPE+1][KTYPE+1] -> FIELD row or -1
  double rcut = 0.0;
  // its device workspace (grown on demand)
  float2 *g_efw32 = nullptr;
  double2 *g_efw64 = nullptr, *g_sfw = nullptr, *g_sc = nullptr;
  int g_cap = 0, g_gcap = 0, g_nptcap = 0, g_efw_ready = 0;
  int *g_nmols = nullptr, *g_ptab = nullptr, *g_fwt = nullptr;
  double *g_xyz = nullptr, *g_fwq = nullptr, *g_out = nullptr, *g_vt = nullptr;
  double *g_dqpart = nullptr;  // reciprocal d/dq partial sums by k range
  size_t g_dqcap = 0;
};

namespace {

template <typename T>
int dalloc(fmc_handle *h, T **p, size_t n) {
  void *q = nullptr;
  if (n == 0) n = 1;
  if (hipMalloc(&q, n * sizeof(T)) != hipSuccess) return FMC_ERR_ALLOC;
  h->allocs.push_back(q);
  *p = (T *)q;
  return 0;
}

// dcell, utility_pack.f:1904-1960 (0-based in and out)
void dcell(const double *a, double *b) {
  b[0] = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  b[1] = std::sqrt(a[3] * a[3] + a[4] * a[4] + a[5] * a[5]);
  b[2] = std::sqrt(a[6] * a[6] + a[7] * a[7] + a[8] * a[8]);
  double axb1 = a[1] * a[5] - a[2] * a[4], axb2 = a[2] * a[3] - a[0] * a[5], axb3 = a[0] * a[4] - a[1] * a[3];
  double bxc1 = a[4] * a[8] - a[5] * a[7], bxc2 = a[5] * a[6] - a[3] * a[8], bxc3 = a[3] * a[7] - a[4] * a[6];
  double cxa1 = a[7] * a[2] - a[1] * a[8], cxa2 = a[0] * a[8] - a[2] * a[6], cxa3 = a[1] * a[6] - a[0] * a[7];
  b[9] = std::fabs(a[0] * bxc1 + a[1] * bxc2 + a[2] * bxc3);
  b[6] = b[9] / std::sqrt(bxc1 * bxc1 + bxc2 * bxc2 + bxc3 * bxc3);
  b[7] = b[9] / std::sqrt(cxa1 * cxa1 + cxa2 * cxa2 + cxa3 * cxa3);
  b[8] = b[9] / std::sqrt(axb1 * axb1 + axb2 * axb2 + axb3 * axb3);
}

// invert, utility_pack.f:1962-2004 (0-based)
void invert(const double *a, double *b) {
  b[0] = a[4] * a[8] - a[5] * a[7];
  b[1] = a[2] * a[7] - a[1] * a[8];
  b[2] = a[1] * a[5] - a[2] * a[4];
  b[3] = a[5] * a[6] - a[3] * a[8];
  b[4] = a[0] * a[8] - a[2] * a[6];
  b[5] = a[2] * a[3] - a[0] * a[5];
  b[6] = a[3] * a[7] - a[4] * a[6];
  b[7] = a[1] * a[6] - a[0] * a[7];
  b[8] = a[0] * a[4] - a[1] * a[3];
  double d = a[0] * b[0] + a[3] * b[1] + a[6] * b[2];
  double r = std::fabs(d) > 0.0 ? 1.0 / d : 0.0;
  for (int i = 0; i < 9; ++i) b[i] *= r;
}

void host_min_image(const DevParams &P, double &dx, double &dy, double &dz) {
  if (P.imcon == 1) {
    dx -= P.cell[0] * std::round(P.inv_cell[0] * dx);
    dy -= P.cell[0] * std::round(P.inv_cell[0] * dy);
    dz -= P.cell[0] * std::round(P.inv_cell[0] * dz);
  } else if (P.imcon == 2) {
    dx -= P.cell[0] * std::round(P.inv_cell[0] * dx);
    dy -= P.cell[4] * std::round(P.inv_cell[1] * dy);
    dz -= P.cell[8] * std::round(P.inv_cell[2] * dz);
  } else {
    const double *r = P.rcell, *c = P.cell;
    double sx = r[0] * dx + r[3] * dy + r[6] * dz, sy = r[1] * dx + r[4] * dy + r[7] * dz, sz = r[2] * dx + r[5] * dy + r[8] * dz;
    sx -= std::round(sx);
    sy -= std::round(sy);
    sz -= std::round(sz);
    dx = c[0] * sx + c[3] * sy + c[6] * sz;
    dy = c[1] * sx + c[4] * sy + c[7] * sz;
    dz = c[2] * sx + c[5] * sy + c[8] * sz;
  }
}

// ewald3 (ewald.f:735-790) of one guest template: the intramolecular
// exclusion energy subtracted once per guest (gcmc.f:2691-2716).
double ewald3_template(const DevParams &P, int mol, int g, int natms) {
  const double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741;
  const double a4 = -1.453152027, a5 = 1.061405429, pp = 0.3275911;
  const double rr3 = 0.333333333333, r10 = 0.1, r42 = 0.02380952381, r216 = 4.62962962963e-3;
  double tot = 0.0, alpha = P.alpha;
  for (int i = 0; i < natms - 1; ++i) {
    double chgea = P.site_q[mol][i] / P.epsq * R4PIE0, e = 0.0;
    for (int j = i + 1; j < natms; ++j) {
      double dx = P.tmpl[g][i][0] - P.tmpl[g][j][0], dy = P.tmpl[g][i][1] - P.tmpl[g][j][1],
             dz = P.tmpl[g][i][2] - P.tmpl[g][j][2];
      host_min_image(P, dx, dy, dz);
      double chgprd = chgea * P.site_q[mol][j];
      double rrr = std::sqrt(dx * dx + dy * dy + dz * dz), alpr = rrr * alpha, alpr2 = alpr * alpr, erfr;
      if (alpr < 1e-2) {
        erfr = 2.0 * chgprd * (alpha / SQRPI) * (1.0 + alpr2 * (-rr3 + alpr2 * (r10 + alpr2 * (-r42 + alpr2 * r216))));
      } else {
        double tt = 1.0 / (1.0 + pp * alpha * rrr);
        double ex = std::exp(-(alpha * rrr) * (alpha * rrr));
        erfr = (1.0 - tt * (a1 + tt * (a2 + tt * (a3 + tt * (a4 + tt * a5)))) * ex) * chgprd / rrr;
      }
      e += erfr;
    }
    tot += e;
  }
  return tot;
}

// Workgroups of a k_run launch: one per walker, or with teams 8 * ceil(nw / 8)
// * team (the members of a walker sit on one XCD, see k_run).
int run_grid(const fmc_handle *h, int team) {
  return team > 1 ? 8 * ((h->nwalkers + 7) / 8) * team : h->nwalkers;
}

// The k_run variant the LDS plan selects (its index in allow_lds's list).
int run_variant(const fmc_handle *h) {
  if (h->variant_ovr >= 0) return h->variant_ovr;
  if (!h->hp.lds_all && h->hp.lj_lds) return 1;
  if (!h->hp.lds_all) return 0;
  return h->nsite == 3 ? 2 : 3;
}

template <int NS, typename T>
const void *run_fn(int variant) {
  switch (variant) {
    case 1: return (const void *)k_run<NS, T, false, 0, true>;
    case 2: return (const void *)k_run<NS, T, true, 3, false>;
    case 3: return (const void *)k_run<NS, T, true, 0, false>;
    default: return (const void *)k_run<NS, T, false, 0, false>;
  }
}

template <int NS, typename T>
void launch_run(fmc_handle *h, int nsteps, int count) {
  const int team = h->team, nw = h->nwalkers;
  dim3 grid(run_grid(h, team)), block(h->block);
  size_t lds = h->hp.lds_bytes;
  switch (run_variant(h)) {
    case 1:
      hipLaunchKernelGGL((k_run<NS, T, false, 0, true>), grid, block, lds, h->stream, h->dp, h->dw, nsteps, count,
                         team, nw);
      break;
    case 2:
      hipLaunchKernelGGL((k_run<NS, T, true, 3, false>), grid, block, lds, h->stream, h->dp, h->dw, nsteps, count,
                         team, nw);
      break;
    case 3:
      hipLaunchKernelGGL((k_run<NS, T, true, 0, false>), grid, block, lds, h->stream, h->dp, h->dw, nsteps, count,
                         team, nw);
      break;
    default:
      hipLaunchKernelGGL((k_run<NS, T, false, 0, false>), grid, block, lds, h->stream, h->dp, h->dw, nsteps, count,
                         team, nw);
  }
}

// Workgroups of the current k_run variant the device holds at once with
// `block` threads each (the bound a team launch must stay within: a member
// waits on the others, so all of them must be resident together).
int run_capacity(const fmc_handle *h, int block) {
  const void *fn = nullptr;
  const bool mixed = h->precision == FMC_MIXED;
  const int v = run_variant(h);
  switch (h->ns) {
    case 2: fn = mixed ? run_fn<2, float>(v) : run_fn<2, double>(v); break;
    case 3: fn = mixed ? run_fn<3, float>(v) : run_fn<3, double>(v); break;
    default: fn = mixed ? run_fn<4, float>(v) : run_fn<4, double>(v); break;
  }
  int per_cu = 0, dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, h->hp.lds_bytes) != hipSuccess) return 0;
  return per_cu * prop.multiProcessorCount;
}

// A team launch must fit on the device at once.
int team_fits(const fmc_handle *h) {
  if (h->team <= 1) return FMC_OK;
  return run_grid(h, h->team) <= run_capacity(h, h->block) ? FMC_OK : FMC_ERR_ARG;
}

// After a team run: a wait that gave up (the walkers it touched carry
// overflow bit FMC_OVF_TEAM) makes the call fail; the flag is cleared.
int team_error(fmc_handle *h) {
  if (h->team <= 1) return FMC_OK;
  int e = 0;
  HIPCHK(hipMemcpy(&e, h->dw.xerr, sizeof(int), hipMemcpyDeviceToHost));
  if (!e) return FMC_OK;
  // members may have stopped at different exchanges: clear every record and
  // count, so that nothing of this run can pass for a later exchange (the
  // flagged walkers' states are not to be trusted)
  HIPCHK(hipMemset(h->dw.xch, 0, sizeof(double) * (size_t)h->nwalkers * 2 * KTEAM * XREC));
  HIPCHK(hipMemset(h->dw.xcnt, 0, sizeof(long long) * (size_t)h->nwalkers));
  HIPCHK(hipMemset(h->dw.xerr, 0, sizeof(int)));
  return FMC_ERR_TEAM;
}

// Before a team run: the records hold only exchanges of the current team (a
// record that another team size left behind could carry a current tag), so
// they are cleared when the team changes.  Outside any graph capture.
int team_records(fmc_handle *h) {
  if (h->team <= 1 || h->team == h->xch_team) return FMC_OK;
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemset(h->dw.xch, 0, sizeof(double) * (size_t)h->nwalkers * 2 * KTEAM * XREC));
  h->xch_team = h->team;
  return FMC_OK;
}

// The private guest arrays of members 1 .. team - 1 (grown on demand).
int team_arrays(fmc_handle *h, int team) {
  if (team <= h->dw.team_alloc) return FMC_OK;
  double *p = nullptr;
  const size_t n = (size_t)h->nwalkers * (team - 1) * 3 * h->hp.gatoms_total;
  if (hipMalloc(&p, n * sizeof(double)) != hipSuccess) return FMC_ERR_ALLOC;
  if (h->dw.gxt) {
    for (auto it = h->allocs.begin(); it != h->allocs.end(); ++it)
      if (*it == (void *)h->dw.gxt) {
        h->allocs.erase(it);
        break;
      }
    (void)hipFree(h->dw.gxt);
  }
  h->allocs.push_back(p);
  h->dw.gxt = p;
  h->dw.team_alloc = team;
  return FMC_OK;
}

size_t smem_size(int ns) {
  switch (ns) {
    case 2: return sizeof(Smem<2>);
    case 3: return sizeof(Smem<3>);
    default: return sizeof(Smem<4>);
  }
}

template <int NS, typename T>
hipError_t allow_lds(int bytes) {
  const void *fns[] = {(const void *)k_run<NS, T, false, 0, false>, (const void *)k_run<NS, T, false, 0, true>,
                       (const void *)k_run<NS, T, true, 3, false>, (const void *)k_run<NS, T, true, 0, false>,
                       (const void *)k_trial<NS, T, false>, (const void *)k_trial<NS, T, true>,
                       (const void *)k_accept<NS>};
  for (const void *f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t allow_lds_dispatch(int ns, int mixed, int bytes) {
  switch (ns) {
    case 2: return mixed ? allow_lds<2, float>(bytes) : allow_lds<2, double>(bytes);
    case 3: return mixed ? allow_lds<3, float>(bytes) : allow_lds<3, double>(bytes);
    default: return mixed ? allow_lds<4, float>(bytes) : allow_lds<4, double>(bytes);
  }
}

void run_dispatch(fmc_handle *h, int nsteps, int count) {
  bool mixed = h->precision == FMC_MIXED;
  switch (h->ns) {
    case 2: mixed ? launch_run<2, float>(h, nsteps, count) : launch_run<2, double>(h, nsteps, count); break;
    case 3: mixed ? launch_run<3, float>(h, nsteps, count) : launch_run<3, double>(h, nsteps, count); break;
    default: mixed ? launch_run<4, float>(h, nsteps, count) : launch_run<4, double>(h, nsteps, count); break;
  }
}

// The seam trial takes the LJ arithmetic of the handle's chain kernel: FP32
// columns (mixed mode) where the chains stage them in LDS, FP64 otherwise.
template <int NS, typename T>
void launch_trial(fmc_handle *h, int w) {
  if (run_variant(h) == 1)
    hipLaunchKernelGGL((k_trial<NS, T, true>), dim3(1), dim3(h->block), h->hp.lds_bytes, h->stream, h->dp, h->dw, w);
  else
    hipLaunchKernelGGL((k_trial<NS, T, false>), dim3(1), dim3(h->block), h->hp.lds_bytes, h->stream, h->dp, h->dw, w);
}

void trial_dispatch(fmc_handle *h, int w) {
  bool mixed = h->precision == FMC_MIXED;
  switch (h->ns) {
    case 2: mixed ? launch_trial<2, float>(h, w) : launch_trial<2, double>(h, w); break;
    case 3: mixed ? launch_trial<3, float>(h, w) : launch_trial<3, double>(h, w); break;
    default: mixed ? launch_trial<4, float>(h, w) : launch_trial<4, double>(h, w); break;
  }
}

void accept_dispatch(fmc_handle *h, int w) {
  switch (h->ns) {
    case 2: hipLaunchKernelGGL((k_accept<2>), dim3(1), dim3(h->block), h->hp.lds_bytes, h->stream, h->dp, h->dw, w); break;
    case 3: hipLaunchKernelGGL((k_accept<3>), dim3(1), dim3(h->block), h->hp.lds_bytes, h->stream, h->dp, h->dw, w); break;
    default: hipLaunchKernelGGL((k_accept<4>), dim3(1), dim3(h->block), h->hp.lds_bytes, h->stream, h->dp, h->dw, w); break;
  }
}

// FP64 re-sum of S(k): every walker (only < 0) or walker `only`
void resum_dispatch(fmc_handle *h, double *drift, int only = -1) {
  dim3 grid((h->hp.nk + 255) / 256, only < 0 ? h->nwalkers : 1);
  const int w0 = only < 0 ? 0 : only;
  switch (h->ns) {
    case 2: hipLaunchKernelGGL((k_resum<2>), grid, dim3(256), 0, h->stream, h->dp, h->dw, drift, w0); break;
    case 3: hipLaunchKernelGGL((k_resum<3>), grid, dim3(256), 0, h->stream, h->dp, h->dw, drift, w0); break;
    default: hipLaunchKernelGGL((k_resum<4>), grid, dim3(256), 0, h->stream, h->dp, h->dw, drift, w0); break;
  }
}

int walker_natoms(fmc_handle *h, int w, std::vector<WState> &ws) {
  int n = h->hp.nfw;
  for (int g = 1; g <= h->hp.ntpguest; ++g) {
    int mol = h->hp.locguest[g];
    n += ws[w].nmols[mol] * h->hp.natoms[mol];
  }
  return n;
}

// The per-walker device arrays of walker w alone: a launch through this view
// addresses walker w as walker 0.
DevWalkers walker_view(const fmc_handle *h, int w) {
  const DevParams &P = h->hp;
  DevWalkers dw1 = h->dw;
  dw1.ws = h->dw.ws + w;
  dw1.gx = h->dw.gx + (size_t)w * P.gatoms_total;
  dw1.gy = h->dw.gy + (size_t)w * P.gatoms_total;
  dw1.gz = h->dw.gz + (size_t)w * P.gatoms_total;
  dw1.sk = h->dw.sk + (size_t)w * (P.maxmls + 1) * P.nk;
  if (P.fw_per_walker) {
    dw1.fwx = h->dw.fwx + (size_t)w * P.nfw;
    dw1.fwy = h->dw.fwy + (size_t)w * P.nfw;
    dw1.fwz = h->dw.fwz + (size_t)w * P.nfw;
  }
  return dw1;
}

// single_point for every walker (only < 0) or walker `only`: species-resolved
// recip, real and vdw sums.  out[w][0..4] as fmc_single_point; also
// refreshes ws[w].energy (gcmc.f:554-582).
int single_points(fmc_handle *h, std::vector<WState> &ws, std::vector<double> &out, bool set_energy,
                  int only = -1) {
  const DevParams &P = h->hp;
  const int NW = h->nwalkers, NC = KMAXC + 1;
  const int wlo = only < 0 ? 0 : only, whi = only < 0 ? NW : only + 1;
  std::vector<double> rec((size_t)NW * NC), pairs((size_t)NW * 2 * NC, 0.0);
  double *d = h->dscratch;
  HIPCHK(hipMemsetAsync(d, 0, sizeof(double) * (size_t)NW * 3 * NC, h->stream));
  {
    const DevWalkers dv = only < 0 ? h->dw : walker_view(h, only);
    double *o = d + (size_t)wlo * NC;
    const dim3 grid(whi - wlo);
    switch (h->ns) {
      case 2: hipLaunchKernelGGL((k_recip_sums<2>), grid, dim3(256), 0, h->stream, h->dp, dv, o); break;
      case 3: hipLaunchKernelGGL((k_recip_sums<3>), grid, dim3(256), 0, h->stream, h->dp, dv, o); break;
      default: hipLaunchKernelGGL((k_recip_sums<4>), grid, dim3(256), 0, h->stream, h->dp, dv, o); break;
    }
  }
  // k_pair_sums walks one walker's own atoms per launch
  for (int w = wlo; w < whi; ++w) {
    int n = walker_natoms(h, w, ws);
    dim3 grid((n + 255) / 256, 1);
    const DevWalkers dw1 = walker_view(h, w);
    double *o = d + (size_t)NW * NC + (size_t)w * 2 * NC;
    switch (h->ns) {
      case 2: hipLaunchKernelGGL((k_pair_sums<2>), grid, dim3(256), 0, h->stream, h->dp, dw1, n, o); break;
      case 3: hipLaunchKernelGGL((k_pair_sums<3>), grid, dim3(256), 0, h->stream, h->dp, dw1, n, o); break;
      default: hipLaunchKernelGGL((k_pair_sums<4>), grid, dim3(256), 0, h->stream, h->dp, dw1, n, o); break;
    }
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(rec.data(), d, sizeof(double) * (size_t)NW * NC, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipMemcpyAsync(pairs.data(), d + (size_t)NW * NC, sizeof(double) * (size_t)NW * 2 * NC,
                        hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  out.assign((size_t)NW * 5, 0.0);
  const int NS = h->ns;
  const double C = 2.0 * P.rvolm * R4PIE0 / P.epsq;
  for (int w = wlo; w < whi; ++w) {
    WState &s = ws[w];
    // ewald1 finalisation, ewald.f:626-643
    double q2 = h->selfq2_all_fw;
    for (int g = 1; g <= P.ntpguest; ++g) q2 += s.nmols[P.locguest[g]] * P.selfq2[g];
    double engsicold = -R4PIE0 / P.epsq * P.alpha * q2 / SQRPI;
    double qfix = -(0.5 * PI * R4PIE0 / P.epsq) * ((s.sumchg / P.alpha) * (s.sumchg / P.alpha) / P.volm);
    double ewald1sum = C * rec[(size_t)w * NC] + engsicold + qfix;
    double e1[KMAXC + 1] = {0};
    for (int i = 1; i <= NS; ++i) {
      double fkk = 1.0;
      for (int j = 1; j <= i; ++j) {
        if (i == j) fkk = 0.5;
        int ik = loc2(i, j);
        double sic = -R4PIE0 / P.epsq * P.alpha * s.engsic[ik] / SQRPI;
        sic = sic - ((fkk * PI * R4PIE0 / P.epsq) * s.chgsum_mol[i] * s.chgsum_mol[j] / (P.alpha * P.alpha * P.volm));
        e1[ik] = C * rec[(size_t)w * NC + ik] + sic;
      }
    }
    double e2sum = 0.0, evsum = 0.0, e3 = 0.0;
    const double *pe2 = pairs.data() + (size_t)w * 2 * NC, *pev = pe2 + NC;
    for (int t = 1; t <= KMAXC; ++t) {
      e2sum += pe2[t];
      evsum += pev[t];
    }
    for (int g = 1; g <= P.ntpguest; ++g) e3 += s.nmols[P.locguest[g]] * P.ewald3en[P.locguest[g]];
    out[(size_t)w * 5 + 0] = (e2sum + ewald1sum - e3) / P.engunit;
    out[(size_t)w * 5 + 1] = ewald1sum / P.engunit;
    out[(size_t)w * 5 + 2] = e2sum / P.engunit;
    out[(size_t)w * 5 + 3] = (evsum + s.elrc) / P.engunit;
    out[(size_t)w * 5 + 4] = e3 / P.engunit;
    if (set_energy) {
      for (int g = 1; g <= P.ntpguest; ++g) {
        int mol = P.locguest[g];
        double sp = 0.0;
        for (int j = 1; j <= NS; ++j) {
          int ik = loc2(mol, j);
          sp += (e1[ik] + pe2[ik] + s.elrc_mol[ik] + pev[ik]) / P.engunit;
        }
        if (s.nmols[mol] > 0) sp -= s.nmols[mol] * P.ewald3en[mol] / P.engunit;
        s.energy[mol] = sp;
      }
    }
  }
  return 0;
}

// lrcorrect counts and tail (vdw_terms.f:468-684) for one walker's state.
void lrcorrect_counts(const DevParams &P, const fmc_system *sys, const std::vector<int> &fwtype,
                      const std::vector<int> &fwspec, const std::vector<int> &fwfrz, WState &s) {
  memset(s.numtyp, 0, sizeof s.numtyp);
  memset(s.numfrz, 0, sizeof s.numfrz);
  memset(s.numtyp_mol, 0, sizeof s.numtyp_mol);
  memset(s.numfrz_mol, 0, sizeof s.numfrz_mol);
  for (size_t a = 0; a < fwtype.size(); ++a) {
    int ka = fwtype[a], mol = fwspec[a];
    s.numtyp[ka] += 1;
    s.numtyp_mol[mol][ka] += 1;
    if (fwfrz[a]) {
      s.numfrz[ka] += 1;
      s.numfrz_mol[mol][ka] += 1;
    }
  }
  for (int g = 1; g <= P.ntpguest; ++g) {
    int mol = P.locguest[g];
    for (int m = 0; m < s.nmols[mol]; ++m)
      for (int i = 0; i < P.natoms[mol]; ++i) {
        int ka = P.site_t[mol][i];
        s.numtyp[ka] += 1;
        s.numtyp_mol[mol][ka] += 1;
        if (P.site_f[mol][i]) {
          s.numfrz[ka] += 1;
          s.numfrz_mol[mol][ka] += 1;
        }
      }
  }
  const double twopi = 2.0 * PI, volm = P.volm;
  s.elrc = 0.0;
  for (int t = 0; t <= KMAXC; ++t) s.elrc_mol[t] = 0.0;
  int ivdw = 0;
  for (int i = 1; i <= P.ntypes; ++i)
    for (int j = 1; j <= i; ++j) {
      ++ivdw;
      double eadd = P.steadd[ivdw];
      double denprd = twopi * ((double)s.numtyp[i] * (double)s.numtyp[j] - (double)s.numfrz[i] * (double)s.numfrz[j]) / (volm * volm);
      s.elrc += volm * denprd * eadd;
      double e0[KMAXC + 1] = {0};
      for (int it = 1; it <= P.maxmls; ++it)
        for (int jt = 1; jt <= P.maxmls; ++jt) {
          int kt = loc2(it, jt);
          e0[kt] += twopi * ((double)s.numtyp_mol[it][i] * (double)s.numtyp_mol[jt][j] -
                             (double)s.numfrz_mol[it][i] * (double)s.numfrz_mol[jt][j]) / (volm * volm);
        }
      for (int t = 1; t <= KMAXC; ++t) s.elrc_mol[t] += volm * eadd * e0[t];
    }
}

// The LDS left after the plan (from `end`) holds the in-kernel re-sum's
// per-atom phase rows: up to 32 KB, never so much that fewer workgroups fit
// on a CU, none below 8 atoms a batch.  Sets off_rs, rs_batch, lds_bytes.
void plan_resum_rows(DevParams &P, size_t end) {
  const size_t budget = 160 * 1024 - 1024, cu = 160 * 1024;
  end = (end + 15) & ~(size_t)15;
  const size_t per_cu = cu / std::max<size_t>(end, 1);
  size_t cap = per_cu >= 1 ? std::min(budget, cu / per_cu) : budget;
  size_t room = cap > end ? cap - end : 0;
  room = std::min(room, (size_t)32 * 1024);
  const size_t per = (size_t)3 * P.kcp * 16 + 12;
  const int nb = (int)(room / per);
  if (nb >= 8 && !getenv("FMC_NO_RESUM_ROWS")) {
    P.off_rs = (int)end;
    P.rs_batch = nb;
    P.lds_bytes = (int)((end + (size_t)nb * per + 15) & ~(size_t)15);
  } else {
    P.off_rs = 0;
    P.rs_batch = 0;
    P.lds_bytes = (int)end;
  }
}

}  // namespace

extern "C" {

int fmc_create(const fmc_system *sys, int nwalkers, const fmc_walker *walkers, int precision, fmc_handle **out) {
  if (!sys || !walkers || !out || nwalkers < 1) return FMC_ERR_ARG;
  if (sys->imcon < 1 || sys->imcon > 3) return FMC_ERR_IMCON;
  if (sys->nspecies < 2 || sys->nspecies > KMAXS || sys->ntypes > KTYPE || sys->ntypes < 1) return FMC_ERR_ARG;
  if (precision != FMC_FP64 && precision != FMC_MIXED) return FMC_ERR_ARG;
  {
    // swap needs a second guest kind (mc_swap, gcmc.f:2309-2322); switch is a
    // step without a move below three guest kinds (gcmc.f:2098) and is not
    // restated for three or more
    int nguest = 0;
    for (int m = 0; m < sys->nspecies; ++m) nguest += sys->species[m].is_guest ? 1 : 0;
    for (int w = 0; w < nwalkers; ++w)
      for (int g = 0; g < nguest && g < FMC_MAX_SPECIES; ++g) {
        const float *c = walkers[w].move_cum[g];
        if (c[4] > c[3] && nguest < 2) return FMC_ERR_ARG;
        if (c[7] > c[6] && nguest >= 3) return FMC_ERR_ARG;
      }
  }
  fmc_handle *h = new fmc_handle();
  h->nwalkers = nwalkers;
  h->precision = precision;
  h->ns = sys->nspecies;
  DevParams &P = h->hp;
  memset(&P, 0, sizeof P);
  P.imcon = sys->imcon;
  P.maxmls = sys->nspecies;
  P.ntypes = sys->ntypes;
  for (int i = 0; i < 9; ++i) P.cell[i] = sys->cell[i];
  invert(P.cell, P.rcell);
  P.inv_cell[0] = 1.0 / P.cell[0];
  P.inv_cell[1] = 1.0 / P.cell[4];
  P.inv_cell[2] = 1.0 / P.cell[8];
  double celprp[10];
  dcell(P.cell, celprp);
  P.volm = celprp[9];
  P.statvolm = P.volm * 1e-30;
  double width = std::min(std::min(celprp[6], celprp[7]), celprp[8]) / 2.0;
  if (sys->rcut > width) {
    delete h;
    return FMC_ERR_CUTOFF_WIDTH;
  }
  P.rcut2 = sys->rcut * sys->rcut;
  P.rclim2 = (sys->rcut + sys->delr) * (sys->rcut + sys->delr);
  P.overlap2 = sys->overlap * sys->overlap;
  P.alpha = sys->alpha;
  P.epsq = sys->epsq > 0 ? sys->epsq : 1.0;
  P.engunit = sys->engunit;
  P.rengunit = 1.0 / sys->engunit;
  P.rvolm = 2.0 * PI / P.volm;
  P.kmax1 = sys->kmax[0];
  P.kmax2 = sys->kmax[1];
  P.kmax3 = sys->kmax[2];
  if (P.kmax1 >= KCAP || P.kmax2 >= KCAP || P.kmax3 >= KCAP || P.kmax1 < 1 || P.kmax2 < 1 || P.kmax3 < 1) {
    delete h;
    return FMC_ERR_ARG;
  }
  // ---- tables: mxegrd (utility_pack.f:996), erfcgen (ewald.f:1-30), forgen (vdw_terms.f:181-234)
  double rvdw = sys->rvdw > 0 ? sys->rvdw : sys->rcut;
  int mxegrd = std::max(1000, (int)(rvdw / 0.01 + 0.5) + 4);
  P.mxegrd = mxegrd;
  P.drewd = sys->rcut / (double)(mxegrd - 4);
  P.rdrewd = 1.0 / P.drewd;
  P.dlrpot = rvdw / (double)(mxegrd - 4);
  P.rdlrpot = 1.0 / P.dlrpot;
  std::vector<double> erc(mxegrd + 3, 0.0);
  {
    const double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741;
    const double a4 = -1.453152027, a5 = 1.061405429, pp = 0.3275911;
    for (int i = 1; i <= mxegrd; ++i) {
      double rrr = (double)i * P.drewd, tt = 1.0 / (1.0 + pp * P.alpha * rrr), ar = P.alpha * rrr;
      double exp1 = std::exp(-(ar * ar));
      erc[i] = tt * (a1 + tt * (a2 + tt * (a3 + tt * (a4 + tt * a5)))) * exp1 / rrr;
    }
  }
  const int npairs = sys->npairs;
  const int ntri = (P.ntypes * (P.ntypes + 1)) / 2;
  const int maxvdw = std::max(npairs, ntri);
  std::vector<int> lstvdw(maxvdw + 2, 0), ltpvdw(maxvdw + 2, -1);
  std::vector<double> peps(maxvdw + 2, 0.0), psig(maxvdw + 2, 0.0);
  for (int p = 1; p <= npairs; ++p) {
    int i = sys->pair_i[p - 1], j = sys->pair_j[p - 1];
    if (i < 1 || j < 1 || i > P.ntypes || j > P.ntypes) {
      delete h;
      return FMC_ERR_ARG;
    }
    lstvdw[loc2(i, j)] = p;
    ltpvdw[p] = sys->pair_kind[p - 1];
    peps[p] = sys->pair_eps[p - 1];
    psig[p] = sys->pair_sig[p - 1];
  }
  if (npairs < ntri)
    for (int i = 1; i <= ntri; ++i)
      if (lstvdw[i] == 0) lstvdw[i] = npairs + 1;
  P.vstride = mxegrd + 3;
  std::vector<double> vvv((size_t)(npairs + 1) * P.vstride, 0.0);
  for (int p = 1; p <= npairs; ++p) {
    if (ltpvdw[p] != 2) continue;  // only the lj form is on this path (vdw_terms.f:222-234)
    for (int i = 1; i <= mxegrd; ++i) {
      double rrr = (double)i * P.dlrpot, x = psig[p] / rrr, x2 = x * x, sr6 = x2 * x2 * x2;
      vvv[(size_t)(p - 1) * P.vstride + i] = 4.0 * peps[p] * sr6 * (sr6 - 1.0);
    }
  }
  h->pair_eps.assign(peps.begin() + 1, peps.begin() + 1 + npairs);
  h->pair_sig.assign(psig.begin() + 1, psig.begin() + 1 + npairs);
  h->pair_row.assign((KTYPE + 1) * (KTYPE + 1), -1);
  for (int ti = 1; ti <= P.ntypes; ++ti)
    for (int tj = 1; tj <= P.ntypes; ++tj) {
      int k = lstvdw[loc2(ti, tj)];
      if (k >= 1 && k <= npairs && ltpvdw[k] == 2) h->pair_row[ti * (KTYPE + 1) + tj] = k - 1;
    }
  for (int ti = 0; ti <= KTYPE; ++ti)
    for (int tj = 0; tj <= KTYPE; ++tj) P.pairidx[ti][tj] = -1;
  for (int ti = 1; ti <= P.ntypes; ++ti)
    for (int tj = 1; tj <= P.ntypes; ++tj) {
      int k = lstvdw[loc2(ti, tj)];
      // srfrce skip rule, vdw_terms.f:431
      if (k >= 1 && k <= npairs && ltpvdw[k] < 100 && std::fabs(vvv[(size_t)(k - 1) * P.vstride + 1]) > 1e-10)
        P.pairidx[ti][tj] = (int8_t)(k - 1);
    }
  // surface_check thresholds (mc_moves.f:26-40): sigma = prmvdw(ivdw,2) of the
  // type pair's VDW row, zero for a pair with no row of its own
  for (int ti = 0; ti <= KTYPE; ++ti)
    for (int tj = 0; tj <= KTYPE; ++tj) {
      double sg = 0.0;
      if (ti >= 1 && tj >= 1 && ti <= P.ntypes && tj <= P.ntypes) {
        int k = lstvdw[loc2(ti, tj)];
        if (k >= 1 && k <= npairs) sg = psig[k];
      }
      P.surf2[ti][tj] = (sys->surftol + sg) * (sys->surftol + sg);
    }
  P.nwindsteps = 0;
  P.accsample = 0;
  P.maxwin = 0;
  P.nplots = 0;
  P.nnumg = 0;
  P.numg_cap = 0;
  // tail constants, lrcorrect eadd (vdw_terms.f:581-645)
  P.ntri = ntri;
  {
    int ivdw = 0;
    double rc = sys->rcut, rc3 = rc * rc * rc, rc9 = rc3 * rc3 * rc3;
    for (int i = 1; i <= P.ntypes; ++i)
      for (int j = 1; j <= i; ++j) {
        ++ivdw;
        int k = lstvdw[ivdw];
        double eadd = 0.0;
        if (k >= 1 && k <= npairs && ltpvdw[k] == 2) {
          double s2 = psig[k] * psig[k], s6 = s2 * s2 * s2;
          eadd = 4.0 * peps[k] * (s6 * s6 / (9.0 * rc9) - s6 / (3.0 * rc3));
        }
        if (i != j) eadd *= 2.0;
        P.steadd[ivdw] = eadd;
      }
  }
  // ---- k-vectors, ewald.f:87-94,151-222 (half space, 1.05 rcpcut sphere)
  std::vector<int> klmn;
  std::vector<double> akk;
  {
    double rb[10];
    dcell(P.rcell, rb);
    double twopi = 2.0 * PI;
    double rcpcut = std::min(std::min((double)P.kmax1 * rb[6], (double)P.kmax2 * rb[7]), (double)P.kmax3 * rb[8]);
    rcpcut = rcpcut * 1.05 * twopi;
    double rcpct2 = rcpcut * rcpcut, ralph = -0.25 / (P.alpha * P.alpha);
    const double *rc = P.rcell;
    int mmin = 0, nmin = 1;
    for (int ll = 0; ll <= P.kmax1; ++ll) {
      double tmp = twopi * (double)ll;
      double rkx1 = tmp * rc[0], rky1 = tmp * rc[3], rkz1 = tmp * rc[6];
      for (int mm = mmin; mm <= P.kmax2; ++mm) {
        tmp = twopi * (double)mm;
        double rkx2 = rkx1 + tmp * rc[1], rky2 = rky1 + tmp * rc[4], rkz2 = rkz1 + tmp * rc[7];
        for (int nn = nmin; nn <= P.kmax3; ++nn) {
          tmp = twopi * (double)nn;
          double rkx3 = rkx2 + tmp * rc[2], rky3 = rky2 + tmp * rc[5], rkz3 = rkz2 + tmp * rc[8];
          double rksq = rkx3 * rkx3 + rky3 * rky3 + rkz3 * rkz3;
          if (rksq <= rcpct2) {
            klmn.push_back(ll | ((mm + 128) << 8) | ((nn + 128) << 16));
            akk.push_back(std::exp(ralph * rksq) * (1.0 / rksq));
          }
        }
        nmin = -P.kmax3;
      }
      mmin = -P.kmax2;
    }
  }
  P.nk = (int)klmn.size();
  // internal order: by a_k, largest first (the sum over k is order-free; S(k)
  // readback restores the reference's enumeration order)
  {
    std::vector<int> order(P.nk);
    for (int i = 0; i < P.nk; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return akk[a] > akk[b]; });
    std::vector<int> k2(P.nk);
    std::vector<double> a2(P.nk);
    for (int i = 0; i < P.nk; ++i) {
      k2[i] = klmn[order[i]];
      a2[i] = akk[order[i]];
    }
    klmn.swap(k2);
    akk.swap(a2);
    h->kperm = order;
    double split = 1e-4;
    if (const char *e = getenv("FMC_MIXED_AK_SPLIT")) split = atof(e);
    int n64 = P.nk;
    if (precision == FMC_MIXED) {
      n64 = 0;
      while (n64 < P.nk && akk[n64] >= akk[0] * split) ++n64;
    }
    P.nk64 = n64;
  }
  // ---- species
  std::vector<double> fwx, fwy, fwz, fwq;
  std::vector<int> fwinfo, fwtype, fwspec, fwfrz;
  int ng = 0;
  double fwq2 = 0.0;
  for (int m = 1; m <= P.maxmls; ++m) {
    const fmc_species &sp = sys->species[m - 1];
    P.natoms[m] = sp.natoms;
    if (sp.is_guest) {
      if (sp.natoms > KSITE || sp.natoms < 1) {
        delete h;
        return FMC_ERR_ARG;
      }
      ++ng;
      P.locguest[ng] = m;
      P.gofspecies[m] = ng;
      double q2 = 0.0, qs = 0.0;
      for (int i = 0; i < sp.natoms; ++i) {
        P.site_q[m][i] = sp.site_charge[i];
        P.site_m[m][i] = sp.site_mass[i];
        P.site_chgea[m][i] = sp.site_charge[i] / P.epsq * R4PIE0;
        P.site_t[m][i] = sp.site_type[i];
        P.site_f[m][i] = sp.site_frozen[i];
        for (int d = 0; d < 3; ++d) P.tmpl[ng][i][d] = sp.template_xyz[3 * i + d];
        q2 += sp.site_charge[i] * sp.site_charge[i];
        qs += sp.site_charge[i];
        P.numtyp_g[ng][sp.site_type[i]] += 1;
        if (sp.site_frozen[i]) P.numfrz_g[ng][sp.site_type[i]] += 1;
      }
      P.selfq2[ng] = q2;
      P.chgtmp[ng] = qs;
    } else {
      for (int im = 0; im < sp.nmols; ++im)
        for (int i = 0; i < sp.natoms; ++i) {
          size_t a = (size_t)im * sp.natoms + i;
          fwx.push_back(sp.xyz ? sp.xyz[3 * a] : 0.0);
          fwy.push_back(sp.xyz ? sp.xyz[3 * a + 1] : 0.0);
          fwz.push_back(sp.xyz ? sp.xyz[3 * a + 2] : 0.0);
          fwq.push_back(sp.site_charge[i]);
          fwq2 += sp.site_charge[i] * sp.site_charge[i];
          fwinfo.push_back(sp.site_type[i] | (m << 8) | ((sp.site_frozen[i] ? 1 : 0) << 16));
          fwtype.push_back(sp.site_type[i]);
          fwspec.push_back(m);
          fwfrz.push_back(sp.site_frozen[i] ? 1 : 0);
        }
    }
  }
  P.ntpguest = ng;
  if (ng < 1) {
    delete h;
    return FMC_ERR_ARG;
  }
  h->selfq2_all_fw = fwq2;
  P.nfw = (int)fwx.size();
  h->fw_type = fwtype;
  h->fw_frz = fwfrz;
  h->fw_spec = fwspec;
  h->rcut = sys->rcut;
  // decision constants: the same expressions as the per-trial code of
  // ewald1_guest (ewald.f:96-105) and mc_insert (gcmc.f:1729)
  P.dc_C = 2.0 * P.rvolm * R4PIE0 / P.epsq;
  for (int g = 0; g <= KMAXS; ++g) P.dc_esic[g] = 0.0;
  for (int g = 1; g <= ng; ++g) P.dc_esic[g] = -R4PIE0 / P.epsq * P.alpha * P.selfq2[g] / SQRPI;
  h->nsite = 3;
  for (int g = 1; g <= ng; ++g)
    if (P.natoms[P.locguest[g]] > 3) h->nsite = 0;
  if (getenv("FMC_NO_NSITE")) h->nsite = 0;
  // mixed mode re-syncs the FP32-updated S(k) with a full FP64 re-sum after
  // this many accepted moves; the FP64 mode follows the reference (never)
  P.resum_every = precision == FMC_MIXED ? 1000 : 0;
  if (getenv("FMC_RESUM_EVERY")) P.resum_every = std::max(0, atoi(getenv("FMC_RESUM_EVERY")));
  P.spec_prop = getenv("FMC_NO_SPEC") ? 0 : 1;  // speculative proposals in k_run (FMC_NO_SPEC: off)
  // guest capacities: one molecule per 40 A^3 of cell, denser than a liquid
  // of small molecules (liquid CO2 holds one per ~66 A^3); a walker that
  // fills its slots stops inserting and sets overflow bit 0
  int slot = 0;
  for (int g = 1; g <= ng; ++g) {
    int mol = P.locguest[g];
    int cap = (int)std::ceil(P.volm / 40.0);
    cap = std::max(64, std::min(cap, 8192));
    cap = std::max(cap, sys->species[mol - 1].nmols + 64);
    if (const char *e = getenv("FMC_GUEST_CAP")) cap = std::max(sys->species[mol - 1].nmols + 1, atoi(e));
    P.gcap[g] = cap;
    P.gbase[g] = slot;
    slot += cap * P.natoms[mol];
  }
  P.gatoms_total = slot;
  // ewald3 exclusions of each guest template
  for (int g = 1; g <= ng; ++g) P.ewald3en[P.locguest[g]] = ewald3_template(P, P.locguest[g], g, P.natoms[P.locguest[g]]);
  // ---- gstlrcorrect (vdw_terms.f:717-765) as linear coefficients in the guest counts
  {
    const double twopi = 2.0 * PI, volm = P.volm;
    int F[KTYPE + 1] = {0}, Ff[KTYPE + 1] = {0};
    int Fm[KMAXS + 1][KTYPE + 1] = {{0}}, Ffm[KMAXS + 1][KTYPE + 1] = {{0}};
    for (size_t a = 0; a < fwtype.size(); ++a) {
      F[fwtype[a]] += 1;
      Fm[fwspec[a]][fwtype[a]] += 1;
      if (fwfrz[a]) {
        Ff[fwtype[a]] += 1;
        Ffm[fwspec[a]][fwtype[a]] += 1;
      }
    }
    for (int g = 1; g <= ng; ++g) {
      int mol = P.locguest[g];
      for (int lex = 0; lex <= 1; ++lex) {
        double sg = lex ? -1.0 : 1.0;
        double c0 = 0.0, c1[KMAXS + 1] = {0}, m0[KMAXS + 1] = {0}, m1[KMAXS + 1] = {0};
        int ivdw = 0;
        for (int i = 1; i <= P.ntypes; ++i)
          for (int j = 1; j <= i; ++j) {
            ++ivdw;
            double w0 = twopi / volm * P.steadd[ivdw];
            double na = P.numtyp_g[g][i], nb = P.numtyp_g[g][j], fa = P.numfrz_g[g][i], fb = P.numfrz_g[g][j];
            c0 += w0 * (sg * (na * F[j] + nb * F[i] - fa * Ff[j] - fb * Ff[i]) + (na * nb - fa * fb));
            for (int g2 = 1; g2 <= ng; ++g2)
              c1[g2] += w0 * sg * (na * P.numtyp_g[g2][j] + nb * P.numtyp_g[g2][i] - fa * P.numfrz_g[g2][j] -
                                   fb * P.numfrz_g[g2][i]);
            for (int jt = 1; jt <= P.maxmls; ++jt) {
              double ssg = (lex && jt == mol) ? -1.0 : 1.0;
              double self = (jt == mol) ? (na * nb - fa * fb) : 0.0;
              int g2 = P.gofspecies[jt];
              if (g2) {
                m0[jt] += w0 * self;
                m1[jt] += w0 * ssg * (na * P.numtyp_g[g2][j] + nb * P.numtyp_g[g2][i] - fa * P.numfrz_g[g2][j] -
                                      fb * P.numfrz_g[g2][i]);
              } else {
                m0[jt] += w0 * (ssg * (na * Fm[jt][j] + nb * Fm[jt][i] - fa * Ffm[jt][j] - fb * Ffm[jt][i]) + self);
              }
            }
          }
        P.tl_c0[g][lex] = c0;
        for (int g2 = 1; g2 <= ng; ++g2) P.tl_c1[g][lex][g2] = c1[g2];
        for (int jt = 1; jt <= P.maxmls; ++jt) {
          P.tm_c0[g][lex][jt] = m0[jt];
          P.tm_c1[g][lex][jt] = m1[jt];
        }
      }
    }
  }
  // ---- LDS plan.  Always: Smem, erfc table, phase tables, constants block.
  // The all-in-LDS variant also needs the framework, the k list and every
  // guest slot (its LJ columns are read from L2, in FP64); without room for
  // all of them (or with FMC_NO_LDS_ALL set) the kernels read those from
  // global memory and stage the LJ columns of the guest-site types in LDS
  // when they fit.
  {
    const size_t budget = 160 * 1024 - 1024;
    auto al = [](size_t x) { return (x + 15) & ~(size_t)15; };
    P.nsmax = 1;
    for (int g = 1; g <= ng; ++g) P.nsmax = std::max(P.nsmax, P.natoms[P.locguest[g]]);
    P.kcp = std::max(P.kmax1, std::max(P.kmax2, P.kmax3)) + 1;
    size_t off = (smem_size(h->ns) + 255) & ~(size_t)255;
    P.off_erc = (int)off;
    off = al(off + (size_t)8 * (mxegrd + 3));
    P.off_ph = (int)off;
    off = al(off + (size_t)16 * 2 * P.nsmax * 3 * P.kcp);
    P.off_ph32 = (int)off;
    off = al(off + (size_t)8 * 2 * P.nsmax * 3 * P.kcp);
    P.off_cst = (int)off;
    off = al(off + (size_t)CST_BYTES);
    if (off > budget) {
      delete h;
      return FMC_ERR_ARG;
    }
    const size_t base_end = off;
    h->lds_base_end = base_end;
    // LJ columns used by guest sites (FP32 mode only)
    int ncol = 0;
    for (int ti = 0; ti <= KTYPE; ++ti)
      for (int tj = 0; tj <= KTYPE; ++tj) P.lpairidx[ti][tj] = -1;
    std::vector<int> colmap(npairs + 1, -1);
    bool cols_ok = true;
    for (int g = 1; g <= ng; ++g) {
      int m = P.locguest[g];
      for (int s2 = 0; s2 < P.natoms[m]; ++s2) {
        int ti = P.site_t[m][s2];
        for (int tj = 0; tj <= KTYPE; ++tj) {
          int c = P.pairidx[ti][tj];
          if (c < 0) continue;
          if (colmap[c] < 0) {
            if (ncol >= 64) {
              cols_ok = false;
              continue;
            }
            colmap[c] = ncol;
            P.vcol_src[ncol++] = c;
          }
          P.lpairidx[ti][tj] = (int8_t)colmap[c];
        }
      }
    }
    P.nvcol = ncol;
    // the LJ wall segment of mixed mode: from just below the overlap distance
    // (closer pairs overlap) up to the last entry of any guest-site column
    // with |V| above 0.5 kcal/mol, where FP32 storage alone moves a trial
    // energy by more than 1e-7 kT
    {
      const double vwall = 0.5 * 418.4;  // internal energy units (10 J/mol)
      P.lj_lo = std::max(0, (int)std::floor(std::sqrt(P.overlap2) * P.rdlrpot) - 1);
      int hi = P.lj_lo + 1;
      for (int g = 1; g <= ng; ++g)
        for (int s2 = 0; s2 < P.natoms[P.locguest[g]]; ++s2)
          for (int tj = 0; tj <= KTYPE; ++tj) {
            const int c = P.pairidx[P.site_t[P.locguest[g]][s2]][tj];
            if (c < 0) continue;
            for (int i = P.vstride - 1; i > hi; --i)
              if (std::fabs(vvv[(size_t)c * P.vstride + i]) > vwall) {
                hi = i + 1;
                break;
              }
          }
      P.lj_hi = std::min(hi, mxegrd);
      P.lj_n64 = P.lj_hi + 2 - P.lj_lo;
    }
    P.lds_all = 0;
    if (!getenv("FMC_NO_LDS_ALL") && cols_ok) {
      size_t o = base_end;
      int ofw = (int)o;
      o = al(o + (size_t)36 * P.nfw);
      int ok = (int)o;
      o = al(o + (size_t)12 * P.nk);
      int og = (int)o;
      o = al(o + (size_t)24 * P.gatoms_total);
      if (o <= budget) {
        P.lds_all = 1;
        P.off_fw = ofw;
        P.off_k = ok;
        P.off_g = og;
        off = o;
      }
    }
    // without the all-in-LDS plan: the LJ columns alone, when they fit
    // (mixed mode: FP64 wall entries then the whole FP32 columns; FP64
    // mode: the whole FP64 columns)
    {
      const size_t nc = (size_t)std::max(ncol, 1);
      size_t o = base_end;
      h->ljo_off64 = (int)o;
      if (precision == FMC_MIXED) {
        o = al(o + (size_t)8 * nc * P.lj_n64);
        h->ljo_off32 = (int)o;
        o = al(o + (size_t)4 * nc * P.vstride);
      } else {
        h->ljo_off32 = (int)o;
        o = al(o + (size_t)8 * nc * P.vstride);
      }
      h->ljo_end = o;
      h->ljo_fits = cols_ok && o <= budget && !getenv("FMC_NO_LJ_LDS");
    }
    P.lj_lds = 0;
    if (!P.lds_all && h->ljo_fits) {
      P.lj_lds = 1;
      P.off_vvv64 = h->ljo_off64;
      P.off_vvv = h->ljo_off32;
      off = h->ljo_end;
    }
    h->plan_end = al(off);
    plan_resum_rows(P, h->plan_end);
    if (getenv("FMC_VERBOSE"))
      fprintf(stderr, "fastmc: LDS plan %d bytes, all-in-LDS %d, LJ alone in LDS %d, nsite %d, LJ columns %d "
              "(wall %d..%d), guest slots %d, re-sum batch %d\n", P.lds_bytes, P.lds_all, P.lj_lds, h->nsite, ncol,
              P.lj_lo, P.lj_hi, P.gatoms_total, P.rs_batch);
  }
  // ---- device buffers
  HIPCHK(hipStreamCreate(&h->stream));
  // the fixed scratch of fmc_widom_noalloc, resolved once per device (a
  // handle without it only refuses that entry point)
  if (scratch_address(&h->scratch) != FMC_OK) h->scratch = nullptr;
  HIPCHK(allow_lds_dispatch(h->ns, precision == FMC_MIXED, P.lds_bytes));
  HIPCHK(hipEventCreate(&h->ev0));
  HIPCHK(hipEventCreate(&h->ev1));
  int rc = 0;
  double *d_erc, *d_vvv, *d_akk, *d_fwx, *d_fwy, *d_fwz, *d_fwq;
  float *d_erc32, *d_vvv32;
  int *d_klmn, *d_fwinfo;
  rc |= dalloc(h, &d_erc, erc.size());
  rc |= dalloc(h, &d_vvv, vvv.size());
  rc |= dalloc(h, &d_erc32, erc.size());
  rc |= dalloc(h, &d_vvv32, vvv.size());
  rc |= dalloc(h, &d_klmn, klmn.size());
  rc |= dalloc(h, &d_akk, akk.size());
  rc |= dalloc(h, &d_fwx, fwx.size());
  rc |= dalloc(h, &d_fwy, fwy.size());
  rc |= dalloc(h, &d_fwz, fwz.size());
  rc |= dalloc(h, &d_fwq, fwq.size());
  rc |= dalloc(h, &d_fwinfo, fwinfo.size());
  rc |= dalloc(h, &h->dp, 1);
  rc |= dalloc(h, &h->dw.ws, (size_t)nwalkers);
  rc |= dalloc(h, &h->dw.gx, (size_t)nwalkers * P.gatoms_total);
  rc |= dalloc(h, &h->dw.gy, (size_t)nwalkers * P.gatoms_total);
  rc |= dalloc(h, &h->dw.gz, (size_t)nwalkers * P.gatoms_total);
  rc |= dalloc(h, &h->dw.sk, (size_t)nwalkers * (P.maxmls + 1) * P.nk);
  rc |= dalloc(h, &h->dw.sscr, (size_t)nwalkers * 2 * P.nk);
  rc |= dalloc(h, &h->dw.sbak, (size_t)nwalkers * (P.maxmls + 1) * P.nk);
  // team exchange records and counters (zero: no record is current)
  rc |= dalloc(h, &h->dw.xch, (size_t)nwalkers * 2 * KTEAM * XREC);
  rc |= dalloc(h, &h->dw.xcnt, (size_t)nwalkers);
  rc |= dalloc(h, &h->dw.xerr, 1);
  h->dw.gxt = nullptr;
  h->dw.team_alloc = 1;
  rc |= dalloc(h, (char **)&h->dw.trial_out, 4096);
  rc |= dalloc(h, &h->dscratch, (size_t)nwalkers * 3 * (KMAXC + 1) + 64);
  bool per_walker_fw = false;
  for (int w = 0; w < nwalkers; ++w)
    if (walkers[w].framework_xyz) per_walker_fw = true;
  P.fw_per_walker = per_walker_fw ? 1 : 0;
  if (per_walker_fw) {
    rc |= dalloc(h, &h->dw.fwx, (size_t)nwalkers * P.nfw);
    rc |= dalloc(h, &h->dw.fwy, (size_t)nwalkers * P.nfw);
    rc |= dalloc(h, &h->dw.fwz, (size_t)nwalkers * P.nfw);
  }
  if (rc) {
    fmc_destroy(h);
    return FMC_ERR_ALLOC;
  }
  HIPCHK(hipMemset(h->dw.xch, 0, sizeof(double) * (size_t)nwalkers * 2 * KTEAM * XREC));
  HIPCHK(hipMemset(h->dw.xcnt, 0, sizeof(long long) * (size_t)nwalkers));
  HIPCHK(hipMemset(h->dw.xerr, 0, sizeof(int)));
  std::vector<float> erc32(erc.begin(), erc.end()), vvv32(vvv.begin(), vvv.end());
  HIPCHK(hipMemcpy(d_erc, erc.data(), erc.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_vvv, vvv.data(), vvv.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_erc32, erc32.data(), erc32.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_vvv32, vvv32.data(), vvv32.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_klmn, klmn.data(), klmn.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_akk, akk.data(), akk.size() * 8, hipMemcpyHostToDevice));
  if (P.nfw) {
    HIPCHK(hipMemcpy(d_fwx, fwx.data(), fwx.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_fwy, fwy.data(), fwy.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_fwz, fwz.data(), fwz.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_fwq, fwq.data(), fwq.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_fwinfo, fwinfo.data(), fwinfo.size() * 4, hipMemcpyHostToDevice));
  }
  P.erc = d_erc;
  P.vvv = d_vvv;
  P.erc32 = d_erc32;
  P.vvv32 = d_vvv32;
  P.klmn = d_klmn;
  P.akk = d_akk;
  P.fwx = d_fwx;
  P.fwy = d_fwy;
  P.fwz = d_fwz;
  P.fwq = d_fwq;
  P.fwinfo = d_fwinfo;
  HIPCHK(hipMemcpy(h->dp, &P, sizeof P, hipMemcpyHostToDevice));
  if (per_walker_fw) {
    for (int w = 0; w < nwalkers; ++w) {
      std::vector<double> x(P.nfw), y(P.nfw), z(P.nfw);
      for (int a = 0; a < P.nfw; ++a) {
        if (walkers[w].framework_xyz) {
          x[a] = walkers[w].framework_xyz[3 * a];
          y[a] = walkers[w].framework_xyz[3 * a + 1];
          z[a] = walkers[w].framework_xyz[3 * a + 2];
        } else {
          x[a] = fwx[a];
          y[a] = fwy[a];
          z[a] = fwz[a];
        }
      }
      HIPCHK(hipMemcpy(h->dw.fwx + (size_t)w * P.nfw, x.data(), 8 * (size_t)P.nfw, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(h->dw.fwy + (size_t)w * P.nfw, y.data(), 8 * (size_t)P.nfw, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(h->dw.fwz + (size_t)w * P.nfw, z.data(), 8 * (size_t)P.nfw, hipMemcpyHostToDevice));
    }
  }
  // ---- walkers
  std::vector<WState> ws(nwalkers);
  std::vector<double> gx((size_t)P.gatoms_total), gy((size_t)P.gatoms_total), gz((size_t)P.gatoms_total);
  for (int w = 0; w < nwalkers; ++w) {
    WState &s = ws[w];
    memset(&s, 0, sizeof s);
    const fmc_walker &wk = walkers[w];
    // duni seeding, utility_pack.f:2816-2836
    {
      int i = wk.seed[0], j = wk.seed[1], k = wk.seed[2], l = wk.seed[3];
      for (int ii = 1; ii <= 97; ++ii) {
        float sacc = 0.0f, t = 0.5f;
        for (int jj = 1; jj <= 24; ++jj) {
          int m = ((i * j) % 179 * k) % 179;
          i = j;
          j = k;
          k = m;
          l = (53 * l + 1) % 169;
          if ((l * m) % 64 >= 32) sacc = sacc + t;
          t = 0.5f * t;
        }
        s.u[ii] = sacc;
      }
      s.mz_made = 0;
      s.mz_used = 0;
      s.mz_ir = 97;
      s.mz_jr = 33;
      s.mz_c = 362436;
    }
    s.temp = wk.temperature;
    s.beta = 1.0 / (KBOLTZ * wk.temperature);
    for (int g = 1; g <= ng; ++g) {
      s.fug[g] = wk.fugacity[g - 1];
      for (int mv = 0; mv < FMC_NMOVES; ++mv) s.mcf[g][mv] = wk.move_cum[g - 1][mv];
      s.delr[g] = wk.delr[g - 1];
      s.delrdisp[g] = wk.delr[g - 1];
      s.tran_delr[g] = wk.delr[g - 1];
    }
    s.rotangle = PI / 3.0;
    s.rota_rotangle = s.rotangle;
    s.sumchg = sys->sumchg;
    for (int m = 1; m <= P.maxmls; ++m) {
      const fmc_species &sp = sys->species[m - 1];
      s.nmols[m] = sp.is_guest ? sp.nmols : 0;
      double cs = 0.0, c2 = 0.0;
      for (int im = 0; im < sp.nmols; ++im)
        for (int i = 0; i < sp.natoms; ++i) {
          cs += sp.site_charge[i];
          c2 += sp.site_charge[i] * sp.site_charge[i];
        }
      s.chgsum_mol[m] = cs;
      s.engsic[loc2(m, m)] = c2;
    }
    lrcorrect_counts(P, sys, fwtype, fwspec, fwfrz, s);
    std::fill(gx.begin(), gx.end(), 0.0);
    std::fill(gy.begin(), gy.end(), 0.0);
    std::fill(gz.begin(), gz.end(), 0.0);
    for (int g = 1; g <= ng; ++g) {
      int mol = P.locguest[g];
      const fmc_species &sp = sys->species[mol - 1];
      for (int a = 0; a < sp.nmols * sp.natoms && sp.xyz; ++a) {
        gx[P.gbase[g] + a] = sp.xyz[3 * a];
        gy[P.gbase[g] + a] = sp.xyz[3 * a + 1];
        gz[P.gbase[g] + a] = sp.xyz[3 * a + 2];
      }
    }
    HIPCHK(hipMemcpy(h->dw.gx + (size_t)w * P.gatoms_total, gx.data(), 8 * gx.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->dw.gy + (size_t)w * P.gatoms_total, gy.data(), 8 * gy.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->dw.gz + (size_t)w * P.gatoms_total, gz.data(), 8 * gz.size(), hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(h->dw.ws, ws.data(), sizeof(WState) * (size_t)nwalkers, hipMemcpyHostToDevice));
  // full S(k) then the single-point energies (gcmc.f:542-582)
  resum_dispatch(h, nullptr);
  HIPCHK(hipGetLastError());
  std::vector<double> sp;
  if (int e = single_points(h, ws, sp, true)) {
    fmc_destroy(h);
    return e;
  }
  HIPCHK(hipMemcpy(h->dw.ws, ws.data(), sizeof(WState) * (size_t)nwalkers, hipMemcpyHostToDevice));
  h->pend_state.assign(nwalkers, 0);
  *out = h;
  return FMC_OK;
}

int fmc_destroy(fmc_handle *h) {
  if (!h) return FMC_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  for (void *p : h->allocs) (void)hipFree(p);
  if (h->dtrace) (void)hipFree(h->dtrace);
  if (h->dw.win) (void)hipFree(h->dw.win);
  if (h->dw.prob) (void)hipFree(h->dw.prob);
  if (h->dw.numg) (void)hipFree(h->dw.numg);
  for (void *p : {(void *)h->g_efw32, (void *)h->g_efw64, (void *)h->g_sfw, (void *)h->g_sc, (void *)h->g_nmols,
                  (void *)h->g_ptab, (void *)h->g_fwt, (void *)h->g_xyz, (void *)h->g_fwq, (void *)h->g_out,
                  (void *)h->g_vt, (void *)h->g_dqpart})
    if (p) (void)hipFree(p);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return FMC_OK;
}

int fmc_info(const fmc_handle *h, int *newld, int *mxegrd, int *nframework, int *maxmls) {
  if (!h) return FMC_ERR_ARG;
  if (newld) *newld = h->hp.nk;
  if (mxegrd) *mxegrd = h->hp.mxegrd;
  if (nframework) *nframework = h->hp.nfw;
  if (maxmls) *maxmls = h->hp.maxmls;
  return FMC_OK;
}

int fmc_single_point(fmc_handle *h, int w, double out[5]) {
  if (!h || w < 0 || w >= h->nwalkers || !out) return FMC_ERR_ARG;
  std::vector<WState> ws(h->nwalkers);
  HIPCHK(hipMemcpy(ws.data(), h->dw.ws, sizeof(WState) * ws.size(), hipMemcpyDeviceToHost));
  std::vector<double> sp;
  if (int e = single_points(h, ws, sp, false)) return e;
  for (int i = 0; i < 5; ++i) out[i] = sp[(size_t)w * 5 + i];
  return FMC_OK;
}

int fmc_full_resum(fmc_handle *h, int w, double *max_drift) {
  if (!h || w < -1 || w >= h->nwalkers) return FMC_ERR_ARG;
  double *d = h->dscratch;
  HIPCHK(hipMemsetAsync(d, 0, sizeof(double) * h->nwalkers, h->stream));
  resum_dispatch(h, d);
  HIPCHK(hipGetLastError());
  std::vector<double> dr(h->nwalkers);
  HIPCHK(hipMemcpyAsync(dr.data(), d, sizeof(double) * h->nwalkers, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (max_drift) {
    double m = 0.0;
    for (int i = 0; i < h->nwalkers; ++i)
      if (w < 0 || i == w) m = std::max(m, dr[i]);
    *max_drift = m;
  }
  return FMC_OK;
}

int fmc_load_guests(fmc_handle *h, int w, int iguest, const double *xyz, int nmols, const double *sumchg) {
  if (!h || w < 0 || w >= h->nwalkers) return FMC_ERR_ARG;
  const DevParams &P = h->hp;
  if (iguest < 1 || iguest > P.ntpguest || nmols < 0 || nmols > P.gcap[iguest]) return FMC_ERR_ARG;
  if (nmols > 0 && !xyz) return FMC_ERR_ARG;
  if (P.wl_nbins > 0 && iguest == 1 && nmols + 1 > P.wl_nbins) return FMC_ERR_ARG;
  const int mol = P.locguest[iguest], na = P.natoms[mol], n = nmols * na;
  std::vector<WState> ws(h->nwalkers);
  HIPCHK(hipMemcpy(ws.data(), h->dw.ws, sizeof(WState) * ws.size(), hipMemcpyDeviceToHost));
  WState &s = ws[w];
  // the guest kind's slots: the given molecules, zeros past them up to the old count
  const int nold = s.nmols[mol] * na, nput = std::max(n, nold);
  if (nput > 0) {
    std::vector<double> x(nput, 0.0), y(nput, 0.0), z(nput, 0.0);
    for (int a = 0; a < n; ++a) {
      x[a] = xyz[3 * a];
      y[a] = xyz[3 * a + 1];
      z[a] = xyz[3 * a + 2];
    }
    const size_t base = (size_t)w * P.gatoms_total + P.gbase[iguest];
    HIPCHK(hipMemcpy(h->dw.gx + base, x.data(), 8 * (size_t)nput, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->dw.gy + base, y.data(), 8 * (size_t)nput, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->dw.gz + base, z.data(), 8 * (size_t)nput, hipMemcpyHostToDevice));
  }
  // counts and sums that follow from the molecule list, as at start-up
  s.nmols[mol] = nmols;
  s.chgsum_mol[mol] = (double)nmols * P.chgtmp[iguest];
  s.engsic[loc2(mol, mol)] = (double)nmols * P.selfq2[iguest];
  if (sumchg) s.sumchg = *sumchg;
  lrcorrect_counts(P, nullptr, h->fw_type, h->fw_spec, h->fw_frz, s);
  s.pend_valid = 0;
  s.pend_kind = 0;
  h->pend_state[w] = 0;
  HIPCHK(hipMemcpy(h->dw.ws + w, &s, sizeof s, hipMemcpyHostToDevice));
  // S(k) of every species and the total from the coordinates, then the energies
  resum_dispatch(h, nullptr, w);
  HIPCHK(hipGetLastError());
  std::vector<double> sp;
  if (int e = single_points(h, ws, sp, true, w)) return e;
  HIPCHK(hipMemcpy(h->dw.ws + w, &s, sizeof s, hipMemcpyHostToDevice));
  return FMC_OK;
}

static int seam_trial(fmc_handle *h, int w, int kind, int iguest, int choice, const double *xyz, fmc_trial *out) {
  if (!h || w < 0 || w >= h->nwalkers || !out) return FMC_ERR_ARG;
  const DevParams &P = h->hp;
  if (iguest < 1 || iguest > P.ntpguest) return FMC_ERR_ARG;
  int mol = P.locguest[iguest], natms = P.natoms[mol];
  WState s;
  HIPCHK(hipMemcpy(&s, h->dw.ws + w, sizeof s, hipMemcpyDeviceToHost));
  if (kind != FMC_MOVE_INSERT && (choice < 1 || choice > s.nmols[mol])) return FMC_ERR_ARG;
  if (kind == FMC_MOVE_INSERT && s.nmols[mol] >= P.gcap[iguest]) return FMC_ERR_ARG;
  s.pend_kind = kind;
  s.pend_iguest = iguest;
  s.pend_choice = choice;
  s.pend_valid = 0;
  for (int i = 0; i < 3 * KSITE; ++i) s.pend_xyz[i] = 0.0;
  if (xyz)
    for (int i = 0; i < 3 * natms; ++i) s.pend_xyz[i] = xyz[i];
  HIPCHK(hipMemcpy(h->dw.ws + w, &s, sizeof s, hipMemcpyHostToDevice));
  trial_dispatch(h, w);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, h->dw.trial_out, sizeof(fmc_trial), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  h->pend_state[w] = out->overlap && kind != FMC_MOVE_DELETE ? 2 : 1;
  return FMC_OK;
}

int fmc_trial_insert(fmc_handle *h, int w, int iguest, const double *xyz, fmc_trial *out) {
  if (!xyz) return FMC_ERR_ARG;
  return seam_trial(h, w, FMC_MOVE_INSERT, iguest, 0, xyz, out);
}

int fmc_trial_delete(fmc_handle *h, int w, int iguest, int choice, fmc_trial *out) {
  return seam_trial(h, w, FMC_MOVE_DELETE, iguest, choice, nullptr, out);
}

int fmc_trial_displace(fmc_handle *h, int w, int iguest, int choice, const double *xyz_new, fmc_trial *out) {
  if (!xyz_new) return FMC_ERR_ARG;
  return seam_trial(h, w, FMC_MOVE_DISPLACE, iguest, choice, xyz_new, out);
}

int fmc_accept(fmc_handle *h, int w) {
  if (!h || w < 0 || w >= h->nwalkers) return FMC_ERR_ARG;
  if (h->pend_state[w] != 1) return FMC_ERR_STATE;
  accept_dispatch(h, w);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->stream));
  h->pend_state[w] = 0;
  return FMC_OK;
}

int fmc_reject(fmc_handle *h, int w) {
  if (!h || w < 0 || w >= h->nwalkers) return FMC_ERR_ARG;
  if (h->pend_state[w] == 0) return FMC_ERR_STATE;
  h->pend_state[w] = 0;  // nothing on the device changed during the trial
  return FMC_OK;
}

int fmc_get_guests(fmc_handle *h, int w, int iguest, double *xyz, int cap_mols, int *nmols) {
  if (!h || w < 0 || w >= h->nwalkers || !xyz || !nmols) return FMC_ERR_ARG;
  const DevParams &P = h->hp;
  if (iguest < 1 || iguest > P.ntpguest) return FMC_ERR_ARG;
  WState s;
  HIPCHK(hipMemcpy(&s, h->dw.ws + w, sizeof s, hipMemcpyDeviceToHost));
  int mol = P.locguest[iguest], na = P.natoms[mol], nm = s.nmols[mol];
  *nmols = nm;
  int n = std::min(nm, cap_mols) * na;
  std::vector<double> x(n), y(n), z(n);
  size_t base = (size_t)w * P.gatoms_total + P.gbase[iguest];
  if (n) {
    HIPCHK(hipMemcpy(x.data(), h->dw.gx + base, 8 * (size_t)n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(y.data(), h->dw.gy + base, 8 * (size_t)n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(z.data(), h->dw.gz + base, 8 * (size_t)n, hipMemcpyDeviceToHost));
  }
  for (int a = 0; a < n; ++a) {
    xyz[3 * a] = x[a];
    xyz[3 * a + 1] = y[a];
    xyz[3 * a + 2] = z[a];
  }
  return FMC_OK;
}

int fmc_get_sk(fmc_handle *h, int w, double *sk) {
  if (!h || w < 0 || w >= h->nwalkers || !sk) return FMC_ERR_ARG;
  const DevParams &P = h->hp;
  std::vector<double2> buf(P.nk);
  HIPCHK(hipMemcpy(buf.data(), h->dw.sk + ((size_t)w * (P.maxmls + 1) + P.maxmls) * P.nk, sizeof(double2) * P.nk,
                   hipMemcpyDeviceToHost));
  for (int i = 0; i < P.nk; ++i) {
    sk[2 * h->kperm[i]] = buf[i].x;
    sk[2 * h->kperm[i] + 1] = buf[i].y;
  }
  return FMC_OK;
}

int fmc_get_stats(fmc_handle *h, int w, fmc_walker_stats *out) {
  if (!h || w < 0 || w >= h->nwalkers || !out) return FMC_ERR_ARG;
  WState s;
  HIPCHK(hipMemcpy(&s, h->dw.ws + w, sizeof s, hipMemcpyDeviceToHost));
  memset(out, 0, sizeof *out);
  out->steps = s.steps;
  for (int i = 0; i < 9; ++i) {
    out->accepted[i] = s.acc[i];
    out->attempted[i] = s.att[i];
  }
  const DevParams &P = h->hp;
  for (int g = 1; g <= P.ntpguest && g <= FMC_MAX_SPECIES; ++g) {
    out->nmols[g - 1] = s.nmols[P.locguest[g]];
    out->sum_n[g - 1] = s.sum_n[g];
    out->sum_n2[g - 1] = s.sum_n2[g];
  }
  for (int m = 1; m <= P.maxmls; ++m) out->energy[m - 1] = s.energy[m];
  out->sum_e = s.sum_e;
  out->sum_e2 = s.sum_e2;
  out->counted = s.counted;
  out->overflow = s.overflow;
  for (int g = 1; g <= P.ntpguest && g <= FMC_MAX_SPECIES; ++g) out->surface[g - 1] = s.surfmols[P.locguest[g]];
  out->prodcount = s.prodcount;
  out->windows = s.nwin;
  out->resums = s.nresum;
  out->resum_drift = s.resum_drift;
  return FMC_OK;
}

int fmc_reset_stats(fmc_handle *h) {
  if (!h) return FMC_ERR_ARG;
  std::vector<WState> ws(h->nwalkers);
  HIPCHK(hipMemcpy(ws.data(), h->dw.ws, sizeof(WState) * ws.size(), hipMemcpyDeviceToHost));
  for (auto &s : ws) {
    s.counted = 0;
    s.sum_e = s.sum_e2 = 0.0;
    for (int i = 0; i <= KMAXS; ++i) s.sum_n[i] = s.sum_n2[i] = 0.0;
    for (int i = 0; i < 9; ++i) s.acc[i] = s.att[i] = 0;
    s.prodcount = s.rollcount = s.nwin = 0;
    memset(s.chain, 0, sizeof s.chain);
  }
  HIPCHK(hipMemcpy(h->dw.ws, ws.data(), sizeof(WState) * ws.size(), hipMemcpyHostToDevice));
  return FMC_OK;
}

int fmc_run(fmc_handle *h, int nsteps, fmc_trace_rec *trace, int count_stats) {
  if (!h || nsteps < 0) return FMC_ERR_ARG;
  if (nsteps == 0) return FMC_OK;
  if (team_fits(h) != FMC_OK) return FMC_ERR_ARG;
  if (team_records(h) != FMC_OK) return FMC_ERR_HIP;
  DevWalkers dw = h->dw;
  if (trace) {
    if (h->dtrace_cap < nsteps) {
      if (h->dtrace) (void)hipFree(h->dtrace);
      h->dtrace = nullptr;
      HIPCHK(hipMalloc(&h->dtrace, sizeof(fmc_trace_rec) * (size_t)nsteps));
      h->dtrace_cap = nsteps;
    }
    dw.trace = h->dtrace;
    dw.trace_w = h->trace_w;
  } else {
    dw.trace = nullptr;
  }
  DevWalkers saved = h->dw;
  h->dw = dw;
  HIPCHK(hipEventRecord(h->ev0, h->stream));
  run_dispatch(h, nsteps, count_stats);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(h->ev1, h->stream));
  h->dw = saved;
  HIPCHK(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms = ms;
  h->last_launches = 1;
  if (trace)
    HIPCHK(hipMemcpy(trace, h->dtrace, sizeof(fmc_trace_rec) * (size_t)nsteps, hipMemcpyDeviceToHost));
  return team_error(h);
}

int fmc_run_graph(fmc_handle *h, int nsweeps, int nsteps_per_sweep, int count_stats) {
  if (!h || nsweeps < 1 || nsteps_per_sweep < 1) return FMC_ERR_ARG;
  if (team_records(h) != FMC_OK) return FMC_ERR_HIP;
  GraphKey key{nsweeps, nsteps_per_sweep, count_stats};
  auto it = h->graphs.find(key);
  if (it == h->graphs.end()) {
    if (team_fits(h) != FMC_OK) return FMC_ERR_ARG;
    hipGraph_t graph;
    DevWalkers saved = h->dw;
    h->dw.trace = nullptr;
    HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < nsweeps; ++i) run_dispatch(h, nsteps_per_sweep, count_stats);
    HIPCHK(hipStreamEndCapture(h->stream, &graph));
    h->dw = saved;
    hipGraphExec_t exec;
    HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(graph));
    it = h->graphs.emplace(key, exec).first;
  }
  HIPCHK(hipEventRecord(h->ev0, h->stream));
  HIPCHK(hipGraphLaunch(it->second, h->stream));
  HIPCHK(hipEventRecord(h->ev1, h->stream));
  HIPCHK(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms = ms;
  h->last_launches = nsweeps;
  return team_error(h);
}

int fmc_set_trace_walker(fmc_handle *h, int w) {
  if (!h || w < 0 || w >= h->nwalkers) return FMC_ERR_ARG;
  h->trace_w = w;
  return FMC_OK;
}

int fmc_last_timing(fmc_handle *h, double *ms, int *launches) {
  if (!h) return FMC_ERR_ARG;
  if (ms) *ms = h->last_ms;
  if (launches) *launches = h->last_launches;
  return FMC_OK;
}

int fmc_set_block(fmc_handle *h, int threads) {
  if (!h) return FMC_ERR_ARG;
  if (threads == 0) {
    // autotune on a snapshot of the walkers: time a short run per block size
    // and restore every array a run writes afterwards (state, guests, S(k),
    // Wang-Landau ln g and visits; records past the snapshot's counts are
    // rewritten by later runs)
    const DevParams &P = h->hp;
    struct Snap { void *dev; size_t bytes; void *copy; };
    std::vector<Snap> snap = {
        {h->dw.ws, sizeof(WState) * (size_t)h->nwalkers, nullptr},
        {h->dw.gx, 8 * (size_t)h->nwalkers * P.gatoms_total, nullptr},
        {h->dw.gy, 8 * (size_t)h->nwalkers * P.gatoms_total, nullptr},
        {h->dw.gz, 8 * (size_t)h->nwalkers * P.gatoms_total, nullptr},
        {h->dw.sk, sizeof(double2) * (size_t)h->nwalkers * (P.maxmls + 1) * P.nk, nullptr}};
    if (P.wl_nbins > 0) {
      snap.push_back({h->dw.wl_lng, sizeof(double) * (size_t)h->nwalkers * P.wl_nbins, nullptr});
      snap.push_back({h->dw.wl_hist, sizeof(int) * (size_t)h->nwalkers * P.wl_nbins, nullptr});
    }
    int rc = FMC_OK;
    for (auto &sn : snap) {
      if (hipMalloc(&sn.copy, sn.bytes) != hipSuccess) {
        sn.copy = nullptr;
        rc = FMC_ERR_ALLOC;
        break;
      }
      if (hipMemcpy(sn.copy, sn.dev, sn.bytes, hipMemcpyDeviceToDevice) != hipSuccess) {
        rc = FMC_ERR_HIP;
        break;
      }
    }
    // the tiles: threads per workgroup x workgroups per walker (each member
    // of a team takes 1 / team of the pair atoms and of the k-vectors), every
    // team the device holds at once, and where the tables are read from: the
    // LDS plan's kernel variant or the one that reads the framework, the k
    // list and the LJ columns from global memory (L2)
    const int prev = h->block, prev_team = h->team, prev_var = h->variant_ovr;
    int best = prev, best_team = prev_team, best_var = prev_var;
    double best_ms = 1e30;
    const int cand[3] = {128, 256, 512};
    const int teams[8] = {1, 2, 3, 4, 6, 8, 12, 16};
    h->variant_ovr = -1;
    const int plan_var = run_variant(h);
    const int vars[2] = {-1, plan_var != 0 ? 0 : -2};
    for (int vi = 0; vi < 2 && rc == FMC_OK; ++vi) {
      if (vars[vi] == -2) continue;
      h->variant_ovr = vars[vi];
      for (int ti = 0; ti < 8 && rc == FMC_OK; ++ti) {
        const int team = teams[ti];
        if (team > 1 && (h->tune_team_max < team || team_arrays(h, team) != FMC_OK)) continue;
        for (int i = 0; i < 3 && rc == FMC_OK; ++i) {
          if (team > 1 && run_grid(h, team) > run_capacity(h, cand[i])) continue;
          h->block = cand[i];
          h->team = team;
          int e = fmc_run(h, 64, nullptr, 0);
          if (e == FMC_OK) e = fmc_run(h, 256, nullptr, 0);
          if (e == FMC_OK && h->last_ms < best_ms) {
            best_ms = h->last_ms;
            best = cand[i];
            best_team = team;
            best_var = vars[vi];
          }
          for (auto &sn : snap)
            if (hipMemcpy(sn.dev, sn.copy, sn.bytes, hipMemcpyDeviceToDevice) != hipSuccess) rc = FMC_ERR_HIP;
          if (e != FMC_OK && rc == FMC_OK) rc = e;
        }
      }
    }
    for (auto &sn : snap)
      if (sn.copy) (void)hipFree(sn.copy);
    h->block = rc == FMC_OK ? best : prev;
    h->team = rc == FMC_OK ? best_team : prev_team;
    h->variant_ovr = rc == FMC_OK ? best_var : prev_var;
    for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
    h->graphs.clear();
    return rc;
  }
  if (threads < 64 || threads > 512 || threads % 64) return FMC_ERR_ARG;
  if (h->team > 1 && run_grid(h, h->team) > run_capacity(h, threads)) return FMC_ERR_ARG;
  h->block = threads;
  for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  h->graphs.clear();
  return FMC_OK;
}

int fmc_get_block(const fmc_handle *h, int *threads) {
  if (!h || !threads) return FMC_ERR_ARG;
  *threads = h->block;
  return FMC_OK;
}

int fmc_set_team(fmc_handle *h, int team) {
  if (!h || team < 0 || team > KTEAM) return FMC_ERR_ARG;
  if (team == 0) {
    // let the autotuner of fmc_set_block(0) consider teams (up to KTEAM)
    h->tune_team_max = KTEAM;
    return FMC_OK;
  }
  if (team > 1 && run_grid(h, team) > run_capacity(h, h->block)) return FMC_ERR_ARG;
  const int rc = team_arrays(h, team);
  if (rc != FMC_OK) return rc;
  h->team = team;
  h->tune_team_max = team;
  for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  h->graphs.clear();
  return FMC_OK;
}

int fmc_get_team(const fmc_handle *h, int *team) {
  if (!h || !team) return FMC_ERR_ARG;
  *team = h->team;
  return FMC_OK;
}

int fmc_wl_window(int win, int nwindows, int maxn, int minn, int *minmol, int *maxmol, int *tail) {
  // divide_jobs, wang_landau.f:1742-1781 (integer arithmetic as written there)
  if (nwindows < 1 || win < 0 || win >= nwindows || maxn < 1 || minn < 0 || minn > maxn) return FMC_ERR_ARG;
  int t = nwindows == 1 ? 0 : 3;
  int nodex = win + 1;
  long long a = (long long)(maxn + 1) * nodex * nodex + (long long)((maxn + 1) / nwindows) * nodex * nodex;
  int mx = (int)(a / ((long long)nodex * nodex + nodex + 1)) + t;
  long long b = (long long)(maxn + 1) * win * win + (long long)((maxn + 1) / nwindows) * win * win;
  int mn = (int)(b / ((long long)win * win + win + 1)) - t;
  if (nodex == 1) mn = minn;
  if (nodex == nwindows) mx = maxn;
  if (mn < 0) mn = 0;
  if (mx > maxn) mx = maxn;
  if (minmol) *minmol = mn;
  if (maxmol) *maxmol = mx;
  if (tail) *tail = t;
  return FMC_OK;
}

int fmc_wl_setup(fmc_handle *h, const fmc_wl_params *p) {
  if (!h || !p) return FMC_ERR_ARG;
  if (p->maxn < 1 || p->minn < 0 || p->minn >= p->maxn || p->nwindows < 1 || p->nwindows > h->nwalkers ||
      p->visittol < 0 || !(p->flatcoeff > 0.0) || !(p->prectol > 0.0) || p->create_molecules < 0)
    return FMC_ERR_ARG;
  if (!(p->wlprec > 1.0)) return 2319;  // readinputs.f:739
  DevParams &P = h->hp;
  if (p->maxn > P.gcap[1]) return FMC_ERR_ARG;
  for (int win = 0; win < p->nwindows; ++win) {
    int mn, mx;
    if (fmc_wl_window(win, p->nwindows, p->maxn, p->minn, &mn, &mx, nullptr) != FMC_OK || mn > mx)
      return FMC_ERR_ARG;
  }
  // every check first, on copies: the handle changes only once all pass
  const int mol1 = P.locguest[1];
  std::vector<WState> ws(h->nwalkers);
  HIPCHK(hipMemcpy(ws.data(), h->dw.ws, sizeof(WState) * ws.size(), hipMemcpyDeviceToHost));
  for (int w = 0; w < h->nwalkers; ++w)
    if (ws[w].nmols[mol1] > p->maxn) return FMC_ERR_ARG;  // more guests than the sampled range
  const int nbins = p->maxn + 2;
  if (h->dw.wl_lng && nbins > h->wl_maxn + 2) return FMC_ERR_ARG;  // re-setup may not grow the range
  // LDS: ln g and visits after the rest of the plan; without room the
  // all-in-LDS variant gives way
  DevParams Pn = P;
  const size_t budget = 160 * 1024 - 1024;
  const size_t need = (size_t)12 * nbins + 32;
  size_t base = h->plan_end;
  if (base + need > budget && Pn.lds_all) {
    // the all-in-LDS plan gives way; the LJ columns alone stay if they fit
    Pn.lds_all = 0;
    Pn.lj_lds = 0;
    base = h->lds_base_end;
    if (h->ljo_fits && h->ljo_end + need <= budget) {
      Pn.lj_lds = 1;
      Pn.off_vvv64 = h->ljo_off64;
      Pn.off_vvv = h->ljo_off32;
      base = h->ljo_end;
    }
  }
  if (base + need > budget && Pn.lj_lds) {
    Pn.lj_lds = 0;
    base = h->lds_base_end;
  }
  if (base + need > budget) return FMC_ERR_ARG;
  Pn.wl_off = (int)((base + 15) & ~(size_t)15);
  plan_resum_rows(Pn, Pn.wl_off + (size_t)12 * nbins);
  Pn.wl_nbins = nbins;
  Pn.wl_flat = p->flatcoeff;
  Pn.wl_visittol = p->visittol;
  Pn.wl_prectol = p->prectol;
  // guest-1 thermal wavelength, wang_landau.f:121-123 (molmass = sum of site masses, readinputs.f:208,247)
  double molmass = 0.0;
  for (int i = 0; i < P.natoms[mol1]; ++i) molmass += P.site_m[mol1][i];
  for (int w = 0; w < h->nwalkers; ++w) {
    WState &s = ws[w];
    int win = (int)((long long)w * p->nwindows / h->nwalkers), mn, mx;
    fmc_wl_window(win, p->nwindows, p->maxn, p->minn, &mn, &mx, nullptr);
    s.wl_on = 1;
    s.wl_done = 0;
    s.wl_error = 0;
    s.wl_min = mn;
    s.wl_max = mx;
    s.wl_preload = std::max(p->create_molecules, mn);
    s.wl_hmin = 0;
    s.wl_hmin_cnt = mx - mn + 1;
    s.wl_hsum = 0;
    s.wl_sweeps = 0;
    s.wl_iter = 0;
    s.wl_f = p->wlprec;
    s.wl_logf = std::log(p->wlprec);
    double lambda = HPLANCK / std::sqrt(2.0 * PI * molmass / 1000.0 * BOLTZ * s.temp / AVO) / 1.0e-10;
    s.wl_lambda3 = lambda * lambda * lambda;
    // step sizes of the Wang-Landau loop, wang_landau.f:105-110
    s.rotangle = PI / 3.0;
    s.rota_rotangle = s.rotangle;
    for (int g = 1; g <= P.ntpguest; ++g) {
      s.delrdisp[g] = s.delr[g];
      s.tran_delr[g] = s.delr[g];
    }
  }
  // commit: the graphs captured with the old plan go first
  for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  h->graphs.clear();
  if (!h->dw.wl_lng) {
    int rc = dalloc(h, &h->dw.wl_lng, (size_t)h->nwalkers * nbins);
    rc |= dalloc(h, &h->dw.wl_hist, (size_t)h->nwalkers * nbins);
    if (rc) return FMC_ERR_ALLOC;
  }
  HIPCHK(hipMemset(h->dw.wl_lng, 0, sizeof(double) * (size_t)h->nwalkers * nbins));
  HIPCHK(hipMemset(h->dw.wl_hist, 0, sizeof(int) * (size_t)h->nwalkers * nbins));
  HIPCHK(allow_lds_dispatch(h->ns, h->precision == FMC_MIXED, Pn.lds_bytes));
  P = Pn;
  h->plan_end = base;
  HIPCHK(hipMemcpy(h->dp, &P, sizeof P, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->dw.ws, ws.data(), sizeof(WState) * ws.size(), hipMemcpyHostToDevice));
  h->wl_maxn = p->maxn;
  h->wl_minn = p->minn;
  h->wl_nwin = p->nwindows;
  return FMC_OK;
}

int fmc_wl_get(fmc_handle *h, int w, double *lng, int *visits, fmc_wl_status *st) {
  if (!h || w < 0 || w >= h->nwalkers || !h->hp.wl_nbins) return FMC_ERR_ARG;
  const int nb = h->hp.wl_nbins;
  if (lng)
    HIPCHK(hipMemcpy(lng, h->dw.wl_lng + (size_t)w * nb, sizeof(double) * (h->wl_maxn + 1), hipMemcpyDeviceToHost));
  if (visits)
    HIPCHK(hipMemcpy(visits, h->dw.wl_hist + (size_t)w * nb, sizeof(int) * (h->wl_maxn + 1), hipMemcpyDeviceToHost));
  if (st) {
    WState s;
    HIPCHK(hipMemcpy(&s, h->dw.ws + w, sizeof s, hipMemcpyDeviceToHost));
    st->minmol = s.wl_min;
    st->maxmol = s.wl_max;
    st->done = s.wl_done;
    st->error = s.wl_error;
    st->sweeps = s.wl_sweeps;
    st->steps = s.steps;
    st->f = s.wl_f;
    st->lnf = s.wl_logf;
    st->lambda3 = s.wl_lambda3;
  }
  return FMC_OK;
}

int fmc_wl_stitch(const double *lng_windows, int nwindows, int maxn, int minn, double *lng_out) {
  // stitch_branches, wang_landau.f:1548-1621, in 0-based N
  if (!lng_windows || !lng_out || nwindows < 1) return FMC_ERR_ARG;
  const int nb = maxn + 1;
  int mn0, mx0, tail;
  if (fmc_wl_window(0, nwindows, maxn, minn, &mn0, &mx0, &tail) != FMC_OK) return FMC_ERR_ARG;
  std::vector<double> dos(lng_windows, lng_windows + nb);
  const double shift = dos[0];
  for (int i = 0; i < mx0 - mn0; ++i) dos[i] -= shift;  // the reference shifts varchunk bins
  for (int win = 1; win < nwindows; ++win) {
    const double *buf = lng_windows + (size_t)win * nb;
    int mn, mx;
    fmc_wl_window(win, nwindows, maxn, minn, &mn, &mx, &tail);
    const double cst = dos[mn + 1] - buf[mn + 1];  // obtain_min_shift: obj(1) - opt(1)
    int ik = 1;
    for (int j = mn; j <= mx; ++j) {
      ++ik;
      if (ik >= 2 && ik <= 2 * tail - 1) dos[j] = (dos[j] + buf[j] + cst) / 2.0;
      else dos[j] = buf[j] + cst;
    }
  }
  for (int i = 0; i < nb; ++i) lng_out[i] = dos[i];
  return FMC_OK;
}

int fmc_wl_result(fmc_handle *h, double *lng_out) {
  if (!h || !lng_out || !h->hp.wl_nbins) return FMC_ERR_ARG;
  const int nb = h->wl_maxn + 1, nwin = h->wl_nwin;
  std::vector<double> all((size_t)h->nwalkers * h->hp.wl_nbins);
  HIPCHK(hipMemcpy(all.data(), h->dw.wl_lng, sizeof(double) * all.size(), hipMemcpyDeviceToHost));
  std::vector<double> wins((size_t)nwin * nb, 0.0);
  std::vector<int> cnt(nwin, 0);
  for (int w = 0; w < h->nwalkers; ++w) {
    int win = (int)((long long)w * nwin / h->nwalkers), mn, mx;
    fmc_wl_window(win, nwin, h->wl_maxn, h->wl_minn, &mn, &mx, nullptr);
    const double *g = all.data() + (size_t)w * h->hp.wl_nbins;
    const double ref = g[mn];
    for (int n = mn; n <= mx; ++n) wins[(size_t)win * nb + n] += g[n] - ref;
    cnt[win] += 1;
  }
  for (int win = 0; win < nwin; ++win)
    for (int n = 0; n < nb; ++n)
      if (cnt[win]) wins[(size_t)win * nb + n] /= cnt[win];
  return fmc_wl_stitch(wins.data(), nwin, h->wl_maxn, h->wl_minn, lng_out);
}

int fmc_wl_isotherm(const double *lng, int maxn, int minn, double temperature, double lambda3,
                    const double *p_pa, int np, int units, double *n_out) {
  // compute_isotherm / grand_canonical_partition, wang_landau.f:1623-1740:
  // Z = sum_{N=0}^{maxn-minn} exp(ln g(N) + beta mu (N - minn)),
  // <N> = sum_{N=0}^{maxn-1} N exp(ln g(N) + beta mu N) / Z, as log-sum-exp
  if (!lng || !p_pa || !n_out || np < 0 || maxn < 1 || minn < 0 || minn > maxn || !(temperature > 0.0) ||
      !(lambda3 > 0.0) || (units != 0 && units != 1))
    return FMC_ERR_ARG;
  const double beta = 1.0 / (KBOLTZ * temperature);
  std::vector<double> zt(maxn - minn + 1), nt(maxn);
  for (int ip = 0; ip < np; ++ip) {
    double bmu;
    if (units == 0) {
      double mu = std::log(p_pa[ip] * lambda3 * beta) / beta;
      bmu = beta * mu;
    } else {
      bmu = std::log(p_pa[ip] * lambda3 * 1.0e-30 / (BOLTZ * temperature));
    }
    double zmax = -INFINITY, nmax = -INFINITY;
    for (int i = 0; i <= maxn - minn; ++i) {
      zt[i] = lng[i] + bmu * (double)(i - minn);
      zmax = std::max(zmax, zt[i]);
    }
    for (int k = 1; k < maxn; ++k) {
      nt[k] = std::log((double)k) + lng[k] + bmu * (double)k;
      nmax = std::max(nmax, nt[k]);
    }
    double zs = 0.0, ns = 0.0;
    for (int i = 0; i <= maxn - minn; ++i) zs += std::exp(zt[i] - zmax);
    for (int k = 1; k < maxn; ++k) ns += std::exp(nt[k] - nmax);
    n_out[ip] = maxn > 1 ? std::exp(std::log(ns) + nmax - std::log(zs) - zmax) : 0.0;
  }
  return FMC_OK;
}

namespace {

// Device buffers of one fmc_widom call: 4 draws per attempt and the staged
// per-attempt outputs for `chunk` attempts, per-block partial sums, maps.
struct WidomBufs {
  float *uni = nullptr;
  double *est = nullptr, *bsum = nullptr, *emap = nullptr, *cmap = nullptr;
  int *ovl = nullptr;
  long long *bcnt = nullptr;
  long long chunk = 0;
  int nblocks = 0;
};

// Grid of widom_grid: spacing griddim, gridfactor unit cells per axis
// (gcmc.f:258-281, 605-610).
void widom_grid_size(const DevParams &P, const int *gf, double griddim, int *ng, int *wng) {
  double celprp[10];
  dcell(P.cell, celprp);
  for (int a = 0; a < 3; ++a) {
    const double gd = celprp[a] / std::ceil(celprp[a] / gf[a] / griddim) / gf[a];
    ng[a] = (int)std::ceil(celprp[a] / (gd * gf[a]));
    wng[a] = (int)std::ceil(celprp[a] / gd);
  }
}

// widom_grid (gcmc.f:2454-2548) on walker w with the given buffers (maps
// already cleared): attempts in the reference's loop order, `chunk` at a
// time, orientations from the walker's duni sequence (which advances by 4
// draws per attempt).
int widom_run(fmc_handle *h, int w, int iguest, int nstep, const int *gf, const int *ng, const int *wng,
              const WidomBufs &B, double *estep_out, int *overlap_out, double *emap_out, double *cmap_out,
              fmc_widom_result *res) {
  const DevParams &P = h->hp;
  const long long total = (long long)ng[0] * ng[1] * ng[2] * nstep;
  const size_t mapsize = (size_t)ng[0] * ng[1] * ng[2];
  WState ws;
  HIPCHK(hipMemcpy(&ws, h->dw.ws + w, sizeof(WState), hipMemcpyDeviceToHost));
  // the unit cell the map folds into (gcmc.f:299-306)
  double ucell[9], rucell[9];
  for (int i = 0; i < 9; ++i) ucell[i] = P.cell[i] / (double)gf[i / 3];
  invert(ucell, rucell);
  const WidomLds L = widom_lds(P);
  const bool sites3 = h->nsite == 3;
  const void *kf = h->precision == FMC_FP64
                       ? (sites3 ? (const void *)k_widom<double, 3> : (const void *)k_widom<double, 0>)
                       : (sites3 ? (const void *)k_widom<float, 3> : (const void *)k_widom<float, 0>);
  if (L.total > 160 * 1024) return FMC_ERR_ARG;
  HIPCHK(hipFuncSetAttribute(kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.total));
  std::vector<double> bsum(B.nblocks);
  std::vector<long long> bcnt(B.nblocks);
  double hsum = 0.0;
  long long accepted = 0;
  for (long long a0 = 0; a0 < total; a0 += B.chunk) {
    const long long cnt = std::min(B.chunk, total - a0);
    hipLaunchKernelGGL(k_draws, dim3(1), dim3(64), 0, h->stream, h->dw, w, 4 * cnt, B.uni);
    WidomArgs A{};
    A.w = w;
    A.g = iguest;
    A.nstep = nstep;
    for (int a = 0; a < 3; ++a) {
      A.ng[a] = ng[a];
      A.wng[a] = wng[a];
    }
    A.a0 = a0;
    A.count = cnt;
    A.uni = B.uni;
    A.estep = estep_out ? B.est : nullptr;
    A.overlap = overlap_out ? B.ovl : nullptr;
    A.bsum = B.bsum;
    A.bcnt = B.bcnt;
    A.kboltz = KBOLTZ;
    A.emap = emap_out ? B.emap : nullptr;
    A.cmap = emap_out ? B.cmap : nullptr;
    for (int i = 0; i < 9; ++i) A.rucell[i] = rucell[i];
    const int nb = (int)std::min<long long>((cnt + WIDOM_WPB - 1) / WIDOM_WPB, B.nblocks);
    if (h->precision == FMC_FP64) {
      if (sites3) hipLaunchKernelGGL((k_widom<double, 3>), dim3(nb), dim3(64 * WIDOM_WPB), L.total, h->stream, h->dp, h->dw, A);
      else hipLaunchKernelGGL((k_widom<double, 0>), dim3(nb), dim3(64 * WIDOM_WPB), L.total, h->stream, h->dp, h->dw, A);
    } else {
      if (sites3) hipLaunchKernelGGL((k_widom<float, 3>), dim3(nb), dim3(64 * WIDOM_WPB), L.total, h->stream, h->dp, h->dw, A);
      else hipLaunchKernelGGL((k_widom<float, 0>), dim3(nb), dim3(64 * WIDOM_WPB), L.total, h->stream, h->dp, h->dw, A);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(bsum.data(), B.bsum, sizeof(double) * nb, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(bcnt.data(), B.bcnt, sizeof(long long) * nb, hipMemcpyDeviceToHost));
    if (estep_out) HIPCHK(hipMemcpy(estep_out + a0, B.est, sizeof(double) * cnt, hipMemcpyDeviceToHost));
    if (overlap_out) HIPCHK(hipMemcpy(overlap_out + a0, B.ovl, sizeof(int) * cnt, hipMemcpyDeviceToHost));
    for (int b = 0; b < nb; ++b) {
      hsum += bsum[b];
      accepted += bcnt[b];
    }
  }
  if (emap_out) {
    HIPCHK(hipMemcpy(emap_out, B.emap, sizeof(double) * mapsize, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cmap_out, B.cmap, sizeof(double) * mapsize, hipMemcpyDeviceToHost));
  }
  res->attempts = total;
  res->accepted = accepted;
  res->avg_boltz = hsum / (double)total;
  res->henry = res->avg_boltz / AVO / BOLTZ / ws.temp * 1.0e5;
  for (int a = 0; a < 3; ++a) {
    res->ngrid[a] = ng[a];
    res->wngrid[a] = wng[a];
  }
  return FMC_OK;
}

int widom_args(fmc_handle *h, int w, int iguest, int nstep, double griddim, const int *gridfactor,
               const double *emap_out, const double *cmap_out, fmc_widom_result *res, int *gf) {
  if (!h || !res || w < 0 || w >= h->nwalkers || iguest < 1 || iguest > h->hp.ntpguest || nstep < 1 ||
      !(griddim > 0.0))
    return FMC_ERR_ARG;
  for (int a = 0; a < 3; ++a) gf[a] = gridfactor ? gridfactor[a] : 1;
  if (gf[0] < 1 || gf[1] < 1 || gf[2] < 1 || (!emap_out) != (!cmap_out)) return FMC_ERR_ARG;
  return FMC_OK;
}

}  // namespace

// widom_grid (gcmc.f:2454-2548) on walker w: grid set-up of gcmc.f:258-281 and
// 605-610, attempts in the reference's loop order, orientations from the
// walker's duni sequence (which advances by 4 draws per attempt).
int fmc_widom(fmc_handle *h, int w, int iguest, int nstep, double griddim, const int *gridfactor,
              double *estep_out, int *overlap_out, double *emap_out, double *cmap_out, fmc_widom_result *res) {
  int gf[3];
  if (int e = widom_args(h, w, iguest, nstep, griddim, gridfactor, emap_out, cmap_out, res, gf)) return e;
  int ng[3], wng[3];
  widom_grid_size(h->hp, gf, griddim, ng, wng);
  const long long total = (long long)ng[0] * ng[1] * ng[2] * nstep;
  const size_t mapsize = (size_t)ng[0] * ng[1] * ng[2];
  WidomBufs B;
  B.chunk = std::min<long long>(total, 1LL << 22);
  B.nblocks = (int)std::min<long long>((B.chunk + WIDOM_WPB - 1) / WIDOM_WPB, 4096);
  int rc = FMC_OK;
  if (hipMalloc(&B.uni, sizeof(float) * 4 * (size_t)B.chunk) != hipSuccess ||
      hipMalloc(&B.bsum, sizeof(double) * B.nblocks) != hipSuccess ||
      hipMalloc(&B.bcnt, sizeof(long long) * B.nblocks) != hipSuccess ||
      (estep_out && hipMalloc(&B.est, sizeof(double) * (size_t)B.chunk) != hipSuccess) ||
      (overlap_out && hipMalloc(&B.ovl, sizeof(int) * (size_t)B.chunk) != hipSuccess) ||
      (emap_out && (hipMalloc(&B.emap, sizeof(double) * mapsize) != hipSuccess ||
                    hipMalloc(&B.cmap, sizeof(double) * mapsize) != hipSuccess ||
                    hipMemset(B.emap, 0, sizeof(double) * mapsize) != hipSuccess ||
                    hipMemset(B.cmap, 0, sizeof(double) * mapsize) != hipSuccess)))
    rc = FMC_ERR_ALLOC;
  if (rc == FMC_OK)
    rc = widom_run(h, w, iguest, nstep, gf, ng, wng, B, estep_out, overlap_out, emap_out, cmap_out, res);
  for (void *p : {(void *)B.uni, (void *)B.est, (void *)B.ovl, (void *)B.bsum, (void *)B.bcnt, (void *)B.emap,
                  (void *)B.cmap})
    if (p) (void)hipFree(p);
  return rc;
}

int fmc_widom_noalloc(fmc_handle *h, int w, int iguest, int nstep, double griddim, const int *gridfactor,
                      double *estep_out, int *overlap_out, double *emap_out, double *cmap_out,
                      fmc_widom_result *res) {
  int gf[3];
  if (int e = widom_args(h, w, iguest, nstep, griddim, gridfactor, emap_out, cmap_out, res, gf)) return e;
  if (!h->scratch) return FMC_ERR_STATE;  // fmc_create could not resolve the scratch array
  int ng[3], wng[3];
  widom_grid_size(h->hp, gf, griddim, ng, wng);
  const size_t mapsize = (size_t)ng[0] * ng[1] * ng[2];
  if (emap_out && mapsize > (size_t)FMC_WIDOM_SCRATCH_MAP) {
    fprintf(stderr, "fastmc: fmc_widom_noalloc: an energy map of %zu grid points exceeds the fixed scratch (%d)\n",
            mapsize, FMC_WIDOM_SCRATCH_MAP);
    return FMC_ERR_SCRATCH;
  }
  // carve the fixed array: draws, per-attempt staging, block sums, maps
  unsigned char *base = (unsigned char *)h->scratch;
  const long long CH = FMC_WIDOM_SCRATCH_CHUNK;
  WidomBufs B;
  B.chunk = CH;
  B.nblocks = 4096;
  size_t off = 0;
  B.uni = (float *)(base + off);
  off += sizeof(float) * 4 * CH;
  B.est = (double *)(base + off);
  off += sizeof(double) * CH;
  B.ovl = (int *)(base + off);
  off += sizeof(int) * CH;
  B.bsum = (double *)(base + off);
  off += sizeof(double) * B.nblocks;
  B.bcnt = (long long *)(base + off);
  off += sizeof(long long) * B.nblocks;
  B.emap = (double *)(base + off);
  off += sizeof(double) * FMC_WIDOM_SCRATCH_MAP;
  B.cmap = (double *)(base + off);
  if (emap_out) {
    // the array keeps the previous call's maps: clear them with a fill kernel
    const int n = (int)mapsize;
    hipLaunchKernelGGL(k_fill_zero, dim3((n + 255) / 256), dim3(256), 0, h->stream, B.emap, n);
    hipLaunchKernelGGL(k_fill_zero, dim3((n + 255) / 256), dim3(256), 0, h->stream, B.cmap, n);
    HIPCHK(hipGetLastError());
  }
  return widom_run(h, w, iguest, nstep, gf, ng, wng, B, estep_out, overlap_out, emap_out, cmap_out, res);
}

int fmc_energy_grad(fmc_handle *h, const fmc_theta *theta, int nconf, const int *nmols, const double *xyz,
                    int gstride, int precision, double *U, double *dU_dq, double *dU_deps, double *dU_dsig) {
  if (!h || !theta || !theta->fw_charge || !theta->type_eps || !theta->type_sig || nconf < 1 || !nmols || !xyz ||
      gstride < 1 || (precision != FMC_FP64 && precision != FMC_MIXED))
    return FMC_ERR_ARG;
  const DevParams &P = h->hp;
  const int nfw = P.nfw, nk = P.nk, ng = P.ntpguest, nt = P.ntypes, TT = KTYPE + 1;
  for (int a = 0; a < nfw; ++a)
    if (!h->fw_frz[a]) return FMC_ERR_ARG;  // framework-framework pairs are left out: a rigid, frozen framework only
  int gmax = 0;
  for (int c = 0; c < nconf; ++c) {
    int n = 0;
    for (int g = 1; g <= ng; ++g) {
      if (nmols[c * ng + g - 1] < 0) return FMC_ERR_ARG;
      n += nmols[c * ng + g - 1] * P.natoms[P.locguest[g]];
    }
    if (n > gstride) return FMC_ERR_ARG;
    gmax = std::max(gmax, n);
  }
  for (int t = 1; t <= nt; ++t)
    if (!(theta->type_eps[t - 1] >= 0.0) || !(theta->type_sig[t - 1] > 0.0)) return FMC_ERR_ARG;
  std::vector<int> isfw(TT, 0);
  for (int a = 0; a < nfw; ++a) isfw[h->fw_type[a]] = 1;
  // ---- LJ tables of the theta pairs (forgen, vdw_terms.f:222-234, and their
  // derivatives with respect to eps and sigma at the same grid points)
  const int vs = P.vstride, mx = P.mxegrd;
  std::vector<int> ptab(TT * TT, -1);
  std::vector<double> vt, peps_t, psig_t;
  for (int i = 1; i <= nt; ++i)
    for (int j = 1; j <= i; ++j) {
      if (!isfw[i] && !isfw[j]) continue;
      double e = std::sqrt(theta->type_eps[i - 1] * theta->type_eps[j - 1]);
      double sg = 0.5 * (theta->type_sig[i - 1] + theta->type_sig[j - 1]);
      size_t o = vt.size();
      vt.resize(o + 3 * (size_t)vs, 0.0);
      for (int r = 1; r <= mx; ++r) {
        double rrr = (double)r * P.dlrpot, x = sg / rrr, x2 = x * x, x6 = x2 * x2 * x2;
        vt[o + r] = 4.0 * e * x6 * (x6 - 1.0);
        vt[o + vs + r] = 4.0 * x6 * (x6 - 1.0);
        vt[o + 2 * vs + r] = 4.0 * e * (12.0 * x6 * x6 - 6.0 * x6) / sg;
      }
      int id = (int)peps_t.size();
      peps_t.push_back(e);
      psig_t.push_back(sg);
      // srfrce skip rule, vdw_terms.f:431; -2 marks a theta pair without LJ
      int v = std::fabs(vt[o + 1]) > 1e-10 ? id : -2;
      ptab[i * TT + j] = v;
      ptab[j * TT + i] = v;
    }
  const int npt = (int)peps_t.size();
  // ---- device workspace
  const size_t NTT = (size_t)TT * TT;
  // outputs: type pairs at stride ntypes + 1; the framework part by KSFW k ranges
  const int ntt = nt + 1, KSFW = 16;
  const size_t NTC = (size_t)ntt * ntt;
  const size_t outn = (size_t)nconf * (1 + nfw + 2 + 2 * NTC) + (size_t)KSFW * nfw;
  if (!h->g_efw32) {
    HIPCHK(hipMalloc(&h->g_efw32, sizeof(float2) * (size_t)nk * nfw));
    HIPCHK(hipMalloc(&h->g_efw64, sizeof(double2) * (size_t)nk * nfw));
    HIPCHK(hipMalloc(&h->g_sfw, sizeof(double2) * nk));
    HIPCHK(hipMalloc(&h->g_fwq, sizeof(double) * std::max(nfw, 1)));
    HIPCHK(hipMalloc(&h->g_fwt, sizeof(int) * std::max(nfw, 1)));
    HIPCHK(hipMalloc(&h->g_ptab, sizeof(int) * NTT));
    HIPCHK(hipMemcpy(h->g_fwt, h->fw_type.data(), sizeof(int) * nfw, hipMemcpyHostToDevice));
  }
  if (nconf > h->g_cap || gstride > h->g_gcap) {
    if (h->g_sc) (void)hipFree(h->g_sc);
    if (h->g_nmols) (void)hipFree(h->g_nmols);
    if (h->g_xyz) (void)hipFree(h->g_xyz);
    if (h->g_out) (void)hipFree(h->g_out);
    h->g_cap = std::max(nconf, h->g_cap);
    h->g_gcap = std::max(gstride, h->g_gcap);
    HIPCHK(hipMalloc(&h->g_sc, sizeof(double2) * (size_t)h->g_cap * nk));
    HIPCHK(hipMalloc(&h->g_nmols, sizeof(int) * (size_t)h->g_cap * ng));
    HIPCHK(hipMalloc(&h->g_xyz, sizeof(double) * 3 * (size_t)h->g_cap * h->g_gcap));
    HIPCHK(hipMalloc(&h->g_out, sizeof(double) * ((size_t)h->g_cap * (1 + nfw + 2 + 2 * NTC) + (size_t)KSFW * nfw)));
  }
  if (npt > h->g_nptcap) {
    if (h->g_vt) (void)hipFree(h->g_vt);
    h->g_nptcap = npt;
    HIPCHK(hipMalloc(&h->g_vt, sizeof(double) * (3 * (size_t)vs + 1) * std::max(npt, 1)));
  }
  GradArgs A{};
  A.nconf = nconf;
  A.nk = nk;
  A.nfw = nfw;
  A.gstride = gstride;
  A.ntpguest = ng;
  A.nmols = h->g_nmols;
  A.gx = h->g_xyz;
  A.fwq = h->g_fwq;
  A.fx = P.fw_per_walker ? h->dw.fwx : P.fwx;
  A.fy = P.fw_per_walker ? h->dw.fwy : P.fwy;
  A.fz = P.fw_per_walker ? h->dw.fwz : P.fwz;
  A.fwt = h->g_fwt;
  A.efw32 = h->g_efw32;
  A.efw64 = h->g_efw64;
  A.sfw = h->g_sfw;
  A.sc = h->g_sc;
  A.urec = h->g_out;
  A.dq = A.urec + nconf;
  A.e2 = A.dq + (size_t)nconf * nfw;
  A.ev = A.e2 + nconf;
  A.dpe = A.ev + nconf;
  A.dps = A.dpe + (size_t)nconf * NTC;
  A.dqfw = A.dps + (size_t)nconf * NTC;
  A.ntt = ntt;
  A.vt = h->g_vt;
  A.rpe = h->g_vt + 3 * (size_t)vs * npt;
  A.ptab = h->g_ptab;
  A.vstride = vs;
  A.npt = npt;
  HIPCHK(hipMemcpyAsync(h->g_nmols, nmols, sizeof(int) * (size_t)nconf * ng, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->g_xyz, xyz, sizeof(double) * 3 * (size_t)nconf * gstride, hipMemcpyHostToDevice,
                        h->stream));
  HIPCHK(hipMemcpyAsync(h->g_fwq, theta->fw_charge, sizeof(double) * nfw, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->g_ptab, ptab.data(), sizeof(int) * NTT, hipMemcpyHostToDevice, h->stream));
  if (npt) {
    for (int i = 0; i < npt; ++i) vt.push_back(1.0 / peps_t[i]);  // theta pairs with LJ have eps > 0
    HIPCHK(hipMemcpyAsync(h->g_vt, vt.data(), sizeof(double) * vt.size(), hipMemcpyHostToDevice, h->stream));
  }
  HIPCHK(hipMemsetAsync(h->g_out, 0, sizeof(double) * outn, h->stream));
  if (!h->g_efw_ready) {
    hipLaunchKernelGGL(k_grad_phases, dim3((nfw + 255) / 256, nk), dim3(256), 0, h->stream, h->dp, A);
    h->g_efw_ready = 1;
  }
  hipLaunchKernelGGL(k_grad_sfw, dim3(nk), dim3(256), 0, h->stream, A);
  {
    // guest sites in batches whose axis phase rows fill at most 48 KB of LDS
    const size_t per = (size_t)3 * P.kcp * sizeof(double2) + sizeof(double);
    const int nb = (int)std::max<size_t>(1, std::min<size_t>(256, (48 * 1024) / per));
    hipLaunchKernelGGL(k_grad_sconf, dim3(nconf), dim3(256), per * nb, h->stream, h->dp, A, nb);
  }
  hipLaunchKernelGGL(k_grad_dq_fw, dim3((nfw + 63) / 64, KSFW), dim3(64), 0, h->stream, h->dp, A, KSFW);
  {
    // k split so that the grid has at least 512 blocks
    const int tiles = ((nfw + 255) / 256) * ((nconf + GCT - 1) / GCT);
    A.ksplit = std::max(1, std::min(nk / 32 + 1, (512 + tiles - 1) / tiles));
    const size_t need = (size_t)A.ksplit * nconf * nfw;
    if (need > h->g_dqcap) {
      if (h->g_dqpart) (void)hipFree(h->g_dqpart);
      HIPCHK(hipMalloc(&h->g_dqpart, sizeof(double) * need));
      h->g_dqcap = need;
    }
    A.dqpart = h->g_dqpart;
    const dim3 gb((nfw + 255) / 256, (nconf + GCT - 1) / GCT, A.ksplit);
    if (precision == FMC_MIXED) hipLaunchKernelGGL((k_grad_dq_block<float>), gb, dim3(256), 0, h->stream, h->dp, A);
    else hipLaunchKernelGGL((k_grad_dq_block<double>), gb, dim3(256), 0, h->stream, h->dp, A);
  }
  hipLaunchKernelGGL(k_grad_pairs, dim3(((nfw + gmax) * GPS + 255) / 256, nconf), dim3(256),
                     sizeof(double) * ((size_t)P.mxegrd + 3), h->stream, h->dp, A);
  HIPCHK(hipGetLastError());
  // ---- closed forms (FP64, host, while the kernels run): self term,
  // charged-system term, ewald3, tail and the tail's type gradients; kept per
  // configuration and combined with the device sums below in the same order
  const double kq = R4PIE0 / P.epsq, C = P.dc_C, alpha = P.alpha, volm = P.volm, twopi = 2.0 * PI;
  const double rc = h->rcut, rc3 = rc * rc * rc, rc9 = rc3 * rc3 * rc3;
  double q2fw = 0.0, qfw = 0.0;
  for (int a = 0; a < nfw; ++a) {
    q2fw += theta->fw_charge[a] * theta->fw_charge[a];
    qfw += theta->fw_charge[a];
  }
  std::vector<int> nfwt(TT, 0), ffwt(TT, 0);
  for (int a = 0; a < nfw; ++a) {
    nfwt[h->fw_type[a]] += 1;
    ffwt[h->fw_type[a]] += h->fw_frz[a] ? 1 : 0;
  }
  std::vector<double> c_sic(nconf), c_fix(nconf), c_e3(nconf), c_lrc(nconf), c_qs(nconf);
  std::vector<double> c_te((size_t)nconf * TT, 0.0), c_ts((size_t)nconf * TT, 0.0);
  for (int c = 0; c < nconf; ++c) {
    double q2 = q2fw, qs = qfw, e3 = 0.0;
    std::vector<double> n(TT, 0.0), f(TT, 0.0);
    for (int t = 0; t < TT; ++t) {
      n[t] = nfwt[t];
      f[t] = ffwt[t];
    }
    for (int g = 1; g <= ng; ++g) {
      int mol = P.locguest[g], nm = nmols[c * ng + g - 1];
      q2 += nm * P.selfq2[g];
      qs += nm * P.chgtmp[g];
      e3 += nm * P.ewald3en[mol];
      for (int i = 0; i < P.natoms[mol]; ++i) {
        n[P.site_t[mol][i]] += nm;
        if (P.site_f[mol][i]) f[P.site_t[mol][i]] += nm;
      }
    }
    const double engsic = -kq * alpha * q2 / SQRPI;
    const double qfix = -(0.5 * PI * kq) * ((qs / alpha) * (qs / alpha) / volm);
    // tail (lrcorrect, vdw_terms.f:581-645) with theta pairs mixed
    double elrc = 0.0;
    double *gte = &c_te[(size_t)c * TT], *gts = &c_ts[(size_t)c * TT];
    for (int i = 1; i <= nt; ++i)
      for (int j = 1; j <= i; ++j) {
        double w = twopi * (n[i] * n[j] - f[i] * f[j]) / volm;
        double two = i != j ? 2.0 : 1.0;
        if (isfw[i] || isfw[j]) {
          double e = std::sqrt(theta->type_eps[i - 1] * theta->type_eps[j - 1]);
          double sg = 0.5 * (theta->type_sig[i - 1] + theta->type_sig[j - 1]);
          double s2 = sg * sg, s6 = s2 * s2 * s2;
          double shape = 4.0 * (s6 * s6 / (9.0 * rc9) - s6 / (3.0 * rc3));
          elrc += w * two * e * shape;
          double deadd_de = w * two * shape;
          double deadd_ds = w * two * 4.0 * e * (12.0 * s6 * s6 / sg / (9.0 * rc9) - 6.0 * s6 / sg / (3.0 * rc3));
          // Lorentz-Berthelot chain rule to the type parameters
          if (i == j) {
            gte[i] += deadd_de;
            gts[i] += deadd_ds;
          } else {
            if (theta->type_eps[i - 1] > 0.0) gte[i] += deadd_de * e / (2.0 * theta->type_eps[i - 1]);
            if (theta->type_eps[j - 1] > 0.0) gte[j] += deadd_de * e / (2.0 * theta->type_eps[j - 1]);
            gts[i] += 0.5 * deadd_ds;
            gts[j] += 0.5 * deadd_ds;
          }
        } else {
          int k = h->pair_row[i * TT + j];
          if (k >= 0) {
            double s2 = h->pair_sig[k] * h->pair_sig[k], s6 = s2 * s2 * s2;
            elrc += w * two * 4.0 * h->pair_eps[k] * (s6 * s6 / (9.0 * rc9) - s6 / (3.0 * rc3));
          }
        }
      }
    c_sic[c] = engsic;
    c_fix[c] = qfix;
    c_e3[c] = e3;
    c_lrc[c] = elrc;
    c_qs[c] = qs;
  }
  // ---- device sums
  std::vector<double> out(outn);
  HIPCHK(hipMemcpyAsync(out.data(), h->g_out, sizeof(double) * outn, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  const double *urec = out.data(), *dq = urec + nconf, *e2 = dq + (size_t)nconf * nfw, *ev = e2 + nconf;
  const double *dpe = ev + nconf, *dps = dpe + (size_t)nconf * NTC, *dqfwp = dps + (size_t)nconf * NTC;
  std::vector<double> dqfw(nfw, 0.0);
  for (int z = 0; z < KSFW; ++z)
    for (int a = 0; a < nfw; ++a) dqfw[a] += dqfwp[(size_t)z * nfw + a];
  for (int c = 0; c < nconf; ++c) {
    const double engsic = c_sic[c], qfix = c_fix[c], e3 = c_e3[c], elrc = c_lrc[c], qs = c_qs[c];
    double *gte = &c_te[(size_t)c * TT], *gts = &c_ts[(size_t)c * TT];
    // pair LJ gradients, chain rule to the type parameters
    for (int i = 1; i <= nt; ++i)
      for (int j = 1; j <= i; ++j) {
        double de = dpe[(size_t)c * NTC + i * ntt + j], ds = dps[(size_t)c * NTC + i * ntt + j];
        if (de == 0.0 && ds == 0.0) continue;
        double e = std::sqrt(theta->type_eps[i - 1] * theta->type_eps[j - 1]);
        if (i == j) {
          gte[i] += de;
          gts[i] += ds;
        } else {
          if (theta->type_eps[i - 1] > 0.0) gte[i] += de * e / (2.0 * theta->type_eps[i - 1]);
          if (theta->type_eps[j - 1] > 0.0) gte[j] += de * e / (2.0 * theta->type_eps[j - 1]);
          gts[i] += 0.5 * ds;
          gts[j] += 0.5 * ds;
        }
      }
    const double eu = P.engunit;
    if (U) U[c] = (C * urec[c] + engsic + qfix + e2[c] - e3 + ev[c] + elrc) / eu;
    if (dU_dq) {
      const double dfix = -(PI * kq) * qs / (alpha * alpha * volm);
      for (int a = 0; a < nfw; ++a) {
        double dself = -2.0 * kq * alpha * theta->fw_charge[a] / SQRPI;
        dU_dq[(size_t)c * nfw + a] = (dqfw[a] + dq[(size_t)c * nfw + a] + dself + dfix) / eu;
      }
    }
    for (int t = 1; t <= nt; ++t) {
      if (dU_deps) dU_deps[(size_t)c * nt + t - 1] = isfw[t] ? gte[t] / eu : 0.0;
      if (dU_dsig) dU_dsig[(size_t)c * nt + t - 1] = isfw[t] ? gts[t] / eu : 0.0;
    }
  }
  return FMC_OK;
}

int fmc_set_resum(fmc_handle *h, int every) {
  if (!h || every < 0) return FMC_ERR_ARG;
  h->hp.resum_every = every;
  HIPCHK(hipMemcpy(h->dp, &h->hp, sizeof(DevParams), hipMemcpyHostToDevice));
  return FMC_OK;
}

int fmc_set_sampling(fmc_handle *h, int nwindsteps, int accepted_only, int max_windows) {
  if (!h || nwindsteps < 0 || max_windows < 1) return FMC_ERR_ARG;
  HIPCHK(hipStreamSynchronize(h->stream));
  // captured graphs hold the old window buffer
  for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  h->graphs.clear();
  if (h->dw.win) HIPCHK(hipFree(h->dw.win));
  h->dw.win = nullptr;
  const size_t n = (size_t)h->nwalkers * (size_t)max_windows;
  HIPCHK(hipMalloc(&h->dw.win, sizeof(fmc_window) * n));
  HIPCHK(hipMemset(h->dw.win, 0, sizeof(fmc_window) * n));
  h->hp.nwindsteps = nwindsteps;
  h->hp.accsample = accepted_only ? 1 : 0;
  h->hp.maxwin = max_windows;
  HIPCHK(hipMemcpy(h->dp, &h->hp, sizeof(DevParams), hipMemcpyHostToDevice));
  std::vector<WState> ws(h->nwalkers);
  HIPCHK(hipMemcpy(ws.data(), h->dw.ws, sizeof(WState) * ws.size(), hipMemcpyDeviceToHost));
  for (auto &s : ws) {
    s.prodcount = s.rollcount = s.nwin = 0;
    memset(s.chain, 0, sizeof s.chain);
    s.overflow &= ~2;
  }
  HIPCHK(hipMemcpy(h->dw.ws, ws.data(), sizeof(WState) * ws.size(), hipMemcpyHostToDevice));
  return FMC_OK;
}

int fmc_get_windows(fmc_handle *h, int w, fmc_window *out, int cap, int *nwin) {
  if (!h || w < 0 || w >= h->nwalkers || cap < 0 || (cap > 0 && !out)) return FMC_ERR_ARG;
  WState s;
  HIPCHK(hipMemcpy(&s, h->dw.ws + w, sizeof s, hipMemcpyDeviceToHost));
  const int n = (int)s.nwin, m = std::min(n, cap);
  if (nwin) *nwin = n;
  if (m > 0)
    HIPCHK(hipMemcpy(out, h->dw.win + (size_t)w * h->hp.maxwin, sizeof(fmc_window) * (size_t)m,
                     hipMemcpyDeviceToHost));
  return FMC_OK;
}

int fmc_set_prob(fmc_handle *h, double griddim, const int gridfactor[3], int nplots, const fmc_prob_plot *plots,
                 int ngrid_out[3]) {
  if (!h || nplots < 0 || nplots > FMC_MAX_PLOTS || (nplots > 0 && (!plots || !(griddim > 0.0)))) return FMC_ERR_ARG;
  DevParams &P = h->hp;
  const int gf[3] = {gridfactor ? gridfactor[0] : 1, gridfactor ? gridfactor[1] : 1, gridfactor ? gridfactor[2] : 1};
  if (gf[0] < 1 || gf[1] < 1 || gf[2] < 1) return FMC_ERR_ARG;
  for (int p = 0; p < nplots; ++p)
    if (plots[p].iguest < 1 || plots[p].iguest > P.ntpguest || plots[p].nsites < 0 ||
        plots[p].nsites > FMC_MAX_SITES + 1)
      return FMC_ERR_ARG;
  HIPCHK(hipStreamSynchronize(h->stream));
  for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  h->graphs.clear();
  if (h->dw.prob) HIPCHK(hipFree(h->dw.prob));
  h->dw.prob = nullptr;
  P.nplots = 0;
  if (nplots > 0) {
    // grid points per axis (gcmc.f:258-281) and the unit cell the positions
    // fold into (gcmc.f:299-306)
    double celprp[10], ucell[9];
    dcell(P.cell, celprp);
    for (int a = 0; a < 3; ++a) {
      const double gd = celprp[a] / std::ceil(celprp[a] / gf[a] / griddim) / gf[a];
      P.pgrid[a] = (int)std::ceil(celprp[a] / (gd * gf[a]));
    }
    for (int i = 0; i < 9; ++i) ucell[i] = P.cell[i] / (double)gf[i / 3];
    invert(ucell, P.prucell);
    P.pgridsize = P.pgrid[0] * P.pgrid[1] * P.pgrid[2];
    for (int p = 0; p < nplots; ++p) {
      P.plot_g[p] = plots[p].iguest;
      P.plot_ns[p] = plots[p].nsites;
      for (int e = 0; e < plots[p].nsites; ++e) P.plot_site[p][e] = plots[p].sites[e];
    }
    const size_t n = (size_t)h->nwalkers * nplots * P.pgridsize;
    HIPCHK(hipMalloc(&h->dw.prob, sizeof(double) * n));
    HIPCHK(hipMemset(h->dw.prob, 0, sizeof(double) * n));
    P.nplots = nplots;
    if (ngrid_out)
      for (int a = 0; a < 3; ++a) ngrid_out[a] = P.pgrid[a];
  }
  HIPCHK(hipMemcpy(h->dp, &P, sizeof(DevParams), hipMemcpyHostToDevice));
  return FMC_OK;
}

int fmc_set_numguests(fmc_handle *h, int nnumg, int cap) {
  if (!h || nnumg < 0 || cap < 0 || (nnumg > 0 && cap < 1)) return FMC_ERR_ARG;
  DevParams &P = h->hp;
  HIPCHK(hipStreamSynchronize(h->stream));
  for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  h->graphs.clear();
  if (h->dw.numg) HIPCHK(hipFree(h->dw.numg));
  h->dw.numg = nullptr;
  P.nnumg = 0;
  P.numg_cap = 0;
  if (nnumg > 0) {
    HIPCHK(hipMalloc(&h->dw.numg, sizeof(fmc_numguests_rec) * (size_t)h->nwalkers * cap));
    HIPCHK(hipMemset(h->dw.numg, 0, sizeof(fmc_numguests_rec) * (size_t)h->nwalkers * cap));
    P.nnumg = nnumg;
    P.numg_cap = cap;
  }
  HIPCHK(hipMemcpy(h->dp, &P, sizeof(DevParams), hipMemcpyHostToDevice));
  std::vector<WState> ws(h->nwalkers);
  HIPCHK(hipMemcpy(ws.data(), h->dw.ws, sizeof(WState) * ws.size(), hipMemcpyDeviceToHost));
  for (auto &s : ws) {
    s.nnumg_n = 0;
    s.overflow &= ~4;
  }
  HIPCHK(hipMemcpy(h->dw.ws, ws.data(), sizeof(WState) * ws.size(), hipMemcpyHostToDevice));
  return FMC_OK;
}

int fmc_get_numguests(fmc_handle *h, int w, fmc_numguests_rec *out, int cap, int *n) {
  if (!h || w < 0 || w >= h->nwalkers || cap < 0 || (cap > 0 && !out)) return FMC_ERR_ARG;
  WState s;
  HIPCHK(hipMemcpy(&s, h->dw.ws + w, sizeof s, hipMemcpyDeviceToHost));
  const int total = (int)s.nnumg_n, m = std::min(total, cap);
  if (n) *n = total;
  if (m > 0)
    HIPCHK(hipMemcpy(out, h->dw.numg + (size_t)w * h->hp.numg_cap, sizeof(fmc_numguests_rec) * (size_t)m,
                     hipMemcpyDeviceToHost));
  return FMC_OK;
}

int fmc_get_prob(fmc_handle *h, int w, int plot, double *out) {
  if (!h || w < 0 || w >= h->nwalkers || plot < 0 || plot >= h->hp.nplots || !out) return FMC_ERR_ARG;
  const DevParams &P = h->hp;
  HIPCHK(hipMemcpy(out, h->dw.prob + ((size_t)w * P.nplots + plot) * P.pgridsize, sizeof(double) * P.pgridsize,
                   hipMemcpyDeviceToHost));
  return FMC_OK;
}

int fmc_production_stats(fmc_handle *h, int w, fmc_guest_stats *out) {
  if (!h || w < 0 || w >= h->nwalkers || !out) return FMC_ERR_ARG;
  WState s;
  HIPCHK(hipMemcpy(&s, h->dw.ws + w, sizeof s, hipMemcpyDeviceToHost));
  std::vector<fmc_window> win((size_t)std::max<long long>(s.nwin, 1));
  int n = 0;
  int rc = fmc_get_windows(h, w, win.data(), (int)win.size(), &n);
  if (rc) return rc;
  return fmc_stats_from_windows(win.data(), n, h->hp.ntpguest, s.temp, out);
}

// Step clock of k_run (fmc_kernels.hpp, FMC_PHASE_CLOCK builds only).
int fmc_phase_clock(unsigned long long *ticks, int n, int clear) {
#ifdef FMC_PHASE_CLOCK
  const int m = fmc::FMC_NPHASE + 3;
  if (!ticks || n < m) return FMC_ERR_ARG;
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpyFromSymbol(ticks, HIP_SYMBOL(fmc::fmc_phase_ticks), sizeof(unsigned long long) * m, 0,
                             hipMemcpyDeviceToHost));
  if (clear) {
    std::vector<unsigned long long> z(m, 0ULL);
    HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(fmc::fmc_phase_ticks), z.data(), sizeof(unsigned long long) * m, 0,
                             hipMemcpyHostToDevice));
  }
  return FMC_OK;
#else
  (void)ticks;
  (void)n;
  (void)clear;
  return FMC_ERR_ARG;  // a library built without the step clock
#endif
}
}  // extern "C"
