// sedcol_api.hip -- host side of the C ABI declared in include/sedcol.h.
//
// A handle owns the device copy of one sediment domain (geometry, forcing,
// state, accumulators, diagnostics) and one HIP stream.  Every entry point
// validates its arguments, returns a sedcol_status code and records a message
// for sedcol_last_error(); nothing aborts the process.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "sedcol_kernels.hpp"

using sc::Dev;
using sc::NV;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
  g_err = msg;
  return code;
}

#define HIPOK(call)                                                                      \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(SEDCOL_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));       \
  } while (0)

constexpr int kLpc[3] = {16, 32, 64};
// steps per launch of the one-workgroup adaptive Euler (at most
// kAeFusedMaxAttempts attempts per launch, i.e. about 1000 per step)
constexpr int kAeFusedSteps = 4096;

// Workspace of sedcol_rhs_noalloc: staging of its input state, its
// right-hand side and top fluxes, and (Fortran layout) of the caller's
// array, so that the call needs no allocation.  One copy per device, held
// from module load until the process ends; host code reaches it only
// through the address hipGetSymbolAddress returns on that device.
constexpr size_t kNoallocBytes = size_t(64) << 20;
__device__ __align__(256) double scratch[kNoallocBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_noalloc_mu[kMaxDevices];  // one call at a time per device

// Device address of `scratch` on device dev (which must be current): looked
// up once per device, at the first sedcol_create there.
hipError_t scratch_address(int dev, double **out)
{
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = static_cast<double *>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

}  // namespace

struct sedcol_handle {
  int device = 0;  // current at sedcol_create; every call runs on it
  int ncol = 0, kmax = 0, layout = 0;
  long long ncell = 0;
  std::vector<int> off;   // ncol+1
  std::vector<int> mask;  // ncol
  std::vector<int> bins[3];
  int *d_off = nullptr;
  int *d_list[3] = {nullptr, nullptr, nullptr};
  double *d_dz = nullptr, *d_dzc = nullptr, *d_phi = nullptr, *d_phiI = nullptr, *d_bfac = nullptr;
  double *d_biot = nullptr;
  double *d_fcap = nullptr;  // distributed-POM flux cap (bcup_part = 4), ncell
  // bioturbation profile 3: parameters, depth terms, the caller's arithmetic
  // form (profile 3 runs in the reference form)
  bool zw_on = false;
  sedcol_bioturbation_zw zwp{};
  double *d_ze1 = nullptr, *d_ze2 = nullptr, *d_zcum = nullptr;
  int32_t user_arith = 0;
  int nrec = 0;
  long long fup_rec_stride = 0;
  double t0 = 0.0, dt_rec = INFINITY;
  double *d_temp = nullptr, *d_cbw = nullptr, *d_fup = nullptr;
  size_t forcing_cap = 0;  // records allocated
  double *d_conc = nullptr, *d_cellacc = nullptr, *d_colacc = nullptr, *d_flux = nullptr;
  sedcol_coldiag *d_diag = nullptr;
  double *d_scr = nullptr;  // NV*ncell scratch (rhs in/out, adaptive Euler rates)
  double *d_scr2 = nullptr; // NV*ncell scratch
  double *d_small = nullptr;  // NV*ncol + ncol scratch
  double *d_warm = nullptr;   // scale, offset (NV*kWarmMaxNodes each) and eps (NV) of a warm start
  double *d_noalloc = nullptr;  // this device's sedcol_rhs_noalloc workspace (not owned)
  int *d_flag = nullptr;     // [0] error flags, [1..3] work-queue counters per bin, [4] windowed columns pending
  int *h_flag = nullptr;     // pinned host copy of d_flag[0..4]
  int *d_work = nullptr;     // ncol: attempts of each column in the last adaptive launch
  int *d_wrec = nullptr;     // ncol: adaptive position between windows (record, time left, attempts)
  double *d_wrem = nullptr;
  long long *d_watt = nullptr;
  sc::EmpCtl *d_ctl = nullptr;
  sc::EmpCtl h_ctl{};
  sedcol_schedule sched{0, 0, 1, 0};
  int lpc[3] = {16, 32, 64};  // lanes per column of each bin in the adaptive kernel
  int blk[3] = {256, 256, 256};  // workgroup size of each bin in the adaptive kernel
  int lpl[3] = {1, 1, 1};        // layers per lane of each bin in the adaptive kernel
  hipStream_t aux[2] = {nullptr, nullptr};  // for concurrent bins
  hipEvent_t evfork = nullptr, evjoin[2] = {nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<int> h_work;   // host copy, used to order the work queues
  Dev *d_devview = nullptr;  // device copy of dev(), read by the adaptive kernel
  sedcol_params *d_params = nullptr;
  Dev h_devview{};           // host staging of the copies (kept alive for async copies)
  sedcol_params h_params{};
  sedcol_params p{};
  bool have_params = false, have_geom = false, have_forcing = false;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  int last_launches = 0;
  double last_min_dt = 1.e20;

  Dev dev() const
  {
    Dev d;
    d.ncol = ncol;
    d.ncell = ncell;
    d.off = d_off;
    d.dz = d_dz;
    d.dzc = d_dzc;
    d.phi = d_phi;
    d.phiI = d_phiI;
    d.bfac = d_bfac;
    d.biot = d_biot;
    d.fcap = d_fcap;
    d.zw.on = zw_on ? 1 : 0;
    d.zw.L1 = zwp.L1;
    d.zw.L2 = zwp.L2;
    d.zw.beta = zwp.beta;
    d.zw.b = zwp.b;
    // mmol C m-3 -> mg C g-1 dry mass (fabm_sediment_driver.F90:383,386)
    d.zw.f_labile = 1.0 / 1.2 * 12.01 / zwp.dry_density / 1000.0;
    d.zw.f_semilabile = 1.0 / 6.0 * 12.01 / zwp.dry_density / 1000.0;
    d.zw.e1 = d_ze1;
    d.zw.e2 = d_ze2;
    d.zw.cum = d_zcum;
    d.nrec = nrec;
    d.t0 = t0;
    d.dt_rec = dt_rec;
    d.temp = d_temp;
    d.cbw = d_cbw;
    d.fup = d_fup;
    d.fup_rec_stride = fup_rec_stride;
    d.conc = d_conc;
    d.cellacc = d_cellacc;
    d.colacc = d_colacc;
    d.flux = d_flux;
    d.diag = d_diag;
    d.wrec = d_wrec;
    d.wrem = d_wrem;
    d.watt = d_watt;
    return d;
  }
};

namespace {

// Fortran layout index of cell (c, k) and state (c, k, n) for a uniform grid
inline long long fidx(const sedcol_handle *h, int c, int k) { return (long long)k * h->ncol + c; }

// host cell array (caller layout) -> packed
void pack_cells(const sedcol_handle *h, const double *src, std::vector<double> &dst)
{
  dst.resize(h->ncell);
  if (h->layout == SEDCOL_LAYOUT_PACKED) {
    std::memcpy(dst.data(), src, sizeof(double) * h->ncell);
    return;
  }
  for (int c = 0; c < h->ncol; ++c)
    for (int k = 0; k < h->kmax; ++k) dst[(long long)c * h->kmax + k] = src[fidx(h, c, k)];
}

void pack_state(const sedcol_handle *h, const double *src, std::vector<double> &dst)
{
  dst.resize((size_t)NV * h->ncell);
  if (h->layout == SEDCOL_LAYOUT_PACKED) {
    std::memcpy(dst.data(), src, sizeof(double) * NV * h->ncell);
    return;
  }
  const long long plane = (long long)h->ncol * h->kmax;
  for (int n = 0; n < NV; ++n)
    for (int k = 0; k < h->kmax; ++k)
      for (int c = 0; c < h->ncol; ++c)
        dst[n * h->ncell + (long long)c * h->kmax + k] = src[n * plane + fidx(h, c, k)];
}

void unpack_state(const sedcol_handle *h, const std::vector<double> &src, double *dst)
{
  if (h->layout == SEDCOL_LAYOUT_PACKED) {
    std::memcpy(dst, src.data(), sizeof(double) * NV * h->ncell);
    return;
  }
  const long long plane = (long long)h->ncol * h->kmax;
  for (int n = 0; n < NV; ++n)
    for (int k = 0; k < h->kmax; ++k)
      for (int c = 0; c < h->ncol; ++c)
        dst[n * plane + fidx(h, c, k)] = src[n * h->ncell + (long long)c * h->kmax + k];
}

// Makes the handle's device current for the duration of one entry point and
// restores the caller's current device afterwards: a handle belongs to the
// device that was current at sedcol_create.
struct DeviceScope {
  int prev = -1, dev;
  explicit DeviceScope(int d) : dev(d)
  {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceScope()
  {
    if (prev >= 0 && prev != dev) (void)hipSetDevice(prev);
  }
  DeviceScope(const DeviceScope &) = delete;
  DeviceScope &operator=(const DeviceScope &) = delete;
};

int check_handle(sedcol_handle *h, bool need_state_inputs)
{
  if (!h) return fail(SEDCOL_EINVAL, "null handle");
  if (need_state_inputs) {
    if (!h->have_params) return fail(SEDCOL_EINVAL, "sedcol_set_params has not been called");
    if (!h->have_geom) return fail(SEDCOL_EINVAL, "sedcol_set_geometry has not been called");
    if (!h->have_forcing) return fail(SEDCOL_EINVAL, "no forcing set (sedcol_set_forcing)");
    if (h->p.bcup_part == 4 && !h->d_fcap)
      return fail(SEDCOL_EINVAL, "bcup_part = 4 needs the flux cap (sedcol_set_flux_cap)");
  }
  return SEDCOL_OK;
}

int alloc_forcing(sedcol_handle *h, int nrec)
{
  if ((size_t)nrec <= h->forcing_cap) return SEDCOL_OK;
  (void)hipFree(h->d_temp);
  (void)hipFree(h->d_cbw);
  (void)hipFree(h->d_fup);
  h->d_temp = h->d_cbw = h->d_fup = nullptr;
  h->forcing_cap = 0;
  HIPOK(hipMalloc(&h->d_temp, sizeof(double) * (size_t)nrec * h->ncol));
  HIPOK(hipMalloc(&h->d_cbw, sizeof(double) * (size_t)nrec * NV * h->ncol));
  HIPOK(hipMalloc(&h->d_fup, sizeof(double) * (size_t)nrec * NV * h->ncol));
  h->forcing_cap = (size_t)nrec;  // every array sized for nrec records
  return SEDCOL_OK;
}

int reset_diag(sedcol_handle *h)
{
  std::vector<sedcol_coldiag> dg(h->ncol);
  for (auto &x : dg) {
    std::memset(&x, 0, sizeof x);
    x.dt_minacc = INFINITY;
    x.p_min = 1.0;
  }
  HIPOK(hipMemcpyAsync(h->d_diag, dg.data(), sizeof(sedcol_coldiag) * h->ncol, hipMemcpyHostToDevice, h->stream));
  HIPOK(hipStreamSynchronize(h->stream));
  return SEDCOL_OK;
}

int read_flag(sedcol_handle *h, int *flag)
{
  HIPOK(hipMemcpyAsync(flag, h->d_flag, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPOK(hipStreamSynchronize(h->stream));
  return SEDCOL_OK;
}

int flag_status(int flag)
{
  if (flag & 1) return fail(SEDCOL_ENAN, "NaN detected in the state");
  if (flag & 2) return fail(SEDCOL_ELIMIT, "adaptive attempt limit reached");
  return SEDCOL_OK;
}

// forcing record for time t (piecewise constant; first/last extend outward)
int record_of(const sedcol_handle *h, double t)
{
  if (h->nrec <= 1) return 0;
  double q = std::floor((t - h->t0) / h->dt_rec);
  if (!(q > 0.0)) return 0;
  if (q >= h->nrec - 1) return h->nrec - 1;
  return (int)q;
}

// smallest lanes per column that holds the largest column of bin b with
// lpl layers per lane (two layers per lane need at least 16 lanes: lanes
// 0..NV-1 own the top-flux accumulators)
int bin_min_lpc(const sedcol_handle *h, int b, int lpl = 1)
{
  int kmax = 1;
  for (int c : h->bins[b]) kmax = std::max(kmax, h->off[c + 1] - h->off[c]);
  const int lanes = (kmax + lpl - 1) / lpl;
  return lanes <= 16 ? 16 : lanes <= 32 ? 32 : 64;
}

int check_tuning(const sedcol_handle *h, const sedcol_tuning *t)
{
  for (int b = 0; b < SEDCOL_NBIN; ++b) {
    const int l = t->lanes_per_column[b], g = t->columns_per_group[b];
    const int lpl = t->layers_per_lane[b] == 0 ? 1 : t->layers_per_lane[b];
    if (lpl != 1 && lpl != 2) return fail(SEDCOL_EINVAL, "layers_per_lane must be 1 or 2");
    if (l != 16 && l != 32 && l != 64) return fail(SEDCOL_EINVAL, "lanes_per_column must be 16, 32 or 64");
    if (l < bin_min_lpc(h, b, lpl)) return fail(SEDCOL_EINVAL, "lanes_per_column below the bin's largest column");
    const int w = l * g;
    if (g < 1 || (w != 64 && w != 128 && w != 256))
      return fail(SEDCOL_EINVAL, "lanes_per_column * columns_per_group must be 64, 128 or 256");
  }
  return SEDCOL_OK;
}

// kernel shape bin b runs with under the current parameters: an installed
// two-layers-per-lane shape falls back to one layer per lane (with enough
// lanes for the bin) where that kernel does not cover the parameters
// (linearly implicit stepper, geometric-mean weights, bioturbation profile
// 3, BcUp 4)
void effective_shape(const sedcol_handle *h, int b, int *lpc, int *blk, int *lpl)
{
  *lpc = h->lpc[b];
  *blk = h->blk[b];
  *lpl = h->lpl[b];
  if (*lpl == 2 && !sc::emp2_supported(h->p)) {
    *lpl = 1;
    *lpc = std::max(*lpc, bin_min_lpc(h, b, 1));
    *blk = std::max(*blk, *lpc);
  }
}

void drop_graph(sedcol_handle *h)
{
  if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->graph_exec = nullptr;
  h->graph = nullptr;
}

// One adaptive launch per non-empty bin (only bin `only` when only >= 0):
// in order on the handle's stream, or as concurrent branches forked from it.
// Work-queue counters and the pending count must be zero beforehand.
int launch_bins(sedcol_handle *h, int only, int *launches)
{
  int list[3], nb = 0;
  for (int b = 0; b < 3; ++b)
    if (!h->bins[b].empty() && (only < 0 || only == b)) list[nb++] = b;
  const bool fork = h->sched.concurrent_bins && nb > 1;
  if (fork) HIPOK(hipEventRecord(h->evfork, h->stream));
  for (int i = 0; i < nb; ++i) {
    const int b = list[i];
    hipStream_t st = h->stream;
    if (fork && i > 0) {
      st = h->aux[i - 1];
      HIPOK(hipStreamWaitEvent(st, h->evfork, 0));
    }
    int lpc, blk, lpl;
    effective_shape(h, b, &lpc, &blk, &lpl);
    if (!sc::launch_emp_adaptive(lpc, blk, lpl, st, h->d_devview, h->d_list[b], (int)h->bins[b].size(), h->p,
                                 h->d_params, h->d_ctl, h->d_flag + 1 + b, h->d_flag + 4, h->d_work, h->d_flag))
      return fail(SEDCOL_EINVAL, "unsupported lanes per column / workgroup size");
    ++*launches;
    if (fork && i > 0) {
      HIPOK(hipEventRecord(h->evjoin[i - 1], st));
      HIPOK(hipStreamWaitEvent(h->stream, h->evjoin[i - 1], 0));
    }
  }
  HIPOK(hipGetLastError());
  return SEDCOL_OK;
}

// pending count and error flags after the work queued so far
int read_progress(sedcol_handle *h, int *flag, int *pending)
{
  HIPOK(hipMemcpyAsync(h->h_flag, h->d_flag, sizeof(int) * 5, hipMemcpyDeviceToHost, h->stream));
  HIPOK(hipStreamSynchronize(h->stream));
  *flag = h->h_flag[0];
  *pending = h->h_flag[4];
  return SEDCOL_OK;
}

// graph of sched.graph_windows windows: (reset counters, one launch per bin)
int capture_graph(sedcol_handle *h)
{
  drop_graph(h);
  // device queries of the launcher (resident blocks) outside the capture
  for (int b = 0; b < 3; ++b) {
    int lpc, blk, lpl;
    effective_shape(h, b, &lpc, &blk, &lpl);
    if (!h->bins[b].empty() && sc::emp_resident_blocks(lpc, blk, lpl, h->p) <= 0)
      return fail(SEDCOL_EINVAL, "unsupported lanes per column / workgroup size");
  }
  int launches = 0, rc = SEDCOL_OK;
  HIPOK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  for (int w = 0; w < h->sched.graph_windows && rc == SEDCOL_OK; ++w) {
    sc::launch_fill_int(h->stream, h->d_flag + 1, 4, 0);
    rc = launch_bins(h, -1, &launches);
  }
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(h->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(SEDCOL_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  h->graph = g;
  HIPOK(hipGraphInstantiate(&h->graph_exec, h->graph, nullptr, nullptr, 0));
  return SEDCOL_OK;
}

// Adaptive EMP over [t_start, t_end) for every bin (or bin `only`), by the
// handle's schedule.  Sets *launches to the adaptive kernel launches made.
int run_adaptive(sedcol_handle *h, double t_start, double t_end, int only, int *launches)
{
  if (h->p.adaptive_method == SEDCOL_ROS2_ADAPTIVE && (h->zw_on || h->p.bcup_part == 4))
    return fail(SEDCOL_EUNSUPPORTED,
                "the linearly implicit stepper does not support bioturbation profile 3 or bcup_part = 4");
  const int r0 = record_of(h, t_start), r1 = record_of(h, t_end);
  const sedcol_schedule &sc_ = h->sched;
  const bool windowed = sc_.window > 0;
  HIPOK(hipMemsetAsync(h->d_flag, 0, sizeof(int) * 8, h->stream));
  h->h_devview = h->dev();
  h->h_params = h->p;
  h->h_ctl.t_start = t_start;
  h->h_ctl.t_end = t_end;
  h->h_ctl.max_attempts = 1ll << 40;
  h->h_ctl.r0 = r0;
  h->h_ctl.r1 = r1;
  h->h_ctl.window = windowed ? sc_.window : 0;
  h->h_ctl.resume = windowed ? 1 : 0;
  HIPOK(hipMemcpyAsync(h->d_devview, &h->h_devview, sizeof(Dev), hipMemcpyHostToDevice, h->stream));
  HIPOK(hipMemcpyAsync(h->d_params, &h->h_params, sizeof(sedcol_params), hipMemcpyHostToDevice, h->stream));
  HIPOK(hipMemcpyAsync(h->d_ctl, &h->h_ctl, sizeof(sc::EmpCtl), hipMemcpyHostToDevice, h->stream));
  *launches = 0;
  if (!windowed) return launch_bins(h, only, launches);
  for (int b = 0; b < 3; ++b)
    if (!h->bins[b].empty() && (only < 0 || only == b))
      sc::launch_emp_begin(h->stream, h->d_list[b], (int)h->bins[b].size(), r0, h->d_wrec, h->d_wrem, h->d_watt);
  HIPOK(hipGetLastError());
  int rc, flag = 0, pending = 0;
  const bool graph = sc_.graph_windows > 0 && only < 0;
  if (graph && !h->graph_exec && (rc = capture_graph(h))) return rc;
  int nb = 0;
  for (int b = 0; b < 3; ++b) nb += !h->bins[b].empty();
  for (;;) {
    if (graph) {
      HIPOK(hipGraphLaunch(h->graph_exec, h->stream));
      *launches += sc_.graph_windows * nb;
    } else {
      for (int w = 0; w < sc_.check_every; ++w) {
        sc::launch_fill_int(h->stream, h->d_flag + 1, 4, 0);
        if ((rc = launch_bins(h, only, launches))) return rc;
      }
    }
    if ((rc = read_progress(h, &flag, &pending))) return rc;
    if (pending == 0 || flag) break;
  }
  return SEDCOL_OK;
}

// NV*ncell scratch arrays, allocated on first use
int ensure_scratch(sedcol_handle *h)
{
  if (!h->d_scr) HIPOK(hipMalloc(&h->d_scr, sizeof(double) * NV * (size_t)h->ncell));
  if (!h->d_scr2) HIPOK(hipMalloc(&h->d_scr2, sizeof(double) * NV * (size_t)h->ncell));
  return SEDCOL_OK;
}

// upload the current bins as the work lists (after a change of the active set)
int upload_bins(sedcol_handle *h)
{
  for (int b = 0; b < 3; ++b)
    if (!h->bins[b].empty())
      HIPOK(hipMemcpyAsync(h->d_list[b], h->bins[b].data(), sizeof(int) * h->bins[b].size(),
                           hipMemcpyHostToDevice, h->stream));
  HIPOK(hipStreamSynchronize(h->stream));
  drop_graph(h);  // the graph holds the list lengths
  return SEDCOL_OK;
}
struct CkHeader {
  char magic[8];
  int64_t ncol, ncell, nvar, version;
};
constexpr char kCkMagic[8] = {'S', 'E', 'D', 'C', 'O', 'L', 'C', 'K'};

// device arrays of a checkpoint, in blob order
std::vector<std::pair<void *, size_t>> ck_parts(sedcol_handle *h)
{
  const size_t nc = (size_t)h->ncol, ncell = (size_t)h->ncell;
  return {{h->d_conc, sizeof(double) * NV * ncell},  {h->d_cellacc, sizeof(double) * 10 * ncell},
          {h->d_colacc, sizeof(double) * 2 * NV * nc}, {h->d_flux, sizeof(double) * NV * nc},
          {h->d_diag, sizeof(sedcol_coldiag) * nc}};
}

int64_t ck_bytes(sedcol_handle *h)
{
  int64_t n = sizeof(CkHeader);
  for (auto &q : ck_parts(h)) n += (int64_t)q.second;
  return n;
}
// steady-state residual of the listed columns (and their top fluxes, NV x
// ncol, when flux is not null) into host arrays
int residual_flux(sedcol_handle *h, const double *cfloor, double *resid, double *flux)
{
  int rc;
  if ((flux || h->ncol < NV) && (rc = ensure_scratch(h))) return rc;
  double *d_cf = h->d_small + (size_t)NV * h->ncol;  // NV slots needed
  double *d_res = h->d_small;
  if (h->ncol < NV) d_cf = h->d_scr2;
  double *d_fl = flux ? h->d_scr : nullptr;  // NV*ncell >= NV*ncol
  HIPOK(hipMemcpyAsync(d_cf, cfloor, sizeof(double) * NV, hipMemcpyHostToDevice, h->stream));
  HIPOK(hipMemsetAsync(d_res, 0, sizeof(double) * h->ncol, h->stream));
  if (d_fl) HIPOK(hipMemsetAsync(d_fl, 0, sizeof(double) * NV * h->ncol, h->stream));
  const Dev d = h->dev();
  for (int b = 0; b < 3; ++b)
    if (!h->bins[b].empty())
      sc::launch_residual(kLpc[b], h->stream, d, h->d_list[b], (int)h->bins[b].size(), h->p, 0, d_cf, d_res, d_fl);
  HIPOK(hipGetLastError());
  HIPOK(hipMemcpyAsync(resid, d_res, sizeof(double) * h->ncol, hipMemcpyDeviceToHost, h->stream));
  if (d_fl) HIPOK(hipMemcpyAsync(flux, d_fl, sizeof(double) * NV * h->ncol, hipMemcpyDeviceToHost, h->stream));
  HIPOK(hipStreamSynchronize(h->stream));
  return SEDCOL_OK;
}
}  // namespace

extern "C" {

const char *sedcol_last_error(void) { return g_err.c_str(); }

int sedcol_create(const sedcol_desc *desc, sedcol_handle **out)
{
  if (!desc || !out) return fail(SEDCOL_EINVAL, "null argument");
  *out = nullptr;
  if (desc->nvar != NV) return fail(SEDCOL_EINVAL, "nvar must be SEDCOL_NVAR (11)");
  if (desc->ncol < 1) return fail(SEDCOL_EINVAL, "ncol must be positive");
  if (desc->kmax < 1 || desc->kmax > SEDCOL_KMAX) return fail(SEDCOL_EINVAL, "kmax outside 1..SEDCOL_KMAX");
  if (desc->layout != SEDCOL_LAYOUT_FORTRAN && desc->layout != SEDCOL_LAYOUT_PACKED)
    return fail(SEDCOL_EINVAL, "unknown layout");
  if (desc->layout == SEDCOL_LAYOUT_FORTRAN && desc->nlayers)
    return fail(SEDCOL_EINVAL, "the Fortran layout needs a uniform number of layers (nlayers = NULL)");
  sedcol_handle *h = new sedcol_handle();
  h->ncol = desc->ncol;
  h->kmax = desc->kmax;
  h->layout = desc->layout;
  h->off.resize(h->ncol + 1);
  h->mask.assign(h->ncol, 0);
  h->off[0] = 0;
  for (int c = 0; c < h->ncol; ++c) {
    int K = desc->nlayers ? desc->nlayers[c] : desc->kmax;
    if (K < 1 || K > desc->kmax) {
      delete h;
      return fail(SEDCOL_EINVAL, "nlayers outside 1..kmax");
    }
    h->off[c + 1] = h->off[c] + K;
    if (desc->mask) h->mask[c] = desc->mask[c] != 0;
  }
  h->ncell = h->off[h->ncol];
  for (int c = 0; c < h->ncol; ++c) {
    if (h->mask[c]) continue;
    int K = h->off[c + 1] - h->off[c];
    h->bins[K <= 16 ? 0 : K <= 32 ? 1 : 2].push_back(c);
  }
  auto cleanup = [&](int code) {
    sedcol_destroy(h);
    return code;
  };
#define HIPOK_C(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return cleanup(fail(SEDCOL_EHIP, std::string(#call) + ": " + hipGetErrorString(e_))); \
  } while (0)
  HIPOK_C(hipGetDevice(&h->device));
  HIPOK_C(scratch_address(h->device, &h->d_noalloc));
  HIPOK_C(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  HIPOK_C(hipEventCreate(&h->ev0));
  HIPOK_C(hipEventCreate(&h->ev1));
  const size_t cb = sizeof(double) * (size_t)h->ncell;
  HIPOK_C(hipMalloc(&h->d_off, sizeof(int) * (h->ncol + 1)));
  HIPOK_C(hipMemcpy(h->d_off, h->off.data(), sizeof(int) * (h->ncol + 1), hipMemcpyHostToDevice));
  for (int b = 0; b < 3; ++b)
    if (!h->bins[b].empty()) {
      HIPOK_C(hipMalloc(&h->d_list[b], sizeof(int) * h->bins[b].size()));
      HIPOK_C(hipMemcpy(h->d_list[b], h->bins[b].data(), sizeof(int) * h->bins[b].size(), hipMemcpyHostToDevice));
    }
  HIPOK_C(hipMalloc(&h->d_dz, cb));
  HIPOK_C(hipMalloc(&h->d_dzc, cb));
  HIPOK_C(hipMalloc(&h->d_phi, cb));
  HIPOK_C(hipMalloc(&h->d_phiI, cb));
  HIPOK_C(hipMalloc(&h->d_bfac, cb));
  HIPOK_C(hipMalloc(&h->d_biot, sizeof(double) * h->ncol));
  HIPOK_C(hipMalloc(&h->d_conc, NV * cb));
  HIPOK_C(hipMalloc(&h->d_cellacc, 10 * cb));
  HIPOK_C(hipMalloc(&h->d_colacc, sizeof(double) * 2 * NV * h->ncol));
  HIPOK_C(hipMalloc(&h->d_flux, sizeof(double) * NV * h->ncol));
  HIPOK_C(hipMalloc(&h->d_diag, sizeof(sedcol_coldiag) * h->ncol));
  HIPOK_C(hipMalloc(&h->d_small, sizeof(double) * (NV + 1) * h->ncol));
  HIPOK_C(hipMalloc(&h->d_flag, sizeof(int) * 8));  // error flag, one work-queue counter per bin, pending
  HIPOK_C(hipMemset(h->d_flag, 0, sizeof(int) * 8));
  HIPOK_C(hipHostMalloc((void **)&h->h_flag, sizeof(int) * 8, hipHostMallocDefault));
  HIPOK_C(hipMalloc(&h->d_work, sizeof(int) * h->ncol));
  HIPOK_C(hipMalloc(&h->d_wrec, sizeof(int) * h->ncol));
  HIPOK_C(hipMalloc(&h->d_wrem, sizeof(double) * h->ncol));
  HIPOK_C(hipMalloc(&h->d_watt, sizeof(long long) * h->ncol));
  HIPOK_C(hipMalloc(&h->d_ctl, sizeof(sc::EmpCtl)));
  for (int a = 0; a < 2; ++a) {
    HIPOK_C(hipStreamCreateWithFlags(&h->aux[a], hipStreamNonBlocking));
    HIPOK_C(hipEventCreateWithFlags(&h->evjoin[a], hipEventDisableTiming));
  }
  HIPOK_C(hipEventCreateWithFlags(&h->evfork, hipEventDisableTiming));
  HIPOK_C(hipMemset(h->d_work, 0, sizeof(int) * h->ncol));
  HIPOK_C(hipMalloc(&h->d_devview, sizeof(Dev)));
  HIPOK_C(hipMalloc(&h->d_params, sizeof(sedcol_params)));
  HIPOK_C(hipMemset(h->d_conc, 0, NV * cb));
  HIPOK_C(hipMemset(h->d_cellacc, 0, 10 * cb));
  HIPOK_C(hipMemset(h->d_colacc, 0, sizeof(double) * 2 * NV * h->ncol));
  HIPOK_C(hipMemset(h->d_flux, 0, sizeof(double) * NV * h->ncol));
#undef HIPOK_C
  int rc = reset_diag(h);
  if (rc) return cleanup(rc);
  *out = h;
  return SEDCOL_OK;
}

int sedcol_destroy(sedcol_handle *h)
{
  if (!h) return SEDCOL_OK;
  const DeviceScope dscope(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  void *ptrs[] = {h->d_off, h->d_list[0], h->d_list[1], h->d_list[2], h->d_dz, h->d_dzc, h->d_phi, h->d_phiI,
                  h->d_bfac, h->d_biot, h->d_temp, h->d_cbw, h->d_fup, h->d_conc, h->d_cellacc, h->d_colacc,
                  h->d_flux, h->d_diag, h->d_scr, h->d_scr2, h->d_small, h->d_flag, h->d_work,
                  h->d_devview, h->d_params, h->d_wrec, h->d_wrem, h->d_watt, h->d_ctl, h->d_fcap,
                  h->d_ze1, h->d_ze2, h->d_zcum, h->d_warm};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  if (h->h_flag) (void)hipHostFree(h->h_flag);
  drop_graph(h);
  for (int a = 0; a < 2; ++a) {
    if (h->aux[a]) (void)hipStreamDestroy(h->aux[a]);
    if (h->evjoin[a]) (void)hipEventDestroy(h->evjoin[a]);
  }
  if (h->evfork) (void)hipEventDestroy(h->evfork);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return SEDCOL_OK;
}

int sedcol_set_params(sedcol_handle *h, const sedcol_params *p)
{
  if (!h || !p) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  if (!(p->bcup_dis >= 1 && p->bcup_dis <= 3)) return fail(SEDCOL_EINVAL, "bcup_dis must be 1, 2 or 3");
  if (!(p->bcup_part >= 1 && p->bcup_part <= 4)) return fail(SEDCOL_EINVAL, "bcup_part must be 1, 2, 3 or 4");
  if (!(p->bcdown >= 1 && p->bcdown <= 3)) return fail(SEDCOL_EINVAL, "bcdown must be 1, 2 or 3");
  if (!(p->emp_dt_min > 0.0) || !(p->emp_dt_max >= p->emp_dt_min))
    return fail(SEDCOL_EINVAL, "need 0 < emp_dt_min <= emp_dt_max");
  if (!(p->emp_rtol >= 0.0) || !(p->emp_atol > 0.0)) return fail(SEDCOL_EINVAL, "need emp_rtol >= 0, emp_atol > 0");
  if (p->arith != SEDCOL_ARITH_REFERENCE && p->arith != SEDCOL_ARITH_FAST)
    return fail(SEDCOL_EINVAL, "arith must be SEDCOL_ARITH_REFERENCE or SEDCOL_ARITH_FAST");
  if (p->emp_weights != 0 && p->emp_weights != 1) return fail(SEDCOL_EINVAL, "emp_weights must be 0 or 1");
  if (p->adaptive_method != 0 && p->adaptive_method != SEDCOL_EMP_ADAPTIVE &&
      p->adaptive_method != SEDCOL_ROS2_ADAPTIVE)
    return fail(SEDCOL_EINVAL, "adaptive_method must be 0, SEDCOL_EMP_ADAPTIVE or SEDCOL_ROS2_ADAPTIVE");
  h->p = *p;
  h->user_arith = p->arith;
  if (h->zw_on) h->p.arith = sc::ARITH_EXT;
  h->have_params = true;
  drop_graph(h);  // the kernel instance depends on the arithmetic form
  return SEDCOL_OK;
}

namespace {
int zw_geometry(sedcol_handle *h)
{
  const Dev d = h->dev();
  for (int b = 0; b < 3; ++b)
    if (!h->bins[b].empty())
      sc::launch_zw_geometry(h->stream, d, h->d_list[b], (int)h->bins[b].size(), h->zwp.k_l, h->d_ze1, h->d_ze2,
                             h->d_zcum);
  HIPOK(hipGetLastError());
  HIPOK(hipStreamSynchronize(h->stream));
  return SEDCOL_OK;
}
}  // namespace

int sedcol_set_bioturbation_zw(sedcol_handle *h, const sedcol_bioturbation_zw *zw)
{
  if (!h) return fail(SEDCOL_EINVAL, "null handle");
  const DeviceScope dscope(h->device);
  if (!zw) {
    h->zw_on = false;
    h->p.arith = h->user_arith;
    drop_graph(h);
    return SEDCOL_OK;
  }
  if (!h->have_geom) return fail(SEDCOL_EINVAL, "sedcol_set_geometry has not been called");
  const double v[6] = {zw->k_l, zw->L1, zw->L2, zw->beta, zw->b, zw->dry_density};
  for (double x : v)
    if (!std::isfinite(x)) return fail(SEDCOL_EINVAL, "profile 3 parameters must be finite");
  if (!(zw->dry_density > 0.0) || !(zw->beta > 0.0)) return fail(SEDCOL_EINVAL, "need dry_density > 0, beta > 0");
  if (!h->d_ze1) {
    HIPOK(hipMalloc(&h->d_ze1, sizeof(double) * (size_t)h->ncell));
    HIPOK(hipMalloc(&h->d_ze2, sizeof(double) * (size_t)h->ncell));
    HIPOK(hipMalloc(&h->d_zcum, sizeof(double) * (size_t)h->ncol));
  }
  h->zwp = *zw;
  h->zw_on = true;
  int rc = zw_geometry(h);
  if (rc) {
    h->zw_on = false;
    return rc;
  }
  h->p.arith = sc::ARITH_EXT;
  drop_graph(h);
  return SEDCOL_OK;
}

int sedcol_set_geometry(sedcol_handle *h, const double *dz, const double *dzc, const double *phi,
                        const double *phi_intf, const double *bioturb_factor, const double *bioturbation)
{
  if (!h || !dz || !dzc || !phi || !phi_intf || !bioturb_factor || !bioturbation)
    return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  std::vector<double> a[5];
  pack_cells(h, dz, a[0]);
  pack_cells(h, dzc, a[1]);
  pack_cells(h, phi, a[2]);
  pack_cells(h, phi_intf, a[3]);
  pack_cells(h, bioturb_factor, a[4]);
  // fabm_sed_check_domain (fabm_sediment_driver.F90:488-545) for wet columns;
  // porosity 1 is excluded too: particulate divergence divides by 1 - phi
  for (int c = 0; c < h->ncol; ++c) {
    if (h->mask[c]) continue;
    if (!(bioturbation[c] >= 0.0) || !std::isfinite(bioturbation[c]))
      return fail(SEDCOL_EDOMAIN, "bioturbation coefficient must be finite and >= 0 (column " + std::to_string(c) + ")");
    for (int i = h->off[c]; i < h->off[c + 1]; ++i) {
      const int k = i - h->off[c];
      if (!(a[2][i] > 0.0 && a[2][i] < 1.0))
        return fail(SEDCOL_EDOMAIN, "sediment porosity outside (0,1) at column " + std::to_string(c) + " layer " + std::to_string(k));
      if (!(a[3][i] > 0.0 && a[3][i] <= 1.0))
        return fail(SEDCOL_EDOMAIN, "interface porosity outside (0,1] at column " + std::to_string(c));
      if (!(a[0][i] > 0.0) || !std::isfinite(a[0][i]))
        return fail(SEDCOL_EDOMAIN, "sediment layer height <= 0 at column " + std::to_string(c) + " layer " + std::to_string(k));
      if (k > 0 && (!(a[1][i] > 0.0) || !std::isfinite(a[1][i])))
        return fail(SEDCOL_EDOMAIN, "sediment central layer difference <= 0 at column " + std::to_string(c) + " layer " + std::to_string(k));
      if (!(a[4][i] >= 0.0) || !std::isfinite(a[4][i]))
        return fail(SEDCOL_EDOMAIN, "bioturbation factor must be finite and >= 0 at column " + std::to_string(c));
    }
  }
  const size_t cb = sizeof(double) * (size_t)h->ncell;
  HIPOK(hipMemcpyAsync(h->d_dz, a[0].data(), cb, hipMemcpyHostToDevice, h->stream));
  HIPOK(hipMemcpyAsync(h->d_dzc, a[1].data(), cb, hipMemcpyHostToDevice, h->stream));
  HIPOK(hipMemcpyAsync(h->d_phi, a[2].data(), cb, hipMemcpyHostToDevice, h->stream));
  HIPOK(hipMemcpyAsync(h->d_phiI, a[3].data(), cb, hipMemcpyHostToDevice, h->stream));
  HIPOK(hipMemcpyAsync(h->d_bfac, a[4].data(), cb, hipMemcpyHostToDevice, h->stream));
  HIPOK(hipMemcpyAsync(h->d_biot, bioturbation, sizeof(double) * h->ncol, hipMemcpyHostToDevice, h->stream));
  HIPOK(hipStreamSynchronize(h->stream));
  h->have_geom = true;
  if (h->zw_on) return zw_geometry(h);
  return SEDCOL_OK;
}

int sedcol_set_forcing_series(sedcol_handle *h, int32_t nrec, double t0, double dt_rec, const double *temp,
                              const double *c_bw, const double *flux_top, int32_t nrec_flux)
{
  if (!h || !temp || !c_bw || !flux_top) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  if (nrec < 1) return fail(SEDCOL_EINVAL, "nrec must be positive");
  if (nrec_flux != 1 && nrec_flux != nrec) return fail(SEDCOL_EINVAL, "nrec_flux must be 1 or nrec");
  if (nrec > 1 && !(dt_rec > 0.0 && std::isfinite(dt_rec))) return fail(SEDCOL_EINVAL, "dt_rec must be positive");
  if (!std::isfinite(t0)) return fail(SEDCOL_EINVAL, "t0 must be finite");
  int rc = alloc_forcing(h, nrec);
  if (rc) return rc;
  HIPOK(hipMemcpyAsync(h->d_temp, temp, sizeof(double) * (size_t)nrec * h->ncol, hipMemcpyHostToDevice, h->stream));
  HIPOK(hipMemcpyAsync(h->d_cbw, c_bw, sizeof(double) * (size_t)nrec * NV * h->ncol, hipMemcpyHostToDevice, h->stream));
  HIPOK(hipMemcpyAsync(h->d_fup, flux_top, sizeof(double) * (size_t)nrec_flux * NV * h->ncol, hipMemcpyHostToDevice,
                       h->stream));
  HIPOK(hipStreamSynchronize(h->stream));
  h->nrec = nrec;
  h->fup_rec_stride = nrec_flux > 1 ? (long long)NV * h->ncol : 0;
  h->t0 = t0;
  h->dt_rec = nrec > 1 ? dt_rec : INFINITY;
  h->have_forcing = true;
  return SEDCOL_OK;
}

int sedcol_set_forcing(sedcol_handle *h, const double *temp, const double *c_bw, const double *flux_top)
{
  return sedcol_set_forcing_series(h, 1, 0.0, INFINITY, temp, c_bw, flux_top, 1);
}

int sedcol_set_flux_cap(sedcol_handle *h, const double *flux_cap)
{
  if (!h || !flux_cap) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  std::vector<double> cap;
  pack_cells(h, flux_cap, cap);
  for (int c = 0; c < h->ncol; ++c) {
    if (h->mask[c]) continue;
    for (int k = h->off[c]; k < h->off[c + 1]; ++k)
      if (!(cap[k] >= 0.0) || !std::isfinite(cap[k])) return fail(SEDCOL_EDOMAIN, "flux cap must be finite and >= 0");
  }
  if (!h->d_fcap) HIPOK(hipMalloc(&h->d_fcap, sizeof(double) * (size_t)h->ncell));
  HIPOK(hipMemcpyAsync(h->d_fcap, cap.data(), sizeof(double) * cap.size(), hipMemcpyHostToDevice, h->stream));
  HIPOK(hipStreamSynchronize(h->stream));
  return SEDCOL_OK;
}

int sedcol_assemble_forcing(sedcol_handle *h, const double *temp_surface, const double *c_surface,
                            const double *w_surface, const int32_t *present, double *flux_out)
{
  if (!h || !present) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  if (!h->have_params) return fail(SEDCOL_EINVAL, "sedcol_set_params has not been called");
  if (!h->have_geom) return fail(SEDCOL_EINVAL, "sedcol_set_geometry has not been called");
  if (h->nrec > 1) return fail(SEDCOL_EINVAL, "a forcing series is active; boundary assembly sets one record");
  unsigned bits = 0;
  bool need_w = false;
  for (int n = 0; n < NV; ++n)
    if (present[n]) {
      bits |= 1u << n;
      need_w |= sc::part(n);
    }
  if (bits && !c_surface) return fail(SEDCOL_EINVAL, "c_surface is needed for the present fields");
  if (need_w && !w_surface) return fail(SEDCOL_EINVAL, "w_surface is needed for particulate fields");
  const size_t nc = (size_t)h->ncol;
  if (!h->have_forcing) {
    int rc = alloc_forcing(h, 1);
    if (rc) return rc;
    HIPOK(hipMemsetAsync(h->d_temp, 0, sizeof(double) * nc, h->stream));
    HIPOK(hipMemsetAsync(h->d_cbw, 0, sizeof(double) * NV * nc, h->stream));
    HIPOK(hipMemsetAsync(h->d_fup, 0, sizeof(double) * NV * nc, h->stream));
    h->nrec = 1;
    h->fup_rec_stride = 0;
    h->t0 = 0.0;
    h->dt_rec = INFINITY;
    h->have_forcing = true;
  }
  int rc = ensure_scratch(h);  // inputs staged in the NV*ncell scratch (>= (2 NV + 1) ncol when K >= 3)
  if (rc) return rc;
  double *d_ts = nullptr, *d_cs = nullptr, *d_ws = nullptr;
  double *stage = h->d_scr;
  std::vector<double> tmp;
  const size_t need = (2 * NV + 1) * nc;
  double *own = nullptr;
  if ((size_t)NV * h->ncell < need) {
    HIPOK(hipMalloc(&own, sizeof(double) * need));
    stage = own;
  }
  hipError_t e = hipSuccess;
  if (temp_surface) {
    d_ts = stage;
    e = hipMemcpyAsync(d_ts, temp_surface, sizeof(double) * nc, hipMemcpyHostToDevice, h->stream);
  }
  if (e == hipSuccess && c_surface) {
    d_cs = stage + nc;
    e = hipMemcpyAsync(d_cs, c_surface, sizeof(double) * NV * nc, hipMemcpyHostToDevice, h->stream);
  }
  if (e == hipSuccess && w_surface) {
    d_ws = stage + (NV + 1) * nc;
    e = hipMemcpyAsync(d_ws, w_surface, sizeof(double) * NV * nc, hipMemcpyHostToDevice, h->stream);
  }
  const Dev d = h->dev();
  for (int b = 0; b < 3 && e == hipSuccess; ++b)
    if (!h->bins[b].empty())
      sc::launch_assemble(h->stream, d, h->d_list[b], (int)h->bins[b].size(), h->p, bits, d_ts, d_cs, d_ws, h->d_temp,
                          h->d_cbw, h->d_fup);
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess && flux_out)
    e = hipMemcpyAsync(flux_out, h->d_fup, sizeof(double) * NV * nc, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (own) (void)hipFree(own);
  if (e != hipSuccess) return fail(SEDCOL_EHIP, std::string("sedcol_assemble_forcing: ") + hipGetErrorString(e));
  return SEDCOL_OK;
}

int sedcol_set_state(sedcol_handle *h, const double *conc)
{
  if (!h || !conc) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  std::vector<double> buf;
  pack_state(h, conc, buf);
  HIPOK(hipMemcpyAsync(h->d_conc, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice, h->stream));
  HIPOK(hipStreamSynchronize(h->stream));
  return SEDCOL_OK;
}

int sedcol_set_state_profiles(sedcol_handle *h, const float *prof, int64_t ld, int32_t nz, const double *scale,
                              const double *offset, const double *eps, double dz1, const double *depth)
{
  if (!h) return fail(SEDCOL_EINVAL, "null handle");
  const DeviceScope dscope(h->device);
  if (!h->have_geom) return fail(SEDCOL_EINVAL, "sedcol_set_geometry has not been called");
  if (!prof || !scale || !offset || !eps) return fail(SEDCOL_EINVAL, "null argument");
  if (nz < 2 || nz > sc::kWarmMaxNodes) return fail(SEDCOL_EINVAL, "nz must be 2..64");
  if (ld < (int64_t)NV * nz) return fail(SEDCOL_EINVAL, "ld must be at least nvar*nz");
  if (!(dz1 > 0.0) || !std::isfinite(dz1)) return fail(SEDCOL_EINVAL, "dz1 must be positive and finite");
  hipPointerAttribute_t at{};
  if (hipPointerGetAttributes(&at, prof) != hipSuccess || at.type != hipMemoryTypeDevice || at.device != h->device) {
    (void)hipGetLastError();
    return fail(SEDCOL_EINVAL, "prof must be device memory of the handle's device");
  }
  if (depth)
    for (int c = 0; c < h->ncol; ++c)
      if (!h->mask[c] && !(depth[c] > 0.0 && std::isfinite(depth[c])))
        return fail(SEDCOL_EINVAL, "depth must be positive and finite for every wet column");
  if (!h->d_warm) HIPOK(hipMalloc(&h->d_warm, sizeof(double) * (2 * NV * sc::kWarmMaxNodes + NV)));
  const size_t nv = (size_t)NV * nz;
  HIPOK(hipMemcpyAsync(h->d_warm, scale, sizeof(double) * nv, hipMemcpyHostToDevice, h->stream));
  HIPOK(hipMemcpyAsync(h->d_warm + nv, offset, sizeof(double) * nv, hipMemcpyHostToDevice, h->stream));
  HIPOK(hipMemcpyAsync(h->d_warm + 2 * nv, eps, sizeof(double) * NV, hipMemcpyHostToDevice, h->stream));
  if (depth) HIPOK(hipMemcpyAsync(h->d_small, depth, sizeof(double) * h->ncol, hipMemcpyHostToDevice, h->stream));
  sc::WarmProfiles w{prof, (long long)ld, nz, dz1, h->d_warm, h->d_warm + nv, h->d_warm + 2 * nv,
                     depth ? h->d_small : nullptr};
  const Dev d = h->dev();
  for (int b = 0; b < 3; ++b)
    if (!h->bins[b].empty()) sc::launch_warm_profiles(kLpc[b], h->stream, d, h->d_list[b], (int)h->bins[b].size(), w);
  HIPOK(hipGetLastError());
  HIPOK(hipStreamSynchronize(h->stream));
  return SEDCOL_OK;
}

int sedcol_get_state(sedcol_handle *h, double *conc)
{
  if (!h || !conc) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  std::vector<double> buf((size_t)NV * h->ncell);
  HIPOK(hipMemcpyAsync(buf.data(), h->d_conc, sizeof(double) * buf.size(), hipMemcpyDeviceToHost, h->stream));
  HIPOK(hipStreamSynchronize(h->stream));
  unpack_state(h, buf, conc);
  return SEDCOL_OK;
}


int sedcol_rhs(sedcol_handle *h, const double *conc, double *rhs, double *flux_top_out)
{
  int rc = check_handle(h, true);
  if (rc) return rc;
  const DeviceScope dscope(h->device);
  if (!conc || !rhs) return fail(SEDCOL_EINVAL, "null argument");
  if (h->nrec > 1) return fail(SEDCOL_EINVAL, "a forcing series is active; get_rhs uses one record (sedcol_set_forcing)");
  if ((rc = ensure_scratch(h))) return rc;
  std::vector<double> buf;
  pack_state(h, conc, buf);
  HIPOK(hipMemcpyAsync(h->d_scr, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice, h->stream));
  HIPOK(hipMemsetAsync(h->d_scr2, 0, sizeof(double) * buf.size(), h->stream));
  HIPOK(hipMemsetAsync(h->d_small, 0, sizeof(double) * NV * h->ncol, h->stream));
  const Dev d = h->dev();
  for (int b = 0; b < 3; ++b)
    if (!h->bins[b].empty())
      sc::launch_rhs(kLpc[b], h->stream, d, h->d_list[b], (int)h->bins[b].size(), h->p, 0, h->d_scr, h->d_scr2,
                     h->d_small);
  HIPOK(hipGetLastError());
  HIPOK(hipMemcpyAsync(buf.data(), h->d_scr2, sizeof(double) * buf.size(), hipMemcpyDeviceToHost, h->stream));
  std::vector<double> fl((size_t)NV * h->ncol);
  HIPOK(hipMemcpyAsync(fl.data(), h->d_small, sizeof(double) * fl.size(), hipMemcpyDeviceToHost, h->stream));
  HIPOK(hipStreamSynchronize(h->stream));
  unpack_state(h, buf, rhs);
  if (flux_top_out) std::memcpy(flux_top_out, fl.data(), sizeof(double) * fl.size());
  return SEDCOL_OK;
}

int sedcol_rhs_noalloc(sedcol_handle *h, const double *conc, double *rhs, double *flux_top_out)
{
  int rc = check_handle(h, true);
  if (rc) return rc;
  const DeviceScope dscope(h->device);
  if (!conc || !rhs) return fail(SEDCOL_EINVAL, "null argument");
  if (h->nrec > 1) return fail(SEDCOL_EINVAL, "a forcing series is active; get_rhs uses one record (sedcol_set_forcing)");
  const bool fortran = h->layout == SEDCOL_LAYOUT_FORTRAN;
  const size_t cells = (size_t)NV * (size_t)h->ncell, fl = (size_t)NV * (size_t)h->ncol;
  const size_t need = sizeof(double) * ((fortran ? 3 : 2) * cells + fl);
  if (need > kNoallocBytes)
    return fail(SEDCOL_EINVAL, "sedcol_rhs_noalloc: the domain needs " + std::to_string(need) +
                                   " bytes of workspace, more than its fixed " + std::to_string(kNoallocBytes) +
                                   " (use sedcol_rhs)");
  if (!h->d_noalloc) return fail(SEDCOL_EHIP, "sedcol_rhs_noalloc: no workspace address on this device");
  std::lock_guard<std::mutex> lk(g_noalloc_mu[h->device]);
  double *in = h->d_noalloc, *out = in + cells, *flux = out + cells, *raw = flux + fl;
  HIPOK(hipMemcpyAsync(fortran ? raw : in, conc, sizeof(double) * cells, hipMemcpyHostToDevice, h->stream));
  if (fortran) sc::launch_layout(h->stream, raw, in, h->ncol, h->kmax, true);
  // land columns keep rhs = 0 and no top flux; fill kernels, not memsets
  sc::launch_fill_dbl(h->stream, out, (long long)cells, 0.0);
  sc::launch_fill_dbl(h->stream, flux, (long long)fl, 0.0);
  const Dev d = h->dev();
  for (int b = 0; b < 3; ++b)
    if (!h->bins[b].empty())
      sc::launch_rhs(kLpc[b], h->stream, d, h->d_list[b], (int)h->bins[b].size(), h->p, 0, in, out, flux);
  if (fortran) sc::launch_layout(h->stream, out, raw, h->ncol, h->kmax, false);
  HIPOK(hipGetLastError());
  HIPOK(hipMemcpyAsync(rhs, fortran ? raw : out, sizeof(double) * cells, hipMemcpyDeviceToHost, h->stream));
  if (flux_top_out)
    HIPOK(hipMemcpyAsync(flux_top_out, flux, sizeof(double) * fl, hipMemcpyDeviceToHost, h->stream));
  HIPOK(hipStreamSynchronize(h->stream));
  return SEDCOL_OK;
}

int sedcol_rates(sedcol_handle *h, double *dic_release, double *n2_release)
{
  int rc = check_handle(h, true);
  if (rc) return rc;
  const DeviceScope dscope(h->device);
  if (!dic_release || !n2_release) return fail(SEDCOL_EINVAL, "null argument");
  if (h->nrec > 1) return fail(SEDCOL_EINVAL, "a forcing series is active; the rates use one record (sedcol_set_forcing)");
  if ((rc = ensure_scratch(h))) return rc;
  const size_t nc = (size_t)h->ncell;
  HIPOK(hipMemsetAsync(h->d_scr, 0, sizeof(double) * 2 * nc, h->stream));
  const Dev d = h->dev();
  for (int b = 0; b < 3; ++b)
    if (!h->bins[b].empty())
      sc::launch_rhs(kLpc[b], h->stream, d, h->d_list[b], (int)h->bins[b].size(), h->p, 0, h->d_conc, nullptr,
                     nullptr, h->d_scr);
  HIPOK(hipGetLastError());
  std::vector<double> buf(2 * nc);
  HIPOK(hipMemcpyAsync(buf.data(), h->d_scr, sizeof(double) * buf.size(), hipMemcpyDeviceToHost, h->stream));
  HIPOK(hipStreamSynchronize(h->stream));
  if (h->layout == SEDCOL_LAYOUT_PACKED) {
    std::memcpy(dic_release, buf.data(), sizeof(double) * nc);
    std::memcpy(n2_release, buf.data() + nc, sizeof(double) * nc);
  } else {
    for (int c = 0; c < h->ncol; ++c)
      for (int k = 0; k < h->kmax; ++k) {
        dic_release[fidx(h, c, k)] = buf[(size_t)c * h->kmax + k];
        n2_release[fidx(h, c, k)] = buf[nc + (size_t)c * h->kmax + k];
      }
  }
  return SEDCOL_OK;
}

// ADAPTIVE_EULER over dt on the resident state (solver_library.F90:104-140)
static int adaptive_euler(sedcol_handle *h, double dt, int rec, int *launches)
{
  int rc = ensure_scratch(h);
  if (rc) return rc;
  const Dev d = h->dev();
  double dt_int = 0.0, dt_red = dt;
  bool need_rhs = true;
  while (dt_int < dt) {
    if (need_rhs) {
      for (int b = 0; b < 3; ++b)
        if (!h->bins[b].empty()) {
          sc::launch_rhs_store(kLpc[b], h->stream, d, h->d_list[b], (int)h->bins[b].size(), h->p, rec, h->d_scr,
                               h->d_scr2);
          ++*launches;
        }
      need_rhs = false;
    }
    HIPOK(hipMemsetAsync(h->d_flag, 0, sizeof(int), h->stream));
    for (int b = 0; b < 3; ++b)
      if (!h->bins[b].empty()) {
        sc::launch_ae_check(kLpc[b], h->stream, d, h->d_list[b], (int)h->bins[b].size(), dt_red,
                            h->p.relative_change_min, h->d_scr, h->d_flag);
        ++*launches;
      }
    HIPOK(hipGetLastError());
    int viol = 0;
    if ((rc = read_flag(h, &viol))) return rc;
    if (viol && dt_red > h->p.dt_min) {
      dt_red = dt_red * 0.25;  // the state is unchanged: the right-hand side is reused
    } else {
      if (dt_red < h->last_min_dt) h->last_min_dt = dt_red;
      for (int b = 0; b < 3; ++b)
        if (!h->bins[b].empty()) {
          sc::launch_ae_commit(kLpc[b], h->stream, d, h->d_list[b], (int)h->bins[b].size(), dt_red, h->d_scr,
                               h->d_scr2);
          ++*launches;
        }
      dt_int = dt_int + dt_red;
      need_rhs = true;
    }
  }
  HIPOK(hipGetLastError());
  return SEDCOL_OK;
}

static int run_fixed(sedcol_handle *h, double dt, int nsteps, int method, int steps_per_launch, int *launches)
{
  const Dev d = h->dev();
  const int rec = 0;
  int left = nsteps;
  const int chunk = steps_per_launch > 0 ? steps_per_launch : nsteps;
  while (left > 0) {
    const int ns = std::min(left, chunk);
    for (int b = 0; b < 3; ++b)
      if (!h->bins[b].empty()) {
        sc::launch_fixed(kLpc[b], method, h->stream, d, h->d_list[b], (int)h->bins[b].size(), h->p, rec, dt, ns,
                         h->d_flag);
        ++*launches;
      }
    HIPOK(hipGetLastError());
    left -= ns;
  }
  return SEDCOL_OK;
}

int sedcol_run(sedcol_handle *h, double dt, int32_t nsteps, int32_t method, int32_t steps_per_launch)
{
  int rc = check_handle(h, true);
  if (rc) return rc;
  const DeviceScope dscope(h->device);
  if (nsteps < 0) return fail(SEDCOL_EINVAL, "nsteps must be >= 0");
  if (!(dt > 0.0) || !std::isfinite(dt)) return fail(SEDCOL_EINVAL, "dt must be positive");
  if (method < SEDCOL_EULER || method > SEDCOL_EMP2)
    return fail(SEDCOL_EINVAL, "sedcol_run takes methods 0..5 (use sedcol_integrate for method 6)");
  if (h->nrec > 1) return fail(SEDCOL_EINVAL, "fixed-step runs use a single forcing record (sedcol_set_forcing)");
  HIPOK(hipMemsetAsync(h->d_flag, 0, sizeof(int), h->stream));
  int launches = 0;
  HIPOK(hipEventRecord(h->ev0, h->stream));
  int one_bin = -1, nbins = 0;
  for (int b = 0; b < 3; ++b)
    if (!h->bins[b].empty()) {
      ++nbins;
      one_bin = b;
    }
  if (method == SEDCOL_ADAPTIVE_EULER && nbins == 1 &&
      (int)h->bins[one_bin].size() <= sc::kAeFusedBlock / kLpc[one_bin]) {
    // every wet column in one workgroup: the steps run without a host round
    // trip per attempt, in launches of kAeFusedSteps steps
    h->h_devview = h->dev();
    h->h_params = h->p;
    HIPOK(hipMemcpyAsync(h->d_devview, &h->h_devview, sizeof(Dev), hipMemcpyHostToDevice, h->stream));
    HIPOK(hipMemcpyAsync(h->d_params, &h->h_params, sizeof(sedcol_params), hipMemcpyHostToDevice, h->stream));
    for (int s0 = 0; s0 < nsteps; s0 += kAeFusedSteps) {
      const int ns = std::min(kAeFusedSteps, nsteps - s0);
      sc::launch_ae_fused(kLpc[one_bin], h->stream, h->d_devview, h->d_list[one_bin], (int)h->bins[one_bin].size(),
                          h->p, h->d_params, 0, dt, ns, h->d_flag, h->d_small);
      ++launches;
      HIPOK(hipGetLastError());
      double mind = INFINITY;
      HIPOK(hipMemcpyAsync(&mind, h->d_small, sizeof(double), hipMemcpyDeviceToHost, h->stream));
      HIPOK(hipStreamSynchronize(h->stream));
      if (mind < h->last_min_dt) h->last_min_dt = mind;
      int flag = 0;
      if ((rc = read_flag(h, &flag))) return rc;
      if (flag) break;
    }
  } else if (method == SEDCOL_ADAPTIVE_EULER) {
    for (int s = 0; s < nsteps; ++s) {
      if ((rc = adaptive_euler(h, dt, 0, &launches))) return rc;
      if (h->p.clamp) {
        const Dev d = h->dev();
        for (int b = 0; b < 3; ++b)
          if (!h->bins[b].empty()) {
            sc::launch_clamp(kLpc[b], h->stream, d, h->d_list[b], (int)h->bins[b].size(), h->d_flag);
            ++launches;
          }
      }
    }
  } else {
    if ((rc = run_fixed(h, dt, nsteps, method, steps_per_launch, &launches))) return rc;
  }
  HIPOK(hipEventRecord(h->ev1, h->stream));
  HIPOK(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIPOK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms = ms;
  h->last_launches = launches;
  int flag = 0;
  if ((rc = read_flag(h, &flag))) return rc;
  return flag_status(flag);
}

int sedcol_integrate(sedcol_handle *h, double t_start, double t_end)
{
  int rc = check_handle(h, true);
  if (rc) return rc;
  const DeviceScope dscope(h->device);
  if (!std::isfinite(t_start) || !std::isfinite(t_end) || !(t_end >= t_start))
    return fail(SEDCOL_EINVAL, "need finite t_start <= t_end");
  int launches = 0;
  HIPOK(hipEventRecord(h->ev0, h->stream));
  if ((rc = run_adaptive(h, t_start, t_end, -1, &launches))) return rc;
  HIPOK(hipEventRecord(h->ev1, h->stream));
  HIPOK(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIPOK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms = ms;
  h->last_launches = launches;
  int flag = 0;
  if ((rc = read_flag(h, &flag))) return rc;
  if ((rc = flag_status(flag))) return rc;
  // Longest-first order for the next launch: the columns that took the most
  // attempts in this one start first, so the launch does not end with a few
  // long columns running on an otherwise idle device.  Results do not depend
  // on the order (columns are independent).
  h->h_work.resize(h->ncol);
  HIPOK(hipMemcpyAsync(h->h_work.data(), h->d_work, sizeof(int) * h->ncol, hipMemcpyDeviceToHost, h->stream));
  HIPOK(hipStreamSynchronize(h->stream));
  for (int b = 0; b < 3; ++b) {
    auto &v = h->bins[b];
    if (v.size() < 2) continue;
    std::stable_sort(v.begin(), v.end(), [&](int a, int c) { return h->h_work[a] > h->h_work[c]; });
    HIPOK(hipMemcpyAsync(h->d_list[b], v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice, h->stream));
  }
  HIPOK(hipStreamSynchronize(h->stream));
  return SEDCOL_OK;
}

int sedcol_set_schedule(sedcol_handle *h, const sedcol_schedule *s)
{
  if (!h || !s) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  if (s->window < 0) return fail(SEDCOL_EINVAL, "window must be >= 0");
  if (s->graph_windows < 0 || s->graph_windows > 4096) return fail(SEDCOL_EINVAL, "graph_windows outside 0..4096");
  if (s->check_every < 1) return fail(SEDCOL_EINVAL, "check_every must be >= 1");
  if (s->concurrent_bins != 0 && s->concurrent_bins != 1) return fail(SEDCOL_EINVAL, "concurrent_bins must be 0 or 1");
  h->sched = *s;
  drop_graph(h);
  return SEDCOL_OK;
}

int sedcol_get_schedule(sedcol_handle *h, sedcol_schedule *s)
{
  if (!h || !s) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  *s = h->sched;
  return SEDCOL_OK;
}

int sedcol_set_tuning(sedcol_handle *h, const sedcol_tuning *t)
{
  if (!h || !t) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  int rc = check_tuning(h, t);
  if (rc) return rc;
  for (int b = 0; b < SEDCOL_NBIN; ++b) {
    h->lpc[b] = t->lanes_per_column[b];
    h->blk[b] = t->lanes_per_column[b] * t->columns_per_group[b];
    h->lpl[b] = t->layers_per_lane[b] == 2 ? 2 : 1;
  }
  drop_graph(h);
  return SEDCOL_OK;
}

int sedcol_get_tuning(sedcol_handle *h, sedcol_tuning *t)
{
  if (!h || !t) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  for (int b = 0; b < SEDCOL_NBIN; ++b) {
    t->lanes_per_column[b] = h->lpc[b];
    t->columns_per_group[b] = h->blk[b] / h->lpc[b];
    t->layers_per_lane[b] = h->lpl[b];
  }
  return SEDCOL_OK;
}

int sedcol_autotune(sedcol_handle *h, double t_start, double t_span, int32_t repeats, sedcol_tuning *best,
                    double *ms_table)
{
  int rc = check_handle(h, true);
  if (rc) return rc;
  const DeviceScope dscope(h->device);
  if (!best) return fail(SEDCOL_EINVAL, "null argument");
  if (!std::isfinite(t_start) || !(t_span > 0.0) || !std::isfinite(t_span))
    return fail(SEDCOL_EINVAL, "need finite t_start and t_span > 0");
  if (repeats < 1 || repeats > 100) return fail(SEDCOL_EINVAL, "repeats outside 1..100");
  static const int kL[3] = {16, 32, 64}, kW[3] = {64, 128, 256};
  if (ms_table)
    for (int i = 0; i < SEDCOL_NBIN * 18; ++i) ms_table[i] = -1.0;
  // snapshot of everything an integration changes; restored before every
  // trial and at the end, so tuning leaves the handle's state as it was
  const size_t nc = (size_t)h->ncol, ncell = (size_t)h->ncell;
  struct Snap { void *dst; const void *src; size_t bytes; };
  std::vector<Snap> snaps = {
      {nullptr, h->d_conc, sizeof(double) * NV * ncell}, {nullptr, h->d_cellacc, sizeof(double) * 10 * ncell},
      {nullptr, h->d_colacc, sizeof(double) * 2 * NV * nc}, {nullptr, h->d_flux, sizeof(double) * NV * nc},
      {nullptr, h->d_diag, sizeof(sedcol_coldiag) * nc}, {nullptr, h->d_work, sizeof(int) * nc}};
  auto release = [&]() {
    for (auto &q : snaps)
      if (q.dst) (void)hipFree(q.dst);
  };
  for (auto &q : snaps) {
    if (hipMalloc(&q.dst, q.bytes) != hipSuccess) {
      release();
      return fail(SEDCOL_EHIP, "sedcol_autotune: out of device memory for the state snapshot");
    }
    if (hipMemcpyAsync(q.dst, q.src, q.bytes, hipMemcpyDeviceToDevice, h->stream) != hipSuccess) {
      release();
      return fail(SEDCOL_EHIP, "sedcol_autotune: snapshot copy failed");
    }
  }
  auto restore = [&]() -> int {
    for (auto &q : snaps) HIPOK(hipMemcpyAsync(const_cast<void *>(q.src), q.dst, q.bytes, hipMemcpyDeviceToDevice,
                                               h->stream));
    return SEDCOL_OK;
  };
  const sedcol_schedule saved_sched = h->sched;
  int saved_lpc[3], saved_blk[3], saved_lpl[3];
  std::memcpy(saved_lpc, h->lpc, sizeof saved_lpc);
  std::memcpy(saved_blk, h->blk, sizeof saved_blk);
  std::memcpy(saved_lpl, h->lpl, sizeof saved_lpl);
  h->sched = sedcol_schedule{0, 0, 1, 0};  // one launch per trial
  int best_l[3], best_w[3], best_p[3];
  std::memcpy(best_l, h->lpc, sizeof best_l);
  std::memcpy(best_w, h->blk, sizeof best_w);
  std::memcpy(best_p, h->lpl, sizeof best_p);
  for (int b = 0; b < 3 && rc == SEDCOL_OK; ++b) {
    if (h->bins[b].empty()) continue;
    double best_ms = INFINITY;
    for (int ip = 0; ip < 2 && rc == SEDCOL_OK; ++ip) {
      const int lpl = ip + 1;
      if (lpl == 2 && !sc::emp2_supported(h->p)) continue;
      const int lmin = bin_min_lpc(h, b, lpl);
      for (int il = 0; il < 3 && rc == SEDCOL_OK; ++il) {
        if (kL[il] < lmin) continue;
        for (int iw = 0; iw < 3 && rc == SEDCOL_OK; ++iw) {
          if (kW[iw] < kL[il]) continue;
          h->lpc[b] = kL[il];
          h->blk[b] = kW[iw];
          h->lpl[b] = lpl;
          double tmin = INFINITY;
          for (int r = 0; r < repeats && rc == SEDCOL_OK; ++r) {
            int launches = 0, flag = 0;
            if ((rc = restore())) break;
            if (hipEventRecord(h->ev0, h->stream) != hipSuccess) { rc = fail(SEDCOL_EHIP, "event record"); break; }
            if ((rc = run_adaptive(h, t_start, t_start + t_span, b, &launches))) break;
            if (hipEventRecord(h->ev1, h->stream) != hipSuccess || hipEventSynchronize(h->ev1) != hipSuccess) {
              rc = fail(SEDCOL_EHIP, "event synchronize");
              break;
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, h->ev0, h->ev1);
            if ((rc = read_flag(h, &flag))) break;
            if ((rc = flag_status(flag))) break;
            tmin = std::min(tmin, (double)ms);
          }
          if (rc) break;
          if (ms_table) ms_table[b * 18 + ip * 9 + il * 3 + iw] = tmin;
          if (tmin < best_ms) {
            best_ms = tmin;
            best_l[b] = kL[il];
            best_w[b] = kW[iw];
            best_p[b] = lpl;
          }
        }
      }
    }
  }
  int rc2 = restore();
  if (rc2 == SEDCOL_OK && hipStreamSynchronize(h->stream) != hipSuccess) rc2 = fail(SEDCOL_EHIP, "synchronize");
  release();
  h->sched = saved_sched;
  if (rc || rc2) {
    std::memcpy(h->lpc, saved_lpc, sizeof saved_lpc);
    std::memcpy(h->blk, saved_blk, sizeof saved_blk);
    std::memcpy(h->lpl, saved_lpl, sizeof saved_lpl);
    return rc ? rc : rc2;
  }
  std::memcpy(h->lpc, best_l, sizeof best_l);
  std::memcpy(h->blk, best_w, sizeof best_w);
  std::memcpy(h->lpl, best_p, sizeof best_p);
  drop_graph(h);
  return sedcol_get_tuning(h, best);
}

int sedcol_step(sedcol_handle *h, double dt, int32_t method, double *conc_inout, double *flux_top_out)
{
  int rc = check_handle(h, true);
  if (rc) return rc;
  const DeviceScope dscope(h->device);
  if (!conc_inout) return fail(SEDCOL_EINVAL, "null argument");
  if (method < SEDCOL_EULER || method > SEDCOL_ROS2_ADAPTIVE) return fail(SEDCOL_EINVAL, "unknown method");
  if ((rc = sedcol_set_state(h, conc_inout))) return rc;
  if (method == SEDCOL_EMP_ADAPTIVE || method == SEDCOL_ROS2_ADAPTIVE) {
    if (h->nrec > 1) return fail(SEDCOL_EINVAL, "sedcol_step uses a single forcing record");
    // the adaptive stepper of this call; the handle's choice is restored
    const int32_t keep = h->p.adaptive_method;
    const int32_t want = method == SEDCOL_ROS2_ADAPTIVE ? SEDCOL_ROS2_ADAPTIVE : 0;
    const bool swap = (keep == SEDCOL_ROS2_ADAPTIVE) != (want == SEDCOL_ROS2_ADAPTIVE);
    if (swap) {
      h->p.adaptive_method = want;
      drop_graph(h);
    }
    rc = sedcol_integrate(h, 0.0, dt);
    if (swap) {
      h->p.adaptive_method = keep;
      drop_graph(h);
    }
  } else {
    rc = sedcol_run(h, dt, 1, method, 0);
  }
  int rc2 = sedcol_get_state(h, conc_inout);
  if (rc) return rc;
  if (rc2) return rc2;
  if (flux_top_out) return sedcol_get_fluxes(h, flux_top_out);
  return SEDCOL_OK;
}

int sedcol_spinup(sedcol_handle *h, double t_start, double chunk, double t_max, double tol, const double *cfloor,
                  const double *flux_ref, double flux_rtol, const double *flux_floor, double *t_conv, double *resid,
                  int32_t *n_unconverged)
{
  int rc = check_handle(h, true);
  if (rc) return rc;
  const DeviceScope dscope(h->device);
  if (!cfloor) return fail(SEDCOL_EINVAL, "null cfloor");
  if (h->nrec > 1)
    return fail(SEDCOL_EINVAL, "a forcing series is active; spin-up runs under one record (sedcol_set_forcing)");
  if (!std::isfinite(t_start) || !(chunk > 0.0) || !std::isfinite(chunk) || !(t_max >= 0.0) ||
      !std::isfinite(t_max) || !(tol > 0.0))
    return fail(SEDCOL_EINVAL, "need finite t_start, chunk > 0, t_max >= 0, tol > 0");
  if (flux_ref && (!flux_floor || !(flux_rtol >= 0.0)))
    return fail(SEDCOL_EINVAL, "a flux criterion needs flux_floor and flux_rtol >= 0");
  const int nc = h->ncol;
  std::vector<double> tc(nc, -1.0), res(nc, 0.0);
  std::vector<double> fl(flux_ref ? (size_t)NV * nc : 0);
  auto converged = [&](int c) {
    if (!(res[c] < tol)) return false;
    if (!flux_ref) return true;
    for (int n = 0; n < NV; ++n) {
      const double fr = flux_ref[(size_t)n * nc + c];
      const double sc = std::max(std::fabs(fr), flux_floor[c]);
      if (!(std::fabs(fl[(size_t)n * nc + c] - fr) <= flux_rtol * sc)) return false;
    }
    return true;
  };
  std::vector<int> full[3];
  for (int b = 0; b < 3; ++b) full[b] = h->bins[b];
  std::vector<char> active(nc, 0);
  for (int b = 0; b < 3; ++b)
    for (int c : full[b]) active[c] = 1;
  double t = t_start;
  const double t_end = t_start + t_max;
  for (;;) {
    if ((rc = residual_flux(h, cfloor, res.data(), flux_ref ? fl.data() : nullptr))) break;
    int left = 0;
    for (int b = 0; b < 3; ++b) {
      std::vector<int> keep;
      keep.reserve(h->bins[b].size());
      for (int c : h->bins[b]) {
        if (converged(c)) {
          tc[c] = t - t_start;
          active[c] = 0;
        } else {
          keep.push_back(c);
        }
      }
      left += (int)keep.size();
      h->bins[b].swap(keep);
    }
    if (left == 0 || !(t < t_end)) break;
    if ((rc = upload_bins(h))) break;
    const double t1 = std::min(t + chunk, t_end);
    if ((rc = sedcol_integrate(h, t, t1))) break;
    t = t1;
  }
  // every wet column back in the work lists, final residual of all
  for (int b = 0; b < 3; ++b) h->bins[b] = full[b];
  int rc2 = upload_bins(h);
  if (!rc && !rc2) rc2 = residual_flux(h, cfloor, res.data(), nullptr);
  if (rc) return rc;
  if (rc2) return rc2;
  int32_t un = 0;
  for (int c = 0; c < nc; ++c) un += active[c] ? 1 : 0;
  if (t_conv) std::memcpy(t_conv, tc.data(), sizeof(double) * nc);
  if (resid) std::memcpy(resid, res.data(), sizeof(double) * nc);
  if (n_unconverged) *n_unconverged = un;
  return SEDCOL_OK;
}

int sedcol_checkpoint_size(sedcol_handle *h, int64_t *bytes)
{
  if (!h || !bytes) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  *bytes = ck_bytes(h);
  return SEDCOL_OK;
}

int sedcol_save_checkpoint(sedcol_handle *h, void *buf, int64_t bytes)
{
  if (!h || !buf) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  if (bytes < ck_bytes(h)) return fail(SEDCOL_EINVAL, "checkpoint buffer too small (sedcol_checkpoint_size)");
  CkHeader hd;
  std::memcpy(hd.magic, kCkMagic, 8);
  hd.ncol = h->ncol;
  hd.ncell = h->ncell;
  hd.nvar = NV;
  hd.version = 1;
  std::memcpy(buf, &hd, sizeof hd);
  char *p = static_cast<char *>(buf) + sizeof hd;
  for (auto &q : ck_parts(h)) {
    HIPOK(hipMemcpyAsync(p, q.first, q.second, hipMemcpyDeviceToHost, h->stream));
    p += q.second;
  }
  HIPOK(hipStreamSynchronize(h->stream));
  return SEDCOL_OK;
}

int sedcol_load_checkpoint(sedcol_handle *h, const void *buf, int64_t bytes)
{
  if (!h || !buf) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  if (bytes < (int64_t)sizeof(CkHeader)) return fail(SEDCOL_EINVAL, "checkpoint too short");
  CkHeader hd;
  std::memcpy(&hd, buf, sizeof hd);
  if (std::memcmp(hd.magic, kCkMagic, 8) != 0 || hd.version != 1)
    return fail(SEDCOL_EINVAL, "not a sedcol checkpoint");
  if (hd.ncol != h->ncol || hd.ncell != h->ncell || hd.nvar != NV)
    return fail(SEDCOL_EINVAL, "checkpoint of a different domain");
  if (bytes < ck_bytes(h)) return fail(SEDCOL_EINVAL, "checkpoint too short");
  const char *p = static_cast<const char *>(buf) + sizeof hd;
  for (auto &q : ck_parts(h)) {
    HIPOK(hipMemcpyAsync(q.first, p, q.second, hipMemcpyHostToDevice, h->stream));
    p += q.second;
  }
  HIPOK(hipStreamSynchronize(h->stream));
  return SEDCOL_OK;
}

int sedcol_get_fluxes(sedcol_handle *h, double *flux_top_out)
{
  if (!h || !flux_top_out) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  HIPOK(hipMemcpyAsync(flux_top_out, h->d_flux, sizeof(double) * NV * h->ncol, hipMemcpyDeviceToHost, h->stream));
  HIPOK(hipStreamSynchronize(h->stream));
  return SEDCOL_OK;
}

int sedcol_get_budget(sedcol_handle *h, double *budget)
{
  if (!h || !budget) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  double *d_b = nullptr;
  HIPOK(hipMalloc(&d_b, sizeof(double) * SEDCOL_NBUD * h->ncol));
  hipError_t e = hipMemsetAsync(d_b, 0, sizeof(double) * SEDCOL_NBUD * h->ncol, h->stream);
  const Dev d = h->dev();
  for (int b = 0; b < 3 && e == hipSuccess; ++b)
    if (!h->bins[b].empty()) sc::launch_budget(kLpc[b], h->stream, d, h->d_list[b], (int)h->bins[b].size(), d_b);
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(budget, d_b, sizeof(double) * SEDCOL_NBUD * h->ncol, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(d_b);
  if (e != hipSuccess) return fail(SEDCOL_EHIP, std::string("sedcol_get_budget: ") + hipGetErrorString(e));
  return SEDCOL_OK;
}

int sedcol_reset_budget(sedcol_handle *h)
{
  if (!h) return fail(SEDCOL_EINVAL, "null handle");
  const DeviceScope dscope(h->device);
  HIPOK(hipMemsetAsync(h->d_cellacc, 0, sizeof(double) * 10 * (size_t)h->ncell, h->stream));
  HIPOK(hipMemsetAsync(h->d_colacc, 0, sizeof(double) * 2 * NV * h->ncol, h->stream));
  HIPOK(hipStreamSynchronize(h->stream));
  return SEDCOL_OK;
}

int sedcol_get_inventory(sedcol_handle *h, double *inventory)
{
  if (!h || !inventory) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  HIPOK(hipMemsetAsync(h->d_small, 0, sizeof(double) * NV * h->ncol, h->stream));
  const Dev d = h->dev();
  for (int b = 0; b < 3; ++b)
    if (!h->bins[b].empty())
      sc::launch_inventory(kLpc[b], h->stream, d, h->d_list[b], (int)h->bins[b].size(), h->d_small);
  HIPOK(hipGetLastError());
  HIPOK(hipMemcpyAsync(inventory, h->d_small, sizeof(double) * NV * h->ncol, hipMemcpyDeviceToHost, h->stream));
  HIPOK(hipStreamSynchronize(h->stream));
  return SEDCOL_OK;
}

int sedcol_get_diagnostics(sedcol_handle *h, sedcol_coldiag *diag)
{
  if (!h || !diag) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  HIPOK(hipMemcpyAsync(diag, h->d_diag, sizeof(sedcol_coldiag) * h->ncol, hipMemcpyDeviceToHost, h->stream));
  HIPOK(hipStreamSynchronize(h->stream));
  return SEDCOL_OK;
}

int sedcol_reset_diagnostics(sedcol_handle *h)
{
  if (!h) return fail(SEDCOL_EINVAL, "null handle");
  const DeviceScope dscope(h->device);
  h->last_min_dt = 1.e20;
  return reset_diag(h);
}

int sedcol_residual(sedcol_handle *h, const double *cfloor, double *resid)
{
  int rc = check_handle(h, true);
  if (rc) return rc;
  const DeviceScope dscope(h->device);
  if (!cfloor || !resid) return fail(SEDCOL_EINVAL, "null argument");
  if (h->nrec > 1) return fail(SEDCOL_EINVAL, "a forcing series is active; the residual uses one record (sedcol_set_forcing)");
  return residual_flux(h, cfloor, resid, nullptr);
}

int sedcol_get_last_min_dt(sedcol_handle *h, double *last_min_dt)
{
  if (!h || !last_min_dt) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  *last_min_dt = h->last_min_dt;
  return SEDCOL_OK;
}

int sedcol_synchronize(sedcol_handle *h)
{
  if (!h) return fail(SEDCOL_EINVAL, "null handle");
  const DeviceScope dscope(h->device);
  HIPOK(hipStreamSynchronize(h->stream));
  return SEDCOL_OK;
}

int sedcol_last_timing(sedcol_handle *h, double *ms, int32_t *launches)
{
  if (!h) return fail(SEDCOL_EINVAL, "null handle");
  const DeviceScope dscope(h->device);
  if (ms) *ms = h->last_ms;
  if (launches) *launches = h->last_launches;
  return SEDCOL_OK;
}

}  // extern "C"
