// api.hip -- the dfargo context and the C ABI of include/dfargo.h.
//
// One context holds one disc resident in HBM: the state fields in the
// reference's ring-major layout, the work grids the sub-step needs, the 1-D
// tables and a device control block.  A sub-step is a fixed list of kernels
// that read dt and the loop condition from the control block, so the list
// is captured once in a HIP graph and replayed; the host only reads the
// control block back between batches of sub-steps.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

#include "kernels.h"

#define HIPCHK(call)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return dfargo_set_error(DFARGO_ERR_HIP, "%s failed: %s (%s:%d)", #call,               \
                              hipGetErrorString(e_), __FILE__, __LINE__);                   \
  } while (0)

struct dfargo_ctx {
  dfargo_params p;
  int prec = DFARGO_FP64, dev = 0;
  hipStream_t stream = nullptr;
  dfargo_tables tab;
  size_t esz = 8, nfield = 0; // element size, elements per field
  std::vector<void *> allocs;
  KArgs a;
  dfargo_ctrl *d_ctrl = nullptr, *h_ctrl = nullptr;
  std::vector<int> seq; // kernel ids of one sub-step
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  long graph_version = 0; // bumped whenever the captured sub-step is dropped
  int initialized = 0, profiling = 0;
  double kt_ms[KID_COUNT];
  long kt_n[KID_COUNT];
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // the second stream and the events of the self-gravity branch (graph
  // capture only; default priority)
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // the next sub-step's row transforms of self-gravity are formed by
  // k_finish (a.sg_rows_in_finish), and once at the start of each call
  bool sg_post = false;
  std::vector<hipEvent_t> evpool;
  std::vector<double> VthetaMed, VradMed, gcs_host;
  std::vector<float> fstage;
  // per-DT diagnostics: ring partials and sums (allocated on first use)
  double *diag_part = nullptr, *diag_out = nullptr;
  // dust particles (dfargo_attach_dust)
  int dust_on = 0;
  dfargo_dust_params dp;
  std::vector<double> dustsize_init;
  // DAMPTOVISCOUS: the 1-D viscous disc's initial cell centres, Sigma and
  // vr, uploaded again by every InitEuler
  std::vector<double> v1d_rmed0, v1d_sigma0, v1d_vrad0;
  // the dust fluid (dfargo_attach_dust_fluid): switches and its initial
  // surface density, which every InitEuler starts the fluid from
  dfargo_dust_fluid_params fp;
  std::vector<double> df_drho0;
};

static int noalloc_setup(int dev);

// ------------------------------------------------------------------------

static int dalloc(dfargo_ctx *c, void **p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess)
    return dfargo_set_error(DFARGO_ERR_ALLOC, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  c->allocs.push_back(*p);
  e = hipMemset(*p, 0, bytes);
  if (e != hipSuccess) return dfargo_set_error(DFARGO_ERR_HIP, "hipMemset failed: %s", hipGetErrorString(e));
  return DFARGO_OK;
}

static int upload_table(dfargo_ctx *c, const double **dst, const double *src, size_t n) {
  void *p;
  int rc = dalloc(c, &p, sizeof(double) * n);
  if (rc) return rc;
  HIPCHK(hipMemcpy(p, src, sizeof(double) * n, hipMemcpyHostToDevice));
  *dst = (const double *)p;
  return DFARGO_OK;
}

// host double array -> device field of the context's precision
static int put_field(dfargo_ctx *c, void *dst, const double *src, size_t n) {
  if (c->prec == DFARGO_FP64) {
    HIPCHK(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  } else {
    c->fstage.resize(n);
    for (size_t i = 0; i < n; i++) c->fstage[i] = (float)src[i];
    HIPCHK(hipMemcpyAsync(dst, c->fstage.data(), sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return DFARGO_OK;
}
static int get_field_raw(dfargo_ctx *c, double *dst, const void *src, size_t n) {
  if (c->prec == DFARGO_FP64) {
    HIPCHK(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  } else {
    c->fstage.resize(n);
    HIPCHK(hipMemcpyAsync(c->fstage.data(), src, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; i++) dst[i] = (double)c->fstage[i];
  }
  return DFARGO_OK;
}

static int ctrl_push(dfargo_ctx *c) {
  HIPCHK(hipMemcpyAsync(c->d_ctrl, c->h_ctrl, sizeof(dfargo_ctrl), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return DFARGO_OK;
}
static int ctrl_pull(dfargo_ctx *c) {
  HIPCHK(hipMemcpyAsync(c->h_ctrl, c->d_ctrl, sizeof(dfargo_ctrl), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return DFARGO_OK;
}

// a corotating frame follows planet 0 (or the outermost), a binary
// frame planets 0 and 1: GetPsysInfo reads them unchecked (src/Psys.c:246-271)
static int check_frame_planets(const dfargo_params &p, const dfargo_planets *s) {
  const int need = (p.corotating == DFARGO_FRAME_BINARY_CENTER) ? 2 : (p.corotating ? 1 : 0);
  if (s->nb < need)
    return dfargo_set_error(DFARGO_ERR_ARG, "FRAME %s needs %d planet(s), got %d",
                            p.corotating == DFARGO_FRAME_BINARY_CENTER ? "B" : "C/G", need, s->nb);
  // the non-reflecting boundary takes its pattern speed from the innermost
  // and outermost planets (src/SideEuler.c:590-600,702-712)
  if (p.boundary == DFARGO_BC_NONREFLECTING && s->nb < 1)
    return dfargo_set_error(DFARGO_ERR_ARG, "the non-reflecting boundary needs a planet");
  // accretion (src/Planet.c:11-126) runs in the control kernel after the
  // force sums were formed; it corrects those sums, which holds for the
  // plain columns with a mass-independent smoothing, and it changes the
  // density that the self-gravity transforms (running beside the control
  // kernel) and the particles read
  int nacc = 0;
  for (int k = 0; k < s->nb; k++) nacc += s->acc[k] > 1e-10;
  if (nacc > 1) return dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "at most one accreting planet is supported");
  if (nacc && (p.self_gravity || p.exclude_hill || p.roche_smoothing != 0.0))
    return dfargo_set_error(DFARGO_ERR_UNSUPPORTED,
                            "planet accretion with self-gravity, EXCLUDEHILL or ROCHESMOOTHING is not implemented");
  return DFARGO_OK;
}

// a change of what the sub-step launches drops the graph, so that the next
// sub-step captures it again
static int drop_graph(dfargo_ctx *c) {
  if (c->gexec) {
    HIPCHK(hipGraphExecDestroy(c->gexec));
    c->gexec = nullptr;
  }
  if (c->graph) {
    HIPCHK(hipGraphDestroy(c->graph));
    c->graph = nullptr;
  }
  c->graph_version++;
  return DFARGO_OK;
}

// the planets' accretion switch is part of the captured sub-step: a change
// drops the graph so that the next call captures it again
static int set_accretion(dfargo_ctx *c, const dfargo_planets *s) {
  int acc = 0;
  for (int k = 0; k < s->nb; k++) acc |= s->acc[k] > 1e-10;
  if (acc && c->dust_on)
    return dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "planet accretion with particles is not implemented");
  if (acc != c->a.accrete) {
    c->a.accrete = acc;
    return drop_graph(c);
  }
  return DFARGO_OK;
}

static void planets_to_ctrl(dfargo_ctrl *k, const dfargo_planets *s) {
  k->nb = s->nb;
  for (int i = 0; i < DF_MAXPL; i++) {
    k->pl_m[i] = s->mass[i]; k->pl_x[i] = s->x[i]; k->pl_y[i] = s->y[i];
    k->pl_vx[i] = s->vx[i]; k->pl_vy[i] = s->vy[i]; k->pl_acc[i] = s->acc[i];
    k->pl_m0[i] = s->initial_mass[i]; k->pl_m1[i] = s->final_mass[i];
    k->pl_feel_disk[i] = s->feel_disk[i]; k->pl_feel_others[i] = s->feel_others[i];
    k->pl_binary[i] = s->binary[i];
  }
}
static void ctrl_to_planets(const dfargo_ctrl *k, dfargo_planets *s) {
  s->nb = k->nb;
  for (int i = 0; i < DF_MAXPL; i++) {
    s->mass[i] = k->pl_m[i]; s->x[i] = k->pl_x[i]; s->y[i] = k->pl_y[i];
    s->vx[i] = k->pl_vx[i]; s->vy[i] = k->pl_vy[i]; s->acc[i] = k->pl_acc[i];
    s->initial_mass[i] = k->pl_m0[i]; s->final_mass[i] = k->pl_m1[i];
    s->feel_disk[i] = k->pl_feel_disk[i]; s->feel_others[i] = k->pl_feel_others[i];
    s->binary[i] = k->pl_binary[i];
  }
}

template <typename T> static void launch_one(dfargo_ctx *c, int id, hipStream_t s) { launch_kernel<T>(id, c->a, s); }

static void launch_id(dfargo_ctx *c, int id, hipStream_t s) {
  if (c->prec == DFARGO_FP64) launch_one<double>(c, id, s);
  else launch_one<float>(c, id, s);
}
static void launch_id(dfargo_ctx *c, int id) { launch_id(c, id, c->stream); }

// the kernels of the self-gravity branch that starts with the sub-step
static bool is_sg_kernel(int id) {
  return id == KID_SG_ROWS_FWD || id == KID_SG_COLS || id == KID_SG_ROWS_INV || id == KID_SGZ_MEAN ||
         id == KID_SGZ_CONV || id == KID_SGZ_FILL;
}

// Start of an AlgoGas call on stream s: the device loop condition, the
// sound speed of the state, and (when k_finish's partner forms them) the
// row spectra of the state's density for the first sub-step.
static void call_start(dfargo_ctx *c, double DT, hipStream_t s) {
  launch_algogas_begin(c->a, DT, s);
  if (c->p.energy_equation) {
    if (c->prec == DFARGO_FP64) launch_eos_pre<double>(c->a, s);
    else launch_eos_pre<float>(c->a, s);
  }
  if (c->sg_post) {
    if (c->prec == DFARGO_FP64) launch_sg_rows_forced<double>(c->a, s);
    else launch_sg_rows_forced<float>(c->a, s);
  }
}

static int ensure_graph(dfargo_ctx *c);

static int run_substep(dfargo_ctx *c) {
  if (c->profiling) {
    // one event pair per kernel from a pool; collected once per sub-step
    const size_t nk = c->seq.size();
    while (c->evpool.size() < 2 * nk) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      c->evpool.push_back(e);
    }
    for (size_t q = 0; q < nk; q++) {
      HIPCHK(hipEventRecord(c->evpool[2 * q], c->stream));
      launch_id(c, c->seq[q]);
      HIPCHK(hipGetLastError());
      HIPCHK(hipEventRecord(c->evpool[2 * q + 1], c->stream));
    }
    HIPCHK(hipEventSynchronize(c->evpool[2 * nk - 1]));
    for (size_t q = 0; q < nk; q++) {
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, c->evpool[2 * q], c->evpool[2 * q + 1]));
      c->kt_ms[c->seq[q]] += ms;
      c->kt_n[c->seq[q]] += 1;
    }
    return DFARGO_OK;
  }
  int rc = ensure_graph(c);
  if (rc) return rc;
  HIPCHK(hipGraphLaunch(c->gexec, c->stream));
  return DFARGO_OK;
}

// The sub-step's kernel list captured once as a graph (c->graph) and
// instantiated (c->gexec).
static int ensure_graph(dfargo_ctx *c) {
  if (!c->gexec) {
    // The self-gravity transforms read only the density the sub-step starts
    // with and feed only the source kernel: in the graph they form a branch
    // of their own, beside the CFL, control and particle kernels, joined
    // before k_source.
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const bool fork = c->p.self_gravity != 0;
    if (fork) {
      HIPCHK(hipEventRecord(c->ev_fork, c->stream));
      HIPCHK(hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
      for (int id : c->seq)
        if (is_sg_kernel(id)) launch_id(c, id, c->stream2);
      HIPCHK(hipEventRecord(c->ev_join, c->stream2));
    }
    bool joined = false;
    for (int id : c->seq) {
      if (fork && is_sg_kernel(id)) continue;
      // the first kernel that reads the accelerations waits for them
      if (fork && !joined && (id == KID_SOURCE || id == KID_DF_SOURCE || id == KID_DF_AXISG)) {
        HIPCHK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
        joined = true;
      }
      launch_id(c, id);
    }
    hipError_t e = hipStreamEndCapture(c->stream, &c->graph);
    if (e != hipSuccess) return dfargo_set_error(DFARGO_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    HIPCHK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
  }
  return DFARGO_OK;
}

// ------------------------------------------------------------------------
// creation

// The kernels of one sub-step, in the order of src/SourceEuler.c:398-716.
// Self-gravity only reads the density at the start of the sub-step: its
// transforms run first, and SubStep1 applies the kick (update_sgvelocity).
// The particles move after the planets (src/SourceEuler.c:504-531), i.e.
// after the control kernel, and before the source terms that read their
// feedback.
//
// With the full self-gravity and an edge that k_finish completes (wall or
// open, no wave-killing zones or other boundary kernel after it), k_finish
// can also form the row transforms for the next sub-step from the density
// it stores (sg_post); the sub-step's self-gravity branch then starts with
// the radial transforms.  Measured on MI355X (512x1536 energy+SG disc) the
// ring-per-workgroup k_finish takes the row transforms' time back (21 us
// against 9 us) and the CFL branch, which the radial transforms now meet
// sooner, slows by as much, so the sub-step is no faster (0.172 against
// 0.170 ms): it stays off (kSgRowsInFinish).
static constexpr bool kSgRowsInFinish = false;
static void build_sequence(dfargo_ctx *c) {
  c->seq.clear();
  const bool wkz = c->p.wave_killing_zones && (c->a.wkz_n_in > 0 || c->a.wkz_i_out < c->p.nrad);
  c->sg_post = kSgRowsInFinish && c->p.self_gravity && !c->p.sg_zero_mode &&
               (c->p.boundary == DFARGO_BC_WALL || c->p.boundary == DFARGO_BC_OPEN) && !wkz &&
               c->p.damp_to != DFARGO_DAMP_TO_VISCOUS;
  if (c->p.self_gravity && c->p.sg_zero_mode)
    for (int id : {KID_SGZ_MEAN, KID_SGZ_CONV, KID_SGZ_FILL}) c->seq.push_back(id);
  else if (c->p.self_gravity && c->sg_post)
    for (int id : {KID_SG_COLS, KID_SG_ROWS_INV}) c->seq.push_back(id);
  else if (c->p.self_gravity)
    for (int id : {KID_SG_ROWS_FWD, KID_SG_COLS, KID_SG_ROWS_INV}) c->seq.push_back(id);
  c->seq.push_back(KID_CFL);
  if (c->a.df_on) c->seq.push_back(KID_DF_CFL);
  c->seq.push_back(KID_CONTROL);
  if (c->dust_on)
    for (int id : {KID_DUST_MOVE, KID_DUST_SCAN, KID_DUST_SCATTER, KID_DUST_SORT, KID_DUST_DEPOSIT,
                   KID_DUST_REORDER})
      c->seq.push_back(id);
  const bool df = c->a.df_on != 0;
  // the dust fluid's source step starts from the state the sub-step starts
  // with, before k_source (whose feedback drag reads its Stokes numbers)
  if (df && c->p.self_gravity && !c->p.sg_zero_mode && c->a.df_feel_sg == DFARGO_DUST_SG_ZERO_MODE)
    c->seq.push_back(KID_DF_AXISG);
  if (df) c->seq.push_back(KID_DF_SOURCE);
  c->seq.push_back(KID_SOURCE);
  if (c->p.energy_equation) c->seq.push_back(KID_ENERGY);
  if (df && c->a.df_visc)
    for (int id : {KID_DF_STRESS, KID_DF_VISC}) c->seq.push_back(id);
  if (df) c->seq.push_back(KID_DF_SUBSTEP2);
  c->a.sg_rows_in_finish = c->sg_post ? 1 : 0;
  for (int id : {KID_TRANSPORT_RAD, KID_TRANSPORT_AZI, KID_FINISH}) c->seq.push_back(id);
  if (df)
    for (int id : {KID_DF_TRANSPORT_RAD, KID_DF_TRANSPORT_AZI, KID_DF_FINISH}) c->seq.push_back(id);
  if (c->p.boundary == DFARGO_BC_NONREFLECTING) {
    if (c->p.energy_equation) c->seq.push_back(KID_EOS_STEP);
    c->seq.push_back(KID_NONREFLECTING);
  }
  // SolveOneDViscousEvolution runs whenever DAMPTOVISCOUS is on
  // (src/SourceEuler.c:540-541); nothing reads it before the damping
  if (c->p.damp_to == DFARGO_DAMP_TO_VISCOUS) c->seq.push_back(KID_VISC1D);
  if (c->p.wave_killing_zones && (c->a.wkz_n_in > 0 || c->a.wkz_i_out < c->p.nrad))
    c->seq.push_back(KID_EVANESCENT);
}

static int build_context(dfargo_ctx *c) {
  const dfargo_params &p = c->p;
  const int nr = p.nrad, ns = p.nsec;
  int rc = dfargo_host_tables(&p, &c->tab);
  if (rc) return rc;
  dfargo_tables &t = c->tab;
  KArgs &a = c->a;
  memset(&a, 0, sizeof(a));
  a.nr = nr;
  a.ns = ns;
  a.energy = p.energy_equation;
  a.visc_on = (p.alpha_viscosity != 0.0) || (p.viscosity != 0.0);
  a.alpha_on = p.alpha_viscosity != 0.0;
  a.fast = p.fast_transport;
  a.floors = p.add_floors;
  a.bc = p.boundary;
  a.sg = p.self_gravity;
  a.beta_cooling = p.beta_cooling;
  a.viscous_heating = p.viscous_heating;
  a.gas_indirect = (!p.discard_gas_indirect_term) && (fabs(p.pmax - p.pmin - 2. * M_PI) < 0.01);
  a.indirect_term = p.indirect_term;
  a.alt_indirect = p.alternative_gas_indirect_term;
  a.exclude_hill = p.exclude_hill;
  a.roche = p.roche_smoothing != 0.0;
  a.forced_circular = p.forced_circular;
  a.frame = p.corotating;
  a.frame_outer = p.corotate_with_outer_planet;
  // the wave-killing zones as ring ranges (Rmed increases with i)
  a.wkz_n_in = 0;
  a.wkz_i_out = nr;
  if (p.wave_killing_zones) {
    while (a.wkz_n_in < nr && t.Rmed[a.wkz_n_in] < p.wkz_rmin) a.wkz_n_in++;
    while (a.wkz_i_out > 0 && t.Rmed[a.wkz_i_out - 1] > p.wkz_rmax) a.wkz_i_out--;
  }
  a.wkz_rmin = p.wkz_rmin;
  a.wkz_rmax = p.wkz_rmax;
  a.damp_ratio = p.damping_ratio;
  a.damp_ini = p.damp_to == DFARGO_DAMP_TO_INI;
  a.zero_vrad = p.zero_vrad_bc;
  a.damp_visc = p.damp_to == DFARGO_DAMP_TO_VISCOUS;
  if (a.damp_visc) {
    // InitializeOneDViscousEvolution (src/SideEuler.c:1444-1487), FP64 on
    // the host: a log grid over [RMIN1D, RMAX1D], Sigma and the viscosity
    // of the initial profiles, vr of the steady viscous drift
    const int n = p.nrad1d;
    std::vector<double> radii(n + 1), rmed(n), sig(n), nu(n, 0.0), vr(n);
    for (int i = 0; i <= n; i++) radii[i] = p.rmin1d * exp((double)i / (double)n * log(p.rmax1d / p.rmin1d));
    for (int i = 0; i < n; i++) {
      rmed[i] = 2.0 / 3.0 * (radii[i + 1] * radii[i + 1] * radii[i + 1] - radii[i] * radii[i] * radii[i]);
      rmed[i] = rmed[i] / (radii[i + 1] * radii[i + 1] - radii[i] * radii[i]);
      sig[i] = dfargo_sigma_profile(&p, rmed[i]);
      const double cs = dfargo_aspect_ratio(&p, rmed[i]) / sqrt(rmed[i]) * pow(rmed[i], p.flaring_index);
      if (p.viscosity != 0.0) nu[i] = p.viscosity;
      if (p.alpha_viscosity != 0.0) {
        nu[i] = p.alpha_viscosity * cs * cs * pow(rmed[i], 1.5);
        vr[i] = -3.0 * nu[i] / rmed[i] * (-p.sigma_slope + 2.0 * p.flaring_index + 1.0);
      } else
        vr[i] = -3.0 * nu[i] / rmed[i] * (-p.sigma_slope + 0.5);
    }
    // the damping interpolates the 1-D disc at every zone ring's Rmed
    // (src/SideEuler.c:868-887); without the zones it only evolves
    if (p.wave_killing_zones)
      for (int i = 0; i < nr; i++)
        if (!(rmed[0] < t.Rmed[i] && t.Rmed[i] <= rmed[n - 1]))
          return dfargo_set_error(DFARGO_ERR_ARG, "the 1-D grid's cell centres must enclose every Rmed");
    c->v1d_rmed0 = rmed;
    c->v1d_sigma0 = sig;
    c->v1d_vrad0 = vr;
    a.n1d = n;
    a.bc1d = p.bc1d;
    if ((rc = upload_table(c, &a.v1d_rmed, rmed.data(), n)) || (rc = upload_table(c, &a.v1d_nu, nu.data(), n)) ||
        (rc = upload_table(c, (const double **)&a.v1d_sigma, sig.data(), n)) ||
        (rc = upload_table(c, (const double **)&a.v1d_vrad, vr.data(), n)))
      return rc;
  }
  a.xcd_remap = 1;
  a.tune_azi_waves = 4;
  a.tune_finish_block = 256;
  a.sg_fft_buffers = 0;
  a.sg_rows_pair = 0;
  {
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, c->dev));
    a.num_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  }
  a.pm = p.pmax - p.pmin;
  a.pmin = p.pmin;
  a.dphi = a.pm / (double)ns;
  a.invdphi = 1.0 / a.dphi;
  a.gamma = p.adiabatic_index;
  a.alpha = p.alpha_viscosity;
  a.visc_const = p.viscosity;
  a.floor_dens = p.density_floor;
  a.sg_eps = t.sg_eps;
  a.sg_rstep = t.sg_rstep;
  a.sg_tstep = t.sg_tstep;
  a.vh_lr1 = t.vh_lr1;
  a.vh_lr2 = t.vh_lr2;
  a.mass_taper_param = p.mass_taper;
  a.thickness_smoothing = p.thickness_smoothing;
  a.roche_smoothing = p.roche_smoothing;
  a.indirect_smoothing = p.indirect_term_smoothing;
  a.exclude_hill_factor = p.exclude_hill_factor;
  a.release_date = p.release_date;
  a.release_radius = p.release_radius;
  a.aspect_ratio = p.aspect_ratio;
  a.flaring = p.flaring_index;
  a.transition_radius = p.transition_radius;
  a.transition_ratio = p.transition_ratio;
  a.transition_width = p.transition_width;
  // tables
  size_t n1 = (size_t)nr + 2;
  if ((rc = upload_table(c, &a.Rinf, t.Rinf, n1)) || (rc = upload_table(c, &a.Rsup, t.Rsup, n1)) ||
      (rc = upload_table(c, &a.Rmed, t.Rmed, n1)) || (rc = upload_table(c, &a.Surf, t.Surf, n1)) ||
      (rc = upload_table(c, &a.InvRmed, t.InvRmed, n1)) || (rc = upload_table(c, &a.InvSurf, t.InvSurf, n1)) ||
      (rc = upload_table(c, &a.InvDiffRmed, t.InvDiffRmed, n1)) ||
      (rc = upload_table(c, &a.InvDiffRsup, t.InvDiffRsup, n1)) ||
      (rc = upload_table(c, &a.InvRinf, t.InvRinf, n1)) || (rc = upload_table(c, &a.CosAz, t.CosAz, ns)) ||
      (rc = upload_table(c, &a.SinAz, t.SinAz, ns)) || (rc = upload_table(c, &a.SigmaMed, t.SigmaMed, n1)) ||
      (rc = upload_table(c, &a.EnergyMed, t.EnergyMed, n1)) || (rc = upload_table(c, &a.cs_iso, t.cs_iso, n1)) ||
      (rc = upload_table(c, &a.powmin2, t.powmin2, n1)) || (rc = upload_table(c, &a.r15, t.r15, n1)) ||
      (rc = upload_table(c, &a.visc_mult, t.visc_mult, n1)) || (rc = upload_table(c, &a.beta, t.beta, n1)) ||
      (rc = upload_table(c, &a.supp_torque, t.supp_torque, n1)) ||
      (rc = upload_table(c, &a.sg_sqrt, t.sg_sqrt, n1)) || (rc = upload_table(c, &a.sg_normr, t.sg_normr, n1)) ||
      (rc = upload_table(c, &a.sg_normt, t.sg_normt, n1)))
    return rc;
  {
    double o1 = sqrt(t.Rmed[2] / t.Rmed[1]), o2 = sqrt(t.Rmed[2] / t.Rmed[0]);
    double o3 = pow(t.Rmed[nr - 1] / t.Rmed[nr - 2], -0.5);
    if ((rc = upload_table(c, &a.open_vt_in1, &o1, 1)) || (rc = upload_table(c, &a.open_vt_in2, &o2, 1)) ||
        (rc = upload_table(c, &a.open_vt_out, &o3, 1)))
      return rc;
  }
  {
    // 1 / dxtheta of SubStep1/2 (src/SourceEuler.c:933-934), per ring
    std::vector<double> inv(n1, 0.0);
    for (int i = 0; i < nr; i++) {
      const double dxtheta = (p.pmax - p.pmin) / (double)ns * t.Rmed[i];
      inv[i] = 1.0 / dxtheta;
    }
    if ((rc = upload_table(c, &a.InvDxTheta, inv.data(), n1))) return rc;
  }
  c->VthetaMed.assign(n1, 0.0);
  if ((rc = upload_table(c, &a.VthetaMed, c->VthetaMed.data(), n1))) return rc;
  c->VradMed.assign(n1, 0.0);
  if ((rc = upload_table(c, &a.VradMed, c->VradMed.data(), n1))) return rc;
  // fields: (nr+2) rings so that ring nr (vr outer face) exists with slack
  c->esz = (c->prec == DFARGO_FP64) ? sizeof(double) : sizeof(float);
  c->nfield = (size_t)(nr + 2) * ns;
  void **fields[] = {&a.rho, &a.vr, &a.vt, &a.en, &a.cs, &a.vri, &a.vti, &a.trr, &a.trp, &a.tpp,
                     &a.divv, &a.vh, &a.t_rho, &a.t_rp, &a.t_rm, &a.t_tp, &a.t_tm, &a.t_en, &a.pot, &a.e_int,
                     &a.z_rho, &a.z_rp, &a.z_rm, &a.z_tp, &a.z_tm, &a.z_en};
  for (void **f : fields)
    if ((rc = dalloc(c, f, c->esz * c->nfield))) return rc;
  void **rings[] = {(void **)&a.vmoy, (void **)&a.ring_dt, (void **)&a.nu_ring, (void **)&a.gcs,
                    (void **)&a.vmed, (void **)&a.lost_col};
  for (void **r : rings)
    if ((rc = dalloc(c, r, sizeof(double) * (size_t)(std::max(nr, ns) + 2)))) return rc;
  if ((rc = dalloc(c, (void **)&a.ring_part, sizeof(double) * (size_t)(nr + 2) * DF_NPART(DF_MAXTG))) ||
      (rc = dalloc(c, (void **)&a.nshift, sizeof(long) * (size_t)(nr + 2))) ||
      (rc = dalloc(c, (void **)&a.vt_part, sizeof(double) * (size_t)nr * ((ns + DF_SRC_COLS - 1) / DF_SRC_COLS))) ||
      (rc = dalloc(c, (void **)&c->d_ctrl, sizeof(dfargo_ctrl))))
    return rc;
  a.ctrl = c->d_ctrl;
  a.sg_zero = p.self_gravity && p.sg_zero_mode;
  if (a.sg_zero) {
    // SELFGRAVITY Z (src/sgzero.c:3-84): the reduced kernel summed over
    // the azimuths, times tstep, on the 2 Nrad log-radius cells, and the
    // ring's own term of the self-force compensation, FP64 on the host
    const int L2 = 2 * nr;
    const double eps = t.sg_eps;
    if ((rc = dalloc(c, &a.sgr, c->esz * c->nfield)) || (rc = dalloc(c, &a.sgt, c->esz * c->nfield)) ||
        (rc = dalloc(c, (void **)&a.sgz_mean, sizeof(double) * (size_t)(nr + 2))) ||
        (rc = dalloc(c, (void **)&a.sgz_acc, sizeof(double) * (size_t)(nr + 2))))
      return rc;
    std::vector<double> kr(L2, 0.0);
    for (int i = 0; i < L2; i++) {
      const double u = (i < nr) ? log(t.Radii[i] / t.Radii[0]) : -log(t.Radii[2 * nr - i] / t.Radii[0]);
      for (int j = 0; j < ns; j++)
        kr[i] += (1.0 + eps * eps - t.CosAz[j] * exp(-u)) * pow(eps * eps * exp(u) + 2.0 * (cosh(u) - t.CosAz[j]), -1.5);
      kr[i] *= t.sg_tstep;
    }
    if ((rc = upload_table(c, &a.sgz_kr, kr.data(), L2))) return rc;
    double integrale = 0.0;
    for (int j = 0; j < ns; j++)
      integrale += (1.0 + eps * eps - t.CosAz[j]) / pow(2.0 * (1.0 - t.CosAz[j]) + eps * eps, 1.5);
    a.sgz_integrale = integrale * t.sg_tstep;
  }
  // self-gravity: twiddles and kernel spectra
  if (p.self_gravity && !a.sg_zero) {
    const int L2 = 2 * nr;
    const size_t csz = 2 * c->esz;
    if ((rc = dalloc(c, &a.sgr, c->esz * c->nfield)) || (rc = dalloc(c, &a.sgt, c->esz * c->nfield)) ||
        (rc = dalloc(c, &a.sg_z, csz * (size_t)nr * ns)) || (rc = dalloc(c, &a.sg_w, csz * (size_t)nr * ns)) ||
        (rc = dalloc(c, &a.sg_kr, csz * (size_t)(ns / 2 + 1) * L2)) ||
        (rc = dalloc(c, &a.sg_kt, csz * (size_t)(ns / 2 + 1) * L2)) ||
        (rc = dalloc(c, &a.tw_ns, csz * ns)) || (rc = dalloc(c, &a.tw_2n, csz * L2)))
      return rc;
    // twiddles exp(-2 pi i m / L), FP64 host values
    std::vector<double> tw(2 * (size_t)std::max(ns, L2));
    auto put_tw = [&](void *dst, int L) -> int {
      for (int m = 0; m < L; m++) {
        tw[2 * m] = cos(2.0 * M_PI * (double)m / (double)L);
        tw[2 * m + 1] = -sin(2.0 * M_PI * (double)m / (double)L);
      }
      return put_field(c, dst, tw.data(), 2 * (size_t)L);
    };
    if ((rc = put_tw(a.tw_ns, ns)) || (rc = put_tw(a.tw_2n, L2))) return rc;
    // kernels (src/sgkernel.c:3-27) in FP64 on the host, spectra on the GPU
    const double eps = t.sg_eps;
    std::vector<double> kr(2 * (size_t)L2 * ns, 0.0), kt(2 * (size_t)L2 * ns, 0.0);
    for (int i = 0; i < L2; i++) {
      double u = (i < nr) ? log(t.Radii[i] / t.Radii[0]) : -log(t.Radii[2 * nr - i] / t.Radii[0]);
      for (int j = 0; j < ns; j++) {
        size_t l = (size_t)i * ns + j;
        double v = 1.0 + eps * eps - t.CosAz[j] * exp(-u);
        v *= pow(eps * eps * exp(u) + 2.0 * (cosh(u) - t.CosAz[j]), -1.5);
        kr[2 * l] = v;
        double w = t.SinAz[j];
        w *= pow(eps * eps * exp(u) + 2.0 * (cosh(u) - t.CosAz[j]), -1.5);
        kt[2 * l] = w;
      }
    }
    void *tmp = nullptr, *twd_ns = nullptr, *twd_2n = nullptr;
    HIPCHK(hipMalloc(&tmp, sizeof(double) * 2 * (size_t)L2 * ns));
    HIPCHK(hipMalloc(&twd_ns, sizeof(double) * 2 * ns));
    HIPCHK(hipMalloc(&twd_2n, sizeof(double) * 2 * L2));
    for (int m = 0; m < ns; m++) {
      tw[2 * m] = cos(2.0 * M_PI * (double)m / (double)ns);
      tw[2 * m + 1] = -sin(2.0 * M_PI * (double)m / (double)ns);
    }
    HIPCHK(hipMemcpy(twd_ns, tw.data(), sizeof(double) * 2 * ns, hipMemcpyHostToDevice));
    for (int m = 0; m < L2; m++) {
      tw[2 * m] = cos(2.0 * M_PI * (double)m / (double)L2);
      tw[2 * m + 1] = -sin(2.0 * M_PI * (double)m / (double)L2);
    }
    HIPCHK(hipMemcpy(twd_2n, tw.data(), sizeof(double) * 2 * L2, hipMemcpyHostToDevice));
    // the k_sg_cols kernel reads wavenumbers 0..Ns/2 only, one contiguous
    // column of 2 Nrad values per wavenumber: keep that half, transposed
    void *dsts[2] = {a.sg_kr, a.sg_kt};
    std::vector<double> *srcs[2] = {&kr, &kt};
    const int nh = ns / 2 + 1;
    std::vector<double> half(2 * (size_t)nh * L2);
    for (int q = 0; q < 2; q++) {
      HIPCHK(hipMemcpy(tmp, srcs[q]->data(), sizeof(double) * 2 * (size_t)L2 * ns, hipMemcpyHostToDevice));
      launch_fft_rows<double>(tmp, L2, ns, twd_ns, -1, c->stream);
      launch_fft_cols<double>(tmp, ns, L2, twd_2n, -1, c->stream);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(c->stream));
      HIPCHK(hipMemcpy(srcs[q]->data(), tmp, sizeof(double) * 2 * (size_t)L2 * ns, hipMemcpyDeviceToHost));
      for (int k2 = 0; k2 < nh; k2++)
        for (int k1 = 0; k1 < L2; k1++) {
          half[2 * ((size_t)k2 * L2 + k1)] = (*srcs[q])[2 * ((size_t)k1 * ns + k2)];
          half[2 * ((size_t)k2 * L2 + k1) + 1] = (*srcs[q])[2 * ((size_t)k1 * ns + k2) + 1];
        }
      if ((rc = put_field(c, dsts[q], half.data(), half.size()))) return rc;
    }
    HIPCHK(hipFree(tmp));
    HIPCHK(hipFree(twd_ns));
    HIPCHK(hipFree(twd_2n));
  }
  // control block
  HIPCHK(hipHostMalloc((void **)&c->h_ctrl, sizeof(dfargo_ctrl), hipHostMallocDefault));
  memset(c->h_ctrl, 0, sizeof(dfargo_ctrl));
  c->h_ctrl->omega_frame = p.omega_frame;
  c->h_ctrl->mindt = 100.0; /* src/var.c:161-162 defaults */
  c->h_ctrl->maxdt = 0.00001;
  c->h_ctrl->active = 0;
  if ((rc = ctrl_push(c))) return rc;
  if (c->prec == DFARGO_FP64) prepare_kernels<double>(a);
  else prepare_kernels<float>(a);
  HIPCHK(hipGetLastError());
  build_sequence(c);
  HIPCHK(hipEventCreate(&c->ev0));
  HIPCHK(hipEventCreate(&c->ev1));
  HIPCHK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));

  return DFARGO_OK;
}

extern "C" int dfargo_create(const dfargo_params *p, int precision, int device, dfargo_ctx **out) {
  if (!out) return dfargo_set_error(DFARGO_ERR_ARG, "null output pointer");
  *out = nullptr;
  int rc = dfargo_check_params(p);
  if (rc) return rc;
  if (precision != DFARGO_FP64 && precision != DFARGO_FP32)
    return dfargo_set_error(DFARGO_ERR_ARG, "precision must be DFARGO_FP64 or DFARGO_FP32");
  if (p->self_gravity) {
    // the radial transforms of one wavenumber pair hold four 2*NRAD complex
    // columns in the 160 KiB of LDS
    size_t L = 2 * (size_t)p->nrad, lds = 4 * (L + L / 16 + 1) * 2 * (precision == DFARGO_FP64 ? 8 : 4);
    if (lds > 160 * 1024)
      return dfargo_set_error(DFARGO_ERR_UNSUPPORTED,
                              "self-gravity with NRAD=%d needs %zu bytes of LDS (> 160 KiB) at this precision",
                              p->nrad, lds);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return dfargo_set_error(DFARGO_ERR_HIP, "no HIP device is visible");
  if (device < 0 || device >= ndev) return dfargo_set_error(DFARGO_ERR_ARG, "device %d out of range", device);
  if (ring_kcheck(p->nsec) < 0)
    return dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "NSEC=%d cannot be split into <=1024-thread rings", p->nsec);
  HIPCHK(hipSetDevice(device));
  if ((rc = noalloc_setup(device))) return rc;
  dfargo_ctx *c = new dfargo_ctx();
  c->p = *p;
  c->prec = precision;
  c->dev = device;
  memset(c->kt_ms, 0, sizeof c->kt_ms);
  memset(c->kt_n, 0, sizeof c->kt_n);
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return dfargo_set_error(DFARGO_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
  }
  rc = build_context(c);
  if (rc) {
    dfargo_destroy(c);
    return rc;
  }
  *out = c;
  return DFARGO_OK;
}

extern "C" int dfargo_destroy(dfargo_ctx *c) {
  if (!c) return DFARGO_OK;
  hipSetDevice(c->dev);
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  if (c->graph) hipGraphDestroy(c->graph);
  for (void *p : c->allocs) hipFree(p);
  if (c->h_ctrl) hipHostFree(c->h_ctrl);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->ev_fork) hipEventDestroy(c->ev_fork);
  if (c->ev_join) hipEventDestroy(c->ev_join);

  if (c->stream2) hipStreamDestroy(c->stream2);
  for (hipEvent_t e : c->evpool) hipEventDestroy(e);
  if (c->stream) hipStreamDestroy(c->stream);
  dfargo_free_tables(&c->tab);
  delete c;
  return DFARGO_OK;
}

// ------------------------------------------------------------------------
// InitEuler

static int put_doubles(dfargo_ctx *c, double *dst, const double *src, size_t n);

// The dust fluid's start-up in InitEuler (src/SourceEuler.c:265-269):
// InitDustStokesNumber on the gas density (zero relative velocity,
// src/Pframeforce.c:836-874) and InitDustVelocities (src/Pframeforce.c:
// 644-830) as src/SourceEuler.c:268 calls it, with the two densities in each
// other's places (in the feedback branch its eps is gas over dust).  The
// routine's earlier assignments are all overwritten by its final loops,
// which alone are restated.  The fluid's state, Stokes grid and
// DVthetaMed / DVradMed go to the device; St returns for the gas velocities.
static int df_init(dfargo_ctx *c, const double *rho, const std::vector<double> &axisg, std::vector<double> &St) {
  const dfargo_params &p = c->p;
  const dfargo_tables &t = c->tab;
  const dfargo_dust_fluid_params &f = c->fp;
  const int nr = p.nrad, ns = p.nsec;
  const size_t n = (size_t)(nr + 1) * ns;
  KArgs &a = c->a;
  const double *drho = c->df_drho0.data();
  const double um = 1.9891e30 * f.factor_unit_mass, ul = 1.49598e11 * f.factor_unit_length;
  const double rho_int = f.rhopart * 1000.0 * pow(ul, 3.) / um, radius = f.sizepart / ul;
  St.assign(n, 0.0);
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      const size_t l = (size_t)j + (size_t)i * ns;
      const double densg_cgs = 0.1 * rho[l] * um * pow(ul, -2.0);
      const double lambda = 3.34e-8 * pow(densg_cgs, -1.) * dfargo_aspect_ratio(&p, t.Rmed[i]) * t.Rmed[i];
      const double Kn = 0.5 * lambda / radius, f_D = 1.0, k_D = 1.0;
      const double Cdrag = pow(3.0 * Kn + 1.0, 2.0) * pow(9.0 * Kn * Kn * f_D + 3.0 * Kn * k_D, -1.0);
      St[l] = 0.5 * M_PI * Cdrag * rho_int * radius / rho[l];
    }
  std::vector<double> dvr(n, 0.0), dvt(n, 0.0), dvtm(nr + 2, 0.0), dvrm(nr + 2, 0.0);
  for (int i = 0; i < nr; i++) {
    const double r = t.Rmed[i];
    double vk = sqrt(1.0 * 1.0 / r);
    const double eta = -0.5 * pow(p.aspect_ratio, 2.0) * pow(r, 2.0 * p.flaring_index) *
                       (-1. - p.sigma_slope + 2.0 * p.flaring_index);
    if (p.self_gravity) vk = sqrt(1.0 * 1.0 / r - r * axisg[i]);
    for (int j = 0; j < ns; j++) {
      const size_t l = (size_t)i * ns + j;
      if (f.dust_feedback) {
        const double eps = rho[l] / drho[l];
        dvr[l] = -2.0 * eta * vk / (St[l] + pow(1.0 + eps, 2.0) / St[l]);
        dvt[l] = vk - eta * vk * (1.0 + eps) / (pow(St[l], 2.0) + pow(1.0 + eps, 2.0));
      } else {
        dvr[l] = -2.0 * eta * vk / (St[l] + pow(St[l], -1.0));
        dvt[l] = vk - eta * vk / (pow(St[l], 2.0) + 1.0);
      }
      dvt[l] -= p.omega_frame * r;
    }
    dvtm[i] = dvt[(size_t)i * ns] + r * p.omega_frame;
    dvrm[i] = dvr[(size_t)i * ns];
  }
  int rc;
  if ((rc = put_doubles(c, a.drho, drho, n)) || (rc = put_doubles(c, a.dvr, dvr.data(), n)) ||
      (rc = put_doubles(c, a.dvt, dvt.data(), n)) || (rc = put_doubles(c, a.dstokes, St.data(), n)) ||
      (rc = put_doubles(c, (double *)a.DVthetaMed, dvtm.data(), nr + 2)) ||
      (rc = put_doubles(c, (double *)a.DVradMed, dvrm.data(), nr + 2)))
    return rc;
  // DVradInt / DVthetaInt of rings 0 and nrad stay zero, as in the reference
  HIPCHK(hipMemset(a.dvri, 0, sizeof(double) * c->nfield));
  HIPCHK(hipMemset(a.dvti, 0, sizeof(double) * c->nfield));
  return DFARGO_OK;
}

extern "C" int dfargo_init_euler(dfargo_ctx *c, const double *rho, double *vrad, double *vtheta,
                                 const double *energy, const dfargo_planets *sys_in) {
  if (!c || !rho || !vrad || !vtheta) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  if (c->p.energy_equation && !energy)
    return dfargo_set_error(DFARGO_ERR_ARG, "the energy equation needs an energy array");
  HIPCHK(hipSetDevice(c->dev));
  const dfargo_params &p = c->p;
  const dfargo_tables &t = c->tab;
  const int nr = p.nrad, ns = p.nsec;
  const size_t n = (size_t)(nr + 1) * ns;
  KArgs &a = c->a;
  int rc;
  dfargo_planets sys;
  if (sys_in) sys = *sys_in;
  else memset(&sys, 0, sizeof sys);
  if ((rc = check_frame_planets(p, &sys)) || (rc = set_accretion(c, &sys))) return rc;
  std::vector<double> zero(n, 0.0);
  if ((rc = put_field(c, a.rho, rho, n)) || (rc = put_field(c, a.vr, zero.data(), n)) ||
      (rc = put_field(c, a.vt, zero.data(), n)))
    return rc;
  if (p.energy_equation && (rc = put_field(c, a.en, energy, n))) return rc;
  // start-up self-gravity (src/main.c:178-186): SG_Accr of the initial
  // density and its ring mean for the planets and InitGasVelocities
  std::vector<double> axisg(nr, 0.0);
  if (p.self_gravity) {
    if (c->prec == DFARGO_FP64) launch_sg_accel<double>(a, c->stream);
    else launch_sg_accel<float>(a, c->stream);
    HIPCHK(hipGetLastError());
    if (a.sg_zero) {
      // the axisymmetric mode's SG_Accr is the profile itself
      // (src/sgsysinit.c:12-15, src/Pframeforce.c:483-486)
      HIPCHK(hipStreamSynchronize(c->stream));
      HIPCHK(hipMemcpy(axisg.data(), a.sgz_acc, sizeof(double) * nr, hipMemcpyDeviceToHost));
    } else {
      std::vector<double> accr(n);
      if ((rc = get_field_raw(c, accr.data(), a.sgr, (size_t)nr * ns))) return rc;
      for (int i = 0; i < nr; i++) {
        double s = 0.;
        for (int j = 0; j < ns; j++) s += accr[(size_t)i * ns + j];
        axisg[i] = s / (double)ns;
      }
    }
    // init_planetarysys_withSG (src/sgsysinit.c:3-49), non-binary
    for (int k = 0; k < sys.nb; k++) {
      double x = sys.x[k], y = sys.y[k], r = sqrt(x * x + y * y), sgacc;
      double dist = (double)(r / (1. + p.eccentricity));
      if (dist < t.Rmed[nr - 1]) {
        int ipl = 0;
        while ((t.Rmed[ipl] <= dist) && (ipl < nr - 2)) ipl++;
        double ri = t.Rmed[ipl], rip1 = t.Rmed[ipl + 1], dr = rip1 - ri;
        sgacc = (dist - ri) * axisg[ipl + 1] + (rip1 - dist) * axisg[ipl];
        sgacc /= dr;
      } else
        sgacc = axisg[nr - 1];
      sys.vy[k] *= (double)sqrt(1. - dist * dist * sgacc / 1.0 / (1.0 + sys.mass[k]));
    }
  }
  // ComputeSoundSpeed and the ring mean GLOBAL_SoundSpeed
  std::vector<double> cs(n, 0.0);
  c->gcs_host.assign(nr + 2, 0.0);
  for (int i = 0; i < nr; i++) {
    for (int j = 0; j < ns; j++) {
      size_t l = (size_t)i * ns + j;
      cs[l] = p.energy_equation
                  ? sqrt(p.adiabatic_index * (p.adiabatic_index - 1.0) * energy[l] / rho[l])
                  : t.cs_iso[i];
    }
    double s = 0.;
    for (int j = 0; j < ns; j++) s += cs[(size_t)i * ns + j];
    c->gcs_host[i] = s / (double)ns;
  }
  if (p.energy_equation && (rc = put_field(c, a.cs, cs.data(), n))) return rc;
  std::vector<double> nu(nr + 2, 0.0);
  for (int i = 0; i < nr; i++) {
    double v = p.viscosity;
    if (p.alpha_viscosity != 0.0) v = p.alpha_viscosity * c->gcs_host[i] * c->gcs_host[i] * t.r15[i];
    nu[i] = ((p.alpha_viscosity != 0.0) || (p.viscosity != 0.0)) ? v * t.visc_mult[i] : 0.0;
  }
  HIPCHK(hipMemcpy(a.nu_ring, nu.data(), sizeof(double) * (nr + 2), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(a.gcs, c->gcs_host.data(), sizeof(double) * (nr + 2), hipMemcpyHostToDevice));
  std::vector<double> dfSt;
  if (a.df_on && (rc = df_init(c, rho, axisg, dfSt))) return rc;
  // InitGasVelocities (src/Pframeforce.c:442-641), no centrifugal balance
  std::vector<double> vr(n, 0.0), vt(n, 0.0);
  for (int i = 0; i < nr; i++) {
    double r = t.Rmed[i], omega = sqrt(1.0 * 1.0 / r / r / r);
    double myvtheta = omega * r * sqrt(1.0 - pow(p.aspect_ratio, 2.0) * pow(r, 2.0 * p.flaring_index) *
                                                 (1. + p.sigma_slope - 2.0 * p.flaring_index));
    for (int j = 0; j < ns; j++) vt[j + (size_t)i * ns] = myvtheta;
  }
  if (p.self_gravity)
    for (int i = 0; i < nr; i++)
      for (int j = 0; j < ns; j++) {
        size_t l = (size_t)i * ns + j;
        double vt2 = vt[l] * vt[l];
        vt2 -= t.Rmed[i] * axisg[i];
        vt[l] = sqrt(vt2);
      }
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) vt[(size_t)i * ns + j] -= p.omega_frame * t.Rmed[i];
  for (int i = 0; i < nr; i++) {
    double r = t.Rmed[i], ri = t.Rinf[i], viscosity = nu[i];
    for (int j = 0; j < ns; j++) {
      size_t l = (size_t)i * ns + j;
      if (p.imposed_disk_drift != 0.0)
        vr[l] = p.imposed_disk_drift * p.sigma0 / t.SigmaInf[i] / ri;
      else if (p.alpha_viscosity != 0.0)
        vr[l] -= 3.0 * viscosity / r * (-p.sigma_slope + 2.0 * p.flaring_index + 1.0);
      else
        vr[l] -= 3.0 * viscosity / r * (-p.sigma_slope + .5);
    }
  }
  // with the dust fluid's feedback, the gas velocities of the drag
  // equilibrium (Nakagawa, Sekiya & Hayashi 1986; src/Pframeforce.c:609-634)
  if (a.df_on && c->fp.dust_feedback)
    for (int i = 0; i < nr; i++) {
      const double r = t.Rmed[i];
      double vk = sqrt(1.0 * 1.0 / r);
      const double eta = -0.5 * pow(p.aspect_ratio, 2.0) * pow(r, 2.0 * p.flaring_index) *
                         (-1. - p.sigma_slope + 2.0 * p.flaring_index);
      if (p.self_gravity) vk = sqrt(1.0 * 1.0 / r - r * axisg[i]);
      for (int j = 0; j < ns; j++) {
        const size_t l = (size_t)i * ns + j;
        const double eps = c->df_drho0[l] / rho[l], St = dfSt[l];
        vr[l] = 2.0 * eps * eta * vk * St / (St * St + pow(1.0 + eps, 2.0));
        vt[l] = vk * (1.0 - eta) + eps * eta * vk * (1.0 + eps) / (pow(St, 2.0) + pow(1.0 + eps, 2.0));
        vt[l] -= p.omega_frame * r;
      }
    }
  // the initial profiles the wave-killing zones may damp towards
  // (src/Pframeforce.c:635-640)
  for (int i = 0; i < nr; i++) {
    c->VthetaMed[i] = vt[(size_t)i * ns] + t.Rmed[i] * p.omega_frame;
    c->VradMed[i] = vr[(size_t)i * ns];
  }
  HIPCHK(hipMemcpy((void *)a.VthetaMed, c->VthetaMed.data(), sizeof(double) * (nr + 2), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy((void *)a.VradMed, c->VradMed.data(), sizeof(double) * (nr + 2), hipMemcpyHostToDevice));
  // the 1-D viscous disc restarts from its initial profiles
  // (InitializeOneDViscousEvolution, src/SideEuler.c:1444-1487)
  if (a.damp_visc) {
    HIPCHK(hipMemcpy((void *)a.v1d_rmed, c->v1d_rmed0.data(), sizeof(double) * a.n1d, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(a.v1d_sigma, c->v1d_sigma0.data(), sizeof(double) * a.n1d, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(a.v1d_vrad, c->v1d_vrad0.data(), sizeof(double) * a.n1d, hipMemcpyHostToDevice));
  }
  memcpy(vrad, vr.data(), sizeof(double) * n);
  memcpy(vtheta, vt.data(), sizeof(double) * n);
  if ((rc = put_field(c, a.vr, vr.data(), n)) || (rc = put_field(c, a.vt, vt.data(), n))) return rc;
  // ComputeViscousTerms at start-up (stress grids for output)
  if (c->prec == DFARGO_FP64) launch_visc_stress_state<double>(a, c->stream);
  else launch_visc_stress_state<float>(a, c->stream);
  HIPCHK(hipGetLastError());
  // control block
  dfargo_ctrl *k = c->h_ctrl;
  planets_to_ctrl(k, &sys);
  k->time = 0.0;
  k->time0 = 0.0;
  k->omega_frame = p.omega_frame;
  if ((rc = ctrl_push(c))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  c->initialized = 1;
  return DFARGO_OK;
}

// ------------------------------------------------------------------------
// AlgoGas

extern "C" int dfargo_algogas_device(dfargo_ctx *c, double DT) {
  if (!c) return dfargo_set_error(DFARGO_ERR_ARG, "null context");
  if (!c->initialized) return dfargo_set_error(DFARGO_ERR_STATE, "dfargo_init_euler has not been called");
  if (!(DT > 0.0)) return dfargo_set_error(DFARGO_ERR_ARG, "DT must be positive");
  HIPCHK(hipSetDevice(c->dev));
  call_start(c, DT, c->stream);
  HIPCHK(hipGetLastError());
  long launched = 0;
  int batch = 1, rc;
  const long cap = 50000000L;
  for (;;) {
    for (int b = 0; b < batch; b++)
      if ((rc = run_substep(c))) return rc;
    launched += batch;
    if ((rc = ctrl_pull(c))) return rc;
    const dfargo_ctrl *k = c->h_ctrl;
    if (!(k->dtemp < 0.999999999 * DT)) break;
    if (launched > cap)
      return dfargo_set_error(DFARGO_ERR_STATE, "more than %ld sub-steps in one AlgoGas call", cap);
    double rem = (DT - k->dtemp) / (k->dt > 0.0 ? k->dt : DT);
    batch = (int)std::min(std::max(ceil(rem), 1.0), 4096.0);
  }
  return DFARGO_OK;
}

extern "C" int dfargo_run_substeps(dfargo_ctx *c, double DT, int nsub) {
  if (!c) return dfargo_set_error(DFARGO_ERR_ARG, "null context");
  if (!c->initialized) return dfargo_set_error(DFARGO_ERR_STATE, "dfargo_init_euler has not been called");
  if (!(DT > 0.0) || nsub < 0) return dfargo_set_error(DFARGO_ERR_ARG, "need DT > 0 and nsub >= 0");
  HIPCHK(hipSetDevice(c->dev));
  call_start(c, DT, c->stream);
  HIPCHK(hipGetLastError());
  int rc;
  for (int b = 0; b < nsub; b++)
    if ((rc = run_substep(c))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return DFARGO_OK;
}

extern "C" int dfargo_upload_state(dfargo_ctx *c, const double *rho, const double *vrad, const double *vtheta,
                                   const double *energy, const dfargo_planets *sys) {
  if (!c) return dfargo_set_error(DFARGO_ERR_ARG, "null context");
  HIPCHK(hipSetDevice(c->dev));
  const size_t n = (size_t)(c->p.nrad + 1) * c->p.nsec;
  int rc;
  if (rho && (rc = put_field(c, c->a.rho, rho, n))) return rc;
  if (vrad && (rc = put_field(c, c->a.vr, vrad, n))) return rc;
  if (vtheta && (rc = put_field(c, c->a.vt, vtheta, n))) return rc;
  if (energy && c->p.energy_equation && (rc = put_field(c, c->a.en, energy, n))) return rc;
  if (sys) {
    if ((rc = check_frame_planets(c->p, sys)) || (rc = set_accretion(c, sys))) return rc;
    if ((rc = ctrl_pull(c))) return rc;
    planets_to_ctrl(c->h_ctrl, sys);
    if ((rc = ctrl_push(c))) return rc;
  }
  return DFARGO_OK;
}

extern "C" int dfargo_download_state(dfargo_ctx *c, double *rho, double *vrad, double *vtheta, double *energy,
                                     dfargo_planets *sys) {
  if (!c) return dfargo_set_error(DFARGO_ERR_ARG, "null context");
  HIPCHK(hipSetDevice(c->dev));
  const size_t n = (size_t)(c->p.nrad + 1) * c->p.nsec;
  int rc;
  if (rho && (rc = get_field_raw(c, rho, c->a.rho, n))) return rc;
  if (vrad && (rc = get_field_raw(c, vrad, c->a.vr, n))) return rc;
  if (vtheta && (rc = get_field_raw(c, vtheta, c->a.vt, n))) return rc;
  if (energy) {
    if (c->p.energy_equation) {
      if ((rc = get_field_raw(c, energy, c->a.en, n))) return rc;
    } else
      memset(energy, 0, sizeof(double) * n);
  }
  if (sys) {
    if ((rc = ctrl_pull(c))) return rc;
    ctrl_to_planets(c->h_ctrl, sys);
  }
  return DFARGO_OK;
}

extern "C" int dfargo_algogas(dfargo_ctx *c, double DT, double *rho, double *vrad, double *vtheta,
                              double *energy, dfargo_planets *sys) {
  int rc = dfargo_upload_state(c, rho, vrad, vtheta, energy, sys);
  if (rc) return rc;
  if ((rc = dfargo_algogas_device(c, DT))) return rc;
  return dfargo_download_state(c, rho, vrad, vtheta, energy, sys);
}

// ------------------------------------------------------------------------
// queries

extern "C" int dfargo_get_status(dfargo_ctx *c, dfargo_status *st) {
  if (!c || !st) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(c->dev));
  int rc = ctrl_pull(c);
  if (rc) return rc;
  const dfargo_ctrl *k = c->h_ctrl;
  memset(st, 0, sizeof(*st));
  st->physical_time = k->time;
  st->physical_time_initial = k->time0;
  st->omega_frame = k->omega_frame;
  st->mindt = k->mindt;
  st->maxdt = k->maxdt;
  st->lost_mass = k->lost_mass;
  st->mass_taper = k->mass_taper;
  st->disk_on_primary_x = k->dop_x;
  st->disk_on_primary_y = k->dop_y;
  st->disk_on_primary_alt_x = k->dopalt_x;
  st->disk_on_primary_alt_y = k->dopalt_y;
  for (int q = 0; q < 3; q++) {
    st->bary_x[q] = k->barx[q];
    st->bary_y[q] = k->bary[q];
  }
  st->substeps_total = k->substeps_total;
  st->last_call_substeps = k->substeps;
  st->gas_timesteps_cfl = k->gas_timesteps_cfl;
  st->crashed_density = k->crashed_d;
  st->crashed_energy = k->crashed_e;
  st->crashed = k->crashed_e;
  {
    const unsigned long long b = k->fargo_margin_bits;
    double m;
    memcpy(&m, &b, sizeof m);
    st->fargo_margin = m;
  }
  return DFARGO_OK;
}

// ------------------------------------------------------------------------
// per-DT diagnostics

static int run_diagnostics(dfargo_ctx *c, const DiagBodies &b, std::vector<double> &out) {
  int rc;
  if (!c->diag_part) {
    if ((rc = dalloc(c, (void **)&c->diag_part, sizeof(double) * DIAG_NPART(DF_MAXPL) * (size_t)c->p.nrad)) ||
        (rc = dalloc(c, (void **)&c->diag_out, sizeof(double) * DIAG_NOUT(DF_MAXPL))))
      return rc;
  }
  if (c->prec == DFARGO_FP64) launch_diagnostics<double>(c->a, b, c->diag_part, c->diag_out, c->stream);
  else launch_diagnostics<float>(c->a, b, c->diag_part, c->diag_out, c->stream);
  HIPCHK(hipGetLastError());
  out.assign(DIAG_NOUT(b.n), 0.0);
  HIPCHK(hipMemcpyAsync(out.data(), c->diag_out, sizeof(double) * out.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return DFARGO_OK;
}

extern "C" int dfargo_gas_budgets(dfargo_ctx *c, double *out) {
  if (!c || !out) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  if (!c->initialized) return dfargo_set_error(DFARGO_ERR_STATE, "dfargo_init_euler has not been called");
  HIPCHK(hipSetDevice(c->dev));
  DiagBodies b;
  memset(&b, 0, sizeof b);
  std::vector<double> r;
  int rc = run_diagnostics(c, b, r);
  if (rc) return rc;
  if ((rc = ctrl_pull(c))) return rc;
  out[0] = r[0];
  out[1] = r[1];
  out[2] = r[2];
  // CheckMomentumConservation's planet term (src/SideEuler.c:196-206)
  const dfargo_ctrl *k = c->h_ctrl;
  double planetsmom = 0.0;
  for (int q = 0; q < k->nb; q++) {
    const double xp = k->pl_x[q], yp = k->pl_y[q];
    const double rpl = sqrt(xp * xp + yp * yp), thetapl = atan2(yp, xp);
    const double vazimpl = -k->pl_vx[q] * sin(thetapl) + k->pl_vy[q] * cos(thetapl);
    planetsmom += k->pl_m[q] * rpl * vazimpl;
  }
  out[3] = planetsmom;
  return DFARGO_OK;
}

// UpdateLog's bodies: the planets of the control block, with UpdateLog's
// smoothing and ComputeForce's Hill radius (src/Force.c:287-292, 118-119)
// on the host, in FP64
static void log_bodies(const dfargo_ctx *c, DiagBodies &b) {
  const dfargo_ctrl *k = c->h_ctrl;
  memset(&b, 0, sizeof b);
  b.n = k->nb;
  for (int q = 0; q < k->nb; q++) {
    const double x = k->pl_x[q], y = k->pl_y[q], m = k->pl_m[q], r = sqrt(x * x + y * y);
    b.x[q] = x;
    b.y[q] = y;
    b.a[q] = sqrt(x * x + y * y);
    b.rsm[q] = (c->p.roche_smoothing != 0.0)
                   ? r * pow(m / 3., 1. / 3.) * c->p.roche_smoothing
                   : c->p.thickness_smoothing * dfargo_aspect_ratio(&c->p, r) * pow(r, 1.0 + c->p.flaring_index);
    b.rh[q] = pow(m / 3., 1. / 3.) * b.a[q] + 1e-15;
  }
}

// the tqwk columns and indtq of each planet (src/Force.c:293-325) from the
// diagnostics' sums r
static void log_torques(const dfargo_ctx *c, const double *r, double *torques, double *indirect) {
  const dfargo_ctrl *k = c->h_ctrl;
  for (int q = 0; q < k->nb; q++) {
    const double x = k->pl_x[q], y = k->pl_y[q], vx = k->pl_vx[q], vy = k->pl_vy[q];
    const double *f = r + 3 + 8 * q; // fx i, fx ex i, fx o, fx ex o, fy i, fy ex i, fy o, fy ex o
    double *o = torques + 8 * q;
    o[0] = x * f[4] - y * f[0];
    o[1] = x * f[6] - y * f[2];
    o[2] = x * f[5] - y * f[1];
    o[3] = x * f[7] - y * f[3];
    o[4] = vx * f[0] + vy * f[4];
    o[5] = vx * f[2] + vy * f[6];
    o[6] = vx * f[1] + vy * f[5];
    o[7] = vx * f[3] + vy * f[7];
    indirect[q] = -x * k->dop_y + y * k->dop_x;
  }
}

extern "C" int dfargo_update_log(dfargo_ctx *c, int max_planets, double *torques, double *indirect, int *nb) {
  if (!c || !torques || !indirect || !nb) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  if (!c->initialized) return dfargo_set_error(DFARGO_ERR_STATE, "dfargo_init_euler has not been called");
  HIPCHK(hipSetDevice(c->dev));
  int rc = ctrl_pull(c);
  if (rc) return rc;
  const dfargo_ctrl *k = c->h_ctrl;
  if (k->nb > max_planets)
    return dfargo_set_error(DFARGO_ERR_ARG, "%d planets, room for %d", k->nb, max_planets);
  DiagBodies b;
  log_bodies(c, b);
  std::vector<double> r;
  if ((rc = run_diagnostics(c, b, r))) return rc;
  log_torques(c, r.data(), torques, indirect);
  *nb = k->nb;
  return DFARGO_OK;
}

// ------------------------------------------------------------------------
// UpdateLog without allocation
//
// The ring partials and sums of the diagnostics kernels live in one
// file-scope device array instead of buffers allocated on first use.  Each
// device holds its own copy until the process ends; host code reaches it
// only through the address hipGetSymbolAddress returns on that device
// (noalloc_scratch[dev], set when a context is created there).  Every call
// on a device shares that device's array: calls that overlap on one device
// race.  Each call writes every slot it reads (the ring partials of its
// rings and bodies, then the sums), so nothing needs clearing between calls.

#define DF_NOALLOC_MAX_NRAD 8192
#define DF_NOALLOC_DOUBLES ((size_t)DIAG_NPART(DF_MAXPL) * DF_NOALLOC_MAX_NRAD + DIAG_NOUT(DF_MAXPL))
static __device__ __align__(256) double scratch[DF_NOALLOC_DOUBLES];
#define DF_NOALLOC_MAX_DEV 64
static double *noalloc_scratch[DF_NOALLOC_MAX_DEV];
static std::mutex noalloc_mutex;

// the device address of this device's scratch array (the current device
// is `dev`); once per device
static int noalloc_setup(int dev) {
  if (dev < 0 || dev >= DF_NOALLOC_MAX_DEV)
    return dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "device %d beyond the %d the library tracks", dev,
                            DF_NOALLOC_MAX_DEV);
  std::lock_guard<std::mutex> lock(noalloc_mutex);
  if (noalloc_scratch[dev]) return DFARGO_OK;
  double *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return dfargo_set_error(DFARGO_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
                            hipGetErrorString(e));
  noalloc_scratch[dev] = ptr;
  return DFARGO_OK;
}

extern "C" int dfargo_update_log_noalloc(dfargo_ctx *c, int max_planets, double *torques, double *indirect,
                                         int *nb) {
  if (!c || !torques || !indirect || !nb) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  if (!c->initialized) return dfargo_set_error(DFARGO_ERR_STATE, "dfargo_init_euler has not been called");
  const int nr = c->p.nrad;
  if (nr > DF_NOALLOC_MAX_NRAD)
    return dfargo_set_error(DFARGO_ERR_ARG, "NRAD=%d exceeds the %d rings of the no-allocation scratch", nr,
                            DF_NOALLOC_MAX_NRAD);
  double *part = (c->dev >= 0 && c->dev < DF_NOALLOC_MAX_DEV) ? noalloc_scratch[c->dev] : nullptr;
  if (!part) return dfargo_set_error(DFARGO_ERR_STATE, "no scratch address for device %d", c->dev);
  HIPCHK(hipSetDevice(c->dev));
  int rc = ctrl_pull(c);
  if (rc) return rc;
  const dfargo_ctrl *k = c->h_ctrl;
  if (k->nb > max_planets)
    return dfargo_set_error(DFARGO_ERR_ARG, "%d planets, room for %d", k->nb, max_planets);
  DiagBodies b;
  log_bodies(c, b);
  // the partials of this disc's rings, then the sums
  double *out_d = part + (size_t)DIAG_NPART(DF_MAXPL) * nr;
  if (c->prec == DFARGO_FP64) launch_diagnostics<double>(c->a, b, part, out_d, c->stream);
  else launch_diagnostics<float>(c->a, b, part, out_d, c->stream);
  HIPCHK(hipGetLastError());
  double r[DIAG_NOUT(DF_MAXPL)];
  HIPCHK(hipMemcpyAsync(r, out_d, sizeof(double) * DIAG_NOUT(b.n), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  log_torques(c, r, torques, indirect);
  *nb = k->nb;
  return DFARGO_OK;
}

extern "C" int dfargo_set_time(dfargo_ctx *c, double t, double t0, double omega_frame) {
  if (!c) return dfargo_set_error(DFARGO_ERR_ARG, "null context");
  HIPCHK(hipSetDevice(c->dev));
  int rc = ctrl_pull(c);
  if (rc) return rc;
  c->h_ctrl->time = t;
  c->h_ctrl->time0 = t0;
  c->h_ctrl->omega_frame = omega_frame;
  return ctrl_push(c);
}

extern "C" int dfargo_get_field(dfargo_ctx *c, int field, double *out) {
  if (!c || !out) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(c->dev));
  const int nr = c->p.nrad, ns = c->p.nsec;
  const size_t n = (size_t)(nr + 1) * ns;
  KArgs &a = c->a;
  int rc;
  memset(out, 0, sizeof(double) * n);
  switch (field) {
    case DFARGO_FIELD_POTENTIAL:
      if (c->prec == DFARGO_FP64) launch_potential<double>(a, c->stream);
      else launch_potential<float>(a, c->stream);
      HIPCHK(hipGetLastError());
      return get_field_raw(c, out, a.pot, (size_t)nr * ns);
    case DFARGO_FIELD_SOUNDSPEED:
      if (c->p.energy_equation) return get_field_raw(c, out, a.cs, (size_t)nr * ns);
      for (int i = 0; i < nr; i++)
        for (int j = 0; j < ns; j++) out[(size_t)i * ns + j] = c->tab.cs_iso[i];
      return DFARGO_OK;
    case DFARGO_FIELD_PRESSURE:
    case DFARGO_FIELD_TEMPERATURE: {
      // from the current state (ComputePressureField / ComputeTemperatureField)
      std::vector<double> rho(n), e(n, 0.0);
      if ((rc = get_field_raw(c, rho.data(), a.rho, n))) return rc;
      if (c->p.energy_equation && (rc = get_field_raw(c, e.data(), a.en, n))) return rc;
      for (int i = 0; i < nr; i++)
        for (int j = 0; j < ns; j++) {
          size_t l = (size_t)i * ns + j;
          double pr = c->p.energy_equation ? (c->p.adiabatic_index - 1.0) * e[l]
                                           : rho[l] * c->tab.cs_iso[i] * c->tab.cs_iso[i];
          out[l] = (field == DFARGO_FIELD_PRESSURE) ? pr : 1.0 / 1.0 * pr / rho[l];
        }
      return DFARGO_OK;
    }
    case DFARGO_FIELD_DIVV: return get_field_raw(c, out, a.divv, (size_t)nr * ns);
    case DFARGO_FIELD_VISCHEAT: return get_field_raw(c, out, a.vh, (size_t)nr * ns);
    case DFARGO_FIELD_SG_ACCR:
      if (!c->p.self_gravity) return DFARGO_OK;
      return get_field_raw(c, out, a.sgr, (size_t)nr * ns);
    case DFARGO_FIELD_SG_ACCT:
      if (!c->p.self_gravity) return DFARGO_OK;
      return get_field_raw(c, out, a.sgt, (size_t)nr * ns);
    case DFARGO_FIELD_TAURR: return get_field_raw(c, out, a.trr, (size_t)nr * ns);
    case DFARGO_FIELD_TAURP: return get_field_raw(c, out, a.trp, (size_t)nr * ns);
    case DFARGO_FIELD_TAUPP: return get_field_raw(c, out, a.tpp, (size_t)nr * ns);
    case DFARGO_FIELD_DUST_DENSITY:
      if (!c->dust_on) return DFARGO_OK;
      return get_field_raw(c, out, a.dustdens, (size_t)nr * ns);
    case DFARGO_FIELD_FB_ACCR:
      if (!c->dust_on) return DFARGO_OK;
      return get_field_raw(c, out, a.radfbacc, (size_t)nr * ns);
    case DFARGO_FIELD_FB_ACCT:
      if (!c->dust_on) return DFARGO_OK;
      return get_field_raw(c, out, a.azifbacc, (size_t)nr * ns);
    case DFARGO_FIELD_FB_EDOT:
      if (!c->dust_on) return DFARGO_OK;
      return get_field_raw(c, out, a.fbedot, (size_t)nr * ns);
  }
  return dfargo_set_error(DFARGO_ERR_ARG, "unknown field %d", field);
}

extern "C" int dfargo_compute_selfgravity(dfargo_ctx *c) {
  if (!c) return dfargo_set_error(DFARGO_ERR_ARG, "null context");
  if (!c->p.self_gravity) return dfargo_set_error(DFARGO_ERR_STATE, "self-gravity is off for this context");
  HIPCHK(hipSetDevice(c->dev));
  if (c->prec == DFARGO_FP64) launch_sg_accel<double>(c->a, c->stream);
  else launch_sg_accel<float>(c->a, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return DFARGO_OK;
}

extern "C" double dfargo_fviscosity(dfargo_ctx *c, double rad) {
  if (!c) return 0.0;
  const dfargo_params &p = c->p;
  double viscosity = p.viscosity;
  if (p.alpha_viscosity != 0.0) {
    std::vector<double> g(p.nrad + 2, 0.0);
    hipSetDevice(c->dev);
    if (hipMemcpy(g.data(), c->a.gcs, sizeof(double) * (p.nrad + 2), hipMemcpyDeviceToHost) != hipSuccess)
      return 0.0;
    int i = 0;
    while (i < p.nrad - 1 && c->tab.Rmed[i] < rad) i++;
    viscosity = p.alpha_viscosity * g[i] * g[i] * pow(rad, 1.5);
  }
  double rmin = p.cavity_radius - p.cavity_width * p.aspect_ratio * pow(p.cavity_radius, 1.0 + p.flaring_index);
  double rmax = p.cavity_radius + p.cavity_width * p.aspect_ratio * pow(p.cavity_radius, 1.0 + p.flaring_index);
  if (rad < rmin) viscosity *= p.cavity_ratio;
  if ((rad >= rmin) && (rad <= rmax)) viscosity *= exp((rmax - rad) / (rmax - rmin) * log(p.cavity_ratio));
  return viscosity;
}

extern "C" int dfargo_get_fargo_shift(dfargo_ctx *c, long *nshift, double *vmed) {
  if (!c) return dfargo_set_error(DFARGO_ERR_ARG, "null context");
  HIPCHK(hipSetDevice(c->dev));
  if (nshift) HIPCHK(hipMemcpy(nshift, c->a.nshift, sizeof(long) * c->p.nrad, hipMemcpyDeviceToHost));
  if (vmed) HIPCHK(hipMemcpy(vmed, c->a.vmed, sizeof(double) * c->p.nrad, hipMemcpyDeviceToHost));
  return DFARGO_OK;
}

extern "C" int dfargo_get_visc1d(dfargo_ctx *c, double *rmed, double *sigma, double *vrad, int *n1d) {
  if (!c || !n1d) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  *n1d = c->a.damp_visc ? c->a.n1d : 0;
  if (!c->a.damp_visc) return DFARGO_OK;
  const size_t bytes = sizeof(double) * c->a.n1d;
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (rmed) HIPCHK(hipMemcpy(rmed, c->a.v1d_rmed, bytes, hipMemcpyDeviceToHost));
  if (sigma) HIPCHK(hipMemcpy(sigma, c->a.v1d_sigma, bytes, hipMemcpyDeviceToHost));
  if (vrad) HIPCHK(hipMemcpy(vrad, c->a.v1d_vrad, bytes, hipMemcpyDeviceToHost));
  return DFARGO_OK;
}

// The restart branch of InitializeOneDViscousEvolution
// (src/SideEuler.c:1489-1502) replaces Rmed1D and Sigma1D with what it
// reads; this takes any of the three profiles from the host.  The cell
// centres must increase, and under the wave-killing zones enclose every
// zone ring's Rmed, the bracket search of k_evanescent relies on both.
extern "C" int dfargo_set_visc1d(dfargo_ctx *c, const double *rmed, const double *sigma, const double *vrad,
                                 int n1d) {
  if (!c) return dfargo_set_error(DFARGO_ERR_ARG, "null context");
  if (!c->a.damp_visc) return dfargo_set_error(DFARGO_ERR_ARG, "the 1-D viscous disc is off (DAMPTOVISCOUS)");
  const int n = c->a.n1d;
  if (n1d != n) return dfargo_set_error(DFARGO_ERR_ARG, "the 1-D disc has NRAD1D = %d cells, got %d", n, n1d);
  for (int i = 0; i < n; i++)
    if ((rmed && !(rmed[i] > 0.0 && isfinite(rmed[i]))) || (sigma && !isfinite(sigma[i])) ||
        (vrad && !isfinite(vrad[i])))
      return dfargo_set_error(DFARGO_ERR_ARG, "1-D profile value %d is not finite (or Rmed1D <= 0)", i);
  if (rmed) {
    for (int i = 1; i < n; i++)
      if (!(rmed[i] > rmed[i - 1])) return dfargo_set_error(DFARGO_ERR_ARG, "Rmed1D must increase (cell %d)", i);
    if (c->p.wave_killing_zones)
      for (int i = 0; i < c->p.nrad; i++)
        if (!(rmed[0] < c->tab.Rmed[i] && c->tab.Rmed[i] <= rmed[n - 1]))
          return dfargo_set_error(DFARGO_ERR_ARG, "the 1-D grid's cell centres must enclose every Rmed");
  }
  const size_t bytes = sizeof(double) * n;
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (rmed) HIPCHK(hipMemcpy((void *)c->a.v1d_rmed, rmed, bytes, hipMemcpyHostToDevice));
  if (sigma) HIPCHK(hipMemcpy(c->a.v1d_sigma, sigma, bytes, hipMemcpyHostToDevice));
  if (vrad) HIPCHK(hipMemcpy(c->a.v1d_vrad, vrad, bytes, hipMemcpyHostToDevice));
  return DFARGO_OK;
}

extern "C" int dfargo_set_profiling(dfargo_ctx *c, int enable) {
  if (!c) return dfargo_set_error(DFARGO_ERR_ARG, "null context");
  c->profiling = enable ? 1 : 0;
  memset(c->kt_ms, 0, sizeof c->kt_ms);
  memset(c->kt_n, 0, sizeof c->kt_n);
  return DFARGO_OK;
}

extern "C" int dfargo_get_kernel_times(dfargo_ctx *c, int max, char (*names)[48], double *total_ms, long *launches,
                                       int *count) {
  if (!c || !count) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  int n = 0;
  for (int id = 0; id < KID_COUNT && n < max; id++) {
    if (c->kt_n[id] == 0) continue;
    if (names) {
      strncpy(names[n], kKernelNames[id], 47);
      names[n][47] = 0;
    }
    if (total_ms) total_ms[n] = c->kt_ms[id];
    if (launches) launches[n] = c->kt_n[id];
    n++;
  }
  *count = n;
  return DFARGO_OK;
}

extern "C" void *dfargo_stream(dfargo_ctx *c) { return c ? (void *)c->stream : nullptr; }

// ------------------------------------------------------------------------
// dust particles

static int put_doubles(dfargo_ctx *c, double *dst, const double *src, size_t n) {
  if (!src || n == 0) return DFARGO_OK;
  HIPCHK(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return DFARGO_OK;
}
static int get_doubles(dfargo_ctx *c, double *dst, const double *src, size_t n) {
  if (!dst || n == 0) return DFARGO_OK;
  HIPCHK(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return DFARGO_OK;
}

// particle arrays of the caller (index = particle id) into a set, ids in
// order; a NULL array leaves that field as it is
static int put_dust_set(dfargo_ctx *c, const DustSet &P, const dfargo_dust_state *st, size_t np) {
  int rc;
  if ((rc = put_doubles(c, P.r, st->r, np)) || (rc = put_doubles(c, P.th, st->th, np)) ||
      (rc = put_doubles(c, P.vr, st->vr, np)) || (rc = put_doubles(c, P.vth, st->vth, np)) ||
      (rc = put_doubles(c, P.l, st->l, np)) || (rc = put_doubles(c, P.size, st->dustsize, np)))
    return rc;
  std::vector<int> ids(np);
  for (size_t k = 0; k < np; k++) ids[k] = (int)k;
  HIPCHK(hipMemcpy(P.id, ids.data(), sizeof(int) * np, hipMemcpyHostToDevice));
  return DFARGO_OK;
}

extern "C" int dfargo_attach_dust(dfargo_ctx *c, const dfargo_dust_params *d, const dfargo_dust_state *st,
                                  double timestep) {
  if (!c || !d || !st) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  if (!c->initialized) return dfargo_set_error(DFARGO_ERR_STATE, "dfargo_init_euler has not been called");
  if (c->dust_on) return dfargo_set_error(DFARGO_ERR_STATE, "particles are already attached to this context");
  if (c->a.df_on && d->nbpart > 0)
    return dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "dust particles beside the dust fluid are not implemented");
  int rc = dfargo_check_dust_params(&c->p, d);
  if (rc) return rc;
  if (c->a.accrete)
    return dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "planet accretion with particles is not implemented");
  const int n = d->nbpart;
  if (n == 0) return DFARGO_OK; // NBPART = 0: no particles, nothing changes
  if (!st->r || !st->th || !st->vr || !st->vth || !st->l || !st->dustsize)
    return dfargo_set_error(DFARGO_ERR_ARG, "the dust state needs r, th, vr, vth, l and dustsize");
  if (!(timestep > 0.0)) return dfargo_set_error(DFARGO_ERR_ARG, "the start-up time step must be positive");
  const dfargo_params &p = c->p;
  const int nr = p.nrad, ns = p.nsec;
  if (ns < 4) return dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "particles need NSEC >= 4");
  HIPCHK(hipSetDevice(c->dev));
  KArgs &a = c->a;
  const dfargo_tables &t = c->tab;
  const size_t np = (size_t)n, ncell = (size_t)nr * ns;
  for (int set = 0; set < 2; set++) {
    DustSet &P = a.pset[set];
    double **pd[] = {&P.r, &P.th, &P.vr, &P.vth, &P.l, &P.size, &P.stokes};
    for (double **q : pd)
      if ((rc = dalloc(c, (void **)q, sizeof(double) * np))) return rc;
    if ((rc = dalloc(c, (void **)&P.id, sizeof(int) * np))) return rc;
  }
  double **pd[] = {&a.p_raddrag, &a.p_azidrag, &a.p_ts, &a.p_dv2};
  for (double **q : pd)
    if ((rc = dalloc(c, (void **)q, sizeof(double) * np))) return rc;
  if ((rc = dalloc(c, (void **)&a.p_sorted, sizeof(unsigned long long) * np))) return rc;
  if ((rc = dalloc(c, (void **)&a.p_w, sizeof(double) * 18 * np)) ||
      (rc = dalloc(c, (void **)&a.p_tsmin, sizeof(unsigned long long) * ((np + 255) / 256))))
    return rc;
  int **pi[] = {&a.p_cell, &a.p_slot, &a.p_base, &a.out_list};
  for (int **q : pi)
    if ((rc = dalloc(c, (void **)q, sizeof(int) * np))) return rc;
  int **ci[] = {&a.cell_count, &a.cell_local, &a.cell_start};
  for (int **q : ci)
    if ((rc = dalloc(c, (void **)q, sizeof(int) * ncell))) return rc;
  if ((rc = dalloc(c, (void **)&a.ring_tot, sizeof(int) * (size_t)(nr + 1))) ||
      (rc = dalloc(c, (void **)&a.ring_off, sizeof(int) * (size_t)(nr + 1))))
    return rc;
  void **grids[] = {&a.radgradp, &a.azigradp, &a.radindacc, &a.aziindacc, &a.radsgacc, &a.azisgacc,
                    &a.radfbacc, &a.azifbacc, &a.fbedot, &a.dustdens, &a.cs_old};
  for (void **g : grids)
    if ((rc = dalloc(c, g, c->esz * c->nfield))) return rc;
  if ((rc = upload_table(c, &a.Azimuth, t.Azimuth, (size_t)ns))) return rc;
  // particle state into set 0, in the caller's order
  if ((rc = put_dust_set(c, a.pset[0], st, np))) return rc;
  c->dustsize_init.assign(st->dustsize_init ? st->dustsize_init : st->dustsize,
                          (st->dustsize_init ? st->dustsize_init : st->dustsize) + np);
  // switches and constants (ComputeCodeUnits, src/LowTasks.c:204-212)
  a.d_on = 1;
  a.d_np = n;
  a.d_interp = d->interpolation;
  a.d_feel_disk = d->dust_feel_disk;
  a.d_feel_sg = d->dust_feel_sg;
  a.d_feel_planets = d->dust_feel_planets;
  a.d_feedback = d->dust_feedback;
  a.d_sfta = d->short_friction_time_approximation;
  a.d_fixed_dt = !d->no_dt_constraint_by_particles;
  a.log_grid = p.log_grid;
  a.d_dust_timestep = d->dust_timestep;
  a.d_unit_mass = 1.9891e30 * d->factor_unit_mass;
  const double unit_length = 1.49598e11 * d->factor_unit_length;
  a.d_len_m2 = pow(unit_length, -2.0);
  a.d_rho_code = d->rhopart * 1000.0 * pow(unit_length, 3.) / a.d_unit_mass;
  const double half = 0.5 * (p.pmax - p.pmin) / (double)ns;
  a.aziinf0 = t.Azimuth[0] - half;
  a.azisup_last = t.Azimuth[ns - 1] + half;
  a.delta_phi = (t.Azimuth[0] + half) - (t.Azimuth[0] - half);
  a.delta_r = p.log_grid ? log(t.Rsup[0] / t.Rinf[0]) : t.Rsup[0] - t.Rinf[0];
  a.rmin = p.rmin;
  a.rmax = p.rmax;
  a.d_ring_step = p.log_grid ? log(t.Rinf[1] / t.Rinf[0]) : t.Rinf[1] - t.Rinf[0];
  // control block: super-particle mass and a fresh minimum stopping time
  if ((rc = ctrl_pull(c))) return rc;
  c->h_ctrl->particles_mass = d->particle_mass;
  c->h_ctrl->particles_mass0 = d->particle_mass;
  c->h_ctrl->dust_mp = d->particle_mass;
  c->h_ctrl->pset = 0;
  c->h_ctrl->out_count = 0;
  if ((rc = ctrl_push(c))) return rc;
  // start-up interpolation (src/main.c:287)
  if (c->prec == DFARGO_FP64) launch_dust_startup<double>(a, timestep, c->h_ctrl->omega_frame, c->stream);
  else launch_dust_startup<float>(a, timestep, c->h_ctrl->omega_frame, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  c->dp = *d;
  c->dust_on = 1;
  // the sub-step now holds the particle kernels: capture it again
  if ((rc = drop_graph(c))) return rc;
  build_sequence(c);
  return DFARGO_OK;
}

// ------------------------------------------------------------------------
// the dust fluid

// DAspectRatio (src/Viscosity.c:123-138), LAMBDADOUBLING = 0
static double df_aspect_ratio(const dfargo_params &p, double dar, double rad) {
  double ar = dar;
  const double rmin = p.transition_radius - p.transition_width * dar;
  const double rmax = p.transition_radius + p.transition_width * dar;
  if (rad < rmin) ar *= p.transition_ratio;
  if ((rad >= rmin) && (rad <= rmax)) ar *= exp((rmax - rad) / (rmax - rmin) * log(p.transition_ratio));
  return ar;
}

extern "C" int dfargo_attach_dust_fluid(dfargo_ctx *c, const dfargo_dust_fluid_params *f, const double *drho) {
  if (!c || !f || !drho) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  if (c->a.df_on) return dfargo_set_error(DFARGO_ERR_STATE, "the dust fluid is already attached to this context");
  if (c->dust_on)
    return dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "dust particles beside the dust fluid are not implemented");
  int rc = dfargo_check_dust_fluid_params(&c->p, f, c->prec);
  if (rc) return rc;
  const dfargo_params &p = c->p;
  const dfargo_tables &t = c->tab;
  const int nr = p.nrad, ns = p.nsec;
  const size_t n = (size_t)(nr + 1) * ns;
  for (size_t l = 0; l < (size_t)nr * ns; l++)
    if (!(drho[l] > 0.0) || !std::isfinite(drho[l]))
      return dfargo_set_error(DFARGO_ERR_ARG, "the dust density must be positive (cell %zu)", l);
  HIPCHK(hipSetDevice(c->dev));
  KArgs &a = c->a;
  double **grids[] = {&a.drho, &a.dvr, &a.dvt, &a.dstokes, &a.dvri, &a.dvti, &a.dtrr, &a.dtrp, &a.dtpp,
                      &a.dt_rho, &a.dt_rp, &a.dt_rm, &a.dt_tp, &a.dt_tm, &a.dz_rho, &a.dz_rp, &a.dz_rm,
                      &a.dz_tp, &a.dz_tm};
  for (double **g : grids)
    if ((rc = dalloc(c, (void **)g, sizeof(double) * c->nfield))) return rc;
  const int ntile = (ns + DF_SRC_COLS - 1) / DF_SRC_COLS;
  if ((rc = dalloc(c, (void **)&a.dvmoy, sizeof(double) * (nr + 2))) ||
      (rc = dalloc(c, (void **)&a.dvmed, sizeof(double) * (nr + 2))) ||
      (rc = dalloc(c, (void **)&a.df_axisg, sizeof(double) * (nr + 2))) ||
      (rc = dalloc(c, (void **)&a.df_ring_dt, sizeof(double) * (nr + 2))) ||
      (rc = dalloc(c, (void **)&a.dnshift, sizeof(long) * (nr + 2))) ||
      (rc = dalloc(c, (void **)&a.dvt_part, sizeof(double) * (size_t)nr * ntile)) ||
      (rc = dalloc(c, (void **)&a.dlost_col, sizeof(double) * ns)))
    return rc;
  // ring tables: the dust's sound speed (ComputeDustSoundSpeedAndPressure,
  // src/SourceEuler.c:2353-2371) and its azimuthal mean summed cell by cell
  // (GLOBAL_DustSoundSpeed), DFViscosity at Rmed (src/Viscosity.c:72-104),
  // pow(Rsup, 1.5) and the cavity factor of FViscosity at Rsup for Diffd,
  // DSigmaMed (FillDSigma, src/Theo.c:95-101)
  const int n1 = nr + 2;
  std::vector<double> dcs(n1, 0.0), gdcs(n1, 0.0), dnu(n1, 0.0), sup15(n1, 0.0), cavsup(n1, 1.0), dsig(n1, 0.0);
  for (int i = 0; i < nr; i++) {
    const double r = t.Rmed[i];
    dcs[i] = df_aspect_ratio(p, f->daspect_ratio, r) * sqrt(1.0 * 1.0 / r) * pow(r, f->dflaring_index);
    double sum = 0.0;
    for (int j = 0; j < ns; j++) sum += dcs[i];
    gdcs[i] = sum / (double)ns;
    dsig[i] = dfargo_sigma_profile(&p, r) * f->dust_to_gas_density_ratio;
  }
  const bool dvisc = (f->dalpha_viscosity != 0.0) || (f->dviscosity != 0.0);
  for (int i = 0; i < nr; i++) {
    const double rad = t.Rmed[i];
    double v = f->dviscosity;
    if (f->dalpha_viscosity != 0.0) {
      int k = 0;
      while (t.Rmed[k] < rad) k++;
      v = f->dalpha_viscosity * gdcs[k] * gdcs[k] * pow(rad, 1.5);
    }
    const double rmin = p.cavity_radius - p.cavity_width * f->daspect_ratio;
    const double rmax = p.cavity_radius + p.cavity_width * f->daspect_ratio;
    if (rad < rmin) v *= p.cavity_ratio;
    if ((rad >= rmin) && (rad <= rmax)) v *= exp((rmax - rad) / (rmax - rmin) * log(p.cavity_ratio));
    dnu[i] = dvisc ? v : 0.0;
  }
  for (int i = 0; i < nr; i++) {
    const double rad = t.Rsup[i];
    sup15[i] = pow(rad, 1.5);
    const double rmin = p.cavity_radius - p.cavity_width * p.aspect_ratio * pow(p.cavity_radius, 1.0 + p.flaring_index);
    const double rmax = p.cavity_radius + p.cavity_width * p.aspect_ratio * pow(p.cavity_radius, 1.0 + p.flaring_index);
    if (rad < rmin) cavsup[i] = p.cavity_ratio;
    if ((rad >= rmin) && (rad <= rmax)) cavsup[i] = exp((rmax - rad) / (rmax - rmin) * log(p.cavity_ratio));
  }
  if ((rc = upload_table(c, &a.df_cs, dcs.data(), n1)) || (rc = upload_table(c, &a.df_nu, dnu.data(), n1)) ||
      (rc = upload_table(c, &a.df_sup15, sup15.data(), n1)) ||
      (rc = upload_table(c, &a.df_cavsup, cavsup.data(), n1)) ||
      (rc = upload_table(c, &a.DSigmaMed, dsig.data(), n1)) ||
      (rc = upload_table(c, &a.DVthetaMed, dsig.data(), n1)) || (rc = upload_table(c, &a.DVradMed, dsig.data(), n1)))
    return rc;
  // the grain in code units (src/SourceEuler.c:812-818; ComputeCodeUnits,
  // src/LowTasks.c:204-212)
  const double um = 1.9891e30 * f->factor_unit_mass, ul = 1.49598e11 * f->factor_unit_length;
  a.df_rho_int = f->rhopart * 1000.0 * pow(ul, 3.) / um;
  a.df_radius = f->sizepart / ul;
  a.df_to_cgs = 0.1 * um * pow(ul, -2.0);
  a.df_diff_inn = f->dust_diff_inn;
  a.df_sfta = f->sfta ? 1 : 0;
  a.df_feel_disk = f->dust_feel_disk ? 1 : 0;
  a.df_feedback = f->dust_feedback ? 1 : 0;
  a.df_feel_sg = f->dust_feel_sg;
  a.df_diffusion = f->dust_diffusion ? 1 : 0;
  a.df_visc = dvisc ? 1 : 0;
  a.df_open = f->open_inner_dust ? 1 : 0;
  c->fp = *f;
  c->df_drho0.assign(drho, drho + n);
  for (int j = 0; j < ns; j++) c->df_drho0[(size_t)nr * ns + j] = 0.0;
  if ((rc = put_doubles(c, a.drho, c->df_drho0.data(), n))) return rc;
  a.df_on = 1;
  if ((rc = drop_graph(c))) return rc;
  build_sequence(c);
  return DFARGO_OK;
}

extern "C" int dfargo_download_dust_fluid(dfargo_ctx *c, double *drho, double *dvrad, double *dvtheta,
                                          double *stokes) {
  if (!c) return dfargo_set_error(DFARGO_ERR_ARG, "null context");
  if (!c->a.df_on) return dfargo_set_error(DFARGO_ERR_STATE, "no dust fluid is attached to this context");
  HIPCHK(hipSetDevice(c->dev));
  const size_t n = (size_t)(c->p.nrad + 1) * c->p.nsec;
  int rc;
  if ((rc = get_doubles(c, drho, c->a.drho, n)) || (rc = get_doubles(c, dvrad, c->a.dvr, n)) ||
      (rc = get_doubles(c, dvtheta, c->a.dvt, n)) || (rc = get_doubles(c, stokes, c->a.dstokes, n)))
    return rc;
  return DFARGO_OK;
}

extern "C" int dfargo_upload_dust_fluid(dfargo_ctx *c, const double *drho, const double *dvrad,
                                        const double *dvtheta) {
  if (!c) return dfargo_set_error(DFARGO_ERR_ARG, "null context");
  if (!c->a.df_on) return dfargo_set_error(DFARGO_ERR_STATE, "no dust fluid is attached to this context");
  HIPCHK(hipSetDevice(c->dev));
  const size_t n = (size_t)(c->p.nrad + 1) * c->p.nsec;
  int rc;
  if ((rc = put_doubles(c, c->a.drho, drho, n)) || (rc = put_doubles(c, c->a.dvr, dvrad, n)) ||
      (rc = put_doubles(c, c->a.dvt, dvtheta, n)))
    return rc;
  return DFARGO_OK;
}

extern "C" int dfargo_download_dust(dfargo_ctx *c, dfargo_dust_state *st) {
  if (!c || !st) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  if (!c->dust_on) return dfargo_set_error(DFARGO_ERR_STATE, "no particles are attached to this context");
  HIPCHK(hipSetDevice(c->dev));
  int rc = ctrl_pull(c);
  if (rc) return rc;
  const DustSet &P = c->a.pset[c->h_ctrl->pset];
  const size_t np = (size_t)c->a.d_np;
  // the device holds the particles sorted by cell: back to the caller's order
  std::vector<int> ids(np);
  HIPCHK(hipMemcpy(ids.data(), P.id, sizeof(int) * np, hipMemcpyDeviceToHost));
  std::vector<double> buf(np);
  const double *src[] = {P.r, P.th, P.vr, P.vth, P.l, P.size, P.stokes};
  double *dst[] = {st->r, st->th, st->vr, st->vth, st->l, st->dustsize, st->stokesnb};
  for (int f = 0; f < 7; f++) {
    if (!dst[f]) continue;
    if ((rc = get_doubles(c, buf.data(), src[f], np))) return rc;
    for (size_t s = 0; s < np; s++) dst[f][ids[s]] = buf[s];
  }
  if (st->dustsize_init) memcpy(st->dustsize_init, c->dustsize_init.data(), sizeof(double) * np);
  return DFARGO_OK;
}

extern "C" int dfargo_upload_dust(dfargo_ctx *c, const dfargo_dust_state *st) {
  if (!c || !st) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  if (!c->dust_on) return dfargo_set_error(DFARGO_ERR_STATE, "no particles are attached to this context");
  HIPCHK(hipSetDevice(c->dev));
  const size_t np = (size_t)c->a.d_np;
  // the current state in the caller's order, overlaid with what is given,
  // written back in that order (the Stokes numbers of the last
  // interpolation follow their particles)
  dfargo_dust_state cur;
  std::vector<double> r(np), th(np), vr(np), vth(np), l(np), sz(np), stk(np);
  cur.r = r.data(); cur.th = th.data(); cur.vr = vr.data(); cur.vth = vth.data(); cur.l = l.data();
  cur.dustsize = sz.data(); cur.stokesnb = stk.data(); cur.dustsize_init = nullptr;
  int rc = dfargo_download_dust(c, &cur);
  if (rc) return rc;
  if (st->r) r.assign(st->r, st->r + np);
  if (st->th) th.assign(st->th, st->th + np);
  if (st->vr) vr.assign(st->vr, st->vr + np);
  if (st->vth) vth.assign(st->vth, st->vth + np);
  if (st->l) l.assign(st->l, st->l + np);
  if (st->dustsize) sz.assign(st->dustsize, st->dustsize + np);
  const DustSet &P = c->a.pset[c->h_ctrl->pset];
  if ((rc = put_dust_set(c, P, &cur, np)) || (rc = put_doubles(c, P.stokes, stk.data(), np))) return rc;
  if (st->dustsize_init) c->dustsize_init.assign(st->dustsize_init, st->dustsize_init + np);
  return DFARGO_OK;
}

// ------------------------------------------------------------------------
// launch-shape tuning

extern "C" int dfargo_get_tuning(dfargo_ctx *c, dfargo_tuning *t) {
  if (!c || !t) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  t->xcd_remap = c->a.xcd_remap;
  t->azi_waves = c->a.tune_azi_waves;
  t->finish_block = c->a.tune_finish_block;
  t->sg_fft_buffers = c->a.sg_fft_buffers;
  t->sg_rows_pair = c->a.sg_rows_pair;
  return DFARGO_OK;
}

extern "C" int dfargo_set_tuning(dfargo_ctx *c, const dfargo_tuning *t) {
  if (!c || !t) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  if ((t->xcd_remap != 0 && t->xcd_remap != 1) || (t->azi_waves != 1 && t->azi_waves != 2 && t->azi_waves != 4) ||
      (t->finish_block != 64 && t->finish_block != 128 && t->finish_block != 256) || t->sg_fft_buffers < 0 ||
      t->sg_fft_buffers > 2 || t->sg_rows_pair < 0 || t->sg_rows_pair > 2)
    return dfargo_set_error(DFARGO_ERR_ARG, "tuning out of range (xcd_remap 0/1, azi_waves 1/2/4, "
                                            "finish_block 64/128/256, sg_fft_buffers 0/1/2, sg_rows_pair 0/1/2)");
  if (t->xcd_remap == c->a.xcd_remap && t->azi_waves == c->a.tune_azi_waves &&
      t->finish_block == c->a.tune_finish_block && t->sg_fft_buffers == c->a.sg_fft_buffers &&
      t->sg_rows_pair == c->a.sg_rows_pair)
    return DFARGO_OK;
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->a.xcd_remap = t->xcd_remap;
  c->a.tune_azi_waves = t->azi_waves;
  c->a.tune_finish_block = t->finish_block;
  c->a.sg_fft_buffers = t->sg_fft_buffers;
  c->a.sg_rows_pair = t->sg_rows_pair;
  return drop_graph(c);
}

// best launch shapes found per disc shape and switches, for the process
typedef std::tuple<int, int, int, int, int, int, int> TuneKey;
static std::map<TuneKey, dfargo_tuning> tune_cache;
static std::mutex tune_mutex;

static TuneKey tune_key(const dfargo_ctx *c) {
  return TuneKey(c->p.nrad, c->p.nsec, c->prec, c->p.energy_equation, c->p.self_gravity, c->dust_on, c->dev);
}

// the average time of one launch of kernel `id` in the scratch disc's
// last profiled sub-steps
static double kernel_avg_ms(const dfargo_ctx *s, int id) { return s->kt_n[id] ? s->kt_ms[id] / s->kt_n[id] : 0.0; }

extern "C" int dfargo_autotune(dfargo_ctx *c, int substeps, dfargo_tuning *best_out) {
  if (!c) return dfargo_set_error(DFARGO_ERR_ARG, "null context");
  if (!c->initialized) return dfargo_set_error(DFARGO_ERR_STATE, "dfargo_init_euler has not been called");
  if (substeps < 1 || substeps > 1000) return dfargo_set_error(DFARGO_ERR_ARG, "substeps must lie in [1, 1000]");
  int rc;
  {
    std::lock_guard<std::mutex> lock(tune_mutex);
    auto it = tune_cache.find(tune_key(c));
    if (it != tune_cache.end()) {
      if (best_out) *best_out = it->second;
      return dfargo_set_tuning(c, &it->second);
    }
  }
  HIPCHK(hipSetDevice(c->dev));
  // a scratch disc of the same parameters from the initial state, so the
  // trials leave this disc's state alone
  const size_t n = (size_t)(c->p.nrad + 1) * c->p.nsec;
  std::vector<double> rho(n, 0.0), en(n, 0.0), vr(n), vt(n);
  if ((rc = dfargo_init_gas_density(&c->p, rho.data()))) return rc;
  if (c->p.energy_equation && (rc = dfargo_init_gas_energy(&c->p, en.data()))) return rc;
  if ((rc = ctrl_pull(c))) return rc;
  dfargo_planets pl;
  memset(&pl, 0, sizeof pl);
  ctrl_to_planets(c->h_ctrl, &pl);
  dfargo_ctx *s = nullptr;
  if ((rc = dfargo_create(&c->p, c->prec, c->dev, &s))) return rc;
  const double DT = 2.0 * M_PI / 20.0;
  dfargo_tuning cur, best;
  dfargo_get_tuning(c, &best);
  // one knob at a time (each shapes different kernels), the others at the
  // best values found so far; a warm-up pass before each timed one
  auto trial = [&](const dfargo_tuning &t, double *ms) -> int {
    int r;
    if ((r = dfargo_set_tuning(s, &t)) || (r = dfargo_set_profiling(s, 1)) ||
        (r = dfargo_run_substeps(s, DT, substeps)) || (r = dfargo_set_profiling(s, 1)) ||
        (r = dfargo_run_substeps(s, DT, substeps)))
      return r;
    ms[0] = kernel_avg_ms(s, KID_CFL) + kernel_avg_ms(s, KID_ENERGY) + kernel_avg_ms(s, KID_EOS_STEP);
    ms[1] = kernel_avg_ms(s, KID_TRANSPORT_AZI);
    ms[2] = kernel_avg_ms(s, KID_FINISH);
    // the transforms run beside the CFL branch: their knob is judged by the
    // captured sub-step (a warm-up call instantiates the graph)
    ms[3] = ms[4] = 0.0;
    if (t.sg_fft_buffers || t.sg_rows_pair) {
      if ((r = dfargo_set_profiling(s, 0)) || (r = dfargo_run_substeps(s, DT, substeps))) return r;
      const auto t0 = std::chrono::steady_clock::now();
      if ((r = dfargo_run_substeps(s, DT, substeps))) return r;
      ms[3] = ms[4] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / substeps;
    }
    return DFARGO_OK;
  };
  rc = dfargo_init_euler(s, rho.data(), vr.data(), vt.data(), c->p.energy_equation ? en.data() : nullptr, &pl);
  const int xcd[] = {1, 0}, waves[] = {4, 2, 1}, blocks[] = {256, 128, 64}, bufs[] = {2, 1}, pairs[] = {2, 1};
  // the transforms' knobs only with the full self-gravity
  const int nknobs = (c->p.self_gravity && !c->p.sg_zero_mode) ? 5 : 3;
  for (int knob = 0; knob < nknobs && !rc; knob++) {
    double best_ms = 1e30;
    const int nv = (knob == 0 || knob >= 3) ? 2 : 3;
    for (int v = 0; v < nv && !rc; v++) {
      cur = best;
      if (knob == 0) cur.xcd_remap = xcd[v];
      else if (knob == 1) cur.azi_waves = waves[v];
      else if (knob == 2) cur.finish_block = blocks[v];
      else if (knob == 3) cur.sg_fft_buffers = bufs[v];
      else cur.sg_rows_pair = pairs[v];
      double ms[5];
      if ((rc = trial(cur, ms))) break;
      if (ms[knob] < best_ms) {
        best_ms = ms[knob];
        if (knob == 0) best.xcd_remap = cur.xcd_remap;
        else if (knob == 1) best.azi_waves = cur.azi_waves;
        else if (knob == 2) best.finish_block = cur.finish_block;
        else if (knob == 3) best.sg_fft_buffers = cur.sg_fft_buffers;
        else best.sg_rows_pair = cur.sg_rows_pair;
      }
    }
  }
  dfargo_destroy(s);
  HIPCHK(hipSetDevice(c->dev));
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lock(tune_mutex);
    tune_cache[tune_key(c)] = best;
  }
  if (best_out) *best_out = best;
  return dfargo_set_tuning(c, &best);
}

// ------------------------------------------------------------------------
// batched discs: one graph holds every disc's sub-step as a child graph,
// without dependencies between discs; each launch of it advances every
// disc whose AlgoGas call is still running by one sub-step (the discs'
// kernels read their own device loop condition, so a finished disc's
// kernels return at once)

struct dfargo_batch {
  int mode = DFARGO_BATCH_STREAMS;
  std::vector<dfargo_ctx *> d;
  std::vector<long> version; // each disc's graph_version at capture
  std::vector<double> DT;
  int dev = 0;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
};

extern "C" int dfargo_batch_destroy(dfargo_batch *b) {
  if (!b) return DFARGO_OK;
  hipSetDevice(b->dev);
  if (b->stream) hipStreamSynchronize(b->stream);
  if (b->gexec) hipGraphExecDestroy(b->gexec);
  if (b->graph) hipGraphDestroy(b->graph);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return DFARGO_OK;
}

extern "C" int dfargo_batch_create(dfargo_ctx *const *discs, int n, dfargo_batch **out) {
  if (!discs || !out || n < 1 || n > 4096) return dfargo_set_error(DFARGO_ERR_ARG, "need 1..4096 discs");
  *out = nullptr;
  for (int i = 0; i < n; i++) {
    if (!discs[i]) return dfargo_set_error(DFARGO_ERR_ARG, "disc %d is null", i);
    if (!discs[i]->initialized)
      return dfargo_set_error(DFARGO_ERR_STATE, "disc %d: dfargo_init_euler has not been called", i);
    if (discs[i]->dev != discs[0]->dev) return dfargo_set_error(DFARGO_ERR_ARG, "the discs must share a device");
    if (discs[i]->profiling) return dfargo_set_error(DFARGO_ERR_STATE, "disc %d has profiling on", i);
    for (int j = 0; j < i; j++)
      if (discs[j] == discs[i]) return dfargo_set_error(DFARGO_ERR_ARG, "disc %d appears twice", i);
  }
  dfargo_batch *b = new dfargo_batch();
  b->dev = discs[0]->dev;
  b->d.assign(discs, discs + n);
  b->DT.assign(n, 0.0);
  int rc = DFARGO_OK;
  hipError_t e = hipSetDevice(b->dev);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipGraphCreate(&b->graph, 0);
  for (int i = 0; i < n && e == hipSuccess && !rc; i++) {
    dfargo_ctx *c = b->d[i];
    if ((rc = ensure_graph(c))) break;
    b->version.push_back(c->graph_version);
    hipGraphNode_t node;
    e = hipGraphAddChildGraphNode(&node, b->graph, nullptr, 0, c->graph);
  }
  if (e == hipSuccess && !rc) e = hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0);
  if (e != hipSuccess && !rc) rc = dfargo_set_error(DFARGO_ERR_HIP, "batch graph: %s", hipGetErrorString(e));
  if (rc) {
    dfargo_batch_destroy(b);
    return rc;
  }
  *out = b;
  return DFARGO_OK;
}

extern "C" int dfargo_batch_set_mode(dfargo_batch *b, int mode) {
  if (!b) return dfargo_set_error(DFARGO_ERR_ARG, "null batch");
  if (mode != DFARGO_BATCH_GRAPH && mode != DFARGO_BATCH_STREAMS)
    return dfargo_set_error(DFARGO_ERR_ARG, "unknown batch mode %d", mode);
  b->mode = mode;
  return DFARGO_OK;
}

// each disc's sub-step graph on the disc's own stream, as many times as the
// disc's own estimate of its remaining sub-steps, the discs interleaved so
// that consecutive launches go to different streams (and hardware queues);
// a finished disc launches nothing
static int batch_streams(dfargo_batch *b) {
  const int n = (int)b->d.size();
  for (int i = 0; i < n; i++) {
    dfargo_ctx *c = b->d[i];
    call_start(c, b->DT[i], c->stream);
  }
  HIPCHK(hipGetLastError());
  std::vector<int> rem(n, 1);
  long launched = 0;
  const long cap = 50000000L;
  for (;;) {
    int most = 0;
    for (int i = 0; i < n; i++) most = std::max(most, rem[i]);
    if (most == 0) break;
    for (int r = 0; r < most; r++)
      for (int i = 0; i < n; i++)
        if (rem[i] > r) HIPCHK(hipGraphLaunch(b->d[i]->gexec, b->d[i]->stream));
    launched += most;
    for (int i = 0; i < n; i++)
      if (rem[i] > 0)
        HIPCHK(hipMemcpyAsync(b->d[i]->h_ctrl, b->d[i]->d_ctrl, sizeof(dfargo_ctrl), hipMemcpyDeviceToHost,
                              b->d[i]->stream));
    for (int i = 0; i < n; i++)
      if (rem[i] > 0) HIPCHK(hipStreamSynchronize(b->d[i]->stream));
    for (int i = 0; i < n; i++) {
      if (rem[i] == 0) continue;
      const dfargo_ctrl *k = b->d[i]->h_ctrl;
      rem[i] = 0;
      if (k->dtemp < 0.999999999 * b->DT[i])
        rem[i] = (int)std::min(std::max(ceil((b->DT[i] - k->dtemp) / (k->dt > 0.0 ? k->dt : b->DT[i])), 1.0),
                               4096.0);
    }
    if (launched > cap)
      return dfargo_set_error(DFARGO_ERR_STATE, "more than %ld sub-steps in one batched AlgoGas call", cap);
  }
  return DFARGO_OK;
}

extern "C" int dfargo_batch_algogas(dfargo_batch *b, const double *DT) {
  if (!b || !DT) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  const int n = (int)b->d.size();
  for (int i = 0; i < n; i++) {
    if (!(DT[i] > 0.0)) return dfargo_set_error(DFARGO_ERR_ARG, "disc %d: DT must be positive", i);
    if (b->d[i]->graph_version != b->version[i])
      return dfargo_set_error(DFARGO_ERR_STATE, "disc %d's sub-step changed since the batch was made", i);
  }
  HIPCHK(hipSetDevice(b->dev));
  if (b->mode == DFARGO_BATCH_STREAMS) {
    for (int i = 0; i < n; i++) b->DT[i] = DT[i];
    return batch_streams(b);
  }
  // each disc's own stream holds no pending work once the batch starts
  for (dfargo_ctx *c : b->d) HIPCHK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; i++) {
    dfargo_ctx *c = b->d[i];
    b->DT[i] = DT[i];
    call_start(c, DT[i], b->stream);
  }
  HIPCHK(hipGetLastError());
  long launched = 0;
  int batch = 1;
  const long cap = 50000000L;
  for (;;) {
    for (int k = 0; k < batch; k++) HIPCHK(hipGraphLaunch(b->gexec, b->stream));
    launched += batch;
    for (dfargo_ctx *c : b->d)
      HIPCHK(hipMemcpyAsync(c->h_ctrl, c->d_ctrl, sizeof(dfargo_ctrl), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    double rem = 0.0;
    for (int i = 0; i < n; i++) {
      const dfargo_ctrl *k = b->d[i]->h_ctrl;
      if (k->dtemp < 0.999999999 * b->DT[i])
        rem = std::max(rem, ceil((b->DT[i] - k->dtemp) / (k->dt > 0.0 ? k->dt : b->DT[i])));
    }
    if (rem <= 0.0) break;
    if (launched > cap)
      return dfargo_set_error(DFARGO_ERR_STATE, "more than %ld sub-steps in one batched AlgoGas call", cap);
    batch = (int)std::min(std::max(rem, 1.0), 4096.0);
  }
  return DFARGO_OK;
}
