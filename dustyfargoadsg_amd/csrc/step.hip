// step.hip -- CDNA4 kernels of one AlgoGas hydro sub-step
// (src/SourceEuler.c:398-716 and its callees).
//
// Storage type T is double or float; arithmetic is double throughout, with
// the reference's operation order, so that the FP64 build reproduces the
// reference pointwise (the library is compiled without contraction).
// Every kernel of a sub-step reads dfargo_ctrl::active and returns at once
// when the host's DT is already covered, so that a fixed batch of sub-steps
// can be launched (or replayed from a HIP graph) without a host round trip.
#include <hip/hip_runtime.h>
#include <math.h>

#include "fft_lds.h"
#include "kernels.h"

const char *kKernelNames[KID_COUNT] = {
    "sg_rows_fwd", "sg_cols", "sg_rows_inv",
    "cfl_ring",    "control",        "source",  "energy",         "transport_rad",
    "transport_azi", "finish",      "dust_move",      "dust_scan",  "dust_scatter",
    "dust_sort",     "dust_deposit", "dust_reorder", "evanescent", "nonreflecting",
    "sgz_mean",      "sgz_conv",     "sgz_fill",     "eos_step",      "visc1d",
    "df_cfl",        "df_axisg",      "df_source",    "df_stress",    "df_visc",       "df_substep2",
    "df_transport_rad", "df_transport_azi", "df_finish"};

// ------------------------------------------------------------------------
// helpers
// ------------------------------------------------------------------------

// The dust fluid runs with FP64 storage only: its branches in the shared
// kernels are compiled out of the FP32 instantiations.
#define DF_ON(T, a) (sizeof(T) == sizeof(double) && (a).df_on)

template <typename T> __device__ __forceinline__ double ld(const void *p, long i) {
  return (double)((const T *)p)[i];
}
template <typename T> __device__ __forceinline__ void st(void *p, long i, double v) {
  ((T *)p)[i] = (T)v;
}
// Unbiased stochastic rounding of the state to FP32 storage.  A value
// between two floats is stored as the upper one with probability equal to
// its distance from the lower one, in ulps; the draw is a hash of the cell,
// the sub-step count and a per-field salt, so the result is deterministic
// (a graph replay and an eager run agree).  Round-to-nearest would drop the
// same sub-ulp increment of a slowly evolving field every sub-step and let
// the error grow linearly; with unbiased rounding it grows as a random
// walk.  FP64 storage stores the value as it is.
__device__ __forceinline__ unsigned mix32(unsigned x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}
__device__ __forceinline__ unsigned sr_key(const dfargo_ctrl *c, long l, unsigned salt) {
  return mix32((unsigned)l * 0x9E3779B9U ^ mix32((unsigned)c->substeps_total * 0x85EBCA6BU + salt));
}
template <typename T> __device__ __forceinline__ T sr_round(double v, unsigned key);
template <> __device__ __forceinline__ double sr_round<double>(double v, unsigned) { return v; }
// The 29 mantissa bits of the double below the float's last bit are the
// distance to the float of smaller magnitude, in units of 2^-29 ulp: adding
// 29 random bits (the top bits of the key) and clearing them carries into
// the float's last bit with exactly that probability, and the remaining
// double converts to float exactly (for normal floats; exact floats, zeros,
// infinities and NaNs come out unchanged).
template <> __device__ __forceinline__ float sr_round<float>(double v, unsigned key) {
  const long long b = __double_as_longlong(v);
  const long long t = (b + (long long)(key >> 3)) & ~0x1FFFFFFFLL;
  return (float)__longlong_as_double(t);
}
template <typename T> __device__ __forceinline__ void st_sr(void *p, long i, double v, unsigned key) {
  ((T *)p)[i] = sr_round<T>(v, key);
}
// salts of the stochastically rounded stores
enum { SR_VR = 1, SR_VT, SR_E, SR_EN, SR_RAD = 8, SR_AZI = 16, SR_EVA = 24, SR_ACC = 32, SR_FIN = 40 };

// a / b from the hardware reciprocal, two Newton steps and one correction
// of the quotient: within an ulp or two of the IEEE quotient for operands
// well inside the exponent range (no scaling steps)
__device__ __forceinline__ double div_nr(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  const double q = a * r;
  return fma(r, fma(-b, q, a), q);
}

// arithmetic type of the transport stencils for a storage type, and its
// division: with FP64 storage the reference's IEEE double quotient; with
// FP32 storage FP64 arithmetic whose quotients need not be correctly
// rounded (their ulp is 2^-29 of the stored values' rounding)
template <typename T> struct TransportArith {
  using type = T;
  static __device__ __forceinline__ double div(double a, double b) { return a / b; }
};
template <> struct TransportArith<float> {
  using type = double;
  static __device__ __forceinline__ double div(double a, double b) { return div_nr(a, b); }
};

__device__ __forceinline__ double dmin2(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmax2(double a, double b) { return (b > a) ? b : a; }

// the sum over a wavefront, the same in every lane (xor butterfly)
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// deterministic block reductions (fixed tree for a given blockDim)
__device__ double block_sum(double v, double *sh) {
  int t = threadIdx.x, TB = blockDim.x, P = 1;
  while (P * 2 <= TB) P *= 2;
  sh[t] = v;
  __syncthreads();
  if (t < TB - P) sh[t] += sh[t + P];
  __syncthreads();
  for (int s = P / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}
__device__ double block_min(double v, double *sh) {
  int t = threadIdx.x, TB = blockDim.x, P = 1;
  while (P * 2 <= TB) P *= 2;
  sh[t] = v;
  __syncthreads();
  if (t < TB - P) sh[t] = dmin2(sh[t], sh[t + P]);
  __syncthreads();
  for (int s = P / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] = dmin2(sh[t], sh[t + s]);
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}
// N sums at once: wavefront shuffle trees, then one pass over the wave
// partials (blocks that are whole wavefronts), else the LDS tree.  sh must
// hold 17*N doubles (and blockDim doubles for the fallback).
template <int N> __device__ void block_sum_n(double (&v)[N], double *sh) {
  const int TB = blockDim.x;
  if (TB % 64 != 0) {
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = block_sum(v[k], sh);
    return;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = TB >> 6;
#pragma unroll
  for (int k = 0; k < N; k++)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < N; k++) sh[wave * N + k] = v[k];
  __syncthreads();
  if (threadIdx.x < N) {
    double s = 0.0;
    for (int w = 0; w < nw; w++) s += sh[w * N + threadIdx.x];
    sh[16 * N + threadIdx.x] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; k++) v[k] = sh[16 * N + k];
  __syncthreads();
}
__device__ double block_sum_fast(double v, double *sh) {
  double w[1] = {v};
  block_sum_n<1>(w, sh);
  return w[0];
}
__device__ double block_min_fast(double v, double *sh) {
  const int TB = blockDim.x;
  if (TB % 64 != 0) return block_min(v, sh);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = TB >> 6;
  for (int off = 32; off > 0; off >>= 1) v = dmin2(v, __shfl_down(v, off, 64));
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = sh[0];
    for (int w = 1; w < nw; w++) m = dmin2(m, sh[w]);
    sh[16] = m;
  }
  __syncthreads();
  double r = sh[16];
  __syncthreads();
  return r;
}

// N sums and one minimum in one pass (same scheme as block_sum_n); sh
// must hold 17*(N+1) doubles (and blockDim doubles for the fallback).
template <int N> __device__ void block_sum_min_n(double (&v)[N], double &mn, double *sh) {
  const int TB = blockDim.x;
  if (TB % 64 != 0) {
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = block_sum(v[k], sh);
    mn = block_min(mn, sh);
    return;
  }
  const int M = N + 1, lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = TB >> 6;
#pragma unroll
  for (int k = 0; k < N; k++)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  for (int off = 32; off > 0; off >>= 1) mn = dmin2(mn, __shfl_down(mn, off, 64));
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < N; k++) sh[wave * M + k] = v[k];
    sh[wave * M + N] = mn;
  }
  __syncthreads();
  if (threadIdx.x < N) {
    double s = 0.0;
    for (int w = 0; w < nw; w++) s += sh[w * M + threadIdx.x];
    sh[16 * M + threadIdx.x] = s;
  } else if (threadIdx.x == N) {
    double m = sh[N];
    for (int w = 1; w < nw; w++) m = dmin2(m, sh[w * M + N]);
    sh[16 * M + N] = m;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; k++) v[k] = sh[16 * M + k];
  mn = sh[16 * M + N];
  __syncthreads();
}

// Neumaier-compensated running sum
struct CSum {
  double s, c;
  __device__ void add(double x) {
    double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
  __device__ double value() const { return s + c; }
};

// Ring kernels hold K cells of a ring per thread.  With K = 1 a block may
// have 1024 threads (128 VGPRs per lane), with K = 2 768 threads (168
// VGPRs), from K = 3 on 512 threads (256 VGPRs), so that the K-wide working
// set stays in registers.
#define RING_LB(K) ((K) == 1 ? 1024 : ((K) == 2 ? 768 : 512))

// AspectRatio (src/Viscosity.c:107-122) and compute_smoothing
// (src/Force.c:255-263), LAMBDADOUBLING = 0
__device__ double d_aspect_ratio(const KArgs &a, double rad) {
  double ar = a.aspect_ratio;
  double rmin = a.transition_radius - a.transition_width * a.aspect_ratio;
  double rmax = a.transition_radius + a.transition_width * a.aspect_ratio;
  if (rad < rmin) ar *= a.transition_ratio;
  if ((rad >= rmin) && (rad <= rmax))
    ar *= exp((rmax - rad) / (rmax - rmin) * log(a.transition_ratio));
  return ar;
}
__device__ double d_smoothing(const KArgs &a, double r) {
  return a.thickness_smoothing * d_aspect_ratio(a, r) * pow(r, 1.0 + a.flaring);
}

// gravitational potential at cell (i, j): FillForcesArrays
// (src/Pframeforce.c:44-129) with the parameters the control kernel stored
__device__ double pot_at(const KArgs &a, const dfargo_ctrl *c, int i, int j) {
  double x = a.Rmed[i] * a.CosAz[j], y = a.Rmed[i] * a.SinAz[j];
  double pot = 0.0;
  for (int k = 0; k < c->nb; k++) {
    double xp = c->pot_x[k], yp = c->pot_y[k], mp = c->pot_m[k];
    double distance = (x - xp) * (x - xp) + (y - yp) * (y - yp);
    double distancesmooth = sqrt(distance + c->pot_smooth[k]);
    double pk = -DF_G * mp / distancesmooth;
    if (a.indirect_term) pk += DF_G * mp * c->pot_invd3[k] * (x * xp + y * yp);
    pot += pk;
  }
  double ps = -DF_G * 1.0 * a.InvRmed[i];
  if (a.gas_indirect) ps -= c->ind_x * x + c->ind_y * y;
  pot += ps;
  return pot;
}

// indirect part of the potential felt by the particles (the IndPot grid of
// FillForcesArrays, src/Pframeforce.c:60-88): the planets' term is
// assigned, not summed, so the last planet's stands, as in the reference
__device__ double indpot_at(const KArgs &a, const dfargo_ctrl *c, int i, int j) {
  const double x = a.Rmed[i] * a.CosAz[j], y = a.Rmed[i] * a.SinAz[j];
  double ip = 0.0;
  if (a.indirect_term && a.d_feel_planets && c->nb > 0) {
    const int k = c->nb - 1;
    ip = DF_G * c->pot_m[k] * c->pot_invd3[k] * (x * c->pot_x[k] + y * c->pot_y[k]);
  }
  if (a.gas_indirect && a.d_feel_disk) ip -= c->ind_x * x + c->ind_y * y;
  return ip;
}

// ------------------------------------------------------------------------
// per-call and per-sub-step control
// ------------------------------------------------------------------------

// GetPsysInfo (src/Psys.c:233-350), actions MARK (get = 0) and GET
// (get = 1).  The frame's centre is planet 0 (the outermost planet with
// COROTATEWITHOUTERPLANET) or, with FRAME B, the barycentre of planets 0 and
// 1, moved with FRAME G to the guiding centre of its orbit unless e < 1e-8.
// GET returns the angle the centre turned through since the mark.  The
// orbital frequency the reference also forms (with its self-gravity
// correction) serves only the FREQUENCY action and the guiding centre's
// velocity, which MARK and GET do not use.
__device__ double psys_info(const KArgs &a, dfargo_ctrl *c, int get) {
  const int k = a.frame_outer ? c->nb - 1 : 0;
  double x = c->pl_x[k], y = c->pl_y[k], vx = c->pl_vx[k], vy = c->pl_vy[k], m = c->pl_m[k] + 1.;
  if (a.frame == DFARGO_FRAME_BINARY_CENTER) {
    double m0 = c->pl_m[0], m1 = c->pl_m[1];
    if (m0 == 0.0) m0 = c->pl_m1[0];
    if (m1 == 0.0) m1 = c->pl_m1[1];
    x = (m0 * c->pl_x[0] + m1 * c->pl_x[1]) / (m0 + m1);
    y = (m0 * c->pl_y[0] + m1 * c->pl_y[1]) / (m0 + m1);
    vx = (m0 * c->pl_vx[0] + m1 * c->pl_vx[1]) / (m0 + m1);
    vy = (m0 * c->pl_vy[0] + m1 * c->pl_vy[1]) / (m0 + m1);
    m = m0 + m1 + 1.;
  }
  double xc = x, yc = y;
  if (a.frame == DFARGO_FRAME_GUIDING_CENTER) {
    const double h = x * vy - y * vx, d = sqrt(x * x + y * y);
    const double Ax = x * vy * vy - y * vx * vy - DF_G * m * x / d;
    const double Ay = y * vx * vx - x * vx * vy - DF_G * m * y / d;
    const double e = sqrt(Ax * Ax + Ay * Ay) / m, sa = h * h / DF_G / m / (1. - e * e);
    const double arg = (e == 0.0) ? 1.0 : (1.0 - d / sa) / e;
    double E = (fabs(arg) >= 1.0) ? DF_PI * (1. - arg / fabs(arg)) / 2. : acos((1.0 - d / sa) / e);
    if ((x * y * (vy * vy - vx * vx) + vx * vy * (x * x - y * y)) < 0) E = -E;
    const double M = E - e * sin(E), pa = atan2(Ay, Ax);
    if (!(e < 1e-8)) {
      xc = sa * cos(M + pa);
      yc = sa * sin(M + pa);
    }
  }
  if (!get) {
    c->mark_x = xc;
    c->mark_y = yc;
    return 0.0;
  }
  const double d2 = sqrt(xc * xc + yc * yc), d1 = sqrt(c->mark_x * c->mark_x + c->mark_y * c->mark_y);
  const double cross = c->mark_x * yc - xc * c->mark_y;
  c->mark_x = xc;
  c->mark_y = yc;
  return asin(cross / (d1 * d2));
}

// start of a sub-step: mass taper, n-body limit and the force-sum targets
// (src/SourceEuler.c:398-440, src/Pframeforce.c:131-161); one thread, run
// once the loop condition holds
__device__ void begin_substep(const KArgs &a, dfargo_ctrl *c) {
  int nb = c->nb;
  if (a.mass_taper_param > 1e-2) {
    double mt = (c->time - c->time0) / (a.mass_taper_param * 2.0 * DF_PI);
    mt = (mt > 1.0 ? 1.0 : pow(sin(mt * DF_PI / 2.0), 2.0));
    c->mass_taper = mt;
    if (mt < 1.0)
      for (int k = 0; k < nb; k++) c->pl_m[k] = c->pl_m0[k] + (c->pl_m1[k] - c->pl_m0[k]) * mt;
    c->particles_mass = c->particles_mass0 * mt;
  } else {
    for (int k = 0; k < nb; k++)
      if (c->pl_acc[k] < 1e-10) c->pl_m[k] = c->pl_m1[k];
    c->particles_mass = c->particles_mass0;
  }
  double dtnbody = 100.0;
  if (nb > 1) {
    dtnbody = 1e5;
    for (int k = 0; k < nb; k++)
      for (int j = k + 1; j < nb; j++) {
        double dx = c->pl_x[k] - c->pl_x[j], dy = c->pl_y[k] - c->pl_y[j];
        double dist = sqrt(dx * dx + dy * dy);
        double buf = 2.0 * DF_PI * sqrt(dist * dist * dist / (c->pl_m[k] + c->pl_m[j] + 1e-8)) / 100.0;
        dtnbody = dmin2(dtnbody, buf);
      }
  }
  c->dtnbody = dtnbody;
  // target 0: the star (ComputeAccel at 0,0, src/SourceEuler.c:482)
  c->tg_x[0] = 0.0;
  c->tg_y[0] = 0.0;
  c->tg_rsm[0] = a.indirect_smoothing;
  c->tg_a[0] = sqrt(0.0 * 0.0 + 0.0 * 0.0);
  c->tg_rh[0] = pow(0.0 / 3., 1. / 3.) * c->tg_a[0] + 1e-15;
  c->tg_planet[0] = -1;
  int nt = 1;
  for (int k = 0; k < nb; k++) {
    if (!c->pl_feel_disk[k]) continue;
    double m = c->pl_m[k], x = c->pl_x[k], y = c->pl_y[k], r = sqrt(x * x + y * y);
    double sm = a.roche ? r * pow(m / 3., 1. / 3.) * a.roche_smoothing : d_smoothing(a, r);
    c->tg_x[nt] = x;
    c->tg_y[nt] = y;
    c->tg_rsm[nt] = sm;
    c->tg_a[nt] = sqrt(x * x + y * y);
    c->tg_rh[nt] = pow(m / 3., 1. / 3.) * c->tg_a[nt] + 1e-15;
    c->tg_planet[nt] = k;
    nt++;
  }
  c->ntarget = nt;
}

// AlgoGas entry: loop variables, and the first sub-step's start.  The
// sub-step kernels read three flags of the control block: `go` (the loop
// condition of the current sub-step: read by the kernels that may run
// before or beside the control kernel, the CFL and self-gravity kernels),
// `active` (set by the control kernel from `go`, read by every kernel after
// it) and `go_next` (the condition of the next sub-step, set by the control
// kernel, which also prepares that sub-step's start; it tells the kernels
// after the control kernel whether theirs is the last sub-step of the
// call).  The last kernel of a sub-step copies go_next into go, so no flag
// changes under a kernel that reads it, even with the self-gravity kernels
// running concurrently with the control kernel.
__global__ void k_algogas_begin(KArgs a, double DT) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    dfargo_ctrl *c = a.ctrl;
    c->DT = DT;
    c->dtemp = 0.0;
    c->substeps = 0;
    c->gas_timesteps_cfl = 1;
    c->active = 0;
    c->go = c->dtemp < 0.999999999 * c->DT;
    c->go_next = c->go;
    c->fargo_margin_bits = 0x3FE0000000000000ULL; // 0.5
    if (c->go) begin_substep(a, c);
  }
}

// ConditionCFL per ring (src/SourceEuler.c:1340-1484) fused with the
// per-ring partial sums of ComputeForce (src/Force.c:155-214), GasTotalMass
// and the barycentre (src/SideEuler.c:15-77) and DetectCrash
// (src/SourceEuler.c:51-69): one read of rho, vr, vt (and cs, e).
template <typename T, int K> __global__ void __launch_bounds__(RING_LB(K)) k_cfl_ring(KArgs a) {
  __shared__ double sh[1024];
  const dfargo_ctrl *c = a.ctrl;
  if (!c->go) return;
  const int i = ring_of_block(a), ns = a.ns, nr = a.nr, t = threadIdx.x, TB = blockDim.x;
  const long base = (long)i * ns;
  // the CFL arithmetic is FP64 whatever the storage type: the sub-step
  // count ceil(DT / dt_min) then follows the FP64 run's, and FP32 storage
  // changes the solution, not the time discretisation
  using R = double;
  const T *vt = (const T *)a.vt + base, *vr = (const T *)a.vr + base;
  R vt_own[K];
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < K; k++) {
    vt_own[k] = (R)vt[t + k * TB];
    s += (double)vt_own[k];
  }
  s = block_sum_fast(s, sh);
  const double vmoy = s / (double)ns;
  if (t == 0) a.vmoy[i] = vmoy;
  double dtmin = 1e30;
  if (i >= 1) {
    const R viscosity = (R)a.nu_ring[i];
    const R dxrad = (R)(a.Rsup[i] - a.Rinf[i]);
    const R dxtheta = (R)(a.Rmed[i] * a.pm / (double)ns);
    const R vm = (R)vmoy, dxmin = (dxrad < dxtheta) ? dxrad : dxtheta;
    const R invdt5 = (viscosity != R(0)) ? viscosity * R(4) / (R)a.powmin2[i] : R(1e-10);
    // the ring's cell sizes divide every cell's terms: one reciprocal each
    // (within an ulp of the quotients, which only moves dt by as much)
    const R idxmin = R(1) / dxmin, idxrad = R(1) / dxrad, idxtheta = R(1) / dxtheta;
    R smax = R(0);
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int j = t + k * TB, jp = (j == ns - 1) ? 0 : j + 1;
      const R cs = a.energy ? (R)((const T *)a.cs)[base + j] : (R)a.cs_iso[i];
      const R vres = a.fast ? vt_own[k] - vm : vt_own[k];
      const R vrl = (R)vr[j];
      const R invdt1 = cs * idxmin;
      const R invdt2 = fabs(vrl) * idxrad;
      const R invdt3 = fabs(vres) * idxtheta;
      R dvr = (R)vr[j + ns] - vrl;
      R dvt = (R)vt[jp] - vt_own[k];
      if (dvr >= R(0)) dvr = R(1e-10); else dvr = -dvr;
      if (dvt >= R(0)) dvt = R(1e-10); else dvt = -dvt;
      const R q4 = dvr * idxrad, q5 = dvt * idxtheta;
      R invdt4 = (q5 > q4) ? q5 : q4;
      invdt4 *= R(4.0 * DF_CVNR * DF_CVNR);
      const R invdt6 = R(1e-10);
      // dt = CFLSECURITY / sqrt(S) decreases with S, and rounding keeps
      // that order: the smallest dt of the ring is that of its largest S
      const R S = invdt1 * invdt1 + invdt2 * invdt2 + invdt3 * invdt3 + invdt4 * invdt4 + invdt5 * invdt5 +
                  invdt6 * invdt6;
      smax = (S > smax) ? S : smax;
    }
    dtmin = R(DF_CFLSECURITY) / sqrt(smax);
  }
  // per-ring partials, stored partial-major (ring_part[p * nr + i]):
  // negative-density and negative-energy counts, disc mass, barycentre.
  // With the energy equation the same pass is ComputeSoundSpeed
  // (src/SourceEuler.c:2340, after the CFL above read the previous cs) and
  // the ring mean of cs that mpi_make1Dprofile feeds to FViscosity
  // (src/Viscosity.c:30-35) for the kernels after this one.
  double *part = a.ring_part + i;
  {
    CSum m = {0, 0}, xb = {0, 0}, yb = {0, 0};
    double crd = 0.0, cre = 0.0, css = 0.0;
    const double g = a.gamma;
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int j = t + k * TB;
      double dens = ld<T>(a.rho, base + j);
      if (dens < 0.0) crd += 1.0;
      if (a.energy) {
        const double en = ld<T>(a.en, base + j);
        if (en < 0.0) cre += 1.0;
        const double cs = sqrt(g * (g - 1.0) * en / dens);
        // the particles of this sub-step still read the previous cs
        if (a.d_on) ((T *)a.cs_old)[base + j] = ((const T *)a.cs)[base + j];
        st<T>(a.cs, base + j, cs);
        css += (double)(T)cs;
      }
      // the particle buckets of this sub-step start empty
      if (a.d_on) a.cell_count[base + j] = 0;
      double cm = a.Surf[i] * dens;
      m.add(cm);
      xb.add(cm * (a.Rmed[i] * a.CosAz[j]));
      yb.add(cm * (a.Rmed[i] * a.SinAz[j]));
    }
    double v[6] = {crd, cre, m.value(), xb.value(), yb.value(), css};
    block_sum_min_n<6>(v, dtmin, sh);
    if (t == 0) {
#pragma unroll
      for (int k = 0; k < 5; k++) part[(long)k * nr] = v[k];
      a.ring_dt[i] = dtmin;
      if (a.energy) {
        const double gcs = v[5] / (double)ns;
        a.gcs[i] = gcs;
        double nu = a.visc_const;
        if (a.alpha_on) nu = a.alpha * gcs * gcs * a.r15[i];
        a.nu_ring[i] = a.visc_on ? nu * a.visc_mult[i] : 0.0;
      }
    }
  }
  // then fx, fx (Hill-cut), fy, fy (Hill-cut) per force target; the
  // inner/outer split of GlobalForce is only ever summed, so it is not kept
  const bool hill = a.exclude_hill;
  for (int tg = 0; tg < c->ntarget; tg++) {
    const double x = c->tg_x[tg], y = c->tg_y[tg], rsm = c->tg_rsm[tg];
    const double rh = c->tg_rh[tg];
    CSum fx[2] = {{0, 0}, {0, 0}}, fy[2] = {{0, 0}, {0, 0}};
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int j = t + k * TB;
      double xc = a.Rmed[i] * a.CosAz[j], yc = a.Rmed[i] * a.SinAz[j];
      double cellmass = a.Surf[i] * ld<T>(a.rho, base + j);
      double dx = xc - x, dy = yc - y;
      double d2 = dx * dx + dy * dy;
      double dist2 = d2 + rsm * rsm;
      double distance = sqrt(dist2);
      double InvDist3 = 1.0 / dist2 / distance;
      fx[0].add(DF_G * cellmass * dx * InvDist3 * 1.0);
      fy[0].add(DF_G * cellmass * dy * InvDist3 * 1.0);
      if (hill) {
        double planet_distance = sqrt(d2);
        double cutoff = a.exclude_hill_factor * rh, hill_cut;
        if (planet_distance / cutoff < 0.5) hill_cut = 0.0;
        else if (planet_distance > cutoff) hill_cut = 1.0;
        else hill_cut = pow(sin((planet_distance / cutoff - .5) * DF_PI), 2.);
        fx[1].add(DF_G * cellmass * dx * InvDist3 * hill_cut);
        fy[1].add(DF_G * cellmass * dy * InvDist3 * hill_cut);
      }
    }
    double *g = part + (long)(5 + 4 * tg) * nr;
    if (hill) {
      double v[4] = {fx[0].value(), fx[1].value(), fy[0].value(), fy[1].value()};
      block_sum_n<4>(v, sh);
      if (t == 0)
#pragma unroll
        for (int k = 0; k < 4; k++) g[(long)k * nr] = v[k];
    } else {
      double v[2] = {fx[0].value(), fy[0].value()};
      block_sum_n<2>(v, sh);
      if (t == 0) { g[0] = v[0]; g[2 * (long)nr] = v[1]; }
    }
  }
}

// Cash-Karp RK5 of the star+planet problem (src/RungeKunta.c:14-91)
__device__ void d_deriv_rk5(const double *q, const double *masses, double *deriv, int n, double dt,
                            const int *feelothers, int indirect) {
  const double *x = q, *y = q + n, *vx = y + n, *vy = vx + n;
  double *dx = deriv, *dy = deriv + n, *dvx = dy + n, *dvy = dvx + n;
  for (int i = 0; i < n; i++) {
    const double Di = sqrt(x[i] * x[i] + y[i] * y[i]);
    dx[i] = vx[i];
    dy[i] = vy[i];
    dvx[i] = -DF_G * 1.0 / Di / Di / Di * x[i];
    dvy[i] = -DF_G * 1.0 / Di / Di / Di * y[i];
    for (int j = 0; j < n; j++) {
      if (indirect) {
        const double Dj = sqrt(x[j] * x[j] + y[j] * y[j]);
        dvx[i] -= DF_G * masses[j] / Dj / Dj / Dj * x[j];
        dvy[i] -= DF_G * masses[j] / Dj / Dj / Dj * y[j];
      }
      if ((j != i) && feelothers[i]) {
        double dist = (x[i] - x[j]) * (x[i] - x[j]) + (y[i] - y[j]) * (y[i] - y[j]);
        dist = sqrt(dist);
        dvx[i] += DF_G * masses[j] / dist / dist / dist * (x[j] - x[i]);
        dvy[i] += DF_G * masses[j] / dist / dist / dist * (y[j] - y[i]);
      }
    }
  }
  for (int i = 0; i < 4 * n; i++) deriv[i] *= dt;
}
__device__ void d_translate(const double *q0, const double *k1, const double *k2, const double *k3,
                            const double *k4, const double *k5, double c1, double c2, double c3,
                            double c4, double c5, double *q1, int n) {
  for (int i = 0; i < 4 * n; i++)
    q1[i] = q0[i] + c1 * k1[i] + c2 * k2[i] + c3 * k3[i] + c4 * k4[i] + c5 * k5[i];
}
// k holds the six stages, 6 x 4*DF_MAXPL doubles (LDS: the control kernel
// keeps per-thread arrays out of scratch memory)
__device__ void d_rk5(const double *q0, double dt, const double *m, double *q1, int n,
                      const int *fo, int ind, double *k) {
  double *k1 = k, *k2 = k1 + 4 * DF_MAXPL, *k3 = k2 + 4 * DF_MAXPL, *k4 = k3 + 4 * DF_MAXPL,
         *k5 = k4 + 4 * DF_MAXPL, *k6 = k5 + 4 * DF_MAXPL;
  for (int i = 0; i < 4 * DF_MAXPL; i++) k1[i] = k2[i] = k3[i] = k4[i] = k5[i] = k6[i] = 0.0;
  d_deriv_rk5(q0, m, k1, n, dt, fo, ind);
  d_translate(q0, k1, k2, k3, k4, k5, 0.2, 0.0, 0.0, 0.0, 0.0, q1, n);
  d_deriv_rk5(q1, m, k2, n, dt, fo, ind);
  d_translate(q0, k1, k2, k3, k4, k5, 0.075, 0.225, 0.0, 0.0, 0.0, q1, n);
  d_deriv_rk5(q1, m, k3, n, dt, fo, ind);
  d_translate(q0, k1, k2, k3, k4, k5, 0.3, -0.9, 1.2, 0.0, 0.0, q1, n);
  d_deriv_rk5(q1, m, k4, n, dt, fo, ind);
  d_translate(q0, k1, k2, k3, k4, k5, -11.0 / 54.0, 2.5, -70.0 / 27.0, 35.0 / 27.0, 0.0, q1, n);
  d_deriv_rk5(q1, m, k5, n, dt, fo, ind);
  d_translate(q0, k1, k2, k3, k4, k5, 1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0,
              44275.0 / 110592.0, 253.0 / 4096.0, q1, n);
  d_deriv_rk5(q1, m, k6, n, dt, fo, ind);
  for (int i = 0; i < 4 * n; i++)
    q1[i] = q0[i] + 37.0 / 378.0 * k1[i] + 250.0 / 621.0 * k3[i] + 125.0 / 594.0 * k4[i] +
            512.0 / 1771.0 * k6[i];
}

// Scalar part of a sub-step, one block: the CFL reduction and sub-step dt
// (src/SourceEuler.c:443-470, 1485-1551), the gas indirect term and the
// barycentre acceleration (src/SideEuler.c:91-112), the potential
// parameters of FillForcesArrays (src/Pframeforce.c:44-90), and the planet
// update (AdvanceSystemFromDisk, AdvanceSystemRK5, RotatePsys).
// AccreteOntoPlanets (src/Planet.c:11-126) by the whole control workgroup
// once dt is known.  Each cell of a planet's window belongs to one thread,
// which updates its density in place.  The planet's mass and momentum gains
// are block sums over the accreted cells, and so are the changes of the
// disc's mass, barycentre and force sums: the CFL kernel formed those
// partials before the accretion, the reference forms them after it, and
// they are linear in the density.  At most one planet accretes (checked on
// the host), so no cell is accreted by two threads.  red holds 16 doubles
// per reduced value.
template <typename T>
__device__ void accrete_planets(const KArgs &a, dfargo_ctrl *c, double dt, double *sums, double *red) {
  const int t = threadIdx.x, TB = blockDim.x, lane = t & 63, wv = t >> 6, nw = TB >> 6;
  const int nr = a.nr, ns = a.ns, ntg = c->ntarget;
  constexpr int NV = 6 + 2 * DF_MAXTG;
  for (int k = 0; k < c->nb; k++) {
    if (!(c->pl_acc[k] > 1e-10)) continue;
    const double facc = dt * c->pl_acc[k], facc1 = 1.0 / 3.0 * facc, facc2 = 2.0 / 3.0 * facc;
    const double Xp = c->pl_x[k], Yp = c->pl_y[k], Mp = c->pl_m[k];
    const double Rp = sqrt(Xp * Xp + Yp * Yp), RRoche = pow((1.0 / 3.0 * Mp), (1.0 / 3.0)) * Rp;
    int i_min = 0, i_max = nr - 1;
    while ((i_min < nr) && (a.Rsup[i_min] < Rp - RRoche)) i_min++;
    while ((i_max > 0) && (a.Rinf[i_max] > Rp + RRoche)) i_max--;
    const double angle = atan2(Yp, Xp);
    const int j_min = (int)((double)ns / a.pm * (angle - 2.0 * RRoche / Rp - a.pmin));
    const int j_max = (int)((double)ns / a.pm * (angle + 2.0 * RRoche / Rp - a.pmin));
    const int wj = j_max - j_min + 1;
    const int ncell = (i_max >= i_min && wj > 0) ? (i_max - i_min + 1) * wj : 0;
    // 0-2: the planet's mass and momentum gains; 3-5: the disc's mass and
    // barycentre changes (per thread); then fx, fy per force target, summed
    // per wavefront after each chunk of TB cells into the wave's own slots
    // of red (few live registers: no spills in the control kernel)
    for (int q = t; q < NV * 16; q += TB) red[q] = 0.0;
    __syncthreads();
    double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int base = 0; base < ncell; base += TB) {
      const int idx = base + t;
      double dcm = 0.0, xc = 0.0, yc = 0.0;
      if (idx < ncell) {
        const int i = i_min + idx / wj, j = j_min + idx % wj;
        int jf = j % ns;
        jf = (jf < 0) ? jf + ns : jf;
        const long l = (long)i * ns + jf, lip = l + ns, ljp = (jf == ns - 1) ? (long)i * ns : l + 1;
        xc = a.Rmed[i] * a.CosAz[jf];
        yc = a.Rmed[i] * a.SinAz[jf];
        const double dx = Xp - xc, dy = Yp - yc, distance = sqrt(dx * dx + dy * dy);
        // the inner radius (0.45) lies inside the outer one (0.75)
        if (distance < 0.75 * RRoche) {
          const double vtcell = 0.5 * (ld<T>(a.vt, l) + ld<T>(a.vt, ljp)) + a.Rmed[i] * c->omega_frame;
          const double vrcell = 0.5 * (ld<T>(a.vr, l) + ld<T>(a.vr, lip));
          const double vxcell = (vrcell * xc - vtcell * yc) / a.Rmed[i];
          const double vycell = (vrcell * yc + vtcell * xc) / a.Rmed[i];
          const double d0 = ld<T>(a.rho, l);
          double dens = d0;
          double deltaM = facc1 * dens * a.Surf[i];
          dens *= (1.0 - facc1);
          v[0] += deltaM;
          v[1] += deltaM * vxcell;
          v[2] += deltaM * vycell;
          if (distance < 0.45 * RRoche) {
            deltaM = facc2 * dens * a.Surf[i];
            dens *= (1.0 - facc2);
            v[0] += deltaM;
            v[1] += deltaM * vxcell;
            v[2] += deltaM * vycell;
          }
          const T dstored = sr_round<T>(dens, sr_key(c, l, SR_ACC));
          ((T *)a.rho)[l] = dstored;
          // the change of the cell's mass as stored, in the partials' terms
          dcm = a.Surf[i] * ((double)dstored - d0);
          v[3] += dcm;
          v[4] += dcm * xc;
          v[5] += dcm * yc;
        }
      }
      for (int tg = 0; tg < ntg; tg++) {
        double fx = 0.0, fy = 0.0;
        if (dcm != 0.0) {
          const double ddx = xc - c->tg_x[tg], ddy = yc - c->tg_y[tg], rsm = c->tg_rsm[tg];
          const double dist2 = ddx * ddx + ddy * ddy + rsm * rsm, dd = sqrt(dist2);
          const double InvDist3 = 1.0 / dist2 / dd;
          fx = DF_G * dcm * ddx * InvDist3 * 1.0;
          fy = DF_G * dcm * ddy * InvDist3 * 1.0;
        }
        fx = wave_sum(fx);
        fy = wave_sum(fy);
        if (lane == 0) {
          red[(6 + 2 * tg) * 16 + wv] += fx;
          red[(7 + 2 * tg) * 16 + wv] += fy;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 6; q++) {
      const double w = wave_sum(v[q]);
      if (lane == 0) red[q * 16 + wv] = w;
    }
    __syncthreads();
    if (t == 0) {
      // the wavefronts' sums of value q, in wavefront order
      auto total = [&](int q) {
        double x = 0.0;
        for (int w = 0; w < nw; w++) x += red[q * 16 + w];
        return x;
      };
      double Px = Mp * c->pl_vx[k], Py = Mp * c->pl_vy[k], M = Mp;
      Px += total(1);
      Py += total(2);
      M += total(0);
      if (c->pl_feel_disk[k]) {
        c->pl_vx[k] = Px / M;
        c->pl_vy[k] = Py / M;
      }
      c->pl_m[k] = M;
      sums[2] += total(3);
      sums[3] += total(4);
      sums[4] += total(5);
      for (int tg = 0; tg < ntg; tg++) {
        sums[5 + 4 * tg] += total(6 + 2 * tg);
        sums[5 + 4 * tg + 2] += total(7 + 2 * tg);
      }
    }
    __syncthreads();
  }
}

template <typename T>
__device__ void control_serial(const KArgs &a, dfargo_ctrl *c, const double *sums, double *sh, double s_dt);

// Two builds: without accretion, 1024 threads; with it, 512, whose larger
// register budget keeps the accretion path and the one-thread n-body code
// free of spills.
template <bool ACC> struct CtrlBlock { static constexpr int n = 1024; };
template <> struct CtrlBlock<true> { static constexpr int n = 512; };
template <typename T, bool ACC> __global__ void __launch_bounds__(CtrlBlock<ACC>::n) k_control(KArgs a) {
  __shared__ double sh[1024];
  __shared__ double sums[DF_NPART(DF_MAXTG)];
  // The control block is staged in LDS for the kernel's run: the one-thread
  // part below reads and writes dozens of its fields in dependent order,
  // which from global memory costs a memory latency each.  It is copied in
  // and back by all threads; no kernel that may run beside this one writes
  // the block (the self-gravity transforms only read its go flag, which
  // this kernel leaves as it is).
  __shared__ dfargo_ctrl cs;
  static_assert(sizeof(dfargo_ctrl) % 8 == 0, "the control block is copied as 8-byte words");
  constexpr int NW = (int)(sizeof(dfargo_ctrl) / 8);
  dfargo_ctrl *const cg = a.ctrl;
  const int t = threadIdx.x, TB = blockDim.x, nr = a.nr, ns = a.ns;
  if (!cg->go) {
    if (t == 0) cg->active = 0;
    return;
  }
  for (int w = t; w < NW; w += TB) ((unsigned long long *)&cs)[w] = ((const unsigned long long *)cg)[w];
  __syncthreads();
  dfargo_ctrl *c = &cs;
  double nd = 1e30;
  for (int i = 1 + t; i < nr; i += TB) {
    double v = a.ring_dt[i];
    // with the dust fluid each cell's limit is max2(min2(dtg, dtd), 1e-10)
    // (src/SourceEuler.c:1462-1465); both maps are monotone, so per ring
    if (a.df_on) v = dmax2(dmin2(v, a.df_ring_dt[i]), 1.e-10);
    nd = dmin2(nd, v);
  }
  for (int i = t; i < nr - 1; i += TB) {
    double dt = a.pm * DF_CFLSECURITY / (double)ns /
                fabs(a.vmoy[i] * a.InvRmed[i] - a.vmoy[i + 1] * a.InvRmed[i + 1]);
    nd = dmin2(nd, dt);
    // the dust fluid's shear limit (src/SourceEuler.c:1486-1491)
    if (a.df_on && !a.df_sfta) {
      dt = a.pm * DF_CFLSECURITY / (double)ns / fabs(a.dvmoy[i] * a.InvRmed[i] - a.dvmoy[i + 1] * a.InvRmed[i + 1]);
      nd = dmin2(nd, dt);
    }
  }
  nd = block_min_fast(nd, sh);
  // minimum stopping time of the last particle interpolation (reset to 1e6
  // at each interpolation, src/Dsys.c:300): the minimum of the per-block
  // minima, compared as bit patterns of positive doubles
  double dust_min_ts = 1e6;
  if (a.d_on && a.d_np > 0 && a.d_feedback) {
    unsigned long long m = (unsigned long long)__double_as_longlong(1e6);
    const int nb = (a.d_np + 255) / 256;
    for (int b = t; b < nb; b += TB) m = (a.p_tsmin[b] < m) ? a.p_tsmin[b] : m;
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = (unsigned long long)__shfl_xor((long)m, off, 64);
      m = (o < m) ? o : m;
    }
    __shared__ unsigned long long tsw[16];
    if ((t & 63) == 0) tsw[t >> 6] = m;
    __syncthreads();
    m = tsw[0];
    for (int w = 1; w < TB / 64; w++) m = (tsw[w] < m) ? tsw[w] : m;
    dust_min_ts = __longlong_as_double((long long)m);
  }
  // each wavefront sums whole partials: lanes stride the rings with a
  // compensated sum, then a fixed shuffle tree (deterministic)
  const int NP = DF_NPART(c->ntarget), lane = t & 63, nw = TB >> 6;
  for (int p = t >> 6; p < NP; p += nw) {
    const double *src = a.ring_part + (long)p * nr;
    CSum s = {0, 0};
    for (int i = lane; i < nr; i += 64) s.add(src[i]);
    double v = s.value();
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) sums[p] = v;
  }
  __syncthreads();
  // sub-step dt
  __shared__ double s_dt;
  if (t == 0) {
    double dt;
    if (!a.d_fixed_dt) {
      double deltaT = c->DT - c->dtemp;
      // the minimum stopping time of the last interpolation, resolved by two
      // steps (src/SourceEuler.c:1501-1507)
      if (a.d_on && a.d_np > 0 && a.d_feedback) {
        double dt_fb = 0.5 * dust_min_ts;
        if (dt_fb == 0.0) dt_fb = 1e10;
        if (dt_fb < nd) nd = dt_fb;
      }
      if (deltaT < nd) nd = deltaT;
      int gts = (int)(ceil(deltaT / nd));
      double dthydro = deltaT / (double)gts;
      dt = dmin2(dthydro, c->dtnbody);
      c->gas_timesteps_cfl = gts;
    } else {
      // the particles' fixed time step (src/SourceEuler.c:457-463)
      dt = a.d_dust_timestep;
      c->mindt = a.d_dust_timestep;
      c->maxdt = a.d_dust_timestep;
    }
    c->dtemp += dt;
    if (dt < c->mindt) c->mindt = dt;
    if (dt > c->maxdt) c->maxdt = dt;
    c->dt = dt;
    s_dt = dt;
  }
  if constexpr (ACC) {
    // AccreteOntoPlanets right after dt (src/SourceEuler.c:454)
    __syncthreads();
    accrete_planets<T>(a, c, s_dt, sums, sh);
  }
  __syncthreads();
  if (t == 0) control_serial<T>(a, c, sums, sh, s_dt);
  __syncthreads();
  for (int w = t; w < NW; w += TB) ((unsigned long long *)cg)[w] = ((const unsigned long long *)&cs)[w];
}

// The one-thread part of the control kernel on the LDS copy c of the
// control block: crash flags, indirect terms, barycentre, FillForcesArrays'
// parameters, AdvanceSystemFromDisk, AdvanceSystemRK5, the frame and the
// loop condition.
template <typename T>
__device__ void control_serial(const KArgs &a, dfargo_ctrl *c, const double *sums, double *sh, double s_dt) {
  const double dt = s_dt;
  c->omega_prev = c->omega_frame;
  c->crashed_d = sums[0] > 0.0;
  c->crashed_e = sums[1] > 0.0;
  // gas indirect term: ComputeAccel(star), src/SideEuler.c:282-298
  const double *g0 = sums + 5;
  c->dop_x = a.exclude_hill ? g0[1] : g0[0];
  c->dop_y = a.exclude_hill ? g0[3] : g0[2];
  // barycentre (src/SideEuler.c:42-112)
  double total = sums[2];
  c->barx[2] = sums[3] / (1.0 + total);
  c->bary[2] = sums[4] / (1.0 + total);
  c->dopalt_x = -(c->barx[2] - 2.0 * c->barx[1] + c->barx[0]) / (dt * dt);
  c->dopalt_y = -(c->bary[2] - 2.0 * c->bary[1] + c->bary[0]) / (dt * dt);
  c->barx[0] = c->barx[1]; c->barx[1] = c->barx[2];
  c->bary[0] = c->bary[1]; c->bary[1] = c->bary[2];
  // FillForcesArrays parameters
  double planetmass_taper = 1.0;
  if (a.mass_taper_param > 1e-2) {
    double mt = (c->time - c->time0) / (a.mass_taper_param * 2.0 * DF_PI);
    planetmass_taper = (mt > 1.0 ? 1.0 : .5 * (1.0 - cos(DF_PI * mt)));
  }
  c->ind_x = -c->dop_x;
  c->ind_y = -c->dop_y;
  if (a.alt_indirect) {
    c->ind_x = -c->dopalt_x;
    c->ind_y = -c->dopalt_y;
  }
  if (!a.indirect_term) {
    c->ind_x = 0.0;
    c->ind_y = 0.0;
  }
  const int nb = c->nb;
  for (int k = 0; k < nb; k++) {
    double xp = c->pl_x[k], yp = c->pl_y[k], mp = c->pl_m[k] * planetmass_taper;
    double PlanetDistance = sqrt(xp * xp + yp * yp);
    double invd3 = 1.0 / PlanetDistance / PlanetDistance / PlanetDistance;
    double RRoche = PlanetDistance * pow((1.0 / 3.0 * mp), 1.0 / 3.0);
    double sm = a.roche ? RRoche * a.roche_smoothing : d_smoothing(a, PlanetDistance);
    c->pot_x[k] = xp;
    c->pot_y[k] = yp;
    c->pot_m[k] = mp;
    c->pot_invd3[k] = invd3;
    c->pot_smooth[k] = sm * sm;
  }
  // the corotating frame's mark (src/SourceEuler.c:476), before the
  // disc's kick changes the velocities
  if (a.frame) psys_info(a, c, 0);
  // AdvanceSystemFromDisk (src/Pframeforce.c:131-161)
  for (int tg = 1; tg < c->ntarget; tg++) {
    int k = c->tg_planet[tg];
    const double *g = sums + 5 + 4 * tg;
    double gx = a.exclude_hill ? g[1] : g[0];
    double gy = a.exclude_hill ? g[3] : g[2];
    c->pl_vx[k] += dt * gx;
    c->pl_vy[k] += dt * gy;
    c->pl_vx[k] += dt * c->ind_x;
    c->pl_vy[k] += dt * c->ind_y;
  }
  // AdvanceSystemRK5 (src/Pframeforce.c:163-266)
  if (nb > 0) {
    double *q0 = sh, *q1 = q0 + 4 * DF_MAXPL, *m = q1 + 4 * DF_MAXPL, *kst = m + DF_MAXPL;
    if (!a.forced_circular) {
      for (int i = 0; i < nb; i++) {
        q0[i] = c->pl_x[i]; q0[i + nb] = c->pl_y[i];
        q0[i + 2 * nb] = c->pl_vx[i]; q0[i + 3 * nb] = c->pl_vy[i];
        m[i] = c->pl_m[i];
      }
      d_rk5(q0, dt, m, q1, nb, c->pl_feel_others, a.indirect_term, kst);
    }
    for (int i = 1 - (c->time >= a.release_date); i < nb; i++) {
      if (!a.forced_circular) {
        c->pl_x[i] = q1[i]; c->pl_y[i] = q1[i + nb];
        c->pl_vx[i] = q1[i + 2 * nb]; c->pl_vy[i] = q1[i + 3 * nb];
      } else {
        double x = c->pl_x[i], y = c->pl_y[i], theta = atan2(y, x);
        double vx = c->pl_vx[i], vy = c->pl_vy[i];
        double r = sqrt(x * x + y * y), v = sqrt(vx * vx + vy * vy);
        double omega = (-y * vx + x * vy) / r / r, dtheta = omega * dt;
        c->pl_x[i] = r * cos(theta + dtheta);
        c->pl_y[i] = r * sin(theta + dtheta);
        c->pl_vx[i] = -v * sin(theta + dtheta);
        c->pl_vy[i] = v * cos(theta + dtheta);
      }
    }
    if (c->time < a.release_date) {
      double x = c->pl_x[0], y = c->pl_y[0], r = sqrt(x * x + y * y), theta = atan2(y, x);
      double rdot = (a.release_radius - r) / (a.release_date - c->time);
      double omega = sqrt((1. + c->pl_m[0]) / r / r / r), new_r = r + rdot * dt, denom = r - new_r;
      double dtheta = (denom != 0.0) ? 2. * dt * r * omega / denom * (sqrt(r / new_r) - 1.) : omega * dt;
      double vx = rdot, vy = new_r * sqrt((1. + c->pl_m[0]) / new_r / new_r / new_r);
      c->pl_x[0] = new_r * cos(dtheta + theta);
      c->pl_y[0] = new_r * sin(dtheta + theta);
      c->pl_vx[0] = vx * cos(dtheta + theta) - vy * sin(dtheta + theta);
      c->pl_vy[0] = vx * sin(dtheta + theta) + vy * cos(dtheta + theta);
    }
    // the frame follows its centre (src/SourceEuler.c:516-524); the
    // source kernel applies CorrectVtheta (vt -= domega Rmed) where it
    // reads vt
    if (a.frame) {
      const double omega_new = psys_info(a, c, 1) / dt;
      c->domega = omega_new - c->omega_frame;
      c->omega_frame = omega_new;
    }
    // the particles of this sub-step feel the planets where the n-body
    // step left them (src/SourceEuler.c:504-513 come before RotatePsys)
    if (a.d_on)
      for (int i = 0; i < nb; i++) {
        c->dust_pl_x[i] = c->pl_x[i];
        c->dust_pl_y[i] = c->pl_y[i];
      }
    // RotatePsys (src/Psys.c:352-370)
    double angle = c->omega_frame * dt, sint = sin(angle), cost = cos(angle);
    for (int i = 0; i < nb; i++) {
      double xt = c->pl_x[i], yt = c->pl_y[i];
      c->pl_x[i] = xt * cost + yt * sint;
      c->pl_y[i] = -xt * sint + yt * cost;
      xt = c->pl_vx[i];
      yt = c->pl_vy[i];
      c->pl_vx[i] = xt * cost + yt * sint;
      c->pl_vy[i] = -xt * sint + yt * cost;
    }
  }
  // PhysicalTime += dt closes the sub-step (src/SourceEuler.c:714); every
  // other use of PhysicalTime in the sub-step is above
  c->time += dt;
  c->substeps += 1;
  c->substeps_total += 1;
  c->active = 1;
  if (a.d_on) {
    // masses of this sub-step for the particle kernels, and a fresh
    // minimum for their interpolation (src/Dsys.c:300)
    c->dust_mp = c->particles_mass;
    for (int k = 0; k < nb; k++) c->dust_pl_m[k] = c->pl_m[k];
    c->out_count = 0;
  }
  // loop condition and start of the next sub-step
  c->go_next = c->dtemp < 0.999999999 * c->DT;
  if (c->go_next) begin_substep(a, c);
}

// ComputeSoundSpeed at the start of AlgoGas (src/SourceEuler.c:2340); in
// the sub-steps the CFL kernel recomputes cs with its ring means
template <typename T, int K> __global__ void __launch_bounds__(RING_LB(K)) k_eos(KArgs a) {
  const int i = ring_of_block(a), ns = a.ns, t = threadIdx.x, TB = blockDim.x;
  const long base = (long)i * ns;
  const double g = a.gamma;
#pragma unroll
  for (int k = 0; k < K; k++) {
    const long l = base + t + k * TB;
    st<T>(a.cs, l, sqrt(g * (g - 1.0) * ld<T>(a.en, l) / ld<T>(a.rho, l)));
  }
}

// ComputeViscousTerms stresses (src/Viscosity.c:159-300) on (vr, vt)
template <typename T>
__device__ void visc_stress_cell(const KArgs &a, const void *vr, const void *vt, int i, int j,
                                 double *Trr, double *Tpp, double *Trp, double *divv) {
  const int ns = a.ns;
  const long l = (long)i * ns + j, lip = l + ns, ljp = (j == ns - 1) ? (long)i * ns : l + 1;
  const double onethird = 1.0 / 3.0, nu = a.nu_ring[i];
  double vrl = ld<T>(vr, l), vrp = ld<T>(vr, lip), vtl = ld<T>(vt, l), vtp = ld<T>(vt, ljp);
  double dv = (vrp * a.Rsup[i] - vrl * a.Rinf[i]) * a.InvDiffRsup[i] * a.InvRmed[i];
  dv += (vtp - vtl) * a.invdphi * a.InvRmed[i];
  double Drr = (vrp - vrl) * a.InvDiffRsup[i];
  double Dpp = (vtp - vtl) * a.invdphi * a.InvRmed[i] + 0.5 * (vrp + vrl) * a.InvRmed[i];
  double rho = ld<T>(a.rho, l);
  *Trr = 2.0 * rho * nu * (Drr - onethird * dv);
  *Tpp = 2.0 * rho * nu * (Dpp - onethird * dv);
  *divv = dv;
  if (i >= 1) {
    const long lim = l - ns, ljm = (j == 0) ? (long)i * ns + ns - 1 : l - 1, ljmim = ljm - ns;
    double Drp = 0.5 * (a.Rinf[i] * (vtl * a.InvRmed[i] - ld<T>(vt, lim) * a.InvRmed[i - 1]) *
                            a.InvDiffRmed[i] +
                        (vrl - ld<T>(vr, ljm)) * a.invdphi * a.InvRinf[i]);
    *Trp = 2.0 * 0.25 * (rho + ld<T>(a.rho, lim) + ld<T>(a.rho, ljm) + ld<T>(a.rho, ljmim)) * nu * Drp;
  } else {
    *Trp = 0.0;
  }
}

// stresses of the current state, for the diagnostics outside a sub-step
template <typename T> __global__ void __launch_bounds__(256) k_visc_stress(KArgs a, int from_state) {
  const int i = blockIdx.y, ns = a.ns, j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ns) return;
  const long l = (long)i * ns + j;
  double Trr, Tpp, Trp, dv;
  visc_stress_cell<T>(a, from_state ? a.vr : a.vri, from_state ? a.vt : a.vti, i, j, &Trr, &Tpp, &Trp, &dv);
  st<T>(a.trr, l, Trr);
  st<T>(a.tpp, l, Tpp);
  st<T>(a.trp, l, Trp);
}

// Source terms of one sub-step on tiles of SRC_TR rings x SRC_TC columns:
// SubStep1 (src/SourceEuler.c:865-994, with the self-gravity kick of
// update_sgvelocity, src/sgupdate.c:30-72), the viscous stresses and
// velocity update (ComputeViscousTerms + UpdateVelocitiesWithViscosity,
// src/Viscosity.c:159-378) and SubStep2 (src/SourceEuler.c:1012-1116), in
// that order, with the intermediate VradInt / VthetaInt and stresses in
// LDS.  Each stage is evaluated on the tile and the halo the next stage
// reads: potential on a 4-cell halo, SubStep1 on 3, stresses on 2, the
// viscous update on 1.  The state is read only; the post-SubStep2
// velocities go to vri / vti and the energy to e_int (the ActualiseGas
// copies of src/SourceEuler.c:598-630 are deferred to the end of the
// sub-step), so no workgroup overwrites what another one still reads.
#define SRC_TR (SrcBlock<T>::tr)
#define SRC_TC DF_SRC_COLS
static_assert(SRC_TC == 64, "the SubStep2 stage sums one tile ring per wavefront");
#define SRC_PR (SRC_TR + 6)
#define SRC_PC (SRC_TC + 6)
#define SRC_VR (SRC_TR + 5)
#define SRC_VC (SRC_TC + 5)
#define SRC_SR (SRC_TR + 3)
#define SRC_SC (SRC_TC + 3)
#define SRC_FC (SRC_TC + 7)  // state staged on rings i0-4 .. i0+TR+1, columns j0-4 .. j0+TC+2
// Rings (tr) and threads (n) per tile.  With FP32 storage 16 rings (78 KiB
// of LDS, two tiles per CU) recompute fewer halo cells per tile cell than
// 8 rings (48 KiB, three tiles): the kernel runs 11 % faster at 1024x3072.
// FP64 storage (64 KiB at 8 rings) stays at 8 rings, where 16 measured
// 15 % slower.  768 threads fill 24 wavefronts per CU either way; the
// wavefronts hide the latency of the stages' LDS and global reads (256,
// 512 and 1024 threads measured slower).  The tile shape does not change
// the results.
template <typename T> struct SrcBlock {
  static constexpr int n = 768, tr = 16;
};
template <> struct SrcBlock<double> {
  static constexpr int n = 768, tr = 8;
};

// the 1-D tables k_source reads, staged per tile
enum { ST_IDRM, ST_RINF, ST_IRINF, ST_RMED, ST_SUPP, ST_IDXT, ST_NU, ST_RSUP, ST_IDRS, ST_IRMED, ST_N };
__device__ __forceinline__ const double *src_table(const KArgs &a, int k) {
  switch (k) {
    case ST_IDRM: return a.InvDiffRmed;
    case ST_RINF: return a.Rinf;
    case ST_IRINF: return a.InvRinf;
    case ST_RMED: return a.Rmed;
    case ST_SUPP: return a.supp_torque;
    case ST_IDXT: return a.InvDxTheta;
    case ST_NU: return a.nu_ring;
    case ST_RSUP: return a.Rsup;
    case ST_IDRS: return a.InvDiffRsup;
    default: return a.InvRmed;
  }
}

template <typename T> __device__ __forceinline__ double src_press(const KArgs &a, long l, int i) {
  if (!a.energy) {
    const double cs = (double)(T)a.cs_iso[i];
    return ld<T>(a.rho, l) * cs * cs;
  }
  return (a.gamma - 1.0) * ld<T>(a.en, l);
}

template <typename T> __global__ void __launch_bounds__(SrcBlock<T>::n) k_source(KArgs a) {
  __shared__ double sPot[SRC_PR][SRC_PC];
  // VradInt / VthetaInt in FP64 whatever the storage type: their sub-step
  // increments are kept in full until the final, unbiased rounding
  __shared__ double sVr[SRC_VR][SRC_VC], sVt[SRC_VR][SRC_VC];
  __shared__ T sTrr[SRC_SR][SRC_SC], sTpp[SRC_SR][SRC_SC], sTrp[SRC_SR][SRC_SC];
  // the state the stages read, staged once: density and vtheta as stored,
  // pressure in FP64 (slot = V slot + 1 = S slot + 2 in both directions)
  __shared__ T sRho[SRC_PR][SRC_FC], sVth[SRC_PR][SRC_FC];
  __shared__ double sP[SRC_PR][SRC_FC];
  __shared__ int sCol[SRC_FC];
  // the 1-D tables on the staged rings (zero outside the grid)
  __shared__ double sTab[ST_N][SRC_PR];
  const dfargo_ctrl *c = a.ctrl;
  if (!c->active) return;
  constexpr int TB = SrcBlock<T>::n;
  const int ns = a.ns, nr = a.nr, tid = threadIdx.x;
  const int j0 = blockIdx.x * SRC_TC, i0 = blockIdx.y * SRC_TR;
  const double dt = c->dt;
#define STAB(k, i) sTab[k][(i) - i0 + 4]
  // the particles of this sub-step were re-sorted into the other set
  if (a.d_on && blockIdx.x == 0 && blockIdx.y == 0 && tid == 0) a.ctrl->pset ^= 1;
  // The global values stage 1 reads at its cells (rings i0-3 .. i0+SRC_TR+1,
  // columns j0-3 .. j0+SRC_TC+1): loaded first, so that their latency
  // overlaps the staging below.
  constexpr int S1_N = SRC_VR * SRC_VC, S1_IT = (S1_N + TB - 1) / TB;
  T pf_vr[S1_IT], pf_sgr[S1_IT], pf_sgrm[S1_IT], pf_sgt[S1_IT], pf_sgtm[S1_IT];
#pragma unroll
  for (int it = 0; it < S1_IT; it++) {
    const int idx = tid + it * TB, q = idx / SRC_VC, p = idx - q * SRC_VC, i = i0 - 3 + q;
    pf_vr[it] = pf_sgr[it] = pf_sgrm[it] = pf_sgt[it] = pf_sgtm[it] = (T)0;
    if (idx < S1_N && i >= 0 && i <= nr - 1) {
      int j = (j0 - 3 + p) % ns, jm = (j0 - 4 + p) % ns;
      j = (j < 0) ? j + ns : j;
      jm = (jm < 0) ? jm + ns : jm;
      const long l = (long)i * ns + j;
      if (i >= 1) {
        pf_vr[it] = ((const T *)a.vr)[l];
        if (a.sg) {
          pf_sgr[it] = ((const T *)a.sgr)[l];
          pf_sgrm[it] = ((const T *)a.sgr)[l - ns];
        }
      }
      if (a.sg) {
        pf_sgt[it] = ((const T *)a.sgt)[l];
        pf_sgtm[it] = ((const T *)a.sgt)[(long)i * ns + jm];
      }
    }
  }
  // columns j0-4 .. j0+SRC_TC+2 of the staged region, wrapped
  for (int p = tid; p < SRC_FC; p += TB) {
    int jj = (j0 - 4 + p) % ns;
    sCol[p] = (jj < 0) ? jj + ns : jj;
  }
  for (int idx = tid; idx < ST_N * SRC_PR; idx += TB) {
    const int k = idx / SRC_PR, q = idx - k * SRC_PR, r = i0 - 4 + q;
    sTab[k][q] = (r >= 0 && r <= nr - 1) ? src_table(a, k)[r] : 0.0;
  }
  __syncthreads();
  // stage 0: the state of rings i0-4 .. i0+SRC_TR+1 into LDS, and the
  // potential there (rings outside the grid hold zeros no valid cell reads)
  const bool frame_trip = a.d_on && a.d_np > 0 && c->omega_prev != 0.0;
  for (int idx = tid; idx < SRC_PR * SRC_FC; idx += TB) {
    const int q = idx / SRC_FC, p = idx - q * SRC_FC, r = i0 - 4 + q;
    const bool in = (r >= 0 && r <= nr - 1);
    const long l = (long)r * ns + sCol[p];
    sRho[q][p] = in ? ((const T *)a.rho)[l] : (T)0;
    if (in && frame_trip) {
      // the particles' interpolation holds vt + Rmed OmegaFrame in the gas
      // grid and subtracts it again (src/Dsys.c:316,1082): SubStep1 reads
      // the round trip's result
      const double rw = sTab[ST_RMED][q] * c->omega_prev;
      sVth[q][p] = (T)((ld<T>(a.vt, l) + rw) - rw);
    } else {
      sVth[q][p] = in ? ((const T *)a.vt)[l] : (T)0;
    }
    sP[q][p] = in ? src_press<T>(a, l, r) : 0.0;
    if (p < SRC_PC) sPot[q][p] = in ? pot_at(a, c, r, sCol[p]) : 0.0;
  }
  __syncthreads();
  // the change of the corotating frame's rate this sub-step
  const double dom = a.frame ? c->domega : 0.0;
  // stage 1: SubStep1 on rings i0-3 .. i0+SRC_TR+1 (VradInt and VthetaInt
  // of ring nrad are zero, as in the reference where they are never set)
#pragma unroll
  for (int it = 0; it < S1_IT; it++) {
    const int idx = tid + it * TB;
    if (idx >= S1_N) break;
    const int q = idx / SRC_VC, p = idx - q * SRC_VC, i = i0 - 3 + q, j = sCol[p + 1];
    if (i < 0 || i > nr - 1) {
      sVr[q][p] = 0.0;
      sVt[q][p] = 0.0;
      continue;
    }
    const long l = (long)i * ns + j;
    const double rl = (double)sRho[q + 1][p + 1], pl = sP[q + 1][p + 1], potl = sPot[q + 1][p + 1];
    // the particles' grids (pressure gradients, indirect and self-gravity
    // accelerations, src/SourceEuler.c:893-948) on the tile's own cells
    const int jt = j0 - 3 + p;
    const bool dgrid = a.d_on && q >= 3 && q < 3 + SRC_TR && jt >= j0 && jt < j0 + SRC_TC && jt < ns;
    double vri = 0.0;
    if (i >= 1) {
      const double plim = sP[q][p + 1], idrm = STAB(ST_IDRM, i);
      double gradp = (pl - plim) * 2.0 / (rl + (double)sRho[q][p + 1]) * idrm;
      double gradphi = (potl - sPot[q][p + 1]) * idrm;
      double vt2;
      if (a.frame) {
        // CorrectVtheta (src/SideEuler.c:1133-1148) on the values read
        const double ci = dom * STAB(ST_RMED, i), cim = dom * STAB(ST_RMED, i - 1);
        vt2 = ((double)sVth[q + 1][p + 1] - ci) + ((double)sVth[q + 1][p + 2] - ci) +
              ((double)sVth[q][p + 1] - cim) + ((double)sVth[q][p + 2] - cim);
      } else {
        vt2 = (double)sVth[q + 1][p + 1] + (double)sVth[q + 1][p + 2] + (double)sVth[q][p + 1] +
              (double)sVth[q][p + 2];
      }
      vt2 = vt2 / 4.0 + STAB(ST_RINF, i) * c->omega_frame;
      vt2 = vt2 * vt2;
      vri = (double)pf_vr[it] + dt * (-gradp - gradphi + vt2 * STAB(ST_IRINF, i));
      if (a.d_on && a.d_feedback) vri += dt * ld<T>(a.radfbacc, l);
      if (DF_ON(T, a) && a.df_feedback) {
        // the dust fluid's drag on the gas, implicit (src/SourceEuler.c:915-921)
        const double ts = a.dstokes[l] * a.r15[i], eps = a.drho[l] / rl;
        vri = (vri * ts + eps * a.dvr[l] * dt) / (ts + dt * eps);
      }
      if (dgrid) {
        st<T>(a.radgradp, l, (pl - plim) * idrm);
        if (a.indirect_term)
          st<T>(a.radindacc, l, -(indpot_at(a, c, i, j) - indpot_at(a, c, i - 1, j)) * idrm);
      }
      if (a.sg) {
        double radsgacc = ((STAB(ST_RINF, i) - STAB(ST_RMED, i - 1)) * (double)pf_sgr[it] +
                           (STAB(ST_RMED, i) - STAB(ST_RINF, i)) * (double)pf_sgrm[it]) * idrm;
        if (dgrid) st<T>(a.radsgacc, l, radsgacc);
        vri = vri + dt * radsgacc;
      }
    }
    const double invdxtheta = STAB(ST_IDXT, i);
    const double pljm = sP[q + 1][p];
    double gradp = (pl - pljm) * 2.0 / (rl + (double)sRho[q + 1][p]) * invdxtheta;
    double gradphi = (potl - sPot[q + 1][p]) * invdxtheta;
    double vti = (double)sVth[q + 1][p + 1];
    if (a.frame) vti = vti - dom * STAB(ST_RMED, i);
    vti = vti - dt * (gradp + gradphi);
    vti += dt * STAB(ST_SUPP, i);
    if (a.d_on && a.d_feedback) vti += dt * ld<T>(a.azifbacc, l);
    if (DF_ON(T, a) && a.df_feedback) {
      // src/SourceEuler.c:991-997, with the dust's vtheta after CorrectVtheta
      const double ts = a.dstokes[l] * a.r15[i], eps = a.drho[l] / rl;
      const double dvtl = a.frame ? a.dvt[l] - dom * STAB(ST_RMED, i) : a.dvt[l];
      vti = (vti * ts + eps * dvtl * dt) / (ts + dt * eps);
    }
    if (dgrid) {
      st<T>(a.azigradp, l, (pl - pljm) * invdxtheta);
      if (a.indirect_term)
        st<T>(a.aziindacc, l, -(indpot_at(a, c, i, j) - indpot_at(a, c, i, sCol[p])) * invdxtheta);
    }
    if (a.sg) {
      const double azisgacc = 0.5 * ((double)pf_sgt[it] + (double)pf_sgtm[it]);
      if (dgrid) st<T>(a.azisgacc, l, azisgacc);
      vti = vti + dt * azisgacc;
    }
    sVr[q][p] = vri;
    sVt[q][p] = vti;
  }
  __syncthreads();
  if (a.visc_on) {
    // stage 2: stresses on rings i0-2 .. i0+SRC_TR (V slot = S slot + 1)
    const double onethird = 1.0 / 3.0;
    for (int idx = tid; idx < SRC_SR * SRC_SC; idx += TB) {
      const int q = idx / SRC_SC, p = idx - q * SRC_SC, i = i0 - 2 + q, j = sCol[p + 2];
      if (i < 0 || i > nr - 1) {
        sTrr[q][p] = sTpp[q][p] = sTrp[q][p] = (T)0;
        continue;
      }
      const double nu = STAB(ST_NU, i), irm = STAB(ST_IRMED, i), idrs = STAB(ST_IDRS, i);
      const double vrl = (double)sVr[q + 1][p + 1], vrp = (double)sVr[q + 2][p + 1];
      const double vtl = (double)sVt[q + 1][p + 1], vtp = (double)sVt[q + 1][p + 2];
      double dv = (vrp * STAB(ST_RSUP, i) - vrl * STAB(ST_RINF, i)) * idrs * irm;
      dv += (vtp - vtl) * a.invdphi * irm;
      const double Drr = (vrp - vrl) * idrs;
      const double Dpp = (vtp - vtl) * a.invdphi * irm + 0.5 * (vrp + vrl) * irm;
      const double rho = (double)sRho[q + 2][p + 2];
      sTrr[q][p] = (T)(2.0 * rho * nu * (Drr - onethird * dv));
      sTpp[q][p] = (T)(2.0 * rho * nu * (Dpp - onethird * dv));
      double Trp = 0.0;
      if (i >= 1) {
        const double Drp = 0.5 * (STAB(ST_RINF, i) * (vtl * irm - (double)sVt[q][p + 1] * STAB(ST_IRMED, i - 1)) *
                                      STAB(ST_IDRM, i) +
                                  (vrl - (double)sVr[q + 1][p]) * a.invdphi * STAB(ST_IRINF, i));
        Trp = 2.0 * 0.25 *
              (rho + (double)sRho[q + 1][p + 2] + (double)sRho[q + 2][p + 1] + (double)sRho[q + 1][p + 1]) * nu *
              Drp;
      }
      sTrp[q][p] = (T)Trp;
    }
    __syncthreads();
    // stage 3: the viscous update in place, VthetaInt on rings i0 ..
    // i0+SRC_TR-1 x columns j0-1 .. j0+SRC_TC, VradInt on rings i0-1 ..
    // i0+SRC_TR x columns j0 .. j0+SRC_TC-1 (S slot = V slot - 1)
    for (int idx = tid; idx < SRC_TR * (SRC_TC + 2); idx += TB) {
      const int q = idx / (SRC_TC + 2), p = idx - q * (SRC_TC + 2), i = i0 + q;
      const int qv = q + 3, pv = p + 2, qs = qv - 1, ps = pv - 1;
      if (i >= 1 && i < nr - 1) {
        const double rho = (double)sRho[qv + 1][pv + 1];
        const double trpl = (double)sTrp[qs][ps], trpp = (double)sTrp[qs + 1][ps];
        double v = (double)sVt[qv][pv];
        v += dt * STAB(ST_IRMED, i) *
             ((STAB(ST_RSUP, i) * trpp - STAB(ST_RINF, i) * trpl) * STAB(ST_IDRS, i) +
              ((double)sTpp[qs][ps] - (double)sTpp[qs][ps - 1]) * a.invdphi + 0.5 * (trpl + trpp)) /
             (0.5 * (rho + (double)sRho[qv + 1][pv]));
        sVt[qv][pv] = v;
      }
    }
    for (int idx = tid; idx < (SRC_TR + 2) * SRC_TC; idx += TB) {
      const int q = idx / SRC_TC, p = idx - q * SRC_TC, i = i0 - 1 + q;
      const int qv = q + 2, pv = p + 3, qs = qv - 1, ps = pv - 1;
      if (i >= 1 && i <= nr - 1) {
        const double rho = (double)sRho[qv + 1][pv + 1];
        double v = (double)sVr[qv][pv];
        v += dt * STAB(ST_IRINF, i) *
             ((STAB(ST_RMED, i) * (double)sTrr[qs][ps] - STAB(ST_RMED, i - 1) * (double)sTrr[qs - 1][ps]) *
                  STAB(ST_IDRM, i) +
              ((double)sTrp[qs][ps + 1] - (double)sTrp[qs][ps]) * a.invdphi -
              0.5 * ((double)sTpp[qs][ps] + (double)sTpp[qs - 1][ps])) /
             (0.5 * (rho + (double)sRho[qv][pv + 1]));
        sVr[qv][pv] = v;
      }
    }
    __syncthreads();
  }
  // stage 4: SubStep2 on the tile (V slot = tile slot + 3).  Each
  // wavefront holds one ring of the tile (SRC_TC = 64 columns) and leaves
  // the tile's sum of the elongation rates vt dt / dt of the stored vtheta
  // (ComputeThetaElongations and ComputeAverageThetaVelocities,
  // src/TransportEuler.c:249-290), from which the azimuthal transport forms
  // the ring mean.
  const double C2 = DF_CVNR * DF_CVNR, invdt = 1.0 / dt;
  const int ntile = (ns + SRC_TC - 1) / SRC_TC;
  for (int b0 = 0; b0 < SRC_TR * SRC_TC; b0 += TB) {
    const int idx = b0 + tid;
    const int q = idx / SRC_TC, p = idx - q * SRC_TC, i = i0 + q, jt = j0 + p;
    double rate = 0.0;
    if (idx < SRC_TR * SRC_TC && i <= nr - 1 && jt < ns) {
      const int qv = q + 3, pv = p + 3, j = sCol[pv + 1];
      const long l = (long)i * ns + j;
      const double rho = (double)sRho[qv + 1][pv + 1];
      const double vrl = (double)sVr[qv][pv], vrp = (double)sVr[qv + 1][pv];
      const double vtl = (double)sVt[qv][pv], vtp = (double)sVt[qv][pv + 1], vtm = (double)sVt[qv][pv - 1];
      double dv = vrp - vrl;
      const double qr = (dv < 0.0) ? C2 * rho * dv * dv : 0.0;
      dv = vtp - vtl;
      const double qt = (dv < 0.0) ? C2 * rho * dv * dv : 0.0;
      const double rhojm = (double)sRho[qv + 1][pv];
      dv = vtl - vtm;
      const double qtm = (dv < 0.0) ? C2 * rhojm * dv * dv : 0.0;
      double vrn = 0.0;
      if (i >= 1) {
        const double rhoim = (double)sRho[qv][pv + 1];
        dv = vrl - (double)sVr[qv - 1][pv];
        const double qrm = (dv < 0.0) ? C2 * rhoim * dv * dv : 0.0;
        vrn = vrl - dt * 2.0 / (rho + rhoim) * (qr - qrm) * STAB(ST_IDRM, i);
      }
      const double invdxtheta = STAB(ST_IDXT, i);
      st_sr<T>(a.vri, l, vrn, sr_key(c, l, SR_VR));
      const T vts = sr_round<T>(vtl - dt * 2.0 / (rho + rhojm) * (qt - qtm) * invdxtheta, sr_key(c, l, SR_VT));
      ((T *)a.vti)[l] = vts;
      rate = ((double)vts * dt) * invdt;
      if (a.energy) {
        double e = ld<T>(a.en, l);
        e = e - dt * qr * (vrp - vrl) * STAB(ST_IDRS, i) - dt * qt * (vtp - vtl) * invdxtheta;
        st_sr<T>(a.e_int, l, e, sr_key(c, l, SR_E));
      }
      if (i == nr - 1) {
        // vr and vt of ring nrad are zero after ActualiseGas
        st<T>(a.vr, l + ns, 0.0);
        st<T>(a.vt, l + ns, 0.0);
      }
    }
    rate = wave_sum(rate);
    if ((tid & 63) == 0 && idx < SRC_TR * SRC_TC && i <= nr - 1) a.vt_part[(long)i * ntile + blockIdx.x] = rate;
  }
}
#undef STAB

// Energy source terms, per ring: ComputeDivergenceVelocity
// (src/SourceEuler.c:2373-2411), the stresses and ComputeViscousHeating
// (src/SourceEuler.c:1554-1606) on the post-SubStep2 velocities, then
// SubStep3 (src/SourceEuler.c:1147-1198) with beta cooling
template <typename T>
__device__ double vischeat_cell(const KArgs &a, int i, int j, double *trr, double *tpp,
                                double *trp, double *divv) {
  visc_stress_cell<T>(a, a.vri, a.vti, i, j, trr, tpp, trp, divv);
  const double viscosity = a.nu_ring[i];
  if (viscosity == 0.0) return 0.0;
  const double dens = ld<T>(a.rho, (long)i * a.ns + j);
  double vh = 0.5 / viscosity / dens * (*trr * *trr + 2.0 * *trp * *trp + *tpp * *tpp);
  vh += (2.0 / 9.0) * viscosity * dens * *divv * *divv;
  return vh;
}

template <typename T, int K> __global__ void __launch_bounds__(RING_LB(K)) k_energy(KArgs a) {
  __shared__ double sh[1024];
  const dfargo_ctrl *c = a.ctrl;
  if (!c->active) return;
  const int i = ring_of_block(a), ns = a.ns, nr = a.nr, t = threadIdx.x, TB = blockDim.x;
  const long base = (long)i * ns;
  const double dt = c->dt;
  // the stresses, divergence and heating are diagnostics read between
  // AlgoGas calls: stored on the last sub-step of a call only
  const bool last = !c->go_next;
  double vh[K], dv[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int j = t + k * TB;
    double trr, tpp, trp, d;
    if (a.viscous_heating && i >= 1) {
      vh[k] = vischeat_cell<T>(a, i, j, &trr, &tpp, &trp, &d);
    } else {
      visc_stress_cell<T>(a, a.vri, a.vti, i, j, &trr, &tpp, &trp, &d);
      vh[k] = 0.0;
    }
    if (a.viscous_heating && i == 0) {
      // power-law extrapolation from rings 1 and 2; the test reads the
      // viscosity left by the ring loop, i.e. that of ring nrad-1
      double t1, t2, t3, d1;
      double q1 = vischeat_cell<T>(a, 1, j, &t1, &t2, &t3, &d1);
      double q2 = vischeat_cell<T>(a, 2, j, &t1, &t2, &t3, &d1);
      vh[k] = (a.nu_ring[nr - 1] != 0.0) ? q1 * exp(log(q1 / q2) * a.vh_lr1 / a.vh_lr2) : 0.0;
    }
    dv[k] = d;
    if (last) {
      st<T>(a.trr, base + j, trr);
      st<T>(a.tpp, base + j, tpp);
      st<T>(a.trp, base + j, trp);
      st<T>(a.divv, base + j, d);
      st<T>(a.vh, base + j, vh[k]);
    }
  }
  double enew[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const long l = base + t + k * TB;
    double num = ld<T>(a.e_int, l) + (a.viscous_heating ? 1 : 0) * vh[k] * dt;
    double den = 1.0 + (a.gamma - 1.0) * dt * dv[k];
    enew[k] = num / den;
  }
  if (a.beta_cooling) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < K; k++) s += (ld<T>(a.vti, base + t + k * TB) + a.Rmed[i] * c->omega_frame);
    s = block_sum_fast(s, sh);
    double omega = s / (double)ns;
    omega /= a.Rmed[i];
    double coolingtime = a.beta[i] / omega;
#pragma unroll
    for (int k = 0; k < K; k++) {
      double num = enew[k], den = 1.0 + dt / coolingtime;
      enew[k] = num / den;
    }
  }
  // heating by the dust feedback (src/SourceEuler.c:1291-1299)
  if (a.d_on && a.d_feedback) {
#pragma unroll
    for (int k = 0; k < K; k++) enew[k] += ld<T>(a.fbedot, base + t + k * TB) * dt;
  }
#pragma unroll
  for (int k = 0; k < K; k++) st_sr<T>(a.en, base + t + k * TB, enew[k], sr_key(c, base + t + k * TB, SR_EN));
}

// ------------------------------------------------------------------------
// transport
// ------------------------------------------------------------------------

// Radial transport (ComputeLRMomenta + OneWindRad, src/TransportEuler.c:
// 89-220) on tiles of RT_R rings x RT_C columns.  The Work values
// Q/(rho+1e-20) of the tile's rings and of two halo rings either side, then
// the ComputeStarRad slopes, go through LDS, so each is formed once per cell
// (plus the halo); every thread then updates one cell.  Quantity slots:
// 0 rp, 1 rm, 2 tp, 3 tm, 4 e, 5 rho (the Work of the density itself),
// 6 rho as its own base for RhoStar.  The arithmetic is that of the storage
// type R (the reference's double with FP64 storage).
#define RT_C 32
#define RT_R 8

// slope of ComputeStarRad at ring r (zero on the first and last ring)
template <typename T, typename R = typename TransportArith<T>::type>
__device__ __forceinline__ R rad_slope(const KArgs &a, int r, R qm, R q0, R qp) {
  if (r <= 0 || r >= a.nr - 1) return R(0);
  R dqm = (q0 - qm) * (R)a.InvDiffRmed[r];
  R dqp = (qp - q0) * (R)a.InvDiffRmed[r + 1];
  // branch-free: the quotient is formed on every lane and selected (where
  // it is not used it may be an infinity or a NaN of a zero denominator,
  // which the select discards; nothing traps)
  const bool use = dqp * dqm > R(0);
  const R d = TransportArith<T>::div(R(2) * dqp * dqm, dqp + dqm);
  return use ? d : R(0);
}
// The upwind side of face r (between rings r-1 and r) for velocity v and
// the weight of its slope in QRStar: (Rmed[r]-Rmed[r-1]-v dt)/2 from ring
// r-1 when v > 0, else -(Rmed[r+1]-Rmed[r]+v dt)/2 from ring r (the same
// products and sums as the reference's two branches); shared by every
// quantity at the face.  Faces 0 and nr have no star (open = false).
template <typename R> struct RadFace {
  bool open, up;
  R w;
};
template <typename R> __device__ __forceinline__ RadFace<R> rad_face(const KArgs &a, int r, R v, R dt) {
  RadFace<R> f;
  f.open = (r > 0 && r < a.nr);
  f.up = v > R(0);
  const int rc = f.open ? r : 1; // keeps the table reads in range on closed faces
  const R wu = ((R)a.Rmed[rc] - (R)a.Rmed[rc - 1] - v * dt) * R(0.5);
  const R wd = ((R)a.Rmed[rc + 1] - (R)a.Rmed[rc] + v * dt) * R(0.5);
  f.w = f.up ? wu : -wd;
  return f;
}
// QRStar at face r from the Work values of rings r-1, r and the slopes there
template <typename R>
__device__ __forceinline__ R rad_star(const RadFace<R> &f, R wm, R w0, R dqm, R dq0) {
  const R s = (f.up ? wm : w0) + f.w * (f.up ? dqm : dq0);
  return f.open ? s : R(0);
}
// the six transported quantities of cell (r, j) before the pass
template <typename T, typename R>
__device__ __forceinline__ void rad_quantities(const KArgs &a, int r, int j, R Of, R (&Q)[6]) {
  const int ns = a.ns;
  const int jp = (j == ns - 1) ? 0 : j + 1;
  const T *rho = (const T *)a.rho + (long)r * ns, *vr = (const T *)a.vri + (long)r * ns;
  const T *vt = (const T *)a.vti + (long)r * ns;
  const R d = (R)rho[j], Rm = (R)a.Rmed[r];
  Q[0] = d * (R)vr[j + ns];
  Q[1] = d * (R)vr[j];
  Q[2] = d * ((R)vt[jp] + Rm * Of) * Rm;
  Q[3] = d * ((R)vt[j] + Rm * Of) * Rm;
  Q[4] = a.energy ? (R)((const T *)a.en)[(long)r * ns + j] : R(0);
  Q[5] = d;
}

template <typename T> __global__ void __launch_bounds__(RT_C * RT_R) k_transport_rad(KArgs a) {
  using R = typename TransportArith<T>::type;
  __shared__ R sW[7][RT_R + 4][RT_C];
  __shared__ R sD[7][RT_R + 2][RT_C];
  const dfargo_ctrl *c = a.ctrl;
  if (!c->active) return;
  const int ns = a.ns, nr = a.nr, tx = threadIdx.x, ty = threadIdx.y;
  const int j = blockIdx.x * RT_C + tx, i0 = blockIdx.y * RT_R;
  const int jc = (j < ns) ? j : ns - 1; // columns past the edge only fill the tile
  const R dt = (R)c->dt, Of = (R)c->omega_frame, dtheta = (R)(a.pm / (double)ns);
  // Work values of rings i0-2 .. i0+RT_R+1
  for (int q = ty; q < RT_R + 4; q += RT_R) {
    const int r = i0 - 2 + q;
    if (r < 0 || r > nr - 1) {
#pragma unroll
      for (int k = 0; k < 7; k++) sW[k][q][tx] = R(0);
      continue;
    }
    R Q[6];
    rad_quantities<T, R>(a, r, jc, Of, Q);
    // the Work values Q / (rho + 1e-20) share one reciprocal per cell
    const R inv = TransportArith<T>::div(R(1), Q[5] + R(1e-20));
#pragma unroll
    for (int k = 0; k < 6; k++) sW[k][q][tx] = Q[k] * inv;
    sW[6][q][tx] = Q[5];
  }
  __syncthreads();
  // slopes at rings i0-1 .. i0+RT_R
  for (int q = ty; q < RT_R + 2; q += RT_R) {
    const int r = i0 - 1 + q;
#pragma unroll
    for (int k = 0; k < 7; k++) sD[k][q][tx] = rad_slope<T>(a, r, sW[k][q][tx], sW[k][q + 1][tx], sW[k][q + 2][tx]);
  }
  __syncthreads();
  const int i = i0 + ty;
  if (j >= ns || i >= nr) return;
  const long l = (long)i * ns + j;
  const R v0 = (R)((const T *)a.vri)[l], v1 = (R)((const T *)a.vri)[l + ns];
  // stars at faces i (Work of rings i-1, i: slots ty+1, ty+2; slopes of
  // rings i-1, i: slots ty, ty+1) and i+1
  const RadFace<R> f0 = rad_face<R>(a, i, v0, dt), f1 = rad_face<R>(a, i + 1, v1, dt);
  const R rs0 = rad_star<R>(f0, sW[6][ty + 1][tx], sW[6][ty + 2][tx], sD[6][ty][tx], sD[6][ty + 1][tx]);
  const R rs1 = rad_star<R>(f1, sW[6][ty + 2][tx], sW[6][ty + 3][tx], sD[6][ty + 1][tx], sD[6][ty + 2][tx]);
  R Q[6];
  rad_quantities<T, R>(a, i, j, Of, Q);
  const R fin = dt * dtheta * (R)a.Rinf[i], fout = dt * dtheta * (R)a.Rsup[i], invs = (R)a.InvSurf[i];
  void *outp[6] = {a.t_rp, a.t_rm, a.t_tp, a.t_tm, a.t_en, a.t_rho};
#pragma unroll
  for (int k = 0; k < 6; k++) {
    if (k == 4 && !a.energy) continue;
    if (k < 4 && DF_ON(T, a) && a.tr_skip_mom) {
      ((T *)outp[k])[l] = (T)Q[k];
      continue;
    }
    const R q0 = rad_star<R>(f0, sW[k][ty + 1][tx], sW[k][ty + 2][tx], sD[k][ty][tx], sD[k][ty + 1][tx]);
    const R q1 = rad_star<R>(f1, sW[k][ty + 2][tx], sW[k][ty + 3][tx], sD[k][ty + 1][tx], sD[k][ty + 2][tx]);
    // VanLeerRadial (src/TransportEuler.c:205-217)
    R varq = fin * q0 * rs0 * v0;
    varq -= fout * q1 * rs1 * v1;
    ((T *)outp[k])[l] = sr_round<T>((double)Q[k] + (double)(varq * invs), sr_key(c, l, SR_RAD + k));
    if (k == 5 && i == 0) a.lost_col[j] = (double)varq;
  }
}

// Azimuthal transport (OneWindTheta, src/TransportEuler.c:225-465) on ring
// segments, one wavefront each: the residual and uniform passes and the
// shift.  The ring quantities of the FARGO scheme come first, in FP64
// whatever the storage type, from the per-tile sums k_source leaves of the
// velocity before transport: the mean VMed
// (ComputeAverageThetaVelocities), the integer shift Nshift and the
// constant residual (ComputeConstantResidual, src/TransportEuler.c:292-348).
// Every wavefront of a ring forms them from the same sums in the same
// order, so all agree on the shift.

// The cells a wavefront owns: its 64 lanes hold 64 consecutive cells of a
// ring, and each of the two passes leaves two fewer valid cells at either
// end (a cell's update reads its neighbours up to two cells away), so the
// middle 56 are complete after both.
#define AZ_LANES 64
#define AZ_HALO 4
#define AZ_OWN (AZ_LANES - 2 * AZ_HALO)

// The neighbouring lanes' values: lane L receives lane L - 1's (prev) or
// lane L + 1's (next); the end lanes receive their own, which only feeds
// cells outside the valid range.  CDNA's whole-wavefront DPP shifts
// (wave_shr:1, wave_shl:1) move each 32-bit half in one VALU operation,
// without the LDS crossbar of a general lane permute.
template <int CTRL> __device__ __forceinline__ double dpp_shift(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xf, 0xf, false),
                          __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xf, 0xf, false));
}
#define DPP_WAVE_SHL1 0x130
#define DPP_WAVE_SHR1 0x138
__device__ __forceinline__ double lane_prev(double v) { return dpp_shift<DPP_WAVE_SHR1>(v); }
__device__ __forceinline__ double lane_next(double v) { return dpp_shift<DPP_WAVE_SHL1>(v); }

// ComputeStarTheta (src/TransportEuler.c:419-444) at this lane's cell (the
// face between the previous cell and this one), for the values x, the
// velocity v of the cell and the time step dt
template <typename T, typename R = typename TransportArith<T>::type>
__device__ __forceinline__ R azi_star_lane(R x, R v, R dt, R dxtheta, R invdxtheta) {
  const R xm = lane_prev(x), xp = lane_next(x);
  const R dqm = x - xm, dqp = xp - x;
  const R dq = (dqp * dqm > R(0)) ? TransportArith<T>::div(dqp * dqm, dqp + dqm) * invdxtheta : R(0);
  const R dqprev = lane_prev(dq);
  const R ksi = v * dt;
  if (ksi > R(0)) return xm + (dxtheta - ksi) * dqprev;
  return x - (dxtheta + ksi) * dq;
}

// QuantitiesAdvection (src/TransportEuler.c:375-400) with velocity v: the
// density star first, then every quantity's star, flux and update.  Q[5]
// is the density.  The flux through this lane's face is dxrad S RS v; the
// cell's update is its own face's flux minus the next lane's.
template <typename T, typename R = typename TransportArith<T>::type>
__device__ __forceinline__ void azi_pass_lane(const KArgs &a, R (&Q)[6], R v, R dt, R dxtheta, R invdxtheta,
                                              R dxrad, R invsurf) {
  const R RI = Q[5];
  const R RS = azi_star_lane<T>(RI, v, dt, dxtheta, invdxtheta);
  // the Work values Q / (rho + 1e-20) share one reciprocal
  const R inv = TransportArith<T>::div(R(1), RI + R(1e-20));
#pragma unroll
  for (int q = 0; q < 6; q++) {
    if (q == 4 && !a.energy) continue;
    if (q < 4 && DF_ON(T, a) && a.tr_skip_mom) continue;
    const R S = azi_star_lane<T>(Q[q] * inv, v, dt, dxtheta, invdxtheta);
    const R f = dxrad * S * RS * v;
    const R fn = lane_next(f);
    R varq = f;
    varq -= fn;
    Q[q] += varq * invsurf;
  }
}

template <typename T> __global__ void __launch_bounds__(256) k_transport_azi(KArgs a) {
  using R = typename TransportArith<T>::type;
  const dfargo_ctrl *c = a.ctrl;
  if (!c->active) return;
  const int ns = a.ns, i = blockIdx.y, lane = threadIdx.x & (AZ_LANES - 1);
  const int seg = blockIdx.x * (blockDim.x / AZ_LANES) + threadIdx.x / AZ_LANES;
  if (seg * AZ_OWN >= ns) return; // a whole wavefront
  int j = (seg * AZ_OWN - AZ_HALO + lane) % ns;
  if (j < 0) j += ns;
  const long base = (long)i * ns;
  const double dt = c->dt, invdt = 1.0 / dt;
  // the ring's mean velocity, shift and residual
  const int ntile = (ns + DF_SRC_COLS - 1) / DF_SRC_COLS;
  double sum = 0.0;
  for (int t = lane; t < ntile; t += AZ_LANES) sum += a.vt_part[(long)i * ntile + t];
  const double vmed = wave_sum(sum) / (double)ns;
  const double dpinvns = a.pm / (double)ns;
  const double Ntilde = vmed * a.InvRmed[i] * dt * (double)ns / a.pm;
  const double Nround = floor(Ntilde + 0.5);
  const long nshift = (long)Nround;
  const double u = (Ntilde - Nround) * a.Rmed[i] * invdt * dpinvns;
  if (seg == 0 && lane == 0) {
    a.vmed[i] = vmed;
    a.nshift[i] = nshift;
    // distance of Ntilde from the nearest rounding tie, min over the call
    // (the gas's shift)
    const double margin = 0.5 - fabs(Ntilde - Nround);
    if (!(DF_ON(T, a) && a.tr_dust))
      atomicMin(&a.ctrl->fargo_margin_bits, (unsigned long long)__double_as_longlong(margin > 0.0 ? margin : 0.0));
  }
  // mass through the inner edge (VanLeerRadial's LostByDisk, ring 0;
  // LostMassd for the dust fluid)
  if (i == 0 && seg == 0) {
    double lost = 0.0;
    for (int jj = lane; jj < ns; jj += AZ_LANES) lost += a.lost_col[jj];
    lost = wave_sum(lost);
    if (lane == 0) {
      if (DF_ON(T, a) && a.tr_dust) a.ctrl->lost_massd += lost;
      else a.ctrl->lost_mass += lost;
    }
  }
  // quantities after the radial pass: rp, rm, tp, tm, [e], rho; the
  // results go to z_*, since other wavefronts of the ring may still read
  // the cells this one stores
  const void *qp[6] = {a.t_rp, a.t_rm, a.t_tp, a.t_tm, a.t_en, a.t_rho};
  void *zp[6] = {a.z_rp, a.z_rm, a.z_tp, a.z_tm, a.z_en, a.z_rho};
  R Q[6];
#pragma unroll
  for (int q = 0; q < 6; q++) Q[q] = (q == 4 && !a.energy) ? R(0) : (R)((const T *)qp[q])[base + j];
  // the residual velocity of the cell (src/TransportEuler.c:280-290)
  double v = (ld<T>(a.vti, base + j) * dt) * invdt - vmed;
  if (!a.fast) v = u + v;
  const double dxth = a.pm / (double)ns * a.Rmed[i];
  const R dxtheta = (R)dxth, invdxtheta = (R)(1.0 / dxth);
  const R dxrad = (R)((a.Rsup[i] - a.Rinf[i]) * dt), invsurf = (R)(1.0 / a.Surf[i]), rdt = (R)dt;
  // residual pass (the only pass of the standard transport), then the
  // uniform pass with the constant residual
  azi_pass_lane<T>(a, Q, (R)v, rdt, dxtheta, invdxtheta, dxrad, invsurf);
  if (a.fast) azi_pass_lane<T>(a, Q, (R)u, rdt, dxtheta, invdxtheta, dxrad, invsurf);
  // AdvectSHIFT (src/TransportEuler.c:350-373): new[j] = old[j - Nshift]
  const int own = lane - AZ_HALO;
  if (own < 0 || own >= AZ_OWN || seg * AZ_OWN + own >= ns) return;
  long sh_mod = nshift % ns;
  if (sh_mod < 0) sh_mod += ns;
  long jn = j + sh_mod;
  if (jn >= ns) jn -= ns;
#pragma unroll
  for (int q = 0; q < 6; q++) {
    if (q == 4 && !a.energy) continue;
    ((T *)zp[q])[base + jn] = sr_round<T>((double)Q[q], sr_key(c, base + jn, SR_AZI + q));
  }
}

// ComputeVelocities (src/TransportEuler.c:515-547) and the floors
// (src/SourceEuler.c:2506-2541) of one cell from the transported momenta,
// rounded to the storage type as the stored state would be.  The
// velocities are quotients formed here, so with FP32 storage they are
// rounded stochastically like every other state store (keyed by the cell,
// so an edge ring that copies a neighbour's value recomputes exactly the
// value stored there); density and energy are already stored values.
struct CellState {
  double rho, vr, vt, e;
};
template <typename T> __device__ CellState vel_cell(const KArgs &a, const dfargo_ctrl *c, int i, int j) {
  const int ns = a.ns;
  const long l = (long)i * ns + j, lim = l - ns, ljm = (j == 0) ? (long)i * ns + ns - 1 : l - 1;
  const double rho = ld<T>(a.z_rho, l);
  CellState s;
  s.vr = 0.0;
  if (i != 0) s.vr = (ld<T>(a.z_rp, lim) + ld<T>(a.z_rm, l)) / (rho + ld<T>(a.z_rho, lim) + 1e-20);
  s.vt = (ld<T>(a.z_tp, ljm) + ld<T>(a.z_tm, l)) / (rho + ld<T>(a.z_rho, ljm) + 1e-20) / a.Rmed[i] -
         a.Rmed[i] * c->omega_frame;
  s.rho = rho;
  if (a.floors && s.rho < a.floor_dens) s.rho = a.floor_dens;
  s.e = 0.0;
  if (a.energy) {
    s.e = ld<T>(a.z_en, l);
    const double e_bottom = 1e-3 * a.EnergyMed[i];
    if (a.floors && s.e < e_bottom) s.e = e_bottom;
  }
  s.rho = (double)(T)s.rho;
  s.vr = (double)sr_round<T>(s.vr, sr_key(c, l, SR_FIN));
  s.vt = (double)sr_round<T>(s.vt, sr_key(c, l, SR_FIN + 1));
  s.e = (double)(T)s.e;
  return s;
}

// End of the sub-step: ComputeVelocities and the floors, then
// ApplyBoundaryCondition (src/SideEuler.c:1102-1131) -- wall or open edges
// and the sub-Keplerian reset of vt in rings 0 and nrad-1, which runs for
// every boundary type (src/SideEuler.c:1110-1112).  Each cell of the edge
// rings takes the values the boundary routines would copy into it from the
// rings they read (2, 3 and nrad-2), recomputed here from the momenta.
// Stores cell (i, j) of rings 0 .. nrad-1 and returns its density as
// stored.
template <typename T> __device__ __forceinline__ double finish_cell(const KArgs &a, const dfargo_ctrl *c, int i,
                                                                    int j) {
  const int nr = a.nr;
  const long l = (long)i * a.ns + j;
  CellState s = vel_cell<T>(a, c, i, j);
  if (a.bc == DFARGO_BC_WALL) {
    if (i == 0) {
      s.vr = -vel_cell<T>(a, c, 2, j).vr;
      s.rho = a.SigmaMed[0];
    } else if (i == 1) {
      s.vr = 0.0;
      s.rho = a.SigmaMed[1];
    }
    if (i == nr - 2) s.rho = a.SigmaMed[nr - 2];
    if (i == nr - 1) {
      s.vr = 0.0;
      s.rho = a.SigmaMed[nr - 1];
    }
  }
  if (i == 0) s.vt = a.VthetaMed[0] - a.Rmed[0] * c->omega_frame;
  if (i == nr - 1) s.vt = a.VthetaMed[nr - 1] - a.Rmed[nr - 1] * c->omega_frame;
  if (a.bc == DFARGO_BC_OPEN) {
    // OpenBoundary (src/SideEuler.c:300-408), standard prescription
    if (i <= 2) {
      const CellState s2 = (i == 2) ? s : vel_cell<T>(a, c, 2, j);
      const double vrp = vel_cell<T>(a, c, 3, j).vr;
      const double vr2 = ((vrp >= 0.0) || (s2.rho < a.SigmaMed[0])) ? 0.0 : vrp;
      if (i == 2) s.vr = vr2;
      if (i == 1) {
        s.rho = s2.rho;
        s.e = s2.e;
        s.vt = s2.vt * a.open_vt_in1[0];
        s.vr = vr2;
      }
      if (i == 0) {
        s.rho = s2.rho;
        s.e = s2.e;
        s.vt = s2.vt * a.open_vt_in2[0];
      }
    }
    if (i == nr - 1) {
      const CellState sm = vel_cell<T>(a, c, nr - 2, j);
      s.rho = sm.rho;
      s.e = sm.e;
      s.vt = sm.vt * a.open_vt_out[0];
      s.vr = ((sm.vr < 0.0) || (sm.rho < a.SigmaMed[nr - 2])) ? 0.0 : sm.vr;
    }
  }
  st<T>(a.rho, l, s.rho);
  st<T>(a.vr, l, s.vr);
  st<T>(a.vt, l, s.vt);
  if (a.energy) st<T>(a.en, l, s.e);
  return (double)(T)s.rho;
}
// the wall's radial velocity at the outer face (ring nrad)
template <typename T> __device__ __forceinline__ void finish_outer_face(const KArgs &a, const dfargo_ctrl *c, int j) {
  if (a.bc == DFARGO_BC_WALL) st<T>(a.vr, (long)a.nr * a.ns + j, -vel_cell<T>(a, c, a.nr - 2, j).vr);
}

// Rings 0 .. nrad, the last only for the wall's vr at the outer face.
template <typename T> __global__ void __launch_bounds__(256) k_finish(KArgs a) {
  const dfargo_ctrl *c = a.ctrl;
  if (!c->active) return;
  // the sub-step ends here: the next one's loop condition takes effect
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) a.ctrl->go = c->go_next;
  const int i = blockIdx.y, ns = a.ns, nr = a.nr, j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ns) return;
  if (i == nr) {
    finish_outer_face<T>(a, c, j);
    return;
  }
  finish_cell<T>(a, c, i, j);
}

// The same with self-gravity, one workgroup per ring: after storing its
// ring the workgroup forms the ring's reduced densities from the values it
// stored and transforms them (compute_fftdensity and the azimuthal FFT of
// the next sub-step, src/sgdens.c:54-70), so that the next sub-step's
// self-gravity starts from the radial transforms.  The workgroup of ring
// nrad-1 also stores the outer face.
template <typename T> __global__ void __launch_bounds__(SG_TB_MAX) k_finish_rows(KArgs a, cx<T> *zt) {
  extern __shared__ unsigned char smem[];
  const dfargo_ctrl *c = a.ctrl;
  if (!c->active) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) a.ctrl->go = c->go_next;
  const int i = ring_of_block(a), ns = a.ns, nr = a.nr;
  cx<T> *A = (cx<T> *)smem, *B = A + fpsz(ns);
  for (int j = threadIdx.x; j < ns; j += blockDim.x) {
    A[fp(j)] = sg_reduced<T>(a, i, finish_cell<T>(a, c, i, j));
    if (i == nr - 1) finish_outer_face<T>(a, c, j);
  }
  sg_row_transform<T>(a, i, A, B, zt);
}

// EvanescentBoundary (src/SideEuler.c:796-950) for the gas, after the
// boundary conditions: in the wave-killing zones Rmed < WKZRMIN and
// Rmed > WKZRMAX each ring relaxes as q <- (q + lambda q0) / (1 + lambda)
// towards its azimuthal mean (DampToAxi, FP64 sums) or the initial profile
// (DampToIni), lambda = x^2 dt / (ratio Rmed^1.5) with x rising from 0 at
// the zone's edge to 1 at the grid's edge.  With ZEROVRADBC the two
// innermost and the outermost ring take vr = 0 and the target density and
// vtheta.  One 256-thread workgroup per ring of the zones; each thread
// reads and rewrites the same cells, so the update is in place.
enum { EV_TB = 256 };

// SolveOneDViscousEvolution (src/SideEuler.c:1518-1565): one explicit step
// of the 1-D viscous disc the DAMPTOVISCOUS zones relax towards -- the
// diffusion dSigma/dt = 3[(nu Sigma)'' + 3/(2R)(nu Sigma)'] on the
// non-uniform 1-D grid, its radial velocity, then the boundary rule.  One
// workgroup; the profile is staged in LDS, updated there, written back.
enum { V1D_MAX = 1024, V1D_TB = 256 };
__global__ void __launch_bounds__(V1D_TB) k_visc1d(KArgs a) {
  const dfargo_ctrl *c = a.ctrl;
  if (!c->active) return;
  __shared__ double sS[V1D_MAX], sV[V1D_MAX], sN[V1D_MAX], sR[V1D_MAX], sT[V1D_MAX];
  const int n = a.n1d, tid = threadIdx.x;
  const double dt = c->dt;
  for (int i = tid; i < n; i += V1D_TB) {
    sS[i] = a.v1d_sigma[i];
    sV[i] = a.v1d_vrad[i];
    sN[i] = a.v1d_nu[i];
    sR[i] = a.v1d_rmed[i];
  }
  __syncthreads();
  for (int i = 1 + tid; i <= n - 2; i += V1D_TB) {
    const double drm = sR[i] - sR[i - 1], drp = sR[i + 1] - sR[i];
    const double drnu = (drm * drm * sN[i + 1] - drp * drp * sN[i - 1] + (drp * drp - drm * drm) * sN[i]) /
                        (drm * drm * drp + drp * drp * drm);
    const double drsigma = (drm * drm * sS[i + 1] - drp * drp * sS[i - 1] + (drp * drp - drm * drm) * sS[i]) /
                           (drm * drm * drp + drp * drp * drm);
    const double d2rnu = (drm * sN[i + 1] + drp * sN[i - 1] - (drm + drp) * sN[i]) /
                         (0.5 * drm * drp * drp + 0.5 * drp * drm * drm);
    const double d2rsigma = (drm * sS[i + 1] + drp * sS[i - 1] - (drm + drp) * sS[i]) /
                            (0.5 * drm * drp * drp + 0.5 * drp * drm * drm);
    const double term1 = 3.0 * sS[i] * (1.5 * drnu / sR[i] + d2rnu);
    const double term2 = 3.0 * drsigma * (1.5 * sN[i] / sR[i] + 2.0 * drnu);
    const double term3 = 3.0 * sN[i] * d2rsigma;
    sT[i] = term1 + term2 + term3;
    sV[i] = -1.5 * sN[i] / sR[i] - 3.0 * drnu - 1.5 * (sN[i] / sS[i]) * drsigma;
  }
  __syncthreads();
  for (int i = 1 + tid; i <= n - 2; i += V1D_TB) sS[i] += dt * sT[i];
  __syncthreads();
  if (tid == 0) {
    if (a.bc1d == 1) {
      sS[0] = sS[1] * (sN[1] / sN[0]) * sqrt(sR[1] / sR[0]);
      sS[n - 1] = sS[n - 2] * (sN[n - 2] / sN[n - 1]) * sqrt(sR[n - 2] / sR[n - 1]);
    } else if (a.bc1d == 2) {
      sS[0] = sS[1] * (sN[1] / sN[0]);
      sS[n - 1] = sS[n - 2] * (sN[n - 2] / sN[n - 1]);
    } else {
      sS[0] = 1e-9;
      sS[n - 1] = 1e-9;
    }
    sV[0] = sV[1];
    sV[n - 1] = sV[n - 2];
  }
  __syncthreads();
  for (int i = tid; i < n; i += V1D_TB) {
    a.v1d_sigma[i] = sS[i];
    a.v1d_vrad[i] = sV[i];
  }
}
template <typename T> __global__ void __launch_bounds__(EV_TB) k_evanescent(KArgs a) {
  const dfargo_ctrl *c = a.ctrl;
  if (!c->active) return;
  const int b = blockIdx.x, nr = a.nr, ns = a.ns, tid = threadIdx.x;
  const int i = (b < a.wkz_n_in) ? b : a.wkz_i_out + (b - a.wkz_n_in);
  const long base = (long)i * ns;
  const double R = a.Rmed[i];
  const bool in_zone = R < a.wkz_rmin;
  const double damping = in_zone ? (R - a.wkz_rmin) / (a.Rmed[0] - a.wkz_rmin)
                                 : (R - a.wkz_rmax) / (a.Rmed[nr - 1] - a.wkz_rmax);
  const double Tz = a.damp_ratio * pow(R, 1.5);
  const double lambda = damping * damping * c->dt / Tz;
  double q0[4];
  if (a.damp_visc) {
    // the ring means of vtheta and e, Sigma and vr interpolated linearly in
    // the 1-D disc (src/SideEuler.c:868-887)
    __shared__ double partv[EV_TB / 64][2];
    double s[2] = {0.0, 0.0};
    for (int j = tid; j < ns; j += EV_TB) {
      s[0] += ld<T>(a.vt, base + j);
      if (a.energy) s[1] += ld<T>(a.en, base + j);
    }
#pragma unroll
    for (int k = 0; k < 2; k++) s[k] = wave_sum(s[k]);
    if ((tid & 63) == 0) {
      partv[tid >> 6][0] = s[0];
      partv[tid >> 6][1] = s[1];
    }
    __syncthreads();
    double sv = 0.0, se = 0.0;
    for (int w = 0; w < EV_TB / 64; w++) {
      sv += partv[w][0];
      se += partv[w][1];
    }
    q0[1] = sv / (double)ns;
    q0[3] = se / (double)ns;
    int m = 0;
    while (a.v1d_rmed[m] < R) m++;
    m--;
    const double r0 = a.v1d_rmed[m], r1 = a.v1d_rmed[m + 1];
    q0[2] = ((r1 - R) * a.v1d_sigma[m] + (R - r0) * a.v1d_sigma[m + 1]) / (r1 - r0);
    q0[0] = ((r1 - R) * a.v1d_vrad[m] + (R - r0) * a.v1d_vrad[m + 1]) / (r1 - r0);
  } else if (a.damp_ini) {
    q0[0] = a.VradMed[i];
    q0[1] = a.VthetaMed[i] - R * c->omega_frame;
    q0[2] = a.SigmaMed[i];
    q0[3] = a.EnergyMed[i];
  } else {
    __shared__ double part[EV_TB / 64][4];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int j = tid; j < ns; j += EV_TB) {
      s[0] += ld<T>(a.vr, base + j);
      s[1] += ld<T>(a.vt, base + j);
      s[2] += ld<T>(a.rho, base + j);
      if (a.energy) s[3] += ld<T>(a.en, base + j);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) s[k] = wave_sum(s[k]);
    if ((tid & 63) == 0)
#pragma unroll
      for (int k = 0; k < 4; k++) part[tid >> 6][k] = s[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; k++) {
      double t = 0.0;
      for (int w = 0; w < EV_TB / 64; w++) t += part[w][k];
      q0[k] = t / (double)ns;
    }
  }
  // the dust fluid's targets (src/SideEuler.c:844-848,887-926): its
  // initial profiles, or with DampToAxi its ring means; with the open dust
  // edge the fluid is not damped
  double dq0[3] = {0.0, 0.0, 0.0};
  const bool ddamp = DF_ON(T, a) && !a.df_open;
  if (ddamp) {
    if (!a.damp_ini && !a.damp_visc) {
      __shared__ double dpart[EV_TB / 64][3];
      double s[3] = {0.0, 0.0, 0.0};
      for (int j = tid; j < ns; j += EV_TB) {
        s[0] += a.dvr[base + j];
        s[1] += a.dvt[base + j];
        s[2] += a.drho[base + j];
      }
#pragma unroll
      for (int k = 0; k < 3; k++) s[k] = wave_sum(s[k]);
      if ((tid & 63) == 0)
#pragma unroll
        for (int k = 0; k < 3; k++) dpart[tid >> 6][k] = s[k];
      __syncthreads();
#pragma unroll
      for (int k = 0; k < 3; k++) {
        double t = 0.0;
        for (int w = 0; w < EV_TB / 64; w++) t += dpart[w][k];
        dq0[k] = t / (double)ns;
      }
    } else {
      dq0[0] = a.DVradMed[i];
      dq0[1] = a.DVthetaMed[i] - R * c->omega_frame;
      dq0[2] = a.DSigmaMed[i];
    }
  }
  const bool edge = a.zero_vrad && ((in_zone && (i == 0 || i == 1)) || (!in_zone && i == nr - 1));
  for (int j = tid; j < ns; j += EV_TB) {
    const long l = base + j;
    if (ddamp) {
      a.dvr[l] = (a.dvr[l] + lambda * dq0[0]) / (1.0 + lambda);
      a.dvt[l] = (a.dvt[l] + lambda * dq0[1]) / (1.0 + lambda);
      a.drho[l] = (a.drho[l] + lambda * dq0[2]) / (1.0 + lambda);
    }
    double vr = (ld<T>(a.vr, l) + lambda * q0[0]) / (1.0 + lambda);
    double vt = (ld<T>(a.vt, l) + lambda * q0[1]) / (1.0 + lambda);
    double rho = (ld<T>(a.rho, l) + lambda * q0[2]) / (1.0 + lambda);
    if (edge) {
      vr = 0.0;
      rho = q0[2];
      vt = q0[1];
    }
    st_sr<T>(a.vr, l, vr, sr_key(c, l, SR_EVA));
    st_sr<T>(a.vt, l, vt, sr_key(c, l, SR_EVA + 1));
    st_sr<T>(a.rho, l, rho, sr_key(c, l, SR_EVA + 2));
    if (a.energy) st_sr<T>(a.en, l, (ld<T>(a.en, l) + lambda * q0[3]) / (1.0 + lambda), sr_key(c, l, SR_EVA + 3));
  }
}

// NonReflectingBoundary (src/SideEuler.c:569-794), barotropic case without
// the accreting variant, after the sub-Keplerian reset (which it leaves
// alone).  Workgroup 0 does the inner edge, workgroup 1 the outer one.
// Each edge takes a pattern speed from the innermost (outermost) planet
// inside the grid, turns it into an integer azimuthal shift i_angle over the
// first (last) ring spacing at the mean sound speed, and
//   inner: rho(0, j + i_angle) = rho(1, j), shifted by a constant so that
//          ring 0's mean is SigmaMed[0]; vr(1, j) = 2 vr_med - vr(2, j) with
//          vr_med = -cs1 (rho(1, j) - <rho(1)>) / <rho(1)>;
//   outer: rho(nr-1, j) = rho(nr-2, j - i_angle) + SigmaMed[nr-1] - <...>;
//          vr(nr-1, j) = 2 vr_med - vr(nr-2, j), vr_med = +cs (...).
// The new edge ring is a cyclic shift of its neighbour, so its mean is the
// neighbour's: the constant is formed from that mean.  The sound speed is
// the isothermal ring value, the mean of its identical cells.
enum { NR_TB = 256 };

// block sums of N values per thread, in a fixed order (wavefront tree, then
// the wavefronts in order); every thread gets the totals
template <int N> __device__ void nr_block_sums(double (&v)[N], double (*part)[NR_TB / 64]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < N; q++) {
    const double w = wave_sum(v[q]);
    if ((tid & 63) == 0) part[q][tid >> 6] = w;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < N; q++) {
    double s = 0.0;
    for (int w = 0; w < NR_TB / 64; w++) s += part[q][w];
    v[q] = s;
  }
}

// the pattern speed: the innermost (outermost) planet inside the grid (the
// loops of src/SideEuler.c:590-600 and 702-712; the reference leaves r
// undefined when planet 0 is outside the grid, here the grid's edge stands
// in)
__device__ double nr_pattern_omega(const KArgs &a, const dfargo_ctrl *c, bool inner) {
  const double r0 = a.Rinf[0], r1 = a.Rsup[a.nr - 1];
  double buf = inner ? r1 : r0, rp_sel = buf, omega = pow(buf, -1.5);
  for (int k = 0; k < c->nb; k++) {
    const double xp = c->pl_x[k], yp = c->pl_y[k], rp = sqrt(xp * xp + yp * yp);
    if (inner ? ((rp < buf) && (rp >= r0)) : ((rp > buf) && (rp <= r1))) {
      rp_sel = rp;
      omega = pow(rp, -1.5);
    }
    buf = rp_sel;
  }
  return omega;
}

__device__ __forceinline__ int nr_wrap(int j, int ns) {
  int jp = j % ns;
  return (jp < 0) ? jp + ns : jp;
}

// NonReflectingBoundary (src/SideEuler.c:569-794) without the accreting
// variant, after the sub-Keplerian reset (which it leaves alone).
// Workgroup 0 does the inner edge, workgroup 1 the outer one.  Each edge
// takes a pattern speed from the innermost (outermost) planet inside the
// grid and turns it into an integer azimuthal shift i_angle over the first
// (last) ring spacing at a mean sound speed.
//   barotropic: rho(0, j + i_angle) = rho(1, j) and
//     rho(nr-1, j) = rho(nr-2, j - i_angle), each shifted by a constant so
//     that the edge ring's mean is SigmaMed; vr at the faces 1 and nr-1 is
//     2 vr_med - vr of the next face in, vr_med = -/+ cs drho / <rho>;
//   adiabatic (inner edge): the energy is copied with the shift, the
//     density follows it acoustically, SigmaMed[0] + (gamma-1)(e -
//     EnergyMed[0]) / csinit0^2, and both are shifted to the initial means;
//     vr_med uses the initial profiles; the outer edge is the barotropic one
//     with the energy copied and shifted to EnergyMed too and vr_med from the
//     cell's own sound speed.
// A shifted edge ring is a cyclic copy of its neighbour, so its mean is the
// neighbour's: the constants are formed from that mean.  The sound speed is
// the isothermal ring value (the mean of its identical cells), or, with the
// energy equation, the grid k_eos_step has just formed from the state the
// boundary sees (src/SideEuler.c:1121-1122).
template <typename T> __global__ void __launch_bounds__(NR_TB) k_nonreflecting(KArgs a) {
  const dfargo_ctrl *c = a.ctrl;
  if (!c->active) return;
  const bool inner = blockIdx.x == 0;
  const int nr = a.nr, ns = a.ns, tid = threadIdx.x;
  // the ring the edge copies from, and the edge ring
  const int src = inner ? 1 : nr - 2, dst = inner ? 0 : nr - 1;
  const long bs = (long)src * ns, bd = (long)dst * ns;
  __shared__ double part[4][NR_TB / 64];
  const double omega = nr_pattern_omega(a, c, inner);
  const double g = a.gamma;
  if (a.energy && inner) {
    const double csinit0 = sqrt(g) * a.aspect_ratio * pow(a.Rmed[0], -0.5 + a.flaring);
    const double csinit1 = sqrt(g) * a.aspect_ratio * pow(a.Rmed[1], -0.5 + a.flaring);
    double dangle = (pow(a.Rinf[1], -1.5) - omega) / (.5 * (csinit0 + csinit1));
    dangle *= (a.Rmed[1] - a.Rmed[0]);
    const int i_angle = (int)(dangle / a.pm * (double)ns + .5);
    double s[2] = {0.0, 0.0};
    for (int j = tid; j < ns; j += NR_TB) {
      const double e = ld<T>(a.en, bs + j);
      s[0] += e;
      s[1] += a.SigmaMed[0] + (g - 1.0) * (e - a.EnergyMed[0]) / csinit0 / csinit0;
    }
    nr_block_sums<2>(s, part);
    const double mean_e = s[0] / (double)ns, mean_rho = s[1] / (double)ns;
    for (int j = tid; j < ns; j += NR_TB) {
      const double e = ld<T>(a.en, bs + j);
      const double rho = a.SigmaMed[0] + (g - 1.0) * (e - a.EnergyMed[0]) / csinit0 / csinit0;
      const int jp = nr_wrap(j + i_angle, ns);
      st<T>(a.en, jp, e + (a.EnergyMed[0] - mean_e));
      st<T>(a.rho, jp, rho + (a.SigmaMed[0] - mean_rho));
      const double vr_med = -csinit1 * (ld<T>(a.rho, bs + j) - a.SigmaMed[1]) / a.SigmaMed[1];
      st<T>(a.vr, bs + j, 2. * vr_med - ld<T>(a.vr, bs + j + ns));
    }
    return;
  }
  // ring means: rho of the source ring; the sound speed of both rings
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int j = tid; j < ns; j += NR_TB) {
    s[0] += ld<T>(a.rho, bs + j);
    if (a.energy) {
      s[1] += ld<T>(a.cs, bd + j);
      s[2] += ld<T>(a.cs, bs + j);
      s[3] += ld<T>(a.en, bs + j);
    }
  }
  nr_block_sums<4>(s, part);
  const double mean = s[0] / (double)ns;
  // isothermal: the ring value, the mean of the ring's identical cells
  const double cs_a = a.energy ? s[1] / (double)ns : a.cs_iso[dst];
  const double cs_b = a.energy ? s[2] / (double)ns : a.cs_iso[src];
  double dangle = (pow(a.Rinf[inner ? 1 : nr - 2], -1.5) - omega) / (.5 * (cs_a + cs_b));
  dangle *= inner ? (a.Rmed[1] - a.Rmed[0]) : (a.Rmed[nr - 1] - a.Rmed[nr - 2]);
  const int i_angle = (int)(dangle / a.pm * (double)ns + .5);
  const double shift = a.SigmaMed[dst] - mean;
  const double eshift = a.EnergyMed[nr - 1] - s[3] / (double)ns;
  for (int j = tid; j < ns; j += NR_TB) {
    const long ls = bs + j;
    const double rs = ld<T>(a.rho, ls);
    if (inner) {
      const double vr_med = -a.cs_iso[1] * (rs - mean) / mean;
      st<T>(a.rho, nr_wrap(j + i_angle, ns), rs + shift);
      st<T>(a.vr, ls, 2. * vr_med - ld<T>(a.vr, ls + ns));
    } else {
      const long l = bd + j;
      const long lp = bs + nr_wrap(j - i_angle, ns);
      // the outer face's own sound speed with the energy equation
      const double csv = a.energy ? ld<T>(a.cs, l) : a.cs_iso[nr - 1];
      const double vr_med = csv * (rs - mean) / mean;
      st<T>(a.rho, l, ld<T>(a.rho, lp) + shift);
      if (a.energy) st<T>(a.en, l, ld<T>(a.en, lp) + eshift);
      st<T>(a.vr, l, 2. * vr_med - ld<T>(a.vr, ls));
    }
  }
}

// ComputeSoundSpeed inside ApplyBoundaryCondition (src/SideEuler.c:
// 1121-1122), for the non-reflecting boundary with the energy equation: cs
// of the state after transport and floors, which the boundary and the next
// sub-step's CFL condition read
template <typename T, int K> __global__ void __launch_bounds__(RING_LB(K)) k_eos_step(KArgs a) {
  if (!a.ctrl->active) return;
  const int i = ring_of_block(a), ns = a.ns, t = threadIdx.x, TB = blockDim.x;
  const long base = (long)i * ns;
  const double g = a.gamma;
#pragma unroll
  for (int k = 0; k < K; k++) {
    const long l = base + t + k * TB;
    st<T>(a.cs, l, sqrt(g * (g - 1.0) * ld<T>(a.en, l) / ld<T>(a.rho, l)));
  }
}

// Potential grid on demand (the field of FillForcesArrays)
template <typename T> __global__ void k_potential(KArgs a) {
  const int i = blockIdx.y, ns = a.ns, j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ns) return;
  st<T>(a.pot, (long)i * ns + j, pot_at(a, a.ctrl, i, j));
}

// ------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------

static int ring_k(int ns) {
  static const int ks[] = {1, 2, 3, 4, 6, 8, 12, 16};
  for (int k : ks)
    if (ns % k == 0 && ns / k <= RING_LB(k)) return k;
  return -1;
}

int ring_kcheck(int ns) { return ring_k(ns); }

// Six cells per thread (four waves per ring) for the ring reductions of
// the CFL kernel: its block reductions then run over fewer waves, measured
// faster than two or three cells per thread at more waves.  Rings of fewer
// than 1536 cells keep the smallest K.
static int wide_k(int ns) {
  if (ns % 6 == 0 && ns / 6 >= 256 && ns / 6 <= RING_LB(6)) return 6;
  return ring_k(ns);
}

#define RING_DISPATCH(KERNEL, T, a, s, lds, ...) RING_DISPATCH_K(ring_k((a).ns), KERNEL, T, a, s, lds, ##__VA_ARGS__)
#define RING_DISPATCH_K(KSEL, KERNEL, T, a, s, lds, ...)                                    \
  do {                                                                                      \
    int K_ = (KSEL);                                                                        \
    dim3 g_((a).nr), b_((a).ns / K_);                                                       \
    switch (K_) {                                                                           \
      case 1: hipLaunchKernelGGL((KERNEL<T, 1>), g_, b_, lds, s, a, ##__VA_ARGS__); break;  \
      case 2: hipLaunchKernelGGL((KERNEL<T, 2>), g_, b_, lds, s, a, ##__VA_ARGS__); break;  \
      case 3: hipLaunchKernelGGL((KERNEL<T, 3>), g_, b_, lds, s, a, ##__VA_ARGS__); break;  \
      case 4: hipLaunchKernelGGL((KERNEL<T, 4>), g_, b_, lds, s, a, ##__VA_ARGS__); break;  \
      case 6: hipLaunchKernelGGL((KERNEL<T, 6>), g_, b_, lds, s, a, ##__VA_ARGS__); break;  \
      case 8: hipLaunchKernelGGL((KERNEL<T, 8>), g_, b_, lds, s, a, ##__VA_ARGS__); break;  \
      case 12: hipLaunchKernelGGL((KERNEL<T, 12>), g_, b_, lds, s, a, ##__VA_ARGS__); break; \
      case 16: hipLaunchKernelGGL((KERNEL<T, 16>), g_, b_, lds, s, a, ##__VA_ARGS__); break; \
    }                                                                                       \
  } while (0)

#include "dust_fluid.h"

static dim3 cell_grid(const KArgs &a, int rings) {
  int bx = a.ns < 256 ? a.ns : 256;
  return dim3((a.ns + bx - 1) / bx, rings);
}
static dim3 cell_block(const KArgs &a) { return dim3(a.ns < 256 ? a.ns : 256); }
// the velocity/boundary kernel's workgroup: the tuned width, at most a ring
static int finish_bx(const KArgs &a) {
  const int w = (a.tune_finish_block == 64 || a.tune_finish_block == 128) ? a.tune_finish_block : 256;
  return a.ns < w ? a.ns : w;
}

template <typename T> void launch_kernel(int id, const KArgs &a, hipStream_t s) {
  switch (id) {
    case KID_CFL: RING_DISPATCH_K(wide_k(a.ns), k_cfl_ring, T, a, s, 0); break;
    case KID_CONTROL:
      if (a.accrete) hipLaunchKernelGGL((k_control<T, true>), dim3(1), dim3(CtrlBlock<true>::n), 0, s, a);
      else hipLaunchKernelGGL((k_control<T, false>), dim3(1), dim3(CtrlBlock<false>::n), 0, s, a);
      break;
    case KID_EOS_STEP: RING_DISPATCH(k_eos_step, T, a, s, 0); break;
    case KID_VISC1D: hipLaunchKernelGGL(k_visc1d, dim3(1), dim3(V1D_TB), 0, s, a); break;
    case KID_NONREFLECTING:
      hipLaunchKernelGGL(k_nonreflecting<T>, dim3(2), dim3(NR_TB), 0, s, a);
      break;
    case KID_EVANESCENT:
      hipLaunchKernelGGL(k_evanescent<T>, dim3(a.wkz_n_in + (a.nr - a.wkz_i_out)), dim3(EV_TB), 0, s, a);
      break;
    case KID_SOURCE:
      hipLaunchKernelGGL(k_source<T>, dim3((a.ns + SRC_TC - 1) / SRC_TC, (a.nr + SRC_TR - 1) / SRC_TR), dim3(SrcBlock<T>::n), 0, s,
                         a);
      break;
    case KID_ENERGY: RING_DISPATCH(k_energy, T, a, s, 0); break;
    case KID_TRANSPORT_RAD: {
      dim3 g((a.ns + RT_C - 1) / RT_C, (a.nr + RT_R - 1) / RT_R);
      hipLaunchKernelGGL(k_transport_rad<T>, g, dim3(RT_C, RT_R), 0, s, a);
      break;
    }
    case KID_TRANSPORT_AZI: {
      const int per = (a.tune_azi_waves == 1 || a.tune_azi_waves == 2) ? a.tune_azi_waves : 4;
      const int nseg = (a.ns + AZ_OWN - 1) / AZ_OWN;
      hipLaunchKernelGGL(k_transport_azi<T>, dim3((nseg + per - 1) / per, a.nr), dim3(per * AZ_LANES), 0, s, a);
      break;
    }
    case KID_FINISH: {
      if (a.sg_rows_in_finish) {
        const size_t lds = sg_rows_lds_bytes(a, sizeof(cx<T>));
        hipLaunchKernelGGL(k_finish_rows<T>, dim3(a.nr), dim3(sg_threads(a, lds, a.nr, a.ns / 2)), lds, s, a,
                           (cx<T> *)a.sg_z);
        break;
      }
      const int bx = finish_bx(a);
      hipLaunchKernelGGL(k_finish<T>, dim3((a.ns + bx - 1) / bx, a.nr + 1), dim3(bx), 0, s, a);
      break;
    }
    case KID_DUST_MOVE:
    case KID_DUST_SCAN:
    case KID_DUST_SCATTER:
    case KID_DUST_SORT:
    case KID_DUST_DEPOSIT:
    case KID_DUST_REORDER: launch_dust_kernel<T>(id, a, s); break;
    case KID_DF_CFL:
    case KID_DF_AXISG:
    case KID_DF_SOURCE:
    case KID_DF_STRESS:
    case KID_DF_VISC:
    case KID_DF_SUBSTEP2:
    case KID_DF_TRANSPORT_RAD:
    case KID_DF_TRANSPORT_AZI:
    case KID_DF_FINISH:
      // the dust fluid runs with FP64 storage only (the context refuses it
      // otherwise)
      if (sizeof(T) == sizeof(double)) launch_df_kernel(id, a, s);
      break;
    default: launch_sg_kernel<T>(id, a, s); break;
  }
}

template <typename T> void launch_eos_pre(const KArgs &a, hipStream_t s) {
  RING_DISPATCH(k_eos, T, a, s, 0);
}
template <typename T> void launch_potential(const KArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_potential<T>, cell_grid(a, a.nr), cell_block(a), 0, s, a);
}
template <typename T> void launch_visc_stress_state(const KArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_visc_stress<T>, cell_grid(a, a.nr), cell_block(a), 0, s, a, 1);
}
void launch_algogas_begin(const KArgs &a, double DT, hipStream_t s) {
  hipLaunchKernelGGL(k_algogas_begin, dim3(1), dim3(64), 0, s, a, DT);
}

template <typename T> void prepare_kernels(const KArgs &a) {
  if (a.sg && !a.sg_zero) {
    sg_prepare<T>(a);
    // dynamic LDS above 64 KiB must be requested (gfx950 has 160 KiB per CU)
    const size_t lds = sg_rows_lds_bytes(a, sizeof(cx<T>));
    if (lds > 65536)
      hipFuncSetAttribute((const void *)k_finish_rows<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  }
}
template void prepare_kernels<double>(const KArgs &);
template void prepare_kernels<float>(const KArgs &);
template void launch_kernel<double>(int, const KArgs &, hipStream_t);
template void launch_kernel<float>(int, const KArgs &, hipStream_t);
template void launch_eos_pre<double>(const KArgs &, hipStream_t);
template void launch_eos_pre<float>(const KArgs &, hipStream_t);
template void launch_potential<double>(const KArgs &, hipStream_t);
template void launch_potential<float>(const KArgs &, hipStream_t);
template void launch_visc_stress_state<double>(const KArgs &, hipStream_t);
template void launch_visc_stress_state<float>(const KArgs &, hipStream_t);

// ------------------------------------------------------------------------
// per-DT diagnostics (src/main.c:291-394): outside AlgoGas, on the state
// the last call left
// ------------------------------------------------------------------------

// Per ring i, part[q * nr + i] for q = 0 GasTotalMass (src/SideEuler.c:
// 15-36), 1 GasMomentum (src/SideEuler.c:114-146), 2 GasTotalEnergy
// (src/SideEuler.c:149-180; the energy grid of an isothermal run holds
// zeros), then for body k, 3 + 4k + {0 fx, 1 fx Hill-cut, 2 fy, 3 fy
// Hill-cut} of ComputeForce (src/Force.c:155-214); the ring's inner/outer
// side is decided in k_diag_reduce.  Compensated FP64 sums, fixed block
// reduction order.
template <typename T, int K> __global__ void __launch_bounds__(RING_LB(K)) k_diag_ring(KArgs a, DiagBodies b,
                                                                                       double *part) {
  __shared__ double sh[1024];
  const dfargo_ctrl *c = a.ctrl;
  const int i = blockIdx.x, ns = a.ns, nr = a.nr, t = threadIdx.x, TB = blockDim.x;
  const long base = (long)i * ns;
  const double Rm = a.Rmed[i], S = a.Surf[i], Of = c->omega_frame;
  const double wr = (a.Rsup[i] - a.Rinf[i]);
  CSum m = {0, 0}, mom = {0, 0}, en = {0, 0};
  for (int k = 0; k < K; k++) {
    const int j = t + k * TB;
    const long l = base + j;
    const double d = ld<T>(a.rho, l);
    const double vtl = ld<T>(a.vt, l), vtp = ld<T>(a.vt, (j < ns - 1) ? l + 1 : base);
    const double vt_cent = 0.5 * (vtl + vtp) + Rm * Of;
    double vr_cent = (Rm - a.Rinf[i]) * ld<T>(a.vr, l + ns) + (a.Rsup[i] - Rm) * ld<T>(a.vr, l);
    vr_cent /= wr;
    const double e = a.energy ? ld<T>(a.en, l) : 0.0;
    m.add(S * d);
    mom.add(S * d * Rm * vt_cent);
    en.add(0.5 * S * d * (vr_cent * vr_cent + vt_cent * vt_cent) + S * e - S * d * pot_at(a, c, i, j));
  }
  double v[3] = {m.value(), mom.value(), en.value()};
  block_sum_n<3>(v, sh);
  if (t == 0)
    for (int q = 0; q < 3; q++) part[(long)q * nr + i] = v[q];
  for (int bd = 0; bd < b.n; bd++) {
    const double x = b.x[bd], y = b.y[bd], rsm = b.rsm[bd], cutoff = a.exclude_hill_factor * b.rh[bd];
    CSum f[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for (int k = 0; k < K; k++) {
      const int j = t + k * TB;
      const double xc = Rm * a.CosAz[j], yc = Rm * a.SinAz[j];
      const double cellmass = S * ld<T>(a.rho, base + j);
      const double dx = xc - x, dy = yc - y;
      const double d2 = dx * dx + dy * dy;
      const double planet_distance = sqrt(d2);
      const double dist2 = d2 + rsm * rsm;
      const double distance = sqrt(dist2);
      const double InvDist3 = 1.0 / dist2 / distance;
      double hill_cut;
      if (planet_distance / cutoff < 0.5) hill_cut = 0.0;
      else if (planet_distance > cutoff) hill_cut = 1.0;
      else hill_cut = pow(sin((planet_distance / cutoff - .5) * DF_PI), 2.);
      f[0].add(DF_G * cellmass * dx * InvDist3 * 1.0);
      f[1].add(DF_G * cellmass * dx * InvDist3 * hill_cut);
      f[2].add(DF_G * cellmass * dy * InvDist3 * 1.0);
      f[3].add(DF_G * cellmass * dy * InvDist3 * hill_cut);
    }
    double w[4] = {f[0].value(), f[1].value(), f[2].value(), f[3].value()};
    block_sum_n<4>(w, sh);
    if (t == 0)
      for (int q = 0; q < 4; q++) part[(long)(3 + 4 * bd + q) * nr + i] = w[q];
  }
}

// Ring sums in ring order (compensated): out[0..2] the budgets, then for
// body k out[3 + 8k ..] = fx inner, fx Hill-cut inner, fx outer, fx
// Hill-cut outer, fy inner, fy Hill-cut inner, fy outer, fy Hill-cut outer
// (GlobalForce's order; a ring is inner when Rmed < a_k).
__global__ void __launch_bounds__(64) k_diag_reduce(KArgs a, DiagBodies b, const double *part, double *out) {
  const int q = threadIdx.x, nr = a.nr;
  if (q < 3) {
    CSum s = {0, 0};
    for (int i = 0; i < nr; i++) s.add(part[(long)q * nr + i]);
    out[q] = s.value();
  }
  const int bd = q / 4, comp = q % 4;
  if (bd < b.n) {
    CSum in = {0, 0}, outer = {0, 0};
    const double *pp = part + (long)(3 + 4 * bd + comp) * nr;
    for (int i = 0; i < nr; i++) {
      if (a.Rmed[i] < b.a[bd]) in.add(pp[i]);
      else outer.add(pp[i]);
    }
    // comp: 0 fx, 1 fx Hill-cut, 2 fy, 3 fy Hill-cut -> GlobalForce slots
    const int axis = comp / 2, hill = comp % 2;
    out[3 + 8 * bd + 4 * axis + hill] = in.value();
    out[3 + 8 * bd + 4 * axis + 2 + hill] = outer.value();
  }
}

template <typename T> void launch_diagnostics(const KArgs &a, const DiagBodies &b, double *part, double *out,
                                              hipStream_t s) {
  RING_DISPATCH(k_diag_ring, T, a, s, 0, b, part);
  hipLaunchKernelGGL(k_diag_reduce, dim3(1), dim3(64), 0, s, a, b, (const double *)part, out);
}
template void launch_diagnostics<double>(const KArgs &, const DiagBodies &, double *, double *, hipStream_t);
template void launch_diagnostics<float>(const KArgs &, const DiagBodies &, double *, double *, hipStream_t);

