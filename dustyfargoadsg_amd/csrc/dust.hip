// dust.hip -- Lagrangian dust particles of one sub-step on CDNA4
// (SemiUpdateDustPositions, src/DustUpdate.c:6-138; UpdateDustVelocities,
// src/DustUpdate.c:267-435; interpolation, src/Dsys.c:265-1149; the
// start-up interpolation of src/main.c:287).
//
// One thread per particle moves it, finds its cell, gathers the gas
// quantities of its three stencils (unrolled per interpolation scheme),
// computes its Stokes number and drag, updates its velocity, stores its
// stencil weights and takes a slot in its cell's bucket.  The feedback
// deposit (the scatter half of interpolation) is turned into a gather so
// that no float atomics decide the summation order: the bucket counts are
// scanned ring by ring, the particles scattered to their slots, each
// bucket sorted by particle id, and each grid cell walks the buckets of
// the cells whose stencils can reach it in a fixed order.  Every sum is
// thus deterministic; against the reference's particle-index order only
// its rounding differs.  Last, the particles are copied in bucket
// order into the other of two particle sets: the next sub-step finds
// neighbouring particles in neighbouring memory, and its gathers and
// deposit read few cache lines per wavefront.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace {

template <typename T> __device__ __forceinline__ double ldd(const void *p, long i) {
  return (double)((const T *)p)[i];
}

// AspectRatio (src/Viscosity.c:107-122)
__device__ double d_aspect(const KArgs &a, double rad) {
  double ar = a.aspect_ratio;
  const double rmin = a.transition_radius - a.transition_width * a.aspect_ratio;
  const double rmax = a.transition_radius + a.transition_width * a.aspect_ratio;
  if (rad < rmin) ar *= a.transition_ratio;
  if ((rad >= rmin) && (rad <= rmax)) ar *= exp((rmax - rad) / (rmax - rmin) * log(a.transition_ratio));
  return ar;
}

// reference azimuth of column j of a stencil: AziInf for the vtheta one,
// the cell centre otherwise; columns past either end continue the grid
// (src/Dsys.c:411-425).  Returns the wrapped column.  Branch-free: the
// table load does not wait on the column's case.
__device__ __forceinline__ int d_col(const KArgs &a, int kind, int j, double &az) {
  const int ns = a.ns;
  const double half = 0.5 * (a.pm) / (double)ns;
  const int jc = j < 0 ? 0 : (j >= ns ? ns - 1 : j);
  const double a0 = (kind == 1) ? a.Azimuth[jc] - half : a.Azimuth[jc];
  az = (j >= 0 && j < ns) ? a0 : (j < 0 ? a0 + j * a.delta_phi : a0 + (j + 1 - ns) * a.delta_phi);
  return (j < 0) ? j + ns : (j >= ns ? j - ns : j);
}

// radial distance of a stencil node: to Rinf for the vr stencil, to Rmed
// otherwise, logarithmic on a logarithmic grid
// (0 outside the grid; computed branch-free so that the table loads of a
// stencil issue together)
__device__ __forceinline__ double d_dr(const KArgs &a, int kind, int i, double rp) {
  const bool inr = (i >= 0) && (i <= a.nr - 1);
  const int ic = i < 0 ? 0 : (i > a.nr - 1 ? a.nr - 1 : i);
  const double ref = (kind == 0) ? a.Rinf[ic] : a.Rmed[ic];
  const double d = a.log_grid ? fabs(log(rp / ref)) : fabs(rp - ref);
  return inr ? d : 0.0;
}

__device__ __forceinline__ bool d_in_grid(const KArgs &a, double r) {
  return (r >= a.Rinf[0]) && (r < a.Rsup[a.nr - 1]);
}

// keep an azimuth in [AziInf[0], AziSup[ns-1]] (src/DustUpdate.c:38-45)
__device__ __forceinline__ double d_wrap(const KArgs &a, double th) {
  if (th < a.aziinf0) th += (a.azisup_last - a.aziinf0);
  if (th > a.azisup_last) th -= (a.azisup_last - a.aziinf0);
  return th;
}

// cell of a particle inside the grid: ip is the last ring with
// Rinf[ip] <= r (the linear search of src/Dsys.c:363-365), found from the
// grid's spacing and corrected against the table, jp from the azimuth.
// False when jp falls outside the grid, which only an azimuth outside
// [AziInf[0], AziSup[ns-1]] can produce.
__device__ bool d_cell(const KArgs &a, double rp, double tp, int &ip, int &jp) {
  const int nr = a.nr;
  double g = a.log_grid ? log(rp / a.Rinf[0]) / a.d_ring_step : (rp - a.Rinf[0]) / a.d_ring_step;
  int i = (int)g;
  i = i < 0 ? 0 : (i > nr - 1 ? nr - 1 : i);
  while (i > 0 && rp < a.Rinf[i]) i--;
  while (i < nr - 1 && rp >= a.Rinf[i + 1]) i++;
  ip = i;
  jp = (int)floor(a.ns * (tp - a.aziinf0) / (a.azisup_last - a.aziinf0));
  if (jp == a.ns) jp = 0;
  return jp >= 0 && jp < a.ns;
}

// ------------------------------------------------------------------------
// stencils, specialised per interpolation scheme S: node counts, first
// node and the separable weights (wr depends on the radial distance only,
// wt on the azimuthal one)

__host__ __device__ constexpr int nodes_r(int S, int kind) {
  return S == DFARGO_INTERP_NGP ? (kind == 0 ? 2 : 1) : (S == DFARGO_INTERP_CIC ? 2 : 3);
}
__host__ __device__ constexpr int nodes_t(int S, int kind) {
  return S == DFARGO_INTERP_NGP ? (kind == 1 ? 2 : 1) : (S == DFARGO_INTERP_CIC ? 2 : 3);
}

// first node of stencil `kind` (0 vr, 1 vtheta, 2 cell-centred) for a
// particle in cell (ip, jp) (src/Dsys.c:371-406,456-488,620-639): the rings
// stay within ip-1 .. ip+1 and the columns within jp-1 .. jp+2, the
// neighbourhood the deposit gathers from
// (az_jp = Azimuth[jp], rmed_ip = Rmed[ip])
template <int S>
__device__ __forceinline__ void stencil_base(int kind, int ip, int jp, double rp, double tp, double az_jp,
                                             double rmed_ip, int &imin, int &jmin) {
  if (S == DFARGO_INTERP_NGP) {
    imin = ip;
    jmin = jp;
  } else if (S == DFARGO_INTERP_CIC) {
    int myip = ip, myjp = jp;
    if (kind != 1) myjp = (tp < az_jp) ? jp - 1 : jp;
    if (kind != 0) myip = (rp < rmed_ip) ? ip - 1 : ip;
    imin = myip;
    jmin = myjp;
  } else {
    int myjp = jp;
    if (kind == 1) myjp = (tp > az_jp) ? jp + 1 : jp;
    imin = ip - 1;
    jmin = myjp - 1;
  }
}

// the weights of src/Dsys.c:407-441,490-524,640-665, one factor at a time
template <int S> __device__ __forceinline__ double w_rad(int kind, double dr, double delta) {
  if (S == DFARGO_INTERP_NGP) return (kind == 0) ? ((dr < 0.5 * delta) ? 1.0 : 0.0) : 1.0;
  if (S == DFARGO_INTERP_CIC) return 1.0 - dr / delta;
  double w = 0.0;
  if (dr < 0.5 * delta) w = 0.75 - dr * dr / delta / delta;
  if ((dr >= 0.5 * delta) && (dr <= 1.5 * delta)) w = 0.5 * (1.5 - dr / delta) * (1.5 - dr / delta);
  if (dr > 1.5 * delta) w = 0.0;
  return w;
}
template <int S> __device__ __forceinline__ double w_azi(int kind, double dphi, double delta) {
  if (S == DFARGO_INTERP_NGP) return (kind == 1) ? ((dphi < 0.5 * delta) ? 1.0 : 0.0) : 1.0;
  if (S == DFARGO_INTERP_CIC) return 1.0 - dphi / delta;
  double w = 0.0;
  if (dphi < 0.5 * delta) w = 0.75 - dphi * dphi / delta / delta;
  if ((dphi >= 0.5 * delta) && (dphi <= 1.5 * delta)) w = 0.5 * (1.5 - dphi / delta) * (1.5 - dphi / delta);
  if (dphi > 1.5 * delta) w = 0.0;
  return w;
}

// weights of stencil KIND for the particle at (rp, tp) in cell (ip, jp),
// stored for the deposit when asked; returns the first node
template <int S, int KIND>
__device__ __forceinline__ void stencil_weights(const KArgs &a, int ip, int jp, double rp, double tp, double az_jp,
                                                double rmed_ip, double (&wr)[3], double (&wt)[3], int (&col)[3],
                                                int &imin, int &jmin, bool store, int k) {
  constexpr int NI = nodes_r(S, KIND), NJ = nodes_t(S, KIND);
  stencil_base<S>(KIND, ip, jp, rp, tp, az_jp, rmed_ip, imin, jmin);
#pragma unroll
  for (int q = 0; q < NI; q++) wr[q] = w_rad<S>(KIND, d_dr(a, KIND, imin + q, rp), a.delta_r);
#pragma unroll
  for (int q = 0; q < NJ; q++) {
    double az;
    col[q] = d_col(a, KIND, jmin + q, az);
    wt[q] = w_azi<S>(KIND, fabs(tp - az), a.delta_phi);
  }
  if (store) {
    const long np = a.d_np;
#pragma unroll
    for (int q = 0; q < NI; q++) a.p_w[(KIND * 6 + q) * np + k] = wr[q];
#pragma unroll
    for (int q = 0; q < NJ; q++) a.p_w[(KIND * 6 + 3 + q) * np + k] = wt[q];
  }
}

// Gathers of stencil KIND: up to four grids at its nodes, all loads issued
// before the sums, which then run in the reference's order (columns outer,
// rings inner) over the nodes inside the grid.  Grid g1 of the
// cell-centred stencil may be the per-ring isothermal sound speed; g2 and
// g3 are read always (a valid grid stands in for an unused one) and summed
// only when use2 / use3.  With a frame rate, g0 is vtheta and is read as
// vt + Rmed frame at each node.
template <typename T, int S, int KIND>
__device__ __forceinline__ void stencil_gather(const KArgs &a, const double (&wr)[3], const double (&wt)[3],
                                               const int (&col)[3], int imin, const void *g0, const void *g1,
                                               const void *g2, const void *g3, bool iso1, bool use2, bool use3,
                                               double &s0, double &s1, double &s2, double &s3,
                                               double frame = 0.0) {
  constexpr int NI = nodes_r(S, KIND), NJ = nodes_t(S, KIND);
  const int nr = a.nr, ns = a.ns;
  bool inr[NI];
  int ic[NI];
#pragma unroll
  for (int p = 0; p < NI; p++) {
    const int i = imin + p;
    inr[p] = (i >= 0) && (i <= nr - 1);
    ic[p] = i < 0 ? 0 : (i > nr - 1 ? nr - 1 : i);
  }
  double x0[NJ][NI], x1[NJ][NI], x2[NJ][NI], x3[NJ][NI];
#pragma unroll
  for (int q = 0; q < NJ; q++)
#pragma unroll
    for (int p = 0; p < NI; p++) {
      const long l = (long)ic[p] * ns + col[q];
      x0[q][p] = ldd<T>(g0, l);
      if (frame != 0.0) x0[q][p] = x0[q][p] + a.Rmed[ic[p]] * frame;
      x1[q][p] = iso1 ? (double)(T)((const double *)g1)[ic[p]] : ldd<T>(g1, l);
      x2[q][p] = ldd<T>(g2, l);
      x3[q][p] = ldd<T>(g3, l);
    }
#pragma unroll
  for (int q = 0; q < NJ; q++)
#pragma unroll
    for (int p = 0; p < NI; p++) {
      const double w = wr[p] * wt[q];
      s0 = inr[p] ? s0 + (w * x0[q][p]) : s0;
      s1 = inr[p] ? s1 + (w * x1[q][p]) : s1;
      if (use2) s2 = inr[p] ? s2 + (w * x2[q][p]) : s2;
      if (use3) s3 = inr[p] ? s3 + (w * x3[q][p]) : s3;
    }
}

// does this sub-step (or the start-up) need the buckets and the deposit?
__device__ __forceinline__ bool dust_deposit_needed(const KArgs &a, const dfargo_ctrl *c, int force) {
  return force || a.d_feedback || !c->go_next;
}

// One particle per thread.  In a sub-step: SemiUpdateDustPositions over
// 2 dt, interpolation (gathers, Stokes number, drag), UpdateDustVelocities
// and RotateDsys by OmegaFrame dt; the particle is counted into its cell's
// bucket at the position before the rotation, where the reference's
// interpolation deposits it.  At start-up: the interpolation only, with the
// given time step.  The gas vtheta is read in the fixed frame, vt + Rmed
// OmegaFrame per cell (src/Dsys.c:316), with the frame rate before this
// sub-step's frame update.  A particle outside the grid keeps its state
// and has no cell (p_cell = -1).
template <typename T, int S>
__global__ void __launch_bounds__(256) k_dust_move(KArgs a, int startup, double timestep0) {
  dfargo_ctrl *c = a.ctrl;
  if (!startup && !c->active) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int nr = a.nr, ns = a.ns;
  const double timestep = startup ? timestep0 : c->dt;
  const bool store01 = a.d_feedback, store2 = startup || !c->go_next || (a.d_feedback && a.energy);
  const DustSet P = a.pset[c->pset];
  unsigned long long ts_bits = ~0ull;
  bool outside = false;
  if (k < a.d_np) {
    double r = P.r[k], th = P.th[k];
    const double vrd = P.vr[k], vtd = P.vth[k];
    if (!startup && d_in_grid(a, r)) {
      const double rnew = r + vrd * 0.5 * (2. * timestep);
      th += vtd * r * 0.25 * (2. * timestep) * (1.0 / (r * r) + 1.0 / (rnew * rnew));
      th = d_wrap(a, th);
      th = d_wrap(a, th);
      r = rnew;
      P.r[k] = r;
    }
    // RotateDsys (src/Dsys.c:1153-1169) by the frame's new rate (in the
    // fixed frame a wrap only); the stencils below use the azimuth before it
    if (!startup) P.th[k] = d_wrap(a, th - c->omega_frame * timestep);
    const double frame = startup ? c->omega_frame : c->omega_prev;
    int ip, jp;
    if (!d_in_grid(a, r) || !d_cell(a, r, th, ip, jp)) {
      a.p_cell[k] = -1;
      outside = true;
    } else {
      const int cell = ip * ns + jp;
      a.p_cell[k] = cell;
      a.p_slot[k] = atomicAdd(&a.cell_count[cell], 1);
      const double rp = r, tp = th;
      const double az_jp = a.Azimuth[jp], rmed_ip = a.Rmed[ip];
      double wr[3], wt[3];
      int col[3], imin, jmin, base = 0;
      // vr stencil: gas vr, pressure gradient, indirect and SG accelerations
      double vrg = 0.0, radgradp = 0.0, radind = 0.0, radsg = 0.0;
      stencil_weights<S, 0>(a, ip, jp, rp, tp, az_jp, rmed_ip, wr, wt, col, imin, jmin, store01, k);
      base |= ((imin - ip + 1) << 0) | ((jmin - jp + 1) << 2);
      stencil_gather<T, S, 0>(a, wr, wt, col, imin, a.vr, a.radgradp, a.radindacc, a.radsgacc, false,
                              a.indirect_term, a.sg, vrg, radgradp, radind, radsg);
      // vtheta stencil
      double vtg = 0.0, azigradp = 0.0, aziind = 0.0, azisg = 0.0;
      stencil_weights<S, 1>(a, ip, jp, rp, tp, az_jp, rmed_ip, wr, wt, col, imin, jmin, store01, k);
      base |= ((imin - ip + 1) << 4) | ((jmin - jp + 1) << 6);
      stencil_gather<T, S, 1>(a, wr, wt, col, imin, a.vt, a.azigradp, a.aziindacc, a.azisgacc, false,
                              a.indirect_term, a.sg, vtg, azigradp, aziind, azisg, frame);
      // cell-centred stencil: density and the sound speed the sub-step
      // started with
      double densg = 0.0, csg = 0.0, unused0 = 0.0, unused1 = 0.0;
      stencil_weights<S, 2>(a, ip, jp, rp, tp, az_jp, rmed_ip, wr, wt, col, imin, jmin, store2, k);
      base |= ((imin - ip + 1) << 8) | ((jmin - jp + 1) << 10);
      stencil_gather<T, S, 2>(a, wr, wt, col, imin, a.rho,
                              a.energy ? ((startup ? a.cs : a.cs_old)) : (const void *)a.cs_iso, a.rho, a.rho,
                              !a.energy, false, false, densg, csg, unused0, unused1);
      a.p_base[k] = base;
      // Stokes number (Paardekooper 2007, src/Dsys.c:733-790)
      double vrp = vrd, vtp = vtd;
      const double densg_cgs = 0.1 * densg * a.d_unit_mass * a.d_len_m2;
      const double lambda = 3.34e-8 * (1.0 / densg_cgs) * d_aspect(a, rp) * rp;
      const double Kn = 0.5 * lambda / P.size[k];
      const double Mach = sqrt((vrp - vrg) * (vrp - vrg) + (vtp - vtg) * (vtp - vtg)) / csg;
      const double f_D = sqrt(1.0 + 9.0 * DF_PI * Mach * Mach / 128.0);
      const double Re = 3.0 * sqrt(DF_PI / 8.0) * Mach / Kn;
      double k_D = 0.0;
      if (Re <= 500.0)
        k_D = 1.0 + 0.15 * pow(Re, 0.687);
      else {
        if ((Re > 500.0) && (Re <= 1500.0)) k_D = 3.96e-6 * pow(Re, 2.4);
        if (Re > 1500.0) k_D = 0.11 * Re;
      }
      const double kn1 = 3.0 * Kn + 1.0;
      const double Cdrag = (kn1 * kn1) * (1.0 / (9.0 * Kn * Kn * f_D + 3.0 * Kn * k_D));
      const double St = 0.5 * DF_PI * Cdrag * P.size[k] * a.d_rho_code / densg;
      const double r15 = rp * sqrt(rp);
      const double Ts = St * r15;
      ts_bits = (unsigned long long)__double_as_longlong(Ts);
      P.stokes[k] = St;
      // drag acceleration (src/Dsys.c:800-830)
      if ((Ts < timestep) && a.d_sfta) {
        vrp = vrg + Ts * radgradp / densg;
        vtp = vtg + Ts * azigradp / densg;
      }
      a.p_raddrag[k] = -(vrp - vrg) / Ts;
      a.p_azidrag[k] = -(vtp - vtg) / Ts;
      a.p_ts[k] = Ts;
      a.p_dv2[k] = (vrp - vrg) * (vrp - vrg) + (vtp - vtg) * (vtp - vtg);
      if (!startup) {
        // UpdateDustVelocities (src/DustUpdate.c:290-424) from the velocity
        // the particle had before the drag estimate; its stopping time is
        // Ts (same St, same radius)
        const double rd = r, td = th, stoptime = Ts;
        int local_feel_disk = a.d_feel_disk;
        if ((rd < a.rmin) || (rd > a.rmax)) local_feel_disk = 0;
        if ((stoptime < timestep) && a.d_sfta) {
          const double vth_new = vtg + stoptime * azigradp / densg;
          P.vr[k] = vrg + stoptime * radgradp / densg;
          P.vth[k] = vth_new;
          P.l[k] = rd * vth_new;
        } else {
          double FgravR = -(1.0 / (rd * rd)), FgravTh = 0.;
          if ((fabs(FgravR + vtd * vtd / rd) < 1e-15) && (!a.d_feel_planets) && (!a.d_feel_disk) &&
              (!a.d_feel_sg))
            FgravR = -vtd * vtd / rd;
          if (a.d_feel_planets) {
            for (int q = 0; q < c->nb; q++) {
              const double xp = c->dust_pl_x[q], yp = c->dust_pl_y[q];
              const double rpl = sqrt(xp * xp + yp * yp), tpl = atan2(yp, xp), mp = c->dust_pl_m[q];
              const double eps = a.thickness_smoothing * a.aspect_ratio * pow(0.5 * rpl, 1.0 + a.flaring) *
                                 pow(2.0, 1.0 + a.flaring);
              const double dd = sqrt(rd * rd + rpl * rpl - 2.0 * rd * rpl * cos(td - tpl) + eps * eps);
              const double id3 = 1.0 / (dd * dd * dd);
              FgravR += mp * (rpl * cos(td - tpl) - rd) * id3;
              FgravTh += -mp * rpl * sin(td - tpl) * id3;
            }
          }
          if (a.indirect_term) {
            FgravR += radind;
            FgravTh += aziind;
          }
          if (a.sg && a.d_feel_sg) {
            FgravR += radsg;
            FgravTh += azisg;
          }
          const double den = 1.0 + local_feel_disk * timestep / stoptime;
          const double numTh = (vtd - vtg) + timestep * (FgravTh - vrd * vtd / rd);
          const double vth_new = vtg + numTh / den;
          P.vth[k] = vth_new;
          const double numR = (vrd - vrg) + timestep * (FgravR + (vth_new * vth_new) / rd);
          P.vr[k] = vrg + numR / den;
        }
      }
    }
  }
  // the particles outside the grid take slots of their own list, one
  // atomic per wavefront
  {
    const unsigned long long m = __ballot(outside);
    const int lane = threadIdx.x & 63;
    if (outside) {
      const int leader = __ffsll((long long)m) - 1;
      int base0 = 0;
      if (lane == leader) base0 = atomicAdd(&c->out_count, __popcll(m));
      base0 = __shfl(base0, leader, 64);
      a.p_slot[k] = base0 + __popcll(m & ((1ull << lane) - 1ull));
    }
  }
  // minimum stopping time of the block (positive doubles order like their
  // bit patterns), one entry per block: the control kernel of the next
  // sub-step takes the minimum over the blocks (no same-address atomics)
  __shared__ unsigned long long wmin[4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = (unsigned long long)__shfl_xor((long)ts_bits, off, 64);
    ts_bits = (o < ts_bits) ? o : ts_bits;
  }
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = ts_bits;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = wmin[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = (wmin[w] < m) ? wmin[w] : m;
    a.p_tsmin[blockIdx.x] = m;
  }
}

// bucket counts zeroed (start-up; in the sub-steps the CFL kernel does it,
// and the control kernel zeroes the count of particles outside the grid)
__global__ void __launch_bounds__(256) k_dust_zero(KArgs a) {
  const long n = (long)a.nr * a.ns;
  for (long l = blockIdx.x * (long)blockDim.x + threadIdx.x; l < n; l += (long)gridDim.x * blockDim.x)
    a.cell_count[l] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) a.ctrl->out_count = 0;
}

// exclusive scan of x over the block (1024 threads max): wavefront
// shuffles, then the wavefront totals; returns the exclusive prefix and
// the block total
__device__ __forceinline__ int block_exclusive_scan(int x, int *wsum, int &total) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, nw = (blockDim.x + 63) >> 6;
  int v = x;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(v, off, 64);
    if (lane >= off) v += y;
  }
  if (lane == 63 || t == (int)blockDim.x - 1) wsum[w] = v;
  __syncthreads();
  if (w == 0) {
    int s = (lane < nw) ? wsum[lane] : 0;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(s, off, 64);
      if (lane >= off) s += y;
    }
    if (lane < nw) wsum[lane] = s;
  }
  __syncthreads();
  total = wsum[nw - 1];
  const int excl = v - x + (w > 0 ? wsum[w - 1] : 0);
  __syncthreads();
  return excl;
}

// Bucket starts, one block per ring: the exclusive scan of the ring's
// counts (ring-local starts) and the ring total
__global__ void __launch_bounds__(1024) k_dust_ring_scan(KArgs a, int force) {
  __shared__ int wsum[16];
  const dfargo_ctrl *c = a.ctrl;
  if (!force && !c->active) return;
  const int i = blockIdx.x, ns = a.ns, t = threadIdx.x, nt = blockDim.x;
  const int per = (ns + nt - 1) / nt, j0 = t * per, j1 = min(j0 + per, ns);
  const long base = (long)i * ns;
  int s = 0;
  for (int j = j0; j < j1; j++) s += a.cell_count[base + j];
  int total;
  int excl = block_exclusive_scan(s, wsum, total);
  for (int j = j0; j < j1; j++) {
    a.cell_local[base + j] = excl;
    excl += a.cell_count[base + j];
  }
  if (t == 0) a.ring_tot[i] = total;
}

// ring offsets: exclusive scan of the ring totals, one block; ring_off[nr]
// is the number of particles inside the grid
__global__ void __launch_bounds__(1024) k_dust_ring_offsets(KArgs a, int force) {
  __shared__ int wsum[16];
  const dfargo_ctrl *c = a.ctrl;
  if (!force && !c->active) return;
  const int nr = a.nr, t = threadIdx.x, nt = blockDim.x;
  int carry = 0;
  for (int b = 0; b < nr; b += nt) {
    const int v = (b + t < nr) ? a.ring_tot[b + t] : 0;
    int tot;
    const int ex = block_exclusive_scan(v, wsum, tot);
    if (b + t < nr) a.ring_off[b + t] = carry + ex;
    carry += tot;
  }
  if (t == 0) a.ring_off[nr] = carry;
}

// particles into their cell's bucket, at the slot their count gave them
// (id and position), or into the list of those outside the grid
__global__ void __launch_bounds__(256) k_dust_scatter(KArgs a, int force) {
  const dfargo_ctrl *c = a.ctrl;
  if (!force && !c->active) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.d_np) return;
  const int cell = a.p_cell[k];
  if (cell < 0) {
    a.out_list[a.p_slot[k]] = k;
    return;
  }
  const unsigned long long id = (unsigned long long)(unsigned)a.pset[c->pset].id[k];
  a.p_sorted[a.ring_off[cell / a.ns] + a.cell_local[cell] + a.p_slot[k]] = (id << 32) | (unsigned)k;
}

// each bucket sorted by particle index (insertion sort: buckets are small),
// and its global start recorded
__global__ void __launch_bounds__(256) k_dust_sort(KArgs a, int force) {
  const dfargo_ctrl *c = a.ctrl;
  if (!force && !c->active) return;
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  if (j >= a.ns) return;
  const long l = (long)i * a.ns + j;
  const int start = a.ring_off[i] + a.cell_local[l];
  a.cell_start[l] = start;
  unsigned long long *b = a.p_sorted + start;
  const int n = a.cell_count[l];
  for (int x = 1; x < n; x++) {
    const unsigned long long v = b[x];
    int y = x - 1;
    while (y >= 0 && b[y] > v) {
      b[y + 1] = b[y];
      y--;
    }
    b[y + 1] = v;
  }
}

// Feedback deposit, one thread per grid cell (i, j): the particles of
// rings i-1 .. i+1 x columns j-2 .. j+1 (every stencil that can reach this
// cell), bucket by bucket in that fixed order and by particle id within a
// bucket, each with the weights its interpolation stored.  The order is
// deterministic (a replay or another run gives the same bits); against the
// reference's particle-index order it changes only the rounding of the
// sums.  The dust density is a diagnostic: computed at start-up and on the
// last sub-step of a call.
#define DEP_BX 256
template <typename T, int S> __global__ void __launch_bounds__(DEP_BX) k_dust_deposit(KArgs a, int force) {
  __shared__ int sStart[3][DEP_BX + 3], sCount[3][DEP_BX + 3];
  const dfargo_ctrl *c = a.ctrl;
  if (!force && !c->active) return;
  if (!dust_deposit_needed(a, c, force)) return;
  const int nr = a.nr, ns = a.ns, bx = blockDim.x, t = threadIdx.x;
  const int j0 = blockIdx.x * bx, i = blockIdx.y, j = j0 + t;
  // bucket starts and counts of the neighbourhood, columns j0-2 .. j0+bx
  for (int idx = t; idx < 3 * (bx + 3); idx += bx) {
    const int q = idx / (bx + 3), p = idx - q * (bx + 3), si = i - 1 + q;
    int cj = j0 - 2 + p;
    cj = (cj < 0) ? cj + ns : ((cj >= ns) ? cj - ns : cj);
    if (si < 0 || si > nr - 1) {
      sStart[q][p] = 0;
      sCount[q][p] = 0;
    } else {
      const long sl = (long)si * ns + cj;
      sStart[q][p] = a.cell_start[sl];
      sCount[q][p] = a.cell_count[sl];
    }
  }
  __syncthreads();
  if (j >= ns) return;
  const bool want_dens = force || !c->go_next;
  const bool want_fb = a.d_feedback, want_e = a.d_feedback && a.energy;
  const long l = (long)i * ns + j, np = a.d_np;
  const double mp = force ? c->particles_mass : c->dust_mp, surf = a.Surf[i];
  const double mg = ldd<T>(a.rho, l) * surf;
  double fbr = 0.0, fbt = 0.0, fbe = 0.0, dd = 0.0;
  for (int b = 0; b < 12; b++) {
    const int e0 = sStart[b >> 2][t + (b & 3)], e1 = e0 + sCount[b >> 2][t + (b & 3)];
    for (int e = e0; e < e1; e++) {
      const int k = (int)(unsigned)(a.p_sorted[e] & 0xffffffffull);
      const int cell = a.p_cell[k], code = a.p_base[k], ip = cell / ns, jp = cell - ip * ns;
#pragma unroll
      for (int kind = 0; kind < 3; kind++) {
        if (kind < 2 && !want_fb) continue;
        if (kind == 2 && !want_dens && !want_e) continue;
        const int imin = ip - 1 + ((code >> (4 * kind)) & 3), jmin = jp - 1 + ((code >> (4 * kind + 2)) & 3);
        const int p = i - imin;
        int q = j - jmin;
        q = (q < 0) ? q + ns : ((q >= ns) ? q - ns : q);
        if (p < 0 || p >= nodes_r(S, kind) || q >= nodes_t(S, kind)) continue;
        const double wr = a.p_w[(kind * 6 + p) * np + k], wt = a.p_w[(kind * 6 + 3 + q) * np + k];
        if (kind == 0) {
          fbr -= (wr * wt * a.p_raddrag[k]) * mp / mg;
        } else if (kind == 1) {
          fbt -= (wr * wt * a.p_azidrag[k]) * mp / mg;
        } else {
          if (want_dens) dd += wr * wt * mp / surf;
          if (want_e) {
            const double rhop = mp / surf;
            fbe += rhop * wr * wt * a.p_dv2[k] / a.p_ts[k];
          }
        }
      }
    }
  }
  if (want_fb) {
    ((T *)a.radfbacc)[l] = (T)fbr;
    ((T *)a.azifbacc)[l] = (T)fbt;
  }
  if (want_e) ((T *)a.fbedot)[l] = (T)fbe;
  if (want_dens) ((T *)a.dustdens)[l] = (T)dd;
}

// The particles re-sorted by cell (bucket order, the particles outside the
// grid after them) into the other set, so that the next sub-step's
// gathers and deposit read neighbouring particles from neighbouring memory.
__global__ void __launch_bounds__(256) k_dust_reorder(KArgs a) {
  const dfargo_ctrl *c = a.ctrl;
  if (!c->active) return;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.d_np) return;
  const int nin = a.ring_off[a.nr];
  const int src = (s < nin) ? (int)(unsigned)(a.p_sorted[s] & 0xffffffffull) : a.out_list[s - nin];
  const DustSet A = a.pset[c->pset], B = a.pset[c->pset ^ 1];
  B.r[s] = A.r[src];
  B.th[s] = A.th[src];
  B.vr[s] = A.vr[src];
  B.vth[s] = A.vth[src];
  B.l[s] = A.l[src];
  B.size[s] = A.size[src];
  B.stokes[s] = A.stokes[src];
  B.id[s] = A.id[src];
}

// The gas vtheta after the start-up interpolation in a rotating frame:
// vt + Rmed OmegaFrame and back (src/Dsys.c:316,1082), one cell per thread
template <typename T> __global__ void __launch_bounds__(256) k_dust_frame_trip(KArgs a) {
  const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.ns) return;
  const long l = (long)i * a.ns + j;
  const double rw = a.Rmed[i] * a.ctrl->omega_frame;
  ((T *)a.vt)[l] = (T)((ldd<T>(a.vt, l) + rw) - rw);
}

}  // namespace

template <typename T, int S> static void dust_launch_s(int id, const KArgs &a, hipStream_t s, int force,
                                                       double timestep) {
  const int np = a.d_np, pb = (np + 255) / 256;
  const int bx = a.ns < DEP_BX ? a.ns : DEP_BX;
  const dim3 cg((a.ns + bx - 1) / bx, a.nr);
  switch (id) {
    case KID_DUST_MOVE:
      if (np > 0) hipLaunchKernelGGL((k_dust_move<T, S>), dim3(pb), dim3(256), 0, s, a, force, timestep);
      break;
    case KID_DUST_SCAN:
      // the bucket counts are zeroed by the CFL kernel of the sub-step (and
      // by the start-up before its move); the ring scan turns them into starts
      hipLaunchKernelGGL(k_dust_ring_scan, dim3(a.nr), dim3(a.ns < 1024 ? ((a.ns + 63) / 64) * 64 : 1024), 0, s,
                         a, force);
      hipLaunchKernelGGL(k_dust_ring_offsets, dim3(1), dim3(1024), 0, s, a, force);
      break;
    case KID_DUST_SCATTER:
      if (np > 0) hipLaunchKernelGGL(k_dust_scatter, dim3(pb), dim3(256), 0, s, a, force);
      break;
    case KID_DUST_SORT: hipLaunchKernelGGL(k_dust_sort, cg, dim3(bx), 0, s, a, force); break;
    case KID_DUST_DEPOSIT: hipLaunchKernelGGL((k_dust_deposit<T, S>), cg, dim3(bx), 0, s, a, force); break;
    case KID_DUST_REORDER:
      if (np > 0) hipLaunchKernelGGL(k_dust_reorder, dim3(pb), dim3(256), 0, s, a);
      break;
  }
}

template <typename T> static void dust_launch(int id, const KArgs &a, hipStream_t s, int force, double timestep) {
  switch (a.d_interp) {
    case DFARGO_INTERP_NGP: dust_launch_s<T, DFARGO_INTERP_NGP>(id, a, s, force, timestep); break;
    case DFARGO_INTERP_CIC: dust_launch_s<T, DFARGO_INTERP_CIC>(id, a, s, force, timestep); break;
    default: dust_launch_s<T, DFARGO_INTERP_TSC>(id, a, s, force, timestep); break;
  }
}

template <typename T> void launch_dust_kernel(int id, const KArgs &a, hipStream_t s) { dust_launch<T>(id, a, s, 0, 0.0); }

template <typename T> void launch_dust_startup(const KArgs &a, double timestep, double omega_frame, hipStream_t s) {
  hipLaunchKernelGGL(k_dust_zero, dim3(1024), dim3(256), 0, s, a);
  for (int id : {KID_DUST_MOVE, KID_DUST_SCAN, KID_DUST_SCATTER, KID_DUST_SORT, KID_DUST_DEPOSIT})
    dust_launch<T>(id, a, s, 1, timestep);
  if (omega_frame != 0.0 && a.d_np > 0) {
    const int bx = a.ns < 256 ? a.ns : 256;
    hipLaunchKernelGGL(k_dust_frame_trip<T>, dim3((a.ns + bx - 1) / bx, a.nr), dim3(bx), 0, s, a);
  }
}

template void launch_dust_kernel<double>(int, const KArgs &, hipStream_t);
template void launch_dust_kernel<float>(int, const KArgs &, hipStream_t);
template void launch_dust_startup<double>(const KArgs &, double, double, hipStream_t);
template void launch_dust_startup<float>(const KArgs &, double, double, hipStream_t);
