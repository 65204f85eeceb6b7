// dust.hip -- Lagrangian dust particles of one sub-step on CDNA4
// (SemiUpdateDustPositions, src/DustUpdate.c:6-138; UpdateDustVelocities,
// src/DustUpdate.c:267-435; interpolation, src/Dsys.c:265-1149; the
// start-up interpolation of src/main.c:287).
//
// One thread per particle moves it, finds its cell, gathers the gas
// quantities of its three stencils, computes its Stokes number and drag
// and updates its velocity.  The feedback deposit (the scatter half of
// interpolation) is turned into a gather so that no float atomics decide
// the summation order: particles are bucketed by cell (counts, exclusive
// scan, scatter, each bucket then sorted by particle index), and each grid
// cell walks the particles of the cells whose stencils can reach it, in
// increasing particle index, recomputing each one's weight for this cell.
// Every cell therefore sums its contributions in the order of the
// reference's sequential particle loop.
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

// weights for a node at distances (dr, dphi); kind is the stencil: 0 the
// vr one (radial NGP cut), 1 the vtheta one (azimuthal NGP cut), 2 the
// cell-centred one (src/Dsys.c:407-441,490-524,640-665)
__device__ void d_weights(int scheme, int kind, double dr, double dphi, double delta_r, double delta_phi,
                          double &wr, double &wt) {
  if (scheme == DFARGO_INTERP_NGP) {
    wr = 1.0;
    wt = 1.0;
    if (kind == 0) wr = (dr < 0.5 * delta_r) ? 1.0 : 0.0;
    if (kind == 1) wt = (dphi < 0.5 * delta_phi) ? 1.0 : 0.0;
  } else if (scheme == DFARGO_INTERP_CIC) {
    wr = 1.0 - dr / delta_r;
    wt = 1.0 - dphi / delta_phi;
  } else {
    if (dr < 0.5 * delta_r) wr = 0.75 - dr * dr / delta_r / delta_r;
    if ((dr >= 0.5 * delta_r) && (dr <= 1.5 * delta_r)) wr = 0.5 * (1.5 - dr / delta_r) * (1.5 - dr / delta_r);
    if (dr > 1.5 * delta_r) wr = 0.0;
    if (dphi < 0.5 * delta_phi) wt = 0.75 - dphi * dphi / delta_phi / delta_phi;
    if ((dphi >= 0.5 * delta_phi) && (dphi <= 1.5 * delta_phi))
      wt = 0.5 * (1.5 - dphi / delta_phi) * (1.5 - dphi / delta_phi);
    if (dphi > 1.5 * delta_phi) wt = 0.0;
  }
}

// node ranges of the three stencils for a particle in cell (ip, jp).  The
// rings stay within ip-1 .. ip+1 and the columns within jp-1 .. jp+2 (the
// TSC vtheta stencil shifts one column up when the particle is past the
// cell centre): the deposit gathers from exactly that neighbourhood.
__device__ void d_stencil(const KArgs &a, int kind, int ip, int jp, double rp, double tp, int &imin, int &imax,
                          int &jmin, int &jmax) {
  int myip = ip, myjp = jp;
  if (a.d_interp == DFARGO_INTERP_NGP) {
    imin = ip;
    imax = ip + (kind == 0 ? 1 : 0);
    jmin = jp;
    jmax = jp + (kind == 1 ? 1 : 0);
  } else if (a.d_interp == DFARGO_INTERP_CIC) {
    if (kind != 1) myjp = (tp < a.Azimuth[jp]) ? jp - 1 : jp;
    if (kind != 0) myip = (rp < a.Rmed[ip]) ? ip - 1 : ip;
    imin = myip;
    imax = myip + 1;
    jmin = myjp;
    jmax = myjp + 1;
  } else {
    if (kind == 1) myjp = (tp > a.Azimuth[jp]) ? jp + 1 : jp;
    imin = ip - 1;
    imax = ip + 1;
    jmin = myjp - 1;
    jmax = myjp + 1;
  }
}

// reference azimuth of column j of a stencil: AziInf for the vtheta one,
// the cell centre otherwise; columns past either end continue the grid
// (src/Dsys.c:411-425).  Returns the wrapped column.
__device__ int d_col(const KArgs &a, int kind, int j, double &az) {
  const int ns = a.ns;
  const double half = 0.5 * (a.pm) / (double)ns;
  if (j >= 0 && j < ns) {
    az = (kind == 1) ? a.Azimuth[j] - half : a.Azimuth[j];
    return j;
  }
  if (j < 0) {
    az = ((kind == 1) ? a.Azimuth[0] - half : a.Azimuth[0]) + j * a.delta_phi;
    return j + ns;
  }
  az = ((kind == 1) ? a.Azimuth[ns - 1] - half : a.Azimuth[ns - 1]) + (j + 1 - ns) * a.delta_phi;
  return j - ns;
}

// radial distance of a stencil node: to Rinf for the vr stencil, to Rmed
// otherwise, logarithmic on a logarithmic grid
__device__ double d_dr(const KArgs &a, int kind, int i, double rp) {
  if (i < 0 || i > a.nr - 1) return 0.0;
  const double ref = (kind == 0) ? a.Rinf[i] : a.Rmed[i];
  return a.log_grid ? fabs(log(rp / ref)) : fabs(rp - ref);
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
// Rinf[ip] <= r (the linear search of src/Dsys.c:363-365, as a bisection),
// jp from the azimuth.  False when jp falls outside the grid, which only an
// azimuth outside [AziInf[0], AziSup[ns-1]] can produce.
__device__ bool d_cell(const KArgs &a, double rp, double tp, int &ip, int &jp) {
  int lo = 0, hi = a.nr;  // Rinf[lo] <= rp < Rinf[hi], Rinf[nr] = Rsup[nr-1]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (rp >= a.Rinf[mid]) lo = mid;
    else hi = mid;
  }
  ip = lo;
  jp = (int)floor(a.ns * (tp - a.aziinf0) / (a.azisup_last - a.aziinf0));
  if (jp == a.ns) jp = 0;
  return jp >= 0 && jp < a.ns;
}

// One particle per thread.  In a sub-step: SemiUpdateDustPositions over
// 2 dt, interpolation (gathers, Stokes number, drag), UpdateDustVelocities
// and RotateDsys by OmegaFrame dt = 0.  At start-up: the interpolation
// only, with the given time step.  A particle outside the grid keeps its
// state and has no cell (p_cell = -1).
template <typename T> __global__ void __launch_bounds__(256) k_dust_move(KArgs a, int startup, double timestep0) {
  dfargo_ctrl *c = a.ctrl;
  if (!startup && !c->active) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.d_np) return;
  const int nr = a.nr, ns = a.ns, scheme = a.d_interp;
  const double timestep = startup ? timestep0 : c->dt;
  double r = a.p_r[k], th = a.p_th[k];
  if (!startup && d_in_grid(a, r)) {
    const double rnew = r + a.p_vr[k] * 0.5 * (2. * timestep);
    th += a.p_vth[k] * r * 0.25 * (2. * timestep) * (pow(r, -2.) + pow(rnew, -2.));
    th = d_wrap(a, th);
    th = d_wrap(a, th);
    r = rnew;
    a.p_r[k] = r;
  }
  // RotateDsys by a zero angle still wraps every particle
  if (!startup) a.p_th[k] = d_wrap(a, th);
  int ip, jp;
  if (!d_in_grid(a, r) || !d_cell(a, r, th, ip, jp)) {
    a.p_cell[k] = -1;
    return;
  }
  a.p_cell[k] = ip * ns + jp;
  const double rp = r, tp = th;
  const void *csp = (startup || !a.energy) ? a.cs : a.cs_old;
  double vrg = 0.0, vtg = 0.0, radgradp = 0.0, azigradp = 0.0, radind = 0.0, aziind = 0.0;
  double radsg = 0.0, azisg = 0.0, densg = 0.0, csg = 0.0, wr = 0.0, wt = 0.0, az;
  int imin, imax, jmin, jmax;
  // vr stencil
  d_stencil(a, 0, ip, jp, rp, tp, imin, imax, jmin, jmax);
  for (int j = jmin; j <= jmax; j++) {
    const int myj = d_col(a, 0, j, az);
    const double dphi = fabs(tp - az);
    for (int i = imin; i <= imax; i++) {
      d_weights(scheme, 0, d_dr(a, 0, i, rp), dphi, a.delta_r, a.delta_phi, wr, wt);
      if (i >= 0 && i <= nr - 1) {
        const long l = (long)i * ns + myj;
        vrg += (wr * wt * ldd<T>(a.vr, l));
        radgradp += (wr * wt * ldd<T>(a.radgradp, l));
        if (a.indirect_term) radind += (wr * wt * ldd<T>(a.radindacc, l));
        if (a.sg) radsg += (wr * wt * ldd<T>(a.radsgacc, l));
      }
    }
  }
  // vtheta stencil
  d_stencil(a, 1, ip, jp, rp, tp, imin, imax, jmin, jmax);
  for (int j = jmin; j <= jmax; j++) {
    const int myj = d_col(a, 1, j, az);
    const double dphi = fabs(tp - az);
    for (int i = imin; i <= imax; i++) {
      d_weights(scheme, 1, d_dr(a, 1, i, rp), dphi, a.delta_r, a.delta_phi, wr, wt);
      if (i >= 0 && i <= nr - 1) {
        const long l = (long)i * ns + myj;
        vtg += (wr * wt * ldd<T>(a.vt, l));
        azigradp += (wr * wt * ldd<T>(a.azigradp, l));
        if (a.indirect_term) aziind += (wr * wt * ldd<T>(a.aziindacc, l));
        if (a.sg) azisg += (wr * wt * ldd<T>(a.azisgacc, l));
      }
    }
  }
  // cell-centred stencil: density and sound speed
  d_stencil(a, 2, ip, jp, rp, tp, imin, imax, jmin, jmax);
  for (int j = jmin; j <= jmax; j++) {
    const int myj = d_col(a, 2, j, az);
    const double dphi = fabs(tp - az);
    for (int i = imin; i <= imax; i++) {
      d_weights(scheme, 2, d_dr(a, 2, i, rp), dphi, a.delta_r, a.delta_phi, wr, wt);
      if (i >= 0 && i <= nr - 1) {
        const long l = (long)i * ns + myj;
        densg += (wr * wt * ldd<T>(a.rho, l));
        csg += (wr * wt * (a.energy ? ldd<T>(csp, l) : (double)(T)a.cs_iso[i]));
      }
    }
  }
  // Stokes number (Paardekooper 2007, src/Dsys.c:733-790)
  double vrp = a.p_vr[k], vtp = a.p_vth[k];
  const double densg_cgs = 0.1 * densg * a.d_unit_mass * pow(a.d_unit_length, -2.0);
  const double lambda = 3.34e-8 * pow(densg_cgs, -1.) * d_aspect(a, rp) * rp;
  const double Kn = 0.5 * lambda / a.p_size[k];
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
  const double Cdrag = pow(3.0 * Kn + 1.0, 2.0) * pow(9.0 * Kn * Kn * f_D + 3.0 * Kn * k_D, -1.0);
  const double St = 0.5 * DF_PI * Cdrag * a.p_size[k] * a.d_rho_code / densg;
  const double Ts = St * pow(rp, 1.5);
  // minimum stopping time: positive doubles order like their bit patterns
  atomicMin(&c->min_ts_bits, (unsigned long long)__double_as_longlong(Ts));
  a.p_stokes[k] = St;
  // drag acceleration (src/Dsys.c:800-830)
  if ((Ts < timestep) && a.d_sfta) {
    vrp = vrg + Ts * radgradp / densg;
    vtp = vtg + Ts * azigradp / densg;
  }
  a.p_raddrag[k] = -(vrp - vrg) / Ts;
  a.p_azidrag[k] = -(vtp - vtg) / Ts;
  a.p_ts[k] = Ts;
  a.p_dv2[k] = (vrp - vrg) * (vrp - vrg) + (vtp - vtg) * (vtp - vtg);
  if (startup) return;
  // UpdateDustVelocities (src/DustUpdate.c:290-424) from the velocity the
  // particle had before the drag estimate
  const double rd = r, td = th, vrd = a.p_vr[k], vtd = a.p_vth[k];
  int local_feel_disk = a.d_feel_disk;
  if ((rd < a.rmin) || (rd > a.rmax)) local_feel_disk = 0;
  const double stoptime = St * pow(rd, 1.5);
  if ((stoptime < timestep) && a.d_sfta) {
    const double vth_new = vtg + stoptime * azigradp / densg;
    a.p_vr[k] = vrg + stoptime * radgradp / densg;
    a.p_vth[k] = vth_new;
    a.p_l[k] = rd * vth_new;
  } else {
    double FgravR = -pow(rd, -2.), FgravTh = 0.;
    if ((fabs(FgravR + vtd * vtd / rd) < 1e-15) && (!a.d_feel_planets) && (!a.d_feel_disk) && (!a.d_feel_sg))
      FgravR = -vtd * vtd / rd;
    if (a.d_feel_planets) {
      for (int q = 0; q < c->nb; q++) {
        const double xp = c->pl_x[q], yp = c->pl_y[q];
        const double rpl = sqrt(xp * xp + yp * yp), tpl = atan2(yp, xp), mp = c->dust_pl_m[q];
        const double eps = a.thickness_smoothing * a.aspect_ratio * pow(0.5 * rpl, 1.0 + a.flaring) *
                           pow(2.0, 1.0 + a.flaring);
        const double dd = sqrt(rd * rd + rpl * rpl - 2.0 * rd * rpl * cos(td - tpl) + eps * eps);
        FgravR += mp * (rpl * cos(td - tpl) - rd) * pow(dd, -3.0);
        FgravTh += -mp * rpl * sin(td - tpl) * pow(dd, -3.0);
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
    a.p_vth[k] = vth_new;
    const double numR = (vrd - vrg) + timestep * (FgravR + (vth_new * vth_new) / rd);
    a.p_vr[k] = vrg + numR / den;
  }
}

// bucket counts of the particles per cell (and a zeroed fill cursor)
__global__ void __launch_bounds__(256) k_dust_zero(KArgs a, int force) {
  if (!force && !a.ctrl->active) return;
  const long n = (long)a.nr * a.ns;
  for (long l = blockIdx.x * (long)blockDim.x + threadIdx.x; l < n; l += (long)gridDim.x * blockDim.x) {
    a.cell_count[l] = 0;
    a.cell_fill[l] = 0;
  }
}
__global__ void __launch_bounds__(256) k_dust_count(KArgs a, int force) {
  if (!force && !a.ctrl->active) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.d_np) return;
  const int cell = a.p_cell[k];
  if (cell >= 0) atomicAdd(&a.cell_count[cell], 1);
}

// exclusive scan of the counts: per-block scans, one block scans the block
// totals, then the offsets are added
#define SCAN_B 1024
__device__ int block_inclusive_scan(int v, int *sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < SCAN_B; off <<= 1) {
    const int add = (t >= off) ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  return sh[t];
}
__global__ void __launch_bounds__(SCAN_B) k_scan_blocks(KArgs a, int force) {
  __shared__ int sh[SCAN_B];
  if (!force && !a.ctrl->active) return;
  const long n = (long)a.nr * a.ns, g = blockIdx.x * (long)SCAN_B + threadIdx.x;
  const int v = (g < n) ? a.cell_count[g] : 0;
  const int inc = block_inclusive_scan(v, sh);
  if (g < n) a.cell_start[g] = inc - v;
  if (threadIdx.x == SCAN_B - 1) a.scan_sums[blockIdx.x] = inc;
}
__global__ void __launch_bounds__(SCAN_B) k_scan_sums(KArgs a, int force, int nblocks) {
  __shared__ int sh[SCAN_B];
  if (!force && !a.ctrl->active) return;
  int carry = 0;
  for (int base = 0; base < nblocks; base += SCAN_B) {
    const int g = base + threadIdx.x;
    const int v = (g < nblocks) ? a.scan_sums[g] : 0;
    const int inc = block_inclusive_scan(v, sh);
    if (g < nblocks) a.scan_sums[g] = carry + inc - v;
    const int total = sh[SCAN_B - 1];
    __syncthreads();
    carry += total;
  }
}
__global__ void __launch_bounds__(SCAN_B) k_scan_add(KArgs a, int force) {
  if (!force && !a.ctrl->active) return;
  const long n = (long)a.nr * a.ns, g = blockIdx.x * (long)SCAN_B + threadIdx.x;
  if (g < n) a.cell_start[g] += a.scan_sums[blockIdx.x];
}

// particles into their cell's bucket, in arbitrary order
__global__ void __launch_bounds__(256) k_dust_scatter(KArgs a, int force) {
  if (!force && !a.ctrl->active) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.d_np) return;
  const int cell = a.p_cell[k];
  if (cell < 0) return;
  a.p_sorted[a.cell_start[cell] + atomicAdd(&a.cell_fill[cell], 1)] = k;
}

// each bucket sorted by particle index (insertion sort: buckets are small)
__global__ void __launch_bounds__(256) k_dust_sort(KArgs a, int force) {
  if (!force && !a.ctrl->active) return;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.ns) return;
  const long l = (long)blockIdx.y * a.ns + j;
  int *b = a.p_sorted + a.cell_start[l];
  const int n = a.cell_count[l];
  for (int x = 1; x < n; x++) {
    const int v = b[x];
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
// cell), merged in increasing particle index, each weighted for this cell
// with its own stencils.  The dust density is a diagnostic: computed at
// start-up and on the last sub-step of a call.
#define DEP_NB 12
template <typename T> __global__ void __launch_bounds__(256) k_dust_deposit(KArgs a, int force) {
  const dfargo_ctrl *c = a.ctrl;
  if (!force && !c->active) return;
  const int nr = a.nr, ns = a.ns, scheme = a.d_interp;
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  if (j >= ns) return;
  const bool want_dens = force || !c->go;
  const bool want_fb = a.d_feedback, want_e = a.d_feedback && a.energy;
  const long l = (long)i * ns + j;
  int cur[DEP_NB], end[DEP_NB];
#pragma unroll
  for (int q = 0; q < DEP_NB; q++) {
    const int si = i - 1 + q / 4, sj0 = j - 2 + q % 4;
    const int sj = (sj0 < 0) ? sj0 + ns : ((sj0 >= ns) ? sj0 - ns : sj0);
    if (si < 0 || si > nr - 1) {
      cur[q] = end[q] = 0;
    } else {
      const long sl = (long)si * ns + sj;
      cur[q] = a.cell_start[sl];
      end[q] = cur[q] + a.cell_count[sl];
    }
  }
  const double mp = force ? c->particles_mass : c->dust_mp, surf = a.Surf[i];
  const double mg = ldd<T>(a.rho, l) * surf;
  double fbr = 0.0, fbt = 0.0, fbe = 0.0, dd = 0.0;
  for (;;) {
    int best = -1, bk = 0x7fffffff;
#pragma unroll
    for (int q = 0; q < DEP_NB; q++)
      if (cur[q] < end[q]) {
        const int kk = a.p_sorted[cur[q]];
        if (kk < bk) {
          bk = kk;
          best = q;
        }
      }
    if (best < 0) break;
#pragma unroll
    for (int q = 0; q < DEP_NB; q++)
      if (q == best) cur[q]++;
    const int k = bk;
    const double rp = a.p_r[k], tp = a.p_th[k];
    const int cell = a.p_cell[k], ip = cell / ns, jp = cell - ip * ns;
    double wr = 0.0, wt = 0.0, az;
    int imin, imax, jmin, jmax;
    for (int kind = 0; kind < 3; kind++) {
      if (kind < 2 && !want_fb) continue;
      if (kind == 2 && !want_dens && !want_e) continue;
      d_stencil(a, kind, ip, jp, rp, tp, imin, imax, jmin, jmax);
      if (i < imin || i > imax) continue;
      for (int jj = jmin; jj <= jmax; jj++) {
        const int myj = d_col(a, kind, jj, az);
        if (myj != j) continue;
        const double dphi = fabs(tp - az);
        d_weights(scheme, kind, d_dr(a, kind, i, rp), dphi, a.delta_r, a.delta_phi, wr, wt);
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

}  // namespace

template <typename T> static void dust_launch(int id, const KArgs &a, hipStream_t s, int force, double timestep) {
  const int np = a.d_np, pb = (np + 255) / 256;
  const long ncell = (long)a.nr * a.ns;
  const int nsb = dust_scan_blocks(ncell);
  const int bx = a.ns < 256 ? a.ns : 256;
  const dim3 cg((a.ns + bx - 1) / bx, a.nr);
  switch (id) {
    case KID_DUST_MOVE:
      if (np > 0) hipLaunchKernelGGL(k_dust_move<T>, dim3(pb), dim3(256), 0, s, a, force, timestep);
      break;
    case KID_DUST_COUNT:
      hipLaunchKernelGGL(k_dust_zero, dim3(1024), dim3(256), 0, s, a, force);
      if (np > 0) hipLaunchKernelGGL(k_dust_count, dim3(pb), dim3(256), 0, s, a, force);
      break;
    case KID_DUST_SCAN:
      hipLaunchKernelGGL(k_scan_blocks, dim3(nsb), dim3(SCAN_B), 0, s, a, force);
      hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_B), 0, s, a, force, nsb);
      hipLaunchKernelGGL(k_scan_add, dim3(nsb), dim3(SCAN_B), 0, s, a, force);
      break;
    case KID_DUST_SCATTER:
      if (np > 0) hipLaunchKernelGGL(k_dust_scatter, dim3(pb), dim3(256), 0, s, a, force);
      hipLaunchKernelGGL(k_dust_sort, cg, dim3(bx), 0, s, a, force);
      break;
    case KID_DUST_DEPOSIT: hipLaunchKernelGGL(k_dust_deposit<T>, cg, dim3(bx), 0, s, a, force); break;
  }
}

int dust_scan_blocks(long ncell) { return (int)((ncell + SCAN_B - 1) / SCAN_B); }

template <typename T> void launch_dust_kernel(int id, const KArgs &a, hipStream_t s) { dust_launch<T>(id, a, s, 0, 0.0); }

template <typename T> void launch_dust_startup(const KArgs &a, double timestep, hipStream_t s) {
  for (int id : {KID_DUST_MOVE, KID_DUST_COUNT, KID_DUST_SCAN, KID_DUST_SCATTER, KID_DUST_DEPOSIT})
    dust_launch<T>(id, a, s, 1, timestep);
}

template void launch_dust_kernel<double>(int, const KArgs &, hipStream_t);
template void launch_dust_kernel<float>(int, const KArgs &, hipStream_t);
template void launch_dust_startup<double>(const KArgs &, double, hipStream_t);
template void launch_dust_startup<float>(const KArgs &, double, hipStream_t);
