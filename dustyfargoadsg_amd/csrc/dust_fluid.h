// dust_fluid.h -- the dust fluid (DUSTFLUID YES) of the sub-step, included
// by step.hip after the gas kernels whose helpers it shares (pot_at,
// src_press, d_aspect_ratio, wave_sum).  The fluid runs with FP64 storage;
// its fields are one cell per thread, in global memory: it is a passive
// second fluid beside the gas (the gas only feels it through the optional
// feedback drag that k_source applies), so nothing here is on the gas's
// critical path when the fluid is off.
//
// Order within a sub-step (src/SourceEuler.c:545-700):
//   k_df_axisg     ring means of SG_Accr for DUSTFEELSG Z
//   k_df_source    Stokes number, the dust's SubStep1 kicks and drag,
//                  self-gravity (before k_source, whose feedback drag reads
//                  the Stokes numbers and the dust state the sub-step starts
//                  with)
//   k_df_stress, k_df_visc   the dust's viscous stresses and update
//   k_df_substep2  artificial viscosity (DVradNew), or in the
//                  short-friction-time approximation DVradInt as it is;
//                  the per-tile sums of the new vtheta for the FARGO mean
//   k_transport_rad / k_transport_azi on the dust fields (Transportd)
//   k_df_finish    ComputeVelocities, Diffd, the floor, the wall,
//                  sub-Keplerian and open-dust boundaries
// and the wave-killing zones damp the fluid in k_evanescent.

// ---- the grain's Stokes number: SubStep1 (src/SourceEuler.c:820-858) ----
// Paardekooper (2007): the drag coefficient from the Knudsen number, the
// Mach number of the relative velocity and the Reynolds number.  lam is the
// ring's 3.34e-8 AspectRatio(Rmed) Rmed / df_to_cgs, so that the mean free
// path is lam / Sigma (pow(x, -1) of the reference taken as 1 / x).
__device__ __forceinline__ double df_stokes_at(const KArgs &a, double lam, double rho, double dvr, double dvt,
                                               double cs) {
  const double lambda = lam / rho;
  const double Kn = 0.5 * lambda / a.df_radius;
  const double Mach = sqrt(dvr * dvr + dvt * dvt) / cs;
  const double f_D = sqrt(1.0 + 9.0 * DF_PI * Mach * Mach / 128.0);
  const double Re = 3.0 * sqrt(DF_PI / 8.0) * Mach / Kn;
  double k_D = 0.0;
  if (Re <= 500.0)
    k_D = 1.0 + 0.15 * pow(Re, 0.687);
  else {
    if ((Re > 500.0) && (Re <= 1500.0)) k_D = 3.96e-6 * pow(Re, 2.4);
    if (Re > 1500.0) k_D = 0.11 * Re;
  }
  const double Cdrag = (3.0 * Kn + 1.0) * (3.0 * Kn + 1.0) / (9.0 * Kn * Kn * f_D + 3.0 * Kn * k_D);
  return 0.5 * DF_PI * Cdrag * a.df_rho_int * a.df_radius / rho;
}

// The dust fluid's part of ConditionCFL (src/SourceEuler.c:1352-1361,
// 1376-1384, 1430-1471): the ring mean DVmoy and the ring's smallest
// CFLSECURITY / sqrt(sum of the six limits squared), one workgroup per
// ring; k_control takes each ring's max2(min2(gas, dust), 1e-10) and the
// dust's shear limit.  Without the short-friction-time approximation the
// sound speed, motion, residual and drag limits enter; the artificial and
// the physical viscosity always.  Like the gas's CFL kernel it multiplies
// by the ring's reciprocal cell sizes.
__global__ void __launch_bounds__(256) k_df_cfl(KArgs a) {
  __shared__ double sh[1024];
  const dfargo_ctrl *c = a.ctrl;
  if (!c->go) return;
  const int i = blockIdx.x, ns = a.ns, tid = threadIdx.x, TB = blockDim.x;
  const long base = (long)i * ns;
  const double *dvr = a.dvr + base, *dvt = a.dvt + base;
  double s = 0.0;
  for (int j = tid; j < ns; j += TB) s += dvt[j];
  s = block_sum_fast(s, sh);
  const double dvmoy = s / (double)ns;
  if (tid == 0) a.dvmoy[i] = dvmoy;
  if (i < 1) {
    if (tid == 0) a.df_ring_dt[i] = 1e30;
    return;
  }
  const double dxrad = a.Rsup[i] - a.Rinf[i], dxtheta = a.Rmed[i] * a.pm / (double)ns;
  const double dxmin = (dxrad < dxtheta) ? dxrad : dxtheta;
  const double idxmin = 1.0 / dxmin, idxrad = 1.0 / dxrad, idxtheta = 1.0 / dxtheta;
  const double dnu = a.df_nu[i];
  const double invdt12 = (dnu != 0.0) ? dnu * 4.0 / a.powmin2[i] : 1e-10;
  const double invdt7 = a.df_sfta ? 1e-10 : a.df_cs[i] * idxmin;
  const double om = sqrt(DF_G * 1.0 / a.Rmed[i]) / a.Rmed[i];
  double smax = 0.0;
  for (int j = tid; j < ns; j += TB) {
    const int jp = (j == ns - 1) ? 0 : j + 1;
    double invdt8 = 1e-10, invdt9 = 1e-10, invdt10 = 1e-10;
    if (!a.df_sfta) {
      invdt8 = fabs(dvr[j] + 1e-15) * idxrad;
      invdt9 = fabs(a.fast ? dvt[j] - dvmoy : dvt[j]) * idxtheta;
      if (a.df_feel_disk) invdt10 = 1. / (a.dstokes[base + j] + 1.e-10) * om;
    }
    double ddvr = dvr[j + ns] - dvr[j], ddvt = dvt[jp] - dvt[j];
    if (ddvr >= 0.0) ddvr = 1e-10; else ddvr = -ddvr;
    if (ddvt >= 0.0) ddvt = 1e-10; else ddvt = -ddvt;
    const double d4 = ddvr * idxrad, d5 = ddvt * idxtheta;
    double invdt11 = (d5 > d4) ? d5 : d4;
    invdt11 *= 4.0 * DF_CVNR * DF_CVNR;
    const double S = invdt7 * invdt7 + invdt8 * invdt8 + invdt9 * invdt9 + invdt10 * invdt10 +
                     invdt11 * invdt11 + invdt12 * invdt12;
    smax = (S > smax) ? S : smax;
  }
  // the largest sum gives the smallest dt (the map is monotone)
  smax = -block_min_fast(-smax, sh);
  if (tid == 0) a.df_ring_dt[i] = DF_CFLSECURITY / sqrt(smax);
}

// ring means of SG_Accr (mpi_make1Dprofile for DUSTFEELSG Z,
// src/sgupdate.c:26-27), one workgroup per ring
__global__ void __launch_bounds__(256) k_df_axisg(KArgs a) {
  __shared__ double sh[1024];
  const dfargo_ctrl *c = a.ctrl;
  if (!c->active) return;
  const int i = blockIdx.x, ns = a.ns;
  double s = 0.0;
  for (int j = threadIdx.x; j < ns; j += blockDim.x) s += ((const double *)a.sgr)[(long)i * ns + j];
  s = block_sum_fast(s, sh);
  if (threadIdx.x == 0) a.df_axisg[i] = s / (double)ns;
}

// The dust's part of SubStep1 (src/SourceEuler.c:805-999) and of
// update_sgvelocity (src/sgupdate.c:29-70): DVradInt on rings >= 1,
// DVthetaInt everywhere.  Tiles of DFS_R rings x DFS_C columns; the
// potential, both pressures, both densities and the dust's vtheta of the
// tile and its inner and preceding neighbours are staged in LDS once (the
// potential, a planet sum per cell, is the costly part).  In a corotating
// frame both fluids' vtheta are read through CorrectVtheta
// (src/SourceEuler.c:516-523).
enum { DFS_R = 4, DFS_C = 64 };
__global__ void __launch_bounds__(DFS_R * DFS_C) k_df_source(KArgs a) {
  constexpr int PR = DFS_R + 1, PC = DFS_C + 2; // rings i0-1 .., columns j0-1 .. j0+DFS_C
  __shared__ double sPot[PR][PC], sP[PR][PC], sRho[PR][PC], sD[PR][PC], sDP[PR][PC], sDvt[PR][PC];
  __shared__ double sLam[PR];
  const dfargo_ctrl *c = a.ctrl;
  if (!c->active) return;
  const int ns = a.ns, nr = a.nr, tid = threadIdx.x;
  const int j0 = blockIdx.x * DFS_C, i0 = blockIdx.y * DFS_R;
  const double dt = c->dt, dom = a.frame ? c->domega : 0.0;
  const double *rho = (const double *)a.rho, *vr = (const double *)a.vr, *vt = (const double *)a.vt;
  const double *drho = a.drho, *dvr = a.dvr, *dvt = a.dvt;
  for (int idx = tid; idx < PR * PC; idx += DFS_R * DFS_C) {
    const int q = idx / PC, p = idx - q * PC, r = i0 - 1 + q;
    int jj = (j0 - 1 + p) % ns;
    if (jj < 0) jj += ns;
    const bool in = (r >= 0 && r <= nr - 1);
    const long m = (long)r * ns + jj;
    sPot[q][p] = in ? pot_at(a, c, r, jj) : 0.0;
    sP[q][p] = in ? src_press<double>(a, m, r) : 0.0;
    sRho[q][p] = in ? rho[m] : 0.0;
    sD[q][p] = in ? drho[m] : 0.0;
    sDP[q][p] = in ? drho[m] * a.df_cs[r] * a.df_cs[r] : 0.0;
    sDvt[q][p] = in ? (a.frame ? dvt[m] - dom * a.Rmed[r] : dvt[m]) : 0.0;
  }
  if (tid < PR) {
    const int r = i0 - 1 + tid;
    sLam[tid] = (r >= 0 && r <= nr - 1) ? 3.34e-8 * d_aspect_ratio(a, a.Rmed[r]) * a.Rmed[r] / a.df_to_cgs : 0.0;
  }
  __syncthreads();
  const int ty = tid / DFS_C, tx = tid - ty * DFS_C, i = i0 + ty, j = j0 + tx;
  if (i > nr - 1 || j >= ns) return;
  const int q = ty + 1, p = tx + 1; // the cell's LDS slot
  const long l = (long)i * ns + j, lim = l - ns;
  const double ci = dom * a.Rmed[i];
  const double rl = sRho[q][p], vtl = a.frame ? vt[l] - ci : vt[l], dvtl = sDvt[q][p];
  const double cs = a.energy ? ((const double *)a.cs)[l] : a.cs_iso[i];
  const double St = df_stokes_at(a, sLam[q], rl, dvr[l] - vr[l], dvtl - vtl, cs);
  a.dstokes[l] = St;
  const double invomega = a.r15[i];
  const double pl = sP[q][p], potl = sPot[q][p], dl = sD[q][p], dpl = sDP[q][p];
  const int dsg = a.sg ? a.df_feel_sg : DFARGO_DUST_SG_NO;
  if (i >= 1) {
    const double idrm = a.InvDiffRmed[i];
    const double gradp = (pl - sP[q - 1][p]) * 2.0 / (rl + sRho[q - 1][p]) * idrm;
    const double gradphi = (potl - sPot[q - 1][p]) * idrm;
    double ts = St * invomega, dvri;
    if (!a.df_sfta) {
      const double dgradp = (dpl - sDP[q - 1][p]) * 2.0 / (dl + sD[q - 1][p]) * idrm;
      double dvt2 = dvtl + sDvt[q][p + 1] + sDvt[q - 1][p] + sDvt[q - 1][p + 1];
      dvt2 = dvt2 / 4.0 + a.Rinf[i] * c->omega_frame;
      dvt2 = dvt2 * dvt2;
      dvri = dvr[l] + dt * (-gradphi - dgradp + dvt2 * a.InvRinf[i]);
      if (a.df_feel_disk) dvri = (dvri * ts + vr[l] * dt) / (ts + dt);
    } else {
      // the short-friction-time velocity (Johansen & Klahr 2005); the
      // option needs DUSTFEELDISK
      if (a.df_feedback) ts /= (1.0 + dl / rl);
      dvri = vr[l] + ts * gradp;
    }
    if (dsg != DFARGO_DUST_SG_NO) {
      const double *sgr = (const double *)a.sgr;
      if (a.sg_zero || dsg == DFARGO_DUST_SG_YES)
        dvri = dvri + dt * (((a.Rinf[i] - a.Rmed[i - 1]) * sgr[l] + (a.Rmed[i] - a.Rinf[i]) * sgr[lim]) * idrm);
      else
        dvri = dvri + dt * (((a.Rinf[i] - a.Rmed[i - 1]) * a.df_axisg[i] + (a.Rmed[i] - a.Rinf[i]) * a.df_axisg[i - 1]) *
                            idrm);
    }
    a.dvri[l] = dvri;
  }
  const double invdxtheta = a.InvDxTheta[i];
  const double gradp = (pl - sP[q][p - 1]) * 2.0 / (rl + sRho[q][p - 1]) * invdxtheta;
  const double gradphi = (potl - sPot[q][p - 1]) * invdxtheta;
  double ts = St * invomega, dvti;
  if (!a.df_sfta) {
    const double dgradp = (dpl - sDP[q][p - 1]) * 2.0 / (dl + sD[q][p - 1]) * invdxtheta;
    dvti = dvtl - dt * (gradphi + dgradp);
    if (a.df_feel_disk) dvti = (dvti * ts + vtl * dt) / (ts + dt);
  } else {
    if (a.df_feedback) ts /= (1.0 + dl / rl);
    dvti = vtl + ts * gradp;
  }
  if (dsg == DFARGO_DUST_SG_YES && !a.sg_zero) {
    const double *sgt = (const double *)a.sgt;
    const long ljm = (long)i * ns + ((j == 0) ? ns - 1 : j - 1);
    dvti = dvti + dt * (0.5 * (sgt[l] + sgt[ljm]));
  }
  a.dvti[l] = dvti;
}

// The dust's stresses of ComputeViscousTerms (src/Viscosity.c:189-298)
// from DVradInt, DVthetaInt, with DFViscosity per ring: TAUrr, TAUpp on
// rings 0 .. nrad-1, TAUrp on rings 1 .. nrad-1.
__global__ void __launch_bounds__(256) k_df_stress(KArgs a) {
  const dfargo_ctrl *c = a.ctrl;
  if (!c->active) return;
  const int i = blockIdx.y, ns = a.ns;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ns) return;
  const long l = (long)i * ns + j, lip = l + ns, lim = l - ns;
  const int jm = (j == 0) ? ns - 1 : j - 1, jp = (j == ns - 1) ? 0 : j + 1;
  const long ljm = (long)i * ns + jm, ljp = (long)i * ns + jp;
  const double *vr = a.dvri, *vt = a.dvti, *drho = a.drho;
  const double onethird = 1.0 / 3.0, invdphi = a.invdphi, nu = a.df_nu[i];
  double div = (vr[lip] * a.Rsup[i] - vr[l] * a.Rinf[i]) * a.InvDiffRsup[i] * a.InvRmed[i];
  div += (vt[ljp] - vt[l]) * invdphi * a.InvRmed[i];
  const double Drr = (vr[lip] - vr[l]) * a.InvDiffRsup[i];
  const double Dpp = (vt[ljp] - vt[l]) * invdphi * a.InvRmed[i] + 0.5 * (vr[lip] + vr[l]) * a.InvRmed[i];
  a.dtrr[l] = 2.0 * drho[l] * nu * (Drr - onethird * div);
  a.dtpp[l] = 2.0 * drho[l] * nu * (Dpp - onethird * div);
  if (i >= 1) {
    const double Drp = 0.5 * (a.Rinf[i] * (vt[l] * a.InvRmed[i] - vt[lim] * a.InvRmed[i - 1]) * a.InvDiffRmed[i] +
                              (vr[l] - vr[ljm]) * invdphi * a.InvRinf[i]);
    a.dtrp[l] = 2.0 * 0.25 * (drho[l] + drho[lim] + drho[ljm] + drho[ljm - ns]) * nu * Drp;
  }
}

// the dust's part of UpdateVelocitiesWithViscosity (src/Viscosity.c:301-380)
__global__ void __launch_bounds__(256) k_df_visc(KArgs a) {
  const dfargo_ctrl *c = a.ctrl;
  if (!c->active) return;
  const int i = blockIdx.y, ns = a.ns, nr = a.nr;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ns || i < 1) return;
  const long l = (long)i * ns + j, lip = l + ns, lim = l - ns;
  const int jm = (j == 0) ? ns - 1 : j - 1, jp = (j == ns - 1) ? 0 : j + 1;
  const long ljm = (long)i * ns + jm, ljp = (long)i * ns + jp;
  const double *Trr = a.dtrr, *Trp = a.dtrp, *Tpp = a.dtpp, *rho = a.drho;
  const double DeltaT = c->dt, invdphi = a.invdphi;
  if (i <= nr - 2)
    a.dvti[l] += DeltaT * a.InvRmed[i] *
                 ((a.Rsup[i] * Trp[lip] - a.Rinf[i] * Trp[l]) * a.InvDiffRsup[i] + (Tpp[l] - Tpp[ljm]) * invdphi +
                  0.5 * (Trp[l] + Trp[lip])) /
                 (0.5 * (rho[l] + rho[ljm]));
  a.dvri[l] += DeltaT * a.InvRinf[i] *
               ((a.Rmed[i] * Trr[l] - a.Rmed[i - 1] * Trr[lim]) * a.InvDiffRmed[i] + (Trp[ljp] - Trp[l]) * invdphi -
                0.5 * (Tpp[l] + Tpp[lim])) /
               (0.5 * (rho[l] + rho[lim]));
}

// The dust's part of SubStep2 (src/SourceEuler.c:1060-1102) into the
// fluid's state: DVradNew, DVthetaNew, or in the short-friction-time
// approximation DVradInt, DVthetaInt as they are (src/SourceEuler.c:600-630).
// vr of rings 0 and nrad is zero (the reference never sets it there).  One
// wavefront per DF_SRC_COLS-column tile of a ring, which leaves the tile's
// sum of vt dt / dt for the ring mean of the FARGO transport.
__global__ void __launch_bounds__(64) k_df_substep2(KArgs a) {
  const dfargo_ctrl *c = a.ctrl;
  if (!c->active) return;
  const int i = blockIdx.y, ns = a.ns, nr = a.nr;
  const int j = blockIdx.x * DF_SRC_COLS + threadIdx.x;
  const double dt = c->dt, invdt = 1.0 / dt;
  const double *vr = a.dvri, *vt = a.dvti, *drho = a.drho;
  double rate = 0.0;
  if (j < ns) {
    const long l = (long)i * ns + j;
    if (i == nr) {
      a.dvr[l] = 0.0;
      return;
    }
    const int jm = (j == 0) ? ns - 1 : j - 1, jp = (j == ns - 1) ? 0 : j + 1;
    const long ljm = (long)i * ns + jm, ljp = (long)i * ns + jp, lim = l - ns;
    double vrn, vtn;
    if (a.df_sfta) {
      vrn = (i == 0) ? 0.0 : vr[l];
      vtn = vt[l];
    } else {
      const double C2 = DF_CVNR * DF_CVNR;
      double dv = vr[l + ns] - vr[l];
      const double qr = (dv < 0.0) ? C2 * drho[l] * dv * dv : 0.0;
      vrn = 0.0;
      if (i >= 1) {
        dv = vr[l] - vr[lim];
        const double qrm = (dv < 0.0) ? C2 * drho[lim] * dv * dv : 0.0;
        vrn = vr[l] - dt * 2.0 / (drho[l] + drho[lim]) * (qr - qrm) * a.InvDiffRmed[i];
      }
      dv = vt[ljp] - vt[l];
      const double qt = (dv < 0.0) ? C2 * drho[l] * dv * dv : 0.0;
      dv = vt[l] - vt[ljm];
      const double qtm = (dv < 0.0) ? C2 * drho[ljm] * dv * dv : 0.0;
      vtn = vt[l] - dt * 2.0 / (drho[l] + drho[ljm]) * (qt - qtm) * a.InvDxTheta[i];
    }
    a.dvr[l] = vrn;
    a.dvt[l] = vtn;
    rate = (vtn * dt) * invdt;
  }
  if (i == nr) return;
  rate = wave_sum(rate);
  if (threadIdx.x == 0) a.dvt_part[(long)i * gridDim.x + blockIdx.x] = rate;
}

// the transport kernels' argument block on the dust fluid (Transportd,
// src/TransportEuler.c:553-568): the dust's state and buffers in the gas's
// places, no energy
__host__ __device__ inline KArgs df_transport_args(const KArgs &a) {
  KArgs d = a;
  d.energy = 0;
  d.rho = a.drho;
  d.vri = a.dvr;
  d.vti = a.dvt;
  d.t_rho = a.dt_rho;
  d.t_rp = a.dt_rp;
  d.t_rm = a.dt_rm;
  d.t_tp = a.dt_tp;
  d.t_tm = a.dt_tm;
  d.z_rho = a.dz_rho;
  d.z_rp = a.dz_rp;
  d.z_rm = a.dz_rm;
  d.z_tp = a.dz_tp;
  d.z_tm = a.dz_tm;
  d.vt_part = a.dvt_part;
  d.vmed = a.dvmed;
  d.nshift = a.dnshift;
  d.lost_col = a.dlost_col;
  d.tr_dust = 1;
  d.tr_skip_mom = a.df_sfta;
  return d;
}

// ComputeVelocities (src/TransportEuler.c:515-547) on the dust's
// transported momenta
__device__ __forceinline__ void df_vel(const KArgs &a, const dfargo_ctrl *c, int i, int j, double &vr, double &vt) {
  const int ns = a.ns;
  const long l = (long)i * ns + j, lim = l - ns, ljm = (j == 0) ? (long)i * ns + ns - 1 : l - 1;
  const double rho = a.dz_rho[l];
  vr = 0.0;
  if (i != 0) vr = (a.dz_rp[lim] + a.dz_rm[l]) / (rho + a.dz_rho[lim] + 1e-20);
  vt = (a.dz_tp[ljm] + a.dz_tm[l]) / (rho + a.dz_rho[ljm] + 1e-20) / a.Rmed[i] - a.Rmed[i] * c->omega_frame;
}

// the gas density after its transport and floor (what Diffd reads)
__device__ __forceinline__ double df_gas_rho(const KArgs &a, long m) {
  double r = ((const double *)a.z_rho)[m];
  if (a.floors && r < a.floor_dens) r = a.floor_dens;
  return r;
}

// Diffd's fluxes (src/SourceEuler.c:2620-2635): the radial one through the
// outer face of cell m (ring im) and the azimuthal one through its
// following face, with the Schmidt number (1 + 4 St^2) / (1 + St^2)^2
__device__ __forceinline__ double df_flux_r(const KArgs &a, int im, long m, double nup) {
  const long mip = m + a.ns;
  const double St = a.dstokes[mip];
  const double q = 1.0 + St * St;
  const double Dturb = nup * (1.0 + 4.0 * St * St) * (1.0 / (q * q));
  const double r0 = df_gas_rho(a, m), r1 = df_gas_rho(a, mip);
  return -Dturb * (r1 + r0) / 2. * (a.dz_rho[mip] / r1 - a.dz_rho[m] / r0) * a.InvDiffRmed[im + 1];
}
__device__ __forceinline__ double df_flux_t(const KArgs &a, int im, long m, long mjp, double nu, double invdtheta) {
  const double St = a.dstokes[m];
  const double q = 1.0 + St * St;
  const double Dturb = nu * (1.0 + 4.0 * St * St) * (1.0 / (q * q));
  const double r0 = df_gas_rho(a, m), r1 = df_gas_rho(a, mjp);
  return -Dturb * (r1 + r0) / 2. * (a.dz_rho[mjp] / r1 - a.dz_rho[m] / r0) * invdtheta / a.Rmed[im];
}
// FViscosity at Rsup[i] (the first Rmed at or beyond it is Rmed[i+1])
__device__ __forceinline__ double df_gas_nu_sup(const KArgs &a, int i) {
  if (!a.visc_on) return 0.0;
  double v = a.visc_const;
  if (a.alpha_on) v = a.alpha * a.gcs[i + 1] * a.gcs[i + 1] * a.df_sup15[i];
  return v * a.df_cavsup[i];
}

// The dust's state of cell (i, j) at the end of the sub-step before
// OpenBoundaryd: ComputeVelocities, Diffd (src/SourceEuler.c:2599-2656),
// the floor (AddFloorDensity), then the dust's part of ApplyWallBoundaries
// (src/SideEuler.c:1066-1091) with DSigmaMed and of
// ApplySubKeplerianBoundary (src/SideEuler.c:1025-1041) for rings 0 and
// nrad-1.  Rings 0 .. nrad-1.
struct DfCell {
  double rho, vr, vt;
};
__device__ DfCell df_cell(const KArgs &a, const dfargo_ctrl *c, int i, int j) {
  const int ns = a.ns, nr = a.nr;
  const long l = (long)i * ns + j;
  DfCell s;
  df_vel(a, c, i, j, s.vr, s.vt);
  s.rho = a.dz_rho[l];
  if (a.df_diffusion && i >= 1 && i <= nr - 2 && a.Rmed[i] > a.df_diff_inn) {
    const double dtheta = 2.0 * DF_PI / (double)ns, invdtheta = 1.0 / dtheta;
    const int jm = (j == 0) ? ns - 1 : j - 1, jp = (j == ns - 1) ? 0 : j + 1;
    const long lim = l - ns, ljm = (long)i * ns + jm, ljp = (long)i * ns + jp;
    const double nu = a.visc_on ? a.nu_ring[i] : 0.0;
    const double frp = df_flux_r(a, i, l, df_gas_nu_sup(a, i));
    const double frm = df_flux_r(a, i - 1, lim, df_gas_nu_sup(a, i - 1));
    const double ftp = df_flux_t(a, i, l, ljp, nu, invdtheta);
    const double ftm = df_flux_t(a, i, ljm, l, nu, invdtheta);
    s.rho = s.rho - (a.Rsup[i] * dtheta * frp - a.Rinf[i] * dtheta * frm + (a.Rsup[i] - a.Rinf[i]) * (ftp - ftm)) *
                        c->dt * a.InvSurf[i];
  }
  if (a.floors && s.rho < a.floor_dens) s.rho = a.floor_dens;
  if (a.bc == DFARGO_BC_WALL) {
    if (i == 0) {
      double v2, t2;
      df_vel(a, c, 2, j, v2, t2);
      s.vr = -v2;
      s.rho = a.DSigmaMed[0];
    } else if (i == 1) {
      s.vr = 0.0;
      s.rho = a.DSigmaMed[1];
    }
    if (i == nr - 2) s.rho = a.DSigmaMed[nr - 2];
    if (i == nr - 1) {
      s.vr = 0.0;
      s.rho = a.DSigmaMed[nr - 1];
    }
  }
  if (i == 0) s.vt = a.DVthetaMed[0] - a.Rmed[0] * c->omega_frame;
  if (i == nr - 1) s.vt = a.DVthetaMed[nr - 1] - a.Rmed[nr - 1] * c->omega_frame;
  return s;
}

// The end of the dust's sub-step: df_cell, then with INNERBOUNDARYDUST
// OPEN the edges of OpenBoundaryd (src/SideEuler.c:411-466): rings 0 and 1
// take ring 2's density, rings 1 and 2 ring 3's vr, ring nrad-1 ring
// nrad-2's density and vr.  Rings 0 .. nrad, the last only for the wall's
// vr of the outer face.
__global__ void __launch_bounds__(256) k_df_finish(KArgs a) {
  const dfargo_ctrl *c = a.ctrl;
  if (!c->active) return;
  const int i = blockIdx.y, ns = a.ns, nr = a.nr;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ns) return;
  const long l = (long)i * ns + j;
  if (i == nr) {
    double vr = 0.0, vt;
    if (a.bc == DFARGO_BC_WALL) {
      df_vel(a, c, nr - 2, j, vr, vt);
      vr = -vr;
    }
    a.dvr[l] = vr;
    return;
  }
  DfCell s = df_cell(a, c, i, j);
  if (a.df_open) {
    if (i <= 1) s.rho = df_cell(a, c, 2, j).rho;
    if (i == 1 || i == 2) s.vr = df_cell(a, c, 3, j).vr;
    if (i == nr - 1) {
      const DfCell m = df_cell(a, c, nr - 2, j);
      s.rho = m.rho;
      s.vr = m.vr;
    }
  }
  a.drho[l] = s.rho;
  a.dvr[l] = s.vr;
  a.dvt[l] = s.vt;
}

static void launch_df_kernel(int id, const KArgs &a, hipStream_t s) {
  const int bx = a.ns < 256 ? a.ns : 256;
  const dim3 g((a.ns + bx - 1) / bx, a.nr), b(bx);
  switch (id) {
    case KID_DF_CFL: hipLaunchKernelGGL(k_df_cfl, dim3(a.nr), dim3(256), 0, s, a); break;
    case KID_DF_AXISG: hipLaunchKernelGGL(k_df_axisg, dim3(a.nr), dim3(256), 0, s, a); break;
    case KID_DF_SOURCE:
      hipLaunchKernelGGL(k_df_source, dim3((a.ns + DFS_C - 1) / DFS_C, (a.nr + DFS_R - 1) / DFS_R),
                         dim3(DFS_R * DFS_C), 0, s, a);
      break;
    case KID_DF_STRESS: hipLaunchKernelGGL(k_df_stress, g, b, 0, s, a); break;
    case KID_DF_VISC: hipLaunchKernelGGL(k_df_visc, g, b, 0, s, a); break;
    case KID_DF_SUBSTEP2: {
      const int ntile = (a.ns + DF_SRC_COLS - 1) / DF_SRC_COLS;
      hipLaunchKernelGGL(k_df_substep2, dim3(ntile, a.nr + 1), dim3(DF_SRC_COLS), 0, s, a);
      break;
    }
    case KID_DF_TRANSPORT_RAD: {
      const KArgs d = df_transport_args(a);
      hipLaunchKernelGGL(k_transport_rad<double>, dim3((a.ns + RT_C - 1) / RT_C, (a.nr + RT_R - 1) / RT_R),
                         dim3(RT_C, RT_R), 0, s, d);
      break;
    }
    case KID_DF_TRANSPORT_AZI: {
      const KArgs d = df_transport_args(a);
      const int per = (a.tune_azi_waves == 1 || a.tune_azi_waves == 2) ? a.tune_azi_waves : 4;
      const int nseg = (a.ns + AZ_OWN - 1) / AZ_OWN;
      hipLaunchKernelGGL(k_transport_azi<double>, dim3((nseg + per - 1) / per, a.nr), dim3(per * AZ_LANES), 0, s, d);
      break;
    }
    case KID_DF_FINISH: hipLaunchKernelGGL(k_df_finish, dim3(g.x, a.nr + 1), b, 0, s, a); break;
    default: break;
  }
}
