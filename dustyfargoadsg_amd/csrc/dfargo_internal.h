/*
 * dfargo_internal.h -- structures shared by the host C code and the HIP
 * sources of libdfargo.  Not part of the public ABI.
 */
#ifndef DFARGO_INTERNAL_H
#define DFARGO_INTERNAL_H

#include <stdarg.h>

#include "../../include/dfargo.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Host copies of the 1-D tables of FillPolar1DArrays
 * (src/SourceEuler.c:71-188) and the per-ring constants the kernels read,
 * all evaluated with the host libm. */
typedef struct dfargo_tables {
  int nr, ns;
  double *Radii, *Rinf, *Rsup, *Rmed, *Surf, *InvRmed, *InvSurf;
  double *InvDiffRmed, *InvDiffRsup, *InvRinf, *CosAz, *SinAz, *Azimuth;
  double *SigmaMed, *SigmaInf, *EnergyMed;
  double *cs_iso;      /* isothermal sound speed per ring                 */
  double *powmin2;     /* pow(min(dr, r dphi), 2.0) of ConditionCFL       */
  double *r15;         /* pow(Rmed, 1.5) of FViscosity                    */
  double *visc_mult;   /* cavity factor of FViscosity (1.0 outside)       */
  double *beta;        /* ComputeBetaCooling(Rmed)                         */
  double *supp_torque; /* imposed-drift torque of SubStep1                */
  double *sg_sqrt;     /* sqrt(Rmed/Rmed[0])                               */
  double *sg_normr, *sg_normt; /* SG normalisation per ring               */
  double sg_eps, sg_rstep, sg_tstep;
  double vh_lr1, vh_lr2; /* log(r0/r1), log(r1/r2) of ComputeViscousHeating */
} dfargo_tables;

int dfargo_host_tables(const dfargo_params *p, dfargo_tables *t);
void dfargo_free_tables(dfargo_tables *t);
double dfargo_sigma_profile(const dfargo_params *p, double r);
double dfargo_energy_profile(const dfargo_params *p, double r);
int dfargo_fft_length_ok(int n);
int dfargo_set_error(int code, const char *fmt, ...);

#ifdef __cplusplus
}
#endif

#define DF_MAXPL DFARGO_MAX_PLANETS
#define DF_MAXTG (DF_MAXPL + 1) /* force-sum targets: the star + planets */
#define DF_NPART(ntg) (5 + 4 * (ntg))

/* Device control block: the scalars of AlgoGas that evolve sub-step by
 * sub-step.  Lives in device memory; one copy per context. */
typedef struct dfargo_ctrl {
  double DT, dtemp, dt, time, time0, omega_frame, mindt, maxdt, lost_mass, mass_taper;
  double lost_massd; /* LostMassd: the dust fluid's mass through the inner edge */
  double dtnbody;
  double dop_x, dop_y, dopalt_x, dopalt_y, ind_x, ind_y;
  double barx[3], bary[3];
  /* planetary system */
  double pl_m[DF_MAXPL], pl_x[DF_MAXPL], pl_y[DF_MAXPL], pl_vx[DF_MAXPL], pl_vy[DF_MAXPL];
  double pl_acc[DF_MAXPL], pl_m0[DF_MAXPL], pl_m1[DF_MAXPL];
  int pl_feel_disk[DF_MAXPL], pl_feel_others[DF_MAXPL], pl_binary[DF_MAXPL];
  int nb;
  /* potential of this sub-step (FillForcesArrays) */
  double pot_x[DF_MAXPL], pot_y[DF_MAXPL], pot_m[DF_MAXPL], pot_smooth[DF_MAXPL];
  double pot_invd3[DF_MAXPL];
  /* force-sum targets of this sub-step: 0 = star, then FeelDisk planets */
  double tg_x[DF_MAXTG], tg_y[DF_MAXTG], tg_rsm[DF_MAXTG], tg_a[DF_MAXTG], tg_rh[DF_MAXTG];
  int tg_planet[DF_MAXTG];
  int ntarget;
  int go, active, gas_timesteps_cfl, substeps, crashed_d, crashed_e;
  int go_next; /* the loop condition for the next sub-step, set by the control kernel */
  long substeps_total;
  /* dust particles: super-particle mass (tapered like the planets) */
  double particles_mass, particles_mass0;
  /* the super-particle and planet masses of the current sub-step, kept by
     the control kernel for the particle kernels that run after it (it
     already tapers the masses for the next sub-step) */
  double dust_mp, dust_pl_m[DF_MAXPL];
  /* ... and the planets' positions the particles feel (after the n-body
     step, before RotatePsys) and the frame rate their interpolation reads
     the gas in (OmegaFrame before this sub-step's frame update) */
  double dust_pl_x[DF_MAXPL], dust_pl_y[DF_MAXPL], omega_prev;
  /* the particle set holding the current state, and the particles outside
     the grid counted by this sub-step's move */
  int pset, out_count;
  /* corotating frame: the change of OmegaFrame this sub-step
     (CorrectVtheta's domega) and GetPsysInfo's mark */
  double domega, mark_x, mark_y;
  /* the smallest 1/2 - |Ntilde - Nshift| of this AlgoGas call, as the bits
     of a non-negative double (they order as unsigned integers) */
  unsigned long long fargo_margin_bits;
} dfargo_ctrl;

#endif
