/*
 * oracle.h -- FP64 CPU restatement of Dusty FARGO-ADSG's gas hydro step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links or loads this
 * code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg use it, as the checker the HIP path is compared against.
 *
 * The reference cannot be built here: its sources include headers that
 * src/varparser.pl generates at build time (param.h, global_ex.h, sg_ex.h),
 * and its self-gravity needs FFTW-MPI.  This file restates the algorithm
 * from the sources instead; every function cites the reference lines it
 * follows.  How each function is pinned is listed in oracle.c's header.
 */
#ifndef DFARGO_ORACLE_H
#define DFARGO_ORACLE_H

#include "../include/dfargo.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx or_ctx;

or_ctx *or_create(const dfargo_params *p);
void or_destroy(or_ctx *c);

/* tables of FillPolar1DArrays: radii (nrad+1), rmed, surf (nrad), azimuth (nsec) */
void or_get_tables(or_ctx *c, double *radii, double *rmed, double *surf,
                   double *azimuth);

/* host initial state (InitGasDensity, InitGasEnergy, InitPlanetarySystem) */
void or_init_gas_density(or_ctx *c, double *rho);
void or_init_gas_energy(or_ctx *c, double *energy);
void or_init_planetary_system(or_ctx *c, int nb, const double *dist,
                              const double *mass, const double *acc,
                              const int *feel_disk, const int *feel_others,
                              dfargo_planets *out);

/* start-up SG call + InitEuler; writes the initial velocities */
int or_init_euler(or_ctx *c, const double *rho, double *vrad, double *vtheta,
                  const double *energy, const dfargo_planets *sys);

void or_set_state(or_ctx *c, const double *rho, const double *vrad,
                  const double *vtheta, const double *energy,
                  const dfargo_planets *sys);
void or_get_state(or_ctx *c, double *rho, double *vrad, double *vtheta,
                  double *energy, dfargo_planets *sys);
void or_set_time(or_ctx *c, double t, double t0, double omega_frame);
/* EvanescentBoundary on the current state, for the pins of its formula */
void or_evanescent_boundary(or_ctx *c, double step);
/* NonReflectingBoundary on the current state (barotropic), for its pins */
void or_nonreflecting_boundary(or_ctx *c);
/* AccreteOntoPlanets on the current state, for its pins */
void or_accrete_onto_planets(or_ctx *c, double dt);
/* compute_SGZeroMode: SG_Accr per ring (nrad values) of the density rho */
void or_sg_zero_mode(or_ctx *c, const double *rho, double *accr);
/* the DAMPTOVISCOUS 1-D disc: one step of dt (none for dt <= 0), then its
 * Sigma and vr (nrad1d values each) */
void or_solve_1d_viscous(or_ctx *c, double dt, double *sigma, double *vrad);
/* the 1-D disc's Rmed1D, Sigma1D, Vrad1D (nrad1d values each; NULL skips):
 * set as the restart read sets them, and read back */
void or_set_visc1d(or_ctx *c, const double *rmed, const double *sigma, const double *vrad);
void or_get_visc1d(or_ctx *c, double *rmed, double *sigma, double *vrad);

/* the dust fluid: switched on with its initial DRho ((nrad + 1) * nsec);
 * or_init_euler then sets its Stokes number and velocities */
int or_dust_fluid_attach(or_ctx *c, const dfargo_dust_fluid_params *f, const double *drho);
void or_dust_fluid_get(or_ctx *c, double *drho, double *dvr, double *dvt, double *stokes);
void or_dust_fluid_set(or_ctx *c, const double *drho, const double *dvr, const double *dvt);

/* AlgoGas; returns the number of sub-steps taken */
int or_algogas(or_ctx *c, double DT);
void or_get_status(or_ctx *c, dfargo_status *st);
void or_get_field(or_ctx *c, int field, double *out);
void or_get_fargo_shift(or_ctx *c, long *nshift, double *vmed);

/* pieces, for pinning tests */
int or_condition_cfl(or_ctx *c, double deltaT);
void or_transport(or_ctx *c, double dt);
void or_substep1(or_ctx *c, double dt);
void or_viscous(or_ctx *c, double dt);     /* ComputeViscousTerms + Update on Vrad/Vtheta */
void or_sg_accel(or_ctx *c, const double *rho, double *accr, double *acct);
void or_sg_accel_direct(or_ctx *c, const double *rho, double *accr, double *acct);
void or_fft(double *re, double *im, int n, int sign);
void or_rk5(const double *q0, double dt, const double *masses, double *q1,
            int n, const int *feelothers, int indirect_term);
double or_fviscosity(or_ctx *c, double r);

/* per-DT diagnostics: GasTotalMass, GasMomentum, GasTotalEnergy, planets'
   angular momentum; UpdateLog's torques/powers (8 per planet) and indirect
   torque (returns the number of planets); FindOrbitalElements */
void or_gas_budgets(or_ctx *c, double out[4]);
int or_update_log(or_ctx *c, double *tq, double *ind);
void or_orbital_elements(double x, double y, double vx, double vy, double m, double out[7]);

/* dust particles */
int or_dust_init_system(const dfargo_params *p, const dfargo_dust_params *d, const double *axi_sg_accr,
                        double *size, double *r, double *th, double *vr, double *vth, double *l);
/* attach after or_init_euler; runs the start-up interpolation with the
   given time step (src/main.c:287 passes DT) */
int or_dust_attach(or_ctx *c, const dfargo_dust_params *d, const double *size, const double *r,
                   const double *th, const double *vr, const double *vth, const double *l, double timestep);
void or_dust_interpolation(or_ctx *c, double timestep);
void or_dust_get(or_ctx *c, double *r, double *th, double *vr, double *vth, double *l, double *stokes);
void or_dust_get_gas_at_particles(or_ctx *c, double *vrg, double *vtg, double *densg, double *csg);
void or_dust_step(or_ctx *c, double dt);
void or_dust_get_drag(or_ctx *c, double *raddrag, double *azidrag);
double or_dust_min_stopping_time(or_ctx *c);
double or_dust_particles_mass(or_ctx *c);

/* ---- test hooks: one stage of the sub-step on the context's grids ------ */
enum or_test_grid_id {
  OR_G_RHO, OR_G_VR, OR_G_VT, OR_G_EN, OR_G_CS, OR_G_PRES, OR_G_POT, OR_G_VRADINT, OR_G_VTHETAINT,
  OR_G_VRADNEW, OR_G_VTHETANEW, OR_G_ENERGYINT, OR_G_ENERGYNEW, OR_G_QR, OR_G_QT, OR_G_DIVV, OR_G_DRR,
  OR_G_DPP, OR_G_DRP, OR_G_TRR, OR_G_TPP, OR_G_TRP, OR_G_VISCHEAT, OR_G_INDPOT, OR_G_COUNT
};
enum or_test_stage_id {
  OR_S_EOS,           /* ComputeSoundSpeed + ComputePressureField          */
  OR_S_SUBSTEP1,      /* SubStep1 (+SG, viscosity) into VradInt/VthetaInt  */
  OR_S_SUBSTEP2,      /* SubStep2 from VradInt/VthetaInt                    */
  OR_S_VISC_TERMS,    /* ComputeViscousTerms on Vrad/Vtheta                 */
  OR_S_VISC_HEATING,  /* ComputeViscousHeating from the stress grids        */
  OR_S_DIVERGENCE,    /* ComputeDivergenceVelocity on Vrad/Vtheta           */
  OR_S_SUBSTEP3,      /* SubStep3 from EnergyInt, ViscHeat, DivV, Vtheta    */
  OR_S_POTENTIAL,     /* FillForcesArrays                                   */
  OR_S_BOUNDARY,      /* ApplyBoundaryCondition                             */
  OR_S_FLOORS,        /* AddFloorDensity + AddFloorEnergy                   */
  OR_S_CRASH          /* DetectCrash(Rho) | DetectCrash(Energy) << 1        */
};
/* copy (nrad+1)*nsec doubles of grid `which` out of (set = 0) or into
 * (set = 1) the context; -1 for an unknown grid */
int or_test_grid(or_ctx *c, int which, double *buf, int set);
/* run one stage; its return value (OR_S_CRASH: the two flags), -1 unknown */
int or_test_stage(or_ctx *c, int stage, double dt);

#ifdef __cplusplus
}
#endif
#endif
