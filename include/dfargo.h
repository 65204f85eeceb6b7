/*
 * dfargo.h -- C ABI of the MI355X-native Dusty FARGO-ADSG gas hydro step.
 *
 * Every entry point takes plain pointers, sizes and the parameter block
 * below; no HIP or torch type appears in a signature.  Field arrays use the
 * reference's polar layout: ring-major doubles, l = j + i*nsec, with
 * (nrad+1)*nsec entries (ring nrad holds the outer-interface radial velocity).
 *
 * Each entry point names the reference function it replaces (file:line in
 * the Dusty FARGO-ADSG sources).  Errors: a non-zero DFARGO_ERR_* return and
 * a message from dfargo_last_error(); nothing calls exit().
 */
#ifndef DFARGO_H
#define DFARGO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DFARGO_MAX_PLANETS 8

enum dfargo_status_code {
  DFARGO_OK = 0,
  DFARGO_ERR_ARG = 1,          /* invalid argument or size                  */
  DFARGO_ERR_UNSUPPORTED = 2,  /* parameter combination not implemented     */
  DFARGO_ERR_HIP = 3,          /* HIP runtime failure (message has detail)   */
  DFARGO_ERR_ALLOC = 4,        /* host or device allocation failed           */
  DFARGO_ERR_STATE = 5         /* call made before the required set-up call  */
};

enum dfargo_precision {
  DFARGO_FP64 = 0,   /* fields stored and stenciled in double            */
  DFARGO_FP32 = 1    /* fields in float; FP64 for ring means, force sums,
                        planets, dt and budgets                          */
};

enum dfargo_boundary {
  DFARGO_BC_WALL = 0,  /* INNERBOUNDARY WALL (default, src/var.c:20)       */
  DFARGO_BC_OPEN = 1,  /* INNERBOUNDARY OPEN (src/SideEuler.c:300-408)     */
  DFARGO_BC_NONREFLECTING = 2 /* INNERBOUNDARY NONREFLECTING
                                 (src/SideEuler.c:569-794)                 */
};

/* FRAME (src/Interpret.c:225-233): fixed (rotating at OMEGAFRAME) or
 * corotating with a planet, its guiding centre or a binary's barycentre
 * (GetPsysInfo, src/Psys.c:233-350). */
enum dfargo_frame {
  DFARGO_FRAME_FIXED = 0,           /* FRAME F                               */
  DFARGO_FRAME_COROTATING = 1,      /* FRAME C                               */
  DFARGO_FRAME_GUIDING_CENTER = 2,  /* FRAME G                               */
  DFARGO_FRAME_BINARY_CENTER = 3    /* FRAME B                               */
};

/* What the wave-killing zones damp towards (src/Interpret.c:344-358). */
enum dfargo_damp {
  DFARGO_DAMP_TO_AXI = 0,      /* the ring's instantaneous mean (default)  */
  DFARGO_DAMP_TO_INI = 1,      /* the initial profiles (DAMPTOINI)         */
  DFARGO_DAMP_TO_VISCOUS = 2   /* a 1-D viscous disc evolved beside the
                                  grid (DAMPTOVISCOUS,
                                  src/SideEuler.c:1444-1565)               */
};

/* Parameter block: the .par keys the hydro step reads (src/var.c:15-202),
 * already decoded the way src/Interpret.c:165-476 decodes them. */
typedef struct dfargo_params {
  /* grid */
  int    nrad, nsec;                 /* NRAD, NSEC                          */
  double rmin, rmax;                 /* RMIN, RMAX                          */
  int    log_grid;                   /* RADIALSPACING starts with L         */
  double pmin, pmax;                 /* PMIN, PMAX                          */
  /* disc */
  double sigma0, sigma_slope;        /* SIGMA0, SIGMASLOPE                  */
  double aspect_ratio, flaring_index;/* ASPECTRATIO, FLARINGINDEX           */
  double viscosity, alpha_viscosity; /* VISCOSITY, ALPHAVISCOSITY           */
  double cavity_radius, cavity_ratio, cavity_width;          /* CAVITY*    */
  double transition_radius, transition_ratio, transition_width;
  double imposed_disk_drift;         /* IMPOSEDDISKDRIFT                    */
  int    add_noise;                  /* ADDNOISE                            */
  long   noise_seed;                 /* value srand48() is seeded with      */
  /* thermodynamics */
  int    energy_equation;            /* ENERGYEQUATION                      */
  int    viscous_heating;            /* VISCOUSHEATING (default YES)        */
  int    beta_cooling;               /* BETACOOLING                         */
  double adiabatic_index;            /* ADIABATICINDEX                      */
  double beta_cooling_time, beta_cooling_slope;
  /* self-gravity */
  int    self_gravity;               /* SELFGRAVITY YES or Z (sg_zero_mode) */
  double sg_thickness_smoothing;     /* SGTHICKNESSSMOOTHING                */
  /* planets and forces */
  double thickness_smoothing;        /* THICKNESSSMOOTHING                  */
  double roche_smoothing;            /* ROCHESMOOTHING                      */
  double indirect_term_smoothing;    /* INDIRECTTERMSMOOTHING               */
  double mass_taper;                 /* MASSTAPER (orbits)                  */
  int    indirect_term;              /* INDIRECTTERM (default YES)          */
  int    discard_gas_indirect_term;  /* DISCARDGASINDIRECTTERM              */
  int    alternative_gas_indirect_term;
  int    exclude_hill;               /* EXCLUDEHILL                         */
  double exclude_hill_factor;        /* EXCLUDEHILLFACTOR                   */
  int    forced_circular;            /* FORCEDCIRCULAR                      */
  double release_date, release_radius;
  double eccentricity;               /* ECCENTRICITY                        */
  /* frame */
  int    corotating;                 /* FRAME, enum dfargo_frame            */
  double omega_frame;                /* OMEGAFRAME, or with a corotating
                                        frame the start-up frequency the host
                                        sets (src/main.c:229-246)           */
  /* numerics */
  int    fast_transport;             /* TRANSPORT FARGO (1) or STANDARD (0) */
  int    add_floors;                 /* ADDFLOORS (default YES)             */
  double density_floor;              /* DENSITYFLOOR                        */
  int    boundary;                   /* enum dfargo_boundary                */
  /* frame (continued) */
  int    corotate_with_outer_planet; /* COROTATEWITHOUTERPLANET             */
  /* wave-killing zones (EvanescentBoundary, src/SideEuler.c:796-950) */
  int    wave_killing_zones;         /* WAVEKILLINGZONES                    */
  double wkz_rmin, wkz_rmax;         /* WKZRMIN, WKZRMAX                    */
  double damping_ratio;              /* DAMPINGTODYNAMICALTIMERATIO (0.3)   */
  int    damp_to;                    /* enum dfargo_damp                    */
  int    zero_vrad_bc;               /* ZEROVRADBC                          */
  /* self-gravity (continued) */
  int    sg_zero_mode;               /* SELFGRAVITY Z: axisymmetric part
                                        only (src/sgzero.c)                 */
  /* the 1-D viscous disc of DAMPTOVISCOUS */
  int    nrad1d;                     /* NRAD1D (100, at most 1024)          */
  double rmin1d, rmax1d;             /* RMIN1D, RMAX1D (0.01, 100)          */
  int    bc1d;                       /* BOUNDARY1DGRID: 0 low density (T),
                                        1 zero-velocity steady state (Z),
                                        2 steady state vr = -3/2 nu/R (D)   */
} dfargo_params;

/* PlanetarySystem (src/types.h:97-111) plus the two mass tables that
 * src/global.h:33 keeps beside it. */
typedef struct dfargo_planets {
  int    nb;
  double mass[DFARGO_MAX_PLANETS], x[DFARGO_MAX_PLANETS], y[DFARGO_MAX_PLANETS];
  double vx[DFARGO_MAX_PLANETS], vy[DFARGO_MAX_PLANETS], acc[DFARGO_MAX_PLANETS];
  int    feel_disk[DFARGO_MAX_PLANETS], feel_others[DFARGO_MAX_PLANETS];
  int    binary[DFARGO_MAX_PLANETS];
  double initial_mass[DFARGO_MAX_PLANETS], final_mass[DFARGO_MAX_PLANETS];
} dfargo_planets;

/* Scalar globals AlgoGas leaves host-coherent (src/global.h:32-42,
 * src/SourceEuler.c:38, src/TransportEuler.c:27). */
typedef struct dfargo_status {
  double physical_time, physical_time_initial, omega_frame;
  double mindt, maxdt, lost_mass, mass_taper;
  double disk_on_primary_x, disk_on_primary_y;         /* DiskOnPrimaryAcceleration     */
  double disk_on_primary_alt_x, disk_on_primary_alt_y; /* DiskOnPrimaryAcceleration_alt */
  double bary_x[3], bary_y[3];       /* stardiscbarycenter_x/y               */
  long   substeps_total;             /* hydro sub-steps since create         */
  int    last_call_substeps;         /* sub-steps of the last AlgoGas call   */
  int    gas_timesteps_cfl;          /* GasTimeStepsCFL of the last sub-step */
  int    crashed_density;            /* DetectCrash(Rho), last sub-step      */
  int    crashed_energy;             /* DetectCrash(Energy), last sub-step   */
  int    crashed;                    /* the reference's Crashed flag (the
                                        energy scan overwrites the density
                                        scan, src/SourceEuler.c:545-548)     */
  double fargo_margin;               /* the smallest 1/2 - |Ntilde - Nshift|
                                        over the rings and sub-steps of the
                                        last AlgoGas call: how far the FARGO
                                        shift's argument stayed from a
                                        rounding tie (src/TransportEuler.c:
                                        310-316); 0.5 before any sub-step   */
} dfargo_status;

/* Derived grids the host may read back (src/Output.c:234-284). */
enum dfargo_field {
  DFARGO_FIELD_POTENTIAL = 0,
  DFARGO_FIELD_SOUNDSPEED = 1,
  DFARGO_FIELD_PRESSURE = 2,
  DFARGO_FIELD_TEMPERATURE = 3,
  DFARGO_FIELD_DIVV = 4,
  DFARGO_FIELD_VISCHEAT = 5,
  DFARGO_FIELD_SG_ACCR = 6,      /* gr     (src/sgacc.c:157)  */
  DFARGO_FIELD_SG_ACCT = 7,      /* gtheta (src/sgacc.c:158)  */
  DFARGO_FIELD_TAURR = 8,
  DFARGO_FIELD_TAURP = 9,
  DFARGO_FIELD_TAUPP = 10,
  DFARGO_FIELD_DUST_DENSITY = 11, /* dustpcdens (src/Dsys.c:708)   */
  DFARGO_FIELD_FB_ACCR = 12,      /* RadFBAcc (src/Dsys.c:877)     */
  DFARGO_FIELD_FB_ACCT = 13,      /* AziFBAcc (src/Dsys.c:972)     */
  DFARGO_FIELD_FB_EDOT = 14       /* FBedot (src/Dsys.c:1066)      */
};

/* ---- Lagrangian dust particles (src/Dsys.c, src/DustUpdate.c) ----------- */

/* Interpolation kernels of `interpolation` (INTERPOLATION, src/Interpret.c:318-336). */
enum dfargo_interp { DFARGO_INTERP_NGP = 0, DFARGO_INTERP_CIC = 1, DFARGO_INTERP_TSC = 2 };

/* The particle switches of the .par file, decoded as src/Interpret.c does;
 * defaults of src/var.c:136-181 (dfargo_dust_default_params).  Options the
 * GPU path does not implement are fields so that they are rejected
 * loudly: dust_feel_turb, dust_growth, remove_dust_from_hill,
 * zz_integrator, dust_feel_sg_zero_mode must be 0. */
typedef struct dfargo_dust_params {
  int    nbpart;                     /* NBPART                                */
  double sizeminpart, sizemaxpart;   /* SIZEMINPART, SIZEMAXPART (m)          */
  double sizepartslope;              /* SIZEPARTSLOPE                         */
  double rhopart;                    /* RHOPART (g/cm^3)                      */
  double dustslope;                  /* DUSTSLOPE                             */
  double rmindust, rmaxdust;         /* RMINDUST, RMAXDUST (0: RMIN, RMAX)    */
  int    interpolation;              /* enum dfargo_interp                    */
  int    dust_feel_disk, dust_feel_sg, dust_feel_planets;
  int    dust_feedback;              /* DUSTFEEDBACK                          */
  int    short_friction_time_approximation; /* SFTAPPROX                     */
  int    no_dt_constraint_by_particles;     /* NODTCONSTRAINTBYPCS          */
  double dust_timestep;              /* DUSTTIMESTEP (fixed dt otherwise)     */
  double factor_unit_mass, factor_unit_length; /* FACTORUNITMASS/LENGTH      */
  double particle_mass;              /* Particles_Mass_Initial (src/main.c:190-223) */
  long   seed;                       /* seed of InitDustSystem's srand48      */
  int    dust_feel_turb, dust_growth, remove_dust_from_hill, zz_integrator;
  int    dust_feel_sg_zero_mode;
} dfargo_dust_params;

/* ---- the dust fluid (DUSTFLUID YES) -------------------------------------- */

/* The dust-fluid switches of the .par file (src/var.c:136-186, decoded as
 * src/Interpret.c:253-342 does; dfargo_dust_fluid_default_params holds the
 * defaults).  The fluid is one grain size, SIZEPART, whose Stokes number
 * follows from the local gas state (src/SourceEuler.c:820-858). */
enum dfargo_dust_feel_sg { DFARGO_DUST_SG_NO = 0, DFARGO_DUST_SG_YES = 1, DFARGO_DUST_SG_ZERO_MODE = 2 };
typedef struct dfargo_dust_fluid_params {
  double sizepart;                   /* SIZEPART (m, 1e-3)                    */
  double rhopart;                    /* RHOPART (g/cm^3, 1)                   */
  double factor_unit_mass, factor_unit_length; /* FACTORUNITMASS/LENGTH (1)  */
  double dust_to_gas_density_ratio;  /* DUSTTOGASDENSITYRATIO (0.01)          */
  double dviscosity, dalpha_viscosity; /* DVISCOSITY, DALPHAVISCOSITY (0)    */
  double daspect_ratio;              /* DASPECTRATIO (0.05)                   */
  double dflaring_index;             /* DFLARINGINDEX (0)                     */
  double dust_diff_inn;              /* DUSTDIFFINN (0)                       */
  int    sfta;                       /* SFTAPPROX (YES): dust velocities from
                                        the short-friction-time approximation */
  int    dust_feel_disk;             /* DUSTFEELDISK (YES)                    */
  int    dust_feedback;              /* DUSTFEEDBACK (NO): drag on the gas    */
  int    dust_feel_sg;               /* DUSTFEELSG, enum dfargo_dust_feel_sg  */
  int    dust_diffusion;             /* DUSTDIFFUSION (NO)                    */
  int    open_inner_dust;            /* INNERBOUNDARYDUST OPEN: OpenBoundaryd
                                        (src/SideEuler.c:411-466) instead of
                                        the wall, and no damping of the fluid
                                        in the wave-killing zones            */
} dfargo_dust_fluid_params;

/* The particle arrays of DustSystem (src/types.h:114-146) the step reads
 * and writes, nbpart entries each. */
typedef struct dfargo_dust_state {
  double *r, *th, *vr, *vth, *l;     /* position, velocity, R*vth            */
  double *dustsize, *dustsize_init;  /* code units                            */
  double *stokesnb;                  /* Stokes number of the last interpolation */
} dfargo_dust_state;

typedef struct dfargo_ctx dfargo_ctx;

/* ---- host-side reference functions (pure CPU, glibc math) ---------------- */

/* Defaults of src/var.c:15-202 for every field of dfargo_params. */
void dfargo_default_params(dfargo_params *p);

/* Validates a parameter block; returns DFARGO_ERR_UNSUPPORTED for options
 * this build does not implement.  Replaces the checks in
 * src/Interpret.c:399-476. */
int dfargo_check_params(const dfargo_params *p);

/* FillPolar1DArrays (src/SourceEuler.c:71-188) without the file output.
 * radii: nrad+1; rmed, surf: nrad; azimuth: nsec. */
int dfargo_fill_polar_1d_arrays(const dfargo_params *p, double *radii,
                                double *rmed, double *surf, double *azimuth);

/* InitGasDensity (src/Pframeforce.c:297-341), standard power law with the
 * optional AddNoise draw (drand48 seeded with p->noise_seed). */
int dfargo_init_gas_density(const dfargo_params *p, double *rho);

/* InitGasEnergy (src/Pframeforce.c:422-440). */
int dfargo_init_gas_energy(const dfargo_params *p, double *energy);

/* InitPlanetarySystem (src/Psys.c:96-196) after the .cfg file has been
 * parsed into per-planet columns. */
int dfargo_init_planetary_system(const dfargo_params *p, int nb,
                                 const double *dist, const double *mass,
                                 const double *acc, const int *feel_disk,
                                 const int *feel_others, dfargo_planets *out);

/* AspectRatio (src/Viscosity.c:107-122) at PhysicalTime = initial time. */
double dfargo_aspect_ratio(const dfargo_params *p, double r);

/* ---- GPU context ----------------------------------------------------------- */

/* Allocates the device state for one disc and uploads the 1-D tables:
 * the device half of FillPolar1DArrays plus the allocations of InitEuler
 * (src/SourceEuler.c:190-291), InitTransport (src/TransportEuler.c:32-59)
 * and InitViscosity (src/Viscosity.c:141-157).  precision: enum
 * dfargo_precision.  device: HIP device ordinal. */
int dfargo_create(const dfargo_params *p, int precision, int device,
                  dfargo_ctx **out);
int dfargo_destroy(dfargo_ctx *ctx);
const char *dfargo_last_error(void);

/* InitEuler (src/SourceEuler.c:190-291) together with the start-up
 * self-gravity call of src/main.c:178-186: computes sound speed, pressure,
 * temperature, the initial velocities (InitGasVelocities,
 * src/Pframeforce.c:442-641, written into vrad/vtheta) and the viscous
 * stresses, and makes rho, vrad, vtheta, energy and sys the device state. */
int dfargo_init_euler(dfargo_ctx *ctx, const double *rho, double *vrad,
                      double *vtheta, const double *energy,
                      const dfargo_planets *sys);

/* AlgoGas (src/SourceEuler.c:326-719): advances the state by DT with as
 * many CFL sub-steps as the reference takes.  The host arrays are uploaded
 * before and written back after the call, as the reference host expects
 * (src/main.c:387).  energy may be NULL when the energy equation is off;
 * sys may be NULL when there are no planets. */
int dfargo_algogas(dfargo_ctx *ctx, double DT, double *rho, double *vrad,
                   double *vtheta, double *energy, dfargo_planets *sys);

/* AlgoGas on the device-resident state, with no host transfer.  Used by a
 * host that keeps the state in HBM between calls. */
int dfargo_algogas_device(dfargo_ctx *ctx, double DT);

/* Measurement entry point (no counterpart in the reference): starts an
 * AlgoGas call of length DT on the device state and runs exactly nsub
 * sub-step passes, fewer only if DT is covered first; the call is left open
 * (dtemp < DT).  Used to time a fixed number of hydro sub-steps. */
int dfargo_run_substeps(dfargo_ctx *ctx, double DT, int nsub);

/* Explicit transfers of the state (rho, vrad, vtheta, energy: (nrad+1)*nsec
 * doubles each, energy may be NULL) and of the planetary system. */
int dfargo_upload_state(dfargo_ctx *ctx, const double *rho, const double *vrad,
                        const double *vtheta, const double *energy,
                        const dfargo_planets *sys);
int dfargo_download_state(dfargo_ctx *ctx, double *rho, double *vrad,
                          double *vtheta, double *energy, dfargo_planets *sys);

/* Scalar globals after the last call; set_time overrides PhysicalTime,
 * PhysicalTimeInitial and OmegaFrame (restart path, src/main.c:262-272). */
int dfargo_get_status(dfargo_ctx *ctx, dfargo_status *st);
int dfargo_set_time(dfargo_ctx *ctx, double physical_time,
                    double physical_time_initial, double omega_frame);

/* Copies a derived grid (enum dfargo_field) to out ((nrad+1)*nsec doubles).
 * Temperature is computed on demand (ComputeTemperatureField,
 * src/SourceEuler.c:2436-2460), the others hold their values from the last
 * sub-step. */
int dfargo_get_field(dfargo_ctx *ctx, int field, double *out);

/* compute_selfgravity(..., SGUpdate = NO) (src/sgmain.c:3-35) on the
 * device-resident density: refreshes gr/gtheta without a velocity kick. */
int dfargo_compute_selfgravity(dfargo_ctx *ctx);

/* FViscosity (src/Viscosity.c:23-69) with the current ring-mean sound
 * speed. */
double dfargo_fviscosity(dfargo_ctx *ctx, double r);

/* The integer FARGO shifts Nshift[i] (src/TransportEuler.c:310-316) and the
 * ring-mean azimuthal velocities VMed[i] of the last sub-step (nrad each). */
int dfargo_get_fargo_shift(dfargo_ctx *ctx, long *nshift, double *vmed);

/* DAMPTOVISCOUS's 1-D viscous disc (src/SideEuler.c:1444-1565): its cell
 * centres Rmed1D, Sigma1D and Vrad1D (n1d values each, NRAD1D) as the last
 * sub-step left them, the state Write1DViscProfiles writes
 * (src/SideEuler.c:1567-1585); any pointer may be NULL.  *n1d is 0 and
 * nothing is written when the option is off. */
int dfargo_get_visc1d(dfargo_ctx *ctx, double *rmed, double *sigma, double *vrad, int *n1d);

/* Sets the 1-D disc's Rmed1D, Sigma1D and/or Vrad1D (n1d must be NRAD1D;
 * a NULL pointer keeps that profile), as the restart branch of
 * InitializeOneDViscousEvolution (src/SideEuler.c:1489-1502) replaces
 * Rmed1D and Sigma1D with the values of gasdens.ascii_rad.<N>.dat.  Call it
 * after dfargo_init_euler, which resets the profiles to their t = 0 values
 * (dfargo_upload_state leaves them).  DFARGO_ERR_ARG when the option is off, on a
 * length mismatch, a non-finite value, Rmed1D not increasing, or (with the
 * wave-killing zones) Rmed1D not enclosing the grid's Rmed. */
int dfargo_set_visc1d(dfargo_ctx *ctx, const double *rmed, const double *sigma, const double *vrad, int n1d);

/* ---- per-DT diagnostics (src/main.c:291-394, between AlgoGas calls) -------- */

/* The MonitorIntegral budgets of the device state, in FP64 on the device:
 * out[0] GasTotalMass (src/SideEuler.c:15-36), out[1] GasMomentum
 * (src/SideEuler.c:114-146), out[2] GasTotalEnergy (src/SideEuler.c:149-180,
 * with the potential of the last sub-step; an isothermal run's energy grid
 * is zero, as CreatePolarGrid leaves it), out[3] the planets' angular
 * momentum of CheckMomentumConservation (src/SideEuler.c:182-210). */
int dfargo_gas_budgets(dfargo_ctx *ctx, double *out);

/* UpdateLog (src/Force.c:266-328): the disc force on each planet
 * (ComputeForce, src/Force.c:77-253, FP64 on the device), turned into the
 * columns of tqwk<k>.dat: torques[8k .. 8k+7] = x fy - y fx of the inner,
 * outer, Hill-excluded inner and outer discs, then vx fx + vy fy of the same
 * four; indirect[k] = the indirect torque of indtq<k>.dat from the last
 * sub-step's DiskOnPrimaryAcceleration.  *nb receives the planet count
 * (DFARGO_ERR_ARG if it exceeds max_planets). */
int dfargo_update_log(dfargo_ctx *ctx, int max_planets, double *torques, double *indirect, int *nb);

/* dfargo_update_log without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (under HIP
 * graph capture, in tight serving loops).  The same kernels and results as
 * dfargo_update_log, with the ring partials and sums in a static device
 * array of the library (one per device, 2.3 MB, kept until the process
 * ends): NRAD must be <= 8192 (DFARGO_ERR_ARG otherwise).  Calls that
 * overlap on one device share that array and race. */
int dfargo_update_log_noalloc(dfargo_ctx *ctx, int max_planets, double *torques, double *indirect, int *nb);

/* FindOrbitalElements (src/Planet.c:129-180) for a body at (x, y) moving
 * with (vx, vy) about a total mass m: out = e, a, M, V, PerihelionPA,
 * lambda, varpi (the columns of orbit<n>.dat after the time). */
int dfargo_orbital_elements(double x, double y, double vx, double vy, double m, double *out);

/* The run-level keys of a .par file that dfargo_params does not carry
 * (src/var.c:15-24). */
typedef struct dfargo_run_params {
  double dt;                  /* DT: the interval of one AlgoGas call     */
  int    ntot, ninterm;       /* NTOT, NINTERM                            */
  char   planet_config[512];  /* PLANETCONFIG                             */
  char   output_dir[512];     /* OUTPUTDIR                                */
} dfargo_run_params;

/* ReadVariables (src/Interpret.c:82-160) and the decoding after it
 * (src/Interpret.c:165-476) for the keys the hydro step reads, with the
 * defaults of src/var.c; then dfargo_check_params.  A missing mandatory key
 * is DFARGO_ERR_ARG; a file that switches on physics not implemented here
 * (thermal cooling, diffusion, the dust fluid, ...) is
 * DFARGO_ERR_UNSUPPORTED.  run may be NULL.  The particle keys are not read:
 * dfargo_dust_params describes the particle set.  Replaces the host's
 * ReadVariables for callers without the reference's host. */
int dfargo_read_par(const char *path, dfargo_params *p, dfargo_run_params *run);

/* InitPlanetarySystem (src/Psys.c:96-196) from a planet .cfg file: lines
 * starting with a letter are "name distance mass accretion FeelDisk
 * FeelOthers [Binary]", read with %f as the reference reads them. */
int dfargo_read_planet_config(const char *path, const dfargo_params *p, dfargo_planets *sys);

/* ---- launch shapes and batched discs -------------------------------------- */

/* Launch shapes of the sub-step kernels that do not change results: the
 * ring kernels' XCD-contiguous block order, the azimuthal transport's
 * wavefronts per workgroup, the velocity/boundary kernel's workgroup
 * width, the LDS buffers of the self-gravity transforms and their rings
 * per workgroup.  No
 * counterpart in the reference (its loops have no launch shape). */
typedef struct dfargo_tuning {
  int xcd_remap;     /* 1: the rings of one XCD are contiguous; 0: round-robin */
  int azi_waves;     /* 1, 2 or 4 wavefronts per workgroup                    */
  int finish_block;  /* 64, 128 or 256 threads per workgroup                  */
  int sg_fft_buffers; /* LDS buffers per transform: 2 (ping-pong), 1 (in
                         place, half the LDS, one more barrier per pass), or
                         0: one where that lets more workgroups share a CU   */
  int sg_rows_pair;   /* rings per workgroup of the self-gravity row transforms:
                         1, 2, or 0: two where the rows run in place         */
} dfargo_tuning;
int dfargo_get_tuning(dfargo_ctx *ctx, dfargo_tuning *t);
int dfargo_set_tuning(dfargo_ctx *ctx, const dfargo_tuning *t);
/* Times each knob's values on a scratch disc of the same parameters
 * (`substeps` profiled sub-steps per trial; this disc's state is not
 * touched), applies the fastest and keeps it for every later disc of the
 * same shape, precision, switches and device in this process. */
int dfargo_autotune(dfargo_ctx *ctx, int substeps, dfargo_tuning *best);

/* Many discs advanced together, each with its own device-side CFL step
 * and loop condition (src/SourceEuler.c:398-453 per disc), every disc's
 * sub-step a captured graph (dispatch: dfargo_batch_set_mode).  dfargo_batch_algogas is one AlgoGas call of length DT[i] for
 * disc i, state left in HBM (as dfargo_algogas_device).  The discs must be
 * initialised, on one device, and keep the sub-step they had at creation
 * (attaching particles or switching accretion needs a new batch;
 * DFARGO_ERR_STATE otherwise).  No counterpart in the reference. */
typedef struct dfargo_batch dfargo_batch;
int dfargo_batch_create(dfargo_ctx *const *discs, int n, dfargo_batch **out);
int dfargo_batch_algogas(dfargo_batch *b, const double *DT);
int dfargo_batch_destroy(dfargo_batch *b);
/* How a batch dispatches: DFARGO_BATCH_GRAPH, one graph holding every
 * disc's sub-step launched once per sub-step for all discs (a finished
 * disc's kernels return at once); DFARGO_BATCH_STREAMS (the default), each
 * disc's sub-step graph on the disc's own stream, as many times as that
 * disc still needs, the discs interleaved across the hardware queues. */
enum dfargo_batch_mode { DFARGO_BATCH_GRAPH = 0, DFARGO_BATCH_STREAMS = 1 };
int dfargo_batch_set_mode(dfargo_batch *b, int mode);

/* ---- surrogate warm start: the MLP (no counterpart in the reference) ------- */

/* A small MLP trained and evaluated on the GPU with the library's own BF16
 * matrix kernels (FP32 accumulation, FP32 master weights, Adam): n_in
 * inputs (<= 32) -> n_hidden layers of `width` GELU units -> n_out linear
 * outputs (width, n_out multiples of 32).  Inputs and targets are host
 * FP32 arrays, row-major, `batch` rows.  Parameters in the flat order
 * W0, b0, W1, b1, ... with W_l row-major (inputs x outputs).  The loss is
 * the mean squared error over the batch's outputs. */
typedef struct dfargo_mlp dfargo_mlp;
int dfargo_mlp_create(int n_in, int width, int n_hidden, int n_out, int max_batch, unsigned long seed, int device,
                      dfargo_mlp **out);
int dfargo_mlp_destroy(dfargo_mlp *m);
size_t dfargo_mlp_param_count(const dfargo_mlp *m);
int dfargo_mlp_get_params(dfargo_mlp *m, float *out);
int dfargo_mlp_set_params(dfargo_mlp *m, const float *in);
int dfargo_mlp_forward(dfargo_mlp *m, int batch, const float *x, float *y);
/* the loss and its gradient with respect to every parameter (flat order) */
int dfargo_mlp_gradients(dfargo_mlp *m, int batch, const float *x, const float *t, float *loss, float *grads);
/* `steps` full-batch Adam steps (lr; beta 0.9, 0.999) on the device; the
 * loss after them if loss is not NULL */
int dfargo_mlp_train(dfargo_mlp *m, int batch, const float *x, const float *t, int steps, float lr, float *loss);

/* Per-kernel timing with HIP events on the context's stream.  While
 * enabled, sub-steps launch eagerly and every kernel is bracketed by a pair
 * of events.  names: up to max entries of at most 47 characters. */
int dfargo_set_profiling(dfargo_ctx *ctx, int enable);
int dfargo_get_kernel_times(dfargo_ctx *ctx, int max, char (*names)[48],
                            double *total_ms, long *launches, int *count);

/* The hipStream_t the context launches on, as an opaque pointer. */
void *dfargo_stream(dfargo_ctx *ctx);

/* ---- dust particles ------------------------------------------------------- */

/* Defaults of src/var.c:136-181 (NODTCONSTRAINTBYPCS YES, DUSTFEEDBACK NO,
 * INTERPOLATION TSC, SFTAPPROX YES) with nbpart particles. */
void dfargo_dust_default_params(dfargo_dust_params *d, int nbpart);
/* Checks the particle switches; DFARGO_ERR_UNSUPPORTED for the options the
 * GPU path does not implement (turbulent kicks, growth, Hill-sphere
 * removal, the ZZ integrator, the axisymmetric-only SG options). */
int dfargo_check_dust_params(const dfargo_params *p, const dfargo_dust_params *d);
/* InitDustSystem (src/Dsys.c:128-216): sizes, radii and azimuths drawn
 * with drand48 seeded by d->seed (all sizes, then all radii, then all
 * azimuths), vr = 0, vth Keplerian.  With self-gravity the axisymmetric
 * radial SG acceleration (nrad values, GLOBAL_AxiSGAccr) corrects vth;
 * pass NULL without self-gravity. */
int dfargo_init_dust_system(const dfargo_params *p, const dfargo_dust_params *d,
                            const double *axi_sg_accr, dfargo_dust_state *out);
/* Attaches particles to a context (after dfargo_init_euler) and runs the
 * start-up `interpolation` (src/main.c:287, with the output time step DT)
 * that sets the gas values at the particles, the feedback grids and the
 * minimum stopping time.  From
 * then on every sub-step of AlgoGas moves the particles as
 * src/SourceEuler.c:504-531 does. */
int dfargo_attach_dust(dfargo_ctx *ctx, const dfargo_dust_params *d, const dfargo_dust_state *st,
                       double timestep);
/* Particle arrays back to the host / from the host (nbpart entries). */
int dfargo_download_dust(dfargo_ctx *ctx, dfargo_dust_state *st);
int dfargo_upload_dust(dfargo_ctx *ctx, const dfargo_dust_state *st);

/* ---- the dust fluid --------------------------------------------------------- */

/* Defaults of src/var.c:136-186 for the dust fluid. */
void dfargo_dust_fluid_default_params(dfargo_dust_fluid_params *f);
/* Checks the dust-fluid switches against the gas parameters;
 * DFARGO_ERR_UNSUPPORTED for what the GPU path does not implement: FP32
 * storage, dust particles beside the fluid, SFTAPPROX YES with DUSTFEELDISK
 * NO (the reference then leaves DVradInt unset). */
int dfargo_check_dust_fluid_params(const dfargo_params *p, const dfargo_dust_fluid_params *f, int precision);
/* InitDustDensity (src/Pframeforce.c:344-380) without its noise options:
 * DSigma(Rmed[i]) = Sigma(Rmed[i]) * DUSTTOGASDENSITYRATIO on every cell
 * ((nrad + 1) * nsec values, ring nrad zero). */
int dfargo_init_dust_fluid_density(const dfargo_params *p, const dfargo_dust_fluid_params *f, double *drho);
/* Switches the dust fluid on for a context (before dfargo_init_euler) with
 * its initial surface density ((nrad + 1) * nsec values, DRho of
 * src/main.c:167-168).  dfargo_init_euler then runs InitDustStokesNumber
 * and InitDustVelocities (src/SourceEuler.c:265-269) and, with
 * DUSTFEEDBACK, the gas velocities of src/Pframeforce.c:609-634; every
 * AlgoGas call advances the fluid with the gas. */
int dfargo_attach_dust_fluid(dfargo_ctx *ctx, const dfargo_dust_fluid_params *f, const double *drho);
/* The dust-fluid switches of a .par file (src/Interpret.c:253-342, with
 * src/var.c's defaults): *dust_fluid takes DUSTFLUID's value.
 * DFARGO_ERR_UNSUPPORTED for RESTARTWITHNEWDUST YES.  dfargo_read_par
 * reads the gas's part of the same file. */
int dfargo_read_dust_fluid_par(const char *path, dfargo_dust_fluid_params *f, int *dust_fluid);
/* The fluid's fields DRho, DVrad, DVtheta and the Stokes grid of the last
 * sub-step ((nrad + 1) * nsec values each; any pointer may be NULL) to the
 * host, and DRho, DVrad, DVtheta from the host. */
int dfargo_download_dust_fluid(dfargo_ctx *ctx, double *drho, double *dvrad, double *dvtheta, double *stokes);
int dfargo_upload_dust_fluid(dfargo_ctx *ctx, const double *drho, const double *dvrad, const double *dvtheta);

#ifdef __cplusplus
}
#endif
#endif /* DFARGO_H */
