/*
 * fargo_shim.c -- the drop-in: Dusty FARGO-ADSG's hydro-step symbols,
 * routed to libdfargo (include/dfargo.h).
 *
 * Link it into the reference in place of the hot-path objects
 * (SourceEuler.o TransportEuler.o Viscosity.o Pframeforce.o sgmain.o
 * sginit.o sgdens.o sgkernel.o sgacc.o sgupdate.o sgsysinit.o Dsys.o
 * DustUpdate.o, src/makefile:54-60).  The reference's host objects then
 * import exactly the 26 symbols defined here: the functions AlgoGas,
 * InitEuler, FillPolar1DArrays, ComputeSoundSpeed, ComputePressureField,
 * ComputeTemperatureField, ComputeViscousTerms, ComputeViscousHeating,
 * ComputeEntropyDiffusion, ComputeRadiativeDiffusion, CircumPlanetaryMass,
 * FViscosity, AspectRatio, DAspectRatio, InitGasDensity, InitGasEnergy,
 * InitDustDensity, InitDustSystem, FreeDust, SolveOrbits,
 * compute_selfgravity, init_planetarysys_withSG, interpolation, and the
 * data DiskOnPrimaryAcceleration, DiskOnPrimaryAcceleration_alt, LostMass.
 *
 * Conventions kept from the reference: void returns, fatal errors end in
 * prs_exit(1) with the library's message, one dot per hydro sub-step on
 * stdout.  The state lives on the GPU between calls; AlgoGas leaves the
 * host grids and globals coherent on return (src/main.c:387).
 * DFARGO_PRECISION=fp32 in the environment selects FP32 field storage.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "dfargo.h"
#include "fargo_ref.h"

void FillSigma(void);  /* src/Theo.c:87-93 */
void FillEnergy(void); /* src/Theo.c:172-206 */
void FillDSigma(void); /* src/Theo.c:95-101 */

/* the functions this file defines, with the parameter types of their
 * definitions in the removed objects (src/proto.h declares them without
 * prototypes) */
void AlgoGas(Force *, PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *,
             PolarGrid *, PolarGrid *, PlanetarySystem *, DustSystem *);
void InitEuler(PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *,
               PlanetarySystem *);
void FillPolar1DArrays(void);
void ComputeSoundSpeed(PolarGrid *, PolarGrid *, PlanetarySystem *);
void ComputePressureField(PolarGrid *, PolarGrid *);
void ComputeTemperatureField(PolarGrid *, PolarGrid *);
void ComputeViscousTerms(PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *);
void ComputeViscousHeating(PolarGrid *);
void ComputeEntropyDiffusion(PolarGrid *, PolarGrid *);
void ComputeRadiativeDiffusion(PolarGrid *, PolarGrid *);
real CircumPlanetaryMass(PolarGrid *, PlanetarySystem *);
real FViscosity(real);
real AspectRatio(real);
real DAspectRatio(real);
void InitGasDensity(PolarGrid *);
void InitGasEnergy(PolarGrid *);
void InitDustDensity(PolarGrid *);
DustSystem *InitDustSystem(void);
void FreeDust(DustSystem *);
void SolveOrbits(PlanetarySystem *);
void compute_selfgravity(PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *, real, boolean);
void init_planetarysys_withSG(PlanetarySystem *);
void interpolation(DustSystem *, PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *, real);

/* the data the removed objects defined (src/SourceEuler.c:38,
 * src/TransportEuler.c:28) */
Pair DiskOnPrimaryAcceleration, DiskOnPrimaryAcceleration_alt;
real LostMass = 0.0;

static dfargo_ctx *ctx = NULL;
static dfargo_dust_params dpar;
static int dust_attached = 0;
static int fluid_attached = 0; /* the dust fluid is on the context */

static void die(int rc) {
  if (rc) {
    mastererr("libdfargo: %s\n", dfargo_last_error());
    prs_exit(1);
  }
}

static void unsupported(const char *what) {
  mastererr("libdfargo: %s is not supported by the GPU hydro step\n", what);
  prs_exit(1);
}

/* the decoded .par switches (src/Interpret.c:165-476) as a parameter block */
static void params_from_globals(dfargo_params *p) {
  dfargo_default_params(p);
  p->nrad = NRAD; p->nsec = NSEC; p->rmin = RMIN; p->rmax = RMAX;
  p->log_grid = LogGrid; p->pmin = PMIN; p->pmax = PMAX;
  p->sigma0 = SIGMA0; p->sigma_slope = SIGMASLOPE;
  p->aspect_ratio = ASPECTRATIO; p->flaring_index = FLARINGINDEX;
  p->viscosity = VISCOSITY; p->alpha_viscosity = ALPHAVISCOSITY;
  p->cavity_radius = CAVITYRADIUS; p->cavity_ratio = CAVITYRATIO; p->cavity_width = CAVITYWIDTH;
  p->transition_radius = TRANSITIONRADIUS; p->transition_ratio = TRANSITIONRATIO;
  p->transition_width = TRANSITIONWIDTH; p->imposed_disk_drift = IMPOSEDDISKDRIFT;
  p->add_noise = AddNoise;
  p->noise_seed = (long)time(NULL); /* the second src/Interpret.c:276 seeds drand48 with */
  p->energy_equation = EnergyEquation; p->viscous_heating = ViscousHeating;
  p->beta_cooling = BetaCooling; p->adiabatic_index = ADIABATICINDEX;
  p->beta_cooling_time = BETACOOLINGTIME; p->beta_cooling_slope = BETACOOLINGSLOPE;
  p->self_gravity = SelfGravity; p->sg_zero_mode = SGZeroMode;
  p->sg_thickness_smoothing = SGTHICKNESSSMOOTHING;
  p->thickness_smoothing = THICKNESSSMOOTHING; p->roche_smoothing = ROCHESMOOTHING;
  p->indirect_term_smoothing = INDIRECTTERMSMOOTHING; p->mass_taper = MASSTAPER;
  p->indirect_term = Indirect_Term; p->discard_gas_indirect_term = Discard_GasIndirect_term;
  p->alternative_gas_indirect_term = Alternative_Gas_Indirect_Term;
  p->exclude_hill = ExcludeHill; p->exclude_hill_factor = EXCLUDEHILLFACTOR;
  p->forced_circular = ForcedCircular; p->release_date = RELEASEDATE;
  p->release_radius = RELEASERADIUS; p->eccentricity = ECCENTRICITY;
  p->corotating = !Corotating ? DFARGO_FRAME_FIXED
                  : GuidingCenter ? DFARGO_FRAME_GUIDING_CENTER
                  : BinaryCenter ? DFARGO_FRAME_BINARY_CENTER : DFARGO_FRAME_COROTATING;
  p->corotate_with_outer_planet = CorotateWithOuterPlanet;
  p->omega_frame = OmegaFrame;
  p->fast_transport = FastTransport; p->add_floors = AddFloors; p->density_floor = floordens;
  p->boundary = OpenInner ? DFARGO_BC_OPEN : NonReflecting ? DFARGO_BC_NONREFLECTING : DFARGO_BC_WALL;
  p->wave_killing_zones = Evanescent; p->wkz_rmin = WKZRMIN; p->wkz_rmax = WKZRMAX;
  p->damping_ratio = DAMPINGTODYNAMICALTIMERATIO;
  p->damp_to = DampToViscous ? DFARGO_DAMP_TO_VISCOUS : DampToIni ? DFARGO_DAMP_TO_INI : DFARGO_DAMP_TO_AXI;
  p->zero_vrad_bc = ImposeZeroRadialVelocityBC;
  p->nrad1d = NRAD1D; p->rmin1d = RMIN1D; p->rmax1d = RMAX1D;
  p->bc1d = BC1D_SS_ZeroVel ? 1 : BC1D_SS_NonZeroVel ? 2 : 0;
}

static void ensure_context(void) {
  dfargo_params p;
  const char *prec;
  if (ctx) return;
  if (LAMBDADOUBLING != 0.0) unsupported("LAMBDADOUBLING");
  params_from_globals(&p);
  prec = getenv("DFARGO_PRECISION");
  die(dfargo_create(&p, (prec && (*prec == 'f' || *prec == 'F') && strstr(prec, "32")) ? DFARGO_FP32 : DFARGO_FP64,
                    0, &ctx));
}

static void planets_to_abi(const PlanetarySystem *s, dfargo_planets *d) {
  int k;
  memset(d, 0, sizeof *d);
  if (!s) return;
  if (s->nb > DFARGO_MAX_PLANETS) unsupported("more than DFARGO_MAX_PLANETS planets");
  d->nb = s->nb;
  for (k = 0; k < s->nb; k++) {
    d->mass[k] = s->mass[k]; d->x[k] = s->x[k]; d->y[k] = s->y[k];
    d->vx[k] = s->vx[k]; d->vy[k] = s->vy[k]; d->acc[k] = s->acc[k];
    d->feel_disk[k] = s->FeelDisk[k]; d->feel_others[k] = s->FeelOthers[k];
    d->binary[k] = s->Binary[k];
    d->initial_mass[k] = InitialPlanetMass[k]; d->final_mass[k] = FinalPlanetMass[k];
  }
}

static void planets_from_abi(const dfargo_planets *d, PlanetarySystem *s) {
  int k;
  if (!s) return;
  for (k = 0; k < s->nb; k++) {
    s->mass[k] = d->mass[k]; s->x[k] = d->x[k]; s->y[k] = d->y[k];
    s->vx[k] = d->vx[k]; s->vy[k] = d->vy[k];
  }
}

static PolarGrid *grid_once(PolarGrid **g, char *name) {
  if (!*g) *g = CreatePolarGrid(NRAD, NSEC, name);
  return *g;
}

/* ---- geometry ------------------------------------------------------------ */

/* FillPolar1DArrays (src/SourceEuler.c:71-188) for one rank: the radii
 * (log or arithmetic; a radii.dat file is refused, the device computes
 * its own), the ring and azimuth tables, and the used_rad.dat,
 * used_azi.dat and minmaxradii.dat files. */
void FillPolar1DArrays(void) {
  char name[1100];
  FILE *f;
  int i, j;
  const real dr = (RMAX - RMIN) / (real)GLOBALNRAD, dphi = (PMAX - PMIN) / (real)NSEC;
  snprintf(name, sizeof name, "%s%s", OUTPUTDIR, "radii.dat");
  if ((f = fopen(name, "r")) != NULL) {
    fclose(f);
    unsupported("a radii.dat grid");
  }
  for (i = 0; i <= GLOBALNRAD; i++)
    Radii[i] = LogGrid ? RMIN * exp((real)i / (real)GLOBALNRAD * log(RMAX / RMIN)) : RMIN + dr * (real)i;
  for (i = 0; i < GLOBALNRAD; i++)
    GlobalRmed[i] = 2.0 / 3.0 * (Radii[i + 1] * Radii[i + 1] * Radii[i + 1] - Radii[i] * Radii[i] * Radii[i]) /
                    (Radii[i + 1] * Radii[i + 1] - Radii[i] * Radii[i]);
  for (i = 0; i < NRAD; i++) {
    Rinf[i] = Radii[i + IMIN];
    Rsup[i] = Radii[i + IMIN + 1];
    Rmed[i] = 2.0 / 3.0 * (Rsup[i] * Rsup[i] * Rsup[i] - Rinf[i] * Rinf[i] * Rinf[i]);
    Rmed[i] = Rmed[i] / (Rsup[i] * Rsup[i] - Rinf[i] * Rinf[i]);
    Surf[i] = 0.5 * (PMAX - PMIN) * (Rsup[i] * Rsup[i] - Rinf[i] * Rinf[i]) / (real)NSEC;
    InvRmed[i] = 1.0 / Rmed[i];
    InvSurf[i] = 1.0 / Surf[i];
    InvDiffRsup[i] = 1.0 / (Rsup[i] - Rinf[i]);
    InvRinf[i] = 1.0 / Rinf[i];
  }
  Rinf[NRAD] = Radii[NRAD + IMIN];
  for (i = 1; i < NRAD; i++) InvDiffRmed[i] = 1.0 / (Rmed[i] - Rmed[i - 1]);
  for (j = 0; j < NSEC; j++) {
    Azimuth[j] = PMIN + (PMAX - PMIN) * (real)j / (real)NSEC;
    CosAzimuth[j] = cos(Azimuth[j]);
    SinAzimuth[j] = sin(Azimuth[j]);
    AziInf[j] = Azimuth[j] - 0.5 * dphi;
    AziSup[j] = Azimuth[j] + 0.5 * dphi;
  }
  if (CPU_Rank != 0) return;
  snprintf(name, sizeof name, "%sminmaxradii.dat", OUTPUTDIR);
  if ((f = fopen(name, "w")) == NULL) unsupported("writing minmaxradii.dat");
  fprintf(f, "%d\t%.18g\t%.18g\n", CPU_Rank, Rinf[Zero_or_active], Rsup[Max_or_active - 1]);
  fclose(f);
  snprintf(name, sizeof name, "%s%s", OUTPUTDIR, Restart ? "newused_rad.dat" : "used_rad.dat");
  if ((f = fopen(name, "w")) == NULL) unsupported("writing used_rad.dat");
  for (i = 0; i <= GLOBALNRAD; i++) fprintf(f, "%.18g\n", Radii[i]);
  fclose(f);
  snprintf(name, sizeof name, "%sused_azi.dat", OUTPUTDIR);
  if ((f = fopen(name, "w")) == NULL) unsupported("writing used_azi.dat");
  for (j = 0; j < NSEC; j++) fprintf(f, "%.18g\t%.18g\t%.18g\n", Azimuth[j], AziInf[j], AziSup[j]);
  fclose(f);
}

/* ---- initial state ------------------------------------------------------- */

void InitGasDensity(PolarGrid *Rho) {
  dfargo_params p;
  FillSigma(); /* SigmaMed / SigmaInf for the host's own readers */
  params_from_globals(&p);
  die(dfargo_init_gas_density(&p, Rho->Field));
}

void InitGasEnergy(PolarGrid *energ) {
  dfargo_params p;
  FillEnergy();
  params_from_globals(&p);
  die(dfargo_init_gas_energy(&p, energ->Field));
}

/* the dust-fluid switches (src/Interpret.c:253-342) as a parameter block */
static void dust_fluid_params_from_globals(dfargo_dust_fluid_params *f) {
  dfargo_dust_fluid_default_params(f);
  f->sizepart = Sizepart; f->rhopart = RHOPART;
  f->factor_unit_mass = FACTORUNITMASS; f->factor_unit_length = FACTORUNITLENGTH;
  f->dust_to_gas_density_ratio = DUSTTOGASDENSITYRATIO;
  f->dviscosity = DVISCOSITY; f->dalpha_viscosity = DALPHAVISCOSITY;
  f->daspect_ratio = DASPECTRATIO; f->dflaring_index = DFLARINGINDEX; f->dust_diff_inn = DUSTDIFFINN;
  f->sfta = ShortFrictionTimeApproximation; f->dust_feel_disk = DustFeelDisk;
  f->dust_feedback = DustFeedback;
  f->dust_feel_sg = !DustFeelSG ? DFARGO_DUST_SG_NO : DustFeelSGZeroMode ? DFARGO_DUST_SG_ZERO_MODE
                                                                          : DFARGO_DUST_SG_YES;
  f->dust_diffusion = DustDiffusion; f->open_inner_dust = OpenInnerDust;
}

/* InitDustDensity (src/Pframeforce.c:344-380) without its noise options */
void InitDustDensity(PolarGrid *DRho) {
  dfargo_params p;
  dfargo_dust_fluid_params f;
  if (AddNoise) unsupported("ADDNOISE with DUSTFLUID (the dust's draws continue the gas's drand48 stream)");
  FillDSigma();
  params_from_globals(&p);
  dust_fluid_params_from_globals(&f);
  die(dfargo_init_dust_fluid_density(&p, &f, DRho->Field));
}

/* InitEuler (src/SourceEuler.c:190-291): the device context, the host
 * grids the host reads back, and the start-up state (sound speed,
 * pressure, temperature, equilibrium velocities, stresses). */
void InitEuler(PolarGrid *Vr, PolarGrid *Vt, PolarGrid *Rho, PolarGrid *Energy, PolarGrid *DVr, PolarGrid *DVt,
               PolarGrid *DRho, PlanetarySystem *sys) {
  dfargo_planets pl;
  ensure_context();
  grid_once(&Potential, "Potential");
  grid_once(&Pressure, "Pressure");
  grid_once(&SoundSpeed, "SoundSpeed");
  grid_once(&Temperature, "Temperature");
  grid_once(&ViscHeat, "ViscousHeating");
  grid_once(&Test, "Test");
  grid_once(&RadFBAcc, "RadFBAcc");
  grid_once(&AziFBAcc, "AziFBAcc");
  grid_once(&FBedot, "FBedot");
  grid_once(&DivergenceVelocity, "DivV");
  grid_once(&TAURR, "TAUrr");
  grid_once(&TAURP, "TAUrp");
  grid_once(&TAUPP, "TAUpp");
  InitComputeAccel();
  die(dfargo_set_time(ctx, PhysicalTime, PhysicalTimeInitial, OmegaFrame));
  planets_to_abi(sys, &pl);
  /* the dust fluid starts from the host's DRho (src/SourceEuler.c:265-269) */
  if (DustFluid) {
    dfargo_dust_fluid_params f;
    dust_fluid_params_from_globals(&f);
    if (!fluid_attached) die(dfargo_attach_dust_fluid(ctx, &f, DRho->Field));
    else die(dfargo_upload_dust_fluid(ctx, DRho->Field, NULL, NULL));
    fluid_attached = 1;
  }
  die(dfargo_init_euler(ctx, Rho->Field, Vr->Field, Vt->Field, EnergyEquation ? Energy->Field : NULL, &pl));
  if (DustFluid) die(dfargo_download_dust_fluid(ctx, DRho->Field, DVr->Field, DVt->Field, NULL));
  /* the planets as the device holds them (with self-gravity their
     start-up velocities are corrected, src/sgsysinit.c:3-49) */
  die(dfargo_download_state(ctx, NULL, NULL, NULL, NULL, &pl));
  planets_from_abi(&pl, sys);
  ComputeSoundSpeed(Rho, Energy, sys);
  ComputePressureField(Rho, Energy);
  ComputeTemperatureField(Rho, Energy);
  ComputeViscousTerms(Vr, Vt, Rho, DVr, DVt, DRho);
  if (EnergyEquation && ViscousHeating) ComputeViscousHeating(Rho);
}

/* ---- the hydro step -------------------------------------------------------- */

void AlgoGas(Force *force, PolarGrid *Rho, PolarGrid *Vrad, PolarGrid *Vtheta, PolarGrid *Energy, PolarGrid *Label,
             PolarGrid *DRho, PolarGrid *dustpcdens, PolarGrid *DVrad, PolarGrid *DVtheta, PlanetarySystem *sys,
             DustSystem *dsys) {
  dfargo_planets pl;
  dfargo_status st;
  int k;
  (void)force; (void)Label;
  planets_to_abi(sys, &pl);
  die(dfargo_set_time(ctx, PhysicalTime, PhysicalTimeInitial, OmegaFrame));
  if (DustFluid) die(dfargo_upload_dust_fluid(ctx, DRho->Field, DVrad->Field, DVtheta->Field));
  die(dfargo_algogas(ctx, DT, Rho->Field, Vrad->Field, Vtheta->Field, EnergyEquation ? Energy->Field : NULL, &pl));
  if (DustFluid) die(dfargo_download_dust_fluid(ctx, DRho->Field, DVrad->Field, DVtheta->Field, NULL));
  planets_from_abi(&pl, sys);
  if (NBPART != 0 && dust_attached && dsys) {
    dfargo_dust_state ds;
    ds.r = dsys->r; ds.th = dsys->th; ds.vr = dsys->vr; ds.vth = dsys->vth; ds.l = dsys->l;
    ds.dustsize = dsys->dustsize; ds.dustsize_init = dsys->dustsize_init; ds.stokesnb = dsys->stokesnb;
    die(dfargo_download_dust(ctx, &ds));
    die(dfargo_get_field(ctx, DFARGO_FIELD_DUST_DENSITY, dustpcdens->Field));
  }
  die(dfargo_get_status(ctx, &st));
  PhysicalTime = st.physical_time;
  OmegaFrame = st.omega_frame;
  MINDT = st.mindt;
  MAXDT = st.maxdt;
  LostMass = st.lost_mass;
  MassTaper = st.mass_taper;
  DiskOnPrimaryAcceleration.x = st.disk_on_primary_x;
  DiskOnPrimaryAcceleration.y = st.disk_on_primary_y;
  DiskOnPrimaryAcceleration_alt.x = st.disk_on_primary_alt_x;
  DiskOnPrimaryAcceleration_alt.y = st.disk_on_primary_alt_y;
  for (k = 0; k < 3; k++) {
    stardiscbarycenter_x[k] = st.bary_x[k];
    stardiscbarycenter_y[k] = st.bary_y[k];
  }
  for (k = 0; k < st.last_call_substeps; k++) masterprint("."); /* src/SourceEuler.c:567 */
  if (st.crashed) masterprint("\nNegative energy after a hydro sub-step (DetectCrash)\n");
}

/* ---- derived grids --------------------------------------------------------- */

/* ComputeSoundSpeed / ComputePressureField / ComputeTemperatureField
 * (src/SourceEuler.c:2323-2460) on the host arrays the host passes
 * (after a restart read they are not on the device yet) */
void ComputeSoundSpeed(PolarGrid *Rho, PolarGrid *Energy, PlanetarySystem *sys) {
  int i, j;
  real *cs = grid_once(&SoundSpeed, "SoundSpeed")->Field;
  (void)sys;
  for (i = 0; i < Rho->Nrad; i++) {
    const real iso = AspectRatio(Rmed[i]) * sqrt(1.0 / Rmed[i]) * pow(Rmed[i], FLARINGINDEX);
    for (j = 0; j < Rho->Nsec; j++) {
      const int l = i * Rho->Nsec + j;
      cs[l] = EnergyEquation ? sqrt(ADIABATICINDEX * (ADIABATICINDEX - 1.0) * Energy->Field[l] / Rho->Field[l]) : iso;
    }
  }
}

void ComputePressureField(PolarGrid *Rho, PolarGrid *Energy) {
  int l;
  const real *cs = grid_once(&SoundSpeed, "SoundSpeed")->Field;
  real *pr = grid_once(&Pressure, "Pressure")->Field;
  for (l = 0; l < Rho->Nrad * Rho->Nsec; l++)
    pr[l] = EnergyEquation ? (ADIABATICINDEX - 1.0) * Energy->Field[l] : Rho->Field[l] * cs[l] * cs[l];
}

void ComputeTemperatureField(PolarGrid *Rho, PolarGrid *Energy) {
  int l;
  const real *pr = grid_once(&Pressure, "Pressure")->Field;
  real *t = grid_once(&Temperature, "Temperature")->Field;
  for (l = 0; l < Rho->Nrad * Rho->Nsec; l++)
    t[l] = EnergyEquation ? (ADIABATICINDEX - 1.0) * Energy->Field[l] / Rho->Field[l] : pr[l] / Rho->Field[l];
}

/* the stress tensor and divergence of the device state (the host's arrays
 * equal it after every AlgoGas) */
void ComputeViscousTerms(PolarGrid *Vr, PolarGrid *Vt, PolarGrid *Rho, PolarGrid *DVr, PolarGrid *DVt,
                         PolarGrid *DRho) {
  (void)Vr; (void)Vt; (void)Rho; (void)DVr; (void)DVt; (void)DRho;
  ensure_context();
  die(dfargo_get_field(ctx, DFARGO_FIELD_TAURR, grid_once(&TAURR, "TAUrr")->Field));
  die(dfargo_get_field(ctx, DFARGO_FIELD_TAURP, grid_once(&TAURP, "TAUrp")->Field));
  die(dfargo_get_field(ctx, DFARGO_FIELD_TAUPP, grid_once(&TAUPP, "TAUpp")->Field));
  die(dfargo_get_field(ctx, DFARGO_FIELD_DIVV, grid_once(&DivergenceVelocity, "DivV")->Field));
}

void ComputeViscousHeating(PolarGrid *Rho) {
  (void)Rho;
  ensure_context();
  die(dfargo_get_field(ctx, DFARGO_FIELD_VISCHEAT, grid_once(&ViscHeat, "ViscousHeating")->Field));
}

void ComputeEntropyDiffusion(PolarGrid *Rho, PolarGrid *Energy) {
  (void)Rho; (void)Energy;
  unsupported("ENTROPYDIFFUSION");
}

void ComputeRadiativeDiffusion(PolarGrid *Rho, PolarGrid *Energy) {
  (void)Rho; (void)Energy;
  unsupported("RADIATIVEDIFFUSION");
}

/* ---- host-side functions of the removed objects -------------------------- */

/* the gas mass inside planet 0's Hill sphere (src/SourceEuler.c:2462-2504),
 * one rank */
real CircumPlanetaryMass(PolarGrid *Rho, PlanetarySystem *sys) {
  const int ns = Rho->Nsec;
  const real xpl = sys->x[0], ypl = sys->y[0], rpl = sqrt(xpl * xpl + ypl * ypl);
  const real rh = rpl * pow(sys->mass[0] / 3., 1. / 3.);
  real m = 0.0;
  int i, j;
  for (i = Zero_or_active; i < Max_or_active; i++)
    for (j = 0; j < ns; j++) {
      const int l = i * ns + j;
      const real dx = CellAbscissa->Field[l] - xpl, dy = CellOrdinate->Field[l] - ypl;
      if (sqrt(dx * dx + dy * dy) < rh) m += Surf[i] * Rho->Field[l];
    }
  return m;
}

real FViscosity(real r) {
  ensure_context();
  return dfargo_fviscosity(ctx, r);
}

real AspectRatio(real r) {
  dfargo_params p;
  if (LAMBDADOUBLING != 0.0) unsupported("LAMBDADOUBLING");
  params_from_globals(&p);
  return dfargo_aspect_ratio(&p, r);
}

/* the dust fluid's aspect ratio (src/Viscosity.c:124-139) */
real DAspectRatio(real r) {
  real h = DASPECTRATIO;
  const real scale = 1.0 + (PhysicalTime - PhysicalTimeInitial) * LAMBDADOUBLING;
  const real rmin = (TRANSITIONRADIUS - TRANSITIONWIDTH * DASPECTRATIO) * scale;
  const real rmax = (TRANSITIONRADIUS + TRANSITIONWIDTH * DASPECTRATIO) * scale;
  if (r < rmin) h *= TRANSITIONRATIO;
  if (r >= rmin && r <= rmax) h *= exp((rmax - r) / (rmax - rmin) * log(TRANSITIONRATIO));
  return h;
}

void SolveOrbits(PlanetarySystem *sys) {
  int k;
  for (k = 0; k < sys->nb; k++)
    FindOrbitalElements(sys->x[k], sys->y[k], sys->vx[k], sys->vy[k], 1.0 + sys->mass[k], k);
}

/* ---- self-gravity ----------------------------------------------------------- */

/* compute_selfgravity (src/sgmain.c:3-35): the acceleration of the host's
 * density into gr / gtheta.  The host calls it at start-up (before
 * InitEuler) and once per DT for its logs; inside AlgoGas the device runs
 * its own. */
void compute_selfgravity(PolarGrid *Rho, PolarGrid *Vr, PolarGrid *Vt, PolarGrid *DVr, PolarGrid *DVt, real DeltaT,
                         boolean SGUpdate) {
  (void)Vr; (void)Vt; (void)DVr; (void)DVt; (void)DeltaT;
  if (SGUpdate) unsupported("a host-side self-gravity velocity update");
  ensure_context();
  grid_once(&gr, "sgaccr");
  grid_once(&gtheta, "sgacctheta");
  grid_once(&torquesg, "torquesg");
  die(dfargo_upload_state(ctx, Rho->Field, NULL, NULL, NULL, NULL));
  die(dfargo_compute_selfgravity(ctx));
  die(dfargo_get_field(ctx, DFARGO_FIELD_SG_ACCR, gr->Field));
  die(dfargo_get_field(ctx, DFARGO_FIELD_SG_ACCT, gtheta->Field));
}

/* init_planetarysys_withSG (src/sgsysinit.c:3-49) runs inside
 * dfargo_init_euler, which InitEuler calls; the corrected velocities reach
 * the host's PlanetarySystem there. */
void init_planetarysys_withSG(PlanetarySystem *sys) { (void)sys; }

/* ---- Lagrangian particles ------------------------------------------------- */

static void dust_params_from_globals(dfargo_dust_params *d) {
  dfargo_dust_default_params(d, NBPART);
  d->sizeminpart = SIZEMINPART; d->sizemaxpart = SIZEMAXPART; d->sizepartslope = SIZEPARTSLOPE;
  d->rhopart = RHOPART; d->dustslope = DUSTSLOPE; d->rmindust = RMINDUST; d->rmaxdust = RMAXDUST;
  d->interpolation = (*INTERPOLATION == 'N' || *INTERPOLATION == 'n')   ? DFARGO_INTERP_NGP
                     : (*INTERPOLATION == 'C' || *INTERPOLATION == 'c') ? DFARGO_INTERP_CIC
                                                                          : DFARGO_INTERP_TSC;
  d->dust_feel_disk = DustFeelDisk; d->dust_feel_sg = DustFeelSG; d->dust_feel_planets = DustFeelPlanets;
  d->dust_feedback = DustFeedback; d->short_friction_time_approximation = ShortFrictionTimeApproximation;
  d->no_dt_constraint_by_particles = NoTimestepConstraintByParticles;
  d->dust_timestep = DUSTTIMESTEP;
  d->factor_unit_mass = FACTORUNITMASS; d->factor_unit_length = FACTORUNITLENGTH;
  d->particle_mass = Particles_Mass_Initial;
  d->seed = (long)time(NULL); /* srand48(time(NULL)), src/Dsys.c:173 */
  d->dust_feel_turb = DustFeelTurb; d->dust_growth = DustGrowth;
  d->remove_dust_from_hill = RemoveDustFromPlanetsHillRadius; d->zz_integrator = ZZIntegrator;
  d->dust_feel_sg_zero_mode = DustFeelSGZeroMode;
}

/* AllocDustSystem + InitDustSystem (src/Dsys.c:7-92, 128-262): the
 * DustSystem's 29 arrays of NBPART + 1 reals, the positions, sizes and
 * velocities drawn as the reference draws them */
DustSystem *InitDustSystem(void) {
  DustSystem *s = (DustSystem *)calloc(1, sizeof(DustSystem));
  real **a;
  size_t k, n = (size_t)NBPART + 1;
  dfargo_params p;
  dfargo_dust_state st;
  real *axi = NULL;
  if (!s) unsupported("allocating the dust system");
  for (a = (real **)s, k = 0; k < sizeof(DustSystem) / sizeof(real *); k++)
    if ((a[k] = (real *)calloc(n, sizeof(real))) == NULL) unsupported("allocating the dust system");
  params_from_globals(&p);
  dust_params_from_globals(&dpar);
  if (SelfGravity) {
    /* the ring means of the start-up radial self-gravity (GLOBAL_AxiSGAccr) */
    int i, j;
    axi = (real *)calloc((size_t)NRAD, sizeof(real));
    for (i = 0; i < NRAD && gr; i++) {
      for (j = 0; j < NSEC; j++) axi[i] += gr->Field[i * NSEC + j];
      axi[i] /= (real)NSEC;
    }
  }
  st.r = s->r; st.th = s->th; st.vr = s->vr; st.vth = s->vth; st.l = s->l;
  st.dustsize = s->dustsize; st.dustsize_init = s->dustsize_init; st.stokesnb = s->stokesnb;
  die(dfargo_init_dust_system(&p, &dpar, axi, &st));
  free(axi);
  return s;
}

void FreeDust(DustSystem *s) {
  real **a = (real **)s;
  size_t k;
  for (k = 0; k < sizeof(DustSystem) / sizeof(real *); k++) free(a[k]);
  free(s);
}

/* interpolation (src/Dsys.c:265-1149): the host calls it once, at start-up
 * (src/main.c:287); the particles then move to the device, which runs the
 * interpolation of every sub-step inside AlgoGas */
void interpolation(DustSystem *sys, PolarGrid *gasvr, PolarGrid *gasvt, PolarGrid *gasdens, PolarGrid *dustpcdens,
                   real timestep) {
  dfargo_dust_state st;
  (void)gasvr; (void)gasvt; (void)gasdens;
  if (dust_attached) unsupported("a second host-side interpolation");
  st.r = sys->r; st.th = sys->th; st.vr = sys->vr; st.vth = sys->vth; st.l = sys->l;
  st.dustsize = sys->dustsize; st.dustsize_init = sys->dustsize_init; st.stokesnb = sys->stokesnb;
  dpar.particle_mass = Particles_Mass_Initial; /* set by src/main.c:190-223 after InitDustSystem */
  die(dfargo_attach_dust(ctx, &dpar, &st, timestep));
  die(dfargo_download_dust(ctx, &st));
  die(dfargo_get_field(ctx, DFARGO_FIELD_DUST_DENSITY, dustpcdens->Field));
  dust_attached = 1;
}
