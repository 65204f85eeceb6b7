/*
 * fargo_host.c -- a small C host for the drop-in shim where the reference's
 * own host (src/main.c, src/Interpret.c, src/var.c, src/LowTasks.c, ...) is
 * not available, e.g. on a GPU box that has only this repository.
 *
 * It defines the host symbols fargo_shim.c imports (fargo_ref.h), reads a
 * .par file and its planet .cfg with dfargo_read_par /
 * dfargo_read_planet_config, and replays the reference's call sequence for
 * one rank (src/main.c:50-397): the grids, FillPolar1DArrays, the initial
 * density and energy, the start-up self-gravity, OmegaFrame, InitEuler,
 * then NTOT AlgoGas calls, writing the gas fields and planet files every
 * NINTERM calls in the reference's formats (src/Output.c:20-218).
 *
 *   fargo_host file.par [ntot]     (OUTPUTDIR of the .par receives the files)
 *
 * Restart, the per-DT logs, particles and the dust fluid are left to the
 * reference's own host.
 */
#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include "dfargo.h"
#include "fargo_ref.h"

/* the shim's functions the host calls (src/proto.h) */
void FillPolar1DArrays(void);
void InitGasDensity(PolarGrid *);
void InitGasEnergy(PolarGrid *);
void InitDustDensity(PolarGrid *);
void InitEuler(PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *,
               PlanetarySystem *);
void AlgoGas(Force *, PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *,
             PolarGrid *, PolarGrid *, PlanetarySystem *, DustSystem *);
void compute_selfgravity(PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *, PolarGrid *, real, boolean);
void init_planetarysys_withSG(PlanetarySystem *);
void SolveOrbits(PlanetarySystem *);
extern real LostMass;

/* ---- the reference host's data (src/var.c parameters, src/global.h,
 *      src/Interpret.c:23-50, src/main.c:10) --------------------------------- */
real DT, SIGMA0, RMIN, RMAX, PMIN, PMAX, SIGMASLOPE, ASPECTRATIO, FLARINGINDEX;
int NRAD, NSEC, NBPART, NRAD1D;
real VISCOSITY, ALPHAVISCOSITY, CAVITYRADIUS, CAVITYRATIO, CAVITYWIDTH;
real TRANSITIONRADIUS, TRANSITIONRATIO, TRANSITIONWIDTH, IMPOSEDDISKDRIFT;
real ADIABATICINDEX, BETACOOLINGTIME, BETACOOLINGSLOPE, SGTHICKNESSSMOOTHING;
real THICKNESSSMOOTHING, ROCHESMOOTHING, INDIRECTTERMSMOOTHING, MASSTAPER;
real EXCLUDEHILLFACTOR, RELEASEDATE, RELEASERADIUS, ECCENTRICITY, OMEGAFRAME;
real WKZRMIN, WKZRMAX, DAMPINGTODYNAMICALTIMERATIO, RMIN1D, RMAX1D;
real SIZEMINPART, SIZEMAXPART, SIZEPARTSLOPE, RHOPART, DUSTSLOPE, RMINDUST, RMAXDUST;
real DUSTTIMESTEP, FACTORUNITMASS = 1.0, FACTORUNITLENGTH = 1.0, DASPECTRATIO, LAMBDADOUBLING;
real DUSTTOGASDENSITYRATIO, DVISCOSITY, DALPHAVISCOSITY, DFLARINGINDEX, DUSTDIFFINN, Sizepart;
real floordens, MINDT = 100.0, MAXDT = 0.00001;
char INTERPOLATION[512] = "TSC", OUTPUTDIR[512];
boolean FastTransport = YES, GuidingCenter, BinaryCenter, Indirect_Term = YES, Discard_GasIndirect_term;
boolean Alternative_Gas_Indirect_Term, NonReflecting, Corotating, Evanescent;
boolean SelfGravity, SGZeroMode, AddNoise, EnergyEquation, ViscousHeating = YES, BetaCooling;
boolean ForcedCircular, DampToIni, DampToViscous, CorotateWithOuterPlanet;
boolean ShortFrictionTimeApproximation = YES, AddFloors = YES, DustFeelDisk = YES, DustFeelSG = YES;
boolean DustFeelPlanets = YES, DustFeelTurb, DustGrowth, ZZIntegrator, NoTimestepConstraintByParticles;
boolean DustFeedback, RemoveDustFromPlanetsHillRadius = YES, DustFeelSGZeroMode;
boolean BC1D_SS_ZeroVel, BC1D_SS_NonZeroVel, ImposeZeroRadialVelocityBC;
boolean OpenInner, Restart, ExcludeHill, LogGrid;
boolean DustFluid, DustDiffusion, OpenInnerDust;
int CPU_Rank = 0, CPU_Number = 1, IMIN = 0, GLOBALNRAD, Zero_or_active, Max_or_active;
real Rinf[MAX1D], Rsup[MAX1D], Rmed[MAX1D], Surf[MAX1D], InvRmed[MAX1D], InvSurf[MAX1D];
real InvDiffRmed[MAX1D], InvDiffRsup[MAX1D], InvRinf[MAX1D], Radii[MAX1D], GlobalRmed[MAX1D];
real Azimuth[MAX1D], CosAzimuth[MAX1D], SinAzimuth[MAX1D], AziInf[MAX1D], AziSup[MAX1D];
real InitialPlanetMass[MAX1D], FinalPlanetMass[MAX1D], SigmaMed[MAX1D], EnergyMed[MAX1D], DSigmaMed[MAX1D];
real OmegaFrame, PhysicalTime, PhysicalTimeInitial, MassTaper;
real stardiscbarycenter_x[3], stardiscbarycenter_y[3];
real Particles_Mass, Particles_Mass_Initial;
PolarGrid *CellAbscissa, *CellOrdinate, *Potential, *Pressure, *SoundSpeed, *Temperature;
PolarGrid *DivergenceVelocity, *TAURR, *TAUPP, *TAURP, *ViscHeat, *gr, *gtheta, *torquesg;
PolarGrid *RadFBAcc, *AziFBAcc, *FBedot, *Test;

static dfargo_params P; /* the decoded .par file */
static dfargo_dust_fluid_params F; /* its dust-fluid switches */

/* ---- the reference host's helpers --------------------------------------- */

/* CreatePolarGrid (src/LowTasks.c:84-115): (Nr+3)(Ns+1)+5 zeroed reals */
PolarGrid *CreatePolarGrid(int Nr, int Ns, char *name) {
  PolarGrid *g = (PolarGrid *)malloc(sizeof(PolarGrid));
  if (!g) prs_exit(1);
  g->Nrad = Nr;
  g->Nsec = Ns;
  g->Name = name;
  g->Field = (real *)calloc((size_t)(Nr + 3) * (size_t)(Ns + 1) + 5, sizeof(real));
  if (!g->Field) prs_exit(1);
  return g;
}

void InitComputeAccel(void) { /* src/SideEuler.c:263-280 */
  int i, j;
  CellAbscissa = CreatePolarGrid(NRAD, NSEC, "abscissa");
  CellOrdinate = CreatePolarGrid(NRAD, NSEC, "ordinate");
  for (i = 0; i < NRAD; i++)
    for (j = 0; j < NSEC; j++) {
      CellAbscissa->Field[i * NSEC + j] = Rmed[i] * CosAzimuth[j];
      CellOrdinate->Field[i * NSEC + j] = Rmed[i] * SinAzimuth[j];
    }
}

void masterprint(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vprintf(fmt, ap);
  va_end(ap);
  fflush(stdout);
}

void mastererr(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
}

void prs_exit(int code) {
  fflush(stdout);
  exit(code);
}

/* FindOrbitalElements (src/Planet.c:129-180): one orbit<n>.dat line */
void FindOrbitalElements(real x, real y, real vx, real vy, real m, int n) {
  double o[7];
  char name[1100];
  FILE *f;
  if (dfargo_orbital_elements(x, y, vx, vy, m, o)) return;
  snprintf(name, sizeof name, "%sorbit%d.dat", OUTPUTDIR, n);
  if ((f = fopen(name, "a")) == NULL) return;
  fprintf(f, "%.12g\t%.12g\t%.12g\t%.12g\t%.12g\t%.12g\t%.12g\n", PhysicalTime, o[0], o[1], o[2], o[3], o[4],
          o[6]);
  fclose(f);
}

/* FillSigma / FillEnergy (src/Theo.c:87-206): the ring profiles, from the
 * initial fields the library computes without noise */
void FillSigma(void) {
  dfargo_params p = P;
  real *f = (real *)calloc((size_t)(NRAD + 1) * NSEC, sizeof(real));
  int i;
  p.add_noise = 0;
  if (f && !dfargo_init_gas_density(&p, f))
    for (i = 0; i < NRAD; i++) SigmaMed[i] = f[i * NSEC];
  free(f);
}

/* FillDSigma (src/Theo.c:95-101): DSigma at Rmed */
void FillDSigma(void) {
  real *f = (real *)calloc((size_t)(NRAD + 1) * NSEC, sizeof(real));
  int i;
  if (f && !dfargo_init_dust_fluid_density(&P, &F, f))
    for (i = 0; i < NRAD; i++) DSigmaMed[i] = f[i * NSEC];
  free(f);
}

void FillEnergy(void) {
  real *f = (real *)calloc((size_t)(NRAD + 1) * NSEC, sizeof(real));
  int i;
  if (f && !dfargo_init_gas_energy(&P, f))
    for (i = 0; i < NRAD; i++) EnergyMed[i] = f[i * NSEC];
  free(f);
}

/* ---- output (src/Output.c:39-58, 178-218) ---------------------------------- */

static void write_field(const PolarGrid *g, const char *base, int number) {
  char name[1100];
  FILE *f;
  snprintf(name, sizeof name, "%s%s%d.dat", OUTPUTDIR, base, number);
  if ((f = fopen(name, "wb")) == NULL) {
    mastererr("cannot write %s\n", name);
    prs_exit(1);
  }
  fwrite(g->Field, sizeof(real), (size_t)g->Nrad * g->Nsec, f);
  fclose(f);
}

static void write_planets(const PlanetarySystem *s, int number) {
  char name[1100];
  FILE *f;
  int k;
  for (k = 0; k < s->nb; k++) {
    snprintf(name, sizeof name, "%splanet%d.dat", OUTPUTDIR, k);
    if ((f = fopen(name, number == 0 ? "w" : "a")) == NULL) prs_exit(1);
    fprintf(f, "%d\t%#.14g\t%#.14g\t%#.14g\t%#.14g\t%#.14g\t%#.14g\t%#.14g\t%#.14g\t%#.14g\t%#.14g\n", number,
            s->x[k], s->y[k], s->vx[k], s->vy[k], s->mass[k], LostMass, PhysicalTime, OmegaFrame, 0.0, 0.0);
    fclose(f);
  }
}

/* ---- the .par decoding into the globals ----------------------------------- */

static void globals_from_params(const dfargo_params *p, const dfargo_run_params *run) {
  DT = run->dt;
  NRAD = p->nrad; NSEC = p->nsec; GLOBALNRAD = p->nrad; RMIN = p->rmin; RMAX = p->rmax;
  LogGrid = p->log_grid; PMIN = p->pmin; PMAX = p->pmax;
  SIGMA0 = p->sigma0; SIGMASLOPE = p->sigma_slope; ASPECTRATIO = p->aspect_ratio;
  FLARINGINDEX = p->flaring_index; VISCOSITY = p->viscosity; ALPHAVISCOSITY = p->alpha_viscosity;
  CAVITYRADIUS = p->cavity_radius; CAVITYRATIO = p->cavity_ratio; CAVITYWIDTH = p->cavity_width;
  TRANSITIONRADIUS = p->transition_radius; TRANSITIONRATIO = p->transition_ratio;
  TRANSITIONWIDTH = p->transition_width; IMPOSEDDISKDRIFT = p->imposed_disk_drift;
  AddNoise = p->add_noise; EnergyEquation = p->energy_equation; ViscousHeating = p->viscous_heating;
  BetaCooling = p->beta_cooling; ADIABATICINDEX = p->adiabatic_index;
  BETACOOLINGTIME = p->beta_cooling_time; BETACOOLINGSLOPE = p->beta_cooling_slope;
  SelfGravity = p->self_gravity; SGZeroMode = p->sg_zero_mode; SGTHICKNESSSMOOTHING = p->sg_thickness_smoothing;
  THICKNESSSMOOTHING = p->thickness_smoothing; ROCHESMOOTHING = p->roche_smoothing;
  INDIRECTTERMSMOOTHING = p->indirect_term_smoothing; MASSTAPER = p->mass_taper;
  Indirect_Term = p->indirect_term; Discard_GasIndirect_term = p->discard_gas_indirect_term;
  Alternative_Gas_Indirect_Term = p->alternative_gas_indirect_term;
  ExcludeHill = p->exclude_hill; EXCLUDEHILLFACTOR = p->exclude_hill_factor;
  ForcedCircular = p->forced_circular; RELEASEDATE = p->release_date; RELEASERADIUS = p->release_radius;
  ECCENTRICITY = p->eccentricity; OMEGAFRAME = p->omega_frame;
  Corotating = p->corotating != DFARGO_FRAME_FIXED;
  GuidingCenter = p->corotating == DFARGO_FRAME_GUIDING_CENTER;
  BinaryCenter = p->corotating == DFARGO_FRAME_BINARY_CENTER;
  CorotateWithOuterPlanet = p->corotate_with_outer_planet;
  FastTransport = p->fast_transport; AddFloors = p->add_floors; floordens = p->density_floor;
  OpenInner = p->boundary == DFARGO_BC_OPEN; NonReflecting = p->boundary == DFARGO_BC_NONREFLECTING;
  Evanescent = p->wave_killing_zones; WKZRMIN = p->wkz_rmin; WKZRMAX = p->wkz_rmax;
  DAMPINGTODYNAMICALTIMERATIO = p->damping_ratio;
  DampToViscous = p->damp_to == DFARGO_DAMP_TO_VISCOUS; DampToIni = p->damp_to == DFARGO_DAMP_TO_INI;
  ImposeZeroRadialVelocityBC = p->zero_vrad_bc;
  NRAD1D = p->nrad1d; RMIN1D = p->rmin1d; RMAX1D = p->rmax1d;
  BC1D_SS_ZeroVel = p->bc1d == 1; BC1D_SS_NonZeroVel = p->bc1d == 2;
  Zero_or_active = 0;
  Max_or_active = NRAD;
  snprintf(OUTPUTDIR, sizeof OUTPUTDIR, "%s", run->output_dir);
  if (OUTPUTDIR[0] && OUTPUTDIR[strlen(OUTPUTDIR) - 1] != '/') strncat(OUTPUTDIR, "/", sizeof OUTPUTDIR - 1);
}

static PlanetarySystem *planetary_system(const dfargo_planets *d) {
  PlanetarySystem *s = (PlanetarySystem *)calloc(1, sizeof(PlanetarySystem));
  const int n = d->nb > 0 ? d->nb : 1;
  int k;
  s->nb = d->nb;
  s->mass = calloc(n, sizeof(real)); s->x = calloc(n, sizeof(real)); s->y = calloc(n, sizeof(real));
  s->vx = calloc(n, sizeof(real)); s->vy = calloc(n, sizeof(real)); s->acc = calloc(n, sizeof(real));
  s->name = calloc(n, sizeof(char *));
  s->FeelDisk = calloc(n, sizeof(boolean)); s->FeelOthers = calloc(n, sizeof(boolean));
  s->Binary = calloc(n, sizeof(boolean));
  for (k = 0; k < d->nb; k++) {
    s->mass[k] = d->mass[k]; s->x[k] = d->x[k]; s->y[k] = d->y[k];
    s->vx[k] = d->vx[k]; s->vy[k] = d->vy[k]; s->acc[k] = d->acc[k];
    s->FeelDisk[k] = d->feel_disk[k]; s->FeelOthers[k] = d->feel_others[k]; s->Binary[k] = d->binary[k];
    InitialPlanetMass[k] = d->initial_mass[k];
    FinalPlanetMass[k] = d->final_mass[k];
  }
  return s;
}

int main(int argc, char **argv) {
  dfargo_run_params run;
  dfargo_planets pl;
  PlanetarySystem *sys;
  PolarGrid *rho, *vr, *vt, *en, *label, *pcdens, *drho, *dvr, *dvt;
  Force force;
  int i, ntot;
  if (argc < 2) {
    fprintf(stderr, "usage: %s file.par [ntot]\n", argv[0]);
    return 1;
  }
  if (dfargo_read_par(argv[1], &P, &run)) {
    fprintf(stderr, "%s\n", dfargo_last_error());
    return 1;
  }
  globals_from_params(&P, &run);
  {
    int on = 0;
    if (dfargo_read_dust_fluid_par(argv[1], &F, &on)) {
      fprintf(stderr, "%s\n", dfargo_last_error());
      return 1;
    }
    /* src/Interpret.c:253-342 */
    DustFluid = on; Sizepart = F.sizepart; RHOPART = F.rhopart;
    FACTORUNITMASS = F.factor_unit_mass; FACTORUNITLENGTH = F.factor_unit_length;
    DUSTTOGASDENSITYRATIO = F.dust_to_gas_density_ratio; DVISCOSITY = F.dviscosity;
    DALPHAVISCOSITY = F.dalpha_viscosity; DASPECTRATIO = F.daspect_ratio; DFLARINGINDEX = F.dflaring_index;
    DUSTDIFFINN = F.dust_diff_inn; ShortFrictionTimeApproximation = F.sfta; DustFeelDisk = F.dust_feel_disk;
    DustFeedback = F.dust_feedback; DustFeelSG = F.dust_feel_sg != DFARGO_DUST_SG_NO;
    DustFeelSGZeroMode = F.dust_feel_sg == DFARGO_DUST_SG_ZERO_MODE; DustDiffusion = F.dust_diffusion;
    OpenInnerDust = F.open_inner_dust;
  }
  ntot = argc > 2 ? atoi(argv[2]) : run.ntot;
  memset(&pl, 0, sizeof pl);
  if (run.planet_config[0] && dfargo_read_planet_config(run.planet_config, &P, &pl)) {
    fprintf(stderr, "%s\n", dfargo_last_error());
    return 1;
  }
  /* src/main.c:145-186 */
  rho = CreatePolarGrid(NRAD, NSEC, "gasdens");
  vr = CreatePolarGrid(NRAD, NSEC, "gasvrad");
  vt = CreatePolarGrid(NRAD, NSEC, "gasvtheta");
  en = CreatePolarGrid(NRAD, NSEC, "gasenergy");
  label = CreatePolarGrid(NRAD, NSEC, "gaslabel");
  pcdens = CreatePolarGrid(NRAD, NSEC, "pcdens");
  drho = CreatePolarGrid(NRAD, NSEC, "dustdens");
  dvr = CreatePolarGrid(NRAD, NSEC, "dustvrad");
  dvt = CreatePolarGrid(NRAD, NSEC, "dustvtheta");
  memset(&force, 0, sizeof force);
  FillPolar1DArrays();
  sys = planetary_system(&pl);
  InitGasDensity(rho);
  if (DustFluid) InitDustDensity(drho);
  if (EnergyEquation) InitGasEnergy(en);
  if (SelfGravity) {
    compute_selfgravity(rho, vr, vt, dvr, dvt, 0.0, NO);
    init_planetarysys_withSG(sys);
  }
  /* src/main.c:228-246: the frame rate (corotating: the planet's) */
  OmegaFrame = OMEGAFRAME;
  if (Corotating) {
    const int k = CorotateWithOuterPlanet ? sys->nb - 1 : 0;
    const real r = sqrt(sys->x[k] * sys->x[k] + sys->y[k] * sys->y[k]);
    OmegaFrame = sqrt(sys->vx[k] * sys->vx[k] + sys->vy[k] * sys->vy[k]) / r;
  }
  InitEuler(vr, vt, rho, en, dvr, dvt, drho, sys);
  write_field(rho, "gasdens", 0);
  write_field(vr, "gasvrad", 0);
  write_field(vt, "gasvtheta", 0);
  if (DustFluid) {
    write_field(drho, "dustdens", 0);
    write_field(dvr, "dustvrad", 0);
    write_field(dvt, "dustvtheta", 0);
  }
  write_planets(sys, 0);
  /* src/main.c:302-397 */
  for (i = 1; i <= ntot; i++) {
    AlgoGas(&force, rho, vr, vt, en, label, drho, pcdens, dvr, dvt, sys, NULL);
    SolveOrbits(sys);
    masterprint("\n");
    if (run.ninterm > 0 && i % run.ninterm == 0) {
      const int out = i / run.ninterm;
      write_field(rho, "gasdens", out);
      write_field(vr, "gasvrad", out);
      write_field(vt, "gasvtheta", out);
      if (EnergyEquation) write_field(en, "gasenergy", out);
      if (DustFluid) {
        write_field(drho, "dustdens", out);
        write_field(dvr, "dustvrad", out);
        write_field(dvt, "dustvtheta", out);
      }
      write_planets(sys, out);
    }
  }
  return 0;
}
