/*
 * fargo_ref.h -- the part of Dusty FARGO-ADSG's host ABI that the drop-in
 * shim (fargo_shim.c) compiles against, declared here so that the shim
 * builds without the reference's generated headers (param.h, global_ex.h,
 * sg_ex.h, which its own build writes with src/varparser.pl).
 *
 * The structure layouts must equal the reference's src/types.h (field
 * order and types); tests/test_shim.py compares them field by field where
 * the reference sources are present.  Every extern below is a symbol the
 * reference's host objects define: parameters of src/var.c (REAL -> real,
 * INT -> int, STRING -> char[512], as src/varparser.pl:40-50 declares
 * them), tables and scalars of src/global.h, switches of
 * src/Interpret.c:23-50 and src/main.c:10-11, and the helper functions of
 * src/LowTasks.c, src/SideEuler.c, src/Planet.c and src/Psys.c.
 */
#ifndef FARGO_REF_H
#define FARGO_REF_H

#include <stdio.h>

typedef int boolean; /* src/types.h:9 */
typedef double real; /* src/types.h:10 */

#define YES 1
#define NO 0
#define MAX1D 32768 /* src/types.h:78 */

/* src/types.h:29-34 */
typedef struct pair {
  real x;
  real y;
} Pair;

/* src/types.h:38-48 */
typedef struct force {
  real fx_inner, fy_inner, fx_ex_inner, fy_ex_inner;
  real fx_outer, fy_outer, fx_ex_outer, fy_ex_outer;
  real *GlobalForce;
} Force;

/* src/types.h:50-57 */
typedef struct polargrid {
  int Nrad;
  int Nsec;
  real *Field;
  char *Name;
} PolarGrid;

/* src/types.h:97-111 */
typedef struct planetary_system {
  int nb;
  real *mass, *x, *y, *vx, *vy, *acc;
  char **name;
  boolean *FeelDisk, *FeelOthers, *Binary;
} PlanetarySystem;

/* src/types.h:114-146: 29 arrays of NBPART reals */
typedef struct DustParticlesSystem {
  real *dustsize, *dustsize_init, *r, *th, *vr, *vth, *l, *dvr, *dvth;
  real *rad_geff_acc, *azi_geff_acc, *rad_gradp, *azi_gradp, *rad_drag_acc, *azi_drag_acc;
  real *rad_ind_acc, *azi_ind_acc, *rad_sg_acc, *azi_sg_acc, *jacobi, *stokesnb;
  real *gasvratpc, *gasvtatpc, *gasdensatpc, *gascsatpc;
  real *senddusttoprevCPU, *senddusttonextCPU, *recvdustfromprevCPU, *recvdustfromnextCPU;
} DustSystem;

/* ---- parameters (src/var.c) ---------------------------------------------- */
extern real DT, SIGMA0, RMIN, RMAX, PMIN, PMAX, SIGMASLOPE, ASPECTRATIO, FLARINGINDEX;
extern int NRAD, NSEC, NBPART, NRAD1D;
extern real VISCOSITY, ALPHAVISCOSITY, CAVITYRADIUS, CAVITYRATIO, CAVITYWIDTH;
extern real TRANSITIONRADIUS, TRANSITIONRATIO, TRANSITIONWIDTH, IMPOSEDDISKDRIFT;
extern real ADIABATICINDEX, BETACOOLINGTIME, BETACOOLINGSLOPE, SGTHICKNESSSMOOTHING;
extern real THICKNESSSMOOTHING, ROCHESMOOTHING, INDIRECTTERMSMOOTHING, MASSTAPER;
extern real EXCLUDEHILLFACTOR, RELEASEDATE, RELEASERADIUS, ECCENTRICITY, OMEGAFRAME;
extern real WKZRMIN, WKZRMAX, DAMPINGTODYNAMICALTIMERATIO, RMIN1D, RMAX1D;
extern real SIZEMINPART, SIZEMAXPART, SIZEPARTSLOPE, RHOPART, DUSTSLOPE, RMINDUST, RMAXDUST;
extern real DUSTTIMESTEP, FACTORUNITMASS, FACTORUNITLENGTH, DASPECTRATIO, LAMBDADOUBLING;
extern real DUSTTOGASDENSITYRATIO, DVISCOSITY, DALPHAVISCOSITY, DFLARINGINDEX, DUSTDIFFINN, Sizepart;
extern real floordens; /* DENSITYFLOOR (src/var.c) */
extern char INTERPOLATION[512], OUTPUTDIR[512];

/* ---- switches (src/Interpret.c:23-50, src/main.c:10, src/global.h:45,58) -- */
extern boolean FastTransport, GuidingCenter, BinaryCenter, Indirect_Term, Discard_GasIndirect_term;
extern boolean Alternative_Gas_Indirect_Term, NonReflecting, Corotating, Evanescent;
extern boolean SelfGravity, SGZeroMode, AddNoise, EnergyEquation, ViscousHeating, BetaCooling;
extern boolean ForcedCircular, DampToIni, DampToViscous, CorotateWithOuterPlanet;
extern boolean ShortFrictionTimeApproximation, AddFloors, DustFeelDisk, DustFeelSG, DustFeelPlanets;
extern boolean DustFeelTurb, DustGrowth, ZZIntegrator, NoTimestepConstraintByParticles;
extern boolean DustFeedback, RemoveDustFromPlanetsHillRadius, DustFeelSGZeroMode;
extern boolean BC1D_SS_ZeroVel, BC1D_SS_NonZeroVel, ImposeZeroRadialVelocityBC;
extern boolean OpenInner, Restart, ExcludeHill, LogGrid;
extern boolean DustFluid, DustDiffusion, OpenInnerDust;

/* ---- tables and scalars (src/global.h) ------------------------------------ */
extern int CPU_Rank, CPU_Number, IMIN, GLOBALNRAD, Zero_or_active, Max_or_active;
extern real Rinf[MAX1D], Rsup[MAX1D], Rmed[MAX1D], Surf[MAX1D], InvRmed[MAX1D], InvSurf[MAX1D];
extern real InvDiffRmed[MAX1D], InvDiffRsup[MAX1D], InvRinf[MAX1D], Radii[MAX1D], GlobalRmed[MAX1D];
extern real Azimuth[MAX1D], CosAzimuth[MAX1D], SinAzimuth[MAX1D], AziInf[MAX1D], AziSup[MAX1D];
extern real InitialPlanetMass[MAX1D], FinalPlanetMass[MAX1D];
extern real DSigmaMed[MAX1D];
extern real OmegaFrame, PhysicalTime, PhysicalTimeInitial, MassTaper;
extern real MINDT, MAXDT; /* MINDT, MAXDT of src/var.c */
extern real stardiscbarycenter_x[3], stardiscbarycenter_y[3];
extern real Particles_Mass, Particles_Mass_Initial;
extern PolarGrid *CellAbscissa, *CellOrdinate, *Potential, *Pressure, *SoundSpeed, *Temperature;
extern PolarGrid *DivergenceVelocity, *TAURR, *TAUPP, *TAURP, *ViscHeat, *gr, *gtheta, *torquesg;
extern PolarGrid *RadFBAcc, *AziFBAcc, *FBedot, *Test;

/* ---- host helpers -------------------------------------------------------- */
PolarGrid *CreatePolarGrid(int Nr, int Ns, char *name); /* src/LowTasks.c:84-115 */
void InitComputeAccel(void);                            /* src/SideEuler.c:263-280 */
void masterprint(const char *fmt, ...);                 /* src/LowTasks.c:39-46 */
void mastererr(const char *fmt, ...);                   /* src/LowTasks.c:48-55 */
void prs_exit(int code);                                /* src/LowTasks.c:32-37 */
void FindOrbitalElements(real x, real y, real vx, real vy, real m, int n); /* src/Planet.c:129 */

#endif
