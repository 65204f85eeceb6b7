/*
 * oracle.c -- FP64 CPU restatement of Dusty FARGO-ADSG's AlgoGas sub-step.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Not linked into the product.
 *
 * Every function follows the reference line by line, with the same
 * operation order, so that pointwise results are bitwise reproducible by a
 * device build compiled without contraction.  Citations are to the
 * Dusty FARGO-ADSG sources (src/...).
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   tables          closed forms of Rmed, Surf, log spacing
 *   init            equilibrium of the initial state (vr stays ~h^2 small)
 *   CFL             closed-form dt for a uniform, static disc
 *   SubStep1        free fall at rest is exactly -dt/(R_i R_{i-1}); a
 *                   linear pressure ramp adds -dt dP/dR/rho; the initial
 *                   equilibrium's kick is small (test_oracle_stages.py)
 *   viscosity       solid-body rotation is stress-free; Keplerian shear
 *                   stress matches -3/2 nu Sigma Omega to O(dr^2); a uniform
 *                   radial flow gives e_phiphi = div v = alpha/r and
 *                   tau_rr, tau_phiphi = -2/3, 4/3 rho nu alpha/r exactly;
 *                   an azimuthal sine gives the forward difference
 *   heating         Keplerian shear on a power-law disc: Q is an exact
 *                   power law of Rmed including the extrapolated ring 0,
 *                   first-order convergent to 9/4 nu Sigma Omega^2
 *   SubStep2        von Neumann-Richtmyer q = C^2 rho dv^2 on single radial
 *                   and azimuthal jumps, faces pushed apart by dt q/(rho dR),
 *                   compressional heating dt q |dv|/dR; expansion untouched
 *   SubStep3        closed-form beta cooling e/(1+dt/t_c); pdV of a uniform
 *                   radial flow with heating, backward Euler in closed form
 *                   and convergent to e0 exp(-(gamma-1) alpha t/r)
 *   potential       Plummer planet + its indirect term in closed form (with
 *                   and without INDIRECTTERM); the gas indirect term is
 *                   the disc's softened pull on the star, summed over the
 *                   sub-step's starting state, as a uniform field
 *   OPEN edge       ghost rings copy the density, scale v_theta as r^-1/2,
 *                   pass v_r only for outflow
 *   particles       a particle feeling only the star and a planet follows
 *                   the softened point-mass acceleration in vector form;
 *                   FBedot integrates to the drag dissipation m |dv|^2/t_s;
 *                   TSC off the nearest node sums the textbook kernel at
 *                   the reference's three nodes
 *   floors/crash    floors lift only what is below; the crash flag
 *                   overwrite, ActualiseGas's zero edges, the unconditional
 *                   sub-Keplerian reset and the double mass taper as spec
 *                   tests of their lines
 *   transport       mass conservation, zero-velocity identity, an integer
 *                   solid-body rotation equals numpy.roll, TVD in 1-D
 *   FARGO shift     Nshift = round(Omega dt Ns / 2pi) on solid-body rings
 *   SG              brute-force direct convolution on tiny grids,
 *                   axisymmetric disc gives g_theta = 0, mirror antisymmetry
 *   FFT             naive DFT on mixed-radix sizes
 *   RK5             Cash-Karp tableau identities, 5th-order convergence,
 *                   Kepler energy and angular momentum
 *   diagnostics     budgets of a uniform Keplerian disc in closed form, a
 *                   one-cell disc's force on a planet, Hill exclusion,
 *                   orbital elements at pericentre and at E = pi/2
 *   frames          (psys_info, correct_vtheta) a circular planet stays put
 *                   in FRAME C at rate sqrt(1+m); a run started inertial
 *                   equals one in a frame fixed at that rate; FRAME G turns
 *                   at the mean motion of an eccentric orbit, also from an
 *                   apocentre start in the anomaly's |arg| >= 1 branch;
 *                   FRAME B (also with massless planets, on the final
 *                   masses) and the outer-planet option follow the right
 *                   body; particles in a rotating frame are the fixed
 *                   frame's turned by OmegaFrame t
 *   accretion       (accrete_onto_planets) mass conservation planet+gas,
 *                   per-cell fractions by distance, momentum conservation
 *   wave-killing    (evanescent_boundary) ring means kept, deviations
 *                   scaled by 1/(1+lambda) with lambda's closed-form
 *                   profile; convergence to the initial profiles; ZEROVRADBC
 *   non-reflecting  (nonreflecting_boundary) edge rings are shifted copies
 *                   with the initial means, vr at the faces follows the
 *                   outgoing acoustic relation, zero turn at the planets'
 *                   pitch radii, the adiabatic density/energy relation
 *   SG zero mode    (or_sg_zero_mode) equals the full field's ring values of
 *                   an axisymmetric disc up to the two self-force terms
 *   1-D disc        (solve_1d_viscous) dSigma/dt of power laws under a
 *                   constant viscosity in closed form, vr = -3/2 nu/R for a
 *                   uniform disc, and vr of power-law discs against the
 *                   reference's formula with its 1.5 on the Sigma' term
 *   (tests/test_oracle_frames.py for the last six)
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define G 1.0
#define CFLSECURITY 0.5 /* src/SourceEuler.c:10 */
#define CVNR 1.41       /* src/SourceEuler.c:13 */
#define MAXPL DFARGO_MAX_PLANETS

typedef struct { double x, y; } pair_t;

struct or_ctx {
  dfargo_params p;
  int nr, ns;
  double fargo_margin; /* smallest 1/2 - |Ntilde - Nshift| of the last AlgoGas call */
  size_t gsize; /* doubles per grid, as CreatePolarGrid (src/LowTasks.c:97) */
  /* 1-D tables (src/global.h:29-37) */
  double *Radii, *Rinf, *Rsup, *Rmed, *Surf, *InvRmed, *InvSurf, *InvDiffRmed;
  double *InvDiffRsup, *InvRinf, *GlobalRmed, *Azimuth, *CosAz, *SinAz;
  double *SigmaMed, *SigmaInf, *EnergyMed, *VthetaMed, *VradMed, *GlobalCs;
  double *AxiSG;
  double *cabs, *cord; /* CellAbscissa / CellOrdinate */
  /* state */
  double *rho, *vr, *vt, *en;
  dfargo_planets sys;
  /* module-static work grids of the reference */
  double *cs, *pres, *temp, *pot, *vradint, *vthetaint, *vradnew, *vthetanew;
  double *energyint, *energynew, *rhoint, *temperint, *divv, *drr, *drp, *dpp;
  double *trr, *trp, *tpp, *vischeat;
  double *rp, *rm, *tp, *tm, *rhostar, *work, *qrstar, *vthetares, *elong;
  double *dq, *tempshift;
  long *nshift;
  double *vmed, *vmoy;
  int *nosplit;
  /* self-gravity */
  int sg_ready;
  double sg_eps, sg_rstep, sg_tstep;
  double *kr_re, *kr_im, *kt_re, *kt_im;
  double *sg_accr, *sg_acct, *gr, *gtheta;
  /* scalars */
  double PhysicalTime, PhysicalTimeInitial, OmegaFrame, MINDT, MAXDT, LostMass;
  double MassTaper;
  double Xplanet, Yplanet; /* GetPsysInfo's mark (src/Psys.c:334-344) */
  /* the 1-D viscous disc of DAMPTOVISCOUS (src/SideEuler.c:1444-1565) */
  int n1d;
  double *rmed1d, *sigma1d, *vrad1d, *nu1d, *t1, *t2, *t3;
  pair_t dop, dop_alt, indirect;
  double barx[3], bary[3];
  int GasTimeStepsCFL, last_substeps, crashed_d, crashed_e, crashed;
  long substeps;
  /* gas grids the particles read and write (src/SourceEuler.c:210-219) */
  double *indpot, *radgradp, *azigradp, *radindacc, *aziindacc, *radsgacc, *azisgacc;
  double *radfbacc, *azifbacc, *fbedot, *dustdens;
  /* dust particles (src/types.h:114-146) */
  int dust_on, np;
  dfargo_dust_params dp;
  double *d_size, *d_size_init, *d_r, *d_th, *d_vr, *d_vth, *d_l;
  double *d_vrg, *d_vtg, *d_densg, *d_csg, *d_radgradp, *d_azigradp, *d_radind, *d_aziind;
  double *d_radsg, *d_azisg, *d_raddrag, *d_azidrag, *d_stokes;
  double *AziInf, *AziSup;
  double particles_mass_initial, particles_mass, min_stopping_time, unit_mass, unit_length;
  /* the dust fluid (DUSTFLUID YES): DRho, DVrad, DVtheta and the module
     grids of src/SourceEuler.c, src/Viscosity.c and src/TransportEuler.c
     that carry its terms */
  int df_on;
  dfargo_dust_fluid_params fp;
  double *drho, *dvr, *dvt, *stokes, *dpres, *dvradint, *dvthetaint, *dvradnew, *dvthetanew;
  double *dqr, *dqt, *ddrr, *ddrp, *ddpp, *ddivv, *dtrr, *dtrp, *dtpp, *fdiffrp, *fdifftp;
  double *dcs1, *DSigmaMed, *DVthetaMed, *DVradMed, *GlobalDCs, *AxiSGd, *dvmed, *dvmoy;
  long *dnshift;
  double LostMassd;
  int df_skip_momenta; /* Transportd in the short-friction-time approximation */
};

static double min2(double a, double b) { return (b < a) ? b : a; }
static void df_stokes(or_ctx *c);
static void df_viscous_terms(or_ctx *c, const double *vr, const double *vt);
static void df_update_viscosity(or_ctx *c, double *vr, double *vt, double DeltaT);
static void df_init_gas_feedback(or_ctx *c);
static double dfviscosity(const or_ctx *c, double rad);
static int df_has_viscosity(const or_ctx *c);
static void df_sound_speed_pressure(or_ctx *c);
static void df_init_stokes(or_ctx *c);
static void df_init_velocities(or_ctx *c);
static void df_wall_boundaries(or_ctx *c);
static void df_subkeplerian_boundary(or_ctx *c);
static void df_open_boundary(or_ctx *c);
static double max2(double a, double b) { return (b > a) ? b : a; }

static double *grid(or_ctx *c) { return (double *)calloc(c->gsize, sizeof(double)); }

/* ------------------------------------------------------------------------ */
/* profiles: Sigma, Energy, AspectRatio, FViscosity, compute_smoothing       */
/* ------------------------------------------------------------------------ */

/* Sigma (src/Theo.c:13-60), power law with optional cavity, ScalingFactor 1 */
static double sigma_profile(const dfargo_params *p, double r) {
  double cavity = 1.0;
  if (r < p->cavity_radius) cavity = 1.0 / p->cavity_ratio;
  return cavity * 1.0 * p->sigma0 * pow(r, -p->sigma_slope);
}

/* Energy (src/Theo.c:156-170) with R = MU = 1 */
static double energy_profile(const dfargo_params *p, double r) {
  double energy0 = 1.0 / 1.0 / (p->adiabatic_index - 1.0) * sigma_profile(p, r) *
                   pow(p->aspect_ratio, 2.0) * pow(r, -1.0 + 2.0 * p->flaring_index);
  return 1.0 * 1.0 * energy0;
}

/* AspectRatio (src/Viscosity.c:107-122), LAMBDADOUBLING = 0 */
static double aspect_ratio(const dfargo_params *p, double rad) {
  double ar = p->aspect_ratio;
  double rmin = p->transition_radius - p->transition_width * p->aspect_ratio;
  double rmax = p->transition_radius + p->transition_width * p->aspect_ratio;
  double scale = 1.0;
  rmin *= scale;
  rmax *= scale;
  if (rad < rmin) ar *= p->transition_ratio;
  if ((rad >= rmin) && (rad <= rmax))
    ar *= exp((rmax - rad) / (rmax - rmin) * log(p->transition_ratio));
  return ar;
}

/* FViscosity (src/Viscosity.c:23-69) */
double or_fviscosity(or_ctx *c, double rad) {
  const dfargo_params *p = &c->p;
  double viscosity = p->viscosity, rmin, rmax;
  int i = 0;
  if (p->alpha_viscosity != 0.0) {
    while (c->GlobalRmed[i] < rad) i++;
    viscosity = p->alpha_viscosity * c->GlobalCs[i] * c->GlobalCs[i] * pow(rad, 1.5);
  }
  rmin = p->cavity_radius -
         p->cavity_width * p->aspect_ratio * pow(p->cavity_radius, 1.0 + p->flaring_index);
  rmax = p->cavity_radius +
         p->cavity_width * p->aspect_ratio * pow(p->cavity_radius, 1.0 + p->flaring_index);
  if (rad < rmin) viscosity *= p->cavity_ratio;
  if ((rad >= rmin) && (rad <= rmax))
    viscosity *= exp((rmax - rad) / (rmax - rmin) * log(p->cavity_ratio));
  return viscosity;
}

static int has_viscosity(const dfargo_params *p) {
  return (p->alpha_viscosity != 0.0) || (p->viscosity != 0.0);
}

/* compute_smoothing (src/Force.c:255-263) */
static double compute_smoothing(const dfargo_params *p, double r) {
  return p->thickness_smoothing * aspect_ratio(p, r) * pow(r, 1.0 + p->flaring_index);
}

/* mpi_make1Dprofile (src/axilib.c:5-57), one rank */
static void make_1d_profile(or_ctx *c, const double *f, double *axi) {
  for (int i = 0; i < c->nr; i++) {
    double s = 0.;
    for (int j = 0; j < c->ns; j++) s += f[i * c->ns + j];
    s /= (double)c->ns;
    axi[i] = s;
  }
}

/* ------------------------------------------------------------------------ */
/* creation: FillPolar1DArrays, InitComputeAccel, allocations                */
/* ------------------------------------------------------------------------ */

or_ctx *or_create(const dfargo_params *p) {
  or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
  c->fargo_margin = 0.5;
  int nr = p->nrad, ns = p->nsec, i, j;
  c->p = *p;
  c->nr = nr;
  c->ns = ns;
  c->gsize = (size_t)(nr + 3) * (size_t)(ns + 1) + 5;
#define T1(n) (double *)calloc((size_t)(n), sizeof(double))
  c->Radii = T1(nr + 2); c->Rinf = T1(nr + 2); c->Rsup = T1(nr + 2);
  c->Rmed = T1(nr + 2); c->Surf = T1(nr + 2); c->InvRmed = T1(nr + 2);
  c->InvSurf = T1(nr + 2); c->InvDiffRmed = T1(nr + 2); c->InvDiffRsup = T1(nr + 2);
  c->InvRinf = T1(nr + 2); c->GlobalRmed = T1(nr + 2); c->Azimuth = T1(ns);
  c->CosAz = T1(ns); c->SinAz = T1(ns); c->SigmaMed = T1(nr + 2);
  c->SigmaInf = T1(nr + 2); c->EnergyMed = T1(nr + 2); c->VthetaMed = T1(nr + 2);
  c->VradMed = T1(nr + 2); c->GlobalCs = T1(nr + 2); c->AxiSG = T1(nr + 2);
  c->vmed = T1(nr + 2); c->vmoy = T1(nr + 2);
#undef T1
  c->nshift = (long *)calloc((size_t)nr + 2, sizeof(long));
  c->nosplit = (int *)calloc((size_t)nr + 2, sizeof(int));
  /* FillPolar1DArrays (src/SourceEuler.c:71-188), no radii.dat */
  double drrsep = (p->rmax - p->rmin) / (double)nr;
  for (i = 0; i <= nr; i++) {
    if (p->log_grid)
      c->Radii[i] = p->rmin * exp((double)i / (double)nr * log(p->rmax / p->rmin));
    else
      c->Radii[i] = p->rmin + drrsep * (double)(i);
  }
  for (i = 0; i < nr; i++) {
    double *R = c->Radii;
    c->GlobalRmed[i] = 2.0 / 3.0 * (R[i + 1] * R[i + 1] * R[i + 1] - R[i] * R[i] * R[i]);
    c->GlobalRmed[i] = c->GlobalRmed[i] / (R[i + 1] * R[i + 1] - R[i] * R[i]);
  }
  for (i = 0; i < nr; i++) {
    c->Rinf[i] = c->Radii[i];
    c->Rsup[i] = c->Radii[i + 1];
    c->Rmed[i] = 2.0 / 3.0 * (c->Rsup[i] * c->Rsup[i] * c->Rsup[i] - c->Rinf[i] * c->Rinf[i] * c->Rinf[i]);
    c->Rmed[i] = c->Rmed[i] / (c->Rsup[i] * c->Rsup[i] - c->Rinf[i] * c->Rinf[i]);
    c->Surf[i] = 0.5 * (p->pmax - p->pmin) * (c->Rsup[i] * c->Rsup[i] - c->Rinf[i] * c->Rinf[i]) / (double)ns;
    c->InvRmed[i] = 1.0 / c->Rmed[i];
    c->InvSurf[i] = 1.0 / c->Surf[i];
    c->InvDiffRsup[i] = 1.0 / (c->Rsup[i] - c->Rinf[i]);
    c->InvRinf[i] = 1.0 / c->Rinf[i];
  }
  c->Rinf[nr] = c->Radii[nr];
  for (i = 1; i < nr; i++) c->InvDiffRmed[i] = 1.0 / (c->Rmed[i] - c->Rmed[i - 1]);
  for (j = 0; j < ns; j++) {
    c->Azimuth[j] = p->pmin + (p->pmax - p->pmin) * (double)j / (double)ns;
    c->CosAz[j] = cos(c->Azimuth[j]);
    c->SinAz[j] = sin(c->Azimuth[j]);
  }
  /* grids */
  c->rho = grid(c); c->vr = grid(c); c->vt = grid(c); c->en = grid(c);
  c->cs = grid(c); c->pres = grid(c); c->temp = grid(c); c->pot = grid(c);
  c->vradint = grid(c); c->vthetaint = grid(c); c->vradnew = grid(c);
  c->vthetanew = grid(c); c->energyint = grid(c); c->energynew = grid(c);
  c->rhoint = grid(c); c->temperint = grid(c); c->divv = grid(c); c->drr = grid(c);
  c->drp = grid(c); c->dpp = grid(c); c->trr = grid(c); c->trp = grid(c);
  c->tpp = grid(c); c->vischeat = grid(c); c->rp = grid(c); c->rm = grid(c);
  c->tp = grid(c); c->tm = grid(c); c->rhostar = grid(c); c->work = grid(c);
  c->qrstar = grid(c); c->vthetares = grid(c); c->elong = grid(c);
  c->dq = grid(c); c->tempshift = grid(c); c->cabs = grid(c); c->cord = grid(c);
  c->sg_accr = grid(c); c->sg_acct = grid(c); c->gr = grid(c); c->gtheta = grid(c);
  c->indpot = grid(c); c->radgradp = grid(c); c->azigradp = grid(c); c->radindacc = grid(c);
  c->aziindacc = grid(c); c->radsgacc = grid(c); c->azisgacc = grid(c); c->radfbacc = grid(c);
  c->azifbacc = grid(c); c->fbedot = grid(c); c->dustdens = grid(c);
  /* InitComputeAccel (src/SideEuler.c:263-280) */
  for (i = 0; i < nr; i++)
    for (j = 0; j < ns; j++) {
      c->cabs[j + i * ns] = c->Rmed[i] * cos(c->Azimuth[j]);
      c->cord[j + i * ns] = c->Rmed[i] * sin(c->Azimuth[j]);
    }
  /* FillSigma (src/Theo.c:85-91); FillEnergy (src/Theo.c:172-176) */
  for (i = 0; i < nr; i++) {
    c->SigmaMed[i] = sigma_profile(p, c->Rmed[i]);
    c->SigmaInf[i] = sigma_profile(p, c->Rinf[i]);
    if (p->energy_equation) c->EnergyMed[i] = energy_profile(p, c->Rmed[i]);
  }
  c->OmegaFrame = p->omega_frame;
  c->MINDT = 100.0; /* src/var.c:161 defaults */
  c->MAXDT = 0.00001;
  return c;
}

void or_destroy(or_ctx *c) {
  if (!c) return;
  free(c->rmed1d); free(c->sigma1d); free(c->vrad1d); free(c->nu1d);
  free(c->t1); free(c->t2); free(c->t3);
  double **all[] = {&c->Radii, &c->Rinf, &c->Rsup, &c->Rmed, &c->Surf, &c->InvRmed,
                    &c->InvSurf, &c->InvDiffRmed, &c->InvDiffRsup, &c->InvRinf,
                    &c->GlobalRmed, &c->Azimuth, &c->CosAz, &c->SinAz, &c->SigmaMed,
                    &c->SigmaInf, &c->EnergyMed, &c->VthetaMed, &c->VradMed,
                    &c->GlobalCs, &c->AxiSG, &c->vmed, &c->vmoy, &c->rho, &c->vr, &c->vt,
                    &c->en, &c->cs, &c->pres, &c->temp, &c->pot, &c->vradint,
                    &c->vthetaint, &c->vradnew, &c->vthetanew, &c->energyint,
                    &c->energynew, &c->rhoint, &c->temperint, &c->divv, &c->drr,
                    &c->drp, &c->dpp, &c->trr, &c->trp, &c->tpp, &c->vischeat,
                    &c->rp, &c->rm, &c->tp, &c->tm, &c->rhostar, &c->work,
                    &c->qrstar, &c->vthetares, &c->elong, &c->dq, &c->tempshift,
                    &c->cabs, &c->cord, &c->sg_accr, &c->sg_acct, &c->gr,
                    &c->gtheta, &c->kr_re, &c->kr_im, &c->kt_re, &c->kt_im,
                    &c->indpot, &c->radgradp, &c->azigradp, &c->radindacc, &c->aziindacc,
                    &c->radsgacc, &c->azisgacc, &c->radfbacc, &c->azifbacc, &c->fbedot,
                    &c->dustdens, &c->d_size, &c->d_size_init, &c->d_r, &c->d_th, &c->d_vr,
                    &c->d_vth, &c->d_l, &c->d_vrg, &c->d_vtg, &c->d_densg, &c->d_csg,
                    &c->d_radgradp, &c->d_azigradp, &c->d_radind, &c->d_aziind, &c->d_radsg,
                    &c->d_azisg, &c->d_raddrag, &c->d_azidrag, &c->d_stokes, &c->AziInf,
                    &c->AziSup, &c->drho, &c->dvr, &c->dvt, &c->stokes, &c->dpres,
                    &c->dvradint, &c->dvthetaint, &c->dvradnew, &c->dvthetanew, &c->dqr,
                    &c->dqt, &c->ddrr, &c->ddrp, &c->ddpp, &c->ddivv, &c->dtrr, &c->dtrp,
                    &c->dtpp, &c->fdiffrp, &c->fdifftp, &c->dcs1, &c->DSigmaMed,
                    &c->DVthetaMed, &c->DVradMed, &c->GlobalDCs, &c->AxiSGd, &c->dvmed,
                    &c->dvmoy};
  for (size_t k = 0; k < sizeof(all) / sizeof(all[0]); k++) free(*all[k]);
  free(c->nshift);
  free(c->dnshift);
  free(c->nosplit);
  free(c);
}

void or_get_tables(or_ctx *c, double *radii, double *rmed, double *surf, double *azimuth) {
  if (radii) memcpy(radii, c->Radii, sizeof(double) * (c->nr + 1));
  if (rmed) memcpy(rmed, c->Rmed, sizeof(double) * c->nr);
  if (surf) memcpy(surf, c->Surf, sizeof(double) * c->nr);
  if (azimuth) memcpy(azimuth, c->Azimuth, sizeof(double) * c->ns);
}

/* ------------------------------------------------------------------------ */
/* host initial state                                                        */
/* ------------------------------------------------------------------------ */

/* InitGasDensity (src/Pframeforce.c:297-341), drand48 stream of srand48(seed) */
void or_init_gas_density(or_ctx *c, double *dens) {
  unsigned short xs[3];
  xs[0] = 0x330E;
  xs[1] = (unsigned short)(c->p.noise_seed & 0xFFFF);
  xs[2] = (unsigned short)((c->p.noise_seed >> 16) & 0xFFFF);
  memset(dens, 0, sizeof(double) * (size_t)(c->nr + 1) * c->ns);
  for (int i = 0; i < c->nr; i++)
    for (int j = 0; j < c->ns; j++) {
      int l = j + i * c->ns;
      dens[l] = c->SigmaMed[i];
      if (c->p.add_noise) {
        double randomnb = 2.0 * erand48(xs) - 1.0;
        dens[l] += 1e-3 * c->SigmaMed[i] * randomnb;
      }
    }
}

/* InitGasEnergy (src/Pframeforce.c:422-440) */
void or_init_gas_energy(or_ctx *c, double *energy) {
  memset(energy, 0, sizeof(double) * (size_t)(c->nr + 1) * c->ns);
  for (int i = 0; i < c->nr; i++)
    for (int j = 0; j < c->ns; j++) energy[j + i * c->ns] = energy_profile(&c->p, c->Rmed[i]);
}

/* InitPlanetarySystem (src/Psys.c:96-196); the .cfg columns are read with
 * %f, so each value is rounded to float first (src/Psys.c:118) */
void or_init_planetary_system(or_ctx *c, int nb, const double *dist_in,
                              const double *mass_in, const double *acc_in,
                              const int *feel_disk, const int *feel_others,
                              dfargo_planets *sys) {
  memset(sys, 0, sizeof(*sys));
  sys->nb = nb;
  for (int i = 0; i < nb; i++) {
    float dist = (float)dist_in[i], mass = (float)mass_in[i], accret = (float)acc_in[i];
    double massinvelocity;
    sys->mass[i] = (double)mass;
    sys->initial_mass[i] = 0.0;
    sys->final_mass[i] = sys->mass[i];
    if (c->p.mass_taper > 1e-3)
      massinvelocity = 0.0;
    else
      massinvelocity = sys->final_mass[i];
    sys->x[i] = (double)dist * (1.0 + c->p.eccentricity);
    sys->y[i] = 0.0;
    sys->vy[i] = (double)sqrt(G * (1.0 + massinvelocity) / dist) *
                 sqrt((1.0 - c->p.eccentricity) / (1.0 + c->p.eccentricity));
    sys->vx[i] = -0.0000000001 * sys->vy[i];
    sys->acc[i] = accret;
    sys->feel_disk[i] = feel_disk[i];
    sys->feel_others[i] = feel_others[i];
    sys->binary[i] = 0;
  }
}

/* ------------------------------------------------------------------------ */
/* equation of state                                                          */
/* ------------------------------------------------------------------------ */

/* ComputeSoundSpeed (src/SourceEuler.c:2323-2347) */
static void compute_sound_speed(or_ctx *c) {
  const dfargo_params *p = &c->p;
  for (int i = 0; i < c->nr; i++) {
    double cs_buf = 0.;
    if (!p->energy_equation)
      cs_buf = aspect_ratio(p, c->Rmed[i]) * sqrt(G * 1.0 / c->Rmed[i]) * pow(c->Rmed[i], p->flaring_index);
    for (int j = 0; j < c->ns; j++) {
      int l = i * c->ns + j;
      if (!p->energy_equation)
        c->cs[l] = cs_buf;
      else
        c->cs[l] = sqrt(p->adiabatic_index * (p->adiabatic_index - 1.0) * c->en[l] / c->rho[l]);
    }
  }
}

/* ComputePressureField (src/SourceEuler.c:2414-2434) */
static void compute_pressure(or_ctx *c) {
  for (int i = 0; i < c->nr; i++)
    for (int j = 0; j < c->ns; j++) {
      int l = i * c->ns + j;
      if (!c->p.energy_equation)
        c->pres[l] = c->rho[l] * c->cs[l] * c->cs[l];
      else
        c->pres[l] = (c->p.adiabatic_index - 1.0) * c->en[l];
    }
}

/* ComputeTemperatureField (src/SourceEuler.c:2436-2460), MU = R = 1 */
static void compute_temperature(or_ctx *c) {
  for (int i = 0; i < c->nr; i++)
    for (int j = 0; j < c->ns; j++) {
      int l = i * c->ns + j;
      if (!c->p.energy_equation)
        c->temp[l] = 1.0 / 1.0 * c->pres[l] / c->rho[l];
      else
        c->temp[l] = 1.0 / 1.0 * (c->p.adiabatic_index - 1.0) * c->en[l] / c->rho[l];
    }
}

/* ------------------------------------------------------------------------ */
/* viscosity                                                                  */
/* ------------------------------------------------------------------------ */

/* ComputeDivergenceVelocity (src/SourceEuler.c:2373-2411) */
static void compute_divergence(or_ctx *c, const double *vr, const double *vt) {
  int nr = c->nr, ns = c->ns;
  double dphi = (c->p.pmax - c->p.pmin) / (double)ns, invdphi = 1.0 / dphi;
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns, lip = l + ns, ljp = (j == ns - 1) ? i * ns : l + 1;
      c->divv[l] = (vr[lip] * c->Rsup[i] - vr[l] * c->Rinf[i]) * c->InvDiffRsup[i] * c->InvRmed[i];
      c->divv[l] += (vt[ljp] - vt[l]) * invdphi * c->InvRmed[i];
    }
}

/* ComputeViscousTerms (src/Viscosity.c:159-300) */
static void compute_viscous_terms(or_ctx *c, const double *vr, const double *vt) {
  int nr = c->nr, ns = c->ns;
  const double *rho = c->rho;
  double dphi = (c->p.pmax - c->p.pmin) / (double)ns, invdphi = 1.0 / dphi;
  double onethird = 1.0 / 3.0, viscosity;
  compute_divergence(c, vr, vt);
  if (c->p.alpha_viscosity != 0.0) make_1d_profile(c, c->cs, c->GlobalCs);
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns, lip = l + ns, ljp = (j == ns - 1) ? i * ns : l + 1;
      c->drr[l] = (vr[lip] - vr[l]) * c->InvDiffRsup[i];
      c->dpp[l] = (vt[ljp] - vt[l]) * invdphi * c->InvRmed[i] + 0.5 * (vr[lip] + vr[l]) * c->InvRmed[i];
    }
  for (int i = 1; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns, ljm = (j == 0) ? i * ns + ns - 1 : l - 1, lim = l - ns;
      c->drp[l] = 0.5 * (c->Rinf[i] * (vt[l] * c->InvRmed[i] - vt[lim] * c->InvRmed[i - 1]) * c->InvDiffRmed[i] +
                         (vr[l] - vr[ljm]) * invdphi * c->InvRinf[i]);
    }
  for (int i = 0; i < nr; i++) {
    viscosity = has_viscosity(&c->p) ? or_fviscosity(c, c->Rmed[i]) : 0.0;
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns;
      c->trr[l] = 2.0 * rho[l] * viscosity * (c->drr[l] - onethird * c->divv[l]);
      c->tpp[l] = 2.0 * rho[l] * viscosity * (c->dpp[l] - onethird * c->divv[l]);
    }
  }
  for (int i = 1; i < nr; i++) {
    viscosity = has_viscosity(&c->p) ? or_fviscosity(c, c->Rmed[i]) : 0.0;
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns, lim = l - ns, ljm = (j == 0) ? i * ns + ns - 1 : l - 1;
      int ljmim = ljm - ns;
      c->trp[l] = 2.0 * 0.25 * (rho[l] + rho[lim] + rho[ljm] + rho[ljmim]) * viscosity * c->drp[l];
    }
  }
}

/* UpdateVelocitiesWithViscosity (src/Viscosity.c:302-378) */
static void update_velocities_viscosity(or_ctx *c, double *vr, double *vt, double DeltaT) {
  int nr = c->nr, ns = c->ns;
  const double *rho = c->rho, *Trr = c->trr, *Trp = c->trp, *Tpp = c->tpp;
  double dphi = (c->p.pmax - c->p.pmin) / (double)ns, invdphi = 1.0 / dphi;
  for (int i = 1; i < nr - 1; i++)
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns, lip = l + ns, ljm = (j == 0) ? i * ns + ns - 1 : l - 1;
      vt[l] += DeltaT * c->InvRmed[i] *
               ((c->Rsup[i] * Trp[lip] - c->Rinf[i] * Trp[l]) * c->InvDiffRsup[i] +
                (Tpp[l] - Tpp[ljm]) * invdphi + 0.5 * (Trp[l] + Trp[lip])) /
               (0.5 * (rho[l] + rho[ljm]));
    }
  for (int i = 1; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns, lim = l - ns, ljp = (j == ns - 1) ? i * ns : l + 1;
      vr[l] += DeltaT * c->InvRinf[i] *
               ((c->Rmed[i] * Trr[l] - c->Rmed[i - 1] * Trr[lim]) * c->InvDiffRmed[i] +
                (Trp[ljp] - Trp[l]) * invdphi - 0.5 * (Tpp[l] + Tpp[lim])) /
               (0.5 * (rho[l] + rho[lim]));
    }
}

void or_viscous(or_ctx *c, double dt) {
  compute_viscous_terms(c, c->vr, c->vt);
  update_velocities_viscosity(c, c->vr, c->vt, dt);
}

/* ComputeViscousHeating (src/SourceEuler.c:1554-1606) */
static void compute_viscous_heating(or_ctx *c) {
  int nr = c->nr, ns = c->ns;
  const double *dens = c->rho;
  double viscosity = 0.0;
  for (int i = 1; i < nr; i++) {
    if (has_viscosity(&c->p)) viscosity = or_fviscosity(c, c->Rmed[i]);
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns;
      if (viscosity != 0.0) {
        c->vischeat[l] = 0.5 / viscosity / dens[l] *
                         (c->trr[l] * c->trr[l] + 2.0 * c->trp[l] * c->trp[l] + c->tpp[l] * c->tpp[l]);
        c->vischeat[l] += (2.0 / 9.0) * viscosity * dens[l] * c->divv[l] * c->divv[l];
      } else
        c->vischeat[l] = 0.0;
    }
  }
  double r = c->Rmed[0], rip = c->Rmed[1], ri2p = c->Rmed[2];
  for (int j = 0; j < ns; j++) {
    int l = j, lip = l + ns, li2p = lip + ns;
    double qpip = c->vischeat[lip], qpi2p = c->vischeat[li2p];
    if (viscosity != 0.0)
      c->vischeat[l] = qpip * exp(log(qpip / qpi2p) * log(r / rip) / log(rip / ri2p));
    else
      c->vischeat[l] = 0.0;
  }
}

/* ------------------------------------------------------------------------ */
/* forces, potential, planets                                                */
/* ------------------------------------------------------------------------ */

/* ComputeForce (src/Force.c:77-253) with dimfxy = 2; gf as GlobalForce */
static void compute_force(or_ctx *c, const double *dens, double x, double y,
                          double rsmoothing, double mass, double gf[8]) {
  int nr = c->nr, ns = c->ns;
  double fxi[2] = {0., 0.}, fxo[2] = {0., 0.}, fyi[2] = {0., 0.}, fyo[2] = {0., 0.};
  double a = sqrt(x * x + y * y);
  double rh = pow(mass / 3., 1. / 3.) * a + 1e-15;
  int binary = (c->sys.nb > 1) && c->sys.binary[0];
  double xb = 0., yb = 0.;
  if (binary) {
    double m0 = c->sys.mass[0], m1 = c->sys.mass[1];
    xb = (m0 * c->sys.x[0] + m1 * c->sys.x[1]) / (m0 + m1);
    yb = (m0 * c->sys.y[0] + m1 * c->sys.y[1]) / (m0 + m1);
  }
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns;
      double xc = c->cabs[l], yc = c->cord[l];
      double cellmass = c->Surf[i] * dens[l];
      double dx = xc - x, dy = yc - y;
      double d2 = dx * dx + dy * dy;
      double planet_distance = sqrt(d2);
      double dist2 = d2 + rsmoothing * rsmoothing;
      double distance = sqrt(dist2);
      double InvDist3 = 1.0 / dist2 / distance;
      if (binary) {
        d2 = (xc - xb) * (xc - xb) + (yc - yb) * (yc - yb);
        planet_distance = sqrt(d2);
      }
      for (int k = 0; k < 2; k++) {
        double hill_cut;
        if (k != 0) {
          double cutoff = c->p.exclude_hill_factor * rh;
          if (planet_distance / cutoff < 0.5)
            hill_cut = 0.0;
          else if (planet_distance > cutoff)
            hill_cut = 1.0;
          else
            hill_cut = pow(sin((planet_distance / cutoff - .5) * M_PI), 2.);
        } else
          hill_cut = 1.;
        if (c->Rmed[i] < a) {
          fxi[k] += G * cellmass * dx * InvDist3 * hill_cut;
          fyi[k] += G * cellmass * dy * InvDist3 * hill_cut;
        } else {
          fxo[k] += G * cellmass * dx * InvDist3 * hill_cut;
          fyo[k] += G * cellmass * dy * InvDist3 * hill_cut;
        }
      }
    }
  gf[0] = fxi[0]; gf[1] = fxi[1]; gf[2] = fxo[0]; gf[3] = fxo[1];
  gf[4] = fyi[0]; gf[5] = fyi[1]; gf[6] = fyo[0]; gf[7] = fyo[1];
}

/* ComputeAccel (src/SideEuler.c:282-298) */
static pair_t compute_accel(or_ctx *c, const double *dens, double x, double y,
                            double rsmoothing, double mass) {
  double gf[8];
  pair_t a;
  compute_force(c, dens, x, y, rsmoothing, mass, gf);
  if (c->p.exclude_hill) {
    a.x = gf[1] + gf[3];
    a.y = gf[5] + gf[7];
  } else {
    a.x = gf[0] + gf[2];
    a.y = gf[4] + gf[6];
  }
  return a;
}

/* GasTotalMass, ComputeBarycenterPosition, ComputeOppositeAcceleration
 * OfBarycenter (src/SideEuler.c:15-112) */
static pair_t opposite_acc_barycenter(or_ctx *c, double timestep) {
  int nr = c->nr, ns = c->ns;
  double total = 0.0, lxb = 0.0, lyb = 0.0;
  pair_t acc;
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) total += c->Surf[i] * c->rho[j + i * ns];
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      int l = i * ns + j;
      lxb += c->Surf[i] * c->rho[l] * c->cabs[l];
      lyb += c->Surf[i] * c->rho[l] * c->cord[l];
    }
  c->barx[2] = lxb / (1.0 + total);
  c->bary[2] = lyb / (1.0 + total);
  acc.x = -(c->barx[2] - 2.0 * c->barx[1] + c->barx[0]) / (timestep * timestep);
  acc.y = -(c->bary[2] - 2.0 * c->bary[1] + c->bary[0]) / (timestep * timestep);
  c->barx[0] = c->barx[1]; c->barx[1] = c->barx[2];
  c->bary[0] = c->bary[1]; c->bary[1] = c->bary[2];
  return acc;
}

/* ComputeIndirectTerm + FillForcesArrays (src/Pframeforce.c:25-129) */
static void fill_forces_arrays(or_ctx *c) {
  const dfargo_params *p = &c->p;
  int nr = c->nr, ns = c->ns, nb = c->sys.nb;
  double planetmass_taper;
  if (p->mass_taper > 1e-2) {
    double mt = (c->PhysicalTime - c->PhysicalTimeInitial) / (p->mass_taper * 2.0 * M_PI);
    planetmass_taper = (mt > 1.0 ? 1.0 : .5 * (1.0 - cos(M_PI * mt)));
  } else
    planetmass_taper = 1.0;
  c->indirect.x = -c->dop.x;
  c->indirect.y = -c->dop.y;
  if (p->alternative_gas_indirect_term) {
    c->indirect.x = -c->dop_alt.x;
    c->indirect.y = -c->dop_alt.y;
  }
  if (!p->indirect_term) {
    c->indirect.x = 0.0;
    c->indirect.y = 0.0;
  }
  for (int l = 0; l < (nr + 1) * ns; l++) c->pot[l] = 0.0;
  for (int l = 0; l < (nr + 1) * ns; l++) c->indpot[l] = 0.0;
  const int dust_planets = c->dust_on && c->dp.dust_feel_planets, dust_disk = c->dust_on && c->dp.dust_feel_disk;
  for (int k = 0; k < nb; k++) {
    double xplanet = c->sys.x[k], yplanet = c->sys.y[k];
    double mplanet = c->sys.mass[k] * planetmass_taper;
    double PlanetDistance = sqrt(xplanet * xplanet + yplanet * yplanet);
    double InvPlanetDistance3 = 1.0 / PlanetDistance / PlanetDistance / PlanetDistance;
    double RRoche = PlanetDistance * pow((1.0 / 3.0 * mplanet), 1.0 / 3.0);
    double smoothing = (p->roche_smoothing != 0.0) ? RRoche * p->roche_smoothing
                                                     : compute_smoothing(p, PlanetDistance);
    double smooth = smoothing * smoothing;
    for (int i = 0; i < nr; i++)
      for (int j = 0; j < ns; j++) {
        int l = j + i * ns;
        double x = c->Rmed[i] * c->CosAz[j], y = c->Rmed[i] * c->SinAz[j];
        double distance = (x - xplanet) * (x - xplanet) + (y - yplanet) * (y - yplanet);
        double distancesmooth = sqrt(distance + smooth);
        double pot = -G * mplanet / distancesmooth;
        if (p->indirect_term) {
          pot += G * mplanet * InvPlanetDistance3 * (x * xplanet + y * yplanet);
          /* assigned, not summed, as in the reference */
          if (dust_planets) c->indpot[l] = G * mplanet * InvPlanetDistance3 * (x * xplanet + y * yplanet);
        }
        c->pot[l] += pot;
      }
  }
  for (int i = 0; i < nr; i++) {
    double InvDistance = 1.0 / c->Rmed[i];
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns;
      double x = c->Rmed[i] * c->CosAz[j], y = c->Rmed[i] * c->SinAz[j];
      double pot = -G * 1.0 * InvDistance;
      if ((!p->discard_gas_indirect_term) && (fabs(p->pmax - p->pmin - 2. * M_PI) < 0.01)) {
        pot -= c->indirect.x * x + c->indirect.y * y;
        if (dust_disk) c->indpot[l] -= c->indirect.x * x + c->indirect.y * y;
      }
      c->pot[l] += pot;
    }
  }
}

/* AdvanceSystemFromDisk (src/Pframeforce.c:131-161) */
static void advance_system_from_disk(or_ctx *c, double dt) {
  for (int k = 0; k < c->sys.nb; k++) {
    if (!c->sys.feel_disk[k]) continue;
    double m = c->sys.mass[k], x = c->sys.x[k], y = c->sys.y[k];
    double r = sqrt(x * x + y * y), smoothing;
    if (c->p.roche_smoothing != 0.0)
      smoothing = r * pow(m / 3., 1. / 3.) * c->p.roche_smoothing;
    else
      smoothing = compute_smoothing(&c->p, r);
    pair_t gamma = compute_accel(c, c->rho, x, y, smoothing, m);
    c->sys.vx[k] += dt * gamma.x;
    c->sys.vy[k] += dt * gamma.y;
    c->sys.vx[k] += dt * c->indirect.x;
    c->sys.vy[k] += dt * c->indirect.y;
  }
}

/* DerivMotionRK5 (src/RungeKunta.c:14-53) */
static void deriv_motion_rk5(const double *q, const double *masses, double *deriv,
                             int n, double dt, const int *feelothers, int indirect) {
  const double *x = q, *y = q + n, *vx = y + n, *vy = vx + n;
  double *dx = deriv, *dy = deriv + n, *dvx = dy + n, *dvy = dvx + n;
  double Dist[MAXPL];
  for (int i = 0; i < n; i++) Dist[i] = sqrt(x[i] * x[i] + y[i] * y[i]);
  for (int i = 0; i < n; i++) {
    dx[i] = vx[i];
    dy[i] = vy[i];
    dvx[i] = -G * 1.0 / Dist[i] / Dist[i] / Dist[i] * x[i];
    dvy[i] = -G * 1.0 / Dist[i] / Dist[i] / Dist[i] * y[i];
    for (int j = 0; j < n; j++) {
      if (indirect) {
        dvx[i] -= G * masses[j] / Dist[j] / Dist[j] / Dist[j] * x[j];
        dvy[i] -= G * masses[j] / Dist[j] / Dist[j] / Dist[j] * y[j];
      }
      if ((j != i) && feelothers[i]) {
        double dist = (x[i] - x[j]) * (x[i] - x[j]) + (y[i] - y[j]) * (y[i] - y[j]);
        dist = sqrt(dist);
        dvx[i] += G * masses[j] / dist / dist / dist * (x[j] - x[i]);
        dvy[i] += G * masses[j] / dist / dist / dist * (y[j] - y[i]);
      }
    }
  }
  for (int i = 0; i < 4 * n; i++) deriv[i] *= dt;
}

static void translate_rk5(const double *qold, const double *k1, const double *k2,
                          const double *k3, const double *k4, const double *k5,
                          double c1, double c2, double c3, double c4, double c5,
                          double *qnew, int n) {
  for (int i = 0; i < 4 * n; i++)
    qnew[i] = qold[i] + c1 * k1[i] + c2 * k2[i] + c3 * k3[i] + c4 * k4[i] + c5 * k5[i];
}

/* RungeKunta (src/RungeKunta.c:64-91), Cash-Karp tableau */
void or_rk5(const double *q0, double dt, const double *masses, double *q1, int n,
            const int *feelothers, int indirect) {
  double k1[4 * MAXPL], k2[4 * MAXPL], k3[4 * MAXPL], k4[4 * MAXPL], k5[4 * MAXPL], k6[4 * MAXPL];
  memset(k1, 0, sizeof k1); memset(k2, 0, sizeof k2); memset(k3, 0, sizeof k3);
  memset(k4, 0, sizeof k4); memset(k5, 0, sizeof k5); memset(k6, 0, sizeof k6);
  deriv_motion_rk5(q0, masses, k1, n, dt, feelothers, indirect);
  translate_rk5(q0, k1, k2, k3, k4, k5, 0.2, 0.0, 0.0, 0.0, 0.0, q1, n);
  deriv_motion_rk5(q1, masses, k2, n, dt, feelothers, indirect);
  translate_rk5(q0, k1, k2, k3, k4, k5, 0.075, 0.225, 0.0, 0.0, 0.0, q1, n);
  deriv_motion_rk5(q1, masses, k3, n, dt, feelothers, indirect);
  translate_rk5(q0, k1, k2, k3, k4, k5, 0.3, -0.9, 1.2, 0.0, 0.0, q1, n);
  deriv_motion_rk5(q1, masses, k4, n, dt, feelothers, indirect);
  translate_rk5(q0, k1, k2, k3, k4, k5, -11.0 / 54.0, 2.5, -70.0 / 27.0, 35.0 / 27.0, 0.0, q1, n);
  deriv_motion_rk5(q1, masses, k5, n, dt, feelothers, indirect);
  translate_rk5(q0, k1, k2, k3, k4, k5, 1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0,
                44275.0 / 110592.0, 253.0 / 4096.0, q1, n);
  deriv_motion_rk5(q1, masses, k6, n, dt, feelothers, indirect);
  for (int i = 0; i < 4 * n; i++)
    q1[i] = q0[i] + 37.0 / 378.0 * k1[i] + 250.0 / 621.0 * k3[i] + 125.0 / 594.0 * k4[i] +
            512.0 / 1771.0 * k6[i];
}

/* AdvanceSystemRK5 (src/Pframeforce.c:163-266) without ForcedInnerCircular */
static void advance_system_rk5(or_ctx *c, double dt) {
  dfargo_planets *s = &c->sys;
  int n = s->nb;
  double q0[4 * MAXPL], q1[4 * MAXPL], masses[MAXPL];
  if (n == 0) return;
  if (!c->p.forced_circular) {
    for (int i = 0; i < n; i++) {
      q0[i] = s->x[i]; q0[i + n] = s->y[i]; q0[i + 2 * n] = s->vx[i]; q0[i + 3 * n] = s->vy[i];
      masses[i] = s->mass[i];
    }
    or_rk5(q0, dt, masses, q1, n, s->feel_others, c->p.indirect_term);
  }
  for (int i = 1 - (c->PhysicalTime >= c->p.release_date); i < n; i++) {
    if (!c->p.forced_circular) {
      s->x[i] = q1[i]; s->y[i] = q1[i + n]; s->vx[i] = q1[i + 2 * n]; s->vy[i] = q1[i + 3 * n];
    } else {
      double x = s->x[i], y = s->y[i], theta = atan2(y, x), vx = s->vx[i], vy = s->vy[i];
      double r = sqrt(x * x + y * y), v = sqrt(vx * vx + vy * vy);
      double omega = (-y * vx + x * vy) / r / r, dtheta = omega * dt;
      s->x[i] = r * cos(theta + dtheta);
      s->y[i] = r * sin(theta + dtheta);
      s->vx[i] = -v * sin(theta + dtheta);
      s->vy[i] = v * cos(theta + dtheta);
    }
  }
  if (c->PhysicalTime < c->p.release_date) {
    double x = s->x[0], y = s->y[0], r = sqrt(x * x + y * y), theta = atan2(y, x);
    double rdot = (c->p.release_radius - r) / (c->p.release_date - c->PhysicalTime);
    double omega = sqrt((1. + s->mass[0]) / r / r / r), new_r = r + rdot * dt, denom = r - new_r;
    double dtheta = (denom != 0.0) ? 2. * dt * r * omega / denom * (sqrt(r / new_r) - 1.) : omega * dt;
    double vx = rdot, vy = new_r * sqrt((1. + s->mass[0]) / new_r / new_r / new_r);
    s->x[0] = new_r * cos(dtheta + theta);
    s->y[0] = new_r * sin(dtheta + theta);
    s->vx[0] = vx * cos(dtheta + theta) - vy * sin(dtheta + theta);
    s->vy[0] = vx * sin(dtheta + theta) + vy * cos(dtheta + theta);
  }
}

/* RotatePsys (src/Psys.c:352-370) */
static void rotate_psys(or_ctx *c, double angle) {
  double sint = sin(angle), cost = cos(angle);
  for (int i = 0; i < c->sys.nb; i++) {
    double xt = c->sys.x[i], yt = c->sys.y[i];
    c->sys.x[i] = xt * cost + yt * sint;
    c->sys.y[i] = -xt * sint + yt * cost;
    xt = c->sys.vx[i];
    yt = c->sys.vy[i];
    c->sys.vx[i] = xt * cost + yt * sint;
    c->sys.vy[i] = -xt * sint + yt * cost;
  }
}

/* GetPsysInfo (src/Psys.c:233-350), actions MARK (get = 0) and GET
 * (get = 1).  The frame's centre is planet 0 (the outermost planet with
 * COROTATEWITHOUTERPLANET) or, with FRAME B, the barycentre of planets 0 and
 * 1, moved with FRAME G to the guiding centre of its orbit unless e < 1e-8.
 * GET returns the angle the centre turned through since the mark.  The
 * orbital frequency the reference also forms (with its self-gravity
 * correction) serves only the FREQUENCY action and the guiding centre's
 * velocity, which MARK and GET do not use. */
static double psys_info(or_ctx *c, int get) {
  const dfargo_planets *s = &c->sys;
  const int k = c->p.corotate_with_outer_planet ? s->nb - 1 : 0;
  double x = s->x[k], y = s->y[k], vx = s->vx[k], vy = s->vy[k], m = s->mass[k] + 1.;
  if (c->p.corotating == DFARGO_FRAME_BINARY_CENTER) {
    double m0 = s->mass[0], m1 = s->mass[1];
    if (m0 == 0.0) m0 = s->final_mass[0];
    if (m1 == 0.0) m1 = s->final_mass[1];
    x = (m0 * s->x[0] + m1 * s->x[1]) / (m0 + m1);
    y = (m0 * s->y[0] + m1 * s->y[1]) / (m0 + m1);
    vx = (m0 * s->vx[0] + m1 * s->vx[1]) / (m0 + m1);
    vy = (m0 * s->vy[0] + m1 * s->vy[1]) / (m0 + m1);
    m = m0 + m1 + 1.;
  }
  double xc = x, yc = y;
  if (c->p.corotating == DFARGO_FRAME_GUIDING_CENTER) {
    const double h = x * vy - y * vx, d = sqrt(x * x + y * y);
    const double Ax = x * vy * vy - y * vx * vy - G * m * x / d;
    const double Ay = y * vx * vx - x * vx * vy - G * m * y / d;
    const double e = sqrt(Ax * Ax + Ay * Ay) / m, a = h * h / G / m / (1. - e * e);
    const double arg = (e == 0.0) ? 1.0 : (1.0 - d / a) / e;
    double E = (fabs(arg) >= 1.0) ? M_PI * (1. - arg / fabs(arg)) / 2. : acos((1.0 - d / a) / e);
    if ((x * y * (vy * vy - vx * vx) + vx * vy * (x * x - y * y)) < 0) E = -E;
    const double M = E - e * sin(E), pa = atan2(Ay, Ax);
    if (!(e < 1e-8)) {
      xc = a * cos(M + pa);
      yc = a * sin(M + pa);
    }
  }
  if (!get) {
    c->Xplanet = xc;
    c->Yplanet = yc;
    return 0.0;
  }
  const double d2 = sqrt(xc * xc + yc * yc), d1 = sqrt(c->Xplanet * c->Xplanet + c->Yplanet * c->Yplanet);
  const double cross = c->Xplanet * yc - xc * c->Yplanet;
  c->Xplanet = xc;
  c->Yplanet = yc;
  return asin(cross / (d1 * d2));
}

/* AccreteOntoPlanets (src/Planet.c:11-126), one process (no ghost rings):
 * gas within 0.75 and 0.45 of the Roche radius loses the fractions dt acc/3
 * and then 2 dt acc/3 of its density; the planet gains that mass, and the
 * momentum it carried at the cell-centred velocity when it feels the disc. */
static void accrete_onto_planets(or_ctx *c, double dt) {
  const dfargo_params *p = &c->p;
  const int nr = c->nr, ns = c->ns;
  double *dens = c->rho, *vrad = c->vr, *vtheta = c->vt;
  dfargo_planets *sys = &c->sys;
  for (int k = 0; k < sys->nb; k++) {
    if (!(sys->acc[k] > 1e-10)) continue;
    double dMplanet = 0.0, dPxPlanet = 0.0, dPyPlanet = 0.0;
    const double facc = dt * sys->acc[k], facc1 = 1.0 / 3.0 * facc, facc2 = 2.0 / 3.0 * facc;
    const double frac1 = 0.75, frac2 = 0.45;
    const double Xplanet = sys->x[k], Yplanet = sys->y[k];
    const double VXplanet = sys->vx[k], VYplanet = sys->vy[k];
    double Mplanet = sys->mass[k];
    const double Rplanet = sqrt(Xplanet * Xplanet + Yplanet * Yplanet);
    const double RRoche = pow((1.0 / 3.0 * Mplanet), (1.0 / 3.0)) * Rplanet;
    int i_min = 0, i_max = nr - 1;
    while ((c->Rsup[i_min] < Rplanet - RRoche) && (i_min < nr)) i_min++;
    while ((c->Rinf[i_max] > Rplanet + RRoche) && (i_max > 0)) i_max--;
    const double angle = atan2(Yplanet, Xplanet);
    const int j_min = (int)((double)ns / (p->pmax - p->pmin) * (angle - 2.0 * RRoche / Rplanet - p->pmin));
    const int j_max = (int)((double)ns / (p->pmax - p->pmin) * (angle + 2.0 * RRoche / Rplanet - p->pmin));
    double PxPlanet = Mplanet * VXplanet, PyPlanet = Mplanet * VYplanet;
    for (int i = i_min; i <= i_max; i++)
      for (int j = j_min; j <= j_max; j++) {
        int jf = j;
        while (jf < 0) jf += ns;
        while (jf >= ns) jf -= ns;
        const int l = jf + i * ns, lip = l + ns;
        const int ljp = (jf == ns - 1) ? i * ns : l + 1;
        const double xc = c->cabs[l], yc = c->cord[l];
        const double dx = Xplanet - xc, dy = Yplanet - yc;
        const double distance = sqrt(dx * dx + dy * dy);
        const double vtcell = 0.5 * (vtheta[l] + vtheta[ljp]) + c->Rmed[i] * c->OmegaFrame;
        const double vrcell = 0.5 * (vrad[l] + vrad[lip]);
        const double vxcell = (vrcell * xc - vtcell * yc) / c->Rmed[i];
        const double vycell = (vrcell * yc + vtcell * xc) / c->Rmed[i];
        if (distance < frac1 * RRoche) {
          const double deltaM = facc1 * dens[l] * c->Surf[i];
          dens[l] *= (1.0 - facc1);
          dPxPlanet += deltaM * vxcell;
          dPyPlanet += deltaM * vycell;
          dMplanet += deltaM;
        }
        if (distance < frac2 * RRoche) {
          const double deltaM = facc2 * dens[l] * c->Surf[i];
          dens[l] *= (1.0 - facc2);
          dPxPlanet += deltaM * vxcell;
          dPyPlanet += deltaM * vycell;
          dMplanet += deltaM;
        }
      }
    PxPlanet += dPxPlanet;
    PyPlanet += dPyPlanet;
    Mplanet += dMplanet;
    if (sys->feel_disk[k]) {
      sys->vx[k] = PxPlanet / Mplanet;
      sys->vy[k] = PyPlanet / Mplanet;
    }
    sys->mass[k] = Mplanet;
  }
}

void or_accrete_onto_planets(or_ctx *c, double dt) { accrete_onto_planets(c, dt); }

/* InitializeOneDViscousEvolution (src/SideEuler.c:1444-1487), without
 * the dims1D/used_rad1D files; its restart read is or_set_visc1d */
static void init_1d_viscous(or_ctx *c) {
  const dfargo_params *p = &c->p;
  const int n = p->nrad1d;
  double *radii = (double *)calloc(n + 1, sizeof(double));
  c->n1d = n;
  c->rmed1d = (double *)calloc(n, sizeof(double));
  c->sigma1d = (double *)calloc(n, sizeof(double));
  c->vrad1d = (double *)calloc(n, sizeof(double));
  c->nu1d = (double *)calloc(n, sizeof(double));
  c->t1 = (double *)calloc(n, sizeof(double));
  c->t2 = (double *)calloc(n, sizeof(double));
  c->t3 = (double *)calloc(n, sizeof(double));
  for (int i = 0; i <= n; i++) radii[i] = p->rmin1d * exp((double)i / (double)n * log(p->rmax1d / p->rmin1d));
  for (int i = 0; i < n; i++) {
    c->rmed1d[i] = 2.0 / 3.0 * (radii[i + 1] * radii[i + 1] * radii[i + 1] - radii[i] * radii[i] * radii[i]);
    c->rmed1d[i] = c->rmed1d[i] / (radii[i + 1] * radii[i + 1] - radii[i] * radii[i]);
    c->sigma1d[i] = sigma_profile(p, c->rmed1d[i]);
    const double cs = aspect_ratio(p, c->rmed1d[i]) / sqrt(c->rmed1d[i]) * pow(c->rmed1d[i], p->flaring_index);
    if (p->viscosity != 0.0) c->nu1d[i] = p->viscosity;
    if (p->alpha_viscosity != 0.0) {
      c->nu1d[i] = p->alpha_viscosity * cs * cs * pow(c->rmed1d[i], 1.5);
      c->vrad1d[i] = -3.0 * c->nu1d[i] / c->rmed1d[i] * (-p->sigma_slope + 2.0 * p->flaring_index + 1.0);
    } else
      c->vrad1d[i] = -3.0 * c->nu1d[i] / c->rmed1d[i] * (-p->sigma_slope + .5);
  }
  free(radii);
}

/* SolveOneDViscousEvolution (src/SideEuler.c:1518-1565) */
static void solve_1d_viscous(or_ctx *c, double dt) {
  const int n = c->n1d;
  double *R = c->rmed1d, *S = c->sigma1d, *V = c->vrad1d, *nu = c->nu1d;
  for (int i = 1; i <= n - 2; i++) {
    const double drm = R[i] - R[i - 1], drp = R[i + 1] - R[i];
    const double drnu = (drm * drm * nu[i + 1] - drp * drp * nu[i - 1] + (drp * drp - drm * drm) * nu[i]) /
                        (drm * drm * drp + drp * drp * drm);
    const double drsigma = (drm * drm * S[i + 1] - drp * drp * S[i - 1] + (drp * drp - drm * drm) * S[i]) /
                           (drm * drm * drp + drp * drp * drm);
    const double d2rnu = (drm * nu[i + 1] + drp * nu[i - 1] - (drm + drp) * nu[i]) /
                         (0.5 * drm * drp * drp + 0.5 * drp * drm * drm);
    const double d2rsigma = (drm * S[i + 1] + drp * S[i - 1] - (drm + drp) * S[i]) /
                            (0.5 * drm * drp * drp + 0.5 * drp * drm * drm);
    c->t1[i] = 3.0 * S[i] * (1.5 * drnu / R[i] + d2rnu);
    c->t2[i] = 3.0 * drsigma * (1.5 * nu[i] / R[i] + 2.0 * drnu);
    c->t3[i] = 3.0 * nu[i] * d2rsigma;
    V[i] = -1.5 * nu[i] / R[i] - 3.0 * drnu - 1.5 * (nu[i] / S[i]) * drsigma;
  }
  for (int i = 1; i <= n - 2; i++) S[i] += dt * (c->t1[i] + c->t2[i] + c->t3[i]);
  if (c->p.bc1d == 1) {
    S[0] = S[1] * (nu[1] / nu[0]) * sqrt(R[1] / R[0]);
    S[n - 1] = S[n - 2] * (nu[n - 2] / nu[n - 1]) * sqrt(R[n - 2] / R[n - 1]);
  }
  if (c->p.bc1d == 2) {
    S[0] = S[1] * (nu[1] / nu[0]);
    S[n - 1] = S[n - 2] * (nu[n - 2] / nu[n - 1]);
  }
  if (c->p.bc1d == 0) {
    S[0] = 1e-9;
    S[n - 1] = 1e-9;
  }
  V[0] = V[1];
  V[n - 1] = V[n - 2];
}

void or_solve_1d_viscous(or_ctx *c, double dt, double *sigma, double *vrad) {
  if (!c->n1d) init_1d_viscous(c);
  if (dt > 0.0) solve_1d_viscous(c, dt);
  memcpy(sigma, c->sigma1d, sizeof(double) * c->n1d);
  memcpy(vrad, c->vrad1d, sizeof(double) * c->n1d);
}

/* The restart branch of InitializeOneDViscousEvolution
 * (src/SideEuler.c:1489-1502) overwrites Rmed1D and Sigma1D with the values
 * it reads, after viscosity1D was computed from the t = 0 cell centres; a
 * NULL pointer keeps that profile. */
void or_set_visc1d(or_ctx *c, const double *rmed, const double *sigma, const double *vrad) {
  if (!c->n1d) init_1d_viscous(c);
  const size_t b = sizeof(double) * c->n1d;
  if (rmed) memcpy(c->rmed1d, rmed, b);
  if (sigma) memcpy(c->sigma1d, sigma, b);
  if (vrad) memcpy(c->vrad1d, vrad, b);
}

void or_get_visc1d(or_ctx *c, double *rmed, double *sigma, double *vrad) {
  if (!c->n1d) init_1d_viscous(c);
  const size_t b = sizeof(double) * c->n1d;
  if (rmed) memcpy(rmed, c->rmed1d, b);
  if (sigma) memcpy(sigma, c->sigma1d, b);
  if (vrad) memcpy(vrad, c->vrad1d, b);
}

/* CorrectVtheta (src/SideEuler.c:1133-1148) */
static void correct_vtheta(or_ctx *c, double domega) {
  for (int i = 0; i < c->nr; i++)
    for (int j = 0; j < c->ns; j++) c->vt[i * c->ns + j] -= domega * c->Rmed[i];
  if (c->df_on) /* CorrectVtheta (DVtheta, domega), src/SourceEuler.c:521 */
    for (int i = 0; i < c->nr; i++)
      for (int j = 0; j < c->ns; j++) c->dvt[i * c->ns + j] -= domega * c->Rmed[i];
}

/* ------------------------------------------------------------------------ */
/* self-gravity                                                               */
/* ------------------------------------------------------------------------ */

/* Mixed-radix (2, 3, else naive) recursive DFT, unnormalized,
 * y[k] = sum_n x[n] exp(sign*2*pi*i*k*n/N).  Pinned against a naive DFT. */
static void fft_rec(const double *xr, const double *xi, int stride, double *yr,
                    double *yi, int n, int sign, const double *twr, const double *twi,
                    int tstride) {
  if (n == 1) {
    yr[0] = xr[0];
    yi[0] = xi[0];
    return;
  }
  int p = (n % 2 == 0) ? 2 : ((n % 3 == 0) ? 3 : n);
  int m = n / p;
  if (p == n) { /* naive */
    for (int k = 0; k < n; k++) {
      double sr = 0., si = 0.;
      for (int t = 0; t < n; t++) {
        long idx = ((long)k * t) % n * tstride;
        double wr = twr[idx], wi = sign * twi[idx];
        sr += xr[t * stride] * wr - xi[t * stride] * wi;
        si += xr[t * stride] * wi + xi[t * stride] * wr;
      }
      yr[k] = sr;
      yi[k] = si;
    }
    return;
  }
  double *tr = (double *)malloc(sizeof(double) * n), *ti = (double *)malloc(sizeof(double) * n);
  for (int r = 0; r < p; r++)
    fft_rec(xr + r * stride, xi + r * stride, stride * p, tr + r * m, ti + r * m, m, sign, twr,
            twi, tstride * p);
  for (int k = 0; k < m; k++)
    for (int q = 0; q < p; q++) {
      int kk = k + q * m;
      double sr = 0., si = 0.;
      for (int r = 0; r < p; r++) {
        long idx = ((long)r * kk) % n * tstride;
        double wr = twr[idx], wi = sign * twi[idx];
        sr += tr[r * m + k] * wr - ti[r * m + k] * wi;
        si += tr[r * m + k] * wi + ti[r * m + k] * wr;
      }
      yr[kk] = sr;
      yi[kk] = si;
    }
  free(tr);
  free(ti);
}

void or_fft(double *re, double *im, int n, int sign) {
  double *twr = (double *)malloc(sizeof(double) * n), *twi = (double *)malloc(sizeof(double) * n);
  double *yr = (double *)malloc(sizeof(double) * n), *yi = (double *)malloc(sizeof(double) * n);
  for (int k = 0; k < n; k++) {
    twr[k] = cos(2.0 * M_PI * (double)k / (double)n);
    twi[k] = sin(2.0 * M_PI * (double)k / (double)n);
  }
  fft_rec(re, im, 1, yr, yi, n, sign, twr, twi, 1);
  memcpy(re, yr, sizeof(double) * n);
  memcpy(im, yi, sizeof(double) * n);
  free(twr); free(twi); free(yr); free(yi);
}

/* 2-D complex transform of an nx-by-ny row-major array */
static void fft2d(double *re, double *im, int nx, int ny, int sign) {
  int mx = nx > ny ? nx : ny;
  double *br = (double *)malloc(sizeof(double) * mx), *bi = (double *)malloc(sizeof(double) * mx);
  for (int i = 0; i < nx; i++) or_fft(re + (size_t)i * ny, im + (size_t)i * ny, ny, sign);
  for (int j = 0; j < ny; j++) {
    for (int i = 0; i < nx; i++) { br[i] = re[(size_t)i * ny + j]; bi[i] = im[(size_t)i * ny + j]; }
    or_fft(br, bi, nx, sign);
    for (int i = 0; i < nx; i++) { re[(size_t)i * ny + j] = br[i]; im[(size_t)i * ny + j] = bi[i]; }
  }
  free(br);
  free(bi);
}

/* init_sg + compute_fftkernel (src/sginit.c:3-13, src/sgkernel.c:3-27):
 * kernels on the (2 Nrad) x Nsec log-polar mesh, then their spectra */
static void sg_init(or_ctx *c) {
  int nr = c->nr, ns = c->ns, n2 = 2 * nr;
  size_t tot = (size_t)n2 * ns;
  c->sg_eps = c->p.sg_thickness_smoothing * c->p.aspect_ratio;
  c->sg_rstep = log(c->Radii[nr] / c->Radii[0]) / (double)nr;
  c->sg_tstep = (c->p.pmax - c->p.pmin) / (double)ns;
  c->kr_re = (double *)calloc(tot, sizeof(double)); c->kr_im = (double *)calloc(tot, sizeof(double));
  c->kt_re = (double *)calloc(tot, sizeof(double)); c->kt_im = (double *)calloc(tot, sizeof(double));
  double eps = c->sg_eps;
  for (int i = 0; i < n2; i++) {
    double u = (i < nr) ? log(c->Radii[i] / c->Radii[0]) : -log(c->Radii[2 * nr - i] / c->Radii[0]);
    for (int j = 0; j < ns; j++) {
      size_t l = (size_t)i * ns + j;
      c->kr_re[l] = 1.0 + eps * eps - c->CosAz[j] * exp(-u);
      c->kr_re[l] *= pow(eps * eps * exp(u) + 2.0 * (cosh(u) - c->CosAz[j]), -1.5);
      c->kt_re[l] = c->SinAz[j];
      c->kt_re[l] *= pow(eps * eps * exp(u) + 2.0 * (cosh(u) - c->CosAz[j]), -1.5);
    }
  }
  fft2d(c->kr_re, c->kr_im, n2, ns, -1);
  fft2d(c->kt_re, c->kt_im, n2, ns, -1);
  c->sg_ready = 1;
}

/* normalisation and self-force (src/sgacc.c:137-161) applied to the raw
 * transform output raw_r/raw_t (the un-normalised inverse transform) */
static void sg_finish(or_ctx *c, const double *rho, const double *raw_r, const double *raw_t,
                      double *accr, double *acct) {
  int nr = c->nr, ns = c->ns;
  for (int i = 0; i < nr; i++) {
    double normaccr = c->sg_rstep * c->sg_tstep / ((double)(2 * nr) * (double)ns);
    double normacct = normaccr;
    normaccr /= sqrt(c->Rmed[i] / c->GlobalRmed[0]);
    normacct /= (c->Rmed[i] / c->GlobalRmed[0] * sqrt(c->Rmed[i] / c->GlobalRmed[0]));
    for (int j = 0; j < ns; j++) {
      int l = i * ns + j;
      accr[l] = raw_r[l] * normaccr;
      acct[l] = raw_t[l] * normacct;
    }
  }
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      int l = i * ns + j;
      accr[l] += G * rho[l] * c->sg_rstep * c->sg_tstep / c->sg_eps;
    }
}

/* compute_fftdensity + compute_sgacc (src/sgdens.c:3-86, src/sgacc.c:3-162) */
void or_sg_accel(or_ctx *c, const double *rho, double *accr, double *acct) {
  int nr = c->nr, ns = c->ns, n2 = 2 * nr;
  size_t tot = (size_t)n2 * ns;
  if (!c->sg_ready) sg_init(c);
  double *sr = (double *)calloc(tot, sizeof(double)), *si = (double *)calloc(tot, sizeof(double));
  double *tr = (double *)calloc(tot, sizeof(double)), *ti = (double *)calloc(tot, sizeof(double));
  double *ar = (double *)calloc(tot, sizeof(double)), *ai = (double *)calloc(tot, sizeof(double));
  double *br = (double *)calloc(tot, sizeof(double)), *bi = (double *)calloc(tot, sizeof(double));
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      size_t l = (size_t)i * ns + j;
      sr[l] = rho[i * ns + j] * sqrt(c->Rmed[i] / c->GlobalRmed[0]);
      tr[l] = sr[l] * c->Rmed[i] / c->GlobalRmed[0];
    }
  fft2d(sr, si, n2, ns, -1);
  fft2d(tr, ti, n2, ns, -1);
  for (size_t l = 0; l < tot; l++) {
    ar[l] = -G * (c->kr_re[l] * sr[l] - c->kr_im[l] * si[l]);
    ai[l] = -G * (c->kr_re[l] * si[l] + c->kr_im[l] * sr[l]);
    br[l] = -G * (c->kt_re[l] * tr[l] - c->kt_im[l] * ti[l]);
    bi[l] = -G * (c->kt_re[l] * ti[l] + c->kt_im[l] * tr[l]);
  }
  fft2d(ar, ai, n2, ns, +1);
  fft2d(br, bi, n2, ns, +1);
  sg_finish(c, rho, ar, br, accr, acct);
  free(sr); free(si); free(tr); free(ti); free(ar); free(ai); free(br); free(bi);
}

/* the same acceleration by direct summation of the cyclic convolution, for
 * tiny grids: raw[i][j] = -G * 2Nr*Ns * sum K[i-i'][j-j'] S[i'][j'] */
void or_sg_accel_direct(or_ctx *c, const double *rho, double *accr, double *acct) {
  int nr = c->nr, ns = c->ns, n2 = 2 * nr;
  double eps, *rawr, *rawt;
  if (!c->sg_ready) sg_init(c);
  eps = c->sg_eps;
  rawr = (double *)calloc((size_t)nr * ns, sizeof(double));
  rawt = (double *)calloc((size_t)nr * ns, sizeof(double));
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      long double sumr = 0., sumt = 0.;
      for (int ip = 0; ip < nr; ip++)
        for (int jp = 0; jp < ns; jp++) {
          int di = ((i - ip) % n2 + n2) % n2, dj = ((j - jp) % ns + ns) % ns;
          double u = (di < nr) ? log(c->Radii[di] / c->Radii[0]) : -log(c->Radii[2 * nr - di] / c->Radii[0]);
          double base = pow(eps * eps * exp(u) + 2.0 * (cosh(u) - c->CosAz[dj]), -1.5);
          double kr = (1.0 + eps * eps - c->CosAz[dj] * exp(-u)) * base;
          double kt = c->SinAz[dj] * base;
          double s_r = rho[ip * ns + jp] * sqrt(c->Rmed[ip] / c->GlobalRmed[0]);
          double s_t = s_r * c->Rmed[ip] / c->GlobalRmed[0];
          sumr += (long double)kr * s_r;
          sumt += (long double)kt * s_t;
        }
      rawr[i * ns + j] = -G * (double)sumr * (double)n2 * (double)ns;
      rawt[i * ns + j] = -G * (double)sumt * (double)n2 * (double)ns;
    }
  sg_finish(c, rho, rawr, rawt, accr, acct);
  free(rawr);
  free(rawt);
}

/* compute_SGZeroMode (src/sgzero.c:3-84): the axisymmetric part of the
 * disc's self-gravity, SG_Accr per ring.  The reduced ring-mean density
 * (zero-padded to 2 Nrad log-radius cells) is convolved cyclically with the
 * reduced kernel summed over the azimuths, through transforms of length
 * 2 Nrad (unnormalised, so the product carries a factor 2 Nrad that the
 * normalisation rstep / (2 Nrad) removes), then the whole ring's own
 * contribution is compensated. */
void or_sg_zero_mode(or_ctx *c, const double *rho, double *accr) {
  const int nr = c->nr, ns = c->ns, n2 = 2 * nr;
  const double eps = c->p.sg_thickness_smoothing * c->p.aspect_ratio;
  const double rstep = log(c->Radii[nr] / c->Radii[0]) / (double)nr;
  const double tstep = (c->p.pmax - c->p.pmin) / (double)ns;
  double *axi = (double *)calloc(nr, sizeof(double));
  double *sr = (double *)calloc(n2, sizeof(double)), *si = (double *)calloc(n2, sizeof(double));
  double *kr = (double *)calloc(n2, sizeof(double)), *ki = (double *)calloc(n2, sizeof(double));
  double *ar = (double *)calloc(n2, sizeof(double)), *ai = (double *)calloc(n2, sizeof(double));
  make_1d_profile(c, rho, axi);
  for (int i = 0; i < nr; i++) sr[i] = axi[i] * sqrt(c->GlobalRmed[i] / c->GlobalRmed[0]);
  for (int i = 0; i < n2; i++) {
    const double u = (i < nr) ? log(c->Radii[i] / c->Radii[0]) : -log(c->Radii[2 * nr - i] / c->Radii[0]);
    for (int j = 0; j < ns; j++)
      kr[i] += (1.0 + eps * eps - c->CosAz[j] * exp(-u)) *
               pow(eps * eps * exp(u) + 2.0 * (cosh(u) - c->CosAz[j]), -1.5);
    kr[i] *= tstep;
  }
  or_fft(sr, si, n2, -1);
  or_fft(kr, ki, n2, -1);
  for (int m = 0; m < n2; m++) {
    ar[m] = -G * (kr[m] * sr[m] - ki[m] * si[m]);
    ai[m] = -G * (kr[m] * si[m] + ki[m] * sr[m]);
  }
  or_fft(ar, ai, n2, +1);
  for (int i = 0; i < nr; i++) {
    double normaccr = rstep / (2.0 * (double)nr);
    normaccr /= sqrt(c->GlobalRmed[i] / c->GlobalRmed[0]);
    accr[i] = ar[i] * normaccr;
  }
  double integrale = 0.0;
  for (int j = 0; j < ns; j++) {
    const double num = 1.0 + eps * eps - c->CosAz[j];
    const double den = pow(2.0 * (1.0 - c->CosAz[j]) + eps * eps, 1.5);
    integrale += num / den;
  }
  integrale *= tstep;
  for (int i = 0; i < nr; i++) accr[i] += G * axi[i] * rstep * integrale;
  free(axi); free(sr); free(si); free(kr); free(ki); free(ar); free(ai);
}

/* compute_selfgravity (src/sgmain.c:3-35) + update_sgvelocity
 * (src/sgupdate.c:3-73).  In the axisymmetric mode SG_Accr is the ring
 * profile (kept in AxiSG, which the start-up uses as it stands) and there
 * is no azimuthal acceleration: the grids hold the ring values and zeros,
 * which the same update reads. */
static void compute_selfgravity(or_ctx *c, double *vrad, double *vtheta, double Dt, int update) {
  int nr = c->nr, ns = c->ns;
  if (c->p.sg_zero_mode) {
    or_sg_zero_mode(c, c->rho, c->AxiSG);
    for (int i = 0; i < nr; i++)
      for (int j = 0; j < ns; j++) {
        c->sg_accr[i * ns + j] = c->AxiSG[i];
        c->sg_acct[i * ns + j] = 0.0;
      }
  } else
    or_sg_accel(c, c->rho, c->sg_accr, c->sg_acct);
  memcpy(c->gr, c->sg_accr, sizeof(double) * nr * ns);
  memcpy(c->gtheta, c->sg_acct, sizeof(double) * nr * ns);
  if (!update) return;
  /* update_sgvelocity (src/sgupdate.c:3-72) with the dust fluid's kicks:
     DUSTFEELSG Y takes the full accelerations, Z the radial one of the
     ring-mean SG_Accr */
  const int dsg = c->df_on ? c->fp.dust_feel_sg : DFARGO_DUST_SG_NO;
  if (dsg == DFARGO_DUST_SG_ZERO_MODE && !c->p.sg_zero_mode) make_1d_profile(c, c->sg_accr, c->AxiSGd);
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      int l = i * ns + j;
      if (i > 0) {
        double radsgacc = ((c->Rinf[i] - c->Rmed[i - 1]) * c->sg_accr[l] +
                           (c->Rmed[i] - c->Rinf[i]) * c->sg_accr[l - ns]) * c->InvDiffRmed[i];
        c->radsgacc[l] = radsgacc;
        vrad[l] += Dt * radsgacc;
        if (c->p.sg_zero_mode) {
          if (dsg != DFARGO_DUST_SG_NO) c->dvradint[l] += Dt * radsgacc;
        } else if (dsg == DFARGO_DUST_SG_YES) {
          c->dvradint[l] += Dt * radsgacc;
        } else if (dsg == DFARGO_DUST_SG_ZERO_MODE) {
          const double axi = ((c->Rinf[i] - c->Rmed[i - 1]) * c->AxiSGd[i] +
                              (c->Rmed[i] - c->Rinf[i]) * c->AxiSGd[i - 1]) * c->InvDiffRmed[i];
          c->radsgacc[l] = axi;
          c->dvradint[l] += Dt * axi;
        }
      }
      int lm1 = i * ns + (j == 0 ? ns - 1 : j - 1);
      double azisgacc = 0.5 * (c->sg_acct[l] + c->sg_acct[lm1]);
      c->azisgacc[l] = azisgacc;
      vtheta[l] += Dt * azisgacc;
      if (!c->p.sg_zero_mode && dsg == DFARGO_DUST_SG_YES) c->dvthetaint[l] += Dt * azisgacc;
    }
}

/* ------------------------------------------------------------------------ */
/* InitEuler + InitGasVelocities                                              */
/* ------------------------------------------------------------------------ */

/* InitGasVelocities (src/Pframeforce.c:442-641), no centrifugal balance */
static void init_gas_velocities(or_ctx *c) {
  const dfargo_params *p = &c->p;
  int nr = c->nr, ns = c->ns;
  double *vr = c->vr, *vt = c->vt;
  for (int i = 0; i < nr; i++) {
    double r = c->Rmed[i], omega = sqrt(G * 1.0 / r / r / r);
    double myvtheta = omega * r * sqrt(1.0 - pow(p->aspect_ratio, 2.0) * pow(r, 2.0 * p->flaring_index) *
                                                 (1. + p->sigma_slope - 2.0 * p->flaring_index));
    for (int j = 0; j < ns; j++) vt[j + i * ns] = myvtheta;
  }
  if (p->self_gravity) {
    if (!p->sg_zero_mode) make_1d_profile(c, c->sg_accr, c->AxiSG);
    for (int i = 0; i < nr; i++) {
      double r = c->Rmed[i];
      for (int j = 0; j < ns; j++) {
        int l = i * ns + j;
        double vt2 = vt[l] * vt[l];
        vt2 -= r * c->AxiSG[i];
        vt[l] = sqrt(vt2);
      }
    }
  }
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) vt[i * ns + j] -= c->OmegaFrame * c->Rmed[i];
  if (p->alpha_viscosity != 0.0) make_1d_profile(c, c->cs, c->GlobalCs);
  for (int i = 0; i < nr; i++) {
    double r = c->Rmed[i], ri = c->Rinf[i];
    double viscosity = has_viscosity(p) ? or_fviscosity(c, r) : 0.0;
    for (int j = 0; j < ns; j++) {
      int l = i * ns + j;
      if (p->imposed_disk_drift != 0.0)
        vr[l] = p->imposed_disk_drift * p->sigma0 / c->SigmaInf[i] / ri;
      else if (p->alpha_viscosity != 0.0)
        vr[l] -= 3.0 * viscosity / r * (-p->sigma_slope + 2.0 * p->flaring_index + 1.0);
      else
        vr[l] -= 3.0 * viscosity / r * (-p->sigma_slope + .5);
    }
  }
  if (c->df_on && c->fp.dust_feedback) df_init_gas_feedback(c);
  for (int i = 0; i < nr; i++) {
    c->VthetaMed[i] = vt[i * ns] + c->Rmed[i] * c->OmegaFrame;
    c->VradMed[i] = vr[i * ns];
  }
}

/* init_planetarysys_withSG (src/sgsysinit.c:3-49), non-binary */
static void init_planetarysys_with_sg(or_ctx *c) {
  if (!c->p.sg_zero_mode) make_1d_profile(c, c->sg_accr, c->AxiSG);
  for (int k = 0; k < c->sys.nb; k++) {
    double x = c->sys.x[k], y = c->sys.y[k], r = sqrt(x * x + y * y), sgacc;
    double dist = (double)(r / (1. + c->p.eccentricity));
    if (dist < c->GlobalRmed[c->nr - 1]) {
      int ipl = 0;
      while ((c->GlobalRmed[ipl] <= dist) && (ipl < c->nr - 2)) ipl++;
      double ri = c->GlobalRmed[ipl], rip1 = c->GlobalRmed[ipl + 1], dr = rip1 - ri;
      sgacc = (dist - ri) * c->AxiSG[ipl + 1] + (rip1 - dist) * c->AxiSG[ipl];
      sgacc /= dr;
    } else
      sgacc = c->AxiSG[c->nr - 1];
    c->sys.vy[k] *= (double)sqrt(1. - dist * dist * sgacc / G / (1.0 + c->sys.mass[k]));
  }
}

int or_init_euler(or_ctx *c, const double *rho, double *vrad, double *vtheta,
                  const double *energy, const dfargo_planets *sys) {
  size_t n = (size_t)(c->nr + 1) * c->ns;
  memcpy(c->rho, rho, sizeof(double) * n);
  if (energy) memcpy(c->en, energy, sizeof(double) * n);
  if (sys) c->sys = *sys; else memset(&c->sys, 0, sizeof(c->sys));
  memset(c->vr, 0, sizeof(double) * c->gsize);
  memset(c->vt, 0, sizeof(double) * c->gsize);
  if (c->p.self_gravity) { /* src/main.c:178-186 */
    compute_selfgravity(c, c->vr, c->vt, 0.0, 0);
    init_planetarysys_with_sg(c);
  }
  if (c->df_on) { /* src/SourceEuler.c:265-269 */
    df_sound_speed_pressure(c);
    df_init_stokes(c);
    df_init_velocities(c);
  }
  compute_sound_speed(c);
  compute_pressure(c);
  compute_temperature(c);
  init_gas_velocities(c);
  compute_viscous_terms(c, c->vr, c->vt);
  if (c->p.energy_equation && c->p.viscous_heating) compute_viscous_heating(c);
  memcpy(vrad, c->vr, sizeof(double) * n);
  memcpy(vtheta, c->vt, sizeof(double) * n);
  return 0;
}

void or_set_state(or_ctx *c, const double *rho, const double *vrad, const double *vtheta,
                  const double *energy, const dfargo_planets *sys) {
  size_t n = (size_t)(c->nr + 1) * c->ns;
  if (rho) memcpy(c->rho, rho, sizeof(double) * n);
  if (vrad) memcpy(c->vr, vrad, sizeof(double) * n);
  if (vtheta) memcpy(c->vt, vtheta, sizeof(double) * n);
  if (energy) memcpy(c->en, energy, sizeof(double) * n);
  if (sys) c->sys = *sys;
}

void or_get_state(or_ctx *c, double *rho, double *vrad, double *vtheta, double *energy,
                  dfargo_planets *sys) {
  size_t n = (size_t)(c->nr + 1) * c->ns;
  if (rho) memcpy(rho, c->rho, sizeof(double) * n);
  if (vrad) memcpy(vrad, c->vr, sizeof(double) * n);
  if (vtheta) memcpy(vtheta, c->vt, sizeof(double) * n);
  if (energy) memcpy(energy, c->en, sizeof(double) * n);
  if (sys) *sys = c->sys;
}

void or_set_time(or_ctx *c, double t, double t0, double omega_frame) {
  c->PhysicalTime = t;
  c->PhysicalTimeInitial = t0;
  c->OmegaFrame = omega_frame;
}

/* ------------------------------------------------------------------------ */
/* CFL                                                                        */
/* ------------------------------------------------------------------------ */

/* ConditionCFL (src/SourceEuler.c:1304-1552), gas only, no diffusion */
int or_condition_cfl(or_ctx *c, double deltaT) {
  int nr = c->nr, ns = c->ns;
  double *Vmoy = c->vmoy, newdt = 1e30, viscosity = 0.0;
  double *Vres = (double *)malloc(sizeof(double) * (ns + 1));
  const double *vt = c->vt, *vr = c->vr, pm = c->p.pmax - c->p.pmin;
  for (int i = 0; i < nr; i++) {
    Vmoy[i] = 0.0;
    for (int j = 0; j < ns; j++) Vmoy[i] += vt[j + i * ns];
    Vmoy[i] /= (double)ns;
  }
  /* the dust fluid's terms (src/SourceEuler.c:1352-1361,1376-1384,1430-1471,1486-1491) */
  const int df = c->df_on, sfta = c->fp.sfta;
  const double *dustvt = c->dvt, *dustvr = c->dvr;
  double *DVres = (double *)malloc(sizeof(double) * (ns + 1)), dviscosity = 0.0;
  if (df)
    for (int i = 0; i < nr; i++) {
      c->dvmoy[i] = 0.0;
      for (int j = 0; j < ns; j++) c->dvmoy[i] += dustvt[j + i * ns];
      c->dvmoy[i] /= (double)ns;
    }
  for (int i = 1; i < nr; i++) {
    if (has_viscosity(&c->p)) viscosity = or_fviscosity(c, c->Rmed[i]);
    if (df_has_viscosity(c)) dviscosity = dfviscosity(c, c->Rmed[i]);
    double dxrad = c->Rsup[i] - c->Rinf[i];
    double dxtheta = c->Rmed[i] * pm / (double)ns;
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns;
      Vres[j] = c->p.fast_transport ? vt[l] - Vmoy[i] : vt[l];
      if (df) DVres[j] = c->p.fast_transport ? dustvt[l] - c->dvmoy[i] : dustvt[l];
    }
    Vres[ns] = Vres[0];
    DVres[ns] = DVres[0];
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns, lip = l + ns, ljp = (j == ns - 1) ? i * ns : l + 1;
      double cs = c->cs[l];
      double invdt1 = cs / (min2(dxrad, dxtheta));
      double invdt2 = fabs(vr[l]) / dxrad;
      double invdt3 = fabs(Vres[j]) / dxtheta;
      double dvr = vr[lip] - vr[l], dvt = vt[ljp] - vt[l], invdt4, invdt5, invdt6 = 1e-10, dt;
      if (dvr >= 0.0) dvr = 1e-10; else dvr = -dvr;
      if (dvt >= 0.0) dvt = 1e-10; else dvt = -dvt;
      invdt4 = max2(dvr / dxrad, dvt / dxtheta);
      invdt4 *= 4.0 * CVNR * CVNR;
      if (viscosity != 0.0)
        invdt5 = viscosity * 4.0 / pow(min2(dxrad, dxtheta), 2.0);
      else
        invdt5 = 1e-10;
      if (df) {
        double invdt7 = 1.e-10, invdt8 = 1.e-10, invdt9 = 1.e-10, invdt10 = 1.e-10, invdt11, invdt12;
        const double csd = c->dcs1[i];
        if (!sfta) {
          invdt7 = csd / (min2(dxrad, dxtheta));
          invdt8 = fabs(dustvr[l] + 1e-15) / dxrad;
          invdt9 = fabs(DVres[j]) / dxtheta;
          invdt10 = c->fp.dust_feel_disk ? 1. / (c->stokes[l] + 1.e-10) * sqrt(G * 1.0 / c->Rmed[i]) / c->Rmed[i]
                                         : 1.e-10;
        }
        double ddvr = dustvr[lip] - dustvr[l], ddvt = dustvt[ljp] - dustvt[l];
        if (ddvr >= 0.0) ddvr = 1e-10; else ddvr = -ddvr;
        if (ddvt >= 0.0) ddvt = 1e-10; else ddvt = -ddvt;
        invdt11 = max2(ddvr / dxrad, ddvt / dxtheta);
        invdt11 *= 4.0 * CVNR * CVNR;
        if (dviscosity != 0.0)
          invdt12 = dviscosity * 4.0 / pow(min2(dxrad, dxtheta), 2.0);
        else
          invdt12 = 1e-10;
        const double dtg = CFLSECURITY / sqrt(invdt1 * invdt1 + invdt2 * invdt2 + invdt3 * invdt3 +
                                              invdt4 * invdt4 + invdt5 * invdt5 + invdt6 * invdt6);
        const double dtd = CFLSECURITY / sqrt(invdt7 * invdt7 + invdt8 * invdt8 + invdt9 * invdt9 +
                                              invdt10 * invdt10 + invdt11 * invdt11 + invdt12 * invdt12);
        dt = min2(dtg, dtd);
        dt = max2(dt, 1.e-10);
      } else
        dt = CFLSECURITY / sqrt(invdt1 * invdt1 + invdt2 * invdt2 + invdt3 * invdt3 +
                                invdt4 * invdt4 + invdt5 * invdt5 + invdt6 * invdt6);
      if (dt < newdt) newdt = dt;
    }
  }
  for (int i = 0; i < nr - 1; i++) {
    double dt = pm * CFLSECURITY / (double)c->ns /
                fabs(Vmoy[i] * c->InvRmed[i] - Vmoy[i + 1] * c->InvRmed[i + 1]);
    if (dt < newdt) newdt = dt;
  }
  if (df && !sfta)
    for (int i = 0; i < nr - 1; i++) {
      double dt = pm * CFLSECURITY / (double)c->ns /
                  fabs(c->dvmoy[i] * c->InvRmed[i] - c->dvmoy[i + 1] * c->InvRmed[i + 1]);
      if (dt < newdt) newdt = dt;
    }
  free(DVres);
  /* the minimum stopping time of the last interpolation resolved by two
     steps (src/SourceEuler.c:1501-1507) */
  if (c->dust_on && c->np != 0 && c->dp.dust_feedback) {
    double dt_fb = 0.5 * c->min_stopping_time;
    if (dt_fb == 0.0) dt_fb = 1e10;
    if (dt_fb < newdt) newdt = dt_fb;
  }
  if (deltaT < newdt) newdt = deltaT;
  free(Vres);
  return (int)(ceil(deltaT / newdt));
}

/* ------------------------------------------------------------------------ */
/* source steps                                                               */
/* ------------------------------------------------------------------------ */

/* SubStep1 (src/SourceEuler.c:765-1009), with the dust fluid's Stokes
   number, kicks and drag (src/SourceEuler.c:805-858,885-922,961-999) */
void or_substep1(or_ctx *c, double dt) {
  const dfargo_params *p = &c->p;
  const dfargo_dust_fluid_params *fp = &c->fp;
  int nr = c->nr, ns = c->ns;
  const double *rho = c->rho, *vrad = c->vr, *vtheta = c->vt, *Pot = c->pot, *Press = c->pres;
  double *vradint = c->vradint, *vthetaint = c->vthetaint;
  const int df = c->df_on;
  const double *drho = c->drho, *dvrad = c->dvr, *dvtheta = c->dvt, *dPress = c->dpres, *St = c->stokes;
  double *dvradint = c->dvradint, *dvthetaint = c->dvthetaint;
  if (df) df_stokes(c);
  for (int i = 1; i < nr; i++) {
    const double invomega = pow(c->Rmed[i], 1.5);
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns, lim = l - ns, ljp = (j == ns - 1) ? i * ns : l + 1;
      double gradp = (Press[l] - Press[lim]) * 2.0 / (rho[l] + rho[lim]) * c->InvDiffRmed[i];
      c->radgradp[l] = (Press[l] - Press[lim]) * c->InvDiffRmed[i];
      double gradphi = (Pot[l] - Pot[lim]) * c->InvDiffRmed[i];
      c->radindacc[l] = -(c->indpot[l] - c->indpot[lim]) * c->InvDiffRmed[i];
      double vt2 = vtheta[l] + vtheta[ljp] + vtheta[lim] + vtheta[ljp - ns];
      vt2 = vt2 / 4.0 + c->Rinf[i] * c->OmegaFrame;
      vt2 = vt2 * vt2;
      vradint[l] = vrad[l] + dt * (-gradp - gradphi + vt2 * c->InvRinf[i]);
      if (c->dust_on && c->dp.dust_feedback) vradint[l] += dt * c->radfbacc[l];
      if (df) {
        double ts = St[l] * invomega;
        if (!fp->sfta) {
          double dgradp = (dPress[l] - dPress[lim]) * 2.0 / (drho[l] + drho[lim]) * c->InvDiffRmed[i];
          double dvt2 = dvtheta[l] + dvtheta[ljp] + dvtheta[lim] + dvtheta[ljp - ns];
          dvt2 = dvt2 / 4.0 + c->Rinf[i] * c->OmegaFrame;
          dvt2 = dvt2 * dvt2;
          dvradint[l] = dvrad[l] + dt * (-gradphi - dgradp + dvt2 * c->InvRinf[i]);
          if (fp->dust_feel_disk) dvradint[l] = (dvradint[l] * ts + vrad[l] * dt) / (ts + dt);
        } else if (fp->dust_feel_disk) {
          if (fp->dust_feedback) ts /= (1.0 + drho[l] / rho[l]);
          dvradint[l] = vrad[l] + ts * gradp;
        }
        if (fp->dust_feedback) {
          ts = St[l] * invomega;
          c->radfbacc[l] = (drho[l] / rho[l]) * (dvrad[l] - vrad[l]) / (ts + dt / 2.);
          vradint[l] = (vradint[l] * ts + (drho[l] / rho[l]) * dvrad[l] * dt) / (ts + dt * (drho[l] / rho[l]));
        }
      }
    }
  }
  for (int i = 0; i < nr; i++) {
    double dxtheta = (p->pmax - p->pmin) / (double)ns * c->Rmed[i];
    double invdxtheta = 1.0 / dxtheta;
    double supp_torque = p->imposed_disk_drift * 0.5 * pow(c->Rmed[i], -2.5) * p->sigma0 / c->SigmaMed[i];
    const double invomega = pow(c->Rmed[i], 1.5);
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns, ljm = (j == 0) ? i * ns + ns - 1 : l - 1;
      double gradp = (Press[l] - Press[ljm]) * 2.0 / (rho[l] + rho[ljm]) * invdxtheta;
      c->azigradp[l] = (Press[l] - Press[ljm]) * invdxtheta;
      double gradphi = (Pot[l] - Pot[ljm]) * invdxtheta;
      c->aziindacc[l] = -(c->indpot[l] - c->indpot[ljm]) * invdxtheta;
      vthetaint[l] = vtheta[l] - dt * (gradp + gradphi);
      vthetaint[l] += dt * supp_torque;
      if (c->dust_on && c->dp.dust_feedback) vthetaint[l] += dt * c->azifbacc[l];
      if (df) {
        double ts = St[l] * invomega;
        if (!fp->sfta) {
          double dgradp = (dPress[l] - dPress[ljm]) * 2.0 / (drho[l] + drho[ljm]) * invdxtheta;
          dvthetaint[l] = dvtheta[l] - dt * (gradphi + dgradp);
          if (fp->dust_feel_disk) dvthetaint[l] = (dvthetaint[l] * ts + vtheta[l] * dt) / (ts + dt);
        } else if (fp->dust_feel_disk) {
          if (fp->dust_feedback) ts /= (1.0 + drho[l] / rho[l]);
          dvthetaint[l] = vtheta[l] + ts * gradp;
        }
        if (fp->dust_feedback) {
          ts = St[l] * invomega;
          c->azifbacc[l] = (drho[l] / rho[l]) * (dvtheta[l] - vtheta[l]) / (ts + dt / 2.);
          vthetaint[l] =
              (vthetaint[l] * ts + (drho[l] / rho[l]) * dvtheta[l] * dt) / (ts + dt * (drho[l] / rho[l]));
        }
      }
    }
  }
  if (p->self_gravity) compute_selfgravity(c, vradint, vthetaint, dt, 1);
  if (has_viscosity(p) || df_has_viscosity(c)) {
    compute_viscous_terms(c, vradint, vthetaint);
    if (df) df_viscous_terms(c, dvradint, dvthetaint);
    update_velocities_viscosity(c, vradint, vthetaint, dt);
    if (df) df_update_viscosity(c, dvradint, dvthetaint, dt);
  }
}

/* SubStep2 (src/SourceEuler.c:1012-1116) */
static void substep2(or_ctx *c, double dt) {
  int nr = c->nr, ns = c->ns;
  const double *rho = c->rho, *vrad = c->vradint, *vtheta = c->vthetaint, *energy = c->en;
  double *qr = c->rhoint, *qt = c->temperint;
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns, lip = l + ns, ljp = (j == ns - 1) ? i * ns : l + 1;
      double dv = vrad[lip] - vrad[l];
      qr[l] = (dv < 0.0) ? CVNR * CVNR * rho[l] * dv * dv : 0.0;
      dv = vtheta[ljp] - vtheta[l];
      qt[l] = (dv < 0.0) ? CVNR * CVNR * rho[l] * dv * dv : 0.0;
    }
  for (int i = 1; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns, lim = l - ns;
      c->vradnew[l] = vrad[l] - dt * 2.0 / (rho[l] + rho[lim]) * (qr[l] - qr[lim]) * c->InvDiffRmed[i];
    }
  for (int i = 0; i < nr; i++) {
    double dxtheta = (c->p.pmax - c->p.pmin) / (double)ns * c->Rmed[i];
    double invdxtheta = 1.0 / dxtheta;
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns, lip = l + ns, ljm = (j == 0) ? i * ns + ns - 1 : l - 1;
      int ljp = (j == ns - 1) ? i * ns : l + 1;
      c->vthetanew[l] = vtheta[l] - dt * 2.0 / (rho[l] + rho[ljm]) * (qt[l] - qt[ljm]) * invdxtheta;
      if (c->p.energy_equation)
        c->energyint[l] = energy[l] - dt * qr[l] * (vrad[lip] - vrad[l]) * c->InvDiffRsup[i] -
                          dt * qt[l] * (vtheta[ljp] - vtheta[l]) * invdxtheta;
    }
  }
}

/* SubStep3 (src/SourceEuler.c:1119-1302), no thermal cooling/TempPresc */
static void substep3(or_ctx *c, double dt) {
  const dfargo_params *p = &c->p;
  int nr = c->nr, ns = c->ns;
  const double *energy = c->energyint, *vtheta = c->vt;
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns;
      double num = energy[l] + (p->viscous_heating ? 1 : 0) * c->vischeat[l] * dt;
      double den = 1.0 + (p->adiabatic_index - 1.0) * dt * c->divv[l];
      c->energynew[l] = num / den;
    }
  if (p->beta_cooling) {
    for (int i = 0; i < nr; i++) {
      double omega = 0.0;
      for (int j = 0; j < ns; j++) omega += (vtheta[j + i * ns] + c->Rmed[i] * c->OmegaFrame);
      omega /= (double)ns;
      omega /= c->Rmed[i];
      double beta = p->beta_cooling_time * pow(c->Rmed[i], -p->beta_cooling_slope);
      double coolingtime = beta / omega;
      for (int j = 0; j < ns; j++) {
        int l = j + i * ns;
        double num = c->energynew[l], den = 1.0 + dt / coolingtime;
        c->energynew[l] = num / den;
      }
    }
  }
  /* heating by the dust feedback (src/SourceEuler.c:1291-1299) */
  if (c->dust_on && c->dp.dust_feedback)
    for (int l = 0; l < nr * ns; l++) c->energynew[l] += c->fbedot[l] * dt;
}

/* ------------------------------------------------------------------------ */
/* transport                                                                  */
/* ------------------------------------------------------------------------ */

/* ComputeLRMomenta (src/TransportEuler.c:89-118) */
static void compute_lr_momenta(or_ctx *c) {
  int nr = c->nr, ns = c->ns;
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns, lip = l + ns, ljp = (j == ns - 1) ? i * ns : l + 1;
      c->rp[l] = c->rho[l] * c->vr[lip];
      c->rm[l] = c->rho[l] * c->vr[l];
      c->tp[l] = c->rho[l] * (c->vt[ljp] + c->Rmed[i] * c->OmegaFrame) * c->Rmed[i];
      c->tm[l] = c->rho[l] * (c->vt[l] + c->Rmed[i] * c->OmegaFrame) * c->Rmed[i];
    }
}

/* ComputeStarRad (src/TransportEuler.c:142-180) */
static void compute_star_rad(or_ctx *c, const double *qb, const double *vr, double *qs, double dt) {
  int nr = c->nr, ns = c->ns;
  double *dq = c->dq;
  for (int j = 0; j < ns; j++) {
    for (int i = 0; i < nr; i++) {
      int l = j + i * ns, lq = i + j * nr, lip = l + ns, lim = l - ns;
      if ((i == 0) || (i == nr - 1))
        dq[lq] = 0.0;
      else {
        double dqm = (qb[l] - qb[lim]) * c->InvDiffRmed[i];
        double dqp = (qb[lip] - qb[l]) * c->InvDiffRmed[i + 1];
        dq[lq] = (dqp * dqm > 0.0) ? 2.0 * dqp * dqm / (dqp + dqm) : 0.0;
      }
    }
    for (int i = 1; i < nr; i++) {
      int l = j + i * ns, lq = i + j * nr, lim = l - ns;
      if (vr[l] > 0.0)
        qs[l] = qb[lim] + (c->Rmed[i] - c->Rmed[i - 1] - vr[l] * dt) * 0.5 * dq[lq - 1];
      else
        qs[l] = qb[l] - (c->Rmed[i + 1] - c->Rmed[i] + vr[l] * dt) * 0.5 * dq[lq];
    }
    qs[j] = qs[j + ns * nr] = 0.0;
  }
}

/* DivisePolarGrid (src/SideEuler.c:244-261) over rings 0..nr-1 */
static void divise(or_ctx *c, const double *num, const double *den, double *res) {
  for (int l = 0; l < c->nr * c->ns; l++) res[l] = num[l] / (den[l] + 1e-20);
}

/* VanLeerRadial (src/TransportEuler.c:182-220) */
static double van_leer_radial(or_ctx *c, double *qb, double dt) {
  int nr = c->nr, ns = c->ns;
  double dtheta = (c->p.pmax - c->p.pmin) / (double)ns, lost = 0.0;
  divise(c, qb, c->rhoint, c->work);
  compute_star_rad(c, c->work, c->vr, c->qrstar, dt);
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns, lip = l + ns;
      double varq = dt * dtheta * c->Rinf[i] * c->qrstar[l] * c->rhostar[l] * c->vr[l];
      varq -= dt * dtheta * c->Rsup[i] * c->qrstar[lip] * c->rhostar[lip] * c->vr[lip];
      qb[l] += varq * c->InvSurf[i];
      if (i == 0) lost += varq;
    }
  return lost;
}

/* ComputeStarTheta (src/TransportEuler.c:419-444) */
static void compute_star_theta(or_ctx *c, const double *qb, const double *vt, double *qs, double dt) {
  int nr = c->nr, ns = c->ns;
  double *dq = c->dq;
  for (int i = 0; i < nr; i++) {
    double dxtheta = (c->p.pmax - c->p.pmin) / (double)ns * c->Rmed[i];
    double invdxtheta = 1.0 / dxtheta;
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns, ljp = (j == ns - 1) ? i * ns : l + 1, ljm = (j == 0) ? i * ns + ns - 1 : l - 1;
      double dqm = (qb[l] - qb[ljm]), dqp = (qb[ljp] - qb[l]);
      dq[l] = (dqp * dqm > 0.0) ? dqp * dqm / (dqp + dqm) * invdxtheta : 0.0;
    }
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns, ljm = (j == 0 ? ns - 1 : j - 1) + i * ns;
      double ksi = vt[l] * dt;
      if (ksi > 0.0)
        qs[l] = qb[ljm] + (dxtheta - ksi) * dq[ljm];
      else
        qs[l] = qb[l] - (dxtheta + ksi) * dq[l];
    }
  }
}

/* VanLeerTheta (src/TransportEuler.c:446-465) */
static void van_leer_theta(or_ctx *c, const double *vt, double *qb, double dt, int uniform) {
  int nr = c->nr, ns = c->ns;
  divise(c, qb, c->rhoint, c->work);
  compute_star_theta(c, c->work, vt, c->qrstar, dt);
  for (int i = 0; i < nr; i++) {
    double dxrad = (c->Rsup[i] - c->Rinf[i]) * dt, invsurf = 1.0 / c->Surf[i];
    if (!uniform || !c->nosplit[i]) {
      for (int j = 0; j < ns; j++) {
        int l = j + i * ns, ljp = (j == ns - 1) ? i * ns : l + 1;
        double varq = dxrad * c->qrstar[l] * c->rhostar[l] * vt[l];
        varq -= dxrad * c->qrstar[ljp] * c->rhostar[ljp] * vt[ljp];
        qb[l] += varq * invsurf;
      }
    }
  }
}

/* QuantitiesAdvection (src/TransportEuler.c:404-417) */
static void quantities_advection(or_ctx *c, const double *vt, double dt, int uniform) {
  compute_star_theta(c, c->rho, vt, c->rhostar, dt);
  memcpy(c->rhoint, c->rho, sizeof(double) * (c->nr + 1) * c->ns);
  if (!c->df_skip_momenta) {
    van_leer_theta(c, vt, c->rp, dt, uniform);
    van_leer_theta(c, vt, c->rm, dt, uniform);
    van_leer_theta(c, vt, c->tp, dt, uniform);
    van_leer_theta(c, vt, c->tm, dt, uniform);
  }
  if (c->p.energy_equation) van_leer_theta(c, vt, c->en, dt, uniform);
  van_leer_theta(c, vt, c->rho, dt, uniform);
}

/* AdvectSHIFT (src/TransportEuler.c:350-373) */
static void advect_shift(or_ctx *c, double *val) {
  int nr = c->nr, ns = c->ns;
  for (int i = 0; i < nr; i++) {
    for (int j = 0; j < ns; j++) {
      long ji = j - c->nshift[i];
      while (ji < 0) ji += ns;
      while (ji >= ns) ji -= ns;
      c->tempshift[j + i * ns] = val[ji + i * ns];
    }
    for (int j = 0; j < ns; j++) val[j + i * ns] = c->tempshift[j + i * ns];
  }
}

/* OneWindTheta with its helpers (src/TransportEuler.c:225-348) */
static void one_wind_theta(or_ctx *c, double dt) {
  int nr = c->nr, ns = c->ns;
  double invdt = 1.0 / dt, pm = c->p.pmax - c->p.pmin, dpinvns = pm / (double)ns;
  for (int l = 0; l < nr * ns; l++) c->elong[l] = c->vt[l] * dt;
  for (int i = 0; i < nr; i++) {
    double moy = 0.0;
    for (int j = 0; j < ns; j++) moy += c->elong[j + i * ns] * invdt;
    c->vmed[i] = moy / (double)ns;
  }
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) c->vthetares[j + i * ns] = c->elong[j + i * ns] * invdt - c->vmed[i];
  for (int i = 0; i < nr; i++) {
    double Ntilde = c->vmed[i] * c->InvRmed[i] * dt * (double)ns / pm;
    double Nround = floor(Ntilde + 0.5);
    c->nshift[i] = (long)Nround;
    /* distance from the rounding tie, the smallest of the call */
    double margin = 0.5 - fabs(Ntilde - Nround);
    if (margin < c->fargo_margin) c->fargo_margin = margin;
    for (int j = 0; j < ns; j++) c->vt[j + i * ns] = (Ntilde - Nround) * c->Rmed[i] * invdt * dpinvns;
    if (!c->p.fast_transport) {
      c->nosplit[i] = 1;
      for (int j = 0; j < ns; j++) {
        int l = j + i * ns;
        c->vthetares[l] = c->vt[l] + c->vthetares[l];
        c->vt[l] = 0.0;
      }
    } else
      c->nosplit[i] = 0;
  }
  quantities_advection(c, c->vthetares, dt, 0);
  quantities_advection(c, c->vt, dt, 1);
  advect_shift(c, c->rp);
  advect_shift(c, c->rm);
  advect_shift(c, c->tp);
  advect_shift(c, c->tm);
  if (c->p.energy_equation) advect_shift(c, c->en);
  advect_shift(c, c->rho);
}

/* ComputeVelocities (src/TransportEuler.c:515-547) */
static void compute_velocities(or_ctx *c) {
  int nr = c->nr, ns = c->ns;
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns, lim = l - ns, ljm = (j == 0) ? i * ns + ns - 1 : l - 1;
      if (i == 0)
        c->vr[l] = 0.0;
      else
        c->vr[l] = (c->rp[lim] + c->rm[l]) / (c->rho[l] + c->rho[lim] + 1e-20);
      c->vt[l] = (c->tp[ljm] + c->tm[l]) / (c->rho[l] + c->rho[ljm] + 1e-20) / c->Rmed[i] -
                 c->Rmed[i] * c->OmegaFrame;
    }
}

/* Transport (src/TransportEuler.c:65-84) */
void or_transport(or_ctx *c, double dt) {
  compute_lr_momenta(c);
  /* OneWindRad (src/TransportEuler.c:123-137) */
  compute_star_rad(c, c->rho, c->vr, c->rhostar, dt);
  memcpy(c->rhoint, c->rho, sizeof(double) * (c->nr + 1) * c->ns);
  van_leer_radial(c, c->rp, dt);
  van_leer_radial(c, c->rm, dt);
  van_leer_radial(c, c->tp, dt);
  van_leer_radial(c, c->tm, dt);
  if (c->p.energy_equation) van_leer_radial(c, c->en, dt);
  c->LostMass += van_leer_radial(c, c->rho, dt);
  one_wind_theta(c, dt);
  compute_velocities(c);
}

/* ------------------------------------------------------------------------ */
/* boundaries, floors, crash                                                  */
/* ------------------------------------------------------------------------ */

/* ApplySubKeplerianBoundary (src/SideEuler.c:1008-1044) */
static void subkeplerian_boundary(or_ctx *c) {
  int nr = c->nr, ns = c->ns;
  for (int j = 0; j < ns; j++) c->vt[j] = c->VthetaMed[0] - c->Rmed[0] * c->OmegaFrame;
  for (int j = 0; j < ns; j++)
    c->vt[(nr - 1) * ns + j] = c->VthetaMed[nr - 1] - c->Rmed[nr - 1] * c->OmegaFrame;
}

/* ApplyWallBoundaries (src/SideEuler.c:1047-1099) */
static void wall_boundaries(or_ctx *c) {
  int nr = c->nr, ns = c->ns, i = 2;
  for (int j = 0; j < ns; j++) {
    c->vr[(i - 1) * ns + j] = 0.0;
    c->vr[(i - 2) * ns + j] = -c->vr[i * ns + j];
    c->rho[(i - 1) * ns + j] = c->SigmaMed[i - 1];
    c->rho[(i - 2) * ns + j] = c->SigmaMed[i - 2];
  }
  i = nr - 3;
  for (int j = 0; j < ns; j++) {
    c->vr[(i + 2) * ns + j] = 0.0;
    c->vr[(i + 3) * ns + j] = -c->vr[(i + 1) * ns + j];
    c->rho[(i + 1) * ns + j] = c->SigmaMed[i + 1];
    c->rho[(i + 2) * ns + j] = c->SigmaMed[i + 2];
  }
}

/* OpenBoundary (src/SideEuler.c:300-408), standard (non-KN) prescription */
static void open_boundary(or_ctx *c) {
  int nr = c->nr, ns = c->ns, i = 2;
  double *rho = c->rho, *vr = c->vr, *vt = c->vt, *energy = c->en;
  for (int j = 0; j < ns; j++) {
    int l = j + i * ns;
    rho[l - ns] = rho[l];
    rho[l - 2 * ns] = rho[l];
    if (c->p.energy_equation) {
      energy[l - ns] = energy[l];
      energy[l - 2 * ns] = energy[l];
    }
    vt[l - ns] = vt[l] * sqrt(c->Rmed[i] / c->Rmed[i - 1]);
    vt[l - 2 * ns] = vt[l] * sqrt(c->Rmed[i] / c->Rmed[i - 2]);
    if ((vr[l + ns] >= 0.0) || (rho[l] < c->SigmaMed[0])) {
      vr[l] = 0.0;
      vr[l - ns] = 0.0;
    } else {
      vr[l] = vr[l + ns];
      vr[l - ns] = vr[l + ns];
    }
  }
  i = nr - 1;
  for (int j = 0; j < ns; j++) {
    int l = j + i * ns;
    rho[l] = rho[l - ns];
    if (c->p.energy_equation) energy[l] = energy[l - ns];
    vt[l] = vt[l - ns] * pow(c->Rmed[i] / c->Rmed[i - 1], -0.5);
    if ((vr[l - ns] < 0.0) || (rho[l] < c->SigmaMed[nr - 2]))
      vr[l] = 0.0;
    else
      vr[l] = vr[l - ns];
  }
}

/* NonReflectingBoundary (src/SideEuler.c:569-794), no accreting variant.
 * The pattern speed is that of the innermost (outermost) planet inside the
 * grid; where the reference reads an undefined radius (planet 0 outside
 * the grid) the grid's edge stands in.  The energy grid of a barotropic
 * disc holds zeros, which the outer edge's copy and mean adjustment leave
 * at zero. */
static void nonreflecting_boundary(or_ctx *c) {
  const dfargo_params *p = &c->p;
  const int nr = c->nr, ns = c->ns;
  double *rho = c->rho, *vr = c->vr, *cs = c->cs, *energy = c->en;
  /* inner edge, adiabatic case (src/SideEuler.c:654-699) */
  if (p->energy_equation) {
    const int i = 1;
    double buf = c->Radii[nr], r_in = c->Radii[nr], omega_in = pow(c->Radii[nr], -1.5);
    for (int k = 0; k < c->sys.nb; k++) {
      const double xp = c->sys.x[k], yp = c->sys.y[k], rp = sqrt(xp * xp + yp * yp);
      if ((rp < buf) && (rp >= c->Radii[0])) {
        r_in = rp;
        omega_in = pow(rp, -1.5);
      }
      buf = r_in;
    }
    const double g = p->adiabatic_index;
    const double csinit0 = sqrt(g) * p->aspect_ratio * pow(c->Rmed[0], -0.5 + p->flaring_index);
    const double csinit1 = sqrt(g) * p->aspect_ratio * pow(c->Rmed[1], -0.5 + p->flaring_index);
    double dangle = (pow(c->Rinf[1], -1.5) - omega_in) / (.5 * (csinit0 + csinit1));
    dangle *= (c->Rmed[1] - c->Rmed[0]);
    const int i_angle = (int)(dangle / (p->pmax - p->pmin) * (double)ns + .5);
    for (int j = 0; j < ns; j++) {
      const int l = j + i * ns;
      int jp = j + i_angle;
      if (jp >= ns) jp -= ns;
      if (jp < 0) jp += ns;
      energy[jp] = energy[l];
      const double vr_med = -csinit1 * (rho[l] - c->SigmaMed[1]) / c->SigmaMed[1];
      vr[l] = 2. * vr_med - vr[l + ns];
    }
    for (int j = 0; j < ns; j++)
      rho[j] = c->SigmaMed[0] + (g - 1.0) * (energy[j] - c->EnergyMed[0]) / csinit0 / csinit0;
    double mean = 0.0;
    for (int j = 0; j < ns; j++) mean += rho[j];
    mean /= (double)ns;
    for (int j = 0; j < ns; j++) rho[j] += (c->SigmaMed[0] - mean);
    mean = 0.0;
    for (int j = 0; j < ns; j++) mean += energy[j];
    mean /= (double)ns;
    for (int j = 0; j < ns; j++) energy[j] += (c->EnergyMed[0] - mean);
  } else {
    const int i = 1;
    double buf = c->Radii[nr], r_in = c->Radii[nr], omega_in = pow(c->Radii[nr], -1.5);
    for (int k = 0; k < c->sys.nb; k++) {
      const double xp = c->sys.x[k], yp = c->sys.y[k], rp = sqrt(xp * xp + yp * yp);
      if ((rp < buf) && (rp >= c->Radii[0])) {
        r_in = rp;
        omega_in = pow(rp, -1.5);
      }
      buf = r_in;
    }
    double cs0 = 0.0, cs1 = 0.0, dens1 = 0.0;
    for (int j = 0; j < ns; j++) {
      cs0 += cs[j];
      cs1 += cs[ns + j];
      dens1 += rho[ns + j];
    }
    cs0 /= (double)ns;
    cs1 /= (double)ns;
    dens1 /= (double)ns;
    double dangle = (pow(c->Rinf[1], -1.5) - omega_in) / (.5 * (cs0 + cs1));
    dangle *= (c->Rmed[1] - c->Rmed[0]);
    const int i_angle = (int)(dangle / (p->pmax - p->pmin) * (double)ns + .5);
    for (int j = 0; j < ns; j++) {
      const int l = j + i * ns;
      int jp = j + i_angle;
      if (jp >= ns) jp -= ns;
      if (jp < 0) jp += ns;
      rho[jp] = rho[l];
      const double vr_med = -cs1 * (rho[l] - dens1) / dens1;
      vr[l] = 2. * vr_med - vr[l + ns];
    }
    double mean = 0.0;
    for (int j = 0; j < ns; j++) mean += rho[j];
    mean /= (double)ns;
    for (int j = 0; j < ns; j++) rho[j] += (c->SigmaMed[0] - mean);
  }
  /* outer edge (both cases) */
  {
    double buf = c->Radii[0], r_out = c->Radii[0], omega_out = pow(c->Radii[0], -1.5);
    for (int k = 0; k < c->sys.nb; k++) {
      const double xp = c->sys.x[k], yp = c->sys.y[k], rp = sqrt(xp * xp + yp * yp);
      if ((rp > buf) && (rp <= c->Radii[nr])) {
        r_out = rp;
        omega_out = pow(rp, -1.5);
      }
      buf = r_out;
    }
    double csnrm2 = 0.0, csnrm1 = 0.0, densnrm2 = 0.0;
    for (int j = 0; j < ns; j++) {
      csnrm2 += cs[(nr - 2) * ns + j];
      csnrm1 += cs[(nr - 1) * ns + j];
      densnrm2 += rho[(nr - 2) * ns + j];
    }
    csnrm1 /= (double)ns;
    csnrm2 /= (double)ns;
    densnrm2 /= (double)ns;
    const int i = nr - 1;
    double dangle = (pow(c->Rinf[nr - 2], -1.5) - omega_out) / (.5 * (csnrm1 + csnrm2));
    dangle *= (c->Rmed[nr - 1] - c->Rmed[nr - 2]);
    const int i_angle = (int)(dangle / (p->pmax - p->pmin) * (double)ns + .5);
    for (int j = 0; j < ns; j++) {
      const int l = j + i * ns;
      int jp = j - i_angle;
      if (jp >= ns) jp -= ns;
      if (jp < 0) jp += ns;
      const int lp = jp + (i - 1) * ns;
      rho[l] = rho[lp];
      energy[l] = energy[lp];
      const double vr_med = (!p->energy_equation ? csnrm1 : cs[l]) * (rho[l - ns] - densnrm2) / densnrm2;
      vr[l] = 2. * vr_med - vr[l - ns];
    }
    double mean = 0.0;
    for (int j = 0; j < ns; j++) mean += rho[j + ns * (nr - 1)];
    mean /= (double)ns;
    for (int j = 0; j < ns; j++) rho[j + (nr - 1) * ns] += c->SigmaMed[nr - 1] - mean;
    mean = 0.0;
    for (int j = 0; j < ns; j++) mean += energy[j + ns * (nr - 1)];
    mean /= (double)ns;
    for (int j = 0; j < ns; j++) energy[j + (nr - 1) * ns] += c->EnergyMed[nr - 1] - mean;
  }
}

/* ApplyBoundaryCondition (src/SideEuler.c:1102-1131) */
static void apply_boundary_condition(or_ctx *c) {
  if (c->p.boundary == DFARGO_BC_WALL) {
    wall_boundaries(c);
    if (c->df_on) df_wall_boundaries(c);
  }
  subkeplerian_boundary(c); /* unconditional: src/SideEuler.c:1110-1112 */
  if (c->df_on) df_subkeplerian_boundary(c);
  if (c->p.boundary == DFARGO_BC_OPEN) open_boundary(c);
  if (c->df_on && c->fp.open_inner_dust) df_open_boundary(c);
  if (c->p.boundary == DFARGO_BC_NONREFLECTING) {
    if (c->p.energy_equation) compute_sound_speed(c); /* src/SideEuler.c:1121-1122 */
    nonreflecting_boundary(c);
    subkeplerian_boundary(c);
    if (c->df_on) df_subkeplerian_boundary(c);
  }
}

void or_nonreflecting_boundary(or_ctx *c) { nonreflecting_boundary(c); }

/* EvanescentBoundary (src/SideEuler.c:796-950) for the gas: in the
 * wave-killing zones Rmed < WKZRMIN and Rmed > WKZRMAX each ring relaxes as
 * q <- (q + lambda q0) / (1 + lambda) towards its azimuthal mean (DampToAxi)
 * or the initial profile (DampToIni), lambda = x^2 step / (ratio Rmed^1.5)
 * with x rising from 0 at the zone's edge to 1 at the grid's edge.  With
 * ZEROVRADBC the two innermost and the outermost ring of the zones take
 * vr = 0 and the target density and vtheta. */
static void evanescent_boundary(or_ctx *c, double step) {
  const dfargo_params *p = &c->p;
  const int nr = c->nr, ns = c->ns;
  double lambda = 0.0;
  for (int i = 0; i < nr; i++) {
    const double R = c->Rmed[i];
    if (!((R < p->wkz_rmin) || (R > p->wkz_rmax))) continue;
    if (R < p->wkz_rmin) {
      const double damping = (R - p->wkz_rmin) / (c->Rmed[0] - p->wkz_rmin);
      const double Tin = p->damping_ratio * pow(R, 1.5);
      lambda = damping * damping * step / Tin;
    }
    if (R > p->wkz_rmax) {
      const double damping = (R - p->wkz_rmax) / (c->Rmed[nr - 1] - p->wkz_rmax);
      const double Tout = p->damping_ratio * pow(R, 1.5);
      lambda = damping * damping * step / Tout;
    }
    double vr0, vt0, d0, e0;
    if (p->damp_to == DFARGO_DAMP_TO_VISCOUS) {
      vt0 = e0 = 0.0;
      for (int j = 0; j < ns; j++) {
        vt0 += c->vt[i * ns + j];
        e0 += c->en[i * ns + j];
      }
      vt0 /= (double)ns;
      e0 /= (double)ns;
      int m = 0;
      while (c->rmed1d[m] < R) m++;
      m--;
      d0 = (c->rmed1d[m + 1] - R) * c->sigma1d[m] + (R - c->rmed1d[m]) * c->sigma1d[m + 1];
      d0 /= (c->rmed1d[m + 1] - c->rmed1d[m]);
      vr0 = (c->rmed1d[m + 1] - R) * c->vrad1d[m] + (R - c->rmed1d[m]) * c->vrad1d[m + 1];
      vr0 /= (c->rmed1d[m + 1] - c->rmed1d[m]);
    } else if (p->damp_to == DFARGO_DAMP_TO_INI) {
      vt0 = c->VthetaMed[i] - R * c->OmegaFrame;
      vr0 = c->VradMed[i];
      d0 = c->SigmaMed[i];
      e0 = c->EnergyMed[i];
    } else {
      vr0 = vt0 = d0 = e0 = 0.0;
      for (int j = 0; j < ns; j++) {
        const int l = i * ns + j;
        vr0 += c->vr[l];
        vt0 += c->vt[l];
        d0 += c->rho[l];
        e0 += c->en[l];
      }
      vr0 /= (double)ns;
      vt0 /= (double)ns;
      d0 /= (double)ns;
      e0 /= (double)ns;
    }
    /* the dust fluid's targets (src/SideEuler.c:844-848,887-926): the
       initial profiles, or with DampToAxi its ring means */
    double dvr0 = 0.0, dvt0 = 0.0, dd0 = 0.0;
    const int ddamp = c->df_on && !c->fp.open_inner_dust;
    if (ddamp) {
      if (p->damp_to == DFARGO_DAMP_TO_AXI) {
        for (int j = 0; j < ns; j++) {
          const int l = i * ns + j;
          dvr0 += c->dvr[l];
          dvt0 += c->dvt[l];
          dd0 += c->drho[l];
        }
        dvr0 /= (double)ns;
        dvt0 /= (double)ns;
        dd0 /= (double)ns;
      } else {
        dvt0 = c->DVthetaMed[i] - R * c->OmegaFrame;
        dvr0 = c->DVradMed[i];
        dd0 = c->DSigmaMed[i];
      }
    }
    const int edge = p->zero_vrad_bc && (((R < p->wkz_rmin) && (i == 0 || i == 1)) ||
                                         ((R > p->wkz_rmax) && (i == nr - 1)));
    for (int j = 0; j < ns; j++) {
      const int l = i * ns + j;
      c->vr[l] = (c->vr[l] + lambda * vr0) / (1.0 + lambda);
      c->vt[l] = (c->vt[l] + lambda * vt0) / (1.0 + lambda);
      c->rho[l] = (c->rho[l] + lambda * d0) / (1.0 + lambda);
      if (p->energy_equation) c->en[l] = (c->en[l] + lambda * e0) / (1.0 + lambda);
      if (ddamp) {
        c->dvr[l] = (c->dvr[l] + lambda * dvr0) / (1.0 + lambda);
        c->dvt[l] = (c->dvt[l] + lambda * dvt0) / (1.0 + lambda);
        c->drho[l] = (c->drho[l] + lambda * dd0) / (1.0 + lambda);
      }
      if (edge) {
        c->vr[l] = 0.0;
        c->rho[l] = d0;
        c->vt[l] = vt0;
      }
    }
  }
}

void or_evanescent_boundary(or_ctx *c, double step) { evanescent_boundary(c, step); }

/* AddFloorDensity, AddFloorEnergy (src/SourceEuler.c:2506-2541) */
static void add_floors(or_ctx *c) {
  for (int i = 0; i < c->nr; i++)
    for (int j = 0; j < c->ns; j++) {
      int l = i * c->ns + j;
      if (c->rho[l] < c->p.density_floor) c->rho[l] = c->p.density_floor;
    }
  for (int i = 0; i < c->nr; i++) {
    double e_bottom = 1e-3 * c->EnergyMed[i];
    for (int j = 0; j < c->ns; j++) {
      int l = i * c->ns + j;
      if (c->en[l] < e_bottom) c->en[l] = e_bottom;
    }
  }
}

/* DetectCrash (src/SourceEuler.c:51-69) */
static int detect_crash(or_ctx *c, const double *a) {
  int crashed = 0;
  for (int l = 0; l < c->nr * c->ns; l++)
    if (a[l] < 0.0) crashed = 1;
  return crashed;
}

/* ------------------------------------------------------------------------ */
/* AlgoGas                                                                    */
/* ------------------------------------------------------------------------ */


/* ------------------------------------------------------------------------ */
/* Lagrangian dust particles (src/Dsys.c, src/DustUpdate.c)                  */
/* ------------------------------------------------------------------------ */

/* ComputeCodeUnits (src/LowTasks.c:204-212): unit mass in kg, unit length
 * in m */
static void dust_units(const dfargo_dust_params *d, double *unit_mass, double *unit_length) {
  *unit_mass = 1.9891e30 * d->factor_unit_mass;
  *unit_length = 1.49598e11 * d->factor_unit_length;
}

/* InitDustSystem (src/Dsys.c:128-216), single process: all sizes, then all
 * radii, then all azimuths from one drand48 stream */
int or_dust_init_system(const dfargo_params *p, const dfargo_dust_params *d, const double *axi_sg_accr,
                        double *size, double *r, double *th, double *vr, double *vth, double *l) {
  or_ctx *c = or_create(p);
  if (!c) return -1;
  int nr = c->nr, ns = c->ns, n = d->nbpart;
  double unit_mass, unit_length, C1, C2;
  dust_units(d, &unit_mass, &unit_length);
  double rmindust = (d->rmindust == 0.0) ? p->rmin : d->rmindust;
  double rmaxdust = (d->rmaxdust == 0.0) ? p->rmax : d->rmaxdust;
  unsigned short xs[3];
  xs[0] = 0x330E;
  xs[1] = (unsigned short)(d->seed & 0xFFFF);
  xs[2] = (unsigned short)((d->seed >> 16) & 0xFFFF);
  if (d->sizeminpart != d->sizemaxpart) {
    if (d->sizepartslope != 1.0) {
      C1 = pow(d->sizemaxpart, 1.0 - d->sizepartslope);
      C2 = pow(d->sizeminpart, 1.0 - d->sizepartslope);
      for (int i = 0; i < n; i++)
        size[i] = pow(C2 + (C1 - C2) * erand48(xs), 1.0 / (1.0 - d->sizepartslope)) / unit_length;
    } else {
      for (int i = 0; i < n; i++)
        size[i] = d->sizeminpart * exp(erand48(xs) * log(d->sizemaxpart / d->sizeminpart)) / unit_length;
    }
  } else {
    for (int i = 0; i < n; i++) size[i] = d->sizeminpart / unit_length;
  }
  if (d->dustslope != 1.0) {
    C1 = pow(rmaxdust, 1.0 - d->dustslope);
    C2 = pow(rmindust, 1.0 - d->dustslope);
    for (int i = 0; i < n; i++) r[i] = pow(C2 + (C1 - C2) * erand48(xs), 1.0 / (1.0 - d->dustslope));
  } else {
    for (int i = 0; i < n; i++) r[i] = rmindust * exp(erand48(xs) * log(rmaxdust / rmindust));
  }
  double aziinf0 = c->Azimuth[0] - 0.5 * (p->pmax - p->pmin) / (double)ns;
  double azisup = c->Azimuth[ns - 1] + 0.5 * (p->pmax - p->pmin) / (double)ns;
  for (int i = 0; i < n; i++) th[i] = aziinf0 + erand48(xs) * (azisup - aziinf0);
  for (int i = 0; i < n; i++) {
    vr[i] = 0.0;
    if (!p->self_gravity) {
      vth[i] = pow(r[i], -0.5);
    } else {
      double dist = r[i];
      int ipl = 0;
      while ((c->Rmed[ipl] <= dist) && (ipl < nr - 2)) ipl++;
      double ri = c->Rmed[ipl], rip1 = c->Rmed[ipl + 1], dr = rip1 - ri;
      double sgacc = (dist - ri) * axi_sg_accr[ipl + 1] + (rip1 - dist) * axi_sg_accr[ipl];
      sgacc /= dr;
      vth[i] = pow(r[i], -0.5) * sqrt(1.0 - dist * dist * sgacc);
    }
    l[i] = r[i] * vth[i];
  }
  or_destroy(c);
  return 0;
}

int or_dust_attach(or_ctx *c, const dfargo_dust_params *d, const double *size, const double *r,
                   const double *th, const double *vr, const double *vth, const double *l, double timestep) {
  int n = d->nbpart, ns = c->ns;
  c->dust_on = 1;
  c->dp = *d;
  c->np = n;
#define DA(x) c->x = (double *)calloc((size_t)n + 1, sizeof(double))
  DA(d_size); DA(d_size_init); DA(d_r); DA(d_th); DA(d_vr); DA(d_vth); DA(d_l);
  DA(d_vrg); DA(d_vtg); DA(d_densg); DA(d_csg); DA(d_radgradp); DA(d_azigradp);
  DA(d_radind); DA(d_aziind); DA(d_radsg); DA(d_azisg); DA(d_raddrag); DA(d_azidrag); DA(d_stokes);
#undef DA
  memcpy(c->d_size, size, sizeof(double) * n);
  memcpy(c->d_size_init, size, sizeof(double) * n);
  memcpy(c->d_r, r, sizeof(double) * n);
  memcpy(c->d_th, th, sizeof(double) * n);
  memcpy(c->d_vr, vr, sizeof(double) * n);
  memcpy(c->d_vth, vth, sizeof(double) * n);
  memcpy(c->d_l, l, sizeof(double) * n);
  c->AziInf = (double *)calloc((size_t)ns, sizeof(double));
  c->AziSup = (double *)calloc((size_t)ns, sizeof(double));
  for (int j = 0; j < ns; j++) {
    c->AziInf[j] = c->Azimuth[j] - 0.5 * (c->p.pmax - c->p.pmin) / (double)ns;
    c->AziSup[j] = c->Azimuth[j] + 0.5 * (c->p.pmax - c->p.pmin) / (double)ns;
  }
  dust_units(d, &c->unit_mass, &c->unit_length);
  c->particles_mass_initial = d->particle_mass;
  c->particles_mass = d->particle_mass;
  /* start-up interpolation (src/main.c:287), with the sound speed of
     InitEuler */
  c->min_stopping_time = 1e6;
  or_dust_interpolation(c, timestep);
  return 0;
}

/* weights of `interpolation` for the cell at distances (dr, dphi); kind is
 * the stencil: 0 the vr one (radial NGP cut), 1 the vtheta one (azimuthal
 * NGP cut), 2 the cell-centred one (src/Dsys.c:407-441,490-524,640-665) */
static void dust_weights(int scheme, int kind, double dr, double dphi, double delta_r, double delta_phi,
                         double *wr, double *wt) {
  if (scheme == DFARGO_INTERP_NGP) {
    *wr = 1.0;
    *wt = 1.0;
    if (kind == 0) *wr = (dr < 0.5 * delta_r) ? 1.0 : 0.0;
    if (kind == 1) *wt = (dphi < 0.5 * delta_phi) ? 1.0 : 0.0;
  } else if (scheme == DFARGO_INTERP_CIC) {
    *wr = 1.0 - dr / delta_r;
    *wt = 1.0 - dphi / delta_phi;
  } else {
    if (dr < 0.5 * delta_r) *wr = 0.75 - dr * dr / delta_r / delta_r;
    if ((dr >= 0.5 * delta_r) && (dr <= 1.5 * delta_r)) *wr = 0.5 * (1.5 - dr / delta_r) * (1.5 - dr / delta_r);
    if (dr > 1.5 * delta_r) *wr = 0.0;
    if (dphi < 0.5 * delta_phi) *wt = 0.75 - dphi * dphi / delta_phi / delta_phi;
    if ((dphi >= 0.5 * delta_phi) && (dphi <= 1.5 * delta_phi))
      *wt = 0.5 * (1.5 - dphi / delta_phi) * (1.5 - dphi / delta_phi);
    if (dphi > 1.5 * delta_phi) *wt = 0.0;
  }
}

/* index ranges of the three stencils for the particle in cell (ip, jp) */
static void dust_stencil(or_ctx *c, int scheme, int kind, int ip, int jp, double rp, double tp, int *imin,
                         int *imax, int *jmin, int *jmax) {
  int myip = ip, myjp = jp;
  if (scheme == DFARGO_INTERP_NGP) {
    *imin = ip; *imax = ip + (kind == 0 ? 1 : 0);
    *jmin = jp; *jmax = jp + (kind == 1 ? 1 : 0);
  } else if (scheme == DFARGO_INTERP_CIC) {
    if (kind != 1) myjp = (tp < c->Azimuth[jp]) ? jp - 1 : jp;
    if (kind != 0) myip = (rp < c->Rmed[ip]) ? ip - 1 : ip;
    *imin = myip; *imax = myip + 1;
    *jmin = myjp; *jmax = myjp + 1;
  } else {
    if (kind == 1) myjp = (tp > c->Azimuth[jp]) ? jp + 1 : jp;
    *imin = ip - 1; *imax = ip + 1;
    *jmin = myjp - 1; *jmax = myjp + 1;
  }
}

/* wrapped azimuthal index and the stencil's reference azimuth of column j */
static int dust_col(or_ctx *c, int kind, int j, double delta_phi, double *az) {
  int ns = c->ns;
  const double *A = (kind == 1) ? c->AziInf : c->Azimuth;
  if (j >= 0 && j < ns) {
    *az = A[j];
    return j;
  }
  if (j < 0) {
    *az = A[0] + j * delta_phi;
    return j + ns;
  }
  *az = A[ns - 1] + (j + 1 - ns) * delta_phi;
  return j - ns;
}

/* radial distance of the stencil: to Rinf for the vr stencil, Rmed else */
static double dust_dr(or_ctx *c, int kind, int i, double rp) {
  if (i < 0 || i > c->nr - 1) return 0.0;
  double ref = (kind == 0) ? c->Rinf[i] : c->Rmed[i];
  return c->p.log_grid ? fabs(log(rp / ref)) : fabs(rp - ref);
}

/* interpolation (src/Dsys.c:265-1149), single process: gas values at the
 * particles, Stokes numbers and drag accelerations, feedback grids and the
 * dust density.  vtheta is read in the fixed frame: the gas grid holds
 * vt + Rmed*OmegaFrame while the particles are processed and vt again
 * afterwards (src/Dsys.c:316,1082), the round trip leaving the rounding of
 * both operations in the gas grid. */
void or_dust_interpolation(or_ctx *c, double timestep) {
  const dfargo_params *p = &c->p;
  const dfargo_dust_params *d = &c->dp;
  int nr = c->nr, ns = c->ns, scheme = d->interpolation;
  const double *vr = c->vr, *vt = c->vt, *dens = c->rho, *cs = c->cs;
  double dust_density_codeunits = d->rhopart * 1000.0 * pow(c->unit_length, 3.) / c->unit_mass;
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) c->vt[j + i * ns] += c->Rmed[i] * c->OmegaFrame;
  for (int l = 0; l < nr * ns; l++) {
    if (d->dust_feedback) {
      c->radfbacc[l] = 0.0;
      c->azifbacc[l] = 0.0;
      c->fbedot[l] = 0.0;
    }
    c->dustdens[l] = 0.0;
  }
  double delta_r = p->log_grid ? log(c->Rsup[0] / c->Rinf[0]) : c->Rsup[0] - c->Rinf[0];
  double delta_phi = c->AziSup[0] - c->AziInf[0];
  c->min_stopping_time = 1e6;
  for (int k = 0; k < c->np; k++) {
    c->d_vrg[k] = c->d_vtg[k] = c->d_radgradp[k] = c->d_azigradp[k] = 0.0;
    c->d_radind[k] = c->d_radsg[k] = c->d_aziind[k] = c->d_azisg[k] = 0.0;
    c->d_densg[k] = c->d_csg[k] = 0.0;
    double rp = c->d_r[k];
    if (!((rp >= c->Rinf[0]) && (rp < c->Rsup[nr - 1]))) continue;
    int ip = 0;
    while (rp >= c->Rinf[ip]) ip++;
    ip--;
    double tp = c->d_th[k];
    int jp = (int)floor(ns * (tp - c->AziInf[0]) / (c->AziSup[ns - 1] - c->AziInf[0]));
    if (jp == ns) jp = 0;
    int imin, imax, jmin, jmax;
    double wr = 0.0, wt = 0.0, az;
    /* 2a) vr stencil */
    dust_stencil(c, scheme, 0, ip, jp, rp, tp, &imin, &imax, &jmin, &jmax);
    for (int j = jmin; j <= jmax; j++) {
      int myj = dust_col(c, 0, j, delta_phi, &az);
      double dphi = fabs(tp - az);
      for (int i = imin; i <= imax; i++) {
        dust_weights(scheme, 0, dust_dr(c, 0, i, rp), dphi, delta_r, delta_phi, &wr, &wt);
        if (i >= 0 && i <= nr - 1) {
          int l = i * ns + myj;
          c->d_vrg[k] += (wr * wt * vr[l]);
          c->d_radgradp[k] += (wr * wt * c->radgradp[l]);
          if (p->indirect_term) c->d_radind[k] += (wr * wt * c->radindacc[l]);
          if (p->self_gravity) c->d_radsg[k] += (wr * wt * c->radsgacc[l]);
        }
      }
    }
    /* 2b) vtheta stencil */
    dust_stencil(c, scheme, 1, ip, jp, rp, tp, &imin, &imax, &jmin, &jmax);
    for (int j = jmin; j <= jmax; j++) {
      int myj = dust_col(c, 1, j, delta_phi, &az);
      double dphi = fabs(tp - az);
      for (int i = imin; i <= imax; i++) {
        dust_weights(scheme, 1, dust_dr(c, 1, i, rp), dphi, delta_r, delta_phi, &wr, &wt);
        if (i >= 0 && i <= nr - 1) {
          int l = i * ns + myj;
          c->d_vtg[k] += (wr * wt * vt[l]);
          c->d_azigradp[k] += (wr * wt * c->azigradp[l]);
          if (p->indirect_term) c->d_aziind[k] += (wr * wt * c->aziindacc[l]);
          if (p->self_gravity) c->d_azisg[k] += (wr * wt * c->azisgacc[l]);
        }
      }
    }
    /* 2c) cell-centred stencil: density, sound speed, dust density */
    dust_stencil(c, scheme, 2, ip, jp, rp, tp, &imin, &imax, &jmin, &jmax);
    for (int j = jmin; j <= jmax; j++) {
      int myj = dust_col(c, 2, j, delta_phi, &az);
      double dphi = fabs(tp - az);
      for (int i = imin; i <= imax; i++) {
        dust_weights(scheme, 2, dust_dr(c, 2, i, rp), dphi, delta_r, delta_phi, &wr, &wt);
        if (i >= 0 && i <= nr - 1) {
          int l = i * ns + myj;
          c->d_densg[k] += (wr * wt * dens[l]);
          c->d_csg[k] += (wr * wt * cs[l]);
          c->dustdens[l] += wr * wt * c->particles_mass / c->Surf[i];
        }
      }
    }
    /* 2d) Stokes number (Paardekooper 2007) */
    double vrp = c->d_vr[k], vtp = c->d_vth[k], vrg = c->d_vrg[k], vtg = c->d_vtg[k];
    double csg = c->d_csg[k], densg = c->d_densg[k];
    double densg_cgs = 0.1 * densg * c->unit_mass * pow(c->unit_length, -2.0);
    double lambda = 3.34e-8 * pow(densg_cgs, -1.) * aspect_ratio(p, rp) * rp;
    double Kn = 0.5 * lambda / c->d_size[k];
    double Mach = sqrt((vrp - vrg) * (vrp - vrg) + (vtp - vtg) * (vtp - vtg)) / csg;
    double f_D = sqrt(1.0 + 9.0 * M_PI * Mach * Mach / 128.0);
    double Re = 3.0 * sqrt(M_PI / 8.0) * Mach / Kn;
    double k_D = 0.0;
    if (Re <= 500.0)
      k_D = 1.0 + 0.15 * pow(Re, 0.687);
    else {
      if ((Re > 500.0) && (Re <= 1500.0)) k_D = 3.96e-6 * pow(Re, 2.4);
      if (Re > 1500.0) k_D = 0.11 * Re;
    }
    double Cdrag = pow(3.0 * Kn + 1.0, 2.0) * pow(9.0 * Kn * Kn * f_D + 3.0 * Kn * k_D, -1.0);
    double St = 0.5 * M_PI * Cdrag * c->d_size[k] * dust_density_codeunits / densg;
    double Ts = St * pow(c->d_r[k], 1.5);
    if (Ts < c->min_stopping_time) c->min_stopping_time = Ts;
    c->d_stokes[k] = St;
    /* 3) drag acceleration */
    if ((Ts < timestep) && d->short_friction_time_approximation) {
      vrp = vrg + Ts * c->d_radgradp[k] / densg;
      vtp = vtg + Ts * c->d_azigradp[k] / densg;
    }
    c->d_raddrag[k] = -(vrp - vrg) / Ts;
    c->d_azidrag[k] = -(vtp - vtg) / Ts;
    if (!d->dust_feedback) continue;
    double mp = c->particles_mass;
    /* 4a) radial feedback, vr stencil */
    dust_stencil(c, scheme, 0, ip, jp, rp, tp, &imin, &imax, &jmin, &jmax);
    for (int j = jmin; j <= jmax; j++) {
      int myj = dust_col(c, 0, j, delta_phi, &az);
      double dphi = fabs(tp - az);
      for (int i = imin; i <= imax; i++) {
        dust_weights(scheme, 0, dust_dr(c, 0, i, rp), dphi, delta_r, delta_phi, &wr, &wt);
        if (i >= 0 && i <= nr - 1) {
          int l = i * ns + myj;
          double mg = dens[l] * c->Surf[i];
          c->radfbacc[l] -= (wr * wt * c->d_raddrag[k]) * mp / mg;
        }
      }
    }
    /* 4b) azimuthal feedback, vtheta stencil */
    dust_stencil(c, scheme, 1, ip, jp, rp, tp, &imin, &imax, &jmin, &jmax);
    for (int j = jmin; j <= jmax; j++) {
      int myj = dust_col(c, 1, j, delta_phi, &az);
      double dphi = fabs(tp - az);
      for (int i = imin; i <= imax; i++) {
        dust_weights(scheme, 1, dust_dr(c, 1, i, rp), dphi, delta_r, delta_phi, &wr, &wt);
        if (i >= 0 && i <= nr - 1) {
          int l = i * ns + myj;
          double mg = dens[l] * c->Surf[i];
          c->azifbacc[l] -= (wr * wt * c->d_azidrag[k]) * mp / mg;
        }
      }
    }
    /* 4c) heating rate, cell-centred stencil */
    if (p->energy_equation) {
      dust_stencil(c, scheme, 2, ip, jp, rp, tp, &imin, &imax, &jmin, &jmax);
      for (int j = jmin; j <= jmax; j++) {
        int myj = dust_col(c, 2, j, delta_phi, &az);
        double dphi = fabs(tp - az);
        for (int i = imin; i <= imax; i++) {
          dust_weights(scheme, 2, dust_dr(c, 2, i, rp), dphi, delta_r, delta_phi, &wr, &wt);
          if (i >= 0 && i <= nr - 1) {
            int l = i * ns + myj;
            double rhop = mp / c->Surf[i];
            double dv2 = (vrp - vrg) * (vrp - vrg) + (vtp - vtg) * (vtp - vtg);
            c->fbedot[l] += rhop * wr * wt * dv2 / Ts;
          }
        }
      }
    }
  }
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) c->vt[j + i * ns] -= c->Rmed[i] * c->OmegaFrame;
}

/* keep an azimuth in [AziInf[0], AziSup[ns-1]] (src/DustUpdate.c:38-45) */
static double dust_wrap(or_ctx *c, double th) {
  int ns = c->ns;
  if (th < c->AziInf[0]) th += (c->AziSup[ns - 1] - c->AziInf[0]);
  if (th > c->AziSup[ns - 1]) th -= (c->AziSup[ns - 1] - c->AziInf[0]);
  return th;
}

/* SemiUpdateDustPositions (src/DustUpdate.c:6-138), single process, no
 * turbulent kicks and no Hill-sphere removal */
static void dust_semi_update_positions(or_ctx *c, double timestep) {
  int nr = c->nr;
  for (int k = 0; k < c->np; k++) {
    double r = c->d_r[k];
    if (!((r >= c->Rinf[0]) && (r < c->Rsup[nr - 1]))) continue;
    c->d_r[k] += c->d_vr[k] * 0.5 * timestep;
    c->d_th[k] += c->d_vth[k] * r * 0.25 * timestep * (pow(r, -2.) + pow(c->d_r[k], -2.));
    c->d_th[k] = dust_wrap(c, c->d_th[k]);
    c->d_th[k] = dust_wrap(c, c->d_th[k]);
  }
}

/* UpdateDustVelocities (src/DustUpdate.c:267-435), semi-implicit
 * integrator or the short-friction-time limit */
static void dust_update_velocities(or_ctx *c, double timestep) {
  const dfargo_params *p = &c->p;
  const dfargo_dust_params *d = &c->dp;
  int nr = c->nr;
  or_dust_interpolation(c, timestep);
  for (int k = 0; k < c->np; k++) {
    double rd = c->d_r[k];
    if (!((rd >= c->Rinf[0]) && (rd < c->Rsup[nr - 1]))) continue;
    double td = c->d_th[k], vrd = c->d_vr[k], vtd = c->d_vth[k];
    double vrg = c->d_vrg[k], vtg = c->d_vtg[k], St = c->d_stokes[k];
    int local_feel_disk = d->dust_feel_disk;
    if ((rd < p->rmin) || (rd > p->rmax)) local_feel_disk = 0;
    double stoptime = St * pow(rd, 1.5);
    if ((stoptime < timestep) && d->short_friction_time_approximation) {
      c->d_vr[k] = vrg + stoptime * c->d_radgradp[k] / c->d_densg[k];
      c->d_vth[k] = vtg + stoptime * c->d_azigradp[k] / c->d_densg[k];
      c->d_l[k] = rd * c->d_vth[k];
    } else {
      double FgravR = -pow(rd, -2.), FgravTh = 0.;
      if ((fabs(FgravR + vtd * vtd / rd) < 1e-15) && (!d->dust_feel_planets) && (!d->dust_feel_disk) &&
          (!d->dust_feel_sg))
        FgravR = -vtd * vtd / rd;
      if (d->dust_feel_planets) {
        for (int i = 0; i < c->sys.nb; i++) {
          double xp = c->sys.x[i], yp = c->sys.y[i];
          double rp = sqrt(xp * xp + yp * yp), tp = atan2(yp, xp), mp = c->sys.mass[i];
          double eps = p->thickness_smoothing * p->aspect_ratio * pow(0.5 * rp, 1.0 + p->flaring_index) *
                       pow(2.0, 1.0 + p->flaring_index);
          double dd = sqrt(rd * rd + rp * rp - 2.0 * rd * rp * cos(td - tp) + eps * eps);
          FgravR += mp * (rp * cos(td - tp) - rd) * pow(dd, -3.0);
          FgravTh += -mp * rp * sin(td - tp) * pow(dd, -3.0);
        }
      }
      if (p->indirect_term) {
        FgravR += c->d_radind[k];
        FgravTh += c->d_aziind[k];
      }
      if (p->self_gravity && d->dust_feel_sg) {
        FgravR += c->d_radsg[k];
        FgravTh += c->d_azisg[k];
      }
      double den = 1.0 + local_feel_disk * timestep / stoptime;
      double numTh = (vtd - vtg) + timestep * (FgravTh - vrd * vtd / rd);
      c->d_vth[k] = vtg + numTh / den;
      double numR = (vrd - vrg) + timestep * (FgravR + (c->d_vth[k] * c->d_vth[k]) / rd);
      c->d_vr[k] = vrg + numR / den;
    }
  }
}

/* RotateDsys (src/Dsys.c:1153-1169) */
static void dust_rotate(or_ctx *c, double angle) {
  for (int k = 0; k < c->np; k++) c->d_th[k] = dust_wrap(c, c->d_th[k] - angle);
}

void or_dust_get(or_ctx *c, double *r, double *th, double *vr, double *vth, double *l, double *stokes) {
  size_t n = sizeof(double) * (size_t)c->np;
  if (r) memcpy(r, c->d_r, n);
  if (th) memcpy(th, c->d_th, n);
  if (vr) memcpy(vr, c->d_vr, n);
  if (vth) memcpy(vth, c->d_vth, n);
  if (l) memcpy(l, c->d_l, n);
  if (stokes) memcpy(stokes, c->d_stokes, n);
}

void or_dust_get_gas_at_particles(or_ctx *c, double *vrg, double *vtg, double *densg, double *csg) {
  size_t n = sizeof(double) * (size_t)c->np;
  if (vrg) memcpy(vrg, c->d_vrg, n);
  if (vtg) memcpy(vtg, c->d_vtg, n);
  if (densg) memcpy(densg, c->d_densg, n);
  if (csg) memcpy(csg, c->d_csg, n);
}

/* the particle part of one sub-step over a frozen gas state, for pinning:
 * positions, velocities, and RotateDsys by OmegaFrame dt (a wrap only in
 * the fixed frame) */
void or_dust_step(or_ctx *c, double dt) {
  dust_semi_update_positions(c, 2. * dt);
  dust_update_velocities(c, dt);
  dust_rotate(c, c->OmegaFrame * dt);
}

void or_dust_get_drag(or_ctx *c, double *raddrag, double *azidrag) {
  size_t n = sizeof(double) * (size_t)c->np;
  if (raddrag) memcpy(raddrag, c->d_raddrag, n);
  if (azidrag) memcpy(azidrag, c->d_azidrag, n);
}

double or_dust_min_stopping_time(or_ctx *c) { return c->min_stopping_time; }
double or_dust_particles_mass(or_ctx *c) { return c->particles_mass; }

/* ------------------------------------------------------------------------ */
/* the dust fluid (DUSTFLUID YES)                                             */
/* ------------------------------------------------------------------------ */

/* DAspectRatio (src/Viscosity.c:123-138), LAMBDADOUBLING = 0 */
static double daspect_ratio(const or_ctx *c, double rad) {
  const dfargo_params *p = &c->p;
  double ar = c->fp.daspect_ratio;
  double rmin = p->transition_radius - p->transition_width * c->fp.daspect_ratio;
  double rmax = p->transition_radius + p->transition_width * c->fp.daspect_ratio;
  if (rad < rmin) ar *= p->transition_ratio;
  if ((rad >= rmin) && (rad <= rmax)) ar *= exp((rmax - rad) / (rmax - rmin) * log(p->transition_ratio));
  return ar;
}

/* DFViscosity (src/Viscosity.c:72-104), RELEASEDATEVISCOSITY = 0 */
static double dfviscosity(const or_ctx *c, double rad) {
  const dfargo_params *p = &c->p;
  double viscosity = c->fp.dviscosity;
  int i = 0;
  if (c->fp.dalpha_viscosity != 0.0) {
    while (c->GlobalRmed[i] < rad) i++;
    viscosity = c->fp.dalpha_viscosity * c->GlobalDCs[i] * c->GlobalDCs[i] * pow(rad, 1.5);
  }
  double rmin = p->cavity_radius - p->cavity_width * c->fp.daspect_ratio;
  double rmax = p->cavity_radius + p->cavity_width * c->fp.daspect_ratio;
  if (rad < rmin) viscosity *= p->cavity_ratio;
  if ((rad >= rmin) && (rad <= rmax)) viscosity *= exp((rmax - rad) / (rmax - rmin) * log(p->cavity_ratio));
  return viscosity;
}

static int df_has_viscosity(const or_ctx *c) {
  return c->df_on && ((c->fp.dalpha_viscosity != 0.0) || (c->fp.dviscosity != 0.0));
}

/* DSigma (src/Theo.c:66-85), no RestartWithNewDust */
static double dsigma(const or_ctx *c, double r) { return sigma_profile(&c->p, r) * c->fp.dust_to_gas_density_ratio; }

/* ComputeDustSoundSpeedAndPressure (src/SourceEuler.c:2353-2371); the
   sound speed is one value per ring.  Its azimuthal mean
   (mpi_make1Dprofile, GLOBAL_DustSoundSpeed of DFViscosity) is summed
   cell by cell as the reference sums it. */
static void df_sound_speed_pressure(or_ctx *c) {
  for (int i = 0; i < c->nr; i++) {
    const double r = c->Rmed[i];
    c->dcs1[i] = daspect_ratio(c, r) * sqrt(G * 1.0 / r) * pow(r, c->fp.dflaring_index);
    double s = 0.0;
    for (int j = 0; j < c->ns; j++) {
      const int l = i * c->ns + j;
      c->dpres[l] = c->drho[l] * c->dcs1[i] * c->dcs1[i];
      s += c->dcs1[i];
    }
    c->GlobalDCs[i] = s / (double)c->ns;
  }
}

/* the code units of ComputeCodeUnits (src/LowTasks.c:204-212) and the
   grain's internal density and radius in them (src/SourceEuler.c:812-818) */
static void df_grain(const or_ctx *c, double *rho_int, double *radius, double *to_cgs) {
  const double um = 1.9891e30 * c->fp.factor_unit_mass, ul = 1.49598e11 * c->fp.factor_unit_length;
  *rho_int = c->fp.rhopart * 1000.0 * pow(ul, 3.) / um;
  *radius = c->fp.sizepart / ul;
  *to_cgs = 0.1 * um * pow(ul, -2.0);
}

/* InitDustStokesNumber (src/Pframeforce.c:836-874): zero relative velocity
   (f_D = k_D = 1) */
static void df_init_stokes(or_ctx *c) {
  double rho_int, radius, to_cgs;
  df_grain(c, &rho_int, &radius, &to_cgs);
  const double um = 1.9891e30 * c->fp.factor_unit_mass, ul = 1.49598e11 * c->fp.factor_unit_length;
  for (int i = 0; i < c->nr; i++)
    for (int j = 0; j < c->ns; j++) {
      const int l = j + i * c->ns;
      double densg_cgs = 0.1 * c->rho[l] * um * pow(ul, -2.0);
      double lambda = 3.34e-8 * pow(densg_cgs, -1.) * aspect_ratio(&c->p, c->Rmed[i]) * c->Rmed[i];
      double Kn = 0.5 * lambda / radius, f_D = 1.0, k_D = 1.0;
      double Cdrag = pow(3.0 * Kn + 1.0, 2.0) * pow(9.0 * Kn * Kn * f_D + 3.0 * Kn * k_D, -1.0);
      c->stokes[l] = 0.5 * M_PI * Cdrag * rho_int * radius / c->rho[l];
    }
}

/* The Stokes number of SubStep1 (src/SourceEuler.c:820-858): the
   Paardekooper (2007) drag coefficient from the Knudsen number, the
   relative Mach number and the Reynolds number, evaluated on the state
   the sub-step starts from */
static void df_stokes(or_ctx *c) {
  double rho_int, radius, to_cgs;
  df_grain(c, &rho_int, &radius, &to_cgs);
  for (int i = 0; i < c->nr; i++)
    for (int j = 0; j < c->ns; j++) {
      const int l = j + i * c->ns;
      const double densg_cgs = c->rho[l] * to_cgs;
      const double lambda = 3.34e-8 * pow(densg_cgs, -1.) * aspect_ratio(&c->p, c->Rmed[i]) * c->Rmed[i];
      const double Kn = 0.5 * lambda / radius;
      const double dvr = c->dvr[l] - c->vr[l], dvt = c->dvt[l] - c->vt[l];
      const double Mach = sqrt(dvr * dvr + dvt * dvt) / c->cs[l];
      const double f_D = sqrt(1.0 + 9.0 * M_PI * Mach * Mach / 128.0);
      const double Re = 3.0 * sqrt(M_PI / 8.0) * Mach / Kn;
      double k_D = 0.0;
      if (Re <= 500.0)
        k_D = 1.0 + 0.15 * pow(Re, 0.687);
      else {
        if ((Re > 500.0) && (Re <= 1500.0)) k_D = 3.96e-6 * pow(Re, 2.4);
        if (Re > 1500.0) k_D = 0.11 * Re;
      }
      const double Cdrag = (3.0 * Kn + 1.0) * (3.0 * Kn + 1.0) / (9.0 * Kn * Kn * f_D + 3.0 * Kn * k_D);
      c->stokes[l] = 0.5 * M_PI * Cdrag * rho_int * radius / c->rho[l];
    }
}

/* InitDustVelocities (src/Pframeforce.c:644-830) as src/SourceEuler.c:268
   calls it, with the gas and dust densities in each other's places: in the
   feedback branch its eps is gas over dust.  Every earlier assignment of
   the routine is overwritten by its final loops, which alone are
   restated. */
static void df_init_velocities(or_ctx *c) {
  const dfargo_params *p = &c->p;
  const int nr = c->nr, ns = c->ns;
  for (int i = 0; i < nr; i++) {
    const double r = c->Rmed[i];
    double vk = sqrt(G * 1.0 / r);
    const double eta = -0.5 * pow(p->aspect_ratio, 2.0) * pow(r, 2.0 * p->flaring_index) *
                       (-1. - p->sigma_slope + 2.0 * p->flaring_index);
    if (p->self_gravity) vk = sqrt(G * 1.0 / r - r * c->AxiSG[i]);
    for (int j = 0; j < ns; j++) {
      const int l = i * ns + j;
      const double St = c->stokes[l];
      if (c->fp.dust_feedback) {
        const double eps = c->rho[l] / c->drho[l];
        c->dvr[l] = -2.0 * eta * vk / (St + pow(1.0 + eps, 2.0) / St);
        c->dvt[l] = vk - eta * vk * (1.0 + eps) / (pow(St, 2.0) + pow(1.0 + eps, 2.0));
      } else {
        c->dvr[l] = -2.0 * eta * vk / (St + pow(St, -1.0));
        c->dvt[l] = vk - eta * vk / (pow(St, 2.0) + 1.0);
      }
      c->dvt[l] -= c->OmegaFrame * r;
    }
  }
  for (int i = 0; i < nr; i++) {
    c->DVthetaMed[i] = c->dvt[i * ns] + c->Rmed[i] * c->OmegaFrame;
    c->DVradMed[i] = c->dvr[i * ns];
  }
}

/* the gas velocities of InitGasVelocities with dust feedback
   (src/Pframeforce.c:609-634) */
static void df_init_gas_feedback(or_ctx *c) {
  const dfargo_params *p = &c->p;
  for (int i = 0; i < c->nr; i++) {
    const double r = c->Rmed[i];
    double vk = sqrt(G * 1.0 / r);
    const double eta = -0.5 * pow(p->aspect_ratio, 2.0) * pow(r, 2.0 * p->flaring_index) *
                       (-1. - p->sigma_slope + 2.0 * p->flaring_index);
    if (p->self_gravity) vk = sqrt(G * 1.0 / r - r * c->AxiSG[i]);
    for (int j = 0; j < c->ns; j++) {
      const int l = i * c->ns + j;
      const double eps = c->drho[l] / c->rho[l], St = c->stokes[l];
      c->vr[l] = 2.0 * eps * eta * vk * St / (St * St + pow(1.0 + eps, 2.0));
      c->vt[l] = vk * (1.0 - eta) + eps * eta * vk * (1.0 + eps) / (pow(St, 2.0) + pow(1.0 + eps, 2.0));
      c->vt[l] -= c->OmegaFrame * r;
    }
  }
}

/* The dust terms of ComputeViscousTerms (src/Viscosity.c:159-298) on the
   dust velocities vr, vt: the divergence of ComputeDivergenceVelocity
   (src/SourceEuler.c:2373-2412), the strain rates, then the stresses with
   DFViscosity */
static void df_viscous_terms(or_ctx *c, const double *vr, const double *vt) {
  const int nr = c->nr, ns = c->ns;
  const double dphi = (c->p.pmax - c->p.pmin) / (double)ns, invdphi = 1.0 / dphi, onethird = 1.0 / 3.0;
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      const int l = j + i * ns, lip = l + ns, ljp = (j == ns - 1) ? i * ns : l + 1;
      c->ddivv[l] = (vr[lip] * c->Rsup[i] - vr[l] * c->Rinf[i]) * c->InvDiffRsup[i] * c->InvRmed[i];
      c->ddivv[l] += (vt[ljp] - vt[l]) * invdphi * c->InvRmed[i];
      c->ddrr[l] = (vr[lip] - vr[l]) * c->InvDiffRsup[i];
      c->ddpp[l] = (vt[ljp] - vt[l]) * invdphi * c->InvRmed[i] + 0.5 * (vr[lip] + vr[l]) * c->InvRmed[i];
    }
  for (int i = 1; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      const int l = j + i * ns, ljm = (j == 0) ? i * ns + ns - 1 : l - 1, lim = l - ns;
      c->ddrp[l] = 0.5 * (c->Rinf[i] * (vt[l] * c->InvRmed[i] - vt[lim] * c->InvRmed[i - 1]) * c->InvDiffRmed[i] +
                          (vr[l] - vr[ljm]) * invdphi * c->InvRinf[i]);
    }
  for (int i = 0; i < nr; i++) {
    const double nu = df_has_viscosity(c) ? dfviscosity(c, c->Rmed[i]) : 0.0;
    for (int j = 0; j < ns; j++) {
      const int l = j + i * ns;
      c->dtrr[l] = 2.0 * c->drho[l] * nu * (c->ddrr[l] - onethird * c->ddivv[l]);
      c->dtpp[l] = 2.0 * c->drho[l] * nu * (c->ddpp[l] - onethird * c->ddivv[l]);
    }
  }
  for (int i = 1; i < nr; i++) {
    const double nu = df_has_viscosity(c) ? dfviscosity(c, c->Rmed[i]) : 0.0;
    for (int j = 0; j < ns; j++) {
      const int l = j + i * ns, lim = l - ns, ljm = (j == 0) ? i * ns + ns - 1 : l - 1, ljmim = ljm - ns;
      c->dtrp[l] = 2.0 * 0.25 * (c->drho[l] + c->drho[lim] + c->drho[ljm] + c->drho[ljmim]) * nu * c->ddrp[l];
    }
  }
}

/* the dust terms of UpdateVelocitiesWithViscosity (src/Viscosity.c:301-380) */
static void df_update_viscosity(or_ctx *c, double *vr, double *vt, double DeltaT) {
  const int nr = c->nr, ns = c->ns;
  const double invdphi = 1.0 / ((c->p.pmax - c->p.pmin) / (double)ns);
  const double *Trr = c->dtrr, *Trp = c->dtrp, *Tpp = c->dtpp, *rho = c->drho;
  for (int i = 1; i < nr - 1; i++)
    for (int j = 0; j < ns; j++) {
      const int l = j + i * ns, lip = l + ns, ljm = (j == 0) ? i * ns + ns - 1 : l - 1;
      vt[l] += DeltaT * c->InvRmed[i] *
               ((c->Rsup[i] * Trp[lip] - c->Rinf[i] * Trp[l]) * c->InvDiffRsup[i] + (Tpp[l] - Tpp[ljm]) * invdphi +
                0.5 * (Trp[l] + Trp[lip])) /
               (0.5 * (rho[l] + rho[ljm]));
    }
  for (int i = 1; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      const int l = j + i * ns, lim = l - ns, ljp = (j == ns - 1) ? i * ns : l + 1;
      vr[l] += DeltaT * c->InvRinf[i] *
               ((c->Rmed[i] * Trr[l] - c->Rmed[i - 1] * Trr[lim]) * c->InvDiffRmed[i] + (Trp[ljp] - Trp[l]) * invdphi -
                0.5 * (Tpp[l] + Tpp[lim])) /
               (0.5 * (rho[l] + rho[lim]));
    }
}

/* the dust terms of SubStep2 (src/SourceEuler.c:1035-1116): artificial
   viscosity on DVradInt, DVthetaInt into DVradNew, DVthetaNew */
static void df_substep2(or_ctx *c, double dt) {
  const int nr = c->nr, ns = c->ns;
  const double *drho = c->drho, *dvrad = c->dvradint, *dvtheta = c->dvthetaint;
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      const int l = j + i * ns, lip = l + ns, ljp = (j == ns - 1) ? i * ns : l + 1;
      double ddv = dvrad[lip] - dvrad[l];
      c->dqr[l] = (ddv < 0.0) ? CVNR * CVNR * drho[l] * ddv * ddv : 0.0;
      ddv = dvtheta[ljp] - dvtheta[l];
      c->dqt[l] = (ddv < 0.0) ? CVNR * CVNR * drho[l] * ddv * ddv : 0.0;
    }
  for (int i = 1; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      const int l = j + i * ns, lim = l - ns;
      c->dvradnew[l] =
          dvrad[l] - dt * 2.0 / (drho[l] + drho[lim]) * (c->dqr[l] - c->dqr[lim]) * c->InvDiffRmed[i];
    }
  for (int i = 0; i < nr; i++) {
    const double dxtheta = (c->p.pmax - c->p.pmin) / (double)ns * c->Rmed[i], invdxtheta = 1.0 / dxtheta;
    for (int j = 0; j < ns; j++) {
      const int l = j + i * ns, ljm = (j == 0) ? i * ns + ns - 1 : l - 1;
      c->dvthetanew[l] = dvtheta[l] - dt * 2.0 / (drho[l] + drho[ljm]) * (c->dqt[l] - c->dqt[ljm]) * invdxtheta;
    }
  }
}

/* Transportd (src/TransportEuler.c:553-568): the gas routines on the dust
   fields, whose static work grids they share.  In the short-friction-time
   approximation OneWindRadd and QuantitiesAdvectiond skip the momenta
   (src/TransportEuler.c:575,778): they only take the integer shift, and
   ComputeVelocities then divides them by the transported density.
   VanLeerRadialdd, VanLeerThetad and VanLeerRadiald compute what
   VanLeerRadial and VanLeerTheta do (their diffusion terms are comments),
   so the gas routines are called on swapped fields; the FARGO mean and
   shift are the dust's own. */
static void swap_ptr(double **a, double **b) {
  double *t = *a;
  *a = *b;
  *b = t;
}
static void df_transport(or_ctx *c, double dt) {
  const int nr = c->nr;
  long *tl;
  const int energy = c->p.energy_equation;
  const double margin = c->fargo_margin;
  swap_ptr(&c->rho, &c->drho);
  swap_ptr(&c->vr, &c->dvr);
  swap_ptr(&c->vt, &c->dvt);
  swap_ptr(&c->vmed, &c->dvmed);
  tl = c->nshift; c->nshift = c->dnshift; c->dnshift = tl;
  c->p.energy_equation = 0;
  compute_lr_momenta(c);
  compute_star_rad(c, c->rho, c->vr, c->rhostar, dt);
  memcpy(c->rhoint, c->rho, sizeof(double) * (nr + 1) * c->ns);
  if (!c->fp.sfta) {
    van_leer_radial(c, c->rp, dt);
    van_leer_radial(c, c->rm, dt);
    van_leer_radial(c, c->tp, dt);
    van_leer_radial(c, c->tm, dt);
  }
  c->LostMassd += van_leer_radial(c, c->rho, dt);
  c->df_skip_momenta = c->fp.sfta;
  one_wind_theta(c, dt);
  c->df_skip_momenta = 0;
  compute_velocities(c);
  c->p.energy_equation = energy;
  c->fargo_margin = margin;
  tl = c->nshift; c->nshift = c->dnshift; c->dnshift = tl;
  swap_ptr(&c->vmed, &c->dvmed);
  swap_ptr(&c->vt, &c->dvt);
  swap_ptr(&c->vr, &c->dvr);
  swap_ptr(&c->rho, &c->drho);
}

/* Diffd (src/SourceEuler.c:2599-2656): turbulent diffusion of the dust
   concentration with the gas viscosity and the Schmidt number
   (1 + 4 St^2) / (1 + St^2)^2 */
static void df_diffd(or_ctx *c, double dt) {
  const int nr = c->nr, ns = c->ns;
  const double *rho = c->rho, *St = c->stokes;
  double *drho = c->drho;
  const double dtheta = 2.0 * M_PI / (double)ns, invdtheta = 1.0 / dtheta;
  double viscosityp = 0.0, viscosity = 0.0;
  for (int i = 0; i < nr - 1; i++) {
    if (has_viscosity(&c->p)) {
      viscosityp = or_fviscosity(c, c->Rsup[i]);
      viscosity = or_fviscosity(c, c->Rmed[i]);
    }
    for (int j = 0; j < ns; j++) {
      const int l = j + i * ns, lip = l + ns, ljp = (j == ns - 1) ? i * ns : l + 1;
      double Dturb = viscosityp * (1.0 + 4.0 * St[lip] * St[lip]) * pow((1.0 + St[lip] * St[lip]), -2.0);
      c->fdiffrp[l] = -Dturb * (rho[lip] + rho[l]) / 2. * (drho[lip] / rho[lip] - drho[l] / rho[l]) *
                      c->InvDiffRmed[i + 1];
      Dturb = viscosity * (1.0 + 4.0 * St[l] * St[l]) * pow((1.0 + St[l] * St[l]), -2.0);
      c->fdifftp[l] = -Dturb * (rho[ljp] + rho[l]) / 2. * (drho[ljp] / rho[ljp] - drho[l] / rho[l]) * invdtheta /
                      c->Rmed[i];
    }
  }
  for (int i = 1; i < nr - 1; i++) {
    if (!(c->Rmed[i] > c->fp.dust_diff_inn)) continue;
    for (int j = 0; j < ns; j++) {
      const int l = j + i * ns, lim = l - ns, ljm = (j == 0) ? i * ns + ns - 1 : l - 1;
      const double frp = c->fdiffrp[l], frm = c->fdiffrp[lim], ftp = c->fdifftp[l], ftm = c->fdifftp[ljm];
      drho[l] = drho[l] - (c->Rsup[i] * dtheta * frp - c->Rinf[i] * dtheta * frm + (c->Rsup[i] - c->Rinf[i]) * (ftp - ftm)) *
                              dt * c->InvSurf[i];
    }
  }
}

/* the dust's part of ApplyWallBoundaries (src/SideEuler.c:1066-1091) */
static void df_wall_boundaries(or_ctx *c) {
  const int nr = c->nr, ns = c->ns;
  int i = 2;
  for (int j = 0; j < ns; j++) {
    c->dvr[(i - 1) * ns + j] = 0.0;
    c->dvr[(i - 2) * ns + j] = -c->dvr[i * ns + j];
    c->drho[(i - 1) * ns + j] = c->DSigmaMed[i - 1];
    c->drho[(i - 2) * ns + j] = c->DSigmaMed[i - 2];
  }
  i = nr - 3;
  for (int j = 0; j < ns; j++) {
    c->dvr[(i + 2) * ns + j] = 0.0;
    c->dvr[(i + 3) * ns + j] = -c->dvr[(i + 1) * ns + j];
    c->drho[(i + 1) * ns + j] = c->DSigmaMed[i + 1];
    c->drho[(i + 2) * ns + j] = c->DSigmaMed[i + 2];
  }
}

/* the dust's part of ApplySubKeplerianBoundary (src/SideEuler.c:1025-1041) */
static void df_subkeplerian_boundary(or_ctx *c) {
  const int nr = c->nr, ns = c->ns;
  for (int j = 0; j < ns; j++) c->dvt[j] = c->DVthetaMed[0] - c->Rmed[0] * c->OmegaFrame;
  for (int j = 0; j < ns; j++) c->dvt[(nr - 1) * ns + j] = c->DVthetaMed[nr - 1] - c->Rmed[nr - 1] * c->OmegaFrame;
}

/* OpenBoundaryd (src/SideEuler.c:411-466), DontApplySubKeplerian off:
   zero-gradient density, vr copied from the next ring inwards */
static void df_open_boundary(or_ctx *c) {
  const int nr = c->nr, ns = c->ns;
  double *rho = c->drho, *vr = c->dvr;
  int i = 2;
  for (int j = 0; j < ns; j++) {
    const int l = j + i * ns;
    rho[l - ns] = rho[l];
    rho[l - 2 * ns] = rho[l];
    vr[l] = vr[l + ns];
    vr[l - ns] = vr[l + ns];
  }
  i = nr - 1;
  for (int j = 0; j < ns; j++) {
    const int l = j + i * ns;
    rho[l] = rho[l - ns];
    vr[l] = vr[l - ns];
  }
}

int or_dust_fluid_attach(or_ctx *c, const dfargo_dust_fluid_params *f, const double *drho) {
  c->df_on = 1;
  c->fp = *f;
  double **g[] = {&c->drho, &c->dvr, &c->dvt, &c->stokes, &c->dpres, &c->dvradint, &c->dvthetaint,
                  &c->dvradnew, &c->dvthetanew, &c->dqr, &c->dqt, &c->ddrr, &c->ddrp, &c->ddpp,
                  &c->ddivv, &c->dtrr, &c->dtrp, &c->dtpp, &c->fdiffrp, &c->fdifftp};
  for (size_t k = 0; k < sizeof(g) / sizeof(g[0]); k++)
    if (!*g[k]) *g[k] = grid(c);
  double **t[] = {&c->dcs1, &c->DSigmaMed, &c->DVthetaMed, &c->DVradMed, &c->GlobalDCs, &c->AxiSGd,
                  &c->dvmed, &c->dvmoy};
  for (size_t k = 0; k < sizeof(t) / sizeof(t[0]); k++)
    if (!*t[k]) *t[k] = (double *)calloc(c->nr + 2, sizeof(double));
  if (!c->dnshift) c->dnshift = (long *)calloc(c->nr + 2, sizeof(long));
  for (int i = 0; i < c->nr; i++) c->DSigmaMed[i] = dsigma(c, c->Rmed[i]);
  memcpy(c->drho, drho, sizeof(double) * (c->nr + 1) * c->ns);
  return 0;
}

void or_dust_fluid_get(or_ctx *c, double *drho, double *dvr, double *dvt, double *stokes) {
  const size_t b = sizeof(double) * (c->nr + 1) * c->ns;
  if (drho) memcpy(drho, c->drho, b);
  if (dvr) memcpy(dvr, c->dvr, b);
  if (dvt) memcpy(dvt, c->dvt, b);
  if (stokes) memcpy(stokes, c->stokes, b);
}

void or_dust_fluid_set(or_ctx *c, const double *drho, const double *dvr, const double *dvt) {
  const size_t b = sizeof(double) * (c->nr + 1) * c->ns;
  if (drho) memcpy(c->drho, drho, b);
  if (dvr) memcpy(c->dvr, dvr, b);
  if (dvt) memcpy(c->dvt, dvt, b);
}


int or_algogas(or_ctx *c, double DT) {
  const dfargo_params *p = &c->p;
  dfargo_planets *sys = &c->sys;
  int nb = sys->nb, nsub = 0;
  double dtemp = 0.0, dt, dthydro, dtnbody;
  size_t n = (size_t)(c->nr + 1) * c->ns;
  if (p->energy_equation) compute_sound_speed(c);
  c->GasTimeStepsCFL = 1;
  c->fargo_margin = 0.5;
  while (dtemp < 0.999999999 * DT) {
    /* mass taper (src/SourceEuler.c:400-418) */
    if (p->mass_taper > 1e-2) {
      c->MassTaper = (c->PhysicalTime - c->PhysicalTimeInitial) / (p->mass_taper * 2.0 * M_PI);
      c->MassTaper = (c->MassTaper > 1.0 ? 1.0 : pow(sin(c->MassTaper * M_PI / 2.0), 2.0));
      if (c->MassTaper < 1.0)
        for (int k = 0; k < nb; k++)
          sys->mass[k] = sys->initial_mass[k] + (sys->final_mass[k] - sys->initial_mass[k]) * c->MassTaper;
      c->particles_mass = c->particles_mass_initial * c->MassTaper;
    } else {
      for (int k = 0; k < nb; k++)
        if (sys->acc[k] < 1e-10) sys->mass[k] = sys->final_mass[k];
      c->particles_mass = c->particles_mass_initial;
    }
    /* n-body limit (src/SourceEuler.c:421-440) */
    if (nb > 1) {
      dtnbody = 1e5;
      for (int k = 0; k < nb; k++)
        for (int j = k + 1; j < nb; j++) {
          double xk = sys->x[k], xj = sys->x[j], yk = sys->y[k], yj = sys->y[j];
          double mk = sys->mass[k], mj = sys->mass[j];
          double dist = sqrt((xk - xj) * (xk - xj) + (yk - yj) * (yk - yj));
          double buf = 2.0 * M_PI * sqrt(dist * dist * dist / (mk + mj + 1e-8)) / 100.0;
          dtnbody = min2(dtnbody, buf);
        }
    } else
      dtnbody = 100.0;
    if (!c->dust_on || c->np == 0 || c->dp.no_dt_constraint_by_particles) {
      c->GasTimeStepsCFL = or_condition_cfl(c, DT - dtemp);
      dthydro = (DT - dtemp) / (double)c->GasTimeStepsCFL;
      dt = min2(dthydro, dtnbody);
      accrete_onto_planets(c, dt); /* src/SourceEuler.c:454 */
    } else {
      /* fixed particle time step (src/SourceEuler.c:457-463) */
      dt = c->dp.dust_timestep;
      c->MINDT = c->dp.dust_timestep;
      c->MAXDT = c->dp.dust_timestep;
    }
    dtemp += dt;
    if (dt < c->MINDT) c->MINDT = dt;
    if (dt > c->MAXDT) c->MAXDT = dt;
    /* forces and planets (src/SourceEuler.c:472-500) */
    if (p->corotating) psys_info(c, 0);
    c->dop = compute_accel(c, c->rho, 0.0, 0.0, p->indirect_term_smoothing, 0.0);
    c->dop_alt = opposite_acc_barycenter(c, dt);
    fill_forces_arrays(c);
    advance_system_from_disk(c, dt);
    advance_system_rk5(c, dt);
    /* particles (src/SourceEuler.c:504-513): positions over dt, then
       velocities over dt */
    if (c->dust_on && c->np != 0) {
      dust_semi_update_positions(c, 2. * dt);
      dust_update_velocities(c, dt);
    }
    /* the frame follows its centre (src/SourceEuler.c:516-524) */
    if (p->corotating) {
      const double omega_new = psys_info(c, 1) / dt;
      correct_vtheta(c, omega_new - c->OmegaFrame);
      c->OmegaFrame = omega_new;
    }
    rotate_psys(c, c->OmegaFrame * dt);
    if (c->dust_on && c->np != 0) dust_rotate(c, c->OmegaFrame * dt);
    /* the 1-D viscous disc (src/SourceEuler.c:540-541) */
    if (p->damp_to == DFARGO_DAMP_TO_VISCOUS) {
      if (!c->n1d) init_1d_viscous(c);
      solve_1d_viscous(c, dt);
    }
    /* crash scan (src/SourceEuler.c:545-548) */
    c->crashed_d = detect_crash(c, c->rho);
    c->crashed_e = detect_crash(c, c->en);
    c->crashed = c->crashed_e;
    compute_sound_speed(c);
    compute_pressure(c);
    if (c->df_on) df_sound_speed_pressure(c);
    or_substep1(c, dt);
    /* ActualiseGas round trips (src/SourceEuler.c:598-612) */
    memcpy(c->vr, c->vradint, sizeof(double) * n);
    memcpy(c->vt, c->vthetaint, sizeof(double) * n);
    substep2(c, dt);
    if (c->df_on) df_substep2(c, dt);
    memcpy(c->vr, c->vradnew, sizeof(double) * n);
    memcpy(c->vt, c->vthetanew, sizeof(double) * n);
    /* the dust velocities: DVradNew, or in the short-friction-time
       approximation DVradInt (src/SourceEuler.c:600-630) */
    if (c->df_on) {
      memcpy(c->dvr, c->fp.sfta ? c->dvradint : c->dvradnew, sizeof(double) * n);
      memcpy(c->dvt, c->fp.sfta ? c->dvthetaint : c->dvthetanew, sizeof(double) * n);
    }
    if (p->energy_equation) memcpy(c->en, c->energyint, sizeof(double) * n);
    compute_divergence(c, c->vr, c->vt);
    if (p->energy_equation) {
      memcpy(c->energyint, c->en, sizeof(double) * n);
      if (p->viscous_heating) {
        compute_viscous_terms(c, c->vr, c->vt);
        compute_viscous_heating(c);
      }
      substep3(c, dt);
      memcpy(c->en, c->energynew, sizeof(double) * n);
    }
    or_transport(c, dt);
    if (p->add_floors) add_floors(c);
    /* the dust fluid's transport, diffusion and floor (src/SourceEuler.c:667-676) */
    if (c->df_on) {
      df_transport(c, dt);
      if (c->fp.dust_diffusion) df_diffd(c, dt);
      if (p->add_floors)
        for (int l = 0; l < c->nr * c->ns; l++)
          if (c->drho[l] < p->density_floor) c->drho[l] = p->density_floor;
    }
    apply_boundary_condition(c);
    if (p->wave_killing_zones) evanescent_boundary(c, dt);
    compute_temperature(c);
    c->PhysicalTime += dt;
    nsub++;
  }
  c->last_substeps = nsub;
  c->substeps += nsub;
  return nsub;
}

void or_get_status(or_ctx *c, dfargo_status *st) {
  memset(st, 0, sizeof(*st));
  st->physical_time = c->PhysicalTime;
  st->physical_time_initial = c->PhysicalTimeInitial;
  st->omega_frame = c->OmegaFrame;
  st->mindt = c->MINDT;
  st->maxdt = c->MAXDT;
  st->lost_mass = c->LostMass;
  st->mass_taper = c->MassTaper;
  st->disk_on_primary_x = c->dop.x;
  st->disk_on_primary_y = c->dop.y;
  st->disk_on_primary_alt_x = c->dop_alt.x;
  st->disk_on_primary_alt_y = c->dop_alt.y;
  for (int k = 0; k < 3; k++) {
    st->bary_x[k] = c->barx[k];
    st->bary_y[k] = c->bary[k];
  }
  st->substeps_total = c->substeps;
  st->last_call_substeps = c->last_substeps;
  st->gas_timesteps_cfl = c->GasTimeStepsCFL;
  st->crashed_density = c->crashed_d;
  st->crashed_energy = c->crashed_e;
  st->crashed = c->crashed;
  st->fargo_margin = c->fargo_margin;
}

void or_get_field(or_ctx *c, int field, double *out) {
  size_t n = (size_t)(c->nr + 1) * c->ns;
  const double *src = NULL;
  switch (field) {
    case DFARGO_FIELD_POTENTIAL: src = c->pot; break;
    case DFARGO_FIELD_SOUNDSPEED: src = c->cs; break;
    case DFARGO_FIELD_PRESSURE: src = c->pres; break;
    case DFARGO_FIELD_TEMPERATURE: src = c->temp; break;
    case DFARGO_FIELD_DIVV: src = c->divv; break;
    case DFARGO_FIELD_VISCHEAT: src = c->vischeat; break;
    case DFARGO_FIELD_SG_ACCR: src = c->gr; break;
    case DFARGO_FIELD_SG_ACCT: src = c->gtheta; break;
    case DFARGO_FIELD_TAURR: src = c->trr; break;
    case DFARGO_FIELD_TAURP: src = c->trp; break;
    case DFARGO_FIELD_TAUPP: src = c->tpp; break;
    case DFARGO_FIELD_DUST_DENSITY: src = c->dustdens; break;
    case DFARGO_FIELD_FB_ACCR: src = c->radfbacc; break;
    case DFARGO_FIELD_FB_ACCT: src = c->azifbacc; break;
    case DFARGO_FIELD_FB_EDOT: src = c->fbedot; break;
    default: memset(out, 0, sizeof(double) * n); return;
  }
  memcpy(out, src, sizeof(double) * n);
}

void or_get_fargo_shift(or_ctx *c, long *nshift, double *vmed) {
  if (nshift) memcpy(nshift, c->nshift, sizeof(long) * c->nr);
  if (vmed) memcpy(vmed, c->vmed, sizeof(double) * c->nr);
}

/* ------------------------------------------------------------------------ */
/* per-DT diagnostics (src/main.c:291-394, outside AlgoGas)                 */
/* ------------------------------------------------------------------------ */

/* out[0] GasTotalMass (src/SideEuler.c:15-36), out[1] GasMomentum
 * (src/SideEuler.c:114-146), out[2] GasTotalEnergy (src/SideEuler.c:149-180;
 * the energy grid of an isothermal run is CreatePolarGrid's zeros,
 * src/LowTasks.c:108-113), out[3] the planets' angular momentum of
 * CheckMomentumConservation (src/SideEuler.c:182-210).  The potential is
 * the Potential grid of the last sub-step. */
void or_gas_budgets(or_ctx *c, double out[4]) {
  int nr = c->nr, ns = c->ns;
  const double *density = c->rho, *vrad = c->vr, *vtheta = c->vt, *pot = c->pot;
  double total = 0.0;
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) total += c->Surf[i] * density[j + i * ns];
  out[0] = total;
  total = 0.0;
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      int l = i * ns + j;
      double vt_cent;
      if (j < ns - 1) vt_cent = 0.5 * (vtheta[l] + vtheta[l + 1]) + c->Rmed[i] * c->OmegaFrame;
      else vt_cent = 0.5 * (vtheta[l] + vtheta[i * ns]) + c->Rmed[i] * c->OmegaFrame;
      total += c->Surf[i] * density[l] * c->Rmed[i] * vt_cent;
    }
  out[1] = total;
  total = 0.0;
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      int l = j + i * ns;
      double vr_cent = (c->Rmed[i] - c->Rinf[i]) * vrad[l + ns] + (c->Rsup[i] - c->Rmed[i]) * vrad[l];
      vr_cent /= (c->Rsup[i] - c->Rinf[i]);
      double vt_cent;
      if (j < ns - 1) vt_cent = 0.5 * (vtheta[l] + vtheta[l + 1]) + c->Rmed[i] * c->OmegaFrame;
      else vt_cent = 0.5 * (vtheta[l] + vtheta[i * ns]) + c->Rmed[i] * c->OmegaFrame;
      double e = c->p.energy_equation ? c->en[l] : 0.0;
      total += 0.5 * c->Surf[i] * density[l] * (vr_cent * vr_cent + vt_cent * vt_cent) + c->Surf[i] * e -
               c->Surf[i] * density[l] * pot[l];
    }
  out[2] = total;
  double planetsmom = 0.0;
  for (int k = 0; k < c->sys.nb; k++) {
    double xp = c->sys.x[k], yp = c->sys.y[k];
    double rpl = sqrt(xp * xp + yp * yp), thetapl = atan2(yp, xp);
    double vazimpl = -c->sys.vx[k] * sin(thetapl) + c->sys.vy[k] * cos(thetapl);
    planetsmom += c->sys.mass[k] * rpl * vazimpl;
  }
  out[3] = planetsmom;
}

/* UpdateLog (src/Force.c:266-328): for planet k, tq[8k..8k+7] are the
 * columns of tqwk<k>.dat (torques x fy - y fx and powers vx fx + vy fy of
 * the inner, outer, Hill-excluded inner and outer discs) and ind[k] the
 * indirect torque of indtq<k>.dat, from the DiskOnPrimaryAcceleration of
 * the last sub-step.  Returns the number of planets. */
int or_update_log(or_ctx *c, double *tq, double *ind) {
  int nb = c->sys.nb;
  for (int k = 0; k < nb; k++) {
    double x = c->sys.x[k], y = c->sys.y[k], vx = c->sys.vx[k], vy = c->sys.vy[k];
    double r = sqrt(x * x + y * y), m = c->sys.mass[k], smoothing, gf[8];
    if (c->p.roche_smoothing != 0.0) smoothing = r * pow(m / 3., 1. / 3.) * c->p.roche_smoothing;
    else smoothing = compute_smoothing(&c->p, r);
    compute_force(c, c->rho, x, y, smoothing, m, gf);
    /* gf: fx_inner, fx_ex_inner, fx_outer, fx_ex_outer, fy_inner, ... */
    double fxi = gf[0], fxei = gf[1], fxo = gf[2], fxeo = gf[3];
    double fyi = gf[4], fyei = gf[5], fyo = gf[6], fyeo = gf[7];
    double *o = tq + 8 * k;
    o[0] = x * fyi - y * fxi;
    o[1] = x * fyo - y * fxo;
    o[2] = x * fyei - y * fxei;
    o[3] = x * fyeo - y * fxeo;
    o[4] = vx * fxi + vy * fyi;
    o[5] = vx * fxo + vy * fyo;
    o[6] = vx * fxei + vy * fyei;
    o[7] = vx * fxeo + vy * fyeo;
    ind[k] = -x * c->dop.y + y * c->dop.x;
  }
  return nb;
}

/* FindOrbitalElements (src/Planet.c:129-180): out = e, a, M, V,
 * PerihelionPA, lambda, varpi (the columns of orbit<n>.dat after the time) */
void or_orbital_elements(double x, double y, double vx, double vy, double m, double out[7]) {
  double Ax, Ay, e, d, h, a, E, M, V, lambda, varpi, PerihelionPA;
  h = x * vy - y * vx;
  d = sqrt(x * x + y * y);
  Ax = x * vy * vy - y * vx * vy - G * m * x / d;
  Ay = y * vx * vx - x * vx * vy - G * m * y / d;
  e = sqrt(Ax * Ax + Ay * Ay) / G / m;
  a = h * h / G / m / (1 - e * e);
  if (e != 0.0) E = acos((1.0 - d / a) / e);
  else E = 0.0;
  if ((x * y * (vy * vy - vx * vx) + vx * vy * (x * x - y * y)) < 0) E = -E;
  M = E - e * sin(E);
  if (e != 0.0) V = acos((a * (1.0 - e * e) / d - 1.0) / e);
  else V = 0.0;
  if (E < 0.0) V = -V;
  lambda = atan2(y, x);
  if (e != 0.0) PerihelionPA = atan2(Ay, Ax);
  else PerihelionPA = atan2(y, x);
  if (e != 0.0) varpi = 2.0 * atan2(Ay, 1. + Ax);
  else varpi = lambda;
  out[0] = e; out[1] = a; out[2] = M; out[3] = V;
  out[4] = PerihelionPA; out[5] = lambda; out[6] = varpi;
}

/* ------------------------------------------------------------------------ */
/* test hooks: single stages of the sub-step on the context's own grids, so
 * that tests can pin each stage against closed forms                        */
/* ------------------------------------------------------------------------ */

static double *test_grid(or_ctx *c, int which) {
  double *g[OR_G_COUNT] = {c->rho,       c->vr,        c->vt,       c->en,       c->cs,      c->pres,
                           c->pot,       c->vradint,   c->vthetaint, c->vradnew, c->vthetanew, c->energyint,
                           c->energynew, c->rhoint,    c->temperint, c->divv,    c->drr,      c->dpp,
                           c->drp,       c->trr,       c->tpp,       c->trp,     c->vischeat, c->indpot};
  return (which >= 0 && which < OR_G_COUNT) ? g[which] : NULL;
}

int or_test_grid(or_ctx *c, int which, double *buf, int set) {
  double *g = test_grid(c, which);
  if (!g) return -1;
  size_t n = (size_t)(c->nr + 1) * c->ns;
  if (set) memcpy(g, buf, sizeof(double) * n);
  else memcpy(buf, g, sizeof(double) * n);
  return 0;
}

int or_test_stage(or_ctx *c, int stage, double dt) {
  switch (stage) {
    case OR_S_EOS: compute_sound_speed(c); compute_pressure(c); return 0;
    case OR_S_SUBSTEP1: or_substep1(c, dt); return 0;
    case OR_S_SUBSTEP2: substep2(c, dt); return 0;
    case OR_S_VISC_TERMS: compute_viscous_terms(c, c->vr, c->vt); return 0;
    case OR_S_VISC_HEATING: compute_viscous_heating(c); return 0;
    case OR_S_DIVERGENCE: compute_divergence(c, c->vr, c->vt); return 0;
    case OR_S_SUBSTEP3: substep3(c, dt); return 0;
    case OR_S_POTENTIAL: fill_forces_arrays(c); return 0;
    case OR_S_BOUNDARY: apply_boundary_condition(c); return 0;
    case OR_S_FLOORS: add_floors(c); return 0;
    case OR_S_CRASH: return detect_crash(c, c->rho) | (detect_crash(c, c->en) << 1);
  }
  return -1;
}
