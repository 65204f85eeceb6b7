/*
 * host.c -- host-side functions of the C ABI: parameter defaults and checks,
 * the 1-D geometry tables, the initial profiles and the per-ring tables the
 * device kernels read.  All transcendental functions are evaluated here with
 * the host libm, once, so that device arithmetic needs only +, -, *, / and
 * sqrt (all correctly rounded on gfx950).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dfargo_internal.h"

#define G 1.0

void dfargo_default_params(dfargo_params *p) {
  memset(p, 0, sizeof(*p));
  /* src/var.c:15-202 */
  p->nrad = 64; p->nsec = 64; p->rmin = 1.0; p->rmax = 1.0; p->log_grid = 0;
  p->pmin = 0.0; p->pmax = 6.2831853071795864;
  p->sigma0 = 173.; p->sigma_slope = 0.0; p->aspect_ratio = 0.05; p->flaring_index = 0.0;
  p->viscosity = 0.0; p->alpha_viscosity = 0.0;
  p->cavity_radius = 0.0; p->cavity_ratio = 1.0; p->cavity_width = 1.0;
  p->transition_radius = 0.0; p->transition_ratio = 1.0; p->transition_width = 1.0;
  p->imposed_disk_drift = 0.0; p->add_noise = 0; p->noise_seed = 0;
  p->energy_equation = 0; p->viscous_heating = 1; p->beta_cooling = 0;
  p->adiabatic_index = 1.4; p->beta_cooling_time = 10.0; p->beta_cooling_slope = 0.0;
  p->self_gravity = 0; p->sg_thickness_smoothing = 0.3;
  p->thickness_smoothing = 0.3; p->roche_smoothing = 0.0; p->indirect_term_smoothing = 0.0;
  p->mass_taper = 0.0000001; p->indirect_term = 1; p->discard_gas_indirect_term = 0;
  p->alternative_gas_indirect_term = 0; p->exclude_hill = 0; p->exclude_hill_factor = 1.0;
  p->forced_circular = 0; p->release_date = 0.0; p->release_radius = 0.0;
  p->eccentricity = 0.0; p->corotating = 0; p->omega_frame = 0.0;
  p->fast_transport = 1; p->add_floors = 1; p->density_floor = 1e-9;
  p->boundary = DFARGO_BC_WALL;
  p->corotate_with_outer_planet = 0;
  p->wave_killing_zones = 0; p->wkz_rmin = 0.0; p->wkz_rmax = 0.0;
  p->damping_ratio = 0.3; p->damp_to = DFARGO_DAMP_TO_AXI; p->zero_vrad_bc = 0;
  p->sg_zero_mode = 0;
  p->nrad1d = 100; p->rmin1d = 0.01; p->rmax1d = 100.0; p->bc1d = 0;
}

int dfargo_check_params(const dfargo_params *p) {
  if (!p) return dfargo_set_error(DFARGO_ERR_ARG, "null parameter block");
  if (p->nrad < 8 || p->nsec < 8)
    return dfargo_set_error(DFARGO_ERR_ARG, "NRAD and NSEC must be at least 8 (got %d, %d)",
                            p->nrad, p->nsec);
  if (p->nrad > 32768 || p->nsec > 32768) /* MAX1D, src/types.h:78 */
    return dfargo_set_error(DFARGO_ERR_ARG, "NRAD and NSEC are capped at MAX1D = 32768");
  if (!(p->rmax > p->rmin) || !(p->rmin > 0.0))
    return dfargo_set_error(DFARGO_ERR_ARG, "need 0 < RMIN < RMAX");
  if ((p->alpha_viscosity != 0.0) && (p->viscosity != 0.0)) /* src/Interpret.c:449-455 */
    return dfargo_set_error(DFARGO_ERR_ARG, "VISCOSITY and ALPHAVISCOSITY cannot both be set");
  if (p->exclude_hill_factor < 0.0 || p->exclude_hill_factor > 1.0) /* src/Interpret.c:444 */
    return dfargo_set_error(DFARGO_ERR_ARG, "EXCLUDEHILLFACTOR must range between 0 and 1");
  if (p->self_gravity && p->sg_thickness_smoothing == 0.0) /* src/Interpret.c:459 */
    return dfargo_set_error(DFARGO_ERR_ARG, "SGTHICKNESSSMOOTHING must not vanish with self-gravity");
  if (p->self_gravity && !p->log_grid) /* src/sginit.c:6-7 */
    return dfargo_set_error(DFARGO_ERR_ARG, "self-gravity needs a logarithmic radial grid");
  if (p->energy_equation && p->adiabatic_index == 1.0)
    return dfargo_set_error(DFARGO_ERR_ARG, "the energy equation needs ADIABATICINDEX != 1");
  if (p->corotating < DFARGO_FRAME_FIXED || p->corotating > DFARGO_FRAME_BINARY_CENTER)
    return dfargo_set_error(DFARGO_ERR_ARG, "unknown FRAME %d", p->corotating);
  if (p->wave_killing_zones && (p->wkz_rmin == 0.0 || p->wkz_rmax == 0.0)) /* src/Interpret.c:471-476 */
    return dfargo_set_error(DFARGO_ERR_ARG, "wave-killing zones need WKZRMIN and WKZRMAX");
  if (p->wave_killing_zones && p->wkz_rmin > p->wkz_rmax)
    return dfargo_set_error(DFARGO_ERR_ARG, "WKZRMIN must not exceed WKZRMAX");
  if (p->damp_to < DFARGO_DAMP_TO_AXI || p->damp_to > DFARGO_DAMP_TO_VISCOUS)
    return dfargo_set_error(DFARGO_ERR_ARG, "unknown damping target %d", p->damp_to);
  if (p->damp_to == DFARGO_DAMP_TO_VISCOUS) {
    /* the 1-D disc's viscosity is set only by VISCOSITY or ALPHAVISCOSITY
       (src/SideEuler.c:1477-1480); with the wave-killing zones its grid
       must cover the 2-D one for the interpolation of
       src/SideEuler.c:879-886, without them it only evolves */
    if (p->viscosity == 0.0 && p->alpha_viscosity == 0.0)
      return dfargo_set_error(DFARGO_ERR_ARG, "DAMPTOVISCOUS needs VISCOSITY or ALPHAVISCOSITY");
    if (p->nrad1d < 3 || p->nrad1d > 1024)
      return dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "NRAD1D must lie in [3, 1024], got %d", p->nrad1d);
    if (!(p->rmin1d > 0.0) || !(p->rmax1d > p->rmin1d))
      return dfargo_set_error(DFARGO_ERR_ARG, "need 0 < RMIN1D < RMAX1D");
    if (p->wave_killing_zones && (!(p->rmin1d < p->rmin) || !(p->rmax1d > p->rmax)))
      return dfargo_set_error(DFARGO_ERR_ARG, "RMIN1D/RMAX1D must enclose [RMIN, RMAX] under the wave-killing zones");
    if (p->bc1d < 0 || p->bc1d > 2) return dfargo_set_error(DFARGO_ERR_ARG, "unknown BOUNDARY1DGRID %d", p->bc1d);
  }
  if (p->boundary != DFARGO_BC_WALL && p->boundary != DFARGO_BC_OPEN && p->boundary != DFARGO_BC_NONREFLECTING)
    return dfargo_set_error(DFARGO_ERR_UNSUPPORTED,
                            "only WALL, OPEN and NONREFLECTING boundaries are implemented");
  /* the two-dimensional transforms; the axisymmetric mode is a direct
     radial sum with no size condition */
  const int sg2d = p->self_gravity && !p->sg_zero_mode;
  if (sg2d && (p->nrad & (p->nrad - 1)))
    return dfargo_set_error(DFARGO_ERR_UNSUPPORTED,
                            "self-gravity needs NRAD a power of two (2*NRAD FFT), got %d", p->nrad);
  if (sg2d && !dfargo_fft_length_ok(p->nsec))
    return dfargo_set_error(DFARGO_ERR_UNSUPPORTED,
                            "self-gravity needs NSEC = 2^a 3^b <= 4096, got %d", p->nsec);
  if (sg2d && 2 * p->nrad > 4096)
    return dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "self-gravity needs 2*NRAD <= 4096");
  if (fabs(p->pmax - p->pmin - 2. * M_PI) >= 0.01 && p->self_gravity)
    return dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "self-gravity needs a full 2 pi azimuth");
  return DFARGO_OK;
}

int dfargo_fft_length_ok(int n) {
  if (n < 1 || n > 4096) return 0;
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  return n == 1;
}

/* Sigma (src/Theo.c:13-60) */
double dfargo_sigma_profile(const dfargo_params *p, double r) {
  double cavity = 1.0;
  if (r < p->cavity_radius) cavity = 1.0 / p->cavity_ratio;
  return cavity * 1.0 * p->sigma0 * pow(r, -p->sigma_slope);
}

/* Energy (src/Theo.c:156-170), R = MU = 1 */
double dfargo_energy_profile(const dfargo_params *p, double r) {
  double e0 = 1.0 / 1.0 / (p->adiabatic_index - 1.0) * dfargo_sigma_profile(p, r) *
              pow(p->aspect_ratio, 2.0) * pow(r, -1.0 + 2.0 * p->flaring_index);
  return 1.0 * 1.0 * e0;
}

/* AspectRatio (src/Viscosity.c:107-122), LAMBDADOUBLING = 0 */
double dfargo_aspect_ratio(const dfargo_params *p, double rad) {
  double ar = p->aspect_ratio;
  double rmin = p->transition_radius - p->transition_width * p->aspect_ratio;
  double rmax = p->transition_radius + p->transition_width * p->aspect_ratio;
  if (rad < rmin) ar *= p->transition_ratio;
  if ((rad >= rmin) && (rad <= rmax))
    ar *= exp((rmax - rad) / (rmax - rmin) * log(p->transition_ratio));
  return ar;
}

/* The cavity factor FViscosity applies after the base viscosity
 * (src/Viscosity.c:39-48); 1.0 outside the cavity, so multiplying by it is
 * exact there. */
static double viscosity_factor(const dfargo_params *p, double rad) {
  double f = 1.0;
  double rmin = p->cavity_radius -
                p->cavity_width * p->aspect_ratio * pow(p->cavity_radius, 1.0 + p->flaring_index);
  double rmax = p->cavity_radius +
                p->cavity_width * p->aspect_ratio * pow(p->cavity_radius, 1.0 + p->flaring_index);
  if (rad < rmin) f *= p->cavity_ratio;
  if ((rad >= rmin) && (rad <= rmax)) f *= exp((rmax - rad) / (rmax - rmin) * log(p->cavity_ratio));
  return f;
}

/* FillPolar1DArrays (src/SourceEuler.c:71-188) into a table block */
int dfargo_host_tables(const dfargo_params *p, dfargo_tables *t) {
  int nr = p->nrad, ns = p->nsec, i, j;
  double drrsep = (p->rmax - p->rmin) / (double)nr;
  memset(t, 0, sizeof(*t));
  t->nr = nr;
  t->ns = ns;
#define A(n) (double *)calloc((size_t)(n), sizeof(double))
  t->Radii = A(nr + 2); t->Rinf = A(nr + 2); t->Rsup = A(nr + 2); t->Rmed = A(nr + 2);
  t->Surf = A(nr + 2); t->InvRmed = A(nr + 2); t->InvSurf = A(nr + 2);
  t->InvDiffRmed = A(nr + 2); t->InvDiffRsup = A(nr + 2); t->InvRinf = A(nr + 2);
  t->CosAz = A(ns); t->SinAz = A(ns); t->Azimuth = A(ns);
  t->SigmaMed = A(nr + 2); t->SigmaInf = A(nr + 2); t->EnergyMed = A(nr + 2);
  t->cs_iso = A(nr + 2); t->powmin2 = A(nr + 2); t->r15 = A(nr + 2); t->visc_mult = A(nr + 2);
  t->beta = A(nr + 2); t->supp_torque = A(nr + 2); t->sg_sqrt = A(nr + 2);
  t->sg_normr = A(nr + 2); t->sg_normt = A(nr + 2);
#undef A
  if (!t->sg_normt) return dfargo_set_error(DFARGO_ERR_ALLOC, "host table allocation failed");
  for (i = 0; i <= nr; i++) {
    if (p->log_grid)
      t->Radii[i] = p->rmin * exp((double)i / (double)nr * log(p->rmax / p->rmin));
    else
      t->Radii[i] = p->rmin + drrsep * (double)(i);
  }
  for (i = 0; i < nr; i++) {
    t->Rinf[i] = t->Radii[i];
    t->Rsup[i] = t->Radii[i + 1];
    t->Rmed[i] = 2.0 / 3.0 * (t->Rsup[i] * t->Rsup[i] * t->Rsup[i] - t->Rinf[i] * t->Rinf[i] * t->Rinf[i]);
    t->Rmed[i] = t->Rmed[i] / (t->Rsup[i] * t->Rsup[i] - t->Rinf[i] * t->Rinf[i]);
    t->Surf[i] = 0.5 * (p->pmax - p->pmin) * (t->Rsup[i] * t->Rsup[i] - t->Rinf[i] * t->Rinf[i]) / (double)ns;
    t->InvRmed[i] = 1.0 / t->Rmed[i];
    t->InvSurf[i] = 1.0 / t->Surf[i];
    t->InvDiffRsup[i] = 1.0 / (t->Rsup[i] - t->Rinf[i]);
    t->InvRinf[i] = 1.0 / t->Rinf[i];
  }
  t->Rinf[nr] = t->Radii[nr];
  for (i = 1; i < nr; i++) t->InvDiffRmed[i] = 1.0 / (t->Rmed[i] - t->Rmed[i - 1]);
  for (j = 0; j < ns; j++) {
    t->Azimuth[j] = p->pmin + (p->pmax - p->pmin) * (double)j / (double)ns;
    t->CosAz[j] = cos(t->Azimuth[j]);
    t->SinAz[j] = sin(t->Azimuth[j]);
  }
  for (i = 0; i < nr; i++) {
    double dxrad = t->Rsup[i] - t->Rinf[i];
    double dxtheta = t->Rmed[i] * (p->pmax - p->pmin) / (double)ns;
    double m = (dxtheta < dxrad) ? dxtheta : dxrad;
    t->SigmaMed[i] = dfargo_sigma_profile(p, t->Rmed[i]);
    t->SigmaInf[i] = dfargo_sigma_profile(p, t->Rinf[i]);
    if (p->energy_equation) t->EnergyMed[i] = dfargo_energy_profile(p, t->Rmed[i]);
    /* ComputeSoundSpeed, isothermal (src/SourceEuler.c:2336) */
    t->cs_iso[i] = dfargo_aspect_ratio(p, t->Rmed[i]) * sqrt(G * 1.0 / t->Rmed[i]) *
                   pow(t->Rmed[i], p->flaring_index);
    t->powmin2[i] = pow(m, 2.0);                  /* ConditionCFL invdt5 denominator */
    t->r15[i] = pow(t->Rmed[i], 1.5);             /* FViscosity alpha term */
    t->visc_mult[i] = viscosity_factor(p, t->Rmed[i]);
    t->beta[i] = p->beta_cooling_time * pow(t->Rmed[i], -p->beta_cooling_slope); /* src/Theo.c:216-222 */
    t->supp_torque[i] = p->imposed_disk_drift * 0.5 * pow(t->Rmed[i], -2.5) * p->sigma0 / t->SigmaMed[i];
  }
  /* self-gravity constants (src/sginit.c:9-11, src/sgacc.c:137-147) */
  t->sg_eps = p->sg_thickness_smoothing * p->aspect_ratio;
  t->sg_rstep = log(t->Radii[nr] / t->Radii[0]) / (double)nr;
  t->sg_tstep = (p->pmax - p->pmin) / (double)ns;
  for (i = 0; i < nr; i++) {
    double normaccr = t->sg_rstep * t->sg_tstep / ((double)(2 * nr) * (double)ns);
    double normacct = normaccr;
    t->sg_sqrt[i] = sqrt(t->Rmed[i] / t->Rmed[0]);
    normaccr /= sqrt(t->Rmed[i] / t->Rmed[0]);
    normacct /= (t->Rmed[i] / t->Rmed[0] * sqrt(t->Rmed[i] / t->Rmed[0]));
    t->sg_normr[i] = normaccr;
    t->sg_normt[i] = normacct;
  }
  /* ring-0 extrapolation exponents of ComputeViscousHeating
   * (src/SourceEuler.c:1589-1605) */
  if (nr >= 3) {
    t->vh_lr1 = log(t->Rmed[0] / t->Rmed[1]);
    t->vh_lr2 = log(t->Rmed[1] / t->Rmed[2]);
  }
  return DFARGO_OK;
}

void dfargo_free_tables(dfargo_tables *t) {
  double **all[] = {&t->Radii, &t->Rinf, &t->Rsup, &t->Rmed, &t->Surf, &t->InvRmed,
                    &t->InvSurf, &t->InvDiffRmed, &t->InvDiffRsup, &t->InvRinf, &t->CosAz,
                    &t->SinAz, &t->Azimuth, &t->SigmaMed, &t->SigmaInf, &t->EnergyMed,
                    &t->cs_iso, &t->powmin2, &t->r15, &t->visc_mult, &t->beta,
                    &t->supp_torque, &t->sg_sqrt, &t->sg_normr, &t->sg_normt};
  for (size_t k = 0; k < sizeof(all) / sizeof(all[0]); k++) {
    free(*all[k]);
    *all[k] = NULL;
  }
}

int dfargo_fill_polar_1d_arrays(const dfargo_params *p, double *radii, double *rmed,
                                double *surf, double *azimuth) {
  dfargo_tables t;
  int rc = dfargo_check_params(p);
  if (rc) return rc;
  rc = dfargo_host_tables(p, &t);
  if (rc) return rc;
  if (radii) memcpy(radii, t.Radii, sizeof(double) * (p->nrad + 1));
  if (rmed) memcpy(rmed, t.Rmed, sizeof(double) * p->nrad);
  if (surf) memcpy(surf, t.Surf, sizeof(double) * p->nrad);
  if (azimuth) memcpy(azimuth, t.Azimuth, sizeof(double) * p->nsec);
  dfargo_free_tables(&t);
  return DFARGO_OK;
}

/* InitGasDensity (src/Pframeforce.c:297-341): SigmaMed plus the AddNoise
 * draw, drand48 stream seeded like srand48(noise_seed) */
int dfargo_init_gas_density(const dfargo_params *p, double *dens) {
  dfargo_tables t;
  unsigned short xs[3];
  int rc = dfargo_check_params(p);
  if (rc) return rc;
  if (!dens) return dfargo_set_error(DFARGO_ERR_ARG, "null density array");
  rc = dfargo_host_tables(p, &t);
  if (rc) return rc;
  xs[0] = 0x330E;
  xs[1] = (unsigned short)(p->noise_seed & 0xFFFF);
  xs[2] = (unsigned short)((p->noise_seed >> 16) & 0xFFFF);
  memset(dens, 0, sizeof(double) * (size_t)(p->nrad + 1) * p->nsec);
  for (int i = 0; i < p->nrad; i++)
    for (int j = 0; j < p->nsec; j++) {
      int l = j + i * p->nsec;
      dens[l] = t.SigmaMed[i];
      if (p->add_noise) {
        double randomnb = 2.0 * erand48(xs) - 1.0;
        dens[l] += 1e-3 * t.SigmaMed[i] * randomnb;
      }
    }
  dfargo_free_tables(&t);
  return DFARGO_OK;
}

/* InitGasEnergy (src/Pframeforce.c:422-440) */
int dfargo_init_gas_energy(const dfargo_params *p, double *energy) {
  dfargo_tables t;
  int rc = dfargo_check_params(p);
  if (rc) return rc;
  if (!energy) return dfargo_set_error(DFARGO_ERR_ARG, "null energy array");
  rc = dfargo_host_tables(p, &t);
  if (rc) return rc;
  memset(energy, 0, sizeof(double) * (size_t)(p->nrad + 1) * p->nsec);
  for (int i = 0; i < p->nrad; i++)
    for (int j = 0; j < p->nsec; j++)
      energy[j + i * p->nsec] = dfargo_energy_profile(p, t.Rmed[i]);
  dfargo_free_tables(&t);
  return DFARGO_OK;
}

/* InitPlanetarySystem (src/Psys.c:96-196); the .cfg columns are parsed with
 * %f (src/Psys.c:118), so they are rounded to float here as well */
int dfargo_init_planetary_system(const dfargo_params *p, int nb, const double *dist_in,
                                 const double *mass_in, const double *acc_in,
                                 const int *feel_disk, const int *feel_others,
                                 dfargo_planets *sys) {
  if (!p || !sys) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  if (nb < 0 || nb > DFARGO_MAX_PLANETS)
    return dfargo_set_error(DFARGO_ERR_ARG, "between 0 and %d planets are supported", DFARGO_MAX_PLANETS);
  memset(sys, 0, sizeof(*sys));
  sys->nb = nb;
  for (int i = 0; i < nb; i++) {
    float dist = (float)dist_in[i], mass = (float)mass_in[i], accret = (float)acc_in[i];
    double massinvelocity;
    if (p->forced_circular && (feel_disk[i] || feel_others[i])) /* src/Psys.c:146-155 */
      return dfargo_set_error(DFARGO_ERR_ARG, "FeelDisk/FeelOthers cannot be Yes with ForcedCircular");
    sys->mass[i] = (double)mass;
    sys->initial_mass[i] = 0.0;
    sys->final_mass[i] = sys->mass[i];
    massinvelocity = (p->mass_taper > 1e-3) ? 0.0 : sys->final_mass[i];
    sys->x[i] = (double)dist * (1.0 + p->eccentricity);
    sys->y[i] = 0.0;
    sys->vy[i] = (double)sqrt(G * (1.0 + massinvelocity) / dist) *
                 sqrt((1.0 - p->eccentricity) / (1.0 + p->eccentricity));
    sys->vx[i] = -0.0000000001 * sys->vy[i];
    sys->acc[i] = accret;
    sys->feel_disk[i] = feel_disk[i];
    sys->feel_others[i] = feel_others[i];
    sys->binary[i] = 0;
  }
  return DFARGO_OK;
}

/* error reporting */
static char g_err[512];

int dfargo_set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

const char *dfargo_last_error(void) { return g_err; }

/* ------------------------------------------------------------------------ */
/* dust particles                                                             */
/* ------------------------------------------------------------------------ */

/* src/var.c:136-181 and the booleans of src/Interpret.c:39-45 after the
 * default .par strings are decoded */
void dfargo_dust_default_params(dfargo_dust_params *d, int nbpart) {
  memset(d, 0, sizeof(*d));
  d->nbpart = nbpart;
  d->sizeminpart = 0.001;
  d->sizemaxpart = 0.001;
  d->sizepartslope = 3.0;
  d->rhopart = 1.0;
  d->dustslope = 0.5;
  d->rmindust = 0.0;
  d->rmaxdust = 0.0;
  d->interpolation = DFARGO_INTERP_TSC;
  d->dust_feel_disk = 1;
  d->dust_feel_sg = 1;
  d->dust_feel_planets = 1;
  d->dust_feedback = 0;
  d->short_friction_time_approximation = 1;
  d->no_dt_constraint_by_particles = 1;
  d->dust_timestep = 0.0;
  d->factor_unit_mass = 1.0;
  d->factor_unit_length = 1.0;
  d->particle_mass = 0.0;
  d->seed = 0;
}

int dfargo_check_dust_params(const dfargo_params *p, const dfargo_dust_params *d) {
  if (!p || !d) return dfargo_set_error(DFARGO_ERR_ARG, "null parameter block");
  int rc = dfargo_check_params(p);
  if (rc) return rc;
  if (d->nbpart < 0) return dfargo_set_error(DFARGO_ERR_ARG, "NBPART must be >= 0");
  if (d->nbpart > 0 && (d->sizeminpart == d->sizemaxpart) && (d->sizepartslope != 0.0)) /* src/Dsys.c:154-158 */
    return dfargo_set_error(DFARGO_ERR_ARG, "SIZEMINPART = SIZEMAXPART needs SIZEPARTSLOPE = 0");
  if (d->interpolation < DFARGO_INTERP_NGP || d->interpolation > DFARGO_INTERP_TSC)
    return dfargo_set_error(DFARGO_ERR_ARG, "INTERPOLATION must be NGP, CIC or TSC");
  if (!(d->sizeminpart > 0.0) || !(d->sizemaxpart > 0.0) || !(d->rhopart > 0.0))
    return dfargo_set_error(DFARGO_ERR_ARG, "particle sizes and RHOPART must be positive");
  if (!d->no_dt_constraint_by_particles && !(d->dust_timestep > 0.0))
    return dfargo_set_error(DFARGO_ERR_ARG, "a particle time step constraint needs DUSTTIMESTEP > 0");
  if (d->dust_feel_turb)
    return dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "DUSTFEELTURB (stochastic kicks) is not implemented");
  if (d->dust_growth) return dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "DUSTGROWTH is not implemented");
  if (d->remove_dust_from_hill)
    return dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "REMOVEDUSTFROMPLANETSHILLRADIUS is not implemented");
  if (d->zz_integrator) return dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "ZZINTEGRATOR is not implemented");
  if (d->dust_feel_sg_zero_mode)
    return dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "DUSTFEELSG Z (axisymmetric SG only) is not implemented");
  if (p->self_gravity && p->sg_zero_mode)
    return dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "particles with SELFGRAVITY Z are not implemented");
  return DFARGO_OK;
}

void dfargo_dust_fluid_default_params(dfargo_dust_fluid_params *f) {
  memset(f, 0, sizeof(*f));
  f->sizepart = 1e-3;
  f->rhopart = 1.0;
  f->factor_unit_mass = 1.0;
  f->factor_unit_length = 1.0;
  f->dust_to_gas_density_ratio = 0.01;
  f->daspect_ratio = 0.05;
  f->sfta = 1;
  f->dust_feel_disk = 1;
  f->dust_feel_sg = DFARGO_DUST_SG_YES;
}

int dfargo_check_dust_fluid_params(const dfargo_params *p, const dfargo_dust_fluid_params *f, int precision) {
  if (!p || !f) return dfargo_set_error(DFARGO_ERR_ARG, "null parameter block");
  int rc = dfargo_check_params(p);
  if (rc) return rc;
  if (!(f->sizepart > 0.0) || !(f->rhopart > 0.0))
    return dfargo_set_error(DFARGO_ERR_ARG, "SIZEPART and RHOPART must be positive");
  if (!(f->factor_unit_mass > 0.0) || !(f->factor_unit_length > 0.0))
    return dfargo_set_error(DFARGO_ERR_ARG, "FACTORUNITMASS and FACTORUNITLENGTH must be positive");
  if (!(f->dust_to_gas_density_ratio > 0.0))
    return dfargo_set_error(DFARGO_ERR_ARG, "DUSTTOGASDENSITYRATIO must be positive");
  if (f->dust_feel_sg < DFARGO_DUST_SG_NO || f->dust_feel_sg > DFARGO_DUST_SG_ZERO_MODE)
    return dfargo_set_error(DFARGO_ERR_ARG, "unknown DUSTFEELSG %d", f->dust_feel_sg);
  if (precision != DFARGO_FP64)
    return dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "the dust fluid runs with FP64 storage");
  if (f->sfta && !f->dust_feel_disk) /* src/SourceEuler.c:906-913 leaves DVradInt unset */
    return dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "SFTAPPROX YES with DUSTFEELDISK NO is not implemented");
  if (p->nrad < 6) return dfargo_set_error(DFARGO_ERR_ARG, "the dust fluid needs NRAD >= 6");
  return DFARGO_OK;
}

/* InitDustDensity (src/Pframeforce.c:344-380) with FillDSigma
 * (src/Theo.c:95-101), no noise options */
int dfargo_init_dust_fluid_density(const dfargo_params *p, const dfargo_dust_fluid_params *f, double *drho) {
  dfargo_tables t;
  int rc = dfargo_check_params(p);
  if (rc) return rc;
  if (!f || !drho) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  rc = dfargo_host_tables(p, &t);
  if (rc) return rc;
  memset(drho, 0, sizeof(double) * (size_t)(p->nrad + 1) * p->nsec);
  for (int i = 0; i < p->nrad; i++) {
    const double dsig = dfargo_sigma_profile(p, t.Rmed[i]) * f->dust_to_gas_density_ratio;
    for (int j = 0; j < p->nsec; j++) drho[j + i * p->nsec] = dsig;
  }
  dfargo_free_tables(&t);
  return DFARGO_OK;
}

/* InitDustSystem (src/Dsys.c:128-216), one process: sizes, then radii, then
 * azimuths from one drand48 stream seeded like srand48(seed) */
int dfargo_init_dust_system(const dfargo_params *p, const dfargo_dust_params *d, const double *axi_sg_accr,
                            dfargo_dust_state *out) {
  int rc = dfargo_check_dust_params(p, d);
  if (rc) return rc;
  if (!out || !out->r || !out->th || !out->vr || !out->vth || !out->l || !out->dustsize || !out->dustsize_init)
    return dfargo_set_error(DFARGO_ERR_ARG, "null particle array");
  if (p->self_gravity && !axi_sg_accr)
    return dfargo_set_error(DFARGO_ERR_ARG, "self-gravity needs the axisymmetric SG acceleration");
  dfargo_tables t;
  rc = dfargo_host_tables(p, &t);
  if (rc) return rc;
  const int n = d->nbpart, nr = p->nrad, ns = p->nsec;
  const double unit_mass = 1.9891e30 * d->factor_unit_mass, unit_length = 1.49598e11 * d->factor_unit_length;
  const double rmindust = (d->rmindust == 0.0) ? p->rmin : d->rmindust;
  const double rmaxdust = (d->rmaxdust == 0.0) ? p->rmax : d->rmaxdust;
  (void)unit_mass;
  unsigned short xs[3];
  xs[0] = 0x330E;
  xs[1] = (unsigned short)(d->seed & 0xFFFF);
  xs[2] = (unsigned short)((d->seed >> 16) & 0xFFFF);
  double C1, C2;
  if (d->sizeminpart != d->sizemaxpart) {
    if (d->sizepartslope != 1.0) {
      C1 = pow(d->sizemaxpart, 1.0 - d->sizepartslope);
      C2 = pow(d->sizeminpart, 1.0 - d->sizepartslope);
      for (int i = 0; i < n; i++)
        out->dustsize[i] = pow(C2 + (C1 - C2) * erand48(xs), 1.0 / (1.0 - d->sizepartslope)) / unit_length;
    } else {
      for (int i = 0; i < n; i++)
        out->dustsize[i] = d->sizeminpart * exp(erand48(xs) * log(d->sizemaxpart / d->sizeminpart)) / unit_length;
    }
  } else {
    for (int i = 0; i < n; i++) out->dustsize[i] = d->sizeminpart / unit_length;
  }
  if (d->dustslope != 1.0) {
    C1 = pow(rmaxdust, 1.0 - d->dustslope);
    C2 = pow(rmindust, 1.0 - d->dustslope);
    for (int i = 0; i < n; i++) out->r[i] = pow(C2 + (C1 - C2) * erand48(xs), 1.0 / (1.0 - d->dustslope));
  } else {
    for (int i = 0; i < n; i++) out->r[i] = rmindust * exp(erand48(xs) * log(rmaxdust / rmindust));
  }
  /* AziInf[0] and AziSup[ns-1] (src/SourceEuler.c:172-173) */
  const double aziinf0 = t.Azimuth[0] - 0.5 * (p->pmax - p->pmin) / (double)ns;
  const double azisup = t.Azimuth[ns - 1] + 0.5 * (p->pmax - p->pmin) / (double)ns;
  for (int i = 0; i < n; i++) out->th[i] = aziinf0 + erand48(xs) * (azisup - aziinf0);
  for (int i = 0; i < n; i++) {
    out->dustsize_init[i] = out->dustsize[i];
    out->vr[i] = 0.0;
    if (!p->self_gravity) {
      out->vth[i] = pow(out->r[i], -0.5);
    } else {
      const double dist = out->r[i];
      int ipl = 0;
      while ((t.Rmed[ipl] <= dist) && (ipl < nr - 2)) ipl++;
      const double ri = t.Rmed[ipl], rip1 = t.Rmed[ipl + 1], dr = rip1 - ri;
      double sgacc = (dist - ri) * axi_sg_accr[ipl + 1] + (rip1 - dist) * axi_sg_accr[ipl];
      sgacc /= dr;
      out->vth[i] = pow(out->r[i], -0.5) * sqrt(1.0 - dist * dist * sgacc);
    }
    out->l[i] = out->r[i] * out->vth[i];
    if (out->stokesnb) out->stokesnb[i] = 0.0;
  }
  dfargo_free_tables(&t);
  return DFARGO_OK;
}

/* FindOrbitalElements (src/Planet.c:129-180), G = 1 */
int dfargo_orbital_elements(double x, double y, double vx, double vy, double m, double *out) {
  if (!out) return dfargo_set_error(DFARGO_ERR_ARG, "null output");
  if (!(m > 0.0)) return dfargo_set_error(DFARGO_ERR_ARG, "mass %g is not positive", m);
  const double h = x * vy - y * vx;
  const double d = sqrt(x * x + y * y);
  const double Ax = x * vy * vy - y * vx * vy - m * x / d;
  const double Ay = y * vx * vx - x * vx * vy - m * y / d;
  const double e = sqrt(Ax * Ax + Ay * Ay) / m;
  const double a = h * h / m / (1 - e * e);
  double E = (e != 0.0) ? acos((1.0 - d / a) / e) : 0.0;
  if ((x * y * (vy * vy - vx * vx) + vx * vy * (x * x - y * y)) < 0) E = -E;
  const double M = E - e * sin(E);
  double V = (e != 0.0) ? acos((a * (1.0 - e * e) / d - 1.0) / e) : 0.0;
  if (E < 0.0) V = -V;
  const double lambda = atan2(y, x);
  out[0] = e;
  out[1] = a;
  out[2] = M;
  out[3] = V;
  out[4] = (e != 0.0) ? atan2(Ay, Ax) : atan2(y, x);
  out[5] = lambda;
  out[6] = (e != 0.0) ? 2.0 * atan2(Ay, 1. + Ax) : lambda;
  return DFARGO_OK;
}
