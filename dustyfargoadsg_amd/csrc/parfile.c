/* Reading the reference's input files into the C ABI's parameter block:
 * the .par file (ReadVariables and the decoding that follows it,
 * src/Interpret.c:82-480, with the keys and defaults of src/var.c:15-202)
 * and the planet .cfg file (InitPlanetarySystem, src/Psys.c:96-196).
 *
 * Only the keys the hydro step reads are kept; the others are accepted and
 * ignored as the reference ignores unknown ones.  A file that switches on
 * physics this library does not implement is refused with
 * DFARGO_ERR_UNSUPPORTED naming the key, rather than run without it.  The
 * particles' keys (NBPART, SIZEMINPART, ...) are not read here: the particle
 * set is described by dfargo_dust_params.  The dust fluid's switches have
 * a reader of their own (dfargo_read_dust_fluid_par). */
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dfargo_internal.h"

enum { K_REAL, K_INT, K_STR };

/* one key: its type, whether the reference requires it, and its default
 * (src/var.c) as the reference's text */
typedef struct {
  const char *name;
  int type, necessary;
  const char *deflt;
} par_key;

static const par_key kKeys[] = {
    {"DT", K_REAL, 1, "1."}, {"SIGMA0", K_REAL, 1, "173."}, {"NINTERM", K_INT, 1, "10."},
    {"NTOT", K_INT, 1, "1501."}, {"OUTPUTDIR", K_STR, 1, "~masset"}, {"INNERBOUNDARY", K_STR, 0, "WALL"},
    {"TRANSPORT", K_STR, 0, "FAST"}, {"PLANETCONFIG", K_STR, 0, "Systems/SolarSystem.cfg"},
    {"MASSTAPER", K_REAL, 0, "0.0000001"}, {"RADIALSPACING", K_STR, 0, "ARITHMETIC"},
    {"NRAD", K_INT, 1, "64.0"}, {"NSEC", K_INT, 1, "64.0"}, {"RMIN", K_REAL, 1, "1.0"},
    {"RMAX", K_REAL, 1, "1.0"}, {"THICKNESSSMOOTHING", K_REAL, 0, "0.3"},
    {"SGTHICKNESSSMOOTHING", K_REAL, 0, "0.3"}, {"ROCHESMOOTHING", K_REAL, 0, "0.0"},
    {"INDIRECTTERMSMOOTHING", K_REAL, 0, "0.0"}, {"ASPECTRATIO", K_REAL, 1, "0.05"},
    {"VISCOSITY", K_REAL, 0, "0.0"}, {"ALPHAVISCOSITY", K_REAL, 0, "0.0"}, {"SIGMASLOPE", K_REAL, 1, "0.0"},
    {"RELEASERADIUS", K_REAL, 0, "0.0"}, {"RELEASEDATE", K_REAL, 0, "0.0"}, {"OMEGAFRAME", K_REAL, 0, "0.0"},
    {"FRAME", K_STR, 0, "FIXED"}, {"INDIRECTTERM", K_STR, 0, "YES"},
    {"DISCARDGASINDIRECTTERM", K_STR, 0, "NO"}, {"ALTERNATIVEGASINDIRECTTERM", K_STR, 0, "NO"},
    {"EXCLUDEHILL", K_STR, 0, "NO"}, {"EXCLUDEHILLFACTOR", K_REAL, 0, "1.0"},
    {"IMPOSEDDISKDRIFT", K_REAL, 0, "0.0"}, {"FLARINGINDEX", K_REAL, 0, "0.0"},
    {"ECCENTRICITY", K_REAL, 0, "0.0"}, {"CAVITYRADIUS", K_REAL, 0, "0.0"}, {"CAVITYRATIO", K_REAL, 0, "1.0"},
    {"CAVITYWIDTH", K_REAL, 0, "1.0"}, {"TRANSITIONRADIUS", K_REAL, 0, "0.0"},
    {"TRANSITIONRATIO", K_REAL, 0, "1.0"}, {"TRANSITIONWIDTH", K_REAL, 0, "1.0"},
    {"SELFGRAVITY", K_STR, 0, "NO"}, {"FORCEDCIRCULAR", K_STR, 0, "NO"}, {"ENERGYEQUATION", K_STR, 0, "NO"},
    {"ADIABATICINDEX", K_REAL, 1, "1.4"}, {"VISCOUSHEATING", K_STR, 0, "YES"}, {"PMIN", K_REAL, 0, "0.0"},
    {"PMAX", K_REAL, 0, "6.2831853071795864"}, {"BETACOOLING", K_STR, 0, "NO"},
    {"BETACOOLINGTIME", K_REAL, 0, "10.0"}, {"BETACOOLINGSLOPE", K_REAL, 0, "0.0"},
    {"ADDNOISE", K_STR, 0, "NO"}, {"WAVEKILLINGZONES", K_STR, 0, "NO"},
    {"DAMPINGTODYNAMICALTIMERATIO", K_REAL, 0, "0.3"}, {"WKZRMIN", K_REAL, 0, "0.0"},
    {"WKZRMAX", K_REAL, 0, "0.0"}, {"DAMPTOINI", K_STR, 0, "NO"}, {"DAMPTOAXI", K_STR, 0, "YES"},
    {"DAMPTOVISCOUS", K_STR, 0, "NO"}, {"COROTATEWITHOUTERPLANET", K_STR, 0, "NO"},
    {"ADDFLOORS", K_STR, 0, "YES"}, {"DENSITYFLOOR", K_REAL, 0, "1e-9"}, {"ZEROVRADBC", K_STR, 0, "NO"},
    {"NRAD1D", K_INT, 0, "100"}, {"RMIN1D", K_REAL, 0, "0.01"}, {"RMAX1D", K_REAL, 0, "100.0"},
    {"BOUNDARY1DGRID", K_STR, 0, "T"},
    /* the dust fluid (dfargo_read_dust_fluid_par) */
    {"DUSTFLUID", K_STR, 0, "NO"}, {"SIZEPART", K_REAL, 0, "1e-3"}, {"RHOPART", K_REAL, 0, "1.0"},
    {"FACTORUNITMASS", K_REAL, 0, "1.0"}, {"FACTORUNITLENGTH", K_REAL, 0, "1.0"},
    {"DUSTTOGASDENSITYRATIO", K_REAL, 0, "0.01"}, {"DVISCOSITY", K_REAL, 0, "0.0"},
    {"DALPHAVISCOSITY", K_REAL, 0, "0.0"}, {"DASPECTRATIO", K_REAL, 0, "0.05"},
    {"DFLARINGINDEX", K_REAL, 0, "0.0"}, {"DUSTDIFFINN", K_REAL, 0, "0.0"}, {"SFTAPPROX", K_STR, 0, "YES"},
    {"DUSTFEELDISK", K_STR, 0, "YES"}, {"DUSTFEEDBACK", K_STR, 0, "NO"}, {"DUSTFEELSG", K_STR, 0, "YES"},
    {"DUSTDIFFUSION", K_STR, 0, "NO"}, {"INNERBOUNDARYDUST", K_STR, 0, "WALL"},
    {"RESTARTWITHNEWDUST", K_STR, 0, "NO"},
    /* switches of physics not implemented here (refused when on) */
    {"OUTERSOURCEMASS", K_STR, 0, "NO"}, {"ZMPLUS", K_STR, 0, "NO"}, {"ENTROPYDIFFUSION", K_STR, 0, "NO"},
    {"RADIATIVEDIFFUSION", K_STR, 0, "NO"}, {"THERMALCOOLING", K_STR, 0, "NO"},
    {"STELLARIRRADIATION", K_STR, 0, "NO"}, {"TEMPPRESC", K_STR, 0, "NO"}, {"ADDMASS", K_STR, 0, "NO"},
    {"DISCEVAPORATION", K_STR, 0, "NO"}, {"PHOTOEVAPORATION", K_STR, 0, "NO"},
    {"CAVITYTORQUE", K_STR, 0, "NO"},
};
enum { NKEYS = sizeof(kKeys) / sizeof(kKeys[0]) };

typedef struct {
  double r[NKEYS];
  char s[NKEYS][290];
  int read[NKEYS];
} par_values;

static int key_index(const char *upper) {
  for (int i = 0; i < NKEYS; i++)
    if (strcmp(upper, kKeys[i].name) == 0) return i;
  return -1;
}

static double R(const par_values *v, const char *k) { return v->r[key_index(k)]; }
static int I(const par_values *v, const char *k) { return (int)v->r[key_index(k)]; }
static char C(const par_values *v, const char *k) { return (char)toupper((unsigned char)v->s[key_index(k)][0]); }

/* ReadVariables (src/Interpret.c:82-140) into *out (freed by the caller) */
static int read_values(const char *path, par_values **out) {
  par_values *v = (par_values *)calloc(1, sizeof(par_values));
  if (!v) return dfargo_set_error(DFARGO_ERR_ALLOC, "out of host memory");
  /* defaults: var() reads the default text with %lf for every type and
     copies it as the string (src/Interpret.c:52-80) */
  for (int i = 0; i < NKEYS; i++) {
    double d = 0.0;
    sscanf(kKeys[i].deflt, "%lf", &d);
    v->r[i] = (kKeys[i].type == K_INT) ? (double)(int)d : d;
    snprintf(v->s[i], sizeof v->s[i], "%s", kKeys[i].deflt);
  }
  FILE *in = fopen(path, "r");
  if (!in) {
    free(v);
    return dfargo_set_error(DFARGO_ERR_ARG, "unable to read '%s'", path);
  }
  /* src/Interpret.c:101-135: the first word is the key (case-insensitive),
     the value follows any of "\t :=>_"; a '#' begins a comment line */
  char s[350], nm[300], sv[290];
  double temp = 0.0;
  while (fgets(s, 349, in) != NULL) {
    if (sscanf(s, "%299s ", nm) != 1 || nm[0] == '#') continue;
    const char *s1 = s + strlen(nm);
    sscanf(s1 + strspn(s1, "\t :=>_"), "%lf", &temp);
    sv[0] = 0;
    sscanf(s1 + strspn(s1, "\t :=>_"), "%289s ", sv);
    for (char *q = nm; *q; q++) *q = (char)toupper((unsigned char)*q);
    const int k = key_index(nm);
    if (k < 0) continue;
    v->read[k] = 1;
    if (kKeys[k].type == K_INT) v->r[k] = (double)(int)temp;
    else if (kKeys[k].type == K_REAL) v->r[k] = temp;
    else snprintf(v->s[k], sizeof v->s[k], "%s", sv);
  }
  fclose(in);
  int rc = DFARGO_OK;
  for (int i = 0; i < NKEYS && rc == DFARGO_OK; i++)
    if (kKeys[i].necessary && !v->read[i])
      rc = dfargo_set_error(DFARGO_ERR_ARG, "'%s' lacks the mandatory variable %s", path, kKeys[i].name);
  if (rc != DFARGO_OK) {
    free(v);
    return rc;
  }
  *out = v;
  return DFARGO_OK;
}

int dfargo_read_par(const char *path, dfargo_params *p, dfargo_run_params *run) {
  if (!path || !p) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  par_values *v = NULL;
  int rc = read_values(path, &v);
  if (rc) return rc;
  /* physics this library does not implement */
  static const char *const kOff[] = {"OUTERSOURCEMASS", "ZMPLUS",   "ENTROPYDIFFUSION", "THERMALCOOLING",
                                     "TEMPPRESC",       "ADDMASS",  "DISCEVAPORATION",  "PHOTOEVAPORATION",
                                     "CAVITYTORQUE"};
  for (size_t i = 0; i < sizeof kOff / sizeof kOff[0] && rc == DFARGO_OK; i++)
    if (C(v, kOff[i]) == 'Y') rc = dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "%s = YES is not implemented", kOff[i]);
  if (rc == DFARGO_OK && (C(v, "RADIATIVEDIFFUSION") == 'E' || C(v, "RADIATIVEDIFFUSION") == 'I'))
    rc = dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "RADIATIVEDIFFUSION is not implemented");
  if (rc == DFARGO_OK && (C(v, "STELLARIRRADIATION") == 'Y' || C(v, "STELLARIRRADIATION") == 'L'))
    rc = dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "STELLARIRRADIATION is not implemented");
  const char bc = C(v, "INNERBOUNDARY");
  if (rc == DFARGO_OK && bc != 'W' && bc != 'O' && bc != 'N')
    rc = dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "INNERBOUNDARY %s is not implemented", v->s[key_index("INNERBOUNDARY")]);
  if (rc != DFARGO_OK) {
    free(v);
    return rc;
  }
  /* the decoding of src/Interpret.c:165-476 */
  dfargo_default_params(p);
  p->nrad = I(v, "NRAD");
  p->nsec = I(v, "NSEC");
  p->rmin = R(v, "RMIN");
  p->rmax = R(v, "RMAX");
  p->log_grid = C(v, "RADIALSPACING") == 'L';
  p->pmin = R(v, "PMIN");
  p->pmax = R(v, "PMAX");
  p->sigma0 = R(v, "SIGMA0");
  p->sigma_slope = R(v, "SIGMASLOPE");
  p->aspect_ratio = R(v, "ASPECTRATIO");
  p->flaring_index = R(v, "FLARINGINDEX");
  p->viscosity = R(v, "VISCOSITY");
  p->alpha_viscosity = R(v, "ALPHAVISCOSITY");
  p->cavity_radius = R(v, "CAVITYRADIUS");
  p->cavity_ratio = R(v, "CAVITYRATIO");
  p->cavity_width = R(v, "CAVITYWIDTH");
  p->transition_radius = R(v, "TRANSITIONRADIUS");
  p->transition_ratio = R(v, "TRANSITIONRATIO");
  p->transition_width = R(v, "TRANSITIONWIDTH");
  p->imposed_disk_drift = R(v, "IMPOSEDDISKDRIFT");
  /* the reference seeds the noise with time(NULL); the caller sets
     noise_seed for a reproducible draw */
  p->add_noise = C(v, "ADDNOISE") == 'Y';
  p->adiabatic_index = R(v, "ADIABATICINDEX");
  /* an energy equation with ADIABATICINDEX 1 falls back to the locally
     isothermal one (src/Interpret.c:399-402) */
  p->energy_equation = (C(v, "ENERGYEQUATION") == 'Y') && !(p->adiabatic_index == 1.0);
  p->viscous_heating = C(v, "VISCOUSHEATING") != 'N';
  p->beta_cooling = C(v, "BETACOOLING") == 'Y';
  p->beta_cooling_time = R(v, "BETACOOLINGTIME");
  p->beta_cooling_slope = R(v, "BETACOOLINGSLOPE");
  p->self_gravity = C(v, "SELFGRAVITY") == 'Y' || C(v, "SELFGRAVITY") == 'Z';
  p->sg_zero_mode = C(v, "SELFGRAVITY") == 'Z';
  p->sg_thickness_smoothing = R(v, "SGTHICKNESSSMOOTHING");
  p->thickness_smoothing = R(v, "THICKNESSSMOOTHING");
  p->roche_smoothing = R(v, "ROCHESMOOTHING");
  p->indirect_term_smoothing = R(v, "INDIRECTTERMSMOOTHING");
  p->mass_taper = R(v, "MASSTAPER");
  p->indirect_term = C(v, "INDIRECTTERM") != 'N';
  p->discard_gas_indirect_term = C(v, "DISCARDGASINDIRECTTERM") == 'Y';
  p->alternative_gas_indirect_term = C(v, "ALTERNATIVEGASINDIRECTTERM") == 'Y';
  p->exclude_hill = C(v, "EXCLUDEHILL") == 'Y';
  p->exclude_hill_factor = R(v, "EXCLUDEHILLFACTOR");
  p->forced_circular = C(v, "FORCEDCIRCULAR") == 'Y';
  p->release_date = R(v, "RELEASEDATE");
  p->release_radius = R(v, "RELEASERADIUS");
  p->eccentricity = R(v, "ECCENTRICITY");
  const char fr = C(v, "FRAME");
  p->corotating = fr == 'C'   ? DFARGO_FRAME_COROTATING
                  : fr == 'G' ? DFARGO_FRAME_GUIDING_CENTER
                  : fr == 'B' ? DFARGO_FRAME_BINARY_CENTER
                              : DFARGO_FRAME_FIXED;
  p->corotate_with_outer_planet = C(v, "COROTATEWITHOUTERPLANET") == 'Y';
  /* OMEGAFRAME; with a corotating frame the host replaces it by the planet's
     frequency before InitEuler (src/main.c:228-246) */
  p->omega_frame = R(v, "OMEGAFRAME");
  p->fast_transport = C(v, "TRANSPORT") != 'S';
  p->add_floors = C(v, "ADDFLOORS") != 'N';
  p->density_floor = R(v, "DENSITYFLOOR");
  p->boundary = bc == 'O' ? DFARGO_BC_OPEN : bc == 'N' ? DFARGO_BC_NONREFLECTING : DFARGO_BC_WALL;
  p->wave_killing_zones = C(v, "WAVEKILLINGZONES") == 'Y';
  p->wkz_rmin = R(v, "WKZRMIN");
  p->wkz_rmax = R(v, "WKZRMAX");
  p->damping_ratio = R(v, "DAMPINGTODYNAMICALTIMERATIO");
  /* src/Interpret.c:344-358, in that order: the last switch set wins */
  p->damp_to = DFARGO_DAMP_TO_AXI;
  if (C(v, "DAMPTOINI") == 'Y') p->damp_to = DFARGO_DAMP_TO_INI;
  if (C(v, "DAMPTOAXI") == 'N') p->damp_to = DFARGO_DAMP_TO_INI;
  if (C(v, "DAMPTOVISCOUS") == 'Y') p->damp_to = DFARGO_DAMP_TO_VISCOUS;
  p->zero_vrad_bc = C(v, "ZEROVRADBC") == 'Y';
  p->nrad1d = I(v, "NRAD1D");
  p->rmin1d = R(v, "RMIN1D");
  p->rmax1d = R(v, "RMAX1D");
  p->bc1d = C(v, "BOUNDARY1DGRID") == 'Z' ? 1 : C(v, "BOUNDARY1DGRID") == 'D' ? 2 : 0;
  if (run) {
    run->dt = R(v, "DT");
    run->ntot = I(v, "NTOT");
    run->ninterm = I(v, "NINTERM");
    snprintf(run->planet_config, sizeof run->planet_config, "%s", v->s[key_index("PLANETCONFIG")]);
    snprintf(run->output_dir, sizeof run->output_dir, "%s", v->s[key_index("OUTPUTDIR")]);
  }
  free(v);
  return dfargo_check_params(p);
}

/* The dust-fluid switches of the same file (src/Interpret.c:253-342 and
 * src/var.c's defaults).  *dust_fluid is DUSTFLUID's value; the switches
 * are decoded either way.  RESTARTWITHNEWDUST YES is refused (not
 * implemented). */
int dfargo_read_dust_fluid_par(const char *path, dfargo_dust_fluid_params *f, int *dust_fluid) {
  if (!path || !f) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  par_values *v = NULL;
  int rc = read_values(path, &v);
  if (rc) return rc;
  dfargo_dust_fluid_default_params(f);
  if (dust_fluid) *dust_fluid = C(v, "DUSTFLUID") == 'Y';
  f->sizepart = R(v, "SIZEPART");
  f->rhopart = R(v, "RHOPART");
  f->factor_unit_mass = R(v, "FACTORUNITMASS");
  f->factor_unit_length = R(v, "FACTORUNITLENGTH");
  f->dust_to_gas_density_ratio = R(v, "DUSTTOGASDENSITYRATIO");
  f->dviscosity = R(v, "DVISCOSITY");
  f->dalpha_viscosity = R(v, "DALPHAVISCOSITY");
  f->daspect_ratio = R(v, "DASPECTRATIO");
  f->dflaring_index = R(v, "DFLARINGINDEX");
  f->dust_diff_inn = R(v, "DUSTDIFFINN");
  f->sfta = C(v, "SFTAPPROX") != 'N';
  f->dust_feel_disk = C(v, "DUSTFEELDISK") != 'N';
  f->dust_feedback = C(v, "DUSTFEEDBACK") == 'Y';
  f->dust_feel_sg = C(v, "DUSTFEELSG") == 'N'   ? DFARGO_DUST_SG_NO
                    : C(v, "DUSTFEELSG") == 'Z' ? DFARGO_DUST_SG_ZERO_MODE
                                                : DFARGO_DUST_SG_YES;
  f->dust_diffusion = C(v, "DUSTDIFFUSION") == 'Y';
  f->open_inner_dust = C(v, "INNERBOUNDARYDUST") == 'O';
  if (C(v, "RESTARTWITHNEWDUST") == 'Y')
    rc = dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "RESTARTWITHNEWDUST is not implemented");
  free(v);
  return rc;
}

int dfargo_read_planet_config(const char *path, const dfargo_params *p, dfargo_planets *sys) {
  if (!path || !p || !sys) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  FILE *in = fopen(path, "r");
  if (!in) return dfargo_set_error(DFARGO_ERR_ARG, "unable to read '%s'", path);
  double dist[DFARGO_MAX_PLANETS], mass[DFARGO_MAX_PLANETS], acc[DFARGO_MAX_PLANETS];
  int fd[DFARGO_MAX_PLANETS], fo[DFARGO_MAX_PLANETS];
  int nb = 0;
  char s[512], nm[512], t1[512], t2[512], t3[512];
  /* a line that starts with a letter is a planet: name, distance, mass,
     accretion, FeelDisk, FeelOthers[, Binary] (src/Psys.c:114-160) */
  while (fgets(s, 510, in) != NULL) {
    if (!isalpha((unsigned char)s[0])) continue;
    if (sscanf(s, "%511s ", nm) != 1) continue;
    if (nb == DFARGO_MAX_PLANETS) {
      fclose(in);
      return dfargo_set_error(DFARGO_ERR_ARG, "more than %d planets in '%s'", DFARGO_MAX_PLANETS, path);
    }
    const char *s1 = s + strlen(nm);
    float fdist = 0.f, fmass = 0.f, facc = 0.f;
    t1[0] = t2[0] = t3[0] = 0;
    sscanf(s1 + strspn(s1, "\t :=>_"), "%f %f %f %511s %511s %511s", &fdist, &fmass, &facc, t1, t2, t3);
    if (tolower((unsigned char)t3[0]) == 'y') {
      fclose(in);
      return dfargo_set_error(DFARGO_ERR_UNSUPPORTED, "binary planets (%s) are not implemented", nm);
    }
    dist[nb] = fdist;
    mass[nb] = fmass;
    acc[nb] = facc;
    fd[nb] = tolower((unsigned char)t1[0]) != 'n';
    fo[nb] = tolower((unsigned char)t2[0]) != 'n';
    nb++;
  }
  fclose(in);
  return dfargo_init_planetary_system(p, nb, dist, mass, acc, fd, fo, sys);
}
