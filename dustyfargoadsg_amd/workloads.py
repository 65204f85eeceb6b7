"""Parameter blocks of the benchmark and parity workloads.

Each workload is a .par-equivalent dfargo_params plus a planet list
(the columns of a planet .cfg file: distance, mass, accretion, FeelDisk,
FeelOthers).  Sizes and physics follow the disc set-ups the hydro step is
measured on; the seed drives the drand48 AddNoise draw.
"""
import math

from ._capi import default_params, BC_WALL

DEFAULT_SEED = 20261202


def gas_planet(nrad=128, nsec=384, seed=DEFAULT_SEED, **over):
    """Locally isothermal gas disc with one q=1e-3 planet on a fixed circular
    orbit at r=1 (smoothing 0.6 H), constant viscosity 1e-5, FARGO on.
    Sigma = 6.3662e-4 r^-0.5 with 1e-3 uniform noise; log grid over
    [0.4, 2.5]; h = 0.05 flat."""
    p = default_params(
        nrad=nrad, nsec=nsec, rmin=0.4, rmax=2.5, log_grid=1,
        sigma0=6.3662e-4, sigma_slope=0.5, aspect_ratio=0.05, flaring_index=0.0,
        viscosity=1e-5, thickness_smoothing=0.6, forced_circular=1,
        add_noise=1, noise_seed=seed, boundary=BC_WALL, fast_transport=1)
    for k, v in over.items():
        setattr(p, k, v)
    planets = dict(dist=[1.0], mass=[1e-3], acc=[0.0], feel_disk=[0], feel_others=[0])
    return p, planets


def energy_sg(nrad=512, nsec=1536, seed=DEFAULT_SEED, rmax=4.0, **over):
    """Adiabatic (gamma 1.4) self-gravitating disc with beta cooling
    (beta = 10), alpha = 1e-3, no planet.  Sigma ~ r^-1 normalised so that
    the Toomre Q computed with the isothermal sound speed h r Omega is 2 at
    the outer edge: Sigma0 = h / (pi Q_min r_max).  T ~ r^-1 (h = 0.05 at
    r = 1, no flaring); SG softening 0.3 H; log grid over [0.4, rmax]."""
    h, qmin = 0.05, 2.0
    p = default_params(
        nrad=nrad, nsec=nsec, rmin=0.4, rmax=rmax, log_grid=1,
        sigma0=h / (math.pi * qmin * rmax), sigma_slope=1.0,
        aspect_ratio=h, flaring_index=0.0, alpha_viscosity=1e-3,
        energy_equation=1, adiabatic_index=1.4, beta_cooling=1,
        beta_cooling_time=10.0, viscous_heating=1,
        self_gravity=1, sg_thickness_smoothing=0.3,
        add_noise=1, noise_seed=seed, boundary=BC_WALL, fast_transport=1)
    for k, v in over.items():
        setattr(p, k, v)
    planets = dict(dist=[], mass=[], acc=[], feel_disk=[], feel_others=[])
    return p, planets


def gas_sg_large(nrad=1024, nsec=3072, seed=DEFAULT_SEED, **over):
    """Gas disc with self-gravity at 1024x3072 over [0.4, 2.5]: alpha 1e-3,
    one q=1e-3 planet at r=1, isothermal h = 0.05, SG softening 0.3 H."""
    p = default_params(
        nrad=nrad, nsec=nsec, rmin=0.4, rmax=2.5, log_grid=1,
        sigma0=6.3662e-4, sigma_slope=0.5, aspect_ratio=0.05,
        alpha_viscosity=1e-3, thickness_smoothing=0.6, forced_circular=1,
        self_gravity=1, sg_thickness_smoothing=0.3,
        add_noise=1, noise_seed=seed, boundary=BC_WALL, fast_transport=1)
    for k, v in over.items():
        setattr(p, k, v)
    planets = dict(dist=[1.0], mass=[1e-3], acc=[0.0], feel_disk=[0], feel_others=[0])
    return p, planets


def dust_particles(params, nbpart, dust_to_gas=0.01, **over):
    """Particle switches for `nbpart` super-particles over [rmindust,
    rmaxdust] (default: the grid), with the super-particle mass of
    src/main.c:190-203: the dust disc mass between RMINDUST and RMAXDUST
    over NBPART, the dust density being dust_to_gas x Sigma
    (DUSTTOGASDENSITYRATIO, src/Theo.c:66-72) taken without its noise."""
    import ctypes

    import numpy as np

    from . import _capi
    d = _capi.dust_default_params(nbpart, sizepartslope=0.0)
    for k, v in over.items():
        if not hasattr(d, k):
            raise KeyError(f"unknown dust parameter {k!r}")
        setattr(d, k, v)
    if "particle_mass" not in over and nbpart > 0:
        lib = _capi.load()
        nr, ns = params.nrad, params.nsec
        radii, rmed, surf, az = np.zeros(nr + 1), np.zeros(nr), np.zeros(nr), np.zeros(ns)
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        lib.dfargo_fill_polar_1d_arrays(ctypes.byref(params), dp(radii), dp(rmed), dp(surf), dp(az))
        rmin = d.rmindust if d.rmindust != 0.0 else params.rmin
        rmax = d.rmaxdust if d.rmaxdust != 0.0 else params.rmax
        mass = 0.0
        for i in range(nr):
            if rmin <= radii[i] <= rmax:
                r = rmed[i]
                sig = params.sigma0 * r ** (-params.sigma_slope)
                if r < params.cavity_radius:
                    sig /= params.cavity_ratio
                mass += sig * dust_to_gas * 0.5 * (params.pmax - params.pmin) * (radii[i + 1] ** 2 - radii[i] ** 2)
        d.particle_mass = mass / nbpart
    return d


# Sigma0 (M_sun / AU^2) at which a 1 cm grain of 2.7 g/cc has the Epstein
# Stokes number pi/2 s rho_int / Sigma = 1 at r = 1, so that 1 um - 1 cm
# grains span St ~ 1e-4 - 1
SIGMA0_ST1_CM = 4.77e-7


def gas_dust(nrad=1024, nsec=3072, seed=DEFAULT_SEED, **over):
    """Gas plus Lagrangian dust: the gas disc of gas_sg_large without
    self-gravity (alpha 1e-3, one q=1e-3 planet at r=1, h = 0.05) at
    1024x3072, with Sigma0 = 4.77e-7 so that the grains of
    gas_dust_particles span Stokes numbers ~1e-4 - 1; the particles are
    given by gas_dust_particles."""
    kw = dict(sigma0=SIGMA0_ST1_CM)
    kw.update(over)
    return gas_sg_large(nrad=nrad, nsec=nsec, seed=seed, self_gravity=0, **kw)


def gas_dust_particles(params, nbpart=1_000_000, seed=DEFAULT_SEED, **over):
    """10^6 particles uniform in r over [0.5, 2.0] (DUSTSLOPE 0) and in
    azimuth, sizes 1 um - 1 cm with dn/ds ~ s^-3.5, internal density
    2.7 g/cc (Stokes numbers ~1e-5 - 1), CIC interpolation, drag feedback
    on the gas, dust-to-gas mass ratio 0.01."""
    kw = dict(sizeminpart=1e-6, sizemaxpart=1e-2, sizepartslope=3.5, rhopart=2.7, dustslope=0.0,
              rmindust=0.5, rmaxdust=2.0, interpolation=1, dust_feedback=1, seed=seed)
    kw.update(over)
    return dust_particles(params, nbpart, **kw)


def batch_discs(n=64, seed=DEFAULT_SEED, nrads=(128, 256, 512)):
    """n independent isothermal planet discs of uneven grids and parameters
    for the batched sweep: Nrad drawn from `nrads`, Nsec = 3 Nrad, log grid
    over [0.4, 2.5]; planet mass ratio log-uniform in [1e-5, 3e-3] on a
    fixed circular orbit at r = 1 (smoothing 0.6 H), alpha log-uniform in
    [1e-4, 1e-2], h uniform in [0.03, 0.08], flaring index in [0, 0.25],
    Sigma ~ r^-p with p in [0.5, 1.5]; all drawn from `seed`.  Returns a
    list of (params, planets)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        nrad = int(rng.choice(nrads))
        q = float(10 ** rng.uniform(math.log10(1e-5), math.log10(3e-3)))
        alpha = float(10 ** rng.uniform(-4.0, -2.0))
        p = default_params(
            nrad=nrad, nsec=3 * nrad, rmin=0.4, rmax=2.5, log_grid=1,
            sigma0=6.3662e-4, sigma_slope=float(rng.uniform(0.5, 1.5)),
            aspect_ratio=float(rng.uniform(0.03, 0.08)), flaring_index=float(rng.uniform(0.0, 0.25)),
            alpha_viscosity=alpha, thickness_smoothing=0.6, forced_circular=1,
            add_noise=1, noise_seed=seed + k, boundary=BC_WALL, fast_transport=1)
        out.append((p, dict(dist=[1.0], mass=[q], acc=[0.0], feel_disk=[0], feel_others=[0])))
    return out


DUST_WORKLOADS = {"gas_dust": gas_dust_particles}

WORKLOADS = {
    "gas_planet": gas_planet,
    "energy_sg": energy_sg,
    "gas_sg_large": gas_sg_large,
    "gas_dust": gas_dust,
}
