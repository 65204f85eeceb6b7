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


WORKLOADS = {
    "gas_planet": gas_planet,
    "energy_sg": energy_sg,
    "gas_sg_large": gas_sg_large,
}
