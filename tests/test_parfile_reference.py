"""The reference's own example inputs (in/gasonly.par, in/template.par,
in/energyeq.par, in/stochturb.par, in/template.cfg: the only fixtures it
ships) through dfargo_read_par / dfargo_read_planet_config.  Each value
checked is the one written in the file (or src/var.c's default where the
file is silent); the two files that switch on physics this library does not
implement are refused, naming the key.  Skipped where the reference sources
are not present (the GPU box)."""
import os

import pytest

from dustyfargoadsg_amd import _capi
from dustyfargoadsg_amd.disc import DfargoError, read_dust_fluid_par, read_par, read_planet_config

IN = "/root/reference/in"
pytestmark = pytest.mark.skipif(not os.path.isdir(IN), reason="the reference's in/ files are not present")


def test_gasonly_par():
    p, run = read_par(os.path.join(IN, "gasonly.par"))
    assert (p.nrad, p.nsec, p.rmin, p.rmax, p.log_grid) == (200, 400, 0.4, 2.2, 0)  # RadialSpacing A
    assert (p.sigma0, p.sigma_slope, p.aspect_ratio, p.flaring_index) == (1e-4, 0.5, 0.05, 0.25)
    assert (p.alpha_viscosity, p.self_gravity, p.add_noise, p.energy_equation) == (1e-3, 0, 0, 0)
    assert p.adiabatic_index == 1.6667 and p.thickness_smoothing == 0.6 and p.mass_taper == 5.0
    assert (p.forced_circular, p.eccentricity, p.fast_transport) == (0, 0.0, 1)
    assert p.corotating == _capi.FRAME_COROTATING  # Frame C
    assert (p.indirect_term, p.discard_gas_indirect_term) == (1, 1)
    assert p.boundary == _capi.BC_NONREFLECTING  # InnerBoundary Non-Reflecting
    assert p.wave_killing_zones == 0 and p.damp_to == _capi.DAMP_TO_INI  # DampToAxi No, DampToViscous No
    assert (p.nrad1d, p.rmin1d, p.rmax1d) == (1000, 0.04, 20.0)
    assert (p.add_floors, p.density_floor) == (0, 1e-9)
    assert (run.dt, run.ntot, run.ninterm) == (0.314159265358, 400, 20)
    assert run.output_dir == b"out1" and run.planet_config == b"in/template.cfg"


def test_template_par():
    p, run = read_par(os.path.join(IN, "template.par"))
    assert (p.nrad, p.nsec, p.log_grid, p.alpha_viscosity, p.mass_taper) == (200, 400, 1, 1e-4, 1.0)
    assert p.corotating == _capi.FRAME_COROTATING and p.boundary == _capi.BC_NONREFLECTING
    assert (run.ntot, run.ninterm, run.output_dir) == (200, 20, b"out1_wall_damp")


def test_template_par_dust_fluid_switches():
    """template.par leaves DustFluid off but sets its switches: they decode
    as written (DustDiffusion Yes, DustToGasDensityRatio 1e-2, DAspectRatio
    0, Rhopart 1.3, FactorUnitLength 10, DustFeelSG No, SFTApprox No,
    InnerBoundaryDust O) with src/var.c's defaults for the rest."""
    f, on = read_dust_fluid_par(os.path.join(IN, "template.par"))
    assert not on and f.open_inner_dust == 1
    assert (f.dust_diffusion, f.dust_to_gas_density_ratio, f.daspect_ratio) == (1, 1e-2, 0.0)
    assert (f.rhopart, f.factor_unit_length, f.factor_unit_mass) == (1.3, 10.0, 1.0)
    assert (f.dust_feel_sg, f.sfta, f.dust_feel_disk, f.dust_feedback) == (_capi.DUST_SG_NO, 0, 1, 0)
    assert (f.sizepart, f.dviscosity, f.dalpha_viscosity, f.dflaring_index) == (1e-3, 0.0, 0.0, 0.0)


def test_template_cfg():
    p, _ = read_par(os.path.join(IN, "gasonly.par"))
    s = read_planet_config(os.path.join(IN, "template.cfg"), p)
    # "MyPlanet 1.0 1e-3 0.0 No No", the columns read with %f
    assert s.nb == 1 and s.x[0] == 1.0 and s.y[0] == 0.0
    assert s.mass[0] == float(__import__("numpy").float32(1e-3))
    assert (s.acc[0], s.feel_disk[0], s.feel_others[0]) == (0.0, 0, 0)
    assert abs(s.vy[0] - 1.0) < 1e-3


@pytest.mark.parametrize("name,key", [("energyeq", "THERMALCOOLING"), ("stochturb", "TEMPPRESC")])
def test_unimplemented_physics_is_refused(name, key):
    with pytest.raises(DfargoError) as e:
        read_par(os.path.join(IN, f"{name}.par"))
    assert e.value.code == _capi.ERR_UNSUPPORTED and key in str(e.value)
