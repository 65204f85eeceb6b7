"""The .par and planet .cfg readers (ReadVariables and its decoding,
src/Interpret.c:82-476 with the defaults of src/var.c:15-202;
InitPlanetarySystem, src/Psys.c:96-196).  Host code only: no GPU call.

The files are written here in the reference's format (a key, any of the
separators "\\t :=>_", a value; '#' comment lines; keys in any case), for the
workloads this library is measured on, and must decode to exactly the
parameter blocks those workloads build directly."""
import math

import pytest

from dustyfargoadsg_amd import _capi
from dustyfargoadsg_amd.disc import DfargoError, read_par, read_planet_config
from dustyfargoadsg_amd.workloads import energy_sg, gas_planet

MANDATORY = "Dt 0.314159265358979\nNtot 100\nNinterm 10\nOutputDir out/\n"


def _write(tmp_path, text, name="run.par"):
    f = tmp_path / name
    f.write_text(text)
    return f


def _same(a, b, skip=("noise_seed",)):
    for name, _ in _capi.Params._fields_:
        if name in skip:
            continue
        assert getattr(a, name) == getattr(b, name), name


def test_energy_sg_file_decodes_to_the_workload(tmp_path):
    """The adiabatic self-gravitating disc, with the reference's mixed-case
    keys and separators; absent keys take src/var.c's defaults.  The key
    is the first whitespace-delimited word, so a separator is only skipped
    after a blank ("Nrad = 512", "NSEC : 1536")."""
    want, _ = energy_sg(nrad=512, nsec=1536)
    text = MANDATORY + f"""### Gas parameters
Sigma0          {want.sigma0!r}
SigmaSlope      1.0
AspectRatio     0.05
FlaringIndex    0.0
AlphaViscosity  1e-3
SelfGravity     Yes
SgThicknessSmoothing 0.3
AddNoise        Yes
EnergyEquation  Yes
AdiabaticIndex  1.4
BetaCooling     Yes
BetaCoolingTime 10.0
# the grid
Nrad = 512
NSEC : 1536
Rmin  0.4
rmax  4.0
RadialSpacing   Logarithmic
InnerBoundary   WALL
"""
    p, run = read_par(_write(tmp_path, text))
    _same(p, want)
    assert run.ntot == 100 and run.ninterm == 10 and run.output_dir == b"out/"
    assert run.dt == 0.314159265358979
    assert run.planet_config == b"Systems/SolarSystem.cfg"  # the default


def test_planet_disc_files(tmp_path):
    """The locally isothermal planet disc and its one-planet .cfg file: the
    planet's mass read with %f (float), FeelDisk/FeelOthers No."""
    want, pl = gas_planet(nrad=128, nsec=384)
    text = MANDATORY + f"""Sigma0 {want.sigma0!r}
SigmaSlope 0.5
AspectRatio 0.05
Viscosity 1e-5
ThicknessSmoothing 0.6
ForcedCircular YES
AddNoise YES
AdiabaticIndex 1.4
Nrad 128
Nsec 384
Rmin 0.4
Rmax 2.5
RadialSpacing L
Transport FARGO
PlanetConfig in/jupiter.cfg
"""
    p, run = read_par(_write(tmp_path, text))
    _same(p, want)
    assert run.planet_config == b"in/jupiter.cfg"
    cfg = _write(tmp_path, "# Planet  Dist  Mass  Accretion  FeelDisk  FeelOthers\n"
                 "Jupiter   1.0   0.001   0.0   NO   NO\n", "jupiter.cfg")
    sys_ = read_planet_config(cfg, p)
    assert sys_.nb == 1
    assert sys_.mass[0] == float(_capi.ctypes.c_float(0.001).value)
    assert sys_.x[0] == 1.0 and sys_.feel_disk[0] == 0 and sys_.feel_others[0] == 0
    # MASSTAPER (default 1e-7) below 1e-3: the velocity carries the mass
    assert sys_.vy[0] == math.sqrt(1.0 + sys_.mass[0])


def test_decoding_rules(tmp_path):
    """FRAME G, SELFGRAVITY Z, DampToAxi No (damping to the initial
    profiles), INNERBOUNDARY NONREFLECTING, TRANSPORT STANDARD, and an
    energy equation with ADIABATICINDEX 1 falling back to isothermal."""
    base = MANDATORY + "Sigma0 1e-3\nSigmaSlope 0.5\nAspectRatio 0.05\nNrad 32\nNsec 96\n" \
                       "Rmin 0.4\nRmax 2.5\nRadialSpacing L\n"
    p, _ = read_par(_write(tmp_path, base + "Frame G\nSelfGravity Z\nWaveKillingZones Yes\n"
                           "WKZRmin 0.5\nWKZRmax 2.2\nDampToAxi No\nInnerBoundary NONREFLECTING\n"
                           "Transport STANDARD\nEnergyEquation Yes\nAdiabaticIndex 1\n"))
    assert p.corotating == _capi.FRAME_GUIDING_CENTER
    assert p.self_gravity == 1 and p.sg_zero_mode == 1
    assert p.wave_killing_zones == 1 and p.damp_to == _capi.DAMP_TO_INI
    assert p.boundary == _capi.BC_NONREFLECTING and p.fast_transport == 0
    assert p.energy_equation == 0 and p.adiabatic_index == 1.0
    assert p.thickness_smoothing == 0.3 and p.density_floor == 1e-9  # defaults


@pytest.mark.parametrize("extra,code", [
    ("ThermalCooling Yes\n", _capi.ERR_UNSUPPORTED),
    ("RadiativeDiffusion Implicit\n", _capi.ERR_UNSUPPORTED),
    ("InnerBoundary Accreting\n", _capi.ERR_UNSUPPORTED),
    ("Viscosity 1e-5\nAlphaViscosity 1e-3\n", _capi.ERR_ARG),
])
def test_refusals(tmp_path, extra, code):
    base = MANDATORY + "Sigma0 1e-3\nSigmaSlope 0.5\nAspectRatio 0.05\nNrad 32\nNsec 96\n" \
                       "Rmin 0.4\nRmax 2.5\nAdiabaticIndex 1.4\n"
    with pytest.raises(DfargoError) as e:
        read_par(_write(tmp_path, base + extra))
    assert e.value.code == code


def test_dust_fluid_switches(tmp_path):
    """DUSTFLUID YES: the gas reader accepts the file, the dust-fluid reader
    decodes its switches (DUSTFEELSG Z, SFTAPPROX N, DUSTFEEDBACK Y,
    INNERBOUNDARYDUST OPEN) and refuses RESTARTWITHNEWDUST YES."""
    from dustyfargoadsg_amd.disc import read_dust_fluid_par
    base = MANDATORY + "Sigma0 1e-3\nSigmaSlope 0.5\nAspectRatio 0.05\nNrad 32\nNsec 96\n" \
                       "Rmin 0.4\nRmax 2.5\nAdiabaticIndex 1.4\n"
    text = base + ("DustFluid Yes\nSizePart 3e-2\nDustFeelSG Z\nSFTApprox N\nDustFeedback Yes\n"
                   "DViscosity 1e-6\nDustDiffusion Yes\nDustDiffInn 0.6\nDAspectRatio 0.01\n")
    read_par(_write(tmp_path, text))
    f, on = read_dust_fluid_par(_write(tmp_path, text))
    assert on and f.sizepart == 3e-2 and f.dust_feel_sg == _capi.DUST_SG_ZERO_MODE
    assert (f.sfta, f.dust_feedback, f.dust_diffusion, f.dust_feel_disk) == (0, 1, 1, 1)
    assert (f.dviscosity, f.dust_diff_inn, f.daspect_ratio, f.rhopart) == (1e-6, 0.6, 0.01, 1.0)
    assert f.open_inner_dust == 0
    f, _ = read_dust_fluid_par(_write(tmp_path, text + "InnerBoundaryDust OPEN\n"))
    assert f.open_inner_dust == 1
    with pytest.raises(DfargoError) as e:
        read_dust_fluid_par(_write(tmp_path, text + "RestartWithNewDust Yes\n"))
    assert e.value.code == _capi.ERR_UNSUPPORTED


def test_missing_mandatory_and_binary_planet(tmp_path):
    with pytest.raises(DfargoError) as e:
        read_par(_write(tmp_path, "Sigma0 1e-3\nNrad 32\n"))
    assert e.value.code == _capi.ERR_ARG and "mandatory" in str(e.value)
    p = _capi.default_params(nrad=32, nsec=96, rmin=0.4, rmax=2.5)
    cfg = _write(tmp_path, "A 1.0 1e-3 0.0 YES YES YES\n", "bin.cfg")
    with pytest.raises(DfargoError) as e:
        read_planet_config(cfg, p)
    assert e.value.code == _capi.ERR_UNSUPPORTED
