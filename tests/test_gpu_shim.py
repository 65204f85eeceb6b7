"""The drop-in end to end: the small C host (shim/fargo_host.c) reads a .par
and a planet .cfg, replays the reference's start-up and DT loop
(src/main.c:50-397) through the shim's InitEuler / AlgoGas / ... and writes
the reference's output files; they must equal an oracle run of the same
parameters to the FP64 parity bar (1e-12 of each field's maximum, the
planet to the planet file's 14 digits).  AddNoise is off in the file: the
shim seeds the noise with time(NULL), as the reference does."""
import math
import os
import subprocess

import numpy as np
import pytest

from dustyfargoadsg_amd import build, output
from dustyfargoadsg_amd.disc import read_par
from tests.oracle_lib import Oracle

pytestmark = pytest.mark.gpu

PAR = """Dt {dt!r}
Ntot 4
Ninterm 2
OutputDir {out}/
PlanetConfig {cfg}
Sigma0 6.3662e-4
SigmaSlope 0.5
AspectRatio 0.05
Viscosity 1e-5
Nrad 32
Nsec 96
Rmin 0.4
Rmax 2.5
RadialSpacing Logarithmic
ThicknessSmoothing 0.6
AdiabaticIndex 1.4
{extra}
"""
# a mass a float holds exactly: the .cfg is read with %f (src/Psys.c:118)
CFG = "Jupiter 1.0 0.0009765625 0.0 {feel} {feel}\n"


@pytest.mark.parametrize("variant", ["planet_fixed_orbit", "feel_disk_energy"])
def test_fargo_host_equals_the_oracle(tmp_path, variant):
    assert os.access(build.SHIM_HOST, os.X_OK), "build() builds shim/fargo_host"
    dt = 2 * math.pi / 20
    if variant == "planet_fixed_orbit":
        extra, feel = "ForcedCircular Yes\n", "No"
    else:
        extra, feel = "EnergyEquation Yes\nAlphaViscosity 1e-3\nViscosity 0\n", "Yes"
    cfg = tmp_path / "planet.cfg"
    cfg.write_text(CFG.format(feel=feel))
    par = tmp_path / "run.par"
    par.write_text(PAR.format(dt=dt, out=tmp_path, cfg=cfg, extra=extra))
    print(f"case fargo_host {variant}", flush=True)
    r = subprocess.run([build.SHIM_HOST, str(par)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    dots = [len(line) for line in r.stdout.splitlines() if line and set(line) == {"."}]
    p, run = read_par(par)
    o = Oracle(p)
    o.initial_state(dict(dist=[1.0], mass=[0.0009765625], acc=[0.0], feel_disk=[feel == "Yes"],
                         feel_others=[feel == "Yes"]))
    n = p.nrad * p.nsec
    for call in range(1, 5):
        assert o.algogas(dt) == dots[call - 1], call
        if call % 2 == 0:
            st = o.state()
            for name, key in (("gasdens", "rho"), ("gasvtheta", "vt"), ("gasvrad", "vr")):
                got = output.read_disk_polar(tmp_path, name, call // 2, p.nrad, p.nsec)[:n]
                scale = np.abs(st["vt" if key == "vr" else key][:n]).max()
                assert np.abs(got - st[key][:n]).max() <= 1e-12 * scale, (call, name)
            x = output.get_from_planet_file(tmp_path, call // 2, 2, 0)
            assert abs(x - float("%.14g" % st["planets"].x[0])) <= 1e-13


@pytest.mark.parametrize("dust_extra", ["DustFluid Yes\nSizePart 1e-2\n",
                                        "DustFluid Yes\nSizePart 1e-2\nSFTApprox No\nDustFeedback Yes\n"
                                        "DustDiffusion Yes\nInnerBoundaryDust O\n"])
def test_fargo_host_dust_fluid(tmp_path, dust_extra):
    """The same host with DUSTFLUID YES: InitDustDensity, InitEuler and
    AlgoGas carry DRho, DVrad, DVtheta through the shim; the dust files
    equal the oracle's fluid to 1e-12."""
    from dustyfargoadsg_amd.disc import init_dust_fluid_density, read_dust_fluid_par
    assert os.access(build.SHIM_HOST, os.X_OK), "build() builds shim/fargo_host"
    dt = 2 * math.pi / 20
    cfg = tmp_path / "planet.cfg"
    cfg.write_text(CFG.format(feel="No"))
    par = tmp_path / "run.par"
    par.write_text(PAR.format(dt=dt, out=tmp_path, cfg=cfg, extra="ForcedCircular Yes\n" + dust_extra))
    print(f"case fargo_host dust fluid {dust_extra!r}", flush=True)
    r = subprocess.run([build.SHIM_HOST, str(par)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    p, run = read_par(par)
    fp, on = read_dust_fluid_par(par)
    assert on
    o = Oracle(p)
    o.attach_dust_fluid(fp, init_dust_fluid_density(p, fp))
    o.initial_state(dict(dist=[1.0], mass=[0.0009765625], acc=[0.0], feel_disk=[0], feel_others=[0]))
    n = p.nrad * p.nsec
    for call in range(1, 5):
        o.algogas(dt)
        if call % 2 == 0:
            st, d = o.state(), o.dust_fluid()
            for name, ref, scale in (("gasdens", st["rho"], st["rho"]), ("dustdens", d["drho"], d["drho"]),
                                     ("dustvrad", d["dvr"], d["dvt"]), ("dustvtheta", d["dvt"], d["dvt"])):
                got = output.read_disk_polar(tmp_path, name, call // 2, p.nrad, p.nsec)[:n]
                assert np.abs(got - ref[:n]).max() <= 1e-12 * np.abs(scale[:n]).max(), (call, name)
