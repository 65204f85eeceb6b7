"""Restart of the GPU path from the reference's output files (src/Output.c
SendOutput / WritePlanetSystemFile, read back as src/Init.c:18-170 and
src/Output.c:110-176 do, by dustyfargoadsg_amd/output.py): a GPU run writes
output 1, a fresh context restarts from those files and takes the next
call; the result is compared with an oracle run that was never interrupted.
Fields travel as raw FP64 (exact); the planets and PhysicalTime pass through
the planet files' %.14g text, so the bar is 1e-11 of each field's maximum
(the FP64 parity bars otherwise being 1e-12)."""
import math

import numpy as np
import pytest

from dustyfargoadsg_amd import output
from dustyfargoadsg_amd.disc import Disc
from dustyfargoadsg_amd.workloads import gas_planet, energy_sg
from tests.oracle_lib import Oracle

pytestmark = pytest.mark.gpu

DT = 2 * math.pi / 20


def _restart(tmp_path, p, pl, number):
    r = Disc(p)
    rst = r.initial_state(pl)
    rmed = Oracle(p).tables()[1]
    back = output.read_restart(tmp_path, number, p, rmed)
    planets = output.restart_planetary_system(tmp_path, number, rst["planets"])
    lost, t, w = output.restart_scalars(tmp_path, number, p.corotating)
    st = dict(rho=back["rho"], vr=back["vr"], vt=back["vt"], planets=planets)
    if p.energy_equation:
        st["energy"] = back["energy"]
    r.set_state(st)
    r.set_time(t, 0.0, 0.0)
    return r


def test_gpu_restart_continues_the_run(tmp_path):
    p, pl = gas_planet(nrad=32, nsec=96, forced_circular=0)
    pl = dict(pl, feel_disk=[1], feel_others=[1])
    print("case restart gas_planet 32x96", flush=True)
    d = Disc(p)
    d.initial_state(pl)
    d.algogas(DT)
    st, s = d.state(), d.status()
    output.write_outputs(tmp_path, 1, st, p.nrad, p.nsec)
    output.write_planet_files(tmp_path, 1, st["planets"], s)
    r = _restart(tmp_path, p, pl, 1)
    r.algogas(DT)
    o = Oracle(p)
    o.initial_state(pl)
    o.algogas(DT)
    n_o = o.algogas(DT)
    assert r.status().last_call_substeps == n_o
    a, b = r.state(), o.state()
    n = p.nrad * p.nsec
    for k in ("rho", "vt"):
        assert np.abs(a[k][:n] - b[k][:n]).max() <= 1e-11 * np.abs(b[k][:n]).max(), k
    assert np.abs(a["vr"][:n] - b["vr"][:n]).max() <= 1e-11 * np.abs(b["vt"][:n]).max()
    assert abs(a["planets"].x[0] - b["planets"].x[0]) <= 1e-11


def test_gpu_restart_energy_from_temperature(tmp_path):
    """With the energy equation the restart rebuilds e = Sigma T / (gamma - 1)
    from the Temperature file the GPU run wrote (src/Init.c:155-166)."""
    p, pl = energy_sg(nrad=32, nsec=96)
    print("case restart energy_sg 32x96", flush=True)
    d = Disc(p)
    d.initial_state(pl)
    d.algogas(DT)
    st, s = d.state(), d.status()
    output.write_outputs(tmp_path, 1, st, p.nrad, p.nsec, temperature=d.field("temperature"))

    class One:  # GetfromPlanetFile(..., 9, 0) reads planet0.dat in every restart
        nb, x, y, vx, vy, mass = 1, [1.0], [0.0], [0.0], [1.0], [0.0]

    output.write_planet_files(tmp_path, 1, One, s)
    r = _restart(tmp_path, p, pl, 1)
    n = p.nrad * p.nsec
    e_back = r.state()["energy"][:n]
    assert np.abs(e_back - st["energy"][:n]).max() <= 1e-14 * np.abs(st["energy"][:n]).max()
    r.algogas(DT)
    o = Oracle(p)
    o.initial_state(pl)
    o.algogas(DT)
    n_o = o.algogas(DT)
    assert r.status().last_call_substeps == n_o
    a, b = r.state(), o.state()
    for k in ("rho", "vt", "energy"):
        assert np.abs(a[k][:n] - b[k][:n]).max() <= 1e-11 * np.abs(b[k][:n]).max(), k
