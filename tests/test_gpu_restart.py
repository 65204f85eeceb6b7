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


def test_gpu_restart_visc1d(tmp_path):
    """DAMPTOVISCOUS with the wave-killing zones: a GPU run writes output 2
    with the 1-D disc's ascii_rad profiles (Write1DViscProfiles,
    src/SideEuler.c:1564-1585); a fresh context restarts from the files as
    src/SideEuler.c:1489-1502 does (Rmed1D and Sigma1D from the 6-digit
    text, Vrad1D at its t = 0 profile) and takes two more calls.  The oracle
    restarted from the same files must agree: fields to the planet files'
    text precision (1e-11), the 1-D profiles to 1e-12."""
    from dustyfargoadsg_amd import _capi
    p, pl = gas_planet(nrad=32, nsec=96, damp_to=_capi.DAMP_TO_VISCOUS, bc1d=1,
                       wave_killing_zones=1, wkz_rmin=0.55, wkz_rmax=2.2)
    print("case restart DAMPTOVISCOUS gas_planet 32x96", flush=True)
    d = Disc(p)
    d.initial_state(pl)
    _, _, vr0 = d.visc1d(rmed=True)
    for _ in range(2):
        d.algogas(DT)
    st, s = d.state(), d.status()
    output.write_outputs(tmp_path, 2, st, p.nrad, p.nsec)
    output.write_planet_files(tmp_path, 2, st["planets"], s)
    output.write_visc1d_profiles(tmp_path, 2, *st["visc1d"])
    rm, sg = output.read_visc1d_restart(tmp_path, 2, p.nrad1d)
    r = _restart(tmp_path, p, pl, 2)
    r.set_visc1d(rm, sg)
    a_r, a_s, a_v = r.visc1d(rmed=True)
    assert np.array_equal(a_r, rm) and np.array_equal(a_s, sg) and np.array_equal(a_v, vr0)
    o = Oracle(p)
    rst = o.initial_state(pl)
    back = output.read_restart(tmp_path, 2, p, o.tables()[1])
    for k in ("rho", "vr", "vt"):
        rst[k][:] = back[k]
    rst["planets"] = output.restart_planetary_system(tmp_path, 2, rst["planets"])
    o.set_state(rst)
    o.set_time(output.restart_scalars(tmp_path, 2, p.corotating)[1], 0.0, 0.0)
    o.set_visc1d(rm, sg)
    for _ in range(2):
        n_o = o.algogas(DT)
        r.algogas(DT)
        assert r.status().last_call_substeps == n_o
    a, b = r.state(), o.state()
    n = p.nrad * p.nsec
    for k in ("rho", "vt"):
        assert np.abs(a[k][:n] - b[k][:n]).max() <= 1e-11 * np.abs(b[k][:n]).max(), k
    assert np.abs(a["vr"][:n] - b["vr"][:n]).max() <= 1e-11 * np.abs(b["vt"][:n]).max()
    g_r, g_s, g_v = a["visc1d"]
    o_r, o_s, o_v = o.visc1d()
    assert np.array_equal(g_r, o_r)
    assert np.abs(g_s - o_s).max() <= 1e-12 * np.abs(o_s).max()
    assert np.abs(g_v - o_v).max() <= 1e-12 * np.abs(o_v).max()


def test_gpu_set_visc1d_rejects_bad_profiles():
    """dfargo_set_visc1d refuses a length other than NRAD1D, cell centres
    that do not increase or do not enclose the grid under the wave-killing
    zones, and a context without the 1-D disc; a refused call changes
    nothing."""
    from dustyfargoadsg_amd import _capi
    from dustyfargoadsg_amd.disc import DfargoError
    p, pl = gas_planet(nrad=32, nsec=96, damp_to=_capi.DAMP_TO_VISCOUS, bc1d=1,
                       wave_killing_zones=1, wkz_rmin=0.55, wkz_rmax=2.2)
    print("case set_visc1d argument checks", flush=True)
    d = Disc(p)
    d.initial_state(pl)
    rm, sg, vr = d.visc1d(rmed=True)
    with pytest.raises(DfargoError):
        d.set_visc1d(rm[:-1], sg[:-1])
    bad = rm.copy()
    bad[5] = bad[4]
    with pytest.raises(DfargoError):
        d.set_visc1d(bad)
    with pytest.raises(DfargoError):
        d.set_visc1d(rm * 50.0)  # Rmed1D[0] ~ 0.52, beyond the grid's inner Rmed ~ 0.41
    bad = sg.copy()
    bad[3] = np.nan
    with pytest.raises(DfargoError):
        d.set_visc1d(sigma=bad)
    a_r, a_s, a_v = d.visc1d(rmed=True)
    assert np.array_equal(a_r, rm) and np.array_equal(a_s, sg) and np.array_equal(a_v, vr)
    d.set_visc1d(vrad=2.0 * vr)
    assert np.array_equal(d.visc1d()[1], 2.0 * vr)
    q, ql = gas_planet(nrad=32, nsec=96)
    e = Disc(q)
    e.initial_state(ql)
    with pytest.raises(DfargoError):
        e.set_visc1d(rm, sg)
