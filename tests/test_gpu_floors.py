"""The density and energy floors and the crash scan on the GPU against the
FP64 oracle: AddFloorDensity / AddFloorEnergy (src/SourceEuler.c:2506-2541,
called after Transport at src/SourceEuler.c:663-666) and DetectCrash
(src/SourceEuler.c:51-69) with the flag overwrite of
src/SourceEuler.c:545-548 (the energy scan replaces the density scan).

Each case plants states that make a branch bind, then takes one AlgoGas
call short enough (DT = 1e-4) for exactly one sub-step, so the crash flags
of that sub-step are the ones reported:
  * a ring at 1e-12 (below DENSITYFLOOR = 1e-9): the inflow of one sub-step
    leaves it below the floor, which lifts it;
  * a cell with a slightly negative density: DetectCrash(Rho) flags it (the
    inflow of the step then makes it positive again);
  * with the energy equation, a ring at 1e-6 of EnergyMed: below the energy
    floor 1e-3 EnergyMed after the step; and a ring at 1e-12 density with
    its energy scaled alike (the same sound speed: a low density at the
    full energy would drive the velocities, and with them the CFL step,
    without bound).
Bars as in the other FP64 parity tests: 1e-12 of each field's maximum, flags
and sub-step counts identical.
"""
import numpy as np
import pytest

from dustyfargoadsg_amd.disc import Disc
from dustyfargoadsg_amd.workloads import gas_planet, energy_sg
from tests.oracle_lib import Oracle

pytestmark = pytest.mark.gpu

DT_ONE = 1e-4


def _pair_from_state(p, pl, plant):
    o = Oracle(p)
    st = o.initial_state(pl)
    plant(st, p)
    o.set_state(st)
    d = Disc(p)
    d.initial_state(pl)
    d.set_state(st)
    return o, d, st


def _check(o, d, p, keys):
    assert o.algogas(DT_ONE) == 1
    d.algogas(DT_ONE)
    sg, so = d.status(), o.status()
    assert sg.last_call_substeps == 1
    assert (sg.crashed_density, sg.crashed_energy, sg.crashed) == \
        (so.crashed_density, so.crashed_energy, so.crashed)
    g, r = d.state(), o.state()
    n = p.nrad * p.nsec
    for k in keys:
        err = np.abs(g[k][:n] - r[k][:n]).max() / np.abs(r[k][:n]).max()
        assert err <= 1e-12, (k, err)
    return g, r, sg


def test_density_floor_and_crash_isothermal():
    p, pl = gas_planet(nrad=32, nsec=96)
    print("case floors isothermal 32x96", flush=True)
    ns = p.nsec

    def plant(st, p):
        st["rho"][10 * ns:11 * ns] = 1e-12
        st["rho"][20 * ns + 7] = -1e-3 * st["rho"][20 * ns + 6]

    o, d, st = _pair_from_state(p, pl, plant)
    g, r, sg = _check(o, d, p, ("rho", "vt", "vr"))
    assert sg.crashed_density == 1
    # the isothermal energy grid stays zero, so the overwriting scan clears
    # the flag (src/SourceEuler.c:548)
    assert sg.crashed == 0
    # the floor bound: the planted ring and cell sit exactly at the floor
    assert np.all(r["rho"][10 * ns:11 * ns] == p.density_floor)
    assert np.array_equal(g["rho"][10 * ns:11 * ns], r["rho"][10 * ns:11 * ns])


def test_energy_floor_adiabatic():
    p, pl = energy_sg(nrad=32, nsec=96)
    print("case floors energy 32x96", flush=True)
    ns = p.nsec

    emed = {}

    def plant(st, p):
        # the initial energy is EnergyMed, ring by ring (InitGasEnergy)
        emed[12] = st["energy"][12 * ns]
        st["energy"][12 * ns:13 * ns] *= 1e-6
        f = 1e-12 / st["rho"][5 * ns:6 * ns]
        st["rho"][5 * ns:6 * ns] *= f
        st["energy"][5 * ns:6 * ns] *= f

    o, d, st = _pair_from_state(p, pl, plant)
    g, r, sg = _check(o, d, p, ("rho", "vt", "vr", "energy"))
    assert sg.crashed == 0 and sg.crashed_density == 0
    assert np.abs(r["energy"][12 * ns:13 * ns] / (1e-3 * emed[12]) - 1.0).max() <= 1e-15
    assert np.array_equal(g["energy"][12 * ns:13 * ns], r["energy"][12 * ns:13 * ns])
    assert np.all(g["rho"][5 * ns:6 * ns] == p.density_floor)
