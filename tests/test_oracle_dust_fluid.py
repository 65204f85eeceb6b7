"""The oracle's dust fluid (DUSTFLUID YES: src/SourceEuler.c SubStep1/2 and
Diffd, src/TransportEuler.c:550-793 Transportd, src/Pframeforce.c:644-874
the start-up, src/SideEuler.c the boundaries), pinned by what the physics
and the reference's own structure fix:

* the Stokes number in the Epstein limit, St = pi rho_int s / (2 Sigma),
  with the units of src/LowTasks.c:204-212 restated here;
* SubStep1's Stokes number at zero relative velocity equals
  InitDustStokesNumber's (two code paths, one drag law);
* the short-friction-time dust velocity of a sub-step against the start-up
  terminal velocity -2 eta v_K / (St + 1/St) (Weidenschilling 1977,
  Takeuchi & Lin 2002) for small St;
* a dust fluid that is a copy of the gas (same sound speed, density,
  velocities and viscosity, no drag) stays a copy through a sub-step: the
  kicks, artificial viscosity, viscous stresses, the transport with its own
  FARGO shift, the diffusion of a uniform concentration (none), the floor
  and the wall all act on it as on the gas, bit for bit, except the
  sub-Keplerian vtheta of the edge rings, which takes the dust's own
  initial profile;
* turbulent diffusion (Diffd) conserves the dust mass away from the edges
  and lowers a concentration peak;
* with feedback at a dust-to-gas ratio of one the start-up velocities are
  the Nakagawa, Sekiya & Hayashi (1986) equilibrium, whose net radial mass
  flux vanishes; at other ratios the dust's start-up uses eps = gas / dust,
  the reference's swapped call (src/SourceEuler.c:268 against
  src/Pframeforce.c:644);
* the CFL limit of the drag term, dt = CFLSECURITY St / Omega, sets the
  sub-step count for small grains without the short-friction-time
  approximation;
* the open dust edge (OpenBoundaryd) leaves the zero-gradient density and
  copied vr of src/SideEuler.c:411-466, and the wave-killing zones then
  leave the fluid alone.
"""
import ctypes
import math

import numpy as np
import pytest

from dustyfargoadsg_amd import _capi
from dustyfargoadsg_amd.disc import init_dust_fluid_density
from dustyfargoadsg_amd.workloads import gas_planet
from tests.oracle_lib import Oracle

UNIT_MASS, UNIT_LENGTH = 1.9891e30, 1.49598e11  # kg, m (Sun, AU)


def _oracle(p, pl, fp, drho=None):
    o = Oracle(p)
    if drho is None:
        drho = init_dust_fluid_density(p, fp)
    o.attach_dust_fluid(fp, drho)
    st = o.initial_state(pl)
    return o, st


def _no_planet():
    return dict(dist=[], mass=[], acc=[], feel_disk=[], feel_others=[])


def test_host_dust_fluid_density_and_checks(oracle_built):
    p, _ = gas_planet(nrad=16, nsec=32)
    fp = _capi.dust_fluid_default_params(dust_to_gas_density_ratio=0.03)
    drho = init_dust_fluid_density(p, fp).reshape(p.nrad + 1, p.nsec)
    rmed = Oracle(p).tables()[1]
    sig = p.sigma0 * rmed ** -p.sigma_slope * 0.03
    assert np.allclose(drho[: p.nrad], sig[:, None], rtol=1e-15, atol=0)
    assert np.all(drho[p.nrad] == 0.0)
    lib = _capi.load()
    assert lib.dfargo_check_dust_fluid_params(ctypes.byref(p), ctypes.byref(fp), _capi.FP64) == _capi.OK
    assert lib.dfargo_check_dust_fluid_params(ctypes.byref(p), ctypes.byref(fp), _capi.FP32) == _capi.ERR_UNSUPPORTED
    bad = _capi.dust_fluid_default_params(dust_feel_disk=0)
    assert lib.dfargo_check_dust_fluid_params(ctypes.byref(p), ctypes.byref(bad), _capi.FP64) == _capi.ERR_UNSUPPORTED
    bad = _capi.dust_fluid_default_params(sizepart=0.0)
    assert lib.dfargo_check_dust_fluid_params(ctypes.byref(p), ctypes.byref(bad), _capi.FP64) == _capi.ERR_ARG


def test_stokes_epstein_limit(oracle_built):
    """Micron grains in a thin disc: Kn >> 1, the drag coefficient tends to
    one and St = pi rho_int s / (2 Sigma) (Epstein) to 1 / (3 Kn)."""
    p, pl = gas_planet(nrad=16, nsec=32, add_noise=0)
    fp = _capi.dust_fluid_default_params(sizepart=1e-6, rhopart=3.0)
    o, st = _oracle(p, pl, fp)
    St = o.dust_fluid()["stokes"][: p.nrad * p.nsec]
    rho_int = 3.0 * 1000.0 * UNIT_LENGTH ** 3 / UNIT_MASS
    s = 1e-6 / UNIT_LENGTH
    epstein = math.pi * rho_int * s / (2.0 * st["rho"][: p.nrad * p.nsec])
    rel = np.abs(St / epstein - 1.0)
    assert rel.max() < 1e-4
    assert rel.min() > 0.0  # the finite-Kn correction is there


def test_substep_stokes_equals_startup_at_zero_relative_velocity(oracle_built):
    """With the dust moving with the gas (Mach 0: f_D = k_D = 1), SubStep1's
    drag law gives InitDustStokesNumber's Stokes numbers."""
    p, pl = gas_planet(nrad=16, nsec=32)
    fp = _capi.dust_fluid_default_params(sizepart=1e-2)
    o, st = _oracle(p, pl, fp)
    St0 = o.dust_fluid()["stokes"].copy()
    o.set_dust_fluid(dvr=st["vr"], dvt=st["vt"])
    o.algogas(1e-4)  # one sub-step; St is formed from the state it starts with
    St1 = o.dust_fluid()["stokes"]
    n = p.nrad * p.nsec
    assert np.abs(St1[:n] / St0[:n] - 1.0).max() < 1e-14


def test_sfta_velocity_matches_terminal_drift(oracle_built):
    """Short-friction-time approximation, small grains: after a sub-step the
    dust's radial velocity relative to the gas is ts (1/Sigma) dP/dr, the
    start-up terminal drift -2 eta v_K / (St + 1/St) to within the
    discretisation (face against cell centre) of a smooth disc: about 3 %
    on this grid.  Both are taken relative to the gas velocity the sub-step
    starts from."""
    p, _ = gas_planet(nrad=64, nsec=32, add_noise=0, viscosity=0.0)
    fp = _capi.dust_fluid_default_params(sizepart=1e-4)
    o, st = _oracle(p, _no_planet(), fp)
    d0 = o.dust_fluid()
    assert d0["stokes"].max() < 1e-2
    o.algogas(1e-5)
    d1 = o.dust_fluid()
    n, ns = p.nrad * p.nsec, p.nsec
    rel = (d1["dvr"][:n] - st["vr"][:n]).reshape(p.nrad, ns)[8:-8]
    want = (d0["dvr"][:n] - st["vr"][:n]).reshape(p.nrad, ns)[8:-8]
    assert np.all(np.sign(rel) == np.sign(want))
    assert np.abs(rel / want - 1.0).max() < 5e-2


def _copy_case(boundary, wkz=0):
    p, pl = gas_planet(nrad=24, nsec=48, boundary=boundary, wave_killing_zones=wkz, wkz_rmin=0.6,
                       wkz_rmax=2.0, damp_to=_capi.DAMP_TO_AXI)
    fp = _capi.dust_fluid_default_params(sfta=0, dust_feel_disk=0, dust_to_gas_density_ratio=1.0,
                                         daspect_ratio=p.aspect_ratio, dflaring_index=p.flaring_index,
                                         dviscosity=p.viscosity, dust_diffusion=1, dust_feel_sg=0)
    return p, pl, fp


@pytest.mark.parametrize("boundary,wkz", [(_capi.BC_WALL, 0), (_capi.BC_WALL, 1), (_capi.BC_OPEN, 0)])
def test_dust_copy_of_the_gas_stays_a_copy(boundary, wkz, oracle_built):
    """One sub-step; the wave-killing zones (DampToAxi) damp both fluids
    towards their own ring means, which are the same.  The open edge is the
    gas's alone (the dust has OpenBoundaryd only with INNERBOUNDARYDUST
    OPEN): there the comparison leaves out the rings it rewrites."""
    p, pl, fp = _copy_case(boundary, wkz)
    o, st = _oracle(p, pl, fp)
    o.set_dust_fluid(drho=st["rho"], dvr=st["vr"], dvt=st["vt"])
    assert o.algogas(1e-4) == 1
    g, d = o.state(), o.dust_fluid()
    n, ns = p.nrad * p.nsec, p.nsec
    lo, hi = (0, n) if boundary == _capi.BC_WALL else (3 * ns, n - ns)
    assert np.array_equal(d["drho"][lo:hi], g["rho"][lo:hi])
    assert np.array_equal(d["dvr"][lo:hi], g["vr"][lo:hi])
    assert np.array_equal(d["dvt"][max(lo, ns): n - ns], g["vt"][max(lo, ns): n - ns])
    # the edge rings' vtheta: the dust's own start-up profile
    assert not np.array_equal(d["dvt"][:ns], g["vt"][:ns])
    if boundary == _capi.BC_OPEN:
        assert not np.array_equal(d["drho"][:ns], g["rho"][:ns])


def test_diffusion_conserves_mass_and_smooths(oracle_built):
    """A concentration bump in the middle of the disc, no drag and no
    motion of either fluid except the diffusion: the dust mass is conserved
    (the bump is far from the edges) and the peak concentration drops."""
    p, _ = gas_planet(nrad=32, nsec=32, add_noise=0, viscosity=1e-3, boundary=_capi.BC_OPEN)
    base = _capi.dust_fluid_default_params(sfta=0, dust_feel_disk=0, dust_feel_sg=0)
    runs = {}
    for diff in (0, 1):
        fp = _capi.dust_fluid_default_params(sfta=0, dust_feel_disk=0, dust_feel_sg=0, dust_diffusion=diff)
        o = Oracle(p)
        drho = init_dust_fluid_density(p, base)
        r = o.tables()[1]
        bump = 1.0 + 0.5 * np.exp(-((r - 1.2) / 0.1) ** 2)
        drho[: p.nrad * p.nsec] *= np.repeat(bump, p.nsec)
        o.attach_dust_fluid(fp, drho)
        st = o.initial_state(_no_planet())
        # both fluids at rest in the rotating sense: the same vtheta, no vr
        o.set_dust_fluid(drho=drho, dvr=np.zeros_like(drho), dvt=st["vt"])
        o.algogas(1e-3)
        runs[diff] = (o.dust_fluid()["drho"], o.state()["rho"], o.tables()[2])
    (d0, g0, surf), (d1, g1, _) = runs[0], runs[1]
    n = p.nrad * p.nsec
    m0 = (d0[:n].reshape(p.nrad, -1).sum(axis=1) * surf).sum()
    m1 = (d1[:n].reshape(p.nrad, -1).sum(axis=1) * surf).sum()
    assert abs(m1 - m0) <= 1e-13 * m0
    c0 = (d0[:n] / g0[:n]).max()
    c1 = (d1[:n] / g1[:n]).max()
    assert c1 < c0
    assert not np.array_equal(d0, d1)


def test_feedback_startup_is_the_nsh_equilibrium(oracle_built):
    """DUSTFEEDBACK start-up: at a dust-to-gas ratio of one the gas and dust
    radial mass fluxes cancel (Nakagawa, Sekiya & Hayashi 1986); at 0.1 the
    gas takes the equilibrium with eps = 0.1 and the dust the one with
    eps = 10 (the reference's swapped densities)."""
    p, _ = gas_planet(nrad=16, nsec=32, add_noise=0, viscosity=0.0)
    fp = _capi.dust_fluid_default_params(dust_feedback=1, dust_to_gas_density_ratio=1.0, sizepart=3e-2)
    o, st = _oracle(p, _no_planet(), fp)
    d = o.dust_fluid()
    n = p.nrad * p.nsec
    flux = st["rho"][:n] * st["vr"][:n] + d["drho"][:n] * d["dvr"][:n]
    assert np.abs(flux).max() <= 1e-14 * np.abs(st["rho"][:n] * st["vr"][:n]).max()
    assert np.abs(st["vr"][:n]).max() > 0.0
    fp = _capi.dust_fluid_default_params(dust_feedback=1, dust_to_gas_density_ratio=0.1, sizepart=3e-2)
    o, st = _oracle(p, _no_planet(), fp)
    d = o.dust_fluid()
    r = np.repeat(o.tables()[1], p.nsec)
    vk = np.sqrt(1.0 / r)
    eta = -0.5 * p.aspect_ratio ** 2 * (-1.0 - p.sigma_slope)
    St = d["stokes"][:n]
    for eps, got in ((0.1, st["vr"][:n]), (10.0, None)):
        if got is not None:
            want = 2.0 * eps * eta * vk * St / (St * St + (1.0 + eps) ** 2)
            assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()
    want_d = -2.0 * eta * vk / (St + (1.0 + 10.0) ** 2 / St)
    assert np.abs(d["dvr"][:n] - want_d).max() <= 1e-12 * np.abs(want_d).max()


def test_drag_time_step_limit(oracle_built):
    """Without the short-friction-time approximation the drag term
    1 / (St + 1e-10) Omega bounds the sub-step: for small grains
    dt = CFLSECURITY min(St / Omega) and the count is ceil(DT / dt)."""
    p, _ = gas_planet(nrad=16, nsec=32, add_noise=0)
    fp = _capi.dust_fluid_default_params(sfta=0, sizepart=1e-5)
    o, st = _oracle(p, _no_planet(), fp)
    St = o.dust_fluid()["stokes"][: p.nrad * p.nsec].reshape(p.nrad, p.nsec)
    r = o.tables()[1]
    omega = np.sqrt(1.0 / r) / r
    dt = 0.5 * (St[1:] / omega[1:, None]).min()  # the CFL loop starts at ring 1
    DT = 37.3 * dt
    assert o.algogas(DT) == math.ceil(DT / dt)


def test_open_dust_edge(oracle_built):
    """INNERBOUNDARYDUST OPEN beside a wall gas edge and damping zones:
    after a call, rings 0 and 1 hold ring 2's dust density, rings 1 and 2
    ring 3's vr, ring nrad-1 ring nrad-2's density and vr; the zones damp
    the gas but not the dust (a dust copy of the gas loses the identity in
    the zones' interior rings only through the damping)."""
    p, pl = gas_planet(nrad=24, nsec=48, wave_killing_zones=1, wkz_rmin=0.6, wkz_rmax=2.0,
                       damp_to=_capi.DAMP_TO_AXI)
    fp = _capi.dust_fluid_default_params(sfta=0, sizepart=1e-2, open_inner_dust=1)
    o, st = _oracle(p, pl, fp)
    o.algogas(2 * math.pi / 40)
    d = o.dust_fluid()
    ns, nr = p.nsec, p.nrad
    R = lambda a, i: a[i * ns:(i + 1) * ns]
    assert np.array_equal(R(d["drho"], 0), R(d["drho"], 2)) and np.array_equal(R(d["drho"], 1), R(d["drho"], 2))
    assert np.array_equal(R(d["dvr"], 1), R(d["dvr"], 3)) and np.array_equal(R(d["dvr"], 2), R(d["dvr"], 3))
    assert np.array_equal(R(d["drho"], nr - 1), R(d["drho"], nr - 2))
    assert np.array_equal(R(d["dvr"], nr - 1), R(d["dvr"], nr - 2))
    # no damping of the dust: one sub-step of an undragged fluid with the
    # zones on and off (the gas differs, the dust does not)
    fp = _capi.dust_fluid_default_params(sfta=0, dust_feel_disk=0, open_inner_dust=1)
    o, _ = _oracle(p, pl, fp)
    assert o.algogas(1e-4) == 1
    d = o.dust_fluid()
    q, _ = gas_planet(nrad=24, nsec=48, wave_killing_zones=0)
    o2, _ = _oracle(q, pl, fp)
    assert o2.algogas(1e-4) == 1
    d2 = o2.dust_fluid()
    zone = slice(3 * ns, 4 * ns)  # Rmed < 0.6, away from the copied edge rings
    assert np.array_equal(d["drho"][zone], d2["drho"][zone])
    assert not np.array_equal(o.state()["rho"][zone], o2.state()["rho"][zone])
