"""Parity of the particle kernels (k_dust_*) with the FP64 oracle, through
the C ABI.

Both sides start from the same gas state (InitEuler) and the same particles
(InitDustSystem on the host), run the start-up interpolation and then
AlgoGas calls of the same DT, and are compared.  Bars:
  * integer work is bit-exact: the cell of every particle (ring and column
    implied by its position), the number of sub-steps per call (which the
    feedback stopping-time limit and the fixed particle step set);
  * FP64: particle positions and velocities, Stokes numbers <= 1e-11
    relative; gas fields <= 1e-11 as in the gas-only tests; feedback
    grids and the dust density <= 1e-10 of their maximum;
  * FP32 field storage (particles stay FP64): positions <= 1e-6, particle
    velocities <= 1e-5 of max |v_theta|, Stokes numbers <= 1e-3 (they go
    through the Mach number of the dust-gas velocity lag).
The device sums each cell's deposit bucket by bucket (a fixed,
deterministic order) where the oracle's sequential loop goes by particle
index; that, the device's libm (log, pow) against glibc's and the gas
fields' reduction order are the only sources of difference.
"""
import math

import numpy as np
import pytest

from dustyfargoadsg_amd import _capi
from dustyfargoadsg_amd.disc import Disc, init_dust_system
from dustyfargoadsg_amd.workloads import gas_planet, energy_sg, dust_particles
from tests.oracle_lib import Oracle

pytestmark = pytest.mark.gpu

DT = 2 * math.pi / 20

INTERPS = {"ngp": _capi.INTERP_NGP, "cic": _capi.INTERP_CIC, "tsc": _capi.INTERP_TSC}


def cells(o, r, th):
    """(ring, column) of each particle from its position, -1 outside."""
    radii, _, _, az = o.tables()
    ns = len(az)
    half = 0.5 * 2 * math.pi / ns
    inf0, sup = az[0] - half, az[-1] + half
    ip = np.searchsorted(radii, r, side="right") - 1
    inside = (r >= radii[0]) & (r < radii[-1])
    jp = np.floor(ns * (th - inf0) / (sup - inf0)).astype(np.int64)
    jp[jp == ns] = 0
    return np.where(inside, ip, -1), np.where(inside, jp, -1)


def setup(factory, nr, ns, nbpart, precision=_capi.FP64, **dust_over):
    p, pl = factory(nrad=nr, nsec=ns)
    print(f"case {factory.__name__} {nr}x{ns} np={nbpart} prec={precision} {dust_over}", flush=True)
    kw = dict(sizeminpart=0.1, sizemaxpart=1.0, sizepartslope=0.0, rhopart=2.7, rmindust=0.6,
              rmaxdust=2.2, interpolation=_capi.INTERP_CIC, dust_feedback=1, seed=5)
    kw.update(dust_over)
    d = dust_particles(p, nbpart, **kw)
    o = Oracle(p)
    o.initial_state(pl)
    g = Disc(p, precision)
    g.initial_state(pl)
    axi = None
    if p.self_gravity:
        axi = g.compute_selfgravity()[0][: nr * ns].reshape(nr, ns).mean(axis=1)
    parts = init_dust_system(p, d, axi)
    o.dust_attach(d, parts, DT)
    g.attach_dust(d, parts, DT)
    return p, d, o, g


def relmax(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def compare_particles(o, g, tol_pos, tol_vel, tol_st=None):
    so, sg = o.dust_state(), g.dust_state()
    co, cg = cells(o, so["r"], so["th"]), cells(o, sg["r"], sg["th"])
    assert np.array_equal(co[0], cg[0]) and np.array_equal(co[1], cg[1])
    vscale = np.abs(so["vth"]).max()
    assert relmax(sg["r"], so["r"]) <= tol_pos
    assert np.abs(sg["th"] - so["th"]).max() <= tol_pos * 2 * math.pi
    assert np.abs(sg["vr"] - so["vr"]).max() <= tol_vel * vscale
    assert np.abs(sg["vth"] - so["vth"]).max() <= tol_vel * vscale
    inside = co[0] >= 0
    tol_st = max(tol_vel, 1e-11) if tol_st is None else tol_st
    assert relmax(sg["stokesnb"][inside], so["stokesnb"][inside]) <= tol_st
    return so, sg


def compare_grids(o, g, p, names, tol):
    n = p.nrad * p.nsec
    for f in names:
        a, b = g.field(f)[:n], o.field(f)[:n]
        assert relmax(a, b) <= tol, (f, relmax(a, b))


@pytest.mark.parametrize("interp", list(INTERPS))
def test_startup_interpolation(interp):
    """The start-up interpolation: Stokes numbers, feedback grids and the
    dust density match the oracle's."""
    p, d, o, g = setup(gas_planet, 32, 96, 3000, interpolation=INTERPS[interp])
    compare_particles(o, g, 0.0, 1e-12)
    compare_grids(o, g, p, ("dust_density", "fb_accr", "fb_acct"), 1e-12)


@pytest.mark.parametrize("interp", list(INTERPS))
def test_dust_substeps_fp64(interp):
    """Two AlgoGas calls with particles feeling the disc, the planet and the
    indirect term, and the gas feeling their feedback."""
    p, d, o, g = setup(gas_planet, 32, 96, 3000, interpolation=INTERPS[interp])
    for c in range(2):
        n_o = o.algogas(DT)
        g.algogas(DT)
        assert g.status().last_call_substeps == n_o
    compare_particles(o, g, 1e-11, 1e-11)
    go, oo = g.state(), o.state()
    n = p.nrad * p.nsec
    for k in ("rho", "vt"):
        assert relmax(go[k][:n], oo[k][:n]) <= 1e-11, k
    assert np.abs(go["vr"][:n] - oo["vr"][:n]).max() <= 1e-11 * np.abs(oo["vt"][:n]).max()
    compare_grids(o, g, p, ("dust_density", "fb_accr", "fb_acct"), 1e-10)


def test_dust_energy_selfgravity_fp64():
    """Adiabatic self-gravitating disc: particles feel the SG field, the
    feedback heats the gas (FBedot), the interpolation reads the sound
    speed of the previous sub-step."""
    p, d, o, g = setup(energy_sg, 32, 96, 2000, interpolation=_capi.INTERP_TSC, rmindust=0.6,
                       rmaxdust=3.5)
    for c in range(2):
        n_o = o.algogas(DT)
        g.algogas(DT)
        assert g.status().last_call_substeps == n_o
    compare_particles(o, g, 1e-11, 1e-11)
    go, oo = g.state(), o.state()
    n = p.nrad * p.nsec
    for k in ("rho", "vt", "energy"):
        assert relmax(go[k][:n], oo[k][:n]) <= 1e-11, k
    compare_grids(o, g, p, ("dust_density", "fb_accr", "fb_acct", "fb_edot"), 1e-10)


def test_dust_fixed_timestep_and_small_grains():
    """A fixed particle time step (NODTCONSTRAINTBYPCS NO) sets dt; grains of
    1e-4..1 m take both the short-friction-time and the semi-implicit
    branches; no feedback."""
    p, d, o, g = setup(gas_planet, 32, 96, 3000, sizeminpart=1e-4, sizemaxpart=1.0, sizepartslope=1.0,
                       dust_feedback=0, no_dt_constraint_by_particles=0, dust_timestep=DT / 13)
    for c in range(2):
        n_o = o.algogas(DT)
        g.algogas(DT)
        assert g.status().last_call_substeps == n_o
    st, so = g.status(), o.status()
    assert st.mindt == so.mindt == st.maxdt == d.dust_timestep
    compare_particles(o, g, 1e-11, 1e-11)


def test_dust_mass_taper():
    """The super-particle mass follows the planets' mass taper."""
    p2, pl = gas_planet(nrad=32, nsec=96, mass_taper=5.0)
    print("case mass taper 32x96 np=2000", flush=True)
    d = dust_particles(p2, 2000, sizeminpart=0.1, sizemaxpart=1.0, rhopart=2.7, rmindust=0.6, rmaxdust=2.2,
                       interpolation=_capi.INTERP_CIC, dust_feedback=1, seed=5)
    o2, g2 = Oracle(p2), Disc(p2)
    o2.initial_state(pl)
    g2.initial_state(pl)
    parts = init_dust_system(p2, d)
    o2.dust_attach(d, parts, DT)
    g2.attach_dust(d, parts, DT)
    for c in range(2):
        n_o = o2.algogas(DT)
        g2.algogas(DT)
        assert g2.status().last_call_substeps == n_o
    compare_particles(o2, g2, 1e-11, 1e-11)
    compare_grids(o2, g2, p2, ("dust_density", "fb_accr", "fb_acct"), 1e-10)


def test_dust_fp32_fields():
    p, d, o, g = setup(gas_planet, 32, 96, 3000, precision=_capi.FP32, interpolation=_capi.INTERP_TSC)
    for c in range(2):
        o.algogas(DT)
        g.algogas(DT)
    # the Stokes number depends on the Mach number of the dust-gas velocity
    # lag, a difference ~1e-3 of v_K: the FP32 gas velocity (~1e-7 of v_K)
    # moves it by ~1e-4
    compare_particles(o, g, 1e-6, 1e-5, tol_st=1e-3)


def test_dust_graph_replay_equals_eager():
    """The captured sub-step and the eagerly launched one give identical
    particles and fields (the deposit order does not depend on timing)."""
    p, d, o, g = setup(gas_planet, 32, 96, 3000, interpolation=_capi.INTERP_TSC)
    p2, d2, o2, g2 = setup(gas_planet, 32, 96, 3000, interpolation=_capi.INTERP_TSC)
    g2.set_profiling(True)
    for c in range(2):
        g.algogas(DT)
        g2.algogas(DT)
    a, b = g.dust_state(), g2.dust_state()
    for k in ("r", "th", "vr", "vth", "stokesnb"):
        assert np.array_equal(a[k], b[k]), k
    for f in ("dust_density", "fb_accr", "fb_acct"):
        assert np.array_equal(g.field(f), g2.field(f)), f
    assert np.array_equal(g.state()["rho"], g2.state()["rho"])
    assert "dust_deposit" in g2.kernel_times()


@pytest.mark.parametrize("interp", ["ngp", "cic"])
def test_deposit_conserves_mass_at_full_size(interp):
    """At 512x1536 with 10^6 particles: the dust density integrates to the
    mass of the particles inside the grid (NGP and CIC weights sum to one)."""
    p, pl = gas_planet(nrad=512, nsec=1536)
    print(f"case full-size deposit {interp}", flush=True)
    d = dust_particles(p, 1_000_000, sizeminpart=0.1, sizemaxpart=1.0, interpolation=INTERPS[interp],
                       dust_feedback=1, rmindust=0.5, rmaxdust=2.3, seed=9)
    g = Disc(p)
    g.initial_state(pl)
    parts = init_dust_system(p, d)
    g.attach_dust(d, parts, DT)
    g.algogas(DT)
    s = g.dust_state()
    radii = np.zeros(p.nrad + 1)
    import ctypes
    lib = _capi.load()
    rm, surf, az = np.zeros(p.nrad), np.zeros(p.nrad), np.zeros(p.nsec)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    lib.dfargo_fill_polar_1d_arrays(ctypes.byref(p), dp(radii), dp(rm), dp(surf), dp(az))
    inside = (s["r"] >= radii[0]) & (s["r"] < radii[-1])
    dd = g.field("dust_density")[: p.nrad * p.nsec].reshape(p.nrad, p.nsec)
    total = (dd * surf[:, None]).sum()
    assert abs(total / (inside.sum() * d.particle_mass) - 1.0) < 1e-10


@pytest.mark.parametrize("prec", [_capi.FP64, _capi.FP32])
def test_gas_dust_workload_reduced(prec):
    """The gas-plus-dust workload itself (gas_dust + gas_dust_particles: alpha
    1e-3, planet q = 1e-3, grains 1 um - 1 cm with dn/ds ~ s^-3.5, 2.7 g/cc,
    CIC, feedback) at 64x192 with 2*10^4 particles, against the oracle over
    two calls of DT = 2e-4: its grains reach St ~ 1e-4, where the minimum
    stopping time (src/SourceEuler.c:1500-1506) sets the sub-step
    (~16 per call) and the short-friction-time update
    (src/DustUpdate.c:323-360) moves them.  Cells are bit-exact in both
    precisions; FP64 bars as the other particle tests; FP32 storage: the
    FP32 particle bars of test_dust_fp32_fields."""
    from dustyfargoadsg_amd.workloads import gas_dust, gas_dust_particles
    p, pl = gas_dust(nrad=64, nsec=192)
    d = gas_dust_particles(p, 20000)
    print(f"case gas_dust 64x192 np=20000 prec={prec}", flush=True)
    o, g = Oracle(p), Disc(p, prec)
    o.initial_state(pl)
    g.initial_state(pl)
    parts = init_dust_system(p, d)
    o.dust_attach(d, parts, DT)
    g.attach_dust(d, parts, DT)
    for c in range(2):
        n_o = o.algogas(2e-4)
        g.algogas(2e-4)
        assert g.status().last_call_substeps == n_o
        assert abs(g.status().mindt - o.status().mindt) <= 1e-9 * o.status().mindt
    if prec == _capi.FP64:
        so, sg = compare_particles(o, g, 1e-11, 1e-11)
        go, oo = g.state(), o.state()
        n = p.nrad * p.nsec
        for k in ("rho", "vt"):
            assert relmax(go[k][:n], oo[k][:n]) <= 1e-11, k
        compare_grids(o, g, p, ("dust_density", "fb_accr", "fb_acct"), 1e-10)
    else:
        so, sg = compare_particles(o, g, 1e-6, 1e-5, tol_st=1e-3)
    assert so["stokesnb"].min() < 1e-3


def test_gas_dust_full_size_fp32_properties():
    """The full workload: 1024x3072 FP32 gas, 10^6 particles (CIC,
    feedback), 20 sub-steps.  No oracle at this size; instead: every
    particle and field stays finite, the dust density integrates to the
    mass of the particles inside the grid (CIC weights sum to one), the
    Stokes numbers span the grains' range, and the sub-step follows the
    minimum stopping time."""
    import ctypes
    from dustyfargoadsg_amd.workloads import gas_dust, gas_dust_particles
    p, pl = gas_dust()
    d = gas_dust_particles(p)
    print("case gas_dust 1024x3072 np=1e6 FP32, 20 sub-steps", flush=True)
    g = Disc(p, _capi.FP32)
    g.initial_state(pl)
    parts = init_dust_system(p, d)
    g.attach_dust(d, parts, DT)
    g.run_substeps(DT, 20)
    st = g.status()
    assert st.last_call_substeps == 20
    s = g.dust_state()
    for k in ("r", "th", "vr", "vth", "stokesnb"):
        assert np.all(np.isfinite(s[k])), k
    gs = g.state()
    n = p.nrad * p.nsec
    for k in ("rho", "vr", "vt"):
        assert np.all(np.isfinite(gs[k][:n])), k
    lib = _capi.load()
    radii, rm, surf, az = np.zeros(p.nrad + 1), np.zeros(p.nrad), np.zeros(p.nrad), np.zeros(p.nsec)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    lib.dfargo_fill_polar_1d_arrays(ctypes.byref(p), dp(radii), dp(rm), dp(surf), dp(az))
    inside = (s["r"] >= radii[0]) & (s["r"] < radii[-1])
    dd = g.field("dust_density")[:n].reshape(p.nrad, p.nsec)
    total = (dd * surf[:, None]).sum()
    # float storage of the deposit: each cell rounded to ~6e-8
    assert abs(total / (inside.sum() * d.particle_mass) - 1.0) < 1e-6
    stk = s["stokesnb"][inside]
    assert stk.min() < 1e-3 and stk.max() > 5e-3
    assert 0.0 < st.maxdt < 1e-3


def _planet_disc_fixed_frame(nrad, nsec):
    return gas_planet(nrad=nrad, nsec=nsec, omega_frame=1.0)


def _planet_disc_corotating(nrad, nsec):
    return gas_planet(nrad=nrad, nsec=nsec, omega_frame=1.0, corotating=1)


@pytest.mark.parametrize("factory", [_planet_disc_fixed_frame, _planet_disc_corotating])
def test_dust_rotating_frame_fp64(factory):
    """Particles in a rotating frame: a fixed OMEGAFRAME, and FRAME C (the
    frame follows the planet, so OmegaFrame changes every sub-step).  The
    interpolation reads the gas in the fixed frame with the rate before the
    frame update and leaves the round trip's rounding in the gas vtheta
    (src/Dsys.c:316,1082), the particles feel the planet before RotatePsys
    and turn with RotateDsys by the new rate (src/SourceEuler.c:504-530,
    src/Dsys.c:1153-1169)."""
    p, d, o, g = setup(factory, 32, 96, 3000, interpolation=_capi.INTERP_TSC)
    for c in range(2):
        n_o = o.algogas(DT)
        g.algogas(DT)
        assert g.status().last_call_substeps == n_o
    compare_particles(o, g, 1e-11, 1e-11)
    go, oo = g.state(), o.state()
    n = p.nrad * p.nsec
    for k in ("rho", "vt"):
        assert relmax(go[k][:n], oo[k][:n]) <= 1e-11, k
    assert np.abs(go["vr"][:n] - oo["vr"][:n]).max() <= 1e-11 * np.abs(oo["vt"][:n]).max()
    assert abs(g.status().omega_frame - o.status().omega_frame) <= 1e-12
    compare_grids(o, g, p, ("dust_density", "fb_accr", "fb_acct"), 1e-10)
