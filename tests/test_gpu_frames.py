"""Corotating frames (FRAME C, G, B and COROTATEWITHOUTERPLANET;
src/SourceEuler.c:476,516-524, src/Psys.c:233-350, src/SideEuler.c:1133-1148),
the wave-killing zones (EvanescentBoundary, src/SideEuler.c:796-950) and the
non-reflecting boundary (src/SideEuler.c:569-794) of the HIP step against
the FP64 oracle, through the C ABI.  Same bars as the
other parity tests: FP64 fields within 1e-12 of the oracle after two AlgoGas
calls, sub-step counts and FARGO shifts exact; FP32 storage within 2e-5
(Sigma) and 2e-6 (velocities).
"""
import math

import numpy as np
import pytest

from dustyfargoadsg_amd import _capi
from dustyfargoadsg_amd.workloads import gas_planet, energy_sg
from tests.test_gpu_parity import DT, compare, pair, rel

pytestmark = pytest.mark.gpu


def _frame_case(name):
    m = 1e-3
    if name == "corotating":
        # starts in the inertial frame: the first sub-step's CorrectVtheta
        # takes the whole OmegaFrame Rmed off v_theta; the planet feels the
        # disc, so the frame's rate moves every sub-step
        p, pl = gas_planet(nrad=32, nsec=96, forced_circular=0, corotating=_capi.FRAME_COROTATING)
        pl = dict(pl, feel_disk=[1], feel_others=[1])
    elif name == "guiding":
        p, pl = gas_planet(nrad=32, nsec=96, forced_circular=0, eccentricity=0.1,
                           corotating=_capi.FRAME_GUIDING_CENTER, omega_frame=math.sqrt(1 + m))
        pl = dict(pl, feel_disk=[1], feel_others=[1])
    else:
        outer = name == "outer"
        frame = _capi.FRAME_COROTATING if outer else _capi.FRAME_BINARY_CENTER
        p, _ = gas_planet(nrad=32, nsec=96, forced_circular=0, corotating=frame,
                          corotate_with_outer_planet=int(outer), omega_frame=math.sqrt(1 + m))
        pl = dict(dist=[1.0, 1.6], mass=[m, 3e-4], acc=[0.0, 0.0], feel_disk=[1, 0],
                  feel_others=[1, 1])
    return p, pl


@pytest.mark.parametrize("name", ["corotating", "guiding", "binary", "outer"])
def test_frames_fp64(name):
    p, pl = _frame_case(name)
    o, so, d, sg = pair(p, pl)
    for _ in range(2):
        n_o = o.algogas(DT)
        d.algogas(DT)
        assert d.status().last_call_substeps == n_o
        assert np.array_equal(d.fargo_shift()[0], o.fargo_shift()[0])
    compare(o, d, p, 1e-12, 1e-12, 1e-12)
    w_g, w_o = d.status().omega_frame, o.status().omega_frame
    assert abs(w_g - w_o) <= 1e-12 * abs(w_o)
    go, oo = d.state()["planets"], o.state()["planets"]
    for k in ("x", "y", "vx", "vy"):
        for i in range(oo.nb):
            assert abs(getattr(go, k)[i] - getattr(oo, k)[i]) <= 1e-13, (k, i)


def _wkz_case(name):
    zones = dict(wave_killing_zones=1, wkz_rmin=0.55, wkz_rmax=2.2)
    if name == "axi":
        return gas_planet(nrad=32, nsec=96, **zones)
    if name == "ini":
        return gas_planet(nrad=32, nsec=96, damp_to=_capi.DAMP_TO_INI, **zones)
    if name == "zero_vrad":
        return gas_planet(nrad=32, nsec=96, zero_vrad_bc=1, **zones)
    if name == "viscous":
        # DAMPTOVISCOUS: the 1-D viscous disc beside the grid, the boundary
        # rule of BOUNDARYGRID1D Z
        return gas_planet(nrad=32, nsec=96, damp_to=_capi.DAMP_TO_VISCOUS, bc1d=1, **zones)
    if name == "viscous_alpha":
        return energy_sg(nrad=32, nsec=96, damp_to=_capi.DAMP_TO_VISCOUS, wave_killing_zones=1,
                         wkz_rmin=0.55, wkz_rmax=3.5)
    # the energy field is damped too; the K = 6 ring width of the large discs
    return energy_sg(nrad=16, nsec=1536, wave_killing_zones=1, wkz_rmin=0.55, wkz_rmax=3.5)


@pytest.mark.parametrize("name", ["axi", "ini", "zero_vrad", "energy_sg", "viscous", "viscous_alpha"])
def test_wave_killing_zones_fp64(name):
    p, pl = _wkz_case(name)
    o, so, d, sg = pair(p, pl)
    for _ in range(2):
        n_o = o.algogas(DT)
        d.algogas(DT)
        assert d.status().last_call_substeps == n_o
        assert np.array_equal(d.fargo_shift()[0], o.fargo_shift()[0])
    compare(o, d, p, 1e-12, 1e-12, 1e-12, 1e-12)


def test_frame_and_zones_fp32():
    p, pl = gas_planet(nrad=128, nsec=384, forced_circular=0, corotating=_capi.FRAME_COROTATING,
                       omega_frame=math.sqrt(1 + 1e-3), wave_killing_zones=1, wkz_rmin=0.55,
                       wkz_rmax=2.2)
    pl = dict(pl, feel_disk=[1], feel_others=[1])
    o, so, d, sg = pair(p, pl, _capi.FP32)
    for _ in range(2):
        n_o = o.algogas(DT)
        d.algogas(DT)
        assert d.status().last_call_substeps == n_o
    compare(o, d, p, 2e-5, 2e-6, 2e-6)
    assert abs(d.status().omega_frame - o.status().omega_frame) <= 1e-9


def test_frame_needs_its_planets():
    p, _ = gas_planet(nrad=32, nsec=96, corotating=_capi.FRAME_BINARY_CENTER)
    pl = dict(dist=[1.0], mass=[1e-3], acc=[0.0], feel_disk=[0], feel_others=[0])
    from dustyfargoadsg_amd.disc import Disc, DfargoError
    d = Disc(p, _capi.FP64)
    with pytest.raises(DfargoError):
        d.initial_state(pl)


@pytest.mark.parametrize("nr,ns,calls,zones", [(32, 96, 2, 0), (16, 1536, 1, 0), (32, 96, 2, 1)])
def test_nonreflecting_fp64(nr, ns, calls, zones):
    """NonReflectingBoundary (src/SideEuler.c:569-794), barotropic, with a
    planet that feels the disc (its pattern speed moves), alone and under
    the wave-killing zones."""
    over = dict(wave_killing_zones=1, wkz_rmin=0.55, wkz_rmax=2.2) if zones else {}
    p, pl = gas_planet(nrad=nr, nsec=ns, forced_circular=0, boundary=_capi.BC_NONREFLECTING, **over)
    pl = dict(pl, feel_disk=[1], feel_others=[1])
    o, so, d, sg = pair(p, pl)
    for _ in range(calls):
        n_o = o.algogas(DT)
        d.algogas(DT)
        assert d.status().last_call_substeps == n_o
        assert np.array_equal(d.fargo_shift()[0], o.fargo_shift()[0])
    compare(o, d, p, 1e-12, 1e-12, 1e-12)


def test_nonreflecting_fp32():
    p, pl = gas_planet(nrad=128, nsec=384, boundary=_capi.BC_NONREFLECTING)
    o, so, d, sg = pair(p, pl, _capi.FP32)
    for _ in range(2):
        n_o = o.algogas(DT)
        d.algogas(DT)
        assert d.status().last_call_substeps == n_o
    compare(o, d, p, 2e-5, 2e-6, 2e-6)


@pytest.mark.parametrize("nr,ns,frame", [(32, 96, 0), (128, 384, 0), (64, 192, 1)])
def test_accretion_fp64(nr, ns, frame):
    """AccreteOntoPlanets (src/Planet.c:11-126) in the control kernel: the
    density near the planet, the planet's mass and velocity, and through the
    corrected force sums everything after, against the oracle."""
    p, pl = gas_planet(nrad=nr, nsec=ns, forced_circular=0,
                       corotating=_capi.FRAME_COROTATING if frame else _capi.FRAME_FIXED,
                       omega_frame=math.sqrt(1 + 1e-3) if frame else 0.0)
    pl = dict(pl, acc=[2.0], feel_disk=[1], feel_others=[1])
    o, so, d, sg = pair(p, pl)
    for _ in range(2):
        n_o = o.algogas(DT)
        d.algogas(DT)
        assert d.status().last_call_substeps == n_o
        assert np.array_equal(d.fargo_shift()[0], o.fargo_shift()[0])
    compare(o, d, p, 1e-12, 1e-12, 1e-12)
    go, oo = d.state()["planets"], o.state()["planets"]
    assert oo.mass[0] > 1e-3
    assert abs(go.mass[0] - oo.mass[0]) <= 1e-12 * oo.mass[0]
    for k in ("x", "y", "vx", "vy"):
        assert abs(getattr(go, k)[0] - getattr(oo, k)[0]) <= 1e-13, k


def test_accretion_fp32_and_limits():
    p, pl = gas_planet(nrad=128, nsec=384, forced_circular=0)
    pl = dict(pl, acc=[2.0], feel_disk=[1], feel_others=[1])
    o, so, d, sg = pair(p, pl, _capi.FP32)
    for _ in range(2):
        n_o = o.algogas(DT)
        d.algogas(DT)
        assert d.status().last_call_substeps == n_o
    compare(o, d, p, 2e-5, 2e-6, 2e-6)
    assert abs(d.state()["planets"].mass[0] / o.state()["planets"].mass[0] - 1) <= 1e-9
    from dustyfargoadsg_amd.disc import Disc, DfargoError
    p, pl = gas_planet(nrad=32, nsec=96, exclude_hill=1)
    with pytest.raises(DfargoError):
        Disc(p, _capi.FP64).initial_state(dict(pl, acc=[1.0]))


@pytest.mark.parametrize("nr,ns,calls", [(32, 96, 2), (40, 120, 1), (16, 1536, 1)])
def test_sg_zero_mode_fp64(nr, ns, calls):
    """SELFGRAVITY Z (src/sgzero.c): the axisymmetric self-gravity, also on a
    grid the two-dimensional transforms do not take (Nrad = 40)."""
    p, pl = energy_sg(nrad=nr, nsec=ns, sg_zero_mode=1)
    o, so, d, sg = pair(p, pl)
    assert rel(sg["vt"], so["vt"], nr, ns) <= 1e-14
    for _ in range(calls):
        n_o = o.algogas(DT)
        d.algogas(DT)
        assert d.status().last_call_substeps == n_o
        assert np.array_equal(d.fargo_shift()[0], o.fargo_shift()[0])
    compare(o, d, p, 1e-12, 1e-12, 1e-12, 1e-12)
    assert rel(d.field("sg_accr"), o.field("sg_accr"), nr, ns) <= 1e-11


def test_sg_zero_mode_planet_and_fp32():
    """A planet on the start-up orbit corrected by the axisymmetric field
    (init_planetarysys_withSG with GLOBAL_AxiSGAccr = SG_Accr), FP64; then
    FP32 storage."""
    p, pl = gas_planet(nrad=32, nsec=96, self_gravity=1, sg_zero_mode=1, sg_thickness_smoothing=0.3)
    o, so, d, sg = pair(p, pl)
    vy_g, vy_o = d.state()["planets"].vy[0], o.state()["planets"].vy[0]
    assert vy_o != so["planets"].vy[0]  # the start-up correction acted
    assert abs(vy_g - vy_o) <= 1e-15
    for _ in range(2):
        n_o = o.algogas(DT)
        d.algogas(DT)
        assert d.status().last_call_substeps == n_o
    compare(o, d, p, 1e-12, 1e-12, 1e-12)
    p, pl = energy_sg(nrad=64, nsec=192, sg_zero_mode=1)
    o, so, d, sg = pair(p, pl, _capi.FP32)
    for _ in range(2):
        n_o = o.algogas(DT)
        d.algogas(DT)
        assert d.status().last_call_substeps == n_o
    compare(o, d, p, 2e-5, 2e-6, 2e-6, 2e-5)


@pytest.mark.parametrize("nr,ns,calls", [(32, 96, 2), (16, 1536, 1)])
def test_nonreflecting_energy_fp64(nr, ns, calls):
    """The adiabatic non-reflecting boundary (src/SideEuler.c:654-699 and
    the outer edge with the cell's own sound speed), the sound speed formed
    inside the boundary call and read by the next CFL condition, with
    self-gravity and a planet that feels the disc."""
    p, _ = energy_sg(nrad=nr, nsec=ns, boundary=_capi.BC_NONREFLECTING)
    pl = dict(dist=[1.5], mass=[1e-4], acc=[0.0], feel_disk=[1], feel_others=[1])
    o, so, d, sg = pair(p, pl)
    for _ in range(calls):
        n_o = o.algogas(DT)
        d.algogas(DT)
        assert d.status().last_call_substeps == n_o
        assert np.array_equal(d.fargo_shift()[0], o.fargo_shift()[0])
    compare(o, d, p, 1e-12, 1e-12, 1e-12, 1e-12)


@pytest.mark.parametrize("bc1d,prec", [(2, _capi.FP64), (0, _capi.FP64), (1, _capi.FP32)])
def test_damp_to_viscous_1d_disc(bc1d, prec):
    """DAMPTOVISCOUS (src/SideEuler.c:1444-1565): the 1-D viscous disc beside
    the grid, read back directly and compared with the oracle's after two
    calls, for BOUNDARY1DGRID D (2), T (0) and, in FP32 storage, Z (1); the
    damped 2-D fields with it.  The 1-D disc is FP64 at either precision."""
    p, pl = gas_planet(nrad=32, nsec=96, damp_to=_capi.DAMP_TO_VISCOUS, bc1d=bc1d,
                       wave_killing_zones=1, wkz_rmin=0.55, wkz_rmax=2.2)
    o, so, d, sg = pair(p, pl, prec)
    s0, v0 = d.visc1d()
    so0, vo0 = o.viscous_1d(0.0)
    assert len(s0) == p.nrad1d and np.array_equal(s0, so0) and np.array_equal(v0, vo0)
    for _ in range(2):
        n_o = o.algogas(DT)
        d.algogas(DT)
        assert d.status().last_call_substeps == n_o
    s1, v1 = d.visc1d()
    so1, vo1 = o.viscous_1d(0.0)
    assert rel(s1, so1, 1, len(so1)) <= 1e-12
    assert np.abs(v1 - vo1).max() <= 1e-12 * np.abs(vo1).max()
    assert rel(s1, s0, 1, len(s0)) > 0.0  # the 1-D disc evolved
    if prec == _capi.FP64:
        compare(o, d, p, 1e-12, 1e-12, 1e-12)
    else:
        compare(o, d, p, 2e-5, 2e-6, 2e-6)
    # a second InitEuler on the same context restarts the 1-D disc from its
    # initial profiles
    d.initial_state(pl)
    s2, v2 = d.visc1d()
    assert np.array_equal(s2, s0) and np.array_equal(v2, v0)


def test_damp_to_viscous_without_zones_accepts_any_1d_grid():
    """Without the wave-killing zones the 1-D disc only evolves (nothing
    interpolates it), so a 1-D grid that does not enclose the 2-D one is
    accepted, as in the reference."""
    p, pl = gas_planet(nrad=32, nsec=96, damp_to=_capi.DAMP_TO_VISCOUS, bc1d=1, rmin1d=0.5, rmax1d=2.0)
    o, so, d, sg = pair(p, pl)
    n_o = o.algogas(DT)
    d.algogas(DT)
    assert d.status().last_call_substeps == n_o
    compare(o, d, p, 1e-12, 1e-12, 1e-12)


def test_accretion_fp32_ten_orbits():
    """An accreting planet in FP32 storage against FP64 over 10 orbits: the
    accretion rewrites the Roche-lobe cells every sub-step through the
    stochastically rounded store.  The planet feels the disc, so its orbit
    integrates the FP32 torque error: measured on MI355X, ring-mean Sigma
    1.4e-5, Sigma L2 7.1e-5, planet position 3.9e-5, planet mass 4.1e-6
    (the non-accreting planet on a forced circular orbit stays within 1e-11);
    the bars leave a factor of about two."""
    import bench
    p, pl = gas_planet(forced_circular=0)
    pl = dict(pl, acc=[2.0], feel_disk=[1], feel_others=[1])
    print("case accretion 128x384, 10 orbits, FP32 vs FP64", flush=True)
    r = bench.precision_error(p, pl, 10, control=False)["fp32_vs_fp64"]
    print(r, flush=True)
    assert r["substeps"][0] == r["substeps"][1]
    assert r["sigma_ring_mean_linf"] <= 3e-5
    assert r["sigma_l2"] <= 1.5e-4
    assert r["planet_position"] <= 1e-4
    assert r["planet_mass"] <= 1e-5
