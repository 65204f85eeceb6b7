"""Pins the oracle's per-DT diagnostics (GasTotalMass, GasMomentum,
GasTotalEnergy, UpdateLog's torques, FindOrbitalElements) by closed forms,
as the rest of the oracle is pinned.  Each test names the mistake it would
catch.
"""
import math

import numpy as np
import pytest

from dustyfargoadsg_amd._capi import default_params, Planets
from tests.oracle_lib import Oracle, lib, orbital_elements


def _uniform_disc(nr=128, ns=256, energy=False):
    p = default_params(nrad=nr, nsec=ns, rmin=0.4, rmax=2.5, log_grid=1, sigma0=1e-3,
                       sigma_slope=0.0, aspect_ratio=0.05, flaring_index=0.0,
                       energy_equation=1 if energy else 0, viscous_heating=0)
    o = Oracle(p)
    st = o.initial_state()
    return p, o, st


def test_budgets_closed_forms(oracle_built):
    """Uniform Sigma0, Keplerian v_theta = r^-1/2, v_r = 0, uniform energy e0,
    and the star's potential -G/r (filled by one AlgoGas call of 1e-9, whose
    boundary reset of rings 0 and Nrad-1 to the initial profile moves
    v_theta there by O(h^2) only):
      mass     Sigma0 pi (rmax^2 - rmin^2)             (exact: Surf tiles the annulus)
      momentum Sigma0 2 pi (2/5) (rmax^5/2 - rmin^5/2)  (a dropped Rmed factor fails)
      energy   kinetic Sigma0 pi (rmax - rmin), internal e0 pi (rmax^2 - rmin^2),
               potential Sigma0 2 pi (rmax - rmin)      (a sign of the potential term fails)
    to O(dr^2) for the centred sums."""
    p, o, st = _uniform_disc(energy=True)
    nr, ns = p.nrad, p.nsec
    radii, rmed, _, _ = o.tables()
    sig, e0 = 1e-3, 2e-6
    st["rho"][:] = sig
    st["vr"][:] = 0.0
    st["vt"][:] = np.repeat(np.append(rmed, rmed[-1]), ns) ** -0.5
    st["energy"][:] = e0
    o.set_state(st)
    o.algogas(1e-9)
    m, J, E, Jp = o.budgets()
    r0, r1 = radii[0], radii[-1]
    assert abs(m / (sig * math.pi * (r1**2 - r0**2)) - 1) < 1e-12
    assert abs(J / (sig * 2 * math.pi * 0.4 * (r1**2.5 - r0**2.5)) - 1) < 1e-3
    kin = sig * math.pi * (r1 - r0)
    inte = e0 * math.pi * (r1**2 - r0**2)
    pot = sig * 2 * math.pi * (r1 - r0)
    assert abs(E / (kin + inte + pot) - 1) < 1e-3
    assert Jp == 0.0


def test_budgets_isothermal_energy_grid_is_zero(oracle_built):
    """Without the energy equation the energy grid holds CreatePolarGrid's
    zeros: the total energy of a Keplerian uniform disc is kinetic plus
    potential only, 3 pi Sigma0 (rmax - rmin) (an internal term from the
    isothermal sound speed, ~1 % of it here, fails)."""
    p, o, st = _uniform_disc(energy=False)
    ns = p.nsec
    radii, rmed, _, _ = o.tables()
    st["rho"][:] = 1e-3
    st["vr"][:] = 0.0
    st["vt"][:] = np.repeat(np.append(rmed, rmed[-1]), ns) ** -0.5
    o.set_state(st)
    o.algogas(1e-9)
    _, _, E, _ = o.budgets()
    r0, r1 = radii[0], radii[-1]
    assert abs(E / (1e-3 * 3 * math.pi * (r1 - r0)) - 1) < 3e-4


def test_planet_momentum(oracle_built):
    """CheckMomentumConservation's planet term m r v_azimuthal, for a planet at
    (0, 2) moving with (-0.7, 0.1): v_azimuthal = 0.7 (a sign or a swapped
    sin/cos fails)."""
    p, o, st = _uniform_disc(nr=16, ns=32)
    pl = Planets()
    pl.nb = 1
    pl.x[0], pl.y[0], pl.vx[0], pl.vy[0], pl.mass[0] = 0.0, 2.0, -0.7, 0.1, 1e-3
    st["planets"] = pl
    o.set_state(st)
    assert abs(o.budgets()[3] - 1e-3 * 2.0 * 0.7) < 1e-17


def test_update_log_single_cell(oracle_built):
    """All the gas in one cell: the disc force on a planet at (a, 0) is
    G m_cell (xc - a, yc) / (d^2 + eps^2)^1.5 with eps = 0.3 h a (the
    thickness smoothing), in the inner or outer column as Rmed < a, and the
    Hill-excluded columns equal it for a cell far from the planet.  Torque
    a fy, power vy fy (a swapped column, a sign or the smoothing fails)."""
    p, o, st = _uniform_disc(nr=32, ns=64)
    nr, ns = p.nrad, p.nsec
    radii, rmed, surf, az = o.tables()
    a_pl, vy = 1.05, 0.97
    for i0, inner in ((10, True), (25, False)):
        j0 = 7
        st["rho"][:] = 0.0
        st["rho"][i0 * ns + j0] = 1e-2
        pl = Planets()
        pl.nb = 1
        pl.x[0], pl.y[0], pl.vx[0], pl.vy[0], pl.mass[0] = a_pl, 0.0, 0.0, vy, 1e-3
        pl.feel_disk[0] = 1
        st["planets"] = pl
        o.set_state(st)
        tq, ind = o.update_log()
        xc, yc = rmed[i0] * math.cos(az[j0]), rmed[i0] * math.sin(az[j0])
        eps = 0.3 * 0.05 * a_pl
        mc = surf[i0] * 1e-2
        d2 = (xc - a_pl) ** 2 + yc**2
        fy = mc * yc / (d2 + eps**2) ** 1.5
        col = 0 if inner else 1
        assert abs(tq[0, col] - a_pl * fy) <= 1e-13 * abs(a_pl * fy)
        assert tq[0, 1 - col] == 0.0
        assert tq[0, 2 + col] == tq[0, col]
        assert abs(tq[0, 4 + col] - vy * fy) <= 1e-13 * abs(vy * fy)
        assert ind[0] == 0.0


def test_update_log_hill_exclusion(oracle_built):
    """A cell inside half the Hill-cut radius (EXCLUDEHILLFACTOR times the
    Hill radius; ComputeForce forms the excluded columns whatever
    ExcludeHill says) contributes nothing to the excluded columns and fully
    to the plain ones (the Hill cut applied to the wrong column fails)."""
    p, o, st = _uniform_disc(nr=64, ns=256)
    nr, ns = p.nrad, p.nsec
    radii, rmed, surf, az = o.tables()
    i0 = int(np.argmin(abs(rmed - 1.0)))
    st["rho"][:] = 0.0
    st["rho"][i0 * ns + 3] = 1e-2
    pl = Planets()
    pl.nb = 1
    a_pl = rmed[i0]
    pl.x[0], pl.y[0], pl.vx[0], pl.vy[0], pl.mass[0] = a_pl, 0.0, 0.0, 1.0, 1e-2
    st["planets"] = pl
    o.set_state(st)
    tq, _ = o.update_log()
    rh = (1e-2 / 3) ** (1 / 3) * a_pl
    d = math.hypot(rmed[i0] * math.cos(az[3]) - a_pl, rmed[i0] * math.sin(az[3]))
    assert d < 0.5 * p.exclude_hill_factor * rh
    plain = tq[0, 0] + tq[0, 1]
    assert plain != 0.0
    assert tq[0, 2] == 0.0 and tq[0, 3] == 0.0


@pytest.mark.parametrize("e", [0.0, 0.3])
def test_orbital_elements_at_pericentre(oracle_built, e):
    """At pericentre r = a(1 - e) with v = sqrt(G m (1 + e) / (a (1 - e))):
    the elements return a and e, and the periastron and mean longitude at
    the position angle (a wrong vis-viva factor fails).  The anomalies there
    (acos of 1 up to rounding, or of 0/0 on a circle) are left to the next
    test."""
    a, m, phi = 1.7, 1.001, 0.4
    r = a * (1 - e)
    v = math.sqrt(m * (1 + e) / (a * (1 - e)))
    x, y = r * math.cos(phi), r * math.sin(phi)
    vx, vy = -v * math.sin(phi), v * math.cos(phi)
    el = orbital_elements(x, y, vx, vy, m)
    assert abs(el[0] - e) < 1e-12
    assert abs(el[1] - a) < 1e-12 * a
    assert abs(el[5] - phi) < 1e-14
    if e > 0:
        assert abs(el[4] - phi) < 1e-12


def test_orbital_elements_mean_anomaly_quarter(oracle_built):
    """On an e = 0.2 ellipse at eccentric anomaly E = pi/2 (position
    (a(cos E - e), b sin E) from the focus): M = E - e sin E, and the true
    anomaly satisfies tan(V/2) = sqrt((1+e)/(1-e)) tan(E/2)."""
    a, e, m = 1.3, 0.2, 1.0
    b = a * math.sqrt(1 - e * e)
    E = math.pi / 2
    x, y = a * (math.cos(E) - e), b * math.sin(E)
    n = math.sqrt(m / a**3)
    r = a * (1 - e * math.cos(E))
    vx, vy = -a * n * math.sin(E) / (1 - e * math.cos(E)), b * n * math.cos(E) / (1 - e * math.cos(E))
    el = orbital_elements(x, y, vx, vy, m)
    assert abs(el[0] - e) < 1e-12 and abs(el[1] - a) < 1e-12
    assert abs(el[2] - (E - e * math.sin(E))) < 1e-10
    V = 2 * math.atan(math.sqrt((1 + e) / (1 - e)) * math.tan(E / 2))
    assert abs(el[3] - V) < 1e-10
    assert abs(math.hypot(x, y) - r) < 1e-14
