"""The dust fluid (DUSTFLUID YES) on the GPU against the FP64 oracle,
through the C ABI: the start-up (InitDustStokesNumber, InitDustVelocities,
the gas's feedback equilibrium), then AlgoGas calls with the dust's
SubStep1/2 terms, viscosity, Transportd, Diffd, the boundaries and the
wave-killing zones.  Bars: sub-step counts equal; gas and dust fields
within 1e-12 of the oracle's maximum (vr on the vtheta scale, as for the
gas); the Stokes grid within 1e-11 relative: St depends on the relative
speed of the two fluids, a difference of nearly equal velocities whose
relative error is larger than theirs (2.3e-12 measured with self-gravity,
whose transforms sum in another order).  The dust fluid runs with FP64
storage."""
import math

import numpy as np
import pytest

from dustyfargoadsg_amd import _capi
from dustyfargoadsg_amd.disc import Disc, DfargoError, init_dust_fluid_density
from dustyfargoadsg_amd.workloads import gas_planet, energy_sg
from tests.oracle_lib import Oracle
from tests.test_gpu_parity import DT, compare

pytestmark = pytest.mark.gpu


def df_pair(p, pl, fp):
    print(f"case dust fluid {p.nrad}x{p.nsec} sfta={fp.sfta} feedback={fp.dust_feedback}", flush=True)
    drho = init_dust_fluid_density(p, fp)
    o = Oracle(p)
    o.attach_dust_fluid(fp, drho)
    so = o.initial_state(pl)
    d = Disc(p)
    d.attach_dust_fluid(fp, drho)
    sg = d.initial_state(pl)
    return o, so, d, sg


def compare_dust(o, d, p, tol):
    a, b = d.dust_fluid(), o.dust_fluid()
    n = p.nrad * p.nsec
    vscale = np.abs(b["dvt"][:n]).max()
    errs = {
        "drho": np.abs(a["drho"][:n] - b["drho"][:n]).max() / np.abs(b["drho"][:n]).max(),
        "dvr": np.abs(a["dvr"][:n] - b["dvr"][:n]).max() / vscale,
        "dvt": np.abs(a["dvt"][:n] - b["dvt"][:n]).max() / vscale,
        "stokes": (np.abs(a["stokes"][:n] - b["stokes"][:n]) / b["stokes"][:n]).max(),
    }
    for k, v in errs.items():
        assert v <= (10 * tol if k == "stokes" else tol), errs
    return errs


def run_pair(p, pl, fp, calls=2, tol=1e-12):
    o, so, d, sg = df_pair(p, pl, fp)
    for k in ("rho", "vr", "vt"):
        assert np.abs(sg[k] - so[k]).max() <= 1e-14 * np.abs(so[k]).max(), k
    compare_dust(o, d, p, 1e-14)
    for _ in range(calls):
        n_o = o.algogas(DT)
        d.algogas(DT)
        assert d.status().last_call_substeps == n_o
    compare(o, d, p, tol, tol, tol)
    errs = compare_dust(o, d, p, tol)
    print(errs, flush=True)
    return o, d


def test_sfta_planet_wall():
    """The defaults: short-friction-time approximation, mm grains."""
    p, pl = gas_planet(nrad=32, nsec=96)
    run_pair(p, pl, _capi.dust_fluid_default_params())


def test_full_momentum_equations_with_viscosity_and_diffusion():
    p, pl = gas_planet(nrad=32, nsec=96)
    fp = _capi.dust_fluid_default_params(sfta=0, sizepart=3e-2, dviscosity=1e-6, dust_diffusion=1)
    run_pair(p, pl, fp)


def test_no_drag_without_sfta():
    p, pl = gas_planet(nrad=32, nsec=96)
    run_pair(p, pl, _capi.dust_fluid_default_params(sfta=0, dust_feel_disk=0))


@pytest.mark.parametrize("sfta", [1, 0])
def test_feedback(sfta):
    """The dust's drag on the gas in k_source (src/SourceEuler.c:915-997) and
    the start-up equilibrium of both fluids."""
    p, pl = gas_planet(nrad=32, nsec=96)
    fp = _capi.dust_fluid_default_params(sfta=sfta, dust_feedback=1, dust_to_gas_density_ratio=0.1,
                                         sizepart=1e-2)
    run_pair(p, pl, fp)


def test_alpha_viscosities_open_edge():
    p, pl = gas_planet(nrad=32, nsec=96, viscosity=0.0, alpha_viscosity=1e-3, boundary=_capi.BC_OPEN)
    fp = _capi.dust_fluid_default_params(sfta=0, sizepart=1e-2, dalpha_viscosity=1e-4, dust_diffusion=1)
    run_pair(p, pl, fp)


@pytest.mark.parametrize("damp", [_capi.DAMP_TO_AXI, _capi.DAMP_TO_INI])
def test_wave_killing_zones(damp):
    p, pl = gas_planet(nrad=32, nsec=96, wave_killing_zones=1, wkz_rmin=0.55, wkz_rmax=2.2, damp_to=damp)
    run_pair(p, pl, _capi.dust_fluid_default_params(sfta=0, sizepart=1e-2))


def test_corotating_frame():
    p, pl = gas_planet(nrad=32, nsec=96, corotating=_capi.FRAME_COROTATING, forced_circular=0)
    pl = dict(pl, feel_disk=[1], feel_others=[1])
    run_pair(p, pl, _capi.dust_fluid_default_params(sfta=0, sizepart=1e-2, dust_feedback=1))


@pytest.mark.parametrize("feel", [_capi.DUST_SG_YES, _capi.DUST_SG_ZERO_MODE])
def test_self_gravity_energy(feel):
    p, pl = energy_sg(nrad=32, nsec=96)
    fp = _capi.dust_fluid_default_params(sfta=0, sizepart=1e-2, dust_feel_sg=feel)
    run_pair(p, pl, fp, calls=1)


@pytest.mark.parametrize("nr,ns", [(32, 96), (1024, 3072)])
def test_copy_of_the_gas_stays_a_copy(nr, ns):
    """One sub-step with the dust a copy of the gas (same sound speed,
    density, velocities, viscosity, no drag): both leave it bit for bit the
    same except the edge rings' vtheta (the dust's own sub-Keplerian
    profile), as in the oracle (tests/test_oracle_dust_fluid.py), also at
    the full 1024x3072 size."""
    p, pl = gas_planet(nrad=nr, nsec=ns)
    fp = _capi.dust_fluid_default_params(sfta=0, dust_feel_disk=0, dust_to_gas_density_ratio=1.0,
                                         daspect_ratio=p.aspect_ratio, dflaring_index=p.flaring_index,
                                         dviscosity=p.viscosity, dust_diffusion=1, dust_feel_sg=0)
    print(f"case dust fluid copy of the gas {nr}x{ns}", flush=True)
    d = Disc(p)
    d.attach_dust_fluid(fp, init_dust_fluid_density(p, fp))
    sg = d.initial_state(pl)
    d.set_dust_fluid(drho=sg["rho"], dvr=sg["vr"], dvt=sg["vt"])
    d.run_substeps(1e-4, 1)
    g, f = d.state(), d.dust_fluid()
    n, ns = p.nrad * p.nsec, p.nsec
    assert np.array_equal(f["drho"][:n], g["rho"][:n])
    assert np.array_equal(f["dvr"][:n], g["vr"][:n])
    assert np.array_equal(f["dvt"][ns:n - ns], g["vt"][ns:n - ns])


def test_graph_matches_eager():
    """The captured sub-step (graph) against the kernels launched one by one
    (profiling mode): bitwise."""
    p, pl = gas_planet(nrad=32, nsec=96)
    fp = _capi.dust_fluid_default_params(sfta=0, sizepart=1e-2, dust_feedback=1, dust_diffusion=1)
    drho = init_dust_fluid_density(p, fp)
    out = []
    for prof in (0, 1):
        d = Disc(p)
        d.attach_dust_fluid(fp, drho)
        d.initial_state(pl)
        d.set_profiling(prof)
        d.algogas(DT)
        out.append((d.state(), d.dust_fluid()))
    for k in ("rho", "vr", "vt"):
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    for k in ("drho", "dvr", "dvt", "stokes"):
        assert np.array_equal(out[0][1][k], out[1][1][k]), k


def test_refusals():
    p, pl = gas_planet(nrad=32, nsec=96)
    fp = _capi.dust_fluid_default_params()
    drho = init_dust_fluid_density(p, fp)
    d = Disc(p, _capi.FP32)
    with pytest.raises(DfargoError):
        d.attach_dust_fluid(fp, drho)
    d = Disc(p)
    d.attach_dust_fluid(fp, drho)
    with pytest.raises(DfargoError):
        d.attach_dust_fluid(fp, drho)
    d.initial_state(pl)
    dp = _capi.dust_default_params(nbpart=10)
    with pytest.raises(DfargoError):
        d.attach_dust(dp, {k: np.ones(10) for k in _capi.DUST_ARRAYS}, DT)


@pytest.mark.parametrize("bc,wkz", [(_capi.BC_WALL, 1), (_capi.BC_OPEN, 0)])
def test_open_dust_edge(bc, wkz):
    """INNERBOUNDARYDUST OPEN (OpenBoundaryd, src/SideEuler.c:411-466) beside
    a wall or open gas edge; with the wave-killing zones the fluid is left
    undamped."""
    p, pl = gas_planet(nrad=32, nsec=96, boundary=bc, wave_killing_zones=wkz, wkz_rmin=0.55, wkz_rmax=2.2)
    run_pair(p, pl, _capi.dust_fluid_default_params(sfta=0, sizepart=1e-2, open_inner_dust=1, dust_diffusion=1))


def test_restart_from_files(tmp_path):
    """Output 1 written from the GPU run (gas and dust fields, planet files);
    the GPU and the oracle both restart from the files and take a call:
    they agree to 1e-12."""
    from dustyfargoadsg_amd import output
    p, pl = gas_planet(nrad=32, nsec=96, forced_circular=0)
    pl = dict(pl, feel_disk=[1], feel_others=[1])
    fp = _capi.dust_fluid_default_params(sizepart=1e-2, dust_feedback=1)
    drho0 = init_dust_fluid_density(p, fp)
    print("case dust fluid restart 32x96", flush=True)
    d = Disc(p)
    d.attach_dust_fluid(fp, drho0)
    d.initial_state(pl)
    d.algogas(DT)
    st, s = d.state(), d.status()
    output.write_outputs(tmp_path, 1, st, p.nrad, p.nsec, dust_fluid=d.dust_fluid())
    output.write_planet_files(tmp_path, 1, st["planets"], s)
    runs = []
    for side in ("gpu", "oracle"):
        r = Disc(p) if side == "gpu" else Oracle(p)
        r.attach_dust_fluid(fp, drho0)
        rst = r.initial_state(pl)
        rmed = Oracle(p).tables()[1]
        back = output.read_restart(tmp_path, 1, p, rmed)
        st2 = dict(rho=back["rho"], vr=back["vr"], vt=back["vt"],
                   planets=output.restart_planetary_system(tmp_path, 1, rst["planets"]))
        if side == "oracle":
            st2["energy"] = rst["energy"]
        r.set_state(st2)
        r.set_time(output.restart_scalars(tmp_path, 1, p.corotating)[1], 0.0, 0.0)
        bd = output.read_restart_dust_fluid(tmp_path, 1, p)
        r.set_dust_fluid(bd["drho"], bd["dvr"], bd["dvt"])
        runs.append(r)
    g, o = runs
    n_o = o.algogas(DT)
    g.algogas(DT)
    assert g.status().last_call_substeps == n_o
    compare(o, g, p, 1e-12, 1e-12, 1e-12)
    compare_dust(o, g, p, 1e-12)



@pytest.mark.parametrize("full", [0, 1])
def test_full_size_finite(full):
    """The planet disc at its full 1024x3072 size: one AlgoGas call in the
    short-friction-time approximation (96 sub-steps), 50 sub-steps
    of the momentum equations with viscosity and diffusion (1 cm grains
    limit their step to ~4e-5): every field stays finite, the densities and
    Stokes numbers positive.  (Neither fluid's mass is conserved to round-off
    in this configuration, on either side: the reference's wall does not
    hold v_r of ring 0 at zero, so mass crosses the inner edge, 2e-5 of the
    disc per call at 64x192.)"""
    p, pl = gas_planet(nrad=1024, nsec=3072)
    fp = _capi.dust_fluid_default_params(sfta=0 if full else 1, sizepart=1e-2,
                                         dviscosity=1e-6 if full else 0.0, dust_diffusion=full)
    print(f"case dust fluid full size 1024x3072 full={full}", flush=True)
    d = Disc(p)
    d.attach_dust_fluid(fp, init_dust_fluid_density(p, fp))
    d.initial_state(pl)
    if full:
        d.run_substeps(DT, 50)
    else:
        d.algogas(DT)
        assert d.status().last_call_substeps > 50
    f, g = d.dust_fluid(), d.state()
    n = p.nrad * p.nsec
    for k in ("drho", "dvr", "dvt", "stokes"):
        assert np.isfinite(f[k][:n]).all(), k
    for k in ("rho", "vr", "vt"):
        assert np.isfinite(g[k][:n]).all(), k
    assert (f["drho"][:n] > 0).all() and (f["stokes"][:n] > 0).all()
