"""CPU tests of the C ABI: the library loads without a GPU, exports every
entry point include/dfargo.h declares, its structures match the ctypes
mirror, and its host-side reference functions (geometry, initial profiles,
planets) agree bit for bit with the oracle's restatement."""
import ctypes
import math
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from dustyfargoadsg_amd import _capi
from dustyfargoadsg_amd.disc import Disc, DfargoError, initial_fields, make_planets
from dustyfargoadsg_amd.workloads import gas_planet, energy_sg
from tests.oracle_lib import Oracle, ROOT

HEADER = os.path.join(ROOT, "include", "dfargo.h")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_capi.LIB_PATH):
        pytest.fail("libdfargo.so is not built: run __graft_entry__.build()")
    return _capi.load()


def test_exports_every_declared_symbol(lib):
    decl = set(re.findall(r"\b(dfargo_\w+)\s*\(", open(HEADER).read()))
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = sorted(decl - exported)
    assert not missing, missing
    assert decl <= set(_capi.EXPORTED) | {"dfargo_last_error"}


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "dfargo.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu\\n\", sizeof(dfargo_params),"
        " sizeof(dfargo_planets), sizeof(dfargo_status), offsetof(dfargo_params, boundary),"
        " offsetof(dfargo_planets, final_mass), offsetof(dfargo_status, crashed));return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    want = [ctypes.sizeof(_capi.Params), ctypes.sizeof(_capi.Planets), ctypes.sizeof(_capi.Status),
            _capi.Params.boundary.offset, _capi.Planets.final_mass.offset, _capi.Status.crashed.offset]
    assert got == want


def test_defaults_match_python_mirror(lib):
    p = _capi.Params()
    lib.dfargo_default_params(ctypes.byref(p))
    q = _capi.default_params()
    assert p.as_dict() == q.as_dict()


@pytest.mark.parametrize("factory", [gas_planet, energy_sg])
def test_host_tables_match_oracle(lib, oracle_built, factory):
    p, _ = factory(nrad=32, nsec=96)
    o = Oracle(p)
    radii, rmed, surf, az = o.tables()
    r2, m2, s2, a2 = (np.zeros(p.nrad + 1), np.zeros(p.nrad), np.zeros(p.nrad), np.zeros(p.nsec))
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert lib.dfargo_fill_polar_1d_arrays(ctypes.byref(p), dp(r2), dp(m2), dp(s2), dp(a2)) == 0
    for a, b in ((radii, r2), (rmed, m2), (surf, s2), (az, a2)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("factory", [gas_planet, energy_sg])
def test_host_initial_state_matches_oracle(lib, oracle_built, factory):
    """InitGasDensity (with the drand48 AddNoise stream), InitGasEnergy and
    InitPlanetarySystem are bit-identical to the oracle's."""
    p, pl = factory(nrad=32, nsec=96)
    o = Oracle(p)
    so = o.initial_state(pl)
    rho, e = initial_fields(p, lib)
    assert np.array_equal(rho, so["rho"])
    if p.energy_equation:
        assert np.array_equal(e, so["energy"])
    sys_ = make_planets(p, pl, lib)
    assert sys_.nb == so["planets"].nb
    for k in ("mass", "x", "y", "vx", "vy", "final_mass"):
        assert list(getattr(sys_, k)) == list(getattr(so["planets"], k))


def test_noise_stream_is_drand48(lib):
    """AddNoise draws follow srand48(seed)/drand48, i-major then j
    (src/Pframeforce.c:309-320): the first draws equal the POSIX 48-bit LCG."""
    p, _ = gas_planet(nrad=8, nsec=16, seed=12345)
    rho, _ = initial_fields(p, lib)
    x = (12345 << 16) | 0x330E
    a, c, m = 0x5DEECE66D, 0xB, 1 << 48
    p.add_noise = 0
    base, _ = initial_fields(p, lib)
    for l in range(5):
        x = (a * x + c) % m
        u = x / float(m)
        assert rho[l] == base[l] + 1e-3 * base[l] * (2.0 * u - 1.0)


def test_parameter_checks(lib):
    def check(**kw):
        base = dict(nrad=16, nsec=32, rmin=0.5, rmax=2.0)
        base.update(kw)
        p = _capi.default_params(**base)
        return lib.dfargo_check_params(ctypes.byref(p))
    assert check() == _capi.OK
    assert check(viscosity=1e-5, alpha_viscosity=1e-3) == _capi.ERR_ARG
    assert check(exclude_hill_factor=1.5) == _capi.ERR_ARG
    assert check(self_gravity=1, log_grid=0) == _capi.ERR_ARG
    assert check(corotating=_capi.FRAME_BINARY_CENTER) == _capi.OK
    assert check(corotating=4) == _capi.ERR_ARG
    assert check(boundary=_capi.BC_NONREFLECTING) == _capi.OK
    assert check(boundary=_capi.BC_NONREFLECTING, energy_equation=1) == _capi.OK
    assert check(boundary=3) == _capi.ERR_UNSUPPORTED
    assert check(wave_killing_zones=1) == _capi.ERR_ARG
    assert check(wave_killing_zones=1, wkz_rmin=0.5, wkz_rmax=0.4) == _capi.ERR_ARG
    assert check(wave_killing_zones=1, wkz_rmin=0.5, wkz_rmax=2.2,
                 damp_to=_capi.DAMP_TO_VISCOUS) == _capi.ERR_ARG  # needs a viscosity
    assert check(wave_killing_zones=1, wkz_rmin=0.5, wkz_rmax=2.2, viscosity=1e-5,
                 damp_to=_capi.DAMP_TO_VISCOUS) == _capi.OK
    assert check(wave_killing_zones=1, wkz_rmin=0.5, wkz_rmax=2.2, damp_to=_capi.DAMP_TO_INI) == _capi.OK
    assert check(self_gravity=1, log_grid=1, nsec=40) == _capi.ERR_UNSUPPORTED
    assert b"VISCOSITY" in (lib.dfargo_check_params(ctypes.byref(_capi.default_params(
        nrad=16, nsec=32, rmin=0.5, rmax=2.0, viscosity=1, alpha_viscosity=1))) and
        lib.dfargo_last_error())


def test_create_fails_loudly_without_device(lib):
    """With no HIP device the context cannot be created and says why: the
    hydro step has no CPU fallback."""
    p, _ = gas_planet(nrad=16, nsec=48)
    try:
        d = Disc(p)
    except DfargoError as e:
        assert "device" in str(e).lower() or "hip" in str(e).lower()
        return
    d.close()
    pytest.skip("a HIP device is visible; covered by the gpu tests")


def test_dust_struct_layout_and_defaults(lib, tmp_path):
    src = tmp_path / "dlayout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "dfargo.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu\\n\", sizeof(dfargo_dust_params),"
        " sizeof(dfargo_dust_state), offsetof(dfargo_dust_params, seed),"
        " offsetof(dfargo_dust_params, dust_feel_sg_zero_mode));return 0;}\n")
    exe = tmp_path / "dlayout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    want = [ctypes.sizeof(_capi.DustParams), ctypes.sizeof(_capi.DustState),
            _capi.DustParams.seed.offset, _capi.DustParams.dust_feel_sg_zero_mode.offset]
    assert got == want
    d = _capi.DustParams()
    lib.dfargo_dust_default_params(ctypes.byref(d), 123)
    q = _capi.dust_default_params(123)
    for name, _ in _capi.DustParams._fields_:
        assert getattr(d, name) == getattr(q, name), name


@pytest.mark.parametrize("case", [
    dict(sizepartslope=3.5, dustslope=0.5),
    dict(sizepartslope=1.0, dustslope=1.0),
    dict(sizeminpart=0.01, sizemaxpart=0.01, sizepartslope=0.0, rmindust=0.0, rmaxdust=0.0),
])
def test_host_dust_init_matches_oracle(lib, oracle_built, case):
    """InitDustSystem on the host (sizes, radii, azimuths from one drand48
    stream, Keplerian vtheta) is bit-identical to the oracle's."""
    from dustyfargoadsg_amd.disc import init_dust_system
    p, _ = gas_planet(nrad=32, nsec=96)
    kw = dict(sizeminpart=1e-3, sizemaxpart=1e-1, rmindust=0.6, rmaxdust=2.2, particle_mass=1e-9, seed=7)
    kw.update(case)
    d = _capi.dust_default_params(2000, **kw)
    o = Oracle(p)
    want = o.dust_init(d)
    got = init_dust_system(p, d, None, lib)
    for k in ("dustsize", "r", "th", "vr", "vth", "l"):
        assert np.array_equal(got[k], want[k]), k
    assert np.array_equal(got["dustsize_init"], want["dustsize"])


def test_host_dust_init_with_selfgravity_matches_oracle(lib, oracle_built):
    """With self-gravity vtheta takes the axisymmetric SG acceleration
    interpolated at the particle (src/Dsys.c:203-212)."""
    from dustyfargoadsg_amd.disc import init_dust_system
    p, _ = energy_sg(nrad=32, nsec=96)
    axi = -1e-3 * np.linspace(1.0, 2.0, p.nrad)
    d = _capi.dust_default_params(500, sizeminpart=1e-3, sizemaxpart=1e-2, seed=3, particle_mass=1e-9)
    want = Oracle(p).dust_init(d, axi)
    got = init_dust_system(p, d, axi, lib)
    for k in ("r", "th", "vth", "l"):
        assert np.array_equal(got[k], want[k]), k
    assert not np.array_equal(got["vth"], got["r"] ** -0.5)


def test_dust_parameter_checks(lib):
    p, _ = gas_planet(nrad=16, nsec=48)

    def check(**kw):
        base = dict(sizepartslope=0.0)
        base.update(kw)
        d = _capi.dust_default_params(10, **base)
        return lib.dfargo_check_dust_params(ctypes.byref(p), ctypes.byref(d))
    assert check() == _capi.OK
    # the defaults of src/var.c:136-138 themselves are inconsistent, and the
    # reference refuses them too (src/Dsys.c:151-155)
    assert check(sizepartslope=3.0) == _capi.ERR_ARG
    assert check(interpolation=3) == _capi.ERR_ARG
    assert check(rhopart=0.0) == _capi.ERR_ARG
    assert check(no_dt_constraint_by_particles=0, dust_timestep=0.0) == _capi.ERR_ARG
    for k in ("dust_feel_turb", "dust_growth", "remove_dust_from_hill", "zz_integrator",
              "dust_feel_sg_zero_mode"):
        assert check(**{k: 1}) == _capi.ERR_UNSUPPORTED, k
    # particles in a rotating frame are supported
    q = p.copy()
    q.omega_frame = 1.0
    d = _capi.dust_default_params(10, sizepartslope=0.0)
    assert lib.dfargo_check_dust_params(ctypes.byref(q), ctypes.byref(d)) == _capi.OK


def test_orbital_elements_match_oracle(lib):
    """dfargo_orbital_elements is FindOrbitalElements, bit for bit the
    oracle's restatement on random bound orbits; a non-positive mass is an
    argument error."""
    from dustyfargoadsg_amd.disc import orbital_elements
    from tests.oracle_lib import orbital_elements as or_elements
    rng = np.random.default_rng(3)
    for _ in range(50):
        r, phi = rng.uniform(0.5, 2.0), rng.uniform(-3, 3)
        v = math.sqrt(1.0 / r) * rng.uniform(0.8, 1.2)
        ang = phi + math.pi / 2 + rng.uniform(-0.3, 0.3)
        args = (r * math.cos(phi), r * math.sin(phi), v * math.cos(ang), v * math.sin(ang), 1.0 + 1e-3)
        a, b = orbital_elements(*args, lib=lib), or_elements(*args)
        assert np.array_equal(a, b) or np.allclose(a, b, rtol=0, atol=0, equal_nan=True)
    with pytest.raises(Exception):
        orbital_elements(1.0, 0.0, 0.0, 1.0, 0.0, lib=lib)
