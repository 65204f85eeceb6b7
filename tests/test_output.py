"""The reference's on-disk formats and restart path (src/Output.c,
src/Init.c:18-166, src/main.c:266-272), restated in
dustyfargoadsg_amd/output.py.  Host code only: the oracle supplies the
states.  Checks: byte layout and sizes of the field files, the planet-file
line format and GetfromPlanetFile's column rules, the text of dims.dat and
used_rad.dat, and a restart that continues a run of the oracle exactly
(fields are binary; the planets pass through %.14g text, so they are
compared to that precision)."""
import ctypes
import math

import numpy as np

from dustyfargoadsg_amd import output
from dustyfargoadsg_amd.workloads import gas_planet, energy_sg
from tests.oracle_lib import Oracle

DT = 2 * math.pi / 20


def test_field_files_round_trip(tmp_path, oracle_built):
    p, pl = gas_planet(nrad=16, nsec=48)
    o = Oracle(p)
    st = o.initial_state(pl)
    nr, ns = p.nrad, p.nsec
    output.write_outputs(tmp_path, 3, st, nr, ns)
    for name in ("gasdens", "gasvrad", "gasvtheta"):
        assert (tmp_path / f"{name}3.dat").stat().st_size == nr * ns * 8
    back = output.read_disk_polar(tmp_path, "gasvrad", 3, nr, ns)
    assert np.array_equal(back[: nr * ns], st["vr"][: nr * ns]) and not back[nr * ns:].any()
    assert output.read_disk_polar(tmp_path, "gasdens", 4, nr, ns) is None


def test_planet_file_lines_and_columns(tmp_path, oracle_built):
    p, pl = gas_planet(nrad=16, nsec=48)
    o = Oracle(p)
    o.initial_state(pl)
    o.algogas(DT)
    s, sys_ = o.status(), o.state()["planets"]
    output.write_planet_files(tmp_path, 0, sys_, s)
    o.algogas(DT)
    s1, sys1 = o.status(), o.state()["planets"]
    output.write_planet_files(tmp_path, 1, sys1, s1)
    lines = (tmp_path / "planet0.dat").read_text().splitlines()
    assert len(lines) == 2
    # the C library's own "%d" and "%#.14g" give the same text
    libc = ctypes.CDLL(None)
    buf = ctypes.create_string_buffer(64)
    cols = lines[1].split("\t")
    assert len(cols) == 11 and cols[0] == "1"
    for c in cols[1:]:
        libc.snprintf(buf, 64, b"%#.14g", ctypes.c_double(float(c)))
        assert buf.value.decode() == c
    # GetfromPlanetFile picks the line of the asked timestep
    assert output.get_from_planet_file(tmp_path, 1, 2, 0) == float("%.14g" % sys1.x[0])
    assert output.get_from_planet_file(tmp_path, 0, 8, 0) == float("%.14g" % s.physical_time)
    assert output.get_from_planet_file(tmp_path, 1, 9, 0) == 0.0


def test_dims_and_used_rad(tmp_path, oracle_built):
    output.write_dims(tmp_path, 2.5, 100, 10, 128, 384)
    assert (tmp_path / "dims.dat").read_text() == "0\t0\t\t0\t0\t2.500000\t10\t128\t384\n"
    p, _ = gas_planet(nrad=16, nsec=48)
    radii = Oracle(p).tables()[0][:17]
    output.write_used_rad(tmp_path, radii)
    back = np.array([float(x) for x in (tmp_path / "used_rad.dat").read_text().split()])
    assert np.array_equal(back, radii)  # %.18g is exact for doubles


def test_restart_continues_the_run(tmp_path, oracle_built):
    """Write output 1 after one call, restart a fresh oracle from it (fields,
    planets, LostMass, PhysicalTime) and take the next call on both: the
    fields agree to the planets' 14-digit text precision."""
    p, pl = gas_planet(nrad=32, nsec=96, forced_circular=0)
    pl = dict(pl, feel_disk=[1], feel_others=[1])
    o = Oracle(p)
    o.initial_state(pl)
    o.algogas(DT)
    st, s = o.state(), o.status()
    output.write_outputs(tmp_path, 1, st, p.nrad, p.nsec)
    output.write_planet_files(tmp_path, 1, st["planets"], s)
    r = Oracle(p)
    rst = r.initial_state(pl)
    back = output.read_restart(tmp_path, 1, p, r.tables()[1])
    planets = output.restart_planetary_system(tmp_path, 1, rst["planets"])
    lost, t, w = output.restart_scalars(tmp_path, 1, p.corotating)
    assert w is None and t == float("%.14g" % s.physical_time)
    for k in ("rho", "vr", "vt"):
        rst[k][:] = back[k]
    rst["planets"] = planets
    r.set_state(rst)
    r.set_time(s.physical_time, 0.0, 0.0)
    assert r.algogas(DT) == o.algogas(DT)
    a, b = r.state(), o.state()
    n = p.nrad * p.nsec
    for k in ("rho", "vt"):
        assert np.abs(a[k][:n] - b[k][:n]).max() <= 1e-11 * np.abs(b[k][:n]).max(), k


def test_restart_energy_from_temperature(tmp_path, oracle_built):
    """With the energy equation the restart reads the Temperature grid and
    forms e = Sigma T / (gamma - 1) (src/Init.c:155-166); a planet0.dat must
    exist, as GetfromPlanetFile(..., 9, 0) is read in every restart."""
    p, pl = energy_sg(nrad=16, nsec=48)
    o = Oracle(p)
    st = o.initial_state(pl)
    temp = o.field("temperature")
    output.write_outputs(tmp_path, 2, st, p.nrad, p.nsec, temperature=temp)

    class S:
        lost_mass = physical_time = omega_frame = 0.0

    class One:
        nb, x, y, vx, vy, mass = 1, [1.0], [0.0], [0.0], [1.0], [0.0]

    output.write_planet_files(tmp_path, 2, One, S)
    back = output.read_restart(tmp_path, 2, p, o.tables()[1])
    n = p.nrad * p.nsec
    assert np.abs(back["energy"][:n] - st["energy"][:n]).max() <= 1e-14 * np.abs(st["energy"][:n]).max()


def _libc_format(fmt, *vals):
    libc = ctypes.CDLL(None)
    buf = ctypes.create_string_buffer(256)
    args = [ctypes.c_double(v) if isinstance(v, float) else ctypes.c_int(v) for v in vals]
    libc.snprintf(buf, 256, fmt, *args)
    return buf.value.decode()


def _visc1d_case():
    return gas_planet(nrad=32, nsec=96, damp_to=2, bc1d=1, wave_killing_zones=1, wkz_rmin=0.55, wkz_rmax=2.2)


def test_visc1d_files_match_the_c_formats(tmp_path, oracle_built):
    """Write1DViscProfiles' "%lg\\t%lg\\n" lines, dims1D.dat's eight "%d\\t"
    and used_rad1D.dat's "%f\\n" (src/SideEuler.c:1504-1513,1564-1585),
    compared byte for byte with the C library's own formatting of the same
    values; the 1-D radii give back the oracle's Rmed1D."""
    p, pl = _visc1d_case()
    o = Oracle(p)
    o.initial_state(pl)
    o.algogas(DT)
    rmed, sig, vr = o.visc1d()
    output.write_visc1d_profiles(tmp_path, 4, rmed, sig, vr)
    for base, vals in (("gasdens", sig), ("gasvrad", vr)):
        want = "".join(_libc_format(b"%lg\t%lg\n", float(a), float(b)) for a, b in zip(rmed, vals))
        assert (tmp_path / f"{base}.ascii_rad.4.dat").read_text() == want
    output.write_dims_1d(tmp_path, p.nrad1d)
    assert (tmp_path / "dims1D.dat").read_text() == _libc_format(b"%d\t" * 8 + b"\n", *([p.nrad1d] * 8))
    radii = output.radii_1d(p.rmin1d, p.rmax1d, p.nrad1d)
    output.write_used_rad_1d(tmp_path, radii)
    assert (tmp_path / "used_rad1D.dat").read_text() == "".join(_libc_format(b"%f\n", float(r)) for r in radii)
    # the cell centres of src/SideEuler.c:1472-1473 from these radii
    a, b = radii[:-1], radii[1:]
    rm = 2.0 / 3.0 * (b * b * b - a * a * a) / (b * b - a * a)
    assert (np.abs(rm - rmed) / rmed).max() <= 4e-15


def test_visc1d_restart_read(tmp_path, oracle_built):
    """The restart branch of InitializeOneDViscousEvolution
    (src/SideEuler.c:1489-1502): Rmed1D and Sigma1D become the values the C
    library's strtod reads from the 6-digit gasdens.ascii_rad file, Vrad1D
    keeps its t = 0 profile (the reference opens the vr file for writing).
    A restarted oracle continues within the text's precision of the run it
    came from."""
    p, pl = _visc1d_case()
    o = Oracle(p)
    o.initial_state(pl)
    for _ in range(2):
        o.algogas(DT)
    rmed, sig, vr = o.visc1d()
    output.write_visc1d_profiles(tmp_path, 2, rmed, sig, vr)
    rm2, sg2 = output.read_visc1d_restart(tmp_path, 2, p.nrad1d)
    libc = ctypes.CDLL(None)
    libc.strtod.restype = ctypes.c_double
    libc.strtod.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
    toks = (tmp_path / "gasdens.ascii_rad.2.dat").read_bytes().split()
    want = np.array([libc.strtod(t, None) for t in toks]).reshape(-1, 2)
    assert np.array_equal(rm2, want[:, 0]) and np.array_equal(sg2, want[:, 1])
    assert np.abs(rm2 - rmed).max() <= 5e-6 * rmed.max() and np.abs(sg2 - sig).max() <= 5e-6 * sig.max()
    assert not np.array_equal(rm2, rmed)  # the text rounds Rmed1D
    r = Oracle(p)
    r.initial_state(pl)
    _, _, vr0 = r.visc1d()
    r.set_state(o.state())
    r.set_visc1d(rm2, sg2)
    a_r, a_s, a_v = r.visc1d()
    assert np.array_equal(a_r, rm2) and np.array_equal(a_s, sg2) and np.array_equal(a_v, vr0)
    assert not np.array_equal(vr0, vr)
    assert r.algogas(DT) == o.algogas(DT)
    rs, os_ = r.state(), o.state()
    n = p.nrad * p.nsec
    for k in ("rho", "vt"):
        assert np.abs(rs[k][:n] - os_[k][:n]).max() <= 1e-4 * np.abs(os_[k][:n]).max(), k
    _, s_r, _ = r.visc1d()
    _, s_o, _ = o.visc1d()
    assert np.abs(s_r - s_o).max() <= 1e-4 * s_o.max()


def test_restart_dust_fluid(tmp_path, oracle_built):
    """Output 1 with the dust fluid's dustdens/dustvrad/dustvtheta, a fresh
    oracle restarted from the files, the next call on both: gas and dust
    agree to the planet files' 14-digit text (short-friction-time
    approximation, where the stored Stokes grid does not enter the time
    step)."""
    from dustyfargoadsg_amd import _capi
    from dustyfargoadsg_amd.disc import init_dust_fluid_density
    p, pl = gas_planet(nrad=32, nsec=96, forced_circular=0)
    pl = dict(pl, feel_disk=[1], feel_others=[1])
    fp = _capi.dust_fluid_default_params(sizepart=1e-2)
    drho0 = init_dust_fluid_density(p, fp)
    o = Oracle(p)
    o.attach_dust_fluid(fp, drho0)
    o.initial_state(pl)
    o.algogas(DT)
    st, s, df = o.state(), o.status(), o.dust_fluid()
    output.write_outputs(tmp_path, 1, st, p.nrad, p.nsec, dust_fluid=df, stokes=True)
    output.write_planet_files(tmp_path, 1, st["planets"], s)
    assert (tmp_path / "Stokes1.dat").stat().st_size == p.nrad * p.nsec * 8
    r = Oracle(p)
    r.attach_dust_fluid(fp, drho0)
    rst = r.initial_state(pl)
    back = output.read_restart(tmp_path, 1, p, r.tables()[1])
    for k in ("rho", "vr", "vt"):
        rst[k][:] = back[k]
    rst["planets"] = output.restart_planetary_system(tmp_path, 1, rst["planets"])
    r.set_state(rst)
    r.set_time(s.physical_time, 0.0, 0.0)
    bd = output.read_restart_dust_fluid(tmp_path, 1, p)
    assert not bd["frame_fixed"]
    r.set_dust_fluid(bd["drho"], bd["dvr"], bd["dvt"])
    assert r.algogas(DT) == o.algogas(DT)
    n = p.nrad * p.nsec
    a, b = r.state(), o.state()
    for k in ("rho", "vt"):
        assert np.abs(a[k][:n] - b[k][:n]).max() <= 1e-11 * np.abs(b[k][:n]).max(), k
    da, db = r.dust_fluid(), o.dust_fluid()
    for k in ("drho", "dvt"):
        assert np.abs(da[k][:n] - db[k][:n]).max() <= 1e-11 * np.abs(db[k][:n]).max(), k
