"""The reference's on-disk formats and its restart path, for hosts that drive
the hydro step without the reference's own output code.

Each function restates one routine of src/Output.c, src/Init.c or
src/SourceEuler.c (file:line in its docstring) and writes or reads the same
bytes and the same text, so that outputs of either code can be read by the
other and a run of one can restart from the other's files:

* fields: raw native-endian FP64, the first Nrad * Nsec values of the grid
  (rings 0 .. Nrad-1; the extra outer face of vr is not written);
* planet<k>.dat: one line per output, "%d" then ten "%#.14g" columns
  (timestep, x, y, vx, vy, mass, LostMass, PhysicalTime, OmegaFrame, mdcp,
  exces_mdcp), tab-separated, appended;
* dims.dat and used_rad.dat as the reference writes them;
* DAMPTOVISCOUS's 1-D disc: gasdens/gasvrad.ascii_rad.<N>.dat ("%g\t%g"
  lines of Rmed1D and Sigma1D / Vrad1D), dims1D.dat and used_rad1D.dat;
* the dust fluid: dustdens, dustvrad, dustvtheta (and Stokes) like the gas
  fields.
"""
import os

import numpy as np


def _name(outdir, base):
    return os.path.join(str(outdir), base)


def write_disk_polar(outdir, name, number, field, nrad, nsec):
    """WriteDiskPolar (src/Output.c:178-218), one process: the first
    nrad * nsec values of `field` to <outdir>/<name><number>.dat."""
    a = np.ascontiguousarray(np.asarray(field, dtype=np.float64).ravel()[: nrad * nsec])
    if a.size != nrad * nsec:
        raise ValueError(f"{name}: {a.size} values for a {nrad}x{nsec} grid")
    with open(_name(outdir, f"{name}{number}.dat"), "wb") as f:
        f.write(a.tobytes())


def read_disk_polar(outdir, name, number, nrad, nsec):
    """ReadfromFile (src/Init.c:18-52), one process: nrad * nsec values into
    a grid of nrad + 1 rings (the state arrays' shape), ring nrad zero.
    Returns None where the reference warns and keeps its t = 0 settings (no
    such file)."""
    path = _name(outdir, f"{name}{number}.dat")
    if not os.path.exists(path):
        return None
    a = np.fromfile(path, dtype=np.float64, count=nrad * nsec)
    if a.size != nrad * nsec:
        raise ValueError(f"{path}: {a.size} values, expected {nrad * nsec}")
    out = np.zeros((nrad + 1) * nsec)
    out[: nrad * nsec] = a
    return out


PLANET_LINE = "%d" + "\t%#.14g" * 10 + "\n"


def write_planet_files(outdir, timestep, planets, status, mdcp=0.0, exces_mdcp=0.0):
    """WritePlanetSystemFile / WritePlanetFile (src/Output.c:39-73): append
    one line per planet to planet<k>.dat.  `status` is a dfargo_status (or
    anything with lost_mass, physical_time and omega_frame); mdcp and
    exces_mdcp are the circumplanetary-disc masses (0 unless computed)."""
    for k in range(planets.nb):
        with open(_name(outdir, f"planet{k}.dat"), "a") as f:
            f.write(PLANET_LINE % (timestep, planets.x[k], planets.y[k], planets.vx[k], planets.vy[k],
                                   planets.mass[k], status.lost_mass, status.physical_time,
                                   status.omega_frame, mdcp, exces_mdcp))


def get_from_planet_file(outdir, timestep, column, n):
    """GetfromPlanetFile (src/Output.c:110-145): the value in `column`
    (1 = the timestep) of the first line of planet<n>.dat whose timestep is
    `timestep`.  Columns are found as the reference finds them: by skipping
    the characters "eE0123456789-." of a number, then the separators
    "\\t :=>_"."""
    if column < 2:
        raise ValueError("invalid column number")
    path = _name(outdir, f"planet{n}.dat")
    with open(path) as f:
        for line in f:
            head = line.split(None, 1)
            if head and _leading_int(head[0]) == timestep:
                break
        else:
            raise ValueError(f"no entry {timestep} in {path}")
    pos = 0
    for _ in range(column - 1):
        pos = _skip(line, pos, "eE0123456789-.")
        pos = _skip(line, pos, "\t :=>_")
    return float(_leading_float(line[pos:]))


def _skip(s, pos, chars):
    while pos < len(s) and s[pos] in chars:
        pos += 1
    return pos


def _leading_int(tok):
    digits = ""
    for ch in tok:
        if ch.isdigit() or (ch == "-" and not digits):
            digits += ch
        else:
            break
    return int(digits) if digits not in ("", "-") else None


def _leading_float(s):
    s = s.strip().split()[0] if s.strip() else "0"
    return float(s)


def restart_planetary_system(outdir, timestep, planets):
    """RestartPlanetarySystem (src/Output.c:147-176): positions, velocities
    and masses of each planet from the `timestep` line of planet<k>.dat; the
    initial and final masses become the read mass.  Updates `planets` (a
    dfargo_planets) in place and returns it."""
    for k in range(planets.nb):
        planets.x[k] = get_from_planet_file(outdir, timestep, 2, k)
        planets.y[k] = get_from_planet_file(outdir, timestep, 3, k)
        vx = get_from_planet_file(outdir, timestep, 4, k)
        vy = get_from_planet_file(outdir, timestep, 5, k)
        m = get_from_planet_file(outdir, timestep, 6, k)
        planets.initial_mass[k] = m
        planets.final_mass[k] = m
        planets.vx[k], planets.vy[k], planets.mass[k] = vx, vy, m
    return planets


def restart_scalars(outdir, timestep, corotating):
    """The globals src/main.c:266-272 restores: LostMass (column 7),
    PhysicalTime (8) and, in a corotating frame, OmegaFrame (9) of
    planet0.dat.  OmegaFrame is None in a fixed frame (the caller keeps its
    OMEGAFRAME)."""
    lost = get_from_planet_file(outdir, timestep, 7, 0)
    t = get_from_planet_file(outdir, timestep, 8, 0)
    w = get_from_planet_file(outdir, timestep, 9, 0) if corotating else None
    return lost, t, w


def write_dims(outdir, rmax, ntot, ninterm, nrad, nsec):
    """WriteDim (src/Output.c:220-231)."""
    with open(_name(outdir, "dims.dat"), "w") as f:
        f.write("%d\t%d\t\t%d\t%d\t%f\t%d\t%d\t%d\n" % (0, 0, 0, 0, rmax, ntot // ninterm, nrad, nsec))


def write_used_rad(outdir, radii):
    """The used_rad.dat of FillPolar1DArrays (src/SourceEuler.c:145-155):
    the Nrad + 1 interface radii, "%.18g" one per line."""
    with open(_name(outdir, "used_rad.dat"), "w") as f:
        for r in radii:
            f.write("%.18g\n" % r)


def radii_1d(rmin1d, rmax1d, nrad1d):
    """The 1-D disc's NRAD1D + 1 interface radii, logarithmically spaced
    (src/SideEuler.c:1465-1468)."""
    i = np.arange(nrad1d + 1, dtype=np.float64)
    return rmin1d * np.exp(i / nrad1d * np.log(rmax1d / rmin1d))


def write_dims_1d(outdir, nrad1d):
    """dims1D.dat of InitializeOneDViscousEvolution (src/SideEuler.c:1504-1508):
    NRAD1D eight times, each followed by a tab."""
    with open(_name(outdir, "dims1D.dat"), "w") as f:
        f.write("%d\t" * 8 % ((nrad1d,) * 8) + "\n")


def write_used_rad_1d(outdir, radii1d):
    """used_rad1D.dat (src/SideEuler.c:1509-1513): the NRAD1D + 1 interface
    radii, "%f" one per line."""
    with open(_name(outdir, "used_rad1D.dat"), "w") as f:
        for r in radii1d:
            f.write("%f\n" % r)


def write_visc1d_profiles(outdir, number, rmed1d, sigma1d, vrad1d):
    """Write1DViscProfiles (src/SideEuler.c:1564-1585), written with the gas
    fields when the wave-killing zones damp towards the 1-D disc
    (src/main.c:341-342): "%g\t%g" lines of (Rmed1D, Sigma1D) to
    gasdens.ascii_rad.<number>.dat and of (Rmed1D, Vrad1D) to
    gasvrad.ascii_rad.<number>.dat."""
    for base, vals in (("gasdens", sigma1d), ("gasvrad", vrad1d)):
        with open(_name(outdir, f"{base}.ascii_rad.{number}.dat"), "w") as f:
            for r, v in zip(rmed1d, vals):
                f.write("%g\t%g\n" % (r, v))


def read_visc1d_restart(outdir, number, nrad1d):
    """The restart branch of InitializeOneDViscousEvolution
    (src/SideEuler.c:1489-1502): the first NRAD1D (Rmed1D, Sigma1D) pairs of
    gasdens.ascii_rad.<number>.dat, which replace the 1-D disc's cell
    centres and surface density.  Vrad1D keeps its t = 0 profile: the
    reference opens gasvrad.ascii_rad.<number>.dat for writing, so its reads
    there return nothing (and the file is emptied; this reader leaves it as
    it is).  Returns (rmed1d, sigma1d)."""
    path = _name(outdir, f"gasdens.ascii_rad.{number}.dat")
    with open(path) as f:
        vals = f.read().split()
    if len(vals) < 2 * nrad1d:
        raise ValueError(f"{path}: {len(vals) // 2} pairs, expected {nrad1d}")
    a = np.array([float(v) for v in vals[: 2 * nrad1d]], dtype=np.float64).reshape(nrad1d, 2)
    return a[:, 0].copy(), a[:, 1].copy()


def write_outputs(outdir, number, state, nrad, nsec, temperature=None, dust_fluid=None, stokes=False):
    """The gas fields of SendOutput (src/Output.c:235-284) at output
    `number`: gasdens, gasvrad, gasvtheta and, given (with the energy
    equation, where Write_Temperature is on), the Temperature grid.  With
    the dust fluid (a dict of drho, dvr, dvt[, stokes] as
    Disc.dust_fluid() returns) also dustdens, dustvrad, dustvtheta and, with
    stokes=True (WRITEDUSTSTOKES), the Stokes grid."""
    write_disk_polar(outdir, "gasdens", number, state["rho"], nrad, nsec)
    if dust_fluid is not None:
        write_disk_polar(outdir, "dustdens", number, dust_fluid["drho"], nrad, nsec)
    write_disk_polar(outdir, "gasvrad", number, state["vr"], nrad, nsec)
    write_disk_polar(outdir, "gasvtheta", number, state["vt"], nrad, nsec)
    if dust_fluid is not None:
        write_disk_polar(outdir, "dustvrad", number, dust_fluid["dvr"], nrad, nsec)
        write_disk_polar(outdir, "dustvtheta", number, dust_fluid["dvt"], nrad, nsec)
    if temperature is not None:
        write_disk_polar(outdir, "Temperature", number, temperature, nrad, nsec)
    if dust_fluid is not None and stokes:
        write_disk_polar(outdir, "Stokes", number, dust_fluid["stokes"], nrad, nsec)


def read_restart(outdir, number, params, rmed):
    """The restart reads of src/Init.c:133-166: Sigma, vr and vtheta of
    output `number`; vtheta shifted back to the fixed frame when planet0.dat
    records a positive OmegaFrame and this run is not corotating; with the
    energy equation, the energy from the Temperature grid,
    e = Sigma T / (gamma - 1).  Missing field files give None (the reference
    keeps its t = 0 values)."""
    nrad, nsec = params.nrad, params.nsec
    st = {"rho": read_disk_polar(outdir, "gasdens", number, nrad, nsec),
          "vr": read_disk_polar(outdir, "gasvrad", number, nrad, nsec),
          "vt": read_disk_polar(outdir, "gasvtheta", number, nrad, nsec)}
    omega_prev = get_from_planet_file(outdir, number, 9, 0)
    if st["vt"] is not None and omega_prev > 0.0 and not params.corotating:
        v = st["vt"][: nrad * nsec].reshape(nrad, nsec)
        v += np.asarray(rmed[:nrad])[:, None] * omega_prev
    if params.energy_equation:
        temp = read_disk_polar(outdir, "Temperature", number, nrad, nsec)
        st["energy"] = None
        if temp is not None and st["rho"] is not None:
            st["energy"] = st["rho"] * temp / (params.adiabatic_index - 1.0)
    return st


def read_restart_dust_fluid(outdir, number, params):
    """The dust fluid's restart reads (src/Init.c:167-185): dustdens,
    dustvrad, dustvtheta of output `number` (None where a file is missing).
    Its vtheta is not shifted back to the fixed frame: the reference's test
    there assigns instead of comparing (`Corotating = NO`), so it is false,
    and when planet0.dat records a positive OmegaFrame it also switches the
    run's frame off; "frame_fixed" is True in that case, and a host that
    wants the reference's run sets params.corotating = 0 before the
    restart's first call."""
    nrad, nsec = params.nrad, params.nsec
    st = {"drho": read_disk_polar(outdir, "dustdens", number, nrad, nsec),
          "dvr": read_disk_polar(outdir, "dustvrad", number, nrad, nsec),
          "dvt": read_disk_polar(outdir, "dustvtheta", number, nrad, nsec)}
    st["frame_fixed"] = get_from_planet_file(outdir, number, 9, 0) > 0.0
    return st
