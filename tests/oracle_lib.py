"""numpy wrapper of oracle/liboracle.so (the FP64 CPU restatement).

Test infrastructure: used only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.
"""
import ctypes
import os
import subprocess

import numpy as np

from dustyfargoadsg_amd import _capi
from dustyfargoadsg_amd._capi import Params, Planets, Status, FIELDS, DustParams

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = ctypes.CDLL(ORACLE_SO)
        P, PP = ctypes.POINTER(Params), ctypes.POINTER(Planets)
        dp, vp, ip = ctypes.POINTER(ctypes.c_double), ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)
        sig = {
            "or_create": (vp, [P]), "or_destroy": (None, [vp]),
            "or_get_tables": (None, [vp, dp, dp, dp, dp]),
            "or_init_gas_density": (None, [vp, dp]),
            "or_init_gas_energy": (None, [vp, dp]),
            "or_init_planetary_system": (None, [vp, ctypes.c_int, dp, dp, dp, ip, ip, PP]),
            "or_init_euler": (ctypes.c_int, [vp, dp, dp, dp, dp, PP]),
            "or_set_state": (None, [vp, dp, dp, dp, dp, PP]),
            "or_get_state": (None, [vp, dp, dp, dp, dp, PP]),
            "or_set_time": (None, [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
            "or_evanescent_boundary": (None, [vp, ctypes.c_double]),
            "or_nonreflecting_boundary": (None, [vp]),
            "or_accrete_onto_planets": (None, [vp, ctypes.c_double]),
            "or_sg_zero_mode": (None, [vp, dp, dp]),
            "or_solve_1d_viscous": (None, [vp, ctypes.c_double, dp, dp]),
            "or_set_visc1d": (None, [vp, dp, dp, dp]),
            "or_dust_fluid_attach": (ctypes.c_int, [vp, ctypes.POINTER(_capi.DustFluidParams), dp]),
            "or_dust_fluid_get": (None, [vp, dp, dp, dp, dp]),
            "or_dust_fluid_set": (None, [vp, dp, dp, dp]),
            "or_get_visc1d": (None, [vp, dp, dp, dp]),
            "or_algogas": (ctypes.c_int, [vp, ctypes.c_double]),
            "or_get_status": (None, [vp, ctypes.POINTER(Status)]),
            "or_get_field": (None, [vp, ctypes.c_int, dp]),
            "or_get_fargo_shift": (None, [vp, ctypes.POINTER(ctypes.c_long), dp]),
            "or_condition_cfl": (ctypes.c_int, [vp, ctypes.c_double]),
            "or_transport": (None, [vp, ctypes.c_double]),
            "or_substep1": (None, [vp, ctypes.c_double]),
            "or_viscous": (None, [vp, ctypes.c_double]),
            "or_sg_accel": (None, [vp, dp, dp, dp]),
            "or_sg_accel_direct": (None, [vp, dp, dp, dp]),
            "or_fft": (None, [dp, dp, ctypes.c_int, ctypes.c_int]),
            "or_rk5": (None, [dp, ctypes.c_double, dp, dp, ctypes.c_int, ip, ctypes.c_int]),
            "or_fviscosity": (ctypes.c_double, [vp, ctypes.c_double]),
            "or_dust_init_system": (ctypes.c_int, [P, ctypes.POINTER(DustParams), dp, dp, dp, dp, dp, dp, dp]),
            "or_dust_attach": (ctypes.c_int, [vp, ctypes.POINTER(DustParams), dp, dp, dp, dp, dp, dp,
                                              ctypes.c_double]),
            "or_dust_interpolation": (None, [vp, ctypes.c_double]),
            "or_dust_get": (None, [vp, dp, dp, dp, dp, dp, dp]),
            "or_dust_get_gas_at_particles": (None, [vp, dp, dp, dp, dp]),
            "or_dust_get_drag": (None, [vp, dp, dp]),
            "or_dust_step": (None, [vp, ctypes.c_double]),
            "or_dust_min_stopping_time": (ctypes.c_double, [vp]),
            "or_dust_particles_mass": (ctypes.c_double, [vp]),
            "or_gas_budgets": (None, [vp, dp]),
            "or_update_log": (ctypes.c_int, [vp, dp, dp]),
            "or_orbital_elements": (None, [ctypes.c_double] * 5 + [dp]),
            "or_test_grid": (ctypes.c_int, [vp, ctypes.c_int, dp, ctypes.c_int]),
            "or_test_stage": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_planets(ctx_lib, ctx, spec):
    nb = len(spec["dist"])
    out = Planets()
    d = np.array(spec["dist"] + [0.0], dtype=np.float64)
    m = np.array(spec["mass"] + [0.0], dtype=np.float64)
    a = np.array(spec["acc"] + [0.0], dtype=np.float64)
    fd = (ctypes.c_int * (nb + 1))(*(spec["feel_disk"] + [0]))
    fo = (ctypes.c_int * (nb + 1))(*(spec["feel_others"] + [0]))
    ctx_lib.or_init_planetary_system(ctx, nb, ptr(d), ptr(m), ptr(a), fd, fo, ctypes.byref(out))
    return out


class Oracle:
    """One disc on the CPU oracle."""

    def __init__(self, params):
        self.p = params.copy()
        self.nr, self.ns = params.nrad, params.nsec
        self.L = lib()
        self.ctx = self.L.or_create(ctypes.byref(self.p))
        self.n = (self.nr + 1) * self.ns

    def __del__(self):
        try:
            self.L.or_destroy(self.ctx)
        except Exception:
            pass

    def zeros(self):
        return np.zeros(self.n, dtype=np.float64)

    def tables(self):
        radii = np.zeros(self.nr + 1); rmed = np.zeros(self.nr)
        surf = np.zeros(self.nr); az = np.zeros(self.ns)
        self.L.or_get_tables(self.ctx, ptr(radii), ptr(rmed), ptr(surf), ptr(az))
        return radii, rmed, surf, az

    def initial_state(self, planet_spec=None):
        """InitGasDensity, InitGasEnergy, InitPlanetarySystem, InitEuler."""
        rho, e = self.zeros(), self.zeros()
        self.L.or_init_gas_density(self.ctx, ptr(rho))
        if self.p.energy_equation:
            self.L.or_init_gas_energy(self.ctx, ptr(e))
        sys_ = make_planets(self.L, self.ctx, planet_spec or dict(
            dist=[], mass=[], acc=[], feel_disk=[], feel_others=[]))
        vr, vt = self.zeros(), self.zeros()
        self.L.or_init_euler(self.ctx, ptr(rho), ptr(vr), ptr(vt), ptr(e), ctypes.byref(sys_))
        return dict(rho=rho, vr=vr, vt=vt, energy=e, planets=sys_)

    def budgets(self):
        """GasTotalMass, GasMomentum, GasTotalEnergy, planets' angular momentum."""
        out = np.zeros(4)
        self.L.or_gas_budgets(self.ctx, ptr(out))
        return out

    def update_log(self):
        """UpdateLog: (nb, 8) torques and powers, (nb,) indirect torques."""
        tq, ind = np.zeros(8 * 8), np.zeros(8)
        nb = self.L.or_update_log(self.ctx, ptr(tq), ptr(ind))
        return tq[: 8 * nb].reshape(nb, 8), ind[:nb]

    def set_state(self, st):
        self.L.or_set_state(self.ctx, ptr(st["rho"]), ptr(st["vr"]), ptr(st["vt"]),
                            ptr(st.get("energy")),
                            ctypes.byref(st["planets"]) if st.get("planets") is not None else None)

    def state(self):
        rho, vr, vt, e = self.zeros(), self.zeros(), self.zeros(), self.zeros()
        sys_ = Planets()
        self.L.or_get_state(self.ctx, ptr(rho), ptr(vr), ptr(vt), ptr(e), ctypes.byref(sys_))
        return dict(rho=rho, vr=vr, vt=vt, energy=e, planets=sys_)

    def algogas(self, DT):
        return self.L.or_algogas(self.ctx, DT)

    def evanescent(self, step):
        self.L.or_evanescent_boundary(self.ctx, step)

    def nonreflecting(self):
        self.L.or_nonreflecting_boundary(self.ctx)

    def accrete(self, dt):
        self.L.or_accrete_onto_planets(self.ctx, dt)

    def viscous_1d(self, dt):
        s, v = np.zeros(self.p.nrad1d), np.zeros(self.p.nrad1d)
        self.L.or_solve_1d_viscous(self.ctx, dt, ptr(s), ptr(v))
        return s, v

    def attach_dust_fluid(self, fp, drho):
        self.fp = fp
        self.L.or_dust_fluid_attach(self.ctx, ctypes.byref(fp), ptr(np.ascontiguousarray(drho, dtype=np.float64)))

    def dust_fluid(self):
        """(DRho, DVrad, DVtheta, Stokes) grids."""
        out = [self.zeros() for _ in range(4)]
        self.L.or_dust_fluid_get(self.ctx, *[ptr(a) for a in out])
        return dict(drho=out[0], dvr=out[1], dvt=out[2], stokes=out[3])

    def set_dust_fluid(self, drho=None, dvr=None, dvt=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (drho, dvr, dvt)]
        self.L.or_dust_fluid_set(self.ctx, *[ptr(a) for a in arrs])

    def set_visc1d(self, rmed=None, sigma=None, vrad=None):
        self.L.or_set_visc1d(self.ctx, *[None if a is None else ptr(np.ascontiguousarray(a, dtype=np.float64))
                                          for a in (rmed, sigma, vrad)])

    def visc1d(self):
        """(Rmed1D, Sigma1D, Vrad1D) with no step taken."""
        r, s, v = (np.zeros(self.p.nrad1d) for _ in range(3))
        self.L.or_get_visc1d(self.ctx, ptr(r), ptr(s), ptr(v))
        return r, s, v

    def sg_zero_mode(self, rho):
        out = np.zeros(self.nr)
        self.L.or_sg_zero_mode(self.ctx, ptr(rho), ptr(out))
        return out

    def set_time(self, t, t0=0.0, omega_frame=0.0):
        self.L.or_set_time(self.ctx, t, t0, omega_frame)

    def status(self):
        st = Status()
        self.L.or_get_status(self.ctx, ctypes.byref(st))
        return st

    def field(self, name):
        out = self.zeros()
        self.L.or_get_field(self.ctx, FIELDS[name], ptr(out))
        return out

    def fargo_shift(self):
        ns = (ctypes.c_long * self.nr)()
        vm = np.zeros(self.nr)
        self.L.or_get_fargo_shift(self.ctx, ns, ptr(vm))
        return np.array(ns[:], dtype=np.int64), vm

    def sg_accel(self, rho, direct=False):
        ar, at = self.zeros(), self.zeros()
        f = self.L.or_sg_accel_direct if direct else self.L.or_sg_accel
        f(self.ctx, ptr(np.ascontiguousarray(rho, dtype=np.float64)), ptr(ar), ptr(at))
        return ar, at

    # ---- dust particles ----
    def dust_init(self, dp, axi_sg_accr=None):
        """InitDustSystem restated: dict of particle arrays."""
        n = dp.nbpart
        a = {k: np.zeros(n + 1) for k in ("dustsize", "r", "th", "vr", "vth", "l")}
        rc = self.L.or_dust_init_system(ctypes.byref(self.p), ctypes.byref(dp), ptr(axi_sg_accr),
                                        ptr(a["dustsize"]), ptr(a["r"]), ptr(a["th"]), ptr(a["vr"]),
                                        ptr(a["vth"]), ptr(a["l"]))
        assert rc == 0
        return {k: v[:n].copy() for k, v in a.items()}

    def dust_attach(self, dp, parts, timestep):
        self.dp = dp.copy()
        c = {k: np.ascontiguousarray(parts[k], dtype=np.float64) for k in ("dustsize", "r", "th", "vr", "vth", "l")}
        self.L.or_dust_attach(self.ctx, ctypes.byref(self.dp), ptr(c["dustsize"]), ptr(c["r"]), ptr(c["th"]),
                              ptr(c["vr"]), ptr(c["vth"]), ptr(c["l"]), timestep)

    def dust_state(self):
        n = self.dp.nbpart
        a = {k: np.zeros(n) for k in ("r", "th", "vr", "vth", "l", "stokesnb")}
        self.L.or_dust_get(self.ctx, ptr(a["r"]), ptr(a["th"]), ptr(a["vr"]), ptr(a["vth"]), ptr(a["l"]),
                           ptr(a["stokesnb"]))
        return a

    def dust_gas_at_particles(self):
        n = self.dp.nbpart
        a = {k: np.zeros(n) for k in ("vrg", "vtg", "densg", "csg")}
        self.L.or_dust_get_gas_at_particles(self.ctx, ptr(a["vrg"]), ptr(a["vtg"]), ptr(a["densg"]), ptr(a["csg"]))
        return a

    def dust_interpolation(self, timestep):
        self.L.or_dust_interpolation(self.ctx, timestep)

    def dust_min_stopping_time(self):
        return self.L.or_dust_min_stopping_time(self.ctx)

    def dust_drag(self):
        n = self.dp.nbpart
        r, t = np.zeros(n), np.zeros(n)
        self.L.or_dust_get_drag(self.ctx, ptr(r), ptr(t))
        return r, t

    def dust_step(self, dt):
        """particle part of a sub-step over the frozen gas state"""
        self.L.or_dust_step(self.ctx, dt)


def orbital_elements(x, y, vx, vy, m):
    """FindOrbitalElements: e, a, M, V, PerihelionPA, lambda, varpi."""
    out = np.zeros(7)
    lib().or_orbital_elements(x, y, vx, vy, m, ptr(out))
    return out


# oracle/oracle.h: or_test_grid_id and or_test_stage_id
TEST_GRIDS = ["rho", "vr", "vt", "energy", "cs", "pres", "pot", "vradint", "vthetaint", "vradnew",
              "vthetanew", "energyint", "energynew", "qr", "qt", "divv", "drr", "dpp", "drp", "trr",
              "tpp", "trp", "vischeat", "indpot"]
TEST_STAGES = ["eos", "substep1", "substep2", "visc_terms", "visc_heating", "divergence", "substep3",
               "potential", "boundary", "floors", "crash"]


def stage_grid(o, name, values=None):
    """A work grid of the oracle context ((nrad+1)*nsec doubles): read, or
    overwritten with `values`."""
    buf = np.zeros((o.nr + 1) * o.ns) if values is None else np.ascontiguousarray(values, dtype=np.float64)
    rc = lib().or_test_grid(o.ctx, TEST_GRIDS.index(name), buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                            0 if values is None else 1)
    assert rc == 0
    return buf


def run_stage(o, name, dt=0.0):
    """One stage of the sub-step on the oracle context's grids."""
    return lib().or_test_stage(o.ctx, TEST_STAGES.index(name), dt)
