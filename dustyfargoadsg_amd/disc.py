"""Host-side mirror of the reference's hydro entry points over the C ABI.

A `Disc` owns one dfargo context: the reference's InitEuler / AlgoGas /
compute_selfgravity calls map to its methods, with the state as numpy FP64
arrays in the reference's polar layout (l = j + i*nsec, (nrad+1)*nsec).
Errors raise DfargoError with the library's message; there is no CPU
fallback, so a missing or unloadable library raises at construction.
"""
import ctypes

import numpy as np

from . import _capi
from ._capi import Params, Planets, Status, FIELDS


class DfargoError(RuntimeError):
    """A failed C-ABI call; `code` is its DFARGO_ERR_* return code."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


def _ptr(a):
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError("fields must be C-contiguous float64 arrays")
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(lib, rc):
    if rc != _capi.OK:
        raise DfargoError(f"dfargo error {rc}: {lib.dfargo_last_error().decode()}", rc)


def make_planets(params, spec, lib=None):
    """InitPlanetarySystem (src/Psys.c:96-196) from .cfg columns."""
    lib = lib or _capi.load()
    nb = len(spec["dist"])
    out = Planets()
    d = np.array(list(spec["dist"]) + [0.0], dtype=np.float64)
    m = np.array(list(spec["mass"]) + [0.0], dtype=np.float64)
    a = np.array(list(spec["acc"]) + [0.0], dtype=np.float64)
    fd = (ctypes.c_int * (nb + 1))(*(list(spec["feel_disk"]) + [0]))
    fo = (ctypes.c_int * (nb + 1))(*(list(spec["feel_others"]) + [0]))
    _check(lib, lib.dfargo_init_planetary_system(ctypes.byref(params), nb, _ptr(d), _ptr(m),
                                                  _ptr(a), fd, fo, ctypes.byref(out)))
    return out


def initial_fields(params, lib=None):
    """InitGasDensity and InitGasEnergy (src/Pframeforce.c:297-440)."""
    lib = lib or _capi.load()
    n = (params.nrad + 1) * params.nsec
    rho, e = np.zeros(n), np.zeros(n)
    _check(lib, lib.dfargo_init_gas_density(ctypes.byref(params), _ptr(rho)))
    if params.energy_equation:
        _check(lib, lib.dfargo_init_gas_energy(ctypes.byref(params), _ptr(e)))
    return rho, e


class Disc:
    """One disc on one GPU."""

    def __init__(self, params, precision=_capi.FP64, device=0, lib_path=None):
        self.lib = _capi.load(lib_path)
        self.p = params.copy()
        self.nr, self.ns = params.nrad, params.nsec
        self.n = (self.nr + 1) * self.ns
        self.precision = precision
        h = ctypes.c_void_p()
        _check(self.lib, self.lib.dfargo_create(ctypes.byref(self.p), precision, device, ctypes.byref(h)))
        self.ctx = h

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.dfargo_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def zeros(self):
        return np.zeros(self.n, dtype=np.float64)

    # -- reference entry points ------------------------------------------
    def init_euler(self, rho, energy=None, planets=None):
        """InitEuler with the start-up self-gravity call; returns (vr, vt)."""
        vr, vt = self.zeros(), self.zeros()
        _check(self.lib, self.lib.dfargo_init_euler(
            self.ctx, _ptr(np.ascontiguousarray(rho, dtype=np.float64)), _ptr(vr), _ptr(vt),
            _ptr(None if energy is None else np.ascontiguousarray(energy, dtype=np.float64)),
            ctypes.byref(planets) if planets is not None else None))
        return vr, vt

    def initial_state(self, planet_spec=None):
        rho, e = initial_fields(self.p, self.lib)
        sys_ = make_planets(self.p, planet_spec or dict(dist=[], mass=[], acc=[], feel_disk=[],
                                                        feel_others=[]), self.lib)
        vr, vt = self.init_euler(rho, e if self.p.energy_equation else None, sys_)
        return dict(rho=rho, vr=vr, vt=vt, energy=e, planets=sys_)

    def algogas(self, DT, state=None):
        """AlgoGas: with `state` the host arrays are uploaded and updated in
        place (the reference's calling convention); without it the device
        state is advanced with no transfer."""
        if state is None:
            _check(self.lib, self.lib.dfargo_algogas_device(self.ctx, DT))
            return
        e = state.get("energy")
        _check(self.lib, self.lib.dfargo_algogas(
            self.ctx, DT, _ptr(state["rho"]), _ptr(state["vr"]), _ptr(state["vt"]),
            _ptr(e if (e is not None and self.p.energy_equation) else None),
            ctypes.byref(state["planets"]) if state.get("planets") is not None else None))

    def run_substeps(self, DT, nsub):
        """Exactly nsub sub-steps of an AlgoGas call of length DT."""
        _check(self.lib, self.lib.dfargo_run_substeps(self.ctx, DT, nsub))

    def set_state(self, st):
        """Upload the fields and planets of `st`; with DAMPTOVISCOUS, an
        entry "visc1d" = (Rmed1D, Sigma1D, Vrad1D) also sets the 1-D disc."""
        e = st.get("energy")
        _check(self.lib, self.lib.dfargo_upload_state(
            self.ctx, _ptr(st["rho"]), _ptr(st["vr"]), _ptr(st["vt"]),
            _ptr(e if (e is not None and self.p.energy_equation) else None),
            ctypes.byref(st["planets"]) if st.get("planets") is not None else None))
        if st.get("visc1d") is not None:
            self.set_visc1d(*st["visc1d"])

    def state(self):
        """The fields and planets; with DAMPTOVISCOUS also "visc1d" =
        (Rmed1D, Sigma1D, Vrad1D)."""
        rho, vr, vt, e = self.zeros(), self.zeros(), self.zeros(), self.zeros()
        sys_ = Planets()
        _check(self.lib, self.lib.dfargo_download_state(self.ctx, _ptr(rho), _ptr(vr), _ptr(vt),
                                                        _ptr(e), ctypes.byref(sys_)))
        out = dict(rho=rho, vr=vr, vt=vt, energy=e, planets=sys_)
        if self.p.damp_to == _capi.DAMP_TO_VISCOUS:
            out["visc1d"] = self.visc1d(rmed=True)
        return out

    def status(self):
        st = Status()
        _check(self.lib, self.lib.dfargo_get_status(self.ctx, ctypes.byref(st)))
        return st

    def set_time(self, t, t0=0.0, omega_frame=0.0):
        _check(self.lib, self.lib.dfargo_set_time(self.ctx, t, t0, omega_frame))

    def field(self, name):
        out = self.zeros()
        _check(self.lib, self.lib.dfargo_get_field(self.ctx, FIELDS[name], _ptr(out)))
        return out

    def compute_selfgravity(self):
        _check(self.lib, self.lib.dfargo_compute_selfgravity(self.ctx))
        return self.field("sg_accr"), self.field("sg_acct")

    def fargo_shift(self):
        ns = (ctypes.c_long * self.nr)()
        vm = np.zeros(self.nr)
        _check(self.lib, self.lib.dfargo_get_fargo_shift(self.ctx, ns, _ptr(vm)))
        return np.array(ns[:], dtype=np.int64), vm

    def visc1d(self, rmed=False):
        """DAMPTOVISCOUS's 1-D viscous disc: (Sigma, vr), NRAD1D values each,
        or (Rmed1D, Sigma, vr) with rmed=True; empty when the option is off."""
        n = max(self.p.nrad1d, 1)
        r, s, v, got = np.zeros(n), np.zeros(n), np.zeros(n), ctypes.c_int()
        _check(self.lib, self.lib.dfargo_get_visc1d(self.ctx, _ptr(r), _ptr(s), _ptr(v), ctypes.byref(got)))
        m = got.value
        return (r[:m], s[:m], v[:m]) if rmed else (s[:m], v[:m])

    def set_visc1d(self, rmed=None, sigma=None, vrad=None):
        """Replace any of the 1-D disc's Rmed1D, Sigma1D, Vrad1D (the restart
        branch of src/SideEuler.c:1489-1502 replaces the first two)."""
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (rmed, sigma, vrad)]
        n = next((len(a) for a in arrs if a is not None), self.p.nrad1d)
        if any(a is not None and len(a) != n for a in arrs):
            raise ValueError("the 1-D profiles differ in length")
        _check(self.lib, self.lib.dfargo_set_visc1d(self.ctx, *[_ptr(a) for a in arrs], n))

    def gas_budgets(self):
        """GasTotalMass, GasMomentum, GasTotalEnergy, planets' angular momentum."""
        out = np.zeros(4)
        _check(self.lib, self.lib.dfargo_gas_budgets(self.ctx, _ptr(out)))
        return out

    def update_log(self):
        """UpdateLog: (nb, 8) torques and powers, (nb,) indirect torques."""
        tq, ind, nb = np.zeros(8 * 8), np.zeros(8), ctypes.c_int()
        _check(self.lib, self.lib.dfargo_update_log(self.ctx, 8, _ptr(tq), _ptr(ind), ctypes.byref(nb)))
        return tq[: 8 * nb.value].reshape(nb.value, 8), ind[: nb.value]

    def fviscosity(self, r):
        return self.lib.dfargo_fviscosity(self.ctx, r)

    # -- dust particles (src/Dsys.c, src/DustUpdate.c) ---------------------
    def attach_dust_fluid(self, fp, drho=None):
        """The dust fluid (DUSTFLUID YES) from its initial surface density
        (by default InitDustDensity's); call before initial_state /
        init_euler, which set its Stokes numbers and velocities."""
        if drho is None:
            drho = init_dust_fluid_density(self.p, fp, self.lib)
        drho = np.ascontiguousarray(drho, dtype=np.float64)
        _check(self.lib, self.lib.dfargo_attach_dust_fluid(self.ctx, ctypes.byref(fp), _ptr(drho)))

    def dust_fluid(self):
        """The fluid's DRho, DVrad, DVtheta and the Stokes grid of the last
        sub-step ((nrad + 1) * nsec values each)."""
        out = [self.zeros() for _ in range(4)]
        _check(self.lib, self.lib.dfargo_download_dust_fluid(self.ctx, *[_ptr(a) for a in out]))
        return dict(drho=out[0], dvr=out[1], dvt=out[2], stokes=out[3])

    def set_dust_fluid(self, drho=None, dvr=None, dvt=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (drho, dvr, dvt)]
        _check(self.lib, self.lib.dfargo_upload_dust_fluid(self.ctx, *[_ptr(a) for a in arrs]))

    def attach_dust(self, d, parts, timestep):
        """Particles from a dict of nbpart-long arrays (r, th, vr, vth, l,
        dustsize) and the start-up interpolation with `timestep`."""
        n = d.nbpart
        arrs = {k: np.ascontiguousarray(parts.get(k, parts["dustsize"] if k == "dustsize_init"
                                                  else np.zeros(n)), dtype=np.float64)
                for k in _capi.DUST_ARRAYS}
        self._dust_keep = arrs
        self.dust_n = n
        _check(self.lib, self.lib.dfargo_attach_dust(self.ctx, ctypes.byref(d), ctypes.byref(_capi.dust_state(arrs)),
                                                     timestep))

    def dust_state(self):
        """The particle arrays (r, th, vr, vth, l, dustsize, dustsize_init,
        stokesnb) as a dict of float64 arrays."""
        arrs = {k: np.zeros(self.dust_n) for k in _capi.DUST_ARRAYS}
        _check(self.lib, self.lib.dfargo_download_dust(self.ctx, ctypes.byref(_capi.dust_state(arrs))))
        return arrs

    def set_dust_state(self, parts):
        arrs = {k: np.ascontiguousarray(parts[k], dtype=np.float64) for k in _capi.DUST_ARRAYS if k in parts}
        st = _capi.DustState()
        for k, v in arrs.items():
            setattr(st, k, v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        _check(self.lib, self.lib.dfargo_upload_dust(self.ctx, ctypes.byref(st)))

    # -- launch shapes --------------------------------------------------------
    def tuning(self):
        t = _capi.Tuning()
        _check(self.lib, self.lib.dfargo_get_tuning(self.ctx, ctypes.byref(t)))
        return dict(xcd_remap=t.xcd_remap, azi_waves=t.azi_waves, finish_block=t.finish_block,
                    sg_fft_buffers=t.sg_fft_buffers, sg_rows_pair=t.sg_rows_pair)

    def set_tuning(self, **kw):
        t = _capi.Tuning(**dict(self.tuning(), **kw))
        _check(self.lib, self.lib.dfargo_set_tuning(self.ctx, ctypes.byref(t)))

    def autotune(self, substeps=5):
        """Times the launch shapes on a scratch disc of the same parameters
        and applies the fastest (kept for later discs of this shape)."""
        t = _capi.Tuning()
        _check(self.lib, self.lib.dfargo_autotune(self.ctx, substeps, ctypes.byref(t)))
        return dict(xcd_remap=t.xcd_remap, azi_waves=t.azi_waves, finish_block=t.finish_block,
                    sg_fft_buffers=t.sg_fft_buffers, sg_rows_pair=t.sg_rows_pair)

    # -- measurement --------------------------------------------------------
    def set_profiling(self, on):
        _check(self.lib, self.lib.dfargo_set_profiling(self.ctx, 1 if on else 0))

    def kernel_times(self):
        maxn = 64
        names = (ctypes.c_char * 48 * maxn)()
        tot = np.zeros(maxn)
        cnt = (ctypes.c_long * maxn)()
        n = ctypes.c_int()
        _check(self.lib, self.lib.dfargo_get_kernel_times(self.ctx, maxn, names, _ptr(tot), cnt,
                                                          ctypes.byref(n)))
        return {bytes(names[k]).split(b"\0")[0].decode(): (tot[k], cnt[k]) for k in range(n.value)}

    def stream(self):
        return self.lib.dfargo_stream(self.ctx)


class Batch:
    """Discs advanced together: one graph launch advances every unfinished
    disc by one sub-step, each with its own CFL step."""

    def __init__(self, discs):
        self.discs = list(discs)
        self.lib = self.discs[0].lib
        arr = (ctypes.c_void_p * len(self.discs))(*[d.ctx for d in self.discs])
        h = ctypes.c_void_p()
        _check(self.lib, self.lib.dfargo_batch_create(arr, len(self.discs), ctypes.byref(h)))
        self.handle = h

    GRAPH, STREAMS = 0, 1

    def set_mode(self, mode):
        """GRAPH: one graph of all discs per sub-step; STREAMS: each disc's
        graph on its own stream as often as it needs (the default)."""
        _check(self.lib, self.lib.dfargo_batch_set_mode(self.handle, mode))

    def algogas(self, DT):
        """One AlgoGas call per disc, of length DT (a number or one per disc)."""
        dts = np.ascontiguousarray(np.broadcast_to(np.asarray(DT, dtype=np.float64), (len(self.discs),)))
        _check(self.lib, self.lib.dfargo_batch_algogas(self.handle, _ptr(dts)))

    def close(self):
        if getattr(self, "handle", None):
            self.lib.dfargo_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def init_dust_fluid_density(params, fp, lib=None):
    """InitDustDensity (src/Pframeforce.c:344-380) on the host, no noise:
    DSigma(Rmed) on every cell, ring nrad zero."""
    lib = lib or _capi.load()
    drho = np.zeros((params.nrad + 1) * params.nsec)
    _check(lib, lib.dfargo_init_dust_fluid_density(ctypes.byref(params), ctypes.byref(fp), _ptr(drho)))
    return drho


def init_dust_system(params, d, axi_sg_accr=None, lib=None):
    """InitDustSystem (src/Dsys.c:128-216) on the host: dict of particle arrays."""
    lib = lib or _capi.load()
    n = d.nbpart
    arrs = {k: np.zeros(max(n, 1)) for k in _capi.DUST_ARRAYS}
    _check(lib, lib.dfargo_init_dust_system(ctypes.byref(params), ctypes.byref(d), _ptr(axi_sg_accr),
                                            ctypes.byref(_capi.dust_state(arrs))))
    return {k: v[:n].copy() for k, v in arrs.items()}


def read_par(path, lib=None):
    """A .par file as (dfargo_params, dfargo_run_params): ReadVariables and
    its decoding (src/Interpret.c:82-476), then the parameter checks."""
    lib = lib or _capi.load()
    p, run = _capi.Params(), _capi.RunParams()
    _check(lib, lib.dfargo_read_par(str(path).encode(), ctypes.byref(p), ctypes.byref(run)))
    return p, run


def read_dust_fluid_par(path, lib=None):
    """The dust-fluid switches of a .par file: (DustFluidParams, DUSTFLUID on)."""
    lib = lib or _capi.load()
    f, on = _capi.DustFluidParams(), ctypes.c_int()
    _check(lib, lib.dfargo_read_dust_fluid_par(str(path).encode(), ctypes.byref(f), ctypes.byref(on)))
    return f, bool(on.value)


def read_planet_config(path, params, lib=None):
    """A planet .cfg file as the initial planetary system
    (InitPlanetarySystem, src/Psys.c:96-196)."""
    lib = lib or _capi.load()
    sys_ = _capi.Planets()
    _check(lib, lib.dfargo_read_planet_config(str(path).encode(), ctypes.byref(params), ctypes.byref(sys_)))
    return sys_


def orbital_elements(x, y, vx, vy, m, lib=None):
    """FindOrbitalElements: e, a, M, V, PerihelionPA, lambda, varpi."""
    lib = lib or _capi.load()
    out = np.zeros(7)
    _check(lib, lib.dfargo_orbital_elements(x, y, vx, vy, m, _ptr(out)))
    return out
