"""ctypes mirror of include/dfargo.h and the loader of the HIP library.

The structures below must match include/dfargo.h field for field; a test
checks their sizes against the C compiler's.
"""
import ctypes
import os

MAX_PLANETS = 8
_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdfargo.so")

OK, ERR_ARG, ERR_UNSUPPORTED, ERR_HIP, ERR_ALLOC, ERR_STATE = range(6)
FP64, FP32 = 0, 1
BC_WALL, BC_OPEN, BC_NONREFLECTING = 0, 1, 2
FRAME_FIXED, FRAME_COROTATING, FRAME_GUIDING_CENTER, FRAME_BINARY_CENTER = 0, 1, 2, 3
DAMP_TO_AXI, DAMP_TO_INI, DAMP_TO_VISCOUS = 0, 1, 2

FIELDS = {
    "potential": 0, "soundspeed": 1, "pressure": 2, "temperature": 3,
    "divv": 4, "vischeat": 5, "sg_accr": 6, "sg_acct": 7,
    "taurr": 8, "taurp": 9, "taupp": 10,
    "dust_density": 11, "fb_accr": 12, "fb_acct": 13, "fb_edot": 14,
}
INTERP_NGP, INTERP_CIC, INTERP_TSC = 0, 1, 2

c_int, c_long, c_double = ctypes.c_int, ctypes.c_long, ctypes.c_double
_D8 = c_double * MAX_PLANETS
_I8 = c_int * MAX_PLANETS


class Params(ctypes.Structure):
    """dfargo_params (src/var.c:15-202 as decoded by src/Interpret.c)."""
    _fields_ = [
        ("nrad", c_int), ("nsec", c_int),
        ("rmin", c_double), ("rmax", c_double),
        ("log_grid", c_int),
        ("pmin", c_double), ("pmax", c_double),
        ("sigma0", c_double), ("sigma_slope", c_double),
        ("aspect_ratio", c_double), ("flaring_index", c_double),
        ("viscosity", c_double), ("alpha_viscosity", c_double),
        ("cavity_radius", c_double), ("cavity_ratio", c_double), ("cavity_width", c_double),
        ("transition_radius", c_double), ("transition_ratio", c_double),
        ("transition_width", c_double),
        ("imposed_disk_drift", c_double),
        ("add_noise", c_int), ("noise_seed", c_long),
        ("energy_equation", c_int), ("viscous_heating", c_int), ("beta_cooling", c_int),
        ("adiabatic_index", c_double), ("beta_cooling_time", c_double),
        ("beta_cooling_slope", c_double),
        ("self_gravity", c_int), ("sg_thickness_smoothing", c_double),
        ("thickness_smoothing", c_double), ("roche_smoothing", c_double),
        ("indirect_term_smoothing", c_double), ("mass_taper", c_double),
        ("indirect_term", c_int), ("discard_gas_indirect_term", c_int),
        ("alternative_gas_indirect_term", c_int),
        ("exclude_hill", c_int), ("exclude_hill_factor", c_double),
        ("forced_circular", c_int),
        ("release_date", c_double), ("release_radius", c_double),
        ("eccentricity", c_double),
        ("corotating", c_int), ("omega_frame", c_double),
        ("fast_transport", c_int), ("add_floors", c_int), ("density_floor", c_double),
        ("boundary", c_int),
        ("corotate_with_outer_planet", c_int),
        ("wave_killing_zones", c_int), ("wkz_rmin", c_double), ("wkz_rmax", c_double),
        ("damping_ratio", c_double), ("damp_to", c_int), ("zero_vrad_bc", c_int),
        ("sg_zero_mode", c_int),
        ("nrad1d", c_int), ("rmin1d", c_double), ("rmax1d", c_double), ("bc1d", c_int),
    ]

    def copy(self):
        q = Params()
        ctypes.pointer(q)[0] = self
        return q

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class RunParams(ctypes.Structure):
    """dfargo_run_params: DT, NTOT, NINTERM, PLANETCONFIG, OUTPUTDIR."""
    _fields_ = [("dt", c_double), ("ntot", c_int), ("ninterm", c_int),
                ("planet_config", ctypes.c_char * 512), ("output_dir", ctypes.c_char * 512)]


class Planets(ctypes.Structure):
    """dfargo_planets (PlanetarySystem, src/types.h:97-111)."""
    _fields_ = [
        ("nb", c_int),
        ("mass", _D8), ("x", _D8), ("y", _D8), ("vx", _D8), ("vy", _D8), ("acc", _D8),
        ("feel_disk", _I8), ("feel_others", _I8), ("binary", _I8),
        ("initial_mass", _D8), ("final_mass", _D8),
    ]


class Status(ctypes.Structure):
    """dfargo_status: the scalar globals AlgoGas leaves host-coherent."""
    _fields_ = [
        ("physical_time", c_double), ("physical_time_initial", c_double),
        ("omega_frame", c_double), ("mindt", c_double), ("maxdt", c_double),
        ("lost_mass", c_double), ("mass_taper", c_double),
        ("disk_on_primary_x", c_double), ("disk_on_primary_y", c_double),
        ("disk_on_primary_alt_x", c_double), ("disk_on_primary_alt_y", c_double),
        ("bary_x", c_double * 3), ("bary_y", c_double * 3),
        ("substeps_total", c_long), ("last_call_substeps", c_int),
        ("gas_timesteps_cfl", c_int), ("crashed_density", c_int),
        ("crashed_energy", c_int), ("crashed", c_int), ("fargo_margin", c_double),
    ]


class Tuning(ctypes.Structure):
    """dfargo_tuning: launch shapes that do not change results."""
    _fields_ = [("xcd_remap", c_int), ("azi_waves", c_int), ("finish_block", c_int), ("sg_fft_buffers", c_int),
                ("sg_rows_pair", c_int)]


class DustParams(ctypes.Structure):
    """dfargo_dust_params (the particle switches of src/var.c:136-181)."""
    _fields_ = [
        ("nbpart", c_int),
        ("sizeminpart", c_double), ("sizemaxpart", c_double), ("sizepartslope", c_double),
        ("rhopart", c_double), ("dustslope", c_double),
        ("rmindust", c_double), ("rmaxdust", c_double),
        ("interpolation", c_int),
        ("dust_feel_disk", c_int), ("dust_feel_sg", c_int), ("dust_feel_planets", c_int),
        ("dust_feedback", c_int), ("short_friction_time_approximation", c_int),
        ("no_dt_constraint_by_particles", c_int), ("dust_timestep", c_double),
        ("factor_unit_mass", c_double), ("factor_unit_length", c_double),
        ("particle_mass", c_double), ("seed", c_long),
        ("dust_feel_turb", c_int), ("dust_growth", c_int), ("remove_dust_from_hill", c_int),
        ("zz_integrator", c_int), ("dust_feel_sg_zero_mode", c_int),
    ]

    def copy(self):
        q = DustParams()
        ctypes.pointer(q)[0] = self
        return q


DUST_SG_NO, DUST_SG_YES, DUST_SG_ZERO_MODE = 0, 1, 2


class DustFluidParams(ctypes.Structure):
    """dfargo_dust_fluid_params (the dust-fluid switches of src/var.c:136-186)."""
    _fields_ = [
        ("sizepart", c_double), ("rhopart", c_double),
        ("factor_unit_mass", c_double), ("factor_unit_length", c_double),
        ("dust_to_gas_density_ratio", c_double),
        ("dviscosity", c_double), ("dalpha_viscosity", c_double),
        ("daspect_ratio", c_double), ("dflaring_index", c_double), ("dust_diff_inn", c_double),
        ("sfta", c_int), ("dust_feel_disk", c_int), ("dust_feedback", c_int),
        ("dust_feel_sg", c_int), ("dust_diffusion", c_int), ("open_inner_dust", c_int),
    ]


def dust_fluid_default_params(**overrides):
    """dfargo_dust_fluid_default_params() restated (src/var.c defaults)."""
    f = DustFluidParams()
    f.sizepart, f.rhopart = 1e-3, 1.0
    f.factor_unit_mass, f.factor_unit_length = 1.0, 1.0
    f.dust_to_gas_density_ratio = 0.01
    f.daspect_ratio = 0.05
    f.sfta, f.dust_feel_disk, f.dust_feel_sg = 1, 1, DUST_SG_YES
    for k, v in overrides.items():
        if not hasattr(f, k):
            raise KeyError(f"unknown dust-fluid parameter {k!r}")
        setattr(f, k, v)
    return f


class DustState(ctypes.Structure):
    """dfargo_dust_state: pointers to the nbpart-long particle arrays."""
    _fields_ = [(n, ctypes.POINTER(c_double)) for n in
                ("r", "th", "vr", "vth", "l", "dustsize", "dustsize_init", "stokesnb")]


DUST_ARRAYS = ("r", "th", "vr", "vth", "l", "dustsize", "dustsize_init", "stokesnb")


def dust_state(arrays):
    """DustState pointing at a dict of float64 numpy arrays (kept alive by the caller)."""
    st = DustState()
    for n in DUST_ARRAYS:
        setattr(st, n, arrays[n].ctypes.data_as(ctypes.POINTER(c_double)))
    return st


def dust_default_params(nbpart=0, **overrides):
    """dfargo_dust_default_params() restated (src/var.c:136-181 defaults)."""
    d = DustParams()
    d.nbpart = nbpart
    d.sizeminpart, d.sizemaxpart, d.sizepartslope = 0.001, 0.001, 3.0
    d.rhopart, d.dustslope = 1.0, 0.5
    d.rmindust, d.rmaxdust = 0.0, 0.0
    d.interpolation = INTERP_TSC
    d.dust_feel_disk, d.dust_feel_sg, d.dust_feel_planets = 1, 1, 1
    d.dust_feedback, d.short_friction_time_approximation = 0, 1
    d.no_dt_constraint_by_particles, d.dust_timestep = 1, 0.0
    d.factor_unit_mass, d.factor_unit_length = 1.0, 1.0
    d.particle_mass, d.seed = 0.0, 0
    for k, v in overrides.items():
        if not hasattr(d, k):
            raise KeyError(f"unknown dust parameter {k!r}")
        setattr(d, k, v)
    return d


def default_params(**overrides):
    """dfargo_default_params() restated in Python so that host-only code and
    CPU tests need no library: the src/var.c:15-202 defaults."""
    p = Params()
    p.nrad, p.nsec = 64, 64
    p.rmin, p.rmax = 1.0, 1.0
    p.log_grid = 0
    p.pmin, p.pmax = 0.0, 6.2831853071795864
    p.sigma0, p.sigma_slope = 173.0, 0.0
    p.aspect_ratio, p.flaring_index = 0.05, 0.0
    p.viscosity, p.alpha_viscosity = 0.0, 0.0
    p.cavity_radius, p.cavity_ratio, p.cavity_width = 0.0, 1.0, 1.0
    p.transition_radius, p.transition_ratio, p.transition_width = 0.0, 1.0, 1.0
    p.imposed_disk_drift = 0.0
    p.add_noise, p.noise_seed = 0, 0
    p.energy_equation, p.viscous_heating, p.beta_cooling = 0, 1, 0
    p.adiabatic_index, p.beta_cooling_time, p.beta_cooling_slope = 1.4, 10.0, 0.0
    p.self_gravity, p.sg_thickness_smoothing = 0, 0.3
    p.thickness_smoothing, p.roche_smoothing = 0.3, 0.0
    p.indirect_term_smoothing, p.mass_taper = 0.0, 0.0000001
    p.indirect_term, p.discard_gas_indirect_term = 1, 0
    p.alternative_gas_indirect_term = 0
    p.exclude_hill, p.exclude_hill_factor = 0, 1.0
    p.forced_circular = 0
    p.release_date, p.release_radius = 0.0, 0.0
    p.eccentricity = 0.0
    p.corotating, p.omega_frame = 0, 0.0
    p.fast_transport, p.add_floors, p.density_floor = 1, 1, 1e-9
    p.boundary = BC_WALL
    p.corotate_with_outer_planet = 0
    p.wave_killing_zones, p.wkz_rmin, p.wkz_rmax = 0, 0.0, 0.0
    p.damping_ratio, p.damp_to, p.zero_vrad_bc = 0.3, DAMP_TO_AXI, 0
    p.sg_zero_mode = 0
    p.nrad1d, p.rmin1d, p.rmax1d, p.bc1d = 100, 0.01, 100.0, 0
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(f"unknown parameter {k!r}")
        setattr(p, k, v)
    return p


_lib = None


def load(path=None):
    """Loads libdfargo.so.  Raises if it is missing: there is no fallback."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{p} is missing: build it with __graft_entry__.build() "
            "(the hydro step has no CPU fallback)")
    lib = ctypes.CDLL(p)
    P, PP = ctypes.POINTER(Params), ctypes.POINTER(Planets)
    dp, vp = ctypes.POINTER(c_double), ctypes.c_void_p
    fp = ctypes.POINTER(ctypes.c_float)
    sig = {
        "dfargo_default_params": (None, [P]),
        "dfargo_check_params": (c_int, [P]),
        "dfargo_fill_polar_1d_arrays": (c_int, [P, dp, dp, dp, dp]),
        "dfargo_init_gas_density": (c_int, [P, dp]),
        "dfargo_init_gas_energy": (c_int, [P, dp]),
        "dfargo_init_planetary_system": (c_int, [P, c_int, dp, dp, dp,
                                                 ctypes.POINTER(c_int), ctypes.POINTER(c_int), PP]),
        "dfargo_aspect_ratio": (c_double, [P, c_double]),
        "dfargo_create": (c_int, [P, c_int, c_int, ctypes.POINTER(vp)]),
        "dfargo_destroy": (c_int, [vp]),
        "dfargo_last_error": (ctypes.c_char_p, []),
        "dfargo_init_euler": (c_int, [vp, dp, dp, dp, dp, PP]),
        "dfargo_algogas": (c_int, [vp, c_double, dp, dp, dp, dp, PP]),
        "dfargo_algogas_device": (c_int, [vp, c_double]),
        "dfargo_run_substeps": (c_int, [vp, c_double, c_int]),
        "dfargo_upload_state": (c_int, [vp, dp, dp, dp, dp, PP]),
        "dfargo_download_state": (c_int, [vp, dp, dp, dp, dp, PP]),
        "dfargo_get_status": (c_int, [vp, ctypes.POINTER(Status)]),
        "dfargo_set_time": (c_int, [vp, c_double, c_double, c_double]),
        "dfargo_get_field": (c_int, [vp, c_int, dp]),
        "dfargo_compute_selfgravity": (c_int, [vp]),
        "dfargo_fviscosity": (c_double, [vp, c_double]),
        "dfargo_get_fargo_shift": (c_int, [vp, ctypes.POINTER(c_long), dp]),
        "dfargo_get_visc1d": (c_int, [vp, dp, dp, dp, ctypes.POINTER(c_int)]),
        "dfargo_set_visc1d": (c_int, [vp, dp, dp, dp, c_int]),
        "dfargo_set_profiling": (c_int, [vp, c_int]),
        "dfargo_get_kernel_times": (c_int, [vp, c_int, ctypes.c_void_p, dp,
                                            ctypes.POINTER(c_long), ctypes.POINTER(c_int)]),
        "dfargo_stream": (vp, [vp]),
        "dfargo_dust_default_params": (None, [ctypes.POINTER(DustParams), c_int]),
        "dfargo_check_dust_params": (c_int, [P, ctypes.POINTER(DustParams)]),
        "dfargo_init_dust_system": (c_int, [P, ctypes.POINTER(DustParams), dp, ctypes.POINTER(DustState)]),
        "dfargo_attach_dust": (c_int, [vp, ctypes.POINTER(DustParams), ctypes.POINTER(DustState), c_double]),
        "dfargo_download_dust": (c_int, [vp, ctypes.POINTER(DustState)]),
        "dfargo_dust_fluid_default_params": (None, [ctypes.POINTER(DustFluidParams)]),
        "dfargo_check_dust_fluid_params": (c_int, [P, ctypes.POINTER(DustFluidParams), c_int]),
        "dfargo_init_dust_fluid_density": (c_int, [P, ctypes.POINTER(DustFluidParams), dp]),
        "dfargo_attach_dust_fluid": (c_int, [vp, ctypes.POINTER(DustFluidParams), dp]),
        "dfargo_download_dust_fluid": (c_int, [vp, dp, dp, dp, dp]),
        "dfargo_upload_dust_fluid": (c_int, [vp, dp, dp, dp]),
        "dfargo_read_dust_fluid_par": (c_int, [ctypes.c_char_p, ctypes.POINTER(DustFluidParams),
                                               ctypes.POINTER(c_int)]),
        "dfargo_upload_dust": (c_int, [vp, ctypes.POINTER(DustState)]),
        "dfargo_gas_budgets": (c_int, [vp, dp]),
        "dfargo_update_log": (c_int, [vp, c_int, dp, dp, ctypes.POINTER(c_int)]),
        "dfargo_update_log_noalloc": (c_int, [vp, c_int, dp, dp, ctypes.POINTER(c_int)]),
        "dfargo_orbital_elements": (c_int, [c_double] * 5 + [dp]),
        "dfargo_read_par": (c_int, [ctypes.c_char_p, P, ctypes.POINTER(RunParams)]),
        "dfargo_get_tuning": (c_int, [vp, ctypes.POINTER(Tuning)]),
        "dfargo_set_tuning": (c_int, [vp, ctypes.POINTER(Tuning)]),
        "dfargo_autotune": (c_int, [vp, c_int, ctypes.POINTER(Tuning)]),
        "dfargo_batch_create": (c_int, [ctypes.POINTER(vp), c_int, ctypes.POINTER(vp)]),
        "dfargo_batch_algogas": (c_int, [vp, dp]),
        "dfargo_batch_destroy": (c_int, [vp]),
        "dfargo_batch_set_mode": (c_int, [vp, c_int]),
        "dfargo_mlp_create": (c_int, [c_int, c_int, c_int, c_int, c_int, ctypes.c_ulong, c_int,
                                      ctypes.POINTER(vp)]),
        "dfargo_mlp_destroy": (c_int, [vp]),
        "dfargo_mlp_param_count": (ctypes.c_size_t, [vp]),
        "dfargo_mlp_get_params": (c_int, [vp, fp]),
        "dfargo_mlp_set_params": (c_int, [vp, fp]),
        "dfargo_mlp_forward": (c_int, [vp, c_int, fp, fp]),
        "dfargo_mlp_gradients": (c_int, [vp, c_int, fp, fp, fp, fp]),
        "dfargo_mlp_train": (c_int, [vp, c_int, fp, fp, c_int, ctypes.c_float, fp]),
        "dfargo_read_planet_config": (c_int, [ctypes.c_char_p, P, PP]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


EXPORTED = [
    "dfargo_default_params", "dfargo_check_params", "dfargo_fill_polar_1d_arrays",
    "dfargo_init_gas_density", "dfargo_init_gas_energy", "dfargo_init_planetary_system",
    "dfargo_aspect_ratio", "dfargo_create", "dfargo_destroy", "dfargo_last_error",
    "dfargo_init_euler", "dfargo_algogas", "dfargo_algogas_device", "dfargo_run_substeps", "dfargo_upload_state",
    "dfargo_download_state", "dfargo_get_status", "dfargo_set_time", "dfargo_get_field",
    "dfargo_compute_selfgravity", "dfargo_fviscosity", "dfargo_get_fargo_shift",
    "dfargo_get_visc1d", "dfargo_set_visc1d", "dfargo_set_profiling", "dfargo_get_kernel_times", "dfargo_stream",
    "dfargo_dust_default_params", "dfargo_check_dust_params", "dfargo_init_dust_system",
    "dfargo_attach_dust", "dfargo_download_dust", "dfargo_upload_dust",
    "dfargo_dust_fluid_default_params", "dfargo_check_dust_fluid_params", "dfargo_init_dust_fluid_density",
    "dfargo_attach_dust_fluid", "dfargo_download_dust_fluid", "dfargo_upload_dust_fluid",
    "dfargo_read_dust_fluid_par",
    "dfargo_gas_budgets", "dfargo_update_log", "dfargo_update_log_noalloc", "dfargo_orbital_elements",
    "dfargo_read_par", "dfargo_read_planet_config",
    "dfargo_get_tuning", "dfargo_set_tuning", "dfargo_autotune",
    "dfargo_batch_create", "dfargo_batch_algogas", "dfargo_batch_destroy", "dfargo_batch_set_mode",
    "dfargo_mlp_create", "dfargo_mlp_destroy", "dfargo_mlp_param_count", "dfargo_mlp_get_params",
    "dfargo_mlp_set_params", "dfargo_mlp_forward", "dfargo_mlp_gradients", "dfargo_mlp_train",
]
