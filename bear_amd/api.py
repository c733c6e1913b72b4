"""Python binding of the C ABI in include/bear_hip.h (ctypes).

Mirrors the pyBeAR surface that the likelihood path uses
(python/src/bear_python_wrapper.cpp:27-373 of the reference):
``GlobalConfig``, ``ObservationInput``, ``Retrieval.computeLikelihood``,
``Retrieval.computeModel`` (returning a ``ForwardModelOutput``) and
``ForwardModel.calcModelGPU`` on device buffers.  The library is the in-tree
``bear_amd/libbear_hip.so``; importing this module on a machine without it
raises, there is no fallback path.
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbear_hip.so")

_ERRORS = {-1: "InvalidInput", -2: "FileNotFound", -3: "DeviceError", -4: "OutOfMemory", -5: "InternalError"}


class BearError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_ERRORS.get(code, code)}: {msg}")
        self.code = code


class _GlobalConfig(C.Structure):
    _fields_ = [("use_gpu", C.c_int), ("device", C.c_int), ("retrieval_folder_path", C.c_char_p),
                ("cross_section_file_path", C.c_char_p), ("wavenumber_file_path", C.c_char_p),
                ("spectral_discretisation", C.c_int), ("spectral_resolution", C.c_double),
                ("use_error_inflation", C.c_int), ("table_storage", C.c_int)]


class _ModelConfig(C.Structure):
    _fields_ = [("nb_grid_points", C.c_size_t), ("atmos_bottom_pressure", C.c_double),
                ("atmos_top_pressure", C.c_double), ("temperature_profile", C.c_char_p),
                ("radiative_transfer", C.c_char_p), ("chemistry", C.c_char_p), ("opacity_species", C.c_char_p),
                ("fit_mean_molecular_weight", C.c_int), ("fit_scale_height", C.c_int),
                ("use_variable_gravity", C.c_int), ("stellar_spectrum_model", C.c_char_p),
                ("cloud_model", C.c_char_p)]


_dp = C.POINTER(C.c_double)


class _ObservationInput(C.Structure):
    _fields_ = [("name", C.c_char_p), ("type", C.c_char_p), ("nb_points", C.c_size_t),
                ("wavelengths", _dp), ("bin_edges", _dp), ("data", _dp), ("data_error", _dp),
                ("likelihood_weight", _dp), ("instrument_profile_fwhm", _dp),
                ("filter_nb_points", C.c_size_t), ("filter_wavelength", _dp), ("filter_response", _dp),
                ("filter_detector_type", C.c_char_p), ("spectrum_modifier", C.c_char_p)]


class EvalStats(C.Structure):
    _fields_ = [("nb_spectral_points", C.c_size_t), ("nb_levels", C.c_size_t), ("nb_species", C.c_size_t),
                ("nb_bins", C.c_size_t), ("distinct_table_rows", C.c_size_t), ("table_value_bytes", C.c_size_t),
                ("kernel_launches", C.c_size_t), ("kernel_ms_total", C.c_double),
                ("table_device_bytes", C.c_size_t)]


class BatchStats(C.Structure):
    _fields_ = [("precision", C.c_int), ("block", C.c_int), ("tile", C.c_int), ("grid", C.c_int),
                ("used_graph", C.c_int), ("evaluations", C.c_size_t), ("gpu_ms_total", C.c_double),
                ("adjoint_chunk", C.c_size_t), ("point_group", C.c_int)]


TABLE_PRECISIONS = {"fp64": 0, "fp32": 1, "fp16": 2, "bf16": 3, "int16": 4}

_lib = None


def lib():
    """Load libbear_hip.so (built by __graft_entry__.build / make -C bear_amd/csrc)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"bear_amd native library is not built: {LIB_PATH}")
        L = C.CDLL(LIB_PATH)
        vp, sz = C.c_void_p, C.c_size_t
        L.bear_last_error.restype = C.c_char_p
        L.bear_version.restype = C.c_char_p
        L.bear_opacity_data_register.argtypes = [C.c_char_p, C.c_double, C.c_int, sz, _dp, _dp, sz,
                                                 C.POINTER(C.c_float)]
        L.bear_retrieval_create.argtypes = [C.POINTER(_GlobalConfig), C.c_char_p, C.POINTER(_ModelConfig),
                                            C.POINTER(_ObservationInput), sz, C.POINTER(C.c_char_p),
                                            C.POINTER(C.c_char_p), sz, C.POINTER(vp)]
        L.bear_retrieval_destroy.argtypes = [vp]
        for f in ("bear_retrieval_nb_parameters", "bear_retrieval_nb_spectral_points",
                  "bear_retrieval_nb_observations"):
            getattr(L, f).restype = sz
            getattr(L, f).argtypes = [vp]
        L.bear_retrieval_nb_observation_points.restype = sz
        L.bear_retrieval_nb_observation_points.argtypes = [vp, sz]
        L.bear_retrieval_spectral_grid.argtypes = [vp, _dp, _dp]
        L.bear_retrieval_compute_likelihood.argtypes = [vp, _dp, sz, _dp]
        L.bear_retrieval_compute_model.argtypes = [vp, _dp, sz, C.c_int, _dp, _dp, C.POINTER(C.c_int)]
        L.bear_retrieval_set_profiling.argtypes = [vp, C.c_int]
        L.bear_retrieval_stats.argtypes = [vp, C.POINTER(EvalStats)]
        L.bear_forward_model_calc_model_gpu.argtypes = [vp, _dp, sz, vp, C.POINTER(C.c_void_p), C.POINTER(C.c_int)]
        L.bear_forward_model_calc_model_gpu_noalloc.argtypes = [vp, _dp, sz, vp, C.POINTER(C.c_void_p),
                                                                C.POINTER(C.c_int)]
        L.bear_retrieval_set_table_precision.argtypes = [vp, C.c_int]
        L.bear_retrieval_set_device_staging.argtypes = [vp, C.c_int]
        L.bear_retrieval_set_batch_options.argtypes = [vp, C.c_int, C.c_int]
        L.bear_retrieval_compute_likelihood_batch.argtypes = [vp, _dp, sz, sz, _dp]
        L.bear_retrieval_compute_model_batch.argtypes = [vp, _dp, sz, sz, _dp, _dp, C.POINTER(C.c_int)]
        L.bear_retrieval_batch_stats.argtypes = [vp, C.POINTER(BatchStats)]
        L.bear_retrieval_gradient_batch.argtypes = [vp, _dp, sz, sz, _dp, _dp]
        L.bear_retrieval_set_adjoint_memory.argtypes = [vp, sz]
        cpp = C.POINTER(C.c_char_p)
        L.bear_spectrum_model_create.argtypes = [C.POINTER(_GlobalConfig), C.c_char_p, C.c_double, C.c_double, sz,
                                                 C.c_char_p, _dp, _dp, sz, C.POINTER(vp)]
        L.bear_spectrum_model_destroy.argtypes = [vp]
        L.bear_spectrum_model_nb_spectral_points.restype = sz
        L.bear_spectrum_model_nb_spectral_points.argtypes = [vp]
        L.bear_spectrum_model_spectral_grid.argtypes = [vp, _dp, _dp]
        L.bear_emission_calc_spectrum.argtypes = [vp, C.c_double, C.c_double, C.c_double, _dp, _dp, sz, cpp, sz, _dp,
                                                  _dp, C.c_int, _dp]
        L.bear_transmission_calc_spectrum.argtypes = [vp, C.c_double, C.c_double, C.c_double, _dp, _dp, sz, cpp, sz,
                                                      _dp, _dp, C.c_int, _dp]
        L.bear_occultation_calc_spectrum.argtypes = [vp, C.c_double, C.c_double, _dp, _dp, sz, cpp, sz, _dp, _dp, _dp]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise BearError(rc, lib().bear_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _b(s):
    return s.encode() if isinstance(s, str) else s


@dataclass
class GlobalConfig:
    """GlobalConfig (config/global_config.h:28-71): the fields the likelihood path reads."""
    retrieval_folder_path: str = ""
    cross_section_file_path: str = ""
    wavenumber_file_path: str = ""
    spectral_disecretisation: str = "const_wavenumber"
    spectral_resolution: float = 1.0
    use_error_inflation: bool = False
    use_gpu: bool = True
    device: int = 0
    # device storage of the cross-section tables: "fp64" (every call), or
    # "fp32" / "fp16" / "bf16" / "int16" only (the batched forward calls at that
    # precision; single-point likelihoods and models as batches of one)
    table_storage: str = "fp64"

    def _c(self):
        kinds = {"const_wavenumber": 0, "const_wavelength": 1, "const_resolution": 2}
        if self.spectral_disecretisation not in kinds:
            raise BearError(-1, f"unknown spectral discretisation {self.spectral_disecretisation}")
        self._keep = [_b(self.retrieval_folder_path), _b(self.cross_section_file_path), _b(self.wavenumber_file_path)]
        if self.table_storage not in TABLE_PRECISIONS:
            raise BearError(-1, f"unknown table storage {self.table_storage}")
        return _GlobalConfig(int(self.use_gpu), self.device, *self._keep, kinds[self.spectral_disecretisation],
                             float(self.spectral_resolution), int(self.use_error_inflation),
                             TABLE_PRECISIONS[self.table_storage])


@dataclass
class ModelConfig:
    """Forward-model configuration in forward_model.config terms."""
    nb_grid_points: int
    atmos_bottom_pressure: float
    atmos_top_pressure: float
    temperature_profile: str                 # "poly 6 1" | "const"
    chemistry: list                          # ["iso H2O CH4", "bg H2He"]
    opacity_species: list                    # [("H2O", "Molecules/H2O"), ...]
    radiative_transfer: str = "scm"
    fit_mean_molecular_weight: bool = False  # transmission
    fit_scale_height: bool = False           # transmission
    use_variable_gravity: bool = False       # transmission
    stellar_spectrum_model: str = "blackbody"  # secondary eclipse
    cloud_model: list = field(default_factory=list)   # ["grey", "power_law 1.0 fb", "KHnongrey 1.0"]

    def _c(self):
        chem = "\n".join(self.chemistry)
        opac = "\n".join(f"{s} {f}" for s, f in self.opacity_species)
        self._keep = [_b(self.temperature_profile), _b(self.radiative_transfer), _b(chem), _b(opac),
                      _b(self.stellar_spectrum_model), _b("\n".join(self.cloud_model))]
        return _ModelConfig(int(self.nb_grid_points), float(self.atmos_bottom_pressure),
                            float(self.atmos_top_pressure), self._keep[0], self._keep[1], self._keep[2], self._keep[3],
                            int(self.fit_mean_molecular_weight), int(self.fit_scale_height),
                            int(self.use_variable_gravity), self._keep[4], self._keep[5])


@dataclass
class ObservationInput:
    """ObservationInput (observations/observations.h:43-83)."""
    name: str
    type: str
    data: np.ndarray
    data_error: np.ndarray
    wavelengths: np.ndarray | None = None
    bin_wavelength_edges: np.ndarray | None = None      # [n, 2]
    likelihood_weight: np.ndarray | None = None
    instrument_profile_fwhm: np.ndarray | None = None
    filter_response: tuple | None = None                # (wavelength[], response[])
    filter_detector_type: str = ""
    spectrum_modifier_id: str = "none"

    def _c(self):
        f64 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        arrs = dict(w=f64(self.wavelengths), e=f64(self.bin_wavelength_edges), d=f64(self.data),
                    s=f64(self.data_error), lw=f64(self.likelihood_weight), fw=f64(self.instrument_profile_fwhm))
        fl = fr = None
        if self.filter_response is not None:
            fl, fr = f64(self.filter_response[0]), f64(self.filter_response[1])
        self._keep = (arrs, fl, fr, _b(self.name), _b(self.type), _b(self.filter_detector_type),
                      _b(self.spectrum_modifier_id))
        return _ObservationInput(self._keep[3], self._keep[4], arrs["d"].size, _ptr(arrs["w"]), _ptr(arrs["e"]),
                                 _ptr(arrs["d"]), _ptr(arrs["s"]), _ptr(arrs["lw"]), _ptr(arrs["fw"]),
                                 0 if fl is None else fl.size, _ptr(fl), _ptr(fr), self._keep[5], self._keep[6])


@dataclass
class ForwardModelOutput:
    neglect_model: bool = False
    spectrum: np.ndarray = field(default_factory=lambda: np.zeros(0))
    spectrum_obs: list = field(default_factory=list)


def register_opacity_data(folder, temperatures, pressures, rows, species_mass=0.0, log_storage=False):
    """Hold one species' table in memory under ``folder`` (instead of a
    filelist.dat folder): rows [nb_rows, nb_wavenumbers] float32 over the full
    wavenumber list, with the (T, p) of each row."""
    rows = np.ascontiguousarray(rows, dtype=np.float32)
    t = np.ascontiguousarray(temperatures, dtype=np.float64)
    p = np.ascontiguousarray(pressures, dtype=np.float64)
    _check(lib().bear_opacity_data_register(_b(folder), float(species_mass), int(log_storage), rows.shape[0],
                                            _ptr(t), _ptr(p), rows.shape[1], rows.ctypes.data_as(C.POINTER(C.c_float))))


def clear_opacity_data():
    _check(lib().bear_opacity_data_clear())


class Retrieval:
    """Retrieval restricted to the likelihood path (retrieval/retrieval.h)."""

    def __init__(self, config: GlobalConfig, model_config: ModelConfig, observations=(),
                 observation_files=(), observation_modifiers=None, forward_model_type="emission"):
        L = lib()
        gc, mc = config._c(), model_config._c()
        obs = list(observations)
        arr = (_ObservationInput * max(len(obs), 1))(*[o._c() for o in obs])
        files = [_b(f) for f in observation_files]
        mods = [_b(m) for m in (observation_modifiers or ["none"] * len(files))]
        f_arr = (C.c_char_p * max(len(files), 1))(*files)
        m_arr = (C.c_char_p * max(len(mods), 1))(*mods)
        h = C.c_void_p()
        _check(L.bear_retrieval_create(C.byref(gc), _b(forward_model_type), C.byref(mc), arr, len(obs), f_arr, m_arr,
                                       len(files), C.byref(h)))
        self._L, self._h = L, h
        self.nb_parameters = L.bear_retrieval_nb_parameters(h)
        self.nb_spectral_points = L.bear_retrieval_nb_spectral_points(h)
        self.nb_observations = L.bear_retrieval_nb_observations(h)
        self.nb_observation_points = [L.bear_retrieval_nb_observation_points(h, i) for i in range(self.nb_observations)]

    def close(self):
        if getattr(self, "_h", None):
            self._L.bear_retrieval_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def nbParameters(self):
        return self.nb_parameters

    def spectral_grid(self):
        nu = np.empty(self.nb_spectral_points)
        lam = np.empty(self.nb_spectral_points)
        _check(self._L.bear_retrieval_spectral_grid(self._h, _ptr(nu), _ptr(lam)))
        return nu, lam

    def computeLikelihood(self, physical_parameters) -> float:
        p = np.ascontiguousarray(physical_parameters, dtype=np.float64)
        out = C.c_double()
        _check(self._L.bear_retrieval_compute_likelihood(self._h, _ptr(p), p.size, C.byref(out)))
        return out.value

    def computeModel(self, physical_parameters, return_high_res_spectrum=True) -> ForwardModelOutput:
        p = np.ascontiguousarray(physical_parameters, dtype=np.float64)
        spec = np.empty(self.nb_spectral_points) if return_high_res_spectrum else None
        obs = np.empty(sum(self.nb_observation_points))
        neg = C.c_int(0)
        _check(self._L.bear_retrieval_compute_model(self._h, _ptr(p), p.size, int(return_high_res_spectrum),
                                                    _ptr(spec), _ptr(obs), C.byref(neg)))
        out, k = [], 0
        for n in self.nb_observation_points:
            out.append(obs[k:k + n].copy())
            k += n
        return ForwardModelOutput(bool(neg.value), spec if spec is not None else np.zeros(0), out)

    # ---- batches: parameters [batch, nb_parameters]
    def set_device_staging(self, enable=True):
        """Form the batched calls' atmospheres and stencils on the device
        (the default) or on the host."""
        _check(self._L.bear_retrieval_set_device_staging(self._h, int(bool(enable))))

    def set_table_precision(self, precision="fp64"):
        """Table storage of the batched path: "fp64", "fp32", "fp16", "bf16" or
        "int16" (16-bit integers scaled per species and wavenumber)."""
        if precision not in TABLE_PRECISIONS:
            raise BearError(-1, f"unknown table precision {precision}")
        _check(self._L.bear_retrieval_set_table_precision(self._h, TABLE_PRECISIONS[precision]))

    def set_batch_options(self, use_graph=True, autotune=True):
        _check(self._L.bear_retrieval_set_batch_options(self._h, int(use_graph), int(autotune)))

    def computeLikelihoodBatch(self, physical_parameters) -> np.ndarray:
        """log-likelihoods of a batch of live points, one launch sequence."""
        p = np.ascontiguousarray(np.atleast_2d(physical_parameters), dtype=np.float64)
        out = np.empty(p.shape[0])
        _check(self._L.bear_retrieval_compute_likelihood_batch(self._h, _ptr(p), p.shape[0], p.shape[1], _ptr(out)))
        return out

    def computeModelBatch(self, physical_parameters, return_high_res_spectrum=True):
        """(spectra [batch, N] or None, per observation bins [batch, points], neglect [batch])."""
        p = np.ascontiguousarray(np.atleast_2d(physical_parameters), dtype=np.float64)
        b = p.shape[0]
        spec = np.empty((b, self.nb_spectral_points)) if return_high_res_spectrum else None
        total = sum(self.nb_observation_points)
        obs = np.empty((b, total))
        neg = np.zeros(b, dtype=np.int32)
        _check(self._L.bear_retrieval_compute_model_batch(self._h, _ptr(p), b, p.shape[1], _ptr(spec), _ptr(obs),
                                                          neg.ctypes.data_as(C.POINTER(C.c_int))))
        out, k = [], 0
        for n in self.nb_observation_points:
            out.append(obs[:, k:k + n].copy())
            k += n
        return spec, out, neg.astype(bool)

    def gradientBatch(self, physical_parameters):
        """(log-likelihoods [batch], gradients [batch, nb_parameters]) from the
        hand-written adjoint (emission, secondary-eclipse and transmission
        models, FP64 tables)."""
        p = np.ascontiguousarray(np.atleast_2d(physical_parameters), dtype=np.float64)
        ll = np.empty(p.shape[0])
        grad = np.empty_like(p)
        _check(self._L.bear_retrieval_gradient_batch(self._h, _ptr(p), p.shape[0], p.shape[1], _ptr(ll), _ptr(grad)))
        return ll, grad

    def set_adjoint_memory(self, nbytes):
        """Device memory the gradient may hold for stored forward intensities;
        0 recomputes them inside the backward kernel."""
        _check(self._L.bear_retrieval_set_adjoint_memory(self._h, int(nbytes)))

    def batch_stats(self) -> dict:
        st = BatchStats()
        _check(self._L.bear_retrieval_batch_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in BatchStats._fields_}

    def set_profiling(self, enable=True):
        _check(self._L.bear_retrieval_set_profiling(self._h, int(enable)))

    def stats(self) -> dict:
        st = EvalStats()
        _check(self._L.bear_retrieval_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in EvalStats._fields_}

    def calcModelGPU(self, physical_parameters, spectrum_ptr, spectrum_obs_ptrs, no_alloc=False):
        """ForwardModel::calcModelGPU on caller-owned device buffers (raw
        addresses, e.g. torch tensors' data_ptr()); returns the neglect flag.
        no_alloc: the variant that keeps its per-evaluation device buffers in
        the library's static scratch array and allocates nothing."""
        p = np.ascontiguousarray(physical_parameters, dtype=np.float64)
        ptrs = (C.c_void_p * max(len(spectrum_obs_ptrs), 1))(*[C.c_void_p(int(x)) for x in spectrum_obs_ptrs])
        neg = C.c_int(0)
        f = self._L.bear_forward_model_calc_model_gpu_noalloc if no_alloc else self._L.bear_forward_model_calc_model_gpu
        _check(f(self._h, _ptr(p), p.size, C.c_void_p(int(spectrum_ptr)) if spectrum_ptr else None, ptrs,
                 C.byref(neg)))
        return bool(neg.value)


# ---------------------------------------------------------------- direct spectra
class SpectralGrid:
    """pyBeAR's SpectralGrid(config, wavelength_min, wavelength_max): the
    wavenumber list of the configuration sampled between two wavelengths
    (micron) for the direct models below (spectral_grid.cpp:52-72).  The
    sampled lists are those of the model built on it (wavenumber_list,
    wavelength_list)."""

    def __init__(self, config: GlobalConfig, wavelength_min: float, wavelength_max: float):
        self.config = config
        self.wavelength_min = float(wavelength_min)
        self.wavelength_max = float(wavelength_max)


class _DirectModel:
    """A forward model whose calcSpectrum takes the atmosphere's structure
    from the caller (bear_spectrum_model in include/bear_hip.h)."""

    def __init__(self, kind, spectral_grid, nb_grid_points, opacity_species_symbol, opacity_species_folder,
                 stellar=None):
        L = lib()
        if len(opacity_species_symbol) != len(opacity_species_folder):
            raise BearError(-1, "one opacity folder per opacity species")
        gc = spectral_grid.config._c()
        opac = "\n".join(f"{s} {f}" for s, f in zip(opacity_species_symbol, opacity_species_folder))
        sw = sf = None
        if stellar is not None:
            sw = np.ascontiguousarray(stellar[0], dtype=np.float64)
            sf = np.ascontiguousarray(stellar[1], dtype=np.float64)
        h = C.c_void_p()
        _check(L.bear_spectrum_model_create(C.byref(gc), _b(kind), spectral_grid.wavelength_min,
                                            spectral_grid.wavelength_max, int(nb_grid_points), _b(opac), _ptr(sw),
                                            _ptr(sf), 0 if sw is None else sw.size, C.byref(h)))
        self._L, self._h = L, h
        self.nb_grid_points = int(nb_grid_points)
        self.nb_spectral_points = L.bear_spectrum_model_nb_spectral_points(h)
        nu, lam = np.empty(self.nb_spectral_points), np.empty(self.nb_spectral_points)
        _check(L.bear_spectrum_model_spectral_grid(h, _ptr(nu), _ptr(lam)))
        self.wavenumber_list, self.wavelength_list = nu, lam

    def close(self):
        if getattr(self, "_h", None):
            self._L.bear_spectrum_model_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _structure(self, pressure, temperature, species_symbol, mixing_ratios, cloud_optical_depth):
        p = np.ascontiguousarray(pressure, dtype=np.float64)
        t = np.ascontiguousarray(temperature, dtype=np.float64)
        x = np.ascontiguousarray(mixing_ratios, dtype=np.float64).reshape(p.size, len(species_symbol))
        names = (C.c_char_p * max(len(species_symbol), 1))(*[_b(s) for s in species_symbol])
        cloud = None
        if cloud_optical_depth is not None and len(cloud_optical_depth) > 0:
            cloud = np.ascontiguousarray(cloud_optical_depth, dtype=np.float64)
            if cloud.size != (p.size - 1) * self.nb_spectral_points:
                raise BearError(-1, "cloud optical depths: one per layer and spectral point")
        keep = (p, t, x, names, cloud)
        return keep, np.empty(self.nb_spectral_points)


class EmissionModel(_DirectModel):
    """pyBeAR EmissionModel(config, spectral_grid, nb_grid_points,
    radiative_transfer_desc, radiative_transfer_param, opacity_species_symbol,
    opacity_species_folder) (emission.h:134-166)."""

    def __init__(self, config: GlobalConfig, spectral_grid: SpectralGrid, nb_grid_points, radiative_transfer_desc="scm",
                 radiative_transfer_param=(), opacity_species_symbol=(), opacity_species_folder=()):
        if radiative_transfer_desc != "scm":
            raise BearError(-1, f"radiative transfer {radiative_transfer_desc} is not available (only scm)")
        spectral_grid.config = config
        super().__init__("emission", spectral_grid, nb_grid_points, opacity_species_symbol, opacity_species_folder)

    def calcSpectrum(self, surface_gravity, radius, distance, pressure, temperature, species_symbol, mixing_ratios,
                     cloud_optical_depth=(), use_variable_gravity=False):
        """The flux in W m^-2 cm times (radius / distance)^2 per wavenumber
        (emission.cpp:268-340); mixing_ratios [levels][species]."""
        (p, t, x, names, cloud), out = self._structure(pressure, temperature, species_symbol, mixing_ratios,
                                                       cloud_optical_depth)
        _check(self._L.bear_emission_calc_spectrum(self._h, float(surface_gravity), float(radius), float(distance),
                                                   _ptr(p), _ptr(t), p.size, names, len(species_symbol), _ptr(x),
                                                   _ptr(cloud), int(use_variable_gravity), _ptr(out)))
        return out


class TransmissionModel(_DirectModel):
    """pyBeAR TransmissionModel(config, spectral_grid, nb_grid_points,
    opacity_species_symbol, opacity_species_folder) (transmission.h:136-190)."""

    def __init__(self, config: GlobalConfig, spectral_grid: SpectralGrid, nb_grid_points, opacity_species_symbol=(),
                 opacity_species_folder=()):
        spectral_grid.config = config
        super().__init__("transmission", spectral_grid, nb_grid_points, opacity_species_symbol, opacity_species_folder)

    def calcSpectrum(self, surface_gravity, planet_radius, radius_ratio, pressure, temperature, species_symbol,
                     mixing_ratios, cloud_optical_depth=(), use_variable_gravity=False):
        """The transit depth in ppm per wavenumber (transmission.cpp:371-465)."""
        (p, t, x, names, cloud), out = self._structure(pressure, temperature, species_symbol, mixing_ratios,
                                                       cloud_optical_depth)
        _check(self._L.bear_transmission_calc_spectrum(self._h, float(surface_gravity), float(planet_radius),
                                                       float(radius_ratio), _ptr(p), _ptr(t), p.size, names,
                                                       len(species_symbol), _ptr(x), _ptr(cloud),
                                                       int(use_variable_gravity), _ptr(out)))
        return out


class OccultationModel(_DirectModel):
    """pyBeAR OccultationModel(config, spectral_grid, nb_grid_points,
    stellar_spectrum_wavelengths, stellar_spectrum_flux, opacity_species_symbol,
    opacity_species_folder) (secondary_eclipse.cpp:79-114): the star's flux in
    W m^-2 micron^-1 at the given wavelengths (micron)."""

    def __init__(self, config: GlobalConfig, spectral_grid: SpectralGrid, nb_grid_points, stellar_spectrum_wavelengths,
                 stellar_spectrum_flux, opacity_species_symbol=(), opacity_species_folder=()):
        spectral_grid.config = config
        super().__init__("secondary_eclipse", spectral_grid, nb_grid_points, opacity_species_symbol,
                         opacity_species_folder, stellar=(stellar_spectrum_wavelengths, stellar_spectrum_flux))

    def calcSpectrum(self, surface_gravity, radius_ratio, pressure, temperature, species_symbol, mixing_ratios,
                     cloud_optical_depth=()):
        """The eclipse depth in ppm per wavenumber (secondary_eclipse.cpp:395-482)."""
        (p, t, x, names, cloud), out = self._structure(pressure, temperature, species_symbol, mixing_ratios,
                                                       cloud_optical_depth)
        _check(self._L.bear_occultation_calc_spectrum(self._h, float(surface_gravity), float(radius_ratio), _ptr(p),
                                                      _ptr(t), p.size, names, len(species_symbol), _ptr(x),
                                                      _ptr(cloud), _ptr(out)))
        return out


def version():
    return lib().bear_version().decode()
