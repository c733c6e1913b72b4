"""Seeded synthetic workloads for the BeAR forward-model engine.

Everything here is test/benchmark input generation: opacity tables built from
seeded Lorentzian line lists, smooth collision-induced-absorption (CIA) tables,
observations (spectroscopy with a Gaussian line-spread function, band
spectroscopy, photometry with bandpass files) and parameter vectors.

A case can be written to disk in the reference's own file formats, so that
the same bytes drive the product, the CPU oracle and the reference build:

* ``wavenumber_full.dat``  count, then one wavenumber per line
  (read by ``spectral_grid.cpp:87-119``);
* ``<species folder>/filelist.dat``  first line ``mass [log]``, then one
  ``pressure temperature file`` line per table row (``opacity_species.cpp:66-122``);
* raw float32 row files over the full wavenumber grid, linear or log10
  (``cross_section_file.cpp:33-62``, ``sampled_data.cpp:48-88``);
* observation files with the ``#type`` / ``#filter transmission function``
  layout (``observations_load_data.cpp:40-330``) and bandpass files with the
  detector type on line 4 (``observations_filter_response.cpp:42-117``);
* ``observations.list`` (``retrieval_load_observations.cpp:103-141``).
"""

from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

# cgs constants (same values as the reference's physical_const.h)
PLANCK_H = 6.62606896e-27
LIGHT_C = 2.99792458e10
BOLTZMANN_K = 1.3806504e-16
RADIUS_JUPITER = 7.1492e9
RADIUS_SUN = 6.96342e10
PARSEC = 3.08567758135e18
AVOGADRO = 6.022140857e23

MOLECULAR_WEIGHT = {
    "H2O": 18.01528, "CO2": 44.01, "CO": 28.0101, "CH4": 16.04246,
    "HCN": 27.0253, "NH3": 17.03052, "C2H2": 26.04, "H2S": 34.09099,
    "K": 39.0983, "Na": 22.98977, "TiO": 63.8664, "VO": 66.9409,
    "FeH": 56.853, "SH": 34.08, "OH": 17.008, "SO2": 64.066,
}


@dataclass
class SpeciesTable:
    """One opacity species: a (T, p) grid of rows over the full wavenumber grid.

    ``values`` holds exactly what the row files contain (float32): linear
    cross-sections in cm^2 (``mass == 0``), opacities in cm^2/g (``mass > 0``)
    or log10 of either when ``log_storage`` is set.
    """
    symbol: str
    folder: str
    temperatures: np.ndarray          # ascending, K
    pressures: np.ndarray             # ascending, bar
    values: np.ndarray                # float32 [nT, nP, N_full]
    mass: float = 0.0
    log_storage: bool = False
    kind: str = "line"                # "line" | "cia"


@dataclass
class ObservationSpec:
    name: str
    type: str                         # spectroscopy | band-spectroscopy | photometry
    wavelengths: np.ndarray | None = None     # spectroscopy: centres (micron)
    bin_edges: np.ndarray | None = None       # band types: [nb, 2] (micron)
    data: np.ndarray | None = None
    error: np.ndarray | None = None
    fwhm: np.ndarray | None = None            # optional instrument profile FWHM (micron)
    weight: np.ndarray | None = None          # optional likelihood weights
    filter: tuple | None = None               # (detector "Photon"|"Energy", wavelength[], response[])
    modifier: str = "none"


@dataclass
class Case:
    name: str
    wavenumber_full: np.ndarray
    spectral_resolution: float
    nb_levels: int
    p_bottom: float
    p_top: float
    temperature_model: tuple           # ("poly", nb_elements, degree) | ("const",)
    chemistry: list                    # [("iso", [species...]), ("bg", ["H2He"])]
    species: list                      # [SpeciesTable]
    rayleigh: list                     # ["H2", ...]  (opacity entries with folder "Rayleigh")
    observations: list                 # [ObservationSpec]
    parameters: np.ndarray             # physical parameter vector
    use_error_inflation: bool = False
    model: str = "emission"
    extra: dict = field(default_factory=dict)
    # GlobalConfig::spectral_disecretisation: 0 const_wavenumber (spectral
    # resolution in cm^-1), 1 const_wavelength (micron), 2 const_resolution (R)
    discretisation: int = 0
    continuum: list = field(default_factory=list)   # ["H-"]: species without tables
    # cloud models as forward_model.config lines ("grey", "power_law 1.0 fb",
    # "KHnongrey 1.0"); their parameters follow the temperature's
    clouds: list = field(default_factory=list)

    @property
    def opacity_entries(self):
        """(symbol, folder) pairs in the order a forward_model.config lists them."""
        out = []
        for s in self.species:
            out.append((s.symbol, s.folder))
        for r in self.rayleigh:
            out.append((r, "Rayleigh"))
        for c in self.continuum:
            out.append((c, "Continuum"))
        return out


# ---------------------------------------------------------------------------
# tables
# ---------------------------------------------------------------------------

def _line_list(rng, n_lines, nu_lo, nu_hi):
    centres = rng.uniform(nu_lo, nu_hi, n_lines)
    strength = 10.0 ** rng.uniform(-24.0, -18.0, n_lines)      # cm / molecule
    lower_energy = rng.uniform(0.0, 4000.0, n_lines)            # cm^-1
    return centres, strength, lower_energy


def _lorentz_rows(nu, centres, weights_per_row, widths):
    """Sum of Lorentzians of one width per row, by FFT on a padded grid.

    ``weights_per_row`` is [R, n_lines] (line strengths at each row's
    temperature), ``widths`` is [R] (HWHM in cm^-1).  Line centres snap to
    the nearest grid point; the profile is the exact Lorentzian transform
    exp(-2 pi gamma |k|), so rows are smooth and strictly positive.
    """
    n = nu.size
    step = (nu[-1] - nu[0]) / (n - 1)
    m = 1 << int(math.ceil(math.log2(2 * n)))
    idx = np.clip(np.rint((centres - nu[0]) / step).astype(np.int64), 0, n - 1)
    freqs = np.fft.rfftfreq(m, d=step)
    out = np.empty((weights_per_row.shape[0], n), dtype=np.float64)
    for r in range(weights_per_row.shape[0]):
        hist = np.bincount(idx, weights=weights_per_row[r], minlength=m)
        spec = np.fft.rfft(hist)
        spec *= np.exp(-2.0 * math.pi * widths[r] * freqs)
        out[r] = np.fft.irfft(spec, n=m)[:n] / step
    return out


def _line_rows_gpu(nu, centres, weights_per_row, widths, mass, log_storage, batch=32):
    """The rows of line_table computed entirely on the GPU (torch), in
    batches of rows: the line histogram, the Lorentzian transform, the
    1e-30 floor, the mass scaling and log10; returns float32 [R, n].  For
    the full-size workloads (millions of values per row), where the numpy
    path takes minutes; the histogram is summed with atomics, so the values
    are reproducible to the last bit of float64 only, not across backends."""
    import torch
    dev = torch.device("cuda")
    n = nu.size
    step = (nu[-1] - nu[0]) / (n - 1)
    m = 1 << int(math.ceil(math.log2(2 * n)))
    idx = torch.as_tensor(np.clip(np.rint((centres - nu[0]) / step).astype(np.int64), 0, n - 1), device=dev)
    freqs = torch.as_tensor(np.fft.rfftfreq(m, d=step), device=dev)
    w = torch.as_tensor(weights_per_row, device=dev)
    g = torch.as_tensor(widths, device=dev)
    R = weights_per_row.shape[0]
    out = np.empty((R, n), dtype=np.float32)
    for r0 in range(0, R, batch):
        r1 = min(R, r0 + batch)
        hist = torch.zeros((r1 - r0, m), dtype=torch.float64, device=dev)
        hist.index_add_(1, idx, w[r0:r1])
        spec = torch.fft.rfft(hist, dim=1)
        del hist
        spec *= torch.exp(-2.0 * math.pi * g[r0:r1, None] * freqs[None, :])
        rows = torch.fft.irfft(spec, n=m, dim=1)[:, :n] / step
        del spec
        sigma = 1e-30 + torch.clamp(rows, min=0.0)
        if mass > 0:
            sigma = sigma / (mass / AVOGADRO)
        if log_storage:
            sigma = torch.log10(sigma)
        out[r0:r1] = sigma.to(torch.float32).cpu().numpy()
    return out


def line_table(rng, symbol, folder, nu, temperatures, pressures, n_lines=3000,
               mass=0.0, log_storage=False, backend="numpy"):
    """log10 sigma = log10(1e-30 + sum_k S_k(T) L(nu; nu_k, gamma(p, T)))."""
    centres, s0, e_low = _line_list(rng, n_lines, nu[0], nu[-1])
    c2 = PLANCK_H * LIGHT_C / BOLTZMANN_K
    t_grid, p_grid = np.meshgrid(temperatures, pressures, indexing="ij")
    t_rows, p_rows = t_grid.ravel(), p_grid.ravel()
    # line strengths at each row temperature (Boltzmann factor around 1000 K)
    w = s0[None, :] * (1000.0 / t_rows[:, None]) ** 1.5 \
        * np.exp(-c2 * e_low[None, :] * (1.0 / t_rows[:, None] - 1.0 / 1000.0))
    gamma = 0.01 * np.sqrt(t_rows / 1000.0) + 0.07 * p_rows * (296.0 / t_rows) ** 0.7
    if backend == "torch":
        values = _line_rows_gpu(nu, centres, w, gamma, mass, log_storage)
        return SpeciesTable(symbol, folder, np.asarray(temperatures, float), np.asarray(pressures, float),
                            values.reshape(temperatures.size, pressures.size, nu.size),
                            mass=mass, log_storage=log_storage, kind="line")
    sigma = 1e-30 + np.maximum(_lorentz_rows(nu, centres, w, gamma), 0.0)
    sigma = sigma.reshape(temperatures.size, pressures.size, nu.size)
    if mass > 0:
        sigma = sigma / (mass / AVOGADRO)            # store as cm^2/g
    if log_storage:
        sigma = np.log10(sigma)
    return SpeciesTable(symbol, folder, np.asarray(temperatures, float),
                        np.asarray(pressures, float), sigma.astype(np.float32),
                        mass=mass, log_storage=log_storage, kind="line")


def cia_table(rng, symbol, folder, nu, temperatures):
    """Smooth, T-dependent CIA cross-section (cm^5), one pressure row per T."""
    a = rng.uniform(0.5, 1.5)
    phase = rng.uniform(0, 2 * math.pi)
    centre = rng.uniform(nu[0], nu[-1])
    x = (nu - centre) / (nu[-1] - nu[0])
    rows = []
    for t in temperatures:
        lg = (-45.5 + 0.6 * a * np.sin(2 * math.pi * nu / 3000.0 + phase)
              - 3.0 * x * x + 0.9 * np.log10(t / 1000.0) - 0.2 * x * np.log10(t / 1000.0))
        rows.append(10.0 ** lg)
    values = np.asarray(rows)[:, None, :].astype(np.float32)
    return SpeciesTable(symbol, folder, np.asarray(temperatures, float),
                        np.array([1.0]), values, mass=0.0, log_storage=False, kind="cia")


# ---------------------------------------------------------------------------
# observations
# ---------------------------------------------------------------------------

def blackbody_flux_lambda(wavelength_um, t_eff, scale):
    """pi B_lambda(T) * scale in W m^-2 micron^-1 (synthetic 'observed' data)."""
    nu = 1e4 / wavelength_um
    b_nu = 2 * PLANCK_H * LIGHT_C ** 2 * nu ** 3 / np.expm1(PLANCK_H * LIGHT_C * nu / BOLTZMANN_K / t_eff)
    # erg/s/cm^2/sr/cm^-1 -> W m^-2 cm (x1e-3), then per micron: / lambda^2 * 1e4
    return math.pi * b_nu * 1e-3 * scale / wavelength_um ** 2 * 1e4


def noisy(rng, model, frac=0.02):
    data = model * (1.0 + frac * rng.standard_normal(model.shape))
    return data, frac * np.abs(data)


def bandpass(rng, lo, hi, n=50):
    """Seeded smooth bandpass over [lo, hi] micron, ascending wavelengths."""
    w = np.linspace(lo, hi, n)
    x = (w - lo) / (hi - lo)
    resp = np.clip(np.sin(math.pi * x) ** rng.uniform(0.5, 2.0)
                   * (1.0 + 0.2 * np.sin(2 * math.pi * rng.uniform(1, 3) * x + rng.uniform(0, 6))), 0, None)
    resp[0] = 0.0
    resp[-1] = 0.0
    return w, resp


# ---------------------------------------------------------------------------
# parameters
# ---------------------------------------------------------------------------

def poly_temperature_params(rng, nb_dof, t_lo=500.0, t_hi=2500.0):
    """Control temperatures drawn in [t_lo, t_hi] and turned into the DOF
    ratios of the piecewise-polynomial profile (T0 = theta0, Ti = Ti-1 thetai)."""
    temps = np.sort(rng.uniform(t_lo, t_hi, nb_dof))[::-1]     # hot at the bottom
    theta = np.empty(nb_dof)
    theta[0] = temps[0]
    theta[1:] = temps[1:] / temps[:-1]
    return theta


def emission_parameters(rng, nb_absorbers, nb_dof):
    log_g = rng.uniform(4.0, 5.5)
    scale_f = 1.0
    distance = 10.0 * PARSEC
    mixing = 10.0 ** rng.uniform(-8.0, -3.0, nb_absorbers)
    theta = poly_temperature_params(rng, nb_dof)
    return np.concatenate([[log_g, scale_f, distance], mixing, theta])


# ---------------------------------------------------------------------------
# the configured cases
# ---------------------------------------------------------------------------

ABSORBERS = ["H2O", "CH4", "CO", "CO2", "NH3", "H2S", "HCN", "K", "TiO", "VO", "FeH", "SO2"]


def _grid(nu_lo, nu_hi, step):
    n = int(round((nu_hi - nu_lo) / step)) + 1
    return nu_lo + step * np.arange(n, dtype=np.float64)


def tiny_case(seed=7):
    """Small emission case exercising every observation path in seconds:
    spectroscopy with an LSF (ascending input), band spectroscopy with a
    shift modifier, photometry with photon and energy bandpasses."""
    rng = np.random.default_rng(seed)
    nu = _grid(2000.0, 6000.0, 0.5)
    temps = np.array([300.0, 700.0, 1100.0, 1500.0, 1900.0, 2300.0, 2700.0])
    press = 10.0 ** np.arange(-5.0, 2.5, 1.5)
    species = [
        line_table(rng, "H2O", "H2O", nu, temps, press, n_lines=400, mass=MOLECULAR_WEIGHT["H2O"]),
        line_table(rng, "CH4", "CH4", nu, temps, press, n_lines=300, log_storage=True),
        cia_table(rng, "CIA-H2-H2", "CIA/H2-H2", nu, temps),
        cia_table(rng, "CIA-H2-He", "CIA/H2-He", nu, temps),
    ]
    scale = (RADIUS_JUPITER / (10 * PARSEC)) ** 2
    obs = []
    # spectroscopy, ascending wavelengths, with an instrument profile (R ~ 150)
    wl = np.linspace(1.80, 4.60, 24)
    model = blackbody_flux_lambda(wl, 1200.0, scale)
    d, e = noisy(rng, model)
    obs.append(ObservationSpec("spec_lsf", "spectroscopy", wavelengths=wl, data=d, error=e,
                               fwhm=wl / 150.0, weight=np.where(np.arange(wl.size) % 5 == 0, 0.5, 1.0)))
    # band spectroscopy, descending order, no profile, shift modifier
    edges = np.linspace(4.9, 2.1, 9)
    be = np.stack([edges[:-1], edges[1:]], axis=1)
    model = blackbody_flux_lambda(be.mean(axis=1), 1200.0, scale)
    d, e = noisy(rng, model)
    obs.append(ObservationSpec("bands", "band-spectroscopy", bin_edges=be, data=d, error=e,
                               modifier="shift_const"))
    # photometry with photon-counting and energy bandpasses
    for i, (lo, hi, det) in enumerate([(2.2, 2.6, "Photon"), (3.3, 3.9, "Energy")]):
        w, r = bandpass(rng, lo, hi, 30)
        model = blackbody_flux_lambda(np.array([(lo + hi) / 2]), 1200.0, scale)
        d, e = noisy(rng, model)
        obs.append(ObservationSpec(f"phot{i}", "photometry", bin_edges=np.array([[lo, hi]]),
                                   data=d, error=e, filter=(det, w, r)))
    params = emission_parameters(rng, 2, 7)
    params = np.concatenate([params, [0.0]])          # shift_const of the band observation
    return Case("tiny", nu, 1.5, 14, 100.0, 1e-5, ("poly", 6, 1),
                [("iso", ["H2O", "CH4"]), ("bg", ["H2He"])], species, ["H2"], obs, params)


def emission_c1(seed=11):
    """Plain CPU reference emission model: 60 levels, 1e2..1e-6 bar, poly 6 1,
    2 absorbers on 30 T x 9 p tables over 1000-10000 cm^-1 (20,001 points),
    one 100-point spectrum with a Gaussian LSF at R = 100, 2 % noise."""
    rng = np.random.default_rng(seed)
    nu = _grid(1000.0, 10000.0, 0.45)
    temps = np.arange(100.0, 3001.0, 100.0)
    press = 10.0 ** np.arange(-6.0, 2.5, 1.0)
    species = [
        line_table(rng, "H2O", "H2O", nu, temps, press, n_lines=3000),
        line_table(rng, "CH4", "CH4", nu, temps, press, n_lines=3000, log_storage=True),
    ]
    r = 100.0
    lam_lo, lam_hi = 1e4 / nu[-1], 1e4 / nu[0]
    # the observation range (bins + 5 sigma of the profile) stays inside the grid
    lo = lam_lo * 1.12
    hi = lam_hi * 0.88
    wl = np.linspace(lo, hi, 100)
    scale = (RADIUS_JUPITER / (10 * PARSEC)) ** 2
    d, e = noisy(rng, blackbody_flux_lambda(wl, 1100.0, scale))
    obs = [ObservationSpec("spectrum_R100", "spectroscopy", wavelengths=wl[::-1].copy(),
                           data=d[::-1].copy(), error=e[::-1].copy(), fwhm=wl[::-1] / r)]
    params = emission_parameters(rng, 2, 7)
    return Case("emission_c1", nu, 0.45, 60, 100.0, 1e-6, ("poly", 6, 1),
                [("iso", ["H2O", "CH4"]), ("bg", ["H2He"])], species, [], obs, params)


def emission_c2(seed=23, backend="numpy", n_wavenumbers=400_000, nb_bins=300):
    """Plain GPU drop-in emission model: 100 levels, 6 absorbers + H2-H2/H2-He
    CIA + H2 Rayleigh, 400,000 wavenumbers at 0.025 cm^-1, 30 T x 9 p tables,
    300 photometric bins each with a seeded 50-point bandpass; the bins tile
    the wavenumber list, so all 400,000 wavenumbers are sampled."""
    rng = np.random.default_rng(seed)
    step = 0.025
    nu = 1000.0 + step * np.arange(n_wavenumbers, dtype=np.float64)
    temps = np.arange(100.0, 3001.0, 100.0)
    press = 10.0 ** np.arange(-6.0, 2.5, 1.0)
    names = ABSORBERS[:6]
    species = [line_table(rng, s, s, nu, temps, press, n_lines=3000, backend=backend) for s in names]
    species += [cia_table(rng, "CIA-H2-H2", "CIA/H2-H2", nu, temps),
                cia_table(rng, "CIA-H2-He", "CIA/H2-He", nu, temps)]
    lam_lo, lam_hi = 1e4 / nu[-1], 1e4 / nu[0]
    # the bands tile the whole list: every wavenumber is sampled
    edges = np.linspace(lam_lo * (1 + 1e-7), lam_hi * (1 - 1e-7), nb_bins + 1)
    scale = (RADIUS_JUPITER / (10 * PARSEC)) ** 2
    obs = []
    for b in range(nb_bins):
        lo, hi = edges[b], edges[b + 1]
        w, r = bandpass(rng, lo, hi, 50)
        d, e = noisy(rng, blackbody_flux_lambda(np.array([(lo + hi) / 2]), 1100.0, scale))
        obs.append(ObservationSpec(f"band{b:03d}", "photometry", bin_edges=np.array([[lo, hi]]),
                                   data=d, error=e, filter=("Photon" if b % 2 else "Energy", w, r)))
    params = emission_parameters(rng, 6, 7)
    return Case("emission_c2", nu, step, 100, 100.0, 1e-6, ("poly", 6, 1),
                [("iso", names), ("bg", ["H2He"])], species, ["H2"], obs, params)


def transmission_parameters(rng, nb_absorbers, isothermal=True):
    log_g = rng.uniform(2.5, 4.0)
    radius = RADIUS_JUPITER * rng.uniform(0.9, 1.1)
    star = RADIUS_SUN
    mixing = 10.0 ** rng.uniform(-8.0, -3.0, nb_absorbers)
    t = [rng.uniform(500.0, 2500.0)] if isothermal else list(poly_temperature_params(rng, 7))
    return np.concatenate([[log_g, radius, star], mixing, t])


def lsf_edges(lam_min, lam_max, nb, r, low_end=True, high_end=True, margin=1e-6):
    """Band edges (micron, descending) of nb equal bins over [lo, hi] whose
    observation range, widened by 5 sigma of a Gaussian line-spread function
    of FWHM = centre / r at the outermost bins (observations.cpp:236-251),
    reaches lam_min (1 + margin) and lam_max (1 - margin) at the ends
    flagged, so that the sampled grid runs to the ends of the wavenumber
    list; an unflagged end is that bound itself."""
    a = 5.0 / (r * 2.355)
    lo, hi = lam_min, lam_max
    for _ in range(50):            # the outer bins' centres depend on both ends
        w = (hi - lo) / nb
        if low_end:
            lo = (lam_min * (1 + margin) + a * w / 2) / (1 - a)
        if high_end:
            hi = (lam_max * (1 - margin) + a * w / 2) / (1 + a)
    edges = np.linspace(hi, lo, nb + 1)
    return np.stack([edges[:-1], edges[1:]], axis=1)


def _transit_data(lo, hi, nb, r, rng):
    edges = np.linspace(hi, lo, nb + 1)
    be = np.stack([edges[:-1], edges[1:]], axis=1)
    centres = be.mean(axis=1)
    depth = 1.2e4 * np.ones(nb)
    return be, depth, 50.0 * np.ones(nb), centres / r


def transmission_small(seed=31, variant="plain"):
    """Transmission case for the tests: 40 levels 1e2..1e-6 bar, isothermal,
    2 absorbers + H2-H2 / H2-He CIA + H2 Rayleigh, 3000-8000 cm^-1, three
    band-spectroscopy instruments with Gaussian line-spread functions.
    variant: "plain", "variable_gravity", "fit_mmw" (mean molecular weight
    as a parameter) or "fit_scale_height" (scale height in km as a parameter)."""
    rng = np.random.default_rng(seed)
    nu = _grid(3000.0, 8000.0, 1.0)
    temps = np.arange(100.0, 3001.0, 300.0)
    press = 10.0 ** np.arange(-6.0, 2.5, 2.0)
    species = [
        line_table(rng, "H2O", "H2O", nu, temps, press, n_lines=600),
        line_table(rng, "CH4", "CH4", nu, temps, press, n_lines=500, log_storage=True),
        cia_table(rng, "CIA-H2-H2", "CIA/H2-H2", nu, temps),
        cia_table(rng, "CIA-H2-He", "CIA/H2-He", nu, temps),
    ]
    obs = []
    for k, (lo, hi, nb, r) in enumerate([(1.30, 1.65, 8, 70.0), (1.75, 2.45, 15, 130.0), (2.55, 3.20, 25, 300.0)]):
        be, d, e, fw = _transit_data(lo, hi, nb, r, rng)
        obs.append(ObservationSpec(f"instrument{k}", "band-spectroscopy", bin_edges=be, data=d, error=e, fwhm=fw,
                                   modifier="shift_const" if k == 1 else "none"))
    params = transmission_parameters(rng, 2)
    extra = {"variant": variant}
    if variant == "fit_mmw":
        params = np.concatenate([params[:3], [2.3], params[3:]])
    if variant == "fit_scale_height":
        params = np.concatenate([params[:3], [400.0], params[3:]])
    params = np.concatenate([params, [0.0]])
    c = Case(f"transmission_{variant}", nu, 2.0, 40, 100.0, 1e-6, ("const",),
             [("iso", ["H2O", "CH4"]), ("bg", ["H2He"])], species, ["H2"], obs, params, model="transmission",
             extra=extra)
    return c


def eclipse_small(seed=41):
    """Secondary-eclipse case for the tests: 30 levels, poly 6 1, 2 absorbers
    + CIA, a blackbody star, three photometric bandpasses and a 20-bin
    band spectrum with a line-spread function."""
    rng = np.random.default_rng(seed)
    nu = _grid(2500.0, 7000.0, 1.0)
    temps = np.arange(100.0, 3001.0, 300.0)
    press = 10.0 ** np.arange(-6.0, 2.5, 2.0)
    species = [
        line_table(rng, "H2O", "H2O", nu, temps, press, n_lines=600),
        line_table(rng, "CO", "CO", nu, temps, press, n_lines=300, mass=MOLECULAR_WEIGHT["CO"]),
        cia_table(rng, "CIA-H2-H2", "CIA/H2-H2", nu, temps),
        cia_table(rng, "CIA-H2-He", "CIA/H2-He", nu, temps),
    ]
    obs = []
    for k, (lo, hi, det) in enumerate([(1.6, 1.9, "Photon"), (2.1, 2.4, "Energy"), (3.2, 3.8, "Photon")]):
        w, r = bandpass(rng, lo, hi, 50)
        obs.append(ObservationSpec(f"phot{k}", "photometry", bin_edges=np.array([[lo, hi]]),
                                   data=np.array([1500.0]), error=np.array([80.0]), filter=(det, w, r)))
    edges = np.linspace(1.55, 3.6, 21)
    be = np.stack([edges[:-1], edges[1:]], axis=1)
    obs.append(ObservationSpec("spectrum", "band-spectroscopy", bin_edges=be, data=np.full(20, 1500.0),
                               error=np.full(20, 80.0), fwhm=be.mean(axis=1) / 200.0))
    rng2 = np.random.default_rng(seed + 1)
    params = np.concatenate([[rng2.uniform(3.0, 4.0), rng2.uniform(0.08, 0.15), rng2.uniform(4500, 6500)],
                             10.0 ** rng2.uniform(-8.0, -3.0, 2), poly_temperature_params(rng2, 7, 800.0, 2800.0)])
    return Case("eclipse_small", nu, 2.0, 30, 100.0, 1e-6, ("poly", 6, 1),
                [("iso", ["H2O", "CO"]), ("bg", ["H2He"])], species, [], obs, params, model="secondary_eclipse",
                extra={"stellar_model": "blackbody"})


def transmission_c3(seed=53, backend="numpy", n_wavenumbers=300_000, nb_levels=200):
    """Batched transmission configuration: 200 levels 1e2..1e-6 bar (200
    tangent paths), isothermal, 8 absorbers + H2-H2/H2-He CIA + H2 Rayleigh
    on 30 T x 9 p tables, 300,000 wavenumbers at 0.025 cm^-1 from 2000 cm^-1,
    three instruments of uneven size: 20, 147 and 400 bins with Gaussian
    line-spread functions at R = 70, 130 and 1000, side by side over the
    whole list (all 300,000 wavenumbers are sampled)."""
    rng = np.random.default_rng(seed)
    step = 0.025
    nu = 2000.0 + step * np.arange(n_wavenumbers, dtype=np.float64)
    temps = np.arange(100.0, 3001.0, 100.0)
    press = 10.0 ** np.arange(-6.0, 2.5, 1.0)
    names = ABSORBERS[:8]
    species = [line_table(rng, s, s, nu, temps, press, n_lines=3000, backend=backend) for s in names]
    species += [cia_table(rng, "CIA-H2-H2", "CIA/H2-H2", nu, temps),
                cia_table(rng, "CIA-H2-He", "CIA/H2-He", nu, temps)]
    obs = []
    lam_lo, lam_hi = 1e4 / nu[-1], 1e4 / nu[0]
    # three instruments side by side over the whole list (the outer ones'
    # line-spread margins reach its ends): every wavenumber is sampled
    for k, (lo, hi, nb, r) in enumerate([(lam_lo, 1.70, 20, 70.0), (1.70, 3.30, 147, 130.0),
                                         (3.30, lam_hi, 400, 1000.0)]):
        be = lsf_edges(lo, hi, nb, r, low_end=k == 0, high_end=k == 2)
        centres = be.mean(axis=1)
        obs.append(ObservationSpec(f"instrument{k}", "band-spectroscopy", bin_edges=be, data=1.2e4 * np.ones(nb),
                                   error=50.0 * np.ones(nb), fwhm=centres / r))
    params = transmission_parameters(rng, 8)
    return Case("transmission_c3", nu, step, nb_levels, 100.0, 1e-6, ("const",),
                [("iso", names), ("bg", ["H2He"])], species, ["H2"], obs, params, model="transmission",
                extra={"variant": "plain"})


def emission_c4(seed=29, backend="numpy"):
    """Adjoint configuration: the emission_c2 model (100 levels, 6 absorbers
    + 2 CIA = 8 tabulated species, 7 temperature control points, 300 bins) at
    200,000 wavenumbers; 16 parameters (log g, radius and distance scale, 6
    log abundances, 7 temperature control points)."""
    c = emission_c2(seed=seed, backend=backend, n_wavenumbers=200_000)
    c.name = "emission_c4"
    return c


def eclipse_c5(seed=61, backend="numpy", n_wavenumbers=1_000_000):
    """Reduced-precision batched secondary-eclipse configuration: 100 levels,
    poly 6 1, 12 absorbers + H2-H2/H2-He CIA on 30 T x 12 p tables over
    1,000,000 wavenumbers at 0.025 cm^-1 from 1000 cm^-1, a blackbody star,
    a 9-bandpass photometric set tiling the whole list (all 1,000,000
    wavenumbers are sampled) and a 400-bin band spectrum over 1.2-5 micron
    with a Gaussian line-spread function at R = 400."""
    rng = np.random.default_rng(seed)
    step = 0.025
    nu = 1000.0 + step * np.arange(n_wavenumbers, dtype=np.float64)
    temps = np.arange(100.0, 3001.0, 100.0)
    press = 10.0 ** np.linspace(-6.0, 2.0, 12)
    names = ABSORBERS[:12]
    species = [line_table(rng, s, s, nu, temps, press, n_lines=3000, backend=backend) for s in names]
    species += [cia_table(rng, "CIA-H2-H2", "CIA/H2-H2", nu, temps),
                cia_table(rng, "CIA-H2-He", "CIA/H2-He", nu, temps)]
    lam_lo, lam_hi = 1e4 / nu[-1], 1e4 / nu[0]
    obs = []
    # nine bandpasses tiling the list in log wavelength (every wavenumber is sampled)
    bounds = np.geomspace(lam_lo * (1 + 1e-7), lam_hi * (1 - 1e-7), 10)
    for k in range(9):
        lo, hi = bounds[k], bounds[k + 1]
        w, r = bandpass(rng, lo, hi, 50)
        obs.append(ObservationSpec(f"phot{k}", "photometry", bin_edges=np.array([[lo, hi]]),
                                   data=np.array([1500.0]), error=np.array([80.0]),
                                   filter=("Photon" if k % 2 else "Energy", w, r)))
    edges = np.linspace(1.2, 5.0, 401)
    be = np.stack([edges[:-1], edges[1:]], axis=1)
    obs.append(ObservationSpec("spectrum", "band-spectroscopy", bin_edges=be, data=np.full(400, 1500.0),
                               error=np.full(400, 80.0), fwhm=be.mean(axis=1) / 400.0))
    rng2 = np.random.default_rng(seed + 1)
    params = np.concatenate([[rng2.uniform(3.0, 4.0), rng2.uniform(0.08, 0.15), rng2.uniform(4500, 6500)],
                             10.0 ** rng2.uniform(-8.0, -3.0, 12), poly_temperature_params(rng2, 7, 800.0, 2800.0)])
    return Case("eclipse_c5", nu, step, 100, 100.0, 1e-6, ("poly", 6, 1),
                [("iso", names), ("bg", ["H2He"])], species, [], obs, params, model="secondary_eclipse",
                extra={"stellar_model": "blackbody"})


def wide_bands(case: Case, seed, bands):
    """The case with broad photometric bands added: [(lo, hi, detector)] in
    micron, each covering thousands of grid points (bins whose folded segment
    is summed in several pieces)."""
    rng = np.random.default_rng(seed)
    for k, (lo, hi, det) in enumerate(bands):
        w, r = bandpass(rng, lo, hi, 50)
        ref = float(np.median(np.abs(case.observations[0].data)))
        case.observations.append(ObservationSpec(f"wide{k}", "photometry", bin_edges=np.array([[lo, hi]]),
                                                 data=np.array([ref]), error=np.array([0.05 * ref]),
                                                 filter=(det, w, r)))
    case.name = case.name + "_wide"
    return case


def resampled(case: Case, discretisation, resolution, name):
    """The case on a grid sampled at constant wavelength steps (micron,
    discretisation 1) or constant resolving power (discretisation 2)."""
    case.discretisation = discretisation
    case.spectral_resolution = resolution
    case.name = name
    return case


def cloudy(case: Case, name, clouds, values):
    """The case with cloud models (forward_model.config lines, e.g. "grey",
    "power_law 1.0 fb", "KHnongrey 1.0") and their parameter values, which
    go after the temperature parameters (emission.cpp:123-153)."""
    nb_mix = sum(len(c[1]) for c in case.chemistry if c[0] == "iso")
    tm = case.temperature_model
    nb_temp = 1 if tm[0] == "const" else tm[1] * tm[2] + 1
    at = nb_general_parameters(case) + nb_mix + nb_temp
    case.parameters = np.concatenate([case.parameters[:at], np.asarray(values, dtype=np.float64),
                                      case.parameters[at:]])
    case.clouds = list(clouds)
    case.name = name
    return case


def hminus_case(seed=23):
    """The H- continuum (bound-free below 1.6419 micron, free-free at every
    wavelength above 0.1823 micron) with H2O and H2-H2 CIA: a grid from
    0.5 to 5 micron (2,000-20,000 cm^-1 at 4 cm^-1), iso chemistry of H2O,
    H-, H and e- over the H2/He background, 16 levels, spectroscopy and
    band spectroscopy across the bound-free edge."""
    rng = np.random.default_rng(seed)
    nu = _grid(2000.0, 20000.0, 4.0)
    temps = np.array([1000.0, 1500.0, 2000.0, 2500.0, 3000.0])
    press = 10.0 ** np.arange(-5.0, 2.5, 2.5)
    species = [line_table(rng, "H2O", "H2O", nu, temps, press, n_lines=300, mass=MOLECULAR_WEIGHT["H2O"]),
               cia_table(rng, "CIA-H2-H2", "CIA/H2-H2", nu, temps)]
    scale = (RADIUS_JUPITER / (10 * PARSEC)) ** 2
    wl = np.linspace(0.6, 4.5, 40)
    d, e = noisy(rng, blackbody_flux_lambda(wl, 2200.0, scale))
    obs = [ObservationSpec("spec", "spectroscopy", wavelengths=wl, data=d, error=e)]
    edges = np.linspace(4.9, 0.55, 12)
    be = np.stack([edges[:-1], edges[1:]], axis=1)
    d, e = noisy(rng, blackbody_flux_lambda(be.mean(axis=1), 2200.0, scale))
    obs.append(ObservationSpec("bands", "band-spectroscopy", bin_edges=be, data=d, error=e))
    log_g = rng.uniform(4.5, 5.2)
    # H2O, H-, H, e- (the electrons and H- of a ~2000-3000 K atmosphere)
    mixing = np.array([10 ** rng.uniform(-4, -3), 10 ** rng.uniform(-9.5, -8.5), 10 ** rng.uniform(-3, -2),
                       10 ** rng.uniform(-6, -5)])
    theta = poly_temperature_params(rng, 7)
    theta[0] = 2900.0                               # hot enough for the continuum to matter
    params = np.concatenate([[log_g, 1.0, 10.0 * PARSEC], mixing, theta])
    case = Case("hminus", nu, 4.0, 16, 100.0, 1e-5, ("poly", 6, 1),
                [("iso", ["H2O", "H-", "H", "e-"]), ("bg", ["H2He"])], species, [], obs, params)
    case.continuum = ["H-"]
    return case


def transmission_hminus_case(seed=29):
    """Transmission through a 2,800 K isothermal atmosphere with the H-
    continuum: the grid of hminus_case (0.5-5 micron at 4 cm^-1), H2O, H2-H2
    CIA and H2 Rayleigh, iso chemistry of H2O, H-, H and e-, 30 levels
    1e2..1e-6 bar, two band-spectroscopy instruments with line-spread
    functions, one each side of the bound-free edge."""
    rng = np.random.default_rng(seed)
    nu = _grid(2000.0, 20000.0, 4.0)
    temps = np.array([1000.0, 1500.0, 2000.0, 2500.0, 3000.0])
    press = 10.0 ** np.arange(-6.0, 2.5, 2.0)
    species = [line_table(rng, "H2O", "H2O", nu, temps, press, n_lines=300, mass=MOLECULAR_WEIGHT["H2O"]),
               cia_table(rng, "CIA-H2-H2", "CIA/H2-H2", nu, temps)]
    obs = []
    for k, (lo, hi, nb, r) in enumerate([(0.65, 1.55, 18, 100.0), (1.8, 4.4, 26, 150.0)]):
        be, d, e, fw = _transit_data(lo, hi, nb, r, rng)
        obs.append(ObservationSpec(f"instrument{k}", "band-spectroscopy", bin_edges=be, data=d, error=e, fwhm=fw))
    log_g = rng.uniform(2.8, 3.4)
    radius = RADIUS_JUPITER * rng.uniform(1.2, 1.4)
    mixing = np.array([10 ** rng.uniform(-4, -3), 10 ** rng.uniform(-9.5, -8.5), 10 ** rng.uniform(-3, -2),
                       10 ** rng.uniform(-6, -5)])
    params = np.concatenate([[log_g, radius, RADIUS_SUN], mixing, [2800.0]])
    case = Case("transmission_hminus", nu, 4.0, 30, 100.0, 1e-6, ("const",),
                [("iso", ["H2O", "H-", "H", "e-"]), ("bg", ["H2He"])], species, ["H2"], obs, params,
                model="transmission", extra={"variant": "plain"})
    case.continuum = ["H-"]
    return case


MANY_ABSORBERS = ABSORBERS + ["C2H2", "Na", "SH", "OH", "MgO", "AlO", "CaO", "SiO"]


def many_species_case(seed=19):
    """Emission with 20 absorbers + H2-H2/H2-He CIA (22 tabulated species,
    beyond the 16 stencil slots of the register-cached kernels: the generic
    single-point kernel), small tables, 20 levels, 2,001 wavenumbers,
    spectroscopy and band spectroscopy."""
    rng = np.random.default_rng(seed)
    nu = _grid(2500.0, 6500.0, 2.0)
    temps = np.array([400.0, 900.0, 1400.0, 1900.0, 2400.0])
    press = 10.0 ** np.arange(-5.0, 2.5, 2.5)
    species = [line_table(rng, s, s, nu, temps, press, n_lines=150) for s in MANY_ABSORBERS]
    species += [cia_table(rng, "CIA-H2-H2", "CIA/H2-H2", nu, temps),
                cia_table(rng, "CIA-H2-He", "CIA/H2-He", nu, temps)]
    scale = (RADIUS_JUPITER / (10 * PARSEC)) ** 2
    wl = np.linspace(1.6, 3.9, 30)
    d, e = noisy(rng, blackbody_flux_lambda(wl, 1300.0, scale))
    obs = [ObservationSpec("spec", "spectroscopy", wavelengths=wl, data=d, error=e)]
    edges = np.linspace(3.95, 1.6, 8)
    be = np.stack([edges[:-1], edges[1:]], axis=1)
    d, e = noisy(rng, blackbody_flux_lambda(be.mean(axis=1), 1300.0, scale))
    obs.append(ObservationSpec("bands", "band-spectroscopy", bin_edges=be, data=d, error=e))
    params = emission_parameters(rng, len(MANY_ABSORBERS), 7)
    return Case("many_species", nu, 2.0, 20, 100.0, 1e-5, ("poly", 6, 1),
                [("iso", list(MANY_ABSORBERS)), ("bg", ["H2He"])], species, ["H2"], obs, params)


CASES = {"tiny": tiny_case, "emission_c1": emission_c1, "emission_c2": emission_c2,
         "many_species": many_species_case, "hminus": hminus_case, "transmission_hminus": transmission_hminus_case,
         # tiny at constant wavelength steps of 2 nm and at R = 1500
         "tiny_const_wavelength": lambda: resampled(tiny_case(), 1, 2e-3, "tiny_const_wavelength"),
         "tiny_const_resolution": lambda: resampled(tiny_case(), 2, 1500.0, "tiny_const_resolution"),
         # emission_c1 grid: 1000-10000 cm^-1 at 0.45; bands of 14,800 and 2,560 points
         "emission_wide": lambda: wide_bands(emission_c1(), 71, [(1.25, 7.5, "Photon"), (2.0, 2.6, "Energy")]),
         # eclipse_small: 2500-7000 cm^-1 sampled every 2 cm^-1; a band of 2,180 points
         "eclipse_wide": lambda: wide_bands(eclipse_small(), 73, [(1.45, 3.95, "Energy")]),
         "transmission_plain": lambda: transmission_small(variant="plain"),
         "transmission_variable_gravity": lambda: transmission_small(variant="variable_gravity"),
         "transmission_fit_mmw": lambda: transmission_small(variant="fit_mmw"),
         "transmission_fit_scale_height": lambda: transmission_small(variant="fit_scale_height"),
         "eclipse_small": eclipse_small,
         # grey cloud [tau, p_top, p_bottom / p_top] and a power-law cloud one
         # scale height thick [tau, exponent, p_top] (lambda_ref 1 micron)
         "cloudy_tiny": lambda: cloudy(tiny_case(), "cloudy_tiny", ["grey", "power_law 1.0 fb"],
                                       [2.0, 0.1, 20.0, 1.5, -2.0, 1e-3]),
         # KHnongrey [tau, q0, a0, particle size (micron), p_top, p_bottom / p_top]
         # (lambda_ref 1 micron) and a grey cloud one scale height thick
         "cloudy_eclipse": lambda: cloudy(eclipse_small(), "cloudy_eclipse", ["KHnongrey 1.0", "grey fb"],
                                          [3.0, 10.0, 4.0, 0.5, 0.05, 30.0, 1.0, 1e-2]),
         "transmission_cloudy": lambda: cloudy(transmission_small(variant="plain"), "transmission_cloudy",
                                               ["grey", "KHnongrey 2.0 fb"],
                                               [5.0, 1e-2, 50.0, 2.0, 8.0, 4.5, 0.3, 1e-4])}


# ---------------------------------------------------------------------------
# writing the reference's file formats
# ---------------------------------------------------------------------------

def _fmt(x):
    return repr(float(x))


def write_case(case: Case, folder: str, tables: bool = True) -> str:
    """Write ``case`` as a retrieval folder in the reference's formats.

    Returns the folder path (with a trailing slash).  Besides the reference
    files, ``case.json`` records the configuration a driver needs
    (forward-model settings, opacity entries, parameter vector).  With
    tables=False the opacity folders are left out (the tables are then
    registered in memory with the engine and the oracle)."""
    folder = os.path.join(folder, "")
    os.makedirs(folder, exist_ok=True)
    opac = os.path.join(folder, "opacity", "")
    with open(folder + "wavenumber_full.dat", "w") as f:
        f.write(f"{case.wavenumber_full.size}\n")
        f.write("\n".join(_fmt(v) for v in case.wavenumber_full))
        f.write("\n")
    for s in (case.species if tables else []):
        d = os.path.join(opac, s.folder)
        os.makedirs(d, exist_ok=True)
        lines = [f"{_fmt(s.mass)} {'log' if s.log_storage else 'lin'}"]
        for it, t in enumerate(s.temperatures):
            for ip, p in enumerate(s.pressures):
                fname = f"row_T{it:02d}_p{ip:02d}.bin"
                s.values[it, ip].astype("<f4").tofile(os.path.join(d, fname))
                lines.append(f"{_fmt(p)} {_fmt(t)} {fname}")
        # rows listed in a shuffled-by-construction order (p-major) to exercise
        # the table ordering of the reader
        head, rows = lines[0], lines[1:]
        rows = rows[1::2] + rows[0::2]
        with open(os.path.join(d, "filelist.dat"), "w") as f:
            f.write(head + "\n" + "\n".join(rows) + "\n")
    names = []
    for i, o in enumerate(case.observations):
        fname = f"obs_{i:03d}.dat"
        names.append((fname, o.modifier))
        with open(folder + fname, "w") as f:
            f.write("#observation description\n#name\n" + o.name + "\n\n#type\n" + o.type + "\n\n")
            f.write("#filter transmission function\n")
            if o.filter is not None:
                fil = f"filter_{i:03d}.dat"
                f.write(fil + "\n\n")
                det, w, r = o.filter
                with open(folder + fil, "w") as ff:
                    ff.write("# synthetic bandpass\n\n#Detector Type\n" +
                             ("Photon counter" if det == "Photon" else "Energy counter") +
                             "\n\n# Wave (microns)     Filter Response\n")
                    for a, b in zip(w, r):
                        ff.write(f"{_fmt(a)}\t{_fmt(b)}\n")
            else:
                f.write("None\n\n")
            if o.type == "spectroscopy":
                f.write("#wl  flux  error  fwhm  weight\n")
                for j in range(o.wavelengths.size):
                    cols = [o.wavelengths[j], o.data[j], o.error[j]]
                    if o.fwhm is not None or o.weight is not None:
                        cols.append(o.fwhm[j] if o.fwhm is not None else 0.0)
                    if o.weight is not None:
                        cols.append(o.weight[j])
                    f.write("  ".join(_fmt(c) for c in cols) + "\n")
            elif o.type == "band-spectroscopy":
                # with a profile column, write the weight column too: the
                # reader takes a 5-column line's last field as a weight
                # (observations_load_data.cpp:180-191), which would be 0 here
                f.write("#left  right  value  error  fwhm  weight\n")
                for j in range(o.bin_edges.shape[0]):
                    cols = [o.bin_edges[j, 0], o.bin_edges[j, 1], o.data[j], o.error[j]]
                    if o.fwhm is not None or o.weight is not None:
                        cols.append(o.fwhm[j] if o.fwhm is not None else 0.0)
                        cols.append(o.weight[j] if o.weight is not None else 1.0)
                    f.write("  ".join(_fmt(c) for c in cols) + "\n")
            else:
                f.write("#lower  upper  value  error\n")
                cols = [o.bin_edges[0, 0], o.bin_edges[0, 1], o.data[0], o.error[0]]
                if o.weight is not None:
                    cols.append(o.weight[0])
                f.write("  ".join(_fmt(c) for c in cols) + "\n")
    with open(folder + "observations.list", "w") as f:
        for fname, mod in names:
            f.write(f"{fname} {mod}\n")
    meta = {
        "name": case.name,
        "model": case.model,
        "spectral_resolution": case.spectral_resolution,
        "nb_levels": case.nb_levels,
        "p_bottom": case.p_bottom,
        "p_top": case.p_top,
        "temperature_model": list(case.temperature_model),
        "chemistry": [[c[0], list(c[1])] for c in case.chemistry],
        "opacity": [[s, f] for s, f in case.opacity_entries],
        "observations": [n for n, _ in names],
        "modifiers": [m for _, m in names],
        "parameters": [float(v) for v in case.parameters],
        "use_error_inflation": case.use_error_inflation,
        "extra": case.extra,
        "discretisation": case.discretisation,
        "clouds": list(case.clouds),
    }
    with open(folder + "case.json", "w") as f:
        json.dump(meta, f, indent=1)
    # the same settings as plain "key values..." lines for C/C++ drivers
    with open(folder + "case.txt", "w") as f:
        f.write(f"model {case.model}\n")
        f.write(f"spectral_resolution {_fmt(case.spectral_resolution)}\n")
        f.write(f"spectral_discretisation {case.discretisation}\n")
        f.write(f"nb_levels {case.nb_levels}\n")
        f.write(f"p_bottom {_fmt(case.p_bottom)}\n")
        f.write(f"p_top {_fmt(case.p_top)}\n")
        f.write("temperature " + " ".join(str(v) for v in case.temperature_model) + "\n")
        for kind, sp in case.chemistry:
            f.write(f"chemistry {kind} " + " ".join(sp) + "\n")
        for s, fo in case.opacity_entries:
            f.write(f"opacity {s} {fo}\n")
        for c in case.clouds:
            f.write(f"cloud {c}\n")
        f.write(f"error_inflation {int(case.use_error_inflation)}\n")
        variant = case.extra.get("variant", "plain")
        f.write(f"use_variable_gravity {int(variant == 'variable_gravity')}\n")
        f.write(f"fit_mean_molecular_weight {int(variant == 'fit_mmw')}\n")
        f.write(f"fit_scale_height {int(variant == 'fit_scale_height')}\n")
        if "stellar_model" in case.extra:
            f.write(f"stellar_model {case.extra['stellar_model']}\n")
        f.write(f"parameters {case.parameters.size} " + " ".join(_fmt(v) for v in case.parameters) + "\n")
    return folder


def write_direct_case(case: Case, folder: str, wavelength_range, stellar=None) -> str:
    """The case's tables and wavenumber list for pyBeAR's direct models: a
    grid sampled over wavelength_range (micron) without observations, and
    for the eclipse a tabulated stellar flux (wavelengths in micron, flux in
    W m^-2 micron^-1) written to stellar_spectrum.dat."""
    folder = write_case(case, folder)
    with open(folder + "case.txt", "a") as f:
        f.write(f"direct_range {_fmt(wavelength_range[0])} {_fmt(wavelength_range[1])}\n")
        if stellar is not None:
            f.write("stellar_spectrum stellar_spectrum.dat\n")
    if stellar is not None:
        with open(folder + "stellar_spectrum.dat", "w") as f:
            for a, b in zip(*stellar):
                f.write(f"{_fmt(a)} {_fmt(b)}\n")
    return folder


def direct_structure(rng, nb_levels, species, p_bottom=100.0, p_top=1e-6):
    """A caller-made atmosphere for the direct models: pressures log-spaced
    from the bottom (bar), a smooth temperature profile, per-level mixing
    ratios of the given species (H2 and He fill most of the rest)."""
    p = np.logspace(np.log10(p_bottom), np.log10(p_top), nb_levels)
    x = np.linspace(0.0, 1.0, nb_levels)
    t = 1800.0 - 900.0 * x + 150.0 * np.sin(3.0 * x + rng.uniform(0, 1))
    trace = [s for s in species if s not in ("H2", "He")]
    mix = np.zeros((nb_levels, len(species)))
    for j, s in enumerate(species):
        if s in trace:
            mix[:, j] = 10.0 ** (rng.uniform(-6.0, -3.5) + 0.5 * np.sin(2.0 * x + j))
    rest = 1.0 - mix.sum(axis=1)
    for j, s in enumerate(species):
        if s == "H2":
            mix[:, j] = rest * 0.85
        if s == "He":
            mix[:, j] = rest * 0.15
    return p, t, mix


# ---------------------------------------------------------------------------
# engine construction from a case
# ---------------------------------------------------------------------------

def _model_config(meta):
    from . import api
    chem = [" ".join([c[0]] + list(c[1])) for c in meta["chemistry"]]
    variant = meta.get("extra", {}).get("variant", "plain")
    return api.ModelConfig(meta["nb_levels"], meta["p_bottom"], meta["p_top"],
                           " ".join(str(v) for v in meta["temperature_model"]), chem,
                           [tuple(o) for o in meta["opacity"]],
                           fit_mean_molecular_weight=variant == "fit_mmw",
                           fit_scale_height=variant == "fit_scale_height",
                           use_variable_gravity=variant == "variable_gravity",
                           cloud_model=list(meta.get("clouds", [])))


def open_retrieval(folder, device=0, table_storage="fp64"):
    """Engine Retrieval over a case folder written by write_case (files only)."""
    from . import api
    folder = os.path.join(folder, "")
    meta = json.load(open(folder + "case.json"))
    cfg = api.GlobalConfig(retrieval_folder_path=folder, cross_section_file_path=folder + "opacity/",
                           wavenumber_file_path=folder + "wavenumber_full.dat",
                           spectral_disecretisation=DISCRETISATIONS[meta.get("discretisation", 0)],
                           spectral_resolution=meta["spectral_resolution"],
                           use_error_inflation=meta["use_error_inflation"], device=device,
                           table_storage=table_storage)
    return api.Retrieval(cfg, _model_config(meta), observation_files=meta["observations"],
                         observation_modifiers=meta["modifiers"], forward_model_type=meta["model"])


DISCRETISATIONS = ("const_wavenumber", "const_wavelength", "const_resolution")


def case_meta(case: Case):
    return {
        "spectral_resolution": case.spectral_resolution, "nb_levels": case.nb_levels,
        "discretisation": case.discretisation,
        "p_bottom": case.p_bottom, "p_top": case.p_top,
        "temperature_model": list(case.temperature_model),
        "chemistry": [[c[0], list(c[1])] for c in case.chemistry],
        "opacity": [[s, f] for s, f in case.opacity_entries],
        "use_error_inflation": case.use_error_inflation,
        "model": case.model,
        "extra": case.extra,
        "clouds": list(case.clouds),
    }


def observation_inputs(case: Case):
    from . import api
    out = []
    for o in case.observations:
        kw = dict(name=o.name, type=o.type, data=o.data, data_error=o.error,
                  likelihood_weight=o.weight, instrument_profile_fwhm=o.fwhm,
                  spectrum_modifier_id=o.modifier)
        if o.type == "spectroscopy":
            kw["wavelengths"] = o.wavelengths
        else:
            kw["bin_wavelength_edges"] = o.bin_edges
        if o.filter is not None:
            kw["filter_response"] = (o.filter[1], o.filter[2])
            kw["filter_detector_type"] = o.filter[0].lower()
        out.append(api.ObservationInput(**kw))
    return out


def open_retrieval_in_memory(case: Case, scratch_folder, device=0):
    """Engine Retrieval with tables registered in memory and observations
    passed as ObservationInput (only the wavenumber list is written)."""
    from . import api
    os.makedirs(scratch_folder, exist_ok=True)
    wn = os.path.join(scratch_folder, "wavenumber_full.dat")
    with open(wn, "w") as f:
        f.write(f"{case.wavenumber_full.size}\n")
        f.write("\n".join(_fmt(v) for v in case.wavenumber_full))
        f.write("\n")
    api.clear_opacity_data()
    for s in case.species:
        t = np.repeat(s.temperatures, s.pressures.size)
        p = np.tile(s.pressures, s.temperatures.size)
        api.register_opacity_data(s.folder, t, p, s.values.reshape(-1, s.values.shape[-1]),
                                  species_mass=s.mass, log_storage=s.log_storage)
    meta = case_meta(case)
    cfg = api.GlobalConfig(retrieval_folder_path=scratch_folder, cross_section_file_path="",
                           wavenumber_file_path=wn, spectral_resolution=meta["spectral_resolution"],
                           spectral_disecretisation=DISCRETISATIONS[case.discretisation],
                           use_error_inflation=meta["use_error_inflation"], device=device)
    r = api.Retrieval(cfg, _model_config(meta), observations=observation_inputs(case), forward_model_type=case.model)
    api.clear_opacity_data()
    return r


def inject_data(case: Case, model_bins, seed=99, frac=0.02):
    """Replace the observed data by the given model bins (one array per
    observation, in each observation's input order) with 2 % Gaussian noise."""
    rng = np.random.default_rng(seed)
    for o, m in zip(case.observations, model_bins):
        o.data, o.error = noisy(rng, np.asarray(m, dtype=np.float64), frac)


def nb_general_parameters(case: Case):
    """General parameters ahead of the chemistry: emission [log g, f, d],
    transmission [log g, R_p, R_*, (mu | H)], eclipse [log g, Rp/Rs, T_star]."""
    if case.model == "transmission":
        return 3 + (1 if case.extra.get("variant") in ("fit_mmw", "fit_scale_height") else 0)
    return 3


def perturbed(params, case: Case, seed=5, scale=1.0):
    """A nearby live point: small shifts of the physical parameters."""
    rng = np.random.default_rng(seed)
    p = np.array(params, dtype=np.float64)
    g0 = nb_general_parameters(case)
    nb_mix = sum(len(c[1]) for c in case.chemistry if c[0] == "iso")
    p[0] += 0.02 * scale * rng.standard_normal()
    if case.model == "transmission":
        p[1] *= 1 + 0.002 * scale * rng.standard_normal()
    if case.model == "secondary_eclipse":
        p[1] *= 1 + 0.01 * scale * rng.standard_normal()
    p[g0:g0 + nb_mix] *= np.exp(0.1 * scale * rng.standard_normal(nb_mix))
    p[g0 + nb_mix] *= 1 + 0.01 * scale * rng.standard_normal()
    return p
