// C ABI (include/bear_hip.h) over the host C++ classes.
#include "bear_hip.h"

#include <cstring>
#include <new>
#include <sstream>
#include <string>

#include "model.h"

struct bear_retrieval {
  std::unique_ptr<bear_amd::Retrieval> impl;
};

struct bear_spectrum_model {
  std::unique_ptr<bear_amd::DirectModel> impl;
};

namespace {

thread_local std::string g_last_error;

template <class F>
int guarded(F&& f) {
  try {
    g_last_error.clear();
    f();
    return BEAR_OK;
  } catch (const bear_amd::Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "out of host memory";
    return BEAR_ERROR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return BEAR_ERROR_INTERNAL;
  }
}

std::string str(const char* s) { return s ? std::string(s) : std::string(); }

std::vector<double> vec(const double* p, size_t n) { return p ? std::vector<double>(p, p + n) : std::vector<double>(); }

bear_amd::GlobalSettings settings(const bear_global_config* config) {
  bear_amd::GlobalSettings g;
  g.device = config->device;
  g.retrieval_folder = str(config->retrieval_folder_path);
  g.cross_section_path = str(config->cross_section_file_path);
  g.wavenumber_file = str(config->wavenumber_file_path);
  g.discretisation = config->spectral_discretisation;
  g.resolution = config->spectral_resolution;
  g.error_inflation = config->use_error_inflation != 0;
  if (config->table_storage < BEAR_TABLE_FP64 || config->table_storage > BEAR_TABLE_INT16)
    bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "unknown table storage " + std::to_string(config->table_storage));
  g.table_storage = static_cast<bear_amd::TablePrecision>(config->table_storage);
  return g;
}

std::vector<std::pair<std::string, std::string>> opacityEntries(const char* text) {
  std::vector<std::pair<std::string, std::string>> out;
  std::istringstream op(str(text));
  std::string line;
  while (std::getline(op, line)) {
    std::istringstream in(line);
    std::string sym, folder;
    if (!(in >> sym)) continue;
    if (!(in >> folder)) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "opacity species " + sym + " without a folder");
    out.emplace_back(sym, folder);
  }
  return out;
}

// the arguments shared by the three calcSpectrum entries, checked; the
// species ids of FixedChemistry
std::vector<int> structureSpecies(const bear_spectrum_model* m, const double* pressure, const double* temperature,
                                  size_t nb_levels, const char* const* species, size_t nb_species,
                                  const double* mixing_ratios, const double* cloud, double* spectrum) {
  if (!m || !pressure || !temperature || !spectrum || (nb_species && (!species || !mixing_ratios)))
    bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "null argument");
  if (nb_levels != m->impl->nbLevels())
    bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "expected " + std::to_string(m->impl->nbLevels()) + " levels, got " +
                                                  std::to_string(nb_levels));
  // EmissionModel::setCloudProperties and its transmission and eclipse
  // forms (emission.cpp:344-367): the caller's layer optical depths, used
  // when any is positive
  m->impl->engine().setFixedCloud(cloud);
  std::vector<int> ids;
  for (size_t j = 0; j < nb_species; ++j) ids.push_back(bear_amd::speciesIdOrFastChem(str(species[j])));
  return ids;
}

}  // namespace

extern "C" {

const char* bear_last_error(void) { return g_last_error.c_str(); }

const char* bear_version(void) { return "bear-mi355x 0.1 (gfx950)"; }

int bear_opacity_data_register(const char* folder, double species_mass, int log_storage, size_t nb_rows,
                               const double* temperatures, const double* pressures, size_t nb_wavenumbers,
                               const float* rows) {
  return guarded([&] {
    if (!folder || !temperatures || !pressures || !rows || nb_rows == 0 || nb_wavenumbers == 0)
      bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "bear_opacity_data_register: empty table");
    bear_amd::registerTable(folder, species_mass, log_storage != 0, nb_rows, temperatures, pressures, nb_wavenumbers,
                            rows);
  });
}

int bear_opacity_data_clear(void) {
  return guarded([] { bear_amd::clearRegisteredTables(); });
}

int bear_retrieval_create(const bear_global_config* config, const char* forward_model_type,
                          const bear_model_config* model_config, const bear_observation_input* observations,
                          size_t nb_observations, const char* const* obs_files, const char* const* obs_modifiers,
                          size_t nb_obs_files, bear_retrieval** out) {
  return guarded([&] {
    if (!config || !model_config || !out) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "null argument");
    *out = nullptr;
    if (!config->use_gpu) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "this engine runs on the GPU only (use_gpu = 1)");
    const bear_amd::GlobalSettings g = settings(config);

    bear_amd::ModelSettings m;
    m.nb_levels = model_config->nb_grid_points;
    m.bottom_pressure = model_config->atmos_bottom_pressure;
    m.top_pressure = model_config->atmos_top_pressure;
    m.temperature = str(model_config->temperature_profile);
    m.radiative_transfer = str(model_config->radiative_transfer);
    m.chemistry = str(model_config->chemistry);
    m.fit_mean_molecular_weight = model_config->fit_mean_molecular_weight != 0;
    m.fit_scale_height = model_config->fit_scale_height != 0;
    m.use_variable_gravity = model_config->use_variable_gravity != 0;
    if (model_config->stellar_spectrum_model) m.stellar_model = model_config->stellar_spectrum_model;
    m.opacity = opacityEntries(model_config->opacity_species);
    m.clouds = bear_amd::parseCloudModels(str(model_config->cloud_model));

    std::vector<bear_amd::ObservationInputView> inputs;
    for (size_t i = 0; i < nb_observations; ++i) {
      const bear_observation_input& o = observations[i];
      bear_amd::ObservationInputView v;
      v.name = str(o.name);
      v.type = str(o.type);
      v.modifier = str(o.spectrum_modifier);
      const size_t n = o.nb_points;
      if (!o.data || !o.data_error || n == 0) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "observation without data");
      v.data = vec(o.data, n);
      v.error = vec(o.data_error, n);
      v.weight = vec(o.likelihood_weight, n);
      v.fwhm = vec(o.instrument_profile_fwhm, n);
      if (v.type == "spectroscopy") {
        if (!o.wavelengths) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "spectroscopy needs wavelengths");
        v.wavelengths = vec(o.wavelengths, n);
      } else {
        if (!o.bin_edges) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, v.type + " needs bin edges");
        for (size_t b = 0; b < n; ++b) v.bin_edges.push_back({o.bin_edges[2 * b], o.bin_edges[2 * b + 1]});
      }
      if (o.filter_nb_points > 0) {
        v.filter_wavelength = vec(o.filter_wavelength, o.filter_nb_points);
        v.filter_response = vec(o.filter_response, o.filter_nb_points);
        v.detector = str(o.filter_detector_type);
      }
      inputs.push_back(std::move(v));
    }
    std::vector<std::string> files, mods;
    for (size_t i = 0; i < nb_obs_files; ++i) {
      files.push_back(str(obs_files[i]));
      mods.push_back(obs_modifiers ? str(obs_modifiers[i]) : std::string());
    }
    auto* r = new bear_retrieval;
    try {
      r->impl.reset(new bear_amd::Retrieval(g, str(forward_model_type), m, inputs, files, mods));
    } catch (...) {
      delete r;
      throw;
    }
    *out = r;
  });
}

void bear_retrieval_destroy(bear_retrieval* r) { delete r; }

size_t bear_retrieval_nb_parameters(const bear_retrieval* r) { return r ? r->impl->nbParameters() : 0; }
size_t bear_retrieval_nb_spectral_points(const bear_retrieval* r) { return r ? r->impl->grid().size() : 0; }
size_t bear_retrieval_nb_observations(const bear_retrieval* r) { return r ? r->impl->observations().size() : 0; }
size_t bear_retrieval_nb_observation_points(const bear_retrieval* r, size_t i) {
  return (r && i < r->impl->observations().size()) ? r->impl->observations()[i].nbPoints() : 0;
}

int bear_retrieval_spectral_grid(const bear_retrieval* r, double* wavenumbers, double* wavelengths) {
  return guarded([&] {
    if (!r) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "null retrieval");
    const auto& g = r->impl->grid();
    if (wavenumbers) std::memcpy(wavenumbers, g.wavenumber.data(), g.size() * sizeof(double));
    if (wavelengths) std::memcpy(wavelengths, g.wavelength.data(), g.size() * sizeof(double));
  });
}

int bear_retrieval_compute_likelihood(bear_retrieval* r, const double* parameters, size_t n, double* log_likelihood) {
  return guarded([&] {
    if (!r || !parameters || !log_likelihood) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "null argument");
    *log_likelihood = r->impl->computeLikelihood(parameters, n);
  });
}

int bear_retrieval_compute_model(bear_retrieval* r, const double* parameters, size_t n, int return_high_res,
                                 double* spectrum, double* spectrum_obs, int* neglect_model) {
  return guarded([&] {
    if (!r || !parameters || !spectrum_obs) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "null argument");
    const bool neglect = r->impl->computeModel(parameters, n, return_high_res != 0 && spectrum, spectrum, spectrum_obs);
    if (neglect_model) *neglect_model = neglect;
  });
}

int bear_forward_model_calc_model_gpu(bear_retrieval* r, const double* parameters, size_t n, double* spectrum_dev,
                                      double* const* spectrum_obs_dev, int* neglect_model) {
  return guarded([&] {
    if (!r || !parameters) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "null argument");
    const bool neglect = r->impl->calcModelGPU(parameters, n, spectrum_dev, spectrum_obs_dev);
    if (neglect_model) *neglect_model = neglect;
  });
}

int bear_forward_model_calc_model_gpu_noalloc(bear_retrieval* r, const double* parameters, size_t n,
                                              double* spectrum_dev, double* const* spectrum_obs_dev,
                                              int* neglect_model) {
  return guarded([&] {
    if (!r || !parameters) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "null argument");
    const bool neglect = r->impl->calcModelGPUNoAlloc(parameters, n, spectrum_dev, spectrum_obs_dev);
    if (neglect_model) *neglect_model = neglect;
  });
}

int bear_retrieval_set_table_precision(bear_retrieval* r, int precision) {
  return guarded([&] {
    if (!r) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "null retrieval");
    if (precision < BEAR_TABLE_FP64 || precision > BEAR_TABLE_INT16)
      bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "unknown table precision " + std::to_string(precision));
    r->impl->setTablePrecision(static_cast<bear_amd::TablePrecision>(precision));
  });
}

int bear_retrieval_set_device_staging(bear_retrieval* r, int enable) {
  return guarded([&] {
    if (!r) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "null retrieval");
    r->impl->setDeviceStaging(enable != 0);
  });
}

int bear_retrieval_set_batch_options(bear_retrieval* r, int use_graph, int autotune) {
  return guarded([&] {
    if (!r) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "null retrieval");
    r->impl->engine().setBatchOptions(use_graph != 0, autotune != 0);
  });
}

int bear_retrieval_set_adjoint_memory(bear_retrieval* r, size_t bytes) {
  return guarded([&] {
    if (!r) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "null retrieval");
    r->impl->engine().setAdjointMemory(bytes);
  });
}

int bear_retrieval_compute_likelihood_batch(bear_retrieval* r, const double* parameters, size_t batch,
                                            size_t nb_parameters, double* log_likelihood) {
  return guarded([&] {
    if (!r || (batch && (!parameters || !log_likelihood))) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "null argument");
    r->impl->computeLikelihoodBatch(parameters, batch, nb_parameters, log_likelihood);
  });
}

int bear_retrieval_compute_model_batch(bear_retrieval* r, const double* parameters, size_t batch,
                                       size_t nb_parameters, double* spectrum, double* spectrum_obs,
                                       int* neglect_model) {
  return guarded([&] {
    if (!r || (batch && (!parameters || !spectrum_obs))) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "null argument");
    r->impl->computeModelBatch(parameters, batch, nb_parameters, spectrum, spectrum_obs, neglect_model);
  });
}

int bear_retrieval_gradient_batch(bear_retrieval* r, const double* parameters, size_t batch, size_t nb_parameters,
                                  double* log_likelihood, double* gradient) {
  return guarded([&] {
    if (!r || (batch && (!parameters || !log_likelihood || !gradient)))
      bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "null argument");
    r->impl->gradientBatch(parameters, batch, nb_parameters, log_likelihood, gradient);
  });
}

int bear_retrieval_batch_stats(const bear_retrieval* r, bear_batch_stats* out) {
  return guarded([&] {
    if (!r || !out) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "null argument");
    const bear_amd::DeviceEngine& e = r->impl->engine();
    const bear_amd::BatchLaunch lc = e.batchLaunch();
    out->precision = int(e.tablePrecision());
    out->block = lc.block;
    out->tile = lc.tile;
    out->grid = lc.grid;
    out->used_graph = e.batchUsedGraph();
    out->evaluations = e.batchEvaluations();
    out->gpu_ms_total = e.batchMilliseconds();
    out->adjoint_chunk = e.adjointChunk();
    out->point_group = lc.group;
  });
}

int bear_spectrum_model_create(const bear_global_config* config, const char* forward_model_type,
                               double wavelength_min, double wavelength_max, size_t nb_grid_points,
                               const char* opacity_species, const double* stellar_wavelengths,
                               const double* stellar_flux, size_t nb_stellar_points, bear_spectrum_model** out) {
  return guarded([&] {
    if (!config || !out) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "null argument");
    *out = nullptr;
    if (!config->use_gpu) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "this engine runs on the GPU only (use_gpu = 1)");
    if (nb_grid_points < 2) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "the atmosphere needs at least two levels");
    const std::string type = str(forward_model_type);
    if ((type == "secondary_eclipse" || type == "se") && (!stellar_wavelengths || !stellar_flux))
      bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "the secondary-eclipse model needs a stellar spectrum");
    auto* m = new bear_spectrum_model;
    try {
      m->impl.reset(new bear_amd::DirectModel(settings(config), type, wavelength_min, wavelength_max, nb_grid_points,
                                              opacityEntries(opacity_species), vec(stellar_wavelengths, nb_stellar_points),
                                              vec(stellar_flux, nb_stellar_points)));
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

void bear_spectrum_model_destroy(bear_spectrum_model* m) { delete m; }

size_t bear_spectrum_model_nb_spectral_points(const bear_spectrum_model* m) { return m ? m->impl->grid().size() : 0; }

int bear_spectrum_model_spectral_grid(const bear_spectrum_model* m, double* wavenumbers, double* wavelengths) {
  return guarded([&] {
    if (!m) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "null model");
    const auto& g = m->impl->grid();
    if (wavenumbers) std::memcpy(wavenumbers, g.wavenumber.data(), g.size() * sizeof(double));
    if (wavelengths) std::memcpy(wavelengths, g.wavelength.data(), g.size() * sizeof(double));
  });
}

int bear_emission_calc_spectrum(bear_spectrum_model* m, double surface_gravity, double radius, double distance,
                                const double* pressure, const double* temperature, size_t nb_levels,
                                const char* const* species, size_t nb_species, const double* mixing_ratios,
                                const double* cloud_optical_depth, int use_variable_gravity, double* spectrum) {
  return guarded([&] {
    const std::vector<int> ids = structureSpecies(m, pressure, temperature, nb_levels, species, nb_species,
                                                  mixing_ratios, cloud_optical_depth, spectrum);
    if (!m->impl->emission()) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "not an emission model");
    m->impl->emission()->calcSpectrum(surface_gravity, radius, distance, pressure, temperature, ids, mixing_ratios,
                                      use_variable_gravity != 0, spectrum);
  });
}

int bear_transmission_calc_spectrum(bear_spectrum_model* m, double surface_gravity, double planet_radius,
                                    double radius_ratio, const double* pressure, const double* temperature,
                                    size_t nb_levels, const char* const* species, size_t nb_species,
                                    const double* mixing_ratios, const double* cloud_optical_depth,
                                    int use_variable_gravity, double* spectrum) {
  return guarded([&] {
    const std::vector<int> ids = structureSpecies(m, pressure, temperature, nb_levels, species, nb_species,
                                                  mixing_ratios, cloud_optical_depth, spectrum);
    if (!m->impl->transmission()) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "not a transmission model");
    m->impl->transmission()->calcSpectrum(surface_gravity, planet_radius, radius_ratio, pressure, temperature, ids,
                                          mixing_ratios, use_variable_gravity != 0, spectrum);
  });
}

int bear_occultation_calc_spectrum(bear_spectrum_model* m, double surface_gravity, double radius_ratio,
                                   const double* pressure, const double* temperature, size_t nb_levels,
                                   const char* const* species, size_t nb_species, const double* mixing_ratios,
                                   const double* cloud_optical_depth, double* spectrum) {
  return guarded([&] {
    const std::vector<int> ids = structureSpecies(m, pressure, temperature, nb_levels, species, nb_species,
                                                  mixing_ratios, cloud_optical_depth, spectrum);
    if (!m->impl->occultation()) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "not a secondary-eclipse model");
    m->impl->occultation()->calcSpectrum(surface_gravity, radius_ratio, pressure, temperature, ids, mixing_ratios,
                                         spectrum);
  });
}

int bear_retrieval_set_profiling(bear_retrieval* r, int enable) {
  return guarded([&] {
    if (!r) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "null retrieval");
    r->impl->engine().setProfiling(enable != 0);
  });
}

int bear_retrieval_stats(const bear_retrieval* r, bear_eval_stats* out) {
  return guarded([&] {
    if (!r || !out) bear_amd::raise(BEAR_ERROR_INVALID_INPUT, "null argument");
    const bear_amd::DeviceEngine& e = r->impl->engine();
    out->nb_spectral_points = e.nbPoints();
    out->nb_levels = e.nbLevels();
    out->nb_species = e.nbSpecies();
    out->nb_bins = e.nbBins();
    out->distinct_table_rows = e.distinctRows();
    out->table_value_bytes = sizeof(double);
    out->kernel_launches = e.timedLaunches();
    out->kernel_ms_total = e.timedMilliseconds();
    out->table_device_bytes = e.tableBytes();
  });
}

}  // extern "C"
