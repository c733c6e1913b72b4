// Forward models and retrieval of the host side.
#include "model.h"

#include <algorithm>
#include <cmath>

namespace bear_amd {

EmissionModel::EmissionModel(const GlobalSettings& global, const ModelSettings& m, const SpectralGrid& grid,
                             const std::vector<ObservationSetup>& observations, DeviceEngine& engine)
    : observations_(observations),
      engine_(engine),
      atmosphere_(m.nb_levels, m.bottom_pressure, m.top_pressure, m.temperature, m.chemistry) {
  if (!m.radiative_transfer.empty() && m.radiative_transfer != "scm")
    raise(kInvalidInput, "radiative transfer " + m.radiative_transfer + " is not available (only scm)");
  // opacity species (TransportCoefficients ctor, transport_coeff.cpp:40-93)
  TableStore store;
  store.nb_points = grid.size();
  for (auto& e : m.opacity) {
    OpacitySpecies s = makeOpacitySpecies(e.first, e.second, grid);
    // emission's short characteristics use absorption only (emission.cpp:210-218),
    // so scattering species do not enter this model's evaluation
    if (s.rayleigh) continue;
    loadSpeciesTable(s, global.cross_section_path, grid, store);
    species_.push_back(std::move(s));
  }
  engine_.uploadTables(store);
  engine_.configure(m.nb_levels, species_.size());
  nb_chem_ = atmosphere_.nbChemistryParameters();
  nb_temp_ = atmosphere_.nbTemperatureParameters();
  for (auto& o : observations_) nb_modifier_ += size_t(o.modifier_parameters);
}

// EmissionModel::calcModelGPU (emission.cpp:234-263): parameters
// [log g, scaling f, distance d, chemistry..., temperature..., modifiers...]
bool EmissionModel::calcModelGPU(const double* p, double inflation, double* spectrum_dev,
                                 double* const* spectrum_obs_dev) {
  const double* chem = p + nb_general_;
  const double* temp = chem + nb_chem_;
  const double* mods = temp + nb_temp_;
  const bool neglect = atmosphere_.compute(std::pow(10, p[0]), temp, chem);

  const size_t L = atmosphere_.nb_levels, S = species_.size(), stride = engine_.speciesStride();
  LevelState* lev = engine_.levels();
  LevelSpecies* spc = engine_.levelSpecies();
  for (size_t l = 0; l < L; ++l) {
    const double t = atmosphere_.temperature[l];
    lev[l] = LevelState{t, 1.0 / t, atmosphere_.altitude[l], 0.0};
    for (size_t s = 0; s < S; ++s) {
      const OpacitySpecies& sp = species_[s];
      double n = atmosphere_.nd(l, sp.species_id);                 // opacity_species.cpp:407-412
      for (int q : sp.cia_partners) n *= atmosphere_.nd(l, q);
      LevelSpecies& d = spc[l * stride + s];
      d.n = n;                                   // rows stay valid when n == 0: the kernel loads them
      const Stencil st = sp.bracket(atmosphere_.pressure[l], t);
      d.f_p_lower = st.f_p_lower;
      d.f_p_upper = st.f_p_upper;
      d.f_t = st.f_t;
      // reload the rows only where they differ from the previous level's
      uint32_t changed = l == 0;
      for (int k = 0; k < 4; ++k) {
        if (l > 0) changed |= spc[(l - 1) * stride + s].row[k] != st.row[k];
        d.row[k] = st.row[k];
      }
      d.reload = changed;
    }
  }
  size_t k = 0;                                                   // forward_model.cpp:184-204
  for (size_t o = 0; o < observations_.size(); ++o) {
    engine_.shifts()[o] = observations_[o].modifier_parameters ? mods[k] : 0.0;
    k += size_t(observations_[o].modifier_parameters);
  }
  double scaling = phys::radius_jupiter / p[2];                   // emission.cpp:108-120
  scaling = scaling * scaling * p[1];
  engine_.evaluateEmission(scaling, inflation, spectrum_dev, spectrum_obs_dev);
  return neglect;
}

// Retrieval set-up: observations, grid sampling, per-observation operators,
// forward model (retrieval.cpp:70-104, retrieval_load_observations.cpp:38-92)
Retrieval::Retrieval(const GlobalSettings& global, const std::string& model_type, const ModelSettings& model,
                     const std::vector<ObservationInputView>& inputs, const std::vector<std::string>& files,
                     const std::vector<std::string>& modifiers)
    : global_(global) {
  if (model_type != "emission")
    raise(kInvalidInput, "forward model " + model_type + " is not available");
  std::string folder = global_.retrieval_folder;
  if (!folder.empty() && folder.back() != '/') folder += '/';
  global_.retrieval_folder = folder;
  grid_.discretisation = global_.discretisation;
  grid_.resolution = global_.resolution;
  if (global_.discretisation < 0 || global_.discretisation > 2) raise(kInvalidInput, "unknown spectral discretisation");
  grid_.loadFullList(global_.wavenumber_file);
  for (auto& in : inputs) observations_.push_back(observationFromInput(in));
  for (size_t i = 0; i < files.size(); ++i)
    observations_.push_back(observationFromFile(folder + files[i], i < modifiers.size() ? modifiers[i] : "", folder));
  grid_.sample(observations_);
  const bool is_flux = true;                                      // emission bins are fluxes
  for (auto& o : observations_) {
    o.setupOnGrid(grid_);
    o.foldOperator(grid_, is_flux);
  }
  engine_.reset(new DeviceEngine(global_.device));
  engine_->uploadGrid(grid_.wavenumber, grid_.wavelength);
  engine_->uploadObservations(observations_);
  model_.reset(new EmissionModel(global_, model, grid_, observations_, *engine_));
}

void Retrieval::checkSize(size_t n) const {
  if (n < nbParameters())
    raise(kInvalidInput, "expected " + std::to_string(nbParameters()) + " parameters, got " + std::to_string(n));
}

double Retrieval::errorInflation(const double* p, size_t n) const {
  return global_.error_inflation ? std::pow(10, p[n - 1]) : 0.0;   // retrieval.cpp:333-336
}

// Retrieval::logLikelihoodGPU (retrieval.cpp:374-420)
double Retrieval::computeLikelihood(const double* p, size_t n) {
  checkSize(n);
  const bool neglect = model_->calcModelGPU(p, errorInflation(p, n), nullptr, nullptr);
  const double ll = engine_->logLikelihood();
  return neglect ? -1e30 : ll;
}

bool Retrieval::calcModelGPU(const double* p, size_t n, double* spectrum_dev, double* const* obs_dev) {
  checkSize(n);
  const bool neglect = model_->calcModelGPU(p, errorInflation(p, n), spectrum_dev, obs_dev);
  engine_->synchronize();
  return neglect;
}

bool Retrieval::computeModel(const double* p, size_t n, bool high_res, double* spectrum, double* spectrum_obs) {
  checkSize(n);
  double* spec_dev = high_res ? engine_->spectrumBuffer() : nullptr;
  const bool neglect = model_->calcModelGPU(p, errorInflation(p, n), spec_dev, nullptr);
  std::vector<double> bins;
  engine_->copyBins(bins);
  if (high_res && spectrum) engine_->copySpectrum(spectrum, spec_dev);
  // per observation, back in input order (forward_model.cpp:119-123)
  size_t k = 0;
  for (auto& o : observations_) {
    const size_t nb = o.nbPoints();
    if (o.ascending) std::reverse_copy(bins.begin() + k, bins.begin() + k + nb, spectrum_obs + k);
    else std::copy(bins.begin() + k, bins.begin() + k + nb, spectrum_obs + k);
    k += nb;
  }
  return neglect;
}

}  // namespace bear_amd
