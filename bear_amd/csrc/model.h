// Forward models and the retrieval's likelihood path, mirroring the
// reference's ForwardModel interface (forward_model/forward_model.h:61-144)
// on top of the device engine.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "engine.h"
#include "host.h"

namespace bear_amd {

struct GlobalSettings {                 // GlobalConfig (config/global_config.h:28-71)
  int device = 0;
  std::string retrieval_folder, cross_section_path, wavenumber_file;
  int discretisation = 0;
  double resolution = 0;
  bool error_inflation = false;
  TablePrecision table_storage = TablePrecision::fp64;   // see DeviceEngine::setTableStorage
};

// EmissionModelConfig (emission.h:51-102), TransmissionModelConfig
// (transmission.h:55-130), OccultationModelConfig (secondary_eclipse.h:55-120)
struct ModelSettings {
  size_t nb_levels = 0;
  double bottom_pressure = 0, top_pressure = 0;
  std::string temperature, radiative_transfer, chemistry;
  std::vector<std::pair<std::string, std::string>> opacity;   // (symbol, folder)
  bool fit_mean_molecular_weight = false, fit_scale_height = false, use_variable_gravity = false;
  std::string stellar_model = "blackbody";
  std::vector<CloudModelSpec> clouds;
};

class ForwardModel {
 public:
  virtual ~ForwardModel() = default;
  // parameters of the model itself (ForwardModel::parametersNumber)
  virtual size_t parametersNumber() const = 0;
  // ForwardModel::calcModelGPU (forward_model.h:76-79) with the retrieval's
  // error inflation forwarded to the likelihood terms; returns "neglect"
  virtual bool calcModelGPU(const double* parameters, double error_inflation, double* spectrum_dev,
                            double* const* spectrum_obs_dev) = 0;
  // the binned spectrum is a flux (emission, eclipse) or not (transmission)
  virtual bool binsAreFluxes() const = 0;
  virtual ModelKind kind() const = 0;
  // live points [0, batch) (parameters [batch][stride]): scalars, atmosphere
  // and stencils into the engine's batch records, in parallel over points;
  // neglect[b] set per point
  virtual void stageBatch(const double* parameters, size_t batch, size_t stride, const double* error_inflation,
                          char* neglect) = 0;
  // [first, end) of each tabulated species' rows in the table store
  virtual std::vector<std::pair<uint32_t, uint32_t>> speciesRows() const = 0;
  // device staging (StagePlan) of this model's batches with parameter rows
  // of nb_params: uploaded once; false where the model has what the device
  // staging does not form (the H- continuum, clouds)
  virtual bool prepareStaging(size_t nb_params, bool error_inflation) {
    (void)nb_params, (void)error_inflation;
    return false;
  }
  // gradient rows [batch][stride] of the model's parameters (and observation
  // modifiers) from the engine's adjoint: level gradients [batch][levels][slots],
  // bin gradients [batch][bins] and the model's terms (eclipse: per bin
  // [batch][bins][2]; transmission: per chord [batch][segments][tangents]);
  // rows of neglected points are left alone
  virtual void gradientBatch(const double* parameters, size_t batch, size_t stride,
                             const std::vector<double>& level_grad, const std::vector<double>& bin_grad,
                             const std::vector<double>& bin_terms, const char* neglect, double* gradient) {
    (void)parameters, (void)batch, (void)stride, (void)level_grad, (void)bin_grad, (void)bin_terms, (void)neglect,
        (void)gradient;
    raise(kInvalidInput, "this model has no gradient");
  }
};

// What the three models share: atmosphere, opacity species with their
// tables on the device, and the per-level stencils of one evaluation.
class TabulatedModel : public ForwardModel {
 public:
  void stageBatch(const double* parameters, size_t batch, size_t stride, const double* error_inflation,
                  char* neglect) override;
  std::vector<std::pair<uint32_t, uint32_t>> speciesRows() const override;
  bool prepareStaging(size_t nb_params, bool error_inflation) override;
  // pyBeAR's direct spectra: the atmosphere set from the caller's structure
  // (Atmosphere::setStructure: pressure, temperature, mixing ratios [levels]
  // [species] of the given species ids) and staged for one evaluation
  void stageStructure(double surface_gravity, double bottom_radius, bool variable_gravity, const double* pressure,
                      const double* temperature, const std::vector<int>& species, const double* mixing_ratios);

 protected:
  TabulatedModel(const GlobalSettings& global, const ModelSettings& model, const SpectralGrid& grid,
                 const std::vector<ObservationSetup>& observations, DeviceEngine& engine, bool transmission);
  // the atmosphere of one parameter vector and the model's scalars; returns "neglect"
  virtual bool setup(const double* parameters, Atmosphere& atmosphere, BatchModel& scalars) const = 0;
  // levels, stencils (FP64 and/or FP32 form) and observation shifts of one
  // computed atmosphere
  void stageLevels(const Atmosphere& atmosphere, const double* modifiers, LevelState* levels, LevelSpecies* species,
                   LevelSpeciesAdjoint* species_adjoint, LevelSpecies32* species32, double* shifts) const;
  // the single-model path: setup + stageLevels into the engine's staging
  bool stage(const double* parameters, BatchModel& scalars);
  const double* clouds(const double* p) const { return p + nb_general_ + nb_chem_ + nb_temp_; }
  const double* modifiers(const double* p) const { return clouds(p) + nb_cloud_; }
  // the cloud layers of one computed atmosphere (cloud parameters cp,
  // surface gravity g): per level LevelState::cloud_w, and the models'
  // spectral factors
  void stageClouds(const Atmosphere& atmosphere, const double* cp, double g, double scale_height_km,
                   LevelState* levels, CloudSpec& spec) const;
  // the fixed scale height (km) of the transmission variant that fits one, else 0
  virtual double fixedScaleHeight(const double* p) const {
    (void)p;
    return 0.0;
  }
  // the altitude variant (StagePlan::alt_kind) and radius-dependent gravity
  virtual int altitudeKind() const { return 0; }
  virtual bool variableGravity() const { return false; }
  size_t staged_params_ = 0;                // parameter row length of the uploaded StagePlan, 0: none
  // emission and eclipse: the chain from the level gradients through the
  // atmosphere to [log g, chemistry..., temperature..., modifiers...];
  // generalGradient fills the other two general parameters
  void chainGradient(const double* parameters, size_t batch, size_t stride, const std::vector<double>& level_grad,
                     const std::vector<double>& bin_grad, const std::vector<double>& bin_terms, const char* neglect,
                     double* gradient);
  // scale_bar: dL/d(flux scaling); bin_terms: this point's [bins][2] (eclipse)
  virtual void generalGradient(const double* p, double scale_bar, const double* bin_terms, double* out) const {
    (void)p, (void)scale_bar, (void)bin_terms, (void)out;
  }
  // dL/dn[level][species id] from the level gradients G [levels][slots]
  // (n_eff dL/dn_eff per stencil slot) through the CIA products
  // per level and species id d logL / d n from the kernels' level slots (the
  // H- continuum's weights also give temperature terms, added to t_bar)
  void densityGradient(const Atmosphere& atm, const double* G, std::vector<double>& n_bar,
                       std::vector<double>& t_bar) const;
  // the adjoint's level slots: NS + 4, 8 more with the H- continuum, 16 more
  // with clouds (their layer sums from cloudSlot() on)
  size_t cloudSlot() const { return engine_.speciesStride() + (hminus_ ? 12 : 4); }
  size_t levelSlots() const { return cloudSlot() + (clouds_.empty() ? 0 : 4 * kMaxClouds); }
  // d logL / d the cloud models' optical depths and spectral parameters
  void cloudGradient(const Atmosphere& atm, const double* p, double g, const double* G, double* out) const;
  // the observations' shift parameters: per observation the sum of its bins' dL/dbin
  void shiftGradient(const double* bin_grad, double* out) const;
  const std::vector<ObservationSetup>& observations_;
  DeviceEngine& engine_;
  Atmosphere atmosphere_;
  std::vector<std::unique_ptr<Atmosphere>> atmosphere_pool_;   // one per staging thread
  std::vector<OpacitySpecies> species_;     // tabulated absorbers
  int rayleigh_species_ = -1;               // species id of the Rayleigh scatterer, -1: none
  bool hminus_ = false;                     // the H- continuum is listed
  int hminus_ids_[3] = {-1, -1, -1};        // species ids of H-, H and e-
  int group_rep_[2] = {-1, -1};             // a species of each stencil group (StencilGroups)
  std::vector<CloudModelSpec> clouds_;
  bool transmission_ = false;
  size_t nb_general_ = 3, nb_chem_ = 0, nb_temp_ = 0, nb_cloud_ = 0, nb_modifier_ = 0;
};

// EmissionModel (forward_model/emission/emission.h:127-250): parameters
// [log g, f, d, chemistry..., temperature..., modifiers...]
class EmissionModel : public TabulatedModel {
 public:
  EmissionModel(const GlobalSettings& global, const ModelSettings& model, const SpectralGrid& grid,
                const std::vector<ObservationSetup>& observations, DeviceEngine& engine);
  size_t parametersNumber() const override { return 3 + nb_chem_ + nb_temp_ + nb_cloud_ + nb_modifier_; }
  bool calcModelGPU(const double* parameters, double error_inflation, double* spectrum_dev,
                    double* const* spectrum_obs_dev) override;
  bool binsAreFluxes() const override { return true; }
  ModelKind kind() const override { return ModelKind::emission; }
  // EmissionModel::calcSpectrum (emission.cpp:268-340): the flux [N] in
  // W m^-2 cm times radius^2 / distance^2
  void calcSpectrum(double surface_gravity, double radius, double distance, const double* pressure,
                    const double* temperature, const std::vector<int>& species, const double* mixing_ratios,
                    bool variable_gravity, double* spectrum);
  void gradientBatch(const double* parameters, size_t batch, size_t stride, const std::vector<double>& level_grad,
                     const std::vector<double>& bin_grad, const std::vector<double>& bin_terms, const char* neglect,
                     double* gradient) override {
    chainGradient(parameters, batch, stride, level_grad, bin_grad, bin_terms, neglect, gradient);
  }

 protected:
  bool setup(const double* parameters, Atmosphere& atmosphere, BatchModel& scalars) const override;
  void generalGradient(const double* p, double scale_bar, const double* bin_terms, double* out) const override;
};

// OccultationModel (forward_model/secondary_eclipse/secondary_eclipse.h) with
// the blackbody star: parameters [log g, Rp/Rs, T_star, chemistry...,
// temperature..., modifiers...]
class OccultationModel : public TabulatedModel {
 public:
  OccultationModel(const GlobalSettings& global, const ModelSettings& model, const SpectralGrid& grid,
                   const std::vector<ObservationSetup>& observations, DeviceEngine& engine);
  size_t parametersNumber() const override { return 3 + nb_chem_ + nb_temp_ + nb_cloud_ + nb_modifier_; }
  bool calcModelGPU(const double* parameters, double error_inflation, double* spectrum_dev,
                    double* const* spectrum_obs_dev) override;
  bool binsAreFluxes() const override { return true; }
  ModelKind kind() const override { return ModelKind::eclipse; }
  // StarSpectrumFile(wavelengths, flux, grid) (star_file_spectrum.cpp:57-72):
  // a stellar flux in W m^-2 micron^-1 at the given wavelengths (micron),
  // converted to W m^-2 cm and interpolated onto the grid
  void setStellarSpectrum(const std::vector<double>& wavelength, const std::vector<double>& flux,
                          const SpectralGrid& grid);
  // OccultationModel::calcSpectrum (secondary_eclipse.cpp:395-482): the
  // eclipse depth [N] in ppm against that stellar flux
  void calcSpectrum(double surface_gravity, double radius_ratio, const double* pressure, const double* temperature,
                    const std::vector<int>& species, const double* mixing_ratios, double* spectrum);
  void gradientBatch(const double* parameters, size_t batch, size_t stride, const std::vector<double>& level_grad,
                     const std::vector<double>& bin_grad, const std::vector<double>& bin_terms, const char* neglect,
                     double* gradient) override {
    chainGradient(parameters, batch, stride, level_grad, bin_grad, bin_terms, neglect, gradient);
  }

 protected:
  bool setup(const double* parameters, Atmosphere& atmosphere, BatchModel& scalars) const override;
  void generalGradient(const double* p, double scale_bar, const double* bin_terms, double* out) const override;
};

// TransmissionModel (forward_model/transmission/transmission.h): parameters
// [log g, R_p (cm), R_* (cm), (mu | H in km), chemistry..., temperature...,
// modifiers...]
class TransmissionModel : public TabulatedModel {
 public:
  TransmissionModel(const GlobalSettings& global, const ModelSettings& model, const SpectralGrid& grid,
                    const std::vector<ObservationSetup>& observations, DeviceEngine& engine);
  size_t parametersNumber() const override { return nb_general_ + nb_chem_ + nb_temp_ + nb_cloud_ + nb_modifier_; }
  bool calcModelGPU(const double* parameters, double error_inflation, double* spectrum_dev,
                    double* const* spectrum_obs_dev) override;
  bool binsAreFluxes() const override { return false; }
  ModelKind kind() const override { return ModelKind::transmission; }
  // TransmissionModel::calcSpectrum (transmission.cpp:371-465): the transit
  // depth [N] in ppm, R_* = planet_radius / radius_ratio
  void calcSpectrum(double surface_gravity, double planet_radius, double radius_ratio, const double* pressure,
                    const double* temperature, const std::vector<int>& species, const double* mixing_ratios,
                    bool variable_gravity, double* spectrum);
  void gradientBatch(const double* parameters, size_t batch, size_t stride, const std::vector<double>& level_grad,
                     const std::vector<double>& bin_grad, const std::vector<double>& chord_grad, const char* neglect,
                     double* gradient) override;

 protected:
  bool setup(const double* parameters, Atmosphere& atmosphere, BatchModel& scalars) const override;

 private:
  AltitudeOptions altitudeOptions(const double* p) const;
  double fixedScaleHeight(const double* p) const override { return fit_scale_height_ ? p[3] : 0.0; }
  int altitudeKind() const override { return fit_mmw_ ? 1 : fit_scale_height_ ? 2 : 0; }
  bool variableGravity() const override { return variable_gravity_; }
  // the chord lengths of the staged atmosphere into the engine's staging
  void stageChords(double bottom_radius);
  bool fit_mmw_ = false, fit_scale_height_ = false, variable_gravity_ = false;
};

// Retrieval (retrieval/retrieval.h) restricted to the likelihood path
class Retrieval {
 public:
  Retrieval(const GlobalSettings& global, const std::string& model_type, const ModelSettings& model,
            const std::vector<ObservationInputView>& inputs, const std::vector<std::string>& files,
            const std::vector<std::string>& modifiers);
  size_t nbParameters() const { return model_->parametersNumber() + (global_.error_inflation ? 1 : 0); }
  const SpectralGrid& grid() const { return grid_; }
  const std::vector<ObservationSetup>& observations() const { return observations_; }
  double computeLikelihood(const double* parameters, size_t n);
  // ForwardModel::calcModel (forward_model.cpp:60-127) with host outputs
  bool computeModel(const double* parameters, size_t n, bool high_res, double* spectrum, double* spectrum_obs);
  bool calcModelGPU(const double* parameters, size_t n, double* spectrum_dev, double* const* spectrum_obs_dev);
  // the same with the per-evaluation device buffers in the static scratch
  // array (DeviceEngine::useScratch): no device memory is allocated
  bool calcModelGPUNoAlloc(const double* parameters, size_t n, double* spectrum_dev, double* const* spectrum_obs_dev);
  // batches: parameters [batch][n] row-major
  void computeLikelihoodBatch(const double* parameters, size_t batch, size_t n, double* log_likelihood);
  void computeModelBatch(const double* parameters, size_t batch, size_t n, double* spectrum, double* spectrum_obs,
                         int* neglect);
  void setTablePrecision(TablePrecision p);
  // batches staged on the device where the model allows it (the default),
  // or on the host
  void setDeviceStaging(bool on) { device_staging_ = on; }
  // log-likelihoods [batch] and their gradients [batch][n] (FP64 tables)
  void gradientBatch(const double* parameters, size_t batch, size_t n, double* log_likelihood, double* gradient);
  DeviceEngine& engine() const { return *engine_; }

 private:
  double errorInflation(const double* parameters, size_t n) const;
  void checkSize(size_t n) const;
  std::vector<char> stageBatch(const double* parameters, size_t batch, size_t n);
  // evaluate a batch, staged on the device when possible; neglect flags out
  void evaluateStaged(const double* parameters, size_t batch, size_t n, std::vector<char>& neglect);
  bool device_staging_ = true;
  GlobalSettings global_;
  SpectralGrid grid_;
  std::vector<ObservationSetup> observations_;
  std::unique_ptr<DeviceEngine> engine_;
  std::unique_ptr<ForwardModel> model_;
};

// pyBeAR's direct forward models (python/src/bear_python_wrapper.cpp:343-372):
// a model on a grid sampled between two wavelengths (micron), without
// observations, whose calcSpectrum takes the atmosphere's structure from
// the caller.  The pressure boundaries of the set-up are the reference's
// placeholders (10 and 1e-5 bar, emission.cpp:86-90); every call brings its
// own pressures.
class DirectModel {
 public:
  DirectModel(const GlobalSettings& global, const std::string& model_type, double wavelength_min,
              double wavelength_max, size_t nb_levels, const std::vector<std::pair<std::string, std::string>>& opacity,
              const std::vector<double>& stellar_wavelength, const std::vector<double>& stellar_flux);
  const SpectralGrid& grid() const { return grid_; }
  size_t nbLevels() const { return nb_levels_; }
  const std::string& type() const { return type_; }
  EmissionModel* emission() const { return emission_; }
  TransmissionModel* transmission() const { return transmission_; }
  OccultationModel* occultation() const { return occultation_; }
  DeviceEngine& engine() const { return *engine_; }

 private:
  GlobalSettings global_;
  std::string type_;
  size_t nb_levels_ = 0;
  SpectralGrid grid_;
  std::vector<ObservationSetup> observations_;   // none
  std::unique_ptr<DeviceEngine> engine_;
  std::unique_ptr<ForwardModel> model_;
  EmissionModel* emission_ = nullptr;
  TransmissionModel* transmission_ = nullptr;
  OccultationModel* occultation_ = nullptr;
};

void registerTable(const std::string& folder, double mass, bool log_storage, size_t nb_rows,
                   const double* temperatures, const double* pressures, size_t nb_wavenumbers, const float* rows);
void clearRegisteredTables();

}  // namespace bear_amd
