/*
 * bear_hip.h -- C ABI of the MI355X forward-model and likelihood engine.
 *
 * A drop-in for the part of BeAR that one likelihood evaluation runs through:
 * the forward models' calcModel / calcModelGPU, the retrieval's
 * computeLikelihood / computeModel, and the pyBeAR bindings of those calls.
 * Everything crosses this boundary as plain C: scalars, NUL-terminated
 * strings, and pointers with explicit sizes.  No C++ types, no exceptions;
 * every call returns a status code and bear_last_error() describes the last
 * failure of the calling thread.
 *
 * Reference citations are relative to the reference repository's src/.
 *
 * Threading: a bear_retrieval binds to one device and one HIP stream.  Calls
 * on one bear_retrieval are not re-entrant; different objects may be used
 * from different threads.
 */
#ifndef BEAR_HIP_H
#define BEAR_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status
 * The reference reports errors by exceptions (additional/exceptions.h:32-49:
 * InvalidInput, FileNotFound), std::bad_alloc (data_management_kernels.cu:85-91)
 * and exit() on device errors (error_check.h:29-37).  Here each becomes a code. */
#define BEAR_OK                    0
#define BEAR_ERROR_INVALID_INPUT  -1   /* InvalidInput */
#define BEAR_ERROR_FILE_NOT_FOUND -2   /* FileNotFound */
#define BEAR_ERROR_DEVICE         -3   /* a HIP runtime error (the reference exits) */
#define BEAR_ERROR_OUT_OF_MEMORY  -4   /* std::bad_alloc / hipErrorOutOfMemory */
#define BEAR_ERROR_INTERNAL       -5

/* Text of the last error raised on the calling thread ("" if none). */
const char* bear_last_error(void);

/* Library version string. */
const char* bear_version(void);

/* ---------------------------------------------------------------- configuration
 * GlobalConfig (config/global_config.h:28-71), the fields the forward-model
 * path reads.  Paths follow the reference: the opacity folders named by the
 * model configuration are relative to cross_section_file_path; observation
 * files and bandpass files are relative to retrieval_folder_path.  A missing
 * wavenumber file selects the default 0..90000 cm^-1 grid at 0.01 cm^-1
 * (spectral_grid.cpp:70-84). */
typedef struct bear_global_config {
  int use_gpu;                      /* must be 1: this engine has no CPU path */
  int device;                       /* HIP device ordinal */
  const char* retrieval_folder_path;
  const char* cross_section_file_path;
  const char* wavenumber_file_path;
  int spectral_discretisation;      /* 0 const_wavenumber, 1 const_wavelength, 2 const_resolution */
  double spectral_resolution;
  int use_error_inflation;          /* last parameter is log10 of the added variance */
  /* Storage of the cross-section tables on the device (BEAR_TABLE_*, defined
   * below; no counterpart in the reference, whose tables are FP64 in
   * memory).  BEAR_TABLE_FP64 (0) keeps the FP64 rows, which every call can
   * use and from which bear_retrieval_set_table_precision splits
   * reduced-precision copies.  FP32/FP16/BF16/INT16 split each species'
   * rows into the per-wavenumber FP32 base (and scale) and the 16- or
   * 32-bit values while the tables load (one species' FP64 rows at a time
   * pass through host and device memory) and keep nothing else: the
   * batched forward calls run at that precision,
   * bear_retrieval_compute_likelihood and bear_retrieval_compute_model run
   * as batches of one, and the calls that need the FP64 rows
   * (bear_forward_model_calc_model_gpu and its no-allocation variant, the
   * gradient, another table precision) return BEAR_ERROR_INVALID_INPUT.  The
   * direct spectrum models accept FP64 only. */
  int table_storage;
} bear_global_config;

/* Forward-model configuration (EmissionModelConfig, forward_model/emission/emission.h:51-102;
 * TransmissionModelConfig and OccultationModelConfig alike).  The string
 * fields use the line syntax of the reference's forward_model.config:
 *   temperature_profile  "poly 6 1" | "const"
 *   radiative_transfer   "scm"
 *   chemistry            one module per line: "iso H2O CH4\nbg H2He"
 *   opacity_species      one "symbol folder" pair per line: "H2O Molecules/H2O\nCIA-H2-H2 CIA/H2-H2\nH2 Rayleigh" */
typedef struct bear_model_config {
  size_t nb_grid_points;
  double atmos_bottom_pressure;     /* bar */
  double atmos_top_pressure;        /* bar */
  const char* temperature_profile;
  const char* radiative_transfer;
  const char* chemistry;
  const char* opacity_species;
  /* transmission (TransmissionModelConfig, transmission.h:55-70): a 4th
   * general parameter fixes the mean molecular weight or the scale height
   * (km); gravity may fall off as 1/r^2 with altitude */
  int fit_mean_molecular_weight;
  int fit_scale_height;
  int use_variable_gravity;
  /* secondary eclipse (OccultationModelConfig): "blackbody" (NULL: blackbody) */
  const char* stellar_spectrum_model;
  /* cloud models (EmissionModelConfig::cloud_model and cloud_model_parameters,
   * emission.h:66-67; select_cloud_model.h:44-99), one per line, NULL or
   * empty for none: "grey", "KHnongrey <lambda_ref>", "power_law <lambda_ref>"
   * (micron), each optionally followed by "fb" (bottom one scale height
   * below the top).  Their parameters follow the temperature's: grey [tau,
   * p_top, p_bottom / p_top], KHnongrey [tau, q0, a0, particle size (micron),
   * p_top, p_bottom / p_top], power_law [tau, exponent, p_top,
   * p_bottom / p_top], without the last with "fb".  At most 4 models; the
   * single-point calls and the batched calls with FP64 tables compute them,
   * the gradient calls take them in emission and eclipse retrievals;
   * transmission gradients and reduced-precision batches with clouds return
   * BEAR_ERROR_INVALID_INPUT. */
  const char* cloud_model;
} bear_model_config;

/* ObservationInput (observations/observations.h:43-83).  For "spectroscopy"
 * give wavelengths[nb_points]; for "band-spectroscopy" and "photometry" give
 * bin_edges[nb_points][2] (micron).  Optional arrays may be NULL. */
typedef struct bear_observation_input {
  const char* name;
  const char* type;                          /* spectroscopy | band-spectroscopy | photometry */
  size_t nb_points;
  const double* wavelengths;
  const double* bin_edges;
  const double* data;
  const double* data_error;
  const double* likelihood_weight;           /* NULL: all 1 */
  const double* instrument_profile_fwhm;     /* NULL: no line-spread function */
  size_t filter_nb_points;                   /* 0: no bandpass */
  const double* filter_wavelength;
  const double* filter_response;
  const char* filter_detector_type;          /* "energy" | "photon" */
  const char* spectrum_modifier;             /* NULL, "none" or "shift_const" */
} bear_observation_input;

/* Tables held in memory instead of on disk.  A registered folder is used in
 * place of <cross_section_file_path><folder>/filelist.dat and its row files
 * (opacity_species.cpp:66-122, cross_section_file.cpp:33-62): mass and
 * log_storage mean what the first filelist.dat line means, and each row holds
 * nb_wavenumbers float32 values over the full wavenumber grid exactly as a row
 * file would.  The data are copied; registration is process-wide. */
int bear_opacity_data_register(const char* folder, double species_mass, int log_storage,
                               size_t nb_rows, const double* temperatures, const double* pressures,
                               size_t nb_wavenumbers, const float* rows);
int bear_opacity_data_clear(void);

/* ---------------------------------------------------------------- retrieval
 * Retrieval (retrieval/retrieval.h) restricted to the likelihood path:
 * observations set up (retrieval_load_observations.cpp:38-92), the
 * spectral grid sampled from them (spectral_grid.cpp:123-304) and a forward
 * model selected (retrieval/select_forward_model.cpp:38-164): "emission",
 * "transmission" or "secondary_eclipse".
 * Observations are the in-memory inputs followed by the files named in
 * obs_files (entries of observations.list, with their modifiers). */
typedef struct bear_retrieval bear_retrieval;

int bear_retrieval_create(const bear_global_config* config,
                          const char* forward_model_type,
                          const bear_model_config* model_config,
                          const bear_observation_input* observations, size_t nb_observations,
                          const char* const* obs_files, const char* const* obs_modifiers, size_t nb_obs_files,
                          bear_retrieval** out);
void bear_retrieval_destroy(bear_retrieval* r);

/* Sizes: physical parameters (ForwardModel::parametersNumber + 1 with error
 * inflation), high-resolution points, observations and their points. */
size_t bear_retrieval_nb_parameters(const bear_retrieval* r);
size_t bear_retrieval_nb_spectral_points(const bear_retrieval* r);
size_t bear_retrieval_nb_observations(const bear_retrieval* r);
size_t bear_retrieval_nb_observation_points(const bear_retrieval* r, size_t observation);

/* Wavenumbers (cm^-1, ascending) and wavelengths (micron) of the sampled grid
 * (SpectralGrid::wavenumber_list / wavelength_list), each [nb_spectral_points]. */
int bear_retrieval_spectral_grid(const bear_retrieval* r, double* wavenumbers, double* wavelengths);

/* Retrieval::computeLikelihood (retrieval.cpp:295-309, 374-420): one
 * evaluation of the forward model on the device and the Gaussian
 * log-likelihood; -1e30 for a neglected model. */
int bear_retrieval_compute_likelihood(bear_retrieval* r, const double* parameters, size_t nb_parameters,
                                      double* log_likelihood);

/* Retrieval::computeModel -> ForwardModel::calcModel (forward_model.cpp:60-127):
 * spectrum [nb_spectral_points] (may be NULL when return_high_res is 0) in the
 * model's high-resolution units; spectrum_obs holds every observation's points,
 * concatenated in observation order, each in the order of its input
 * (ascending inputs are returned ascending, forward_model.cpp:119-123). */
int bear_retrieval_compute_model(bear_retrieval* r, const double* parameters, size_t nb_parameters,
                                 int return_high_res, double* spectrum, double* spectrum_obs,
                                 int* neglect_model);

/* ForwardModel::calcModelGPU (forward_model.h:76-79): device buffers owned by
 * the caller.  spectrum_dev [nb_spectral_points] doubles; spectrum_obs_dev[i]
 * [nb_observation_points(i)] doubles, in the internal descending-wavelength
 * order of each observation.  Returns after the work is complete. */
int bear_forward_model_calc_model_gpu(bear_retrieval* r, const double* parameters, size_t nb_parameters,
                                      double* spectrum_dev, double* const* spectrum_obs_dev,
                                      int* neglect_model);

/* bear_forward_model_calc_model_gpu for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same arguments and
 * results; the per-evaluation device buffers (staged level and stencil
 * records, the flux, bin piece sums, bins, the log-likelihood) live in a
 * static device array of 16 MiB per GPU, held until the process ends, so
 * the call allocates nothing (the caller's output buffers and the
 * retrieval's tables are set up beforehand as usual).  A model whose buffers
 * exceed 16 MiB (more than about 1.9 million spectral points) is refused
 * with BEAR_ERROR_INVALID_INPUT.  Every call on one GPU uses that GPU's
 * array: calls that overlap on one GPU (from several threads or retrievals)
 * race. */
int bear_forward_model_calc_model_gpu_noalloc(bear_retrieval* r, const double* parameters, size_t nb_parameters,
                                              double* spectrum_dev, double* const* spectrum_obs_dev,
                                              int* neglect_model);

/* ---------------------------------------------------------------- batches
 * Many live points per launch sequence.  The reference evaluates one point
 * per call (the sampler callback, retrieval.h:118-129 and
 * multinest_loglike.cpp:131-186, calls Retrieval::computeLikelihood once per
 * point); these calls take `batch` parameter vectors at once, row-major
 * [batch][nb_parameters] (nb_parameters >= bear_retrieval_nb_parameters), and
 * return what that many single calls would.
 *
 * Table precision of the batched kernels (the single-point calls above
 * always use the FP64 tables):
 *   BEAR_TABLE_FP64  FP64 tables and arithmetic: the single-point results
 *   BEAR_TABLE_FP32  per (species, wavenumber) an FP32 base plus per-row
 *   BEAR_TABLE_FP16  residuals log10 sigma - base stored in FP32, FP16 or
 *   BEAR_TABLE_BF16  BF16; interpolation and exponentials in FP32,
 *                    radiative transfer, transit-depth integral, bins and
 *                    log-likelihood in FP64.  The residuals' precision
 *                    follows each species' range of log10 sigma over its
 *                    rows at a wavenumber.
 *   BEAR_TABLE_INT16 16-bit integers q with log10 sigma = base + scale q,
 *                    base and scale FP32 per (species, wavenumber): the
 *                    same 2 bytes per value as FP16 with an absolute
 *                    error of at most range / 131068 at every value
 *                    (FP16 residuals: up to range / 4096); arithmetic as
 *                    above. */
#define BEAR_TABLE_FP64 0
#define BEAR_TABLE_FP32 1
#define BEAR_TABLE_FP16 2
#define BEAR_TABLE_BF16 3
#define BEAR_TABLE_INT16 4

int bear_retrieval_set_table_precision(bear_retrieval* r, int precision);

/* use_graph: replay each (model, precision, batch size)'s launch sequence
 * from a captured hipGraph.  autotune: time the candidate launch shapes on
 * the first batch of each size and keep the fastest (otherwise a fixed
 * shape).  Both default to 1. */
int bear_retrieval_set_batch_options(bear_retrieval* r, int use_graph, int autotune);

/* Where the batched forward calls form each live point's atmosphere
 * (temperature profile, chemistry, altitudes) and table stencils: on the
 * device from the parameters (enable = 1, the default; the parameters are
 * the only per-batch upload), or on the host in parallel over points with
 * the records uploaded (enable = 0).  No counterpart in the reference, which
 * has no batches.  The gradient calls stage the same way (their chain rule
 * then computes each point's atmosphere on the host).  Models with the H-
 * continuum or clouds stage on the host. */
int bear_retrieval_set_device_staging(bear_retrieval* r, int enable);

/* log_likelihood [batch]; -1e30 for neglected points. */
int bear_retrieval_compute_likelihood_batch(bear_retrieval* r, const double* parameters, size_t batch,
                                            size_t nb_parameters, double* log_likelihood);

/* spectrum [batch][nb_spectral_points] (NULL: not returned), spectrum_obs
 * [batch][total observation points] (each point's observations as in
 * bear_retrieval_compute_model), neglect_model [batch] (may be NULL). */
int bear_retrieval_compute_model_batch(bear_retrieval* r, const double* parameters, size_t batch,
                                       size_t nb_parameters, double* spectrum, double* spectrum_obs,
                                       int* neglect_model);

/* Value and gradient of the log-likelihood for a batch (emission,
 * secondary-eclipse and transmission models, FP64 tables).  No counterpart
 * in the reference, which has no derivatives: a hand-written adjoint of the
 * whole chain (parameters -> temperature profile, chemistry, altitudes ->
 * opacity interpolation (transmission: with Rayleigh scattering; the H-
 * continuum when the adjoint memory holds the kept forward quantities) ->
 * short characteristics, or the tangent paths' optical depths, chord lengths
 * and transit-depth integral -> folded observation operator, for the eclipse
 * against the blackbody star -> Gaussian likelihood; the cloud models of
 * emission and eclipse retrievals likewise).  Transmission models with
 * clouds return BEAR_ERROR_INVALID_INPUT.
 * log_likelihood [batch] as
 * bear_retrieval_compute_likelihood_batch; gradient [batch][nb_parameters]
 * (rows of neglected points are zero). */
int bear_retrieval_gradient_batch(bear_retrieval* r, const double* parameters, size_t batch, size_t nb_parameters,
                                  double* log_likelihood, double* gradient);

/* Device memory (bytes) the gradient may hold for the forward quantities of
 * every level that its backward pass reads back (4 bytes x (2 angles'
 * intensity less the Planck function + opacity) x levels x spectral points
 * per live point, plus per-point partial sums; transmission: the same 12
 * bytes per level and spectral point for the extinction and the tangent
 * paths' adjoints, plus 8 bytes x (levels - 1) x (17 (levels - 1) + the
 * levels - 1 rounded up to 32) per point).
 * Batches are processed in chunks of as many points as fit; with 0, or when
 * a single point does not fit, the emission and eclipse backward kernels
 * recompute the forward pass instead, and the transmission gradient takes
 * one point at a time with its spectral grid in windows (at least 1/16 of
 * the grid each, the forward rerun per window; the same result bit for
 * bit), returning BEAR_ERROR_INVALID_INPUT when not even that fits.
 * Default 96 GiB, never more than 80 % of the free memory.
 * No counterpart in the reference. */
int bear_retrieval_set_adjoint_memory(bear_retrieval* r, size_t bytes);

typedef struct bear_batch_stats {
  int precision;                 /* BEAR_TABLE_* */
  int block;                     /* launch shape of the last batch: threads per workgroup */
  int tile;                      /* wavenumbers per work item */
  int grid;                      /* persistent workgroups */
  int used_graph;
  size_t evaluations;            /* live points evaluated in batches */
  double gpu_ms_total;           /* device time of those batches (with profiling on) */
  size_t adjoint_chunk;          /* last gradient: points per chunk with stored intensities, 0: recomputed */
  int point_group;               /* launch shape: points per group of the work-item order */
} bear_batch_stats;

int bear_retrieval_batch_stats(const bear_retrieval* r, bear_batch_stats* out);

/* ---------------------------------------------------------------- direct spectra
 * pyBeAR's direct forward models (python/src/bear_python_wrapper.cpp:343-372):
 * EmissionModel, TransmissionModel and OccultationModel built on a spectral
 * grid sampled between two wavelengths (SpectralGrid(config, wavelength_min,
 * wavelength_max), spectral_grid.cpp:52-72), without observations, whose
 * calcSpectrum takes the atmosphere's structure from the caller
 * (Atmosphere::setAtmosphericStructure with FixedChemistry,
 * atmosphere.cpp:245-283, fixed_chemistry.cpp:33-89): per level the
 * pressure (bar, bottom first), the temperature and the mixing ratios of the
 * named species (chemistry symbols or FastChem symbols); n = p 1e6 / (k_B T)
 * per species, the mean molecular weight over all species, the hydrostatic
 * altitude.  No background gas is added: CIA partners (H2, He) must be
 * among the species.
 * forward_model_type: "emission", "transmission" or "secondary_eclipse";
 * opacity_species as in bear_model_config; the eclipse takes the star's
 * flux (W m^-2 micron^-1) at the given wavelengths (micron) (the reference's
 * StarSpectrumFile, star_file_spectrum.cpp:57-72), the other models NULL. */
typedef struct bear_spectrum_model bear_spectrum_model;

int bear_spectrum_model_create(const bear_global_config* config, const char* forward_model_type,
                               double wavelength_min, double wavelength_max, size_t nb_grid_points,
                               const char* opacity_species, const double* stellar_wavelengths,
                               const double* stellar_flux, size_t nb_stellar_points, bear_spectrum_model** out);
void bear_spectrum_model_destroy(bear_spectrum_model* m);
size_t bear_spectrum_model_nb_spectral_points(const bear_spectrum_model* m);
int bear_spectrum_model_spectral_grid(const bear_spectrum_model* m, double* wavenumbers, double* wavelengths);

/* EmissionModel::calcSpectrum (emission.cpp:268-340): spectrum [points] is
 * the flux in W m^-2 cm times radius^2 / distance^2 (same units for both).
 * surface_gravity in cm s^-2; radius is also the bottom radius of the
 * variable gravity.  mixing_ratios [nb_levels][nb_species].
 * cloud_optical_depth [nb_levels - 1][points] (layer above level l in row
 * l, points in the grid's order) may be NULL; when any value is positive it
 * is the clouds' optical depth of each layer (FixedCloudModel,
 * fixed_cloud_model.cpp:36-75; emission.cpp:344-367), added to the gas's
 * (emission, eclipse) or, divided by the layer's thickness, to the upper
 * level's extinction (transmission, cloud_model.h:67-88). */
int bear_emission_calc_spectrum(bear_spectrum_model* m, double surface_gravity, double radius, double distance,
                                const double* pressure, const double* temperature, size_t nb_levels,
                                const char* const* species, size_t nb_species, const double* mixing_ratios,
                                const double* cloud_optical_depth, int use_variable_gravity, double* spectrum);

/* TransmissionModel::calcSpectrum (transmission.cpp:371-465): spectrum
 * [points] is the transit depth in ppm; the star's radius is
 * planet_radius / radius_ratio. */
int bear_transmission_calc_spectrum(bear_spectrum_model* m, double surface_gravity, double planet_radius,
                                    double radius_ratio, const double* pressure, const double* temperature,
                                    size_t nb_levels, const char* const* species, size_t nb_species,
                                    const double* mixing_ratios, const double* cloud_optical_depth,
                                    int use_variable_gravity, double* spectrum);

/* OccultationModel::calcSpectrum (secondary_eclipse.cpp:395-482): spectrum
 * [points] is the eclipse depth in ppm against the model's stellar flux. */
int bear_occultation_calc_spectrum(bear_spectrum_model* m, double surface_gravity, double radius_ratio,
                                   const double* pressure, const double* temperature, size_t nb_levels,
                                   const char* const* species, size_t nb_species, const double* mixing_ratios,
                                   const double* cloud_optical_depth, double* spectrum);

/* ---------------------------------------------------------------- measurement
 * No counterpart in the reference (it has no timing code).  With profiling
 * on, HIP events on the retrieval's stream time every launch of the fused
 * opacity + radiative-transfer kernel; the statistics also report how many
 * distinct table rows the last evaluation's stencils touched, which fixes
 * the kernel's compulsory HBM traffic. */
typedef struct bear_eval_stats {
  size_t nb_spectral_points;
  size_t nb_levels;
  size_t nb_species;             /* absorbers with tables in the evaluation */
  size_t nb_bins;                /* observation points, all observations */
  size_t distinct_table_rows;    /* rows referenced by the last evaluation */
  size_t table_value_bytes;      /* bytes per stored table value */
  size_t kernel_launches;        /* timed launches since profiling was enabled */
  double kernel_ms_total;        /* their summed duration */
  size_t table_device_bytes;     /* device memory of the resident tables (FP64 rows, reduced copies) */
} bear_eval_stats;

int bear_retrieval_set_profiling(bear_retrieval* r, int enable);
int bear_retrieval_stats(const bear_retrieval* r, bear_eval_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* BEAR_HIP_H */
