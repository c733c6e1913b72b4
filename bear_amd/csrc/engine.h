// Device side of the engine: memory owned on one GPU and the kernels of one
// likelihood evaluation.  No HIP types appear here, so host-only C++ can
// drive it.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "host.h"

namespace bear_amd {

// Cloud layers (cloud_model/*.cpp): at most kMaxClouds parametric cloud
// models per forward model.  Model m's optical depth in a layer is its
// per-layer weight (LevelState::cloud_w) times a spectral factor f_m(lambda)
// that the kernels form per wavenumber: 1 (grey), lambda^c0 c1 (power law:
// c0 the exponent, c1 = lambda_ref^-c0), c3 / (c0 x^-c1 + x^0.2) with
// x = c2 / lambda (KHnongrey: q0, a0, 2 pi a, the reference efficiency).
constexpr int kMaxClouds = 4;
enum CloudKind : int { kCloudGrey = 0, kCloudKH = 1, kCloudPowerLaw = 2 };
struct CloudSpec {
  int count = 0;
  int kind[kMaxClouds] = {};
  double c[kMaxClouds][4] = {};
};

// What the single-point kernels add to the tabulated opacity: the H-
// continuum (rows [7][N], or none), parametric cloud layers and caller-given
// cloud optical depths [levels - 1][N] (or none).
struct Extras {
  const double* hminus;
  const double* cloud_tau;
  CloudSpec clouds;
};

// Per level: temperature and altitude of the model atmosphere, and whether
// any stencil slot's table rows differ from the previous level's (always at
// level 0): the kernels then reload every slot's rows at once, so one scalar
// decides and the loads issue together.
struct LevelState {
  double temperature;
  double inv_temperature;
  double altitude;
  double rayleigh_density;   // number density of the Rayleigh scatterer (transmission)
  uint32_t reload;           // bit s: slot s's rows changed
  uint32_t pad;
  // bilinear weights x log2(10) (as LevelSpecies::w) shared by every slot of
  // stencil group 0 and 1 (StencilGroups): slots whose species share a table
  // grid have identical weights at every level
  double group_w[2][4];
  // LevelSpecies::log2_n of slots 0-15, contiguous: one scalar load per level
  double log2_n[16];
  // the same two for the reduced-precision tables (LevelSpecies32 values)
  float group_w32[2][4];
  float log2_n32[16];
  // LevelSpeciesAdjoint::u shared by the slots of stencil group 0 and 1
  double group_u[2][4];
  // the H- continuum's weights of its seven per-wavenumber rows: the H-
  // density, then 1e-29 (5040/T)^((i+1)/2) p_e n_H for i = 1..6 (all zero
  // where the level holds no H-)
  double hminus_c[7];
  // clouds: per cloud model its optical-depth weight in the layer below this
  // level (zero at level 0 and outside the cloud), and the factor that turns
  // that layer's cloud optical depth into this level's term: 1 for emission,
  // 1 / (z_l - z_{l-1}) for transmission (an extinction, cloud_model.h:67-88)
  double cloud_w[kMaxClouds];
  double cloud_scale;
};

// log2 of a zero effective density in the FP64 stencil records: finite, so
// that the kernels' exponentials need no clamp, and low enough that
// 2^(kZeroDensityLog2 + log2 sigma) underflows to exactly 0 for any
// cross-section below 2^900 cm^2.
constexpr double kZeroDensityLog2 = -2000.0;

// Transmission geometry (transmission_spectrum.cpp:107-154): tangent path t
// (t = 0 .. L-2) crosses the segments q = t .. L-2 (level q to q+1) with
// chord lengths c[t][q].  The device keeps them transposed, chord[q][t]
// (zero for q < t), with the segment count padded to kSegmentStep and the
// tangent count to kTangentBlock, zeros in the padding.
constexpr int kTangentBlock = 8;       // tangent paths per thread
constexpr int kSegmentStep = 4;        // segments per step
constexpr int chordSegments(int nb_levels) {
  return (nb_levels - 1 + kSegmentStep - 1) / kSegmentStep * kSegmentStep;
}
constexpr int chordTangents(int nb_levels) {
  return (nb_levels - 1 + kTangentBlock - 1) / kTangentBlock * kTangentBlock;
}

// Which stencil slots take the level's shared weights: grouped != 0 when the
// tabulated species use at most two distinct (T, p) grids; slot s then uses
// LevelState::group_w[(group1 >> s) & 1], else its own LevelSpecies::w.
// Kernels read the shared weights with one scalar load per level instead of
// one per slot.
struct StencilGroups {
  uint32_t grouped = 0;
  uint32_t group1 = 0;
};

// Per (level, opacity species) for the forward kernels: the table stencil
// and the effective number density n (n_s, or n_s times the CIA partners'
// densities) in the form the kernels use,
//   sigma n = 2^(log2_n + sum_k w_k v_k),
// v the four stencil rows' log10 cross-sections, w the bilinear weights
// log2(10) {(1-f_t)(1-f_p-), (1-f_t) f_p-, f_t (1-f_p+), f_t f_p+}
// (opacity_species.cpp:277-345), log2_n = kZeroDensityLog2 for n == 0, whose
// term 2^(log2_n + ...) is exactly 0, which removes the species' term as
// opacity_species.cpp:409-412 skips it.  block[k] locates
// stencil row k in the tables: row * stride / 64 (the stride is a multiple of
// 64), so an address is one shift.  reload != 0 when the stencil rows differ
// from the previous level's (always at level 0); the kernels keep the four
// row values in registers until LevelState::reload says a level brings new
// rows.  64 bytes, aligned.
struct alignas(64) LevelSpecies {
  double w[4];
  double log2_n;
  uint32_t block[4];
  uint32_t reload;
  uint32_t pad;
};

// What the adjoint adds to a LevelSpecies: the weights of
// d log2 sigma / dT = sum_k u_k v_k,
// u = log2(10) (d f_t / dT) {-(1-f_p-), -f_p-, 1-f_p+, f_p+}, d f_t / dT =
// 1 / (T+ - T-) (0 when the T rows coincide or T is clamped to the table).
struct alignas(64) LevelSpeciesAdjoint {
  double u[4];
  double pad[4];
};

// Reduced-precision form of LevelSpecies: the bilinear weights
// (1-f_t)(1-f_p-), (1-f_t) f_p-, f_t (1-f_p+), f_t f_p+ times log2(10), and
// log2 of the effective density (-inf when it is zero), so that
// sigma n = 2^(log2_n + base + sum_k w[k] v[k]) for the stored residuals v.
struct alignas(16) LevelSpecies32 {
  float w[4];
  float log2_n;
  uint32_t block[4];                 // row * stride / 64, as in LevelSpecies
  uint32_t reload;
  uint32_t pad[2];
};

// Storage of the log10 cross-section tables read by the batched kernels.
// fp64: the tables as loaded, FP64 arithmetic throughout (the single-model
// path's arithmetic).  fp32 / fp16 / bf16: per (species, wavenumber) an FP32
// base, the midrange of that species' rows there, and per row the residual
// log10 sigma - base in the named type; interpolation and exponentials in
// FP32, radiative transfer, bins and likelihood in FP64.
// fp32 / fp16 / bf16: residuals against a per-(species, wavenumber) FP32
// base; int16: 16-bit integers scaled per (species, wavenumber), value =
// base + scale q
enum class TablePrecision : int { fp64 = 0, fp32 = 1, fp16 = 2, bf16 = 3, int16 = 4 };
inline size_t tableElementBytes(TablePrecision p) {
  return p == TablePrecision::fp64 ? 8 : p == TablePrecision::fp32 ? 4 : 2;
}

enum class ModelKind : int { emission = 0, eclipse = 1, transmission = 2 };

// Scalars of one live point of a batch.
struct BatchModel {
  double scaling;                // emission: (R_J / d)^2 f; eclipse: 1
  double star_inv_temperature;   // eclipse: 1 / T_*
  double depth_scale;            // eclipse: (Rp / Rs)^2 1e6
  double bottom_radius;          // transmission: R_p (cm)
  double star_radius;            // transmission: R_* (cm)
  double inflation;              // added variance (error inflation), or 0
  double pad[2];
  CloudSpec clouds;              // the point's cloud models (their weights per level are in its LevelState)
};

// Device staging of a batch: each live point's atmosphere (temperature
// profile, chemistry, altitudes), stencils and scalars formed on the device
// from its parameters, the model's fixed data uploaded once as one blob
// (this header's counts and byte offsets into it).  The same steps as the
// host's Atmosphere::compute and TabulatedModel::stageLevels.
struct StageModule { int background, sodium_free, first, count; };        // chemistry modules in order
struct StageSpecies { int species_id, partner0, partner1, grid, row_offset, pad[3]; };
// a table grid: temperature rows (their temperatures at t_offset), where
// each row's pressures start in the [t][p] flattening (tstart_offset), the
// rows' temperatures in that flattening (temps_offset) and the picks [L][nt]
struct StageGrid { int nt, t_offset, picks_offset, tstart_offset, temps_offset, pad[3]; };
struct StagePick { uint32_t lower, upper; double fraction; };           // per level and temperature row
struct StagePlan {
  int L = 0, S = 0, NS = 0, kind = 0;             // kind: ModelKind
  int isothermal = 0, basis_entries = 0, nb_temp = 0, nb_chem = 0;
  int g0 = 0, nb_params = 0, error_inflation = 0, nb_obs = 0;
  int alt_kind = 0;                               // 0 hydrostatic, 1 fixed mean molecular weight, 2 scale height
  int variable_gravity = 0, grouped = 0, rayleigh_id = -1;
  uint32_t group1 = 0, blocks_per_row = 0;
  int group_grid[2] = {-1, -1};
  int nb_modules = 0, nb_active = 0, n_ids = 0, mods_offset = 0;
  int na_id = 0, k_id = 0, h2_id = 0, he_id = 0;
  // byte offsets into the blob
  size_t pressure = 0, log_ratio = 0, basis_index = 0, basis_weight = 0, modules = 0, module_species = 0,
         active = 0, id_slot = 0, mol_weight = 0, species = 0, grids = 0, grid_t = 0, grid_tstart = 0, row_ids = 0,
         row_temps = 0, picks = 0, obs_mods = 0;
};

// Launch shape of the batched spectrum kernel, picked by the autotuner.
struct BatchLaunch {
  int block = 256;               // threads per workgroup
  int tile = 256;                // wavenumbers per work item
  int grid = 0;                  // persistent workgroups
  int group = 32;                // points per group of the work-item order (transmission)
};

// Per bin of every observation, in observation order: the folded operator
// segment and the likelihood inputs.  The segment is summed in pieces of at
// most kBinChunk weights, BinChunk records first_chunk .. first_chunk +
// ceil(length / kBinChunk) - 1, so that one wide band (a photometric filter
// over a million wavenumbers) is summed by many workgroups.
constexpr uint32_t kBinChunk = 2048;
struct BinDesc {
  uint64_t weight_offset;
  uint32_t start;
  uint32_t length;
  uint32_t observation;
  uint32_t first_chunk;
  double data, error, weight;
};

struct BinChunk {
  uint64_t weight_offset;
  uint32_t start;
  uint32_t length;
};

class DeviceEngine {
 public:
  explicit DeviceEngine(int device);
  ~DeviceEngine();
  DeviceEngine(const DeviceEngine&) = delete;
  DeviceEngine& operator=(const DeviceEngine&) = delete;

  void uploadGrid(const std::vector<double>& wavenumber, const std::vector<double>& wavelength);
  // Storage of the cross-section tables, set before they are loaded: FP64
  // (every path; reduced precisions split from it on the device by
  // setTablePrecision), or FP32/FP16/BF16 only (the batched forward kernels
  // at that precision; no FP64 rows are kept, so the single-point and
  // adjoint kernels are unavailable).
  void setTableStorage(TablePrecision p);
  TablePrecision tableStorage() const { return storage_; }
  // device tables [rows][stride]; species `slot`'s rows [first, first + count)
  // from host FP64 [count][N], split into base and residuals on the device
  // when the storage is reduced
  void allocateTables(size_t nb_rows, size_t nb_species);
  void uploadSpeciesRows(int slot, size_t first, size_t count, const double* rows);
  size_t tableBytes() const;                     // device bytes of the resident tables
  void uploadRayleigh(const std::vector<double>& cross_section);   // [N] cm^2, empty: no scatterer
  // the H- continuum's rows [7][N] (bound-free cross-section, six free-free
  // polynomials); empty: no H-.  The single-point kernels only.
  void uploadHminus(const std::vector<double>& rows);
  bool hasHminus() const { return hminus_ != nullptr; }
  // parametric cloud models of the configured forward model (their number),
  // and the spectral factors of the next single-point evaluation
  void setCloudModels(int count) { cloud_models_ = count; }
  void setCloudSpec(const CloudSpec& spec) { cloud_spec_ = spec; }
  // the direct models' caller-given layer optical depths [levels - 1][N]
  // for the next evaluations (FixedCloudModel); NULL or all zero: none
  void setFixedCloud(const double* tau);
  void uploadObservations(const std::vector<ObservationSetup>& observations);
  void configure(size_t nb_levels, size_t nb_species, bool transmission);
  void setStencilGroups(StencilGroups g) { groups_ = g; }
  StencilGroups stencilGroups() const { return groups_; }

  // pinned staging of the per-evaluation inputs; filled by the caller
  LevelState* levels() { return levels_host_; }
  LevelSpecies* levelSpecies() { return species_host_; }    // [level][speciesStride()]
  size_t speciesStride() const { return species_stride_; }
  size_t tableStride() const { return stride_; }          // elements per table row, a multiple of 64
  double* shifts() { return shift_host_; }
  double* pathLengths() { return path_host_; }              // chord[segment][tangent] (transmission)

  // One evaluation: the high-resolution spectrum on the grid (kept on the
  // device; also written to spectrum_dev when given), the bins (also to
  // bins_dev[observation] when given), the per-bin likelihood terms and
  // their fixed-order sum.
  //   emission: flux, scaled; spectrum_dev in W m^-2 micron^-1
  //   eclipse:  planet flux (scale 1) against a blackbody star, depths in ppm
  //   transmission: transit depth in ppm
  void evaluateEmission(double scaling, double error_inflation, double* spectrum_dev, double* const* bins_dev);
  void evaluateEclipse(double star_temperature, double radius_ratio, double error_inflation, double* spectrum_dev,
                       double* const* bins_dev);
  void evaluateTransmission(double bottom_radius, double star_radius, double error_inflation, double* spectrum_dev,
                            double* const* bins_dev);
  // eclipse depth against the tabulated stellar flux of uploadStellarSpectrum
  // (no bins): spectrum_dev = F / F_star (Rp/Rs)^2 1e6
  void evaluateEclipseTabulatedStar(double radius_ratio, double* spectrum_dev);
  void uploadStellarSpectrum(const std::vector<double>& flux);   // [N], W m^-2 cm
  void copyFlux(double* host);                   // the last evaluation's flux [N] (emission, eclipse)
  double logLikelihood();                        // waits for the evaluation
  void copyBins(std::vector<double>& out);       // internal (descending) order, all observations
  void copySpectrum(double* host, const double* spectrum_dev);
  void synchronize();
  double* spectrumBuffer();                      // engine-owned [N], allocated on first use

  // No-allocation mode of the single-model path: while on, the evaluate*
  // calls keep their per-evaluation device buffers (staged inputs, flux, bin
  // piece sums, bins, log-likelihood) in the library's static scratch array
  // on this device instead of the engine's own, and allocate nothing.  On
  // raises InvalidInput when this model's buffers exceed the array
  // (scratchBytes()).  Calls that overlap on one GPU share the array and race.
  void useScratch(bool on);
  size_t scratchRequired() const;
  static size_t scratchBytes();

  // ---- batched evaluation: many live points per launch
  // Table storage of the batched kernels; species_rows[s] = [first, end) of
  // species s's rows in the table store.
  void setTablePrecision(TablePrecision p, const std::vector<std::pair<uint32_t, uint32_t>>& species_rows);
  TablePrecision tablePrecision() const { return precision_; }
  void setBatchOptions(bool use_graph, bool autotune) { use_graph_ = use_graph; autotune_ = autotune; }
  // pinned staging for up to `batch` live points, one record per point
  void reserveBatch(size_t batch);
  BatchModel* batchModel(size_t b) { return reinterpret_cast<BatchModel*>(record(b)); }
  LevelState* batchLevels(size_t b) { return reinterpret_cast<LevelState*>(record(b) + rec_levels_); }
  LevelSpecies* batchSpecies(size_t b);          // fp64 tables
  LevelSpeciesAdjoint* batchSpeciesAdjoint(size_t b);   // fp64 tables, for the adjoint
  LevelSpecies32* batchSpecies32(size_t b);      // reduced-precision tables
  double* batchShifts(size_t b) { return reinterpret_cast<double*>(record(b) + rec_shifts_); }
  // spectra, bins and log-likelihoods of live points [0, batch)
  void evaluateBatch(ModelKind kind, size_t batch);
  // Device staging (StagePlan): the model's fixed data once, then batches
  // whose records the device forms from the parameters placed in
  // stagedParameters() ([batch][nb_params], pinned); evaluateBatch(kind,
  // batch, true) stages and evaluates, and stagedNeglect() holds each
  // point's neglect flag afterwards.
  void uploadStagePlan(const StagePlan& plan, const std::vector<unsigned char>& blob);
  bool hasStagePlan() const { return stage_blob_ != nullptr; }
  void reserveStage(size_t batch);
  double* stagedParameters() { return stage_params_host_; }
  const int* stagedNeglect() const { return stage_neglect_host_; }
  void evaluateBatch(ModelKind kind, size_t batch, bool staged);
  const double* batchLogLikelihood() const { return batch_ll_host_; }
  void copyBatchBins(size_t batch, std::vector<double>& out);          // [batch][bins], internal order
  void copyBatchSpectrum(ModelKind kind, size_t batch, double* host);  // [batch][N], output units
  // Value and adjoint of the staged batch (emission, FP64 tables): the
  // log-likelihoods (batchLogLikelihood()) and per point dL/d(level
  // quantities) [batch][levels][speciesStride() + 4]: n dL/dn for the
  // effective density n of each stencil slot, the temperature through the table interpolation
  // (recomputing: and through the Planck function), the altitude, the flux
  // scaling (level 0 only; 0 elsewhere), the temperature through the Planck
  // function (0 when recomputing); also dL/d(bin) and the bins,
  // [batch][bins].  The forward pass keeps every level's intensities and
  // opacity for the backward pass, in chunks of points that fit the budget;
  // a point that does not fit alone is recomputed inside the backward kernel
  // instead.  kind: emission, or eclipse (planet flux at scale 1 against the
  // blackbody star; bin_terms then gets per (point, bin) dL/dbin times the
  // depth without shift, and dL/dbin times d depth / dT_star,
  // [batch][bins][2]), or transmission (always kept; the level slots are
  // n dL/dn, the temperature through the tables, dL/dc_t for the trapezoid
  // coefficient c_t = (R + z_t)(z_{t+1} - z_{t-1}) of tangent path t (level
  // L-1: dL/dH for the R^2 term), at level L-1 only the sum of dL/ddepth x
  // depth, and dL/dn_R of the Rayleigh scatterer; bin_terms gets dL/dchord
  // [batch][L-1 segments][L-1 tangents], zero where the segment lies below
  // the tangent level).
  // (staged: the records formed on the device from stagedParameters(), the
  // neglect flags in stagedNeglect() afterwards)
  void evaluateAdjointBatch(ModelKind kind, size_t batch, std::vector<double>& level_grad,
                            std::vector<double>& bin_grad, std::vector<double>& bins, std::vector<double>& bin_terms,
                            bool staged = false);
  // HBM budget (bytes) of the kept forward quantities; 0: always recompute
  void setAdjointMemory(size_t bytes) { adjoint_budget_ = bytes; }
  size_t adjointMemory() const { return adjoint_budget_; }
  size_t adjointChunk() const { return adj_last_chunk_; }   // points per chunk of the last call; 0: recomputed
  BatchLaunch batchLaunch() const { return last_launch_; }
  size_t batchEvaluations() const { return batch_evaluations_; }
  double batchMilliseconds() const { return batch_ms_; }
  bool batchUsedGraph() const { return last_used_graph_; }

  void setProfiling(bool on);
  size_t timedLaunches() const { return timed_launches_; }
  double timedMilliseconds() const { return timed_ms_; }
  size_t distinctRows() const { return distinct_rows_; }
  size_t nbLevels() const { return nb_levels_; }
  size_t nbSpecies() const { return nb_species_; }
  size_t nbPoints() const { return n_points_; }
  size_t nbBins() const { return n_bins_; }
  int device() const { return device_; }

 private:
  // per-evaluation device buffers of the single-model path
  struct EvalBuffers {
    unsigned char* staging;
    double* flux;
    double* chunk_sums;
    double* bins;
    double* ll;
  };
  EvalBuffers buffers() const;
  bool use_scratch_ = false;
  unsigned char* scratch_ = nullptr;             // this device's static scratch array
  void upload();                                 // staging -> device, on the stream
  void beginTimed();
  void endTimed();
  void launchBins(int mode, double star_temperature, double depth_scale, double inflation, double* const* bins_dev);

  int device_ = 0;
  void* stream_ = nullptr;
  size_t n_points_ = 0, stride_ = 0, n_bins_ = 0, n_obs_ = 0, n_rows_ = 0, n_chunks_ = 0;
  size_t nb_levels_ = 0, nb_species_ = 0, species_stride_ = 0;
  bool transmission_ = false;
  StencilGroups groups_;
  double* wavenumber_ = nullptr;
  double* wavelength_ = nullptr;
  double* tables_ = nullptr;
  double* rayleigh_ = nullptr;
  double* hminus_ = nullptr;                     // [7][N], or none
  int cloud_models_ = 0;
  CloudSpec cloud_spec_;
  double* cloud_tau_ = nullptr;                  // [L-1][N] caller-given cloud optical depths
  bool fixed_cloud_ = false;
  Extras extras() const { return Extras{hminus_, fixed_cloud_ ? cloud_tau_ : nullptr, cloud_spec_}; }
  double* flux_ = nullptr;
  double* spectrum_ = nullptr;
  double* star_ = nullptr;                       // tabulated stellar flux [N], or none
  double* weights_ = nullptr;
  BinDesc* bins_desc_ = nullptr;
  double* bins_ = nullptr;
  BinChunk* bin_chunks_ = nullptr;               // [n_chunks_]
  double* chunk_sums_ = nullptr;                 // [n_chunks_][2]
  double* ll_ = nullptr;
  std::vector<size_t> obs_bin_offset_;
  // per-evaluation inputs: one pinned block, one device block
  unsigned char* staging_host_ = nullptr;
  unsigned char* staging_dev_ = nullptr;
  size_t staging_bytes_ = 0;
  size_t off_species_ = 0, off_shift_ = 0, off_path_ = 0;
  LevelState* levels_host_ = nullptr;
  LevelSpecies* species_host_ = nullptr;
  double* shift_host_ = nullptr;
  double* path_host_ = nullptr;
  double* ll_host_ = nullptr;
  double* ll_host_dev_ = nullptr;                // ll_host_ as the device addresses it
  // measurement
  bool profiling_ = false, event_pending_ = false;
  void* event_start_ = nullptr;
  void* event_stop_ = nullptr;
  size_t timed_launches_ = 0, distinct_rows_ = 0;
  double timed_ms_ = 0;
  std::vector<uint8_t> row_seen_;
  void collectTiming();
  void countRows();

  // batched evaluation
  unsigned char* record(size_t b) { return batch_host_ + b * rec_bytes_; }
  // points [first, first + batch) of the staged records; keep: also store
  // every level's I - B of both angles and opacity, FP32
  // [batch][64-wavenumber tile][levels][3][64] (emission; transmission: the extinction in
  // the first of the three values)
  // nu_count > 0 (transmission): only wavenumbers [nu_first, nu_first +
  // nu_count) of one point, kept from the start of `keep` (nu_first a
  // multiple of the launch's tile)
  // (transmission with keep: also each kept tile's reach, int32
  // [point][tiles of 64][paths], the largest last segment of every path)
  void launchBatchSpectrum(ModelKind kind, size_t batch, const BatchLaunch& launch, size_t first = 0,
                           float* keep = nullptr, size_t nu_first = 0, size_t nu_count = 0, int* reach = nullptr);
  void launchBatchRest(ModelKind kind, size_t batch, size_t first = 0);
  BatchLaunch tuned(ModelKind kind, size_t batch);
  BatchLaunch tune(ModelKind kind, size_t batch);
  void releaseGraph();
  void releaseBatch();
  int nb_cus_ = 0;
  size_t lds_limit_ = 64 * 1024;                 // LDS bytes per workgroup (device attribute)
  TablePrecision precision_ = TablePrecision::fp64;
  TablePrecision storage_ = TablePrecision::fp64;
  size_t base_rows_ = 0;                         // rows of table_base_
  void allocateReduced(TablePrecision p, size_t nb_rows, size_t slots);
  void splitSpecies(const double* rows, size_t count, int slot, size_t first);
  void requireFP64Tables(const char* what) const;
  void* tables_low_ = nullptr;                   // residuals [rows][stride] (fp32 / fp16 / bf16)
  float* table_base_ = nullptr;                  // [species stride][stride]
  bool use_graph_ = true, autotune_ = true;
  size_t batch_capacity_ = 0, rec_bytes_ = 0, rec_levels_ = 0, rec_species_ = 0, rec_adjoint_ = 0, rec_shifts_ = 0;
  StagePlan stage_plan_;
  unsigned char* stage_blob_ = nullptr;
  double* stage_params_ = nullptr;               // [capacity][nb_params]
  double* stage_params_host_ = nullptr;          // pinned
  int* stage_neglect_ = nullptr;
  int* stage_neglect_host_ = nullptr;            // pinned
  size_t stage_capacity_ = 0;
  void launchStage(size_t batch);
  void releaseStage();
  unsigned char* batch_host_ = nullptr;          // pinned records
  unsigned char* batch_dev_ = nullptr;
  double* batch_flux_ = nullptr;                 // [capacity][stride]
  double* batch_bins_ = nullptr;                 // [capacity][bins]
  double* batch_chunk_sums_ = nullptr;           // [capacity][pieces][2]
  double* batch_ll_ = nullptr;                   // [capacity]
  double* batch_ll_host_ = nullptr;              // pinned [capacity]
  double* batch_chord_ = nullptr;                // transmission: [capacity][L][L] FP64
  float* batch_chord32_ = nullptr;               // and FP32
  struct TuneKey { int kind, precision; size_t batch; };
  std::vector<std::pair<TuneKey, BatchLaunch>> tuned_;
  void* graph_exec_ = nullptr;
  TuneKey graph_key_{-1, -1, 0};
  bool graph_staged_ = false;
  BatchLaunch last_launch_;
  bool last_used_graph_ = false;
  size_t batch_evaluations_ = 0;
  double batch_ms_ = 0;
  // adjoint: the observation operator transposed (per wavenumber: bins and
  // weights) and work buffers
  uint64_t* adj_ptr_ = nullptr;                  // [N + 1]
  uint32_t* adj_bin_ = nullptr;
  double* adj_weight_ = nullptr;
  double* adj_fbar_ = nullptr;                   // [capacity][stride]
  double* adj_gbar_ = nullptr;                   // [capacity][bins]
  double* adj_fscale_ = nullptr;                 // eclipse: dL/d(planet bin) [capacity][bins]
  double* adj_terms_ = nullptr;                  // eclipse: [capacity][bins][2]
  double* adj_slope_ = nullptr;                  // eclipse: [capacity][pieces]
  double* adj_intensity_ = nullptr;              // kept: FP32 [chunk][tile][levels][3][64]; recomputing: [grid][levels][2][64]
  double* adj_partial_ = nullptr;                // [chunk][levels][slots][tiles]
  double* adj_out_ = nullptr;                    // [capacity][levels][slots]
  size_t adj_capacity_ = 0, adj_grid_ = 0, adj_partial_size_ = 0, adj_intensity_size_ = 0;
  size_t adjoint_budget_ = size_t(96) << 30, adj_last_chunk_ = 0;
  void releaseAdjoint();
};

}  // namespace bear_amd
