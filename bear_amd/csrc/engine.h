// Device side of the engine: memory owned on one GPU and the kernels of one
// likelihood evaluation.  No HIP types appear here, so host-only C++ can
// drive it.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "host.h"

namespace bear_amd {

// Per level: temperature and altitude of the model atmosphere.
struct LevelState {
  double temperature;
  double inv_temperature;
  double altitude;
  double rayleigh_density;   // number density of the Rayleigh scatterer (transmission)
};

// Per (level, opacity species): effective number density (n_s, or n_s times
// the CIA partners' densities) and the table stencil.  n == 0 removes the
// species' term, as opacity_species.cpp:409-412 skips it.  reload != 0 when
// the stencil rows differ from the previous level's (always at level 0):
// the kernels keep the four row values in registers otherwise.
struct LevelSpecies {
  double n;
  double f_p_lower, f_p_upper, f_t;
  uint32_t row[4];
  uint32_t reload;
  uint32_t pad;
};

// Per bin of every observation, in observation order: the folded operator
// segment and the likelihood inputs.
struct BinDesc {
  uint64_t weight_offset;
  uint32_t start;
  uint32_t length;
  uint32_t observation;
  uint32_t pad;
  double data, error, weight;
};

class DeviceEngine {
 public:
  explicit DeviceEngine(int device);
  ~DeviceEngine();
  DeviceEngine(const DeviceEngine&) = delete;
  DeviceEngine& operator=(const DeviceEngine&) = delete;

  void uploadGrid(const std::vector<double>& wavenumber, const std::vector<double>& wavelength);
  void uploadTables(const TableStore& store);
  void uploadRayleigh(const std::vector<double>& cross_section);   // [N] cm^2, empty: no scatterer
  void uploadObservations(const std::vector<ObservationSetup>& observations);
  void configure(size_t nb_levels, size_t nb_species, bool transmission);

  // pinned staging of the per-evaluation inputs; filled by the caller
  LevelState* levels() { return levels_host_; }
  LevelSpecies* levelSpecies() { return species_host_; }    // [level][speciesStride()]
  size_t speciesStride() const { return species_stride_; }
  double* shifts() { return shift_host_; }
  double* pathLengths() { return path_host_; }              // [level][level] (transmission)

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
  double logLikelihood();                        // waits for the evaluation
  void copyBins(std::vector<double>& out);       // internal (descending) order, all observations
  void copySpectrum(double* host, const double* spectrum_dev);
  void synchronize();
  double* spectrumBuffer();                      // engine-owned [N], allocated on first use

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
  void upload();                                 // staging -> device, on the stream
  void beginTimed();
  void endTimed();
  void launchBins(int mode, double star_temperature, double depth_scale, double inflation, double* const* bins_dev);

  int device_ = 0;
  void* stream_ = nullptr;
  size_t n_points_ = 0, stride_ = 0, n_bins_ = 0, n_obs_ = 0, n_rows_ = 0;
  size_t nb_levels_ = 0, nb_species_ = 0, species_stride_ = 0;
  bool transmission_ = false;
  double* wavenumber_ = nullptr;
  double* wavelength_ = nullptr;
  double* tables_ = nullptr;
  double* rayleigh_ = nullptr;
  double* flux_ = nullptr;
  double* spectrum_ = nullptr;
  double* weights_ = nullptr;
  BinDesc* bins_desc_ = nullptr;
  double* bins_ = nullptr;
  double* ll_terms_ = nullptr;
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
  // measurement
  bool profiling_ = false, event_pending_ = false;
  void* event_start_ = nullptr;
  void* event_stop_ = nullptr;
  size_t timed_launches_ = 0, distinct_rows_ = 0;
  double timed_ms_ = 0;
  std::vector<uint8_t> row_seen_;
  void collectTiming();
  void countRows();
};

}  // namespace bear_amd
