// Host side of the engine: spectral grid, observations, opacity tables,
// atmosphere and the forward models that drive the device kernels.
//
// Reference citations are relative to the reference repository's src/.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace bear_amd {

// ------------------------------------------------------------------ errors
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] void raise(int code, const std::string& msg);

constexpr int kInvalidInput = -1, kFileNotFound = -2, kDevice = -3, kOutOfMemory = -4, kInternal = -5;

// ------------------------------------------------------------------ constants
// cgs constants as in additional/physical_const.h:28-52
namespace phys {
constexpr double planck_h = 6.62606896e-27;
constexpr double light_c = 2.99792458e10;
constexpr double boltzmann_k = 1.3806504e-16;
constexpr double gas_constant = 8.314472e7;
constexpr double radius_jupiter = 7.1492e9;
constexpr double pi = 3.14159265358979323846;
constexpr double avogadro = 6.022140857e23;
}  // namespace phys

// Chemical species ids: indices into a level's number densities, in the
// order of chemistry/chem_species.h:30-90 (with its molecular weights).
int speciesId(const std::string& symbol);           // throws InvalidInput
int speciesIdOrFastChem(const std::string& symbol); // also FastChem's symbols (H2O1, C1O2, ...)
int nbSpeciesIds();
double molecularWeight(int id);
namespace sid { constexpr int total = 0, H = 1, He = 2, H2 = 10, Na = 19, K = 20; }

// ------------------------------------------------------------------ grid
// SpectralGrid (spectral_grid/spectral_grid.cpp): the full list the tables
// are tabulated on and the sampled subset the models are computed on.
class ObservationSetup;

class SpectralGrid {
 public:
  std::vector<double> full_wavenumber, full_wavelength;
  std::vector<size_t> index;               // positions of the sampled points in the full list
  std::vector<double> wavenumber, wavelength;
  int discretisation = 0;
  double resolution = 0;

  void loadFullList(const std::string& path);
  void sample(std::vector<ObservationSetup>& observations);
  // SpectralGrid(config, wavelength_min, wavelength_max) (spectral_grid.cpp:52-72):
  // the list sampled between two wavelengths (micron), no observations
  void sampleRange(double wavelength_min, double wavelength_max);
  size_t size() const { return wavenumber.size(); }
  // nearest element of a monotone list, searching from position start
  static size_t nearest(double x, const std::vector<double>& list, size_t start);

 private:
  // createHighResGrid* over wavenumber ranges (findBinEdges + the sampling rule)
  std::vector<char> markRanges(const std::vector<std::array<double, 2>>& ranges) const;
  void takePoints(const std::vector<char>& keep);
};

// linear interpolation of (x, y) onto the descending new_x, zero outside x's
// range (spectral_grid_interpolate.cpp:60-106)
std::vector<double> interpolateDescending(std::vector<double> x, std::vector<double> y, const std::vector<double>& nx);

// ------------------------------------------------------------------ observations
enum class BandType { spectroscopy, band_spectroscopy, photometry };

// One observation after set-up, plus its folded observation operator: each
// bin's value is a dot product of the high-resolution spectrum with a
// contiguous weight segment.  The operator composes the bandpass
// (observations_filter_response.cpp:166-181), the Gaussian line-spread
// convolution (spectral_band_convolve.cpp:100-144) and the trapezoid band
// integral (spectral_band_integrate.cpp:39-107), which are linear maps of a
// non-negative spectrum.
class ObservationSetup {
 public:
  std::string name;
  BandType type = BandType::spectroscopy;
  std::vector<double> data, error, weight, fwhm;
  std::vector<std::array<double, 2>> edges_wavelength;   // descending, [left > right]
  std::vector<std::array<double, 2>> edges_wavenumber;
  std::vector<double> centres;
  bool ascending = false;
  double range_wavelength[2] = {0, 0};
  int modifier_parameters = 0;
  std::string detector;                                   // "energy" | "photon" | "" (no bandpass)
  std::vector<double> filter_wavelength, filter_response;

  // products of set-up on the sampled grid
  std::vector<std::array<size_t, 2>> edge_index;
  size_t range_index[2] = {0, 0};
  std::vector<double> profile_sigma;                      // [N] or empty
  std::vector<std::array<size_t, 2>> convolution_window;  // [range] or empty
  std::vector<double> filter, filter_weight;              // [N] or empty
  double filter_normalisation = 0;

  // folded operator: bin b covers grid points [segment_start[b], segment_start[b] + segment_length[b])
  std::vector<size_t> segment_start;
  std::vector<size_t> segment_length;
  std::vector<double> segment_weight;                     // concatenated
  std::vector<size_t> segment_offset;                     // into segment_weight

  size_t nbPoints() const { return data.size(); }
  void setupOnGrid(const SpectralGrid& grid);
  void foldOperator(const SpectralGrid& grid, bool is_flux);
};

struct ObservationInputView {                 // plain copy of bear_observation_input
  std::string name, type, detector, modifier;
  std::vector<double> wavelengths, data, error, weight, fwhm, filter_wavelength, filter_response;
  std::vector<std::array<double, 2>> bin_edges;
};
ObservationSetup observationFromInput(const ObservationInputView& in);
ObservationSetup observationFromFile(const std::string& file, const std::string& modifier,
                                     const std::string& retrieval_folder);

// ------------------------------------------------------------------ opacity
// One table row's place in the device table store plus its grid point.
struct TableRow { double temperature, pressure, log_pressure; uint32_t id; };

// Bracketing of a (p, T) point inside one species' table
// (opacity_species.cpp:160-273) and the interpolation weights of
// opacity_species.cpp:277-345: rows r[0..3] = (T-, p-), (T-, p+), (T+, p-), (T+, p+).
struct Stencil {
  uint32_t row[4];
  double f_p_lower, f_p_upper, f_t;
  double dft;                  // d f_t / dT: 1 / (T+ - T-), 0 when the T rows coincide
  // where the rows sit in by_temperature: [t_pos[k / 2]][p_pos[k]] for row k,
  // so that a species with the same grid takes its own rows at once
  uint16_t t_pos[2];
  uint16_t p_pos[4];
};

class OpacitySpecies {
 public:
  std::string symbol, folder;
  int species_id = 0;
  std::vector<int> cia_partners;
  bool rayleigh = false;
  std::vector<double> rayleigh_cross_section;        // [N], cm^2
  // the H- continuum (no tables): per wavenumber the bound-free cross
  // section (cm^2) and the six free-free wavelength polynomials, [7][N]
  bool hminus = false;
  std::vector<double> hminus_rows;
  std::vector<std::vector<TableRow>> by_temperature;  // ascending T, each ascending p

  bool hasTable() const { return !by_temperature.empty(); }
  Stencil bracket(double pressure, double temperature) const;
  // the pressure half of bracket() done once for the atmosphere's fixed
  // pressures (per level and temperature row), then bracket(pressures[level],
  // temperature) by bracketLevel with the same results
  void preparePressures(const std::vector<double>& pressures);
  Stencil bracketLevel(size_t level, double temperature) const;
  // the rows of this species at a stencil bracketed on a species with the same grid
  void takeRows(Stencil& st) const {
    for (int k = 0; k < 4; ++k) st.row[k] = by_temperature[st.t_pos[k / 2]][st.p_pos[k]].id;
  }
  // same temperatures and, per temperature, the same pressures: bracket()
  // then gives the same fractions for every (p, T)
  bool sameGrid(const OpacitySpecies& other) const;
  // preparePressures' picks: per level and temperature row the two pressure
  // rows and the log-pressure fraction
  struct PressurePick { uint16_t lower, upper; double fraction; };
  const std::vector<PressurePick>& pressurePicks() const { return pressure_picks_; }

 private:
  std::vector<PressurePick> pressure_picks_;          // [level][temperature row]
  void bracketTemperature(double t, size_t& lo_i, size_t& hi_i) const;
};

// One species' table before sampling: where its rows come from (row files
// of a filelist.dat folder, or a registered in-memory table), in listed
// order, and the first id they take in the device store.
struct TableSource {
  std::string symbol, dir;
  double mass = 0;
  bool log_storage = false;
  std::vector<std::string> files;          // per row; empty for registered data
  const float* registered = nullptr;       // [rows][registered_points]
  size_t registered_points = 0;
  uint32_t first_row = 0;
  size_t nb_rows = 0;
};

// filelist.dat or the registered table of s: fills s.by_temperature with
// row ids from first_row on (opacity_species.cpp:66-157)
TableSource scanSpeciesTable(OpacitySpecies& s, const std::string& cross_section_path, uint32_t first_row);
// the rows sampled at the grid's index list as log10 FP64, [rows][N]
// (cross_section_file.cpp:33-62, sampled_data.cpp:48-87)
void sampleSpeciesTable(const TableSource& src, const SpectralGrid& grid, std::vector<double>& out);
OpacitySpecies makeOpacitySpecies(const std::string& symbol, const std::string& folder, const SpectralGrid& grid);
// the H- continuum's per-level coefficients of those rows: [0] the H-
// density, [1..6] 1e-29 (5040/T)^((i+1)/2) p_e n_H (h_m.cpp:38-53, 86-137);
// false when the level has no H- (the reference then skips the species)
bool hminusCoefficients(double temperature, double n_hminus, double n_h, double n_e, double (&c)[7]);

// ------------------------------------------------------------------ atmosphere
// Piecewise Lagrange polynomial over log10(p) (additional/piecewise_poly.cpp:49-229)
class PiecewisePolynomial {
 public:
  PiecewisePolynomial(size_t nb_elements, size_t order, double bottom_pressure, double top_pressure);
  void setValues(const std::vector<double>& dof_values);
  double value(double x) const;
  size_t nbDof() const { return vertices_.size(); }
  // value(x) = sum_k dof[index[k]] weight[k], the same products summed in
  // the same order as value() does, for an x fixed in advance
  struct Basis { std::vector<size_t> index; std::vector<double> weight; };
  Basis basis(double x) const;
 private:
  struct Element { std::vector<double> ref, vertex, value; };
  std::vector<Element> elements_;
  std::vector<double> vertices_, values_;
};

struct ChemistryModule { bool background = false; std::vector<int> species; bool sodium_free = false; };

// A cloud model of forward_model.config (select_cloud_model.h:44-99):
// "grey", "KHnongrey <lambda_ref>" or "power_law <lambda_ref>" (micron),
// each optionally followed by "fb" (the bottom one scale height below the
// top instead of a fitted bottom pressure).  Parameters: grey [tau, p_top,
// p_bottom / p_top], KHnongrey [tau, q0, a0, particle size (micron), p_top,
// p_bottom / p_top], power law [tau, exponent, p_top, p_bottom / p_top],
// without the last with "fb" (grey_cloud_model.cpp:40-52,
// kh_cloud_model.cpp:41-62, power_law_cloud_model.cpp:39-60).
struct CloudModelSpec {
  int kind = 0;                     // CloudKind: 0 grey, 1 KHnongrey, 2 power law
  double reference_wavelength = 0;
  bool fixed_bottom = false;
  size_t nbParameters() const {
    const size_t n = kind == 0 ? 3 : kind == 1 ? 6 : 4;
    return fixed_bottom ? n - 1 : n;
  }
};
// one model per line; throws InvalidInput for an unknown model
std::vector<CloudModelSpec> parseCloudModels(const std::string& text);

// Variants of the altitude computation (atmosphere.cpp:63-241): radius-
// dependent gravity, a fixed mean molecular weight, or a fixed scale height.
struct AltitudeOptions {
  double bottom_radius = 1.0;
  bool variable_gravity = false;
  double fixed_mean_molecular_weight = 0;   // > 0: replaces the chemistry's
  double scale_height_km = 0;               // > 0: z = -H ln(p / p0)
};

// Atmosphere (forward_model/atmosphere/atmosphere.cpp) with the temperature
// profiles "poly" and "const" and the chemistry modules "iso" and "bg H2He".
class Atmosphere {
 public:
  Atmosphere(size_t nb_levels, double bottom_pressure, double top_pressure,
             const std::string& temperature_desc, const std::string& chemistry_desc);
  size_t nb_levels;
  std::vector<double> pressure, temperature, altitude, mean_molecular_weight;
  std::vector<double> number_density;       // [L][nbSpeciesIds()]
  size_t nbTemperatureParameters() const { return nb_temperature_param_; }
  size_t nbChemistryParameters() const { return nb_chemistry_param_; }
  // Atmosphere::calcAtmosphereStructure (atmosphere.cpp:63-122); returns "neglect"
  bool compute(double surface_gravity, const double* temperature_param, const double* chemistry_param,
               const AltitudeOptions& options = AltitudeOptions());
  // Atmosphere::setAtmosphericStructure (atmosphere.cpp:245-283) with
  // FixedChemistry (fixed_chemistry.cpp:71-89): the caller's pressure (bar)
  // and temperature per level, n_tot = p 1e6 / (k_B T), n_s = n_tot x_s for
  // the given species ids (mixing_ratios [levels][species]), the mean
  // molecular weight over all species, and the hydrostatic altitude (with
  // gravity falling off as 1/r^2 from bottom_radius when asked).  No
  // temperature floor, no background gas, no neglect.
  void setStructure(double surface_gravity, double bottom_radius, bool variable_gravity, const double* pressure,
                    const double* temperature, const std::vector<int>& species, const double* mixing_ratios);
  double nd(size_t level, int id) const { return number_density[level * n_ids_ + id]; }
  // Atmosphere::calcScaleHeight (atmosphere.cpp:353-360) of the last
  // compute(): R T / (mu g) with the surface gravity, or the fixed scale
  // height (km) of that variant (atmosphere.cpp:232)
  double scaleHeight(size_t level, double surface_gravity, double scale_height_km) const {
    if (scale_height_km > 0) return scale_height_km * 100000.0;
    return phys::gas_constant * temperature[level] / (mean_molecular_weight[level] * surface_gravity);
  }
  // Reverse-mode derivative of the last compute(), made with `options`:
  // from dL/dT[l] (direct effects), dL/dn[l][id] and dL/dz[l] (consumed) to
  // dL/dg and dL/d(parameters), adding to *radius_bar (variable gravity: the
  // bottom radius) and *option_bar (the fixed mean molecular weight, or the
  // scale height in km) when given.
  void adjoint(double surface_gravity, const double* temperature_param, const double* chemistry_param,
               std::vector<double>& temperature_bar, std::vector<double>& number_density_bar,
               std::vector<double>& altitude_bar, double& gravity_bar, double* temperature_param_bar,
               double* chemistry_param_bar, const AltitudeOptions& options = AltitudeOptions(),
               double* radius_bar = nullptr, double* option_bar = nullptr) const;
  // copies are independent (each owns its temperature profile's state)
  Atmosphere(const Atmosphere& other);
  // the profile and chemistry as plain data (device staging)
  bool isothermal() const { return isothermal_; }
  const std::vector<PiecewisePolynomial::Basis>& temperatureBasis() const { return temperature_basis_; }
  const std::vector<ChemistryModule>& chemistryModules() const { return chemistry_; }
  const std::vector<int>& activeIds() const { return active_ids_; }
  size_t nbIds() const { return n_ids_; }
 private:
  size_t n_ids_;
  std::unique_ptr<PiecewisePolynomial> poly_;
  bool isothermal_ = false;
  size_t nb_temperature_param_ = 0, nb_chemistry_param_ = 0;
  std::vector<ChemistryModule> chemistry_;
  // ascending species ids (from 1) that a chemistry module can make non-zero:
  // the sums over species skip the others, whose terms are exact zeros
  std::vector<int> active_ids_;
  std::vector<PiecewisePolynomial::Basis> temperature_basis_;   // per level, for the fixed pressures
  std::vector<char> clamped_;               // levels floored at 50 K by the last compute()
  bool bg_clamped_ = false;
  // z from the mean molecular weight, temperature and pressure (atmosphere.cpp:288-350)
  void altitudes(double surface_gravity, const AltitudeOptions& options);
};

}  // namespace bear_amd
