// CPU oracle (test infrastructure only).
//
// An FP64 C++ restatement of the reference's CPU path for one likelihood
// evaluation of the emission model, from the retrieval folder on disk to the
// log-likelihood.  It follows the reference's arithmetic step by step so that
// it agrees with the reference build to rounding, and it is the checker for
// the HIP engine.  Only the tests, the smoke check and the CPU-baseline leg
// of the benchmark load it; the product never does.
//
// Each block cites the reference file:line it restates (paths relative to the
// reference's src/ directory).  Pinning (see tests/test_oracle_*.py):
//   * every function below is compared with the reference build (oracle/ref.mk)
//     on the seeded cases, intermediate arrays included;
//   * closed forms: isothermal emission = pi B scale 1e-3, trapezoid exactness,
//     bilinear exactness of affine tables, the Gaussian log-likelihood.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

// physical constants in cgs (additional/physical_const.h:28-52)
constexpr double kPlanckH = 6.62606896e-27;
constexpr double kLightC = 2.99792458e10;
constexpr double kBoltzmannK = 1.3806504e-16;
constexpr double kGasConstant = 8.314472e7;
constexpr double kRadiusJupiter = 7.1492e9;
constexpr double kPi = 3.14159265358979323846;
constexpr double kAvogadro = 6.022140857e23;   // sampled_data.cpp:70

// species ids and molecular weights, in the reference's enum order
// (chemistry/chem_species.h:30-90); the ids index the number-density rows.
struct SpeciesInfo { const char* symbol; double weight; };
const SpeciesInfo kSpecies[] = {
  {"Total", 0.0}, {"H", 1.00784}, {"He", 4.002602}, {"C", 12.0107}, {"O", 15.999},
  {"Fe", 55.845}, {"Fe+", 55.845}, {"Ca", 40.078}, {"Ti", 47.867}, {"Ti+", 47.867},
  {"H2", 2.01588}, {"H2O", 18.01528}, {"CO2", 44.01}, {"CO", 28.0101}, {"CH4", 16.04246},
  {"HCN", 27.0253}, {"NH3", 17.03052}, {"C2H2", 26.04}, {"N2", 28.0134}, {"Na", 22.98977},
  {"K", 39.0983}, {"H2S", 34.09099}, {"H-", 1.00784}, {"TiO", 63.8664}, {"VO", 66.9409},
  {"FeH", 56.853}, {"SH", 34.08}, {"MgO", 40.3044}, {"AlO", 42.981}, {"CaO", 56.0774},
  {"CrH", 54.0040}, {"MgH", 26.3209}, {"CaH", 41.0859}, {"TiH", 48.87484}, {"OH", 17.008},
  {"e-", 5.4857990907e-4}, {"V", 50.9415}, {"V+", 50.9415}, {"Mn", 54.938044}, {"Si", 28.085},
  {"Cr", 51.996}, {"Cr+", 51.996}, {"SiO", 44.08}, {"SiO2", 60.08}, {"SO2", 64.066}, {"CS2", 76.139}};
constexpr int kNbSpecies = sizeof(kSpecies) / sizeof(kSpecies[0]);
constexpr int kTotal = 0, kH = 1, kHe = 2, kH2 = 10, kNa = 19, kK = 20;

int speciesId(const std::string& s) {
  for (int i = 0; i < kNbSpecies; ++i) if (s == kSpecies[i].symbol) return i;
  throw std::runtime_error("unknown species " + s);
}

// FixedChemistry::findSpecies (fixed_chemistry.cpp:42-63): the symbol or FastChem's symbol
const char* const kFastChemSymbol[] = {
  "Total", "H", "He", "C", "O", "Fe", "Fe+", "Ca", "Ti", "Ti+", "H2", "H2O1", "C1O2", "C1O1", "C1H4", "C1H1N1_1",
  "H3N1", "C2H2", "N2", "Na", "K", "H2S1", "H1-", "O1Ti1", "O1V1", "H1Fe1", "H1S1", "Mg1O1", "Al1O1", "Ca1O1",
  "Cr1H1", "H1Mg1", "Ca1H1", "H1Ti1", "H1O1", "e-", "V", "V1+", "Mn", "Si", "Cr", "Cr1+", "O1Si1", "O2Si1", "O2S1",
  "C1S2"};
int fixedChemistryId(const std::string& s) {
  for (int i = 0; i < kNbSpecies; ++i) if (s == kSpecies[i].symbol || s == kFastChemSymbol[i]) return i;
  throw std::runtime_error("Chemical species " + s + " not found in the list of species");
}

double lerp(double x1, double x2, double y1, double y2, double x) {
  return y1 + (y2 - y1) * (x - x1) / (x2 - x1);
}

// trapezoid rules (additional/quadrature.h:32-73)
double trapz(const std::vector<double>& x, const std::vector<double>& y) {
  if (x.size() == 1) return y[0];
  double sum = 0.0;
  for (size_t i = 1; i < x.size(); ++i) sum += (x[i] - x[i - 1]) * (y[i] + y[i - 1]);
  return sum * 0.5;
}
double trapzRange(const std::vector<double>& x, const std::vector<double>& y, size_t a, size_t b) {
  if (x.size() == 1) return y[0];
  double sum = 0.0;
  for (size_t i = a + 1; i < b + 1; ++i) sum += (x[i] - x[i - 1]) * (y[i] + y[i - 1]);
  return sum * 0.5;
}

// Planck function per wavenumber (additional/aux_functions.cpp:272-278)
double planck(double t, double nu) {
  return 2. * kPlanckH * kLightC * kLightC * nu * nu * nu
         / (std::exp(kPlanckH * nu * kLightC / kBoltzmannK / t) - 1.0);
}

// Gaussian (additional/aux_functions.cpp:344-351)
double normalDistribution(double sigma, double x) {
  const double c = std::sqrt(1.0 / (2.0 * kPi));
  return c / sigma * std::exp(-x * x / (2.0 * sigma * sigma));
}

// ---------------------------------------------------------------- grid
// spectral_grid/spectral_grid.cpp
struct Grid {
  std::vector<double> nu_full, lam_full, nu, lam;
  std::vector<size_t> index;
  int discretisation = 0;
  double resolution = 0;

  void load(const std::string& file) {                           // :87-119, :70-84
    std::ifstream f(file);
    if (!f) {
      nu_full.assign(size_t(90000.0 / 0.01 + 1), 0.0);
      for (size_t i = 0; i < nu_full.size(); ++i) nu_full[i] = i * 0.01;
    } else {
      size_t n;
      f >> n;
      nu_full.resize(n);
      for (auto& v : nu_full) f >> v;
    }
    lam_full.resize(nu_full.size());
    for (size_t i = 0; i < nu_full.size(); ++i) lam_full[i] = 1.0 / nu_full[i] * 1e4;   // convert.cpp:38-52
  }

  // :323-385 nearest-index searches on ascending or descending data
  static size_t closest(double x, const std::vector<double>& d, size_t start) {
    if (d.size() < 2) return 0;
    const bool asc = d[0] < d[1];
    if (!asc && !(d[0] > d[1])) return 0;
    auto it = asc ? std::lower_bound(d.begin() + start, d.end(), x)
                  : std::lower_bound(d.begin() + start, d.end(), x, std::greater<double>());
    if (it == d.begin() + start) return start;
    const double a = *(it - 1), b = *it;
    if (std::fabs(x - a) < std::fabs(x - b)) return (it - d.begin()) - 1;
    return it - d.begin();
  }

  // :388-412
  std::vector<std::vector<size_t>> binEdges(const std::vector<std::vector<double>>& edges) const {
    std::vector<std::vector<size_t>> out(edges.size(), std::vector<size_t>{0, 0});
    for (size_t i = 0; i < edges.size(); ++i) {
      auto l = std::lower_bound(nu_full.begin(), nu_full.end(), edges[i][0]);
      auto r = std::lower_bound(nu_full.begin(), nu_full.end(), edges[i][1]);
      // beyond the list the reference indexes out of bounds; refuse instead
      if (l == nu_full.end() || r == nu_full.end() || (l == nu_full.begin() && *l > edges[i][0]))
        throw std::runtime_error("observation range outside the wavenumber list");
      out[i][0] = l - nu_full.begin();
      if (nu_full[out[i][0]] > edges[i][0]) out[i][0] -= 1;
      out[i][1] = r - nu_full.begin();
    }
    return out;
  }
};

// ---------------------------------------------------------------- observations
struct Observation {
  std::string name, type;
  std::vector<double> data, error, weight, fwhm;
  std::vector<std::vector<double>> edges_lam, edges_nu, edge_idx_v;
  std::vector<double> centres;
  bool ascending = true;
  double lam_range[2] = {0, 0};
  std::string filter_path, detector;
  std::vector<double> filter_lam, filter_resp;
  int nb_modifier = 0;
  // set-up products on the sampled grid
  std::vector<std::vector<size_t>> edge_idx;
  size_t idx_first = 0, idx_second = 0;
  std::vector<double> sigma;                     // instrument profile sigma on the grid
  std::vector<std::vector<size_t>> conv;         // convolution intervals
  std::vector<double> filter, filter_weight;
  double filter_norm = 0;
};

// observations.cpp:127-184
bool reverseIfAscending(Observation& o, std::vector<double>& w) {
  if (w.size() > 1) {
    if (w[0] < w[1]) {
      std::reverse(w.begin(), w.end());
      std::reverse(o.data.begin(), o.data.end());
      std::reverse(o.error.begin(), o.error.end());
      std::reverse(o.fwhm.begin(), o.fwhm.end());
      std::reverse(o.weight.begin(), o.weight.end());
      return true;
    }
    return false;
  }
  return false;
}
bool reverseIfAscending(Observation& o, std::vector<std::vector<double>>& e) {
  bool asc = false;
  if (e.size() > 1 && e.front()[0] < e.back()[0]) {
    asc = true;
    std::reverse(e.begin(), e.end());
    std::reverse(o.data.begin(), o.data.end());
    std::reverse(o.error.begin(), o.error.end());
    std::reverse(o.fwhm.begin(), o.fwhm.end());
    std::reverse(o.weight.begin(), o.weight.end());
  }
  for (auto& b : e) if (b[0] < b[1]) std::reverse(b.begin(), b.end());
  return asc;
}

// observations.cpp:188-221, 224-232, 236-251
std::vector<std::vector<double>> binEdgesFromCentres(const std::vector<double>& w) {
  std::vector<std::vector<double>> e(w.size(), std::vector<double>(2, 0));
  for (size_t i = 0; i < w.size(); ++i) {
    if (i == 0) {
      e.front()[1] = w[0] - (w[i] - w[i + 1]) * 0.5;
      e.front()[0] = w[0] + (w[i] - w[i + 1]) * 0.5;
    } else if (i == w.size() - 1) {
      e.back()[1] = w[i] - (w[i - 1] - w[i]) * 0.5;
      e.back()[0] = w[i] + (w[i - 1] - w[i]) * 0.5;
    } else {
      e[i][0] = w[i - 1] - (w[i - 1] - w[i]) * 0.5;
      e[i][1] = w[i] - (w[i] - w[i + 1]) * 0.5;
    }
  }
  return e;
}
std::vector<double> centresFromEdges(const std::vector<std::vector<double>>& e) {
  std::vector<double> c(e.size());
  for (size_t i = 0; i < e.size(); ++i) c[i] = e[i][0] - (e[i][0] - e[i][1]) * 0.5;
  return c;
}
void observationEdges(Observation& o, const std::vector<std::vector<double>>& e) {
  if (!o.fwhm.empty()) {
    o.lam_range[0] = e.front()[0] + 5.0 * o.fwhm.front() / 2.355;
    o.lam_range[1] = e.back()[1] - 5.0 * o.fwhm.back() / 2.355;
  } else {
    o.lam_range[0] = e.front()[0];
    o.lam_range[1] = e.back()[1];
  }
}

// bandpass file (observations_filter_response.cpp:42-117)
void readFilter(Observation& o, const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("filter file not found: " + path);
  std::string line;
  for (int i = 0; i < 4; ++i) std::getline(f, line);
  std::stringstream ls(line);
  std::string det;
  ls >> det;
  if (det == "Energy") det = "energy";
  if (det == "Photon") det = "photon";
  if (det != "energy" && det != "photon") throw std::runtime_error("unsupported detector type " + det);
  o.detector = det;
  while (std::getline(f, line)) {
    std::stringstream s(line);
    double a, b;
    if (!(s >> a >> b)) continue;
    o.filter_lam.push_back(a);
    o.filter_resp.push_back(b);
  }
  if (o.filter_lam.front() < o.filter_lam.back()) {
    std::reverse(o.filter_lam.begin(), o.filter_lam.end());
    std::reverse(o.filter_resp.begin(), o.filter_resp.end());
  }
}

// observation files (observations_load_data.cpp:40-330) and modifiers (observations.cpp:367-392)
Observation loadObservation(const std::string& file, const std::string& modifier, const std::string& folder) {
  Observation o;
  if (modifier == "none" || modifier.empty()) o.nb_modifier = 0;
  else if (modifier == "shift_const") o.nb_modifier = 1;
  else throw std::runtime_error("unknown observation modifier " + modifier);
  std::ifstream f(file);
  if (!f) throw std::runtime_error("observation file not found: " + file);
  std::string line;
  std::getline(f, line);
  std::getline(f, line);
  std::getline(f, o.name);
  std::getline(f, line);
  std::getline(f, line);
  f >> o.type;
  if (o.type != "photometry" && o.type != "spectroscopy" && o.type != "band-spectroscopy")
    throw std::runtime_error("unknown band type " + o.type);
  for (int i = 0; i < 3; ++i) std::getline(f, line);
  f >> o.filter_path;
  if (o.filter_path == "None" || o.filter_path == "none") o.filter_path.clear();
  else o.filter_path = folder + o.filter_path;
  for (int i = 0; i < 3; ++i) std::getline(f, line);

  auto countColumns = [](const std::string& l) {
    std::stringstream ss(l);
    std::string item;
    unsigned n = 0;
    while (ss >> item) ++n;
    return n;
  };

  if (o.type == "photometry") {
    std::getline(f, line);
    const unsigned cols = countColumns(line);
    if (cols < 4) throw std::runtime_error("photometry needs four columns");
    std::istringstream in(line);
    double lo, hi, v, e, w;
    in >> lo >> hi >> v >> e >> w;
    std::vector<std::vector<double>> edges(1, {lo, hi});
    o.data.push_back(v);
    o.error.push_back(e);
    o.weight.push_back(cols == 5 ? w : 1.0);
    o.ascending = reverseIfAscending(o, edges);
    o.centres = centresFromEdges(edges);
    observationEdges(o, edges);
    o.edges_lam = edges;
  } else if (o.type == "band-spectroscopy") {
    unsigned cols = 0;
    std::vector<std::vector<double>> edges;
    while (std::getline(f, line)) {
      if (cols == 0) {
        cols = countColumns(line);
        if (cols < 4) throw std::runtime_error("band spectroscopy needs four columns");
      }
      std::istringstream in(line);
      double l = 0, r = 0, v = 0, e = 0, fw = 0, w = 0;
      in >> l >> r >> v >> e >> fw >> w;
      if (l != 0.0 && e != 0.0) {
        edges.push_back({l, r});
        o.data.push_back(v);
        o.error.push_back(e);
      }
      if (fw != 0.0 && cols > 4) o.fwhm.push_back(fw);
      o.weight.push_back(cols == 5 ? w : 1.0);
    }
    o.ascending = reverseIfAscending(o, edges);
    o.centres = centresFromEdges(edges);
    observationEdges(o, edges);
    o.edges_lam = edges;
  } else {
    unsigned cols = 0;
    std::vector<double> wl;
    while (std::getline(f, line)) {
      if (cols == 0) {
        cols = countColumns(line);
        if (cols < 3) throw std::runtime_error("spectroscopy needs three columns");
      }
      std::istringstream in(line);
      double l = 0, v = 0, e = 0, fw = 0, w = 0;
      in >> l >> v >> e >> fw >> w;
      if (l != 0.0 && e != 0.0) {
        wl.push_back(l);
        o.data.push_back(v);
        o.error.push_back(e);
      }
      if (fw != 0.0 && cols > 3) o.fwhm.push_back(fw);
      o.weight.push_back(cols == 5 ? w : 1.0);
    }
    o.ascending = reverseIfAscending(o, wl);
    auto edges = binEdgesFromCentres(wl);
    observationEdges(o, edges);
    o.edges_lam = edges;
    o.centres = wl;
  }
  if (!o.filter_path.empty()) readFilter(o, o.filter_path);
  // SpectralBands::init (spectral_band.cpp:38-63)
  o.edges_nu = o.edges_lam;
  for (auto& w : o.edges_nu) { w[0] = 1.0 / w[0] * 1e4; w[1] = 1.0 / w[1] * 1e4; }
  return o;
}

std::vector<int> markRanges(const Grid& g, const std::vector<std::vector<size_t>>& edge_indices);
void takeIndices(Grid& g, const std::vector<int>& inc);

// SpectralGrid::sampleSpectralGrid + createHighResGrid* (spectral_grid.cpp:123-304)
void sampleGrid(Grid& g, std::vector<Observation>& obs) {
  std::vector<std::vector<double>> oe;
  for (auto& o : obs) oe.push_back({1.0 / o.lam_range[0] * 1e4, 1.0 / o.lam_range[1] * 1e4});
  std::sort(oe.begin(), oe.end(), [](std::vector<double> a, std::vector<double> b) { return a[0] < b[0]; });
  std::vector<std::vector<double>> we{{oe[0][0], oe[0][1]}};
  for (size_t i = 1; i < oe.size(); ++i) {
    if (oe[i][0] > we.back()[0] && oe[i][1] < we.back()[1]) continue;
    if (oe[i][0] < we.back()[1] && oe[i][1] > we.back()[1]) we.back()[1] = oe[i][1];
    else we.push_back({oe[i][0], oe[i][1]});
  }
  std::vector<int> inc = markRanges(g, g.binEdges(we));
  for (auto& o : obs) {                                 // band edges are always included
    size_t start = 0;
    double last_nu = 0;
    for (auto& b : o.edges_nu) {
      if (last_nu > b[0]) start = 0;
      const size_t i1 = Grid::closest(b[0], g.nu_full, start);
      const size_t i2 = Grid::closest(b[1], g.nu_full, start);
      inc[i1] = 1;
      inc[i2] = 1;
      last_nu = b[1];
      start = i2;
    }
  }
  takeIndices(g, inc);

  // SpectralBands::setBandEdgeIndices + init (spectral_band.cpp:66-121)
  for (auto& o : obs) {
    o.edge_idx.clear();
    size_t start = 0;
    double last_nu = 0;
    for (auto& b : o.edges_nu) {
      if (last_nu > b[0]) start = 0;
      const size_t i1 = Grid::closest(b[0], g.nu, start);
      const size_t i2 = Grid::closest(b[1], g.nu, start);
      if (i1 > g.nu.size() - 1 || i2 > g.nu.size() - 1) throw std::runtime_error("edge index not found");
      if (i1 == i2) throw std::runtime_error("spectral grid too coarse to resolve a bin");
      o.edge_idx.push_back({i1, i2});
      last_nu = b[1];
      start = i2;
    }
    o.idx_first = Grid::closest(1.0 / o.lam_range[0] * 1e4, g.nu, 0);
    o.idx_second = Grid::closest(1.0 / o.lam_range[1] * 1e4, g.nu, 0);
  }
}

// SpectralGrid(config, wavelength_min, wavelength_max) (spectral_grid.cpp:52-72)
void sampleRangeGrid(Grid& g, double wavelength_min, double wavelength_max) {
  const std::vector<std::vector<double>> edges{{1.0 / wavelength_max * 1e4, 1.0 / wavelength_min * 1e4}};
  takeIndices(g, markRanges(g, g.binEdges(edges)));
}

// createHighResGridConst* (spectral_grid.cpp:158-236) over bin-edge index pairs
std::vector<int> markRanges(const Grid& g, const std::vector<std::vector<size_t>>& edge_indices) {
  std::vector<int> inc(g.nu_full.size(), 0);
  for (auto& e : edge_indices) {
    inc[e[0]] = 1;
    size_t last = e[0];
    for (size_t j = e[0]; j < e[1]; ++j) {
      bool take = false;
      if (g.discretisation == 0) {
        const double next = g.nu_full[last] + g.resolution;
        take = next == g.nu_full[j] || g.nu_full[j + 1] > next;
      } else if (g.discretisation == 1) {
        const double next = g.lam_full[last] - g.resolution;
        take = next == g.lam_full[j] || g.lam_full[j + 1] < next;
      } else {
        const double next = g.lam_full[last] * (1 - 1. / g.resolution);
        take = next == g.lam_full[j] || g.lam_full[j + 1] < next;
      }
      if (take) { inc[j] = 1; last = j; }
      inc[e[1]] = 1;
    }
  }
  return inc;
}

void takeIndices(Grid& g, const std::vector<int>& inc) {
  g.index.clear();
  for (size_t i = 0; i < inc.size(); ++i) if (inc[i] == 1) g.index.push_back(i);
  g.nu.resize(g.index.size());
  g.lam.resize(g.index.size());
  for (size_t i = 0; i < g.index.size(); ++i) { g.nu[i] = g.nu_full[g.index[i]]; g.lam[i] = g.lam_full[g.index[i]]; }
}

// SpectralGrid::interpolateToWavelengthGrid (spectral_grid_interpolate.cpp:60-172)
std::vector<double> interpolateTo(std::vector<double> x, std::vector<double> y, const std::vector<double>& nx) {
  if (x[0] < x[1]) {
    std::reverse(x.begin(), x.end());
    std::reverse(y.begin(), y.end());
  }
  std::vector<double> out(nx.size(), 0.0);
  for (size_t i = 0; i < nx.size(); ++i) {
    if (nx[i] > x.front()) continue;
    if (nx[i] < x.back()) break;
    size_t idx = 0;
    while (idx < x.size() && !(x[idx] <= nx[i])) ++idx;
    if (x[idx] == nx[i]) out[i] = y[idx];
    else out[i] = lerp(x[idx - 1], x[idx], y[idx - 1], y[idx], nx[i]);
  }
  return out;
}

// instrument profile, convolution windows, filters
// (spectral_band.cpp:127-154, spectral_band_convolve.cpp:44-95,
//  observations_filter_response.cpp:122-160)
void setupObservation(const Grid& g, Observation& o) {
  if (!o.fwhm.empty() && o.fwhm.size() == o.centres.size()) {
    std::vector<double> hr(g.lam.begin() + o.idx_first, g.lam.begin() + o.idx_second + 1);
    std::vector<double> fw = interpolateTo(o.centres, o.fwhm, hr);
    o.sigma.assign(g.lam.size(), 0.0);
    for (size_t i = 0; i < hr.size(); ++i) o.sigma[i + o.idx_first] = fw[i] / 2.355;
    const size_t n = o.idx_second - o.idx_first + 1;
    o.conv.assign(n, std::vector<size_t>(2, 0));
    for (size_t i = 0; i < n; ++i) {
      const size_t k = i + o.idx_first;
      o.conv[i][0] = o.idx_first;
      o.conv[i][1] = o.idx_second;
      const double cut = 5.0 * o.sigma[k];
      if (cut == 0) {
        o.conv[i][0] = k;
        o.conv[i][1] = k;
      } else {
        o.conv[i][0] = Grid::closest(g.lam[k] + cut, g.lam, 0);
        o.conv[i][1] = Grid::closest(g.lam[k] - cut, g.lam, o.conv[i][0]);
      }
    }
  }
  if (!o.filter_lam.empty()) {
    o.filter = interpolateTo(o.filter_lam, o.filter_resp, g.lam);
    o.filter_weight.assign(o.filter.size(), 1.0);
    if (o.detector == "photon")
      for (size_t i = 0; i < g.lam.size(); ++i) o.filter_weight[i] = g.lam[i];
    std::vector<double> y = o.filter_resp;
    if (o.detector == "photon") for (size_t i = 0; i < y.size(); ++i) y[i] *= o.filter_lam[i];
    double norm = trapz(o.filter_lam, y);
    if (norm < 0) norm = std::fabs(norm);
    o.filter_norm = norm;
  }
}

// Observation::processModelSpectrum (observations.cpp:290-304) with its three
// stages: filter (observations_filter_response.cpp:166-181), convolution
// (spectral_band_convolve.cpp:100-144) and band integration
// (spectral_band_integrate.cpp:39-107)
std::vector<double> processSpectrum(const Grid& g, const Observation& o, const std::vector<double>& s, bool is_flux) {
  std::vector<double> sf = s;
  if (!o.filter.empty())
    for (size_t i = 0; i < g.nu.size(); ++i) sf[i] = s[i] * o.filter[i] * o.filter_weight[i] / o.filter_norm;
  std::vector<double> sc = sf;
  if (!o.sigma.empty()) {
    sc.assign(sf.size(), 0.0);
    #pragma omp parallel for schedule(dynamic, 64)
    for (size_t i = o.idx_first; i < o.idx_second + 1; ++i) {
      const size_t c = i - o.idx_first;
      if (o.conv[c][0] == o.conv[c][1] || o.sigma[i] == 0) { sc[i] = sf[i]; continue; }
      std::vector<double> x(g.lam.begin() + o.conv[c][0], g.lam.begin() + o.conv[c][1]);
      std::vector<double> y(sf.begin() + o.conv[c][0], sf.begin() + o.conv[c][1]);
      for (size_t j = 0; j < y.size(); ++j) y[j] *= normalDistribution(o.sigma[i], std::abs(g.lam[i] - x[j]));
      sc[i] = std::abs(trapz(x, y));
    }
  }
  std::vector<double> bins(o.edge_idx.size(), 0.0);
  for (size_t b = 0; b < bins.size(); ++b) {
    const size_t i0 = o.edge_idx[b][0], i1 = o.edge_idx[b][1];
    if (is_flux) {
      const double l = 1.0 / g.nu[i0] * 1e4, r = 1.0 / g.nu[i1] * 1e4;
      bins[b] = trapzRange(g.nu, sc, i0, i1) / (l - r);
    } else {
      double v = std::abs(trapzRange(g.lam, sc, i0, i1));
      if (o.filter.empty()) v /= (g.lam[i0] - g.lam[i1]);
      bins[b] = v;
    }
  }
  return bins;
}

// ---------------------------------------------------------------- opacity
// Tables held in memory instead of a filelist.dat folder (the oracle's
// counterpart of the engine's bear_opacity_data_register): the row values are
// what the row files would hold, float32 over the full wavenumber list.  The
// caller keeps the arrays alive while oracles built from them exist.
struct MemoryTable {
  double mass = 0;
  bool log_storage = false;
  std::vector<double> temperature, pressure;
  size_t nb_points = 0;
  const float* rows = nullptr;             // [rows][nb_points]
};
std::map<std::string, MemoryTable>& memoryTables() {
  static std::map<std::string, MemoryTable> tables;
  return tables;
}

// one tabulated row after sampling (sampled_data.cpp:48-87, cross_section_file.cpp:33-68);
// logsigma is filled on first use, as the reference loads a row's data the
// first time an evaluation touches it (opacity_species.cpp:349-365)
struct Row {
  double temperature, pressure, log_pressure;
  std::vector<double> logsigma;
  std::string file;                        // row file, or empty
  const float* memory = nullptr;           // in-memory row over the full list
  size_t memory_points = 0;
};

struct OpacitySpecies {
  std::string symbol;
  int index = 0;
  std::vector<int> partners;
  bool rayleigh = false;
  bool hminus = false;                            // the H- continuum (GasHm, no tables)
  std::vector<double> hminus_bf;                  // bound-free cross section per wavenumber
  double mass = 0;
  bool log_storage = false;
  std::vector<Row> rows;
  std::vector<std::vector<Row*>> ordered;        // by temperature, then pressure
  std::vector<double> rayleigh_sigma;

  void loadTables(const std::string& dir, const std::string& folder_key) {  // opacity_species.cpp:66-157
    auto mem = memoryTables().find(folder_key);
    if (mem != memoryTables().end()) {
      const MemoryTable& m = mem->second;
      mass = m.mass;
      log_storage = m.log_storage;
      for (size_t r = 0; r < m.temperature.size(); ++r) {
        rows.push_back(Row{m.temperature[r], m.pressure[r], std::log10(m.pressure[r]), {}, "", nullptr, 0});
        rows.back().memory = m.rows + r * m.nb_points;
        rows.back().memory_points = m.nb_points;
      }
    } else {
      std::ifstream f(dir + "/filelist.dat");
      if (!f) throw std::runtime_error("no opacity data for " + symbol);
      std::string line, storage;
      std::getline(f, line);
      std::istringstream head(line);
      head >> mass >> storage;
      log_storage = storage == "log";
      while (std::getline(f, line)) {
        std::istringstream in(line);
        double t, p;
        std::string name;
        in >> p >> t >> name;
        if (name.empty()) continue;
        rows.push_back(Row{t, p, std::log10(p), {}, dir + "/" + name, nullptr, 0});
      }
    }
    for (auto& r : rows) {
      auto it = std::find_if(ordered.begin(), ordered.end(),
                             [&](const std::vector<Row*>& a) { return a[0]->temperature == r.temperature; });
      if (it == ordered.end()) ordered.push_back({&r});
      else it->push_back(&r);
    }
    std::sort(ordered.begin(), ordered.end(),
              [](const std::vector<Row*>& a, const std::vector<Row*>& b) { return a[0]->temperature < b[0]->temperature; });
    for (auto& t : ordered)
      std::sort(t.begin(), t.end(), [](Row* a, Row* b) { return a->pressure < b->pressure; });
  }

  // sample one row at the grid's indices (sampled_data.cpp:48-87): float32
  // value -> double, 10^x for log storage, x mass / N_A, 1e-200 floor, log10
  void ensureRow(Row& row, const std::vector<size_t>& index) const {
    if (!row.logsigma.empty() || index.empty()) return;
    std::vector<float> raw;
    const float* data = row.memory;
    size_t n = row.memory_points;
    if (!data) {
      std::ifstream bf(row.file, std::ios::binary | std::ios::ate);
      if (!bf) throw std::runtime_error("missing row file " + row.file);
      n = size_t(bf.tellg()) / sizeof(float);
      bf.seekg(0);
      raw.resize(n);
      bf.read(reinterpret_cast<char*>(raw.data()), n * sizeof(float));
      data = raw.data();
    }
    std::vector<double>& cs = row.logsigma;
    cs.assign(index.size(), 0.0);
    const double factor = mass / kAvogadro;
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < index.size(); ++i) {
      if (index[i] > n - 1) continue;
      double v = data[index[i]];
      if (log_storage) v = std::pow(10.0, v);
      cs[i] = v;
    }
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < cs.size(); ++i) {
      double v = cs[i];
      if (mass > 0) v *= factor;
      if (v < 1e-200) v = 1e-200;
      cs[i] = std::log10(v);
    }
  }

  // opacity_species.cpp:160-273
  void bracket(double p, double t, Row* out[4]) const {
    const std::vector<Row*>* lo = &ordered.front();
    const std::vector<Row*>* hi = &ordered.front();
    if (t < ordered.front().front()->temperature) {
    } else if (t > ordered.back().front()->temperature) {
      lo = hi = &ordered.back();
    } else {
      for (size_t i = 0; i < ordered.size(); ++i) {
        if (ordered[i].front()->temperature == t) { lo = hi = &ordered[i]; break; }
        if (ordered[i].front()->temperature < t && ordered[i + 1].front()->temperature > t) {
          lo = &ordered[i];
          hi = &ordered[i + 1];
          break;
        }
      }
    }
    auto pick = [p](const std::vector<Row*>& l, Row*& a, Row*& b) {
      if (p < l.front()->pressure) { a = b = l.front(); return; }
      if (p > l.back()->pressure) { a = b = l.back(); return; }
      for (size_t i = 0; i < l.size(); ++i) {
        if (l[i]->pressure == p) { a = b = l[i]; return; }
        if (l[i]->pressure < p && l[i + 1]->pressure > p) { a = l[i]; b = l[i + 1]; return; }
      }
    };
    pick(*lo, out[0], out[1]);
    pick(*hi, out[2], out[3]);
  }

  // opacity_species.cpp:277-345
  void crossSections(double p, double t, std::vector<double>& cs) const {
    Row* d[4] = {nullptr, nullptr, nullptr, nullptr};
    bracket(p, t, d);
    const size_t n = cs.size();
    if (d[0] == d[1] && d[0] == d[2] && d[0] == d[3]) {
      for (size_t i = 0; i < n; ++i) cs[i] = std::pow(10.0, d[0]->logsigma[i]);
      return;
    }
    std::vector<double> lower = d[0]->logsigma, upper = d[2]->logsigma;
    const double lp = std::log10(p);
    if (d[0] != d[1])
      for (size_t i = 0; i < n; ++i)
        lower[i] = lerp(d[0]->log_pressure, d[1]->log_pressure, d[0]->logsigma[i], d[1]->logsigma[i], lp);
    if (d[2] != d[3])
      for (size_t i = 0; i < n; ++i)
        upper[i] = lerp(d[2]->log_pressure, d[3]->log_pressure, d[2]->logsigma[i], d[3]->logsigma[i], lp);
    if (d[0] != d[2])
      for (size_t i = 0; i < n; ++i) cs[i] = lerp(d[0]->temperature, d[2]->temperature, lower[i], upper[i], t);
    else
      cs = lower;
    for (size_t i = 0; i < n; ++i) cs[i] = std::pow(10.0, cs[i]);
  }
};

// Rayleigh cross sections (opacity_species.cpp:528-537, species_rayleigh_cross_sections.cpp:22-45)
double rayleighGeneral(double ref_density, double n_refr, double king, double nu) {
  return 24. * std::pow(kPi, 3) * std::pow(nu, 4) / (ref_density * ref_density)
         * std::pow((n_refr * n_refr - 1) / (n_refr * n_refr + 2), 2) * king;
}

// The H- continuum (h_m.cpp:38-137, John 1988 fits): the bound-free
// photo-detachment cross section (cm^2) at wavelength lam (micron) and the
// free-free absorption per H atom and unit electron pressure (cm^4 / dyne)
// at temperature t.
double hminusBoundFree(double lam) {                        // h_m.cpp:56-82
  const double lambda_0 = 1.6419;
  const double c[7] = {0.0, 152.519, 49.534, -118.858, 92.536, -34.194, 4.982};
  if (!(lam <= lambda_0 && lam >= 0.125)) return 0.0;
  double f = 0;
  for (unsigned i = 1; i < 7; ++i) f += c[i] * std::pow(1.0 / lam - 1.0 / lambda_0, (i - 1) / 2.0);
  return 1e-18 * lam * lam * lam * std::pow(1.0 / lam - 1.0 / lambda_0, 1.5) * f;
}

double hminusFreeFree(double lam, double t) {              // h_m.cpp:86-137
  static const double a1[7] = {0.0, 0.0, 2483.3460, -3449.8890, 2200.0400, -696.2710, 88.2830};
  static const double b1[7] = {0.0, 0.0, 285.8270, -1158.3820, 2427.7190, -1841.4000, 444.5170};
  static const double c1[7] = {0.0, 0.0, -2054.2910, 8746.5230, -13651.1050, 8624.9700, -1863.8650};
  static const double d1[7] = {0.0, 0.0, 2827.7760, -11485.6320, 16755.5240, -10051.5300, 2095.2880};
  static const double e1[7] = {0.0, 0.0, -1341.5370, 5303.6090, -7510.4940, 4400.0670, -901.7880};
  static const double f1[7] = {0.0, 0.0, 208.9520, -812.9390, 1132.7380, -655.0200, 132.9850};
  static const double a2[7] = {0.0, 518.1021, 473.2636, -482.2089, 115.5291, 0.0, 0.0};
  static const double b2[7] = {0.0, -734.8666, 1443.4137, -737.1616, 169.6374, 0.0, 0.0};
  static const double c2[7] = {0.0, 1021.1775, -1977.3395, 1096.8827, -245.6490, 0.0, 0.0};
  static const double d2[7] = {0.0, -479.0721, 922.3575, -521.1341, 114.2430, 0.0, 0.0};
  static const double e2[7] = {0.0, 93.1373, -178.9275, 101.7963, -21.9972, 0.0, 0.0};
  static const double f2[7] = {0.0, -6.4285, 12.3600, -7.0571, 1.5097, 0.0, 0.0};
  const double *a, *b, *c, *d, *e, *f;
  if (lam > 0.3645) { a = a1; b = b1; c = c1; d = d1; e = e1; f = f1; }
  else if (lam >= 0.1823) { a = a2; b = b2; c = c2; d = d2; e = e2; f = f2; }
  else return 0.0;
  double x = 0;
  for (unsigned i = 1; i < 7; ++i)
    x += std::pow(5040.0 / t, (i + 1) / 2.0) *
         (lam * lam * a[i] + b[i] + c[i] / lam + d[i] / lam / lam + e[i] / lam / lam / lam +
          f[i] / lam / lam / lam / lam);
  return x * 1e-29;
}

// One tangent path's optical depth (transmission_spectrum.cpp:110-140):
// segment k (level k to k+1 of the path, n segments) adds path[k] (ext[k] +
// ext[k+1]); the sum stops after the first partial sum above 10, or after
// segment `frozen` when that is >= 0.  *last: the last segment summed.
double tangentSum(const double* path, const double* ext, size_t n, int32_t frozen, int32_t* last) {
  double sum = 0;
  *last = int32_t(n) - 1;
  for (size_t k = 0; k < n; ++k) {
    sum += path[k] * (ext[k] + ext[k + 1]);
    if (frozen >= 0 ? int32_t(k) == frozen : sum > 10) { *last = int32_t(k); break; }
  }
  return sum;
}

// ---------------------------------------------------------------- atmosphere
struct PiecewisePoly {                             // additional/piecewise_poly.cpp:49-229
  struct Element { std::vector<double> ref, vert, val; };
  std::vector<Element> el;
  std::vector<double> vert, val;

  PiecewisePoly(size_t nb_el, size_t order, double bottom, double top) {
    static const std::vector<std::vector<double>> nodes = {
      {0.0}, {-1.0, 1.0}, {-1.0, 0.0, 1.0}, {-1.0, -0.447214, 0.447214, 1.0},
      {-1.0, -0.654654, 0.654654, 1.0}, {-1.0, -0.654654, 0.0, 0.654654, 1.0}};
    if (order > 5) throw std::runtime_error("polynomial order above 5 is not supported");
    if (nb_el < 1) nb_el = 1;
    if (order < 1) order = 1;
    const double b0 = std::log10(bottom), b1 = std::log10(top);
    const double size = (b0 - b1) / nb_el;
    std::vector<double> edges(nb_el + 1);
    edges.front() = b0;
    edges.back() = b1;
    for (size_t i = 1; i < nb_el; ++i) edges[i] = edges[i - 1] - size;
    for (size_t i = 0; i < nb_el; ++i) {
      Element e;
      e.ref = nodes[order];
      e.vert.resize(e.ref.size());
      e.val.assign(e.ref.size(), 0.0);
      for (size_t j = 0; j < e.ref.size(); ++j)
        e.vert[j] = edges[i] + (1.0 + e.ref[j]) / 2.0 * (edges[i + 1] - edges[i]);
      el.push_back(e);
    }
    for (auto& e : el) for (size_t j = 0; j + 1 < e.vert.size(); ++j) vert.push_back(e.vert[j]);
    vert.push_back(el.back().vert.back());
    val.assign(vert.size(), 0.0);
  }
  void set(const std::vector<double>& v) {
    val = v;
    size_t g = 0;
    for (auto& e : el) {
      for (size_t j = 0; j < e.val.size(); ++j) e.val[j] = val[g++];
      --g;
    }
  }
  double value(double x) const {
    if (x > vert.front() || x < vert.back()) return 0.0;
    for (size_t i = 0; i < vert.size(); ++i) if (vert[i] == x) return val[i];
    size_t k = 0;
    for (size_t i = 0; i < el.size(); ++i)
      if (el[i].vert.front() > x && el[i].vert.back() < x) { k = i; break; }
    const Element& e = el[k];
    const double r = 2.0 * (x - e.vert.front()) / (e.vert.back() - e.vert.front()) - 1.0;
    double y = 0;
    for (size_t i = 0; i < e.ref.size(); ++i) {
      double l = 1;
      for (size_t j = 0; j < e.ref.size(); ++j) {
        if (i == j) continue;
        l *= (r - e.ref[j]) / (e.ref[i] - e.ref[j]);
      }
      y += e.val[i] * l;
    }
    return y;
  }
};

struct Chemistry { bool background = false; std::vector<int> species; bool sodium_free = false; };

// Cloud layers (cloud_model/grey_cloud_model.cpp, power_law_cloud_model.cpp,
// kh_cloud_model.cpp; select_cloud_model.h:44-99): kind 0 "grey" [tau,
// p_top, p_bottom / p_top], 1 "KHnongrey" [tau, q0, a0, particle size,
// p_top, p_bottom / p_top], 2 "power_law" [tau, exponent, p_top,
// p_bottom / p_top]; "fb" drops the last parameter and puts the bottom one
// scale height below the top.
struct CloudLayer {
  int kind = 0;
  double reference_wavelength = 0;
  bool fixed_bottom = false;
  size_t nbParameters() const {
    const size_t n = kind == 0 ? 3 : kind == 1 ? 6 : 4;
    return fixed_bottom ? n - 1 : n;
  }
};

// ---------------------------------------------------------------- the model
struct EmissionOracle {
  std::string model = "emission";
  bool variable_gravity = false, fit_mmw = false, fit_scale_height = false;
  Grid grid;
  std::vector<Observation> obs;
  size_t nb_levels = 0;
  double p_bottom = 0, p_top = 0;
  std::vector<double> pressure;
  std::unique_ptr<PiecewisePoly> tpoly;
  bool const_temperature = false;
  size_t nb_temp_param = 0;
  std::vector<Chemistry> chemistry;
  size_t nb_chem_param = 0, nb_modifier = 0;
  std::vector<CloudLayer> clouds;
  size_t nb_cloud_param = 0;
  std::vector<OpacitySpecies> species;
  bool error_inflation = false;
  std::vector<double> default_parameters;
  // pyBeAR's direct models: a grid sampled over [direct_range[0], direct_range[1]]
  // micron, no observations; the eclipse's tabulated stellar flux on the grid
  bool direct = false;
  double direct_range[2] = {0, 0};
  std::vector<double> stellar_flux;

  // per-evaluation state
  std::vector<double> temperature, altitude, mmw, scale_height;
  // layer optical depths of the clouds [L-1][N] (empty: no cloud), the sum
  // over the cloud models (opacity_calc.h:195-240), or the direct models'
  // caller-given ones (fixed_cloud_model.cpp:36-75)
  std::vector<std::vector<double>> cloud_tau;
  // transmission: the same as level extinction [L][N], 0 at the bottom,
  // tau[l-1] / (z_l - z_{l-1}) above (cloud_model.h:67-88)
  std::vector<std::vector<double>> cloud_ext;
  std::vector<std::vector<double>> nd;                 // [L][46]
  std::vector<std::vector<double>> kappa;              // [L][N]
  std::vector<std::vector<double>> scattering;         // [L][N]
  std::vector<double> spectrum_flux, spectrum;
  std::vector<std::vector<double>> bins;
  bool neglect = false;
  double log_like = 0;
  // transmission: stopping segments of the last evaluation (when recorded) and
  // stopping segments imposed on the next ones (when set), [N][L-1]
  bool record_stops = false;
  std::vector<int32_t> recorded_stops, frozen_stops;
  double last_bottom_radius = 0;

  // StarSpectrumFile(wavelengths, flux, grid) (star_file_spectrum.cpp:57-72):
  // two columns, micron and W m^-2 micron^-1, to W m^-2 cm on the grid
  void readStellarSpectrum(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("stellar spectrum not found: " + path);
    std::vector<double> lam, flux;
    double a, b;
    while (f >> a >> b) { lam.push_back(a); flux.push_back(b); }
    for (size_t i = 0; i < flux.size(); ++i) flux[i] = flux[i] * lam[i] * lam[i] / 10000.;
    stellar_flux = interpolateTo(lam, flux, grid.lam);
  }

  void load(const std::string& folder_in) {
    std::string folder = folder_in;
    if (folder.back() != '/') folder += '/';
    std::string stellar_file;
    std::ifstream cf(folder + "case.txt");
    if (!cf) throw std::runtime_error("case.txt not found in " + folder);
    std::string line;
    std::vector<std::string> temp_desc;
    std::vector<std::pair<std::string, std::string>> opac;
    while (std::getline(cf, line)) {
      std::istringstream in(line);
      std::string key, tok;
      in >> key;
      std::vector<std::string> r;
      while (in >> tok) r.push_back(tok);
      if (key == "model") {
        model = r[0];
        if (model != "emission" && model != "transmission" && model != "secondary_eclipse")
          throw std::runtime_error("forward model not restated: " + model);
      }
      if (key == "use_variable_gravity") variable_gravity = r[0] == "1";
      if (key == "fit_mean_molecular_weight") fit_mmw = r[0] == "1";
      if (key == "fit_scale_height") fit_scale_height = r[0] == "1";
      if (key == "stellar_model" && r[0] != "blackbody") throw std::runtime_error("stellar model not restated: " + r[0]);
      if (key == "spectral_resolution") grid.resolution = std::stod(r[0]);
      if (key == "spectral_discretisation") grid.discretisation = std::stoi(r[0]);   // global_config.cpp:147-168
      if (key == "nb_levels") nb_levels = std::stoul(r[0]);
      if (key == "p_bottom") p_bottom = std::stod(r[0]);
      if (key == "p_top") p_top = std::stod(r[0]);
      if (key == "temperature") temp_desc = r;
      if (key == "chemistry") {
        Chemistry c;
        if (r[0] == "iso") {
          for (size_t i = 1; i < r.size(); ++i) c.species.push_back(speciesId(r[i]));
          c.sodium_free = std::find(c.species.begin(), c.species.end(), kNa) != c.species.end();
          nb_chem_param += c.species.size();
        } else if (r[0] == "bg" && r[1] == "H2He") {
          c.background = true;
        } else throw std::runtime_error("chemistry model not restated: " + r[0]);
        chemistry.push_back(c);
      }
      if (key == "opacity") opac.push_back({r[0], r[1]});
      if (key == "error_inflation") error_inflation = r[0] == "1";
      if (key == "cloud") {                               // forward_model.config cloud lines
        CloudLayer c;
        const std::vector<std::string> a(r.begin() + 1, r.end());
        if (r[0] == "grey") {
          c.kind = 0;
          c.fixed_bottom = a.size() == 1 && a[0] == "fb";
        } else if (r[0] == "KHnongrey" || r[0] == "power_law") {
          c.kind = r[0] == "KHnongrey" ? 1 : 2;
          if (a.empty()) throw std::runtime_error("cloud model " + r[0] + " needs a reference wavelength");
          c.reference_wavelength = std::stod(a[0]);
          c.fixed_bottom = a.size() == 2 && a[1] == "fb";
        } else throw std::runtime_error("cloud model not restated: " + r[0]);
        clouds.push_back(c);
        nb_cloud_param += c.nbParameters();
      }
      if (key == "parameters")
        for (size_t i = 0; i < std::stoul(r[0]); ++i) default_parameters.push_back(std::stod(r[1 + i]));
      if (key == "direct_range") {
        direct = true;
        direct_range[0] = std::stod(r[0]);
        direct_range[1] = std::stod(r[1]);
      }
      if (key == "stellar_spectrum") stellar_file = folder + r[0];
    }
    if (grid.discretisation < 0 || grid.discretisation > 2) throw std::runtime_error("unknown spectral discretisation");
    grid.load(folder + "wavenumber_full.dat");

    if (direct) {
      sampleRangeGrid(grid, direct_range[0], direct_range[1]);
      if (!stellar_file.empty()) readStellarSpectrum(stellar_file);
    } else {
      std::ifstream lf(folder + "observations.list");
      while (std::getline(lf, line)) {
        std::istringstream in(line);
        std::string a, b;
        in >> a >> b;
        if (a.empty()) continue;
        obs.push_back(loadObservation(folder + a, b, folder));
        nb_modifier += obs.back().nb_modifier;
      }
      sampleGrid(grid, obs);
      for (auto& o : obs) setupObservation(grid, o);
    }

    // atmosphere (atmosphere.cpp:365-383)
    pressure.assign(nb_levels, 0.0);
    pressure.front() = std::log10(p_bottom);
    const double step = (std::log10(p_bottom) - std::log10(p_top)) / (nb_levels - 1.0);
    for (size_t i = 1; i < nb_levels - 1; ++i) pressure[i] = pressure[i - 1] - step;
    for (size_t i = 0; i < nb_levels - 1; ++i) pressure[i] = std::pow(10.0, pressure[i]);
    pressure.back() = p_top;

    if (temp_desc[0] == "poly") {
      tpoly.reset(new PiecewisePoly(std::stoul(temp_desc[1]), std::stoul(temp_desc[2]), p_bottom, p_top));
      nb_temp_param = std::stoul(temp_desc[1]) * std::stoul(temp_desc[2]) + 1;
    } else if (temp_desc[0] == "const") {
      const_temperature = true;
      nb_temp_param = 1;
    } else throw std::runtime_error("temperature profile not restated: " + temp_desc[0]);

    // opacity species (transport_coeff.cpp:98-237)
    for (auto& e : opac) {
      OpacitySpecies s;
      s.symbol = e.first;
      if (e.first == "CIA-H2-H2") { s.index = kH2; s.partners = {kH2}; }
      else if (e.first == "CIA-H2-He") { s.index = kH2; s.partners = {kHe}; }
      else if (e.first == "CIA-H-He") { s.index = kH; s.partners = {kHe}; }
      else if (e.first == "H2" && e.second == "Rayleigh") { s.index = kH2; s.rayleigh = true; }
      else if (e.first == "H-") { s.index = speciesId("H-"); s.hminus = true; }   // transport_coeff.cpp:145-150
      else s.index = speciesId(e.first);
      if (s.hminus) {
        s.hminus_bf.resize(grid.nu.size());
        for (size_t j = 0; j < grid.nu.size(); ++j) s.hminus_bf[j] = hminusBoundFree(grid.lam[j]);
      } else if (s.rayleigh) {
        s.rayleigh_sigma.resize(grid.nu.size());
        for (size_t j = 0; j < grid.nu.size(); ++j) {
          const double n = (13.58e-5 * (1. + 7.52e-3 / std::pow(grid.lam[j], 2))) + 1.;
          s.rayleigh_sigma[j] = rayleighGeneral(2.651629e19, n, 1.0, grid.nu[j]);
        }
      } else if (e.second == "Rayleigh") {
        throw std::runtime_error("Rayleigh species not restated: " + e.first);
      } else {
        s.loadTables(folder + "opacity/" + e.second, e.second);
      }
      species.push_back(std::move(s));
    }
  }

  // general parameters: emission [log g, f, d] (emission.cpp:68-69); transmission
  // [log g, R_p, R_*, (mu or H)] (transmission.cpp:69-72); eclipse [log g, Rp/Rs]
  // then the blackbody star's T_eff (secondary_eclipse.cpp:72, star_blackbody.h)
  size_t nbGeneral() const {
    if (model == "transmission") return 3 + ((fit_mmw || fit_scale_height) ? 1 : 0);
    if (model == "secondary_eclipse") return 3;
    return 3;
  }
  size_t nbParameters() const {
    return nbGeneral() + nb_chem_param + nb_temp_param + nb_cloud_param + nb_modifier + (error_inflation ? 1 : 0);
  }

  // Atmosphere::calcAtmosphereStructure, its three overloads (atmosphere.cpp:63-241):
  // fixed_mmw > 0 replaces the mean molecular weight, scale_height_km > 0
  // replaces the hydrostatic altitude by z = -H ln(p / p0)
  bool atmosphereStructure(double g, const std::vector<double>& tp, const std::vector<double>& cp,
                           double bottom_radius = 1.0, double fixed_mmw = 0.0, double scale_height_km = 0.0) {
    bool neg = false;
    const size_t L = nb_levels;
    // temperature (piecewise_poly_temperature.cpp:49-84, constant_temperature.h:34-46)
    temperature.assign(L, 0.0);
    if (const_temperature) {
      temperature.assign(L, tp[0]);
    } else {
      std::vector<double> dof(nb_temp_param, 0.0);
      dof[0] = tp[0];
      for (size_t i = 1; i < nb_temp_param; ++i) dof[i] = dof[i - 1] * tp[i];
      tpoly->set(dof);
      for (size_t i = 0; i < L; ++i) temperature[i] = tpoly->value(std::log10(pressure[i]));
    }
    // the piecewise polynomial floors at 50 K and flags the model
    // (piecewise_poly_temperature.cpp:77-81); the isothermal profile does not
    if (!const_temperature)
      for (auto& t : temperature) if (t < 50) { t = 50; neg = true; }
    // chemistry (isoprofile_chemistry.cpp:87-119, background_chemistry.cpp:72-137, chemistry.h:47-76)
    nd.assign(L, std::vector<double>(kNbSpecies, 0.0));
    mmw.assign(L, 0.0);
    auto meanWeight = [&]() {
      for (size_t i = 0; i < L; ++i) {
        double mu = 0;
        for (int j = 0; j < kNbSpecies; ++j) mu += nd[i][j] / nd[i][kTotal] * kSpecies[j].weight;
        mmw[i] = mu;
      }
    };
    size_t k = 0;
    for (auto& c : chemistry) {
      if (!c.background) {
        for (size_t i = 0; i < L; ++i) {
          nd[i][kTotal] = pressure[i] * 1.e6 / kBoltzmannK / temperature[i];
          for (size_t j = 0; j < c.species.size(); ++j) nd[i][c.species[j]] = nd[i][kTotal] * cp[k + j];
          if (!c.sodium_free) nd[i][kNa] = nd[i][kK] * 16.2181;
        }
        k += c.species.size();
        meanWeight();
        for (size_t i = 0; i < L; ++i) {
          double sum = 0;
          for (int j = 1; j < kNbSpecies; ++j) sum += nd[i][j] / nd[i][kTotal];
          if (sum > 1 + 1e-10) { neg = true; break; }
        }
      } else {
        const double h2 = 0.5, he = 0.085114, h2he = h2 + he;
        for (size_t i = 0; i < L; ++i) nd[i][kTotal] = pressure[i] * 1.e6 / kBoltzmannK / temperature[i];
        for (size_t i = 0; i < L; ++i) {
          double bg = 1.0;
          for (int j = 1; j < kNbSpecies; ++j)
            if (j != kH2 && j != kHe) bg -= nd[i][j] / nd[i][kTotal];
          if (bg < 0) { neg = true; bg = 0; }
          nd[i][kH2] = nd[i][kTotal] * (bg * h2 / h2he);
          nd[i][kHe] = nd[i][kTotal] * (bg * he / h2he);
        }
        meanWeight();
      }
    }
    altitude.assign(L, 0.0);
    if (scale_height_km > 0) {                                 // atmosphere.cpp:338-350
      const double h = scale_height_km * 100000.0;
      for (size_t i = 1; i < L; ++i) altitude[i] = -h * std::log(pressure[i] / pressure[0]);
      scale_height.assign(L, h);                               // atmosphere.cpp:232
      return neg;
    }
    if (fixed_mmw > 0) mmw.assign(L, fixed_mmw);               // atmosphere.cpp:168
    hydrostaticAltitude(g, bottom_radius, variable_gravity);
    scaleHeight(g);
    return neg;
  }

  // Atmosphere::setAtmosphericStructure with FixedChemistry (atmosphere.cpp:245-283,
  // fixed_chemistry.cpp:71-89, chemistry.h:47-58)
  void setStructure(double g, double bottom_radius, bool var_gravity, const std::vector<double>& p,
                    const std::vector<double>& t, const std::vector<int>& ids, const std::vector<double>& x) {
    const size_t L = nb_levels, S = ids.size();
    pressure = p;
    temperature = t;
    nd.assign(L, std::vector<double>(kNbSpecies, 0.0));
    for (size_t i = 0; i < L; ++i) {
      nd[i][kTotal] = pressure[i] * 1.e6 / kBoltzmannK / temperature[i];
      for (size_t j = 0; j < S; ++j) nd[i][ids[j]] = nd[i][kTotal] * x[i * S + j];
    }
    mmw.assign(L, 0.0);
    for (size_t i = 0; i < L; ++i) {
      double mu = 0;
      for (int j = 0; j < kNbSpecies; ++j) mu += nd[i][j] / nd[i][kTotal] * kSpecies[j].weight;
      mmw[i] = mu;
    }
    altitude.assign(L, 0.0);
    hydrostaticAltitude(g, bottom_radius, var_gravity);
    scaleHeight(g);
  }

  // Atmosphere::calcScaleHeight (atmosphere.cpp:353-360): R T / (mu g) with
  // the surface gravity
  void scaleHeight(double g) {
    scale_height.assign(nb_levels, 0.0);
    for (size_t i = 0; i < nb_levels; ++i) scale_height[i] = kGasConstant * temperature[i] / (mmw[i] * g);
  }

  // GreyCloudModel::cloudPosition, both forms (grey_cloud_model.cpp:118-197):
  // the levels bracketing the cloud top (and bottom) pressure, or the bottom
  // one scale height below the top; at least two layers thick.  Unsigned
  // arithmetic as in the reference.
  void cloudPosition(double top_pressure, double bottom_pressure, bool fixed_bottom, unsigned& top,
                     unsigned& bottom) const {
    const size_t L = nb_levels;
    top = 0;
    bottom = 0;
    for (size_t i = 0; i < L - 1; ++i)
      if ((pressure[i] > top_pressure && pressure[i + 1] < top_pressure) || pressure[i] == top_pressure)
        top = unsigned(i);
    if (!fixed_bottom) {
      for (size_t i = 0; i < L - 1; ++i)
        if ((pressure[i] > bottom_pressure && pressure[i + 1] < bottom_pressure) || pressure[i] == bottom_pressure)
          bottom = unsigned(i);
      if (top_pressure <= pressure.back()) top = unsigned(L - 1);
      if (bottom_pressure >= pressure[0]) bottom = 0;
    } else {
      if (top_pressure < pressure.back()) top = unsigned(L - 1);
      double bottom_altitude = altitude[top] - scale_height[top];
      if (bottom_altitude < 0) bottom_altitude = 0;
      for (size_t i = 0; i < L - 1; ++i)
        if ((altitude[i] < bottom_altitude && altitude[i + 1] > bottom_altitude) || altitude[i] == bottom_altitude)
          bottom = unsigned(i);
    }
    if (top == bottom) bottom -= 2;
    if (top < 2) {
      top = 2;
      bottom = 0;
    }
  }

  // the cloud models' layer optical depths, summed (opacity_calc.h:195-240;
  // grey_cloud_model.cpp:56-98, power_law_cloud_model.cpp:70-121,
  // kh_cloud_model.cpp:69-133): tau / (top - bottom) on the layers
  // bottom+1 .. top, times (lambda / lambda_ref)^exponent for the power law
  // and Q(x_ref) / (q0 x^-a0 + x^0.2), x = 2 pi a / lambda, for KHnongrey.
  // A top at the uppermost level would name the layer above the grid (the
  // reference writes past its array there); that layer is left out.
  void cloudOpticalDepths(const std::vector<double>& cp) {
    const size_t N = grid.nu.size(), L = nb_levels;
    cloud_tau.clear();
    if (clouds.empty()) return;
    cloud_tau.assign(L - 1, std::vector<double>(N, 0.0));
    size_t k = 0;
    for (const CloudLayer& c : clouds) {
      const double* q = cp.data() + k;
      k += c.nbParameters();
      double tau = q[0];
      const size_t at_top = c.kind == 0 ? 1 : c.kind == 1 ? 4 : 2;
      const double top_pressure = q[at_top];
      const double bottom_pressure = c.fixed_bottom ? 0.0 : top_pressure * q[at_top + 1];
      if (tau < 0) tau = 0;
      unsigned top = 0, bottom = 0;
      cloudPosition(top_pressure, bottom_pressure, c.fixed_bottom, top, bottom);
      const double layer = tau / static_cast<double>(top - bottom);
      std::vector<double> d(N, layer);                    // this model's depth per wavenumber
      if (c.kind == 2) {
        const double reference_value = std::pow(c.reference_wavelength, q[1]);
        for (size_t j = 0; j < N; ++j) d[j] = layer * std::pow(grid.lam[j], q[1]) / reference_value;
      } else if (c.kind == 1) {
        const double q0 = q[1], a0 = q[2], size = q[3];
        const double x_ref = 2 * kPi * size / c.reference_wavelength;
        const double q_ref = q0 * std::pow(x_ref, -a0) + std::pow(x_ref, 0.2);
        for (size_t j = 0; j < N; ++j) {
          const double x = 2 * kPi * size / grid.lam[j];
          d[j] = layer * q_ref / (q0 * std::pow(x, -a0) + std::pow(x, 0.2));
        }
      }
      for (unsigned i = top; i > bottom; --i) {
        if (i >= L - 1) continue;
        for (size_t j = 0; j < N; ++j) cloud_tau[i][j] += d[j];
      }
    }
  }

  // hydrostatic altitude from mmw, T and p (atmosphere.cpp:288-334)
  void hydrostaticAltitude(double g, double bottom_radius, bool var_gravity) {
    const size_t L = nb_levels;
    std::vector<double> rho(L);
    for (size_t i = 0; i < L; ++i) rho[i] = mmw[i] * pressure[i] * 1e6 / (kGasConstant * temperature[i]);
    for (size_t i = 1; i < L; ++i) {
      double gl = g;
      if (var_gravity) {
        const double local_radius = bottom_radius + altitude[i - 1];
        gl = g * std::pow(bottom_radius / local_radius, 2.0);
      }
      const double dz = (1.0 / (rho[i] * gl) + 1.0 / (rho[i - 1] * gl)) * 0.5 * (pressure[i - 1] * 1e6 - pressure[i] * 1e6);
      altitude[i] = altitude[i - 1] + dz;
    }
  }

  // the direct models' calcSpectrum: emission (emission.cpp:268-340, a =
  // radius, b = distance), transmission (transmission.cpp:371-465, a = planet
  // radius, b = radius ratio), eclipse (secondary_eclipse.cpp:395-482, a =
  // radius ratio).  cloud: the caller's layer optical depths [L-1][N] or
  // empty; used when any is positive (emission.cpp:344-367,
  // fixed_cloud_model.cpp:36-75), summed onto zeros (opacity_calc.h:221-237)
  void directSpectrum(double g, double a, double b, const std::vector<double>& p, const std::vector<double>& t,
                      const std::vector<int>& ids, const std::vector<double>& x, bool var_gravity,
                      const std::vector<double>& cloud = {}) {
    const size_t N = grid.nu.size(), L = nb_levels;
    cloud_tau.clear();
    if (std::any_of(cloud.begin(), cloud.end(), [](double v) { return v > 0; })) {
      cloud_tau.assign(L - 1, std::vector<double>(N, 0.0));
      for (size_t l = 0; l + 1 < L; ++l)
        for (size_t i = 0; i < N; ++i) cloud_tau[l][i] += cloud[l * N + i];
    }
    if (model == "transmission") {
      setStructure(g, a, var_gravity, p, t, ids, x);
      opacity();
      last_bottom_radius = a;
      transmissionSpectrum(a, a / b);
      spectrum = spectrum_flux;
    } else if (model == "secondary_eclipse") {
      setStructure(g, 1.0, false, p, t, ids, x);
      opacity();
      radiativeTransfer(1.0);
      spectrum.resize(spectrum_flux.size());
      for (size_t i = 0; i < spectrum.size(); ++i) spectrum[i] = spectrum_flux[i] / stellar_flux[i] * a * a * 1e6;
    } else {
      setStructure(g, a, var_gravity, p, t, ids, x);
      opacity();
      radiativeTransfer(1.0);
      spectrum = spectrum_flux;
      for (size_t i = 0; i < spectrum.size(); ++i) spectrum[i] *= a * a / b / b;
    }
  }

  // OpacityCalculation::calculate + TransportCoefficients::calculate
  // (opacity_calc.h:154-182, transport_coeff.cpp:242-260, opacity_species.cpp:400-455)
  void opacity() {
    const size_t N = grid.nu.size(), L = nb_levels;
    // rows the evaluation touches, sampled on first use
    for (size_t l = 0; l < L; ++l)
      for (auto& s : species) {
        if (s.rows.empty()) continue;
        Row* d[4] = {nullptr, nullptr, nullptr, nullptr};
        s.bracket(pressure[l], temperature[l], d);
        for (Row* r : d) s.ensureRow(*r, grid.index);
      }
    kappa.assign(L, std::vector<double>(N, 0.0));
    scattering.assign(L, std::vector<double>(N, 0.0));
    #pragma omp parallel for schedule(dynamic, 1)
    for (size_t l = 0; l < L; ++l) {
      std::vector<double> cs(N);
      for (auto& s : species) {
        double n = nd[l][s.index];
        for (int q : s.partners) n *= nd[l][q];
        if (n == 0) continue;
        if (!s.rows.empty()) {
          s.crossSections(pressure[l], temperature[l], cs);
          for (size_t i = 0; i < N; ++i) kappa[l][i] += cs[i] * n;
        }
        if (s.rayleigh)
          for (size_t i = 0; i < N; ++i) scattering[l][i] += s.rayleigh_sigma[i] * nd[l][s.index];
        if (s.hminus) {                            // h_m.cpp:38-53
          const double pe = nd[l][speciesId("e-")] * kBoltzmannK * temperature[l];
          const double nh = nd[l][kH];
          for (size_t i = 0; i < N; ++i)
            kappa[l][i] += s.hminus_bf[i] * n + hminusFreeFree(grid.lam[i], temperature[l]) * pe * nh;
        }
      }
    }
  }

  // ShortCharacteristics::calcSpectrum (short_characteristics.cpp:39-125)
  void radiativeTransfer(double scaling) {
    const size_t N = grid.nu.size(), L = nb_levels;
    const double mu[2] = {0.211324865405187, 0.788675134594813};
    const double w[2] = {0.5, 0.5};
    spectrum_flux.assign(N, 0.0);
    #pragma omp parallel for schedule(dynamic, 256)
    for (size_t i = 0; i < N; ++i) {
      std::vector<double> tau(L - 1), b(L);
      for (size_t l = 0; l + 1 < L; ++l) tau[l] = (altitude[l + 1] - altitude[l]) * (kappa[l + 1][i] + kappa[l][i]) / 2.;
      if (!cloud_tau.empty())                                   // short_characteristics.cpp:75-76
        for (size_t l = 0; l + 1 < L; ++l) tau[l] += cloud_tau[l][i];
      for (size_t l = 0; l < L; ++l) b[l] = planck(temperature[l], grid.nu[i]);
      double i1 = b[0], i2 = b[0];
      for (size_t l = 0; l + 1 < L; ++l) {
        if (tau[l] == 0) continue;
        const double d1 = tau[l] / mu[0];
        const double a1 = std::exp(-d1);
        i1 = i1 * a1;
        i1 += (1.0 + (a1 - 1.0) / d1) * b[l + 1] + (-a1 - (a1 - 1.0) / d1) * b[l];
        const double d2 = tau[l] / mu[1];
        const double a2 = std::exp(-d2);
        i2 = i2 * a2;
        i2 += (1.0 + (a2 - 1.0) / d2) * b[l + 1] + (-a2 - (a2 - 1.0) / d2) * b[l];
      }
      spectrum_flux[i] = 2.0 * kPi * (i1 * mu[0] * w[0] + i2 * mu[1] * w[1]) * 1e-3;
    }
    for (auto& s : spectrum_flux) s *= scaling;
  }

  // TransmissionModel::calcTransmissionSpectrum .. distanceToTangentCenter
  // (transmission_spectrum.cpp:32-154): per wavenumber, the tangent optical
  // depth of every level (trapezoid along the chord, both hemispheres, stop
  // once tau > 10), the effective tangent height and the transit depth in ppm
  // Tangent optical depths at wavenumber w (transmission_spectrum.cpp:84-140):
  // tau[t] summed outward from level t, segment by segment, stopping after the
  // first partial sum above 10; stop[t] is the last segment summed (L - 2 when
  // the sum never passes 10).  With frozen stops ([N][L-1], test
  // infrastructure for derivatives with the stopping segments held fixed) each
  // path stops at its given segment instead.
  // Segment lengths of every tangent path, path[t][k] = d(t, t+k+1) - d(t, t+k)
  // with d(t, l) = sqrt((R + z_l)^2 - (R + z_t)^2) (transmission_spectrum.cpp:117-118,
  // 145-154); the same for every wavenumber, so computed once per evaluation.
  std::vector<std::vector<double>> pathLengths(double bottom_radius) const {
    const size_t L = nb_levels;
    auto dist = [&](size_t tangent, size_t level) {
      const double a = bottom_radius + altitude[level];
      const double b = bottom_radius + altitude[tangent];
      return std::sqrt(a * a - b * b);
    };
    std::vector<std::vector<double>> path(L - 1);
    for (size_t t = 0; t + 1 < L; ++t) {
      path[t].resize(L - 1 - t);
      for (size_t k = 0; k < path[t].size(); ++k) path[t][k] = dist(t, t + k + 1) - dist(t, t + k);
    }
    return path;
  }

  void tangentDepths(size_t w, const std::vector<std::vector<double>>& path, double* tau, int32_t* stop) const {
    const size_t L = nb_levels;
    const int32_t* frozen = frozen_stops.empty() ? nullptr : frozen_stops.data() + w * (L - 1);
    std::vector<double> ext(L);
    for (size_t i = 0; i < L; ++i) ext[i] = kappa[i][w] + scattering[i][w];
    if (!cloud_ext.empty())                                     // transmission_spectrum.cpp:121-130
      for (size_t i = 0; i < L; ++i) ext[i] += cloud_ext[i][w];
    for (int t = int(L) - 2; t > -1; --t) {
      int32_t last = 0;
      tau[t] = tangentSum(path[size_t(t)].data(), ext.data() + t, path[size_t(t)].size(), frozen ? frozen[t] - t : -1,
                          &last);
      if (stop) stop[t] = last + t;
    }
  }

  void cloudExtinction() {
    const size_t N = grid.nu.size(), L = nb_levels;
    cloud_ext.clear();
    if (cloud_tau.empty()) return;
    cloud_ext.assign(L, std::vector<double>(N, 0.0));
    for (size_t j = 1; j < L; ++j) {
      const double dz = altitude[j] - altitude[j - 1];
      for (size_t i = 0; i < N; ++i) cloud_ext[j][i] = cloud_tau[j - 1][i] / dz;
    }
  }

  void transmissionSpectrum(double bottom_radius, double star_radius) {
    const size_t N = grid.nu.size(), L = nb_levels;
    cloudExtinction();
    spectrum_flux.assign(N, 0.0);
    if (record_stops) recorded_stops.assign(N * (L - 1), 0);
    const std::vector<std::vector<double>> path = pathLengths(bottom_radius);
    #pragma omp parallel for schedule(dynamic, 256)
    for (size_t w = 0; w < N; ++w) {
      std::vector<double> trans(L, 0.0), tau(L - 1, 0.0);
      tangentDepths(w, path, tau.data(), record_stops ? recorded_stops.data() + w * (L - 1) : nullptr);
      trans.back() = 1;
      for (size_t t = 0; t + 1 < L; ++t) trans[t] = std::exp(-tau[t]);
      double h = 0;
      for (size_t i = 0; i < L - 1; i++)
        h += 2. * ((bottom_radius + altitude[i]) * (1. - trans[i]) + (bottom_radius + altitude[i + 1]) * (1. - trans[i + 1]))
             * (altitude[i + 1] - altitude[i]) * 0.5;
      h = std::sqrt(h + bottom_radius * bottom_radius) - bottom_radius;
      const double r = h + bottom_radius;
      spectrum_flux[w] = r * r / star_radius / star_radius * 1e6;
    }
  }

  void applyModifiers(const std::vector<double>& p, size_t m) {      // forward_model.cpp:161-180
    for (size_t i = 0; i < obs.size(); ++i)
      if (obs[i].nb_modifier != 0) {
        const double shift = p[m];
        if (shift != 0) for (auto& v : bins[i]) v += shift;
        m += obs[i].nb_modifier;
      }
  }

  // EmissionModel::calcModelCPU (emission.cpp:194-228), TransmissionModel::calcModelCPU
  // (transmission.cpp:199-251), OccultationModel::calcModelCPU (secondary_eclipse.cpp:181-256)
  // + Retrieval::logLikelihood (retrieval.cpp:314-370)
  void evaluate(const std::vector<double>& p) {
    if (p.size() < nbParameters()) throw std::runtime_error("too few parameters");
    const size_t g0 = nbGeneral();
    const std::vector<double> cp(p.begin() + g0, p.begin() + g0 + nb_chem_param);
    const std::vector<double> tp(p.begin() + g0 + nb_chem_param, p.begin() + g0 + nb_chem_param + nb_temp_param);
    const size_t clp = g0 + nb_chem_param + nb_temp_param;
    const std::vector<double> cl(p.begin() + clp, p.begin() + clp + nb_cloud_param);
    const size_t mods = clp + nb_cloud_param;
    bins.resize(obs.size());
    if (model == "transmission") {
      const double fit = (fit_mmw || fit_scale_height) ? p[3] : 0.0;
      neglect = atmosphereStructure(std::pow(10, p[0]), tp, cp, p[1], fit_mmw ? fit : 0.0,
                                    fit_scale_height ? fit : 0.0);
      opacity();
      cloudOpticalDepths(cl);
      last_bottom_radius = p[1];
      transmissionSpectrum(p[1], p[2]);
      for (size_t i = 0; i < obs.size(); ++i) bins[i] = processSpectrum(grid, obs[i], spectrum_flux, false);
      applyModifiers(p, mods);
      spectrum = spectrum_flux;
    } else if (model == "secondary_eclipse") {
      neglect = atmosphereStructure(std::pow(10, p[0]), tp, cp);
      opacity();
      cloudOpticalDepths(cl);
      radiativeTransfer(1.0);
      std::vector<double> star(grid.nu.size());                     // star_blackbody.cpp:40-54
      for (size_t i = 0; i < star.size(); ++i) star[i] = planck(p[2], grid.nu[i]) * kPi * 1e-3;
      const double ratio = p[1];
      for (size_t i = 0; i < obs.size(); ++i) {
        const std::vector<double> planet_bins = processSpectrum(grid, obs[i], spectrum_flux, true);
        const std::vector<double> star_bins = processSpectrum(grid, obs[i], star, true);
        bins[i].assign(planet_bins.size(), 0.0);
        for (size_t j = 0; j < planet_bins.size(); ++j)
          bins[i][j] = planet_bins[j] / star_bins[j] * ratio * ratio * 1e6;
      }
      applyModifiers(p, mods);
      spectrum.resize(star.size());
      for (size_t i = 0; i < star.size(); ++i) spectrum[i] = spectrum_flux[i] / star[i] * ratio * ratio * 1e6;
    } else {
      neglect = atmosphereStructure(std::pow(10, p[0]), tp, cp);
      opacity();
      cloudOpticalDepths(cl);
      double scaling = kRadiusJupiter / p[2];                          // emission.cpp:108-120
      scaling = scaling * scaling * p[1];
      radiativeTransfer(scaling);
      for (size_t i = 0; i < obs.size(); ++i) bins[i] = processSpectrum(grid, obs[i], spectrum_flux, true);
      applyModifiers(p, mods);
      spectrum.resize(spectrum_flux.size());                           // radiative_transfer.h:58-60
      for (size_t i = 0; i < spectrum.size(); ++i) spectrum[i] = spectrum_flux[i] / grid.lam[i] / grid.lam[i] * 10000.0;
    }
    const double inflation = error_inflation ? std::pow(10, p.back()) : 0.0;
    log_like = 0;
    for (size_t i = 0; i < obs.size(); ++i)
      for (size_t j = 0; j < obs[i].data.size(); ++j) {
        const double e2 = obs[i].error[j] * obs[i].error[j] + inflation;
        const double d = obs[i].data[j] - bins[i][j];
        log_like += (-0.5 * std::log(e2 * 2.0 * kPi) - 0.5 * d * d / e2) * obs[i].weight[j];
      }
    if (neglect) log_like = -1e30;
  }
};

thread_local std::string g_error;

}  // namespace

// ---------------------------------------------------------------- C interface for ctypes
extern "C" {

const char* oracle_last_error() { return g_error.c_str(); }

void* oracle_emission_create(const char* folder) {
  try {
    auto* m = new EmissionOracle();
    m->load(folder);
    return m;
  } catch (std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}

void oracle_emission_destroy(void* h) { delete static_cast<EmissionOracle*>(h); }

size_t oracle_nb_points(void* h) { return static_cast<EmissionOracle*>(h)->grid.nu.size(); }
size_t oracle_nb_levels(void* h) { return static_cast<EmissionOracle*>(h)->nb_levels; }
size_t oracle_nb_parameters(void* h) { return static_cast<EmissionOracle*>(h)->nbParameters(); }
size_t oracle_nb_observations(void* h) { return static_cast<EmissionOracle*>(h)->obs.size(); }
size_t oracle_nb_bins(void* h, size_t i) { return static_cast<EmissionOracle*>(h)->obs[i].data.size(); }
size_t oracle_nb_conv(void* h, size_t i) { return static_cast<EmissionOracle*>(h)->obs[i].conv.size(); }

void oracle_grid(void* h, double* nu, double* lam, int64_t* index) {
  auto* m = static_cast<EmissionOracle*>(h);
  for (size_t i = 0; i < m->grid.nu.size(); ++i) { nu[i] = m->grid.nu[i]; lam[i] = m->grid.lam[i]; index[i] = m->grid.index[i]; }
}

// per observation: edge indices [bins][2], index range [2], profile sigma [N] (0 when absent),
// convolution intervals [range][2], filter [N] and filter weight [N] (0 when absent), norm
int oracle_observation(void* h, size_t i, int64_t* edges, int64_t* range, double* sigma, int64_t* conv,
                       double* filter, double* filter_weight, double* norm, double* data, double* error,
                       double* weight, int* ascending) {
  auto* m = static_cast<EmissionOracle*>(h);
  const Observation& o = m->obs[i];
  for (size_t b = 0; b < o.edge_idx.size(); ++b) { edges[2 * b] = o.edge_idx[b][0]; edges[2 * b + 1] = o.edge_idx[b][1]; }
  range[0] = o.idx_first;
  range[1] = o.idx_second;
  const size_t N = m->grid.nu.size();
  for (size_t k = 0; k < N; ++k) {
    sigma[k] = o.sigma.empty() ? 0.0 : o.sigma[k];
    filter[k] = o.filter.empty() ? 0.0 : o.filter[k];
    filter_weight[k] = o.filter.empty() ? 0.0 : o.filter_weight[k];
  }
  for (size_t k = 0; k < o.conv.size(); ++k) { conv[2 * k] = o.conv[k][0]; conv[2 * k + 1] = o.conv[k][1]; }
  *norm = o.filter_norm;
  for (size_t k = 0; k < o.data.size(); ++k) { data[k] = o.data[k]; error[k] = o.error[k]; weight[k] = o.weight[k]; }
  *ascending = o.ascending;
  return int(o.conv.size());
}

int oracle_emission_eval(void* h, const double* params, size_t n, double* spectrum_flux, double* spectrum,
                         double* bins, double* log_like, int* neglect) {
  try {
    auto* m = static_cast<EmissionOracle*>(h);
    m->evaluate(std::vector<double>(params, params + n));
    if (spectrum_flux) std::copy(m->spectrum_flux.begin(), m->spectrum_flux.end(), spectrum_flux);
    if (spectrum) std::copy(m->spectrum.begin(), m->spectrum.end(), spectrum);
    size_t k = 0;
    for (auto& b : m->bins) for (double v : b) bins[k++] = v;
    *log_like = m->log_like;
    *neglect = m->neglect;
    return 0;
  } catch (std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

// atmosphere of the last evaluation: T, z, p [L]; number densities [L][46]; mean molecular weight [L]
void oracle_atmosphere(void* h, double* t, double* z, double* p, double* nd, double* mmw) {
  auto* m = static_cast<EmissionOracle*>(h);
  for (size_t l = 0; l < m->nb_levels; ++l) {
    t[l] = m->temperature[l];
    z[l] = m->altitude[l];
    p[l] = m->pressure[l];
    mmw[l] = m->mmw[l];
    for (int j = 0; j < kNbSpecies; ++j) nd[l * kNbSpecies + j] = m->nd[l][j];
  }
}

// absorption and scattering coefficients of the last evaluation, [L][N]
void oracle_coefficients(void* h, double* kappa, double* scattering) {
  auto* m = static_cast<EmissionOracle*>(h);
  const size_t N = m->grid.nu.size();
  for (size_t l = 0; l < m->nb_levels; ++l)
    for (size_t i = 0; i < N; ++i) {
      kappa[l * N + i] = m->kappa[l][i];
      scattering[l * N + i] = m->scattering[l][i];
    }
}

int oracle_nb_species_ids() { return kNbSpecies; }

// the last evaluation's cloud layer optical depths [L-1][N]; 1 when there is
// a cloud, 0 (nothing written) when not
int oracle_cloud_optical_depths(void* h, double* tau) {
  auto* m = static_cast<EmissionOracle*>(h);
  if (m->cloud_tau.empty()) return 0;
  const size_t N = m->grid.nu.size();
  for (size_t l = 0; l < m->cloud_tau.size(); ++l)
    for (size_t i = 0; i < N; ++i) tau[l * N + i] = m->cloud_tau[l][i];
  return 1;
}

// the last transmission evaluation's cloud extinction [L][N]; 1 when there
// is a cloud, 0 (nothing written) when not
int oracle_cloud_extinction(void* h, double* ext) {
  auto* m = static_cast<EmissionOracle*>(h);
  if (m->cloud_ext.empty()) return 0;
  const size_t N = m->grid.nu.size();
  for (size_t l = 0; l < m->cloud_ext.size(); ++l)
    for (size_t i = 0; i < N; ++i) ext[l * N + i] = m->cloud_ext[l][i];
  return 1;
}

// the last evaluation's scale height per level (cm)
void oracle_scale_height(void* h, double* out) {
  auto* m = static_cast<EmissionOracle*>(h);
  for (size_t l = 0; l < m->scale_height.size(); ++l) out[l] = m->scale_height[l];
}

// the direct models' calcSpectrum (EmissionOracle::directSpectrum): species
// as newline-separated symbols, mixing ratios [levels][species], spectrum [N]
int oracle_direct_spectrum(void* h, double g, double a, double b, const double* p, const double* t,
                           const char* species, const double* x, int variable_gravity, const double* cloud,
                           double* spectrum) {
  try {
    auto* m = static_cast<EmissionOracle*>(h);
    if (!m->direct) throw std::runtime_error("not a direct model (no direct_range)");
    std::vector<int> ids;
    std::istringstream in(species ? species : "");
    std::string sym;
    while (in >> sym) ids.push_back(fixedChemistryId(sym));
    const size_t L = m->nb_levels;
    std::vector<double> c;
    if (cloud) c.assign(cloud, cloud + (L - 1) * m->grid.nu.size());
    m->directSpectrum(g, a, b, std::vector<double>(p, p + L), std::vector<double>(t, t + L), ids,
                      std::vector<double>(x, x + L * ids.size()), variable_gravity != 0, c);
    std::copy(m->spectrum.begin(), m->spectrum.end(), spectrum);
    return 0;
  } catch (std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

// one tangent path's sum (tangentSum) on given segment lengths [n] and extinctions [n + 1]
double oracle_tangent_sum(const double* path, const double* ext, size_t n, int32_t frozen, int32_t* last) {
  return tangentSum(path, ext, n, frozen, last);
}

// tables in memory (see MemoryTable); the arrays are referenced, not copied
int oracle_register_table(const char* folder, double mass, int log_storage, size_t nb_rows, const double* temperatures,
                          const double* pressures, size_t nb_points, const float* rows) {
  MemoryTable t;
  t.mass = mass;
  t.log_storage = log_storage != 0;
  t.temperature.assign(temperatures, temperatures + nb_rows);
  t.pressure.assign(pressures, pressures + nb_rows);
  t.nb_points = nb_points;
  t.rows = rows;
  memoryTables()[folder] = t;
  return 0;
}

void oracle_clear_tables() { memoryTables().clear(); }

// transmission: record the stopping segments of the following evaluations
void oracle_transmission_record_stops(void* h, int enable) {
  auto* m = static_cast<EmissionOracle*>(h);
  m->record_stops = enable != 0;
  if (!enable) m->recorded_stops.clear();
}

// the last evaluation's stopping segments [N][L-1] (recording on)
int oracle_transmission_stops(void* h, int32_t* out) {
  auto* m = static_cast<EmissionOracle*>(h);
  if (m->recorded_stops.empty()) { g_error = "no recorded stopping segments"; return -1; }
  std::copy(m->recorded_stops.begin(), m->recorded_stops.end(), out);
  return 0;
}

// impose stopping segments [N][L-1] on the following evaluations; NULL: the reference's rule
void oracle_transmission_freeze_stops(void* h, const int32_t* stops) {
  auto* m = static_cast<EmissionOracle*>(h);
  if (!stops) { m->frozen_stops.clear(); return; }
  const size_t n = m->grid.nu.size() * (m->nb_levels - 1);
  m->frozen_stops.assign(stops, stops + n);
}

// tangent optical depths tau[L-1] and stopping segments stop[L-1] at
// wavenumber index w, from the last evaluation's extinction and altitudes
int oracle_transmission_paths(void* h, size_t w, double* tau, int32_t* stop) {
  auto* m = static_cast<EmissionOracle*>(h);
  if (m->model != "transmission" || m->kappa.empty() || w >= m->grid.nu.size()) {
    g_error = "no transmission evaluation at this wavenumber";
    return -1;
  }
  m->tangentDepths(w, m->pathLengths(m->last_bottom_radius), tau, stop);
  return 0;
}

// closed-form helpers exposed for the known-answer tests
double oracle_planck(double t, double nu) { return planck(t, nu); }

}  // extern "C"
