// Spectral grid and species table of the host side.
#include <algorithm>
#include <cmath>
#include <fstream>
#include <functional>

#include "host.h"

namespace bear_amd {

void raise(int code, const std::string& msg) { throw Error(code, msg); }

namespace {
// chemistry/chem_species.h:30-90: symbol, FastChem symbol and molecular weight per id
struct SpeciesEntry { const char* symbol; const char* fastchem; double weight; };
const SpeciesEntry kSpeciesTable[] = {
  {"Total", "Total", 0.0}, {"H", "H", 1.00784}, {"He", "He", 4.002602}, {"C", "C", 12.0107},
  {"O", "O", 15.999}, {"Fe", "Fe", 55.845}, {"Fe+", "Fe+", 55.845}, {"Ca", "Ca", 40.078},
  {"Ti", "Ti", 47.867}, {"Ti+", "Ti+", 47.867}, {"H2", "H2", 2.01588}, {"H2O", "H2O1", 18.01528},
  {"CO2", "C1O2", 44.01}, {"CO", "C1O1", 28.0101}, {"CH4", "C1H4", 16.04246}, {"HCN", "C1H1N1_1", 27.0253},
  {"NH3", "H3N1", 17.03052}, {"C2H2", "C2H2", 26.04}, {"N2", "N2", 28.0134}, {"Na", "Na", 22.98977},
  {"K", "K", 39.0983}, {"H2S", "H2S1", 34.09099}, {"H-", "H1-", 1.00784}, {"TiO", "O1Ti1", 63.8664},
  {"VO", "O1V1", 66.9409}, {"FeH", "H1Fe1", 56.853}, {"SH", "H1S1", 34.08}, {"MgO", "Mg1O1", 40.3044},
  {"AlO", "Al1O1", 42.981}, {"CaO", "Ca1O1", 56.0774}, {"CrH", "Cr1H1", 54.0040}, {"MgH", "H1Mg1", 26.3209},
  {"CaH", "Ca1H1", 41.0859}, {"TiH", "H1Ti1", 48.87484}, {"OH", "H1O1", 17.008}, {"e-", "e-", 5.4857990907e-4},
  {"V", "V", 50.9415}, {"V+", "V1+", 50.9415}, {"Mn", "Mn", 54.938044}, {"Si", "Si", 28.085},
  {"Cr", "Cr", 51.996}, {"Cr+", "Cr1+", 51.996}, {"SiO", "O1Si1", 44.08}, {"SiO2", "O2Si1", 60.08},
  {"SO2", "O2S1", 64.066}, {"CS2", "C1S2", 76.139}};
}  // namespace

int nbSpeciesIds() { return int(sizeof(kSpeciesTable) / sizeof(kSpeciesTable[0])); }
double molecularWeight(int id) { return kSpeciesTable[id].weight; }
int speciesId(const std::string& symbol) {
  for (int i = 0; i < nbSpeciesIds(); ++i)
    if (symbol == kSpeciesTable[i].symbol) return i;
  raise(kInvalidInput, "chemical species " + symbol + " is not known");
}
// FixedChemistry::findSpecies (fixed_chemistry.cpp:42-63): symbol or FastChem symbol
int speciesIdOrFastChem(const std::string& symbol) {
  for (int i = 0; i < nbSpeciesIds(); ++i)
    if (symbol == kSpeciesTable[i].symbol || symbol == kSpeciesTable[i].fastchem) return i;
  raise(kInvalidInput, "chemical species " + symbol + " not found in the list of species");
}

// spectral_grid.cpp:70-119
void SpectralGrid::loadFullList(const std::string& path) {
  std::ifstream f(path);
  if (!f) {
    const double step = 0.01, top = 90000.0;
    full_wavenumber.resize(size_t(top / step + 1));
    for (size_t i = 0; i < full_wavenumber.size(); ++i) full_wavenumber[i] = i * step;
  } else {
    size_t n = 0;
    f >> n;
    full_wavenumber.resize(n);
    for (auto& v : full_wavenumber) f >> v;
    if (!f) raise(kInvalidInput, "malformed wavenumber list " + path);
  }
  full_wavelength.resize(full_wavenumber.size());
  std::transform(full_wavenumber.begin(), full_wavenumber.end(), full_wavelength.begin(),
                 [](double nu) { return 1.0 / nu * 1e4; });
}

// nearest-index search on monotone data (spectral_grid.cpp:323-385): a
// lower_bound from `start`, then the closer of the two neighbours, the upper
// one on a tie
size_t SpectralGrid::nearest(double x, const std::vector<double>& list, size_t start) {
  if (list.size() < 2) return 0;
  std::vector<double>::const_iterator hit;
  if (list[0] < list[1]) hit = std::lower_bound(list.begin() + start, list.end(), x);
  else if (list[0] > list[1]) hit = std::lower_bound(list.begin() + start, list.end(), x, std::greater<double>());
  else return 0;
  const size_t k = size_t(hit - list.begin());
  if (k == start) return start;
  return std::fabs(x - list[k - 1]) < std::fabs(x - list[k]) ? k - 1 : k;
}

// SpectralGrid::sampleSpectralGrid / createHighResGrid* / findBinEdges
// (spectral_grid.cpp:123-304, 388-412) followed by SpectralBands::setBandEdgeIndices
// and SpectralBands::init (spectral_band.cpp:66-121)
std::vector<char> SpectralGrid::markRanges(const std::vector<std::array<double, 2>>& merged) const {
  std::vector<char> keep(full_wavenumber.size(), 0);
  for (auto& m : merged) {
    size_t lo = size_t(std::lower_bound(full_wavenumber.begin(), full_wavenumber.end(), m[0]) - full_wavenumber.begin());
    const size_t hi = size_t(std::lower_bound(full_wavenumber.begin(), full_wavenumber.end(), m[1]) - full_wavenumber.begin());
    if (lo >= full_wavenumber.size() || hi >= full_wavenumber.size())
      raise(kInvalidInput, "observation range outside the wavenumber list");
    if (full_wavenumber[lo] > m[0]) {
      if (lo == 0) raise(kInvalidInput, "observation range outside the wavenumber list");
      lo -= 1;
    }
    keep[lo] = 1;
    size_t last = lo;
    for (size_t j = lo; j < hi; ++j) {
      bool take;
      if (discretisation == 0) {
        const double next = full_wavenumber[last] + resolution;
        take = next == full_wavenumber[j] || full_wavenumber[j + 1] > next;
      } else if (discretisation == 1) {
        const double next = full_wavelength[last] - resolution;
        take = next == full_wavelength[j] || full_wavelength[j + 1] < next;
      } else {
        const double next = full_wavelength[last] * (1 - 1. / resolution);
        take = next == full_wavelength[j] || full_wavelength[j + 1] < next;
      }
      if (take) { keep[j] = 1; last = j; }
    }
    keep[hi] = 1;
  }
  return keep;
}

void SpectralGrid::takePoints(const std::vector<char>& keep) {
  index.clear();
  for (size_t i = 0; i < keep.size(); ++i) if (keep[i]) index.push_back(i);
  wavenumber.resize(index.size());
  wavelength.resize(index.size());
  for (size_t i = 0; i < index.size(); ++i) {
    wavenumber[i] = full_wavenumber[index[i]];
    wavelength[i] = full_wavelength[index[i]];
  }
}

void SpectralGrid::sampleRange(double wavelength_min, double wavelength_max) {
  if (!(wavelength_min > 0) || !(wavelength_max > wavelength_min))
    raise(kInvalidInput, "the spectral range needs 0 < wavelength_min < wavelength_max");
  takePoints(markRanges({{1.0 / wavelength_max * 1e4, 1.0 / wavelength_min * 1e4}}));
  if (size() < 2) raise(kInvalidInput, "the spectral range holds fewer than two wavenumbers");
}

void SpectralGrid::sample(std::vector<ObservationSetup>& observations) {
  if (observations.empty()) raise(kInvalidInput, "no observations to sample the spectral grid for");
  // merge the observations' wavenumber ranges
  std::vector<std::array<double, 2>> ranges;
  for (auto& o : observations)
    ranges.push_back({1.0 / o.range_wavelength[0] * 1e4, 1.0 / o.range_wavelength[1] * 1e4});
  std::sort(ranges.begin(), ranges.end(),
            [](const std::array<double, 2>& a, const std::array<double, 2>& b) { return a[0] < b[0]; });
  std::vector<std::array<double, 2>> merged{ranges[0]};
  for (size_t i = 1; i < ranges.size(); ++i) {
    auto& m = merged.back();
    if (ranges[i][0] > m[0] && ranges[i][1] < m[1]) continue;
    if (ranges[i][0] < m[1] && ranges[i][1] > m[1]) m[1] = ranges[i][1];
    else merged.push_back(ranges[i]);
  }
  std::vector<char> keep = markRanges(merged);
  // every band edge is a grid point
  for (auto& o : observations) {
    size_t start = 0;
    double previous = 0;
    for (auto& b : o.edges_wavenumber) {
      if (previous > b[0]) start = 0;
      const size_t i1 = nearest(b[0], full_wavenumber, start);
      const size_t i2 = nearest(b[1], full_wavenumber, start);
      keep[i1] = 1;
      keep[i2] = 1;
      previous = b[1];
      start = i2;
    }
  }
  takePoints(keep);
  for (auto& o : observations) {
    o.edge_index.clear();
    size_t start = 0;
    double previous = 0;
    for (auto& b : o.edges_wavenumber) {
      if (previous > b[0]) start = 0;
      const size_t i1 = nearest(b[0], wavenumber, start);
      const size_t i2 = nearest(b[1], wavenumber, start);
      if (i1 == i2)
        raise(kInvalidInput, "the spectral grid is too coarse to resolve the bin " +
                             std::to_string(1.0 / b[0] * 1e4) + " - " + std::to_string(1.0 / b[1] * 1e4) + " micron");
      o.edge_index.push_back({i1, i2});
      previous = b[1];
      start = i2;
    }
    o.range_index[0] = nearest(1.0 / o.range_wavelength[0] * 1e4, wavenumber, 0);
    o.range_index[1] = nearest(1.0 / o.range_wavelength[1] * 1e4, wavenumber, 0);
  }
}

}  // namespace bear_amd
