// Opacity species of the host side: table ingest into the device-ordered
// row store, (p, T) bracketing, and Rayleigh cross-sections.
#include <algorithm>
#include <cmath>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

#include "host.h"

namespace bear_amd {

// ---------------------------------------------------------------- in-memory tables
namespace {
struct RegisteredTable {
  double mass = 0;
  bool log_storage = false;
  std::vector<double> temperature, pressure;
  size_t nb_wavenumbers = 0;
  std::vector<float> rows;
};
std::mutex g_registry_mutex;
std::map<std::string, RegisteredTable>& registry() {
  static std::map<std::string, RegisteredTable> r;
  return r;
}
std::string normalise(std::string folder) {
  while (!folder.empty() && folder.back() == '/') folder.pop_back();
  return folder;
}
}  // namespace

void registerTable(const std::string& folder, double mass, bool log_storage, size_t nb_rows,
                   const double* temperatures, const double* pressures, size_t nb_wavenumbers, const float* rows) {
  RegisteredTable t;
  t.mass = mass;
  t.log_storage = log_storage;
  t.temperature.assign(temperatures, temperatures + nb_rows);
  t.pressure.assign(pressures, pressures + nb_rows);
  t.nb_wavenumbers = nb_wavenumbers;
  t.rows.assign(rows, rows + nb_rows * nb_wavenumbers);
  std::lock_guard<std::mutex> lock(g_registry_mutex);
  registry()[normalise(folder)] = std::move(t);
}

void clearRegisteredTables() {
  std::lock_guard<std::mutex> lock(g_registry_mutex);
  registry().clear();
}

// ---------------------------------------------------------------- species
// TransportCoefficients::addOpacitySpecies (transport_coeff.cpp:98-237):
// CIA pairs, H2 Rayleigh and generic line absorbers
// The H- continuum (h_m.cpp, T. L. John 1988, A&A 193, 189): the bound-free
// photo-detachment cross section up to the 1.6419 micron threshold
// (h_m.cpp:56-82) and, per wavenumber, the six wavelength polynomials
// lam^2 A_i + B_i + C_i / lam + ... + F_i / lam^4 of the free-free fit
// (h_m.cpp:86-137: one coefficient set above 0.3645 micron, another down to
// 0.1823 micron), which the kernels weight per level with
// 1e-29 (5040 / T)^((i+1)/2) p_e n_H.
namespace {
void hminusRows(const SpectralGrid& grid, std::vector<double>& rows) {
  const double lambda_0 = 1.6419;
  const double cb[7] = {0.0, 152.519, 49.534, -118.858, 92.536, -34.194, 4.982};
  // [set][coefficient A..F][i]
  const double ff[2][6][7] = {
      {{0.0, 0.0, 2483.3460, -3449.8890, 2200.0400, -696.2710, 88.2830},
       {0.0, 0.0, 285.8270, -1158.3820, 2427.7190, -1841.4000, 444.5170},
       {0.0, 0.0, -2054.2910, 8746.5230, -13651.1050, 8624.9700, -1863.8650},
       {0.0, 0.0, 2827.7760, -11485.6320, 16755.5240, -10051.5300, 2095.2880},
       {0.0, 0.0, -1341.5370, 5303.6090, -7510.4940, 4400.0670, -901.7880},
       {0.0, 0.0, 208.9520, -812.9390, 1132.7380, -655.0200, 132.9850}},
      {{0.0, 518.1021, 473.2636, -482.2089, 115.5291, 0.0, 0.0},
       {0.0, -734.8666, 1443.4137, -737.1616, 169.6374, 0.0, 0.0},
       {0.0, 1021.1775, -1977.3395, 1096.8827, -245.6490, 0.0, 0.0},
       {0.0, -479.0721, 922.3575, -521.1341, 114.2430, 0.0, 0.0},
       {0.0, 93.1373, -178.9275, 101.7963, -21.9972, 0.0, 0.0},
       {0.0, -6.4285, 12.3600, -7.0571, 1.5097, 0.0, 0.0}}};
  const size_t n = grid.size();
  rows.assign(7 * n, 0.0);
  for (size_t j = 0; j < n; ++j) {
    const double lam = grid.wavelength[j];
    if (lam <= lambda_0 && lam >= 0.125) {
      const double x = 1.0 / lam - 1.0 / lambda_0;
      double f = 0;
      for (int i = 1; i < 7; ++i) f += cb[i] * std::pow(x, (i - 1) / 2.0);
      rows[j] = 1e-18 * lam * lam * lam * std::pow(x, 1.5) * f;
    }
    const int set = lam > 0.3645 ? 0 : lam >= 0.1823 ? 1 : -1;
    if (set < 0) continue;
    const double(&c)[6][7] = ff[set];
    for (int i = 1; i < 7; ++i)
      rows[size_t(i) * n + j] = lam * lam * c[0][i] + c[1][i] + c[2][i] / lam + c[3][i] / lam / lam +
                                c[4][i] / lam / lam / lam + c[5][i] / lam / lam / lam / lam;
  }
}
}  // namespace

std::vector<CloudModelSpec> parseCloudModels(const std::string& text) {
  std::vector<CloudModelSpec> out;
  std::istringstream lines(text);
  std::string line;
  while (std::getline(lines, line)) {
    std::istringstream in(line);
    std::string type, tok;
    if (!(in >> type)) continue;
    std::vector<std::string> args;
    while (in >> tok) args.push_back(tok);
    CloudModelSpec c;
    if (type == "grey") {
      c.kind = 0;
      c.fixed_bottom = args.size() == 1 && args[0] == "fb";
    } else if (type == "KHnongrey" || type == "power_law") {
      c.kind = type == "KHnongrey" ? 1 : 2;
      if (args.empty()) raise(kInvalidInput, "the " + type + " cloud model needs a reference wavelength");
      c.reference_wavelength = std::stod(args[0]);
      c.fixed_bottom = args.size() == 2 && args[1] == "fb";
    } else if (type == "none") {
      continue;
    } else {
      raise(kInvalidInput, "cloud model " + type + " unknown");
    }
    out.push_back(c);
  }
  return out;
}

bool hminusCoefficients(double temperature, double n_hminus, double n_h, double n_e, double (&c)[7]) {
  for (double& v : c) v = 0.0;
  if (n_hminus == 0) return false;                   // opacity_species.cpp:407-412
  const double pe_nh = n_e * phys::boltzmann_k * temperature * n_h;   // electron pressure x H density
  c[0] = n_hminus;
  for (int i = 1; i < 7; ++i) c[i] = 1e-29 * std::pow(5040.0 / temperature, (i + 1) / 2.0) * pe_nh;
  return true;
}

OpacitySpecies makeOpacitySpecies(const std::string& symbol, const std::string& folder, const SpectralGrid& grid) {
  OpacitySpecies s;
  s.symbol = symbol;
  s.folder = folder;
  if (symbol == "CIA-H2-H2") { s.species_id = sid::H2; s.cia_partners = {sid::H2}; }
  else if (symbol == "CIA-H2-He") { s.species_id = sid::H2; s.cia_partners = {sid::He}; }
  else if (symbol == "CIA-H-He") { s.species_id = sid::H; s.cia_partners = {sid::He}; }
  else if (folder == "Rayleigh") {
    if (symbol != "H2") raise(kInvalidInput, "Rayleigh scattering of " + symbol + " is not available");
    s.species_id = sid::H2;
    s.rayleigh = true;
    // species_rayleigh_cross_sections.cpp:22-45 with opacity_species.cpp:528-537
    s.rayleigh_cross_section.resize(grid.size());
    for (size_t j = 0; j < grid.size(); ++j) {
      const double lam = grid.wavelength[j], nu = grid.wavenumber[j];
      const double n = 13.58e-5 * (1. + 7.52e-3 / (lam * lam)) + 1.;
      const double n_ref = 2.651629e19;
      const double f = (n * n - 1) / (n * n + 2);
      s.rayleigh_cross_section[j] = 24. * phys::pi * phys::pi * phys::pi * nu * nu * nu * nu / (n_ref * n_ref) * f * f;
    }
  } else if (symbol == "H-") {                         // transport_coeff.cpp:145-150
    s.species_id = speciesId("H-");
    s.hminus = true;
    hminusRows(grid, s.hminus_rows);
  } else {
    s.species_id = speciesId(symbol);
  }
  return s;
}

// filelist.dat (opacity_species.cpp:66-157) or a registered table: the rows'
// (T, p) and sources; rows ordered by temperature, then pressure
// (opacity_species.cpp:129-157), take ids first_row, first_row + 1, ... in
// listed order.
TableSource scanSpeciesTable(OpacitySpecies& s, const std::string& path, uint32_t first_row) {
  TableSource src;
  src.symbol = s.symbol;
  src.dir = path + s.folder;
  src.first_row = first_row;
  struct Entry { double t, p; };
  std::vector<Entry> entries;
  const RegisteredTable* reg = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_registry_mutex);
    auto it = registry().find(normalise(s.folder));
    if (it == registry().end()) it = registry().find(normalise(src.dir));
    if (it != registry().end()) reg = &it->second;
  }
  if (reg) {
    src.mass = reg->mass;
    src.log_storage = reg->log_storage;
    src.registered = reg->rows.data();
    src.registered_points = reg->nb_wavenumbers;
    for (size_t r = 0; r < reg->temperature.size(); ++r) entries.push_back({reg->temperature[r], reg->pressure[r]});
  } else {
    std::ifstream f(src.dir + "/filelist.dat");
    if (!f) raise(kFileNotFound, "no cross-section data for " + s.symbol + " (" + src.dir + "/filelist.dat)");
    std::string line, storage;
    std::getline(f, line);
    std::istringstream head(line);
    head >> src.mass >> storage;
    src.log_storage = storage == "log";
    while (std::getline(f, line)) {
      std::istringstream in(line);
      Entry e{0, 0};
      std::string file;
      in >> e.p >> e.t >> file;
      if (file.empty()) continue;
      entries.push_back(e);
      src.files.push_back(file);
    }
  }
  if (entries.empty()) raise(kInvalidInput, "empty cross-section table for " + s.symbol);
  src.nb_rows = entries.size();
  std::map<double, std::vector<TableRow>> by_t;
  for (size_t r = 0; r < entries.size(); ++r)
    by_t[entries[r].t].push_back(TableRow{entries[r].t, entries[r].p, std::log10(entries[r].p), uint32_t(first_row + r)});
  s.by_temperature.clear();
  for (auto& kv : by_t) {
    auto rows = kv.second;
    std::sort(rows.begin(), rows.end(), [](const TableRow& a, const TableRow& b) { return a.pressure < b.pressure; });
    s.by_temperature.push_back(rows);
  }
  return src;
}

// every row sampled at the grid's index list, scaled by mass/N_A when a mass
// is given, floored at 1e-200 and stored as log10 in FP64; rows in parallel
void sampleSpeciesTable(const TableSource& src, const SpectralGrid& grid, std::vector<double>& out) {
  const size_t N = grid.size();
  out.resize(src.nb_rows * N);
  bool missing = false;
  std::string missing_name;
  #pragma omp parallel for schedule(dynamic, 1)
  for (size_t r = 0; r < src.nb_rows; ++r) {
    std::vector<float> raw;
    const float* data;
    size_t n_full;
    if (src.registered) {
      data = src.registered + r * src.registered_points;
      n_full = src.registered_points;
    } else {
      std::ifstream bf(src.dir + "/" + src.files[r], std::ios::binary | std::ios::ate);
      if (!bf) {
        #pragma omp critical
        { missing = true; missing_name = src.files[r]; }
        continue;
      }
      n_full = size_t(bf.tellg()) / sizeof(float);
      raw.resize(n_full);
      bf.seekg(0);
      bf.read(reinterpret_cast<char*>(raw.data()), std::streamsize(n_full * sizeof(float)));
      data = raw.data();
    }
    double* dst = out.data() + r * N;
    const double factor = src.mass / phys::avogadro;
    for (size_t i = 0; i < N; ++i) {
      double v = 0.0;
      if (grid.index[i] < n_full) {
        v = data[grid.index[i]];
        if (src.log_storage) v = std::pow(10.0, v);
      }
      if (src.mass > 0) v *= factor;
      if (v < 1e-200) v = 1e-200;
      dst[i] = std::log10(v);
    }
  }
  if (missing) raise(kFileNotFound, "cross-section file not found: " + src.dir + "/" + missing_name);
}

// OpacitySpecies::findClosestDataPoints (opacity_species.cpp:160-273) and
// the weights of calcAbsorptionCrossSections (:277-345).  A weight is zero
// where the reference skips that interpolation (equal rows), so the device
// formula  lo + (hi - lo) f_t  with  lo = r0 + (r1 - r0) f_p_lower  and
// hi = r2 + (r3 - r2) f_p_upper  reproduces every branch.
bool OpacitySpecies::sameGrid(const OpacitySpecies& o) const {
  if (by_temperature.size() != o.by_temperature.size()) return false;
  for (size_t i = 0; i < by_temperature.size(); ++i) {
    const auto& a = by_temperature[i];
    const auto& b = o.by_temperature[i];
    if (a.size() != b.size()) return false;
    for (size_t k = 0; k < a.size(); ++k)
      if (a[k].temperature != b[k].temperature || a[k].pressure != b[k].pressure ||
          a[k].log_pressure != b[k].log_pressure)
        return false;
  }
  return true;
}

void OpacitySpecies::bracketTemperature(double t, size_t& lo_i, size_t& hi_i) const {
  lo_i = hi_i = 0;
  if (t < by_temperature.front().front().temperature) {
  } else if (t > by_temperature.back().front().temperature) {
    lo_i = hi_i = by_temperature.size() - 1;
  } else {
    for (size_t i = 0; i < by_temperature.size(); ++i) {
      const double ti = by_temperature[i].front().temperature;
      if (ti == t) { lo_i = hi_i = i; break; }
      if (ti < t && by_temperature[i + 1].front().temperature > t) {
        lo_i = i;
        hi_i = i + 1;
        break;
      }
    }
  }
}

namespace {
// the two rows of one temperature bracketing p, and the log-pressure fraction
void pickPressure(const std::vector<TableRow>& rows, double p, double lp, size_t& a, size_t& b, double& f) {
  if (p < rows.front().pressure) a = b = 0;
  else if (p > rows.back().pressure) a = b = rows.size() - 1;
  else {
    a = b = 0;
    for (size_t i = 0; i < rows.size(); ++i) {
      if (rows[i].pressure == p) { a = b = i; break; }
      if (rows[i].pressure < p && rows[i + 1].pressure > p) { a = i; b = i + 1; break; }
    }
  }
  f = a != b ? (lp - rows[a].log_pressure) / (rows[b].log_pressure - rows[a].log_pressure) : 0.0;
}
}  // namespace

Stencil OpacitySpecies::bracket(double p, double t) const {
  size_t lo_i, hi_i;
  bracketTemperature(t, lo_i, hi_i);
  const double lp = std::log10(p);
  size_t p0, p1, p2, p3;
  Stencil s;
  pickPressure(by_temperature[lo_i], p, lp, p0, p1, s.f_p_lower);
  pickPressure(by_temperature[hi_i], p, lp, p2, p3, s.f_p_upper);
  const std::vector<TableRow>& lo = by_temperature[lo_i];
  const std::vector<TableRow>& hi = by_temperature[hi_i];
  const TableRow *r0 = &lo[p0], *r2 = &hi[p2];
  s.t_pos[0] = uint16_t(lo_i);
  s.t_pos[1] = uint16_t(hi_i);
  s.p_pos[0] = uint16_t(p0);
  s.p_pos[1] = uint16_t(p1);
  s.p_pos[2] = uint16_t(p2);
  s.p_pos[3] = uint16_t(p3);
  s.row[0] = r0->id;
  s.row[1] = lo[p1].id;
  s.row[2] = r2->id;
  s.row[3] = hi[p3].id;
  s.f_t = r0 != r2 ? (t - r0->temperature) / (r2->temperature - r0->temperature) : 0.0;
  s.dft = r0 != r2 ? 1.0 / (r2->temperature - r0->temperature) : 0.0;
  return s;
}

void OpacitySpecies::preparePressures(const std::vector<double>& pressures) {
  const size_t nt = by_temperature.size();
  pressure_picks_.assign(pressures.size() * nt, PressurePick{0, 0, 0.0});
  for (size_t l = 0; l < pressures.size(); ++l) {
    const double p = pressures[l], lp = std::log10(p);
    for (size_t i = 0; i < nt; ++i) {
      size_t a, b;
      double f;
      pickPressure(by_temperature[i], p, lp, a, b, f);
      pressure_picks_[l * nt + i] = PressurePick{uint16_t(a), uint16_t(b), f};
    }
  }
}

Stencil OpacitySpecies::bracketLevel(size_t level, double t) const {
  size_t lo_i, hi_i;
  bracketTemperature(t, lo_i, hi_i);
  const size_t nt = by_temperature.size();
  const PressurePick& pl = pressure_picks_[level * nt + lo_i];
  const PressurePick& ph = pressure_picks_[level * nt + hi_i];
  const std::vector<TableRow>& lo = by_temperature[lo_i];
  const std::vector<TableRow>& hi = by_temperature[hi_i];
  const TableRow *r0 = &lo[pl.lower], *r2 = &hi[ph.lower];
  Stencil s;
  s.t_pos[0] = uint16_t(lo_i);
  s.t_pos[1] = uint16_t(hi_i);
  s.p_pos[0] = pl.lower;
  s.p_pos[1] = pl.upper;
  s.p_pos[2] = ph.lower;
  s.p_pos[3] = ph.upper;
  s.row[0] = r0->id;
  s.row[1] = lo[pl.upper].id;
  s.row[2] = r2->id;
  s.row[3] = hi[ph.upper].id;
  s.f_p_lower = pl.fraction;
  s.f_p_upper = ph.fraction;
  s.f_t = r0 != r2 ? (t - r0->temperature) / (r2->temperature - r0->temperature) : 0.0;
  s.dft = r0 != r2 ? 1.0 / (r2->temperature - r0->temperature) : 0.0;
  return s;
}

}  // namespace bear_amd
