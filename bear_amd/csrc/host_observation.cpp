// Observations of the host side: input, file formats, set-up on the sampled
// grid, and the folded per-bin operator the device applies.
#include <algorithm>
#include <functional>
#include <cmath>
#include <fstream>
#include <sstream>

#include "host.h"

namespace bear_amd {

namespace {

BandType bandType(const std::string& t) {            // spectral_grid/spectral_band_type.h:51-75
  if (t == "photometry") return BandType::photometry;
  if (t == "spectroscopy") return BandType::spectroscopy;
  if (t == "band-spectroscopy") return BandType::band_spectroscopy;
  raise(kInvalidInput, "unsupported observation type " + t);
}

int modifierParameters(const std::string& m) {       // observations.cpp:367-392
  if (m.empty() || m == "none") return 0;
  if (m == "shift_const") return 1;
  raise(kInvalidInput, "observation modifier " + m + " unknown");
}

std::string detectorName(std::string d) {
  if (d == "Energy") d = "energy";
  if (d == "Photon") d = "photon";
  if (d != "energy" && d != "photon") raise(kInvalidInput, "unsupported detector type " + d);
  return d;
}

// Internal order is descending wavelength; ascending inputs are reversed
// together with their per-point vectors (observations.cpp:127-184).
void toDescending(ObservationSetup& o, std::vector<double>* centres, std::vector<std::array<double, 2>>* edges) {
  bool flip = false;
  if (centres && centres->size() > 1) flip = (*centres)[0] < (*centres)[1];
  if (edges && edges->size() > 1) flip = edges->front()[0] < edges->back()[0];
  o.ascending = flip;
  if (flip) {
    if (centres) std::reverse(centres->begin(), centres->end());
    if (edges) std::reverse(edges->begin(), edges->end());
    std::reverse(o.data.begin(), o.data.end());
    std::reverse(o.error.begin(), o.error.end());
    std::reverse(o.fwhm.begin(), o.fwhm.end());
    std::reverse(o.weight.begin(), o.weight.end());
  }
  if (edges)
    for (auto& b : *edges) if (b[0] < b[1]) std::swap(b[0], b[1]);
}

// bins from centres (observations.cpp:188-221)
std::vector<std::array<double, 2>> edgesFromCentres(const std::vector<double>& w) {
  const size_t n = w.size();
  if (n < 2) raise(kInvalidInput, "a spectroscopy observation needs at least two points");
  std::vector<std::array<double, 2>> e(n);
  for (size_t i = 0; i < n; ++i) {
    if (i == 0) {
      e[0][1] = w[0] - (w[0] - w[1]) * 0.5;
      e[0][0] = w[0] + (w[0] - w[1]) * 0.5;
    } else if (i == n - 1) {
      e[i][1] = w[i] - (w[i - 1] - w[i]) * 0.5;
      e[i][0] = w[i] + (w[i - 1] - w[i]) * 0.5;
    } else {
      e[i][0] = w[i - 1] - (w[i - 1] - w[i]) * 0.5;
      e[i][1] = w[i] - (w[i] - w[i + 1]) * 0.5;
    }
  }
  return e;
}

void finishInput(ObservationSetup& o, std::vector<double> centres, std::vector<std::array<double, 2>> edges) {
  if (o.type == BandType::spectroscopy) {
    toDescending(o, &centres, nullptr);
    edges = edgesFromCentres(centres);
  } else {
    toDescending(o, nullptr, &edges);
    centres.resize(edges.size());
    for (size_t i = 0; i < edges.size(); ++i) centres[i] = edges[i][0] - (edges[i][0] - edges[i][1]) * 0.5;
  }
  o.centres = centres;
  o.edges_wavelength = edges;
  // observation range, widened by 5 sigma of the profile (observations.cpp:236-251)
  if (!o.fwhm.empty()) {
    o.range_wavelength[0] = edges.front()[0] + 5.0 * o.fwhm.front() / 2.355;
    o.range_wavelength[1] = edges.back()[1] - 5.0 * o.fwhm.back() / 2.355;
  } else {
    o.range_wavelength[0] = edges.front()[0];
    o.range_wavelength[1] = edges.back()[1];
  }
  o.edges_wavenumber.resize(edges.size());
  for (size_t i = 0; i < edges.size(); ++i)
    o.edges_wavenumber[i] = {1.0 / edges[i][0] * 1e4, 1.0 / edges[i][1] * 1e4};
  if (!o.filter_wavelength.empty() && o.filter_wavelength.front() < o.filter_wavelength.back()) {
    std::reverse(o.filter_wavelength.begin(), o.filter_wavelength.end());
    std::reverse(o.filter_response.begin(), o.filter_response.end());
  }
}

// bandpass file: four header lines, detector type first word of line 4,
// then "wavelength response" pairs (observations_filter_response.cpp:42-117)
void readBandpass(ObservationSetup& o, const std::string& path) {
  std::ifstream f(path);
  if (!f) raise(kFileNotFound, "bandpass file not found: " + path);
  std::string line;
  for (int i = 0; i < 4; ++i) std::getline(f, line);
  std::string det;
  std::stringstream(line) >> det;
  o.detector = detectorName(det);
  while (std::getline(f, line)) {
    std::stringstream s(line);
    double a, b;
    if (!(s >> a >> b)) continue;
    o.filter_wavelength.push_back(a);
    o.filter_response.push_back(b);
  }
  if (o.filter_wavelength.size() < 2) raise(kInvalidInput, "bandpass file without data: " + path);
}

unsigned countColumns(const std::string& line) {
  std::stringstream ss(line);
  std::string item;
  unsigned n = 0;
  while (ss >> item) ++n;
  return n;
}

}  // namespace

ObservationSetup observationFromInput(const ObservationInputView& in) {     // Observation::init (observations.cpp:40-111)
  ObservationSetup o;
  o.name = in.name;
  o.type = bandType(in.type);
  o.modifier_parameters = modifierParameters(in.modifier);
  o.data = in.data;
  o.error = in.error;
  o.fwhm = in.fwhm;
  o.weight = in.weight.empty() ? std::vector<double>(in.data.size(), 1.0) : in.weight;
  if (!in.filter_wavelength.empty()) {
    o.detector = detectorName(in.detector);
    o.filter_wavelength = in.filter_wavelength;
    o.filter_response = in.filter_response;
  }
  finishInput(o, in.wavelengths, in.bin_edges);
  return o;
}

// observation files (observations_load_data.cpp:40-330)
ObservationSetup observationFromFile(const std::string& file, const std::string& modifier,
                                     const std::string& folder) {
  ObservationSetup o;
  o.modifier_parameters = modifierParameters(modifier);
  std::ifstream f(file);
  if (!f) raise(kFileNotFound, "observation file not found: " + file);
  std::string line, type, filter_path;
  std::getline(f, line);
  std::getline(f, line);
  std::getline(f, o.name);
  std::getline(f, line);
  std::getline(f, line);
  f >> type;
  o.type = bandType(type);
  for (int i = 0; i < 3; ++i) std::getline(f, line);
  f >> filter_path;
  const bool has_filter = !(filter_path == "None" || filter_path == "none" || filter_path.empty());
  for (int i = 0; i < 3; ++i) std::getline(f, line);

  std::vector<double> centres;
  std::vector<std::array<double, 2>> edges;
  if (o.type == BandType::photometry) {
    std::getline(f, line);
    const unsigned cols = countColumns(line);
    if (cols < 4) raise(kInvalidInput, "expected at least four data columns in " + file);
    std::istringstream in(line);
    double lo = 0, hi = 0, v = 0, e = 0, w = 0;
    in >> lo >> hi >> v >> e >> w;
    edges.push_back({lo, hi});
    o.data.push_back(v);
    o.error.push_back(e);
    o.weight.push_back(cols == 5 ? w : 1.0);
  } else {
    const bool bands = o.type == BandType::band_spectroscopy;
    const unsigned min_cols = bands ? 4 : 3;
    unsigned cols = 0;
    while (std::getline(f, line)) {
      if (cols == 0) {
        cols = countColumns(line);
        if (cols < min_cols) raise(kInvalidInput, "too few data columns in " + file);
      }
      std::istringstream in(line);
      double a = 0, b = 0, v = 0, e = 0, fw = 0, w = 0;
      if (bands) in >> a >> b >> v >> e >> fw >> w;
      else in >> a >> v >> e >> fw >> w;
      if (a != 0.0 && e != 0.0) {
        if (bands) edges.push_back({a, b});
        else centres.push_back(a);
        o.data.push_back(v);
        o.error.push_back(e);
      }
      if (fw != 0.0 && cols > min_cols) o.fwhm.push_back(fw);
      o.weight.push_back(cols == 5 ? w : 1.0);
    }
  }
  if (has_filter) readBandpass(o, folder + filter_path);
  finishInput(o, centres, edges);
  return o;
}

// profile, convolution windows and bandpass on the sampled grid
// (spectral_band.cpp:127-154, spectral_band_convolve.cpp:44-95,
//  spectral_grid_interpolate.cpp:123-172, observations_filter_response.cpp:122-160)
// linear interpolation of (x, y) onto descending new_x; zero outside [x_min, x_max]
// (the first x at or below each new point, found by bisection on the
// descending x: the element the reference's linear search finds)
std::vector<double> interpolateDescending(std::vector<double> x, std::vector<double> y, const std::vector<double>& nx) {
  if (x[0] < x[1]) {
    std::reverse(x.begin(), x.end());
    std::reverse(y.begin(), y.end());
  }
  std::vector<double> out(nx.size(), 0.0);
  for (size_t i = 0; i < nx.size(); ++i) {
    if (nx[i] > x.front()) continue;
    if (nx[i] < x.back()) break;
    const size_t k = size_t(std::lower_bound(x.begin(), x.end(), nx[i], std::greater<double>()) - x.begin());
    out[i] = x[k] == nx[i] ? y[k] : y[k - 1] + (y[k] - y[k - 1]) * (nx[i] - x[k - 1]) / (x[k] - x[k - 1]);
  }
  return out;
}

void ObservationSetup::setupOnGrid(const SpectralGrid& grid) {
  const size_t N = grid.size();
  if (!fwhm.empty() && fwhm.size() == centres.size()) {
    const std::vector<double> local(grid.wavelength.begin() + range_index[0], grid.wavelength.begin() + range_index[1] + 1);
    const std::vector<double> f = interpolateDescending(centres, fwhm, local);
    profile_sigma.assign(N, 0.0);
    for (size_t i = 0; i < local.size(); ++i) profile_sigma[i + range_index[0]] = f[i] / 2.355;
    convolution_window.resize(local.size());
    for (size_t i = 0; i < local.size(); ++i) {
      const size_t k = i + range_index[0];
      const double cut = 5.0 * profile_sigma[k];
      if (cut == 0) {
        convolution_window[i] = {k, k};
      } else {
        const size_t a = SpectralGrid::nearest(grid.wavelength[k] + cut, grid.wavelength, 0);
        const size_t b = SpectralGrid::nearest(grid.wavelength[k] - cut, grid.wavelength, a);
        convolution_window[i] = {a, b};
      }
    }
  }
  if (!filter_wavelength.empty()) {
    filter = interpolateDescending(filter_wavelength, filter_response, grid.wavelength);
    filter_weight.assign(N, 1.0);
    std::vector<double> y = filter_response;
    if (detector == "photon") {
      filter_weight = grid.wavelength;
      for (size_t i = 0; i < y.size(); ++i) y[i] *= filter_wavelength[i];
    }
    double norm = 0.0;
    for (size_t i = 1; i < y.size(); ++i)
      norm += (filter_wavelength[i] - filter_wavelength[i - 1]) * (y[i] + y[i - 1]);
    norm = std::fabs(norm * 0.5);
    filter_normalisation = norm;
  }
}

// Fold bandpass, convolution and band integral into one weight segment per
// bin.  All three stages are linear and every weight is non-negative for a
// non-negative spectrum, so the absolute values the reference takes
// (spectral_band_convolve.cpp:143, spectral_band_integrate.cpp:95) only fix
// the sign of each stage.
void ObservationSetup::foldOperator(const SpectralGrid& grid, bool is_flux) {
  const std::vector<double>& nu = grid.wavenumber;
  const std::vector<double>& lam = grid.wavelength;
  const bool lsf = !profile_sigma.empty();
  const bool bandpass = !filter.empty();
  segment_start.clear();
  segment_length.clear();
  segment_weight.clear();
  segment_offset.clear();

  std::vector<double> acc;
  for (auto& e : edge_index) {
    const size_t i0 = e[0], i1 = e[1];
    // grid span the bin can reach through the profile windows
    size_t span_lo = i0, span_hi = i1;
    if (lsf)
      for (size_t i = std::max(i0, range_index[0]); i <= std::min(i1, range_index[1]); ++i) {
        const auto& w = convolution_window[i - range_index[0]];
        span_lo = std::min(span_lo, w[0]);
        span_hi = std::max(span_hi, w[1]);
      }
    acc.assign(span_hi - span_lo + 1, 0.0);
    auto add = [&](size_t j, double v) { acc[j - span_lo] += v; };
    // trapezoid weights of the band integral
    const std::vector<double>& x = is_flux ? nu : lam;
    double denominator;
    if (is_flux) denominator = 1.0 / nu[i0] * 1e4 - 1.0 / nu[i1] * 1e4;
    else denominator = bandpass ? 1.0 : (lam[i0] - lam[i1]);
    for (size_t i = i0; i <= i1; ++i) {
      double t = 0.0;
      if (i > i0) t += 0.5 * (x[i] - x[i - 1]);
      if (i < i1) t += 0.5 * (x[i + 1] - x[i]);
      t = std::fabs(t) / std::fabs(denominator);
      if (t == 0.0) continue;
      // row i of the convolution
      if (!lsf) { add(i, t); continue; }
      if (i < range_index[0] || i > range_index[1]) continue;     // zero outside the observation range
      const auto& w = convolution_window[i - range_index[0]];
      if (w[0] == w[1] || profile_sigma[i] == 0) { add(i, t); continue; }
      const size_t a = w[0], b = w[1];                             // window [a, b), end excluded
      const double sigma = profile_sigma[i];
      const double c = std::sqrt(1.0 / (2.0 * phys::pi)) / sigma;
      if (b - a == 1) {
        add(a, t * c * std::exp(-(lam[i] - lam[a]) * (lam[i] - lam[a]) / (2.0 * sigma * sigma)));
        continue;
      }
      for (size_t j = a; j < b; ++j) {
        double u = 0.0;
        if (j > a) u += 0.5 * (lam[j] - lam[j - 1]);
        if (j + 1 < b) u += 0.5 * (lam[j + 1] - lam[j]);
        const double d = std::fabs(lam[i] - lam[j]);
        add(j, t * std::fabs(u) * c * std::exp(-d * d / (2.0 * sigma * sigma)));
      }
    }
    if (bandpass)
      for (size_t j = span_lo; j <= span_hi; ++j)
        acc[j - span_lo] *= filter[j] * filter_weight[j] / filter_normalisation;
    size_t lo = span_hi + 1, hi = span_lo;
    for (size_t j = span_lo; j <= span_hi; ++j)
      if (acc[j - span_lo] != 0.0) { lo = std::min(lo, j); hi = std::max(hi, j); }
    if (lo > hi) { lo = i0; hi = i0; }
    segment_start.push_back(lo);
    segment_length.push_back(hi - lo + 1);
    segment_offset.push_back(segment_weight.size());
    for (size_t j = lo; j <= hi; ++j) segment_weight.push_back(acc[j - span_lo]);
  }
}

}  // namespace bear_amd
