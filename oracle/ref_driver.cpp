// Test infrastructure: a driver around the reference's own CPU sources.
//
// Compiled only by oracle/ref.mk against the reference tree; it is never part
// of the product.  It composes the reference's CPU emission path in the same
// order as EmissionModel::calcModelCPU (forward_model/emission/emission.cpp:194-228)
// and Retrieval::logLikelihood (retrieval/retrieval.cpp:314-370), after the
// set-up sequence of Retrieval::loadObservations
// (retrieval/retrieval_load_observations.cpp:68-92).  The log-likelihood sum is
// restated here because retrieval.cpp needs Boost headers this image lacks.
//
// The secondary-eclipse case composes the same pieces with the blackbody
// star as OccultationModel::calcModelCPU does (secondary_eclipse.cpp:181-256),
// restating its 10-line ratio loop.  For the transmission case the driver
// runs the reference's atmosphere variants and opacity
// (TransmissionModel::calcAtmosphereStructure, transmission.cpp:155-195) and
// writes them out; the tangent-path integral is a TransmissionModel member
// whose class cannot be built here.
//
// usage: bear_ref_emission <case folder> <output folder> [n_timed_evals]
// Outputs are .npy arrays (little-endian float64 / int64) in the output folder.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "config/global_config.h"
#include "spectral_grid/spectral_grid.h"
#include "observations/observations.h"
#include "forward_model/atmosphere/atmosphere.h"
#include "transport_coeff/opacity_calc.h"
#include "temperature/piecewise_poly_temperature.h"
#include "temperature/constant_temperature.h"
#include "chemistry/isoprofile_chemistry.h"
#include "chemistry/background_chemistry.h"
#include "radiative_transfer/short_characteristics.h"
#include "cloud_model/select_cloud_model.h"
#include "cloud_model/fixed_cloud_model.h"
#include "forward_model/stellar_spectrum/star_blackbody.h"
#include "forward_model/stellar_spectrum/star_file_spectrum.h"
#include "additional/physical_const.h"

namespace {

void writeNpy(const std::string& path, const std::vector<double>& v, size_t rows = 0) {
  std::string shape = rows ? "(" + std::to_string(rows) + ", " + std::to_string(v.size() / rows) + ")"
                           : "(" + std::to_string(v.size()) + ",)";
  std::string header = "{'descr': '<f8', 'fortran_order': False, 'shape': " + shape + ", }";
  while ((10 + header.size() + 1) % 64 != 0) header += ' ';
  header += '\n';
  std::ofstream f(path, std::ios::binary);
  const char magic[] = "\x93NUMPY\x01\x00";
  f.write(magic, 8);
  uint16_t len = static_cast<uint16_t>(header.size());
  f.write(reinterpret_cast<const char*>(&len), 2);
  f.write(header.data(), header.size());
  f.write(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(double));
}

template <class T>
std::vector<double> asDouble(const std::vector<T>& v) { return std::vector<double>(v.begin(), v.end()); }

struct CaseText {
  std::string model = "emission";
  bool variable_gravity = false, fit_mmw = false, fit_scale_height = false;
  size_t nb_levels = 0;
  double resolution = 0, p_bottom = 0, p_top = 0;
  int discretisation = 0;
  std::vector<std::string> temperature;
  std::vector<std::vector<std::string>> chemistry;
  std::vector<std::string> opac_symbol, opac_folder;
  std::vector<std::vector<std::string>> clouds;       // "type args..." of each cloud model
  bool error_inflation = false;
  std::vector<double> parameters;
  bool direct = false;                     // pyBeAR's direct models over direct_range (micron)
  double direct_range[2] = {0, 0};
  std::string stellar_spectrum;
};

// The direct models' inputs (direct_input.txt): gravity, a, b,
// variable_gravity, species symbols, then one "p T x_1 .. x_S" line per level.
struct DirectInput {
  double g = 0, a = 0, b = 0;
  bool variable_gravity = false;
  std::vector<std::string> species;
  std::vector<double> pressure, temperature;
  std::vector<std::vector<double>> mixing_ratios;
  std::vector<std::vector<double>> cloud;      // "cloud" lines: one layer's optical depth per grid point
};

DirectInput readDirectInput(const std::string& path) {
  DirectInput d;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream in(line);
    std::string key;
    in >> key;
    if (key == "gravity") in >> d.g;
    else if (key == "a") in >> d.a;
    else if (key == "b") in >> d.b;
    else if (key == "variable_gravity") { int v = 0; in >> v; d.variable_gravity = v != 0; }
    else if (key == "species") { std::string s; while (in >> s) d.species.push_back(s); }
    else if (key == "level") {
      double p, t;
      in >> p >> t;
      std::vector<double> x(d.species.size());
      for (auto& v : x) in >> v;
      d.pressure.push_back(p);
      d.temperature.push_back(t);
      d.mixing_ratios.push_back(x);
    } else if (key == "cloud") {
      std::vector<double> row;
      double v;
      while (in >> v) row.push_back(v);
      d.cloud.push_back(row);
    }
  }
  return d;
}

CaseText readCase(const std::string& folder) {
  CaseText c;
  std::ifstream f(folder + "case.txt");
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream in(line);
    std::string key;
    in >> key;
    std::vector<std::string> rest;
    std::string tok;
    while (in >> tok) rest.push_back(tok);
    if (key == "model") c.model = rest[0];
    else if (key == "use_variable_gravity") c.variable_gravity = rest[0] == "1";
    else if (key == "fit_mean_molecular_weight") c.fit_mmw = rest[0] == "1";
    else if (key == "fit_scale_height") c.fit_scale_height = rest[0] == "1";
    else if (key == "spectral_resolution") c.resolution = std::stod(rest[0]);
    else if (key == "spectral_discretisation") c.discretisation = std::stoi(rest[0]);
    else if (key == "nb_levels") c.nb_levels = std::stoul(rest[0]);
    else if (key == "p_bottom") c.p_bottom = std::stod(rest[0]);
    else if (key == "p_top") c.p_top = std::stod(rest[0]);
    else if (key == "temperature") c.temperature = rest;
    else if (key == "chemistry") c.chemistry.push_back(rest);
    else if (key == "opacity") { c.opac_symbol.push_back(rest[0]); c.opac_folder.push_back(rest[1]); }
    else if (key == "error_inflation") c.error_inflation = rest[0] == "1";
    else if (key == "cloud") c.clouds.push_back(rest);
    else if (key == "direct_range") {
      c.direct = true;
      c.direct_range[0] = std::stod(rest[0]);
      c.direct_range[1] = std::stod(rest[1]);
    } else if (key == "stellar_spectrum") c.stellar_spectrum = rest[0];
    else if (key == "parameters") {
      size_t n = std::stoul(rest[0]);
      for (size_t i = 0; i < n; ++i) c.parameters.push_back(std::stod(rest[1 + i]));
    }
  }
  return c;
}

}  // namespace

// pyBeAR's direct models, composed as EmissionModel::calcSpectrum
// (emission.cpp:268-340) and OccultationModel::calcSpectrum
// (secondary_eclipse.cpp:395-482) do on the CPU; for transmission the
// atmosphere and the coefficients (the tangent-path integral is a
// TransmissionModel member, whose class cannot be built here).
int runDirect(const CaseText& cs, bear::GlobalConfig& config, const std::string& folder, const std::string& out) {
  using namespace bear;
  SpectralGrid grid(&config, cs.direct_range[0], cs.direct_range[1]);
  const std::vector<double> boundaries = {10, 1e-5};
  Atmosphere atmosphere(cs.nb_levels, boundaries, false);
  OpacityCalculation opacity(&config, &grid, &atmosphere, cs.opac_symbol, cs.opac_folder, false, true);
  ShortCharacteristics rt(&grid);
  const DirectInput d = readDirectInput(folder + "direct_input.txt");
  const bool transmission = cs.model == "transmission", eclipse = cs.model == "secondary_eclipse";
  const double radius = transmission || !eclipse ? d.a : 1.0;
  atmosphere.setAtmosphericStructure(d.g, radius, eclipse ? false : d.variable_gravity, d.pressure, d.temperature,
                                     d.species, d.mixing_ratios);
  // EmissionModel::setCloudProperties (emission.cpp:344-367): a fixed cloud
  // when any optical depth is positive
  std::vector<CloudModel*> clouds;
  bool use_cloud = false;
  for (auto& row : d.cloud)
    for (double v : row) use_cloud = use_cloud || v > 0;
  if (use_cloud) {
    const std::vector<std::vector<double>> zeros(d.cloud.size(), std::vector<double>(d.cloud[0].size(), 0.0));
    clouds.push_back(new FixedCloudModel(d.cloud, zeros, zeros));
  }
  opacity.calculate(clouds, std::vector<double>{});
  std::vector<double> spectrum;
  if (!transmission) {
    spectrum.assign(grid.nbSpectralPoints(), 0.0);
    rt.calcSpectrum(atmosphere, opacity.absorption_coeff, opacity.absorption_coeff, opacity.cloud_optical_depths,
                    opacity.cloud_single_scattering, opacity.cloud_asym_param, 1.0, spectrum);
    if (eclipse) {
      std::vector<double> lam, flux;
      std::ifstream f(folder + cs.stellar_spectrum);
      double a, b;
      while (f >> a >> b) { lam.push_back(a); flux.push_back(b); }
      StarSpectrumFile star(lam, flux, &grid);
      const std::vector<double> stellar = star.calcFlux(std::vector<double>{});
      for (size_t i = 0; i < spectrum.size(); ++i) spectrum[i] = spectrum[i] / stellar[i] * d.a * d.a * 1e6;
    } else {
      for (size_t i = 0; i < spectrum.size(); ++i) spectrum[i] *= d.a * d.a / d.b / d.b;
    }
  }
  writeNpy(out + "wavenumber.npy", grid.wavenumber_list);
  writeNpy(out + "wavelength.npy", grid.wavelength_list);
  writeNpy(out + "temperature.npy", atmosphere.temperature);
  writeNpy(out + "pressure.npy", atmosphere.pressure);
  writeNpy(out + "altitude.npy", atmosphere.altitude);
  std::vector<double> nd;
  for (auto& l : atmosphere.number_densities) nd.insert(nd.end(), l.begin(), l.end());
  writeNpy(out + "number_densities.npy", nd, atmosphere.number_densities.size());
  std::vector<double> k, sc;
  for (auto& row : opacity.absorption_coeff) k.insert(k.end(), row.begin(), row.end());
  for (auto& row : opacity.scattering_coeff) sc.insert(sc.end(), row.begin(), row.end());
  writeNpy(out + "absorption_coeff.npy", k, opacity.absorption_coeff.size());
  writeNpy(out + "scattering_coeff.npy", sc, opacity.scattering_coeff.size());
  if (!spectrum.empty()) writeNpy(out + "spectrum.npy", spectrum);
  if (use_cloud) {
    std::vector<std::vector<double>> ext(grid.nbSpectralPoints(), std::vector<double>(cs.nb_levels, 0.0));
    clouds[0]->convertOpticalDepth(opacity.cloud_optical_depths, ext, atmosphere.altitude);
    std::vector<double> flat;
    for (size_t l = 0; l < cs.nb_levels; ++l)
      for (size_t i = 0; i < grid.nbSpectralPoints(); ++i) flat.push_back(ext[i][l]);
    writeNpy(out + "cloud_extinction.npy", flat, cs.nb_levels);
  }
  std::printf("direct %s N=%zu\n", cs.model.c_str(), grid.nbSpectralPoints());
  return 0;
}

int main(int argc, char** argv) {
  using namespace bear;
  if (argc < 3) { std::cerr << "usage: bear_ref_emission <case> <out> [n_timed]\n"; return 2; }
  std::string folder = argv[1];
  if (folder.back() != '/') folder += '/';
  std::string out = argv[2];
  if (out.back() != '/') out += '/';
  const int n_timed = argc > 3 ? std::atoi(argv[3]) : 0;

  CaseText cs = readCase(folder);

  GlobalConfig config;
  config.use_gpu = false;
  config.retrieval_folder_path = folder;
  config.wavenumber_file_path = folder + "wavenumber_full.dat";
  config.cross_section_file_path = folder + "opacity/";
  config.spectral_disecretisation = cs.discretisation;
  config.spectral_resolution = cs.resolution;
  config.use_error_inflation = cs.error_inflation;

  if (cs.direct) return runDirect(cs, config, folder, out);

  SpectralGrid grid(&config);

  // observations, in the order of Retrieval::loadObservations
  std::vector<std::string> files, modifiers;
  {
    std::ifstream f(folder + "observations.list");
    std::string line;
    while (std::getline(f, line)) {
      std::istringstream in(line);
      std::string a, b;
      in >> a >> b;
      if (a.empty()) continue;
      files.push_back(a);
      modifiers.push_back(b);
    }
  }
  std::vector<Observation> observations(files.size(), Observation(&config, &grid));
  for (size_t i = 0; i < files.size(); ++i) observations[i].init(folder + files[i], modifiers[i]);
  grid.sampleSpectralGrid(observations);
  for (auto& o : observations) {
    o.spectral_bands.setInstrumentProfileFWHW(o.instrument_profile_fwhm);
    o.setFilterResponseFunction();
  }

  // the emission model's modules (EmissionModel ctor + initModules)
  const std::vector<double> boundaries = {cs.p_bottom, cs.p_top};
  Atmosphere atmosphere(cs.nb_levels, boundaries, false);
  OpacityCalculation opacity(&config, &grid, &atmosphere, cs.opac_symbol, cs.opac_folder, false,
                             !cs.clouds.empty());
  ShortCharacteristics rt(&grid);
  // the cloud modules (EmissionModel::initModules, emission_init.cpp:72-89)
  std::vector<CloudModel*> cloud_models;
  size_t nb_cloud_param = 0;
  for (auto& c : cs.clouds) {
    cloud_models.push_back(selectCloudModel(c[0], std::vector<std::string>(c.begin() + 1, c.end())));
    nb_cloud_param += cloud_models.back()->nbParameters();
  }
  std::unique_ptr<Temperature> temperature_profile;
  if (cs.temperature[0] == "const") temperature_profile.reset(new ConstantTemperature());
  else temperature_profile.reset(new PiecewisePolynomialTemperature(std::stoul(cs.temperature[1]),
                                                                    std::stoul(cs.temperature[2]), boundaries));
  Temperature& temperature = *temperature_profile;
  std::vector<Chemistry*> chemistry;
  size_t nb_chem_param = 0;
  for (auto& c : cs.chemistry) {
    if (c[0] == "iso") chemistry.push_back(new IsoprofileChemistry(std::vector<std::string>(c.begin() + 1, c.end())));
    else chemistry.push_back(new BackgroundChemistry(c[1]));
    nb_chem_param += chemistry.back()->nbParameters();
  }
  size_t nb_modifier = 0;
  for (auto& o : observations) nb_modifier += o.nb_modifier_param;

  const size_t nb_temp_param = temperature.nbParameters();
  const bool transmission = cs.model == "transmission", eclipse = cs.model == "secondary_eclipse";
  // general parameters (emission.cpp:68-69, transmission.cpp:69-72, secondary_eclipse.cpp:72 + star)
  const size_t nb_general = transmission ? 3 + ((cs.fit_mmw || cs.fit_scale_height) ? 1 : 0) : 3;
  const std::vector<double>& p = cs.parameters;
  const std::vector<double> model_param(p.begin(), p.begin() + nb_general);
  const std::vector<double> chem_param(p.begin() + nb_general, p.begin() + nb_general + nb_chem_param);
  const std::vector<double> temp_param(p.begin() + nb_general + nb_chem_param,
                                       p.begin() + nb_general + nb_chem_param + nb_temp_param);
  const size_t at_cloud = nb_general + nb_chem_param + nb_temp_param;
  const std::vector<double> cloud_param(p.begin() + at_cloud, p.begin() + at_cloud + nb_cloud_param);
  const std::vector<double> modifier_param(p.begin() + at_cloud + nb_cloud_param,
                                           p.begin() + at_cloud + nb_cloud_param + nb_modifier);
  std::vector<std::vector<double>> cloud_extinction;
  StarBlackBody star(&grid);

  std::vector<double> spectrum, spectrum_high_res_flux;
  std::vector<std::vector<double>> spectrum_obs(observations.size());
  bool neglect = false;
  double log_like = 0;

  auto evaluate = [&]() {
    const double surface_gravity = std::pow(10, model_param[0]);
    if (transmission) {
      // TransmissionModel::calcAtmosphereStructure (transmission.cpp:155-195)
      if (cs.fit_mmw)
        neglect = atmosphere.calcAtmosphereStructure(surface_gravity, model_param[1], cs.variable_gravity,
                                                     &temperature, temp_param, chemistry, chem_param, model_param[3]);
      else if (cs.fit_scale_height)
        neglect = atmosphere.calcAtmosphereStructure(surface_gravity, model_param[3], &temperature, temp_param,
                                                     chemistry, chem_param);
      else
        neglect = atmosphere.calcAtmosphereStructure(surface_gravity, model_param[1], cs.variable_gravity,
                                                     &temperature, temp_param, chemistry, chem_param);
      opacity.calculate(cloud_models, cloud_param);
      // TransmissionModel::calcModelCPU's cloud extinction (transmission.cpp:213-223)
      cloud_extinction.assign(grid.nbSpectralPoints(), std::vector<double>(cs.nb_levels, 0.0));
      if (!cloud_models.empty())
        cloud_models[0]->convertOpticalDepth(opacity.cloud_optical_depths, cloud_extinction, atmosphere.altitude);
      spectrum.clear();
      return;
    }
    // EmissionModel::calcAtmosphereStructure (emission.cpp:158-189); the eclipse
    // model's is the same call (secondary_eclipse.cpp:159-176)
    neglect = atmosphere.calcAtmosphereStructure(surface_gravity, 1.0, false, &temperature, temp_param,
                                                 chemistry, chem_param);
    opacity.calculate(cloud_models, cloud_param);
    spectrum.assign(grid.nbSpectralPoints(), 0.0);
    double scaling = 1.0;
    if (!eclipse) {
      // EmissionModel::radiusDistanceScaling (emission.cpp:108-120)
      scaling = constants::radius_jupiter / model_param[2];
      scaling = scaling * scaling * model_param[1];
    }
    rt.calcSpectrum(atmosphere, opacity.absorption_coeff, opacity.absorption_coeff, opacity.cloud_optical_depths,
                    opacity.cloud_single_scattering, opacity.cloud_asym_param, scaling, spectrum);
    spectrum_high_res_flux = spectrum;
    if (eclipse) {
      std::vector<double> stellar = star.calcFlux(std::vector<double>{model_param[2]});
      const double ratio = model_param[1];
      for (size_t i = 0; i < observations.size(); ++i) {
        std::vector<double> planet_obs = observations[i].processModelSpectrum(spectrum, true);
        std::vector<double> star_obs = observations[i].processModelSpectrum(stellar, true);
        spectrum_obs[i].assign(observations[i].nbPoints(), 0.0);
        for (size_t j = 0; j < spectrum_obs[i].size(); ++j)
          spectrum_obs[i][j] = planet_obs[j] / star_obs[j] * ratio * ratio * 1e6;
      }
      for (size_t i = 0; i < spectrum.size(); ++i) spectrum[i] = spectrum[i] / stellar[i] * ratio * ratio * 1e6;
    } else {
      for (size_t i = 0; i < observations.size(); ++i)
        spectrum_obs[i] = observations[i].processModelSpectrum(spectrum, true);
    }
    // ForwardModel::applyObservationModifier (forward_model.cpp:161-180)
    auto it = modifier_param.begin();
    for (size_t i = 0; i < observations.size(); ++i)
      if (observations[i].nb_modifier_param != 0) {
        const double shift = *it;
        if (shift != 0) for (auto& s : spectrum_obs[i]) s += shift;
        it += observations[i].nb_modifier_param;
      }
    if (!eclipse) rt.changeSpectrumUnits(spectrum);
    // Gaussian log-likelihood, restating retrieval.cpp:330-362
    const double inflation = config.use_error_inflation ? std::pow(10, p.back()) : 0.0;
    log_like = 0;
    for (size_t i = 0; i < observations.size(); ++i)
      for (size_t j = 0; j < observations[i].nbPoints(); ++j) {
        const double e2 = observations[i].data_error[j] * observations[i].data_error[j] + inflation;
        const double d = observations[i].data[j] - spectrum_obs[i][j];
        log_like += (-0.5 * std::log(e2 * 2.0 * constants::pi) - 0.5 * d * d / e2) * observations[i].likelihood_weight[j];
      }
    if (neglect) log_like = -1e30;
  };

  evaluate();

  std::vector<double> timings;
  for (int k = 0; k < n_timed; ++k) {
    auto t0 = std::chrono::steady_clock::now();
    evaluate();
    auto t1 = std::chrono::steady_clock::now();
    timings.push_back(std::chrono::duration<double>(t1 - t0).count());
  }

  // outputs
  writeNpy(out + "wavenumber.npy", grid.wavenumber_list);
  writeNpy(out + "wavelength.npy", grid.wavelength_list);
  writeNpy(out + "index_list.npy", asDouble(grid.spectralIndexList()));
  writeNpy(out + "temperature.npy", atmosphere.temperature);
  writeNpy(out + "pressure.npy", atmosphere.pressure);
  writeNpy(out + "altitude.npy", atmosphere.altitude);
  writeNpy(out + "scale_height.npy", atmosphere.scale_height);
  if (!cloud_models.empty()) {
    // [layer][wavenumber], and for transmission the level extinction [level][wavenumber]
    const size_t N = grid.nbSpectralPoints();
    std::vector<double> tau((cs.nb_levels - 1) * N), ext;
    for (size_t i = 0; i < N; ++i)
      for (size_t l = 0; l + 1 < cs.nb_levels; ++l) tau[l * N + i] = opacity.cloud_optical_depths[i][l];
    writeNpy(out + "cloud_optical_depth.npy", tau, cs.nb_levels - 1);
    if (transmission) {
      ext.resize(cs.nb_levels * N);
      for (size_t i = 0; i < N; ++i)
        for (size_t l = 0; l < cs.nb_levels; ++l) ext[l * N + i] = cloud_extinction[i][l];
      writeNpy(out + "cloud_extinction.npy", ext, cs.nb_levels);
    }
  }
  {
    std::vector<double> nd;
    for (auto& l : atmosphere.number_densities) nd.insert(nd.end(), l.begin(), l.end());
    writeNpy(out + "number_densities.npy", nd, atmosphere.number_densities.size());
  }
  if (grid.nbSpectralPoints() * cs.nb_levels <= 4000000) {
    std::vector<double> k;
    for (auto& row : opacity.absorption_coeff) k.insert(k.end(), row.begin(), row.end());
    writeNpy(out + "absorption_coeff.npy", k, opacity.absorption_coeff.size());
    std::vector<double> s;
    for (auto& row : opacity.scattering_coeff) s.insert(s.end(), row.begin(), row.end());
    writeNpy(out + "scattering_coeff.npy", s, opacity.scattering_coeff.size());
  }
  writeNpy(out + "spectrum_flux.npy", spectrum_high_res_flux);
  writeNpy(out + "spectrum.npy", spectrum);
  for (size_t i = 0; i < observations.size(); ++i) {
    const std::string tag = std::to_string(i);
    std::vector<double> bins = spectrum_obs[i];
    writeNpy(out + "obs" + tag + "_bins.npy", bins);
    std::vector<double> edges;
    for (auto& e : observations[i].spectral_bands.edge_indices) { edges.push_back(e[0]); edges.push_back(e[1]); }
    writeNpy(out + "obs" + tag + "_edge_indices.npy", edges, observations[i].spectral_bands.edge_indices.size());
    auto& sb = observations[i].spectral_bands;
    writeNpy(out + "obs" + tag + "_index_range.npy", {double(sb.obs_index_range.first), double(sb.obs_index_range.second)});
    if (!sb.instrument_profile_sigma.empty()) {
      writeNpy(out + "obs" + tag + "_sigma.npy", sb.instrument_profile_sigma);
      std::vector<double> iv;
      for (auto& e : sb.convolution_quadrature_intervals) { iv.push_back(e[0]); iv.push_back(e[1]); }
      writeNpy(out + "obs" + tag + "_conv_intervals.npy", iv, sb.convolution_quadrature_intervals.size());
    }
    if (!observations[i].filter_response.empty()) {
      writeNpy(out + "obs" + tag + "_filter.npy", observations[i].filter_response);
      writeNpy(out + "obs" + tag + "_filter_norm.npy", {observations[i].filter_response_normalisation});
    }
    writeNpy(out + "obs" + tag + "_data.npy", observations[i].data);
    writeNpy(out + "obs" + tag + "_error.npy", observations[i].data_error);
    writeNpy(out + "obs" + tag + "_weight.npy", observations[i].likelihood_weight);
    writeNpy(out + "obs" + tag + "_ascending.npy", {double(observations[i].ascending_wavelengths)});
  }
  writeNpy(out + "loglike.npy", {log_like, double(neglect)});
  if (!timings.empty()) writeNpy(out + "timings.npy", timings);

  for (auto c : chemistry) delete c;
  std::printf("N=%zu loglike=%.17g neglect=%d\n", grid.nbSpectralPoints(), log_like, int(neglect));
  if (!timings.empty()) {
    std::sort(timings.begin(), timings.end());
    std::printf("median_eval_s=%.6g\n", timings[timings.size() / 2]);
  }
  return 0;
}
