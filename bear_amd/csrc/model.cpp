// Forward models and retrieval of the host side.
#include "model.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace bear_amd {

// opacity species and tables (TransportCoefficients ctor, transport_coeff.cpp:40-93)
TabulatedModel::TabulatedModel(const GlobalSettings& global, const ModelSettings& m, const SpectralGrid& grid,
                               const std::vector<ObservationSetup>& observations, DeviceEngine& engine,
                               bool transmission)
    : observations_(observations),
      engine_(engine),
      atmosphere_(m.nb_levels, m.bottom_pressure, m.top_pressure, m.temperature, m.chemistry) {
  std::vector<double> rayleigh, hminus;
  std::vector<TableSource> sources;
  uint32_t nb_rows = 0;
  for (auto& e : m.opacity) {
    OpacitySpecies s = makeOpacitySpecies(e.first, e.second, grid);
    if (s.hminus) {                                  // absorption of every model (h_m.cpp:30-54)
      if (hminus_) raise(kInvalidInput, "the H- continuum is listed twice");
      hminus_ = true;
      hminus_ids_[0] = s.species_id;
      hminus_ids_[1] = sid::H;
      hminus_ids_[2] = speciesId("e-");
      hminus = std::move(s.hminus_rows);
      continue;
    }
    if (s.rayleigh) {
      // scattering enters the transmission model only: emission's short
      // characteristics take the absorption coefficients alone
      // (emission.cpp:210-218, secondary_eclipse.cpp:200-209)
      if (!transmission) continue;
      if (rayleigh_species_ >= 0) raise(kInvalidInput, "one Rayleigh scatterer at most");
      rayleigh_species_ = s.species_id;
      rayleigh = s.rayleigh_cross_section;
      continue;
    }
    sources.push_back(scanSpeciesTable(s, global.cross_section_path, nb_rows));
    nb_rows += uint32_t(sources.back().nb_rows);
    species_.push_back(std::move(s));
  }
  // one species at a time: host memory holds one species' FP64 rows (and,
  // with reduced-precision storage, the device one species' FP64 staging)
  engine_.allocateTables(nb_rows, sources.size());
  std::vector<double> rows;
  for (size_t q = 0; q < sources.size(); ++q) {
    sampleSpeciesTable(sources[q], grid, rows);
    engine_.uploadSpeciesRows(int(q), sources[q].first_row, sources[q].nb_rows, rows.data());
  }
  rows.clear();
  rows.shrink_to_fit();
  engine_.uploadRayleigh(rayleigh);
  engine_.uploadHminus(hminus);
  for (OpacitySpecies& s : species_) s.preparePressures(atmosphere_.pressure);   // fixed pressure grid
  engine_.configure(m.nb_levels, species_.size(), transmission);
  // stencil groups: the species' table grids, at most two distinct ones
  StencilGroups groups;
  groups.grouped = 1;
  for (size_t s = 0; s < species_.size() && groups.grouped; ++s) {
    int g = -1;
    for (int k = 0; k < 2 && g < 0; ++k) {
      if (group_rep_[k] < 0) { group_rep_[k] = int(s); g = k; }
      else if (species_[s].sameGrid(species_[size_t(group_rep_[k])])) g = k;
    }
    if (g < 0) groups.grouped = 0;
    else groups.group1 |= uint32_t(g) << s;
  }
  if (!groups.grouped) groups.group1 = 0;
  engine_.setStencilGroups(groups);
  nb_chem_ = atmosphere_.nbChemistryParameters();
  nb_temp_ = atmosphere_.nbTemperatureParameters();
  // cloud modules (emission_init.cpp:72-89); their parameters follow the
  // temperature's (emission.cpp:123-153)
  clouds_ = m.clouds;
  if (clouds_.size() > size_t(kMaxClouds))
    raise(kInvalidInput, "at most " + std::to_string(kMaxClouds) + " cloud models");
  for (const auto& c : clouds_) nb_cloud_ += c.nbParameters();
  transmission_ = transmission;
  engine_.setCloudModels(int(clouds_.size()));
  for (auto& o : observations_) nb_modifier_ += size_t(o.modifier_parameters);
}

void TabulatedModel::stageLevels(const Atmosphere& atm, const double* mods, LevelState* lev, LevelSpecies* spc,
                                 LevelSpeciesAdjoint* adj, LevelSpecies32* spc32, double* shifts) const {
  const size_t L = atm.nb_levels, S = species_.size(), stride = engine_.speciesStride();
  const double log2_ten = 3.321928094887362347870319429489390175864831393;
  const size_t blocks_per_row = engine_.tableStride() / 64;     // row -> block offset of the device tables
  const StencilGroups groups = engine_.stencilGroups();
  for (size_t l = 0; l < L; ++l) {
    Stencil group_st[2] = {};
    const double t = atm.temperature[l];
    // opacity_species.cpp:443-446: scattering uses the scatterer's own density
    const double n_r = rayleigh_species_ >= 0 ? atm.nd(l, rayleigh_species_) : 0.0;
    lev[l] = LevelState{};
    lev[l].temperature = t;
    lev[l].inv_temperature = 1.0 / t;
    lev[l].altitude = atm.altitude[l];
    lev[l].rayleigh_density = n_r;
    // a layer's cloud optical depth as this level's term (cloud_model.h:67-88)
    if (l > 0) lev[l].cloud_scale = transmission_ ? 1.0 / (atm.altitude[l] - atm.altitude[l - 1]) : 1.0;
    if (hminus_)                                    // H-, H and electron densities
      hminusCoefficients(t, atm.nd(l, hminus_ids_[0]), atm.nd(l, hminus_ids_[1]), atm.nd(l, hminus_ids_[2]),
                         lev[l].hminus_c);
    for (int g = 0; g < 2; ++g) {
      if (group_rep_[g] < 0) continue;
      const Stencil st = species_[size_t(group_rep_[g])].bracketLevel(l, t);
      group_st[g] = st;
      const double ft = st.f_t, fl = st.f_p_lower, fu = st.f_p_upper;
      lev[l].group_w[g][0] = (1.0 - ft) * (1.0 - fl) * log2_ten;
      lev[l].group_w[g][1] = (1.0 - ft) * fl * log2_ten;
      lev[l].group_w[g][2] = ft * (1.0 - fu) * log2_ten;
      lev[l].group_w[g][3] = ft * fu * log2_ten;
      for (int k = 0; k < 4; ++k) lev[l].group_w32[g][k] = float(lev[l].group_w[g][k]);
      const double du = st.dft * log2_ten;
      lev[l].group_u[g][0] = -(1.0 - fl) * du;
      lev[l].group_u[g][1] = -fl * du;
      lev[l].group_u[g][2] = (1.0 - fu) * du;
      lev[l].group_u[g][3] = fu * du;
    }
    for (size_t s = 0; s < stride; ++s) {
      // padding slots: zero density, row 0 (loaded once at level 0)
      double n = 0.0;
      Stencil st{{0, 0, 0, 0}, 0.0, 0.0, 0.0, 0.0, {0, 0}, {0, 0, 0, 0}};
      if (s < S) {
        const OpacitySpecies& sp = species_[s];
        n = atm.nd(l, sp.species_id);                                 // opacity_species.cpp:407-412
        for (int q : sp.cia_partners) n *= atm.nd(l, q);
        if (groups.grouped) {
          // same grid as the group's representative: same bracket, own rows
          st = group_st[(groups.group1 >> s) & 1u];
          sp.takeRows(st);
        } else {
          st = sp.bracketLevel(l, t);                                // rows stay valid when n == 0
        }
      }
      uint32_t block[4];
      for (int k = 0; k < 4; ++k) block[k] = uint32_t(st.row[k] * blocks_per_row);
      // reload the rows only where they differ from the previous level's
      uint32_t changed = l == 0;
      if (l > 0) {
        const uint32_t* prev = spc ? spc[(l - 1) * stride + s].block : spc32[(l - 1) * stride + s].block;
        for (int k = 0; k < 4; ++k) changed |= prev[k] != block[k];
      }
      if (adj) {
        LevelSpeciesAdjoint& a = adj[l * stride + s];
        a = LevelSpeciesAdjoint{};
        const double du = st.dft * log2_ten;
        a.u[0] = -(1.0 - st.f_p_lower) * du;
        a.u[1] = -st.f_p_lower * du;
        a.u[2] = (1.0 - st.f_p_upper) * du;
        a.u[3] = st.f_p_upper * du;
      }
      if (spc) {
        LevelSpecies& d = spc[l * stride + s];
        const double ft = st.f_t, fl = st.f_p_lower, fu = st.f_p_upper;
        d.w[0] = (1.0 - ft) * (1.0 - fl) * log2_ten;
        d.w[1] = (1.0 - ft) * fl * log2_ten;
        d.w[2] = ft * (1.0 - fu) * log2_ten;
        d.w[3] = ft * fu * log2_ten;
        d.log2_n = n > 0 ? std::log2(n) : kZeroDensityLog2;
        if (s < 16) lev[l].log2_n[s] = d.log2_n;
        for (int k = 0; k < 4; ++k) d.block[k] = block[k];
        d.reload = changed;
        d.pad = 0;
      }
      if (spc32) {
        LevelSpecies32& d = spc32[l * stride + s];
        const double ft = st.f_t, fl = st.f_p_lower, fu = st.f_p_upper;
        d.w[0] = float((1.0 - ft) * (1.0 - fl) * log2_ten);
        d.w[1] = float((1.0 - ft) * fl * log2_ten);
        d.w[2] = float(ft * (1.0 - fu) * log2_ten);
        d.w[3] = float(ft * fu * log2_ten);
        d.log2_n = n > 0 ? float(std::log2(n)) : -std::numeric_limits<float>::infinity();
        if (s < 16) lev[l].log2_n32[s] = d.log2_n;
        for (int k = 0; k < 4; ++k) d.block[k] = block[k];
        d.reload = changed;
        d.pad[0] = d.pad[1] = 0;
      }
      lev[l].reload |= changed << s;
    }
  }
  size_t k = 0;                                                   // forward_model.cpp:184-204
  for (size_t o = 0; o < observations_.size(); ++o) {
    shifts[o] = observations_[o].modifier_parameters ? mods[k] : 0.0;
    k += size_t(observations_[o].modifier_parameters);
  }
}

void TabulatedModel::stageStructure(double g, double radius, bool variable_gravity, const double* p, const double* t,
                                    const std::vector<int>& species, const double* x) {
  atmosphere_.setStructure(g, radius, variable_gravity, p, t, species, x);
  for (OpacitySpecies& s : species_) s.preparePressures(atmosphere_.pressure);   // this call's pressures
  stageLevels(atmosphere_, nullptr, engine_.levels(), engine_.levelSpecies(), nullptr, nullptr, engine_.shifts());
}

bool TabulatedModel::stage(const double* p, BatchModel& scalars) {
  scalars = BatchModel{};
  const bool neglect = setup(p, atmosphere_, scalars);
  stageLevels(atmosphere_, modifiers(p), engine_.levels(), engine_.levelSpecies(), nullptr, nullptr,
              engine_.shifts());
  CloudSpec spec;
  if (!clouds_.empty()) stageClouds(atmosphere_, clouds(p), std::pow(10, p[0]), fixedScaleHeight(p), engine_.levels(), spec);
  engine_.setCloudSpec(spec);
  return neglect;
}

// GreyCloudModel::cloudPosition, both forms (grey_cloud_model.cpp:118-197):
// the levels bracketing the top (and bottom) pressure, or the bottom one
// scale height below the top; at least two layers thick.  Unsigned
// arithmetic as in the reference.
namespace {
void cloudPosition(const Atmosphere& atm, double top_pressure, double bottom_pressure, bool fixed_bottom,
                   double g, double scale_height_km, unsigned& top, unsigned& bottom) {
  const size_t L = atm.nb_levels;
  const std::vector<double>& p = atm.pressure;
  top = 0;
  bottom = 0;
  for (size_t i = 0; i + 1 < L; ++i)
    if ((p[i] > top_pressure && p[i + 1] < top_pressure) || p[i] == top_pressure) top = unsigned(i);
  if (!fixed_bottom) {
    for (size_t i = 0; i + 1 < L; ++i)
      if ((p[i] > bottom_pressure && p[i + 1] < bottom_pressure) || p[i] == bottom_pressure) bottom = unsigned(i);
    if (top_pressure <= p.back()) top = unsigned(L - 1);
    if (bottom_pressure >= p[0]) bottom = 0;
  } else {
    if (top_pressure < p.back()) top = unsigned(L - 1);
    const std::vector<double>& z = atm.altitude;
    double bottom_altitude = z[top] - atm.scaleHeight(top, g, scale_height_km);
    if (bottom_altitude < 0) bottom_altitude = 0;
    for (size_t i = 0; i + 1 < L; ++i)
      if ((z[i] < bottom_altitude && z[i + 1] > bottom_altitude) || z[i] == bottom_altitude) bottom = unsigned(i);
  }
  if (top == bottom) bottom -= 2;
  if (top < 2) {
    top = 2;
    bottom = 0;
  }
}
}  // namespace

// The cloud models' layers (grey_cloud_model.cpp:56-98,
// power_law_cloud_model.cpp:70-121, kh_cloud_model.cpp:69-133): tau / (top -
// bottom) on the layers bottom+1 .. top, in LevelState::cloud_w of the level
// above each layer, times the model's spectral factor, which the kernels form
// from spec.  A top at the uppermost level names a layer above the grid (the
// reference writes past its array there); that layer is left out.
void TabulatedModel::stageClouds(const Atmosphere& atm, const double* cp, double g, double scale_height_km,
                                 LevelState* lev, CloudSpec& spec) const {
  const size_t L = atm.nb_levels;
  const double two_pi = 2.0 * phys::pi;
  spec = CloudSpec{};
  spec.count = int(clouds_.size());
  for (size_t m = 0; m < clouds_.size(); ++m) {
    const CloudModelSpec& c = clouds_[m];
    const double* q = cp;
    cp += c.nbParameters();
    const size_t at_top = c.kind == kCloudGrey ? 1 : c.kind == kCloudKH ? 4 : 2;
    const double top_pressure = q[at_top];
    const double bottom_pressure = c.fixed_bottom ? 0.0 : top_pressure * q[at_top + 1];
    const double tau = q[0] < 0 ? 0.0 : q[0];
    unsigned top = 0, bottom = 0;
    cloudPosition(atm, top_pressure, bottom_pressure, c.fixed_bottom, g, scale_height_km, top, bottom);
    const double layer = tau / static_cast<double>(top - bottom);
    for (unsigned i = top; i > bottom; --i)
      if (i + 1 < L) lev[i + 1].cloud_w[m] = layer;
    spec.kind[m] = c.kind;
    if (c.kind == kCloudPowerLaw) {
      spec.c[m][0] = q[1];
      spec.c[m][1] = 1.0 / std::pow(c.reference_wavelength, q[1]);
      spec.c[m][2] = std::log(c.reference_wavelength);               // for the exponent's gradient
    } else if (c.kind == kCloudKH) {
      const double x_ref = two_pi * q[3] / c.reference_wavelength;
      spec.c[m][0] = q[1];
      spec.c[m][1] = q[2];
      spec.c[m][2] = two_pi * q[3];
      spec.c[m][3] = q[1] * std::pow(x_ref, -q[2]) + std::pow(x_ref, 0.2);
    }
  }
}

void TabulatedModel::stageBatch(const double* params, size_t batch, size_t stride, const double* inflation,
                                char* neglect) {
  const int threads = std::max(1, std::min(omp_get_max_threads(), int(batch)));
  while (atmosphere_pool_.size() < size_t(threads)) atmosphere_pool_.emplace_back(new Atmosphere(atmosphere_));
  const bool wide = engine_.tablePrecision() == TablePrecision::fp64;
  std::string failure;
  int failure_code = 0;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
  for (size_t b = 0; b < batch; ++b) {
    try {
      Atmosphere& atm = *atmosphere_pool_[omp_get_thread_num()];
      const double* p = params + b * stride;
      BatchModel& scalars = *engine_.batchModel(b);
      scalars = BatchModel{};
      neglect[b] = setup(p, atm, scalars);
      scalars.inflation = inflation[b];
      stageLevels(atm, modifiers(p), engine_.batchLevels(b), wide ? engine_.batchSpecies(b) : nullptr,
                  wide ? engine_.batchSpeciesAdjoint(b) : nullptr, wide ? nullptr : engine_.batchSpecies32(b),
                  engine_.batchShifts(b));
      if (!clouds_.empty())
        stageClouds(atm, clouds(p), std::pow(10, p[0]), fixedScaleHeight(p), engine_.batchLevels(b), scalars.clouds);
    } catch (const Error& e) {
#pragma omp critical(bear_stage_failure)
      if (failure.empty()) { failure = e.what(); failure_code = e.code; }
    } catch (const std::exception& e) {
#pragma omp critical(bear_stage_failure)
      if (failure.empty()) { failure = e.what(); failure_code = kInternal; }
    }
  }
  if (!failure.empty()) raise(failure_code, failure);
}

// The model's fixed data for the device staging (StagePlan): pressures,
// the temperature profile's basis per level, the chemistry modules, the
// active species, and per table grid its temperature rows, pressure picks
// and row ids, laid out as one blob.
bool TabulatedModel::prepareStaging(size_t nb_params, bool error_inflation) {
  if (staged_params_ == nb_params) return true;
  if (hminus_ || !clouds_.empty()) return false;
  const Atmosphere& atm = atmosphere_;
  const size_t L = atm.nb_levels, S = species_.size(), NS = engine_.speciesStride();
  StagePlan P;
  std::vector<unsigned char> blob;
  auto put = [&](const void* data, size_t bytes) {
    const size_t off = (blob.size() + 15) / 16 * 16;
    blob.resize(off + std::max<size_t>(bytes, 1), 0);
    if (bytes) std::memcpy(blob.data() + off, data, bytes);
    return off;
  };
  auto putv = [&](const auto& v) { return put(v.data(), v.size() * sizeof(v[0])); };
  P.L = int(L);
  P.S = int(S);
  P.NS = int(NS);
  P.kind = int(kind());
  P.isothermal = atm.isothermal();
  P.nb_temp = int(nb_temp_);
  P.nb_chem = int(nb_chem_);
  P.g0 = int(nb_general_);
  P.nb_params = int(nb_params);
  P.error_inflation = error_inflation;
  P.nb_obs = int(observations_.size());
  P.mods_offset = int(nb_general_ + nb_chem_ + nb_temp_ + nb_cloud_);
  P.alt_kind = altitudeKind();
  P.variable_gravity = variableGravity();
  const StencilGroups groups = engine_.stencilGroups();
  P.grouped = int(groups.grouped);
  P.group1 = groups.group1;
  P.rayleigh_id = rayleigh_species_;
  P.blocks_per_row = uint32_t(engine_.tableStride() / 64);
  P.na_id = sid::Na;
  P.k_id = sid::K;
  P.h2_id = sid::H2;
  P.he_id = sid::He;
  P.pressure = putv(atm.pressure);
  std::vector<double> log_ratio(L, 0.0);
  for (size_t l = 1; l < L; ++l) log_ratio[l] = std::log(atm.pressure[l] / atm.pressure[0]);
  P.log_ratio = putv(log_ratio);
  if (!atm.isothermal()) {
    const auto& basis = atm.temperatureBasis();
    size_t e = 1;
    for (const auto& b : basis) e = std::max(e, b.index.size());
    std::vector<int> bi(L * e, -1);
    std::vector<double> bw(L * e, 0.0);
    for (size_t l = 0; l < L; ++l)
      for (size_t k = 0; k < basis[l].index.size(); ++k) {
        bi[l * e + k] = int(basis[l].index[k]);
        bw[l * e + k] = basis[l].weight[k];
      }
    P.basis_entries = int(e);
    P.basis_index = putv(bi);
    P.basis_weight = putv(bw);
  }
  std::vector<StageModule> modules;
  std::vector<int> module_species;
  for (const auto& m : atm.chemistryModules()) {
    modules.push_back(StageModule{m.background, m.sodium_free, int(module_species.size()), int(m.species.size())});
    module_species.insert(module_species.end(), m.species.begin(), m.species.end());
  }
  P.nb_modules = int(modules.size());
  P.modules = putv(modules);
  P.module_species = putv(module_species);
  const std::vector<int>& active = atm.activeIds();
  P.nb_active = int(active.size());
  P.n_ids = int(atm.nbIds());
  std::vector<int> id_slot(atm.nbIds(), -1);
  std::vector<double> weight(atm.nbIds());
  for (size_t a = 0; a < active.size(); ++a) id_slot[size_t(active[a])] = int(a);
  for (size_t j = 0; j < atm.nbIds(); ++j) weight[j] = molecularWeight(int(j));
  P.active = putv(active);
  P.id_slot = putv(id_slot);
  P.mol_weight = putv(weight);
  // table grids: the stencil groups' representatives, or each species its own
  std::vector<int> grid_rep, species_grid(S, -1);
  if (groups.grouped) {
    for (int g = 0; g < 2; ++g)
      if (group_rep_[g] >= 0) {
        P.group_grid[g] = int(grid_rep.size());
        grid_rep.push_back(group_rep_[g]);
      }
    for (size_t s = 0; s < S; ++s) species_grid[s] = P.group_grid[(groups.group1 >> s) & 1u];
  } else {
    for (size_t s = 0; s < S; ++s) {
      species_grid[s] = int(grid_rep.size());
      grid_rep.push_back(int(s));
    }
  }
  std::vector<StageGrid> grids;
  std::vector<double> grid_t, row_temps;
  std::vector<int> tstart;
  std::vector<StagePick> picks;
  for (int r : grid_rep) {
    const OpacitySpecies& sp = species_[size_t(r)];
    const size_t nt = sp.by_temperature.size();
    StageGrid G{int(nt), int(grid_t.size()), int(picks.size()), int(tstart.size()), int(row_temps.size()), {0, 0, 0}};
    int start = 0;
    for (size_t i = 0; i < nt; ++i) {
      grid_t.push_back(sp.by_temperature[i].front().temperature);
      tstart.push_back(start);
      for (const TableRow& row : sp.by_temperature[i]) row_temps.push_back(row.temperature);
      start += int(sp.by_temperature[i].size());
    }
    for (const auto& pk : sp.pressurePicks()) picks.push_back(StagePick{pk.lower, pk.upper, pk.fraction});
    if (picks.size() != size_t(G.picks_offset) + L * nt) raise(kInternal, "pressure picks do not match the levels");
    grids.push_back(G);
  }
  std::vector<StageSpecies> species(S);
  std::vector<uint32_t> row_ids;
  for (size_t s = 0; s < S; ++s) {
    const OpacitySpecies& sp = species_[s];
    const int p0 = sp.cia_partners.size() > 0 ? sp.cia_partners[0] : -1;
    const int p1 = sp.cia_partners.size() > 1 ? sp.cia_partners[1] : -1;
    if (sp.cia_partners.size() > 2) return false;
    species[s] = StageSpecies{sp.species_id, p0, p1, species_grid[s], int(row_ids.size()), {0, 0, 0}};
    for (const auto& row : sp.by_temperature)
      for (const TableRow& r : row) row_ids.push_back(r.id);
  }
  P.grids = putv(grids);
  P.grid_t = putv(grid_t);
  P.grid_tstart = putv(tstart);
  P.row_temps = putv(row_temps);
  P.picks = putv(picks);
  P.species = putv(species);
  P.row_ids = putv(row_ids);
  std::vector<int> obs_mods;
  for (const auto& o : observations_) obs_mods.push_back(o.modifier_parameters);
  P.obs_mods = putv(obs_mods);
  if ((L * (5 + active.size()) + nb_temp_) * sizeof(double) > 60 * 1024) return false;   // the kernel's LDS
  engine_.uploadStagePlan(P, blob);
  staged_params_ = nb_params;
  return true;
}

std::vector<std::pair<uint32_t, uint32_t>> TabulatedModel::speciesRows() const {
  std::vector<std::pair<uint32_t, uint32_t>> out;
  for (const auto& s : species_) {
    uint32_t lo = std::numeric_limits<uint32_t>::max(), hi = 0;
    for (const auto& row : s.by_temperature)
      for (const auto& r : row) {
        lo = std::min(lo, r.id);
        hi = std::max(hi, r.id + 1);
      }
    out.push_back({lo, hi});
  }
  return out;
}

// ---------------------------------------------------------------- emission
EmissionModel::EmissionModel(const GlobalSettings& global, const ModelSettings& m, const SpectralGrid& grid,
                             const std::vector<ObservationSetup>& observations, DeviceEngine& engine)
    : TabulatedModel(global, m, grid, observations, engine, false) {
  if (!m.radiative_transfer.empty() && m.radiative_transfer != "scm")
    raise(kInvalidInput, "radiative transfer " + m.radiative_transfer + " is not available (only scm)");
}

// parameters and atmosphere (emission.cpp:108-178); scale (R_J / d)^2 f
bool EmissionModel::setup(const double* p, Atmosphere& atm, BatchModel& scalars) const {
  const double* chem = p + 3;
  const double* temp = chem + nb_chem_;
  const bool neglect = atm.compute(std::pow(10, p[0]), temp, chem);
  double scaling = phys::radius_jupiter / p[2];                   // emission.cpp:108-120
  scalars.scaling = scaling * scaling * p[1];
  return neglect;
}

// EmissionModel::calcModelGPU (emission.cpp:234-263)
bool EmissionModel::calcModelGPU(const double* p, double inflation, double* spectrum_dev,
                                 double* const* spectrum_obs_dev) {
  BatchModel scalars;
  const bool neglect = stage(p, scalars);
  engine_.evaluateEmission(scalars.scaling, inflation, spectrum_dev, spectrum_obs_dev);
  return neglect;
}

void EmissionModel::calcSpectrum(double g, double radius, double distance, const double* p, const double* t,
                                 const std::vector<int>& species, const double* x, bool variable_gravity,
                                 double* spectrum) {
  stageStructure(g, radius, variable_gravity, p, t, species, x);
  engine_.evaluateEmission(radius * radius / distance / distance, 0.0, nullptr, nullptr);   // emission.cpp:335-336
  engine_.copyFlux(spectrum);
}

// Chain rule from the device's level gradients to the parameters
// [log g, p1, p2, chemistry..., temperature..., modifiers...] of the emission
// and eclipse models: effective densities to species densities (CIA
// products), then the atmosphere's reverse sweep, g = 10^log g, the model's
// own p1, p2 (generalGradient) and the observations' shifts.
void TabulatedModel::chainGradient(const double* params, size_t batch, size_t stride,
                                   const std::vector<double>& level_grad, const std::vector<double>& bin_grad,
                                   const std::vector<double>& bin_terms, const char* neglect, double* gradient) {
  const size_t L = atmosphere_.nb_levels, NS = engine_.speciesStride(), K = levelSlots();
  const size_t ids = size_t(nbSpeciesIds());
  const size_t nb = engine_.nbBins();
  const int threads = std::max(1, std::min(omp_get_max_threads(), int(batch)));
  while (atmosphere_pool_.size() < size_t(threads)) atmosphere_pool_.emplace_back(new Atmosphere(atmosphere_));
  std::string failure;
  int failure_code = 0;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
  for (size_t b = 0; b < batch; ++b) {
    if (neglect[b]) continue;
    try {
      Atmosphere& atm = *atmosphere_pool_[omp_get_thread_num()];
      const double* p = params + b * stride;
      const double* chem = p + 3;
      const double* temp = chem + nb_chem_;
      const double g = std::pow(10, p[0]);
      atm.compute(g, temp, chem);
      const double* G = level_grad.data() + b * L * K;
      std::vector<double> t_bar(L), z_bar(L), n_bar(L * ids, 0.0);
      for (size_t l = 0; l < L; ++l) {
        t_bar[l] = G[l * K + NS] + G[l * K + NS + 3];
        z_bar[l] = G[l * K + NS + 1];
      }
      densityGradient(atm, G, n_bar, t_bar);
      double* out = gradient + b * stride;
      double g_bar = 0;
      atm.adjoint(g, temp, chem, t_bar, n_bar, z_bar, g_bar, out + 3 + nb_chem_, out + 3);
      out[0] = g_bar * g * std::log(10.0);
      generalGradient(p, G[NS + 2], bin_terms.empty() ? nullptr : bin_terms.data() + b * nb * 2, out);
      shiftGradient(bin_grad.data() + b * nb, out);
      if (!clouds_.empty()) cloudGradient(atm, p, g, G, out);
    } catch (const Error& e) {
#pragma omp critical(bear_gradient_failure)
      if (failure.empty()) { failure = e.what(); failure_code = e.code; }
    } catch (const std::exception& e) {
#pragma omp critical(bear_gradient_failure)
      if (failure.empty()) { failure = e.what(); failure_code = kInternal; }
    }
  }
  if (!failure.empty()) raise(failure_code, failure);
}

// The clouds' layer weights are w_m = tau / (top - bottom) on their layers
// (stageClouds; tau < 0 counts as 0): d logL / d tau = sum over layers of
// (d logL / d w_m) / (top - bottom).  The spectral parameters from the
// device's per-layer sums S_k (slots 4 m + k, rtAdjointKernel<true>) times
// w_m: power law, d / d a = sum w S_1; KHnongrey, f = c3 / u with
// c3 = q0 x_ref^-a0 + x_ref^0.2 (x_ref = 2 pi r / lambda_ref), its own
// derivatives times S_0 / c3 beside the device's u terms.  The cloud's
// position in pressure (or one scale height below its top) moves in whole
// levels: no gradient.
void TabulatedModel::cloudGradient(const Atmosphere& atm, const double* p, double g, const double* G,
                                   double* out) const {
  const size_t L = atm.nb_levels, K = levelSlots(), c0 = cloudSlot();
  const double* cp = clouds(p);
  std::vector<LevelState> lev(L), unit(L);
  CloudSpec spec;
  stageClouds(atm, cp, g, fixedScaleHeight(p), lev.data(), spec);
  std::vector<double> q(cp, cp + nb_cloud_);                       // unit optical depths: the weights' slopes
  size_t off = 0;
  for (const CloudModelSpec& c : clouds_) {
    q[off] = 1.0;
    off += c.nbParameters();
  }
  stageClouds(atm, q.data(), g, fixedScaleHeight(p), unit.data(), spec);
  double* gc = out + (cp - p);
  off = 0;
  for (size_t m = 0; m < clouds_.size(); ++m) {
    double d_tau = 0.0, s[4] = {0.0, 0.0, 0.0, 0.0};
    for (size_t l = 1; l < L; ++l) {
      const double* g = G + l * K + c0 + 4 * m;
      d_tau += g[0] * unit[l].cloud_w[m];
      for (int k = 0; k < 4; ++k) s[k] += g[k] * lev[l].cloud_w[m];
    }
    const double* q = cp + off;
    if (q[0] >= 0) gc[off] += d_tau;
    if (clouds_[m].kind == kCloudPowerLaw) {
      gc[off + 1] += s[1];
    } else if (clouds_[m].kind == kCloudKH) {
      const double q0 = q[1], a0 = q[2], r = q[3];
      const double x_ref = 2.0 * phys::pi * r / clouds_[m].reference_wavelength;
      const double xm = std::pow(x_ref, -a0), x2 = std::pow(x_ref, 0.2), c3 = q0 * xm + x2;
      gc[off + 1] += s[0] * xm / c3 - s[1];
      gc[off + 2] += -s[0] * q0 * xm * std::log(x_ref) / c3 + s[2];
      gc[off + 3] += (s[0] * (0.2 * x2 - a0 * q0 * xm) / c3 - s[3]) / r;
    }
    off += clouds_[m].nbParameters();
  }
}

void TabulatedModel::densityGradient(const Atmosphere& atm, const double* G, std::vector<double>& n_bar,
                                     std::vector<double>& t_bar) const {
  const size_t L = atm.nb_levels, NS = engine_.speciesStride(), K = levelSlots(), ids = size_t(nbSpeciesIds());
  // the H- continuum: slots NS+4 .. NS+10 hold d logL / d c_i for the
  // weights of hminusCoefficients, c_0 = n_H- and, for i >= 1,
  // c_i = a_i T n_e n_H with a_i = 1e-29 (5040 / T)^((i+1)/2) k
  if (hminus_)
    for (size_t l = 0; l < L; ++l) {
      if (atm.nd(l, hminus_ids_[0]) == 0) continue;                   // no H- term at this level
      const double* gh = G + l * K + NS + 4;
      const double t = atm.temperature[l], n_h = atm.nd(l, hminus_ids_[1]), n_e = atm.nd(l, hminus_ids_[2]);
      n_bar[l * ids + size_t(hminus_ids_[0])] += gh[0];
      for (int i = 1; i < 7; ++i) {
        const double a = 1e-29 * std::pow(5040.0 / t, (i + 1) / 2.0) * phys::boltzmann_k;
        n_bar[l * ids + size_t(hminus_ids_[1])] += gh[i] * a * t * n_e;
        n_bar[l * ids + size_t(hminus_ids_[2])] += gh[i] * a * t * n_h;
        t_bar[l] += gh[i] * a * n_e * n_h * (1.0 - (i + 1) / 2.0);
      }
    }
  for (size_t l = 0; l < L; ++l)
    for (size_t s = 0; s < species_.size(); ++s) {
      const OpacitySpecies& sp = species_[s];
      // the kernels give n dL/dn
      double n_eff = atm.nd(l, sp.species_id);
      for (int q : sp.cia_partners) n_eff *= atm.nd(l, q);
      const double gs = n_eff > 0 ? G[l * K + s] / n_eff : 0.0;
      double others = 1.0;
      for (int q : sp.cia_partners) others *= atm.nd(l, q);
      n_bar[l * ids + sp.species_id] += gs * others;
      for (size_t a = 0; a < sp.cia_partners.size(); ++a) {
        double rest = atm.nd(l, sp.species_id);
        for (size_t c = 0; c < sp.cia_partners.size(); ++c)
          if (c != a) rest *= atm.nd(l, sp.cia_partners[c]);
        n_bar[l * ids + sp.cia_partners[a]] += gs * rest;
      }
    }
}

// shift_const: one parameter per observation, added to its bins
// (forward_model.cpp:161-180)
void TabulatedModel::shiftGradient(const double* bin_grad, double* out) const {
  size_t k = nb_general_ + nb_chem_ + nb_temp_ + nb_cloud_, j = 0;
  for (const auto& o : observations_) {
    double sum = 0;
    for (size_t q = 0; q < o.nbPoints(); ++q) sum += bin_grad[j + q];
    if (o.modifier_parameters) out[k++] = sum;
    j += o.nbPoints();
  }
}

// emission: scale = (R_J / d)^2 f (emission.cpp:123-153)
void EmissionModel::generalGradient(const double* p, double scale_bar, const double*, double* out) const {
  const double base = (phys::radius_jupiter / p[2]) * (phys::radius_jupiter / p[2]);
  out[1] = scale_bar * base;
  out[2] = scale_bar * (-2.0 * base * p[1] / p[2]);
}

// ---------------------------------------------------------------- secondary eclipse
OccultationModel::OccultationModel(const GlobalSettings& global, const ModelSettings& m, const SpectralGrid& grid,
                                   const std::vector<ObservationSetup>& observations, DeviceEngine& engine)
    : TabulatedModel(global, m, grid, observations, engine, false) {
  if (!m.radiative_transfer.empty() && m.radiative_transfer != "scm")
    raise(kInvalidInput, "radiative transfer " + m.radiative_transfer + " is not available (only scm)");
  if (m.stellar_model != "blackbody")
    raise(kInvalidInput, "stellar model " + m.stellar_model + " is not available (only blackbody)");
}

// parameters [log g, Rp/Rs, T_*, ...] (secondary_eclipse.cpp:117-153)
bool OccultationModel::setup(const double* p, Atmosphere& atm, BatchModel& scalars) const {
  const double* chem = p + 3;
  const double* temp = chem + nb_chem_;
  const bool neglect = atm.compute(std::pow(10, p[0]), temp, chem);
  scalars.scaling = 1.0;
  scalars.star_inv_temperature = 1.0 / p[2];
  scalars.depth_scale = p[1] * p[1] * 1e6;
  return neglect;
}

// eclipse: depth = P / S (Rp/Rs)^2 1e6 per bin, S the binned blackbody star;
// bin_terms per bin {dL/ddepth (depth - shift), dL/ddepth d depth / dT_*}
void OccultationModel::generalGradient(const double* p, double, const double* bin_terms, double* out) const {
  double ratio_bar = 0, t_bar = 0;
  const size_t nb = engine_.nbBins();
  for (size_t j = 0; bin_terms && j < nb; ++j) {
    ratio_bar += bin_terms[2 * j];
    t_bar += bin_terms[2 * j + 1];
  }
  out[1] = ratio_bar * 2.0 / p[1];
  out[2] = t_bar;
}

void OccultationModel::setStellarSpectrum(const std::vector<double>& wavelength, const std::vector<double>& flux,
                                          const SpectralGrid& grid) {
  if (wavelength.size() < 2 || wavelength.size() != flux.size())
    raise(kInvalidInput, "the stellar spectrum needs matching wavelengths and fluxes (two at least)");
  std::vector<double> f = flux;
  for (size_t i = 0; i < f.size(); ++i) f[i] = f[i] * wavelength[i] * wavelength[i] / 10000.;   // per cm
  engine_.uploadStellarSpectrum(interpolateDescending(wavelength, f, grid.wavelength));
}

void OccultationModel::calcSpectrum(double g, double radius_ratio, const double* p, const double* t,
                                    const std::vector<int>& species, const double* x, double* spectrum) {
  stageStructure(g, 1.0, false, p, t, species, x);
  double* dev = engine_.spectrumBuffer();
  engine_.evaluateEclipseTabulatedStar(radius_ratio, dev);
  engine_.copySpectrum(spectrum, dev);
}

// OccultationModel::calcModelGPU (secondary_eclipse.cpp:262-355): planet
// emission at scale 1, blackbody star, depth = F_p / F_* (Rp/Rs)^2 1e6 per bin
bool OccultationModel::calcModelGPU(const double* p, double inflation, double* spectrum_dev,
                                    double* const* spectrum_obs_dev) {
  BatchModel scalars;
  const bool neglect = stage(p, scalars);
  engine_.evaluateEclipse(p[2], p[1], inflation, spectrum_dev, spectrum_obs_dev);
  return neglect;
}

// ---------------------------------------------------------------- transmission
TransmissionModel::TransmissionModel(const GlobalSettings& global, const ModelSettings& m, const SpectralGrid& grid,
                                     const std::vector<ObservationSetup>& observations, DeviceEngine& engine)
    : TabulatedModel(global, m, grid, observations, engine, true),
      fit_mmw_(m.fit_mean_molecular_weight),
      fit_scale_height_(m.fit_scale_height),
      variable_gravity_(m.use_variable_gravity) {
  if (fit_mmw_ && fit_scale_height_)
    raise(kInvalidInput, "fit either the mean molecular weight or the scale height, not both");
  nb_general_ = 3 + ((fit_mmw_ || fit_scale_height_) ? 1 : 0);    // transmission.cpp:69-72
}

// altitude variants of calcAtmosphereStructure (transmission.cpp:150-195)
AltitudeOptions TransmissionModel::altitudeOptions(const double* p) const {
  AltitudeOptions opt;
  opt.bottom_radius = p[1];
  opt.variable_gravity = variable_gravity_;
  if (fit_mmw_) opt.fixed_mean_molecular_weight = p[3];
  if (fit_scale_height_) opt.scale_height_km = p[3];
  return opt;
}

bool TransmissionModel::setup(const double* p, Atmosphere& atm, BatchModel& scalars) const {
  const double* chem = p + nb_general_;
  const double* temp = chem + nb_chem_;
  scalars.bottom_radius = p[1];
  scalars.star_radius = p[2];
  return atm.compute(std::pow(10, p[0]), temp, chem, altitudeOptions(p));
}

// Chain rule of the transmission model, per point, from the device's level
// gradients and dL/dchord (transmission_spectrum.cpp:32-154):
//   depth = (H + R^2) 1e6 / R_*^2      -> R_*: -2 / R_* sum dL/ddepth depth
//   H = sum_t c_t (1 - T_t), c_t = (R + z_t)(z_{t+1} - z_{t-1}), z_{-1} = z_0
//   chord[q][t] = D(q+1, t) - D(q, t), D(a, t) = sqrt((R + z_a)^2 - (R + z_t)^2)
//     (D(t, t) = 0 whatever the altitudes)
// then the atmosphere's reverse sweep with the model's altitude variant
// (radius through variable gravity, the fixed weight or scale height).
void TransmissionModel::gradientBatch(const double* params, size_t batch, size_t stride,
                                      const std::vector<double>& level_grad, const std::vector<double>& bin_grad,
                                      const std::vector<double>& chord_grad, const char* neglect, double* gradient) {
  const size_t L = atmosphere_.nb_levels, NS = engine_.speciesStride(), K = levelSlots(), P = L - 1;
  const size_t ids = size_t(nbSpeciesIds());
  const size_t nb = engine_.nbBins();
  const int threads = std::max(1, std::min(omp_get_max_threads(), int(batch)));
  while (atmosphere_pool_.size() < size_t(threads)) atmosphere_pool_.emplace_back(new Atmosphere(atmosphere_));
  std::string failure;
  int failure_code = 0;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
  for (size_t b = 0; b < batch; ++b) {
    if (neglect[b]) continue;
    try {
      Atmosphere& atm = *atmosphere_pool_[omp_get_thread_num()];
      const double* p = params + b * stride;
      const double* chem = p + nb_general_;
      const double* temp = chem + nb_chem_;
      const double g = std::pow(10, p[0]);
      const AltitudeOptions opt = altitudeOptions(p);
      atm.compute(g, temp, chem, opt);
      const double* G = level_grad.data() + b * L * K;
      std::vector<double> t_bar(L), z_bar(L, 0.0), n_bar(L * ids, 0.0);
      for (size_t l = 0; l < L; ++l) t_bar[l] = G[l * K + NS];
      densityGradient(atm, G, n_bar, t_bar);
      if (rayleigh_species_ >= 0)
        for (size_t l = 0; l < L; ++l) n_bar[l * ids + size_t(rayleigh_species_)] += G[l * K + NS + 3];
      const std::vector<double>& z = atm.altitude;
      const double r0 = p[1];
      double r_bar = 2.0 * r0 * G[(L - 1) * K + NS + 1];
      for (size_t t = 0; t < P; ++t) {
        const size_t lo = t > 0 ? t - 1 : 0;
        const double e1 = G[t * K + NS + 1], w = z[t + 1] - z[lo], rt = r0 + z[t];
        r_bar += e1 * w;
        z_bar[t] += e1 * w;
        z_bar[t + 1] += e1 * rt;
        z_bar[lo] -= e1 * rt;
      }
      const double* C = chord_grad.data() + b * P * P;
      for (size_t t = 0; t < P; ++t) {
        const double bt = r0 + z[t];
        for (size_t a = t + 1; a < L; ++a) {
          const double ar = r0 + z[a];
          const double d = std::sqrt(ar * ar - bt * bt);
          const double d_bar = C[(a - 1) * P + t] - (a < P ? C[a * P + t] : 0.0);
          if (!(d > 0)) continue;
          z_bar[a] += d_bar * ar / d;
          z_bar[t] -= d_bar * bt / d;
          r_bar += d_bar * (ar - bt) / d;
        }
      }
      double* out = gradient + b * stride;
      double g_bar = 0, option_bar = 0;
      atm.adjoint(g, temp, chem, t_bar, n_bar, z_bar, g_bar, out + nb_general_ + nb_chem_, out + nb_general_, opt,
                  &r_bar, &option_bar);
      out[0] = g_bar * g * std::log(10.0);
      out[1] = r_bar;
      out[2] = -2.0 / p[2] * G[(L - 1) * K + NS + 2];
      if (nb_general_ > 3) out[3] = option_bar;
      shiftGradient(bin_grad.data() + b * nb, out);
    } catch (const Error& e) {
#pragma omp critical(bear_gradient_failure)
      if (failure.empty()) { failure = e.what(); failure_code = e.code; }
    } catch (const std::exception& e) {
#pragma omp critical(bear_gradient_failure)
      if (failure.empty()) { failure = e.what(); failure_code = kInternal; }
    }
  }
  if (!failure.empty()) raise(failure_code, failure);
}

// chord lengths between successive levels for every tangent level
// (transmission_spectrum.cpp:117-118, 145-154), transposed: chord[i][t] =
// sqrt((R+z_{i+1})^2 - (R+z_t)^2) - sqrt((R+z_i)^2 - (R+z_t)^2), i >= t
// (the entries with i < t and the padding stay zero)
void TransmissionModel::stageChords(double r0) {
  const size_t L = atmosphere_.nb_levels;
  const size_t tangents = size_t(chordTangents(int(L)));
  const std::vector<double>& z = atmosphere_.altitude;
  double* chord = engine_.pathLengths();
  for (size_t t = 0; t + 1 < L; ++t) {
    const double b = r0 + z[t];
    double previous = std::sqrt(b * b - b * b);
    for (size_t i = t; i + 1 < L; ++i) {
      const double a = r0 + z[i + 1];
      const double next = std::sqrt(a * a - b * b);
      chord[i * tangents + t] = next - previous;
      previous = next;
    }
  }
}

// TransmissionModel::calcModelGPU (transmission.cpp:257-321)
bool TransmissionModel::calcModelGPU(const double* p, double inflation, double* spectrum_dev,
                                     double* const* spectrum_obs_dev) {
  BatchModel scalars;
  const bool neglect = stage(p, scalars);
  stageChords(p[1]);
  engine_.evaluateTransmission(p[1], p[2], inflation, spectrum_dev, spectrum_obs_dev);
  return neglect;
}

void TransmissionModel::calcSpectrum(double g, double planet_radius, double radius_ratio, const double* p,
                                     const double* t, const std::vector<int>& species, const double* x,
                                     bool variable_gravity, double* spectrum) {
  stageStructure(g, planet_radius, variable_gravity, p, t, species, x);
  stageChords(planet_radius);
  double* dev = engine_.spectrumBuffer();
  engine_.evaluateTransmission(planet_radius, planet_radius / radius_ratio, 0.0, dev, nullptr);   // transmission.cpp:391-392
  engine_.copySpectrum(spectrum, dev);
}

// ---------------------------------------------------------------- retrieval
// Set-up: observations, grid sampling, per-observation operators, forward
// model (retrieval.cpp:70-104, retrieval_load_observations.cpp:38-92,
// select_forward_model.cpp:38-164)
Retrieval::Retrieval(const GlobalSettings& global, const std::string& model_type, const ModelSettings& model,
                     const std::vector<ObservationInputView>& inputs, const std::vector<std::string>& files,
                     const std::vector<std::string>& modifiers)
    : global_(global) {
  const bool emission = model_type == "emission";
  const bool eclipse = model_type == "secondary_eclipse" || model_type == "se";
  const bool transmission = model_type == "transmission";
  if (!emission && !eclipse && !transmission)
    raise(kInvalidInput, "forward model " + model_type + " is not available");
  std::string folder = global_.retrieval_folder;
  if (!folder.empty() && folder.back() != '/') folder += '/';
  global_.retrieval_folder = folder;
  grid_.discretisation = global_.discretisation;
  grid_.resolution = global_.resolution;
  if (global_.discretisation < 0 || global_.discretisation > 2) raise(kInvalidInput, "unknown spectral discretisation");
  grid_.loadFullList(global_.wavenumber_file);
  for (auto& in : inputs) observations_.push_back(observationFromInput(in));
  for (size_t i = 0; i < files.size(); ++i)
    observations_.push_back(observationFromFile(folder + files[i], i < modifiers.size() ? modifiers[i] : "", folder));
  grid_.sample(observations_);
  for (auto& o : observations_) {
    o.setupOnGrid(grid_);
    o.foldOperator(grid_, !transmission);                         // transmission bins are not fluxes
  }
  engine_.reset(new DeviceEngine(global_.device));
  engine_->setTableStorage(global_.table_storage);
  engine_->uploadGrid(grid_.wavenumber, grid_.wavelength);
  engine_->uploadObservations(observations_);
  if (emission) model_.reset(new EmissionModel(global_, model, grid_, observations_, *engine_));
  if (eclipse) model_.reset(new OccultationModel(global_, model, grid_, observations_, *engine_));
  if (transmission) model_.reset(new TransmissionModel(global_, model, grid_, observations_, *engine_));
}

void Retrieval::checkSize(size_t n) const {
  if (n < nbParameters())
    raise(kInvalidInput, "expected " + std::to_string(nbParameters()) + " parameters, got " + std::to_string(n));
}

double Retrieval::errorInflation(const double* p, size_t n) const {
  return global_.error_inflation ? std::pow(10, p[n - 1]) : 0.0;   // retrieval.cpp:333-336
}

// Retrieval::logLikelihoodGPU (retrieval.cpp:374-420)
double Retrieval::computeLikelihood(const double* p, size_t n) {
  checkSize(n);
  if (engine_->tableStorage() != TablePrecision::fp64) {   // reduced-precision tables: a batch of one
    double ll = 0;
    computeLikelihoodBatch(p, 1, n, &ll);
    return ll;
  }
  const bool neglect = model_->calcModelGPU(p, errorInflation(p, n), nullptr, nullptr);
  const double ll = engine_->logLikelihood();
  return neglect ? -1e30 : ll;
}

bool Retrieval::calcModelGPU(const double* p, size_t n, double* spectrum_dev, double* const* obs_dev) {
  checkSize(n);
  const bool neglect = model_->calcModelGPU(p, errorInflation(p, n), spectrum_dev, obs_dev);
  engine_->synchronize();
  return neglect;
}

bool Retrieval::calcModelGPUNoAlloc(const double* p, size_t n, double* spectrum_dev, double* const* obs_dev) {
  checkSize(n);
  engine_->useScratch(true);
  bool neglect = false;
  try {
    neglect = model_->calcModelGPU(p, errorInflation(p, n), spectrum_dev, obs_dev);
    engine_->synchronize();
  } catch (...) {
    engine_->useScratch(false);
    throw;
  }
  engine_->useScratch(false);
  return neglect;
}

std::vector<char> Retrieval::stageBatch(const double* params, size_t batch, size_t n) {
  checkSize(n);
  engine_->reserveBatch(batch);
  std::vector<char> neglect(batch);
  std::vector<double> inflation(batch);
  for (size_t b = 0; b < batch; ++b) inflation[b] = errorInflation(params + b * n, n);
  model_->stageBatch(params, batch, n, inflation.data(), neglect.data());
  return neglect;
}

void Retrieval::evaluateStaged(const double* params, size_t batch, size_t n, std::vector<char>& neglect) {
  checkSize(n);
  if (device_staging_ && model_->prepareStaging(n, global_.error_inflation)) {
    engine_->reserveBatch(batch);
    engine_->reserveStage(batch);
    std::memcpy(engine_->stagedParameters(), params, batch * n * sizeof(double));
    engine_->evaluateBatch(model_->kind(), batch, true);
    neglect.resize(batch);
    for (size_t b = 0; b < batch; ++b) neglect[b] = engine_->stagedNeglect()[b] != 0;
    return;
  }
  neglect = stageBatch(params, batch, n);
  engine_->evaluateBatch(model_->kind(), batch);
}

void Retrieval::computeLikelihoodBatch(const double* params, size_t batch, size_t n, double* ll) {
  if (batch == 0) return;
  std::vector<char> neglect;
  evaluateStaged(params, batch, n, neglect);
  const double* v = engine_->batchLogLikelihood();
  for (size_t b = 0; b < batch; ++b) ll[b] = neglect[b] ? -1e30 : v[b];
}

void Retrieval::computeModelBatch(const double* params, size_t batch, size_t n, double* spectrum,
                                  double* spectrum_obs, int* neglect_out) {
  if (batch == 0) return;
  std::vector<char> neglect;
  evaluateStaged(params, batch, n, neglect);
  std::vector<double> bins;
  engine_->copyBatchBins(batch, bins);
  const size_t nb = engine_->nbBins();
  for (size_t b = 0; b < batch; ++b) {
    size_t k = 0;
    for (auto& o : observations_) {                               // input order (forward_model.cpp:119-123)
      const size_t m = o.nbPoints();
      const double* src = bins.data() + b * nb + k;
      double* dst = spectrum_obs + b * nb + k;
      if (o.ascending) std::reverse_copy(src, src + m, dst);
      else std::copy(src, src + m, dst);
      k += m;
    }
    if (neglect_out) neglect_out[b] = neglect[b];
  }
  if (spectrum) engine_->copyBatchSpectrum(model_->kind(), batch, spectrum);
}

void Retrieval::setTablePrecision(TablePrecision p) { engine_->setTablePrecision(p, model_->speciesRows()); }

// Value and gradient of the log-likelihood (retrieval.cpp:314-370) for a
// batch: the forward batch, the device adjoint, the model's chain rule, and
// the error-inflation term d/de sum_j w_j (-ln(2 pi s_j^2) - r_j^2 / s_j^2) / 2
// with s_j^2 = sigma_j^2 + 10^e.
void Retrieval::gradientBatch(const double* params, size_t batch, size_t n, double* ll, double* gradient) {
  if (batch == 0) return;
  if (engine_->tablePrecision() != TablePrecision::fp64)
    raise(kInvalidInput, "the gradient needs the FP64 tables (table precision fp64)");
  // the records staged on the device where the model allows it; the chain
  // rule below computes each point's atmosphere on the host either way
  std::vector<char> neglect;
  std::vector<double> level_grad, bin_grad, bins, bin_terms;
  checkSize(n);
  if (device_staging_ && model_->prepareStaging(n, global_.error_inflation)) {
    engine_->reserveBatch(batch);
    engine_->reserveStage(batch);
    std::memcpy(engine_->stagedParameters(), params, batch * n * sizeof(double));
    engine_->evaluateAdjointBatch(model_->kind(), batch, level_grad, bin_grad, bins, bin_terms, true);
    neglect.resize(batch);
    for (size_t b = 0; b < batch; ++b) neglect[b] = engine_->stagedNeglect()[b] != 0;
  } else {
    neglect = stageBatch(params, batch, n);
    engine_->evaluateAdjointBatch(model_->kind(), batch, level_grad, bin_grad, bins, bin_terms);
  }
  const double* v = engine_->batchLogLikelihood();
  for (size_t b = 0; b < batch; ++b) ll[b] = neglect[b] ? -1e30 : v[b];
  std::fill(gradient, gradient + batch * n, 0.0);
  model_->gradientBatch(params, batch, n, level_grad, bin_grad, bin_terms, neglect.data(), gradient);
  if (!global_.error_inflation) return;
  const size_t nb = engine_->nbBins();
  for (size_t b = 0; b < batch; ++b) {
    if (neglect[b]) continue;
    const double inflation = errorInflation(params + b * n, n);
    double sum = 0;
    size_t j = 0;
    for (const auto& o : observations_)
      for (size_t q = 0; q < o.nbPoints(); ++q, ++j) {
        const double e2 = o.error[q] * o.error[q] + inflation;
        const double r = o.data[q] - bins[b * nb + j];
        sum += o.weight[q] * (-0.5 / e2 + 0.5 * r * r / (e2 * e2));
      }
    gradient[b * n + n - 1] = sum * inflation * std::log(10.0);
  }
}

bool Retrieval::computeModel(const double* p, size_t n, bool high_res, double* spectrum, double* spectrum_obs) {
  checkSize(n);
  if (engine_->tableStorage() != TablePrecision::fp64) {   // reduced-precision tables: a batch of one
    int neglect = 0;
    computeModelBatch(p, 1, n, high_res ? spectrum : nullptr, spectrum_obs, &neglect);
    return neglect != 0;
  }
  double* spec_dev = high_res ? engine_->spectrumBuffer() : nullptr;
  const bool neglect = model_->calcModelGPU(p, errorInflation(p, n), spec_dev, nullptr);
  std::vector<double> bins;
  engine_->copyBins(bins);
  if (high_res && spectrum) engine_->copySpectrum(spectrum, spec_dev);
  // per observation, back in input order (forward_model.cpp:119-123)
  size_t k = 0;
  for (auto& o : observations_) {
    const size_t nb = o.nbPoints();
    if (o.ascending) std::reverse_copy(bins.begin() + k, bins.begin() + k + nb, spectrum_obs + k);
    else std::copy(bins.begin() + k, bins.begin() + k + nb, spectrum_obs + k);
    k += nb;
  }
  return neglect;
}

}  // namespace bear_amd

namespace bear_amd {

DirectModel::DirectModel(const GlobalSettings& global, const std::string& model_type, double wavelength_min,
                         double wavelength_max, size_t nb_levels,
                         const std::vector<std::pair<std::string, std::string>>& opacity,
                         const std::vector<double>& stellar_wavelength, const std::vector<double>& stellar_flux)
    : global_(global), type_(model_type), nb_levels_(nb_levels) {
  const bool emission = model_type == "emission";
  const bool eclipse = model_type == "secondary_eclipse" || model_type == "se";
  const bool transmission = model_type == "transmission";
  if (!emission && !eclipse && !transmission) raise(kInvalidInput, "forward model " + model_type + " is not available");
  if (global_.discretisation < 0 || global_.discretisation > 2) raise(kInvalidInput, "unknown spectral discretisation");
  if (global_.table_storage != TablePrecision::fp64)
    raise(kInvalidInput, "the direct models keep FP64 tables (table_storage fp64)");
  grid_.discretisation = global_.discretisation;
  grid_.resolution = global_.resolution;
  grid_.loadFullList(global_.wavenumber_file);
  grid_.sampleRange(wavelength_min, wavelength_max);
  ModelSettings m;
  m.nb_levels = nb_levels;
  m.bottom_pressure = 10.0;
  m.top_pressure = 1e-5;
  m.temperature = "const";
  m.radiative_transfer = "scm";
  m.opacity = opacity;
  engine_.reset(new DeviceEngine(global_.device));
  engine_->uploadGrid(grid_.wavenumber, grid_.wavelength);
  engine_->uploadObservations(observations_);
  if (emission) {
    emission_ = new EmissionModel(global_, m, grid_, observations_, *engine_);
    model_.reset(emission_);
  } else if (transmission) {
    transmission_ = new TransmissionModel(global_, m, grid_, observations_, *engine_);
    model_.reset(transmission_);
  } else {
    occultation_ = new OccultationModel(global_, m, grid_, observations_, *engine_);
    model_.reset(occultation_);
    occultation_->setStellarSpectrum(stellar_wavelength, stellar_flux, grid_);
  }
}

}  // namespace bear_amd
