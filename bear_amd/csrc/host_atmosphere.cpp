// Atmosphere of the host side: pressure grid, temperature profile,
// chemistry and hydrostatic altitude, evaluated per model in FP64.
#include <cmath>
#include <sstream>

#include "host.h"

namespace bear_amd {

// ---------------------------------------------------------------- piecewise polynomial
// Gauss-Lobatto nodes on [-1, 1] per element order (additional/piecewise_poly.cpp:35-43;
// orders above 5 are rejected: the reference's order-6 list is malformed)
PiecewisePolynomial::PiecewisePolynomial(size_t nb_elements, size_t order, double bottom, double top) {
  static const std::vector<std::vector<double>> nodes = {
    {0.0}, {-1.0, 1.0}, {-1.0, 0.0, 1.0}, {-1.0, -0.447214, 0.447214, 1.0},
    {-1.0, -0.654654, 0.654654, 1.0}, {-1.0, -0.654654, 0.0, 0.654654, 1.0}};
  if (order > 5) raise(kInvalidInput, "piecewise polynomials above order 5 are not available");
  if (nb_elements < 1) nb_elements = 1;
  if (order < 1) order = 1;
  // element edges from log10(p_bottom) downwards (piecewise_poly.cpp:89-132)
  const double x0 = std::log10(bottom), x1 = std::log10(top);
  const double width = (x0 - x1) / nb_elements;
  std::vector<double> edge(nb_elements + 1);
  edge.front() = x0;
  edge.back() = x1;
  for (size_t i = 1; i < nb_elements; ++i) edge[i] = edge[i - 1] - width;
  for (size_t i = 0; i < nb_elements; ++i) {
    Element e;
    e.ref = nodes[order];
    e.vertex.resize(e.ref.size());
    e.value.assign(e.ref.size(), 0.0);
    for (size_t j = 0; j < e.ref.size(); ++j) e.vertex[j] = edge[i] + (1.0 + e.ref[j]) / 2.0 * (edge[i + 1] - edge[i]);
    elements_.push_back(e);
  }
  for (auto& e : elements_)
    for (size_t j = 0; j + 1 < e.vertex.size(); ++j) vertices_.push_back(e.vertex[j]);
  vertices_.push_back(elements_.back().vertex.back());
  values_.assign(vertices_.size(), 0.0);
}

void PiecewisePolynomial::setValues(const std::vector<double>& v) {   // :149-173, shared element ends
  values_ = v;
  size_t g = 0;
  for (auto& e : elements_) {
    for (auto& x : e.value) x = values_[g++];
    --g;
  }
}

// :176-200 and the element's Lagrange sum (:210-229); outside the domain 0
double PiecewisePolynomial::value(double x) const {
  if (x > vertices_.front() || x < vertices_.back()) return 0.0;
  for (size_t i = 0; i < vertices_.size(); ++i) if (vertices_[i] == x) return values_[i];
  size_t k = 0;
  for (size_t i = 0; i < elements_.size(); ++i)
    if (elements_[i].vertex.front() > x && elements_[i].vertex.back() < x) { k = i; break; }
  const Element& e = elements_[k];
  const double r = 2.0 * (x - e.vertex.front()) / (e.vertex.back() - e.vertex.front()) - 1.0;
  double y = 0;
  for (size_t i = 0; i < e.ref.size(); ++i) {
    double l = 1;
    for (size_t j = 0; j < e.ref.size(); ++j)
      if (i != j) l *= (r - e.ref[j]) / (e.ref[i] - e.ref[j]);
    y += e.value[i] * l;
  }
  return y;
}

PiecewisePolynomial::Basis PiecewisePolynomial::basis(double x) const {
  Basis b;
  if (x > vertices_.front() || x < vertices_.back()) return b;                  // outside: 0
  for (size_t i = 0; i < vertices_.size(); ++i)
    if (vertices_[i] == x) {
      b.index.push_back(i);
      b.weight.push_back(1.0);
      return b;
    }
  size_t k = 0;
  for (size_t i = 0; i < elements_.size(); ++i)
    if (elements_[i].vertex.front() > x && elements_[i].vertex.back() < x) { k = i; break; }
  const Element& e = elements_[k];
  size_t first = 0;                                  // the element's first degree of freedom (shared ends)
  for (size_t i = 0; i < k; ++i) first += elements_[i].value.size() - 1;
  const double r = 2.0 * (x - e.vertex.front()) / (e.vertex.back() - e.vertex.front()) - 1.0;
  for (size_t i = 0; i < e.ref.size(); ++i) {
    double l = 1;
    for (size_t j = 0; j < e.ref.size(); ++j)
      if (i != j) l *= (r - e.ref[j]) / (e.ref[i] - e.ref[j]);
    b.index.push_back(first + i);
    b.weight.push_back(l);
  }
  return b;
}

// ---------------------------------------------------------------- atmosphere
Atmosphere::Atmosphere(size_t L, double bottom, double top, const std::string& temperature_desc,
                       const std::string& chemistry_desc)
    : nb_levels(L), n_ids_(size_t(nbSpeciesIds())) {
  if (L < 2) raise(kInvalidInput, "the atmosphere needs at least two levels");
  // log-equidistant pressures from the bottom (atmosphere.cpp:365-383)
  pressure.assign(L, 0.0);
  pressure.front() = std::log10(bottom);
  const double step = (std::log10(bottom) - std::log10(top)) / (L - 1.0);
  for (size_t i = 1; i < L - 1; ++i) pressure[i] = pressure[i - 1] - step;
  for (size_t i = 0; i < L - 1; ++i) pressure[i] = std::pow(10.0, pressure[i]);
  pressure.back() = top;

  std::istringstream t(temperature_desc);
  std::string kind;
  t >> kind;
  if (kind == "poly") {                              // temperature/select_temperature_profile.h:86-104
    size_t nb_el = 0, order = 0;
    if (!(t >> nb_el >> order)) raise(kInvalidInput, "the poly temperature profile needs two parameters");
    poly_.reset(new PiecewisePolynomial(nb_el, order, bottom, top));
    nb_temperature_param_ = nb_el * order + 1;
    for (size_t l = 0; l < L; ++l) temperature_basis_.push_back(poly_->basis(std::log10(pressure[l])));
    if (poly_->nbDof() != nb_temperature_param_)
      raise(kInvalidInput, "inconsistent piecewise-polynomial degrees of freedom");
  } else if (kind == "const") {
    isothermal_ = true;
    nb_temperature_param_ = 1;
  } else {
    raise(kInvalidInput, "temperature profile " + kind + " is not available");
  }

  std::istringstream c(chemistry_desc);
  std::string line;
  while (std::getline(c, line)) {
    std::istringstream in(line);
    std::string model, sp;
    if (!(in >> model)) continue;
    ChemistryModule m;
    if (model == "iso" || model == "isoprofile") {   // isoprofile_chemistry.cpp:33-80
      while (in >> sp) m.species.push_back(speciesId(sp));
      for (int s : m.species) if (s == sid::Na) m.sodium_free = true;
      nb_chemistry_param_ += m.species.size();
    } else if (model == "bg" || model == "background") {
      in >> sp;
      if (sp != "H2He") raise(kInvalidInput, "background chemistry " + sp + " is not available");
      m.background = true;
    } else {
      raise(kInvalidInput, "chemistry model " + model + " is not available");
    }
    chemistry_.push_back(m);
  }
  std::vector<char> can_be_set(n_ids_, 0);
  for (const ChemistryModule& m : chemistry_) {
    for (int s : m.species) can_be_set[size_t(s)] = 1;
    if (!m.background && !m.sodium_free) can_be_set[sid::Na] = 1;
    if (m.background) can_be_set[sid::H2] = can_be_set[sid::He] = 1;
  }
  for (size_t j = 1; j < n_ids_; ++j)
    if (can_be_set[j]) active_ids_.push_back(int(j));
  temperature.assign(L, 0.0);
  altitude.assign(L, 0.0);
  mean_molecular_weight.assign(L, 0.0);
  number_density.assign(L * n_ids_, 0.0);
}

Atmosphere::Atmosphere(const Atmosphere& o)
    : nb_levels(o.nb_levels),
      pressure(o.pressure),
      temperature(o.temperature),
      altitude(o.altitude),
      mean_molecular_weight(o.mean_molecular_weight),
      number_density(o.number_density),
      n_ids_(o.n_ids_),
      poly_(o.poly_ ? new PiecewisePolynomial(*o.poly_) : nullptr),
      isothermal_(o.isothermal_),
      nb_temperature_param_(o.nb_temperature_param_),
      nb_chemistry_param_(o.nb_chemistry_param_),
      chemistry_(o.chemistry_),
      active_ids_(o.active_ids_),
      temperature_basis_(o.temperature_basis_) {}

bool Atmosphere::compute(double g, const double* tp, const double* cp, const AltitudeOptions& opt) {
  const size_t L = nb_levels;
  bool neglect = false;
  // temperature (piecewise_poly_temperature.cpp:49-84; constant_temperature.h:34-46)
  if (isothermal_) {
    temperature.assign(L, tp[0]);
  } else {
    std::vector<double> dof(nb_temperature_param_);
    dof[0] = tp[0];
    for (size_t i = 1; i < dof.size(); ++i) dof[i] = dof[i - 1] * tp[i];
    for (size_t i = 0; i < L; ++i) {
      const PiecewisePolynomial::Basis& b = temperature_basis_[i];
      double y = 0;                                  // as PiecewisePolynomial::value, weights precomputed
      for (size_t k = 0; k < b.index.size(); ++k) y += dof[b.index[k]] * b.weight[k];
      temperature[i] = y;
    }
  }
  // the polynomial profile floors at 50 K and flags the model
  // (piecewise_poly_temperature.cpp:77-81); the isothermal one does not
  clamped_.assign(L, 0);
  bg_clamped_ = false;
  if (!isothermal_)
    for (size_t l = 0; l < L; ++l)
      if (temperature[l] < 50) { temperature[l] = 50; clamped_[l] = 1; neglect = true; }

  // chemistry modules in order (isoprofile_chemistry.cpp:87-119,
  // background_chemistry.cpp:72-137, chemistry.h:47-76)
  std::fill(number_density.begin(), number_density.end(), 0.0);
  auto n = [&](size_t l, int id) -> double& { return number_density[l * n_ids_ + id]; };
  // the species sums run over active_ids_ (and the total at j = 0 for the
  // mean weight): every other term is an exact zero
  const double w_total = molecularWeight(sid::total);
  auto meanWeight = [&]() {
    for (size_t l = 0; l < L; ++l) {
      double mu = n(l, sid::total) / n(l, sid::total) * w_total;
      for (int j : active_ids_) mu += n(l, j) / n(l, sid::total) * molecularWeight(j);
      mean_molecular_weight[l] = mu;
    }
  };
  size_t k = 0;
  for (auto& m : chemistry_) {
    if (!m.background) {
      for (size_t l = 0; l < L; ++l) {
        n(l, sid::total) = pressure[l] * 1.e6 / phys::boltzmann_k / temperature[l];
        for (size_t j = 0; j < m.species.size(); ++j) n(l, m.species[j]) = n(l, sid::total) * cp[k + j];
        if (!m.sodium_free) n(l, sid::Na) = n(l, sid::K) * 16.2181;
      }
      k += m.species.size();
      meanWeight();
      for (size_t l = 0; l < L; ++l) {
        double sum = 0;
        for (int j : active_ids_) sum += n(l, j) / n(l, sid::total);
        if (sum > 1 + 1e-10) { neglect = true; break; }
      }
    } else {
      const double h2 = 0.5, he = 0.085114, h2he = h2 + he;
      for (size_t l = 0; l < L; ++l) n(l, sid::total) = pressure[l] * 1.e6 / phys::boltzmann_k / temperature[l];
      for (size_t l = 0; l < L; ++l) {
        double bg = 1.0;
        for (int j : active_ids_)
          if (j != sid::H2 && j != sid::He) bg -= n(l, j) / n(l, sid::total);
        if (bg < 0) { neglect = true; bg = 0; bg_clamped_ = true; }
        n(l, sid::H2) = n(l, sid::total) * (bg * h2 / h2he);
        n(l, sid::He) = n(l, sid::total) * (bg * he / h2he);
      }
      meanWeight();
    }
  }
  altitudes(g, opt);
  return neglect;
}

void Atmosphere::altitudes(double g, const AltitudeOptions& opt) {
  const size_t L = nb_levels;
  altitude[0] = 0.0;
  if (opt.scale_height_km > 0) {                   // constant scale height (atmosphere.cpp:338-350)
    const double h = opt.scale_height_km * 100000.0;
    for (size_t l = 1; l < L; ++l) altitude[l] = -h * std::log(pressure[l] / pressure[0]);
    return;
  }
  if (opt.fixed_mean_molecular_weight > 0)          // atmosphere.cpp:168
    std::fill(mean_molecular_weight.begin(), mean_molecular_weight.end(), opt.fixed_mean_molecular_weight);
  // hydrostatic altitude, trapezoid in 1/(rho g), with g falling off as
  // 1/r^2 when asked (atmosphere.cpp:288-334)
  std::vector<double> rho(L);
  for (size_t l = 0; l < L; ++l)
    rho[l] = mean_molecular_weight[l] * pressure[l] * 1e6 / (phys::gas_constant * temperature[l]);
  for (size_t l = 1; l < L; ++l) {
    double gl = g;
    if (opt.variable_gravity) {
      const double local_radius = opt.bottom_radius + altitude[l - 1];
      gl = g * std::pow(opt.bottom_radius / local_radius, 2.0);
    }
    altitude[l] = altitude[l - 1] +
                  (1.0 / (rho[l] * gl) + 1.0 / (rho[l - 1] * gl)) * 0.5 * (pressure[l - 1] * 1e6 - pressure[l] * 1e6);
  }
}

void Atmosphere::setStructure(double g, double bottom_radius, bool variable_gravity, const double* p,
                              const double* t, const std::vector<int>& species, const double* x) {
  const size_t L = nb_levels, S = species.size();
  pressure.assign(p, p + L);
  temperature.assign(t, t + L);
  std::fill(number_density.begin(), number_density.end(), 0.0);
  auto n = [&](size_t l, int id) -> double& { return number_density[l * n_ids_ + id]; };
  for (size_t l = 0; l < L; ++l) {
    n(l, sid::total) = pressure[l] * 1.e6 / phys::boltzmann_k / temperature[l];
    for (size_t j = 0; j < S; ++j) n(l, species[j]) = n(l, sid::total) * x[l * S + j];
  }
  // chemistry.h:47-58: every species in id order (the absent ones add zeros)
  for (size_t l = 0; l < L; ++l) {
    double mu = 0;
    for (size_t j = 0; j < n_ids_; ++j) mu += n(l, int(j)) / n(l, sid::total) * molecularWeight(int(j));
    mean_molecular_weight[l] = mu;
  }
  AltitudeOptions opt;
  opt.bottom_radius = bottom_radius;
  opt.variable_gravity = variable_gravity;
  altitudes(g, opt);
}

// Reverse sweep through compute(), last operation first:
//   z_l = z_{l-1} + (u_l + u_{l-1}) / (2 g_l) (p_{l-1} - p_l) 1e6, u = 1/rho = R T / (mu p 1e6),
//     g_l = g, or g (R / (R + z_{l-1}))^2 with variable gravity; mu the
//     chemistry's or the fixed one; or z_l = -H ln(p_l / p_0)
//   mu_l = sum_j n_j m_j / n_tot
//   background: n_H2,He = n_tot bg r_H2,He, bg = 1 - sum_{j != H2,He} n_j / n_tot
//   isoprofile: n_j = n_tot x_j, then Na = 16.2181 K unless Na is free
//   n_tot = p 1e6 / (k T);  T_l = poly(log10 p_l; dof), dof_i = dof_{i-1} theta_i
void Atmosphere::adjoint(double g, const double* tp, const double* cp, std::vector<double>& t_bar,
                         std::vector<double>& n_bar, std::vector<double>& z_bar, double& g_bar, double* tp_bar,
                         double* cp_bar, const AltitudeOptions& opt, double* radius_bar, double* option_bar) const {
  const size_t L = nb_levels;
  const int S = int(n_ids_);
  auto n = [&](size_t l, int id) { return number_density[l * n_ids_ + id]; };
  auto nb = [&](size_t l, int id) -> double& { return n_bar[l * n_ids_ + id]; };
  std::vector<double> mu_bar(L, 0.0), ntot_bar(L, 0.0);
  for (size_t l = 0; l < L; ++l) ntot_bar[l] = nb(l, sid::total);
  g_bar = 0;
  std::vector<double> u_bar(L, 0.0);
  if (opt.scale_height_km > 0) {
    double h_bar = 0;
    for (size_t l = 1; l < L; ++l) h_bar -= z_bar[l] * std::log(pressure[l] / pressure[0]);
    if (option_bar) *option_bar += h_bar * 100000.0;
  } else {
    // altitude: zb[k] = dL/dz_k with the layers above carried down
    std::vector<double> zb(z_bar.begin(), z_bar.begin() + L);
    for (size_t k = L - 1; k >= 1; --k) {
      const double rho_k = mean_molecular_weight[k] * pressure[k] * 1e6 / (phys::gas_constant * temperature[k]);
      const double rho_p = mean_molecular_weight[k - 1] * pressure[k - 1] * 1e6 /
                           (phys::gas_constant * temperature[k - 1]);
      const double dp = 0.5 * (pressure[k - 1] * 1e6 - pressure[k] * 1e6);
      double gk = g, radius = 0;
      if (opt.variable_gravity) {
        radius = opt.bottom_radius + altitude[k - 1];
        gk = g * std::pow(opt.bottom_radius / radius, 2.0);
      }
      const double h = (1.0 / (rho_k * gk) + 1.0 / (rho_p * gk)) * dp;
      u_bar[k] += zb[k] * dp / gk;
      u_bar[k - 1] += zb[k] * dp / gk;
      const double gk_bar = -zb[k] * h / gk;
      zb[k - 1] += zb[k];
      if (opt.variable_gravity) {
        g_bar += gk_bar * gk / g;
        if (radius_bar) *radius_bar += gk_bar * gk * (2.0 / opt.bottom_radius - 2.0 / radius);
        zb[k - 1] -= gk_bar * 2.0 * gk / radius;
      } else {
        g_bar += gk_bar;
      }
    }
  }
  for (size_t l = 0; l < L; ++l) {
    const double u = phys::gas_constant * temperature[l] / (mean_molecular_weight[l] * pressure[l] * 1e6);
    t_bar[l] += u_bar[l] * u / temperature[l];
    mu_bar[l] -= u_bar[l] * u / mean_molecular_weight[l];
  }
  if (opt.scale_height_km <= 0 && opt.fixed_mean_molecular_weight > 0) {
    // the fixed weight replaced the chemistry's in the altitude
    double m_bar = 0;
    for (size_t l = 0; l < L; ++l) m_bar += mu_bar[l];
    if (option_bar) *option_bar += m_bar;
    mu_bar.assign(L, 0.0);
  }
  // the last mean-weight evaluation
  for (size_t l = 0; l < L; ++l) {
    const double nt = n(l, sid::total);
    for (int j = 1; j < S; ++j) nb(l, j) += mu_bar[l] * molecularWeight(j) / nt;
    ntot_bar[l] -= mu_bar[l] * mean_molecular_weight[l] / nt;
  }
  // chemistry modules in reverse order
  size_t k_end = nb_chemistry_param_;
  for (auto m = chemistry_.rbegin(); m != chemistry_.rend(); ++m) {
    if (m->background) {
      const double h2 = 0.5, he = 0.085114, h2he = h2 + he;
      for (size_t l = 0; l < L; ++l) {
        const double nt = n(l, sid::total);
        double bg = 1.0;
        for (int j = 1; j < S; ++j)
          if (j != sid::H2 && j != sid::He) bg -= n(l, j) / nt;
        if (bg < 0) bg = 0;
        ntot_bar[l] += nb(l, sid::H2) * bg * h2 / h2he + nb(l, sid::He) * bg * he / h2he;
        const double bg_bar = bg_clamped_ ? 0.0 : nt * (nb(l, sid::H2) * h2 + nb(l, sid::He) * he) / h2he;
        nb(l, sid::H2) = 0;
        nb(l, sid::He) = 0;
        for (int j = 1; j < S; ++j)
          if (j != sid::H2 && j != sid::He) {
            nb(l, j) -= bg_bar / nt;
            ntot_bar[l] += bg_bar * n(l, j) / (nt * nt);
          }
      }
    } else {
      const size_t k0 = k_end - m->species.size();
      for (size_t l = 0; l < L; ++l) {
        if (!m->sodium_free) {
          nb(l, sid::K) += 16.2181 * nb(l, sid::Na);
          nb(l, sid::Na) = 0;
        }
        const double nt = n(l, sid::total);
        for (size_t j = 0; j < m->species.size(); ++j) {
          double& v = nb(l, m->species[j]);
          cp_bar[k0 + j] += v * nt;
          ntot_bar[l] += v * cp[k0 + j];
          v = 0;
        }
      }
      k_end = k0;
    }
  }
  // n_tot = p 1e6 / (k T)
  for (size_t l = 0; l < L; ++l) t_bar[l] -= ntot_bar[l] * n(l, sid::total) / temperature[l];
  // temperature profile
  if (isothermal_) {
    double s = 0;
    for (size_t l = 0; l < L; ++l) s += t_bar[l];
    tp_bar[0] += s;
    return;
  }
  const size_t nd = nb_temperature_param_;
  std::vector<double> dof(nd), dof_bar(nd, 0.0), unit(nd, 0.0);
  dof[0] = tp[0];
  for (size_t i = 1; i < nd; ++i) dof[i] = dof[i - 1] * tp[i];
  PiecewisePolynomial basis(*poly_);                  // T_l is linear in the degrees of freedom
  for (size_t i = 0; i < nd; ++i) {
    unit.assign(nd, 0.0);
    unit[i] = 1.0;
    basis.setValues(unit);
    for (size_t l = 0; l < L; ++l)
      if (!clamped_[l]) dof_bar[i] += t_bar[l] * basis.value(std::log10(pressure[l]));
  }
  // dof_i = tp_0 prod_{k=1..i} tp_k
  double carry = 0;
  for (size_t i = nd; i-- > 0;) {
    carry += dof_bar[i] * dof[i];                     // sum_{j >= i} dof_bar_j dof_j
    tp_bar[i] += carry / tp[i];
  }
}

}  // namespace bear_amd
