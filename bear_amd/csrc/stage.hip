// Device staging of a batch: each live point's atmosphere, stencils and
// scalars formed on the device from its parameters, in the records the
// batched kernels read (the steps of Atmosphere::compute and
// TabulatedModel::stageLevels, which stage on the host; the model's fixed
// data come from the host once, as a StagePlan and its blob).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "device_math.h"
#include "engine.h"

namespace bear_amd {

namespace {

void check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  raise(e == hipErrorOutOfMemory ? kOutOfMemory : kDevice, std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
T* deviceAlloc(size_t count, const char* what) {
  void* p = nullptr;
  if (count == 0) count = 1;
  check(hipMalloc(&p, count * sizeof(T)), what);
  return static_cast<T*>(p);
}

void deviceFree(void* p) {
  if (p) (void)hipFree(p);
}

constexpr int kStageBlock = 256;
constexpr double kLog2Ten = 3.321928094887362347870319429489390175864831393;
constexpr double kBoltzmann = 1.3806504e-16;     // cgs, as the host's phys::boltzmann_k
constexpr double kGas = 8.314472e7;              // phys::gas_constant
constexpr double kJupiter = 7.1492e9;            // phys::radius_jupiter

template <class T>
__device__ __forceinline__ const T* at(const unsigned char* blob, size_t offset) {
  return reinterpret_cast<const T*>(blob + offset);
}

// One stencil of grid G at level l and temperature t (OpacitySpecies::
// bracketLevel, opacity_species.cpp:160-273): temperature rows lo, hi (the
// bracketing pair, one row past the ends of the table), their pressure
// picks, the temperature fraction and its derivative.
struct DeviceStencil {
  int t_pos[2];
  int p_pos[4];
  double f_p_lower, f_p_upper, f_t, dft;
};

__device__ DeviceStencil bracket(const StagePlan& P, const unsigned char* blob, int g, int l, double t) {
  const StageGrid G = at<StageGrid>(blob, P.grids)[g];
  const double* tn = at<double>(blob, P.grid_t) + G.t_offset;       // each temperature row's temperature
  int lo = 0, hi = 0;
  if (t < tn[0]) {
  } else if (t > tn[G.nt - 1]) {
    lo = hi = G.nt - 1;
  } else {
    for (int i = 0; i < G.nt; ++i) {
      if (tn[i] == t) { lo = hi = i; break; }
      if (tn[i] < t && tn[i + 1] > t) { lo = i; hi = i + 1; break; }
    }
  }
  const StagePick* picks = at<StagePick>(blob, P.picks) + G.picks_offset + size_t(l) * G.nt;
  const StagePick pl = picks[lo], ph = picks[hi];
  const int* tstart = at<int>(blob, P.grid_tstart) + G.tstart_offset;
  const double* temps = at<double>(blob, P.row_temps) + G.temps_offset;
  const double t0 = temps[tstart[lo] + int(pl.lower)], t2 = temps[tstart[hi] + int(ph.lower)];
  DeviceStencil s;
  s.t_pos[0] = lo;
  s.t_pos[1] = hi;
  s.p_pos[0] = int(pl.lower);
  s.p_pos[1] = int(pl.upper);
  s.p_pos[2] = int(ph.lower);
  s.p_pos[3] = int(ph.upper);
  s.f_p_lower = pl.fraction;
  s.f_p_upper = ph.fraction;
  // the fraction between rows (lo, pl.lower) and (hi, ph.lower), 0 when they are one row
  const bool same = lo == hi && pl.lower == ph.lower;
  s.f_t = same ? 0.0 : (t - t0) / (t2 - t0);
  s.dft = same ? 0.0 : 1.0 / (t2 - t0);
  return s;
}

__device__ __forceinline__ uint32_t rowId(const StagePlan& P, const unsigned char* blob, const StageSpecies& sp,
                                          int t_pos, int p_pos) {
  const StageGrid G = at<StageGrid>(blob, P.grids)[sp.grid];
  const int t0 = at<int>(blob, P.grid_tstart)[G.tstart_offset + t_pos];
  return at<uint32_t>(blob, P.row_ids)[sp.row_offset + t0 + p_pos];
}

// One workgroup per live point.  LDS: per level the temperature, total
// density, mean molecular weight, density rho and altitude, then the active
// species' densities [L][nb_active] and the temperature degrees of freedom.
__global__ __launch_bounds__(kStageBlock) void stageKernel(StagePlan P, const unsigned char* __restrict__ blob,
                                                           const double* __restrict__ params, int batch,
                                                           unsigned char* __restrict__ records, size_t rec_bytes,
                                                           size_t rec_levels, size_t rec_species, size_t rec_adjoint,
                                                           size_t rec_shifts, int wide, int* __restrict__ neglect) {
  extern __shared__ double sh[];
  __shared__ int s_neglect;
  const int m = int(blockIdx.x);
  if (m >= batch) return;
  const int L = P.L, NA = P.nb_active, tid = int(threadIdx.x);
  double* T = sh;
  double* ntot = T + L;
  double* mmw = ntot + L;
  double* rho = mmw + L;
  double* z = rho + L;
  double* nd = z + L;                               // [L][NA]
  double* dof = nd + size_t(L) * NA;                // [nb_temp]
  const double* p = params + size_t(m) * P.nb_params;
  const double* cp = p + P.g0;
  const double* tp = cp + P.nb_chem;
  const double* pres = at<double>(blob, P.pressure);
  const int* slot = at<int>(blob, P.id_slot);       // species id -> active slot, -1: never set
  const int* active = at<int>(blob, P.active);
  const double* mw = at<double>(blob, P.mol_weight);
  if (tid == 0) {
    s_neglect = 0;
    if (!P.isothermal) {                            // piecewise_poly_temperature.cpp:49-84
      dof[0] = tp[0];
      for (int i = 1; i < P.nb_temp; ++i) dof[i] = dof[i - 1] * tp[i];
    }
  }
  __syncthreads();

  // temperature, chemistry, mean molecular weight (Atmosphere::compute)
  int neg = 0;
  for (int l = tid; l < L; l += kStageBlock) {
    double t;
    if (P.isothermal) {
      t = tp[0];
    } else {
      const int* bi = at<int>(blob, P.basis_index) + size_t(l) * P.basis_entries;
      const double* bw = at<double>(blob, P.basis_weight) + size_t(l) * P.basis_entries;
      double y = 0;
      for (int k = 0; k < P.basis_entries && bi[k] >= 0; ++k) y += dof[bi[k]] * bw[k];
      t = y;
      if (t < 50) { t = 50; neg = 1; }              // floor and flag (piecewise_poly_temperature.cpp:77-81)
    }
    T[l] = t;
    double* n = nd + size_t(l) * NA;
    for (int a = 0; a < NA; ++a) n[a] = 0.0;
    const double nt = pres[l] * 1.e6 / kBoltzmann / t;
    double mu = 0.0;
    auto meanWeight = [&] {
      double w = nt / nt * mw[0];
      for (int a = 0; a < NA; ++a) w += n[a] / nt * mw[active[a]];
      return w;
    };
    const StageModule* mods = at<StageModule>(blob, P.modules);
    const int* msp = at<int>(blob, P.module_species);
    int k = 0;
    for (int q = 0; q < P.nb_modules; ++q) {
      const StageModule md = mods[q];
      if (!md.background) {                         // isoprofile_chemistry.cpp:87-119
        for (int j = 0; j < md.count; ++j) n[slot[msp[md.first + j]]] = nt * cp[k + j];
        if (!md.sodium_free && slot[P.na_id] >= 0)
          n[slot[P.na_id]] = (slot[P.k_id] >= 0 ? n[slot[P.k_id]] : 0.0) * 16.2181;
        k += md.count;
        mu = meanWeight();
        double sum = 0;
        for (int a = 0; a < NA; ++a) sum += n[a] / nt;
        if (sum > 1 + 1e-10) neg = 1;
      } else {                                      // background_chemistry.cpp:72-137
        const double h2 = 0.5, he = 0.085114, h2he = h2 + he;
        double bg = 1.0;
        for (int a = 0; a < NA; ++a)
          if (active[a] != P.h2_id && active[a] != P.he_id) bg -= n[a] / nt;
        if (bg < 0) { neg = 1; bg = 0; }
        if (slot[P.h2_id] >= 0) n[slot[P.h2_id]] = nt * (bg * h2 / h2he);
        if (slot[P.he_id] >= 0) n[slot[P.he_id]] = nt * (bg * he / h2he);
        mu = meanWeight();
      }
    }
    ntot[l] = nt;
    if (P.alt_kind == 1 && p[3] > 0) mu = p[3];     // fixed mean molecular weight (atmosphere.cpp:168)
    mmw[l] = mu;
    rho[l] = mu * pres[l] * 1e6 / (kGas * t);
  }
  if (neg) atomicOr(&s_neglect, 1);
  __syncthreads();

  // altitudes (atmosphere.cpp:288-350), in order over the levels
  const double g = pow(10.0, p[0]);
  if (tid == 0) {
    z[0] = 0.0;
    if (P.alt_kind == 2 && p[3] > 0) {
      const double h = p[3] * 100000.0;
      const double* lr = at<double>(blob, P.log_ratio);
      for (int l = 1; l < L; ++l) z[l] = -h * lr[l];
    } else {
      const double radius = p[1];
      for (int l = 1; l < L; ++l) {
        double gl = g;
        if (P.variable_gravity) {
          const double x = radius / (radius + z[l - 1]);
          gl = g * (x * x);
        }
        z[l] = z[l - 1] + (1.0 / (rho[l] * gl) + 1.0 / (rho[l - 1] * gl)) * 0.5 * (pres[l - 1] * 1e6 - pres[l] * 1e6);
      }
    }
  }
  __syncthreads();

  // per level: LevelState and the stencils of every slot (stageLevels)
  unsigned char* rec = records + size_t(m) * rec_bytes;
  LevelState* lev = reinterpret_cast<LevelState*>(rec + rec_levels);
  const StageSpecies* sps = at<StageSpecies>(blob, P.species);
  for (int l = tid; l < L; l += kStageBlock) {
    LevelState& ls = lev[l];
    {
      uint64_t* w = reinterpret_cast<uint64_t*>(&ls);
      for (size_t q = 0; q < sizeof(LevelState) / 8; ++q) w[q] = 0;
    }
    const double t = T[l];
    const double* n = nd + size_t(l) * NA;
    ls.temperature = t;
    ls.inv_temperature = 1.0 / t;
    ls.altitude = z[l];
    ls.rayleigh_density = (P.rayleigh_id >= 0 && slot[P.rayleigh_id] >= 0) ? n[slot[P.rayleigh_id]] : 0.0;
    if (l > 0) ls.cloud_scale = P.kind == int(ModelKind::transmission) ? 1.0 / (z[l] - z[l - 1]) : 1.0;
    DeviceStencil gst[2];
    for (int q = 0; q < 2; ++q) {
      if (P.group_grid[q] < 0) continue;
      const DeviceStencil st = bracket(P, blob, P.group_grid[q], l, t);
      gst[q] = st;
      const double ft = st.f_t, fl = st.f_p_lower, fu = st.f_p_upper;
      ls.group_w[q][0] = (1.0 - ft) * (1.0 - fl) * kLog2Ten;
      ls.group_w[q][1] = (1.0 - ft) * fl * kLog2Ten;
      ls.group_w[q][2] = ft * (1.0 - fu) * kLog2Ten;
      ls.group_w[q][3] = ft * fu * kLog2Ten;
      for (int k = 0; k < 4; ++k) ls.group_w32[q][k] = float(ls.group_w[q][k]);
      const double du = st.dft * kLog2Ten;
      ls.group_u[q][0] = -(1.0 - fl) * du;
      ls.group_u[q][1] = -fl * du;
      ls.group_u[q][2] = (1.0 - fu) * du;
      ls.group_u[q][3] = fu * du;
    }
    for (int s = 0; s < P.NS; ++s) {
      double dens = 0.0;
      DeviceStencil st{{0, 0}, {0, 0, 0, 0}, 0.0, 0.0, 0.0, 0.0};
      uint32_t row[4] = {0, 0, 0, 0};
      if (s < P.S) {
        const StageSpecies sp = sps[s];
        dens = slot[sp.species_id] >= 0 ? n[slot[sp.species_id]] : 0.0;       // opacity_species.cpp:407-412
        if (sp.partner0 >= 0) dens *= slot[sp.partner0] >= 0 ? n[slot[sp.partner0]] : 0.0;
        if (sp.partner1 >= 0) dens *= slot[sp.partner1] >= 0 ? n[slot[sp.partner1]] : 0.0;
        st = P.grouped ? gst[(P.group1 >> s) & 1u] : bracket(P, blob, sp.grid, l, t);
        for (int k = 0; k < 4; ++k) row[k] = rowId(P, blob, sp, st.t_pos[k / 2], st.p_pos[k]);
      }
      uint32_t block[4];
      for (int k = 0; k < 4; ++k) block[k] = uint32_t(row[k] * P.blocks_per_row);
      const double ft = st.f_t, fl = st.f_p_lower, fu = st.f_p_upper;
      if (wide) {
        LevelSpecies& d = reinterpret_cast<LevelSpecies*>(rec + rec_species)[size_t(l) * P.NS + s];
        d.w[0] = (1.0 - ft) * (1.0 - fl) * kLog2Ten;
        d.w[1] = (1.0 - ft) * fl * kLog2Ten;
        d.w[2] = ft * (1.0 - fu) * kLog2Ten;
        d.w[3] = ft * fu * kLog2Ten;
        d.log2_n = dens > 0 ? log2(dens) : kZeroDensityLog2;
        if (s < 16) ls.log2_n[s] = d.log2_n;
        for (int k = 0; k < 4; ++k) d.block[k] = block[k];
        d.reload = 0;
        d.pad = 0;
        LevelSpeciesAdjoint& a = reinterpret_cast<LevelSpeciesAdjoint*>(rec + rec_adjoint)[size_t(l) * P.NS + s];
        const double du = st.dft * kLog2Ten;
        a.u[0] = -(1.0 - fl) * du;
        a.u[1] = -fl * du;
        a.u[2] = (1.0 - fu) * du;
        a.u[3] = fu * du;
        for (int k = 0; k < 4; ++k) a.pad[k] = 0.0;
      } else {
        LevelSpecies32& d = reinterpret_cast<LevelSpecies32*>(rec + rec_species)[size_t(l) * P.NS + s];
        d.w[0] = float((1.0 - ft) * (1.0 - fl) * kLog2Ten);
        d.w[1] = float((1.0 - ft) * fl * kLog2Ten);
        d.w[2] = float(ft * (1.0 - fu) * kLog2Ten);
        d.w[3] = float(ft * fu * kLog2Ten);
        d.log2_n = dens > 0 ? float(log2(dens)) : -__builtin_huge_valf();
        if (s < 16) ls.log2_n32[s] = d.log2_n;
        for (int k = 0; k < 4; ++k) d.block[k] = block[k];
        d.reload = 0;
        d.pad[0] = d.pad[1] = 0;
      }
    }
  }
  __syncthreads();

  // reload flags: a slot's rows differ from the previous level's (always at level 0)
  for (int l = tid; l < L; l += kStageBlock) {
    uint32_t bits = 0;
    for (int s = 0; s < P.NS; ++s) {
      uint32_t changed = l == 0;
      const size_t e = size_t(l) * P.NS + s;
      if (wide) {
        LevelSpecies* d = reinterpret_cast<LevelSpecies*>(rec + rec_species);
        if (l > 0)
          for (int k = 0; k < 4; ++k) changed |= d[e - P.NS].block[k] != d[e].block[k];
        d[e].reload = changed;
      } else {
        LevelSpecies32* d = reinterpret_cast<LevelSpecies32*>(rec + rec_species);
        if (l > 0)
          for (int k = 0; k < 4; ++k) changed |= d[e - P.NS].block[k] != d[e].block[k];
        d[e].reload = changed;
      }
      bits |= changed << s;
    }
    lev[l].reload = bits;
  }

  // the point's scalars and observation shifts (emission.cpp:108-120,
  // secondary_eclipse.cpp:117-153, transmission.cpp:150-195,
  // forward_model.cpp:184-204, retrieval.cpp:333-336)
  if (tid == 0) {
    BatchModel& bm = *reinterpret_cast<BatchModel*>(rec);
    bm = BatchModel{};
    if (P.kind == int(ModelKind::emission)) {
      const double sc = kJupiter / p[2];
      bm.scaling = sc * sc * p[1];
    } else if (P.kind == int(ModelKind::eclipse)) {
      bm.scaling = 1.0;
      bm.star_inv_temperature = 1.0 / p[2];
      bm.depth_scale = p[1] * p[1] * 1e6;
    } else {
      bm.bottom_radius = p[1];
      bm.star_radius = p[2];
    }
    bm.inflation = P.error_inflation ? pow(10.0, p[P.nb_params - 1]) : 0.0;
    double* shifts = reinterpret_cast<double*>(rec + rec_shifts);
    const int* om = at<int>(blob, P.obs_mods);
    int k = 0;
    for (int o = 0; o < P.nb_obs; ++o) {
      shifts[o] = om[o] ? p[P.mods_offset + k] : 0.0;
      k += om[o];
    }
    neglect[m] = s_neglect;
  }
}

size_t stageLds(const StagePlan& plan) {
  return (size_t(plan.L) * (5 + plan.nb_active) + size_t(plan.nb_temp)) * sizeof(double);
}

}  // namespace

void DeviceEngine::uploadStagePlan(const StagePlan& plan, const std::vector<unsigned char>& blob) {
  check(hipSetDevice(device_), "hipSetDevice");
  deviceFree(stage_blob_);
  stage_blob_ = nullptr;
  stage_plan_ = plan;
  stage_blob_ = deviceAlloc<unsigned char>(std::max<size_t>(1, blob.size()), "staging plan");
  hipStream_t s = static_cast<hipStream_t>(stream_);
  check(hipMemcpyAsync(stage_blob_, blob.data(), blob.size(), hipMemcpyHostToDevice, s), "copy staging plan");
  check(hipStreamSynchronize(s), "copy staging plan");
}

void DeviceEngine::reserveStage(size_t batch) {
  const size_t n = size_t(stage_plan_.nb_params);
  if (batch <= stage_capacity_) return;
  check(hipSetDevice(device_), "hipSetDevice");
  check(hipStreamSynchronize(static_cast<hipStream_t>(stream_)), "batch");
  releaseGraph();
  deviceFree(stage_params_);
  deviceFree(stage_neglect_);
  if (stage_params_host_) (void)hipHostFree(stage_params_host_);
  if (stage_neglect_host_) (void)hipHostFree(stage_neglect_host_);
  stage_params_ = nullptr;
  stage_neglect_ = nullptr;
  stage_params_host_ = nullptr;
  stage_neglect_host_ = nullptr;
  stage_capacity_ = 0;
  stage_params_ = deviceAlloc<double>(batch * std::max<size_t>(1, n), "staged parameters");
  stage_neglect_ = deviceAlloc<int>(batch, "staged neglect flags");
  void* h = nullptr;
  check(hipHostMalloc(&h, batch * std::max<size_t>(1, n) * sizeof(double), hipHostMallocDefault),
        "pinned staged parameters");
  stage_params_host_ = static_cast<double*>(h);
  check(hipHostMalloc(&h, batch * sizeof(int), hipHostMallocDefault), "pinned neglect flags");
  stage_neglect_host_ = static_cast<int*>(h);
  stage_capacity_ = batch;
}

void DeviceEngine::launchStage(size_t batch) {
  hipStream_t s = static_cast<hipStream_t>(stream_);
  const size_t n = size_t(stage_plan_.nb_params);
  check(hipMemcpyAsync(stage_params_, stage_params_host_, batch * n * sizeof(double), hipMemcpyHostToDevice, s),
        "upload staged parameters");
  hipLaunchKernelGGL(stageKernel, dim3(unsigned(batch)), dim3(kStageBlock), stageLds(stage_plan_), s, stage_plan_,
                     stage_blob_, stage_params_, int(batch), batch_dev_, rec_bytes_, rec_levels_, rec_species_,
                     rec_adjoint_, rec_shifts_, int(precision_ == TablePrecision::fp64), stage_neglect_);
  check(hipGetLastError(), "staging kernel");
}

void DeviceEngine::releaseStage() {
  for (void* p : {(void*)stage_blob_, (void*)stage_params_, (void*)stage_neglect_}) deviceFree(p);
  if (stage_params_host_) (void)hipHostFree(stage_params_host_);
  if (stage_neglect_host_) (void)hipHostFree(stage_neglect_host_);
  stage_blob_ = nullptr;
  stage_params_ = nullptr;
  stage_neglect_ = nullptr;
  stage_params_host_ = nullptr;
  stage_neglect_host_ = nullptr;
  stage_capacity_ = 0;
}

}  // namespace bear_amd
