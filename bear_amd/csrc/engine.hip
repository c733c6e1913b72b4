// Device kernels and memory of one likelihood evaluation on gfx950.
//
// Layout in HBM (one GPU):
//   wavenumber, wavelength   [N] FP64
//   tables                   [rows][stride] FP64 log10 cross-sections, every
//                            species' (T, p) rows, wavenumber contiguous
//   flux                     [N] FP64, the emergent flux of the last evaluation
//   weights                  folded observation operator, one contiguous
//                            segment per bin
//   bins, ll_terms           [bins] FP64
// Per evaluation the host fills one pinned block (levels, stencils, shifts)
// that is copied in one transfer.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>

#include "engine.h"

namespace bear_amd {

namespace {

void check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  const int code = (e == hipErrorOutOfMemory) ? kOutOfMemory : kDevice;
  raise(code, std::string(what) + ": " + hipGetErrorString(e));
}

constexpr int kBlock = 256;

// Short-characteristics Gauss nodes and weights (short_characteristics.h:70-71)
constexpr double kMu1 = 0.211324865405187;
constexpr double kMu2 = 0.788675134594813;
constexpr double kW = 0.5;

// Planck constants in the reference's association order
// (additional/aux_functions.cpp:272-278): 2 h c c nu^3 / (exp(h nu c / k / T) - 1)
constexpr double kPlanckC1 = 2. * 6.62606896e-27 * 2.99792458e10 * 2.99792458e10;

__device__ __forceinline__ double levelKappa(const double* __restrict__ tab, size_t stride,
                                             const LevelSpecies* __restrict__ d, int nb_species) {
  double k = 0.0;
  for (int s = 0; s < nb_species; ++s) {
    const double n = d[s].n;
    if (n == 0.0) continue;
    const uint32_t r0 = d[s].row[0], r1 = d[s].row[1], r2 = d[s].row[2], r3 = d[s].row[3];
    const double c0 = tab[size_t(r0) * stride];
    double lo = c0;
    if (r1 != r0) lo = c0 + (tab[size_t(r1) * stride] - c0) * d[s].f_p_lower;
    double x = lo;
    if (r2 != r0) {
      const double c2 = tab[size_t(r2) * stride];
      double hi = c2;
      if (r3 != r2) hi = c2 + (tab[size_t(r3) * stride] - c2) * d[s].f_p_upper;
      x = lo + (hi - lo) * d[s].f_t;
    }
    k += exp10(x) * n;
  }
  return k;
}

// Fused opacity + emission radiative transfer, one wavenumber per thread.
// Per level: kappa_l(nu) = sum_s n_s 10^(bilinear log10 sigma_s)
// (opacity_species.cpp:277-345, 400-433); per layer the two-angle short
// characteristics of short_characteristics.cpp:39-125, bottom to top.
// kappa never leaves registers.
__global__ __launch_bounds__(kBlock) void emissionKernel(
    const double* __restrict__ wavenumber, const double* __restrict__ wavelength,
    const double* __restrict__ tables, size_t stride,
    const LevelState* __restrict__ levels, const LevelSpecies* __restrict__ species,
    int nb_levels, int nb_species, int n, double scaling,
    double* __restrict__ flux, double* __restrict__ spectrum) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double nu = wavenumber[i];
  const double* tab = tables + i;
  const double b_num = kPlanckC1 * nu * nu * nu;
  const double x_nu = 6.62606896e-27 * nu * 2.99792458e10 / 1.3806504e-16;

  double k_prev = levelKappa(tab, stride, species, nb_species);
  double b_prev = b_num / (exp(x_nu * levels[0].inv_temperature) - 1.0);
  double z_prev = levels[0].altitude;
  double i1 = b_prev, i2 = b_prev;
  for (int l = 1; l < nb_levels; ++l) {
    const double k_next = levelKappa(tab, stride, species + size_t(l) * nb_species, nb_species);
    const double b_next = b_num / (exp(x_nu * levels[l].inv_temperature) - 1.0);
    const double z_next = levels[l].altitude;
    const double tau = (z_next - z_prev) * (k_next + k_prev) / 2.;
    if (tau != 0) {
      const double r = 1.0 / tau;
      const double a1 = exp(-tau / kMu1);
      const double q1 = (a1 - 1.0) * (kMu1 * r);
      i1 = i1 * a1 + ((1.0 + q1) * b_next + (-a1 - q1) * b_prev);
      const double a2 = exp(-tau / kMu2);
      const double q2 = (a2 - 1.0) * (kMu2 * r);
      i2 = i2 * a2 + ((1.0 + q2) * b_next + (-a2 - q2) * b_prev);
    }
    k_prev = k_next;
    b_prev = b_next;
    z_prev = z_next;
  }
  const double f = 2.0 * 3.14159265358979323846 * (i1 * kMu1 * kW + i2 * kMu2 * kW) * 1e-3 * scaling;
  flux[i] = f;
  if (spectrum) {                      // radiative_transfer.h:58-60, per micron
    const double lam = wavelength[i];
    spectrum[i] = f / lam / lam * 10000.0;
  }
}

__device__ __forceinline__ double waveSum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// One workgroup per bin: dot product of the flux with the bin's folded
// weights (fixed order: strided per thread, then a fixed tree), the
// observation shift (forward_model.cpp:161-180), and the bin's Gaussian
// log-likelihood term (retrieval.cpp:330-358).
__global__ __launch_bounds__(kBlock) void binKernel(
    const double* __restrict__ flux, const double* __restrict__ weights, const BinDesc* __restrict__ bins,
    const double* __restrict__ shifts, double error_inflation, double* __restrict__ out,
    double* __restrict__ ll_terms) {
  __shared__ double partial[kBlock / 64];
  const BinDesc b = bins[blockIdx.x];
  const double* w = weights + b.weight_offset;
  const double* f = flux + b.start;
  double acc = 0.0;
  for (uint32_t j = threadIdx.x; j < b.length; j += kBlock) acc += w[j] * f[j];
  acc = waveSum(acc);
  if ((threadIdx.x & 63) == 0) partial[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = (partial[0] + partial[1]) + (partial[2] + partial[3]);
    const double shift = shifts[b.observation];
    if (shift != 0) v += shift;
    out[blockIdx.x] = v;
    const double e2 = b.error * b.error + error_inflation;
    const double d = b.data - v;
    ll_terms[blockIdx.x] = (-0.5 * log(e2 * 2.0 * 3.14159265358979323846) - 0.5 * d * d / e2) * b.weight;
  }
}

// Fixed-order sum of the per-bin terms in one workgroup.
__global__ __launch_bounds__(1024) void sumKernel(const double* __restrict__ terms, int n, double* __restrict__ out) {
  __shared__ double partial[16];
  double acc = 0.0;
  for (int j = threadIdx.x; j < n; j += 1024) acc += terms[j];
  acc = waveSum(acc);
  if ((threadIdx.x & 63) == 0) partial[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < 16; ++k) s += partial[k];
    *out = s;
  }
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

}  // namespace

DeviceEngine::DeviceEngine(int device) : device_(device) {
  int count = 0;
  check(hipGetDeviceCount(&count), "hipGetDeviceCount");
  if (device < 0 || device >= count) raise(kDevice, "no HIP device " + std::to_string(device));
  check(hipSetDevice(device), "hipSetDevice");
  hipStream_t s;
  check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
  stream_ = s;
  ll_ = deviceAlloc<double>(1, "log-likelihood");
}

DeviceEngine::~DeviceEngine() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(static_cast<hipStream_t>(stream_));
  for (void* p : {(void*)wavenumber_, (void*)wavelength_, (void*)tables_, (void*)flux_, (void*)weights_,
                  (void*)bins_desc_, (void*)bins_, (void*)ll_terms_, (void*)ll_, (void*)staging_dev_,
                  (void*)spectrum_})
    deviceFree(p);
  if (staging_host_) (void)hipHostFree(staging_host_);
  if (stream_) (void)hipStreamDestroy(static_cast<hipStream_t>(stream_));
}

void DeviceEngine::uploadGrid(const std::vector<double>& nu, const std::vector<double>& lam) {
  check(hipSetDevice(device_), "hipSetDevice");
  n_points_ = nu.size();
  stride_ = (n_points_ + 63) / 64 * 64;
  deviceFree(wavenumber_);
  deviceFree(wavelength_);
  deviceFree(flux_);
  wavenumber_ = deviceAlloc<double>(n_points_, "wavenumbers");
  wavelength_ = deviceAlloc<double>(n_points_, "wavelengths");
  flux_ = deviceAlloc<double>(n_points_, "flux");
  check(hipMemcpy(wavenumber_, nu.data(), n_points_ * sizeof(double), hipMemcpyHostToDevice), "copy wavenumbers");
  check(hipMemcpy(wavelength_, lam.data(), n_points_ * sizeof(double), hipMemcpyHostToDevice), "copy wavelengths");
}

void DeviceEngine::uploadTables(const TableStore& store) {
  check(hipSetDevice(device_), "hipSetDevice");
  if (store.nb_rows > 0 && store.nb_points != n_points_) raise(kInvalidInput, "table rows do not match the grid");
  deviceFree(tables_);
  tables_ = deviceAlloc<double>(size_t(store.nb_rows) * stride_, "cross-section tables");
  if (store.nb_rows > 0)
    check(hipMemcpy2D(tables_, stride_ * sizeof(double), store.log_cross_sections.data(), n_points_ * sizeof(double),
                      n_points_ * sizeof(double), store.nb_rows, hipMemcpyHostToDevice),
          "copy cross-section tables");
}

void DeviceEngine::uploadObservations(const std::vector<ObservationSetup>& obs) {
  check(hipSetDevice(device_), "hipSetDevice");
  std::vector<BinDesc> desc;
  std::vector<double> weights;
  obs_bin_offset_.clear();
  for (size_t o = 0; o < obs.size(); ++o) {
    obs_bin_offset_.push_back(desc.size());
    const ObservationSetup& s = obs[o];
    for (size_t b = 0; b < s.segment_start.size(); ++b) {
      BinDesc d{};
      d.weight_offset = weights.size();
      d.start = uint32_t(s.segment_start[b]);
      d.length = uint32_t(s.segment_length[b]);
      d.observation = uint32_t(o);
      d.data = s.data[b];
      d.error = s.error[b];
      d.weight = s.weight[b];
      weights.insert(weights.end(), s.segment_weight.begin() + s.segment_offset[b],
                     s.segment_weight.begin() + s.segment_offset[b] + s.segment_length[b]);
      desc.push_back(d);
    }
  }
  obs_bin_offset_.push_back(desc.size());
  n_bins_ = desc.size();
  n_obs_ = obs.size();
  deviceFree(weights_);
  deviceFree(bins_desc_);
  deviceFree(bins_);
  deviceFree(ll_terms_);
  weights_ = deviceAlloc<double>(weights.size(), "observation operator");
  bins_desc_ = deviceAlloc<BinDesc>(desc.size(), "bin descriptors");
  bins_ = deviceAlloc<double>(n_bins_, "bins");
  ll_terms_ = deviceAlloc<double>(n_bins_, "likelihood terms");
  check(hipMemcpy(weights_, weights.data(), weights.size() * sizeof(double), hipMemcpyHostToDevice), "copy weights");
  check(hipMemcpy(bins_desc_, desc.data(), desc.size() * sizeof(BinDesc), hipMemcpyHostToDevice), "copy bins");
}

void DeviceEngine::configure(size_t nb_levels, size_t nb_species) {
  check(hipSetDevice(device_), "hipSetDevice");
  nb_levels_ = nb_levels;
  nb_species_ = nb_species;
  const size_t lv = nb_levels * sizeof(LevelState);
  const size_t sp = nb_levels * nb_species * sizeof(LevelSpecies);
  const size_t sh = (n_obs_ + 1) * sizeof(double);
  staging_bytes_ = lv + sp + sh;
  if (staging_host_) (void)hipHostFree(staging_host_);
  deviceFree(staging_dev_);
  void* h = nullptr;
  check(hipHostMalloc(&h, staging_bytes_ + sizeof(double), hipHostMallocDefault), "pinned staging");
  staging_host_ = static_cast<unsigned char*>(h);
  std::memset(staging_host_, 0, staging_bytes_ + sizeof(double));
  staging_dev_ = deviceAlloc<unsigned char>(staging_bytes_, "per-evaluation inputs");
  levels_host_ = reinterpret_cast<LevelState*>(staging_host_);
  species_host_ = reinterpret_cast<LevelSpecies*>(staging_host_ + lv);
  shift_host_ = reinterpret_cast<double*>(staging_host_ + lv + sp);
  ll_host_ = reinterpret_cast<double*>(staging_host_ + staging_bytes_);
}

void DeviceEngine::evaluateEmission(double scaling, double inflation, double* spectrum_dev, double* const* bins_dev) {
  check(hipSetDevice(device_), "hipSetDevice");
  hipStream_t s = static_cast<hipStream_t>(stream_);
  check(hipMemcpyAsync(staging_dev_, staging_host_, staging_bytes_, hipMemcpyHostToDevice, s), "upload inputs");
  const size_t lv = nb_levels_ * sizeof(LevelState);
  const size_t sp = nb_levels_ * nb_species_ * sizeof(LevelSpecies);
  const LevelState* lev = reinterpret_cast<const LevelState*>(staging_dev_);
  const LevelSpecies* spc = reinterpret_cast<const LevelSpecies*>(staging_dev_ + lv);
  const double* shifts = reinterpret_cast<const double*>(staging_dev_ + lv + sp);
  const int n = int(n_points_);
  const int blocks = (n + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(emissionKernel, dim3(blocks), dim3(kBlock), 0, s, wavenumber_, wavelength_, tables_, stride_,
                     lev, spc, int(nb_levels_), int(nb_species_), n, scaling, flux_, spectrum_dev);
  check(hipGetLastError(), "emission kernel");
  if (n_bins_ > 0) {
    hipLaunchKernelGGL(binKernel, dim3(unsigned(n_bins_)), dim3(kBlock), 0, s, flux_, weights_, bins_desc_, shifts,
                       inflation, bins_, ll_terms_);
    check(hipGetLastError(), "bin kernel");
    hipLaunchKernelGGL(sumKernel, dim3(1), dim3(1024), 0, s, ll_terms_, int(n_bins_), ll_);
    check(hipGetLastError(), "likelihood kernel");
  }
  if (bins_dev)
    for (size_t o = 0; o < n_obs_; ++o) {
      const size_t a = obs_bin_offset_[o], b = obs_bin_offset_[o + 1];
      if (b > a && bins_dev[o])
        check(hipMemcpyAsync(bins_dev[o], bins_ + a, (b - a) * sizeof(double), hipMemcpyDeviceToDevice, s),
              "copy bins");
    }
}

double DeviceEngine::logLikelihood() {
  hipStream_t s = static_cast<hipStream_t>(stream_);
  check(hipMemcpyAsync(ll_host_, ll_, sizeof(double), hipMemcpyDeviceToHost, s), "read log-likelihood");
  check(hipStreamSynchronize(s), "evaluation");
  return *ll_host_;
}

void DeviceEngine::copyBins(std::vector<double>& out) {
  hipStream_t s = static_cast<hipStream_t>(stream_);
  out.resize(n_bins_);
  if (n_bins_) check(hipMemcpyAsync(out.data(), bins_, n_bins_ * sizeof(double), hipMemcpyDeviceToHost, s), "read bins");
  check(hipStreamSynchronize(s), "evaluation");
}

void DeviceEngine::copySpectrum(double* host, const double* spectrum_dev) {
  hipStream_t s = static_cast<hipStream_t>(stream_);
  check(hipMemcpyAsync(host, spectrum_dev, n_points_ * sizeof(double), hipMemcpyDeviceToHost, s), "read spectrum");
  check(hipStreamSynchronize(s), "evaluation");
}

double* DeviceEngine::spectrumBuffer() {
  if (!spectrum_) {
    check(hipSetDevice(device_), "hipSetDevice");
    spectrum_ = deviceAlloc<double>(n_points_, "high-resolution spectrum");
  }
  return spectrum_;
}

void DeviceEngine::synchronize() {
  check(hipStreamSynchronize(static_cast<hipStream_t>(stream_)), "evaluation");
}

}  // namespace bear_amd
