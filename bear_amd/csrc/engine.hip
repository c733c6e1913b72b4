// Device kernels and memory of one likelihood evaluation on gfx950.
//
// Layout in HBM (one GPU):
//   wavenumber, wavelength   [N] FP64
//   tables                   [rows][stride] FP64 log10 cross-sections, every
//                            species' (T, p) rows, wavenumber contiguous
//   rayleigh                 [N] FP64 Rayleigh cross-section (transmission)
//   flux                     [N] FP64, the high-resolution spectrum of the
//                            last evaluation (flux, or depth in ppm)
//   weights                  folded observation operator, one contiguous
//                            segment per bin
//   bins, ll_terms           [bins] FP64
// Per evaluation the host fills one pinned block (levels, stencils, shifts,
// chord lengths) that is copied in one transfer.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "device_math.h"
#include "engine.h"

namespace bear_amd {

using namespace dev;

// The single-model path's per-evaluation scratch for callers that cannot
// allocate device memory (DeviceEngine::useScratch): one array per GPU, held
// until the process ends.  Host code never passes `scratch` itself to a
// kernel: its device address on each GPU comes from hipGetSymbolAddress at
// set-up (scratchAddress).
constexpr size_t kScratchBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

void check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  const int code = (e == hipErrorOutOfMemory) ? kOutOfMemory : kDevice;
  raise(code, std::string(what) + ": " + hipGetErrorString(e));
}

// Single-point emission: one wavefront per workgroup.  The grid holds more
// wavefronts than the chip's slots at the kernel's occupancy (4 per SIMD at
// 103 VGPRs), and single-wavefront workgroups let the dispatcher refill each
// freed slot at once (4 % faster than 256 at 400,000 wavenumbers).
constexpr int kBlock = 64;
// Single-point transmission: workgroups of 8 wavefronts share one tile's
// segment extinction sums in LDS (S[segment][tile], rows padded to
// chordSegments) and the partial effective heights [kTransmissionBlock /
// tile][tile], FP64
constexpr int kTransmissionBlock = 512;
size_t transmissionLds(size_t nb_levels, size_t tile) {
  return (size_t(chordSegments(int(nb_levels))) + kTransmissionBlock / tile) * tile * sizeof(double);
}

// Emission spectrum epilogue: the flux F = 2 pi (mu1 w1 I1 + mu2 w2 I2) 1e-3 s
// (short_characteristics.cpp:117-122) is kept for the bins; the optional
// high-resolution output is F per micron (radiative_transfer.h:58-60) for
// emission (mode 0) or the eclipse depth F / F_star (Rp/Rs)^2 1e6
// (secondary_eclipse.cpp:249-251, mode 1).
// Mode 2: the eclipse depth against a tabulated stellar flux on the grid,
// F / F_star[i] (Rp/Rs)^2 1e6 (secondary_eclipse.cpp:462-464).
struct SpectrumOut {
  int mode;
  double star_inv_temperature;
  double depth_scale;
  const double* star;
};

__device__ __forceinline__ void writeFlux(int i, double nu, double i1, double i2, double scaling,
                                          const double* wavelength, const SpectrumOut& out, double* flux,
                                          double* spectrum) {
  const double f = 2.0 * kPi * (i1 * kMu1 * kW + i2 * kMu2 * kW) * 1e-3 * scaling;
  flux[i] = f;
  if (spectrum) {
    if (out.mode == 1) {
      spectrum[i] = f * recip(starFlux(nu, out.star_inv_temperature)) * out.depth_scale;
    } else if (out.mode == 2) {
      spectrum[i] = f / out.star[i] * out.depth_scale;
    } else {
      const double lam = wavelength[i];
      spectrum[i] = f / lam / lam * 10000.0;
    }
  }
}

// Fused opacity + emission radiative transfer, one wavenumber per thread,
// levels bottom to top (short_characteristics.cpp:39-125).  The four table
// values of every stencil slot stay in registers from level to level and are
// reloaded only where the rows change, so each table row is read about once
// per thread; kappa and the intensities never leave registers.  The next
// level's rows are requested as soon as this level's kappa is formed, so
// that their latency passes under this layer's Planck function and
// attenuation factors.  The stencil fields are wave-uniform (scalar loads).  EX: the extras (H- continuum,
// clouds) are formed once per thread and added per level and layer: the H-
// term to kappa, the cloud optical depth to the layer's
// (short_characteristics.cpp:75-76).
template <int NS, bool EX>
__global__ __launch_bounds__(kBlock) void emissionKernelCached(
    const double* __restrict__ wavenumber, const double* __restrict__ wavelength,
    const double* __restrict__ tables, size_t stride, Extras ex,
    const LevelState* __restrict__ levels, const LevelSpecies* __restrict__ species, StencilGroups groups,
    int nb_levels, int n, double scaling, SpectrumOut out, double* __restrict__ flux, double* __restrict__ spectrum) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double nu = wavenumber[i];
  const double* tab = tables + i;
  const double b_num = kPlanckC1 * nu * nu * nu;
  const double x_nu = -(6.62606896e-27 * nu * 2.99792458e10 / 1.3806504e-16) * kLog2E;
  double v[NS][4];
  double hm[7], cf[kMaxClouds];
  if (EX) loadExtras(ex, i, n, wavelength[i], hm, cf);
  loadRows<NS>(tab, species, v);
  double k_prev = kappaFrom<NS>(v, species, levels[0], groups);
  if (EX && ex.hminus) k_prev += hminusKappa(hm, levels[0]);
  if (nb_levels > 1 && levels[1].reload) loadRows<NS>(tab, species + NS, v);
  double b_prev = planckAt(b_num, x_nu, levels[0].inv_temperature);
  double z_prev = levels[0].altitude;
  double i1 = b_prev, i2 = b_prev;
  for (int l = 1; l < nb_levels; ++l) {
    const LevelSpecies* d = species + size_t(l) * NS;
    double k_next = kappaFrom<NS>(v, d, levels[l], groups);
    if (EX && ex.hminus) k_next += hminusKappa(hm, levels[l]);
    if (l + 1 < nb_levels && levels[l + 1].reload) loadRows<NS>(tab, d + NS, v);
    const double b_next = planckAt(b_num, x_nu, levels[l].inv_temperature);
    const double z_next = levels[l].altitude;
    double tau = (z_next - z_prev) * (k_next + k_prev) * 0.5;
    if (EX) tau += cloudLayer(ex, cf, levels[l], l, i, n);
    shortCharacteristicsLayer(tau, b_prev, b_next, i1, i2);
    k_prev = k_next;
    b_prev = b_next;
    z_prev = z_next;
  }
  writeFlux(i, nu, i1, i2, scaling, wavelength, out, flux, spectrum);
}

// The same for any number of species, without the register cache.
__global__ __launch_bounds__(kBlock) void emissionKernel(
    const double* __restrict__ wavenumber, const double* __restrict__ wavelength,
    const double* __restrict__ tables, size_t stride, Extras ex,
    const LevelState* __restrict__ levels, const LevelSpecies* __restrict__ species,
    int nb_levels, int nb_chunks, int n, double scaling, SpectrumOut out,
    double* __restrict__ flux, double* __restrict__ spectrum) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double nu = wavenumber[i];
  const double* tab = tables + i;
  const double b_num = kPlanckC1 * nu * nu * nu;
  const double x_nu = -(6.62606896e-27 * nu * 2.99792458e10 / 1.3806504e-16) * kLog2E;
  const int stride_species = nb_chunks * kChunk;
  const bool clouds = ex.clouds.count > 0 || ex.cloud_tau;
  double hm[7], cf[kMaxClouds];
  loadExtras(ex, i, n, wavelength[i], hm, cf);
  double k_prev = chunkedKappa(tab, stride, species, levels[0], nb_chunks);
  if (ex.hminus) k_prev += hminusKappa(hm, levels[0]);
  double b_prev = planckAt(b_num, x_nu, levels[0].inv_temperature);
  double z_prev = levels[0].altitude;
  double i1 = b_prev, i2 = b_prev;
  for (int l = 1; l < nb_levels; ++l) {
    double k_next = chunkedKappa(tab, stride, species + size_t(l) * stride_species, levels[l], nb_chunks);
    if (ex.hminus) k_next += hminusKappa(hm, levels[l]);
    const double b_next = planckAt(b_num, x_nu, levels[l].inv_temperature);
    const double z_next = levels[l].altitude;
    double tau = (z_next - z_prev) * (k_next + k_prev) * 0.5;
    if (clouds) tau += cloudLayer(ex, cf, levels[l], l, i, n);
    shortCharacteristicsLayer(tau, b_prev, b_next, i1, i2);
    k_prev = k_next;
    b_prev = b_next;
    z_prev = z_next;
  }
  writeFlux(i, nu, i1, i2, scaling, wavelength, out, flux, spectrum);
}

// Transmission: transit depth per wavenumber (transmission_spectrum.cpp:32-154).
// A workgroup owns `tile` consecutive wavenumbers (a multiple of 64):
//  1. extinction: thread (level group g, wavenumber j) computes
//     k_l = kappa_abs + sigma_R n_R for its contiguous range of levels, with
//     the register-cached stencils, and stores the segment sums
//     S[q] = k_q + k_{q+1} in LDS (the sum across a group boundary after the
//     first barrier, from the next group's first level);
//  2. tangent optical depths: each thread group takes blocks of eight
//     tangent levels (tangentBlock: chords by scalar loads, summed outward in
//     the reference's order, stopping at the first sum above 10) and adds
//     c_t (1 - e^-tau) to its thread's FP64 partial, where
//     c_t = (R+z_t)(z_t+1 - z_t-1) is the coefficient of (1 - T_t) in the
//     reference's trapezoid (one-sided at t = 0; T = 1 at the top level);
//  3. per wavenumber, the partials in a fixed order, h = sqrt(H + R^2) - R and
//     the depth (R_p + h)^2 / R_*^2 1e6 in FP64.
// The transposed chord matrix chord[q][t] comes from the host.  EX: the H-
// continuum joins the absorption and the clouds' layer optical depths join
// as level extinction (transmission_spectrum.cpp:121-130).
template <int NS, bool EX>
__global__ __launch_bounds__(kTransmissionBlock) void transmissionKernel(
    const double* __restrict__ tables, size_t stride, const double* __restrict__ rayleigh,
    const double* __restrict__ wavelength, Extras ex,
    const LevelState* __restrict__ levels, const LevelSpecies* __restrict__ species, StencilGroups stencil_groups,
    const double* __restrict__ chord, int nb_levels, int n, int tile, double bottom_radius, double star_radius,
    double* __restrict__ depth, double* __restrict__ spectrum) {
  extern __shared__ double lds[];
  const int L = nb_levels;
  const int segments = chordSegments(L), tangents = chordTangents(L);
  double* S = lds;
  double* part = lds + size_t(segments) * tile;             // also the group-boundary levels of step 1
  const int groups = kTransmissionBlock / tile;
  const int j = threadIdx.x % tile;
  const int g = __builtin_amdgcn_readfirstlane(int(threadIdx.x) / tile);   // tile: a multiple of 64
  const int i = blockIdx.x * tile + j;
  const int chunk = (L + groups - 1) / groups;
  const int l0 = g * chunk;
  const int l1 = min(L, l0 + chunk);
  double last = 0.0;
  if (i < n) {
    const double* tab = tables + i;
    const double sigma_r = rayleigh ? rayleigh[i] : 0.0;
    double v[NS][4];
    double hm[7], cf[kMaxClouds];
    if (EX) loadExtras(ex, i, n, wavelength[i], hm, cf);
    for (int l = l0; l < l1; ++l) {
      double k = kappaTerms<NS>(tab, stride, species + size_t(l) * NS, levels[l], stencil_groups,
                                l == l0 || levels[l].reload != 0, v) +
                 sigma_r * levels[l].rayleigh_density;
      if (EX) {
        if (ex.hminus) k += hminusKappa(hm, levels[l]);
        if (l > 0) k += cloudLayer(ex, cf, levels[l], l, i, n) * levels[l].cloud_scale;
      }
      if (l == l0) part[g * tile + j] = k;
      else S[(l - 1) * tile + j] = last + k;
      last = k;
    }
  } else {
    for (int l = l0; l < l1; ++l) {
      if (l == l0) part[g * tile + j] = 0.0;
      else S[(l - 1) * tile + j] = 0.0;
    }
  }
  __syncthreads();
  if (l0 < l1 && l1 < L) S[(l1 - 1) * tile + j] = last + part[(g + 1) * tile + j];
  if (g == 0)
    for (int q = L - 1; q < segments; ++q) S[q * tile + j] = 0.0;
  __syncthreads();
  const double r0 = bottom_radius;
  double h = 0.0;
  for (int t0 = kTangentBlock * g; t0 < tangents; t0 += kTangentBlock * groups) {
    double tau[kTangentBlock];
    tangentBlock<double>(S + j, tile, chord, tangents, t0, L - 1, tau);
#pragma unroll
    for (int k = 0; k < kTangentBlock; ++k) {
      const int t = t0 + k;
      if (t < L - 1) {
        const double zl = levels[t > 0 ? t - 1 : 0].altitude;
        const double zu = levels[t + 1].altitude;
        h += (r0 + levels[t].altitude) * (zu - zl) * (1.0 - pow2(-tau[k] * kLog2E));
      }
    }
  }
  part[g * tile + j] = h;
  __syncthreads();
  if (int(threadIdx.x) < tile && i < n) {
    double hs = 0.0;
    for (int q = 0; q < groups; ++q) hs += part[q * tile + j];
    const double hh = sqrt(hs + r0 * r0) - r0;
    const double r = hh + r0;
    const double d = r * r / star_radius / star_radius * 1e6;
    depth[i] = d;
    if (spectrum) spectrum[i] = d;
  }
}

// One workgroup per piece of a bin's segment (binChunkBody).
__global__ __launch_bounds__(kBlockBins) void binKernel(const double* __restrict__ flux,
                                                        const double* __restrict__ wavenumber,
                                                        const double* __restrict__ weights,
                                                        const BinChunk* __restrict__ chunks, int mode,
                                                        double star_inv_temperature, double* __restrict__ chunk_sums) {
  binChunkBody(flux, wavenumber, weights, chunks[blockIdx.x], mode, star_inv_temperature,
               chunk_sums + size_t(blockIdx.x) * 2);
}

// Bins and the log-likelihood in one workgroup (binFinalizeBody); the sum is
// also stored to the pinned host word the caller reads after the stream's
// synchronisation
__global__ __launch_bounds__(1024) void sumKernel(const double* __restrict__ chunk_sums,
                                                  const BinDesc* __restrict__ bins, int n, int mode,
                                                  double depth_scale, const double* __restrict__ shifts,
                                                  double error_inflation, double* __restrict__ out,
                                                  double* __restrict__ ll, double* __restrict__ ll_host) {
  binFinalizeBody(chunk_sums, bins, n, mode, depth_scale, shifts, error_inflation, out, ll, ll_host);
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

constexpr size_t kMaxSpecies = 16;                    // register-cached kernels up to 16 species

// scratch's address on `device` (current), resolved once per device
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch_address[kMaxDevices] = {};

unsigned char* scratchAddress(int device) {
  if (device < 0 || device >= kMaxDevices) raise(kDevice, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_address[device]) {
    void* ptr = nullptr;
    check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!ptr) raise(kDevice, "no device address for the scratch array");
    g_scratch_address[device] = static_cast<unsigned char*>(ptr);
  }
  return g_scratch_address[device];
}

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

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
  hipDeviceProp_t prop;
  check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
  nb_cus_ = prop.multiProcessorCount;
  lds_limit_ = prop.sharedMemPerBlock;
  scratch_ = scratchAddress(device);
}

size_t DeviceEngine::scratchBytes() { return kScratchBytes; }

// staged inputs, flux [stride], piece sums [pieces][2], bins, log-likelihood
size_t DeviceEngine::scratchRequired() const {
  return align256(staging_bytes_) + align256(stride_ * sizeof(double)) + align256(n_chunks_ * 2 * sizeof(double)) +
         align256(n_bins_ * sizeof(double)) + align256(sizeof(double));
}

void DeviceEngine::useScratch(bool on) {
  if (on && scratchRequired() > kScratchBytes)
    raise(kInvalidInput, "the no-allocation path holds at most " + std::to_string(kScratchBytes >> 20) +
                             " MiB of per-evaluation buffers; this model needs " + std::to_string(scratchRequired()) +
                             " bytes");
  use_scratch_ = on;
}

DeviceEngine::EvalBuffers DeviceEngine::buffers() const {
  if (!use_scratch_) return EvalBuffers{staging_dev_, flux_, chunk_sums_, bins_, ll_};
  unsigned char* p = scratch_;
  EvalBuffers b;
  b.staging = p;
  p += align256(staging_bytes_);
  b.flux = reinterpret_cast<double*>(p);
  p += align256(stride_ * sizeof(double));
  b.chunk_sums = reinterpret_cast<double*>(p);
  p += align256(n_chunks_ * 2 * sizeof(double));
  b.bins = reinterpret_cast<double*>(p);
  p += align256(n_bins_ * sizeof(double));
  b.ll = reinterpret_cast<double*>(p);
  return b;
}

DeviceEngine::~DeviceEngine() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(static_cast<hipStream_t>(stream_));
  releaseBatch();
  releaseAdjoint();
  releaseStage();
  for (void* p : {(void*)wavenumber_, (void*)wavelength_, (void*)tables_, (void*)rayleigh_, (void*)hminus_,
                  (void*)flux_, (void*)weights_, (void*)bins_desc_, (void*)bin_chunks_, (void*)bins_,
                  (void*)chunk_sums_, (void*)ll_, (void*)staging_dev_, (void*)cloud_tau_, (void*)spectrum_, (void*)adj_ptr_, (void*)adj_bin_, (void*)adj_weight_,
                  (void*)star_})
    deviceFree(p);
  if (staging_host_) (void)hipHostFree(staging_host_);
  if (event_start_) (void)hipEventDestroy(static_cast<hipEvent_t>(event_start_));
  if (event_stop_) (void)hipEventDestroy(static_cast<hipEvent_t>(event_stop_));
  if (stream_) (void)hipStreamDestroy(static_cast<hipStream_t>(stream_));
}

void DeviceEngine::setProfiling(bool on) {
  check(hipSetDevice(device_), "hipSetDevice");
  if (on && !event_start_) {
    hipEvent_t a, b;
    check(hipEventCreate(&a), "hipEventCreate");
    check(hipEventCreate(&b), "hipEventCreate");
    event_start_ = a;
    event_stop_ = b;
  }
  profiling_ = on;
  timed_launches_ = 0;
  timed_ms_ = 0;
  batch_evaluations_ = 0;
  batch_ms_ = 0;
  event_pending_ = false;
}

// after the stream has drained: add the last timed launch
void DeviceEngine::collectTiming() {
  if (!event_pending_) return;
  float ms = 0;
  check(hipEventElapsedTime(&ms, static_cast<hipEvent_t>(event_start_), static_cast<hipEvent_t>(event_stop_)),
        "hipEventElapsedTime");
  timed_ms_ += ms;
  ++timed_launches_;
  event_pending_ = false;
}

// distinct table rows referenced by the staged stencils
void DeviceEngine::countRows() {
  row_seen_.assign(n_rows_, 0);
  size_t count = 0;
  for (size_t k = 0; k < nb_levels_ * species_stride_; ++k) {
    const LevelSpecies& d = species_host_[k];
    if (d.log2_n <= kZeroDensityLog2) continue;                     // zero density: term removed
    uint32_t r[4];
    for (int q = 0; q < 4; ++q) r[q] = uint32_t(size_t(d.block[q]) * 64 / stride_);
    const bool use[4] = {true, r[1] != r[0], r[2] != r[0], r[2] != r[0] && r[3] != r[2]};
    for (int q = 0; q < 4; ++q)
      if (use[q] && r[q] < n_rows_ && !row_seen_[r[q]]) { row_seen_[r[q]] = 1; ++count; }
  }
  distinct_rows_ = count;
}

void DeviceEngine::beginTimed() {
  if (!profiling_) return;
  countRows();
  check(hipEventRecord(static_cast<hipEvent_t>(event_start_), static_cast<hipStream_t>(stream_)), "hipEventRecord");
}

void DeviceEngine::endTimed() {
  if (!profiling_) return;
  check(hipEventRecord(static_cast<hipEvent_t>(event_stop_), static_cast<hipStream_t>(stream_)), "hipEventRecord");
  event_pending_ = true;
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
  // set-up copies go through the engine's own (non-blocking) stream and are
  // waited for: a pageable hipMemcpy may return before its DMA has landed,
  // and nothing would order it before kernels on this stream
  hipStream_t s = static_cast<hipStream_t>(stream_);
  check(hipMemcpyAsync(wavenumber_, nu.data(), n_points_ * sizeof(double), hipMemcpyHostToDevice, s),
        "copy wavenumbers");
  check(hipMemcpyAsync(wavelength_, lam.data(), n_points_ * sizeof(double), hipMemcpyHostToDevice, s),
        "copy wavelengths");
  check(hipStreamSynchronize(s), "copy grid");
}

void DeviceEngine::requireFP64Tables(const char* what) const {
  if (storage_ != TablePrecision::fp64)
    raise(kInvalidInput, std::string(what) +
                             " needs the FP64 tables; these were loaded at reduced precision only (use the batched "
                             "forward calls)");
}

void DeviceEngine::setTableStorage(TablePrecision p) {
  if (tables_ || tables_low_) raise(kInternal, "table storage set after the tables were loaded");
  storage_ = p;
  precision_ = p;
}

void DeviceEngine::allocateTables(size_t nb_rows, size_t nb_species) {
  check(hipSetDevice(device_), "hipSetDevice");
  if (nb_rows * (stride_ / 64) > size_t(UINT32_MAX))   // stencils hold 32-bit block offsets
    raise(kInvalidInput, "tables too large: rows x wavenumbers exceeds 2^38 values");
  deviceFree(tables_);
  tables_ = nullptr;
  n_rows_ = nb_rows;
  if (storage_ == TablePrecision::fp64) {
    tables_ = deviceAlloc<double>(nb_rows * stride_, "cross-section tables");
    return;
  }
  // residual rows and the per-species base only; every species passes
  // through an FP64 staging buffer of its own rows (uploadSpeciesRows)
  const size_t slots = std::max<size_t>(kChunk, (nb_species + kChunk - 1) / kChunk * kChunk);
  allocateReduced(storage_, nb_rows, slots);
}

void DeviceEngine::uploadSpeciesRows(int slot, size_t first, size_t count, const double* rows) {
  if (first + count > n_rows_) raise(kInternal, "table rows beyond the allocated store");
  if (count == 0) return;
  hipStream_t s = static_cast<hipStream_t>(stream_);
  double* dst = tables_;
  if (storage_ == TablePrecision::fp64) dst += first * stride_;
  else dst = deviceAlloc<double>(count * stride_, "table staging");
  try {
    check(hipMemcpy2DAsync(dst, stride_ * sizeof(double), rows, n_points_ * sizeof(double), n_points_ * sizeof(double),
                           count, hipMemcpyHostToDevice, s),
          "copy cross-section tables");
    if (storage_ != TablePrecision::fp64) splitSpecies(dst, count, slot, first);
    check(hipStreamSynchronize(s), "copy cross-section tables");
  } catch (...) {
    if (storage_ != TablePrecision::fp64) deviceFree(dst);
    throw;
  }
  if (storage_ != TablePrecision::fp64) deviceFree(dst);
}

size_t DeviceEngine::tableBytes() const {
  size_t bytes = tables_ ? n_rows_ * stride_ * sizeof(double) : 0;
  if (tables_low_) bytes += n_rows_ * stride_ * tableElementBytes(precision_) + base_rows_ * stride_ * sizeof(float);
  return bytes;
}

void DeviceEngine::uploadRayleigh(const std::vector<double>& sigma) {
  check(hipSetDevice(device_), "hipSetDevice");
  deviceFree(rayleigh_);
  rayleigh_ = nullptr;
  if (sigma.empty()) return;
  if (sigma.size() != n_points_) raise(kInvalidInput, "Rayleigh cross-sections do not match the grid");
  rayleigh_ = deviceAlloc<double>(n_points_, "Rayleigh cross-sections");
  hipStream_t s = static_cast<hipStream_t>(stream_);
  check(hipMemcpyAsync(rayleigh_, sigma.data(), n_points_ * sizeof(double), hipMemcpyHostToDevice, s),
        "copy Rayleigh cross-sections");
  check(hipStreamSynchronize(s), "copy Rayleigh cross-sections");
}

void DeviceEngine::uploadHminus(const std::vector<double>& rows) {
  check(hipSetDevice(device_), "hipSetDevice");
  deviceFree(hminus_);
  hminus_ = nullptr;
  if (rows.empty()) return;
  if (rows.size() != 7 * n_points_) raise(kInvalidInput, "H- rows do not match the grid");
  hminus_ = deviceAlloc<double>(rows.size(), "H- continuum rows");
  hipStream_t s = static_cast<hipStream_t>(stream_);
  check(hipMemcpyAsync(hminus_, rows.data(), rows.size() * sizeof(double), hipMemcpyHostToDevice, s),
        "copy H- continuum rows");
  check(hipStreamSynchronize(s), "copy H- continuum rows");
}

void DeviceEngine::setFixedCloud(const double* tau) {
  fixed_cloud_ = false;
  if (!tau || nb_levels_ < 2) return;
  const size_t count = (nb_levels_ - 1) * n_points_;
  bool any = false;                               // emission.cpp:344-360: an all-zero cloud is no cloud
  for (size_t i = 0; i < count && !any; ++i) any = tau[i] > 0;
  if (!any) return;
  check(hipSetDevice(device_), "hipSetDevice");
  if (!cloud_tau_) cloud_tau_ = deviceAlloc<double>(count, "cloud optical depths");
  hipStream_t s = static_cast<hipStream_t>(stream_);
  check(hipMemcpyAsync(cloud_tau_, tau, count * sizeof(double), hipMemcpyHostToDevice, s), "copy cloud optical depths");
  check(hipStreamSynchronize(s), "copy cloud optical depths");
  fixed_cloud_ = true;
}

void DeviceEngine::uploadObservations(const std::vector<ObservationSetup>& obs) {
  check(hipSetDevice(device_), "hipSetDevice");
  std::vector<BinDesc> desc;
  std::vector<BinChunk> chunks;
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
      d.first_chunk = uint32_t(chunks.size());
      for (uint32_t k = 0; k < d.length; k += kBinChunk)
        chunks.push_back(BinChunk{d.weight_offset + k, d.start + k, std::min(kBinChunk, d.length - k)});
      weights.insert(weights.end(), s.segment_weight.begin() + s.segment_offset[b],
                     s.segment_weight.begin() + s.segment_offset[b] + s.segment_length[b]);
      desc.push_back(d);
    }
  }
  obs_bin_offset_.push_back(desc.size());
  n_bins_ = desc.size();
  n_chunks_ = chunks.size();
  n_obs_ = obs.size();
  deviceFree(weights_);
  deviceFree(bins_desc_);
  deviceFree(bin_chunks_);
  deviceFree(bins_);
  deviceFree(chunk_sums_);
  batch_capacity_ = 0;                                     // the batch's piece sums are sized by the pieces
  weights_ = deviceAlloc<double>(weights.size(), "observation operator");
  bins_desc_ = deviceAlloc<BinDesc>(desc.size(), "bin descriptors");
  bin_chunks_ = deviceAlloc<BinChunk>(chunks.size(), "bin pieces");
  bins_ = deviceAlloc<double>(n_bins_, "bins");
  chunk_sums_ = deviceAlloc<double>(n_chunks_ * 2, "bin piece sums");
  hipStream_t s = static_cast<hipStream_t>(stream_);
  if (!weights.empty())
    check(hipMemcpyAsync(weights_, weights.data(), weights.size() * sizeof(double), hipMemcpyHostToDevice, s),
          "copy weights");
  if (!desc.empty())
    check(hipMemcpyAsync(bins_desc_, desc.data(), desc.size() * sizeof(BinDesc), hipMemcpyHostToDevice, s),
          "copy bins");
  if (!chunks.empty())
    check(hipMemcpyAsync(bin_chunks_, chunks.data(), chunks.size() * sizeof(BinChunk), hipMemcpyHostToDevice, s),
          "copy bin pieces");
  // the operator transposed, per wavenumber (bins in increasing order), for the adjoint
  std::vector<uint64_t> ptr(n_points_ + 1, 0);
  for (const BinDesc& d : desc)
    for (uint32_t k = 0; k < d.length; ++k) ++ptr[d.start + k + 1];
  for (size_t i = 0; i < n_points_; ++i) ptr[i + 1] += ptr[i];
  std::vector<uint32_t> adj_bin(ptr.back());
  std::vector<double> adj_w(ptr.back());
  std::vector<uint64_t> fill(ptr.begin(), ptr.end() - 1);
  for (size_t j = 0; j < desc.size(); ++j)
    for (uint32_t k = 0; k < desc[j].length; ++k) {
      const uint64_t e = fill[desc[j].start + k]++;
      adj_bin[e] = uint32_t(j);
      adj_w[e] = weights[desc[j].weight_offset + k];
    }
  releaseAdjoint();
  deviceFree(adj_ptr_);
  deviceFree(adj_bin_);
  deviceFree(adj_weight_);
  adj_ptr_ = deviceAlloc<uint64_t>(ptr.size(), "transposed operator");
  adj_bin_ = deviceAlloc<uint32_t>(adj_bin.size(), "transposed operator");
  adj_weight_ = deviceAlloc<double>(adj_w.size(), "transposed operator");
  check(hipMemcpyAsync(adj_ptr_, ptr.data(), ptr.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s),
        "copy transposed operator");
  if (!adj_bin.empty()) {
    check(hipMemcpyAsync(adj_bin_, adj_bin.data(), adj_bin.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s),
          "copy transposed operator");
    check(hipMemcpyAsync(adj_weight_, adj_w.data(), adj_w.size() * sizeof(double), hipMemcpyHostToDevice, s),
          "copy transposed operator");
  }
  check(hipStreamSynchronize(s), "copy observations");
}

void DeviceEngine::configure(size_t nb_levels, size_t nb_species, bool transmission) {
  check(hipSetDevice(device_), "hipSetDevice");
  nb_levels_ = nb_levels;
  nb_species_ = nb_species;
  transmission_ = transmission;
  species_stride_ = (nb_species + kChunk - 1) / kChunk * kChunk;   // padding entries keep n = 0, row 0
  if (species_stride_ == 0) species_stride_ = kChunk;
  if (transmission && species_stride_ > kMaxSpecies)
    raise(kInvalidInput, "the transmission model takes at most 16 tabulated species");
  if (transmission && transmissionLds(nb_levels, 64) > lds_limit_)
    raise(kInvalidInput, "too many levels for the transmission kernel");
  const size_t lv = (nb_levels * sizeof(LevelState) + 63) / 64 * 64;      // stencil records start 64-B aligned
  const size_t sp = nb_levels * species_stride_ * sizeof(LevelSpecies);
  const size_t sh = (n_obs_ + 1) * sizeof(double);
  const size_t pa = transmission ? size_t(chordSegments(int(nb_levels))) * chordTangents(int(nb_levels)) *
                                       sizeof(double) : 0;
  off_species_ = lv;
  off_shift_ = lv + sp;
  off_path_ = lv + sp + sh;
  staging_bytes_ = lv + sp + sh + pa;
  if (staging_host_) (void)hipHostFree(staging_host_);
  deviceFree(staging_dev_);
  void* h = nullptr;
  check(hipHostMalloc(&h, staging_bytes_ + sizeof(double), hipHostMallocDefault), "pinned staging");
  staging_host_ = static_cast<unsigned char*>(h);
  std::memset(staging_host_, 0, staging_bytes_ + sizeof(double));
  staging_dev_ = deviceAlloc<unsigned char>(staging_bytes_, "per-evaluation inputs");
  levels_host_ = reinterpret_cast<LevelState*>(staging_host_);
  species_host_ = reinterpret_cast<LevelSpecies*>(staging_host_ + off_species_);
  for (size_t s = 0; s < species_stride_; ++s) species_host_[s].reload = 1;   // level 0 always loads
  shift_host_ = reinterpret_cast<double*>(staging_host_ + off_shift_);
  path_host_ = transmission ? reinterpret_cast<double*>(staging_host_ + off_path_) : nullptr;
  ll_host_ = reinterpret_cast<double*>(staging_host_ + staging_bytes_);
  void* mapped = nullptr;
  check(hipHostGetDevicePointer(&mapped, ll_host_, 0), "hipHostGetDevicePointer");
  ll_host_dev_ = static_cast<double*>(mapped);
}

void DeviceEngine::upload() {
  check(hipSetDevice(device_), "hipSetDevice");
  check(hipMemcpyAsync(buffers().staging, staging_host_, staging_bytes_, hipMemcpyHostToDevice,
                       static_cast<hipStream_t>(stream_)),
        "upload inputs");
}

void DeviceEngine::launchBins(int mode, double star_temperature, double depth_scale, double inflation,
                              double* const* bins_dev) {
  hipStream_t s = static_cast<hipStream_t>(stream_);
  const EvalBuffers buf = buffers();
  const double* shifts = reinterpret_cast<const double*>(buf.staging + off_shift_);
  if (n_bins_ > 0) {
    hipLaunchKernelGGL(binKernel, dim3(unsigned(n_chunks_)), dim3(kBlockBins), 0, s, buf.flux, wavenumber_, weights_,
                       bin_chunks_, mode, mode == 1 ? 1.0 / star_temperature : 0.0, buf.chunk_sums);
    check(hipGetLastError(), "bin kernel");
    hipLaunchKernelGGL(sumKernel, dim3(1), dim3(1024), 0, s, buf.chunk_sums, bins_desc_, int(n_bins_), mode,
                       depth_scale, shifts, inflation, buf.bins, buf.ll, ll_host_dev_);
    check(hipGetLastError(), "likelihood kernel");
  }
  if (bins_dev)
    for (size_t o = 0; o < n_obs_; ++o) {
      const size_t a = obs_bin_offset_[o], b = obs_bin_offset_[o + 1];
      if (b > a && bins_dev[o])
        check(hipMemcpyAsync(bins_dev[o], buf.bins + a, (b - a) * sizeof(double), hipMemcpyDeviceToDevice, s),
              "copy bins");
    }
}

namespace {
bool hasExtras(const Extras& ex) { return ex.hminus || ex.cloud_tau || ex.clouds.count > 0; }

void launchEmission(hipStream_t s, size_t species_stride, int blocks, const double* nu, const double* lam,
                    const double* tables, size_t stride, const Extras& ex, const LevelState* lev,
                    const LevelSpecies* spc, StencilGroups groups, int L, int n, double scaling, SpectrumOut out,
                    double* flux, double* spectrum) {
  switch (species_stride) {
#define BEAR_CACHED(NS)                                                                                             \
  case NS:                                                                                                          \
    if (hasExtras(ex))                                                                                              \
      hipLaunchKernelGGL((emissionKernelCached<NS, true>), dim3(blocks), dim3(kBlock), 0, s, nu, lam, tables,      \
                         stride, ex, lev, spc, groups, L, n, scaling, out, flux, spectrum);                         \
    else                                                                                                            \
      hipLaunchKernelGGL((emissionKernelCached<NS, false>), dim3(blocks), dim3(kBlock), 0, s, nu, lam, tables,     \
                         stride, ex, lev, spc, groups, L, n, scaling, out, flux, spectrum);                         \
    break;
    BEAR_CACHED(4)
    BEAR_CACHED(8)
    BEAR_CACHED(12)
    BEAR_CACHED(16)
#undef BEAR_CACHED
    default:
      hipLaunchKernelGGL(emissionKernel, dim3(blocks), dim3(kBlock), 0, s, nu, lam, tables, stride, ex, lev,
                         spc, L, int(species_stride / kChunk), n, scaling, out, flux, spectrum);
  }
}
}  // namespace

void DeviceEngine::evaluateEmission(double scaling, double inflation, double* spectrum_dev, double* const* bins_dev) {
  if (transmission_) raise(kInvalidInput, "the engine is configured for transmission");
  requireFP64Tables("the single-model path");
  upload();
  const EvalBuffers buf = buffers();
  hipStream_t s = static_cast<hipStream_t>(stream_);
  const LevelState* lev = reinterpret_cast<const LevelState*>(buf.staging);
  const LevelSpecies* spc = reinterpret_cast<const LevelSpecies*>(buf.staging + off_species_);
  const int n = int(n_points_);
  beginTimed();
  launchEmission(s, species_stride_, (n + kBlock - 1) / kBlock, wavenumber_, wavelength_, tables_, stride_, extras(),
                 lev, spc, groups_, int(nb_levels_), n, scaling, SpectrumOut{0, 0.0, 0.0, nullptr}, buf.flux, spectrum_dev);
  check(hipGetLastError(), "emission kernel");
  endTimed();
  launchBins(0, 0.0, 0.0, inflation, bins_dev);
}

void DeviceEngine::evaluateEclipse(double star_temperature, double radius_ratio, double inflation,
                                   double* spectrum_dev, double* const* bins_dev) {
  if (transmission_) raise(kInvalidInput, "the engine is configured for transmission");
  requireFP64Tables("the single-model path");
  upload();
  const EvalBuffers buf = buffers();
  hipStream_t s = static_cast<hipStream_t>(stream_);
  const LevelState* lev = reinterpret_cast<const LevelState*>(buf.staging);
  const LevelSpecies* spc = reinterpret_cast<const LevelSpecies*>(buf.staging + off_species_);
  const int n = int(n_points_);
  const double depth_scale = radius_ratio * radius_ratio * 1e6;
  beginTimed();
  launchEmission(s, species_stride_, (n + kBlock - 1) / kBlock, wavenumber_, wavelength_, tables_, stride_, extras(),
                 lev, spc, groups_, int(nb_levels_), n, 1.0, SpectrumOut{1, 1.0 / star_temperature, depth_scale, nullptr}, buf.flux, spectrum_dev);
  check(hipGetLastError(), "emission kernel");
  endTimed();
  launchBins(1, star_temperature, depth_scale, inflation, bins_dev);
}

void DeviceEngine::evaluateEclipseTabulatedStar(double radius_ratio, double* spectrum_dev) {
  if (transmission_) raise(kInvalidInput, "the engine is configured for transmission");
  if (!star_) raise(kInvalidInput, "no stellar spectrum on the grid");
  requireFP64Tables("the single-model path");
  upload();
  const EvalBuffers buf = buffers();
  hipStream_t s = static_cast<hipStream_t>(stream_);
  const LevelState* lev = reinterpret_cast<const LevelState*>(buf.staging);
  const LevelSpecies* spc = reinterpret_cast<const LevelSpecies*>(buf.staging + off_species_);
  const int n = int(n_points_);
  const double depth_scale = radius_ratio * radius_ratio * 1e6;
  beginTimed();
  launchEmission(s, species_stride_, (n + kBlock - 1) / kBlock, wavenumber_, wavelength_, tables_, stride_, extras(),
                 lev, spc, groups_, int(nb_levels_), n, 1.0, SpectrumOut{2, 0.0, depth_scale, star_}, buf.flux, spectrum_dev);
  check(hipGetLastError(), "emission kernel");
  endTimed();
}

void DeviceEngine::uploadStellarSpectrum(const std::vector<double>& flux) {
  check(hipSetDevice(device_), "hipSetDevice");
  if (flux.size() != n_points_) raise(kInvalidInput, "stellar spectrum does not match the grid");
  deviceFree(star_);
  star_ = nullptr;
  star_ = deviceAlloc<double>(n_points_, "stellar spectrum");
  hipStream_t s = static_cast<hipStream_t>(stream_);
  check(hipMemcpyAsync(star_, flux.data(), n_points_ * sizeof(double), hipMemcpyHostToDevice, s),
        "copy stellar spectrum");
  check(hipStreamSynchronize(s), "copy stellar spectrum");
}

void DeviceEngine::copyFlux(double* host) {
  hipStream_t s = static_cast<hipStream_t>(stream_);
  check(hipMemcpyAsync(host, buffers().flux, n_points_ * sizeof(double), hipMemcpyDeviceToHost, s), "read flux");
  check(hipStreamSynchronize(s), "evaluation");
  collectTiming();
}

void DeviceEngine::evaluateTransmission(double bottom_radius, double star_radius, double inflation,
                                        double* spectrum_dev, double* const* bins_dev) {
  if (!transmission_) raise(kInvalidInput, "the engine is not configured for transmission");
  requireFP64Tables("the single-model path");
  upload();
  const EvalBuffers buf = buffers();
  hipStream_t s = static_cast<hipStream_t>(stream_);
  const LevelState* lev = reinterpret_cast<const LevelState*>(buf.staging);
  const LevelSpecies* spc = reinterpret_cast<const LevelSpecies*>(buf.staging + off_species_);
  const double* path = reinterpret_cast<const double*>(buf.staging + off_path_);
  const int n = int(n_points_), L = int(nb_levels_);
  // tiles of 128 or 64 wavenumbers, the wider one when its LDS (the segment
  // sums and the partials) fits the workgroup
  const int tile = transmissionLds(L, 128) <= lds_limit_ ? 128 : 64;
  const size_t lds = transmissionLds(L, tile);
  const int blocks = (n + tile - 1) / tile;
  const Extras ex = extras();
  beginTimed();
  switch (species_stride_) {
#define BEAR_TRANSMISSION(NS)                                                                                       \
  case NS:                                                                                                          \
    if (hasExtras(ex))                                                                                              \
      hipLaunchKernelGGL((transmissionKernel<NS, true>), dim3(blocks), dim3(kTransmissionBlock), lds, s, tables_,  \
                         stride_, rayleigh_, wavelength_, ex, lev, spc, groups_, path, L, n, tile, bottom_radius,  \
                         star_radius, buf.flux, spectrum_dev);                                                      \
    else                                                                                                            \
      hipLaunchKernelGGL((transmissionKernel<NS, false>), dim3(blocks), dim3(kTransmissionBlock), lds, s, tables_, \
                         stride_, rayleigh_, wavelength_, ex, lev, spc, groups_, path, L, n, tile, bottom_radius,  \
                         star_radius, buf.flux, spectrum_dev);                                                      \
    break;
    BEAR_TRANSMISSION(4)
    BEAR_TRANSMISSION(8)
    BEAR_TRANSMISSION(12)
    BEAR_TRANSMISSION(16)
#undef BEAR_TRANSMISSION
    default:
      raise(kInvalidInput, "unsupported number of species for transmission");
  }
  check(hipGetLastError(), "transmission kernel");
  endTimed();
  launchBins(0, 0.0, 0.0, inflation, bins_dev);
}

double DeviceEngine::logLikelihood() {
  hipStream_t s = static_cast<hipStream_t>(stream_);
  check(hipStreamSynchronize(s), "evaluation");                      // sumKernel stored *ll_host_
  collectTiming();
  return n_bins_ > 0 ? *ll_host_ : 0.0;
}

void DeviceEngine::copyBins(std::vector<double>& out) {
  hipStream_t s = static_cast<hipStream_t>(stream_);
  out.resize(n_bins_);
  if (n_bins_) check(hipMemcpyAsync(out.data(), bins_, n_bins_ * sizeof(double), hipMemcpyDeviceToHost, s), "read bins");
  check(hipStreamSynchronize(s), "evaluation");
  collectTiming();
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
  collectTiming();
}

}  // namespace bear_amd
