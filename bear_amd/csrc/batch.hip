// Batched likelihood evaluation on gfx950: many live points of one retrieval
// per launch sequence.
//
// One work item is (wavenumber tile, live point).  The spectrum kernels are
// persistent: a grid sized by the autotuner walks the items, ordered so that
// every live point of one tile runs on the same XCD (workgroups go to XCDs
// round-robin by id, and the grid is a multiple of 8), where the tile's
// table rows stay in that XCD's L2 while the points reuse them.
//
// Per batch, in one hipGraph when enabled: the pinned per-point records go to
// the device in one copy; (transmission) the chord lengths of every point are
// computed; the spectrum kernel writes [point][wavenumber] FP64; one
// workgroup per (bin, point) folds the observation operator and the Gaussian
// term; one workgroup per point sums its terms; the log-likelihoods come back
// in one copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>

#include "device_math.h"
#include "tile_window.h"
#include "engine.h"

namespace bear_amd {

using namespace dev;

namespace {

void check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  const int code = (e == hipErrorOutOfMemory) ? kOutOfMemory : kDevice;
  raise(code, std::string(what) + ": " + hipGetErrorString(e));
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

constexpr int kBatchBlock = 256;
constexpr int kTransmissionBlock = 512;          // 8 wavefronts share one tile's extinction in LDS
constexpr float kLog2TenF = 3.32192809488736234787f;

// ---------------------------------------------------------------- storage types
// raw(): the stored value as loaded; value(): its FP32 value
struct TabF32 {
  using T = float;
  using R = float;
  static __device__ __forceinline__ R raw(const T* p) { return *p; }
  static __device__ __forceinline__ float value(R r) { return r; }
  static __device__ __forceinline__ T store(double v) { return float(v); }
  static constexpr bool kScaled = false;
};
struct TabF16 {
  using T = _Float16;
  using R = _Float16;
  static __device__ __forceinline__ R raw(const T* p) { return *p; }
  static __device__ __forceinline__ float value(R r) { return float(r); }
  static __device__ __forceinline__ T store(double v) { return _Float16(float(v)); }
  static constexpr bool kScaled = false;
};
struct TabBF16 {
  using T = uint16_t;
  using R = float;                                 // widened when a row is read
  static __device__ __forceinline__ R raw(const T* p) { return __uint_as_float(uint32_t(*p) << 16); }
  static __device__ __forceinline__ float value(R r) { return r; }
  static __device__ __forceinline__ T store(double v) {   // round to nearest even
    const uint32_t u = __float_as_uint(float(v));
    return uint16_t((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
  }
  static constexpr bool kScaled = false;
};
// 16-bit integers q in [-32767, 32767] with value = base + scale q, base and
// scale per (species, wavenumber): the 16 bits cover each wavenumber's own
// range uniformly (an absolute error of at most range / 131068), where FP16
// residuals spend bits on exponents.  The base array then holds per species
// slot the base row and the scale row.
struct TabS16 {
  using T = int16_t;
  using R = float;                                 // scale q, formed when a row is read
  static __device__ __forceinline__ float value(R r) { return r; }
  static constexpr bool kScaled = true;
  static constexpr int kMax = 32767;
};

// One species' FP64 rows [0, count) (row-major, `stride` apart) at
// wavenumber i: base = FP32 midrange of the values, residual = value - base
// into the species' rows of `low` (same layout).  Nothing is floored:
// log-linear interpolation between a row at the reference's 1e-200 floor and
// a row far above it yields values in between that can dominate a level's
// opacity.
template <class Tab>
__global__ __launch_bounds__(kBatchBlock) void splitTableKernel(const double* __restrict__ rows, size_t stride, int n,
                                                                uint32_t count, int slot, float* __restrict__ base,
                                                                typename Tab::T* __restrict__ low) {
  const int i = blockIdx.x * kBatchBlock + threadIdx.x;
  if (i >= n) return;
  double hi = -INFINITY, lo = INFINITY;
  for (uint32_t r = 0; r < count; ++r) {
    const double v = rows[size_t(r) * stride + i];
    hi = fmax(hi, v);
    lo = fmin(lo, v);
  }
  const float b = float(0.5 * (hi + lo));
  if constexpr (Tab::kScaled) {
    // scale: the smallest FP32 value with kMax scale >= the largest residual
    const double reach = fmax(hi - double(b), double(b) - lo);
    float sc = reach > 0.0 ? float(reach / Tab::kMax) : 1.0f;
    if (double(sc) * Tab::kMax < reach) sc = __uint_as_float(__float_as_uint(sc) + 1u);   // next FP32 up
    base[size_t(2 * slot) * stride + i] = b;
    base[size_t(2 * slot + 1) * stride + i] = sc;
    for (uint32_t r = 0; r < count; ++r) {
      const double q = rint((rows[size_t(r) * stride + i] - double(b)) / double(sc));
      low[size_t(r) * stride + i] = typename Tab::T(fmin(fmax(q, -double(Tab::kMax)), double(Tab::kMax)));
    }
  } else {
    base[size_t(slot) * stride + i] = b;
    for (uint32_t r = 0; r < count; ++r)
      low[size_t(r) * stride + i] = Tab::store(rows[size_t(r) * stride + i] - double(b));
  }
}

// ---------------------------------------------------------------- opacity per level
// FP64 tables and arithmetic: the single-model path's interpolation.
template <int NS>
struct Opacity64 {
  static constexpr int kNS = NS;
  using Species = LevelSpecies;
  using K = double;
  const double* tab;
  size_t stride;
  double v[NS][4];
  __device__ __forceinline__ void init(const void* tables, const float*, size_t s, int i, int) {
    tab = static_cast<const double*>(tables) + i;
    stride = s;
  }
  __device__ __forceinline__ double kappa(const LevelSpecies* d, const LevelState& lev, StencilGroups groups,
                                          bool reload) {
    return kappaTerms<NS>(tab, stride, d, lev, groups, reload, v);
  }
};

// Reduced-precision tables: x = log2(10) base + log2 n + sum_k w_k v_k in
// FP32 and sigma n = 2^x with the hardware exponential.
template <int NS, class Tab>
struct Opacity32 {
  static constexpr int kNS = NS;
  using Species = LevelSpecies32;
  using K = float;
  const typename Tab::T* tab;
  size_t stride;
  float bz[NS];
  float sc[Tab::kScaled ? NS : 1];                 // scales of TabS16
  typename Tab::R v[NS][4];       // as stored; converted where used (FP16 folds into v_fma_mix_f32)
  // slot0: the first stencil slot this thread handles (its base row)
  __device__ __forceinline__ void init(const void* tables, const float* base, size_t s, int i, int slot0) {
    tab = static_cast<const typename Tab::T*>(tables) + i;
    stride = s;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      if constexpr (Tab::kScaled) {
        bz[q] = base[size_t(2 * (slot0 + q)) * s + i] * kLog2TenF;
        sc[q] = base[size_t(2 * (slot0 + q) + 1) * s + i];
      } else {
        bz[q] = base[size_t(slot0 + q) * s + i] * kLog2TenF;
      }
    }
  }
  // x = bz + log2 n + sum_k w_k v_k (TabS16: v = scale q, formed at the row's read)
  __device__ __forceinline__ float exponent(int s, float start, const float* w) const {
    float x = fmaf(w[0], Tab::value(v[s][0]), start);
    x = fmaf(w[1], Tab::value(v[s][1]), x);
    x = fmaf(w[2], Tab::value(v[s][2]), x);
    return fmaf(w[3], Tab::value(v[s][3]), x);
  }
  __device__ __forceinline__ typename Tab::R load(const typename Tab::T* p, int s) const {
    if constexpr (Tab::kScaled) return sc[s] * float(*p);
    else return Tab::raw(p);
  }
  // grouped stencils: the level's shared weights and contiguous densities
  // (FP32 copies), as kappaTerms does for the FP64 tables
  __device__ __forceinline__ float kappa(const LevelSpecies32* d, const LevelState& lev, StencilGroups groups,
                                         bool reload) {
    if (reload) {
#pragma unroll
      for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int q = 0; q < 4; ++q) v[s][q] = load(tab + (size_t(d[s].block[q]) << 6), s);
    }
    float k = 0.0f;
    if (groups.grouped) {
      float a[4], b[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        a[q] = lev.group_w32[0][q];
        b[q] = lev.group_w32[1][q];
      }
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const bool one = (groups.group1 >> s) & 1u;
        const float w[4] = {one ? b[0] : a[0], one ? b[1] : a[1], one ? b[2] : a[2], one ? b[3] : a[3]};
        k += __builtin_amdgcn_exp2f(exponent(s, bz[s] + lev.log2_n32[s], w));
      }
    } else {
#pragma unroll
      for (int s = 0; s < NS; ++s) k += __builtin_amdgcn_exp2f(exponent(s, bz[s] + d[s].log2_n, d[s].w));
    }
    return k;
  }
};

// ---------------------------------------------------------------- emission
// Fused opacity + two-angle short characteristics per (wavenumber, point),
// as the single-point kernel; a work item is blockDim wavenumbers of one
// point and the flux F = 2 pi sum mu w I 1e-3 scaling goes to flux[point][i].
// With `keep`, both angles' intensity less the level's Planck function,
// I - B, and the opacity at every level also go to
// keep (keptColumn layout) in FP32, for the adjoint's backward sweep
// (which recomputes B in FP64: the difference it needs keeps its own
// relative precision, and the gradient's bar is 1e-4).
// EX (FP64 tables): the H- continuum joins each level's opacity and the
// point's cloud models (BatchModel::clouds, weights in its LevelState) each
// layer's optical depth, as in the single-point kernel.
template <class Op, bool EX = false>
__global__ __launch_bounds__(kBatchBlock) void emissionBatchKernel(
    const double* __restrict__ wavenumber, const double* __restrict__ wavelength, const void* __restrict__ tables,
    const float* __restrict__ base, size_t stride, Extras ex, Records rec, int nb_levels, int n, int tiles, int batch,
    double* __restrict__ flux, float* __restrict__ keep) {
  using Species = typename Op::Species;
  const size_t total = nbItems(tiles, batch);
  for (size_t item = blockIdx.x; item < total; item += gridDim.x) {
    int m, tile;
    if (!workItem(item, batch, tiles, m, tile)) continue;
    const int i = tile * int(blockDim.x) + int(threadIdx.x);
    if (i >= n) continue;
    const unsigned char* r = rec.at(m);
    const BatchModel* bm = reinterpret_cast<const BatchModel*>(r);
    const LevelState* levels = reinterpret_cast<const LevelState*>(r + rec.levels);
    const Species* species = reinterpret_cast<const Species*>(r + rec.species);
    Op op;
    op.init(tables, base, stride, i, 0);
    const double nu = wavenumber[i];
    const double b_num = kPlanckC1 * nu * nu * nu;
    const double x_nu = -(6.62606896e-27 * nu * 2.99792458e10 / 1.3806504e-16) * kLog2E;
    Extras e = ex;
    double hm[7], cf[kMaxClouds];
    if constexpr (EX) {
      e.clouds = bm->clouds;
      loadExtras(e, i, n, wavelength[i], hm, cf);
    }
    double k_prev = double(op.kappa(species, levels[0], rec.groups, true));
    if (EX && e.hminus) k_prev += hminusKappa(hm, levels[0]);
    double b_prev = planckAt(b_num, x_nu, levels[0].inv_temperature);
    double z_prev = levels[0].altitude;
    double i1 = b_prev, i2 = b_prev;
    float* kept = keep ? keep + keptColumn(m, stride, nb_levels, i) : nullptr;
    if (kept) {
      kept[0] = 0.0f;                                      // I = B at the bottom
      kept[kKeptValue] = 0.0f;
      kept[2 * kKeptValue] = float(k_prev);
    }
    for (int l = 1; l < nb_levels; ++l) {
      double k_next = double(op.kappa(species + size_t(l) * Op::kNS, levels[l], rec.groups, levels[l].reload != 0));
      if (EX && e.hminus) k_next += hminusKappa(hm, levels[l]);
      const double b_next = planckAt(b_num, x_nu, levels[l].inv_temperature);
      const double z_next = levels[l].altitude;
      double tau = (z_next - z_prev) * (k_next + k_prev) * 0.5;
      if (EX) tau += cloudLayer(e, cf, levels[l], l, i, n);
      shortCharacteristicsLayer(tau, b_prev, b_next, i1, i2);
      if (kept) {
        float* at = kept + size_t(l) * kKeptLevel;
        at[0] = float(i1 - b_next);
        at[kKeptValue] = float(i2 - b_next);
        at[2 * kKeptValue] = float(k_next);
      }
      k_prev = k_next;
      b_prev = b_next;
      z_prev = z_next;
    }
    flux[size_t(m) * stride + i] = 2.0 * kPi * (i1 * kMu1 * kW + i2 * kMu2 * kW) * 1e-3 * bm->scaling;
  }
}

// ---------------------------------------------------------------- transmission
// Chord lengths of every point, transposed (chord[point][q][t], zero for
// q < t and in the padding): sqrt((R+z_{q+1})^2 - (R+z_t)^2) -
// sqrt((R+z_q)^2 - (R+z_t)^2), q >= t (transmission_spectrum.cpp:117-118,
// 145-154); one workgroup per (segment q, point), FP64 and/or FP32.
// Evaluated without contraction, as the host does: fused, a*a - b*b at a == b
// leaves the rounding error of b*b, which can be negative.
__global__ __launch_bounds__(kBatchBlock) void chordKernel(Records rec, int nb_levels, double* __restrict__ chord,
                                                           float* __restrict__ chord32) {
#pragma clang fp contract(off)
  const int q = blockIdx.x, m = blockIdx.y;
  const int tangents = chordTangents(nb_levels);
  const unsigned char* r = rec.at(m);
  const double r0 = reinterpret_cast<const BatchModel*>(r)->bottom_radius;
  const LevelState* levels = reinterpret_cast<const LevelState*>(r + rec.levels);
  const size_t row = (size_t(m) * gridDim.x + q) * tangents;
  for (int t = int(threadIdx.x); t < tangents; t += kBatchBlock) {
    double c = 0.0;
    if (q < nb_levels - 1 && t <= q) {
      const double b = r0 + levels[t].altitude;
      const double a1 = r0 + levels[q + 1].altitude, a0 = r0 + levels[q].altitude;
      c = sqrt(a1 * a1 - b * b) - sqrt(a0 * a0 - b * b);
    }
    if (chord) chord[row + t] = c;
    if (chord32) chord32[row + t] = float(c);
  }
}

__device__ __forceinline__ double transmittance(double tau) { return pow2(-tau * kLog2E); }
__device__ __forceinline__ double transmittance(float tau) {
  return double(__builtin_amdgcn_exp2f(-tau * 1.44269504088896340736f));
}

// Transit depth per (wavenumber, point) (transmission_spectrum.cpp:32-154).
// One work item = `tile` wavenumbers of one point:
//  1. extinction k_l (Op::K: FP64 or FP32), threads split over level ranges
//     with register-cached stencils, stored as the segment sums
//     S[q] = k_q + k_{q+1} in LDS (across a group boundary after the first
//     barrier, from the next group's first level);
//  2. blocks of eight tangent levels per thread group (tangentBlock: tau
//     summed outward in the reference's order, stopping at the first sum
//     above 10); each path adds c_t (1 - e^-tau) to its thread's FP64
//     partial, where c_t = (R+z_t)(z_{t+1} - z_{t-1}) is the coefficient of
//     (1 - T_t) in the reference's trapezoid (one-sided at t = 0; T = 1 at
//     the top level);
//  3. per wavenumber, the partials in a fixed order, h = sqrt(H + R^2) - R
//     and depth (R + h)^2 / R_*^2 1e6 in FP64.
// kKeep (FP64 tables, for the adjoint): also, in the keptColumn layout,
// {k_l, tau_t, last segment summed by path t} per level in FP32.
// EX (FP64 tables): the H- continuum and the point's cloud extinction
// (BatchModel::clouds, the layer below each level over its thickness) join
// each level's extinction, as in the single-point kernel.
template <class Op, bool kKeep = false, bool EX = false>
__global__ __launch_bounds__(kTransmissionBlock) void transmissionBatchKernel(
    const void* __restrict__ tables, const float* __restrict__ base, size_t stride, const double* __restrict__ rayleigh,
    const double* __restrict__ wavelength, Extras ex, Records rec, const double* __restrict__ chord64,
    const float* __restrict__ chord32, int nb_levels, int n, TileWindow win, int tile, int batch, int point_group,
    double* __restrict__ depth, float* __restrict__ keep, int* __restrict__ reach) {
  using K = typename Op::K;
  using Species = typename Op::Species;
  extern __shared__ double lds[];
  const int L = nb_levels;
  const int segments = chordSegments(L), tangents = chordTangents(L);
  K* S = reinterpret_cast<K*>(lds);
  double* part = lds + (size_t(segments) * tile * sizeof(K) + 7) / 8;   // also the group-boundary levels
  const int groups = kTransmissionBlock / tile;
  const int j = int(threadIdx.x) % tile;
  const int g = __builtin_amdgcn_readfirstlane(int(threadIdx.x) / tile);   // tile: a multiple of 64
  const int chunk = (L + groups - 1) / groups;
  const int l0 = g * chunk;
  const int l1 = min(L, l0 + chunk);
  const size_t total = nbItems(win.count, batch);
  for (size_t item = blockIdx.x; item < total; item += gridDim.x) {
    int m, tl;
    if (!windowItem(item, batch, win, m, tl, point_group)) continue;      // uniform over the workgroup
    const unsigned char* r = rec.at(m);
    const BatchModel* bm = reinterpret_cast<const BatchModel*>(r);
    const LevelState* levels = reinterpret_cast<const LevelState*>(r + rec.levels);
    const Species* species = reinterpret_cast<const Species*>(r + rec.species);
    const int i = tl * tile + j;
    K last = K(0);
    if (i < n) {
      Op op;
      op.init(tables, base, stride, i, 0);
      const double sigma_r = rayleigh ? rayleigh[i] : 0.0;
      Extras e = ex;
      double hm[7], cf[kMaxClouds];
      if constexpr (EX) {
        e.clouds = bm->clouds;
        loadExtras(e, i, n, wavelength[i], hm, cf);
      }
      for (int l = l0; l < l1; ++l) {
        const double kd = double(op.kappa(species + size_t(l) * Op::kNS, levels[l], rec.groups,
                                          l == l0 || levels[l].reload != 0));
        double ke = kd + sigma_r * levels[l].rayleigh_density;
        if constexpr (EX) {
          if (e.hminus) ke += hminusKappa(hm, levels[l]);
          if (l > 0) ke += cloudLayer(e, cf, levels[l], l, i, n) * levels[l].cloud_scale;
        }
        const K k = K(ke);
        if constexpr (kKeep) keep[keptAt(win, m, stride, L, i) + size_t(l) * kKeptLevel] = float(k);
        if (l == l0) part[g * tile + j] = double(k);
        else S[(l - 1) * tile + j] = last + k;
        last = k;
      }
    } else {
      for (int l = l0; l < l1; ++l) {
        if (l == l0) part[g * tile + j] = 0.0;
        else S[(l - 1) * tile + j] = K(0);
      }
    }
    __syncthreads();
    if (l0 < l1 && l1 < L) S[(l1 - 1) * tile + j] = last + K(part[(g + 1) * tile + j]);
    if (g == 0)
      for (int q = L - 1; q < segments; ++q) S[q * tile + j] = K(0);
    __syncthreads();
    const double r0 = bm->bottom_radius;
    const K* chords;
    if constexpr (sizeof(K) == 8) chords = chord64;
    else chords = chord32;
    chords += size_t(m) * segments * tangents;
    double h = 0.0;
    for (int t0 = kTangentBlock * g; t0 < tangents; t0 += kTangentBlock * groups) {
      K tau[kTangentBlock];
      if constexpr (kKeep) {
        int stop[kTangentBlock];
        tangentBlock<K, true>(S + j, tile, chords, tangents, t0, L - 1, tau, stop);
        // per tile of 64 wavenumbers (a wavefront) and path: the largest
        // last segment, with which the adjoint passes over paths and steps
        // that end below its segments
        const int first64 = tl * tile + (j & ~63);
        int* rl = reach + (size_t(m) * ((n + 63) / 64) + first64 / 64) * (L - 1);
#pragma unroll
        for (int k = 0; k < kTangentBlock; ++k) {
          const int t = t0 + k;
          if (t < L - 1) {                                                 // uniform
            const int mx = waveMaxInt(i < n ? stop[k] : -1);
            if ((j & 63) == 0 && first64 < n) rl[t] = mx;
          }
        }
        if (i < n)
#pragma unroll
          for (int k = 0; k < kTangentBlock; ++k) {
            const int t = t0 + k;
            if (t < L - 1) {
              float* kt = keep + keptAt(win, m, stride, L, i) + size_t(t) * kKeptLevel;
              kt[kKeptValue] = float(tau[k]);
              kt[2 * kKeptValue] = float(stop[k]);
            }
          }
      } else {
        tangentBlock<K>(S + j, tile, chords, tangents, t0, L - 1, tau);
      }
#pragma unroll
      for (int k = 0; k < kTangentBlock; ++k) {
        const int t = t0 + k;
        if (t < L - 1) {
          const double zl = levels[t > 0 ? t - 1 : 0].altitude;
          const double zu = levels[t + 1].altitude;
          h += (r0 + levels[t].altitude) * (zu - zl) * (1.0 - transmittance(tau[k]));
        }
      }
    }
    part[g * tile + j] = h;
    __syncthreads();
    if (int(threadIdx.x) < tile && i < n) {
      double hs = 0.0;
      for (int q = 0; q < groups; ++q) hs += part[q * tile + j];
      const double hh = sqrt(hs + r0 * r0) - r0;
      const double rr = hh + r0;
      depth[size_t(m) * stride + i] = rr * rr / bm->star_radius / bm->star_radius * 1e6;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- bins and likelihood
// One workgroup per (piece of a bin's segment, point) (binChunkBody); mode 1
// (eclipse) also sums the folded blackbody star of that point.
__global__ __launch_bounds__(kBlockBins) void binBatchKernel(
    const double* __restrict__ flux, size_t stride, const double* __restrict__ wavenumber,
    const double* __restrict__ weights, const BinChunk* __restrict__ chunks, int nb_chunks, Records rec, int mode,
    double* __restrict__ chunk_sums) {
  const int m = blockIdx.y;
  const BatchModel* bm = reinterpret_cast<const BatchModel*>(rec.at(m));
  binChunkBody(flux + size_t(m) * stride, wavenumber, weights, chunks[blockIdx.x], mode, bm->star_inv_temperature,
               chunk_sums + (size_t(m) * nb_chunks + blockIdx.x) * 2);
}

// Bins and log-likelihood of one point per workgroup (binFinalizeBody)
__global__ __launch_bounds__(1024) void sumBatchKernel(const double* __restrict__ chunk_sums, int nb_chunks,
                                                       const BinDesc* __restrict__ bins, int nb_bins, Records rec,
                                                       int mode, double* __restrict__ out, double* __restrict__ ll) {
  const int m = blockIdx.x;
  const unsigned char* r = rec.at(m);
  const BatchModel* bm = reinterpret_cast<const BatchModel*>(r);
  const double* shifts = reinterpret_cast<const double*>(r + rec.shifts);
  binFinalizeBody(chunk_sums + size_t(m) * nb_chunks * 2, bins, nb_bins, mode, bm->depth_scale, shifts,
                  bm->inflation, out + size_t(m) * nb_bins, ll + m, nullptr);
}

// high-resolution output per point: emission F per micron, eclipse depth
// F / F_* (Rp/Rs)^2 1e6, transmission depth
__global__ __launch_bounds__(kBatchBlock) void batchSpectrumKernel(
    const double* __restrict__ flux, size_t stride, const double* __restrict__ wavenumber,
    const double* __restrict__ wavelength, Records rec, int mode, int n, double* __restrict__ out) {
  const int i = blockIdx.x * kBatchBlock + threadIdx.x;
  const int m = blockIdx.y;
  if (i >= n) return;
  const BatchModel* bm = reinterpret_cast<const BatchModel*>(rec.at(m));
  const double f = flux[size_t(m) * stride + i];
  double v = f;
  if (mode == 0) {
    const double lam = wavelength[i];
    v = f / lam / lam * 10000.0;
  } else if (mode == 1) {
    v = f * recip(starFlux(wavenumber[i], bm->star_inv_temperature)) * bm->depth_scale;
  }
  out[size_t(m) * n + i] = v;
}

// ---------------------------------------------------------------- launch helpers
// ex: the H- rows and whether the points carry cloud models (the extras
// variants, FP64 tables only; the callers refuse the others)
template <int NS>
void launchEmission(TablePrecision p, dim3 grid, dim3 block, hipStream_t s, const double* nu, const double* lam,
                    const void* tab64, const void* tab_low, const float* base, size_t stride, const Extras& ex,
                    bool extras, Records rec, int L, int n, int tiles, int batch, double* flux, float* keep) {
  switch (p) {
    case TablePrecision::fp64:
      if (extras)
        hipLaunchKernelGGL((emissionBatchKernel<Opacity64<NS>, true>), grid, block, 0, s, nu, lam, tab64, base, stride,
                           ex, rec, L, n, tiles, batch, flux, keep);
      else
        hipLaunchKernelGGL(emissionBatchKernel<Opacity64<NS>>, grid, block, 0, s, nu, lam, tab64, base, stride, ex,
                           rec, L, n, tiles, batch, flux, keep);
      break;
    case TablePrecision::fp32:
      hipLaunchKernelGGL((emissionBatchKernel<Opacity32<NS, TabF32>>), grid, block, 0, s, nu, lam, tab_low, base,
                         stride, ex, rec, L, n, tiles, batch, flux, keep);
      break;
    case TablePrecision::fp16:
      hipLaunchKernelGGL((emissionBatchKernel<Opacity32<NS, TabF16>>), grid, block, 0, s, nu, lam, tab_low, base,
                         stride, ex, rec, L, n, tiles, batch, flux, keep);
      break;
    case TablePrecision::bf16:
      hipLaunchKernelGGL((emissionBatchKernel<Opacity32<NS, TabBF16>>), grid, block, 0, s, nu, lam, tab_low, base,
                         stride, ex, rec, L, n, tiles, batch, flux, keep);
      break;
    case TablePrecision::int16:
      hipLaunchKernelGGL((emissionBatchKernel<Opacity32<NS, TabS16>>), grid, block, 0, s, nu, lam, tab_low, base,
                         stride, ex, rec, L, n, tiles, batch, flux, keep);
      break;
  }
}

template <int NS>
void launchTransmission(TablePrecision p, dim3 grid, size_t lds, hipStream_t s, const void* tab64,
                        const void* tab_low, const float* base, size_t stride, const double* rayleigh,
                        const double* lam, const Extras& ex, bool extras, Records rec, const double* chord,
                        const float* chord32, int L, int n, TileWindow win, int tile, int batch, int group,
                        double* depth, float* keep, int* reach) {
  const dim3 block(kTransmissionBlock);
  switch (p) {
    case TablePrecision::fp64:
      if (keep && extras)
        hipLaunchKernelGGL((transmissionBatchKernel<Opacity64<NS>, true, true>), grid, block, lds, s, tab64, base,
                           stride, rayleigh, lam, ex, rec, chord, chord32, L, n, win, tile, batch, group, depth, keep,
                           reach);
      else if (keep)
        hipLaunchKernelGGL((transmissionBatchKernel<Opacity64<NS>, true>), grid, block, lds, s, tab64, base, stride,
                           rayleigh, lam, ex, rec, chord, chord32, L, n, win, tile, batch, group, depth, keep, reach);
      else if (extras)
        hipLaunchKernelGGL((transmissionBatchKernel<Opacity64<NS>, false, true>), grid, block, lds, s, tab64, base,
                           stride, rayleigh, lam, ex, rec, chord, chord32, L, n, win, tile, batch, group, depth,
                           nullptr, nullptr);
      else
        hipLaunchKernelGGL(transmissionBatchKernel<Opacity64<NS>>, grid, block, lds, s, tab64, base, stride, rayleigh,
                           lam, ex, rec, chord, chord32, L, n, win, tile, batch, group, depth, nullptr, nullptr);
      break;
    case TablePrecision::fp32:
      hipLaunchKernelGGL((transmissionBatchKernel<Opacity32<NS, TabF32>>), grid, block, lds, s, tab_low, base,
                         stride, rayleigh, lam, ex, rec, chord, chord32, L, n, win, tile, batch, group, depth, nullptr,
                         nullptr);
      break;
    case TablePrecision::fp16:
      hipLaunchKernelGGL((transmissionBatchKernel<Opacity32<NS, TabF16>>), grid, block, lds, s, tab_low, base,
                         stride, rayleigh, lam, ex, rec, chord, chord32, L, n, win, tile, batch, group, depth, nullptr,
                         nullptr);
      break;
    case TablePrecision::bf16:
      hipLaunchKernelGGL((transmissionBatchKernel<Opacity32<NS, TabBF16>>), grid, block, lds, s, tab_low, base,
                         stride, rayleigh, lam, ex, rec, chord, chord32, L, n, win, tile, batch, group, depth, nullptr,
                         nullptr);
      break;
    case TablePrecision::int16:
      hipLaunchKernelGGL((transmissionBatchKernel<Opacity32<NS, TabS16>>), grid, block, lds, s, tab_low, base,
                         stride, rayleigh, lam, ex, rec, chord, chord32, L, n, win, tile, batch, group, depth, nullptr,
                         nullptr);
      break;
  }
}

size_t align(size_t x, size_t a) { return (x + a - 1) / a * a; }

size_t transmissionLds(int L, int tile, bool wide) {
  const size_t k = size_t(chordSegments(L)) * tile * (wide ? 8 : 4);
  return align(k, 8) + size_t(kTransmissionBlock / tile) * tile * sizeof(double);
}

}  // namespace

// ---------------------------------------------------------------- precision
// residual tables [rows][stride] and the base [slots][stride], zero base in
// the padding slots (their density is zero)
void DeviceEngine::allocateReduced(TablePrecision p, size_t nb_rows, size_t slots) {
  void* low = nullptr;
  check(hipMalloc(&low, std::max<size_t>(1, nb_rows) * stride_ * tableElementBytes(p)), "reduced-precision tables");
  tables_low_ = low;
  base_rows_ = slots * (p == TablePrecision::int16 ? 2 : 1);          // int16: base and scale per slot
  table_base_ = deviceAlloc<float>(base_rows_ * stride_, "table base");
  check(hipMemsetAsync(table_base_, 0, base_rows_ * stride_ * sizeof(float), static_cast<hipStream_t>(stream_)),
        "clear table base");
}

// species `slot`'s FP64 rows (device, [count][stride]) split into the base and
// the residual rows [first, first + count)
void DeviceEngine::splitSpecies(const double* rows, size_t count, int slot, size_t first) {
  hipStream_t s = static_cast<hipStream_t>(stream_);
  const int n = int(n_points_);
  const dim3 grid((n + kBatchBlock - 1) / kBatchBlock), block(kBatchBlock);
  const uint32_t c = uint32_t(count);
  switch (precision_) {
    case TablePrecision::fp32:
      hipLaunchKernelGGL(splitTableKernel<TabF32>, grid, block, 0, s, rows, stride_, n, c, slot, table_base_,
                         static_cast<float*>(tables_low_) + first * stride_);
      break;
    case TablePrecision::fp16:
      hipLaunchKernelGGL(splitTableKernel<TabF16>, grid, block, 0, s, rows, stride_, n, c, slot, table_base_,
                         static_cast<_Float16*>(tables_low_) + first * stride_);
      break;
    case TablePrecision::bf16:
      hipLaunchKernelGGL(splitTableKernel<TabBF16>, grid, block, 0, s, rows, stride_, n, c, slot, table_base_,
                         static_cast<uint16_t*>(tables_low_) + first * stride_);
      break;
    case TablePrecision::int16:
      hipLaunchKernelGGL(splitTableKernel<TabS16>, grid, block, 0, s, rows, stride_, n, c, slot, table_base_,
                         static_cast<int16_t*>(tables_low_) + first * stride_);
      break;
    default:
      raise(kInternal, "no split for FP64 tables");
  }
  check(hipGetLastError(), "table split kernel");
}

void DeviceEngine::setTablePrecision(TablePrecision p,
                                     const std::vector<std::pair<uint32_t, uint32_t>>& species_rows) {
  check(hipSetDevice(device_), "hipSetDevice");
  if (storage_ != TablePrecision::fp64) {
    if (p != storage_)
      raise(kInvalidInput, "the tables were loaded at reduced precision only; no other precision is available");
    return;
  }
  releaseGraph();
  deviceFree(tables_low_);
  deviceFree(table_base_);
  tables_low_ = nullptr;
  table_base_ = nullptr;
  precision_ = p;
  batch_capacity_ = 0;                                  // record layout depends on the precision
  tuned_.clear();
  if (p == TablePrecision::fp64) return;
  if (species_rows.size() > species_stride_) raise(kInvalidInput, "more species than stencil slots");
  allocateReduced(p, n_rows_, species_stride_);
  for (size_t q = 0; q < species_rows.size(); ++q) {
    const uint32_t r0 = species_rows[q].first, r1 = species_rows[q].second;
    if (r1 <= r0 || r1 > n_rows_) raise(kInternal, "species rows out of range");
    splitSpecies(tables_ + size_t(r0) * stride_, r1 - r0, int(q), r0);
  }
  check(hipStreamSynchronize(static_cast<hipStream_t>(stream_)), "table split");
}

// ---------------------------------------------------------------- staging
LevelSpecies* DeviceEngine::batchSpecies(size_t b) {
  if (precision_ != TablePrecision::fp64) raise(kInternal, "FP64 stencils requested with reduced-precision tables");
  return reinterpret_cast<LevelSpecies*>(record(b) + rec_species_);
}

LevelSpeciesAdjoint* DeviceEngine::batchSpeciesAdjoint(size_t b) {
  if (precision_ != TablePrecision::fp64) raise(kInternal, "adjoint stencils requested with reduced-precision tables");
  return reinterpret_cast<LevelSpeciesAdjoint*>(record(b) + rec_adjoint_);
}

LevelSpecies32* DeviceEngine::batchSpecies32(size_t b) {
  if (precision_ == TablePrecision::fp64) raise(kInternal, "FP32 stencils requested with FP64 tables");
  return reinterpret_cast<LevelSpecies32*>(record(b) + rec_species_);
}

void DeviceEngine::reserveBatch(size_t batch) {
  if (batch == 0) raise(kInvalidInput, "empty batch");
  if (batch > 65535) raise(kInvalidInput, "at most 65535 live points per batch");
  if (batch <= batch_capacity_) return;
  check(hipSetDevice(device_), "hipSetDevice");
  check(hipStreamSynchronize(static_cast<hipStream_t>(stream_)), "batch");
  releaseGraph();
  const size_t L = nb_levels_, NS = species_stride_;
  const size_t spc = precision_ == TablePrecision::fp64 ? sizeof(LevelSpecies) : sizeof(LevelSpecies32);
  const bool wide = precision_ == TablePrecision::fp64;
  rec_levels_ = align(sizeof(BatchModel), 64);
  rec_species_ = align(rec_levels_ + L * sizeof(LevelState), 64);
  rec_adjoint_ = wide ? align(rec_species_ + L * NS * spc, 64) : 0;
  rec_shifts_ = align(wide ? rec_adjoint_ + L * NS * sizeof(LevelSpeciesAdjoint) : rec_species_ + L * NS * spc, 64);
  rec_bytes_ = align(rec_shifts_ + (n_obs_ + 1) * sizeof(double), 256);
  if (batch_host_) (void)hipHostFree(batch_host_);
  if (batch_ll_host_) (void)hipHostFree(batch_ll_host_);
  batch_host_ = nullptr;
  batch_ll_host_ = nullptr;
  for (void* p : {(void*)batch_dev_, (void*)batch_flux_, (void*)batch_bins_, (void*)batch_chunk_sums_, (void*)batch_ll_,
                  (void*)batch_chord_, (void*)batch_chord32_})
    deviceFree(p);
  batch_dev_ = nullptr;
  batch_flux_ = batch_bins_ = batch_chunk_sums_ = batch_ll_ = batch_chord_ = nullptr;
  batch_chord32_ = nullptr;
  batch_capacity_ = 0;
  void* h = nullptr;
  check(hipHostMalloc(&h, batch * rec_bytes_, hipHostMallocDefault), "pinned batch records");
  batch_host_ = static_cast<unsigned char*>(h);
  std::memset(batch_host_, 0, batch * rec_bytes_);
  check(hipHostMalloc(&h, batch * sizeof(double), hipHostMallocDefault), "pinned batch log-likelihoods");
  batch_ll_host_ = static_cast<double*>(h);
  batch_dev_ = deviceAlloc<unsigned char>(batch * rec_bytes_, "batch records");
  batch_flux_ = deviceAlloc<double>(batch * stride_, "batch spectra");
  batch_bins_ = deviceAlloc<double>(batch * std::max<size_t>(1, n_bins_), "batch bins");
  batch_chunk_sums_ = deviceAlloc<double>(batch * std::max<size_t>(1, n_chunks_) * 2, "batch bin piece sums");
  batch_ll_ = deviceAlloc<double>(batch, "batch log-likelihoods");
  if (transmission_) {
    const size_t chord_size = size_t(chordSegments(int(L))) * chordTangents(int(L));
    if (precision_ == TablePrecision::fp64) batch_chord_ = deviceAlloc<double>(batch * chord_size, "batch chord lengths");
    else batch_chord32_ = deviceAlloc<float>(batch * chord_size, "batch chord lengths");
  }
  batch_capacity_ = batch;
}

// ---------------------------------------------------------------- launches
void DeviceEngine::launchBatchSpectrum(ModelKind kind, size_t batch, const BatchLaunch& lc, size_t first,
                                       float* keep, size_t nu_first, size_t nu_count, int* reach) {
  hipStream_t s = static_cast<hipStream_t>(stream_);
  const Records rec{batch_dev_ + first * rec_bytes_, rec_bytes_, rec_levels_, rec_species_, rec_adjoint_,
                    rec_shifts_, groups_};
  double* flux = batch_flux_ + first * stride_;
  const int n = int(n_points_), L = int(nb_levels_), B = int(batch);
  const bool wide = precision_ == TablePrecision::fp64;
  const Extras ex{hminus_, nullptr, CloudSpec{}};       // the clouds' parameters come with each point's record
  const bool extras = hminus_ != nullptr || cloud_models_ > 0;
  if (kind == ModelKind::transmission) {
    hipLaunchKernelGGL(chordKernel, dim3(unsigned(chordSegments(L)), unsigned(B)), dim3(kBatchBlock), 0, s, rec, L,
                       wide ? batch_chord_ : nullptr, wide ? nullptr : batch_chord32_);
    check(hipGetLastError(), "chord kernel");
    const int tiles = (n + lc.tile - 1) / lc.tile;
    TileWindow win = wholeGrid(tiles);
    if (nu_count > 0) {                                   // wavenumbers [nu_first, nu_first + nu_count)
      if (nu_first % size_t(lc.tile) != 0 || nu_first >= size_t(n)) raise(kInternal, "misaligned wavenumber window");
      win.first = int(nu_first / size_t(lc.tile));
      win.count = int((std::min(size_t(n), nu_first + nu_count) + size_t(lc.tile) - 1) / size_t(lc.tile)) - win.first;
      win.kept_origin = nu_first / kKeptValue * size_t(L) * kKeptLevel;
    }
    const size_t lds = transmissionLds(L, lc.tile, wide);
    switch (species_stride_) {
#define BEAR_T(NS)                                                                                                  \
  case NS:                                                                                                          \
    launchTransmission<NS>(precision_, dim3(unsigned(lc.grid)), lds, s, tables_, tables_low_, table_base_, stride_, \
                           rayleigh_, wavelength_, ex, extras, rec, batch_chord_, batch_chord32_, L, n, win,       \
                           lc.tile, B, lc.group, flux, keep, reach);                                               \
    break;
      BEAR_T(4)
      BEAR_T(8)
      BEAR_T(12)
      BEAR_T(16)
#undef BEAR_T
      default:
        raise(kInvalidInput, "the batched kernels take at most 16 tabulated species");
    }
    check(hipGetLastError(), "transmission batch kernel");
    return;
  }
  const int tiles = (n + lc.block - 1) / lc.block;
  switch (species_stride_) {
#define BEAR_E(NS)                                                                                                   \
  case NS:                                                                                                           \
    launchEmission<NS>(precision_, dim3(unsigned(lc.grid)), dim3(unsigned(lc.block)), s, wavenumber_, wavelength_,   \
                       tables_, tables_low_, table_base_, stride_, ex, extras, rec, L, n, tiles, B, flux, keep);    \
    break;
    BEAR_E(4)
    BEAR_E(8)
    BEAR_E(12)
    BEAR_E(16)
#undef BEAR_E
    default:
      raise(kInvalidInput, "the batched kernels take at most 16 tabulated species");
  }
  check(hipGetLastError(), "emission batch kernel");
}

void DeviceEngine::launchBatchRest(ModelKind kind, size_t batch, size_t first) {
  hipStream_t s = static_cast<hipStream_t>(stream_);
  const Records rec{batch_dev_ + first * rec_bytes_, rec_bytes_, rec_levels_, rec_species_, rec_adjoint_,
                    rec_shifts_, groups_};
  if (n_bins_ > 0) {
    const int mode = kind == ModelKind::eclipse ? 1 : 0;
    double* sums = batch_chunk_sums_ + first * n_chunks_ * 2;
    hipLaunchKernelGGL(binBatchKernel, dim3(unsigned(n_chunks_), unsigned(batch)), dim3(kBlockBins), 0, s,
                       batch_flux_ + first * stride_, stride_, wavenumber_, weights_, bin_chunks_, int(n_chunks_), rec,
                       mode, sums);
    check(hipGetLastError(), "batch bin kernel");
    hipLaunchKernelGGL(sumBatchKernel, dim3(unsigned(batch)), dim3(1024), 0, s, sums, int(n_chunks_), bins_desc_,
                       int(n_bins_), rec, mode, batch_bins_ + first * n_bins_, batch_ll_ + first);
    check(hipGetLastError(), "batch likelihood kernel");
    check(hipMemcpyAsync(batch_ll_host_ + first, batch_ll_ + first, batch * sizeof(double), hipMemcpyDeviceToHost,
                         s),
          "read batch log-likelihoods");
  }
}

// Autotuner: times the spectrum kernel of the first min(batch, 64) staged
// points for each candidate launch shape and keeps the fastest.
//   emission / eclipse: workgroup of 64, 128 or 256 wavenumbers; persistent
//     grid of 8, 16 or 32 waves per CU
//   transmission: tile of 64 or 128 wavenumbers within the device's LDS per
//     workgroup; grid of 1, 2 or 4 workgroups per CU; points taken in groups
//     of 1 or 4 (each point brings its own chord matrix)
BatchLaunch DeviceEngine::tune(ModelKind kind, size_t batch) {
  hipStream_t s = static_cast<hipStream_t>(stream_);
  const size_t b = std::min<size_t>(batch, 64);
  std::vector<BatchLaunch> cand;
  if (kind == ModelKind::transmission) {
    const bool wide = precision_ == TablePrecision::fp64;
    for (int tile : {64, 128}) {
      if (transmissionLds(int(nb_levels_), tile, wide) > lds_limit_) continue;
      for (int group : {1, 4})
        for (int per_cu : {1, 2, 4})
          cand.push_back(BatchLaunch{kTransmissionBlock, tile, std::max(8, nb_cus_ * per_cu), group});
    }
    if (cand.empty()) raise(kInvalidInput, "too many levels for the transmission batch kernel");
  } else {
    for (int block : {64, 128, 256})
      for (int waves : {8, 16, 32}) cand.push_back(BatchLaunch{block, block, std::max(8, nb_cus_ * waves / (block / 64))});
  }
  for (auto& c : cand) c.grid = int(align(size_t(c.grid), 8));
  if (!autotune_) return cand[cand.size() / 2];
  check(hipMemcpyAsync(batch_dev_, batch_host_, b * rec_bytes_, hipMemcpyHostToDevice, s), "upload batch");
  hipEvent_t e0, e1;
  check(hipEventCreate(&e0), "hipEventCreate");
  check(hipEventCreate(&e1), "hipEventCreate");
  BatchLaunch best = cand.front();
  float best_ms = std::numeric_limits<float>::max();
  try {
    for (const auto& c : cand) {
      launchBatchSpectrum(kind, b, c);                    // warm-up
      check(hipEventRecord(e0, s), "hipEventRecord");
      launchBatchSpectrum(kind, b, c);
      check(hipEventRecord(e1, s), "hipEventRecord");
      check(hipEventSynchronize(e1), "tuning");
      float ms = 0;
      check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
      if (ms < best_ms) { best_ms = ms; best = c; }
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

// The launch shape of (kind, table precision, batch size): tuned on first use
BatchLaunch DeviceEngine::tuned(ModelKind kind, size_t batch) {
  const TuneKey key{int(kind), int(precision_), batch};
  for (auto& t : tuned_)
    if (t.first.kind == key.kind && t.first.precision == key.precision && t.first.batch == key.batch) {
      last_launch_ = t.second;
      return t.second;
    }
  const BatchLaunch launch = tune(kind, batch);
  tuned_.push_back({key, launch});
  last_launch_ = launch;
  return launch;
}

void DeviceEngine::releaseGraph() {
  if (graph_exec_) (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec_));
  graph_exec_ = nullptr;
  graph_key_ = TuneKey{-1, -1, 0};
}

void DeviceEngine::evaluateBatch(ModelKind kind, size_t batch) { evaluateBatch(kind, batch, false); }

void DeviceEngine::evaluateBatch(ModelKind kind, size_t batch, bool staged) {
  if (staged && (!stage_blob_ || batch > stage_capacity_)) raise(kInternal, "device staging not prepared");
  if (species_stride_ > 16) raise(kInvalidInput, "the batched kernels take at most 16 tabulated species");
  if ((hminus_ || cloud_models_ > 0) && precision_ != TablePrecision::fp64)
    raise(kInvalidInput, "the H- continuum and cloud layers take FP64 tables");
  if (batch == 0 || batch > batch_capacity_) raise(kInternal, "batch larger than the staged records");
  if ((kind == ModelKind::transmission) != transmission_) raise(kInvalidInput, "model kind does not match the engine");
  check(hipSetDevice(device_), "hipSetDevice");
  hipStream_t s = static_cast<hipStream_t>(stream_);
  const TuneKey key{int(kind), int(precision_), batch};
  auto same = [](const TuneKey& a, const TuneKey& b) {
    return a.kind == b.kind && a.precision == b.precision && a.batch == b.batch;
  };
  const BatchLaunch launch = tuned(kind, batch);
  auto sequence = [&] {
    if (staged) launchStage(batch);
    else check(hipMemcpyAsync(batch_dev_, batch_host_, batch * rec_bytes_, hipMemcpyHostToDevice, s), "upload batch");
    launchBatchSpectrum(kind, batch, launch);
    launchBatchRest(kind, batch);
    if (staged)
      check(hipMemcpyAsync(stage_neglect_host_, stage_neglect_, batch * sizeof(int), hipMemcpyDeviceToHost, s),
            "read neglect flags");
  };
  if (profiling_) check(hipEventRecord(static_cast<hipEvent_t>(event_start_), s), "hipEventRecord");
  if (use_graph_) {
    if (!graph_exec_ || !same(graph_key_, key) || graph_staged_ != staged) {
      releaseGraph();
      check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin graph capture");
      hipGraph_t graph = nullptr;
      try {
        sequence();
      } catch (...) {
        (void)hipStreamEndCapture(s, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
      }
      check(hipStreamEndCapture(s, &graph), "end graph capture");
      hipGraphExec_t exec = nullptr;
      const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      check(e, "instantiate graph");
      graph_exec_ = exec;
      graph_key_ = key;
      graph_staged_ = staged;
    }
    check(hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec_), s), "launch graph");
  } else {
    sequence();
  }
  last_used_graph_ = use_graph_;
  if (profiling_) check(hipEventRecord(static_cast<hipEvent_t>(event_stop_), s), "hipEventRecord");
  check(hipStreamSynchronize(s), "batch evaluation");
  if (profiling_) {
    float ms = 0;
    check(hipEventElapsedTime(&ms, static_cast<hipEvent_t>(event_start_), static_cast<hipEvent_t>(event_stop_)),
          "hipEventElapsedTime");
    batch_ms_ += ms;
  }
  batch_evaluations_ += batch;
}

void DeviceEngine::copyBatchBins(size_t batch, std::vector<double>& out) {
  hipStream_t s = static_cast<hipStream_t>(stream_);
  out.resize(batch * n_bins_);
  if (!out.empty())
    check(hipMemcpyAsync(out.data(), batch_bins_, out.size() * sizeof(double), hipMemcpyDeviceToHost, s),
          "read batch bins");
  check(hipStreamSynchronize(s), "read batch bins");
}

void DeviceEngine::copyBatchSpectrum(ModelKind kind, size_t batch, double* host) {
  hipStream_t s = static_cast<hipStream_t>(stream_);
  const Records rec{batch_dev_, rec_bytes_, rec_levels_, rec_species_, rec_adjoint_, rec_shifts_, groups_};
  const int n = int(n_points_);
  double* out = deviceAlloc<double>(batch * n_points_, "batch spectrum output");
  try {
    hipLaunchKernelGGL(batchSpectrumKernel, dim3(unsigned((n + kBatchBlock - 1) / kBatchBlock), unsigned(batch)),
                       dim3(kBatchBlock), 0, s, batch_flux_, stride_, wavenumber_, wavelength_, rec,
                       kind == ModelKind::emission ? 0 : (kind == ModelKind::eclipse ? 1 : 2), n, out);
    check(hipGetLastError(), "batch spectrum kernel");
    check(hipMemcpyAsync(host, out, batch * n_points_ * sizeof(double), hipMemcpyDeviceToHost, s),
          "read batch spectrum");
    check(hipStreamSynchronize(s), "read batch spectrum");
  } catch (...) {
    deviceFree(out);
    throw;
  }
  deviceFree(out);
}

void DeviceEngine::releaseBatch() {
  releaseGraph();
  if (batch_host_) (void)hipHostFree(batch_host_);
  if (batch_ll_host_) (void)hipHostFree(batch_ll_host_);
  for (void* p : {(void*)batch_dev_, (void*)batch_flux_, (void*)batch_bins_, (void*)batch_chunk_sums_, (void*)batch_ll_,
                  (void*)batch_chord_, (void*)batch_chord32_, tables_low_, (void*)table_base_})
    deviceFree(p);
}

}  // namespace bear_amd
