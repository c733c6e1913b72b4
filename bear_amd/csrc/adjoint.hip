// Adjoint of the batched emission, secondary-eclipse and transmission
// likelihoods on gfx950: the gradient of log L with respect to every level's
// temperature, altitude and effective species densities (and the flux
// scaling; transmission: the chord lengths and trapezoid coefficients of the
// tangent paths), for every live point of a batch.  The host carries it on
// to the retrieval parameters.  The transmission kernels are described at
// tangentAdjointKernel.
//
// Reverse order of the forward chain (emission and eclipse):
//   d logL / d bin_j = w_j (d_j - m_j) / s_j^2                  (gbarKernel)
//     eclipse: also d logL / d P_j for the planet's binned flux and the
//     star's terms                          (eclipseSlopeKernel, eclipseGbarKernel)
//   d logL / d F(nu) = sum_j d logL / d P_j  W_j(nu)            (fluxAdjointKernel, the
//                      folded operator transposed, per wavenumber)
//   short characteristics and Planck function                   (rtAdjointKernel)
//   table interpolation                                         (opacityAdjointKernel)
//   sums over wavenumbers                                       (adjointReduceKernel)
//
// The forward batch kernel keeps both angles' intensity less the Planck
// function and the opacity at every level (keptColumn layout,
// FP32) for as many points at a time as the budget holds.  rtAdjointKernel
// runs the short-characteristics recurrence backwards over them (no tables,
// few registers, many waves) and
// overwrites each level's opacity with d logL / d kappa; opacityAdjointKernel
// then recomputes each level's cross-sections from the tables as the forward
// kernels do and forms d logL / d n_eff and d logL / dT through the tables.
// When one point's kept quantities exceed the budget, emissionAdjointKernel
// does all of it for each work item, recomputing its forward sweep into a
// per-workgroup scratch first.
//
// Each kernel takes one wavefront per work item (64 wavenumbers of one point);
// once a level has all its contributions the wavefront sums them with
// permlane swaps and DPP moves and the last lane of each row stores its
// row's totals to partial[point][level][tile][slot]; a last kernel sums the
// tiles in a fixed order.  Slots per level: [0, NS) n dlogL/dn for the
// stencil slots' effective densities n, NS temperature through the tables (recomputing: also
// through the Planck function), NS+1 altitude, NS+2 flux scaling (level 0),
// NS+3 temperature through the Planck function.
#include <hip/hip_runtime.h>

#include <algorithm>
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

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kAdjointWavesPerCu = 16;                 // persistent grids of the table kernels
constexpr int kRtWavesPerCu = 16;                      // and of the recurrence kernel
constexpr size_t kPartialBudget = size_t(4) << 30;     // partial sums per chunk when recomputing
constexpr double kLn2 = 0.693147180559945309417232121458176568075500134;
constexpr double kInvMu1 = 1.0 / kMu1;
constexpr double kInvMu2 = 1.0 / kMu2;

// d logL / d bin_j for every (bin, point)
__global__ __launch_bounds__(kBlock) void gbarKernel(const double* __restrict__ bins, const BinDesc* __restrict__ desc,
                                                     int nb_bins, Records rec, double* __restrict__ gbar) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  const int m = blockIdx.y;
  if (j >= nb_bins) return;
  const double inflation = reinterpret_cast<const BatchModel*>(rec.at(m))->inflation;
  const BinDesc d = desc[j];
  const double e2 = d.error * d.error + inflation;
  gbar[size_t(m) * nb_bins + j] = d.weight * (d.data - bins[size_t(m) * nb_bins + j]) / e2;
}

// Secondary eclipse (secondary_eclipse.cpp:231-246, 262-355): depth_j =
// P_j / S_j (Rp/Rs)^2 1e6 (+ shift) with P_j = sum W_j F and S_j = sum W_j
// pi B(T_*) 1e-3, both summed per piece of the bin's folded segment by the
// forward bin kernel (chunk_sums[point][piece][2]).  Two steps:
//  1. eclipseSlopeKernel, one workgroup per (piece, point): the piece's sum
//     of W_j dB(T_*)/dT_* pi 1e-3 (so that a wide photometric band is spread
//     over as many workgroups as it has pieces);
//  2. eclipseGbarKernel, one thread per (bin, point): the pieces' sums in
//     order, then
//       gbar[j]    d logL / d depth_j
//       fscale[j]  d logL / d P_j = gbar_j (Rp/Rs)^2 1e6 / S_j
//       terms[j]   {gbar_j (depth_j - shift), -gbar_j (depth_j - shift) dS_j/dT_* / S_j}
__global__ __launch_bounds__(kBlockBins) void eclipseSlopeKernel(const double* __restrict__ wavenumber,
                                                                 const double* __restrict__ weights,
                                                                 const BinChunk* __restrict__ chunks, int nb_chunks,
                                                                 Records rec, double* __restrict__ slope) {
  __shared__ double partial[kBlockBins / 64];
  const int m = blockIdx.y;
  const BatchModel* bm = reinterpret_cast<const BatchModel*>(rec.at(m));
  const BinChunk c = chunks[blockIdx.x];
  const double* w = weights + c.weight_offset;
  double acc = 0.0;
  for (uint32_t k = threadIdx.x; k < c.length; k += kBlockBins) {
    double s, ds;
    starFluxWithDerivative(wavenumber[c.start + k], bm->star_inv_temperature, s, ds);
    acc = fma(w[k], ds, acc);
  }
  acc = waveSum(acc);
  if ((threadIdx.x & 63) == 0) partial[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0)
    slope[size_t(m) * nb_chunks + blockIdx.x] = (partial[0] + partial[1]) + (partial[2] + partial[3]);
}

__global__ __launch_bounds__(kBlock) void eclipseGbarKernel(
    const double* __restrict__ chunk_sums, const double* __restrict__ slope, int nb_chunks,
    const BinDesc* __restrict__ desc, int nb_bins, Records rec, const double* __restrict__ bins,
    double* __restrict__ gbar, double* __restrict__ fscale, double* __restrict__ terms) {
  const int j = blockIdx.x * kBlock + threadIdx.x, m = blockIdx.y;
  if (j >= nb_bins) return;
  const BatchModel* bm = reinterpret_cast<const BatchModel*>(rec.at(m));
  const BinDesc d = desc[j];
  const uint32_t pieces = (d.length + kBinChunk - 1) / kBinChunk;
  const double* cs = chunk_sums + (size_t(m) * nb_chunks + d.first_chunk) * 2;
  const double* sl = slope + size_t(m) * nb_chunks + d.first_chunk;
  double pj = 0.0, sj = 0.0, dsj = 0.0;
  for (uint32_t k = 0; k < pieces; ++k) {
    pj += cs[2 * k];
    sj += cs[2 * k + 1];
    dsj += sl[k];
  }
  const size_t o = size_t(m) * nb_bins + j;
  const double e2 = d.error * d.error + bm->inflation;
  const double g = d.weight * (d.data - bins[o]) / e2;
  const double depth = pj / sj * bm->depth_scale;
  gbar[o] = g;
  fscale[o] = g * bm->depth_scale / sj;
  terms[2 * o] = g * depth;
  terms[2 * o + 1] = -g * depth * dsj / sj;
}

// d logL / d F(nu) through the transposed operator
__global__ __launch_bounds__(kBlock) void fluxAdjointKernel(const uint64_t* __restrict__ ptr,
                                                            const uint32_t* __restrict__ bin,
                                                            const double* __restrict__ weight,
                                                            const double* __restrict__ gbar, int nb_bins, int n,
                                                            size_t stride, double* __restrict__ fbar) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const int m = blockIdx.y;
  if (i >= n) return;
  const double* g = gbar + size_t(m) * nb_bins;
  double acc = 0.0;
  for (uint64_t e = ptr[i]; e < ptr[i + 1]; ++e) acc = fma(g[bin[e]], weight[e], acc);
  fbar[size_t(m) * stride + i] = acc;
}

// Per slot of one level: the opacity term t_s = n_s sigma_s =
// 2^(log2 n_s + sum_k w_k v_k), as the forward kernels form it (with
// kGradientExp the exponential to 1.2e-7, pow2Gradient), and
// y_s = d log2 sigma_s / dT = sum_k u_k v_k; f(s, t_s, y_s) for each.  Weights
// from the level's shared sets with grouped stencils, else from the slots'
// records.
template <int NS, bool kGradientExp = false, class F>
__device__ __forceinline__ void slotTerms(const LevelSpecies* __restrict__ d, const LevelSpeciesAdjoint* __restrict__ a,
                                          const LevelState& lev, StencilGroups groups, const double (&v)[NS][4],
                                          F&& f) {
  auto term = [](double x) { return kGradientExp ? pow2Gradient(x) : pow2Opacity(x); };
  if (groups.grouped) {
    double wa[4], wb[4], ua[4], ub[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      wa[q] = lev.group_w[0][q];
      wb[q] = lev.group_w[1][q];
      ua[q] = lev.group_u[0][q];
      ub[q] = lev.group_u[1][q];
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const bool one = (groups.group1 >> s) & 1u;
      double x = fma(one ? wb[0] : wa[0], v[s][0], lev.log2_n[s]);
      x = fma(one ? wb[1] : wa[1], v[s][1], x);
      x = fma(one ? wb[2] : wa[2], v[s][2], x);
      x = fma(one ? wb[3] : wa[3], v[s][3], x);
      double y = (one ? ub[0] : ua[0]) * v[s][0];
      y = fma(one ? ub[1] : ua[1], v[s][1], y);
      y = fma(one ? ub[2] : ua[2], v[s][2], y);
      y = fma(one ? ub[3] : ua[3], v[s][3], y);
      f(s, term(x), y);
    }
  } else {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      double x = fma(d[s].w[0], v[s][0], d[s].log2_n);
      x = fma(d[s].w[1], v[s][1], x);
      x = fma(d[s].w[2], v[s][2], x);
      x = fma(d[s].w[3], v[s][3], x);
      double y = a[s].u[0] * v[s][0];
      y = fma(a[s].u[1], v[s][1], y);
      y = fma(a[s].u[2], v[s][2], y);
      y = fma(a[s].u[3], v[s][3], y);
      f(s, term(x), y);
    }
  }
}

// Opacity of one level for the backward sweep: kappa = sum_s t_s and
// d kappa / dT = ln 2 sum_s t_s y_s (slotTerms); t_s goes to t[s * 64] (this
// lane's LDS column).  Every slot's rows are reloaded when `reload` (the
// level processed before this one, the one above, brought new rows; or the
// first level of the sweep).
template <int NS>
__device__ __forceinline__ void levelOpacity(const double* __restrict__ tab, const LevelSpecies* __restrict__ d,
                                             const LevelSpeciesAdjoint* __restrict__ a, const LevelState& lev,
                                             StencilGroups groups, bool reload, double (&v)[NS][4],
                                             double* __restrict__ t, double& kappa, double& dkdt) {
  if (reload) {
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
      for (int q = 0; q < 4; ++q) v[s][q] = tab[size_t(d[s].block[q]) << 6];
  }
  double k = 0.0, dk = 0.0;
  slotTerms<NS>(d, a, lev, groups, v, [&](int s, double ts, double ys) {
    t[s * kWave] = ts;
    k += ts;
    dk = fma(ts, ys, dk);
  });
  kappa = k;
  dkdt = dk * kLn2;
}

// B(T) as planckAt and dB/dT = B X / (T (1 - q)), X = h c nu / (k T), q = e^-X
__device__ __forceinline__ void planckAdjoint(double b_num, double x_nu, double inv_t, double& b, double& dbdt) {
  const double q = pow2Lean(x_nu * inv_t);
  const double r = recip(1.0 - q);
  b = b_num * q * r;
  dbdt = b * (-x_nu * inv_t * kLn2) * inv_t * r;
}

// One angle of one layer backwards.  Forward: a = e^(-tau/mu),
// q = (a - 1) mu / tau, I_l = a I_{l-1} + (1 + q) B_l + (-a - q) B_{l-1}.
// dq/dtau = (mu (1 - a) - a tau) / tau^2, from its series in delta = tau/mu
// below 1e-2 (the closed form cancels there).  r = 1 / tau, inv_mu = 1 / mu.
__device__ __forceinline__ void angleAdjoint(double tau, double r, double mu, double inv_mu, double d_prev,
                                             double b_lo, double b_hi, double& ib, double& tbar, double& bbar_hi,
                                             double& bbar_lo) {
  const double a = pow2(tau * (-kLog2E / mu));
  const double q = (a - 1.0) * (mu * r);
  const double delta = tau * inv_mu;
  double dq;
  if (delta < 1e-2)
    dq = fma(delta, fma(delta, fma(delta, fma(delta, 1.0 / 144.0, -1.0 / 30.0), 1.0 / 8.0), -1.0 / 3.0), 0.5) *
         inv_mu;
  else
    dq = (mu * (1.0 - a) - a * tau) * r * r;
  tbar = fma(ib, d_prev * (-a * inv_mu) + (b_hi - b_lo) * dq, tbar);
  bbar_hi = fma(ib, 1.0 + q, bbar_hi);
  bbar_lo = fma(ib, -a - q, bbar_lo);
  ib *= a;
}

// The wavefront's sums of values x[0, K) (K a multiple of 4) to
// out[slot0 + k] for k < valid: the last lane of each row stores the totals
// its row holds.
template <int K>
__device__ __forceinline__ void emitSums(double* __restrict__ out, int lane, double (&x)[K], int slot0, int valid) {
  waveTotals<K>(x);
  if ((lane & 15) == 15) {
    const int r = rowSlot(lane >> 4);
#pragma unroll
    for (int b = 0; b < K / 4; ++b)
      if (4 * b + r < valid) out[slot0 + 4 * b + r] = x[b];
  }
}

// All slots of one level from the fused kernel: each slot's opacity term
// t_s = n_s sigma_s times d logL / d kappa (so n_s d logL / d n_s), the
// temperature (tables and Planck function), the altitude, the flux scaling
// (0 except at level 0) and 0 in the last slot.
template <int NS>
__device__ __forceinline__ void emitLevel(double* __restrict__ out, int lane, double kbar,
                                          const double* __restrict__ t, double tbar, double zbar,
                                          double scaling) {
  double x[NS + 4];
#pragma unroll
  for (int s = 0; s < NS; ++s) x[s] = kbar * t[s * kWave];
  x[NS] = tbar;
  x[NS + 1] = zbar;
  x[NS + 2] = scaling;
  x[NS + 3] = 0.0;
  emitSums<NS + 4>(out, lane, x, 0, NS + 4);
}

// Forward sweep into scratch = [grid][levels][2][64], then the backward
// sweep, per work item.
template <int NS>
__global__ __launch_bounds__(kWave) void emissionAdjointKernel(
    const double* __restrict__ wavenumber, const double* __restrict__ tables, size_t stride, Records rec,
    int nb_levels, int n, int tiles, int batch, const double* __restrict__ fbar, double* __restrict__ scratch,
    double* __restrict__ partial) {
  constexpr int K = NS + 4;
  __shared__ double sig[2 * NS * kWave];
  const int L = nb_levels;
  const int lane = int(threadIdx.x);
  double* my_sig = sig + lane;
  const size_t total = nbItems(tiles, batch);
  for (size_t item = blockIdx.x; item < total; item += gridDim.x) {
    int m, tile;
    if (!workItem(item, batch, tiles, m, tile)) continue;          // uniform over the wavefront
    const int i = tile * kWave + lane;
    const bool valid = i < n;
    const int ic = valid ? i : n - 1;                                // idle lanes: zero adjoints
    const unsigned char* r = rec.at(m);
    const BatchModel* bm = reinterpret_cast<const BatchModel*>(r);
    const LevelState* levels = reinterpret_cast<const LevelState*>(r + rec.levels);
    const LevelSpecies* species = reinterpret_cast<const LevelSpecies*>(r + rec.species);
    const LevelSpeciesAdjoint* extra = reinterpret_cast<const LevelSpeciesAdjoint*>(r + rec.adjoint);
    double* out = partial + (size_t(m) * L * size_t(tiles) + tile) * K;         // level l: + l * tiles * K
    const double* tab = tables + ic;
    const double nu = wavenumber[ic];
    const double b_num = kPlanckC1 * nu * nu * nu;
    const double x_nu = -(6.62606896e-27 * nu * 2.99792458e10 / 1.3806504e-16) * kLog2E;
    double v[NS][4];

    // forward sweep; I_l of angle a at I[l * 2 * 64 + a * 64]
    double* store = scratch + size_t(blockIdx.x) * size_t(L) * 2 * kWave + lane;
    {
      double k_prev = kappaTerms<NS>(tab, stride, species, levels[0], rec.groups, true, v);
      double b_prev = planckAt(b_num, x_nu, levels[0].inv_temperature);
      double z_prev = levels[0].altitude;
      double i1 = b_prev, i2 = b_prev;
      store[0] = i1;
      store[kWave] = i2;
      for (int l = 1; l < L; ++l) {
        const double k_next = kappaTerms<NS>(tab, stride, species + size_t(l) * NS, levels[l], rec.groups,
                                             levels[l].reload != 0, v);
        const double b_next = planckAt(b_num, x_nu, levels[l].inv_temperature);
        const double z_next = levels[l].altitude;
        shortCharacteristicsLayer((z_next - z_prev) * (k_next + k_prev) * 0.5, b_prev, b_next, i1, i2);
        store[size_t(l) * 2 * kWave] = i1;
        store[(size_t(l) * 2 + 1) * kWave] = i2;
        k_prev = k_next;
        b_prev = b_next;
        z_prev = z_next;
      }
    }
    const double* I = store;
    constexpr size_t level_step = 2 * kWave, angle_step = kWave;

    // F = 2 pi (mu1 w I1 + mu2 w I2) 1e-3 scaling at the top level
    const double fb = valid ? fbar[size_t(m) * stride + i] : 0.0;
    const double unit = 2.0 * kPi * 1e-3;
    const double* top = I + size_t(L - 1) * level_step;
    const double s_bar = fb * unit * (top[0] * (kMu1 * kW) + top[angle_step] * (kMu2 * kW));
    double ib1 = fb * unit * bm->scaling * (kMu1 * kW);
    double ib2 = fb * unit * bm->scaling * (kMu2 * kW);

    // backward sweep; level l's cross-sections in LDS half (l & 1)
    double k_hi, dk_hi, b_hi, db_hi;
    levelOpacity<NS>(tab, species + size_t(L - 1) * NS, extra + size_t(L - 1) * NS, levels[L - 1], rec.groups, true, v,
                     my_sig + ((L - 1) & 1) * NS * kWave, k_hi, dk_hi);
    planckAdjoint(b_num, x_nu, levels[L - 1].inv_temperature, b_hi, db_hi);
    double z_hi = levels[L - 1].altitude;
    double kbar_hi = 0.0, zbar_hi = 0.0, bbar_hi = 0.0;
    for (int l = L - 1; l >= 1; --l) {
      const double* below = I + size_t(l - 1) * level_step;
      const double i1p = below[0];
      const double i2p = below[angle_step];
      double k_lo, dk_lo, b_lo, db_lo;
      levelOpacity<NS>(tab, species + size_t(l - 1) * NS, extra + size_t(l - 1) * NS, levels[l - 1], rec.groups,
                       levels[l].reload != 0, v,
                       my_sig + ((l - 1) & 1) * NS * kWave, k_lo, dk_lo);
      planckAdjoint(b_num, x_nu, levels[l - 1].inv_temperature, b_lo, db_lo);
      const double z_lo = levels[l - 1].altitude;
      const double dz = z_hi - z_lo;
      const double k_mean = (k_hi + k_lo) * 0.5;
      const double tau = dz * k_mean;
      double tbar = 0.0, bbar_lo = 0.0;
      if (tau != 0) {
        const double rt = recip(tau);
        angleAdjoint(tau, rt, kMu1, kInvMu1, i1p - b_lo, b_lo, b_hi, ib1, tbar, bbar_hi, bbar_lo);
        angleAdjoint(tau, rt, kMu2, kInvMu2, i2p - b_lo, b_lo, b_hi, ib2, tbar, bbar_hi, bbar_lo);
      }
      const double kbar_lo = tbar * dz * 0.5;
      kbar_hi += kbar_lo;
      zbar_hi = fma(tbar, k_mean, zbar_hi);
      // every contribution to level l is in: emit it
      emitLevel<NS>(out + size_t(l) * tiles * K, lane, kbar_hi, my_sig + (l & 1) * NS * kWave,
                    fma(bbar_hi, db_hi, kbar_hi * dk_hi), zbar_hi, 0.0);
      k_hi = k_lo;
      dk_hi = dk_lo;
      b_hi = b_lo;
      db_hi = db_lo;
      z_hi = z_lo;
      kbar_hi = kbar_lo;
      zbar_hi = -tbar * k_mean;
      bbar_hi = bbar_lo;
    }
    bbar_hi += ib1 + ib2;                                            // I_0 = B_0
    emitLevel<NS>(out, lane, kbar_hi, my_sig, fma(bbar_hi, db_hi, kbar_hi * dk_hi), zbar_hi, s_bar);
  }
}

// Backward short characteristics over the kept forward quantities of a
// chunk (keptColumn layout: I_1 - B, I_2 - B, kappa per level).  Emits per level
// the altitude, the flux scaling (level 0) and the temperature through the
// Planck function, and overwrites each level's kappa with d logL / d kappa
// for opacityAdjointKernel.  ns: stencil slots per level.
// CL: the point's cloud models (BatchModel::clouds, layer weights
// LevelState::cloud_w) add their optical depth to each layer's, and per
// layer and model four sums of d logL / d tau times g_k go to slots
// cloud_slot0 + 4 m + k: g_0 = f, the model's spectral factor; power law
// f = (lambda / lambda_ref)^a: g_1 = f (ln lambda - ln lambda_ref);
// KHnongrey f = c3 / u, u = q0 x^-a0 + x^0.2, x = 2 pi r / lambda:
// g_1 = f x^-a0 / u, g_2 = f q0 x^-a0 ln x / u,
// g_3 = f (0.2 x^0.2 - a0 q0 x^-a0) / u (the host adds c3's own terms).
template <bool CL = false>
__global__ __launch_bounds__(kWave) void rtAdjointKernel(const double* __restrict__ wavenumber, Records rec,
                                                         int nb_levels, int n, size_t stride, int tiles, int batch,
                                                         int ns, int kslots, const double* __restrict__ fbar,
                                                         float* __restrict__ keep, double* __restrict__ partial,
                                                         const double* __restrict__ wavelength, int cloud_slot0) {
  const int L = nb_levels;
  const int K = kslots;
  const int lane = int(threadIdx.x);
  const size_t level_step = kKeptLevel;
  const size_t total = nbItems(tiles, batch);
  for (size_t item = blockIdx.x; item < total; item += gridDim.x) {
    int m, tile;
    if (!workItem(item, batch, tiles, m, tile)) continue;          // uniform over the wavefront
    const int i = tile * kWave + lane;
    const bool valid = i < n;
    const int ic = valid ? i : n - 1;                                // idle lanes: zero adjoints, no stores
    const unsigned char* r = rec.at(m);
    const BatchModel* bm = reinterpret_cast<const BatchModel*>(r);
    const LevelState* levels = reinterpret_cast<const LevelState*>(r + rec.levels);
    float* kp = keep + keptColumn(m, stride, L, ic);
    double* out = partial + (size_t(m) * L * size_t(tiles) + tile) * K;         // level l: + l * tiles * K
    const double nu = wavenumber[ic];
    const double b_num = kPlanckC1 * nu * nu * nu;
    const double x_nu = -(6.62606896e-27 * nu * 2.99792458e10 / 1.3806504e-16) * kLog2E;
    const double fb = valid ? fbar[size_t(m) * stride + i] : 0.0;
    const double unit = 2.0 * kPi * 1e-3;
    double cf[kMaxClouds], cg[kMaxClouds][3];
    if constexpr (CL) {
      double hm[7];
      const Extras e{nullptr, nullptr, bm->clouds};
      const double lam = wavelength[ic];
      loadExtras(e, ic, n, lam, hm, cf);
#pragma unroll
      for (int c = 0; c < kMaxClouds; ++c) {
        cg[c][0] = cg[c][1] = cg[c][2] = 0.0;
        if (c < e.clouds.count && e.clouds.kind[c] == kCloudPowerLaw) {
          cg[c][0] = cf[c] * (log(lam) - e.clouds.c[c][2]);
        } else if (c < e.clouds.count && e.clouds.kind[c] == kCloudKH) {
          const double* k = e.clouds.c[c];
          const double x = k[2] / lam, xm = pow(x, -k[1]), x2 = pow(x, 0.2);
          const double fu = cf[c] / (k[0] * xm + x2);
          cg[c][0] = fu * xm;
          cg[c][1] = fu * k[0] * xm * log(x);
          cg[c][2] = fu * (0.2 * x2 - k[1] * k[0] * xm);
        }
      }
    }
    // kept: d = I - B of both angles and kappa per level, FP32
    const float* top = kp + size_t(L - 1) * level_step;
    double b_hi, db_hi;
    planckAdjoint(b_num, x_nu, levels[L - 1].inv_temperature, b_hi, db_hi);
    const double s_bar = fb * unit * ((double(top[0]) + b_hi) * (kMu1 * kW) + (double(top[kKeptValue]) + b_hi) * (kMu2 * kW));
    double ib1 = fb * unit * bm->scaling * (kMu1 * kW);
    double ib2 = fb * unit * bm->scaling * (kMu2 * kW);
    double k_hi = top[2 * kKeptValue];
    double z_hi = levels[L - 1].altitude;
    double kbar_hi = 0.0, zbar_hi = 0.0, bbar_hi = 0.0;
    const float* below = kp + size_t(L - 2) * level_step;          // the level below, read one step ahead
    double d1n = below[0], d2n = below[kKeptValue], kn = below[2 * kKeptValue];
    for (int l = L - 1; l >= 1; --l) {
      const double d1p = d1n, d2p = d2n, k_lo = kn;
      if (l >= 2) {
        below = kp + size_t(l - 2) * level_step;
        d1n = below[0];
        d2n = below[kKeptValue];
        kn = below[2 * kKeptValue];
      }
      double b_lo, db_lo;
      planckAdjoint(b_num, x_nu, levels[l - 1].inv_temperature, b_lo, db_lo);
      const double z_lo = levels[l - 1].altitude;
      const double dz = z_hi - z_lo;
      const double k_mean = (k_hi + k_lo) * 0.5;
      double tau = dz * k_mean;
      if constexpr (CL)
#pragma unroll
        for (int c = 0; c < kMaxClouds; ++c) tau = fma(levels[l].cloud_w[c], cf[c], tau);
      double tbar = 0.0, bbar_lo = 0.0;
      if (tau != 0) {
        const double rt = recip(tau);
        angleAdjoint(tau, rt, kMu1, kInvMu1, d1p, b_lo, b_hi, ib1, tbar, bbar_hi, bbar_lo);
        angleAdjoint(tau, rt, kMu2, kInvMu2, d2p, b_lo, b_hi, ib2, tbar, bbar_hi, bbar_lo);
      }
      const double kbar_lo = tbar * dz * 0.5;
      kbar_hi += kbar_lo;
      zbar_hi = fma(tbar, k_mean, zbar_hi);
      // every contribution to level l is in
      if (valid) kp[size_t(l) * level_step + 2 * kKeptValue] = float(kbar_hi);
      double x[4] = {zbar_hi, 0.0, bbar_hi * db_hi, 0.0};
      emitSums<4>(out + size_t(l) * tiles * K, lane, x, ns + 1, 3);
      if constexpr (CL) {
        double xc[4 * kMaxClouds];
#pragma unroll
        for (int c = 0; c < kMaxClouds; ++c) {
          xc[4 * c] = tbar * cf[c];
#pragma unroll
          for (int k = 0; k < 3; ++k) xc[4 * c + 1 + k] = tbar * cg[c][k];
        }
        emitSums<4 * kMaxClouds>(out + size_t(l) * tiles * K, lane, xc, cloud_slot0, 4 * kMaxClouds);
      }
      k_hi = k_lo;
      b_hi = b_lo;
      db_hi = db_lo;
      z_hi = z_lo;
      kbar_hi = kbar_lo;
      zbar_hi = -tbar * k_mean;
      bbar_hi = bbar_lo;
    }
    bbar_hi += ib1 + ib2;                                            // I_0 = B_0
    if (valid) kp[2 * kKeptValue] = float(kbar_hi);
    double x[4] = {zbar_hi, s_bar, bbar_hi * db_hi, 0.0};
    emitSums<4>(out, lane, x, ns + 1, 3);
    if constexpr (CL) {                                              // no layer below level 0
      double xc[4 * kMaxClouds] = {};
      emitSums<4 * kMaxClouds>(out, lane, xc, cloud_slot0, 4 * kMaxClouds);
    }
  }
}

// Per level, from d logL / d kappa (value 2 of each level's kept triple, written by
// rtAdjointKernel): the opacity terms recomputed from the tables as the
// forward kernels do (register-cached rows, reloaded when they change),
// n_s d logL / d n_s = sum over wavenumbers of kappa-bar t_s and the
// temperature's share through the tables, kappa-bar d kappa / dT.
// HM: also, in slots NS+4 .. NS+10, kappa-bar times each of the H-
// continuum's seven rows (hminus [7][n]): the gradients of its per-level
// weights LevelState::hminus_c.
template <int NS, bool HM = false>
__global__ __launch_bounds__(kWave) void opacityAdjointKernel(const double* __restrict__ tables, size_t stride,
                                                              Records rec, int nb_levels, int n, int tiles,
                                                              TileWindow win, int batch,
                                                              const float* __restrict__ keep,
                                                              const double* __restrict__ hminus, int kslots,
                                                              double* __restrict__ partial) {
  constexpr int KX = NS + 4;
  const int K = kslots;                                             // partial row: every level slot
  const int L = nb_levels;
  const int lane = int(threadIdx.x);
  const size_t total = nbItems(win.count, batch);
  for (size_t item = blockIdx.x; item < total; item += gridDim.x) {
    int m, tile;
    if (!windowItem(item, batch, win, m, tile)) continue;          // uniform over the wavefront
    const int i = tile * kWave + lane;
    const bool valid = i < n;
    const int ic = valid ? i : n - 1;
    const unsigned char* r = rec.at(m);
    const LevelState* levels = reinterpret_cast<const LevelState*>(r + rec.levels);
    const LevelSpecies* species = reinterpret_cast<const LevelSpecies*>(r + rec.species);
    const LevelSpeciesAdjoint* extra = reinterpret_cast<const LevelSpeciesAdjoint*>(r + rec.adjoint);
    const float* kb = keep + keptAt(win, m, stride, L, ic) + 2 * kKeptValue;
    double* out = partial + (size_t(m) * L * size_t(tiles) + tile) * K;         // level l: + l * tiles * K
    const double* tab = tables + ic;
    double v[NS][4];
    double hm[7];
    if constexpr (HM)
#pragma unroll
      for (int r = 0; r < 7; ++r) hm[r] = hminus[size_t(r) * n + ic];
    // d logL / d kappa streams from HBM: read two levels ahead
    double kb1 = kb[0], kb2 = L > 1 ? kb[kKeptLevel] : 0.0;
    for (int l = 0; l < L; ++l) {
      const double kbar = valid ? kb1 : 0.0;
      const LevelSpecies* d = species + size_t(l) * NS;
      const LevelSpeciesAdjoint* a = extra + size_t(l) * NS;
      if (l == 0 || levels[l].reload != 0) {
#pragma unroll
        for (int s = 0; s < NS; ++s)
#pragma unroll
          for (int q = 0; q < 4; ++q) v[s][q] = tab[size_t(d[s].block[q]) << 6];
      }
      kb1 = kb2;
      if (l + 2 < L) kb2 = kb[size_t(l + 2) * kKeptLevel];
      double x[KX];
      double dk = 0.0;
      // the terms to 1.2e-7 (pow2Gradient): they only weight d logL / d kappa
      slotTerms<NS, true>(d, a, levels[l], rec.groups, v, [&](int s, double ts, double ys) {
        x[s] = kbar * ts;
        dk = fma(ts, ys, dk);
      });
      x[NS] = kbar * dk * kLn2;
#pragma unroll
      for (int s = NS + 1; s < KX; ++s) x[s] = 0.0;
      emitSums<KX>(out + size_t(l) * tiles * K, lane, x, 0, NS + 1);
      if constexpr (HM) {
        double xh[8];
#pragma unroll
        for (int r = 0; r < 7; ++r) xh[r] = kbar * hm[r];
        xh[7] = 0.0;
        emitSums<8>(out + size_t(l) * tiles * K, lane, xh, NS + 4, 8);
      }
    }
  }
}

// out[point][level][slot] = sum over tiles of partial[point][level][tile][slot]:
// per thread a strided run of tiles (each a contiguous row of K slots), then
// the wavefronts' sums, in a fixed order
template <int K>
__global__ __launch_bounds__(kBlock) void adjointReduceKernel(const double* __restrict__ partial, int nb_levels,
                                                              int tiles, double* __restrict__ out) {
  __shared__ double part[kBlock / 64][K];
  const int l = blockIdx.x, m = blockIdx.y;
  const size_t row = size_t(m) * nb_levels + l;
  const double* p = partial + row * size_t(tiles) * K;
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  for (int t = threadIdx.x; t < tiles; t += kBlock) {
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] += p[size_t(t) * K + k];
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double v = waveSum(acc[k]);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    const int k = int(threadIdx.x);
    out[row * K + k] = (part[0][k] + part[1][k]) + (part[2][k] + part[3][k]);
  }
}

// hminus: the H- continuum's rows (kept forward only), its sums in slots
// NS+4 .. NS+10 of kslots per level
template <int NS>
void launchAdjoint(bool kept, dim3 grid, dim3 rt_grid, hipStream_t s, const double* nu, const double* lam,
                   const double* tables, size_t stride, Records rec, int L, int n, int tiles, int batch,
                   const double* fbar, double* buffer, double* partial, const double* hminus, int kslots,
                   int cloud_slot0) {
  if (kept) {
    float* kept_values = reinterpret_cast<float*>(buffer);
    if (cloud_slot0 >= 0)
      hipLaunchKernelGGL(rtAdjointKernel<true>, rt_grid, dim3(kWave), 0, s, nu, rec, L, n, stride, tiles, batch, NS,
                         kslots, fbar, kept_values, partial, lam, cloud_slot0);
    else
      hipLaunchKernelGGL(rtAdjointKernel<false>, rt_grid, dim3(kWave), 0, s, nu, rec, L, n, stride, tiles, batch, NS,
                         kslots, fbar, kept_values, partial, lam, cloud_slot0);
    if (hminus)
      hipLaunchKernelGGL((opacityAdjointKernel<NS, true>), grid, dim3(kWave), 0, s, tables, stride, rec, L, n, tiles,
                         wholeGrid(tiles), batch, kept_values, hminus, kslots, partial);
    else
      hipLaunchKernelGGL(opacityAdjointKernel<NS>, grid, dim3(kWave), 0, s, tables, stride, rec, L, n, tiles,
                         wholeGrid(tiles), batch, kept_values, hminus, kslots, partial);
  } else {
    hipLaunchKernelGGL(emissionAdjointKernel<NS>, grid, dim3(kWave), 0, s, nu, tables, stride, rec, L, n, tiles,
                       batch, fbar, buffer, partial);
  }
}

// ---------------------------------------------------------------- transmission
// Adjoint of the transit depth (transmission_spectrum.cpp:32-154) over the
// kept forward quantities (keptColumn layout, FP32) with value 0 the
// extinction k_l of the forward batch kernel.  With e = d logL / d depth
// 1e6 / R_*^2 (= d logL / dH, depth = (H + R^2) 1e6 / R_*^2) and
//   H = sum_t c_t (1 - e^-tau_t),  c_t = (R + z_t)(z_{t+1} - z_{t-1}),
//   tau_t = sum_{q = t}^{last_t} chord[q][t] S_q,  S_q = k_q + k_{q+1},
// (last_t: the segment at which the sum first passes 10, else L - 2):
//  0. the kept forward (transmissionBatchKernel) stores tau_t in slot 1 and
//     last_t in slot 2 of level t;
//     The kept forward also stores per (point, tile of 64 wavenumbers, path)
//     the tile's largest last_t (reach), with which steps 2 and 3 pass over
//     paths and wavenumber steps that end below their segments without
//     loading them;
//  1. tangentAdjointKernel: tau-bar_t = e c_t e^-tau_t over slot 1; sums over wavenumbers of
//     e (1 - e^-tau_t) (d logL / d c_t) in slot NS+1 of level t, of e (the
//     R^2 term) in slot NS+1 of level L-1 and of d logL / d depth x depth
//     (the R_* term) in slot NS+2 of level L-1;
//  2. chordAdjointKernel + chordReduceKernel: d logL / d chord[q][t] =
//     sum over wavenumbers of tau-bar_t S_q where q <= last_t;
//  3. segmentAdjointKernel: S-bar_q = sum_{t <= q, q <= last_t} tau-bar_t
//     chord[q][t] and d logL / d k_l = S-bar_{l-1} + S-bar_l over slot 2 (in
//     decreasing q: slot 2 of level q+1 is read for the last time by
//     S-bar_q), and the Rayleigh sum sigma_R d logL / d k_l in slot NS+3;
//  4. opacityAdjointKernel as for emission (slot 2 = d logL / d k).
// The host carries d logL / d chord and d logL / d c_t to the altitudes and
// the radius.
__global__ __launch_bounds__(kWave) void tangentAdjointKernel(Records rec, int nb_levels, int n, size_t stride, int tiles,
                                                              TileWindow win, int batch, int ns, int kslots,
                                                              const double* __restrict__ fbar,
                                                              const double* __restrict__ depth,
                                                              float* __restrict__ keep, double* __restrict__ partial) {
  const int L = nb_levels, K = kslots;
  const int lane = int(threadIdx.x);
  const size_t level_step = kKeptLevel;
  const size_t total = nbItems(win.count, batch);
  for (size_t item = blockIdx.x; item < total; item += gridDim.x) {
    int m, tile;
    if (!windowItem(item, batch, win, m, tile)) continue;          // uniform over the wavefront
    const int i = tile * kWave + lane;
    const bool valid = i < n;
    const int ic = valid ? i : n - 1;                                // idle lanes: zero adjoints, no stores
    const unsigned char* r = rec.at(m);
    const BatchModel* bm = reinterpret_cast<const BatchModel*>(r);
    const LevelState* levels = reinterpret_cast<const LevelState*>(r + rec.levels);
    float* kp = keep + keptAt(win, m, stride, L, ic);
    double* out = partial + (size_t(m) * L * size_t(tiles) + tile) * K;         // level l: + l * tiles * K
    const double fb = valid ? fbar[size_t(m) * stride + i] : 0.0;
    const double e = fb * 1e6 / (bm->star_radius * bm->star_radius);
    const double r0 = bm->bottom_radius;
    for (int t = 0; t < L - 1; ++t) {
      float* kt = kp + size_t(t) * level_step + kKeptValue;
      const double tr = pow2(-double(*kt) * kLog2E);
      const double zl = levels[t > 0 ? t - 1 : 0].altitude, zu = levels[t + 1].altitude;
      const double ct = (r0 + levels[t].altitude) * (zu - zl);
      if (valid) *kt = float(e * ct * tr);
      double x4[4] = {e * (1.0 - tr), 0.0, 0.0, 0.0};
      emitSums<4>(out + size_t(t) * tiles * K, lane, x4, ns + 1, 2);
    }
    const double d = valid ? depth[size_t(m) * stride + i] : 0.0;
    double x4[4] = {e, fb * d, 0.0, 0.0};
    emitSums<4>(out + size_t(L - 1) * tiles * K, lane, x4, ns + 1, 2);
  }
}

// Step 2: one workgroup of 256 threads per (64 x 64 block of (q, t) with
// t-block <= q-block, point, range of wavenumbers), 4 x 4 pairs per thread
// (four consecutive segments, so each wavefront holds 16 consecutive ones,
// times four paths), the wavenumbers in steps of 32 through LDS; a
// wavefront skips a step in which no path of the block reaches its
// segments.  partial[point][range][q][t] for P = L - 1 segments and tangent
// paths.
constexpr int kChordBlock = 64, kChordNu = 32, kChordRanges = 16;
// chord ranges (and hence adjoint windows) start at multiples of this many
// wavenumbers: the widest tile of the batched transmission kernel
constexpr int kWindowAlign = 128;

__global__ __launch_bounds__(256) void chordAdjointKernel(const float* __restrict__ keep, int nb_levels, int n,
                                                          size_t stride, int ranges, int range_first,
                                                          int nu_per_range, TileWindow win, int tiles,
                                                          const int* __restrict__ reach,
                                                          double* __restrict__ partial) {
  __shared__ float s_sh[kChordBlock][kChordNu + 1];           // S_q in FP32 (the gradient's inputs are FP32)
  __shared__ float a_sh[kChordBlock][kChordNu + 1];
  __shared__ int l_sh[kChordBlock][kChordNu + 1];             // last segment of each path (-1: none)
  __shared__ int s_reach[2];                                   // bit w: some path reaches wavefront w's segments
  const int L = nb_levels, P = L - 1;
  int bq = 0, bt = int(blockIdx.x);                                 // triangular block index
  while (bt > bq) {
    bt -= bq + 1;
    ++bq;
  }
  const int m = int(blockIdx.y), range = range_first + int(blockIdx.z);
  const int q0 = bq * kChordBlock, t0 = bt * kChordBlock;
  const int tid = int(threadIdx.x), tq = tid >> 4, tt = tid & 15, wave = tid >> 6;
  const int q_first = q0 + 4 * tq;                                  // this thread's segments q_first .. + 3
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  if (tid < 2) s_reach[tid] = 0;
  __syncthreads();
  const int nu0 = range * nu_per_range, nu1 = min(n, nu0 + nu_per_range);
  const int* rl = reach + size_t(m) * tiles * P;
  for (int v0 = nu0, step = 0; v0 < nu1; v0 += kChordNu, ++step) {
    // a step (within one tile of 64) in which no path of the block reaches
    // its first segment adds nothing: nothing is loaded for it
    const int tile_v = v0 / kWave;
    const bool live = tid < kChordBlock && t0 + tid < P && rl[size_t(tile_v) * P + t0 + tid] >= q0;
    if (!__syncthreads_or(live)) continue;                         // uniform over the workgroup
    int reach = 0;
    for (int e = tid; e < kChordBlock * kChordNu; e += 256) {
      const int row = e / kChordNu, col = e % kChordNu;
      const int v = v0 + col, q = q0 + row, t = t0 + row;
      float sv = 0.0f, av = 0.0f;
      int lv = -1;
      if (v < nu1) {
        const float* col = keep + keptAt(win, m, stride, L, v);
        if (q < P) sv = float(double(col[size_t(q) * kKeptLevel]) + double(col[size_t(q + 1) * kKeptLevel]));
        if (t < P) {
          av = col[size_t(t) * kKeptLevel + kKeptValue];
          lv = int(col[size_t(t) * kKeptLevel + 2 * kKeptValue]);
        }
      }
      s_sh[row][col] = sv;
      a_sh[row][col] = av;
      l_sh[row][col] = lv;
      // wavefronts 0 .. (lv - q0) / 16, whose first segments this path reaches
      if (lv >= q0) reach |= (2 << min(3, (lv - q0) / 16)) - 1;
    }
    if (reach) atomicOr(&s_reach[step & 1], reach);
    __syncthreads();
    const int bits = s_reach[step & 1];
    if ((bits >> wave) & 1) {                                       // uniform over the wavefront
      for (int v = 0; v < kChordNu; ++v) {
        double sv[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) sv[a] = double(s_sh[4 * tq + a][v]);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const double av = double(a_sh[tt + 16 * b][v]);
          const int lv = l_sh[tt + 16 * b][v];
          // past the path's last segment only the high word of tau-bar is
          // cleared: a subnormal below 2.2e-314 whose products with S (FP32
          // extinction sums) vanish in the sums; one select per product
          const int hi = __double2hiint(av), lo = __double2loint(av);
#pragma unroll
          for (int a = 0; a < 4; ++a)
            acc[a][b] = fma(__hiloint2double(q_first + a <= lv ? hi : 0, lo), sv[a], acc[a][b]);
        }
      }
    }
    __syncthreads();
    // every thread has read this step's bits: clear them for the step after
    // next (the next step ORs into the other word, cleared a step ago)
    if (tid == 0) s_reach[step & 1] = 0;
  }
  double* out = partial + (size_t(m) * ranges + range) * size_t(P) * P;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int q = q_first + a;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int t = t0 + tt + 16 * b;
      if (q < P && t < P) out[size_t(q) * P + t] = t <= q ? acc[a][b] : 0.0;
    }
  }
}

// out[point][q][t] = the ranges' partial sums in order (q >= t; else 0)
__global__ __launch_bounds__(kBlock) void chordReduceKernel(const double* __restrict__ partial, int P, int ranges,
                                                            double* __restrict__ out) {
  const int e = int(blockIdx.x) * kBlock + int(threadIdx.x), m = int(blockIdx.y);
  if (e >= P * P) return;
  const int q = e / P, t = e % P;
  double sum = 0.0;
  if (t <= q)
    for (int k = 0; k < ranges; ++k) sum += partial[(size_t(m) * ranges + k) * size_t(P) * P + e];
  out[size_t(m) * P * P + e] = sum;
}

// chord_t[point][t][q] = chord[point][q][t] for t < L - 1, q < segmentRows(L)
// (zero past the last segment), so that step 3 reads a path's chords over
// consecutive segments with one scalar load
constexpr int kAdjointSegments = 32;
__host__ __device__ constexpr int segmentRows(int nb_levels) {
  return (nb_levels - 1 + kAdjointSegments - 1) / kAdjointSegments * kAdjointSegments;
}

__global__ __launch_bounds__(kBlock) void chordTransposeKernel(const double* __restrict__ chord, int nb_levels,
                                                               double* __restrict__ chord_t) {
  const int L = nb_levels, P = L - 1, rows = segmentRows(L);
  const int segments = chordSegments(L), tangents = chordTangents(L);
  const int e = int(blockIdx.x) * kBlock + int(threadIdx.x), m = int(blockIdx.y);
  if (e >= P * rows) return;
  const int t = e / rows, q = e % rows;
  chord_t[size_t(m) * P * rows + e] = q < P ? chord[(size_t(m) * segments + q) * tangents + t] : 0.0;
}

__global__ __launch_bounds__(kWave) void segmentAdjointKernel(Records rec, const double* __restrict__ chord_t,
                                                              const double* __restrict__ rayleigh, int nb_levels,
                                                              int n, size_t stride, int tiles, TileWindow win,
                                                              int batch, int ns, int kslots, float* __restrict__ keep,
                                                              double* __restrict__ partial,
                                                              const int* __restrict__ reach) {
  const int L = nb_levels, K = kslots;
  const int rows = segmentRows(L);
  const int lane = int(threadIdx.x);
  const size_t level_step = kKeptLevel;
  const size_t total = nbItems(win.count, batch);
  for (size_t item = blockIdx.x; item < total; item += gridDim.x) {
    int m, tile;
    if (!windowItem(item, batch, win, m, tile, 1)) continue;   // uniform; tiles of one point together (its chords)
    const int i = tile * kWave + lane;
    const bool valid = i < n;
    const int ic = valid ? i : n - 1;
    float* kp = keep + keptAt(win, m, stride, L, ic);
    const double* c = chord_t + size_t(m) * (L - 1) * rows;
    double* out = partial + (size_t(m) * L * size_t(tiles) + tile) * K;
    const double sr = (rayleigh && valid) ? rayleigh[i] : 0.0;
    const int* rl = reach + (size_t(m) * tiles + tile) * (L - 1);   // per path: the wavefront's largest last_t
    double s_hi = 0.0;                                               // S-bar of the segment above
    // kAdjointSegments segments per pass over the paths (one load of
    // tau-bar_t and last_t for all of them); chord[q][t] = 0 for t > q and
    // past the last segment.  Per path, tests uniform over the wavefront:
    // every lane's path ends below the pass (skipped), or reaches past it
    // (no masks); otherwise the products are masked lane by lane.
    for (int qb = (L - 2) / kAdjointSegments * kAdjointSegments; qb >= 0; qb -= kAdjointSegments) {
      double sb[kAdjointSegments];
#pragma unroll
      for (int u = 0; u < kAdjointSegments; ++u) sb[u] = 0.0;
      const int t_end = min(qb + kAdjointSegments, L - 1);
      // this lane's tau-bar and last segment of path t, for the paths that
      // reach the pass (the next one's are loaded while a path's products
      // run); the others end below every segment of the pass
      auto column = [&](int t) { return kp + size_t(t) * level_step; };
      auto next = [&](int t) {                         // the first path from t on that reaches the pass
        while (t < t_end && rl[t] < qb) ++t;
        return t;
      };
      int t = next(0);
      double tb = 0.0;
      int last = -1;
      if (t < t_end) {
        tb = double(column(t)[kKeptValue]);
        last = int(column(t)[2 * kKeptValue]);
      }
      while (t < t_end) {
        const int t_next = next(t + 1);
        double tb_next = 0.0;
        int last_next = -1;
        if (t_next < t_end) {                          // uniform
          tb_next = double(column(t_next)[kKeptValue]);
          last_next = int(column(t_next)[2 * kKeptValue]);
        }
        {
          const double* ct = c + size_t(t) * rows + qb;
          const int k = last - qb;                     // this lane sums segments qb .. qb + k
          // Past segment qb + k only the high word of tau-bar is cleared:
          // what remains is a subnormal below 2.2e-314, whose products
          // with chords (below 1e10 cm) stay below 3e-304 and vanish in
          // the sums.  One select per product instead of two.
          const int hi = __double2hiint(tb), lo = __double2loint(tb);
#pragma unroll
          for (int u = 0; u < kAdjointSegments; ++u)
            sb[u] = fma(__hiloint2double(u <= k ? hi : 0, lo), ct[u], sb[u]);
        }
        tb = tb_next;
        last = last_next;
        t = t_next;
      }
#pragma unroll
      for (int u = kAdjointSegments - 1; u >= 0; --u) {
        const int q = qb + u;
        if (q > L - 2) continue;
        const double kb = sb[u] + s_hi;                              // d logL / d k_{q+1}
        if (valid) kp[size_t(q + 1) * level_step + 2 * kKeptValue] = float(kb);
        double x4[4] = {kb * sr, 0.0, 0.0, 0.0};
        emitSums<4>(out + size_t(q + 1) * tiles * K, lane, x4, ns + 3, 1);
        s_hi = sb[u];
      }
    }
    {
      const double kb = s_hi;                                        // d logL / d k_0
      if (valid) kp[2 * kKeptValue] = float(kb);
      double x4[4] = {kb * sr, 0.0, 0.0, 0.0};
      emitSums<4>(out, lane, x4, ns + 3, 1);
    }
  }
}

// The transmission adjoint over the wavenumbers of chord ranges [r0, r1)
// (win: their tiles of 64; keep holds those tiles' kept quantities): path,
// chord, segment and table adjoints.  The chord products of the ranges go
// to chord_partial[point][range]; chordReduceKernel sums the ranges once
// every window is done.  chord_t (chordTransposeKernel) is made before the
// first window.
template <int NS>
void launchTransmissionAdjoint(dim3 grid, dim3 rt_grid, hipStream_t s, const double* tables, size_t stride,
                               const double* rayleigh, Records rec, int L, int n, int tiles, TileWindow win,
                               int batch, const double* fbar, const double* depth, float* keep, double* partial,
                               int ranges, int r0, int r1, int nu_per_range, double* chord_partial,
                               const double* chord_t, int* reach, const double* hminus, int kslots) {
  hipLaunchKernelGGL(tangentAdjointKernel, rt_grid, dim3(kWave), 0, s, rec, L, n, stride, tiles, win, batch, NS,
                     kslots, fbar, depth, keep, partial);
  const int P = L - 1, blocks = (P + kChordBlock - 1) / kChordBlock;
  hipLaunchKernelGGL(chordAdjointKernel,
                     dim3(unsigned(blocks * (blocks + 1) / 2), unsigned(batch), unsigned(r1 - r0)), dim3(256), 0, s,
                     keep, L, n, stride, ranges, r0, nu_per_range, win, tiles, reach, chord_partial);
  hipLaunchKernelGGL(segmentAdjointKernel, rt_grid, dim3(kWave), 0, s, rec, chord_t, rayleigh, L, n, stride, tiles,
                     win, batch, NS, kslots, keep, partial, reach);
  if (hminus)
    hipLaunchKernelGGL((opacityAdjointKernel<NS, true>), grid, dim3(kWave), 0, s, tables, stride, rec, L, n, tiles,
                       win, batch, keep, hminus, kslots, partial);
  else
    hipLaunchKernelGGL(opacityAdjointKernel<NS>, grid, dim3(kWave), 0, s, tables, stride, rec, L, n, tiles, win,
                       batch, keep, hminus, kslots, partial);
}

size_t align8(size_t x) { return (x + 7) / 8 * 8; }

}  // namespace

void DeviceEngine::releaseAdjoint() {
  for (void* p : {(void*)adj_fbar_, (void*)adj_gbar_, (void*)adj_fscale_, (void*)adj_terms_, (void*)adj_slope_,
                  (void*)adj_intensity_, (void*)adj_partial_, (void*)adj_out_})
    deviceFree(p);
  adj_fbar_ = adj_gbar_ = adj_fscale_ = adj_terms_ = adj_slope_ = adj_intensity_ = adj_partial_ = adj_out_ = nullptr;
  adj_capacity_ = adj_grid_ = adj_partial_size_ = adj_intensity_size_ = 0;
}

void DeviceEngine::evaluateAdjointBatch(ModelKind kind, size_t batch, std::vector<double>& level_grad,
                                        std::vector<double>& bin_grad, std::vector<double>& bins,
                                        std::vector<double>& bin_terms, bool staged) {
  if (staged && (!stage_blob_ || batch > stage_capacity_)) raise(kInternal, "device staging not prepared");
  const bool transmission = kind == ModelKind::transmission;
  if (transmission != transmission_) raise(kInternal, "adjoint model kind differs from the configured model");
  const bool eclipse = kind == ModelKind::eclipse;
  if (precision_ != TablePrecision::fp64) raise(kInvalidInput, "the adjoint needs the FP64 tables");
  if (cloud_models_ > 0 && kind == ModelKind::transmission)
    raise(kInvalidInput, "the transmission adjoint does not take cloud layers");
  if (batch == 0 || batch > batch_capacity_) raise(kInternal, "adjoint batch larger than the staged records");
  const int NS = int(species_stride_);
  if (NS != 4 && NS != 8 && NS != 12 && NS != 16) raise(kInvalidInput, "the adjoint takes at most 16 tabulated species");
  check(hipSetDevice(device_), "hipSetDevice");
  hipStream_t s = static_cast<hipStream_t>(stream_);
  // level slots: NS + 4, 8 more for the H- continuum's weights, 16 more for
  // the cloud models' layer sums
  const int cloud_slot0 = cloud_models_ > 0 ? NS + (hminus_ ? 12 : 4) : -1;
  const int n = int(n_points_), L = int(nb_levels_), nb = int(n_bins_),
            K = NS + (hminus_ ? 12 : 4) + (cloud_models_ > 0 ? 4 * kMaxClouds : 0);
  const int tiles = (n + kWave - 1) / kWave;
  const int grid = int(align8(std::max<size_t>(8, size_t(nb_cus_) * kAdjointWavesPerCu)));
  const int rt_grid = int(align8(std::max<size_t>(8, size_t(nb_cus_) * kRtWavesPerCu)));

  // points per chunk: kept forward quantities within the budget (and 80 % of
  // the free memory, checked when the buffers grow), else recomputed
  // (emission and eclipse); transmission also holds per point the chord
  // gradients and their per-range partial sums, and when one point's kept
  // quantities exceed the budget it takes one point at a time with its grid
  // in windows of whole chord ranges, running the forward kernel again with
  // the kept quantities for each window
  const int P = L - 1;
  // chord ranges: fixed by the grid, none empty, each a multiple of 128
  // wavenumbers (the batched transmission kernel's widest tile)
  int ranges = transmission ? std::max(1, std::min(kChordRanges, (n + 511) / 512)) : 1;
  int nu_per_range = 0;
  for (;;) {
    nu_per_range = int(((n + ranges - 1) / ranges + kWindowAlign - 1) / kWindowAlign * kWindowAlign);
    if (ranges == 1 || (ranges - 1) * nu_per_range < n) break;
    --ranges;
  }
  // chord partial sums, chord gradients, transposed chords and the reach
  // table (int32 [tiles][P]) of each point
  const size_t chord_point = transmission ? size_t(P) * P * (kChordRanges + 1) + size_t(P) * segmentRows(L) +
                                                (size_t(tiles) * P + 1) / 2
                                          : 0;                                   // doubles per point
  const size_t partial_point = size_t(L) * K * size_t(tiles);               // doubles per point
  const size_t kept_values = size_t(L) * 3 * stride_;                          // FP32 values per point
  const size_t kept_point = (kept_values + 1) / 2 + partial_point + stride_ + chord_point;  // doubles per point
  const size_t kept_range = size_t(L) * 3 * size_t(nu_per_range);             // FP32 values per chord range
  size_t budget = adjoint_budget_;
  size_t chunk = 0, buffer_size = 0, partial_size = 0;
  int window_ranges = ranges;                                                  // chord ranges per window
  bool kept = false;
  auto plan = [&] {
    chunk = std::min(batch, budget / (kept_point * sizeof(double)));
    kept = chunk > 0;
    window_ranges = ranges;
    if (!kept && transmission) {
      const size_t fixed = partial_point + stride_ + chord_point;             // doubles, one point
      const size_t room = budget / sizeof(double);
      window_ranges = room > fixed ? int(std::min<size_t>(ranges, (room - fixed) / ((kept_range + 1) / 2))) : 0;
      if (window_ranges > 0) {
        chunk = 1;
        kept = true;
      }
    }
    if (!kept)
      chunk = std::min(batch, std::max<size_t>(1, kPartialBudget / ((partial_point + stride_) * sizeof(double))));
    const size_t chunks = (batch + chunk - 1) / chunk;                      // equal chunks, no small tail
    chunk = (batch + chunks - 1) / chunks;
    if (!kept) buffer_size = size_t(grid) * L * 2 * kWave;                   // doubles
    else if (window_ranges < ranges) buffer_size = (size_t(window_ranges) * kept_range + 1) / 2;
    else buffer_size = (chunk * kept_values + 1) / 2;
    partial_size = chunk * (partial_point + chord_point);
  };
  plan();
  if (transmission && !kept)
    raise(kInvalidInput, "the transmission adjoint keeps one point's opacity and path adjoints for at least 1/" +
                             std::to_string(ranges) + " of its grid (" +
                             std::to_string((partial_point + stride_ + chord_point + (kept_range + 1) / 2) *
                                            sizeof(double)) +
                             " bytes): raise the adjoint memory");
  const bool grow = batch_capacity_ > adj_capacity_ || size_t(grid) != adj_grid_ ||
                    partial_size > adj_partial_size_ || buffer_size > adj_intensity_size_;
  if (grow) {
    size_t free_bytes = 0, total_bytes = 0;
    check(hipMemGetInfo(&free_bytes, &total_bytes), "hipMemGetInfo");
    const size_t held = (adj_intensity_size_ + adj_partial_size_) * sizeof(double);
    budget = std::min(adjoint_budget_, (free_bytes + held) / 10 * 8);
    plan();
    if (transmission && !kept) raise(kOutOfMemory, "not enough free device memory for the transmission adjoint");
    releaseAdjoint();
    adj_capacity_ = batch_capacity_;
    adj_grid_ = size_t(grid);
    adj_partial_size_ = partial_size;
    adj_intensity_size_ = buffer_size;
    adj_gbar_ = deviceAlloc<double>(adj_capacity_ * std::max<size_t>(1, n_bins_), "adjoint bins");
    adj_fscale_ = deviceAlloc<double>(adj_capacity_ * std::max<size_t>(1, n_bins_), "adjoint planet bins");
    adj_terms_ = deviceAlloc<double>(adj_capacity_ * std::max<size_t>(1, n_bins_) * 2, "adjoint eclipse terms");
    adj_slope_ = deviceAlloc<double>(adj_capacity_ * std::max<size_t>(1, n_chunks_), "adjoint eclipse star slopes");
    adj_out_ = deviceAlloc<double>(adj_capacity_ * L * K, "adjoint level gradients");
    adj_fbar_ = deviceAlloc<double>(adj_capacity_ * stride_, "adjoint flux");
    adj_intensity_ = deviceAlloc<double>(buffer_size, "adjoint forward quantities");
    adj_partial_ = deviceAlloc<double>(partial_size, "adjoint partial sums");
  }
  if (hminus_ && !kept)
    raise(kInvalidInput, "the H- continuum's gradient needs the kept forward quantities: raise the adjoint memory");
  if (cloud_models_ > 0 && !kept)
    raise(kInvalidInput, "the clouds' gradient needs the kept forward quantities: raise the adjoint memory");
  const BatchLaunch launch = tuned(transmission ? ModelKind::transmission : ModelKind::emission, chunk);
  bin_terms.assign(transmission ? batch * P * P : eclipse ? batch * n_bins_ * 2 : 0, 0.0);
  if (profiling_) check(hipEventRecord(static_cast<hipEvent_t>(event_start_), s), "hipEventRecord");
  if (staged) {
    launchStage(batch);
    check(hipMemcpyAsync(stage_neglect_host_, stage_neglect_, batch * sizeof(int), hipMemcpyDeviceToHost, s),
          "read neglect flags");
  } else {
    check(hipMemcpyAsync(batch_dev_, batch_host_, batch * rec_bytes_, hipMemcpyHostToDevice, s), "upload batch");
  }
  for (size_t first = 0; first < batch; first += chunk) {
    const int cb = int(std::min(chunk, batch - first));
    const Records rec{batch_dev_ + first * rec_bytes_, rec_bytes_, rec_levels_, rec_species_, rec_adjoint_,
                      rec_shifts_, groups_};
    float* keep = kept ? reinterpret_cast<float*>(adj_intensity_) : nullptr;
    const bool windowed = transmission && window_ranges < ranges;   // the kept forward runs per window
    // transmission: chord partial sums, chord gradients, transposed chords, reach
    double* chord_partial = adj_partial_ + size_t(cb) * partial_point;
    double* chord_out = chord_partial + size_t(cb) * P * P * kChordRanges;
    double* chord_t = chord_out + size_t(cb) * P * P;
    int* reach = reinterpret_cast<int*>(chord_t + size_t(cb) * P * segmentRows(L));
    launchBatchSpectrum(transmission ? ModelKind::transmission : ModelKind::emission, size_t(cb), launch, first,
                        windowed ? nullptr : keep, 0, 0, transmission && !windowed ? reach : nullptr);
    launchBatchRest(kind, size_t(cb), first);
    double* gbar = adj_gbar_ + first * n_bins_;
    double* fscale = eclipse ? adj_fscale_ + first * n_bins_ : gbar;
    if (nb > 0 && eclipse) {
      const int pieces = int(n_chunks_);
      hipLaunchKernelGGL(eclipseSlopeKernel, dim3(unsigned(pieces), unsigned(cb)), dim3(kBlockBins), 0, s,
                         wavenumber_, weights_, bin_chunks_, pieces, rec, adj_slope_);
      check(hipGetLastError(), "adjoint eclipse star kernel");
      hipLaunchKernelGGL(eclipseGbarKernel, dim3(unsigned((nb + kBlock - 1) / kBlock), unsigned(cb)), dim3(kBlock),
                         0, s, batch_chunk_sums_ + first * n_chunks_ * 2, adj_slope_, pieces, bins_desc_, nb, rec,
                         batch_bins_ + first * n_bins_, gbar, fscale, adj_terms_ + first * n_bins_ * 2);
      check(hipGetLastError(), "adjoint eclipse bin kernel");
    } else if (nb > 0) {
      hipLaunchKernelGGL(gbarKernel, dim3(unsigned((nb + kBlock - 1) / kBlock), unsigned(cb)), dim3(kBlock), 0, s,
                         batch_bins_ + first * n_bins_, bins_desc_, nb, rec, gbar);
      check(hipGetLastError(), "adjoint bin kernel");
    }
    hipLaunchKernelGGL(fluxAdjointKernel, dim3(unsigned((n + kBlock - 1) / kBlock), unsigned(cb)), dim3(kBlock), 0,
                       s, adj_ptr_, adj_bin_, adj_weight_, fscale, nb, n, stride_, adj_fbar_);
    check(hipGetLastError(), "adjoint flux kernel");
    if (transmission) {
      const int rows = segmentRows(L);
      hipLaunchKernelGGL(chordTransposeKernel, dim3(unsigned((P * rows + kBlock - 1) / kBlock), unsigned(cb)),
                         dim3(kBlock), 0, s, batch_chord_, L, chord_t);
      for (int r0 = 0; r0 < ranges; r0 += window_ranges) {
        const int r1 = std::min(ranges, r0 + window_ranges);
        TileWindow win = wholeGrid(tiles);
        if (windowed) {                                   // one point: rerun its forward over the window, kept
          const size_t nu_first = size_t(r0) * nu_per_range;
          const size_t nu_end = std::min(size_t(n), size_t(r1) * nu_per_range);
          win.first = int(nu_first / kWave);
          win.count = int((nu_end + kWave - 1) / kWave) - win.first;
          win.kept_origin = nu_first / kWave * size_t(L) * kKeptLevel;
          launchBatchSpectrum(ModelKind::transmission, 1, launch, first, keep, nu_first, nu_end - nu_first, reach);
        }
        switch (NS) {
#define BEAR_TADJ(NSV)                                                                                             \
  case NSV:                                                                                                        \
    launchTransmissionAdjoint<NSV>(dim3(unsigned(grid)), dim3(unsigned(rt_grid)), s, tables_, stride_, rayleigh_,  \
                                   rec, L, n, tiles, win, cb, adj_fbar_, batch_flux_ + first * stride_, keep,     \
                                   adj_partial_, ranges, r0, r1, nu_per_range, chord_partial, chord_t, reach,     \
                                   hminus_, K);                                                                   \
    break;
          BEAR_TADJ(4)
          BEAR_TADJ(8)
          BEAR_TADJ(12)
          BEAR_TADJ(16)
#undef BEAR_TADJ
        }
        check(hipGetLastError(), "transmission adjoint kernels");
      }
      hipLaunchKernelGGL(chordReduceKernel, dim3(unsigned((P * P + kBlock - 1) / kBlock), unsigned(cb)),
                         dim3(kBlock), 0, s, chord_partial, P, ranges, chord_out);
      check(hipGetLastError(), "chord reduction kernel");
      check(hipMemcpyAsync(bin_terms.data() + first * P * P, chord_out, size_t(cb) * P * P * sizeof(double),
                           hipMemcpyDeviceToHost, s),
            "read chord gradients");
    } else {
    switch (NS) {
#define BEAR_ADJ(NSV)                                                                                              \
  case NSV:                                                                                                        \
    launchAdjoint<NSV>(kept, dim3(unsigned(grid)), dim3(unsigned(rt_grid)), s, wavenumber_, wavelength_, tables_,   \
                       stride_, rec, L, n, tiles, cb, adj_fbar_, adj_intensity_, adj_partial_, hminus_, K,         \
                       cloud_slot0);                                                                               \
    break;
      BEAR_ADJ(4)
      BEAR_ADJ(8)
      BEAR_ADJ(12)
      BEAR_ADJ(16)
#undef BEAR_ADJ
    }
    check(hipGetLastError(), "adjoint kernels");
    }
    double* level_out = adj_out_ + first * L * K;
    switch (K) {
#define BEAR_RED(KV)                                                                                               \
  case KV:                                                                                                         \
    hipLaunchKernelGGL(adjointReduceKernel<KV>, dim3(unsigned(L), unsigned(cb)), dim3(kBlock), 0, s, adj_partial_, \
                       L, tiles, level_out);                                                                       \
    break;
      BEAR_RED(8)
      BEAR_RED(12)
      BEAR_RED(16)
      BEAR_RED(20)
      BEAR_RED(24)
      BEAR_RED(28)
      BEAR_RED(32)
      BEAR_RED(36)
      BEAR_RED(40)
      BEAR_RED(44)
#undef BEAR_RED
    }
    check(hipGetLastError(), "adjoint reduction kernel");
  }
  if (profiling_) check(hipEventRecord(static_cast<hipEvent_t>(event_stop_), s), "hipEventRecord");
  level_grad.resize(batch * L * K);
  bin_grad.resize(batch * n_bins_);
  bins.resize(batch * n_bins_);
  check(hipMemcpyAsync(level_grad.data(), adj_out_, level_grad.size() * sizeof(double), hipMemcpyDeviceToHost, s),
        "read level gradients");
  if (eclipse && nb > 0)
    check(hipMemcpyAsync(bin_terms.data(), adj_terms_, bin_terms.size() * sizeof(double), hipMemcpyDeviceToHost, s),
          "read eclipse bin terms");
  if (nb > 0) {
    check(hipMemcpyAsync(bin_grad.data(), adj_gbar_, bin_grad.size() * sizeof(double), hipMemcpyDeviceToHost, s),
          "read bin gradients");
    check(hipMemcpyAsync(bins.data(), batch_bins_, bins.size() * sizeof(double), hipMemcpyDeviceToHost, s),
          "read bins");
  }
  check(hipStreamSynchronize(s), "adjoint");
  if (profiling_) {
    float ms = 0;
    check(hipEventElapsedTime(&ms, static_cast<hipEvent_t>(event_start_), static_cast<hipEvent_t>(event_stop_)),
          "hipEventElapsedTime");
    batch_ms_ += ms;
  }
  batch_evaluations_ += batch;
  last_used_graph_ = false;
  adj_last_chunk_ = kept ? chunk : 0;
}

}  // namespace bear_amd
