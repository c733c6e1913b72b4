// Device arithmetic shared by the kernels: FP64 exp2 and reciprocal, the
// Planck function, the blackbody star and one short-characteristics layer.
#pragma once

#include <hip/hip_runtime.h>

#include "engine.h"

namespace bear_amd {
namespace dev {

// Short-characteristics Gauss nodes and weights (short_characteristics.h:70-71)
constexpr double kMu1 = 0.211324865405187;
constexpr double kMu2 = 0.788675134594813;
constexpr double kW = 0.5;

// Planck constants in the reference's association order
// (additional/aux_functions.cpp:272-278): 2 h c c nu^3 / (exp(h nu c / k / T) - 1)
constexpr double kPlanckC1 = 2. * 6.62606896e-27 * 2.99792458e10 * 2.99792458e10;
constexpr double kPi = 3.14159265358979323846;
constexpr double kLog2Ten = 3.321928094887362347870319429489390175864831393;
constexpr double kLog2E = 1.442695040888963407359924681001892137426645954;
constexpr double kLn2 = 0.693147180559945309417232121458176568075500134;

// 2^t in FP64 with relative error below 3.1e-15: t = n + f, |f| <= 1/2,
// 2^f from its degree-10 Chebyshev interpolant on [-1/2, 1/2] (coefficients
// in monomial form), then an exponent shift.  For t <= 0 (the attenuation
// and transmittance factors e^-tau); t is clamped below so that the shift
// stays in range: results below 2^-1100 flush to 0.
__device__ __forceinline__ double pow2(double t) {
  t = fmax(t, -1100.0);
  const double n = __builtin_rint(t);
  const double f = t - n;
  double p = 6.767963482574983e-09;
  p = fma(p, f, 1.0206977921453389e-07);
  p = fma(p, f, 1.3217210256367588e-06);
  p = fma(p, f, 1.525266622014565e-05);
  p = fma(p, f, 0.00015403526882354606);
  p = fma(p, f, 0.0013333558214738293);
  p = fma(p, f, 0.00961812911053657);
  p = fma(p, f, 0.055504108664544756);
  p = fma(p, f, 0.24022650695901254);
  p = fma(p, f, 0.6931471805599453);
  p = fma(p, f, 1.0);
  return __builtin_amdgcn_ldexp(p, int(n));
}

// 2^t with two polynomial terms fewer: degree-8 Chebyshev interpolant of 2^f
// on [-1/2, 1/2] (relative error 1.1e-12), exponent shift.  For quantities
// that enter the result linearly: the opacity terms (they add, nothing
// cancels) and the Planck function.  The attenuation factors e^-delta of the
// radiative transfer keep pow2: (e^-delta - 1) / delta cancels for thin
// layers and magnifies the exponential's error by 1 / delta.
// t >= -1100 (so -inf, a zero density, gives 0); t is bounded above by the
// physics (sigma n < 2^1000, Planck arguments <= 0).
__device__ __forceinline__ double pow2Lean(double t) {
  t = fmax(t, -1100.0);
  const double n = __builtin_rint(t);
  const double f = t - n;
  double p = 1.3255225648015585e-06;
  p = fma(p, f, 1.5310079150997826e-05);
  p = fma(p, f, 0.00015403455848948458);
  p = fma(p, f, 0.001333345056925233);
  p = fma(p, f, 0.009618129159399412);
  p = fma(p, f, 0.055504109412040106);
  p = fma(p, f, 0.24022650695813164);
  p = fma(p, f, 0.6931471805459343);
  p = fma(p, f, 0.9999999999999999);
  return __builtin_amdgcn_ldexp(p, int(n));
}

// pow2Lean for the opacity terms 2^(log2 n + sum_k w_k v_k), whose argument
// is finite and above -2300 (kZeroDensityLog2 for a zero density): no clamp
__device__ __forceinline__ double pow2Opacity(double t) {
  const double n = __builtin_rint(t);
  const double f = t - n;
  double p = 1.3255225648015585e-06;
  p = fma(p, f, 1.5310079150997826e-05);
  p = fma(p, f, 0.00015403455848948458);
  p = fma(p, f, 0.001333345056925233);
  p = fma(p, f, 0.009618129159399412);
  p = fma(p, f, 0.055504109412040106);
  p = fma(p, f, 0.24022650695813164);
  p = fma(p, f, 0.6931471805459343);
  p = fma(p, f, 0.9999999999999999);
  return __builtin_amdgcn_ldexp(p, int(n));
}

// 2^t with relative error about 1.2e-7, for the gradient's opacity terms
// (whose bar is 1e-4): t = n + f in FP64, 2^f by the hardware's FP32
// exponential, then the exponent shift; t is clamped below as in pow2Lean
__device__ __forceinline__ double pow2Gradient(double t) {
  t = fmax(t, -1100.0);
  const double n = __builtin_rint(t);
  const float f = float(t - n);
  return __builtin_amdgcn_ldexp(double(__builtin_amdgcn_exp2f(f)), int(n));
}

// 1/d for finite non-zero d: hardware estimate and two Newton steps
__device__ __forceinline__ double recip(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(fma(-d, r, 1.0), r, r);
  r = fma(fma(-d, r, 1.0), r, r);
  return r;
}

// Planck function per wavenumber written as c1 nu^3 q / (1 - q),
// q = e^(-h nu c / k T): the reference's expression without exp overflow
__device__ __forceinline__ double planckAt(double b_num, double x_nu, double inv_t) {
  const double q = pow2Lean(x_nu * inv_t);
  return b_num * q * recip(1.0 - q);
}

// blackbody star flux pi B(T) 1e-3 (star_blackbody.cpp:40-54)
__device__ __forceinline__ double starFlux(double nu, double inv_t) {
  const double x_nu = -(6.62606896e-27 * nu * 2.99792458e10 / 1.3806504e-16) * kLog2E;
  return planckAt(kPlanckC1 * nu * nu * nu, x_nu, inv_t) * kPi * 1e-3;
}

// starFlux and its derivative with respect to the star's temperature:
// d B / dT = B ln2 (-x) / T^2 / (1 - q) for q = 2^(x / T)
__device__ __forceinline__ void starFluxWithDerivative(double nu, double inv_t, double& flux, double& d_flux) {
  const double x_nu = -(6.62606896e-27 * nu * 2.99792458e10 / 1.3806504e-16) * kLog2E;
  const double q = pow2Lean(x_nu * inv_t);
  const double r = recip(1.0 - q);
  flux = kPlanckC1 * nu * nu * nu * q * r * kPi * 1e-3;
  d_flux = flux * (-x_nu * kLn2) * inv_t * inv_t * r;
}

// One layer of the two-angle short characteristics (short_characteristics.cpp:88-114)
__device__ __forceinline__ void shortCharacteristicsLayer(double tau, double b_prev, double b_next,
                                                          double& i1, double& i2) {
  if (tau != 0) {
    const double r = recip(tau);
    const double a1 = pow2(tau * (-kLog2E / kMu1));
    const double q1 = (a1 - 1.0) * (kMu1 * r);
    i1 = fma(i1, a1, (1.0 + q1) * b_next + (-a1 - q1) * b_prev);
    const double a2 = pow2(tau * (-kLog2E / kMu2));
    const double q2 = (a2 - 1.0) * (kMu2 * r);
    i2 = fma(i2, a2, (1.0 + q2) * b_next + (-a2 - q2) * b_prev);
  }
}

// The opacity terms of NH consecutive stencil slots at one level,
// sum_s 2^(log2 n_s + sum_k w_sk v_sk) with the slots' four table values v
// kept in registers between levels and all reloaded when `reload` (the
// level brings new rows for some slot, or the first level)
// (opacity_species.cpp:277-345, 400-433: the bilinear log10 interpolation in
// T and log p, times the effective density).  Rows are always valid; a zero
// density (log2 n = -inf) removes a term exactly.
// With grouped stencils (StencilGroups) the weights come from the level's two
// shared sets, group1 holding these NH slots' group bits, and the densities
// from the level's contiguous copy (NH <= 16, the first NH slots); otherwise
// both from each slot's record.
// The two halves are also available apart (loadRows, kappaFrom), so that a
// kernel can issue the next level's row loads before this level's other
// work and find them arrived when it needs them.
template <int NH>
__device__ __forceinline__ void loadRows(const double* __restrict__ tab, const LevelSpecies* __restrict__ d,
                                         double (&v)[NH][4]) {
#pragma unroll
  for (int s = 0; s < NH; ++s)
#pragma unroll
    for (int q = 0; q < 4; ++q) v[s][q] = tab[size_t(d[s].block[q]) << 6];
}

template <int NH>
__device__ __forceinline__ double kappaFrom(const double (&v)[NH][4], const LevelSpecies* __restrict__ d,
                                            const LevelState& lev, StencilGroups groups);

template <int NH>
__device__ __forceinline__ double kappaTerms(const double* __restrict__ tab, size_t stride,
                                             const LevelSpecies* __restrict__ d, const LevelState& lev,
                                             StencilGroups groups, bool reload, double (&v)[NH][4]) {
  if (reload) loadRows<NH>(tab, d, v);
  return kappaFrom<NH>(v, d, lev, groups);
}

template <int NH>
__device__ __forceinline__ double kappaFrom(const double (&v)[NH][4], const LevelSpecies* __restrict__ d,
                                            const LevelState& lev, StencilGroups groups) {
  double k = 0.0;
  if (groups.grouped) {
    double a[4], b[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a[q] = lev.group_w[0][q];
      b[q] = lev.group_w[1][q];
    }
#pragma unroll
    for (int s = 0; s < NH; ++s) {
      const bool one = (groups.group1 >> s) & 1u;
      double x = fma(one ? b[0] : a[0], v[s][0], lev.log2_n[s]);
      x = fma(one ? b[1] : a[1], v[s][1], x);
      x = fma(one ? b[2] : a[2], v[s][2], x);
      x = fma(one ? b[3] : a[3], v[s][3], x);
      k += pow2Opacity(x);
    }
  } else {
#pragma unroll
    for (int s = 0; s < NH; ++s) {
      double x = fma(d[s].w[0], v[s][0], d[s].log2_n);
      x = fma(d[s].w[1], v[s][1], x);
      x = fma(d[s].w[2], v[s][2], x);
      x = fma(d[s].w[3], v[s][3], x);
      k += pow2Opacity(x);
    }
  }
  return k;
}

// The H- continuum (h_m.cpp:38-53, 56-82, 86-137): per wavenumber the
// bound-free cross-section and the six free-free wavelength polynomials
// (rows [7][n], read once per thread by loadExtras), per level their weights
// LevelState::hminus_c, so kappa_H- = c0 sigma_bf + sum_i c_i P_i.
__device__ __forceinline__ double hminusKappa(const double (&hm)[7], const LevelState& lev) {
  double k = lev.hminus_c[0] * hm[0];
#pragma unroll
  for (int r = 1; r < 7; ++r) k = fma(lev.hminus_c[r], hm[r], k);
  return k;
}

// per thread, once: the H- rows and the cloud models' spectral factors at
// wavelength lam (micron) (power_law_cloud_model.cpp:114-120,
// kh_cloud_model.cpp:116-132); unused slots are zero
__device__ __forceinline__ void loadExtras(const Extras& ex, int i, int n, double lam, double (&hm)[7],
                                           double (&cf)[kMaxClouds]) {
#pragma unroll
  for (int r = 0; r < 7; ++r) hm[r] = ex.hminus ? ex.hminus[size_t(r) * n + i] : 0.0;
#pragma unroll
  for (int m = 0; m < kMaxClouds; ++m) {
    double f = 0.0;
    if (m < ex.clouds.count) {
      const double* c = ex.clouds.c[m];
      if (ex.clouds.kind[m] == kCloudPowerLaw) {
        f = pow(lam, c[0]) * c[1];
      } else if (ex.clouds.kind[m] == kCloudKH) {
        const double x = c[2] / lam;
        f = c[3] / (c[0] * pow(x, -c[1]) + pow(x, 0.2));
      } else {
        f = 1.0;
      }
    }
    cf[m] = f;
  }
}

// the clouds' optical depth in the layer below level l (l >= 1): the cloud
// models' weights times their factors, plus the caller-given depth
// (opacity_calc.h:221-237 sums them)
__device__ __forceinline__ double cloudLayer(const Extras& ex, const double (&cf)[kMaxClouds], const LevelState& lev,
                                             int l, int i, int n) {
  double t = 0.0;
#pragma unroll
  for (int m = 0; m < kMaxClouds; ++m) t = fma(lev.cloud_w[m], cf[m], t);
  if (ex.cloud_tau) t += ex.cloud_tau[size_t(l - 1) * n + i];
  return t;
}

// The same without a register cache, for any number of species: chunks of
// kChunk slots with every row of a chunk loaded before any is used; weights
// and densities from each slot's record.
constexpr int kChunk = 4;

__device__ __forceinline__ double chunkedKappa(const double* __restrict__ tab, size_t stride,
                                               const LevelSpecies* __restrict__ d, const LevelState& lev,
                                               int nb_chunks) {
  double k = 0.0;
  for (int c = 0; c < nb_chunks; ++c, d += kChunk) {
    double v[kChunk][4];
    k += kappaTerms<kChunk>(tab, stride, d, lev, StencilGroups{0u, 0u}, true, v);      // per-slot records
  }
  return k;
}

// sum over the 64 lanes of a wavefront (the result is in lane 0)
__device__ __forceinline__ double waveSum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// v of another lane through DPP moves of both halves (every row, every bank)
template <int kCtrl>
__device__ __forceinline__ double dppMove(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_mov_dpp(int(unsigned(b)), kCtrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(int(unsigned(b >> 32)), kCtrl, 0xF, 0xF, false);
  return __builtin_bit_cast(double, (unsigned long long)unsigned(lo) | ((unsigned long long)unsigned(hi) << 32));
}

__device__ __forceinline__ double joinHalves(unsigned lo, unsigned hi) {
  return __builtin_bit_cast(double, (unsigned long long)lo | ((unsigned long long)hi << 32));
}

// the largest of an int over the 64 lanes of a wavefront, in every lane:
// v_permlane32_swap and v_permlane16_swap pair lanes j, j+32 and then j,
// j+16, DPP moves finish within rows of 16 (as waveTotals does for sums)
__device__ __forceinline__ int waveMaxInt(int v) {
  const auto a = __builtin_amdgcn_permlane32_swap(unsigned(v), unsigned(v), false, false);
  v = max(int(a[0]), int(a[1]));
  const auto b = __builtin_amdgcn_permlane16_swap(unsigned(v), unsigned(v), false, false);
  v = max(int(b[0]), int(b[1]));
  v = max(v, __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false));      // quad_perm [1,0,3,2]
  v = max(v, __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false));      // quad_perm [2,3,0,1]
  v = max(v, __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, false));     // row_half_mirror
  v = max(v, __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, false));     // row_mirror
  return v;
}

// Sums over the 64 lanes of a wavefront of K values at once (K a multiple of
// 4), all in VALU registers.  Two halving steps first: v_permlane32_swap
// exchanges the upper 32 lanes of value 2a with the lower 32 of value 2a+1,
// so one add leaves value 2a summed over lane pairs (j, j+32) in lanes 0-31
// and value 2a+1 in lanes 32-63; v_permlane16_swap does the same between
// rows.  Then each remaining value is summed within its rows of 16 (DPP
// quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror, row_mirror).  On return,
// every lane of row r holds in x[b] the total of value 4b + kRowSlot[r].
template <int K>
__device__ __forceinline__ void waveTotals(double (&x)[K]) {
  static_assert(K % 4 == 0, "values in groups of four");
#pragma unroll
  for (int a = 0; a < K / 2; ++a) {
    const unsigned long long p = __builtin_bit_cast(unsigned long long, x[2 * a]);
    const unsigned long long q = __builtin_bit_cast(unsigned long long, x[2 * a + 1]);
    const auto lo = __builtin_amdgcn_permlane32_swap(unsigned(p), unsigned(q), false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(unsigned(p >> 32), unsigned(q >> 32), false, false);
    x[a] = joinHalves(lo[0], hi[0]) + joinHalves(lo[1], hi[1]);
  }
#pragma unroll
  for (int b = 0; b < K / 4; ++b) {
    const unsigned long long p = __builtin_bit_cast(unsigned long long, x[2 * b]);
    const unsigned long long q = __builtin_bit_cast(unsigned long long, x[2 * b + 1]);
    const auto lo = __builtin_amdgcn_permlane16_swap(unsigned(p), unsigned(q), false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(unsigned(p >> 32), unsigned(q >> 32), false, false);
    x[b] = joinHalves(lo[0], hi[0]) + joinHalves(lo[1], hi[1]);
  }
#pragma unroll
  for (int b = 0; b < K / 4; ++b) {
    double v = x[b];
    v += dppMove<0xB1>(v);        // quad_perm [1,0,3,2]
    v += dppMove<0x4E>(v);        // quad_perm [2,3,0,1]
    v += dppMove<0x141>(v);       // row_half_mirror
    v += dppMove<0x140>(v);       // row_mirror
    x[b] = v;
  }
}
// value of row r after waveTotals: 4b + kRowSlot[r]
__device__ __forceinline__ int rowSlot(int row) { return row == 1 ? 2 : row == 2 ? 1 : row; }

// Optical depths of the tangent paths t0 .. t0+7 at this lane's wavenumber
// (transmission_spectrum.cpp:107-139): tau_t = sum over q >= t of
// c[t][q] (k_q + k_{q+1}) in increasing q, stopping at the first partial sum
// above 10.  seg[q * tile] holds k_q + k_{q+1} for this lane's column (zero
// in the padded rows), chord the transposed chord matrix with `tangents`
// columns; t0 (a multiple of 8) and chord are uniform over the wavefront, so
// the chords arrive by scalar loads.  All eight paths take four segments per
// step from one set of LDS reads; a path that first passes 10 within a step
// is summed again one segment at a time from its value before the step (the
// same fused operations in the same order), so it keeps the partial sum at
// which the reference stops.  Paths already past 10 keep accumulating, which
// leaves their stored value untouched.  kStop: also the last segment each
// path summed (nb_segments - 1 when it never passed 10) to stop[].
template <class K, bool kStop = false>
__device__ __forceinline__ void tangentBlock(const K* __restrict__ seg, int tile, const K* __restrict__ chord,
                                             int tangents, int t0, int nb_segments, K (&out)[kTangentBlock],
                                             int* stop = nullptr) {
  // fin[k]: the stopping sum once path k has passed 10 (always above 10), 0 before
  K tau[kTangentBlock], fin[kTangentBlock];
  int last[kTangentBlock];
#pragma unroll
  for (int k = 0; k < kTangentBlock; ++k) {
    tau[k] = K(0);
    fin[k] = K(0);
    last[k] = nb_segments - 1;
  }
  for (int q0 = t0; q0 < nb_segments; q0 += kSegmentStep) {
    K sq[kSegmentStep];
#pragma unroll
    for (int u = 0; u < kSegmentStep; ++u) sq[u] = seg[(q0 + u) * tile];
    const K* c = chord + size_t(q0) * tangents + t0;
    K prev[kTangentBlock];
#pragma unroll
    for (int k = 0; k < kTangentBlock; ++k) prev[k] = tau[k];
#pragma unroll
    for (int u = 0; u < kSegmentStep; ++u)
#pragma unroll
      for (int k = 0; k < kTangentBlock; ++k) tau[k] = fma(c[u * tangents + k], sq[u], tau[k]);
    bool all = true;
#pragma unroll
    for (int k = 0; k < kTangentBlock; ++k) {
      const bool past = tau[k] > K(10);
      const bool first = past && fin[k] == K(0);
      if (__builtin_amdgcn_ballot_w64(first) != 0) {       // uniform branch: exact re-sum of path k
        K x = prev[k], hit = K(0);
        int hit_q = 0;
        bool found = false;
#pragma unroll
        for (int u = 0; u < kSegmentStep; ++u) {
          x = fma(c[u * tangents + k], sq[u], x);
          const bool now = !found && x > K(10);
          hit = now ? x : hit;
          if constexpr (kStop) hit_q = now ? q0 + u : hit_q;
          found = found || now;
        }
        fin[k] = first ? hit : fin[k];
        if constexpr (kStop) last[k] = first ? hit_q : last[k];
      }
      all = all && past;
    }
    if (__builtin_amdgcn_ballot_w64(!all) == 0) break;
  }
#pragma unroll
  for (int k = 0; k < kTangentBlock; ++k) out[k] = fin[k] != K(0) ? fin[k] : tau[k];
  if constexpr (kStop)
#pragma unroll
    for (int k = 0; k < kTangentBlock; ++k) stop[k] = last[k];
}

// Bins (secondary_eclipse.cpp:231-246, forward_model.cpp:161-180,
// retrieval.cpp:330-358), in two steps:
//  1. binChunkBody, one workgroup of kBlockBins threads per piece of a bin's
//     folded segment: the dot product of the spectrum with the weights (four
//     products per thread per round, so that a piece costs one or two memory
//     latencies), and with mode 1 (eclipse) the folded blackbody star; a
//     fixed tree to out[0], out[1];
//  2. binFinalizeBody, one workgroup of 1024 threads per evaluation: per bin
//     its pieces' sums in order, the value (eclipse: planet / star
//     (Rp/Rs)^2 1e6), the observation's shift and the Gaussian
//     log-likelihood term, and the terms' fixed-order sum.
constexpr int kBlockBins = 256;

__device__ __forceinline__ void binChunkBody(const double* __restrict__ flux, const double* __restrict__ wavenumber,
                                             const double* __restrict__ weights, const BinChunk& c, int mode,
                                             double star_inv_temperature, double* __restrict__ out) {
  __shared__ double partial[2][kBlockBins / 64];
  const double* w = weights + c.weight_offset;
  const double* f = flux + c.start;
  double a[4] = {0.0, 0.0, 0.0, 0.0}, acc_star = 0.0;
  uint32_t j = threadIdx.x;
  for (; j + 3 * kBlockBins < c.length; j += 4 * kBlockBins) {
#pragma unroll
    for (int u = 0; u < 4; ++u) a[u] = fma(w[j + u * kBlockBins], f[j + u * kBlockBins], a[u]);
  }
  for (; j < c.length; j += kBlockBins) a[0] = fma(w[j], f[j], a[0]);
  double acc = (a[0] + a[1]) + (a[2] + a[3]);
  if (mode == 1)
    for (uint32_t k = threadIdx.x; k < c.length; k += kBlockBins)
      acc_star += w[k] * starFlux(wavenumber[c.start + k], star_inv_temperature);
  acc = waveSum(acc);
  acc_star = waveSum(acc_star);
  if ((threadIdx.x & 63) == 0) {
    partial[0][threadIdx.x >> 6] = acc;
    partial[1][threadIdx.x >> 6] = acc_star;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = (partial[0][0] + partial[0][1]) + (partial[0][2] + partial[0][3]);
    out[1] = (partial[1][0] + partial[1][1]) + (partial[1][2] + partial[1][3]);
  }
}

// Step 2 for n bins with pieces' sums chunk_sums[piece][2]: bin values to
// out[bin], the log-likelihood to *ll (and *ll_host when given).
__device__ __forceinline__ void binFinalizeBody(const double* __restrict__ chunk_sums,
                                                const BinDesc* __restrict__ bins, int n, int mode,
                                                double depth_scale, const double* __restrict__ shifts,
                                                double error_inflation, double* __restrict__ out,
                                                double* __restrict__ ll, double* __restrict__ ll_host) {
  __shared__ double partial[16];
  double acc = 0.0;
  for (int j = threadIdx.x; j < n; j += 1024) {
    const BinDesc b = bins[j];
    const uint32_t pieces = (b.length + kBinChunk - 1) / kBinChunk;
    const double* cs = chunk_sums + size_t(b.first_chunk) * 2;
    double v = pieces > 0 ? cs[0] : 0.0, star = pieces > 0 ? cs[1] : 0.0;       // an empty segment sums to 0
    for (uint32_t k = 1; k < pieces; ++k) {
      v += cs[2 * k];
      star += cs[2 * k + 1];
    }
    if (mode == 1) v = v / star * depth_scale;
    const double shift = shifts[b.observation];
    if (shift != 0) v += shift;
    out[j] = v;
    const double e2 = b.error * b.error + error_inflation;
    const double d = b.data - v;
    acc += (-0.5 * log(e2 * 2.0 * kPi) - 0.5 * d * d / e2) * b.weight;
  }
  acc = waveSum(acc);
  if ((threadIdx.x & 63) == 0) partial[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < 16; ++k) s += partial[k];
    *ll = s;
    if (ll_host) *ll_host = s;
  }
}

// The adjoint's kept forward quantities, FP32 [point][64-wavenumber tile]
// [level][3][64] (per level three values: emission I_1 - B, I_2 - B, kappa;
// transmission k, tau, last segment), so that a wavefront's store or load
// of one value of one level is one contiguous 256-byte run next to the
// level's other values.  keptColumn: offset of (point m, level 0, value 0,
// wavenumber i); level l at + kKeptLevel l, value v at + kKeptValue v.
constexpr int kKeptValue = 64, kKeptLevel = 3 * kKeptValue;
__device__ __forceinline__ size_t keptColumn(int m, size_t stride, int levels, int i) {
  return (size_t(m) * (stride / kKeptValue) + size_t(i / kKeptValue)) * size_t(levels) * kKeptLevel +
         size_t(i % kKeptValue);
}

// ---------------------------------------------------------------- batch work items
// Per-point records of a batch: [BatchModel][LevelState L][stencils L x NS]
// [adjoint stencils L x NS (FP64 tables)][shifts]
struct Records {
  const unsigned char* base;
  size_t bytes, levels, species, adjoint, shifts;
  StencilGroups groups;
  __device__ __forceinline__ const unsigned char* at(int m) const { return base + size_t(m) * bytes; }
};

// item -> (tile, point): XCD x = item mod 8 owns tiles x, x + 8, ...; its
// items take the points in groups of `group` (kPointGroup unless the launch
// says otherwise) and run through every point of a group for one tile before
// the next tile, so that the L2 of each XCD holds the group's per-level
// records (about 50 KB per point) together with the table rows of the tiles
// in flight, which every point of the group reuses.  Transmission adds each
// point's chord matrix (320 KB at 200 levels), for which smaller groups pay.
constexpr int kPointGroup = 32;

__device__ __forceinline__ bool workItem(size_t item, int batch, int tiles, int& point, int& tile,
                                         int group = kPointGroup) {
  const size_t k = item / 8;                                   // this XCD's sequence
  const size_t per_xcd = size_t((tiles + 7) / 8);              // tiles per XCD, padded
  const size_t span = size_t(group) * per_xcd;                 // items of a full group
  const size_t g = k / span;
  const int first = int(g) * group;
  const int points = min(group, batch - first);                // the last group may be smaller
  const size_t w = k - g * span;
  point = first + int(w % size_t(points));
  tile = int(w / size_t(points)) * 8 + int(item % 8);
  return tile < tiles;
}

__host__ __device__ inline size_t nbItems(int tiles, int batch) { return size_t((tiles + 7) / 8) * 8 * size_t(batch); }

}  // namespace dev
}  // namespace bear_amd
