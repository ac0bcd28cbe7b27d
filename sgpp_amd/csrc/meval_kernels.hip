// gfx950 kernels of the multi-evaluation operator and the device runtime
// that launches them.
//
// Arithmetic per (data point x, grid point j):
//   phi_j(x) = prod_d max(0, 1 - |L_jd * x_d - I_jd|)
// evaluated as the two-instruction step
//   u = fma(L, x, -I);  p = clamp01(fma(-p, |u|, p))      (p in [0,1])
// which the compiler emits as v_fma + v_fma(neg, abs, clamp). This is VALU
// work (no contraction), so there is no MFMA anywhere in this file.
//
// Layout: data points SoA [D][Mpad] in Morton order (dperm maps a stored row
// to the caller's row); grid points point-major (L, I) rows [Npad][D][2],
// sorted by their set of non-constant dimensions and the Morton code of their
// support centre (perm maps a stored point to the caller's index), with
// per-64-point-tile support boxes and per-point dimension masks.
//
// Both kernels give lanes K data points each (coordinates in registers). A
// wave takes the box of its 64*K points, skips grid tiles whose box misses
// it, tests the remaining tiles' points against it with one ballot and
// evaluates the points that may be nonzero, J at a time, over the union of
// their non-constant dimensions (ModLinear level-1 factors are exactly 1).
// Every dropped term is exactly zero (cull_device.hpp), so results equal a
// full evaluation.
//   mult (r = B alpha): blocks = data tiles (x) x G grid chunks (y); the G
//     FP64 partial rows are summed in a fixed order by reduce_parts.
//   multTranspose (a = B^T y): blocks = data segments (x) x grid chunks of
//     kChunkT points (y); per taken grid point the wave's sum over its lanes
//     (DPP) goes to a per-tile register, and each tile is added to the
//     wave's row of FP64 LDS accumulators; the four rows and the S partial
//     rows are summed in a fixed order. No global atomics: results are
//     deterministic.
// Precision: FP64 throughout, or FP32 basis products and per-tile sums with
//   the per-tile partials accumulated in FP64 registers and FP64 partials.
//
// Reference semantics: datadriven/src/sgpp/datadriven/operation/hash/
//   OperationMultiEvalStreaming/OperationMultiEvalStreaming_multImpl.cpp:576-623 (mult),
//   ..._multTransposeImpl.cpp:483-508 (multTranspose),
//   OperationMultiEvalModMaskStreaming/..._multImpl.cpp:604-656 (ModLinear).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <sstream>

#include "cull_device.hpp"
#include "meval_device.hpp"

namespace sgpp_hip {

namespace {

constexpr int kBlock = 256;  // 4 waves of 64
constexpr int kTG = 64;      // grid points per tile (culling unit, one point per lane)
constexpr int kTM = 128;     // data points per step of the generic multTranspose
constexpr int kMaxTemplD = 20;
// Default launch sizes. Culled blocks cost very different amounts, so many
// short blocks balance better: mult ~128 workgroups per CU and grid chunks of
// at most 24 tiles; multTranspose ~256 workgroups per CU (each block covers
// kChunkT grid points). Measured on the d=10 ModLinear (1e5 x 1e6) and d=8
// Linear (3.2e4 x 5e5) workloads.
constexpr int kBlocksPerCU = 128;
constexpr int kMaxChunkTiles = 24;
constexpr int kBlocksPerCUT = 256;
// cap on the FP64 partials of one launch (G x Mpad or S x Npad doubles)
constexpr int64_t kPartialBudget = int64_t(512) << 20;
// data rows past M kept in the padded data (a wave's last group of 64*K
// points may start just below M)
constexpr size_t kDataSlack = 1024;

// Points per lane (K) and grid points evaluated together (J). Values from
// variant sweeps on the d=10 ModLinear, d=8 Linear and d=20 Linear workloads.
// Kernels for d > 10 are held to 128 registers (4 waves per SIMD) although
// some values then spill: at d = 20 that was 4-21 % faster than 3 waves
// without spills; at d = 10 (FP64) it was 10-15 % slower, so d <= 10 keeps
// its natural allocation.
template <typename T>
constexpr int mult_k(int D) {
  return sizeof(T) == 8 ? (D <= 10 ? 4 : 2) : (D <= 10 ? 8 : 4);
}
template <typename T>
constexpr int mult_j(int D) {
  return sizeof(T) == 4 && D > 10 ? 4 : 2;
}
template <typename T>
constexpr int multT_k(int D) {
  return sizeof(T) == 8 ? (D <= 10 ? 4 : 2) : (D <= 10 ? 8 : 4);
}
template <typename T>
constexpr int multT_j(int D) {
  return sizeof(T) == 4 && D > 10 ? 4 : 2;
}
// grid points per multTranspose block (FP64 accumulators in LDS, one row per wave)
constexpr int kChunkT = 1024;

// (L, I) of grid point j and dimension d: LI[(j*D + d)*2 + {0,1}] (point-major:
// one point's 2*D values are contiguous, so a wave fetches a taken point's
// whole row with one wide scalar load; tile t = points [64t, 64t+64)).
__device__ __forceinline__ int64_t li_at(int64_t j, int d, int D) { return (j * D + d) * 2; }

// ---------------------------------------------------------------- mult
// Lanes own data points (K per lane, coordinates in registers). The data are
// stored in Morton order, so a wave's 64*K points fill a small box. For every
// 64-point grid tile of its chunk a wave
//   1. skips the tile if the tile's box of supports misses its data box;
//   2. tests the tile's points against the box (one point per lane, ballot);
//   3. evaluates only points that may be nonzero in the box, J at a time,
//      over the union of their non-constant dimensions.
// Every dropped term is exactly zero, so the sums equal a full evaluation.
// X: [D][ldx] SoA, LI: point-major pairs, TB: [Npad/64][D][2] FP32 tile
// boxes, A: [Npad] alpha*scale (zero padded), part: [G][ldx]. Block: kBlock*K
// data points (each wave 64*K consecutive ones) x grid chunk [y*chunk,
// +chunk). kSkip: test for constant (ModLinear level-1) dimensions.
template <int D, int K, int J, typename T, bool kSkip>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(D > 10 ? 4 : 1)))
void mult_kernel(const T* __restrict__ X, int64_t ldx,
                                                      const T* __restrict__ LI,
                                                      const float* __restrict__ TB,
                                                      const T* __restrict__ A,
                                                      double* __restrict__ part, int chunk,
                                                      int jend_all) {
  // a wave's 64*K points are contiguous (a small box in Morton order): point
  // k of a lane is m0 + 64*k
  const int lane = threadIdx.x & 63;
  const int64_t m0 = static_cast<int64_t>(blockIdx.x) * (kBlock * K) + (threadIdx.x >> 6) * (64 * K) + lane;
  T x[K][D];
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int d = 0; d < D; ++d) x[k][d] = X[d * ldx + m0 + k * 64];
  WaveBox<D, T> box;
  wave_box<D, D, K, T>(x, box);
  double accd[K];
#pragma unroll
  for (int k = 0; k < K; ++k) accd[k] = 0.0;
  const int jb = blockIdx.y * chunk;
  const int je = min(jb + chunk, jend_all);
  for (int j0 = jb; j0 < je; j0 += kTG) {
    const int64_t t = j0 / kTG;
    if (!tile_may_hit<D, D, T>(TB, t, box)) continue;
    const T* tile = LI + t * D * kTG * 2;
    unsigned long long bal = tile_hits<D, D, T>(tile, lane, box);
    T acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = T(0);
    while (bal) {
      int jj[J];
      bool used[J];
      T a[J];
      take_hits<J, true, T>(bal, A, j0, jj, used, a);
      T p[J][K];
      eval_group<D, J, K, T, kSkip>(tile, jj, x, p);
#pragma unroll
      for (int q = 0; q < J; ++q)
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = fmaT(a[q], p[q][k], acc[k]);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) accd[k] += static_cast<double>(acc[k]);
  }
  double* out = part + static_cast<int64_t>(blockIdx.y) * ldx;
#pragma unroll
  for (int k = 0; k < K; ++k) out[m0 + k * 64] = accd[k];
}

// Generic-dimension mult (D > kMaxTemplD): one data point per lane,
// coordinates re-read from global memory (cached), runtime loops, no culling.
template <typename T>
__global__ __launch_bounds__(kBlock) void mult_kernel_generic(const T* __restrict__ X, int64_t ldx,
                                                              const T* __restrict__ LI,
                                                              const T* __restrict__ A,
                                                              double* __restrict__ part, int chunk,
                                                              int jend_all, int D) {
  const int64_t m = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  const int jb = blockIdx.y * chunk;
  const int je = min(jb + chunk, jend_all);
  double accd = 0.0;
  for (int j0 = jb; j0 < je; j0 += kTG) {
    T acc = T(0);
    for (int j = j0; j < j0 + kTG; ++j) {
      T p = T(1);
      for (int d = 0; d < D; ++d) {
        const int64_t e = li_at(j, d, D);
        p = hat_step(p, LI[e], X[d * ldx + m], LI[e + 1]);
      }
      acc = fmaT(A[j], p, acc);
    }
    accd += static_cast<double>(acc);
  }
  part[static_cast<int64_t>(blockIdx.y) * ldx + m] = accd;
}

// ------------------------------------------------------- multTranspose
// Same orientation as mult: lanes own data points (K per lane, coordinates
// and y in registers), grid tiles are culled against the wave's data box and
// the surviving points are evaluated J at a time. A point's contribution
// sum_k y_k phi(x_k) is summed over the wave's lanes (DPP) into lane jj of a
// per-tile register (a point is taken at most once per group and tile); the
// tile is then added, in FP64, to the wave's own row of LDS accumulators. At
// the end the four rows are added in a fixed order. Block: data segment [blockIdx.x*seg, +seg) x
// grid chunk [blockIdx.y*kChunkT, +kChunkT), the segment split into four
// contiguous wave shares. part: [S][ldp].
template <int D, int K, int J, typename T, bool kSkip>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(D > 10 ? 4 : 1)))
void multT_kernel(const T* __restrict__ X, int64_t ldx,
                                                       const T* __restrict__ Y,
                                                       const T* __restrict__ LI,
                                                       const float* __restrict__ TB,
                                                       double* __restrict__ part, int64_t ldp,
                                                       int jend_all, int64_t seg, int64_t mend) {
  __shared__ double accs[kBlock / 64][kChunkT];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int jb = blockIdx.y * kChunkT;
  const int je = min(jb + kChunkT, jend_all);
  for (int i = threadIdx.x; i < (kBlock / 64) * kChunkT; i += kBlock) (&accs[0][0])[i] = 0.0;
  __syncthreads();
  const int64_t mb = static_cast<int64_t>(blockIdx.x) * seg;
  const int64_t me = min(mb + seg, mend);
  // per-wave share, a multiple of 64*K; points past the share weigh 0
  const int64_t share = ((me - mb) / (kBlock / 64) + 64 * K - 1) / (64 * K) * (64 * K);
  const int64_t wb = mb + wave * share;
  const int64_t we = min(wb + share, me);
  double* acc = accs[wave];
  for (int64_t m0 = wb; m0 < we; m0 += 64 * K) {
    T x[K][D], y[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int64_t m = m0 + k * 64 + lane;
#pragma unroll
      for (int d = 0; d < D; ++d) x[k][d] = X[d * ldx + m];
      y[k] = (m < we) ? Y[m] : T(0);
    }
    WaveBox<D, T> box;
    wave_box<D, D, K, T>(x, box);
    for (int j0 = jb; j0 < je; j0 += kTG) {
      const int64_t t = j0 / kTG;
      if (!tile_may_hit<D, D, T>(TB, t, box)) continue;
      const T* tile = LI + t * D * kTG * 2;
      unsigned long long bal = tile_hits<D, D, T>(tile, lane, box);
      // a grid point is taken at most once per (group, tile): its sum goes
      // to lane jj of tsum, and the tile is added to the wave's row once
      T tsum = T(0);
      while (bal) {
        int jj[J];
        bool used[J];
        T unused_w[J];
        take_hits<J, false, T>(bal, static_cast<const T*>(nullptr), j0, jj, used, unused_w);
        T p[J][K];
        eval_group<D, J, K, T, kSkip>(tile, jj, x, p);
#pragma unroll
        for (int q = 0; q < J; ++q) {
          if (!used[q]) continue;
          T c = T(0);
#pragma unroll
          for (int k = 0; k < K; ++k) c = fmaT(y[k], p[q][k], c);
          const T sq = lane63(sum_to_lane63(c));
          tsum = lane == jj[q] ? sq : tsum;
        }
      }
      // one vector add per tile into the wave's own accumulator row (in
      // issue order: deterministic; untaken lanes add +0)
      atomicAdd(&acc[j0 - jb + lane], static_cast<double>(tsum));
    }
  }
  __syncthreads();
  double* out = part + static_cast<int64_t>(blockIdx.x) * ldp;
  for (int i = threadIdx.x; i < je - jb; i += kBlock) {
    double s = accs[0][i];
#pragma unroll
    for (int w = 1; w < kBlock / 64; ++w) s += accs[w][i];
    out[jb + i] = s;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void multT_kernel_generic(const T* __restrict__ X, int64_t ldx,
                                                               const T* __restrict__ Y,
                                                               const T* __restrict__ LI,
                                                               double* __restrict__ part,
                                                               int64_t ldp, int64_t seg,
                                                               int64_t mend, int D) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  const int64_t mb = static_cast<int64_t>(blockIdx.y) * seg;
  const int64_t me = min(mb + seg, mend);
  double accd = 0.0;
  for (int64_t m0 = mb; m0 < me; m0 += kTM) {
    T acc = T(0);
    for (int64_t m = m0; m < m0 + kTM; ++m) {
      T p = T(1);
      for (int d = 0; d < D; ++d) {
        const int64_t e = li_at(j, d, D);
        p = hat_step(p, LI[e], X[d * ldx + m], LI[e + 1]);
      }
      acc = fmaT(Y[m], p, acc);
    }
    accd += static_cast<double>(acc);
  }
  part[static_cast<int64_t>(blockIdx.y) * ldp + j] = accd;
}

// ------------------------------------------------------------- helpers
// A[s] = alpha[perm[s]] * scale[s] for stored position s (perm may be null)
template <typename T>
__global__ void prep_alpha_kernel(const double* __restrict__ alpha, const double* __restrict__ scale,
                                  const int32_t* __restrict__ perm, T* __restrict__ A, int64_t n,
                                  int64_t npad) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= npad) return;
  if (s < n) {
    const int64_t j = perm ? perm[s] : s;
    A[s] = static_cast<T>(alpha[j] * scale[s]);
  } else {
    A[s] = T(0);
  }
}

// Y[m] = y[dperm[m]] for stored position m (data in storage order), 0 past n
template <typename T>
__global__ void prep_y_kernel(const double* __restrict__ y, const int32_t* __restrict__ dperm,
                              T* __restrict__ Y, int64_t n, int64_t npad) {
  const int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (m < npad) Y[m] = (m < n) ? static_cast<T>(y[dperm[m]]) : T(0);
}

// out[perm ? perm[i] : i] = (scale ? scale[i] : 1) * sum_{p < nparts} part[p*ld + i],
// summed in the fixed order p = 0, 1, ... Each thread owns the pair i0, i0 + 1
// (16-byte loads; ld is even and the rows are padded, so i0 + 1 < ld) and
// issues eight rows' loads before their adds.
__global__ void reduce_parts_kernel(const double* __restrict__ part, int64_t ld, int nparts,
                                    const double* __restrict__ scale,
                                    const int32_t* __restrict__ perm, double* __restrict__ out,
                                    int64_t n) {
  const int64_t i0 = 2 * (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x);
  if (i0 >= n) return;
  const double2* col = reinterpret_cast<const double2*>(part + i0);
  const int64_t ld2 = ld / 2;
  double s0 = 0.0, s1 = 0.0;
  int p = 0;
  for (; p + 8 <= nparts; p += 8) {
    double2 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = col[(p + u) * ld2];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      s0 += v[u].x;
      s1 += v[u].y;
    }
  }
  for (; p < nparts; ++p) {
    const double2 v = col[p * ld2];
    s0 += v.x;
    s1 += v.y;
  }
  out[perm ? perm[i0] : i0] = scale ? s0 * scale[i0] : s0;
  if (i0 + 1 < n) out[perm ? perm[i0 + 1] : i0 + 1] = scale ? s1 * scale[i0 + 1] : s1;
}

// ------------------------------------------------------------ dispatch
// kSkip: test for constant dimensions (ModLinear); Linear grids have none
template <typename T, int D, bool kSkip>
void launch_mult_D(const T* X, int64_t ldx, const T* LI, const float* TB, const T* A,
                   double* part, int chunk, int jend, int G, int64_t Mpad,
                   hipStream_t s) {
  constexpr int K = mult_k<T>(D);
  dim3 grid(static_cast<unsigned>(Mpad / (kBlock * K)), static_cast<unsigned>(G));
  clear_launch_error();
  hipLaunchKernelGGL((mult_kernel<D, K, mult_j<T>(D), T, kSkip>), grid, dim3(kBlock), 0, s, X,
                     ldx, LI, TB, A, part, chunk, jend);
}

template <typename T, int D, bool kSkip>
void launch_multT_D(const T* X, int64_t ldx, const T* Y, const T* LI, const float* TB,
                    double* part, int64_t ldp, int jend, int64_t seg,
                    int64_t mend, int S, hipStream_t s) {
  constexpr int K = multT_k<T>(D);
  // data segments vary fastest: consecutive blocks share one grid chunk
  // (its rows stay hot in every XCD's caches) and read different data
  dim3 grid(static_cast<unsigned>(S), static_cast<unsigned>((jend + kChunkT - 1) / kChunkT));
  clear_launch_error();
  hipLaunchKernelGGL((multT_kernel<D, K, multT_j<T>(D), T, kSkip>), grid, dim3(kBlock), 0, s, X,
                     ldx, Y, LI, TB, part, ldp, jend, seg, mend);
}

template <typename T>
using MultFn = void (*)(const T*, int64_t, const T*, const float*, const T*, double*, int, int,
                        int, int64_t, hipStream_t);
template <typename T>
using MultTFn = void (*)(const T*, int64_t, const T*, const T*, const float*, double*, int64_t,
                         int, int64_t, int64_t, int, hipStream_t);

template <typename T, int... Ds>
struct Table {
  static MultFn<T> mult(int D, bool skip) {
    MultFn<T> f = nullptr;
    ((D == Ds ? (f = skip ? &launch_mult_D<T, Ds, true> : &launch_mult_D<T, Ds, false>, 0) : 0),
     ...);
    return f;
  }
  static MultTFn<T> multT(int D, bool skip) {
    MultTFn<T> f = nullptr;
    ((D == Ds ? (f = skip ? &launch_multT_D<T, Ds, true> : &launch_multT_D<T, Ds, false>, 0)
              : 0),
     ...);
    return f;
  }
};
template <typename T>
using DimTable = Table<T, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20>;

template <typename T>
int mult_k_rt(int D) {
  return D > kMaxTemplD ? 1 : mult_k<T>(D);
}
template <typename T>
int multT_k_rt(int D) {
  return D > kMaxTemplD ? 1 : multT_k<T>(D);
}

int ceil_div(int64_t a, int64_t b) { return static_cast<int>((a + b - 1) / b); }

}  // namespace

// Morton order of n points in [0,1]^d (row-major): b = min(16, 64/d) bits
// per coordinate, interleaved from the top bit; ties keep the input order.
std::vector<int32_t> morton_order(const double* pts, size_t n, size_t d) {
  const int b = static_cast<int>(std::max<size_t>(1, std::min<size_t>(16, 64 / d)));
  const size_t dd = std::min<size_t>(d, 64 / static_cast<size_t>(b));
  const double scale = static_cast<double>(1u << b);
  const uint32_t top = (1u << b) - 1u;
  std::vector<std::pair<uint64_t, int32_t>> key(n);
  std::vector<uint32_t> q(dd);
  for (size_t i = 0; i < n; ++i) {
    for (size_t k = 0; k < dd; ++k) {
      const double v = pts[i * d + k] * scale;
      q[k] = v <= 0.0 ? 0u : std::min<uint32_t>(top, static_cast<uint32_t>(v));
    }
    uint64_t c = 0;
    for (int bit = b - 1; bit >= 0; --bit)
      for (size_t k = 0; k < dd; ++k) c = (c << 1) | ((q[k] >> bit) & 1u);
    key[i] = {c, static_cast<int32_t>(i)};
  }
  std::sort(key.begin(), key.end());
  std::vector<int32_t> order(n);
  for (size_t i = 0; i < n; ++i) order[i] = key[i].second;
  return order;
}

// Grid storage order: by the point's set of non-constant dimensions, then by
// the Morton code of its support's centre (li: point-major (L, I) pairs).
std::vector<int32_t> grid_order(const std::vector<double>& li, const std::vector<uint32_t>& pmask,
                                size_t n, size_t d) {
  std::vector<double> centre(n * d);
  for (size_t j = 0; j < n; ++j)
    for (size_t k = 0; k < d; ++k) {
      const double L = li[(j * d + k) * 2], I = li[(j * d + k) * 2 + 1];
      const double c = L > 0.0 ? I / L : 0.5;
      centre[j * d + k] = std::min(1.0, std::max(0.0, c));
    }
  const std::vector<int32_t> mo = morton_order(centre.data(), n, d);
  std::vector<int32_t> rank(n);
  for (size_t i = 0; i < n; ++i) rank[mo[i]] = static_cast<int32_t>(i);
  std::vector<int32_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
    return pmask[a] != pmask[b] ? pmask[a] < pmask[b] : rank[a] < rank[b];
  });
  return order;
}

// Per 64-point tile and dimension the union of the points' supports
// [(I - 1) / L, (I + 1) / L] (whole line for L = 0), from point-major (L, I)
// pairs given as doubles, stored in T rounded outwards.
template <typename T>
std::vector<T> tile_boxes(const std::vector<double>& li, size_t npad, size_t d) {
  const size_t ntiles = npad / kTG;
  std::vector<T> tb(ntiles * d * 2);
  for (size_t t = 0; t < ntiles; ++t)
    for (size_t k = 0; k < d; ++k) {
      double lo = 2.0, hi = -1.0;
      for (size_t p = 0; p < static_cast<size_t>(kTG); ++p) {
        const size_t e = ((t * kTG + p) * d + k) * 2;
        const double L = li[e], I = li[e + 1];
        if (L > 0.0) {
          lo = std::min(lo, (I - 1.0) / L);
          hi = std::max(hi, (I + 1.0) / L);
        } else {
          lo = -1.0;
          hi = 2.0;
        }
      }
      T tl = static_cast<T>(lo), th = static_cast<T>(hi);
      if (static_cast<double>(tl) > lo) tl = std::nextafter(tl, static_cast<T>(-3));
      if (static_cast<double>(th) < hi) th = std::nextafter(th, static_cast<T>(3));
      tb[(t * d + k) * 2] = tl;
      tb[(t * d + k) * 2 + 1] = th;
    }
  return tb;
}

template std::vector<float> tile_boxes<float>(const std::vector<double>&, size_t, size_t);
template std::vector<double> tile_boxes<double>(const std::vector<double>&, size_t, size_t);

// =================================================================== runtime

void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::ostringstream os;
    os << what << ": " << hipGetErrorString(e) << " (" << static_cast<int>(e) << ")";
    throw device_error(os.str());
  }
}

void encode_grid(Basis basis, size_t dim, size_t n, const uint32_t* level, const uint32_t* index,
                 std::vector<double>& li, std::vector<double>& scale) {
  li.assign(n * dim * 2, 0.0);
  scale.assign(n, 1.0);
  for (size_t j = 0; j < n; ++j) {
    double s = 1.0;
    for (size_t d = 0; d < dim; ++d) {
      const uint32_t l = level[j * dim + d];
      const uint32_t i = index[j * dim + d];
      if (l < 1 || l > 30 || (i & 1u) == 0 || i >= (1u << l))
        throw input_error("invalid grid point (level/index out of range for a no-boundary grid)");
      const double h = static_cast<double>(1u << l);
      double L = h, I = static_cast<double>(i);
      if (basis == Basis::ModLinear) {
        if (l == 1) {
          L = 0.0;
          I = 0.0;
        } else if (i == 1) {
          L = 0.5 * h;
          I = 0.0;
          s *= 2.0;
        } else if (i == (1u << l) - 1) {
          L = 0.5 * h;
          I = 0.5 * h;
          s *= 2.0;
        }
      }
      li[(j * dim + d) * 2] = L;
      li[(j * dim + d) * 2 + 1] = I;
    }
    scale[j] = s;
  }
}

DeviceMultiEval::DeviceMultiEval(const double* data, size_t M, size_t dim, Basis basis,
                                 unsigned precisions, int device, const LaunchConfig& cfg)
    : M_(M), D_(dim), basis_(basis), precisions_(precisions), device_(device), cfg_(cfg) {
  if (M == 0 || dim == 0) throw input_error("empty dataset");
  if (precisions == 0 || (precisions & ~3u)) throw input_error("invalid precision mask");
  if (dim > 64) throw input_error("dimension above 64 is not supported");
  check_hip(hipSetDevice(device_), "hipSetDevice");
  hipDeviceProp_t prop;
  check_hip(hipGetDeviceProperties(&prop, device_), "hipGetDeviceProperties");
  num_cus_ = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  check_hip(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  check_hip(hipEventCreate(&ev0_), "hipEventCreate");
  check_hip(hipEventCreate(&ev1_), "hipEventCreate");

  for (size_t i = 0; i < M_ * D_; ++i)
    if (!(data[i] >= 0.0 && data[i] <= 1.0)) throw input_error("data points must lie in [0,1]^d");
  // at least kDataSlack rows past the data: a wave's last group of points
  // may start just below the end of the data
  Mpad_ = ((M_ + kDataSlack + kDataPad - 1) / kDataPad) * kDataPad;
  // storage order: Morton order of the points, so that consecutive points
  // (one wave's) fill small boxes; dperm[stored position] = caller's row
  std::vector<int32_t> dperm = morton_order(data, M_, D_);
  // pad by repeating the last stored row, transpose to [D][Mpad]
  std::vector<double> soa(D_ * Mpad_);
  for (size_t d = 0; d < D_; ++d) {
    double* col = soa.data() + d * Mpad_;
    for (size_t m = 0; m < M_; ++m) col[m] = data[static_cast<size_t>(dperm[m]) * D_ + d];
    const double last = col[M_ - 1];
    for (size_t m = M_; m < Mpad_; ++m) col[m] = last;
  }
  check_hip(hipMalloc(&dperm_, M_ * sizeof(int32_t)), "hipMalloc(data order)");
  check_hip(hipMemcpy(dperm_, dperm.data(), M_ * sizeof(int32_t), hipMemcpyHostToDevice),
            "upload data order");
  if (serves(Precision::FP64)) {
    check_hip(hipMalloc(&X64_, soa.size() * sizeof(double)), "hipMalloc(X64)");
    check_hip(hipMemcpy(X64_, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice),
              "upload X64");
  }
  if (serves(Precision::FP32)) {
    std::vector<float> s32(soa.begin(), soa.end());
    check_hip(hipMalloc(&X32_, s32.size() * sizeof(float)), "hipMalloc(X32)");
    check_hip(hipMemcpy(X32_, s32.data(), s32.size() * sizeof(float), hipMemcpyHostToDevice),
              "upload X32");
  }
}

DeviceMultiEval::~DeviceMultiEval() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  void* bufs[] = {X64_, X32_, LI64_, LI32_, TB64_, TB32_, scale_, perm_, dperm_, A_, Y_,
                  part_, hin_, hout_};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (auto* pool : {&tm_mult_, &tm_multT_})
    for (auto& e : *pool) {
      (void)hipEventDestroy(e.a);
      (void)hipEventDestroy(e.b);
    }
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void DeviceMultiEval::ensure(void** buf, size_t* cap, size_t bytes) {
  if (*cap >= bytes && *buf) return;
  if (*buf) {
    check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    check_hip(hipFree(*buf), "hipFree");
    *buf = nullptr;
    *cap = 0;
  }
  check_hip(hipMalloc(buf, bytes), "hipMalloc(scratch)");
  *cap = bytes;
}

void DeviceMultiEval::set_grid(const uint32_t* level, const uint32_t* index, size_t N) {
  if (N == 0) throw input_error("empty grid");
  if (N > (size_t(1) << 30)) throw input_error("grid too large");
  // non-constant dimensions of each point (ModLinear level-1 factors are 1)
  const bool mod = basis_ == Basis::ModLinear;
  std::vector<uint32_t> pmask(N);
  for (size_t j = 0; j < N; ++j) {
    uint32_t m = 0;
    for (size_t d = 0; d < D_ && d < 32; ++d)
      if (!mod || level[j * D_ + d] > 1) m |= 1u << d;
    pmask[j] = m;
  }
  std::vector<double> li0, sc0;
  encode_grid(basis_, D_, N, level, index, li0, sc0);
  // storage order: by the set of non-constant dimensions, then by the Morton
  // code of the support's centre, so that a 64-point tile holds points with
  // the same dimensions in a small region
  std::vector<int32_t> order = grid_order(li0, pmask, N, D_);
  std::vector<double> li(N * D_ * 2), sc(N);
  for (size_t s = 0; s < N; ++s) {
    const size_t j = static_cast<size_t>(order[s]);
    for (size_t e = 0; e < D_ * 2; ++e) li[s * D_ * 2 + e] = li0[j * D_ * 2 + e];
    sc[s] = sc0[j];
  }
  const size_t Npad = ((N + kGridPad - 1) / kGridPad) * kGridPad;
  // padded points: (L, I) = (0, 0) -> factor 1, weight 0 through A / ignored rows
  li.resize(Npad * D_ * 2, 0.0);
  sc.resize(Npad, 1.0);
  check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  N_ = N;
  Npad_ = Npad;
  ensure(reinterpret_cast<void**>(&scale_), &gridcap_scale_, Npad_ * sizeof(double));
  check_hip(hipMemcpy(scale_, sc.data(), Npad_ * sizeof(double), hipMemcpyHostToDevice),
            "upload scale");
  ensure(&perm_, &perm_cap_, N_ * sizeof(int32_t));
  check_hip(hipMemcpy(perm_, order.data(), N_ * sizeof(int32_t), hipMemcpyHostToDevice),
            "upload permutation");
  if (serves(Precision::FP64)) {
    ensure(&LI64_, &gridcap_li64_, li.size() * sizeof(double));
    check_hip(hipMemcpy(LI64_, li.data(), li.size() * sizeof(double), hipMemcpyHostToDevice),
              "upload LI64");
    // FP32 boxes of the FP64 supports, rounded outwards
    const std::vector<float> tb = tile_boxes<float>(li, Npad_, D_);
    ensure(&TB64_, &tbcap64_, tb.size() * sizeof(float));
    check_hip(hipMemcpy(TB64_, tb.data(), tb.size() * sizeof(float), hipMemcpyHostToDevice),
              "upload tile boxes");
  }
  if (serves(Precision::FP32)) {
    std::vector<float> l32(li.begin(), li.end());
    ensure(&LI32_, &gridcap_li32_, l32.size() * sizeof(float));
    check_hip(hipMemcpy(LI32_, l32.data(), l32.size() * sizeof(float), hipMemcpyHostToDevice),
              "upload LI32");
    // boxes of the FP32 (L, I) the FP32 kernels evaluate
    const std::vector<double> l32d(l32.begin(), l32.end());
    const std::vector<float> tb = tile_boxes<float>(l32d, Npad_, D_);
    ensure(&TB32_, &tbcap32_, tb.size() * sizeof(float));
    check_hip(hipMemcpy(TB32_, tb.data(), tb.size() * sizeof(float), hipMemcpyHostToDevice),
              "upload tile boxes");
  }
}

void DeviceMultiEval::record_start() { check_hip(hipEventRecord(ev0_, stream_), "hipEventRecord"); }
void DeviceMultiEval::record_stop() { check_hip(hipEventRecord(ev1_, stream_), "hipEventRecord"); }

template <typename T>
void DeviceMultiEval::run_mult(const T* X, const T* LI, const float* TB, const double* d_alpha,
                               double* d_result, hipEvent_t* tev) {
  hipStream_t s = stream_;
  ensure(&A_, &A_cap_, Npad_ * sizeof(T));
  clear_launch_error();
  hipLaunchKernelGGL(prep_alpha_kernel<T>, dim3(ceil_div(Npad_, 256)), dim3(256), 0, s, d_alpha,
                     scale_, static_cast<const int32_t*>(perm_), static_cast<T*>(A_),
                     static_cast<int64_t>(N_), static_cast<int64_t>(Npad_));
  check_hip(hipGetLastError(), "prep_alpha");
  const int K = mult_k_rt<T>(static_cast<int>(D_));
  const int ntiles = ceil_div(N_, kTG);
  const int bx = static_cast<int>(Mpad_ / (kBlock * K));
  int G = cfg_.mult_grid_chunks;
  if (G <= 0) {
    // many short blocks: the dispatcher then balances the uneven per-block
    // work (culling makes blocks cost very different amounts) across all CUs
    const int target = cfg_.target_blocks > 0 ? cfg_.target_blocks : kBlocksPerCU * num_cus_;
    G = std::max(1, ceil_div(target, bx));
    G = std::max(G, ceil_div(ntiles, kMaxChunkTiles));
    G = std::min(G, std::max(1, ntiles / 4));  // at least 4 tiles per chunk
    // partials G x Mpad doubles within the budget
    G = std::min<int64_t>(G, std::max<int64_t>(1, kPartialBudget / (Mpad_ * sizeof(double))));
  }
  G = std::max(1, std::min(G, ntiles));
  const int tiles_per = ceil_div(ntiles, G);
  G = ceil_div(ntiles, tiles_per);
  const int chunk = tiles_per * kTG;
  const int jend = ntiles * kTG;
  ensure(&part_, &part_cap_, static_cast<size_t>(G) * Mpad_ * sizeof(double));
  double* P = static_cast<double*>(part_);
  if (tev) check_hip(hipEventRecord(tev[0], s), "hipEventRecord");
  if (D_ <= static_cast<size_t>(kMaxTemplD)) {
    MultFn<T> f = DimTable<T>::mult(static_cast<int>(D_), basis_ == Basis::ModLinear);
    f(X, static_cast<int64_t>(Mpad_), LI, TB, static_cast<const T*>(A_),
      P, chunk, jend, G, static_cast<int64_t>(Mpad_), s);
  } else {
    dim3 grid(static_cast<unsigned>(Mpad_ / kBlock), static_cast<unsigned>(G));
    clear_launch_error();
    hipLaunchKernelGGL(mult_kernel_generic<T>, grid, dim3(kBlock), 0, s, X,
                       static_cast<int64_t>(Mpad_), LI, static_cast<const T*>(A_), P, chunk, jend,
                       static_cast<int>(D_));
  }
  check_hip(hipGetLastError(), "mult kernel");
  if (tev) check_hip(hipEventRecord(tev[1], s), "hipEventRecord");
  // results back to the caller's row order
  clear_launch_error();
  hipLaunchKernelGGL(reduce_parts_kernel, dim3(ceil_div(M_, 512)), dim3(256), 0, s, P,
                     static_cast<int64_t>(Mpad_), G, static_cast<const double*>(nullptr),
                     static_cast<const int32_t*>(dperm_), d_result, static_cast<int64_t>(M_));
  check_hip(hipGetLastError(), "reduce (mult)");
  last_G_ = G;
}

template <typename T>
void DeviceMultiEval::run_multT(const T* X, const T* LI, const float* TB, const double* d_y,
                                double* d_result, hipEvent_t* tev) {
  hipStream_t s = stream_;
  ensure(&Y_, &Y_cap_, Mpad_ * sizeof(T));
  clear_launch_error();
  hipLaunchKernelGGL(prep_y_kernel<T>, dim3(ceil_div(Mpad_, 256)), dim3(256), 0, s, d_y,
                     static_cast<const int32_t*>(dperm_), static_cast<T*>(Y_),
                     static_cast<int64_t>(M_), static_cast<int64_t>(Mpad_));
  check_hip(hipGetLastError(), "prep_y");
  const bool templ = D_ <= static_cast<size_t>(kMaxTemplD);
  const int ntiles = ceil_div(N_, kTG);
  const int jend = ntiles * kTG;
  // grid blocks: chunks of kChunkT points (templated kernel) or kBlock points
  const int bx = templ ? ceil_div(jend, kChunkT) : ceil_div(N_, kBlock);
  // Data are split into groups (templated kernel: one wave's 64*K points;
  // generic kernel: kTM points); a segment is 4 waves x gpw groups, so the
  // four waves of a block get equal shares. mend = whole groups past M (the
  // padding holds at least kDataSlack rows past M).
  const int64_t group = templ ? 64 * multT_k_rt<T>(static_cast<int>(D_)) : kTM;
  const int64_t ngroups = (static_cast<int64_t>(M_) + group - 1) / group;
  const int64_t mend = ngroups * group;
  const int64_t waves = templ ? kBlock / 64 : 1;
  int S = cfg_.multT_data_segs;
  if (S <= 0) {
    const int target = cfg_.target_blocks > 0 ? cfg_.target_blocks
                                              : (templ ? kBlocksPerCUT : kBlocksPerCU) * num_cus_;
    S = std::max(1, ceil_div(target, bx));
    S = std::min<int64_t>(S, std::max<int64_t>(1, kPartialBudget / (Npad_ * sizeof(double))));
  }
  const int64_t gpw = std::max<int64_t>(1, (ngroups + waves * S - 1) / (waves * S));
  const int64_t seg = waves * gpw * group;
  S = ceil_div(mend, seg);
  ensure(&part_, &part_cap_, static_cast<size_t>(S) * Npad_ * sizeof(double));
  double* P = static_cast<double*>(part_);
  if (tev) check_hip(hipEventRecord(tev[0], s), "hipEventRecord");
  if (templ) {
    MultTFn<T> f = DimTable<T>::multT(static_cast<int>(D_), basis_ == Basis::ModLinear);
    f(X, static_cast<int64_t>(Mpad_), static_cast<const T*>(Y_), LI, TB,
      P, static_cast<int64_t>(Npad_), jend, seg, mend, S,
      s);
  } else {
    dim3 grid(static_cast<unsigned>(bx), static_cast<unsigned>(S));
    clear_launch_error();
    hipLaunchKernelGGL(multT_kernel_generic<T>, grid, dim3(kBlock), 0, s, X,
                       static_cast<int64_t>(Mpad_), static_cast<const T*>(Y_), LI, P,
                       static_cast<int64_t>(Npad_), seg, mend, static_cast<int>(D_));
  }
  check_hip(hipGetLastError(), "multTranspose kernel");
  if (tev) check_hip(hipEventRecord(tev[1], s), "hipEventRecord");
  // results back to the caller's grid order, times the per-point scale
  clear_launch_error();
  hipLaunchKernelGGL(reduce_parts_kernel, dim3(ceil_div(N_, 512)), dim3(256), 0, s, P,
                     static_cast<int64_t>(Npad_), S, scale_, static_cast<const int32_t*>(perm_),
                     d_result, static_cast<int64_t>(N_));
  check_hip(hipGetLastError(), "reduce (multTranspose)");
  last_S_ = S;
}

void DeviceMultiEval::mult_device(const double* d_alpha, double* d_result, Precision p) {
  if (N_ == 0) throw input_error("grid not set (call prepare)");
  if (!serves(p)) throw input_error("precision not enabled for this instance");
  check_hip(hipSetDevice(device_), "hipSetDevice");
  EventPair* tp = timing_on_ ? next_timer(tm_mult_, used_mult_) : nullptr;
  hipEvent_t tev[2] = {tp ? tp->a : nullptr, tp ? tp->b : nullptr};
  record_start();
  if (p == Precision::FP64)
    run_mult<double>(static_cast<const double*>(X64_), static_cast<const double*>(LI64_),
                     static_cast<const float*>(TB64_), d_alpha, d_result, tp ? tev : nullptr);
  else
    run_mult<float>(static_cast<const float*>(X32_), static_cast<const float*>(LI32_),
                    static_cast<const float*>(TB32_), d_alpha, d_result, tp ? tev : nullptr);
  record_stop();
}

void DeviceMultiEval::mult_transpose_device(const double* d_y, double* d_result, Precision p) {
  if (N_ == 0) throw input_error("grid not set (call prepare)");
  if (!serves(p)) throw input_error("precision not enabled for this instance");
  check_hip(hipSetDevice(device_), "hipSetDevice");
  EventPair* tp = timing_on_ ? next_timer(tm_multT_, used_multT_) : nullptr;
  hipEvent_t tev[2] = {tp ? tp->a : nullptr, tp ? tp->b : nullptr};
  record_start();
  if (p == Precision::FP64)
    run_multT<double>(static_cast<const double*>(X64_), static_cast<const double*>(LI64_),
                      static_cast<const float*>(TB64_), d_y, d_result, tp ? tev : nullptr);
  else
    run_multT<float>(static_cast<const float*>(X32_), static_cast<const float*>(LI32_),
                     static_cast<const float*>(TB32_), d_y, d_result, tp ? tev : nullptr);
  record_stop();
}

DeviceMultiEval::EventPair* DeviceMultiEval::next_timer(std::vector<EventPair>& pool,
                                                       size_t& used) {
  if (used == pool.size()) {
    EventPair e;
    check_hip(hipEventCreate(&e.a), "hipEventCreate");
    check_hip(hipEventCreate(&e.b), "hipEventCreate");
    pool.push_back(e);
  }
  return &pool[used++];
}

void DeviceMultiEval::set_kernel_timing(bool on) {
  timing_on_ = on;
  used_mult_ = used_multT_ = 0;
}

void DeviceMultiEval::kernel_timing(int* n_mult, double* ms_mult, int* n_multT,
                                    double* ms_multT) {
  check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  auto sum = [](std::vector<EventPair>& pool, size_t used) {
    double t = 0.0;
    for (size_t i = 0; i < used; ++i) {
      float ms = 0.0f;
      check_hip(hipEventElapsedTime(&ms, pool[i].a, pool[i].b), "hipEventElapsedTime");
      t += ms;
    }
    return t;
  };
  if (n_mult) *n_mult = static_cast<int>(used_mult_);
  if (ms_mult) *ms_mult = sum(tm_mult_, used_mult_);
  if (n_multT) *n_multT = static_cast<int>(used_multT_);
  if (ms_multT) *ms_multT = sum(tm_multT_, used_multT_);
}

void DeviceMultiEval::synchronize() {
  check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, ev0_, ev1_) == hipSuccess) last_kernel_ms_ = ms;
}

void DeviceMultiEval::mult_host(const double* alpha, double* result, Precision p) {
  const auto t0 = std::chrono::steady_clock::now();
  ensure(&hin_, &hin_cap_, N_ * sizeof(double));
  ensure(&hout_, &hout_cap_, M_ * sizeof(double));
  check_hip(hipMemcpyAsync(hin_, alpha, N_ * sizeof(double), hipMemcpyHostToDevice, stream_),
            "copy alpha");
  mult_device(static_cast<const double*>(hin_), static_cast<double*>(hout_), p);
  check_hip(hipMemcpyAsync(result, hout_, M_ * sizeof(double), hipMemcpyDeviceToHost, stream_),
            "copy result");
  synchronize();
  last_wall_s_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void DeviceMultiEval::mult_transpose_host(const double* y, double* result, Precision p) {
  const auto t0 = std::chrono::steady_clock::now();
  ensure(&hin_, &hin_cap_, M_ * sizeof(double));
  ensure(&hout_, &hout_cap_, N_ * sizeof(double));
  check_hip(hipMemcpyAsync(hin_, y, M_ * sizeof(double), hipMemcpyHostToDevice, stream_),
            "copy source");
  mult_transpose_device(static_cast<const double*>(hin_), static_cast<double*>(hout_), p);
  check_hip(hipMemcpyAsync(result, hout_, N_ * sizeof(double), hipMemcpyDeviceToHost, stream_),
            "copy result");
  synchronize();
  last_wall_s_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace sgpp_hip
