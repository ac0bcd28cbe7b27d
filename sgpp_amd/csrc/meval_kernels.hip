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
// mult (r = B alpha): lanes own data points (K per lane, coordinates held in
//   registers for the whole launch); grid tiles of 64 points are staged in
//   LDS and read as wave-uniform broadcasts. Work is split over data tiles
//   (blockIdx.x) and G grid chunks (blockIdx.y); the G FP64 partial rows are
//   summed in a fixed order by reduce_parts.
// multTranspose (a = B^T y): lanes own grid points (K per lane, L/I in
//   registers); data tiles of 128 points (coordinates + y) are staged in LDS
//   and broadcast. Work is split over grid tiles (blockIdx.x) and S data
//   segments (blockIdx.y); the S FP64 partial rows are summed in a fixed
//   order. No atomics are used anywhere: results are deterministic.
// Constant factors: a ModLinear level-1 factor is exactly 1. For ModLinear
//   grids the points are kept sorted by their set of non-constant dimensions
//   (an internal permutation, undone on input and output), every 64-point
//   tile carries the union of its points' sets, and mult skips the other
//   dimensions with a wave-uniform scalar branch. Skipping a factor of
//   exactly 1 leaves every result bit-identical.
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

#include "meval_device.hpp"

namespace sgpp_hip {

namespace {

constexpr int kBlock = 256;  // 4 waves of 64
constexpr int kTG = 64;      // grid points per LDS tile (mult) and per mask tile
constexpr int kTM = 128;     // data points per LDS tile (multTranspose)
constexpr int kMaxTemplD = 20;
// default launch size: workgroups per CU and launch (measured optimum region
// 40-300 on the d=10 grids; below ~16 the last wave of blocks leaves CUs idle)
constexpr int kBlocksPerCU = 64;

__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float absT(float a) { return __builtin_fabsf(a); }
__device__ __forceinline__ double absT(double a) { return __builtin_fabs(a); }
__device__ __forceinline__ float clamp01(float v) { return __builtin_amdgcn_fmed3f(v, 0.0f, 1.0f); }
__device__ __forceinline__ double clamp01(double v) { return __builtin_fmin(__builtin_fmax(v, 0.0), 1.0); }

// running product: p * max(0, 1 - |u|) for p in [0,1]
template <typename T>
__device__ __forceinline__ T hat_step(T p, T L, T x, T I) {
  const T u = fmaT(L, x, -I);
  return clamp01(fmaT(-p, absT(u), p));
}

// Points per lane (K) and points per sub-tile (J). One wave-uniform branch
// per (sub-tile, dimension) decides whether a dimension is evaluated, so the
// branch cost is shared by J*K factor updates per lane. K keeps the per-lane
// operands (coordinates for mult, (L, I) for multTranspose) within ~40 FP32
// or ~80 FP64 registers. Values from launch-shape and variant sweeps on the
// d=10 ModLinear headline grid.
template <typename T>
constexpr int mult_k(int D) {
  return sizeof(T) == 8 ? (D <= 10 ? 4 : 2) : (D <= 5 ? 8 : (D <= 10 ? 4 : 2));
}
template <typename T>
constexpr int mult_j(int D) {
  return sizeof(T) == 8 ? 4 : (mult_k<T>(D) == 8 ? 4 : 8);
}
// multTranspose: with the dimension skip one tile per wave (K = 1) keeps the
// wave's mask as small as the tile's; without it 2-4 points per lane.
template <typename T>
constexpr int multT_k(int D, bool skip) {
  return skip ? 1 : (sizeof(T) == 8 ? (D <= 10 ? 2 : 1) : (D <= 5 ? 4 : 2));
}
template <typename T>
constexpr int multT_j(int D, bool skip) {
  return skip ? 8 : (sizeof(T) == 8 ? (multT_k<T>(D, false) == 1 ? 8 : 4) : (D <= 10 ? 4 : 8));
}

// (L, I) of a 64-point tile t and dimension d: LI[((t*D + d)*64 + jj)*2 + {0,1}]
// (tile-major, then dimension, then point), so the J points of a sub-tile are
// contiguous for every dimension and a wave's 64 points are contiguous too.
__device__ __forceinline__ int64_t li_at(int64_t j, int d, int D) {
  return (((j >> 6) * D + d) * 64 + (j & 63)) * 2;
}

// ---------------------------------------------------------------- mult
// X: [D][ldx] SoA, LI: tile-major pairs (li_at), A: [Npad] (alpha*scale, zero
// padded), qmask: [Npad/4] non-constant dimensions of every 4 consecutive
// points (kSkip only; a sub-tile ORs its J/4 masks), part: [G][ldx]. Block
// covers kBlock*K data points and grid chunk [blockIdx.y*chunk, min(+chunk, jend_all)).
template <int D, int K, int J, typename T, bool kSkip>
__global__ __launch_bounds__(kBlock) void mult_kernel(const T* __restrict__ X, int64_t ldx,
                                                      const T* __restrict__ LI,
                                                      const T* __restrict__ A,
                                                      const uint32_t* __restrict__ qmask,
                                                      double* __restrict__ part, int chunk,
                                                      int jend_all) {
  static_assert(kTG % J == 0 && J % 4 == 0, "sub-tile must divide the tile, 4 must divide it");
  __shared__ T sLI[kTG * D * 2];
  __shared__ T sA[kTG];
  __shared__ uint32_t sM[kTG / 4];
  const int64_t m0 = static_cast<int64_t>(blockIdx.x) * (kBlock * K) + threadIdx.x;
  T x[K][D];
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int d = 0; d < D; ++d) x[k][d] = X[d * ldx + m0 + k * kBlock];

  double accd[K];
#pragma unroll
  for (int k = 0; k < K; ++k) accd[k] = 0.0;

  const int jb = blockIdx.y * chunk;
  const int je = min(jb + chunk, jend_all);
  for (int j0 = jb; j0 < je; j0 += kTG) {
    __syncthreads();
    for (int e = threadIdx.x; e < kTG * D * 2; e += kBlock)
      sLI[e] = LI[static_cast<int64_t>(j0) * D * 2 + e];
    if (threadIdx.x < kTG) sA[threadIdx.x] = A[j0 + threadIdx.x];
    if (kSkip && threadIdx.x < kTG / 4) sM[threadIdx.x] = qmask[j0 / 4 + threadIdx.x];
    __syncthreads();
    T acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = T(0);
    for (int j = 0; j < kTG; j += J) {
      uint32_t mask = 0xffffffffu;
      if (kSkip) {
        mask = 0u;
#pragma unroll
        for (int c = 0; c < J / 4; ++c) mask |= sM[j / 4 + c];
        mask = __builtin_amdgcn_readfirstlane(mask);
      }
      T p[J][K];
#pragma unroll
      for (int q = 0; q < J; ++q)
#pragma unroll
        for (int k = 0; k < K; ++k) p[q][k] = T(1);
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (kSkip && !((mask >> d) & 1u)) continue;
        // (L, I) of the sub-tile's J points in dimension d: one batch of LDS reads
        T li[2 * J];
#pragma unroll
        for (int c = 0; c < 2 * J; ++c) li[c] = sLI[(d * kTG + j) * 2 + c];
#pragma unroll
        for (int q = 0; q < J; ++q)
#pragma unroll
          for (int k = 0; k < K; ++k) p[q][k] = hat_step(p[q][k], li[2 * q], x[k][d], li[2 * q + 1]);
      }
#pragma unroll
      for (int q = 0; q < J; ++q) {
        const T a = sA[j + q];
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = fmaT(a, p[q][k], acc[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) accd[k] += static_cast<double>(acc[k]);
  }
  double* out = part + static_cast<int64_t>(blockIdx.y) * ldx;
#pragma unroll
  for (int k = 0; k < K; ++k) out[m0 + k * kBlock] = accd[k];
}

// Generic-dimension mult (D > kMaxTemplD): one data point per lane,
// coordinates re-read from global memory (cached), runtime loops.
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
// X: [D][ldx] SoA, Y: [ldx] (zero padded), LI: tile-major pairs (li_at),
// part: [S][ldp]. Block covers grid points [blockIdx.x*kBlock*K, +kBlock*K)
// and data segment [blockIdx.y*seg, min(+seg, mend)). The data tile sits in
// LDS as [D+1][kTM] (coordinates, then y). With kSkip each wave evaluates
// only the union of the non-constant dimensions of its K tiles.
template <int D, int K, int J, typename T, bool kSkip>
__global__ __launch_bounds__(kBlock) void multT_kernel(const T* __restrict__ X, int64_t ldx,
                                                       const T* __restrict__ Y,
                                                       const T* __restrict__ LI,
                                                       const uint32_t* __restrict__ tmask,
                                                       double* __restrict__ part, int64_t ldp,
                                                       int64_t seg, int64_t mend) {
  static_assert(kTM % J == 0, "sub-tile must divide the tile");
  __shared__ T sX[(D + 1) * kTM];
  const int64_t jblock = static_cast<int64_t>(blockIdx.x) * (kBlock * K);
  const int64_t j0 = jblock + threadIdx.x;
  uint32_t mask = 0xffffffffu;
  if (kSkip) {
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
    mask = 0u;
#pragma unroll
    for (int k = 0; k < K; ++k) mask |= tmask[(jblock + wave * 64 + k * kBlock) / kTG];
  }
  T L[K][D], I[K][D];
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int64_t e = li_at(j0 + k * kBlock, d, D);
      L[k][d] = LI[e];
      I[k][d] = LI[e + 1];
    }
  double accd[K];
#pragma unroll
  for (int k = 0; k < K; ++k) accd[k] = 0.0;

  const int64_t mb = static_cast<int64_t>(blockIdx.y) * seg;
  const int64_t me = min(mb + seg, mend);
  for (int64_t m0 = mb; m0 < me; m0 += kTM) {
    __syncthreads();
    for (int e = threadIdx.x; e < kTM * (D + 1); e += kBlock) {
      const int d = e / kTM;
      const int m = e - d * kTM;
      sX[e] = (d < D) ? X[d * ldx + m0 + m] : Y[m0 + m];
    }
    __syncthreads();
    T acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = T(0);
    for (int m = 0; m < kTM; m += J) {
      T p[J][K];
#pragma unroll
      for (int q = 0; q < J; ++q)
#pragma unroll
        for (int k = 0; k < K; ++k) p[q][k] = T(1);
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (kSkip && !((mask >> d) & 1u)) continue;
        T xv[J];
#pragma unroll
        for (int q = 0; q < J; ++q) xv[q] = sX[d * kTM + m + q];
#pragma unroll
        for (int q = 0; q < J; ++q)
#pragma unroll
          for (int k = 0; k < K; ++k) p[q][k] = hat_step(p[q][k], L[k][d], xv[q], I[k][d]);
      }
#pragma unroll
      for (int q = 0; q < J; ++q) {
        const T y = sX[D * kTM + m + q];
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = fmaT(y, p[q][k], acc[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) accd[k] += static_cast<double>(acc[k]);
  }
  double* out = part + static_cast<int64_t>(blockIdx.y) * ldp;
#pragma unroll
  for (int k = 0; k < K; ++k) out[j0 + k * kBlock] = accd[k];
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

template <typename T>
__global__ void prep_y_kernel(const double* __restrict__ y, T* __restrict__ Y, int64_t n,
                              int64_t npad) {
  const int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (m < npad) Y[m] = (m < n) ? static_cast<T>(y[m]) : T(0);
}

// out[perm ? perm[i] : i] = (scale ? scale[i] : 1) * sum_{p < nparts} part[p*ld + i],
// summed in a fixed order
__global__ void reduce_parts_kernel(const double* __restrict__ part, int64_t ld, int nparts,
                                    const double* __restrict__ scale,
                                    const int32_t* __restrict__ perm, double* __restrict__ out,
                                    int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int p = 0; p < nparts; ++p) s += part[p * ld + i];
  out[perm ? perm[i] : i] = scale ? s * scale[i] : s;
}

// ------------------------------------------------------------ dispatch
template <typename T, int D>
void launch_mult_D(bool skip, const T* X, int64_t ldx, const T* LI, const T* A,
                   const uint32_t* tm, double* part, int chunk, int jend, int G, int64_t Mpad,
                   hipStream_t s) {
  constexpr int K = mult_k<T>(D);
  constexpr int J = mult_j<T>(D);
  dim3 grid(static_cast<unsigned>(Mpad / (kBlock * K)), static_cast<unsigned>(G));
  clear_launch_error();
  if constexpr (D >= 2) {
    if (skip) {
      hipLaunchKernelGGL((mult_kernel<D, K, J, T, true>), grid, dim3(kBlock), 0, s, X, ldx, LI, A,
                         tm, part, chunk, jend);
      return;
    }
  }
  hipLaunchKernelGGL((mult_kernel<D, K, J, T, false>), grid, dim3(kBlock), 0, s, X, ldx, LI, A,
                     tm, part, chunk, jend);
}

template <typename T, int D>
void launch_multT_D(bool skip, const T* X, int64_t ldx, const T* Y, const T* LI,
                    const uint32_t* tm, double* part, int64_t ldp, int64_t seg, int64_t mend,
                    int S, int64_t N, hipStream_t s) {
  clear_launch_error();
  // one dimension leaves nothing to skip (and its skip variant compiles badly)
  if constexpr (D >= 2) {
    if (skip) {
      constexpr int K = multT_k<T>(D, true);
      const int64_t cover = static_cast<int64_t>(kBlock) * K;
      dim3 grid(static_cast<unsigned>((N + cover - 1) / cover), static_cast<unsigned>(S));
      hipLaunchKernelGGL((multT_kernel<D, K, multT_j<T>(D, true), T, true>), grid, dim3(kBlock), 0, s,
                         X, ldx, Y, LI, tm, part, ldp, seg, mend);
      return;
    }
  }
  constexpr int K = multT_k<T>(D, false);
  const int64_t cover = static_cast<int64_t>(kBlock) * K;
  dim3 grid(static_cast<unsigned>((N + cover - 1) / cover), static_cast<unsigned>(S));
  hipLaunchKernelGGL((multT_kernel<D, K, multT_j<T>(D, false), T, false>), grid, dim3(kBlock), 0, s, X,
                     ldx, Y, LI, tm, part, ldp, seg, mend);
}

template <typename T>
using MultFn = void (*)(bool, const T*, int64_t, const T*, const T*, const uint32_t*, double*, int,
                        int, int, int64_t, hipStream_t);
template <typename T>
using MultTFn = void (*)(bool, const T*, int64_t, const T*, const T*, const uint32_t*, double*,
                         int64_t, int64_t, int64_t, int, int64_t, hipStream_t);

template <typename T, int... Ds>
struct Table {
  static MultFn<T> mult(int D) {
    MultFn<T> f = nullptr;
    ((D == Ds ? (f = &launch_mult_D<T, Ds>, 0) : 0), ...);
    return f;
  }
  static MultTFn<T> multT(int D) {
    MultTFn<T> f = nullptr;
    ((D == Ds ? (f = &launch_multT_D<T, Ds>, 0) : 0), ...);
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
int multT_k_rt(int D, bool skip) {
  return D > kMaxTemplD ? 1 : multT_k<T>(D, skip && D >= 2);
}

int ceil_div(int64_t a, int64_t b) { return static_cast<int>((a + b - 1) / b); }

}  // namespace

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

  Mpad_ = ((M_ + kDataPad - 1) / kDataPad) * kDataPad;
  // pad by repeating the last row, transpose to [D][Mpad]
  std::vector<double> soa(D_ * Mpad_);
  for (size_t d = 0; d < D_; ++d) {
    double* col = soa.data() + d * Mpad_;
    for (size_t m = 0; m < M_; ++m) col[m] = data[m * D_ + d];
    const double last = data[(M_ - 1) * D_ + d];
    for (size_t m = M_; m < Mpad_; ++m) col[m] = last;
  }
  for (double v : soa)
    if (!(v >= 0.0 && v <= 1.0)) throw input_error("data points must lie in [0,1]^d");
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
  void* bufs[] = {X64_, X32_, LI64_, LI32_, scale_, perm_, tmask_, qmask_, A_, Y_, part_, hin_,
                  hout_};
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
  // non-constant dimensions of each point; ModLinear level-1 factors are 1
  std::vector<uint32_t> pmask(N);
  const uint32_t full = D_ >= 32 ? 0xffffffffu : ((1u << D_) - 1u);
  const bool skip = basis_ == Basis::ModLinear && D_ <= static_cast<size_t>(kMaxTemplD);
  for (size_t j = 0; j < N; ++j) {
    uint32_t m = 0;
    for (size_t d = 0; d < D_; ++d)
      if (!skip || level[j * D_ + d] > 1) m |= 1u << d;
    pmask[j] = skip ? m : full;
  }
  // storage order: sorted by the set of non-constant dimensions (stable)
  std::vector<int32_t> order(N);
  std::iota(order.begin(), order.end(), 0);
  if (skip)
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return pmask[a] < pmask[b]; });
  std::vector<uint32_t> lv(N * D_), ix(N * D_);
  for (size_t s = 0; s < N; ++s)
    for (size_t d = 0; d < D_; ++d) {
      lv[s * D_ + d] = level[order[s] * D_ + d];
      ix[s * D_ + d] = index[order[s] * D_ + d];
    }
  std::vector<double> li, sc;
  encode_grid(basis_, D_, N, lv.data(), ix.data(), li, sc);
  const size_t Npad = ((N + kGridPad - 1) / kGridPad) * kGridPad;
  // padded points: (L, I) = (0, 0) -> factor 1, weight 0 through A / ignored rows
  li.resize(Npad * D_ * 2, 0.0);
  sc.resize(Npad, 1.0);
  // point-major [Npad][D][2] -> tile-major [Npad/64][D][64][2]
  {
    std::vector<double> t(li.size());
    for (size_t j = 0; j < Npad; ++j)
      for (size_t d = 0; d < D_; ++d) {
        const size_t dst = (((j / kTG) * D_ + d) * kTG + (j % kTG)) * 2;
        t[dst] = li[(j * D_ + d) * 2];
        t[dst + 1] = li[(j * D_ + d) * 2 + 1];
      }
    li.swap(t);
  }
  std::vector<uint32_t> tm(Npad / kTG, 0u);
  for (size_t s = 0; s < N; ++s) tm[s / kTG] |= pmask[order[s]];
  check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  N_ = N;
  Npad_ = Npad;
  skip_ = skip;
  ensure(reinterpret_cast<void**>(&scale_), &gridcap_scale_, Npad_ * sizeof(double));
  check_hip(hipMemcpy(scale_, sc.data(), Npad_ * sizeof(double), hipMemcpyHostToDevice),
            "upload scale");
  ensure(&tmask_, &tmask_cap_, tm.size() * sizeof(uint32_t));
  check_hip(hipMemcpy(tmask_, tm.data(), tm.size() * sizeof(uint32_t), hipMemcpyHostToDevice),
            "upload tile masks");
  std::vector<uint32_t> qm(Npad / 4, 0u);
  for (size_t s = 0; s < N; ++s) qm[s / 4] |= pmask[order[s]];
  ensure(&qmask_, &qmask_cap_, qm.size() * sizeof(uint32_t));
  check_hip(hipMemcpy(qmask_, qm.data(), qm.size() * sizeof(uint32_t), hipMemcpyHostToDevice),
            "upload sub-tile masks");
  if (skip_) {
    ensure(&perm_, &perm_cap_, N_ * sizeof(int32_t));
    check_hip(hipMemcpy(perm_, order.data(), N_ * sizeof(int32_t), hipMemcpyHostToDevice),
              "upload permutation");
  }
  if (serves(Precision::FP64)) {
    ensure(&LI64_, &gridcap_li64_, li.size() * sizeof(double));
    check_hip(hipMemcpy(LI64_, li.data(), li.size() * sizeof(double), hipMemcpyHostToDevice),
              "upload LI64");
  }
  if (serves(Precision::FP32)) {
    std::vector<float> l32(li.begin(), li.end());
    ensure(&LI32_, &gridcap_li32_, l32.size() * sizeof(float));
    check_hip(hipMemcpy(LI32_, l32.data(), l32.size() * sizeof(float), hipMemcpyHostToDevice),
              "upload LI32");
  }
}

void DeviceMultiEval::record_start() { check_hip(hipEventRecord(ev0_, stream_), "hipEventRecord"); }
void DeviceMultiEval::record_stop() { check_hip(hipEventRecord(ev1_, stream_), "hipEventRecord"); }

template <typename T>
void DeviceMultiEval::run_mult(const T* X, const T* LI, const double* d_alpha, double* d_result,
                               hipEvent_t* tev) {
  hipStream_t s = stream_;
  const int32_t* perm = skip_ ? static_cast<const int32_t*>(perm_) : nullptr;
  ensure(&A_, &A_cap_, Npad_ * sizeof(T));
  clear_launch_error();
  hipLaunchKernelGGL(prep_alpha_kernel<T>, dim3(ceil_div(Npad_, 256)), dim3(256), 0, s, d_alpha,
                     scale_, perm, static_cast<T*>(A_), static_cast<int64_t>(N_),
                     static_cast<int64_t>(Npad_));
  check_hip(hipGetLastError(), "prep_alpha");
  const int K = mult_k_rt<T>(static_cast<int>(D_));
  const int ntiles = ceil_div(N_, kTG);
  const int bx = static_cast<int>(Mpad_ / (kBlock * K));
  int G = cfg_.mult_grid_chunks;
  if (G <= 0) {
    // many short blocks: the dispatcher then balances the uneven per-block
    // work (the skip makes tiles cost different amounts) across all CUs
    const int target = cfg_.target_blocks > 0 ? cfg_.target_blocks : kBlocksPerCU * num_cus_;
    G = std::max(1, ceil_div(target, bx));
    G = std::min(G, std::max(1, ntiles / 4));  // at least 4 tiles per chunk
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
    MultFn<T> f = DimTable<T>::mult(static_cast<int>(D_));
    f(skip_, X, static_cast<int64_t>(Mpad_), LI, static_cast<const T*>(A_),
      static_cast<const uint32_t*>(qmask_), P, chunk, jend, G, static_cast<int64_t>(Mpad_), s);
  } else {
    dim3 grid(static_cast<unsigned>(Mpad_ / kBlock), static_cast<unsigned>(G));
    clear_launch_error();
    hipLaunchKernelGGL(mult_kernel_generic<T>, grid, dim3(kBlock), 0, s, X,
                       static_cast<int64_t>(Mpad_), LI, static_cast<const T*>(A_), P, chunk, jend,
                       static_cast<int>(D_));
  }
  check_hip(hipGetLastError(), "mult kernel");
  if (tev) check_hip(hipEventRecord(tev[1], s), "hipEventRecord");
  clear_launch_error();
  hipLaunchKernelGGL(reduce_parts_kernel, dim3(ceil_div(M_, 256)), dim3(256), 0, s, P,
                     static_cast<int64_t>(Mpad_), G, static_cast<const double*>(nullptr),
                     static_cast<const int32_t*>(nullptr), d_result, static_cast<int64_t>(M_));
  check_hip(hipGetLastError(), "reduce (mult)");
  last_G_ = G;
}

template <typename T>
void DeviceMultiEval::run_multT(const T* X, const T* LI, const double* d_y, double* d_result,
                                hipEvent_t* tev) {
  hipStream_t s = stream_;
  ensure(&Y_, &Y_cap_, Mpad_ * sizeof(T));
  clear_launch_error();
  hipLaunchKernelGGL(prep_y_kernel<T>, dim3(ceil_div(Mpad_, 256)), dim3(256), 0, s, d_y,
                     static_cast<T*>(Y_), static_cast<int64_t>(M_), static_cast<int64_t>(Mpad_));
  check_hip(hipGetLastError(), "prep_y");
  const int K = multT_k_rt<T>(static_cast<int>(D_), skip_);
  const int64_t cover = static_cast<int64_t>(kBlock) * K;
  const int bx = ceil_div(N_, cover);
  // data end: only tiles that contain real points (the rest of Mpad is padding)
  const int64_t mtiles = ceil_div(M_, kTM);
  const int64_t mend = mtiles * kTM;
  int S = cfg_.multT_data_segs;
  if (S <= 0) {
    const int target = cfg_.target_blocks > 0 ? cfg_.target_blocks : kBlocksPerCU * num_cus_;
    S = std::max(1, ceil_div(target, bx));
    S = static_cast<int>(std::min<int64_t>(S, std::max<int64_t>(1, mtiles / 4)));
  }
  S = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(S, mtiles)));
  const int64_t tiles_per = (mtiles + S - 1) / S;
  S = static_cast<int>((mtiles + tiles_per - 1) / tiles_per);
  const int64_t seg = tiles_per * kTM;
  ensure(&part_, &part_cap_, static_cast<size_t>(S) * Npad_ * sizeof(double));
  double* P = static_cast<double*>(part_);
  if (tev) check_hip(hipEventRecord(tev[0], s), "hipEventRecord");
  if (D_ <= static_cast<size_t>(kMaxTemplD)) {
    MultTFn<T> f = DimTable<T>::multT(static_cast<int>(D_));
    f(skip_, X, static_cast<int64_t>(Mpad_), static_cast<const T*>(Y_), LI,
      static_cast<const uint32_t*>(tmask_), P, static_cast<int64_t>(Npad_), seg, mend, S,
      static_cast<int64_t>(N_), s);
  } else {
    dim3 grid(static_cast<unsigned>(ceil_div(N_, kBlock)), static_cast<unsigned>(S));
    clear_launch_error();
    hipLaunchKernelGGL(multT_kernel_generic<T>, grid, dim3(kBlock), 0, s, X,
                       static_cast<int64_t>(Mpad_), static_cast<const T*>(Y_), LI, P,
                       static_cast<int64_t>(Npad_), seg, mend, static_cast<int>(D_));
  }
  check_hip(hipGetLastError(), "multTranspose kernel");
  if (tev) check_hip(hipEventRecord(tev[1], s), "hipEventRecord");
  clear_launch_error();
  hipLaunchKernelGGL(reduce_parts_kernel, dim3(ceil_div(N_, 256)), dim3(256), 0, s, P,
                     static_cast<int64_t>(Npad_), S, scale_,
                     skip_ ? static_cast<const int32_t*>(perm_) : nullptr, d_result,
                     static_cast<int64_t>(N_));
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
    run_mult<double>(static_cast<const double*>(X64_), static_cast<const double*>(LI64_), d_alpha,
                     d_result, tp ? tev : nullptr);
  else
    run_mult<float>(static_cast<const float*>(X32_), static_cast<const float*>(LI32_), d_alpha,
                    d_result, tp ? tev : nullptr);
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
    run_multT<double>(static_cast<const double*>(X64_), static_cast<const double*>(LI64_), d_y,
                      d_result, tp ? tev : nullptr);
  else
    run_multT<float>(static_cast<const float*>(X32_), static_cast<const float*>(LI32_), d_y,
                     d_result, tp ? tev : nullptr);
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
