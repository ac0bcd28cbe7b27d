// gfx950 kernels of the multi-evaluation operator (mult, multTranspose, the
// generic-dimension fall-backs, prep and reduction kernels) and their launch
// tables. Included by meval_kernels.hip (the runtime, plain kernels) and by
// meval_count.hip (the work-counting instantiations), so the two sets of
// template instantiations compile in parallel. Everything here has internal
// linkage; the counting launchers cross the file boundary through the
// function-pointer types at the end.
//
// See meval_kernels.hip for the arithmetic, layout and culling scheme.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

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
// at most 16 tiles (FP32) or 24 (FP64: as fast per launch, fewer partials to
// reduce); multTranspose up to ~1024 workgroups per CU (each block
// covers kChunkT grid points; the data segments stop at one 64*K-point group
// per wave). Measured on the d=10 ModLinear (1e5 x 1e6), d=8 Linear
// (3.2e4 x 5e5) and d=20 Linear (1e6 x 1e7) workloads (on the first, FP32
// mult with 16-tile chunks took 1.97 ms against 2.08 with 24 and 1.98 with
// 12; on the last, multTranspose with 264 data segments took 312 ms against
// 362 with 66).
constexpr int kBlocksPerCU = 128;
template <typename T>
constexpr int max_chunk_tiles() {
  return sizeof(T) == 4 ? 16 : 24;
}
constexpr int kBlocksPerCUT = 1024;
// cap on the FP64 partials of one launch (G x Mpad or S x Npad doubles)
// (on the d=20 Linear workload, 1e6 x 1e7, 512 MiB held mult to 6 grid
// chunks and multTranspose to 66 data segments; 2 GiB allows 26 and 268,
// 13-14 % faster)
constexpr int64_t kPartialBudget = int64_t(2048) << 20;
// data rows past M kept in the padded data (a wave's last group of 64*K
// points may start just below M)
constexpr size_t kDataSlack = 1024;

// Points per lane (K) and grid points evaluated together (J). Values from
// variant sweeps on the d=10 ModLinear, d=8 Linear and d=20 Linear workloads.
// Kernels for d > 10, and FP32 kernels of any d, are held to 128 registers
// (4 waves per SIMD) although some values then spill: at d = 20 that was
// 4-21 % faster than 3 waves without spills, and the FP32 d = 10 mult kernel
// (129 registers left to itself) 12 % faster; FP64 at d = 10 was 10-40 %
// slower, so FP64 kernels with d <= 10 keep their natural allocation.
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

// The box for a tile test. FP64 boxes of up to 10 dimensions (2*D register
// pairs) are re-read with scalar loads for each test rather than held across
// the loops: held, they are spilled to vector lanes and read back with one
// v_readlane per value. (Above 10 dimensions the kernels are held to 128
// vector registers, and re-reading the box made them spill more vector
// registers instead.) The empty asm makes g opaque, so the loads are not
// hoisted out of the loop.
template <int D, typename T>
__device__ __forceinline__ WaveBox<D, T> test_box(const T* __restrict__ WB, int64_t g,
                                                  const WaveBox<D, T>& held) {
  if constexpr (sizeof(T) == 8 && D <= 10) {
    asm volatile("" : "+s"(g));
    WaveBox<D, T> b;
    load_box<D, T>(WB, g, b);
    return b;
  } else {
    return held;
  }
}

// ---------------------------------------------------------------- mult
// Lanes own data points (K per lane, coordinates in registers). The data are
// stored in Morton order, so a wave's 64*K points fill a small box. For every
// 64-point grid tile of its chunk a wave
//   1. skips the tile if the tile's box of supports misses its data box;
//   2. tests the tile's points against the box (one point per lane, ballot);
//   3. evaluates only points that may be nonzero in the box, J at a time,
//      over the union of their non-constant dimensions.
// Every dropped term is exactly zero, so the sums equal a full evaluation.
// X: [D][ldx] SoA, WB: [ldx/(64*K)][D][2] boxes of the data groups, LI:
// point-major pairs, TB: [Npad/64][D][2] FP32 tile boxes, A: [Npad]
// alpha*scale (zero padded), part: [G][ldx]. Block: kBlock*K
// data points (each wave 64*K consecutive ones) x grid chunk [y*chunk,
// +chunk). chunk_fastest: the grid is (chunks, data blocks) instead of
// (data blocks, chunks), so a data block's chunks run one after the other and
// its coordinates stay in the caches (the partials are the same either way).
// kSkip: test for constant (ModLinear level-1) dimensions.
// kCount: also add the wave's executed hat steps and pairs to cnt[0], cnt[1].
template <int D, int K, int J, typename T, bool kSkip, bool kCount>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(D > 10 || sizeof(T) == 4 ? 4 : 1)))
void mult_kernel(const T* __restrict__ X, int64_t ldx, const T* __restrict__ WB,
                 const T* __restrict__ LI, const float* __restrict__ TB,
                 const uint32_t* __restrict__ PM, const T* __restrict__ A,
                 double* __restrict__ part, int chunk, int jend_all, int chunk_fastest,
                 unsigned long long* __restrict__ cnt) {
  const unsigned bdata = chunk_fastest ? blockIdx.y : blockIdx.x;
  const unsigned bchunk = chunk_fastest ? blockIdx.x : blockIdx.y;
  // a wave's 64*K points are contiguous (a small box in Morton order): point
  // k of a lane is m0 + 64*k; the wave is data group g
  const int lane = threadIdx.x & 63;
  const int64_t g = static_cast<int64_t>(bdata) * (kBlock / 64) +
                    __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t m0 = g * (64 * K) + lane;
  WaveBox<D, T> box;
  load_box<D, T>(WB, g, box);
  // wave-uniform base per dimension, per-lane 32-bit offset (the loads take
  // the scalar-base form: no per-load address arithmetic)
  const T* xg = X + g * (64 * K);
  T x[K][D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const T* xd = xg + d * ldx;
#pragma unroll
    for (int k = 0; k < K; ++k) x[k][d] = xd[static_cast<unsigned>(k * 64 + lane)];
  }
  double accd[K];
#pragma unroll
  for (int k = 0; k < K; ++k) accd[k] = 0.0;
  const int jb = static_cast<int>(bchunk) * chunk;
  const int je = min(jb + chunk, jend_all);
  unsigned long long n_steps = 0, n_pairs = 0;
  // windows of up to 64 tiles: one lane per tile tests the tile's box
  for (int w0 = jb; w0 < je; w0 += 64 * kTG) {
  unsigned long long tmask = tiles_may_hit<D, D, T>(
      TB, w0 / kTG, min(64, (je - w0 + kTG - 1) / kTG), lane, test_box(WB, g, box));
  while (tmask) {
    const int j0 = w0 + __builtin_ctzll(tmask) * kTG;
    tmask &= tmask - 1;
    const int64_t t = j0 / kTG;
    const T* tile = LI + t * D * kTG * 2;
    // this lane's point's dimensions, loaded with the tile test's rows (the
    // empty asm uses it there, so the load is not sunk to its first use)
    const uint32_t pm = kSkip ? PM[j0 + lane] : 0u;
    unsigned long long bal = tile_hits<D, D, T>(tile, lane, test_box(WB, g, box));
    if (kSkip) asm volatile("" ::"v"(pm));
    T acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = T(0);
    while (bal) {
      int jj[J];
      bool used[J];
      T a[J];
      take_hits<J, true, T>(bal, A, j0, jj, used, a);
      T p[J][K];
      const int ev =
          eval_group<D, J, K, T, kSkip>(tile, jj, x, p, kSkip ? group_dims<J>(pm, jj) : ~0u);
      if (kCount) {
        n_steps += static_cast<unsigned long long>(ev) * (J * K * 64);
        n_pairs += J * K * 64;
      }
#pragma unroll
      for (int q = 0; q < J; ++q)
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = fmaT(a[q], p[q][k], acc[k]);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) accd[k] += static_cast<double>(acc[k]);
  }
  }
  double* out = part + static_cast<int64_t>(bchunk) * ldx;
#pragma unroll
  for (int k = 0; k < K; ++k) out[m0 + k * 64] = accd[k];
  if (kCount && lane == 0) {
    atomicAdd(cnt, n_steps);
    atomicAdd(cnt + 1, n_pairs);
  }
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
// contiguous wave shares (seg and the shares are multiples of 64*K, so every
// group of 64*K points is one of WB's data groups). part: [S][ldp].
template <int D, int K, int J, typename T, bool kSkip, bool kCount>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(D > 10 || sizeof(T) == 4 ? 4 : 1)))
void multT_kernel(const T* __restrict__ X, int64_t ldx, const T* __restrict__ WB,
                  const T* __restrict__ Y, const T* __restrict__ LI, const float* __restrict__ TB,
                  const uint32_t* __restrict__ PM, double* __restrict__ part, int64_t ldp,
                  int jend_all, int64_t seg, int64_t mend, int chunk_fastest,
                  unsigned long long* __restrict__ cnt) {
  __shared__ double accs[kBlock / 64][kChunkT];
  // chunk_fastest: the grid is (chunks, segments), as in mult_kernel
  const unsigned bseg = chunk_fastest ? blockIdx.y : blockIdx.x;
  const unsigned bchunk = chunk_fastest ? blockIdx.x : blockIdx.y;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int jb = static_cast<int>(bchunk) * kChunkT;
  const int je = min(jb + kChunkT, jend_all);
  for (int i = threadIdx.x; i < (kBlock / 64) * kChunkT; i += kBlock) (&accs[0][0])[i] = 0.0;
  __syncthreads();
  const int64_t mb = static_cast<int64_t>(bseg) * seg;
  const int64_t me = min(mb + seg, mend);
  // per-wave share, a multiple of 64*K; points past the share weigh 0
  const int64_t share = ((me - mb) / (kBlock / 64) + 64 * K - 1) / (64 * K) * (64 * K);
  const int64_t wb = mb + wave * share;
  const int64_t we = min(wb + share, me);
  double* acc = accs[wave];
  unsigned long long n_steps = 0, n_pairs = 0;
  for (int64_t m0 = wb; m0 < we; m0 += 64 * K) {
    WaveBox<D, T> box;
    load_box<D, T>(WB, m0 / (64 * K), box);
    // wave-uniform base per dimension, per-lane 32-bit offset (scalar-base
    // loads)
    T x[K][D], y[K];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const T* xd = X + d * ldx + m0;
#pragma unroll
      for (int k = 0; k < K; ++k) x[k][d] = xd[static_cast<unsigned>(k * 64 + lane)];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int64_t m = m0 + k * 64 + lane;
      y[k] = (m < we) ? Y[m] : T(0);
    }
    // the chunk's tiles (kChunkT / kTG <= 64), one lane per tile
    static_assert(kChunkT / kTG <= 64, "one window of tiles per chunk");
    unsigned long long tmask =
        tiles_may_hit<D, D, T>(TB, jb / kTG, (je - jb + kTG - 1) / kTG, lane, box);
    while (tmask) {
      const int j0 = jb + __builtin_ctzll(tmask) * kTG;
      tmask &= tmask - 1;
      const int64_t t = j0 / kTG;
      const T* tile = LI + t * D * kTG * 2;
      // this lane's point's dimensions, loaded with the tile test's rows (the
      // empty asm uses it there, so the load is not sunk to its first use)
      const uint32_t pm = kSkip ? PM[j0 + lane] : 0u;
      unsigned long long bal = tile_hits<D, D, T>(tile, lane, test_box(WB, m0 / (64 * K), box));
      if (kSkip) asm volatile("" ::"v"(pm));
      // a grid point is taken at most once per (group, tile): its sum goes
      // to lane jj of tsum, and the tile is added to the wave's row once
      T tsum = T(0);
      while (bal) {
        int jj[J];
        bool used[J];
        T unused_w[J];
        take_hits<J, false, T>(bal, static_cast<const T*>(nullptr), j0, jj, used, unused_w);
        T p[J][K];
        const int ev =
            eval_group<D, J, K, T, kSkip>(tile, jj, x, p, kSkip ? group_dims<J>(pm, jj) : ~0u);
        if (kCount) {
          n_steps += static_cast<unsigned long long>(ev) * (J * K * 64);
          n_pairs += J * K * 64;
        }
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
  double* out = part + static_cast<int64_t>(bseg) * ldp;
  for (int i = threadIdx.x; i < je - jb; i += kBlock) {
    double s = accs[0][i];
#pragma unroll
    for (int w = 1; w < kBlock / 64; ++w) s += accs[w][i];
    out[jb + i] = s;
  }
  if (kCount && lane == 0) {
    atomicAdd(cnt + 2, n_steps);
    atomicAdd(cnt + 3, n_pairs);
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
// kSkip: test for constant dimensions (ModLinear); Linear grids have none.
// kCount: the work-counting instantiation (cnt: 4 counters).
template <typename T, int D, bool kSkip, bool kCount = false>
void launch_mult_D(const T* X, int64_t ldx, const T* WB, const T* LI, const float* TB,
                   const uint32_t* PM, const T* A, double* part, int chunk, int jend, int G,
                   int64_t Mpad, int chunk_fastest, unsigned long long* cnt, hipStream_t s) {
  constexpr int K = mult_k<T>(D);
  const unsigned bx = static_cast<unsigned>(Mpad / (kBlock * K));
  const dim3 grid = chunk_fastest ? dim3(static_cast<unsigned>(G), bx)
                                  : dim3(bx, static_cast<unsigned>(G));
  clear_launch_error();
  hipLaunchKernelGGL((mult_kernel<D, K, mult_j<T>(D), T, kSkip, kCount>), grid, dim3(kBlock), 0,
                     s, X, ldx, WB, LI, TB, PM, A, part, chunk, jend, chunk_fastest, cnt);
}

template <typename T, int D, bool kSkip, bool kCount = false>
void launch_multT_D(const T* X, int64_t ldx, const T* WB, const T* Y, const T* LI, const float* TB,
                    const uint32_t* PM, double* part, int64_t ldp, int jend, int64_t seg,
                    int64_t mend, int S, int chunk_fastest, unsigned long long* cnt,
                    hipStream_t s) {
  constexpr int K = multT_k<T>(D);
  // data segments fastest: consecutive blocks share one grid chunk (its rows
  // stay hot in every XCD's caches) and read different data; or chunks
  // fastest: consecutive blocks share a data segment
  const unsigned nch = static_cast<unsigned>((jend + kChunkT - 1) / kChunkT);
  const dim3 grid = chunk_fastest ? dim3(nch, static_cast<unsigned>(S))
                                  : dim3(static_cast<unsigned>(S), nch);
  clear_launch_error();
  hipLaunchKernelGGL((multT_kernel<D, K, multT_j<T>(D), T, kSkip, kCount>), grid, dim3(kBlock),
                     0, s, X, ldx, WB, Y, LI, TB, PM, part, ldp, jend, seg, mend, chunk_fastest,
                     cnt);
}

template <typename T>
using MultFn = void (*)(const T*, int64_t, const T*, const T*, const float*, const uint32_t*,
                        const T*, double*, int, int, int, int64_t, int, unsigned long long*,
                        hipStream_t);
template <typename T>
using MultTFn = void (*)(const T*, int64_t, const T*, const T*, const T*, const float*,
                         const uint32_t*, double*, int64_t, int, int64_t, int64_t, int, int,
                         unsigned long long*, hipStream_t);

template <typename T, bool kCount, int... Ds>
struct Table {
  static MultFn<T> mult(int D, bool skip) {
    MultFn<T> f = nullptr;
    ((D == Ds ? (f = skip ? &launch_mult_D<T, Ds, true, kCount>
                          : &launch_mult_D<T, Ds, false, kCount>, 0)
              : 0),
     ...);
    return f;
  }
  static MultTFn<T> multT(int D, bool skip) {
    MultTFn<T> f = nullptr;
    ((D == Ds ? (f = skip ? &launch_multT_D<T, Ds, true, kCount>
                          : &launch_multT_D<T, Ds, false, kCount>, 0)
              : 0),
     ...);
    return f;
  }
};
template <typename T, bool kCount = false>
using DimTable =
    Table<T, kCount, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20>;

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

// Launchers of the work-counting instantiations (meval_count.hip): the same
// kernels, which also add per wave the executed hat steps (pair x evaluated
// dimension) and pairs (J slots x K points x 64 lanes per group) to cnt[0],
// cnt[1] (mult) or cnt[2], cnt[3] (multTranspose).
template <typename T>
using CountMultFn = MultFn<T>;
template <typename T>
using CountMultTFn = MultTFn<T>;
template <typename T>
CountMultFn<T> count_mult_launcher(int D, bool skip);
template <typename T>
CountMultTFn<T> count_multT_launcher(int D, bool skip);
extern template CountMultFn<float> count_mult_launcher<float>(int, bool);
extern template CountMultFn<double> count_mult_launcher<double>(int, bool);
extern template CountMultTFn<float> count_multT_launcher<float>(int, bool);
extern template CountMultTFn<double> count_multT_launcher<double>(int, bool);

}  // namespace sgpp_hip
