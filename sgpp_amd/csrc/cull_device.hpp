// Device building blocks shared by the multi-evaluation kernels and the
// evaluation server: the hat-function step, DPP wave reductions, and the
// culling tests of grid tiles and grid points against a wave's box of
// data points.
//
// Culling is exact: a grid point is dropped only when, for some dimension,
// fma(L, hi, -I) < -1 or fma(L, lo, -I) > 1 over the wave's box [lo, hi].
// fma rounds monotonically in x (L >= 0), so the kernel's own u = fma(L, x, -I)
// then has |u| >= 1 for every x in the box, and its factor
// clamp01(p - p*|u|) is exactly 0.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sgpp_hip {

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

// ------------------------------------------------------- wave helpers
// Cross-lane moves with DPP row controls (VALU operand modifiers, no LDS
// traffic). Lanes whose source is outside the row, or whose row is not in
// RowMask, receive `old`.
template <int Ctrl, int RowMask = 0xf>
__device__ __forceinline__ float dpp(float old, float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), Ctrl, RowMask, 0xf, false));
}
template <int Ctrl, int RowMask = 0xf>
__device__ __forceinline__ double dpp(double old, double v) {
  const long long o = __double_as_longlong(old), s = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(static_cast<int>(o), static_cast<int>(s), Ctrl, RowMask,
                                             0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(static_cast<int>(o >> 32), static_cast<int>(s >> 32),
                                             Ctrl, RowMask, 0xf, false);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
// Cross-lane move without an `old` operand: lanes without a source lane read
// 0 (bound_ctrl); lanes of rows outside RowMask receive an unspecified value.
template <int Ctrl, int RowMask = 0xf>
__device__ __forceinline__ float dpp_z(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), Ctrl, RowMask, 0xf, true));
}
template <int Ctrl, int RowMask = 0xf>
__device__ __forceinline__ double dpp_z(double v) {
  const long long s = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(s), Ctrl, RowMask, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(s >> 32), Ctrl, RowMask, 0xf, true);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
// Reduction over the 64 lanes; the result is complete in lane 63 only.
// Steps: swap within pairs and halves of each quad (every lane then holds its
// quad's value), shift right by 4 and 8 within each row of 16 (lane 15 of a
// row holds the row's value; lanes without a source add 0), broadcast lane 15
// into the next row (lanes 31 and 63 hold two rows each), broadcast lane 31
// into rows 2-3 (lane 63). The two broadcasts leave rows 0, 2 (then 0, 1)
// with unspecified sums; neither is read again: step 5 reads lanes 15 and 47
// before it writes, and step 6 reads lane 31, which step 5 completed.
template <typename T>
__device__ __forceinline__ T sum_to_lane63(T v) {
  v += dpp_z<0xb1>(v);        // quad_perm [1,0,3,2]
  v += dpp_z<0x4e>(v);        // quad_perm [2,3,0,1]
  v += dpp_z<0x114>(v);       // row_shr:4
  v += dpp_z<0x118>(v);       // row_shr:8
  v += dpp_z<0x142, 0xa>(v);  // row_bcast:15 into rows 1, 3
  v += dpp_z<0x143, 0xc>(v);  // row_bcast:31 into rows 2, 3
  return v;
}
// lane 63's value, moved to a scalar register
__device__ __forceinline__ float lane63(float v) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ double lane63(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffLL), 63);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), 63);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
// The box [lo, hi] of a wave's 64*K data points (wave-uniform: scalar
// registers).
template <int DC, typename T>
struct WaveBox {
  T l[DC], h[DC];
};

// The box of data group g, computed once when the data are set (the min and
// max of the group's points per dimension, in T): WB[g][d] = {lo, hi}. g is
// wave-uniform, so the loads are scalar.
template <int D, typename T>
__device__ __forceinline__ void load_box(const T* __restrict__ WB, int64_t g, WaveBox<D, T>& box) {
  const T* b = WB + g * (D * 2);
#pragma unroll
  for (int d = 0; d < D; ++d) {
    box.l[d] = b[2 * d];
    box.h[d] = b[2 * d + 1];
  }
}

// dimensions per batch of bound loads in the tile and point tests
constexpr int kTestBatch = 8;

// Tile pre-test for up to 64 tiles at once, one tile per lane: bit t of the
// result is set when tile t0 + t (t < nt <= 64) may hit the box, i.e. when
// the tile's union of supports TB[t][d] = [lo, hi] (FP32, rounded outwards;
// compared in T, which is exact) meets the box in every tested dimension.
// The bounds are loaded in batches of kTestBatch dimensions and all tests of
// a batch run after its loads, so the loads are issued together. One vector
// load round trip covers a whole window of tiles. TB rows hold D dimensions;
// the first DC are tested (testing fewer keeps more tiles, never drops one
// that can be nonzero).
template <int DC, int D, typename T>
__device__ __forceinline__ unsigned long long tiles_may_hit(const float* __restrict__ TB,
                                                            int64_t t0, int nt, int lane,
                                                            const WaveBox<DC, T>& box) {
  unsigned h = 0u;
  if (lane < nt) {
    const float* tb = TB + (t0 + lane) * D * 2;
    h = 1u;
#pragma unroll
    for (int d0 = 0; d0 < DC; d0 += kTestBatch) {
      float b[2 * kTestBatch];
#pragma unroll
      for (int e = 0; e < 2 * kTestBatch; ++e)
        if (2 * d0 + e < 2 * DC) b[e] = tb[2 * d0 + e];
#pragma unroll
      for (int dd = 0; dd < kTestBatch; ++dd)
        if (d0 + dd < DC)
          h &= static_cast<unsigned>(static_cast<T>(b[2 * dd]) < box.h[d0 + dd]) &
               static_cast<unsigned>(static_cast<T>(b[2 * dd + 1]) > box.l[d0 + dd]);
    }
  }
  return __ballot(h != 0u);
}

// One lane per point of the tile: which points may be nonzero in the box
// (the lane's (L, I) row is read in batches of kTestBatch dimensions).
// Rows hold D dimensions; the first DC are tested.
// FP32 rows of up to 12 dimensions are read in one batch (one wait).
// A point is dropped only if its factor max(0, 1 - |L x - I|) (L >= 0) is
// exactly zero for every x in the box [lo, hi] in some dimension: in exact
// arithmetic when L*hi - I <= -1 or L*lo - I >= 1. fma rounds monotonically,
// so fma(L, hi, -I) < -1 implies L*hi - I < -1, and the kernel's own
// u = fma(L, x, -I) then has |u| >= 1 for all x in the box. The test over all
// dimensions is two reductions (min3 / max3 instructions and two compares per
// point): dropped when min_d fma(L, hi, -I) < -1 or max_d fma(L, lo, -I) > 1.
template <int DC, int D, typename T>
__device__ __forceinline__ unsigned long long tile_hits(const T* __restrict__ tile, int lane,
                                                        const WaveBox<DC, T>& box) {
  constexpr int B = sizeof(T) == 4 && DC <= 12 ? DC : kTestBatch;
  const T* row = tile + lane * D * 2;
  T lo_min = T(0), hi_max = T(0);
#pragma unroll
  for (int d0 = 0; d0 < DC; d0 += B) {
    T r[2 * B];
#pragma unroll
    for (int e = 0; e < 2 * B; ++e)
      if (2 * d0 + e < 2 * DC) r[e] = row[2 * d0 + e];
#pragma unroll
    for (int dd = 0; dd < B; ++dd)
      if (d0 + dd < DC) {
        const T a = fmaT(r[2 * dd], box.h[d0 + dd], -r[2 * dd + 1]);
        const T b = fmaT(r[2 * dd], box.l[d0 + dd], -r[2 * dd + 1]);
        lo_min = d0 + dd == 0 ? a : fmin(lo_min, a);
        hi_max = d0 + dd == 0 ? b : fmax(hi_max, b);
      }
  }
  return __ballot(!(lo_min < T(-1)) && !(hi_max > T(1)));
}

// Takes up to J points from the hit mask (wave-uniform). Missing slots
// repeat slot 0 and are flagged unused; w = the points' weights W[j0 + jj]
// (0 for unused slots; none without kW), loaded before the products so that
// their latency overlaps the evaluation.
template <int J, bool kW, typename T>
__device__ __forceinline__ void take_hits(unsigned long long& bal, const T* __restrict__ W, int j0,
                                          int (&jj)[J], bool (&used)[J], T (&w)[J]) {
#pragma unroll
  for (int q = 0; q < J; ++q) {
    if (bal) {
      const int b = __builtin_ctzll(bal);
      bal &= bal - 1;
      jj[q] = b;
      used[q] = true;
      w[q] = kW ? W[j0 + b] : T(0);
    } else {
      jj[q] = jj[0];
      used[q] = false;
      w[q] = T(0);
    }
  }
}

// the union of the non-constant-dimension masks of the J taken points: lane
// jj[q] holds point jj[q]'s mask in pm (unused slots repeat slot 0)
template <int J>
__device__ __forceinline__ uint32_t group_dims(uint32_t pm, const int (&jj)[J]) {
  uint32_t m = 0u;
#pragma unroll
  for (int q = 0; q < J; ++q)
    m |= static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(pm), jj[q]));
  return m;
}

// An (L, I) pair, loaded as one 8- or 16-byte vector.
template <typename T>
struct alignas(2 * sizeof(T)) LIPair {
  T L, I;
};

// Products p[q][k] = phi_{jj[q]}(x_k) of the J taken points of a tile. The
// points' (L, I) pairs are read in batches of DB dimensions (wave-uniform
// addresses: wide scalar loads). With kSkip, only the dimensions of `live`
// are evaluated: the union of the taken points' non-constant dimensions (a
// ModLinear level-1 factor, (L, I) = (0, 0), is exactly 1); Linear grids
// have no constant dimension. Returns the number of dimensions evaluated
// (wave-uniform).
template <int D, int J, int K, typename T, bool kSkip>
__device__ __forceinline__ int eval_group(const T* __restrict__ tile, const int (&jj)[J],
                                          const T (&x)[K][D], T (&p)[J][K], uint32_t live_all) {
  // dimensions per batch: the batch's pairs take at most 48 scalar registers
  constexpr int kWordsPerDim = J * 2 * static_cast<int>(sizeof(T)) / 4;
  constexpr int DB = D < 48 / kWordsPerDim ? D : 48 / kWordsPerDim;
  const LIPair<T>* rows = reinterpret_cast<const LIPair<T>*>(tile);
#pragma unroll
  for (int q = 0; q < J; ++q)
#pragma unroll
    for (int k = 0; k < K; ++k) p[q][k] = T(1);
  int evaluated = 0;
#pragma unroll
  for (int d0 = 0; d0 < D; d0 += DB) {
    LIPair<T> li[J][DB];
#pragma unroll
    for (int q = 0; q < J; ++q)
#pragma unroll
      for (int dd = 0; dd < DB; ++dd)
        if (d0 + dd < D) li[q][dd] = rows[jj[q] * D + d0 + dd];
    if constexpr (kSkip) {
      // every pair of the batch is used here, so the loads are issued together
      // and waited for once (left to itself the compiler sinks each load into
      // the branch of its dimension's mask bit: one wait per branch)
#pragma unroll
      for (int q = 0; q < J; ++q)
#pragma unroll
        for (int dd = 0; dd < DB; ++dd)
          if (d0 + dd < D) asm volatile("" ::"s"(li[q][dd].L), "s"(li[q][dd].I));
    }
    const uint32_t live = kSkip ? (live_all >> d0) : ~0u;
    evaluated += __builtin_popcount(live & ((1u << (D - d0 < DB ? D - d0 : DB)) - 1u));
#pragma unroll
    for (int dd = 0; dd < DB; ++dd) {
      const int d = d0 + dd;
      if (d >= D || !((live >> dd) & 1u)) continue;
#pragma unroll
      for (int q = 0; q < J; ++q)
#pragma unroll
        for (int k = 0; k < K; ++k) p[q][k] = hat_step(p[q][k], li[q][dd].L, x[k][d], li[q][dd].I);
    }
  }
  return evaluated;
}

}  // namespace sgpp_hip
