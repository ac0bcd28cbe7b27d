// Resident evaluation server (continuous batching across models).
// See server.hpp for the design; kernels and the scheduler live here.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

#include "cull_device.hpp"
#include "server.hpp"
#include "subspace_device.hpp"

namespace sgpp_hip {

namespace {

constexpr int kBlock = 256;
constexpr int kTG = 64;
constexpr int kMaxD = 20;
constexpr int kMaxModels = 4096;

struct ModelDesc {
  const void* LI;         // tile-major (L, I): [Npad/64][D][64][2]
  const void* A;          // [Npad] alpha * scale, zero padded
  const uint32_t* qmask;  // [Npad/4] dimensions evaluated per 4 consecutive points
  const uint32_t* umask;  // [Npad/J] dimensions whose (L, I) a sub-tile's J points share
  const uint32_t* hmask;  // [Npad/J] dimensions in which a sub-tile point has L >= kLoCorrectL
  int D;
  int Npad;               // multiple of kTG
  // subspace algorithm (null / 0 when the model has no subspace layout)
  const SubspaceDesc* SR;  // records sorted by level vector (+ kSubBatch - 1 padding)
  const double* TW;        // per table entry: alpha of its grid point, 0 where none
  int nsub;                // subspaces
  int mod;                 // ModLinear basis
};

// FP32 serving: a query coordinate x (FP64) is carried as hi = fl32(x) and
// lo = fl32(x - hi). fma(L, hi, -I) is exact inside the support (Sterbenz for
// I >= 2; for I <= 1 the one rounding is below 2^-25), so the only input
// error of u = L x - I is L * lo, up to L * 2^-25. Factors with
// L >= kLoCorrectL add it back, u = fma(L, lo, fma(L, hi, -I)); below that
// the dropped term is at most 32 * 2^-25 = 2^-20 per factor. (The threshold
// trades served throughput against that bound: on the 64-model serving
// workload L >= 32 served 4.9e7 evals/s, L >= 64 5.1e7, L >= 128 5.3e7,
// no correction 5.7e7.)
constexpr float kLoCorrectL = 64.0f;

// The queries of one model in a batch ("span"): stored [D][ldq] at xoff and
// cut into ldq / tile query tiles. Its work items are the (query tile, grid
// chunk) pairs, tile-major, numbered from item0; item (t, c) writes partial
// row pbase + (t * nchunks + c) * tile. Its reduction slots (one per padded
// query) are numbered from slot0; ldq, and so slot0, are multiples of kBlock.
struct Span {
  int model;   // model table index
  int D;
  int tile;    // queries per tile (kBlock * K)
  int nchunks;
  int chunk;   // grid points per chunk, a multiple of kTG
  int nq;      // valid queries
  int ldq;     // padded query count, a multiple of tile
  int item0;
  int slot0;
  int lo;      // 1: the low parts of the queries follow at xoff + D * ldq
  long long xoff;
  long long pbase;
  long long out0;  // output offset of the span's first query
};
static_assert(sizeof(Span) == 64, "one span is 64 bytes");

// Subspace-algorithm spans: the same layout, with `chunk` subspaces per
// item, `nchunks` = ceil(nsub / chunk) and one query per lane (tile = kBlock;
// two per lane at d <= 10 spilled registers and served no more);
// their items are numbered apart from the streaming items and follow them in
// the work list (counts below).
// counts[0] = streaming items, [1] = reduction slots, [2] = streaming spans
// (spans [0, counts[2])), [3] = subspace items, [4] = all spans (the subspace
// spans are [counts[2], counts[4])).
constexpr int kCounts = 8;

// one work item, decoded from its span
struct Item {
  int q0;          // first query of the tile (span-relative)
  int j0, j1;      // grid chunk, multiples of kTG
  int ldq;
  int lo;          // low parts present (FP32 queries of a model with hmask bits)
  long long xoff;
  long long poff;  // double offset of the item's partial row
};

// index of the span that holds position v, given each span's first
// position first(k) (increasing); wave-uniform arguments give scalar code
template <typename F>
__device__ __forceinline__ int find_span(int nspans, int v, F first) {
  int lo = 0, hi = nspans;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (first(mid) <= v)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

template <typename T>
constexpr int serve_k(int D) {
  return sizeof(T) == 8 ? (D <= 5 ? 4 : (D <= 10 ? 2 : 1)) : (D <= 4 ? 8 : (D <= 10 ? 4 : 2));
}
// grid points per sub-tile (wave-uniform branches per sub-tile and
// dimension): J*K = 32 (FP32) or 16 (FP64) running products per lane, J at
// most 16 (FP32) or 8 (FP64) so that every dimension case of the persistent
// kernel stays within one register budget (FP64 with J = 16 needs 180
// registers). Larger sub-tiles spread a shared dimension's per-query factor
// over more points.
template <typename T>
constexpr int serve_j(int D) {
  return (sizeof(T) == 8 ? 16 : 32) / serve_k<T>(D) < (sizeof(T) == 8 ? 8 : 16)
             ? (sizeof(T) == 8 ? 16 : 32) / serve_k<T>(D)
             : (sizeof(T) == 8 ? 8 : 16);
}

// LDS staging buffer of one 64-point grid tile: weights A (kTG values),
// per-4-point masks (kTG/4 words), shared-dimension and low-part masks per
// J-point sub-tile (kTG/J <= kTG/4 words each), then the tile-major (L, I)
// pairs (D*kTG*2 values); every part starts on a 16-byte boundary.
template <typename T>
constexpr int stage_a_off() { return 0; }
template <typename T>
constexpr int stage_m_off() { return kTG * static_cast<int>(sizeof(T)); }
template <typename T>
constexpr int stage_u_off() { return stage_m_off<T>() + kTG; }
template <typename T>
constexpr int stage_h_off() { return stage_u_off<T>() + kTG; }
template <typename T>
constexpr int stage_li_off() { return stage_h_off<T>() + kTG; }
template <typename T>
constexpr int stage_bytes(int D) { return stage_li_off<T>() + kTG * D * 2 * static_cast<int>(sizeof(T)); }

// Block-wide copy of n 16-byte chunks from global memory straight into LDS
// (global_load_lds_dwordx4: no registers; completion via vmcnt). Each wave
// copies 64 consecutive chunks per instruction; dst is wave-uniform.
__device__ __forceinline__ void copy_to_lds(const uint4* __restrict__ src, uint4* dst, int n,
                                            int wave, int lane) {
  for (int c0 = wave * 64; c0 < n; c0 += kBlock)
    if (c0 + lane < n) __builtin_amdgcn_global_load_lds(src + c0 + lane, dst + c0, 16, 0, 0);
}

// Starts the copy of grid tile j0 of model md (sub-tiles of J points) into the
// staging buffer buf.
template <int D, int J, typename T>
__device__ __forceinline__ void stage_tile(const ModelDesc& md, int j0, unsigned char* buf) {
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const unsigned char* A = static_cast<const unsigned char*>(md.A);
  const unsigned char* M = reinterpret_cast<const unsigned char*>(md.qmask);
  const unsigned char* LI = static_cast<const unsigned char*>(md.LI);
  copy_to_lds(reinterpret_cast<const uint4*>(A + static_cast<long long>(j0) * sizeof(T)),
              reinterpret_cast<uint4*>(buf + stage_a_off<T>()), kTG * sizeof(T) / 16, wave, lane);
  copy_to_lds(reinterpret_cast<const uint4*>(M + j0), reinterpret_cast<uint4*>(buf + stage_m_off<T>()),
              kTG / 16, wave, lane);
  copy_to_lds(reinterpret_cast<const uint4*>(md.umask + j0 / J),
              reinterpret_cast<uint4*>(buf + stage_u_off<T>()), kTG / J / 4, wave, lane);
  copy_to_lds(reinterpret_cast<const uint4*>(md.hmask + j0 / J),
              reinterpret_cast<uint4*>(buf + stage_h_off<T>()), kTG / J / 4, wave, lane);
  copy_to_lds(reinterpret_cast<const uint4*>(LI + static_cast<long long>(j0) * D * 2 * sizeof(T)),
              reinterpret_cast<uint4*>(buf + stage_li_off<T>()), kTG * D * 2 * sizeof(T) / 16,
              wave, lane);
}

// one work item: a tile of kBlock*K queries of one request against one grid
// chunk. The chunk's 64-point grid tiles are staged in LDS, double-buffered:
// the copy of tile t+1 runs while tile t is evaluated, one barrier per tile.
// A tile is walked in sub-tiles of J points; per (sub-tile, dimension)
// wave-uniform tests of the sub-tile's masks decide whether the dimension is
// skipped (ModLinear level-1 factors are exactly 1; Linear models carry full
// masks), evaluated once per query because all J points share its (L, I)
// (the factor then scales the sub-tile's weighted sum), or evaluated per
// point. FP32 items with low parts add L * lo in the dimensions of the
// sub-tile's low-part mask (lo is read from the batch buffer there, so it
// holds no registers across the loop).
template <int D, int K, int J, typename T>
__device__ __forceinline__ void serve_item(const Item& item, const ModelDesc& md, const T* X,
                                           double* part, unsigned char* sbuf, int bufbytes) {
  static_assert(kTG % J == 0 && J % 4 == 0, "sub-tile must divide the tile, 4 must divide it");
  constexpr bool kLo = sizeof(T) == 4;
  const T* Xr = X + item.xoff;
  const T* Xl = Xr + static_cast<long long>(D) * item.ldq;  // low parts (when item.lo)
  const bool lo = kLo && item.lo != 0;
  T x[K][D];
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int d = 0; d < D; ++d)
      x[k][d] = Xr[static_cast<long long>(d) * item.ldq + item.q0 + k * kBlock + threadIdx.x];
  double accd[K];
#pragma unroll
  for (int k = 0; k < K; ++k) accd[k] = 0.0;
  int cur = 0;
  stage_tile<D, J, T>(md, item.j0, sbuf);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (int j0 = item.j0; j0 < item.j1; j0 += kTG) {
    if (j0 + kTG < item.j1) stage_tile<D, J, T>(md, j0 + kTG, sbuf + (cur ^ 1) * bufbytes);
    const unsigned char* buf = sbuf + cur * bufbytes;
    const T* sA = reinterpret_cast<const T*>(buf + stage_a_off<T>());
    const uint32_t* sM = reinterpret_cast<const uint32_t*>(buf + stage_m_off<T>());
    const uint32_t* sU = reinterpret_cast<const uint32_t*>(buf + stage_u_off<T>());
    const uint32_t* sH = reinterpret_cast<const uint32_t*>(buf + stage_h_off<T>());
    const T* sLI = reinterpret_cast<const T*>(buf + stage_li_off<T>());
    T acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = T(0);
    for (int j = 0; j < kTG; j += J) {
      uint32_t mask = 0u;
#pragma unroll
      for (int c = 0; c < J / 4; ++c) mask |= sM[j / 4 + c];
      mask = __builtin_amdgcn_readfirstlane(mask);
      const uint32_t shared = __builtin_amdgcn_readfirstlane(sU[j / J]);
      const uint32_t hm = lo ? __builtin_amdgcn_readfirstlane(sH[j / J]) : 0u;
      T p[J][K];
      T g[K];  // product of the factors the J points share
#pragma unroll
      for (int k = 0; k < K; ++k) {
        g[k] = T(1);
#pragma unroll
        for (int q = 0; q < J; ++q) p[q][k] = T(1);
      }
      // shared dimensions first, in a loop of their own (skipped when the
      // sub-tile shares none), then the per-point ones: two simple loops
      // compile to fewer scalar instructions than one with both cases
      const uint32_t msh = mask & shared, mrest = mask & ~shared;
      const uint32_t msh_plain = msh & ~hm, mplain = mrest & ~hm;
      if (msh_plain != 0u) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
          if (!((msh_plain >> d) & 1u)) continue;
          const T L = sLI[(d * kTG + j) * 2], I = sLI[(d * kTG + j) * 2 + 1];
#pragma unroll
          for (int k = 0; k < K; ++k) g[k] *= clamp01(T(1) - absT(fmaT(L, x[k][d], -I)));
        }
      }
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (!((mplain >> d) & 1u)) continue;
        T li[2 * J];
#pragma unroll
        for (int c = 0; c < 2 * J; ++c) li[c] = sLI[(d * kTG + j) * 2 + c];
#pragma unroll
        for (int q = 0; q < J; ++q)
#pragma unroll
          for (int k = 0; k < K; ++k) {
            const T u = fmaT(li[2 * q], x[k][d], -li[2 * q + 1]);
            p[q][k] = clamp01(fmaT(-p[q][k], absT(u), p[q][k]));
          }
      }
      // dimensions that add the low parts: a scalar loop over the mask bits
      // (few per sub-tile), hi and lo re-read from the batch buffer, two
      // queries per lane at a time (registers)
      if (kLo && (mask & hm) != 0u) {
        constexpr int KC = K < 2 ? K : 2;
        for (uint32_t mc = mask & hm; mc != 0u; mc &= mc - 1u) {
          const int d = __builtin_ctz(mc);
          const long long off = static_cast<long long>(d) * item.ldq + item.q0 + threadIdx.x;
          const bool sh = (shared >> d) & 1u;
#pragma unroll
          for (int k0 = 0; k0 < K; k0 += KC) {
            T xh[KC], xl[KC];
#pragma unroll
            for (int k = 0; k < KC; ++k) {
              xh[k] = Xr[off + (k0 + k) * kBlock];
              xl[k] = Xl[off + (k0 + k) * kBlock];
            }
            if (sh) {
              const T L = sLI[(d * kTG + j) * 2], I = sLI[(d * kTG + j) * 2 + 1];
#pragma unroll
              for (int k = 0; k < KC; ++k)
                g[k0 + k] *= clamp01(T(1) - absT(fmaT(L, xl[k], fmaT(L, xh[k], -I))));
            } else {
#pragma unroll
              for (int q = 0; q < J; ++q) {
                const T L = sLI[(d * kTG + j + q) * 2], I = sLI[(d * kTG + j + q) * 2 + 1];
#pragma unroll
                for (int k = 0; k < KC; ++k) {
                  const T u = fmaT(L, xl[k], fmaT(L, xh[k], -I));
                  p[q][k0 + k] = clamp01(fmaT(-p[q][k0 + k], absT(u), p[q][k0 + k]));
                }
              }
            }
          }
        }
      }
      T sub[K];
#pragma unroll
      for (int k = 0; k < K; ++k) sub[k] = T(0);
#pragma unroll
      for (int q = 0; q < J; ++q) {
        const T a = sA[j + q];
#pragma unroll
        for (int k = 0; k < K; ++k) sub[k] = fmaT(a, p[q][k], sub[k]);
      }
#pragma unroll
      for (int k = 0; k < K; ++k) acc[k] = fmaT(g[k], sub[k], acc[k]);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) accd[k] += static_cast<double>(acc[k]);
    // the next tile's copy has landed and every wave is done with this one
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    cur ^= 1;
  }
  double* out = part + item.poff;
#pragma unroll
  for (int k = 0; k < K; ++k) out[k * kBlock + threadIdx.x] = accd[k];
}

int serve_k_rt(Precision p, int D) {
  return p == Precision::FP64 ? serve_k<double>(D) : serve_k<float>(D);
}

int serve_j_rt(Precision p, int D) {
  return p == Precision::FP64 ? serve_j<double>(D) : serve_j<float>(D);
}

template <typename T, int D>
__device__ __forceinline__ void dispatch_one(const Item& it, const ModelDesc& md, const T* X,
                                             double* part, unsigned char* sbuf) {
  serve_item<D, serve_k<T>(D), serve_j<T>(D), T>(it, md, X, part, sbuf, stage_bytes<T>(kMaxD));
}

// One subspace item: queries [q0, q0 + kBlock) of a span (one per lane)
// against subspaces [sb, se) of its model, in FP64: x = hi (+ lo for FP32
// queries), per subspace the changed dimensions, the weight gathered from the
// table (the next record prefetched); the first subspace of the
// item starts from scratch. Padding queries (zero coordinates) give values
// the reduction ignores.
template <int D, bool kMod, typename T>
__device__ void serve_sub_item(const ModelDesc& md, const T* __restrict__ X, long long xoff,
                               int ldq, int lo, int q0, int sb, int se, double* __restrict__ out) {
  const long long q = q0 + static_cast<int>(threadIdx.x);
  double x[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    x[d] = static_cast<double>(X[xoff + static_cast<long long>(d) * ldq + q]);
    if (sizeof(T) == 4 && lo)
      x[d] += static_cast<double>(X[xoff + static_cast<long long>(D + d) * ldq + q]);
  }
  double prod[D + 1];
  unsigned flat[D + 1];
  prod[0] = 1.0;
  flat[0] = 0u;
  double acc = 0.0;
  // Records are read as eight words (one wide scalar load, constant address
  // space). The next record's load is issued once the current record has
  // arrived (the empty asm uses all of it, and the scheduling barrier keeps
  // the load below), so it overlaps this subspace's work; record se exists
  // (padding). The weights are gathered through a global-address-space
  // pointer (global rather than flat loads).
  typedef const __attribute__((address_space(4))) unsigned* RecPtr;
  typedef const __attribute__((address_space(1))) double* WPtr;
  const RecPtr R = (RecPtr)(md.SR);
  const WPtr TW = (WPtr)(md.TW);
  static_assert(sizeof(SubspaceDesc) == 8 * sizeof(unsigned), "eight words per record");
  unsigned cur[8];
#pragma unroll
  for (int w = 0; w < 8; ++w) cur[w] = R[8 * sb + w];
  for (int s = sb; s < se; ++s) {
    asm volatile("" ::"s"(cur[0]), "s"(cur[1]), "s"(cur[2]), "s"(cur[3]), "s"(cur[4]),
                 "s"(cur[5]), "s"(cur[6]), "s"(cur[7]));
    __builtin_amdgcn_sched_barrier(0);
    unsigned nxt[8];
#pragma unroll
    for (int w = 0; w < 8; ++w) nxt[w] = R[8 * (s + 1) + w];
    SubspaceDesc sd;
    __builtin_memcpy(&sd, cur, sizeof(SubspaceDesc));
    walk_subspace<D, kMod>(x, sd, s == sb ? 0 : sd.recalc, prod, flat);
    acc = __builtin_fma(TW[sd.off + flat[D]], prod[D], acc);
#pragma unroll
    for (int w = 0; w < 8; ++w) cur[w] = nxt[w];
  }
  out[threadIdx.x] = acc;
}

template <int D, typename T>
__device__ __forceinline__ void dispatch_sub(const ModelDesc& md, const T* X, const Span& sp,
                                             int q0, int sb, int se, double* out) {
  if (md.mod)
    serve_sub_item<D, true, T>(md, X, sp.xoff, sp.ldq, sp.lo, q0, sb, se, out);
  else
    serve_sub_item<D, false, T>(md, X, sp.xoff, sp.ldq, sp.lo, q0, sb, se, out);
}

// Persistent work-list kernel: blocks pull items from a counter until the
// list is exhausted (every block reaches the exit test). Items [0, counts[0])
// are streaming items of the spans [0, counts[2]); items from counts[0] on are
// the counts[3] subspace items of the spans [counts[2], counts[4]). One kernel
// runs both kinds, so one tail drains the batch. FP32 is held to 128
// registers (4 waves per SIMD); the FP64 dimension cases need more and are
// held to 168 (3 waves per SIMD).
template <typename T>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 3)))
void serve_kernel(const ModelDesc* __restrict__ models, const Span* __restrict__ spans,
                  const int* __restrict__ counts, unsigned* counter, const T* __restrict__ X,
                  double* __restrict__ part) {
  // two staging buffers (double-buffered grid tiles)
  __shared__ __align__(16) unsigned char sbuf[2 * stage_bytes<T>(kMaxD)];
  __shared__ unsigned s_item;
  const int ns = counts[0];
  const int n = ns + counts[3];
  const int nspans = counts[2];
  const Span* sspans = spans + counts[2];
  const int nsubspans = counts[4] - counts[2];
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1u);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(static_cast<int>(s_item));
    __syncthreads();
    if (it >= n || it < 0) break;
    if (it >= ns) {
      // subspace item
      const int is = it - ns;
      const Span sp = sspans[find_span(nsubspans, is, [&](int k) { return sspans[k].item0; })];
      const ModelDesc md = models[sp.model];
      const int local = is - sp.item0;
      const int t = local / sp.nchunks;
      const int c = local - t * sp.nchunks;
      const int sb = c * sp.chunk;
      const int se = min(sb + sp.chunk, md.nsub);
      double* out = part + sp.pbase + static_cast<long long>(local) * sp.tile;
      switch (sp.D) {
#define SGPP_CASE(DD) \
  case DD:            \
    dispatch_sub<DD, T>(md, X, sp, t * sp.tile, sb, se, out); \
    break;
        SGPP_CASE(1) SGPP_CASE(2) SGPP_CASE(3) SGPP_CASE(4) SGPP_CASE(5) SGPP_CASE(6) SGPP_CASE(7)
        SGPP_CASE(8) SGPP_CASE(9) SGPP_CASE(10) SGPP_CASE(11) SGPP_CASE(12) SGPP_CASE(13)
        SGPP_CASE(14) SGPP_CASE(15) SGPP_CASE(16) SGPP_CASE(17) SGPP_CASE(18) SGPP_CASE(19)
        SGPP_CASE(20)
#undef SGPP_CASE
        default:
          break;
      }
      continue;
    }
    const Span sp = spans[find_span(nspans, it, [&](int k) { return spans[k].item0; })];
    const ModelDesc md = models[sp.model];
    const int local = it - sp.item0;
    const int t = local / sp.nchunks;
    Item item;
    item.q0 = t * sp.tile;
    item.j0 = (local - t * sp.nchunks) * sp.chunk;
    item.j1 = min(item.j0 + sp.chunk, md.Npad);
    item.ldq = sp.ldq;
    item.lo = sp.lo;
    item.xoff = sp.xoff;
    item.poff = sp.pbase + static_cast<long long>(local) * sp.tile;
    switch (sp.D) {
#define SGPP_CASE(DD) \
  case DD:            \
    dispatch_one<T, DD>(item, md, X, part, sbuf); \
    break;
      SGPP_CASE(1) SGPP_CASE(2) SGPP_CASE(3) SGPP_CASE(4) SGPP_CASE(5) SGPP_CASE(6) SGPP_CASE(7)
      SGPP_CASE(8) SGPP_CASE(9) SGPP_CASE(10) SGPP_CASE(11) SGPP_CASE(12) SGPP_CASE(13)
      SGPP_CASE(14) SGPP_CASE(15) SGPP_CASE(16) SGPP_CASE(17) SGPP_CASE(18) SGPP_CASE(19)
      SGPP_CASE(20)
#undef SGPP_CASE
      default:
        break;
    }
  }
}

// out[query] = the sum of the query's partials over its model's grid chunks,
// in chunk order. Each block takes kBlock consecutive reduction slots (all in
// one span); each thread keeps kRedBatch partial loads in flight and adds
// them in order. counts[1] = slots, counts[4] = spans.
constexpr int kRedBatch = 8;
constexpr int kReduceBlocksPerCU = 8;
__global__ __launch_bounds__(kBlock) void serve_reduce(const Span* __restrict__ spans,
                                                       const int* __restrict__ counts,
                                                       const double* __restrict__ part,
                                                       double* __restrict__ out) {
  const int nslots = counts[1];
  const int nspans = counts[4];
  for (int s0 = blockIdx.x * kBlock; s0 < nslots; s0 += gridDim.x * kBlock) {
    const Span sp = spans[find_span(nspans, s0, [&](int k) { return spans[k].slot0; })];
    const int local = s0 - sp.slot0 + static_cast<int>(threadIdx.x);
    if (local >= sp.nq) continue;  // padding query
    const int t = local / sp.tile;
    const long long step = sp.tile;
    const double* p = part + sp.pbase + static_cast<long long>(t) * sp.nchunks * step +
                      (local - t * sp.tile);
    double sum = 0.0;
    int c = 0;
    for (; c + kRedBatch <= sp.nchunks; c += kRedBatch) {
      double v[kRedBatch];
#pragma unroll
      for (int u = 0; u < kRedBatch; ++u) v[u] = p[(c + u) * step];
#pragma unroll
      for (int u = 0; u < kRedBatch; ++u) sum += v[u];
    }
    for (; c < sp.nchunks; ++c) sum += p[c * step];
    out[sp.out0 + local] = sum;
  }
}

__global__ void reset_counter(unsigned* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] = 0u;
}

// Span of q queries of one model (stored [D][ldq] at xoff, ldq a multiple of
// tile, followed by their [D][ldq] low parts when lo), its items numbered from
// item0, its slots from slot0, its partials at pbase and its results at out0.
Span make_span(int table_index, int D, size_t Npad, size_t tile, size_t chunk, size_t q,
               size_t ldq, long long xoff, size_t item0, size_t slot0, size_t pbase,
               long long out0, bool lo) {
  Span sp{};
  sp.lo = lo ? 1 : 0;
  sp.model = table_index;
  sp.D = D;
  sp.tile = static_cast<int>(tile);
  sp.nchunks = static_cast<int>((Npad + chunk - 1) / chunk);
  sp.chunk = static_cast<int>(chunk);
  sp.nq = static_cast<int>(q);
  sp.ldq = static_cast<int>(ldq);
  sp.item0 = static_cast<int>(item0);
  sp.slot0 = static_cast<int>(slot0);
  sp.xoff = xoff;
  sp.pbase = static_cast<long long>(pbase);
  sp.out0 = out0;
  return sp;
}

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// autotuner: synthetic request size, candidate grid chunks, partial cap
constexpr size_t kTuneQueries = 16384;
constexpr size_t kTuneBatchCopies = 8;
constexpr int kTuneChunks[] = {256, 512, 1024, 2048, 4096, 8192, 16384, 32768};
constexpr size_t kTunePartDoubles = size_t(1) << 25;

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

struct Server::Model {
  int id = 0;
  int table_index = 0;
  Basis basis = Basis::Linear;
  int D = 0;
  size_t N = 0, Npad = 0;
  void* LI = nullptr;
  void* A = nullptr;
  void* pm = nullptr;  // [Npad/4] dimensions evaluated per 4 consecutive points
  void* um = nullptr;  // [Npad/J] dimensions whose (L, I) the J points of a sub-tile share
  void* hm = nullptr;  // [Npad/J] dimensions of a sub-tile that add the queries' low parts
  bool lo = false;     // FP32 server and (some hm bit set or sub): queries carry low parts
  // subspace algorithm: records and table weights (when the layout exists),
  // whether the model is served by it, and subspaces per work item
  void* sr = nullptr;
  void* tw = nullptr;
  int nsub = 0;
  bool sub = false;
  int sub_chunk = 0;
  size_t bytes = 0;
  int active = 0;    // requests queued or running
  int inflight = 0;  // pieces of its requests in formed or running batches
  void free_device() {
    void* p[] = {LI, A, pm, um, hm, sr, tw};
    for (void* q : p)
      if (q) (void)hipFree(q);
    LI = A = pm = um = hm = sr = tw = nullptr;
  }
};

struct Server::Request {
  uint64_t id = 0;
  int model = 0;
  size_t q = 0;
  size_t next = 0;      // first query not yet placed in a batch
  size_t done = 0;      // queries whose results have arrived
  std::vector<char> X;  // [D][q] in the compute type; FP32: then the [D][q] low parts
  std::vector<double> out;
  double t_submit = 0.0, t_done = 0.0, deadline = 0.0;
  int status = REQ_QUEUED;  // QUEUED while queries remain unplaced
  std::string error;
};

// Queries [q0, q0 + nq) of a request, placed in one batch: columns
// [qoff, qoff + nq) of its model's block at xoff (leading dimension ldq);
// results at out_off of the batch output.
struct Server::Piece {
  std::shared_ptr<Request> r;
  size_t q0 = 0, nq = 0, qoff = 0, ldq = 0;
  long long xoff = 0, out_off = 0;
};

struct Server::Slot {
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_done = nullptr;
  hipGraphExec_t graph = nullptr;
  void* dX = nullptr;
  Span* dSpans = nullptr;
  int* dCounts = nullptr;
  unsigned* dCounter = nullptr;
  double* dPart = nullptr;
  double* dOut = nullptr;
  char* hX = nullptr;
  Span* hSpans = nullptr;
  int* hCounts = nullptr;
  double* hOut = nullptr;
  size_t capX = 0, capItems = 0, capSpans = 0, capPart = 0, capOut = 0;
  // current batch
  bool busy = false;
  bool inject = false;
  std::vector<Piece> pieces;
  struct Block {
    long long xoff;
    int D;
    size_t q, ldq;
    bool lo;
  };
  std::vector<Block> blocks;  // one per model in the batch (padding columns)
  size_t usedX = 0, nItems = 0, nSlots = 0, nSpans = 0, usedPart = 0, usedOut = 0;
  size_t nSubItems = 0, nStreamSpans = 0;  // subspace items; spans of the streaming kernel
};

// reset, work-list kernel, reduction of slot s's batch on stream st (also
// what the slot's graph holds)
void Server::enqueue_batch_kernels(Slot& s, hipStream_t st) {
  enqueue_kernels(st, s.dSpans, s.dCounts, s.dCounter, s.dX, s.dPart, s.dOut);
}

void Server::enqueue_kernels(hipStream_t st, const void* spans, const int* counts,
                             unsigned* counter, const void* X, double* part, double* out) {
  const ModelDesc* models = static_cast<const ModelDesc*>(model_table_);
  const Span* sp = static_cast<const Span*>(spans);
  clear_launch_error();
  hipLaunchKernelGGL(reset_counter, dim3(1), dim3(64), 0, st, counter);
  if (cfg_.precision == Precision::FP64) {
    hipLaunchKernelGGL(serve_kernel<double>, dim3(cfg_.blocks), dim3(kBlock), 0, st, models, sp,
                       counts, counter, static_cast<const double*>(X), part);
  } else {
    hipLaunchKernelGGL(serve_kernel<float>, dim3(cfg_.blocks), dim3(kBlock), 0, st, models, sp,
                       counts, counter, static_cast<const float*>(X), part);
  }
  hipLaunchKernelGGL(serve_reduce, dim3(kReduceBlocksPerCU * num_cus_), dim3(kBlock), 0, st, sp,
                     counts, part, out);
}

Server::Server(const ServerConfig& cfg) : cfg_(cfg) {
  if (cfg_.max_inflight < 1 || cfg_.max_inflight > 8) throw input_error("max_inflight in [1, 8]");
  if (cfg_.algorithm < kServeAuto || cfg_.algorithm > kServeSubspace)
    throw input_error("unknown serving algorithm");
  // item and slot numbers are ints on the device
  if (cfg_.max_items > (size_t(1) << 30) || cfg_.max_batch_queries > (size_t(1) << 28))
    throw input_error("max_items <= 2^30 and max_batch_queries <= 2^28");
  check_hip(hipSetDevice(cfg_.device), "hipSetDevice");
  hipDeviceProp_t prop;
  check_hip(hipGetDeviceProperties(&prop, cfg_.device), "hipGetDeviceProperties");
  num_cus_ = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  if (cfg_.blocks <= 0) cfg_.blocks = 8 * num_cus_;
  const size_t esz = cfg_.precision == Precision::FP64 ? 8 : 4;
  check_hip(hipMalloc(&model_table_, kMaxModels * sizeof(ModelDesc)), "hipMalloc(models)");
  used_bytes_ += kMaxModels * sizeof(ModelDesc);
  for (int i = 0; i < cfg_.max_inflight; ++i) {
    std::unique_ptr<Slot> s(new Slot());
    s->capOut = cfg_.max_batch_queries;
    s->capX = cfg_.max_batch_queries * kMaxD + size_t(kMaxD) * 4096 * 64;
    s->capItems = cfg_.max_items;
    s->capSpans = kMaxModels;  // one span per model and batch
    s->capPart = cfg_.max_partial_doubles;
    const size_t dev_bytes = s->capX * esz + s->capSpans * sizeof(Span) + kCounts * sizeof(int) +
                             s->capPart * sizeof(double) + s->capOut * sizeof(double);
    if (used_bytes_ + dev_bytes > cfg_.memory_budget)
      throw input_error("memory budget too small for the batch buffers");
    used_bytes_ += dev_bytes;
    check_hip(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
    check_hip(hipEventCreate(&s->ev_start), "hipEventCreate");
    check_hip(hipEventCreate(&s->ev_done), "hipEventCreate");
    check_hip(hipMalloc(&s->dX, s->capX * esz), "hipMalloc(batch X)");
    check_hip(hipMalloc(reinterpret_cast<void**>(&s->dSpans), s->capSpans * sizeof(Span)),
              "hipMalloc(spans)");
    check_hip(hipMalloc(reinterpret_cast<void**>(&s->dCounts), kCounts * sizeof(int)),
              "hipMalloc(counts)");
    check_hip(hipMalloc(reinterpret_cast<void**>(&s->dCounter), sizeof(unsigned)),
              "hipMalloc(counter)");
    check_hip(hipMalloc(reinterpret_cast<void**>(&s->dPart), s->capPart * sizeof(double)),
              "hipMalloc(partials)");
    check_hip(hipMalloc(reinterpret_cast<void**>(&s->dOut), s->capOut * sizeof(double)),
              "hipMalloc(out)");
    check_hip(hipHostMalloc(reinterpret_cast<void**>(&s->hX), s->capX * esz), "hipHostMalloc");
    check_hip(hipHostMalloc(reinterpret_cast<void**>(&s->hSpans), s->capSpans * sizeof(Span)),
              "hipHostMalloc");
    check_hip(hipHostMalloc(reinterpret_cast<void**>(&s->hCounts), kCounts * sizeof(int)),
              "hipHostMalloc");
    check_hip(hipHostMalloc(reinterpret_cast<void**>(&s->hOut), s->capOut * sizeof(double)),
              "hipHostMalloc");
    if (cfg_.use_graph) {
      // reset + the two persistent work-list kernels + reduction, captured once
      // per slot; the item counts are read from device memory, so one graph
      // serves every batch
      hipGraph_t g;
      check_hip(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal),
                "hipStreamBeginCapture");
      enqueue_batch_kernels(*s, s->stream);
      check_hip(hipStreamEndCapture(s->stream, &g), "hipStreamEndCapture");
      check_hip(hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0), "hipGraphInstantiate");
      check_hip(hipGraphDestroy(g), "hipGraphDestroy");
    }
    slots_.push_back(std::move(s));
  }
  load_cache();
  thread_ = std::thread(&Server::loop, this);
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
    for (auto& r : queue_)
      if (r->status == REQ_QUEUED) r->status = REQ_CANCELLED;
    queue_.clear();
  }
  cv_.notify_all();
  if (thread_.joinable()) thread_.join();
  for (auto& s : slots_) {
    (void)hipStreamSynchronize(s->stream);
    if (s->graph) (void)hipGraphExecDestroy(s->graph);
    void* dev[] = {s->dX, s->dSpans, s->dCounts, s->dCounter, s->dPart, s->dOut};
    for (void* p : dev)
      if (p) (void)hipFree(p);
    void* host[] = {s->hX, s->hSpans, s->hCounts, s->hOut};
    for (void* p : host)
      if (p) (void)hipHostFree(p);
    (void)hipEventDestroy(s->ev_start);
    (void)hipEventDestroy(s->ev_done);
    (void)hipStreamDestroy(s->stream);
  }
  for (auto& kv : models_) kv.second->free_device();
  if (model_table_) (void)hipFree(model_table_);
}

int Server::admit_model(Basis basis, size_t dim, size_t N, const uint32_t* level,
                        const uint32_t* index, const double* alpha) {
  if (dim < 1 || dim > static_cast<size_t>(kMaxD)) throw input_error("model dimension in [1, 20]");
  if (N == 0) throw input_error("empty model");
  // Points are sorted by their set of dimensions above level 1, then by their
  // encoded (L, I) pairs, dimension 0 first. Consecutive points then share
  // whole dimensions: ModLinear level-1 factors are constant and skipped per
  // 4 points; a dimension whose (L, I) all J points of a sub-tile share is
  // evaluated once per query. The model's output is a sum over all points,
  // so the order changes nothing but the summation order.
  const bool skip = basis == Basis::ModLinear && dim >= 2;
  std::vector<double> lip0, sc0;
  encode_grid(basis, dim, N, level, index, lip0, sc0);
  std::vector<uint32_t> cls(N), pmask(N);
  for (size_t j = 0; j < N; ++j) {
    uint32_t c = 0;
    for (size_t d = 0; d < dim; ++d)
      if (level[j * dim + d] > 1) c |= 1u << d;
    cls[j] = c;
    pmask[j] = skip ? c : (1u << dim) - 1u;
  }
  std::vector<size_t> order(N);
  for (size_t j = 0; j < N; ++j) order[j] = j;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    if (cls[a] != cls[b]) return cls[a] < cls[b];
    for (size_t e = 0; e < 2 * dim; ++e)
      if (lip0[a * 2 * dim + e] != lip0[b * 2 * dim + e])
        return lip0[a * 2 * dim + e] < lip0[b * 2 * dim + e];
    return false;
  });
  std::vector<double> lip(N * dim * 2), sc(N);
  for (size_t s2 = 0; s2 < N; ++s2) {
    for (size_t e = 0; e < 2 * dim; ++e) lip[s2 * 2 * dim + e] = lip0[order[s2] * 2 * dim + e];
    sc[s2] = sc0[order[s2]];
  }
  auto m = std::make_shared<Model>();
  m->basis = basis;
  m->D = static_cast<int>(dim);
  m->N = N;
  m->Npad = ((N + kTG - 1) / kTG) * kTG;
  const bool f64 = cfg_.precision == Precision::FP64;
  const size_t esz = f64 ? 8 : 4;
  const size_t J = static_cast<size_t>(serve_j_rt(cfg_.precision, m->D));
  m->bytes = m->Npad * dim * 2 * esz + m->Npad * esz +
             (m->Npad / 4 + 2 * (m->Npad / J)) * sizeof(uint32_t);
  std::vector<double> A(m->Npad, 0.0);
  for (size_t s2 = 0; s2 < N; ++s2) A[s2] = alpha[order[s2]] * sc[s2];
  // pairs as the kernel sees them (rounded to the compute type)
  auto pair_at = [&](size_t j, size_t d, int c) {
    const double v = lip[(j * dim + d) * 2 + c];
    return f64 ? v : static_cast<double>(static_cast<float>(v));
  };
  // tile-major (L, I); padded points (0, 0) evaluate to 1 with weight 0
  std::vector<double> li(m->Npad * dim * 2, 0.0);
  for (size_t j = 0; j < N; ++j)
    for (size_t d = 0; d < dim; ++d) {
      const size_t dst = (((j / kTG) * dim + d) * kTG + (j % kTG)) * 2;
      li[dst] = lip[(j * dim + d) * 2];
      li[dst + 1] = lip[(j * dim + d) * 2 + 1];
    }
  // padded points keep mask 0: nothing evaluated, weight 0
  std::vector<uint32_t> qm(m->Npad / 4, 0u);
  for (size_t s2 = 0; s2 < N; ++s2) qm[s2 / 4] |= pmask[order[s2]];
  // per sub-tile of J points: dimensions in which every real point has the
  // (L, I) of the sub-tile's first point (padded points carry weight 0, so
  // their factor does not matter; a sub-tile of padding only is never
  // evaluated)
  std::vector<uint32_t> um(m->Npad / J, 0u);
  for (size_t t = 0; t * J < N; ++t) {
    const size_t j0 = t * J, j1 = std::min(N, j0 + J);
    uint32_t u = 0;
    for (size_t d = 0; d < dim; ++d) {
      bool same = true;
      for (size_t j = j0 + 1; j < j1 && same; ++j)
        same = pair_at(j, d, 0) == pair_at(j0, d, 0) && pair_at(j, d, 1) == pair_at(j0, d, 1);
      if (same) u |= 1u << d;
    }
    um[t] = u;
  }
  // per sub-tile (FP32 only): dimensions in which some real point has
  // L >= kLoCorrectL, where the queries' low parts are added
  std::vector<uint32_t> hmk(m->Npad / J, 0u);
  if (!f64)
    for (size_t t = 0; t * J < N; ++t) {
      const size_t j0 = t * J, j1 = std::min(N, j0 + J);
      uint32_t h = 0;
      for (size_t d = 0; d < dim; ++d)
        for (size_t j = j0; j < j1; ++j)
          if (pair_at(j, d, 0) >= static_cast<double>(kLoCorrectL)) h |= 1u << d;
      hmk[t] = h;
      if (h) m->lo = true;
    }
  // subspace layout (records and FP64 table weights in the caller's point
  // order) unless the server runs the streaming kernel only; a grid whose
  // tables would be too large has none
  SubspaceLayout lay;
  std::vector<double> tw;
  if (cfg_.algorithm != kServeStreaming) {
    try {
      lay = build_subspace_layout(dim, N, level, index);
      tw.assign(lay.ti.size(), 0.0);
      for (size_t e = 0; e < tw.size(); ++e)
        if (lay.ti[e] >= 0) tw[e] = alpha[lay.ti[e]];
      m->nsub = lay.nsub;
      m->bytes += lay.desc.size() * sizeof(SubspaceDesc) + tw.size() * sizeof(double);
    } catch (const input_error&) {
      lay = SubspaceLayout();
      tw.clear();
    }
  }
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (used_bytes_ + m->bytes > cfg_.memory_budget)
      throw input_error("model does not fit in the memory budget");
    used_bytes_ += m->bytes;
    int idx = -1;
    std::vector<bool> used(kMaxModels, false);
    for (auto& kv : models_) used[kv.second->table_index] = true;
    for (int i = 0; i < kMaxModels; ++i)
      if (!used[i]) {
        idx = i;
        break;
      }
    if (idx < 0) {
      used_bytes_ -= m->bytes;
      throw input_error("too many resident models");
    }
    m->table_index = idx;
    m->id = next_model_++;
    models_[m->id] = m;
  }
  try {
    check_hip(hipSetDevice(cfg_.device), "hipSetDevice");
    check_hip(hipMalloc(&m->LI, m->Npad * dim * 2 * esz), "hipMalloc(model LI)");
    check_hip(hipMalloc(&m->A, m->Npad * esz), "hipMalloc(model A)");
    check_hip(hipMalloc(&m->pm, qm.size() * sizeof(uint32_t)), "hipMalloc(model masks)");
    check_hip(hipMemcpy(m->pm, qm.data(), qm.size() * sizeof(uint32_t), hipMemcpyHostToDevice),
              "upload masks");
    check_hip(hipMalloc(&m->um, um.size() * sizeof(uint32_t)), "hipMalloc(model shared masks)");
    check_hip(hipMemcpy(m->um, um.data(), um.size() * sizeof(uint32_t), hipMemcpyHostToDevice),
              "upload shared masks");
    check_hip(hipMalloc(&m->hm, hmk.size() * sizeof(uint32_t)), "hipMalloc(model low-part masks)");
    check_hip(hipMemcpy(m->hm, hmk.data(), hmk.size() * sizeof(uint32_t), hipMemcpyHostToDevice),
              "upload low-part masks");
    if (f64) {
      check_hip(hipMemcpy(m->LI, li.data(), li.size() * 8, hipMemcpyHostToDevice), "upload LI");
      check_hip(hipMemcpy(m->A, A.data(), A.size() * 8, hipMemcpyHostToDevice), "upload A");
    } else {
      std::vector<float> l32(li.begin(), li.end()), a32(A.begin(), A.end());
      check_hip(hipMemcpy(m->LI, l32.data(), l32.size() * 4, hipMemcpyHostToDevice), "upload LI");
      check_hip(hipMemcpy(m->A, a32.data(), a32.size() * 4, hipMemcpyHostToDevice), "upload A");
    }
    if (m->nsub > 0) {
      check_hip(hipMalloc(&m->sr, lay.desc.size() * sizeof(SubspaceDesc)),
                "hipMalloc(model subspaces)");
      check_hip(hipMemcpy(m->sr, lay.desc.data(), lay.desc.size() * sizeof(SubspaceDesc),
                          hipMemcpyHostToDevice),
                "upload subspaces");
      check_hip(hipMalloc(&m->tw, tw.size() * sizeof(double)), "hipMalloc(model tables)");
      check_hip(hipMemcpy(m->tw, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice),
                "upload tables");
    }
    ModelDesc md{m->LI,
                 m->A,
                 static_cast<const uint32_t*>(m->pm),
                 static_cast<const uint32_t*>(m->um),
                 static_cast<const uint32_t*>(m->hm),
                 m->D,
                 static_cast<int>(m->Npad),
                 static_cast<const SubspaceDesc*>(m->sr),
                 static_cast<const double*>(m->tw),
                 m->nsub,
                 basis == Basis::ModLinear ? 1 : 0};
    check_hip(hipMemcpy(static_cast<ModelDesc*>(model_table_) + m->table_index, &md,
                        sizeof(ModelDesc), hipMemcpyHostToDevice),
              "upload model descriptor");
  } catch (...) {
    std::lock_guard<std::mutex> lk(mu_);
    models_.erase(m->id);
    used_bytes_ -= m->bytes;
    m->free_device();
    throw;
  }
  {
    bool have = true;
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (cfg_.autotune) have = tuned_.count(tune_key(basis, m->D)) > 0;
      m->active++;  // not evictable while its launches run
    }
    try {
      if (!have) tune(*m);
      choose_algorithm(*m);
    } catch (...) {
      std::lock_guard<std::mutex> lk(mu_);
      m->active--;
      throw;
    }
    std::lock_guard<std::mutex> lk(mu_);
    // the subspace kernel reads FP64 queries: an FP32 server carries their low parts
    if (m->sub && !f64) m->lo = true;
    m->active--;
  }
  return m->id;
}


void Server::evict_model(int id) {
  std::shared_ptr<Model> m;
  {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = models_.find(id);
    if (it == models_.end()) throw input_error("unknown model");
    // pieces of a cancelled request can still sit in a formed or running batch
    if (it->second->active > 0 || it->second->inflight > 0)
      throw input_error("model has queued or running requests");
    m = it->second;
    models_.erase(it);
    used_bytes_ -= m->bytes;
  }
  m->free_device();
}

size_t Server::model_bytes(int id) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = models_.find(id);
  if (it == models_.end()) throw input_error("unknown model");
  return it->second->bytes;
}

size_t Server::memory_used() {
  std::lock_guard<std::mutex> lk(mu_);
  return used_bytes_;
}

uint64_t Server::submit(int model, size_t q, const double* queries, double deadline_s) {
  std::shared_ptr<Model> m;
  if (q == 0) throw input_error("empty request");
  {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = models_.find(model);
    if (it == models_.end()) throw input_error("unknown model");
    m = it->second;
  }
  const double t_enter = now_s();
  auto r = std::make_shared<Request>();
  r->model = model;
  r->q = q;
  r->out.assign(q, 0.0);
  const size_t D = m->D;
  const bool f64 = cfg_.precision == Precision::FP64;
  const bool lo = !f64 && m->lo;
  // [D][q] in the compute type; FP32: hi = fl32(x) and, for models that add
  // them, lo = fl32(x - hi) in rows [D, 2D). Transposed in blocks of kCvt
  // queries (each output row written contiguously), checked on the way.
  r->X.resize(D * q * (f64 ? 8 : (lo ? 8 : 4)));
  constexpr size_t kCvt = 64;
  bool in_cube = true;
  for (size_t i0 = 0; i0 < q; i0 += kCvt) {
    const size_t n = std::min(kCvt, q - i0);
    for (size_t d = 0; d < D; ++d) {
      const double* src = queries + i0 * D + d;
      if (f64) {
        double* h = reinterpret_cast<double*>(r->X.data()) + d * q + i0;
        for (size_t k = 0; k < n; ++k) {
          const double v = src[k * D];
          in_cube &= (v >= 0.0) & (v <= 1.0);
          h[k] = v;
        }
      } else {
        float* h = reinterpret_cast<float*>(r->X.data()) + d * q + i0;
        float* l = lo ? reinterpret_cast<float*>(r->X.data()) + (D + d) * q + i0 : nullptr;
        for (size_t k = 0; k < n; ++k) {
          const double v = src[k * D];
          in_cube &= (v >= 0.0) & (v <= 1.0);
          const float hv = static_cast<float>(v);
          h[k] = hv;
          if (lo) l[k] = static_cast<float>(v - static_cast<double>(hv));
        }
      }
    }
  }
  if (!in_cube) throw input_error("query points must lie in [0,1]^d");
  r->t_submit = t_enter;  // latency includes the conversion of the queries
  r->deadline = deadline_s > 0 ? r->t_submit + deadline_s : 0.0;
  {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = models_.find(model);
    if (it == models_.end()) throw input_error("model was evicted");
    it->second->active++;
    r->id = next_request_++;
    requests_[r->id] = r;
    queue_.push_back(r);
  }
  cv_.notify_all();
  return r->id;
}

int Server::poll(uint64_t id, double* out, size_t out_capacity, double wait_s,
                 double* latency_s) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = requests_.find(id);
  if (it == requests_.end()) return -1;
  std::shared_ptr<Request> r = it->second;
  auto finished = [&] { return r->status != REQ_QUEUED && r->status != REQ_RUNNING; };
  if (!finished() && wait_s != 0.0) {
    if (wait_s < 0)
      done_cv_.wait(lk, finished);
    else
      done_cv_.wait_for(lk, std::chrono::duration<double>(wait_s), finished);
  }
  if (!finished()) return r->status;
  if (r->status == REQ_DONE && out) {
    // the request stays available for a poll with a large enough buffer
    if (out_capacity < r->q) throw input_error("output buffer holds fewer values than the request");
    std::memcpy(out, r->out.data(), r->q * sizeof(double));
  }
  if (latency_s) *latency_s = r->t_done - r->t_submit;
  requests_.erase(id);
  return r->status;
}

bool Server::cancel(uint64_t id) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = requests_.find(id);
  if (it == requests_.end() || it->second->status != REQ_QUEUED) return false;
  fail_request(it->second, REQ_CANCELLED, "cancelled");
  return true;
}

void Server::inject_kernel_error() { inject_error_ = true; }

ServerStats Server::stats() {
  std::lock_guard<std::mutex> lk(mu_);
  return stats_;
}

void Server::drain() {
  std::unique_lock<std::mutex> lk(mu_);
  done_cv_.wait(lk, [&] { return queue_.empty() && running_ == 0; });
}

// caller holds mu_
void Server::fail_request(const std::shared_ptr<Request>& r, int status, const char* why) {
  r->status = status;
  r->error = why;
  r->t_done = now_s();
  auto m = models_.find(r->model);
  if (m != models_.end()) m->second->active--;
  stats_.requests_failed++;
  done_cv_.notify_all();
}

// caller holds mu_: the batch of slot s failed on the device; its requests
// fail (a request split over several batches fails once, at its first failed
// piece) and the slot is free again
void Server::fail_pieces(Slot& s, const char* why) {
  for (const Piece& pc : s.pieces) {
    if (pc.r->status == REQ_QUEUED || pc.r->status == REQ_RUNNING)
      fail_request(pc.r, REQ_FAILED_DEVICE, why);
    auto m = models_.find(pc.r->model);
    if (m != models_.end()) m->second->inflight--;
  }
  s.pieces.clear();
}

// caller holds mu_: move requests from the queue into slot s (FIFO). The
// queries of all requests for one model form one block [D][ldq] of the batch
// (in queue order), tiled and chunked together: a query's result depends only
// on the model and its grid chunks, so packing changes no result bit, and a
// request pays for at most one partially filled tile per model and batch.
// When the rest of the front request does not fit, its longest prefix that
// does is placed and the request stays at the front for the next batch.
bool Server::form_batch(Slot& s, std::vector<Piece>& taken) {
  s.usedX = s.nItems = s.nSlots = s.nSpans = s.usedPart = s.usedOut = 0;
  s.nSubItems = s.nStreamSpans = 0;
  s.blocks.clear();
  const double t = now_s();
  struct Group {
    const Model* m;
    size_t tile, chunk, nchunks, q;
    std::vector<size_t> pieces;  // indices into taken
  };
  std::vector<Group> groups;
  std::map<int, size_t> gidx;
  // resources of a block of q queries
  auto need = [](const Group& g, size_t q, size_t* xel, size_t* items, size_t* part) {
    const size_t ntiles = (q + g.tile - 1) / g.tile;
    *xel = static_cast<size_t>(g.m->D) * ntiles * g.tile * (g.m->lo ? 2 : 1);
    *items = ntiles * g.nchunks;
    *part = *items * g.tile;
  };
  size_t useX = 0, useItems = 0, usePart = 0, useOut = 0;
  while (!queue_.empty() && taken.size() < static_cast<size_t>(cfg_.max_requests_per_batch)) {
    std::shared_ptr<Request> r = queue_.front();
    if (r->status != REQ_QUEUED) {  // cancelled or failed meanwhile
      queue_.pop_front();
      continue;
    }
    if (r->next == 0 && r->deadline > 0 && t > r->deadline) {
      queue_.pop_front();
      fail_request(r, REQ_FAILED_TIMEOUT, "deadline passed before the request was scheduled");
      continue;
    }
    const Model& m = *models_.at(r->model);
    // subspace models: one query per lane, sub_chunk subspaces per item
    Group probe{&m,
                m.sub ? static_cast<size_t>(kBlock)
                      : static_cast<size_t>(kBlock) * serve_k_rt(cfg_.precision, m.D),
                m.sub ? static_cast<size_t>(m.sub_chunk) : chunk_for(m),
                0,
                0,
                {}};
    probe.nchunks = m.sub ? (static_cast<size_t>(m.nsub) + probe.chunk - 1) / probe.chunk
                          : (m.Npad + probe.chunk - 1) / probe.chunk;
    size_t x1, i1, p1;
    need(probe, 1, &x1, &i1, &p1);
    if (i1 > s.capItems || p1 > s.capPart || x1 > s.capX) {
      queue_.pop_front();
      fail_request(r, REQ_FAILED_BUDGET, "one query tile of this model exceeds the batch buffers");
      continue;
    }
    auto gi = gidx.find(r->model);
    const Group& g = gi == gidx.end() ? probe : groups[gi->second];
    size_t x0 = 0, i0 = 0, p0 = 0;
    if (gi != gidx.end()) need(g, g.q, &x0, &i0, &p0);
    // does a piece of n more queries fit? (monotone in n)
    auto fits = [&](size_t n) {
      need(g, g.q + n, &x1, &i1, &p1);
      return useX + x1 - x0 <= s.capX && useItems + i1 - i0 <= s.capItems &&
             usePart + p1 - p0 <= s.capPart && useOut + n <= s.capOut;
    };
    const size_t rest = r->q - r->next;
    size_t n = rest;
    if (!fits(rest)) {
      size_t lo = 0, hi = rest;  // fits(lo), !fits(hi)
      while (hi - lo > 1) {
        const size_t mid = lo + (hi - lo) / 2;
        (fits(mid) ? lo : hi) = mid;
      }
      n = lo;
      if (n == 0) break;  // full: the rest waits for the next slot (FIFO)
    }
    need(g, g.q + n, &x1, &i1, &p1);
    useX += x1 - x0;
    useItems += i1 - i0;
    usePart += p1 - p0;
    useOut += n;
    if (gi == gidx.end()) {
      gidx[r->model] = groups.size();
      groups.push_back(probe);
    }
    Group& gg = groups[gidx[r->model]];
    Piece pc;
    pc.r = r;
    pc.q0 = r->next;
    pc.nq = n;
    pc.qoff = gg.q;
    gg.q += n;
    gg.pieces.push_back(taken.size());
    taken.push_back(pc);
    models_.at(r->model)->inflight++;
    r->next += n;
    if (r->next == r->q) {
      queue_.pop_front();
      r->status = REQ_RUNNING;
    }
    if (n < rest) break;  // the batch is full
  }
  // the streaming kernel's spans first, then the subspace kernel's; each
  // kind numbers its items from 0, slots and partials run on across both
  for (int pass = 0; pass < 2; ++pass) {
    for (Group& g : groups) {
      if (g.m->sub != (pass == 1)) continue;
      const size_t ntiles = (g.q + g.tile - 1) / g.tile;
      const size_t ldq = ntiles * g.tile;
      const long long xoff = static_cast<long long>(s.usedX);
      const long long out0 = static_cast<long long>(s.usedOut);
      for (size_t k : g.pieces) {
        Piece& pc = taken[k];
        pc.xoff = xoff;
        pc.ldq = ldq;
        pc.out_off = out0 + static_cast<long long>(pc.qoff);
      }
      size_t& item0 = g.m->sub ? s.nSubItems : s.nItems;
      s.hSpans[s.nSpans++] = make_span(
          g.m->table_index, g.m->D, g.m->sub ? static_cast<size_t>(g.m->nsub) : g.m->Npad, g.tile,
          g.chunk, g.q, ldq, xoff, item0, s.nSlots, s.usedPart, out0, g.m->lo);
      s.blocks.push_back(Slot::Block{xoff, g.m->D, g.q, ldq, g.m->lo});
      item0 += ntiles * g.nchunks;
      s.nSlots += ldq;
      s.usedPart += ntiles * g.nchunks * g.tile;
      s.usedX += static_cast<size_t>(g.m->D) * ldq * (g.m->lo ? 2 : 1);
      s.usedOut += g.q;
    }
    if (pass == 0) s.nStreamSpans = s.nSpans;
  }
  return !taken.empty();
}

void Server::launch_batch(Slot& s) {
  const size_t esz = cfg_.precision == Precision::FP64 ? 8 : 4;
  // each request's rows into its model's block, padding columns zero
  // (FP32 models with low parts: rows d in [D, 2D) are the low parts)
  for (const Piece& pc : s.pieces) {
    const Request& r = *pc.r;
    const size_t rows = r.X.size() / (r.q * esz);
    for (size_t d = 0; d < rows; ++d)
      std::memcpy(s.hX + (pc.xoff + d * pc.ldq + pc.qoff) * esz,
                  r.X.data() + (d * r.q + pc.q0) * esz, pc.nq * esz);
  }
  for (const Slot::Block& b : s.blocks)
    for (int d = 0; d < b.D * (b.lo ? 2 : 1); ++d)
      std::memset(s.hX + (b.xoff + d * b.ldq + b.q) * esz, 0, (b.ldq - b.q) * esz);
  s.hCounts[0] = static_cast<int>(s.nItems);
  s.hCounts[1] = static_cast<int>(s.nSlots);
  s.hCounts[2] = static_cast<int>(s.nStreamSpans);
  s.hCounts[3] = static_cast<int>(s.nSubItems);
  s.hCounts[4] = static_cast<int>(s.nSpans);
  for (int k = 5; k < kCounts; ++k) s.hCounts[k] = 0;
  check_hip(hipEventRecord(s.ev_start, s.stream), "hipEventRecord");
  check_hip(hipMemcpyAsync(s.dX, s.hX, s.usedX * esz, hipMemcpyHostToDevice, s.stream), "H2D X");
  check_hip(hipMemcpyAsync(s.dSpans, s.hSpans, s.nSpans * sizeof(Span), hipMemcpyHostToDevice,
                           s.stream),
            "H2D spans");
  check_hip(hipMemcpyAsync(s.dCounts, s.hCounts, kCounts * sizeof(int), hipMemcpyHostToDevice,
                           s.stream),
            "H2D counts");
  if (s.graph) {
    check_hip(hipGraphLaunch(s.graph, s.stream), "hipGraphLaunch");
  } else {
    enqueue_batch_kernels(s, s.stream);
    check_hip(hipGetLastError(), "serve launch");
  }
  check_hip(hipMemcpyAsync(s.hOut, s.dOut, s.usedOut * sizeof(double), hipMemcpyDeviceToHost,
                           s.stream),
            "D2H out");
  check_hip(hipEventRecord(s.ev_done, s.stream), "hipEventRecord");
}

// caller holds mu_
void Server::finish_batch(Slot& s) {
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, s.ev_start, s.ev_done) == hipSuccess)
    stats_.busy_seconds += ms * 1e-3;
  const double t = now_s();
  for (const Piece& pc : s.pieces) {
    Request& r = *pc.r;
    auto mp = models_.find(r.model);
    if (mp != models_.end()) mp->second->inflight--;
    if (r.status != REQ_QUEUED && r.status != REQ_RUNNING) continue;  // failed or cancelled
    if (s.inject) {
      fail_request(pc.r, REQ_FAILED_DEVICE, "injected kernel error");
      continue;
    }
    std::memcpy(r.out.data() + pc.q0, s.hOut + pc.out_off, pc.nq * sizeof(double));
    r.done += pc.nq;
    if (r.done < r.q) continue;
    r.X.clear();
    r.X.shrink_to_fit();
    r.status = REQ_DONE;
    r.t_done = t;
    auto m = models_.find(r.model);
    if (m != models_.end()) m->second->active--;
    stats_.requests_done++;
    stats_.queries_done += r.q;
  }
  stats_.batches++;
  stats_.kernel_launches += cfg_.use_graph ? 1 : 4;
  s.pieces.clear();
  s.busy = false;
  s.inject = false;
  running_--;
  done_cv_.notify_all();
}

int Server::tune_key(Basis basis, int D) { return (basis == Basis::ModLinear ? 64 : 0) + D; }

// Grid chunks per query of a streaming model at most: each chunk is one FP64
// partial per query, written by the kernel, read back by the reduction and
// counted against the batch's partial buffer. (The tuned chunk is timed on one
// model per basis and dimension; a model with many more grid points would get
// proportionally more chunks.) On the 64-model stream at 16000 req/s a cap of
// 256 raised the served rate from 8.3e7 to 8.8e7 evals/s (64: the same) and
// lowered p99 at 2000 req/s.
constexpr size_t kMaxServeChunks = 256;

// caller holds mu_: grid points per work item of model m
size_t Server::chunk_for(const Model& m) const {
  size_t chunk;
  auto it = tuned_.find(tune_key(m.basis, m.D));
  if (it != tuned_.end()) {
    chunk = static_cast<size_t>(it->second.chunk);
  } else {
    const double per_point =
        static_cast<double>(kBlock) * serve_k_rt(cfg_.precision, m.D) * (2 * m.D + 1);
    chunk = static_cast<size_t>(std::ceil(cfg_.item_work / per_point / kTG)) * kTG;
  }
  const size_t floor_chunk = (m.Npad + kMaxServeChunks * kTG - 1) / (kMaxServeChunks * kTG) * kTG;
  chunk = std::max(chunk, floor_chunk);
  return std::max<size_t>(kTG, std::min(chunk, m.Npad));
}

int Server::model_algorithm(int id) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = models_.find(id);
  if (it == models_.end()) throw input_error("unknown model");
  return it->second->sub ? kServeSubspace : kServeStreaming;
}

bool Server::tuned_shape(Basis basis, int dim, TunedShape* out) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = tuned_.find(tune_key(basis, dim));
  if (it == tuned_.end()) return false;
  if (out) *out = it->second;
  return true;
}

// Times every candidate grid chunk (grid points per work item) on the model
// with synthetic uniform queries, alone on its own stream, and keeps the
// fastest with its cost per (query x grid point); the time covers the
// work-list kernel and the reduction of its partials. The load follows the
// server's batching: one request of kTuneQueries when every launch holds one
// request (latency of a lone request), kTuneBatchCopies times as many when
// batches pack many requests (throughput of a saturated device). The result
// applies to every model of the same basis and dimension on this server and
// is written to the cache file when one is configured.
void Server::tune(Model& m) {
  const bool f64 = cfg_.precision == Precision::FP64;
  const size_t esz = f64 ? 8 : 4;
  const size_t D = static_cast<size_t>(m.D);
  const size_t nq = kTuneQueries * (cfg_.max_requests_per_batch == 1 ? 1 : kTuneBatchCopies);
  const size_t tile = static_cast<size_t>(kBlock) * serve_k_rt(cfg_.precision, m.D);
  const size_t ntiles = nq / tile;
  std::vector<char> hx(nq * D * esz * (m.lo ? 2 : 1));
  uint64_t st = 0x5EED0000ull + static_cast<uint64_t>(D);
  for (size_t d = 0; d < D; ++d)
    for (size_t i = 0; i < nq; ++i) {
      const double v = static_cast<double>(splitmix64(st) >> 11) * 0x1.0p-53;
      if (f64) {
        reinterpret_cast<double*>(hx.data())[d * nq + i] = v;
      } else {
        const float hi = static_cast<float>(v);
        reinterpret_cast<float*>(hx.data())[d * nq + i] = hi;
        if (m.lo)
          reinterpret_cast<float*>(hx.data())[(D + d) * nq + i] =
              static_cast<float>(v - static_cast<double>(hi));
      }
    }
  std::vector<int> cands;
  for (int c : kTuneChunks)
    if (static_cast<size_t>(c) < m.Npad && ntiles * ((m.Npad + c - 1) / c) * tile <= kTunePartDoubles)
      cands.push_back(c);
  cands.push_back(static_cast<int>(m.Npad));
  check_hip(hipSetDevice(cfg_.device), "hipSetDevice");
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  void* dX = nullptr;
  Span* dSpans = nullptr;
  int* dCounts = nullptr;
  unsigned* dCounter = nullptr;
  double* dPart = nullptr;
  double* dOut = nullptr;
  auto release = [&] {
    if (stream) (void)hipStreamSynchronize(stream);
    void* p[] = {dX, dSpans, dCounts, dCounter, dPart, dOut};
    for (void* x : p)
      if (x) (void)hipFree(x);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (stream) (void)hipStreamDestroy(stream);
  };
  TunedShape res{};
  try {
    check_hip(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    check_hip(hipEventCreate(&e0), "hipEventCreate");
    check_hip(hipEventCreate(&e1), "hipEventCreate");
    check_hip(hipMalloc(&dX, hx.size()), "hipMalloc(tune X)");
    check_hip(hipMalloc(reinterpret_cast<void**>(&dSpans), sizeof(Span)), "hipMalloc(tune span)");
    check_hip(hipMalloc(reinterpret_cast<void**>(&dOut), nq * sizeof(double)),
              "hipMalloc(tune out)");
    check_hip(hipMalloc(reinterpret_cast<void**>(&dCounts), kCounts * sizeof(int)), "hipMalloc");
    check_hip(hipMalloc(reinterpret_cast<void**>(&dCounter), sizeof(unsigned)), "hipMalloc");
    check_hip(hipMalloc(reinterpret_cast<void**>(&dPart),
                        std::max(kTunePartDoubles, ntiles * tile) * sizeof(double)),
              "hipMalloc(tune partials)");
    check_hip(hipMemcpy(dX, hx.data(), hx.size(), hipMemcpyHostToDevice), "upload tune X");
    double best = -1.0;
    for (int c : cands) {
      const size_t nchunks = (m.Npad + c - 1) / c;
      const size_t nitems = ntiles * nchunks;
      const Span sp = make_span(m.table_index, m.D, m.Npad, tile, static_cast<size_t>(c), nq, nq,
                                0, 0, 0, 0, 0, m.lo);
      const int counts[kCounts] = {static_cast<int>(nitems), static_cast<int>(nq), 1, 0, 1};
      check_hip(hipMemcpy(dSpans, &sp, sizeof(Span), hipMemcpyHostToDevice), "upload tune span");
      check_hip(hipMemcpy(dCounts, counts, sizeof(counts), hipMemcpyHostToDevice),
                "upload tune counts");
      float tmin = -1.0f;
      for (int rep = 0; rep < 4; ++rep) {  // the first run warms up
        check_hip(hipEventRecord(e0, stream), "hipEventRecord");
        enqueue_kernels(stream, dSpans, dCounts, dCounter, dX, dPart, dOut);
        check_hip(hipGetLastError(), "tune launch");
        check_hip(hipEventRecord(e1, stream), "hipEventRecord");
        check_hip(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.0f;
        check_hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        if (rep > 0 && (tmin < 0.0f || ms < tmin)) tmin = ms;
      }
      const double cost = tmin * 1e-3 / (static_cast<double>(nq) * static_cast<double>(m.Npad));
      if (best < 0.0 || cost < best) {
        best = cost;
        res.chunk = c;
        res.cost = cost;
      }
    }
  } catch (...) {
    release();
    throw;
  }
  release();
  std::lock_guard<std::mutex> lk(mu_);
  tuned_[tune_key(m.basis, m.D)] = res;
  save_cache();
}

// Subspaces per work item when the subspace algorithm runs untuned.
constexpr int kSubDefaultChunks = 32;
// synthetic load of the algorithm choice on a batching server (x kTuneQueries)
constexpr size_t kChoiceBatchCopies = 16;

// Chooses between the streaming kernel and the subspace kernel for model m
// (which has a subspace layout) and the subspaces per work item: with the
// autotuner, both are timed on this model with synthetic uniform queries (the
// same load as the shape tuner; streaming at its tuned chunk, the subspace
// kernel at nsub / 2^k subspaces per item) and the faster kept; without it,
// the configured algorithm decides (auto = streaming) and the subspace kernel
// takes ceil(nsub / 32) subspaces per item. The caller holds a reference that
// keeps the model resident.
void Server::choose_algorithm(Model& m) {
  const int want = cfg_.algorithm;
  if (!m.sr || want == kServeStreaming) return;
  const int untuned = (m.nsub + kSubDefaultChunks - 1) / kSubDefaultChunks;
  if (!cfg_.autotune) {
    if (want == kServeSubspace) {
      m.sub = true;
      m.sub_chunk = untuned;
    }
    return;
  }
  const bool f64 = cfg_.precision == Precision::FP64;
  const size_t esz = f64 ? 8 : 4;
  const size_t D = static_cast<size_t>(m.D);
  // a lone request when every launch holds one; otherwise enough queries that
  // even one subspace chunk per query tile fills the device, as the many
  // models of a saturated batch do together
  const size_t nq = kTuneQueries * (cfg_.max_requests_per_batch == 1 ? 1 : kChoiceBatchCopies);
  // queries with low parts (FP32): rows [D, 2D)
  std::vector<char> hx(nq * D * esz * 2);
  uint64_t st = 0xC0FFEE00ull + static_cast<uint64_t>(D);
  for (size_t d = 0; d < D; ++d)
    for (size_t i = 0; i < nq; ++i) {
      const double v = static_cast<double>(splitmix64(st) >> 11) * 0x1.0p-53;
      if (f64) {
        reinterpret_cast<double*>(hx.data())[d * nq + i] = v;
      } else {
        const float hi = static_cast<float>(v);
        reinterpret_cast<float*>(hx.data())[d * nq + i] = hi;
        reinterpret_cast<float*>(hx.data())[(D + d) * nq + i] =
            static_cast<float>(v - static_cast<double>(hi));
      }
    }
  size_t stream_chunk;
  {
    std::lock_guard<std::mutex> lk(mu_);
    stream_chunk = chunk_for(m);
  }
  const size_t stile = static_cast<size_t>(kBlock) * serve_k_rt(cfg_.precision, m.D);
  const size_t s_items = (nq / stile) * ((m.Npad + stream_chunk - 1) / stream_chunk);
  const size_t partcap = std::max(kTunePartDoubles, s_items * stile);
  check_hip(hipSetDevice(cfg_.device), "hipSetDevice");
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  void* dX = nullptr;
  Span* dSpans = nullptr;
  int* dCounts = nullptr;
  unsigned* dCounter = nullptr;
  double* dPart = nullptr;
  double* dOut = nullptr;
  auto release = [&] {
    if (stream) (void)hipStreamSynchronize(stream);
    void* p[] = {dX, dSpans, dCounts, dCounter, dPart, dOut};
    for (void* x : p)
      if (x) (void)hipFree(x);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (stream) (void)hipStreamDestroy(stream);
  };
  try {
    check_hip(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    check_hip(hipEventCreate(&e0), "hipEventCreate");
    check_hip(hipEventCreate(&e1), "hipEventCreate");
    check_hip(hipMalloc(&dX, hx.size()), "hipMalloc(choice X)");
    check_hip(hipMalloc(reinterpret_cast<void**>(&dSpans), sizeof(Span)), "hipMalloc(span)");
    check_hip(hipMalloc(reinterpret_cast<void**>(&dOut), nq * sizeof(double)), "hipMalloc(out)");
    check_hip(hipMalloc(reinterpret_cast<void**>(&dCounts), kCounts * sizeof(int)), "hipMalloc");
    check_hip(hipMalloc(reinterpret_cast<void**>(&dCounter), sizeof(unsigned)), "hipMalloc");
    check_hip(hipMalloc(reinterpret_cast<void**>(&dPart), partcap * sizeof(double)),
              "hipMalloc(choice partials)");
    check_hip(hipMemcpy(dX, hx.data(), hx.size(), hipMemcpyHostToDevice), "upload choice X");
    // best of three timed runs after a warm-up, kernels and reduction
    auto time_ms = [&](const Span& sp, const int* counts) {
      check_hip(hipMemcpy(dSpans, &sp, sizeof(Span), hipMemcpyHostToDevice), "upload span");
      check_hip(hipMemcpy(dCounts, counts, kCounts * sizeof(int), hipMemcpyHostToDevice),
                "upload counts");
      float tmin = -1.0f;
      for (int rep = 0; rep < 4; ++rep) {
        check_hip(hipEventRecord(e0, stream), "hipEventRecord");
        enqueue_kernels(stream, dSpans, dCounts, dCounter, dX, dPart, dOut);
        check_hip(hipGetLastError(), "choice launch");
        check_hip(hipEventRecord(e1, stream), "hipEventRecord");
        check_hip(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.0f;
        check_hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        if (rep > 0 && (tmin < 0.0f || ms < tmin)) tmin = ms;
      }
      return static_cast<double>(tmin);
    };
    double best = -1.0;
    if (want != kServeSubspace) {
      // the streaming kernel at its chunk; its queries' low parts (FP32) are
      // rows [D, 2D) as here when the model uses them
      const Span sp = make_span(m.table_index, m.D, m.Npad, stile, stream_chunk, nq, nq, 0, 0, 0,
                                0, 0, m.lo);
      const int counts[kCounts] = {static_cast<int>(s_items), static_cast<int>(nq), 1, 0, 1};
      best = time_ms(sp, counts);
    }
    int best_chunk = 0;
    for (int k = 1; k <= 256; k *= 2) {
      const int c = (m.nsub + k - 1) / k;
      const size_t nchunks = (static_cast<size_t>(m.nsub) + c - 1) / c;
      if (nq * nchunks > partcap) break;
      if (k > 1 && c == (m.nsub + k / 2 - 1) / (k / 2)) continue;  // same chunk as before
      const size_t items = (nq / kBlock) * nchunks;
      const Span sp = make_span(m.table_index, m.D, static_cast<size_t>(m.nsub), kBlock,
                                static_cast<size_t>(c), nq, nq, 0, 0, 0, 0, 0, !f64);
      const int counts[kCounts] = {0, static_cast<int>(nq), 0, static_cast<int>(items), 1};
      const double t = time_ms(sp, counts);
      if (best < 0.0 || t < best) {
        best = t;
        best_chunk = c;
      }
    }
    if (best_chunk > 0) {
      m.sub = true;
      m.sub_chunk = best_chunk;
    }
  } catch (...) {
    release();
    throw;
  }
  release();
}

// Cache file: one line per tuned shape,
//   sgpp_hip_server_shape <precision> <basis> <dim> <chunk> <cost>
// (precision 0 = FP64, 1 = FP32; basis 0 = Linear, 1 = ModLinear).
void Server::load_cache() {
  if (cfg_.autotune_cache.empty()) return;
  std::ifstream f(cfg_.autotune_cache);
  std::string tag;
  int prec, basis, dim;
  TunedShape t;
  const int my_prec = cfg_.precision == Precision::FP64 ? 0 : 1;
  while (f >> tag >> prec >> basis >> dim >> t.chunk >> t.cost) {
    if (tag != "sgpp_hip_server_shape" || prec != my_prec || dim < 1 || dim > kMaxD) continue;
    if (t.chunk < kTG || t.chunk % kTG) continue;
    tuned_[tune_key(basis ? Basis::ModLinear : Basis::Linear, dim)] = t;
  }
}

// caller holds mu_; keeps the other precision's lines
void Server::save_cache() {
  if (cfg_.autotune_cache.empty()) return;
  const int my_prec = cfg_.precision == Precision::FP64 ? 0 : 1;
  std::vector<std::string> keep;
  {
    std::ifstream f(cfg_.autotune_cache);
    std::string line;
    while (std::getline(f, line)) {
      std::istringstream is(line);
      std::string tag;
      int prec = -1;
      is >> tag >> prec;
      if (tag == "sgpp_hip_server_shape" && prec != my_prec) keep.push_back(line);
    }
  }
  const std::string tmp = cfg_.autotune_cache + ".tmp";
  {
    std::ofstream f(tmp);
    if (!f) return;  // a cache that cannot be written is not an error
    for (const std::string& l : keep) f << l << "\n";
    f.precision(9);
    for (const auto& kv : tuned_)
      f << "sgpp_hip_server_shape " << my_prec << " " << (kv.first >= 64 ? 1 : 0) << " "
        << (kv.first % 64) << " " << kv.second.chunk << " " << kv.second.cost << "\n";
  }
  (void)std::rename(tmp.c_str(), cfg_.autotune_cache.c_str());
}

void Server::loop() {
  (void)hipSetDevice(cfg_.device);
  std::unique_lock<std::mutex> lk(mu_);
  for (;;) {
    bool any_busy = false;
    for (auto& sp : slots_) {
      Slot& s = *sp;
      if (!s.busy) continue;
      const hipError_t q = hipEventQuery(s.ev_done);
      if (q == hipSuccess) {
        finish_batch(s);
      } else if (q == hipErrorNotReady) {
        any_busy = true;
      } else {
        fail_pieces(s, hipGetErrorString(q));
        s.busy = false;
        running_--;
      }
    }
    for (auto& sp : slots_) {
      Slot& s = *sp;
      if (s.busy || queue_.empty()) continue;
      std::vector<Piece> taken;
      if (!form_batch(s, taken)) continue;
      s.pieces = std::move(taken);
      s.busy = true;
      s.inject = inject_error_.exchange(false);
      running_++;
      lk.unlock();
      bool ok = true;
      std::string err;
      try {
        launch_batch(s);
      } catch (const std::exception& e) {
        ok = false;
        err = e.what();
      }
      lk.lock();
      if (!ok) {
        fail_pieces(s, err.c_str());
        s.busy = false;
        running_--;
      } else {
        any_busy = true;
      }
    }
    if (queue_.empty() && running_ == 0) done_cv_.notify_all();
    if (stop_ && !any_busy) break;
    if (any_busy)
      cv_.wait_for(lk, std::chrono::microseconds(20));
    else
      cv_.wait(lk, [&] { return stop_ || !queue_.empty(); });
  }
}

}  // namespace sgpp_hip
