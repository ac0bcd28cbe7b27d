// Subspace-indexed evaluation: the second algorithm of the operator, after
// the reference's SubspaceLinear operator
// (datadriven/src/sgpp/datadriven/operation/hash/OperationMultipleEvalSubspace/
// combined/OperationMultipleEvalSubspaceCombined.hpp:92-204, inner loop
// ..._uncachedMultTransposeInner.cpp:13-173).
//
// A sparse grid is a union of subspaces W_l (level vector l); in each W_l the
// supports of the basis functions tile [0,1]^d, so a data point x lies in the
// support of exactly one function of W_l: per dimension the cell
// c = floor(2^l x), index i = c | 1 (made odd). Instead of testing grid
// points against data points, every data point visits every subspace, finds
// its one function there and looks it up in the subspace's dense index table
// (the grid point, or none when the adaptive grid lacks it).
//
// Layout: subspaces sorted by level vector (dimension 0 first), so
// consecutive ones share leading levels; a subspace records the first
// dimension where it differs from its predecessor, and the per-dimension
// partial products and flat table indices before that dimension are reused
// (the reference's "nextIterationToRecalc"). Tables are row-major over the
// dimensions (dimension 0 most significant), Pi 2^(l_d - 1) entries each.
//   mult:          r(x) = sum_W w[table entry of x in W] * phi(x); the
//                  weights are scattered into table order first (0 where the
//                  grid has no point).
//   multTranspose: a[entry] += y(x) * phi(x). Blocks = data segments (x) x
//                  chunks of consecutive subspaces whose tables fit in LDS
//                  (y); a block sums its segment's contributions to its chunk
//                  in an LDS copy of the chunk's tables (LDS FP64 atomic
//                  adds; a wave whose points share one entry adds once, after
//                  a DPP sum) and stores the chunk into its segment's partial
//                  row; the S partial rows are summed in a fixed order while
//                  gathering to grid order. A subspace whose table alone
//                  exceeds the LDS chunk is a chunk of its own that adds into
//                  the partial row with global FP64 atomics. The order of
//                  atomic adds, and so the last bits of the sums, varies
//                  between runs.
// Lanes own data points (data in Morton order, as in the streaming kernels).
// FP64 throughout.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <vector>

#include "cull_device.hpp"
#include "meval_device.hpp"
#include "subspace_device.hpp"

namespace sgpp_hip {

namespace {

constexpr int kSubBlock = 256;
// LDS table entries of one multTranspose chunk (32 KiB of FP64)
constexpr int kSubLds = 4096;

// A multTranspose chunk: subspaces [s0, s1), table entries [off0, off0 + n);
// big: one subspace whose table exceeds kSubLds (global atomics)
struct alignas(16) SubChunk {
  int s0, s1;
  long long off0;
  int n, big;
  int pad[2];
};
static_assert(sizeof(SubChunk) == 32, "one chunk record is 32 bytes");

// mult: lanes own data points X[:, m] (stored order), out[m] = r(x_m).
template <int D, bool kMod>
__global__ __launch_bounds__(kSubBlock) void subspace_mult_kernel(
    const double* __restrict__ X, int64_t ldx, const SubspaceDesc* __restrict__ desc, int S,
    const double* __restrict__ TW, double* __restrict__ out) {
  const int64_t m = static_cast<int64_t>(blockIdx.x) * kSubBlock + threadIdx.x;
  double x[D];
#pragma unroll
  for (int d = 0; d < D; ++d) x[d] = X[d * ldx + m];
  double prod[D + 1];
  unsigned flat[D + 1];
  prod[0] = 1.0;
  flat[0] = 0u;
  double acc = 0.0;
  // a batch's weights are gathered before any is used, and used after the
  // next batch's gathers are issued (the gathers' latency overlaps the
  // next batch's products)
  double wp[kSubBatch], pp[kSubBatch];
#pragma unroll
  for (int k = 0; k < kSubBatch; ++k) wp[k] = pp[k] = 0.0;
  for (int s = 0; s < S; s += kSubBatch) {
    SubspaceDesc sd[kSubBatch];
#pragma unroll
    for (int k = 0; k < kSubBatch; ++k) sd[k] = desc[s + k];
    double wn[kSubBatch], pn[kSubBatch];
#pragma unroll
    for (int k = 0; k < kSubBatch; ++k) {
      wn[k] = pn[k] = 0.0;
      if (s + k < S) {
        walk_subspace<D, kMod>(x, sd[k], sd[k].recalc, prod, flat);
        wn[k] = TW[sd[k].off + flat[D]];
        pn[k] = prod[D];
      }
    }
#pragma unroll
    for (int k = 0; k < kSubBatch; ++k) {
      acc = __builtin_fma(wp[k], pp[k], acc);
      wp[k] = wn[k];
      pp[k] = pn[k];
    }
  }
#pragma unroll
  for (int k = 0; k < kSubBatch; ++k) acc = __builtin_fma(wp[k], pp[k], acc);
  out[m] = acc;
}

// multTranspose, one block per (data segment, chunk): the segment's rows
// [m0, m0 + seg_len) visit the chunk's subspaces; the chunk's sums go to
// part[segment][off0 .. off0 + n).
template <int D, bool kMod>
__global__ __launch_bounds__(kSubBlock) void subspace_multT_kernel(
    const double* __restrict__ X, int64_t ldx, const double* __restrict__ Y,
    const SubspaceDesc* __restrict__ desc, const SubChunk* __restrict__ chunks, int64_t seg_len,
    int64_t mpad, double* __restrict__ part, int64_t ntable) {
  __shared__ double acc[kSubLds];
  const SubChunk ch = chunks[blockIdx.y];
  const int64_t m0 = static_cast<int64_t>(blockIdx.x) * seg_len;
  // seg_len and mpad are multiples of the block size
  const int64_t m1 = m0 + seg_len < mpad ? m0 + seg_len : mpad;
  double* __restrict__ prow = part + static_cast<int64_t>(blockIdx.x) * ntable + ch.off0;
  if (ch.big) {
    for (int i = threadIdx.x; i < ch.n; i += kSubBlock) prow[i] = 0.0;
    __threadfence();  // the zeros reach L2 before this block's atomics
  } else {
    for (int i = threadIdx.x; i < ch.n; i += kSubBlock) acc[i] = 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  for (int64_t mb = m0; mb < m1; mb += kSubBlock) {
    const int64_t m = mb + threadIdx.x;
    double x[D];
#pragma unroll
    for (int d = 0; d < D; ++d) x[d] = X[d * ldx + m];
    const double y = Y[m];  // 0 for padding rows
    double prod[D + 1];
    unsigned flat[D + 1];
    prod[0] = 1.0;
    flat[0] = 0u;
    for (int sb = ch.s0; sb < ch.s1; sb += kSubBatch) {
      SubspaceDesc sdb[kSubBatch];
#pragma unroll
      for (int k = 0; k < kSubBatch; ++k) sdb[k] = desc[sb + k];
#pragma unroll
      for (int k = 0; k < kSubBatch; ++k) {
      const int s = sb + k;
      if (s >= ch.s1) break;
      const SubspaceDesc& sd = sdb[k];
      // a chunk starts from scratch
      walk_subspace<D, kMod>(x, sd, s == ch.s0 ? 0 : sd.recalc, prod, flat);
      const double v = y * prod[D];
      const unsigned e = static_cast<unsigned>(sd.off - ch.off0) + flat[D];
      const unsigned e0 =
          static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(e)));
      const bool one = __ballot(e != e0) == 0ull;
      // (separate LDS and global forms: a pointer that may be either would
      // compile to flat atomics)
      if (one) {
        // the whole wave is in one entry: one add of the wave's sum
        const double sum = lane63(sum_to_lane63(v));
        if (lane == 0 && sum != 0.0) {
          if (ch.big)
            atomicAdd(&prow[e0], sum);
          else
            atomicAdd(&acc[e0], sum);
        }
      } else if (v != 0.0) {
        if (ch.big)
          atomicAdd(&prow[e], v);
        else
          atomicAdd(&acc[e], v);
      }
      }
    }
  }
  if (!ch.big) {
    __syncthreads();
    for (int i = threadIdx.x; i < ch.n; i += kSubBlock) prow[i] = acc[i];
  }
}

// Y[m] = y[dperm[m]] for stored row m, 0 for the padding rows
__global__ void subspace_prep_y_kernel(const double* __restrict__ y,
                                       const int32_t* __restrict__ dperm, double* __restrict__ Y,
                                       int64_t n, int64_t npad) {
  const int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (m < npad) Y[m] = m < n ? y[dperm[m]] : 0.0;
}

// TW[e] = alpha[TI[e]] (0 where TI[e] < 0)
__global__ void subspace_weights_kernel(const double* __restrict__ alpha,
                                        const int32_t* __restrict__ TI, double* __restrict__ TW,
                                        int64_t n) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < n) TW[e] = TI[e] >= 0 ? alpha[TI[e]] : 0.0;
}

// a[j] = sum_{q < nseg} part[q][TE[j]] (TE[j]: the table entry of grid point
// j), summed in the fixed order q = 0, 1, ...
__global__ void subspace_gather_kernel(const double* __restrict__ part, int64_t ntable, int nseg,
                                       const int64_t* __restrict__ TE, double* __restrict__ a,
                                       int64_t n) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double* p = part + TE[j];
  double s = 0.0;
  for (int q = 0; q < nseg; ++q) s += p[static_cast<int64_t>(q) * ntable];
  a[j] = s;
}

// out[dperm[m]] = v[m] for the M data rows (stored order back to the caller's)
__global__ void subspace_scatter_rows_kernel(const double* __restrict__ v,
                                             const int32_t* __restrict__ dperm,
                                             double* __restrict__ out, int64_t n) {
  const int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (m < n) out[dperm[m]] = v[m];
}

unsigned blocks(int64_t n, int b) { return static_cast<unsigned>((n + b - 1) / b); }

template <int D>
void launch_sub_mult(bool mod, const double* X, int64_t ldx, const SubspaceDesc* desc, int S,
                     const double* TW, double* out, int64_t Mpad, hipStream_t s) {
  clear_launch_error();
  if (mod)
    hipLaunchKernelGGL((subspace_mult_kernel<D, true>), dim3(blocks(Mpad, kSubBlock)),
                       dim3(kSubBlock), 0, s, X, ldx, desc, S, TW, out);
  else
    hipLaunchKernelGGL((subspace_mult_kernel<D, false>), dim3(blocks(Mpad, kSubBlock)),
                       dim3(kSubBlock), 0, s, X, ldx, desc, S, TW, out);
}

template <int D>
void launch_sub_multT(bool mod, const double* X, int64_t ldx, const double* Y,
                      const SubspaceDesc* desc, const SubChunk* chunks, int nchunk, int nseg,
                      int64_t seg_len, double* part, int64_t ntable, hipStream_t s) {
  const int64_t mpad = ldx;  // X holds the padded rows, dimension-major
  clear_launch_error();
  const dim3 grid(static_cast<unsigned>(nseg), static_cast<unsigned>(nchunk));
  if (mod)
    hipLaunchKernelGGL((subspace_multT_kernel<D, true>), grid, dim3(kSubBlock), 0, s, X, ldx, Y,
                       desc, chunks, seg_len, mpad, part, ntable);
  else
    hipLaunchKernelGGL((subspace_multT_kernel<D, false>), grid, dim3(kSubBlock), 0, s, X, ldx, Y,
                       desc, chunks, seg_len, mpad, part, ntable);
}

using SubMultFn = void (*)(bool, const double*, int64_t, const SubspaceDesc*, int, const double*,
                           double*, int64_t, hipStream_t);
using SubMultTFn = void (*)(bool, const double*, int64_t, const double*, const SubspaceDesc*,
                            const SubChunk*, int, int, int64_t, double*, int64_t, hipStream_t);

template <int... Ds>
struct SubTable {
  static SubMultFn mult(int D) {
    SubMultFn f = nullptr;
    ((D == Ds ? (f = &launch_sub_mult<Ds>, 0) : 0), ...);
    return f;
  }
  static SubMultTFn multT(int D) {
    SubMultTFn f = nullptr;
    ((D == Ds ? (f = &launch_sub_multT<Ds>, 0) : 0), ...);
    return f;
  }
};
using SubDims = SubTable<1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20>;

}  // namespace

int64_t subspace_table_limit(size_t N) {
  return std::max<int64_t>(int64_t(1) << 26, 64 * static_cast<int64_t>(N));
}

SubspaceLayout build_subspace_layout(size_t D, size_t N, const uint32_t* level,
                                     const uint32_t* index) {
  if (D > static_cast<size_t>(kSubMaxD))
    throw input_error("the subspace algorithm supports dimensions up to 20");
  // group the points by level vector; std::map orders level vectors
  // lexicographically, dimension 0 first
  std::map<std::vector<uint8_t>, std::vector<uint32_t>> groups;
  std::vector<uint8_t> lv(D);
  for (size_t j = 0; j < N; ++j) {
    for (size_t d = 0; d < D; ++d) {
      const uint32_t l = level[j * D + d], i = index[j * D + d];
      if (l < 1 || l > 30 || (i & 1u) == 0 || i >= (1u << l))
        throw input_error("invalid grid point (level/index out of range for a no-boundary grid)");
      lv[d] = static_cast<uint8_t>(l);
    }
    groups[lv].push_back(static_cast<uint32_t>(j));
  }
  SubspaceLayout lay;
  lay.desc.reserve(groups.size() + kSubBatch - 1);
  lay.te.resize(N);
  const std::vector<uint8_t>* prev = nullptr;
  for (const auto& g : groups) {
    int levelsum = 0;
    for (size_t d = 0; d < D; ++d) levelsum += g.first[d] - 1;
    if (levelsum > 30) throw input_error("a subspace's index table would exceed 2^30 entries");
    SubspaceDesc sd{};
    sd.off = lay.total;
    int r = 0;
    if (prev)
      while (r < static_cast<int>(D) && (*prev)[r] == g.first[r]) ++r;
    sd.recalc = prev ? r : 0;
    for (size_t d = 0; d < D; ++d) sd.level[d] = g.first[d];
    lay.desc.push_back(sd);
    lay.total += int64_t(1) << levelsum;
    prev = &g.first;
  }
  // a table entry per cell of every subspace; refuse tables far larger than the grid
  if (lay.total > subspace_table_limit(N))
    throw input_error("the subspace index tables of this grid would hold more than 64 entries "
                      "per grid point");
  lay.ti.assign(static_cast<size_t>(lay.total), -1);
  size_t s = 0;
  for (const auto& g : groups) {
    for (uint32_t j : g.second) {
      int64_t flat = 0;
      for (size_t d = 0; d < D; ++d) flat = (flat << (g.first[d] - 1)) + (index[j * D + d] >> 1);
      lay.ti[static_cast<size_t>(lay.desc[s].off + flat)] = static_cast<int32_t>(j);
      lay.te[j] = lay.desc[s].off + flat;
    }
    ++s;
  }
  lay.nsub = static_cast<int>(lay.desc.size());
  lay.desc.resize(lay.desc.size() + kSubBatch - 1, SubspaceDesc{});  // batch padding
  return lay;
}

void DeviceMultiEval::set_grid_subspace(const uint32_t* level, const uint32_t* index, size_t N) {
  check_hip(hipSetDevice(device_), "hipSetDevice");
  SubspaceLayout lay = build_subspace_layout(D_, N, level, index);
  std::vector<SubspaceDesc>& desc = lay.desc;
  const int64_t total = lay.total;
  const int nsub = lay.nsub;
  // multTranspose chunks: runs of consecutive subspaces whose tables fit in
  // kSubLds entries; a larger table is a chunk of its own
  std::vector<SubChunk> chunks;
  for (int q = 0; q < nsub; ++q) {
    const int64_t n = (q + 1 < nsub ? desc[q + 1].off : total) - desc[q].off;
    if (n > kSubLds) {
      chunks.push_back(SubChunk{q, q + 1, desc[q].off, static_cast<int>(n), 1, {0, 0}});
      continue;
    }
    if (chunks.empty() || chunks.back().big || chunks.back().n + n > kSubLds)
      chunks.push_back(SubChunk{q, q, desc[q].off, 0, 0, {0, 0}});
    chunks.back().s1 = q + 1;
    chunks.back().n += static_cast<int>(n);
  }
  if (chunks.size() > 65535) throw input_error("too many subspace chunks for one launch");
  check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  N_ = N;
  Npad_ = N;
  nchunk_ = static_cast<int>(chunks.size());
  nsub_ = nsub;
  ntable_ = total;
  // evaluated dimensions per data row: those not shared with the previous
  // subspace (multTranspose: nor at a chunk's start), ModLinear level 1 excluded
  const bool mod = basis_ == Basis::ModLinear;
  auto evaluated = [&](const SubspaceDesc& sd, int from) {
    uint64_t n = 0;
    for (int d = from; d < static_cast<int>(D_); ++d) n += (mod && sd.level[d] == 1) ? 0 : 1;
    return n;
  };
  sub_recompute_ = sub_recompute_t_ = 0;
  for (int q = 0; q < nsub; ++q) sub_recompute_ += evaluated(desc[q], desc[q].recalc);
  for (const SubChunk& c : chunks)
    for (int q = c.s0; q < c.s1; ++q)
      sub_recompute_t_ += evaluated(desc[q], q == c.s0 ? 0 : desc[q].recalc);
  ensure(&sub_desc_, &sub_desc_cap_, desc.size() * sizeof(SubspaceDesc));
  check_hip(hipMemcpy(sub_desc_, desc.data(), desc.size() * sizeof(SubspaceDesc),
                      hipMemcpyHostToDevice),
            "upload subspaces");
  ensure(&sub_ti_, &sub_ti_cap_, lay.ti.size() * sizeof(int32_t));
  check_hip(hipMemcpy(sub_ti_, lay.ti.data(), lay.ti.size() * sizeof(int32_t),
                      hipMemcpyHostToDevice),
            "upload subspace tables");
  ensure(&sub_te_, &sub_te_cap_, lay.te.size() * sizeof(int64_t));
  check_hip(hipMemcpy(sub_te_, lay.te.data(), lay.te.size() * sizeof(int64_t),
                      hipMemcpyHostToDevice),
            "upload grid point entries");
  ensure(&sub_chunk_, &sub_chunk_cap_, chunks.size() * sizeof(SubChunk));
  check_hip(hipMemcpy(sub_chunk_, chunks.data(), chunks.size() * sizeof(SubChunk),
                      hipMemcpyHostToDevice),
            "upload subspace chunks");
}

void DeviceMultiEval::run_subspace_mult(const double* d_alpha, double* d_result, hipEvent_t* tev) {
  hipStream_t s = stream_;
  ensure(&sub_tw_, &sub_tw_cap_, static_cast<size_t>(ntable_) * sizeof(double));
  ensure(&part_, &part_cap_, Mpad_ * sizeof(double));
  double* TW = static_cast<double*>(sub_tw_);
  double* P = static_cast<double*>(part_);
  clear_launch_error();
  hipLaunchKernelGGL(subspace_weights_kernel, dim3(blocks(ntable_, 256)), dim3(256), 0, s, d_alpha,
                     static_cast<const int32_t*>(sub_ti_), TW, ntable_);
  check_hip(hipGetLastError(), "subspace weights");
  if (tev) check_hip(hipEventRecord(tev[0], s), "hipEventRecord");
  SubDims::mult(static_cast<int>(D_))(basis_ == Basis::ModLinear, static_cast<const double*>(X64_),
                                      static_cast<int64_t>(Mpad_),
                                      static_cast<const SubspaceDesc*>(sub_desc_), nsub_, TW, P,
                                      static_cast<int64_t>(Mpad_), s);
  check_hip(hipGetLastError(), "subspace mult kernel");
  if (tev) check_hip(hipEventRecord(tev[1], s), "hipEventRecord");
  clear_launch_error();
  hipLaunchKernelGGL(subspace_scatter_rows_kernel, dim3(blocks(M_, 256)), dim3(256), 0, s, P,
                     static_cast<const int32_t*>(dperm_), d_result, static_cast<int64_t>(M_));
  check_hip(hipGetLastError(), "subspace scatter");
}

void DeviceMultiEval::run_subspace_multT(const double* d_y, double* d_result, hipEvent_t* tev) {
  hipStream_t s = stream_;
  ensure(&Y_, &Y_cap_, Mpad_ * sizeof(double));
  // data segments: about target_blocks blocks in all, whole 256-row steps,
  // partial rows (one table copy per segment) held under 512 MiB
  const int64_t steps = static_cast<int64_t>(Mpad_) / kSubBlock;
  const int64_t target = cfg_.target_blocks > 0 ? cfg_.target_blocks : 8 * num_cus_;
  int64_t nseg = std::max<int64_t>(1, (target + nchunk_ - 1) / nchunk_);
  if (cfg_.multT_data_segs > 0) nseg = cfg_.multT_data_segs;
  nseg = std::min<int64_t>(nseg, steps);
  nseg = std::min<int64_t>(nseg, std::max<int64_t>(1, (int64_t(512) << 20) / (8 * ntable_)));
  const int64_t len = (steps + nseg - 1) / nseg * kSubBlock;  // rows per segment
  nseg = (static_cast<int64_t>(Mpad_) + len - 1) / len;       // the last may be shorter
  ensure(&part_, &part_cap_, static_cast<size_t>(nseg * ntable_) * sizeof(double));
  clear_launch_error();
  hipLaunchKernelGGL(subspace_prep_y_kernel, dim3(blocks(Mpad_, 256)), dim3(256), 0, s, d_y,
                     static_cast<const int32_t*>(dperm_), static_cast<double*>(Y_),
                     static_cast<int64_t>(M_), static_cast<int64_t>(Mpad_));
  check_hip(hipGetLastError(), "subspace prep");
  if (tev) check_hip(hipEventRecord(tev[0], s), "hipEventRecord");
  SubDims::multT(static_cast<int>(D_))(
      basis_ == Basis::ModLinear, static_cast<const double*>(X64_), static_cast<int64_t>(Mpad_),
      static_cast<const double*>(Y_), static_cast<const SubspaceDesc*>(sub_desc_),
      static_cast<const SubChunk*>(sub_chunk_), nchunk_, static_cast<int>(nseg), len,
      static_cast<double*>(part_), ntable_, s);
  check_hip(hipGetLastError(), "subspace multTranspose kernel");
  if (tev) check_hip(hipEventRecord(tev[1], s), "hipEventRecord");
  clear_launch_error();
  hipLaunchKernelGGL(subspace_gather_kernel, dim3(blocks(N_, 256)), dim3(256), 0, s,
                     static_cast<const double*>(part_), ntable_, static_cast<int>(nseg),
                     static_cast<const int64_t*>(sub_te_), d_result, static_cast<int64_t>(N_));
  check_hip(hipGetLastError(), "subspace gather");
  last_S_ = static_cast<int>(nseg);
}

}  // namespace sgpp_hip
