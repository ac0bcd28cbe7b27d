// Subspace-indexed evaluation, shared by the operator's subspace kernels
// (subspace_kernels.hip) and the evaluation server (server.hip): the subspace
// record, the per-dimension cell factor, the walk over a subspace's changed
// dimensions, and the host-side construction of records and index tables.
//
// After the reference's SubspaceLinear operator
// (datadriven/src/sgpp/datadriven/operation/hash/OperationMultipleEvalSubspace/
// combined/OperationMultipleEvalSubspaceCombined.hpp:92-204): in each
// subspace W_l (level vector l) the supports of the basis functions tile
// [0,1]^d, so a point x lies in the support of exactly one function of W_l,
// found per dimension from the cell of x; a dense index table per subspace
// gives that function's grid point (or none).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace sgpp_hip {

constexpr int kSubMaxD = 20;
// Subspace records the kernels load at once (one wait on the scalar loads per
// batch); record arrays carry kSubBatch - 1 records of padding.
constexpr int kSubBatch = 4;

// One subspace: offset of its table, the first dimension to recompute (the
// dimensions before it have the previous record's levels), and its levels (a
// 32-byte record: one wide scalar load).
struct alignas(16) SubspaceDesc {
  long long off;
  int recalc;
  unsigned char level[kSubMaxD];
};
static_assert(sizeof(SubspaceDesc) == 32, "one subspace record is 32 bytes");

// Per dimension: the point's cell in a level-l subspace and the factor of its
// basis function there. With t = 2^(l-1) x, the support of the function with
// index i = 2h + 1 is [h, h + 1) in t: h = floor(t) (the table position),
// u = t - h in [0, 1), and
//   Linear hat:            min(2u, 2 - 2u)  (1 - |2^l x - i|, exact in FP64)
//   ModLinear, level 1:    1
//              h = 0:      2 - 2u           (max(0, 2 - 2^l x) on its cell)
//              h = last:   2u               (max(0, 2^l x - i + 1))
//              otherwise:  the hat          (LinearModifiedBasis.hpp:36-53).
// x = 1 belongs to the last cell (h = 2^(l-1) - 1, u = 1), where the hat is 0
// and the ModLinear boundary function 2.
template <bool kMod>
__device__ __forceinline__ double cell_factor(double x, int l, unsigned& half_index) {
  const double t = __builtin_ldexp(x, l - 1);  // exact
  const unsigned last = (1u << (l - 1)) - 1u;
  const unsigned c = static_cast<unsigned>(t);
  const unsigned h = c < last ? c : last;
  half_index = h;
  const double u = t - static_cast<double>(h);  // exact
  const double a = u + u;
  const double b = 2.0 - a;
  if (!kMod) return __builtin_fmin(a, b);
  if (l == 1) return 1.0;
  return h == 0u ? b : (h == last ? a : __builtin_fmin(a, b));
}

// Dimensions [recalc, D) of subspace sd: prod[d + 1] = prod[d] * factor,
// flat[d + 1] = flat[d] * 2^(l - 1) + h; those below recalc keep the values
// of the previous subspace.
template <int D, bool kMod>
__device__ __forceinline__ void walk_subspace(const double (&x)[D], const SubspaceDesc& sd,
                                              int recalc, double (&prod)[D + 1],
                                              unsigned (&flat)[D + 1]) {
#pragma unroll
  for (int d = 0; d < D; ++d) {
    if (d < recalc) continue;
    const int l = sd.level[d];
    if (kMod && l == 1) {
      prod[d + 1] = prod[d];
      flat[d + 1] = flat[d];
      continue;
    }
    unsigned h;
    const double f = cell_factor<kMod>(x[d], l, h);
    prod[d + 1] = prod[d] * f;
    flat[d + 1] = (flat[d] << (l - 1)) + h;
  }
}

// Records and index tables of a grid (host side). desc: one record per
// subspace, sorted by level vector (dimension 0 first), then kSubBatch - 1
// zero records; ti[e]: the grid point of table entry e, or -1; te[j]: the
// table entry of grid point j; total: table entries. Tables are row-major
// over the dimensions (dimension 0 most significant), prod 2^(l_d - 1)
// entries each.
struct SubspaceLayout {
  std::vector<SubspaceDesc> desc;
  std::vector<int32_t> ti;
  std::vector<int64_t> te;
  int64_t total = 0;
  int nsub = 0;
};

// Subspace tables larger than this are refused (entries).
int64_t subspace_table_limit(size_t N);

// Builds the layout; throws input_error for points outside a no-boundary
// grid, d > kSubMaxD, a subspace above 2^30 cells or tables beyond
// subspace_table_limit(N).
SubspaceLayout build_subspace_layout(size_t D, size_t N, const uint32_t* level,
                                     const uint32_t* index);

}  // namespace sgpp_hip
