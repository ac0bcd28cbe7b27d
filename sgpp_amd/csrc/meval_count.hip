// Work-counting instantiations of the multi-evaluation kernels: the kernels
// of meval_kernels.hpp with kCount set, which also add each wave's executed
// hat steps and pairs to four counters. The runtime launches them once,
// outside any timed region, to report the work the culled kernels execute.
#include "meval_kernels.hpp"

namespace sgpp_hip {

template <typename T>
CountMultFn<T> count_mult_launcher(int D, bool skip) {
  return DimTable<T, true>::mult(D, skip);
}

template <typename T>
CountMultTFn<T> count_multT_launcher(int D, bool skip) {
  return DimTable<T, true>::multT(D, skip);
}

template CountMultFn<float> count_mult_launcher<float>(int, bool);
template CountMultFn<double> count_mult_launcher<double>(int, bool);
template CountMultTFn<float> count_multT_launcher<float>(int, bool);
template CountMultTFn<double> count_multT_launcher<double>(int, bool);

}  // namespace sgpp_hip
