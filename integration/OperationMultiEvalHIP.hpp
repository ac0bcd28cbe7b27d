// SG++-side adapter of the gfx950 multi-evaluation operator.
//
// Implements sgpp::base::OperationMultipleEval
// (base/src/sgpp/base/operation/hash/OperationMultipleEval.hpp:27-130) on top of
// the C-ABI of libsgpp_hip.so (include/sgpp_hip.h). It is compiled against
// SG++'s own headers: a maintainer places it under
// datadriven/src/sgpp/datadriven/operation/hash/OperationMultiEvalHIP/, adds
// the HIP value to OperationMultipleEvalSubType
// (datadriven/src/sgpp/datadriven/operation/hash/DatadrivenOperationCommon.hpp:26-37)
// and one branch to op_factory::createOperationMultipleEval
// (datadriven/src/sgpp/datadriven/DatadrivenOpFactory.cpp:409-552):
//
//   if (configuration.getSubType() == datadriven::kOperationMultipleEvalSubTypeHIP)
//     return datadriven::createOperationMultiEvalHIP(grid, dataset, configuration);
//
// Parameters follow the OCL precedent
// (OperationMultipleEvalStreamingOCLMultiPlatform/OperatorFactory.cpp:17-55): an
// OperationConfiguration (JSON) in the configuration, with
//   "INTERNAL_PRECISION": "double" (default) or "float" (FP32 products with
//                         FP64 accumulation),
//   "DEVICE":             HIP device ordinal (default 0).
// The factory's SUBSPACELINEAR type with the HIP sub-type selects the
// subspace-indexed kernels (the SubspaceLinear semantics, FP64 only):
//
//   if (configuration.getType() == datadriven::OperationMultipleEvalType::SUBSPACELINEAR &&
//       configuration.getSubType() == datadriven::kOperationMultipleEvalSubTypeHIP)
//     return datadriven::createOperationMultiEvalHIP(grid, dataset, configuration);
#pragma once

#include <sgpp/base/datatypes/DataMatrix.hpp>
#include <sgpp/base/datatypes/DataVector.hpp>
#include <sgpp/base/exception/data_exception.hpp>
#include <sgpp/base/exception/factory_exception.hpp>
#include <sgpp/base/exception/not_implemented_exception.hpp>
#include <sgpp/base/exception/operation_exception.hpp>
#include <sgpp/base/grid/Grid.hpp>
#include <sgpp/base/operation/hash/OperationMultipleEval.hpp>
#include <sgpp/base/tools/OperationConfiguration.hpp>
#include <sgpp/datadriven/operation/hash/DatadrivenOperationCommon.hpp>

#include <sgpp_hip.h>

#include <algorithm>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace sgpp {
namespace datadriven {

// The sub-type value a maintainer appends to OperationMultipleEvalSubType.
// Until then a configuration names it by this value.
constexpr OperationMultipleEvalSubType kOperationMultipleEvalSubTypeHIP =
    static_cast<OperationMultipleEvalSubType>(64);

class OperationMultiEvalHIP : public base::OperationMultipleEval {
 public:
  // Copies the dataset into HBM (the streaming constructor's copy, pad and
  // transpose, OperationMultiEvalStreaming.cpp:13-25) and reads the grid.
  OperationMultiEvalHIP(base::Grid& grid, base::DataMatrix& dataset, int basis, int precision,
                        int device, int algorithm = SGPP_HIP_ALGO_STREAMING)
      : OperationMultipleEval(grid, dataset), algorithm_(algorithm) {
    sgpp_hip_grid_desc desc{basis, grid.getDimension()};
    sgpp_hip_params p{precision, /*keep_fp64=*/0, device, 0, 0, 0, algorithm};
    check(sgpp_hip_meval_create(&desc, dataset.getPointer(), dataset.getNrows(),
                                dataset.getNcols(), &p, &h_));
    try {
      prepare();
    } catch (...) {
      sgpp_hip_meval_destroy(h_);
      throw;
    }
  }
  ~OperationMultiEvalHIP() override { sgpp_hip_meval_destroy(h_); }
  OperationMultiEvalHIP(const OperationMultiEvalHIP&) = delete;
  OperationMultiEvalHIP& operator=(const OperationMultiEvalHIP&) = delete;

  // OperationMultiEvalStreaming.cpp:181: the grid's level/index arrays, in its
  // sequence order
  void prepare() override {
    base::GridStorage& st = grid.getStorage();
    const size_t n = st.getSize(), d = st.getDimension();
    std::vector<uint32_t> lv(n * d), ix(n * d);
    for (size_t s = 0; s < n; ++s)
      for (size_t k = 0; k < d; ++k) {
        base::HashGridPoint::level_type l;
        base::HashGridPoint::index_type i;
        st.getPoint(s).get(k, l, i);
        lv[s * d + k] = static_cast<uint32_t>(l);
        ix[s * d + k] = static_cast<uint32_t>(i);
      }
    check(sgpp_hip_meval_prepare(h_, lv.data(), ix.data(), n));
  }

  // result = B alpha; alpha holds one value per grid point, result ends with
  // one value per data point (as OperationMultiEvalStreaming::mult leaves it)
  void mult(base::DataVector& alpha, base::DataVector& result) override {
    if (alpha.getSize() != sgpp_hip_meval_num_grid(h_))
      throw base::data_exception("OperationMultiEvalHIP::mult: alpha size != grid size");
    result.resize(sgpp_hip_meval_num_data(h_));
    check(sgpp_hip_meval_mult(h_, alpha.getPointer(), result.getPointer()));
  }

  // rows [start, end) of B alpha (the range form of OperationMultipleEval.hpp:70-73):
  // the full product, of which the range is kept
  void mult(base::DataVector& alpha, base::DataVector& result, const size_t start,
            const size_t end) override {
    const size_t M = sgpp_hip_meval_num_data(h_);
    if (start > end || end > M)
      throw base::data_exception("OperationMultiEvalHIP::mult: invalid data range");
    base::DataVector full(M);
    mult(alpha, full);
    result.resize(end - start);
    std::copy(full.getPointer() + start, full.getPointer() + end, result.getPointer());
  }

  // result = B^T source; source holds at least one value per data point (the
  // first M are used; it is never modified), result ends with one value per
  // grid point
  void multTranspose(base::DataVector& source, base::DataVector& result) override {
    if (source.getSize() < sgpp_hip_meval_num_data(h_))
      throw base::data_exception(
          "OperationMultiEvalHIP::multTranspose: source shorter than the dataset");
    result.resize(sgpp_hip_meval_num_grid(h_));
    check(sgpp_hip_meval_mult_transpose(h_, source.getPointer(), result.getPointer()));
  }

  // entries [start, end) of B^T source (OperationMultipleEval.hpp:97-100)
  void multTranspose(base::DataVector& source, base::DataVector& result, const size_t start,
                     const size_t end) override {
    const size_t N = sgpp_hip_meval_num_grid(h_);
    if (start > end || end > N)
      throw base::data_exception("OperationMultiEvalHIP::multTranspose: invalid grid range");
    base::DataVector full(N);
    multTranspose(source, full);
    result.resize(end - start);
    std::copy(full.getPointer() + start, full.getPointer() + end, result.getPointer());
  }

  double getDuration() override { return sgpp_hip_meval_duration(h_); }

  std::string getImplementationName() override {
    return algorithm_ == SGPP_HIP_ALGO_SUBSPACE ? "HIP_SUBSPACE" : "HIP_STREAMING";
  }

 private:
  // status codes back to SG++'s exception types (sgpp_hip.h)
  static void check(sgpp_hip_status s) {
    switch (s) {
      case SGPP_HIP_OK:
        return;
      case SGPP_HIP_ERR_FACTORY:
        throw base::factory_exception(sgpp_hip_last_error());
      case SGPP_HIP_ERR_INVALID_ARGUMENT:
        throw base::data_exception(sgpp_hip_last_error());
      case SGPP_HIP_ERR_NOT_IMPLEMENTED:
        throw base::not_implemented_exception(sgpp_hip_last_error());
      default:
        throw base::operation_exception(sgpp_hip_last_error());
    }
  }

  sgpp_hip_meval* h_ = nullptr;
  int algorithm_;
};

// The factory branch: Linear and ModLinear grids; any other grid type is a
// factory_exception, as the factory treats back-ends it was built without.
inline base::OperationMultipleEval* createOperationMultiEvalHIP(
    base::Grid& grid, base::DataMatrix& dataset,
    OperationMultipleEvalConfiguration& configuration) {
  int basis;
  if (grid.getType() == base::GridType::Linear)
    basis = SGPP_HIP_BASIS_LINEAR;
  else if (grid.getType() == base::GridType::ModLinear)
    basis = SGPP_HIP_BASIS_MODLINEAR;
  else
    throw base::factory_exception(
        "OperationMultiEvalHIP: only Linear and ModLinear grids are supported");
  int precision = SGPP_HIP_PREC_FP64;
  int device = 0;
  std::shared_ptr<base::OperationConfiguration> params = configuration.getParameters();
  if (params) {
    if (params->contains("INTERNAL_PRECISION")) {
      const std::string& p = (*params)["INTERNAL_PRECISION"].get();
      if (p == "float")
        precision = SGPP_HIP_PREC_FP32_FP64ACC;
      else if (p != "double")
        throw base::factory_exception(
            "OperationMultiEvalHIP: INTERNAL_PRECISION must be \"double\" or \"float\"");
    }
    if (params->contains("DEVICE")) device = static_cast<int>((*params)["DEVICE"].getInt());
  }
  if (configuration.getType() == OperationMultipleEvalType::SUBSPACELINEAR) {
    if (precision != SGPP_HIP_PREC_FP64)
      throw base::factory_exception(
          "OperationMultiEvalHIP: the SUBSPACELINEAR kernels compute in double precision");
    return new OperationMultiEvalHIP(grid, dataset, basis, precision, device,
                                     SGPP_HIP_ALGO_SUBSPACE);
  }
  return new OperationMultiEvalHIP(grid, dataset, basis, precision, device);
}

}  // namespace datadriven
}  // namespace sgpp
