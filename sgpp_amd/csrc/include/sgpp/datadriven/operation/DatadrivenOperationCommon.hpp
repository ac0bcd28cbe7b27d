// Operator selection, as datadriven/src/sgpp/datadriven/operation/hash/
// DatadrivenOperationCommon.hpp:17-95, with one new sub-type: HIP.
// A HIP operator is requested as (STREAMING, HIP) and tuned by a
// HIPOperationConfiguration carried in the parameter slot.
#pragma once

#include <memory>
#include <string>

namespace sgpp {
namespace base {

// parameter bag slot (base/src/sgpp/base/tools/OperationConfiguration.hpp:18-26)
class OperationConfiguration {
 public:
  virtual ~OperationConfiguration() {}
  virtual OperationConfiguration* clone() const = 0;
};

}  // namespace base

namespace datadriven {

enum class OperationMultipleEvalType {
  DEFAULT,
  STREAMING,
  SUBSPACELINEAR,
  ADAPTIVE,
  MORTONORDER,
  SCALAPACK
};

enum class OperationMultipleEvalSubType {
  DEFAULT,
  SIMPLE,
  COMBINED,
  OCL,
  OCLFASTMP,
  OCLMP,
  OCLMASKMP,
  OCLOPT,
  OCLUNIFIED,
  CUDA,
  HIP  // new: gfx950 streaming kernels of this library
};

enum class OperationMultipleEvalMPIType { NONE, MASTERSLAVE, HPX };

// Arithmetic of the HIP operator.
//   FP64:          every operation in FP64 (plain kernel)
//   FP32_FP64ACC:  basis products and per-tile sums in FP32, per-tile
//                  partials and all reductions accumulated in FP64
enum class HIPPrecision { FP64 = 0, FP32_FP64ACC = 1 };

// Evaluation scheme of the HIP operator.
//   STREAMING: every (data point, grid point) pair, culled by supports
//              (the StreamingDefault / ModMaskStreaming semantics)
//   SUBSPACE:  per data point and subspace the one basis function whose
//              support holds the point, from a dense index table (the
//              SubspaceLinear semantics; FP64, dimension <= 20). The factory
//              selects it for OperationMultipleEvalType::SUBSPACELINEAR.
enum class HIPAlgorithm { STREAMING = 0, SUBSPACE = 1 };

class HIPOperationConfiguration : public base::OperationConfiguration {
 public:
  HIPPrecision precision = HIPPrecision::FP64;
  // also upload FP64 data when precision is FP32_FP64ACC (needed by the
  // mixed-precision solver's FP64 residuals)
  bool keepFP64 = false;
  int device = 0;
  int multGridChunks = 0;   // 0 = automatic
  int multTDataSegs = 0;    // 0 = automatic
  int targetBlocks = 0;     // 0 = automatic
  HIPAlgorithm algorithm = HIPAlgorithm::STREAMING;
  base::OperationConfiguration* clone() const override {
    return new HIPOperationConfiguration(*this);
  }
};

class OperationMultipleEvalConfiguration {
 private:
  OperationMultipleEvalType type = OperationMultipleEvalType::DEFAULT;
  OperationMultipleEvalSubType subType = OperationMultipleEvalSubType::DEFAULT;
  OperationMultipleEvalMPIType mpiType = OperationMultipleEvalMPIType::NONE;
  std::shared_ptr<base::OperationConfiguration> parameters;
  std::string name;

 public:
  OperationMultipleEvalConfiguration(
      OperationMultipleEvalType type = OperationMultipleEvalType::DEFAULT,
      OperationMultipleEvalSubType subType = OperationMultipleEvalSubType::DEFAULT,
      OperationMultipleEvalMPIType mpiType = OperationMultipleEvalMPIType::NONE,
      std::string name = "unnamed")
      : type(type), subType(subType), mpiType(mpiType), name(name) {}
  OperationMultipleEvalConfiguration(OperationMultipleEvalType type,
                                     OperationMultipleEvalSubType subType,
                                     const base::OperationConfiguration& params,
                                     std::string name = "unnamed")
      : type(type), subType(subType), parameters(params.clone()), name(name) {}

  OperationMultipleEvalMPIType getMPIType() { return mpiType; }
  OperationMultipleEvalType getType() { return type; }
  OperationMultipleEvalSubType getSubType() { return subType; }
  std::shared_ptr<base::OperationConfiguration> getParameters() { return parameters; }
  std::string& getName() { return name; }
};

}  // namespace datadriven
}  // namespace sgpp
