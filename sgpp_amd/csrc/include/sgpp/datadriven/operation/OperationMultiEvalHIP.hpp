// HIP implementation of OperationMultipleEval for Linear and ModLinear
// no-boundary grids on gfx950.
//
// Lifecycle as OperationMultiEvalStreaming (datadriven/src/sgpp/datadriven/
// operation/hash/OperationMultiEvalStreaming/OperationMultiEvalStreaming.cpp:13-25, 92-156):
//   construction copies the dataset (padded and transposed on the device),
//   prepare() re-reads the grid (call it after every grid change),
//   mult / multTranspose are synchronous, getDuration() returns the seconds
//   of the last call. An instance is not reentrant (one HIP stream).
// Differences: inputs are validated (alpha.size() == N, result.size() == N
// for multTranspose) and the caller's source vector is never modified.
#pragma once

#include <memory>
#include <string>

#include "sgpp/base/operation/OperationMultipleEval.hpp"
#include "sgpp/datadriven/operation/DatadrivenOperationCommon.hpp"

namespace sgpp_hip {
class DeviceMultiEval;
}

namespace sgpp {
namespace datadriven {

class OperationMultiEvalHIP : public base::OperationMultipleEval {
 public:
  OperationMultiEvalHIP(base::Grid& grid, base::DataMatrix& dataset,
                        const HIPOperationConfiguration& config);
  ~OperationMultiEvalHIP() override;

  void mult(base::DataVector& alpha, base::DataVector& result) override;
  void multTranspose(base::DataVector& source, base::DataVector& result) override;
  // Range forms (OperationMultipleEval.hpp:70-73, 97-100) with the semantics of
  // the streaming OCL operator (OperationMultiEvalStreamingOCLMultiPlatform.hpp:178-242):
  // result is resized to end - start and holds rows [start, end) of B alpha
  // (entries [start, end) of B^T source). The full product is computed and the
  // range kept: one GPU evaluates the whole operator faster than a partition.
  void mult(base::DataVector& alpha, base::DataVector& result, size_t startIndexData,
            size_t endIndexData) override;
  void multTranspose(base::DataVector& source, base::DataVector& result, size_t startIndexGrid,
                     size_t endIndexGrid) override;
  void prepare() override;
  double getDuration() override { return duration_; }
  std::string getImplementationName() override;

  // Same products on device pointers (FP64, inputs already in HBM);
  // asynchronous on this operator's stream.
  void multDevice(const double* d_alpha, double* d_result);
  void multTransposeDevice(const double* d_source, double* d_result);
  void multDevice(const double* d_alpha, double* d_result, HIPPrecision p);
  void multTransposeDevice(const double* d_source, double* d_result, HIPPrecision p);
  void synchronize();
  // kernel-only milliseconds of the last launch sequence (HIP events)
  double getKernelMilliseconds();

  HIPPrecision getPrecision() const { return config_.precision; }
  const HIPOperationConfiguration& getConfiguration() const { return config_; }
  sgpp_hip::DeviceMultiEval& device() { return *dev_; }
  size_t getNumData() const { return numData_; }
  size_t getNumGrid() const { return numGrid_; }

 private:
  HIPOperationConfiguration config_;
  std::unique_ptr<sgpp_hip::DeviceMultiEval> dev_;
  size_t numData_ = 0;
  size_t numGrid_ = 0;
  double duration_ = -1.0;
};

}  // namespace datadriven

namespace op_factory {

// createOperationMultipleEval(grid, dataset, configuration)
// (datadriven/src/sgpp/datadriven/DatadrivenOpFactory.hpp:185-187).
// This library builds the (STREAMING | DEFAULT, HIP) branch; every other
// combination throws factory_exception, as the reference does for
// back-ends it was compiled without (DatadrivenOpFactory.cpp:449-551).
base::OperationMultipleEval* createOperationMultipleEval(
    base::Grid& grid, base::DataMatrix& dataset,
    datadriven::OperationMultipleEvalConfiguration& configuration);

}  // namespace op_factory
}  // namespace sgpp
