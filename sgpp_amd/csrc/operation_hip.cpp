// OperationMultiEvalHIP and the factory branch that creates it.

#include <chrono>
#include <cstring>
#include <vector>

#include "meval_device.hpp"
#include "sgpp/datadriven/operation/OperationMultiEvalHIP.hpp"

namespace sgpp {
namespace datadriven {

namespace {
sgpp_hip::Precision to_dev(HIPPrecision p) {
  return p == HIPPrecision::FP64 ? sgpp_hip::Precision::FP64 : sgpp_hip::Precision::FP32;
}
}  // namespace

OperationMultiEvalHIP::OperationMultiEvalHIP(base::Grid& grid, base::DataMatrix& dataset,
                                             const HIPOperationConfiguration& config)
    : base::OperationMultipleEval(grid, dataset), config_(config) {
  if (grid.getType() != base::GridType::Linear && grid.getType() != base::GridType::ModLinear)
    throw base::factory_exception("OperationMultiEvalHIP supports Linear and ModLinear grids");
  if (dataset.getNcols() != grid.getDimension())
    throw base::data_exception("dataset dimension does not match the grid dimension");
  if (dataset.getNrows() == 0) throw base::data_exception("dataset is empty");
  unsigned mask = 1u << static_cast<int>(to_dev(config.precision));
  if (config.keepFP64) mask |= 1u << static_cast<int>(sgpp_hip::Precision::FP64);
  sgpp_hip::LaunchConfig lc;
  lc.mult_grid_chunks = config.multGridChunks;
  lc.multT_data_segs = config.multTDataSegs;
  lc.target_blocks = config.targetBlocks;
  const sgpp_hip::Basis basis = grid.getType() == base::GridType::Linear
                                    ? sgpp_hip::Basis::Linear
                                    : sgpp_hip::Basis::ModLinear;
  numData_ = dataset.getNrows();
  dev_.reset(new sgpp_hip::DeviceMultiEval(
      dataset.getPointer(), numData_, dataset.getNcols(), basis, mask, config.device, lc,
      config.algorithm == HIPAlgorithm::SUBSPACE ? sgpp_hip::Algorithm::Subspace
                                                 : sgpp_hip::Algorithm::Streaming));
  this->prepare();
}

OperationMultiEvalHIP::~OperationMultiEvalHIP() {}

void OperationMultiEvalHIP::prepare() {
  std::vector<uint32_t> level, index;
  grid.getStorage().getLevelIndexArrays(level, index);
  numGrid_ = grid.getSize();
  dev_->set_grid(level.data(), index.data(), numGrid_);
  isPrepared = true;
}

void OperationMultiEvalHIP::mult(base::DataVector& alpha, base::DataVector& result) {
  const auto t0 = std::chrono::steady_clock::now();
  if (alpha.size() != numGrid_)
    throw base::data_exception("mult: alpha.size() must equal the number of grid points");
  std::vector<double> r(numData_);
  dev_->mult_host(alpha.data(), r.data(), to_dev(config_.precision));
  // result keeps the caller's size (OperationMultiEvalStreaming.cpp:96-112)
  for (size_t i = 0; i < result.size(); ++i) result[i] = i < numData_ ? r[i] : 0.0;
  duration_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void OperationMultiEvalHIP::multTranspose(base::DataVector& source, base::DataVector& result) {
  const auto t0 = std::chrono::steady_clock::now();
  if (result.size() != numGrid_)
    throw base::data_exception(
        "multTranspose: result.size() must equal the number of grid points");
  // entries beyond source.size() count as zero (OperationMultiEvalStreaming.cpp:120-128)
  const double* y = source.data();
  std::vector<double> ycopy;
  if (source.size() < numData_) {
    ycopy.assign(numData_, 0.0);
    std::memcpy(ycopy.data(), source.data(), source.size() * sizeof(double));
    y = ycopy.data();
  }
  dev_->mult_transpose_host(y, result.data(), to_dev(config_.precision));
  duration_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void OperationMultiEvalHIP::mult(base::DataVector& alpha, base::DataVector& result,
                                 size_t startIndexData, size_t endIndexData) {
  if (startIndexData > endIndexData || endIndexData > numData_)
    throw base::data_exception("mult: invalid data range");
  base::DataVector full(numData_);
  mult(alpha, full);
  result.resize(endIndexData - startIndexData);
  for (size_t i = startIndexData; i < endIndexData; ++i) result[i - startIndexData] = full[i];
}

void OperationMultiEvalHIP::multTranspose(base::DataVector& source, base::DataVector& result,
                                          size_t startIndexGrid, size_t endIndexGrid) {
  if (startIndexGrid > endIndexGrid || endIndexGrid > numGrid_)
    throw base::data_exception("multTranspose: invalid grid range");
  base::DataVector full(numGrid_);
  multTranspose(source, full);
  result.resize(endIndexGrid - startIndexGrid);
  for (size_t i = startIndexGrid; i < endIndexGrid; ++i) result[i - startIndexGrid] = full[i];
}

void OperationMultiEvalHIP::multDevice(const double* a, double* r) {
  dev_->mult_device(a, r, to_dev(config_.precision));
}
void OperationMultiEvalHIP::multTransposeDevice(const double* y, double* r) {
  dev_->mult_transpose_device(y, r, to_dev(config_.precision));
}
void OperationMultiEvalHIP::multDevice(const double* a, double* r, HIPPrecision p) {
  dev_->mult_device(a, r, to_dev(p));
}
void OperationMultiEvalHIP::multTransposeDevice(const double* y, double* r, HIPPrecision p) {
  dev_->mult_transpose_device(y, r, to_dev(p));
}
void OperationMultiEvalHIP::synchronize() { dev_->synchronize(); }
double OperationMultiEvalHIP::getKernelMilliseconds() {
  dev_->synchronize();
  return dev_->last_kernel_ms();
}

std::string OperationMultiEvalHIP::getImplementationName() {
  if (config_.algorithm == HIPAlgorithm::SUBSPACE) return "HIP_SUBSPACE_FP64";
  return config_.precision == HIPPrecision::FP64 ? "HIP_STREAMING_FP64"
                                                 : "HIP_STREAMING_FP32_FP64ACC";
}

}  // namespace datadriven

namespace op_factory {

base::OperationMultipleEval* createOperationMultipleEval(
    base::Grid& grid, base::DataMatrix& dataset,
    datadriven::OperationMultipleEvalConfiguration& configuration) {
  using datadriven::OperationMultipleEvalMPIType;
  using datadriven::OperationMultipleEvalSubType;
  using datadriven::OperationMultipleEvalType;
  if (configuration.getMPIType() != OperationMultipleEvalMPIType::NONE)
    throw base::factory_exception(
        "Error creating function: the library wasn't compiled with MPI/HPX support");
  // (STREAMING or DEFAULT, HIP): the streaming kernels
  // (SUBSPACELINEAR, HIP): the subspace kernels, where the reference's factory
  //   creates OperationMultipleEvalSubspaceCombined (DatadrivenOpFactory.cpp:453-468)
  const bool subspace = configuration.getType() == OperationMultipleEvalType::SUBSPACELINEAR;
  if (configuration.getSubType() == OperationMultipleEvalSubType::HIP &&
      (subspace || configuration.getType() == OperationMultipleEvalType::STREAMING ||
       configuration.getType() == OperationMultipleEvalType::DEFAULT)) {
    if (grid.getType() != base::GridType::Linear && grid.getType() != base::GridType::ModLinear)
      throw base::factory_exception("OperationMultiEval is not implemented for this grid type.");
    datadriven::HIPOperationConfiguration cfg;
    auto params = configuration.getParameters();
    if (params) {
      auto* hp = dynamic_cast<datadriven::HIPOperationConfiguration*>(params.get());
      if (!hp)
        throw base::factory_exception("HIP operator needs a HIPOperationConfiguration parameter");
      cfg = *hp;
    }
    if (subspace) cfg.algorithm = datadriven::HIPAlgorithm::SUBSPACE;
    return new datadriven::OperationMultiEvalHIP(grid, dataset, cfg);
  }
  throw base::factory_exception(
      "OperationMultiEval: this library provides the (STREAMING, HIP) and (SUBSPACELINEAR, HIP) "
      "operators only");
}

}  // namespace op_factory
}  // namespace sgpp
