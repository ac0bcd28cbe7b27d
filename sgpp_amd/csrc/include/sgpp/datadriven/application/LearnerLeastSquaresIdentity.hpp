// Regression learner on the HIP operator: the caller of the multi-evaluation
// path that drives refine -> prepare -> b = B^T y -> CG.
//
// Interface and step order follow
//   datadriven/src/sgpp/datadriven/application/LearnerBase.cpp:113-327 (train),
//   :347-354 (predict), :369-411 (getAccuracy), :81-98 (InitializeGrid) and
//   LearnerLeastSquaresIdentity.cpp:32-39 (createDMSystem), :61-68 (predict).
// The system A = B^T B + M*lambda*I (SystemMatrixLeastSquaresIdentity.cpp:32-47)
// and the CG recurrence (solver/src/sgpp/solver/sle/ConjugateGradients.cpp:22-151)
// run device-resident: only the grid refinement runs on the host.
//
// Differences from the reference:
//   - the implementation configuration must select the HIP operator
//     (STREAMING or DEFAULT type, HIP sub-type);
//   - with HIPPrecision::FP32_FP64ACC the solve is FP64 iterative refinement
//     around CG on FP32 products (same stopping rule ||r|| <= eps ||b||);
//   - only CG is provided (BiCGSTAB throws application_exception);
//   - errorBasedRefinement_ throws not_implemented_exception: the reference's
//     LearnerLeastSquaresIdentity::multTranspose (LearnerLeastSquaresIdentity.cpp:70-77)
//     calls mult() on an M-vector for an N-vector result, so that branch has
//     no well-defined result to reproduce.
#pragma once

#include <cstddef>
#include <memory>
#include <vector>

#include "sgpp/base/datatypes/DataMatrix.hpp"
#include "sgpp/base/datatypes/DataVector.hpp"
#include "sgpp/base/grid/Grid.hpp"
#include "sgpp/datadriven/operation/DatadrivenOperationCommon.hpp"

namespace sgpp {
namespace base {

// base/src/sgpp/base/grid/Grid.hpp: GeneralGridConfiguration (type_, dim_, level_)
struct RegularGridConfiguration {
  GridType type_ = GridType::Linear;
  size_t dim_ = 0;
  int level_ = 3;
};

// base/src/sgpp/base/grid/Grid.hpp:140-165 (the members the learner reads)
struct AdaptivityConfiguration {
  size_t numRefinements_ = 1;
  double refinementThreshold_ = 0.0;
  size_t numRefinementPoints_ = 5;
  bool errorBasedRefinement_ = false;
};

}  // namespace base

namespace solver {

// solver/src/sgpp/solver/TypesSolver.hpp:17-25
enum class SLESolverType { CG, BiCGSTAB, FISTA };

struct SLESolverConfiguration {
  SLESolverType type_ = SLESolverType::CG;
  double eps_ = 1e-10;
  size_t maxIterations_ = 1000;
  double threshold_ = 0.0;
  bool verbose_ = false;
};

}  // namespace solver

namespace datadriven {

// datadriven/src/sgpp/datadriven/tools/TypesDatadriven.hpp:37-54
struct LearnerTiming {
  double timeComplete_ = 0.0;
  double timeMultComplete_ = 0.0;
  double timeMultCompute_ = 0.0;
  double timeMultTransComplete_ = 0.0;
  double timeMultTransCompute_ = 0.0;
  double timeRegularization_ = 0.0;
  double GFlop_ = 0.0;
  double GByte_ = 0.0;
};

// datadriven/src/sgpp/datadriven/tools/LearnerVectorizedPerformanceCalculator.hpp:23-28
struct LearnerVectorizedPerformance {
  double GFlop_ = 0.0;
  double GByte_ = 0.0;
};

// Operation count convention of the reference's vectorised learners
// (LearnerVectorizedPerformanceCalculator.cpp:17-113): per CG run,
// iterations + 1.5 (+1 with reused coefficients, +1 per 50 iterations)
// products; Linear 2 * N * M * d * 6 flop each, ModLinear 8 / 10 / 12 flop
// per (grid point, dimension, instance) for left / right / interior factors
// and none for level 1; BiCGSTAB doubles both counts.
class LearnerVectorizedPerformanceCalculator {
 public:
  static LearnerVectorizedPerformance getGFlopAndGByte(base::Grid& grid, size_t numInstances,
                                                       solver::SLESolverType solver,
                                                       size_t numIterations, size_t sizeDatatype,
                                                       bool reuseAlpha = false,
                                                       bool verbose = false);
};

// One pass of the refinement loop (for reporting).
struct LearnerStep {
  size_t gridSize = 0;
  size_t iterations = 0;        // CG iterations (inner iterations in mixed mode)
  size_t refinements = 0;       // outer refinement steps in mixed mode
  double residual = 0.0;        // ||b - A alpha|| / ||b|| at the end of the solve
  double trainAccuracy = -1.0;  // MSE (regression) or accuracy, when tested
  double seconds = 0.0;
};

class LearnerLeastSquaresIdentity {
 public:
  explicit LearnerLeastSquaresIdentity(bool isRegression, bool verbose = false);
  ~LearnerLeastSquaresIdentity();

  // LearnerBase.hpp:292 (implementationConfiguration); must select the HIP operator
  void setImplementation(const OperationMultipleEvalConfiguration& operationConfiguration);

  LearnerTiming train(base::DataMatrix& trainDataset, base::DataVector& classes,
                      const base::RegularGridConfiguration& gridConfig,
                      const solver::SLESolverConfiguration& SolverConfigRefine,
                      const solver::SLESolverConfiguration& SolverConfigFinal,
                      const base::AdaptivityConfiguration& adaptivityConfig,
                      bool testAccDuringAdapt, double lambdaRegularization,
                      base::DataMatrix* testDataset = nullptr,
                      base::DataVector* testClasses = nullptr);

  void predict(base::DataMatrix& testDataset, base::DataVector& classesComputed);
  double getAccuracy(base::DataMatrix& testDataset, const base::DataVector& classesReference,
                     double threshold = 0.0);
  double getAccuracy(const base::DataVector& classesComputed,
                     const base::DataVector& classesReference, double threshold = 0.0) const;

  bool getIsRegression() const { return isRegression_; }
  bool isTrained() const { return isTrained_; }
  base::Grid& getGrid();
  const base::DataVector& getAlpha() const;
  const std::vector<LearnerStep>& getSteps() const { return steps_; }

 private:
  void initializeGrid(const base::RegularGridConfiguration& gridConfig);

  bool isRegression_;
  bool isVerbose_;
  bool reuseCoefficients_ = true;  // LearnerBase.cpp:30
  bool isTrained_ = false;
  OperationMultipleEvalConfiguration implementationConfiguration_;
  std::unique_ptr<base::Grid> grid_;
  std::unique_ptr<base::DataVector> alpha_;
  std::vector<LearnerStep> steps_;
  double execTime_ = 0.0;
  double GFlop_ = 0.0;
  double GByte_ = 0.0;
};

}  // namespace datadriven
}  // namespace sgpp
