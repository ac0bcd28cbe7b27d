// Least-squares model fitter of the reference's datamining pipeline
// (datadriven/src/sgpp/datadriven/datamining/modules/fitting/
// ModelFittingLeastSquares.cpp:30-142) on the HIP operator: the same system
// (SystemMatrixLeastSquaresIdentity, B^T B alpha + M lambda alpha = B^T y) and
// CG (warm-started, reuse = true) as LearnerLeastSquaresIdentity.
//
//   fit(data, y)    regular grid of the configured level and type, alpha = 0,
//                   solve with the final solver configuration (:49-61);
//   adapt()         one surplus refinement (numRefinementPoints_,
//                   refinementThreshold_), alpha zero-extended, solve with the
//                   refinement configuration; false once numRefinements_ are
//                   done or when refinement adds no point (:63-100);
//   update(data, y) new data on the current grid, solve from the current
//                   alpha with the final configuration (see below);
//   evaluate        B alpha on new points through the operator (:37-47).
//
// Deviation: the reference's update() calls reset(), which frees the grid
// (:124-127), and then builds the system on it (:103-111, :129-140). Here
// update keeps the grid and warm-starts from the current alpha; the
// refinement count restarts at zero as reset() does.
#pragma once

#include <cstddef>
#include <memory>

#include "sgpp/base/datatypes/DataMatrix.hpp"
#include "sgpp/base/datatypes/DataVector.hpp"
#include "sgpp/base/grid/Grid.hpp"
#include "sgpp/datadriven/application/LearnerLeastSquaresIdentity.hpp"
#include "sgpp/datadriven/operation/DatadrivenOperationCommon.hpp"

namespace sgpp {
namespace datadriven {

// FitterConfigurationLeastSquares: the members the fitter reads; defaults of
// FitterConfiguration::setupDefaults (FitterConfiguration.cpp:150-220).
struct FitterConfigurationLeastSquares {
  base::RegularGridConfiguration grid;  // type_, level_; dim_ comes from the data
  base::AdaptivityConfiguration refinement{1, 0.0, 1, false};
  solver::SLESolverConfiguration solverRefine{solver::SLESolverType::CG, 1e-12, 100, 1e-12, false};
  solver::SLESolverConfiguration solverFinal{solver::SLESolverType::CG, 1e-12, 100, 1e-12, false};
  double lambda = 0.01;
  // must select the HIP operator (OperationMultipleEvalSubType::HIP)
  OperationMultipleEvalConfiguration multipleEval;
};

class ModelFittingLeastSquares {
 public:
  explicit ModelFittingLeastSquares(const FitterConfigurationLeastSquares& config);
  ~ModelFittingLeastSquares();

  void fit(const base::DataMatrix& data, const base::DataVector& targets);
  bool adapt();
  void update(const base::DataMatrix& data, const base::DataVector& targets);
  double evaluate(const base::DataVector& sample);
  void evaluate(base::DataMatrix& samples, base::DataVector& results);

  bool isFitted() const { return grid_ != nullptr; }
  base::Grid& getGrid();
  const base::DataVector& getSurpluses() const { return alpha_; }
  size_t getRefinementsPerformed() const { return refinementsPerformed_; }
  size_t getLastIterations() const { return lastIterations_; }

 private:
  void assembleSystemAndSolve(const solver::SLESolverConfiguration& sc);

  FitterConfigurationLeastSquares config_;
  std::unique_ptr<base::Grid> grid_;
  base::DataVector alpha_;
  std::unique_ptr<base::DataMatrix> data_;
  std::unique_ptr<base::DataVector> targets_;
  size_t refinementsPerformed_ = 0;
  size_t lastIterations_ = 0;
};

}  // namespace datadriven
}  // namespace sgpp
