// Regression learner on the HIP operator (see LearnerLeastSquaresIdentity.hpp).
#include "sgpp/datadriven/application/LearnerLeastSquaresIdentity.hpp"
#include "sgpp/datadriven/datamining/ModelFittingLeastSquares.hpp"

#include <chrono>
#include <cstdint>
#include <iostream>
#include <string>

#include "sgpp/base/exception/exceptions.hpp"
#include "sgpp/datadriven/operation/OperationMultiEvalHIP.hpp"
#include "solver_device.hpp"

namespace sgpp {
namespace datadriven {

using base::DataMatrix;
using base::DataVector;

// LearnerVectorizedPerformanceCalculator.cpp:17-113
LearnerVectorizedPerformance LearnerVectorizedPerformanceCalculator::getGFlopAndGByte(
    base::Grid& grid, size_t numInstances, solver::SLESolverType solver, size_t numIterations,
    size_t sizeDatatype, bool reuseAlpha, bool /*verbose*/) {
  LearnerVectorizedPerformance r;
  const double M = static_cast<double>(numInstances);
  const double N = static_cast<double>(grid.getSize());
  const double dim = static_cast<double>(grid.getDimension());
  const double bytes = static_cast<double>(sizeDatatype);
  // products per CG run: the iterations, b (counted as half), the start residual,
  // one more with a reused start vector, one per residual recomputation
  double its = static_cast<double>(numIterations) + 1.5;
  if (reuseAlpha) its += 1.0;
  its += static_cast<double>(numIterations / 50);
  if (grid.getType() == base::GridType::ModLinear) {
    base::HashGridStorage& st = grid.getStorage();
    const size_t d = grid.getDimension();
    for (size_t g = 0; g < grid.getSize(); ++g) {
      const base::HashGridPoint& p = st.getPoint(g);
      for (size_t h = 0; h < d; ++h) {
        base::level_t l;
        base::index_t i;
        p.get(h, l, i);
        if (l == 1) continue;
        double flop, words;
        if (i == 1) {
          flop = 8.0;
          words = 4.0;
        } else if (static_cast<uint64_t>(i) == (uint64_t(1) << l) - 1) {
          flop = 10.0;
          words = 6.0;
        } else {
          flop = 12.0;
          words = 6.0;
        }
        r.GFlop_ += 1e-9 * flop * its * M;
        r.GByte_ += 1e-9 * words * its * bytes * M;
      }
    }
    r.GByte_ += 1e-9 * its * (N * (M + 1.0) * bytes);   // coefficients, transposed product
    r.GByte_ += 1e-9 * its * ((N + 1.0) * M * bytes);   // coefficients, product
  } else {
    r.GFlop_ += 2.0 * 1e-9 * N * M * dim * 6.0 * its;
    r.GByte_ += 2.0 * 1e-9 * N * M * dim * 3.0 * its * bytes;
    r.GByte_ += 1e-9 * its * (N * M * bytes);
    r.GByte_ += 1e-9 * its * (N * M * bytes);
  }
  if (solver == solver::SLESolverType::BiCGSTAB) {
    r.GFlop_ *= 2.0;
    r.GByte_ *= 2.0;
  }
  return r;
}

namespace {

// The configuration must name the HIP operator; the learner's copy also keeps
// FP64 data on the device when products run in FP32 (the refinement's FP64
// residuals need it).
HIPOperationConfiguration hip_parameters(OperationMultipleEvalConfiguration& c) {
  if (c.getSubType() != OperationMultipleEvalSubType::HIP ||
      (c.getType() != OperationMultipleEvalType::STREAMING &&
       c.getType() != OperationMultipleEvalType::DEFAULT))
    throw base::factory_exception(
        "LearnerLeastSquaresIdentity: the implementation configuration must select the "
        "(STREAMING, HIP) operator");
  HIPOperationConfiguration p;
  auto params = c.getParameters();
  if (params) {
    auto* hp = dynamic_cast<HIPOperationConfiguration*>(params.get());
    if (!hp) throw base::factory_exception("HIP operator needs a HIPOperationConfiguration");
    p = *hp;
  }
  return p;
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

LearnerLeastSquaresIdentity::LearnerLeastSquaresIdentity(bool isRegression, bool verbose)
    : isRegression_(isRegression),
      isVerbose_(verbose),
      implementationConfiguration_(OperationMultipleEvalType::STREAMING,
                                   OperationMultipleEvalSubType::HIP, HIPOperationConfiguration()) {}

LearnerLeastSquaresIdentity::~LearnerLeastSquaresIdentity() {}

void LearnerLeastSquaresIdentity::setImplementation(
    const OperationMultipleEvalConfiguration& operationConfiguration) {
  OperationMultipleEvalConfiguration c = operationConfiguration;
  hip_parameters(c);  // validates
  implementationConfiguration_ = c;
}

base::Grid& LearnerLeastSquaresIdentity::getGrid() {
  if (!grid_) throw base::application_exception("LearnerLeastSquaresIdentity: no grid yet");
  return *grid_;
}

const DataVector& LearnerLeastSquaresIdentity::getAlpha() const {
  if (!alpha_) throw base::application_exception("LearnerLeastSquaresIdentity: no alpha yet");
  return *alpha_;
}

// LearnerBase.cpp:81-98
void LearnerLeastSquaresIdentity::initializeGrid(const base::RegularGridConfiguration& gc) {
  if (gc.type_ != base::GridType::Linear && gc.type_ != base::GridType::ModLinear)
    throw base::application_exception(
        "LearnerBase::InitializeGrid: An unsupported grid type was chosen!");
  if (gc.level_ < 1) throw base::application_exception("LearnerBase::InitializeGrid: level < 1");
  grid_.reset(new base::Grid(gc.type_, gc.dim_));
  grid_->getGenerator().regular(static_cast<size_t>(gc.level_));
  alpha_.reset(new DataVector(grid_->getSize(), 0.0));
}

// LearnerBase.cpp:113-327
LearnerTiming LearnerLeastSquaresIdentity::train(
    DataMatrix& trainDataset, DataVector& classes, const base::RegularGridConfiguration& gridConfig,
    const solver::SLESolverConfiguration& SolverConfigRefine,
    const solver::SLESolverConfiguration& SolverConfigFinal,
    const base::AdaptivityConfiguration& adaptivityConfig, bool testAccDuringAdapt,
    double lambdaRegularization, DataMatrix* testDataset, DataVector* testClasses) {
  LearnerTiming result;
  if (trainDataset.getNrows() != classes.getSize())
    throw base::application_exception(
        "LearnerBase::train: length of classes vector does not match to dataset!");
  if (SolverConfigRefine.type_ != solver::SLESolverType::CG ||
      SolverConfigFinal.type_ != solver::SLESolverType::CG)
    throw base::application_exception(
        "LearnerBase::train: An unsupported SLE solver type was chosen!");
  if (adaptivityConfig.errorBasedRefinement_)
    throw base::not_implemented_exception(
        "LearnerLeastSquaresIdentity: error-based refinement is not provided");
  if (gridConfig.dim_ != trainDataset.getNcols())
    throw base::application_exception("LearnerBase::train: grid and data dimension differ");

  execTime_ = GFlop_ = GByte_ = 0.0;
  steps_.clear();
  isTrained_ = false;
  initializeGrid(gridConfig);

  // system matrix operator (SystemMatrixLeastSquaresIdentity.cpp:17-26, 60-65)
  HIPOperationConfiguration hp = hip_parameters(implementationConfiguration_);
  const bool mixed = hp.precision == HIPPrecision::FP32_FP64ACC;
  if (mixed) hp.keepFP64 = true;
  OperationMultipleEvalConfiguration cfg(OperationMultipleEvalType::STREAMING,
                                         OperationMultipleEvalSubType::HIP, hp, "learner");
  std::unique_ptr<base::OperationMultipleEval> B(
      op_factory::createOperationMultipleEval(*grid_, trainDataset, cfg));
  auto* hop = dynamic_cast<OperationMultiEvalHIP*>(B.get());
  if (!hop) throw base::factory_exception("LearnerLeastSquaresIdentity: not a HIP operator");
  hop->device().set_kernel_timing(true);

  double oldAcc = 0.0;
  for (size_t i = 0; i < adaptivityConfig.numRefinements_ + 1; ++i) {
    const auto t0 = std::chrono::steady_clock::now();
    if (i > 0) {
      base::SurplusRefinementFunctor f(*alpha_, adaptivityConfig.numRefinementPoints_,
                                       adaptivityConfig.refinementThreshold_);
      grid_->getGenerator().refine(f);
      B->prepare();
      alpha_->resize(grid_->getSize(), 0.0);
    }
    const solver::SLESolverConfiguration& sc =
        (i == adaptivityConfig.numRefinements_) ? SolverConfigFinal : SolverConfigRefine;
    sgpp_hip::DeviceRegression system(hop->device());
    sgpp_hip::SolveOptions o;
    o.mixed = mixed;
    o.lambda = lambdaRegularization;
    o.max_iterations = sc.maxIterations_;
    o.epsilon = sc.eps_;
    o.reuse = reuseCoefficients_;
    const sgpp_hip::SolveResult r = system.solve(classes.getPointer(), alpha_->getPointer(), o);
    const double stepTime = seconds_since(t0);
    execTime_ += stepTime;

    // LearnerLeastSquaresIdentity::postProcessing (LearnerLeastSquaresIdentity.cpp:41-59)
    const LearnerVectorizedPerformance perf =
        LearnerVectorizedPerformanceCalculator::getGFlopAndGByte(
            *grid_, trainDataset.getNrows(), sc.type_, r.iterations, sizeof(double),
            reuseCoefficients_, true);
    GFlop_ += perf.GFlop_;
    GByte_ += perf.GByte_;

    int nm = 0, nt = 0;
    double msm = 0.0, mst = 0.0;
    hop->device().kernel_timing(&nm, &msm, &nt, &mst);
    result.timeComplete_ = execTime_;
    // products run on device-resident vectors: no transfers inside the solve
    result.timeMultCompute_ = result.timeMultComplete_ = msm * 1e-3;
    result.timeMultTransCompute_ = result.timeMultTransComplete_ = mst * 1e-3;
    result.timeRegularization_ = 0.0;
    result.GFlop_ = GFlop_;
    result.GByte_ = GByte_;

    LearnerStep step;
    step.gridSize = grid_->getSize();
    step.iterations = r.iterations;
    step.refinements = r.refinements;
    step.residual = r.residual_rel;
    step.seconds = stepTime;
    if (isVerbose_)
      std::cout << "refinement " << i << ": grid " << step.gridSize << ", iterations "
                << step.iterations << ", residual " << step.residual << std::endl;

    if (testAccDuringAdapt) {
      // the system operator holds the training data: the same values as a
      // fresh operator from predict()
      DataVector computed(trainDataset.getNrows());
      B->mult(*alpha_, computed);
      const double acc = getAccuracy(computed, classes);
      step.trainAccuracy = acc;
      if (testDataset != nullptr && testClasses != nullptr) {
        const double testAcc = getAccuracy(*testDataset, *testClasses);
        if (isVerbose_) std::cout << "test: " << testAcc << std::endl;
      }
      steps_.push_back(step);
      if (isRegression_) {
        if (i > 0 && oldAcc <= acc) break;
      } else {
        if (i > 0 && oldAcc >= acc) break;
      }
      oldAcc = acc;
    } else {
      steps_.push_back(step);
    }
  }
  isTrained_ = true;
  return result;
}

// LearnerLeastSquaresIdentity.cpp:61-68
void LearnerLeastSquaresIdentity::predict(DataMatrix& testDataset, DataVector& classesComputed) {
  if (!grid_ || !alpha_) throw base::application_exception("LearnerBase::predict: not trained");
  classesComputed.resize(testDataset.getNrows());
  std::unique_ptr<base::OperationMultipleEval> op(
      op_factory::createOperationMultipleEval(*grid_, testDataset, implementationConfiguration_));
  op->mult(*alpha_, classesComputed);
}

// LearnerBase.cpp:369-378
double LearnerLeastSquaresIdentity::getAccuracy(DataMatrix& testDataset,
                                                const DataVector& classesReference,
                                                double threshold) {
  DataVector computed(testDataset.getNrows());
  predict(testDataset, computed);
  return getAccuracy(computed, classesReference, threshold);
}

// LearnerBase.cpp:380-411
double LearnerLeastSquaresIdentity::getAccuracy(const DataVector& computed,
                                                const DataVector& reference,
                                                double threshold) const {
  if (computed.getSize() != reference.getSize())
    throw base::application_exception(
        "LearnerBase::getAccuracy: lengths of classes vectors do not match!");
  if (isRegression_) {
    double s = 0.0;
    for (size_t i = 0; i < computed.getSize(); ++i) {
      const double t = computed[i] - reference[i];
      s += t * t;
    }
    return s / static_cast<double>(computed.getSize());
  }
  size_t correct = 0;
  for (size_t i = 0; i < computed.getSize(); ++i)
    if ((computed[i] >= threshold && reference[i] >= 0.0) ||
        (computed[i] < threshold && reference[i] < 0.0))
      ++correct;
  return static_cast<double>(correct) / static_cast<double>(computed.getSize());
}

// ------------------------------------------------ ModelFittingLeastSquares

ModelFittingLeastSquares::ModelFittingLeastSquares(const FitterConfigurationLeastSquares& config)
    : config_(config) {
  if (config_.multipleEval.getSubType() == OperationMultipleEvalSubType::DEFAULT &&
      config_.multipleEval.getType() == OperationMultipleEvalType::DEFAULT)
    config_.multipleEval = OperationMultipleEvalConfiguration(
        OperationMultipleEvalType::STREAMING, OperationMultipleEvalSubType::HIP,
        HIPOperationConfiguration(), "fitter");
  hip_parameters(config_.multipleEval);  // validates
  if (config_.grid.type_ != base::GridType::Linear &&
      config_.grid.type_ != base::GridType::ModLinear)
    throw base::application_exception("ModelFittingLeastSquares: unsupported grid type");
  if (config_.grid.level_ < 1)
    throw base::application_exception("ModelFittingLeastSquares: grid level < 1");
  if (config_.solverRefine.type_ != solver::SLESolverType::CG ||
      config_.solverFinal.type_ != solver::SLESolverType::CG)
    throw base::application_exception("ModelFittingLeastSquares: only CG is provided");
}

ModelFittingLeastSquares::~ModelFittingLeastSquares() {}

base::Grid& ModelFittingLeastSquares::getGrid() {
  if (!grid_) throw base::application_exception("ModelFittingLeastSquares: no grid yet");
  return *grid_;
}

// ModelFittingLeastSquares.cpp:49-61
void ModelFittingLeastSquares::fit(const DataMatrix& data, const DataVector& targets) {
  if (data.getNrows() != targets.getSize() || data.getNrows() == 0)
    throw base::application_exception("ModelFittingLeastSquares: data and targets differ in size");
  grid_.reset();
  refinementsPerformed_ = 0;
  data_.reset(new DataMatrix(data));
  targets_.reset(new DataVector(targets));
  grid_.reset(new base::Grid(config_.grid.type_, data.getNcols()));
  grid_->getGenerator().regular(static_cast<size_t>(config_.grid.level_));
  alpha_ = DataVector(grid_->getSize(), 0.0);
  assembleSystemAndSolve(config_.solverFinal);
}

// ModelFittingLeastSquares.cpp:63-100 (no interaction stencils)
bool ModelFittingLeastSquares::adapt() {
  if (!grid_)
    throw base::application_exception(
        "ModelFittingLeastSquares: Can't refine before initial grid is created");
  if (refinementsPerformed_ >= config_.refinement.numRefinements_) return false;
  base::SurplusRefinementFunctor f(alpha_, config_.refinement.numRefinementPoints_,
                                   config_.refinement.refinementThreshold_);
  const size_t before = grid_->getSize();
  grid_->getGenerator().refine(f);
  if (grid_->getSize() <= before) return false;
  alpha_.resize(grid_->getSize(), 0.0);
  assembleSystemAndSolve(config_.solverRefine);
  refinementsPerformed_++;
  return true;
}

// ModelFittingLeastSquares.cpp:102-111; keeps the grid (see the header)
void ModelFittingLeastSquares::update(const DataMatrix& data, const DataVector& targets) {
  if (!grid_) {
    fit(data, targets);
    return;
  }
  if (data.getNrows() != targets.getSize() || data.getNrows() == 0)
    throw base::application_exception("ModelFittingLeastSquares: data and targets differ in size");
  if (data.getNcols() != grid_->getDimension())
    throw base::application_exception("ModelFittingLeastSquares: data dimension differs");
  refinementsPerformed_ = 0;
  data_.reset(new DataMatrix(data));
  targets_.reset(new DataVector(targets));
  assembleSystemAndSolve(config_.solverFinal);
}

// ModelFittingLeastSquares.cpp:37-40, through the operator on one point
double ModelFittingLeastSquares::evaluate(const DataVector& sample) {
  DataMatrix one(sample.data(), 1, sample.getSize());
  DataVector r(1);
  evaluate(one, r);
  return r[0];
}

// ModelFittingLeastSquares.cpp:43-47
void ModelFittingLeastSquares::evaluate(DataMatrix& samples, DataVector& results) {
  if (!grid_) throw base::application_exception("ModelFittingLeastSquares: not fitted");
  if (samples.getNcols() != grid_->getDimension())
    throw base::application_exception("ModelFittingLeastSquares: sample dimension differs");
  results.resize(samples.getNrows());
  std::unique_ptr<base::OperationMultipleEval> op(
      op_factory::createOperationMultipleEval(*grid_, samples, config_.multipleEval));
  op->mult(alpha_, results);
}

// ModelFittingLeastSquares.cpp:129-140: system on the current grid and data,
// b = B^T y, CG from the current alpha (reuse = true)
void ModelFittingLeastSquares::assembleSystemAndSolve(const solver::SLESolverConfiguration& sc) {
  HIPOperationConfiguration hp = hip_parameters(config_.multipleEval);
  const bool mixed = hp.precision == HIPPrecision::FP32_FP64ACC;
  if (mixed) hp.keepFP64 = true;
  OperationMultipleEvalConfiguration cfg(OperationMultipleEvalType::STREAMING,
                                         OperationMultipleEvalSubType::HIP, hp, "fitter");
  std::unique_ptr<base::OperationMultipleEval> B(
      op_factory::createOperationMultipleEval(*grid_, *data_, cfg));
  auto* hop = dynamic_cast<OperationMultiEvalHIP*>(B.get());
  if (!hop) throw base::factory_exception("ModelFittingLeastSquares: not a HIP operator");
  sgpp_hip::DeviceRegression system(hop->device());
  sgpp_hip::SolveOptions o;
  o.mixed = mixed;
  o.lambda = config_.lambda;
  o.max_iterations = sc.maxIterations_;
  o.epsilon = sc.eps_;
  o.reuse = true;
  const sgpp_hip::SolveResult r = system.solve(targets_->getPointer(), alpha_.getPointer(), o);
  lastIterations_ = r.iterations;
}

}  // namespace datadriven
}  // namespace sgpp
