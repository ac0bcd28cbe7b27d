// ORACLE -- test infrastructure only (never shipped, never on the product path).
//
// A driver program around the REAL reference learner chain, compiled from the
// reference's own sources by oracle/Makefile (libsgpp_ref_learn.so):
//   LearnerLeastSquaresIdentity / LearnerBase::train, predict, getAccuracy
//     (datadriven/src/sgpp/datadriven/application/LearnerBase.cpp:113-411,
//      LearnerLeastSquaresIdentity.cpp:32-68),
//   SystemMatrixLeastSquaresIdentity (algorithm/SystemMatrixLeastSquaresIdentity.cpp:18-65),
//   solver::ConjugateGradients, op_factory::createOperationMultipleEval
//     (DatadrivenOpFactory.cpp:409-552),
//   ModelFittingLeastSquares::fit / adapt / evaluate
//     (datamining/modules/fitting/ModelFittingLeastSquares.cpp:30-142),
//   RegressionLearner::train / predict / getMSE with DMSystemMatrix
//     (application/RegressionLearner.cpp:62-128, algorithm/DMSystemMatrix.cpp:16-86),
//   DMWeightMatrix (algorithm/DMWeightMatrix.cpp:24-60) in ConjugateGradients.
//
// It also proves the drop-in: this program defines
// op_factory::createOperationMultipleEval itself. A definition in the program
// takes precedence over the library's, so every call the reference classes
// make to the factory arrives here. A configuration with the HIP sub-type gets
// the SG++-side adapter (integration/OperationMultiEvalHIP.hpp) over
// libsgpp_hip.so; every other configuration is passed on to the reference's
// own factory (found with dlsym(RTLD_NEXT)). This stands in for the one branch
// a maintainer adds to the factory; no reference file is changed.
// RegressionLearner, DMSystemMatrix and DMWeightMatrix take their operator
// from the base factory's createOperationMultipleEval(grid, dataset)
// (base/src/sgpp/base/operation/BaseOpFactory.cpp:595-625), which has no
// configuration; with op "hipdefault" this program's definition of that
// function returns the adapter (FP64) for Linear and ModLinear grids, the
// branch a maintainer adds there to make HIP the default.
//
//   learner_check <learner|fitter|regression|weighted>
//                 <streaming|default|subspace|hip64|hip32|hipsubspace|hipdefault> <dir> key=value...
// dir holds X.bin (M x d), y.bin (M), Xt.bin (Mt x d), raw little-endian
// float64. Keys: basis (0 Linear, 1 ModLinear), dim, level, M, Mt, lambda,
// eps_refine, it_refine, eps_final, it_final, refinements, points, threshold,
// test_acc, regularization (0 Identity, 1 Laplace, 2 Diagonal; regression
// mode). weighted mode also reads w.bin (M weights). Outputs in dir:
// <stage>_level.bin, <stage>_index.bin (uint32, N x d, sequence order),
// <stage>_alpha.bin, <stage>_pred.bin (float64) for stage "train" (learner,
// regression: alpha empty, the learner keeps its weights private; weighted),
// or "fit" and "adapt" (fitter), and info.txt.

#include <dlfcn.h>

#include <sgpp/base/datatypes/DataMatrix.hpp>
#include <sgpp/base/datatypes/DataVector.hpp>
#include <sgpp/base/grid/Grid.hpp>
#include <sgpp/base/tools/OperationConfiguration.hpp>
#include <sgpp/base/operation/BaseOpFactory.hpp>
#include <sgpp/datadriven/DatadrivenOpFactory.hpp>
#include <sgpp/datadriven/algorithm/DMWeightMatrix.hpp>
#include <sgpp/datadriven/application/LearnerLeastSquaresIdentity.hpp>
#include <sgpp/datadriven/application/RegressionLearner.hpp>
#include <sgpp/datadriven/configuration/RegularizationConfiguration.hpp>
#include <sgpp/solver/sle/ConjugateGradients.hpp>
#include <sgpp/datadriven/datamining/modules/fitting/FitterConfigurationLeastSquares.hpp>
#include <sgpp/datadriven/datamining/modules/fitting/ModelFittingLeastSquares.hpp>
#include <sgpp/datadriven/operation/hash/DatadrivenOperationCommon.hpp>
#include <sgpp/datadriven/tools/Dataset.hpp>
#include <sgpp/solver/TypesSolver.hpp>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "OperationMultiEvalHIP.hpp"

using sgpp::base::DataMatrix;
using sgpp::base::DataVector;
using sgpp::base::Grid;
using sgpp::datadriven::OperationMultipleEvalConfiguration;

namespace {
int g_hip_ops = 0;         // operators the factory returned from the HIP branch
int g_other_ops = 0;       // operators it passed on to the reference's factory
bool g_base_hip = false;   // op "hipdefault": the base factory returns the adapter
int g_base_hip_ops = 0;    // operators the base factory returned from the HIP branch
int g_base_other_ops = 0;  // operators it passed on to the reference's base factory
}  // namespace

namespace sgpp {
namespace op_factory {

base::OperationMultipleEval* createOperationMultipleEval(
    base::Grid& grid, base::DataMatrix& dataset,
    datadriven::OperationMultipleEvalConfiguration& configuration) {
  if (configuration.getSubType() == datadriven::kOperationMultipleEvalSubTypeHIP) {
    ++g_hip_ops;
    return datadriven::createOperationMultiEvalHIP(grid, dataset, configuration);
  }
  using Fn = base::OperationMultipleEval* (*)(base::Grid&, base::DataMatrix&,
                                              datadriven::OperationMultipleEvalConfiguration&);
  static Fn next = reinterpret_cast<Fn>(dlsym(
      RTLD_NEXT,
      "_ZN4sgpp10op_factory27createOperationMultipleEvalERNS_4base4GridERNS1_"
      "10DataMatrixERNS_10datadriven34OperationMultipleEvalConfigurationE"));
  if (!next) throw std::runtime_error("the reference's createOperationMultipleEval not found");
  ++g_other_ops;
  return next(grid, dataset, configuration);
}

base::OperationMultipleEval* createOperationMultipleEval(base::Grid& grid,
                                                         base::DataMatrix& dataset) {
  if (g_base_hip && (grid.getType() == base::GridType::Linear ||
                     grid.getType() == base::GridType::ModLinear)) {
    ++g_base_hip_ops;
    datadriven::OperationMultipleEvalConfiguration cfg(
        datadriven::OperationMultipleEvalType::STREAMING,
        datadriven::kOperationMultipleEvalSubTypeHIP);
    return datadriven::createOperationMultiEvalHIP(grid, dataset, cfg);
  }
  using Fn = base::OperationMultipleEval* (*)(base::Grid&, base::DataMatrix&);
  static Fn next = reinterpret_cast<Fn>(dlsym(
      RTLD_NEXT, "_ZN4sgpp10op_factory27createOperationMultipleEvalERNS_4base4GridERNS1_10DataMatrixE"));
  if (!next) throw std::runtime_error("the reference's base createOperationMultipleEval not found");
  ++g_base_other_ops;
  return next(grid, dataset);
}

}  // namespace op_factory
}  // namespace sgpp

namespace {

std::vector<double> read_f64(const std::string& path, size_t n) {
  std::vector<double> v(n);
  std::ifstream f(path, std::ios::binary);
  if (!f.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(n * 8)))
    throw std::runtime_error("cannot read " + path);
  return v;
}

template <typename T>
void write_raw(const std::string& path, const T* p, size_t n) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(p), static_cast<std::streamsize>(n * sizeof(T)));
  if (!f) throw std::runtime_error("cannot write " + path);
}

void write_grid(const std::string& dir, const std::string& stage, Grid& grid,
                const DataVector& alpha, const DataVector& pred) {
  sgpp::base::GridStorage& st = grid.getStorage();
  const size_t n = st.getSize(), d = st.getDimension();
  std::vector<uint32_t> lv(n * d), ix(n * d);
  for (size_t s = 0; s < n; ++s)
    for (size_t k = 0; k < d; ++k) {
      sgpp::base::HashGridPoint::level_type l;
      sgpp::base::HashGridPoint::index_type i;
      st.getPoint(s).get(k, l, i);
      lv[s * d + k] = l;
      ix[s * d + k] = i;
    }
  write_raw(dir + "/" + stage + "_level.bin", lv.data(), lv.size());
  write_raw(dir + "/" + stage + "_index.bin", ix.data(), ix.size());
  write_raw(dir + "/" + stage + "_alpha.bin", alpha.data(), alpha.getSize());
  write_raw(dir + "/" + stage + "_pred.bin", pred.data(), pred.getSize());
}

OperationMultipleEvalConfiguration config_for(const std::string& op) {
  using sgpp::datadriven::OperationMultipleEvalSubType;
  using sgpp::datadriven::OperationMultipleEvalType;
  if (op == "default" || op == "hipdefault") return OperationMultipleEvalConfiguration();
  if (op == "streaming")
    return OperationMultipleEvalConfiguration(OperationMultipleEvalType::STREAMING,
                                              OperationMultipleEvalSubType::DEFAULT);
  if (op == "subspace")  // the reference's OperationMultipleEvalSubspaceCombined
    return OperationMultipleEvalConfiguration(OperationMultipleEvalType::SUBSPACELINEAR,
                                              OperationMultipleEvalSubType::COMBINED);
  if (op == "hip64" || op == "hip32" || op == "hipsubspace") {
    sgpp::base::OperationConfiguration params;
    params.addTextAttr("INTERNAL_PRECISION", op == "hip32" ? "float" : "double");
    return OperationMultipleEvalConfiguration(op == "hipsubspace"
                                                  ? OperationMultipleEvalType::SUBSPACELINEAR
                                                  : OperationMultipleEvalType::STREAMING,
                                              sgpp::datadriven::kOperationMultipleEvalSubTypeHIP,
                                              params, "HIP");
  }
  throw std::runtime_error("unknown operator " + op);
}

sgpp::solver::SLESolverConfiguration cg_config(double eps, size_t it) {
  sgpp::solver::SLESolverConfiguration c;
  c.type_ = sgpp::solver::SLESolverType::CG;
  c.eps_ = eps;
  c.maxIterations_ = it;
  c.threshold_ = -1.0;
  c.verbose_ = false;
  return c;
}

int run(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: learner_check <learner|fitter> <op> <dir> key=value...\n");
    return 2;
  }
  const std::string mode = argv[1], op = argv[2], dir = argv[3];
  std::map<std::string, std::string> kv;
  for (int a = 4; a < argc; ++a) {
    const std::string s = argv[a];
    const size_t e = s.find('=');
    if (e == std::string::npos) throw std::runtime_error("bad argument " + s);
    kv[s.substr(0, e)] = s.substr(e + 1);
  }
  auto num = [&](const char* k) {
    auto it = kv.find(k);
    if (it == kv.end()) throw std::runtime_error(std::string("missing ") + k);
    return std::stod(it->second);
  };
  const int basis = static_cast<int>(num("basis"));
  const size_t d = static_cast<size_t>(num("dim"));
  const size_t M = static_cast<size_t>(num("M")), Mt = static_cast<size_t>(num("Mt"));
  const int level = static_cast<int>(num("level"));
  const double lambda = num("lambda");
  const std::vector<double> X = read_f64(dir + "/X.bin", M * d);
  const std::vector<double> y = read_f64(dir + "/y.bin", M);
  const std::vector<double> Xt = read_f64(dir + "/Xt.bin", Mt * d);
  DataMatrix train(X.data(), M, d), test(Xt.data(), Mt, d);
  DataVector classes(y);
  OperationMultipleEvalConfiguration cfg = config_for(op);
  g_base_hip = op == "hipdefault";
  const auto gtype = basis == 0 ? sgpp::base::GridType::Linear : sgpp::base::GridType::ModLinear;
  std::ofstream info(dir + "/info.txt");

  if (mode == "learner") {
    sgpp::datadriven::LearnerLeastSquaresIdentity learner(true, false);
    learner.setImplementation(cfg);
    sgpp::base::RegularGridConfiguration gc;
    gc.type_ = gtype;
    gc.dim_ = d;
    gc.level_ = level;
    sgpp::base::AdaptivityConfiguration ac;
    ac.numRefinements_ = static_cast<size_t>(num("refinements"));
    ac.refinementThreshold_ = num("threshold");
    ac.numRefinementPoints_ = static_cast<size_t>(num("points"));
    ac.errorBasedRefinement_ = false;
    const auto timing = learner.train(
        train, classes, gc,
        cg_config(num("eps_refine"), static_cast<size_t>(num("it_refine"))),
        cg_config(num("eps_final"), static_cast<size_t>(num("it_final"))), ac,
        num("test_acc") != 0.0, lambda);
    DataVector pred(Mt);
    learner.predict(test, pred);
    const double mse = learner.getAccuracy(train, classes);
    write_grid(dir, "train", learner.getGrid(), learner.getAlpha(), pred);
    info.precision(17);
    info << "train_mse " << mse << "\n"
         << "gflop " << timing.GFlop_ << "\n"
         << "gbyte " << timing.GByte_ << "\n";
  } else if (mode == "fitter") {
    sgpp::datadriven::FitterConfigurationLeastSquares fc;
    fc.setupDefaults();
    fc.getGridConfig().type_ = gtype;
    fc.getGridConfig().level_ = level;
    fc.getRefinementConfig().numRefinements_ = static_cast<size_t>(num("refinements"));
    fc.getRefinementConfig().numRefinementPoints_ = static_cast<size_t>(num("points"));
    fc.getRefinementConfig().refinementThreshold_ = num("threshold");
    fc.getSolverRefineConfig() = cg_config(num("eps_refine"), static_cast<size_t>(num("it_refine")));
    fc.getSolverFinalConfig() = cg_config(num("eps_final"), static_cast<size_t>(num("it_final")));
    fc.getRegularizationConfig().lambda_ = lambda;
    fc.getMultipleEvalConfig() = cfg;
    sgpp::datadriven::ModelFittingLeastSquares model(fc);
    sgpp::datadriven::Dataset ds(M, d);
    ds.getData() = train;
    ds.getTargets() = classes;
    model.fit(ds);
    DataVector pred(Mt);
    model.evaluate(test, pred);
    write_grid(dir, "fit", model.getGrid(), model.getSurpluses(), pred);
    const bool refined = model.adapt();
    model.evaluate(test, pred);
    write_grid(dir, "adapt", model.getGrid(), model.getSurpluses(), pred);
    info << "refined " << (refined ? 1 : 0) << "\n";
  } else if (mode == "regression") {
    sgpp::base::RegularGridConfiguration gc;
    gc.type_ = gtype;
    gc.dim_ = d;
    gc.level_ = level;
    sgpp::base::AdaptivityConfiguration ac;
    ac.numRefinements_ = static_cast<size_t>(num("refinements"));
    ac.refinementThreshold_ = num("threshold");
    ac.numRefinementPoints_ = static_cast<size_t>(num("points"));
    sgpp::datadriven::RegularizationConfiguration rc;
    const int reg = static_cast<int>(num("regularization"));
    rc.type_ = reg == 0   ? sgpp::datadriven::RegularizationType::Identity
               : reg == 1 ? sgpp::datadriven::RegularizationType::Laplace
                          : sgpp::datadriven::RegularizationType::Diagonal;
    rc.lambda_ = lambda;
    rc.exponentBase_ = 2.0;
    sgpp::datadriven::RegressionLearner learner(
        gc, ac, cg_config(num("eps_refine"), static_cast<size_t>(num("it_refine"))),
        cg_config(num("eps_final"), static_cast<size_t>(num("it_final"))), rc);
    learner.train(train, classes);
    DataVector pred = learner.predict(test);
    const double mse = learner.getMSE(train, classes);
    write_grid(dir, "train", learner.getGrid(), DataVector(0), pred);
    info.precision(17);
    info << "train_mse " << mse << "\n";
  } else if (mode == "weighted") {
    // (B^T W B + lambda C) alpha = B^T W y with C the identity, solved by CG
    std::unique_ptr<Grid> grid(basis == 0 ? Grid::createLinearGrid(d)
                                          : Grid::createModLinearGrid(d));
    grid->getGenerator().regular(level);
    DataVector w(read_f64(dir + "/w.bin", M));
    std::unique_ptr<sgpp::base::OperationMatrix> C(sgpp::op_factory::createOperationIdentity(*grid));
    sgpp::datadriven::DMWeightMatrix A(*grid, train, *C, lambda, w);
    DataVector b(grid->getSize()), alpha(grid->getSize(), 0.0);
    A.generateb(classes, b);
    sgpp::solver::ConjugateGradients cg(static_cast<size_t>(num("it_final")), num("eps_final"));
    cg.solve(A, alpha, b, false, false, -1.0);
    DataVector pred(Mt);
    std::unique_ptr<sgpp::base::OperationMultipleEval> op_t(
        sgpp::op_factory::createOperationMultipleEval(*grid, test));
    op_t->mult(alpha, pred);
    write_grid(dir, "train", *grid, alpha, pred);
    info << "iterations " << cg.getNumberIterations() << "\n";
  } else {
    throw std::runtime_error("unknown mode " + mode);
  }
  info << "hip_ops " << g_hip_ops << "\n"
       << "reference_factory_ops " << g_other_ops << "\n"
       << "base_hip_ops " << g_base_hip_ops << "\n"
       << "base_reference_ops " << g_base_other_ops << "\n";
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "learner_check: %s\n", e.what());
    return 1;
  }
}
