// ORACLE -- test infrastructure only (never shipped, never on the product path).
//
// A thin extern "C" driver around the REAL reference classes, compiled from
// the reference's own sources by oracle/Makefile into oracle/_ref/. It lets
// the tests and the CPU baseline call, unchanged:
//   Grid::createLinearGrid / createModLinearGrid + regular + refine
//     (base/src/sgpp/base/grid/Grid.hpp, HashGenerator.hpp, HashRefinement.cpp)
//   op_factory::createOperationHierarchisation (base/src/sgpp/base/operation/BaseOpFactory.hpp:56)
//   op_factory::createOperationMultipleEval(grid, data)  -- the DEFAULT operator (:145)
//   datadriven::OperationMultiEvalStreaming / OperationMultiEvalModMaskStreaming
//     (StreamingDefault) and OperationMultipleEvalSubspaceCombined (SubspaceLinear)
//   solver::ConjugateGradients on (B^T B + M*lambda*I), as
//     SystemMatrixLeastSquaresIdentity::mult does it (SystemMatrixLeastSquaresIdentity.cpp:32-47).
// Only this file is ours; everything it calls is the reference, untouched.

#include <sgpp/base/datatypes/DataMatrix.hpp>
#include <sgpp/base/datatypes/DataVector.hpp>
#include <sgpp/base/grid/Grid.hpp>
#include <sgpp/base/grid/generation/functors/SurplusRefinementFunctor.hpp>
#include <sgpp/base/operation/BaseOpFactory.hpp>
#include <sgpp/base/operation/hash/OperationHierarchisation.hpp>
#include <sgpp/base/operation/hash/OperationMatrix.hpp>
#include <sgpp/base/operation/hash/OperationMultipleEval.hpp>
#include <sgpp/datadriven/operation/hash/OperationMultiEvalModMaskStreaming/OperationMultiEvalModMaskStreaming.hpp>
#include <sgpp/datadriven/operation/hash/OperationMultiEvalStreaming/OperationMultiEvalStreaming.hpp>
#include <sgpp/datadriven/operation/hash/OperationMultipleEvalSubspace/combined/OperationMultipleEvalSubspaceCombined.hpp>
#include <sgpp/datadriven/tools/LearnerVectorizedPerformanceCalculator.hpp>
#include <sgpp/solver/sle/ConjugateGradients.hpp>

#include <chrono>
#include <cstdint>
#include <cstring>
#include <exception>
#include <memory>
#include <string>

using sgpp::base::DataMatrix;
using sgpp::base::DataVector;
using sgpp::base::Grid;

namespace {
std::string g_err;

struct RefGrid {
  std::unique_ptr<Grid> grid;
};
struct RefOp {
  DataMatrix data;
  std::unique_ptr<sgpp::base::OperationMultipleEval> op;
  double seconds = 0.0;
};

class LeastSquares : public sgpp::base::OperationMatrix {
 public:
  LeastSquares(sgpp::base::OperationMultipleEval& B, size_t M, double lambda)
      : B(B), M(M), lambda(lambda) {}
  void mult(DataVector& alpha, DataVector& result) override {
    DataVector temp(M);
    B.mult(alpha, temp);
    B.multTranspose(temp, result);
    result.axpy(static_cast<double>(M) * lambda, alpha);
  }

 private:
  sgpp::base::OperationMultipleEval& B;
  size_t M;
  double lambda;
};

template <typename F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  } catch (...) {
    g_err = "unknown";
    return 1;
  }
}
}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

void* ref_grid_new(int basis, size_t dim) {
  RefGrid* g = new RefGrid();
  g->grid.reset(basis == 0 ? Grid::createLinearGrid(dim) : Grid::createModLinearGrid(dim));
  return g;
}
void ref_grid_free(void* g) { delete static_cast<RefGrid*>(g); }
// a grid holding exactly the given points, inserted in order with the
// reference's HashGridStorage::insert (for operators that only read level/index)
void* ref_grid_from_arrays(int basis, size_t dim, size_t n, const uint32_t* lv,
                           const uint32_t* ix) {
  RefGrid* g = new RefGrid();
  g->grid.reset(basis == 0 ? Grid::createLinearGrid(dim) : Grid::createModLinearGrid(dim));
  auto& st = g->grid->getStorage();
  sgpp::base::HashGridPoint p(dim);
  for (size_t s = 0; s < n; ++s) {
    for (size_t k = 0; k < dim; ++k) p.set(k, lv[s * dim + k], ix[s * dim + k]);
    st.insert(p);
  }
  return g;
}
int ref_grid_regular(void* g, size_t level) {
  return guard([&] { static_cast<RefGrid*>(g)->grid->getGenerator().regular(level); });
}
size_t ref_grid_size(void* g) { return static_cast<RefGrid*>(g)->grid->getSize(); }
int ref_grid_level_index(void* g, uint32_t* lv, uint32_t* ix, uint8_t* leaf) {
  return guard([&] {
    auto& st = static_cast<RefGrid*>(g)->grid->getStorage();
    const size_t d = st.getDimension();
    for (size_t s = 0; s < st.getSize(); ++s) {
      for (size_t k = 0; k < d; ++k) {
        uint32_t l, i;
        st.getPoint(s).get(k, l, i);
        lv[s * d + k] = l;
        ix[s * d + k] = i;
      }
      if (leaf) leaf[s] = st.getPoint(s).isLeaf() ? 1 : 0;
    }
  });
}
int ref_grid_refine_surplus(void* g, const double* alpha, size_t n, size_t num, double thr) {
  return guard([&] {
    DataVector a(n);
    std::memcpy(a.getPointer(), alpha, n * sizeof(double));
    sgpp::base::SurplusRefinementFunctor f(a, num, thr);
    static_cast<RefGrid*>(g)->grid->getGenerator().refine(f);
  });
}
int ref_grid_hierarchise(void* g, const double* values, double* alpha) {
  return guard([&] {
    Grid& grid = *static_cast<RefGrid*>(g)->grid;
    DataVector v(grid.getSize());
    std::memcpy(v.getPointer(), values, v.getSize() * sizeof(double));
    std::unique_ptr<sgpp::base::OperationHierarchisation> h(
        sgpp::op_factory::createOperationHierarchisation(grid));
    h->doHierarchisation(v);
    std::memcpy(alpha, v.getPointer(), v.getSize() * sizeof(double));
  });
}

// kind: 0 DEFAULT, 1 StreamingDefault (Linear or ModMask), 2 SubspaceLinear
void* ref_op_new(void* g, int kind, const double* X, size_t M) {
  RefOp* o = new RefOp();
  if (guard([&] {
        Grid& grid = *static_cast<RefGrid*>(g)->grid;
        const size_t d = grid.getDimension();
        o->data = DataMatrix(X, M, d);
        if (kind == 0) {
          o->op.reset(sgpp::op_factory::createOperationMultipleEval(grid, o->data));
        } else if (kind == 1) {
          if (grid.getType() == sgpp::base::GridType::Linear)
            o->op.reset(new sgpp::datadriven::OperationMultiEvalStreaming(grid, o->data));
          else
            o->op.reset(new sgpp::datadriven::OperationMultiEvalModMaskStreaming(grid, o->data));
        } else {
          o->op.reset(new sgpp::datadriven::OperationMultipleEvalSubspaceCombined(grid, o->data));
        }
      }) != 0) {
    delete o;
    return nullptr;
  }
  return o;
}
void ref_op_free(void* o) { delete static_cast<RefOp*>(o); }
int ref_op_prepare(void* o) { return guard([&] { static_cast<RefOp*>(o)->op->prepare(); }); }
int ref_op_mult(void* o, const double* alpha, size_t N, double* result, size_t M) {
  return guard([&] {
    RefOp* r = static_cast<RefOp*>(o);
    DataVector a(N), res(M);
    std::memcpy(a.getPointer(), alpha, N * sizeof(double));
    const auto t0 = std::chrono::steady_clock::now();
    r->op->mult(a, res);
    r->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::memcpy(result, res.getPointer(), M * sizeof(double));
  });
}
int ref_op_mult_transpose(void* o, const double* y, size_t M, double* result, size_t N) {
  return guard([&] {
    RefOp* r = static_cast<RefOp*>(o);
    DataVector src(M), res(N);
    std::memcpy(src.getPointer(), y, M * sizeof(double));
    const auto t0 = std::chrono::steady_clock::now();
    r->op->multTranspose(src, res);
    r->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::memcpy(result, res.getPointer(), N * sizeof(double));
  });
}
double ref_op_seconds(void* o) { return static_cast<RefOp*>(o)->seconds; }

int ref_cg_solve(void* o, size_t M, size_t N, const double* y, double lambda, size_t max_it,
                 double eps, double* alpha, size_t* iterations, double* residuum) {
  return guard([&] {
    RefOp* r = static_cast<RefOp*>(o);
    LeastSquares A(*r->op, M, lambda);
    DataVector yy(M), b(N), a(N, 0.0);
    std::memcpy(yy.getPointer(), y, M * sizeof(double));
    r->op->multTranspose(yy, b);
    sgpp::solver::ConjugateGradients cg(max_it, eps);
    cg.solve(A, a, b, false, false, 0.0);
    std::memcpy(alpha, a.getPointer(), N * sizeof(double));
    if (iterations) *iterations = cg.getNumberIterations();
    if (residuum) *residuum = cg.getResiduum();
  });
}

// LearnerVectorizedPerformanceCalculator::getGFlopAndGByte, unchanged
int ref_gflop_gbyte(void* g, size_t M, int solver, size_t iterations, size_t size_dt, int reuse,
                    double* gflop, double* gbyte) {
  return guard([&] {
    auto p = sgpp::datadriven::LearnerVectorizedPerformanceCalculator::getGFlopAndGByte(
        *static_cast<RefGrid*>(g)->grid, M,
        solver == 0 ? sgpp::solver::SLESolverType::CG : sgpp::solver::SLESolverType::BiCGSTAB,
        iterations, size_dt, reuse != 0, false);
    *gflop = p.GFlop_;
    *gbyte = p.GByte_;
  });
}

// The steps of LearnerBase::train (LearnerBase.cpp:113-327) for a regression
// learner with surplus refinement, on reference parts: the regular grid of
// InitializeGrid (:81-98), the operator kind chosen as in ref_op_new, the
// system of SystemMatrixLeastSquaresIdentity (LeastSquares above; b = B^T y),
// ConjugateGradients with reused coefficients (LearnerBase.cpp:30) and
// SurplusRefinementFunctor + refine. g: an empty grid from ref_grid_new.
// Per pass: grid size, CG iterations, train MSE (when test_acc). Returns the
// passes in *n_steps (arrays sized num_refinements + 1) and alpha in
// alpha_out (capacity alpha_cap, at least the final grid size).
int ref_learner_train(void* g, int kind, const double* X, size_t M, const double* y, int level,
                      double lambda, double eps_refine, size_t it_refine, double eps_final,
                      size_t it_final, size_t num_refinements, double threshold,
                      size_t num_points, int test_acc, double* alpha_out, size_t alpha_cap,
                      size_t* n_steps, size_t* step_size, size_t* step_iterations,
                      double* step_mse) {
  return guard([&] {
    Grid& grid = *static_cast<RefGrid*>(g)->grid;
    const size_t d = grid.getDimension();
    grid.getGenerator().regular(level);
    DataVector alpha(grid.getSize());
    alpha.setAll(0.0);
    DataMatrix data(X, M, d);
    DataVector classes(M);
    std::memcpy(classes.getPointer(), y, M * sizeof(double));
    std::unique_ptr<sgpp::base::OperationMultipleEval> B;
    if (kind == 0)
      B.reset(sgpp::op_factory::createOperationMultipleEval(grid, data));
    else if (grid.getType() == sgpp::base::GridType::Linear)
      B.reset(new sgpp::datadriven::OperationMultiEvalStreaming(grid, data));
    else
      B.reset(new sgpp::datadriven::OperationMultiEvalModMaskStreaming(grid, data));
    LeastSquares A(*B, M, lambda);
    sgpp::solver::ConjugateGradients cg(it_refine, eps_refine);
    double oldAcc = 0.0;
    size_t k = 0;
    for (size_t i = 0; i < num_refinements + 1; ++i) {
      if (i > 0) {
        sgpp::base::SurplusRefinementFunctor f(alpha, num_points, threshold);
        grid.getGenerator().refine(f);
        B->prepare();
        alpha.resizeZero(grid.getSize());
      }
      DataVector b(alpha.getSize());
      DataVector myClasses(classes);
      B->multTranspose(myClasses, b);
      if (i == num_refinements) {
        cg.setMaxIterations(it_final);
        cg.setEpsilon(eps_final);
      }
      cg.solve(A, alpha, b, true, false, 0.0);
      step_size[k] = grid.getSize();
      step_iterations[k] = cg.getNumberIterations();
      step_mse[k] = -1.0;
      ++k;
      if (test_acc) {
        DataVector computed(M);
        B->mult(alpha, computed);
        computed.sub(classes);
        computed.sqr();
        const double acc = computed.sum() / static_cast<double>(M);
        step_mse[k - 1] = acc;
        if (i > 0 && oldAcc <= acc) break;
        oldAcc = acc;
      }
    }
    *n_steps = k;
    if (alpha.getSize() > alpha_cap) throw std::runtime_error("alpha_out too small");
    std::memcpy(alpha_out, alpha.getPointer(), alpha.getSize() * sizeof(double));
  });
}

}  // extern "C"
