// extern "C" boundary of libsgpp_hip.so (declarations in include/sgpp_hip.h).
// Every entry point catches all exceptions, maps them to a status code and
// keeps the message in a thread-local string for sgpp_hip_last_error().

#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/sgpp_hip.h"
#include "meval_device.hpp"
#include "sgpp/base/grid/Grid.hpp"
#include "sgpp/datadriven/application/LearnerLeastSquaresIdentity.hpp"
#include "sgpp/datadriven/datamining/ModelFittingLeastSquares.hpp"
#include "sgpp/datadriven/operation/OperationMultiEvalHIP.hpp"
#include "solver_device.hpp"

using sgpp::base::DataMatrix;
using sgpp::base::DataVector;
using sgpp::base::Grid;

struct sgpp_hip_grid {
  std::unique_ptr<Grid> grid;
  std::unique_ptr<sgpp::base::IncrementalHierarchisation> hier;
};

struct sgpp_hip_learner {
  std::unique_ptr<sgpp::datadriven::LearnerLeastSquaresIdentity> learner;
};

struct sgpp_hip_fitter {
  std::unique_ptr<sgpp::datadriven::ModelFittingLeastSquares> fitter;
};

struct sgpp_hip_meval {
  // either an owned device operator (array interface) ...
  std::unique_ptr<sgpp_hip::DeviceMultiEval> own;
  // ... or an operator created through the C++ factory for a grid handle
  std::unique_ptr<sgpp::base::OperationMultipleEval> op;
  DataMatrix data;  // the dataset the factory operator refers to
  sgpp_hip::DeviceMultiEval* dev = nullptr;
  sgpp_hip::Precision prec = sgpp_hip::Precision::FP64;
  double duration = -1.0;
};

namespace {

thread_local std::string g_last_error;

sgpp_hip_status fail(sgpp_hip_status s, const std::string& msg) {
  g_last_error = msg;
  return s;
}

template <typename F>
sgpp_hip_status guarded(F&& f) {
  try {
    g_last_error.clear();
    return f();
  } catch (const sgpp::base::factory_exception& e) {
    return fail(SGPP_HIP_ERR_FACTORY, e.what());
  } catch (const sgpp::base::data_exception& e) {
    return fail(SGPP_HIP_ERR_INVALID_ARGUMENT, e.what());
  } catch (const sgpp::base::not_implemented_exception& e) {
    return fail(SGPP_HIP_ERR_NOT_IMPLEMENTED, e.what());
  } catch (const sgpp::base::generation_exception& e) {
    return fail(SGPP_HIP_ERR_GENERATION, e.what());
  } catch (const sgpp::base::operation_exception& e) {
    return fail(SGPP_HIP_ERR_OPERATION, e.what());
  } catch (const sgpp::base::solver_exception& e) {
    return fail(SGPP_HIP_ERR_SOLVER, e.what());
  } catch (const sgpp::base::application_exception& e) {
    return fail(SGPP_HIP_ERR_INVALID_ARGUMENT, e.what());
  } catch (const sgpp_hip::input_error& e) {
    return fail(SGPP_HIP_ERR_INVALID_ARGUMENT, e.what());
  } catch (const sgpp_hip::device_error& e) {
    return fail(SGPP_HIP_ERR_DEVICE, e.what());
  } catch (const std::bad_alloc&) {
    return fail(SGPP_HIP_ERR_DEVICE, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(SGPP_HIP_ERR_INTERNAL, e.what());
  } catch (...) {
    return fail(SGPP_HIP_ERR_INTERNAL, "unknown exception");
  }
}

sgpp_hip::Precision prec_of(int p) {
  if (p == SGPP_HIP_PREC_FP64) return sgpp_hip::Precision::FP64;
  if (p == SGPP_HIP_PREC_FP32_FP64ACC) return sgpp_hip::Precision::FP32;
  throw sgpp::base::data_exception("unknown precision");
}

sgpp::datadriven::HIPOperationConfiguration hip_config_of(const sgpp_hip_params* params) {
  sgpp::datadriven::HIPOperationConfiguration hc;
  if (params) {
    prec_of(params->precision);
    hc.precision = params->precision == SGPP_HIP_PREC_FP64
                       ? sgpp::datadriven::HIPPrecision::FP64
                       : sgpp::datadriven::HIPPrecision::FP32_FP64ACC;
    hc.keepFP64 = params->keep_fp64 != 0;
    hc.device = params->device;
    hc.multGridChunks = params->mult_grid_chunks;
    hc.multTDataSegs = params->multT_data_segs;
    hc.targetBlocks = params->target_blocks;
    if (params->algorithm != SGPP_HIP_ALGO_STREAMING && params->algorithm != SGPP_HIP_ALGO_SUBSPACE)
      throw sgpp::base::factory_exception("unknown algorithm");
    hc.algorithm = params->algorithm == SGPP_HIP_ALGO_SUBSPACE
                       ? sgpp::datadriven::HIPAlgorithm::SUBSPACE
                       : sgpp::datadriven::HIPAlgorithm::STREAMING;
  }
  return hc;
}

sgpp::base::GridType basis_of(int b) {
  if (b == SGPP_HIP_BASIS_LINEAR) return sgpp::base::GridType::Linear;
  if (b == SGPP_HIP_BASIS_MODLINEAR) return sgpp::base::GridType::ModLinear;
  throw sgpp::base::factory_exception("unknown basis / grid type");
}

#define REQUIRE(cond, msg) \
  if (!(cond)) throw sgpp::base::data_exception(msg)

}  // namespace

namespace sgpp_hip {
void set_last_error(const std::string& msg) { g_last_error = msg; }
}  // namespace sgpp_hip

extern "C" {

const char* sgpp_hip_last_error(void) { return g_last_error.c_str(); }
int sgpp_hip_version(void) { return 10000; }

// ------------------------------------------------------------------ grids

sgpp_hip_status sgpp_hip_grid_create(int basis, size_t dim, sgpp_hip_grid** out) {
  return guarded([&] {
    REQUIRE(out, "out is null");
    REQUIRE(dim >= 1 && dim <= 64, "dimension must be in [1, 64]");
    auto* g = new sgpp_hip_grid();
    g->grid.reset(new Grid(basis_of(basis), dim));
    *out = g;
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_grid_destroy(sgpp_hip_grid* g) {
  return guarded([&] {
    delete g;
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_grid_regular(sgpp_hip_grid* g, size_t level) {
  return guarded([&] {
    REQUIRE(g, "grid is null");
    if (g->grid->getSize() != 0)
      throw sgpp::base::generation_exception("regular() needs an empty grid");
    g->grid->getGenerator().regular(level);
    return SGPP_HIP_OK;
  });
}

size_t sgpp_hip_grid_size(const sgpp_hip_grid* g) { return g ? g->grid->getSize() : 0; }
size_t sgpp_hip_grid_dim(const sgpp_hip_grid* g) { return g ? g->grid->getDimension() : 0; }

sgpp_hip_status sgpp_hip_grid_get_level_index(const sgpp_hip_grid* g, uint32_t* level,
                                              uint32_t* index) {
  return guarded([&] {
    REQUIRE(g && level && index, "null argument");
    std::vector<uint32_t> l, i;
    g->grid->getStorage().getLevelIndexArrays(l, i);
    std::memcpy(level, l.data(), l.size() * sizeof(uint32_t));
    std::memcpy(index, i.data(), i.size() * sizeof(uint32_t));
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_grid_get_coordinates(const sgpp_hip_grid* g, double* coords) {
  return guarded([&] {
    REQUIRE(g && coords, "null argument");
    std::vector<double> c;
    g->grid->getStorage().getCoordinateArrays(c);
    std::memcpy(coords, c.data(), c.size() * sizeof(double));
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_grid_get_leaf(const sgpp_hip_grid* g, uint8_t* leaf) {
  return guarded([&] {
    REQUIRE(g && leaf, "null argument");
    auto& st = g->grid->getStorage();
    for (size_t s = 0; s < st.getSize(); ++s) leaf[s] = st.getPoint(s).isLeaf() ? 1 : 0;
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_grid_refine_surplus(sgpp_hip_grid* g, const double* alpha,
                                             size_t n_alpha, size_t refinements_num,
                                             double threshold, size_t* n_added) {
  return guarded([&] {
    REQUIRE(g && alpha, "null argument");
    REQUIRE(n_alpha == g->grid->getSize(), "alpha size must equal the grid size");
    DataVector a(alpha, n_alpha);
    sgpp::base::SurplusRefinementFunctor f(a, refinements_num, threshold);
    const size_t before = g->grid->getSize();
    g->grid->getGenerator().refine(f);
    if (n_added) *n_added = g->grid->getSize() - before;
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_grid_hierarchise(sgpp_hip_grid* g, const double* node_values,
                                          double* alpha) {
  return guarded([&] {
    REQUIRE(g && node_values && alpha, "null argument");
    const size_t n = g->grid->getSize();
    DataVector v(node_values, n);
    sgpp::base::doHierarchisation(g->grid->getType(), g->grid->getStorage(), v);
    std::memcpy(alpha, v.data(), n * sizeof(double));
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_grid_hierarchise_incremental(sgpp_hip_grid* g,
                                                      const double* node_values, double* alpha) {
  return guarded([&] {
    REQUIRE(g && node_values && alpha, "null argument");
    const size_t n = g->grid->getSize();
    if (!g->hier)
      g->hier.reset(
          new sgpp::base::IncrementalHierarchisation(g->grid->getType(), g->grid->getStorage()));
    DataVector v(node_values, n), a;
    g->hier->update(v, a);
    std::memcpy(alpha, a.data(), n * sizeof(double));
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_grid_serialize(const sgpp_hip_grid* g, char* buf, size_t cap,
                                        size_t* len) {
  return guarded([&] {
    REQUIRE(g, "grid is null");
    const std::string s = g->grid->serialize();
    if (len) *len = s.size() + 1;
    if (buf) {
      REQUIRE(cap >= s.size() + 1, "buffer too small");
      std::memcpy(buf, s.c_str(), s.size() + 1);
    }
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_grid_unserialize(const char* text, sgpp_hip_grid** out) {
  return guarded([&] {
    REQUIRE(text && out, "null argument");
    auto* g = new sgpp_hip_grid();
    try {
      g->grid.reset(Grid::unserialize(text));
    } catch (...) {
      delete g;
      throw;
    }
    *out = g;
    return SGPP_HIP_OK;
  });
}

// --------------------------------------------------------------- operator

namespace {
sgpp_hip::LaunchConfig launch_of(const sgpp_hip_params* p) {
  sgpp_hip::LaunchConfig lc;
  if (p) {
    lc.mult_grid_chunks = p->mult_grid_chunks;
    lc.multT_data_segs = p->multT_data_segs;
    lc.target_blocks = p->target_blocks;
  }
  return lc;
}
}  // namespace

sgpp_hip_status sgpp_hip_meval_create(const sgpp_hip_grid_desc* grid, const double* data,
                                      size_t M, size_t d, const sgpp_hip_params* params,
                                      sgpp_hip_meval** out) {
  return guarded([&] {
    REQUIRE(grid && data && out, "null argument");
    REQUIRE(grid->dim == d, "grid dimension does not match the data dimension");
    basis_of(grid->basis);
    const int precision = params ? params->precision : SGPP_HIP_PREC_FP64;
    const sgpp_hip::Precision pr = prec_of(precision);
    unsigned mask = 1u << static_cast<int>(pr);
    if (params && params->keep_fp64) mask |= 1u;
    const int algorithm = params ? params->algorithm : SGPP_HIP_ALGO_STREAMING;
    if (algorithm != SGPP_HIP_ALGO_STREAMING && algorithm != SGPP_HIP_ALGO_SUBSPACE)
      throw sgpp::base::factory_exception("unknown algorithm");
    auto h = std::unique_ptr<sgpp_hip_meval>(new sgpp_hip_meval());
    h->own.reset(new sgpp_hip::DeviceMultiEval(
        data, M, d,
        grid->basis == SGPP_HIP_BASIS_LINEAR ? sgpp_hip::Basis::Linear
                                             : sgpp_hip::Basis::ModLinear,
        mask, params ? params->device : 0, launch_of(params),
        algorithm == SGPP_HIP_ALGO_SUBSPACE ? sgpp_hip::Algorithm::Subspace
                                            : sgpp_hip::Algorithm::Streaming));
    h->dev = h->own.get();
    h->prec = pr;
    *out = h.release();
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_meval_create_for_grid(sgpp_hip_grid* grid, const double* data, size_t M,
                                               const sgpp_hip_params* params,
                                               sgpp_hip_meval** out) {
  return guarded([&] {
    REQUIRE(grid && data && out, "null argument");
    auto h = std::unique_ptr<sgpp_hip_meval>(new sgpp_hip_meval());
    const size_t d = grid->grid->getDimension();
    h->data = DataMatrix(data, M, d);
    sgpp::datadriven::HIPOperationConfiguration hc = hip_config_of(params);
    // the subspace algorithm through the factory's SUBSPACELINEAR branch
    const bool subspace = hc.algorithm == sgpp::datadriven::HIPAlgorithm::SUBSPACE;
    hc.algorithm = sgpp::datadriven::HIPAlgorithm::STREAMING;
    sgpp::datadriven::OperationMultipleEvalConfiguration cfg(
        subspace ? sgpp::datadriven::OperationMultipleEvalType::SUBSPACELINEAR
                 : sgpp::datadriven::OperationMultipleEvalType::STREAMING,
        sgpp::datadriven::OperationMultipleEvalSubType::HIP, hc, "HIP");
    h->op.reset(sgpp::op_factory::createOperationMultipleEval(*grid->grid, h->data, cfg));
    auto* hop = dynamic_cast<sgpp::datadriven::OperationMultiEvalHIP*>(h->op.get());
    h->dev = &hop->device();
    h->prec = hc.precision == sgpp::datadriven::HIPPrecision::FP64 ? sgpp_hip::Precision::FP64
                                                                    : sgpp_hip::Precision::FP32;
    *out = h.release();
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_meval_prepare(sgpp_hip_meval* h, const uint32_t* level,
                                       const uint32_t* index, size_t N) {
  return guarded([&] {
    REQUIRE(h && level && index, "null argument");
    if (h->op)
      throw sgpp::base::operation_exception(
          "operator is bound to a grid handle: use sgpp_hip_meval_prepare_grid");
    h->dev->set_grid(level, index, N);
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_meval_prepare_grid(sgpp_hip_meval* h) {
  return guarded([&] {
    REQUIRE(h, "null argument");
    if (!h->op) throw sgpp::base::operation_exception("operator has no grid handle");
    h->op->prepare();
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_meval_mult(sgpp_hip_meval* h, const double* alpha, double* result) {
  return guarded([&] {
    REQUIRE(h && alpha && result, "null argument");
    if (h->op) {
      DataVector a(alpha, h->dev->num_grid());
      DataVector r(h->dev->num_data());
      h->op->mult(a, r);
      std::memcpy(result, r.data(), r.size() * sizeof(double));
      h->duration = h->op->getDuration();
    } else {
      h->dev->mult_host(alpha, result, h->prec);
      h->duration = h->dev->last_wall_s();
    }
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_meval_mult_transpose(sgpp_hip_meval* h, const double* source,
                                              double* result) {
  return guarded([&] {
    REQUIRE(h && source && result, "null argument");
    if (h->op) {
      DataVector y(source, h->dev->num_data());
      DataVector a(h->dev->num_grid());
      h->op->multTranspose(y, a);
      std::memcpy(result, a.data(), a.size() * sizeof(double));
      h->duration = h->op->getDuration();
    } else {
      h->dev->mult_transpose_host(source, result, h->prec);
      h->duration = h->dev->last_wall_s();
    }
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_meval_mult_range(sgpp_hip_meval* h, const double* alpha, size_t start,
                                          size_t end, double* result) {
  return guarded([&] {
    REQUIRE(h && alpha && (result || start == end), "null argument");
    const size_t M = h->dev->num_data();
    REQUIRE(start <= end && end <= M, "invalid data range");
    DataVector r;
    if (h->op) {
      DataVector a(alpha, h->dev->num_grid());
      h->op->mult(a, r, start, end);
      h->duration = h->op->getDuration();
    } else {
      DataVector full(M);
      h->dev->mult_host(alpha, full.data(), h->prec);
      h->duration = h->dev->last_wall_s();
      r.assign(full.begin() + start, full.begin() + end);
    }
    if (end > start) std::memcpy(result, r.data(), (end - start) * sizeof(double));
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_meval_mult_transpose_range(sgpp_hip_meval* h, const double* source,
                                                    size_t start, size_t end, double* result) {
  return guarded([&] {
    REQUIRE(h && source && (result || start == end), "null argument");
    const size_t N = h->dev->num_grid();
    REQUIRE(start <= end && end <= N, "invalid grid range");
    DataVector r;
    if (h->op) {
      DataVector y(source, h->dev->num_data());
      h->op->multTranspose(y, r, start, end);
      h->duration = h->op->getDuration();
    } else {
      DataVector full(N);
      h->dev->mult_transpose_host(source, full.data(), h->prec);
      h->duration = h->dev->last_wall_s();
      r.assign(full.begin() + start, full.begin() + end);
    }
    if (end > start) std::memcpy(result, r.data(), (end - start) * sizeof(double));
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_meval_mult_device(sgpp_hip_meval* h, const double* d_alpha,
                                           double* d_result, int precision) {
  return guarded([&] {
    REQUIRE(h && d_alpha && d_result, "null argument");
    h->dev->mult_device(d_alpha, d_result, prec_of(precision));
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_meval_mult_device_noalloc(sgpp_hip_meval* h, const double* d_alpha,
                                                   double* d_result, int precision) {
  return guarded([&] {
    REQUIRE(h && d_alpha && d_result, "null argument");
    h->dev->mult_device_noalloc(d_alpha, d_result, prec_of(precision));
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_meval_mult_transpose_device(sgpp_hip_meval* h, const double* d_source,
                                                     double* d_result, int precision) {
  return guarded([&] {
    REQUIRE(h && d_source && d_result, "null argument");
    h->dev->mult_transpose_device(d_source, d_result, prec_of(precision));
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_meval_synchronize(sgpp_hip_meval* h) {
  return guarded([&] {
    REQUIRE(h, "null argument");
    h->dev->synchronize();
    return SGPP_HIP_OK;
  });
}

double sgpp_hip_meval_duration(const sgpp_hip_meval* h) { return h ? h->duration : -1.0; }

double sgpp_hip_meval_kernel_ms(sgpp_hip_meval* h) {
  if (!h) return -1.0;
  try {
    h->dev->synchronize();
    return h->dev->last_kernel_ms();
  } catch (...) {
    return -1.0;
  }
}

void* sgpp_hip_meval_stream(sgpp_hip_meval* h) {
  return h ? reinterpret_cast<void*>(h->dev->stream()) : nullptr;
}
size_t sgpp_hip_meval_num_data(const sgpp_hip_meval* h) { return h ? h->dev->num_data() : 0; }
size_t sgpp_hip_meval_num_grid(const sgpp_hip_meval* h) { return h ? h->dev->num_grid() : 0; }

sgpp_hip_status sgpp_hip_meval_last_shape(const sgpp_hip_meval* h, int* mult_chunks,
                                          int* multT_segs) {
  return guarded([&] {
    REQUIRE(h, "null argument");
    if (mult_chunks) *mult_chunks = h->dev->last_mult_chunks();
    if (multT_segs) *multT_segs = h->dev->last_multT_segs();
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_meval_kernel_timing_enable(sgpp_hip_meval* h, int on) {
  return guarded([&] {
    REQUIRE(h, "null argument");
    h->dev->set_kernel_timing(on != 0);
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_meval_kernel_timing(sgpp_hip_meval* h, int* n_mult, double* ms_mult,
                                             int* n_multT, double* ms_multT) {
  return guarded([&] {
    REQUIRE(h, "null argument");
    h->dev->kernel_timing(n_mult, ms_mult, n_multT, ms_multT);
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_meval_count_work(sgpp_hip_meval* h, int precision, uint64_t* out4) {
  return guarded([&] {
    REQUIRE(h && out4, "null argument");
    REQUIRE(precision == SGPP_HIP_PREC_FP64 || precision == SGPP_HIP_PREC_FP32_FP64ACC,
            "unknown precision");
    h->dev->count_work(precision == SGPP_HIP_PREC_FP64 ? sgpp_hip::Precision::FP64
                                                       : sgpp_hip::Precision::FP32,
                       out4);
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_meval_destroy(sgpp_hip_meval* h) {
  return guarded([&] {
    delete h;
    return SGPP_HIP_OK;
  });
}

// ----------------------------------------------------------------- solver

sgpp_hip_status sgpp_hip_regression_solve(sgpp_hip_meval* h, const double* y,
                                          const sgpp_hip_solve_params* p, double* alpha,
                                          sgpp_hip_solve_info* info) {
  return guarded([&] {
    REQUIRE(h && y && p && alpha, "null argument");
    REQUIRE(p->mode == SGPP_HIP_SOLVE_FP64 || p->mode == SGPP_HIP_SOLVE_MIXED ||
                p->mode == SGPP_HIP_SOLVE_FP32,
            "unknown mode");
    if (p->mode != SGPP_HIP_SOLVE_FP32)
      REQUIRE(h->dev->serves(sgpp_hip::Precision::FP64),
              "the solver needs an operator that holds FP64 data (precision FP64 or keep_fp64)");
    if (p->mode != SGPP_HIP_SOLVE_FP64)
      REQUIRE(h->dev->serves(sgpp_hip::Precision::FP32),
              "this mode needs an operator that holds FP32 data");
    sgpp_hip::SolveOptions o;
    o.mixed = p->mode == SGPP_HIP_SOLVE_MIXED;
    o.fp32_products = p->mode == SGPP_HIP_SOLVE_FP32;
    o.lambda = p->lambda;
    o.max_iterations = p->max_iterations;
    o.epsilon = p->epsilon;
    o.inner_epsilon = p->inner_epsilon > 0 ? p->inner_epsilon : 1e-5;
    o.max_refinements = p->max_refinements > 0 ? p->max_refinements : 20;
    o.reuse = p->reuse != 0;
    sgpp_hip::DeviceRegression solver(*h->dev);
    const sgpp_hip::SolveResult r = solver.solve(y, alpha, o);
    if (info) {
      info->iterations = r.iterations;
      info->refinements = r.refinements;
      info->fp64_matvecs = r.fp64_matvecs;
      info->fp32_matvecs = r.fp32_matvecs;
      info->residual_rel = r.residual_rel;
      info->seconds = r.seconds;
      info->converged = r.converged ? 1 : 0;
    }
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_system_matrix_mult(sgpp_hip_meval* h, double lambda, const double* v,
                                            double* q, int precision) {
  return guarded([&] {
    REQUIRE(h && v && q, "null argument");
    sgpp_hip::DeviceRegression solver(*h->dev);
    solver.system_mult_host(v, q, lambda, prec_of(precision));
    return SGPP_HIP_OK;
  });
}

// ---------------------------------------------------------------- learner

namespace {

sgpp::solver::SLESolverConfiguration solver_config_of(const sgpp_hip_solver_config* c) {
  sgpp::solver::SLESolverConfiguration s;
  if (c->type == SGPP_HIP_SLE_CG)
    s.type_ = sgpp::solver::SLESolverType::CG;
  else if (c->type == SGPP_HIP_SLE_BICGSTAB)
    s.type_ = sgpp::solver::SLESolverType::BiCGSTAB;
  else
    throw sgpp::base::data_exception("unknown solver type");
  s.eps_ = c->eps;
  s.maxIterations_ = c->max_iterations;
  s.threshold_ = c->threshold;
  return s;
}

}  // namespace

sgpp_hip_status sgpp_hip_learner_create(int is_regression, const sgpp_hip_params* params,
                                        sgpp_hip_learner** out) {
  return guarded([&] {
    REQUIRE(out, "out is null");
    auto l = std::unique_ptr<sgpp_hip_learner>(new sgpp_hip_learner());
    l->learner.reset(new sgpp::datadriven::LearnerLeastSquaresIdentity(is_regression != 0));
    l->learner->setImplementation(sgpp::datadriven::OperationMultipleEvalConfiguration(
        sgpp::datadriven::OperationMultipleEvalType::STREAMING,
        sgpp::datadriven::OperationMultipleEvalSubType::HIP, hip_config_of(params), "learner"));
    *out = l.release();
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_learner_destroy(sgpp_hip_learner* l) {
  return guarded([&] {
    delete l;
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_learner_train(sgpp_hip_learner* l, const double* data, size_t M,
                                       size_t d, const double* classes,
                                       const sgpp_hip_grid_config* grid,
                                       const sgpp_hip_solver_config* solver_refine,
                                       const sgpp_hip_solver_config* solver_final,
                                       const sgpp_hip_adaptivity_config* adaptivity,
                                       int test_acc_during_adapt, double lambda,
                                       sgpp_hip_learner_timing* timing) {
  return guarded([&] {
    REQUIRE(l && data && classes && grid && solver_refine && solver_final && adaptivity,
            "null argument");
    REQUIRE(M > 0 && d >= 1 && d <= 64, "bad data shape");
    DataMatrix X(data, M, d);
    DataVector y(classes, M);
    sgpp::base::RegularGridConfiguration gc;
    gc.type_ = basis_of(grid->basis);
    gc.dim_ = grid->dim;
    gc.level_ = grid->level;
    sgpp::base::AdaptivityConfiguration ac;
    ac.numRefinements_ = adaptivity->num_refinements;
    ac.refinementThreshold_ = adaptivity->refinement_threshold;
    ac.numRefinementPoints_ = adaptivity->num_refinement_points;
    ac.errorBasedRefinement_ = adaptivity->error_based_refinement != 0;
    const sgpp::datadriven::LearnerTiming t = l->learner->train(
        X, y, gc, solver_config_of(solver_refine), solver_config_of(solver_final), ac,
        test_acc_during_adapt != 0, lambda);
    if (timing) {
      timing->time_complete = t.timeComplete_;
      timing->time_mult_complete = t.timeMultComplete_;
      timing->time_mult_compute = t.timeMultCompute_;
      timing->time_mult_trans_complete = t.timeMultTransComplete_;
      timing->time_mult_trans_compute = t.timeMultTransCompute_;
      timing->time_regularization = t.timeRegularization_;
      timing->gflop = t.GFlop_;
      timing->gbyte = t.GByte_;
    }
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_learner_predict(sgpp_hip_learner* l, const double* data, size_t M,
                                         size_t d, double* out) {
  return guarded([&] {
    REQUIRE(l && data && out, "null argument");
    REQUIRE(l->learner->isTrained(), "learner is not trained");
    REQUIRE(d == l->learner->getGrid().getDimension(), "data dimension differs from the grid's");
    DataMatrix X(data, M, d);
    DataVector r;
    l->learner->predict(X, r);
    std::memcpy(out, r.data(), M * sizeof(double));
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_learner_accuracy(sgpp_hip_learner* l, const double* data, size_t M,
                                          size_t d, const double* classes, double threshold,
                                          double* accuracy) {
  return guarded([&] {
    REQUIRE(l && data && classes && accuracy, "null argument");
    REQUIRE(l->learner->isTrained(), "learner is not trained");
    REQUIRE(d == l->learner->getGrid().getDimension(), "data dimension differs from the grid's");
    DataMatrix X(data, M, d);
    DataVector y(classes, M);
    *accuracy = l->learner->getAccuracy(X, y, threshold);
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_learner_grid_size(sgpp_hip_learner* l, size_t* n) {
  return guarded([&] {
    REQUIRE(l && n, "null argument");
    *n = l->learner->isTrained() ? l->learner->getGrid().getSize() : 0;
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_learner_level_index(sgpp_hip_learner* l, uint32_t* level,
                                             uint32_t* index) {
  return guarded([&] {
    REQUIRE(l && level && index, "null argument");
    REQUIRE(l->learner->isTrained(), "learner is not trained");
    std::vector<uint32_t> lv, ix;
    l->learner->getGrid().getStorage().getLevelIndexArrays(lv, ix);
    std::memcpy(level, lv.data(), lv.size() * sizeof(uint32_t));
    std::memcpy(index, ix.data(), ix.size() * sizeof(uint32_t));
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_learner_alpha(sgpp_hip_learner* l, double* alpha) {
  return guarded([&] {
    REQUIRE(l && alpha, "null argument");
    REQUIRE(l->learner->isTrained(), "learner is not trained");
    const DataVector& a = l->learner->getAlpha();
    std::memcpy(alpha, a.data(), a.size() * sizeof(double));
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_learner_num_steps(sgpp_hip_learner* l, size_t* n) {
  return guarded([&] {
    REQUIRE(l && n, "null argument");
    *n = l->learner->getSteps().size();
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_learner_step_info(sgpp_hip_learner* l, size_t k,
                                           sgpp_hip_learner_step* out) {
  return guarded([&] {
    REQUIRE(l && out, "null argument");
    const auto& steps = l->learner->getSteps();
    REQUIRE(k < steps.size(), "step index out of range");
    out->grid_size = steps[k].gridSize;
    out->iterations = steps[k].iterations;
    out->refinements = steps[k].refinements;
    out->residual = steps[k].residual;
    out->train_accuracy = steps[k].trainAccuracy;
    out->seconds = steps[k].seconds;
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_fitter_create(const sgpp_hip_fitter_config* cfg,
                                       const sgpp_hip_params* params, sgpp_hip_fitter** out) {
  return guarded([&] {
    REQUIRE(cfg && out, "null argument");
    REQUIRE(cfg->refinement.error_based_refinement == 0, "error-based refinement is not provided");
    sgpp::datadriven::FitterConfigurationLeastSquares c;
    c.grid.type_ = basis_of(cfg->grid.basis);
    c.grid.level_ = cfg->grid.level;
    c.refinement.numRefinements_ = cfg->refinement.num_refinements;
    c.refinement.refinementThreshold_ = cfg->refinement.refinement_threshold;
    c.refinement.numRefinementPoints_ = cfg->refinement.num_refinement_points;
    c.solverRefine = solver_config_of(&cfg->solver_refine);
    c.solverFinal = solver_config_of(&cfg->solver_final);
    c.lambda = cfg->lambda;
    c.multipleEval = sgpp::datadriven::OperationMultipleEvalConfiguration(
        sgpp::datadriven::OperationMultipleEvalType::STREAMING,
        sgpp::datadriven::OperationMultipleEvalSubType::HIP, hip_config_of(params), "fitter");
    auto f = std::unique_ptr<sgpp_hip_fitter>(new sgpp_hip_fitter());
    f->fitter.reset(new sgpp::datadriven::ModelFittingLeastSquares(c));
    *out = f.release();
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_fitter_destroy(sgpp_hip_fitter* f) {
  return guarded([&] {
    delete f;
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_fitter_fit(sgpp_hip_fitter* f, const double* data, size_t M, size_t d,
                                    const double* targets) {
  return guarded([&] {
    REQUIRE(f && data && targets, "null argument");
    REQUIRE(M > 0 && d >= 1 && d <= 64, "bad data shape");
    f->fitter->fit(DataMatrix(data, M, d), DataVector(targets, M));
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_fitter_adapt(sgpp_hip_fitter* f, int* refined) {
  return guarded([&] {
    REQUIRE(f && refined, "null argument");
    REQUIRE(f->fitter->isFitted(), "fitter has no model yet");
    *refined = f->fitter->adapt() ? 1 : 0;
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_fitter_update(sgpp_hip_fitter* f, const double* data, size_t M,
                                       size_t d, const double* targets) {
  return guarded([&] {
    REQUIRE(f && data && targets, "null argument");
    REQUIRE(M > 0 && d >= 1 && d <= 64, "bad data shape");
    REQUIRE(!f->fitter->isFitted() || d == f->fitter->getGrid().getDimension(),
            "data dimension differs from the grid's");
    f->fitter->update(DataMatrix(data, M, d), DataVector(targets, M));
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_fitter_evaluate(sgpp_hip_fitter* f, const double* data, size_t M,
                                         size_t d, double* out) {
  return guarded([&] {
    REQUIRE(f && data && out, "null argument");
    REQUIRE(f->fitter->isFitted(), "fitter has no model yet");
    REQUIRE(M > 0 && d == f->fitter->getGrid().getDimension(),
            "data dimension differs from the grid's");
    DataMatrix X(data, M, d);
    DataVector r;
    f->fitter->evaluate(X, r);
    std::memcpy(out, r.data(), M * sizeof(double));
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_fitter_grid_size(sgpp_hip_fitter* f, size_t* n) {
  return guarded([&] {
    REQUIRE(f && n, "null argument");
    *n = f->fitter->isFitted() ? f->fitter->getGrid().getSize() : 0;
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_fitter_level_index(sgpp_hip_fitter* f, uint32_t* level,
                                            uint32_t* index) {
  return guarded([&] {
    REQUIRE(f && level && index, "null argument");
    REQUIRE(f->fitter->isFitted(), "fitter has no model yet");
    std::vector<uint32_t> lv, ix;
    f->fitter->getGrid().getStorage().getLevelIndexArrays(lv, ix);
    std::memcpy(level, lv.data(), lv.size() * sizeof(uint32_t));
    std::memcpy(index, ix.data(), ix.size() * sizeof(uint32_t));
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_fitter_alpha(sgpp_hip_fitter* f, double* alpha) {
  return guarded([&] {
    REQUIRE(f && alpha, "null argument");
    REQUIRE(f->fitter->isFitted(), "fitter has no model yet");
    const DataVector& a = f->fitter->getSurpluses();
    std::memcpy(alpha, a.data(), a.size() * sizeof(double));
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_fitter_info(sgpp_hip_fitter* f, size_t* last_iterations,
                                     size_t* refinements) {
  return guarded([&] {
    REQUIRE(f && last_iterations && refinements, "null argument");
    *last_iterations = f->fitter->getLastIterations();
    *refinements = f->fitter->getRefinementsPerformed();
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_learner_performance(sgpp_hip_grid* grid, size_t num_instances,
                                             int solver_type, size_t num_iterations,
                                             size_t size_datatype, int reuse_alpha,
                                             double* gflop, double* gbyte) {
  return guarded([&] {
    REQUIRE(grid && gflop && gbyte, "null argument");
    sgpp_hip_solver_config c{solver_type, 0.0, 0, 0.0};
    const auto p = sgpp::datadriven::LearnerVectorizedPerformanceCalculator::getGFlopAndGByte(
        *grid->grid, num_instances, solver_config_of(&c).type_, num_iterations, size_datatype,
        reuse_alpha != 0);
    *gflop = p.GFlop_;
    *gbyte = p.GByte_;
    return SGPP_HIP_OK;
  });
}

}  // extern "C"
