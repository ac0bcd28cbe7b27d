// Device-resident regression solve on top of DeviceMultiEval.
#pragma once

#include <cstddef>

#include "meval_device.hpp"

namespace sgpp_hip {

struct SolveOptions {
  bool mixed = false;
  // with mixed == false: CG with FP32 products (b included), FP64 vectors
  bool fp32_products = false;
  double lambda = 0.0;
  size_t max_iterations = 1000;
  double epsilon = 1e-10;
  double inner_epsilon = 1e-5;
  size_t max_refinements = 20;
  bool reuse = false;
};

struct SolveResult {
  size_t iterations = 0;
  size_t refinements = 0;
  size_t fp64_matvecs = 0;
  size_t fp32_matvecs = 0;
  double residual_rel = 0.0;
  double seconds = 0.0;
  // the stopping test was met (FP64/FP32 modes: CG's delta_new <= eps^2
  // delta_0; mixed: the FP64 residual ||r|| <= eps ||b||), not a budget
  bool converged = false;
};

class DeviceRegression {
 public:
  explicit DeviceRegression(DeviceMultiEval& op);
  ~DeviceRegression();
  DeviceRegression(const DeviceRegression&) = delete;
  DeviceRegression& operator=(const DeviceRegression&) = delete;

  SolveResult solve(const double* y_host, double* alpha_host, const SolveOptions& o);
  void system_mult_host(const double* v, double* q, double lambda, Precision p);

 private:
  void apply(const double* v, double* q, double lambda, Precision p);
  double dot(const double* x, const double* y);
  void dot_dev(const double* x, const double* y, double* out);
  const double* read_scalars(const double* s, int n);
  void one_thread(void (*kernel)(double*));
  void axpy(double a, const double* x, double* y);
  void xpby(const double* x, double b, double* y);
  void sub(const double* x, const double* y, double* z);
  void fill(double* y, double v);
  void copy(const double* x, double* y);
  size_t cg(double* x, const double* rhs, double lambda, Precision p, size_t max_it, double eps,
            bool reuse, double* final_delta, double* start_delta, bool skip_zero_apply);

  DeviceMultiEval& op_;
  size_t N_ = 0, M_ = 0;
  void* vec_ = nullptr;
  double* tempM_ = nullptr;
  void* dotpart_ = nullptr;
  void* hostscal_ = nullptr;
  double *alpha_, *b_, *r_, *d_, *q_, *c_, *w_, *corr_, *rhs_;
  double* scal_ = nullptr;  // CG scalars in device memory
  size_t fp64_matvecs_ = 0, fp32_matvecs_ = 0;
};

}  // namespace sgpp_hip
