// Device-resident least-squares solve for (B^T B + lambda*M*I) alpha = B^T y.
//
// FP64 mode follows ConjugateGradients::solve
// (solver/src/sgpp/solver/sle/ConjugateGradients.cpp:22-151) step by step:
// start alpha = 0 unless reused, stop when delta_new <= eps^2 * delta_0,
// recompute the true residual every 50 iterations, break on d.q == 0.
// The system matrix is SystemMatrixLeastSquaresIdentity::mult
// (datadriven/src/sgpp/datadriven/algorithm/SystemMatrixLeastSquaresIdentity.cpp:32-47):
// temp = B v, q = B^T temp, q += M*lambda*v.
//
// Mixed mode: FP64 iterative refinement around FP32-matvec CG. Vectors, dot
// products and the residual r = b - A alpha are FP64; only the products
// inside the inner CG use the FP32-with-FP64-accumulation kernels:
//   r = b - A64 alpha; repeat { solve A32 c = r to inner_eps (CG, FP64
//   vectors); alpha += c; r = b - A64 alpha } until ||r|| <= eps ||b||.
// The inner solves share one budget of max_iterations CG iterations.
// Dot products use a fixed two-stage reduction (deterministic). The CG
// scalars (d.q, the step a, beta) stay on the device; the host reads d.q and
// delta_new once per iteration, for the stopping tests.

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <vector>

#include "meval_device.hpp"
#include "solver_device.hpp"

namespace sgpp_hip {

namespace {

constexpr int kRB = 256;
constexpr int kDotBlocks = 256;

__global__ void axpy_kernel(double a, const double* __restrict__ x, double* __restrict__ y,
                            int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}
// y = x + b*y
__global__ void xpby_kernel(const double* __restrict__ x, double b, double* __restrict__ y,
                            int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i] + b * y[i];
}
// z = x - y
__global__ void sub_kernel(const double* __restrict__ x, const double* __restrict__ y,
                           double* __restrict__ z, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) z[i] = x[i] - y[i];
}
__global__ void fill_kernel(double* __restrict__ y, double v, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] = v;
}
// y += sign * (*a) * x, the step read from device memory
__global__ void axpy_dev_kernel(double sign, const double* __restrict__ a,
                                const double* __restrict__ x, double* __restrict__ y, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] += (sign * a[0]) * x[i];
}
// y = x + (*b) * y
__global__ void xpby_dev_kernel(const double* __restrict__ x, const double* __restrict__ b,
                                double* __restrict__ y, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i] + b[0] * y[i];
}
// CG scalars in device memory
enum { kDeltaNew = 0, kDq = 1, kStep = 2, kDeltaOld = 3, kBeta = 4, kNumScalars = 8 };
// step a = delta_new / (d.q); 0 when d.q == 0 (the host then stops, and the
// iteration changes no iterate: ConjugateGradients breaks before its update)
__global__ void cg_step_kernel(double* __restrict__ s) {
  if (threadIdx.x == 0) s[kStep] = s[kDq] != 0.0 ? s[kDeltaNew] / s[kDq] : 0.0;
}
__global__ void cg_keep_delta_kernel(double* __restrict__ s) {
  if (threadIdx.x == 0) s[kDeltaOld] = s[kDeltaNew];
}
__global__ void cg_beta_kernel(double* __restrict__ s) {
  if (threadIdx.x == 0) s[kBeta] = s[kDeltaNew] / s[kDeltaOld];
}

// stage 1: block b sums the contiguous chunk [b*chunk, (b+1)*chunk)
__global__ __launch_bounds__(kRB) void dot_stage1(const double* __restrict__ x,
                                                  const double* __restrict__ y, int64_t n,
                                                  int64_t chunk, double* __restrict__ part) {
  __shared__ double s[kRB];
  const int64_t b0 = static_cast<int64_t>(blockIdx.x) * chunk;
  const int64_t b1 = min(b0 + chunk, n);
  double acc = 0.0;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += kRB) acc += x[i] * y[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kRB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}
__global__ __launch_bounds__(kRB) void dot_stage2(const double* __restrict__ part, int nb,
                                                  double* __restrict__ out) {
  __shared__ double s[kRB];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += kRB) acc += part[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kRB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + kRB - 1) / kRB); }

}  // namespace

DeviceRegression::DeviceRegression(DeviceMultiEval& op) : op_(op) {
  N_ = op.num_grid();
  M_ = op.num_data();
  if (N_ == 0) throw device_error("solver: grid not set");
  const size_t nvec = 9;
  check_hip(hipMalloc(&vec_, nvec * N_ * sizeof(double)), "hipMalloc(solver vectors)");
  check_hip(hipMalloc(&tempM_, M_ * sizeof(double)), "hipMalloc(solver temp)");
  check_hip(hipMalloc(&dotpart_, (kDotBlocks + 1 + kNumScalars) * sizeof(double)),
            "hipMalloc(dot)");
  check_hip(hipHostMalloc(&hostscal_, kNumScalars * sizeof(double)), "hipHostMalloc");
  alpha_ = static_cast<double*>(vec_);
  b_ = alpha_ + N_;
  r_ = b_ + N_;
  d_ = r_ + N_;
  q_ = d_ + N_;
  c_ = q_ + N_;
  w_ = c_ + N_;
  corr_ = w_ + N_;
  rhs_ = corr_ + N_;
  scal_ = static_cast<double*>(dotpart_) + kDotBlocks + 1;
}

DeviceRegression::~DeviceRegression() {
  (void)hipStreamSynchronize(op_.stream());
  if (vec_) (void)hipFree(vec_);
  if (tempM_) (void)hipFree(tempM_);
  if (dotpart_) (void)hipFree(dotpart_);
  if (hostscal_) (void)hipHostFree(hostscal_);
}

void DeviceRegression::apply(const double* v, double* q, double lambda, Precision p) {
  op_.mult_device(v, tempM_, p);
  op_.mult_transpose_device(tempM_, q, p);
  clear_launch_error();
  hipLaunchKernelGGL(axpy_kernel, dim3(blocks_for(N_)), dim3(kRB), 0, op_.stream(),
                     static_cast<double>(M_) * lambda, v, q, static_cast<int64_t>(N_));
  check_hip(hipGetLastError(), "axpy");
  if (p == Precision::FP64)
    ++fp64_matvecs_;
  else
    ++fp32_matvecs_;
}

// x . y into device memory (out), fixed-order two-stage reduction
void DeviceRegression::dot_dev(const double* x, const double* y, double* out) {
  const int64_t n = static_cast<int64_t>(N_);
  const int64_t chunk = (n + kDotBlocks - 1) / kDotBlocks;
  double* part = static_cast<double*>(dotpart_);
  clear_launch_error();
  hipLaunchKernelGGL(dot_stage1, dim3(kDotBlocks), dim3(kRB), 0, op_.stream(), x, y, n, chunk,
                     part);
  hipLaunchKernelGGL(dot_stage2, dim3(1), dim3(kRB), 0, op_.stream(), part, kDotBlocks, out);
  check_hip(hipGetLastError(), "dot");
}

// copies n device scalars to the host and waits for them (one synchronisation)
const double* DeviceRegression::read_scalars(const double* s, int n) {
  check_hip(hipMemcpyAsync(hostscal_, s, n * sizeof(double), hipMemcpyDeviceToHost, op_.stream()),
            "scalar copy");
  check_hip(hipStreamSynchronize(op_.stream()), "scalar sync");
  return static_cast<const double*>(hostscal_);
}

double DeviceRegression::dot(const double* x, const double* y) {
  double* out = static_cast<double*>(dotpart_) + kDotBlocks;
  dot_dev(x, y, out);
  return read_scalars(out, 1)[0];
}

void DeviceRegression::one_thread(void (*kernel)(double*)) {
  clear_launch_error();
  hipLaunchKernelGGL(kernel, dim3(1), dim3(64), 0, op_.stream(), scal_);
  check_hip(hipGetLastError(), "cg scalars");
}

void DeviceRegression::axpy(double a, const double* x, double* y) {
  clear_launch_error();
  hipLaunchKernelGGL(axpy_kernel, dim3(blocks_for(N_)), dim3(kRB), 0, op_.stream(), a, x, y,
                     static_cast<int64_t>(N_));
  check_hip(hipGetLastError(), "axpy");
}
void DeviceRegression::xpby(const double* x, double b, double* y) {
  clear_launch_error();
  hipLaunchKernelGGL(xpby_kernel, dim3(blocks_for(N_)), dim3(kRB), 0, op_.stream(), x, b, y,
                     static_cast<int64_t>(N_));
  check_hip(hipGetLastError(), "xpby");
}
void DeviceRegression::sub(const double* x, const double* y, double* z) {
  clear_launch_error();
  hipLaunchKernelGGL(sub_kernel, dim3(blocks_for(N_)), dim3(kRB), 0, op_.stream(), x, y, z,
                     static_cast<int64_t>(N_));
  check_hip(hipGetLastError(), "sub");
}
void DeviceRegression::fill(double* y, double v) {
  clear_launch_error();
  hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(N_)), dim3(kRB), 0, op_.stream(), y, v,
                     static_cast<int64_t>(N_));
  check_hip(hipGetLastError(), "fill");
}
void DeviceRegression::copy(const double* x, double* y) {
  check_hip(hipMemcpyAsync(y, x, N_ * sizeof(double), hipMemcpyDeviceToDevice, op_.stream()),
            "copy");
}

// ConjugateGradients::solve on device vectors; solves A x = rhs in place.
// Returns the iteration count; delta values as in the reference.
size_t DeviceRegression::cg(double* x, const double* rhs, double lambda, Precision p,
                            size_t max_it, double eps, bool reuse, double* final_delta,
                            double* start_delta, bool skip_zero_apply) {
  const double eps2 = eps * eps;
  double delta0 = 0.0;
  copy(rhs, r_);
  if (reuse) {
    fill(q_, 0.0);
    apply(q_, w_, lambda, p);
    sub(r_, w_, r_);
    delta0 = dot(r_, r_) * eps2;
  } else {
    fill(x, 0.0);
  }
  // The reference applies A to the zero start vector too (ConjugateGradients.cpp:62);
  // the inner solves of the mixed mode skip that product, whose result is exactly 0.
  if (reuse || !skip_zero_apply) {
    apply(x, w_, lambda, p);
    sub(r_, w_, r_);
  }
  copy(r_, d_);
  dot_dev(r_, r_, scal_ + kDeltaNew);
  double delta_new = read_scalars(scal_ + kDeltaNew, 1)[0];
  if (!reuse) delta0 = delta_new * eps2;
  if (start_delta) *start_delta = reuse ? delta0 / eps2 : delta_new;
  size_t it = 0;
  const unsigned nb = blocks_for(N_);
  const int64_t n = static_cast<int64_t>(N_);
  while (it < max_it && delta_new > delta0) {
    apply(d_, q_, lambda, p);
    dot_dev(d_, q_, scal_ + kDq);
    one_thread(cg_step_kernel);
    clear_launch_error();
    hipLaunchKernelGGL(axpy_dev_kernel, dim3(nb), dim3(kRB), 0, op_.stream(), 1.0,
                       scal_ + kStep, d_, x, n);
    check_hip(hipGetLastError(), "axpy");
    if ((it % 50) == 0 && it > 0) {
      apply(x, w_, lambda, p);
      sub(rhs, w_, r_);
    } else {
      clear_launch_error();
      hipLaunchKernelGGL(axpy_dev_kernel, dim3(nb), dim3(kRB), 0, op_.stream(), -1.0,
                         scal_ + kStep, q_, r_, n);
      check_hip(hipGetLastError(), "axpy");
    }
    one_thread(cg_keep_delta_kernel);
    dot_dev(r_, r_, scal_ + kDeltaNew);
    one_thread(cg_beta_kernel);
    clear_launch_error();
    hipLaunchKernelGGL(xpby_dev_kernel, dim3(nb), dim3(kRB), 0, op_.stream(), r_,
                       scal_ + kBeta, d_, n);
    check_hip(hipGetLastError(), "xpby");
    // d.q and delta_new (kDeltaNew = 0, kDq = 1): one synchronisation per
    // iteration. d.q == 0 stops before the count (the reference's break);
    // the step was 0, so x is unchanged.
    const double* h = read_scalars(scal_, 2);
    if (h[kDq] == 0.0) break;
    delta_new = h[kDeltaNew];
    ++it;
  }
  if (final_delta) *final_delta = delta_new;
  return it;
}

SolveResult DeviceRegression::solve(const double* y_host, double* alpha_host,
                                    const SolveOptions& o) {
  const auto t0 = std::chrono::steady_clock::now();
  fp64_matvecs_ = fp32_matvecs_ = 0;
  SolveResult res;
  const bool plain32 = !o.mixed && o.fp32_products;
  const Precision bp = plain32 ? Precision::FP32 : Precision::FP64;
  // reporting residual: FP64 when the operator holds FP64 data
  const Precision rp = op_.serves(Precision::FP64) ? Precision::FP64 : Precision::FP32;
  // b = B^T y (SystemMatrixLeastSquaresIdentity::generateb)
  check_hip(hipMemcpyAsync(tempM_, y_host, M_ * sizeof(double), hipMemcpyHostToDevice,
                           op_.stream()),
            "copy y");
  op_.mult_transpose_device(tempM_, b_, bp);
  if (o.reuse)
    check_hip(hipMemcpyAsync(alpha_, alpha_host, N_ * sizeof(double), hipMemcpyHostToDevice,
                             op_.stream()),
              "copy alpha");
  const double bnorm = std::sqrt(dot(b_, b_));
  if (!o.mixed) {
    double fd = 0.0, sd = 0.0;
    res.iterations = cg(alpha_, b_, o.lambda, bp, o.max_iterations, o.epsilon, o.reuse, &fd,
                        &sd, false);
    // CG's own test (ConjugateGradients.cpp: delta_new <= eps^2 delta_0)
    res.converged = fd <= o.epsilon * o.epsilon * sd;
  } else {
    // r = b - A64 alpha (alpha = 0 or reused)
    // the outer residual lives in c_, which cg() does not touch
    if (o.reuse) {
      apply(alpha_, q_, o.lambda, Precision::FP64);
      sub(b_, q_, c_);
    } else {
      fill(alpha_, 0.0);
      copy(b_, c_);
    }
    double rnorm = std::sqrt(dot(c_, c_));
    size_t steps = 0;
    // the inner solves share the caller's CG budget (max_iterations in all)
    while (rnorm > o.epsilon * bnorm && steps < o.max_refinements &&
           res.iterations < o.max_iterations) {
      copy(c_, rhs_);
      res.iterations += cg(corr_, rhs_, o.lambda, Precision::FP32,
                           o.max_iterations - res.iterations, o.inner_epsilon, false, nullptr,
                           nullptr, true);
      axpy(1.0, corr_, alpha_);
      apply(alpha_, q_, o.lambda, Precision::FP64);
      sub(b_, q_, c_);
      rnorm = std::sqrt(dot(c_, c_));
      ++steps;
    }
    res.refinements = steps;
    res.converged = rnorm <= o.epsilon * bnorm;
  }
  // final residual ||b - A alpha|| / ||b||
  apply(alpha_, q_, o.lambda, rp);
  sub(b_, q_, c_);
  const double rn = std::sqrt(dot(c_, c_));
  res.residual_rel = bnorm > 0 ? rn / bnorm : rn;
  // the reporting residual is not part of the solve
  if (rp == Precision::FP64)
    --fp64_matvecs_;
  else
    --fp32_matvecs_;
  check_hip(hipMemcpyAsync(alpha_host, alpha_, N_ * sizeof(double), hipMemcpyDeviceToHost,
                           op_.stream()),
            "copy alpha out");
  check_hip(hipStreamSynchronize(op_.stream()), "sync");
  res.fp64_matvecs = fp64_matvecs_;
  res.fp32_matvecs = fp32_matvecs_;
  res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return res;
}

void DeviceRegression::system_mult_host(const double* v, double* q, double lambda, Precision p) {
  check_hip(hipMemcpyAsync(d_, v, N_ * sizeof(double), hipMemcpyHostToDevice, op_.stream()),
            "copy v");
  apply(d_, q_, lambda, p);
  check_hip(hipMemcpyAsync(q, q_, N_ * sizeof(double), hipMemcpyDeviceToHost, op_.stream()),
            "copy q");
  check_hip(hipStreamSynchronize(op_.stream()), "sync");
}

}  // namespace sgpp_hip
