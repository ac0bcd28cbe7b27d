/*
 * ORACLE -- test infrastructure only. Used by tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg as the checker / CPU baseline; never part
 * of the shipped product path.
 *
 * CPU restatement (plain C, OpenMP) of SG++'s multi-evaluation operators:
 *
 *  oracle_mult / oracle_mult_transpose
 *      The definition B_ij = prod_d phi(l_jd, i_jd, x_id) with
 *      phi = LinearBasis::eval (base/src/sgpp/base/operation/hash/common/
 *      basis/LinearBasis.hpp:37-41) or LinearModifiedBasis::eval
 *      (LinearModifiedBasis.hpp:36-53), summed in long double. This equals
 *      the reference's DEFAULT operator (AlgorithmMultipleEvaluation /
 *      AlgorithmDGEMV) on data inside the unit cube.
 *
 *  port_streaming_mult / port_streaming_mult_transpose
 *      The reference's CPU StreamingDefault operators
 *      (datadriven/src/sgpp/datadriven/operation/hash/OperationMultiEvalStreaming/
 *      OperationMultiEvalStreaming.cpp:92-143, _multImpl.cpp:576-623,
 *      _multTransposeImpl.cpp:483-508; ModLinear:
 *      OperationMultiEvalModMaskStreaming_multImpl.cpp:604-656 with the
 *      encoding of HashGridStorage.cpp:519-560): FP64, 12 grid points x
 *      DATA_CHUNK data points blocking, OpenMP over data chunks for mult and
 *      over grid points for multTranspose. Timed as the CPU baseline when
 *      the reference build is unavailable.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

enum { BASIS_LINEAR = 0, BASIS_MODLINEAR = 1 };

/* LinearBasis::eval */
static double phi_linear(uint32_t l, uint32_t i, double x) {
  const double v = 1.0 - fabs((double)(1u << l) * x - (double)i);
  return v > 0.0 ? v : 0.0;
}

/* LinearModifiedBasis::eval */
static double phi_modlinear(uint32_t l, uint32_t i, double x) {
  const uint32_t hinv = 1u << l;
  const double h = (double)hinv;
  if (l == 1) return 1.0;
  if (i == 1) return (x <= 2.0 / h) ? (2.0 - h * x) : 0.0;
  if (i == hinv - 1) return (x >= 1.0 - 2.0 / h) ? (h * x - (double)i + 1.0) : 0.0;
  {
    const double v = 1.0 - fabs(h * x - (double)i);
    return v > 0.0 ? v : 0.0;
  }
}

static double basis_value(int basis, size_t d, const double* x, const uint32_t* lv,
                          const uint32_t* ix) {
  double v = 1.0;
  size_t k;
  for (k = 0; k < d; ++k) {
    v *= (basis == BASIS_LINEAR) ? phi_linear(lv[k], ix[k], x[k]) : phi_modlinear(lv[k], ix[k], x[k]);
    if (v == 0.0) break;
  }
  return v;
}

/* r_i = sum_j alpha_j B_ij; X is M x d row-major; lv/ix N x d row-major */
void oracle_mult(int basis, size_t d, size_t M, const double* X, size_t N, const uint32_t* lv,
                 const uint32_t* ix, const double* alpha, double* r) {
  long i;
#pragma omp parallel for schedule(dynamic, 64)
  for (i = 0; i < (long)M; ++i) {
    long double acc = 0.0L;
    size_t j;
    for (j = 0; j < N; ++j) {
      const double b = basis_value(basis, d, X + (size_t)i * d, lv + j * d, ix + j * d);
      if (b != 0.0) acc += (long double)alpha[j] * (long double)b;
    }
    r[i] = (double)acc;
  }
}

/* a_j = sum_i y_i B_ij */
void oracle_mult_transpose(int basis, size_t d, size_t M, const double* X, size_t N,
                           const uint32_t* lv, const uint32_t* ix, const double* y, double* a) {
  long j;
#pragma omp parallel for schedule(dynamic, 16)
  for (j = 0; j < (long)N; ++j) {
    long double acc = 0.0L;
    size_t i;
    for (i = 0; i < M; ++i) {
      const double b = basis_value(basis, d, X + i * d, lv + (size_t)j * d, ix + (size_t)j * d);
      if (b != 0.0) acc += (long double)y[i] * (long double)b;
    }
    a[j] = (double)acc;
  }
}

/* single basis value, for tests of the closed forms */
double oracle_phi(int basis, uint32_t l, uint32_t i, double x) {
  return basis == BASIS_LINEAR ? phi_linear(l, i, x) : phi_modlinear(l, i, x);
}

/* ------------------------------------------------------ StreamingDefault port */

#define GRID_CHUNK 12
#ifndef DATA_CHUNK
#define DATA_CHUNK 24
#endif

typedef struct {
  int basis;
  size_t d, M, Mpad, N;
  double* data; /* d x Mpad, padded by repeating the last row */
  double* L;    /* N x d */
  double* I;    /* N x d */
  uint64_t* mask;
  double* offset;
} port_op;

void* port_create(int basis, size_t d, size_t M, const double* X, size_t N, const uint32_t* lv,
                  const uint32_t* ix) {
  port_op* op = (port_op*)calloc(1, sizeof(port_op));
  size_t i, k, j;
  op->basis = basis;
  op->d = d;
  op->M = M;
  op->Mpad = ((M + DATA_CHUNK - 1) / DATA_CHUNK) * DATA_CHUNK;
  op->N = N;
  op->data = (double*)malloc(sizeof(double) * d * op->Mpad);
  for (k = 0; k < d; ++k)
    for (i = 0; i < op->Mpad; ++i) op->data[k * op->Mpad + i] = X[(i < M ? i : M - 1) * d + k];
  op->L = (double*)malloc(sizeof(double) * N * d);
  op->I = (double*)malloc(sizeof(double) * N * d);
  op->mask = (uint64_t*)malloc(sizeof(uint64_t) * N * d);
  op->offset = (double*)malloc(sizeof(double) * N * d);
  for (j = 0; j < N; ++j)
    for (k = 0; k < d; ++k) {
      const uint32_t l = lv[j * d + k], ii = ix[j * d + k];
      const size_t e = j * d + k;
      if (basis == BASIS_LINEAR) {
        op->L[e] = (double)(1u << l);
        op->I[e] = (double)ii;
        op->mask[e] = 0x8000000000000000ull;
        op->offset[e] = 1.0;
      } else if (l == 1) {
        op->L[e] = 0.0;
        op->I[e] = 0.0;
        op->mask[e] = 0;
        op->offset[e] = 1.0;
      } else if (ii == 1) {
        op->L[e] = -(double)(1u << l);
        op->I[e] = 0.0;
        op->mask[e] = 0;
        op->offset[e] = 2.0;
      } else if (ii == (1u << l) - 1) {
        op->L[e] = (double)(1u << l);
        op->I[e] = (double)ii;
        op->mask[e] = 0;
        op->offset[e] = 1.0;
      } else {
        op->L[e] = (double)(1u << l);
        op->I[e] = (double)ii;
        op->mask[e] = 0x8000000000000000ull;
        op->offset[e] = 1.0;
      }
    }
  return op;
}

void port_destroy(void* p) {
  port_op* op = (port_op*)p;
  if (!op) return;
  free(op->data);
  free(op->L);
  free(op->I);
  free(op->mask);
  free(op->offset);
  free(op);
}

/* factor: Linear  max(1 - |L x - I|, 0); ModLinear max((Lx - I | mask) + offset, 0) */
static inline double port_factor(const port_op* op, size_t e, double x) {
  double t = op->L[e] * x - op->I[e];
  uint64_t u;
  memcpy(&u, &t, 8);
  u |= op->mask[e];
  memcpy(&t, &u, 8);
  t += op->offset[e];
  return t > 0.0 ? t : 0.0;
}

/* OpenMP partition of [0, n) into contiguous blocks of `block` (getPartitionSegment) */
static void partition(size_t n, size_t block, size_t* s, size_t* e) {
  size_t threads = 1, me = 0;
#ifdef _OPENMP
  threads = (size_t)omp_get_num_threads();
  me = (size_t)omp_get_thread_num();
#endif
  {
    const size_t blocks = n / block;
    size_t per = blocks / threads;
    const size_t rem = blocks - per * threads;
    size_t off;
    if (me < rem) {
      per++;
      off = per * me;
    } else {
      off = rem * (per + 1) + (me - rem) * per;
    }
    *s = off * block;
    *e = *s + per * block;
  }
}

void port_mult(void* p, const double* alpha, double* result) {
  port_op* op = (port_op*)p;
  double* r = (double*)calloc(op->Mpad, sizeof(double));
#pragma omp parallel
  {
    size_t s, e, c, m, i, j, k;
    partition(op->Mpad, DATA_CHUNK, &s, &e);
    for (c = s; c < e; c += DATA_CHUNK) {
      for (m = 0; m < op->N; m += GRID_CHUNK) {
        const size_t gend = m + GRID_CHUNK < op->N ? m + GRID_CHUNK : op->N;
        for (j = m; j < gend; ++j) {
          double sup[DATA_CHUNK];
          for (i = 0; i < DATA_CHUNK; ++i) sup[i] = alpha[j];
          for (k = 0; k < op->d; ++k) {
            const size_t eidx = j * op->d + k;
            const double* xs = op->data + k * op->Mpad + c;
            for (i = 0; i < DATA_CHUNK; ++i) sup[i] *= port_factor(op, eidx, xs[i]);
          }
          for (i = 0; i < DATA_CHUNK; ++i) r[c + i] += sup[i];
        }
      }
    }
  }
  memcpy(result, r, op->M * sizeof(double));
  free(r);
}

void port_mult_transpose(void* p, const double* source, double* result) {
  port_op* op = (port_op*)p;
  double* y = (double*)calloc(op->Mpad, sizeof(double));
  memcpy(y, source, op->M * sizeof(double));
#pragma omp parallel
  {
    size_t s, e, j, i, k;
    partition(op->N, 1, &s, &e);
    for (j = s; j < e; ++j) {
      double acc = 0.0;
      for (i = 0; i < op->Mpad; i += DATA_CHUNK) {
        double sup[DATA_CHUNK];
        size_t q;
        for (q = 0; q < DATA_CHUNK; ++q) sup[q] = y[i + q];
        for (k = 0; k < op->d; ++k) {
          const size_t eidx = j * op->d + k;
          const double* xs = op->data + k * op->Mpad + i;
          for (q = 0; q < DATA_CHUNK; ++q) sup[q] *= port_factor(op, eidx, xs[q]);
        }
        for (q = 0; q < DATA_CHUNK; ++q) acc += sup[q];
      }
      result[j] = acc;
    }
  }
  free(y);
}

int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
