/*
 * C-ABI of the gfx950 multi-evaluation library (libsgpp_hip.so).
 *
 * Plain pointers and sizes only. Every call returns an sgpp_hip_status;
 * on failure sgpp_hip_last_error() returns a thread-local message. A C++
 * wrapper maps the codes back to SG++'s exception types
 * (base/src/sgpp/base/exception/).
 *
 * Array conventions: data points are M x d row-major FP64 (the layout of
 * sgpp::base::DataMatrix); grid level/index arrays are N x d row-major
 * uint32 in the grid storage's sequence order; all vectors are FP64.
 */
#ifndef SGPP_HIP_H_
#define SGPP_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum sgpp_hip_status {
  SGPP_HIP_OK = 0,
  SGPP_HIP_ERR_INVALID_ARGUMENT = 1, /* sgpp::base::data_exception, bad size/pointer */
  SGPP_HIP_ERR_FACTORY = 2,          /* sgpp::base::factory_exception */
  SGPP_HIP_ERR_OPERATION = 3,        /* sgpp::base::operation_exception */
  SGPP_HIP_ERR_NOT_IMPLEMENTED = 4,  /* sgpp::base::not_implemented_exception */
  SGPP_HIP_ERR_GENERATION = 5,       /* sgpp::base::generation_exception */
  SGPP_HIP_ERR_DEVICE = 6,           /* HIP runtime failure */
  SGPP_HIP_ERR_BUDGET = 7,           /* served loop: request exceeds the memory budget */
  SGPP_HIP_ERR_NOT_FOUND = 8,        /* served loop: unknown model or request id */
  SGPP_HIP_ERR_CANCELLED = 9,        /* served loop: request cancelled */
  SGPP_HIP_ERR_TIMEOUT = 10,         /* served loop: request deadline passed */
  SGPP_HIP_ERR_SOLVER = 11,          /* solver did not reach its tolerance */
  SGPP_HIP_ERR_INTERNAL = 99
} sgpp_hip_status;

enum { SGPP_HIP_BASIS_LINEAR = 0, SGPP_HIP_BASIS_MODLINEAR = 1 };
enum { SGPP_HIP_PREC_FP64 = 0, SGPP_HIP_PREC_FP32_FP64ACC = 1 };

const char* sgpp_hip_last_error(void);
int sgpp_hip_version(void);

/* ---------------------------------------------------------------- grids
 * Input builder with SG++-identical sequence numbers:
 * Grid::createLinearGrid / createModLinearGrid (base/src/sgpp/base/grid/Grid.hpp:244,339),
 * getGenerator().regular(level) (HashGenerator.hpp:282-342),
 * getGenerator().refine(SurplusRefinementFunctor(alpha, n, thr)) (HashRefinement.cpp:124-141),
 * OperationHierarchisation{Linear,ModLinear}::doHierarchisation.
 */
typedef struct sgpp_hip_grid sgpp_hip_grid;

sgpp_hip_status sgpp_hip_grid_create(int basis, size_t dim, sgpp_hip_grid** out);
sgpp_hip_status sgpp_hip_grid_destroy(sgpp_hip_grid* g);
sgpp_hip_status sgpp_hip_grid_regular(sgpp_hip_grid* g, size_t level);
size_t sgpp_hip_grid_size(const sgpp_hip_grid* g);
size_t sgpp_hip_grid_dim(const sgpp_hip_grid* g);
sgpp_hip_status sgpp_hip_grid_get_level_index(const sgpp_hip_grid* g, uint32_t* level,
                                              uint32_t* index);
sgpp_hip_status sgpp_hip_grid_get_coordinates(const sgpp_hip_grid* g, double* coords);
sgpp_hip_status sgpp_hip_grid_get_leaf(const sgpp_hip_grid* g, uint8_t* leaf);
/* surplus refinement of the refinements_num largest |alpha| with |alpha| >= threshold */
sgpp_hip_status sgpp_hip_grid_refine_surplus(sgpp_hip_grid* g, const double* alpha,
                                             size_t n_alpha, size_t refinements_num,
                                             double threshold, size_t* n_added);
/* full hierarchisation of node values (N in, N surpluses out) */
sgpp_hip_status sgpp_hip_grid_hierarchise(sgpp_hip_grid* g, const double* node_values,
                                          double* alpha);
/* hierarchisation that reuses the per-sweep state of earlier calls on the
 * same grid: only points appended since the last call are processed. The
 * result is identical to sgpp_hip_grid_hierarchise. */
sgpp_hip_status sgpp_hip_grid_hierarchise_incremental(sgpp_hip_grid* g,
                                                      const double* node_values, double* alpha);
/* Grid::serialize text ("linear"/"modlinear" + storage, version 9);
 * *len receives the length incl. the terminating zero; buf may be NULL */
sgpp_hip_status sgpp_hip_grid_serialize(const sgpp_hip_grid* g, char* buf, size_t cap,
                                        size_t* len);
sgpp_hip_status sgpp_hip_grid_unserialize(const char* text, sgpp_hip_grid** out);

/* ------------------------------------------------------------- operator
 * Replaces the operator a reference caller obtains from
 * sgpp::op_factory::createOperationMultipleEval(grid, dataset, cfg)
 * (datadriven/src/sgpp/datadriven/DatadrivenOpFactory.hpp:185-187) and its
 * OperationMultipleEval interface (base/src/sgpp/base/operation/hash/
 * OperationMultipleEval.hpp:27-130).
 */
typedef struct sgpp_hip_meval sgpp_hip_meval;

typedef struct sgpp_hip_grid_desc {
  int basis;  /* SGPP_HIP_BASIS_* */
  size_t dim;
} sgpp_hip_grid_desc;

/* Callers compiled against a header without the `algorithm` member must be
 * rebuilt: the library reads every member below. */
typedef struct sgpp_hip_params {
  int precision;        /* SGPP_HIP_PREC_* used by mult / mult_transpose */
  int keep_fp64;        /* also hold FP64 data when precision is FP32 */
  int device;           /* HIP device ordinal */
  int mult_grid_chunks; /* 0 = automatic */
  int multT_data_segs;  /* 0 = automatic */
  int target_blocks;    /* 0 = automatic */
  int algorithm;        /* SGPP_HIP_ALGO_* (0 = streaming) */
} sgpp_hip_params;

/* STREAMING: every (data point, grid point) pair, culled by supports
 *   (OperationMultiEvalStreaming / OperationMultiEvalModMaskStreaming);
 * SUBSPACE: per data point and subspace the one basis function whose support
 *   holds the point, looked up in a dense per-subspace index table
 *   (OperationMultipleEvalSubspaceCombined, datadriven/src/sgpp/datadriven/
 *   operation/hash/OperationMultipleEvalSubspace/combined/
 *   OperationMultipleEvalSubspaceCombined.hpp:92-204). FP64 only
 *   (precision SGPP_HIP_PREC_FP64, keep_fp64 0), dimension <= 20, Linear and
 *   ModLinear; multiTranspose sums with FP64 atomics, so its last bits vary
 *   between runs. Grids whose index tables would exceed
 *   max(2^26, 64 N) entries are refused. */
enum { SGPP_HIP_ALGO_STREAMING = 0, SGPP_HIP_ALGO_SUBSPACE = 1 };

/* constructor: copies, pads and transposes the dataset into HBM
 * (OperationMultiEvalStreaming.cpp:13-25) */
sgpp_hip_status sgpp_hip_meval_create(const sgpp_hip_grid_desc* grid, const double* data_rowmajor,
                                      size_t M, size_t d, const sgpp_hip_params* params,
                                      sgpp_hip_meval** out);
/* operator bound to a grid handle, created through the C++ factory
 * (createOperationMultipleEval with (STREAMING, HIP)); the grid must
 * outlive the operator and prepare_grid() re-reads it after refinement */
sgpp_hip_status sgpp_hip_meval_create_for_grid(sgpp_hip_grid* grid, const double* data_rowmajor,
                                               size_t M, const sgpp_hip_params* params,
                                               sgpp_hip_meval** out);
/* prepare() (OperationMultiEvalStreaming.cpp:181): new grid arrays */
sgpp_hip_status sgpp_hip_meval_prepare(sgpp_hip_meval* h, const uint32_t* level,
                                       const uint32_t* index, size_t N);
sgpp_hip_status sgpp_hip_meval_prepare_grid(sgpp_hip_meval* h);
/* mult (OperationMultipleEval.hpp:55): alpha N -> result M (host pointers) */
sgpp_hip_status sgpp_hip_meval_mult(sgpp_hip_meval* h, const double* alpha, double* result);
/* multTranspose (OperationMultipleEval.hpp:81): source M -> result N */
sgpp_hip_status sgpp_hip_meval_mult_transpose(sgpp_hip_meval* h, const double* source,
                                              double* result);
/* range forms (OperationMultipleEval.hpp:70-73, 97-100): rows [start, end) of
 * B alpha into result (end - start values), entries [start, end) of B^T source
 * into result (end - start values); start > end or end past the data / grid
 * size is SGPP_HIP_ERR_INVALID_ARGUMENT */
sgpp_hip_status sgpp_hip_meval_mult_range(sgpp_hip_meval* h, const double* alpha, size_t start,
                                          size_t end, double* result);
sgpp_hip_status sgpp_hip_meval_mult_transpose_range(sgpp_hip_meval* h, const double* source,
                                                    size_t start, size_t end, double* result);
/* same on device pointers, asynchronous on the operator's stream */
sgpp_hip_status sgpp_hip_meval_mult_device(sgpp_hip_meval* h, const double* d_alpha,
                                           double* d_result, int precision);
sgpp_hip_status sgpp_hip_meval_mult_transpose_device(sgpp_hip_meval* h, const double* d_source,
                                                     double* d_result, int precision);
/* sgpp_hip_meval_mult_device that allocates nothing (no hipMalloc, no
 * hipFree, no allocating library call). No counterpart in the reference: it
 * is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation). The alpha weights and the
 * FP64 partials live in a fixed 64 MiB device array per GPU, held until the
 * process ends and shared by every call on that GPU: calls that overlap on
 * one GPU race. The grid is cut into fewer chunks so that the partials fit;
 * inputs needing more than 64 MiB at one chunk (N x 4 or 8 bytes plus about
 * M x 8 bytes) are refused with SGPP_HIP_ERR_INVALID_ARGUMENT. Asynchronous
 * on the operator's stream, like sgpp_hip_meval_mult_device. */
sgpp_hip_status sgpp_hip_meval_mult_device_noalloc(sgpp_hip_meval* h, const double* d_alpha,
                                                   double* d_result, int precision);
sgpp_hip_status sgpp_hip_meval_synchronize(sgpp_hip_meval* h);
/* getDuration() (OperationMultipleEval.hpp:120): seconds of the last host call */
double sgpp_hip_meval_duration(const sgpp_hip_meval* h);
/* kernel-only milliseconds of the last launch sequence (HIP events) */
double sgpp_hip_meval_kernel_ms(sgpp_hip_meval* h);
/* the operator's hipStream_t, as an opaque pointer */
void* sgpp_hip_meval_stream(sgpp_hip_meval* h);
size_t sgpp_hip_meval_num_data(const sgpp_hip_meval* h);
size_t sgpp_hip_meval_num_grid(const sgpp_hip_meval* h);
/* G and S chosen by the last mult / mult_transpose launch */
sgpp_hip_status sgpp_hip_meval_last_shape(const sgpp_hip_meval* h, int* mult_chunks,
                                          int* multT_segs);
/* HIP-event timing of the main mult / multTranspose kernel of every launch
 * on the operator's stream (prep and reduction kernels excluded):
 * enable (resets the totals), then read count and total milliseconds. */
sgpp_hip_status sgpp_hip_meval_kernel_timing_enable(sgpp_hip_meval* h, int on);
sgpp_hip_status sgpp_hip_meval_kernel_timing(sgpp_hip_meval* h, int* n_mult, double* ms_mult,
                                             int* n_multT, double* ms_multT);
/* Executed work of one mult and one multTranspose in the given precision,
 * counted by counting instantiations of the same kernels (launched untimed,
 * on zero inputs; culling does not depend on the values). out[0], out[1]:
 * mult hat steps (data point x grid point x evaluated dimension) and pairs
 * (every lane of every evaluated group of grid points); out[2], out[3]: the
 * same for multTranspose. No reference counterpart: the reference's operators
 * evaluate every pair, so LearnerVectorizedPerformanceCalculator
 * (datadriven/src/sgpp/datadriven/tools/LearnerVectorizedPerformanceCalculator.cpp:89-94)
 * counts all M x N pairs. */
sgpp_hip_status sgpp_hip_meval_count_work(sgpp_hip_meval* h, int precision, uint64_t* out4);
sgpp_hip_status sgpp_hip_meval_destroy(sgpp_hip_meval* h);

/* --------------------------------------------------------------- solver
 * (B^T B + lambda * M * I) alpha = B^T y, the system of
 * SystemMatrixLeastSquaresIdentity (datadriven/src/sgpp/datadriven/algorithm/
 * SystemMatrixLeastSquaresIdentity.cpp:32-56) solved by
 * ConjugateGradients::solve (solver/src/sgpp/solver/sle/ConjugateGradients.cpp:22-151).
 */
/* FP64:  every product in FP64 (the reference's arithmetic);
 * MIXED: FP64 iterative refinement around CG on FP32 products, stopping on
 *        the FP64 residual ||b - A alpha|| <= epsilon ||b||;
 * FP32:  the reference's CG with FP32 products (b included) and FP64
 *        vectors, no refinement. */
enum { SGPP_HIP_SOLVE_FP64 = 0, SGPP_HIP_SOLVE_MIXED = 1, SGPP_HIP_SOLVE_FP32 = 2 };

typedef struct sgpp_hip_solve_params {
  int mode;                 /* SGPP_HIP_SOLVE_* */
  double lambda;            /* regularisation lambda (system uses M * lambda) */
  size_t max_iterations;    /* CG iterations (all inner solves together in mixed mode) */
  double epsilon;           /* stop when ||r|| <= epsilon * ||r0|| (FP64 residual) */
  double inner_epsilon;     /* mixed mode: inner CG reduction per refinement step */
  size_t max_refinements;   /* mixed mode: outer iterative-refinement steps */
  int reuse;                /* start from the alpha passed in (FP64 mode) */
} sgpp_hip_solve_params;

typedef struct sgpp_hip_solve_info {
  size_t iterations;        /* CG iterations (all inner iterations in mixed mode) */
  size_t refinements;       /* outer refinement steps (mixed mode) */
  size_t fp64_matvecs;      /* applications of A in FP64 */
  size_t fp32_matvecs;      /* applications of A in FP32 */
  double residual_rel;      /* final ||b - A alpha|| / ||b|| in FP64 */
  double seconds;           /* wall-clock of the solve */
  int converged;            /* the stopping test was met (0: iteration or refinement budget spent) */
} sgpp_hip_solve_info;

/* y: M values, alpha: N values (in: start value when reuse, out: solution).
 * The operator must serve FP64 (FP64 mode), FP32 and FP64 (mixed mode) or
 * FP32 (FP32 mode; residual_rel is then FP64 only if FP64 data is held). */
sgpp_hip_status sgpp_hip_regression_solve(sgpp_hip_meval* h, const double* y,
                                          const sgpp_hip_solve_params* p, double* alpha,
                                          sgpp_hip_solve_info* info);
/* b = B^T y (SystemMatrixLeastSquaresIdentity::generateb) and q = A v, FP64 */
sgpp_hip_status sgpp_hip_system_matrix_mult(sgpp_hip_meval* h, double lambda, const double* v,
                                            double* q, int precision);

/* --------------------------------------------------------- served loop
 * Resident models + continuous batching. The reference has no serving loop:
 * its prediction path builds one operator per request
 * (LearnerLeastSquaresIdentity::predict, datadriven/src/sgpp/datadriven/
 * application/LearnerLeastSquaresIdentity.cpp:61-68); these calls replace that
 * pattern for callers that evaluate many models.
 */
typedef struct sgpp_hip_server sgpp_hip_server;

typedef struct sgpp_hip_server_config {
  int device;
  int precision;                /* SGPP_HIP_PREC_* of the served products */
  size_t memory_budget;         /* bytes for models + batch buffers; 0 = 64 GiB */
  size_t max_batch_queries;     /* queries per batch; 0 = 4 Mi, at most 2^28 */
  size_t max_items;             /* work items per batch; 0 = 256 Ki, at most 2^30 */
  size_t max_partial_doubles;   /* partial-sum capacity per batch; 0 = 128 Mi */
  int max_inflight;             /* batches on the GPU at once; 0 = 2 */
  int max_requests_per_batch;   /* 1 = one launch per request; 0 = 65536 */
  int blocks;                   /* persistent grid size; 0 = 8 per CU */
  int use_graph;                /* replay the batch kernels from a hipGraph */
  double item_work;             /* untuned: lane-operations per work item; 0 = 2^25 */
  int autotune;                 /* time tile shapes per (basis, dimension) at admission */
  const char* autotune_cache;   /* file of tuned shapes (read, then updated); NULL = none */
  int algorithm;                /* SGPP_HIP_SERVE_*; 0 = per model the faster */
} sgpp_hip_server_config;

/* How a served model is evaluated. STREAMING: every (query, grid point)
 * pair, in the configured precision. SUBSPACE: per query, each subspace of
 * the model once, the one basis function whose support holds the query taken
 * from a dense index table, in FP64 (the SubspaceLinear algorithm,
 * OperationMultipleEvalSubspaceCombined.hpp:92-204); models whose tables
 * would be too large stay on STREAMING. AUTO: with autotune, the faster of
 * the two, timed per model at admission; without it, STREAMING. Callers
 * compiled against a header without the `algorithm` member must be rebuilt. */
enum { SGPP_HIP_SERVE_AUTO = 0, SGPP_HIP_SERVE_STREAMING = 1, SGPP_HIP_SERVE_SUBSPACE = 2 };

/* A tuned shape: grid points per work item and the measured seconds per
 * (query x grid point). A model with more than 256 x chunk grid points uses
 * larger items, so that it has at most 256 grid chunks per query. */
typedef struct sgpp_hip_server_shape {
  int chunk;
  double cost;
} sgpp_hip_server_shape;

enum {
  SGPP_HIP_REQ_QUEUED = 0,
  SGPP_HIP_REQ_RUNNING = 1,
  SGPP_HIP_REQ_DONE = 2,
  SGPP_HIP_REQ_FAILED_BUDGET = 3,  /* one query tile of the model exceeds a batch's buffers */
  SGPP_HIP_REQ_FAILED_TIMEOUT = 4,
  SGPP_HIP_REQ_CANCELLED = 5,
  SGPP_HIP_REQ_FAILED_DEVICE = 6
};

typedef struct sgpp_hip_server_statistics {
  uint64_t requests_done;
  uint64_t requests_failed;
  uint64_t queries_done;
  uint64_t batches;
  uint64_t kernel_launches;
  double busy_seconds;          /* summed device time of the batches */
} sgpp_hip_server_statistics;

sgpp_hip_status sgpp_hip_server_create(const sgpp_hip_server_config* cfg, sgpp_hip_server** out);
sgpp_hip_status sgpp_hip_server_destroy(sgpp_hip_server* s);
/* level/index N x d, alpha N; SGPP_HIP_ERR_BUDGET when it does not fit */
sgpp_hip_status sgpp_hip_server_admit_model(sgpp_hip_server* s, int basis, size_t dim, size_t N,
                                            const uint32_t* level, const uint32_t* index,
                                            const double* alpha, int* model_id);
sgpp_hip_status sgpp_hip_server_evict_model(sgpp_hip_server* s, int model_id);
sgpp_hip_status sgpp_hip_server_memory_used(sgpp_hip_server* s, size_t* bytes);
/* queries: q x d row-major in [0,1]^d (d = the model's dimension); deadline_s
 * <= 0 means none. An FP32 server carries each coordinate as fl32(x) plus,
 * for models with levels where the rounding matters (L = 2^l >= 64),
 * fl32(x - fl32(x)). */
sgpp_hip_status sgpp_hip_server_submit(sgpp_hip_server* s, int model_id, size_t q,
                                       const double* queries, double deadline_s,
                                       uint64_t* request_id);
/* *status receives SGPP_HIP_REQ_*; a finished request is released (its q
 * results copied to out when DONE). out holds out_capacity values; a DONE
 * request with more than out_capacity results is refused with
 * SGPP_HIP_ERR_INVALID_ARGUMENT and kept. wait_s < 0 blocks, 0 polls. */
sgpp_hip_status sgpp_hip_server_poll(sgpp_hip_server* s, uint64_t request_id, double* out,
                                     size_t out_capacity, double wait_s, int* status,
                                     double* latency_s);
sgpp_hip_status sgpp_hip_server_cancel(sgpp_hip_server* s, uint64_t request_id);
/* host-forced fault: the next batch reports a device error on its requests */
sgpp_hip_status sgpp_hip_server_inject_error(sgpp_hip_server* s);
sgpp_hip_status sgpp_hip_server_drain(sgpp_hip_server* s);
sgpp_hip_status sgpp_hip_server_stats(sgpp_hip_server* s, sgpp_hip_server_statistics* out);
/* the tuned shape of (basis, dim); *found = 0 when that pair is not tuned.
 * No reference counterpart (the reference has no serving loop); its OpenCL
 * operator tuner is StaticParameterTuner (datadriven/src/sgpp/datadriven/
 * application/StaticParameterTuner.cpp:234-367). */
sgpp_hip_status sgpp_hip_server_tuned_shape(sgpp_hip_server* s, int basis, int dim, int* found,
                                            sgpp_hip_server_shape* out);
/* *algorithm = SGPP_HIP_SERVE_STREAMING or SGPP_HIP_SERVE_SUBSPACE: the kernel
 * that serves the model (no reference counterpart) */
sgpp_hip_status sgpp_hip_server_model_algorithm(sgpp_hip_server* s, int model_id,
                                                int* algorithm);

/* --------------------------------------------------------- learner
 * The regression learner that drives the path: regular grid -> (refine ->
 * prepare -> b = B^T y -> CG)* (LearnerBase::train, datadriven/src/sgpp/
 * datadriven/application/LearnerBase.cpp:113-327, with
 * LearnerLeastSquaresIdentity.cpp:32-39 for the system), predict
 * (LearnerLeastSquaresIdentity.cpp:61-68) and getAccuracy
 * (LearnerBase.cpp:369-411). Only CG is provided; error-based refinement
 * is refused (SGPP_HIP_ERR_NOT_IMPLEMENTED).
 */
typedef struct sgpp_hip_learner sgpp_hip_learner;

/* RegularGridConfiguration (base/src/sgpp/base/grid/Grid.hpp) */
typedef struct sgpp_hip_grid_config {
  int basis;  /* SGPP_HIP_BASIS_* */
  size_t dim;
  int level;
} sgpp_hip_grid_config;

/* SLESolverConfiguration (solver/src/sgpp/solver/TypesSolver.hpp:19-25) */
enum { SGPP_HIP_SLE_CG = 0, SGPP_HIP_SLE_BICGSTAB = 1 };
typedef struct sgpp_hip_solver_config {
  int type;               /* SGPP_HIP_SLE_*; only CG is provided */
  double eps;
  size_t max_iterations;
  double threshold;
} sgpp_hip_solver_config;

/* AdaptivityConfiguration (base/src/sgpp/base/grid/Grid.hpp:140-165) */
typedef struct sgpp_hip_adaptivity_config {
  size_t num_refinements;
  double refinement_threshold;
  size_t num_refinement_points;
  int error_based_refinement;  /* must be 0 */
} sgpp_hip_adaptivity_config;

/* LearnerTiming (datadriven/src/sgpp/datadriven/tools/TypesDatadriven.hpp:37-54) */
typedef struct sgpp_hip_learner_timing {
  double time_complete;
  double time_mult_complete;
  double time_mult_compute;
  double time_mult_trans_complete;
  double time_mult_trans_compute;
  double time_regularization;
  double gflop;
  double gbyte;
} sgpp_hip_learner_timing;

/* one pass of the refinement loop */
typedef struct sgpp_hip_learner_step {
  size_t grid_size;
  size_t iterations;
  size_t refinements;
  double residual;        /* ||b - A alpha|| / ||b|| after the solve */
  double train_accuracy;  /* MSE or accuracy when tested, else -1 */
  double seconds;
} sgpp_hip_learner_step;

/* params selects the operator's precision (FP32_FP64ACC => mixed solve) and device */
sgpp_hip_status sgpp_hip_learner_create(int is_regression, const sgpp_hip_params* params,
                                        sgpp_hip_learner** out);
sgpp_hip_status sgpp_hip_learner_destroy(sgpp_hip_learner* l);
/* data: M x d row-major, classes: M values */
sgpp_hip_status sgpp_hip_learner_train(sgpp_hip_learner* l, const double* data, size_t M,
                                       size_t d, const double* classes,
                                       const sgpp_hip_grid_config* grid,
                                       const sgpp_hip_solver_config* solver_refine,
                                       const sgpp_hip_solver_config* solver_final,
                                       const sgpp_hip_adaptivity_config* adaptivity,
                                       int test_acc_during_adapt, double lambda,
                                       sgpp_hip_learner_timing* timing);
sgpp_hip_status sgpp_hip_learner_predict(sgpp_hip_learner* l, const double* data, size_t M,
                                         size_t d, double* out);
sgpp_hip_status sgpp_hip_learner_accuracy(sgpp_hip_learner* l, const double* data, size_t M,
                                          size_t d, const double* classes, double threshold,
                                          double* accuracy);
sgpp_hip_status sgpp_hip_learner_grid_size(sgpp_hip_learner* l, size_t* n);
/* level/index: N x d, alpha: N */
sgpp_hip_status sgpp_hip_learner_level_index(sgpp_hip_learner* l, uint32_t* level,
                                             uint32_t* index);
sgpp_hip_status sgpp_hip_learner_alpha(sgpp_hip_learner* l, double* alpha);
sgpp_hip_status sgpp_hip_learner_num_steps(sgpp_hip_learner* l, size_t* n);
sgpp_hip_status sgpp_hip_learner_step_info(sgpp_hip_learner* l, size_t k,
                                           sgpp_hip_learner_step* out);

/* GFlop / GByte of one CG run on a grid under the reference's convention
 * (LearnerVectorizedPerformanceCalculator::getGFlopAndGByte,
 * datadriven/src/sgpp/datadriven/tools/LearnerVectorizedPerformanceCalculator.cpp:17-113) */
sgpp_hip_status sgpp_hip_learner_performance(sgpp_hip_grid* grid, size_t num_instances,
                                             int solver_type, size_t num_iterations,
                                             size_t size_datatype, int reuse_alpha,
                                             double* gflop, double* gbyte);

/* --------------------------------------------------------- fitter
 * The least-squares model fitter of the datamining pipeline
 * (ModelFittingLeastSquares, datadriven/src/sgpp/datadriven/datamining/
 * modules/fitting/ModelFittingLeastSquares.cpp:30-142): fit (regular grid,
 * solve), adapt (one surplus refinement, warm-started solve), update (new
 * data on the current grid, warm-started solve; the reference's update frees
 * its grid before using it, this one keeps it) and evaluate. Only CG is
 * provided. The fitter owns copies of the data it was given. */
typedef struct sgpp_hip_fitter sgpp_hip_fitter;

/* FitterConfigurationLeastSquares (FitterConfiguration.cpp:150-220 defaults:
 * 1 refinement of 1 point, CG eps 1e-12 / 100 iterations, lambda 0.01);
 * grid.dim is ignored (taken from the data) */
typedef struct sgpp_hip_fitter_config {
  sgpp_hip_grid_config grid;
  sgpp_hip_adaptivity_config refinement;
  sgpp_hip_solver_config solver_refine;
  sgpp_hip_solver_config solver_final;
  double lambda;
} sgpp_hip_fitter_config;

sgpp_hip_status sgpp_hip_fitter_create(const sgpp_hip_fitter_config* cfg,
                                       const sgpp_hip_params* params, sgpp_hip_fitter** out);
sgpp_hip_status sgpp_hip_fitter_destroy(sgpp_hip_fitter* f);
/* data: M x d row-major; targets: M */
sgpp_hip_status sgpp_hip_fitter_fit(sgpp_hip_fitter* f, const double* data, size_t M, size_t d,
                                    const double* targets);
/* *refined = 1 when the grid grew and the model was solved again */
sgpp_hip_status sgpp_hip_fitter_adapt(sgpp_hip_fitter* f, int* refined);
sgpp_hip_status sgpp_hip_fitter_update(sgpp_hip_fitter* f, const double* data, size_t M,
                                       size_t d, const double* targets);
sgpp_hip_status sgpp_hip_fitter_evaluate(sgpp_hip_fitter* f, const double* data, size_t M,
                                         size_t d, double* out);
sgpp_hip_status sgpp_hip_fitter_grid_size(sgpp_hip_fitter* f, size_t* n);
sgpp_hip_status sgpp_hip_fitter_level_index(sgpp_hip_fitter* f, uint32_t* level,
                                            uint32_t* index);
sgpp_hip_status sgpp_hip_fitter_alpha(sgpp_hip_fitter* f, double* alpha);
/* CG iterations of the last solve and refinements since fit / update */
sgpp_hip_status sgpp_hip_fitter_info(sgpp_hip_fitter* f, size_t* last_iterations,
                                     size_t* refinements);

#ifdef __cplusplus
}
#endif

#endif /* SGPP_HIP_H_ */
