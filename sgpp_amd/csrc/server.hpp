// Resident evaluation server: many trained sparse-grid models kept in HBM,
// queries of many requests packed into one persistent launch.
//
// The reference has no serving loop. Its prediction path constructs one
// operator per call (LearnerLeastSquaresIdentity::predict,
// datadriven/src/sgpp/datadriven/application/LearnerLeastSquaresIdentity.cpp:61-68),
// i.e. one copy + one mult per request. Here:
//   * admit_model uploads the encoded grid and alpha once (memory budget);
//   * submit enqueues a request (model id, q query points);
//   * a scheduler thread forms batches from the queue (FIFO) whenever one of
//     `max_inflight` batch slots is free, so requests that arrive while others
//     run are admitted and overlap on the device (continuous batching);
//   * a batch is a work list of items (model, query tile, grid chunk) run by
//     one persistent kernel whose blocks pull items from a counter, followed
//     by a reduction kernel; both are replayed from a captured hipGraph;
//   * the kernel stages 64-point grid tiles in LDS and evaluates them in
//     sub-tiles, skipping per sub-tile the dimensions in which every point is
//     constant (ModLinear level 1) and evaluating once per query those in
//     which every point has the same (L, I); models are stored sorted by
//     their refined-dimension sets and (L, I) so that sub-tiles share both;
//   * the queries of all requests of a batch for one model form one block,
//     tiled together (a request pays for at most one partial query tile per
//     model and batch); the grid chunk per work item comes from an autotuner
//     that times the candidates per (basis, dimension) on the first admitted
//     model and can persist them in a cache file;
//   * partials of a query are summed over its model's grid chunks in chunk
//     order, and a query's result depends only on its model and that
//     chunking, so a result is bit-identical whether the request runs alone
//     or in any mix;
//   * a model whose grid has few points per subspace evaluation can instead be
//     served by a second work-list kernel in the same launch that visits,
//     per query, each subspace (level vector) once and takes the one basis
//     function whose support holds the query from a dense index table, in
//     FP64 (the SubspaceLinear algorithm); with the autotuner the faster of
//     the two is chosen per model at admission;
//   * a request larger than a batch's buffers is split by queries into
//     pieces over successive batches (results unchanged, by the above); only
//     a model whose single query tile exceeds them fails "over budget".
// Per-request failures (over budget, deadline passed, cancelled, an injected
// kernel-error flag) are reported on the request and never stop the loop.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "meval_device.hpp"

namespace sgpp_hip {

// Evaluation algorithm of the served models (ServerConfig::algorithm).
constexpr int kServeAuto = 0;       // per model, the faster (timed; streaming untuned)
constexpr int kServeStreaming = 1;  // every (query, grid point) pair
constexpr int kServeSubspace = 2;   // subspace lookups wherever the model has a layout

struct ServerConfig {
  int device = 0;
  Precision precision = Precision::FP32;  // FP32 products + FP64 accumulation, or FP64
  size_t memory_budget = size_t(64) << 30;  // models + batch buffers, bytes
  size_t max_batch_queries = size_t(1) << 22;
  size_t max_items = size_t(1) << 18;       // work items per batch
  size_t max_partial_doubles = size_t(1) << 27;
  int max_inflight = 2;       // batch slots on the GPU at once
  int max_requests_per_batch = 1 << 16;
  int blocks = 0;             // persistent grid size; 0 = 8 per CU
  bool use_graph = true;
  double item_work = 33554432.0;  // untuned: lane-operations per work item (tile x chunk x (2d+1))
  bool autotune = false;          // time tile shapes per (basis, dimension) at admission
  std::string autotune_cache;     // file of tuned shapes, read at start and updated; "" = none
  int algorithm = kServeAuto;
};

// Tuned shape of one (basis, dimension): grid points per work item and the
// measured cost in seconds per (query x grid point).
struct TunedShape {
  int chunk = 0;
  double cost = 0.0;
};

enum RequestStatus : int {
  REQ_QUEUED = 0,
  REQ_RUNNING = 1,
  REQ_DONE = 2,
  REQ_FAILED_BUDGET = 3,
  REQ_FAILED_TIMEOUT = 4,
  REQ_CANCELLED = 5,
  REQ_FAILED_DEVICE = 6,
  REQ_FAILED_INVALID = 7,
};

struct ServerStats {
  uint64_t requests_done = 0;
  uint64_t requests_failed = 0;
  uint64_t queries_done = 0;
  uint64_t batches = 0;
  uint64_t kernel_launches = 0;
  double busy_seconds = 0.0;  // sum of batch device times
};

class Server {
 public:
  explicit Server(const ServerConfig& cfg);
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  // level/index: N x d row-major; alpha: N surpluses. Returns the model id.
  int admit_model(Basis basis, size_t dim, size_t N, const uint32_t* level, const uint32_t* index,
                  const double* alpha);
  void evict_model(int id);
  size_t model_bytes(int id);
  size_t memory_used();

  // queries: q x d row-major FP64 in [0,1]^d. deadline_s <= 0: none
  // (seconds after submission). Returns a request id (never blocks).
  uint64_t submit(int model, size_t q, const double* queries, double deadline_s);
  // status of a request; when done, copies the q results into out (if not
  // null; out_capacity must be at least q, else input_error and the request
  // stays) and releases the request. wait_s < 0 waits forever, 0 polls.
  int poll(uint64_t id, double* out, size_t out_capacity, double wait_s, double* latency_s);
  bool cancel(uint64_t id);
  // makes the next batch report a device error on all its requests
  void inject_kernel_error();
  // the tuned shape of (basis, dim), if tuned or loaded from the cache
  bool tuned_shape(Basis basis, int dim, TunedShape* out);
  // kServeStreaming or kServeSubspace: the kernel that serves model id
  int model_algorithm(int id);
  ServerStats stats();
  void drain();  // blocks until no request is queued or running

 private:
  struct Model;
  struct Request;
  struct Piece;
  struct Slot;
  size_t chunk_for(const Model& m) const;
  void tune(Model& m);
  void choose_algorithm(Model& m);
  void enqueue_batch_kernels(Slot& s, hipStream_t st);
  void enqueue_kernels(hipStream_t st, const void* spans, const int* counts, unsigned* counter,
                       const void* X, double* part, double* out);
  void load_cache();
  void save_cache();
  static int tune_key(Basis basis, int D);
  void loop();
  bool form_batch(Slot& s, std::vector<Piece>& taken);
  void launch_batch(Slot& s);
  void finish_batch(Slot& s);
  void fail_request(const std::shared_ptr<Request>& r, int status, const char* why);
  void fail_pieces(Slot& s, const char* why);

  ServerConfig cfg_;
  int num_cus_ = 256;
  size_t used_bytes_ = 0;
  std::mutex mu_;
  std::condition_variable cv_;       // scheduler wake-up
  std::condition_variable done_cv_;  // request completion
  std::deque<std::shared_ptr<Request>> queue_;
  std::map<uint64_t, std::shared_ptr<Request>> requests_;
  std::map<int, std::shared_ptr<Model>> models_;
  std::map<int, TunedShape> tuned_;  // by tune_key
  std::vector<std::unique_ptr<Slot>> slots_;
  void* model_table_ = nullptr;  // device array of model descriptors
  int next_model_ = 0;
  uint64_t next_request_ = 1;
  int running_ = 0;
  bool stop_ = false;
  std::atomic<bool> inject_error_{false};
  ServerStats stats_;
  std::thread thread_;
};

}  // namespace sgpp_hip
