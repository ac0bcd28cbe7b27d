// Resident evaluation server (continuous batching across models).
// See server.hpp for the design; kernels and the scheduler live here.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "server.hpp"

namespace sgpp_hip {

namespace {

constexpr int kBlock = 256;
constexpr int kTG = 64;
constexpr int kMaxD = 20;
constexpr int kMaxModels = 4096;

struct ModelDesc {
  const void* LI;         // tile-major (L, I): [Npad/64][D][64][2]
  const void* A;          // [Npad] alpha * scale, zero padded
  const uint32_t* qmask;  // [Npad/4] dimensions evaluated per 4 consecutive points
  int D;
  int Npad;               // multiple of kTG
};

struct Item {
  int model;
  int D;
  int q0;          // first query of the tile (request-relative)
  int nq;          // valid queries in the tile
  int j0, j1;      // grid chunk, multiples of kTG
  long long xoff;  // element offset of the request's query block [D][ldq]
  int ldq;         // padded query count of the request (multiple of the tile)
  int pad;
  long long poff;  // double offset of this item's partial row (tile doubles)
};

struct RedItem {
  long long pbase;  // partial offset of chunk 0 of this tile
  int nchunks;
  int nq;
  int tile;
  int pad;
  long long out;    // output offset of the tile's first query
};

template <typename T>
constexpr int serve_k(int D) {
  return sizeof(T) == 8 ? (D <= 5 ? 4 : (D <= 10 ? 2 : 1)) : (D <= 4 ? 8 : (D <= 10 ? 4 : 2));
}
// grid points per sub-tile (one wave-uniform branch per sub-tile and
// dimension); J*K running products per lane, capped so that every dimension
// case of the persistent kernel stays within the same register budget
template <typename T>
constexpr int serve_j(int D) {
  return (sizeof(T) == 8 ? 16 : 32) / serve_k<T>(D) < 8 ? (sizeof(T) == 8 ? 16 : 32) / serve_k<T>(D)
                                                          : 8;
}
template <typename T>
int serve_k_rt(int D) {
  return sizeof(T) == 8 ? (D <= 5 ? 4 : (D <= 10 ? 2 : 1)) : (D <= 4 ? 8 : (D <= 10 ? 4 : 2));
}

__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float absT(float a) { return __builtin_fabsf(a); }
__device__ __forceinline__ double absT(double a) { return __builtin_fabs(a); }
__device__ __forceinline__ float clamp01(float v) { return __builtin_amdgcn_fmed3f(v, 0.0f, 1.0f); }
__device__ __forceinline__ double clamp01(double v) { return __builtin_fmin(__builtin_fmax(v, 0.0), 1.0); }

// one work item: a tile of kBlock*K queries of one request against one grid
// chunk. The grid tile is walked in sub-tiles of J points; per (sub-tile,
// dimension) one wave-uniform test of the sub-tile's mask decides whether the
// dimension is evaluated (ModLinear level-1 factors are exactly 1; Linear
// models carry full masks).
template <int D, int K, int J, typename T>
__device__ __forceinline__ void serve_item(const Item& item, const ModelDesc& md, const T* X,
                                           double* part, T* sLI, T* sA, uint32_t* sM) {
  static_assert(kTG % J == 0 && J % 4 == 0, "sub-tile must divide the tile, 4 must divide it");
  const T* LI = static_cast<const T*>(md.LI);
  const T* A = static_cast<const T*>(md.A);
  const T* Xr = X + item.xoff;
  T x[K][D];
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int d = 0; d < D; ++d)
      x[k][d] = Xr[static_cast<long long>(d) * item.ldq + item.q0 + k * kBlock + threadIdx.x];
  double accd[K];
#pragma unroll
  for (int k = 0; k < K; ++k) accd[k] = 0.0;
  for (int j0 = item.j0; j0 < item.j1; j0 += kTG) {
    __syncthreads();
    for (int e = threadIdx.x; e < kTG * D * 2; e += kBlock)
      sLI[e] = LI[static_cast<long long>(j0) * D * 2 + e];
    if (threadIdx.x < kTG) sA[threadIdx.x] = A[j0 + threadIdx.x];
    if (threadIdx.x < kTG / 4) sM[threadIdx.x] = md.qmask[j0 / 4 + threadIdx.x];
    __syncthreads();
    T acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = T(0);
    for (int j = 0; j < kTG; j += J) {
      uint32_t mask = 0u;
#pragma unroll
      for (int c = 0; c < J / 4; ++c) mask |= sM[j / 4 + c];
      mask = __builtin_amdgcn_readfirstlane(mask);
      T p[J][K];
#pragma unroll
      for (int q = 0; q < J; ++q)
#pragma unroll
        for (int k = 0; k < K; ++k) p[q][k] = T(1);
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (!((mask >> d) & 1u)) continue;
        T li[2 * J];
#pragma unroll
        for (int c = 0; c < 2 * J; ++c) li[c] = sLI[(d * kTG + j) * 2 + c];
#pragma unroll
        for (int q = 0; q < J; ++q)
#pragma unroll
          for (int k = 0; k < K; ++k) {
            const T u = fmaT(li[2 * q], x[k][d], -li[2 * q + 1]);
            p[q][k] = clamp01(fmaT(-p[q][k], absT(u), p[q][k]));
          }
      }
#pragma unroll
      for (int q = 0; q < J; ++q) {
        const T a = sA[j + q];
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = fmaT(a, p[q][k], acc[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) accd[k] += static_cast<double>(acc[k]);
  }
  double* out = part + item.poff;
#pragma unroll
  for (int k = 0; k < K; ++k) out[k * kBlock + threadIdx.x] = accd[k];
}

template <typename T, int D>
__device__ __forceinline__ void dispatch_one(const Item& it, const ModelDesc& md, const T* X,
                                             double* part, T* sLI, T* sA, uint32_t* sM) {
  serve_item<D, serve_k<T>(D), serve_j<T>(D), T>(it, md, X, part, sLI, sA, sM);
}

// Persistent work-list kernel: blocks pull items from a counter until the
// list is exhausted (every block reaches the exit test).
template <typename T>
__global__ __launch_bounds__(kBlock) void serve_kernel(const ModelDesc* __restrict__ models,
                                                       const Item* __restrict__ items,
                                                       const int* __restrict__ counts,
                                                       unsigned* counter, const T* __restrict__ X,
                                                       double* __restrict__ part) {
  __shared__ T sLI[kTG * kMaxD * 2];
  __shared__ T sA[kTG];
  __shared__ uint32_t sM[kTG / 4];
  __shared__ unsigned s_item;
  const unsigned n = static_cast<unsigned>(counts[0]);
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1u);
    __syncthreads();
    const unsigned it = s_item;
    __syncthreads();
    if (it >= n) break;
    const Item item = items[it];
    const ModelDesc md = models[item.model];
    switch (item.D) {
#define SGPP_CASE(DD) \
  case DD:            \
    dispatch_one<T, DD>(item, md, X, part, sLI, sA, sM); \
    break;
      SGPP_CASE(1) SGPP_CASE(2) SGPP_CASE(3) SGPP_CASE(4) SGPP_CASE(5) SGPP_CASE(6) SGPP_CASE(7)
      SGPP_CASE(8) SGPP_CASE(9) SGPP_CASE(10) SGPP_CASE(11) SGPP_CASE(12) SGPP_CASE(13)
      SGPP_CASE(14) SGPP_CASE(15) SGPP_CASE(16) SGPP_CASE(17) SGPP_CASE(18) SGPP_CASE(19)
      SGPP_CASE(20)
#undef SGPP_CASE
      default:
        break;
    }
  }
}

// out[tile query i] = sum over chunks in chunk order
__global__ __launch_bounds__(kBlock) void serve_reduce(const RedItem* __restrict__ red,
                                                       const int* __restrict__ counts,
                                                       const double* __restrict__ part,
                                                       double* __restrict__ out) {
  const int n = counts[1];
  for (int r = blockIdx.x; r < n; r += gridDim.x) {
    const RedItem ri = red[r];
    for (int i = threadIdx.x; i < ri.nq; i += kBlock) {
      double s = 0.0;
      for (int c = 0; c < ri.nchunks; ++c) s += part[ri.pbase + static_cast<long long>(c) * ri.tile + i];
      out[ri.out + i] = s;
    }
  }
}

__global__ void reset_counter(unsigned* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] = 0u;
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

struct Server::Model {
  int id = 0;
  int table_index = 0;
  Basis basis = Basis::Linear;
  int D = 0;
  size_t N = 0, Npad = 0;
  void* LI = nullptr;
  void* A = nullptr;
  void* tmask = nullptr;
  size_t bytes = 0;
  int tile = 0;    // queries per work item
  int chunk = 0;   // grid points per work item
  int active = 0;  // requests queued or running
};

struct Server::Request {
  uint64_t id = 0;
  int model = 0;
  size_t q = 0;
  size_t ldq = 0;
  std::vector<char> X;  // [D][ldq] in the compute type, zero padded
  std::vector<double> out;
  double t_submit = 0.0, t_done = 0.0, deadline = 0.0;
  int status = REQ_QUEUED;
  std::string error;
  // batch placement
  long long xoff = 0, out_off = 0;
};

struct Server::Slot {
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_done = nullptr;
  hipGraphExec_t graph = nullptr;
  void* dX = nullptr;
  Item* dItems = nullptr;
  RedItem* dRed = nullptr;
  int* dCounts = nullptr;
  unsigned* dCounter = nullptr;
  double* dPart = nullptr;
  double* dOut = nullptr;
  char* hX = nullptr;
  Item* hItems = nullptr;
  RedItem* hRed = nullptr;
  int* hCounts = nullptr;
  double* hOut = nullptr;
  size_t capX = 0, capItems = 0, capRed = 0, capPart = 0, capOut = 0;
  // current batch
  bool busy = false;
  bool inject = false;
  std::vector<std::shared_ptr<Request>> reqs;
  size_t usedX = 0, nItems = 0, nRed = 0, usedPart = 0, usedOut = 0;
};

Server::Server(const ServerConfig& cfg) : cfg_(cfg) {
  if (cfg_.max_inflight < 1 || cfg_.max_inflight > 8) throw input_error("max_inflight in [1, 8]");
  check_hip(hipSetDevice(cfg_.device), "hipSetDevice");
  hipDeviceProp_t prop;
  check_hip(hipGetDeviceProperties(&prop, cfg_.device), "hipGetDeviceProperties");
  num_cus_ = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  if (cfg_.blocks <= 0) cfg_.blocks = 8 * num_cus_;
  const size_t esz = cfg_.precision == Precision::FP64 ? 8 : 4;
  check_hip(hipMalloc(&model_table_, kMaxModels * sizeof(ModelDesc)), "hipMalloc(models)");
  used_bytes_ += kMaxModels * sizeof(ModelDesc);
  for (int i = 0; i < cfg_.max_inflight; ++i) {
    std::unique_ptr<Slot> s(new Slot());
    s->capOut = cfg_.max_batch_queries;
    s->capX = cfg_.max_batch_queries * kMaxD + size_t(kMaxD) * 4096 * 64;
    s->capItems = cfg_.max_items;
    s->capRed = cfg_.max_batch_queries / kBlock + size_t(cfg_.max_requests_per_batch) + 1;
    s->capPart = cfg_.max_partial_doubles;
    const size_t dev_bytes = s->capX * esz + s->capItems * sizeof(Item) +
                             s->capRed * sizeof(RedItem) + 4 * sizeof(int) +
                             s->capPart * sizeof(double) + s->capOut * sizeof(double);
    if (used_bytes_ + dev_bytes > cfg_.memory_budget)
      throw input_error("memory budget too small for the batch buffers");
    used_bytes_ += dev_bytes;
    check_hip(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
    check_hip(hipEventCreate(&s->ev_start), "hipEventCreate");
    check_hip(hipEventCreate(&s->ev_done), "hipEventCreate");
    check_hip(hipMalloc(&s->dX, s->capX * esz), "hipMalloc(batch X)");
    check_hip(hipMalloc(reinterpret_cast<void**>(&s->dItems), s->capItems * sizeof(Item)),
              "hipMalloc(items)");
    check_hip(hipMalloc(reinterpret_cast<void**>(&s->dRed), s->capRed * sizeof(RedItem)),
              "hipMalloc(red)");
    check_hip(hipMalloc(reinterpret_cast<void**>(&s->dCounts), 4 * sizeof(int)),
              "hipMalloc(counts)");
    check_hip(hipMalloc(reinterpret_cast<void**>(&s->dCounter), sizeof(unsigned)),
              "hipMalloc(counter)");
    check_hip(hipMalloc(reinterpret_cast<void**>(&s->dPart), s->capPart * sizeof(double)),
              "hipMalloc(partials)");
    check_hip(hipMalloc(reinterpret_cast<void**>(&s->dOut), s->capOut * sizeof(double)),
              "hipMalloc(out)");
    check_hip(hipHostMalloc(reinterpret_cast<void**>(&s->hX), s->capX * esz), "hipHostMalloc");
    check_hip(hipHostMalloc(reinterpret_cast<void**>(&s->hItems), s->capItems * sizeof(Item)),
              "hipHostMalloc");
    check_hip(hipHostMalloc(reinterpret_cast<void**>(&s->hRed), s->capRed * sizeof(RedItem)),
              "hipHostMalloc");
    check_hip(hipHostMalloc(reinterpret_cast<void**>(&s->hCounts), 4 * sizeof(int)),
              "hipHostMalloc");
    check_hip(hipHostMalloc(reinterpret_cast<void**>(&s->hOut), s->capOut * sizeof(double)),
              "hipHostMalloc");
    if (cfg_.use_graph) {
      // reset + persistent work-list kernel + reduction, captured once per slot;
      // the item counts are read from device memory, so one graph serves every batch
      hipGraph_t g;
      check_hip(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal),
                "hipStreamBeginCapture");
      clear_launch_error();
      hipLaunchKernelGGL(reset_counter, dim3(1), dim3(64), 0, s->stream, s->dCounter);
      if (cfg_.precision == Precision::FP64)
        hipLaunchKernelGGL(serve_kernel<double>, dim3(cfg_.blocks), dim3(kBlock), 0, s->stream,
                           static_cast<const ModelDesc*>(model_table_), s->dItems, s->dCounts,
                           s->dCounter, static_cast<const double*>(s->dX), s->dPart);
      else
        hipLaunchKernelGGL(serve_kernel<float>, dim3(cfg_.blocks), dim3(kBlock), 0, s->stream,
                           static_cast<const ModelDesc*>(model_table_), s->dItems, s->dCounts,
                           s->dCounter, static_cast<const float*>(s->dX), s->dPart);
      hipLaunchKernelGGL(serve_reduce, dim3(2 * num_cus_), dim3(kBlock), 0, s->stream, s->dRed,
                         s->dCounts, s->dPart, s->dOut);
      check_hip(hipStreamEndCapture(s->stream, &g), "hipStreamEndCapture");
      check_hip(hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0), "hipGraphInstantiate");
      check_hip(hipGraphDestroy(g), "hipGraphDestroy");
    }
    slots_.push_back(std::move(s));
  }
  thread_ = std::thread(&Server::loop, this);
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
    for (auto& r : queue_)
      if (r->status == REQ_QUEUED) r->status = REQ_CANCELLED;
    queue_.clear();
  }
  cv_.notify_all();
  if (thread_.joinable()) thread_.join();
  for (auto& s : slots_) {
    (void)hipStreamSynchronize(s->stream);
    if (s->graph) (void)hipGraphExecDestroy(s->graph);
    void* dev[] = {s->dX, s->dItems, s->dRed, s->dCounts, s->dCounter, s->dPart, s->dOut};
    for (void* p : dev)
      if (p) (void)hipFree(p);
    void* host[] = {s->hX, s->hItems, s->hRed, s->hCounts, s->hOut};
    for (void* p : host)
      if (p) (void)hipHostFree(p);
    (void)hipEventDestroy(s->ev_start);
    (void)hipEventDestroy(s->ev_done);
    (void)hipStreamDestroy(s->stream);
  }
  for (auto& kv : models_) {
    if (kv.second->LI) (void)hipFree(kv.second->LI);
    if (kv.second->A) (void)hipFree(kv.second->A);
    if (kv.second->tmask) (void)hipFree(kv.second->tmask);
  }
  if (model_table_) (void)hipFree(model_table_);
}

int Server::admit_model(Basis basis, size_t dim, size_t N, const uint32_t* level,
                        const uint32_t* index, const double* alpha) {
  if (dim < 1 || dim > static_cast<size_t>(kMaxD)) throw input_error("model dimension in [1, 20]");
  if (N == 0) throw input_error("empty model");
  // ModLinear: points sorted by their set of non-constant dimensions (level
  // above 1), so that tiles share masks; the model's output is a sum over all
  // points, so the order changes nothing but the summation order
  const bool skip = basis == Basis::ModLinear && dim >= 2;
  std::vector<uint32_t> pmask(N);
  for (size_t j = 0; j < N; ++j) {
    uint32_t mk = 0;
    for (size_t d = 0; d < dim; ++d)
      if (!skip || level[j * dim + d] > 1) mk |= 1u << d;
    pmask[j] = mk;
  }
  std::vector<size_t> order(N);
  for (size_t j = 0; j < N; ++j) order[j] = j;
  if (skip)
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return pmask[a] < pmask[b]; });
  std::vector<uint32_t> lv(N * dim), ix(N * dim);
  for (size_t s2 = 0; s2 < N; ++s2)
    for (size_t d = 0; d < dim; ++d) {
      lv[s2 * dim + d] = level[order[s2] * dim + d];
      ix[s2 * dim + d] = index[order[s2] * dim + d];
    }
  std::vector<double> lip, sc;
  encode_grid(basis, dim, N, lv.data(), ix.data(), lip, sc);
  auto m = std::make_shared<Model>();
  m->basis = basis;
  m->D = static_cast<int>(dim);
  m->N = N;
  m->Npad = ((N + kTG - 1) / kTG) * kTG;
  const bool f64 = cfg_.precision == Precision::FP64;
  const size_t esz = f64 ? 8 : 4;
  m->bytes = m->Npad * dim * 2 * esz + m->Npad * esz + (m->Npad / 4) * sizeof(uint32_t);
  const int K = f64 ? serve_k_rt<double>(m->D) : serve_k_rt<float>(m->D);
  m->tile = kBlock * K;
  const double per_point = static_cast<double>(m->tile) * (2 * dim + 1);
  size_t chunk = static_cast<size_t>(std::ceil(cfg_.item_work / per_point / kTG)) * kTG;
  chunk = std::max<size_t>(kTG, std::min(chunk, m->Npad));
  m->chunk = static_cast<int>(chunk);
  std::vector<double> A(m->Npad, 0.0);
  for (size_t s2 = 0; s2 < N; ++s2) A[s2] = alpha[order[s2]] * sc[s2];
  // tile-major (L, I); padded points (0, 0) evaluate to 1 with weight 0
  std::vector<double> li(m->Npad * dim * 2, 0.0);
  for (size_t j = 0; j < N; ++j)
    for (size_t d = 0; d < dim; ++d) {
      const size_t dst = (((j / kTG) * dim + d) * kTG + (j % kTG)) * 2;
      li[dst] = lip[(j * dim + d) * 2];
      li[dst + 1] = lip[(j * dim + d) * 2 + 1];
    }
  // padded points keep mask 0: nothing evaluated, weight 0
  std::vector<uint32_t> qm(m->Npad / 4, 0u);
  for (size_t s2 = 0; s2 < N; ++s2) qm[s2 / 4] |= pmask[order[s2]];
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (used_bytes_ + m->bytes > cfg_.memory_budget)
      throw input_error("model does not fit in the memory budget");
    used_bytes_ += m->bytes;
    int idx = -1;
    std::vector<bool> used(kMaxModels, false);
    for (auto& kv : models_) used[kv.second->table_index] = true;
    for (int i = 0; i < kMaxModels; ++i)
      if (!used[i]) {
        idx = i;
        break;
      }
    if (idx < 0) {
      used_bytes_ -= m->bytes;
      throw input_error("too many resident models");
    }
    m->table_index = idx;
    m->id = next_model_++;
    models_[m->id] = m;
  }
  try {
    check_hip(hipSetDevice(cfg_.device), "hipSetDevice");
    check_hip(hipMalloc(&m->LI, m->Npad * dim * 2 * esz), "hipMalloc(model LI)");
    check_hip(hipMalloc(&m->A, m->Npad * esz), "hipMalloc(model A)");
    check_hip(hipMalloc(&m->tmask, qm.size() * sizeof(uint32_t)), "hipMalloc(model masks)");
    check_hip(hipMemcpy(m->tmask, qm.data(), qm.size() * sizeof(uint32_t), hipMemcpyHostToDevice),
              "upload masks");
    if (f64) {
      check_hip(hipMemcpy(m->LI, li.data(), li.size() * 8, hipMemcpyHostToDevice), "upload LI");
      check_hip(hipMemcpy(m->A, A.data(), A.size() * 8, hipMemcpyHostToDevice), "upload A");
    } else {
      std::vector<float> l32(li.begin(), li.end()), a32(A.begin(), A.end());
      check_hip(hipMemcpy(m->LI, l32.data(), l32.size() * 4, hipMemcpyHostToDevice), "upload LI");
      check_hip(hipMemcpy(m->A, a32.data(), a32.size() * 4, hipMemcpyHostToDevice), "upload A");
    }
    ModelDesc md{m->LI, m->A, static_cast<const uint32_t*>(m->tmask), m->D,
                 static_cast<int>(m->Npad)};
    check_hip(hipMemcpy(static_cast<ModelDesc*>(model_table_) + m->table_index, &md,
                        sizeof(ModelDesc), hipMemcpyHostToDevice),
              "upload model descriptor");
  } catch (...) {
    std::lock_guard<std::mutex> lk(mu_);
    models_.erase(m->id);
    used_bytes_ -= m->bytes;
    if (m->LI) (void)hipFree(m->LI);
    if (m->A) (void)hipFree(m->A);
    if (m->tmask) (void)hipFree(m->tmask);
    throw;
  }
  return m->id;
}

void Server::evict_model(int id) {
  std::shared_ptr<Model> m;
  {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = models_.find(id);
    if (it == models_.end()) throw input_error("unknown model");
    if (it->second->active > 0) throw input_error("model has queued or running requests");
    m = it->second;
    models_.erase(it);
    used_bytes_ -= m->bytes;
  }
  check_hip(hipFree(m->LI), "hipFree");
  check_hip(hipFree(m->A), "hipFree");
  check_hip(hipFree(m->tmask), "hipFree");
}

size_t Server::model_bytes(int id) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = models_.find(id);
  if (it == models_.end()) throw input_error("unknown model");
  return it->second->bytes;
}

size_t Server::memory_used() {
  std::lock_guard<std::mutex> lk(mu_);
  return used_bytes_;
}

uint64_t Server::submit(int model, size_t q, const double* queries, double deadline_s) {
  std::shared_ptr<Model> m;
  {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = models_.find(model);
    if (it == models_.end()) throw input_error("unknown model");
    m = it->second;
  }
  if (q == 0) throw input_error("empty request");
  const double t_enter = now_s();
  auto r = std::make_shared<Request>();
  r->model = model;
  r->q = q;
  r->ldq = ((q + m->tile - 1) / m->tile) * m->tile;
  const size_t D = m->D;
  const bool f64 = cfg_.precision == Precision::FP64;
  r->X.assign(D * r->ldq * (f64 ? 8 : 4), 0);
  for (size_t i = 0; i < q; ++i)
    for (size_t d = 0; d < D; ++d) {
      const double v = queries[i * D + d];
      if (!(v >= 0.0 && v <= 1.0)) throw input_error("query points must lie in [0,1]^d");
      if (f64)
        reinterpret_cast<double*>(r->X.data())[d * r->ldq + i] = v;
      else
        reinterpret_cast<float*>(r->X.data())[d * r->ldq + i] = static_cast<float>(v);
    }
  r->t_submit = t_enter;  // latency includes the conversion of the queries
  r->deadline = deadline_s > 0 ? r->t_submit + deadline_s : 0.0;
  {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = models_.find(model);
    if (it == models_.end()) throw input_error("model was evicted");
    it->second->active++;
    r->id = next_request_++;
    requests_[r->id] = r;
    queue_.push_back(r);
  }
  cv_.notify_all();
  return r->id;
}

int Server::poll(uint64_t id, double* out, double wait_s, double* latency_s) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = requests_.find(id);
  if (it == requests_.end()) return -1;
  std::shared_ptr<Request> r = it->second;
  auto finished = [&] { return r->status != REQ_QUEUED && r->status != REQ_RUNNING; };
  if (!finished() && wait_s != 0.0) {
    if (wait_s < 0)
      done_cv_.wait(lk, finished);
    else
      done_cv_.wait_for(lk, std::chrono::duration<double>(wait_s), finished);
  }
  if (!finished()) return r->status;
  if (r->status == REQ_DONE && out) std::memcpy(out, r->out.data(), r->q * sizeof(double));
  if (latency_s) *latency_s = r->t_done - r->t_submit;
  requests_.erase(id);
  return r->status;
}

bool Server::cancel(uint64_t id) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = requests_.find(id);
  if (it == requests_.end() || it->second->status != REQ_QUEUED) return false;
  fail_request(it->second, REQ_CANCELLED, "cancelled");
  return true;
}

void Server::inject_kernel_error() { inject_error_ = true; }

ServerStats Server::stats() {
  std::lock_guard<std::mutex> lk(mu_);
  return stats_;
}

void Server::drain() {
  std::unique_lock<std::mutex> lk(mu_);
  done_cv_.wait(lk, [&] { return queue_.empty() && running_ == 0; });
}

// caller holds mu_
void Server::fail_request(const std::shared_ptr<Request>& r, int status, const char* why) {
  r->status = status;
  r->error = why;
  r->t_done = now_s();
  auto m = models_.find(r->model);
  if (m != models_.end()) m->second->active--;
  stats_.requests_failed++;
  done_cv_.notify_all();
}

// caller holds mu_: move requests from the queue into slot s (FIFO)
bool Server::form_batch(Slot& s, std::vector<std::shared_ptr<Request>>& taken) {
  s.usedX = s.nItems = s.nRed = s.usedPart = s.usedOut = 0;
  const double t = now_s();
  const size_t esz = cfg_.precision == Precision::FP64 ? 8 : 4;
  while (!queue_.empty() && taken.size() < static_cast<size_t>(cfg_.max_requests_per_batch)) {
    std::shared_ptr<Request> r = queue_.front();
    if (r->status != REQ_QUEUED) {  // cancelled meanwhile
      queue_.pop_front();
      continue;
    }
    if (r->deadline > 0 && t > r->deadline) {
      queue_.pop_front();
      fail_request(r, REQ_FAILED_TIMEOUT, "deadline passed before the request was scheduled");
      continue;
    }
    const Model& m = *models_.at(r->model);
    const size_t ntiles = r->ldq / m.tile;
    const size_t nchunks = (m.Npad + m.chunk - 1) / m.chunk;
    const size_t items = ntiles * nchunks;
    const size_t part = items * m.tile;
    const size_t xel = static_cast<size_t>(m.D) * r->ldq;
    if (items > s.capItems || part > s.capPart || xel > s.capX || r->q > s.capOut ||
        ntiles > s.capRed) {
      queue_.pop_front();
      fail_request(r, REQ_FAILED_BUDGET, "request exceeds the batch buffer capacity");
      continue;
    }
    if (s.nItems + items > s.capItems || s.usedPart + part > s.capPart ||
        s.usedX + xel > s.capX || s.usedOut + r->q > s.capOut || s.nRed + ntiles > s.capRed)
      break;  // full: the rest waits for the next slot (FIFO)
    queue_.pop_front();
    r->status = REQ_RUNNING;
    r->xoff = static_cast<long long>(s.usedX);
    r->out_off = static_cast<long long>(s.usedOut);
    for (size_t tI = 0; tI < ntiles; ++tI) {
      RedItem& ri = s.hRed[s.nRed++];
      ri.pbase = static_cast<long long>(s.usedPart + tI * nchunks * m.tile);
      ri.nchunks = static_cast<int>(nchunks);
      ri.nq = static_cast<int>(std::min<size_t>(m.tile, r->q - tI * m.tile));
      ri.tile = m.tile;
      ri.pad = 0;
      ri.out = r->out_off + static_cast<long long>(tI * m.tile);
      for (size_t c = 0; c < nchunks; ++c) {
        Item& it = s.hItems[s.nItems++];
        it.model = m.table_index;
        it.D = m.D;
        it.q0 = static_cast<int>(tI * m.tile);
        it.nq = ri.nq;
        it.j0 = static_cast<int>(c * m.chunk);
        it.j1 = static_cast<int>(std::min<size_t>((c + 1) * m.chunk, m.Npad));
        it.xoff = r->xoff;
        it.ldq = static_cast<int>(r->ldq);
        it.pad = 0;
        it.poff = ri.pbase + static_cast<long long>(c * m.tile);
      }
    }
    s.usedPart += part;
    s.usedX += xel;
    s.usedOut += r->q;
    (void)esz;
    taken.push_back(r);
  }
  return !taken.empty();
}

void Server::launch_batch(Slot& s) {
  const size_t esz = cfg_.precision == Precision::FP64 ? 8 : 4;
  for (auto& r : s.reqs) std::memcpy(s.hX + r->xoff * esz, r->X.data(), r->X.size());
  s.hCounts[0] = static_cast<int>(s.nItems);
  s.hCounts[1] = static_cast<int>(s.nRed);
  s.hCounts[2] = s.hCounts[3] = 0;
  check_hip(hipEventRecord(s.ev_start, s.stream), "hipEventRecord");
  check_hip(hipMemcpyAsync(s.dX, s.hX, s.usedX * esz, hipMemcpyHostToDevice, s.stream), "H2D X");
  check_hip(hipMemcpyAsync(s.dItems, s.hItems, s.nItems * sizeof(Item), hipMemcpyHostToDevice,
                           s.stream),
            "H2D items");
  check_hip(hipMemcpyAsync(s.dRed, s.hRed, s.nRed * sizeof(RedItem), hipMemcpyHostToDevice,
                           s.stream),
            "H2D red");
  check_hip(hipMemcpyAsync(s.dCounts, s.hCounts, 4 * sizeof(int), hipMemcpyHostToDevice,
                           s.stream),
            "H2D counts");
  if (s.graph) {
    check_hip(hipGraphLaunch(s.graph, s.stream), "hipGraphLaunch");
  } else {
    clear_launch_error();
    hipLaunchKernelGGL(reset_counter, dim3(1), dim3(64), 0, s.stream, s.dCounter);
    if (cfg_.precision == Precision::FP64)
      hipLaunchKernelGGL(serve_kernel<double>, dim3(cfg_.blocks), dim3(kBlock), 0, s.stream,
                         static_cast<const ModelDesc*>(model_table_), s.dItems, s.dCounts,
                         s.dCounter, static_cast<const double*>(s.dX), s.dPart);
    else
      hipLaunchKernelGGL(serve_kernel<float>, dim3(cfg_.blocks), dim3(kBlock), 0, s.stream,
                         static_cast<const ModelDesc*>(model_table_), s.dItems, s.dCounts,
                         s.dCounter, static_cast<const float*>(s.dX), s.dPart);
    hipLaunchKernelGGL(serve_reduce, dim3(2 * num_cus_), dim3(kBlock), 0, s.stream, s.dRed,
                       s.dCounts, s.dPart, s.dOut);
    check_hip(hipGetLastError(), "serve launch");
  }
  check_hip(hipMemcpyAsync(s.hOut, s.dOut, s.usedOut * sizeof(double), hipMemcpyDeviceToHost,
                           s.stream),
            "D2H out");
  check_hip(hipEventRecord(s.ev_done, s.stream), "hipEventRecord");
}

// caller holds mu_
void Server::finish_batch(Slot& s) {
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, s.ev_start, s.ev_done) == hipSuccess)
    stats_.busy_seconds += ms * 1e-3;
  const double t = now_s();
  for (auto& r : s.reqs) {
    if (s.inject) {
      fail_request(r, REQ_FAILED_DEVICE, "injected kernel error");
      continue;
    }
    r->out.assign(s.hOut + r->out_off, s.hOut + r->out_off + r->q);
    r->X.clear();
    r->X.shrink_to_fit();
    r->status = REQ_DONE;
    r->t_done = t;
    auto m = models_.find(r->model);
    if (m != models_.end()) m->second->active--;
    stats_.requests_done++;
    stats_.queries_done += r->q;
  }
  stats_.batches++;
  stats_.kernel_launches += cfg_.use_graph ? 1 : 3;
  s.reqs.clear();
  s.busy = false;
  s.inject = false;
  running_--;
  done_cv_.notify_all();
}

void Server::loop() {
  (void)hipSetDevice(cfg_.device);
  std::unique_lock<std::mutex> lk(mu_);
  for (;;) {
    bool any_busy = false;
    for (auto& sp : slots_) {
      Slot& s = *sp;
      if (!s.busy) continue;
      const hipError_t q = hipEventQuery(s.ev_done);
      if (q == hipSuccess) {
        finish_batch(s);
      } else if (q == hipErrorNotReady) {
        any_busy = true;
      } else {
        for (auto& r : s.reqs) fail_request(r, REQ_FAILED_DEVICE, hipGetErrorString(q));
        s.reqs.clear();
        s.busy = false;
        running_--;
      }
    }
    for (auto& sp : slots_) {
      Slot& s = *sp;
      if (s.busy || queue_.empty()) continue;
      std::vector<std::shared_ptr<Request>> taken;
      if (!form_batch(s, taken)) continue;
      s.reqs = std::move(taken);
      s.busy = true;
      s.inject = inject_error_.exchange(false);
      running_++;
      lk.unlock();
      bool ok = true;
      std::string err;
      try {
        launch_batch(s);
      } catch (const std::exception& e) {
        ok = false;
        err = e.what();
      }
      lk.lock();
      if (!ok) {
        for (auto& r : s.reqs) fail_request(r, REQ_FAILED_DEVICE, err.c_str());
        s.reqs.clear();
        s.busy = false;
        running_--;
      } else {
        any_busy = true;
      }
    }
    if (queue_.empty() && running_ == 0) done_cv_.notify_all();
    if (stop_ && !any_busy) break;
    if (any_busy)
      cv_.wait_for(lk, std::chrono::microseconds(20));
    else
      cv_.wait(lk, [&] { return stop_ || !queue_.empty(); });
  }
}

}  // namespace sgpp_hip
