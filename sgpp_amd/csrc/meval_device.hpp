// Device side of the multi-evaluation operator: buffers, launches, timing.
//
// The operator evaluates B (M x N, B_ij = phi_j(x_i)) on a sparse grid with
// the Linear or ModLinear hat basis:
//   mult:          r = B * alpha   (grid -> data)
//   multTranspose: a = B^T * y     (data -> grid)
// Everything here is owned by one DeviceMultiEval object; the C++ operator
// class (OperationMultiEvalHIP) and the C-ABI sit on top of it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace sgpp_hip {

enum class Basis : int { Linear = 0, ModLinear = 1 };
enum class Precision : int { FP64 = 0, FP32 = 1 };
// Streaming: every (data point, grid point) pair, culled by supports.
// Subspace: per data point and subspace (level vector), the one basis function
//   whose support holds the point, looked up in a dense index table (FP64).
enum class Algorithm : int { Streaming = 0, Subspace = 1 };

// Error thrown for any failing HIP call or invalid launch shape.
struct device_error : public std::runtime_error {
  explicit device_error(const std::string& m) : std::runtime_error(m) {}
};

// Error thrown for inputs the operator cannot evaluate (bad sizes, data
// outside the unit cube, level/index pairs that are not no-boundary points).
struct input_error : public std::runtime_error {
  explicit input_error(const std::string& m) : std::runtime_error(m) {}
};

void check_hip(hipError_t e, const char* what);

// hipGetLastError() also returns an error that an earlier HIP call left
// behind in this thread, including calls of other code in the process that
// ignores its own errors. Every launch here is preceded by this call, so the
// check after the launch reports that launch only.
inline void clear_launch_error() { (void)hipGetLastError(); }

// Launch tuning knobs; 0 means "choose automatically".
struct LaunchConfig {
  int mult_grid_chunks = 0;   // G: grid-point chunks for mult (partials over G)
  int multT_data_segs = 0;    // S: data segments for multTranspose (partials over S)
  int target_blocks = 0;      // workgroups to aim for per launch
};

// Encodes level/index pairs of one grid into the (L, I) pair per point and
// dimension that the kernels consume, plus a per-point power-of-two scale:
//   Linear:    phi = max(0, 1 - |L x - I|),      L = 2^l, I = i, scale 1
//   ModLinear: level 1            -> (0, 0)       (phi == 1)
//              i == 1             -> (2^l/2, 0)   (phi/2 = 1 - 2^l x / 2)
//              i == 2^l - 1       -> (2^l/2, 2^l/2)
//              otherwise          -> (2^l, i)
//              scale = 2^(number of modified boundary factors)
// Both bases then share the kernel step p <- clamp(p - p*|L x - I|).
void encode_grid(Basis basis, size_t dim, size_t n, const uint32_t* level, const uint32_t* index,
                 std::vector<double>& li, std::vector<double>& scale);

// Morton order of n points in [0,1]^d (row-major); ties keep input order.
std::vector<int32_t> morton_order(const double* pts, size_t n, size_t d);

// Grid storage order for culling: by the point's set of non-constant
// dimensions (pmask), then by the Morton code of its support's centre.
// li: point-major (L, I) pairs from encode_grid.
std::vector<int32_t> grid_order(const std::vector<double>& li, const std::vector<uint32_t>& pmask,
                                size_t n, size_t d);

// Per 64-point tile and dimension, the union of the supports of the tile's
// points below n, [(I - 1) / L, (I + 1) / L] (the whole line for L = 0),
// rounded outwards to T (an empty box when the tile holds none). li:
// point-major (L, I) pairs of npad points (npad a multiple of 64).
// Returns [npad/64][d][2].
template <typename T>
std::vector<T> tile_boxes(const std::vector<double>& li, size_t npad, size_t n, size_t d);

class DeviceMultiEval {
 public:
  // data_rowmajor: M x d, FP64 row-major (the caller's DataMatrix layout).
  // precisions: bitmask of (1 << Precision) that the instance must serve.
  // The subspace algorithm computes in FP64 only (precisions must be FP64)
  // and supports dim <= 20.
  DeviceMultiEval(const double* data_rowmajor, size_t M, size_t dim, Basis basis,
                  unsigned precisions, int device, const LaunchConfig& cfg,
                  Algorithm algorithm = Algorithm::Streaming);
  ~DeviceMultiEval();
  DeviceMultiEval(const DeviceMultiEval&) = delete;
  DeviceMultiEval& operator=(const DeviceMultiEval&) = delete;

  // level/index: N x d row-major uint32 (the grid storage's point order).
  void set_grid(const uint32_t* level, const uint32_t* index, size_t N);

  // Device-pointer entry points (inputs/outputs already in HBM, FP64).
  // alpha: N values, result: M values; y: M values, result: N values.
  void mult_device(const double* d_alpha, double* d_result, Precision p);
  // mult_device without any allocation: the alpha weights and the FP64
  // partials live in the library's fixed device scratch (kNoAllocScratchBytes
  // per GPU, shared by every call on that GPU: calls that overlap on one GPU
  // race). The grid chunks are capped so that the partials fit; inputs whose
  // weights and one row of partials do not fit are refused (input_error).
  void mult_device_noalloc(const double* d_alpha, double* d_result, Precision p);
  void mult_transpose_device(const double* d_y, double* d_result, Precision p);

  // Host-pointer entry points (copy in, compute, copy out, synchronise).
  void mult_host(const double* alpha, double* result, Precision p);
  void mult_transpose_host(const double* y, double* result, Precision p);

  void synchronize();
  hipStream_t stream() const { return stream_; }
  size_t num_data() const { return M_; }
  size_t num_data_padded() const { return Mpad_; }
  size_t num_grid() const { return N_; }
  size_t dim() const { return D_; }
  Basis basis() const { return basis_; }
  Algorithm algorithm() const { return algorithm_; }
  // subspace algorithm: number of subspaces and of index table entries
  int num_subspaces() const { return nsub_; }
  int64_t num_table_entries() const { return ntable_; }
  bool serves(Precision p) const { return (precisions_ >> static_cast<int>(p)) & 1u; }

  // Kernel-only milliseconds of the last mult / multTranspose (HIP events).
  double last_kernel_ms() const { return last_kernel_ms_; }
  // Wall-clock seconds of the last host-pointer call (copies included).
  double last_wall_s() const { return last_wall_s_; }

  // Shapes the last launches used (for reporting).
  int last_mult_chunks() const { return last_G_; }
  int last_multT_segs() const { return last_S_; }

  // HIP-event timing of the main mult / multTranspose kernel of every launch
  // (prep and reduction kernels excluded), on this operator's stream.
  // Enabling resets the totals; timing() synchronises and sums the launches
  // since then.
  void set_kernel_timing(bool on);
  void kernel_timing(int* n_mult, double* ms_mult, int* n_multT, double* ms_multT);

  // Work the culled kernels execute for one mult and one multTranspose in
  // precision p, counted by the kernels' counting instantiations (launched
  // here, untimed, on zero inputs; culling does not depend on the values):
  // out[0], out[1] = mult hat steps (pair x evaluated dimension) and pairs
  // (every lane of every evaluated group); out[2], out[3] = the same for
  // multTranspose. Synchronises the stream.
  void count_work(Precision p, uint64_t out[4]);

 private:
  void ensure(void** buf, size_t* cap, size_t bytes);
  void record_start();
  void record_stop();
  // Caller-held scratch of a mult launch: weights in the compute type, then
  // the FP64 partials.
  struct MultScratch {
    void* base;
    size_t bytes;
  };
  template <typename T>
  void run_mult(const T* X, const T* LI, const float* TB, const double* d_alpha, double* d_result,
                hipEvent_t* tev, const MultScratch* ext = nullptr);
  template <typename T>
  void run_multT(const T* X, const T* LI, const float* TB, const double* d_y, double* d_result,
                 hipEvent_t* tev);
  // data-group boxes of the templated kernels (set at construction)
  template <typename T>
  void upload_group_boxes(const std::vector<T>& soa, void** dst);
  // block order of the templated mult / multTranspose kernels per precision
  // (index 0 FP64, 1 FP32): 1 = grid chunks fastest; timed when the grid is
  // set
  int mult_cf_[2] = {0, 0};
  int multT_cf_[2] = {0, 0};
  template <typename T>
  void choose_mult_order(const T* X, const T* LI, const float* TB);
  template <typename T>
  const T* group_boxes() const {
    return static_cast<const T*>(sizeof(T) == sizeof(double) ? WB64_ : WB32_);
  }
  // the subspace algorithm (subspace_kernels.hip)
  void set_grid_subspace(const uint32_t* level, const uint32_t* index, size_t N);
  void run_subspace_mult(const double* d_alpha, double* d_result, hipEvent_t* tev);
  void run_subspace_multT(const double* d_y, double* d_result, hipEvent_t* tev);

  // Storage orders. Data: Morton order, dperm_[stored row] = caller's row.
  // Grid: by each point's set of non-constant dimensions, then Morton order
  // of the support centres; perm_[stored position] = caller's index.
  void* dperm_ = nullptr;
  void* perm_ = nullptr;   size_t perm_cap_ = 0;
  void* pm_ = nullptr;     size_t pm_cap_ = 0;   // [Npad] non-constant dimensions per point

  size_t M_ = 0, Mpad_ = 0, D_ = 0, N_ = 0, Npad_ = 0;
  Basis basis_;
  unsigned precisions_;
  Algorithm algorithm_ = Algorithm::Streaming;
  int device_ = 0;
  LaunchConfig cfg_;
  int num_cus_ = 256;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;

  // data in storage order, SoA dim-major [D][Mpad]
  void* X64_ = nullptr;
  void* X32_ = nullptr;
  // per data group (one wave's 64*K consecutive points in the templated
  // kernels, d <= 20) the box of its points, [Mpad/(64*K)][D][2] {lo, hi}, in
  // the data's precision
  void* WB64_ = nullptr;
  void* WB32_ = nullptr;
  // grid pairs, point-major [Npad][D][2], per-point scale [Npad] and
  // per-tile (64 points) boxes of the supports [Npad/64][D][2], FP32 rounded
  // outwards for both precisions
  void* LI64_ = nullptr;
  void* LI32_ = nullptr;
  void* TB64_ = nullptr;
  void* TB32_ = nullptr;
  double* scale_ = nullptr;
  size_t gridcap_li64_ = 0, gridcap_li32_ = 0, gridcap_scale_ = 0, tbcap64_ = 0, tbcap32_ = 0;

  // subspace algorithm: subspace records (sorted by level vector), the index
  // tables (grid point per cell, -1 where the grid has none), each grid
  // point's table entry, and a table-shaped FP64 buffer (mult weights);
  // sub_recompute_(t_) = dimensions mult (multTranspose) evaluates per data row
  int nsub_ = 0;
  int64_t ntable_ = 0;
  uint64_t sub_recompute_ = 0, sub_recompute_t_ = 0;
  void* sub_desc_ = nullptr;  size_t sub_desc_cap_ = 0;
  void* sub_ti_ = nullptr;    size_t sub_ti_cap_ = 0;
  void* sub_te_ = nullptr;    size_t sub_te_cap_ = 0;
  void* sub_tw_ = nullptr;    size_t sub_tw_cap_ = 0;
  void* sub_chunk_ = nullptr; size_t sub_chunk_cap_ = 0;  // multTranspose chunks
  int nchunk_ = 0;

  // scratch
  void* A_ = nullptr;     size_t A_cap_ = 0;      // alpha*scale in compute type
  void* Y_ = nullptr;     size_t Y_cap_ = 0;      // padded y in compute type
  void* part_ = nullptr;  size_t part_cap_ = 0;   // FP64 partials
  void* hin_ = nullptr;   size_t hin_cap_ = 0;    // device staging for host calls
  void* hout_ = nullptr;  size_t hout_cap_ = 0;

  struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
  };
  EventPair* next_timer(std::vector<EventPair>& pool, size_t& used);
  bool timing_on_ = false;
  unsigned long long* count_ = nullptr;  // set while count_work launches
  std::vector<EventPair> tm_mult_, tm_multT_;
  size_t used_mult_ = 0, used_multT_ = 0;

  double last_kernel_ms_ = 0.0;
  double last_wall_s_ = 0.0;
  int last_G_ = 0, last_S_ = 0;
};

// Size of the fixed device scratch of mult_device_noalloc (per GPU).
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;

// Padding units (points): data is padded to a multiple of kDataPad by
// repeating the last stored row, the grid to a multiple of kGridPad with
// points that evaluate to a constant and carry zero weight.
constexpr size_t kDataPad = 4096;
constexpr size_t kGridPad = 2048;

}  // namespace sgpp_hip
