// gfx950 kernels of the multi-evaluation operator and the device runtime
// that launches them.
//
// Arithmetic per (data point x, grid point j):
//   phi_j(x) = prod_d max(0, 1 - |L_jd * x_d - I_jd|)
// evaluated as the two-instruction step
//   u = fma(L, x, -I);  p = clamp01(fma(-p, |u|, p))      (p in [0,1])
// which the compiler emits as v_fma + v_fma(neg, abs, clamp). This is VALU
// work (no contraction), so there is no MFMA anywhere in this file.
//
// Layout: data points SoA [D][Mpad] in Morton order (dperm maps a stored row
// to the caller's row); grid points point-major (L, I) rows [Npad][D][2],
// sorted by their set of non-constant dimensions and the Morton code of their
// support centre (perm maps a stored point to the caller's index), with
// per-64-point-tile support boxes and per-point dimension masks.
//
// Both kernels give lanes K data points each (coordinates in registers). A
// wave takes the box of its 64*K points, skips grid tiles whose box misses
// it, tests the remaining tiles' points against it with one ballot and
// evaluates the points that may be nonzero, J at a time, over the union of
// their non-constant dimensions (ModLinear level-1 factors are exactly 1).
// Every dropped term is exactly zero (cull_device.hpp), so results equal a
// full evaluation.
//   mult (r = B alpha): blocks = data tiles (x) x G grid chunks (y); the G
//     FP64 partial rows are summed in a fixed order by reduce_parts.
//   multTranspose (a = B^T y): blocks = data segments (x) x grid chunks of
//     kChunkT points (y); per taken grid point the wave's sum over its lanes
//     (DPP) goes to a per-tile register, and each tile is added to the
//     wave's row of FP64 LDS accumulators; the four rows and the S partial
//     rows are summed in a fixed order. No global atomics: results are
//     deterministic.
// Precision: FP64 throughout, or FP32 basis products and per-tile sums with
//   the per-tile partials accumulated in FP64 registers and FP64 partials.
//
// Reference semantics: datadriven/src/sgpp/datadriven/operation/hash/
//   OperationMultiEvalStreaming/OperationMultiEvalStreaming_multImpl.cpp:576-623 (mult),
//   ..._multTransposeImpl.cpp:483-508 (multTranspose),
//   OperationMultiEvalModMaskStreaming/..._multImpl.cpp:604-656 (ModLinear).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <sstream>

#include "meval_kernels.hpp"

namespace sgpp_hip {


// Morton order of n points in [0,1]^d (row-major): b = min(16, 64/d) bits
// per coordinate, interleaved from the top bit; ties keep the input order.
std::vector<int32_t> morton_order(const double* pts, size_t n, size_t d) {
  const int b = static_cast<int>(std::max<size_t>(1, std::min<size_t>(16, 64 / d)));
  const size_t dd = std::min<size_t>(d, 64 / static_cast<size_t>(b));
  const double scale = static_cast<double>(1u << b);
  const uint32_t top = (1u << b) - 1u;
  std::vector<std::pair<uint64_t, int32_t>> key(n);
  std::vector<uint32_t> q(dd);
  for (size_t i = 0; i < n; ++i) {
    for (size_t k = 0; k < dd; ++k) {
      const double v = pts[i * d + k] * scale;
      q[k] = v <= 0.0 ? 0u : std::min<uint32_t>(top, static_cast<uint32_t>(v));
    }
    uint64_t c = 0;
    for (int bit = b - 1; bit >= 0; --bit)
      for (size_t k = 0; k < dd; ++k) c = (c << 1) | ((q[k] >> bit) & 1u);
    key[i] = {c, static_cast<int32_t>(i)};
  }
  std::sort(key.begin(), key.end());
  std::vector<int32_t> order(n);
  for (size_t i = 0; i < n; ++i) order[i] = key[i].second;
  return order;
}

// Grid storage order: by the point's set of non-constant dimensions, then by
// the Morton code of its support's centre (li: point-major (L, I) pairs).
std::vector<int32_t> grid_order(const std::vector<double>& li, const std::vector<uint32_t>& pmask,
                                size_t n, size_t d) {
  std::vector<double> centre(n * d);
  for (size_t j = 0; j < n; ++j)
    for (size_t k = 0; k < d; ++k) {
      const double L = li[(j * d + k) * 2], I = li[(j * d + k) * 2 + 1];
      const double c = L > 0.0 ? I / L : 0.5;
      centre[j * d + k] = std::min(1.0, std::max(0.0, c));
    }
  const std::vector<int32_t> mo = morton_order(centre.data(), n, d);
  std::vector<int32_t> rank(n);
  for (size_t i = 0; i < n; ++i) rank[mo[i]] = static_cast<int32_t>(i);
  std::vector<int32_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
    return pmask[a] != pmask[b] ? pmask[a] < pmask[b] : rank[a] < rank[b];
  });
  return order;
}

// Per 64-point tile and dimension the union of the supports of its points
// below n, [(I - 1) / L, (I + 1) / L] (whole line for L = 0), from
// point-major (L, I) pairs given as doubles, stored in T rounded outwards; a
// tile with no such point gets an empty box.
template <typename T>
std::vector<T> tile_boxes(const std::vector<double>& li, size_t npad, size_t n, size_t d) {
  const size_t ntiles = npad / kTG;
  std::vector<T> tb(ntiles * d * 2);
  for (size_t t = 0; t < ntiles; ++t)
    for (size_t k = 0; k < d; ++k) {
      double lo = 2.0, hi = -1.0;
      for (size_t p = 0; p < static_cast<size_t>(kTG) && t * kTG + p < n; ++p) {
        const size_t e = ((t * kTG + p) * d + k) * 2;
        const double L = li[e], I = li[e + 1];
        if (L > 0.0) {
          lo = std::min(lo, (I - 1.0) / L);
          hi = std::max(hi, (I + 1.0) / L);
        } else {
          lo = -1.0;
          hi = 2.0;
        }
      }
      T tl = static_cast<T>(lo), th = static_cast<T>(hi);
      if (static_cast<double>(tl) > lo) tl = std::nextafter(tl, static_cast<T>(-3));
      if (static_cast<double>(th) < hi) th = std::nextafter(th, static_cast<T>(3));
      tb[(t * d + k) * 2] = tl;
      tb[(t * d + k) * 2 + 1] = th;
    }
  return tb;
}

template std::vector<float> tile_boxes<float>(const std::vector<double>&, size_t, size_t, size_t);
template std::vector<double> tile_boxes<double>(const std::vector<double>&, size_t, size_t,
                                                size_t);

// =================================================================== runtime

void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::ostringstream os;
    os << what << ": " << hipGetErrorString(e) << " (" << static_cast<int>(e) << ")";
    throw device_error(os.str());
  }
}

void encode_grid(Basis basis, size_t dim, size_t n, const uint32_t* level, const uint32_t* index,
                 std::vector<double>& li, std::vector<double>& scale) {
  li.assign(n * dim * 2, 0.0);
  scale.assign(n, 1.0);
  for (size_t j = 0; j < n; ++j) {
    double s = 1.0;
    for (size_t d = 0; d < dim; ++d) {
      const uint32_t l = level[j * dim + d];
      const uint32_t i = index[j * dim + d];
      if (l < 1 || l > 30 || (i & 1u) == 0 || i >= (1u << l))
        throw input_error("invalid grid point (level/index out of range for a no-boundary grid)");
      const double h = static_cast<double>(1u << l);
      double L = h, I = static_cast<double>(i);
      if (basis == Basis::ModLinear) {
        if (l == 1) {
          L = 0.0;
          I = 0.0;
        } else if (i == 1) {
          L = 0.5 * h;
          I = 0.0;
          s *= 2.0;
        } else if (i == (1u << l) - 1) {
          L = 0.5 * h;
          I = 0.5 * h;
          s *= 2.0;
        }
      }
      li[(j * dim + d) * 2] = L;
      li[(j * dim + d) * 2 + 1] = I;
    }
    scale[j] = s;
  }
}

DeviceMultiEval::DeviceMultiEval(const double* data, size_t M, size_t dim, Basis basis,
                                 unsigned precisions, int device, const LaunchConfig& cfg,
                                 Algorithm algorithm)
    : M_(M),
      D_(dim),
      basis_(basis),
      precisions_(precisions),
      algorithm_(algorithm),
      device_(device),
      cfg_(cfg) {
  if (M == 0 || dim == 0) throw input_error("empty dataset");
  if (precisions == 0 || (precisions & ~3u)) throw input_error("invalid precision mask");
  if (dim > 64) throw input_error("dimension above 64 is not supported");
  if (algorithm == Algorithm::Subspace && precisions != 1u)
    throw input_error("the subspace algorithm computes in FP64 only");
  if (algorithm == Algorithm::Subspace && dim > 20)
    throw input_error("the subspace algorithm supports dimensions up to 20");
  check_hip(hipSetDevice(device_), "hipSetDevice");
  hipDeviceProp_t prop;
  check_hip(hipGetDeviceProperties(&prop, device_), "hipGetDeviceProperties");
  num_cus_ = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  check_hip(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  check_hip(hipEventCreate(&ev0_), "hipEventCreate");
  check_hip(hipEventCreate(&ev1_), "hipEventCreate");

  for (size_t i = 0; i < M_ * D_; ++i)
    if (!(data[i] >= 0.0 && data[i] <= 1.0)) throw input_error("data points must lie in [0,1]^d");
  // at least kDataSlack rows past the data: a wave's last group of points
  // may start just below the end of the data
  Mpad_ = ((M_ + kDataSlack + kDataPad - 1) / kDataPad) * kDataPad;
  // storage order: Morton order of the points, so that consecutive points
  // (one wave's) fill small boxes; dperm[stored position] = caller's row
  std::vector<int32_t> dperm = morton_order(data, M_, D_);
  // pad by repeating the last stored row, transpose to [D][Mpad]
  std::vector<double> soa(D_ * Mpad_);
  for (size_t d = 0; d < D_; ++d) {
    double* col = soa.data() + d * Mpad_;
    for (size_t m = 0; m < M_; ++m) col[m] = data[static_cast<size_t>(dperm[m]) * D_ + d];
    const double last = col[M_ - 1];
    for (size_t m = M_; m < Mpad_; ++m) col[m] = last;
  }
  check_hip(hipMalloc(&dperm_, M_ * sizeof(int32_t)), "hipMalloc(data order)");
  check_hip(hipMemcpy(dperm_, dperm.data(), M_ * sizeof(int32_t), hipMemcpyHostToDevice),
            "upload data order");
  const bool templ = D_ <= static_cast<size_t>(kMaxTemplD);
  if (serves(Precision::FP64)) {
    check_hip(hipMalloc(&X64_, soa.size() * sizeof(double)), "hipMalloc(X64)");
    check_hip(hipMemcpy(X64_, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice),
              "upload X64");
    if (templ) upload_group_boxes(soa, &WB64_);
  }
  if (serves(Precision::FP32)) {
    std::vector<float> s32(soa.begin(), soa.end());
    check_hip(hipMalloc(&X32_, s32.size() * sizeof(float)), "hipMalloc(X32)");
    check_hip(hipMemcpy(X32_, s32.data(), s32.size() * sizeof(float), hipMemcpyHostToDevice),
              "upload X32");
    if (templ) upload_group_boxes(s32, &WB32_);
  }
}

// Boxes of the data groups of the templated kernels (rows [g*64K, (g+1)*64K)
// of the stored data; mult and multTranspose use the same K), uploaded to
// *dst as [Mpad/(64K)][D][2] {lo, hi}. Rounding to FP32 is monotone, so the
// FP32 boxes are the rounded FP64 ones.
template <typename T>
void DeviceMultiEval::upload_group_boxes(const std::vector<T>& soa, void** dst) {
  const int K = mult_k_rt<T>(static_cast<int>(D_));
  if (K != multT_k_rt<T>(static_cast<int>(D_)))
    throw device_error("mult and multTranspose data groups differ");
  const size_t group = 64 * static_cast<size_t>(K);
  if (Mpad_ % group != 0) throw device_error("padded data is not a whole number of groups");
  const size_t ng = Mpad_ / group;
  std::vector<T> wb(ng * D_ * 2);
  for (size_t d = 0; d < D_; ++d) {
    const T* col = soa.data() + d * Mpad_;
    for (size_t g = 0; g < ng; ++g) {
      T lo = col[g * group], hi = lo;
      for (size_t m = g * group + 1; m < (g + 1) * group; ++m) {
        lo = std::min(lo, col[m]);
        hi = std::max(hi, col[m]);
      }
      wb[(g * D_ + d) * 2] = lo;
      wb[(g * D_ + d) * 2 + 1] = hi;
    }
  }
  check_hip(hipMalloc(dst, wb.size() * sizeof(T)), "hipMalloc(data group boxes)");
  check_hip(hipMemcpy(*dst, wb.data(), wb.size() * sizeof(T), hipMemcpyHostToDevice),
            "upload data group boxes");
}

DeviceMultiEval::~DeviceMultiEval() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  void* bufs[] = {X64_,   X32_,   WB64_,  WB32_,  LI64_,    LI32_,   TB64_,
                  TB32_,  scale_, perm_,  pm_,    dperm_,   A_,      Y_,
                  part_,  hin_,   hout_,  sub_desc_, sub_ti_, sub_te_, sub_tw_,
                  sub_chunk_};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (auto* pool : {&tm_mult_, &tm_multT_})
    for (auto& e : *pool) {
      (void)hipEventDestroy(e.a);
      (void)hipEventDestroy(e.b);
    }
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void DeviceMultiEval::ensure(void** buf, size_t* cap, size_t bytes) {
  if (*cap >= bytes && *buf) return;
  if (*buf) {
    check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    check_hip(hipFree(*buf), "hipFree");
    *buf = nullptr;
    *cap = 0;
  }
  check_hip(hipMalloc(buf, bytes), "hipMalloc(scratch)");
  *cap = bytes;
}

void DeviceMultiEval::set_grid(const uint32_t* level, const uint32_t* index, size_t N) {
  if (N == 0) throw input_error("empty grid");
  if (N > (size_t(1) << 30)) throw input_error("grid too large");
  if (algorithm_ == Algorithm::Subspace) {
    set_grid_subspace(level, index, N);
    return;
  }
  // non-constant dimensions of each point (ModLinear level-1 factors are 1)
  const bool mod = basis_ == Basis::ModLinear;
  std::vector<uint32_t> pmask(N);
  for (size_t j = 0; j < N; ++j) {
    uint32_t m = 0;
    for (size_t d = 0; d < D_ && d < 32; ++d)
      if (!mod || level[j * D_ + d] > 1) m |= 1u << d;
    pmask[j] = m;
  }
  std::vector<double> li0, sc0;
  encode_grid(basis_, D_, N, level, index, li0, sc0);
  // storage order: by the set of non-constant dimensions, then by the Morton
  // code of the support's centre, so that a 64-point tile holds points with
  // the same dimensions in a small region
  std::vector<int32_t> order = grid_order(li0, pmask, N, D_);
  std::vector<double> li(N * D_ * 2), sc(N);
  for (size_t s = 0; s < N; ++s) {
    const size_t j = static_cast<size_t>(order[s]);
    for (size_t e = 0; e < D_ * 2; ++e) li[s * D_ * 2 + e] = li0[j * D_ * 2 + e];
    sc[s] = sc0[j];
  }
  const size_t Npad = ((N + kGridPad - 1) / kGridPad) * kGridPad;
  // padded points: (L, I) = (1, 4) in every dimension, a factor that is
  // exactly 0 on [0, 1] (u = x - 4), so the culling tests drop them; weight 0
  // through A, rows ignored by the reductions
  li.resize(Npad * D_ * 2, 0.0);
  for (size_t s = N; s < Npad; ++s)
    for (size_t d = 0; d < D_; ++d) {
      li[(s * D_ + d) * 2] = 1.0;
      li[(s * D_ + d) * 2 + 1] = 4.0;
    }
  sc.resize(Npad, 1.0);
  // non-constant dimensions per stored point (0 for padding)
  std::vector<uint32_t> pm(Npad, 0u);
  for (size_t s = 0; s < N; ++s) pm[s] = pmask[order[s]];
  check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  N_ = N;
  Npad_ = Npad;
  ensure(&pm_, &pm_cap_, Npad_ * sizeof(uint32_t));
  check_hip(hipMemcpy(pm_, pm.data(), Npad_ * sizeof(uint32_t), hipMemcpyHostToDevice),
            "upload dimension masks");
  ensure(reinterpret_cast<void**>(&scale_), &gridcap_scale_, Npad_ * sizeof(double));
  check_hip(hipMemcpy(scale_, sc.data(), Npad_ * sizeof(double), hipMemcpyHostToDevice),
            "upload scale");
  ensure(&perm_, &perm_cap_, N_ * sizeof(int32_t));
  check_hip(hipMemcpy(perm_, order.data(), N_ * sizeof(int32_t), hipMemcpyHostToDevice),
            "upload permutation");
  if (serves(Precision::FP64)) {
    ensure(&LI64_, &gridcap_li64_, li.size() * sizeof(double));
    check_hip(hipMemcpy(LI64_, li.data(), li.size() * sizeof(double), hipMemcpyHostToDevice),
              "upload LI64");
    // FP32 boxes of the FP64 supports, rounded outwards
    const std::vector<float> tb = tile_boxes<float>(li, Npad_, N_, D_);
    ensure(&TB64_, &tbcap64_, tb.size() * sizeof(float));
    check_hip(hipMemcpy(TB64_, tb.data(), tb.size() * sizeof(float), hipMemcpyHostToDevice),
              "upload tile boxes");
  }
  if (serves(Precision::FP32)) {
    std::vector<float> l32(li.begin(), li.end());
    ensure(&LI32_, &gridcap_li32_, l32.size() * sizeof(float));
    check_hip(hipMemcpy(LI32_, l32.data(), l32.size() * sizeof(float), hipMemcpyHostToDevice),
              "upload LI32");
    // boxes of the FP32 (L, I) the FP32 kernels evaluate
    const std::vector<double> l32d(l32.begin(), l32.end());
    const std::vector<float> tb = tile_boxes<float>(l32d, Npad_, N_, D_);
    ensure(&TB32_, &tbcap32_, tb.size() * sizeof(float));
    check_hip(hipMemcpy(TB32_, tb.data(), tb.size() * sizeof(float), hipMemcpyHostToDevice),
              "upload tile boxes");
  }
  if (D_ <= static_cast<size_t>(kMaxTemplD) && !count_) {
    if (serves(Precision::FP64))
      choose_mult_order<double>(static_cast<const double*>(X64_),
                                static_cast<const double*>(LI64_),
                                static_cast<const float*>(TB64_));
    if (serves(Precision::FP32))
      choose_mult_order<float>(static_cast<const float*>(X32_), static_cast<const float*>(LI32_),
                               static_cast<const float*>(TB32_));
  }
}

// Block order of the templated mult and multTranspose kernels (the same
// choice for each): data blocks fastest (consecutive
// workgroups share a grid chunk, whose rows stay in the caches) or grid
// chunks fastest (a data block's coordinates stay in the caches). Both give
// the same partials, bit for bit. Which is faster depends on how densely the
// data hit the grid: on a d=10 ModLinear grid (1e5 points, 1e6 sparse hits)
// chunks fastest was 8 % faster in FP32, on a d=8 regular grid (dense hits)
// 3-8 % slower, and on a d=20 grid with 1e7 points (long blocks) 5 % slower.
// So both orders are timed once per grid, on zero weights (the culled work
// does not depend on them), and the faster kept; a first launch longer than
// kOrderProbeMs keeps data blocks fastest without further runs.
constexpr float kOrderProbeMs = 20.0f;

template <typename T>
void DeviceMultiEval::choose_mult_order(const T* X, const T* LI, const float* TB) {
  const int ix = sizeof(T) == sizeof(double) ? 0 : 1;
  mult_cf_[ix] = multT_cf_[ix] = 0;
  double* dA = nullptr;
  double* dR = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  auto release = [&]() {
    if (dA) (void)hipFree(dA);
    if (dR) (void)hipFree(dR);
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  };
  try {
    check_hip(hipMalloc(reinterpret_cast<void**>(&dA), N_ * sizeof(double)), "hipMalloc(probe)");
    check_hip(hipMalloc(reinterpret_cast<void**>(&dR), M_ * sizeof(double)), "hipMalloc(probe)");
    check_hip(hipMemsetAsync(dA, 0, N_ * sizeof(double), stream_), "hipMemsetAsync(probe)");
    check_hip(hipMemsetAsync(dR, 0, M_ * sizeof(double), stream_), "hipMemsetAsync(probe)");
    check_hip(hipEventCreate(&ev[0]), "hipEventCreate");
    check_hip(hipEventCreate(&ev[1]), "hipEventCreate");
    // per kernel and order a first launch (a long first one ends that
    // kernel's probe: large problems pay one extra launch), then the timed
    // one; multTranspose runs on y = dR (zero: set above, and mult's output
    // on zero weights) and writes dA
    for (int kern = 0; kern < 2; ++kern) {
      int& cf = kern == 0 ? mult_cf_[ix] : multT_cf_[ix];
      auto timed = [&]() {
        if (kern == 0)
          run_mult<T>(X, LI, TB, dA, dR, ev);
        else
          run_multT<T>(X, LI, TB, dR, dA, ev);
        check_hip(hipEventSynchronize(ev[1]), "hipEventSynchronize");
        float t = 0.0f;
        check_hip(hipEventElapsedTime(&t, ev[0], ev[1]), "hipEventElapsedTime");
        return t;
      };
      float ms[2] = {0.0f, 0.0f};
      for (int o = 0; o < 2; ++o) {
        cf = o;
        ms[o] = timed();
        if (o == 0 && ms[0] > kOrderProbeMs) break;
        ms[o] = timed();
      }
      cf = (ms[0] <= kOrderProbeMs && ms[1] < ms[0]) ? 1 : 0;
    }
  } catch (...) {
    mult_cf_[ix] = multT_cf_[ix] = 0;
    release();
    throw;
  }
  release();
}

void DeviceMultiEval::record_start() { check_hip(hipEventRecord(ev0_, stream_), "hipEventRecord"); }
void DeviceMultiEval::record_stop() { check_hip(hipEventRecord(ev1_, stream_), "hipEventRecord"); }

template <typename T>
void DeviceMultiEval::run_mult(const T* X, const T* LI, const float* TB, const double* d_alpha,
                               double* d_result, hipEvent_t* tev, const MultScratch* ext) {
  hipStream_t s = stream_;
  // weights, then partials (with ext: in the caller's scratch, 256-byte aligned)
  const size_t a_bytes = (Npad_ * sizeof(T) + 255) / 256 * 256;
  void* A = nullptr;
  if (ext) {
    if (a_bytes + Mpad_ * sizeof(double) > ext->bytes) {
      std::ostringstream os;
      os << "input too large for the fixed scratch: " << N_ << " grid points and " << M_
         << " data points need " << (a_bytes + Mpad_ * sizeof(double))
         << " bytes at one grid chunk, the scratch holds " << ext->bytes;
      throw input_error(os.str());
    }
    A = ext->base;
  } else {
    ensure(&A_, &A_cap_, Npad_ * sizeof(T));
    A = A_;
  }
  clear_launch_error();
  hipLaunchKernelGGL(prep_alpha_kernel<T>, dim3(ceil_div(Npad_, 256)), dim3(256), 0, s, d_alpha,
                     scale_, static_cast<const int32_t*>(perm_), static_cast<T*>(A),
                     static_cast<int64_t>(N_), static_cast<int64_t>(Npad_));
  check_hip(hipGetLastError(), "prep_alpha");
  const int K = mult_k_rt<T>(static_cast<int>(D_));
  const int ntiles = ceil_div(N_, kTG);
  const int bx = static_cast<int>(Mpad_ / (kBlock * K));
  int G = cfg_.mult_grid_chunks;
  if (G <= 0) {
    // many short blocks: the dispatcher then balances the uneven per-block
    // work (culling makes blocks cost very different amounts) across all CUs
    const int target = cfg_.target_blocks > 0 ? cfg_.target_blocks : kBlocksPerCU * num_cus_;
    G = std::max(1, ceil_div(target, bx));
    G = std::max(G, ceil_div(ntiles, max_chunk_tiles<T>()));
    G = std::min(G, std::max(1, ntiles / 4));  // at least 4 tiles per chunk
    // partials G x Mpad doubles within the budget
    G = std::min<int64_t>(G, std::max<int64_t>(1, kPartialBudget / (Mpad_ * sizeof(double))));
  }
  if (ext)  // partials G x Mpad doubles within the scratch
    G = std::min<int64_t>(G, (ext->bytes - a_bytes) / (Mpad_ * sizeof(double)));
  G = std::max(1, std::min(G, ntiles));
  const int tiles_per = ceil_div(ntiles, G);
  G = ceil_div(ntiles, tiles_per);
  const int chunk = tiles_per * kTG;
  const int jend = ntiles * kTG;
  double* P;
  if (ext) {
    P = reinterpret_cast<double*>(static_cast<unsigned char*>(ext->base) + a_bytes);
  } else {
    ensure(&part_, &part_cap_, static_cast<size_t>(G) * Mpad_ * sizeof(double));
    P = static_cast<double*>(part_);
  }
  if (tev) check_hip(hipEventRecord(tev[0], s), "hipEventRecord");
  if (D_ <= static_cast<size_t>(kMaxTemplD)) {
    const bool skip = basis_ == Basis::ModLinear;
    MultFn<T> f = count_ ? count_mult_launcher<T>(static_cast<int>(D_), skip)
                         : DimTable<T>::mult(static_cast<int>(D_), skip);
    // grid chunks fastest only where the data blocks fit the grid's y extent
    const int cf = mult_cf_[sizeof(T) == sizeof(double) ? 0 : 1] && bx <= 65535 ? 1 : 0;
    f(X, static_cast<int64_t>(Mpad_), group_boxes<T>(), LI, TB, static_cast<const uint32_t*>(pm_),
      static_cast<const T*>(A), P, chunk, jend, G, static_cast<int64_t>(Mpad_), cf, count_, s);
  } else {
    dim3 grid(static_cast<unsigned>(Mpad_ / kBlock), static_cast<unsigned>(G));
    clear_launch_error();
    hipLaunchKernelGGL(mult_kernel_generic<T>, grid, dim3(kBlock), 0, s, X,
                       static_cast<int64_t>(Mpad_), LI, static_cast<const T*>(A), P, chunk, jend,
                       static_cast<int>(D_));
  }
  check_hip(hipGetLastError(), "mult kernel");
  if (tev) check_hip(hipEventRecord(tev[1], s), "hipEventRecord");
  // results back to the caller's row order
  clear_launch_error();
  hipLaunchKernelGGL(reduce_parts_kernel, dim3(ceil_div(M_, 512)), dim3(256), 0, s, P,
                     static_cast<int64_t>(Mpad_), G, static_cast<const double*>(nullptr),
                     static_cast<const int32_t*>(dperm_), d_result, static_cast<int64_t>(M_));
  check_hip(hipGetLastError(), "reduce (mult)");
  last_G_ = G;
}

template <typename T>
void DeviceMultiEval::run_multT(const T* X, const T* LI, const float* TB, const double* d_y,
                                double* d_result, hipEvent_t* tev) {
  hipStream_t s = stream_;
  ensure(&Y_, &Y_cap_, Mpad_ * sizeof(T));
  clear_launch_error();
  hipLaunchKernelGGL(prep_y_kernel<T>, dim3(ceil_div(Mpad_, 256)), dim3(256), 0, s, d_y,
                     static_cast<const int32_t*>(dperm_), static_cast<T*>(Y_),
                     static_cast<int64_t>(M_), static_cast<int64_t>(Mpad_));
  check_hip(hipGetLastError(), "prep_y");
  const bool templ = D_ <= static_cast<size_t>(kMaxTemplD);
  const int ntiles = ceil_div(N_, kTG);
  const int jend = ntiles * kTG;
  // grid blocks: chunks of kChunkT points (templated kernel) or kBlock points
  const int bx = templ ? ceil_div(jend, kChunkT) : ceil_div(N_, kBlock);
  // Data are split into groups (templated kernel: one wave's 64*K points;
  // generic kernel: kTM points); a segment is 4 waves x gpw groups, so the
  // four waves of a block get equal shares. mend = whole groups past M (the
  // padding holds at least kDataSlack rows past M).
  const int64_t group = templ ? 64 * multT_k_rt<T>(static_cast<int>(D_)) : kTM;
  const int64_t ngroups = (static_cast<int64_t>(M_) + group - 1) / group;
  const int64_t mend = ngroups * group;
  const int64_t waves = templ ? kBlock / 64 : 1;
  int S = cfg_.multT_data_segs;
  if (S <= 0) {
    const int target = cfg_.target_blocks > 0 ? cfg_.target_blocks
                                              : (templ ? kBlocksPerCUT : kBlocksPerCU) * num_cus_;
    S = std::max(1, ceil_div(target, bx));
    S = std::min<int64_t>(S, std::max<int64_t>(1, kPartialBudget / (Npad_ * sizeof(double))));
  }
  const int64_t gpw = std::max<int64_t>(1, (ngroups + waves * S - 1) / (waves * S));
  const int64_t seg = waves * gpw * group;
  S = ceil_div(mend, seg);
  ensure(&part_, &part_cap_, static_cast<size_t>(S) * Npad_ * sizeof(double));
  double* P = static_cast<double*>(part_);
  if (tev) check_hip(hipEventRecord(tev[0], s), "hipEventRecord");
  if (templ) {
    const bool skip = basis_ == Basis::ModLinear;
    MultTFn<T> f = count_ ? count_multT_launcher<T>(static_cast<int>(D_), skip)
                          : DimTable<T>::multT(static_cast<int>(D_), skip);
    // grid chunks fastest only where the segments fit the grid's y extent
    const int cf = multT_cf_[sizeof(T) == sizeof(double) ? 0 : 1] && S <= 65535 ? 1 : 0;
    f(X, static_cast<int64_t>(Mpad_), group_boxes<T>(), static_cast<const T*>(Y_), LI, TB,
      static_cast<const uint32_t*>(pm_), P, static_cast<int64_t>(Npad_), jend, seg, mend, S,
      cf, count_, s);
  } else {
    dim3 grid(static_cast<unsigned>(bx), static_cast<unsigned>(S));
    clear_launch_error();
    hipLaunchKernelGGL(multT_kernel_generic<T>, grid, dim3(kBlock), 0, s, X,
                       static_cast<int64_t>(Mpad_), static_cast<const T*>(Y_), LI, P,
                       static_cast<int64_t>(Npad_), seg, mend, static_cast<int>(D_));
  }
  check_hip(hipGetLastError(), "multTranspose kernel");
  if (tev) check_hip(hipEventRecord(tev[1], s), "hipEventRecord");
  // results back to the caller's grid order, times the per-point scale
  clear_launch_error();
  hipLaunchKernelGGL(reduce_parts_kernel, dim3(ceil_div(N_, 512)), dim3(256), 0, s, P,
                     static_cast<int64_t>(Npad_), S, scale_, static_cast<const int32_t*>(perm_),
                     d_result, static_cast<int64_t>(N_));
  check_hip(hipGetLastError(), "reduce (multTranspose)");
  last_S_ = S;
}

void DeviceMultiEval::mult_device(const double* d_alpha, double* d_result, Precision p) {
  if (N_ == 0) throw input_error("grid not set (call prepare)");
  if (!serves(p)) throw input_error("precision not enabled for this instance");
  check_hip(hipSetDevice(device_), "hipSetDevice");
  EventPair* tp = timing_on_ ? next_timer(tm_mult_, used_mult_) : nullptr;
  hipEvent_t tev[2] = {tp ? tp->a : nullptr, tp ? tp->b : nullptr};
  record_start();
  if (algorithm_ == Algorithm::Subspace)
    run_subspace_mult(d_alpha, d_result, tp ? tev : nullptr);
  else if (p == Precision::FP64)
    run_mult<double>(static_cast<const double*>(X64_), static_cast<const double*>(LI64_),
                     static_cast<const float*>(TB64_), d_alpha, d_result, tp ? tev : nullptr);
  else
    run_mult<float>(static_cast<const float*>(X32_), static_cast<const float*>(LI32_),
                    static_cast<const float*>(TB32_), d_alpha, d_result, tp ? tev : nullptr);
  record_stop();
}

namespace {
// Fixed device scratch of mult_device_noalloc: one copy per GPU, held until
// the process ends. In host code this name is only a stand-in for the device
// symbol; the address to use is the one hipGetSymbolAddress reports for the
// current device (device_scratch).
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

// this device's address of the scratch array (looked up once per device,
// with the device current)
void* device_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) throw input_error("device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[device]) {
    void* ptr = nullptr;
    check_hip(hipSetDevice(device), "hipSetDevice");
    check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!ptr) throw device_error("hipGetSymbolAddress(scratch) returned a null address");
    g_scratch[device] = ptr;
  }
  return g_scratch[device];
}
}  // namespace

void DeviceMultiEval::mult_device_noalloc(const double* d_alpha, double* d_result, Precision p) {
  if (N_ == 0) throw input_error("grid not set (call prepare)");
  if (!serves(p)) throw input_error("precision not enabled for this instance");
  if (algorithm_ != Algorithm::Streaming)
    throw input_error("the no-allocation entry point serves the streaming algorithm only");
  const MultScratch sc{device_scratch(device_), kNoAllocScratchBytes};
  check_hip(hipSetDevice(device_), "hipSetDevice");
  record_start();
  if (p == Precision::FP64)
    run_mult<double>(static_cast<const double*>(X64_), static_cast<const double*>(LI64_),
                     static_cast<const float*>(TB64_), d_alpha, d_result, nullptr, &sc);
  else
    run_mult<float>(static_cast<const float*>(X32_), static_cast<const float*>(LI32_),
                    static_cast<const float*>(TB32_), d_alpha, d_result, nullptr, &sc);
  record_stop();
}

void DeviceMultiEval::mult_transpose_device(const double* d_y, double* d_result, Precision p) {
  if (N_ == 0) throw input_error("grid not set (call prepare)");
  if (!serves(p)) throw input_error("precision not enabled for this instance");
  check_hip(hipSetDevice(device_), "hipSetDevice");
  EventPair* tp = timing_on_ ? next_timer(tm_multT_, used_multT_) : nullptr;
  hipEvent_t tev[2] = {tp ? tp->a : nullptr, tp ? tp->b : nullptr};
  record_start();
  if (algorithm_ == Algorithm::Subspace)
    run_subspace_multT(d_y, d_result, tp ? tev : nullptr);
  else if (p == Precision::FP64)
    run_multT<double>(static_cast<const double*>(X64_), static_cast<const double*>(LI64_),
                      static_cast<const float*>(TB64_), d_y, d_result, tp ? tev : nullptr);
  else
    run_multT<float>(static_cast<const float*>(X32_), static_cast<const float*>(LI32_),
                     static_cast<const float*>(TB32_), d_y, d_result, tp ? tev : nullptr);
  record_stop();
}

void DeviceMultiEval::count_work(Precision p, uint64_t out[4]) {
  if (N_ == 0) throw input_error("grid not set (call prepare)");
  if (!serves(p)) throw input_error("precision not enabled for this instance");
  if (algorithm_ == Algorithm::Subspace) {
    // every data row (padding included) visits every subspace and evaluates
    // the dimensions it does not share with the previous one
    out[1] = out[3] = static_cast<uint64_t>(Mpad_) * static_cast<uint64_t>(nsub_);
    out[0] = static_cast<uint64_t>(Mpad_) * sub_recompute_;
    out[2] = static_cast<uint64_t>(Mpad_) * sub_recompute_t_;
    return;
  }
  check_hip(hipSetDevice(device_), "hipSetDevice");
  const size_t n = std::max(N_, M_) * sizeof(double);
  ensure(&hin_, &hin_cap_, n);
  ensure(&hout_, &hout_cap_, n);
  check_hip(hipMemsetAsync(hin_, 0, n, stream_), "clear count input");
  unsigned long long* dcnt = nullptr;
  check_hip(hipMalloc(reinterpret_cast<void**>(&dcnt), 4 * sizeof(unsigned long long)),
            "hipMalloc(counters)");
  const bool timing = timing_on_;
  timing_on_ = false;  // counting launches are never timed
  try {
    check_hip(hipMemsetAsync(dcnt, 0, 4 * sizeof(unsigned long long), stream_), "clear counters");
    count_ = dcnt;
    mult_device(static_cast<const double*>(hin_), static_cast<double*>(hout_), p);
    mult_transpose_device(static_cast<const double*>(hin_), static_cast<double*>(hout_), p);
    count_ = nullptr;
    unsigned long long h[4];
    check_hip(hipMemcpyAsync(h, dcnt, sizeof(h), hipMemcpyDeviceToHost, stream_), "read counters");
    check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    for (int i = 0; i < 4; ++i) out[i] = h[i];
  } catch (...) {
    count_ = nullptr;
    timing_on_ = timing;
    (void)hipStreamSynchronize(stream_);
    (void)hipFree(dcnt);
    throw;
  }
  timing_on_ = timing;
  check_hip(hipFree(dcnt), "hipFree(counters)");
  if (D_ > static_cast<size_t>(kMaxTemplD)) {
    // the generic kernels evaluate every pair of the padded arrays in every dimension
    const uint64_t jend = static_cast<uint64_t>(ceil_div(N_, kTG)) * kTG;
    out[1] = static_cast<uint64_t>(Mpad_) * jend;
    out[0] = out[1] * D_;
    out[3] = out[1];
    out[2] = out[0];
  }
}

DeviceMultiEval::EventPair* DeviceMultiEval::next_timer(std::vector<EventPair>& pool,
                                                       size_t& used) {
  if (used == pool.size()) {
    EventPair e;
    check_hip(hipEventCreate(&e.a), "hipEventCreate");
    check_hip(hipEventCreate(&e.b), "hipEventCreate");
    pool.push_back(e);
  }
  return &pool[used++];
}

void DeviceMultiEval::set_kernel_timing(bool on) {
  timing_on_ = on;
  used_mult_ = used_multT_ = 0;
}

void DeviceMultiEval::kernel_timing(int* n_mult, double* ms_mult, int* n_multT,
                                    double* ms_multT) {
  check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  auto sum = [](std::vector<EventPair>& pool, size_t used) {
    double t = 0.0;
    for (size_t i = 0; i < used; ++i) {
      float ms = 0.0f;
      check_hip(hipEventElapsedTime(&ms, pool[i].a, pool[i].b), "hipEventElapsedTime");
      t += ms;
    }
    return t;
  };
  if (n_mult) *n_mult = static_cast<int>(used_mult_);
  if (ms_mult) *ms_mult = sum(tm_mult_, used_mult_);
  if (n_multT) *n_multT = static_cast<int>(used_multT_);
  if (ms_multT) *ms_multT = sum(tm_multT_, used_multT_);
}

void DeviceMultiEval::synchronize() {
  check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, ev0_, ev1_) == hipSuccess) last_kernel_ms_ = ms;
}

void DeviceMultiEval::mult_host(const double* alpha, double* result, Precision p) {
  const auto t0 = std::chrono::steady_clock::now();
  ensure(&hin_, &hin_cap_, N_ * sizeof(double));
  ensure(&hout_, &hout_cap_, M_ * sizeof(double));
  check_hip(hipMemcpyAsync(hin_, alpha, N_ * sizeof(double), hipMemcpyHostToDevice, stream_),
            "copy alpha");
  mult_device(static_cast<const double*>(hin_), static_cast<double*>(hout_), p);
  check_hip(hipMemcpyAsync(result, hout_, M_ * sizeof(double), hipMemcpyDeviceToHost, stream_),
            "copy result");
  synchronize();
  last_wall_s_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void DeviceMultiEval::mult_transpose_host(const double* y, double* result, Precision p) {
  const auto t0 = std::chrono::steady_clock::now();
  ensure(&hin_, &hin_cap_, M_ * sizeof(double));
  ensure(&hout_, &hout_cap_, N_ * sizeof(double));
  check_hip(hipMemcpyAsync(hin_, y, M_ * sizeof(double), hipMemcpyHostToDevice, stream_),
            "copy source");
  mult_transpose_device(static_cast<const double*>(hin_), static_cast<double*>(hout_), p);
  check_hip(hipMemcpyAsync(result, hout_, N_ * sizeof(double), hipMemcpyDeviceToHost, stream_),
            "copy result");
  synchronize();
  last_wall_s_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace sgpp_hip
