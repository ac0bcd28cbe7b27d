"""ctypes binding of libsgpp_hip.so (declarations: include/sgpp_hip.h).

Python mirror of the C++ operator interface for tests and benchmarks:
``Grid`` (createLinearGrid / createModLinearGrid, regular, refine,
hierarchise), ``MultiEval`` (OperationMultipleEval: mult, multTranspose,
prepare, getDuration, the regression solve), ``Server`` (resident models,
continuous batching) and ``Learner`` (LearnerLeastSquaresIdentity: train,
predict, getAccuracy). There is no fallback: if the
library is missing, importing this module raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsgpp_hip.so")

OK = 0
STATUS_NAMES = {0: "OK", 1: "INVALID_ARGUMENT", 2: "FACTORY", 3: "OPERATION", 4: "NOT_IMPLEMENTED",
                5: "GENERATION", 6: "DEVICE", 7: "BUDGET", 8: "NOT_FOUND", 9: "CANCELLED",
                10: "TIMEOUT", 11: "SOLVER", 99: "INTERNAL"}
BASIS_LINEAR, BASIS_MODLINEAR = 0, 1
PREC_FP64, PREC_FP32 = 0, 1
ALGO_STREAMING, ALGO_SUBSPACE = 0, 1
SERVE_AUTO, SERVE_STREAMING, SERVE_SUBSPACE = 0, 1, 2
SOLVE_FP64, SOLVE_MIXED, SOLVE_FP32 = 0, 1, 2
SLE_CG, SLE_BICGSTAB = 0, 1


class SGppHIPError(RuntimeError):
    def __init__(self, status, message):
        self.status = status
        self.status_name = STATUS_NAMES.get(status, str(status))
        super().__init__("%s: %s" % (self.status_name, message))


if not os.path.exists(LIB_PATH):
    raise ImportError("libsgpp_hip.so is not built (run __graft_entry__.build() or "
                      "python -m sgpp_amd.build); there is no CPU fallback")

# One HIP runtime per process: PyTorch ships its own libamdhip64 (loaded by
# the unversioned name), the library needs libamdhip64.so.7. Loading torch
# first makes the library bind to torch's copy instead of loading a second
# runtime, which would then fail to find the device.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - torch is optional for the C-ABI
    pass

lib = ctypes.CDLL(LIB_PATH)

_p = ctypes.c_void_p
_sz = ctypes.c_size_t
_st = ctypes.c_int
_dp = ctypes.POINTER(ctypes.c_double)
_up = ctypes.POINTER(ctypes.c_uint32)


class GridDesc(ctypes.Structure):
    _fields_ = [("basis", ctypes.c_int), ("dim", ctypes.c_size_t)]


class Params(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("keep_fp64", ctypes.c_int), ("device", ctypes.c_int),
                ("mult_grid_chunks", ctypes.c_int), ("multT_data_segs", ctypes.c_int),
                ("target_blocks", ctypes.c_int), ("algorithm", ctypes.c_int)]


class SolveParams(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("lambda_", ctypes.c_double),
                ("max_iterations", ctypes.c_size_t), ("epsilon", ctypes.c_double),
                ("inner_epsilon", ctypes.c_double), ("max_refinements", ctypes.c_size_t),
                ("reuse", ctypes.c_int)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_size_t), ("refinements", ctypes.c_size_t),
                ("fp64_matvecs", ctypes.c_size_t), ("fp32_matvecs", ctypes.c_size_t),
                ("residual_rel", ctypes.c_double), ("seconds", ctypes.c_double),
                ("converged", ctypes.c_int)]


class GridConfig(ctypes.Structure):
    _fields_ = [("basis", ctypes.c_int), ("dim", ctypes.c_size_t), ("level", ctypes.c_int)]


class SolverConfig(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("eps", ctypes.c_double),
                ("max_iterations", ctypes.c_size_t), ("threshold", ctypes.c_double)]


class AdaptivityConfig(ctypes.Structure):
    _fields_ = [("num_refinements", ctypes.c_size_t), ("refinement_threshold", ctypes.c_double),
                ("num_refinement_points", ctypes.c_size_t),
                ("error_based_refinement", ctypes.c_int)]


class LearnerTiming(ctypes.Structure):
    _fields_ = [("time_complete", ctypes.c_double), ("time_mult_complete", ctypes.c_double),
                ("time_mult_compute", ctypes.c_double),
                ("time_mult_trans_complete", ctypes.c_double),
                ("time_mult_trans_compute", ctypes.c_double),
                ("time_regularization", ctypes.c_double), ("gflop", ctypes.c_double),
                ("gbyte", ctypes.c_double)]


class FitterConfig(ctypes.Structure):
    _fields_ = [("grid", GridConfig), ("refinement", AdaptivityConfig),
                ("solver_refine", SolverConfig), ("solver_final", SolverConfig),
                ("lambda_", ctypes.c_double)]


class LearnerStep(ctypes.Structure):
    _fields_ = [("grid_size", ctypes.c_size_t), ("iterations", ctypes.c_size_t),
                ("refinements", ctypes.c_size_t), ("residual", ctypes.c_double),
                ("train_accuracy", ctypes.c_double), ("seconds", ctypes.c_double)]


def _decl(name, restype, argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = argtypes
    return f


# every symbol declared in include/sgpp_hip.h, with its ctypes signature
SIGNATURES = {
    "sgpp_hip_last_error": (ctypes.c_char_p, []),
    "sgpp_hip_version": (ctypes.c_int, []),
    "sgpp_hip_grid_create": (_st, [ctypes.c_int, _sz, ctypes.POINTER(_p)]),
    "sgpp_hip_grid_destroy": (_st, [_p]),
    "sgpp_hip_grid_regular": (_st, [_p, _sz]),
    "sgpp_hip_grid_size": (_sz, [_p]),
    "sgpp_hip_grid_dim": (_sz, [_p]),
    "sgpp_hip_grid_get_level_index": (_st, [_p, _up, _up]),
    "sgpp_hip_grid_get_coordinates": (_st, [_p, _dp]),
    "sgpp_hip_grid_get_leaf": (_st, [_p, ctypes.POINTER(ctypes.c_uint8)]),
    "sgpp_hip_grid_refine_surplus": (_st, [_p, _dp, _sz, _sz, ctypes.c_double,
                                           ctypes.POINTER(_sz)]),
    "sgpp_hip_grid_hierarchise": (_st, [_p, _dp, _dp]),
    "sgpp_hip_grid_hierarchise_incremental": (_st, [_p, _dp, _dp]),
    "sgpp_hip_grid_serialize": (_st, [_p, ctypes.c_char_p, _sz, ctypes.POINTER(_sz)]),
    "sgpp_hip_grid_unserialize": (_st, [ctypes.c_char_p, ctypes.POINTER(_p)]),
    "sgpp_hip_meval_create": (_st, [ctypes.POINTER(GridDesc), _dp, _sz, _sz,
                                    ctypes.POINTER(Params), ctypes.POINTER(_p)]),
    "sgpp_hip_meval_create_for_grid": (_st, [_p, _dp, _sz, ctypes.POINTER(Params),
                                             ctypes.POINTER(_p)]),
    "sgpp_hip_meval_prepare": (_st, [_p, _up, _up, _sz]),
    "sgpp_hip_meval_prepare_grid": (_st, [_p]),
    "sgpp_hip_meval_mult": (_st, [_p, _dp, _dp]),
    "sgpp_hip_meval_mult_transpose": (_st, [_p, _dp, _dp]),
    "sgpp_hip_meval_mult_range": (_st, [_p, _dp, ctypes.c_size_t, ctypes.c_size_t, _dp]),
    "sgpp_hip_meval_mult_transpose_range": (_st, [_p, _dp, ctypes.c_size_t, ctypes.c_size_t,
                                                  _dp]),
    "sgpp_hip_meval_mult_device": (_st, [_p, _p, _p, ctypes.c_int]),
    "sgpp_hip_meval_mult_transpose_device": (_st, [_p, _p, _p, ctypes.c_int]),
    "sgpp_hip_meval_synchronize": (_st, [_p]),
    "sgpp_hip_meval_duration": (ctypes.c_double, [_p]),
    "sgpp_hip_meval_kernel_ms": (ctypes.c_double, [_p]),
    "sgpp_hip_meval_stream": (_p, [_p]),
    "sgpp_hip_meval_num_data": (_sz, [_p]),
    "sgpp_hip_meval_num_grid": (_sz, [_p]),
    "sgpp_hip_meval_last_shape": (_st, [_p, ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_int)]),
    "sgpp_hip_meval_kernel_timing_enable": (_st, [_p, ctypes.c_int]),
    "sgpp_hip_meval_mult_device_noalloc": (_st, [_p, _p, _p, ctypes.c_int]),
    "sgpp_hip_meval_count_work": (_st, [_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]),
    "sgpp_hip_meval_kernel_timing": (_st, [_p, ctypes.POINTER(ctypes.c_int), _dp,
                                           ctypes.POINTER(ctypes.c_int), _dp]),
    "sgpp_hip_meval_destroy": (_st, [_p]),
    "sgpp_hip_regression_solve": (_st, [_p, _dp, ctypes.POINTER(SolveParams), _dp,
                                        ctypes.POINTER(SolveInfo)]),
    "sgpp_hip_system_matrix_mult": (_st, [_p, ctypes.c_double, _dp, _dp, ctypes.c_int]),
    "sgpp_hip_server_create": (_st, [ctypes.c_void_p, ctypes.POINTER(_p)]),
    "sgpp_hip_server_destroy": (_st, [_p]),
    "sgpp_hip_server_admit_model": (_st, [_p, ctypes.c_int, _sz, _sz, _up, _up, _dp,
                                          ctypes.POINTER(ctypes.c_int)]),
    "sgpp_hip_server_evict_model": (_st, [_p, ctypes.c_int]),
    "sgpp_hip_server_memory_used": (_st, [_p, ctypes.POINTER(_sz)]),
    "sgpp_hip_server_submit": (_st, [_p, ctypes.c_int, _sz, _dp, ctypes.c_double,
                                     ctypes.POINTER(ctypes.c_uint64)]),
    "sgpp_hip_server_poll": (_st, [_p, ctypes.c_uint64, _dp, _sz, ctypes.c_double,
                                   ctypes.POINTER(ctypes.c_int), _dp]),
    "sgpp_hip_server_cancel": (_st, [_p, ctypes.c_uint64]),
    "sgpp_hip_server_inject_error": (_st, [_p]),
    "sgpp_hip_server_drain": (_st, [_p]),
    "sgpp_hip_server_stats": (_st, [_p, ctypes.c_void_p]),
    "sgpp_hip_server_tuned_shape": (_st, [_p, ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_int), ctypes.c_void_p]),
    "sgpp_hip_server_model_algorithm": (_st, [_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "sgpp_hip_learner_create": (_st, [ctypes.c_int, ctypes.POINTER(Params), ctypes.POINTER(_p)]),
    "sgpp_hip_learner_destroy": (_st, [_p]),
    "sgpp_hip_learner_train": (_st, [_p, _dp, _sz, _sz, _dp, ctypes.POINTER(GridConfig),
                                     ctypes.POINTER(SolverConfig), ctypes.POINTER(SolverConfig),
                                     ctypes.POINTER(AdaptivityConfig), ctypes.c_int,
                                     ctypes.c_double, ctypes.POINTER(LearnerTiming)]),
    "sgpp_hip_learner_predict": (_st, [_p, _dp, _sz, _sz, _dp]),
    "sgpp_hip_learner_accuracy": (_st, [_p, _dp, _sz, _sz, _dp, ctypes.c_double, _dp]),
    "sgpp_hip_learner_grid_size": (_st, [_p, ctypes.POINTER(_sz)]),
    "sgpp_hip_learner_level_index": (_st, [_p, _up, _up]),
    "sgpp_hip_learner_alpha": (_st, [_p, _dp]),
    "sgpp_hip_learner_num_steps": (_st, [_p, ctypes.POINTER(_sz)]),
    "sgpp_hip_learner_step_info": (_st, [_p, _sz, ctypes.POINTER(LearnerStep)]),
    "sgpp_hip_learner_performance": (_st, [_p, _sz, ctypes.c_int, _sz, _sz, ctypes.c_int, _dp,
                                           _dp]),
    "sgpp_hip_fitter_create": (_st, [ctypes.POINTER(FitterConfig), ctypes.POINTER(Params),
                                     ctypes.POINTER(_p)]),
    "sgpp_hip_fitter_destroy": (_st, [_p]),
    "sgpp_hip_fitter_fit": (_st, [_p, _dp, _sz, _sz, _dp]),
    "sgpp_hip_fitter_adapt": (_st, [_p, ctypes.POINTER(ctypes.c_int)]),
    "sgpp_hip_fitter_update": (_st, [_p, _dp, _sz, _sz, _dp]),
    "sgpp_hip_fitter_evaluate": (_st, [_p, _dp, _sz, _sz, _dp]),
    "sgpp_hip_fitter_grid_size": (_st, [_p, ctypes.POINTER(_sz)]),
    "sgpp_hip_fitter_level_index": (_st, [_p, _up, _up]),
    "sgpp_hip_fitter_alpha": (_st, [_p, _dp]),
    "sgpp_hip_fitter_info": (_st, [_p, ctypes.POINTER(_sz), ctypes.POINTER(_sz)]),
}

F = {}
for _name, (_res, _args) in SIGNATURES.items():
    F[_name] = _decl(_name, _res, _args)


def last_error():
    return F["sgpp_hip_last_error"]().decode(errors="replace")


def check(status):
    if status != OK:
        raise SGppHIPError(status, last_error())


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Grid:
    """Linear / ModLinear no-boundary sparse grid with SG++ sequence numbers."""

    def __init__(self, basis, dim, _handle=None):
        self.basis = basis
        if _handle is not None:
            self.h = _handle
        else:
            h = _p()
            check(F["sgpp_hip_grid_create"](basis, dim, ctypes.byref(h)))
            self.h = h

    @classmethod
    def linear(cls, dim):
        return cls(BASIS_LINEAR, dim)

    @classmethod
    def modlinear(cls, dim):
        return cls(BASIS_MODLINEAR, dim)

    def __del__(self):
        h = getattr(self, "h", None)
        if h and F is not None:
            F["sgpp_hip_grid_destroy"](h)
            self.h = None

    @property
    def size(self):
        return F["sgpp_hip_grid_size"](self.h)

    @property
    def dim(self):
        return F["sgpp_hip_grid_dim"](self.h)

    def regular(self, level):
        check(F["sgpp_hip_grid_regular"](self.h, level))
        return self

    def level_index(self):
        n, d = self.size, self.dim
        lv = np.empty((n, d), dtype=np.uint32)
        ix = np.empty((n, d), dtype=np.uint32)
        check(F["sgpp_hip_grid_get_level_index"](self.h, lv.ctypes.data_as(_up),
                                                 ix.ctypes.data_as(_up)))
        return lv, ix

    def coordinates(self):
        c = np.empty((self.size, self.dim), dtype=np.float64)
        check(F["sgpp_hip_grid_get_coordinates"](self.h, _dptr(c)))
        return c

    def leaf(self):
        a = np.empty(self.size, dtype=np.uint8)
        check(F["sgpp_hip_grid_get_leaf"](self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return a

    def refine_surplus(self, alpha, refinements_num, threshold=0.0):
        a = _f64(alpha)
        n = _sz(0)
        check(F["sgpp_hip_grid_refine_surplus"](self.h, _dptr(a), a.size, refinements_num,
                                                threshold, ctypes.byref(n)))
        return n.value

    def hierarchise(self, node_values, incremental=False):
        v = _f64(node_values)
        out = np.empty_like(v)
        fn = "sgpp_hip_grid_hierarchise_incremental" if incremental else "sgpp_hip_grid_hierarchise"
        check(F[fn](self.h, _dptr(v), _dptr(out)))
        return out

    def serialize(self):
        n = _sz(0)
        check(F["sgpp_hip_grid_serialize"](self.h, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        check(F["sgpp_hip_grid_serialize"](self.h, buf, n.value, ctypes.byref(n)))
        return buf.value.decode()

    @classmethod
    def unserialize(cls, text):
        h = _p()
        check(F["sgpp_hip_grid_unserialize"](text.encode(), ctypes.byref(h)))
        basis = BASIS_LINEAR if text.lstrip().startswith("linear") else BASIS_MODLINEAR
        return cls(basis, 0, _handle=h)


    def performance(self, num_instances, iterations, solver=SLE_CG, size_datatype=8,
                    reuse_alpha=False):
        """(GFlop, GByte) of one CG run under the reference learners' convention
        (LearnerVectorizedPerformanceCalculator::getGFlopAndGByte)"""
        gf, gb = ctypes.c_double(0), ctypes.c_double(0)
        check(F["sgpp_hip_learner_performance"](self.h, num_instances, solver, iterations,
                                                size_datatype, int(reuse_alpha),
                                                ctypes.byref(gf), ctypes.byref(gb)))
        return gf.value, gb.value


class MultiEval:
    """OperationMultipleEval on the GPU (mult = B alpha, multTranspose = B^T y).

    algorithm=ALGO_STREAMING: the culled streaming kernels (FP64 or FP32 with
    FP64 accumulation); ALGO_SUBSPACE: the subspace-indexed kernels
    (SubspaceLinear semantics, FP64, d <= 20)."""

    def __init__(self, basis, data, precision=PREC_FP64, keep_fp64=False, device=0,
                 mult_grid_chunks=0, multT_data_segs=0, target_blocks=0, grid=None,
                 algorithm=ALGO_STREAMING):
        data = _f64(data)
        if data.ndim != 2:
            raise ValueError("data must be M x d")
        self.M, self.d = data.shape
        self.precision = precision
        prm = Params(precision, int(keep_fp64), device, mult_grid_chunks, multT_data_segs,
                     target_blocks, algorithm)
        h = _p()
        if grid is not None:
            self.grid = grid  # keep alive
            check(F["sgpp_hip_meval_create_for_grid"](grid.h, _dptr(data), self.M,
                                                      ctypes.byref(prm), ctypes.byref(h)))
        else:
            desc = GridDesc(basis, self.d)
            check(F["sgpp_hip_meval_create"](ctypes.byref(desc), _dptr(data), self.M, self.d,
                                             ctypes.byref(prm), ctypes.byref(h)))
        self.h = h
        self.N = 0

    def __del__(self):
        h = getattr(self, "h", None)
        if h and F is not None:
            F["sgpp_hip_meval_destroy"](h)
            self.h = None

    def prepare(self, level=None, index=None):
        if level is None:
            check(F["sgpp_hip_meval_prepare_grid"](self.h))
        else:
            lv = np.ascontiguousarray(level, dtype=np.uint32)
            ix = np.ascontiguousarray(index, dtype=np.uint32)
            if lv.shape != ix.shape or lv.ndim != 2 or lv.shape[1] != self.d:
                raise ValueError("level/index must both be N x d")
            check(F["sgpp_hip_meval_prepare"](self.h, lv.ctypes.data_as(_up),
                                              ix.ctypes.data_as(_up), lv.shape[0]))
        self.N = F["sgpp_hip_meval_num_grid"](self.h)
        return self

    def mult(self, alpha):
        a = _f64(alpha)
        if a.size != self.N:
            raise ValueError("alpha must have N entries")
        r = np.empty(self.M, dtype=np.float64)
        check(F["sgpp_hip_meval_mult"](self.h, _dptr(a), _dptr(r)))
        return r

    def mult_transpose(self, source):
        y = _f64(source)
        if y.size != self.M:
            raise ValueError("source must have M entries")
        a = np.empty(self.N, dtype=np.float64)
        check(F["sgpp_hip_meval_mult_transpose"](self.h, _dptr(y), _dptr(a)))
        return a

    def mult_range(self, alpha, start, end):
        """rows [start, end) of B alpha"""
        a = _f64(alpha)
        if a.size != self.N:
            raise ValueError("alpha must have N entries")
        r = np.empty(max(0, end - start), dtype=np.float64)
        check(F["sgpp_hip_meval_mult_range"](self.h, _dptr(a), start, end, _dptr(r)))
        return r

    def mult_transpose_range(self, source, start, end):
        """entries [start, end) of B^T source"""
        y = _f64(source)
        if y.size != self.M:
            raise ValueError("source must have M entries")
        r = np.empty(max(0, end - start), dtype=np.float64)
        check(F["sgpp_hip_meval_mult_transpose_range"](self.h, _dptr(y), start, end, _dptr(r)))
        return r

    def mult_device(self, d_alpha_ptr, d_result_ptr, precision=None):
        p = self.precision if precision is None else precision
        check(F["sgpp_hip_meval_mult_device"](self.h, d_alpha_ptr, d_result_ptr, p))

    def mult_transpose_device(self, d_source_ptr, d_result_ptr, precision=None):
        p = self.precision if precision is None else precision
        check(F["sgpp_hip_meval_mult_transpose_device"](self.h, d_source_ptr, d_result_ptr, p))

    def synchronize(self):
        check(F["sgpp_hip_meval_synchronize"](self.h))

    def duration(self):
        return F["sgpp_hip_meval_duration"](self.h)

    def kernel_ms(self):
        return F["sgpp_hip_meval_kernel_ms"](self.h)

    def stream(self):
        return F["sgpp_hip_meval_stream"](self.h)

    def last_shape(self):
        g, s = ctypes.c_int(0), ctypes.c_int(0)
        check(F["sgpp_hip_meval_last_shape"](self.h, ctypes.byref(g), ctypes.byref(s)))
        return g.value, s.value

    def kernel_timing_enable(self, on=True):
        check(F["sgpp_hip_meval_kernel_timing_enable"](self.h, int(on)))

    def kernel_timing(self):
        """(n_mult, ms_mult, n_multT, ms_multT) of the main kernels since enabling."""
        a, c = ctypes.c_int(0), ctypes.c_int(0)
        b, d = ctypes.c_double(0), ctypes.c_double(0)
        check(F["sgpp_hip_meval_kernel_timing"](self.h, ctypes.byref(a), ctypes.byref(b),
                                                ctypes.byref(c), ctypes.byref(d)))
        return a.value, b.value, c.value, d.value

    def count_work(self, precision):
        """Executed work of one mult and one multTranspose (counting kernels):
        {"mult_hat_steps", "mult_pairs", "multT_hat_steps", "multT_pairs"}."""
        out = (ctypes.c_uint64 * 4)()
        check(F["sgpp_hip_meval_count_work"](self.h, precision, out))
        return {"mult_hat_steps": out[0], "mult_pairs": out[1], "multT_hat_steps": out[2],
                "multT_pairs": out[3]}

    def system_mult(self, v, lam, precision=PREC_FP64):
        v = _f64(v)
        q = np.empty_like(v)
        check(F["sgpp_hip_system_matrix_mult"](self.h, lam, _dptr(v), _dptr(q), precision))
        return q

    def solve(self, y, lam, mode=SOLVE_FP64, max_iterations=1000, epsilon=1e-10,
              inner_epsilon=1e-5, max_refinements=20, alpha0=None):
        y = _f64(y)
        alpha = np.zeros(self.N) if alpha0 is None else _f64(alpha0).copy()
        prm = SolveParams(mode, lam, max_iterations, epsilon, inner_epsilon, max_refinements,
                          int(alpha0 is not None))
        info = SolveInfo()
        check(F["sgpp_hip_regression_solve"](self.h, _dptr(y), ctypes.byref(prm), _dptr(alpha),
                                             ctypes.byref(info)))
        return alpha, {k: getattr(info, k) for k, _ in SolveInfo._fields_}


# ------------------------------------------------------------- served loop

REQ_QUEUED, REQ_RUNNING, REQ_DONE = 0, 1, 2
REQ_FAILED_BUDGET, REQ_FAILED_TIMEOUT, REQ_CANCELLED, REQ_FAILED_DEVICE = 3, 4, 5, 6
REQ_NAMES = {0: "QUEUED", 1: "RUNNING", 2: "DONE", 3: "FAILED_BUDGET", 4: "FAILED_TIMEOUT",
             5: "CANCELLED", 6: "FAILED_DEVICE"}


class ServerConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("precision", ctypes.c_int),
                ("memory_budget", ctypes.c_size_t), ("max_batch_queries", ctypes.c_size_t),
                ("max_items", ctypes.c_size_t), ("max_partial_doubles", ctypes.c_size_t),
                ("max_inflight", ctypes.c_int), ("max_requests_per_batch", ctypes.c_int),
                ("blocks", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("item_work", ctypes.c_double), ("autotune", ctypes.c_int),
                ("autotune_cache", ctypes.c_char_p), ("algorithm", ctypes.c_int)]


class ServerShape(ctypes.Structure):
    _fields_ = [("chunk", ctypes.c_int), ("cost", ctypes.c_double)]


class ServerStatistics(ctypes.Structure):
    _fields_ = [("requests_done", ctypes.c_uint64), ("requests_failed", ctypes.c_uint64),
                ("queries_done", ctypes.c_uint64), ("batches", ctypes.c_uint64),
                ("kernel_launches", ctypes.c_uint64), ("busy_seconds", ctypes.c_double)]


class Server:
    """Resident models + continuous batching (sgpp_hip_server_*).

    algorithm: SERVE_AUTO (per model the faster of the streaming and the
    subspace kernel when autotune is on, else streaming), SERVE_STREAMING or
    SERVE_SUBSPACE (wherever the model has a subspace layout)."""

    def __init__(self, precision=PREC_FP32, memory_budget=0, max_batch_queries=0, max_items=0,
                 max_partial_doubles=0, max_inflight=2, max_requests_per_batch=0, blocks=0,
                 use_graph=True, item_work=0.0, device=0, autotune=False, autotune_cache=None,
                 algorithm=0):
        self._cache = autotune_cache.encode() if autotune_cache else None
        cfg = ServerConfig(device, precision, memory_budget, max_batch_queries, max_items,
                           max_partial_doubles, max_inflight, max_requests_per_batch, blocks,
                           int(use_graph), item_work, int(autotune), self._cache, algorithm)
        h = _p()
        check(F["sgpp_hip_server_create"](ctypes.cast(ctypes.byref(cfg), ctypes.c_void_p),
                                          ctypes.byref(h)))
        self.h = h
        self.dims = {}
        self._q = {}  # query count of each request not yet released

    def close(self):
        h = getattr(self, "h", None)
        if h and F is not None:
            F["sgpp_hip_server_destroy"](h)
            self.h = None

    def __del__(self):
        self.close()

    def admit_model(self, basis, level, index, alpha):
        lv = np.ascontiguousarray(level, dtype=np.uint32)
        ix = np.ascontiguousarray(index, dtype=np.uint32)
        a = _f64(alpha)
        mid = ctypes.c_int(-1)
        check(F["sgpp_hip_server_admit_model"](self.h, basis, lv.shape[1], lv.shape[0],
                                               lv.ctypes.data_as(_up), ix.ctypes.data_as(_up),
                                               _dptr(a), ctypes.byref(mid)))
        self.dims[mid.value] = lv.shape[1]
        return mid.value

    def evict_model(self, model_id):
        check(F["sgpp_hip_server_evict_model"](self.h, model_id))
        self.dims.pop(model_id, None)

    def memory_used(self):
        n = _sz(0)
        check(F["sgpp_hip_server_memory_used"](self.h, ctypes.byref(n)))
        return n.value

    def submit(self, model_id, queries, deadline_s=0.0):
        """queries: q x d array, d = the model's dimension; returns (request id, q)"""
        X = _f64(queries)
        d = self.dims.get(model_id)
        if X.ndim != 2 or (d is not None and X.shape[1] != d):
            raise ValueError("queries must be a q x %s array for model %d, got shape %s"
                             % (d, model_id, X.shape))
        rid = ctypes.c_uint64(0)
        check(F["sgpp_hip_server_submit"](self.h, model_id, X.shape[0], _dptr(X), deadline_s,
                                          ctypes.byref(rid)))
        self._q[rid.value] = X.shape[0]
        return rid.value, X.shape[0]

    def poll(self, rid, q=None, wait_s=-1.0):
        """(status, results or None, latency seconds); a finished request is
        released. The result buffer is sized from the request's own query
        count (q, if given, must equal it)."""
        n = self._q.get(rid)
        if n is None:
            raise SGppHIPError(8, "unknown request id %d" % rid)
        if q is not None and q != n:
            raise ValueError("request %d has %d queries, not %d" % (rid, n, q))
        out = np.empty(n)
        st = ctypes.c_int(-1)
        lat = ctypes.c_double(0)
        check(F["sgpp_hip_server_poll"](self.h, rid, _dptr(out), n, wait_s, ctypes.byref(st),
                                        ctypes.byref(lat)))
        if st.value not in (REQ_QUEUED, REQ_RUNNING):
            self._q.pop(rid, None)
        if st.value == REQ_DONE:
            return st.value, out, lat.value
        return st.value, None, lat.value

    def cancel(self, rid):
        return F["sgpp_hip_server_cancel"](self.h, rid) == OK

    def inject_error(self):
        check(F["sgpp_hip_server_inject_error"](self.h))

    def drain(self):
        check(F["sgpp_hip_server_drain"](self.h))

    def stats(self):
        s = ServerStatistics()
        check(F["sgpp_hip_server_stats"](self.h, ctypes.cast(ctypes.byref(s), ctypes.c_void_p)))
        return {k: getattr(s, k) for k, _ in ServerStatistics._fields_}

    def model_algorithm(self, model_id):
        """SERVE_STREAMING or SERVE_SUBSPACE: the kernel that serves the model."""
        a = ctypes.c_int(0)
        check(F["sgpp_hip_server_model_algorithm"](self.h, model_id, ctypes.byref(a)))
        return a.value

    def tuned_shape(self, basis, dim):
        """{"chunk", "cost"} of the tuned shape of (basis, dim), or None."""
        found = ctypes.c_int(0)
        sh = ServerShape()
        check(F["sgpp_hip_server_tuned_shape"](self.h, basis, dim, ctypes.byref(found),
                                               ctypes.cast(ctypes.byref(sh), ctypes.c_void_p)))
        if not found.value:
            return None
        return {"chunk": sh.chunk, "cost": sh.cost}


# ------------------------------------------------------------- learner

class Learner:
    """LearnerLeastSquaresIdentity on the GPU operator (sgpp_hip_learner_*)."""

    def __init__(self, is_regression=True, precision=PREC_FP64, device=0):
        prm = Params(precision, 0, device, 0, 0, 0)
        h = _p()
        check(F["sgpp_hip_learner_create"](int(is_regression), ctypes.byref(prm),
                                           ctypes.byref(h)))
        self.h = h
        self.d = 0

    def __del__(self):
        h = getattr(self, "h", None)
        if h and F is not None:
            F["sgpp_hip_learner_destroy"](h)
            self.h = None

    def close(self):
        self.__del__()

    def train(self, X, y, basis, level, lam, refine=(1e-10, 1000), final=None,
              num_refinements=1, refinement_threshold=0.0, num_refinement_points=5,
              test_acc_during_adapt=False, error_based_refinement=False, solver=SLE_CG):
        """refine / final: (eps, max_iterations) of the CG runs during and
        after refinement; returns the LearnerTiming fields as a dict"""
        X = _f64(X)
        y = _f64(y)
        M, d = X.shape
        if y.shape != (M,):
            # LearnerBase::train: "length of classes vector does not match to dataset!"
            raise ValueError("y must have one value per data point")
        final = refine if final is None else final
        gc = GridConfig(basis, d, level)
        sr = SolverConfig(solver, refine[0], refine[1], 0.0)
        sf = SolverConfig(solver, final[0], final[1], 0.0)
        ac = AdaptivityConfig(num_refinements, refinement_threshold, num_refinement_points,
                              int(error_based_refinement))
        t = LearnerTiming()
        check(F["sgpp_hip_learner_train"](self.h, _dptr(X), M, d, _dptr(y), ctypes.byref(gc),
                                          ctypes.byref(sr), ctypes.byref(sf), ctypes.byref(ac),
                                          int(test_acc_during_adapt), lam, ctypes.byref(t)))
        self.d = d
        return {k: getattr(t, k) for k, _ in LearnerTiming._fields_}

    def predict(self, X):
        X = _f64(X)
        out = np.empty(X.shape[0])
        check(F["sgpp_hip_learner_predict"](self.h, _dptr(X), X.shape[0], X.shape[1], _dptr(out)))
        return out

    def accuracy(self, X, y, threshold=0.0):
        X = _f64(X)
        y = _f64(y)
        a = ctypes.c_double(0)
        check(F["sgpp_hip_learner_accuracy"](self.h, _dptr(X), X.shape[0], X.shape[1], _dptr(y),
                                             threshold, ctypes.byref(a)))
        return a.value

    @property
    def grid_size(self):
        n = _sz(0)
        check(F["sgpp_hip_learner_grid_size"](self.h, ctypes.byref(n)))
        return n.value

    def level_index(self):
        n = self.grid_size
        lv = np.empty((n, self.d), dtype=np.uint32)
        ix = np.empty((n, self.d), dtype=np.uint32)
        check(F["sgpp_hip_learner_level_index"](self.h, lv.ctypes.data_as(_up),
                                                ix.ctypes.data_as(_up)))
        return lv, ix

    def alpha(self):
        a = np.empty(self.grid_size)
        check(F["sgpp_hip_learner_alpha"](self.h, _dptr(a)))
        return a

    def steps(self):
        n = _sz(0)
        check(F["sgpp_hip_learner_num_steps"](self.h, ctypes.byref(n)))
        out = []
        for k in range(n.value):
            st = LearnerStep()
            check(F["sgpp_hip_learner_step_info"](self.h, k, ctypes.byref(st)))
            out.append({f: getattr(st, f) for f, _ in LearnerStep._fields_})
        return out


class Fitter:
    """ModelFittingLeastSquares on the GPU operator (sgpp_hip_fitter_*).

    Defaults follow FitterConfiguration::setupDefaults: one refinement of one
    point, CG to eps 1e-12 in at most 100 iterations, lambda 0.01."""

    def __init__(self, basis=BASIS_LINEAR, level=3, num_refinements=1, num_points=1,
                 threshold=0.0, refine=(1e-12, 100), final=(1e-12, 100), lam=0.01,
                 precision=PREC_FP64, device=0):
        cfg = FitterConfig(GridConfig(basis, 0, level),
                           AdaptivityConfig(num_refinements, threshold, num_points, 0),
                           SolverConfig(0, refine[0], refine[1], 0.0),
                           SolverConfig(0, final[0], final[1], 0.0), lam)
        prm = Params(precision, 0, device, 0, 0, 0)
        h = _p()
        check(F["sgpp_hip_fitter_create"](ctypes.byref(cfg), ctypes.byref(prm), ctypes.byref(h)))
        self.h = h
        self.d = 0

    def __del__(self):
        h = getattr(self, "h", None)
        if h and F is not None:
            F["sgpp_hip_fitter_destroy"](h)
            self.h = None

    def fit(self, X, y):
        X, y = _f64(X), _f64(y)
        check(F["sgpp_hip_fitter_fit"](self.h, _dptr(X), X.shape[0], X.shape[1], _dptr(y)))
        self.d = X.shape[1]
        return self

    def adapt(self):
        r = ctypes.c_int(0)
        check(F["sgpp_hip_fitter_adapt"](self.h, ctypes.byref(r)))
        return bool(r.value)

    def update(self, X, y):
        X, y = _f64(X), _f64(y)
        check(F["sgpp_hip_fitter_update"](self.h, _dptr(X), X.shape[0], X.shape[1], _dptr(y)))
        self.d = X.shape[1]
        return self

    def evaluate(self, X):
        X = _f64(X)
        out = np.empty(X.shape[0])
        check(F["sgpp_hip_fitter_evaluate"](self.h, _dptr(X), X.shape[0], X.shape[1], _dptr(out)))
        return out

    @property
    def grid_size(self):
        n = _sz(0)
        check(F["sgpp_hip_fitter_grid_size"](self.h, ctypes.byref(n)))
        return n.value

    def level_index(self):
        n = self.grid_size
        lv = np.empty((n, self.d), dtype=np.uint32)
        ix = np.empty((n, self.d), dtype=np.uint32)
        check(F["sgpp_hip_fitter_level_index"](self.h, lv.ctypes.data_as(_up),
                                               ix.ctypes.data_as(_up)))
        return lv, ix

    def alpha(self):
        a = np.empty(self.grid_size)
        check(F["sgpp_hip_fitter_alpha"](self.h, _dptr(a)))
        return a

    def info(self):
        it, ref = _sz(0), _sz(0)
        check(F["sgpp_hip_fitter_info"](self.h, ctypes.byref(it), ctypes.byref(ref)))
        return {"last_iterations": it.value, "refinements": ref.value}
