"""ORACLE -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker or the CPU baseline; the product
(sgpp_amd) never imports it.

Two layers:
  * ``_oracle.so`` (sgpp_oracle.c): plain-C restatement of SG++'s operator
    definition (LinearBasis / LinearModifiedBasis products summed in long
    double) and of the StreamingDefault CPU operators ("port").
  * ``_ref/libsgpp_ref_v{3,4}.so``: the reference's own grid, DEFAULT,
    StreamingDefault, SubspaceLinear and CG code, compiled from its sources by
    oracle/Makefile (plus our thin ref_driver.cpp). Present wherever it was
    built; the v4 (AVX-512) build is used only on hosts whose CPU reports
    the AVX-512 subset it needs.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_dp = ctypes.POINTER(ctypes.c_double)
_up = ctypes.POINTER(ctypes.c_uint32)
_sz = ctypes.c_size_t
_vp = ctypes.c_void_p

BASIS_LINEAR, BASIS_MODLINEAR = 0, 1
KIND_DEFAULT, KIND_STREAMING, KIND_SUBSPACE = 0, 1, 2

_lib = None
_ref = None
_ref_variant = None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _d(a):
    return a.ctypes.data_as(_dp)


def _u(a):
    return a.ctypes.data_as(_up)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "_oracle.so")
        if not os.path.exists(path):
            raise ImportError("oracle/_oracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(path)
        L.oracle_mult.argtypes = [ctypes.c_int, _sz, _sz, _dp, _sz, _up, _up, _dp, _dp]
        L.oracle_mult_transpose.argtypes = [ctypes.c_int, _sz, _sz, _dp, _sz, _up, _up, _dp, _dp]
        L.oracle_phi.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double]
        L.oracle_phi.restype = ctypes.c_double
        L.port_create.argtypes = [ctypes.c_int, _sz, _sz, _dp, _sz, _up, _up]
        L.port_create.restype = _vp
        L.port_destroy.argtypes = [_vp]
        L.port_mult.argtypes = [_vp, _dp, _dp]
        L.port_mult_transpose.argtypes = [_vp, _dp, _dp]
        L.oracle_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def phi(basis, level, index, x):
    return lib().oracle_phi(basis, level, index, x)


def mult(basis, X, level, index, alpha):
    """r = B alpha by the definition (long double sums)."""
    X, lv, ix, a = _f64(X), _u32(level), _u32(index), _f64(alpha)
    M, d = X.shape
    r = np.empty(M)
    lib().oracle_mult(basis, d, M, _d(X), lv.shape[0], _u(lv), _u(ix), _d(a), _d(r))
    return r


def mult_transpose(basis, X, level, index, y):
    """a = B^T y by the definition (long double sums)."""
    X, lv, ix, yy = _f64(X), _u32(level), _u32(index), _f64(y)
    M, d = X.shape
    a = np.empty(lv.shape[0])
    lib().oracle_mult_transpose(basis, d, M, _d(X), lv.shape[0], _u(lv), _u(ix), _d(yy), _d(a))
    return a


class Port:
    """StreamingDefault restatement (FP64, OpenMP), the 'port' CPU baseline."""

    def __init__(self, basis, X, level, index):
        X, lv, ix = _f64(X), _u32(level), _u32(index)
        self.M, self.d = X.shape
        self.N = lv.shape[0]
        self.h = lib().port_create(basis, self.d, self.M, _d(X), self.N, _u(lv), _u(ix))

    def __del__(self):
        if getattr(self, "h", None):
            lib().port_destroy(self.h)
            self.h = None

    def mult(self, alpha):
        a = _f64(alpha)
        r = np.empty(self.M)
        lib().port_mult(self.h, _d(a), _d(r))
        return r

    def mult_transpose(self, y):
        yy = _f64(y)
        a = np.empty(self.N)
        lib().port_mult_transpose(self.h, _d(yy), _d(a))
        return a


def num_threads():
    return lib().oracle_num_threads()


# ------------------------------------------------------------ reference build

def _cpu_has_avx512():
    try:
        with open("/proc/cpuinfo") as f:
            flags = set()
            for line in f:
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    break
        return {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"} <= flags
    except OSError:
        return False


def ref_variant():
    ref()
    return _ref_variant


def ref():
    """The reference build, or None where it was not built."""
    global _ref, _ref_variant
    if _ref is not None:
        return _ref
    variants = (["v4", "v3"] if _cpu_has_avx512() else ["v3"])
    for v in variants:
        path = os.path.join(HERE, "_ref", "libsgpp_ref_%s.so" % v)
        if os.path.exists(path):
            # lazy binding: the reference library has a few unresolved symbols of
            # basis types that are never called on this path
            R = ctypes.CDLL(path, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
            R.ref_last_error.restype = ctypes.c_char_p
            R.ref_grid_new.argtypes = [ctypes.c_int, _sz]
            R.ref_grid_new.restype = _vp
            R.ref_grid_free.argtypes = [_vp]
            R.ref_grid_regular.argtypes = [_vp, _sz]
            R.ref_grid_from_arrays.argtypes = [ctypes.c_int, _sz, _sz, _up, _up]
            R.ref_grid_from_arrays.restype = _vp
            R.ref_grid_size.argtypes = [_vp]
            R.ref_grid_size.restype = _sz
            R.ref_grid_level_index.argtypes = [_vp, _up, _up, ctypes.POINTER(ctypes.c_uint8)]
            R.ref_grid_refine_surplus.argtypes = [_vp, _dp, _sz, _sz, ctypes.c_double]
            R.ref_grid_hierarchise.argtypes = [_vp, _dp, _dp]
            R.ref_op_new.argtypes = [_vp, ctypes.c_int, _dp, _sz]
            R.ref_op_new.restype = _vp
            R.ref_op_free.argtypes = [_vp]
            R.ref_op_prepare.argtypes = [_vp]
            R.ref_op_mult.argtypes = [_vp, _dp, _sz, _dp, _sz]
            R.ref_op_mult_transpose.argtypes = [_vp, _dp, _sz, _dp, _sz]
            R.ref_op_seconds.argtypes = [_vp]
            R.ref_op_seconds.restype = ctypes.c_double
            R.ref_cg_solve.argtypes = [_vp, _sz, _sz, _dp, ctypes.c_double, _sz, ctypes.c_double,
                                       _dp, ctypes.POINTER(_sz), ctypes.POINTER(ctypes.c_double)]
            R.ref_gflop_gbyte.argtypes = [_vp, _sz, ctypes.c_int, _sz, _sz, ctypes.c_int, _dp, _dp]
            R.ref_learner_train.argtypes = [_vp, ctypes.c_int, _dp, _sz, _dp, ctypes.c_int,
                                            ctypes.c_double, ctypes.c_double, _sz,
                                            ctypes.c_double, _sz, _sz, ctypes.c_double, _sz,
                                            ctypes.c_int, _dp, _sz, ctypes.POINTER(_sz),
                                            ctypes.POINTER(_sz), ctypes.POINTER(_sz), _dp]
            _ref, _ref_variant = R, v
            return R
    return None


def _rcheck(rc):
    if rc != 0:
        raise RuntimeError("reference error: " + ref().ref_last_error().decode())


class RefGrid:
    """The reference's own Grid (Linear / ModLinear)."""

    def __init__(self, basis, dim, _handle=None):
        R = ref()
        if R is None:
            raise RuntimeError("reference build not available")
        self.basis, self.dim = basis, dim
        self.h = _handle if _handle is not None else R.ref_grid_new(basis, dim)

    @classmethod
    def from_arrays(cls, basis, level, index):
        """Reference grid holding exactly these points in this order."""
        lv, ix = _u32(level), _u32(index)
        n, dim = lv.shape
        R = ref()
        if R is None:
            raise RuntimeError("reference build not available")
        return cls(basis, dim, _handle=R.ref_grid_from_arrays(basis, dim, n, _u(lv), _u(ix)))

    def __del__(self):
        if getattr(self, "h", None) and ref() is not None:
            ref().ref_grid_free(self.h)
            self.h = None

    @property
    def size(self):
        return ref().ref_grid_size(self.h)

    def regular(self, level):
        _rcheck(ref().ref_grid_regular(self.h, level))
        return self

    def level_index(self, with_leaf=False):
        n = self.size
        lv = np.empty((n, self.dim), dtype=np.uint32)
        ix = np.empty((n, self.dim), dtype=np.uint32)
        lf = np.empty(n, dtype=np.uint8)
        _rcheck(ref().ref_grid_level_index(self.h, _u(lv), _u(ix),
                                           lf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return (lv, ix, lf) if with_leaf else (lv, ix)

    def refine_surplus(self, alpha, num, threshold=0.0):
        a = _f64(alpha)
        _rcheck(ref().ref_grid_refine_surplus(self.h, _d(a), a.size, num, threshold))

    def hierarchise(self, values):
        v = _f64(values)
        out = np.empty_like(v)
        _rcheck(ref().ref_grid_hierarchise(self.h, _d(v), _d(out)))
        return out

    def performance(self, num_instances, iterations, solver=0, size_datatype=8,
                    reuse_alpha=False):
        """the reference's LearnerVectorizedPerformanceCalculator::getGFlopAndGByte"""
        gf, gb = ctypes.c_double(0), ctypes.c_double(0)
        _rcheck(ref().ref_gflop_gbyte(self.h, num_instances, solver, iterations, size_datatype,
                                      int(reuse_alpha), ctypes.byref(gf), ctypes.byref(gb)))
        return gf.value, gb.value


def learner_train(basis, X, y, level, lam, refine=(1e-10, 1000), final=None, num_refinements=1,
                  threshold=0.0, num_points=5, test_acc=False, kind=1):
    """LearnerBase::train's steps on reference parts (ref_learner_train).
    Returns (RefGrid, alpha, [(grid size, CG iterations, train MSE or -1)])."""
    X = _f64(X)
    y = _f64(y)
    M, d = X.shape
    final = refine if final is None else final
    g = RefGrid(basis, d)
    cap = 4 * (regular_size_bound(d, level) + (num_refinements + 1) * num_points * (2 * d + 2) ** 2)
    alpha = np.empty(cap)
    k = num_refinements + 1
    sizes = np.zeros(k, dtype=np.uint64)
    iters = np.zeros(k, dtype=np.uint64)
    mse = np.zeros(k)
    n = _sz(0)
    P = ctypes.POINTER(_sz)
    _rcheck(ref().ref_learner_train(g.h, kind, _d(X), M, _d(y), level, lam, refine[0],
                                    refine[1], final[0], final[1], num_refinements, threshold,
                                    num_points, int(test_acc), _d(alpha), cap, ctypes.byref(n),
                                    sizes.ctypes.data_as(P), iters.ctypes.data_as(P), _d(mse)))
    steps = [(int(sizes[j]), int(iters[j]), float(mse[j])) for j in range(n.value)]
    return g, alpha[:g.size].copy(), steps


def regular_size_bound(d, level):
    """points of the regular no-boundary grid of this level"""
    import math
    return sum(2 ** k * math.comb(d - 1 + k, d - 1) for k in range(level))


class RefOp:
    """A reference OperationMultipleEval (DEFAULT / StreamingDefault / SubspaceLinear)."""

    def __init__(self, grid, kind, X):
        X = _f64(X)
        self.grid = grid
        self.M = X.shape[0]
        self.h = ref().ref_op_new(grid.h, kind, _d(X), self.M)
        if not self.h:
            raise RuntimeError("reference error: " + ref().ref_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None) and ref() is not None:
            ref().ref_op_free(self.h)
            self.h = None

    def prepare(self):
        _rcheck(ref().ref_op_prepare(self.h))

    def mult(self, alpha):
        a = _f64(alpha)
        r = np.empty(self.M)
        _rcheck(ref().ref_op_mult(self.h, _d(a), a.size, _d(r), self.M))
        return r

    def mult_transpose(self, y):
        yy = _f64(y)
        n = self.grid.size
        a = np.empty(n)
        _rcheck(ref().ref_op_mult_transpose(self.h, _d(yy), self.M, _d(a), n))
        return a

    def seconds(self):
        return ref().ref_op_seconds(self.h)

    def cg_solve(self, y, lam, max_it, eps):
        yy = _f64(y)
        n = self.grid.size
        a = np.empty(n)
        it = _sz(0)
        res = ctypes.c_double(0)
        _rcheck(ref().ref_cg_solve(self.h, self.M, n, _d(yy), lam, max_it, eps, _d(a),
                                   ctypes.byref(it), ctypes.byref(res)))
        return a, it.value, res.value


# ----------------------------------------------------- reference learner chain

def learner_check_path():
    """_ref/learner_check (the reference's LearnerLeastSquaresIdentity and
    ModelFittingLeastSquares, compiled from its sources, with the HIP adapter
    behind its factory), or None where it was not built."""
    p = os.path.join(HERE, "_ref", "learner_check")
    return p if os.path.exists(p) else None


def run_learner_check(mode, op, basis, X, y, level, lam, Xt=None, refine=(1e-10, 1000),
                      final=None, num_refinements=1, threshold=0.0, num_points=5,
                      test_acc=False, regularization=0, weights=None, timeout=600):
    """Runs the reference's own learner (mode "learner": LearnerLeastSquaresIdentity::train,
    predict, getAccuracy), fitter (mode "fitter": ModelFittingLeastSquares fit, adapt,
    evaluate), RegressionLearner (mode "regression": train, predict, getMSE, with
    regularization 0 Identity, 1 Laplace, 2 Diagonal) or DMWeightMatrix in CG (mode
    "weighted", data weights `weights`) in a child process, with the operator op:
    "streaming" (StreamingDefault), "default", "hip64" / "hip32" (the HIP operator
    through the SG++-side adapter, selected by the configuration) or "hipdefault" (the
    base factory returns the adapter). Returns {stage: (level, index, alpha,
    predictions)} and the info values."""
    import subprocess
    import tempfile
    exe = learner_check_path()
    if exe is None:
        raise RuntimeError("oracle/_ref/learner_check is not built (make -C oracle ref)")
    X, y = _f64(X), _f64(y)
    M, d = X.shape
    Xt = X[:1] if Xt is None else _f64(Xt)
    final = refine if final is None else final
    with tempfile.TemporaryDirectory() as tmp:
        X.tofile(os.path.join(tmp, "X.bin"))
        y.tofile(os.path.join(tmp, "y.bin"))
        Xt.tofile(os.path.join(tmp, "Xt.bin"))
        if weights is not None:
            _f64(weights).tofile(os.path.join(tmp, "w.bin"))
        args = [exe, mode, op, tmp, "basis=%d" % basis, "dim=%d" % d, "level=%d" % level,
                "M=%d" % M, "Mt=%d" % Xt.shape[0], "lambda=%r" % lam,
                "eps_refine=%r" % refine[0], "it_refine=%d" % refine[1],
                "eps_final=%r" % final[0], "it_final=%d" % final[1],
                "refinements=%d" % num_refinements, "points=%d" % num_points,
                "threshold=%r" % threshold, "test_acc=%d" % int(test_acc),
                "regularization=%d" % regularization]
        r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError("learner_check failed (%d): %s" % (r.returncode, r.stderr.strip()))
        info = {}
        with open(os.path.join(tmp, "info.txt")) as f:
            for line in f:
                k, v = line.split()
                info[k] = float(v)
        out = {}
        for stage in ("train", "fit", "adapt"):
            p = os.path.join(tmp, stage + "_alpha.bin")
            if not os.path.exists(p):
                continue
            lv = np.fromfile(os.path.join(tmp, stage + "_level.bin"), dtype=np.uint32)
            ix = np.fromfile(os.path.join(tmp, stage + "_index.bin"), dtype=np.uint32)
            out[stage] = (lv.reshape(-1, d), ix.reshape(-1, d), np.fromfile(p),
                          np.fromfile(os.path.join(tmp, stage + "_pred.bin")))
    return out, info
