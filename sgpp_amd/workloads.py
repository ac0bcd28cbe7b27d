"""Seeded synthetic workloads of the multi-evaluation path.

Random numbers come from one documented generator, independent of numpy's
and the C++ library's distributions, so every machine produces the same
bits:

  splitmix64 finaliser over a counter:  for stream s and draw k,
      z = s * 0xD1B54A32D192ED03 + (k + 1) * 0x9E3779B97F4A7C15   (mod 2^64)
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9
      z = (z ^ (z >> 27)) * 0x94D049BB133111EB
      z =  z ^ (z >> 31)
  uniform [0,1):  (z >> 11) * 2^-53
  normal:         Box-Muller, sqrt(-2 ln u1) * cos(2 pi u2), u1 in (0,1]

Every named workload derives its streams from its seed (seed * 16 + tag).
"""
import math

import numpy as np

_M1 = np.uint64(0xD1B54A32D192ED03)
_G = np.uint64(0x9E3779B97F4A7C15)
_C1 = np.uint64(0xBF58476D1CE4E5B9)
_C2 = np.uint64(0x94D049BB133111EB)


def _u64(stream, n, offset=0):
    k = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(stream) * _M1 + k * _G
        z = (z ^ (z >> np.uint64(30))) * _C1
        z = (z ^ (z >> np.uint64(27))) * _C2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(stream, n):
    return (_u64(stream, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(stream, n):
    z = _u64(stream, 2 * n)
    u1 = ((z[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)
    u2 = (z[1::2] >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def stream(seed, tag):
    return seed * 16 + tag


def uniform_points(seed, n, d):
    """n points uniform on [0,1)^d, row-major."""
    return uniform(stream(seed, 0), n * d).reshape(n, d)


def gaussian_mixture(seed, n, d, components, sigma, lo=0.2, hi=0.8, chunk=1 << 20):
    """Equal-weight isotropic mixture, means uniform in [lo,hi]^d, clipped to [0,1]."""
    means = lo + (hi - lo) * uniform(stream(seed, 1), components * d).reshape(components, d)
    out = np.empty((n, d), dtype=np.float64)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        c = (_u64(stream(seed, 2), e - s, offset=s) >> np.uint64(11)).astype(np.float64)
        comp = np.minimum((c * (components / 9007199254740992.0)).astype(np.int64), components - 1)
        z = _normal_block(stream(seed, 3), (e - s) * d, offset=s * d).reshape(e - s, d)
        out[s:e] = np.clip(means[comp] + sigma * z, 0.0, 1.0)
    return out


def _normal_block(strm, n, offset):
    z = _u64(strm, 2 * n, offset=2 * offset)
    u1 = ((z[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)
    u2 = (z[1::2] >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def gaussian_bump(c):
    """f(x) = exp(-sum_j (x_j - 0.5)^2 / 0.05)"""
    return np.exp(-((c - 0.5) ** 2).sum(axis=1) / 0.05)


def cosine_product(c):
    """f(x) = prod_j (1 + 0.5 cos(pi x_j))"""
    return np.prod(1.0 + 0.5 * np.cos(np.pi * c), axis=1)


def damped_cosine_product(c):
    """f(x) = prod_j 4 x_j (1 - x_j) (1 + 0.5 cos(pi x_j)); zero on the
    boundary of the cube. (Without the damping factor the boundary values
    cannot be represented by a grid without boundary points, and surplus
    refinement keeps halving the cells next to the faces until it runs out of
    levels.)"""
    return np.prod(4.0 * c * (1.0 - c) * (1.0 + 0.5 * np.cos(np.pi * c)), axis=1)


def adaptive_grid(basis, dim, start_level, f, per_step, target, max_steps=10000, progress=None):
    """Regular grid, then repeat {hierarchise f at the grid points; refine the
    per_step largest |surplus|} until the grid has >= target points.
    progress(steps, size) is called every 10 steps when given.
    Returns (grid, alpha_of_f, steps)."""
    from . import capi
    g = capi.Grid(basis, dim).regular(start_level)
    coords = g.coordinates()
    vals = f(coords)
    alpha = g.hierarchise(vals, incremental=True)
    steps = 0
    while g.size < target and steps < max_steps:
        n0 = g.size
        g.refine_surplus(alpha, per_step)
        if g.size == n0:
            break
        lv, ix = g.level_index()
        new = ix[n0:] / (2.0 ** lv[n0:])
        vals = np.concatenate([vals, f(new)])
        alpha = g.hierarchise(vals, incremental=True)
        steps += 1
        if progress is not None and steps % 10 == 0:
            progress(steps, g.size)
    return g, alpha, steps


# ------------------------------------------------------------- named configs

def config_small_linear():
    """Linear d=5 level 5 regular grid (N = 1471); 10,000 points U[0,1]^5 (seed 0),
    alpha ~ N(0,1) (seed 1), y ~ N(0,1) (seed 2)."""
    from . import capi
    g = capi.Grid(capi.BASIS_LINEAR, 5).regular(5)
    X = uniform_points(0, 10000, 5)
    alpha = normal(stream(1, 0), g.size)
    y = normal(stream(2, 0), X.shape[0])
    return dict(name="linear_d5_l5_regular_m10k", basis=capi.BASIS_LINEAR, grid=g, X=X,
                alpha=alpha, y=y)


def config_modlinear_adaptive(target=100000, M=1000000):
    """ModLinear d=10: regular level 4, +200 refinements per step of the
    hierarchised Gaussian bump until >= target points; M points from a
    5-component mixture (sigma 0.08, means in [0.2,0.8]^10, clipped, seed 3).
    alpha ~ N(0,1) (seed 1), y ~ N(0,1) (seed 2)."""
    from . import capi
    g, falpha, steps = adaptive_grid(capi.BASIS_MODLINEAR, 10, 4, gaussian_bump, 200, target)
    X = gaussian_mixture(3, M, 10, 5, 0.08)
    alpha = normal(stream(1, 0), g.size)
    y = normal(stream(2, 0), M)
    return dict(name="modlinear_d10_adaptive_n%d_m%d" % (g.size, M),
                basis=capi.BASIS_MODLINEAR, grid=g, X=X, alpha=alpha, y=y, steps=steps)


def config_regression(M=500000, level=6, dim=8):
    """Linear d=8 level 6 (N = 31,745); M points U[0,1]^8 (seed 4);
    y = sum_j sin(2 pi x_j) x_{(j+1) mod 8} + N(0, 0.01^2); lambda = 1e-4."""
    from . import capi
    g = capi.Grid(capi.BASIS_LINEAR, dim).regular(level)
    X = uniform_points(4, M, dim)
    y = np.zeros(M)
    for j in range(dim):
        y += np.sin(2.0 * np.pi * X[:, j]) * X[:, (j + 1) % dim]
    y += 0.01 * normal(stream(4, 5), M)
    return dict(name="linear_d%d_l%d_regression_m%d" % (dim, level, M),
                basis=capi.BASIS_LINEAR, grid=g, X=X, y=y, lam=1e-4)


def config_large_streaming(target=1000000, M=10000000, progress=None):
    """Linear d=20: regular level 4, +200 refinements per step of the
    hierarchised damped cosine product until >= target points; M points from
    an 8-component mixture (sigma 0.05, means in [0.2,0.8]^20, clipped, seed 6),
    stored as FP32 values (the FP64 path evaluates the same values);
    y ~ N(0,1) (seed 6, tag 7); lambda 1e-5."""
    from . import capi
    g, falpha, steps = adaptive_grid(capi.BASIS_LINEAR, 20, 4, damped_cosine_product, 200,
                                     target, progress=progress)
    X = gaussian_mixture(6, M, 20, 8, 0.05).astype(np.float32).astype(np.float64)
    y = normal(stream(6, 7), M)
    return dict(name="linear_d20_adaptive_n%d_m%d" % (g.size, M), basis=capi.BASIS_LINEAR,
                grid=g, X=X, y=y, lam=1e-5, steps=steps)


def regular_size(d, level):
    """number of points of the regular no-boundary sparse grid"""
    return sum(2 ** k * math.comb(d - 1 + k, d - 1) for k in range(level))


def config_serving(n_models=64, rate=2000.0, duration=4.0, qmax=100000, nmin=1000, nmax=500000,
                   seed=5):
    """64 resident models (d uniform in 2..20; regular or adaptive Linear /
    ModLinear grids with N log-uniform in [nmin, nmax]; alpha ~ N(0,1)) and a
    Poisson request stream at `rate` req/s for `duration` seconds, each request
    for a uniform model with q log-uniform in [1, qmax] points uniform on
    [0,1]^d. Query points are slices of one uniform pool per dimension."""
    from . import capi
    u = uniform(stream(seed, 0), n_models * 5).reshape(n_models, 5)
    models = []
    for m in range(n_models):
        d = 2 + min(18, int(u[m, 0] * 19))
        basis = capi.BASIS_LINEAR if u[m, 1] < 0.5 else capi.BASIS_MODLINEAR
        adaptive = u[m, 2] >= 0.5
        target = math.exp(math.log(nmin) + u[m, 3] * (math.log(nmax) - math.log(nmin)))
        levels = [L for L in range(1, 31) if regular_size(d, L) <= nmax]
        if adaptive:
            # regular start, then surplus refinement driven by a random indicator,
            # with steps small enough not to overshoot the target by more than 2d points
            L = max([L for L in levels if regular_size(d, L) <= target / 2] or [1])
            g = capi.Grid(basis, d).regular(L)
            step = 0
            while g.size < target:
                n0 = g.size
                num = max(1, min(n0 // 5, int(math.ceil((target - n0) / (2 * d)))))
                g.refine_surplus(1.0 + uniform(stream(seed * 1000 + m, 16 + step), n0), num)
                step += 1
                if g.size == n0:
                    break
        else:
            inside = [L for L in levels if regular_size(d, L) >= nmin] or levels[-1:]
            L = min(inside, key=lambda L: abs(math.log(regular_size(d, L)) - math.log(target)))
            g = capi.Grid(basis, d).regular(L)
        lv, ix = g.level_index()
        alpha = normal(stream(seed * 1000 + m, 2), g.size)
        models.append(dict(basis=basis, dim=d, level=lv, index=ix, alpha=alpha,
                           kind="adaptive" if adaptive else "regular"))
    n_req = int(rate * duration * 1.5) + 16
    v = uniform(stream(seed, 3), 4 * n_req).reshape(n_req, 4)
    t = np.cumsum(-np.log(1.0 - v[:, 0]) / rate)
    keep = t < duration
    t = t[keep]
    model = np.minimum((v[keep, 1] * n_models).astype(np.int64), n_models - 1)
    q = np.maximum(1, np.minimum(qmax, np.floor(np.exp(v[keep, 2] * math.log(qmax))))).astype(np.int64)
    pool_rows = qmax + 4096
    dims = sorted({mm["dim"] for mm in models})
    pools = {d: uniform(stream(seed, 8 + d), pool_rows * d).reshape(pool_rows, d) for d in dims}
    offs = (v[keep, 3] * (pool_rows - q)).astype(np.int64)
    return dict(models=models, arrival=t, model=model, q=q, offset=offs, pools=pools,
                rate=rate, duration=duration)
