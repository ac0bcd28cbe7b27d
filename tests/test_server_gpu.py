"""Served loop: results against the oracle, identical alone and in a mix,
and per-request failure handling (budget, deadline, cancel, injected device
error) that never stops the loop.

Query points are drawn in FP64 and submitted as they are; the FP32 server
rounds them itself. Tolerances: FP32-with-FP64-accumulation server 1e-5
relative, FP64 server 1e-12 relative (||got - ref||_2 / ||ref||_2 per
request). Every behaviour is checked for both evaluation algorithms of the
server (streaming and subspace lookups); the subspace kernel computes in FP64
from the queries' FP64 values (hi + lo parts on an FP32 server)."""
import time

import numpy as np
import pytest

import oracle
from sgpp_amd import workloads

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def small_models(capi):
    specs = []
    g = capi.Grid.linear(2).regular(5)
    specs.append((capi.BASIS_LINEAR, g))
    g = capi.Grid.modlinear(5).regular(3)
    g.refine_surplus(np.linspace(1.0, 2.0, g.size), 20)
    specs.append((capi.BASIS_MODLINEAR, g))
    g = capi.Grid.linear(10).regular(3)
    specs.append((capi.BASIS_LINEAR, g))
    g = capi.Grid.modlinear(20).regular(2)
    specs.append((capi.BASIS_MODLINEAR, g))
    out = []
    for i, (basis, g) in enumerate(specs):
        lv, ix = g.level_index()
        alpha = workloads.normal(workloads.stream(50 + i, 0), g.size)
        out.append((basis, lv, ix, alpha))
    return out


def small_server(capi, precision, **kw):
    kw.setdefault("max_batch_queries", 1 << 18)
    kw.setdefault("max_items", 1 << 14)
    kw.setdefault("max_partial_doubles", 1 << 22)
    kw.setdefault("memory_budget", 4 << 30)
    return capi.Server(precision=precision, **kw)


ALGOS = [1, 2]  # SERVE_STREAMING, SERVE_SUBSPACE


@pytest.mark.parametrize("algo", ALGOS)
@pytest.mark.parametrize("prec,tol", [(1, 1e-5), (0, 1e-12)])
def test_results_match_oracle(gpu, prec, tol, algo):
    capi = gpu
    srv = small_server(capi, prec, algorithm=algo)
    models = small_models(capi)
    ids = [srv.admit_model(b, lv, ix, a) for b, lv, ix, a in models]
    assert all(srv.model_algorithm(m) == algo for m in ids)
    if algo == capi.SERVE_SUBSPACE:
        tol = 1e-12  # FP64 lookups on either server
    reqs = []
    for k, q in enumerate([1, 7, 300, 1025, 5000]):
        for m, (b, lv, ix, a) in zip(ids, models):
            X = workloads.uniform_points(60 + k, q, lv.shape[1])
            rid, n = srv.submit(m, X)
            reqs.append((rid, n, X, b, lv, ix, a))
    for rid, n, X, b, lv, ix, a in reqs:
        st, r, lat = srv.poll(rid, n)
        assert st == capi.REQ_DONE, capi.REQ_NAMES[st]
        assert rel(r, oracle.mult(b, X, lv, ix, a)) <= tol
        assert lat > 0
    s = srv.stats()
    assert s["requests_done"] == len(reqs) and s["requests_failed"] == 0
    assert s["batches"] >= 1
    srv.close()


@pytest.mark.parametrize("algo", ALGOS)
def test_alone_and_in_a_mix_are_identical(gpu, algo):
    capi = gpu
    srv = small_server(capi, 1, algorithm=algo)
    models = small_models(capi)
    ids = [srv.admit_model(b, lv, ix, a) for b, lv, ix, a in models]
    probes = [(ids[j], workloads.uniform_points(70 + j, q, models[j][1].shape[1]))
              for j, q in ((1, 2049), (2, 333), (3, 5))]
    alone = []
    for m, X in probes:
        rid, n = srv.submit(m, X)
        st, r, _ = srv.poll(rid, n)
        assert st == capi.REQ_DONE
        alone.append(r)
    # the same requests interleaved with traffic for every model, several times
    for rep in range(3):
        rids = []
        for k in range(12):
            m = ids[k % len(ids)]
            rids.append(srv.submit(m, workloads.uniform_points(80 + k + 13 * rep, 100 + 37 * k,
                                                               models[k % len(ids)][1].shape[1])))
            if k % 4 == 1:
                rids.append(("probe", srv.submit(*probes[(k // 4) % 3]), (k // 4) % 3))
        for item in rids:
            if item[0] == "probe":
                (rid, n), j = item[1], item[2]
                st, r, _ = srv.poll(rid, n)
                assert st == capi.REQ_DONE
                assert np.array_equal(r, alone[j]), "result differs in a mix"
            else:
                rid, n = item
                assert srv.poll(rid, n)[0] == capi.REQ_DONE
    srv.close()


@pytest.mark.parametrize("algo", ALGOS)
@pytest.mark.parametrize("prec,tol", [(1, 1e-5), (0, 1e-12)])
def test_large_requests_are_split(gpu, prec, tol, algo):
    """Requests larger than a batch's buffers (queries, work items, partials)
    run as pieces over several batches, interleaved with other traffic; every
    result is bit-identical to the same request on a server whose batches
    hold it whole, and within tolerance of the oracle."""
    capi = gpu
    models = small_models(capi)
    whole = small_server(capi, prec, algorithm=algo)
    split = small_server(capi, prec, max_batch_queries=1 << 12, max_items=64,
                         max_partial_doubles=1 << 16, algorithm=algo)
    ids_w = [whole.admit_model(b, lv, ix, a) for b, lv, ix, a in models]
    ids_s = [split.admit_model(b, lv, ix, a) for b, lv, ix, a in models]
    reqs = []
    for k, q in enumerate([20000, 9000]):
        for j, (b, lv, ix, a) in enumerate(models):
            X = workloads.uniform_points(110 + 4 * k + j, q, lv.shape[1])
            small = split.submit(ids_s[j], X[:q // 50])  # traffic between the pieces
            reqs.append((split.submit(ids_s[j], X), whole.submit(ids_w[j], X), small, X, j))
    n_split = 0
    for (rs, n), (rw, _), small, X, j in reqs:
        st, got, _ = split.poll(rs, n)
        assert st == capi.REQ_DONE, capi.REQ_NAMES[st]
        st, ref, _ = whole.poll(rw, n)
        assert st == capi.REQ_DONE
        assert np.array_equal(got, ref), "a split request differs from the whole one"
        b, lv, ix, a = models[j]
        assert rel(got, oracle.mult(b, X, lv, ix, a)) <= tol
        assert split.poll(*small)[0] == capi.REQ_DONE
        n_split += n > (1 << 12)
    s = split.stats()
    assert s["requests_failed"] == 0 and s["requests_done"] == 2 * len(reqs)
    assert s["batches"] >= n_split * 3  # every large request took several batches
    split.close()
    whole.close()


def test_autotune_shapes_and_cache(gpu, tmp_path):
    """Tuned shapes per (basis, dimension): valid chunks and costs, results
    unchanged within tolerance, written to the cache file and read back by a
    new server without tuning again."""
    capi = gpu
    cache = str(tmp_path / "shapes.txt")
    srv = small_server(capi, 1, autotune=True, autotune_cache=cache)
    models = small_models(capi)
    ids = [srv.admit_model(b, lv, ix, a) for b, lv, ix, a in models]
    # with the tuner the faster kernel serves each model: a regular Linear
    # d=4 level-9 grid (N = 65,537 points in 495 subspaces) is far cheaper by
    # subspace lookups
    big = capi.Grid.linear(4).regular(9)
    blv, bix = big.level_index()
    balpha = workloads.normal(workloads.stream(99, 0), big.size)
    bid = srv.admit_model(capi.BASIS_LINEAR, blv, bix, balpha)
    assert srv.model_algorithm(bid) == capi.SERVE_SUBSPACE
    Xb = workloads.uniform_points(98, 3000, 4)
    rid, n = srv.submit(bid, Xb)
    st, r, _ = srv.poll(rid, n)
    assert st == capi.REQ_DONE
    assert rel(r, oracle.mult(0, Xb, blv, bix, balpha)) <= 1e-12
    assert all(srv.model_algorithm(m) in (capi.SERVE_STREAMING, capi.SERVE_SUBSPACE)
               for m in ids)
    shapes = {}
    for b, lv, ix, a in models:
        sh = srv.tuned_shape(b, lv.shape[1])
        assert sh is not None
        assert sh["chunk"] >= 64 and sh["chunk"] % 64 == 0 and sh["cost"] > 0
        shapes[(b, lv.shape[1])] = sh
    for k, q in enumerate([3, 700, 4000]):
        for m, (b, lv, ix, a) in zip(ids, models):
            X = workloads.uniform_points(90 + k, q, lv.shape[1])
            rid, n = srv.submit(m, X)
            st, r, _ = srv.poll(rid, n)
            assert st == capi.REQ_DONE
            assert rel(r, oracle.mult(b, X, lv, ix, a)) <= 1e-5
    srv.close()
    with open(cache) as f:
        assert f.read().count("sgpp_hip_server_shape") == len(models) + 1
    srv2 = small_server(capi, 1, autotune=True, autotune_cache=cache)
    for (b, d), sh in shapes.items():
        got = srv2.tuned_shape(b, d)
        assert got["chunk"] == sh["chunk"]
        assert abs(got["cost"] - sh["cost"]) <= 1e-6 * sh["cost"]
    srv2.close()


def test_failures_are_per_request(gpu):
    capi = gpu
    srv = small_server(capi, 1, max_inflight=1, max_requests_per_batch=1,
                       max_batch_queries=1 << 16, max_partial_doubles=1 << 20)
    models = small_models(capi)
    ids = [srv.admit_model(b, lv, ix, a) for b, lv, ix, a in models]
    d0 = models[0][1].shape[1]
    # a model over the memory budget is refused
    big = capi.Grid.linear(12).regular(5)
    lv, ix = big.level_index()
    tiny = small_server(capi, 1, memory_budget=1 << 27, max_batch_queries=1 << 12,
                        max_items=1 << 10, max_partial_doubles=1 << 16)
    with pytest.raises(capi.SGppHIPError) as e:
        tiny.admit_model(capi.BASIS_LINEAR, np.repeat(lv, 64, axis=0), np.repeat(ix, 64, axis=0),
                         np.ones(lv.shape[0] * 64))
    assert e.value.status_name == "BUDGET"
    tiny.close()
    # a request larger than a batch is split and completes
    rid, n = srv.submit(ids[0], workloads.uniform_points(90, 70000, d0))
    assert srv.poll(rid, n)[0] == capi.REQ_DONE
    # a model whose single query tile needs more work items than a batch
    # holds (64-point chunks, one item per batch) fails alone, over budget
    one = small_server(capi, 1, max_items=1, item_work=1.0)
    m0 = one.admit_model(*models[0])
    rid, n = one.submit(m0, workloads.uniform_points(90, 10, d0))
    assert one.poll(rid, n)[0] == capi.REQ_FAILED_BUDGET
    one.close()
    # a deadline that has passed before scheduling
    heavy = srv.submit(ids[2], workloads.uniform_points(91, 60000, models[2][1].shape[1]))
    late = srv.submit(ids[0], workloads.uniform_points(92, 10, d0), deadline_s=1e-7)
    assert srv.poll(*late)[0] == capi.REQ_FAILED_TIMEOUT
    # cancel a request queued behind the heavy one
    queued = srv.submit(ids[0], workloads.uniform_points(93, 10, d0))
    if srv.cancel(queued[0]):
        assert srv.poll(*queued)[0] == capi.REQ_CANCELLED
    else:
        assert srv.poll(*queued)[0] == capi.REQ_DONE
    assert srv.poll(*heavy)[0] == capi.REQ_DONE
    # an injected device error fails one batch; the loop keeps serving
    srv.inject_error()
    bad = srv.submit(ids[1], workloads.uniform_points(94, 50, models[1][1].shape[1]))
    assert srv.poll(*bad)[0] == capi.REQ_FAILED_DEVICE
    good = srv.submit(ids[1], workloads.uniform_points(95, 50, models[1][1].shape[1]))
    st, r, _ = srv.poll(*good)
    assert st == capi.REQ_DONE and np.all(np.isfinite(r))
    # eviction frees the model; unknown ids and out-of-cube points are rejected
    used = srv.memory_used()
    srv.evict_model(ids[3])
    assert srv.memory_used() < used
    with pytest.raises(capi.SGppHIPError):
        srv.submit(ids[3], workloads.uniform_points(96, 5, 20))
    with pytest.raises(capi.SGppHIPError):
        srv.submit(ids[0], np.full((3, d0), 1.5))
    srv.close()


@pytest.mark.parametrize("prec,tol", [(1, 1e-5), (0, 1e-12)])
def test_shared_subtile_dimensions(gpu, prec, tol):
    """Adaptive models whose sorted points share some dimensions within a
    sub-tile and not others, at dimensions that give every sub-tile size
    (FP32: 4 points at d <= 4, 8 at d <= 10, 16 above; FP64: 4 and 8)."""
    capi = gpu
    srv = small_server(capi, prec)
    for i, (basis, d, level) in enumerate([(0, 3, 4), (1, 3, 4), (0, 8, 3), (1, 8, 3),
                                            (0, 14, 2), (1, 14, 3)]):
        g = capi.Grid(basis, d).regular(level)
        g.refine_surplus(np.linspace(1.0, 3.0, g.size), 7)
        lv, ix = g.level_index()
        alpha = workloads.normal(workloads.stream(120 + i, 0), g.size)
        mid = srv.admit_model(basis, lv, ix, alpha)
        X = workloads.uniform_points(130 + i, 3000, d)
        rid, n = srv.submit(mid, X)
        st, r, _ = srv.poll(rid, n)
        assert st == capi.REQ_DONE
        assert rel(r, oracle.mult(basis, X, lv, ix, alpha)) <= tol, (basis, d)
    srv.close()


def test_fp64_queries_at_high_levels(gpu):
    """Models whose levels make FP32 rounding of the queries matter (up to
    2^(l-25) per factor: level 12 gives 1.2e-4). FP64-drawn queries through
    the FP32 server stay within 1e-5 of the oracle on the same FP64 points:
    the server adds each coordinate's low part where L >= 64."""
    capi = gpu
    srv = small_server(capi, 1)
    specs = [(capi.BASIS_LINEAR, 2, 12, 0), (capi.BASIS_MODLINEAR, 3, 10, 0),
             (capi.BASIS_LINEAR, 4, 8, 12), (capi.BASIS_MODLINEAR, 6, 6, 40),
             (capi.BASIS_LINEAR, 12, 4, 60)]
    worst = 0.0
    for i, (basis, d, level, refine) in enumerate(specs):
        g = capi.Grid(basis, d).regular(level)
        for step in range(3 if refine else 0):
            g.refine_surplus(1.0 + workloads.uniform(workloads.stream(140 + i, step), g.size),
                             refine)
        lv, ix = g.level_index()
        assert lv.max() >= 5
        alpha = workloads.normal(workloads.stream(150 + i, 0), g.size)
        mid = srv.admit_model(basis, lv, ix, alpha)
        for k, q in enumerate((1, 333, 2500)):
            X = workloads.uniform_points(160 + 4 * i + k, q, d)
            rid, n = srv.submit(mid, X)
            st, r, _ = srv.poll(rid, n)
            assert st == capi.REQ_DONE
            err = rel(r, oracle.mult(basis, X, lv, ix, alpha))
            worst = max(worst, err)
            assert err <= 1e-5, (basis, d, level, q, err)
    print("worst request error of the FP32 server on FP64 queries: %.2e" % worst)
    srv.close()


def test_served_64_model_poisson_stream(gpu):
    """The served workload: 64 models (d 2..20, Linear/ModLinear, regular or
    adaptive, 1e3..5e5 points) and 0.5 s of its Poisson request stream at
    2000 req/s (q 1..1e5, FP64-drawn queries), on an FP32 and an FP64 server.

    Tolerance. FP32 evaluation of a sum r = sum_j a_j phi_j(x) has a forward
    error bounded by c * d * 2^-24 * sum_j |a_j| phi_j(x) (each term carries
    the rounding of its d factors), so for every request
        ||r32 - r64|| <= 1e-5 * ||B |alpha| ||      (B|alpha| from the FP64 server),
    and, for requests whose sum does not cancel (kappa = ||B|alpha||| / ||r64||
    <= 10), the plain relative error ||r32 - r64|| / ||r64|| <= 1e-5. Requests
    of one or a few queries whose terms cancel (kappa up to ~10^3) can exceed
    1e-5 in the plain relative error at the same forward error; the worst
    are printed. Sampled requests are checked against the oracle (first 64
    queries of every 32nd request: FP64 1e-12, FP32 by the same two rules)."""
    capi = gpu
    c = workloads.config_serving(duration=0.5)
    dims = [m["dim"] for m in c["models"]]
    servers = {1: capi.Server(precision=1), 0: capi.Server(precision=0)}
    ids = {p: [s.admit_model(m["basis"], m["level"], m["index"], m["alpha"])
               for m in c["models"]] for p, s in servers.items()}
    ids_abs = [servers[0].admit_model(m["basis"], m["level"], m["index"], np.abs(m["alpha"]))
               for m in c["models"]]
    reqs = []
    for k in range(len(c["q"])):
        m, q, off = int(c["model"][k]), int(c["q"][k]), int(c["offset"][k])
        X = c["pools"][dims[m]][off:off + q]
        rid = {p: s.submit(ids[p][m], X) for p, s in servers.items()}
        rid["abs"] = servers[0].submit(ids_abs[m], X)
        reqs.append((k, m, X, rid))
    worst_fwd = worst_plain = 0.0
    above = []
    checked = 0
    for k, m, X, rid in reqs:
        out = {}
        for p, s in servers.items():
            st, r, _ = s.poll(*rid[p])
            assert st == capi.REQ_DONE, (k, p, capi.REQ_NAMES[st])
            out[p] = r
        st, rabs, _ = servers[0].poll(*rid["abs"])
        assert st == capi.REQ_DONE
        dn = np.linalg.norm(out[1] - out[0])
        fwd = dn / np.linalg.norm(rabs)
        plain = dn / max(np.linalg.norm(out[0]), 1e-300)
        kappa = np.linalg.norm(rabs) / max(np.linalg.norm(out[0]), 1e-300)
        worst_fwd = max(worst_fwd, fwd)
        worst_plain = max(worst_plain, plain)
        assert fwd <= 1e-5, (k, m, X.shape, fwd)
        if kappa <= 10.0:
            assert plain <= 1e-5, (k, m, X.shape, plain, kappa)
        if plain > 1e-5:
            above.append((k, X.shape[0], plain, kappa))
        if k % 32 == 0:
            mm = c["models"][m]
            rows = X[:64]
            ref = oracle.mult(mm["basis"], rows, mm["level"], mm["index"], mm["alpha"])
            refa = oracle.mult(mm["basis"], rows, mm["level"], mm["index"], np.abs(mm["alpha"]))
            assert rel(out[0][:64], ref) <= 1e-12, (k, m)
            assert np.linalg.norm(out[1][:64] - ref) <= 1e-5 * np.linalg.norm(refa), (k, m)
            checked += 1
    print("%d requests: worst forward error %.2e of ||B|alpha|||, worst plain relative error "
          "%.2e; above 1e-5 plain (request, q, error, kappa): %s; %d checked against the oracle"
          % (len(reqs), worst_fwd, worst_plain, above, checked))
    assert len(reqs) > 800 and checked >= 25
    for s in servers.values():
        assert s.stats()["requests_failed"] == 0
        s.close()


def test_cancelled_split_request_blocks_eviction_until_its_pieces_retire(gpu):
    """A request split over many batches is cancelled while pieces of it may
    sit in formed or running batches. Eviction of its model is refused until
    those pieces have retired (the kernels would read the model's freed
    arrays otherwise); after a drain it succeeds and the server keeps
    serving other models correctly."""
    capi = gpu
    models = small_models(capi)
    srv = small_server(capi, 1, max_batch_queries=1 << 12, max_items=64,
                       max_partial_doubles=1 << 16)
    ids = [srv.admit_model(b, lv, ix, a) for b, lv, ix, a in models]
    d2 = models[2][1].shape[1]
    refused = 0
    for rep in range(4):
        victim = srv.admit_model(*models[2])
        rid, n = srv.submit(victim, workloads.uniform_points(170 + rep, 60000, d2))
        time.sleep(0.0005 * rep)
        cancelled = srv.cancel(rid)
        try:
            srv.evict_model(victim)
            evicted = True
        except capi.SGppHIPError as e:
            assert "queued or running" in str(e)
            evicted = False
            refused += 1
        st = srv.poll(rid, n)[0]
        assert st == (capi.REQ_CANCELLED if cancelled else capi.REQ_DONE)
        srv.drain()
        if not evicted:
            srv.evict_model(victim)
    b, lv, ix, a = models[1]
    X = workloads.uniform_points(180, 500, lv.shape[1])
    st, r, _ = srv.poll(*srv.submit(ids[1], X))
    assert st == capi.REQ_DONE and rel(r, oracle.mult(b, X, lv, ix, a)) <= 1e-5
    print("eviction refused %d of 4 times" % refused)
    srv.close()


def test_poll_refuses_a_short_output_buffer(gpu):
    """sgpp_hip_server_poll copies a finished request's results only into a
    buffer that holds all of them; a short one is refused and the request
    stays available. submit refuses query arrays of the wrong width."""
    import ctypes

    capi = gpu
    models = small_models(capi)
    srv = small_server(capi, 1)
    b, lv, ix, a = models[1]
    mid = srv.admit_model(b, lv, ix, a)
    X = workloads.uniform_points(190, 100, lv.shape[1])
    rid, n = srv.submit(mid, X)
    srv.drain()
    out = np.zeros(n - 1)
    st = ctypes.c_int(-1)
    lat = ctypes.c_double(0)
    code = capi.F["sgpp_hip_server_poll"](srv.h, rid, out.ctypes.data_as(capi._dp), n - 1, -1.0,
                                          ctypes.byref(st), ctypes.byref(lat))
    assert code == 1 and np.all(out == 0.0)
    st, r, _ = srv.poll(rid, n)
    assert st == capi.REQ_DONE and rel(r, oracle.mult(b, X, lv, ix, a)) <= 1e-5
    with pytest.raises(ValueError):
        srv.submit(mid, X[:, :-1])
    with pytest.raises(ValueError):
        srv.submit(mid, X[0])
    srv.close()


def test_served_workload_on_the_subspace_kernel(gpu):
    """The 64 models of the served workload on a server that evaluates every
    model by subspace lookups (FP64) against a streaming FP64 server, over
    the first 0.1 s of the request stream: every request within 1e-12 of the
    streaming result; sampled requests against the oracle."""
    capi = gpu
    c = workloads.config_serving(duration=0.1)
    dims = [m["dim"] for m in c["models"]]
    sub = capi.Server(precision=0, algorithm=capi.SERVE_SUBSPACE)
    ref = capi.Server(precision=0, algorithm=capi.SERVE_STREAMING)
    ids = {}
    for name, srv in (("sub", sub), ("ref", ref)):
        ids[name] = [srv.admit_model(m["basis"], m["level"], m["index"], m["alpha"])
                     for m in c["models"]]
    n_sub = sum(sub.model_algorithm(i) == capi.SERVE_SUBSPACE for i in ids["sub"])
    print("models on the subspace kernel: %d of %d" % (n_sub, len(ids["sub"])), flush=True)
    assert n_sub >= len(ids["sub"]) - 4  # only grids with oversized tables stay streaming
    reqs = []
    for k in range(len(c["q"])):
        m, q, off = int(c["model"][k]), int(c["q"][k]), int(c["offset"][k])
        X = c["pools"][dims[m]][off:off + q]
        reqs.append((k, m, X, sub.submit(ids["sub"][m], X), ref.submit(ids["ref"][m], X)))
    worst = 0.0
    for k, m, X, (rs, n), (rr, _) in reqs:
        st, a, _ = sub.poll(rs, n)
        assert st == capi.REQ_DONE, (k, capi.REQ_NAMES[st])
        st, b, _ = ref.poll(rr, n)
        assert st == capi.REQ_DONE
        e = rel(a, b)
        worst = max(worst, e)
        assert e <= 1e-12, (k, m, e)
        if k % 32 == 0:
            mm = c["models"][m]
            o = oracle.mult(mm["basis"], X[:64], mm["level"], mm["index"], mm["alpha"])
            assert rel(a[:64], o) <= 1e-12
    print("requests %d, worst relative difference %.2e" % (len(reqs), worst), flush=True)
    sub.close()
    ref.close()
