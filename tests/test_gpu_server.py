"""GPU tests of the continuous-batching solve server (hodlr_server_*).

Bars:
  * every served column reaches the FP64-certified relative residual 1e-8
    (checked on the host against an FP64 operator assembled directly on
    the same points, which is the same operator: ACA is deterministic);
  * a request's result alone equals its result in a mix of other requests
    and operators submitted before, with and after it, to 1e-12 relative
    (the column arithmetic does not depend on the other columns; the test
    also reports whether the two are bit-identical);
  * HBM admission: with a budget for about one operator, building a second
    evicts the idle first; a later request on the first rebuilds it and
    gives the same result; an operator larger than the budget fails with
    HODLR_E_NOMEM and so do its requests;
  * zero right-hand sides converge at entry with x = 0.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = pytest.importorskip("hodlrnd_amd")
from hodlrnd_amd import workloads as Wl  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def pts3(n, seed):
    return Wl.problem_points(seed, n, 3)


def kern(n):
    return H.Kernel(H.KERNEL_INV_R, 3, 1.0 / n, 1.0)


def direct_op(pts):
    n = len(pts)
    T = H.Tree(pts, -np.ones(3), np.ones(3), 64, 0)
    return H.Operator(T, kern(n), 1e-8)


def rhs(n, k, seed):
    return np.stack([Wl.splitmix_uniform(seed + c, n) for c in range(k)], axis=1)


def check_resid(op, B, X, tol=1e-8):
    R = B - op.matvec(X)
    for c in range(B.shape[1]):
        r = np.linalg.norm(R[:, c]) / np.linalg.norm(B[:, c])
        assert r < tol * 1.01, (c, r)


def test_server_solves_to_tolerance():
    say("case server_basic")
    srv = H.Server(slots=8, tol=1e-8, hbm_budget=8e9)
    P = pts3(4000, 1)
    oid = srv.submit_build(P, kern(4000))
    B = rhs(4000, 3, 10)
    rid = srv.submit_solve(oid, B)
    X, info = srv.wait(rid, timeout=120)
    say(info)
    assert info["rc"] == 0 and info["status"] == 2 and info["converged"] == 1
    assert info["max_resid"] < 1e-8
    check_resid(direct_op(P), B, X)
    st = srv.stats()
    assert st["builds"] == 1 and st["requests_done"] == 1 and st["columns_done"] == 3
    srv.close()


def test_server_alone_equals_in_mix():
    say("case server_alone_vs_mix")
    P1, P2, P3 = pts3(5000, 2), pts3(3000, 3), pts3(7000, 4)
    B = rhs(5000, 4, 20)
    alone = H.Server(slots=8, hbm_budget=8e9)
    o1 = alone.submit_build(P1, kern(5000))
    X_alone, i_alone = alone.wait(alone.submit_solve(o1, B), timeout=120)
    alone.close()
    assert i_alone["converged"] == 1

    mix = H.Server(slots=8, hbm_budget=8e9)
    m1 = mix.submit_build(P1, kern(5000))
    m2 = mix.submit_build(P2, kern(3000))
    m3 = mix.submit_build(P3, kern(7000))
    ids = [mix.submit_solve(m2, rhs(3000, 8, 30)), mix.submit_solve(m1, rhs(5000, 3, 40)),
           mix.submit_solve(m3, rhs(7000, 5, 50))]
    rid = mix.submit_solve(m1, B)  # shares its engine with the 3 columns above
    ids += [mix.submit_solve(m1, rhs(5000, 6, 60)), mix.submit_solve(m2, rhs(3000, 2, 70))]
    X_mix, i_mix = mix.wait(rid, timeout=120)
    for q in ids:
        _, inf = mix.wait(q, timeout=120)
        assert inf["rc"] == 0 and inf["converged"] == 1
    st = mix.stats()
    mix.close()
    e = relerr(X_mix, X_alone)
    say("alone vs mix rel diff", e, "bit-identical", bool(np.array_equal(X_mix, X_alone)), "ticks", st["ticks"])
    assert i_mix["converged"] == 1
    assert e < 1e-12
    assert i_mix["iterations"] == i_alone["iterations"] and i_mix["outer"] == i_alone["outer"]
    check_resid(direct_op(P1), B, X_mix)


def test_server_admission_eviction_rebuild():
    say("case server_admission")
    P1, P2 = pts3(6000, 5), pts3(6000, 6)
    probe = H.Server(slots=4, hbm_budget=8e9)
    o = probe.submit_build(P1, kern(6000))
    probe.wait(probe.submit_solve(o, rhs(6000, 1, 80)), timeout=120)
    one = probe.stats()["used_bytes"]
    probe.close()
    srv = H.Server(slots=4, hbm_budget=1.5 * one)
    o1 = srv.submit_build(P1, kern(6000))
    B = rhs(6000, 2, 90)
    X1, i1 = srv.wait(srv.submit_solve(o1, B), timeout=120)
    o2 = srv.submit_build(P2, kern(6000))
    X2, i2 = srv.wait(srv.submit_solve(o2, rhs(6000, 2, 100)), timeout=120)
    st = srv.stats()
    say("after second build", st)
    assert i1["converged"] == 1 and i2["converged"] == 1
    assert st["evictions"] >= 1 and st["used_bytes"] <= st["budget_bytes"] * 1.0001
    X1b, i1b = srv.wait(srv.submit_solve(o1, B), timeout=120)  # rebuilt from the kept points
    assert i1b["built"] == 1 and i1b["converged"] == 1
    assert np.array_equal(X1b, X1)
    assert srv.stats()["builds"] == 3
    srv.close()


def test_server_rejections_and_zero_rhs():
    say("case server_errors")
    srv = H.Server(slots=2, hbm_budget=1 << 20)
    o = srv.submit_build(pts3(20000, 7), kern(20000))
    try:  # the build may already have failed when the solve is submitted
        rid = srv.submit_solve(o, rhs(20000, 1, 5))
        _, inf = srv.wait(rid, timeout=120)
        assert inf["status"] == H.E_NOMEM
        code = inf["rc"]
    except H.HodlrError as e:
        code = e.code
    assert code == H.E_NOMEM
    with pytest.raises(H.HodlrError):
        srv.submit_solve(12345, np.zeros(10))
    import ctypes as C
    z, rq = np.zeros(10), C.c_int64()
    assert H.lib.hodlr_server_submit_solve(srv.h, 12345, 1, H._p(z), C.byref(rq)) == H.E_ARG
    srv.close()
    srv = H.Server(slots=2, hbm_budget=8e9)
    o = srv.submit_build(pts3(2000, 8), kern(2000))
    B = np.zeros((2000, 2))
    B[:, 1] = Wl.splitmix_uniform(3, 2000)
    X, inf = srv.wait(srv.submit_solve(o, B), timeout=120)
    assert inf["converged"] == 1
    assert not np.any(X[:, 0])
    srv.close()


def test_server_32_slots():
    """32 columns per engine: the FP32 and FP64 32-RHS products (matrix-core
    forms) serve one request of 32 right-hand sides in one pass."""
    say("case server_32_slots")
    srv = H.Server(slots=32, tol=1e-8, hbm_budget=8e9)
    P = pts3(5000, 11)
    oid = srv.submit_build(P, kern(5000))
    B = rhs(5000, 32, 110)
    X, info = srv.wait(srv.submit_solve(oid, B), timeout=120)
    say(info)
    assert info["rc"] == 0 and info["converged"] == 1 and info["max_resid"] < 1e-8
    check_resid(direct_op(P), B, X)
    srv.close()


def test_server_nonfinite_rhs_fails_only_its_request():
    """A right-hand side holding a NaN fails its own request with
    HODLR_E_BREAKDOWN; a request sharing the engine with it completes with
    the same result as alone, and the slots it used serve later requests
    (cleared of its non-finite values) with the same results as alone."""
    say("case server_nan_rhs")
    P = pts3(4000, 12)
    B_good = rhs(4000, 3, 120)
    B_later = rhs(4000, 4, 130)
    alone = H.Server(slots=8, hbm_budget=8e9)
    o = alone.submit_build(P, kern(4000))
    X_good, _ = alone.wait(alone.submit_solve(o, B_good), timeout=120)
    X_later, _ = alone.wait(alone.submit_solve(o, B_later), timeout=120)
    alone.close()

    srv = H.Server(slots=8, hbm_budget=8e9)
    o = srv.submit_build(P, kern(4000))
    B_bad = rhs(4000, 6, 140)
    B_bad[17, 2] = np.nan
    r_bad = srv.submit_solve(o, B_bad)
    r_good = srv.submit_solve(o, B_good)
    _, i_bad = srv.wait(r_bad, timeout=120)
    X1, i_good = srv.wait(r_good, timeout=120)
    say("bad", i_bad, "good", i_good)
    assert i_bad["status"] == H.E_BREAKDOWN and i_bad["rc"] == H.E_BREAKDOWN
    assert i_good["rc"] == 0 and i_good["converged"] == 1
    assert relerr(X1, X_good) < 1e-12
    X2, i_later = srv.wait(srv.submit_solve(o, B_later), timeout=120)  # reuses the failed slots
    assert i_later["rc"] == 0 and i_later["converged"] == 1
    assert np.all(np.isfinite(X2)) and relerr(X2, X_later) < 1e-12
    st = srv.stats()
    srv.close()
    assert st["requests_done"] == 2


def test_server_poisson_stream_vs_oracle():
    """A serving stream shaped like the 200 req/s workload, scaled for test
    time: Poisson arrivals at 200 req/s, N log-uniform in [1000, 8000], 30%
    of requests build a new operator, 70% reuse a resident one with 1-8 new
    right-hand sides.  Every column is certified below 1e-8; the columns of
    requests with N <= 5000 are compared with the FP64 oracle's GMRES
    solution on its own operator (GMRES.hpp:160-303), within 100 tol."""
    say("case server_poisson_stream")
    import time
    from oracle import pyoracle as O
    count = 14
    arrivals = Wl.poisson_arrivals(200.0, count, seed=9)
    u = Wl.splitmix_uniform(77, 3 * count).reshape(count, 3)
    srv = H.Server(slots=16, tol=1e-8, hbm_budget=16e9, builders=2)
    ops, reqs = [], []
    t0 = time.time()
    for q in range(count):
        while time.time() - t0 < arrivals[q]:
            time.sleep(0.0005)
        if not ops or u[q, 0] < -0.4:
            n = Wl.log_uniform_int(u[q, 1], 1000, 8000)
            P = pts3(n, 300 + q)
            ops.append((srv.submit_build(P, kern(n)), P))
            oid, P = ops[-1]
            k = 1
        else:
            oid, P = ops[int((u[q, 1] + 1) / 2 * len(ops)) % len(ops)]
            k = 1 + int((u[q, 2] + 1) / 2 * 8) % 8
        B = rhs(len(P), k, 400 + 10 * q)
        reqs.append((srv.submit_solve(oid, B), P, B))
    checked = 0
    for rid, P, B in reqs:
        X, info = srv.wait(rid, timeout=120)
        assert info["rc"] == 0 and info["converged"] == 1 and info["max_resid"] < 1e-8, info
        n = len(P)
        if n > 5000:
            continue
        T = O.Tree(P, -np.ones(3), np.ones(3), 64, 0)
        ref = O.Operator(T, O.KernelDesc(O.K_INV, 3, 1.0 / n, 1.0), 1e-8, nthreads=16)
        for c in range(B.shape[1]):
            xr, ir = ref.gmres(B[:, c], tol=1e-10, nthreads=16)
            e = relerr(X[:, c], xr)
            assert e < 100 * 1e-8, (rid, c, e)
            checked += 1
    st = srv.stats()
    srv.close()
    say("stream: ops", len(ops), "requests", count, "columns vs oracle", checked, "stats", st)
    assert len(ops) >= 3 and checked >= 5 and st["requests_done"] == count


def test_server_largest_operator_vs_oracle():
    """The serving workload's upper end: an operator of N = 200,000 points
    (3-D uniform, A = I + (1/N)/r, eps_ACA 1e-8, Nmax 64) built by the server
    next to a small one, 4 right-hand sides on it and 2 on the small one in
    flight together.  Every column is certified below 1e-8, and each column
    of the large request matches the FP64 oracle's GMRES solution on the
    oracle's own assembly of that operator (GMRES.hpp:160-303) within 100 tol."""
    say("case server_largest_operator N=200000")
    from oracle import pyoracle as O
    n, n2 = 200000, 3000
    P, P2 = pts3(n, 901), pts3(n2, 902)
    srv = H.Server(slots=8, tol=1e-8, builders=2)
    big = srv.submit_build(P, kern(n))
    small = srv.submit_build(P2, kern(n2))
    B, B2 = rhs(n, 4, 910), rhs(n2, 2, 920)
    r_big = srv.submit_solve(big, B)
    r_small = srv.submit_solve(small, B2)
    X, info = srv.wait(r_big, timeout=300)
    X2, info2 = srv.wait(r_small, timeout=300)
    st = srv.stats()
    srv.close()
    say("served", info, info2, st)
    assert info["rc"] == 0 and info["converged"] == 1 and info["max_resid"] < 1e-8
    assert info2["rc"] == 0 and info2["converged"] == 1 and info2["max_resid"] < 1e-8
    T = O.Tree(P, -np.ones(3), np.ones(3), 64, 0)
    ref = O.Operator(T, O.KernelDesc(O.K_INV, 3, 1.0 / n, 1.0), 1e-8, nthreads=16)
    for c in range(B.shape[1]):
        xr, ir = ref.gmres(B[:, c], tol=1e-10, nthreads=16)
        e = relerr(X[:, c], xr)
        say("column", c, "vs oracle", e, "oracle iterations", ir["iterations"])
        assert e < 100 * 1e-8
    del ref, T
