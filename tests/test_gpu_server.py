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
