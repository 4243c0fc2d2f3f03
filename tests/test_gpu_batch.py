"""GPU parity of batched problems: one block-diagonal operator over many
independent problems of uneven size (hodlr_op_assemble_batch) and the
lockstep per-instance GMRES (hodlr_gmres_batch).

Bars:
  * batch matvec, each problem's rows: bit-identical to the same problem
    assembled alone (the batch changes numbering, not arithmetic);
  * batch GMRES, each instance: bit-identical solution and identical
    iteration count alone and inside any batch (FP64 and mixed modes);
  * batch GMRES instance vs the single-problem solver (hodlr_gmres): same
    iteration count (+-1), solutions within 100 * tol;
  * mixed FP32-inner / FP64-refined batch solves: every instance's true
    FP64 residual below 1e-8.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = pytest.importorskip("hodlrnd_amd")
from oracle import pyoracle as O  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


# (ndim, n, family, nmax) of a small uneven batch: mixed dimensions, kernels
# and leaf sizes, one problem that is a single dense leaf
SPECS = [(3, 3000, O.K_EXP, 64), (2, 5200, O.K_INV, 32), (3, 1800, O.K_INV, 64),
         (2, 40, O.K_EXP, 64), (3, 4100, O.K_EXP, 32)]


def make_problem(i, nd, n, fam, nmax):
    pts = O.splitmix_uniform(900 + i, n * nd).reshape(n, nd)
    tree = H.Tree(pts, -np.ones(nd), np.ones(nd), nmax, 0)
    # identity plus a smooth far field: diag 1 + 1/N, scale 1/N
    kern = H.Kernel(fam, nd, 1.0 / n, 1.0 + 1.0 / n)
    return pts, tree, kern


def same_forms(src, dst):
    """Give dst the matvec kernel forms src uses: each operator picks its
    defaults from its own segment lists, and sums agree bit for bit only
    between the same forms."""
    for prec in (H.PREC_F64, H.PREC_F32):
        for k in (1, 2, 4, 8, 16, 32):
            dst.set_variant(prec, k, src.get_variant(prec, k))


@pytest.fixture(scope="module")
def batch():
    probs = [make_problem(i, *s) for i, s in enumerate(SPECS)]
    trees = [p[1] for p in probs]
    kerns = [p[2] for p in probs]
    op = H.Operator(trees, kerns, 1e-8)
    singles = [H.Operator(t, k, 1e-8) for t, k in zip(trees, kerns)]
    for s in singles:
        same_forms(op, s)
    return probs, op, singles


def test_batch_offsets(batch):
    probs, op, singles = batch
    off = op.offsets()
    assert off[0] == 0 and off[-1] == op.n
    assert np.array_equal(np.diff(off), [s[1] for s in SPECS])


def test_batch_matvec_bit_identical(batch):
    say("case batch_matvec")
    probs, op, singles = batch
    off = op.offsets()
    X = np.stack([O.splitmix_uniform(31 + c, op.n) for c in range(4)], axis=1)
    for prec, K in ((H.PREC_F64, 1), (H.PREC_F64, 4), (H.PREC_F32, 4)):
        Y = op.matvec(X[:, :K], prec=prec)
        for p, s in enumerate(singles):
            Ys = s.matvec(X[off[p]:off[p + 1], :K], prec=prec)
            assert np.array_equal(Y[off[p]:off[p + 1]], Ys), (prec, K, p)
    # one problem against the oracle's HODLR matvec
    pts = probs[0][0]
    nd, n, fam, nmax = SPECS[0]
    T = O.Tree(pts, -np.ones(nd), np.ones(nd), nmax, 0)
    ref = O.Operator(T, O.KernelDesc(fam, nd, 1.0 / n, 1.0 + 1.0 / n, None), 1e-8, sys_size=-1)
    y = op.matvec(X[:, 0])[:n]
    assert relerr(y, ref.matvec(X[:n, 0])) < 1e-12


@pytest.mark.parametrize("mode,K", [(H.SOLVE_F64, 1), (H.SOLVE_F64, 4), (H.SOLVE_MIXED, 2)])
def test_batch_gmres_alone_equals_in_batch(batch, mode, K):
    say("case batch_alone_vs_mix", mode, K)
    probs, op, singles = batch
    off = op.offsets()
    B = np.stack([O.splitmix_uniform(70 + c, op.n) for c in range(K)], axis=1)
    x0 = "zero" if mode == H.SOLVE_MIXED else "b"
    X, infos = op.gmres_batch(B, tol=1e-10 if mode == H.SOLVE_F64 else 1e-8, mode=mode, x0=x0, restart=40)
    for p, s in enumerate(singles):
        Xs, inf_s = s.gmres_batch(B[off[p]:off[p + 1]], tol=1e-10 if mode == H.SOLVE_F64 else 1e-8,
                                  mode=mode, x0=x0, restart=40)
        assert np.array_equal(X[off[p]:off[p + 1]], Xs), p
        for c in range(K):
            a, b = infos[p * K + c], inf_s[c]
            assert a["converged"] and b["converged"]
            assert a["iterations"] == b["iterations"], (p, c)
            assert a["resid"] == b["resid"]


def test_batch_gmres_vs_single_solver(batch):
    say("case batch_vs_single_solver")
    probs, op, singles = batch
    off = op.offsets()
    B = O.splitmix_uniform(5, op.n)
    X, infos = op.gmres_batch(B, tol=1e-10, restart=30)
    for p, s in enumerate(singles):
        b = B[off[p]:off[p + 1]]
        x, info = s.gmres(b, tol=1e-10, restart=30)
        say(p, infos[p]["iterations"], info["iterations"])
        assert abs(infos[p]["iterations"] - info["iterations"]) <= 1
        assert relerr(X[off[p]:off[p + 1], 0], x) < 1e-8


def test_batch_mixed_reaches_1e8(batch):
    say("case batch_mixed_1e-8")
    probs, op, singles = batch
    off = op.offsets()
    K = 8
    B = np.stack([O.splitmix_uniform(200 + c, op.n) for c in range(K)], axis=1)
    X, infos = op.gmres_batch(B, tol=1e-8, mode=H.SOLVE_MIXED, x0="zero")
    R = B - op.matvec(X)
    for p in range(len(SPECS)):
        for c in range(K):
            sl = slice(off[p], off[p + 1])
            r = np.linalg.norm(R[sl, c]) / np.linalg.norm(B[sl, c])
            assert infos[p * K + c]["converged"]
            assert r < 1e-8, (p, c, r)
            assert abs(r - infos[p * K + c]["true_resid"]) < 1e-10


def test_batch_device_entry_zero_rhs(batch):
    """Zero right-hand sides are converged at entry (GMRES.hpp:198-213: the
    relative residual of b = 0 is 0); their instances cost no iterations."""
    say("case batch_zero_rhs")
    probs, op, singles = batch
    off = op.offsets()
    B = np.zeros((op.n, 2))
    B[off[1]:off[2], 1] = O.splitmix_uniform(3, off[2] - off[1])
    X, infos = op.gmres_batch(B, tol=1e-10)
    for p in range(len(SPECS)):
        assert infos[2 * p]["iterations"] == 0
        assert infos[2 * p + 1]["iterations"] == (0 if p != 1 else infos[3]["iterations"])
    assert infos[3]["iterations"] > 0
    assert not np.any(X[:, 0])


def test_batch_step_graphs_replay_exactly(batch):
    """hipGraph replays of the captured Arnoldi steps (hodlr_op_set_solver_graphs)
    against the same solves with every kernel launched separately: bit-identical
    solutions and iteration counts on every replay, with eager work and
    synchronisations between the replays, after a recapture forced by grown
    work buffers, and with a skeleton-form batch operator."""
    say("case step_graphs")
    probs, op, singles = batch
    K = 4
    B = np.stack([O.splitmix_uniform(700 + c, op.n) for c in range(K)], axis=1)
    for mode, x0 in ((H.SOLVE_F64, "b"), (H.SOLVE_MIXED, "zero")):
        op.set_graphs(False)
        Xe, ie = op.gmres_batch(B, tol=1e-8, mode=mode, x0=x0)
        op.set_graphs(True)
        for rep in range(3):
            Xg, ig = op.gmres_batch(B, tol=1e-8, mode=mode, x0=x0)
            say("mode", mode, "replay", rep, "max diff", float(np.max(np.abs(Xg - Xe))))
            assert np.array_equal(Xg, Xe), (mode, rep)
            assert [i["iterations"] for i in ig] == [i["iterations"] for i in ie]
            op.matvec(B[:, 0])  # eager work between replays
        op.matvec(np.concatenate([B] * 4, axis=1))  # 16 RHS: work buffers regrow, graphs recapture
        Xg, _ = op.gmres_batch(B, tol=1e-8, mode=mode, x0=x0)
        assert np.array_equal(Xg, Xe), (mode, "after regrow")
    trees = [p[1] for p in probs]
    kerns = [p[2] for p in probs]
    sk = H.Operator(trees, kerns, 1e-8, sys_size=64)
    assert sk.stats()["n_skeleton"] > 0
    sk.set_graphs(False)
    Xe, _ = sk.gmres_batch(B, tol=1e-8, mode=H.SOLVE_MIXED, x0="zero")
    sk.set_graphs(True)
    for rep in range(2):
        Xg, _ = sk.gmres_batch(B, tol=1e-8, mode=H.SOLVE_MIXED, x0="zero")
        assert np.array_equal(Xg, Xe), ("skeleton", rep)
