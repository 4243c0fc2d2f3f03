"""GPU parity of skeleton blocks (hodlr_op_assemble_skeleton): compressed
blocks with more than sys_size columns keep only their pivot rows and
columns and the pseudo-inverse of their core, and every matvec evaluates
their entries again (LowRank_matrix.hpp:22-36,55-58,685-698).

Bars:
  * ranks, dense flags and the skeleton count equal the oracle's (1/r);
  * FP64 matvec vs the oracle's skeleton-form matvec (the same pivots and
    cores, applied by a CPQR solve on the CPU): 1e-12 relative when the
    cores are well conditioned (eps_ACA 1e-4); 10 eps_ACA otherwise -- the
    pivots of an eps_ACA-accurate cross run down to eps_ACA, so cond(C) ~
    |A| / eps_ACA and any two roundings of A(Is, J) x differ by that much
    after C^+;
  * FP64 matvec vs the dense product: at most 1.5x the oracle's own
    skeleton-form error + 10 eps_ACA.  The skeleton form is the reference's
    ACA_MEM_EFF approximant, which is not the stored cross approximant when
    ACA retried a column (LowRank_matrix.hpp:245-316 keeps the old gamma):
    at sys_size <= Nmax it is ~60x less accurate, in the oracle as here;
  * FP32 operator vs FP64: 1e-5; multi-RHS columns vs single-RHS: 1e-13;
  * a block-diagonal batch equals each problem assembled alone, bit for bit;
  * solves reach the tolerance; operator files round-trip bit for bit;
  * every schedule of the skeleton phases (after the stored passes on the
    calling stream, beside them on a second stream, beside them with capped
    occupancy, the cost model's or the tuner's pick) gives the same bits.
"""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = pytest.importorskip("hodlrnd_amd")
from oracle import pyoracle as O  # noqa: E402
from hodlrnd_amd import workloads as Wl  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


CASES = [
    # name, nd, n, nmax, dprime, eps, family, sys_size, bar vs the oracle's skeleton form
    ("3d_inv_loose", 3, 6000, 32, 0, 1e-4, O.K_INV, 64, 1e-12),
    ("3d_inv_2nmax", 3, 6000, 32, 0, 1e-10, O.K_INV, 64, 1e-9),
    ("3d_inv_all", 3, 4000, 32, 0, 1e-8, O.K_INV, 0, 1e-7),
    ("2d_inv_2nmax", 2, 5000, 32, 0, 1e-10, O.K_INV, 64, 1e-9),
    ("3d_exp_2nmax", 3, 4000, 32, 0, 1e-8, O.K_EXP, 64, None),
    ("2d_log_2nmax", 2, 4000, 32, 0, 1e-10, O.K_LOG, 64, None),
]


@pytest.mark.parametrize("name,nd,n,nmax,dprime,eps,family,sys_size,bar", CASES)
def test_skeleton_matvec_parity(name, nd, n, nmax, dprime, eps, family, sys_size, bar):
    say("case", name)
    scale = -1.0 / (2 * math.pi) / n if family == O.K_LOG else 1.0 / n
    pts = O.splitmix_uniform(300 + n, n * nd).reshape(n, nd)
    lo, hi = -np.ones(nd), np.ones(nd)
    T = H.Tree(pts, lo, hi, nmax, dprime)
    kern = H.Kernel(family, nd, scale, 1.0)
    A = H.Operator(T, kern, eps, sys_size=sys_size)
    ok = O.KernelDesc(family, nd, scale, 1.0, None)
    ref = O.Operator(O.Tree(pts, lo, hi, nmax, dprime), ok, eps, sys_size=sys_size)
    st, rst = A.stats(), ref.stats()
    say(name, "skeleton blocks", st["n_skeleton"], "oracle", rst["n_skeleton"], "e_skel", st["e_skel"],
        "e_far stored", st["e_far"], "blocks", st["n_blocks"])
    assert st["n_skeleton"] > 0
    r_gpu, d_gpu = A.block_ranks()
    r_ref, d_ref = ref.block_ranks()
    same = np.mean((r_gpu == r_ref) & (d_gpu == d_ref))
    x = O.splitmix_uniform(9, n)
    y = A.matvec(x)
    y_or = ref.matvec(x)
    yd = O.dense_matvec(ok, pts, x)
    e_dense, e_ref, e_or_dense = relerr(y, yd), relerr(y, y_or), relerr(y_or, yd)
    say(name, "rank agreement", same, "vs oracle skeleton form", e_ref, "vs dense", e_dense,
        "oracle skeleton form vs dense", e_or_dense)
    if family == O.K_INV:
        assert same == 1.0 and st["n_skeleton"] == rst["n_skeleton"]
        assert e_ref < bar
    else:
        assert same >= 0.99
    assert e_dense < 1.5 * e_or_dense + 10 * eps
    y32 = A.matvec(x, prec=H.PREC_F32)
    say(name, "fp32 vs fp64", relerr(y32, y))
    assert relerr(y32, y) < 1e-5


def test_skeleton_multi_rhs_columns():
    say("case multi_rhs")
    n, nd = 5000, 3
    pts = O.splitmix_uniform(41, n * nd).reshape(n, nd)
    T = H.Tree(pts, -np.ones(nd), np.ones(nd), 32, 0)
    A = H.Operator(T, H.Kernel(H.KERNEL_INV_R, nd, 1.0 / n, 1.0), 1e-8, sys_size=64)
    X = np.stack([O.splitmix_uniform(50 + c, n) for c in range(16)], axis=1)
    for prec, ks in ((H.PREC_F64, (2, 4, 8, 16)), (H.PREC_F32, (4, 16, 32))):
        y1 = np.stack([A.matvec(X[:, c], prec=prec) for c in range(X.shape[1])], axis=1)
        for k in ks:
            Xk = np.concatenate([X] * (k // 16 + 1), axis=1)[:, :k]
            Yk = A.matvec(Xk, prec=prec)
            ref = np.concatenate([y1] * (k // 16 + 1), axis=1)[:, :k]
            e = max(relerr(Yk[:, c], ref[:, c]) for c in range(k))
            say("prec", prec, "k", k, "max column error", e)
            assert e < (1e-13 if prec == H.PREC_F64 else 1e-6)


def test_skeleton_weighted_4d_gauss_legendre():
    say("case weighted_4d")
    gx, gw = O.gauss_legendre(8)
    pts = O.tensor_grid(4, gx)
    wts = O.tensor_weights(4, gw)
    n = len(pts)
    lo, hi = -np.ones(4), np.ones(4)
    T = H.Tree(pts, lo, hi, 16, 0)
    A = H.Operator(T, H.Kernel(H.KERNEL_INV_R, 4, 1.0, 1.0, wts), 1e-8, sys_size=32)
    ok = O.KernelDesc(O.K_INV, 4, 1.0, 1.0, wts)
    ref = O.Operator(O.Tree(pts, lo, hi, 16, 0), ok, 1e-8, sys_size=32)
    st = A.stats()
    say("skeleton blocks", st["n_skeleton"], "oracle", ref.stats()["n_skeleton"])
    assert st["n_skeleton"] > 0 and st["n_skeleton"] == ref.stats()["n_skeleton"]
    x = O.splitmix_uniform(12, n)
    y = A.matvec(x)
    y_or = ref.matvec(x)
    yd = O.dense_matvec(ok, pts, x)
    e_ref, e_dense, e_or_dense = relerr(y, y_or), relerr(y, yd), relerr(y_or, yd)
    say("vs oracle skeleton form", e_ref, "vs dense", e_dense, "oracle skeleton form vs dense", e_or_dense)
    assert e_ref < 10 * 1e-8 and e_dense < 1.5 * e_or_dense + 10 * 1e-8


def test_skeleton_solves():
    say("case solves")
    n, nd = 8000, 3
    pts = O.splitmix_uniform(61, n * nd).reshape(n, nd)
    lo, hi = -np.ones(nd), np.ones(nd)
    kern = H.Kernel(H.KERNEL_INV_R, nd, 1.0 / n, 1.0)
    A = H.HODLRdD_matrix(kern, pts, lo, hi, nmax=32, dprime=0, eps_aca=1e-10, sys_size=64)
    A.Assemble_matrix_operators()
    assert A.op.stats()["n_skeleton"] > 0
    ref = O.Operator(O.Tree(pts, lo, hi, 32, 0), O.KernelDesc(O.K_INV, nd, 1.0 / n, 1.0, None), 1e-10,
                     sys_size=64)
    b = O.splitmix_uniform(62, n)
    xs = A.solve(b)
    xr, ir = ref.gmres(b, tol=1e-10)
    say("gmres gpu", A.last_info["iterations"], "oracle", ir["iterations"], "diff", relerr(xs, xr))
    assert A.last_info["converged"] and abs(A.last_info["iterations"] - ir["iterations"]) <= 1
    assert relerr(xs, xr) < 100 * 1e-10
    x, info = A.op.gmres(b, tol=1e-8, mode=H.SOLVE_MIXED, x0=np.zeros(n))
    say("mixed", info["iterations"], info["outer"], info["true_resid"])
    assert info["converged"] and info["true_resid"] < 1e-8


def test_skeleton_batch_equals_alone():
    say("case batch")
    specs = [(3, 3000, H.KERNEL_INV_R), (2, 4500, H.KERNEL_EXP), (3, 2000, H.KERNEL_EXP)]
    trees, kerns, alone = [], [], []
    for i, (nd, n, fam) in enumerate(specs):
        P = Wl.problem_points(80 + i, n, nd)
        trees.append(H.Tree(P, -np.ones(nd), np.ones(nd), 32, 0))
        kerns.append(H.Kernel(fam, nd, 1.0 / n, 1.0 + 1.0 / n))
        alone.append(H.Operator(trees[-1], kerns[-1], 1e-8, sys_size=64))
    B = H.Operator(trees, kerns, 1e-8, sys_size=64)
    for a in alone:  # each operator picks its default forms from its own segment lists
        for prec in (H.PREC_F64, H.PREC_F32):
            for k in (1, 4):
                a.set_variant(prec, k, B.get_variant(prec, k))
    off = B.offsets()
    assert B.stats()["n_skeleton"] == sum(a.stats()["n_skeleton"] for a in alone) > 0
    for prec in (H.PREC_F64, H.PREC_F32):
        for k in (1, 4):
            X = np.stack([Wl.splitmix_uniform(90 + c, B.n) for c in range(k)], axis=1)
            Y = B.matvec(X if k > 1 else X[:, 0], prec=prec).reshape(B.n, k)
            for p, a in enumerate(alone):
                Xp = X[off[p]:off[p + 1]]
                Yp = a.matvec(Xp if k > 1 else Xp[:, 0], prec=prec).reshape(-1, k)
                assert np.array_equal(Y[off[p]:off[p + 1]], Yp), (prec, k, p)


def test_skeleton_operator_file(tmp_path):
    say("case opfile")
    n, nd = 5000, 3
    pts = O.splitmix_uniform(71, n * nd).reshape(n, nd)
    T = H.Tree(pts, -np.ones(nd), np.ones(nd), 32, 0)
    A = H.Operator(T, H.Kernel(H.KERNEL_INV_R, nd, 1.0 / n, 1.0), 1e-8, sys_size=64)
    path = tmp_path / "sk.op"
    A.save(path)
    B = H.Operator.load(path)
    sa, sb = A.stats(), B.stats()
    for key in ("n_skeleton", "e_skel", "e_far", "bytes_core", "n_blocks"):
        assert sa[key] == sb[key], key
    x = O.splitmix_uniform(72, n)
    for prec in (H.PREC_F64, H.PREC_F32):
        assert np.array_equal(A.matvec(x, prec=prec), B.matvec(x, prec=prec))
    path2 = tmp_path / "sk2.op"
    B.save(path2)
    assert path.read_bytes() == path2.read_bytes()


def test_skeleton_schedules_bit_identical():
    say("case schedules")
    n, nd = 6000, 3
    pts = O.splitmix_uniform(81, n * nd).reshape(n, nd)
    T = H.Tree(pts, -np.ones(nd), np.ones(nd), 32, 0)
    A = H.Operator(T, H.Kernel(H.KERNEL_INV_R, nd, 1.0 / n, 1.0), 1e-8, sys_size=64)
    assert A.stats()["n_skeleton"] > 0
    X = np.stack([O.splitmix_uniform(82 + c, n) for c in range(32)], axis=1)
    modes = (H.SKEL_SERIAL, H.SKEL_OVERLAP, H.SKEL_CAPPED, H.SKEL_AUTO)
    for prec, ks in ((H.PREC_F64, (1, 4, 16)), (H.PREC_F32, (1, 8, 32))):
        for k in ks:
            Xk = X[:, :k] if k > 1 else X[:, 0]
            ys = {}
            for m in modes:
                A.set_skeleton_schedule(m)
                ys[m] = [A.matvec(Xk, prec=prec) for _ in range(2)]
            A.set_skeleton_schedule(H.SKEL_AUTO)
            say("prec", prec, "k", k, "auto schedule", A.skeleton_schedule(prec, k))
            assert A.skeleton_schedule(prec, k) in (H.SKEL_OVERLAP, H.SKEL_CAPPED)
            for m in modes:
                for y in ys[m]:
                    assert np.array_equal(y, ys[H.SKEL_SERIAL][0]), (prec, k, m)
    b = O.splitmix_uniform(90, n)
    sols = []
    for m in (H.SKEL_SERIAL, H.SKEL_CAPPED):
        A.set_skeleton_schedule(m)
        x, info = A.gmres(b, tol=1e-8, mode=H.SOLVE_MIXED, x0=np.zeros(n))
        assert info["converged"] and info["true_resid"] < 1e-8
        sols.append((x, info["iterations"]))
    assert sols[0][1] == sols[1][1] and np.array_equal(sols[0][0], sols[1][0])
    A.set_skeleton_schedule(H.SKEL_AUTO)
    A.tune(reps=1)
    for prec, k in ((H.PREC_F64, 1), (H.PREC_F32, 1), (H.PREC_F32, 8)):
        m = A.skeleton_schedule(prec, k)
        say("tuned schedule prec", prec, "k", k, m)
        assert m in (H.SKEL_SERIAL, H.SKEL_OVERLAP, H.SKEL_CAPPED)
    assert np.array_equal(A.matvec(X[:, :8], prec=H.PREC_F32), A.matvec(X[:, :8], prec=H.PREC_F32))
    A.set_skeleton_schedule(H.SKEL_SERIAL)
    y_serial = A.matvec(X[:, :8], prec=H.PREC_F32)
    A.set_skeleton_schedule(H.SKEL_AUTO)
    assert np.array_equal(A.matvec(X[:, :8], prec=H.PREC_F32), y_serial)
    with pytest.raises(H.HodlrError):
        A.set_skeleton_schedule(7)
