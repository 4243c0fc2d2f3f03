"""GPU round trip of operator files (hodlr_op_save / hodlr_op_load).

Bar: a loaded operator is the stored one -- block ranks and statistics
equal, FP64 and FP32 matvecs bit-identical for 1 and 8 right-hand sides,
GMRES solutions bit-identical -- for a single problem with column weights
(4D Gauss-Legendre grid) and for a batch of uneven problems; a corrupted
arena fails the checksum.
"""
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


def same_operator(a, b, n):
    ra, da = a.block_ranks()
    rb, db = b.block_ranks()
    assert np.array_equal(ra, rb) and np.array_equal(da, db)
    sa, sb = a.stats(), b.stats()
    for k in ("n_flop_ref", "max_rank", "e_near", "e_far", "n_blocks", "n_dense", "bytes_f64", "bytes_f32"):
        assert sa[k] == sb[k], k
    X = np.stack([Wl.splitmix_uniform(40 + c, n) for c in range(8)], axis=1)
    for prec in (H.PREC_F64, H.PREC_F32):
        assert np.array_equal(a.matvec(X[:, 0], prec=prec), b.matvec(X[:, 0], prec=prec))
        assert np.array_equal(a.matvec(X, prec=prec), b.matvec(X, prec=prec))


def test_operator_file_single_weighted(tmp_path):
    say("case opfile_weighted")
    gx, gw = O.gauss_legendre(6)
    pts = O.tensor_grid(4, gx)
    wts = O.tensor_weights(4, gw)
    n = len(pts)
    T = H.Tree(pts, -np.ones(4), np.ones(4), 16, 0)
    A = H.Operator(T, H.Kernel(H.KERNEL_INV_R, 4, 1.0, 1.0, wts), 1e-8)
    path = tmp_path / "w.op"
    A.save(path)
    B = H.Operator.load(path)
    assert B.n == n and np.array_equal(B.trees[0].perm(), T.perm())
    same_operator(A, B, n)
    b = Wl.splitmix_uniform(9, n)
    xa, ia = A.gmres(b, tol=1e-10)
    xb, ib = B.gmres(b, tol=1e-10)
    assert ia["iterations"] == ib["iterations"] and np.array_equal(xa, xb)
    # the file was written from the loaded operator too: identical bytes
    path2 = tmp_path / "w2.op"
    B.save(path2)
    assert path.read_bytes() == path2.read_bytes()


def test_operator_file_batch_and_checksum(tmp_path):
    say("case opfile_batch")
    specs = [(3, 2500, H.KERNEL_EXP), (2, 4000, H.KERNEL_INV_R), (3, 1200, H.KERNEL_INV_R)]
    trees, kerns = [], []
    for i, (nd, n, fam) in enumerate(specs):
        P = Wl.problem_points(60 + i, n, nd)
        trees.append(H.Tree(P, -np.ones(nd), np.ones(nd), 32, 0))
        kerns.append(H.Kernel(fam, nd, 1.0 / n, 1.0 + 1.0 / n))
    A = H.Operator(trees, kerns, 1e-8)
    path = tmp_path / "b.op"
    A.save(path)
    B = H.Operator.load(path)
    assert len(B.trees) == 3 and np.array_equal(B.offsets(), A.offsets())
    same_operator(A, B, A.n)
    Bm = np.stack([Wl.splitmix_uniform(70 + c, A.n) for c in range(4)], axis=1)
    Xa, _ = A.gmres_batch(Bm, tol=1e-8, mode=H.SOLVE_MIXED, x0="zero")
    Xb, _ = B.gmres_batch(Bm, tol=1e-8, mode=H.SOLVE_MIXED, x0="zero")
    assert np.array_equal(Xa, Xb)
    raw = bytearray(path.read_bytes())
    raw[len(raw) // 2] ^= 0x40  # flip one bit inside the arenas
    bad = tmp_path / "bad.op"
    bad.write_bytes(bytes(raw))
    with pytest.raises(H.HodlrError) as e:
        H.Operator.load(bad)
    assert e.value.code == H.E_ARG
