"""Pins of the FP64 CPU oracle (CPU only).

Every oracle function is checked against something independent of its own
formula: the reference's golden fixtures (data/x_10000.bin, rhs_10000.bin),
numpy/LAPACK, closed forms and brute force on small inputs.
"""
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as O


# ---------------------------------------------------------------- fixtures

def test_park_miller_matches_x_fixture(golden_dir):
    """data/x_10000.bin (generate_rhs.cpp:37,43) is Eigen's Random() on the
    Park-Miller minimal-standard rand() from seed 1: bit-exact."""
    x = O.read_cereal_vec(os.path.join(golden_dir, "x_10000.bin"))
    assert x.shape == (10000,)
    assert np.array_equal(x, O.park_miller(10000))


def test_cereal_header(golden_dir):
    raw = open(os.path.join(golden_dir, "rhs_10000.bin"), "rb").read()
    assert len(raw) == 1 + 8 + 8 * 10000 and raw[0] == 1
    assert int(np.frombuffer(raw[1:9], "<u8")[0]) == 10000


def kernel_4d(num_points=10):
    h = 1.0 / num_points
    kii = 1.0 + O.quadruple_integral([0] * 4, [h * 0.5] * 4, 32)
    pts = O.tensor_grid(4, O.uniform_nodes(-1, 1, num_points))
    return O.KernelDesc(O.K_EXP, 4, h ** 4 * (-1.0 / (4 * math.pi)), kii), pts, kii


def test_rhs_fixture_dense_product(golden_dir):
    """data/rhs_10000.bin = dense K x for kernel_4d_test at numPoints = 10
    (generate_rhs.cpp:44-53, user_kernel.hpp:117-177)."""
    k, pts, kii = kernel_4d()
    assert kii == pytest.approx(0.9999995297281162, abs=2e-16)
    x = O.read_cereal_vec(os.path.join(golden_dir, "x_10000.bin"))
    rhs = O.read_cereal_vec(os.path.join(golden_dir, "rhs_10000.bin"))
    y = O.dense_matvec(k, pts, x, nthreads=8)
    assert np.linalg.norm(y - rhs) / np.linalg.norm(rhs) < 1e-14


def test_hodlr_matvec_on_fixture(golden_dir):
    """include/main.cpp flow (NDIM 4, Nmax 1000, d' 1, eps 1e-10) through the
    oracle's HODLR operator: A x_10000 against rhs_10000."""
    k, pts, _ = kernel_4d()
    x = O.read_cereal_vec(os.path.join(golden_dir, "x_10000.bin"))
    rhs = O.read_cereal_vec(os.path.join(golden_dir, "rhs_10000.bin"))
    T = O.Tree(pts, -np.ones(4), np.ones(4), 1000, 1)
    assert T.nlevels == 2
    op = O.Operator(T, k, 1e-10, sys_size=-1)
    y = op.matvec(x)
    assert np.linalg.norm(y - rhs) / np.linalg.norm(rhs) < 1e-8


# ---------------------------------------------------------------- quadrature

@pytest.mark.parametrize("n", [1, 2, 3, 7, 16, 17, 32, 64])
def test_gauss_legendre_vs_numpy(n):
    x, w = O.gauss_legendre(n)
    o = np.argsort(x)
    nx, nw = np.polynomial.legendre.leggauss(n)
    assert np.abs(x[o] - nx).max() < 4e-16
    # numpy's leggauss weights carry ~1e-13 relative error at n = 64; the
    # oracle's are checked bit-exactly against the reference table below
    assert np.abs((w[o] - nw) / nw).max() < 1e-11
    # exact for polynomials of degree 2n - 1
    for deg in range(0, 2 * n, max(1, n // 3)):
        exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
        assert abs(np.dot(w, x ** deg) - exact) < 1e-13


def test_gauss_legendre_reference_order():
    """Gauss_Legendre_Nodes_and_Weights.hpp order: (-x1, +x1, -x2, +x2, ...)
    from the centre outward; odd n starts with the centre node.  Values from
    the table's n = 16 entries."""
    x, w = O.gauss_legendre(16)
    assert x[0] == pytest.approx(-0.0950125098376374401853193354249580631, abs=1e-16)
    assert x[1] == -x[0]
    assert x[15] == pytest.approx(0.989400934991649932596154173450332627, abs=1e-16)
    assert w[0] == pytest.approx(0.18945061045506849628539672320828310515, abs=1e-16)
    # bit-exact against the table's 128-digit values (correctly rounded)
    assert x[0] == -0.095012509837637440185319335424958063130353055689065456697
    assert w[15] == 0.027152459411754094851780572456018103512267375566760797990
    x32, w32 = O.gauss_legendre(32)
    assert x32[31] == 0.99726386184948156354498112866504
    x7, _ = O.gauss_legendre(7)
    assert x7[0] == 0.0 and x7[1] < 0 < x7[2]
    assert np.all(np.diff(np.abs(x[::2])) > 0)


def test_quadruple_integral_known_answer():
    """integral_equation_4D/test.cpp: the 32^4 rule over [-1,1]^4.  Compared
    with the same rule summed pairwise by numpy (the reference sums in one
    sequential loop, hence the 1e-11 tolerance)."""
    v = O.quadruple_integral([-1] * 4, [1] * 4, 32)
    nx, nw = np.polynomial.legendre.leggauss(32)
    X = np.stack(np.meshgrid(nx, nx, nx, nx, indexing="ij"), -1).reshape(-1, 4)
    W = np.einsum("i,j,k,l->ijkl", nw, nw, nw, nw).ravel()
    ref = np.sum(W * (-1 / (4 * math.pi)) * np.exp(-np.sqrt((X ** 2).sum(1))))
    assert v == pytest.approx(ref, rel=1e-11)
    assert v == pytest.approx(-0.4309127131308739, rel=1e-11)


# ---------------------------------------------------------------- entries

def test_entries_closed_form():
    pts = O.splitmix_uniform(5, 30).reshape(10, 3)
    for fam, f in [(O.K_LOG, np.log), (O.K_INV, lambda r: 1 / r), (O.K_EXP, lambda r: np.exp(-r))]:
        w = np.linspace(0.5, 1.5, 10)
        k = O.KernelDesc(fam, 3, 0.7, 2.5, w)
        for i in range(10):
            for j in range(10):
                e = O.entry(k, pts, i, j)
                if i == j:
                    assert e == 2.5
                else:
                    r = np.linalg.norm(pts[i] - pts[j])
                    assert e == pytest.approx(0.7 * f(r) * w[j], rel=1e-14)


def test_uniform_nodes_and_grid_order():
    """myHeaders.hpp:111-120 and user_kernel.hpp:79-95 (dimension 0 outermost)."""
    u = O.uniform_nodes(-1, 1, 5)
    assert np.array_equal(u, np.array([-1.0, -0.5, 0.0, 0.5, 1.0]))
    g = O.tensor_grid(3, u)
    assert g.shape == (125, 3)
    assert np.array_equal(g[1], [-1.0, -1.0, -0.5]) and np.array_equal(g[25], [-0.5, -1.0, -1.0])


def test_splitmix_stream():
    """SplitMix64 (Vigna): first output for seed 0 is 0xE220A8397B1DCDAF."""
    z = 0xE220A8397B1DCDAF
    v = O.splitmix_uniform(0, 1)[0]
    assert v == 2.0 * ((z >> 11) * 2.0 ** -53) - 1.0


# ---------------------------------------------------------------- tree and lists

def cell_index(pts, lo, hi, levels):
    """Independent check of the tree: the cell of each point by integer
    arithmetic on a dyadic grid (valid when no point sits on a cell face)."""
    h = (hi - lo) / (1 << levels)
    return np.floor((pts - lo) / h).astype(np.int64)


@pytest.mark.parametrize("nd,n,nmax,dprime", [(2, 3000, 16, 0), (3, 6000, 32, 0), (3, 4000, 32, -1),
                                               (3, 4000, 32, 1), (3, 4000, 32, 2), (4, 4096, 16, 3),
                                               (4, 5000, 20, -1), (2, 2000, 8, 1)])
def test_tree_closed_form(nd, n, nmax, dprime):
    pts = O.splitmix_uniform(nd * 7919 + n, n * nd).reshape(n, nd)
    lo, hi = -np.ones(nd), np.ones(nd)
    T = O.Tree(pts, lo, hi, nmax, dprime)
    L = T.nlevels - 1
    # uniform depth: the smallest depth whose largest box holds <= nmax
    for depth in range(0, L + 1):
        cells = cell_index(pts, lo, hi, depth)
        keys = np.zeros(n, dtype=np.int64)
        for d in range(nd):
            keys = keys * (1 << depth) + cells[:, d]
        biggest = np.bincount(keys).max() if depth else n
        assert (biggest > nmax) == (depth < L)
    # leaf membership = Morton index of the cell, dimension 0 most significant
    cells = cell_index(pts, lo, hi, L)
    morton = np.zeros(n, dtype=np.int64)
    for bit in range(L - 1, -1, -1):
        for d in range(nd):
            morton = (morton << 1) | ((cells[:, d] >> bit) & 1)
    for j in range(T.level_size(L)):
        p = T.points(L, j)
        assert np.all(np.diff(p) > 0)  # ascending ids (HDD_clusters.hpp:112-124)
        assert np.all(morton[p] == j)
    # lists in closed form: neighbour iff Chebyshev distance 1 and fewer than
    # nd - d' coordinates differ; the interaction list holds the other children
    # of the parent and of the parent's neighbours (HDD_node.hpp:150-180)
    nch = 1 << nd
    for l in range(1, L + 1):
        side = 1 << l
        for j in range(T.level_size(l)):
            c = np.array([(j >> (nd * (l - 1 - b))) for b in range(l)])  # not used directly
            coords = np.zeros(nd, dtype=np.int64)
            for b in range(l):
                for d in range(nd):
                    coords[d] = coords[d] * 2 + ((j >> (nd * (l - 1 - b) + (nd - 1 - d))) & 1)
            parent = j // nch
            cand = [parent * nch + q for q in range(nch) if parent * nch + q != j]
            for pn in T.neighbours(l - 1, parent):
                cand += [pn * nch + q for q in range(nch)]
            nbr, il = [], []
            for o in cand:
                oc = np.zeros(nd, dtype=np.int64)
                for b in range(l):
                    for d in range(nd):
                        oc[d] = oc[d] * 2 + ((o >> (nd * (l - 1 - b) + (nd - 1 - d))) & 1)
                diff = np.abs(oc - coords)
                k = int(np.count_nonzero(diff))
                if diff.max() <= 1 and k < nd - dprime:
                    nbr.append(o)
                else:
                    il.append(o)
            assert list(T.neighbours(l, j)) == nbr
            assert list(T.interactions(l, j)) == il
            assert side > 0 and c is not None


# ---------------------------------------------------------------- ACA

def test_aca_separable_kernel_rank_one():
    """exp(-|x - y|) in 1-D with every row point left of every column point is
    exp(x) exp(-y): rank one; ACA stops at rank 1 after the pivot floor."""
    xs = -1.0 + 0.5 * O.splitmix_uniform(3, 40) * 0.5 - 0.25
    ys = 1.0 - 0.25 * (O.splitmix_uniform(4, 50) + 1) / 2
    pts = np.concatenate([xs, ys]).reshape(-1, 1)
    k = O.KernelDesc(O.K_EXP, 1, 1.0, 1.0)
    rows, cols = np.arange(40), np.arange(40, 90)
    a = O.aca(k, pts, rows, cols, 1e-12)
    assert a["rank"] == 1 and not a["dense"]
    K = O.dense_matrix(k, pts, rows, cols)
    assert np.abs(K - a["U"] @ a["V"].T).max() < 1e-14 * np.abs(K).max()
    # entries evaluated (LowRank_matrix.hpp:126-225): row 0 and its pivot
    # column; the next pivot row's residual is below the floor, and its 10
    # retries (:144-216) each evaluate a row before the loop breaks: 12 rows
    # of n = 50 and 1 column of m = 40
    assert a["retries"] == 10 and a["entries"] == 12 * 50 + 40


def test_aca_cross_reproduces_pivots_and_error():
    pts = O.splitmix_uniform(11, 600).reshape(200, 3)
    pts[:100, 0] -= 3.0  # two separated clusters
    k = O.KernelDesc(O.K_INV, 3, 1.0, 0.0)
    rows, cols = np.arange(100), np.arange(100, 200)
    a = O.aca(k, pts, rows, cols, 1e-10)
    K = O.dense_matrix(k, pts, rows, cols)
    A = a["U"] @ a["V"].T
    assert not a["dense"] and 3 <= a["rank"] < 99
    # residual rows and columns at the pivots vanish (cross interpolation)
    for i in a["row_piv"]:
        assert np.abs(K[i] - A[i]).max() < 1e-12 * np.abs(K).max()
    for j in a["col_piv"]:
        assert np.abs(K[:, j] - A[:, j]).max() < 1e-12 * np.abs(K).max()
    # the absolute pivot floor bounds the error (Frobenius, generous factor)
    assert np.abs(K - A).max() < 1e-7
    # comparable to the SVD truncation at that rank
    s = np.linalg.svd(K, compute_uv=False)
    assert s[a["rank"]] < 1e-8 * s[0] * 100


def _fused(a, u, v):
    """a - u*v rounded once (what an fma would give), via exact rationals."""
    from fractions import Fraction
    return float(Fraction(a) - Fraction(u) * Fraction(v))


def test_aca_residual_rounds_product_then_difference():
    """The reference's residual update `row - u[i](ri) * v[i]` (Eigen, g++
    without -mfma: LowRank_matrix.hpp:130,240) rounds the product, then the
    difference.  The second step's factors are checked bit for bit against
    that two-rounding form computed here by numpy, and the check is shown to
    separate it from a fused (one-rounding) update on this block."""
    pts = O.splitmix_uniform(31, 3 * 120).reshape(120, 3)
    pts[:60, 0] -= 4.0
    k = O.KernelDesc(O.K_INV, 3, 1.0, 0.0)
    rows, cols = np.arange(60), np.arange(60, 120)
    a = O.aca(k, pts, rows, cols, 1e-8)
    # entries ~0.25, far above the pivot floor: the first steps take no retry
    assert a["rank"] >= 3 and not a["dense"]
    K = O.dense_matrix(k, pts, rows, cols)
    U, V = a["U"], a["V"]
    ri0, cj0 = a["row_piv"][0], a["col_piv"][0]
    assert ri0 == 0  # LowRank_matrix.hpp:96-97
    assert np.array_equal(V[:, 0], K[ri0])
    assert np.array_equal(U[:, 0], (1.0 / K[ri0, cj0]) * K[:, cj0])
    ri1, cj1 = a["row_piv"][1], a["col_piv"][1]
    row1 = K[ri1] - U[ri1, 0] * V[:, 0]          # numpy: product, then difference
    assert np.array_equal(V[:, 1], row1)
    col1 = K[:, cj1] - V[cj1, 0] * U[:, 0]
    assert np.array_equal(U[:, 1], (1.0 / row1[cj1]) * col1)
    fused = np.array([_fused(K[ri1, j], U[ri1, 0], V[j, 0]) for j in range(len(cols))])
    assert np.count_nonzero(fused != row1) > 0


def test_aca_dense_fallback_and_tiny_blocks():
    pts = O.splitmix_uniform(12, 30).reshape(10, 3)
    pts[:5, 0] -= 3.0
    k = O.KernelDesc(O.K_INV, 3, 1.0, 0.0)
    a = O.aca(k, pts, np.arange(5), np.arange(5, 10), 1e-14)
    assert a["dense"] and a["rank"] == 5  # r >= min(m, n) - 1
    a1 = O.aca(k, pts, np.arange(1), np.arange(5, 10), 1e-14)
    assert a1["dense"] and a1["rank"] == 1  # one row: no column pivot step
    assert a1["entries"] == 5  # that row only (no row left to pivot: :220-225)


def test_aca_all_below_floor_gives_rank_zero():
    """A block whose entries all sit below eps_ACA: every retry fails, the
    loop breaks at rank 0 (LowRank_matrix.hpp:144-225)."""
    pts = O.splitmix_uniform(13, 300).reshape(100, 3)
    pts[:50, 0] -= 3.0
    k = O.KernelDesc(O.K_INV, 3, 1e-14, 0.0)
    a = O.aca(k, pts, np.arange(50), np.arange(50, 100), 1e-8)
    assert a["rank"] == 0 and not a["dense"] and a["retries"] == 10
    assert a["entries"] == 11 * 50  # row 0 and its 10 retries, no column


# ---------------------------------------------------------------- operator

def small_problem(nd=3, n=2500, seed=21, fam=O.K_INV):
    pts = O.splitmix_uniform(seed, n * nd).reshape(n, nd)
    return pts, O.KernelDesc(fam, nd, 1.0 / n, 1.0)


@pytest.mark.parametrize("eps", [1e-6, 1e-8, 1e-10])
def test_hodlr_matvec_vs_dense(eps):
    pts, k = small_problem()
    T = O.Tree(pts, -np.ones(3), np.ones(3), 64, 0)
    op = O.Operator(T, k, eps)
    x = O.splitmix_uniform(1, len(pts))
    y = op.matvec(x)
    yd = O.dense_matvec(k, pts, x)
    assert np.linalg.norm(y - yd) / np.linalg.norm(yd) < 100 * eps


def test_skeleton_form_equals_stored_factors():
    """LowRank_matrix.hpp:55-58 (K(I, Js) K(Is, Js)^-1 K(Is, J)) against the
    stored cross factors: the same operator in exact arithmetic."""
    pts, k = small_problem(n=3000)
    T = O.Tree(pts, -np.ones(3), np.ones(3), 32, 0)
    x = O.splitmix_uniform(2, len(pts))
    a = O.Operator(T, k, 1e-10, sys_size=-1).matvec(x)
    b_op = O.Operator(T, k, 1e-10, sys_size=64)
    assert b_op.stats()["n_skeleton"] > 0
    b = b_op.matvec(x)
    assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-9


def test_stats_reference_formula():
    """n_FLOP = sum_leaf n (n + sum_nbr n_b) + sum_IL rank (m + n)
    (HDD_node.hpp:183-202); dense fallback counted as min(m, n) (m + n)."""
    pts, k = small_problem(n=2000)
    T = O.Tree(pts, -np.ones(3), np.ones(3), 64, 0)
    op = O.Operator(T, k, 1e-8)
    s = op.stats()
    L = T.nlevels - 1
    e_near = 0.0
    for j in range(T.level_size(L)):
        n = T.npts(L, j)
        if n:
            e_near += n * (n + sum(T.npts(L, b) for b in T.neighbours(L, j)))
    assert s["e_near"] == e_near
    r, d = op.block_ranks()
    assert s["n_blocks"] == len(r) and s["n_dense"] == int(d.sum())


def test_stats_far_field_count_independent():
    """The far-field part of n_FLOP (HDD_node.hpp:194-199: rank (m + n) over
    every interaction-list pair with points on both sides, dense fallback at
    min(m, n)) counted here by walking the tree's lists and compressing each
    pair on its own, against the operator's statistics."""
    pts, k = small_problem(n=1500, seed=23)
    T = O.Tree(pts, -np.ones(3), np.ones(3), 48, 0)
    eps = 1e-8
    op = O.Operator(T, k, eps)
    s = op.stats()
    far, nblk, ndense, rmax = 0.0, 0, 0, 0
    for l in range(1, T.nlevels):
        for j in range(T.level_size(l)):
            rows = T.points(l, j)
            if len(rows) == 0:
                continue
            for b in T.interactions(l, j):
                cols = T.points(l, b)
                if len(cols) == 0:
                    continue
                a = O.aca(k, pts, rows, cols, eps, seed=O.lib.or_block_seed(l, j, int(b)))
                assert a["rank"] == (min(len(rows), len(cols)) if a["dense"] else a["steps"])
                far += a["rank"] * (len(rows) + len(cols))
                nblk += 1
                ndense += a["dense"]
                rmax = max(rmax, a["rank"])
    assert nblk > 100 and 0 < ndense < nblk
    assert s["n_blocks"] == nblk and s["n_dense"] == ndense and s["max_rank"] == rmax
    assert s["n_flop_ref"] - s["e_near"] == far


# ---------------------------------------------------------------- GMRES and QR

def test_gmres_vs_lapack():
    rng = np.random.default_rng(3)
    n = 60
    A = np.eye(n) + 0.3 * rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    x, info = O.gmres_dense(A, b, restart=100, tol=1e-12)
    assert info["resid"] < 1e-12
    assert np.linalg.norm(x - np.linalg.solve(A, b)) / np.linalg.norm(x) < 1e-10
    # restarts (restart < iterations needed) still converge
    x2, info2 = O.gmres_dense(A, b, restart=5, tol=1e-10)
    assert info2["iterations"] > 5 and np.linalg.norm(A @ x2 - b) / np.linalg.norm(b) < 1e-9


def test_gmres_edge_cases():
    n = 20
    b = np.arange(1.0, n + 1)
    # A = I with x0 = b: exits before any progress (GMRES.hpp:198-213)
    x, info = O.gmres_dense(np.eye(n), b)
    assert info["iterations"] == 0 and np.array_equal(x, b)
    # zero right-hand side: normb := 1, x0 = 0 is exact
    x, info = O.gmres_dense(np.eye(n) * 2.0, np.zeros(n))
    assert info["iterations"] == 0
    # max_iter exhausted returns max_iter (GMRES.hpp:298-302)
    rng = np.random.default_rng(4)
    A = rng.standard_normal((n, n))
    x, info = O.gmres_dense(A, rng.standard_normal(n), x0=np.zeros(n), restart=3, max_iter=4, tol=1e-14)
    assert info["iterations"] == 4


def test_cpqr_solve_vs_lapack():
    rng = np.random.default_rng(5)
    for n in (1, 5, 30):
        A = rng.standard_normal((n, n))
        b = rng.standard_normal(n)
        x, rank = O.cpqr_solve(A, b)
        assert rank == n
        assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)


def test_oracle_gmres_on_hodlr():
    pts, k = small_problem(n=2000)
    T = O.Tree(pts, -np.ones(3), np.ones(3), 64, 0)
    op = O.Operator(T, k, 1e-10)
    b = O.splitmix_uniform(8, len(pts))
    x, info = op.gmres(b, tol=1e-10)
    assert info["resid"] < 1e-10
    assert np.linalg.norm(op.matvec(x) - b) / np.linalg.norm(b) < 1e-9


def test_custom_family_pinned():
    """The oracle's custom family (f(r) = Kernel_Fun through a function
    pointer) against what the mathematics fixes: Matern-3/2 with s = 0 is the
    all-ones kernel, so every admissible block is rank 1 and the HODLR
    product is exact; Yukawa entries equal c exp(-r) / r from numpy; a
    Gaussian of width l2 = 1e12 is 1 to 1e-12, so its ranks are those of the
    all-ones kernel."""
    import math
    n = 900
    pts = O.splitmix_uniform(77, n * 3).reshape(n, 3)
    lo, hi = -np.ones(3), np.ones(3)
    T = O.Tree(pts, lo, hi, 32, 0)
    k1 = O.KernelDesc(O.K_CUSTOM, 3, 1.0 / n, 2.0, fn="matern32", params=[0.0])
    A = O.Operator(T, k1, 1e-10)
    r, d = A.block_ranks()
    assert len(r) > 100 and np.all(r[d == 0] == 1) and not d.any()
    x = O.splitmix_uniform(78, n)
    y = A.matvec(x)
    assert np.allclose(y, 2.0 * x + (x.sum() - x) / n, rtol=0, atol=1e-13)
    ky = O.KernelDesc(O.K_CUSTOM, 3, 0.25, 1.0, fn="yukawa", params=[1.0 / (4 * math.pi)])
    for i, j in [(0, 1), (5, 700), (899, 3)]:
        rr = math.sqrt(sum((pts[i, q] - pts[j, q]) ** 2 for q in range(3)))
        want = 0.25 * (1.0 / (4 * math.pi) * math.exp(-rr) / rr)
        assert O.entry(ky, pts, i, j) == want
    assert O.entry(ky, pts, 4, 4) == 1.0
    kg = O.KernelDesc(O.K_CUSTOM, 3, 1.0 / n, 1.0, fn="gauss", params=[1e12])
    rg, dg = O.Operator(T, kg, 1e-10).block_ranks()
    assert np.array_equal(rg, r) and np.array_equal(dg, d)
