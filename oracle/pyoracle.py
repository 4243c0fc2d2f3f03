"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

K_LOG, K_INV, K_EXP, K_CUSTOM = 0, 1, 2, 3


class Kernel(C.Structure):
    _fields_ = [("family", C.c_int32), ("ndim", C.c_int32), ("scale", C.c_double),
                ("diag", C.c_double), ("weight", C.c_void_p), ("fn", C.c_void_p),
                ("params", C.c_void_p)]


class AcaInfo(C.Structure):
    _fields_ = [("rank", C.c_int32), ("dense", C.c_int32), ("steps", C.c_int32),
                ("retries", C.c_int32), ("entries", C.c_int64)]


class OpStats(C.Structure):
    _fields_ = [("n_flop_ref", C.c_double), ("max_rank", C.c_int64), ("e_near", C.c_double),
                ("e_far", C.c_double), ("n_blocks", C.c_int64), ("n_dense", C.c_int64),
                ("n_skeleton", C.c_int64), ("init_seconds", C.c_double), ("aca_entries", C.c_double)]


class GmresInfo(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("matvecs", C.c_int32), ("resid", C.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB):
        build()
    lib = C.CDLL(_LIB)
    P = C.c_void_p
    d, i32, i64, u64 = C.c_double, C.c_int32, C.c_int64, C.c_uint64
    sig = {
        "or_uniform_nodes": (None, [d, d, C.c_int, P]),
        "or_tensor_grid": (None, [C.c_int, C.c_int, P, P]),
        "or_tensor_weights": (None, [C.c_int, C.c_int, P, P]),
        "or_park_miller": (None, [i64, P]),
        "or_splitmix_uniform": (None, [u64, i64, P]),
        "or_gauss_legendre": (C.c_int, [C.c_int, P, P]),
        "or_quadruple_integral": (d, [P, P, C.c_int]),
        "or_entry": (d, [P, P, i64, i64]),
        "or_dense_matvec": (None, [P, P, i64, P, P, C.c_int]),
        "or_tree_build": (P, [C.c_int, i64, P, P, P, C.c_int, C.c_int]),
        "or_tree_free": (None, [P]),
        "or_tree_nlevels": (C.c_int, [P]),
        "or_tree_level_size": (i64, [P, C.c_int]),
        "or_tree_node_npts": (i64, [P, C.c_int, i64]),
        "or_tree_node_points": (None, [P, C.c_int, i64, P]),
        "or_tree_list": (i64, [P, C.c_int, i64, C.c_int, P]),
        "or_interaction_type": (C.c_int, [P, C.c_int, i64, i64]),
        "or_block_seed": (u64, [C.c_int, i64, i64]),
        "or_aca": (C.c_int, [P, P, P, C.c_int, P, C.c_int, d, u64, C.c_int, P, P, P, P, P]),
        "or_op_assemble": (P, [P, P, P, d, C.c_int, C.c_int]),
        "or_op_assemble_sample": (P, [P, P, P, d, C.c_int, C.c_int, P, P]),
        "or_op_free": (None, [P]),
        "or_op_matvec": (None, [P, P, P, C.c_int]),
        "or_op_get_stats": (None, [P, P]),
        "or_op_block_ranks": (i64, [P, P, P]),
        "or_gmres": (C.c_int, [P, P, P, C.c_int, C.c_int, d, C.c_int, P, C.c_int, P]),
        "or_gmres_dense": (C.c_int, [P, i64, P, P, C.c_int, C.c_int, d, P]),
        "or_cpqr_solve": (C.c_int, [C.c_int, P, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def uniform_nodes(a, b, n):
    out = np.empty(n)
    lib.or_uniform_nodes(a, b, n, _p(out))
    return out


def tensor_grid(ndim, nodes1d):
    nodes1d = f64(nodes1d)
    n = len(nodes1d)
    pts = np.empty((n ** ndim, ndim))
    lib.or_tensor_grid(ndim, n, _p(nodes1d), _p(pts))
    return pts


def tensor_weights(ndim, w1d):
    w1d = f64(w1d)
    n = len(w1d)
    w = np.empty(n ** ndim)
    lib.or_tensor_weights(ndim, n, _p(w1d), _p(w))
    return w


def park_miller(n):
    out = np.empty(n)
    lib.or_park_miller(n, _p(out))
    return out


def splitmix_uniform(seed, n):
    out = np.empty(n)
    lib.or_splitmix_uniform(seed, n, _p(out))
    return out


def gauss_legendre(n):
    x, w = np.empty(n), np.empty(n)
    assert lib.or_gauss_legendre(n, _p(x), _p(w)) == 0
    return x, w


def quadruple_integral(a, b, ngauss=32):
    a, b = f64(a), f64(b)
    return lib.or_quadruple_integral(_p(a), _p(b), ngauss)


class KernelDesc:
    """entry(i, j) = diag if i == j else scale * f(|x_i - x_j|) * w_j;
    family K_CUSTOM takes f by name: one of CUSTOM_F ("yukawa", "matern32",
    "gauss") with its parameters"""

    def __init__(self, family, ndim, scale, diag, weight=None, fn=None, params=None):
        self.weight = None if weight is None else f64(weight)
        self.params = None if params is None else f64(params)
        fptr = None
        if family == K_CUSTOM:
            fptr = C.cast(getattr(lib, CUSTOM_F[fn]), C.c_void_p).value
        self.c = Kernel(family, ndim, scale, diag,
                        None if self.weight is None else self.weight.ctypes.data, fptr,
                        None if self.params is None else self.params.ctypes.data)

    @property
    def ref(self):
        return C.byref(self.c)


CUSTOM_F = {"yukawa": "or_f_yukawa", "matern32": "or_f_matern32", "gauss": "or_f_gauss"}


def entry(k, pts, i, j):
    return lib.or_entry(k.ref, _p(pts), i, j)


def dense_matvec(k, pts, x, nthreads=8):
    pts, x = f64(pts), f64(x)
    y = np.empty(len(x))
    lib.or_dense_matvec(k.ref, _p(pts), len(x), _p(x), _p(y), nthreads)
    return y


def dense_matrix(k, pts, rows=None, cols=None):
    n = len(pts)
    rows = np.arange(n) if rows is None else rows
    cols = np.arange(n) if cols is None else cols
    return np.array([[lib.or_entry(k.ref, _p(pts), int(i), int(j)) for j in cols] for i in rows])


class Tree:
    def __init__(self, pts, lo, hi, nmax, dprime):
        self.pts = f64(pts)
        self.ndim = self.pts.shape[1]
        lo, hi = f64(lo), f64(hi)
        self.h = lib.or_tree_build(self.ndim, len(self.pts), _p(self.pts), _p(lo), _p(hi),
                                   nmax, dprime)
        assert self.h

    def __del__(self):
        if getattr(self, "h", None):
            lib.or_tree_free(self.h)
            self.h = None

    @property
    def nlevels(self):
        return lib.or_tree_nlevels(self.h)

    def level_size(self, l):
        return lib.or_tree_level_size(self.h, l)

    def points(self, l, j):
        n = lib.or_tree_node_npts(self.h, l, j)
        out = np.empty(n, dtype=np.int64)
        if n:
            lib.or_tree_node_points(self.h, l, j, _p(out))
        return out

    def npts(self, l, j):
        return lib.or_tree_node_npts(self.h, l, j)

    def lst(self, l, j, which):
        n = lib.or_tree_list(self.h, l, j, which, None)
        out = np.empty(n, dtype=np.int64)
        if n:
            lib.or_tree_list(self.h, l, j, which, _p(out))
        return out

    def neighbours(self, l, j):
        return self.lst(l, j, 0)

    def interactions(self, l, j):
        return self.lst(l, j, 1)


def aca(k, pts, rows, cols, eps, seed=0):
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    m, n = len(rows), len(cols)
    cap = min(m, n)
    U = np.zeros((cap, m))
    V = np.zeros((cap, n))
    rp = np.zeros(cap + 1, dtype=np.int32)
    cp = np.zeros(cap + 1, dtype=np.int32)
    info = AcaInfo()
    lib.or_aca(k.ref, _p(f64(pts)), _p(rows), m, _p(cols), n, eps, seed, cap,
               _p(U), _p(V), _p(rp), _p(cp), C.byref(info))
    r = info.steps
    return dict(rank=info.rank, dense=bool(info.dense), steps=r, retries=info.retries, entries=info.entries,
                U=U[:r].T.copy(), V=V[:r].T.copy(), row_piv=rp[:r].copy(), col_piv=cp[:r].copy())


class Operator:
    def __init__(self, tree, k, eps, sys_size=-1, nthreads=8, sample=None):
        self.tree, self.k = tree, k
        if sample is None:
            self.h = lib.or_op_assemble(tree.h, k.ref, _p(tree.pts), eps, sys_size, nthreads)
        else:
            stride = np.ascontiguousarray(sample[0], dtype=np.int64)
            offset = np.ascontiguousarray(sample[1], dtype=np.int64)
            self.h = lib.or_op_assemble_sample(tree.h, k.ref, _p(tree.pts), eps, sys_size,
                                               nthreads, _p(stride), _p(offset))
        self.n = len(tree.pts)

    def __del__(self):
        if getattr(self, "h", None):
            lib.or_op_free(self.h)
            self.h = None

    def matvec(self, x, nthreads=8):
        x = f64(x)
        y = np.empty(self.n)
        lib.or_op_matvec(self.h, _p(x), _p(y), nthreads)
        return y

    def stats(self):
        s = OpStats()
        lib.or_op_get_stats(self.h, C.byref(s))
        return {f: getattr(s, f) for f, _ in OpStats._fields_}

    def block_ranks(self):
        nb = lib.or_op_block_ranks(self.h, None, None)
        r = np.empty(nb, dtype=np.int32)
        dn = np.empty(nb, dtype=np.int32)
        lib.or_op_block_ranks(self.h, _p(r), _p(dn))
        return r, dn

    def gmres(self, b, x0=None, restart=100, max_iter=None, tol=1e-10, nthreads=8):
        b = f64(b)
        x = b.copy() if x0 is None else f64(x0).copy()
        max_iter = self.n if max_iter is None else max_iter
        hist = np.zeros(max_iter + 2)
        info = GmresInfo()
        lib.or_gmres(self.h, _p(b), _p(x), restart, max_iter, tol, nthreads, _p(hist),
                     len(hist), C.byref(info))
        return x, dict(iterations=info.iterations, matvecs=info.matvecs, resid=info.resid)


def gmres_dense(A, b, x0=None, restart=100, max_iter=None, tol=1e-10):
    A = f64(A)
    b = f64(b)
    n = len(b)
    x = b.copy() if x0 is None else f64(x0).copy()
    info = GmresInfo()
    lib.or_gmres_dense(_p(A), n, _p(b), _p(x), restart, n if max_iter is None else max_iter,
                       tol, C.byref(info))
    return x, dict(iterations=info.iterations, matvecs=info.matvecs, resid=info.resid)


def cpqr_solve(A, b):
    A = np.asfortranarray(A, dtype=np.float64)
    b = f64(b)
    x = np.empty_like(b)
    rank = lib.or_cpqr_solve(len(b), A.ctypes.data_as(C.c_void_p), _p(b), _p(x))
    return x, rank


def read_cereal_vec(path):
    """cereal portable-binary std::vector<double>: 1-byte endianness flag,
    u64 element count, little-endian doubles (generate_rhs.cpp:43,53)."""
    raw = open(path, "rb").read()
    assert raw[0] == 1
    n = int(np.frombuffer(raw[1:9], dtype="<u8")[0])
    return np.frombuffer(raw[9:9 + 8 * n], dtype="<f8").copy()
