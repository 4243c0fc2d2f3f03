"""Python host binding of libhodlrnd (MI355X / gfx950 HODLRnD hot path).

Mirrors the reference's operator API (include/HDD_matrix.hpp:20-85,
include/GMRES.hpp:44-50) over the C ABI declared in include/hodlrnd.h:

    A = HODLRdD_matrix(kernel, points, x1, x2, nmax=..., dprime=..., eps_aca=...)
    A.Assemble_matrix_operators()
    y = A * x
    x = A.solve(b)
    A.print_matrix_statistics()

Every compute call goes to the HIP library; there is no CPU fallback.  If
libhodlrnd.so has not been built the import raises.
"""
import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libhodlrnd.so")

OK, E_ARG, E_DEVICE, E_NOMEM, E_RANK, E_NOCONV, E_BREAKDOWN, E_STATE = 0, -1, -2, -3, -4, -5, -6, -7
KERNEL_LOG, KERNEL_INV_R, KERNEL_EXP, KERNEL_CUSTOM = 0, 1, 2, 3
PREC_F64, PREC_F32 = 1, 2
SOLVE_F64, SOLVE_MIXED = 0, 1


class HodlrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("hodlrnd error %d: %s" % (code, msg))
        self.code = code


MV_PROJ_MFMA, MV_PANEL_MFMA, MV_PANEL_FLAT, MV_PANEL_VEC, MV_PROJ_PACK, MV_PANEL_WIDE, MV_PROJ_WAVE = \
    1, 2, 4, 8, 16, 32, 64
# where a matvec evaluates its skeleton blocks: after the stored passes on the
# calling stream, beside them on a second stream, or beside them with capped
# occupancy; AUTO = the tuner's pick, else a cost model (same bits for all)
SKEL_AUTO, SKEL_SERIAL, SKEL_OVERLAP, SKEL_CAPPED = 0, 1, 2, 3
# Arnoldi orthogonalisation of single-problem solves: two classical
# Gram-Schmidt passes (default) or modified Gram-Schmidt as the reference
ORTHO_CGS2, ORTHO_MGS = 0, 1


class KernelDesc(C.Structure):
    _fields_ = [("family", C.c_int32), ("ndim", C.c_int32), ("scale", C.c_double),
                ("diag", C.c_double), ("weight", C.c_void_p)]


class TreeInfo(C.Structure):
    _fields_ = [("nlevels", C.c_int32), ("ndim", C.c_int32), ("npoints", C.c_int64),
                ("nnodes", C.c_int64), ("nleaves", C.c_int64), ("max_leaf", C.c_int64),
                ("n_admissible", C.c_int64), ("n_near", C.c_int64)]


class OpStats(C.Structure):
    _fields_ = [("n_flop_ref", C.c_double), ("max_rank", C.c_int64), ("e_near", C.c_double),
                ("e_far", C.c_double), ("n_blocks", C.c_int64), ("n_dense", C.c_int64),
                ("bytes_f64", C.c_double), ("bytes_f32", C.c_double),
                ("matvec_bytes_f64", C.c_double), ("matvec_bytes_f32", C.c_double),
                ("assemble_seconds", C.c_double), ("rank_seconds", C.c_double),
                ("fill_seconds", C.c_double), ("n_skeleton", C.c_int64), ("e_skel", C.c_double),
                ("bytes_core", C.c_double), ("aca_entries", C.c_double)]


class GmresParams(C.Structure):
    _fields_ = [("restart", C.c_int32), ("max_iter", C.c_int32), ("tol", C.c_double),
                ("mode", C.c_int32), ("x0_is_b", C.c_int32), ("inner_tol", C.c_double),
                ("max_outer", C.c_int32), ("verify", C.c_int32), ("hist_cap", C.c_int32),
                ("hist", C.c_void_p), ("ortho", C.c_int32)]


class GmresInfo(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("matvecs_f64", C.c_int32), ("matvecs_f32", C.c_int32),
                ("outer", C.c_int32), ("converged", C.c_int32), ("hist_len", C.c_int32),
                ("resid", C.c_double), ("true_resid", C.c_double), ("seconds", C.c_double)]


SIGNATURES = {
    "hodlr_version": (C.c_char_p, []),
    "hodlr_ctx_create": (C.c_int, [C.c_int, C.c_void_p]),
    "hodlr_ctx_set_assembly_staging": (C.c_int, [C.c_void_p, C.c_int64]),
    "hodlr_ctx_destroy": (None, [C.c_void_p]),
    "hodlr_last_error": (C.c_char_p, [C.c_void_p]),
    "hodlr_tree_build": (C.c_int, [C.c_void_p, C.c_int, C.c_int64, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "hodlr_tree_get_info": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hodlr_tree_perm": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hodlr_tree_nodes": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "hodlr_tree_lists": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "hodlr_tree_destroy": (None, [C.c_void_p]),
    "hodlr_op_assemble": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int,
                                    C.c_void_p]),
    "hodlr_op_assemble_batch": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_double,
                                          C.c_int, C.c_void_p]),
    "hodlr_op_assemble_skeleton": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_int,
                                             C.c_int, C.c_void_p]),
    "hodlr_op_problem_offsets": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "hodlr_gmres_batch": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "hodlr_gmres_batch_device": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p]),
    "hodlr_op_tune_matvec": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "hodlr_op_set_matvec_variant": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "hodlr_op_get_matvec_variant": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "hodlr_op_set_skeleton_schedule": (C.c_int, [C.c_void_p, C.c_int]),
    "hodlr_op_get_skeleton_schedule": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "hodlr_server_create": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p]),
    "hodlr_server_destroy": (None, [C.c_void_p]),
    "hodlr_server_submit_build": (C.c_int, [C.c_void_p, C.c_int, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p]),
    "hodlr_server_submit_solve": (C.c_int, [C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p]),
    "hodlr_server_wait": (C.c_int, [C.c_void_p, C.c_int64, C.c_double, C.c_void_p, C.c_void_p]),
    "hodlr_server_get_stats": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hodlr_server_last_error": (C.c_char_p, [C.c_void_p]),
    "hodlr_dense_rows": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    "hodlr_op_save": (C.c_int, [C.c_void_p, C.c_char_p]),
    "hodlr_op_load": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_void_p]),
    "hodlr_op_get_tree": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "hodlr_op_assemble_custom": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                           C.c_double, C.c_int, C.c_void_p]),
    "hodlr_vec_save": (C.c_int, [C.c_char_p, C.c_int64, C.c_void_p]),
    "hodlr_vec_load": (C.c_int, [C.c_char_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "hodlr_op_get_stats": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hodlr_op_set_solver_graphs": (C.c_int, [C.c_void_p, C.c_int]),
    "hodlr_op_block_ranks": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    "hodlr_op_destroy": (None, [C.c_void_p]),
    "hodlr_matvec": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "hodlr_matvec_device": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    "hodlr_matvec_device_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                              C.c_void_p]),
    "hodlr_gmres": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "hodlr_gmres_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "hodlr_matvec_timing": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]),
    "hodlr_matvec_kernel_bytes": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "hodlr_op_profile": (C.c_int, [C.c_void_p, C.c_int]),
    "hodlr_op_profile_read": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "hodlr_stream_copy": (C.c_int, [C.c_void_p, C.c_int64, C.c_int, C.c_void_p]),
    "hodlr_stream_read": (C.c_int, [C.c_void_p, C.c_int64, C.c_int, C.c_void_p]),
    "hodlr_device_alloc": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p]),
    "hodlr_device_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hodlr_copy_to_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    "hodlr_copy_to_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    "hodlr_to_tree_order": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "hodlr_from_tree_order": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
}


def load_library(path=LIBPATH):
    if not os.path.exists(path):
        raise ImportError("libhodlrnd.so is not built (%s); run __graft_entry__.build()" % path)
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = load_library()


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_double), ("slots", C.c_int32), ("restart", C.c_int32),
                ("max_outer", C.c_int32), ("tol", C.c_double), ("inner_floor", C.c_double),
                ("eps_aca", C.c_double), ("nmax", C.c_int32), ("dprime", C.c_int32), ("builders", C.c_int32)]


class RequestInfo(C.Structure):
    _fields_ = [("op_id", C.c_int64), ("nrhs", C.c_int32), ("status", C.c_int32), ("converged", C.c_int32),
                ("iterations", C.c_int32), ("outer", C.c_int32), ("built", C.c_int32),
                ("max_resid", C.c_double), ("t_submit", C.c_double), ("t_start", C.c_double),
                ("t_done", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("builds", C.c_int64), ("evictions", C.c_int64), ("requests_done", C.c_int64),
                ("columns_done", C.c_int64), ("ticks", C.c_int64), ("resident_ops", C.c_int64),
                ("used_bytes", C.c_double), ("peak_bytes", C.c_double), ("budget_bytes", C.c_double),
                ("build_seconds", C.c_double), ("tick_seconds", C.c_double)]


class Context:
    def __init__(self, device=0):
        self.h = C.c_void_p()
        rc = lib.hodlr_ctx_create(device, C.byref(self.h))
        if rc != OK:
            raise HodlrError(rc, "context creation failed (device %d)" % device)

    def check(self, rc, what=""):
        if rc != OK:
            raise HodlrError(rc, "%s: %s" % (what, lib.hodlr_last_error(self.h).decode()))

    def set_assembly_staging(self, max_bytes=-1):
        """Cap on the buffer that keeps the ACA rank pass's factors for the
        fill step: -1 default (a quarter of free memory), 0 = recompute every
        block in a second ACA pass, > 0 = at most that many bytes."""
        self.check(lib.hodlr_ctx_set_assembly_staging(self.h, int(max_bytes)), "set_assembly_staging")

    def close(self):
        if self.h:
            lib.hodlr_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()


_default_ctx = None


def stream_copy_gbps(nbytes=1 << 30, reps=5, ctx=None):
    """Measured device copy bandwidth (read + write bytes per second, GB/s)."""
    ctx = ctx or default_context()
    g = C.c_double()
    ctx.check(lib.hodlr_stream_copy(ctx.h, int(nbytes), reps, C.byref(g)), "stream_copy")
    return g.value


def stream_read_gbps(nbytes=4 << 30, reps=5, ctx=None):
    """Measured device streaming-read bandwidth (GB/s)."""
    ctx = ctx or default_context()
    g = C.c_double()
    ctx.check(lib.hodlr_stream_read(ctx.h, int(nbytes), reps, C.byref(g)), "stream_read")
    return g.value


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class DeviceArray:
    """Device buffer owned by a context (hodlr_device_alloc)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        ctx.check(lib.hodlr_device_alloc(ctx.h, self.nbytes, C.byref(self.ptr)), "device_alloc")

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        self.ctx.check(lib.hodlr_copy_to_device(self.ctx.h, self.ptr, _p(a), a.nbytes), "to_device")

    def download(self, a):
        self.ctx.check(lib.hodlr_copy_to_host(self.ctx.h, _p(a), self.ptr, a.nbytes), "to_host")
        return a

    def free(self):
        if self.ptr:
            lib.hodlr_device_free(self.ctx.h, self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        self.free()


class Tree:
    """2^d tree and lists (HDD_tree.hpp:23-74, HDD_node.hpp:150-180)."""

    def __init__(self, points, x1, x2, nmax, dprime, ctx=None):
        self.ctx = ctx or default_context()
        self.points = _f64(points)
        if self.points.ndim != 2:
            raise ValueError("points must be n x ndim")
        self.n, self.ndim = self.points.shape
        lo, hi = _f64(x1), _f64(x2)
        self.h = C.c_void_p()
        self.ctx.check(lib.hodlr_tree_build(self.ctx.h, self.ndim, self.n, _p(self.points), _p(lo),
                                            _p(hi), nmax, dprime, C.byref(self.h)), "tree_build")
        self.info = TreeInfo()
        lib.hodlr_tree_get_info(self.h, C.byref(self.info))

    @classmethod
    def _borrowed(cls, ctx, handle, owner):
        """View of a tree owned by a loaded operator (not destroyed here)."""
        t = cls.__new__(cls)
        t.ctx, t.h, t._owner = ctx, handle, owner
        t.info = TreeInfo()
        lib.hodlr_tree_get_info(t.h, C.byref(t.info))
        t.n, t.ndim = t.info.npoints, t.info.ndim
        return t

    def __del__(self):
        if getattr(self, "h", None) and getattr(self, "_owner", None) is None:
            lib.hodlr_tree_destroy(self.h)
        self.h = None

    @property
    def nlevels(self):
        return self.info.nlevels

    def perm(self):
        out = np.empty(self.n, dtype=np.int64)
        lib.hodlr_tree_perm(self.h, _p(out))
        return out

    def nodes(self):
        nn = self.info.nnodes
        lo = np.empty(self.nlevels + 1, dtype=np.int64)
        st = np.empty(nn, dtype=np.int64)
        ct = np.empty(nn, dtype=np.int64)
        lib.hodlr_tree_nodes(self.h, _p(lo), _p(st), _p(ct))
        return lo, st, ct

    def lists(self, which):
        ptr = np.empty(self.info.nnodes + 1, dtype=np.int64)
        lib.hodlr_tree_lists(self.h, which, _p(ptr), None)
        idx = np.empty(max(int(ptr[-1]), 1), dtype=np.int64)
        lib.hodlr_tree_lists(self.h, which, _p(ptr), _p(idx))
        return ptr, idx[:int(ptr[-1])]


def vec_save(path, v):
    """cereal portable-binary std::vector<double> (the reference's data/*.bin)."""
    v = _f64(v).ravel()
    rc = lib.hodlr_vec_save(str(path).encode(), len(v), _p(v))
    if rc != OK:
        raise HodlrError(rc, "vec_save: %s" % path)


def vec_load(path):
    n = C.c_int64()
    rc = lib.hodlr_vec_load(str(path).encode(), 0, None, C.byref(n))
    if rc != OK:
        raise HodlrError(rc, "vec_load: %s" % path)
    v = np.empty(n.value)
    rc = lib.hodlr_vec_load(str(path).encode(), len(v), _p(v), C.byref(n))
    if rc != OK:
        raise HodlrError(rc, "vec_load: %s" % path)
    return v


class Kernel:
    """Device-evaluable kernel: A_ij = diag if i == j else scale * f(r) * w_j."""

    def __init__(self, family, ndim, scale, diag, weight=None):
        self.family, self.ndim, self.scale, self.diag = family, ndim, scale, diag
        self.weight = None if weight is None else _f64(weight)

    def desc(self):
        return KernelDesc(self.family, self.ndim, self.scale, self.diag,
                          None if self.weight is None else self.weight.ctypes.data)


# reference test kernels (user_kernel.hpp:10-179)
def kernel_2d_test():
    return Kernel(KERNEL_LOG, 2, -1.0 / (2 * math.pi), 0.0)


def kernel_3d_test():
    return Kernel(KERNEL_INV_R, 3, 1.0 / (4 * math.pi), 0.0)


def kernel_4d_test(num_points, kii):
    h = 1.0 / num_points
    return Kernel(KERNEL_EXP, 4, h ** 4 * (-1.0 / (4 * math.pi)), kii)


class Operator:
    """Assembled HODLR operator (HDD_matrix.hpp:39-47, GMRES.hpp:160-303).
    With lists of trees and kernels: the block-diagonal operator of a batch
    of independent problems (hodlr_op_assemble_batch).  sys_size (an int >= 0)
    keeps every compressed block whose column node holds more than sys_size
    points as a skeleton evaluated on the fly in each matvec
    (hodlr_op_assemble_skeleton; the reference's SYS_SIZE = 2 Nmax,
    LowRank_matrix.hpp:22-36)."""

    def __init__(self, tree, kernel, eps_aca, prec=PREC_F64 | PREC_F32, sys_size=None):
        self.h = C.c_void_p()
        self.sys_size = -1 if sys_size is None else int(sys_size)
        if self.sys_size >= 0:
            self.trees = list(tree) if isinstance(tree, (list, tuple)) else [tree]
            self.kernels = list(kernel) if isinstance(kernel, (list, tuple)) else [kernel]
            self.ctx = self.trees[0].ctx
            arr = (C.c_void_p * len(self.trees))(*[t.h.value for t in self.trees])
            kds = (KernelDesc * len(self.kernels))(*[k.desc() for k in self.kernels])
            self.ctx.check(lib.hodlr_op_assemble_skeleton(self.ctx.h, len(self.trees), arr, kds, eps_aca,
                                                          self.sys_size, prec, C.byref(self.h)),
                           "op_assemble_skeleton")
            self.tree, self.kernel = self.trees[0], self.kernels[0]
        elif isinstance(tree, (list, tuple)):
            self.trees, self.kernels = list(tree), list(kernel)
            self.ctx = self.trees[0].ctx
            arr = (C.c_void_p * len(self.trees))(*[t.h.value for t in self.trees])
            kds = (KernelDesc * len(self.kernels))(*[k.desc() for k in self.kernels])
            self.ctx.check(lib.hodlr_op_assemble_batch(self.ctx.h, len(self.trees), arr, kds, eps_aca, prec,
                                                       C.byref(self.h)), "op_assemble_batch")
            self.tree, self.kernel = self.trees[0], self.kernels[0]
        else:
            self.trees, self.kernels = [tree], [kernel]
            self.tree, self.kernel, self.ctx = tree, kernel, tree.ctx
            kd = kernel.desc()
            self.ctx.check(lib.hodlr_op_assemble(self.ctx.h, tree.h, C.byref(kd), eps_aca, prec,
                                                 C.byref(self.h)), "op_assemble")
        self.n = sum(t.n for t in self.trees)
        self.prec = prec

    def dense_rows(self, rows, x):
        """FP64 direct-sum rows of the dense operator (point ids) times x."""
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        x = _f64(x)
        y = np.empty(len(rows))
        self.ctx.check(lib.hodlr_dense_rows(self.h, len(rows), _p(rows), _p(x), _p(y)), "dense_rows")
        return y

    def save(self, path):
        """Operator file: points, kernel, ACA ranks and FP64 arenas."""
        self.ctx.check(lib.hodlr_op_save(self.h, str(path).encode()), "op_save")

    @classmethod
    def load(cls, path, ctx=None, prec=PREC_F64 | PREC_F32):
        """Operator from a file written by save (no ACA is run)."""
        ctx = ctx or default_context()
        self = cls.__new__(cls)
        self.h = C.c_void_p()
        self.ctx = ctx
        ctx.check(lib.hodlr_op_load(ctx.h, str(path).encode(), prec, C.byref(self.h)), "op_load")
        self.trees, p = [], 0
        while True:
            th = C.c_void_p()
            if lib.hodlr_op_get_tree(self.h, p, C.byref(th)) != OK:
                break
            self.trees.append(Tree._borrowed(ctx, th, self))
            p += 1
        self.kernels = [None] * len(self.trees)
        self.tree, self.kernel = self.trees[0], None
        self.n = sum(t.n for t in self.trees)
        self.prec = prec
        return self

    def offsets(self):
        out = np.zeros(len(self.trees) + 1, dtype=np.int64)
        lib.hodlr_op_problem_offsets(self.h, _p(out), len(out))
        return out

    def gmres_batch(self, B, restart=100, max_iter=None, tol=1e-8, mode=SOLVE_F64, x0="b",
                    inner_tol=0.0, max_outer=20, X0=None):
        """Every (problem, column) of B (N x nrhs, problems concatenated in
        point order) solved independently; returns X and per-instance infos."""
        B = np.asfortranarray(np.asarray(B, dtype=np.float64).reshape(self.n, -1))
        nrhs = B.shape[1]
        x0code = {"b": 1, "zero": 2, "given": 0}[x0]
        X = np.asfortranarray(B.copy() if X0 is None else np.asarray(X0, dtype=np.float64).reshape(self.n, -1))
        p = GmresParams(restart, 0 if max_iter is None else max_iter, tol, mode, x0code, inner_tol,
                        max_outer, 0, 0, None)
        infos = (GmresInfo * (len(self.trees) * nrhs))()
        rc = lib.hodlr_gmres_batch(self.h, nrhs, _p(B), _p(X), C.byref(p), infos)
        if rc not in (OK, E_NOCONV):
            self.ctx.check(rc, "gmres_batch")
        return X, [{f: getattr(i, f) for f, _ in GmresInfo._fields_} for i in infos]

    def gmres_batch_device(self, b_dev, x_dev, nrhs, restart=100, max_iter=None, tol=1e-8,
                           mode=SOLVE_F64, x0="b", inner_tol=0.0, max_outer=20):
        x0code = {"b": 1, "zero": 2, "given": 0}[x0]
        p = GmresParams(restart, 0 if max_iter is None else max_iter, tol, mode, x0code, inner_tol,
                        max_outer, 0, 0, None)
        infos = (GmresInfo * (len(self.trees) * nrhs))()
        rc = lib.hodlr_gmres_batch_device(self.h, nrhs, b_dev, x_dev, C.byref(p), infos)
        if rc not in (OK, E_NOCONV):
            self.ctx.check(rc, "gmres_batch_device")
        return [{f: getattr(i, f) for f, _ in GmresInfo._fields_} for i in infos]

    def __del__(self):
        if getattr(self, "h", None):
            lib.hodlr_op_destroy(self.h)
            self.h = None

    def stats(self):
        s = OpStats()
        lib.hodlr_op_get_stats(self.h, C.byref(s))
        return {f: getattr(s, f) for f, _ in OpStats._fields_}

    def block_ranks(self):
        nb = self.stats()["n_blocks"]
        r = np.empty(max(nb, 1), dtype=np.int32)
        d = np.empty(max(nb, 1), dtype=np.int32)
        self.ctx.check(lib.hodlr_op_block_ranks(self.h, _p(r), _p(d), max(nb, 1)), "block_ranks")
        return r[:nb], d[:nb]

    def matvec(self, x, prec=PREC_F64):
        x = np.asarray(x, dtype=np.float64)
        single = x.ndim == 1
        X = np.asfortranarray(x.reshape(self.n, -1))
        Y = np.empty_like(X, order="F")
        self.ctx.check(lib.hodlr_matvec(self.h, prec, X.shape[1], _p(X), _p(Y)), "matvec")
        return Y[:, 0].copy() if single else Y

    def gmres(self, b, x0=None, restart=100, max_iter=None, tol=1e-10, mode=SOLVE_F64,
              inner_tol=0.0, max_outer=20, verify=True, hist_cap=512, ortho=ORTHO_CGS2):
        b = _f64(b)
        x = b.copy() if x0 is None else _f64(x0).copy()
        hist = np.zeros(hist_cap)
        p = GmresParams(restart, self.n if max_iter is None else max_iter, tol, mode,
                        1 if x0 is None else 0, inner_tol, max_outer, 1 if verify else 0,
                        hist_cap, hist.ctypes.data, ortho)
        info = GmresInfo()
        rc = lib.hodlr_gmres(self.h, _p(b), _p(x), C.byref(p), C.byref(info))
        if rc not in (OK, E_NOCONV):
            self.ctx.check(rc, "gmres")
        out = {f: getattr(info, f) for f, _ in GmresInfo._fields_}
        out["hist"] = hist[:info.hist_len].copy()
        return x, out

    def gmres_device(self, b_dev, x_dev, restart=100, max_iter=None, tol=1e-10, mode=SOLVE_F64,
                     inner_tol=0.0, max_outer=20, verify=False, x0="b", ortho=ORTHO_CGS2):
        """Solve with b and x resident on the device (tree order, double).
        x0: "b" (the reference's choice), "zero", or "given" (x holds it)."""
        x0code = {"b": 1, "zero": 2, "given": 0}[x0]
        p = GmresParams(restart, self.n if max_iter is None else max_iter, tol, mode,
                        x0code, inner_tol, max_outer, 1 if verify else 0, 0, None, ortho)
        info = GmresInfo()
        rc = lib.hodlr_gmres_device(self.h, b_dev, x_dev, C.byref(p), C.byref(info))
        if rc not in (OK, E_NOCONV):
            self.ctx.check(rc, "gmres_device")
        return {f: getattr(info, f) for f, _ in GmresInfo._fields_}

    def tune(self, prec_mask=PREC_F64 | PREC_F32, reps=3):
        """Pick the fastest kernel form per matvec pass for 1..32 RHS."""
        self.ctx.check(lib.hodlr_op_tune_matvec(self.h, prec_mask, reps), "tune_matvec")

    def set_variant(self, prec, nrhs, variant):
        self.ctx.check(lib.hodlr_op_set_matvec_variant(self.h, prec, nrhs, variant), "set_variant")

    def get_variant(self, prec, nrhs):
        v = C.c_int()
        self.ctx.check(lib.hodlr_op_get_matvec_variant(self.h, prec, nrhs, C.byref(v)), "get_variant")
        return v.value

    def set_skeleton_schedule(self, mode):
        self.ctx.check(lib.hodlr_op_set_skeleton_schedule(self.h, mode), "set_skeleton_schedule")

    def skeleton_schedule(self, prec=PREC_F64, nrhs=1):
        """The schedule a (prec, nrhs) product of this operator takes."""
        v = C.c_int()
        self.ctx.check(lib.hodlr_op_get_skeleton_schedule(self.h, prec, nrhs, C.byref(v)), "get_skeleton_schedule")
        return v.value

    def set_graphs(self, enable=True):
        """Replay captured Arnoldi steps as hipGraphs in the batched solver."""
        self.ctx.check(lib.hodlr_op_set_solver_graphs(self.h, 1 if enable else 0), "set_solver_graphs")

    def matvec_timing(self, prec=PREC_F64, nrhs=1, reps=10):
        ms = np.zeros(4)
        self.ctx.check(lib.hodlr_matvec_timing(self.h, prec, nrhs, reps, _p(ms)), "matvec_timing")
        return ms

    def profile(self, enable=True):
        self.ctx.check(lib.hodlr_op_profile(self.h, 1 if enable else 0), "profile")

    def profile_read(self, reset=True):
        """Summed ms {f64 proj, f64 panel, f32 proj, f32 panel, f64 skeleton,
        f32 skeleton} and matvec counts {f64, f32} since the last reset."""
        ms = np.zeros(6)
        cnt = np.zeros(2, dtype=np.int64)
        self.ctx.check(lib.hodlr_op_profile_read(self.h, _p(ms), _p(cnt), 1 if reset else 0),
                       "profile_read")
        return ms, cnt

    def kernel_bytes(self, prec=PREC_F64, nrhs=1):
        b = np.zeros(3)
        lib.hodlr_matvec_kernel_bytes(self.h, prec, nrhs, _p(b))
        return b


class HODLRdD_matrix:
    """Reference-shaped operator object (HDD_matrix.hpp:10-89)."""

    def __init__(self, kernel, points, x1, x2, nmax=64, dprime=0, eps_aca=1e-10, ctx=None,
                 prec=PREC_F64 | PREC_F32, sys_size=None):
        """sys_size: None stores every factor; an int (the reference's
        SYS_SIZE is 2 * nmax) keeps larger blocks as on-the-fly skeletons."""
        self.kernel = kernel
        self.eps_aca = eps_aca
        self.prec = prec
        self.sys_size = sys_size
        self.tree = Tree(points, x1, x2, nmax, dprime, ctx)
        self.N = self.tree.n
        self.op = None

    def get_size(self):
        return self.N

    def Assemble_matrix_operators(self):
        self.op = Operator(self.tree, self.kernel, self.eps_aca, self.prec, sys_size=self.sys_size)

    def __mul__(self, x):
        return self.op.matvec(x)

    def solve(self, b, tol=None, mode=SOLVE_F64, **kw):
        """GMRES(100) from x0 = b to tol = eps_ACA (HDD_matrix.hpp:51-62)."""
        x, info = self.op.gmres(b, restart=100, max_iter=self.N,
                                tol=self.eps_aca if tol is None else tol, mode=mode, **kw)
        self.last_info = info
        return x

    def print_matrix_statistics(self):
        s = self.op.stats()
        print("Memory (in GB) : %g" % (8e-9 * s["n_flop_ref"]))
        print("Number of FLOP : %g" % s["n_flop_ref"])
        print("Compression Ratio : %g" % (s["n_flop_ref"] / (float(self.N) ** 2)))
        print("Maximum rank across the Tree : %d" % s["max_rank"])
        print("Time to Initialize (in s) : %g" % s["assemble_seconds"])
        return s


class Server:
    """Continuous-batching solve server (hodlr_server_*): operators built by a
    builder thread with HBM-footprint admission, solve columns batched per
    operator as FP32-inner / FP64-refined GMRES instances."""

    def __init__(self, device=0, hbm_budget=0.0, slots=8, restart=100, max_outer=20, tol=1e-8,
                 inner_floor=0.0, eps_aca=1e-8, nmax=64, dprime=0, builders=2):
        self.cfg = ServerConfig(hbm_budget, slots, restart, max_outer, tol, inner_floor, eps_aca, nmax, dprime,
                                builders)
        self.h = C.c_void_p()
        rc = lib.hodlr_server_create(device, C.byref(self.cfg), C.byref(self.h))
        if rc != OK:
            raise HodlrError(rc, "server creation failed")
        self.n_of = {}

    def _check(self, rc, what):
        if rc != OK:
            raise HodlrError(rc, "%s: %s" % (what, lib.hodlr_server_last_error(self.h).decode()))

    def submit_build(self, points, kernel, lo=None, hi=None):
        pts = _f64(points)
        n, nd = pts.shape
        lo = -np.ones(nd) if lo is None else _f64(lo)
        hi = np.ones(nd) if hi is None else _f64(hi)
        kd = kernel.desc()
        oid = C.c_int64()
        self._check(lib.hodlr_server_submit_build(self.h, nd, n, _p(pts), _p(lo), _p(hi), C.byref(kd),
                                                  C.byref(oid)), "submit_build")
        self.n_of[oid.value] = n
        return oid.value

    def submit_solve(self, op_id, B):
        if op_id not in self.n_of:
            raise HodlrError(E_ARG, "submit_solve: unknown operator id %r" % (op_id,))
        n = self.n_of[op_id]
        B = np.asfortranarray(np.asarray(B, dtype=np.float64).reshape(n, -1))
        rid = C.c_int64()
        self._check(lib.hodlr_server_submit_solve(self.h, op_id, B.shape[1], _p(B), C.byref(rid)),
                    "submit_solve")
        self._pending = getattr(self, "_pending", {})
        self._pending[rid.value] = (n, B.shape[1])
        return rid.value

    def wait(self, req_id, timeout=-1.0, want_x=True):
        n, k = self._pending[req_id]
        X = np.empty((n, k), order="F") if want_x else None
        info = RequestInfo()
        rc = lib.hodlr_server_wait(self.h, req_id, timeout, _p(X) if want_x else None, C.byref(info))
        out = {f: getattr(info, f) for f, _ in RequestInfo._fields_}
        out["rc"] = rc
        return X, out

    def stats(self):
        st = ServerStats()
        self._check(lib.hodlr_server_get_stats(self.h, C.byref(st)), "stats")
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            lib.hodlr_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()

