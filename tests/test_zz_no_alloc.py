"""GPU test of the allocation-free matvec (hodlr_matvec_device_noalloc)
against hodlr_matvec_device, through the C ABI.

Both run the same kernels; only where the intermediates live differs (the
operator's own buffers vs one fixed 64 MiB device array per GPU), so the
bar is bit-identity.  Two allocation-free calls run back to back, the
second on a smaller operator with fewer right-hand sides (the array keeps
the first call's larger contents), then a skeleton operator (on-the-fly
blocks: their slot values live in the array too); an operator whose
intermediates exceed the array is refused with HODLR_E_NOMEM.
"""
import ctypes as C
import re
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = pytest.importorskip("hodlrnd_amd")
from hodlrnd_amd import workloads as Wl  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def operator(n, seed, sys_size=None, nmax=64):
    pts = Wl.problem_points(seed, n, 3)
    T = H.Tree(pts, -np.ones(3), np.ones(3), nmax, 0)
    return T, H.Operator(T, H.Kernel(H.KERNEL_INV_R, 3, 1.0 / n, 1.0), 1e-8, sys_size=sys_size)


class Case:
    """Device input and the two outputs of one (operator, precision, nrhs)."""

    def __init__(self, A, prec, nrhs, seed):
        self.A, self.prec, self.nrhs = A, prec, nrhs
        dt = np.float64 if prec == H.PREC_F64 else np.float32
        self.x = Wl.splitmix_uniform(seed, A.n * nrhs).astype(dt)
        nb = self.x.nbytes
        self.dx = H.DeviceArray(A.ctx, nb)
        self.dy_ref = H.DeviceArray(A.ctx, nb)
        self.dy = H.DeviceArray(A.ctx, nb)
        self.dx.upload(self.x)
        self.dy_ref.upload(np.full_like(self.x, np.nan))
        self.dy.upload(np.full_like(self.x, -7.0))

    def reference(self):
        rc = H.lib.hodlr_matvec_device(self.A.h, self.prec, self.nrhs, self.dx.ptr, self.dy_ref.ptr, None)
        assert rc == H.OK, H.lib.hodlr_last_error(self.A.ctx.h)

    def noalloc(self):
        rc = H.lib.hodlr_matvec_device_noalloc(self.A.h, self.prec, self.nrhs, self.dx.ptr, self.dy.ptr, None)
        assert rc == H.OK, H.lib.hodlr_last_error(self.A.ctx.h)

    def check(self, name):
        y_ref = self.dy_ref.download(np.empty_like(self.x))
        y = self.dy.download(np.empty_like(self.x))
        say(name, "n", self.A.n, "nrhs", self.nrhs, "max |y|", float(np.abs(y_ref).max()),
            "identical", bool(np.array_equal(y, y_ref)))
        assert np.all(np.isfinite(y_ref))
        assert np.array_equal(y, y_ref)


def test_noalloc_matvec_equals_matvec_device():
    say("case noalloc_two_calls")
    T1, A1 = operator(6000, 1)
    T2, A2 = operator(2500, 2)
    big = Case(A1, H.PREC_F64, 8, 11)
    small = Case(A2, H.PREC_F32, 2, 12)
    big.reference()
    small.reference()
    big.noalloc()    # first call: the larger operator, 8 RHS
    small.noalloc()  # second call, smaller input, right after it
    big.check("first call (f64)")
    small.check("second call (f32)")
    # the same two products again in the other order and precisions
    b2, s2 = Case(A1, H.PREC_F32, 4, 13), Case(A2, H.PREC_F64, 1, 14)
    b2.reference()
    s2.reference()
    b2.noalloc()
    s2.noalloc()
    b2.check("third call (f32)")
    s2.check("fourth call (f64)")


def test_noalloc_skeleton_operator_and_refusal():
    say("case noalloc_skeleton")
    T3, A3 = operator(4000, 3, sys_size=64, nmax=32)
    assert A3.stats()["n_skeleton"] > 0
    for prec, k, seed in ((H.PREC_F64, 2, 21), (H.PREC_F32, 1, 22)):
        c = Case(A3, prec, k, seed)
        c.reference()
        c.noalloc()
        c.check("skeleton operator")
    say("case noalloc_refusal")
    T4, A4 = operator(60000, 4)
    c = Case(A4, H.PREC_F32, 32, 23)
    rc = H.lib.hodlr_matvec_device_noalloc(A4.h, H.PREC_F32, 32, c.dx.ptr, c.dy.ptr, None)
    msg = H.lib.hodlr_last_error(A4.ctx.h).decode()
    say("refusal rc", rc, msg)
    assert rc == H.E_NOMEM and "allocation-free" in msg
    need = int(re.search(r"needs (\d+) bytes", msg).group(1))
    assert need > 64 << 20
    c.reference()  # the allocating entry point still serves it
    assert np.all(np.isfinite(c.dy_ref.download(np.empty_like(c.x))))
    # a refused call left the array usable: a small product right after it
    d = Case(A3, H.PREC_F64, 1, 24)
    d.reference()
    d.noalloc()
    d.check("after refusal")
    # argument errors
    assert H.lib.hodlr_matvec_device_noalloc(A3.h, H.PREC_F64, 3, d.dx.ptr, d.dy.ptr, None) == H.E_ARG
    assert H.lib.hodlr_matvec_device_noalloc(None, H.PREC_F64, 1, d.dx.ptr, d.dy.ptr, None) == H.E_ARG
