"""CPU tests of libhodlrnd's host side through the C ABI (no GPU needed).

A host-only context (device -1) builds trees and lists; these must match the
FP64 oracle bit-exactly (ids, point order, list order).  Everything that
needs the GPU must fail with HODLR_E_STATE, never fall back to the CPU.
"""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import hodlrnd_amd as H
from oracle import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "hodlrnd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(hodlr_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 25
    lib = C.CDLL(H.LIBPATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding knows every one of them
    assert set(names) <= set(H.SIGNATURES), set(names) - set(H.SIGNATURES)


def test_version_string():
    assert b"gfx950" in H.lib.hodlr_version()


@pytest.fixture(scope="module")
def hctx():
    return H.Context(-1)


CASES = [(3, 10000, 64, 0), (2, 5000, 32, 0), (2, 3000, 16, -1), (4, 10000, 64, 0),
         (3, 8000, 32, 1), (4, 4096, 16, 3), (3, 8000, 32, 2), (3, 8000, 32, -1), (1, 500, 8, 0)]


@pytest.mark.parametrize("nd,n,nmax,dprime", CASES)
def test_tree_bit_exact_vs_oracle(hctx, nd, n, nmax, dprime):
    pts = O.splitmix_uniform(nd * 1000 + n, n * nd).reshape(n, nd)
    lo, hi = -np.ones(nd), np.ones(nd)
    T = H.Tree(pts, lo, hi, nmax, dprime, hctx)
    R = O.Tree(pts, lo, hi, nmax, dprime)
    assert T.nlevels == R.nlevels
    perm = T.perm()
    assert np.array_equal(np.sort(perm), np.arange(n))
    lo_, st, ct = T.nodes()
    nb, ib = T.lists(0)
    ni, ii = T.lists(1)
    for l in range(T.nlevels):
        for j in range(R.level_size(l)):
            g = lo_[l] + j
            seg = perm[st[g]:st[g] + ct[g]]
            ref = R.points(l, j)
            assert np.array_equal(np.sort(seg), ref)
            if l == T.nlevels - 1:
                assert np.array_equal(seg, ref)  # leaves keep ascending ids
            assert np.array_equal(ib[nb[g]:nb[g + 1]], R.neighbours(l, j))
            assert np.array_equal(ii[ni[g]:ni[g + 1]], R.interactions(l, j))


def test_tree_grid_ties_and_empty_boxes(hctx):
    """Tensor grids put points exactly on split planes (x < mid goes low) and
    Gauss-Legendre grids leave boxes empty: both as the reference."""
    for pts, nmax, dp in [(O.tensor_grid(3, O.uniform_nodes(-1, 1, 17)), 32, 0),
                          (O.tensor_grid(4, O.gauss_legendre(8)[0]), 16, 0),
                          (O.tensor_grid(2, O.uniform_nodes(-1, 1, 40)), 16, -1)]:
        nd = pts.shape[1]
        T = H.Tree(pts, -np.ones(nd), np.ones(nd), nmax, dp, hctx)
        R = O.Tree(pts, -np.ones(nd), np.ones(nd), nmax, dp)
        lo_, st, ct = T.nodes()
        L = T.nlevels - 1
        assert np.array_equal(ct[lo_[L]:lo_[L + 1]], [R.npts(L, j) for j in range(R.level_size(L))])
        ni, ii = T.lists(1)
        for j in range(R.level_size(L)):
            g = lo_[L] + j
            assert np.array_equal(ii[ni[g]:ni[g + 1]], R.interactions(L, j))


def test_depth_zero_tree(hctx):
    pts = O.splitmix_uniform(1, 90).reshape(30, 3)
    T = H.Tree(pts, -np.ones(3), np.ones(3), 64, 0, hctx)
    assert T.nlevels == 1 and T.info.n_admissible == 0


def test_host_only_context_refuses_device_work(hctx):
    pts = O.splitmix_uniform(2, 600).reshape(200, 3)
    T = H.Tree(pts, -np.ones(3), np.ones(3), 32, 0, hctx)
    k = H.Kernel(H.KERNEL_INV_R, 3, 1.0 / 200, 1.0)
    with pytest.raises(H.HodlrError) as e:
        H.Operator(T, k, 1e-8)
    assert e.value.code == H.E_STATE


def test_argument_errors(hctx):
    pts = np.zeros((4, 9))
    with pytest.raises(H.HodlrError) as e:
        H.Tree(pts, -np.ones(9), np.ones(9), 8, 0, hctx)
    assert e.value.code == H.E_ARG
    with pytest.raises(H.HodlrError) as e:
        H.Tree(np.zeros((4, 2)), -np.ones(2), np.ones(2), 0, 0, hctx)
    assert e.value.code == H.E_ARG
    pts = O.splitmix_uniform(3, 300).reshape(100, 3)
    T = H.Tree(pts, -np.ones(3), np.ones(3), 32, 0, hctx)
    kd = H.KernelDesc(H.KERNEL_INV_R, 2, 1.0, 1.0, None)  # ndim mismatch
    h = C.c_void_p()
    assert H.lib.hodlr_op_assemble(hctx.h, T.h, C.byref(kd), 1e-8, 3, C.byref(h)) == H.E_ARG
    kd = H.KernelDesc(7, 3, 1.0, 1.0, None)  # unknown family
    assert H.lib.hodlr_op_assemble(hctx.h, T.h, C.byref(kd), 1e-8, 3, C.byref(h)) == H.E_ARG
    kd = H.KernelDesc(H.KERNEL_INV_R, 3, 1.0, 1.0, None)
    assert H.lib.hodlr_op_assemble(hctx.h, T.h, C.byref(kd), 2.0, 3, C.byref(h)) == H.E_ARG
    assert H.lib.hodlr_op_assemble(None, T.h, C.byref(kd), 1e-8, 3, C.byref(h)) == H.E_ARG


def test_tree_order_round_trip(hctx):
    pts = O.splitmix_uniform(4, 1500).reshape(500, 3)
    T = H.Tree(pts, -np.ones(3), np.ones(3), 16, 0, hctx)
    X = np.asfortranarray(np.stack([O.splitmix_uniform(9 + c, 500) for c in range(3)], 1))
    tree_order = np.empty(500 * 3)
    back = np.empty_like(X, order="F")
    assert H.lib.hodlr_to_tree_order(T.h, 3, H._p(X), H._p(tree_order)) == 0
    assert H.lib.hodlr_from_tree_order(T.h, 3, H._p(tree_order), H._p(back)) == 0
    assert np.array_equal(back, X)
    perm = T.perm()
    assert np.array_equal(tree_order.reshape(500, 3)[:, 1], X[perm, 1])


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_cpp_api_example_compiles(tmp_path):
    """The reference-shaped C++ API (include/hodlrnd/HODLRnD.hpp) builds with
    a plain host compiler against the C ABI."""
    out = tmp_path / "main_4d"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "hodlrnd_amd", "examples", "main_4d.cpp"), "-o", str(out),
                        "-L" + os.path.join(ROOT, "hodlrnd_amd"), "-lhodlrnd"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]


def test_server_config_validation():
    """hodlr_server_create checks its configuration before touching a GPU."""
    import ctypes as C
    bad = [dict(slots=3), dict(restart=0), dict(restart=200), dict(tol=0.0), dict(eps_aca=1.5), dict(nmax=0),
           dict(max_outer=0), dict(builders=0), dict(builders=17)]
    for b in bad:
        kw = dict(hbm_budget=0.0, slots=8, restart=100, max_outer=20, tol=1e-8, inner_floor=0.0, eps_aca=1e-8,
                  nmax=64, dprime=0, builders=2)
        kw.update(b)
        cfg = H.ServerConfig(**kw)
        h = C.c_void_p()
        assert H.lib.hodlr_server_create(0, C.byref(cfg), C.byref(h)) == H.E_ARG, b
    assert H.lib.hodlr_server_create(-1, C.byref(H.ServerConfig()), None) == H.E_ARG


def test_cereal_vector_files(tmp_path, golden_dir):
    """hodlr_vec_load reads the reference's fixtures (cereal portable binary,
    data/x_10000.bin) exactly as the oracle's parser; save/load round trip."""
    import os
    x = H.vec_load(os.path.join(golden_dir, "x_10000.bin"))
    assert np.array_equal(x, O.read_cereal_vec(os.path.join(golden_dir, "x_10000.bin")))
    p = tmp_path / "v.bin"
    v = O.splitmix_uniform(3, 777)
    H.vec_save(p, v)
    raw = p.read_bytes()
    assert raw[0] == 1 and int.from_bytes(raw[1:9], "little") == 777 and len(raw) == 9 + 8 * 777
    assert np.array_equal(H.vec_load(p), v)
    assert np.array_equal(O.read_cereal_vec(str(p)), v)
    p.write_bytes(b"\x00" + raw[1:])
    with pytest.raises(H.HodlrError):
        H.vec_load(p)


def test_operator_file_rejects_bad_files(hctx, tmp_path):
    """hodlr_op_load checks magic, version and length before any device work."""
    import ctypes as C
    h = C.c_void_p()
    bad = tmp_path / "bad.op"
    bad.write_bytes(b"NOTANOPERATORFILE")
    assert H.lib.hodlr_op_load(hctx.h, str(bad).encode(), 3, C.byref(h)) == H.E_ARG
    trunc = tmp_path / "trunc.op"
    trunc.write_bytes(b"HODLROP1\x01\x00")
    assert H.lib.hodlr_op_load(hctx.h, str(trunc).encode(), 3, C.byref(h)) == H.E_ARG
    assert H.lib.hodlr_op_load(hctx.h, str(tmp_path / "missing.op").encode(), 3, C.byref(h)) == H.E_ARG


def test_reference_shaped_user_kernels_identified(tmp_path):
    """Kernel classes in the reference's shape (user_kernel.hpp:10-179: Eigen
    vector names, Kernel_Fun, getMatrixEntry, no device description) compile
    unchanged against include/hodlrnd/HODLRnD.hpp with g++; their entries are
    identified from getMatrixEntry (a uniform 1/(4 pi r), per-column weights
    w_j / r), a Yukawa kernel exp(-r)/r is refused with a message, and the
    generic iterSolver<Op> (GMRES.hpp:31-303) solves a dense host operator
    (restart 100 and 5) and exits at once on A = I with x0 = b."""
    import math
    import re
    out = tmp_path / "user_kernels"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "hodlrnd_amd", "examples", "user_kernels.cpp"), "-o", str(out),
                        "-L" + os.path.join(ROOT, "hodlrnd_amd"), "-lhodlrnd",
                        "-Wl,-rpath," + os.path.join(ROOT, "hodlrnd_amd")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([str(out), "identify"], capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    o = r.stdout
    m = re.search(r"identified green_3d family=inv_r scale=(\S+) diag=(\S+) weights=0", o)
    assert m and abs(float(m.group(1)) - 1 / (4 * math.pi)) < 1e-16 and float(m.group(2)) == 0.0
    assert "identified weighted_second_kind family=inv_r scale=1 diag=1 weights=700" in o
    assert re.search(r"refused yukawa_3d code=-1: kernel entries are not", o)
    for rs in (100, 5):
        g = re.search(r"generic gmres restart=%d iterations=(\d+) resid=(\S+) error=(\S+)" % rs, o)
        assert g and 0 < int(g.group(1)) < 343 and float(g.group(2)) < 1e-12 and float(g.group(3)) < 1e-10
    assert "generic gmres identity iterations=0" in o
    log = (tmp_path / "gmres_generic_r5.txt").read_text()
    assert log.startswith("Restarted GMRES Parameters") and "Iter[6]" in log
    assert "Reached Solution before Max_Iterations" in log and "Mat-Vec Time :" in log
    assert "Exited before any Progress" in (tmp_path / "gmres_generic_exact.txt").read_text()


def test_custom_entry_checks_its_launcher_table(hctx):
    """hodlr_op_assemble_custom refuses another kernel family, a launcher
    table of another layout or with a null launcher, before any device work."""
    import ctypes as C

    class Table(C.Structure):
        _fields_ = [("version", C.c_uint32), ("layout", C.c_uint32)] + [(n, C.c_void_p) for n in
                                                                         ("aca", "dense_fill", "near_fill",
                                                                          "dense_rows")]

    pts = O.splitmix_uniform(5, 300 * 3).reshape(300, 3)
    T = H.Tree(pts, -np.ones(3), np.ones(3), 32, 0, hctx)
    h = C.c_void_p()
    obj = (C.c_double * 2)(1.0, 2.0)
    bad = Table(0, 0, 1, 1, 1, 1)

    def call(family, table, nbytes=16):
        kd = H.KernelDesc(family, 3, 1.0, 1.0, None)
        return H.lib.hodlr_op_assemble_custom(hctx.h, T.h, C.byref(kd), obj, nbytes, C.byref(table), 1e-8, 3,
                                              C.byref(h))

    assert call(H.KERNEL_INV_R, bad) == H.E_ARG
    assert "HODLR_KERNEL_CUSTOM" in H.lib.hodlr_last_error(hctx.h).decode()
    assert call(H.KERNEL_CUSTOM, bad) == H.E_ARG
    assert "another layout" in H.lib.hodlr_last_error(hctx.h).decode()
    assert call(H.KERNEL_CUSTOM, bad, nbytes=-1) == H.E_ARG
    assert H.lib.hodlr_op_assemble_custom(hctx.h, T.h, None, obj, 16, C.byref(bad), 1e-8, 3, C.byref(h)) == H.E_ARG


def test_device_kernel_classes_compile_with_hipcc(tmp_path):
    """Kernel classes in the reference's shape between
    hodlrnd/device_kernels_begin.hpp and _end.hpp compile unchanged as host
    and device code (hipcc, -Werror), the library's entry kernels
    instantiated for each Kernel_Fun; a class with a host-only Kernel_Fun
    used in such a program is a compile error, not a silent host fallback."""
    src = os.path.join(ROOT, "hodlrnd_amd", "examples", "device_kernels.cpp")
    flags = ["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-arch=gfx950", "-O1", "-std=c++17", "-ffp-contract=off",
             "-Wall", "-Werror", "-Wno-unused-command-line-argument", "-I" + os.path.join(ROOT, "include")]
    r = subprocess.run(flags + ["-c", src, "-o", str(tmp_path / "dk.o")], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    bad = tmp_path / "host_only.cpp"
    bad.write_text('#include "hodlrnd/device_kernels_begin.hpp"\n#include "hodlrnd/device_kernels_end.hpp"\n'
                   "struct host_only {\n  std::vector<ptsnD> *gridPoints;\n"
                   "  double Kernel_Fun(double x) { return std::exp(-x) / x; }\n"
                   "  double getMatrixEntry(int i, int j) { return i == j ? 1.0 : 0.5; }\n};\n"
                   "int main() { std::vector<ptsnD> *p = new std::vector<ptsnD>(4); host_only *k = new host_only;\n"
                   "  k->gridPoints = p; Vec a(3), b(3); HODLRdD_matrix<host_only> M(k, p, a, b);\n"
                   "  M.Assemble_matrix_operators(); }\n")
    r = subprocess.run(flags + ["-c", str(bad), "-o", str(tmp_path / "bad.o")], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode != 0 and "Kernel_Fun" in r.stderr


def test_tree_build_refuses_runaway_inputs(hctx):
    """Points outside the bounding box, an empty or non-finite box, and more
    than nmax coincident points would make the uniform-depth split loop
    (HDD_tree.hpp:43-69) deepen without end: each is refused with E_ARG."""
    pts = O.splitmix_uniform(9, 300 * 3).reshape(300, 3)
    lo, hi = -np.ones(3), np.ones(3)
    bad = pts.copy()
    bad[17, 2] = 1.5
    for p, a, b, msg in [(bad, lo, hi, "point 17 lies outside"),
                         (pts, lo, lo, "finite lo < hi"),
                         (pts, lo, np.array([1.0, np.inf, 1.0]), "finite lo < hi"),
                         (np.vstack([pts, np.zeros((40, 3))]), lo, hi, "coincident points")]:
        with pytest.raises(H.HodlrError) as e:
            H.Tree(p, a, b, 32, 0, hctx)
        assert e.value.code == H.E_ARG and msg in str(e.value), str(e.value)
    nan = pts.copy()
    nan[3, 0] = np.nan
    with pytest.raises(H.HodlrError):
        H.Tree(nan, lo, hi, 32, 0, hctx)
