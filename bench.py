#!/usr/bin/env python3
"""Headline benchmark: solves/s and p50/p99 latency at 1e-8 relative residual,
FP32-inner / FP64-refined versus plain FP64 HODLR GMRES, with the fast
matvec's achieved HBM bandwidth.

Workload (one MI355X): NDIM = 3, N = 1,000,000 points on the 100^3 uniform
grid of [-1,1]^3 (uniform_nodes, dimension 0 outermost), second-kind kernel
A_ij = delta_ij + (1/N) / |x_i - x_j|, Nmax = 128, eps_ACA = 1e-8, d' = 0,
one uniform [-1,1] right-hand side (SplitMix64), GMRES(100) from x0 = b to a
relative residual of 1e-8.  A step is one solve with the compressed operator
resident in HBM and b already on the device.

The line also carries:
  roofline      the matvec kernel with the most time in the timed mixed
                solves; achieved = its algorithmic bytes per launch / its
                mean launch time from HIP events recorded on the solver's
                stream during the timed region;
  cpu_baseline  the FP64 CPU restatement (oracle/) in the reference's
                execution model (one thread; skeleton blocks for |cols| >
                2 Nmax re-evaluate entries per matvec) on a stratified sample
                of the same operator's row nodes, scaled to one solve.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/s + p99 latency at 1e-8 residual, mixed vs FP64 HODLR; matvec HBM GB/s"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_points(n1d, a=-1.0, b=1.0):
    """uniform_nodes(a, b, n) (myHeaders.hpp:111-120) on every axis, tensor
    grid with dimension 0 outermost (user_kernel.hpp:79-95)."""
    nodes = a + (b - a) / float(n1d - 1) * np.arange(n1d, dtype=np.float64)
    g = np.meshgrid(nodes, nodes, nodes, indexing="ij")
    return np.ascontiguousarray(np.stack([c.ravel() for c in g], axis=1))


def splitmix_uniform(seed, n):
    """SplitMix64 -> u = (z >> 11) * 2^-53 -> 2u - 1 (same stream as the tests)."""
    out = np.empty(n, dtype=np.float64)
    s = np.uint64(seed)
    gold = np.uint64(0x9E3779B97F4A7C15)
    m1, m2 = np.uint64(0xBF58476D1CE4E5B9), np.uint64(0x94D049BB133111EB)
    idx = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = s + idx * gold
        z = (z ^ (z >> np.uint64(30))) * m1
        z = (z ^ (z >> np.uint64(27))) * m2
        z = z ^ (z >> np.uint64(31))
    out[:] = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * out - 1.0


def measured_traffic(kernel, alg_bytes):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE passes scaled by the copy-kernel factors),
    used only if it was measured on this exact operator (same algorithmic
    bytes per launch); otherwise None."""
    path = os.path.join(ROOT, "profiles", "matvec_traffic_grid100_N1e6.json")
    try:
        d = json.load(open(path))
        e = d["kernels"][kernel]
        if abs(e["algorithmic_bytes_per_launch"] - alg_bytes) <= 1e-3 * alg_bytes:
            return float(e["hbm_bytes_per_launch"]), "profiles/matvec_traffic_grid100_N1e6.json"
    except (OSError, KeyError, ValueError):
        pass
    return None, None


def cpu_baseline(pts, n_mv, nmax, eps, budget_s=20.0):
    """Times the FP64 CPU restatement on a stratified sample of row nodes."""
    from oracle import pyoracle as O
    n = len(pts)
    t0 = time.time()
    T = O.Tree(pts, -np.ones(3), np.ones(3), nmax, 0)
    k = O.KernelDesc(O.K_INV, 3, 1.0 / n, 1.0)
    L = T.nlevels
    want = {1: 1, 2: 2, 3: 4, 4: 16, 5: 64}
    est_mv, est_asm, sampled = 0.0, 0.0, []
    x = splitmix_uniform(2, n)
    for l in range(1, L):
        ns = T.level_size(l)
        w = min(want.get(l, 64), ns)
        stride = np.zeros(L, dtype=np.int64)
        off = np.zeros(L, dtype=np.int64)
        stride[l] = max(1, ns // w)
        off[l] = stride[l] // 2
        cnt = len(range(int(off[l]), ns, int(stride[l])))
        ta = time.time()
        op = O.Operator(T, k, eps, sys_size=2 * nmax, nthreads=1, sample=(stride, off))
        tb = time.time()
        reps, tm = 0, 0.0
        while reps < 1 or (tm < 0.3 and reps < 20):
            t1 = time.time()
            op.matvec(x, nthreads=1)
            tm += time.time() - t1
            reps += 1
        scale = ns / cnt
        est_mv += tm / reps * scale
        est_asm += (tb - ta) * scale
        sampled.append("L%d:%d/%d" % (l, cnt, ns))
        del op
        if time.time() - t0 > budget_s:
            break
    solve_s = est_mv * n_mv
    return {
        "value": 1.0 / solve_s,
        "unit": "solves/s",
        "cores": 1,
        "kind": "port",
        "sample": ("oracle FP64 restatement, 1 thread, reference execution model (skeleton blocks for "
                   "|cols| > 2*Nmax); row nodes sampled per level %s, matvec and assembly times scaled "
                   "by nodes/sampled; solve = est. matvec %.3f s x %d matvecs of the FP64 GPU solve"
                   % (",".join(sampled), est_mv, n_mv)),
        "est_matvec_s": est_mv,
        "est_assembly_s": est_asm,
        "wall_s": time.time() - t0,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n1d", type=int, default=100)
    ap.add_argument("--nmax", type=int, default=128)
    ap.add_argument("--eps", type=float, default=1e-8)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--inner-tol", type=float, default=0.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the library's plain copy kernel (1 GiB, 3 launches): the "
                         "known-size kernel for scaling FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every box has one "
              "MI355X); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import hodlrnd_amd as H

    if args.copy_calibration:
        g = H.stream_copy_gbps(1 << 30, 3)
        print(json.dumps({"copy_kernel": "hodlr::stream_copy_kernel", "bytes_read_per_launch": 1 << 30,
                          "bytes_written_per_launch": 1 << 30, "launches": 4, "gbps": g}))
        return

    copy_gbs = H.stream_copy_gbps(1 << 30, 5)
    n1 = args.n1d
    pts = make_points(n1)
    n = len(pts)
    kern = H.Kernel(H.KERNEL_INV_R, 3, 1.0 / n, 1.0)
    t0 = time.time()
    tree = H.Tree(pts, -np.ones(3), np.ones(3), args.nmax, 0)
    t_tree = time.time() - t0
    log("tree: N=%d levels=%d admissible=%d (%.2fs)" % (n, tree.nlevels, tree.info.n_admissible, t_tree))
    t0 = time.time()
    op = H.Operator(tree, kern, args.eps, H.PREC_F64 | H.PREC_F32)
    t_asm = time.time() - t0
    st = op.stats()
    log("assembled in %.2fs: e_near %.3g e_far %.3g, %.1f GB FP64 + %.1f GB FP32"
        % (t_asm, st["e_near"], st["e_far"], st["bytes_f64"] * 1e-9, st["bytes_f32"] * 1e-9))

    ctx = tree.ctx
    perm = tree.perm()
    b = splitmix_uniform(1, n)
    bt = np.ascontiguousarray(b[perm])
    db = H.DeviceArray(ctx, 8 * n)
    dx = H.DeviceArray(ctx, 8 * n)
    db.upload(bt)

    def run(mode, k):
        # FP64: the reference's solve (x0 = b, HDD_matrix.hpp:52-53); mixed:
        # iterative refinement from x0 = 0, whose first residual is b exactly
        x0 = "b" if mode == H.SOLVE_F64 else "zero"
        lat, infos = [], []
        for _ in range(k):
            t1 = time.perf_counter()
            info = op.gmres_device(db.ptr, dx.ptr, restart=100, max_iter=n, tol=args.tol, mode=mode,
                                   inner_tol=args.inner_tol, x0=x0)
            lat.append(time.perf_counter() - t1)
            infos.append(info)
        return np.array(lat), infos

    run(H.SOLVE_F64, args.warmup)
    lat64, inf64 = run(H.SOLVE_F64, args.steps)
    log("fp64: %.2f ms/solve, iterations %d" % (lat64.mean() * 1e3, inf64[-1]["iterations"]))
    run(H.SOLVE_MIXED, args.warmup)
    op.profile(True)
    op.profile_read(reset=True)
    latm, infm = run(H.SOLVE_MIXED, args.steps)
    ms, cnt = op.profile_read(reset=True)
    op.profile(False)
    log("mixed: %.2f ms/solve, iterations %d, outer %d" % (latm.mean() * 1e3, infm[-1]["iterations"],
                                                          infm[-1]["outer"]))

    # accuracy of the last mixed solution against the FP64 operator
    xt = dx.download(np.empty(n))
    x = np.empty(n)
    x[perm] = xt
    r = b - op.matvec(x, prec=H.PREC_F64)
    true_res = float(np.linalg.norm(r) / np.linalg.norm(b))

    # dominant kernel of the timed mixed solves
    kb64 = op.kernel_bytes(H.PREC_F64, 1)
    kb32 = op.kernel_bytes(H.PREC_F32, 1)
    cands = [("matvec_proj_f64", ms[0], cnt[0], kb64[1]), ("matvec_panel_f64", ms[1], cnt[0], kb64[2]),
             ("matvec_proj_f32", ms[2], cnt[1], kb32[1]), ("matvec_panel_f32", ms[3], cnt[1], kb32[2])]
    name, tot_ms, c, bytes_ = max(cands, key=lambda t: t[1])
    avg_ms = tot_ms / max(c, 1)
    achieved = bytes_ / (avg_ms * 1e-3) * 1e-9
    traffic, traffic_src = measured_traffic(name, bytes_)
    mv32 = (ms[2] + ms[3]) / max(cnt[1], 1)
    mv64 = (ms[0] + ms[1]) / max(cnt[0], 1)

    n_mv64 = inf64[-1]["matvecs_f64"]
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle sample)...")
        cpu = cpu_baseline(pts, n_mv64, args.nmax, args.eps)
        log("cpu baseline: %.4g solves/s (%.1fs of CPU work)" % (cpu["value"], cpu["wall_s"]))

    val = args.steps / latm.sum()
    out = {
        "metric": METRIC,
        "value": val,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": latm.mean() * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 matvec + fp64 refinement",
        "data": "synthetic",
        "config": {
            "workload": "3D %d^3 uniform grid, N=%d, A=I+(1/N)/r, Nmax=%d, eps_ACA=%g, d'=0, "
                        "GMRES(100) to rel. residual %g, 1 RHS (FP64: x0=b; mixed: x0=0)"
                        % (n1, n, args.nmax, args.eps, args.tol),
            "n": n, "nmax": args.nmax, "eps_aca": args.eps, "dprime": 0, "tol": args.tol,
        },
        "p50_ms": float(np.percentile(latm, 50) * 1e3),
        "p99_ms": float(np.percentile(latm, 99) * 1e3),
        "fp64_solves_per_s": args.steps / lat64.sum(),
        "fp64_p50_ms": float(np.percentile(lat64, 50) * 1e3),
        "fp64_p99_ms": float(np.percentile(lat64, 99) * 1e3),
        "speedup_mixed_vs_fp64": (args.steps / latm.sum()) / (args.steps / lat64.sum()),
        "iterations_fp64": inf64[-1]["iterations"],
        "iterations_mixed_inner": infm[-1]["iterations"],
        "refinement_sweeps": infm[-1]["outer"],
        "matvecs_mixed": {"f32": infm[-1]["matvecs_f32"], "f64": infm[-1]["matvecs_f64"]},
        "true_resid_mixed": true_res,
        "resid_fp64_estimate": inf64[-1]["resid"],
        "matvec_ms": {"f64": mv64, "f32": mv32},
        "matvec_hbm_gbs": {"f64": kb64[0] / (mv64 * 1e-3) * 1e-9 if mv64 else None,
                           "f32": kb32[0] / (mv32 * 1e-3) * 1e-9 if mv32 else None},
        "assemble_s": t_asm,
        "aca_rank_pass_s": st["rank_seconds"],
        "fill_s": st["fill_seconds"],
        "operator_gb": {"f64": st["bytes_f64"] * 1e-9, "f32": st["bytes_f32"] * 1e-9},
        "roofline": {
            "bound": "hbm",
            "kernel": name,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "measured_copy_gbs": copy_gbs,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "algorithmic_bytes_per_launch": bytes_,
            "avg_launch_ms": avg_ms,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
