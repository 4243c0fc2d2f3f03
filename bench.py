#!/usr/bin/env python3
"""Headline benchmark: solves/s and p50/p99 latency at 1e-8 relative residual,
FP32-inner / FP64-refined versus plain FP64 HODLR GMRES, with the fast
matvec's achieved HBM bandwidth.

Workload (one MI355X): NDIM = 3, N = 1,000,000 points on the 100^3 uniform
grid of [-1,1]^3 (uniform_nodes, dimension 0 outermost), second-kind kernel
A_ij = delta_ij + (1/N) / |x_i - x_j|, Nmax = 128, eps_ACA = 1e-8, d' = 0,
one uniform [-1,1] right-hand side (SplitMix64), GMRES(100) from x0 = b to a
relative residual of 1e-8.  The operator is the reference's default one:
every compressed block with more than SYS_SIZE = 2 Nmax = 256 columns keeps
only its pivots and core pseudo-inverse and has its entries evaluated on the
fly in every matvec (myHeaders.hpp:48, LowRank_matrix.hpp:22-36); the same
solve with every factor stored is the stored_factors leg.  A step is one
solve with the compressed operator resident in HBM and b already on the
device.

The line also carries:
  roofline      the HBM-streaming matvec kernel with the most time in the
                timed mixed solves; achieved = its algorithmic bytes per
                launch / its mean launch time from HIP events recorded on the
                solver's stream during the timed region;
  skeleton      the on-the-fly phase of the skeleton blocks: the schedule the
                solves used (on a second stream beside the stored passes,
                compute-bound beside memory-bound work), the whole matvec in
                it and on one stream, and the phase alone on one stream: its
                entries/s and fraction of the FP64 vector-ALU issue peak;
  cpu_baseline  the FP64 CPU restatement (oracle/) in the reference's
                execution model (skeleton blocks for |cols| > 2 Nmax
                re-evaluate entries per matvec) on the host cores: the whole
                operator assembled and one full matvec timed, one solve =
                that matvec x the FP64 GPU solve's matvec count;
                cpu_baseline_1t is the same on one thread, estimated from a
                stratified sample of the operator's row nodes.
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
# FP64 vector issue rate of MI355X: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
# (78.6 TFLOP/s counts each of these FMAs as 2 flops)
VALU_F64_LANE_OPS = 256 * 4 * 16 * 2.4e9
# issue slots per on-the-fly entry of the 3-D 1/r skeleton kernels: FP64
# operator 3 v_add_f64 + 3 v_fma_f64 (distance from differences), FP32
# operator 1 v_add_f64 + 3 v_fma_f64 (distance from centred coordinates and
# norms); v_rsq_f64 (a quarter-rate transcendental: 4 slots); 5 for the
# refinement (cubic step, FP64 operator) or 4 (Newton step, FP32 operator);
# 1 FMA per RHS
SKEL_SLOTS = {"f64": 16, "f32": 13}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_points(n1d):
    """uniform_nodes(-1, 1, n1d) (myHeaders.hpp:111-120) on every axis, tensor
    grid with dimension 0 outermost (user_kernel.hpp:79-95)."""
    from hodlrnd_amd import workloads as Wl
    return Wl.grid_points(n1d)


def splitmix_uniform(seed, n):
    from hodlrnd_amd import workloads as Wl
    return Wl.splitmix_uniform(seed, n)


def measured_traffic(kernel, alg_bytes, name="matvec_traffic_grid100_N1e6.json", kernel_name=None):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE passes scaled by the copy-kernel factors),
    used only if it was measured on this exact operator (same algorithmic
    bytes per launch) and, when kernel_name is given, with the same kernel
    form; otherwise None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        d = json.load(open(path))
        e = d["kernels"][kernel]
        if kernel_name is not None and e["kernel"] != kernel_name:
            return None, None
        if abs(e["algorithmic_bytes_per_launch"] - alg_bytes) <= 1e-3 * alg_bytes:
            return float(e["hbm_bytes_per_launch"]), "profiles/" + name
    except (OSError, KeyError, ValueError):
        pass
    return None, None


def host_cores():
    """Cores this process may use (capped at the 16 of one GPU's share) and
    the host CPU model, read at run time."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return max(1, min(n, 16)), model


def cpu_baseline_full(pts, n_mv, nmax, eps):
    """The FP64 CPU restatement of the whole headline operator on the host
    cores (OpenMP): tree, assembly in the reference's execution model
    (skeleton blocks for |cols| > 2 Nmax, LowRank_matrix.hpp:22-36) and one
    full matvec, each timed; a solve is the timed matvec times the FP64 GPU
    solve's matvec count (GMRES's vector work is a small part on the CPU)."""
    from oracle import pyoracle as O
    n = len(pts)
    nthreads, model = host_cores()
    t0 = time.time()
    T = O.Tree(pts, -np.ones(3), np.ones(3), nmax, 0)
    op = O.Operator(T, O.KernelDesc(O.K_INV, 3, 1.0 / n, 1.0), eps, sys_size=2 * nmax, nthreads=nthreads)
    t_asm = time.time() - t0
    x = splitmix_uniform(2, n)
    t1 = time.time()
    op.matvec(x, nthreads=nthreads)
    t_mv = time.time() - t1
    del op, T
    return {
        "value": 1.0 / (t_mv * n_mv),
        "unit": "solves/s",
        "cores": nthreads,
        "cpu_model": model,
        "kind": "port",
        "sample": ("oracle FP64 restatement of the whole operator on %d threads, reference execution model "
                   "(skeleton blocks for |cols| > 2*Nmax): tree + assembly %.1f s and one full matvec %.2f s "
                   "timed; solve = that matvec x %d matvecs of the FP64 GPU solve" % (nthreads, t_asm, t_mv, n_mv)),
        "assembly_s": t_asm,
        "matvec_s": t_mv,
    }


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


def to_tree_interleaved(trees, off, M):
    """Host N x K (problems concatenated, point order) -> tree order with the
    K columns interleaved, the device layout of the batch entry points."""
    out = np.empty_like(M)
    for p, t in enumerate(trees):
        out[off[p]:off[p + 1]] = M[off[p]:off[p + 1]][t.perm()]
    return np.ascontiguousarray(out)


def from_tree_interleaved(trees, off, M):
    out = np.empty_like(M)
    for p, t in enumerate(trees):
        out[off[p]:off[p + 1]][t.perm()] = M[off[p]:off[p + 1]]
    return out


def cpu_baseline_batch(specs, nmax, eps, tol, budget_s=25.0, omp=False):
    """The FP64 CPU restatement solving problems of the batch one at a time
    (one thread, reference execution model), over problems spread across the
    batch until the budget is used; solves/s = problems solved / time."""
    from oracle import pyoracle as O
    from hodlrnd_amd import workloads as Wl
    t0 = time.time()
    done, t_solve = [], 0.0
    order = list(range(0, len(specs), 16)) + list(range(8, len(specs), 16))
    for p in order:
        n, nd = specs[p]
        X = Wl.problem_points(p, n, nd)
        T = O.Tree(X, -np.ones(nd), np.ones(nd), nmax, 0)
        k = O.KernelDesc(O.K_EXP, nd, 1.0 / n, 1.0 + 1.0 / n)
        op = O.Operator(T, k, eps, sys_size=2 * nmax, nthreads=1)
        b = Wl.splitmix_uniform(11_000 + p, n)
        t1 = time.time()
        op.gmres(b, tol=tol, nthreads=1)
        t_solve += time.time() - t1
        done.append(p)
        del op
        if time.time() - t0 > budget_s:
            break
    if omp:
        # all host cores: independent problems solved concurrently, one per
        # thread (the oracle's calls release the interpreter lock); assembly
        # untimed as above
        from concurrent.futures import ThreadPoolExecutor
        nthreads, model = host_cores()
        probs = order[:max(len(done), 2 * nthreads)]  # enough problems to keep every thread busy
        ops = []
        for p in probs:
            n, nd = specs[p]
            X = Wl.problem_points(p, n, nd)
            T = O.Tree(X, -np.ones(nd), np.ones(nd), nmax, 0)
            k = O.KernelDesc(O.K_EXP, nd, 1.0 / n, 1.0 + 1.0 / n)
            ops.append((O.Operator(T, k, eps, sys_size=2 * nmax, nthreads=nthreads),
                        Wl.splitmix_uniform(11_000 + p, n)))
        t1 = time.time()
        with ThreadPoolExecutor(nthreads) as ex:
            list(ex.map(lambda ob: ob[0].gmres(ob[1], tol=tol, nthreads=1), ops))
        t_omp = time.time() - t1
        return {"value": len(probs) / t_omp, "unit": "solves/s", "cores": nthreads, "cpu_model": model,
                "kind": "port",
                "sample": "%d problems spread over the batch (the 1-thread sample's first) solved "
                          "concurrently, one per thread on %d threads, assembly untimed" % (len(probs), nthreads)}
    return {
        "value": len(done) / t_solve,
        "unit": "solves/s",
        "cores": 1,
        "kind": "port",
        "sample": ("oracle FP64 restatement, 1 thread, reference execution model (skeleton blocks for "
                   "|cols| > 2*Nmax), GMRES(100) from x0=b to %g (assembly untimed) for 1 RHS of each of %d "
                   "problems spread over the batch (%s); solves/s = problems / time" %
                   (tol, len(done), ",".join(str(q) for q in done))),
        "wall_s": time.time() - t0,
    }


def bench_batch(args, H):
    """Batched uneven problems: 256 independent systems, N log-uniform in
    [2000, 64000], NDIM from {2, 3}, points uniform in [-1,1]^d,
    A = I + (1/N) exp(-r) (A_ii = 1 + 1/N), Nmax 64, eps_ACA 1e-8, d' = 0,
    8 uniform right-hand sides per system, all solved to 1e-8 as one batch."""
    from hodlrnd_amd import workloads as Wl
    K = 8
    specs = Wl.batch_specs(args.nprob)
    trees, kerns = [], []
    t0 = time.time()
    for p, (n, nd) in enumerate(specs):
        X = Wl.problem_points(p, n, nd)
        trees.append(H.Tree(X, -np.ones(nd), np.ones(nd), args.nmax, 0))
        kerns.append(H.Kernel(H.KERNEL_EXP, nd, 1.0 / n, 1.0 + 1.0 / n))
    t_tree = time.time() - t0
    t0 = time.time()
    op = H.Operator(trees, kerns, args.eps, H.PREC_F64 | H.PREC_F32)
    t_asm = time.time() - t0
    st = op.stats()
    N = op.n
    off = op.offsets()
    NI = len(specs) * K
    log("batch: %d problems, N total %d, trees %.2fs, assembled in %.2fs: %.1f GB FP64 + %.1f GB FP32"
        % (len(specs), N, t_tree, t_asm, st["bytes_f64"] * 1e-9, st["bytes_f32"] * 1e-9))
    t0 = time.time()
    op.tune(reps=2)
    t_tune = time.time() - t0
    variants = {("f64" if pr == H.PREC_F64 else "f32") + "_k%d" % k: op.get_variant(pr, k)
                for pr in (H.PREC_F64, H.PREC_F32) for k in (4, 8, 16)}
    log("kernel forms tuned in %.1fs: %s" % (t_tune, variants))
    if args.matvec_only:
        res = {}
        for prec in (H.PREC_F64, H.PREC_F32):
            ms = op.matvec_timing(prec, K, args.matvec_only)
            kb = op.kernel_bytes(prec, K)
            res["f64" if prec == H.PREC_F64 else "f32"] = {"ms": list(ms), "bytes": list(kb)}
            log("matvec prec %d K %d: %.3f ms (proj %.3f, panel %.3f)" % (prec, K, ms[0], ms[1], ms[2]))
        return {"matvec_only": res, "kernel_forms": variants}
    ctx = op.ctx
    B = np.empty((N, K))
    for p in range(len(specs)):
        for c in range(K):
            B[off[p]:off[p + 1], c] = Wl.splitmix_uniform(11_000 + 1000 * c + p, off[p + 1] - off[p])
    db = H.DeviceArray(ctx, 8 * N * K)
    dx = H.DeviceArray(ctx, 8 * N * K)
    db.upload(to_tree_interleaved(trees, off, B))

    def run(mode, k):
        x0 = "b" if mode == H.SOLVE_F64 else "zero"
        lat, infos = [], None
        for _ in range(k):
            t1 = time.perf_counter()
            infos = op.gmres_batch_device(db.ptr, dx.ptr, K, restart=100, tol=args.tol, mode=mode, x0=x0)
            lat.append(time.perf_counter() - t1)
        return np.array(lat), infos

    run(H.SOLVE_MIXED, args.warmup)
    op.profile(True)
    op.profile_read(reset=True)
    latm, infm = run(H.SOLVE_MIXED, args.steps)
    ms, cnt = op.profile_read(reset=True)
    op.profile(False)
    log("batched mixed: %.1f ms per batch of %d solves" % (latm.mean() * 1e3, NI))
    Xt = dx.download(np.empty((N, K)))
    X = from_tree_interleaved(trees, off, Xt)
    R = B - op.matvec(X)
    res = [np.linalg.norm(R[off[p]:off[p + 1], c]) / np.linalg.norm(B[off[p]:off[p + 1], c])
           for p in range(len(specs)) for c in range(K)]
    run(H.SOLVE_F64, 1)
    lat64b, inf64b = run(H.SOLVE_F64, max(1, args.steps // 2))
    log("batched fp64: %.1f ms per batch" % (lat64b.mean() * 1e3))

    # the plain FP64 GPU path: every (problem, RHS) solved alone with the
    # single-problem solver on an FP64 operator of that problem (x0 = b)
    del db, dx
    t_plain, n_plain, it_plain = 0.0, 0, []
    nsample = len(specs) if args.plain_all else min(len(specs), 64)
    picks = np.linspace(0, len(specs) - 1, nsample).round().astype(int)
    for p in picks:
        n = specs[p][0]
        o1 = H.Operator(trees[p], kerns[p], args.eps, H.PREC_F64)
        perm = trees[p].perm()
        d1b = H.DeviceArray(ctx, 8 * n)
        d1x = H.DeviceArray(ctx, 8 * n)
        # untimed first solve: the solver's work buffers are allocated once
        # per operator, as the batched solver's are during its warm-up
        d1b.upload(np.ascontiguousarray(B[off[p]:off[p + 1], 0][perm]))
        o1.gmres_device(d1b.ptr, d1x.ptr, restart=100, max_iter=n, tol=args.tol, mode=H.SOLVE_F64, x0="b")
        for c in range(K):
            d1b.upload(np.ascontiguousarray(B[off[p]:off[p + 1], c][perm]))
            t1 = time.perf_counter()
            inf = o1.gmres_device(d1b.ptr, d1x.ptr, restart=100, max_iter=n, tol=args.tol, mode=H.SOLVE_F64,
                                  x0="b")
            t_plain += time.perf_counter() - t1
            n_plain += 1
            it_plain.append(inf["iterations"])
        del o1, d1b, d1x
    plain_sps = n_plain / t_plain
    log("plain fp64 single solves: %.3f ms each (%d problems x %d RHS)" % (t_plain / n_plain * 1e3, nsample, K))

    kb32 = op.kernel_bytes(H.PREC_F32, K)
    kb64 = op.kernel_bytes(H.PREC_F64, K)
    cands = [("matvec_proj_f64", ms[0], cnt[0], kb64[1]), ("matvec_panel_f64", ms[1], cnt[0], kb64[2]),
             ("matvec_proj_f32", ms[2], cnt[1], kb32[1]), ("matvec_panel_f32", ms[3], cnt[1], kb32[2])]
    name, tot_ms, c, bytes_ = max(cands, key=lambda t: t[1])
    avg_ms = tot_ms / max(c, 1)
    achieved = bytes_ / (avg_ms * 1e-3) * 1e-9
    pt = name[-3:]
    vk = variants.get("%s_k%d" % (pt, K), 0)
    kname = (panel_kernel_name if "panel" in name else proj_kernel_name)(vk, K, "float" if pt == "f32" else "double")
    traffic_b, traffic_b_src = measured_traffic(name, bytes_, "matvec_traffic_batch256_k8.json", kname)
    mv32 = (ms[2] + ms[3]) / max(cnt[1], 1)
    copy_gbs = H.stream_copy_gbps(1 << 30, 5)
    read_gbs = H.stream_read_gbps(8 << 30, 5)
    cpu = None
    cpu_omp = None
    if not args.no_cpu:
        log("cpu baseline (oracle, problems one at a time)...")
        cpu = cpu_baseline_batch(specs, args.nmax, args.eps, args.tol, budget_s=args.cpu_budget)
        log("cpu baseline on all host cores...")
        cpu_omp = cpu_baseline_batch(specs, args.nmax, args.eps, args.tol, budget_s=args.cpu_budget, omp=True)
        log("cpu baseline: %.3f solves/s on %d cores" % (cpu_omp["value"], cpu_omp["cores"]))
        log("cpu baseline: %.4g solves/s" % cpu["value"])
    val = NI / latm.mean()
    iters = [i["iterations"] for i in infm]
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
            "workload": "%d independent problems, N log-uniform [2000, 64000] (total %d), NDIM in {2,3}, "
                        "points uniform [-1,1]^d, A=I+(1/N)exp(-r) (A_ii=1+1/N), Nmax=%d, eps_ACA=%g, "
                        "d'=0, %d RHS each, GMRES(100) to rel. residual %g; a step solves the whole batch"
                        % (len(specs), N, args.nmax, args.eps, K, args.tol),
            "problems": len(specs), "rhs_per_problem": K, "n_total": int(N),
        },
        "p50_ms": float(np.percentile(latm, 50) * 1e3),
        "p99_ms": float(np.percentile(latm, 99) * 1e3),
        "instances_per_step": NI,
        "max_true_resid": float(max(res)),
        "all_converged": bool(all(i["converged"] for i in infm)),
        "inner_iterations": {"max": int(max(iters)), "mean": float(np.mean(iters))},
        "refinement_sweeps": int(infm[0]["outer"]),
        "batched_fp64_solves_per_s": NI / lat64b.mean(),
        "plain_fp64_solves_per_s": plain_sps,
        "plain_fp64_sample": "%d problems x %d RHS, single-problem FP64 operator and solver, x0=b" % (nsample, K),
        "plain_fp64_iterations_mean": float(np.mean(it_plain)),
        "speedup_vs_plain_fp64": val / plain_sps,
        "speedup_vs_batched_fp64": val * lat64b.mean() / NI,
        "matvec_ms_f32_k8": mv32,
        "matvec_hbm_gbs_f32_k8": kb32[0] / (mv32 * 1e-3) * 1e-9 if mv32 else None,
        "assemble_s": t_asm,
        "tune_s": t_tune,
        "kernel_forms": variants,
        "operator_gb": {"f64": st["bytes_f64"] * 1e-9, "f32": st["bytes_f32"] * 1e-9},
        "roofline": {
            "bound": "hbm", "kernel": name, "kernel_form": kname, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "measured_copy_gbs": copy_gbs,
            "measured_read_gbs": read_gbs, "frac_of_measured_read": achieved / read_gbs if read_gbs else None,
            "traffic": traffic_b,
            "traffic_source": traffic_b_src,
            "algorithmic_bytes_per_launch": bytes_, "avg_launch_ms": avg_ms,
        },
        "mfma": mfma_counters(variants, K),
        "cpu_baseline": cpu,
        "cpu_baseline_omp": cpu_omp,
    }
    return out


def panel_kernel_name(v, K, t="float"):
    """The panel-product kernel that kernel form v runs at K right-hand sides."""
    if v & 2 and K >= 4:
        return "%s<%s, %d>" % ("panel_mfma_w_kernel" if v & 32 else "panel_mfma_kernel", t, K)
    if v & 8 and K <= 8:
        return "panel_vec_kernel<%s, %d>" % (t, K)
    return "%s<%s, %d>" % ("panel_kernel" if v & 4 else "panel_seg_kernel", t, K)


def proj_kernel_name(v, K, t="float"):
    """The projection kernel that kernel form v runs at K right-hand sides."""
    if v & 1 and K >= 4:
        return "proj_mfma_kernel<%s, %d>" % (t, K)
    return "%s<%s, %d>" % ("proj_pack_kernel" if v & 16 and K <= 2 else "proj_kernel", t, K)


def mfma_counters(variants, K):
    """MFMA utilisation of the FP32 K-RHS matrix-core kernels the tuner picked,
    from the rocprofv3 counter summary committed for this workload (None when
    no matrix-core form was picked or it was not measured)."""
    v = variants.get("f32_k%d" % K, 0)
    names = (["proj_mfma_kernel<float, %d>" % K] if v & 1 else []) + \
            ([panel_kernel_name(v, K)] if v & 2 else [])
    path = os.path.join(ROOT, "profiles", "mfma_util_batch256_k8.json")
    if not names or not os.path.exists(path):
        return None
    with open(path) as f:
        tab = json.load(f)
    out = {n: {"util": tab[n]["mfma_util"], "tflops_issued": tab[n]["mfma_tflops"],
               "peak_tflops": tab[n]["peak_tflops"]} for n in names if n in tab}
    return {"kernels": out, "source": "profiles/mfma_util_batch256_k8.json (rocprofv3 SQ_VALU_MFMA_BUSY_CYCLES, "
                                      "GRBM_GUI_ACTIVE, SQ_INSTS_VALU_MFMA_MOPS_F32)"} if out else None


def bench_skeleton(args, H):
    """One N = n1d^3 solve per step (the headline's problem) with the
    reference's memory-efficient operator (LowRank_matrix.hpp:22-36,55-58,
    685-698): every compressed block whose column node holds more than
    sys_size points (default 2 Nmax, the reference's SYS_SIZE) keeps only its
    pivot rows and columns and the pseudo-inverse of its core; each matvec
    evaluates its entries again on the GPU."""
    n1 = args.n1d
    pts = make_points(n1)
    n = len(pts)
    sys_size = args.sys_size if args.sys_size >= 0 else 2 * args.nmax
    kern = H.Kernel(H.KERNEL_INV_R, 3, 1.0 / n, 1.0)
    tree = H.Tree(pts, -np.ones(3), np.ones(3), args.nmax, 0)
    t0 = time.time()
    op = H.Operator(tree, kern, args.eps, H.PREC_F64 | H.PREC_F32, sys_size=sys_size)
    t_asm = time.time() - t0
    st = op.stats()
    # the stored form of the same blocks: their r (m + n) factor entries in
    # place of the cores and pivot positions
    stored64 = st["bytes_f64"] - st["bytes_core"] + 8.0 * st["e_skel"]
    stored32 = st["bytes_f32"] + 4.0 * st["e_skel"]
    log("skeleton: N=%d sys_size=%d, assembled in %.2fs: %d skeleton blocks, %.3g entries on the fly per "
        "matvec, %.1f GB FP64 + %.1f GB FP32 (stored form: %.1f + %.1f GB)"
        % (n, sys_size, t_asm, st["n_skeleton"], st["e_skel"], st["bytes_f64"] * 1e-9, st["bytes_f32"] * 1e-9,
           stored64 * 1e-9, stored32 * 1e-9))
    # the whole product in the schedule the solves use (skeleton phases
    # beside the stored passes), then on one stream, where the skeleton
    # phases' own time is the "skeleton" figure
    mvt = {}
    for prec, key in ((H.PREC_F64, "f64"), (H.PREC_F32, "f32")):
        sched = op.skeleton_schedule(prec, 1)
        ms = op.matvec_timing(prec, 1, 10)
        op.set_skeleton_schedule(H.SKEL_SERIAL)
        m1 = op.matvec_timing(prec, 1, 10)
        op.set_skeleton_schedule(H.SKEL_AUTO)
        mvt[key] = {"total": ms[0], "proj": ms[1], "panel": ms[2], "skeleton_span": ms[3], "schedule": sched,
                    "total_one_stream": m1[0], "skeleton": m1[3]}
        log("matvec %s: %.3f ms in schedule %d (stored part: proj %.3f, panel %.3f; skeleton span %.3f); "
            "one stream %.3f ms (skeleton alone %.3f)" % (key, ms[0], sched, ms[1], ms[2], ms[3], m1[0], m1[3]))
    ctx = tree.ctx
    perm = tree.perm()
    b = splitmix_uniform(1, n)
    db = H.DeviceArray(ctx, 8 * n)
    dx = H.DeviceArray(ctx, 8 * n)
    db.upload(np.ascontiguousarray(b[perm]))

    def run(mode, k):
        x0 = "b" if mode == H.SOLVE_F64 else "zero"
        lat, infos = [], []
        for _ in range(k):
            t1 = time.perf_counter()
            infos.append(op.gmres_device(db.ptr, dx.ptr, restart=100, max_iter=n, tol=args.tol, mode=mode,
                                         inner_tol=args.inner_tol, x0=x0))
            lat.append(time.perf_counter() - t1)
        return np.array(lat), infos

    run(H.SOLVE_F64, args.warmup)
    lat64, inf64 = run(H.SOLVE_F64, args.steps)
    run(H.SOLVE_MIXED, args.warmup)
    latm, infm = run(H.SOLVE_MIXED, args.steps)
    log("skeleton fp64: %.2f ms/solve (%d iterations), mixed: %.2f ms/solve (%d inner, %d sweeps)"
        % (lat64.mean() * 1e3, inf64[-1]["iterations"], latm.mean() * 1e3, infm[-1]["iterations"],
           infm[-1]["outer"]))
    xt = dx.download(np.empty(n))
    x = np.empty(n)
    x[perm] = xt
    rows = np.unique((np.abs(splitmix_uniform(77, 256)) * n).astype(np.int64))
    xr = splitmix_uniform(78, n)
    yd = op.dense_rows(rows, xr)
    mv_err = float(np.linalg.norm(op.matvec(xr)[rows] - yd) / np.linalg.norm(yd))
    dres = float(np.linalg.norm(b[rows] - op.dense_rows(rows, x)) / np.linalg.norm(b[rows]))
    true_res = float(np.linalg.norm(b - op.matvec(x)) / np.linalg.norm(b))
    log("sampled dense rows: matvec error %.3g, residual of the mixed solution %.3g" % (mv_err, dres))
    sk = mvt["f64"]["skeleton"]
    ent = st["e_skel"] / (sk * 1e-3) if sk > 0 else 0.0
    achieved = ent * SKEL_SLOTS["f64"]
    return {
        "metric": "solves/s at 1e-8 residual, mixed vs FP64, on-the-fly skeleton operator",
        "value": args.steps / latm.sum(),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": latm.mean() * 1e3,
        "higher_is_better": True,
        "dtype": "fp32 matvec (fp64 on-the-fly entries) + fp64 refinement",
        "data": "synthetic",
        "config": {"workload": "3D %d^3 uniform grid, N=%d, A=I+(1/N)/r, Nmax=%d, eps_ACA=%g, d'=0, "
                               "skeleton blocks for |cols| > %d, GMRES(100) to rel. residual %g, 1 RHS"
                               % (n1, n, args.nmax, args.eps, sys_size, args.tol),
                   "n": n, "nmax": args.nmax, "sys_size": sys_size},
        "p50_ms": float(np.percentile(latm, 50) * 1e3),
        "p99_ms": float(np.percentile(latm, 99) * 1e3),
        "fp64_solves_per_s": args.steps / lat64.sum(),
        "fp64_p50_ms": float(np.percentile(lat64, 50) * 1e3),
        "fp64_p99_ms": float(np.percentile(lat64, 99) * 1e3),
        "iterations_fp64": inf64[-1]["iterations"],
        "iterations_mixed_inner": infm[-1]["iterations"],
        "refinement_sweeps": infm[-1]["outer"],
        "true_resid_mixed": true_res,
        "dense_check": {"rows": int(len(rows)), "matvec_rel_err": mv_err, "mixed_solution_resid": dres},
        "matvec_ms": mvt,
        "skeleton_blocks": st["n_skeleton"],
        "entries_on_the_fly_per_matvec": st["e_skel"],
        "entries_per_s_f64": ent,
        "entries_per_s_f32": st["e_skel"] / (mvt["f32"]["skeleton"] * 1e-3) if mvt["f32"]["skeleton"] else 0.0,
        "operator_gb": {"f64": st["bytes_f64"] * 1e-9, "f32": st["bytes_f32"] * 1e-9,
                        "cores": st["bytes_core"] * 1e-9},
        "stored_form_gb": {"f64": stored64 * 1e-9, "f32": stored32 * 1e-9},
        "assemble_s": t_asm,
        "roofline": {
            "bound": "valu",
            "kernel": "skel_proj + skel_reduce + skel_core + skel_expand (FP64 operator, 1 RHS)",
            "achieved": achieved,
            "peak": VALU_F64_LANE_OPS,
            "unit": "FP64 VALU lane-ops/s",
            "frac": achieved / VALU_F64_LANE_OPS,
            "slots_per_entry": SKEL_SLOTS["f64"],
            "avg_phase_ms": sk,
            "traffic": None,
        },
    }


def timed_solves(H, op, tree, b, tol, steps, warmup, inner_tol=0.0):
    """FP64 solves (x0 = b, HDD_matrix.hpp:52-53) and FP32-inner / FP64-refined
    solves (x0 = 0) with b resident on the device; the matvec kernels' HIP
    events are recorded on the solver's stream during each timed set."""
    n = tree.n
    perm = tree.perm()
    db = H.DeviceArray(tree.ctx, 8 * n)
    dx = H.DeviceArray(tree.ctx, 8 * n)
    db.upload(np.ascontiguousarray(b[perm]))

    def run(mode, k):
        x0 = "b" if mode == H.SOLVE_F64 else "zero"
        lat, infos = [], []
        for _ in range(k):
            t1 = time.perf_counter()
            infos.append(op.gmres_device(db.ptr, dx.ptr, restart=100, max_iter=n, tol=tol, mode=mode,
                                         inner_tol=inner_tol, x0=x0))
            lat.append(time.perf_counter() - t1)
        return np.array(lat), infos

    out = {}
    for mode, key in ((H.SOLVE_F64, "f64"), (H.SOLVE_MIXED, "mixed")):
        run(mode, warmup)
        op.profile(True)
        op.profile_read(reset=True)
        lat, inf = run(mode, steps)
        ms, cnt = op.profile_read(reset=True)
        op.profile(False)
        xt = dx.download(np.empty(n))
        x = np.empty(n)
        x[perm] = xt
        out[key] = {"lat": lat, "info": inf[-1], "ms": ms, "cnt": cnt, "x": x}
    return out


def kernel_roofline(op, H, ms, cnt):
    """The matvec kernel with the most time in a timed set of solves:
    algorithmic bytes per launch over its mean HIP-event launch time."""
    kb64 = op.kernel_bytes(H.PREC_F64, 1)
    kb32 = op.kernel_bytes(H.PREC_F32, 1)
    cands = [("matvec_proj_f64", ms[0], cnt[0], kb64[1]), ("matvec_panel_f64", ms[1], cnt[0], kb64[2]),
             ("matvec_proj_f32", ms[2], cnt[1], kb32[1]), ("matvec_panel_f32", ms[3], cnt[1], kb32[2])]
    name, tot, c, by = max(cands, key=lambda t: t[1])
    avg = tot / max(c, 1)
    ach = by / (avg * 1e-3) * 1e-9 if avg > 0 else 0.0
    return {"bound": "hbm", "kernel": name, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": by, "avg_launch_ms": avg,
            "launches": int(c)}


def solve_summary(r, steps):
    f, m = r["f64"], r["mixed"]
    return {
        "fp64_solves_per_s": steps / f["lat"].sum(),
        "fp64_p50_ms": float(np.percentile(f["lat"], 50) * 1e3),
        "fp64_p99_ms": float(np.percentile(f["lat"], 99) * 1e3),
        "fp64_iterations": f["info"]["iterations"],
        "mixed_solves_per_s": steps / m["lat"].sum(),
        "mixed_p50_ms": float(np.percentile(m["lat"], 50) * 1e3),
        "mixed_p99_ms": float(np.percentile(m["lat"], 99) * 1e3),
        "mixed_inner_iterations": m["info"]["iterations"],
        "mixed_sweeps": m["info"]["outer"],
        "mixed_true_resid": m["info"]["true_resid"],
    }


def bench_random3d(args, H):
    """NDIM 3, N = 10,000 points uniform iid in [-1,1]^3 (SplitMix64 seed 1),
    A = I + (1/N)/r, Nmax 64, eps_ACA 1e-10, d' 0, uniform RHS, GMRES(100) to
    1e-10: the GPU FP64 and mixed solves beside the FP64 CPU restatement run
    end to end (tree, assembly with the reference's skeleton blocks, one
    matvec, the full solve) on one thread and on the host cores; errors
    against the dense FP64 operator evaluated by direct summation on the
    device."""
    n, nd, nmax, eps, tol = 10000, 3, 64, 1e-10, 1e-10
    pts = splitmix_uniform(1, n * nd).reshape(n, nd)
    kern = H.Kernel(H.KERNEL_INV_R, nd, 1.0 / n, 1.0)
    t0 = time.time()
    tree = H.Tree(pts, -np.ones(nd), np.ones(nd), nmax, 0)
    op = H.Operator(tree, kern, eps, H.PREC_F64 | H.PREC_F32)
    t_build = time.time() - t0
    st = op.stats()
    b = splitmix_uniform(3, n)
    r = timed_solves(H, op, tree, b, tol, args.steps, args.warmup)
    mv = op.matvec_timing(H.PREC_F64, 1, 20)
    x = splitmix_uniform(2, n)
    y = op.matvec(x)
    yd = op.dense_rows(np.arange(n), x)
    mv_err = float(np.linalg.norm(y - yd) / np.linalg.norm(yd))
    xs = r["f64"]["x"]
    res_dense = float(np.linalg.norm(b - op.dense_rows(np.arange(n), xs)) / np.linalg.norm(b))
    log("random3d: gpu fp64 %.2f ms/solve (%d it), matvec %.3f ms, vs dense: matvec %.3g, solution residual %.3g"
        % (r["f64"]["lat"].mean() * 1e3, r["f64"]["info"]["iterations"], mv[0], mv_err, res_dense))
    out = {
        "metric": "solves/s at 1e-10 residual, GPU FP64 vs the FP64 CPU path",
        "value": args.steps / r["f64"]["lat"].sum(),
        "unit": "solves/s",
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["f64"]["lat"].mean() * 1e3,
        "dtype": "fp64",
        "config": {"workload": "3D N=10000 uniform iid [-1,1]^3 (seed 1), A=I+(1/N)/r, Nmax=64, eps_ACA=1e-10, "
                               "d'=0, GMRES(100) to rel. residual 1e-10, x0=b, 1 uniform RHS; every factor stored",
                   "n": n, "nmax": nmax, "eps_aca": eps, "tol": tol, "sys_size": -1},
        "build_s": t_build,
        "matvec_ms_f64": mv[0],
        "dense_check": {"rows": n, "matvec_rel_err": mv_err, "fp64_solution_resid_dense": res_dense},
        "operator_stats": {"n_flop_ref": st["n_flop_ref"], "max_rank": int(st["max_rank"]),
                           "blocks": int(st["n_blocks"]), "dense_blocks": int(st["n_dense"])},
        "roofline": kernel_roofline(op, H, r["f64"]["ms"], r["f64"]["cnt"]),
    }
    out.update(solve_summary(r, args.steps))
    if not args.no_cpu:
        from oracle import pyoracle as O
        nthreads, model = host_cores()
        ok = O.KernelDesc(O.K_INV, nd, 1.0 / n, 1.0)
        for key, nt in (("cpu_1t", 1), ("cpu_all", nthreads)):
            t0 = time.time()
            T = O.Tree(pts, -np.ones(nd), np.ones(nd), nmax, 0)
            ref = O.Operator(T, ok, eps, sys_size=2 * nmax, nthreads=nt)
            t_asm = time.time() - t0
            t1 = time.time()
            ref.matvec(x, nthreads=nt)
            t_mv = time.time() - t1
            t1 = time.time()
            xr, ir = ref.gmres(b, tol=tol, nthreads=nt)
            t_solve = time.time() - t1
            out[key] = {"value": 1.0 / t_solve, "unit": "solves/s", "cores": nt, "cpu_model": model,
                        "kind": "port", "build_s": t_asm, "matvec_s": t_mv, "solve_s": t_solve,
                        "iterations": ir["iterations"], "resid": ir["resid"],
                        "sample": "oracle FP64 restatement end to end on the whole problem, reference execution "
                                  "model (skeleton blocks for |cols| > 2*Nmax), %d thread(s)" % nt}
            log("random3d %s: build %.2fs, matvec %.3fs, solve %.3fs (%d it)" % (key, t_asm, t_mv, t_solve,
                                                                           ir["iterations"]))
        out["cpu_baseline"] = out["cpu_1t"]
    return out


def gauss_legendre_ref_order(m):
    """m-point Gauss-Legendre rule on [-1, 1] in the order of the reference's
    table (Gauss_Legendre_Nodes_and_Weights.hpp: -x1, +x1, -x2, +x2, ... from
    the centre outward; odd m starts with the centre node)."""
    x, w = np.polynomial.legendre.leggauss(m)
    h = m // 2
    order = [h] if m % 2 else []
    for k in range(h):
        lo, hi = (h - 1 - k, h + k) if m % 2 == 0 else (h - 1 - k, h + 1 + k)
        order += [lo, hi]
    return x[order], w[order]


def bench_gl4d(args, H):
    """4-D second-kind integral equation: the tensor product of 16
    Gauss-Legendre nodes per axis on [-1,1]^4 (N = 65,536), A_ij = delta_ij +
    w_j / |x_i - x_j| with w_j the product weight, Nmax 64, eps_ACA 1e-8,
    d' in {-1, 0, 3}; mixed (FP32 matvec, FP64 refinement) and FP64 solves to
    1e-8 per d'."""
    g, gw = gauss_legendre_ref_order(16)
    G = np.meshgrid(*([g] * 4), indexing="ij")
    pts = np.ascontiguousarray(np.stack([c.ravel() for c in G], axis=1))
    W = np.meshgrid(*([gw] * 4), indexing="ij")
    wts = np.ascontiguousarray(W[0].ravel() * W[1].ravel() * W[2].ravel() * W[3].ravel())
    n = len(pts)
    b = splitmix_uniform(7, n)
    rows = np.unique((np.abs(splitmix_uniform(6, 512)) * n).astype(np.int64))
    xr = splitmix_uniform(5, n)
    legs = {}
    for dp in (-1, 0, 3):
        t0 = time.time()
        tree = H.Tree(pts, -np.ones(4), np.ones(4), 64, dp)
        op = H.Operator(tree, H.Kernel(H.KERNEL_INV_R, 4, 1.0, 1.0, wts), 1e-8, H.PREC_F64 | H.PREC_F32)
        t_build = time.time() - t0
        st = op.stats()
        r = timed_solves(H, op, tree, b, args.tol, args.steps, args.warmup)
        mv64 = op.matvec_timing(H.PREC_F64, 1, 10)
        mv32 = op.matvec_timing(H.PREC_F32, 1, 10)
        yd = op.dense_rows(rows, xr)
        mv_err = float(np.linalg.norm(op.matvec(xr)[rows] - yd) / np.linalg.norm(yd))
        d = {"build_s": t_build, "blocks": int(st["n_blocks"]), "dense_blocks": int(st["n_dense"]),
             "max_rank": int(st["max_rank"]), "operator_gb_f64": st["bytes_f64"] * 1e-9,
             "matvec_ms": {"f64": mv64[0], "f32": mv32[0]},
             "dense_check": {"rows": int(len(rows)), "matvec_rel_err": mv_err},
             "roofline": kernel_roofline(op, H, r["mixed"]["ms"], r["mixed"]["cnt"])}
        d.update(solve_summary(r, args.steps))
        legs["dprime_%d" % dp] = d
        log("gl4d d'=%d: %d blocks, build %.2fs, mixed %.2f ms/solve, fp64 %.2f ms/solve, matvec f32 %.3f ms"
            % (dp, st["n_blocks"], t_build, r["mixed"]["lat"].mean() * 1e3, r["f64"]["lat"].mean() * 1e3,
               mv32[0]))
        del op, tree
    return {
        "metric": "solves/s at 1e-8 residual, mixed vs FP64, per d'",
        "value": legs["dprime_0"]["mixed_solves_per_s"],
        "unit": "solves/s",
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 / legs["dprime_0"]["mixed_solves_per_s"],
        "dtype": "fp32 matvec + fp64 refinement",
        "config": {"workload": "4D 16^4 Gauss-Legendre tensor nodes on [-1,1]^4, N=65536, A=I+w_j/r, Nmax=64, "
                               "eps_ACA=1e-8, d' in {-1,0,3}, GMRES(100) to rel. residual 1e-8, 1 uniform RHS; "
                               "value = d'=0 mixed; every factor stored",
                   "n": n, "sys_size": -1},
        "legs": legs,
    }


def bench_serve(args, H):
    """Serving stream: Poisson arrivals of solve requests, NDIM = 3, N
    log-uniform in [1000, 200000], points uniform in [-1,1]^3, A = I + (1/N)/r
    (A_ii = 1), eps_ACA 1e-8, Nmax 128 (the --nmax default), d' 0 (Nmax and d'
    are not fixed by the workload).  30% of requests bring new
    points (build + solve), 70% reuse an operator built earlier in the stream
    (uniformly), each with 1..32 right-hand sides.  Every column is solved to
    a FP64-certified relative residual of 1e-8.

    The offered stream (args.rate) runs first; the served request rate over
    its wall time is the server's saturation point when arrivals outpace it.
    A second stream offered at 0.8 x that rate (args.sub_requests requests)
    then shows the latency below saturation, where it does not grow with the
    stream's length."""
    from hodlrnd_amd import workloads as Wl

    def plan_for(nreq, rate, seed):
        arr = Wl.poisson_arrivals(rate, nreq, seed=seed) if rate > 0 else np.zeros(nreq)
        u = (Wl.splitmix_uniform(50 + seed, 4 * nreq).reshape(nreq, 4) + 1.0) / 2.0
        plan = []
        for q in range(nreq):
            build = q == 0 or u[q, 0] < 0.3
            n = Wl.log_uniform_int(2 * u[q, 1] - 1, 1000, 200000) if build else None
            plan.append((build, n, 1 + min(31, int(u[q, 2] * 32)), u[q, 3]))
        return plan, arr

    def one_stream(nreq, rate, seed):
        plan, arr = plan_for(nreq, rate, seed)
        srv = H.Server(slots=args.slots, tol=args.tol, eps_aca=args.eps, nmax=args.nmax, dprime=0,
                       hbm_budget=args.hbm_budget * 1e9 if args.hbm_budget > 0 else 0.0, builders=args.builders)
        # warm the kernels once (first launches load code objects)
        w = srv.submit_build(Wl.problem_points(999, 2000, 3), H.Kernel(H.KERNEL_INV_R, 3, 1.0 / 2000, 1.0))
        srv.wait(srv.submit_solve(w, Wl.splitmix_uniform(998, 2000)), timeout=120)
        log("serve: %d requests offered at %s req/s, %d builds" % (nreq, "%.1f" % rate if rate > 0 else "inf",
                                                                  sum(p[0] for p in plan)))
        try:
            return serve_stream(srv, plan, arr, nreq, rate, args, H, Wl)
        finally:
            srv.close()

    out = one_stream(args.requests, args.rate, 5)
    out["offered_req_per_s"] = args.rate
    if args.sub_requests > 0 and out["requests_per_s"] < 0.95 * args.rate:
        sub_rate = 0.8 * out["requests_per_s"]
        sub = one_stream(args.sub_requests, sub_rate, 6)
        out["below_saturation"] = {k: sub[k] for k in ("requests_per_s", "value", "p50_ms", "p99_ms",
                                                       "p50_ms_reuse", "p99_ms_reuse", "p50_ms_build",
                                                       "p99_ms_build", "p99_ms_second_half", "columns",
                                                       "wall_s", "all_converged", "max_resid", "server")}
        out["below_saturation"]["offered_req_per_s"] = sub_rate
        out["below_saturation"]["requests"] = args.sub_requests
    return out


def serve_stream(srv, plan, arr, nreq, rate, args, H, Wl):
    ops, nof, rids = [], {}, []
    t0 = time.perf_counter()
    for q, (build, n, nrhs, pick) in enumerate(plan):
        dt = arr[q] - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        if build:
            pts = Wl.problem_points(5000 + q, n, 3)
            oid = srv.submit_build(pts, H.Kernel(H.KERNEL_INV_R, 3, 1.0 / n, 1.0))
            ops.append(oid)
            nof[oid] = n
        else:
            oid = ops[min(len(ops) - 1, int(pick * len(ops)))]
            n = nof[oid]
        B = np.stack([Wl.splitmix_uniform(100000 + 64 * q + c, n) for c in range(nrhs)], axis=1)
        rids.append((srv.submit_solve(oid, B), build, nrhs, n))
        if (q + 1) % 200 == 0:
            log("submitted %d/%d" % (q + 1, nreq))
    t_sub = time.perf_counter() - t0
    lat, lat_b, lat_r, cols, worst, okc = [], [], [], 0, 0.0, True
    t_first, t_last = None, 0.0
    for k, (rid, build, nrhs, n) in enumerate(rids):
        _, inf = srv.wait(rid, timeout=1200, want_x=False)
        if inf["rc"] != 0:
            raise RuntimeError("request %d failed: %s" % (rid, inf))
        L = inf["t_done"] - inf["t_submit"]
        lat.append(L)
        (lat_b if build else lat_r).append(L)
        cols += nrhs
        worst = max(worst, inf["max_resid"])
        okc = okc and inf["converged"] == 1
        t_first = inf["t_submit"] if t_first is None else min(t_first, inf["t_submit"])
        t_last = max(t_last, inf["t_done"])
        if (k + 1) % 200 == 0:
            log("done %d/%d" % (k + 1, nreq))
    st = srv.stats()
    wall = t_last - t_first
    lat = np.array(lat)
    out = {
        "metric": METRIC,
        "value": cols / wall,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": nreq,
        "warmup": 1,
        "ms_per_step": float(np.mean(lat) * 1e3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 matvec + fp64 refinement",
        "data": "synthetic",
        "config": {
            "workload": "serving stream: %d requests, Poisson %s req/s, 3D, N log-uniform [1000, 200000], "
                        "A=I+(1/N)/r, eps_ACA=%g, Nmax=%d, d'=0, 30%% new operators / 70%% reuse, 1-32 RHS, "
                        "tol %g; a step is one request" % (nreq, rate if rate > 0 else "inf", args.eps, args.nmax,
                                                            args.tol),
            "requests": nreq, "rate": rate, "slots": args.slots, "builders": args.builders,
        },
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        # later half of the stream (in submission order): under saturation
        # it exceeds the first half, as the queue keeps growing
        "p99_ms_first_half": float(np.percentile(lat[: nreq // 2], 99) * 1e3),
        "p99_ms_second_half": float(np.percentile(lat[nreq // 2:], 99) * 1e3),
        "p50_ms_reuse": float(np.percentile(lat_r, 50) * 1e3) if lat_r else None,
        "p99_ms_reuse": float(np.percentile(lat_r, 99) * 1e3) if lat_r else None,
        "p50_ms_build": float(np.percentile(lat_b, 50) * 1e3) if lat_b else None,
        "p99_ms_build": float(np.percentile(lat_b, 99) * 1e3) if lat_b else None,
        "requests_per_s": nreq / wall,
        "columns": cols,
        "wall_s": wall,
        "submit_s": t_sub,
        "all_converged": bool(okc),
        "max_resid": worst,
        "server": {k: st[k] for k in ("builds", "evictions", "ticks", "peak_bytes", "budget_bytes",
                                      "build_seconds", "tick_seconds")},
    }
    return out


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
    ap.add_argument("--cpu-budget", type=float, default=25.0, help="seconds of CPU baseline work per leg")
    ap.add_argument("--no-extra", action="store_true",
                    help="grid100: skip the batched and serving legs reported beside the headline")
    ap.add_argument("--workload", choices=["grid100", "batch256", "serve", "skeleton", "random3d", "gl4d"],
                    default="grid100",
                    help="grid100: one N=1e6 solve per step; batch256: 256 uneven problems x 8 RHS "
                         "solved as one batch per step; serve: a Poisson stream of solve requests; "
                         "skeleton: the grid100 solve with on-the-fly skeleton blocks")
    ap.add_argument("--sys-size", type=int, default=-1,
                    help="grid100 / skeleton: blocks with more columns are skeletons (default 2 Nmax, the "
                         "reference's SYS_SIZE)")
    ap.add_argument("--operator", choices=["skeleton", "stored"], default="skeleton",
                    help="grid100: the reference's default operator (skeleton blocks for |cols| > sys_size, "
                         "entries evaluated on the fly) or every factor stored")
    ap.add_argument("--requests", type=int, default=100, help="serve: requests in the stream")
    ap.add_argument("--sub-requests", type=int, default=100,
                    help="serve: requests of the second stream, offered at 0.8 x the served rate (0: none)")
    ap.add_argument("--rate", type=float, default=200.0, help="serve: arrivals per second (0: all at once)")
    ap.add_argument("--slots", type=int, default=32, help="serve: solve columns per operator engine")
    ap.add_argument("--hbm-budget", type=float, default=0.0, help="serve: GB for operators (0: 90%% of free)")
    ap.add_argument("--builders", type=int, default=4, help="serve: concurrent operator builds")
    ap.add_argument("--nprob", type=int, default=256)
    ap.add_argument("--plain-all", action="store_true",
                    help="batch256: time the plain FP64 path on every problem, not 64 spread ones")
    ap.add_argument("--matvec-only", type=int, default=0, metavar="R",
                    help="grid100: assemble, then only R FP64 and R FP32 matvecs (counter runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the library's plain copy kernel (1 GiB, 3 launches): the "
                         "known-size kernel for scaling FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (independent problems "
              "would shard by problem, one operator by row ownership); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import hodlrnd_amd as H

    if args.copy_calibration:
        g = H.stream_copy_gbps(1 << 30, 3)
        print(json.dumps({"copy_kernel": "hodlr::stream_copy_kernel", "bytes_read_per_launch": 1 << 30,
                          "bytes_written_per_launch": 1 << 30, "launches": 4, "gbps": g}))
        return

    if args.workload in ("batch256", "serve"):
        if args.nmax == 128:
            args.nmax = 64
        print(json.dumps((bench_batch if args.workload == "batch256" else bench_serve)(args, H)))
        return
    if args.workload == "skeleton":
        print(json.dumps(bench_skeleton(args, H)))
        return
    if args.workload in ("random3d", "gl4d"):
        print(json.dumps((bench_random3d if args.workload == "random3d" else bench_gl4d)(args, H)))
        return

    copy_gbs = H.stream_copy_gbps(1 << 30, 5)
    read_gbs = H.stream_read_gbps(8 << 30, 5)
    n1 = args.n1d
    pts = make_points(n1)
    n = len(pts)
    kern = H.Kernel(H.KERNEL_INV_R, 3, 1.0 / n, 1.0)
    t0 = time.time()
    tree = H.Tree(pts, -np.ones(3), np.ones(3), args.nmax, 0)
    t_tree = time.time() - t0
    log("tree: N=%d levels=%d admissible=%d (%.2fs)" % (n, tree.nlevels, tree.info.n_admissible, t_tree))
    # the reference's default operator keeps skeleton blocks for every
    # compressed block with more than SYS_SIZE = 2 Nmax columns
    # (myHeaders.hpp:48, LowRank_matrix.hpp:22-36); --operator stored keeps
    # every factor
    sys_size = (args.sys_size if args.sys_size >= 0 else 2 * args.nmax) if args.operator == "skeleton" else -1
    t0 = time.time()
    op = H.Operator(tree, kern, args.eps, H.PREC_F64 | H.PREC_F32, sys_size=sys_size)
    t_asm = time.time() - t0
    st = op.stats()
    log("assembled in %.2fs (sys_size %d): e_near %.3g e_far %.3g, %d skeleton blocks (%.3g entries on the fly), "
        "%.1f GB FP64 + %.1f GB FP32; ACA rank pass %.2fs, %.3g entries (%.3g/s)"
        % (t_asm, sys_size, st["e_near"], st["e_far"], st["n_skeleton"], st["e_skel"], st["bytes_f64"] * 1e-9,
           st["bytes_f32"] * 1e-9, st["rank_seconds"], st["aca_entries"],
           st["aca_entries"] / st["rank_seconds"] if st["rank_seconds"] else 0.0))

    if args.matvec_only:
        for prec in (H.PREC_F64, H.PREC_F32):
            ms = op.matvec_timing(prec, 1, args.matvec_only)
            log("matvec prec %d: %.3f ms (proj %.3f, panel %.3f, skeleton %.3f)"
                % (prec, ms[0], ms[1], ms[2], ms[3]))
        return

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
    # against the dense operator itself, on 256 sampled rows (device direct
    # sums): the HODLR approximation error of the matvec, and the residual
    # of the mixed solution for the exact problem
    rows = np.unique((np.abs(splitmix_uniform(77, 256)) * n).astype(np.int64))
    xr = splitmix_uniform(78, n)
    yd = op.dense_rows(rows, xr)
    mv_err = float(np.linalg.norm(op.matvec(xr)[rows] - yd) / np.linalg.norm(yd))
    dres = float(np.linalg.norm(b[rows] - op.dense_rows(rows, x)) / np.linalg.norm(b[rows]))
    log("sampled dense rows: matvec error %.3g, residual of the mixed solution %.3g" % (mv_err, dres))

    # dominant kernel of the timed mixed solves
    kb64 = op.kernel_bytes(H.PREC_F64, 1)
    kb32 = op.kernel_bytes(H.PREC_F32, 1)
    cands = [("matvec_proj_f64", ms[0], cnt[0], kb64[1]), ("matvec_panel_f64", ms[1], cnt[0], kb64[2]),
             ("matvec_proj_f32", ms[2], cnt[1], kb32[1]), ("matvec_panel_f32", ms[3], cnt[1], kb32[2])]
    name, tot_ms, c, bytes_ = max(cands, key=lambda t: t[1])
    avg_ms = tot_ms / max(c, 1)
    achieved = bytes_ / (avg_ms * 1e-3) * 1e-9
    traffic_file = "matvec_traffic_grid100_N1e6%s.json" % ("_sys%d" % sys_size if sys_size >= 0 else "")
    traffic, traffic_src = measured_traffic(name, bytes_, traffic_file)
    # whole matvec (stored part + skeleton blocks) in the schedule the solves
    # used, timed on its own; the stored part's kernels as timed in the solves
    precs = (("f64", H.PREC_F64), ("f32", H.PREC_F32))
    t_whole = {k: op.matvec_timing(p, 1, 10) for k, p in precs}
    mv32, mv64 = t_whole["f32"][0], t_whole["f64"][0]
    st32 = (ms[2] + ms[3]) / max(cnt[1], 1)
    st64 = (ms[0] + ms[1]) / max(cnt[0], 1)
    skeleton = None
    one_stream = None
    if st["n_skeleton"]:
        # the skeleton phases (skel_proj, skel_reduce, skel_core and one
        # skel_expand launch per level) run beside the stored passes on a
        # second stream; alone on one stream they are timed against the FP64
        # vector-ALU issue peak
        sched = {k: op.skeleton_schedule(p, 1) for k, p in precs}
        op.set_skeleton_schedule(H.SKEL_SERIAL)
        t_one = {k: op.matvec_timing(p, 1, 10) for k, p in precs}
        op.set_skeleton_schedule(H.SKEL_AUTO)
        sk = {k: t_one[k][3] for k, _ in precs}
        ent = {k: st["e_skel"] / (v * 1e-3) if v else 0.0 for k, v in sk.items()}
        skeleton = {
            "sys_size": sys_size, "blocks": int(st["n_skeleton"]), "entries_per_matvec": st["e_skel"],
            "cores_gb": st["bytes_core"] * 1e-9,
            "schedule": {k: {1: "one stream", 2: "beside the stored passes",
                             3: "beside the stored passes, occupancy capped"}[v] for k, v in sched.items()},
            "span_beside_ms": {"f64": ms[4] / max(cnt[0], 1), "f32": ms[5] / max(cnt[1], 1)},
            "matvec_ms_one_stream": {k: t_one[k][0] for k, _ in precs},
            "phase_ms": sk, "entries_per_s": ent,
            "roofline": {"bound": "valu", "unit": "FP64 VALU lane-ops/s", "peak": VALU_F64_LANE_OPS,
                         "slots_per_entry": SKEL_SLOTS,
                         "achieved": {k: ent[k] * SKEL_SLOTS[k] for k in ent},
                         "frac": {k: ent[k] * SKEL_SLOTS[k] / VALU_F64_LANE_OPS for k in ent}},
        }
        # the dominant kernel when the skeleton phases are not beside it
        t_alone = t_one[name.split("_")[-1]][1 if "_proj_" in name else 2]
        one_stream = {"avg_launch_ms": t_alone, "achieved": bytes_ / (t_alone * 1e-3) * 1e-9,
                      "frac": bytes_ / (t_alone * 1e-3) * 1e-9 / HBM_PEAK_GBS}
        log("skeleton phase alone per matvec: f64 %.3f ms, f32 %.3f ms (%.3g / %.3g entries/s); schedule %s; "
            "whole matvec f64 %.3f ms (one stream %.3f), f32 %.3f ms (one stream %.3f)"
            % (sk["f64"], sk["f32"], ent["f64"], ent["f32"], sched, mv64, t_one["f64"][0], mv32, t_one["f32"][0]))

    n_mv64 = inf64[-1]["matvecs_f64"]
    cpu = cpu_1t = None
    if not args.no_cpu:
        del op  # the CPU restatement of the whole operator needs the host memory, not the GPU's
        log("cpu baseline (oracle, whole operator on the host cores)...")
        cpu = cpu_baseline_full(pts, n_mv64, args.nmax, args.eps)
        log("cpu baseline: assembly %.1fs, matvec %.2fs on %d threads: %.4g solves/s"
            % (cpu["assembly_s"], cpu["matvec_s"], cpu["cores"], cpu["value"]))
        cpu_1t = cpu_baseline(pts, n_mv64, args.nmax, args.eps)
        log("cpu baseline, 1 thread (row-node sample): %.4g solves/s" % cpu_1t["value"])

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
                        "GMRES(100) to rel. residual %g, 1 RHS (FP64: x0=b; mixed: x0=0); %s"
                        % (n1, n, args.nmax, args.eps, args.tol,
                           "the reference's default operator: skeleton blocks for |cols| > SYS_SIZE = %d "
                           "(2 Nmax), their entries evaluated on the fly" % sys_size if sys_size >= 0
                           else "every factor stored (sys_size -1)"),
            "n": n, "nmax": args.nmax, "eps_aca": args.eps, "dprime": 0, "tol": args.tol, "sys_size": sys_size,
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
        "dense_check": {"rows": int(len(rows)), "matvec_rel_err": mv_err, "mixed_solution_resid": dres},
        "resid_fp64_estimate": inf64[-1]["resid"],
        "matvec_ms": {"f64": mv64, "f32": mv32},
        # the stored factors' bytes over the time of the kernels that stream them
        "matvec_hbm_gbs": {"f64": kb64[0] / (st64 * 1e-3) * 1e-9 if st64 else None,
                           "f32": kb32[0] / (st32 * 1e-3) * 1e-9 if st32 else None},
        "skeleton": skeleton,
        "assemble_s": t_asm,
        "aca_rank_pass_s": st["rank_seconds"],
        # the rank pass's kernel entries (each row and column the ACA
        # evaluates, retries included) over its wall time
        "aca_entries": st["aca_entries"],
        "aca_entries_per_s": st["aca_entries"] / st["rank_seconds"] if st["rank_seconds"] else None,
        "fill_s": st["fill_seconds"],
        "operator_gb": {"f64": st["bytes_f64"] * 1e-9, "f32": st["bytes_f32"] * 1e-9},
        "stored_form_gb": {"f64": (st["bytes_f64"] - st["bytes_core"] + 8.0 * st["e_skel"]) * 1e-9,
                           "f32": (st["bytes_f32"] + 4.0 * st["e_skel"]) * 1e-9},
        # the reference's print_matrix_statistics figures (HDD_matrix.hpp:69-85)
        "operator_stats": {"n_flop_ref": st["n_flop_ref"], "max_rank": int(st["max_rank"]),
                           "blocks": int(st["n_blocks"]), "dense_blocks": int(st["n_dense"]),
                           "e_near": st["e_near"], "e_far": st["e_far"]},
        "roofline": {
            "bound": "hbm",
            "kernel": name,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "measured_copy_gbs": copy_gbs,
            # a plain streaming read of 8 GiB on the same card: the practical
            # ceiling of a kernel that only reads
            "measured_read_gbs": read_gbs,
            "frac_of_measured_read": achieved / read_gbs if read_gbs else None,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "algorithmic_bytes_per_launch": bytes_,
            "avg_launch_ms": avg_ms,
            # the same kernel timed with the skeleton phases after it on one
            # stream instead of beside it (the solves run them beside it:
            # the kernel shares its CUs, the whole product is faster)
            "one_stream": one_stream,
        },
        # every pass of the timed mixed solves (the FP32 projection and panel
        # product take nearly the same time, so which one is dominant varies)
        "matvec_kernels": {
            kn: {"avg_launch_ms": tm / max(cn, 1), "launches": int(cn),
                 "algorithmic_bytes_per_launch": by,
                 "achieved_gbs": by / (tm / max(cn, 1) * 1e-3) * 1e-9 if tm else None,
                 "frac": by / (tm / max(cn, 1) * 1e-3) * 1e-9 / HBM_PEAK_GBS if tm else None,
                 "traffic": measured_traffic(kn, by, traffic_file)[0]}
            for kn, tm, cn, by in cands},
        "cpu_baseline": cpu,
        "cpu_baseline_1t": cpu_1t,
    }
    if not args.no_extra:
        # the batched and served forms of the same solve, each on its own
        # workload; the same solve with every factor stored; and a 160^3
        # grid (N = 4.1M) whose stored FP64 + FP32 operator (~330 GB)
        # exceeds HBM, in skeleton form -- reported beside the headline.
        # Each runs in a child process of its own (a fresh device-memory
        # state), after this process has released its operator.
        op = None
        del db, dx
        import subprocess
        legs = (("batched", ["--workload", "batch256", "--steps", "3", "--warmup", "1", "--cpu-budget", "10"]),
                ("serving", ["--workload", "serve", "--requests", "100", "--sub-requests", "100"]),
                ("random3d_N1e4", ["--workload", "random3d", "--steps", "20", "--warmup", "2"]),
                ("gauss_legendre_4d", ["--workload", "gl4d", "--steps", "10", "--warmup", "2"]),
                ("stored_factors", ["--workload", "grid100", "--operator", "stored", "--no-extra", "--no-cpu",
                                    "--steps", "5", "--warmup", "1"]),
                ("skeleton_beyond_hbm", ["--workload", "skeleton", "--n1d", "160", "--sys-size", "128",
                                         "--steps", "3", "--warmup", "1"]))
        for key, extra in legs:
            cmd = [sys.executable, "-u", os.path.abspath(__file__)] + extra + (["--no-cpu"] if args.no_cpu else [])
            try:
                r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, text=True, timeout=900)
                line = [l for l in r.stdout.splitlines() if l.startswith("{")]
                if r.returncode != 0 or not line:
                    raise RuntimeError("exit status %d" % r.returncode)
                r2 = json.loads(line[-1])
                out[key] = {k: r2[k] for k in r2 if k not in ("metric", "higher_is_better", "scaling",
                                                               "vs_baseline", "n_gpus", "data")}
            except Exception as ex:  # the headline stands on its own
                out[key] = {"error": repr(ex)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
