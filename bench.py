"""HPL benchmark of the MI355X LU solve (P=Q=1 HPL_pdgesv path).

One step = one timed HPL_pdgesv (factor + back-solve) of the seeded HPL
matrix, timed as the reference times it (src/HPL_pdtest.cpp:128-136): the
matrix is generated before the timed region and the residual check runs after
it.  Every step must pass the HPL scaled-residual test (< 16).

Default: the FP64 solve at N=193280, NB=896 on one GPU (the [A|b] matrix
fills 278.4 GiB of HBM) -- the line's `value` -- followed by the two
mixed-precision forms, reported in the same JSON line:
  * "bf16": N=131072, NB=512, FP32 factors, BF16 MFMA update (operands split
    into --terms BF16 numbers), FP32 panel, FP64 refinement;
  * "fp8":  N=229376, NB=1024, FP16 factors, per-128x128-tile power-of-two
    scaled e4m3 MFMA update, FP32 panel, FP64 refinement,
    A = U(-0.5,0.5) + (N/4) I.
Their timed region is the whole hplx_mxp_pdgesv (factor, first solve and
every refinement step); HPL-equivalent TFLOP/s = (2/3 N^3 + 2 N^2) / t.
The default run then times the adjoint solve ("adjoint": getrf, getrs and
the adjoint of X = A^{-1} B at N=16384, A = U(-0.5,0.5) + 2 sqrt(N) I, 64
right-hand sides; --no-adjoint skips it).
--mode fp64|bf16|fp8 runs one form alone and makes it the line's value;
--mode adjoint runs the adjoint leg alone (value: its getrf's rate).
Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

# wall clock of the whole run starts here: every leg checks what is left
_T0 = time.monotonic()


def _host_threads():
    """(threads the CPU baseline uses, CPUs this process may run on).  All
    CPUs in the affinity mask, capped by OMP_NUM_THREADS when the host sets
    it (the CPU share a shared host grants one GPU job)."""
    try:
        cpus = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        cpus = os.cpu_count() or 1
    use = cpus
    cap = os.environ.get("OMP_NUM_THREADS", "")
    if cap.isdigit() and int(cap) > 0:
        use = min(use, int(cap))
    if os.environ.get("HPLX_CPU_CORES", "").isdigit():
        use = int(os.environ["HPLX_CPU_CORES"])
    return max(use, 1), cpus


# host BLAS threads for the CPU baseline; set before numpy/scipy load it
_CORES, _HOST_CPUS = _host_threads()
os.environ["OPENBLAS_NUM_THREADS"] = str(_CORES)

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("HPL-equiv TFLOP/s at residual<16, 1 GPU: FP64 LU vs BF16/FP8 "
          "LU+FP64 IR")
FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 matrix peak
BF16_PEAK_TFLOPS = 2500.0  # dense BF16 MFMA
FP8_PEAK_TFLOPS = 5000.0   # dense FP8 MFMA
HBM_PEAK_GBS = 8000.0
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "trailing_update_traffic.json")

FORMS = {
    # name: (N, NB, storage, update, diag shift as a fraction of N)
    "bf16": (131072, 512, "fp32", "bf16", 0.0),
    "fp8": (229376, 1024, "fp16", "fp8", 0.25),
}


def log(msg):
    print("[bench] %6.1f s  %s" % (time.monotonic() - _T0, msg),
          file=sys.stderr, flush=True)


def remaining(args):
    """Seconds left of the run's wall-clock budget."""
    return args.budget_s - (time.monotonic() - _T0)


# Expected wall time of each optional leg on one MI355X (allocation,
# generation, the timed run and its checks), what must be left of the budget
# for the leg to start.
LEG_SECONDS = {"bf16": 60.0, "fp8": 45.0, "adjoint": 30.0, "cpu": 20.0}


def hpl_flops(n):
    return (2.0 / 3.0) * n ** 3 + 2.0 * n ** 2


def ref_flops(n):
    return (2.0 / 3.0) * n ** 3 + 1.5 * n ** 2


def cpu_baseline(n_sample, nb_sample, cores, host_cpus, min_s, max_solves):
    """The oracle (a C restatement of the same P=Q=1 algorithm) on host
    OpenBLAS: the CPU configuration (N=4096, NB=256 by default) solved
    repeatedly until min_s seconds of solve time have accumulated (at most
    max_solves solves), each checked with the HPL residual.  value = the
    mean rate over the timed solves."""
    from oracle import oracle as o
    desc = o.use_host_blas(True)
    times, resids = [], []
    try:
        while len(times) < max_solves and (not times or sum(times) < min_s):
            t0 = time.perf_counter()
            r = o.pdgesv(n_sample, nb_sample)
            times.append(time.perf_counter() - t0)
            resids.append(o.residual(n_sample, r["x"])["resid"])
            if not resids[-1] < 16.0:
                raise SystemExit("CPU baseline residual check FAILED: %r"
                                 % resids[-1])
    finally:
        o.use_host_blas(False)
    t = sum(times) / len(times)
    return {"value": round(hpl_flops(n_sample) / t / 1e12, 4),
            "unit": "TFLOP/s", "cores": cores, "host_cpus": host_cpus,
            "kind": "port",
            "sample": "%d P=Q=1 solves of N=%d NB=%d NBMIN=32 NDIV=2 (oracle "
                      "hpl_oracle.c, dgemm/dtrsm from %s, %d threads), "
                      "%.3f s per solve, %.1f s in all, max resid %.4f"
                      % (len(times), n_sample, nb_sample, desc, cores, t,
                         sum(times), max(resids))}


def load_traffic(n, nb):
    """(bytes per launch of this run's shape or None, the committed counter
    measurement or None)."""
    if not os.path.exists(PROFILE_TRAFFIC):
        return None, None
    try:
        d = json.load(open(PROFILE_TRAFFIC))
    except ValueError:
        return None, None
    probe = {k: d[k] for k in ("update_probe", "N", "NB", "bytes_per_launch",
                               "algorithmic_bytes_per_launch",
                               "traffic_over_algorithmic") if k in d}
    if d.get("N") == n and d.get("NB") == nb:
        return d.get("bytes_per_launch"), probe
    return None, probe


def prefix_matrix(ctx, rp, n, k, nb, shift):
    """Columns 0..k-1 of the seeded N x N matrix (plus shift on the
    diagonal), generated on the device by the library's matgen and copied to
    the host in FP64."""
    import numpy as np
    d = ctx.alloc(8 * n * k)
    try:
        ctx.pdrandmat(n, k, nb, d.ptr, n, rp.HPL_ISEED, shift)
        return np.asfortranarray(d.download((n, k), np.float64))
    finally:
        d.free()


def gold_prefix(n, shift):
    """Committed LAPACK pivots of the longest N x K column prefix (K = 4096
    or 64) of the seeded matrix plus shift on the diagonal, or None."""
    import numpy as np
    path = os.path.join(ROOT, "tests", "golden", "hpl_golden.npz")
    if not os.path.exists(path):
        return None
    fx = np.load(path, allow_pickle=False)
    for k in (4096, 64):
        key = "prefix_ipiv_%d_%d" % (n, k) + ("_shift%d" % shift if shift else "")
        if key in fx.files:
            return fx[key]
    return None


def x_fp64(ctx, rp, n, nb):
    """x of the FP64 solve (hplx_pdgesv) of the same seeded system."""
    import numpy as np
    mat = ctx.pmat_alloc(n, nb)
    try:
        ctx.pdrandmat(n, n + 1, nb, mat.A, mat.ld, rp.HPL_ISEED, 0.0)
        info = ctx.pdgesv(rp.palg_default(), mat)
        if info != 0:
            raise SystemExit("FP64 reference solve: zero pivot at %d" % info)
        chk = ctx.pdtest_check(mat)
        return ctx.download_vec(mat.X, n, np.float64), chk["resid"]
    finally:
        ctx.pmat_free(mat)


def run_fp64(ctx, rp, args):
    import numpy as np
    n, nb = args.n, args.nb
    palg = rp.palg_default()
    palg.frac = args.frac

    # kernel pre-warm on a small matrix (cf. HPL_pdwarmup); not a step.  Only
    # for small N: at large N the one-off code-object load is lost in a
    # minutes-long solve, and without it every trailing-update launch in a
    # kernel trace of this command belongs to the timed solve.
    prewarm = args.prewarm
    if prewarm is None:
        prewarm = min(n, 4096) if n <= 32768 else 0
    if prewarm > 0:
        wmat = ctx.pmat_alloc(prewarm, min(nb, prewarm))
        ctx.pdrandmat(prewarm, prewarm + 1, wmat.nb, wmat.A, wmat.ld,
                      rp.HPL_ISEED, 0.0)
        ctx.pdgesv(palg, wmat)
        ctx.pmat_free(wmat)

    log("fp64: allocating N=%d NB=%d (%.1f GiB)" % (
        n, nb, rp.hpl_ld(n) * (n + 1) * 8 / 2 ** 30))
    mat = ctx.pmat_alloc(n, nb)
    times, checks, stats, kstats, swaps, pans = [], [], [], [], [], []
    # Warm-up solves only where they are short: above N = 32768 one solve
    # takes tens of seconds and the code objects are loaded by the first
    # panel, so the requested warm-up steps are not run there.
    warmup = args.warmup if n <= 32768 else 0
    step, warm_done, last = 0, 0, 0.0
    while True:
        timed = warm_done >= warmup
        if timed and len(times) >= max(args.steps, 1):
            break
        if step > 0:
            # the next solve (matgen + solve + check, as long as the last
            # one) must fit the FP64 share and leave the other legs theirs
            if timed and times and sum(times) + last > args.max_timed_s:
                break
            if remaining(args) - args.reserve_s < 1.1 * last:
                break
        s0 = time.monotonic()
        ctx.pdrandmat(n, n + 1, nb, mat.A, mat.ld, rp.HPL_ISEED, 0.0)
        ctx.synchronize()
        t0 = time.perf_counter()
        info = ctx.pdgesv(palg, mat)
        t = time.perf_counter() - t0
        if info != 0:
            raise SystemExit("pdgesv: exact zero pivot at column %d" % info)
        st = ctx.update_stats()
        kinds = {k: ctx.update_timing(k) for k in (1, 2, 3)}
        swp = ctx.swap_timing()
        pan = ctx.panel_timing()
        pk_runs = ctx.panel_kernel_runs()
        chk = ctx.pdtest_check(mat)
        last = time.monotonic() - s0
        log("fp64 %s step %d: %.3f s  %.2f TFLOP/s  resid %.4f  trailing GEMM "
            "%.2f TFLOP/s over %d launches" % (
                "timed" if timed else "warmup", step, t,
                hpl_flops(n) / t / 1e12, chk["resid"],
                st[2] / (st[0] * 1e-3) / 1e12 if st[0] > 0 else 0.0, st[1]))
        if not chk["resid"] < 16.0:
            raise SystemExit("residual check FAILED: %r" % chk)
        step += 1
        if timed:
            times.append(t)
            checks.append(chk)
            stats.append(st)
            kstats.append(kinds)
            swaps.append(swp)
            pans.append(pan)
        else:
            warm_done += 1

    pivots = None
    if args.check_pivots == 0:
        # the leading pivots against LAPACK dgetrf on the N x K column
        # prefix of the same matrix (committed fixture, longest K available)
        gold = gold_prefix(n, 0.0)
        if gold is not None:
            k = len(gold)
            gpiv = ctx.download_vec(mat.ipiv, k, np.int32)
            pivots = {"prefix": k, "equal": bool(np.array_equal(gpiv, gold)),
                      "against": "LAPACK dgetrf on the N x %d column prefix "
                                 "(committed fixture)" % k}
    if args.check_pivots > 0:
        import scipy.linalg.lapack as la
        k = min(args.check_pivots, n)
        log("prefix pivot check: first %d pivots vs LAPACK dgetrf" % k)
        gpiv = ctx.download_vec(mat.ipiv, k, np.int32)
        pref = prefix_matrix(ctx, rp, n, k, nb, 0.0)
        _, lpiv, _ = la.dgetrf(pref, overwrite_a=1)
        pivots = {"prefix": k, "equal": bool(np.array_equal(gpiv, lpiv[:k])),
                  "against": "LAPACK dgetrf on the N x %d column prefix" % k}
    ctx.pmat_free(mat)

    t_mean = sum(times) / len(times)
    traffic, traffic_probe = load_traffic(n, nb)
    gemm_ms = sum(s[0] for s in stats)
    gemm_launch = sum(s[1] for s in stats)
    gemm_flops = sum(s[2] for s in stats)
    # per kind of update launch (hplx_update_timing): 1 one launch per
    # update (the look-ahead's next-panel columns; every step when serial),
    # 2 the claimed part beside the panel, 3 the resume part on every CU.
    # The roofline is that of the kernel with the most time.
    names = {1: "update_kernel<false> (one launch per update)",
             2: "update_claimed_kernel (beside the panel)",
             3: "update_resume_kernel (rest of the update, every CU)"}
    per_kind = {}
    for k in (1, 2, 3):
        ms = sum(ks[k][0] for ks in kstats)
        cnt = sum(ks[k][1] for ks in kstats)
        fl = sum(ks[k][2] for ks in kstats)
        per_kind[k] = {"kernel": names[k], "launches": cnt,
                       "ms": round(ms, 2),
                       "avg_launch_ms": round(ms / cnt, 4) if cnt else None,
                       "tflops": round(fl / (ms * 1e-3) / 1e12, 3)
                       if ms > 0 else None, "flops": fl}
    dom = max(per_kind, key=lambda k: per_kind[k]["ms"])
    d = per_kind[dom]
    achieved = d["flops"] / (d["ms"] * 1e-3) / 1e12 if d["ms"] > 0 else 0.0
    # the row interchanges of the trailing columns (hplx_swap_timing): HBM
    # bound, 4 x 8 B per interchanged element
    sw_ms = sum(w[0] for w in swaps)
    sw_cnt = sum(w[1] for w in swaps)
    sw_bytes = sum(w[2] for w in swaps)
    sw_gbs = sw_bytes / (sw_ms * 1e-3) / 1e9 if sw_ms > 0 else None
    swap = {"kernel": "swap_gather_kernel (row interchanges of the trailing "
                      "columns)",
            "launches": sw_cnt, "ms": round(sw_ms, 2),
            "avg_launch_ms": round(sw_ms / sw_cnt, 4) if sw_cnt else None,
            "algorithmic_gbs": round(sw_gbs, 1) if sw_gbs else None,
            "bytes_per_element": 32,
            "hbm_frac": round(sw_gbs / HBM_PEAK_GBS, 4) if sw_gbs else None}
    # the panel factorisations (hplx_panel_timing): latency bound (the
    # per-column pivot exchange); the rate is that of the panel read and
    # written once, 16 B per element
    pn_ms = sum(w[0] for w in pans)
    pn_cnt = sum(w[1] for w in pans)
    pn_bytes = sum(w[2] for w in pans)
    pn_gbs = pn_bytes / (pn_ms * 1e-3) / 1e9 if pn_ms > 0 else None
    panel = {"panels": pn_cnt, "persistent_kernel_panels": pk_runs,
             "ms": round(pn_ms, 2),
             "avg_panel_ms": round(pn_ms / pn_cnt, 3) if pn_cnt else None,
             "algorithmic_gbs": round(pn_gbs, 1) if pn_gbs else None,
             "bytes_per_element": 16,
             "hbm_frac": round(pn_gbs / HBM_PEAK_GBS, 4) if pn_gbs else None,
             "note": "beside the update with the look-ahead; bound by the "
                     "leaves' per-column pivot exchange, not by HBM"}
    return {
        "value": hpl_flops(n) / t_mean / 1e12, "t_mean": t_mean,
        "config": {"workload": "hpl_pdgesv_fp64_p1q1", "N": n, "NB": nb,
                   "P": 1, "Q": 1, "pfact": "right", "rfact": "right",
                   "nbmin": 32, "ndiv": 2, "depth": 1, "swap": "long",
                   "frac": args.frac,
                   "panel": "persistent kernel on %d of %d panels, the "
                            "launch sequence on the rest (%s)"
                            % (pk_runs, (n + nb - 1) // nb, args.panel)},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3),
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                     "traffic": traffic,
                     "traffic_measured": traffic_probe,
                     "kernel": "trailing-update FP64 MFMA GEMM "
                               "(A22 -= L21*U12): %s, %d launches, %.1f ms"
                               % (d["kernel"], d["launches"], d["ms"]),
                     "avg_launch_ms": d["avg_launch_ms"],
                     "flops_per_launch": d["flops"] / max(d["launches"], 1),
                     "all_update_launches": {
                         "launches": gemm_launch, "ms": round(gemm_ms, 2),
                         "tflops": round(gemm_flops / (gemm_ms * 1e-3) / 1e12,
                                         3) if gemm_ms > 0 else None},
                     "by_kernel": {str(k): {kk: vv for kk, vv in v.items()
                                            if kk != "flops"}
                                   for k, v in per_kind.items()}},
        "swap": swap, "panel": panel,
        "residual": round(max(c["resid"] for c in checks), 6),
        "residual_r0_form": round(max(c["resid_r0_form"] for c in checks), 6),
        "tflops_ref_formula": round(ref_flops(n) / t_mean / 1e12, 3),
        "pivot_prefix_check": pivots,
        "steps_run": len(times), "warmup_run": warm_done,
        "step_seconds": [round(t, 3) for t in times],
    }


def update_traffic(n, nb, update, terms, storage_bytes):
    """Algorithmic MFMA operations and HBM bytes of all trailing-update
    launches of an n x n factorisation with panels of nb: C read and written
    once in the storage type, the packed operands read once."""
    ops = hbm = 0.0
    prods = terms * (terms + 1) // 2 if update == "bf16" else 1
    opb = 2 * terms if update == "bf16" else 1  # packed bytes per element
    for j in range(0, n, nb):
        jb = min(nb, n - j)
        m = n - j - jb
        if m <= 0:
            continue
        ops += 2.0 * m * m * jb * prods
        hbm += 2.0 * m * m * storage_bytes + 2.0 * m * jb * opb
    return ops, hbm


def run_mixed(ctx, rp, form, args):
    import numpy as np
    n, nb, storage, update, shift_frac = FORMS[form]
    if form == "bf16" and args.bf16_n:
        n = args.bf16_n
    if form == "fp8" and args.fp8_n:
        n = args.fp8_n
    shift = shift_frac * n
    st_code = rp.HPLX_STORE_FP32 if storage == "fp32" else rp.HPLX_STORE_FP16
    up_code = rp.HPLX_UPDATE_BF16 if update == "bf16" else rp.HPLX_UPDATE_FP8
    terms = args.terms if update == "bf16" else 1
    opts = rp.mxp_opts_default(up_code, terms)
    palg = rp.palg_default()
    log("%s: allocating N=%d NB=%d (%s factors %.1f GiB)" % (
        form, n, nb, storage, rp.hpl_ld(n) * n * (4 if storage == "fp32" else 2)
        / 2 ** 30))
    mat = ctx.mxp_alloc(n, nb, st_code)
    try:
        runs = []
        for step in range(args.mixed_warmup + 1):
            ctx.mxp_matgen(mat, rp.HPL_ISEED, shift)
            t0 = time.perf_counter()
            rc, res = ctx.mxp_pdgesv(palg, opts, mat)
            t = time.perf_counter() - t0
            log("%s step %d: %.3f s  %.2f TFLOP/s  its %d  resid %.4f  factor "
                "%.1f ms  update %.1f ms over %d launches" % (
                    form, step, t, hpl_flops(n) / t / 1e12, res["its"],
                    res["resid"], res["factor_ms"], res["update_ms"],
                    res["update_launches"]))
            if rc != 0 or not res["resid"] < 16.0:
                raise SystemExit("%s: refinement did not reach residual < 16: "
                                 "%r" % (form, res))
            runs.append((t, res))
        t, res = runs[-1]
        # the last run's update launches by kind (1 whole update, 2 claimed
        # beside the panel, 3 resumed): (ms, launches, flops)
        kinds = {k: ctx.update_timing(k) for k in (1, 2, 3)}
        ipiv = ctx.download_vec(mat.ipiv, n, np.int32)
        ident = int(np.sum(ipiv != np.arange(n)))
        x = ctx.download_vec(mat.X, n, np.float64) if form == "bf16" else None
    finally:
        ctx.mxp_free(mat)
    gold = gold_prefix(n, shift)
    pivots = None
    if gold is not None:
        k = len(gold)
        agree = int(np.sum(ipiv[:k] == gold))
        pivots = {"prefix": k, "agree": agree, "equal": agree == k,
                  "against": "LAPACK dgetrf (FP64) on the N x %d column "
                             "prefix (committed fixture)" % k}
    xcmp = None
    if x is not None and args.x_check:
        log("%s: FP64 reference solve at N=%d for the x comparison" % (form, n))
        x64, r64 = x_fp64(ctx, rp, n, args.nb)
        d = x - x64
        xcmp = {"rel_inf": float(np.max(np.abs(d)) / np.max(np.abs(x64))),
                "rel_2": float(np.linalg.norm(d) / np.linalg.norm(x64)),
                "fp64_residual": round(r64, 6),
                "fp64_nb": args.nb}
    sb = 4 if storage == "fp32" else 2
    ops, hbm = update_traffic(n, nb, update, terms, sb)
    peak = BF16_PEAK_TFLOPS if update == "bf16" else FP8_PEAK_TFLOPS
    prods = terms * (terms + 1) // 2 if update == "bf16" else 1
    total_fl = sum(v[2] for v in kinds.values()) or 1.0
    # the roofline of the kernel kind with the most time (the resumed part
    # of the split updates with the look-ahead): its MFMA operations and its
    # share of the algorithmic bytes over its own launches
    dom = max(kinds, key=lambda k: kinds[k][0])
    d_ms, d_cnt, d_fl = kinds[dom]
    sec = d_ms * 1e-3
    mfma = d_fl * prods / sec / 1e12 if sec > 0 else 0.0
    gbs = hbm * d_fl / total_fl / sec / 1e9 if sec > 0 else 0.0
    fr_m, fr_h = mfma / peak, gbs / HBM_PEAK_GBS
    names = {1: "update of one launch (the next panel's columns, or a whole "
                "update)", 2: "claimed part beside the panel",
             3: "resumed part, every CU"}
    by_kernel = {str(k): {"kernel": names[k], "launches": v[1],
                          "ms": round(v[0], 1),
                          "mfma_tflops": round(v[2] * prods / v[0] / 1e9, 2)
                          if v[0] > 0 else None}
                 for k, v in kinds.items() if v[1] > 0}
    return {
        "value": round(hpl_flops(n) / t / 1e12, 3), "ms": round(t * 1e3, 1),
        "config": {"workload": "hpl_mxp_%s_p1q1" % form, "N": n, "NB": nb,
                   "storage": storage, "update": update, "terms": terms,
                   "panel": "fp32", "refinement": "fp64 residual and update, fp32 solves",
                   "diag_shift": shift},
        "refinement_its": res["its"], "residual": round(res["resid"], 6),
        "residual_history": [float("%.4g" % h) for h in res["history"]],
        "factor_ms": round(res["factor_ms"], 1),
        "solve_ms": round(res["solve_ms"], 1),
        "pivots_off_diagonal": ident,
        "pivot_prefix_check": pivots,
        "x_vs_fp64": xcmp,
        "update_roofline": {
            "bound": "mfma" if fr_m >= fr_h else "hbm",
            "kernel": names[dom],
            "mfma_achieved": round(mfma, 2), "mfma_peak": peak,
            "mfma_frac": round(fr_m, 4), "hbm_achieved_gbs": round(gbs, 1),
            "hbm_peak_gbs": HBM_PEAK_GBS, "hbm_frac": round(fr_h, 4),
            "launches": d_cnt, "ms": round(d_ms, 1),
            "all_update_launches": {"launches": res["update_launches"],
                                    "ms": round(res["update_ms"], 1),
                                    "mfma_tflops": round(ops / res["update_ms"]
                                                         / 1e9, 2)},
            "by_kernel": by_kernel},
    }


def copy_probe(gib):
    """A plain device copy of known size (gib GiB read, gib GiB written), the
    calibration of the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = int(gib * 2 ** 30) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 8, "copies": 3}), flush=True)


def _pmc_per_dispatch(path, counter, match):
    import csv
    vals, names = {}, {}
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter or not match(r["Kernel_Name"]):
                continue
            d = r["Dispatch_Id"]
            vals[d] = vals.get(d, 0.0) + float(r["Counter_Value"])
            names[d] = r["Kernel_Name"]
    return vals, names


def update_probe(m, n, jb):
    """The FP64 trailing update (HPL_pdupdateTT: swaps, U12 = L11^-1 A12,
    A22 -= L21 U12) alone on an m x (jb + n) block, three times, with the
    row interchanges of a random panel: the shape the HBM counters are read
    on (under rocprofv3 --pmc the whole solve ends
    with a host-side crash in the profiler right after the matrix
    allocation, at N = 193280 and at N = 163840)."""
    import numpy as np
    import rochpl_amd as rp
    ctx = rp.Context(0)
    lda = m + 32
    dA = ctx.alloc(8 * lda * (jb + n))
    ctx.pdrandmat(m, jb + n, jb, dA.ptr, lda, rp.HPL_ISEED, 0.0)
    # the interchanges of a panel of a random matrix: pivot i drawn from the
    # rows at or below i (composed on the device, as in the solve)
    rng = np.random.default_rng(7)
    ipiv = np.array([rng.integers(i, m) for i in range(jb)], dtype=np.int32)
    dpiv = ctx.alloc(4 * jb)
    dpiv.upload(ipiv)
    dperm = ctx.alloc(8 * jb)
    ctx.pipid(jb, dpiv.ptr, dperm.ptr)
    dW = ctx.alloc(8 * jb * (n + 4))
    for _ in range(3):
        ctx.pdupdate(dA.ptr, lda, m, jb, n, dperm.ptr, dW.ptr)
    ctx.synchronize()
    ctx.close()
    print(json.dumps({"update_probe": {"m": m - jb, "n": n, "k": jb},
                      "launches": 3}), flush=True)


def mxp_probe(form, m, n, k):
    """The low-precision trailing update alone (packing + GEMM, through
    hplx_mxp_update), three times on an m x n x k shape: FP8 with FP16 C (the
    FP8 form) or 3-term BF16 with FP32 C (the BF16 form); the shape the
    kernels' counters are read on."""
    import numpy as np
    import rochpl_amd as rp
    ctx = rp.Context(0)
    fp8 = form == "fp8"
    storage = rp.HPLX_STORE_FP16 if fp8 else rp.HPLX_STORE_FP32
    csize = 2 if fp8 else 4
    dL = ctx.alloc(4 * m * k)
    dU = ctx.alloc(4 * k * n)
    dC = ctx.alloc(csize * m * n)
    rng = np.random.default_rng(3)
    # operands of the size the factorisation produces (|L| <= 1, U of the
    # scale of A's entries), generated block by block on the host
    cdt = np.float16 if fp8 else np.float32
    for dst, rows, cols, scale, dt in ((dL, m, k, 1.0, np.float32),
                                       (dU, k, n, 0.5, np.float32),
                                       (dC, m, n, 0.5, cdt)):
        size = np.dtype(dt).itemsize
        step = max(1, (64 << 20) // (size * rows))
        for c0 in range(0, cols, step):
            blk = rng.uniform(-scale, scale, (rows, min(step, cols - c0)))
            dst.upload(np.asfortranarray(blk, dtype=dt),
                       byte_offset=size * rows * c0)
    for _ in range(3):
        ctx.mxp_update(rp.HPLX_UPDATE_FP8 if fp8 else rp.HPLX_UPDATE_BF16,
                       1 if fp8 else 3, storage, m, n, k, dL.ptr, m, dU.ptr, k,
                       dC.ptr, m)
    ctx.synchronize()
    for d in (dL, dU, dC):
        d.free()
    ctx.close()
    print(json.dumps({"mxp_probe": {"form": form, "m": m, "n": n, "k": k},
                      "launches": 3}), flush=True)


def traffic_summary(fetch_csv, write_csv, cfetch_csv, cwrite_csv, n, nb,
                    copy_bytes, probe=None):
    """HBM bytes per trailing-update launch from two rocprofv3 --pmc runs
    (FETCH_SIZE, WRITE_SIZE; KiB per dispatch), each scaled by what the same
    counter reads for the copy probe of known size.  probe = (m, n, k) of the
    update probe, or None for the FP64 bench at (n, nb)."""
    # the trailing update's instance (update_kernel<true> runs the panel
    # products)
    upd = lambda k: "update_kernel<false>" in k
    cpy = lambda k: "copy" in k.lower()
    f, _ = _pmc_per_dispatch(fetch_csv, "FETCH_SIZE", upd)
    w, _ = _pmc_per_dispatch(write_csv, "WRITE_SIZE", upd)
    cf, _ = _pmc_per_dispatch(cfetch_csv, "FETCH_SIZE", cpy)
    cw, _ = _pmc_per_dispatch(cwrite_csv, "WRITE_SIZE", cpy)
    if not (f and w and cf and cw):
        raise SystemExit("traffic summary: missing dispatches")
    cf_kib = sum(cf.values()) / len(cf)
    cw_kib = sum(cw.values()) / len(cw)
    ff = copy_bytes / (cf_kib * 1024.0)
    wf = copy_bytes / (cw_kib * 1024.0)
    fb = sum(f.values()) / len(f) * 1024.0 * ff
    wb = sum(w.values()) / len(w) * 1024.0 * wf
    # algorithmic bytes of the average launch: C (m x nt) read and written,
    # L21 and U12^T read once
    alg = 0.0
    launches = 0
    if probe is not None:
        pm, pn, pk = probe
        alg, launches = 8.0 * (2.0 * pm * pn + pm * pk + pn * pk), 1
    else:
        for j in range(0, n, nb):
            jb = min(nb, n - j)
            m = n - j - jb
            if m <= 0:
                continue
            alg += 8.0 * (2.0 * m * m + 2.0 * m * jb)
            launches += 1
    shape = ({"update_probe": {"m": probe[0], "n": probe[1], "k": probe[2]}}
             if probe is not None else {"N": n, "NB": nb})
    d = dict(shape, **{"kernel": "update_kernel<false>",
         "launches": len(f), "fetch_kib_per_launch_raw": sum(f.values()) / len(f),
         "write_kib_per_launch_raw": sum(w.values()) / len(w),
         "copy_probe_bytes": copy_bytes, "copy_fetch_kib_raw": cf_kib,
         "copy_write_kib_raw": cw_kib, "fetch_factor": round(ff, 4),
         "write_factor": round(wf, 4),
         "fetch_bytes_per_launch": fb, "write_bytes_per_launch": wb,
         "bytes_per_launch": fb + wb,
         "algorithmic_bytes_per_launch": alg / max(launches, 1)})
    d["traffic_over_algorithmic"] = round(d["bytes_per_launch"] /
                                          d["algorithmic_bytes_per_launch"], 4)
    os.makedirs(os.path.dirname(PROFILE_TRAFFIC), exist_ok=True)
    with open(PROFILE_TRAFFIC, "w") as fh:
        json.dump(d, fh, indent=1)
    print(json.dumps(d), flush=True)


def run_adjoint(ctx, rp, args):
    """The adjoint solve: X = A^{-1} B and, for an upstream gradient G,
    gradB = A^{-T} G and gradA = -gradB X^T, at N = 16384 with
    A = U(-0.5,0.5) + 2 sqrt(N) I from the HPL generator and 64 right-hand
    sides from a seeded normal generator.  getrf, getrs and the adjoint are
    each the median of `reps` timed calls after one untimed call; the
    adjoint is timed with and without gradA, whose difference is the
    HBM-bound outer product that writes the N x N gradA.  Results are
    checked on the host against the matrix as generated: the residuals of
    both solves and sampled gradA columns against -gradB X^T."""
    import numpy as np
    n, nb, nrhs, reps = args.adj_n, args.adj_nb, 64, 5
    shift = 2.0 * n ** 0.5
    palg = rp.palg_default()
    log("adjoint: N=%d NB=%d nrhs=%d" % (n, nb, nrhs))
    mat = ctx.pmat_alloc(n, nb)
    rng = np.random.default_rng(20240501)
    B = np.asfortranarray(rng.standard_normal((n, nrhs)))
    G = np.asfortranarray(rng.standard_normal((n, nrhs)))
    dX, dG, dgB = (ctx.alloc(8 * n * nrhs) for _ in range(3))
    dgA = ctx.alloc(8 * n * n)
    dG.upload(G)

    def timed(fn):
        ctx.synchronize()
        t0 = time.perf_counter()
        fn()
        ctx.synchronize()
        return time.perf_counter() - t0

    def getrf():
        info = ctx.pdgetrf(palg, mat)
        if info != 0:
            raise SystemExit("adjoint: exact zero pivot at column %d" % info)

    try:
        ctx.pdrandmat(n, n, nb, mat.A, mat.ld, rp.HPL_ISEED, shift)
        A0 = np.array(ctx.download_matrix(mat, rows=n, cols=n), order="F")
        t_lu = []
        for r in range(reps + 1):
            ctx.pdrandmat(n, n, nb, mat.A, mat.ld, rp.HPL_ISEED, shift)
            t = timed(getrf)
            if r:
                t_lu.append(t)
        t_rs = []
        for r in range(reps + 1):
            dX.upload(B)
            t = timed(lambda: ctx.pdgetrs(mat, 0, nrhs, dX.ptr, n))
            if r:
                t_rs.append(t)
        t_adj, t_adj_b = [], []
        for r in range(reps + 1):
            t = timed(lambda: ctx.pdgesv_adjoint(mat, nrhs, dX.ptr, n, dG.ptr,
                                                 n, dgB.ptr, n, dgA.ptr, n))
            if r:
                t_adj.append(t)
            t = timed(lambda: ctx.pdgesv_adjoint(mat, nrhs, dX.ptr, n, dG.ptr,
                                                 n, dgB.ptr, n, None, n))
            if r:
                t_adj_b.append(t)
        X = dX.download((n, nrhs), np.float64)
        gB = dgB.download((n, nrhs), np.float64)
        cols = np.linspace(0, n - 1, 16).astype(int)
        gA = np.stack([dgA.download((n,), np.float64, byte_offset=8 * n * int(j))
                       for j in cols], axis=1)
    finally:
        for d in (dX, dG, dgB, dgA):
            d.free()
        ctx.pmat_free(mat)

    an = np.linalg.norm(A0)
    res_x = float(np.linalg.norm(A0 @ X - B) / (an * np.linalg.norm(X)))
    res_gb = float(np.linalg.norm(A0.T @ gB - G) / (an * np.linalg.norm(gB)))
    ref = -(gB @ X[cols].T)
    err_ga = float(np.abs(gA - ref).max() / np.abs(ref).max())
    del A0
    checks = {"solve_residual": res_x, "adjoint_residual": res_gb,
              "gradA_vs_outer_product": err_ga}
    if not (res_x < 1e-13 and res_gb < 1e-13 and err_ga < 1e-12):
        raise SystemExit("adjoint check FAILED: %r" % checks)
    med = lambda v: sorted(v)[len(v) // 2]
    lu, rs, adj, adj_b = med(t_lu), med(t_rs), med(t_adj), med(t_adj_b)
    ga = max(adj - adj_b, 1e-9)
    out = {"config": {"workload": "hpl_getrf_getrs_adjoint_p1q1", "N": n,
                      "NB": nb, "nrhs": nrhs, "diag_shift": "2*sqrt(N)",
                      "reps": reps},
           "getrf_ms": round(lu * 1e3, 3),
           "getrf_tflops": round(2.0 / 3.0 * n ** 3 / lu / 1e12, 2),
           "getrs_ms": round(rs * 1e3, 3),
           "getrs_tflops": round(2.0 * n * n * nrhs / rs / 1e12, 3),
           "adjoint_ms": round(adj * 1e3, 3),
           "adjoint_without_gradA_ms": round(adj_b * 1e3, 3),
           "gradA_ms": round(ga * 1e3, 3),
           "gradA_gbs": round(8.0 * n * n / ga / 1e9, 1),
           "gradA_hbm_frac": round(8.0 * n * n / ga / 1e9 / HBM_PEAK_GBS, 4),
           "checks": checks}
    log("adjoint: getrf %.2f ms (%.1f TF/s), getrs %.2f ms, adjoint %.2f ms "
        "(gradA %.2f ms, %.0f GB/s)" % (lu * 1e3, out["getrf_tflops"],
                                        rs * 1e3, adj * 1e3, ga * 1e3,
                                        out["gradA_gbs"]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--mode", choices=["all", "fp64", "bf16", "fp8", "adjoint"],
                    default="all")
    ap.add_argument("--n", type=int, default=193280)
    ap.add_argument("--nb", type=int, default=896)
    ap.add_argument("--frac", type=float, default=0.6)
    ap.add_argument("--terms", type=int, default=3,
                    help="BF16 terms per operand of the bf16 form (1-3)")
    ap.add_argument("--bf16-n", type=int, default=0)
    ap.add_argument("--fp8-n", type=int, default=0)
    ap.add_argument("--mixed-warmup", type=int, default=0)
    ap.add_argument("--cpu-n", type=int, default=4096,
                    help="N of the CPU baseline's solves (the CPU "
                         "configuration: N=4096, NB=256)")
    ap.add_argument("--cpu-nb", type=int, default=256)
    ap.add_argument("--cpu-min-s", type=float, default=10.0,
                    help="repeat the CPU baseline's solve until this much "
                         "solve time has accumulated")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--budget-s", type=float, default=480.0,
                    help="wall-clock budget of the whole run from process "
                         "start: FP64 steps beyond the first and the other "
                         "legs run only while it allows")
    ap.add_argument("--max-timed-s", type=float, default=120.0,
                    help="timed FP64 solve seconds after which no further "
                         "step starts")
    ap.add_argument("--no-adjoint", dest="adjoint", action="store_false",
                    help="skip the adjoint-solve leg of the default run")
    ap.add_argument("--adj-n", type=int, default=16384)
    ap.add_argument("--adj-nb", type=int, default=512)
    ap.add_argument("--prewarm", type=int, default=None,
                    help="N of an untimed warm-up solve (default: min(N,4096) "
                         "for N <= 32768, none above)")
    ap.add_argument("--check-pivots", type=int, default=0,
                    help="compare the first K pivots with LAPACK dgetrf on the "
                         "N x K column prefix")
    ap.add_argument("--panel", choices=["auto", "kernel", "launches"],
                    default="auto",
                    help="FP64 panels: the persistent panel kernel for "
                         "panels of >= 65536 rows (auto, the library "
                         "default), for every panel (kernel), or none "
                         "(launches): HPLX_OPT_PANEL_KERNEL 2, 1, 0")
    ap.add_argument("--no-x-check", dest="x_check", action="store_false",
                    help="skip the FP64 solve that the bf16 form's x is "
                         "compared with")
    ap.add_argument("--copy-probe-gib", type=float, default=0.0,
                    help="only run a device copy of this size (counter "
                         "calibration) and exit")
    ap.add_argument("--mxp-probe", nargs=4, default=None,
                    metavar=("FORM", "M", "N", "K"),
                    help="run the BF16 or FP8 trailing update alone (fp8|bf16 "
                         "m n k), for counter collection")
    ap.add_argument("--update-probe", nargs=2, type=int, default=None,
                    metavar=("M", "N"),
                    help="only run the FP64 trailing update on an M x (NB+N) "
                         "block (counter shape) and exit")
    ap.add_argument("--traffic-summary", nargs=4, default=None,
                    metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV",
                             "COPY_WRITE_CSV"),
                    help="write %s from rocprofv3 --pmc counter files of this "
                         "bench (--mode fp64) and of --copy-probe-gib, then "
                         "exit" % os.path.relpath(PROFILE_TRAFFIC, ROOT))
    args = ap.parse_args()

    if args.traffic_summary:
        probe = None
        if args.update_probe:
            probe = (args.update_probe[0] - args.nb, args.update_probe[1],
                     args.nb)
        traffic_summary(*args.traffic_summary, args.n, args.nb,
                        int(args.copy_probe_gib * 2 ** 30) // 8 * 8, probe)
        return
    if args.update_probe:
        update_probe(args.update_probe[0], args.update_probe[1], args.nb)
        return
    if args.mxp_probe:
        f, m, n, k = args.mxp_probe
        mxp_probe(f, int(m), int(n), int(k))
        return
    if args.copy_probe_gib > 0:
        copy_probe(args.copy_probe_gib)
        return
    if args.gpus != 1:
        print("bench.py: runs on exactly one GPU (got --gpus %d); "
              "multi-GPU runs are not part of it" % args.gpus, flush=True)
        sys.exit(2)

    import rochpl_amd as rp

    ctx = rp.Context(0)
    ctx.panel_kernel = {"auto": 2, "kernel": 1, "launches": 0}[args.panel]
    out = {"metric": METRIC}
    # legs after the first run only while the wall-clock budget leaves them
    # their expected time; a skipped leg is named in the line
    legs = [f for f in ("bf16", "fp8") if args.mode in ("all", f)]
    if args.mode == "adjoint" or (args.mode == "all" and args.adjoint):
        legs.append("adjoint")
    if args.mode != "adjoint" and not args.no_cpu:
        legs.append("cpu")
    skipped = []

    def leg_fits(leg):
        if remaining(args) >= LEG_SECONDS[leg]:
            return True
        log("%s leg skipped: %.0f s of the %.0f s budget left" % (
            leg, remaining(args), args.budget_s))
        skipped.append(leg)
        return False

    fp64 = None
    if args.mode in ("all", "fp64"):
        args.reserve_s = sum(LEG_SECONDS[leg] for leg in legs)
        fp64 = run_fp64(ctx, rp, args)
    mixed = {}
    for form in ("bf16", "fp8"):
        if form in legs and (args.mode == form or leg_fits(form)):
            try:
                mixed[form] = run_mixed(ctx, rp, form, args)
            except (SystemExit, rp.HplxError) as e:
                # a mixed form that fails is reported, not silently dropped;
                # the FP64 line stands on its own
                if args.mode == form:
                    raise
                log("%s form failed: %s" % (form, e))
                mixed[form] = {"error": str(e)}
    adjoint = None
    if "adjoint" in legs and (args.mode == "adjoint" or leg_fits("adjoint")):
        try:
            adjoint = run_adjoint(ctx, rp, args)
        except (SystemExit, rp.HplxError) as e:
            # reported in the line like a failing mixed form
            if args.mode == "adjoint":
                raise
            log("adjoint leg failed: %s" % e)
            adjoint = {"error": str(e)}
    ctx.close()

    if args.mode == "adjoint":
        # the adjoint leg alone: value = HPL-equivalent rate of its getrf
        a = adjoint
        out.update({
            "value": round(hpl_flops(args.adj_n) / (a["getrf_ms"] * 1e-3) / 1e12, 3),
            "unit": "TFLOP/s", "n_gpus": 1, "steps": a["config"]["reps"],
            "warmup": 1, "ms_per_step": round(a["getrf_ms"] + a["getrs_ms"] +
                                              a["adjoint_ms"], 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64",
            "data": "synthetic: HPL 64-bit LCG matgen, seed 100, plus "
                    "2 sqrt(N) I; B, G seeded normal",
            "config": a["config"],
            "roofline": {"bound": "hbm", "achieved": a["gradA_gbs"],
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": a["gradA_hbm_frac"], "traffic": None,
                         "kernel": "gradA = -gradB X^T (writes N x N)"},
            "adjoint": a, "cpu_baseline": None})
        print(json.dumps(out), flush=True)
        return

    if fp64 is not None:
        head = fp64
        out.update({
            "value": round(fp64["value"], 3), "unit": "TFLOP/s",
            "n_gpus": 1, "steps": fp64["steps_run"],
            "warmup": fp64["warmup_run"],
            "steps_requested": args.steps, "warmup_requested": args.warmup,
            "step_seconds": fp64["step_seconds"],
            "ms_per_step": round(fp64["t_mean"] * 1e3, 2),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64",
            "data": "synthetic: HPL 64-bit LCG matgen, seed 100, uniform "
                    "[-0.5,0.5], b = column N",
            "config": head["config"], "roofline": head["roofline"]})
        for k in ("swap", "panel", "residual", "residual_r0_form",
                  "tflops_ref_formula", "pivot_prefix_check"):
            out[k] = head[k]
    else:
        form = args.mode
        m = mixed[form]
        ur = m["update_roofline"]
        bound = ur["bound"]
        out.update({
            "value": m["value"], "unit": "TFLOP/s", "n_gpus": 1, "steps": 1,
            "warmup": args.mixed_warmup, "ms_per_step": m["ms"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": {"bf16": "bf16+f32+f64", "fp8": "fp8+f16+f32+f64"}[form],
            "data": "synthetic: HPL 64-bit LCG matgen, seed 100",
            "config": m["config"],
            "roofline": {
                "bound": bound,
                "achieved": ur["mfma_achieved"] if bound == "mfma"
                else ur["hbm_achieved_gbs"],
                "peak": ur["mfma_peak"] if bound == "mfma" else HBM_PEAK_GBS,
                "unit": "TFLOP/s" if bound == "mfma" else "GB/s",
                "frac": ur["mfma_frac"] if bound == "mfma" else ur["hbm_frac"],
                "traffic": None,
                "kernel": "low-precision trailing-update GEMM: "
                          + ur.get("kernel", "")}})
    for form, m in mixed.items():
        if fp64 is not None and "value" in m:
            m["vs_fp64"] = round(m["value"] / fp64["value"], 3)
        out[form] = m
    if adjoint is not None:
        out["adjoint"] = adjoint
    out["cpu_baseline"] = None
    if "cpu" in legs and leg_fits("cpu"):
        log("CPU baseline: N=%d NB=%d on %d threads (%d CPUs in the affinity "
            "mask)" % (args.cpu_n, args.cpu_nb, _CORES, _HOST_CPUS))
        out["cpu_baseline"] = cpu_baseline(
            args.cpu_n, args.cpu_nb, _CORES, _HOST_CPUS,
            min(args.cpu_min_s, max(remaining(args) - 5.0, 1.0)), 200)
    out["legs_skipped"] = skipped
    out["wall_s"] = round(time.monotonic() - _T0, 1)
    out["budget_s"] = args.budget_s
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
