#!/usr/bin/env python3
"""Benchmark of the dense FP64 path on one MI355X (gfx950).

Workload (one step): the blocked Cholesky DecomposeCHOL(A, LOWER, nb=256)
followed by SolveCHOL(B, A, LOWER) with 64 right-hand sides, n = 8192,
A = G*G^T/n + I with G of i.i.d. N(0,1) entries, B of i.i.d. N(0,1) entries,
column-major FP64. Inputs are generated on the GPU from fixed seeds and all
step copies are staged in HBM before the timed region, so the timed steps do
nothing but the factorization and the solve.

Prints one JSON line. `value` is the FP64 throughput of the step (LAPACK
flop counts: n^3/3 + n^2/2 + n/6 for DPOTRF plus 2*n^2*nrhs for DPOTRS).
`roofline` describes the dominant kernel (the SYRK trailing update of the
factorization), timed with HIP events on its own stream inside the timed
region. `cpu_baseline` times the CPU oracle (the reference's algorithm,
built from oracle/) on a bounded sample of the same workload.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("FP64 DGEMM/DPOTRF TFLOP/s vs peak; mixed-vs-FP64 solve time at 1e-12 bwd err")
PEAK_FP64_MFMA = 78.6  # TFLOP/s, MI355X dense FP64 matrix
# the trailing-update kernel of the Cholesky (triangle-only GEMM instance)
SYRK_KERNEL = "k_gemm<double, 64, 64, 16, 2, 2, 1, false, true, true, true>"
LOWER = 0x4


def potrf_flops(n):
    return n ** 3 / 3.0 + n ** 2 / 2.0 + n / 6.0


def potrs_flops(n, nrhs):
    return 2.0 * n * n * nrhs


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--nb", type=int, default=256)
    ap.add_argument("--nrhs", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the other workloads (DGEMM 1024^3, LU and mixed LU n=16384, "
                         "Cholesky adjoint n=4096, LDL and Bunch-Kaufman n=4096)")
    ap.add_argument("--profile-run", action="store_true",
                    help="one warm-up and one timed step plus a 256 MiB reference copy, for "
                         "rocprofv3 counter passes; prints no JSON")
    ap.add_argument("--only", choices=["dgemm", "lu", "mixed", "adjoint", "sympiv"],
                    help="run only this other workload (for profiling) and print its JSON")
    ap.add_argument("--kernel-stats", metavar="DB",
                    help="print the per-kernel duration table of a rocprofv3 --kernel-trace "
                         "database and exit")
    ap.add_argument("--counter-table", nargs="+", metavar="DB",
                    help="print, per kernel, the dispatch count and the average of every "
                         "counter in rocprofv3 --pmc databases, and exit")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_DB", "WRITE_DB"),
                    help="turn two rocprofv3 --pmc databases of a --profile-run (FETCH_SIZE, "
                         "then WRITE_SIZE) into profiles/traffic.json and exit")
    return ap.parse_args()


def main():
    args = parse()
    if args.traffic_from:
        return traffic_summary(*args.traffic_from)
    if args.kernel_stats:
        return kernel_stats(args.kernel_stats)
    if args.counter_table:
        return counter_table(args.counter_table)
    if args.profile_run:
        args.steps, args.warmup, args.no_cpu, args.no_extra = 1, 1, True, True
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path runs on a single MI355X)",
              file=sys.stderr)
        return 2
    import torch
    from matops_amd import _lib

    L = _lib.lib()
    if L.mops_device_count() < 1 or not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        return 1
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    n, nb, nrhs = args.n, args.nb, args.nrhs

    def colmajor(r, c):
        return torch.empty((c, r), dtype=torch.float64, device=dev).T

    def ok(rc, what):
        if rc != 0:
            raise RuntimeError(f"{what}: {rc} {_lib.last_error()}")

    gen = torch.Generator(device=dev)
    if args.only:
        if args.only == "dgemm":
            res = extra_dgemm(L, ok, colmajor, gen, stream, sp)
        elif args.only == "lu":
            res = extra_lu(L, ok, colmajor, torch, dev, sp)[0]
        elif args.only == "mixed":
            lu, _ = extra_lu(L, ok, colmajor, torch, dev, sp, reps=0)
            res = extra_mixed(L, ok, colmajor, torch, dev, sp, lapack_solution(torch, dev, colmajor),
                              lu["total_ms"], reps=0)
        elif args.only == "sympiv":
            res = extra_sym_piv(L, ok, colmajor, torch, dev, sp)
        else:
            res = extra_adjoint(L, ok, colmajor, torch, dev, sp)
        print(json.dumps({args.only: res}), flush=True)
        return 0

    # ---- inputs: A = G G^T / n + I (G = H^T, H ~ N(0,1) seed 3), B seed 4
    gen.manual_seed(3)
    H = colmajor(n, n)
    H.normal_(generator=gen)
    A0 = colmajor(n, n)
    ok(L.mops_d_dgemm(0x1, n, n, n, 1.0 / n, H.data_ptr(), n, H.data_ptr(), n, 0.0,
                      A0.data_ptr(), n, sp), "dgemm")
    A0 += torch.eye(n, dtype=torch.float64, device=dev)
    A0.copy_(torch.tril(A0) + torch.tril(A0, -1).T)
    del H
    gen.manual_seed(4)
    B0 = colmajor(n, nrhs)
    B0.normal_(generator=gen)
    nsteps = args.warmup + args.steps + 1  # + one traced step after the timed ones
    As = []
    Bs = []
    for _ in range(nsteps):
        a = colmajor(n, n)
        a.copy_(A0)
        b = colmajor(n, nrhs)
        b.copy_(B0)
        As.append(a)
        Bs.append(b)
    info = torch.full((nsteps,), -7, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()

    def step(i):
        ok(L.mops_d_dpotrf(LOWER, n, nb, As[i].data_ptr(), n,
                           ctypes.c_void_p(info.data_ptr() + 4 * i), sp), "dpotrf")
        ok(L.mops_d_dpotrs(LOWER, n, nrhs, As[i].data_ptr(), n, Bs[i].data_ptr(), n, sp),
           "dpotrs")

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.warmup, nsteps - 1):
        step(i)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # one more step, untimed, with HIP events around every trailing update
    # launch (on the stream it runs on): the dominant kernel's own rate
    L.mops_trace_enable(1)
    torch.cuda.synchronize()
    step(nsteps - 1)
    torch.cuda.synchronize()
    cnt = ctypes.c_int()
    ms = ctypes.c_double()
    fl = ctypes.c_double()
    ok(L.mops_trace_collect(0, ctypes.byref(cnt), ctypes.byref(ms), ctypes.byref(fl)), "trace")
    L.mops_trace_enable(0)

    if args.profile_run:
        # reference transfer of known size, 256 MiB read + 256 MiB written, by
        # a kernel that appears nowhere else in the run (elementwise negation)
        src = torch.empty(32 << 20, dtype=torch.float64, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        print("profile run done", flush=True)
        return 0

    # ---- correctness of the timed results (size-independent properties)
    inf = info.cpu().tolist()
    if any(v != -1 for v in inf):
        raise RuntimeError(f"factorization reported info {inf}")
    Lf = torch.tril(As[nsteps - 2])  # last timed step
    fres = (torch.linalg.norm(A0 - Lf @ Lf.T) / torch.linalg.norm(A0)).item()
    X = Bs[nsteps - 2]
    eta = (torch.linalg.norm(A0 @ X - B0) /
           (torch.linalg.norm(A0) * torch.linalg.norm(X) + torch.linalg.norm(B0))).item()
    eps = 2.0 ** -52
    if not (fres <= 10 * n * eps and eta <= 10 * n * eps):
        raise RuntimeError(f"residual check failed: factor {fres:.3e} solve {eta:.3e}")

    elapsed = t1 - t0
    step_flops = potrf_flops(n) + potrs_flops(n, nrhs)
    value = step_flops * args.steps / elapsed / 1e12
    achieved = fl.value / (ms.value * 1e-3) / 1e12 if ms.value > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"DecomposeCHOL(A, LOWER, nb={nb}) + SolveCHOL(B, A, LOWER), "
                        f"n={n}, {nrhs} RHS, A=G*G^T/n+I, G and B i.i.d. N(0,1)",
            "n": n, "nb": nb, "nrhs": nrhs, "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "mfma",
            "kernel": f"SYRK trailing update, {SYRK_KERNEL} (lower triangle), every launch",
            "achieved": round(achieved, 3),
            "peak": PEAK_FP64_MFMA,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64_MFMA, 4),
            "traffic": None,
            "launches": cnt.value,
            "avg_launch_us": round(ms.value * 1e3 / max(cnt.value, 1), 2),
        },
        "check": {"factor_residual": fres, "solve_backward_error": eta},
    }
    traffic = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(traffic):
        with open(traffic) as f:
            tj = json.load(f)
        if tj.get("kernel") == SYRK_KERNEL:  # measured on this kernel
            out["roofline"]["traffic"] = tj.get("syrk_update_bytes_per_launch")

    if not args.no_extra:
        del As, Bs
        torch.cuda.empty_cache()
        out["dgemm_1024"] = extra_dgemm(L, ok, colmajor, gen, stream, sp)
        out["lu_16384"], xlu = extra_lu(L, ok, colmajor, torch, dev, sp)
        xref = lapack_solution(torch, dev, colmajor)
        out["lu_16384"]["forward_error_vs_lapack"] = _fwd(xlu, xref)
        out["mixed_lu_16384"] = extra_mixed(L, ok, colmajor, torch, dev, sp, xref,
                                            out["lu_16384"]["total_ms"])
        out["chol_adjoint_4096"] = extra_adjoint(L, ok, colmajor, torch, dev, sp)
        out["ldl_bk_4096"] = extra_sym_piv(L, ok, colmajor, torch, dev, sp)

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(nb, n, nrhs)
    print(json.dumps(out), flush=True)
    return 0


def _events(torch):
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def extra_dgemm(L, ok, colmajor, gen, stream, sp):
    """DGEMM 1024^3, C = A*B, alpha = 1, beta = 0, A and B U[-1,1] (seeds 1, 2):
    time per launch of 20 launches back to back between two HIP events
    (the host's launch of one overlaps the previous one), median and minimum
    of 10 such rounds; beside it the median of 50 single launches, each
    between its own events (which adds the host's launch latency)."""
    import torch
    m = 1024
    gen.manual_seed(1)
    A = colmajor(m, m)
    A.uniform_(-1, 1, generator=gen)
    gen.manual_seed(2)
    Bm = colmajor(m, m)
    Bm.uniform_(-1, 1, generator=gen)
    C = colmajor(m, m)
    for _ in range(5):
        ok(L.mops_d_dgemm(0, m, m, m, 1.0, A.data_ptr(), m, Bm.data_ptr(), m, 0.0,
                          C.data_ptr(), m, sp), "dgemm")
    e0, e1 = _events(torch)
    burst = 20
    times = []
    for _ in range(10):
        e0.record(stream)
        for _ in range(burst):
            ok(L.mops_d_dgemm(0, m, m, m, 1.0, A.data_ptr(), m, Bm.data_ptr(), m, 0.0,
                              C.data_ptr(), m, sp), "dgemm")
        e1.record(stream)
        e1.synchronize()
        times.append(e0.elapsed_time(e1) / burst)
    single = []
    for _ in range(50):
        e0.record(stream)
        ok(L.mops_d_dgemm(0, m, m, m, 1.0, A.data_ptr(), m, Bm.data_ptr(), m, 0.0,
                          C.data_ptr(), m, sp), "dgemm")
        e1.record(stream)
        e1.synchronize()
        single.append(e0.elapsed_time(e1))
    times.sort()
    single.sort()
    med = times[len(times) // 2]
    return {"median_us": round(med * 1e3, 2), "min_us": round(times[0] * 1e3, 2),
            "tflops_median": round(2 * m ** 3 / (med * 1e-3) / 1e12, 3),
            "frac_of_peak": round(2 * m ** 3 / (med * 1e-3) / 1e12 / PEAK_FP64_MFMA, 4),
            "timing": f"per launch, {burst} launches back to back between HIP events",
            "single_launch_median_us": round(single[len(single) // 2] * 1e3, 2)}


def _lu_problem(torch, dev, colmajor, n, nrhs):
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    A = colmajor(n, n)
    A.normal_(generator=g)
    g.manual_seed(6)
    B = colmajor(n, nrhs)
    B.normal_(generator=g)
    return A, B


def _eta(torch, A, X, B):
    R = B - A @ X
    an = A.abs().sum(dim=1).max()
    return (R.abs().amax(dim=0) / (an * X.abs().amax(dim=0) + B.abs().amax(dim=0))).max().item()


def extra_lu(L, ok, colmajor, torch, dev, sp, n=16384, nrhs=16, reps=3):
    """DecomposeLU(A, piv, 256) + SolveLU, n = 16384, 16 RHS, N(0,1) (seeds 5, 6):
    best of `reps` wall times of the device calls (factor + solve)."""
    A0, B0 = _lu_problem(torch, dev, colmajor, n, nrhs)
    A = colmajor(n, n)
    ipiv = torch.zeros(n, dtype=torch.int64, device=dev)
    perm = torch.zeros(n, dtype=torch.int32, device=dev)
    st = torch.zeros(1, dtype=torch.int32, device=dev)
    best = None
    for _ in range(reps + 1):
        A.copy_(A0)
        X = B0.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ok(L.mops_d_dgetrf(n, n, 256, A.data_ptr(), n, ipiv.data_ptr(), perm.data_ptr(),
                           st.data_ptr(), sp), "dgetrf")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ok(L.mops_d_dgetrs(0, n, nrhs, A.data_ptr(), n, ipiv.data_ptr(), n, perm.data_ptr(),
                           X.data_ptr(), n, sp), "dgetrs")
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if best is None or t2 - t0 < best[0] + best[1]:
            best = (t1 - t0, t2 - t1)
    if st.item() != 0:
        raise RuntimeError("dgetrf: grid barrier timed out")
    eta = _eta(torch, A0, X, B0)
    flops = 2 * n ** 3 / 3 - n ** 2 / 2 + 5 * n / 6
    res = {"n": n, "nrhs": nrhs, "factor_ms": round(best[0] * 1e3, 2),
           "solve_ms": round(best[1] * 1e3, 2),
           "total_ms": round((best[0] + best[1]) * 1e3, 2),
           "factor_tflops": round(flops / best[0] / 1e12, 3),
           "backward_error": eta}
    if reps > 0:
        # one more factorization with HIP events around each panel
        # factorization and each trailing GEMM (on the streams they run on)
        A.copy_(A0)
        L.mops_trace_enable(1)
        torch.cuda.synchronize()
        ok(L.mops_d_dgetrf(n, n, 256, A.data_ptr(), n, ipiv.data_ptr(), perm.data_ptr(),
                           st.data_ptr(), sp), "dgetrf")
        torch.cuda.synchronize()
        for key, cls in (("trailing_gemm", 1), ("panel_factor", 2)):
            cnt, ms, fl = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
            ok(L.mops_trace_collect(cls, ctypes.byref(cnt), ctypes.byref(ms), ctypes.byref(fl)),
               "trace")
            res[key] = {"launches": cnt.value, "sum_ms": round(ms.value, 2)}
            if fl.value > 0 and ms.value > 0:
                res[key]["tflops"] = round(fl.value / ms.value / 1e9, 3)
        L.mops_trace_enable(0)
    return (res, X)


def _fwd(X, xref):
    return ((X - xref).abs().amax() / xref.abs().amax()).item()


def lapack_solution(torch, dev, colmajor, n=16384, nrhs=16):
    """x = A^-1 B of the LU workload by LAPACK (scipy dgesv, FP64, on the
    host cores), the reference solution for the forward errors."""
    import scipy.linalg as sla
    A0, B0 = _lu_problem(torch, dev, colmajor, n, nrhs)
    A = A0.cpu().numpy()
    B = B0.cpu().numpy()
    x = sla.solve(A, B, overwrite_a=True, check_finite=False)
    return torch.from_numpy(x).to(dev)


def extra_mixed(L, ok, colmajor, torch, dev, sp, xref, lu_total_ms, n=16384, nrhs=16, reps=2):
    """The same system by LU in FP32 (and in BF16) + FP64 refinement to
    eta <= 1e-12 (mops_d_dgesv_ir), next to the same entry point in FP64;
    the speed-up is taken against the plain FP64 DecomposeLU + SolveLU wall
    time (lu_total_ms). Forward errors are against LAPACK's FP64 solution."""
    from matops_amd import _lib
    A0, B0 = _lu_problem(torch, dev, colmajor, n, nrhs)
    res = {}
    for name, prec in (("fp64", 0), ("fp32_ir", 1), ("bf16_ir", 2)):
        best = None
        for _ in range(reps + 1):
            X = colmajor(n, nrhs)
            rep = _lib.IrReport()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ok(L.mops_d_dgesv_ir(n, nrhs, A0.data_ptr(), n, B0.data_ptr(), n, X.data_ptr(), n,
                                 prec, 1e-12, 10, ctypes.byref(rep), sp), "dgesv_ir")
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            if best is None or t < best[0]:
                best = (t, rep.iterations, rep.backward_error, rep.converged, rep.factor_ms,
                        rep.refine_ms)
        res[name] = {"wall_ms": round(best[0] * 1e3, 2), "iterations": best[1],
                     "backward_error": best[2], "converged": bool(best[3]),
                     "factor_ms": round(best[4], 2), "refine_ms": round(best[5], 2),
                     "forward_error_vs_lapack": _fwd(X, xref),
                     "backward_error_check": _eta(torch, A0, X, B0)}
    res["plain_fp64_lu_total_ms"] = lu_total_ms
    res["speedup_fp32_ir_vs_plain_fp64_lu"] = round(lu_total_ms / res["fp32_ir"]["wall_ms"], 3)
    return res


def extra_adjoint(L, ok, colmajor, torch, dev, sp, n=4096):
    """Cholesky adjoint n = 4096: A = Q diag(logspace(0, 4)) Q^T, W lower
    N(0,1); time of mops_d_dpotrf_adjoint with its default block order.
    "tflops" counts the 2n^3/3 flops of the blocked reverse sweep (the
    cheapest known form, so the rate is comparable across algorithms);
    "tflops_executed" the 5n^3/3 of the closed form the GPU runs."""
    g = torch.Generator(device="cpu").manual_seed(7)
    X = torch.randn((n, n), dtype=torch.float64, generator=g).to(dev)
    Q, _ = torch.linalg.qr(X)
    A = (Q * torch.logspace(0, 4, n, dtype=torch.float64, device=dev)) @ Q.T
    A = 0.5 * (A + A.T)
    W = torch.tril(torch.randn((n, n), dtype=torch.float64, generator=g)).to(dev)
    Lc = colmajor(n, n)
    Lc.copy_(A)
    ok(L.mops_d_dpotrf(LOWER, n, 256, Lc.data_ptr(), n, None, sp), "dpotrf")
    Lc.copy_(torch.tril(Lc))
    Ab = colmajor(n, n)
    best = None
    for _ in range(4):
        Ab.copy_(W)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ok(L.mops_d_dpotrf_adjoint(n, 0, Lc.data_ptr(), n, Ab.data_ptr(), n, sp), "adjoint")
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        best = t if best is None else min(best, t)
    return {"n": n, "ms": round(best * 1e3, 2),
            "tflops": round(2 * n ** 3 / 3 / best / 1e12, 3),
            "tflops_executed": round(5 * n ** 3 / 3 / best / 1e12, 3),
            "symmetric": bool(torch.equal(Ab, Ab.T))}


def extra_sym_piv(L, ok, colmajor, torch, dev, sp, n=4096, nrhs=16, reps=3):
    """DecomposeLDL(A, W, ipiv, LOWER, 0) + SolveLDL and DecomposeBK(A, W, ipiv,
    LOWER, 0) + SolveBK at n = 4096, 16 RHS, A = G + G^T with G ~ N(0,1)
    (diagonal scaled by 0.05 for Bunch-Kaufman, which then takes 2x2 pivots):
    best of `reps` device times (HIP events on the launch stream) of the
    factorization and of the solve; TFLOP/s with n^3/3 flops (the
    factorization) and 2 n^2 nrhs (the solve). Checked by the solve's
    backward error."""
    g = torch.Generator(device="cpu").manual_seed(9)
    G = torch.randn((n, n), dtype=torch.float64, generator=g).to(dev)
    S = G + G.T
    del G
    B0 = torch.randn((n, nrhs), dtype=torch.float64, generator=g).to(dev)
    strm = torch.cuda.current_stream()  # the stream sp names
    out = {"n": n, "nrhs": nrhs}
    for name, scale, fac, sol in (("ldl", 1.0, "mops_d_dsytrf_ldl", "mops_d_dsytrs_ldl"),
                                  ("bk", 0.05, "mops_d_dsytrf_bk", "mops_d_dsytrs_bk")):
        A0 = S.clone()
        A0.diagonal().mul_(scale)
        A = colmajor(n, n)
        X = colmajor(n, nrhs)
        ipiv = torch.zeros(n, dtype=torch.int64, device=dev)
        tf = ts = None
        for _ in range(reps):
            A.copy_(A0)
            X.copy_(B0)
            torch.cuda.synchronize()
            e0, e1 = _events(torch)
            e2 = torch.cuda.Event(enable_timing=True)
            e0.record(strm)
            if name == "ldl":
                ok(getattr(L, fac)(LOWER, n, 0, A.data_ptr(), n, ipiv.data_ptr(), n, None, sp),
                   fac)
            else:
                ok(getattr(L, fac)(LOWER, n, 0, A.data_ptr(), n, ipiv.data_ptr(), n, sp), fac)
            e1.record(strm)
            ok(getattr(L, sol)(LOWER, n, nrhs, A.data_ptr(), n, ipiv.data_ptr(), n,
                               X.data_ptr(), n, sp), sol)
            e2.record(strm)
            torch.cuda.synchronize()
            f, v = e0.elapsed_time(e1), e1.elapsed_time(e2)
            tf = f if tf is None else min(tf, f)
            ts = v if ts is None else min(ts, v)
        R = A0 @ X - B0
        eta = (R.abs().max() / (A0.abs().sum(dim=1).max() * X.abs().max() +
                                B0.abs().max())).item()
        out[name] = {"factor_ms": round(tf, 2), "solve_ms": round(ts, 3),
                     "factor_tflops": round(n ** 3 / 3 / (tf * 1e-3) / 1e12, 3),
                     "solve_tflops": round(2 * n * n * nrhs / (ts * 1e-3) / 1e12, 3),
                     "two_by_two_blocks": int((ipiv < 0).sum().item()) // 2,
                     "solve_backward_error": eta}
    return out


def kernel_stats(db):
    """Per-kernel calls / total / average / min / max duration (us) and share,
    from the kernels view of a rocprofv3 database."""
    import sqlite3
    con = sqlite3.connect(db)
    rows = con.execute(
        "select name, count(*), sum(end - start), avg(end - start), min(end - start), "
        "max(end - start) from kernels group by name order by sum(end - start) desc").fetchall()
    tot = sum(r[2] for r in rows) or 1
    print(f"{'calls':>6} {'total_us':>12} {'avg_us':>10} {'min_us':>10} {'max_us':>10} {'pct':>6}  kernel")
    for name, cnt, sm, av, mn, mx in rows:
        short = name.replace("mops::(anonymous namespace)::", "")
        if len(short) > 110:
            short = short[:107] + "..."
        print(f"{cnt:6d} {sm / 1e3:12.1f} {av / 1e3:10.2f} {mn / 1e3:10.2f} {mx / 1e3:10.2f} "
              f"{100 * sm / tot:6.2f}  {short}")
    return 0


def counter_table(dbs):
    """Per kernel and counter: dispatches, average and total value, from the
    counters_collection view of rocprofv3 --pmc databases."""
    import sqlite3
    print(f"{'counter':28} {'calls':>6} {'average':>16} {'total':>18}  kernel")
    for db in dbs:
        con = sqlite3.connect(db)
        rows = con.execute(
            "select counter_name, kernel_name, count(*), avg(value), sum(value) "
            "from counters_collection group by counter_name, kernel_name "
            "order by counter_name, sum(value) desc").fetchall()
        for cname, kname, cnt, av, tot in rows:
            short = kname.replace("mops::(anonymous namespace)::", "")
            if len(short) > 90:
                short = short[:87] + "..."
            print(f"{cname:28} {cnt:6d} {av:16.1f} {tot:18.1f}  {short}")
    return 0


def traffic_summary(fetch_db, write_db):
    """Per-dispatch FETCH_SIZE / WRITE_SIZE (KiB) of the SYRK trailing update
    and of the reference copy, from two rocprofv3 --pmc databases of a
    --profile-run; HBM bytes per SYRK launch = counter * (copy bytes / copy
    counter). Writes profiles/traffic.json."""
    import sqlite3

    def per_kernel(db, counter):
        con = sqlite3.connect(db)
        return con.execute("select kernel_name, value from counters_collection "
                           "where counter_name = ?", (counter,)).fetchall()

    out = {}
    for db, counter in ((fetch_db, "FETCH_SIZE"), (write_db, "WRITE_SIZE")):
        rows = per_kernel(db, counter)
        syrk = [v for name, v in rows if SYRK_KERNEL in name]
        copy = [v for name, v in rows if "neg" in name.lower()]
        if not syrk or not copy:
            raise RuntimeError(f"{counter}: kernels not found in {db}")
        out[counter] = {"syrk_kib_avg": sum(syrk) / len(syrk), "syrk_launches": len(syrk),
                        "copy_kib_avg": sum(copy) / len(copy), "copy_launches": len(copy)}
    copy_bytes = 256 << 20
    fs = out["FETCH_SIZE"]
    ws = out["WRITE_SIZE"]
    f_scale = copy_bytes / (fs["copy_kib_avg"] * 1024)
    w_scale = copy_bytes / (ws["copy_kib_avg"] * 1024)
    traffic = fs["syrk_kib_avg"] * 1024 * f_scale + ws["syrk_kib_avg"] * 1024 * w_scale
    res = {"kernel": SYRK_KERNEL, "syrk_update_bytes_per_launch": traffic, "fetch_scale": f_scale,
           "write_scale": w_scale, "counters": out}
    with open(os.path.join(ROOT, "profiles", "traffic.json"), "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res, indent=1))
    return 0


def cpu_baseline(nb, n, nrhs):
    """The CPU oracle (the reference's calgo kernels under the restated Go
    glue) on the host: the Cholesky workload itself on one core (value), the
    DGEMM 1024^3 and a bounded LU sample (n = 4096, 16 RHS) with one and
    with all host cores as Mult workers, and LDL / Bunch-Kaufman with their
    solves at n = 4096 on one core (the workload of the GPU leg)."""
    try:
        from oracle import oracle as orc
    except Exception as e:  # the baseline is reported, never required
        return {"value": None, "unit": "TFLOP/s", "error": f"oracle unavailable: {e}"}
    out = orc.bench_cholesky_sample(nb=nb, n=n, nrhs=nrhs)
    ncores, _ = orc.host_cores()
    out["other_workloads"] = [orc.bench_mult_and_lu(1)]
    if ncores > 1:
        out["other_workloads"].append(orc.bench_mult_and_lu(ncores))
    out["ldl_bk_4096"] = orc.bench_sym_piv_sample(n=4096)
    return out


def _release():
    """Free this thread's device workspace and streams before the runtime
    shuts down."""
    try:
        import torch
        from matops_amd import _lib
        if torch.cuda.is_available():
            torch.cuda.synchronize()
            _lib.lib().mops_thread_release()
    except Exception:
        pass


if __name__ == "__main__":
    rc = main()
    _release()
    sys.exit(rc)
