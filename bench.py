"""Benchmark of the DAFMM Lippmann-Schwinger solver on one GPU.

Default line (the metric's configuration): C3, uniform depth 7 (589,824
unknowns), kappa 160, 8 seeded Gaussian bumps, 64 plane waves at 2 pi j / 64.
A step is one batch of the 64 systems solved by FP32-inner batched GMRES inside
FP64 iterative refinement, every system verified by its FP64 true residual;
`value` is solves/s, beside plain FP64 GMRES(50) timed in the same run on the
same GPU. Operators, right-hand sides and solutions are resident in HBM before
timing starts. Other workloads: --workload c2 | learned | ragged | setup | readme.
Prints one JSON line.
"""
import argparse
import copy
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import lippmann_schwinger_solver_amd as lsg  # noqa: E402

METRIC = "FP64-verified solves/s at 1e-8 residual vs plain GPU GMRES; matvec HBM GB/s"
HBM_PEAK_GBS = 8000.0


def c2_inputs(seed=0):
    inp = lsg.InputsToDFMM(kappa=80.0, nLevelsUniform=6, L=0.5, nChebNodes=6, degreeOfBases=6, TOL_POW=10,
                           nCones_LFR=16)
    inp.contrast = lsg.bumps_from_seed(seed, 5)
    # plane-wave angle: first draw of the SplitMix64 stream seeded with 1000 + seed
    s = (1000 + seed + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = s
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    z ^= z >> 31
    inp.inc_angle = 2 * np.pi * ((z >> 11) * 2.0 ** -53)
    return inp


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cpu_timed_solve(inp, tol, restart=50):
    """The oracle (CPU restatement of the reference, FP64, OpenMP) on the host
    cores: its operator build (timed on its own), the median of 5 matvecs after
    one warm-up, and one restarted GMRES(restart) solve from x0 = 0, cycle by
    cycle until the true FP64 relative residual (one more oracle matvec per
    cycle) is at tol, timed as a whole (GMRES/GMRES.cpp:307-319 with the
    reference's timers, :383-399)."""
    import oracle
    t0 = time.time()
    o = oracle.Oracle(inp.to_c())
    setup = time.time() - t0
    x = np.random.default_rng(0).standard_normal(o.N) + 0j
    o.matvec(x)
    mv = []
    for _ in range(5):
        t1 = time.perf_counter()
        o.matvec(x)
        mv.append(time.perf_counter() - t1)
    b = o.rhs()
    bn = np.linalg.norm(b)
    t1 = time.perf_counter()
    xk, cycles, its = np.zeros(o.N, dtype=complex), 0, 0
    rel = 1.0
    while cycles < 40:
        xk, it = o.gmres_cycle(b, xk, restart)
        cycles += 1
        its += it
        rel = float(np.linalg.norm(b - o.matvec(xk)) / bn)
        if rel <= tol:
            break
    t_solve = time.perf_counter() - t1
    o.close()
    return {"value": 1.0 / t_solve if rel <= tol else 0.0, "unit": "solves/s", "cores": _threads(), "kind": "port",
            "solve_s": t_solve, "solve_cycles": cycles, "solve_arnoldi_steps": its, "solve_rel_residual": rel,
            "matvec_ms_median_of_5": 1e3 * float(np.median(mv)), "setup_s": setup}


def measured_traffic(workload, pass_name):
    """HBM bytes per launch of the dominant pass from the committed FETCH_SIZE /
    WRITE_SIZE counter runs of this workload (scaled by a plain copy's
    counters); None if absent."""
    f = os.path.join(ROOT, "profiles", f"{workload}_leaf_m2l_pmc_traffic.json")
    try:
        with open(f) as fh:
            d = json.load(fh)
    except (OSError, ValueError):
        return None
    return d.get("traffic_bytes") if d.get("pass") == pass_name else None


class DeviceBlock:
    """Column-major complex128 block in device memory (plain hipMalloc), so
    that timed solves start from inputs already resident in HBM."""

    def __init__(self, host):
        import ctypes as C
        self.hip = C.CDLL("libamdhip64.so")
        self.shape = host.shape
        self.nbytes = host.nbytes
        self.ptr = C.c_void_p()
        if self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
        self.upload(host)

    def upload(self, host):
        import ctypes as C
        h = np.ascontiguousarray(host)
        if self.hip.hipMemcpy(self.ptr, h.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes), 1) != 0:
            raise RuntimeError("hipMemcpy H2D failed")

    def zero(self):
        import ctypes as C
        if self.hip.hipMemset(self.ptr, 0, C.c_size_t(self.nbytes)) != 0:
            raise RuntimeError("hipMemset failed")

    def zero_range(self, offset, nbytes):
        import ctypes as C
        if self.hip.hipMemset(C.c_void_p(self.ptr.value + offset), 0, C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMemset failed")

    def download(self):
        import ctypes as C
        out = np.empty(self.shape, dtype=np.complex128)
        if self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.nbytes), 2) != 0:
            raise RuntimeError("hipMemcpy D2H failed")
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


MFMA_F64_PEAK_TFLOPS = 78.6
MFMA_F32_PEAK_TFLOPS = 157.3
MFMA_BF16_PEAK_TFLOPS = 2500.0   # dense BF16 MFMA (no sparsity)


def batched_leg(S, info, args, st_plain, plain_sps, restart=50, restart32=30):
    """64 plane waves at 2 pi j / 64 on the same geometry, solved together by
    batched GMRES (one FP64 MFMA block matvec per Arnoldi step); the right-hand
    sides are resident in HBM before timing; every system's FP64 true residual
    must reach tol. `restart` is the FP64 batched restart length, `restart32`
    the inner solves' (FP32 pass state, split-BF16 operators, FP16 basis; one
    cycle per refinement): the Gram-Schmidt of 64 systems streams 64 basis
    vectors per column, so the low-precision batch prefers shorter cycles
    (restart sweeps are recorded with the workload results under profiles/)."""
    angles = 2 * np.pi * np.arange(64) / 64
    Bh = np.ascontiguousarray(S.rhs_angles(angles).T)      # 64 x N: column q at q*N
    Bd, Xd = DeviceBlock(Bh), DeviceBlock(np.zeros_like(Bh))
    ptrs = (Bd.ptr.value, Xd.ptr.value)
    Xd.zero()
    S.gmres_batch(None, device_ptrs=ptrs, nrhs=64, ld=S.N, restart_len=restart, max_restarts=40, tol=args.tol)  # warm-up
    log("batched FP64: warm-up done")
    t = 0.0
    for _ in range(args.steps):
        Xd.zero()
        t0 = time.perf_counter()
        _, st, rel = S.gmres_batch(None, device_ptrs=ptrs, nrhs=64, ld=S.N, restart_len=restart, max_restarts=40,
                                   tol=args.tol)
        t += time.perf_counter() - t0
        log(f"batched FP64 batch: {st}")
        if st["status"] != 0 or st["converged"] != 64 or not rel.max() <= args.tol:
            raise RuntimeError(f"batched solve did not verify: {st}")
    X = Xd.download().T
    # the same batch with the low-precision inner solves inside FP64 refinement
    Xd.zero()
    S.gmres_batch(None, device_ptrs=ptrs, nrhs=64, ld=S.N, restart_len=restart32, max_restarts=1, tol=args.tol,
                  mode="fp32ir", inner_tol=1e-5, max_outer=40)   # warm-up
    log("batched FP32-IR: warm-up done")
    t32 = 0.0
    for _ in range(args.steps):
        Xd.zero()
        t0 = time.perf_counter()
        _, st32, rel32 = S.gmres_batch(None, device_ptrs=ptrs, nrhs=64, ld=S.N, restart_len=restart32,
                                       max_restarts=1, tol=args.tol, mode="fp32ir", inner_tol=1e-5, max_outer=40)
        t32 += time.perf_counter() - t0
        log(f"batched FP32-IR batch: {st32}")
        if st32["status"] != 0 or st32["converged"] != 64 or not rel32.max() <= args.tol:
            raise RuntimeError(f"batched FP32-IR solve did not verify: {st32}")
    X32 = Xd.download().T
    Bd.free()
    Xd.free()
    # independent FP64 residuals of a few systems through the single-vector matvec
    spot = max(float(np.linalg.norm(S.MatVecProduct(X[:, j]) - Bh[j]) / np.linalg.norm(Bh[j])) for j in (0, 21, 42, 63))
    mv_block = S.time_matvec_block(reps=5)
    flops = 8.0 * 64 * (info["E_far"] + info["E_near"])
    sps = 64 * args.steps / t
    sps32 = 64 * args.steps / t32
    spot32 = max(float(np.linalg.norm(S.MatVecProduct(X32[:, j]) - Bh[j]) / np.linalg.norm(Bh[j]))
                 for j in (0, 21, 42, 63))
    return {"solves_per_s": sps, "speedup_vs_plain_fp64": sps / plain_sps, "ms_per_batch": 1e3 * t / args.steps,
            "restart_len": restart,
            "arnoldi_steps": st["iterations"], "cycles": st["restarts"], "max_rel_residual": float(rel.max()),
            "spot_check_rel_residual": spot, "block_matvec_ms": mv_block,
            "block_matvec_us_per_rhs": 1e3 * mv_block / 64,
            "block_matvec_mfma": {"achieved_tflops": flops / (mv_block * 1e-3) / 1e12, "peak": MFMA_F64_PEAK_TFLOPS,
                                  "frac": flops / (mv_block * 1e-3) / 1e12 / MFMA_F64_PEAK_TFLOPS,
                                  "flops_per_matvec": flops},
            "fp32_inner": {"solves_per_s": sps32, "speedup_vs_plain_fp64": sps32 / plain_sps, "restart_len": restart32,
                           "speedup_vs_batched_fp64": sps32 / sps,
                           "block_matvec_fp32_ms": S.time_matvec_block(reps=5, fp32=True),
                           "ms_per_batch": 1e3 * t32 / args.steps, "inner_arnoldi_steps": st32["iterations"],
                           "refinements": st32["refinements"], "max_rel_residual": float(rel32.max()),
                           "spot_check_rel_residual": spot32,
                           "max_rel_diff_vs_fp64_batch": max(float(np.linalg.norm(X32[:, j] - X[:, j]) /
                                                                   np.linalg.norm(X[:, j])) for j in range(64))}}


def pmc_probe(S):
    """A plain device-to-device copy of 256 MiB (its FETCH_SIZE / WRITE_SIZE
    calibrate the counters: 262144 KiB each) followed by 3 matvecs."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    nb = 256 << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nb)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nb)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nb)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nb), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    ms = S.time_matvec(reps=3)
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"pmc probe: copy {nb >> 20} MiB, matvec {ms * 1e3:.1f} us")


def pmc_probe_block(S):
    """3 block matvecs (64 right-hand sides) on FP64 MFMA, then 3 on FP32 MFMA,
    then 3 in split BF16, each after one warm-up, for counter runs; the
    algorithmic flops of one are printed for the ratio."""
    ms = S.time_matvec_block(reps=3)
    S.set_inner_arith("fp32")
    ms32 = S.time_matvec_block(reps=3, fp32=True)
    S.set_inner_arith("split_bf16")
    mssb = S.time_matvec_block(reps=3, fp32=True)
    info = S.info()
    flops = 8.0 * 64 * (info["E_far"] + info["E_near"])
    log(f"pmc block probe: block matvec {ms:.3f} ms FP64, {ms32:.3f} ms FP32, {mssb:.3f} ms split BF16, "
        f"{flops:.4e} algorithmic flops each")


def setup_record(wall_s, info, inp):
    """Operator build time and its phases as the library reports them: the
    host build (OpenMP; the operator every timed solve uses) and, beside it,
    the same build with NCA skeletons and operator entries on the GPU."""
    ph = info["setup_phase_seconds"]
    rec = {"wall_s": wall_s, "host_build_s": info["setup_seconds"], "far_tables_s": ph[0], "nca_skeletons_s": ph[1],
           "operator_fill_s": ph[2], "near_tables_s": ph[3],
           "threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count()))}
    dev_inp = copy.copy(inp)
    dev_inp.setup_device = 1
    d = lsg.HostBuild(dev_inp).info()
    dp = d["setup_phase_seconds"]
    rec["device_build"] = {"build_s": d["setup_seconds"], "nca_skeletons_s": dp[1], "operator_fill_s": dp[2],
                           "E_far": int(d["E_far"])}
    return rec


def seeded_angle(seed):
    """First draw of the SplitMix64 stream seeded with 1000 + seed, scaled to [0, 2 pi)."""
    z = (1000 + seed + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    z ^= z >> 31
    return 2 * np.pi * ((z >> 11) * 2.0 ** -53)


C3_WORKLOAD = ("C3, 64 angles: uniform depth 7 (16384 leaves x 36 Chebyshev nodes = 589824 unknowns), kappa 160, "
               "8 seeded Gaussian bumps, NCA tol 1e-10; 64 plane waves at 2 pi j/64 solved by batched GMRES(24) "
               "inner solves (FP32 pass state, split-BF16 operators on BF16 MFMA, FP16 basis) inside an FP64 "
               "flexible-GMRES(30) outer iteration to FP64-verified relative residual {tol:g}, against plain FP64 "
               "GMRES(50) on the same GPU")


def _relres(S, x, b):
    return float(np.linalg.norm(S.MatVecProduct(x) - b) / np.linalg.norm(b))


def bench_c3(args):
    """Config C3 with the 64 incident angles of the batched configuration: one
    step is one batch of 64 systems (plane waves at 2 pi j / 64, x0 = 0) solved
    by batched GMRES with the FP32 block operator on FP32 MFMA inside an FP64
    flexible GMRES outer iteration (each outer step: one FP32 GMRES(restart)
    solve, one FP64 block matvec, the corrections re-weighted by the outer
    least-squares problem); every system must reach the FP64 true relative
    residual tol. The baseline, timed in the same run on the same GPU, is plain
    FP64 GMRES(50) on single systems of the batch. Right-hand sides and
    solutions stay in HBM for the timed region. Beside the line: the same batch
    by plain FP64 refinement (x += d) around the same inner solves, and by FP64
    batched GMRES."""
    inp = c3_inputs(0)
    t0 = time.time()
    S = lsg.DFMM(inp)
    S.FMMTreeInitialisation()
    info = S.info()
    setup_s = time.time() - t0
    N = S.N
    log(f"C3 setup {setup_s:.1f} s: N={N} E_far={info['E_far']} bytes/matvec={info['bytes_matvec']}")
    n_changed, report = S.autotune(reps=3)
    log(f"autotune ({n_changed} changed):\n{report}")
    if args.pmc_probe_block:
        pmc_probe_block(S)
        return S.close()
    if args.pmc_probe:
        pmc_probe(S)
        return S.close()
    angles = 2 * np.pi * np.arange(64) / 64
    Bh = np.ascontiguousarray(S.rhs_angles(angles).T)          # 64 x N: system q at q*N
    Bd, Xd = DeviceBlock(Bh), DeviceBlock(np.zeros_like(Bh))
    col = lambda blk, q: blk.ptr.value + 16 * N * q

    # plain FP64 GMRES(50), one system at a time (systems 0, 16, 32, 48)
    plain_q = (0, 16, 32, 48)

    def plain(q):
        Xd.zero_range(16 * N * q, 16 * N)
        _, st = S.gmres_ex(None, mode="fp64", restart_len=50, max_restarts=40, tol=args.tol,
                           device_ptrs=(col(Bd, q), col(Xd, q)))
        if st["status"] != 0 or st["converged"] != 1 or not st["rel_residual"] <= args.tol:
            raise RuntimeError(f"plain solve of system {q} did not verify: {st}")
        return st

    plain(0)
    S._check(S._lib.lsg_sync(S._ctx))
    t0 = time.perf_counter()
    plain_stats = [plain(q) for q in plain_q]
    S._check(S._lib.lsg_sync(S._ctx))
    t_plain = time.perf_counter() - t0
    plain_sps = len(plain_q) / t_plain
    log(f"plain FP64 GMRES(50): {len(plain_q)} solves in {t_plain:.3f} s ({plain_sps:.3f} solves/s); "
        f"steps {[st['iterations'] for st in plain_stats]}")

    def batch(mode):
        Xd.zero()
        kw = {"fp32fgmres": dict(mode="fp32fgmres", inner_tol=1e-5, max_outer=args.outer, max_restarts=1),
              "fp32ir": dict(mode="fp32ir", inner_tol=1e-5, max_outer=60, max_restarts=1),
              "fp64": dict(mode="fp64", max_restarts=40)}[mode]
        _, st, rel = S.gmres_batch(None, device_ptrs=(Bd.ptr.value, Xd.ptr.value), nrhs=64, ld=N,
                                   restart_len=args.restart, tol=args.tol, **kw)
        if st["status"] != 0 or st["converged"] != 64 or not rel.max() <= args.tol:
            raise RuntimeError(f"batched {mode} solve did not verify: {st}")
        return st, rel

    for _ in range(args.warmup):
        st, _ = batch("fp32fgmres")
        log(f"warm-up batch: {st}")
    S._check(S._lib.lsg_sync(S._ctx))
    t0 = time.perf_counter()
    steps = []
    for _ in range(args.steps):
        steps.append(batch("fp32fgmres"))
    S._check(S._lib.lsg_sync(S._ctx))
    elapsed = time.perf_counter() - t0
    value = 64 * args.steps / elapsed
    st32, rel32 = steps[-1]
    log(f"timed: {args.steps} batches of 64 in {elapsed:.3f} s ({value:.2f} solves/s); last {st32}")
    log("block operator choices:\n" + "\n".join(l for l in S.tuning().splitlines() if "block matvec" in l))
    X32 = Xd.download()
    recomputed = [_relres(S, X32[q], Bh[q]) for q in range(64)]
    log(f"recomputed FP64 residuals: max {max(recomputed):.3e}")

    # the same inner solves in plain FP64 refinement (x += d), one timed batch
    batch("fp32ir")
    S._check(S._lib.lsg_sync(S._ctx))
    t0 = time.perf_counter()
    st_ir, rel_ir = batch("fp32ir")
    S._check(S._lib.lsg_sync(S._ctx))
    t_ir = time.perf_counter() - t0

    # the line's solver with FP32 MFMA operators in the inner solves, one timed batch
    S.set_inner_arith("fp32")
    S._check(S._lib.lsg_sync(S._ctx))
    t0 = time.perf_counter()
    st_f32, rel_f32 = batch("fp32fgmres")
    S._check(S._lib.lsg_sync(S._ctx))
    t_f32 = time.perf_counter() - t0
    S.set_inner_arith("split_bf16")

    # the same batch with the FP64 block operator (no refinement), one timed batch
    batch("fp64")
    S._check(S._lib.lsg_sync(S._ctx))
    t0 = time.perf_counter()
    st64, rel64 = batch("fp64")
    S._check(S._lib.lsg_sync(S._ctx))
    t64 = time.perf_counter() - t0
    X64 = Xd.download()
    Bd.free()
    Xd.free()

    # plain matvec: whole-matvec rate and per-pass device time (HIP events on
    # the operator's stream, passes serialised)
    total_ms = S.time_matvec(reps=args.profile_reps)
    passes = S.profile_passes(reps=args.profile_reps)
    dom = max(passes, key=lambda p: p[1])
    achieved = dom[2] / (dom[1] * 1e-3) / 1e9
    mv_gbs = info["bytes_matvec"] / (total_ms * 1e-3) / 1e9
    log(f"matvec {total_ms * 1e3:.1f} us ({mv_gbs:.0f} GB/s algorithmic)")
    for p in passes:
        log(f"  pass {p[0]:10s} {p[1] * 1e3:9.1f} us  {p[2] / 1e6:9.2f} MB  {p[2] / (p[1] * 1e-3) / 1e9:8.1f} GB/s")
    mv32 = S.time_matvec_fp32(reps=args.profile_reps)
    blk64 = S.time_matvec_block(reps=5)
    S.set_inner_arith("fp32")
    blk32 = S.time_matvec_block(reps=5, fp32=True)
    S.set_inner_arith("split_bf16")
    blksb = S.time_matvec_block(reps=5, fp32=True)
    flops = 8.0 * 64 * (info["E_far"] + info["E_near"])
    out = {"metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "c64 inner (FP32 pass state, split-BF16 operators on BF16 MFMA, FP16 basis) + c128 outer",
           "data": "synthetic",
           "config": {"workload": C3_WORKLOAD.format(tol=args.tol), "N": N, "systems_per_step": 64,
                      "restart_len": args.restart, "outer_krylov_dim": args.outer, "inner_tol": 1e-5,
                      "inner_arith": "split BF16 (a = ah + al in bf16, products (ah + al)(bh + bl) as four BF16 MFMA terms, FP32 sums)",
                      "inner_arnoldi_steps_per_batch": st32["iterations"], "outer_steps": st32["refinements"],
                      "max_rel_residual": float(rel32.max()), "recomputed_max_rel_residual": max(recomputed),
                      "inputs": "B, X resident in HBM (device pointers)"},
           "speedup_vs_plain_fp64": value / plain_sps,
           "plain_fp64": {"solves_per_s": plain_sps, "systems": list(plain_q), "restart_len": 50,
                          "arnoldi_steps": [st["iterations"] for st in plain_stats],
                          "max_rel_residual": max(st["rel_residual"] for st in plain_stats)},
           "fp32_inner_plain_refinement": {"solves_per_s": 64 / t_ir, "speedup_vs_plain_fp64": 64 / t_ir / plain_sps,
                                           "inner_arnoldi_steps": st_ir["iterations"],
                                           "refinements": st_ir["refinements"],
                                           "max_rel_residual": float(rel_ir.max())},
           "fp32_mfma_inner": {"solves_per_s": 64 / t_f32, "speedup_vs_plain_fp64": 64 / t_f32 / plain_sps,
                               "inner_arnoldi_steps": st_f32["iterations"], "outer_steps": st_f32["refinements"],
                               "max_rel_residual": float(rel_f32.max())},
           "batched_fp64": {"solves_per_s": 64 / t64, "speedup_vs_plain_fp64": 64 / t64 / plain_sps,
                            "arnoldi_steps": st64["iterations"], "max_rel_residual": float(rel64.max()),
                            "max_rel_diff_fp32_inner_vs_fp64": max(float(np.linalg.norm(X32[q] - X64[q]) /
                                                                        np.linalg.norm(X64[q])) for q in range(64))},
           "setup": setup_record(setup_s, info, inp),
           "matvec": {"ms": total_ms, "algorithmic_bytes": int(info["bytes_matvec"]), "GBps": mv_gbs,
                      "frac_hbm": mv_gbs / HBM_PEAK_GBS, "E_far": int(info["E_far"]), "fp32_ms": mv32},
           "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("c3", dom[0]),
                        "algorithmic_bytes_per_launch": int(dom[2]), "launch_ms": dom[1]},
           "block_matvec": {"fp64_ms": blk64, "fp32_ms": blk32, "flops": flops,
                            "fp64_tflops": flops / (blk64 * 1e-3) / 1e12, "fp64_mfma_peak": MFMA_F64_PEAK_TFLOPS,
                            "fp64_frac": flops / (blk64 * 1e-3) / 1e12 / MFMA_F64_PEAK_TFLOPS,
                            "fp32_tflops": flops / (blk32 * 1e-3) / 1e12, "fp32_mfma_peak": MFMA_F32_PEAK_TFLOPS,
                            "fp32_frac": flops / (blk32 * 1e-3) / 1e12 / MFMA_F32_PEAK_TFLOPS,
                            "split_bf16_ms": blksb, "split_bf16_tflops": flops / (blksb * 1e-3) / 1e12,
                            "split_bf16_mfma_peak": MFMA_BF16_PEAK_TFLOPS / 4,
                            "split_bf16_frac": 4 * flops / (blksb * 1e-3) / 1e12 / MFMA_BF16_PEAK_TFLOPS},
           "cpu_baseline": None}
    S.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_c3(inp, X32, Bh, args.tol)
        out["config"]["max_rel_residual_vs_oracle"] = out["cpu_baseline"].pop("results_max_rel_residual_vs_oracle")
    print(json.dumps(out), flush=True)


def cpu_baseline_c3(inp, X, B, tol):
    """CPU side of the default line: the oracle of C3 (operator build and the
    median of 5 matvecs, on the host cores) checks every one of the 64 results
    (FP64 residual through the oracle's own operator), then a timed oracle
    GMRES(50) solve of C2 to tol gives the CPU solve rate (a C3 solve, about
    450 oracle matvecs, does not fit the run)."""
    import oracle
    t0 = time.time()
    o = oracle.Oracle(inp.to_c())
    setup = time.time() - t0
    log(f"oracle C3 setup {setup:.1f} s")
    x = np.random.default_rng(0).standard_normal(o.N) + 0j
    o.matvec(x)
    mv = []
    for _ in range(5):
        t1 = time.perf_counter()
        o.matvec(x)
        mv.append(time.perf_counter() - t1)
    res = [float(np.linalg.norm(o.matvec(X[q]) - B[q]) / np.linalg.norm(B[q])) for q in range(X.shape[0])]
    o.close()
    log(f"oracle C3: matvec {1e3 * np.median(mv):.1f} ms; results vs oracle max {max(res):.3e}")
    rec = cpu_timed_solve(c2_inputs(0), tol)
    log(f"oracle C2 timed solve: {rec}")
    rec.update({"sample": f"oracle GMRES(50) solve of C2 (N = 147456, kappa 80) from x0 = 0 to {tol:g}, timed as a "
                          f"whole ({rec['solve_s']:.2f} s, {rec['solve_cycles']} cycles); the oracle of C3 built "
                          f"({setup:.0f} s) and its matvec timed (median of 5) beside it",
                "c3_setup_s": setup, "c3_matvec_ms_median_of_5": 1e3 * float(np.median(mv)),
                "c2_solve_s": rec.pop("solve_s"), "results_max_rel_residual_vs_oracle": max(res)})
    rec["c2_setup_s"] = rec.pop("setup_s")
    rec["c2_matvec_ms_median_of_5"] = rec.pop("matvec_ms_median_of_5")
    return rec


def c3_inputs(seed=0):
    inp = lsg.InputsToDFMM(kappa=160.0, nLevelsUniform=7, L=0.5, nChebNodes=6, degreeOfBases=6, TOL_POW=10,
                           nCones_LFR=16)
    inp.contrast = lsg.bumps_from_seed(seed, 8)
    inp.inc_angle = seeded_angle(seed)
    return inp


def bench_setup(args):
    """Operator build (tree, tables, NCA skeletons, fused operators) of the C2
    and C3 geometries on the host (OpenMP) and with skeletons and entries on the
    GPU; both builds report their phases."""
    out = {"workload": "operator build: C2 (N = 147456, kappa 80) and C3 (N = 589824, kappa 160), host vs GPU",
           "threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))}
    for name, mk in (("C2", c2_inputs), ("C3", c3_inputs)):
        for dev in (1, 0):
            inp = mk(0)
            inp.setup_device = dev
            t0 = time.time()
            h = lsg.HostBuild(inp)
            wall = time.time() - t0
            i = h.info()
            ph = i["setup_phase_seconds"]
            rec = {"wall_s": wall, "far_tables_s": ph[0], "nca_skeletons_s": ph[1], "operator_fill_s": ph[2],
                   "near_tables_s": ph[3], "E_far": int(i["E_far"]), "max_rank": int(i["max_rank"])}
            log(f"{name} {'GPU' if dev else 'host'} build: {rec}")
            out[f"{name}_{'gpu' if dev else 'host'}"] = rec
            del h
        out[f"{name}_speedup"] = out[f"{name}_host"]["wall_s"] / out[f"{name}_gpu"]["wall_s"]
    print(json.dumps(out), flush=True)


# GMRES/Readme.md:43-115: the reference's sample run, the only timings it
# publishes for this path (hardware and thread count unstated)
README_ARGS = dict(nCones_LFR=16, nChebNodes=6, treeAdaptivity=5, L=0.5, kappa=40.0, yes2DFMM=1, degreeOfBases=6,
                   TOL_POW=8, m=10, restart=1, nLevelsUniform=5, Qchoice=0, tree_mode=1)
README_SWEEP = [((1, 10), 0.277782, 4.59412e-06), ((1, 50), 1.47135, 1.40294e-15), ((1, 100), 2.90013, 1.30732e-15),
                ((1, 200), 6.22192, 1.39836e-15), ((1, 500), 19.1974, 1.40164e-15), ((2, 10), 0.751761, 1.28443e-10),
                ((2, 50), 2.63477, 8.96364e-16), ((2, 100), 5.29393, 7.94185e-16), ((2, 200), 11.7579, 8.37373e-16),
                ((2, 500), 38.9819, 7.82761e-16)]
README_SETUP_S, README_MATVEC_V_S = 285.211, 0.011503


def bench_readme(args):
    """The reference's own sample run (adaptive tree, 5328 unknowns, kappa 40,
    TOL_POW 8; `testFMM2D_GMRES 16 6 5 0.5 40.0 1 6 8 10 1 5 0`) on this GPU:
    setup, every entry of its GMRES sweep (fixed cycles from x0 = 0, true FP64
    residual) and the volume-potential matvec, beside the README's times."""
    inp = lsg.InputsToDFMM(**README_ARGS)
    t0 = time.time()
    S = lsg.DFMM(inp)
    S.FMMTreeInitialisation()
    setup_s = time.time() - t0
    info = S.info()
    n_changed, report = S.autotune(reps=3)
    log(f"autotune ({n_changed} changed):\n{report}")
    b = S.rhs()
    bn = np.linalg.norm(b)
    sweep = []
    for (restart, m), t_ref, r_ref in README_SWEEP:
        x = lsg.RestartedGMRES(restart, S, m, inp, b)     # warm-up (workspace of this m)
        S._check(S._lib.lsg_sync(S._ctx))
        reps = 5
        t1 = time.perf_counter()
        for _ in range(reps):
            x = lsg.RestartedGMRES(restart, S, m, inp, b)
        S._check(S._lib.lsg_sync(S._ctx))
        t = (time.perf_counter() - t1) / reps
        r = float(np.linalg.norm(b - S.MatVecProduct(x)) / bn)
        rec = {"restart": restart, "m": m, "seconds": t, "readme_seconds": t_ref, "speedup": t_ref / t,
               "true_rel_residual": r, "readme_rel_residual": r_ref}
        log(f"GMRES({restart},{m}): {t * 1e3:.2f} ms (README {t_ref} s, {t_ref / t:.0f}x), residual {r:.5e} "
            f"(README {r_ref:.5e})")
        sweep.append(rec)
    V = lsg.DFMM(lsg.InputsToDFMM(**dict(README_ARGS, find_phi=0)))
    V.FMMTreeInitialisation()
    V.autotune(reps=3)
    mv_v = V.time_matvec(reps=50) * 1e-3
    # the same setup with the operator build and the table quadrature on the GPU
    t0 = time.time()
    Sd = lsg.DFMM(lsg.InputsToDFMM(**dict(README_ARGS, setup_device=1)))
    Sd.FMMTreeInitialisation()
    setup_dev_s = time.time() - t0
    Sd.close()
    out = {"workload": "README run: adaptive tree (nLevels 5, 148 leaves, 5328 unknowns), kappa 40, Gaussian "
                       "contrast, TOL_POW 8, GMRES sweep of GMRES.cpp main",
           "N": int(info["N"]), "nLevels": int(info["nLevels"]), "level_LFR": int(info["level_LFR"]),
           "setup_s": setup_s, "readme_setup_s": README_SETUP_S, "setup_speedup": README_SETUP_S / setup_s,
           "setup_gpu_build_s": setup_dev_s, "setup_gpu_build_speedup": README_SETUP_S / setup_dev_s,
           "matvec_ms": S.time_matvec(reps=50), "matvec_volume_potential_ms": mv_v * 1e3,
           "readme_matvec_volume_potential_ms": README_MATVEC_V_S * 1e3,
           "matvec_volume_potential_speedup": README_MATVEC_V_S / mv_v, "sweep": sweep,
           "readme_hardware": "unstated CPU (GMRES/Readme.md)"}
    print(json.dumps(out), flush=True)
    S.close()
    V.close()


def bench_learned(args):
    """The learned preconditioner at the configuration it is specified for
    (learned.c5_experiment): trained on 4096 solves the solver generates,
    evaluated on held-out systems (256 by default) against plain GMRES.
    --nets lists the dilation sets to train on the same data."""
    from lippmann_schwinger_solver_amd import learned as LN
    nets_spec = tuple(tuple(int(d) for d in spec.split(",")) for spec in args.nets.split(";"))
    for res in LN.c5_experiment(nets_spec, n_eval=args.n_eval, epochs=args.epochs, tol=args.tol, log=log):
        res.pop("per_system")
        log(f"learned {res['dilations']}: {res}")
        print(json.dumps({"workload": "learned preconditioner (C5)", "result": res}), flush=True)


def bench_ragged(args):
    """Ragged batch: 48 independent scatterers, tree depth uniform in {4, 5, 6},
    kappa uniform in [20, 80], 3-8 seeded bumps and a seeded angle each. The
    fused batch (one plan, one launch sequence per matvec, hipGraph-captured,
    autotuned, per-system Krylov bases) against the same systems solved one
    after the other and by concurrent contexts (one host thread each)."""
    rng = np.random.default_rng(2024)
    hosts, solvers, rhs, cases = [], [], [], []
    t0 = time.time()
    for i in range(48):
        depth, kappa, nb = int(rng.integers(4, 7)), float(rng.uniform(20, 80)), int(rng.integers(3, 9))
        inp = lsg.InputsToDFMM(kappa=kappa, nLevelsUniform=depth, L=0.5, nChebNodes=6, degreeOfBases=6, TOL_POW=10,
                               nCones_LFR=16)
        inp.contrast = lsg.bumps_from_seed(100 + i, nb)
        inp.inc_angle = seeded_angle(100 + i)
        h = lsg.HostBuild(inp)
        S = lsg.DFMM(inp)
        S.FMMTreeInitialisation(host=h)
        S.autotune(reps=2)
        S.use_graph(True)
        hosts.append(h)
        solvers.append(S)
        rhs.append(S.rhs())
        cases.append((depth, round(kappa, 2), nb))
    log(f"48 scatterers built in {time.time() - t0:.0f} s")
    R = lsg.RaggedBatch(hosts)
    R.autotune(reps=2)
    R.use_graph(True)
    R.gmres(rhs, tol=args.tol)                              # warm-up
    t0 = time.perf_counter()
    xs, stats, rc = R.gmres(rhs, tol=args.tol)
    t_fused = time.perf_counter() - t0
    if rc != 0:
        raise RuntimeError(f"fused ragged batch did not verify: {[s['rel_residual'] for s in stats]}")
    worst = max(float(np.linalg.norm(S.MatVecProduct(x) - b) / np.linalg.norm(b)) for S, b, x in zip(solvers, rhs, xs))
    lsg.gmres_ragged(solvers, rhs, tol=args.tol)            # warm-up
    t0 = time.perf_counter()
    _, st_con, rc_con = lsg.gmres_ragged(solvers, rhs, tol=args.tol)
    t_con = time.perf_counter() - t0
    t0 = time.perf_counter()
    for S, b in zip(solvers, rhs):
        S.gmres(b, restart_len=50, max_restarts=20, tol=args.tol)
    t_seq = time.perf_counter() - t0
    out = {"workload": "C4b ragged batch: 48 scatterers, depth {4,5,6}, kappa [20,80], 3-8 bumps, tol %g" % args.tol,
           "solves_per_s_fused": 48 / t_fused, "solves_per_s_concurrent_contexts": 48 / t_con,
           "solves_per_s_sequential": 48 / t_seq, "fused_speedup_vs_sequential": t_seq / t_fused,
           "concurrent_speedup_vs_sequential": t_seq / t_con, "fused_matvec_ms": R.time_matvec(reps=5),
           "max_rel_residual": max(s["rel_residual"] for s in stats), "recomputed_max_rel_residual": worst,
           "total_unknowns": int(sum(R.sizes)), "iterations": [s["iterations"] for s in stats], "cases": cases}
    print(json.dumps(out), flush=True)
    R.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 2 batches of 64 angles for c3, 10 solves for c2)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps before the timed ones (default 1; c2: 2)")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--restart", type=int, default=24, help="restart length of the batched solves (FP32 inner)")
    ap.add_argument("--outer", type=int, default=30, help="outer Krylov dimension of the FP64 flexible GMRES")
    ap.add_argument("--nets", default="1,2,4,8,4,1",
                    help="learned workload: dilation sets of the networks to train, ';'-separated")
    ap.add_argument("--n-eval", type=int, default=256, help="learned workload: held-out systems")
    ap.add_argument("--epochs", type=int, default=30, help="learned workload: training epochs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-reps", type=int, default=20)
    ap.add_argument("--pmc-probe-block", action="store_true",
                    help="only run block (64-column, MFMA) matvecs, for rocprofv3 --pmc")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="only run a 256 MiB device copy (counter calibration) and 3 matvecs, for rocprofv3 --pmc")
    ap.add_argument("--workload", choices=["c3", "c2", "learned", "ragged", "setup", "readme"], default="c3",
                    help="c3: the default line (64 angles of C3, FP32-inner batched against plain FP64); "
                         "c2: plain FP64 GMRES on C2 with its FP32-inner and batched legs; "
                         "learned: train the network on 4096 generated solves and evaluate on 256 held-out; "
                         "ragged: 48 independent scatterers; setup: operator build on the host vs the GPU; "
                         "readme: the reference's README run (adaptive tree) beside its published times")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the solver is single-device (sharding over several GPUs is "
              "designed but not implemented); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "setup":
        return bench_setup(args)
    if args.workload == "learned":
        return bench_learned(args)
    if args.workload == "ragged":
        return bench_ragged(args)
    if args.workload == "readme":
        return bench_readme(args)
    if args.workload == "c2":
        args.steps = 10 if args.steps is None else args.steps
        args.warmup = 2 if args.warmup is None else args.warmup
        return bench_c2(args)
    args.steps = 2 if args.steps is None else args.steps
    args.warmup = 1 if args.warmup is None else args.warmup
    return bench_c3(args)


def bench_c2(args):
    """Config C2 (depth 6, 147,456 unknowns, kappa 80, 5 seeded bumps): a step is
    one plain FP64 GMRES(50) solve from x0 = 0 to 1e-8 with b and x in HBM."""
    inp = c2_inputs(0)
    t0 = time.time()
    S = lsg.DFMM(inp)
    S.FMMTreeInitialisation()
    info = S.info()
    setup_s = time.time() - t0
    log(f"setup {setup_s:.1f} s: N={info['N']} E_far={info['E_far']} bytes/matvec={info['bytes_matvec']}")
    n_changed, report = S.autotune(reps=3)
    log(f"autotune ({n_changed} changed):\n{report}")
    b = S.rhs()

    # b and x resident in HBM before timing (plain hipMalloc); x is zeroed per solve
    bd, xd = DeviceBlock(b.reshape(1, -1)), DeviceBlock(np.zeros((1, S.N), dtype=np.complex128))

    def solve():
        xd.zero()
        _, st = S.gmres_ex(None, mode="fp64", restart_len=50, max_restarts=20, tol=args.tol,
                           device_ptrs=(bd.ptr.value, xd.ptr.value))
        if st["status"] != 0 or st["converged"] != 1 or not st["rel_residual"] <= args.tol:
            raise RuntimeError(f"solve did not verify: {st}")
        return st

    for _ in range(args.warmup):
        st = solve()
        log(f"warmup: {st}")
    S._check(S._lib.lsg_sync(S._ctx))
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(solve())
    S._check(S._lib.lsg_sync(S._ctx))
    elapsed = time.perf_counter() - t0
    value = args.steps / elapsed
    log(f"timed: {args.steps} solves in {elapsed:.3f} s; last {stats[-1]}")
    xres = xd.download().reshape(-1)
    verify = float(np.linalg.norm(S.MatVecProduct(xres) - b) / np.linalg.norm(b))
    # the same solve through host buffers (PCIe copies through the context's
    # pinned staging), median of three
    host_runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        S.gmres(b, restart_len=50, max_restarts=20, tol=args.tol)
        host_runs.append(time.perf_counter() - t0)
    host_ms = 1e3 * float(np.median(host_runs))
    log(f"recomputed residual {verify:.3e}; host-buffer solve {host_ms:.2f} ms")

    # the mixed-precision solver on the same system: FP32-operator GMRES inside
    # FP64 iterative refinement, verified by the FP64 residual
    x64 = S.gmres_ex(b, mode="fp64", restart_len=50, max_restarts=20, tol=args.tol)[0]
    S.gmres_ex(b, mode="fp32ir", restart_len=50, max_restarts=1, tol=args.tol, inner_tol=1e-5)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        x32, st32 = S.gmres_ex(b, mode="fp32ir", restart_len=50, max_restarts=1, tol=args.tol, inner_tol=1e-5)
        if st32["status"] != 0 or not st32["rel_residual"] <= args.tol:
            raise RuntimeError(f"FP32-IR solve did not verify: {st32}")
    el32 = time.perf_counter() - t1
    fp32_ir = {"solves_per_s": args.steps / el32, "ms_per_solve": 1e3 * el32 / args.steps,
               "inner_iterations": st32["iterations"], "refinements": st32["refinements"],
               "final_rel_residual": st32["rel_residual"],
               "rel_diff_vs_fp64_solution": float(np.linalg.norm(x32 - x64) / np.linalg.norm(x64)),
               "matvec_fp32_ms": S.time_matvec_fp32(reps=args.profile_reps)}
    log(f"fp32-IR: {fp32_ir}")

    batched = batched_leg(S, info, args, st_plain=stats[-1], plain_sps=value)
    log(f"batched: {batched}")

    # full matvec (both streams) and per-pass device time (HIP events on the
    # operator's stream, passes serialised)
    total_ms = S.time_matvec(reps=args.profile_reps)
    passes = S.profile_passes(reps=args.profile_reps)
    dom = max(passes, key=lambda p: p[1])
    achieved = dom[2] / (dom[1] * 1e-3) / 1e9
    mv_gbs = info["bytes_matvec"] / (total_ms * 1e-3) / 1e9
    log(f"matvec {total_ms * 1e3:.1f} us ({mv_gbs:.0f} GB/s algorithmic); serialised passes sum "
        f"{sum(p[1] for p in passes) * 1e3:.1f} us")
    for p in passes:
        log(f"  pass {p[0]:10s} {p[1] * 1e3:9.1f} us  {p[2] / 1e6:9.2f} MB  {p[2] / (p[1] * 1e-3) / 1e9:8.1f} GB/s")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {"workload": "C2: plain FP64 DAFMM-GMRES(50), uniform depth 6 (4096 leaves x 36 Chebyshev "
                               "nodes = 147456 unknowns), kappa 80, 5 seeded Gaussian bumps, NCA tol 1e-10, "
                               f"GMRES tol {args.tol:g}",
                   "N": int(info["N"]), "iterations_per_solve": stats[-1]["iterations"],
                   "restarts_per_solve": stats[-1]["restarts"], "final_rel_residual": stats[-1]["rel_residual"],
                   "recomputed_rel_residual": verify, "inputs": "b, x resident in HBM (device pointers)",
                   "host_buffer_solve_ms": host_ms},
        "setup": setup_record(setup_s, info, inp),
        "matvec": {"ms": total_ms, "algorithmic_bytes": int(info["bytes_matvec"]), "GBps": mv_gbs,
                   "frac_hbm": mv_gbs / HBM_PEAK_GBS, "E_far": int(info["E_far"])},
        "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("c2", dom[0])},
        "cpu_baseline": None,
        "fp32_ir": fp32_ir,
        "batched_64_angles": batched,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_timed_solve(inp, args.tol)
    print(json.dumps(out), flush=True)
    S.close()


if __name__ == "__main__":
    main()
