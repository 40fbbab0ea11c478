"""Benchmark: FP64-verified GMRES solves/s of the DAFMM Lippmann-Schwinger
operator on one GPU (config C2: uniform depth 6, 147,456 unknowns, kappa 80,
5 seeded Gaussian bumps, plane wave at a seeded angle, NCA tol 1e-10).

A step is one GMRES(50) solve from x0 = 0 to relative residual 1e-8, whose
FP64 true residual is checked on the device; the operator and b are resident
in HBM before timing starts. Prints one JSON line.
"""
import argparse
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


def cpu_baseline(inp, iters_per_solve, restarts_per_solve):
    """Oracle (CPU restatement of the reference) on the host cores: build the
    same operator, time one GMRES(10) cycle (11 matvecs + MGS), scale to a solve
    with the GPU's iteration count."""
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.time()
    o = oracle.Oracle(inp.to_c())
    setup = time.time() - t0
    b = lsg.DFMM  # noqa: F841  (keep the product module imported)
    x = np.random.default_rng(0).standard_normal(o.N) + 0j
    m = 10
    t0 = time.time()
    o.gmres(x, 1, m)
    t_cycle = time.time() - t0
    t_iter = t_cycle / (m + 1)
    est = t_iter * (iters_per_solve + restarts_per_solve + 1)
    return {"value": 1.0 / est, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"oracle GMRES(10) cycle on config C2 ({t_cycle:.2f} s, {m + 1} matvecs + MGS), "
                      f"scaled to {iters_per_solve} Arnoldi steps + {restarts_per_solve + 1} residual matvecs; "
                      f"oracle setup {setup:.1f} s not timed"}


def measured_traffic(pass_name):
    """HBM bytes per launch of the dominant pass from the committed FETCH_SIZE /
    WRITE_SIZE counter runs (scaled by a plain copy's counters); None if absent."""
    f = os.path.join(ROOT, "profiles", "c2_leaf_m2l_pmc_traffic.json")
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

    def download(self):
        import ctypes as C
        out = np.empty(self.shape, dtype=np.complex128)
        if self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.nbytes), 2) != 0:
            raise RuntimeError("hipMemcpy D2H failed")
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


MFMA_F64_PEAK_TFLOPS = 78.6


def batched_leg(S, info, args, st_plain, plain_sps):
    """64 plane waves at 2 pi j / 64 on the same geometry, solved together by
    batched GMRES (one FP64 MFMA block matvec per Arnoldi step); the right-hand
    sides are resident in HBM before timing; every system's FP64 true residual
    must reach tol."""
    angles = 2 * np.pi * np.arange(64) / 64
    Bh = np.ascontiguousarray(S.rhs_angles(angles).T)      # 64 x N: column q at q*N
    Bd, Xd = DeviceBlock(Bh), DeviceBlock(np.zeros_like(Bh))
    ptrs = (Bd.ptr.value, Xd.ptr.value)
    Xd.zero()
    S.gmres_batch(None, device_ptrs=ptrs, nrhs=64, ld=S.N, restart_len=50, max_restarts=20, tol=args.tol)  # warm-up
    t = 0.0
    for _ in range(args.steps):
        Xd.zero()
        t0 = time.perf_counter()
        _, st, rel = S.gmres_batch(None, device_ptrs=ptrs, nrhs=64, ld=S.N, restart_len=50, max_restarts=20,
                                   tol=args.tol)
        t += time.perf_counter() - t0
        if st["status"] != 0 or st["converged"] != 64 or not rel.max() <= args.tol:
            raise RuntimeError(f"batched solve did not verify: {st}")
    X = Xd.download().T
    Bd.free()
    Xd.free()
    # independent FP64 residuals of a few systems through the single-vector matvec
    spot = max(float(np.linalg.norm(S.MatVecProduct(X[:, j]) - Bh[j]) / np.linalg.norm(Bh[j])) for j in (0, 21, 42, 63))
    mv_block = S.time_matvec_block(reps=5)
    flops = 8.0 * 64 * (info["E_far"] + info["E_near"])
    sps = 64 * args.steps / t
    return {"solves_per_s": sps, "speedup_vs_plain_fp64": sps / plain_sps, "ms_per_batch": 1e3 * t / args.steps,
            "arnoldi_steps": st["iterations"], "cycles": st["restarts"], "max_rel_residual": float(rel.max()),
            "spot_check_rel_residual": spot, "block_matvec_ms": mv_block,
            "block_matvec_us_per_rhs": 1e3 * mv_block / 64,
            "block_matvec_mfma": {"achieved_tflops": flops / (mv_block * 1e-3) / 1e12, "peak": MFMA_F64_PEAK_TFLOPS,
                                  "frac": flops / (mv_block * 1e-3) / 1e12 / MFMA_F64_PEAK_TFLOPS,
                                  "flops_per_matvec": flops}}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-reps", type=int, default=20)
    ap.add_argument("--pmc-probe", action="store_true",
                    help="only run a 256 MiB device copy (counter calibration) and 3 matvecs, for rocprofv3 --pmc")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the solver is single-device (multi-GPU sharding is a design "
              "note, not implemented); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    inp = c2_inputs(0)
    t0 = time.time()
    S = lsg.DFMM(inp)
    S.FMMTreeInitialisation()
    info = S.info()
    log(f"setup {time.time() - t0:.1f} s: N={info['N']} E_far={info['E_far']} bytes/matvec={info['bytes_matvec']}")
    b = S.rhs()
    if args.pmc_probe:
        pmc_probe(S)
        S.close()
        return

    def solve():
        x, st = S.gmres(b, restart_len=50, max_restarts=20, tol=args.tol)
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
                   "restarts_per_solve": stats[-1]["restarts"], "final_rel_residual": stats[-1]["rel_residual"]},
        "matvec": {"ms": total_ms, "algorithmic_bytes": int(info["bytes_matvec"]), "GBps": mv_gbs,
                   "frac_hbm": mv_gbs / HBM_PEAK_GBS, "E_far": int(info["E_far"])},
        "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom[0])},
        "cpu_baseline": None,
        "fp32_ir": fp32_ir,
        "batched_64_angles": batched,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(inp, stats[-1]["iterations"], stats[-1]["restarts"])
    print(json.dumps(out), flush=True)
    S.close()


if __name__ == "__main__":
    main()
