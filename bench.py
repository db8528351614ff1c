#!/usr/bin/env python3
"""Headline benchmark: n=16384 FP64 solve time, mixed-precision LU + FP64
iterative refinement against the FP64 LU, at equal backward error.

Workload (one "step" = one full solve A X = B from device-resident FP64 data):
  A = U diag(s) V^T, U and V Haar orthogonal (QR of a Gaussian matrix with the
  sign fix), s log-spaced from 1 to 1e6 (cond_2(A) = 1e6), n = 16384; B n x nrhs
  i.i.d. N(0,1).  Inputs are generated on the GPU and are not timed.

Printed JSON line: value = milliseconds per mixed-precision solve (lower is
better) with the FP64 solve time, the speedup and both backward errors in
`config` (1 right-hand side; the 64-column case is reported beside it);
`roofline` for the kernel that dominates the critical path (the resident panel
leaf) and `roofline_gemm` for the trailing-matrix update beside it, both from
HIP events recorded on the library's streams during the timed steps;
`cpu_baseline` from the CPU oracle on a bounded sample.  `config.c1` and
`config.c2` time the plain FP64 getrf + getrs at n=512 (one right-hand side)
and n=8192 (16 right-hand sides, autotuned panel width).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 16384] [--nrhs 1]
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "n=16384 solve time: BF16/FP32-MFMA LU+IR vs FP64 LU at equal backward error"
PEAK = {"hbm_gbs": 8000.0, "f64": 78.6, "f32": 157.3, "bf16": 2500.0}
# Mixed-precision variants: factorization arithmetic, and "+tp" for tournament
# pivoting of the FP32 panel (LAIR_PIVOT_TOURNAMENT) instead of partial pivoting.
PREC_NAMES = {"fp32": 1, "bf16x6": 2, "fp32+tp": 1 | 0x100, "bf16x6+tp": 2 | 0x100}
EPS = float(np.finfo(np.float64).eps)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def gen_lib():
    path = os.path.join(ROOT, "lair_amd", "lib", "liblair_gen.so")
    lib = ctypes.CDLL(path)
    lib.lair_gen_gaussian.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                      ctypes.c_uint64]
    lib.lair_gen_gaussian.restype = ctypes.c_int
    lib.lair_gen_kappa.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_double,
                                   ctypes.c_uint64]
    lib.lair_gen_kappa.restype = ctypes.c_int
    lib.lair_gen_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
    lib.lair_gen_copy.restype = ctypes.c_int
    lib.lair_gen_batch_shift.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    lib.lair_gen_batch_shift.restype = ctypes.c_int
    lib.lair_gen_to_f32.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
    lib.lair_gen_to_f32.restype = ctypes.c_int
    lib.lair_gen_shift_diag.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_double]
    lib.lair_gen_shift_diag.restype = ctypes.c_int
    return lib


def bench_batched(ctx, gl, dtype, steps, count=20000, seed=4):
    """C4: `count` systems, n ~ U{8..256}, A ~ N(0,1) + n I, nrhs ~ U{1,2,4},
    solved in one graph-captured persistent launch (plan autotuned once).
    Inputs are regenerated on the device before every timed launch (untimed)."""
    import lair_amd as L
    rng = np.random.default_rng(seed)
    ns = rng.integers(8, 257, count).astype(np.int32)
    rs = rng.choice(np.array([1, 2, 4], np.int32), count)
    offs = np.concatenate([[0], np.cumsum(ns.astype(np.int64) ** 2)[:-1]]).astype(np.int64)
    na, nb_ = int((ns.astype(np.int64) ** 2).sum()), int((ns.astype(np.int64) * rs).sum())
    w = np.dtype(dtype).itemsize
    d_ns, d_offs = L.DeviceArray.from_host(ns.reshape(-1, 1)), L.DeviceArray.from_host(offs.reshape(-1, 1))
    a64, b64 = L.DeviceArray(na, 1), L.DeviceArray(nb_, 1)
    da, db = L.DeviceArray(na, 1, dtype), L.DeviceArray(nb_, 1, dtype)
    dp, di = L.DeviceArray(int(ns.sum()), 1, np.int64), L.DeviceArray(count, 1, np.int32)

    def regen():
        rc = gl.lair_gen_gaussian(a64.ptr, na, 1, 1, seed)
        rc |= gl.lair_gen_batch_shift(a64.ptr, d_ns.ptr, d_offs.ptr, count)
        rc |= gl.lair_gen_gaussian(b64.ptr, nb_, 1, 1, seed + 1)
        if dtype == np.float64:
            ctx.lib.lair_memcpy_d2d(da.ptr, a64.ptr, na * 8)
            ctx.lib.lair_memcpy_d2d(db.ptr, b64.ptr, nb_ * 8)
        else:
            rc |= gl.lair_gen_to_f32(a64.ptr, da.ptr, na)
            rc |= gl.lair_gen_to_f32(b64.ptr, db.ptr, nb_)
        if rc:
            raise RuntimeError("batched input generation failed")

    regen()
    plan = L.BatchedPlan(ctx, ns, rs, da, db, dp, di, precision=0 if dtype == np.float64 else 1)
    times = []
    for _ in range(steps + 1):
        regen()
        ctx.synchronize()
        t = time.perf_counter()
        plan.run()
        ctx.synchronize()
        times.append(time.perf_counter() - t)
    ms = float(np.mean(times[1:])) * 1e3
    info = di.download().ravel()
    # Spot check: backward error of three systems against their inputs.
    def fetch(arr, off, cnt, dt):
        h = np.empty(cnt, dt)
        ctx.lib.lair_memcpy_d2h(ctypes.c_void_p(h.ctypes.data), ctypes.c_void_p(arr.ptr.value + off * h.itemsize),
                                h.nbytes)
        return h
    boffs = np.concatenate([[0], np.cumsum(ns.astype(np.int64) * rs)[:-1]])
    bes = []
    for i in (0, count // 2, count - 1):
        n, r = int(ns[i]), int(rs[i])
        A = fetch(a64, int(offs[i]), n * n, np.float64).reshape(n, n)
        B = fetch(b64, int(boffs[i]), n * r, np.float64).reshape(n, r)
        if dtype == np.float32:
            A, B = A.astype(np.float32).astype(np.float64), B.astype(np.float32).astype(np.float64)
        X = fetch(db, int(boffs[i]), n * r, dtype).reshape(n, r).astype(np.float64)
        bes.append(float(np.abs(B - A @ X).max() / (np.abs(A).sum(1).max() * np.abs(X).max())))
    cfg = plan.config()
    plan.close()
    nf = ns.astype(np.float64)
    flops = float((2 * nf ** 3 / 3 + 2 * nf ** 2 * rs).sum())
    byts = float((w * (2 * nf ** 2 + 2 * nf * rs)).sum())
    for d in (d_ns, d_offs, a64, b64, da, db, dp, di):
        d.free()
    return {"ms": round(ms, 3), "systems": count, "systems_per_s": round(count / (ms * 1e-3), 1),
            "singular": int((info >= 0).sum()), "sample_backward_error": max(bes),
            "algorithmic_bytes": byts, "algorithmic_flops": flops,
            "hbm_gbs": round(byts / (ms * 1e-3) / 1e9, 1),
            "hbm_frac": round(byts / (ms * 1e-3) / 1e9 / PEAK["hbm_gbs"], 4),
            "plan": cfg}


def bench_adjoint(ctx, gl, steps, n=4096, nrhs=64):
    """C5: A ~ N(0,1) + 4 sqrt(n) I, B and XBAR ~ N(0,1); times
    lair_dsolve_backward_device (BBAR = A^-T XBAR by a transposed getrs on the
    stored factors, ABAR = -BBAR X^T by the FP64 MFMA GEMM)."""
    import lair_amd as L
    ld = n + 64
    da, db, dxb = L.DeviceArray(n, n, ld=ld), L.DeviceArray(n, nrhs), L.DeviceArray(n, nrhs)
    rc = gl.lair_gen_gaussian(da.ptr, n, n, ld, 4096)
    rc |= gl.lair_gen_shift_diag(da.ptr, n, ld, 4.0 * np.sqrt(n))
    rc |= gl.lair_gen_gaussian(db.ptr, n, nrhs, nrhs, 4097)
    rc |= gl.lair_gen_gaussian(dxb.ptr, n, nrhs, nrhs, 4098)
    if rc:
        raise RuntimeError("adjoint input generation failed")
    lib = ctx.lib
    dlu = L.DeviceArray(n, n, ld=ld)
    lib.lair_memcpy_d2d(dlu.ptr, da.ptr, n * ld * 8)
    dp = L.DeviceArray(n, 1, np.int64)
    sing = ctypes.c_int64(-1)
    L.check(lib.lair_dgetrf_device(ctx.h, n, n, dlu.ptr, ld, dp.ptr, 0, ctypes.byref(sing)), "getrf")
    dx = L.DeviceArray(n, nrhs)
    lib.lair_memcpy_d2d(dx.ptr, db.ptr, n * nrhs * 8)
    L.check(lib.lair_dgetrs_device(ctx.h, n, nrhs, dlu.ptr, ld, dp.ptr, dx.ptr, nrhs, 0), "getrs")
    dbb, dab = L.DeviceArray(n, nrhs), L.DeviceArray(n, n)
    times = []
    for _ in range(steps + 1):
        ctx.synchronize()
        t = time.perf_counter()
        L.check(lib.lair_dsolve_backward_device(ctx.h, n, nrhs, dlu.ptr, ld, dp.ptr, dx.ptr, nrhs, dxb.ptr,
                                                nrhs, dbb.ptr, nrhs, dab.ptr, n), "solve_backward")
        ctx.synchronize()
        times.append(time.perf_counter() - t)
    ms = float(np.mean(times[1:])) * 1e3
    # Closed-form check on the host: A^T BBAR = XBAR.
    a = da.download()
    bbar, xbar = dbb.download(), dxb.download()
    rel = float(np.abs(a.T @ bbar - xbar).max() / (np.abs(a).sum(0).max() * np.abs(bbar).max()))
    byts = 2.0 * n * n * 8
    for d in (da, db, dxb, dlu, dp, dx, dbb, dab):
        d.free()
    return {"ms": round(ms, 3), "n": n, "nrhs": nrhs, "bbar_backward_error": rel,
            "algorithmic_flops": 2.0 * n * n * nrhs * 2, "algorithmic_bytes": byts,
            "hbm_gbs": round(byts / (ms * 1e-3) / 1e9, 1),
            "hbm_frac": round(byts / (ms * 1e-3) / 1e9 / PEAK["hbm_gbs"], 4)}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "kernel_traffic.json")


def measured_traffic(key):
    """HBM bytes per launch of a kernel class ("panel", or the trailing update
    of "fp32" / "bf16x6"), from the committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the same counters' reading
    for a copy kernel of known size); None when no summary for it exists."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if key != "panel":
        key = "bf16x6" if key.startswith("bf16x6") else "fp32"
    entry = t.get(key)
    return entry.get("hbm_bytes_per_launch") if entry else None


def prof_read(ctx, kname):
    from lair_amd import _capi
    ms, cnt, fl, by = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
    ctx.lib.lair_profile_read(ctx.h, _capi.KCLASS[kname], ctypes.byref(ms), ctypes.byref(cnt),
                              ctypes.byref(fl), ctypes.byref(by))
    return ms.value, cnt.value, fl.value, by.value


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_solve_ms(n, seed=99):
    """The oracle's restatement of getrf_row_major + getrs on one core for an
    n x n N(0,1) system (milliseconds)."""
    from lair_amd import inputs
    from oracle import oracle as O
    a = inputs.gaussian_matrix(seed, n, n)
    b = inputs.gaussian_matrix(seed + 1, n, 1)
    t = time.perf_counter()
    piv, _ = O.getrf(a, nthreads=1)
    O.getrs(a, piv, b)
    return (time.perf_counter() - t) * 1e3


def cpu_baseline(n_target, n_sample):
    """The oracle's restatement of getrf_row_major + getrs (single thread) on an
    n_sample Gaussian system, scaled to n_target by the n^3 flop ratio."""
    ms = cpu_solve_ms(n_sample)
    scale = (n_target / n_sample) ** 3
    return {"value": round(ms * scale, 1), "unit": "ms",
            "cores": 1, "kind": "port",
            "sample": f"oracle getrf_row_major+getrs on one core, n={n_sample} N(0,1), "
                      f"{ms / 1e3:.2f} s measured, scaled by (n/{n_sample})^3 to n={n_target}",
            "cpu": cpu_model(), "nproc": os.cpu_count()}


def bench_plain(ctx, gl, n, nrhs, steps, nb, seed):
    """Plain FP64 getrf + getrs from device-resident N(0,1) data (the C1 and
    C2 shapes): ms per factor-and-solve, with its backward error."""
    import lair_amd as L
    lib = ctx.lib
    ld = n + 64 if n % 4096 == 0 else ((n + 63) // 64) * 64
    da, dlu = L.DeviceArray(n, n, ld=ld), L.DeviceArray(n, n, ld=ld)
    db, dx = L.DeviceArray(n, nrhs), L.DeviceArray(n, nrhs)
    dp = L.DeviceArray(n, 1, np.int64)
    rc = gl.lair_gen_gaussian(da.ptr, n, n, ld, seed)
    rc |= gl.lair_gen_gaussian(db.ptr, n, nrhs, nrhs, seed + 1)
    if rc:
        raise RuntimeError("input generation failed")
    sing = ctypes.c_int64(-1)

    def step():
        lib.lair_memcpy_d2d(dlu.ptr, da.ptr, n * ld * 8)
        lib.lair_memcpy_d2d(dx.ptr, db.ptr, n * nrhs * 8)
        L.check(lib.lair_dgetrf_device(ctx.h, n, n, dlu.ptr, ld, dp.ptr, nb, ctypes.byref(sing)), "getrf")
        L.check(lib.lair_dgetrs_device(ctx.h, n, nrhs, dlu.ptr, ld, dp.ptr, dx.ptr, nrhs, 0), "getrs")

    step()  # warm-up (and the autotuner's one-time search when nb < 0)
    times = []
    for _ in range(steps):
        lib.lair_memcpy_d2d(dlu.ptr, da.ptr, n * ld * 8)
        lib.lair_memcpy_d2d(dx.ptr, db.ptr, n * nrhs * 8)
        ctx.synchronize()
        t = time.perf_counter()
        L.check(lib.lair_dgetrf_device(ctx.h, n, n, dlu.ptr, ld, dp.ptr, nb, ctypes.byref(sing)), "getrf")
        L.check(lib.lair_dgetrs_device(ctx.h, n, nrhs, dlu.ptr, ld, dp.ptr, dx.ptr, nrhs, 0), "getrs")
        ctx.synchronize()
        times.append(time.perf_counter() - t)
    v = ctypes.c_double()
    L.check(lib.lair_dbackward_error(ctx.h, n, nrhs, da.ptr, ld, db.ptr, nrhs, dx.ptr, nrhs, ctypes.byref(v)),
            "berr")
    for d in (da, dlu, db, dx, dp):
        d.free()
    ms = float(np.median(times)) * 1e3
    flops = 2.0 * n ** 3 / 3 + 2.0 * n * n * nrhs
    return {"ms": round(ms, 3), "n": n, "nrhs": nrhs, "panel_width": "autotuned" if nb < 0 else nb,
            "backward_error": v.value, "tflops": round(flops / (ms * 1e-3) / 1e12, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--nrhs", type=int, default=1)
    ap.add_argument("--kappa", type=float, default=1e6)
    ap.add_argument("--precisions", default="bf16x6+tp,fp32+tp,bf16x6,fp32")
    ap.add_argument("--cpu-sample-n", type=int, default=5120)
    ap.add_argument("--only-mixed", action="store_true",
                    help="skip the FP64 solve (counter-collection runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a 256 MiB copy kernel (known bytes for counter calibration)")
    ap.add_argument("--no-side", action="store_true",
                    help="skip the batched (C4) and adjoint (C5) measurements")
    ap.add_argument("--gaussian", action="store_true",
                    help="A ~ N(0,1) from the library's own generator kernel instead of the kappa=1e6 "
                         "construction (same factorization shapes; for counter-collection runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain FP64 getrf + getrs timings at n=512 and n=8192")
    ap.add_argument("--no-nrhs64", action="store_true",
                    help="skip the 64-right-hand-side run of the headline workload")
    ap.add_argument("--only-batched", action="store_true",
                    help="only the batched (C4) measurements, FP64 then FP32, one JSON line each "
                         "(profiling runs of the batched kernel)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the n=16384 LU "
              "fits one MI355X); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import lair_amd as L
    from lair_amd import _capi

    if args.only_batched:
        ctx = L.Context()
        gl = gen_lib()
        for dt in (np.float64, np.float32):
            print(json.dumps({"batched_c4": np.dtype(dt).name, **bench_batched(ctx, gl, dt, args.steps)}),
                  flush=True)
        return

    n, nrhs, K, W = args.n, args.nrhs, args.steps, args.warmup
    ld = n + 64 if (n % 4096 == 0) else ((n + 63) // 64) * 64
    ctx = L.Context()
    lib = ctx.lib
    gl = gen_lib()
    da = L.DeviceArray(n, n, np.float64, ld=ld)
    db = L.DeviceArray(n, nrhs)
    dx = L.DeviceArray(n, nrhs)
    t = time.perf_counter()
    if args.gaussian:
        rc = gl.lair_gen_gaussian(da.ptr, n, n, ld, 16384)
    else:
        rc = gl.lair_gen_kappa(da.ptr, n, ld, args.kappa, 16384)
    rc |= gl.lair_gen_gaussian(db.ptr, n, nrhs, nrhs, 16385)
    if rc:
        raise RuntimeError(f"input generation failed ({rc})")
    log(f"generated A (n={n}, kappa={args.kappa:g}) and B ({nrhs} rhs) in {time.perf_counter() - t:.1f} s")

    # Host-to-device transfer of A, reported separately (never the headline).
    host = np.empty((n, ld), np.float64)
    lib.lair_memcpy_d2h(ctypes.c_void_p(host.ctypes.data), da.ptr, host.nbytes)
    t = time.perf_counter()
    lib.lair_memcpy_h2d(da.ptr, ctypes.c_void_p(host.ctypes.data), host.nbytes)
    h2d_ms = (time.perf_counter() - t) * 1e3
    del host

    if args.copy_calibration:
        cb = 256 << 20
        src, dst = L.DeviceArray(cb // 8, 1), L.DeviceArray(cb // 8, 1)
        if gl.lair_gen_copy(dst.ptr, src.ptr, cb):
            raise RuntimeError("calibration copy failed")
        log(f"calibration copy of {cb} bytes done")
        src.free()
        dst.free()

    def berr():
        v = ctypes.c_double()
        L.check(lib.lair_dbackward_error(ctx.h, n, nrhs, da.ptr, ld, db.ptr, nrhs, dx.ptr, nrhs,
                                         ctypes.byref(v)), "berr")
        return v.value

    # FP64 LU + getrs.
    def fp64_step():
        L.check(lib.lair_dgesv_device(ctx.h, n, nrhs, da.ptr, ld, db.ptr, nrhs, dx.ptr, nrhs), "dgesv")

    fp64_ms, fp64_berr = None, None
    if not args.only_mixed:
        for _ in range(W):
            fp64_step()
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(K):
            fp64_step()
        ctx.synchronize()
        fp64_ms = (time.perf_counter() - t) * 1e3 / K
        fp64_berr = berr()
        log(f"fp64 LU: {fp64_ms:.2f} ms/solve, backward error {fp64_berr:.2e}")

    results = {}
    for name in [p for p in args.precisions.split(",") if p]:
        prec = PREC_NAMES[name]
        info = _capi.IrInfo()

        def step():
            st = lib.lair_dgesv_mixed_device(ctx.h, n, nrhs, da.ptr, ld, db.ptr, nrhs, dx.ptr, nrhs, prec,
                                             20, 0.0, ctypes.byref(info))
            if st not in (0, _capi.LAIR_ERR_NOCONV):
                L.check(st, "dgesv_mixed")
            return st

        for _ in range(W):
            step()
        lib.lair_profile_reset(ctx.h)
        # Events around the panel leaves (the critical path's dominant kernel)
        # and the trailing updates, on the streams they are launched on.
        lib.lair_profile_enable(ctx.h, (1 << _capi.KCLASS["gemm"]) | (1 << _capi.KCLASS["panel"]))
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(K):
            st = step()
        ctx.synchronize()
        ms = (time.perf_counter() - t) * 1e3 / K
        lib.lair_profile_enable(ctx.h, 0)
        g_ms, g_cnt, g_fl, g_by = prof_read(ctx, "gemm")
        p_ms, p_cnt, p_fl, p_by = prof_read(ctx, "panel")
        be = berr()
        results[name] = {"ms": ms, "status": st, "iterations": info.iterations,
                         "converged": info.converged, "fell_back": info.fell_back,
                         "backward_error": be, "gemm": (g_ms / K, g_cnt // K, g_fl / K, g_by / K),
                         "panel": (p_ms / K, p_cnt // K, p_fl / K, p_by / K)}
        log(f"{name}: {ms:.2f} ms/solve, {info.iterations} IR steps, backward error {be:.2e}, "
            f"fell_back={info.fell_back}, trailing gemm {g_ms / K:.2f} ms over {g_cnt // K} launches")

    # Per-class time breakdown of one extra profiled solve (headline precision).
    ok = {k: v for k, v in results.items() if v["converged"] and not v["fell_back"]}
    head = min(ok, key=lambda k: ok[k]["ms"]) if ok else min(results, key=lambda k: results[k]["ms"])
    hr = results[head]
    lib.lair_profile_reset(ctx.h)
    lib.lair_profile_enable(ctx.h, 0xFF)
    info = _capi.IrInfo()
    lib.lair_dgesv_mixed_device(ctx.h, n, nrhs, da.ptr, ld, db.ptr, nrhs, dx.ptr, nrhs, PREC_NAMES[head], 20,
                                0.0, ctypes.byref(info))
    lib.lair_profile_enable(ctx.h, 0)
    breakdown = {k: round(prof_read(ctx, k)[0], 3) for k in _capi.KCLASS}

    # Roofline of the panel leaves (the kernel class that dominates the
    # critical path): algorithmic bytes = the leaf read once and written once,
    # 2 * rows * width * 4 bytes per launch, over the launches' event time.
    p_ms, p_cnt, p_fl, p_by = hr["panel"]
    p_ach = p_by / (p_ms * 1e-3) / 1e9 if p_ms > 0 else 0.0
    roof = {"bound": "hbm", "achieved": round(p_ach, 2), "peak": PEAK["hbm_gbs"], "unit": "GB/s",
            "frac": round(p_ach / PEAK["hbm_gbs"], 5), "traffic": measured_traffic("panel"),
            "kernel": f"resident panel leaf rpanel_kernel ({head}); latency-bound by the per-column "
                      f"exchange", "launches_per_step": p_cnt,
            "algorithmic_bytes_per_launch": round(p_by / max(p_cnt, 1)),
            "us_per_launch": round(p_ms * 1e3 / max(p_cnt, 1), 1),
            "us_per_column": round(p_ms * 1e3 / n, 3), "ms_per_step": round(p_ms, 3)}
    # Roofline of the trailing update: the lower of the MFMA and HBM ceilings.
    g_ms, g_cnt, g_fl, g_by = hr["gemm"]
    mfma_peak = PEAK["bf16"] if head.startswith("bf16") else PEAK["f32"]
    products = 6 if head.startswith("bf16x6") else 1
    t_mfma = g_fl * products / (mfma_peak * 1e12)
    t_hbm = g_by / (PEAK["hbm_gbs"] * 1e9)
    if t_hbm >= t_mfma:
        achieved = g_by / (g_ms * 1e-3) / 1e9
        roof_g = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK["hbm_gbs"], "unit": "GB/s",
                  "frac": round(achieved / PEAK["hbm_gbs"], 4)}
    else:
        achieved = g_fl / (g_ms * 1e-3) / 1e12
        roof_g = {"bound": "mfma", "achieved": round(achieved, 2), "peak": mfma_peak, "unit": "TFLOP/s",
                  "frac": round(achieved / mfma_peak, 4)}
    roof_g.update({"traffic": measured_traffic(head), "kernel": f"trailing update ({head})",
                   "launches_per_step": g_cnt, "algorithmic_flops_per_step": g_fl,
                   "algorithmic_bytes_per_step": g_by, "ms_per_step": round(g_ms, 3)})

    # The same workload with 64 right-hand sides: FP64 and the fastest variant.
    c3_64 = None
    if not args.no_nrhs64 and not args.only_mixed and nrhs == 1:
        r64 = 64
        db64, dx64 = L.DeviceArray(n, r64), L.DeviceArray(n, r64)
        if gl.lair_gen_gaussian(db64.ptr, n, r64, r64, 16386):
            raise RuntimeError("input generation failed")

        def berr64():
            v = ctypes.c_double()
            L.check(lib.lair_dbackward_error(ctx.h, n, r64, da.ptr, ld, db64.ptr, r64, dx64.ptr, r64,
                                             ctypes.byref(v)), "berr")
            return v.value

        def timed(fn):
            fn()
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(K):
                fn()
            ctx.synchronize()
            return (time.perf_counter() - t0) * 1e3 / K

        f64 = timed(lambda: L.check(lib.lair_dgesv_device(ctx.h, n, r64, da.ptr, ld, db64.ptr, r64, dx64.ptr,
                                                          r64), "dgesv"))
        f64_be = berr64()
        info64 = _capi.IrInfo()
        mx = timed(lambda: lib.lair_dgesv_mixed_device(ctx.h, n, r64, da.ptr, ld, db64.ptr, r64, dx64.ptr, r64,
                                                       PREC_NAMES[head], 20, 0.0, ctypes.byref(info64)))
        c3_64 = {"nrhs": r64, "fp64_ms": round(f64, 3), "mixed_ms": round(mx, 3), "precision": head,
                 "speedup_vs_fp64": round(f64 / mx, 3), "ir_iterations": info64.iterations,
                 "fell_back": info64.fell_back, "backward_error": berr64(), "fp64_backward_error": f64_be}
        log(f"nrhs=64: fp64 {f64:.2f} ms, {head} {mx:.2f} ms ({info64.iterations} IR steps)")
        db64.free()
        dx64.free()

    side = {}
    if not args.no_side and not args.only_mixed:
        for d in (da, db, dx):
            d.free()
        for name, dt in (("fp64", np.float64), ("fp32", np.float32)):
            side[f"batched_{name}"] = bench_batched(ctx, gl, dt, max(K, 3))
            log(f"batched {name}: {side[f'batched_{name}']['ms']:.3f} ms for 20000 systems "
                f"({side[f'batched_{name}']['hbm_gbs']} GB/s algorithmic)")
        side["adjoint"] = bench_adjoint(ctx, gl, max(K, 3))
        log(f"adjoint n=4096 nrhs=64: {side['adjoint']['ms']:.3f} ms")

    plain = {}
    if not args.no_plain and not args.only_mixed:
        plain["c1"] = bench_plain(ctx, gl, 512, 1, max(K, 5), 0, 512)
        plain["c2"] = bench_plain(ctx, gl, 8192, 16, max(K, 3), -1, 8192)
        log(f"plain FP64 getrf+getrs: n=512 {plain['c1']['ms']:.3f} ms, n=8192 nrhs=16 {plain['c2']['ms']:.2f} ms")

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(n, args.cpu_sample_n)
            if "c1" in plain:
                c1_cpu = cpu_solve_ms(512, 512)
                plain["c1"]["cpu_ms_one_core"] = round(c1_cpu, 2)
                plain["c1"]["speedup_vs_cpu"] = round(c1_cpu / plain["c1"]["ms"], 1)
            if "c2" in plain:
                plain["c2"]["cpu_ms_one_core_extrapolated"] = round(cpu["value"] * (8192 / n) ** 3, 1)
        except Exception as e:  # the oracle build is optional on a box without gcc
            log(f"cpu baseline skipped: {e}")

    line = {
        "metric": METRIC,
        "value": round(hr["ms"], 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(hr["ms"], 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": ("bf16x6-mfma+f64-ir" if head.startswith("bf16x6") else "f32-mfma+f64-ir"),
        "data": "synthetic",
        "config": {
            "workload": f"dgesv n={n} nrhs={nrhs} kappa={args.kappa:g} (Haar U,V; log-spaced s)",
            "n": n, "nrhs": nrhs, "kappa": args.kappa, "precision": head,
            "fp64_ms": round(fp64_ms, 3) if fp64_ms else None,
            "speedup_vs_fp64": round(fp64_ms / hr["ms"], 3) if fp64_ms else None,
            "backward_error": hr["backward_error"],
            "fp64_backward_error": fp64_berr,
            "tolerance": n * EPS,
            "ir_iterations": hr["iterations"],
            "variants": {k: {"ms": round(v["ms"], 3), "iterations": v["iterations"],
                             "backward_error": v["backward_error"], "fell_back": v["fell_back"]}
                         for k, v in results.items()},
            "h2d_ms_A": round(h2d_ms, 2),
            "breakdown_ms": breakdown,
            "c3_nrhs64": c3_64,
            "c1": plain.get("c1"),
            "c2": plain.get("c2"),
            "batched_c4": side.get("batched_fp64"),
            "batched_c4_fp32": side.get("batched_fp32"),
            "adjoint_c5": side.get("adjoint"),
        },
        "roofline": roof,
        "roofline_gemm": roof_g,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    if not side:
        for d in (da, db, dx):
            d.free()
    ctx.close()


if __name__ == "__main__":
    main()
