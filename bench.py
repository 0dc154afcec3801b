"""Benchmark of the rVPM particle-interaction step on one MI355X.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]

Workload: the direct UJ at scale -- N = 1,048,576 particles, X ~ U[-1,1]^3,
Gamma ~ N(0, 1e-3), sigma ~ U[0.02, 0.05], vol = sigma^3, gaussian-erf kernel.
One step is one RK3 substep on the HBM-resident field: the all-pairs U + J
evaluation (mixed FP32 pair math with FP64 accumulation) followed by the fused
FP64 stretching / rVPM right-hand side / low-storage update, so the default
three steps are one full RK3 time step.  `value` is N^2 pair interactions per
second of the timed steps.

Also reported on the same input: the FP64 path's rate and the mixed path's
normwise error against it (U, J, dGamma/dt), the FP64 path against the CPU
oracle on a target sample, the Winckelmans kernel's rate, the roofline of the
pair kernel (timed with HIP events on the library's stream) and the CPU
baseline (the oracle's OpenMP UJ_direct on a bounded target sample).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle interactions/s (UJ direct+FMM, N=1M) mixed vs FP64 at stated rel. err"


def flops_header():
    txt = open(os.path.join(ROOT, "include", "vpmx_flops.h")).read()
    return {m.group(1): float(m.group(2))
            for m in re.finditer(r"#define\s+VPMX_(\w+)\s+([0-9.eE+-]+)", txt)}


def normwise(a, ref):
    d = np.linalg.norm(a - ref, axis=1)
    n = np.linalg.norm(ref, axis=1)
    return float(d.max() / n.max()), float(np.sqrt(np.mean(d ** 2)) / np.sqrt(np.mean(n ** 2)))


def dgamma_dt(U, J, G, sig, transposed=True, f=0.0, g=0.2):
    """rVPM right-hand side of Gamma (docs/src/theory/rvpm.md:196-236), FP64
    numpy, for the error report only."""
    Jm = J.reshape(-1, 3, 3)   # Jm[:, l, k] = J[k + 3 l] = dU_k/dx_l
    S = np.einsum("plk,pl->pk", Jm, G) if transposed else np.einsum("plk,pk->pl", Jm, G)
    Z = (g + f) / (1 + 3 * f) * np.sum(S * G, axis=1) / np.sum(G * G, axis=1)
    return S - 3 * Z[:, None] * G


class Field:
    """Particle matrix + resident device state through the C ABI."""

    def __init__(self, ctx, X, G, s, v):
        from flowunsteady_amd import vpm
        self.vpm = vpm
        self.ctx = ctx
        self.n = len(X)
        self.pf = vpm.ParticleField(self.n)
        vpm.add_particles(self.pf, X, G, s, vol=v)
        self.lay = self.pf.layout

    def upload(self):
        from flowunsteady_amd import _native as N
        self.ctx.check(N.lib().vpmx_upload(self.ctx.ptr, self.pf._ptr(), self.n,
                                           ctypes.byref(self.lay)))

    def download(self):
        from flowunsteady_amd import _native as N
        self.ctx.check(N.lib().vpmx_download(self.ctx.ptr, self.pf._ptr(), self.n,
                                             ctypes.byref(self.lay)))
        return self.vpm.get_U(self.pf).copy(), self.vpm.get_J(self.pf).copy()

    def uj(self, kernel, precision):
        from flowunsteady_amd import _native as N
        P = N.UJParams()
        P.kernel = kernel
        P.precision = precision
        P.transposed = 1
        P.reset = 1
        self.ctx.check(N.lib().vpmx_uj_resident(self.ctx.ptr, ctypes.byref(P)))

    def substep(self, kernel, precision, dt):
        """One resident Euler step: UJ + the fused FP64 stretching / rVPM
        right-hand side / update pass -- the same work as one RK3 substep."""
        from flowunsteady_amd import _native as N
        P = N.UJParams()
        P.kernel = kernel
        P.precision = precision
        P.transposed = 1
        P.reset = 1
        S = N.StepParams()
        S.f, S.g = 0.0, 0.2
        self.ctx.check(N.lib().vpmx_euler_resident(self.ctx.ptr, ctypes.byref(P), ctypes.byref(S),
                                                   dt, 0))


def cpu_baseline(X, G, s, budget_s=12.0):
    """Oracle UJ_direct (OpenMP, FP64) on a bounded target sample against all N
    sources; interactions/s = sample * N / t."""
    from oracle import oracle as O
    n = len(X)
    threads = O.num_threads()
    m = 64
    t0 = time.perf_counter()
    O.uj_direct("gaussianerf", X[:m], X, G, s)
    t = time.perf_counter() - t0
    m2 = int(min(n, max(m, m * budget_s / max(t, 1e-3))))
    rng = np.random.default_rng(0)
    idx = np.sort(rng.choice(n, m2, replace=False))
    t0 = time.perf_counter()
    O.uj_direct("gaussianerf", X[idx], X, G, s)
    t = time.perf_counter() - t0
    return {"value": m2 * n / t, "unit": "interactions/s", "cores": threads, "kind": "port",
            "sample": f"{m2} targets x {n} sources, gaussian-erf U+J, FP64 OpenMP oracle "
                      f"(C restatement of UJ_direct), {t:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and oracle checks")
    ap.add_argument("--fp64-steps", type=int, default=1)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is "
                                   "described in DESIGN.md but not built", "n_gpus": args.gpus}))
        sys.exit(2)

    from flowunsteady_amd import _native as N
    from flowunsteady_amd import workloads

    fl = flops_header()
    n = args.n
    X, G, s, v = workloads.uniform_cube(n=n, seed=2)
    ctx = N.Context(0)
    fld = Field(ctx, X, G, s, v)
    pairs = float(n) * float(n)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)

    # ---- accuracy on the initial state: mixed vs FP64 GPU (full N) --------------
    fld.upload()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.fp64_steps):
        fld.uj(N.GAUSSIANERF, N.FP64)
    t_fp64 = (time.perf_counter() - t0) / args.fp64_steps
    k64_ms, k64_n = ctx.timing(N.T_UJ)
    U64, J64 = fld.download()
    log(f"fp64 UJ: {t_fp64:.3f} s/eval ({pairs / t_fp64:.3e} interactions/s)")
    ctx.set_timing(False)
    fld.uj(N.GAUSSIANERF, N.MIXED)
    Umx, Jmx = fld.download()
    err = {"U": normwise(Umx, U64), "J": normwise(Jmx, J64),
           "dGamma_dt": normwise(dgamma_dt(Umx, Jmx, G, s), dgamma_dt(U64, J64, G, s))}
    log("mixed vs fp64:", err)

    # ---- Winckelmans rate (one evaluation each) -------------------------------------
    wk = {}
    for prec, name in ((N.MIXED, "mixed"), (N.FP64, "fp64")):
        t0 = time.perf_counter()
        fld.uj(N.WINCKELMANS, prec)
        wk[name] = pairs / (time.perf_counter() - t0)
    log("winckelmans:", wk)

    # ---- timed steps (mixed, gaussian-erf) ----------------------------------------------
    fld.upload()
    for _ in range(args.warmup):
        fld.substep(N.GAUSSIANERF, N.MIXED, 1e-4)
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fld.substep(N.GAUSSIANERF, N.MIXED, 1e-4)
    elapsed = time.perf_counter() - t0
    kms, kn = ctx.timing(N.T_UJ)
    ums, un = ctx.timing(N.T_UPDATE)
    ctx.set_timing(False)
    value = pairs * args.steps / elapsed
    kernel_s = kms / 1e3 / max(kn, 1)
    achieved = fl["FLOPS_UJ_PAIR_GAUSSIANERF"] * pairs / kernel_s / 1e12
    peak = fl["PEAK_FP32_VECTOR_TFLOPS"]
    log(f"mixed step: {elapsed / args.steps:.3f} s, kernel {kernel_s:.3f} s, {achieved:.1f} TFLOP/s")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair math, f64 accumulation",
        "data": "synthetic",
        "config": {"workload": "direct UJ, uniform cube N=1048576 (X~U[-1,1]^3, "
                               "Gamma~N(0,1e-3), sigma~U[0.02,0.05]), gaussian-erf, U+J + "
                               "stretching/rVPM update per step",
                   "N": n, "kernel": "gaussianerf", "scheme": "direct",
                   "step": "one RK substep: UJ + fused FP64 update",
                   "parallelism": "single GPU"},
        "fp64": {"value": pairs / t_fp64, "unit": "interactions/s",
                 "kernel_ms": k64_ms / max(k64_n, 1)},
        "speedup_mixed_vs_fp64": value / (pairs / t_fp64),
        "rel_err_mixed_vs_fp64": {k: {"max": e[0], "rms": e[1]} for k, e in err.items()},
        "winckelmans": wk,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None,
                     "kernel": "uj_mixed_kernel<gaussianerf>",
                     "kernel_ms": 1e3 * kernel_s,
                     "flops_per_pair": fl["FLOPS_UJ_PAIR_GAUSSIANERF"]},
        "update_kernel_ms": ums / max(un, 1),
    }

    if not args.no_cpu:
        from oracle import oracle as O
        idx = np.sort(np.random.default_rng(1).choice(n, 256, replace=False))
        Ur, Jr = O.uj_direct("gaussianerf", X[idx], X, G, s)
        out["fp64_vs_oracle"] = {"U_max": normwise(U64[idx], Ur)[0],
                                 "J_max": normwise(J64[idx], Jr)[0], "targets": 256}
        out["mixed_vs_oracle"] = {"U_max": normwise(Umx[idx], Ur)[0],
                                  "J_max": normwise(Jmx[idx], Jr)[0], "targets": 256}
        log("oracle checks:", out["fp64_vs_oracle"], out["mixed_vs_oracle"])
        out["cpu_baseline"] = cpu_baseline(X, G, s)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
