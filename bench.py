#!/usr/bin/env python3
"""Benchmark: one TEBD sweep on one MI355X.

Workload (the configuration the headline metric is quoted on): open MPS,
n = 64 sites, d = 2, initial |0...0>, depth-20 brickwork of Haar-random
4x4 gates (seed 20261203, serpentine order: 630 gates), simple_update! with
maxdim = 256 and absolute threshold 1e-10.  A "step" is one full sweep of
630 gates; the state starts on the device and only the kept bond dimension
k crosses back to the host per gate (the plain form).

Forms: fp64 = contractions in FP64 MFMA; split_* = contractions in 8-bit
split slices on bf16/fp16 MFMA with exact FP32 accumulation.  QR and SVD are
FP64 in every form.  Accuracy: relative l2 error of 1024 seeded amplitudes
<x|psi> against the FP64 oracle's amplitudes (tests/golden fixture).

Prints one JSON line (rank 0, N = 1).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "tenet.jl_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

N_SITES, DEPTH, CHI, THR = 64, 20, 256, 1e-10
SEED = 20261202 + 1
FIXTURE = os.path.join(ROOT, "tests", "golden", "tebd_n64_d20_chi256.npz")
PEAK_FP64_TFLOPS = 78.6

FORMS = {"fp64": 0, "fp32": 1, "split_bf16": 2, "split_fp16": 3}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_form(ctx, T, sched, gate_views, prec, steps, warmup, bits):
    import numpy as np
    from tenet_hip.tebd import tebd_sweep

    def one():
        psi = tebd_sweep(ctx, N_SITES, sched, gate_views, CHI, THR, prec=prec)
        ctx.sync()
        return psi

    for w in range(warmup):
        t0 = time.perf_counter()
        one()
        log(f"prec {prec} warmup {w} {time.perf_counter() - t0:.3f}s")
    ctx.sync()
    ms = []
    psi = None
    # live per-launch timing of the dominant kernel (HIP events on its stream)
    ctx.stats_enable(T._lib.KSTAT_SVD_ROUND)
    ctx.timer_start()
    t_all = time.perf_counter()
    for s in range(steps):
        t0 = time.perf_counter()
        psi = one()
        ms.append((time.perf_counter() - t0) * 1e3)
        log(f"prec {prec} step {s} {ms[-1]:.1f} ms")
    wall = time.perf_counter() - t_all
    ev_ms = ctx.timer_stop()
    kst = ctx.stats_read()
    ctx.stats_enable(T._lib.KSTAT_OFF)
    amps = psi.amplitudes(bits)
    flops = sum(psi.flops.values())
    return {"ms_per_step": wall * 1e3 / steps, "event_ms_per_step": ev_ms / steps, "step_ms": ms, "amps": amps,
            "flops": flops, "flops_split": dict(psi.flops), "bonds": psi.bondsizes(), "kstats": kst}


def cpu_baseline(sched, gates, budget_s):
    """The oracle (numpy/LAPACK restatement of Tenet's simple_update! loop)
    on the host cores, on the first gates of the same sweep until the time
    budget is spent."""
    import numpy as np
    from oracle import tenet_oracle as O

    psi = O.product_state_mps([[1.0, 0.0]] * N_SITES)
    flops = 0.0
    t0 = time.perf_counter()
    done = 0
    for a, U in zip(sched, gates):
        # nominal work of this gate with the same conventions as the device path
        lo, hi = psi.form[1], psi.form[2]
        for i in range(lo, a):
            chil, d, chi = psi.A[i].shape
            if chi > 1:
                flops += _qr_fl(chil * d, chi) + 8.0 * min(chil * d, chi) * chi * psi.A[i + 1].shape[1] * psi.A[i + 1].shape[2]
        for i in range(hi, a + 1, -1):
            chi, d, chir = psi.A[i].shape
            if chi > 1:
                flops += _qr_fl(d * chir, chi) + 8.0 * psi.A[i - 1].shape[0] * psi.A[i - 1].shape[1] * chi * min(d * chir, chi)
        O.canonize_mixed(psi, a, a + 1)
        chil, d1, chi = psi.A[a].shape
        _, d2, chir = psi.A[a + 1].shape
        flops += 8.0 * chil * d1 * chi * d2 * chir + 8.0 * chil * chir * (d1 * d2) ** 2 + _svd_fl(chil * d1, d2 * chir)
        O.simple_update_mps(psi, U, a, CHI, THR)
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return flops / dt / 1e9, done, dt


def _qr_fl(m, n):
    p, q = max(m, n), min(m, n)
    return 4.0 * (4.0 * p * q * q - 4.0 * q ** 3 / 3.0)


def _svd_fl(m, n):
    p, q = max(m, n), min(m, n)
    return 4.0 * (4.0 * p * p * q + 8.0 * p * q * q + 9.0 * q ** 3)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--forms", default="fp64,split_bf16")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: this engine has no multi-GPU path (the TEBD sweep is "
              "sequential in MixedCanonical form); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import tenet_hip as T
    from tenet_hip.tebd import brickwork_schedule, haar_gates, upload_gates

    fx = np.load(FIXTURE)
    bits = np.unpackbits(fx["bits"], axis=1)[:, :N_SITES]
    ref_amps = fx["amps"]
    sched = brickwork_schedule(N_SITES, DEPTH)
    gates = haar_gates(SEED, len(sched))
    ctx = T.Context(0)
    gate_views = upload_gates(ctx, gates)
    results = {}
    for form in [f for f in args.forms.split(",") if f]:
        log(f"form {form}")
        r = run_form(ctx, T, sched, gate_views, FORMS[form], args.steps, args.warmup, bits)
        r["relerr"] = float(np.linalg.norm(r["amps"] - ref_amps) / np.linalg.norm(ref_amps))
        r["bonds_match"] = bool(list(r["bonds"]) == list(fx["bonds"]))
        results[form] = r
        log(f"form {form}: {r['ms_per_step']:.1f} ms/sweep, relerr {r['relerr']:.2e}, "
            f"{r['flops'] / 1e9:.1f} GFLOP nominal")
    main_form = min(results, key=lambda f: results[f]["ms_per_step"] if results[f]["relerr"] <= 1e-8 else 1e30)
    r = results[main_form]
    gflops = r["flops"] / (r["ms_per_step"] * 1e-3) / 1e9

    k = r["kstats"]
    roof = None
    if k["launches"] > 0 and k["ms"] > 0:
        ach = k["flops"] / (k["ms"] * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": round(ach, 4), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / PEAK_FP64_TFLOPS, 5), "traffic": None,
                "kernel": "svd_round_kernel (one-sided Jacobi round: pair Grams, rotations, apply)",
                "launches": k["launches"], "avg_launch_us": round(k["ms"] * 1e3 / k["launches"], 2),
                "kernel_share_of_step": round(k["ms"] / (r["ms_per_step"] * args.steps), 3),
                "algorithmic_bytes_per_launch": round(k["bytes"] / k["launches"])}
    cpu = None
    if not args.no_cpu:
        threads = os.environ.get("OPENBLAS_NUM_THREADS") or str(os.cpu_count())
        cg, ng, dt = cpu_baseline(sched, gates, args.cpu_seconds)
        cpu = {"value": round(cg, 3), "unit": "GFLOP/s", "cores": int(threads), "kind": "port",
               "sample": f"oracle simple_update! loop on the first {ng} of 630 gates of the same sweep "
                         f"({dt:.1f} s, numpy/OpenBLAS LAPACK, {threads} threads)"}

    line = {
        "metric": "TEBD sweep time & GFLOP/s split-prec vs FP64 @relerr<=1e-8; req/s + p99 latency",
        "value": round(gflops, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128" if main_form == "fp64" else f"c128 ({main_form} contractions, fp64 qr/svd)",
        "data": "synthetic (seeded Haar gates, |0...0>)",
        "config": {"workload": "TEBD brickwork n=64 d=2 depth=20 chi_max=256 threshold=1e-10 (630 gates)",
                   "form": main_form, "relerr_vs_fp64_oracle": r["relerr"],
                   "forms": {f: {"ms_per_sweep": round(v["ms_per_step"], 3), "relerr": v["relerr"],
                                 "gflop_nominal": round(v["flops"] / 1e9, 2), "bonds_match_oracle": v["bonds_match"]}
                             for f, v in results.items()}},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
