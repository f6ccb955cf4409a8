#!/usr/bin/env python3
"""Benchmark: the TEBD sweep on one MI355X (and, with --workload, the other
configurations: plain, heff, grouped, serving).

Default workload (the configuration the headline metric is quoted on): open
MPS, n = 64 sites, d = 2, initial |0...0>, depth-20 brickwork of Haar-random
4x4 gates (seed 20261203, serpentine order: 630 gates), maxdim = 256 and
absolute threshold 1e-10.  A "step" is one full sweep of 630 gates; the
gates and the state are resident on the device before the timed region and
only kept bond dimensions cross back to the host.

Forms (QR and SVD are FP64 in every form):
  vidal       the batched form: each brickwork layer's 31-32 gates act on
              disjoint bonds, so a layer is a few grouped contraction
              launches and one grouped Jacobi SVD (Vidal gauge, Hastings'
              inverse-free update); the headline line.  Accuracy against the
              oracle's Vidal-gauge restatement on the same gates.
  vidal_graph the same batched form with the steady layers (bond
              dimensions saturated) replayed from captured HIP graphs.
  fp64        the plain FP64 GPU form: simple_update! gate by gate in
              MixedCanonical form, one SVD at a time with the kept bond
              dimension read back per gate (no grouping, no graph).
  split_bf16 / split_fp16   the plain form with its contractions on 8-bit
              bf16 / fp16 split slices (exact FP32 accumulation).
  fp32        the plain form with ComplexF32-grade (FP32 MFMA) contractions.
The plain forms are checked against the oracle's MixedCanonical sweep.
Accuracy: relative l2 error of 1024 seeded amplitudes <x|psi> against the
FP64 oracle's (tests/golden fixtures), and every bond dimension equal.

Prints one JSON line (rank 0, N = 1).
"""
import argparse
import json
import os
import platform
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "tenet.jl_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

N_SITES, DEPTH, CHI, THR = 64, 20, 256, 1e-10
SEED = 20261202 + 1
FIXTURE = os.path.join(ROOT, "tests", "golden", "tebd_n64_d20_chi256.npz")
VIDAL_FIXTURE = os.path.join(ROOT, "tests", "golden", "tebd_vidal_n64_d20_chi256.npz")
PEAK_FP64_TFLOPS = 78.6
# HBM bytes per round-kernel dispatch (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 KiB,
# factors from the scale kernel over 256 MiB) for each TEBD form, from
# rocprofv3 --pmc passes over one sweep; None where not measured.
ROUND_TRAFFIC_BYTES = {"fp64": 17_005_568, "vidal": 216_834_819, "vidal_graph": 216_834_819}

FORMS = {"fp64": 0, "fp32": 1, "split_bf16": 2, "split_fp16": 3, "vidal": 0, "vidal_graph": 0}
DEFAULT_FORMS = "vidal,vidal_graph,fp64,split_bf16,split_fp16,fp32"
VIDAL_FORMS = ("vidal", "vidal_graph")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _timed_sweeps(ctx, T, one, steps, warmup, tag):
    """warmup untimed sweeps, then `steps` timed ones with the round kernels'
    per-launch timing on (HIP events on their stream); returns the last
    result, the wall time per sweep and the round-kernel statistics."""
    for w in range(warmup):
        t0 = time.perf_counter()
        r = one()
        log(f"{tag} warmup {w} {time.perf_counter() - t0:.3f}s")
        if hasattr(r, "close"):
            r.close()
    ctx.sync()
    ctx.stats_enable(T._lib.KSTAT_SVD_ROUND)
    t_all = time.perf_counter()
    res = None
    for s in range(steps):
        if res is not None and hasattr(res, "close"):
            res.close()
        t0 = time.perf_counter()
        res = one()
        log(f"{tag} step {s} {(time.perf_counter() - t0) * 1e3:.1f} ms")
    wall = (time.perf_counter() - t_all) / steps
    kst = ctx.stats_read()
    ctx.stats_enable(T._lib.KSTAT_OFF)
    return res, wall, kst


def run_form(ctx, T, form, sched, layers, gate_views, steps, warmup, bits):
    from tenet_hip.tebd import tebd_sweep
    from tenet_hip.vidal import tebd_vidal_sweep

    if form in VIDAL_FORMS:
        def one():
            v = tebd_vidal_sweep(ctx, N_SITES, layers, gate_views, CHI, THR, graphs=(form == "vidal_graph"))
            ctx.sync()
            return v
    else:
        def one():
            psi = tebd_sweep(ctx, N_SITES, sched, gate_views, CHI, THR, prec=FORMS[form])
            ctx.sync()
            return psi

    res, wall, kst = _timed_sweeps(ctx, T, one, steps, warmup, form)
    psi = res.to_mps() if form in VIDAL_FORMS else res
    amps = psi.amplitudes(bits)
    out = {"ms_per_step": wall * 1e3, "amps": amps, "flops": sum(res.flops.values()),
           "flops_split": dict(res.flops), "bonds": res.bondsizes(), "kstats": kst, "steps": steps,
           "warmup": warmup}
    if form in VIDAL_FORMS:
        res.close()
    return out


def _cpu_info():
    model = platform.processor() or ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def _saturated_gate_sample(form, rng):
    """Host tensors of one gate of the saturated sweep: sites (256, 2, 256),
    a Haar gate, and (Vidal) the Lambda of the left bond."""
    from oracle import tenet_oracle as O

    A = O.complex_gaussian(rng, (CHI, 2, CHI))
    B = O.complex_gaussian(rng, (CHI, 2, CHI))
    lam = np.sort(rng.random(CHI))[::-1] if form == "vidal" else None
    return A, B, O.haar_unitary(rng, 4), lam


def _cpu_gate(form, A, B, U, lam):
    """One saturated gate as the oracle runs it: the Vidal-gauge update
    (oracle.vidal_gate), or for the plain form the canonize! move onto the
    gate (a 512 x 256 thin QR and R's contraction into the next site) and
    Muscle.simple_update (oracle.simple_update: theta, gate, 512 x 512 SVD,
    truncation to 256, sqrt(s) into both sides)."""
    from oracle import tenet_oracle as O

    if form == "vidal":
        # sites 1, 2 of a three-site chain: X = Lambda_0 theta~ is 512 x 512
        O.vidal_gate(O.OVidal([A[:1].copy(), A, B], [lam, np.ones(CHI)]), U, 1, CHI, THR)
        return
    ta = O.OTensor(A, ["l", "pa", "b"])
    q, r = O.tensor_qr_thin(ta, ["l", "pa"], ["b"], "v")
    tb = O.binary_einsum(r, O.OTensor(B, ["b", "pb", "r"]), dims=["b"]).rename({"v": "b"})
    G = O.OTensor(U.reshape(2, 2, 2, 2), ["oa", "ob", "pa", "pb"])
    O.simple_update(q.rename({"v": "b"}), "pa", tb, "pb", "b", G, "oa", "ob", maxdim=CHI, threshold=THR)


def _gate_flops(form):
    m = 2 * CHI
    fl = 8.0 * m * CHI * m + 8.0 * CHI * CHI * 16 * 4 + _svd_fl(m, m)  # theta, gate, SVD
    if form == "vidal":
        return fl + 8.0 * m * m * CHI  # B_a = theta~ V
    return fl + _qr_fl(m, CHI) + 8.0 * CHI * CHI * m  # the canonize! move's QR and R . B


def cpu_baseline_tebd(form, budget_s):
    """The oracle (numpy / OpenBLAS LAPACK: gesdd, Householder QR) on the
    host cores, on a bounded sample of the same workload: gates at the
    sweep's saturated shape (chi = 256: 512 x 512 bond matrices, 49 of the
    63 bonds end there), one warm-up block, then the median of 5 timed
    blocks, at all host threads and at 1 thread (threadpoolctl)."""
    from threadpoolctl import threadpool_limits

    rng = np.random.default_rng(5)
    samples = [_saturated_gate_sample(form, rng) for _ in range(4)]
    model, ncpu = _cpu_info()
    out = {}
    for label, threads in (("all", None), ("one", 1)):
        ctx_lim = threadpool_limits(limits=threads) if threads else threadpool_limits(limits=None)
        with ctx_lim:
            t0 = time.perf_counter()
            _cpu_gate(form, *samples[0])  # warm-up
            per_gate = time.perf_counter() - t0
            # block size from the warm-up: ~budget / 12 per block
            nblk = max(1, min(8, int(budget_s / 12.0 / max(per_gate, 1e-3))))
            times = []
            for rep in range(5):
                t0 = time.perf_counter()
                for g in range(nblk):
                    _cpu_gate(form, *samples[g % len(samples)])
                times.append((time.perf_counter() - t0) / nblk)
            med = statistics.median(times)
        out[label] = {"s_per_gate": med, "gflops": _gate_flops(form) / med / 1e9, "gates_per_block": nblk}
    threads = int(os.environ.get("OPENBLAS_NUM_THREADS") or ncpu)
    return out, model, threads


def _qr_fl(m, n):
    p, q = max(m, n), min(m, n)
    return 4.0 * (4.0 * p * q * q - 4.0 * q ** 3 / 3.0)


def _svd_fl(m, n):
    p, q = max(m, n), min(m, n)
    return 4.0 * (4.0 * p * p * q + 8.0 * p * q * q + 9.0 * q ** 3)


# ---------------------------------------------------------------- DMRG H_eff
HEFF_CHI, HEFF_W, HEFF_SEED = 2048, 5, 20261202 + 3
HEFF_FORMS = {"fp64": 0, "fp32": 1, "split_bf16": 2, "split_fp16": 3, "split_i8": 4, "split_i8_fast": 5}
HEFF_SAMPLE_ROWS = 256


def heff_flops(chi, w):
    lt = 8.0 * (chi * w) * (4 * chi) * chi
    op = 8.0 * (chi * chi) * (w * 4) * (w * 4)
    rt = 8.0 * (chi * 4) * chi * (w * chi)
    return lt + op + rt


def bench_heff(args):
    """Two-site effective-Hamiltonian apply (DMRG.jl:10-18, 297-300):
    L, R (2048, 5, 2048) and theta (2048, 2, 2, 2048) complex128 Gaussians
    generated on the device (theta normalised), XXZ MPO (J = 1, Delta = 0.5).
    A step is one apply.  Accuracy: relative l2 error of the first 256
    output rows against the FP64 oracle on the same rows."""
    import numpy as np
    import tenet_hip as T
    from tenet_hip import _lib as L
    from tenet_hip.dmrg import heff_apply, xxz_mpo_tensor

    chi, w = HEFF_CHI, HEFF_W
    ctx = T.Context(0)
    Ld = T.DeviceArray(ctx, (chi, w, chi)).gaussian_(HEFF_SEED)
    Rd = T.DeviceArray(ctx, (chi, w, chi)).gaussian_(HEFF_SEED + 1)
    th = T.DeviceArray(ctx, (chi, 2, 2, chi)).gaussian_(HEFF_SEED + 2)
    nrm = np.sqrt(th.nrm2sq())
    th.scale_(1.0 / nrm)
    Wh = xxz_mpo_tensor(1.0, 0.5)
    tL = T.Tensor(Ld, ["l", "w", "lk"])
    tR = T.Tensor(Rd, ["r", "wr", "rk"])
    tth = T.Tensor(th, ["lk", "i1", "i2", "rk"])
    tW1 = T.Tensor.from_host(ctx, Wh, ["w", "m", "o1", "i1"])
    tW2 = T.Tensor.from_host(ctx, Wh, ["m", "wr", "o2", "i2"])
    flops = heff_flops(chi, w)

    ref_rows = None
    cpu = None
    if not args.no_cpu:
        from oracle import synth
        from oracle import tenet_oracle as O

        S = HEFF_SAMPLE_ROWS
        Lh = synth.gaussian(HEFF_SEED, chi * w * chi).reshape(chi, w, chi, order="F")[:S].copy()
        Rh = synth.gaussian(HEFF_SEED + 1, chi * w * chi).reshape(chi, w, chi, order="F")
        thh = synth.gaussian(HEFF_SEED + 2, chi * 4 * chi).reshape(chi, 2, 2, chi, order="F") / nrm
        O.heff_apply(Lh[:8], Wh, Wh, Rh, thh)  # warm BLAS
        t0 = time.perf_counter()
        ref_rows = O.heff_apply(Lh, Wh, Wh, Rh, thh)
        dt = time.perf_counter() - t0
        threads = os.environ.get("OPENBLAS_NUM_THREADS") or str(os.cpu_count())
        cpu = {"value": round(heff_flops(chi, w) * S / chi / dt / 1e9, 3), "unit": "GFLOP/s", "cores": int(threads),
               "kind": "port",
               "sample": f"oracle H_eff apply on the first {S} of {chi} output rows ({dt:.1f} s, numpy/OpenBLAS, "
                         f"{threads} threads)"}
        del Rh, thh

    results = {}
    for form in [f for f in args.forms.split(",") if f]:
        prec = HEFF_FORMS[form]
        for _ in range(args.warmup):
            y = heff_apply(tL, tW1, tW2, tR, tth, prec=prec)
            ctx.sync()
            del y
        ctx.stats_sample_every(1)
        ctx.stats_enable(L.KSTAT_GEMM)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            y = heff_apply(tL, tW1, tW2, tR, tth, prec=prec)
            ctx.sync()
        wall = (time.perf_counter() - t0) / args.steps
        kst = ctx.stats_read()
        ctx.stats_enable(L.KSTAT_OFF)
        ctx.stats_sample_every(17)
        rel = None
        if ref_rows is not None:
            yh = y.to_host()[:HEFF_SAMPLE_ROWS]
            rel = float(np.linalg.norm(yh - ref_rows) / np.linalg.norm(ref_rows))
        del y
        results[form] = {"ms": wall * 1e3, "relerr": rel, "kstats": kst}
        log(f"heff {form}: {wall * 1e3:.2f} ms/apply, {flops / wall / 1e12:.2f} TFLOP/s, relerr {rel}")

    main_form = min(results, key=lambda f: results[f]["ms"])
    r = results[main_form]
    k = r["kstats"]
    roof = None
    if k["launches"] > 0 and k["ms"] > 0:
        ach = k["flops"] / (k["ms"] * 1e-3) / 1e12
        # peak of the format the result is accurate to: FP64 for the split
        # forms (their executed slice products against the matrix-core peak
        # of their slice type are reported beside it)
        peak = 157.3 if main_form == "fp32" else PEAK_FP64_TFLOPS
        roof = {"bound": "mfma", "achieved": round(ach, 4), "peak": peak, "unit": "TFLOP/s",
                "frac": round(ach / peak, 5), "traffic": None, "kernel": "contraction kernels of the apply",
                "launches": k["launches"], "avg_launch_us": round(k["ms"] * 1e3 / k["launches"], 2)}
        pairs = {"split_bf16": 21, "split_fp16": 21, "split_i8": 28, "split_i8_fast": 15}.get(main_form)
        if pairs:
            mpeak = 5000.0 if "i8" in main_form else 2500.0
            ex = ach * pairs
            roof.update({"executed_slice_products_tops": round(ex, 2), "slice_mfma_peak": mpeak,
                         "executed_frac": round(ex / mpeak, 4),
                         "peak_basis": "FP64 matrix peak, the format the split result is accurate to; frac above 1 "
                                       "is FP64-accurate work faster than native FP64 can run; executed_frac "
                                       "prices the slice products against their own matrix-core peak",
                         "executed_basis": f"{pairs} slice-pair products, each the apply's nominal work"})
    line = {
        "metric": "DMRG two-site H_eff apply throughput (nominal complex flops)",
        "value": round(flops / (r["ms"] * 1e-3) / 1e9, 3), "unit": "GFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(r["ms"], 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c128",
        "data": "synthetic (device-generated seeded Gaussians)",
        "config": {"workload": "DMRG H_eff apply L,R (2048,5,2048) theta (2048,2,2,2048) XXZ J=1 Delta=0.5",
                   "form": main_form, "gflop_nominal": round(flops / 1e9, 1),
                   "forms": {f: {"ms_per_apply": round(v["ms"], 3), "relerr_vs_fp64_oracle": v["relerr"],
                                 "tflops": round(flops / (v["ms"] * 1e-3) / 1e12, 3)} for f, v in results.items()}},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


# ---------------------------------------------------------------- grouped batch
GROUPED_COUNT, GROUPED_SEED = 8192, 20261202 + 2


def bench_grouped(args):
    """8192 uneven pairwise contractions (ranks 3-5, extents log-uniform
    2..128, 1-3 contracted labels, operands and result <= 2^20 elements,
    complex64 Gaussians generated on the device) in one persistent grouped
    launch, plus 8192 truncated SVDs (sides log-uniform 16..512, complex64
    input, FP64 factorisation, maxdim = ceil(min(m, n) / 2), threshold 1e-10)
    in one grouped SVD.  A step is both batches; inputs stay resident."""
    import ctypes as C

    import numpy as np
    import tenet_hip as T
    from tenet_hip import _lib as L
    from tenet_hip.grouped import GroupedPlan, autotune_tiles, grouped_svd
    from tenet_hip.workloads import contraction_work, grouped_contractions, grouped_svds, svd_work

    ctx = T.Context(0)
    t0 = time.perf_counter()
    specs = grouped_contractions(GROUPED_SEED, GROUPED_COUNT)
    fl_c, by_c = contraction_work(specs)
    sizes = [sp.sizes for sp in specs]
    in_off, out_off = [], []
    ai = ao = 0
    for a, b, c in sizes:
        in_off.append((ai, ai + a))
        ai += a + b
        out_off.append(ao)
        ao += c
    arena_in = T.DeviceArray(ctx, (ai,), np.complex64).gaussian_(GROUPED_SEED)
    arena_out = T.DeviceArray(ctx, (ao,), np.complex64)
    costs = autotune_tiles(ctx, np.complex64)
    probs = [(sp.shape(sp.la), sp.la, sp.shape(sp.lb), sp.lb, sp.shape(sp.lc), sp.lc) for sp in specs]
    plan = GroupedPlan(ctx, probs, np.complex64)
    info = plan.info()
    ptrs = np.zeros(3 * len(specs), dtype=np.uint64)
    for i, ((oa, ob), oc) in enumerate(zip(in_off, out_off)):
        ptrs[3 * i] = arena_in.ptr + 8 * oa
        ptrs[3 * i + 1] = arena_in.ptr + 8 * ob
        ptrs[3 * i + 2] = arena_out.ptr + 8 * oc
    ptrs_dev = T.DeviceArray(ctx, (len(ptrs) // 2 + 1,), np.complex128)
    L.check(L.lib.tenet_memcpy_h2d(ctx.h, C.c_void_p(ptrs_dev.ptr), ptrs.ctypes.data_as(C.c_void_p), ptrs.nbytes,
                                   None))
    log(f"grouped: plan built in {time.perf_counter() - t0:.1f}s, {info}, tile costs {costs}")

    ms, ns, maxdims = grouped_svds(GROUPED_SEED + 1000, GROUPED_COUNT)
    fl_s = svd_work(ms, ns)
    svd_in = T.DeviceArray(ctx, (int(np.sum(ms * ns)),), np.complex64).gaussian_(GROUPED_SEED + 1000)
    mats = []
    o = 0
    for m, n in zip(ms, ns):
        v = T.DeviceArray(ctx, (int(m), int(n)), np.complex64, ptr=svd_in.ptr + 8 * o, owner=False)
        v._base = svd_in
        mats.append(v)
        o += int(m * n)

    def contract():
        plan.exec_dev(ptrs_dev.ptr)
        ctx.sync()

    for _ in range(args.warmup):
        contract()
    ctx.stats_sample_every(1)
    ctx.stats_enable(L.KSTAT_GEMM)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        contract()
    t_c = (time.perf_counter() - t0) / args.steps
    kst = ctx.stats_read()
    ctx.stats_enable(L.KSTAT_OFF)
    log(f"grouped contractions: {t_c * 1e3:.2f} ms, {fl_c / t_c / 1e12:.2f} TFLOP/s, {by_c / t_c / 1e9:.0f} GB/s")

    svd_res = None
    ctx.stats_sample_every(17)
    ctx.stats_enable(L.KSTAT_SVD_ROUND)
    t0 = time.perf_counter()
    for _ in range(max(1, args.steps)):
        svd_res = grouped_svd(ctx, mats, maxdims, 1e-10)
        ctx.sync()
    t_s = (time.perf_counter() - t0) / max(1, args.steps)
    kss = ctx.stats_read()
    ctx.stats_enable(L.KSTAT_OFF)
    log(f"grouped SVDs: {t_s * 1e3:.1f} ms, {fl_s / t_s / 1e12:.2f} TFLOP/s nominal")

    # accuracy on a sample, against the FP64 oracle on regenerated inputs
    rel_c = rel_s = None
    cpu = None
    if not args.no_cpu:
        from oracle import synth
        from oracle import tenet_oracle as O

        rng = np.random.default_rng(1)
        errs = []
        out_h = None
        for i in rng.choice(len(specs), 48, replace=False):
            sp = specs[int(i)]
            a, b, c = sizes[int(i)]
            oa, ob = in_off[int(i)]
            A = synth.gaussian(GROUPED_SEED, a, oa).reshape(sp.shape(sp.la), order="F")
            B = synth.gaussian(GROUPED_SEED, b, ob).reshape(sp.shape(sp.lb), order="F")
            ref = O.binary_einsum(O.OTensor(A, sp.la), O.OTensor(B, sp.lb), blas=True).permute(sp.lc).data
            got = np.empty(c, dtype=np.complex64)
            L.check(L.lib.tenet_memcpy_d2h(ctx.h, got.ctypes.data_as(C.c_void_p),
                                           C.c_void_p(arena_out.ptr + 8 * out_off[int(i)]), got.nbytes, None))
            errs.append(np.linalg.norm(got.reshape(sp.shape(sp.lc), order="F") - ref) / np.linalg.norm(ref))
        rel_c = float(max(errs))
        serr = []
        offs = np.concatenate([[0], np.cumsum(ms * ns)])
        for i in rng.choice(len(ms), 16, replace=False):
            m, n = int(ms[i]), int(ns[i])
            Ah = synth.gaussian(GROUPED_SEED + 1000, m * n, int(offs[i]), np.complex64).reshape(m, n, order="F")
            s_ref = np.linalg.svd(Ah.astype(np.complex128), compute_uv=False)  # the same complex64 input
            k = svd_res[int(i)][3]
            serr.append(float(np.max(np.abs(svd_res[int(i)][1][:k] - s_ref[:k])) / s_ref[0]))
        rel_s = max(serr)
        # CPU: the oracle (BLAS-routed einsum; LAPACK gesdd) on a bounded sample
        fl_done = 0.0
        done = 0
        t_cpu_c = 0.0
        budget = min(args.cpu_seconds, 20.0)
        for sp, (a, b, c), (oa, ob) in zip(specs, sizes, in_off):
            A = synth.gaussian(GROUPED_SEED, a, oa, np.complex64).reshape(sp.shape(sp.la), order="F")
            B = synth.gaussian(GROUPED_SEED, b, ob, np.complex64).reshape(sp.shape(sp.lb), order="F")
            ts = time.perf_counter()
            O.binary_einsum(O.OTensor(A, sp.la), O.OTensor(B, sp.lb), blas=True)
            t_cpu_c += time.perf_counter() - ts
            fl_done += 8.0 * float(np.prod(sp.mnk))
            done += 1
            if t_cpu_c > budget / 2:
                break
        sv_fl = 0.0
        sv_done = 0
        t_cpu_s = 0.0
        for i in range(len(ms)):
            m, n = int(ms[i]), int(ns[i])
            Ah = synth.gaussian(GROUPED_SEED + 1000, m * n, int(offs[i])).reshape(m, n, order="F")
            ts = time.perf_counter()
            np.linalg.svd(Ah, full_matrices=False)
            t_cpu_s += time.perf_counter() - ts
            sv_fl += svd_work([m], [n])
            sv_done += 1
            if t_cpu_s > budget / 2:
                break
        threads = os.environ.get("OPENBLAS_NUM_THREADS") or str(os.cpu_count())
        cpu = {"value": round((fl_done + sv_fl) / (t_cpu_c + t_cpu_s) / 1e9, 3), "unit": "GFLOP/s",
               "cores": int(threads), "kind": "port",
               "sample": f"oracle: first {done} contractions (complex64 numpy einsum via BLAS, {t_cpu_c:.1f} s) "
                         f"and first {sv_done} SVDs (complex128 LAPACK gesdd, {t_cpu_s:.1f} s), {threads} threads"}

    total = t_c + t_s
    # the dominant kernel is the grouped SVD round (~97% of the step); the
    # contraction kernel's figures ride along
    roof = None
    contraction_kernel = None
    if kst["launches"]:
        ach = kst["bytes"] / (kst["ms"] * 1e-3) / 1e9
        contraction_kernel = {
            "kernel": "grouped_gemm_c64_kernel (8192 contractions, one persistent launch)", "bound": "hbm",
            "achieved": round(ach, 2), "peak": 8000.0, "unit": "GB/s", "frac": round(ach / 8000.0, 4),
            "launches": kst["launches"], "avg_launch_us": round(kst["ms"] * 1e3 / kst["launches"], 1),
            "fp32_mfma_tflops": round(fl_c / (kst["ms"] / kst["launches"] * 1e-3) / 1e12, 3),
            "fp32_mfma_frac": round(fl_c / (kst["ms"] / kst["launches"] * 1e-3) / 1e12 / 157.3, 4)}
    if kss["launches"] and kss["ms"] > 0:
        ach = fl_s * max(1, args.steps) / (kss["ms"] * 1e-3) / 1e12
        executed = kss["flops"] / (kss["ms"] * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": round(ach, 4), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / PEAK_FP64_TFLOPS, 5), "traffic": None,
                "achieved_basis": "nominal SVD flops 4(4p^2q+8pq^2+9q^3) of the batch / round-kernel time",
                "executed_tflops": round(executed, 3), "executed_frac": round(executed / PEAK_FP64_TFLOPS, 4),
                "kernel": "svd_round_mfma_kernel (grouped: a round of every unconverged SVD in each launch; Gram "
                          "and rotation on v_mfma_f64_16x16x4)",
                "launches": kss["launches"], "avg_launch_us": round(kss["ms"] * 1e3 / kss["launches"], 2),
                "kernel_share_of_step": round(kss["ms"] / (total * 1e3 * max(1, args.steps)), 3),
                "contraction_kernel": contraction_kernel}
    line = {
        "metric": "grouped batch throughput (8192 contractions + 8192 truncated SVDs)",
        "value": round((fl_c + fl_s) / total / 1e9, 3), "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(total * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64 (fp32 mfma contractions, fp64 svd)",
        "data": "synthetic (device-generated seeded Gaussians)",
        "config": {"workload": "grouped 8192 contractions (rank 3-5, extents 2..128, <=2^20 elems) + 8192 SVDs "
                               "(16..512, maxdim ceil(min/2), threshold 1e-10)",
                   "contractions": {"ms": round(t_c * 1e3, 3), "gflop": round(fl_c / 1e9, 2),
                                    "gbytes": round(by_c / 1e9, 2), "plan": info, "tile_costs": costs,
                                    "max_relerr_vs_fp64_oracle_48_sampled": rel_c},
                   "svds": {"ms": round(t_s * 1e3, 3), "gflop_nominal": round(fl_s / 1e9, 1),
                            "max_sigma_err_rel_sigma1_16_sampled": rel_s}},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    plan.close()
    ctx.close()


# ---------------------------------------------------------------- serving
SERVE_STATES, SERVE_SEED, SERVE_P99_BOUND_MS = 32, 20261202 + 4, 50.0


def bench_serving(args):
    """32 device-resident MPS states (n uniform 16..128, d = 2, chi
    log-uniform 16..512 capped at 2^min(b, n-b), complex64 Gaussians,
    canonize!(site 1) + normalize!), every left environment prefilled.
    Requests: 50% amplitudes <x|psi> (uniform x), 50% Pauli strings of weight
    1-4 on consecutive sites.  Poisson arrivals at increasing rates; the
    reported rate is the highest whose p99 latency stays within 50 ms."""
    import numpy as np
    import tenet_hip as T
    from tenet_hip import _lib as L
    from tenet_hip.serving import Server, latency_summary
    from tenet_hip.workloads import build_serving_states, serving_requests

    ctx = T.Context(0)
    t0 = time.perf_counter()
    states = build_serving_states(ctx, SERVE_SEED, SERVE_STATES)
    ctx.sync()
    log(f"serving: {len(states)} states built in {time.perf_counter() - t0:.1f}s, "
        f"{sum(s.bytes for s in states) / 1e9:.2f} GB, n {[s.n for s in states]}")
    t0 = time.perf_counter()
    srv = Server(ctx, states, n_streams=4, max_inflight=1024, max_amp_batch=512, prefill_envs=True)
    log(f"serving: {len(srv.env_fixed)} environments prefilled in {time.perf_counter() - t0:.1f}s")
    warm = serving_requests(SERVE_SEED + 1, states, 400)
    srv.run(warm)
    def leg(reqs, arr):
        """one serving run with the amplitude kernels timed; amp_bytes: the
        slice bytes the requests' chains read (one selected slice per
        request and site: the per-request kernels' algorithmic bytes)"""
        ctx.stats_sample_every(1)
        ctx.stats_enable(L.KSTAT_AMPLITUDE)
        srv.run(reqs, arr)
        kst = ctx.stats_read()
        ctx.stats_enable(L.KSTAT_OFF)
        lat = latency_summary(reqs)
        req_bytes = 0.0
        for q in reqs:
            if q.kind == "amp":
                c = states[q.state].chis
                req_bytes += 8.0 * float(np.sum(c[:-1] * c[1:]))
        return {"requests": len(reqs), "served_rps": round(len(reqs) / max(q.t_done for q in reqs), 1),
                **{k: round(v, 3) for k, v in lat.items()}, "amp_kernel_ms": round(kst["ms"], 2),
                "amp_launches": kst["launches"], "amp_bytes": req_bytes}

    sweep = []
    best = None
    rates = [float(x) for x in args.rates.split(",")]
    for i, rate in enumerate(rates):
        nreq = int(min(max(rate * args.serve_seconds, 200), 40000))
        reqs, arr = serving_requests(SERVE_SEED + 10 + i, states, nreq, rate=rate)
        ent = {"offered_rps": rate, **leg(reqs, arr)}
        sweep.append(ent)
        log(f"serving rate {rate}: {ent}")
        if ent["p99_ms"] <= SERVE_P99_BOUND_MS:
            best = ent
        else:
            break
    # saturated amplitude leg: every request present at t = 0, so the batches
    # are full (max_amp_batch) and the kernels run at the state set's size
    sat_reqs = serving_requests(SERVE_SEED + 77, states, args.sat_requests, amp_fraction=1.0)
    sat = leg(sat_reqs, None)
    log(f"serving saturated amplitudes: {sat}")
    ctx.stats_sample_every(17)
    stats = dict(srv.stats)
    srv.close()
    r = best if best is not None else sweep[0]
    roof = None
    if sat["amp_launches"] and sat["amp_kernel_ms"] > 0:
        ach = sat["amp_bytes"] / (sat["amp_kernel_ms"] * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 2), "peak": 8000.0, "unit": "GB/s", "frac": round(ach / 8000, 5),
                "traffic": None, "kernel": "amp_site_kernel (site-parallel amplitudes: one launch per site position, each "
                                           "request's slice product over ceil(chi/16) workgroups)",
                "measured_on": f"saturated leg: {sat['requests']} amplitude requests present at t = 0 "
                               f"(batches of up to 512), {sat['served_rps']} req/s",
                "achieved_basis": "one selected slice per request and site (c64) / amplitude kernel time; a batch "
                                  "of 512 requests over 32 states reads each slice ~8 times, the repeats from "
                                  "L2 / Infinity Cache",
                "launches": sat["amp_launches"], "avg_launch_us": round(sat["amp_kernel_ms"] * 1e3 / sat["amp_launches"], 1)}
    cpu = None
    if not args.no_cpu:
        from oracle import tenet_oracle as O

        hosts = [O.OMPS([a.to_host().astype(np.complex64) for a in st.mps.sites]) for st in states[:8]]
        reqs = serving_requests(SERVE_SEED + 99, states[:8], 400)
        t0 = time.perf_counter()
        done = 0
        for q in reqs:
            if q.kind == "amp":
                O.amplitude(q.bits, hosts[q.state])
            else:
                O.expectation_pauli(hosts[q.state], q.start, q.paulis)
            done += 1
            if time.perf_counter() - t0 > min(args.cpu_seconds, 15.0):
                break
        dt = time.perf_counter() - t0
        threads = os.environ.get("OPENBLAS_NUM_THREADS") or str(os.cpu_count())
        cpu = {"value": round(done / dt, 2), "unit": "req/s", "cores": int(threads), "kind": "port",
               "sample": f"oracle amplitude / full-sweep expectation on {done} requests over the first 8 states, "
                         f"one at a time ({dt:.1f} s, numpy/OpenBLAS, {threads} threads)"}
    line = {
        "metric": "serving req/s at p99 <= 50 ms (amplitudes + Pauli expectations, continuous batching)",
        "value": r["served_rps"], "unit": "req/s", "n_gpus": 1, "steps": len(sweep), "warmup": 1,
        "ms_per_step": round(1e3 / max(r["served_rps"], 1e-9), 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64", "data": "synthetic (device-generated seeded Gaussian MPS)",
        "config": {"workload": "serving 32 MPS states n 16..128 chi 16..512 c64; 50% amplitude / 50% Pauli weight 1-4;"
                               " Poisson arrivals", "p99_bound_ms": SERVE_P99_BOUND_MS, "p50_ms": r["p50_ms"],
                   "p99_ms": r["p99_ms"], "sweep": sweep, "saturated_amplitudes": sat, "server_stats": stats,
                   "state_gb": round(sum(s.bytes for s in states) / 1e9, 3)},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    for st in states:
        st.close()
    ctx.close()


# ---------------------------------------------------------------- TEBD (default)
def layers_of(sched, n):
    """Split the serpentine brickwork schedule into its layers (order kept,
    so gate g still acts on bond sched[g])."""
    out, i, parity = [], 0, 0
    while i < len(sched):
        cnt = len(range(parity, n - 1, 2))
        out.append(list(sched[i:i + cnt]))
        i += cnt
        parity ^= 1
    return out


def bench_tebd(args):
    import tenet_hip as T
    from tenet_hip.tebd import brickwork_schedule, haar_gates, upload_gates

    fx = {"plain": np.load(FIXTURE), "vidal": np.load(VIDAL_FIXTURE)}
    bits = np.unpackbits(fx["plain"]["bits"], axis=1)[:, :N_SITES]
    assert np.array_equal(fx["plain"]["bits"], fx["vidal"]["bits"])
    sched = brickwork_schedule(N_SITES, DEPTH)
    gates = haar_gates(SEED, len(sched))
    layers = layers_of(list(sched), N_SITES)
    forms = [f for f in args.forms.split(",") if f]
    head = forms[0]
    ctx = T.Context(0)
    gate_views = upload_gates(ctx, gates)
    results = {}
    for form in forms:
        steps = args.steps if form == head else min(args.steps, args.cmp_steps)
        warm = args.warmup if form == head else min(args.warmup, 1)
        log(f"form {form}: {warm} warmup + {steps} timed sweeps")
        r = run_form(ctx, T, form, sched, layers, gate_views, steps, warm, bits)
        ref = fx["vidal" if form in VIDAL_FORMS else "plain"]
        r["relerr"] = float(np.linalg.norm(r["amps"] - ref["amps"]) / np.linalg.norm(ref["amps"]))
        r["bonds_match"] = bool(list(r["bonds"]) == list(ref["bonds"]))
        if form in VIDAL_FORMS:  # distance to the MixedCanonical oracle (a different truncation order)
            pa = fx["plain"]["amps"]
            r["dev_vs_mixedcanonical_oracle"] = float(np.linalg.norm(r["amps"] - pa) / np.linalg.norm(pa))
        results[form] = r
        log(f"form {form}: {r['ms_per_step']:.1f} ms/sweep, relerr {r['relerr']:.2e}, bonds match "
            f"{r['bonds_match']}, {r['flops'] / 1e9:.1f} GFLOP nominal")
    r = results[head]
    base = results.get("fp64")
    # the round kernel: nominal SVD work of the sweep over the kernel's time
    k = r["kstats"]
    roof = None
    if k["launches"] > 0 and k["ms"] > 0:
        svd_nominal = r["flops_split"]["svd"] * r["steps"]
        ach = svd_nominal / (k["ms"] * 1e-3) / 1e12
        executed = k["flops"] / (k["ms"] * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": round(ach, 4), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / PEAK_FP64_TFLOPS, 5), "traffic": ROUND_TRAFFIC_BYTES.get(head),
                "kernel": ("svd_round_mfma_kernel (grouped one-sided Jacobi rounds: every SVD of a brickwork layer "
                           "per launch; Gram and rotation on v_mfma_f64_16x16x4, inner 32x32 solve on FP64 VALU)"
                           if head in VIDAL_FORMS else
                           "svd_round_kernel (one-sided Jacobi round of one SVD: pair Grams, rotations, apply; "
                           "FP64 VALU)"),
                "achieved_basis": "nominal SVD flops 4(4p^2q+8pq^2+9q^3) of the sweep / round-kernel time",
                "peak_basis": "FP64 (vector and matrix peaks are both 78.6 TFLOP/s on MI355X)",
                "executed_tflops": round(executed, 3), "executed_frac": round(executed / PEAK_FP64_TFLOPS, 4),
                "launches": k["launches"], "avg_launch_us": round(k["ms"] * 1e3 / k["launches"], 2),
                "kernel_share_of_step": round(k["ms"] / (r["ms_per_step"] * r["steps"]), 3),
                "algorithmic_flops_per_launch": round(svd_nominal / k["launches"])}
    cpu = None
    if not args.no_cpu:
        cform = "vidal" if head in VIDAL_FORMS else "fp64"
        res, model, threads = cpu_baseline_tebd(cform, args.cpu_seconds)
        est_sweep = r["flops"] / (res["all"]["gflops"] * 1e9)
        cpu = {"value": round(res["all"]["gflops"], 3), "unit": "GFLOP/s", "cores": threads, "kind": "port",
               "sample": (f"oracle {'Vidal-gauge update' if cform == 'vidal' else 'canonize! move + simple_update'} "
                          f"at the sweep's saturated shape (chi 256, 512x512 SVD), median of 5 blocks of "
                          f"{res['all']['gates_per_block']} gates after a warm-up; numpy/OpenBLAS LAPACK on "
                          f"{threads} threads ({model}); at this rate the full sweep's nominal work takes "
                          f"{est_sweep:.1f} s"),
               "one_thread": {"value": round(res["one"]["gflops"], 3), "unit": "GFLOP/s",
                              "gates_per_block": res["one"]["gates_per_block"]},
               "cpu_model": model, "host_cpus": os.cpu_count()}
    line = {
        "metric": "TEBD sweep time & GFLOP/s split-prec vs FP64 @relerr<=1e-8; req/s + p99 latency",
        "value": round(r["flops"] / (r["ms_per_step"] * 1e-3) / 1e9, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": r["steps"],
        "warmup": r["warmup"],
        "ms_per_step": round(r["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic (seeded Haar gates, |0...0>)",
        "config": {
            "workload": "TEBD brickwork n=64 d=2 depth=20 chi_max=256 threshold=1e-10 (630 gates)",
            "form": head,
            "relerr_vs_fp64_oracle": r["relerr"],
            "bonds_match_oracle": r["bonds_match"],
            "speedup_vs_plain_fp64": round(base["ms_per_step"] / r["ms_per_step"], 3) if base else None,
            "forms": {f: {"ms_per_sweep": round(v["ms_per_step"], 3), "timed_sweeps": v["steps"],
                          "relerr_vs_own_fp64_oracle": v["relerr"], "bonds_match_oracle": v["bonds_match"],
                          "gflop_nominal": round(v["flops"] / 1e9, 2),
                          "round_kernel_share": (round(v["kstats"]["ms"] / (v["ms_per_step"] * v["steps"]), 3)
                                                 if v["kstats"]["launches"] else None),
                          **({"dev_vs_mixedcanonical_oracle": v["dev_vs_mixedcanonical_oracle"]}
                             if "dev_vs_mixedcanonical_oracle" in v else {})}
                      for f, v in results.items()},
            "oracle": {"vidal": "oracle Vidal-gauge sweep (tests/golden/tebd_vidal_n64_d20_chi256.npz)",
                       "plain forms": "oracle MixedCanonical sweep (tests/golden/tebd_n64_d20_chi256.npz)"},
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


# ---------------------------------------------------------------- traffic calibration
def bench_copy(args):
    """x *= alpha over --copy-mib MiB of complex128, five times (tenet_scale:
    every element read once and written once).  Its FETCH_SIZE / WRITE_SIZE
    counters per dispatch against the known bytes calibrate the traffic
    figures of the other kernels."""
    import tenet_hip as T

    ctx = T.Context(0)
    n = args.copy_mib * (1 << 20) // 16
    a = T.DeviceArray(ctx, (n,)).gaussian_(1)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(5):
        a.scale_(1.0)
    ctx.sync()
    dt = (time.perf_counter() - t0) / 5
    print(json.dumps({"metric": "scale kernel traffic calibration", "read_bytes": a.nbytes, "write_bytes": a.nbytes,
                      "ms": round(dt * 1e3, 3), "gbps": round(2 * a.nbytes / dt / 1e9, 1)}), flush=True)
    ctx.close()


# ---------------------------------------------------------------- plain small
PLAIN_N, PLAIN_CHI, PLAIN_SEED = 20, 32, 20261202 + 0


def plain_arrays():
    """n = 20, d = 2, uniform chi = 32 (edges included: site 1 is (1, 2, 32)
    as a 3-way array), standard complex Gaussian entries from numpy PCG64."""
    from tenet_hip.workloads import plain_state_arrays

    return plain_state_arrays(PLAIN_SEED, PLAIN_N, PLAIN_CHI)


def bench_plain(args):
    """canonize!(psi, site 1) from MixedCanonical(all sites) (19 QR moves,
    canonize.jl:128-149), then <psi|psi> by the transfer-matrix sweep
    (overlap.jl:36-50) and norm(psi) (norm.jl:17-37).  A step is that
    sequence on a fresh copy of the state (copies prepared untimed);
    latency-bound, reported per step."""
    import numpy as np
    import tenet_hip as T
    from tenet_hip import _lib as L
    from tenet_hip.mps import MPS, qr_flops

    arrs = plain_arrays()
    ctx = T.Context(0)
    n, chi = PLAIN_N, PLAIN_CHI
    # nominal work: 19 QR moves (QR of (chi_l d) x chi, R . next site), overlap sweep
    fl_qr = 0.0
    for i in range(n - 1, 0, -1):
        chil, d, chir = arrs[i].shape
        fl_qr += qr_flops(d * chir, chil) + 8.0 * arrs[i - 1].shape[0] * 2 * chil * min(d * chir, chil)
    fl_ov = 0.0
    for a in arrs:
        chil, d, chir = a.shape
        fl_ov += 8.0 * chil * chil * d * chir + 8.0 * chil * d * chir * chir
    states = [MPS.from_host(ctx, arrs) for _ in range(args.warmup + args.steps)]
    ctx.sync()
    vals = []

    def one(psi):
        psi.canonize_(0)
        v = psi.overlap(psi)
        nr = psi.norm()
        return v, nr

    ctx.stats_sample_every(1)
    ctx.stats_enable(L.KSTAT_QR)  # during the warmup too: its timing events are created there
    for w in range(args.warmup):
        one(states[w])
    ctx.sync()
    ctx.stats_enable(L.KSTAT_QR)  # counters restart for the timed steps
    t0 = time.perf_counter()
    for s in range(args.steps):
        vals.append(one(states[args.warmup + s]))
    wall = (time.perf_counter() - t0) / args.steps
    kst = ctx.stats_read()
    ctx.stats_enable(L.KSTAT_OFF)
    ctx.stats_sample_every(17)
    ov, nr = vals[-1]

    # the same step as one captured graph (canonize.jl:136-145 sweep, overlap,
    # norm): every replay re-reads the source state and rewrites the same
    # output buffers; its values equal the eager step's bit for bit
    from tenet_hip.graphs import CanonizeNormGraph

    src = MPS.from_host(ctx, arrs)
    g = CanonizeNormGraph(src, 0)
    for _ in range(max(1, args.warmup)):
        g.run()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gv = g.run()
    wall_graph = (time.perf_counter() - t0) / args.steps
    graph_equal = gv == (ov, nr)
    g.close()
    log(f"plain: eager {wall * 1e6:.1f} us/step, graph {wall_graph * 1e6:.1f} us/step, equal {graph_equal}")

    rel = None
    cpu = None
    if not args.no_cpu:
        from oracle import tenet_oracle as O

        ref = O.overlap_mps(O.OMPS(arrs), O.OMPS(arrs))
        rel = abs(ov - ref) / abs(ref)
        reps = 0
        t0 = time.perf_counter()
        while True:
            p = O.OMPS([a.copy() for a in arrs])
            O.canonize_mixed(p, 0, 0)
            O.overlap_mps(p, p)
            O.norm(p)
            reps += 1
            if time.perf_counter() - t0 > min(args.cpu_seconds, 10.0):
                break
        dt = (time.perf_counter() - t0) / reps
        threads = os.environ.get("OPENBLAS_NUM_THREADS") or str(os.cpu_count())
        cpu = {"value": round(dt * 1e6, 1), "unit": "us/step", "cores": int(threads), "kind": "port",
               "sample": f"oracle canonize!+overlap+norm on the same state, {reps} repetitions (numpy/OpenBLAS, "
                         f"{threads} threads)"}
    roof = None
    if kst["launches"]:
        ach = kst["flops"] / (kst["ms"] * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": round(ach, 5), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / PEAK_FP64_TFLOPS, 6), "traffic": None,
                "kernel": "qr_reg_kernel (one canonize move per launch: 64x32 / 2x32 Householder QR of A^H with "
                          "R^H absorbed into the neighbour, one workgroup)",
                "launches": kst["launches"], "avg_launch_us": round(kst["ms"] * 1e3 / kst["launches"], 2),
                "note": "latency-bound: 19 dependent moves, 32 barrier-separated reflector steps each"}
    best = min(wall, wall_graph)
    line = {
        "metric": "plain canonize!+norm latency (n=20, chi=32)", "value": round(best * 1e6, 1), "unit": "us/step",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(best * 1e3, 4),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "c128",
        "data": "synthetic (seeded Gaussian MPS)",
        "config": {"workload": "open MPS n=20 d=2 chi=32 Gaussian: canonize!(site 1) + overlap + norm",
                   "gflop_nominal": round((fl_qr + fl_ov) / 1e9, 4), "overlap": [ov.real, ov.imag], "norm": nr,
                   "overlap_relerr_vs_fp64_oracle": rel,
                   "forms": {"eager": {"us_per_step": round(wall * 1e6, 1)},
                             "graph": {"us_per_step": round(wall_graph * 1e6, 1),
                                       "bit_identical_to_eager": bool(graph_equal)}}},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps of the headline form (default 5; 20 plain)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (default 1; 3 for plain)")
    ap.add_argument("--forms", default=None,
                    help=f"TEBD forms, headline first (default {DEFAULT_FORMS}); heff: fp64,fp32,split_bf16,split_fp16")
    ap.add_argument("--cmp-steps", type=int, default=2, help="timed sweeps of each comparison TEBD form")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-mib", type=int, default=1024, help="traffic calibration buffer (workload copy)")
    ap.add_argument("--workload", default="tebd", choices=["tebd", "heff", "plain", "grouped", "serving", "copy"])
    ap.add_argument("--rates", default="250,500,1000,2000,4000,8000,16000")
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--sat-requests", type=int, default=4096, help="amplitude requests of the saturated serving leg")
    args = ap.parse_args()
    latency = args.workload == "plain"  # a ~1 ms step: more samples
    if args.steps is None:
        args.steps = 20 if latency else 5
    if args.warmup is None:
        args.warmup = 3 if latency else 1
    if args.gpus != 1:
        print("bench.py measures one MI355X and runs no multi-GPU path (every GPU box here has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "heff":
        args.forms = args.forms or "fp64,fp32,split_bf16,split_fp16,split_i8,split_i8_fast"
        return bench_heff(args)
    args.forms = args.forms or DEFAULT_FORMS
    if args.workload == "plain":
        return bench_plain(args)
    if args.workload == "grouped":
        return bench_grouped(args)
    if args.workload == "serving":
        return bench_serving(args)
    if args.workload == "copy":
        return bench_copy(args)
    return bench_tebd(args)


if __name__ == "__main__":
    main()
