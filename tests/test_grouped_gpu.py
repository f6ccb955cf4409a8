"""Grouped launches against the oracle: many uneven contractions in one
persistent launch (every output against a per-problem binary_einsum of the
oracle), K-chunked tiles (deterministic chunk-order sums, repeated launches
of one plan), the tile autotuner, and grouped truncated SVDs.

Tolerances: complex64 problems within 2e-5 relative (norm-wise, FP32 MFMA
accumulation over K <= ~1e5 terms); complex128 within 1e-12.
"""
import numpy as np
import pytest

from oracle import tenet_oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def _run_grouped(ctx, specs, dtype, seed):
    import tenet_hip as T
    from tenet_hip.grouped import GroupedPlan

    rng = np.random.default_rng(seed)
    probs, hosts, outs, ptrs = [], [], [], []
    for sp in specs:
        a = O.complex_gaussian(rng, sp.shape(sp.la), dtype)
        b = O.complex_gaussian(rng, sp.shape(sp.lb), dtype)
        da, db = T.DeviceArray.from_host(ctx, a, dtype), T.DeviceArray.from_host(ctx, b, dtype)
        dc = T.DeviceArray(ctx, sp.shape(sp.lc), dtype)
        probs.append((sp.shape(sp.la), sp.la, sp.shape(sp.lb), sp.lb, sp.shape(sp.lc), sp.lc))
        hosts.append((a, b))
        outs.append((da, db, dc))
        ptrs += [da.ptr, db.ptr, dc.ptr]
    plan = GroupedPlan(ctx, probs, dtype)
    return plan, hosts, outs, ptrs


def _check(specs, hosts, outs, tol):
    for sp, (a, b), (_, _, dc) in zip(specs, hosts, outs):
        dims = [l for l in sp.la if l in sp.lb and l not in sp.lc]
        ref = O.binary_einsum(O.OTensor(a.astype(np.complex128), sp.la), O.OTensor(b.astype(np.complex128), sp.lb),
                              dims)
        ref = ref.permute(sp.lc).data
        assert _rel(dc.to_host(), ref) < tol, (sp.la, sp.lb, sp.extent)


@pytest.mark.parametrize("dtype,tol", [(np.complex64, 2e-5), (np.complex128, 1e-12)])
def test_grouped_random_problems(ctx, dtype, tol):
    """300 problems from the grouped workload generator with a small size
    cap: ranks 3-5, 1-3 contracted labels, extents log-uniform 2..128."""
    from tenet_hip.workloads import grouped_contractions

    specs = grouped_contractions(11, 300, cap=1 << 12)
    plan, hosts, outs, ptrs = _run_grouped(ctx, specs, dtype, 5)
    info = plan.info()
    assert sum(info["tiles_64_32_16"]) + info["smallk"] == len(specs)
    plan.exec(ptrs)
    _check(specs, hosts, outs, tol)
    plan.close()


def test_grouped_k_chunks_deterministic(ctx):
    """A few long-K problems among many small ones force K chunking; the
    chunk-order reduction makes repeated launches of one plan bit-identical."""
    from tenet_hip.workloads import ContractionSpec, grouped_contractions

    specs = grouped_contractions(12, 400, cap=1 << 12)
    for j in range(6):
        specs.append(ContractionSpec(["m", "k1", "k2"], ["k1", "k2", "n"], ["m", "n"],
                                     {"m": 3 + 17 * j, "n": 5 + 9 * j, "k1": 97, "k2": 523}))
    plan, hosts, outs, ptrs = _run_grouped(ctx, specs, np.complex64, 6)
    info = plan.info()
    assert info["split_tiles"] > 0, info
    plan.exec(ptrs)
    first = [o[2].to_host() for o in outs]
    _check(specs, hosts, outs, 2e-5)
    for o in outs:
        o[2].zero_()
    plan.exec(ptrs)
    for f, o in zip(first, outs):
        assert np.array_equal(f, o[2].to_host())
    plan.close()


def test_grouped_batch_labels_and_empty(ctx):
    """Batch (Hadamard) labels kept in the output, an output with permuted
    labels, and a contraction over an empty extent (zero result)."""
    from tenet_hip.workloads import ContractionSpec

    specs = [ContractionSpec(["i", "k", "z"], ["k", "j", "z"], ["i", "j", "z"], {"i": 20, "j": 33, "k": 40, "z": 7}),
             ContractionSpec(["i", "k"], ["k", "j"], ["j", "i"], {"i": 70, "j": 3, "k": 1000}),
             ContractionSpec(["i", "k"], ["k", "j"], ["i", "j"], {"i": 4, "j": 5, "k": 0})]
    plan, hosts, outs, ptrs = _run_grouped(ctx, specs, np.complex128, 7)
    plan.exec(ptrs)
    _check(specs, hosts, outs, 1e-12)
    assert np.all(outs[2][2].to_host() == 0)
    plan.close()


def test_autotune_then_plan(ctx):
    from tenet_hip.grouped import autotune_tiles
    from tenet_hip.workloads import grouped_contractions

    costs = autotune_tiles(ctx, np.complex64)
    assert len(costs) == 3 and costs[0] == pytest.approx(1.0) and all(c > 0 for c in costs)
    specs = grouped_contractions(13, 100, cap=1 << 12)
    plan, hosts, outs, ptrs = _run_grouped(ctx, specs, np.complex64, 8)
    plan.exec(ptrs)
    _check(specs, hosts, outs, 2e-5)
    plan.close()


def test_grouped_svd_vs_lapack(ctx):
    """Grouped truncated SVDs: singular values against LAPACK, the kept
    count from the truncation rule (maxdim = ceil(min(m, n) / 2)), the
    discarded weight, and U diag(S) Vh against the best rank-k product."""
    import tenet_hip as T
    from tenet_hip.grouped import grouped_svd

    rng = np.random.default_rng(9)
    shapes = [(16, 16), (40, 17), (17, 40), (100, 64), (64, 100), (33, 33), (200, 20), (20, 200)]
    hosts = [O.complex_gaussian(rng, s) for s in shapes]
    mats = [T.DeviceArray.from_host(ctx, h) for h in hosts]
    maxdims = [(min(s) + 1) // 2 for s in shapes]
    res = grouped_svd(ctx, mats, maxdims, 1e-10)
    for h, md, (U, S, Vh, k, disc) in zip(hosts, maxdims, res):
        s_ref = np.linalg.svd(h, compute_uv=False)
        assert k == md
        assert np.allclose(S[:k], s_ref[:k], rtol=1e-12, atol=1e-12 * s_ref[0])
        assert disc == pytest.approx(np.sum(s_ref[k:] ** 2), rel=1e-10)
        Uh, Vhh = U.to_host()[:, :k], Vh.to_host()[:k, :]
        u, s, vh = np.linalg.svd(h, full_matrices=False)
        best = (u[:, :k] * s[:k]) @ vh[:k]
        assert _rel((Uh * S[:k]) @ Vhh, best) < 1e-11


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_grouped_svd_every_small_width(ctx, dtype):
    """Every width 1..33 in one grouped batch: widths up to 32 are one column
    pair of 2b = 2, 4, ..., 32 columns solved by the full inner sweep alone
    (33 takes the block rounds).  Singular values against LAPACK and
    U diag(S) Vh against the input, no truncation."""
    import tenet_hip as T
    from tenet_hip.grouped import grouped_svd

    rng = np.random.default_rng(21)
    shapes = [(n + 3, n) for n in range(1, 34)] + [(n, n + 5) for n in (2, 7, 16, 31)]
    hosts = [O.complex_gaussian(rng, s).astype(dtype) for s in shapes]
    mats = [T.DeviceArray.from_host(ctx, h) for h in hosts]
    res = grouped_svd(ctx, mats, None, 0.0)
    tol = 1e-12 if dtype == np.complex128 else 2e-5
    for h, (U, S, Vh, k, disc) in zip(hosts, res):
        h64 = h.astype(np.complex128)
        s_ref = np.linalg.svd(h64, compute_uv=False)
        assert k == min(h.shape)
        assert np.allclose(S[:k], s_ref, rtol=tol, atol=tol * s_ref[0])
        Uh, Vhh = U.to_host()[:, :k].astype(np.complex128), Vh.to_host()[:k, :].astype(np.complex128)
        assert _rel((Uh * S[:k]) @ Vhh, h64) < 10 * tol
        assert _rel(Uh.conj().T @ Uh, np.eye(k)) < 10 * tol


@pytest.mark.parametrize("want", [(True, True, True), (True, True, False), (False, False, False)],
                         ids=["all_outputs", "no_disc", "no_s_k_disc"])
def test_grouped_svd_graded_problems_converge(ctx, want):
    """A batch mixing Gaussian matrices with strongly graded ones (sigma over
    12 decades, random singular vectors), which the plain sweeps do not
    finish within the sweep limit: those problems are redone one at a time
    by the QR-preconditioned single-matrix path, and every result matches
    LAPACK under the truncation rule (maxdim = min(m, n), threshold 1e-10).
    The graded problems sit at indices > 0 and the optional outputs (S, k,
    discarded weight) are also passed as NULL, as the Vidal-gauge layer
    passes the discarded weight: the fallback must honour the NULLs."""
    import tenet_hip as T
    from tenet_hip.grouped import grouped_svd

    want_s, want_k, want_disc = want
    rng = np.random.default_rng(12)

    def graded(m, n):
        k = min(m, n)
        qa, _ = np.linalg.qr(O.complex_gaussian(rng, (m, k)))
        qb, _ = np.linalg.qr(O.complex_gaussian(rng, (n, k)))
        return (qa * np.logspace(0, -12, k)) @ qb.conj().T

    hosts = [O.complex_gaussian(rng, (100, 64)), graded(600, 300), O.complex_gaussian(rng, (40, 17)),
             graded(200, 320)]
    mats = [T.DeviceArray.from_host(ctx, h) for h in hosts]
    maxdims = [min(h.shape) for h in hosts]
    res = grouped_svd(ctx, mats, maxdims, 1e-10, want_s=want_s, want_k=want_k, want_disc=want_disc)
    for h, (U, S, Vh, k, disc) in zip(hosts, res):
        u, s_ref, vh = np.linalg.svd(h, full_matrices=False)
        kref = int(np.sum(s_ref > 1e-10))
        if want_k:
            assert k == kref
        if want_disc:
            assert disc == pytest.approx(np.sum(s_ref[kref:] ** 2), rel=1e-8, abs=1e-20)
        Vk = Vh.to_host()[:kref, :]
        best = (u[:, :kref] * s_ref[:kref]) @ vh[:kref]
        # A projected on the kept right singular subspace is the best rank-k product
        assert _rel(h @ Vk.conj().T @ Vk, best) < 1e-11
        if want_s:
            assert np.allclose(S[:kref], s_ref[:kref], rtol=1e-12, atol=1e-13 * s_ref[0])
            assert _rel((U.to_host()[:, :kref] * S[:kref]) @ Vk, best) < 1e-11


@pytest.mark.parametrize("dtype,tol", [(np.complex64, 2e-6), (np.complex128, 1e-14)])
def test_grouped_small_k_elementwise_units(ctx, dtype, tol):
    """Contractions with K <= 8 and a thin output run as element-wise units
    (no MFMA tile): the Vidal layer's Lambda scaling (dims = [] with the
    label kept, a batch of 1 x N products), its gate application (K = 4
    over the physical pair), a strided copy through a scalar, and a
    conjugated K = 2 product; each against the oracle's binary_einsum."""
    import tenet_hip as T
    from tenet_hip.grouped import GroupedPlan

    rng = np.random.default_rng(21)
    l, r = 40, 36
    cases = [
        # (shape_a, labels_a, shape_b, labels_b, shape_c, labels_c, strides_a, conj)
        ((l,), "l", (l, 2, 2, r), "lpqr", (l, 2, 2, r), "lpqr", None, (0, 0)),
        ((l, 2, 2, r), "lxyr", (2, 2, 2, 2), "pqxy", (l, 2, 2, r), "lpqr", None, (0, 0)),
        ((5, 2, r), "kqr", (), "", (5, 2, r), "kqr", (1, 9, 18), (0, 0)),
        ((3, 2, 7), "akb", (2, 5), "kc", (3, 7, 5), "abc", None, (1, 1)),
    ]
    probs, hosts, outs, ptrs, conj = [], [], [], [], []
    for sa, la, sb, lb, sc, lc, st_a, cj in cases:
        la, lb, lc = list(la), list(lb), list(lc)
        if st_a is not None:  # a view of the first 5 rows of a 9-row buffer
            full = O.complex_gaussian(rng, (9, 2, r), dtype)
            a = full[:5]
            da = T.DeviceArray.from_host(ctx, full, dtype)
        else:
            a = O.complex_gaussian(rng, sa, dtype)
            da = T.DeviceArray.from_host(ctx, a, dtype)
        b = O.complex_gaussian(rng, sb, dtype)
        db = T.DeviceArray.from_host(ctx, b, dtype)
        dc = T.DeviceArray(ctx, sc, dtype)
        pr = (sa, la, sb, lb, sc, lc)
        if st_a is not None:
            pr = pr + (st_a, None)
        probs.append(pr)
        hosts.append((a, b, la, lb, lc, cj))
        outs.append((da, db, dc))
        ptrs += [da.ptr, db.ptr, dc.ptr]
        conj.append(cj)
    plan = GroupedPlan(ctx, probs, dtype, conj=conj)
    try:
        assert plan.info()["smallk"] == len(cases)
        for rep in range(2):
            plan.exec(ptrs)
            for (a, b, la, lb, lc, cj), (_, _, dc) in zip(hosts, outs):
                a2 = np.conj(a) if cj[0] else a
                b2 = np.conj(b) if cj[1] else b
                dims = [x for x in la if x in lb and x not in lc]
                ref = O.binary_einsum(O.OTensor(np.asarray(a2, np.complex128), la),
                                      O.OTensor(np.asarray(b2, np.complex128), lb), dims).permute(lc).data
                assert _rel(dc.to_host(), ref) < tol, (la, lb, lc)
    finally:
        plan.close()


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-11), (np.complex64, 2e-5)])
def test_grouped_svd_without_u_preconditioned(ctx, dtype, tol):
    """Grouped SVDs with no U output: the problems with 64 <= n <= m <= 512
    take the QR-preconditioned Jacobi (columns sorted by norm, Householder
    QR, Jacobi on R^H, V^H written through the column order); the others
    (wide, m > 512, n < 64) the plain one.  Gaussian, graded (sigma over 12
    decades), exactly rank-deficient and Vidal-like (rows and columns
    scaled by decaying Schmidt values) matrices, with a maxdim cut on two.
    Against LAPACK on the same (rounded) input: the kept count, S, V^H's
    kept rows orthonormal, and A projected on them equal to the best rank-k
    product."""
    import tenet_hip as T
    from tenet_hip.grouped import grouped_svd

    rng = np.random.default_rng(31)

    def graded(m, n, decades=12):
        k = min(m, n)
        qa, _ = np.linalg.qr(O.complex_gaussian(rng, (m, k)))
        qb, _ = np.linalg.qr(O.complex_gaussian(rng, (n, k)))
        return (qa * np.logspace(0, -decades, k)) @ qb.conj().T

    def lowrank(m, n, r):
        return O.complex_gaussian(rng, (m, r)) @ O.complex_gaussian(rng, (r, n))

    def vidal_like(chi):
        lam_l = np.sort(rng.exponential(size=chi))[::-1] ** 3
        lam_r = np.sort(rng.exponential(size=chi))[::-1] ** 3
        th = O.complex_gaussian(rng, (chi, 2, 2, chi)) * lam_r[None, None, None, :]
        return (lam_l[:, None, None, None] * th).reshape(2 * chi, 2 * chi)

    hosts = [O.complex_gaussian(rng, (512, 512)), graded(512, 256), O.complex_gaussian(rng, (300, 64)),
             O.complex_gaussian(rng, (64, 100)), O.complex_gaussian(rng, (600, 100)), O.complex_gaussian(rng, (40, 17)),
             lowrank(130, 130, 40), vidal_like(256), vidal_like(96)]
    hosts = [h / np.linalg.norm(h) for h in hosts]
    hosts = [h.astype(dtype).astype(np.complex128) for h in hosts]
    maxdims = [min(h.shape) for h in hosts]
    maxdims[0], maxdims[7] = 256, 200
    mats = [T.DeviceArray.from_host(ctx, h.astype(dtype)) for h in hosts]
    res = grouped_svd(ctx, mats, maxdims, 1e-10, want_u=False)
    for i, (h, md, (U, S, Vh, k, disc)) in enumerate(zip(hosts, maxdims, res)):
        assert U is None
        u, s_ref, vh = np.linalg.svd(h, full_matrices=False)
        kref = min(md, int(np.sum(s_ref > 1e-10)))
        assert k == kref, i
        assert np.allclose(S[:k], s_ref[:k], rtol=1e-12, atol=1e-13 * s_ref[0]), i
        assert disc == pytest.approx(np.sum(s_ref[k:] ** 2), rel=1e-8, abs=1e-20), i
        Vk = Vh.to_host()[:k, :].astype(np.complex128)
        assert np.abs(Vk @ Vk.conj().T - np.eye(k)).max() < tol, i
        best = (u[:, :k] * s_ref[:k]) @ vh[:k]
        assert _rel(h @ Vk.conj().T @ Vk, best) < tol, i


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 2e-5)])
def test_grouped_plain_gemm_groups(ctx, dtype, tol):
    """Problems whose operands are plain column-major matrices run as batched
    library GEMMs from tenet_grouped_exec (grouped by shape and operand op):
    A.B, A^T.B, A.B^T, conj(A)^T.B, A.conj(B)^T (a strided view with ld
    above the rows, as the Vidal output step reads V^H[:k]), next to a
    permuted (gathered) problem and a small one that stay in the persistent
    kernel.  Each against the oracle's binary_einsum, and the device-table
    launch (every problem in the kernel) against the same reference."""
    import ctypes as C

    import tenet_hip as T
    from tenet_hip import _lib as L
    from tenet_hip.grouped import GroupedPlan

    rng = np.random.default_rng(33)
    m, n, k = 256, 192, 160  # 63 MFLOP per problem: two make a library group
    # (shape_a, labels_a, shape_b, labels_b, shape_c, labels_c, strides_b, conj, copies)
    cases = [
        ((m, k), "ik", (k, n), "kj", (m, n), "ij", None, (0, 0), 3),
        ((k, m), "ki", (k, n), "kj", (m, n), "ij", None, (0, 0), 2),
        ((m, k), "ik", (n, k), "jk", (m, n), "ij", None, (0, 0), 2),
        ((k, m), "ki", (k, n), "kj", (m, n), "ij", None, (1, 0), 2),
        ((m, 2, k // 2), "iqr", (n, 2, k // 2), "jqr", (m, n), "ij", (1, n + 7, 2 * (n + 7)), (0, 1), 2),
        ((k, m), "ki", (n, k), "jk", (n, m), "ji", None, (0, 0), 2),  # C transposed: kernel
        ((6, 5), "ik", (5, 7), "kj", (6, 7), "ij", None, (0, 0), 3),  # small: kernel
    ]
    probs, hosts, outs, ptrs, conj = [], [], [], [], []
    for sa, la, sb, lb, sc, lc, st_b, cj, copies in cases:
        for _ in range(copies):
            la_, lb_, lc_ = list(la), list(lb), list(lc)
            a = O.complex_gaussian(rng, sa, dtype)
            da = T.DeviceArray.from_host(ctx, a, dtype)
            if st_b is not None:  # a view of the first n rows of an (n + 7)-row buffer
                full = O.complex_gaussian(rng, (sb[0] + 7,) + tuple(sb[1:]), dtype)
                b = full[:sb[0]]
                db = T.DeviceArray.from_host(ctx, full, dtype)
            else:
                b = O.complex_gaussian(rng, sb, dtype)
                db = T.DeviceArray.from_host(ctx, b, dtype)
            dc = T.DeviceArray(ctx, sc, dtype)
            pr = (sa, la_, sb, lb_, sc, lc_)
            if st_b is not None:
                pr = pr + (None, st_b)
            probs.append(pr)
            hosts.append((a, b, la_, lb_, lc_, cj))
            outs.append((da, db, dc))
            ptrs += [da.ptr, db.ptr, dc.ptr]
            conj.append(cj)
    plan = GroupedPlan(ctx, probs, dtype, conj=conj)

    def check():
        for (a, b, la_, lb_, lc_, cj), (_, _, dc) in zip(hosts, outs):
            a2 = np.conj(a) if cj[0] else a
            b2 = np.conj(b) if cj[1] else b
            dims = [x for x in la_ if x in lb_ and x not in lc_]
            ref = O.binary_einsum(O.OTensor(np.asarray(a2, np.complex128), la_),
                                  O.OTensor(np.asarray(b2, np.complex128), lb_), dims).permute(lc_).data
            assert _rel(dc.to_host(), ref) < tol, (la_, lb_, lc_, cj)

    try:
        for rep in range(2):
            for _, _, dc in outs:
                dc.zero_()
            plan.exec(ptrs)
            check()
        for _, _, dc in outs:
            dc.zero_()
        words = np.zeros(len(ptrs) + len(ptrs) % 2, dtype=np.uint64)
        words[:len(ptrs)] = ptrs
        tab = T.DeviceArray.from_host(ctx, words.view(np.complex128), np.complex128)
        plan.exec_dev(tab.ptr)
        check()
    finally:
        plan.close()


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 2e-5)])
def test_grouped_plain_gemm_strided(ctx, dtype, tol):
    """A library group whose operands sit evenly spaced in one pool block each
    (as a Vidal layer carves them) runs as one strided-batched GEMM: A.B and
    A.conj(B)^T through a strided view, against the oracle."""
    import ctypes as C

    import tenet_hip as T
    from tenet_hip import _lib as L
    from tenet_hip.device import carve
    from tenet_hip.grouped import GroupedPlan

    rng = np.random.default_rng(34)
    m, n, k, P = 256, 192, 160, 4
    for cj, sb, lb, st_b in [((0, 0), (k, n), "kj", None), ((0, 1), (n, k), "jk", (1, n + 5))]:
        a = [O.complex_gaussian(rng, (m, k), dtype) for _ in range(P)]
        bfull = [O.complex_gaussian(rng, (sb[0] + (5 if st_b else 0), sb[1]), dtype) for _ in range(P)]
        b = [x[:sb[0]] for x in bfull]
        da = carve(ctx, [(m, k)] * P, dtype)
        db = carve(ctx, [x.shape for x in bfull], dtype)
        dc = carve(ctx, [(m, n)] * P, dtype)
        for x, h in zip(da, a):
            hh = np.asfortranarray(h)
            L.check(L.lib.tenet_memcpy_h2d(ctx.h, C.c_void_p(x.ptr), hh.ctypes.data_as(C.c_void_p), hh.nbytes, None))
        for x, h in zip(db, bfull):
            hh = np.asfortranarray(h)
            L.check(L.lib.tenet_memcpy_h2d(ctx.h, C.c_void_p(x.ptr), hh.ctypes.data_as(C.c_void_p), hh.nbytes, None))
        pr = ((m, k), list("ik"), sb, list(lb), (m, n), list("ij"))
        if st_b is not None:
            pr = pr + (None, st_b)
        plan = GroupedPlan(ctx, [pr] * P, dtype, conj=[cj] * P)
        try:
            plan.exec([p for j in range(P) for p in (da[j].ptr, db[j].ptr, dc[j].ptr)])
            for j in range(P):
                b2 = np.conj(b[j]) if cj[1] else b[j]
                ref = O.binary_einsum(O.OTensor(np.asarray(a[j], np.complex128), list("ik")),
                                      O.OTensor(np.asarray(b2, np.complex128), list(lb)), ["k"]).permute(list("ij")).data
                assert _rel(dc[j].to_host(), ref) < tol, (cj, j)
        finally:
            plan.close()
