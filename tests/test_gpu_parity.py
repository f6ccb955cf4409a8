"""GPU parity of the HIP path against the CPU oracle (oracle/tenet_oracle.py)
on seeded inputs, through the C-ABI.

Tolerances: FP64 results within 1e-12 relative (norm-wise) of the oracle;
integer-valued FP64 contractions bit-exact; FP32 (c64) within 2e-5 relative.
Gauge-dependent factors (Q, R, U, V) are compared through gauge-invariant
quantities only (products, projectors, isometry), as the reference's tests do.
"""
import numpy as np
import pytest

from oracle import tenet_oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


# ---------------------------------------------------------------- K1: einsum
def _random_case(rng, max_rank=5, max_ext=6):
    labels = list("abcdefghij")
    while True:
        ra, rb = rng.integers(1, max_rank + 1, size=2)
        la = list(rng.choice(labels, ra, replace=False))
        lb = list(rng.choice(labels, rb, replace=False))
        shared = [l for l in la if l in lb]
        if not shared:
            continue
        size = {l: int(rng.integers(1, max_ext + 1)) for l in labels}
        nd = int(rng.integers(0, len(shared) + 1))
        dims = list(rng.choice(shared, nd, replace=False)) if nd else []
        return la, lb, size, dims


@pytest.mark.parametrize("seed", range(40))
def test_einsum_random_labels_c128(ctx, seed):
    import tenet_hip as T

    rng = np.random.default_rng(1000 + seed)
    la, lb, size, dims = _random_case(rng)
    a = O.complex_gaussian(rng, [size[l] for l in la])
    b = O.complex_gaussian(rng, [size[l] for l in lb])
    ref = O.binary_einsum(O.OTensor(a, la), O.OTensor(b, lb), dims)
    ta, tb = T.Tensor.from_host(ctx, a, la), T.Tensor.from_host(ctx, b, lb)
    got = T.binary_einsum(ta, tb, dims)
    assert got.labels == ref.labels
    assert _rel(got.to_host(), ref.data) < 1e-12


@pytest.mark.parametrize("seed", range(10))
def test_einsum_integer_inputs_bit_exact(ctx, seed):
    import tenet_hip as T

    rng = np.random.default_rng(2000 + seed)
    la, lb, size, dims = _random_case(rng)
    a = rng.integers(-8, 9, [size[l] for l in la]) + 1j * rng.integers(-8, 9, [size[l] for l in la])
    b = rng.integers(-8, 9, [size[l] for l in lb]) + 1j * rng.integers(-8, 9, [size[l] for l in lb])
    ref = O.binary_einsum(O.OTensor(a, la), O.OTensor(b, lb), dims)
    got = T.binary_einsum(T.Tensor.from_host(ctx, a, la), T.Tensor.from_host(ctx, b, lb), dims)
    assert np.array_equal(got.to_host(), ref.data)


@pytest.mark.parametrize("seed", range(8))
def test_einsum_random_labels_c64(ctx, seed):
    import tenet_hip as T

    rng = np.random.default_rng(3000 + seed)
    la, lb, size, dims = _random_case(rng)
    a = O.complex_gaussian(rng, [size[l] for l in la], np.complex64)
    b = O.complex_gaussian(rng, [size[l] for l in lb], np.complex64)
    ref = O.binary_einsum(O.OTensor(a.astype(np.complex128), la), O.OTensor(b.astype(np.complex128), lb), dims)
    got = T.binary_einsum(T.Tensor.from_host(ctx, a, la), T.Tensor.from_host(ctx, b, lb), dims)
    assert _rel(got.to_host(), ref.data) < 2e-5


@pytest.mark.parametrize("shape", [(64, 96, 80), (257, 130, 70), (512, 512, 256), (1, 300, 7), (300, 1, 5)])
def test_einsum_matrix_shapes(ctx, shape):
    import tenet_hip as T

    m, n, k = shape
    rng = np.random.default_rng(m + n + k)
    a = O.complex_gaussian(rng, (m, k))
    b = O.complex_gaussian(rng, (k, n))
    got = T.binary_einsum(T.Tensor.from_host(ctx, a, ["i", "k"]), T.Tensor.from_host(ctx, b, ["k", "j"]))
    assert _rel(got.to_host(), a @ b) < 1e-13


def test_einsum_hadamard_dims_empty(ctx):
    """dims=[] absorbs a vector into a tensor (canonize.jl:44)."""
    import tenet_hip as T

    rng = np.random.default_rng(5)
    s = rng.random(7)
    v = O.complex_gaussian(rng, (7, 3, 5))
    got = T.binary_einsum(T.Tensor.from_host(ctx, s.astype(complex), ["k"]), T.Tensor.from_host(ctx, v, ["k", "o", "r"]),
                          dims=[])
    assert _rel(got.to_host(), s[:, None, None] * v) < 1e-15


def test_einsum_shape_errors(ctx):
    import tenet_hip as T

    a = T.Tensor.from_host(ctx, np.ones((2, 3)), ["i", "k"])
    b = T.Tensor.from_host(ctx, np.ones((4, 5)), ["k", "j"])
    with pytest.raises(T.TenetError) as e:
        T.binary_einsum(a, b)
    assert e.value.status == 2  # ERR_SHAPE


# ---------------------------------------------------------------- K2: QR
# single-workgroup path for m, n <= 64 (64x64 at its limit); blocked otherwise (65x63, 66x64, 128x32, ...), with
# two register rows per panel thread up to 512 rows and four up to 1024 (1024x512, 777x300); taller: unblocked
@pytest.mark.parametrize("m,n", [(6, 4), (4, 6), (64, 32), (2, 32), (512, 256), (1, 5), (7, 1), (130, 130), (64, 64),
                                 (65, 63), (128, 32), (4096, 1), (1, 4096), (8, 1000), (66, 64), (1024, 512),
                                 (777, 300), (1500, 40)])
def test_qr_invariants_and_r_diagonal(ctx, m, n):
    import tenet_hip as T

    rng = np.random.default_rng(m * 1000 + n)
    A = O.complex_gaussian(rng, (m, n))
    Q, R = T.tensor_qr_thin(T.Tensor.from_host(ctx, A, ["i", "j"]), ["i"], ["j"], "v")
    q, r = Q.to_host(), R.to_host()
    k = min(m, n)
    assert q.shape == (m, k) and r.shape == (k, n)
    assert np.linalg.norm(q.conj().T @ q - np.eye(k)) < 1e-13 * max(1, k)
    assert _rel(q @ r, A) < 1e-13
    assert np.allclose(np.tril(r, -1), 0)
    _, r_ref = np.linalg.qr(A)
    assert np.allclose(np.abs(np.diag(r)), np.abs(np.diag(r_ref)), rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("m,n", [(64, 32), (17, 9), (3, 50), (64, 64), (1, 1), (40, 1), (9, 64)])
@pytest.mark.parametrize("flags", [0, 1, 2, 4, 7])
@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_qr_small_flags_strides_dtypes(ctx, m, n, flags, dt):
    """tenet_qr_thin_ex on the single-workgroup path (m, n <= 64) with every
    adjoint flag, leading dimensions larger than the matrix and complex64:
    A^H read in place for QR_IN_ADJ, Q^H / R^H written for QR_Q_ADJ / QR_R_ADJ."""
    import ctypes as C

    import tenet_hip as T
    from tenet_hip import _lib as L
    from tenet_hip.device import DTYPES

    rng = np.random.default_rng(m * 131 + n * 7 + flags)
    A = O.complex_gaussian(rng, (m, n))
    k = min(m, n)
    in_adj, q_adj, r_adj = flags & 1, flags & 2, flags & 4
    pad = 3
    src = A.conj().T if in_adj else A  # stored matrix
    lda = src.shape[0] + pad
    Ah = np.zeros((lda, src.shape[1]), dtype=dt)
    Ah[: src.shape[0]] = src
    qshape = (k, m) if q_adj else (m, k)
    rshape = (n, k) if r_adj else (k, n)
    ldq, ldr = qshape[0] + pad, rshape[0] + pad
    dA = T.DeviceArray.from_host(ctx, Ah)
    dQ = T.DeviceArray(ctx, (ldq, qshape[1]), dt).zero_()
    dR = T.DeviceArray(ctx, (ldr, rshape[1]), dt).zero_()
    L.check(L.lib.tenet_qr_thin_ex(ctx.h, None, DTYPES[np.dtype(dt)], flags, m, n, C.c_void_p(dA.ptr), lda,
                                   C.c_void_p(dQ.ptr), ldq, C.c_void_p(dR.ptr), ldr), "qr_thin_ex")
    qh, rh = dQ.to_host(), dR.to_host()
    assert not np.any(qh[qshape[0]:]) and not np.any(rh[rshape[0]:])  # padding rows untouched
    q = qh[: qshape[0]].astype(complex)
    r = rh[: rshape[0]].astype(complex)
    q = q.conj().T if q_adj else q
    r = r.conj().T if r_adj else r
    tol = 1e-13 if dt == np.complex128 else 1e-5
    a_in = Ah[: src.shape[0]].astype(complex)
    a_in = a_in.conj().T if in_adj else a_in
    assert np.linalg.norm(q.conj().T @ q - np.eye(k)) < tol * max(1, k)
    assert _rel(q @ r, a_in) < tol
    assert not np.any(np.tril(r, -1))
    _, r_ref = np.linalg.qr(a_in)
    assert np.allclose(np.abs(np.diag(r)), np.abs(np.diag(r_ref)), rtol=10 * tol, atol=tol)


@pytest.mark.parametrize("m,n,nb", [(64, 32, 64), (2, 2, 8), (17, 9, 5), (64, 64, 130), (8, 40, 3), (1, 1, 1)])
@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_qr_absorb_matches_qr_then_contraction(ctx, m, n, nb, side, dt):
    """tenet_qr_absorb = tensor_qr_thin + binary_einsum(R, B) (canonize.jl:58-70)
    in one launch: Bout equals R B (side 0) or B R^H (side 1) with the R of
    tenet_qr_thin_ex on the same input, and Q Bout = A B / Bout Q^H = B A."""
    import ctypes as C

    import tenet_hip as T
    from tenet_hip import _lib as L
    from tenet_hip.device import DTYPES

    rng = np.random.default_rng(m * 31 + n * 17 + nb + side)
    k = min(m, n)
    code = DTYPES[np.dtype(dt)]
    tol = 1e-13 if dt == np.complex128 else 1e-5
    if side == 0:  # right move: A is m x n, Q m x k, B n x nb
        flags = 0
        A = O.complex_gaussian(rng, (m, n)).astype(dt)
        Bm = O.complex_gaussian(rng, (n, nb)).astype(dt)
        lda, ldq, ldb, ldbo = m, m, n, k
        qshape, oshape = (m, k), (k, nb)
    else:  # left move: A^H is factorised (A stored n x m), Q^H written (k x m), B nb x n
        flags = L.QR_IN_ADJ | L.QR_Q_ADJ
        A = O.complex_gaussian(rng, (n, m)).astype(dt)
        Bm = O.complex_gaussian(rng, (nb, n)).astype(dt)
        lda, ldq, ldb, ldbo = n, k, nb, nb
        qshape, oshape = (k, m), (nb, k)
    dA, dB = T.DeviceArray.from_host(ctx, A), T.DeviceArray.from_host(ctx, Bm)
    dQ, dO = T.DeviceArray(ctx, qshape, dt), T.DeviceArray(ctx, oshape, dt)
    L.check(L.lib.tenet_qr_absorb(ctx.h, None, code, flags, m, n, C.c_void_p(dA.ptr), lda, C.c_void_p(dQ.ptr), ldq,
                                  side, nb, C.c_void_p(dB.ptr), ldb, C.c_void_p(dO.ptr), ldbo), "qr_absorb")
    q, out = dQ.to_host().astype(complex), dO.to_host().astype(complex)
    # the same factorisation through tenet_qr_thin_ex: its R, contracted on the host
    rshape = (n, k) if side else (k, n)
    dQ2, dR2 = T.DeviceArray(ctx, qshape, dt), T.DeviceArray(ctx, rshape, dt)
    L.check(L.lib.tenet_qr_thin_ex(ctx.h, None, code, flags | (L.QR_R_ADJ if side else 0), m, n, C.c_void_p(dA.ptr),
                                   lda, C.c_void_p(dQ2.ptr), ldq, C.c_void_p(dR2.ptr), rshape[0]), "qr_thin_ex")
    r2 = dR2.to_host().astype(complex)
    assert _rel(dQ2.to_host(), dQ.to_host()) < tol  # same reflectors (rows per lane may differ: rounding only)
    A64, B64 = A.astype(complex), Bm.astype(complex)
    if side == 0:
        assert _rel(out, r2 @ B64) < tol
        assert _rel(q @ out, A64 @ B64) < 10 * tol
        assert np.linalg.norm(q.conj().T @ q - np.eye(k)) < tol * max(1, k)
    else:
        assert _rel(out, B64 @ r2) < tol  # r2 holds R^H (n x k)
        assert _rel(out @ q, B64 @ A64) < 10 * tol
        assert np.linalg.norm(q @ q.conj().T - np.eye(k)) < tol * max(1, k)


def test_qr_absorb_rejects_large_and_overlap(ctx):
    import ctypes as C

    import tenet_hip as T
    from tenet_hip import _lib as L

    a = T.DeviceArray(ctx, (65, 8)).zero_()
    q = T.DeviceArray(ctx, (65, 8))
    b = T.DeviceArray(ctx, (8, 4)).zero_()
    o = T.DeviceArray(ctx, (8, 4))
    st = L.lib.tenet_qr_absorb(ctx.h, None, L.C128, 0, 65, 8, C.c_void_p(a.ptr), 65, C.c_void_p(q.ptr), 65, 0, 4,
                               C.c_void_p(b.ptr), 8, C.c_void_p(o.ptr), 8)
    assert st == L.TENET_ERR_UNSUPPORTED
    st = L.lib.tenet_qr_absorb(ctx.h, None, L.C128, 0, 64, 8, C.c_void_p(a.ptr), 65, C.c_void_p(q.ptr), 65, 0, 4,
                               C.c_void_p(b.ptr), 8, C.c_void_p(b.ptr), 8)
    assert st == L.TENET_ERR_ARG


def test_qr_rank_deficient_integer(ctx):
    """reshape(1:64, 16, 4) has rank 2: Q must still be an isometry."""
    import tenet_hip as T

    A = np.arange(1, 65, dtype=float).reshape(16, 4, order="F").astype(complex)
    Q, R = T.tensor_qr_thin(T.Tensor.from_host(ctx, A, ["i", "j"]), ["i"], ["j"], "v")
    q, r = Q.to_host(), R.to_host()
    assert np.linalg.norm(q.conj().T @ q - np.eye(4)) < 1e-13
    assert _rel(q @ r, A) < 1e-14


# ---------------------------------------------------------------- K3: SVD
# (1500, 1100): 70 blocks, row split capped at 16 chunks of 94 rows (two staged passes per workgroup), no QR
# preconditioning (p > 512)
@pytest.mark.parametrize("m,n", [(6, 4), (4, 6), (33, 20), (64, 64), (100, 300), (512, 512), (1, 1), (1, 9), (9, 1),
                                 (40, 70), (1500, 1100)])
def test_svd_values_and_factors(ctx, m, n):
    import tenet_hip as T

    rng = np.random.default_rng(m * 7919 + n)
    A = O.complex_gaussian(rng, (m, n))
    U, s, V = T.tensor_svd_thin(T.Tensor.from_host(ctx, A, ["i", "j"]), ["i"], "s")
    u, v = U.to_host(), V.to_host()
    s_ref = np.linalg.svd(A, compute_uv=False)
    k = min(m, n)
    assert np.allclose(s, s_ref, rtol=1e-12, atol=1e-13 * s_ref[0])
    assert np.all(np.diff(s) <= 0)
    assert np.linalg.norm(u.conj().T @ u - np.eye(k)) < 1e-12 * max(1, k)
    assert np.linalg.norm(v @ v.conj().T - np.eye(k)) < 1e-12 * max(1, k)
    assert _rel((u * s) @ v, A) < 2e-14 * max(8, k)


def test_svd_rank_deficient_and_truncation(ctx):
    """Rank-1 first tensor of mps.jl:465-472: one nonzero singular value."""
    import tenet_hip as T

    A = np.ones((4, 16), dtype=complex)
    U, s, V = T.tensor_svd_thin(T.Tensor.from_host(ctx, A, ["i", "j"]), ["i"], "s", threshold=1e-10)
    assert len(s) == 1 and abs(s[0] - 8.0) < 1e-12
    rec = (U.to_host().reshape(4, 1) * s) @ V.to_host().reshape(1, 16)
    assert _rel(rec, A) < 1e-14
    rng = np.random.default_rng(4)
    B = O.complex_gaussian(rng, (50, 40))
    U, s, V = T.tensor_svd_thin(T.Tensor.from_host(ctx, B, ["i", "j"]), ["i"], "s", maxdim=7)
    assert len(s) == 7 and U.shape == (50, 7) and V.shape == (7, 40)
    assert np.allclose(s, np.linalg.svd(B, compute_uv=False)[:7], rtol=1e-12)


def test_svd_graded_spectrum(ctx):
    """Singular values spanning 1 .. 1e-12: absolute error ~ eps * s_max."""
    import tenet_hip as T

    rng = np.random.default_rng(9)
    n = 96
    qa, _ = np.linalg.qr(O.complex_gaussian(rng, (n, n)))
    qb, _ = np.linalg.qr(O.complex_gaussian(rng, (n, n)))
    sv = np.logspace(0, -12, n)
    A = (qa * sv) @ qb.conj().T
    _, s, _ = T.tensor_svd_thin(T.Tensor.from_host(ctx, A, ["i", "j"]), ["i"], "s")
    assert np.max(np.abs(s - sv)) < 1e-13


@pytest.mark.parametrize("m,n", [(1024, 512), (1500, 300)])
def test_svd_graded_large(ctx, m, n):
    """sigma over 12 decades with random singular vectors: plain one-sided
    Jacobi does not finish within the sweep limit here; up to 1024 rows the
    QR-preconditioned path runs first, above it is the fallback after the
    plain sweeps report no convergence."""
    import tenet_hip as T

    rng = np.random.default_rng(m + n)
    k = min(m, n)
    qa, _ = np.linalg.qr(O.complex_gaussian(rng, (m, k)))
    qb, _ = np.linalg.qr(O.complex_gaussian(rng, (n, k)))
    sv = np.logspace(0, -12, k)
    A = (qa * sv) @ qb.conj().T
    U, s, V = T.tensor_svd_thin(T.Tensor.from_host(ctx, A, ["i", "j"]), ["i"], "s")
    s_ref = np.linalg.svd(A, compute_uv=False)
    assert np.max(np.abs(np.asarray(s) - s_ref)) < 1e-13
    u, v = U.to_host(), V.to_host()
    assert _rel((u * s) @ v, A) < 1e-12


# ---------------------------------------------------------------- canonize / compress
def _int_arrays():
    def cm(n, shape):
        return np.arange(1, n + 1, dtype=float).reshape(shape, order="F")
    return [cm(16, (4, 4)), cm(64, (4, 4, 4)), cm(64, (4, 4, 4)), cm(64, (4, 4, 4)), cm(16, (4, 4))]


def _dense(sites):
    v = sites[0].reshape(sites[0].shape[1], sites[0].shape[2])
    for a in sites[1:]:
        v = np.tensordot(v, a, axes=([v.ndim - 1], [0])).reshape(-1, a.shape[2])
    return v.reshape(-1)


@pytest.mark.parametrize("lo,hi", [(2, 2), (1, 2), (0, 0), (4, 4), (0, 4)])
def test_canonize_integer_mps(ctx, lo, hi):
    """test/unit/mps.jl:62-443: isometry pattern + contract(psi_c) ~ contract(psi)."""
    import tenet_hip as T

    psi = T.MPS.from_arrays(ctx, _int_arrays())
    psi.form = ("none",)
    ref = O.OMPS.from_tenet_arrays(_int_arrays()).to_dense()
    psi.canonize_(lo, hi)
    host = psi.to_host()
    om = O.OMPS(host, ("mixed", lo, hi))
    assert O.checkform(om, atol=1e-12)
    assert _rel(_dense(host), ref) < 1e-13


def test_canonize_sweep_and_bond_transitions(ctx):
    import tenet_hip as T

    psi = T.MPS.from_arrays(ctx, _int_arrays())
    ref = O.OMPS.from_tenet_arrays(_int_arrays()).to_dense()
    for step in [("m", 4), ("m", 2), ("b", 1), ("b", 3), ("b", 0), ("m", 3)]:
        if step[0] == "m":
            psi.canonize_(step[1])
        else:
            psi.canonize_bond_(step[1])
        om = O.OMPS(psi.to_host(), psi.form)
        assert O.checkform(om, atol=1e-12), step
    assert psi.form == ("mixed", 3, 3)
    assert _rel(_dense(psi.to_host()), ref) < 1e-13


def test_compress_reference_cases(ctx):
    """mps.jl:465-522."""
    import tenet_hip as T

    arrays = _int_arrays()
    arrays[0] = np.ones((4, 4))
    ref = O.OMPS.from_tenet_arrays(arrays).to_dense()
    for kw in ({"maxdim": 1}, {"threshold": 0.1}):
        psi = T.MPS.from_arrays(ctx, arrays)
        psi.compress_bond_(0, **kw)
        assert psi.bondsizes()[0] == 1
        host = psi.to_host()
        host[0] = host[0] * psi.lam[None, None, :]
        assert _rel(_dense(host), ref) < 1.49e-8
    psi = T.MPS.from_arrays(ctx, arrays)
    psi.compress_(maxdim=1)
    assert psi.bondsizes() == [1, 1, 1, 1]
    # threshold = 1.1 x second Schmidt value of bond 1-2, taken from the same
    # implementation as the reference test does (mps.jl:510-511); for this
    # rank-1 bond that value is rounding noise, so it cannot come from LAPACK
    thr = T.MPS.from_arrays(ctx, arrays).schmidt_values(0)[1] * 1.1
    psi = T.MPS.from_arrays(ctx, arrays)
    psi.compress_(threshold=thr)
    assert psi.bondsizes()[0] == 1
    host = psi.to_host()
    host[-2] = host[-2] * psi.lam[None, None, :]
    assert _rel(_dense(host), ref) < 1.49e-8


def test_truncate_bondsizes(ctx):
    """mps.jl:524-531: [2,4,8,16,8,4,2] then truncate!(maxdim=10) -> max 10."""
    import tenet_hip as T

    rng = np.random.default_rng(1)
    om = O.random_gaussian_mps(rng, 8, 2, 128, cap_edges=True)
    psi = T.MPS.from_host(ctx, om.A)
    assert psi.bondsizes() == [2, 4, 8, 16, 8, 4, 2]
    psi.truncate_(maxdim=10)
    assert max(psi.bondsizes()) == 10
    O.compress(om, maxdim=10)
    assert psi.bondsizes() == om.bondsizes()


# ---------------------------------------------------------------- gates / TEBD
def test_apply_x_and_cx(ctx):
    """test/unit/simple_update.jl:4-43."""
    import tenet_hip as T

    X = np.array([[0, 1], [1, 0]], dtype=complex)
    p = T.MPS.product_state(ctx, [[1, 0], [0, 1]])
    p.apply_one_site_(X, 0)
    assert np.array_equal(p.to_host()[0].reshape(-1), [0, 1])
    CX = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=complex)
    p = T.MPS.product_state(ctx, [[0, 1], [0, 1]])
    p.simple_update_(CX, 0)
    a, b = p.to_host()
    assert np.allclose(np.abs(a[0, :, 0]), [0, 1])
    assert np.allclose(np.abs(b[0, :, 0]), [1, 0])
    if a.shape[2] > 1:
        assert np.allclose(np.abs(a[0, :, 1]), 0)


@pytest.mark.parametrize("n,depth,maxdim", [(8, 6, None), (10, 8, 8), (12, 10, 16)])
def test_tebd_vs_oracle(ctx, n, depth, maxdim):
    import tenet_hip as T
    from tenet_hip.tebd import tebd_sweep, upload_gates

    sched, gates = O.tebd_gates(77 + n, n, depth)
    ref = O.tebd_run(n, sched, gates, maxdim, 1e-10)
    psi = tebd_sweep(ctx, n, sched, upload_gates(ctx, gates), maxdim, 1e-10)
    assert psi.bondsizes() == ref.bondsizes()
    assert _rel(_dense(psi.to_host()), ref.to_dense()) < 1e-10


# ---------------------------------------------------------------- sweeps
def test_overlap_norm_amplitudes_expectation(ctx):
    import tenet_hip as T

    rng = np.random.default_rng(11)
    om = O.random_gaussian_mps(rng, 9, 2, 6)
    vec = om.to_dense()
    psi = T.MPS.from_host(ctx, om.A)
    ov = psi.overlap(psi)
    assert abs(ov - np.vdot(vec, vec)) < 1e-12 * abs(np.vdot(vec, vec))
    bits = rng.integers(0, 2, (5, 9))
    amps = psi.amplitudes(bits)
    for b, a in zip(bits, amps):
        idx = int("".join(map(str, b)), 2)
        assert abs(a - vec[idx]) < 1e-12 * np.abs(vec).max()
    nrm = psi.copy().norm()
    assert abs(nrm ** 2 - np.vdot(vec, vec).real) < 1e-12 * nrm ** 2
    # expectation on a canonical copy
    p2 = psi.copy()
    p2.canonize_(0)
    o2 = O.OMPS([a.copy() for a in om.A])
    O.canonize_mixed(o2, 0, 0)
    for start, ps in [(0, "X"), (3, "ZZ"), (5, "XYZ"), (2, "YXZI")]:
        got = p2.expect_pauli(start, ps)
        ref = O.expectation_pauli(o2, start, ps)
        assert abs(got - ref) < 1e-10 * abs(np.vdot(vec, vec))


# bonds <= 32 with d chi^2 <= 2048 take the one-workgroup sweep; chi 40 the per-site launches
@pytest.mark.parametrize("n,d,chi_a,chi_b", [(9, 2, 6, 5), (6, 3, 8, 7), (7, 2, 32, 32), (7, 2, 40, 24),
                                             (1, 2, 1, 1)])
def test_overlap_of_two_states(ctx, n, d, chi_a, chi_b):
    """overlap(a, b) = <b|a> (overlap.jl:36-50) against the dense vectors."""
    import tenet_hip as T

    rng = np.random.default_rng(100 * n + chi_a)
    oa = O.random_gaussian_mps(rng, n, d, chi_a)
    ob = O.random_gaussian_mps(rng, n, d, chi_b)
    a, b = T.MPS.from_host(ctx, oa.A), T.MPS.from_host(ctx, ob.A)
    ref = np.vdot(ob.to_dense(), oa.to_dense())
    got = a.overlap(b)
    scale = np.linalg.norm(oa.to_dense()) * np.linalg.norm(ob.to_dense())
    assert abs(got - ref) < 1e-12 * scale


def test_overlap_large_dynamic_range(ctx):
    """Unnormalised Gaussian MPS (n=20, chi=32): <psi|psi> ~ 1e36 needs the
    running log-scale."""
    import tenet_hip as T

    rng = np.random.default_rng(20261202)
    om = O.random_gaussian_mps(rng, 20, 2, 32)
    psi = T.MPS.from_host(ctx, om.A)
    m, lg = psi.overlap(psi, log=True)
    mr, lgr = O.overlap_mps_logscaled(om, om)
    assert abs(np.log(abs(m)) + lg - (np.log(abs(mr)) + lgr)) < 1e-10


# ---------------------------------------------------------------- split precision
# relative error bounds per split form: bf16 / fp16 (six 8-bit slices) and
# int8 (seven 7-bit slices) are FP64 grade; int8 with five slices is 2^-35
# relative to the row / column maxima
SPLIT_TOL = {2: 5e-9, 3: 5e-9, 4: 1e-12, 5: 5e-9}


@pytest.mark.parametrize("prec", [2, 3, 4, 5])
@pytest.mark.parametrize("seed", range(6))
def test_einsum_split_precision_random_labels(ctx, prec, seed):
    """Split slices with exact block accumulation (bf16 / fp16 MFMA with FP32
    blocks, int8 MFMA with int32 blocks) against the FP64 oracle."""
    import tenet_hip as T

    rng = np.random.default_rng(4000 + seed)
    la, lb, size, dims = _random_case(rng)
    a = O.complex_gaussian(rng, [size[l] for l in la])
    b = O.complex_gaussian(rng, [size[l] for l in lb])
    ref = O.binary_einsum(O.OTensor(a, la), O.OTensor(b, lb), dims)
    got = T.binary_einsum(T.Tensor.from_host(ctx, a, la), T.Tensor.from_host(ctx, b, lb), dims, prec=prec)
    assert _rel(got.to_host(), ref.data) < SPLIT_TOL[prec]


@pytest.mark.parametrize("prec", [2, 3, 4, 5])
@pytest.mark.parametrize("shape", [(96, 80, 2048), (300, 7, 129), (1, 64, 1000), (70, 90, 9000)])
def test_einsum_split_precision_long_k(ctx, prec, shape):
    """Several 128-long exact blocks summed in FP64, with rows scaled over
    20 orders of magnitude (per-row / per-column power-of-two scales)."""
    import tenet_hip as T

    m, n, k = shape
    rng = np.random.default_rng(m * 3 + n + k)
    a = O.complex_gaussian(rng, (m, k)) * np.logspace(-10, 10, m)[:, None]
    b = O.complex_gaussian(rng, (k, n)) * np.logspace(5, -5, n)[None, :]
    ref = a @ b
    got = T.binary_einsum(T.Tensor.from_host(ctx, a, ["i", "k"]), T.Tensor.from_host(ctx, b, ["k", "j"]), prec=prec)
    g = got.to_host()
    # error relative to each output's own row/column magnitude
    rel = np.abs(g - ref) / (np.abs(a).max(1)[:, None] * np.abs(b).max(0)[None, :] * np.sqrt(k))
    assert rel.max() < (1e-9 if prec != 4 else 1e-13)
    assert _rel(g, ref) < SPLIT_TOL[prec]


@pytest.mark.parametrize("prec", [2, 4, 5])
def test_einsum_split_integer_exact(ctx, prec):
    """Small integers are represented exactly by the first slice: the split
    product of integer matrices is exact."""
    import tenet_hip as T

    rng = np.random.default_rng(8)
    a = rng.integers(-100, 101, (70, 300)) + 1j * rng.integers(-100, 101, (70, 300))
    b = rng.integers(-100, 101, (300, 50)) + 1j * rng.integers(-100, 101, (300, 50))
    got = T.binary_einsum(T.Tensor.from_host(ctx, a, ["i", "k"]), T.Tensor.from_host(ctx, b, ["k", "j"]), prec=prec)
    assert np.array_equal(got.to_host(), a @ b)


@pytest.mark.parametrize("conj", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_einsum_split_i8_batched_conj_alpha_beta(ctx, conj):
    """The int8 split form through a plan with conjugated operands, a batch
    label and C = alpha A B + beta C, against numpy (complex128, seven
    slices: 1e-13), and complex64 operands (four slices: FP32 grade)."""
    import ctypes as C

    import tenet_hip as T
    from tenet_hip import _lib as L
    from tenet_hip.device import tdesc

    rng = np.random.default_rng(31 + 2 * conj[0] + conj[1])
    for dt, tol in ((np.complex128, 1e-13), (np.complex64, 2e-6)):
        a = O.complex_gaussian(rng, (37, 130, 3)).astype(dt)
        b = O.complex_gaussian(rng, (130, 45, 3)).astype(dt)
        c0 = O.complex_gaussian(rng, (37, 45, 3)).astype(dt)
        A, B = T.DeviceArray.from_host(ctx, a, dt), T.DeviceArray.from_host(ctx, b, dt)
        Cd = T.DeviceArray.from_host(ctx, c0, dt)
        h = C.c_void_p()
        L.check(L.lib.tenet_einsum_plan(ctx.h, A.code, L.PREC_SPLIT_I8, C.byref(tdesc(a.shape, [1, 2, 3])),
                                        C.byref(tdesc(b.shape, [2, 4, 3])), C.byref(tdesc(c0.shape, [1, 4, 3])),
                                        C.byref(h)), "plan")
        al = (C.c_double * 2)(0.5, -1.25)
        be = (C.c_double * 2)(-0.75, 0.5)
        L.check(L.lib.tenet_einsum_exec(h, None, C.c_void_p(A.ptr), C.c_void_p(B.ptr), C.c_void_p(Cd.ptr), conj[0],
                                        conj[1], al, be), "exec")
        L.lib.tenet_plan_destroy(h)
        aa = np.conj(a) if conj[0] else a
        bb = np.conj(b) if conj[1] else b
        ref = (0.5 - 1.25j) * np.einsum("ikz,kjz->ijz", aa.astype(np.complex128), bb.astype(np.complex128)) + \
            (-0.75 + 0.5j) * c0.astype(np.complex128)
        assert _rel(Cd.to_host(), ref) < tol


@pytest.mark.parametrize("conj", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("layout", ["nn", "tt"])
def test_einsum_plain_gemm_library_path(ctx, conj, layout):
    """Native-precision contractions that are plain column-major GEMMs of at
    least 1e8 flops run through the vendor library (strided-batched, with a
    batch label, C = alpha op(A) op(B) + beta C): A.B with both operands
    untransposed (a conjugated one then stays in the tile kernel) and with
    both transposed (conjugation folds into op C), against numpy."""
    import ctypes as C

    import tenet_hip as T
    from tenet_hip import _lib as L
    from tenet_hip.device import tdesc

    rng = np.random.default_rng(51 + 2 * conj[0] + conj[1] + (10 if layout == "tt" else 0))
    m, n, k, z = 160, 144, 200, 3  # 8 m n k z = 1.1e8 flops
    for dt, tol in ((np.complex128, 1e-13), (np.complex64, 2e-6)):
        if layout == "nn":
            a = O.complex_gaussian(rng, (m, k, z)).astype(dt)
            b = O.complex_gaussian(rng, (k, n, z)).astype(dt)
            la, lb, spec = [1, 2, 3], [2, 4, 3], "ikz,kjz->ijz"
        else:
            a = O.complex_gaussian(rng, (k, m, z)).astype(dt)
            b = O.complex_gaussian(rng, (n, k, z)).astype(dt)
            la, lb, spec = [2, 1, 3], [4, 2, 3], "kiz,jkz->ijz"
        c0 = O.complex_gaussian(rng, (m, n, z)).astype(dt)
        A, B = T.DeviceArray.from_host(ctx, a, dt), T.DeviceArray.from_host(ctx, b, dt)
        Cd = T.DeviceArray.from_host(ctx, c0, dt)
        h = C.c_void_p()
        L.check(L.lib.tenet_einsum_plan(ctx.h, A.code, L.PREC_NATIVE, C.byref(tdesc(a.shape, la)),
                                        C.byref(tdesc(b.shape, lb)), C.byref(tdesc(c0.shape, [1, 4, 3])),
                                        C.byref(h)), "plan")
        al = (C.c_double * 2)(0.5, -1.25)
        be = (C.c_double * 2)(-0.75, 0.5)
        L.check(L.lib.tenet_einsum_exec(h, None, C.c_void_p(A.ptr), C.c_void_p(B.ptr), C.c_void_p(Cd.ptr), conj[0],
                                        conj[1], al, be), "exec")
        L.lib.tenet_plan_destroy(h)
        aa = np.conj(a) if conj[0] else a
        bb = np.conj(b) if conj[1] else b
        ref = (0.5 - 1.25j) * np.einsum(spec, aa.astype(np.complex128), bb.astype(np.complex128)) + \
            (-0.75 + 0.5j) * c0.astype(np.complex128)
        assert _rel(Cd.to_host(), ref) < tol


# ---------------------------------------------------------------- DMRG H_eff
@pytest.mark.parametrize("prec,tol", [(0, 1e-13), (1, 1e-5), (2, 1e-12), (3, 1e-12), (4, 1e-12),
                                      (5, 5e-9)])
def test_heff_apply_vs_oracle(ctx, prec, tol):
    """EffectiveHamiltonian apply (DMRG.jl:16-18) on an XXZ MPO with random
    environments, every precision against the FP64 oracle."""
    import tenet_hip as T
    from tenet_hip.dmrg import heff_apply, xxz_mpo_tensor

    rng = np.random.default_rng(77)
    chi = 24
    W = xxz_mpo_tensor(1.0, 0.5)
    assert np.allclose(W, O.xxz_mpo_W(1.0, 0.5))
    Lh = O.complex_gaussian(rng, (chi, 5, chi))
    Rh = O.complex_gaussian(rng, (chi, 5, chi))
    th = O.complex_gaussian(rng, (chi, 2, 2, chi))
    th /= np.linalg.norm(th)
    ref = O.heff_apply(Lh, W, W, Rh, th)
    y = heff_apply(T.Tensor.from_host(ctx, Lh, ["l", "w", "lk"]), T.Tensor.from_host(ctx, W, ["w", "m", "o1", "i1"]),
                   T.Tensor.from_host(ctx, W, ["m", "wr", "o2", "i2"]), T.Tensor.from_host(ctx, Rh, ["r", "wr", "rk"]),
                   T.Tensor.from_host(ctx, th, ["lk", "i1", "i2", "rk"]), prec=prec)
    assert y.labels == ["l", "o1", "o2", "r"]
    assert _rel(y.to_host(), ref) < tol


def test_heff_ising_spectrum(ctx):
    """The Ising ground energy pinned by test/unit/dmrg.jl:38 enters through
    the MPO; on a 4-site chain with unitary outer sites H_eff has exactly the
    spectrum of H (eigvalsh of the 16 x 16 H_eff built column by column)."""
    import tenet_hip as T
    from tenet_hip.dmrg import heff_apply

    W = O.ising_mpo_W(1.0, 1.0)
    lv = np.zeros(3); lv[2] = 1
    rv = np.zeros(3); rv[0] = 1
    rng = np.random.default_rng(2)
    A1, A4 = O.haar_unitary(rng, 2), O.haar_unitary(rng, 2)
    lW = np.einsum("l,lroi->roi", lv, W)
    rW = np.einsum("lroi,r->loi", W, rv)
    Lenv = np.einsum("oa,woi,ib->awb", A1.conj(), lW, A1)
    Renv = np.einsum("do,woi,ci->dwc", A4.conj(), rW, A4)
    M = np.zeros((16, 16), complex)
    tl = [T.Tensor.from_host(ctx, Lenv, ["l", "w", "lk"]), T.Tensor.from_host(ctx, W, ["w", "m", "o1", "i1"]),
          T.Tensor.from_host(ctx, W, ["m", "wr", "o2", "i2"]), T.Tensor.from_host(ctx, Renv, ["r", "wr", "rk"])]
    for col in range(16):
        e = np.zeros(16, complex); e[col] = 1
        y = heff_apply(*tl, T.Tensor.from_host(ctx, e.reshape(2, 2, 2, 2, order="F"), ["lk", "i1", "i2", "rk"]))
        M[:, col] = y.to_host().reshape(-1, order="F")
    H = O.mpo_dense([W] * 4, lv, rv)
    assert np.allclose(np.linalg.eigvalsh(M), np.linalg.eigvalsh(H), atol=1e-12)


# ---------------------------------------------------------------- synthetic inputs and helpers
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_fill_gaussian_matches_host_mirror(ctx, dtype):
    """tenet_fill_gaussian against oracle/synth.py element for element,
    including an offset slice of the same stream."""
    import tenet_hip as T
    from oracle import synth

    d = T.DeviceArray(ctx, (1000, 37), dtype).gaussian_(11)
    ref = synth.gaussian(11, 37000, dtype=dtype).reshape(1000, 37, order="F")
    tol = 1e-13 if dtype == np.complex128 else 1e-6
    assert _rel(d.to_host(), ref) < tol
    s = T.DeviceArray(ctx, (500,), dtype).gaussian_(11, offset=1234)
    assert _rel(s.to_host(), ref.reshape(-1, order="F")[1234:1734]) < tol


def test_scale_and_nrm2sq(ctx):
    import tenet_hip as T

    rng = np.random.default_rng(5)
    a = O.complex_gaussian(rng, (333, 77))
    d = T.DeviceArray.from_host(ctx, a)
    assert abs(d.nrm2sq() - np.linalg.norm(a) ** 2) < 1e-12 * np.linalg.norm(a) ** 2
    d.scale_(0.5 - 2.0j)
    assert _rel(d.to_host(), (0.5 - 2.0j) * a) < 1e-15
    assert T.DeviceArray(ctx, (0,)).nrm2sq() == 0.0


def test_normalize_and_norm_c64_and_c128(ctx):
    """normalize!(::MPS) (normalize.jl:15-37) scales the canonical centre by a
    0-d device scalar: norm 1 afterwards, state direction unchanged."""
    import tenet_hip as T

    rng = np.random.default_rng(31)
    for dtype, tol in ((np.complex128, 1e-12), (np.complex64, 2e-6)):
        arrs = [O.complex_gaussian(rng, s) for s in [(1, 2, 4), (4, 2, 5), (5, 2, 3), (3, 2, 1)]]
        psi = T.MPS.from_host(ctx, arrs, dtype)
        ref = O.OMPS([a.astype(np.complex128) for a in arrs])
        nref = O.norm(ref.copy())
        assert abs(psi.norm() - nref) < tol * nref
        psi.normalize_()
        assert abs(psi.norm() - 1.0) < tol
        v = O.OMPS([a.astype(np.complex128) for a in psi.to_host()]).to_dense()
        assert _rel(v, ref.to_dense() / nref) < tol


@pytest.mark.parametrize("m,n,plain", [(100, 64, True), (64, 100, True), (300, 128, True), (128, 300, True),
                                       (512, 512, False), (200, 512, False)])
def test_svd_preconditioned_matches_plain(ctx, m, n, plain):
    """The QR-preconditioned Jacobi (Jacobi on R^H) on a spectrum graded from
    1 to 1e-13: σ, the truncation count and the rank-k product against the
    construction, and against the plain one-sided Jacobi where that one
    converges (at 512 columns and this grading the plain sweeps do not
    settle within 40 sweeps: the reason for the QR step)."""
    import tenet_hip as T

    rng = np.random.default_rng(m * 7 + n)
    k = min(m, n)
    qa, _ = np.linalg.qr(O.complex_gaussian(rng, (m, k)))
    qb, _ = np.linalg.qr(O.complex_gaussian(rng, (n, k)))
    sv = np.logspace(0, -13, k)
    A = (qa * sv) @ qb.conj().T
    res = {}
    for on in ((True, False) if plain else (True,)):
        ctx.svd_precondition(on)
        U, S, V = T.tensor_svd_thin(T.Tensor.from_host(ctx, A, ["i", "j"]), ["i"], "s", ["j"], maxdim=k - 3,
                                    threshold=1e-10)
        res[on] = (U.to_host(), S, V.to_host())
    ctx.svd_precondition(True)
    U1, S1, V1 = res[True]
    assert len(S1) == int(np.count_nonzero(sv[:k - 3] > 1e-10))
    assert np.allclose(S1, sv[:len(S1)], rtol=1e-10, atol=1e-15)
    if plain:
        S0 = res[False][1]
        assert len(S0) == len(S1)
        assert np.allclose(S1, S0, rtol=1e-10, atol=1e-15)
    best = (qa[:, :len(S1)] * sv[:len(S1)]) @ qb[:, :len(S1)].conj().T
    assert _rel((U1 * S1) @ V1, best) < 1e-11
    # vectors formed as X / σ (one side or the other, per path) are orthogonal
    # to ~eps σ1 / σ: check the well-separated part, σ > 1e-5
    g = int(np.count_nonzero(S1 > 1e-5))
    assert np.allclose(U1[:, :g].conj().T @ U1[:, :g], np.eye(g), atol=1e-11)
    assert np.allclose(V1[:g] @ V1[:g].conj().T, np.eye(g), atol=1e-11)


def test_gate_bond_diagonal_move_matches_qr(ctx):
    """After simple_update! (AbsorbEqually) the next canonize! move across the
    gate's bond is a diagonal rescaling; it must give the same state and the
    same canonical form as the Householder move."""
    import tenet_hip as T
    from tenet_hip.tebd import brickwork_schedule, haar_gates

    n = 10
    sched = brickwork_schedule(n, 6)
    gates = haar_gates(123, len(sched))
    psi_fast = T.MPS.product_state(ctx, [[1.0, 0.0]] * n)
    psi_ref = T.MPS.product_state(ctx, [[1.0, 0.0]] * n)
    for a, U in zip(sched, gates):
        psi_fast.simple_update_(U, a, maxdim=16, threshold=1e-12)
        psi_ref.simple_update_(U, a, maxdim=16, threshold=1e-12)
        psi_ref._gate_bond = None  # force the Householder move
    for psi in (psi_fast, psi_ref):
        psi.canonize_(0)
    assert psi_fast.bondsizes() == psi_ref.bondsizes()
    vf = O.OMPS([a for a in psi_fast.to_host()]).to_dense()
    vr = O.OMPS([a for a in psi_ref.to_host()]).to_dense()
    assert _rel(vf, vr) < 1e-12
    for a in psi_fast.to_host()[1:]:
        assert O.is_right_isometry(a, atol=1e-12)


def _isometry_pattern_ok(psi, lo, hi, atol=1e-11):
    """Sites left of lo are left isometries and sites right of hi right
    isometries (the MixedCanonical(lo..hi) pattern, canonize.jl:128-149)."""
    sites = psi.to_host()
    return (all(O.is_left_isometry(a, atol=atol) for a in sites[:lo])
            and all(O.is_right_isometry(a, atol=atol) for a in sites[hi + 1:]))


@pytest.mark.parametrize("case", ["left_across_gate", "widen_then_shrink", "right_across_gate"])
def test_gate_bond_diagonal_move_directions(ctx, case):
    """The diagonal shortcut may only replace the canonize! move across the
    gate's own bond, in either direction.  Each case ends in a form whose
    every site outside the centre is an isometry, with the same state as the
    Householder moves give."""
    import tenet_hip as T
    from tenet_hip.tebd import haar_gates

    n = 8
    gates = haar_gates(321, 6)
    fast = T.MPS.product_state(ctx, [[1.0, 0.0]] * n)
    ref = T.MPS.product_state(ctx, [[1.0, 0.0]] * n)
    # entangle the chain so every bond is > 1
    for g, a in enumerate([0, 2, 4, 6, 1, 3]):
        for psi in (fast, ref):
            psi.simple_update_(gates[g], a, maxdim=8, threshold=1e-12)
            if psi is ref:
                psi._gate_bond = None
    # the last gate acted on (3, 4)
    steps = {"left_across_gate": [(3, 3), (2, 2)],
             "widen_then_shrink": [(3, 6), (3, 4)],
             "right_across_gate": [(4, 4), (5, 5)]}[case]
    for lo, hi in steps:
        for psi in (fast, ref):
            psi.canonize_(lo, hi)
            if psi is ref:
                psi._gate_bond = None
        assert fast.form == ref.form == ("mixed", lo, hi)
        assert _isometry_pattern_ok(fast, lo, hi)
        vf = O.OMPS(fast.to_host()).to_dense()
        vr = O.OMPS(ref.to_host()).to_dense()
        assert _rel(vf, vr) < 1e-12


def test_free_async_and_workspace_query(ctx):
    """tenet_free_async queues the block's return behind the stream's earlier
    work: once that work is done the next malloc of its size class reuses
    it; a pointer the pool does not own is refused.
    tenet_workspace_size matches what the calls take: after reserving it, an
    SVD on that stream does not grow the workspace (checked through the
    level-2 log)."""
    import ctypes as C
    import os
    import tempfile

    import tenet_hip as T
    from tenet_hip import _lib as L

    nbytes = (67 << 20) + 16  # a size class (68 MiB) no other test uses
    a = ctx.malloc(nbytes)
    L.check(L.lib.tenet_fill_gaussian(ctx.h, None, L.C128, nbytes // 16, C.c_void_p(a), C.c_uint64(3),
                                      C.c_uint64(0)))
    L.check(L.lib.tenet_free_async(ctx.h, C.c_void_p(a), None), "free_async")
    ctx.sync()
    b = ctx.malloc(nbytes)
    assert b == a
    ctx.free(b)
    assert L.lib.tenet_free_async(ctx.h, C.c_void_p(a + 256), None) == L.TENET_ERR_ARG
    for m, k in ((512, 256), (300, 300)):
        sz = C.c_size_t()
        for op in (L.OP_QR, L.OP_SVD, L.OP_SIMPLE_UPDATE):
            L.check(L.lib.tenet_workspace_size(ctx.h, op, L.C128, m, k, C.byref(sz)), "workspace_size")
            assert sz.value > 0
    sz = C.c_size_t()
    L.check(L.lib.tenet_workspace_size(ctx.h, L.OP_SVD, L.C128, 400, 300, C.byref(sz)))
    s = ctx.stream()
    L.check(L.lib.tenet_workspace_reserve(ctx.h, s, sz.value), "reserve")
    fd, path = tempfile.mkstemp()
    os.close(fd)
    try:
        L.check(L.lib.tenet_set_log(2, path.encode()))
        rng = np.random.default_rng(2)
        A = T.DeviceArray.from_host(ctx, O.complex_gaussian(rng, (400, 300)))
        U, Vh = T.DeviceArray(ctx, (400, 300)), T.DeviceArray(ctx, (300, 300))
        S = ctx.malloc(300 * 8)
        L.check(L.lib.tenet_svd_thin(ctx.h, s, L.C128, 400, 300, C.c_void_p(A.ptr), 400, C.c_void_p(U.ptr), 400,
                                     C.c_void_p(S), C.c_void_p(Vh.ptr), 300, 0, 0.0, None, None), "svd")
        ctx.sync(s)
        ctx.free(S)
    finally:
        L.check(L.lib.tenet_set_log(0, None))
    log = open(path).read()
    os.unlink(path)
    assert "workspace 0 grows" not in log, log
    L.check(L.lib.tenet_stream_destroy(ctx.h, s))
