"""tenet_svd_thin_noalloc, the allocation-free SVD (per-GPU static scratch),
against tenet_svd_thin through the C-ABI.

Two calls in a row, the second with a smaller input (the scratch array keeps
the first call's contents).  Shapes whose tenet_svd_thin also runs the plain
Jacobi (complex64, or min(m, n) < 64) take the same kernels in the same
order: bit-identical outputs.  A complex128 shape with min(m, n) >= 64 is
QR-preconditioned in tenet_svd_thin and not here: singular values within
1e-12 relative, k and the discarded weight equal, U diag(S) V^H equal
within 1e-12 relative.  A shape whose workspace exceeds the scratch array
is refused before any launch.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import tenet_oracle as O

pytestmark = pytest.mark.gpu


def _run(ctx, fname, A_host, dtype, maxdim, thr):
    import tenet_hip as T
    from tenet_hip import _lib as L

    m, n = A_host.shape
    k = min(m, n)
    A = T.DeviceArray.from_host(ctx, A_host, dtype)
    U = T.DeviceArray(ctx, (m, k), dtype)
    Vh = T.DeviceArray(ctx, (k, n), dtype)
    sdev = ctx.malloc(k * 8 + 64)
    try:
        L.check(getattr(L.lib, fname)(ctx.h, None, A.code, m, n, C.c_void_p(A.ptr), m, C.c_void_p(U.ptr), m,
                                      C.c_void_p(sdev), C.c_void_p(Vh.ptr), k, int(maxdim), float(thr),
                                      C.c_void_p(sdev + k * 8), C.c_void_p(sdev + k * 8 + 8)), fname)
        S = np.zeros(k)
        kd = np.zeros(2, dtype=np.int64)
        L.check(L.lib.tenet_memcpy_d2h(ctx.h, S.ctypes.data_as(C.c_void_p), C.c_void_p(sdev), k * 8, None))
        L.check(L.lib.tenet_memcpy_d2h(ctx.h, kd.ctypes.data_as(C.c_void_p), C.c_void_p(sdev + k * 8), 16, None))
    finally:
        ctx.free(sdev)
    disc = float(kd[1:].view(np.float64)[0])
    return U.to_host(), S, Vh.to_host(), int(kd[0]), disc


def _pair(ctx, A, dtype, maxdim, thr, exact):
    ref = _run(ctx, "tenet_svd_thin", A, dtype, maxdim, thr)
    got = _run(ctx, "tenet_svd_thin_noalloc", A, dtype, maxdim, thr)
    Ur, Sr, Vr, kr, dr = ref
    Ug, Sg, Vg, kg, dg = got
    assert kg == kr
    if exact:
        assert np.array_equal(Sg, Sr) and np.array_equal(Ug, Ur) and np.array_equal(Vg, Vr) and dg == dr
        return
    assert np.allclose(Sg, Sr, rtol=1e-12, atol=1e-14 * Sr[0])
    assert abs(dg - dr) <= 1e-12 * max(dr, Sr[0] ** 2 * 1e-16)
    k = kr
    rec_r = (Ur[:, :k] * Sr[:k]) @ Vr[:k]
    rec_g = (Ug[:, :k] * Sg[:k]) @ Vg[:k]
    assert np.linalg.norm(rec_g - rec_r) <= 1e-12 * np.linalg.norm(rec_r)
    # and against LAPACK
    s_ref = np.linalg.svd(A.astype(np.complex128), compute_uv=False)
    assert np.allclose(Sg, s_ref, rtol=1e-10 if dtype == np.complex64 else 1e-12, atol=1e-13 * s_ref[0])


@pytest.mark.parametrize("shapes,dtype,exact", [
    (((40, 24), (20, 12)), np.complex128, True),     # one block pair: plain Jacobi in both
    (((300, 48), (24, 100)), np.complex128, True),    # min < 64: plain in both; wide second call
    (((200, 150), (96, 64)), np.complex64, True),     # complex64 is never preconditioned
    (((512, 512), (300, 200)), np.complex128, False),  # preconditioned in tenet_svd_thin only
])
def test_noalloc_svd_matches_svd_thin_twice(ctx, shapes, dtype, exact):
    rng = np.random.default_rng(11)
    for (m, n), (maxdim, thr) in zip(shapes, ((0, 0.0), (max(1, min(shapes[1]) // 2), 1e-10))):
        A = O.complex_gaussian(rng, (m, n)).astype(dtype)
        print(f"noalloc svd {m}x{n} {np.dtype(dtype).name} maxdim={maxdim}", flush=True)
        _pair(ctx, A, dtype, maxdim, thr, exact)


def test_noalloc_svd_refuses_oversized(ctx):
    import tenet_hip as T
    from tenet_hip import _lib as L

    m, n = 4096, 2048  # X alone is 128 MiB of complex128
    A = T.DeviceArray(ctx, (m, n), np.complex128)
    U = T.DeviceArray(ctx, (m, n), np.complex128)
    Vh = T.DeviceArray(ctx, (n, n), np.complex128)
    st = L.lib.tenet_svd_thin_noalloc(ctx.h, None, L.C128, m, n, C.c_void_p(A.ptr), m, C.c_void_p(U.ptr), m, None,
                                      C.c_void_p(Vh.ptr), n, 0, 0.0, None, None)
    assert st == L.TENET_ERR_UNSUPPORTED
    assert b"scratch" in L.lib.tenet_last_error()
