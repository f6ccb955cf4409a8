"""Graph-captured sweeps (tenet_graph_*, tenet_hip.graphs) and the
site-parallel amplitudes, through the C-ABI.

Captured canonize! + overlap + norm (canonize.jl:136-145, overlap.jl:36-50,
norm.jl:17-37) on the n = 20, chi = 32 state: every replay equals the eager
run bit for bit (same kernels, same order, same inputs), with eager work and
synchronisations between the replays, and both equal the oracle within
1e-12.  A captured tenet_svd_thin equals the eager call bit for bit and
leaves the device status word clear; a matrix its fixed sweep budget does
not finish sets it.  tenet_amplitudes_sites (overlap.jl:22-32) equals the
oracle and the one-workgroup kernel.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import synth  # noqa: F401
from oracle import tenet_oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_canonize_norm_graph_replays_equal_eager(ctx):
    import tenet_hip as T
    from tenet_hip.graphs import CanonizeNormGraph
    from tenet_hip.workloads import plain_state_arrays

    arrs = plain_state_arrays(20261202, 20, 32)
    src = T.MPS.from_host(ctx, arrs)
    p = src.copy()
    p.canonize_(0)
    ov, nr = p.overlap(p), p.norm()
    eager_sites = p.to_host()
    g = CanonizeNormGraph(src, 0)
    try:
        for rep in range(3):
            print(f"graph replay {rep}", flush=True)
            ovg, nrg = g.run()
            assert ovg == ov and nrg == nr
            for a, b in zip(g.state.to_host(), eager_sites):
                assert np.array_equal(a, b)
            # eager work on the same stream between replays
            q = src.copy()
            q.canonize_(5)
            ctx.sync()
            del q
    finally:
        g.close()
    ref = O.OMPS([a.copy() for a in arrs])
    O.canonize_mixed(ref, 0, 0)
    assert abs(ov - O.overlap_mps(ref, ref)) <= 1e-12 * abs(ov)
    assert abs(nr - O.norm(ref)) <= 1e-12 * nr


def _svd_call(ctx, A, m, n, U, S, Vh, kd):
    from tenet_hip import _lib as L

    k = min(m, n)
    return L.lib.tenet_svd_thin(ctx.h, None, L.C128, m, n, C.c_void_p(A.ptr), m, C.c_void_p(U.ptr), m,
                                C.c_void_p(S), C.c_void_p(Vh.ptr), k, k // 2, 1e-10, C.c_void_p(kd), None)


def test_captured_svd_equals_eager(ctx):
    import tenet_hip as T
    from tenet_hip import _lib as L

    m, n = 200, 120
    rng = np.random.default_rng(3)
    A = T.DeviceArray.from_host(ctx, O.complex_gaussian(rng, (m, n)))
    outs = []
    for _ in range(2):
        outs.append((T.DeviceArray(ctx, (m, n)), ctx.malloc(8 * n), T.DeviceArray(ctx, (n, n)), ctx.malloc(64)))
    U0, S0, V0, k0 = outs[0]
    L.check(_svd_call(ctx, A, m, n, U0, S0, V0, k0), "svd (eager)")
    ctx.sync()
    U1, S1, V1, k1 = outs[1]
    with ctx.capture() as g:
        L.check(_svd_call(ctx, A, m, n, U1, S1, V1, k1), "svd (captured)")
    try:
        for rep in range(2):
            g.launch()
            assert ctx.graph_status() == 0
            assert np.array_equal(U1.to_host(), U0.to_host()) and np.array_equal(V1.to_host(), V0.to_host())
            s0, s1, ka, kb = np.zeros(n), np.zeros(n), np.zeros(1, np.int64), np.zeros(1, np.int64)
            for h, d, nb in ((s0, S0, 8 * n), (s1, S1, 8 * n), (ka, k0, 8), (kb, k1, 8)):
                L.check(L.lib.tenet_memcpy_d2h(ctx.h, h.ctypes.data_as(C.c_void_p), C.c_void_p(d), nb, None))
            assert np.array_equal(s0, s1) and ka[0] == kb[0] == n // 2
            assert np.allclose(s0, np.linalg.svd(A.to_host(), compute_uv=False), rtol=1e-12)
    finally:
        g.close()
        for _, S, _, kd in outs:
            ctx.free(S)
            ctx.free(kd)


def test_captured_svd_sets_status_when_sweeps_do_not_suffice(ctx):
    """A graded 800 x 600 matrix (singular values over 12 decades, mixing
    singular vectors) needs more than the captured budget of 30 plain
    sweeps (the eager limit of 40 does not suffice either): the replay
    completes and the status word reports it."""
    import tenet_hip as T
    from tenet_hip import _lib as L

    m, n = 800, 600
    rng = np.random.default_rng(7)
    Q1, _ = np.linalg.qr(O.complex_gaussian(rng, (m, n)))
    Q2, _ = np.linalg.qr(O.complex_gaussian(rng, (n, n)))
    A_h = (Q1 * np.logspace(0, -12, n)) @ Q2.conj().T
    A = T.DeviceArray.from_host(ctx, A_h)
    U, Vh = T.DeviceArray(ctx, (m, n)), T.DeviceArray(ctx, (n, n))
    S, kd = ctx.malloc(8 * n), ctx.malloc(64)
    ctx.svd_precondition(False)
    try:
        # the capture cannot grow the workspace: reserve what this SVD needs
        need = C.c_size_t()
        L.check(L.lib.tenet_workspace_size(ctx.h, L.OP_SVD, L.C128, m, n, C.byref(need)), "workspace_size")
        L.check(L.lib.tenet_workspace_reserve(ctx.h, None, need.value), "workspace_reserve")
        assert ctx.graph_status() == 0
        with ctx.capture() as g:
            L.check(_svd_call(ctx, A, m, n, U, S, Vh, kd), "svd (captured)")
        g.launch()
        assert ctx.graph_status() & 1
        assert ctx.graph_status() == 0  # read clears it
        g.close()
    finally:
        ctx.svd_precondition(True)
        ctx.free(S)
        ctx.free(kd)


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 2e-5)])
def test_site_parallel_amplitudes(ctx, dtype, tol):
    import tenet_hip as T

    rng = np.random.default_rng(11)
    for n, chi in ((14, 100), (9, 70), (6, 8)):
        host = O.random_gaussian_mps(rng, n, 2, chi, cap_edges=True)
        arrs = [a / np.sqrt(2 * a.shape[2]) for a in host.A]
        psi = T.MPS.from_host(ctx, arrs, dtype)
        bits = rng.integers(0, 2, (9, n))
        print(f"amplitudes n={n} chi={chi} {np.dtype(dtype).name}", flush=True)
        a_site = psi.amplitudes(bits, site_parallel=True)
        a_one = psi.amplitudes(bits, site_parallel=False)
        ref_psi = O.OMPS([a.astype(dtype).astype(np.complex128) for a in arrs])
        ref = np.array([O.amplitude(b, ref_psi) for b in bits])
        assert _rel(a_site, ref) <= tol
        assert _rel(a_site, a_one) <= tol


def test_site_parallel_amplitudes_report_nan_beyond_bounds(ctx):
    """A request with more sites than n_max (or a bond above chi_max) gets
    NaN instead of a partial product."""
    import tenet_hip as T
    from tenet_hip import _lib as L

    rng = np.random.default_rng(12)
    n = 8
    arrs = O.random_gaussian_mps(rng, n, 2, 16, cap_edges=True).A
    psi = T.MPS.from_host(ctx, arrs)
    chis = np.asarray(psi.chis(), dtype=np.int64)
    ptrs = np.asarray([s.ptr for s in psi.sites], dtype=np.uint64)
    bits = np.zeros(n, dtype=np.uint8)
    buf = ctx.malloc(4096)
    try:
        p_ptrs, p_chis, p_bits, p_tab, p_out = buf, buf + 512, buf + 1024, buf + 1536, buf + 2048
        tab = np.array([n, p_ptrs, p_chis, p_bits], dtype=np.uint64)
        for src, dst in ((ptrs, p_ptrs), (chis, p_chis), (bits, p_bits), (tab, p_tab)):
            L.check(L.lib.tenet_memcpy_h2d(ctx.h, C.c_void_p(dst), src.ctypes.data_as(C.c_void_p), src.nbytes, None))
        for n_max, chi_max in ((n - 1, 16), (n, 8)):
            L.check(L.lib.tenet_amplitudes_sites(ctx.h, None, L.C128, 1, n_max, chi_max, C.c_void_p(p_tab),
                                                 C.c_void_p(p_tab + 8), C.c_void_p(p_tab + 16),
                                                 C.c_void_p(p_tab + 24), 2, C.c_void_p(p_out)), "amplitudes_sites")
            h = np.zeros(3)
            L.check(L.lib.tenet_memcpy_d2h(ctx.h, h.ctypes.data_as(C.c_void_p), C.c_void_p(p_out), 24, None))
            assert np.isnan(h[0]) and np.isnan(h[1])
    finally:
        ctx.free(buf)
