"""Vidal-gauge TEBD on the device against the oracle's restatement of the
published algorithm (the reference's Vidal update is a TODO stub,
simple_update.jl:94-112).  The oracle itself is pinned in test_oracle.py by
the dense state vector, the sequential simple_update! TEBD and the Schmidt
spectra.

Tolerances: complex128 throughout, FP64 SVD and contractions.  States are
compared through gauge-invariant quantities (dense vectors or amplitudes,
Lambda spectra, bond sizes): relative 1e-10 with truncation (one-sided Jacobi
and LAPACK agree to ~1e-13 on each kept singular value, and the kept sets are
identical), 1e-12 without.
"""
import numpy as np
import pytest

from oracle import tenet_oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _host_vidal(v):
    return O.OVidal([b.to_host() for b in v.B], [x.copy() for x in v.lam_host])


def _gates(rng, layers):
    return np.stack([O.haar_unitary(rng, 4) for _ in range(sum(map(len, layers)))])


@pytest.mark.parametrize("n,depth", [(8, 6), (11, 5)])
def test_vidal_layers_untruncated_equal_dense(ctx, n, depth):
    import tenet_hip as T

    rng = np.random.default_rng(100 + n)
    layers = O.brickwork_layers(n, depth)
    gates = _gates(rng, layers)
    v = T.tebd_vidal_sweep(ctx, n, layers, list(gates), None, None)
    ref = O.vidal_tebd_run(n, layers, gates, None, None)
    hv = _host_vidal(v)
    assert hv.bondsizes() == ref.bondsizes()
    assert _rel(hv.to_mps().to_dense(), ref.to_mps().to_dense()) < 1e-12
    for a, b in zip(hv.lam, ref.lam):
        assert np.allclose(a, b, rtol=1e-12, atol=1e-14)
    assert all(O.is_right_isometry(b, 1e-11) for b in hv.B)
    v.close()


@pytest.mark.parametrize("maxdim,threshold", [(4, 1e-10), (6, 1e-3), (16, 1e-10)])
def test_vidal_layers_truncated_match_oracle(ctx, maxdim, threshold):
    """Same gates, same truncation rule: bond sizes identical, amplitudes and
    Lambda spectra within 1e-10 of the oracle."""
    import tenet_hip as T

    n, depth = 12, 8
    rng = np.random.default_rng(7 * maxdim)
    layers = O.brickwork_layers(n, depth)
    gates = _gates(rng, layers)
    v = T.tebd_vidal_sweep(ctx, n, layers, list(gates), maxdim, threshold)
    ref = O.vidal_tebd_run(n, layers, gates, maxdim, threshold)
    hv = _host_vidal(v)
    assert hv.bondsizes() == ref.bondsizes()
    assert max(hv.bondsizes()) <= maxdim
    assert _rel(hv.to_mps().to_dense(), ref.to_mps().to_dense()) < 1e-10
    for a, b in zip(hv.lam, ref.lam):
        assert np.allclose(a, b, rtol=1e-10, atol=1e-13)
    # amplitudes through the device transfer sweep of the MPS view
    bits = np.random.default_rng(3).integers(0, 2, size=(64, n))
    amps = v.to_mps().amplitudes(bits)
    dense = ref.to_mps().to_dense()
    idx = [int("".join(map(str, b)), 2) for b in bits]
    assert _rel(amps, dense[idx]) < 1e-10
    v.close()


def test_vidal_from_mps_and_layer_on_a_random_state(ctx):
    """from_mps gives the Schmidt spectra of a random MPS; a layer on that
    state matches the oracle's."""
    import tenet_hip as T

    rng = np.random.default_rng(21)
    n = 9
    psi = O.random_gaussian_mps(rng, n, 2, 8)
    O.normalize(psi)
    dpsi = T.MPS.from_host(ctx, psi.A)
    v = T.VidalMPS.from_mps(dpsi)
    ref = O.vidal_from_mps(psi)
    assert v.bondsizes() == ref.bondsizes()
    for a, b in zip(v.lam_host, ref.lam):
        assert np.allclose(a, b, rtol=1e-12, atol=1e-14)
    assert _rel(_host_vidal(v).to_mps().to_dense(), psi.to_dense()) < 1e-12
    gates = [O.haar_unitary(rng, 4) for _ in range(4)]
    bonds = [1, 3, 5, 7]
    ks = v.apply_layer_(gates, bonds, maxdim=6, threshold=1e-10)
    for U, a in zip(gates, bonds):
        O.vidal_gate(ref, U, a, maxdim=6, threshold=1e-10)
    assert ks == [ref.bondsizes()[a] for a in bonds]
    assert _rel(_host_vidal(v).to_mps().to_dense(), ref.to_mps().to_dense()) < 1e-10
    v.close()


def test_vidal_layer_rejects_overlapping_bonds(ctx):
    import tenet_hip as T

    v = T.VidalMPS.product_state(ctx, [[1.0, 0.0]] * 6)
    g = O.haar_unitary(np.random.default_rng(0), 4)
    with pytest.raises(ValueError):
        v.apply_layer_([g, g], [1, 2])


def test_vidal_sweep_with_captured_layers_equals_eager(ctx):
    """n = 14, depth 18, maxdim 8: the bonds saturate after a few layers and
    the later layers of each parity replay a captured graph
    (apply_layer_graph_); sites, Lambdas and bond sizes equal the eager
    sweep's bit for bit, and the state equals the oracle's Vidal sweep."""
    import tenet_hip as T
    from tenet_hip.tebd import upload_gates
    from tenet_hip.vidal import VidalMPS

    n, depth, chi, thr = 14, 18, 8, 1e-10
    rng = np.random.default_rng(77)
    layers = [list(b) if i % 2 == 0 else list(reversed(b)) for i, b in enumerate(O.brickwork_layers(n, depth))]
    gates = _gates(rng, layers)
    views = upload_gates(ctx, gates)
    eager = T.tebd_vidal_sweep(ctx, n, layers, views, chi, thr)
    used = []
    orig = VidalMPS.apply_layer_graph_

    def counting(self, *a, **k):
        used.append(1)
        return orig(self, *a, **k)

    VidalMPS.apply_layer_graph_ = counting
    try:
        graph = T.tebd_vidal_sweep(ctx, n, layers, views, chi, thr, graphs=True)
    finally:
        VidalMPS.apply_layer_graph_ = orig
    print(f"captured layers replayed: {len(used)} of {depth}", flush=True)
    assert len(used) >= 6
    assert graph.bondsizes() == eager.bondsizes()
    for a, b in zip(graph.B, eager.B):
        assert np.array_equal(a.to_host(), b.to_host())
    for a, b in zip(graph.lam_host, eager.lam_host):
        assert np.array_equal(a, b)
    ref = O.vidal_tebd_run(n, layers, gates, chi, thr)
    hv = _host_vidal(graph)
    assert hv.bondsizes() == ref.bondsizes()
    assert _rel(hv.to_mps().to_dense(), ref.to_mps().to_dense()) < 1e-10
    graph.close()
    eager.close()


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-13), (np.complex64, 3e-6)])
@pytest.mark.parametrize("d", [2, 3, 4])
def test_grouped_gate_against_oracle(ctx, dtype, tol, d):
    """tenet_grouped_gate (the gate contraction of simple_update, batched and
    fused with the Lambda scaling) against the oracle's contraction and
    Hadamard product: 37 problems (two launch batches), ragged bonds
    including 1, every third without Lambda, every other X computed in place
    over theta."""
    import ctypes as C

    import tenet_hip as T
    from tenet_hip import _lib as L

    rng = np.random.default_rng(40 + d)
    count = 37
    shapes = [(int(rng.integers(1, 30)), int(rng.integers(1, 30))) for _ in range(count)]
    shapes[0] = (1, 1)
    ths = [O.complex_gaussian(rng, (l, d, d, r)).astype(dtype) for l, r in shapes]
    gs = [O.complex_gaussian(rng, (d, d, d, d)).astype(dtype) for _ in range(count)]
    lams = [O.complex_gaussian(rng, (l,)).astype(dtype) if i % 3 else None for i, (l, r) in enumerate(shapes)]
    th_dev = [T.DeviceArray.from_host(ctx, x, dtype) for x in ths]
    g_dev = [T.DeviceArray.from_host(ctx, x, dtype) for x in gs]
    lam_dev = [T.DeviceArray.from_host(ctx, x, dtype) if x is not None else None for x in lams]
    out_dev = [T.DeviceArray(ctx, x.shape, dtype) for x in ths]
    x_dev = []
    for i, x in enumerate(ths):
        if lams[i] is None:
            x_dev.append(None)
        elif i % 2:
            x_dev.append(th_dev[i])  # in place
        else:
            x_dev.append(T.DeviceArray(ctx, x.shape, dtype))
    Pp = C.c_void_p
    chil = np.array([s[0] for s in shapes], dtype=np.int64)
    chir = np.array([s[1] for s in shapes], dtype=np.int64)
    L.check(L.lib.tenet_grouped_gate(
        ctx.h, None, L.C128 if dtype == np.complex128 else L.C64, count, d,
        chil.ctypes.data_as(C.POINTER(C.c_int64)), chir.ctypes.data_as(C.POINTER(C.c_int64)),
        (Pp * count)(*[x.ptr for x in th_dev]), (Pp * count)(*[x.ptr for x in g_dev]),
        (Pp * count)(*[x.ptr if x is not None else None for x in lam_dev]),
        (Pp * count)(*[x.ptr for x in out_dev]), (Pp * count)(*[x.ptr if x is not None else None for x in x_dev])),
        "grouped_gate")
    for i in range(count):
        ref = O.binary_einsum(O.OTensor(gs[i].astype(np.complex128), ["p", "q", "x", "y"]),
                              O.OTensor(ths[i].astype(np.complex128), ["l", "x", "y", "r"]), dims=["x", "y"])
        ref = np.transpose(ref.data, (2, 0, 1, 3))  # (p, q, l, r) -> (l, p, q, r)
        assert _rel(out_dev[i].to_host(), ref) < tol, i
        if lams[i] is not None:
            xref = O.binary_einsum(O.OTensor(lams[i].astype(np.complex128), ["l"]),
                                   O.OTensor(ref, ["l", "p", "q", "r"]), dims=[]).data
            assert _rel(x_dev[i].to_host(), xref) < tol, i


def test_grouped_gate_rejects_overlap(ctx):
    """theta_out may not overlap theta; X and Lambda come together."""
    import ctypes as C

    import tenet_hip as T
    from tenet_hip import _lib as L

    th = T.DeviceArray(ctx, (4, 2, 2, 4), np.complex128).zero_()
    g = T.DeviceArray(ctx, (2, 2, 2, 2), np.complex128).zero_()
    lam = T.DeviceArray(ctx, (4,), np.complex128).zero_()
    Pp = C.c_void_p
    one = np.array([4], dtype=np.int64)
    pi = one.ctypes.data_as(C.POINTER(C.c_int64))
    st = L.lib.tenet_grouped_gate(ctx.h, None, L.C128, 1, 2, pi, pi, (Pp * 1)(th.ptr), (Pp * 1)(g.ptr),
                                  (Pp * 1)(lam.ptr), (Pp * 1)(th.ptr), (Pp * 1)(th.ptr))
    assert st == L.TENET_ERR_ARG
    out = T.DeviceArray(ctx, (4, 2, 2, 4), np.complex128)
    st = L.lib.tenet_grouped_gate(ctx.h, None, L.C128, 1, 2, pi, pi, (Pp * 1)(th.ptr), (Pp * 1)(g.ptr),
                                  (Pp * 1)(lam.ptr), (Pp * 1)(out.ptr), (Pp * 1)(None))
    assert st == L.TENET_ERR_ARG
