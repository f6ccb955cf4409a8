"""Pins for the CPU oracle (oracle/tenet_oracle.py) against the reference's own
tests, closed forms and brute force.  CPU only."""
import itertools

import numpy as np
import pytest

from oracle import tenet_oracle as O


def _brute_einsum(a, b, dims):
    shared = [l for l in a.labels if l in b.labels]
    out = [l for l in a.labels if l not in dims] + [l for l in b.labels if l not in a.labels]
    size = {}
    for t in (a, b):
        for l, s in zip(t.labels, t.data.shape):
            size[l] = s
    res = np.zeros([size[l] for l in out], dtype=np.result_type(a.data, b.data))
    alll = list(dict.fromkeys(a.labels + b.labels))
    for idx in itertools.product(*[range(size[l]) for l in alll]):
        env = dict(zip(alll, idx))
        res[tuple(env[l] for l in out)] += a.data[tuple(env[l] for l in a.labels)] * b.data[tuple(env[l] for l in b.labels)]
    return O.OTensor(res, out)


@pytest.mark.parametrize("seed", range(6))
def test_binary_einsum_vs_brute_force(seed):
    rng = np.random.default_rng(seed)
    labels = list("abcdefg")
    ra, rb = rng.integers(1, 4, size=2)
    la = list(rng.choice(labels, ra, replace=False))
    lb = list(rng.choice(labels, rb, replace=False))
    size = {l: int(rng.integers(1, 4)) for l in labels}
    a = O.OTensor(rng.integers(-3, 4, [size[l] for l in la]).astype(float), la)
    b = O.OTensor(rng.integers(-3, 4, [size[l] for l in lb]).astype(float), lb)
    shared = [l for l in la if l in lb]
    for dims in (shared, shared[:1], []):
        got = O.binary_einsum(a, b, dims)
        ref = _brute_einsum(a, b, dims)
        assert got.labels == ref.labels
        assert np.array_equal(got.data, ref.data)  # integer-valued: exact


def test_binary_einsum_hadamard_closed_form():
    s = O.OTensor(np.array([2.0, 3.0]), ["k"])
    v = O.OTensor(np.arange(6.0).reshape(2, 3), ["k", "r"])
    out = O.binary_einsum(s, v, dims=[])
    assert out.labels == ["k", "r"]
    assert np.array_equal(out.data, np.array([[0, 2, 4], [9, 12, 15.0]]))


@pytest.mark.parametrize("m,n", [(6, 4), (4, 6), (5, 5), (1, 3), (8, 1)])
def test_qr_invariants(m, n):
    rng = np.random.default_rng(m * 10 + n)
    A = O.OTensor(O.complex_gaussian(rng, (m, n)), ["i", "j"])
    Q, R = O.tensor_qr_thin(A, ["i"], ["j"], "v")
    k = min(m, n)
    assert Q.data.shape == (m, k) and R.data.shape == (k, n)
    assert np.allclose(Q.data.conj().T @ Q.data, np.eye(k), atol=1e-13)
    assert np.allclose(Q.data @ R.data, A.data, atol=1e-13)
    assert np.allclose(np.tril(R.data, -1), 0)


def test_svd_closed_forms():
    # rank-1 matrix: one nonzero singular value = |u| |v|
    u = np.array([1.0, 2.0, 2.0])
    v = np.array([3.0, 4.0])
    A = O.OTensor(np.outer(u, v).astype(np.complex128), ["i", "j"])
    U, S, V = O.tensor_svd_thin(A, ["i"], "s")
    assert np.allclose(S.data, [15.0, 0.0], atol=1e-13)
    # diagonal matrix: singular values are the sorted magnitudes
    D = O.OTensor(np.diag([1.0, -5.0, 3.0j]), ["i", "j"])
    U, S, V = O.tensor_svd_thin(D, ["i"], "s", ["j"])
    assert np.allclose(S.data, [5.0, 3.0, 1.0])
    rec = np.einsum("is,s,sj->ij", U.data, S.data, V.data)
    assert np.allclose(rec, D.data)


def test_truncation_rule():
    s = np.array([5.0, 3.0, 1e-3, 1e-12])
    assert O.truncation_keep(s, maxdim=2) == 2
    assert O.truncation_keep(s, threshold=1e-10) == 3
    assert O.truncation_keep(s, maxdim=10, threshold=2.0) == 2
    assert O.truncation_keep(s, maxdim=1, threshold=10.0) == 0


def _int_mps():
    """The 5-site integer MPS of test/unit/mps.jl:63-69 (column-major reshape)."""
    def colmajor(n, shape):
        return np.arange(1, n + 1, dtype=float).reshape(shape, order="F")
    arrays = [colmajor(16, (4, 4)), colmajor(64, (4, 4, 4)), colmajor(64, (4, 4, 4)),
              colmajor(64, (4, 4, 4)), colmajor(16, (4, 4))]
    return O.OMPS.from_tenet_arrays(arrays)


@pytest.mark.parametrize("lo,hi", [(2, 2), (1, 2), (0, 0), (4, 4)])
def test_canonize_mixed_pattern_and_state(lo, hi):
    """mps.jl:62-443: isometry pattern + contract(psi_c) ~ contract(psi)."""
    psi = _int_mps()
    psi.form = ("none",)
    ref = psi.to_dense()
    p = psi.copy()
    O.canonize_mixed(p, lo, hi)
    assert O.checkform(p)
    assert p.form == ("mixed", lo, hi)
    for i in range(p.n):
        if i < lo:
            assert O.is_left_isometry(p.A[i])
        if i > hi:
            assert O.is_right_isometry(p.A[i])
    if lo == hi and 0 < lo < p.n - 1:
        assert not O.is_left_isometry(p.A[lo])
        assert not O.is_right_isometry(p.A[lo])
    got = p.to_dense()
    assert np.linalg.norm(got - ref) <= 1.49e-8 * np.linalg.norm(ref)


def test_canonize_sweep_and_bond_transitions():
    psi = _int_mps()
    ref = psi.to_dense()
    O.canonize_mixed(psi, 4, 4)
    O.canonize_mixed(psi, 2, 2)
    assert O.checkform(psi)
    O.canonize_bond(psi, 1)  # Mixed -> Bond(2-3)
    assert psi.form == ("bond", 1) and O.checkform(psi)
    O.canonize_bond(psi, 3)  # Bond -> Bond to the right
    assert psi.form == ("bond", 3) and O.checkform(psi)
    O.canonize_bond(psi, 0)  # Bond -> Bond to the left
    assert psi.form == ("bond", 0) and O.checkform(psi)
    O.canonize_mixed(psi, 3, 3)  # Bond -> Mixed
    assert psi.form == ("mixed", 3, 3) and O.checkform(psi)
    got = psi.to_dense()
    assert np.linalg.norm(got - ref) <= 1.49e-8 * np.linalg.norm(ref)


def _compress_mps():
    """mps.jl:465-472: first tensor ones(4,4) has rank 1."""
    psi = _int_mps()
    psi.A[0] = np.ones((1, 4, 4))
    return psi


def test_compress_single_bond_maxdim_and_threshold():
    """mps.jl:465-495."""
    for kw in ({"maxdim": 1}, {"threshold": 0.1}):
        psi = _compress_mps()
        ref = psi.to_dense()
        O.compress_bond(psi, 0, **kw)
        assert psi.A[0].shape[2] == 1
        O.absorb_lambda_into_right(psi)
        got = psi.to_dense()
        assert np.linalg.norm(got - ref) <= 1.49e-8 * np.linalg.norm(ref)


def test_compress_all_bonds():
    """mps.jl:497-522."""
    psi = _compress_mps()
    O.compress(psi, maxdim=1)
    assert all(b == 1 for b in psi.bondsizes())
    psi = _compress_mps()
    ref = psi.to_dense()
    thr = O.schmidt_values(psi, 0)[1] * 1.1
    O.compress(psi, threshold=thr)
    assert psi.bondsizes()[0] == 1
    O.absorb_lambda_into_right(psi)
    got = psi.to_dense()
    assert np.linalg.norm(got - ref) <= 1.49e-8 * np.linalg.norm(ref)


def test_bondsizes_truncate():
    """mps.jl:524-531 with a Tenet-shaped random MPS (right isometries via LQ)."""
    rng = np.random.default_rng(1)
    n = 8
    psi = O.random_gaussian_mps(rng, n, 2, 128, cap_edges=True)
    assert psi.bondsizes() == [2, 4, 8, 16, 8, 4, 2]
    O.compress(psi, maxdim=10)
    assert max(psi.bondsizes()) == 10


def test_apply_x_and_cx():
    """test/unit/simple_update.jl:4-43."""
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    p = O.product_state_mps([[1, 0], [0, 1]])
    O.apply_one_site(p, X, 0)
    assert np.array_equal(p.A[0].reshape(-1), [0, 1])
    O.apply_one_site(p, X, 1)
    assert np.array_equal(p.A[1].reshape(-1), [1, 0])

    CX = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=complex)
    p = O.product_state_mps([[0, 1], [0, 1]])
    O.simple_update_mps(p, CX, 0)
    a, b = p.A[0], p.A[1]  # (1, d, k), (k, d, 1)
    assert np.allclose(np.abs(a[0, :, 0]), [0, 1])
    assert np.allclose(np.abs(b[0, :, 0]), [1, 0])
    if a.shape[2] > 1:
        assert np.allclose(np.abs(a[0, :, 1]), 0)
        assert np.allclose(np.abs(b[1, :, 0]), 0)


def _dense_apply(vec, U4, a, n):
    v = vec.reshape([2] * n)
    v = np.moveaxis(v, [a, a + 1], [0, 1]).reshape(4, -1)
    v = (U4 @ v).reshape([2, 2] + [2] * (n - 2))
    return np.moveaxis(v, [0, 1], [a, a + 1]).reshape(-1)


def test_tebd_vs_dense_statevector():
    n, depth = 8, 6
    sched, gates = O.tebd_gates(7, n, depth)
    psi = O.tebd_run(n, sched, gates, maxdim=None, threshold=1e-14)
    vec = np.zeros(2 ** n, complex)
    vec[0] = 1
    for a, U in zip(sched, gates):
        vec = _dense_apply(vec, U, a, n)
    got = psi.to_dense()
    assert np.linalg.norm(got - vec) < 1e-12
    rng = np.random.default_rng(3)
    for _ in range(5):
        bits = rng.integers(0, 2, n)
        idx = int("".join(map(str, bits)), 2)
        assert abs(O.amplitude(bits, psi) - vec[idx]) < 1e-12
    assert abs(O.norm(psi.copy()) - 1.0) < 1e-12


def test_overlap_norm_expectation_vs_dense():
    rng = np.random.default_rng(5)
    psi = O.random_gaussian_mps(rng, 6, 2, 4)
    vec = psi.to_dense()
    assert abs(O.overlap_mps(psi, psi) - np.vdot(vec, vec)) < 1e-10 * np.vdot(vec, vec).real
    val, lg = O.overlap_mps_logscaled(psi, psi)
    assert abs(val * np.exp(lg) - np.vdot(vec, vec)) < 1e-10 * np.vdot(vec, vec).real
    nrm = O.norm(psi.copy())
    assert abs(nrm ** 2 - np.vdot(vec, vec).real) < 1e-10 * nrm ** 2
    # <x|psi> without conjugation of x
    bits = [1, 0, 1, 1, 0, 0]
    idx = int("".join(map(str, bits)), 2)
    assert abs(O.amplitude(bits, psi) - vec[idx]) < 1e-12 * abs(vec).max()
    m, lg = O.amplitude_logscaled(bits, psi)
    assert abs(m * np.exp(lg) - vec[idx]) < 1e-12 * abs(vec).max()
    # Pauli expectation against the dense operator
    start, ps = 2, "XZY"
    op = np.array([[1.0]])
    for i in range(6):
        P = O.PAULI[ps[i - start]] if start <= i < start + 3 else O.PAULI["I"]
        op = np.kron(op, P)
    ref = np.vdot(vec, op @ vec)
    assert abs(O.expectation_pauli(psi, start, ps) - ref) < 1e-10 * abs(np.vdot(vec, vec))


def test_ising_ed_energy_via_heff():
    """test/unit/dmrg.jl:38 pins the Ising n=10 ground energy per site
    (-1.2381489999654788); ED of the MPO reproduces it, and the two-site
    H_eff built from the same W tensors is Hermitian with <theta|H_eff|theta>
    equal to <psi|H|psi> for a random state."""
    n = 10
    W = O.ising_mpo_W(1.0, 1.0)
    lv = np.zeros(3); lv[2] = 1
    rv = np.zeros(3); rv[0] = 1
    H = O.mpo_dense([W] * n, lv, rv)
    assert np.allclose(H, H.conj().T)
    e0 = np.linalg.eigvalsh(H)[0]
    assert abs(e0 / n - (-1.2381489999654788)) < 1e-10
    # two-site H_eff on a 4-site chain whose outer sites are unitaries: the
    # embedding is unitary, so H_eff (16x16) has exactly H's spectrum
    rng = np.random.default_rng(2)
    A1 = O.haar_unitary(rng, 2)  # (phys, bond)
    A4 = O.haar_unitary(rng, 2)  # (bond, phys)
    lW = np.einsum("l,lroi->roi", lv, W)  # (w, o, i)
    rW = np.einsum("lroi,r->loi", W, rv)  # (w, o, i)
    Lenv = np.einsum("oa,woi,ib->awb", A1.conj(), lW, A1)  # (a, w, a')
    Renv = np.einsum("do,woi,ci->dwc", A4.conj(), rW, A4)  # (d, w, c')
    H2 = O.mpo_dense([W] * 4, lv, rv)
    M = np.zeros((16, 16), complex)
    for col in range(16):
        e = np.zeros(16, complex); e[col] = 1
        M[:, col] = O.heff_apply(Lenv, W, W, Renv, e.reshape(2, 2, 2, 2)).reshape(-1)
    assert np.allclose(M, M.conj().T, atol=1e-12)
    assert np.allclose(np.linalg.eigvalsh(M), np.linalg.eigvalsh(H2), atol=1e-12)


def test_xxz_mpo_hermitian_and_spectrum():
    n = 6
    W = O.xxz_mpo_W(1.0, 0.5)
    lv = np.zeros(5); lv[4] = 1
    rv = np.zeros(5); rv[0] = 1
    H = O.mpo_dense([W] * n, lv, rv)
    assert np.allclose(H, H.conj().T)
    # closed form for the 2-site XXZ: H = (S+S- + S-S+)/2 + Delta Sz Sz
    H2 = O.mpo_dense([W] * 2, lv, rv)
    sz = np.diag([0.5, -0.5])
    sp = np.array([[0, 1.0], [0, 0]])
    ref = 0.5 * (np.kron(sp, sp.T) + np.kron(sp.T, sp)) + 0.5 * np.kron(sz, sz)
    assert np.allclose(H2, ref)


def test_haar_unitary():
    rng = np.random.default_rng(0)
    U = O.haar_unitary(rng, 4)
    assert np.allclose(U.conj().T @ U, np.eye(4), atol=1e-13)


# ---------------------------------------------------------------- synthetic inputs
def test_synth_gaussian_moments_and_determinism():
    """The host mirror of the device generator: deterministic per (seed, i),
    offset slices agree with the full stream, and the first two moments are
    those of a standard complex Gaussian."""
    from oracle import synth

    z = synth.gaussian(7, 200000)
    assert np.array_equal(synth.gaussian(7, 1000, offset=5000), z[5000:6000])
    assert not np.array_equal(synth.gaussian(8, 1000), z[:1000])
    assert abs(z.mean()) < 0.01
    assert abs(np.mean(np.abs(z) ** 2) - 1.0) < 0.01
    assert abs(np.mean(z * z)) < 0.01  # circular: E[z^2] = 0
    u1, u2 = synth.uniforms(7, 0, 100000)
    assert 0.0 <= u1.min() and u1.max() < 1.0 and abs(u2.mean() - 0.5) < 0.01


def test_synth_splitmix_known_values():
    """splitmix64's finaliser on the published first outputs of the
    generator seeded with 0: mix64(k * golden) for k = 1, 2, 3."""
    from oracle import synth

    k = np.arange(1, 4, dtype=np.uint64)
    with np.errstate(over="ignore"):
        got = synth._mix64(k * synth._G)
    assert [int(v) for v in got] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


# ---- Vidal-gauge TEBD (restated algorithm; the reference's update is a stub)
def _dense_brickwork(n, layers, gates):
    psi = np.zeros(2 ** n, complex)
    psi[0] = 1.0
    g = 0
    for bonds in layers:
        for a in bonds:
            psi = np.kron(np.kron(np.eye(2 ** a), gates[g]), np.eye(2 ** (n - a - 2))) @ psi
            g += 1
    return psi


def test_vidal_tebd_untruncated_equals_dense_and_sequential():
    """Without truncation the Vidal-gauge TEBD is exact: its state equals the
    dense state vector and the sequential simple_update! TEBD's; every Lambda
    equals the Schmidt spectrum of that bond; every B is right-isometric."""
    rng = np.random.default_rng(11)
    n, depth = 8, 6
    layers = O.brickwork_layers(n, depth)
    gates = np.stack([O.haar_unitary(rng, 4) for _ in range(sum(map(len, layers)))])
    v = O.vidal_tebd_run(n, layers, gates, None, None)
    dense = _dense_brickwork(n, layers, gates)
    assert np.allclose(v.to_mps().to_dense(), dense, atol=1e-13)
    seq = O.tebd_run(n, [a for b in layers for a in b], gates, None, None)
    assert np.allclose(seq.to_dense(), dense, atol=1e-13)
    for i in range(n - 1):
        sv = np.sort(O.schmidt_values(seq.copy(), i))[::-1]
        assert np.allclose(sv[:len(v.lam[i])], v.lam[i], atol=1e-12)
        assert np.allclose(sv[len(v.lam[i]):], 0.0, atol=1e-12)
    assert all(O.is_right_isometry(b, 1e-12) for b in v.B)


def test_vidal_truncation_rule_and_from_mps():
    """A layer keeps min(maxdim, #S > threshold) per bond (compress.jl:46-59);
    vidal_from_mps reproduces the state and its Schmidt spectra."""
    rng = np.random.default_rng(12)
    n = 8
    psi = O.random_gaussian_mps(rng, n, 2, 6)
    O.normalize(psi)
    v = O.vidal_from_mps(psi)
    assert np.allclose(v.to_mps().to_dense(), psi.to_dense(), atol=1e-12)
    for i in range(n - 1):
        sv = np.sort(O.schmidt_values(psi.copy(), i))[::-1]
        assert np.allclose(sv[:len(v.lam[i])], v.lam[i], atol=1e-12)
    gates = [O.haar_unitary(rng, 4) for _ in range(4)]
    bonds = [0, 2, 4, 6]
    for U, a in zip(gates, bonds):
        s = O.vidal_gate(v, U, a, maxdim=3, threshold=1e-3)
        assert len(s) == min(3, int(np.count_nonzero(s > 1e-3))) and v.bondsizes()[a] == len(s)


# ---- MPO application (evolve!(::MPS, ::AbstractMPO), evolve.jl:19-138)
def test_evolve_mpo_direct_and_zipup_equal_dense_product():
    rng = np.random.default_rng(31)
    n = 7
    psi = O.random_gaussian_mps(rng, n, 2, 5)
    Ws = O.mpo_sites(O.xxz_mpo_W(1.0, 0.5), n, np.eye(5)[4], np.eye(5)[0])
    H = O.mpo_sites_dense(Ws)
    assert np.allclose(H, O.mpo_dense([O.xxz_mpo_W(1.0, 0.5)] * n, np.eye(5)[4], np.eye(5)[0]))
    ref = H @ psi.to_dense()
    d = O.evolve_mpo_direct(psi, Ws)
    assert d.bondsizes() == [5 * x for x in psi.bondsizes()] and d.form == ("mixed", 0, n - 1)
    assert np.allclose(d.to_dense(), ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    z = O.evolve_mpo_zipup(psi, Ws, None, None)
    assert np.allclose(z.to_dense(), ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    assert all(O.is_left_isometry(a, 1e-12) for a in z.A[:-1])
    # truncation: each cut keeps min(maxdim, #S > threshold)
    zt = O.evolve_mpo_zipup(psi, Ws, 3, 1e-8)
    assert max(zt.bondsizes()) <= 3
    assert np.linalg.norm(zt.to_dense() - ref) < np.linalg.norm(ref)


# ---- sampling (sample.jl:20-72)
def boolean_mps_0010_1101():
    """The reference's test MPS (test/unit/mps.jl:445-463): |0010> + |1101>."""
    s1 = np.array([[1.0, 0.0], [0.0, 1.0]])  # (r, o)
    s2 = np.zeros((2, 2, 2))  # (l, r, o)
    s2[0, 0, 0] = s2[1, 1, 1] = 1.0
    s3 = np.zeros((2, 2, 2))
    s3[0, 1, 1] = s3[1, 0, 0] = 1.0
    s4 = np.array([[0.0, 1.0], [1.0, 0.0]]) / np.sqrt(2.0)  # (l, o)
    return O.OMPS.from_tenet_arrays([s1, s2, s3, s4], order=("l", "r", "o"))


def test_sample_boolean_mps_support():
    psi = boolean_mps_0010_1101()
    dense = psi.to_dense()
    assert np.isclose(abs(dense[0b0010]), 1 / np.sqrt(2)) and np.isclose(abs(dense[0b1101]), 1 / np.sqrt(2))
    smp = O.sample(psi, 128, seed=5)
    rows = {tuple(r) for r in smp}
    assert rows <= {(0, 0, 1, 0), (1, 1, 0, 1)} and len(rows) == 2


def test_sample_frequencies_match_born_rule():
    rng = np.random.default_rng(17)
    n = 5
    psi = O.random_gaussian_mps(rng, n, 2, 3)
    O.normalize(psi)
    prob = np.abs(psi.to_dense()) ** 2
    prob /= prob.sum()
    N = 20000
    smp = O.sample(psi, N, seed=9)
    idx = smp @ (1 << np.arange(n - 1, -1, -1))
    counts = np.bincount(idx, minlength=2 ** n)
    exp = N * prob
    m = exp > 5
    chi2 = np.sum((counts[m] - exp[m]) ** 2 / exp[m])
    dof = int(m.sum()) - 1
    assert chi2 < dof + 5 * np.sqrt(2 * dof)


# ---------------------------------------------------------------- Muscle.simple_update (tensor level)
def test_simple_update_tensor_cx_known_answer():
    """The CX case of test/unit/simple_update.jl:19-43 at the tensor level:
    |1>|1> under CX is |1>|0> (site a the control, most significant digit)."""
    A = O.OTensor(np.array([[0.0], [1.0]], dtype=complex), ["pa", "b"])   # (ia, bond)
    B = O.OTensor(np.array([[0.0, 1.0]], dtype=complex), ["b", "pb"])     # (bond, ib)
    CX = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=complex)
    G = O.OTensor(CX.reshape(2, 2, 2, 2), ["oa", "ob", "pa", "pb"])
    nA, nB, s = O.simple_update(A, "pa", B, "pb", "b", G, "oa", "ob")
    assert nA.labels == ["pa", "b"] and nB.labels == ["b", "pb"]
    theta = np.einsum("ik,kj->ij", nA.data, nB.data)
    assert np.allclose(theta, [[0, 0], [1, 0]], atol=1e-15)
    assert np.allclose(s[0], 1.0) and np.all(s[1:] < 1e-15)


@pytest.mark.parametrize("maxdim", [None, 3])
def test_simple_update_tensor_equals_dense_and_best_rank_k(maxdim):
    """theta~ = G.(A.B) brute-forced with loops; the factors multiply back to
    theta~ (no cut) or to its best rank-k approximation (Eckart-Young), and
    AbsorbEqually puts sqrt(s) on both sides: nA^H nA = diag(s) = nB nB^H."""
    rng = np.random.default_rng(5)
    chil, d, chi, chir = 3, 2, 4, 3
    a = O.complex_gaussian(rng, (chil, d, chi))
    b = O.complex_gaussian(rng, (chi, d, chir))
    g = O.haar_unitary(rng, 4).reshape(d, d, d, d)  # (oa, ob, ia, ib)
    dense = np.zeros((chil, d, d, chir), dtype=complex)
    for l, oa, ob, r, ia, ib, c in itertools.product(range(chil), range(d), range(d), range(chir), range(d),
                                                     range(d), range(chi)):
        dense[l, oa, ob, r] += g[oa, ob, ia, ib] * a[l, ia, c] * b[c, ib, r]
    A = O.OTensor(a, ["l", "pa", "b"])
    B = O.OTensor(b, ["b", "pb", "r"])
    G = O.OTensor(g, ["oa", "ob", "pa", "pb"])
    nA, nB, s = O.simple_update(A, "pa", B, "pb", "b", G, "oa", "ob", maxdim=maxdim)
    k = len(s)
    assert k == (min(chil * d, d * chir) if maxdim is None else maxdim)
    mat = dense.reshape(chil * d, d * chir)
    u, sv, vh = np.linalg.svd(mat)
    best = (u[:, :k] * sv[:k]) @ vh[:k]
    got = np.einsum("lxk,kyr->lxyr", nA.data, nB.data).reshape(chil * d, d * chir)
    assert np.linalg.norm(got - best) < 1e-12 * np.linalg.norm(mat)
    na = nA.data.reshape(chil * d, k)
    nb = nB.data.reshape(k, d * chir)
    assert np.allclose(na.conj().T @ na, np.diag(s), atol=1e-12)
    assert np.allclose(nb @ nb.conj().T, np.diag(s), atol=1e-12)


@pytest.mark.parametrize("absorb", ["left", "right"])
def test_simple_update_tensor_absorb_sides(absorb):
    """AbsorbLeft keeps V^H an exact right isometry and AbsorbRight keeps U a
    left isometry; the product is unchanged by where s goes."""
    rng = np.random.default_rng(6)
    a = O.complex_gaussian(rng, (2, 2, 3))
    b = O.complex_gaussian(rng, (3, 2, 2))
    g = O.haar_unitary(rng, 4).reshape(2, 2, 2, 2)
    A, B = O.OTensor(a, ["l", "pa", "b"]), O.OTensor(b, ["b", "pb", "r"])
    G = O.OTensor(g, ["oa", "ob", "pa", "pb"])
    eA, eB, _ = O.simple_update(A, "pa", B, "pb", "b", G, "oa", "ob")
    nA, nB, s = O.simple_update(A, "pa", B, "pb", "b", G, "oa", "ob", absorb=absorb)
    k = len(s)
    na, nb = nA.data.reshape(4, k), nB.data.reshape(k, 4)
    if absorb == "left":
        assert np.allclose(nb @ nb.conj().T, np.eye(k), atol=1e-12)
    else:
        assert np.allclose(na.conj().T @ na, np.eye(k), atol=1e-12)
    assert np.allclose(na @ nb, eA.data.reshape(4, k) @ eB.data.reshape(k, 4), atol=1e-12)


# ---------------------------------------------------------------- overlap(ProductState, MPS)
def test_overlap_product_vs_dense_contraction():
    """<x|psi> for general (non-basis, complex) product vectors, without
    conjugating x (overlap.jl:22-32), against the dense state contracted with
    the Kronecker product of the vectors."""
    rng = np.random.default_rng(7)
    psi = O.random_gaussian_mps(rng, 6, 2, 4)
    xs = [O.complex_gaussian(rng, (2,)) for _ in range(6)]
    kron = xs[0]
    for x in xs[1:]:
        kron = np.kron(kron, x)
    ref = np.dot(kron, psi.to_dense())  # no conjugation
    assert abs(O.overlap_product(xs, psi) - ref) < 1e-12 * abs(ref) + 1e-14
    # basis vectors reduce to the amplitude of that bitstring
    bits = [1, 0, 0, 1, 1, 0]
    e = [np.eye(2)[b] for b in bits]
    idx = int("".join(map(str, bits)), 2)
    assert abs(O.overlap_product(e, psi) - psi.to_dense()[idx]) < 1e-12
