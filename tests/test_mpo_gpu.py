"""MPO application on the device (evolve!(::MPS, ::AbstractMPO),
src/Operations/evolve.jl:19-138) against the oracle restatement, which is
pinned by the dense MPO x state product (test_oracle.py).

Tolerances: complex128 throughout; dense states within 1e-12 relative for
the exact products, 1e-10 with truncation (the kept sets are identical; the
Jacobi and LAPACK factors agree to ~1e-13 per singular value).
"""
import numpy as np
import pytest

from oracle import tenet_oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _case(seed, n=7, chi=5):
    rng = np.random.default_rng(seed)
    psi = O.random_gaussian_mps(rng, n, 2, chi)
    Wb = O.xxz_mpo_W(1.0, 0.5)
    Ws = O.mpo_sites(Wb, n, np.eye(5)[4], np.eye(5)[0])
    return psi, Wb, Ws


def test_evolve_direct_equals_dense_product(ctx):
    import tenet_hip as T

    psi, Wb, Ws = _case(41)
    n = psi.n
    d = T.MPS.from_host(ctx, psi.A)
    op = T.MPO.from_bulk(ctx, Wb, n, np.eye(5)[4], np.eye(5)[0])
    assert all(np.allclose(a, b) for a, b in zip(op.to_host(), Ws))
    T.evolve_direct_(d, op)
    assert d.bondsizes() == [5 * x for x in psi.bondsizes()] and d.form == ("mixed", 0, n - 1)
    ref = O.mpo_sites_dense(Ws) @ psi.to_dense()
    assert _rel(O.OMPS(d.to_host()).to_dense(), ref) < 1e-12
    # <psi|H|psi> through the device transfer sweep
    e = d.overlap(T.MPS.from_host(ctx, psi.A))
    assert abs(e - np.vdot(psi.to_dense(), ref)) < 1e-11 * abs(np.vdot(psi.to_dense(), ref))


@pytest.mark.parametrize("maxdim,threshold", [(None, None), (6, 1e-8), (3, 1e-12)])
def test_evolve_zipup_matches_oracle(ctx, maxdim, threshold):
    import tenet_hip as T

    psi, Wb, Ws = _case(43)
    n = psi.n
    d = T.MPS.from_host(ctx, psi.A)
    op = T.MPO.from_host(ctx, Ws)
    T.evolve_zipup_(d, op, maxdim, threshold)
    ref = O.evolve_mpo_zipup(psi, Ws, maxdim, threshold)
    assert d.bondsizes() == ref.bondsizes()
    assert d.form == ("mixed", n - 1, n - 1)
    host = d.to_host()
    if threshold is not None:
        # without a threshold the cuts keep numerically zero singular values;
        # their U columns (X / sigma) are not completed to an orthonormal
        # basis as LAPACK's are, so the isometry is checked where the
        # threshold drops them (the state itself is checked in every case)
        assert all(O.is_left_isometry(a, 1e-11) for a in host[:-1])
    assert _rel(O.OMPS(host).to_dense(), ref.to_dense()) < (1e-12 if maxdim is None else 1e-10)
    if maxdim is None:
        assert _rel(O.OMPS(host).to_dense(), O.mpo_sites_dense(Ws) @ psi.to_dense()) < 1e-12


def test_evolve_direct_then_compress(ctx):
    """The usual pairing: exact product, then compress! back to a bond cap;
    bond sizes and the norm equal the oracle's compress of the oracle's
    product (compress! leaves the MPS BondCanonical with a Lambda)."""
    import tenet_hip as T

    psi, Wb, Ws = _case(47, n=8, chi=4)
    d = T.MPS.from_host(ctx, psi.A)
    T.evolve_direct_(d, T.MPO.from_host(ctx, Ws))
    d.compress_(maxdim=8, threshold=1e-10)
    ref = O.evolve_mpo_direct(psi, Ws)
    O.compress(ref, maxdim=8, threshold=1e-10)
    assert d.bondsizes() == ref.bondsizes()
    nref = O.norm(ref.copy())
    assert abs(d.norm() - nref) < 1e-10 * nref


# ---- canonize! / compress! on an MPO (canonize.jl:255, compress.jl:72) --------
# The twelve form transitions of test/unit/mpo.jl:81-462 on its integer MPO
# (reshape(1:64, 4, 4, 4) at the edges, reshape(1:256, 4, 4, 4, 4) inside,
# default order (:l, :r, :o, :i)); 0-based sites, bond i = sites (i, i+1).
def _ref_mpo_arrays():
    edge = np.arange(1, 65, dtype=np.float64).reshape(4, 4, 4, order="F")
    mid = np.arange(1, 257, dtype=np.float64).reshape(4, 4, 4, 4, order="F")
    return [edge, mid, mid, mid, edge]


def _fused(sites4):
    """(chi_l, o, i, chi_r) host sites as an oracle MPS over the fused
    column-major physical index o + d_o i."""
    return O.OMPS([s.reshape(s.shape[0], s.shape[1] * s.shape[2], s.shape[3], order="F") for s in sites4])


def _dense_absorbed(fused: "O.OMPS", form, lam):
    """Dense operator with a BondCanonical Lambda absorbed into the bond's
    left site (hadamard on its right bond, as the reference tests do)."""
    A = [x.copy() for x in fused.A]
    if form[0] == "bond":
        A[form[1]] = A[form[1]] * np.asarray(lam)[None, None, :]
    return O.OMPS(A).to_dense()


def _apply(obj, step, oracle):
    kind, a = step[0], step[1:]
    if kind == "none":
        obj.form = ("none",)
    elif kind == "mixed":
        (O.canonize_mixed(obj, *a) if oracle else obj.canonize_(*a))
    else:
        (O.canonize_bond(obj, *a) if oracle else obj.canonize_bond_(*a))


_TRANSITIONS = [
    (("none",), ("mixed", 2, 2)),
    (("none",), ("mixed", 1, 2)),
    (("mixed", 4, 4), ("mixed", 2, 2)),
    (("mixed", 0, 4), ("mixed", 2, 2)),
    (("mixed", 0, 4), ("mixed", 1, 2)),
    (("none",), ("bond", 1)),
    (("mixed", 4, 4), ("bond", 1)),
    (("mixed", 2, 4), ("bond", 0)),
    (("mixed", 2, 4), ("bond", 1)),
    (("bond", 3), ("bond", 1)),
    (("bond", 1), ("mixed", 2, 2)),
    (("bond", 3), ("mixed", 2, 2)),
]


@pytest.mark.parametrize("setup,target", _TRANSITIONS)
def test_mpo_canonize_transitions(ctx, setup, target):
    import tenet_hip as T

    op = T.MPO.from_arrays(ctx, _ref_mpo_arrays())
    dense0 = _fused(op.to_host_sites()).to_dense()
    _apply(op, setup, False)
    # the oracle restatement starts from the same (set-up) device state
    ref = _fused(op.to_host_sites())
    ref.form, ref.lam = op.form, None if op.lam is None else np.array(op.lam)
    assert _rel(_dense_absorbed(ref, op.form, op.lam), dense0) < 1e-12
    _apply(op, target, False)
    _apply(ref, target, True)
    assert op.form == ref.form == target
    assert op.bondsizes() == ref.bondsizes()
    fused = _fused(op.to_host_sites())
    lo, hi = (target[1], target[2]) if target[0] == "mixed" else (target[1], target[1] + 1)
    for i, a in enumerate(fused.A):
        if i < lo:
            assert O.is_left_isometry(a, 1e-10), i
        if i > hi:
            assert O.is_right_isometry(a, 1e-10), i
    # the centre is no isometry toward the outside (the reference's !isisometry checks)
    assert not O.is_right_isometry(fused.A[lo], 1e-6)
    assert not O.is_left_isometry(fused.A[hi], 1e-6)
    got = _dense_absorbed(fused, op.form, op.lam)
    assert _rel(got, dense0) < 1e-11
    assert _rel(got, _dense_absorbed(ref, ref.form, ref.lam)) < 1e-11
    if target[0] == "bond":
        assert np.allclose(op.lam, ref.lam, rtol=1e-11, atol=1e-11 * ref.lam[0])


def test_mpo_compress_xxz_sum(ctx):
    """compress!(::MPO) (compress.jl:72): the MPO of H + H (bond dimension
    10) compresses to the XXZ MPO's bond dimension 5 with the operator
    unchanged; then maxdim = 3 keeps the oracle's bond sizes and Schmidt
    values."""
    import tenet_hip as T

    n = 6
    Wb = O.xxz_mpo_W(1.0, 0.5)
    Ws = O.mpo_sites(Wb, n, np.eye(5)[4], np.eye(5)[0])
    # H + H as a direct sum of the bond spaces (bond dimension 10)
    two = []
    for j, W in enumerate(Ws):
        l, r, o, i = W.shape
        L2 = 1 if j == 0 else 2 * l
        R2 = 1 if j == n - 1 else 2 * r
        X = np.zeros((L2, R2, o, i), dtype=np.complex128)
        if j == 0:
            X[:, :r] = W
            X[:, r:] = W
        elif j == n - 1:
            X[:l] = W
            X[l:] = W
        else:
            X[:l, :r] = W
            X[l:, r:] = W
        two.append(X)
    op = T.MPO.from_host(ctx, two)
    dense0 = _fused(op.to_host_sites()).to_dense()
    assert _rel(dense0, 2.0 * _fused([np.transpose(W, (0, 2, 3, 1)) for W in Ws]).to_dense()) < 1e-14
    op.compress_(threshold=1e-10)
    assert max(op.bondsizes()) <= 5 and op.form == ("bond", n - 2)
    assert _rel(_dense_absorbed(_fused(op.to_host_sites()), op.form, op.lam), dense0) < 1e-12
    ref = _fused(T.MPO.from_host(ctx, two).to_host_sites())
    O.compress(ref, threshold=1e-10)
    assert op.bondsizes() == ref.bondsizes()
    op2 = T.MPO.from_host(ctx, two)
    op2.compress_(maxdim=3)
    ref2 = _fused(T.MPO.from_host(ctx, two).to_host_sites())
    O.compress(ref2, maxdim=3)
    assert op2.bondsizes() == ref2.bondsizes() == [min(3, b) for b in ref.bondsizes()]
    assert np.allclose(op2.lam, ref2.lam, rtol=1e-11)
