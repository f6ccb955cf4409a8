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
