"""Sampling on the device (Tenet's sample, src/Operations/sample.jl:20-72)
against the oracle restatement, which draws from the same counter-based
uniform stream (uniform of sample s at site i = u1 of generator element
s * n + i), so complex128 samples agree value for value.  The oracle is
pinned in test_oracle.py by the reference's boolean-MPS test and by the
Born rule."""
import numpy as np
import pytest

from oracle import tenet_oracle as O

pytestmark = pytest.mark.gpu


def _boolean_mps():
    s1 = np.array([[1.0, 0.0], [0.0, 1.0]])
    s2 = np.zeros((2, 2, 2))
    s2[0, 0, 0] = s2[1, 1, 1] = 1.0
    s3 = np.zeros((2, 2, 2))
    s3[0, 1, 1] = s3[1, 0, 0] = 1.0
    s4 = np.array([[0.0, 1.0], [1.0, 0.0]]) / np.sqrt(2.0)
    return O.OMPS.from_tenet_arrays([s1, s2, s3, s4], order=("l", "r", "o"))


@pytest.mark.parametrize("n,chi,seed", [(10, 8, 123), (7, 16, 5), (16, 4, 77)])
def test_sample_matches_oracle_value_for_value(ctx, n, chi, seed):
    import tenet_hip as T

    rng = np.random.default_rng(seed)
    psi = O.random_gaussian_mps(rng, n, 2, chi)
    O.normalize(psi)
    d = T.MPS.from_host(ctx, psi.A)
    got = d.sample(256, seed=seed)
    want = O.sample(psi, 256, seed)
    assert np.array_equal(got, want)


def test_sample_boolean_mps(ctx):
    """The reference's test: every sample of |0010> + |1101> is one of the two."""
    import tenet_hip as T

    psi = _boolean_mps()
    d = T.MPS.from_host(ctx, psi.A)
    smp = d.sample(128, seed=1)
    rows = {tuple(r) for r in smp}
    assert rows <= {(0, 0, 1, 0), (1, 1, 0, 1)} and len(rows) == 2


def test_sample_born_rule_and_c64(ctx):
    import tenet_hip as T

    rng = np.random.default_rng(8)
    n = 6
    psi = O.random_gaussian_mps(rng, n, 2, 4)
    O.normalize(psi)
    prob = np.abs(psi.to_dense()) ** 2
    prob /= prob.sum()
    N = 40000
    smp = T.MPS.from_host(ctx, psi.A).sample(N, seed=3)
    idx = smp @ (1 << np.arange(n - 1, -1, -1))
    counts = np.bincount(idx, minlength=2 ** n)
    exp = N * prob
    m = exp > 5
    chi2 = np.sum((counts[m] - exp[m]) ** 2 / exp[m])
    dof = int(m.sum()) - 1
    assert chi2 < dof + 5 * np.sqrt(2 * dof)
    # complex64 state: same stream; decisions can differ only where a running
    # sum lies within float rounding of u * total
    d64 = T.MPS.from_host(ctx, [a.astype(np.complex64) for a in psi.A], dtype=np.complex64)
    got = d64.sample(512, seed=4)
    want = O.sample(O.OMPS([a.astype(np.complex64).astype(np.complex128) for a in psi.A]), 512, 4)
    assert np.mean(np.all(got == want, axis=1)) > 0.99
