"""CPU oracle: a numpy/LAPACK restatement of Tenet.jl v0.10.3's contraction,
QR/SVD-truncation and transfer-matrix path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library or the
``tenet_hip`` host package) imports, calls or links this module; only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and there only as the checker / the timed CPU path.

The arithmetic that Tenet delegates to Muscle.jl (``binary_einsum``,
``tensor_qr_thin``, ``tensor_svd_thin``, ``simple_update``) is restated here
with numpy's einsum and LAPACK (Householder ``geqrf``/``ungqr`` for QR,
``gesdd`` for the SVD), the same library families Julia's LinearAlgebra
uses.  Muscle.jl (compat 0.3.9, Project.toml:38) is not vendored with the
reference; its published behaviour is restated from Tenet's call sites.

Parity pins (what checks each function; see tests/test_oracle.py):
  binary_einsum      brute-force loops on tiny inputs; integer inputs exact;
                     dims=[] Hadamard closed form
  tensor_qr_thin     Q^H Q = I, QR = A, R upper-triangular, k = min(m, n)
  tensor_svd_thin    singular values vs closed forms (rank-1, diagonal),
                     U diag(s) V = A, descending order
  truncation rule    the reference's compress tests (rank-1 exact cut,
                     threshold = 1.1 x second Schmidt value, bondsizes)
  canonize/compress  isometry pattern + full-state equality
                     (reference test/unit/mps.jl:62-531)
  simple_update      X / CX known answers (test/unit/simple_update.jl:4-43),
                     dense state-vector brute force for small n
  overlap/norm       dense state-vector brute force, norm^2 = overlap
  heff_apply         exact diagonalisation energy of the Ising chain
                     (test/unit/dmrg.jl:38) through a two-site H_eff
  sample             the reference's boolean-MPS test (test/unit/mps.jl:445-463:
                     every sample in {|0010>, |1101>}), frequencies against
                     |psi(x)|^2 (chi-square) on a dense-checkable state
  evolve_mpo_*       dense MPO x state product (direct exact; zip-up exact
                     without truncation), truncation rule, left isometries
  vidal_* (Vidal-gauge TEBD; the reference's Vidal update is a TODO stub,
                     so the published algorithm is restated)
                     dense state-vector brute force without truncation,
                     equality with the sequential simple_update! TEBD when
                     nothing is truncated, Lambda = the Schmidt values of
                     the MixedCanonical state, right-isometric B
Parity unpinned (no reference-side check exists without Julia): Muscle's
output label order of binary_einsum, Muscle.simple_update's internal
normalisation, raw Q/R/U/V gauges (only gauge-invariant quantities are
compared).
"""
from __future__ import annotations

import math
import string
from dataclasses import dataclass, field

import numpy as np

# --------------------------------------------------------------------------
# labelled tensors and Muscle's kernels
# --------------------------------------------------------------------------


@dataclass
class OTensor:
    """A dense array whose axes carry hashable labels (Muscle's ``Tensor``)."""

    data: np.ndarray
    labels: list = field(default_factory=list)

    def __post_init__(self):
        self.labels = list(self.labels)
        if self.data.ndim != len(self.labels):
            raise ValueError("rank/label mismatch")
        if len(set(self.labels)) != len(self.labels):
            raise ValueError("repeated label")

    def size(self, label):
        return self.data.shape[self.labels.index(label)]

    def permute(self, labels):
        return OTensor(np.transpose(self.data, [self.labels.index(l) for l in labels]), labels)

    def rename(self, mapping):
        return OTensor(self.data, [mapping.get(l, l) for l in self.labels])


def binary_einsum(a: OTensor, b: OTensor, dims=None, blas=False) -> OTensor:
    """Pairwise contraction over ``dims`` (default: all shared labels).

    Shared labels not in ``dims`` are kept once and multiplied elementwise
    (batch/Hadamard); ``dims=[]`` is how Tenet absorbs singular values
    (canonize.jl:44, evolve.jl:76).  Output labels are a's labels in order
    followed by b's labels not in a (Muscle's own order is unpinned; callers
    realign by label, DMRG.jl:320).  ``blas=True`` lets numpy route the
    product through BLAS (permute + gemm, as Muscle's CPU path does) instead
    of its plain summation loop; the checker uses the loop.
    """
    shared = [l for l in a.labels if l in b.labels]
    dims = shared if dims is None else list(dims)
    for d in dims:
        if d not in shared:
            raise ValueError(f"contracted label {d!r} not shared")
    out = [l for l in a.labels if l not in dims] + [l for l in b.labels if l not in a.labels]
    letters = {}
    for l in a.labels + b.labels:
        if l not in letters:
            letters[l] = string.ascii_letters[len(letters)]
    spec = "".join(letters[l] for l in a.labels) + "," + "".join(letters[l] for l in b.labels) \
        + "->" + "".join(letters[l] for l in out)
    return OTensor(np.einsum(spec, a.data, b.data, optimize="greedy" if blas else False), out)


def _matricize(A: OTensor, rows, cols):
    t = A.permute(list(rows) + list(cols))
    m = int(np.prod([t.data.shape[i] for i in range(len(rows))], dtype=np.int64))
    n = int(np.prod([t.data.shape[len(rows) + i] for i in range(len(cols))], dtype=np.int64))
    return t.data.reshape(m, n), t.data.shape


def tensor_qr_thin(A: OTensor, inds_q, inds_r, ind_virtual):
    """Thin Householder QR (Muscle.tensor_qr_thin at canonize.jl:58).

    Rows = ``inds_q``, cols = ``inds_r``; k = min(m, n), so a bond may shrink.
    Q carries (inds_q..., ind_virtual), R carries (ind_virtual, inds_r...).
    """
    mat, shp = _matricize(A, inds_q, inds_r)
    q, r = np.linalg.qr(mat, mode="reduced")
    k = q.shape[1]
    nq = len(inds_q)
    Q = OTensor(q.reshape(tuple(shp[:nq]) + (k,)), list(inds_q) + [ind_virtual])
    R = OTensor(r.reshape((k,) + tuple(shp[nq:])), [ind_virtual] + list(inds_r))
    return Q, R


def tensor_svd_thin(A: OTensor, inds_u, ind_s, inds_v=None):
    """Thin SVD (Muscle.tensor_svd_thin at canonize.jl:41,97; evolve.jl:62).

    ``inds_v`` defaults to the complement of ``inds_u`` (evolve.jl:62).
    Returns U (inds_u..., ind_s), s (ind_s,) real descending (DMRG.jl:340
    relies on the order), V (ind_s, inds_v...) with A = U diag(s) V.
    """
    if inds_v is None:
        inds_v = [l for l in A.labels if l not in inds_u]
    mat, shp = _matricize(A, inds_u, inds_v)
    u, s, vh = np.linalg.svd(mat, full_matrices=False)  # LAPACK gesdd
    k = s.shape[0]
    nu = len(inds_u)
    U = OTensor(u.reshape(tuple(shp[:nu]) + (k,)), list(inds_u) + [ind_s])
    S = OTensor(s.copy(), [ind_s])
    V = OTensor(vh.reshape((k,) + tuple(shp[nu:])), [ind_s] + list(inds_v))
    return U, S, V


def truncation_keep(s: np.ndarray, maxdim=None, threshold=None) -> int:
    """Number of singular values kept (compress.jl:46-59).

    First cut to ``min(maxdim, len(s))`` (compress.jl:47-51), then keep the
    entries with ``abs(s) > threshold`` (ABSOLUTE threshold, compress.jl:54-59).
    For a descending ``s`` the kept set is a prefix.
    """
    k = len(s) if maxdim is None else min(int(maxdim), len(s))
    if threshold is not None:
        k = int(np.count_nonzero(np.abs(s[:k]) > threshold))
    return k


def _slice_label(t: OTensor, label, k):
    ax = t.labels.index(label)
    sl = [slice(None)] * t.data.ndim
    sl[ax] = slice(0, k)
    return OTensor(t.data[tuple(sl)], t.labels)


def simple_update(A, ia, B, ib, bond, G, oa, ob, maxdim=None, threshold=None, absorb="equal"):
    """Muscle.simple_update with ``absorb=AbsorbEqually()`` (simple_update.jl:53-64).

    theta = A.B over ``bond``; theta.G maps the physical inputs (ia, ib) to the
    gate outputs (oa, ob), which are renamed back to (ia, ib) (simple_update.jl:66-69);
    SVD with rows = (A's other labels, ia) and cols = (B's other labels, ib);
    truncation rule of ``truncation_keep``; sqrt(s) absorbed into both sides.
    """
    theta = binary_einsum(A, B, dims=[bond])
    theta = binary_einsum(theta, G, dims=[ia, ib])
    theta = theta.rename({oa: ia, ob: ib})
    rows = [l for l in A.labels if l != bond]
    cols = [l for l in B.labels if l != bond]
    U, S, V = tensor_svd_thin(theta, rows, bond, cols)
    k = truncation_keep(S.data, maxdim, threshold)
    U, V = _slice_label(U, bond, k), _slice_label(V, bond, k)
    s = S.data[:k]
    if absorb == "equal":
        r = np.sqrt(s)
        U = binary_einsum(U, OTensor(r.astype(U.data.dtype), [bond]), dims=[])
        V = binary_einsum(OTensor(r.astype(V.data.dtype), [bond]), V, dims=[])
    elif absorb == "left":
        U = binary_einsum(U, OTensor(s.astype(U.data.dtype), [bond]), dims=[])
    elif absorb == "right":
        V = binary_einsum(OTensor(s.astype(V.data.dtype), [bond]), V, dims=[])
    newA = U.permute([l for l in A.labels])
    newB = V.permute([l for l in B.labels])
    return newA, newB, s


# --------------------------------------------------------------------------
# MPS in the (l, o, r) site layout
# --------------------------------------------------------------------------


class OMPS:
    """Open-boundary MPS.  Every site is stored as a 3-way array (chi_l, d, chi_r)
    with chi_l = 1 on site 1 and chi_r = 1 on site n (Tenet stores the edge
    sites as 2-way arrays, MPS.jl:63-69; the size-1 edge legs here are never
    treated as bonds).  ``form`` follows CanonicalForm.jl:10-87:
    ('mixed', lo, hi) with 0-based sites, ('bond', i) for bond i-(i+1), or
    ('none',).  ``lam`` holds the Lambda vector of a BondCanonical MPS
    (LambdaSite, VidalMPS.jl:8-17).
    """

    def __init__(self, arrays, form=None):
        self.A = [np.asarray(a) for a in arrays]
        n = len(self.A)
        self.form = form if form is not None else ("mixed", 0, n - 1)  # MPS.jl:18
        self.lam = None

    @property
    def n(self):
        return len(self.A)

    def copy(self):
        o = OMPS([a.copy() for a in self.A], self.form)
        o.lam = None if self.lam is None else self.lam.copy()
        return o

    def bondsizes(self):
        return [a.shape[2] for a in self.A[:-1]]

    @staticmethod
    def from_tenet_arrays(arrays, order=("l", "r", "o")):
        """Tenet's ``MPS(arrays; order)`` (MPS.jl:50-94): edge arrays are 2-way."""
        order = list(order)
        n = len(arrays)
        out = []
        for i, arr in enumerate(arrays):
            loc = [x for x in order if not ((i == 0 and x == "l") or (i == n - 1 and x == "r"))]
            a = np.asarray(arr)
            if i == 0:
                a = np.expand_dims(a, axis=0)
                loc = ["l"] + loc
            if i == n - 1:
                a = np.expand_dims(a, axis=-1)
                loc = loc + ["r"]
            a = np.transpose(a, [loc.index("l"), loc.index("o"), loc.index("r")])
            out.append(np.ascontiguousarray(a))
        return OMPS(out)

    def to_dense(self):
        """Full state vector, site 1 as the most significant digit."""
        v = self.A[0].reshape(self.A[0].shape[1], self.A[0].shape[2])
        for a in self.A[1:]:
            v = np.tensordot(v, a, axes=([v.ndim - 1], [0]))
            v = v.reshape(-1, a.shape[2])
        v = v.reshape(-1)
        if self.form[0] == "bond" and self.lam is not None:
            raise RuntimeError("absorb Lambda before densifying")
        return v


def product_state_mps(vectors):
    """``convert(MPS, ProductState(...))`` (MPS.jl:167-201): all bonds size 1."""
    return OMPS([np.asarray(v, dtype=np.complex128).reshape(1, -1, 1) for v in vectors])


def _qr_move_right(psi: OMPS, i):
    """generic_canonize_site!(site i, bond i-(i+1); method=:qr), canonize.jl:23-76."""
    a, b = psi.A[i], psi.A[i + 1]
    chil, d, chi = a.shape
    if chi == 1:  # canonize.jl:38 skips size-1 bonds
        return
    q, r = np.linalg.qr(a.reshape(chil * d, chi), mode="reduced")
    k = q.shape[1]
    psi.A[i] = q.reshape(chil, d, k)
    psi.A[i + 1] = (r @ b.reshape(chi, -1)).reshape(k, b.shape[1], b.shape[2])


def _qr_move_left(psi: OMPS, i):
    """generic_canonize_site!(site i, bond (i-1)-i; method=:qr), canonize.jl:23-76."""
    a, b = psi.A[i], psi.A[i - 1]
    chi, d, chir = a.shape
    if chi == 1:
        return
    # rows = (o, r), cols = l: QR of a^T-shaped matrix
    mat = a.reshape(chi, d * chir).T
    q, r = np.linalg.qr(mat, mode="reduced")
    k = q.shape[1]
    psi.A[i] = q.T.reshape(k, d, chir)
    psi.A[i - 1] = (b.reshape(-1, chi) @ r.T).reshape(b.shape[0], b.shape[1], k)


def canonize_mixed(psi: OMPS, lo, hi=None):
    """Mixed -> Mixed sweep (canonize.jl:128-149) with 0-based centre [lo, hi]."""
    hi = lo if hi is None else hi
    if psi.form[0] == "none":  # canonize.jl:123-126
        psi.form = ("mixed", 0, psi.n - 1)
    if psi.form[0] == "bond":
        _bond_to_mixed(psi, lo, hi)
        return psi
    _, slo, shi = psi.form
    if (slo, shi) == (lo, hi):
        return psi
    for i in range(slo, lo):  # left-to-right QR sweep
        _qr_move_right(psi, i)
    for i in range(shi, hi, -1):  # right-to-left QR sweep
        _qr_move_left(psi, i)
    psi.form = ("mixed", lo, hi)
    return psi


def _bond_canonize_site(psi: OMPS, i):
    """generic_bond_canonize_site! at site i across bond i-(i+1) (canonize.jl:78-116)."""
    a, b = psi.A[i], psi.A[i + 1]
    chil, d, chi = a.shape
    if chi == 1:
        psi.lam = np.ones(1)
        return
    u, s, vh = np.linalg.svd(a.reshape(chil * d, chi), full_matrices=False)
    k = s.shape[0]
    psi.A[i] = u.reshape(chil, d, k)
    psi.A[i + 1] = (vh @ b.reshape(chi, -1)).reshape(k, b.shape[1], b.shape[2])
    psi.lam = s.copy()


def canonize_bond(psi: OMPS, i):
    """Mixed/Bond -> BondCanonical(bond i-(i+1)) (canonize.jl:151-179, 221-246)."""
    if psi.form[0] == "bond":
        if psi.form[1] == i:
            return psi
        j = psi.form[1]
        # absorb Lambda toward the target side, then proceed from Mixed
        if j < i:
            _absorb_lambda(psi, j + 1, left=True)
            psi.form = ("mixed", j + 1, j + 1)
        else:
            _absorb_lambda(psi, j, left=False)
            psi.form = ("mixed", j, j)
    if psi.form[0] == "none":
        psi.form = ("mixed", 0, psi.n - 1)
    _, lo, hi = psi.form
    if hi <= i:  # canonize to the right
        canonize_mixed(psi, i, i)
    elif i + 1 <= lo:  # canonize to the left: centre at new_max = i+1, site i+1 factorised
        canonize_mixed(psi, i + 1, i + 1)
        _bond_canonize_left(psi, i + 1)
        psi.form = ("bond", i)
        return psi
    elif lo <= i and i + 1 <= hi:
        canonize_mixed(psi, i, i)
    else:
        raise ValueError("cannot canonize")
    _bond_canonize_site(psi, i)
    psi.form = ("bond", i)
    return psi


def _bond_canonize_left(psi: OMPS, i):
    """Bond-canonize at site i across bond (i-1)-i (canonize.jl:78-116, site = new_max)."""
    a, b = psi.A[i], psi.A[i - 1]
    chi, d, chir = a.shape
    if chi == 1:
        psi.lam = np.ones(1)
        return
    u, s, vh = np.linalg.svd(a.reshape(chi, d * chir).T, full_matrices=False)
    k = s.shape[0]
    psi.A[i] = u.T.reshape(k, d, chir)
    psi.A[i - 1] = (b.reshape(-1, chi) @ vh.T).reshape(b.shape[0], b.shape[1], k)
    psi.lam = s.copy()


def _absorb_lambda(psi: OMPS, site, left):
    """hadamard(tensor, s) absorption of Lambda (canonize.jl:181-219)."""
    lam = psi.lam
    if left:  # Lambda sits on the left bond of ``site``
        psi.A[site] = psi.A[site] * lam[:, None, None]
    else:
        psi.A[site] = psi.A[site] * lam[None, None, :]
    psi.lam = None


def bond_to_mixed(psi: OMPS, lo, hi):
    _bond_to_mixed(psi, lo, hi)
    return psi


def _bond_to_mixed(psi: OMPS, lo, hi):
    """Bond -> Mixed (canonize.jl:181-219)."""
    j = psi.form[1]
    old_min, old_max = j, j + 1
    if old_max <= lo:
        _absorb_lambda(psi, old_max, left=True)
        psi.form = ("mixed", old_max, old_max)
    elif hi <= old_min:
        _absorb_lambda(psi, old_min, left=False)
        psi.form = ("mixed", old_min, old_min)
    else:
        raise ValueError("cannot canonize from Bond to this Mixed form")
    canonize_mixed(psi, lo, hi)


def compress_bond(psi: OMPS, i, maxdim=None, threshold=None):
    """generic_mps_compress!(tn, bond i-(i+1)) (compress.jl:29-69)."""
    if maxdim is None and threshold is None:
        raise ValueError("Either `maxdim` or `threshold` must be specified")
    canonize_bond(psi, i)
    k = truncation_keep(psi.lam, maxdim, threshold)
    psi.A[i] = psi.A[i][:, :, :k]
    psi.A[i + 1] = psi.A[i + 1][:k]
    psi.lam = psi.lam[:k]
    return psi


def compress(psi: OMPS, maxdim=None, threshold=None):
    """All bonds, sequentially (compress.jl:21-27); ``truncate!`` is the same."""
    for i in range(psi.n - 1):
        compress_bond(psi, i, maxdim, threshold)
    return psi


def absorb_lambda_into_right(psi: OMPS):
    """Test helper mirroring mps.jl:479-483: fold Lambda into a neighbour."""
    j = psi.form[1]
    psi.A[j] = psi.A[j] * psi.lam[None, None, :]
    psi.lam = None
    psi.form = ("mixed", j, j + 1)


def simple_update_mps(psi: OMPS, U4: np.ndarray, a: int, maxdim=None, threshold=None):
    """simple_update!(::MPS, gate on sites (a, a+1)) (simple_update.jl:82-92).

    ``U4`` is the 4x4 gate with site a as the most significant digit:
    U4[2*oa + ob, 2*ia + ib] (test/unit/simple_update.jl:19-33 fixes this
    convention).  Moves the centre onto (a, a+1), then Muscle.simple_update
    with AbsorbEqually.  Returns the kept singular values.
    """
    canonize_mixed(psi, a, a + 1)
    A, B = psi.A[a], psi.A[a + 1]
    chil, d, chi = A.shape
    _, d2, chir = B.shape
    G = U4.reshape(d, d2, d, d2)  # (oa, ob, ia, ib)
    theta = np.tensordot(A, B, axes=([2], [0]))  # (l, ia, ib, r)
    theta = np.einsum("xyij,lijr->lxyr", G, theta, optimize=False)
    mat = theta.reshape(chil * d, d2 * chir)
    u, s, vh = np.linalg.svd(mat, full_matrices=False)
    k = truncation_keep(s, maxdim, threshold)
    r = np.sqrt(s[:k])
    psi.A[a] = (u[:, :k] * r[None, :]).reshape(chil, d, k)
    psi.A[a + 1] = (r[:, None] * vh[:k]).reshape(k, d2, chir)
    psi.form = ("mixed", a, a + 1)
    return s[:k]


def apply_one_site(psi: OMPS, U2: np.ndarray, a: int):
    """1-site shortcut of generic_simple_update! (simple_update.jl:28-37)."""
    psi.A[a] = np.einsum("xi,lir->lxr", U2, psi.A[a], optimize=False)


def overlap_product(bits_vectors, psi: OMPS):
    """overlap(::ProductState, ::MPS) (overlap.jl:22-32): no conjugation of x."""
    v = np.ones(1, dtype=np.complex128)
    for x, a in zip(bits_vectors, psi.A):
        v = v @ np.einsum("i,lir->lr", np.asarray(x, dtype=a.dtype), a, optimize=False)
    return v[0]


def amplitude(bits, psi: OMPS):
    """<x|psi> for a 0/1 bitstring: a slice per site, then a vector-matrix chain."""
    v = np.ones(1, dtype=psi.A[0].dtype)
    for x, a in zip(bits, psi.A):
        v = v @ a[:, int(x), :]
    return v[0]


def amplitude_logscaled(bits, psi: OMPS):
    """Same chain with a per-site rescale; returns (mantissa, log_scale)."""
    v = np.ones(1, dtype=psi.A[0].dtype)
    lg = 0.0
    for x, a in zip(bits, psi.A):
        v = v @ a[:, int(x), :]
        m = np.max(np.abs(v))
        if m > 0:
            v = v / m
            lg += math.log(m)
    return v[0], lg


def overlap_mps(a: OMPS, b: OMPS):
    """overlap(a::MPS, b::MPS) = <b|a> (overlap.jl:36-50), b conjugated (:39)."""
    env = np.ones((1, 1), dtype=np.complex128)
    for x, y in zip(a.A, b.A):
        t = np.tensordot(env, x, axes=([0], [0]))  # (lb, o, ra)
        env = np.tensordot(t, y.conj(), axes=([0, 1], [0, 1]))  # (ra, rb)
    return env[0, 0]


def overlap_mps_logscaled(a: OMPS, b: OMPS):
    """<b|a> with a per-site rescale of the environment; returns (value, log_scale)."""
    env = np.ones((1, 1), dtype=np.complex128)
    lg = 0.0
    for x, y in zip(a.A, b.A):
        t = np.tensordot(env, x, axes=([0], [0]))
        env = np.tensordot(t, y.conj(), axes=([0, 1], [0, 1]))
        m = np.max(np.abs(env))
        if m > 0:
            env = env / m
            lg += math.log(m)
    return env[0, 0], lg


def norm(psi: OMPS):
    """norm(::MPS) (norm.jl:17-37): canonize to the min centre, Frobenius norm."""
    if psi.form[0] == "bond":
        return float(np.linalg.norm(psi.lam))
    if psi.form[0] == "none":
        psi.form = ("mixed", 0, psi.n - 1)
    lo = psi.form[1]
    canonize_mixed(psi, lo, lo)
    return float(np.linalg.norm(psi.A[lo]))


def normalize(psi: OMPS):
    """normalize!(::MPS) (normalize.jl:15-37)."""
    if psi.form[0] == "bond":
        psi.lam = psi.lam / np.linalg.norm(psi.lam)
        return psi
    if psi.form[0] == "none":
        psi.form = ("mixed", 0, psi.n - 1)
    lo = psi.form[1]
    canonize_mixed(psi, lo, lo)
    psi.A[lo] = psi.A[lo] / np.linalg.norm(psi.A[lo])
    return psi


def is_left_isometry(a, atol=1e-12):
    chil, d, chir = a.shape
    m = a.reshape(chil * d, chir)
    return np.allclose(m.conj().T @ m, np.eye(chir), atol=atol)


def is_right_isometry(a, atol=1e-12):
    chil, d, chir = a.shape
    m = a.reshape(chil, d * chir)
    return np.allclose(m @ m.conj().T, np.eye(chil), atol=atol)


def checkform(psi: OMPS, atol=1e-12):
    """checkform for MixedCanonical / BondCanonical (MPO.jl:41-83)."""
    if psi.form[0] == "mixed":
        lo, hi = psi.form[1], psi.form[2]
    elif psi.form[0] == "bond":
        lo, hi = psi.form[1], psi.form[1] + 1
    else:
        return True
    for i in range(psi.n):
        if i < lo and not is_left_isometry(psi.A[i], atol):
            return False
        if i > hi and not is_right_isometry(psi.A[i], atol):
            return False
    return True


def schmidt_values(psi: OMPS, i):
    """Tenet.schmidt_values (entropy.jl:25-31): BondCanonical Lambda."""
    p = psi.copy()
    canonize_bond(p, i)
    return p.lam.copy()


# --------------------------------------------------------------------------
# transfer-matrix expectation of a Pauli string
# --------------------------------------------------------------------------

PAULI = {
    "I": np.eye(2, dtype=np.complex128),
    "X": np.array([[0, 1], [1, 0]], dtype=np.complex128),
    "Y": np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    "Z": np.array([[1, 0], [0, -1]], dtype=np.complex128),
}


def expectation_pauli(psi: OMPS, start: int, paulis: str):
    """<psi|P|psi> / 1 for a Pauli string on sites start..start+w-1.

    Composition from the reference's pieces: apply one-site
    Paulis by the 1-site shortcut (simple_update.jl:28-37), then the
    transfer-matrix overlap (overlap.jl:36-50).  Pauli matrices follow
    ext/TenetPythonCallExt/pytket.jl:23-31.
    """
    env = np.ones((1, 1), dtype=np.complex128)
    for i, a in enumerate(psi.A):
        op = PAULI[paulis[i - start]] if start <= i < start + len(paulis) else None
        pa = a if op is None else np.einsum("xi,lir->lxr", op, a, optimize=False)
        t = np.tensordot(env, pa, axes=([0], [0]))
        env = np.tensordot(t, a.conj(), axes=([0, 1], [0, 1]))
    return env[0, 0]


# --------------------------------------------------------------------------
# seeded generators (numpy PCG64)
# --------------------------------------------------------------------------

BASE_SEED = 20261202


def sample(psi: OMPS, nsamples: int, seed: int):
    """sample(psi, nsamples) (sample.jl:20-72) restated per sample: on a copy
    canonised to the first site, v = [1]; at site i t = v . A_i (x, r),
    p_x = sum_r |t[x, r]|^2, x* = the first x whose running sum exceeds
    u * sum(p) (the last x with p_x > 0 if none), v = t[x*] / sqrt(p_x*).
    The uniform u of (sample s, site i) is u1 of generator element s n + i
    (synth.uniforms; Julia's RNG stream is not reproducible here, so this is
    the device's stream).  Returns 0-based values (Tenet's are 1-based)."""
    from .synth import uniforms

    p = psi.copy()
    canonize_mixed(p, 0, 0)
    n = p.n
    u = uniforms(seed, 0, nsamples * n)[0].reshape(nsamples, n)
    out = np.zeros((nsamples, n), dtype=np.int32)
    for s in range(nsamples):
        v = np.ones(1, dtype=p.A[0].dtype)
        for i in range(n):
            t = np.tensordot(v, p.A[i], axes=([0], [0]))  # (x, r)
            pr = np.sum(np.abs(t) ** 2, axis=1)
            target = u[s, i] * pr.sum()
            cum, pick, last = 0.0, -1, 0
            for x in range(len(pr)):
                if pr[x] > 0.0:
                    last = x
                cum += pr[x]
                if pick < 0 and pr[x] > 0.0 and cum > target:
                    pick = x
            pick = last if pick < 0 else pick
            out[s, i] = pick
            v = t[pick] / np.sqrt(pr[pick])
    return out


def complex_gaussian(rng, shape, dtype=np.complex128):
    """Standard complex Gaussian (N(0,1) + i N(0,1)) / sqrt(2)."""
    re = rng.standard_normal(shape)
    im = rng.standard_normal(shape)
    return ((re + 1j * im) / math.sqrt(2)).astype(dtype)


def haar_unitary(rng, n=4):
    """Haar-random unitary: QR of a Ginibre matrix with the phase fix of
    Mezzadri, Notices AMS 54 (2007) 592-604."""
    z = complex_gaussian(rng, (n, n))
    q, r = np.linalg.qr(z)
    ph = np.diag(r) / np.abs(np.diag(r))
    return q * ph[None, :]


def random_gaussian_mps(rng, n, d, chi, dtype=np.complex128, cap_edges=False):
    """Open MPS with i.i.d. Gaussian entries.  With ``cap_edges`` each bond is
    capped at d^min(b, n-b) as Tenet's rand(MPS) does (MPS.jl:120-152)."""
    dims = []
    for b in range(1, n):
        c = chi
        if cap_edges:
            c = min(chi, d ** min(b, n - b)) if min(b, n - b) < 63 else chi
        dims.append(c)
    arrays = []
    for i in range(n):
        chil = 1 if i == 0 else dims[i - 1]
        chir = 1 if i == n - 1 else dims[i]
        arrays.append(complex_gaussian(rng, (chil, d, chir), dtype))
    return OMPS(arrays)


def brickwork_schedule(n, depth):
    """Depth-``depth`` brickwork: odd layers on bonds (1,2),(3,4),... left to
    right; even layers on (2,3),(4,5),... right to left (serpentine).
    Yields 0-based left sites."""
    for layer in range(depth):
        if layer % 2 == 0:
            yield from range(0, n - 1, 2)
        else:
            yield from reversed(range(1, n - 1, 2))


def tebd_gates(seed, n, depth):
    rng = np.random.default_rng(seed)
    sched = list(brickwork_schedule(n, depth))
    gates = np.stack([haar_unitary(rng, 4) for _ in sched])
    return sched, gates


def tebd_run(n, sched, gates, maxdim, threshold, progress=None):
    """TEBD as the user loop of simple_update! over the brickwork (simple_update.jl:82-92 per gate)."""
    psi = product_state_mps([[1.0, 0.0]] * n)
    for g, (a, U) in enumerate(zip(sched, gates)):
        simple_update_mps(psi, U, a, maxdim, threshold)
        if progress is not None:
            progress(g)
    return psi


# --------------------------------------------------------------------------
# Vidal-gauge TEBD (Gamma / Lambda form)
# --------------------------------------------------------------------------
# Tenet holds the Gamma / Lambda containers (VidalMPS.jl:19-204) but its
# Vidal update and conversions are unfinished (simple_update.jl:94-112 and
# canonize.jl:258-310 are TODO stubs; no test covers them).  This restates
# the published algorithm: G. Vidal, Phys. Rev. Lett. 91, 147902 (2003), in
# the inverse-free form of M. B. Hastings, J. Math. Phys. 50, 095207 (2009):
# sites are kept as B_i = Gamma_i Lambda_i (right-canonical), Lambda_i is the
# Schmidt spectrum of bond i-(i+1).  A gate on (a, a+1):
#   theta~ = G . (B_a B_{a+1}),  X = Lambda_{a-1} theta~ = U S V^H,
#   keep k by the compress! rule (compress.jl:46-59: maxdim, then the
#   absolute threshold on raw S),
#   B_{a+1} <- V^H[:k],  Lambda_a <- S[:k],  B_a <- theta~ V[:, :k]
# (B_a = Lambda_{a-1}^{-1} U S without dividing by Lambda).  The gates of a
# brickwork layer act on disjoint bonds and commute, so a layer's updates
# are independent.


class OVidal:
    """Vidal-gauge MPS: B[i] = Gamma_i Lambda_i as (chi_l, d, chi_r) arrays,
    lam[i] the Lambda vector of bond i-(i+1) (n - 1 of them)."""

    def __init__(self, B, lam):
        self.B = [np.asarray(b) for b in B]
        self.lam = [np.asarray(x, dtype=np.float64) for x in lam]

    @property
    def n(self):
        return len(self.B)

    def bondsizes(self):
        return [b.shape[2] for b in self.B[:-1]]

    def to_mps(self) -> OMPS:
        """The B sites are right-canonical: an MPS with its centre at site 0."""
        return OMPS([b.copy() for b in self.B], form=("mixed", 0, 0))


def vidal_product_state(vectors) -> OVidal:
    B = [np.asarray(v, dtype=np.complex128).reshape(1, -1, 1) for v in vectors]
    nrm = [np.linalg.norm(b) for b in B]
    B = [b / x for b, x in zip(B, nrm)]
    B[0] = B[0] * float(np.prod(nrm))
    return OVidal(B, [np.ones(1) for _ in range(len(B) - 1)])


def vidal_from_mps(psi: OMPS) -> OVidal:
    """Right-canonicalise (centre at site 0), then one left-to-right SVD
    sweep: X_i = Lambda_{i-1} Bt_i = U S V^H with Bt_0 = A_0 and
    Bt_{i+1} = V^H A_{i+1}; B_i = Bt_i V, Lambda_i = S (all nonzero kept)."""
    p = psi.copy()
    canonize_mixed(p, 0, 0)
    n = p.n
    B, lam = [], []
    bt = p.A[0]
    prev = np.ones(1)
    for i in range(n - 1):
        l, d, r = bt.shape
        u, s, vh = np.linalg.svd((prev[:, None, None] * bt).reshape(l * d, r), full_matrices=False)
        k = int(np.count_nonzero(s > 0.0))
        vh = vh[:k]
        B.append(np.tensordot(bt, vh.conj(), axes=([2], [1])))  # Bt V
        lam.append(s[:k])
        bt = np.tensordot(vh, p.A[i + 1], axes=([1], [0]))
        prev = s[:k]
    B.append(bt)
    return OVidal(B, lam)


def vidal_gate(v: OVidal, U4: np.ndarray, a: int, maxdim=None, threshold=None):
    """One two-site update on (a, a+1); U4 as in simple_update_mps."""
    A, Bn = v.B[a], v.B[a + 1]
    chil, d, _ = A.shape
    _, d2, chir = Bn.shape
    G = U4.reshape(d, d2, d, d2)  # (oa, ob, ia, ib)
    th = np.tensordot(A, Bn, axes=([2], [0]))  # (l, ia, ib, r)
    th = np.einsum("xyij,lijr->lxyr", G, th, optimize=False)
    left = v.lam[a - 1] if a > 0 else np.ones(1)
    X = (left[:, None, None, None] * th).reshape(chil * d, d2 * chir)
    _, s, vh = np.linalg.svd(X, full_matrices=False)
    k = truncation_keep(s, maxdim, threshold)
    vh = vh[:k]
    v.B[a] = (th.reshape(chil * d, d2 * chir) @ vh.conj().T).reshape(chil, d, k)
    v.B[a + 1] = vh.reshape(k, d2, chir)
    v.lam[a] = s[:k]
    return s[:k]


def brickwork_layers(n, depth):
    """The brickwork schedule as layers of disjoint bonds (left sites)."""
    return [list(range(0, n - 1, 2)) if layer % 2 == 0 else list(range(1, n - 1, 2)) for layer in range(depth)]


def vidal_tebd_run(n, layers, gates, maxdim, threshold):
    """Vidal-gauge TEBD from |0...0>: gates consumed layer by layer in the
    order of ``layers`` (one gate per bond entry)."""
    v = vidal_product_state([[1.0, 0.0]] * n)
    g = 0
    for bonds in layers:
        for a in bonds:
            vidal_gate(v, gates[g], a, maxdim, threshold)
            g += 1
    return v


# --------------------------------------------------------------------------
# MPO application (evolve!(::MPS, ::AbstractMPO))
# --------------------------------------------------------------------------
# MPO sites as W[l, r, o, i] arrays with size-1 edge bonds (MPO.jl:7 order;
# Tenet stores the edge sites without the l / r leg).  The MPS's physical
# index is contracted with the MPO's input i, the output o becomes the new
# physical index (align!(mps, :outputs, op, :inputs), evolve.jl:24-25).


def mpo_sites(W, n, left_vec, right_vec):
    """Site tensors of a uniform MPO from its bulk W[l, r, o, i] and boundary vectors."""
    first = np.einsum("l,lroi->roi", np.asarray(left_vec, W.dtype), W, optimize=False)[None]
    last = np.einsum("lroi,r->loi", W, np.asarray(right_vec, W.dtype), optimize=False)[:, None]
    return [first] + [W.copy() for _ in range(n - 2)] + [last]


def mpo_sites_dense(Ws):
    """Dense operator (site 1 the most significant digit) from MPO site tensors."""
    M = Ws[0][0]  # (r, o, i)
    for W in Ws[1:]:
        M = np.einsum("aOI,abxy->bOxIy", M, W, optimize=False)
        D = M.shape[1] * M.shape[2]
        M = M.reshape(M.shape[0], D, D)
    return M[0]


def evolve_mpo_direct(psi: OMPS, Ws) -> OMPS:
    """generic_evolve_mps_mpo_direct! (evolve.jl:19-46) + the canonical-form
    reset of evolve!(::MPS, ::AbstractMPO) (evolve.jl:131-138): per site
    A'[(l, L), o, (r, R)] = sum_p A[l, p, r] W[L, R, o, p], the virtual
    pairs fused (mps bond major, MPO bond minor), form MixedCanonical(all)."""
    out = []
    for A, W in zip(psi.A, Ws):
        c = np.einsum("lpr,LRop->lLorR", A, W, optimize=False)
        l, L_, o, r, R = c.shape
        out.append(c.reshape(l * L_, o, r * R))
    return OMPS(out, form=("mixed", 0, psi.n - 1))


def evolve_mpo_zipup(psi: OMPS, Ws, maxdim=None, threshold=None) -> OMPS:
    """generic_evolve_mps_mpo_zipup! (evolve.jl:49-129): site 1 gives
    C[o, (r, R)] = U S V, truncated by the compress! rule, site 1 <- U,
    carry R = S V; site i: R . A_i . W_i as [(o, s), (r, R)] = U S V, site i
    <- U as (s, o, s'), carry S V; the last site takes R . A_n . W_n.  The
    reference spells the keyword ``threhold`` (evolve.jl:49).  The result's
    centre is the last site (the U's are left-isometric); the reference
    leaves the canonical-form trait as it was."""
    n = psi.n
    A, W = psi.A[0], Ws[0]
    C = np.einsum("pr,Rop->orR", A[0], W[0], optimize=False)  # (o, r, R)
    d, r, R = C.shape
    u, s, vh = np.linalg.svd(C.reshape(d, r * R), full_matrices=False)
    k = truncation_keep(s, maxdim, threshold)
    sites = [u[:, :k].reshape(1, d, k)]
    carry = (s[:k, None] * vh[:k]).reshape(k, r, R)
    for i in range(1, n - 1):
        t = np.einsum("srR,rpq->sRpq", carry, psi.A[i], optimize=False)
        t = np.einsum("sRpq,RQop->osqQ", t, Ws[i], optimize=False)  # (o, s, r2, R2)
        o, sd, q, Q = t.shape
        u, s, vh = np.linalg.svd(t.reshape(o * sd, q * Q), full_matrices=False)
        k = truncation_keep(s, maxdim, threshold)
        sites.append(np.transpose(u[:, :k].reshape(o, sd, k), (1, 0, 2)))
        carry = (s[:k, None] * vh[:k]).reshape(k, q, Q)
    t = np.einsum("srR,rp->sRp", carry, psi.A[n - 1][:, :, 0], optimize=False)
    t = np.einsum("sRp,Rop->so", t, Ws[n - 1][:, 0], optimize=False)
    sites.append(t[:, :, None])
    return OMPS(sites, form=("mixed", n - 1, n - 1))


# --------------------------------------------------------------------------
# DMRG pieces: MPOs and the two-site effective Hamiltonian
# --------------------------------------------------------------------------


def ising_mpo_W(h=1.0, J=1.0):
    """Bulk tensor of ising_1d_mpo (Models/Ising.jl:12-30) as W[l, r, o, i]."""
    Id = np.eye(2)
    sx = np.array([[0.0, 1.0], [1.0, 0.0]])
    sz = np.array([[1.0, 0.0], [0.0, -1.0]])
    W = np.zeros((2, 2, 3, 3), dtype=np.complex128)  # (i, o, l, r) as in Ising.jl
    W[:, :, 0, 0] = Id
    W[:, :, 1, 0] = sz
    W[:, :, 2, 0] = -h * sx
    W[:, :, 2, 1] = -J * sz
    W[:, :, 2, 2] = Id
    return np.transpose(W, (2, 3, 1, 0))  # (l, r, o, i)


def xxz_mpo_W(J=1.0, delta=0.5):
    """Heisenberg XXZ bulk MPO tensor W[l, r, o, i] (5x5 lower-triangular form):
    H = J sum (S+S- + S-S+)/2 + J*Delta sum Sz Sz with S = sigma/2."""
    sp = np.array([[0.0, 1.0], [0.0, 0.0]])
    sm = sp.T.copy()
    sz = 0.5 * np.array([[1.0, 0.0], [0.0, -1.0]])
    Id = np.eye(2)
    W = np.zeros((5, 5, 2, 2), dtype=np.complex128)
    W[0, 0] = Id
    W[1, 0] = sp
    W[2, 0] = sm
    W[3, 0] = sz
    W[4, 1] = 0.5 * J * sm
    W[4, 2] = 0.5 * J * sp
    W[4, 3] = J * delta * sz
    W[4, 4] = Id
    return W  # (l, r, o, i): W[l, r] is the operator block (o, i)


def heff_apply(L, W1, W2, R, theta):
    """EffectiveHamiltonian(le, re, op)(psi) = ((le.psi).op).re (DMRG.jl:16-18)
    with op = W1.W2 pre-contracted (DMRG.jl:297-299).

    Shapes: L (chi_l, w, chi_l'), theta (chi_l', i1, i2, chi_r'), W (w, w, o, i),
    R (chi_r, w, chi_r').  Returns y (chi_l, o1, o2, chi_r).  Applied as three
    pairwise contractions (L.theta, .W1W2 over (w, i1, i2), .R) like the reference.
    """
    t = np.tensordot(L, theta, axes=([2], [0]))  # (a, w, i1, i2, c)
    op = np.einsum("ab xi, bc yj->a c x y i j".replace(" ", ""), W1, W2, optimize=False)  # (wl, wr, o1, o2, i1, i2)
    t = np.einsum("awijc,wexyij->axyec", t, op, optimize=False)  # (a, o1, o2, wr, c)
    y = np.tensordot(t, R, axes=([3, 4], [1, 2]))  # (a, o1, o2, d)
    return y


def mpo_dense(Ws, left_vec, right_vec):
    """Dense Hamiltonian from bulk W (l, r, o, i) tensors with boundary vectors."""
    n = len(Ws)
    M = np.einsum("l,lroi->roi", left_vec, Ws[0], optimize=False)
    for W in Ws[1:]:
        M = np.einsum("aOI,abxy->bOxIy", M, W, optimize=False)
        D = M.shape[1] * M.shape[2]
        M = M.reshape(M.shape[0], D, D)
    return np.einsum("roi,r->oi", M, right_vec, optimize=False)
