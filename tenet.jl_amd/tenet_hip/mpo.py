"""MPO on the device and its application to an MPS (evolve!(::MPS, ::AbstractMPO)).

Sites are column-major (chi_l, d_out, d_in, chi_r) DeviceArrays with size-1
edge bonds (Tenet's MPO order is (:l, :r, :o, :i), MPO.jl:7, with the edge
sites stored without the l / r leg; here every site keeps both legs so the
layout matches the MPS's (chi_l, d, chi_r)).  The MPS's physical index is
contracted with the MPO's input and the output becomes the new physical
index (align!(mps, :outputs, op, :inputs), src/Operations/evolve.jl:24-25).

  evolve_direct_   generic_evolve_mps_mpo_direct! + the form reset of
                   evolve!(::MPS, ::AbstractMPO)   evolve.jl:19-46, 131-138
  evolve_zipup_    generic_evolve_mps_mpo_zipup!   evolve.jl:49-129

The direct product runs every site's contraction in one grouped launch; the
fused virtual pairs (mps bond major ... in column-major order: the mps index
fastest) need no copy, since (l, L, o, r, R) contiguous is (l L, o, r R).
"""
from __future__ import annotations

import numpy as np

from .device import Context, DeviceArray
from .grouped import GroupedPlan
from .mps import MPS, svd_flops
from .vidal import _leading_rows


class MPO:
    """Canonical forms: Tenet runs the MPS code on an MPO (canonize!(::MPO),
    canonize.jl:255; compress!(::MPO), compress.jl:72), with the site's
    non-bond legs (o, i) as the matricised rows.  Here the column-major
    (chi_l, o, i, chi_r) site is the (chi_l, o i, chi_r) MPS site with a
    fused physical index, so the MPS methods run on views and the result
    is viewed back; ``form`` and ``lam`` follow the MPS conventions."""

    def __init__(self, ctx: Context, sites, dtype=np.complex128, form=None):
        self.ctx = ctx
        self.sites = list(sites)
        self.dtype = np.dtype(dtype)
        self.form = form if form is not None else ("mixed", 0, len(self.sites) - 1)  # MPO.jl: MixedCanonical(all)
        self.lam = None

    @property
    def n(self):
        return len(self.sites)

    def bondsizes(self):
        return [s.shape[3] for s in self.sites[:-1]]

    def copy(self):
        o = MPO(self.ctx, [s.copy() for s in self.sites], self.dtype, self.form)
        o.lam = None if self.lam is None else self.lam.copy()
        return o

    # ---- canonical forms on the fused (o, i) index -------------------------------
    def _as_mps(self) -> MPS:
        m = MPS(self.ctx, [s.view((s.shape[0], s.shape[1] * s.shape[2], s.shape[3])) for s in self.sites],
                self.form, self.dtype)
        m.lam = self.lam
        return m

    def _take(self, m: MPS):
        dims = [(s.shape[1], s.shape[2]) for s in self.sites]
        self.sites = [a.view((a.shape[0], o, i, a.shape[2])) for a, (o, i) in zip(m.sites, dims)]
        self.form = m.form
        self.lam = m.lam
        return self

    def canonize_(self, lo, hi=None):
        """canonize!(::MPO, MixedCanonical) (canonize.jl:121-255)."""
        return self._take(self._as_mps().canonize_(lo, hi))

    def canonize_bond_(self, i):
        """canonize!(::MPO, BondCanonical(bond i-(i+1)))."""
        return self._take(self._as_mps().canonize_bond_(i))

    def compress_bond_(self, i, maxdim=None, threshold=None):
        """compress!(::MPO, bond) (compress.jl:29-72)."""
        return self._take(self._as_mps().compress_bond_(i, maxdim, threshold))

    def compress_(self, maxdim=None, threshold=None):
        """compress!(::MPO) over all bonds (compress.jl:21-27, 72)."""
        return self._take(self._as_mps().compress_(maxdim, threshold))

    truncate_ = compress_

    def norm(self):
        """Frobenius norm of the operator (norm.jl:17-37 on an MPO)."""
        m = self._as_mps()
        v = m.norm()
        self._take(m)
        return v

    @staticmethod
    def from_host(ctx: Context, arrays_lroi, dtype=np.complex128, form=None):
        """From host W[l, r, o, i] site arrays (size-1 edge bonds)."""
        sites = [DeviceArray.from_host(ctx, np.transpose(np.asarray(W), (0, 2, 3, 1)), dtype) for W in arrays_lroi]
        return MPO(ctx, sites, dtype, form)

    @staticmethod
    def from_arrays(ctx: Context, arrays, order=("l", "r", "o", "i"), dtype=np.complex128):
        """Tenet's ``MPO(arrays; order)`` (MPO.jl:41-132): the edge arrays
        lack the l (first site) or r (last site) leg."""
        order = list(order)
        n = len(arrays)
        out = []
        for j, arr in enumerate(arrays):
            loc = [x for x in order if not ((j == 0 and x == "l") or (j == n - 1 and x == "r"))]
            a = np.asarray(arr)
            if j == 0:
                a = a[None]
                loc = ["l"] + loc
            if j == n - 1:
                a = a[..., None]
                loc = loc + ["r"]
            out.append(np.transpose(a, [loc.index(x) for x in ("l", "r", "o", "i")]))
        return MPO.from_host(ctx, out, dtype)

    @staticmethod
    def from_bulk(ctx: Context, W, n, left_vec, right_vec, dtype=np.complex128):
        """A uniform MPO from its bulk W[l, r, o, i] and boundary vectors."""
        W = np.asarray(W)
        first = np.einsum("l,lroi->roi", np.asarray(left_vec, W.dtype), W, optimize=False)[None]
        last = np.einsum("lroi,r->loi", W, np.asarray(right_vec, W.dtype), optimize=False)[:, None]
        return MPO.from_host(ctx, [first] + [W] * (n - 2) + [last], dtype)

    def to_host_sites(self):
        """Site arrays in the device layout (chi_l, o, i, chi_r)."""
        return [s.to_host() for s in self.sites]

    def to_host(self):
        """Site arrays as W[l, r, o, i]."""
        return [np.transpose(s.to_host(), (0, 3, 1, 2)) for s in self.sites]


def evolve_direct_(psi: MPS, op: MPO) -> MPS:
    """A_i <- A_i . W_i over the physical index, virtual pairs fused; the
    bond dimensions multiply.  All sites in one grouped launch; the form
    becomes MixedCanonical(all sites) (the product is not canonical)."""
    if psi.n != op.n:
        raise ValueError("MPS and MPO must have the same number of sites")
    probs, ptrs, outs = [], [], []
    for A, W in zip(psi.sites, op.sites):
        l, p, r = A.shape
        L_, o, i, R = W.shape
        if i != p:
            raise ValueError("MPO input dimension differs from the MPS physical dimension")
        c = DeviceArray(psi.ctx, (l, L_, o, r, R), psi.dtype)
        probs.append(((l, p, r), "apb", (L_, o, p, R), "AopB", (l, L_, o, r, R), "aAobB"))
        ptrs += [A.ptr, W.ptr, c.ptr]
        outs.append(c)
        psi.flops["contract"] += 8.0 * l * p * r * L_ * o * R
    plan = GroupedPlan(psi.ctx, probs, psi.dtype)
    try:
        plan.exec(ptrs)
    finally:
        plan.close()
    psi.sites = [c.view((c.shape[0] * c.shape[1], c.shape[2], c.shape[3] * c.shape[4])) for c in outs]
    psi.form = ("mixed", 0, psi.n - 1)
    psi.lam = None
    psi._gate_bond = None
    return psi


def evolve_zipup_(psi: MPS, op: MPO, maxdim=None, threshold=None) -> MPS:
    """Zip-up: C_1 = A_1 . W_1 = U S V (rows: the physical output), site 1
    <- U, carry S V; site i: (S V) . A_i . W_i with rows (output, left bond)
    = U S V, site i <- U, carry S V; the last site takes the carry's
    product.  Each cut keeps min(maxdim, #S > threshold) (the reference's
    ``threhold`` keyword).  The U's are left-isometric, so the centre ends
    on the last site."""
    if psi.n != op.n:
        raise ValueError("MPS and MPO must have the same number of sites")
    n = psi.n
    md = 0 if maxdim is None else int(maxdim)
    thr = 0.0 if threshold is None else float(threshold)
    ctx, dt = psi.ctx, psi.dtype

    class _V:  # carrier for _leading_rows
        pass

    hv = _V()
    hv.ctx, hv.dtype, hv.code = ctx, dt, psi.code
    hv._one = DeviceArray.from_host(ctx, np.ones((), dtype=dt), dt)

    def cut(M, m, ncol):
        """Truncated SVD of M (m x ncol): (U[:, :k], S[:k] V^H[:k], k)."""
        U, s, Vh, k = psi._svd(M, m, ncol, md, thr)
        psi.flops["svd"] += svd_flops(m, ncol)
        Uk = DeviceArray(ctx, (m, k), dt, ptr=U.ptr, owner=False)
        Uk._owner = U
        Vk = Vh if k == Vh.shape[0] else _leading_rows(hv, Vh, k)
        sv = DeviceArray.from_host(ctx, np.asarray(s[:k]).astype(dt), dt)
        carry = psi._einsum(sv, [1], Vk, [1, 2], (k, ncol), [1, 2])
        return Uk, carry, k

    # site 1: C[o, r, R] = sum_p A(p, r) W(o, p, R)
    A, W = psi.sites[0], op.sites[0]
    _, p, r = A.shape
    _, o, _, R = W.shape
    C = psi._einsum(A.view((p, r)), [2, 3], W.view((o, p, R)), [5, 2, 6], (o, r, R), [5, 3, 6])
    psi.flops["contract"] += 8.0 * p * r * o * R
    Uk, carry, k = cut(C, o, r * R)
    new_sites = [Uk.view((1, o, k))]
    cd = (k, r, R)
    for i in range(1, n - 1):
        A, W = psi.sites[i], op.sites[i]
        _, p, q = A.shape
        _, o, _, Q = W.shape
        cv = carry.view(cd)
        t = psi._einsum(cv, [1, 2, 3], A, [2, 4, 5], (cd[0], cd[2], p, q), [1, 3, 4, 5])  # (s, R, p, q)
        t2 = psi._einsum(t, [1, 3, 4, 5], W, [3, 6, 4, 7], (o, cd[0], q, Q), [6, 1, 5, 7])  # (o, s, q, Q)
        psi.flops["contract"] += 8.0 * cd[0] * cd[1] * cd[2] * p * q + 8.0 * cd[0] * cd[2] * p * q * o * Q
        Uk, carry, k = cut(t2, o * cd[0], q * Q)
        # site i as (s, o, s'): U's rows are (o, s), o fastest
        new_sites.append(psi._einsum(Uk.view((o, cd[0], k)), [1, 2, 3], hv._one, [], (cd[0], o, k), [2, 1, 3]))
        cd = (k, q, Q)
    # last site: (S V) . A_n . W_n
    A, W = psi.sites[n - 1], op.sites[n - 1]
    _, p, _ = A.shape
    L_, o, _, _ = W.shape
    cv = carry.view(cd)
    t = psi._einsum(cv, [1, 2, 3], A.view((cd[1], p)), [2, 4], (cd[0], cd[2], p), [1, 3, 4])
    last = psi._einsum(t, [1, 3, 4], W.view((L_, o, p)), [3, 6, 4], (cd[0], o), [1, 6])
    psi.flops["contract"] += 8.0 * cd[0] * cd[1] * cd[2] * p + 8.0 * cd[0] * cd[2] * p * o
    new_sites.append(last.view((cd[0], o, 1)))
    psi.sites = new_sites
    psi.form = ("mixed", n - 1, n - 1)
    psi.lam = None
    psi._gate_bond = None
    return psi
