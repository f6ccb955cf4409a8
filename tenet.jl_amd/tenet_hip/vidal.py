"""Vidal-gauge MPS and TEBD on the device (Gamma / Lambda form).

Tenet keeps the Gamma / Lambda containers (src/Components/VidalMPS.jl:19-204)
but its Vidal update and conversions are unfinished (TODO stubs at
src/Operations/simple_update.jl:94-112 and canonize.jl:258-310).  This
follows the published algorithm: G. Vidal, Phys. Rev. Lett. 91, 147902
(2003), in the inverse-free form of M. B. Hastings, J. Math. Phys. 50,
095207 (2009).

Sites are kept as B_i = Gamma_i Lambda_i, right-canonical column-major
(chi_l, d, chi_r) DeviceArrays; Lambda_i (the Schmidt spectrum of bond
i-(i+1)) is a device vector with zero imaginary part, so the Lambda_{a-1}
scaling is a Hadamard contraction (binary_einsum with the label kept).  A
gate on (a, a+1):

  theta~ = G . (B_a B_{a+1}),   X = Lambda_{a-1} theta~ = U S V^H,
  keep k by compress!'s rule (compress.jl:46-59: maxdim, then the absolute
  threshold on raw S),
  B_{a+1} <- V^H[:k],  Lambda_a <- S[:k],  B_a <- theta~ V[:, :k]

(B_a = Lambda_{a-1}^{-1} U S without dividing by Lambda).  The gates of a
brickwork layer act on disjoint bonds and commute, so ``apply_layer_`` runs
a layer as four grouped contraction launches and one grouped SVD:

  1. theta_k = B_a B_{a+1}                 (contract the bond)
  2. theta~_k = theta_k . G_k              (contract the gate inputs)
  3. X_k = Lambda_{a-1} (.) theta~_k       (Hadamard; at a = 0 X = theta~)
  4. grouped one-sided Jacobi SVD of X_k, truncation on the device
  5. B_a <- theta~_k conj(V^H[:k])^T and B_{a+1} <- V^H[:k] (a conjugated
     contraction and a strided copy, one launch)

Plans are cached on the context by the layer's shapes (after the bonds
saturate every layer of a parity reuses its plans, and so does every later
sweep).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .device import DTYPES, Context, DeviceArray, tdesc
from .grouped import GroupedPlan
from .mps import MPS, _gate_tensor, svd_flops


class VidalMPS:
    def __init__(self, ctx: Context, B, lam_host, dtype=np.complex128):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        self.code = DTYPES[self.dtype]
        self.B = list(B)
        self.lam_host = [np.asarray(x, dtype=np.float64) for x in lam_host]
        self.lam = self._upload_lams(self.lam_host)
        self._plans = {}
        self._one = DeviceArray.from_host(ctx, np.ones((), dtype=self.dtype), self.dtype)
        self.flops = {"svd": 0.0, "contract": 0.0}

    # ---- construction -------------------------------------------------
    def _upload_lams(self, lams):
        """Device Lambda vectors (complex, one upload for all)."""
        if not lams:
            return []
        flat = np.concatenate([np.asarray(x, dtype=np.float64) for x in lams]).astype(self.dtype)
        owner = DeviceArray.from_host(self.ctx, flat, self.dtype)
        views, off = [], 0
        for x in lams:
            v = DeviceArray(self.ctx, (len(x),), self.dtype, ptr=owner.ptr + off * self.dtype.itemsize, owner=False)
            v._owner = owner
            views.append(v)
            off += len(x)
        return views

    @staticmethod
    def product_state(ctx: Context, vectors, dtype=np.complex128):
        """|v_1 ... v_n> with every Lambda = [1] (the norm sits in B_0)."""
        vs = [np.asarray(v, dtype=np.complex128).reshape(1, -1, 1) for v in vectors]
        nrm = [float(np.linalg.norm(v)) for v in vs]
        vs = [v / x for v, x in zip(vs, nrm)]
        vs[0] = vs[0] * float(np.prod(nrm))
        B = [DeviceArray.from_host(ctx, v, dtype) for v in vs]
        return VidalMPS(ctx, B, [np.ones(1) for _ in range(len(vs) - 1)], dtype)

    @staticmethod
    def from_mps(psi: MPS):
        """Right-canonicalise (centre at site 0), then one left-to-right SVD
        sweep: X_i = Lambda_{i-1} Bt_i = U S V^H with Bt_0 = A_0 and
        Bt_{i+1} = V^H A_{i+1}; B_i = Bt_i V, Lambda_i = S (nonzero part)."""
        p = psi.copy()
        p.canonize_(0)
        ctx, dt = p.ctx, p.dtype
        v = VidalMPS.__new__(VidalMPS)
        v.ctx, v.dtype, v.code = ctx, dt, DTYPES[dt]
        v._plans = {}
        v._one = DeviceArray.from_host(ctx, np.ones((), dtype=dt), dt)
        v.flops = {"svd": 0.0, "contract": 0.0}
        B, lam = [], []
        bt = p.sites[0]
        prev = None
        for i in range(p.n - 1):
            l, d, r = bt.shape
            X = bt if prev is None else p._einsum(prev, [1], bt, [1, 2, 3], bt.shape, [1, 2, 3])
            _, S, Vh, k = p._svd(X, l * d, r, 0, 0.0)
            k = int(np.count_nonzero(S[:k] > 0.0))
            Vv = Vh if k == Vh.shape[0] else _leading_rows(v, Vh, k)
            B.append(p._einsum(bt, [1, 2, 3], Vv, [4, 3], (l, d, k), [1, 2, 4], 0, 1))
            lam.append(np.asarray(S[:k], dtype=np.float64))
            nxt = p.sites[i + 1]
            bt = p._einsum(Vv, [4, 3], nxt, [3, 5, 6], (k, nxt.shape[1], nxt.shape[2]), [4, 5, 6])
            prev = DeviceArray.from_host(ctx, lam[-1].astype(dt), dt)
        B.append(bt)
        v.B = B
        v.lam_host = lam
        v.lam = v._upload_lams(lam)
        return v

    # ---- views ----------------------------------------------------------
    @property
    def n(self):
        return len(self.B)

    def bondsizes(self):
        return [b.shape[2] for b in self.B[:-1]]

    def to_mps(self) -> MPS:
        """The B sites are right-canonical: an MPS with its centre at site 0."""
        return MPS(self.ctx, [b.copy() for b in self.B], form=("mixed", 0, 0), dtype=self.dtype)

    # ---- one brickwork layer ------------------------------------------------
    def _plan(self, key, problems, conj=None):
        # cached on the context: every sweep (and every VidalMPS) with the
        # same layer shapes reuses the plans
        key = ("vidal", self.dtype.str) + key
        p = self.ctx.plan_cache.get(key)
        if p is None:
            p = GroupedPlan(self.ctx, problems, self.dtype, L.PREC_NATIVE, conj)
            self.ctx.plan_cache[key] = p
        return p

    def apply_layer_(self, gates, bonds, maxdim=None, threshold=None):
        """Apply the two-site gates ``gates[j]`` (4x4, site a as the most
        significant digit, U4[d*oa+ob, d*ia+ib] as in simple_update!) on the
        disjoint bonds ``bonds[j]`` = (a, a+1).  Returns the kept counts."""
        ctx, dt = self.ctx, self.dtype
        P = len(bonds)
        if P == 0:
            return []
        if len(set(bonds)) != P or any(abs(a - b) < 2 for a in bonds for b in bonds if a != b):
            raise ValueError("a layer's bonds must be disjoint")
        md = 0 if maxdim is None else int(maxdim)
        thr = 0.0 if threshold is None else float(threshold)
        Gs = []
        for U in gates:
            if isinstance(U, DeviceArray):
                Gs.append(U)
            else:
                d = self.B[0].shape[1]
                Gs.append(DeviceArray.from_host(ctx, _gate_tensor(np.asarray(U), d), dt))
        shp = []
        for a in bonds:
            l, d1, b = self.B[a].shape
            _, d2, r = self.B[a + 1].shape
            shp.append((l, d1, b, d2, r))
        # 1. theta = B_a B_{a+1}
        th = [DeviceArray(ctx, (l, d1, d2, r), dt) for (l, d1, b, d2, r) in shp]
        p1 = self._plan(("theta", tuple(shp)), [((l, d1, b), "lxb", (b, d2, r), "byr", (l, d1, d2, r), "lxyr")
                                                 for (l, d1, b, d2, r) in shp])
        p1.exec([x for j, a in enumerate(bonds) for x in (self.B[a].ptr, self.B[a + 1].ptr, th[j].ptr)])
        # 2. theta~ = theta . G   (G[oa, ob, ia, ib])
        tt = [DeviceArray(ctx, (l, d1, d2, r), dt) for (l, d1, b, d2, r) in shp]
        p2 = self._plan(("gate", tuple(shp)), [((l, d1, d2, r), "lxyr", (d1, d2, d1, d2), "pqxy", (l, d1, d2, r),
                                                "lpqr") for (l, d1, b, d2, r) in shp])
        p2.exec([x for j in range(P) for x in (th[j].ptr, Gs[j].ptr, tt[j].ptr)])
        # 3. X = Lambda_{a-1} (.) theta~
        idx = [j for j, a in enumerate(bonds) if a > 0]
        X = list(tt)
        if idx:
            for j in idx:
                X[j] = th[j]  # theta is no longer needed: reuse its memory
            p3 = self._plan(("lam", tuple(shp[j] for j in idx)),
                            [((shp[j][0],), "l", shp[j][:2] + shp[j][3:], "lpqr", shp[j][:2] + shp[j][3:], "lpqr")
                             for j in idx])
            p3.exec([x for j in idx for x in (self.lam[bonds[j] - 1].ptr, tt[j].ptr, X[j].ptr)])
        # 4. grouped SVD of X as (l d) x (d r)
        ms = np.array([l * d1 for (l, d1, b, d2, r) in shp], dtype=np.int64)
        ns = np.array([d2 * r for (l, d1, b, d2, r) in shp], dtype=np.int64)
        kmax = np.minimum(ms, ns)
        # U is never used (B_a is formed from theta~ and V): no U pointers, so
        # the SVD of an m <= n problem runs on X^H and accumulates no rotations
        Vs = [DeviceArray(ctx, (int(k), int(n)), dt) for k, n in zip(kmax, ns)]
        Ss = [ctx.malloc(int(k) * 8) for k in kmax]
        kd = ctx.malloc(P * 8)
        mds = np.full(P, md, dtype=np.int64)
        Pp = C.c_void_p
        try:
            L.check(L.lib.tenet_grouped_svd(
                ctx.h, None, self.code, P, ms.ctypes.data_as(C.POINTER(C.c_int64)),
                ns.ctypes.data_as(C.POINTER(C.c_int64)), (Pp * P)(*[x.ptr for x in X]),
                (Pp * P)(*([None] * P)), (Pp * P)(*Ss), (Pp * P)(*[v.ptr for v in Vs]),
                mds.ctypes.data_as(C.POINTER(C.c_int64)), thr, Pp(kd), None), "grouped_svd")
            ks = np.zeros(P, dtype=np.int64)
            L.check(L.lib.tenet_memcpy_d2h(ctx.h, ks.ctypes.data_as(Pp), Pp(kd), ks.nbytes, None))
            svals = []
            for j in range(P):
                s = np.zeros(int(ks[j]))
                if len(s):
                    L.check(L.lib.tenet_memcpy_d2h(ctx.h, s.ctypes.data_as(Pp), Pp(Ss[j]), s.nbytes, None))
                svals.append(s)
        finally:
            for p in Ss:
                ctx.free(p)
            ctx.free(kd)
        for (l, d1, b, d2, r), m, n in zip(shp, ms, ns):
            self.flops["svd"] += svd_flops(int(m), int(n))
            self.flops["contract"] += 8.0 * l * d1 * b * d2 * r + 8.0 * l * d1 * d2 * r * d1 * d2
        # 5. B_a <- theta~ conj(V^H[:k])^T ; B_{a+1} <- V^H[:k] (strided view of V^H)
        newA = [DeviceArray(ctx, (l, d1, int(k)), dt) for (l, d1, b, d2, r), k in zip(shp, ks)]
        newB = [DeviceArray(ctx, (int(k), d2, r), dt) for (l, d1, b, d2, r), k in zip(shp, ks)]
        probs, conj, ptrs = [], [], []
        for j, ((l, d1, b, d2, r), k, km) in enumerate(zip(shp, ks, kmax)):
            k, km = int(k), int(km)
            vst = (1, km, km * d2)
            probs.append(((l, d1, d2, r), "lpqr", (k, d2, r), "kqr", (l, d1, k), "lpk", None, vst))
            conj.append((0, 1))
            ptrs += [tt[j].ptr, Vs[j].ptr, newA[j].ptr]
            probs.append(((k, d2, r), "kqr", (), "", (k, d2, r), "kqr", vst, None))
            conj.append((0, 0))
            ptrs += [Vs[j].ptr, self._one.ptr, newB[j].ptr]
        key = ("out", tuple(shp), tuple(int(k) for k in ks))
        self._plan(key, probs, conj).exec(ptrs)
        self.flops["contract"] += sum(8.0 * l * d1 * d2 * r * int(k) for (l, d1, b, d2, r), k in zip(shp, ks))
        for j, a in enumerate(bonds):
            self.B[a] = newA[j]
            self.B[a + 1] = newB[j]
            self.lam_host[a] = svals[j]
        fresh = self._upload_lams([svals[j] for j in range(P)])
        for j, a in enumerate(bonds):
            self.lam[a] = fresh[j]
        del Vs, th, tt, X
        return [int(k) for k in ks]

    def close(self):
        """Plans live on the context (Context.close releases them)."""
        self._plans = {}


def _leading_rows(v: VidalMPS, M: DeviceArray, k: int) -> DeviceArray:
    """Contiguous copy of the first k rows of a column-major matrix."""
    m, n = M.shape
    out = DeviceArray(v.ctx, (k, n), v.dtype)
    da, db, dc = tdesc((k, n), [1, 2], [1, m]), tdesc((), []), tdesc((k, n), [1, 2])
    L.check(L.lib.tenet_einsum(v.ctx.h, None, v.code, L.PREC_NATIVE, C.byref(da), C.c_void_p(M.ptr), C.byref(db),
                               C.c_void_p(v._one.ptr), C.byref(dc), C.c_void_p(out.ptr), 0, 0), "rows")
    return out


def brickwork_layers(n, depth):
    """The brickwork schedule as layers of disjoint bonds (0-based left sites):
    odd layers (1,2),(3,4),...; even layers (2,3),(4,5),..."""
    return [list(range(0, n - 1, 2)) if layer % 2 == 0 else list(range(1, n - 1, 2)) for layer in range(depth)]


def tebd_vidal_sweep(ctx: Context, n, layers, gate_views, maxdim, threshold, psi=None, dtype=np.complex128):
    """Vidal-gauge TEBD from |0...0> (or ``psi``): gate_views consumed in
    order, one per bond entry of ``layers``."""
    v = psi if psi is not None else VidalMPS.product_state(ctx, [[1.0, 0.0]] * n, dtype)
    g = 0
    for bonds in layers:
        v.apply_layer_(gate_views[g:g + len(bonds)], bonds, maxdim, threshold)
        g += len(bonds)
    return v
