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
a layer as three grouped launches (two contractions and the fused gate) and
one grouped SVD:

  1. theta_k = B_a B_{a+1}                 (contract the bond)
  2. theta~_k = theta_k . G_k              (contract the gate inputs)
  3. X_k = Lambda_{a-1} (.) theta~_k       (Hadamard; at a = 0 X = theta~;
                                            2 and 3 in one tenet_grouped_gate launch)
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
from .device import DTYPES, Context, DeviceArray, carve, tdesc
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
        self._graphs = {}  # (parity, layer signature) -> _LayerGraph
        self._lam_bufs = None  # fixed per-bond Lambda buffers once layers are captured
        self._lam_stale = False  # lam_host behind the device Lambdas (captured layers ran)
        self._kept = {}  # (bonds, shapes, maxdim, threshold) -> the kept counts of the last such layer

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
        v._graphs = {}
        v._lam_bufs = None
        v._lam_stale = False
        v._kept = {}
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
        """The cached grouped plan for ``key``; ``problems`` (a list, or a
        callable returning (problems, conj)) is only built on a miss."""
        # cached on the context: every sweep (and every VidalMPS) with the
        # same layer shapes reuses the plans
        key = ("vidal", self.dtype.str) + key
        p = self.ctx.plan_cache.get(key)
        if p is None:
            if callable(problems):
                problems, conj = problems()
            p = GroupedPlan(self.ctx, problems, self.dtype, L.PREC_NATIVE, conj)
            self.ctx.plan_cache[key] = p
        return p

    def _out_step(self, shp, kmax, ks, tt, Vs):
        """Step 5 of a layer for kept counts ``ks``: the new sites (one pool
        block) and the grouped plan B_a <- theta~ conj(V^H[:k])^T,
        B_{a+1} <- V^H[:k] (a strided view of V^H), with its pointer list."""
        ctx, dt, P = self.ctx, self.dtype, len(shp)
        sites = carve(ctx, [(l, d1, k) for (l, d1, b, d2, r), k in zip(shp, ks)] +
                      [(k, d2, r) for (l, d1, b, d2, r), k in zip(shp, ks)], dt)
        newA, newB = sites[:P], sites[P:]

        def problems():
            probs, conj = [], []
            for (l, d1, b, d2, r), k, km in zip(shp, ks, kmax):
                km = int(km)
                vst = (1, km, km * d2)
                probs.append(((l, d1, d2, r), "lpqr", (k, d2, r), "kqr", (l, d1, k), "lpk", None, vst))
                conj.append((0, 1))
                probs.append(((k, d2, r), "kqr", (), "", (k, d2, r), "kqr", vst, None))
                conj.append((0, 0))
            return probs, conj

        plan = self._plan(("out", tuple(shp), ks), problems)
        one = self._one.ptr
        ptrs = [x for j in range(P) for x in (tt[j].ptr, Vs[j].ptr, newA[j].ptr, Vs[j].ptr, one, newB[j].ptr)]
        return newA, newB, plan, ptrs

    def apply_layer_(self, gates, bonds, maxdim=None, threshold=None):
        """Apply the two-site gates ``gates[j]`` (4x4, site a as the most
        significant digit, U4[d*oa+ob, d*ia+ib] as in simple_update!) on the
        disjoint bonds ``bonds[j]`` = (a, a+1).  Returns the kept counts.

        The host waits on the device once per layer, for the kept counts;
        everything else is queued behind.  While the counts come back the
        output step is prepared for the counts the same layer kept last time
        (or min(kmax, maxdim)), so the stream idles only for the check."""
        ctx, dt = self.ctx, self.dtype
        P = len(bonds)
        if P == 0:
            return []
        sb = sorted(bonds)
        if any(y - x < 2 for x, y in zip(sb, sb[1:])):
            raise ValueError("a layer's bonds must be disjoint")
        if list(bonds) != sb:
            # the gates commute: run them in site order, so that every layer
            # lays its operands out in the same order and the bond matrices
            # of equal shape sit evenly spaced (one strided-batched GEMM)
            order = sorted(range(P), key=lambda j: bonds[j])
            ks = self.apply_layer_([gates[j] for j in order], sb, maxdim, threshold)
            out = [0] * P
            for pos, j in enumerate(order):
                out[j] = ks[pos]
            return out
        # an eager layer replaces sites and Lambdas: captured layers bound to
        # the old buffers are dropped
        self._drop_graphs()
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
        shp = tuple(shp)
        # 1. theta = B_a B_{a+1}
        th = carve(ctx, [(l, d1, d2, r) for (l, d1, b, d2, r) in shp], dt)
        p1 = self._plan(("theta", shp), lambda: ([((l, d1, b), "lxb", (b, d2, r), "byr", (l, d1, d2, r), "lxyr")
                                                  for (l, d1, b, d2, r) in shp], None))
        p1.exec([x for j, a in enumerate(bonds) for x in (self.B[a].ptr, self.B[a + 1].ptr, th[j].ptr)])
        # 2. theta~ = theta . G (G[oa, ob, ia, ib]) and 3. X = Lambda_{a-1} (.) theta~,
        # one launch; X in theta's memory (theta is no longer needed)
        tt = carve(ctx, [(l, d1, d2, r) for (l, d1, b, d2, r) in shp], dt)
        X = [th[j] if a > 0 else tt[j] for j, a in enumerate(bonds)]
        _gate_step(self, shp, bonds, th, Gs, tt)
        # 4. grouped SVD of X as (l d) x (d r)
        ms = np.array([l * d1 for (l, d1, b, d2, r) in shp], dtype=np.int64)
        ns = np.array([d2 * r for (l, d1, b, d2, r) in shp], dtype=np.int64)
        kmax = np.minimum(ms, ns)
        # U is never used (B_a is formed from theta~ and V): no U pointers, so
        # the SVD of an m <= n problem runs on X^H and accumulates no rotations
        Vs = carve(ctx, [(int(k), int(n)) for k, n in zip(kmax, ns)], dt)
        # the kept counts and every S in one buffer: one download
        s_off = np.concatenate([[0], np.cumsum(kmax)]).astype(np.int64)
        nk = P + int(s_off[-1])
        kd = ctx.malloc(nk * 8)
        sbase = kd + P * 8
        mds = np.full(P, md, dtype=np.int64)
        Pp = C.c_void_p
        lam_off = (nk * 8 + 255) // 256 * 256
        stg = ctx.staging(lam_off + int(s_off[-1]) * dt.itemsize)
        try:
            L.check(L.lib.tenet_grouped_svd(
                ctx.h, None, self.code, P, ms.ctypes.data_as(C.POINTER(C.c_int64)),
                ns.ctypes.data_as(C.POINTER(C.c_int64)), (Pp * P)(*[x.ptr for x in X]),
                (Pp * P)(*([None] * P)), (Pp * P)(*[sbase + 8 * int(o) for o in s_off[:-1]]),
                (Pp * P)(*[v.ptr for v in Vs]), mds.ctypes.data_as(C.POINTER(C.c_int64)), thr, Pp(kd), None),
                "grouped_svd")
            # the download is queued behind the SVD's last kernels
            L.check(L.lib.tenet_memcpy_async(ctx.h, Pp(stg.ptr), Pp(kd), C.c_size_t(nk * 8), 1, None), "d2h")
            guess = self._kept.get((tuple(bonds), shp, md, thr))
            if guess is None:
                guess = tuple(int(min(k, md)) if md > 0 else int(k) for k in kmax)
            out = self._out_step(shp, kmax, guess, tt, Vs)
            ctx.sync()
            host = stg.view(np.float64, nk)
            ks = tuple(int(k) for k in host[:P].view(np.int64))
            svals = [host[P + int(s_off[j]):P + int(s_off[j]) + ks[j]].copy() for j in range(P)]
        finally:
            ctx.free(kd)
        if ks != guess:
            out = self._out_step(shp, kmax, ks, tt, Vs)
        self._kept[(tuple(bonds), shp, md, thr)] = ks
        newA, newB, plan, ptrs = out
        plan.exec(ptrs)
        for (l, d1, b, d2, r), m, n, k in zip(shp, ms, ns, ks):
            self.flops["svd"] += svd_flops(int(m), int(n))
            self.flops["contract"] += 8.0 * l * d1 * b * d2 * r + 8.0 * l * d1 * d2 * r * d1 * d2 + \
                8.0 * l * d1 * d2 * r * k
        for j, a in enumerate(bonds):
            self.B[a] = newA[j]
            self.B[a + 1] = newB[j]
            self.lam_host[a] = svals[j]
        # the new Lambdas: complex copies of the kept S, uploaded from the
        # staging buffer behind the output step (no wait)
        tot = sum(ks)
        if tot:
            lam = stg.view(dt, tot, offset=lam_off)
            lam[:] = np.concatenate(svals)
            owner = DeviceArray(ctx, (tot,), dt)
            L.check(L.lib.tenet_memcpy_async(ctx.h, Pp(owner.ptr), Pp(stg.ptr + lam_off),
                                             C.c_size_t(tot * dt.itemsize), 0, None), "h2d")
            off = 0
            for j, a in enumerate(bonds):
                v = DeviceArray(ctx, (ks[j],), dt, ptr=owner.ptr + off * dt.itemsize, owner=False)
                v._owner = owner
                self.lam[a] = v
                off += ks[j]
        else:
            for j, a in enumerate(bonds):
                self.lam[a] = DeviceArray(ctx, (0,), dt)
        del Vs, th, tt, X
        return list(ks)

    # ---- graph-captured layers -------------------------------------------------
    def layer_signature(self, bonds):
        """Shapes a layer on ``bonds`` reads: (l, d1, b, d2, r) per gate."""
        return tuple((self.B[a].shape[0], self.B[a].shape[1], self.B[a].shape[2], self.B[a + 1].shape[1],
                      self.B[a + 1].shape[2]) for a in bonds)

    def _stabilise_lams(self):
        """Give every Lambda a fixed device buffer (capacity d * max bond),
        the current values at its start; captured layers write S there."""
        if self._lam_bufs is not None:
            return
        ctx, dt = self.ctx, self.dtype
        cap = max(2, max(b.shape[1] for b in self.B) * max(self.bondsizes() + [1]))
        bufs = []
        for a, lam in enumerate(self.lam):
            k = lam.shape[0]
            buf = DeviceArray(ctx, (max(cap, k),), dt)
            L.check(L.lib.tenet_memcpy_d2d(ctx.h, C.c_void_p(buf.ptr), C.c_void_p(lam.ptr), k * dt.itemsize, None),
                    "d2d")
            self.lam[a] = buf.view((k,))
            bufs.append(buf)
        self._lam_bufs = bufs

    def _drop_graphs(self):
        for g in self._graphs.values():
            g.close()
        self._graphs = {}
        if self._lam_stale:
            self.refresh_lams()
        self._lam_bufs = None

    def refresh_lams(self):
        """Download the device Lambdas into lam_host (after captured layers)."""
        for a, lam in enumerate(self.lam):
            self.lam_host[a] = lam.to_host().real.astype(np.float64).copy()
        self._lam_stale = False

    def apply_layer_graph_(self, gates, bonds, maxdim=None, threshold=None):
        """The same layer as apply_layer_, replayed from a HIP graph captured
        for the layer's shapes (once per parity and shape signature).  Valid
        in the steady state, where every gate keeps the bond dimension it
        found (the caller establishes that with an eager layer of the same
        signature); the graph writes sites, Lambda and the SVD's S in place.
        After the replay the kept counts are read back: a gate that kept
        fewer (its threshold cut deeper) has its sites and Lambda cut to k,
        as the eager layer would, and the graphs are dropped.  Returns the
        kept counts."""
        sig = self.layer_signature(bonds)
        key = (tuple(bonds), sig, maxdim, threshold)
        g = self._graphs.get(key)
        if g is None:
            self._stabilise_lams()
            g = _LayerGraph(self, list(bonds), sig, maxdim, threshold)
            self._graphs[key] = g
        ks = g.replay(gates)
        self._lam_stale = True
        for (l, d1, b, d2, r), m, n in zip(sig, g.ms, g.ns):
            self.flops["svd"] += svd_flops(int(m), int(n))
            self.flops["contract"] += 8.0 * l * d1 * b * d2 * r + 8.0 * l * d1 * d2 * r * d1 * d2 + \
                8.0 * l * d1 * d2 * r * b
        short = [j for j, (k, sh) in enumerate(zip(ks, sig)) if k != sh[2]]
        if short:
            for j in short:
                a, k = bonds[j], int(ks[j])
                l, d1, b, d2, r = sig[j]
                if k > b:
                    raise RuntimeError("captured Vidal layer: a gate kept more singular values than its bond holds")
                self.B[a] = self.B[a].view((l, d1, k)).copy()
                self.B[a + 1] = _leading_rows(self, self.B[a + 1].view((b, d2 * r)), k).view((k, d2, r))
                self.lam[a] = self.lam[a].view((k,))
            self._drop_graphs()
        return [int(k) for k in ks]

    def close(self):
        """Plans live on the context (Context.close releases them)."""
        self._drop_graphs()
        self._plans = {}


def _gate_dim(shp):
    """The common physical dimension of a layer's bonds when the fused gate
    launch applies (every d1 = d2 = d with 2 <= d <= 4), else None."""
    ds = {s[1] for s in shp} | {s[3] for s in shp}
    d = next(iter(ds))
    return d if len(ds) == 1 and 2 <= d <= 4 else None


def _gate_fused(v: VidalMPS, shp, bonds, th_ptrs, gate_ptrs, tt_ptrs):
    """theta~_j = theta_j . G_j into tt_ptrs[j] and, for a > 0, X_j =
    Lambda_{a-1} (.) theta~_j into th_ptrs[j] (in place), one
    tenet_grouped_gate launch (capturable: no table upload)."""
    ctx, P = v.ctx, len(bonds)
    Pp = C.c_void_p
    chil = np.array([s[0] for s in shp], dtype=np.int64)
    chir = np.array([s[4] for s in shp], dtype=np.int64)
    lam = [v.lam[a - 1].ptr if a > 0 else None for a in bonds]
    xo = [th_ptrs[j] if a > 0 else None for j, a in enumerate(bonds)]
    L.check(L.lib.tenet_grouped_gate(
        ctx.h, None, v.code, P, _gate_dim(shp), chil.ctypes.data_as(C.POINTER(C.c_int64)),
        chir.ctypes.data_as(C.POINTER(C.c_int64)), (Pp * P)(*th_ptrs), (Pp * P)(*gate_ptrs), (Pp * P)(*lam),
        (Pp * P)(*tt_ptrs), (Pp * P)(*xo)), "grouped_gate")


def _gate_step(v: VidalMPS, shp, bonds, th, gates, tt):
    """Steps 2 and 3 of a layer: the fused launch, or (bonds of mixed or
    larger physical dimensions) the two grouped contractions."""
    P = len(bonds)
    if _gate_dim(shp) is not None:
        _gate_fused(v, shp, bonds, [x.ptr for x in th], [g.ptr for g in gates], [x.ptr for x in tt])
        return
    p2 = v._plan(("gate", shp), lambda: ([((l, d1, d2, r), "lxyr", (d1, d2, d1, d2), "pqxy", (l, d1, d2, r), "lpqr")
                                          for (l, d1, b, d2, r) in shp], None))
    p2.exec([x for j in range(P) for x in (th[j].ptr, gates[j].ptr, tt[j].ptr)])
    idx = [j for j, a in enumerate(bonds) if a > 0]
    if idx:
        sub = tuple(shp[j] for j in idx)
        p3 = v._plan(("lam", sub), lambda: ([((s[0],), "l", s[:2] + s[3:], "lpqr", s[:2] + s[3:], "lpqr")
                                             for s in sub], None))
        p3.exec([x for j in idx for x in (v.lam[bonds[j] - 1].ptr, tt[j].ptr, th[j].ptr)])


def _leading_rows(v: VidalMPS, M: DeviceArray, k: int) -> DeviceArray:
    """Contiguous copy of the first k rows of a column-major matrix."""
    m, n = M.shape
    out = DeviceArray(v.ctx, (k, n), v.dtype)
    da, db, dc = tdesc((k, n), [1, 2], [1, m]), tdesc((), []), tdesc((k, n), [1, 2])
    L.check(L.lib.tenet_einsum(v.ctx.h, None, v.code, L.PREC_NATIVE, C.byref(da), C.c_void_p(M.ptr), C.byref(db),
                               C.c_void_p(v._one.ptr), C.byref(dc), C.c_void_p(out.ptr), 0, 0), "rows")
    return out


class _LayerGraph:
    """One brickwork layer at fixed shapes as a HIP graph: the contraction
    launches of apply_layer_ (same grouped plans, on device pointer tables),
    the grouped SVD bound by tenet_grouped_svd_plan writing S as complex
    values straight into each bond's Lambda buffer, and the site updates
    written in place (the steady state keeps every shape).  The layer's
    gates are copied into a staging buffer before each replay."""

    def __init__(self, v: VidalMPS, bonds, sig, maxdim, threshold):
        ctx, dt = v.ctx, v.dtype
        self.ctx, self.v, self.bonds = ctx, v, bonds
        P = len(bonds)
        self.P = P
        md = 0 if maxdim is None else int(maxdim)
        thr = 0.0 if threshold is None else float(threshold)
        es = dt.itemsize
        d0 = sig[0][1]
        self.gsize = d0 ** 4
        self.stage = DeviceArray(ctx, (P * self.gsize,), dt)
        shp = list(sig)
        th = [DeviceArray(ctx, (l, d1, d2, r), dt) for (l, d1, b, d2, r) in shp]
        tt = [DeviceArray(ctx, (l, d1, d2, r), dt) for (l, d1, b, d2, r) in shp]
        self.ms = np.array([l * d1 for (l, d1, b, d2, r) in shp], dtype=np.int64)
        self.ns = np.array([d2 * r for (l, d1, b, d2, r) in shp], dtype=np.int64)
        kmax = np.minimum(self.ms, self.ns)
        Vs = [DeviceArray(ctx, (int(k), int(n)), dt) for k, n in zip(kmax, self.ns)]
        for j, a in enumerate(bonds):
            if v._lam_bufs[a].size < kmax[j]:
                raise RuntimeError("Lambda buffer below the layer's kmax")
        self._keep = [th, tt, Vs]
        # 1. theta = B_a B_{a+1}; 2. theta~ = theta . G; 3. X = Lambda_{a-1} (.) theta~ (theta's memory)
        p1 = v._plan(("theta", tuple(shp)), [((l, d1, b), "lxb", (b, d2, r), "byr", (l, d1, d2, r), "lxyr")
                                             for (l, d1, b, d2, r) in shp])
        t1 = [x for j, a in enumerate(bonds) for x in (v.B[a].ptr, v.B[a + 1].ptr, th[j].ptr)]
        gate_ptrs = [self.stage.ptr + j * self.gsize * es for j in range(P)]
        idx = [j for j, a in enumerate(bonds) if a > 0]
        X = list(tt)
        for j in idx:
            X[j] = th[j]
        fused = _gate_dim(tuple(shp)) is not None
        p2 = p3 = t2 = t3 = None
        if not fused:
            p2 = v._plan(("gate", tuple(shp)), [((l, d1, d2, r), "lxyr", (d1, d2, d1, d2), "pqxy", (l, d1, d2, r),
                                                 "lpqr") for (l, d1, b, d2, r) in shp])
            t2 = [x for j in range(P) for x in (th[j].ptr, gate_ptrs[j], tt[j].ptr)]
            if idx:
                p3 = v._plan(("lam", tuple(shp[j] for j in idx)),
                             [((shp[j][0],), "l", shp[j][:2] + shp[j][3:], "lpqr", shp[j][:2] + shp[j][3:], "lpqr")
                              for j in idx])
                t3 = [x for j in idx for x in (v.lam[bonds[j] - 1].ptr, tt[j].ptr, X[j].ptr)]
        # 4. the grouped SVD, bound once: S as complex into the Lambda buffers
        self.kd = ctx.malloc(P * 8)
        Pp = C.c_void_p
        mds = np.full(P, md, dtype=np.int64)
        h = Pp()
        L.check(L.lib.tenet_grouped_svd_plan(
            ctx.h, v.code, P, self.ms.ctypes.data_as(C.POINTER(C.c_int64)),
            self.ns.ctypes.data_as(C.POINTER(C.c_int64)), (Pp * P)(*[x.ptr for x in X]), (Pp * P)(*([None] * P)),
            (Pp * P)(*[v._lam_bufs[a].ptr for a in bonds]), (Pp * P)(*[x.ptr for x in Vs]),
            mds.ctypes.data_as(C.POINTER(C.c_int64)), thr, Pp(self.kd), None, L.S_DTYPE, C.byref(h)),
            "grouped_svd_plan")
        self.svd = h
        # 5. B_a <- theta~ conj(V^H[:k])^T and B_{a+1} <- V^H[:k], in place (k = b)
        probs, conj, t5 = [], [], []
        for j, ((l, d1, b, d2, r), km) in enumerate(zip(shp, kmax)):
            k, km = int(b), int(km)
            vst = (1, km, km * d2)
            probs.append(((l, d1, d2, r), "lpqr", (k, d2, r), "kqr", (l, d1, k), "lpk", None, vst))
            conj.append((0, 1))
            t5 += [tt[j].ptr, Vs[j].ptr, v.B[bonds[j]].ptr]
            probs.append(((k, d2, r), "kqr", (), "", (k, d2, r), "kqr", vst, None))
            conj.append((0, 0))
            t5 += [Vs[j].ptr, v._one.ptr, v.B[bonds[j] + 1].ptr]
        p5 = v._plan(("out", tuple(shp), tuple(int(x[2]) for x in shp)), probs, conj)
        self.tables = []
        for t in (t1, t2, t3, t5):
            if t is None:
                self.tables.append(None)
                continue
            arr = np.asarray(t, dtype=np.uint64)
            dev = ctx.malloc(arr.nbytes)
            L.check(L.lib.tenet_memcpy_h2d(ctx.h, C.c_void_p(dev), arr.ctypes.data_as(C.c_void_p), arr.nbytes,
                                           None))
            self.tables.append(dev)
        with ctx.capture() as g:
            p1.exec_dev(self.tables[0])
            if fused:
                _gate_fused(v, tuple(shp), bonds, [x.ptr for x in th], gate_ptrs, [x.ptr for x in tt])
            else:
                p2.exec_dev(self.tables[1])
                if p3 is not None:
                    p3.exec_dev(self.tables[2])
            L.check(L.lib.tenet_grouped_svd_exec(self.svd, None), "grouped_svd_exec")
            p5.exec_dev(self.tables[3])
        self.graph = g

    def replay(self, gates):
        """Stage the layer's gates (one copy when they are consecutive in one
        buffer, as upload_gates lays them out), replay, and return the kept
        counts (synchronises)."""
        ctx, es = self.ctx, self.v.dtype.itemsize
        nb = self.gsize * es
        if all(g.ptr == gates[0].ptr + j * nb for j, g in enumerate(gates)):
            L.check(L.lib.tenet_memcpy_d2d(ctx.h, C.c_void_p(self.stage.ptr), C.c_void_p(gates[0].ptr),
                                           self.P * nb, None), "d2d")
        else:
            for j, g in enumerate(gates):
                L.check(L.lib.tenet_memcpy_d2d(ctx.h, C.c_void_p(self.stage.ptr + j * nb), C.c_void_p(g.ptr), nb,
                                               None), "d2d")
        self.graph.launch()
        ks = np.zeros(self.P, dtype=np.int64)
        L.check(L.lib.tenet_memcpy_d2h(ctx.h, ks.ctypes.data_as(C.c_void_p), C.c_void_p(self.kd), ks.nbytes, None))
        if ctx.graph_status() & 1:
            raise L.TenetError(L.TENET_ERR_NOCONV, "captured Vidal layer: a Jacobi SVD did not converge")
        return ks

    def close(self):
        if self.graph is not None:
            self.graph.close()
            self.graph = None
        if self.svd is not None:
            L.lib.tenet_grouped_svd_destroy(self.svd)
            self.svd = None
        for t in self.tables:
            if t:
                self.ctx.free(t)
        self.tables = []
        if self.kd:
            self.ctx.free(self.kd)
            self.kd = 0
        self._keep = []


def brickwork_layers(n, depth):
    """The brickwork schedule as layers of disjoint bonds (0-based left sites):
    odd layers (1,2),(3,4),...; even layers (2,3),(4,5),..."""
    return [list(range(0, n - 1, 2)) if layer % 2 == 0 else list(range(1, n - 1, 2)) for layer in range(depth)]


def tebd_vidal_sweep(ctx: Context, n, layers, gate_views, maxdim, threshold, psi=None, dtype=np.complex128,
                     graphs=False):
    """Vidal-gauge TEBD from |0...0> (or ``psi``): gate_views consumed in
    order, one per bond entry of ``layers``.  With ``graphs``, a layer whose
    shape signature an earlier eager layer left unchanged (the steady state,
    once the bonds have saturated) is replayed from a captured graph
    (VidalMPS.apply_layer_graph_); the results equal the eager sweep's."""
    v = psi if psi is not None else VidalMPS.product_state(ctx, [[1.0, 0.0]] * n, dtype)
    steady = set()
    g = 0
    for bonds in layers:
        gates = gate_views[g:g + len(bonds)]
        sig = (tuple(bonds), v.layer_signature(bonds))
        if graphs and sig in steady:
            v.apply_layer_graph_(gates, bonds, maxdim, threshold)
        else:
            ks = v.apply_layer_(gates, bonds, maxdim, threshold)
            if all(k == shp[2] for k, shp in zip(ks, sig[1])):
                steady.add(sig)
        g += len(bonds)
    if v._lam_stale:
        v.refresh_lams()
    return v
