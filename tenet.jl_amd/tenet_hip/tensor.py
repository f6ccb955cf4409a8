"""Muscle-level Tensor API on device tensors: the functions Tenet.jl calls for
its hot path (binary_einsum, tensor_qr_thin, tensor_svd_thin, simple_update),
with the same argument meaning and label semantics, executed by
libtenet_hip.so.

References (Tenet.jl v0.10.3 call sites; Muscle.jl itself is not vendored):
  binary_einsum    src/Operations/canonize.jl:44,61; overlap.jl:42-47
  tensor_qr_thin   src/Operations/canonize.jl:58
  tensor_svd_thin  src/Operations/canonize.jl:41,97; evolve.jl:62
  simple_update    src/Operations/simple_update.jl:53-64
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .device import DTYPES, Context, DeviceArray, tdesc


class Tensor:
    """Dense device tensor whose axes carry hashable labels (Muscle's Tensor)."""

    def __init__(self, data: DeviceArray, labels):
        self.data = data
        self.labels = list(labels)
        if len(self.labels) != len(data.shape):
            raise ValueError("rank/label mismatch")
        if len(set(self.labels)) != len(self.labels):
            raise ValueError("repeated label")

    @property
    def ctx(self) -> Context:
        return self.data.ctx

    @property
    def shape(self):
        return self.data.shape

    @staticmethod
    def from_host(ctx: Context, array, labels, dtype=None):
        return Tensor(DeviceArray.from_host(ctx, array, dtype), labels)

    def to_host(self):
        return self.data.to_host()

    def size(self, label):
        return self.shape[self.labels.index(label)]

    def rename(self, mapping):
        return Tensor(self.data, [mapping.get(l, l) for l in self.labels])

    def __repr__(self):
        return f"Tensor(shape={self.shape}, labels={self.labels})"


class _Codes:
    """Map arbitrary hashable labels to int64 codes for one call."""

    def __init__(self):
        self.m = {}

    def __call__(self, lab):
        if lab not in self.m:
            self.m[lab] = len(self.m) + 1
        return self.m[lab]


def _one(ctx: Context, dtype):
    key = ("_one", np.dtype(dtype).str)
    cache = ctx.__dict__.setdefault("_consts", {})
    if key not in cache:
        cache[key] = DeviceArray.from_host(ctx, np.ones((), dtype=dtype).reshape(()), dtype)
    return cache[key]


def permute(t: Tensor, labels, prec=L.PREC_NATIVE) -> Tensor:
    """Materialise ``t`` with its axes in ``labels`` order (a gather copy)."""
    labels = list(labels)
    if labels == t.labels:
        return t
    code = _Codes()
    one = _one(t.ctx, t.data.dtype)
    shape = [t.size(l) for l in labels]
    out = DeviceArray(t.ctx, shape, t.data.dtype)
    da = tdesc(t.shape, [code(l) for l in t.labels])
    db = tdesc((), [])
    dc = tdesc(shape, [code(l) for l in labels])
    L.check(L.lib.tenet_einsum(t.ctx.h, None, t.data.code, prec, C.byref(da), C.c_void_p(t.data.ptr), C.byref(db),
                               C.c_void_p(one.ptr), C.byref(dc), C.c_void_p(out.ptr), 0, 0), "permute")
    return Tensor(out, labels)


def binary_einsum(a: Tensor, b: Tensor, dims=None, prec=L.PREC_NATIVE, conj_a=False, conj_b=False,
                  out_labels=None) -> Tensor:
    """Pairwise contraction over ``dims`` (default: all shared labels).

    Shared labels not in ``dims`` are kept once and treated elementwise
    (dims=[] absorbs singular-value vectors, canonize.jl:44).  Output labels:
    a's labels not contracted, then b's labels not in a, unless
    ``out_labels`` gives another order."""
    if a.data.dtype != b.data.dtype:
        raise TypeError("operands must share a dtype")
    shared = [l for l in a.labels if l in b.labels]
    dims = shared if dims is None else list(dims)
    for d in dims:
        if d not in shared:
            raise ValueError(f"contracted label {d!r} not shared")
    out = [l for l in a.labels if l not in dims] + [l for l in b.labels if l not in a.labels]
    if out_labels is not None:
        if sorted(map(repr, out_labels)) != sorted(map(repr, out)):
            raise ValueError("out_labels must be a permutation of the output labels")
        out = list(out_labels)
    size = dict(zip(a.labels, a.shape))
    size.update(zip(b.labels, b.shape))
    shape = [size[l] for l in out]
    c = DeviceArray(a.ctx, shape, a.data.dtype)
    code = _Codes()
    da = tdesc(a.shape, [code(l) for l in a.labels])
    db = tdesc(b.shape, [code(l) for l in b.labels])
    dc = tdesc(shape, [code(l) for l in out])
    L.check(L.lib.tenet_einsum(a.ctx.h, None, a.data.code, prec, C.byref(da), C.c_void_p(a.data.ptr), C.byref(db),
                               C.c_void_p(b.data.ptr), C.byref(dc), C.c_void_p(c.ptr), int(conj_a), int(conj_b)),
            "binary_einsum")
    return Tensor(c, out)


def _matricize(A: Tensor, rows, cols):
    t = permute(A, list(rows) + list(cols))
    m = int(np.prod([A.size(l) for l in rows], dtype=np.int64))
    n = int(np.prod([A.size(l) for l in cols], dtype=np.int64))
    return t, m, n


def tensor_qr_thin(A: Tensor, inds_q, inds_r, ind_virtual):
    """Thin Householder QR; Q carries (inds_q..., ind_virtual), R carries
    (ind_virtual, inds_r...); k = min(m, n)."""
    t, m, n = _matricize(A, inds_q, inds_r)
    k = min(m, n)
    Q = DeviceArray(A.ctx, [A.size(l) for l in inds_q] + [k], A.data.dtype)
    R = DeviceArray(A.ctx, [k] + [A.size(l) for l in inds_r], A.data.dtype)
    L.check(L.lib.tenet_qr_thin(A.ctx.h, None, A.data.code, m, n, C.c_void_p(t.data.ptr), m, C.c_void_p(Q.ptr), m,
                                C.c_void_p(R.ptr), k), "tensor_qr_thin")
    return Tensor(Q, list(inds_q) + [ind_virtual]), Tensor(R, [ind_virtual] + list(inds_r))


def tensor_svd_thin(A: Tensor, inds_u, ind_s, inds_v=None, maxdim=None, threshold=None):
    """Thin SVD: U (inds_u..., ind_s), s (host float64, descending),
    V (ind_s, inds_v...).  ``inds_v`` defaults to the complement
    (evolve.jl:62).  With maxdim / threshold the factors are cut by the rule
    of compress.jl:46-59."""
    if inds_v is None:
        inds_v = [l for l in A.labels if l not in inds_u]
    t, m, n = _matricize(A, inds_u, inds_v)
    kmax = min(m, n)
    ctx = A.ctx
    U = DeviceArray(ctx, (m, kmax), A.data.dtype)
    Vh = DeviceArray(ctx, (kmax, n), A.data.dtype)
    S = np.zeros(kmax, dtype=np.float64)
    sdev = ctx.malloc(kmax * 8 + 64)
    try:
        L.check(L.lib.tenet_svd_thin(ctx.h, None, A.data.code, m, n, C.c_void_p(t.data.ptr), m, C.c_void_p(U.ptr), m,
                                     C.c_void_p(sdev), C.c_void_p(Vh.ptr), kmax, int(maxdim or 0),
                                     float(threshold or 0.0), C.c_void_p(sdev + kmax * 8), None), "tensor_svd_thin")
        kbuf = np.zeros(1, dtype=np.int64)
        L.check(L.lib.tenet_memcpy_d2h(ctx.h, S.ctypes.data_as(C.c_void_p), C.c_void_p(sdev), kmax * 8, None))
        L.check(L.lib.tenet_memcpy_d2h(ctx.h, kbuf.ctypes.data_as(C.c_void_p), C.c_void_p(sdev + kmax * 8), 8, None))
    finally:
        ctx.free(sdev)
    k = int(kbuf[0]) if (maxdim or threshold) else kmax
    ushape = [A.size(l) for l in inds_u]
    vshape = [A.size(l) for l in inds_v]
    Ut = Tensor(U.reshape((m, kmax)), ["__rows__", ind_s])
    Vt = Tensor(Vh.reshape((kmax, n)), [ind_s, "__cols__"])
    if k < kmax:
        Ut = _slice_cols(Ut, ind_s, k)
        Vt = _slice_rows(Vt, ind_s, k)
    Uo = Tensor(Ut.data.reshape(ushape + [k]), list(inds_u) + [ind_s])
    Vo = Tensor(Vt.data.reshape([k] + vshape), [ind_s] + list(inds_v))
    return Uo, S[:k].copy(), Vo


def _slice_cols(t: Tensor, label, k):
    """Keep the first k entries of the LAST axis (a contiguous prefix)."""
    assert t.labels[-1] == label
    shp = list(t.shape[:-1]) + [k]
    return Tensor(t.data.view(shp).copy(), t.labels)


def _slice_rows(t: Tensor, label, k):
    """Keep the first k entries of the FIRST axis (a strided gather)."""
    assert t.labels[0] == label
    code = _Codes()
    one = _one(t.ctx, t.data.dtype)
    src_shape = [k] + list(t.shape[1:])
    src_strides = t.data.strides_el
    out = DeviceArray(t.ctx, src_shape, t.data.dtype)
    da = tdesc(src_shape, [code(l) for l in t.labels], src_strides)
    db = tdesc((), [])
    dc = tdesc(src_shape, [code(l) for l in t.labels])
    L.check(L.lib.tenet_einsum(t.ctx.h, None, t.data.code, L.PREC_NATIVE, C.byref(da), C.c_void_p(t.data.ptr),
                               C.byref(db), C.c_void_p(one.ptr), C.byref(dc), C.c_void_p(out.ptr), 0, 0), "slice")
    return Tensor(out, t.labels)


def simple_update(A: Tensor, ia, B: Tensor, ib, bond, G: Tensor, oa, ob, maxdim=None, threshold=None,
                  absorb=L.ABSORB_EQUAL, prec=L.PREC_NATIVE):
    """Muscle.simple_update(A, ia, B, ib, bond, G, oa, ob; maxdim, absorb):
    theta = A.B over ``bond``; theta.G maps (ia, ib) to the gate outputs
    (oa, ob), renamed back to (ia, ib); SVD with rows (A's others, ia) and
    columns (ib, B's others); truncation; singular values absorbed.  Returns
    (A', B', s)."""
    a_other = [l for l in A.labels if l not in (ia, bond)]
    b_other = [l for l in B.labels if l not in (ib, bond)]
    Ap = permute(A, a_other + [ia, bond])
    Bp = permute(B, [bond, ib] + b_other)
    Gp = permute(G, [oa, ob, ia, ib])
    chil = int(np.prod([A.size(l) for l in a_other], dtype=np.int64))
    chir = int(np.prod([B.size(l) for l in b_other], dtype=np.int64))
    d1, d2, chi = A.size(ia), B.size(ib), A.size(bond)
    ctx = A.ctx
    m, n = chil * d1, d2 * chir
    kmax = min(m, n)
    Aout = DeviceArray(ctx, (m * kmax,), A.data.dtype)
    Bout = DeviceArray(ctx, (kmax * n,), A.data.dtype)
    sdev = ctx.malloc(kmax * 8)
    k = C.c_int64(0)
    try:
        L.check(L.lib.tenet_simple_update(ctx.h, None, A.data.code, prec, chil, d1, chi, d2, chir,
                                          C.c_void_p(Ap.data.ptr), C.c_void_p(Bp.data.ptr), C.c_void_p(Gp.data.ptr),
                                          C.c_void_p(Aout.ptr), C.c_void_p(Bout.ptr), C.c_void_p(sdev),
                                          int(maxdim or 0), float(threshold or 0.0), int(absorb), None, None,
                                          C.byref(k)), "simple_update")
        s = np.zeros(kmax)
        L.check(L.lib.tenet_memcpy_d2h(ctx.h, s.ctypes.data_as(C.c_void_p), C.c_void_p(sdev), kmax * 8, None))
    finally:
        ctx.free(sdev)
    k = int(k.value)
    a_shape = [A.size(l) for l in a_other] + [d1, k]
    b_shape = [k, d2] + [B.size(l) for l in b_other]
    An = Aout.view(a_shape)
    Bn = Bout.view(b_shape)
    return Tensor(An, a_other + [ia, bond]), Tensor(Bn, [bond, ib] + b_other), s[:k]


__all__ = ["Tensor", "binary_einsum", "tensor_qr_thin", "tensor_svd_thin", "simple_update", "permute", "DTYPES"]
