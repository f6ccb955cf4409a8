"""Grouped launches: many independent contractions (or thin SVDs) in one
persistent launch.

A ``GroupedPlan`` is built once from the problems' shapes and labels
(offset tables, per-problem tile shape, K chunking, cost-sorted work queue,
all in one device allocation) and executed many times on fresh operand
pointers.  Each problem has the semantics of ``binary_einsum``
(Muscle.binary_einsum at src/Operations/canonize.jl:44-61, evolve.jl:29-119):
labels in a and b but not in c are contracted, labels in all three are
batched.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .device import DTYPES, Context, DeviceArray, tdesc


def _codes(problems):
    m = {}

    def code(lab):
        if lab not in m:
            m[lab] = len(m) + 1
        return m[lab]

    return code


class GroupedPlan:
    """problems: sequence of (shape_a, labels_a, shape_b, labels_b,
    shape_c, labels_c) with column-major contiguous operands, or with an
    eighth and ninth entry: strides of a and b (elements).  conj: optional
    per-problem (conj_a, conj_b) pairs."""

    def __init__(self, ctx: Context, problems, dtype=np.complex64, prec=L.PREC_NATIVE, conj=None):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        self.code = DTYPES[self.dtype]
        self.count = len(problems)
        descs = (L.TDesc * max(3 * self.count, 1))()
        for i, pr in enumerate(problems):
            sa, la, sb, lb, sc, lc = pr[:6]
            st_a = pr[6] if len(pr) > 6 else None
            st_b = pr[7] if len(pr) > 7 else None
            code = _codes(None)
            descs[3 * i] = tdesc(sa, [code(x) for x in la], st_a)
            descs[3 * i + 1] = tdesc(sb, [code(x) for x in lb], st_b)
            descs[3 * i + 2] = tdesc(sc, [code(x) for x in lc])
        self.shapes = [(tuple(p[0]), tuple(p[2]), tuple(p[4])) for p in problems]
        h = C.c_void_p()
        if conj is None:
            L.check(L.lib.tenet_grouped_plan(ctx.h, self.code, prec, self.count, descs, C.byref(h)), "grouped_plan")
        else:
            flags = (C.c_int32 * max(2 * self.count, 1))(*[int(bool(x)) for pair in conj for x in pair])
            L.check(L.lib.tenet_grouped_plan_conj(ctx.h, self.code, prec, self.count, descs, flags, C.byref(h)),
                    "grouped_plan_conj")
        self.h = h

    def info(self):
        u, st = C.c_int64(), C.c_int64()
        shapes = (C.c_int64 * 4)()
        fl, by = C.c_double(), C.c_double()
        L.check(L.lib.tenet_grouped_plan_info(self.h, C.byref(u), C.byref(st), shapes, C.byref(fl), C.byref(by)))
        return {"units": u.value, "split_tiles": st.value, "tiles_64_32_16": list(shapes)[:3], "smallk": shapes[3],
                "flops": fl.value, "bytes": by.value}

    def exec(self, ptrs, stream=None):
        """ptrs: 3 * count device addresses (A, B, C per problem).  The
        launch is asynchronous on the stream (the addresses are copied on
        the call); later work on the same stream sees the results."""
        arr = (C.c_void_p * (3 * self.count))(*[int(p) for p in ptrs])
        L.check(L.lib.tenet_grouped_exec(self.h, stream, arr), "grouped_exec")

    def exec_dev(self, ptrs_dev: int, stream=None):
        """ptrs_dev: device address of 3 * count uint64 pointers; asynchronous."""
        L.check(L.lib.tenet_grouped_exec_dev(self.h, stream, C.c_void_p(int(ptrs_dev))), "grouped_exec_dev")

    def close(self):
        if getattr(self, "h", None) is not None:
            L.lib.tenet_grouped_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune_tiles(ctx: Context, dtype=np.complex64):
    """Measure the per-(tile, K step) cost of the 64x64, 32x32 and 16x16 tile
    shapes on the device; later plans choose shapes with these costs."""
    out = (C.c_double * 3)()
    L.check(L.lib.tenet_autotune_tiles(ctx.h, DTYPES[np.dtype(dtype)], out), "autotune_tiles")
    return list(out)


def grouped_svd(ctx: Context, mats, maxdims=None, threshold=0.0, want_s=True, want_k=True, want_disc=True,
                want_u=True):
    """Thin SVDs of device matrices (DeviceArray, column-major m x n) in one
    grouped launch with per-problem truncation (compress.jl:46-59 rule).
    Returns per problem (U (m x kmax), S (kmax,), Vh (kmax x n), k, discarded
    weight) with only the first k columns / entries / rows meaningful.
    want_s / want_k / want_disc / want_u = False pass NULL for that output
    (the returned entry is then None); without U, problems with
    64 <= n <= m <= 512 take the QR-preconditioned Jacobi."""
    cnt = len(mats)
    if cnt == 0:
        return []
    dt = mats[0].dtype
    code = DTYPES[dt]
    ms = np.array([a.shape[0] for a in mats], dtype=np.int64)
    ns = np.array([a.shape[1] for a in mats], dtype=np.int64)
    kmax = np.minimum(ms, ns)
    md = np.zeros(cnt, dtype=np.int64) if maxdims is None else np.asarray(maxdims, dtype=np.int64)
    # one arena each for the U's, the V^H's and the singular values (views per
    # problem), so the outputs cost three allocations and one download
    item = np.dtype(dt).itemsize
    u_off = np.concatenate([[0], np.cumsum(ms * kmax)])
    v_off = np.concatenate([[0], np.cumsum(kmax * ns)])
    s_off = np.concatenate([[0], np.cumsum(kmax)])
    u_arena = DeviceArray(ctx, (int(u_off[-1]) if want_u else 1,), dt)
    v_arena = DeviceArray(ctx, (int(v_off[-1]),), dt)
    Us, Vs = [], []
    for i in range(cnt):
        u = DeviceArray(ctx, (int(ms[i]), int(kmax[i])), dt, ptr=u_arena.ptr + item * int(u_off[i]), owner=False)
        u._owner = u_arena
        v = DeviceArray(ctx, (int(kmax[i]), int(ns[i])), dt, ptr=v_arena.ptr + item * int(v_off[i]), owner=False)
        v._owner = v_arena
        Us.append(u)
        Vs.append(v)
    s_all = ctx.malloc(int(s_off[-1]) * 8 + 64)
    kd = ctx.malloc(cnt * 8)
    dd = ctx.malloc(cnt * 8)
    try:
        P = C.c_void_p
        s_ptrs = (P * cnt)(*[s_all + 8 * int(o) for o in s_off[:-1]]) if want_s else None
        L.check(L.lib.tenet_grouped_svd(
            ctx.h, None, code, cnt, ms.ctypes.data_as(C.POINTER(C.c_int64)), ns.ctypes.data_as(C.POINTER(C.c_int64)),
            (P * cnt)(*[a.ptr for a in mats]), (P * cnt)(*[u.ptr if want_u else None for u in Us]), s_ptrs,
            (P * cnt)(*[v.ptr for v in Vs]), md.ctypes.data_as(C.POINTER(C.c_int64)), float(threshold),
            P(kd) if want_k else None, P(dd) if want_disc else None),
            "grouped_svd")
        kh = np.zeros(cnt, dtype=np.int64)
        dh = np.zeros(cnt)
        sh = np.zeros(int(s_off[-1]))
        L.check(L.lib.tenet_memcpy_d2h(ctx.h, kh.ctypes.data_as(P), P(kd), kh.nbytes, None))
        L.check(L.lib.tenet_memcpy_d2h(ctx.h, dh.ctypes.data_as(P), P(dd), dh.nbytes, None))
        L.check(L.lib.tenet_memcpy_d2h(ctx.h, sh.ctypes.data_as(P), P(s_all), sh.nbytes, None))
        return [(Us[i] if want_u else None, sh[s_off[i]:s_off[i + 1]].copy() if want_s else None, Vs[i],
                 int(kh[i]) if want_k else None, float(dh[i]) if want_disc else None) for i in range(cnt)]
    finally:
        ctx.free(s_all)
        ctx.free(kd)
        ctx.free(dd)
