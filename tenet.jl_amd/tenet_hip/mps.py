"""Device-resident MPS with Tenet.jl's canonical-form semantics.

Each site is a column-major (chi_l, d, chi_r) DeviceArray (chi_l = 1 on the
first site, chi_r = 1 on the last; Tenet stores the edge sites as 2-way
arrays, MPS.jl:63-69, and those size-1 legs are never treated as bonds).
Sites are 0-based here; Tenet's site"i" is site i-1.

Python names carry a trailing underscore where Tenet's carry a '!':
  canonize_        canonize!            src/Operations/canonize.jl:121-255
  canonize_bond_   canonize!(Bond)      canonize.jl:151-179, 221-246
  compress_        compress!/truncate!  src/Operations/compress.jl:13-72
  simple_update_   simple_update!(MPS)  src/Operations/simple_update.jl:82-92
  evolve_          evolve!(tn, ::Tensor) src/Operations/evolve.jl:140-143
  overlap          overlap(a, b) = <b|a>  src/Operations/overlap.jl:36-50
  norm / normalize_  norm.jl:17-37 / normalize.jl:15-37
  amplitudes       overlap(::ProductState, ::MPS), overlap.jl:22-32
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as L
from .device import DTYPES, Context, DeviceArray, tdesc


def qr_flops(m, n):
    p, q = max(m, n), min(m, n)
    return 4.0 * (4.0 * p * q * q - 4.0 * q ** 3 / 3.0)


def svd_flops(m, n):
    p, q = max(m, n), min(m, n)
    return 4.0 * (4.0 * p * p * q + 8.0 * p * q * q + 9.0 * q ** 3)


def _qr_absorb_fits(m, n, nb):
    """tenet_qr_absorb's shape range (the single-workgroup QR) and a neighbour
    small enough for its four product waves (a wider one contracts faster as
    a separate many-workgroup GEMM)."""
    return m <= 64 and n <= 64 and nb <= 128


def _gate_tensor(U4, d):
    """4x4 gate with site a as the most significant digit, U4[d*oa+ob, d*ia+ib]
    (test/unit/simple_update.jl:19-33) -> column-major G[oa, ob, ia, ib]."""
    return np.asfortranarray(np.asarray(U4).reshape(d, d, d, d))


class MPS:
    def __init__(self, ctx: Context, sites, form=None, dtype=np.complex128):
        self.ctx = ctx
        self.sites = list(sites)
        self.dtype = np.dtype(dtype)
        self.code = DTYPES[self.dtype]
        self.form = form if form is not None else ("mixed", 0, len(self.sites) - 1)  # MPS.jl:18
        self.lam = None  # host Lambda of a BondCanonical MPS
        self.prec = L.PREC_NATIVE
        # nominal work (real flops): complex MAC = 8; thin QR (p >= q) =
        # 4(4pq^2 - 4q^3/3); thin SVD (p >= q) = 4(4p^2q + 8pq^2 + 9q^3)
        self.flops = {"qr": 0.0, "svd": 0.0, "contract": 0.0}
        # after simple_update_ (AbsorbEqually) at bond (a, a+1): sites a = U sqrt(S),
        # a+1 = sqrt(S) V^H.  (a, site a, site a+1, sqrt(S), 1/sqrt(S)) while both
        # sites are unchanged; the next canonize! move across that bond is then a
        # diagonal rescaling (see _diag_move)
        self._gate_bond = None

    # ---- construction ---------------------------------------------------
    @staticmethod
    def from_host(ctx: Context, arrays_lor, dtype=np.complex128, form=None):
        """From (chi_l, d, chi_r) host arrays."""
        sites = [DeviceArray.from_host(ctx, np.asarray(a).reshape(np.shape(a)), dtype) for a in arrays_lor]
        return MPS(ctx, sites, form, dtype)

    @staticmethod
    def from_arrays(ctx: Context, arrays, order=("l", "r", "o"), dtype=np.complex128):
        """Tenet's ``MPS(arrays; order)`` (MPS.jl:50-94): edge arrays are 2-way."""
        order = list(order)
        n = len(arrays)
        out = []
        for i, arr in enumerate(arrays):
            loc = [x for x in order if not ((i == 0 and x == "l") or (i == n - 1 and x == "r"))]
            a = np.asarray(arr)
            if i == 0:
                a = a[None]
                loc = ["l"] + loc
            if i == n - 1:
                a = a[..., None]
                loc = loc + ["r"]
            out.append(np.transpose(a, [loc.index("l"), loc.index("o"), loc.index("r")]))
        return MPS.from_host(ctx, out, dtype)

    @staticmethod
    def product_state(ctx: Context, vectors, dtype=np.complex128):
        """``convert(MPS, ProductState(...))`` (MPS.jl:167-201)."""
        return MPS.from_host(ctx, [np.asarray(v).reshape(1, -1, 1) for v in vectors], dtype)

    def to_host(self):
        return [s.to_host() for s in self.sites]

    def copy(self):
        o = MPS(self.ctx, [s.copy() for s in self.sites], self.form, self.dtype)
        o.lam = None if self.lam is None else self.lam.copy()
        o.prec = self.prec
        return o

    @property
    def n(self):
        return len(self.sites)

    def bondsizes(self):
        return [s.shape[2] for s in self.sites[:-1]]

    def chis(self):
        return [1] + self.bondsizes() + [1]

    # ---- primitive moves --------------------------------------------------
    def _einsum(self, a, la, b, lb, c_shape, lc, conj_a=0, conj_b=0):
        c = DeviceArray(self.ctx, c_shape, self.dtype)
        da, db, dc = tdesc(a.shape, la), tdesc(b.shape, lb), tdesc(c_shape, lc)
        L.check(L.lib.tenet_einsum(self.ctx.h, None, self.code, self.prec, C.byref(da), C.c_void_p(a.ptr),
                                   C.byref(db), C.c_void_p(b.ptr), C.byref(dc), C.c_void_p(c.ptr), conj_a, conj_b),
                "einsum")
        return c

    def _diag_move(self, i, right):
        """The canonize! move across the bond of the last gate, when its two
        sites are still U sqrt(S) and sqrt(S) V^H: the thin QR of U sqrt(S) is
        U times sqrt(S) (Householder's Q and R differ only by diagonal phases,
        a gauge), so the move rescales the two sites (for a right move:
        Q = U, R = sqrt(S), next site sqrt(S) sqrt(S) V^H; mirrored for a
        left move).  ``i`` is the left site of the bond the move crosses
        (bond i-(i+1)) in both directions.  Returns False when the shortcut
        does not apply."""
        rec = self._gate_bond
        if rec is None:
            return False
        a, sa, sb, sq, isq = rec
        if i != a or self.sites[a] is not sa or self.sites[a + 1] is not sb:
            return False
        A, B = self.sites[a], self.sites[a + 1]
        if right:  # site a <- A / sqrt(S) on its right bond; site a+1 <- sqrt(S) B on its left bond
            self.sites[a] = self._einsum(A, [1, 2, 3], isq, [3], A.shape, [1, 2, 3])
            self.sites[a + 1] = self._einsum(sq, [1], B, [1, 2, 3], B.shape, [1, 2, 3])
        else:
            self.sites[a + 1] = self._einsum(isq, [1], B, [1, 2, 3], B.shape, [1, 2, 3])
            self.sites[a] = self._einsum(A, [1, 2, 3], sq, [3], A.shape, [1, 2, 3])
        self._gate_bond = None
        return True

    def _qr_right(self, i):
        """generic_canonize_site!(site i, bond i-(i+1); :qr), canonize.jl:23-76."""
        if self._diag_move(i, True):
            return
        A, B = self.sites[i], self.sites[i + 1]
        chil, d, chi = A.shape
        if chi == 1:  # canonize.jl:38
            return
        m = chil * d
        k = min(m, chi)
        Q = DeviceArray(self.ctx, (chil, d, k), self.dtype)
        _, d2, chir = B.shape
        self.flops["qr"] += qr_flops(m, chi)
        self.flops["contract"] += 8.0 * k * chi * d2 * chir
        if _qr_absorb_fits(m, chi, d2 * chir):  # QR and R's contraction into B in one launch
            Bn = DeviceArray(self.ctx, (k, d2, chir), self.dtype)
            L.check(L.lib.tenet_qr_absorb(self.ctx.h, None, self.code, 0, m, chi, C.c_void_p(A.ptr), m,
                                          C.c_void_p(Q.ptr), m, L.ABSORB_R_B, d2 * chir, C.c_void_p(B.ptr), chi,
                                          C.c_void_p(Bn.ptr), k), "qr_absorb")
            self.sites[i + 1] = Bn
        else:
            R = DeviceArray(self.ctx, (k, chi), self.dtype)
            L.check(L.lib.tenet_qr_thin(self.ctx.h, None, self.code, m, chi, C.c_void_p(A.ptr), m,
                                        C.c_void_p(Q.ptr), m, C.c_void_p(R.ptr), k), "qr")
            self.sites[i + 1] = self._einsum(R, [1, 2], B, [2, 3, 4], (k, d2, chir), [1, 3, 4])
        self.sites[i] = Q

    def _qr_left(self, i):
        """generic_canonize_site!(site i, bond (i-1)-i; :qr): QR of A^H, so
        A = R^H Q^H with Q^H a right isometry."""
        if self._diag_move(i - 1, False):
            return
        A, B = self.sites[i], self.sites[i - 1]
        chi, d, chir = A.shape
        if chi == 1:
            return
        m = d * chir  # rows of A^H
        k = min(m, chi)
        Qh = DeviceArray(self.ctx, (k, d, chir), self.dtype)
        chil2, d2, _ = B.shape
        self.flops["qr"] += qr_flops(m, chi)
        self.flops["contract"] += 8.0 * chil2 * d2 * chi * k
        if _qr_absorb_fits(m, chi, chil2 * d2):  # QR of A^H and B R^H in one launch
            Bn = DeviceArray(self.ctx, (chil2, d2, k), self.dtype)
            L.check(L.lib.tenet_qr_absorb(self.ctx.h, None, self.code, L.QR_IN_ADJ | L.QR_Q_ADJ, m, chi,
                                          C.c_void_p(A.ptr), chi, C.c_void_p(Qh.ptr), k, L.ABSORB_B_RH, chil2 * d2,
                                          C.c_void_p(B.ptr), chil2 * d2, C.c_void_p(Bn.ptr), chil2 * d2),
                    "qr_absorb (adjoint)")
            self.sites[i - 1] = Bn
        else:
            Rh = DeviceArray(self.ctx, (chi, k), self.dtype)
            L.check(L.lib.tenet_qr_thin_ex(self.ctx.h, None, self.code, L.QR_IN_ADJ | L.QR_Q_ADJ | L.QR_R_ADJ, m,
                                           chi, C.c_void_p(A.ptr), chi, C.c_void_p(Qh.ptr), k, C.c_void_p(Rh.ptr),
                                           chi), "qr (adjoint)")
            self.sites[i - 1] = self._einsum(B, [1, 2, 3], Rh, [3, 4], (chil2, d2, k), [1, 2, 4])
        self.sites[i] = Qh

    def _svd(self, M: DeviceArray, m, n, maxdim=0, threshold=0.0):
        kmax = min(m, n)
        U = DeviceArray(self.ctx, (m, kmax), self.dtype)
        Vh = DeviceArray(self.ctx, (kmax, n), self.dtype)
        sdev = self.ctx.malloc(kmax * 8 + 16)
        try:
            st = L.lib.tenet_svd_thin(self.ctx.h, None, self.code, m, n, C.c_void_p(M.ptr), m, C.c_void_p(U.ptr), m,
                                      C.c_void_p(sdev), C.c_void_p(Vh.ptr), kmax, int(maxdim), float(threshold),
                                      C.c_void_p(sdev + kmax * 8), None)
            L.check(st, "svd")
            s = np.zeros(kmax)
            kb = np.zeros(1, dtype=np.int64)
            L.check(L.lib.tenet_memcpy_d2h(self.ctx.h, s.ctypes.data_as(C.c_void_p), C.c_void_p(sdev), kmax * 8, None))
            L.check(L.lib.tenet_memcpy_d2h(self.ctx.h, kb.ctypes.data_as(C.c_void_p), C.c_void_p(sdev + kmax * 8), 8,
                                           None))
        finally:
            self.ctx.free(sdev)
        return U, s, Vh, int(kb[0])

    def _bond_right(self, i):
        """generic_bond_canonize_site! at site i across bond i-(i+1) (canonize.jl:78-116)."""
        A, B = self.sites[i], self.sites[i + 1]
        chil, d, chi = A.shape
        if chi == 1:
            self.lam = np.ones(1)
            return
        U, s, Vh, _ = self._svd(A, chil * d, chi)
        k = len(s)
        _, d2, chir = B.shape
        self.sites[i + 1] = self._einsum(Vh, [1, 2], B, [2, 3, 4], (k, d2, chir), [1, 3, 4])
        self.sites[i] = U.view((chil, d, k))
        self.lam = s

    def _bond_left(self, i):
        """Bond-canonize at site i across bond (i-1)-i: SVD of A (chi x d chi_r)
        = U s Vh; A <- Vh, A_{i-1} <- A_{i-1} U."""
        A, B = self.sites[i], self.sites[i - 1]
        chi, d, chir = A.shape
        if chi == 1:
            self.lam = np.ones(1)
            return
        U, s, Vh, _ = self._svd(A, chi, d * chir)
        k = len(s)
        chil2, d2, _ = B.shape
        self.sites[i - 1] = self._einsum(B, [1, 2, 3], U, [3, 4], (chil2, d2, k), [1, 2, 4])
        self.sites[i] = Vh.view((k, d, chir))
        self.lam = s

    def _scale_bond(self, site, vec, left):
        """hadamard(tensor, s): scale the left (or right) bond of ``site``."""
        t = DeviceArray.from_host(self.ctx, np.asarray(vec, dtype=self.dtype).reshape(-1), self.dtype)
        A = self.sites[site]
        if left:
            self.sites[site] = self._einsum(t, [1], A, [1, 2, 3], A.shape, [1, 2, 3])
        else:
            self.sites[site] = self._einsum(A, [1, 2, 3], t, [3], A.shape, [1, 2, 3])

    # ---- canonical forms ----------------------------------------------------
    def canonize_(self, lo, hi=None):
        """Mixed -> Mixed (canonize.jl:128-149); from NonCanonical first set
        MixedCanonical(all sites) (:123-126); from Bond absorb Lambda (:181-219)."""
        hi = lo if hi is None else hi
        if self.form[0] == "none":
            self.form = ("mixed", 0, self.n - 1)
        if self.form[0] == "bond":
            j = self.form[1]
            if j + 1 <= lo:
                self._scale_bond(j + 1, self.lam, left=True)
                self.form = ("mixed", j + 1, j + 1)
            elif hi <= j:
                self._scale_bond(j, self.lam, left=False)
                self.form = ("mixed", j, j)
            else:
                raise ValueError("cannot canonize from this BondCanonical form to the requested MixedCanonical")
            self.lam = None
        _, slo, shi = self.form
        if (slo, shi) == (lo, hi):
            return self
        for i in range(slo, lo):
            self._qr_right(i)
        for i in range(shi, hi, -1):
            self._qr_left(i)
        self.form = ("mixed", lo, hi)
        return self

    def canonize_bond_(self, i):
        """-> BondCanonical(bond i-(i+1)) (canonize.jl:151-179, 221-246)."""
        if self.form[0] == "bond" and self.form[1] == i:
            return self
        if self.form[0] == "bond":
            j = self.form[1]
            if j < i:
                self.canonize_(i)
            else:
                self.canonize_(i + 1)
        if self.form[0] == "none":
            self.form = ("mixed", 0, self.n - 1)
        _, lo, hi = self.form
        if hi <= i:
            self.canonize_(i)
            self._bond_right(i)
        elif i + 1 <= lo:
            self.canonize_(i + 1)
            self._bond_left(i + 1)
        elif lo <= i and i + 1 <= hi:
            self.canonize_(i)
            self._bond_right(i)
        else:
            raise ValueError("cannot canonize")
        self.form = ("bond", i)
        return self

    def schmidt_values(self, i):
        """Tenet.schmidt_values (src/Operations/entropy.jl:25-31): Lambda of
        BondCanonical(bond i-(i+1)) on a copy."""
        p = self.copy()
        p.canonize_bond_(i)
        return p.lam.copy()

    def compress_bond_(self, i, maxdim=None, threshold=None):
        """generic_mps_compress!(tn, bond) (compress.jl:29-69)."""
        if maxdim is None and threshold is None:
            raise ValueError("Either `maxdim` or `threshold` must be specified")
        if maxdim is not None and maxdim <= 0:
            raise ValueError("maxdim must be a positive integer")
        if threshold is not None and threshold <= 0:
            raise ValueError("Threshold must be positive")
        self.canonize_bond_(i)
        s = self.lam
        k = len(s) if maxdim is None else min(int(maxdim), len(s))
        if threshold is not None:
            k = int(np.count_nonzero(np.abs(s[:k]) > threshold))
        A, B = self.sites[i], self.sites[i + 1]
        chil, d, chi = A.shape
        if k < chi:
            self.sites[i] = A.view((chil, d, k)).copy()
            _, d2, chir = B.shape
            src = tdesc((k, d2, chir), [1, 2, 3], [1, chi, chi * d2])
            self.sites[i + 1] = self._gather(B, src, (k, d2, chir))
        self.lam = s[:k].copy()
        return self

    def _gather(self, src_arr, src_desc, shape):
        from .tensor import _one
        one = _one(self.ctx, self.dtype)
        out = DeviceArray(self.ctx, shape, self.dtype)
        db = tdesc((), [])
        dc = tdesc(shape, [1, 2, 3])
        L.check(L.lib.tenet_einsum(self.ctx.h, None, self.code, L.PREC_NATIVE, C.byref(src_desc),
                                   C.c_void_p(src_arr.ptr), C.byref(db), C.c_void_p(one.ptr), C.byref(dc),
                                   C.c_void_p(out.ptr), 0, 0), "gather")
        return out

    def compress_(self, maxdim=None, threshold=None):
        """All bonds sequentially (compress.jl:21-27)."""
        for i in range(self.n - 1):
            self.compress_bond_(i, maxdim, threshold)
        return self

    truncate_ = compress_

    # ---- gates ----------------------------------------------------------------
    def apply_one_site_(self, U2, a):
        """1-site shortcut of generic_simple_update! (simple_update.jl:28-37)."""
        g = DeviceArray.from_host(self.ctx, np.asarray(U2), self.dtype)
        A = self.sites[a]
        self.sites[a] = self._einsum(g, [4, 2], A, [1, 2, 3], A.shape, [1, 4, 3])
        return self

    def simple_update_(self, gate, a, maxdim=None, threshold=None, prec=None, absorb=L.ABSORB_EQUAL):
        """simple_update!(::MPS, gate on sites (a, a+1); maxdim) (simple_update.jl:82-92):
        move the orthogonality centre onto (a, a+1), then the fused gate update
        (Muscle.simple_update, AbsorbEqually).  ``gate`` is a 4x4 host matrix or
        a device (d, d, d, d) G[oa, ob, ia, ib] array.  Returns the kept bond dim."""
        self.canonize_(a, a + 1)
        A, B = self.sites[a], self.sites[a + 1]
        chil, d1, chi = A.shape
        _, d2, chir = B.shape
        if isinstance(gate, DeviceArray):
            G = gate
        else:
            G = DeviceArray.from_host(self.ctx, _gate_tensor(gate, d1), self.dtype)
        m, n = chil * d1, d2 * chir
        kmax = min(m, n)
        Aout = DeviceArray(self.ctx, (m * kmax,), self.dtype)
        Bout = DeviceArray(self.ctx, (kmax * n,), self.dtype)
        # the diagonal shortcut needs every kept S > 0: guaranteed by a positive threshold
        keep_s = absorb == L.ABSORB_EQUAL and (threshold or 0.0) > 0.0
        S = self.ctx.malloc(8 * kmax) if keep_s else None
        k = C.c_int64(0)
        st = L.lib.tenet_simple_update(self.ctx.h, None, self.code, self.prec if prec is None else prec, chil, d1, chi,
                                       d2, chir, C.c_void_p(A.ptr), C.c_void_p(B.ptr), C.c_void_p(G.ptr),
                                       C.c_void_p(Aout.ptr), C.c_void_p(Bout.ptr), None if S is None else C.c_void_p(S),
                                       int(maxdim or 0), float(threshold or 0.0), int(absorb), None, None, C.byref(k))
        if st != L.TENET_OK and S is not None:
            self.ctx.free(S)
        L.check(st, "simple_update")
        k = int(k.value)
        self.flops["contract"] += 8.0 * m * chi * n + 8.0 * chil * chir * (d1 * d2) ** 2
        self.flops["svd"] += svd_flops(m, n)
        self.sites[a] = Aout.view((chil, d1, k))
        self.sites[a + 1] = Bout.view((k, d2, chir))
        self.form = ("mixed", a, a + 1)
        self._gate_bond = None
        if S is not None:
            sq = DeviceArray(self.ctx, (k,), self.dtype)
            isq = DeviceArray(self.ctx, (k,), self.dtype)
            L.check(L.lib.tenet_sigma_scales(self.ctx.h, None, self.code, k, C.c_void_p(S), C.c_void_p(sq.ptr),
                                             C.c_void_p(isq.ptr)), "sigma_scales")
            # single-stream use: any later reuse of S's pool block is ordered after the kernel above
            self.ctx.free(S)
            self._gate_bond = (a, self.sites[a], self.sites[a + 1], sq, isq)
        return k

    def evolve_(self, gate, a):
        """evolve!(tn, op::Tensor) forwards no keyword arguments (evolve.jl:140-143)."""
        return self.simple_update_(gate, a)

    # ---- transfer-matrix sweeps ----------------------------------------------------
    def _site_ptrs(self):
        return (C.c_void_p * self.n)(*[s.ptr for s in self.sites])

    def overlap(self, other: "MPS", prec=L.PREC_NATIVE, log=False):
        """overlap(a, b) = <b|a> (overlap.jl:36-50).  With log=True returns
        (mantissa, log_scale)."""
        out = self.ctx.malloc(64)
        try:
            self._overlap_into(other, out, prec)
            h = np.zeros(3)
            L.check(L.lib.tenet_memcpy_d2h(self.ctx.h, h.ctypes.data_as(C.c_void_p), C.c_void_p(out), 24, None))
        finally:
            self.ctx.free(out)
        val = complex(h[0], h[1])
        if log:
            return val, float(h[2])
        return val * math.exp(h[2])

    def _overlap_into(self, other: "MPS", out: int, prec=L.PREC_NATIVE):
        """<other|self> into device doubles out[0:3] = (re, im, log-scale),
        with no download (graph-capturable)."""
        if self.n != other.n:
            raise ValueError("site count mismatch")
        d = self.sites[0].shape[1]
        ca = (C.c_int64 * (self.n + 1))(*self.chis())
        cb = (C.c_int64 * (other.n + 1))(*other.chis())
        L.check(L.lib.tenet_overlap(self.ctx.h, None, self.code, prec, self.n, d, ca, self._site_ptrs(), cb,
                                    other._site_ptrs(), C.c_void_p(out)), "overlap")

    def _norm2_dev(self):
        """|A_c|^2 of the (MixedCanonical, single-site) centre tensor as a 0-d
        device array, with no download (graph-capturable)."""
        lo = self.form[1]
        A = self.sites[lo]
        return self._einsum(A, [1, 2, 3], A, [1, 2, 3], (), [], conj_b=1)

    def norm(self):
        """norm(::MPS) (norm.jl:17-37): canonize to the minimum centre, then
        the Frobenius norm of that tensor (or of Lambda when BondCanonical)."""
        if self.form[0] == "bond":
            return float(np.linalg.norm(self.lam))
        if self.form[0] == "none":
            self.form = ("mixed", 0, self.n - 1)
        lo = self.form[1]
        self.canonize_(lo, lo)
        return math.sqrt(abs(self._norm2_dev().to_host().reshape(()).item()))

    def normalize_(self):
        """normalize!(::MPS) (normalize.jl:15-37)."""
        if self.form[0] == "bond":
            self.lam = self.lam / np.linalg.norm(self.lam)
            return self
        nrm = self.norm()
        lo = self.form[1]
        A = self.sites[lo]
        t = DeviceArray.from_host(self.ctx, np.full((), 1.0 / nrm, dtype=self.dtype), self.dtype)
        self.sites[lo] = self._einsum(A, [1, 2, 3], t, [], A.shape, [1, 2, 3])
        return self

    def amplitudes(self, bitstrings, site_parallel=None):
        """<x|psi> for each 0/1 bitstring (overlap.jl:22-32).  Returns
        complex128 values.  ``site_parallel`` (default: when a bond reaches
        64) takes tenet_amplitudes_sites (one launch per site, each slice
        product spread over many workgroups) instead of one workgroup per
        request."""
        bits = np.asarray(bitstrings, dtype=np.uint8).reshape(-1, self.n)
        nreq = bits.shape[0]
        chis = np.asarray(self.chis(), dtype=np.int64)
        ptrs = np.asarray([s.ptr for s in self.sites], dtype=np.uint64)
        ctx = self.ctx
        buf = ctx.malloc(ptrs.nbytes + chis.nbytes + bits.nbytes + nreq * (8 * 4 + 24) + 1024)
        try:
            p_ptrs = buf
            p_chis = p_ptrs + ptrs.nbytes
            p_bits = p_chis + chis.nbytes
            p_tab = (p_bits + bits.nbytes + 63) // 64 * 64
            p_out = p_tab + nreq * 32
            tab = np.zeros((nreq, 4), dtype=np.uint64)
            tab[:, 0] = self.n
            tab[:, 1] = p_ptrs
            tab[:, 2] = p_chis
            tab[:, 3] = p_bits + np.arange(nreq, dtype=np.uint64) * self.n
            # columns are gathered into the four device arrays the kernel reads
            n_arr = np.ascontiguousarray(tab[:, 0].astype(np.int64))
            sp = np.ascontiguousarray(tab[:, 1])
            ch = np.ascontiguousarray(tab[:, 2])
            bt = np.ascontiguousarray(tab[:, 3])
            blob = b"".join(x.tobytes() for x in (n_arr, sp, ch, bt))
            hostblob = np.frombuffer(blob, dtype=np.uint8)
            for src, dst in ((ptrs, p_ptrs), (chis, p_chis), (bits, p_bits)):
                L.check(L.lib.tenet_memcpy_h2d(ctx.h, C.c_void_p(dst), src.ctypes.data_as(C.c_void_p), src.nbytes,
                                               None))
            L.check(L.lib.tenet_memcpy_h2d(ctx.h, C.c_void_p(p_tab), hostblob.ctypes.data_as(C.c_void_p),
                                           hostblob.nbytes, None))
            chi_max = int(chis.max())
            if site_parallel is None:
                site_parallel = chi_max >= 64
            if site_parallel:
                L.check(L.lib.tenet_amplitudes_sites(ctx.h, None, self.code, nreq, self.n, chi_max, C.c_void_p(p_tab),
                                                     C.c_void_p(p_tab + 8 * nreq), C.c_void_p(p_tab + 16 * nreq),
                                                     C.c_void_p(p_tab + 24 * nreq), self.sites[0].shape[1],
                                                     C.c_void_p(p_out)), "amplitudes_sites")
            else:
                L.check(L.lib.tenet_amplitudes(ctx.h, None, self.code, nreq, C.c_void_p(p_tab),
                                               C.c_void_p(p_tab + 8 * nreq), C.c_void_p(p_tab + 16 * nreq),
                                               C.c_void_p(p_tab + 24 * nreq), self.sites[0].shape[1],
                                               C.c_void_p(p_out)), "amplitudes")
            h = np.zeros((nreq, 3))
            L.check(L.lib.tenet_memcpy_d2h(ctx.h, h.ctypes.data_as(C.c_void_p), C.c_void_p(p_out), h.nbytes, None))
        finally:
            ctx.free(buf)
        return (h[:, 0] + 1j * h[:, 1]) * np.exp(h[:, 2])

    def sample(self, nsamples, seed=0):
        """Tenet's ``sample(psi, nsamples)`` (src/Operations/sample.jl:20-72):
        on a copy canonised to the first site, every sample walks the chain
        with a projected vector V (1 at the first site): T = V . A_i for all
        samples in one contraction, then tenet_sample_select draws each
        sample's value from p_x = sum_r |T[x, r]|^2 with the uniform of
        generator element (sample * n + site) and renormalises its
        projection.  Returns an (nsamples, n) int array of 0-based values
        (Tenet's are 1-based)."""
        p = self.copy()
        p.canonize_(0)
        ctx, n = p.ctx, p.n
        ns = int(nsamples)
        out = np.zeros((ns, n), dtype=np.int32)
        if ns == 0:
            return out
        bits = ctx.malloc(ns * n * 4)
        try:
            V = DeviceArray.from_host(ctx, np.ones((1, ns)), self.dtype)
            for i in range(n):
                A = p.sites[i]
                l, d, r = A.shape
                T = p._einsum(V, [1, 4], A, [1, 2, 3], (d, r, ns), [2, 3, 4])
                Vn = DeviceArray(ctx, (r, ns), self.dtype) if i < n - 1 else None
                L.check(L.lib.tenet_sample_select(ctx.h, None, self.code, C.c_void_p(T.ptr), d, r, ns, n, i,
                                                  C.c_uint64(int(seed)), C.c_void_p(bits),
                                                  C.c_void_p(Vn.ptr) if Vn is not None else None), "sample_select")
                V = Vn
            L.check(L.lib.tenet_memcpy_d2h(ctx.h, out.ctypes.data_as(C.c_void_p), C.c_void_p(bits), out.nbytes, None))
        finally:
            ctx.free(bits)
        return out

    def expect_pauli(self, start, paulis, prec=L.PREC_NATIVE):
        """<psi|P|psi> for a Pauli string on sites start..start+w-1, using the
        canonical form: sites left of the centre are left isometries and
        sites right of it right isometries, so the sweep only spans the
        centre and the string."""
        if self.form[0] != "mixed":
            raise ValueError("expect_pauli needs a MixedCanonical MPS")
        lo, hi = self.form[1], self.form[2]
        w = len(paulis)
        s0 = min(lo, start)
        e0 = max(hi, start + w - 1)
        ps = "".join(paulis[i - start] if start <= i < start + w else "I" for i in range(s0, e0 + 1))
        chis = (C.c_int64 * (self.n + 1))(*self.chis())
        out = self.ctx.malloc(64)
        try:
            L.check(L.lib.tenet_expect_pauli(self.ctx.h, None, self.code, prec, s0, len(ps), 2, chis,
                                             self._site_ptrs(), None, ps.encode(), C.c_void_p(out)), "expect_pauli")
            h = np.zeros(3)
            L.check(L.lib.tenet_memcpy_d2h(self.ctx.h, h.ctypes.data_as(C.c_void_p), C.c_void_p(out), 24, None))
        finally:
            self.ctx.free(out)
        return complex(h[0], h[1]) * math.exp(h[2])
