"""Two-site DMRG effective-Hamiltonian apply on the device.

EffectiveHamiltonian(le, re, op)(psi) = binary_einsum(binary_einsum(
binary_einsum(le, psi), op), re)  (src/Algorithms/DMRG.jl:10-18), with
op = binary_einsum(W_{i-1}, W_i) pre-contracted for Dmrg2 (DMRG.jl:297-299)
and psi = theta = A_{i-1} A_i (DMRG.jl:300).  The three contractions run on
the K1 kernels in the requested precision; the result is realigned to
theta's label order as DMRG.jl:320 does with permutedims.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from .device import Context
from .tensor import Tensor, binary_einsum, permute


def xxz_mpo_tensor(J=1.0, delta=0.5):
    """Bulk tensor W[l, r, o, i] of the Heisenberg XXZ chain in the standard
    5 x 5 lower-triangular form with {I, S+, S-, Sz} (S = sigma / 2):
    H = J sum (S+_i S-_{i+1} + S-_i S+_{i+1}) / 2 + J Delta sum Sz_i Sz_{i+1}."""
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
    return W


def heff_apply(L_env: Tensor, W1: Tensor, W2: Tensor, R_env: Tensor, theta: Tensor, prec=L.PREC_NATIVE,
               labels=("l", "o1", "o2", "r")):
    """y = ((L . theta) . (W1 W2)) . R.

    Expected labels:  L_env (l, w, lk)   theta (lk, i1, i2, rk)
                      W1 (w, m, o1, i1)  W2 (m, wr, o2, i2)   R_env (r, wr, rk)
    The two chi^3-scale contractions use ``prec``; the MPO contraction and
    the tiny W1.W2 product run in the tensors' native precision.  Returns y
    with labels ``labels`` = (l, o1, o2, r)."""
    op = binary_einsum(W1, W2)  # (w, o1, i1, wr, o2, i2)
    t1 = binary_einsum(L_env, theta, prec=prec)  # (l, w, i1, i2, rk)
    # the labels R contracts last, in R's order (l, o1, o2, wr, rk): the last
    # contraction's operands are then plain matrices, a library GEMM
    nat = [x for x in t1.labels if x not in op.labels] + [x for x in op.labels if x not in t1.labels]
    kl = [x for x in R_env.labels if x in nat]
    t2 = binary_einsum(t1, op, out_labels=[x for x in nat if x not in kl] + kl)
    y = binary_einsum(t2, R_env, prec=prec)  # (l, o1, o2, r)
    return permute(y, list(labels))


# ---------------------------------------------------------------------------
# Two-site DMRG sweeps (Dmrg2, src/Algorithms/DMRG.jl:214-459) on the device
# ---------------------------------------------------------------------------
import ctypes as C  # noqa: E402

from .device import DeviceArray, tdesc  # noqa: E402
from .mps import MPS, svd_flops  # noqa: E402


class _Plan:
    """A cached einsum plan executed with alpha / beta (C = alpha A.B + beta C)."""

    def __init__(self, ctx, dtype, da, db, dc):
        self.ctx = ctx
        h = C.c_void_p()
        L.check(L.lib.tenet_einsum_plan(ctx.h, dtype, L.PREC_NATIVE, C.byref(da), C.byref(db), C.byref(dc),
                                        C.byref(h)), "einsum_plan")
        self.h = h

    def run(self, A, B, Cm, conj_a=0, conj_b=0, alpha=1.0, beta=0.0):
        al = (C.c_double * 2)(float(np.real(alpha)), float(np.imag(alpha)))
        be = (C.c_double * 2)(float(np.real(beta)), float(np.imag(beta)))
        L.check(L.lib.tenet_einsum_exec(self.h, None, C.c_void_p(A), C.c_void_p(B), C.c_void_p(Cm), conj_a, conj_b,
                                        al, be), "einsum_exec")

    def close(self):
        if self.h is not None:
            L.lib.tenet_plan_destroy(self.h)
            self.h = None


class _Krylov:
    """Device Krylov basis (N x (kdim + 1), one column per vector) and the
    few vector operations Lanczos needs, each one contraction:
    c = V[:, :m]^H w, w += alpha V[:, :m] c, x = V[:, :m] y."""

    def __init__(self, ctx, dtype, N, kdim):
        self.ctx, self.dtype, self.N, self.kdim = ctx, np.dtype(dtype), N, kdim
        self.code = L.C128 if self.dtype == np.complex128 else L.C64
        self.V = DeviceArray(ctx, (N, kdim + 1), dtype)
        self.c = DeviceArray(ctx, (kdim + 1,), dtype)
        self.one = DeviceArray.from_host(ctx, np.ones((), dtype=dtype), dtype)
        self._plans = {}

    def col(self, k):
        v = DeviceArray(self.ctx, (self.N,), self.dtype, ptr=self.V.ptr + k * self.N * self.dtype.itemsize,
                        owner=False)
        v._base = self.V
        return v

    def _plan(self, key, da, db, dc):
        p = self._plans.get(key)
        if p is None:
            p = _Plan(self.ctx, self.code, da, db, dc)
            self._plans[key] = p
        return p

    def overlaps(self, m, w):
        """V[:, :m]^H w -> host (m,)."""
        p = self._plan(("ov", m), tdesc((self.N, m), [1, 2]), tdesc((self.N,), [1]), tdesc((m,), [2]))
        p.run(self.V.ptr, w.ptr, self.c.ptr, conj_a=1)
        out = np.zeros(m, dtype=self.dtype)
        L.check(L.lib.tenet_memcpy_d2h(self.ctx.h, out.ctypes.data_as(C.c_void_p), C.c_void_p(self.c.ptr),
                                       out.nbytes, None))
        return out

    def combine(self, m, coef, out, alpha=1.0, beta=0.0):
        """out = alpha V[:, :m] coef + beta out (coef on the host)."""
        L.check(L.lib.tenet_memcpy_h2d(self.ctx.h, C.c_void_p(self.c.ptr),
                                       np.ascontiguousarray(coef, dtype=self.dtype).ctypes.data_as(C.c_void_p),
                                       m * self.dtype.itemsize, None))
        p = self._plan(("cb", m), tdesc((self.N, m), [1, 2]), tdesc((m,), [2]), tdesc((self.N,), [1]))
        p.run(self.V.ptr, self.c.ptr, out.ptr, alpha=alpha, beta=beta)

    def scale_into(self, src, k, s):
        """V[:, k] = s * src."""
        p = self._plan(("sc",), tdesc((self.N,), [1]), tdesc((), []), tdesc((self.N,), [1]))
        p.run(src.ptr, self.one.ptr, self.col(k).ptr, alpha=s)

    def close(self):
        for p in self._plans.values():
            p.close()
        self._plans = {}


def lanczos_ground(apply, x0: DeviceArray, kdim=8, tol=1e-10, maxiter=60):
    """Lowest eigenpair of the Hermitian map ``apply`` (device vector ->
    device vector of the same size) by restarted Lanczos with full
    reorthogonalisation (two passes), restarting from the Ritz vector.
    Returns (eigenvalue, DeviceArray of x0's shape)."""
    ctx, dt = x0.ctx, x0.dtype
    N = x0.size
    K = _Krylov(ctx, dt, N, kdim)
    try:
        x = DeviceArray(ctx, x0.shape, dt)
        nrm = np.sqrt(x0.nrm2sq())
        K.scale_into(x0.view((N,)), 0, 1.0 / nrm)
        theta = 0.0
        for _ in range(maxiter):
            Hm = np.zeros((kdim + 1, kdim + 1), dtype=np.complex128)
            m = 0
            beta = 0.0
            for k in range(kdim):
                w = apply(K.col(k).view(x0.shape)).view((N,))
                c1 = K.overlaps(k + 1, w)
                K.combine(k + 1, c1, w, alpha=-1.0, beta=1.0)
                c2 = K.overlaps(k + 1, w)
                K.combine(k + 1, c2, w, alpha=-1.0, beta=1.0)
                Hm[:k + 1, k] = c1 + c2
                beta = float(np.sqrt(w.nrm2sq()))
                m = k + 1
                if beta <= 1e-14 * max(1.0, abs(Hm[k, k])) or k == kdim - 1:
                    break
                K.scale_into(w, k + 1, 1.0 / beta)
                Hm[k + 1, k] = beta
            Hh = 0.5 * (Hm[:m, :m] + Hm[:m, :m].conj().T)
            ev, evec = np.linalg.eigh(Hh)
            theta, y = float(ev[0]), evec[:, 0]
            resid = abs(beta * y[m - 1])
            K.combine(m, y, x.view((N,)))
            xn = np.sqrt(x.nrm2sq())
            K.scale_into(x.view((N,)), 0, 1.0 / xn)
            if resid <= tol * max(1.0, abs(theta)) or m < kdim:
                break
        out = DeviceArray(ctx, x0.shape, dt)
        L.check(L.lib.tenet_memcpy_d2d(ctx.h, C.c_void_p(out.ptr), C.c_void_p(K.col(0).ptr),
                                       N * np.dtype(dt).itemsize, None), "d2d")
        return theta, out
    finally:
        K.close()


def _env_left(psi: MPS, Lenv, A, W):
    """L' (b', w', k') = sum L(b, w, k) A(k, p, k') W(w, o, p, w') conj(A)(b, o, b')
    (DMRG.jl:72-82, 106-116)."""
    b, w, k = Lenv.shape
    _, p, k2 = A.shape
    _, o, _, w2 = W.shape
    t1 = psi._einsum(Lenv, [1, 2, 3], A, [3, 4, 5], (b, w, p, k2), [1, 2, 4, 5])
    t2 = psi._einsum(t1, [1, 2, 4, 5], W, [2, 6, 4, 7], (b, k2, o, w2), [1, 5, 6, 7])
    return psi._einsum(t2, [1, 5, 6, 7], A, [1, 6, 8], (A.shape[2], w2, k2), [8, 7, 5], 0, 1)


def _env_right(psi: MPS, Renv, A, W):
    """R' (b, w, k) = sum A(k, p, k2) W(w, o, p, w2) conj(A)(b, o, b2) R(b2, w2, k2)."""
    b2, w2, k2 = Renv.shape
    k, p, _ = A.shape
    w, o, _, _ = W.shape
    t1 = psi._einsum(A, [1, 2, 3], Renv, [4, 5, 3], (k, p, b2, w2), [1, 2, 4, 5])
    t2 = psi._einsum(t1, [1, 2, 4, 5], W, [6, 7, 2, 5], (k, b2, w, o), [1, 4, 6, 7])
    return psi._einsum(t2, [1, 4, 6, 7], A, [8, 7, 4], (A.shape[0], w, k), [8, 6, 1], 0, 1)


def dmrg2_(psi: MPS, H, nsweeps=4, maxdim=128, kdim=8, tol=1e-10, maxiter=60, cut=1e-14):
    """Two-site DMRG (Dmrg2, src/Algorithms/DMRG.jl:214-459) on ``psi`` with
    the MPO ``H`` (tenet_hip.mpo.MPO, sites (l, o, i, r)).  Right
    environments first (centre at site 0), then ``nsweeps`` left-to-right and
    right-to-left passes: per window, the lowest eigenpair of H_eff
    (heff_apply) by restarted Lanczos on the device, the SVD split keeping
    min(maxdim, #S), normalize!(S) and S absorbed on the side the sweep moves
    to (DMRG.jl:336-356, 435-455).  Singular values at or below
    cut * S[0] are dropped as well (the reference keeps numerically zero
    ones; their U columns are not orthonormal in the one-sided Jacobi, and
    they carry no weight).  Returns the last eigenvalue."""
    from .mpo import MPO  # noqa: F401

    n = psi.n
    ctx, dt = psi.ctx, psi.dtype
    psi.canonize_(0)
    ones = DeviceArray.from_host(ctx, np.ones((1, 1, 1)), dt)
    R = [None] * (n + 1)
    Lf = [None] * (n + 1)
    R[n] = ones
    Lf[0] = ones
    for j in range(n - 1, 1, -1):
        R[j] = _env_right(psi, R[j + 1], psi.sites[j], H.sites[j])
    energy = 0.0

    def solve(j):
        A1, A2 = psi.sites[j], psi.sites[j + 1]
        l, d1, _ = A1.shape
        _, d2, r = A2.shape
        theta = psi._einsum(A1, [1, 2, 3], A2, [3, 4, 5], (l, d1, d2, r), [1, 2, 4, 5])
        Le = Tensor(Lf[j], ["l", "w", "lk"])
        Re = Tensor(R[j + 2], ["r", "wr", "rk"])
        W1 = Tensor(H.sites[j], ["w", "o1", "i1", "m"])
        W2 = Tensor(H.sites[j + 1], ["m", "o2", "i2", "wr"])

        def apply(x):
            return heff_apply(Le, W1, W2, Re, Tensor(x, ["lk", "i1", "i2", "rk"])).data

        e, y = lanczos_ground(apply, theta, kdim, tol, maxiter)
        U, s, Vh, kk = psi._svd(y, l * d1, d2 * r, int(maxdim), 0.0)
        psi.flops["svd"] += svd_flops(l * d1, d2 * r)
        k = int(np.count_nonzero(s[:kk] > cut * s[0])) if kk else 0
        k = max(k, 1)
        sk = s[:k] / np.linalg.norm(s[:k])
        return e, U, sk, Vh, k, (l, d1, d2, r)

    def take_cols(U, m, k):
        v = DeviceArray(ctx, (m, k), dt, ptr=U.ptr, owner=False)
        v._base = U
        return v

    from .vidal import _leading_rows

    class _H:
        pass

    hv = _H()
    hv.ctx, hv.dtype, hv.code = ctx, dt, psi.code
    hv._one = DeviceArray.from_host(ctx, np.ones((), dtype=dt), dt)
    for _ in range(nsweeps):
        for j in range(0, n - 1):  # left to right: centre moves to j + 1
            energy, U, sk, Vh, k, (l, d1, d2, r) = solve(j)
            Vk = Vh if k == Vh.shape[0] else _leading_rows(hv, Vh, k)
            sv = DeviceArray.from_host(ctx, sk.astype(dt), dt)
            psi.sites[j] = take_cols(U, l * d1, k).view((l, d1, k))
            psi.sites[j + 1] = psi._einsum(sv, [1], Vk, [1, 2], (k, d2 * r), [1, 2]).view((k, d2, r))
            Lf[j + 1] = _env_left(psi, Lf[j], psi.sites[j], H.sites[j])
        for j in range(n - 2, -1, -1):  # right to left: centre moves to j
            energy, U, sk, Vh, k, (l, d1, d2, r) = solve(j)
            Vk = Vh if k == Vh.shape[0] else _leading_rows(hv, Vh, k)
            sv = DeviceArray.from_host(ctx, sk.astype(dt), dt)
            psi.sites[j] = psi._einsum(take_cols(U, l * d1, k), [1, 2], sv, [2], (l * d1, k), [1, 2]).view((l, d1, k))
            psi.sites[j + 1] = Vk.view((k, d2, r))
            R[j + 1] = _env_right(psi, R[j + 2], psi.sites[j + 1], H.sites[j + 1])
    psi.form = ("mixed", 0, 0)
    psi.lam = None
    psi._gate_bond = None
    return energy


def heff1_apply(L_env: Tensor, W: Tensor, R_env: Tensor, x: Tensor, prec=L.PREC_NATIVE, labels=("l", "o", "r")):
    """One-site EffectiveHamiltonian(le, re, op)(x) (DMRG.jl:10-18 with op =
    the site's MPO tensor, as Dmrg1 builds it at DMRG.jl:117, 171):
    y = ((L . x) . W) . R.  Labels: L_env (l, w, lk), x (lk, i, rk),
    W (w, o, i, wr), R_env (r, wr, rk); y comes back as ``labels``."""
    t1 = binary_einsum(L_env, x, prec=prec)  # (l, w, i, rk)
    t2 = binary_einsum(t1, W)  # (l, rk, o, wr)
    y = binary_einsum(t2, R_env, prec=prec)  # (l, o, r)
    return permute(y, list(labels))


def dmrg1_(psi: MPS, H, nsweeps=4, kdim=8, tol=1e-10, maxiter=60):
    """One-site DMRG (Dmrg1, src/Algorithms/DMRG.jl:44-212) on ``psi`` with
    the MPO ``H``: right environments from a right-canonical state (centre
    at site 0), then per sweep a left-to-right pass over sites 1..n-1 and a
    right-to-left pass over sites n-2..0 (0-based; each pass skips its first
    site, DMRG.jl:90, 145).  At each site: canonize! onto it (one QR move),
    update the environment left behind, and replace the site by the lowest
    eigenvector of the one-site H_eff (restarted Lanczos on the device).
    Bond dimensions never change.  Returns the last eigenvalue."""
    n = psi.n
    ctx, dt = psi.ctx, psi.dtype
    psi.canonize_(0)
    ones = DeviceArray.from_host(ctx, np.ones((1, 1, 1)), dt)
    R = [None] * (n + 1)
    Lf = [None] * (n + 1)
    R[n] = ones
    Lf[0] = ones
    for j in range(n - 1, 0, -1):
        R[j] = _env_right(psi, R[j + 1], psi.sites[j], H.sites[j])
    energy = 0.0

    def solve(i):
        x = psi.sites[i]
        Le = Tensor(Lf[i], ["l", "w", "lk"])
        Re = Tensor(R[i + 1], ["r", "wr", "rk"])
        W = Tensor(H.sites[i], ["w", "o", "i", "wr"])

        def apply(v):
            return heff1_apply(Le, W, Re, Tensor(v, ["lk", "i", "rk"])).data

        return lanczos_ground(apply, x, kdim, tol, maxiter)

    for _ in range(nsweeps):
        for i in range(1, n):  # left to right
            psi.canonize_(i)
            Lf[i] = _env_left(psi, Lf[i - 1], psi.sites[i - 1], H.sites[i - 1])
            energy, y = solve(i)
            psi.sites[i] = y
        for i in range(n - 2, -1, -1):  # right to left
            psi.canonize_(i)
            R[i + 1] = _env_right(psi, R[i + 2], psi.sites[i + 1], H.sites[i + 1])
            energy, y = solve(i)
            psi.sites[i] = y
    psi._gate_bond = None
    return energy
