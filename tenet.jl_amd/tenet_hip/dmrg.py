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
    t2 = binary_einsum(t1, op)  # (l, rk, o1, wr, o2)
    y = binary_einsum(t2, R_env, prec=prec)  # (l, o1, o2, r)
    return permute(y, list(labels))
