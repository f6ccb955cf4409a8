"""TEBD brickwork driver: the user loop of simple_update! over a depth-D
brickwork (Tenet has no TEBD function; examples/mps.ipynb applies gates this
way).  Odd layers act on bonds (1,2),(3,4),... left to right, even layers on
(2,3),(4,5),... right to left (serpentine), so each gate needs about two QR
moves of the orthogonality centre."""
from __future__ import annotations

import numpy as np

from .device import Context, DeviceArray
from .mps import MPS


def brickwork_schedule(n, depth):
    """0-based left site of every gate, in application order."""
    out = []
    for layer in range(depth):
        if layer % 2 == 0:
            out.extend(range(0, n - 1, 2))
        else:
            out.extend(reversed(range(1, n - 1, 2)))
    return out


def haar_gates(seed, count, dim=4):
    """Haar-random unitaries by QR of complex Ginibre matrices with the phase
    fix of Mezzadri (Notices AMS 54, 2007); entries (N(0,1) + i N(0,1))/sqrt(2)
    from numpy's PCG64 stream ``seed`` (the workload's generator)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        z = ((rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))) / np.sqrt(2)).astype(np.complex128)
        q, r = np.linalg.qr(z)
        ph = np.diag(r) / np.abs(np.diag(r))
        out.append(q * ph[None, :])
    return np.stack(out)


def upload_gates(ctx: Context, gates, d=2, dtype=np.complex128):
    """Stack 4x4 gates (site a most significant) as column-major
    G[oa, ob, ia, ib] tensors in one device buffer; returns per-gate views."""
    g = np.asarray(gates).reshape(-1, d, d, d, d)  # (gate, oa, ob, ia, ib) row-major
    stacked = np.transpose(g, (1, 2, 3, 4, 0))  # F-order: G[oa,ob,ia,ib,gate]
    dev = DeviceArray.from_host(ctx, stacked, dtype)
    per = d ** 4
    views = []
    for i in range(g.shape[0]):
        v = DeviceArray(ctx, (d, d, d, d), dtype, ptr=dev.ptr + i * per * np.dtype(dtype).itemsize, owner=False)
        v._base = dev
        views.append(v)
    return views


def tebd_sweep(ctx: Context, n, sched, gate_views, maxdim, threshold, prec=0, dtype=np.complex128, psi=None,
               on_gate=None):
    """Apply every gate of the schedule with simple_update!; returns the MPS."""
    if psi is None:
        psi = MPS.product_state(ctx, [[1.0, 0.0]] * n, dtype)
    for g, (a, G) in enumerate(zip(sched, gate_views)):
        psi.simple_update_(G, a, maxdim, threshold, prec=prec)
        if on_gate is not None:
            on_gate(g, psi)
    return psi
