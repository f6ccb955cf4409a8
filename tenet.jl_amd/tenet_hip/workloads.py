"""Seeded synthetic workloads: shapes drawn on the host (numpy PCG64), entries
generated on the device (``DeviceArray.gaussian_``), so that multi-GB
batches never cross PCIe and any slice can be regenerated on the host.

plain_state_arrays     the n = 20, chi = 32 Gaussian MPS of the plain
                       canonize!/norm workload
grouped_contractions   8192 pairwise contractions of rank-3..5 tensors,
                       extents log-uniform in 2..128, 1-3 contracted labels
grouped_svds           8192 truncated SVDs, sides log-uniform in 16..512
serving_states         random MPS shapes for the request loop
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

CAP_ELEMS = 1 << 20


@dataclass
class ContractionSpec:
    la: list
    lb: list
    lc: list
    extent: dict

    def shape(self, labels):
        return tuple(self.extent[l] for l in labels)

    @property
    def sizes(self):
        return (int(np.prod(self.shape(self.la))), int(np.prod(self.shape(self.lb))),
                int(np.prod(self.shape(self.lc))))

    @property
    def mnk(self):
        shared = [l for l in self.la if l in self.lb]
        K = int(np.prod([self.extent[l] for l in shared])) if shared else 1
        M = int(np.prod([self.extent[l] for l in self.la if l not in shared]))
        N = int(np.prod([self.extent[l] for l in self.lb if l not in shared]))
        return M, N, K


def plain_state_arrays(seed: int, n: int = 20, chi: int = 32, d: int = 2):
    """n sites, uniform bond dimension chi (edges included: site 1 is
    (1, d, chi) as a 3-way array), standard complex Gaussian entries
    (N(0,1) + i N(0,1)) / sqrt(2) from numpy's PCG64 stream ``seed``."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        shp = (1 if i == 0 else chi, d, 1 if i == n - 1 else chi)
        out.append((rng.standard_normal(shp) + 1j * rng.standard_normal(shp)) / np.sqrt(2.0))
    return out


def _log_uniform_int(rng, lo, hi):
    return int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))


def grouped_contractions(seed: int, count: int, cap: int = CAP_ELEMS):
    """rA, rB ~ U{3,4,5}; s ~ U{1,2,3} capped at min(rA, rB) shared labels,
    all contracted; extents round(exp(U(ln 2, ln 128))); a draw is rejected
    and redrawn until both operands and the result hold <= ``cap`` elements.
    Output labels: A's free labels then B's free labels."""
    rng = np.random.default_rng(seed)
    specs = []
    while len(specs) < count:
        ra, rb = int(rng.integers(3, 6)), int(rng.integers(3, 6))
        s = min(int(rng.integers(1, 4)), ra, rb)
        shared = [f"s{j}" for j in range(s)]
        fa = [f"a{j}" for j in range(ra - s)]
        fb = [f"b{j}" for j in range(rb - s)]
        la = fa + shared
        lb = shared + fb
        perm_a = rng.permutation(ra)
        perm_b = rng.permutation(rb)
        la = [la[i] for i in perm_a]
        lb = [lb[i] for i in perm_b]
        ext = {l: _log_uniform_int(rng, 2, 128) for l in fa + shared + fb}
        spec = ContractionSpec(la, lb, [l for l in la if l not in shared] + [l for l in lb if l not in shared], ext)
        if max(spec.sizes) <= cap:
            specs.append(spec)
    return specs


def grouped_svds(seed: int, count: int):
    """m, n independently log-uniform in [16, 512]; maxdim = ceil(min(m, n) / 2)."""
    rng = np.random.default_rng(seed)
    ms = np.array([_log_uniform_int(rng, 16, 512) for _ in range(count)], dtype=np.int64)
    ns = np.array([_log_uniform_int(rng, 16, 512) for _ in range(count)], dtype=np.int64)
    maxdims = (np.minimum(ms, ns) + 1) // 2
    return ms, ns, maxdims


def contraction_work(specs):
    """(flops, bytes) with complex MAC = 8 flops and every operand and the
    result moved once (c64: 8 bytes per element)."""
    fl = 0.0
    by = 0.0
    for sp in specs:
        M, N, K = sp.mnk
        fl += 8.0 * M * N * K
        a, b, c = sp.sizes
        by += 8.0 * (a + b + c)
    return fl, by


def svd_work(ms, ns):
    fl = 0.0
    for m, n in zip(ms, ns):
        p, q = max(m, n), min(m, n)
        fl += 4.0 * (4.0 * p * p * q + 8.0 * p * q * q + 9.0 * q ** 3)
    return fl


def serving_state_shapes(rng, n_lo=16, n_hi=128, chi_lo=16, chi_hi=512, d=2):
    """n ~ U{n_lo..n_hi}; chi log-uniform in [chi_lo, chi_hi], each bond b
    capped at d^min(b, n-b) as rand(MPS) caps it (MPS.jl:120-152).
    Returns (n, chi, bond list)."""
    n = int(rng.integers(n_lo, n_hi + 1))
    chi = _log_uniform_int(rng, chi_lo, chi_hi)
    bonds = []
    for b in range(1, n):
        e = min(b, n - b)
        bonds.append(chi if e >= 20 else min(chi, d ** e))
    return n, chi, bonds


def build_serving_states(ctx, seed: int, count: int, n_lo=16, n_hi=128, chi_lo=16, chi_hi=512,
                         dtype=np.complex64):
    """``count`` device-resident MPS states: shapes from serving_state_shapes,
    complex Gaussian entries generated on the device (state j uses seed
    seed + j, its sites consecutive offsets of that stream), each site
    scaled by 1 / sqrt(d chi_right), then canonize!(psi, site 1) and
    normalize!.  The scale leaves the normalised state unchanged; without it
    E<psi|psi> = d^n prod(chi_b) (about 1e66 at n = 32, chi = 213, and
    1e330 at n = 128, chi = 512), so the centre site that the canonize
    sweep accumulates the norm into overflows complex64; with it
    E<psi|psi> = 1.  Returns ServedState objects."""
    from .device import DeviceArray
    from .mps import MPS
    from .serving import ServedState

    rng = np.random.default_rng(seed)
    out = []
    for j in range(count):
        n, chi, bonds = serving_state_shapes(rng, n_lo, n_hi, chi_lo, chi_hi)
        dims = [1] + bonds + [1]
        sites = []
        off = 0
        for i in range(n):
            a = DeviceArray(ctx, (dims[i], 2, dims[i + 1]), dtype).gaussian_(seed + j, off)
            a.scale_(1.0 / math.sqrt(2.0 * dims[i + 1]))
            off += a.size
            sites.append(a)
        psi = MPS(ctx, sites, dtype=dtype)
        psi.canonize_(0)
        psi.normalize_()
        out.append(ServedState(ctx, psi))
    return out


def serving_requests(seed: int, states, count: int, amp_fraction=0.5, rate=None):
    """``count`` requests over the states: with probability amp_fraction a
    bitstring amplitude (uniform x), else a Pauli string of weight 1-4 on
    consecutive sites at a uniform start with letters uniform in X, Y, Z.
    With ``rate`` (requests/s) also returns Poisson arrival times."""
    from .serving import Request

    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        sid = int(rng.integers(0, len(states)))
        n = states[sid].n
        if rng.random() < amp_fraction:
            reqs.append(Request(i, sid, "amp", bits=rng.integers(0, 2, n).astype(np.uint8)))
        else:
            w = int(rng.integers(1, 5))
            start = int(rng.integers(0, n - w + 1))
            reqs.append(Request(i, sid, "pauli", start=start, paulis="".join(rng.choice(list("XYZ"), w))))
    if rate is None:
        return reqs
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
    return reqs, arrivals
