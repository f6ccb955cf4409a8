"""Graph-captured sweeps: a sequence of library calls with fixed shapes is
recorded once into a HIP graph (tenet_graph_begin / end) and replayed with
one launch, so the host issues no per-kernel work and no per-move Python.

CanonizeNormGraph   canonize!(psi, site lo) (the QR sweep of
                    src/Operations/canonize.jl:136-145) on a copy of psi,
                    then <psi|psi> (overlap.jl:36-50) and norm(psi)
                    (norm.jl:17-37), all on the device; a replay re-reads
                    the source state and rewrites the same output buffers.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as L
from .device import DeviceArray
from .mps import MPS


class CanonizeNormGraph:
    """``run()`` returns (overlap <psi|psi>, norm(psi)) of the canonised copy;
    ``state`` is that copy (its sites are rewritten by every replay).  The
    source MPS must keep its sites (it is read by every replay)."""

    def __init__(self, psi: MPS, lo: int = 0):
        ctx = psi.ctx
        self.ctx = ctx
        self.src = psi
        self._out = ctx.malloc(64)

        def body():
            p = psi.copy()
            p.canonize_(lo)
            p._overlap_into(p, self._out)
            return p, p._norm2_dev()

        # eager run first: contraction plans and workspaces are created
        # outside the capture (neither may grow inside it)
        p, nr = body()
        ctx.sync()
        del p, nr
        with ctx.capture() as g:
            self.state, self._nr = body()
        g.keep(psi, self.state, self._nr)
        self.graph = g

    def launch(self):
        self.graph.launch()

    def read(self):
        """Download the last replay's (overlap, norm) (synchronises)."""
        h = np.zeros(3)
        L.check(L.lib.tenet_memcpy_d2h(self.ctx.h, h.ctypes.data_as(C.c_void_p), C.c_void_p(self._out), 24, None))
        nr = self._nr.to_host().reshape(()).item()
        return complex(h[0], h[1]) * math.exp(h[2]), math.sqrt(abs(nr))

    def run(self):
        self.launch()
        return self.read()

    def close(self):
        self.graph.close()
        if self._out:
            self.ctx.free(self._out)
            self._out = 0
