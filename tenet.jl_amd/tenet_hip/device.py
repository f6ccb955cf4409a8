"""Device context and device-resident column-major arrays (the Julia layout
the C-ABI expects).  Memory comes from the context's caching pool."""
from __future__ import annotations

import ctypes as C
import math
import weakref

import numpy as np

from . import _lib as L

DTYPES = {np.dtype(np.complex128): L.C128, np.dtype(np.complex64): L.C64}
NP_OF = {L.C128: np.complex128, L.C64: np.complex64}


class Context:
    """One device (tenet_create); owns the pool, plan cache and scratch."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        L.check(L.lib.tenet_create(int(device), C.byref(h)), "tenet_create")
        self.h = h
        self.device = device
        self._closed = False
        # host-side plan objects shared by everything on this context (grouped
        # plans of the Vidal-gauge layers, ...), keyed by the caller; closed
        # before the context
        self.plan_cache = {}
        # while a graph is being captured: device blocks freed meanwhile
        # (the graph's replays still write them), released with the graph
        self._defer = None
        self._staging = None

    def close(self):
        if not self._closed:
            for p in self.plan_cache.values():
                try:
                    p.close()
                except Exception:
                    pass
            self.plan_cache = {}
            if self._staging is not None:
                self.sync()
                self._staging.free()
                self._staging = None
            L.lib.tenet_destroy(self.h)
            self._closed = True

    def sync(self, stream=None):
        L.check(L.lib.tenet_stream_sync(self.h, stream), "sync")

    def mem_info(self):
        f, t, p = C.c_size_t(), C.c_size_t(), C.c_size_t()
        L.check(L.lib.tenet_mem_info(self.h, C.byref(f), C.byref(t), C.byref(p)))
        return f.value, t.value, p.value

    def malloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        L.check(L.lib.tenet_malloc(self.h, C.c_size_t(max(int(nbytes), 1)), C.byref(p)), "tenet_malloc")
        return p.value

    def free(self, ptr: int):
        if ptr and not self._closed:
            if self._defer is not None:
                self._defer.append(ptr)
                return
            L.check(L.lib.tenet_free(self.h, C.c_void_p(ptr)), "tenet_free")

    def graph_status(self) -> int:
        """Device status word of captured work (bit 0: a captured Jacobi SVD
        ended unconverged), read and cleared (tenet_graph_status)."""
        v = C.c_int()
        L.check(L.lib.tenet_graph_status(self.h, C.byref(v)), "graph_status")
        return v.value

    def capture(self):
        """``with ctx.capture() as g:`` records every library call issued on
        the context's stream into a HIP graph (tenet_graph_begin / end);
        ``g.launch()`` replays it.  Run the sequence once eagerly first, so
        that workspaces and contraction plans exist (neither may grow during
        a capture).  No host synchronisation or download may happen inside
        the block."""
        return Graph(self)

    def staging(self, nbytes: int) -> "PinnedBuffer":
        """The context's pinned host staging buffer, at least ``nbytes``.
        Copies through it are queued on the context's stream; a caller
        rewrites a region only after a wait that follows the previous copy
        out of that region on the stream.  Growing it waits for the stream."""
        if self._staging is None or self._staging.nbytes < nbytes:
            if self._staging is not None:
                self.sync()
                self._staging.free()
            self._staging = PinnedBuffer(self, max(int(nbytes), 1 << 16) * 2)
        return self._staging

    def scalar_buffer(self) -> int:
        """A small device buffer (64 doubles) owned by the context, for
        scalar results read back right away."""
        if getattr(self, "_scalar", None) is None:
            self._scalar = self.malloc(512)
        return self._scalar

    def stream(self):
        s = C.c_void_p()
        L.check(L.lib.tenet_stream_create(self.h, C.byref(s)), "stream_create")
        return s

    def svd_precondition(self, on: bool):
        """QR-preconditioned Jacobi SVD for complex128 bond matrices (default on)."""
        L.check(L.lib.tenet_svd_set_precondition(self.h, 1 if on else 0), "svd_set_precondition")

    def stats_enable(self, family):
        L.check(L.lib.tenet_stats_enable(self.h, int(family)), "stats_enable")

    def stats_sample_every(self, every: int):
        L.check(L.lib.tenet_stats_sample_every(self.h, int(every)), "stats_sample_every")

    def stats_read(self):
        n, ms, fl, by = C.c_int64(), C.c_double(), C.c_double(), C.c_double()
        L.check(L.lib.tenet_stats_read(self.h, C.byref(n), C.byref(ms), C.byref(fl), C.byref(by)), "stats_read")
        return {"launches": n.value, "ms": ms.value, "flops": fl.value, "bytes": by.value}

    def timer_start(self, stream=None):
        L.check(L.lib.tenet_timer_start(self.h, stream))

    def timer_stop(self, stream=None) -> float:
        ms = C.c_float()
        L.check(L.lib.tenet_timer_stop(self.h, stream, C.byref(ms)))
        return ms.value


class PinnedBuffer:
    """A pinned host buffer (tenet_host_alloc) viewed as numpy arrays."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        p = C.c_void_p()
        L.check(L.lib.tenet_host_alloc(ctx.h, C.c_size_t(max(nbytes, 8)), C.byref(p)), "host_alloc")
        self.ptr = p.value
        self.nbytes = max(nbytes, 8)

    def view(self, dtype, count, offset=0):
        buf = (C.c_char * self.nbytes).from_address(self.ptr)
        return np.frombuffer(buf, dtype=dtype, count=count, offset=offset)

    def free(self):
        if self.ptr:
            L.lib.tenet_host_free(self.ctx.h, C.c_void_p(self.ptr))
            self.ptr = None


def carve(ctx: Context, shapes, dtype):
    """Device arrays of the given shapes cut from one pool block (each at a
    256-byte boundary): one allocation for a layer's worth of operands.  The
    block is released when the last of them is dropped."""
    dt = np.dtype(dtype)
    offs, tot = [], 0
    for s in shapes:
        offs.append(tot)
        tot += (math.prod(s) * dt.itemsize + 255) // 256 * 256
    owner = DeviceArray(ctx, (max(tot // dt.itemsize, 1),), dt)
    out = []
    for s, o in zip(shapes, offs):
        v = DeviceArray(ctx, s, dt, ptr=owner.ptr + o, owner=False)
        v._base = owner
        out.append(v)
    return out


def _free(ctx_ref, ptr):
    ctx = ctx_ref()
    if ctx is not None and not ctx._closed:
        if ctx._defer is not None:
            ctx._defer.append(ptr)
            return
        L.lib.tenet_free(ctx.h, C.c_void_p(ptr))


class Graph:
    """A HIP graph of the library calls made on the context's stream inside
    ``with ctx.capture() as g:``.  Every device buffer the captured calls
    touch must outlive the graph: blocks freed during the capture are held
    until ``close()``, and objects passed to ``keep()`` (the inputs and
    outputs the caller reads after a replay) are referenced by the graph."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.h = None
        self._held = []
        self._keep = []

    def __enter__(self):
        if self.ctx._defer is not None:
            raise RuntimeError("a capture is already active on this context")
        self.ctx._defer = self._held
        try:
            L.check(L.lib.tenet_graph_begin(self.ctx.h, None), "graph_begin")
        except Exception:
            self.ctx._defer = None
            raise
        return self

    def __exit__(self, exc_type, exc, tb):
        h = C.c_void_p()
        st = L.lib.tenet_graph_end(self.ctx.h, None, C.byref(h))
        self.ctx._defer = None
        if exc_type is None:
            L.check(st, "graph_end")
            self.h = h
        elif st == L.TENET_OK:
            L.lib.tenet_graph_destroy(h)
        return False

    def keep(self, *objs):
        self._keep.extend(objs)
        return self

    def launch(self):
        """Replay on the capture stream (ordered after earlier calls on it)."""
        L.check(L.lib.tenet_graph_launch(self.h, None), "graph_launch")

    def close(self):
        if self.h is not None:
            L.lib.tenet_graph_destroy(self.h)
            self.h = None
        held, self._held = self._held, []
        for p in held:
            self.ctx.free(p)
        self._keep = []


class DeviceArray:
    """A column-major complex array in device memory."""

    def __init__(self, ctx: Context, shape, dtype=np.complex128, ptr=None, owner=True):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        if self.dtype not in DTYPES:
            raise TypeError("only complex128 / complex64 device arrays")
        self.code = DTYPES[self.dtype]
        self.size = math.prod(self.shape)
        self.nbytes = self.size * self.dtype.itemsize
        self.ptr = ptr if ptr is not None else ctx.malloc(self.nbytes)
        self._fin = weakref.finalize(self, _free, weakref.ref(ctx), self.ptr) if owner else None

    @property
    def strides_el(self):
        st, acc = [], 1
        for s in self.shape:
            st.append(acc)
            acc *= s
        return st

    @staticmethod
    def from_host(ctx: Context, arr, dtype=None):
        a = np.asarray(arr)
        dt = np.dtype(dtype) if dtype is not None else (a.dtype if a.dtype in DTYPES else np.dtype(np.complex128))
        a = a.astype(dt, copy=False)
        # np.asfortranarray promotes a 0-d array to shape (1,): keep scalars 0-d
        a = np.asfortranarray(a) if a.ndim else a.reshape(())
        d = DeviceArray(ctx, a.shape, dt)
        if d.nbytes:
            L.check(L.lib.tenet_memcpy_h2d(ctx.h, C.c_void_p(d.ptr), a.ctypes.data_as(C.c_void_p),
                                           C.c_size_t(d.nbytes), None), "h2d")
        return d

    def to_host(self):
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        if self.nbytes:
            L.check(L.lib.tenet_memcpy_d2h(self.ctx.h, out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
                                           C.c_size_t(self.nbytes), None), "d2h")
        return out

    def zero_(self):
        L.check(L.lib.tenet_memset_zero(self.ctx.h, C.c_void_p(self.ptr), C.c_size_t(self.nbytes), None))
        return self

    def gaussian_(self, seed: int, offset: int = 0, stream=None):
        """Fill with the library's seeded standard complex Gaussians
        (element i keyed by (seed, offset + i)); nothing crosses PCIe."""
        L.check(L.lib.tenet_fill_gaussian(self.ctx.h, stream, self.code, self.size, C.c_void_p(self.ptr),
                                          C.c_uint64(seed), C.c_uint64(offset)), "fill_gaussian")
        return self

    def scale_(self, alpha, stream=None):
        al = (C.c_double * 2)(float(np.real(alpha)), float(np.imag(alpha)))
        L.check(L.lib.tenet_scale(self.ctx.h, stream, self.code, self.size, C.c_void_p(self.ptr), al), "scale")
        return self

    def nrm2sq(self, stream=None) -> float:
        """sum |x|^2 accumulated in FP64 on the device."""
        out = self.ctx.scalar_buffer()
        L.check(L.lib.tenet_nrm2sq(self.ctx.h, stream, self.code, self.size, C.c_void_p(self.ptr),
                                   C.c_void_p(out)), "nrm2sq")
        v = np.zeros(1)
        L.check(L.lib.tenet_memcpy_d2h(self.ctx.h, v.ctypes.data_as(C.c_void_p), C.c_void_p(out), C.c_size_t(8),
                                       stream), "d2h")
        self.ctx.sync(stream)
        return float(v[0])

    def copy(self):
        d = DeviceArray(self.ctx, self.shape, self.dtype)
        L.check(L.lib.tenet_memcpy_d2d(self.ctx.h, C.c_void_p(d.ptr), C.c_void_p(self.ptr),
                                       C.c_size_t(self.nbytes), None), "d2d")
        return d

    def reshape(self, shape):
        """Column-major reshape sharing the buffer (keeps this array alive)."""
        if int(np.prod(shape, dtype=np.int64)) != self.size:
            raise ValueError("reshape size mismatch")
        v = DeviceArray(self.ctx, shape, self.dtype, ptr=self.ptr, owner=False)
        v._base = self
        return v

    def view(self, shape):
        """Column-major view of the first prod(shape) elements (keeps this array alive)."""
        if int(np.prod(shape, dtype=np.int64)) > self.size:
            raise ValueError("view larger than the buffer")
        v = DeviceArray(self.ctx, shape, self.dtype, ptr=self.ptr, owner=False)
        v._base = self
        return v

    def free(self):
        if self._fin is not None and self._fin.alive:
            self._fin()

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype})"


def tdesc(shape, labels, strides=None):
    d = L.TDesc()
    d.rank = len(shape)
    if d.rank > L.MAX_RANK:
        raise ValueError("rank above 8")
    if strides is None:
        strides, acc = [], 1
        for s in shape:
            strides.append(acc)
            acc *= int(s)
    for i, (e, st, lab) in enumerate(zip(shape, strides, labels)):
        d.extent[i] = int(e)
        d.stride[i] = int(st)
        d.label[i] = int(lab)
    return d
