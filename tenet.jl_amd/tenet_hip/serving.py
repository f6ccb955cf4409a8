"""Request loop over device-resident MPS states: bitstring amplitudes
<x|psi> (overlap(::ProductState, ::MPS), overlap.jl:22-32), Pauli-string
expectations <psi|P|psi> (the overlap.jl:36-50 sweep with one-site Paulis,
ext/TenetPythonCallExt/pytket.jl:23-31 conventions) and batches of
samples (sample(psi, nsamples), src/Operations/sample.jl:20-72) with
continuous batching.

* States stay on the device in MixedCanonical(site 1) form (canonize!, then
  normalize!), so every site right of a Pauli string is a right isometry and
  the sweep stops at the string's end, and a sample request walks the
  chain from site 1 without re-canonising.
* Expectations reuse a cached left environment E(start) per (state, start)
  (tenet_left_env); the first request that needs one computes it on its own
  stream and later requests on other streams wait on its event.
* Continuous batching: requests are admitted as they arrive whenever the
  in-flight slots and the HBM budget (bytes of the states, the environment
  cache and every in-flight request's buffers) allow; all amplitude requests
  admitted in one pass of the loop share one batched launch (one workgroup
  per request), each expectation runs on the next of several streams, and
  completions are polled with events, so new work is issued while earlier
  work is still on the device.  Results do not depend on the mix: every
  kernel computes a request the same way whatever else is in flight.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from collections import OrderedDict, deque
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .device import Context, DeviceArray, PinnedBuffer, tdesc


@dataclass
class Request:
    rid: int
    state: int
    kind: str  # "amp", "pauli" or "sample"
    bits: np.ndarray = None
    start: int = 0
    paulis: str = ""
    nsamples: int = 0
    seed: int = 0
    samples: np.ndarray = None  # (nsamples, n) 0-based values of a "sample" request
    t_arrive: float = 0.0
    t_done: float = -1.0
    value: complex = None
    log_scale: float = 0.0
    t_admit: float = -1.0

    @property
    def result(self):
        if self.kind == "sample":
            return self.samples
        return self.value * math.exp(self.log_scale)


class ServedState:
    """One device-resident MPS with the device tables the batched amplitude
    kernel reads (site pointers, bond sizes)."""

    def __init__(self, ctx: Context, mps):
        self.ctx = ctx
        self.mps = mps
        self.n = mps.n
        self.chis = np.asarray(mps.chis(), dtype=np.int64)
        self.dtype = mps.dtype
        self.code = mps.code
        ptrs = np.asarray([s.ptr for s in mps.sites], dtype=np.uint64)
        self.site_tab = ctx.malloc(ptrs.nbytes)
        self.chi_tab = ctx.malloc(self.chis.nbytes)
        L.check(L.lib.tenet_memcpy_h2d(ctx.h, C.c_void_p(self.site_tab), ptrs.ctypes.data_as(C.c_void_p), ptrs.nbytes,
                                       None))
        L.check(L.lib.tenet_memcpy_h2d(ctx.h, C.c_void_p(self.chi_tab), self.chis.ctypes.data_as(C.c_void_p),
                                       self.chis.nbytes, None))
        self.site_ptrs = (C.c_void_p * self.n)(*[int(p) for p in ptrs])
        self.chi_arr = (C.c_int64 * (self.n + 1))(*[int(c) for c in self.chis])
        self.bytes = sum(s.nbytes for s in mps.sites)

    def env_bytes(self, start):
        return int(self.chis[start]) ** 2 * self.dtype.itemsize + 64

    def close(self):
        self.ctx.free(self.site_tab)
        self.ctx.free(self.chi_tab)


AMP_REQ_BYTES = 64  # device table entries + result per amplitude request
SITE_PARALLEL_CHI = 64  # batches with a bond this large take tenet_amplitudes_sites


def request_footprint(st: ServedState, r) -> int:
    """Device bytes a request holds while in flight (its tables and result,
    or the sweep's two environments and two site buffers)."""
    if r.kind == "amp":
        # tables, bits, result, and the site-parallel path's two vector rows
        return AMP_REQ_BYTES + st.n + 32 * int(np.max(st.chis))
    if r.kind == "sample":
        mx = int(np.max(st.chis))
        d = max(s.shape[1] for s in st.mps.sites)
        return (2 * mx + d * mx) * r.nsamples * st.dtype.itemsize + 4 * r.nsamples * st.n + 3 * 256
    w = len(r.paulis)
    chi = st.chis[r.start:r.start + w + 1]
    mx_env = int(np.max(chi)) ** 2
    mx_site = int(np.max(chi[:-1] * 2 * chi[1:]))
    return (2 * mx_env + 2 * mx_site) * st.dtype.itemsize + 64


class _PinnedPool:
    """Pinned buffers by power-of-two size class, reused across requests
    (pinning pages is far slower than a request)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.free = {}

    def get(self, nbytes: int) -> PinnedBuffer:
        cls = 1 << max(6, (int(nbytes) - 1).bit_length())
        lst = self.free.get(cls)
        if lst:
            return lst.pop()
        return PinnedBuffer(self.ctx, cls)

    def put(self, p: PinnedBuffer):
        self.free.setdefault(p.nbytes, []).append(p)

    def close(self):
        for lst in self.free.values():
            for p in lst:
                p.free()
        self.free = {}


class _Event:
    def __init__(self, ctx: Context):
        self.ctx = ctx
        e = C.c_void_p()
        L.check(L.lib.tenet_event_create(ctx.h, C.byref(e)))
        self.h = e

    def record(self, stream):
        L.check(L.lib.tenet_event_record(self.ctx.h, self.h, stream))

    def done(self) -> bool:
        d = C.c_int()
        L.check(L.lib.tenet_event_query(self.ctx.h, self.h, C.byref(d)))
        return d.value == 1

    def destroy(self):
        L.lib.tenet_event_destroy(self.ctx.h, self.h)


@dataclass
class _Batch:
    reqs: list
    kind: str
    event: _Event
    pinned: PinnedBuffer
    dev: list = field(default_factory=list)
    footprint: int = 0


class Server:
    """Continuous-batching request loop.

    n_streams      device streams that in-flight work is spread over
    max_inflight   admitted-but-unfinished request slots
    hbm_budget     bytes the states, the environment cache and in-flight
                   buffers may hold together (None: 90% of free memory)
    """

    def __init__(self, ctx: Context, states, n_streams=4, max_inflight=512, max_amp_batch=1024, hbm_budget=None,
                 env_cache_bytes=1 << 30, prefill_envs=False):
        self.ctx = ctx
        self.states = list(states)
        self.streams = [ctx.stream() for _ in range(n_streams)]
        self.max_inflight = max_inflight
        self.max_amp_batch = max_amp_batch
        free, total, _ = ctx.mem_info()
        self.state_bytes = sum(s.bytes for s in self.states)
        self.budget = int(0.9 * free) + self.state_bytes if hbm_budget is None else int(hbm_budget)
        if self.state_bytes > self.budget:
            raise ValueError("the states alone exceed the HBM budget")
        self.used = self.state_bytes
        self.env_cache = OrderedDict()  # (state, start) -> dict(env, log pinned, event, bytes)
        self.env_cache_bytes = env_cache_bytes
        self.env_used = 0
        self._rr = 0
        self.events_free = []
        self.pins = _PinnedPool(ctx)
        self.env_fixed = {}  # (state, start) -> {"env": ptr, "logval": float}, resident for the server's life
        if prefill_envs:
            self._prefill()
        self.stats = {"amp_batches": 0, "pauli_launches": 0, "sample_launches": 0, "env_built": 0, "env_hits": 0,
                      "admission_waits": 0}

    def _prefill(self):
        """Every left environment of every state, one sweep per state
        (tenet_left_envs), kept resident: HBM holds them easily, and then no
        expectation request pays for the prefix of its sweep."""
        for sid, st in enumerate(self.states):
            need = sum(st.env_bytes(j) for j in range(1, st.n))
            if self.used + need > self.budget:
                break
            ptrs = [None] + [self.ctx.malloc(st.env_bytes(j)) for j in range(1, st.n)]
            logs = self.ctx.malloc(8 * st.n)
            arr = (C.c_void_p * st.n)(*ptrs)
            L.check(L.lib.tenet_left_envs(self.ctx.h, None, st.code, L.PREC_NATIVE, st.n - 1, 2, st.chi_arr,
                                          st.site_ptrs, arr, C.c_void_p(logs)), "left_envs")
            lg = np.zeros(st.n)
            L.check(L.lib.tenet_memcpy_d2h(self.ctx.h, lg.ctypes.data_as(C.c_void_p), C.c_void_p(logs), lg.nbytes,
                                           None))
            self.ctx.free(logs)
            for j in range(1, st.n):
                self.env_fixed[(sid, j)] = {"env": ptrs[j], "logval": float(lg[j]), "bytes": st.env_bytes(j)}
            self.used += need

    # ---- helpers -------------------------------------------------------------
    def _stream(self):
        s = self.streams[self._rr % len(self.streams)]
        self._rr += 1
        return s

    def _event(self):
        return self.events_free.pop() if self.events_free else _Event(self.ctx)

    def footprint(self, r: Request) -> int:
        return request_footprint(self.states[r.state], r)

    # ---- launches --------------------------------------------------------------
    def _launch_amps(self, reqs, now):
        ctx = self.ctx
        nreq = len(reqs)
        nb = sum(self.states[r.state].n for r in reqs)
        tab_bytes = 32 * nreq
        pin = self.pins.get(tab_bytes + nb + 24 * nreq + 64)
        tab = pin.view(np.uint64, 4 * nreq)
        bits = pin.view(np.uint8, nb, offset=tab_bytes)
        dev = ctx.malloc(tab_bytes + nb + 24 * nreq + 256)
        p_bits = dev + tab_bytes
        p_out = (p_bits + nb + 63) // 64 * 64
        o = 0
        for i, r in enumerate(reqs):
            st = self.states[r.state]
            tab[i] = st.n
            tab[nreq + i] = st.site_tab
            tab[2 * nreq + i] = st.chi_tab
            tab[3 * nreq + i] = p_bits + o
            bits[o:o + st.n] = r.bits
            o += st.n
        s = self._stream()
        L.check(L.lib.tenet_memcpy_async(ctx.h, C.c_void_p(dev), C.c_void_p(pin.ptr), tab_bytes + nb, 0, s))
        code = self.states[reqs[0].state].code
        n_max = max(self.states[r.state].n for r in reqs)
        chi_max = max(int(np.max(self.states[r.state].chis)) for r in reqs)
        if chi_max >= SITE_PARALLEL_CHI:
            # one launch per site position, every request's slice product
            # spread over ceil(chi / 16) workgroups
            L.check(L.lib.tenet_amplitudes_sites(ctx.h, s, code, nreq, n_max, chi_max, C.c_void_p(dev),
                                                 C.c_void_p(dev + 8 * nreq), C.c_void_p(dev + 16 * nreq),
                                                 C.c_void_p(dev + 24 * nreq), 2, C.c_void_p(p_out)),
                    "amplitudes_sites")
        else:
            L.check(L.lib.tenet_amplitudes(ctx.h, s, code, nreq, C.c_void_p(dev), C.c_void_p(dev + 8 * nreq),
                                           C.c_void_p(dev + 16 * nreq), C.c_void_p(dev + 24 * nreq), 2,
                                           C.c_void_p(p_out)), "amplitudes")
        # results land after the request tables in the same pinned buffer
        L.check(L.lib.tenet_memcpy_async(ctx.h, C.c_void_p(pin.ptr + tab_bytes + nb), C.c_void_p(p_out),
                                         24 * nreq, 1, s))
        ev = self._event()
        ev.record(s)
        self.stats["amp_batches"] += 1
        return _Batch(reqs, "amp", ev, pin, [dev], sum(self.footprint(r) for r in reqs))

    def _env(self, st_id, start, stream):
        key = (st_id, start)
        ent = self.env_fixed.get(key)
        if ent is not None:
            self.stats["env_hits"] += 1
            return ent
        ent = self.env_cache.get(key)
        if ent is not None:
            self.env_cache.move_to_end(key)
            self.stats["env_hits"] += 1
            L.check(L.lib.tenet_stream_wait_event(self.ctx.h, stream, ent["event"].h))
            return ent
        st = self.states[st_id]
        nbytes = st.env_bytes(start)
        while self.env_cache and self.env_used + nbytes > self.env_cache_bytes:
            _, old = self.env_cache.popitem(last=False)
            self._pending_env_free.append(old)
            self.env_used -= old["bytes"]
            self.used -= old["bytes"]
        env = self.ctx.malloc(nbytes)
        out = self.ctx.malloc(64)
        pin = self.pins.get(24)
        L.check(L.lib.tenet_left_env(self.ctx.h, stream, st.code, L.PREC_NATIVE, start, 2, st.chi_arr, st.site_ptrs,
                                     C.c_void_p(env), C.c_void_p(out)), "left_env")
        L.check(L.lib.tenet_memcpy_async(self.ctx.h, C.c_void_p(pin.ptr), C.c_void_p(out), 24, 1, stream))
        ev = _Event(self.ctx)
        ev.record(stream)
        ent = {"env": env, "out": out, "log": pin, "event": ev, "bytes": nbytes}
        self.env_cache[key] = ent
        self.env_used += nbytes
        self.used += nbytes
        self.stats["env_built"] += 1
        return ent

    def _launch_pauli(self, r):
        ctx = self.ctx
        st = self.states[r.state]
        s = self._stream()
        env_ptr = None
        ent = None
        if r.start > 0:
            ent = self._env(r.state, r.start, s)
            env_ptr = C.c_void_p(ent["env"])
        out = ctx.malloc(64)
        pin = self.pins.get(24)
        L.check(L.lib.tenet_expect_pauli(ctx.h, s, st.code, L.PREC_NATIVE, r.start, len(r.paulis), 2, st.chi_arr,
                                         st.site_ptrs, env_ptr, r.paulis.encode(), C.c_void_p(out)), "expect_pauli")
        L.check(L.lib.tenet_memcpy_async(ctx.h, C.c_void_p(pin.ptr), C.c_void_p(out), 24, 1, s))
        ev = self._event()
        ev.record(s)
        b = _Batch([r], "pauli", ev, pin, [out], self.footprint(r))
        b.env = ent
        self.stats["pauli_launches"] += 1
        return b

    def _launch_sample(self, r):
        """The sampling sweep of MPS.sample on one stream: V (1 at site 1) .
        A_i for all samples, then tenet_sample_select per site; the values
        land in the request's pinned buffer."""
        ctx = self.ctx
        st = self.states[r.state]
        if st.mps.form != ("mixed", 0, 0):
            raise ValueError("sample requests need the served state canonised to its first site")
        ns, n = int(r.nsamples), st.n
        item = st.dtype.itemsize
        mx = int(np.max(st.chis))
        d = max(x.shape[1] for x in st.mps.sites)
        vbuf = [ctx.malloc(mx * ns * item + 64), ctx.malloc(mx * ns * item + 64)]
        tbuf = ctx.malloc(d * mx * ns * item + 64)
        bits = ctx.malloc(4 * ns * n + 64)
        pin = self.pins.get(max(4 * ns * n, ns * item) + 64)
        s = self._stream()
        ones = pin.view(st.dtype, ns)
        ones[:] = 1.0
        L.check(L.lib.tenet_memcpy_async(ctx.h, C.c_void_p(vbuf[0]), C.c_void_p(pin.ptr), ns * item, 0, s))
        v = vbuf[0]
        for i in range(n):
            l, dd, rr = st.mps.sites[i].shape
            da, db, dc = tdesc((l, ns), [1, 4]), tdesc((l, dd, rr), [1, 2, 3]), tdesc((dd, rr, ns), [2, 3, 4])
            L.check(L.lib.tenet_einsum(ctx.h, s, st.code, L.PREC_NATIVE, C.byref(da), C.c_void_p(v), C.byref(db),
                                       C.c_void_p(st.mps.sites[i].ptr), C.byref(dc), C.c_void_p(tbuf), 0, 0),
                    "einsum")
            vn = vbuf[(i + 1) % 2] if i < n - 1 else None
            L.check(L.lib.tenet_sample_select(ctx.h, s, st.code, C.c_void_p(tbuf), dd, rr, ns, n, i,
                                              C.c_uint64(int(r.seed)), C.c_void_p(bits),
                                              C.c_void_p(vn) if vn is not None else None), "sample_select")
            v = vn
        # the values overwrite the ones in the pinned buffer (same stream, after the upload)
        L.check(L.lib.tenet_memcpy_async(ctx.h, C.c_void_p(pin.ptr), C.c_void_p(bits), 4 * ns * n, 1, s))
        ev = self._event()
        ev.record(s)
        self.stats["sample_launches"] += 1
        return _Batch([r], "sample", ev, pin, vbuf + [tbuf, bits], self.footprint(r))

    def _finish(self, b: _Batch, now):
        if b.kind == "sample":
            r = b.reqs[0]
            r.samples = b.pinned.view(np.int32, r.nsamples * self.states[r.state].n).reshape(r.nsamples, -1).copy()
            r.t_done = now
        elif b.kind == "amp":
            nreq = len(b.reqs)
            nb = sum(self.states[r.state].n for r in b.reqs)
            res = b.pinned.view(np.float64, 3 * nreq, offset=32 * nreq + nb).reshape(nreq, 3)
            for r, (re, im, lg) in zip(b.reqs, res):
                r.value, r.log_scale, r.t_done = complex(re, im), float(lg), now
        else:
            r = b.reqs[0]
            re, im, lg = b.pinned.view(np.float64, 3)
            extra = 0.0
            if b.env is not None:
                extra = b.env["logval"] if "logval" in b.env else float(b.env["log"].view(np.float64, 3)[2])
            r.value, r.log_scale, r.t_done = complex(re, im), float(lg) + extra, now
        for p in b.dev:
            self.ctx.free(p)
        self.pins.put(b.pinned)
        self.events_free.append(b.event)
        self.used -= b.footprint

    # ---- the loop --------------------------------------------------------------
    def run(self, requests, arrivals=None, clock=time.perf_counter):
        """Serve ``requests``; ``arrivals`` (seconds from the start, sorted)
        releases each request at its time (None: all at once).  Returns the
        wall time of the run."""
        reqs = list(requests)
        arr = np.zeros(len(reqs)) if arrivals is None else np.asarray(arrivals, dtype=np.float64)
        order = np.argsort(arr, kind="stable")
        self._pending_env_free = []
        ready = deque()
        inflight = []
        nxt = 0
        t0 = clock()
        done = 0
        while done < len(reqs):
            now = clock() - t0
            while nxt < len(reqs) and arr[order[nxt]] <= now:
                r = reqs[order[nxt]]
                r.t_arrive = float(arr[order[nxt]])
                ready.append(r)
                nxt += 1
            # completions
            still = []
            for b in inflight:
                if b.event.done():
                    self._finish(b, clock() - t0)
                    done += len(b.reqs)
                else:
                    still.append(b)
            inflight = still
            # admission + launch
            n_inflight = sum(len(b.reqs) for b in inflight)
            amps = []
            while ready and n_inflight + len(amps) < self.max_inflight:
                r = ready[0]
                fp = self.footprint(r)
                if r.kind == "pauli" and r.start > 0 and (r.state, r.start) not in self.env_fixed \
                        and (r.state, r.start) not in self.env_cache:
                    fp += self.states[r.state].env_bytes(r.start)
                if self.used + fp > self.budget:
                    if not inflight and not amps:
                        raise RuntimeError(f"request {r.rid} needs {fp} bytes and can never be admitted "
                                           f"({self.used} of {self.budget} held by states and cached environments)")
                    self.stats["admission_waits"] += 1
                    break
                ready.popleft()
                r.t_admit = clock() - t0
                if r.kind == "amp":
                    self.used += self.footprint(r)
                    amps.append(r)
                    if len(amps) >= self.max_amp_batch:
                        inflight.append(self._launch_amps(amps, now))
                        n_inflight += len(amps)
                        amps = []
                elif r.kind == "sample":
                    self.used += self.footprint(r)
                    inflight.append(self._launch_sample(r))
                    n_inflight += 1
                else:
                    self.used += self.footprint(r)
                    inflight.append(self._launch_pauli(r))
                    n_inflight += 1
            if amps:
                inflight.append(self._launch_amps(amps, now))
            if not inflight and not ready and nxt < len(reqs):
                # idle until the next arrival
                wait = arr[order[nxt]] - (clock() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.001))
        for s in self.streams:
            self.ctx.sync(s)
        for ent in self._pending_env_free:
            self._free_env(ent)
        self._pending_env_free = []
        return clock() - t0

    def _free_env(self, ent):
        self.ctx.free(ent["env"])
        self.ctx.free(ent["out"])
        self.pins.put(ent["log"])
        ent["event"].destroy()

    def close(self):
        for s in self.streams:
            self.ctx.sync(s)
        for ent in self.env_cache.values():
            self._free_env(ent)
        self.env_cache.clear()
        for ent in self.env_fixed.values():
            self.ctx.free(ent["env"])
        self.env_fixed.clear()
        for e in self.events_free:
            e.destroy()
        self.events_free = []
        self.pins.close()
        for s in self.streams:
            L.lib.tenet_stream_destroy(self.ctx.h, s)
        self.streams = []


def latency_summary(reqs):
    lat = np.array([r.t_done - r.t_arrive for r in reqs])
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "max_ms": float(lat.max() * 1e3)}
