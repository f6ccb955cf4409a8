"""ctypes binding of libtenet_hip.so (declarations mirror include/tenet_hip.h).

The library is loaded from the package tree (``tenet.jl_amd/lib``), or from
the path in the environment variable TENET_HIP_LIB (another build of the
same sources, e.g. ``make OUT=...``).  There is no CPU fallback: a missing
library raises at import of this module, and creating a context on a host
without a gfx950 device raises TenetError.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("TENET_HIP_LIB") or os.path.normpath(os.path.join(_HERE, "..", "lib", "libtenet_hip.so"))

TENET_OK = 0
TENET_ERR_HIP = 1
TENET_ERR_SHAPE = 2
TENET_ERR_OOM = 3
TENET_ERR_NOCONV = 4
TENET_ERR_UNSUPPORTED = 5
TENET_ERR_ARG = 6
TENET_ERR_NODEVICE = 7
STATUS_NAMES = {
    0: "OK", 1: "ERR_HIP", 2: "ERR_SHAPE", 3: "ERR_OOM", 4: "ERR_NOCONV",
    5: "ERR_UNSUPPORTED", 6: "ERR_ARG", 7: "ERR_NODEVICE",
}

C128, C64 = 0, 1
PREC_NATIVE, PREC_FP32, PREC_SPLIT_BF16, PREC_SPLIT_FP16, PREC_SPLIT_I8, PREC_SPLIT_I8_FAST = 0, 1, 2, 3, 4, 5
ABSORB_NONE, ABSORB_EQUAL, ABSORB_LEFT, ABSORB_RIGHT = 0, 1, 2, 3
QR_IN_ADJ, QR_Q_ADJ, QR_R_ADJ = 1, 2, 4
ABSORB_R_B, ABSORB_B_RH = 0, 1
KSTAT_OFF, KSTAT_SVD_ROUND, KSTAT_GEMM, KSTAT_QR, KSTAT_AMPLITUDE = 0, 1, 2, 3, 4
OP_QR, OP_SVD, OP_SIMPLE_UPDATE = 1, 2, 3
S_REAL, S_DTYPE = 0, 1
MAX_RANK = 8


class TDesc(C.Structure):
    _fields_ = [
        ("rank", C.c_int32),
        ("reserved", C.c_int32),
        ("extent", C.c_int64 * MAX_RANK),
        ("stride", C.c_int64 * MAX_RANK),
        ("label", C.c_int64 * MAX_RANK),
    ]


class TenetError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libtenet_hip.so not built ({LIB_PATH}); run __graft_entry__.build() "
            "or `make -C tenet.jl_amd/csrc`")
    return C.CDLL(LIB_PATH)


lib = _load()

_p = C.c_void_p
_i = C.c_int
_i64 = C.c_int64
_d = C.c_double
_sz = C.c_size_t
_pi64 = C.POINTER(C.c_int64)
_pd = C.POINTER(C.c_double)

# name -> argtypes (all return int unless noted)
SIGNATURES = {
    "tenet_device_count": [C.POINTER(C.c_int)],
    "tenet_create": [_i, C.POINTER(_p)],
    "tenet_destroy": [_p],
    "tenet_stream_create": [_p, C.POINTER(_p)],
    "tenet_stream_destroy": [_p, _p],
    "tenet_stream_sync": [_p, _p],
    "tenet_device_sync": [_p],
    "tenet_malloc": [_p, _sz, C.POINTER(_p)],
    "tenet_free": [_p, _p],
    "tenet_memcpy_h2d": [_p, _p, _p, _sz, _p],
    "tenet_memcpy_d2h": [_p, _p, _p, _sz, _p],
    "tenet_memcpy_d2d": [_p, _p, _p, _sz, _p],
    "tenet_memset_zero": [_p, _p, _sz, _p],
    "tenet_mem_info": [_p, C.POINTER(_sz), C.POINTER(_sz), C.POINTER(_sz)],
    "tenet_free_async": [_p, _p, _p],
    "tenet_workspace_size": [_p, _i, _i, _i64, _i64, C.POINTER(_sz)],
    "tenet_workspace_reserve": [_p, _p, _sz],
    "tenet_set_log": [_i, C.c_char_p],
    "tenet_host_alloc": [_p, _sz, C.POINTER(_p)],
    "tenet_host_free": [_p, _p],
    "tenet_memcpy_async": [_p, _p, _p, _sz, _i, _p],
    "tenet_event_create": [_p, C.POINTER(_p)],
    "tenet_event_destroy": [_p, _p],
    "tenet_event_record": [_p, _p, _p],
    "tenet_event_query": [_p, _p, C.POINTER(C.c_int)],
    "tenet_stream_wait_event": [_p, _p, _p],
    "tenet_einsum_plan": [_p, _i, _i, C.POINTER(TDesc), C.POINTER(TDesc), C.POINTER(TDesc), C.POINTER(_p)],
    "tenet_einsum_exec": [_p, _p, _p, _p, _p, _i, _i, _pd, _pd],
    "tenet_plan_destroy": [_p],
    "tenet_einsum": [_p, _p, _i, _i, C.POINTER(TDesc), _p, C.POINTER(TDesc), _p, C.POINTER(TDesc), _p, _i, _i],
    "tenet_qr_thin": [_p, _p, _i, _i64, _i64, _p, _i64, _p, _i64, _p, _i64],
    "tenet_qr_thin_ex": [_p, _p, _i, _i, _i64, _i64, _p, _i64, _p, _i64, _p, _i64],
    "tenet_qr_absorb": [_p, _p, _i, _i, _i64, _i64, _p, _i64, _p, _i64, _i, _i64, _p, _i64, _p, _i64],
    "tenet_svd_thin": [_p, _p, _i, _i64, _i64, _p, _i64, _p, _i64, _p, _p, _i64, _i64, _d, _p, _p],
    "tenet_svd_thin_noalloc": [_p, _p, _i, _i64, _i64, _p, _i64, _p, _i64, _p, _p, _i64, _i64, _d, _p, _p],
    "tenet_svd_set_precondition": [_p, _i],
    "tenet_simple_update": [_p, _p, _i, _i, _i64, _i64, _i64, _i64, _i64, _p, _p, _p, _p, _p, _p, _i64, _d, _i,
                            _p, _p, _pi64],
    "tenet_overlap": [_p, _p, _i, _i, _i64, _i64, _pi64, C.POINTER(_p), _pi64, C.POINTER(_p), _p],
    "tenet_amplitudes": [_p, _p, _i, _i64, _p, _p, _p, _p, _i64, _p],
    "tenet_amplitudes_sites": [_p, _p, _i, _i64, _i64, _i64, _p, _p, _p, _p, _i64, _p],
    "tenet_expect_pauli": [_p, _p, _i, _i, _i64, _i64, _i64, _pi64, C.POINTER(_p), _p, C.c_char_p, _p],
    "tenet_left_env": [_p, _p, _i, _i, _i64, _i64, _pi64, C.POINTER(_p), _p, _p],
    "tenet_left_envs": [_p, _p, _i, _i, _i64, _i64, _pi64, C.POINTER(_p), C.POINTER(_p), _p],
    "tenet_grouped_einsum": [_p, _p, _i, _i, _i64, C.POINTER(TDesc), C.POINTER(_p)],
    "tenet_grouped_plan": [_p, _i, _i, _i64, C.POINTER(TDesc), C.POINTER(_p)],
    "tenet_grouped_plan_conj": [_p, _i, _i, _i64, C.POINTER(TDesc), C.POINTER(C.c_int32), C.POINTER(_p)],
    "tenet_grouped_exec": [_p, _p, C.POINTER(_p)],
    "tenet_grouped_exec_dev": [_p, _p, _p],
    "tenet_grouped_plan_info": [_p, _pi64, _pi64, _pi64, _pd, _pd],
    "tenet_grouped_plan_destroy": [_p],
    "tenet_autotune_tiles": [_p, _i, _pd],
    "tenet_grouped_svd": [_p, _p, _i, _i64, _pi64, _pi64, C.POINTER(_p), C.POINTER(_p), C.POINTER(_p),
                          C.POINTER(_p), _pi64, _d, _p, _p],
    "tenet_grouped_svd_plan": [_p, _i, _i64, _pi64, _pi64, C.POINTER(_p), C.POINTER(_p), C.POINTER(_p),
                               C.POINTER(_p), _pi64, _d, _p, _p, _i, C.POINTER(_p)],
    "tenet_grouped_svd_exec": [_p, _p],
    "tenet_grouped_svd_destroy": [_p],
    "tenet_fill_gaussian": [_p, _p, _i, _i64, _p, C.c_uint64, C.c_uint64],
    "tenet_sample_select": [_p, _p, _i, _p, _i, _i, _i64, _i, _i, C.c_uint64, _p, _p],
    "tenet_scale": [_p, _p, _i, _i64, _p, _pd],
    "tenet_nrm2sq": [_p, _p, _i, _i64, _p, _p],
    "tenet_sigma_scales": [_p, _p, _i, _i64, _p, _p, _p],
    "tenet_grouped_gate": [_p, _p, _i, _i64, _i, _pi64, _pi64, C.POINTER(_p), C.POINTER(_p), C.POINTER(_p),
                           C.POINTER(_p), C.POINTER(_p)],
    "tenet_graph_begin": [_p, _p],
    "tenet_graph_end": [_p, _p, C.POINTER(_p)],
    "tenet_graph_launch": [_p, _p],
    "tenet_graph_status": [_p, C.POINTER(C.c_int)],
    "tenet_graph_destroy": [_p],
    "tenet_stats_enable": [_p, _i],
    "tenet_stats_read": [_p, _pi64, _pd, _pd, _pd],
    "tenet_stats_sample_every": [_p, _i],
    "tenet_timer_start": [_p, _p],
    "tenet_timer_stop": [_p, _p, C.POINTER(C.c_float)],
}

for _name, _args in SIGNATURES.items():
    _f = getattr(lib, _name)
    _f.argtypes = _args
    _f.restype = C.c_int
lib.tenet_version.restype = C.c_char_p
lib.tenet_version.argtypes = []
lib.tenet_last_error.restype = C.c_char_p
lib.tenet_last_error.argtypes = []

EXPORTED = sorted(list(SIGNATURES) + ["tenet_version", "tenet_last_error"])


def check(status, what=""):
    if status != TENET_OK:
        msg = lib.tenet_last_error().decode(errors="replace")
        raise TenetError(status, f"{what}: {msg}" if what else msg)
    return status


def version():
    return lib.tenet_version().decode()
