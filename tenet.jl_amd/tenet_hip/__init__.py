"""tenet_hip: host side of the MI355X engine for Tenet.jl's contraction,
QR/SVD-truncation and transfer-matrix hot path.  All arithmetic runs in
libtenet_hip.so (HIP, gfx950); this package only plans and calls it."""
from . import _lib
from ._lib import (ABSORB_EQUAL, ABSORB_LEFT, ABSORB_NONE, ABSORB_RIGHT, C64, C128, PREC_FP32, PREC_NATIVE,
                   PREC_SPLIT_BF16, PREC_SPLIT_FP16, TenetError, version)
from .device import Context, DeviceArray
from .mps import MPS
from .tensor import Tensor, binary_einsum, permute, simple_update, tensor_qr_thin, tensor_svd_thin
from .vidal import VidalMPS, brickwork_layers, tebd_vidal_sweep
from .mpo import MPO, evolve_direct_, evolve_zipup_

__all__ = [
    "Context", "DeviceArray", "MPS", "Tensor", "binary_einsum", "permute", "simple_update", "tensor_qr_thin",
    "tensor_svd_thin", "TenetError", "version", "C128", "C64", "PREC_NATIVE", "PREC_FP32", "PREC_SPLIT_BF16",
    "PREC_SPLIT_FP16", "ABSORB_NONE", "ABSORB_EQUAL", "ABSORB_LEFT", "ABSORB_RIGHT", "VidalMPS", "brickwork_layers",
    "tebd_vidal_sweep", "MPO", "evolve_direct_", "evolve_zipup_",
]
