"""MI355X (gfx950) multi-prime NTT multiplication engine.

Drop-in for the reference crate's MulPlan / MulWorkspace / ConvPlan /
ConvWorkspace: exact big-integer products, squares and dense convolutions,
computed by hand-written HIP kernels behind the C ABI of include/nttanafis.h.
"""
from .api import (  # noqa: F401
    Batch,
    ConvPlan,
    ConvWorkspace,
    DeviceBuffer,
    MulPlan,
    MulWorkspace,
    Profile,
    Server,
    clock,
    device_available,
    pinned_array,
    last_error,
    mul_device_noalloc,
    mul_geometry,
    mul_schoolbook,
    mul_split,
    row_len,
    row_shape,
    synchronize,
    transform_form,
    transform_split,
)
from ._ffi import LIB_PATH, load  # noqa: F401
from .field import CrtPair, PrimeField  # noqa: F401

__all__ = [
    "Batch",
    "CrtPair",
    "PrimeField",
    "MulPlan",
    "MulWorkspace",
    "ConvPlan",
    "ConvWorkspace",
    "DeviceBuffer",
    "Server",
    "clock",
    "pinned_array",
    "Profile",
    "device_available",
    "last_error",
    "mul_device_noalloc",
    "mul_geometry",
    "mul_schoolbook",
    "mul_split",
    "row_len",
    "row_shape",
    "synchronize",
    "transform_form",
    "transform_split",
    "load",
    "LIB_PATH",
]
