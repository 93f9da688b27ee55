"""ctypes loader for libsphgrav_hip.so (the HIP gravity library).

The library is built in-tree by ``make -C starsmasher_amd/csrc`` (or
``__graft_entry__.build()``).  Loading fails loudly when it is missing: there is
no CPU fallback for the gravity path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsphgrav_hip.so")

SG_OK = 0
SG_ERR_ARG = 1
SG_ERR_NKERNEL = 2
SG_ERR_MASS = 3
SG_ERR_RANGE2 = 4
SG_ERR_NONFINITE = 5
SG_ERR_HIP = 6
SG_ERR_NO_DEVICE = 7
SG_ERR_STATE = 8
SG_ERR_CONFIG = 9
SG_ERR_SIZE = 10

# sg_compute_device_static: largest n, and the bit flags it writes to *err
SG_STATIC_MAX_N = 32768
SG_VALID_MASS = 1
SG_VALID_RANGE2 = 2
SG_VALID_NONFINITE = 4

SG_PREC_FP32C = 0
SG_PREC_FP64 = 1

# every symbol include/sphgrav.h declares
EXPORTS = (
    "firsthalf_grav_forces_", "lasthalf_grav_forces_", "gpu_init_dev_",
    "sg_status_string", "sg_last_status", "sg_device_count", "sg_create", "sg_destroy",
    "sg_set_stream", "sg_compute_device", "sg_compute_batch_device", "sg_set_batch_config",
    "sg_get_batch_config", "sg_synchronize", "sg_last_kernel_ms", "sg_compute_host",
    "sg_set_config", "sg_get_config", "sg_last_timings", "sg_last_tile_stats", "sg_init_device",
    "sg_firsthalf", "sg_lasthalf", "sg_vjp_device", "sg_vjp_host", "sg_compute_psi_device",
    "sg_compute_psi_host", "sg_psi_device", "sg_psi_host", "sg_grav_psi", "sg_potential_energy",
    "sg_last_precision_stats", "sg_compute_device_static", "sg_last_config",
)


class SphgravError(RuntimeError):
    def __init__(self, status: int, where: str = ""):
        self.status = status
        msg = status_string(status) if _lib is not None else f"status {status}"
        super().__init__(f"{where}: {msg} (status {status})" if where else msg)


_lib = None

_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


def load() -> C.CDLL:
    """Load the library (once).  Raises FileNotFoundError when not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"{LIB_PATH} is not built; run `make -C starsmasher_amd/csrc` "
            "(the gravity path has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    L.firsthalf_grav_forces_.argtypes = [_ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _ip, _ip]
    L.firsthalf_grav_forces_.restype = None
    L.lasthalf_grav_forces_.argtypes = [_ip, _dp, _dp, _dp, _dp, _ip]
    L.lasthalf_grav_forces_.restype = None
    L.gpu_init_dev_.argtypes = [_ip, _dp]
    L.gpu_init_dev_.restype = None
    L.sg_status_string.argtypes = [C.c_int]
    L.sg_status_string.restype = C.c_char_p
    L.sg_last_status.argtypes = []
    L.sg_device_count.argtypes = [_ip]
    L.sg_create.argtypes = [C.c_int, C.POINTER(_vp)]
    L.sg_destroy.argtypes = [_vp]
    L.sg_compute_device.argtypes = [_vp, C.c_int, C.c_int, C.c_int] + [_vp] * 5 + [C.c_int, C.c_int] + [_vp] * 4
    L.sg_compute_host.argtypes = [_vp, C.c_int, C.c_int, C.c_int] + [_dp] * 5 + [C.c_int, C.c_int] + [_dp] * 4
    L.sg_synchronize.argtypes = [_vp]
    L.sg_set_stream.argtypes = [_vp, _vp]
    L.sg_compute_batch_device.argtypes = ([_vp, C.c_int, C.POINTER(C.c_longlong)] + [_vp] * 5
                                          + [C.c_int, C.c_int, C.c_int] + [_vp] * 4)
    L.sg_compute_batch_device.restype = C.c_int
    L.sg_set_batch_config.argtypes = [_vp, C.c_int, C.c_int, C.c_int]
    L.sg_set_batch_config.restype = C.c_int
    L.sg_get_batch_config.argtypes = [_vp, C.c_int, _ip, _ip]
    L.sg_get_batch_config.restype = C.c_int
    L.sg_set_stream.restype = C.c_int
    L.sg_last_kernel_ms.argtypes = [_vp, C.POINTER(C.c_float)]
    L.sg_set_config.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.sg_get_config.argtypes = [_vp, C.c_int, _ip, _ip, _ip]
    L.sg_last_config.argtypes = [_vp, _ip, _ip, _ip]
    L.sg_last_timings.argtypes = [_vp] + [C.POINTER(C.c_float)] * 3
    L.sg_last_tile_stats.argtypes = [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.sg_last_precision_stats.argtypes = [_vp] + [C.POINTER(C.c_longlong)] * 4
    L.sg_init_device.argtypes = [C.c_int]
    L.sg_firsthalf.argtypes = [C.c_int, C.c_int, C.c_int] + [_dp] * 5 + [C.c_int]
    L.sg_lasthalf.argtypes = [_dp] * 4
    L.sg_compute_psi_device.argtypes = ([_vp, C.c_int, C.c_int, C.c_int] + [_vp] * 5
                                        + [C.c_int, C.c_int] + [_vp] * 5)
    L.sg_compute_psi_host.argtypes = ([_vp, C.c_int, C.c_int, C.c_int] + [_dp] * 5
                                      + [C.c_int, C.c_int] + [_dp] * 5)
    L.sg_psi_device.argtypes = [_vp, C.c_int, C.c_int, C.c_int] + [_vp] * 5 + [C.c_int, _vp]
    L.sg_psi_host.argtypes = [_vp, C.c_int, C.c_int, C.c_int] + [_dp] * 5 + [C.c_int, _dp]
    L.sg_grav_psi.argtypes = [C.c_int, C.c_int, C.c_int] + [_dp] * 5 + [C.c_int, _dp]
    L.sg_vjp_device.argtypes = [_vp, C.c_int] + [_vp] * 5 + [C.c_int, C.c_int] + [_vp] * 9
    L.sg_vjp_host.argtypes = [_vp, C.c_int] + [_dp] * 5 + [C.c_int, C.c_int] + [_dp] * 9
    L.sg_potential_energy.argtypes = [_vp, C.c_int, _vp, _vp, C.POINTER(C.c_double)]
    L.sg_compute_device_static.argtypes = ([C.c_int, C.c_int, C.c_int] + [_vp] * 5
                                           + [C.c_int, C.c_int] + [_vp] * 4 + [_vp, _vp])
    for name in ("sg_last_status", "sg_device_count", "sg_create", "sg_destroy", "sg_compute_device",
                 "sg_compute_host", "sg_synchronize", "sg_last_kernel_ms", "sg_set_config",
                 "sg_get_config", "sg_last_timings", "sg_last_tile_stats", "sg_init_device",
                 "sg_firsthalf", "sg_lasthalf", "sg_vjp_device", "sg_vjp_host",
                 "sg_compute_psi_device", "sg_compute_psi_host", "sg_psi_device",
                 "sg_psi_host", "sg_grav_psi", "sg_potential_energy", "sg_last_precision_stats",
                 "sg_compute_device_static", "sg_last_config"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def status_string(status: int) -> str:
    return load().sg_status_string(int(status)).decode()


def check(status: int, where: str = "") -> None:
    if status != SG_OK:
        raise SphgravError(status, where)
