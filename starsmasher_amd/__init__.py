"""starsmasher_amd -- MI355X (gfx950) direct-summation softened self-gravity.

A from-scratch HIP implementation of StarSmasher's gravity library
(pbe/src/SPHgrav_lib2): per-particle acceleration and potential from
positions, masses and smoothing lengths with per-pair cubic-spline (or
Wendland) softening on H_i/H_j.  The product is ``libsphgrav_hip.so`` (C ABI
in ``include/sphgrav.h``), a drop-in for the reference's three Fortran
symbols; this package adds the Python mirror of that interface
(:mod:`starsmasher_amd.gravity`) and the workload generators
(:mod:`starsmasher_amd.ics`).
"""
from . import _lib
from .gravity import (SG_PREC_FP32C, SG_PREC_FP64, Gravity, SphgravError, bonet_psi,
                      device_count, firsthalf_grav_forces, gpu_init_dev, gravforce,
                      lasthalf_grav_forces)

__all__ = [
    "Gravity", "SphgravError", "SG_PREC_FP32C", "SG_PREC_FP64", "device_count",
    "firsthalf_grav_forces", "lasthalf_grav_forces", "gpu_init_dev", "gravforce", "bonet_psi",
    "_lib",
]
