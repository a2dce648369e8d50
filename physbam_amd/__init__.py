"""physbam_amd: MI355X-native explicit APIC MPM step (PhysBAM Hybrid_Methods).

The product is libmpmx.so (HIP kernels for gfx950 behind the C ABI in
include/mpmx.h); ``physbam_amd.mpmx`` is its ctypes binding.
"""
from . import mpmx  # noqa: F401
