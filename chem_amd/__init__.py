"""MI355X-native DF Fock build and SOS-AO-MP2 engine (host layer).

The compute path is libmchip.so (HIP kernels for gfx950 behind the C ABI in
include/mchip.h); this package binds it with ctypes and provides the
descriptors and drivers around it.
"""
__all__ = ["desc", "ints"]
