"""uqcreator_amd -- MI355X-native IPM moment-solver hot path of UQCreator.

The product is the native library libuqk.so (HIP kernels for gfx950 behind the
C ABI of include/uqk.h, plus the C++ host layer).  ``native`` binds it with
ctypes; ``workloads`` builds the synthetic benchmark inputs.
"""
__all__ = ["native", "workloads"]
