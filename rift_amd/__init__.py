"""rift_amd -- MI355X-native Kurganov-Petrova shallow-water step.

The product is the C ABI in include/rift_kp.h and include/rift_host.h,
implemented by librift_kp.so (hand-written gfx950 HIP kernels + C++ host
code).  This package holds the sources (csrc/, host/), the build (Makefile),
a ctypes binding (kp.py) and the synthetic-input generators (synth.py).
"""
