"""SLEEK for MI355X: the SLEEK float32/float64 codec as hand-written HIP
kernels for gfx950, behind a C ABI (include/sleek_gpu.h) and the
reference's C++ drop-in headers (include/sleek_lossless_gpu.h,
include/sleek_lossy_gpu.h).

Modules: ``codec`` (torch-tensor front end), ``batch`` (persistent batched
codec), ``serve`` (continuous-batching request loop), ``datagen`` (seeded
synthetic workloads).  Importing ``codec`` needs the built library
``libsleek_amd.so``; there is no CPU fallback.
"""
__version__ = "0.1.0"
