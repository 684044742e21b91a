"""ferrix_amd -- MI355X (gfx950) backend for FerriX's static-step hot path.

Element stiffness / internal force (C3D4, C3D20), deterministic CSR
assembly, penalty boundary conditions and Jacobi-PCG (FP64 and mixed
FP32/FP64) run as hand-written HIP kernels in ``libferrix_hip.so`` behind the
C ABI of ``include/ferrix_hip.h``.  This package holds the build script, the
ctypes binding and the host-side model ingestion (``.inp`` parser, synthetic
workloads).  Importing it does not touch the GPU.
"""
__all__ = ["inp", "model", "meshgen", "device", "build"]
