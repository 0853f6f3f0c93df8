"""hipacc_amd -- MI355X (gfx950) execution path for Hipacc's data-parallel
operators.

The product is the C-ABI library ``libhipacc_gfx950.so`` (include/
hipacc_hip.h) and the C++ runtime headers under ``runtime/`` that generated
host code calls.  This Python package binds the same C ABI for tests,
benchmarks and torch autograd integration:

* ``capi``       ctypes declarations of the C ABI, library loader
* ``operators``  descriptor builders over torch device tensors
* ``masks``      mask tables of the workloads
* ``build``      in-tree build of the library
"""
__version__ = "0.1"
