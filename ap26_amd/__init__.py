"""ap26_amd -- MI355X-native (gfx950) AP26 search: hand-written HIP kernels
behind a C ABI (include/ap26_gpu.h), a C host with the reference's operator
surface (AP26 KMIN KMAX SHIFT, state files, SOL records, result hash), and
Python mirrors of both for tests and benchmarks.
"""
__version__ = "0.1.0"
