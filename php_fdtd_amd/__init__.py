"""php_fdtd_amd - MI355X-native per-timestep update of the PhP_FDTD solver.

The compute path is libphpfdtd_hip.so (hand-written gfx950 HIP kernels behind
the C ABI of include/phpfdtd.h, sources in ``csrc/``); this package holds its
ctypes binding (``_abi``), a thin Python mirror of that ABI (``sim``:
``Simulation``, ``Batch``), the continuous-batching server (``serve``), the
medium generator and ``medium_map.bin`` reader/writer (``medium``), the seeded
workload builders with reference-format writers (``workloads``), and the
reference-side C binding for fdtd.c (``binding/fdtd_binding.c``).  Importing
the package does not load the library.
"""
__all__ = ["_abi"]
