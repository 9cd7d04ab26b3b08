"""php_fdtd_amd - MI355X-native per-timestep update of the PhP_FDTD solver.

The compute path is libphpfdtd_hip.so (hand-written gfx950 HIP kernels behind
the C ABI of include/phpfdtd.h); this package holds its ctypes binding
(``_abi``), a thin Python ``Simulation`` mirror of that ABI (``sim``), the
reference-format readers/writers (``io``) and the synthetic workload builders
(``workloads``).  Importing the package does not load the library.
"""
__all__ = ["_abi"]
