"""stonefish_amd: Stonefish's geometry-based hydrodynamics on MI355X (gfx950).

The compute path is libsfhydro.so (HIP kernels behind the C ABI declared in
include/sfhydro.h); this package holds its Python binding, the host-side mirror
of the reference interface (hydro.py), input generators (meshes, waves,
workloads) and the build helper.
"""
__version__ = "0.1.0"
