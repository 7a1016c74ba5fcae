"""MI355X (gfx950) parallel toplesets propagation geodesics.

The compute path is libgproshan_ptp.so (HIP kernels + C ABI, include/gproshan_ptp.h);
``gproshan_amd.geodesics`` is its Python binding and a mirror of gproshan's
``geodesics`` PTP front end; ``gproshan_amd.meshgen`` holds the seeded mesh
generators used by the tests and benchmarks.
"""
from . import meshgen  # noqa: F401
