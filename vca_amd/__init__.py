"""vca_amd -- MI355X-native Variational Cluster Approximation solver core.

The product is the HIP library ``libvca_gpu.so`` (C ABI in
``include/vca_gpu.h``); this package is its Python host mirror with the
reference's entry-point names, plus input generators for the workloads.
"""
from .api import (VcaParams, VcaSolver, Sector, VcaError,  # noqa: F401
                  vca_init_solver)
from . import configs  # noqa: F401
