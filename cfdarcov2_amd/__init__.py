"""MI355X-native explicit finite-volume path of cfdARCO.

The product is the HIP library ``libcfdarco_hip.so`` (C-ABI in
``include/cfdarco.h``) plus the C++ operator API under ``cpp/``; this Python
package loads the library (``capi``), builds it (``build``), generates the
synthetic benchmark inputs (``workloads``) and wraps the adjoint for torch
(``adjoint``).
"""
__all__ = ["capi", "workloads", "build"]
