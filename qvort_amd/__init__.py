"""MI355X-native filament velocity for Qvort (octree + Biot-Savart + LIA + AB).

Native library: qvort_amd/lib/libqvort_gpu.so (C ABI in include/qvort_gpu.h),
bound here by :mod:`qvort_amd.gpu`; Fortran bindings in qvort_amd/fortran.
"""
from . import geometry  # noqa: F401
from .gpu import Batch, Context, QvortError, params  # noqa: F401

__all__ = ["Batch", "Context", "QvortError", "params", "geometry"]
