"""MI355X-native electron-transport engine with CudaMONSEL's host API.

The classes re-exported here (SEmaterial, MONSEL_MaterialScatterModel,
NormalMultiPlaneShape, Region, GaussianBeam, MonteCarloSS, BackscatterStats,
...) are pybind11 bindings of the C++ host API in host/monsel_host.hpp.  They
serialise the scene and run it on the GPU through the C-ABI of
include/monsel_gpu.h (libmonsel_gpu.so).  There is no CPU transport path:
importing this package fails if the native libraries are not built.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
DATA_DIR = os.path.join(_HERE, "data")
GPU_LIB = os.path.join(_HERE, "libmonsel_gpu.so")

try:
    from . import _monsel  # noqa: F401
except ImportError as exc:  # pragma: no cover - exercised only without a build
    raise ImportError(
        "cudamonsel_amd native extension is not built; run `python -m cudamonsel_amd.build` "
        f"(or __graft_entry__.build()) first: {exc}") from exc

from ._monsel import *  # noqa: F401,F403
from ._monsel import (NShapes, NISTMottRS, ToSI, FromSI, Elements,  # noqa: F401
                      BrowningEmpiricalCrossSection, ScreenedRutherfordScatteringAngle)

_monsel.setDataDirectory(DATA_DIR)

_gpu = None


def gpu_lib():
    """ctypes handle of libmonsel_gpu.so (the C-ABI), loaded once."""
    global _gpu
    if _gpu is None:
        _gpu = ctypes.CDLL(GPU_LIB)
        _gpu.monsel_last_error.restype = ctypes.c_char_p
    return _gpu


def last_error():
    return gpu_lib().monsel_last_error().decode()
