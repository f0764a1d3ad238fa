"""polybin3d_amd: MI355X-native binned power-spectrum / bispectrum estimators.

Drop-in for the reference's estimator path (pb.PolyBin3D / pb.PSpec / pb.BSpec) with
backend='hip'.  The compute path is libpb3d.so (hand-written HIP kernels for gfx950), bound with
ctypes in _capi.py; it must be built in-tree (python -m polybin3d_amd.build) before use.
"""
from .base import PolyBin3D
from .pspec import PSpec, compute_fisher_mixed
from .bspec import BSpec
from .graph import FisherGraph
from . import autograd

__all__ = ["PolyBin3D", "PSpec", "BSpec", "FisherGraph", "autograd", "compute_fisher_mixed"]
