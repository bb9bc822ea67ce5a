"""jet_amd -- MI355X-native pairwise tensor-contraction path of Jet.

The GPU work runs in libjet_hip.so (HIP kernels for gfx950 behind the C ABI in
include/jet_hip.h); this package is the host side, mirroring Jet's Tensor,
Permuter, TensorNetwork, PathInfo and TaskBasedContractor interfaces.
"""
from .abort import JetError
from ._native import JetHipError, Context, default_context
from .tensor import (Tensor, contract_tensors, add_tensors, slice_index, transpose, reshape,
                     conj)
from .permuter import Permuter
from . import utilities

__version__ = "0.1.0"
