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
from .path_info import PathInfo, PathStepInfo
from .tensor_network import TensorNetwork
from .task_based_contractor import TaskBasedContractor
from .tensor_network_io import TensorFileException, TensorNetworkFile, TensorNetworkSerializer
from .factory import (TensorC64, TensorC128, TensorNetworkC64, TensorNetworkC128,
                      TaskBasedContractorC64, TaskBasedContractorC128, TensorNetworkFileC64,
                      TensorNetworkFileC128, TensorNetworkSerializerC64,
                      TensorNetworkSerializerC128)
from . import factory, utilities

__version__ = "0.1.0"
