"""jaccpot_amd -- MI355X-native FMM evaluation path with jaccpot's API.

The fixed-topology, real-basis (Dehnen) FMM of TobiBu/jaccpot
(``FastMultipoleMethod.compute_accelerations / prepare_state /
evaluate_prepared_state``) on hand-written HIP kernels for gfx950, reached
through the C ABI of ``libjfm.so`` (include/jfm.h).
"""

from .config import (FarFieldConfig, FMMAdvancedConfig, GradConfig, NearFieldConfig, RuntimePolicyConfig,
                     TreeConfig)
from .direct import direct_sum_gravitational_acceleration
from .mutual import BlockStepFMM
from .solver import FastMultipoleMethod, FMMPreparedState
from .topology import Topology, build_topology

__all__ = [
    "FarFieldConfig",
    "FMMAdvancedConfig",
    "GradConfig",
    "NearFieldConfig",
    "RuntimePolicyConfig",
    "TreeConfig",
    "BlockStepFMM",
    "FastMultipoleMethod",
    "FMMPreparedState",
    "Topology",
    "build_topology",
    "direct_sum_gravitational_acceleration",
]
