"""gpuocean_amd: MI355X-native CDKLM16 rotating shallow-water step.

Drop-in for the reference's CDKLM16 path (SWEsimulators/CDKLM16.py,
ensembles/OceanModelEnsemble.py) built on hand-written gfx950 HIP kernels.
"""

from . import Common  # noqa: F401
from .CDKLM16 import CDKLM16  # noqa: F401
from .forcing import AtmosphericPressure, WindStress  # noqa: F401
from .GPUDrifterCollection import GPUDrifterCollection  # noqa: F401
from .OceanStateNoise import OceanStateNoise  # noqa: F401
from .OceanModelEnsemble import Autotuner, BatchedEnsemble, OceanModelEnsemble  # noqa: F401
from .DataAssimilationUtils import ObservationType  # noqa: F401
from .Observation import Observation  # noqa: F401
from .ParticleInfo import ParticleInfo  # noqa: F401

__all__ = ["CDKLM16", "Common", "WindStress", "AtmosphericPressure", "GPUDrifterCollection", "OceanStateNoise",
           "OceanModelEnsemble", "BatchedEnsemble", "Autotuner", "ObservationType", "Observation", "ParticleInfo"]
