"""MI355X photon-transport engine with GGEMS's interfaces.

``from ggems_amd import *`` exposes GGEMS's Python classes (GGEMS,
GGEMSXRaySource, GGEMSVoxelizedPhantom, GGEMSCTSystem,
GGEMSDosimetryCalculator, the managers, ...) over libggems.so, whose
transport runs in the HIP engine (ggems_amd.engine).
"""

from .ggems import *  # noqa: F401,F403
from .ggems import GGEMSError  # noqa: F401
