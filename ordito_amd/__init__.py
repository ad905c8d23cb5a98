"""
ordito_amd: ordito's heat-method geodesic path, rebuilt for the AMD Instinct MI355X (gfx950).

``ordito_amd.heat`` keeps ``ordito.heat``'s signatures (``heat_geodesic``, ``heat_operators``,
``HeatSolver``); ``ordito_amd.mesh.Trimesh`` is the per-mesh cache that keeps factorizations resident;
``ordito_amd.batch`` and ``ordito_amd.serve`` pack many meshes per launch and serve query streams.
The compute runs in ``libheatgeo.so`` (hand-written HIP for CDNA4) through a ctypes C ABI.
"""

from . import creation, heat, laplacian, mesh  # noqa: F401
from .heat import HeatSolver, heat_geodesic, heat_operators  # noqa: F401
from .mesh import Trimesh  # noqa: F401

__version__ = "0.1.0"
