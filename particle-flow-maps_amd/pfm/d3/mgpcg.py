"""3D/mgpcg.py surface: the MGPCG solver classes."""
from ..mgpcg import MGPCG, MGPCG_2, MGPCG_3  # noqa: F401
