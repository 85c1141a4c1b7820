"""solidfem_amd: MI355X-native matrix-free FEM elasticity (FemPhysicsMatrixFree
Newton-CG path of SolidFEM), HIP kernels behind the C-ABI of include/sfem_capi.h."""
from . import mesh  # noqa: F401

__all__ = ["mesh", "capi"]
