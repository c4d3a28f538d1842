"""MI355X (gfx950) kernel-independent FMM matvec with 3bem's operator surface.

The package directory name starts with a digit, so import it with
``importlib.import_module("3bem_amd")``. The Python names mirror tbempy.ThreeD.
"""
from .ThreeD import (FMMConfig, NBodyData, FMMOperator, DirectOperator, Kernel,  # noqa: F401
                     LaplaceSingle, LaplaceDouble, LaplaceHypersingular, ElasticDisplacement,
                     ElasticTraction, ElasticAdjointTraction, ElasticHypersingular, IdentityScalar,
                     IdentityTensor, make_kernel, make_direct_nbody_operator, nbody_eval,
                     kernel_eval, surface_points, pinv, SparseOperator, IntegralOperator)
from ._lib import LIB_PATH, TbemError, load  # noqa: F401
