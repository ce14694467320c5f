"""hindemith_amd -- MI355X-native backend for Hindemith's Horn-Schunck hot path.

``from hindemith_amd.hindemith import *`` gives the reference's composition
front end (``hm``, ``defines``, ``Stencil``, ``SpMat2D``, ``sum2d`` and numpy);
``hindemith_amd.kernels`` exposes the kernels on torch device tensors.
"""
from ._lib import HindemithError  # noqa: F401

__version__ = "0.1.0"
