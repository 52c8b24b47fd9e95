"""jax_fem_amd: MI355X-native element assembly -> linear solve -> adjoint path of jax-fem.

Modules mirror the reference package: ``generate_mesh`` (Mesh, box_mesh),
``problem`` (Problem), ``solver`` (solver, ad_wrapper), ``fe`` and ``basis``.
Device work runs in libfem_mi355x.so (hand-written HIP for gfx950).
"""
import logging

logger = logging.getLogger('jax_fem_amd')
if not logger.handlers:
    _h = logging.StreamHandler()
    _h.setFormatter(logging.Formatter('[%(asctime)s][%(levelname)s] %(name)s: %(message)s'))
    logger.addHandler(_h)
    logger.setLevel(logging.WARNING)
    logger.propagate = False

from .generate_mesh import Mesh, box_mesh, box_mesh_hex27, get_meshio_cell_type  # noqa: E402
from .problem import Problem  # noqa: E402
from .solver import solver, ad_wrapper  # noqa: E402
from . import materials  # noqa: E402

__all__ = ['Mesh', 'box_mesh', 'box_mesh_hex27', 'get_meshio_cell_type', 'Problem', 'solver', 'ad_wrapper',
           'materials', 'logger']
