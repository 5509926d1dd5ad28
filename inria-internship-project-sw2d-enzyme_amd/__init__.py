"""MI355X-native SW2D single-porosity forward step and 4D-Var adjoint.

Host mirror of the reference's solver/assimilation entry points over the
C-ABI of ``lib/libsw2dg.so`` (HIP, gfx950, FP64).  Import this package by
path (its directory name is not a Python identifier), e.g.::

    spec = importlib.util.spec_from_file_location(
        "sw2d_amd", "<repo>/inria-internship-project-sw2d-enzyme_amd/__init__.py",
        submodule_search_locations=["<repo>/inria-internship-project-sw2d-enzyme_amd"])
"""
from . import _native
from .cases import Case, Geometry, dambreak_case, read_result_map, result_name, scale_case, write_case
from .solver import SpSolver

__all__ = ["Case", "Geometry", "SpSolver", "dambreak_case", "scale_case", "write_case",
           "read_result_map", "result_name", "_native"]
