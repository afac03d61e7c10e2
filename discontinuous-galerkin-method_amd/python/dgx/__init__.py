"""dgx: Python side of the MI355X nodal-DG elastic solver (mesh generators, reference formats and a
ctypes binding of libdgx.so)."""
from . import lib, mesh, formats, problem, configs  # noqa: F401
from .mesh import Mesh, structured  # noqa: F401
from .problem import Model, SourceSpec, Sponge, Setup, Solver, build_setup  # noqa: F401
