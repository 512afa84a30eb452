"""dolfin_amd: MI355X-native cell-integral assembly engine with DOLFIN 2019.1's assembler
interface (Assembler / SystemAssembler / assemble / assemble_system) for P1 Poisson and
vector-P2 elasticity forms, an adjoint (reverse-mode) assembly and batched assembly.

The compute path is libdolfin_amd.so (HIP for gfx950, C ABI in include/dolfin_amd.h);
torch provides device memory and streams.
"""
from . import _lib
from .mesh import Mesh, MeshColoring, UnitCubeMesh, UnitIntervalMesh, UnitSquareMesh, jitter_interior
from .function import (Function, FunctionSpace, VectorFunctionSpace, node_reordering,
                       parameters)
from .la import Matrix, Vector, solve
from .fem import (Assembler, DirichletBC, FacetForm, Form, FormSum, SystemAssembler, assemble,
                  assemble_system, body_force_form, elasticity_form, engine_for, exterior_facets,
                  facet_load_form, facet_mass_form, facet_traction_form, load_form,
                  mark_exterior_facets, mass_form, poisson_form)

SCATTER_AUTO = _lib.SCATTER_AUTO
SCATTER_ATOMIC = _lib.SCATTER_ATOMIC
SCATTER_BLOCKED = _lib.SCATTER_BLOCKED
SCATTER_COLORED = _lib.SCATTER_COLORED

__all__ = [
    "Mesh", "MeshColoring", "UnitCubeMesh", "UnitIntervalMesh", "UnitSquareMesh", "jitter_interior",
    "Function", "FunctionSpace", "VectorFunctionSpace", "node_reordering", "parameters",
    "Matrix", "Vector", "solve",
    "Assembler", "DirichletBC", "Form", "SystemAssembler", "assemble", "assemble_system",
    "body_force_form", "elasticity_form", "engine_for", "load_form", "mass_form", "poisson_form",
    "FacetForm", "FormSum", "exterior_facets", "facet_load_form", "facet_mass_form",
    "facet_traction_form", "mark_exterior_facets",
    "SCATTER_AUTO", "SCATTER_ATOMIC", "SCATTER_BLOCKED", "SCATTER_COLORED",
]
