"""smearfem -- MI355X (gfx950) assemble-and-solve for smearFEM.jl's hot path.

Python mirror of the reference's entry points over the C ABI in
include/smearfem.h (library: ``libsmearfem_hip.so`` in this directory).
"""
from .errors import (ArgumentError, BoundsError, DeviceError, DeviceMemoryError, DimensionMismatch,
                     NotConvergedError, SingularException, SmearFEMError, StateError, UndefVarError,
                     UnsupportedError)
from .fem import (Context, Mesh, StiffnessMatrix, assemble_system, basis_function, default_context,
                  gaussian_quadrature)
from .mesh import StructuredMesh, clamp_x0, face_load, random_face_load, structured_mesh
from .solve import BoundaryConditions, System, solve_system, stream_copy_bench
from .batch import Batch, Problem, autotune, make_problem, uneven_hex8_problems, uneven_hex8_specs
from .serve import Server, random_request, run_poisson
from .spec import ProblemSpec, cantilever_spec, random_request_spec
from .surface import (DeviceMesh, PlaneRule, apply_boundary_conditions, bc_planes, csr_add, displacement_sweep,
                      face_traction_loads, meshgrid, setboundaryCond)
from .vtk import write_scene, write_vtk

__all__ = [
    "ArgumentError", "BoundsError", "DeviceError", "DeviceMemoryError", "DimensionMismatch",
    "NotConvergedError", "SingularException", "SmearFEMError", "StateError", "UndefVarError",
    "UnsupportedError", "Context", "Mesh", "StiffnessMatrix", "assemble_system", "basis_function",
    "default_context", "gaussian_quadrature", "StructuredMesh", "clamp_x0", "face_load",
    "random_face_load", "structured_mesh", "BoundaryConditions", "System", "solve_system",
    "stream_copy_bench", "Batch", "Problem", "autotune", "make_problem", "uneven_hex8_problems", "Server",
    "DeviceMesh", "PlaneRule", "apply_boundary_conditions", "bc_planes", "csr_add", "displacement_sweep", "face_traction_loads",
    "meshgrid", "setboundaryCond", "write_scene", "write_vtk",
    "random_request", "run_poisson", "uneven_hex8_specs", "ProblemSpec", "cantilever_spec",
    "random_request_spec",
]
