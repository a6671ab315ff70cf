"""The reference's own call sequence through the C ABI, as the Julia shim makes it.

``assemble_system(ne, NodeList, IEN, ndim, "Q1", nDof, ID, Young, nu)``
(src/fem.jl:135) maps to one ``sfem_assemble`` call with the Julia arrays
as they are (column-major, 1-based); the example's elimination and solve
(examples/vector3D.jl:314-322) to ``sfem_bc_create`` + ``sfem_eliminate`` +
``sfem_system_set_grid`` + ``sfem_system_solve`` + ``sfem_system_solution``.
The checks:
* sfem_assemble detects a meshgrid-numbered mesh and assembles it from the
  grid: the K it returns has the sort-based K's pattern and, its element
  matrices being multiples of the first, values summed from the element
  scales within 1e-12 of the sort-based K's (bitwise with
  SFEM_ASM_EXACT_SUMS); it reports its grid,
  and the multigrid built without element counts applies the finest level
  matrix-free (``matrix_free`` 2), the path the benchmark measures;
* the system holds its own reference on K: K destroyed right after
  elimination, the solve and the reactions still read it;
* the solution meets the 1e-8 bar against the oracle's direct solve.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import smearfem
from smearfem import _lib
from smearfem.mesh import clamp_x0, element_young_loguniform, random_face_load, structured_mesh

pytestmark = pytest.mark.gpu


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def sfem_assemble(ctx, nodes, ien, ndim, ndof, idm, young, nu, n_elem):
    lib = ctx._lib
    nodes = np.asfortranarray(nodes, dtype=np.float64)
    ien = np.asfortranarray(ien, dtype=np.int64)
    idm = np.asfortranarray(idm, dtype=np.int64)
    young = np.ascontiguousarray(np.atleast_1d(young), dtype=np.float64)
    K = C.c_void_p()
    ctx.check(lib.sfem_assemble(ctx.handle, ndim, 1, ndof, _p(nodes, C.c_double), nodes.shape[1],
                                _p(ien, C.c_int64), ien.shape[0], ien.shape[1], n_elem, _p(idm, C.c_int64),
                                idm.shape[0], _p(young, C.c_double), young.size, float(nu), C.byref(K)))
    return K


def export(ctx, K):
    lib = ctx._lib
    m, n, nnz = C.c_int64(), C.c_int64(), C.c_int64()
    ctx.check(lib.sfem_csr_info(K, C.byref(m), C.byref(n), C.byref(nnz), None))
    cp = np.zeros(n.value + 1, dtype=np.int64)
    rv = np.zeros(nnz.value, dtype=np.int64)
    nz = np.zeros(nnz.value)
    ctx.check(lib.sfem_export_csc1(ctx.handle, K, _p(cp, C.c_int64), _p(rv, C.c_int64), _p(nz, C.c_double)))
    return cp, rv, nz


@pytest.mark.parametrize("case", ["cube_meshgrid_ref", "box_per_element_E"])
def test_reference_call_sequence_reaches_matrix_free_multigrid(ctx, case):
    print(f"[case] drop-in {case}", flush=True)
    lib = ctx._lib
    if case == "cube_meshgrid_ref":
        # the example's own mesh (examples/vector3D.jl:10-130): ne = 6 cube
        ne = 6
        nodes, ien, idm = O.meshgrid_ref(0, 1, 0, 1, 0, 1, ne, 3)
        dims, young, nu, n_elem = (ne, ne, ne), 40.0, 0.4, ne ** 3
    else:
        dims = (10, 6, 4)
        m = structured_mesh(dims, lengths=(2.5, 1.5, 1.0))
        nodes, ien, idm, n_elem = m.NodeList, m.IEN, m.ID, m.n_elem
        young, nu = element_young_loguniform(n_elem, 0.1, 10.0, seed=12), 0.3
    K = sfem_assemble(ctx, nodes, ien, 3, 3, idm, young, nu, n_elem)
    nd = C.c_int()
    ne3 = np.zeros(3, dtype=np.int64)
    ctx.check(lib.sfem_csr_grid(K, C.byref(nd), _p(ne3, C.c_int64)))
    assert nd.value == 3 and tuple(ne3) == dims
    got = export(ctx, K)
    # the sort-based K of the same inputs (no grid declared)
    mesh = smearfem.Mesh(ctx, n_elem, nodes, ien, 3, "Q1", 3, idm, young, nu)
    Ks = mesh.assemble()
    assert Ks.grid is None
    ref = Ks.csc_arrays()
    Ks.close()
    mesh.close()
    fs, dev = C.c_int(), C.c_double()
    ctx.check(lib.sfem_csr_grid_values(K, C.byref(fs), C.byref(dev)))
    assert fs.value == 1 and dev.value <= 1e-13
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    assert np.abs(got[2] - ref[2]).max() <= 1e-12 * np.abs(ref[2]).max()
    # loads and constraints: z = 0 face clamped, random loads elsewhere
    n = got[0].size - 1
    z = np.asarray(nodes)[2]
    fixed = np.repeat((z == 0.0).astype(np.uint8), 3)
    f = np.where(fixed == 0, np.random.default_rng(3).standard_normal(n), 0.0)
    bc, sys_ = C.c_void_p(), C.c_void_p()
    ctx.check(lib.sfem_bc_create(ctx.handle, n, _p(f, C.c_double), _p(fixed, C.c_uint8), None, C.byref(bc)))
    ctx.check(lib.sfem_eliminate(ctx.handle, K, bc, C.byref(sys_)))
    lib.sfem_bc_destroy(bc)
    lib.sfem_csr_destroy(K)  # the system keeps K alive
    ctx.check(lib.sfem_system_set_grid(ctx.handle, sys_, 0, None, None))
    info = _lib.MgInfo()
    ctx.check(lib.sfem_system_mg_info(sys_, C.byref(info)))
    assert info.matrix_free == 2, info.matrix_free
    o = _lib.SolveOpts()
    lib.sfem_solve_opts_default(C.byref(o))
    o.rtol = 1e-12
    o.preconditioner = 2
    st = _lib.SolveStats()
    ctx.check(lib.sfem_system_solve(ctx.handle, sys_, C.byref(o), C.byref(st)))
    q = np.zeros(n)
    R = np.zeros(n)
    ctx.check(lib.sfem_system_solution(ctx.handle, sys_, _p(q, C.c_double), _p(R, C.c_double)))
    lib.sfem_system_destroy(sys_)  # releases K too
    Kr = O.assemble_system(0, nodes, ien, 3, "Q1", 3, idm, young, nu, n_elem=n_elem)
    Kff, rhs = O.eliminate(Kr, fixed, f)
    import scipy.sparse.linalg as spla

    u_ref = spla.spsolve(Kff.tocsc(), rhs)
    q_ref = O.recombine(fixed, None, u_ref)
    assert np.linalg.norm(q - q_ref) / np.linalg.norm(q_ref) <= 1e-8
    R_ref = Kr @ q_ref - f
    assert np.abs(R - np.where(fixed == 1, R_ref, 0.0)).max() <= 1e-8 * np.abs(R_ref).max()


def test_non_grid_mesh_keeps_the_sorted_path(ctx):
    """A mesh whose numbering is not the meshgrid's is assembled by sorting,
    and multigrid without element counts is refused with ArgumentError."""
    print("[case] drop-in non-grid mesh", flush=True)
    lib = ctx._lib
    m = structured_mesh((4, 3, 2))
    ien = m.IEN.copy()
    ien[[3, 5]] = ien[[5, 3]]  # two elements swapped: same K, not the meshgrid order
    K = sfem_assemble(ctx, m.NodeList, ien, 3, 3, m.ID, 1.0, 0.3, m.n_elem)
    nd = C.c_int(-1)
    ctx.check(lib.sfem_csr_grid(K, C.byref(nd), None))
    assert nd.value == 0
    Kr = O.assemble_system(0, m.NodeList, ien, 3, "Q1", 3, m.ID, 1.0, 0.3, n_elem=m.n_elem)
    cp, rv, nz = export(ctx, K)
    assert np.array_equal(cp - 1, Kr.indptr) and np.array_equal(rv - 1, Kr.indices)
    assert np.abs(nz - Kr.data).max() <= 1e-12 * np.abs(Kr.data).max()
    fixed = clamp_x0(m)
    f = random_face_load(m, 2, 1, fixed, seed=2)
    bc, sys_ = C.c_void_p(), C.c_void_p()
    n = cp.size - 1
    ctx.check(lib.sfem_bc_create(ctx.handle, n, _p(f, C.c_double), _p(fixed, C.c_uint8), None, C.byref(bc)))
    ctx.check(lib.sfem_eliminate(ctx.handle, K, bc, C.byref(sys_)))
    lib.sfem_bc_destroy(bc)
    assert lib.sfem_system_set_grid(ctx.handle, sys_, 0, None, None) == smearfem.ArgumentError.code
    lib.sfem_system_destroy(sys_)
    lib.sfem_csr_destroy(K)
