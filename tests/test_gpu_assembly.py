"""GPU parity of assemble_system (src/fem.jl:135-256) against the CPU oracle.

Bar: CSR/CSC pattern bit-exact (colptr, rowval identical), values within
1e-12 normwise (max |dK| / max |K|), K bitwise symmetric, and bitwise
identical run to run and between the node-block and scalar key paths.
"""
import numpy as np
import pytest

import oracle as O
import smearfem
from smearfem.mesh import element_young_loguniform, structured_mesh

pytestmark = pytest.mark.gpu

TOL = 1e-12


def gpu_csc(ctx, m, ndim, cls="Q1", ndof=None, young=1.0, nu=0.3, force_scalar=False, coords=None, ID="mesh",
            detect_grid=True):
    ndof = ndim if ndof is None else ndof
    idm = (m.ID if ndof > 1 else None) if isinstance(ID, str) else ID
    K = smearfem.assemble_system(0, m.NodeList if coords is None else coords, m.IEN, ndim, cls, ndof, idm, young,
                                 nu, ctx=ctx, device=True, force_scalar=force_scalar, n_elem=m.n_elem,
                                 detect_grid=detect_grid)
    # the drop-in assembles meshgrid-numbered vector meshes from their grid
    expect_grid = detect_grid and not force_scalar and cls == "Q1" and ndof == ndim and isinstance(ID, str)
    assert (K.grid is not None) == expect_grid, (K.grid, expect_grid)
    arrays = K.csc_arrays()
    shape, bs = K.shape, K.block
    K.close()
    return shape, bs, arrays


def oracle_csc(m, ndim, cls="Q1", ndof=None, young=1.0, nu=0.3, coords=None, ID="mesh"):
    ndof = ndim if ndof is None else ndof
    idm = (m.ID if ndof > 1 else None) if isinstance(ID, str) else ID
    E, J, V = O.assemble_coo(m.n_elem, m.NodeList if coords is None else coords, m.IEN, ndim, cls, ndof, idm,
                             young, nu)
    mm, nn, colptr, rowval, nzval = O.sparse_csc(E, J, V)
    return (mm, nn), (colptr, rowval, nzval)


def assert_parity(g, o):
    (gshape, _, (gcp, grv, gnz)), (oshape, (ocp, orv, onz)) = g, o
    assert gshape == oshape
    assert np.array_equal(gcp, ocp), "colptr differs"
    assert np.array_equal(grv, orv), "rowval differs"
    scale = np.abs(onz).max()
    assert np.abs(gnz - onz).max() <= TOL * scale


def symmetric_bitwise(shape, cp, rv, nz):
    import scipy.sparse as sp

    K = sp.csc_matrix((nz, rv - 1, cp - 1), shape=shape)
    D = (K - K.T).tocoo()
    return D.nnz == 0 or np.abs(D.data).max() == 0.0


CASES = [
    ("cfg1_q4_32x32_plane_strain", (32, 32), "Q1", 2),
    ("q4_laplace_7x5", (7, 5), "Q1", 1),
    ("hex8_elastic_6x3x2", (6, 3, 2), "Q1", 3),
    ("hex8_laplace_5x4x3", (5, 4, 3), "Q1", 1),
]


@pytest.mark.parametrize("detect_grid", [True, False], ids=["grid", "sorted"])
@pytest.mark.parametrize("name,dims,cls,ndof", CASES, ids=[c[0] for c in CASES])
def test_assembly_parity(ctx, name, dims, cls, ndof, detect_grid):
    print(f"[case] {name} detect_grid={detect_grid}", flush=True)
    m = structured_mesh(dims, lengths=[1.0] * len(dims))
    ndim = len(dims)
    E, nu = 1.0, 0.3
    if name.startswith("cfg1"):  # plane strain through the plane-stress code
        E, nu = 1.0 / (1 - 0.3 ** 2), 0.3 / (1 - 0.3)
    g = gpu_csc(ctx, m, ndim, cls, ndof, E, nu, detect_grid=detect_grid)
    assert_parity(g, oracle_csc(m, ndim, cls, ndof, E, nu))
    assert symmetric_bitwise(g[0], *g[2])


def test_q2_laplace_parity(ctx):
    # 9-node mesh of 3x2 elements, nodes numbered on a (2*nx+1) x (2*ny+1) grid
    nx, ny = 3, 2
    gx, gy = 2 * nx + 1, 2 * ny + 1
    X, Y = np.meshgrid(np.linspace(0, 1.5, gx), np.linspace(0, 1, gy))
    nodes = np.vstack([X.ravel(), Y.ravel()])
    nodes += 0.02 * np.random.default_rng(2).standard_normal(nodes.shape)
    nid = lambda i, j: j * gx + i + 1
    ien = []
    for ey in range(ny):
        for ex in range(nx):
            i, j = 2 * ex, 2 * ey
            ien.append([nid(i, j), nid(i + 2, j), nid(i + 2, j + 2), nid(i, j + 2), nid(i + 1, j), nid(i + 2, j + 1),
                        nid(i + 1, j + 2), nid(i, j + 1), nid(i + 1, j + 1)])
    ien = np.array(ien, dtype=np.int64)
    K = smearfem.assemble_system(0, nodes, ien, 2, "Q2", 1, None, 1.0, 0.3, ctx=ctx, device=True, n_elem=len(ien))
    g = (K.shape, K.block, K.csc_arrays())
    K.close()
    E, J, V = O.assemble_coo(len(ien), nodes, ien, 2, "Q2", 1)
    mm, nn, cp, rv, nz = O.sparse_csc(E, J, V)
    assert_parity(g, ((mm, nn), (cp, rv, nz)))


def test_distorted_per_element_young_and_block_vs_scalar_paths(ctx):
    m = structured_mesh((5, 4, 3), lengths=(2.0, 1.0, 0.7))
    rng = np.random.default_rng(9)
    coords = m.NodeList + 0.03 * rng.standard_normal(m.NodeList.shape)
    young = element_young_loguniform(m.n_elem, 0.1, 10.0, seed=4)
    gb = gpu_csc(ctx, m, 3, young=young, coords=coords, detect_grid=False)
    gs = gpu_csc(ctx, m, 3, young=young, coords=coords, force_scalar=True)
    gg = gpu_csc(ctx, m, 3, young=young, coords=coords)  # the grid the mesh is detected as
    assert gb[1] == 3 and gs[1] == 1 and gg[1] == 3
    for a, b, c in zip(gb[2], gs[2], gg[2]):  # the three paths sum in the same order
        assert np.array_equal(a, b) and np.array_equal(a, c)
    assert_parity(gb, oracle_csc(m, 3, young=young, coords=coords))
    assert symmetric_bitwise(gb[0], *gb[2])


def test_run_to_run_bitwise(ctx):
    m = structured_mesh((8, 6, 5))
    young = element_young_loguniform(m.n_elem, 0.1, 10.0, seed=1)
    a = gpu_csc(ctx, m, 3, young=young)
    b = gpu_csc(ctx, m, 3, young=young)
    for x, y in zip(a[2], b[2]):
        assert np.array_equal(x, y)


def test_non_blocked_id_and_shared_dofs(ctx):
    # dof-major numbering (ID[m,l] = (l-1)*nNodes + m) and two nodes tied to
    # the same dofs: only the scalar key path applies; sums follow COO order
    m = structured_mesh((3, 2))
    n = m.n_nodes
    ID = np.stack([np.arange(1, n + 1), n + np.arange(1, n + 1)], axis=1).astype(np.int64)
    g = gpu_csc(ctx, m, 2, ID=ID)
    assert g[1] == 1
    assert_parity(g, oracle_csc(m, 2, ID=ID))
    ID2 = m.ID.copy()
    ID2[5] = ID2[2]  # node 6 shares node 3's dofs
    g2 = gpu_csc(ctx, m, 2, ID=ID2)
    assert_parity(g2, oracle_csc(m, 2, ID=ID2))


def test_cube_drop_in_signature(ctx):
    # assemble_system(ne, NodeList, IEN, ndim, "Q1", nDof, ID, Young, nu) with
    # ne^ndim elements, exactly as examples/vector3D.jl:305 calls it
    NL, IEN, ID = O.meshgrid_ref(0, 1, 0, 1, 0, 1, 3, 3)
    K = smearfem.assemble_system(3, NL, IEN, 3, "Q1", 3, ID, 40, 0.4, ctx=ctx)
    R = O.assemble_system(3, NL, IEN, 3, "Q1", 3, ID, 40, 0.4)
    assert K.shape == R.shape
    assert np.array_equal(K.indptr, R.indptr) and np.array_equal(K.indices, R.indices)
    assert np.abs(K.data - R.data).max() <= TOL * np.abs(R.data).max()


def test_error_behaviour_matches_reference(ctx):
    m = structured_mesh((3, 3))

    def both(nodes, ien, ndim, cls, ndof, idm, n_elem=None):
        n_elem = len(ien) if n_elem is None else n_elem
        try:
            E, J, V = O.assemble_coo(n_elem, nodes, ien, ndim, cls, ndof, idm)
            O.sparse_csc(E, J, V)
            ocode = 0
        except O.OracleError as e:
            ocode = e.code
        try:
            smearfem.assemble_system(0, nodes, ien, ndim, cls, ndof, idm, ctx=ctx, n_elem=n_elem)
            gcode = 0
        except smearfem.SmearFEMError as e:
            gcode = e.code
        return ocode, gcode

    bad = m.IEN.copy(); bad[4, 2] = 0
    assert both(m.NodeList, bad, 2, "Q1", 2, m.ID) == (-2, -2)
    coords = m.NodeList.copy(); coords[:, m.IEN[3] - 1] = 0.25
    assert both(coords, m.IEN, 2, "Q1", 2, m.ID) == (-4, -4)
    # singular element after a bad index: the earlier element wins
    bad2 = m.IEN.copy(); bad2[6, 0] = 10_000
    assert both(coords, bad2, 2, "Q1", 2, m.ID) == (-4, -4)
    ID0 = m.ID.copy(); ID0[7, 0] = 0
    assert both(m.NodeList, m.IEN, 2, "Q1", 2, ID0) == (-1, -1)
    short = m.ID[:5]
    assert both(m.NodeList, m.IEN, 2, "Q1", 2, short) == (-2, -2)
    assert both(m.NodeList, m.IEN, 2, "Q1", 3, np.hstack([m.ID, m.ID[:, :1]])) == (-3, -3)
    assert both(m.NodeList, m.IEN, 2, "Q3", 1, None) == (-5, -5)
    assert both(m.NodeList[:1], m.IEN[:, :2], 1, "Q1", 1, None) == (-3, -3)
    assert both(m.NodeList, m.IEN, 2, "Q1", 1, None, n_elem=12) == (-2, -2)  # ne^ndim > size(IEN,1)


def _mesh_assemble(ctx, m, young, coords=None, grid=None, nu=0.3, exact_sums=False, values=None):
    mesh = smearfem.Mesh(ctx, m.n_elem, m.NodeList if coords is None else coords, m.IEN, m.ndim, "Q1", m.ndim,
                         m.ID, young, nu)
    if grid is not None:
        mesh.set_grid(grid)
    K = mesh.assemble(exact_sums=exact_sums)
    out = K.csc_arrays()
    if values is not None:
        values.update(K.grid_values())
    K.close()
    mesh.close()
    return out


@pytest.mark.parametrize("dims", [(12, 6, 5), (3, 1, 2), (9, 7), (2, 2)])
def test_structured_assembly_from_scales(ctx, dims):
    """On a declared uniform grid every element matrix is s_e Ke(0) to
    rounding, so K's values are summed from the scales: the pattern of the
    sorted path, values within 1e-12 of the element sums (max |dK| / max |K|;
    bitwise those sums with exact_sums), K bitwise symmetric and the same bits
    run to run.  Distorted elements keep the element sums."""
    print(f"[case] structured assembly from scales {dims}", flush=True)
    m = structured_mesh(dims, lengths=(1.7, 1.1, 0.9)[: len(dims)])
    young = element_young_loguniform(m.n_elem, 0.1, 10.0, seed=41)
    ref = _mesh_assemble(ctx, m, young)
    info = {}
    got = _mesh_assemble(ctx, m, young, grid=dims, values=info)
    assert info["from_scales"] and info["ke_deviation"] <= 1e-13, info
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    assert np.abs(got[2] - ref[2]).max() <= 1e-12 * np.abs(ref[2]).max()
    shape = (got[0].size - 1,) * 2
    assert symmetric_bitwise(shape, *got)
    again = _mesh_assemble(ctx, m, young, grid=dims)
    assert all(np.array_equal(a, b) for a, b in zip(got, again))
    info = {}
    exact = _mesh_assemble(ctx, m, young, grid=dims, exact_sums=True, values=info)
    assert not info["from_scales"]
    assert all(np.array_equal(a, b) for a, b in zip(exact, ref))
    # distorted elements: not multiples of Ke(0), the element sums
    rng = np.random.default_rng(41)
    coords = m.NodeList + 0.05 / max(dims) * rng.uniform(-1, 1, m.NodeList.shape)
    info = {}
    dist = _mesh_assemble(ctx, m, young, coords, grid=dims, values=info)
    assert not info["from_scales"] and info["ke_deviation"] > 1e-13
    assert all(np.array_equal(a, b) for a, b in zip(dist, _mesh_assemble(ctx, m, young, coords)))


@pytest.mark.parametrize("dims", [(6, 4, 3), (2, 1, 1), (1, 3, 2), (7, 5), (1, 1)])
def test_structured_assembly_bitwise_equals_sorted(ctx, dims):
    """Declaring the grid replaces keys / sort / segment sums by a gather from
    the grid; the pattern and every value must be bitwise those of the
    sorted path (per-element E, distorted coordinates)."""
    print(f"[case] structured assembly {dims}", flush=True)
    m = structured_mesh(dims)
    rng = np.random.default_rng(len(dims) * 100 + sum(dims))
    young = element_young_loguniform(m.n_elem, 0.1, 10.0, seed=7)
    coords = m.NodeList + 0.1 / max(dims) * rng.uniform(-1, 1, m.NodeList.shape)
    ref = _mesh_assemble(ctx, m, young, coords)
    got = _mesh_assemble(ctx, m, young, coords, grid=dims)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_structured_assembly_meshgrid_and_errors(ctx):
    print("[case] structured assembly meshgrid + errors", flush=True)
    from smearfem.surface import meshgrid
    # sfem_meshgrid meshes are declared grids: equal to the same mesh uploaded and sorted
    dm, _ = meshgrid(0, 1, 0, 1, 0, 1, 5, 3, inflate=True, Young=40.0, nu=0.4, ctx=ctx)
    nodes, ien, idm = dm.export()
    Kg = dm.assemble()
    got = Kg.csc_arrays()
    Kg.close()
    dm.close()
    up = smearfem.Mesh(ctx, ien.shape[0], nodes, ien, 3, "Q1", 3, idm, 40.0, 0.4)
    Ku = up.assemble()
    ref = Ku.csc_arrays()
    Ku.close()
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    # not the grid: wrong counts, a permuted element, a permuted numbering
    m = structured_mesh((4, 3, 2))
    with pytest.raises(smearfem.ArgumentError):
        up.set_grid((4, 3, 3))
    up.close()
    bad = m.IEN.copy()
    bad[[3, 5]] = bad[[5, 3]]  # elements 4 and 6 swapped
    mesh = smearfem.Mesh(ctx, m.n_elem, m.NodeList, bad, 3, "Q1", 3, m.ID, 1.0, 0.3)
    with pytest.raises(smearfem.ArgumentError):
        mesh.set_grid((4, 3, 2))
    mesh.close()
    idp = m.ID.copy()
    idp[[0, 1]] = idp[[1, 0]]  # nodes 1 and 2 exchange their dofs
    mesh = smearfem.Mesh(ctx, m.n_elem, m.NodeList, m.IEN, 3, "Q1", 3, idp, 1.0, 0.3)
    with pytest.raises(smearfem.ArgumentError):
        mesh.set_grid((4, 3, 2))
    mesh.close()


def test_grid_declared_non_blocked_numbering(ctx):
    """A declared grid whose numbering is node-interleaved in component 0 only
    (one node's components 1 and 2 swapped): the grid check accepts it, the
    block path does not apply, so the element matrices stored by node column
    are computed again in COO order for the scalar path: the oracle's K."""
    print("[case] grid declared, non-blocked numbering", flush=True)
    m = structured_mesh((4, 3, 2))
    idp = m.ID.copy()
    idp[5, [1, 2]] = idp[5, [2, 1]]
    young = element_young_loguniform(m.n_elem, 0.1, 10.0, seed=5)
    mesh = smearfem.Mesh(ctx, m.n_elem, m.NodeList, m.IEN, 3, "Q1", 3, idp, young, 0.3)
    mesh.set_grid((4, 3, 2))
    K = mesh.assemble()
    assert K.block == 1 and K.grid is None
    g = (K.shape, K.block, K.csc_arrays())
    K.close()
    mesh.close()
    assert_parity(g, oracle_csc(m, 3, young=young, ID=idp))
