"""Pins the CPU oracle (oracle/sfem_oracle.c) before it is trusted as the
checker of the GPU path.  Every check comes from outside the oracle's own
formulas: the reference's known answers (test/runtests.jl:15-35), a published
element matrix, sympy exact integrals, invariants the mathematics fixes, a
closed-form pattern count, an independent sparse() restatement and scipy's
direct solver.
"""
import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O
from smearfem.mesh import clamp_x0, face_load, structured_mesh

GOLD = Path(__file__).resolve().parent / "golden" / "element_fixtures.json"


# ---- reference known answers: test/runtests.jl:15-35 ----------------------
def test_runtests_basis_kronecker_delta():
    assert list(O.basis_function(-1)[0]) == [1.0, 0.0]
    assert list(O.basis_function(1)[0]) == [0.0, 1.0]
    corners2 = [(-1, -1), (1, -1), (1, 1), (-1, 1)]
    for k, (x, y) in enumerate(corners2):
        assert list(O.basis_function(x, y)[0]) == [1.0 if i == k else 0.0 for i in range(4)]
    corners3 = [(x, y, z) for z in (-1, 1) for (x, y) in corners2]
    for k, (x, y, z) in enumerate(corners3):
        assert list(O.basis_function(x, y, z)[0]) == [1.0 if i == k else 0.0 for i in range(8)]


def test_runtests_gaussian_quadrature_exact():
    assert O.gaussian_quadrature(-1, 1, 2) == ([-1 / np.sqrt(3), 1 / np.sqrt(3)], [1.0, 1.0])
    assert O.gaussian_quadrature(-1, 1, 3) == ([-np.sqrt(3 / 5), 0.0, np.sqrt(3 / 5)], [5 / 9, 8 / 9, 5 / 9])


def test_quadrature_undefined_order_errors():
    with pytest.raises(O.OracleError) as e:
        O.gaussian_quadrature(-1, 1, 4)
    assert e.value.code == -5


@pytest.mark.parametrize("n,a,b", [(2, -0.3, 1.7), (3, -1.3, 1.3)])
def test_quadrature_integrates_polynomials(n, a, b):
    # n-point Gauss-Legendre is exact to degree 2n-1
    xs, ws = O.gaussian_quadrature(a, b, n)
    for deg in range(2 * n):
        exact = (b ** (deg + 1) - a ** (deg + 1)) / (deg + 1)
        assert abs(sum(w * x ** deg for x, w in zip(xs, ws)) - exact) < 1e-13


def test_quadrature3_middle_node_is_zero_as_in_reference():
    # src/fem.jl:27 writes the middle node as 0, not (b+a)/2, so on an
    # interval not centred at 0 the rule is shifted; the restatement keeps it.
    xs, ws = O.gaussian_quadrature(0.0, 2.0, 3)
    assert xs[1] == 0.0 and xs[0] == -(2.0) / (2 * np.sqrt(5 / 3)) + 1.0
    assert abs(sum(ws) - 2.0) < 1e-15


@pytest.mark.parametrize("args,cls", [((0.3, -0.2), "Q1"), ((0.3, -0.2, 0.7), "Q1"), ((0.3, -0.2), "Q2")])
def test_basis_partition_of_unity_and_gradients(args, cls):
    N, dN = O.basis_function(*args, FunctionClass=cls)
    assert abs(N.sum() - 1.0) < 1e-14
    assert np.abs(dN.sum(axis=0)).max() < 1e-14
    h = 1e-6  # central differences pin each derivative column
    for d in range(len(args)):
        ap = list(args); am = list(args)
        ap[d] += h; am[d] -= h
        fd = (O.basis_function(*ap, FunctionClass=cls)[0] - O.basis_function(*am, FunctionClass=cls)[0]) / (2 * h)
        assert np.abs(fd - dN[:, d]).max() < 1e-8


def test_q2_kronecker_delta_at_its_nodes():
    nodes = [(-1, -1), (1, -1), (1, 1), (-1, 1), (0, -1), (1, 0), (0, 1), (-1, 0), (0, 0)]
    for k, (x, y) in enumerate(nodes):
        N, _ = O.basis_function(x, y, FunctionClass="Q2")
        assert np.allclose(N, np.eye(9)[k], atol=0, rtol=0) or np.abs(N - np.eye(9)[k]).max() < 1e-15


# ---- element stiffness --------------------------------------------------
def test_q4_unit_square_matches_published_topopt_element():
    # Sigmund's 99-line topology optimisation code, plane stress unit square:
    # KE = E/(1-nu^2) * [k1..k8] pattern, first row = k
    nu = 0.3
    Ke = O.element_stiffness(2, "Q1", 2, np.array([[0, 1, 1, 0], [0, 0, 1, 1.0]]), 1.0, nu)
    k = np.array([1 / 2 - nu / 6, 1 / 8 + nu / 8, -1 / 4 - nu / 12, -1 / 8 + 3 * nu / 8, -1 / 4 + nu / 12,
                  -1 / 8 - nu / 8, nu / 6, 1 / 8 - 3 * nu / 8])
    assert np.abs(Ke[0] * (1 - nu ** 2) - k).max() < 1e-15


@pytest.mark.parametrize("fx", json.loads(GOLD.read_text())["fixtures"], ids=lambda f: f["name"])
def test_element_matches_sympy_fixture(fx):
    Ke = O.element_stiffness(fx["ndim"], fx["cls"], fx["ndof"], np.array(fx["coords"]), fx["E"], fx["nu"])
    ref = np.array(fx["Ke"])
    assert np.abs(Ke - ref).max() <= 1e-13 * np.abs(ref).max()


def _distorted(ndim, seed=3):
    rng = np.random.default_rng(seed)
    if ndim == 2:
        c = np.array([[0, 1, 1, 0], [0, 0, 1, 1.0]])
    else:
        c = np.array([[0, 1, 1, 0, 0, 1, 1, 0], [0, 0, 1, 1, 0, 0, 1, 1], [0, 0, 0, 0, 1, 1, 1, 1.0]])
    return c + 0.15 * rng.standard_normal(c.shape)


def _element_volume(coords):
    ndim = coords.shape[0]
    # exact measure by the divergence theorem is overkill: 2x2(x2) Gauss on det J
    # is exact for bilinear / trilinear maps; use a fine midpoint rule instead
    n = 40 if ndim == 2 else 16
    g = (np.arange(n) + 0.5) / n * 2 - 1
    vol = 0.0
    for idx in np.ndindex(*(n,) * ndim):
        p = [g[i] for i in idx]
        _, dN = O.basis_function(*p)
        vol += abs(np.linalg.det(coords @ dN)) * (2 / n) ** ndim
    return vol


@pytest.mark.parametrize("ndim", [2, 3])
def test_linear_field_energy_on_distorted_element(ndim):
    # An isoparametric element reproduces linear displacement fields exactly,
    # so u'Ke u = volume * eps'D eps for u = A x (constant strain); pins B,
    # D, the Voigt order, the Jacobian and its inverse on a non-box element.
    coords = _distorted(ndim)
    E, nu = 2.0, 0.27
    Ke = O.element_stiffness(ndim, "Q1", ndim, coords, E, nu)
    A = np.random.default_rng(7).standard_normal((ndim, ndim))
    u = (A @ coords).T.reshape(-1)  # node-major interleaved
    eps_t = (A + A.T) / 2
    if ndim == 2:
        c = E / (1 - nu ** 2)
        sig = c * np.array([[eps_t[0, 0] + nu * eps_t[1, 1], (1 - nu) * eps_t[0, 1]],
                            [(1 - nu) * eps_t[0, 1], eps_t[1, 1] + nu * eps_t[0, 0]]])
    else:
        lam = E * nu / ((1 + nu) * (1 - 2 * nu))
        mu = E / (2 * (1 + nu))
        sig = lam * np.trace(eps_t) * np.eye(3) + 2 * mu * eps_t
    energy = np.sum(sig * eps_t) * _element_volume(coords)
    assert abs(u @ Ke @ u - energy) <= 2e-3 * abs(energy)  # midpoint volume error only
    # rigid-body modes are in the null space
    x = coords
    modes = [np.tile(np.eye(ndim)[d], x.shape[1]) for d in range(ndim)]
    if ndim == 2:
        modes.append(np.stack([-x[1], x[0]], 1).reshape(-1))
    else:
        modes += [np.stack([np.zeros_like(x[0]), -x[2], x[1]], 1).reshape(-1),
                  np.stack([x[2], np.zeros_like(x[0]), -x[0]], 1).reshape(-1),
                  np.stack([-x[1], x[0], np.zeros_like(x[0])], 1).reshape(-1)]
    for r in modes:
        assert np.abs(Ke @ r).max() < 1e-12 * np.abs(Ke).max()
    w = np.linalg.eigvalsh((Ke + Ke.T) / 2)
    assert (w > -1e-12 * w.max()).all() and (w > 1e-10 * w.max()).sum() == Ke.shape[0] - len(modes)


def test_young_linearity():
    c = _distorted(3)
    K1 = O.element_stiffness(3, "Q1", 3, c, 1.0, 0.3)
    K7 = O.element_stiffness(3, "Q1", 3, c, 7.0, 0.3)
    assert np.abs(K7 - 7 * K1).max() <= 1e-14 * np.abs(K7).max()


# ---- sparse() semantics -------------------------------------------------
def test_sparse_combine_order_and_explicit_zeros():
    I = [2, 1, 2, 2, 3, 1]
    J = [1, 1, 1, 1, 3, 1]
    V = [1e16, 5.0, 1.0, -1e16, 0.0, -5.0]
    m, n, colptr, rowval, nzval = O.sparse_csc(I, J, V)
    ref, dims = O.sparse_ref(I, J, V)
    assert (m, n) == dims == (3, 3)
    got = {(int(rowval[k]), j + 1): nzval[k] for j in range(n) for k in range(colptr[j] - 1, colptr[j + 1] - 1)}
    assert got.keys() == ref.keys()
    for k in ref:
        assert got[k] == ref[k]  # bitwise: fold order (1e16 + 1) - 1e16 = 0, not 1
    assert got[(2, 1)] == 0.0 and got[(1, 1)] == 0.0 and got[(3, 3)] == 0.0  # explicit zeros kept
    assert list(rowval[colptr[0] - 1:colptr[1] - 1]) == [1, 2]  # rows ascend within a column


def test_sparse_rejects_index_below_one():
    with pytest.raises(O.OracleError) as e:
        O.sparse_csc([1, 0], [1, 1], [1.0, 2.0])
    assert e.value.code == -1


def test_sparse_random_against_pure_python():
    rng = np.random.default_rng(11)
    I = rng.integers(1, 20, 500)
    J = rng.integers(1, 17, 500)
    V = rng.standard_normal(500) * 10.0 ** rng.integers(-8, 8, 500)
    m, n, colptr, rowval, nzval = O.sparse_csc(I, J, V)
    ref, dims = O.sparse_ref(I, J, V)
    assert (m, n) == dims
    got = {(int(rowval[k]), j + 1): nzval[k] for j in range(n) for k in range(colptr[j] - 1, colptr[j + 1] - 1)}
    assert got == ref


# ---- meshes and assembly ------------------------------------------------
@pytest.mark.parametrize("ndim,ne", [(2, 3), (3, 2)])
def test_structured_mesh_matches_reference_meshgrid(ndim, ne):
    NL, IEN, ID = O.meshgrid_ref(0, 1, 0, 1, 0, 1, ne, ndim)
    m = structured_mesh([ne] * ndim)
    assert np.array_equal(m.IEN, IEN) and np.array_equal(m.ID, ID)
    assert np.abs(m.NodeList - NL).max() == 0.0


def coupled_pairs(dims):
    p = 1
    for d in dims:
        p *= 3 * (d + 1) - 2
    return p


@pytest.mark.parametrize("dims", [(4, 3), (3, 2, 2), (5, 1, 2)])
def test_assembled_pattern_closed_form(dims):
    m = structured_mesh(dims)
    ndim = len(dims)
    K = O.assemble_system(0, m.NodeList, m.IEN, ndim, "Q1", ndim, m.ID, 1.0, 0.3, n_elem=m.n_elem)
    assert K.shape == (m.n_dof, m.n_dof)
    assert K.nnz == coupled_pairs(dims) * ndim ** 2
    L = O.assemble_system(0, m.NodeList, m.IEN, ndim, "Q1", 1, None, 1.0, 0.3, n_elem=m.n_elem)
    assert L.nnz == coupled_pairs(dims)
    # Laplace: constants are in the null space; elasticity: translations
    assert np.abs(L @ np.ones(m.n_nodes)).max() < 1e-13 * abs(L).max()
    for d in range(ndim):
        t = np.zeros(m.n_dof); t[d::ndim] = 1
        assert np.abs(K @ t).max() < 1e-13 * abs(K).max()


def test_assembled_energy_of_linear_field():
    # sum over elements of the constant-strain energy = box volume * eps'D eps
    m = structured_mesh((3, 2, 2), lengths=(1.5, 1.0, 0.5))
    E, nu = 3.0, 0.25
    K = O.assemble_system(0, m.NodeList, m.IEN, 3, "Q1", 3, m.ID, E, nu, n_elem=m.n_elem)
    A = np.array([[0.1, 0.2, -0.3], [0.05, -0.1, 0.2], [0.3, 0.0, 0.15]])
    u = (A @ m.NodeList).T.reshape(-1)
    eps = (A + A.T) / 2
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    sig = lam * np.trace(eps) * np.eye(3) + 2 * mu * eps
    assert abs(u @ (K @ u) - np.sum(sig * eps) * 1.5 * 0.5) < 1e-12


def test_assembly_error_codes():
    m = structured_mesh((2, 2))
    bad = m.IEN.copy(); bad[2, 1] = m.n_nodes + 1
    with pytest.raises(O.OracleError) as e:
        O.assemble_coo(4, m.NodeList, bad, 2, "Q1", 2, m.ID)
    assert e.value.code == -2 and e.value.where == 2
    coords = m.NodeList.copy(); coords[:, m.IEN[1] - 1] = 0.5  # collapse element 2
    with pytest.raises(O.OracleError) as e:
        O.assemble_coo(4, coords, m.IEN, 2, "Q1", 2, m.ID)
    assert e.value.code == -4 and e.value.where == 1
    with pytest.raises(O.OracleError) as e:
        O.assemble_coo(4, m.NodeList, m.IEN, 2, "Q1", 3, np.hstack([m.ID, m.ID[:, :1]]))
    assert e.value.code == -3
    with pytest.raises(O.OracleError) as e:
        O.assemble_coo(4, m.NodeList, m.IEN, 2, "Q3", 1)
    assert e.value.code == -5
    ID0 = m.ID.copy(); ID0[4, 1] = 0
    E_, J_, V_ = O.assemble_coo(4, m.NodeList, m.IEN, 2, "Q1", 2, ID0)
    with pytest.raises(O.OracleError) as e:
        O.sparse_csc(E_, J_, V_)
    assert e.value.code == -1


# ---- elimination and solve ----------------------------------------------
def test_elimination_and_dense_solve_match_scipy_direct():
    m = structured_mesh((6, 2, 2), lengths=(3.0, 1.0, 1.0))
    K = O.assemble_system(0, m.NodeList, m.IEN, 3, "Q1", 3, m.ID, 1.0, 0.3, n_elem=m.n_elem)
    fixed = clamp_x0(m)
    f = face_load(m, 0, 1, 2, -1.0)
    Kff, rhs = O.eliminate(K, fixed, f)
    free = fixed == 0
    assert abs(Kff - K[free][:, free]).max() == 0.0  # C'KC picks the submatrix exactly
    assert np.array_equal(rhs, f[free])
    x = O.dense_solve(Kff, rhs)
    xs = spla.spsolve(sp.csc_matrix(K[free][:, free]), f[free])
    assert np.linalg.norm(x - xs) <= 1e-12 * np.linalg.norm(xs)
    xp, it, rr = O.pcg(Kff, rhs, rtol=1e-12)
    assert rr <= 1e-12 and np.linalg.norm(xp - xs) <= 1e-8 * np.linalg.norm(xs)
    q = O.recombine(fixed, None, x)
    R = K @ q - f  # equilibrium: reactions balance the applied load
    assert abs(R[fixed == 1].reshape(-1, 3).sum(0) + f.reshape(-1, 3).sum(0)).max() < 1e-10


def test_patch_test_prescribed_linear_field():
    # every boundary dof prescribed to a linear field: interior solution exact
    m = structured_mesh((4, 3, 3), lengths=(2.0, 1.0, 1.5))
    rng = np.random.default_rng(5)
    coords = m.NodeList.copy()
    gi = m.node_grid_index()
    interior = (gi[0] > 0) & (gi[0] < 4) & (gi[1] > 0) & (gi[1] < 3) & (gi[2] > 0) & (gi[2] < 3)
    coords[:, interior] += 0.05 * rng.standard_normal((3, interior.sum()))  # distorted interior
    K = O.assemble_system(0, coords, m.IEN, 3, "Q1", 3, m.ID, 1.0, 0.3, n_elem=m.n_elem)
    A = rng.standard_normal((3, 3)) * 0.1
    u_lin = (A @ coords).T.reshape(-1)
    fixed = np.repeat(~interior, 3).astype(np.uint8)
    Kff, rhs = O.eliminate(K, fixed, None, np.where(fixed == 1, u_lin, 0.0))
    x = O.dense_solve(Kff, rhs)
    assert np.abs(x - u_lin[fixed == 0]).max() < 1e-12


def test_threaded_pcg_matches_serial():
    # the all-thread CPU baseline runs the same algorithm as the serial oracle
    m = structured_mesh((12, 4, 4))
    fixed = clamp_x0(m)
    f = face_load(m, 0, 1, 2, -1.0)
    K = O.assemble_system(0, m.NodeList, m.IEN, 3, "Q1", 3, m.ID, 1.0, 0.3, n_elem=m.n_elem)
    Kff, rhs = O.eliminate(K, fixed, f)
    x1, i1, r1 = O.pcg(Kff, rhs, rtol=1e-11)
    x4, i4, r4 = O.pcg_threads(Kff, rhs, 4, rtol=1e-11)
    assert abs(i1 - i4) <= 2 and r4 <= 1e-11
    assert np.linalg.norm(x1 - x4) <= 1e-9 * np.linalg.norm(x1)
