"""The example's boundary machinery and displacement sweep on the GPU against
the oracle (examples/vector3D.jl:10-338, src/PostProcess.jl:30-46).

Bars: face penalty b pattern and values bitwise equal to the oracle's
restatement (same products, same order, no contraction); device meshgrid /
inflate_sphere bitwise equal to the oracle's; K + beta*b entrywise equal to
scipy's sum; each sweep solution within 1e-8 relative of a direct FP64 solve
of the oracle's K_free; a sweep step after set_bc bitwise equal to a fresh
elimination with the same data.
"""
import numpy as np
import pytest

import oracle as O
import smearfem
from smearfem.surface import (PlaneRule, apply_boundary_conditions, bc_planes, csr_add, displacement_sweep,
                              meshgrid, setboundaryCond)
from smearfem.solve import System

pytestmark = pytest.mark.gpu
ERR = 1e-8


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def same_csc(A, B):
    A, B = A.tocsc(), B.tocsc()
    A.sort_indices()
    B.sort_indices()
    return (A.shape == B.shape and np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)
            and np.array_equal(A.data, B.data))


@pytest.mark.parametrize("ne", [2, 5])
def test_surface_penalty_bitwise_vs_oracle(ctx, ne):
    print(f"[case] surface penalty ne={ne}", flush=True)
    NL, IEN, ID, top, btm = O.meshgrid_ref(0, 1, 0, 1, 0, 1, ne, 3, with_faces=True)
    cyl = O.inflate_sphere_ref(NL, 0, 1, 0, 1)
    want = O.apply_boundary_conditions(ne, cyl, IEN, top, btm, 3, "Q1", ID)
    got = apply_boundary_conditions(ne, cyl, IEN, top, btm, 3, "Q1", ID, ctx=ctx)
    assert same_csc(got, want)


def test_surface_penalty_rectangle_closed_form(ctx):
    print("[case] surface penalty rectangle", flush=True)
    hx, hy = 0.3, 0.7
    pts = []
    for z in (0.0, 1.0):
        pts += [(0.3, -0.2, z), (0.3 + hx, -0.2, z), (0.3 + hx, -0.2 + hy, z), (0.3, -0.2 + hy, z)]
    NL = np.array(pts).T
    ID = np.array([[3 * m + l + 1 for l in range(3)] for m in range(8)], dtype=np.int64)
    b = apply_boundary_conditions(1, NL, None, [[5, 6, 7, 8]], [[1, 2, 3, 4]], 3, "Q1", ID, ctx=ctx).toarray()
    mass = hx * hy * np.array([[4, 2, 1, 2], [2, 4, 2, 1], [1, 2, 4, 2], [2, 1, 2, 4]]) / 36.0
    for face in (0, 1):
        for i in range(3):
            blk = b[np.ix_(12 * face + 3 * np.arange(4) + i, 12 * face + 3 * np.arange(4) + i)]
            assert np.allclose(blk, mass, rtol=1e-14, atol=0)


def test_surface_penalty_errors(ctx):
    print("[case] surface penalty errors", flush=True)
    NL, IEN, ID, top, btm = O.meshgrid_ref(0, 1, 0, 1, 0, 1, 3, 3, with_faces=True)
    bad = top.copy()
    bad[4, 2] = NL.shape[1] + 7  # face 5's third node outside NodeList
    with pytest.raises(smearfem.BoundsError, match="face 5"):
        apply_boundary_conditions(3, NL, IEN, bad, btm, 3, "Q1", ID, ctx=ctx)
    badid = ID[:-3]  # ID rows end before the last top nodes
    with pytest.raises(smearfem.BoundsError):
        apply_boundary_conditions(3, NL, IEN, top, btm, 3, "Q1", badid, ctx=ctx)
    with pytest.raises(smearfem.DimensionMismatch):
        apply_boundary_conditions(3, NL, IEN, top, btm, 3, "Q1", ID, nDof=2, ctx=ctx)
    NL2, IEN2, ID2, top2, btm2 = O.meshgrid_ref(0, 1, 0, 1, 0, 0, 3, 2, with_faces=True)
    with pytest.raises(smearfem.DimensionMismatch):
        apply_boundary_conditions(3, NL2, IEN2, top2, btm2, 2, "Q1", ID2, nDof=2, ctx=ctx)


@pytest.mark.parametrize("ne,inflate", [(4, False), (4, True), (7, True)])
def test_device_meshgrid_vs_oracle(ctx, ne, inflate):
    print(f"[case] meshgrid ne={ne} inflate={inflate}", flush=True)
    NL, IEN, ID, top, btm = O.meshgrid_ref(0, 1, 0, 1, 0, 1, ne, 3, with_faces=True)
    if inflate:
        NL = O.inflate_sphere_ref(NL, 0, 1, 0, 1)
    mesh, faces = meshgrid(0, 1, 0, 1, 0, 1, ne, 3, inflate=inflate, ctx=ctx)
    nodes, ien, idm = mesh.export()
    assert np.array_equal(nodes, NL) and np.array_equal(ien, IEN) and np.array_equal(idm, ID)
    fn, fien, fid = faces.export()
    assert np.array_equal(fien, np.vstack([btm, top])) and np.array_equal(fn, NL)
    mesh.close()
    faces.close()
    # 2D and non-cubic shapes follow the same numbering
    NL2, IEN2, ID2 = O.meshgrid_ref(-1, 2, 0.5, 1.5, 0, 0, 6, 2)
    m2, f2 = meshgrid(-1, 2, 0.5, 1.5, 0, 0, 6, 2, ctx=ctx)
    n2, i2, d2 = m2.export()
    assert f2 is None and np.array_equal(n2, NL2) and np.array_equal(i2, IEN2) and np.array_equal(d2, ID2)
    m2.close()
    m3, _ = meshgrid(0, 3, 0, 1, 0, 2, (6, 2, 4), 3, faces=False, ctx=ctx)
    n3, i3, _ = m3.export()
    host = smearfem.structured_mesh((6, 2, 4), lengths=(3.0, 1.0, 2.0))
    assert np.array_equal(i3, host.IEN) and np.allclose(n3, host.NodeList, rtol=0, atol=1e-15)
    m3.close()


def test_csr_add_vs_scipy(ctx):
    print("[case] K + beta b", flush=True)
    ne = 4
    mesh, faces = meshgrid(0, 1, 0, 1, 0, 1, ne, 3, inflate=True, Young=40.0, nu=0.4, ctx=ctx)
    K, b = mesh.assemble(), faces.assemble()
    Kbar = csr_add(K, 100.0, b)
    ks, bs_, kb = K.to_scipy(), b.to_scipy(), Kbar.to_scipy()
    want = (ks + 100.0 * bs_).tocsc()
    diff = (kb - want)
    assert diff.count_nonzero() == 0
    assert kb.nnz == ks.nnz  # b's couplings are a subset of K's on this mesh
    with pytest.raises(smearfem.DimensionMismatch):
        m2, f2 = meshgrid(0, 1, 0, 1, 0, 1, 3, 3, ctx=ctx)
        b2 = f2.assemble()
        csr_add(K, 1.0, b2)
    for h in (K, b, Kbar, b2):
        h.close()
    for m in (mesh, faces, m2, f2):
        m.close()


def oracle_sweep(ne, deltas, young=40.0, nu=0.4, beta=100.0):
    NL, IEN, ID, top, btm = O.meshgrid_ref(0, 1, 0, 1, 0, 1, ne, 3, with_faces=True)
    cyl = O.inflate_sphere_ref(NL, 0, 1, 0, 1)
    K = O.assemble_system(ne, cyl, IEN, 3, "Q1", 3, ID, young, nu)
    b = O.apply_boundary_conditions(ne, cyl, IEN, top, btm, 3, "Q1", ID)
    Kbar = (K + beta * b).tocsc()
    out = []
    for d in deltas:
        q_d, fixed = O.setboundaryCond_ref(cyl, ne, 3, d)
        Kff, rhs = O.eliminate(Kbar, fixed, None, q_d)
        out.append(O.recombine(fixed, q_d, O.dense_solve(Kff, rhs)))
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_displacement_sweep_vs_oracle(ctx, precision):
    print(f"[case] sweep ne=4 {precision}", flush=True)
    deltas = [0.001, 0.011, 0.021, 0.2]
    res = displacement_sweep(4, deltas, precision=precision, rtol=1e-12 if precision == "fp64" else 1e-11,
                             ctx=ctx)
    want = oracle_sweep(4, deltas)
    for q, w, st in zip(res.q, want, res.stats):
        assert st["stop_reason"] == 1
        assert relerr(q, w) <= ERR
    # q is linear in d
    assert relerr(res.q[3] / 0.2, res.q[1] / 0.011) <= 1e-8
    if precision == "fp64":
        tog = displacement_sweep(4, deltas, rtol=1e-12, together=True, ctx=ctx)
        for q, w, st in zip(tog.q, want, tog.stats):
            assert st["stop_reason"] == 1 and relerr(q, w) <= ERR


def test_displacement_sweep_multigrid_vs_oracle(ctx):
    """The sweep with the multigrid hierarchy of the 8^3 cylinder grid (K_bar
    includes the surface penalty; the z-planes fix only u_z)."""
    print("[case] sweep ne=8 multigrid", flush=True)
    deltas = [0.001, 0.011, 0.3]
    res = displacement_sweep(8, deltas, rtol=1e-12, preconditioner="multigrid", ctx=ctx)
    want = oracle_sweep(8, deltas)
    for q, w, st in zip(res.q, want, res.stats):
        assert st["stop_reason"] == 1 and st["iterations"] <= 40
        assert relerr(q, w) <= ERR


def test_set_bc_matches_fresh_elimination(ctx):
    print("[case] set_bc vs fresh", flush=True)
    mesh, faces = meshgrid(0, 1, 0, 1, 0, 1, 3, 3, inflate=True, Young=40.0, nu=0.4, ctx=ctx)
    K, b = mesh.assemble(), faces.assemble()
    Kbar = csr_add(K, 100.0, b)
    bc1, bc2 = setboundaryCond(mesh, 0.05), setboundaryCond(mesh, 0.3)
    S = System(Kbar, bc1)
    S.solve("fp64", rtol=1e-12)
    S.set_bc(bc2)
    S.solve("fp64", rtol=1e-12, x0_zero=True)
    q_upd = S.solution()
    F = System(Kbar, bc2)
    F.solve("fp64", rtol=1e-12)
    assert np.array_equal(q_upd, F.solution())
    # a warm start from the converged solution of the same data needs no iteration
    S.set_bc(bc2)
    st = S.solve("fp64", rtol=1e-12, x0_zero=False)
    assert st["iterations"] <= 1 and st["stop_reason"] == 1
    # a different fixed set is refused
    bc3 = bc_planes(mesh, mesh.n_nodes * 3, [PlaneRule(1, 0.0, 1, 0.0)])
    with pytest.raises(smearfem.ArgumentError):
        S.set_bc(bc3)
    for h in (S, F, bc1, bc2, bc3, K, b, Kbar, mesh, faces):
        h.close()


def test_bc_planes_rules(ctx):
    print("[case] bc planes", flush=True)
    mesh, _ = meshgrid(0, 2, 0, 1, 0, 1, (4, 2, 2), 3, faces=False, ctx=ctx)
    n = mesh.n_nodes * 3
    # clamp x = 0 (three rules), then a conflicting later rule on the same dof is ignored
    rules = [PlaneRule(1, 0.0, d, 0.0) for d in (1, 2, 3)] + [PlaneRule(1, 0.0, 3, 5.0), PlaneRule(1, 2.0, 1, 0.25)]
    bc = bc_planes(mesh, n, rules)
    K = mesh.assemble()
    S = System(K, bc)
    # x = 0 nodes: all dofs fixed; x = 2 nodes: u_x fixed
    nodes, _, _ = mesh.export()
    n_x0 = np.sum(nodes[0] == 0.0)
    n_x2 = np.sum(nodes[0] == 2.0)
    assert S.n_free == n - 3 * n_x0 - n_x2
    S.solve("fp64", rtol=1e-12)
    q = S.solution()
    assert np.all(q[0::3][nodes[0] == 2.0] == 0.25)
    assert np.all(q.reshape(-1, 3)[nodes[0] == 0.0] == 0.0)
    with pytest.raises(smearfem.ArgumentError):
        bc_planes(mesh, n, [PlaneRule(4, 0.0, 1, 0.0)])
    for h in (S, bc, K, mesh):
        h.close()


@pytest.mark.parametrize("k", [1, 3, 5, 50])
def test_solve_many_vs_oracle(ctx, k):
    """k displacements of the sweep at once; every column within 1e-8 of the
    oracle's direct solve, each column's rhs the one set_bc / eliminate form."""
    print(f"[case] solve_many k={k}", flush=True)
    ne = 4
    deltas = [0.001 + 0.01 * i for i in range(k)]
    mesh, faces = meshgrid(0, 1, 0, 1, 0, 1, ne, 3, inflate=True, Young=40.0, nu=0.4, ctx=ctx)
    K, b = mesh.assemble(), faces.assemble()
    Kbar = csr_add(K, 100.0, b)
    bc = setboundaryCond(mesh, deltas[0])
    S = System(Kbar, bc)
    nodes, _, _ = mesh.export()
    qd = np.zeros((S.n, k))
    top = nodes[2] == 1.0
    for j, d in enumerate(deltas):
        qd[2::3, j][top] = -d
    q, per, ms = S.solve_many(q_d=qd, rtol=1e-11)
    want = oracle_sweep(ne, deltas) if k <= 5 else oracle_sweep(ne, deltas[:: max(1, k // 5)])
    cols = range(k) if k <= 5 else range(0, k, max(1, k // 5))
    for j, w in zip(cols, want):
        assert per[j]["stop_reason"] == 1 and per[j]["rel_residual"] <= 1e-11
        assert relerr(q[:, j], w) <= ERR
    # one column alone equals the single-rhs solve to the tolerance
    st = S.solve("fp64", rtol=1e-11)
    assert relerr(q[:, 0], S.solution()) <= 1e-9
    assert abs(per[0]["iterations"] - st["iterations"]) <= 0.1 * st["iterations"] + 32
    for h in (S, bc, K, b, Kbar, mesh, faces):
        h.close()


def test_solve_many_loads_and_errors(ctx):
    print("[case] solve_many loads", flush=True)
    m = smearfem.structured_mesh((6, 3, 3))
    fixed = smearfem.clamp_x0(m)
    K = smearfem.assemble_system(0, m.NodeList, m.IEN, 3, "Q1", 3, m.ID, 1.0, 0.3, n_elem=m.n_elem, ctx=ctx,
                                 device=True)
    rng = np.random.default_rng(0)
    F = rng.standard_normal((m.n_dof, 4))
    S = System(K, smearfem.BoundaryConditions(ctx, m.n_dof, None, fixed))
    q, per, _ = S.solve_many(f=F, rtol=1e-12)
    Ko = O.assemble_system(0, m.NodeList, m.IEN, 3, "Q1", 3, m.ID, 1.0, 0.3, n_elem=m.n_elem)
    for j in range(4):
        Kff, rhs = O.eliminate(Ko, fixed, F[:, j])
        assert relerr(q[:, j], O.recombine(fixed, None, O.dense_solve(Kff, rhs))) <= ERR
    # the scalar-row multi-rhs SpMV (a K assembled with dof-pair keys, block 1)
    Ks = smearfem.assemble_system(0, m.NodeList, m.IEN, 3, "Q1", 3, m.ID, 1.0, 0.3, n_elem=m.n_elem, ctx=ctx,
                                  device=True, force_scalar=True)
    assert Ks.block == 1
    Ss = System(Ks, smearfem.BoundaryConditions(ctx, m.n_dof, None, fixed))
    qs, per_s, _ = Ss.solve_many(f=F, rtol=1e-12)
    assert all(c["stop_reason"] == 1 for c in per_s)
    assert max(relerr(qs[:, j], q[:, j]) for j in range(4)) <= 1e-9
    Ss.close()
    Ks.close()
    with pytest.raises(smearfem.ArgumentError):
        S.solve_many(f=np.zeros((m.n_dof, 65)))
    with pytest.raises(smearfem.NotConvergedError):
        S.solve_many(f=F, rtol=1e-12, max_iter=5)
    S.close()
    K.close()


def test_apply_and_face_traction_loads(ctx):
    """y = K x on the device equals scipy's product; uniform traction loads on
    the bottom and top faces of a box are the closed-form nodal areas times t."""
    print("[case] apply and traction loads", flush=True)
    from smearfem.surface import face_traction_loads

    nx, ny, nz = 4, 3, 2
    L = (2.0, 1.5, 1.0)
    mesh, faces = meshgrid(0, L[0], 0, L[1], 0, L[2], (nx, ny, nz), 3, ctx=ctx)
    K = mesh.assemble()
    x = np.random.default_rng(3).standard_normal(K.shape[0])
    y = K.apply(x)
    assert relerr(y, K.to_scipy() @ x) <= 1e-14
    t = np.array([0.5, -2.0, 3.0])
    f = face_traction_loads(faces, t).reshape(-1, 3)
    nodes, _, _ = mesh.export()
    hx, hy = L[0] / nx, L[1] / ny
    on_face = (nodes[2] == 0.0) | (nodes[2] == L[2])
    wx = np.where((nodes[0] == 0.0) | (nodes[0] == L[0]), 0.5, 1.0)
    wy = np.where((nodes[1] == 0.0) | (nodes[1] == L[1]), 0.5, 1.0)
    area = np.where(on_face, wx * wy * hx * hy, 0.0)
    assert np.allclose(f, area[:, None] * t[None, :], rtol=1e-13, atol=1e-15)
    assert np.allclose(f.sum(axis=0), 2 * L[0] * L[1] * t, rtol=1e-13)
    for h in (K, mesh, faces):
        h.close()
