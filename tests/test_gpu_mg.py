"""Geometric multigrid preconditioner (structured grids) against the oracle.

Coarse operators: every level's stored operator equals the Galerkin product
P'AP of the level above, with P the multilinear interpolation restricted to
free dofs, built here independently with scipy from the oracle's K (1e-12
relative, Frobenius); the coarse fixed flags are those of the coincident fine
dofs.  Solutions: within 1e-8 relative of a direct FP64 solve of the oracle's
K_ff (solved to relative residual 1e-12), in every supported precision, with
prescribed values, partially fixed nodes and 2D Q4 meshes; the iteration
count stays bounded as the mesh is refined; solves are bitwise reproducible.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg  # noqa: F401  (sp.linalg.norm)

import oracle as O
import smearfem
from smearfem.mesh import clamp_x0, element_young_loguniform, random_face_load, structured_mesh
from smearfem.solve import BoundaryConditions, System, solve_system

pytestmark = pytest.mark.gpu
ERR = 1e-8


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def device_K(ctx, m, young, nu=0.3, detect_grid=False):
    # the sort-based K by default: the assembled finest level (FP16 / FP32
    # sliced smoother); the grid-assembled K of the drop-in takes the
    # matrix-free finest level, which the tests below declare explicitly
    return smearfem.assemble_system(0, m.NodeList, m.IEN, m.ndim, "Q1", m.ndim, m.ID, young, nu, ctx=ctx,
                                    device=True, n_elem=m.n_elem, detect_grid=detect_grid)


def oracle_K(m, young, nu=0.3):
    return sp.csr_matrix(O.assemble_system(0, m.NodeList, m.IEN, m.ndim, "Q1", m.ndim, m.ID, young, nu,
                                           n_elem=m.n_elem))


def oracle_direct(K, fixed, f, q_d=None):
    Kff, rhs = O.eliminate(K, fixed, f, q_d)
    return O.recombine(fixed, q_d, O.dense_solve(Kff, rhs))


def roller(m, axis=2):
    """x = 0 clamped, u_axis fixed on the plane x_axis = 0 (partially fixed nodes)."""
    fixed = clamp_x0(m)
    gi = m.node_grid_index()
    fixed[m.ndim * np.nonzero(gi[axis] == 0)[0] + axis] = 1
    return fixed


def interp_1d(e, f):
    """Linear interpolation from the coarse nodes at fine positions
    min(2c, e), c = 0..ceil(e/2), to the e + 1 fine nodes."""
    if f == 1:
        return sp.identity(e + 1, format="csr")
    ec = (e + 1) // 2
    pos = [min(2 * c, e) for c in range(ec + 1)]
    rows, cols, vals = [], [], []
    for i in range(e + 1):
        c = next(k for k in range(ec + 1) if pos[k] >= i)
        if pos[c] == i:
            rows.append(i)
            cols.append(c)
            vals.append(1.0)
        else:  # between pos[c-1] and pos[c]
            t = (i - pos[c - 1]) / (pos[c] - pos[c - 1])
            rows += [i, i]
            cols += [c - 1, c]
            vals += [1.0 - t, t]
    return sp.csr_matrix((vals, (rows, cols)), shape=(e + 1, ec + 1))


def stencil_to_csr(st, dims):
    """[slot, a, b, node] stencil of a structured level -> sparse matrix."""
    ns, bs, _, nn = st.shape
    n = [d + 1 for d in dims]
    nd = len(dims)
    node = np.arange(nn)
    c = [node % n[0], (node // n[0]) % n[1]] + ([node // (n[0] * n[1])] if nd == 3 else [])
    rows, cols, vals = [], [], []
    for s in range(ns):
        off = [s % 3 - 1, (s // 3) % 3 - 1] + ([s // 9 - 1] if nd == 3 else [])
        ok = np.ones(nn, bool)
        for d in range(nd):
            ok &= (c[d] + off[d] >= 0) & (c[d] + off[d] < n[d])
        J = node + off[0] + n[0] * (off[1] + (n[1] * off[2] if nd == 3 else 0))
        for a in range(bs):
            for b in range(bs):
                rows.append(node[ok] * bs + a)
                cols.append(J[ok] * bs + b)
                vals.append(st[s, a, b, ok])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(nn * bs,) * 2)


def galerkin_chain(K, fixed, dims, levels):
    """Expected level operators and fixed flags, level 1 .. levels-1."""
    bs = len(dims)
    free = ~fixed.astype(bool)
    A = sp.diags(free.astype(float)) @ K @ sp.diags(free.astype(float))
    fx = fixed.astype(bool)
    cur = list(dims)
    out = []
    for _ in range(1, levels):
        f = [2 if e >= 2 else 1 for e in cur]
        P1 = [interp_1d(e, fi) for e, fi in zip(cur, f)]
        Pn = P1[0]
        for Pd in P1[1:]:
            Pn = sp.kron(Pd, Pn)
        P = sp.kron(Pn, sp.identity(bs)).tocsr()
        new = [(e + 1) // 2 if fi == 2 else e for e, fi in zip(cur, f)]
        nc = [e + 1 for e in new]
        idx = np.arange(int(np.prod(nc)))
        cc = [idx % nc[0], (idx // nc[0]) % nc[1]] + ([idx // (nc[0] * nc[1])] if bs == 3 else [])
        at = [np.minimum(cc[d] * f[d], cur[d]) for d in range(bs)]  # fine position of each coarse node
        fn = at[0] + (cur[0] + 1) * (at[1] + ((cur[1] + 1) * at[2] if bs == 3 else 0))
        cfx = fx.reshape(-1, bs)[fn].reshape(-1)
        Pm = sp.diags((~fx).astype(float)) @ P @ sp.diags((~cfx).astype(float))
        A = (Pm.T @ A @ Pm).tocsr()
        out.append((A, cfx, tuple(new)))
        fx, cur = cfx, new
    return out


@pytest.mark.parametrize("dims", [(8, 4, 4), (12, 6, 4), (16, 8), (7, 5, 3), (9, 6)])
def test_coarse_operators_are_galerkin_products(ctx, dims):
    print(f"[case] mg galerkin {dims}", flush=True)
    m = structured_mesh(dims)
    young = element_young_loguniform(m.n_elem, 0.1, 10.0, seed=3)
    fixed = roller(m, axis=m.ndim - 1)
    Kd = device_K(ctx, m, young)
    bc = BoundaryConditions(ctx, m.n_dof, np.zeros(m.n_dof), fixed)
    S = System(Kd, bc)
    bc.close()
    info = S.set_grid(dims, coarse_nodes=1)
    assert info["levels"] >= 2 and info["coarse_dofs"] == info["nodes"][-1] * m.ndim
    expect = galerkin_chain(oracle_K(m, young), fixed, dims, info["levels"])
    for lvl, (A, cfx, new) in enumerate(expect, start=1):
        assert info["dims"][lvl] == new
        st, fx = S.mg_stencil(lvl)
        assert np.array_equal(fx.reshape(-1).astype(bool), cfx)
        G = stencil_to_csr(st, new)
        assert abs(G - G.T).max() == 0.0  # symmetrised exactly
        assert sp.linalg.norm(G - A) <= 1e-12 * sp.linalg.norm(A), lvl
    S.close()
    Kd.close()


@pytest.mark.parametrize("precision", ["fp64", "split", "fp32_matrix"])
def test_mg_solve_vs_oracle(ctx, precision):
    print(f"[case] mg solve 16x8x8 {precision}", flush=True)
    m = structured_mesh((16, 8, 8), lengths=(2.0, 1.0, 1.0))
    young = element_young_loguniform(m.n_elem, 0.1, 10.0, seed=5)
    fixed = clamp_x0(m)
    f = random_face_load(m, 2, 1, fixed, seed=5)
    Kd = device_K(ctx, m, young)
    q, st = solve_system(Kd, f, fixed, precision=precision, rtol=1e-12, return_stats=True,
                                        preconditioner="multigrid", grid=(16, 8, 8))
    Kd.close()
    assert st["converged"] == 1 and st["stop_reason"] == 1 and st["rel_residual"] <= 1e-12
    print(f"[mg] {precision}: {st['iterations']} iterations, {st['fp64_updates']} FP64 residuals", flush=True)
    # an FP32 outer operator reaches 1e-12 through FP64 residual restarts
    assert st["iterations"] <= (40 if precision != "fp32_matrix" else 150), st["iterations"]
    assert relerr(q, oracle_direct(oracle_K(m, young), fixed, f)) <= ERR


def test_mg_fp32_smoother_matches_fp16_smoother_solution(ctx):
    """Every choice of smoother values (FP16 scaled or FP32, on the finest
    level and on the levels below it) reaches the oracle's solution; FP16
    values change the preconditioner only."""
    print("[case] mg fp16 vs fp32 smoother", flush=True)
    m = structured_mesh((16, 4, 8), lengths=(2.0, 0.5, 1.0))
    young = element_young_loguniform(m.n_elem, 0.1, 10.0, seed=9)
    fixed = roller(m)
    f = random_face_load(m, 2, 1, fixed, seed=9)
    Kd = device_K(ctx, m, young)
    ref = oracle_direct(oracle_K(m, young), fixed, f)
    its = {}
    for fp16, cfp16 in ((True, True), (True, False), (False, True), (False, False)):
        bc = BoundaryConditions(ctx, m.n_dof, f, fixed)
        S = System(Kd, bc)
        bc.close()
        info = S.set_grid(m.dims, fine_fp16=fp16, coarse_fp16=cfp16)
        assert info["fine_fp16"] == fp16
        assert info["levels"] >= 3  # a smoothed level below the finest
        assert info["coarse_fp16"] == cfp16
        st = S.solve("fp64", rtol=1e-12, preconditioner="multigrid")
        assert st["stop_reason"] == 1
        assert relerr(S.solution(), ref) <= ERR
        its[fp16, cfp16] = st["iterations"]
        S.close()
    Kd.close()
    print(f"[mg] iterations (fine fp16, coarse fp16): {its}", flush=True)
    assert max(its.values()) <= its[False, False] + 3


def test_mg_partial_nodes_prescribed_values_and_set_bc(ctx):
    print("[case] mg roller + q_d + set_bc", flush=True)
    m = structured_mesh((8, 4, 6), lengths=(2.0, 1.0, 1.5))
    young = element_young_loguniform(m.n_elem, 0.1, 10.0, seed=7)
    fixed = roller(m)
    gi = m.node_grid_index()
    top = np.nonzero(gi[2] == m.dims[2])[0]
    fixed[3 * top + 2] = 1
    q_d = np.zeros(m.n_dof)
    q_d[3 * top + 2] = -0.02
    Kd = device_K(ctx, m, young)
    K = oracle_K(m, young)
    bc = BoundaryConditions(ctx, m.n_dof, np.zeros(m.n_dof), fixed, q_d)
    S = System(Kd, bc)
    bc.close()
    S.set_grid(m.dims)
    st = S.solve("fp64", rtol=1e-12, preconditioner="multigrid")
    q = S.solution()
    assert st["stop_reason"] == 1 and st["iterations"] <= 40
    assert relerr(q, oracle_direct(K, fixed, np.zeros(m.n_dof), q_d)) <= ERR
    assert np.array_equal(q[fixed == 1], q_d[fixed == 1])
    # new loads and values, same fixed set: the hierarchy is kept
    f2 = random_face_load(m, 0, 1, fixed, seed=8)
    bc2 = BoundaryConditions(ctx, m.n_dof, f2, fixed, 0.5 * q_d)
    S.set_bc(bc2)
    bc2.close()
    st = S.solve("split", rtol=1e-12, preconditioner="multigrid")
    q2 = S.solution()
    assert st["stop_reason"] == 1
    assert relerr(q2, oracle_direct(K, fixed, f2, 0.5 * q_d)) <= ERR
    S.close()
    Kd.close()


def test_mg_q4_and_bounded_iterations(ctx):
    print("[case] mg q4 + refinement", flush=True)
    its = []
    for ne in (16, 32, 64):
        m = structured_mesh((ne, ne // 2))
        young = element_young_loguniform(m.n_elem, 0.1, 10.0, seed=ne)
        fixed = clamp_x0(m)
        f = random_face_load(m, 1, 1, fixed, seed=ne)
        Kd = device_K(ctx, m, young)
        q, st = solve_system(Kd, f, fixed, rtol=1e-12, return_stats=True,
                                            preconditioner="multigrid", grid=m.dims)
        Kd.close()
        its.append(st["iterations"])
        if ne <= 32:
            assert relerr(q, oracle_direct(oracle_K(m, young), fixed, f)) <= ERR
    assert max(its) <= 40 and its[-1] <= its[0] + 6, its


def test_mg_bitwise_reproducible_and_errors(ctx):
    print("[case] mg reproducible + errors", flush=True)
    m = structured_mesh((8, 4, 4))
    young = element_young_loguniform(m.n_elem, 0.1, 10.0, seed=1)
    fixed = clamp_x0(m)
    f = random_face_load(m, 2, 1, fixed, seed=1)
    Kd = device_K(ctx, m, young)
    bc = BoundaryConditions(ctx, m.n_dof, f, fixed)
    S = System(Kd, bc)
    bc.close()
    with pytest.raises(smearfem.StateError):
        S.solve("fp64", preconditioner="multigrid")
    with pytest.raises(smearfem.ArgumentError):
        S.set_grid((8, 4, 5))  # node count differs
    with pytest.raises(smearfem.ArgumentError):
        S.set_grid((4, 8, 4))  # same node count, K's couplings are not this grid's
    S.set_grid((8, 4, 4))
    with pytest.raises(smearfem.UnsupportedError):
        S.solve("mixed", preconditioner="multigrid")
    S.solve("fp64", rtol=1e-12, preconditioner="multigrid")
    a = S.solution()
    S.solve("fp64", rtol=1e-12, preconditioner="multigrid")
    assert np.array_equal(a, S.solution())
    S.close()
    Kd.close()
    m3 = structured_mesh((1, 1, 1))
    K3 = device_K(ctx, m3, 1.0)
    bc = BoundaryConditions(ctx, m3.n_dof, np.zeros(m3.n_dof), clamp_x0(m3))
    S3 = System(K3, bc)
    bc.close()
    with pytest.raises(smearfem.UnsupportedError):
        S3.set_grid((1, 1, 1))  # nothing to coarsen
    S3.close()
    K3.close()


# ------------------------------------------------ matrix-free finest level --
# Assembled on the grid (sfem_mesh_set_grid), K carries how far every element
# matrix is from a multiple of the first one; the hierarchy then applies the
# finest level from the element scales: the smoother (deviation <= 1e-6) and
# the FP64 outer operator (<= 1e-12).  The deviation is pinned against the
# oracle's element matrices; the operator through solves (with the FP64 outer
# operator matrix-free the solution is the oracle's only if that operator is
# K_ff: the stop tests the assembled K's residual).

def grid_K(ctx, m, young, nu=0.3):
    mesh = smearfem.Mesh(ctx, m.n_elem, m.NodeList, m.IEN, m.ndim, "Q1", m.ndim, m.ID, young, nu)
    mesh.set_grid(m.dims)
    K = mesh.assemble()
    mesh.close()
    return K


def jittered(m, amount, seed):
    """The mesh with interior nodes moved by up to amount x the element size
    (grid topology kept): element matrices no longer multiples of one."""
    rng = np.random.default_rng(seed)
    h = np.array(m.lengths) / np.array(m.dims)
    X = m.NodeList.copy()
    gi = m.node_grid_index()
    inner = np.all([(g > 0) & (g < d) for g, d in zip(gi, m.dims)], axis=0)
    X[:, inner] += amount * h[:, None] * rng.uniform(-1, 1, (m.ndim, int(inner.sum())))
    return type(m)(m.ndim, m.dims, m.lengths, X, m.IEN, m.ID)


def oracle_ke_deviation(m, young, nu=0.3):
    """max over elements and entries of |Ke(e) / s_e - Ke(0)| / max |Ke(0)|,
    s_e = Ke(e)[0, 0] / Ke(0)[0, 0], from the oracle's element matrices."""
    kes = [O.element_stiffness(m.ndim, "Q1", m.ndim, m.NodeList[:, m.IEN[e] - 1], young[e], nu)
           for e in range(m.n_elem)]
    k0 = kes[0]
    return max(np.abs(k / (k[0, 0] / k0[0, 0]) - k0).max() for k in kes) / np.abs(k0).max()


@pytest.mark.parametrize("dims,amount,mode", [((6, 4, 5), 0.0, 2), ((6, 4, 5), 1e-9, 1), ((6, 4, 5), 1e-3, 0),
                                              ((9, 7), 0.0, 2), ((9, 7), 1e-3, 0)])
def test_mg_ke_deviation_matches_oracle(ctx, dims, amount, mode):
    print(f"[case] mg ke deviation {dims} jitter {amount}", flush=True)
    m = structured_mesh(dims, lengths=(1.5, 1.0, 1.25)[: len(dims)])
    if amount:
        m = jittered(m, amount, seed=3)
    young = element_young_loguniform(m.n_elem, 0.1, 10.0, seed=4)
    ref = oracle_ke_deviation(m, young)
    K = grid_K(ctx, m, young)
    S = System(K, BoundaryConditions(ctx, m.n_dof, np.zeros(m.n_dof), clamp_x0(m)))
    info = S.set_grid(m.dims)
    dev = info["ke_deviation"]
    print(f"[mg] deviation device {dev:.3e} oracle {ref:.3e} -> matrix_free {info['matrix_free']}", flush=True)
    if ref > 1e-12:
        assert abs(dev - ref) <= 1e-6 * ref
    else:  # rounding of the coordinates only: both at the level of FP64 rounding
        assert ref <= 1e-13 and dev <= 1e-13
    assert info["matrix_free"] == mode
    S.close()
    K.close()


@pytest.mark.parametrize("dims", [(8, 4, 6), (7, 5, 3), (12, 6)])
def test_mg_matrix_free_solves_match_oracle(ctx, dims):
    print(f"[case] mg matrix-free solves {dims}", flush=True)
    m = structured_mesh(dims, lengths=(2.0, 1.0, 1.5)[: len(dims)])
    young = element_young_loguniform(m.n_elem, 0.1, 10.0, seed=11)
    fixed = roller(m, axis=m.ndim - 1)
    f = random_face_load(m, m.ndim - 1, 1, fixed, seed=11)
    q_d = np.zeros(m.n_dof)
    gi = m.node_grid_index()
    top = np.nonzero(gi[m.ndim - 1] == m.dims[-1])[0]
    fixed[m.ndim * top + m.ndim - 1] = 1  # partially fixed nodes with a prescribed value
    q_d[m.ndim * top + m.ndim - 1] = -0.01
    ref = oracle_direct(oracle_K(m, young), fixed, f, q_d)
    K = grid_K(ctx, m, young)
    its, coarse = {}, {}
    for mode in (2, 1, 0):
        bc = BoundaryConditions(ctx, m.n_dof, f, fixed, q_d)
        S = System(K, bc)
        bc.close()
        info = S.set_grid(m.dims, matrix_free=mode)
        assert info["matrix_free"] == mode
        # the hierarchy starts from the element scales (modes 1, 2) or from K
        coarse[mode] = S.mg_stencil(1)[0]
        st = S.solve("fp64", rtol=1e-12, preconditioner="multigrid")
        assert st["stop_reason"] == 1
        assert relerr(S.solution(), ref) <= ERR
        its[mode] = st["iterations"]
        S.close()
    print(f"[mg] iterations by matrix_free mode {its}", flush=True)
    assert its[1] <= its[0] + 3 and its[2] <= its[0] + 3
    for mode in (1, 2):
        assert np.linalg.norm(coarse[mode] - coarse[0]) <= 1e-12 * np.linalg.norm(coarse[0])
    # the split outer operator reads the sliced values (built on first use by
    # the solve); FP32 coarse operators under the matrix-free finest level
    for precision, cfp16 in (("split", True), ("fp64", False)):
        bc = BoundaryConditions(ctx, m.n_dof, f, fixed, q_d)
        S = System(K, bc)
        bc.close()
        info = S.set_grid(m.dims, coarse_fp16=cfp16)
        assert info["matrix_free"] == 2 and info["coarse_fp16"] == (cfp16 and info["levels"] > 2)
        st = S.solve(precision, rtol=1e-12, preconditioner="multigrid")
        assert st["stop_reason"] == 1
        assert relerr(S.solution(), ref) <= ERR
        S.close()
    K.close()


@pytest.mark.parametrize("dims", [(8, 4, 6), (7, 5, 3), (12, 6), (9, 7)])
def test_first_coarse_level_by_cells(ctx, dims, monkeypatch):
    """Under a matrix-free finest level the first coarse operator is summed
    cell by cell from the element scales: regular cells from the per
    sub-element table, cells with a fixed dof or a single fine element (odd
    element counts) element by element with masked weights.  It equals
    scipy's P'AP (1e-12) and the one-direction products
    (SFEM_MG_GALERKIN_LINES=1) to FP64 rounding, with fixed dofs scattered
    through the grid (single components of a node too)."""
    print(f"[case] mg first level by cells {dims}", flush=True)
    m = structured_mesh(dims)
    young = element_young_loguniform(m.n_elem, 0.1, 10.0, seed=29)
    fixed = roller(m, axis=m.ndim - 1)
    rng = np.random.default_rng(29)
    fixed[rng.choice(m.n_dof, size=max(3, m.n_dof // 40), replace=False)] = 1
    K = grid_K(ctx, m, young)
    ops = {}
    for lines in ("0", "1"):
        monkeypatch.setenv("SFEM_MG_GALERKIN_LINES", lines)
        bc = BoundaryConditions(ctx, m.n_dof, np.zeros(m.n_dof), fixed)
        S = System(K, bc)
        bc.close()
        info = S.set_grid(m.dims, coarse_nodes=1)
        assert info["matrix_free"] == 2
        ops[lines] = S.mg_stencil(1)
        S.close()
    K.close()
    A, cfx, new = galerkin_chain(oracle_K(m, young), fixed, dims, 2)[0]
    st, fx = ops["0"]
    assert np.array_equal(fx.reshape(-1).astype(bool), cfx)
    G = stencil_to_csr(st, new)
    assert abs(G - G.T).max() == 0.0
    assert sp.linalg.norm(G - A) <= 1e-12 * sp.linalg.norm(A)
    assert np.linalg.norm(st - ops["1"][0]) <= 1e-13 * np.linalg.norm(ops["1"][0])


def test_mg_matrix_free_needs_grid_assembly(ctx):
    """K assembled without the grid (sorted keys): no element data, the
    finest level reads K."""
    m = structured_mesh((8, 4, 4))
    young = element_young_loguniform(m.n_elem, 0.1, 10.0, seed=2)
    Kd = device_K(ctx, m, young)
    S = System(Kd, BoundaryConditions(ctx, m.n_dof, np.zeros(m.n_dof), clamp_x0(m)))
    info = S.set_grid(m.dims)
    assert info["matrix_free"] == 0 and info["ke_deviation"] == float("inf")
    S.close()
    Kd.close()


@pytest.mark.parametrize("dims", [(20, 9, 6), (33, 12, 10)])
def test_mg_matrix_free_mfma_matches_valu(ctx, dims, monkeypatch):
    """The matrix-free smoothing steps on the matrix cores (split FP16
    products, tiles of 16 x 4 x 4 nodes: these grids leave partial tiles in
    every direction) against the VALU kernel: both preconditioners reach the
    oracle's solution, in the same number of iterations up to one, and the
    MFMA one is bitwise reproducible."""
    print(f"[case] mg mfma vs valu {dims}", flush=True)
    m = structured_mesh(dims, lengths=(2.0, 1.0, 0.8))
    young = element_young_loguniform(m.n_elem, 0.1, 10.0, seed=17)
    fixed = roller(m, axis=2)
    f = random_face_load(m, 2, 1, fixed, seed=17)
    q_d = np.zeros(m.n_dof)
    gi = m.node_grid_index()
    top = np.nonzero(gi[2] == m.dims[2])[0]
    fixed[3 * top + 2] = 1  # partially fixed nodes with a prescribed value
    q_d[3 * top + 2] = -0.01
    import scipy.sparse.linalg as spla

    Kff, rhs = O.eliminate(oracle_K(m, young), fixed, f, q_d)
    ref = O.recombine(fixed, q_d, spla.spsolve(Kff.tocsc(), rhs))
    K = grid_K(ctx, m, young)
    its, sols = {}, {}
    for flag in ("0", "1", "1"):
        monkeypatch.setenv("SFEM_MF_MFMA", flag)
        bc = BoundaryConditions(ctx, m.n_dof, f, fixed, q_d)
        S = System(K, bc)
        bc.close()
        assert S.set_grid(m.dims)["matrix_free"] == 2
        st = S.solve("fp64", rtol=1e-12, preconditioner="multigrid")
        q = S.solution()
        assert st["stop_reason"] == 1 and relerr(q, ref) <= ERR, flag
        if flag in sols:
            assert np.array_equal(q, sols[flag])  # run to run
        its[flag], sols[flag] = st["iterations"], q
        S.close()
    K.close()
    print(f"[mg] iterations valu {its['0']} mfma {its['1']}", flush=True)
    assert abs(its["1"] - its["0"]) <= 1


@pytest.mark.parametrize("dims", [(40, 20, 14), (96, 50), (23, 9, 7)])
def test_mg_coarse_kernel_forms_same_bits(ctx, dims, monkeypatch):
    """The coarse levels' two kernel forms -- one thread per node on a
    slot-major stencil, a lane group per node on a node-major one -- sum each
    stencil row by the same tree, so solutions and iteration counts are
    bitwise the same whichever form runs on which level; and they meet the
    oracle."""
    print(f"[case] mg coarse forms {dims}", flush=True)
    m = structured_mesh(dims)
    young = element_young_loguniform(m.n_elem, 0.1, 10.0, seed=23)
    fixed = clamp_x0(m)
    f = random_face_load(m, 0, 1, fixed, seed=23)
    q_d = np.zeros(m.n_dof)
    K = device_K(ctx, m, young, detect_grid=True)
    forms = {"threads": "0", "groups": "1000000000", "default": None}
    sols, its = {}, {}
    for name, below in forms.items():
        if below is None:
            monkeypatch.delenv("SFEM_MG_GROUP_BELOW", raising=False)
        else:
            monkeypatch.setenv("SFEM_MG_GROUP_BELOW", below)
        bc = BoundaryConditions(ctx, m.n_dof, f, fixed, q_d)
        S = System(K, bc)
        bc.close()
        S.set_grid(dims)
        st = S.solve("fp64", rtol=1e-11, preconditioner="multigrid")
        sols[name], its[name] = S.solution(), st["iterations"]
        S.close()
    K.close()
    print(f"[mg] coarse forms iterations {its}", flush=True)
    for name in forms:
        assert its[name] == its["threads"] and np.array_equal(sols[name], sols["threads"]), name
    if m.n_dof <= 6000:
        assert relerr(sols["default"], oracle_direct(oracle_K(m, young), fixed, f)) <= ERR
