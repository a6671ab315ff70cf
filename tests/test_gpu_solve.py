"""GPU Dirichlet elimination + solve against direct solves of the oracle's K.

Bar (every precision mode): displacement within 1e-8 relative
(||u - u_ref||_2 / ||u_ref||_2) of a direct FP64 solve of the reference's
K_ff u = C'(f - K q_d).  Small cases are solved to relative residual 1e-12;
the 64x16x16 cantilever to 1e-10 (its FP64 attainable residual is about
3e-11: scipy's direct solve itself reaches 2.7e-11).  Solutions are bitwise
reproducible run to run.
"""
import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as O
import smearfem
from smearfem.mesh import clamp_x0, element_young_loguniform, face_load, random_face_load, structured_mesh
from smearfem.solve import BoundaryConditions, System, solve_system

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
ERR = 1e-8


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def device_K(ctx, m, young=1.0, nu=0.3, coords=None):
    return smearfem.assemble_system(0, m.NodeList if coords is None else coords, m.IEN, m.ndim, "Q1", m.ndim, m.ID,
                                    young, nu, ctx=ctx, device=True, n_elem=m.n_elem)


def oracle_direct(m, fixed, f, q_d=None, young=1.0, nu=0.3, coords=None):
    K = O.assemble_system(0, m.NodeList if coords is None else coords, m.IEN, m.ndim, "Q1", m.ndim, m.ID, young,
                          nu, n_elem=m.n_elem)
    Kff, rhs = O.eliminate(K, fixed, f, q_d)
    u = O.dense_solve(Kff, rhs) if Kff.shape[0] <= 4000 else spla.spsolve(Kff.tocsc(), rhs)
    return O.recombine(fixed, q_d, u), K


@pytest.mark.parametrize("precision,fmt,pc", [("fp64", "csr", "jacobi"), ("fp64", "bsr", "jacobi"),
                                              ("mixed", "bsr", "jacobi"), ("mixed", "csr", "jacobi"),
                                              ("fp32_matrix", "bsr", "jacobi"), ("fp32_matrix", "csr", "jacobi"),
                                              ("split", "bsr", "jacobi"), ("fp64", "bsr", "block_jacobi"),
                                              ("mixed", "bsr", "block_jacobi"), ("split", "bsr", "block_jacobi")])
def test_cfg1_q4_plane_strain_cantilever(ctx, precision, fmt, pc):
    # unit square 32x32, E=1, nu=0.3 plane strain via (E/(1-nu^2), nu/(1-nu)),
    # left edge clamped, unit traction (+x) on the right edge
    print(f"[case] cfg1 {precision} {fmt} {pc}", flush=True)
    m = structured_mesh((32, 32))
    E, nu = 1.0 / (1 - 0.09), 0.3 / 0.7
    fixed = clamp_x0(m)
    f = face_load(m, 0, 1, 0, 1.0)
    K = device_K(ctx, m, E, nu)
    q, R, st = solve_system(K, f, fixed, precision=precision, fmt=fmt, rtol=1e-12, reactions=True,
                            return_stats=True, preconditioner=pc)
    K.close()
    q_ref, Kref = oracle_direct(m, fixed, f, young=E, nu=nu)
    assert st["converged"] and st["rel_residual"] <= 1e-12
    assert relerr(q, q_ref) <= ERR
    assert np.abs(q[fixed == 1]).max() == 0.0
    # equilibrium: reactions balance the applied load
    assert np.abs(R[fixed == 1].reshape(-1, 2).sum(0) + f.reshape(-1, 2).sum(0)).max() < 1e-9
    assert np.abs(R[fixed == 0]).max() == 0.0


@pytest.mark.parametrize("precision,pc", [("fp64", "jacobi"), ("mixed", "jacobi"), ("fp32_matrix", "jacobi"),
                                          ("split", "jacobi"), ("fp64", "block_jacobi"), ("mixed", "block_jacobi"),
                                          ("split", "block_jacobi")])
def test_hex8_cantilever_per_element_young(ctx, precision, pc):
    print(f"[case] hex8 12x4x4 {precision} {pc}", flush=True)
    m = structured_mesh((12, 4, 4), lengths=(3.0, 1.0, 1.0))
    young = element_young_loguniform(m.n_elem, 0.1, 10.0, seed=7)
    fixed = clamp_x0(m)
    f = random_face_load(m, 2, 1, fixed, seed=3)
    K = device_K(ctx, m, young)
    q, st = solve_system(K, f, fixed, precision=precision, rtol=1e-12, return_stats=True, preconditioner=pc)
    K.close()
    q_ref, _ = oracle_direct(m, fixed, f, young=young)
    assert st["converged"]
    assert relerr(q, q_ref) <= ERR


def test_cfg2_against_direct_fixture(ctx):
    gold = json.loads((GOLD / "cfg2_direct.json").read_text())
    m = structured_mesh(gold["dims"], lengths=gold["lengths"])
    fixed = clamp_x0(m)
    f = face_load(m, 0, 1, 2, -1.0)
    K = device_K(ctx, m)
    assert K.nnz == gold["nnz_K"]
    bc = BoundaryConditions(ctx, K.shape[0], f, fixed)
    S = System(K, bc)
    bc.close()
    assert S.n_free == gold["n_free"] and S.nnz == gold["nnz_Kff"]
    for precision, fmt, pc in (("fp64", "csr", "jacobi"), ("fp64", "bsr", "jacobi"), ("mixed", "bsr", "jacobi"),
                               ("fp32_matrix", "bsr", "jacobi"), ("split", "bsr", "jacobi"),
                               ("split", "bsr", "block_jacobi"), ("mixed", "bsr", "block_jacobi")):
        print(f"[case] cfg2 {precision} {fmt} {pc}", flush=True)
        st = S.solve(precision, fmt, rtol=1e-10, preconditioner=pc)
        q, R = S.solution(reactions=True)
        u = q[fixed == 0]
        assert abs(np.linalg.norm(u) - gold["norm_u"]) <= ERR * gold["norm_u"]
        su = np.array(gold["sample_u"])
        assert np.abs(u[gold["sample_idx"]] - su).max() <= ERR * np.abs(su).max()
        gi = m.node_grid_index()
        tip = np.repeat(gi[0] == 64, 3)
        assert abs(q[tip][2::3].mean() - gold["tip_mean_uz"]) <= ERR * abs(gold["tip_mean_uz"])
        assert np.allclose(R[fixed == 1].reshape(-1, 3).sum(0), gold["reaction_sum"], atol=1e-9)
        print(f"  iterations={st['iterations']} updates={st['fp64_updates']} ms={st['solve_ms']:.1f}", flush=True)
    S.close()
    K.close()


def test_patch_test_prescribed_linear_field(ctx):
    m = structured_mesh((6, 4, 5), lengths=(2.0, 1.0, 1.5))
    rng = np.random.default_rng(5)
    gi = m.node_grid_index()
    interior = (gi[0] > 0) & (gi[0] < 6) & (gi[1] > 0) & (gi[1] < 4) & (gi[2] > 0) & (gi[2] < 5)
    coords = m.NodeList.copy()
    coords[:, interior] += 0.05 * rng.standard_normal((3, interior.sum()))
    A = rng.standard_normal((3, 3)) * 0.1
    u_lin = (A @ coords).T.reshape(-1)
    fixed = np.repeat(~interior, 3).astype(np.uint8)
    K = device_K(ctx, m, coords=coords)
    for precision in ("fp64", "mixed", "fp32_matrix", "split"):
        q = solve_system(K, None, fixed, u_lin, precision=precision, rtol=1e-13)
        assert np.abs(q - u_lin).max() <= 1e-10 * np.abs(u_lin).max()
    K.close()


def test_partial_dof_clamp_every_path(ctx):
    # setboundaryCond-style constraints: only the z dof of the bottom and top
    # nodes prescribed (examples/vector3D.jl:159-168) plus one pinned corner
    m = structured_mesh((4, 4, 4))
    gi = m.node_grid_index()
    fixed = np.zeros(m.n_dof, dtype=np.uint8)
    q_d = np.zeros(m.n_dof)
    bottom, top = gi[2] == 0, gi[2] == 4
    fixed[2::3] = bottom | top
    q_d[2::3] = np.where(top, -0.05, 0.0)
    fixed[0:3] = 1  # pin node 1 to remove rigid motion in x, y
    fixed[3 * 4 + 1] = 1  # and node 5's y (rotation about z)
    K = device_K(ctx, m, 40.0, 0.4)
    bc = BoundaryConditions(ctx, K.shape[0], None, fixed, q_d)
    S = System(K, bc)
    bc.close()
    # nodes with some dofs fixed: padded node blocks (unit diagonal on the
    # fixed dofs) carry every block path; the scalar rows remain available
    assert S.block == 3
    q_ref, _ = oracle_direct(m, fixed, None, q_d, young=40.0, nu=0.4)
    for prec, fmt, pc, rtol in (("fp64", "csr", "jacobi", 1e-12), ("fp64", "bsr", "jacobi", 1e-12),
                                ("fp64", "bsr", "block_jacobi", 1e-12), ("split", "bsr", "block_jacobi", 1e-12),
                                ("mixed", "bsr", "block_jacobi", 1e-11), ("mixed", "csr", "jacobi", 1e-11)):
        st = S.solve(prec, fmt, rtol=rtol, preconditioner=pc)
        assert st["format"] == fmt and st["stop_reason"] == 1
        q = S.solution()
        assert relerr(q, q_ref) <= ERR, (prec, fmt, pc)
        assert np.all(q[fixed == 1] == q_d[fixed == 1])
    # a warm start from a solution converged to the same tolerance needs no iteration
    S.solve("fp64", "bsr", rtol=1e-12)
    st = S.solve("fp64", "bsr", rtol=1e-12, x0_zero=False)
    assert st["iterations"] <= 1
    S.close()
    K.close()


def test_solve_bitwise_reproducible(ctx):
    m = structured_mesh((16, 6, 6), lengths=(3.0, 1.0, 1.0))
    fixed = clamp_x0(m)
    f = random_face_load(m, 2, 1, fixed, seed=1)
    K = device_K(ctx, m, element_young_loguniform(m.n_elem, 0.1, 10.0, seed=2))
    for precision, pc in (("fp64", "jacobi"), ("mixed", "jacobi"), ("fp32_matrix", "jacobi"), ("split", "jacobi"),
                          ("split", "block_jacobi"), ("mixed", "block_jacobi")):
        a, sa = solve_system(K, f, fixed, precision=precision, rtol=1e-10, return_stats=True, preconditioner=pc)
        b, sb = solve_system(K, f, fixed, precision=precision, rtol=1e-10, return_stats=True, preconditioner=pc)
        assert np.array_equal(a, b) and sa["iterations"] == sb["iterations"]
    K.close()


def test_zero_load_and_iteration_cap(ctx):
    m = structured_mesh((6, 3, 3))
    fixed = clamp_x0(m)
    K = device_K(ctx, m)
    q, st = solve_system(K, None, fixed, return_stats=True)
    assert st["converged"] and st["iterations"] == 0 and not q.any()
    f = face_load(m, 0, 1, 2, -1.0)
    with pytest.raises(smearfem.NotConvergedError):
        solve_system(K, f, fixed, rtol=1e-14, max_iter=3)
    K.close()


def test_unattainable_tolerance_stops_on_stagnation(ctx):
    # 1e-15 is below what FP64 arithmetic can reach for a bending cantilever:
    # the solver must stop by itself (stagnation), not run to max_iter
    m = structured_mesh((16, 4, 4), lengths=(4.0, 1.0, 1.0))
    fixed = clamp_x0(m)
    f = face_load(m, 0, 1, 2, -1.0)
    K = device_K(ctx, m)
    bc = BoundaryConditions(ctx, K.shape[0], f, fixed)
    S = System(K, bc)
    bc.close()
    for precision in ("fp64", "mixed"):
        st = S.solve(precision, rtol=1e-15, max_iter=100000, raise_on_fail=False)
        assert st["status"] == smearfem.NotConvergedError.code
        assert st["stop_reason"] == 3 and st["iterations"] < 20000
        assert st["rel_residual"] < 1e-9  # still the best FP64-accurate answer
        # it stopped at the rounding floor eps*|| |rhs| + |K||x| || / ||rhs||
        assert 0 < st["floor_residual"] < 1e-10 and st["rel_residual"] <= 20 * st["floor_residual"]
    S.close()
    K.close()
