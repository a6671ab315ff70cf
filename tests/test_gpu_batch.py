"""Batched assemble-and-solve (one persistent launch) against the oracle.

Bar: every problem's displacement within 1e-8 relative of a direct FP64 solve
of the oracle's K for that problem alone (rtol 1e-12), in every precision.
Invariance: a problem's solution is bitwise the same alone, inside a batch,
for any persistent grid size and for any split of the solve over launches
(same item_slices).
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as O
import smearfem
from smearfem.batch import Batch, make_problem, uneven_hex8_problems
from smearfem.mesh import clamp_x0, face_load, random_face_load, structured_mesh

pytestmark = pytest.mark.gpu
ERR = 1e-8


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def oracle_solution(p):
    m = p.mesh
    # Ke is linear in E and nu is constant within each generated problem
    assert np.all(p.nu == p.nu[0])
    K = O.assemble_system(0, m.NodeList, m.IEN, m.ndim, "Q1", m.ndim, m.ID, p.young, float(p.nu[0]),
                          n_elem=m.n_elem)
    Kff, rhs = O.eliminate(K, p.fixed, p.f, p.q_d)
    u = O.dense_solve(Kff, rhs) if Kff.shape[0] <= 3000 else spla.spsolve(Kff.tocsc(), rhs)
    return O.recombine(p.fixed, p.q_d, u)


def small_hex8_batch(n, seed):
    return uneven_hex8_problems(n, seed, lo=2, hi=7)


@pytest.mark.parametrize("precision,pc", [("fp64", "jacobi"), ("mixed", "jacobi"), ("fp32_matrix", "jacobi"),
                                          ("split", "jacobi"), ("fp64", "block_jacobi"), ("mixed", "block_jacobi"),
                                          ("split", "block_jacobi")])
def test_batch_hex8_against_oracle(ctx, precision, pc):
    print(f"[case] batch hex8 x10 {precision} {pc}", flush=True)
    probs = small_hex8_batch(10, seed=11)
    B = Batch(probs, ctx)
    st, per = B.solve(precision, rtol=1e-12, item_slices=2, preconditioner=pc)
    qs = B.solution()
    B.close()
    assert st["not_converged"] == 0 and st["launches"] == 1
    for p, q, s in zip(probs, qs, per):
        assert s["stop_reason"] == 1 and s["rel_residual"] <= 1e-12
        assert relerr(q, oracle_solution(p)) <= ERR
        assert np.abs(q[p.fixed == 1]).max() == 0.0


def test_batch_q4_with_prescribed_values(ctx):
    print("[case] batch q4 x6 fp64/mixed", flush=True)
    rng = np.random.default_rng(3)
    probs = []
    for k in range(6):
        dims = (int(rng.integers(3, 12)), int(rng.integers(2, 8)))
        m = structured_mesh(dims, lengths=(dims[0] / 4, dims[1] / 4))
        fixed = clamp_x0(m)
        q_d = np.where(fixed == 1, 0.01 * rng.standard_normal(m.n_dof), 0.0)
        f = random_face_load(m, 0, 1, fixed, seed=k)
        young = 10.0 ** rng.uniform(-1, 1, m.n_elem)
        probs.append(make_problem(m, young, rng.uniform(0.2, 0.4), f, fixed, q_d))
    B = Batch(probs, ctx)
    # 1e-11: a 10x2 strip here cannot reach 1e-12 in FP64 (its floor is about 4e-12)
    for precision, pc in (("fp64", "jacobi"), ("mixed", "jacobi"), ("split", "block_jacobi")):
        st, per = B.solve(precision, rtol=1e-11, preconditioner=pc)
        for p, q in zip(probs, B.solution()):
            assert relerr(q, oracle_solution(p)) <= ERR
            qd = p.q_d if p.q_d is not None else np.zeros(p.mesh.n_dof)
        assert np.array_equal(q[p.fixed == 1], qd[p.fixed == 1])
    # below the floor: the solve ends on stagnation, not at max_iter
    st, per = B.solve("fp64", rtol=1e-13, raise_on_fail=False)
    assert st["status"] == smearfem.NotConvergedError.code
    assert all(s["stop_reason"] in (1, 3) for s in per) and max(s["iterations"] for s in per) < 5000
    assert all(s["rel_residual"] <= max(1e-13, 20 * s["floor_residual"]) for s in per)
    B.close()


@pytest.mark.parametrize("precision,pc", [("mixed", "jacobi"), ("split", "block_jacobi")])
def test_batch_problem_bitwise_alone_and_in_mix(ctx, precision, pc):
    print(f"[case] batch invariance {precision} {pc}", flush=True)
    probs = small_hex8_batch(7, seed=5)
    target = probs[3]
    A = Batch([target], ctx)
    A.solve(precision, rtol=1e-11, item_slices=1, preconditioner=pc)
    q_alone = A.solution()[0]
    A.close()
    B = Batch(probs, ctx)
    for bpc, check_every in ((1, 0), (4, 0), (2, 7)):
        st, per = B.solve(precision, rtol=1e-11, item_slices=1, blocks_per_cu=bpc, check_every=check_every,
                          preconditioner=pc)
        q_mix = B.solution()[3]
        assert np.array_equal(q_mix, q_alone), (bpc, check_every)
        if check_every:
            assert st["launches"] > 1
    B.close()


def _context_with_env(monkeypatch, **env):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    c = smearfem.Context(0)
    for k in env:
        monkeypatch.delenv(k)
    return c


def test_barrier_forms_bitwise(ctx, monkeypatch):
    """The per-XCD grouped grid barrier (default) and the per-block form give
    bitwise the same solutions; the default launch uses the grouped form."""
    print("[case] batch barrier forms", flush=True)
    probs = small_hex8_batch(5, seed=8)
    B = Batch(probs, ctx)
    st, _ = B.solve("fp64", rtol=1e-11, item_slices=1, blocks_per_cu=2)
    q_grouped = B.solution()
    B.close()
    assert st["barrier_grouped"] == 1
    c = _context_with_env(monkeypatch, SFEM_BARRIER_GROUPED="0")
    B = Batch(probs, c)
    st, _ = B.solve("fp64", rtol=1e-11, item_slices=1, blocks_per_cu=2)
    q_block = B.solution()
    B.close()
    c.close()
    assert st["barrier_grouped"] == 0
    for a, b in zip(q_grouped, q_block):
        assert np.array_equal(a, b)


def test_barrier_wait_limit_ends_launch_with_error(monkeypatch):
    """A grid-barrier wait longer than the limit aborts the launch (every block
    returns at its next barrier) and the solve reports a device error.  The
    limit is forced to 0 from the host; no wait is unbounded."""
    print("[case] batch barrier wait limit", flush=True)
    c = _context_with_env(monkeypatch, SFEM_BARRIER_TIMEOUT_US="0")
    B = Batch(small_hex8_batch(3, seed=2), c)
    with pytest.raises(smearfem.DeviceError):
        B.solve("fp64", rtol=1e-10, item_slices=1)
    B.close()
    c.close()


def test_batch_iteration_cap_and_errors(ctx):
    probs = small_hex8_batch(3, seed=9)
    B = Batch(probs, ctx)
    with pytest.raises(smearfem.NotConvergedError):
        B.solve("fp64", rtol=1e-12, max_iter=3)
    st, per = B.solve("fp64", rtol=1e-12, max_iter=3, raise_on_fail=False)
    assert all(s["stop_reason"] == 4 and s["iterations"] == 3 for s in per)
    assert all(s["rel_residual"] > 1e-6 for s in per)
    B.close()
    # an IEN entry outside its problem's nodes
    bad = structured_mesh((2, 2, 2))
    bad.IEN[1, 3] = bad.n_nodes + 1
    with pytest.raises(smearfem.BoundsError):
        Batch([make_problem(bad, 1.0, 0.3, np.zeros(bad.n_dof), clamp_x0(bad))], ctx)


def _partial_clamp_problems():
    """Nodes with only some dofs fixed: a roller plane (u_z = 0 on z = 0) next
    to a clamped face, a single fixed dof on a loaded node, and a plain
    cantilever in between."""
    out = []
    m = structured_mesh((3, 3, 3))
    fixed = clamp_x0(m)
    fixed[3 * 21 + 2] = 1  # node 21 is at x = 1: only u_z fixed
    out.append(make_problem(m, 1.0, 0.3, face_load(m, 0, 1, 2, -1.0), fixed))
    out.append(uneven_hex8_problems(1, 5, lo=3, hi=4)[0])
    m = structured_mesh((4, 2, 3))
    fixed = clamp_x0(m)
    x = m.NodeList
    for a in np.nonzero(x[2] == 0.0)[0]:
        fixed[3 * a + 2] = 1
    q_d = np.zeros(m.n_dof)
    top = np.nonzero(x[2] == x[2].max())[0]
    fixed[3 * top] = 1
    q_d[3 * top] = 0.01  # prescribed u_x on the top face, u_y and u_z free
    out.append(make_problem(m, 2.0, 0.25, face_load(m, 0, 1, 2, -0.5), fixed, q_d))
    return out


@pytest.mark.parametrize("precision,pc", [("fp64", "jacobi"), ("mixed", "jacobi"), ("split", "block_jacobi"),
                                          ("fp64", "block_jacobi")])
def test_batch_partially_fixed_nodes(ctx, precision, pc):
    print(f"[case] batch partial clamp {precision} {pc}", flush=True)
    probs = _partial_clamp_problems()
    B = Batch(probs, ctx)
    st, per = B.solve(precision, rtol=1e-12, item_slices=2, preconditioner=pc)
    qs = B.solution()
    B.close()
    assert st["not_converged"] == 0
    for p, q, s in zip(probs, qs, per):
        assert s["n_free"] == int((p.fixed == 0).sum())
        assert s["stop_reason"] == 1 and s["rel_residual"] <= 1e-12
        assert relerr(q, oracle_solution(p)) <= ERR
        qd = p.q_d if p.q_d is not None else np.zeros(p.mesh.n_dof)
        assert np.array_equal(q[p.fixed == 1], qd[p.fixed == 1])
    # a problem's solution is bitwise the same alone
    A = Batch([probs[2]], ctx)
    A.solve(precision, rtol=1e-12, item_slices=2, preconditioner=pc)
    alone = A.solution()[0]
    A.close()
    assert np.array_equal(alone, qs[2])


def test_autotune_picks_a_measured_candidate(ctx, tmp_path, monkeypatch):
    monkeypatch.setenv("SMEARFEM_TUNE_CACHE", str(tmp_path / "tune.json"))
    probs = small_hex8_batch(6, seed=2)
    B = Batch(probs, ctx)
    t = smearfem.autotune(B, "mixed", item_slices=(1, 4), blocks_per_cu=(1, 2), probe_iters=10)
    assert not t["cached"] and len(t["trials"]) == 4
    assert (t["item_slices"], t["blocks_per_cu"]) in {(x["item_slices"], x["blocks_per_cu"]) for x in t["trials"]}
    t2 = smearfem.autotune(B, "mixed", item_slices=(1, 4), blocks_per_cu=(1, 2), probe_iters=10)
    assert t2["cached"] and t2["item_slices"] == t["item_slices"]
    st, per = B.solve("mixed", rtol=1e-10, item_slices=t["item_slices"], blocks_per_cu=t["blocks_per_cu"])
    assert st["not_converged"] == 0
    B.close()
