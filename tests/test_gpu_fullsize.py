"""The workloads at their full sizes, checked by size-independent properties
and by direct-solve fixtures.

* 64x16x16 Hex8 cantilever: K entrywise against the oracle's assembly
  (pattern bit-exact, values within 1e-12 normwise), on the grid path the
  drop-in takes and on the sort-based path.
* 256x64x64 Hex8 cantilever (3,244,800 free dofs) on the measured path
  (grid assembly, elimination, matrix-free multigrid, FP64 outer CG):
  closed-form nnz of K and K_ff, K times each rigid translation = 0,
  an FP64 residual computed on the host with scipy from the exported K
  (independent of every device kernel after assembly), the solution against
  an FP64 block-Jacobi solve carried on to 1e-12 and against the mixed form,
  reactions balancing the loads, and the one-shot drop-in reaching the same
  K bitwise.
* The 512-problem batch (seed 4, dims U{4..32}^3): the three largest
  problems against SuperLU solves of the oracle's K (fixtures from
  tests/golden/make_large_direct.py), every problem against an FP64 solve of
  the whole batch to 1e-13, and problems bitwise the same alone and in the
  batch.
* Served requests of 1e5-6e5 dofs against SuperLU fixtures, by error.

Bar: displacement within 1e-8 relative error.  The fixtures keep ||u||, 128
sampled entries and 16 projections g_k . u on seeded N(0, 1) vectors: for an
error e, g_k . e has standard deviation ||e||, so |g_k . (u - u_ref)| <=
4e-8 ||u_ref|| for all 16 fails with high probability once ||e|| > 1e-8 ||u_ref||.
"""
import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
import smearfem
from smearfem.batch import Batch, uneven_hex8_problems
from smearfem.mesh import clamp_x0, element_young_loguniform, face_load, random_face_load, structured_mesh
from smearfem.serve import Server, random_request
from smearfem.solve import BoundaryConditions, System

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
ERR = 1e-8

CFG3_DIMS = (256, 64, 64)
CFG3_NNZ_K = 257_800_329      # prod_d (3 (n_d + 1) - 2) * 9
CFG3_NNZ_KFF = 256_794_606
CFG3_NFREE = 3_244_800


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def closed_form_nnz(dims, ndof):
    return int(np.prod([3 * (d + 1) - 2 for d in dims])) * ndof * ndof


def projections(u, seed, k):
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.array([rng.standard_normal(u.size) @ u for _ in range(k)])


def check_against_fixture(u, fx, n_proj):
    """u (free dofs) against a direct-solve summary."""
    assert u.size == fx["n_free"]
    nu = fx["norm_u"]
    assert abs(np.linalg.norm(u) - nu) <= ERR * nu
    su = np.array(fx["sample_u"])
    assert np.abs(u[fx["sample_idx"]] - su).max() <= ERR * nu
    dp = np.abs(projections(u, fx["proj_seed"], n_proj) - np.array(fx["proj"]))
    assert dp.max() <= 4 * ERR * nu, (dp.max() / nu)
    return float(dp.max() / nu)


# ----------------------------------------------------------- 64x16x16 --
@pytest.mark.parametrize("detect_grid", [True, False], ids=["grid", "sorted"])
def test_cfg2_K_entrywise_vs_oracle(ctx, detect_grid):
    print(f"[case] cfg2 K entrywise detect_grid={detect_grid}", flush=True)
    m = structured_mesh((64, 16, 16), lengths=(4.0, 1.0, 1.0))
    K = smearfem.assemble_system(0, m.NodeList, m.IEN, 3, "Q1", 3, m.ID, 1.0, 0.3, ctx=ctx, device=True,
                                 n_elem=m.n_elem, detect_grid=detect_grid)
    assert (K.grid == (64, 16, 16)) if detect_grid else K.grid is None
    cp, rv, nz = K.csc_arrays()
    K.close()
    E, J, V = O.assemble_coo(m.n_elem, m.NodeList, m.IEN, 3, "Q1", 3, m.ID, 1.0, 0.3)
    _, _, ocp, orv, onz = O.sparse_csc(E, J, V)
    assert cp.size - 1 == 56355 and nz.size == closed_form_nnz((64, 16, 16), 3) == onz.size
    assert np.array_equal(cp, ocp) and np.array_equal(rv, orv)
    assert np.abs(nz - onz).max() <= 1e-12 * np.abs(onz).max()


# ---------------------------------------------------------- 256x64x64 --
@pytest.fixture(scope="module")
def cfg3(ctx):
    """The bench's headline inputs and path: grid assembly + elimination."""
    print("[case] cfg3 256x64x64 inputs", flush=True)
    m = structured_mesh(CFG3_DIMS, lengths=(4.0, 1.0, 1.0))
    young = element_young_loguniform(m.n_elem, 0.1, 10.0, seed=1234)
    fixed = clamp_x0(m)
    f = random_face_load(m, 2, 1, fixed, seed=5678)
    mesh = smearfem.Mesh(ctx, m.n_elem, m.NodeList, m.IEN, 3, "Q1", 3, m.ID, young, 0.3)
    mesh.set_grid(CFG3_DIMS)
    K = mesh.assemble()
    bc = BoundaryConditions(ctx, m.n_dof, f, fixed)
    S = System(K, bc)
    bc.close()
    state = {"m": m, "young": young, "fixed": fixed, "f": f, "mesh": mesh, "K": K, "S": S}
    yield state
    S.close()
    K.close()
    mesh.close()


def test_cfg3_counts(cfg3):
    K, S = cfg3["K"], cfg3["S"]
    assert closed_form_nnz(CFG3_DIMS, 3) == CFG3_NNZ_K
    assert K.shape == (3 * 257 * 65 * 65,) * 2 and K.nnz == CFG3_NNZ_K and K.block == 3
    assert K.grid == CFG3_DIMS
    assert S.n_free == CFG3_NFREE and S.nnz == CFG3_NNZ_KFF


def test_cfg3_multigrid_solution_host_residual(cfg3):
    """The measured path's solution, checked on the host against the exported K."""
    print("[case] cfg3 multigrid solve + host FP64 residual", flush=True)
    S, f, fixed = cfg3["S"], cfg3["f"], cfg3["fixed"]
    info = S.set_grid(CFG3_DIMS)
    assert info["matrix_free"] == 2 and info["levels"] == 7
    st = S.solve("fp64", "bsr", rtol=1e-10, preconditioner="multigrid")
    assert st["converged"] and st["iterations"] <= 40
    q, R = S.solution(reactions=True)
    cfg3["q_mg"] = q
    print(f"  iterations {st['iterations']}, solve {st['solve_ms']:.1f} ms", flush=True)
    Kh = cfg3["K"].to_scipy()
    cfg3["csc"] = (Kh.indptr.copy(), Kh.indices.copy(), Kh.data.copy())
    free = fixed == 0
    r = f - Kh @ q
    rel = np.linalg.norm(r[free]) / np.linalg.norm(f[free])
    print(f"  host FP64 relative residual {rel:.3e} (device {st['rel_residual']:.3e})", flush=True)
    assert rel <= 1.001e-10
    assert np.abs(q[fixed == 1]).max() == 0.0
    # rigid translations are in K's null space (before any Dirichlet data)
    kmax = np.abs(Kh.data).max()
    for d in range(3):
        t = np.zeros(Kh.shape[0])
        t[d::3] = 1.0
        assert np.abs(Kh @ t).max() <= 1e-13 * kmax
    # reactions balance the loads, component by component
    fsum = f.reshape(-1, 3).sum(0)
    rsum = R[fixed == 1].reshape(-1, 3).sum(0)
    assert np.abs(rsum + fsum).max() <= 1e-9 * np.abs(f).sum()
    assert np.abs(R[free]).max() == 0.0


def test_cfg3_against_fp64_jacobi_and_mixed(cfg3):
    print("[case] cfg3 fp64/bj to 1e-12 and mixed/bj", flush=True)
    S = cfg3["S"]
    q_mg = cfg3.get("q_mg")
    st = S.solve("fp64", "bsr", rtol=1e-12, preconditioner="block_jacobi", raise_on_fail=False)
    assert st["stop_reason"] in (1, 3)
    q_ref = S.solution()
    print(f"  fp64/bj {st['iterations']} iterations, relres {st['rel_residual']:.2e}", flush=True)
    if q_mg is not None:
        e = relerr(q_mg, q_ref)
        print(f"  multigrid vs fp64/bj: {e:.2e}", flush=True)
        assert e <= ERR
    st = S.solve("mixed", "bsr", rtol=1e-10, preconditioner="block_jacobi")
    e = relerr(S.solution(), q_ref)
    print(f"  mixed/bj {st['iterations']} iterations, error {e:.2e}", flush=True)
    assert st["converged"] and e <= ERR


def test_cfg3_drop_in_reaches_the_grid_path(ctx, cfg3):
    """assemble_system's own signature (the one-shot sfem_assemble) detects
    the grid: bitwise the same K, and the solve reaches the matrix-free
    multigrid levels without being told the element counts."""
    print("[case] cfg3 drop-in one-shot assembly", flush=True)
    m = cfg3["m"]
    K2 = smearfem.assemble_system(0, m.NodeList, m.IEN, 3, "Q1", 3, m.ID, cfg3["young"], 0.3, ctx=ctx,
                                  device=True, n_elem=m.n_elem)
    assert K2.grid == CFG3_DIMS
    ref = cfg3.get("csc")
    if ref is not None:
        got = K2.to_scipy()
        assert np.array_equal(got.indptr, ref[0]) and np.array_equal(got.indices, ref[1])
        assert np.array_equal(got.data, ref[2])
        del got
    q, st = smearfem.solve_system(K2, cfg3["f"], cfg3["fixed"], precision="fp64", rtol=1e-10,
                                  preconditioner="auto", return_stats=True)
    K2.close()
    assert st["preconditioner"] == "multigrid" and st["converged"]
    if cfg3.get("q_mg") is not None:
        assert np.array_equal(q, cfg3["q_mg"])  # same K, same hierarchy, same arithmetic


# -------------------------------------------------------- 512 problems --
@pytest.fixture(scope="module")
def large_direct():
    return json.loads((GOLD / "large_direct.json").read_text())


def test_cfg4_full_batch(ctx, large_direct):
    print("[case] cfg4 512 problems", flush=True)
    probs = uneven_hex8_problems(512, 4)
    fx = large_direct["batch"]["problems"]
    B = Batch(probs, ctx)
    assert B.nprob == 512
    # FP64 reference of every problem: the whole batch to 1e-13 (stops at the floor)
    B.solve("fp64", rtol=1e-13, item_slices=8, blocks_per_cu=2, raise_on_fail=False,
            preconditioner="block_jacobi")
    ref = B.solution()
    for k, s in fx.items():
        p = probs[int(k)]
        check_against_fixture(ref[int(k)][p.fixed == 0], s, large_direct["n_proj"])
    for precision in ("mixed", "split"):
        st, per = B.solve(precision, rtol=1e-10, item_slices=8, blocks_per_cu=2, preconditioner="block_jacobi")
        qs = B.solution()
        assert st["not_converged"] == 0 and st["launches"] == 1
        worst = max(relerr(q, r) for q, r in zip(qs, ref))
        dp = max(check_against_fixture(qs[int(k)][probs[int(k)].fixed == 0], s, large_direct["n_proj"])
                 for k, s in fx.items())
        print(f"  {precision}/bj: {st['solve_ms']:.0f} ms, max error vs FP64 batch {worst:.2e}, "
              f"largest three vs SuperLU (projection) {dp:.2e}", flush=True)
        assert worst <= ERR
        if precision == "mixed":
            q_mix = qs
    B.close()
    # the same problems alone: bitwise the in-batch result
    for i in (4, 0, 511):
        A = Batch([probs[i]], ctx)
        A.solve("mixed", rtol=1e-10, item_slices=8, blocks_per_cu=2, preconditioner="block_jacobi")
        assert np.array_equal(A.solution()[0], q_mix[i]), i
        A.close()


# ------------------------------------------------------ served requests --
@pytest.mark.parametrize("pc", ["multigrid", "block_jacobi"])
def test_cfg5_large_requests_vs_direct(ctx, large_direct, pc):
    print(f"[case] cfg5 large requests {pc}", flush=True)
    specs = large_direct["serve"]
    reqs = [random_request(np.random.Generator(np.random.PCG64(s["spec"][0])), *s["spec"][1:]) for s in specs]
    if pc == "block_jacobi":  # the table path: the 3D request (2D ones need thousands of Jacobi iterations)
        keep = [i for i, r in enumerate(reqs) if r.mesh.ndim == 3]
        reqs, specs = [reqs[i] for i in keep], [specs[i] for i in keep]
    S = Server(ctx, precision="fp64", rtol=1e-10, hbm_budget=32 << 30, preconditioner=pc)
    ids = [S.submit(r) for r in reqs]
    S.drain()
    for rid, r, s in zip(ids, reqs, specs):
        p = S.poll(rid)
        assert p["state"] == "done" and p["stop_reason"] == 1, p
        q = S.take(rid, r.n_dof)
        e = check_against_fixture(q[r.fixed == 0], s, large_direct["n_proj"])
        print(f"  {r.mesh.dims}: {p['iterations']} iterations, projection error {e:.2e}", flush=True)
    S.close()
