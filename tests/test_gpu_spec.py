"""Problems generated on the device from specs, against the oracle's
restatement of the generator (tests/test_oracle_spec.py pins that one by
closed forms), and served by the multigrid workers.

Bar: coordinates, connectivity, clamp and the consistent traction / body
loads bitwise equal to the restatement (both evaluate the same roundings);
the moduli and the random loads (exp / log / cos) within a few ulps;
solutions within 1e-8 of the oracle's direct solve of the restated inputs;
a served request bitwise the same alone and in a mix, with any number of
workers.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as O
import smearfem
from smearfem.batch import Batch, uneven_hex8_problems, uneven_hex8_specs
from smearfem.serve import Server
from smearfem.solve import System
from smearfem.spec import ProblemSpec, cantilever_spec, random_request_spec
from smearfem.vtk import read_vtu, write_solution

pytestmark = pytest.mark.gpu
ERR = 1e-8


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def direct(sp: ProblemSpec):
    r = O.spec_inputs(sp.as_dict())
    K = O.assemble_system(0, r["NodeList"], r["IEN"], sp.ndim, "Q1", sp.ndim, r["ID"], r["young"], sp.nu,
                          n_elem=r["n_elem"])
    Kff, rhs = O.eliminate(K, r["fixed"], r["f"])
    u = O.dense_solve(Kff, rhs) if Kff.shape[0] <= 3000 else spla.spsolve(Kff.tocsc(), rhs)
    return O.recombine(r["fixed"], None, u)


SPECS = [
    ProblemSpec(3, (6, 4, 5), (0.0, 1.5, 0.0, 1.0, -0.5, 0.75), (0.1, 10.0), 0.3, 11, (0, 0), (0, 1),
                (0.0, 0.0, -2.0), (0.0, 0.5, 0.0), 0.7),
    ProblemSpec(3, (5, 7, 3), (0.0, 1.0, 0.0, 2.0, 0.0, 0.5), (1.0, 1.0), 0.25, 12, (1, 1), (2, 0),
                (1.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.0),
    ProblemSpec(2, (9, 4), (0.0, 3.0, 0.0, 1.0), (0.5, 2.0), 0.35, 13, (0, 0), (1, 1), (0.0, -1.0),
                (0.3, 0.0), 0.2),
]


@pytest.mark.parametrize("k", range(len(SPECS)))
def test_spec_inputs_match_oracle(ctx, k):
    sp = SPECS[k]
    print(f"[case] spec inputs {k} {sp.dims}", flush=True)
    got = sp.export(ctx)
    ref = O.spec_inputs(sp.as_dict())
    assert np.array_equal(got["NodeList"], ref["NodeList"])
    assert np.array_equal(got["fixed"], ref["fixed"])
    assert np.allclose(got["young"], ref["young"], rtol=4e-16 * 8, atol=0)
    if sp.noise == 0.0:
        assert np.array_equal(got["f"], ref["f"])
    else:
        assert np.allclose(got["f"], ref["f"], rtol=0, atol=1e-14 * max(1.0, np.abs(ref["f"]).max()))
    mesh, bc = sp.build(ctx)
    nodes, ien, idm = mesh.export()
    assert np.array_equal(ien, ref["IEN"]) and np.array_equal(idm, ref["ID"])
    bc.close()
    mesh.close()


@pytest.mark.parametrize("k", range(len(SPECS)))
def test_spec_problem_solves_to_oracle(ctx, k):
    """build -> grid assembly -> elimination -> multigrid (matrix-free finest
    level when the moduli allow it) against the oracle's direct solve."""
    sp = SPECS[k]
    print(f"[case] spec solve {k} {sp.dims}", flush=True)
    mesh, bc = sp.build(ctx)
    K = mesh.assemble()
    assert K.grid == sp.dims
    S = System(K, bc)
    bc.close()
    info = S.set_grid()
    st = S.solve("fp64", rtol=1e-11, preconditioner="multigrid")
    q = S.solution()
    assert st["converged"] and relerr(q, direct(sp)) <= ERR
    if sp.young[0] == sp.young[1]:
        assert info["matrix_free"] == 2
    S.close()
    K.close()
    mesh.close()


def test_batch_from_specs_equals_host_batch(ctx):
    """The cfg4 generator's problems (same draws) from specs and from host
    arrays give the same solutions (their inputs differ in the last bits:
    consistent loads integrated vs area shares)."""
    print("[case] batch from specs", flush=True)
    n, seed = 24, 31
    specs = uneven_hex8_specs(n, seed, lo=2, hi=7)
    probs = uneven_hex8_problems(n, seed, lo=2, hi=7)
    Bs = Batch.from_specs(specs, ctx)
    Bh = Batch(probs, ctx)
    assert Bs.n == Bh.n and Bs.slices == Bh.slices
    for B in (Bs, Bh):
        st, per = B.solve("split", rtol=1e-12, item_slices=2, preconditioner="block_jacobi")
        assert st["not_converged"] == 0
    qs, qh = Bs.solution(), Bh.solution()
    for i, (a, b) in enumerate(zip(qs, qh)):
        assert relerr(a, b) <= 1e-10, i
    for i in (0, n - 1):
        assert relerr(qs[i], direct(specs[i])) <= ERR
    Bs.close()
    Bh.close()


@pytest.mark.parametrize("workers", [1, 4])
def test_spec_server_alone_equals_mix(ctx, workers, tmp_path):
    """Spec requests served by the multigrid workers: each result is bitwise
    the same alone and in a mix of concurrently solved requests, and meets
    the oracle bar."""
    print(f"[case] spec server workers={workers}", flush=True)
    rng = np.random.Generator(np.random.PCG64(5))
    reqs = [random_request_spec(rng, 1e3, 3e4) for _ in range(10)]
    reqs.append(cantilever_spec((1, 1, 1), young=(1.0, 1.0), end_load=(0.0, 0.0, -1.0)))  # no hierarchy
    S = Server(ctx, precision="fp64", rtol=1e-11, preconditioner="multigrid", mg_workers=workers,
               item_slices=2, iters_per_step=32)
    alone = {}
    for i in (0, 3):
        rid = S.submit(reqs[i])
        S.drain()
        alone[i] = S.take(rid, reqs[i].n_dof)
    ids = [S.submit(r) for r in reqs]
    peak = 0
    while True:
        info = S.info()
        peak = max(peak, info["active"])
        if not info["queued"] and not info["active"]:
            break
        S.step()
    for k, (rid, sp) in enumerate(zip(ids, reqs)):
        st = S.poll(rid)
        assert st["state"] == "done", st
        q = S.take(rid, sp.n_dof)
        if k in alone:
            assert np.array_equal(q, alone[k]), k
        if sp.n_dof <= 20000:
            assert relerr(q, direct(sp)) <= ERR, k
        if k < 3:  # the served field as a VTK file, read back
            g = read_vtu(write_solution(str(tmp_path / f"req{k}"), sp, q))
            assert np.array_equal(g["fields"]["u"], q.reshape(-1, sp.ndim))
            assert g["points"].shape[0] == sp.n_nodes
    S.close()
    if workers > 1:
        assert peak >= 2  # requests were solved at the same time
