"""Serving loop: continuous batching, HBM-budget admission, alone == in mix.

A request's solution is bitwise the same when it is served alone and when it
arrives while other requests are already iterating on the device (and they
keep iterating while it is admitted); each is within 1e-8 relative of a
direct FP64 solve of the oracle's K.
"""
import numpy as np
import pytest

import oracle as O
import smearfem
from smearfem.batch import make_problem
from smearfem.mesh import clamp_x0, random_face_load, structured_mesh
from smearfem.serve import Server, random_request, run_poisson

pytestmark = pytest.mark.gpu
ERR = 1e-8


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def request(dims, seed):
    rng = np.random.default_rng(seed)
    m = structured_mesh(dims, lengths=tuple(d / max(dims) for d in dims))
    fixed = clamp_x0(m)
    f = random_face_load(m, 0, 1, fixed, seed=seed)
    return make_problem(m, 10.0 ** rng.uniform(-1, 1, m.n_elem), 0.3, f, fixed)


def oracle_solution(p):
    m = p.mesh
    K = O.assemble_system(0, m.NodeList, m.IEN, m.ndim, "Q1", m.ndim, m.ID, p.young, 0.3, n_elem=m.n_elem)
    Kff, rhs = O.eliminate(K, p.fixed, p.f)
    return O.recombine(p.fixed, None, O.dense_solve(Kff, rhs))


@pytest.mark.parametrize("precision,pc", [("mixed", "jacobi"), ("fp64", "jacobi"), ("split", "block_jacobi")])
def test_alone_equals_in_mix(ctx, precision, pc):
    print(f"[case] serve alone==mix {precision} {pc}", flush=True)
    target = request((6, 3, 3), seed=1)
    others = [request((8, 4, 3), 2), request((12, 5), 3), request((5, 5, 5), 4), request((20, 6), 5)]
    S = Server(ctx, precision=precision, rtol=1e-11, iters_per_step=16, item_slices=2, preconditioner=pc)
    rid = S.submit(target)
    S.drain()
    q_alone = S.take(rid, target.n_dof)
    # the mix: two requests start first, the target arrives while they run,
    # the rest arrive after it
    ids = [S.submit(others[0]), S.submit(others[1])]
    S.step()
    S.step()
    assert S.info()["active"] == 2
    rid = S.submit(target)
    S.step()
    assert S.info()["active"] == 3
    ids += [S.submit(others[2]), S.submit(others[3])]
    S.drain()
    q_mix = S.take(rid, target.n_dof)
    assert np.array_equal(q_alone, q_mix)
    assert relerr(q_mix, oracle_solution(target)) <= ERR
    for i, p in zip(ids, others):
        st = S.poll(i)
        assert st["state"] == "done" and st["stop_reason"] == 1
        assert relerr(S.take(i, p.n_dof), oracle_solution(p)) <= ERR
    S.close()


def test_partially_fixed_request_in_mix(ctx):
    """A request with a roller plane (only u_z fixed on z = 0) is served next
    to plain cantilevers: same bits alone and in the mix, oracle parity."""
    print("[case] serve partial clamp in mix", flush=True)
    m = structured_mesh((5, 3, 3), lengths=(1.0, 0.6, 0.6))
    fixed = clamp_x0(m)
    fixed[3 * np.nonzero(m.NodeList[2] == 0.0)[0] + 2] = 1
    f = random_face_load(m, 0, 1, fixed, seed=7)
    target = make_problem(m, 10.0 ** np.random.default_rng(7).uniform(-1, 1, m.n_elem), 0.3, f, fixed)
    S = Server(ctx, precision="fp64", rtol=1e-11, iters_per_step=16, item_slices=2, preconditioner="block_jacobi")
    rid = S.submit(target)
    S.drain()
    q_alone = S.take(rid, target.n_dof)
    others = [request((8, 4, 3), 2), request((12, 5), 3)]
    ids = [S.submit(others[0])]
    S.step()
    rid = S.submit(target)
    ids.append(S.submit(others[1]))
    S.drain()
    q_mix = S.take(rid, target.n_dof)
    assert np.array_equal(q_alone, q_mix)
    assert relerr(q_mix, oracle_solution(target)) <= ERR
    assert np.all(q_mix[fixed == 1] == 0.0)
    for i, p in zip(ids, others):
        assert relerr(S.take(i, p.n_dof), oracle_solution(p)) <= ERR
    S.close()


def test_multigrid_server(ctx):
    """Requests with a grid are solved one at a time with the multigrid
    preconditioner (a grid with no even element count falls back to the
    persistent launch); every result meets the oracle bar, and a request's
    result is bitwise the same alone and in a mix."""
    print("[case] serve multigrid", flush=True)
    reqs = [request((8, 4, 4), 21), request((12, 6), 22), request((5, 3, 3), 23), request((16, 4, 6), 24),
            request((1, 1, 1), 25)]
    S = Server(ctx, precision="fp64", rtol=1e-11, iters_per_step=16, item_slices=2, preconditioner="multigrid")
    rid = S.submit(reqs[0])
    S.drain()
    alone = S.take(rid, reqs[0].n_dof)
    ids = [S.submit(p) for p in reqs]
    S.step()
    S.drain()
    for i, p in zip(ids, reqs):
        st = S.poll(i)
        assert st["state"] == "done" and st["stop_reason"] == 1, st
        if p.mesh.dims != (1, 1, 1):  # one element: no hierarchy, node-block Jacobi
            assert st["iterations"] <= 40, st
        q = S.take(i, p.n_dof)
        assert relerr(q, oracle_solution(p)) <= ERR
        if i == ids[0]:
            assert np.array_equal(q, alone)
    S.close()


def test_hbm_budget_admission(ctx):
    print("[case] serve budget", flush=True)
    reqs = [request((6, 4, 4), s) for s in range(4)]
    # a budget that holds one of these requests at a time
    S = Server(ctx, precision="mixed", rtol=1e-10, iters_per_step=8, hbm_budget=3 << 20)
    ids = [S.submit(p) for p in reqs]
    peak = 0
    while True:
        i = S.info()
        peak = max(peak, i["active"])
        assert i["bytes"] <= 3 << 20
        if not i["queued"] and not i["active"]:
            break
        S.step()
    assert peak == 1
    for rid, p in zip(ids, reqs):
        assert S.poll(rid)["state"] == "done"
        assert relerr(S.take(rid, p.n_dof), oracle_solution(p)) <= ERR
    # a request that can never fit is failed, not left waiting
    big = request((30, 10, 10), 9)
    rid = S.submit(big)
    S.step()
    assert S.poll(rid)["state"] == "failed"
    with pytest.raises(smearfem.StateError):
        S.take(rid, big.n_dof)
    S.close()


def test_poisson_stream_small(ctx):
    print("[case] serve poisson", flush=True)
    rng = np.random.Generator(np.random.PCG64(7))
    reqs = [random_request(rng, 1e3, 2e4) for _ in range(12)]
    S = Server(ctx, precision="mixed", rtol=1e-10, iters_per_step=32)
    errs = {}

    def check(idx, q, st):
        if reqs[idx].n_dof <= 6000:
            errs[idx] = relerr(q, oracle_solution(reqs[idx]))

    res = run_poisson(S, reqs, rate_per_s=200.0, seed=1, check=check)
    S.close()
    s = res.summary()
    assert s["requests"] == 12 and s["failed"] == 0
    assert s["p99_ms"] >= s["p50_ms"] > 0
    assert errs and max(errs.values()) <= ERR
