"""sfem_system_solve_noalloc (the solve that allocates nothing) against sfem_system_solve.

The no-allocation entry point runs the same kernels with the same launch
shapes on a per-GPU workspace array, so its solution must be bitwise the
allocating solve's, and both meet the 1e-8 bar against the oracle's direct
solve.  Two systems are solved one after the other, the second smaller, so a
stale workspace from the first call would show in the second.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as O
import smearfem
from smearfem import _lib
from smearfem.mesh import clamp_x0, element_young_loguniform, random_face_load, structured_mesh
from smearfem.solve import PRECISIONS, PRECONDITIONERS, BoundaryConditions, System

pytestmark = pytest.mark.gpu


def _opts(lib, precision, pc, rtol=1e-10):
    o = _lib.SolveOpts()
    lib.sfem_solve_opts_default(C.byref(o))
    o.precision = PRECISIONS[precision]
    o.format = 2  # node-block format
    o.rtol = rtol
    o.preconditioner = PRECONDITIONERS[pc]
    return o


def _system(ctx, dims, seed):
    m = structured_mesh(dims, lengths=tuple(d / max(dims) for d in dims))
    young = element_young_loguniform(m.n_elem, 0.1, 10.0, seed=seed)
    fixed = clamp_x0(m)
    f = random_face_load(m, m.ndim - 1, 1, fixed, seed=seed + 1)
    K = smearfem.assemble_system(0, m.NodeList, m.IEN, m.ndim, "Q1", m.ndim, m.ID, young, 0.3, ctx=ctx,
                                 device=True, n_elem=m.n_elem)
    bc = BoundaryConditions(ctx, K.shape[0], f, fixed)
    S = System(K, bc)
    bc.close()
    return m, young, fixed, f, K, S


def _direct(m, young, fixed, f):
    Kr = O.assemble_system(0, m.NodeList, m.IEN, m.ndim, "Q1", m.ndim, m.ID, young, 0.3, n_elem=m.n_elem)
    Kff, rhs = O.eliminate(Kr, fixed, f)
    return O.recombine(fixed, None, spla.spsolve(Kff.tocsc(), rhs))


def test_noalloc_solve_equals_allocating_solve(ctx):
    lib = ctx._lib
    for dims, seed in (((40, 12, 12), 21), ((9, 4, 3), 22)):  # the second call has the smaller input
        print(f"[case] noalloc {dims}", flush=True)
        m, young, fixed, f, K, S = _system(ctx, dims, seed)
        q_ref = _direct(m, young, fixed, f)
        for precision, pc in (("fp64", "jacobi"), ("mixed", "block_jacobi"), ("split", "block_jacobi")):
            o = _opts(lib, precision, pc)
            st1 = _lib.SolveStats()
            ctx.check(lib.sfem_system_solve(ctx.handle, S.handle, C.byref(o), C.byref(st1)))
            q1 = S.solution()
            st2 = _lib.SolveStats()
            ctx.check(lib.sfem_system_solve_noalloc(ctx.handle, S.handle, C.byref(o), C.byref(st2)))
            q2 = S.solution()
            assert st2.converged and st2.iterations == st1.iterations, (st1.iterations, st2.iterations)
            assert np.array_equal(q1, q2), (precision, pc)
            assert np.linalg.norm(q2 - q_ref) / np.linalg.norm(q_ref) <= 1e-8
        S.close()
        K.close()


def test_noalloc_solve_refusals(ctx):
    print("[case] noalloc refusals", flush=True)
    lib = ctx._lib
    # format not built yet: a fresh system refuses, the allocating solve builds it
    m, young, fixed, f, K, S = _system(ctx, (6, 3, 3), 5)
    o = _opts(lib, "fp64", "block_jacobi")
    st = _lib.SolveStats()
    assert lib.sfem_system_solve_noalloc(ctx.handle, S.handle, C.byref(o), C.byref(st)) == smearfem.StateError.code
    ctx.check(lib.sfem_system_solve(ctx.handle, S.handle, C.byref(o), C.byref(st)))
    ctx.check(lib.sfem_system_solve_noalloc(ctx.handle, S.handle, C.byref(o), C.byref(st)))
    # multigrid is refused
    mg = _opts(lib, "fp64", "multigrid")
    S.set_grid()
    assert lib.sfem_system_solve_noalloc(ctx.handle, S.handle, C.byref(mg), C.byref(st)) == \
        smearfem.UnsupportedError.code
    S.close()
    K.close()
    # beyond the workspace: 2^20 solver dofs at most (checked before any device work)
    m, young, fixed, f, K, S = _system(ctx, (100, 60, 60), 6)
    assert S.n_free > (1 << 20)
    assert lib.sfem_system_solve_noalloc(ctx.handle, S.handle, C.byref(o), C.byref(st)) == \
        smearfem.UnsupportedError.code
    S.close()
    K.close()
