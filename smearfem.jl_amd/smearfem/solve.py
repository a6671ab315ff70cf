"""Dirichlet elimination and the K u = f solve on the GPU.

The reference solves only inside its example driver
(examples/vector3D.jl:314-322):

    q_d, C = setboundaryCond(...)            # prescribed values, C = I[:, free]
    K_free = C' * K_bar * C
    q_f    = inv(Matrix(K_free)) * C' * (-K_bar * q_d)
    q      = q_d + C * q_f

``solve_system`` keeps that meaning (plus an optional load vector f:
rhs = C'(f - K q_d)) and replaces the dense inverse by Jacobi-PCG on the
device, in FP64 or in mixed FP32/FP64 precision.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from .fem import Context, StiffnessMatrix, _ptr, default_context

PRECISIONS = {"fp64": 0, "mixed": 1, "fp32_matrix": 2, "split": 3}
PRECONDITIONERS = {"jacobi": 0, "block_jacobi": 1, "multigrid": 2}
FORMATS = {"auto": 0, "csr": 1, "bsr": 2}
FORMAT_NAMES = {v: k for k, v in FORMATS.items()}


class BoundaryConditions:
    """Loads and Dirichlet data on the device (``sfem_bc``)."""

    def __init__(self, ctx: Context, n: int, f=None, fixed=None, q_d=None):
        self.ctx = ctx
        fv = None if f is None else np.ascontiguousarray(f, dtype=np.float64)
        mv = None if fixed is None else np.ascontiguousarray(np.asarray(fixed) != 0, dtype=np.uint8)
        qv = None if q_d is None else np.ascontiguousarray(np.ravel(q_d), dtype=np.float64)
        for name, v in (("f", fv), ("fixed", mv), ("q_d", qv)):
            if v is not None and v.size != n:
                raise ValueError(f"{name} has {v.size} entries, size(K,1) = {n}")
        h = C.c_void_p()
        ctx.check(ctx._lib.sfem_bc_create(ctx.handle, int(n), None if fv is None else _ptr(fv, C.c_double),
                                          None if mv is None else _ptr(mv, C.c_uint8),
                                          None if qv is None else _ptr(qv, C.c_double), C.byref(h)))
        self.handle = h
        self.n = int(n)

    @classmethod
    def _wrap(cls, ctx: Context, handle: C.c_void_p, n: int) -> "BoundaryConditions":
        obj = cls.__new__(cls)
        obj.ctx, obj.handle, obj.n = ctx, handle, int(n)
        return obj

    def close(self) -> None:
        if self.handle:
            self.ctx._lib.sfem_bc_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        if _lib.finalizing():  # the HIP runtime may be gone: leave it to process exit
            return
        try:
            self.close()
        except Exception:
            pass


class System:
    """K_ff, rhs and the PCG state on the device (``sfem_system``).  The
    stiffness matrix it was built from must stay alive (reactions use it)."""

    def __init__(self, K: StiffnessMatrix, bc: BoundaryConditions):
        self.ctx = K.ctx
        self.K = K
        h = C.c_void_p()
        self.ctx.check(self.ctx._lib.sfem_eliminate(self.ctx.handle, K.handle, bc.handle, C.byref(h)))
        self.handle = h
        nf, bs = C.c_int64(), C.c_int()
        self.ctx.check(self.ctx._lib.sfem_system_info(h, C.byref(nf), None, C.byref(bs)))
        ms = C.c_double()
        self.ctx.check(self.ctx._lib.sfem_system_eliminate_ms(h, C.byref(ms)))
        self.eliminate_ms = ms.value
        self.n = K.shape[0]
        self.n_free = nf.value
        self.block = bs.value

    @property
    def nnz(self) -> int:
        """Entries of K_ff (the scalar row counts are formed on first request)."""
        nnz = C.c_int64()
        self.ctx.check(self.ctx._lib.sfem_system_info(self.handle, None, C.byref(nnz), None))
        return nnz.value

    def solve(self, precision="fp64", fmt="auto", rtol=1e-10, max_iter=200000, reliable_delta=0.1,
              check_every=64, time_kernels=False, x0_zero=True, raise_on_fail=True,
              preconditioner="jacobi") -> dict:
        lib = self.ctx._lib
        o = _lib.SolveOpts()
        lib.sfem_solve_opts_default(C.byref(o))
        o.precision = PRECISIONS[precision]
        o.format = FORMATS[fmt]
        o.rtol = float(rtol)
        o.max_iter = int(max_iter)
        o.reliable_delta = float(reliable_delta)
        o.check_every = int(check_every)
        o.time_kernels = 1 if time_kernels else 0
        o.x0_zero = 1 if x0_zero else 0
        o.preconditioner = PRECONDITIONERS[preconditioner]
        s = _lib.SolveStats()
        st = lib.sfem_system_solve(self.ctx.handle, self.handle, C.byref(o), C.byref(s))
        stats = {name: getattr(s, name) for name, _ in _lib.SolveStats._fields_}
        stats["format"] = FORMAT_NAMES.get(stats["format"], stats["format"])
        stats["precision"] = precision
        stats["preconditioner"] = preconditioner
        if st and raise_on_fail:
            self.ctx.check(st)
        stats["status"] = st
        return stats

    def set_bc(self, bc: BoundaryConditions) -> None:
        """New loads / prescribed values with the same fixed dofs
        (``sfem_system_set_bc``); the current solution stays as the start of
        the next ``solve(x0_zero=False)``."""
        self.ctx.check(self.ctx._lib.sfem_system_set_bc(self.ctx.handle, self.handle, bc.handle))

    def set_grid(self, dims=None, degree=2, lo=0.1, hi=1.1, power_iters=8, coarse_nodes=64, fine_fp16=True,
                 coarse_degree=0, coarse_fp16=True, matrix_free=2) -> dict:
        """Declare the nodes a structured grid with ``dims`` elements per
        direction (meshgrid numbering, x fastest) and build the geometric
        multigrid hierarchy (``sfem_system_set_grid``); then
        ``solve(preconditioner="multigrid")``.  ``dims=None``: the grid K was
        assembled on.  Returns ``mg_info()``."""
        if dims is None:
            dims = self.K.grid
            if dims is None:
                raise ValueError("set_grid: K was not assembled on a grid; pass dims")
        d = np.asarray(dims, dtype=np.int64).ravel()
        o = _lib.MgOpts()
        self.ctx._lib.sfem_mg_opts_default(C.byref(o))
        o.degree, o.lo, o.hi = int(degree), float(lo), float(hi)
        o.coarse_degree = int(coarse_degree)
        o.power_iters, o.coarse_nodes = int(power_iters), int(coarse_nodes)
        o.fine_fp16 = 1 if fine_fp16 else 0
        o.coarse_fp16 = 1 if coarse_fp16 else 0
        o.matrix_free = int(matrix_free)
        self.ctx.check(self.ctx._lib.sfem_system_set_grid(self.ctx.handle, self.handle, int(d.size),
                                                          _ptr(d, C.c_int64), C.byref(o)))
        self.grid = tuple(int(v) for v in d)
        return self.mg_info()

    def mg_info(self) -> dict:
        info = _lib.MgInfo()
        self.ctx.check(self.ctx._lib.sfem_system_mg_info(self.handle, C.byref(info)))
        nl = info.levels
        nd = len(getattr(self, "grid", (0, 0, 0)))
        return {"levels": nl, "degree": info.degree, "fine_fp16": bool(info.fine_fp16),
                "coarse_fp16": bool(info.coarse_fp16), "nodes": [info.nodes[i] for i in range(nl)],
                "dims": [tuple(info.dims[i][d] for d in range(nd)) for i in range(nl)],
                "lambda_max": [info.lambda_max[i] for i in range(nl)], "coarse_dofs": info.coarse_dofs,
                "setup_ms": info.setup_ms, "matrix_free": info.matrix_free, "ke_deviation": info.ke_deviation}

    def mg_stencil(self, level: int):
        """Level >= 1 Galerkin operator as stored (FP64, fixed rows / columns
        zero): array [slot, a, b, node] with slot = (dx+1) + 3(dy+1) [+ 9(dz+1)],
        and the fixed-dof flags [node, a]."""
        info = self.mg_info()
        nn = info["nodes"][level]
        bs = self.block
        ns = 3 ** len(self.grid)
        st = np.zeros(ns * bs * bs * nn)
        fx = np.zeros(nn * bs, dtype=np.uint8)
        self.ctx.check(self.ctx._lib.sfem_system_mg_stencil(self.ctx.handle, self.handle, int(level),
                                                            _ptr(st, C.c_double), _ptr(fx, C.c_uint8)))
        return st.reshape(ns, bs, bs, nn), fx.reshape(nn, bs)

    def solve_many(self, q_d=None, f=None, rtol=1e-10, max_iter=200000, check_every=32, raise_on_fail=True):
        """k right-hand sides at once (``sfem_system_solve_many``): column j of
        the n x k arrays q_d / f gives q[:, j] = q_d + C K_ff^-1 C'(f - K q_d).
        Returns (q n x k, per-column stats list, device ms)."""
        cols = [a for a in (q_d, f) if a is not None]
        if not cols:
            raise ValueError("solve_many needs q_d or f")
        k = np.asarray(cols[0]).reshape(self.n, -1).shape[1]
        qv = None if q_d is None else np.asfortranarray(np.asarray(q_d, dtype=np.float64).reshape(self.n, k))
        fv = None if f is None else np.asfortranarray(np.asarray(f, dtype=np.float64).reshape(self.n, k))
        o = _lib.SolveOpts()
        self.ctx._lib.sfem_solve_opts_default(C.byref(o))
        o.rtol = float(rtol)
        o.max_iter = int(max_iter)
        o.check_every = int(check_every)
        q = np.zeros((self.n, k), order="F")
        per = (_lib.BatchProblemStats * k)()
        ms = C.c_double()
        st = self.ctx._lib.sfem_system_solve_many(self.ctx.handle, self.handle, k,
                                                  None if qv is None else _ptr(qv, C.c_double),
                                                  None if fv is None else _ptr(fv, C.c_double), C.byref(o),
                                                  _ptr(q, C.c_double), per, C.byref(ms))
        if st and raise_on_fail:
            self.ctx.check(st)
        stats = [{name: getattr(p, name) for name, _ in _lib.BatchProblemStats._fields_} for p in per]
        return q, stats, ms.value

    def solution(self, reactions: bool = False):
        q = np.zeros(self.n)
        R = np.zeros(self.n) if reactions else None
        self.ctx.check(self.ctx._lib.sfem_system_solution(self.ctx.handle, self.handle, _ptr(q, C.c_double),
                                                          None if R is None else _ptr(R, C.c_double)))
        return (q, R) if reactions else q

    def spmv_bench(self, precision="fp64", fmt="auto", reps=20):
        ms, nbytes = C.c_double(), C.c_double()
        self.ctx.check(self.ctx._lib.sfem_system_spmv_bench(self.ctx.handle, self.handle, PRECISIONS[precision],
                                                            FORMATS[fmt], int(reps), C.byref(ms), C.byref(nbytes)))
        return ms.value, nbytes.value

    def close(self) -> None:
        if self.handle:
            self.ctx._lib.sfem_system_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        if _lib.finalizing():  # the HIP runtime may be gone: leave it to process exit
            return
        try:
            self.close()
        except Exception:
            pass


def solve_system(K: StiffnessMatrix, f=None, fixed=None, q_d=None, *, precision="fp64", fmt="auto",
                 rtol=1e-10, max_iter=200000, reliable_delta=0.1, reactions=False, return_stats=False,
                 preconditioner="jacobi", grid=None):
    """q = q_d + C K_ff^{-1} C'(f - K q_d) on the GPU for a device matrix K.
    preconditioner="multigrid" needs the structured grid the nodes come from:
    ``grid`` (elements per direction), else the grid K was assembled on.
    preconditioner="auto": multigrid when K was assembled on a grid (or
    ``grid`` is given) and the precision allows it, else point Jacobi."""
    if grid is None:
        grid = K.grid
    if preconditioner == "auto":
        preconditioner = "multigrid" if grid is not None and precision in ("fp64", "split", "fp32_matrix") \
            else "jacobi"
    if preconditioner == "multigrid" and grid is None:
        raise ValueError("multigrid needs grid= (elements per direction) or K assembled on a grid")
    bc = BoundaryConditions(K.ctx, K.shape[0], f, fixed, q_d)
    try:
        S = System(K, bc)
    finally:
        bc.close()
    try:
        if preconditioner == "multigrid":
            S.set_grid(grid)
        stats = S.solve(precision, fmt, rtol, max_iter, reliable_delta, preconditioner=preconditioner)
        out = S.solution(reactions=reactions)
    finally:
        S.close()
    if reactions:
        q, R = out
        return (q, R, stats) if return_stats else (q, R)
    return (out, stats) if return_stats else out


def stream_copy_bench(ctx: Optional[Context] = None, nbytes: int = 1 << 30, reps: int = 20) -> float:
    """Average ms of one device-to-device copy of nbytes with the library's copy kernel."""
    ctx = ctx or default_context()
    ms = C.c_double()
    ctx.check(ctx._lib.sfem_stream_copy_bench(ctx.handle, int(nbytes), int(reps), C.byref(ms)))
    return ms.value
