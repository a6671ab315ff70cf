"""Many meshes assembled and solved together by one persistent launch.

No counterpart in the reference: every problem of a batch is the maths of one
``assemble_system`` call (src/fem.jl:135-256) followed by the example's
elimination and solve (examples/vector3D.jl:314-322).  The device side is
``sfem_batch_*`` (batch.hip): the meshes are concatenated, assembled by the
same sort-by-key pipeline, eliminated into one padded node-block operator and
iterated by a cooperative kernel that keeps one CG state per problem.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import time
from dataclasses import dataclass
from pathlib import Path
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from .fem import Context, _ptr, default_context
from .mesh import StructuredMesh, clamp_x0, face_load, structured_mesh
from .solve import PRECISIONS, PRECONDITIONERS


@dataclass
class Problem:
    """One Q4 / Hex8 vector-elasticity problem with node-interleaved dofs."""
    mesh: StructuredMesh
    young: np.ndarray   # per element
    nu: np.ndarray      # per element
    f: np.ndarray       # n_dof
    fixed: np.ndarray   # n_dof, uint8
    q_d: Optional[np.ndarray] = None

    @property
    def n_dof(self) -> int:
        return self.mesh.n_dof


def make_problem(mesh: StructuredMesh, young, nu, f, fixed, q_d=None) -> Problem:
    ne = mesh.n_elem
    y = np.broadcast_to(np.asarray(young, dtype=np.float64), (ne,)).copy()
    v = np.broadcast_to(np.asarray(nu, dtype=np.float64), (ne,)).copy()
    return Problem(mesh, y, v, np.asarray(f, dtype=np.float64), np.asarray(fixed, dtype=np.uint8),
                   None if q_d is None else np.asarray(q_d, dtype=np.float64))


def uneven_hex8_problems(n: int, seed: int, lo: int = 4, hi: int = 32) -> List[Problem]:
    """Independent Hex8 cantilevers: nx, ny, nz uniform integers in [lo, hi],
    E log-uniform in [0.1, 10] and nu uniform in [0.2, 0.4] per problem,
    x = 0 clamped, unit downward end load on x = L."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(n):
        dims = tuple(int(d) for d in rng.integers(lo, hi + 1, size=3))
        E = 10.0 ** rng.uniform(-1.0, 1.0)
        nu = rng.uniform(0.2, 0.4)
        m = structured_mesh(dims, lengths=tuple(d / max(dims) for d in dims))
        out.append(make_problem(m, E, nu, face_load(m, 0, 1, 2, -1.0), clamp_x0(m)))
    return out


def uneven_hex8_specs(n: int, seed: int, lo: int = 4, hi: int = 32):
    """The problems of ``uneven_hex8_problems`` (the same draws: dims, E and nu
    per problem) as device specs: the unit end load is the uniform traction
    -1 / (face area) in z, whose consistent nodal loads are face_load's."""
    from .spec import cantilever_spec

    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(n):
        dims = tuple(int(d) for d in rng.integers(lo, hi + 1, size=3))
        E = 10.0 ** rng.uniform(-1.0, 1.0)
        nu = rng.uniform(0.2, 0.4)
        out.append(cantilever_spec(dims, young=(E, E), nu=nu, end_load=(0.0, 0.0, -1.0)))
    return out


class Batch:
    """``sfem_batch``: the problems' padded operator, rhs and solution on the device."""

    def __init__(self, problems: Sequence[Problem], ctx: Optional[Context] = None):
        if not problems:
            raise ValueError("empty batch")
        self.ctx = ctx or default_context()
        ndim = problems[0].mesh.ndim
        if any(p.mesh.ndim != ndim for p in problems):
            raise ValueError("a batch holds problems of one dimension (2D or 3D)")
        nen = 1 << ndim
        nn = np.array([p.mesh.n_nodes for p in problems], dtype=np.int64)
        ne = np.array([p.mesh.n_elem for p in problems], dtype=np.int64)
        node_off = np.concatenate([[0], np.cumsum(nn)]).astype(np.int64)
        elem_off = np.concatenate([[0], np.cumsum(ne)]).astype(np.int64)
        nodes = np.asfortranarray(np.concatenate([p.mesh.NodeList for p in problems], axis=1))
        ien = np.asfortranarray(np.concatenate([p.mesh.IEN for p in problems], axis=0).astype(np.int64))
        assert ien.shape[1] == nen
        young = np.ascontiguousarray(np.concatenate([p.young for p in problems]))
        nu = np.ascontiguousarray(np.concatenate([p.nu for p in problems]))
        f = np.ascontiguousarray(np.concatenate([p.f for p in problems]))
        fixed = np.ascontiguousarray(np.concatenate([p.fixed for p in problems]).astype(np.uint8))
        has_qd = any(p.q_d is not None for p in problems)
        qd = (np.ascontiguousarray(np.concatenate([p.q_d if p.q_d is not None else np.zeros(p.n_dof)
                                                   for p in problems])) if has_qd else None)
        h = C.c_void_p()
        self.ctx.check(self.ctx._lib.sfem_batch_create(
            self.ctx.handle, len(problems), ndim, ndim, _ptr(nodes, C.c_double), _ptr(node_off, C.c_int64),
            _ptr(ien, C.c_int64), nen, _ptr(elem_off, C.c_int64), _ptr(young, C.c_double), _ptr(nu, C.c_double),
            _ptr(f, C.c_double), _ptr(fixed, C.c_uint8), None if qd is None else _ptr(qd, C.c_double),
            C.byref(h)))
        self.handle = h
        self.nprob = len(problems)
        self.ndim = ndim
        self.dof_off = node_off * ndim
        n, npad, sl = C.c_int64(), C.c_int64(), C.c_int64()
        am, em = C.c_double(), C.c_double()
        self.ctx.check(self.ctx._lib.sfem_batch_info(h, C.byref(n), C.byref(npad), C.byref(sl), C.byref(am),
                                                     C.byref(em)))
        self.n, self.n_padded, self.slices = n.value, npad.value, sl.value
        self.assemble_ms, self.eliminate_ms = am.value, em.value
        self.signature = {"ndim": ndim, "nprob": self.nprob, "slices": self.slices}

    @classmethod
    def from_specs(cls, specs, ctx: Optional[Context] = None) -> "Batch":
        """A batch generated on the device from ProblemSpecs
        (``sfem_batch_create_specs``): no host mesh arrays."""
        if not specs:
            raise ValueError("empty batch")
        self = cls.__new__(cls)
        self.ctx = ctx or default_context()
        arr = (_lib.ProblemSpec * len(specs))(*[sp.to_c() for sp in specs])
        h = C.c_void_p()
        self.ctx.check(self.ctx._lib.sfem_batch_create_specs(self.ctx.handle, len(specs), arr, C.byref(h)))
        self.handle = h
        self.nprob = len(specs)
        self.ndim = specs[0].ndim
        self.dof_off = np.concatenate([[0], np.cumsum([sp.n_dof for sp in specs])]).astype(np.int64)
        self._info()
        return self

    def _info(self) -> None:
        n, npad, sl = C.c_int64(), C.c_int64(), C.c_int64()
        am, em = C.c_double(), C.c_double()
        self.ctx.check(self.ctx._lib.sfem_batch_info(self.handle, C.byref(n), C.byref(npad), C.byref(sl),
                                                     C.byref(am), C.byref(em)))
        self.n, self.n_padded, self.slices = n.value, npad.value, sl.value
        self.assemble_ms, self.eliminate_ms = am.value, em.value
        self.signature = {"ndim": self.ndim, "nprob": self.nprob, "slices": self.slices}

    def solve(self, precision="mixed", rtol=1e-10, max_iter=100000, item_slices=8, blocks_per_cu=0,
              check_every=0, reliable_delta=0.1, raise_on_fail=True, preconditioner="jacobi"):
        """Returns (stats dict, list of per-problem stats dicts)."""
        lib = self.ctx._lib
        o = _lib.SolveOpts()
        lib.sfem_solve_opts_default(C.byref(o))
        o.precision = PRECISIONS[precision]
        o.rtol = float(rtol)
        o.max_iter = int(max_iter)
        o.reliable_delta = float(reliable_delta)
        o.check_every = int(check_every)
        o.x0_zero = 1
        o.preconditioner = PRECONDITIONERS[preconditioner]
        per = (_lib.BatchProblemStats * self.nprob)()
        s = _lib.BatchStats()
        st = lib.sfem_batch_solve(self.ctx.handle, self.handle, C.byref(o), int(item_slices), int(blocks_per_cu),
                                  per, C.byref(s))
        stats = {name: getattr(s, name) for name, _ in _lib.BatchStats._fields_}
        stats.update(status=st, precision=precision, item_slices=int(item_slices),
                     blocks_per_cu=int(blocks_per_cu), preconditioner=preconditioner)
        plist = [{name: getattr(p, name) for name, _ in _lib.BatchProblemStats._fields_} for p in per]
        if st and raise_on_fail:
            self.ctx.check(st)
        return stats, plist

    def solution(self) -> List[np.ndarray]:
        q = np.zeros(self.n)
        self.ctx.check(self.ctx._lib.sfem_batch_solution(self.ctx.handle, self.handle, _ptr(q, C.c_double)))
        return [q[self.dof_off[i]:self.dof_off[i + 1]].copy() for i in range(self.nprob)]

    def close(self) -> None:
        if self.handle:
            self.ctx._lib.sfem_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        if _lib.finalizing():  # the HIP runtime may be gone: leave it to process exit
            return
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ autotuner --
def _cache_path() -> Path:
    return Path(os.environ.get("SMEARFEM_TUNE_CACHE", Path.home() / ".cache" / "smearfem_tune.json"))


def _bucket(v: int) -> int:
    return int(np.floor(np.log2(max(v, 1))))


def tune_key(ctx: Context, batch: Batch, precision: str, preconditioner: str = "jacobi") -> str:
    info = ctx.device_info()
    return (f"{info['name'].strip()}|cu{info['n_cu']}|{precision}|{preconditioner}|d{batch.ndim}"
            f"|p{_bucket(batch.nprob)}|s{_bucket(batch.slices)}")


def autotune(batch: Batch, precision="mixed", item_slices=(2, 4, 8, 16, 32), blocks_per_cu=None,
             probe_iters=40, use_cache=True, verbose=False, preconditioner="jacobi") -> dict:
    """Pick (item_slices, blocks_per_cu) for a batch by timing short solves.

    Every candidate runs ``probe_iters`` CG iterations of every problem from
    x = 0 (tolerance 0, so no problem drops out) and is scored by device ms
    per iteration.  Results are cached in a JSON file keyed by device,
    precision, dimension and the batch's size class.
    """
    ctx = batch.ctx
    key = tune_key(ctx, batch, precision, preconditioner)
    path = _cache_path()
    cache = {}
    if use_cache and path.exists():
        try:
            cache = json.loads(path.read_text())
        except (OSError, ValueError):
            cache = {}
        if key in cache:
            return dict(cache[key], cached=True)
    maxb = C.c_int()
    ctx.check(ctx._lib.sfem_batch_max_blocks_per_cu(ctx.handle, PRECISIONS[precision], C.byref(maxb)))
    if blocks_per_cu is None:
        blocks_per_cu = sorted({b for b in (1, 2, 3, 4, 6, 8) if b <= maxb.value} | {maxb.value})
    trials = []
    t0 = time.perf_counter()
    for isl in item_slices:
        for bpc in blocks_per_cu:
            st, _ = batch.solve(precision, rtol=0.0, max_iter=probe_iters, item_slices=isl, blocks_per_cu=bpc,
                                raise_on_fail=False, preconditioner=preconditioner)
            ms_it = st["solve_ms"] / max(st["max_iterations"], 1)
            trials.append({"item_slices": isl, "blocks_per_cu": bpc, "ms_per_iter": ms_it})
            if verbose:
                print(f"[tune] item_slices={isl} blocks_per_cu={bpc}: {ms_it:.4f} ms/iter", flush=True)
    best = min(trials, key=lambda t: t["ms_per_iter"])
    out = {"item_slices": best["item_slices"], "blocks_per_cu": best["blocks_per_cu"],
           "ms_per_iter": best["ms_per_iter"], "trials": trials, "tune_s": time.perf_counter() - t0,
           "max_blocks_per_cu": maxb.value}
    if use_cache:
        cache[key] = {k: v for k, v in out.items() if k != "trials"}
        try:
            path.parent.mkdir(parents=True, exist_ok=True)
            path.write_text(json.dumps(cache, indent=1, sort_keys=True))
        except OSError:
            pass
    return dict(out, cached=False)
