"""Serving: continuous batching of solve requests on one GPU.

No counterpart in the reference.  ``Server`` wraps ``sfem_server_*``
(serve.hip): requests queue on the host, each ``step`` admits queued
requests in arrival order while a slot is free and their device footprint
fits the HBM budget, runs ``iters_per_step`` CG iterations of every active
request in one persistent launch and retires the finished ones.  A request's
result is bitwise the same alone or in any mix.

``run_poisson`` drives a server with a Poisson arrival stream in wall-clock
time and reports per-request latency (arrival to solution on the host).
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib
from .batch import Problem, make_problem
from .fem import Context, _ptr, default_context
from .mesh import clamp_x0, structured_mesh
from .solve import PRECISIONS, PRECONDITIONERS
from .spec import ProblemSpec

REQ_STATES = {0: "queued", 1: "active", 2: "done", 3: "failed"}


class Server:
    def __init__(self, ctx: Optional[Context] = None, precision="mixed", rtol=1e-10, max_iter=100000,
                 reliable_delta=0.1, item_slices=8, blocks_per_cu=0, hbm_budget: Optional[int] = None,
                 max_active=256, iters_per_step=64, preconditioner="jacobi", mg_workers=4):
        self.ctx = ctx or default_context()
        lib = self.ctx._lib
        o = _lib.ServerOpts()
        lib.sfem_server_opts_default(C.byref(o))
        o.precision = PRECISIONS[precision]
        o.rtol = float(rtol)
        o.max_iter = int(max_iter)
        o.reliable_delta = float(reliable_delta)
        o.item_slices = int(item_slices)
        o.blocks_per_cu = int(blocks_per_cu)
        if hbm_budget is not None:
            o.hbm_budget = int(hbm_budget)
        o.max_active = int(max_active)
        o.iters_per_step = int(iters_per_step)
        o.preconditioner = PRECONDITIONERS[preconditioner]
        o.mg_workers = int(mg_workers)
        h = C.c_void_p()
        self.ctx.check(lib.sfem_server_create(self.ctx.handle, C.byref(o), C.byref(h)))
        self.handle = h
        self.opts = {name: getattr(o, name) for name, _ in _lib.ServerOpts._fields_}

    def submit(self, p) -> int:
        """Queue a Problem (host arrays) or a ProblemSpec (generated on the device)."""
        if isinstance(p, ProblemSpec):
            rid = C.c_int64()
            sp = p.to_c()
            self.ctx.check(self.ctx._lib.sfem_server_submit_spec(self.handle, C.byref(sp), C.byref(rid)))
            return rid.value
        m = p.mesh
        nodes = np.asfortranarray(m.NodeList, dtype=np.float64)
        ien = np.asfortranarray(m.IEN, dtype=np.int64)
        young = np.ascontiguousarray(p.young, dtype=np.float64)
        nu = np.ascontiguousarray(p.nu, dtype=np.float64)
        f = np.ascontiguousarray(p.f, dtype=np.float64)
        fixed = np.ascontiguousarray(p.fixed, dtype=np.uint8)
        qd = None if p.q_d is None else np.ascontiguousarray(p.q_d, dtype=np.float64)
        rid = C.c_int64()
        args = (_ptr(nodes, C.c_double), m.n_nodes, _ptr(ien, C.c_int64), m.n_elem, _ptr(young, C.c_double),
                _ptr(nu, C.c_double), _ptr(f, C.c_double), _ptr(fixed, C.c_uint8),
                None if qd is None else _ptr(qd, C.c_double), C.byref(rid))
        if self.opts["preconditioner"] == PRECONDITIONERS["multigrid"] and getattr(m, "dims", None):
            # structured mesh: solved on its own with the multigrid preconditioner
            dims = np.ascontiguousarray(m.dims, dtype=np.int64)
            self.ctx.check(self.ctx._lib.sfem_server_submit_grid(self.handle, m.ndim, _ptr(dims, C.c_int64),
                                                                 *args))
        else:
            self.ctx.check(self.ctx._lib.sfem_server_submit(self.handle, m.ndim, *args))
        return rid.value

    def step(self):
        """One admit / iterate / retire step: (requests finished, launch ms)."""
        n = C.c_int()
        ms = C.c_double()
        self.ctx.check(self.ctx._lib.sfem_server_step(self.handle, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def poll(self, rid: int) -> dict:
        state = C.c_int()
        st = _lib.BatchProblemStats()
        t = (C.c_double * 3)()
        self.ctx.check(self.ctx._lib.sfem_server_poll(self.handle, int(rid), C.byref(state), C.byref(st), t))
        out = {name: getattr(st, name) for name, _ in _lib.BatchProblemStats._fields_}
        out.update(state=REQ_STATES[state.value], t_submit_ms=t[0], t_admit_ms=t[1], t_finish_ms=t[2])
        return out

    def take(self, rid: int, n_dof: int) -> np.ndarray:
        q = np.zeros(n_dof)
        self.ctx.check(self.ctx._lib.sfem_server_take(self.handle, int(rid), _ptr(q, C.c_double)))
        return q

    def info(self) -> dict:
        q, a, b = C.c_int64(), C.c_int64(), C.c_int64()
        self.ctx.check(self.ctx._lib.sfem_server_info(self.handle, C.byref(q), C.byref(a), C.byref(b)))
        return {"queued": q.value, "active": a.value, "bytes": b.value}

    def drain(self) -> None:
        """Step until nothing is queued or active."""
        while True:
            i = self.info()
            if not i["queued"] and not i["active"]:
                return
            self.step()

    def close(self) -> None:
        if self.handle:
            self.ctx._lib.sfem_server_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        if _lib.finalizing():  # the HIP runtime may be gone: leave it to process exit
            return
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------- workloads --
def _dims_for(ndof: float, ndim: int, rng) -> tuple:
    """Structured grid dims with about ndof free dofs and a random aspect."""
    cells = max(ndof / ndim, 4.0)
    aspect = rng.uniform(1.0, 4.0)
    if ndim == 2:
        ny = max(2, int(round(np.sqrt(cells / aspect))))
        nx = max(2, int(round(cells / ny)))
        return (nx, ny)
    s = max(2, int(round((cells / aspect) ** (1.0 / 3.0))))
    nx = max(2, int(round(cells / (s * s))))
    return (nx, s, s)


def random_request(rng, dof_lo=1e3, dof_hi=1e6, p3d=0.5) -> Problem:
    """A Q4 or Hex8 cantilever with DOF count log-uniform in [dof_lo, dof_hi],
    per-element E log-uniform in [0.1, 10], nu = 0.3, x = 0 clamped and
    iid N(0, 1) nodal loads on the free dofs of the x = L face."""
    ndim = 3 if rng.uniform() < p3d else 2
    ndof = 10.0 ** rng.uniform(np.log10(dof_lo), np.log10(dof_hi))
    dims = _dims_for(ndof, ndim, rng)
    m = structured_mesh(dims, lengths=tuple(d / max(dims) for d in dims))
    fixed = clamp_x0(m)
    gi = m.node_grid_index()
    on = np.repeat(gi[0] == dims[0], ndim)
    f = np.where(on & (fixed == 0), rng.normal(0.0, 1.0, m.n_dof), 0.0)
    young = 10.0 ** rng.uniform(-1.0, 1.0, m.n_elem)
    return make_problem(m, young, 0.3, f, fixed)


@dataclass
class StreamResult:
    latencies_ms: List[float] = field(default_factory=list)
    dofs: List[int] = field(default_factory=list)
    iterations: List[int] = field(default_factory=list)
    wall_s: float = 0.0
    steps: int = 0
    launch_ms: float = 0.0
    failed: int = 0
    at_floor: int = 0
    rel_residuals: List[float] = field(default_factory=list)
    max_active: int = 0
    errors: List[str] = field(default_factory=list)

    def summary(self) -> dict:
        lat = np.array(self.latencies_ms) if self.latencies_ms else np.array([np.nan])
        return {
            "requests": len(self.latencies_ms), "failed": self.failed, "at_floor": self.at_floor,
            "max_rel_residual": float(max(self.rel_residuals)) if self.rel_residuals else None,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "wall_s": self.wall_s,
            "solved_dof_per_s": float(np.sum(self.dofs) / self.wall_s) if self.wall_s > 0 else 0.0,
            "requests_per_s": len(self.latencies_ms) / self.wall_s if self.wall_s > 0 else 0.0,
            "steps": self.steps, "launch_ms": self.launch_ms, "max_active": self.max_active,
            # device work over wall time (above 1: requests solved concurrently by the workers)
            "busy_ratio": self.launch_ms / (1e3 * self.wall_s) if self.wall_s > 0 else 0.0,
            "errors": self.errors[:5],
        }


def run_poisson(server: Server, requests: Sequence, rate_per_s: float, seed: int,
                check=None) -> StreamResult:
    """Submit ``requests`` (Problems or ProblemSpecs) at Poisson arrival
    times (rate per second, wall clock) while stepping the server.  Latency
    = finish - arrival: the wait from the arrival to the submission on the
    host plus the server's own submit-to-finish time (stamped by whichever
    thread finished the request), so polling does not add to it."""
    rng = np.random.Generator(np.random.PCG64(seed))
    arrivals = np.cumsum(rng.exponential(1.0 / rate_per_s, len(requests)))
    res = StreamResult()
    pending: Dict[int, tuple] = {}
    t0 = time.perf_counter()
    nxt = 0
    while nxt < len(requests) or pending:
        now = time.perf_counter() - t0
        while nxt < len(requests) and arrivals[nxt] <= now:
            rid = server.submit(requests[nxt])
            pending[rid] = (nxt, time.perf_counter() - t0 - arrivals[nxt])
            nxt += 1
        info = server.info()
        res.max_active = max(res.max_active, info["active"])
        if not info["queued"] and not info["active"]:
            if nxt < len(requests):
                time.sleep(max(0.0, min(arrivals[nxt] - now, 0.01)))
            continue
        _, ms = server.step()
        res.steps += 1
        res.launch_ms += ms
        for rid in list(pending):
            p = server.poll(rid)
            if p["state"] in ("done", "failed"):
                idx, sub_delay = pending.pop(rid)
                # solved: converged to rtol (1), or stopped at the FP64
                # rounding floor of the system (3) when rtol lies below it
                if p["state"] == "failed" or p["stop_reason"] not in (1, 3):
                    res.failed += 1
                    try:
                        server.take(rid, requests[idx].n_dof)
                        res.errors.append(f"request {idx}: stop reason {p['stop_reason']}")
                    except Exception as e:  # noqa: BLE001 - the server's message for the failure
                        res.errors.append(f"request {idx}: {e}")
                    continue
                res.at_floor += p["stop_reason"] == 3
                res.rel_residuals.append(p["rel_residual"])
                q = server.take(rid, requests[idx].n_dof)
                if check is not None:
                    check(idx, q, p)
                res.latencies_ms.append(1e3 * sub_delay + (p["t_finish_ms"] - p["t_submit_ms"]))
                res.dofs.append(int(p["n_free"]))
                res.iterations.append(int(p["iterations"]))
    res.wall_s = time.perf_counter() - t0
    return res
