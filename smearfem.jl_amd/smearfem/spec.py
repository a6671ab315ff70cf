"""Problems generated on the device from a compact spec (``sfem_problem_spec``).

The reference builds its meshes and constraints on the host (meshgrid,
examples/vector3D.jl:10-130; setboundaryCond, :133-173) and passes loads as
arrays.  A ``ProblemSpec`` names a structured Q4 / Hex8 elasticity problem
instead -- the box and element counts, a log-uniform Young's modulus field
drawn from a seed, Poisson's ratio, the clamped face, and consistent loads
(uniform traction on a face, a body force, seeded N(0, sigma^2) nodal loads on
the loaded face) -- and the device forms the mesh, the moduli, the loads and
the constraints in HBM (``sfem_spec_build``).  Servers (``submit_spec``) and
batches (``Batch.from_specs``) take specs, so nothing is prepared or copied on
the host per request.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .fem import Context, _ptr, default_context


@dataclass
class ProblemSpec:
    ndim: int
    dims: Tuple[int, ...]
    box: Tuple[float, ...]
    young: Tuple[float, float] = (1.0, 1.0)   # (lo, hi): log-uniform per element; lo == hi: constant
    nu: float = 0.3
    seed: int = 0
    clamp: Tuple[int, int] = (0, 0)          # (axis, side): every dof of that face fixed
    load: Tuple[int, int] = (0, 1)           # (axis, side) of the loaded face
    traction: Tuple[float, ...] = (0.0, 0.0, 0.0)
    body: Tuple[float, ...] = (0.0, 0.0, 0.0)
    noise: float = 0.0

    @property
    def n_nodes(self) -> int:
        return int(np.prod([d + 1 for d in self.dims[: self.ndim]]))

    @property
    def n_elem(self) -> int:
        return int(np.prod(self.dims[: self.ndim]))

    @property
    def n_dof(self) -> int:
        return self.ndim * self.n_nodes

    @property
    def n_free(self) -> int:
        ax = self.clamp[0]
        face = int(np.prod([self.dims[d] + 1 for d in range(self.ndim) if d != ax]))
        return self.ndim * (self.n_nodes - face)

    def as_dict(self) -> dict:
        """The fields of sfem_problem_spec (the oracle's restatement reads this)."""
        pad = lambda v, n: tuple(v) + (0,) * (n - len(v))  # noqa: E731
        return {"ndim": self.ndim, "dims": pad(self.dims, 3), "box": pad(self.box, 6),
                "young_lo": float(self.young[0]), "young_hi": float(self.young[1]), "nu": float(self.nu),
                "seed": int(self.seed), "clamp_axis": self.clamp[0], "clamp_side": self.clamp[1],
                "load_axis": self.load[0], "load_side": self.load[1], "traction": pad(self.traction, 3),
                "body": pad(self.body, 3), "noise": float(self.noise)}

    def to_c(self) -> _lib.ProblemSpec:
        d = self.as_dict()
        s = _lib.ProblemSpec()
        s.ndim = d["ndim"]
        for k in range(3):
            s.dims[k] = int(d["dims"][k])
            s.traction[k] = float(d["traction"][k])
            s.body[k] = float(d["body"][k])
        for k in range(6):
            s.box[k] = float(d["box"][k])
        s.young_lo, s.young_hi, s.nu = d["young_lo"], d["young_hi"], d["nu"]
        s.seed = d["seed"] & (2 ** 64 - 1)
        s.clamp_axis, s.clamp_side = d["clamp_axis"], d["clamp_side"]
        s.load_axis, s.load_side = d["load_axis"], d["load_side"]
        s.noise = d["noise"]
        return s

    def export(self, ctx: Optional[Context] = None) -> dict:
        """The device-generated inputs copied to the host: NodeList (ndim x
        n_nodes), young (per element), f and fixed (per dof)."""
        ctx = ctx or default_context()
        nodes = np.zeros((self.ndim, self.n_nodes), order="F")
        young = np.zeros(self.n_elem)
        f = np.zeros(self.n_dof)
        fixed = np.zeros(self.n_dof, dtype=np.uint8)
        sp = self.to_c()
        ctx.check(ctx._lib.sfem_spec_export(ctx.handle, C.byref(sp), _ptr(nodes, C.c_double),
                                            _ptr(young, C.c_double), _ptr(f, C.c_double),
                                            _ptr(fixed, C.c_uint8)))
        return {"NodeList": nodes, "young": young, "f": f, "fixed": fixed}

    def build(self, ctx: Optional[Context] = None):
        """(Mesh, BoundaryConditions) on the device (``sfem_spec_build``); the
        mesh is declared its grid, so it assembles from the grid."""
        from .solve import BoundaryConditions
        from .surface import DeviceMesh

        ctx = ctx or default_context()
        m, b = C.c_void_p(), C.c_void_p()
        sp = self.to_c()
        ctx.check(ctx._lib.sfem_spec_build(ctx.handle, C.byref(sp), C.byref(m), C.byref(b)))
        return DeviceMesh(ctx, m), BoundaryConditions._wrap(ctx, b, self.n_dof)


def cantilever_spec(dims, lengths=None, young=(1.0, 1.0), nu=0.3, seed=0, end_load=None, noise=0.0,
                    body=None) -> ProblemSpec:
    """x = 0 face clamped, loads on the x = L face: ``end_load`` is the total
    force there (a uniform traction of end_load / face area), ``noise`` the
    sigma of seeded N(0, sigma^2) nodal loads on its free dofs."""
    nd = len(dims)
    lengths = lengths or tuple(d / max(dims) for d in dims)
    box = tuple(v for L in lengths for v in (0.0, float(L)))
    trac = (0.0,) * nd
    if end_load is not None:
        area = float(np.prod(lengths[1:]))
        trac = tuple(float(t) / area for t in end_load)
    return ProblemSpec(nd, tuple(int(d) for d in dims), box, (float(young[0]), float(young[1])), nu, seed,
                       (0, 0), (0, 1), trac, tuple(body) if body is not None else (0.0,) * nd, noise)


def random_request_spec(rng, dof_lo=1e3, dof_hi=1e6, p3d=0.5) -> ProblemSpec:
    """The serving workload as a spec: a Q4 or Hex8 cantilever with DOF count
    log-uniform in [dof_lo, dof_hi], per-element E log-uniform in [0.1, 10],
    nu = 0.3, x = 0 clamped and N(0, 1) nodal loads on the free dofs of the
    x = L face."""
    from .serve import _dims_for

    ndim = 3 if rng.uniform() < p3d else 2
    ndof = 10.0 ** rng.uniform(np.log10(dof_lo), np.log10(dof_hi))
    dims = _dims_for(ndof, ndim, rng)
    seed = int(rng.integers(0, 2 ** 62))
    return cantilever_spec(dims, young=(0.1, 10.0), nu=0.3, seed=seed, noise=1.0)
