"""The reference example's boundary machinery and displacement sweep on the GPU.

Mirrors examples/vector3D.jl (and src/PostProcess.jl:30-46):

* ``meshgrid``                  -- :10-130, on the device (any nx, ny, nz), with
                                   the bottom / top face lists and the cylinder
                                   map of ``inflate_sphere``;
* ``apply_boundary_conditions`` -- :175-264, the surface penalty b;
* ``csr_add``                   -- K_bar = K + beta*b, :308;
* ``bc_planes`` / ``setboundaryCond`` -- :133-173, Dirichlet data from
                                   coordinate planes;
* ``displacement_sweep``        -- :310-322, one K_free, many prescribed
                                   displacements, each solve started from the
                                   previous solution.

Everything stays in HBM; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from .fem import Context, Mesh, StiffnessMatrix, _ptr, default_context
from .solve import BoundaryConditions, System

INFLATE = 1


class DeviceMesh(Mesh):
    """A mesh generated on the device (``sfem_meshgrid``) or a surface mesh
    of face lists (``sfem_surface_mesh_create``)."""

    def __init__(self, ctx: Context, handle: C.c_void_p):  # noqa: super().__init__ uploads host arrays
        self.ctx = ctx
        self.handle = handle
        ndim, nen, ndof = C.c_int(), C.c_int(), C.c_int()
        nn, ne = C.c_int64(), C.c_int64()
        ctx.check(ctx._lib.sfem_mesh_info(handle, C.byref(ndim), C.byref(nen), C.byref(ndof), C.byref(nn),
                                          C.byref(ne)))
        self.ndim, self.nen, self.n_dof = ndim.value, nen.value, ndof.value
        self.n_nodes, self.n_elem = nn.value, ne.value

    def export(self):
        """(NodeList ndim x nNodes, IEN nElem x nen, ID nNodes x nDof), 1-based."""
        nodes = np.zeros((self.ndim, self.n_nodes), order="F")
        ien = np.zeros((self.n_elem, self.nen), dtype=np.int64, order="F")
        idm = np.zeros((self.n_nodes, self.n_dof), dtype=np.int64, order="F")
        self.ctx.check(self.ctx._lib.sfem_mesh_export(self.ctx.handle, self.handle, _ptr(nodes, C.c_double),
                                                      _ptr(ien, C.c_int64), _ptr(idm, C.c_int64)))
        return nodes, ien, idm


def meshgrid(x0, x1, y0, y1, z0, z1, ne, ndim, *, inflate=False, Young=1.0, nu=0.3, faces=True,
             ctx: Optional[Context] = None):
    """Structured Q4 / Hex8 mesh on the device, as meshgrid(x0,x1,y0,y1,z0,z1,
    ne,ndim) builds it (``ne`` may be a tuple of per-direction counts).
    Returns (mesh, faces) with faces the bottom-then-top surface mesh (3D;
    None in 2D or with ``faces=False``).  ``inflate`` applies inflate_sphere
    to (x, y) -- in the example the cylinder is the mesh K is built on, since
    inflate_sphere overwrites NodeList (examples/vector3D.jl:292)."""
    ctx = ctx or default_context()
    dims = np.ascontiguousarray(np.broadcast_to(np.asarray(ne, dtype=np.int64), (ndim,)).copy())
    box = np.ascontiguousarray([x0, x1, y0, y1, z0, z1][: 2 * ndim], dtype=np.float64)
    young = np.ascontiguousarray(np.atleast_1d(np.asarray(Young, dtype=np.float64)))
    want_faces = faces and ndim == 3
    hm, hf = C.c_void_p(), C.c_void_p()
    ctx.check(ctx._lib.sfem_meshgrid(ctx.handle, int(ndim), _ptr(dims, C.c_int64), _ptr(box, C.c_double),
                                     INFLATE if inflate else 0, _ptr(young, C.c_double), young.size, float(nu),
                                     C.byref(hm), C.byref(hf) if want_faces else None))
    mesh = DeviceMesh(ctx, hm)
    mesh.box = tuple(float(v) for v in box)
    return mesh, (DeviceMesh(ctx, hf) if want_faces else None)


def surface_mesh(NodeList, IEN_top, IEN_btm, ID, ndim=3, nDof=3, ctx: Optional[Context] = None) -> DeviceMesh:
    """Upload a bottom / top face list pair as a surface mesh."""
    ctx = ctx or default_context()
    nodes = np.asfortranarray(np.asarray(NodeList, dtype=np.float64))
    top = np.asfortranarray(np.asarray(IEN_top, dtype=np.int64))
    btm = np.asfortranarray(np.asarray(IEN_btm, dtype=np.int64))
    idm = np.asfortranarray(np.asarray(ID, dtype=np.int64))
    if top.shape != btm.shape:
        raise ValueError("IEN_top and IEN_btm must have the same shape")
    h = C.c_void_p()
    ctx.check(ctx._lib.sfem_surface_mesh_create(
        ctx.handle, int(ndim), int(nDof), _ptr(nodes, C.c_double), nodes.shape[1], _ptr(btm, C.c_int64),
        _ptr(top, C.c_int64), btm.shape[0], btm.shape[1], _ptr(idm, C.c_int64), idm.shape[0], C.byref(h)))
    return DeviceMesh(ctx, h)


def apply_boundary_conditions(ne, NodeList, IEN, IEN_top, IEN_btm, ndim, FunctionClass="Q1", ID=None, nDof=3, *,
                              ctx: Optional[Context] = None, device: bool = False):
    """The surface penalty b of examples/vector3D.jl:175-264 on the GPU.
    Same signature (IEN and FunctionClass are unused there too).  Returns a
    scipy csc_matrix, or the device matrix with ``device=True``."""
    ctx = ctx or default_context()
    m = surface_mesh(NodeList, IEN_top, IEN_btm, ID, ndim, nDof, ctx)
    try:
        b = m.assemble()
    finally:
        m.close()
    if device:
        return b
    try:
        return b.to_scipy()
    finally:
        b.close()


def csr_add(K: StiffnessMatrix, beta: float, b: StiffnessMatrix) -> StiffnessMatrix:
    """K + beta*b on the device (union pattern; exact-zero sums stay stored)."""
    h = C.c_void_p()
    K.ctx.check(K.ctx._lib.sfem_csr_add(K.ctx.handle, K.handle, float(beta), b.handle, C.byref(h)))
    return StiffnessMatrix(K.ctx, h)


def face_traction_loads(faces: DeviceMesh, traction) -> np.ndarray:
    """Consistent nodal loads of a uniform traction t on a face mesh:
    f[3(m-1)+i] = t_i * int N_m dA over the faces, formed on the device as
    b (1 (x) t) with b the faces' penalty matrix (sum_b N_b = 1).  For one
    loaded face, build the face mesh from that face list only."""
    t = np.asarray(traction, dtype=np.float64).ravel()
    if t.size != faces.n_dof:
        raise ValueError(f"traction needs {faces.n_dof} components")
    b = faces.assemble()
    try:
        return b.apply(np.tile(t, b.shape[1] // t.size))
    finally:
        b.close()


@dataclass
class PlaneRule:
    """Fix dof ``dof`` (1-based) of the nodes with coordinate ``axis``
    (1-based) equal to ``coord``, to ``value``."""
    axis: int
    coord: float
    dof: int
    value: float = 0.0


def bc_planes(mesh: DeviceMesh, n_total: int, rules: Sequence[PlaneRule], f=None) -> BoundaryConditions:
    """Dirichlet data from plane rules (first matching rule per node and dof)."""
    ctx = mesh.ctx
    ax = np.ascontiguousarray([r.axis for r in rules], dtype=np.int32)
    co = np.ascontiguousarray([r.coord for r in rules], dtype=np.float64)
    dd = np.ascontiguousarray([r.dof for r in rules], dtype=np.int32)
    va = np.ascontiguousarray([r.value for r in rules], dtype=np.float64)
    fv = None if f is None else np.ascontiguousarray(f, dtype=np.float64)
    if fv is not None and fv.size != n_total:
        raise ValueError(f"f has {fv.size} entries, n_total = {n_total}")
    h = C.c_void_p()
    ctx.check(ctx._lib.sfem_bc_planes(ctx.handle, mesh.handle, int(n_total), len(rules), _ptr(ax, C.c_int),
                                      _ptr(co, C.c_double), _ptr(dd, C.c_int), _ptr(va, C.c_double),
                                      None if fv is None else _ptr(fv, C.c_double), C.byref(h)))
    return BoundaryConditions._wrap(ctx, h, n_total)


def setboundaryCond(mesh: DeviceMesh, d: float, z0: float = 0.0, z1: float = 1.0) -> BoundaryConditions:
    """examples/vector3D.jl:133-173: u_z = 0 on z == z0, u_z = -d on z == z1."""
    return bc_planes(mesh, mesh.n_nodes * mesh.n_dof, [PlaneRule(3, z0, 3, 0.0), PlaneRule(3, z1, 3, -float(d))])


@dataclass
class SweepResult:
    deltas: List[float]
    q: List[np.ndarray] = field(default_factory=list)
    stats: List[dict] = field(default_factory=list)
    build_ms: float = 0.0
    solve_ms: float = 0.0
    n_free: int = 0


def displacement_sweep(ne, deltas, *, Young=40.0, nu=0.4, beta=100.0, precision="fp64", rtol=1e-10,
                       preconditioner="jacobi", warm_start=True, keep_solutions=True, together=False,
                       ctx: Optional[Context] = None) -> SweepResult:
    """The example's solve loop (examples/vector3D.jl:270-322) on the GPU:
    unit cube, ``ne`` elements per direction, inflated to the cylinder, K on
    the cylinder, b on its bottom / top faces, K_bar = K + beta*b, then for
    each d: u_z = 0 at z = 0, u_z = -d at z = 1, q = q_d + C K_free^-1 C'(-K_bar q_d).
    The fixed set does not depend on d, so K_free is built once; each d only
    replaces the rhs (``System.set_bc``) and, with ``warm_start``, starts
    from the previous solution.  ``together`` solves the displacements at
    once instead (up to 64 per pass over K_free, FP64 point Jacobi).
    ``preconditioner="multigrid"`` builds the geometric multigrid hierarchy
    of the ne^3 grid once (it depends on K_bar and the fixed set only) and
    preconditions every solve with it."""
    import time

    ctx = ctx or default_context()
    res = SweepResult(deltas=[float(d) for d in deltas])
    t0 = time.perf_counter()
    mesh, faces = meshgrid(0, 1, 0, 1, 0, 1, ne, 3, inflate=True, Young=Young, nu=nu, ctx=ctx)
    K = mesh.assemble()
    b = faces.assemble()
    Kbar = csr_add(K, beta, b)
    K.close()
    b.close()
    faces.close()
    bc = setboundaryCond(mesh, res.deltas[0])
    S = System(Kbar, bc)
    bc.close()
    if preconditioner == "multigrid" and not together:
        S.set_grid((ne, ne, ne))
    ctx.synchronize()
    res.build_ms = (time.perf_counter() - t0) * 1e3
    res.n_free = S.n_free
    try:
        if together:
            nodes, _, _ = mesh.export()
            top = nodes[2] == 1.0
            for g0 in range(0, len(res.deltas), 64):
                group = res.deltas[g0:g0 + 64]
                qd = np.zeros((S.n, len(group)))
                for j, d in enumerate(group):
                    qd[2::3, j][top] = -d
                q, per, ms = S.solve_many(q_d=qd, rtol=rtol)
                res.solve_ms += ms
                res.stats += per
                if keep_solutions:
                    res.q += [q[:, j].copy() for j in range(len(group))]
            return res
        for i, d in enumerate(res.deltas):
            if i > 0:
                bc = setboundaryCond(mesh, d)
                S.set_bc(bc)
                bc.close()
            st = S.solve(precision, rtol=rtol, x0_zero=not (warm_start and i > 0), preconditioner=preconditioner)
            res.solve_ms += st["solve_ms"]
            res.stats.append(st)
            if keep_solutions:
                res.q.append(S.solution())
    finally:
        S.close()
        Kbar.close()
        mesh.close()
    return res
