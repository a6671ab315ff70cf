"""Structured Q4 / Hex8 meshes in the reference's numbering, plus the loads
and clamps of the benchmark configurations.

Numbering follows ``meshgrid`` (smearFEM examples/vector3D.jl:10-130),
generalised from ne^ndim cubes to nx x ny (x nz) boxes:

* nodes x-fastest, then y, then z (:28-40, :67-86), 1-based;
* Q4 IEN (:45-48) and Hex8 IEN (:94-101), counter-clockwise bottom then top;
* ID[m, l] = ndim*(m-1) + l (:33, :74), i.e. dofs interleaved per node.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np


@dataclass
class StructuredMesh:
    ndim: int
    dims: tuple          # elements per direction
    lengths: tuple       # box edge lengths
    NodeList: np.ndarray  # ndim x nNodes
    IEN: np.ndarray       # nElem x 2^ndim, 1-based
    ID: np.ndarray        # nNodes x ndim, 1-based

    @property
    def n_nodes(self) -> int:
        return self.NodeList.shape[1]

    @property
    def n_elem(self) -> int:
        return self.IEN.shape[0]

    @property
    def n_dof(self) -> int:
        return self.ndim * self.n_nodes

    def node_grid_index(self):
        """0-based (i, j[, k]) grid index of every node."""
        nx1 = self.dims[0] + 1
        m = np.arange(self.n_nodes)
        if self.ndim == 2:
            return m % nx1, m // nx1
        ny1 = self.dims[1] + 1
        return m % nx1, (m // nx1) % ny1, m // (nx1 * ny1)


def structured_mesh(dims: Sequence[int], lengths: Optional[Sequence[float]] = None,
                    origin: Optional[Sequence[float]] = None) -> StructuredMesh:
    dims = tuple(int(d) for d in dims)
    ndim = len(dims)
    if ndim not in (2, 3):
        raise ValueError("dims must have 2 or 3 entries")
    lengths = tuple(float(v) for v in (lengths or [1.0] * ndim))
    origin = tuple(float(v) for v in (origin or [0.0] * ndim))
    axes = [np.linspace(origin[d], origin[d] + lengths[d], dims[d] + 1) for d in range(ndim)]
    if ndim == 2:
        nx, ny = dims
        X, Y = np.meshgrid(axes[0], axes[1], indexing="xy")  # shape (ny+1, nx+1), x fastest
        NodeList = np.vstack([X.ravel(), Y.ravel()])
        j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
        i = i.ravel() + 1
        j = j.ravel() + 1
        IEN = np.stack([(j - 1) * (nx + 1) + i, (j - 1) * (nx + 1) + i + 1, j * (nx + 1) + i + 1,
                        j * (nx + 1) + i], axis=1)
    else:
        nx, ny, nz = dims
        Z, Y, X = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
        NodeList = np.vstack([X.ravel(), Y.ravel(), Z.ravel()])
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        i = i.ravel() + 1
        j = j.ravel() + 1
        k = k.ravel() + 1
        p = (nx + 1) * (ny + 1)
        IEN = np.stack([(k - 1) * p + (j - 1) * (nx + 1) + i, (k - 1) * p + (j - 1) * (nx + 1) + i + 1,
                        (k - 1) * p + j * (nx + 1) + i + 1, (k - 1) * p + j * (nx + 1) + i,
                        k * p + (j - 1) * (nx + 1) + i, k * p + (j - 1) * (nx + 1) + i + 1,
                        k * p + j * (nx + 1) + i + 1, k * p + j * (nx + 1) + i], axis=1)
    n_nodes = NodeList.shape[1]
    m = np.arange(1, n_nodes + 1)
    ID = np.stack([ndim * (m - 1) + l for l in range(1, ndim + 1)], axis=1)
    return StructuredMesh(ndim, dims, lengths, np.ascontiguousarray(NodeList),
                          np.ascontiguousarray(IEN.astype(np.int64)), np.ascontiguousarray(ID.astype(np.int64)))


# ------------------------------------------------------ boundary data --
def clamp_x0(mesh: StructuredMesh) -> np.ndarray:
    """Fixed-dof mask (uint8, n_dof) of every dof of the nodes on x = x_min."""
    gi = mesh.node_grid_index()
    fixed_nodes = gi[0] == 0
    mask = np.repeat(fixed_nodes, mesh.ndim).astype(np.uint8)
    return mask


def face_load(mesh: StructuredMesh, axis: int, side: int, direction: int, total: float) -> np.ndarray:
    """Consistent nodal loads of a uniform traction on the face x_axis = min/max.

    Bilinear face elements (2D: linear edges) give each node the share of the
    face area its basis function integrates to; the loads sum to ``total``
    along ``direction``.
    """
    gi = mesh.node_grid_index()
    end = 0 if side == 0 else mesh.dims[axis]
    on = gi[axis] == end
    w = np.ones(mesh.n_nodes)
    for d in range(mesh.ndim):
        if d == axis:
            continue
        nd = mesh.dims[d]
        edge = (gi[d] == 0) | (gi[d] == nd)
        w = w * np.where(edge, 0.5, 1.0) / nd
    w = np.where(on, w, 0.0)
    f = np.zeros(mesh.n_dof)
    f[direction::mesh.ndim] = total * w / w.sum()
    return f


def random_face_load(mesh: StructuredMesh, axis: int, side: int, fixed: np.ndarray, seed: int,
                     sigma: float = 1.0) -> np.ndarray:
    """iid N(0, sigma^2) nodal loads on every component of the free nodes of a face."""
    gi = mesh.node_grid_index()
    end = 0 if side == 0 else mesh.dims[axis]
    on = np.repeat(gi[axis] == end, mesh.ndim)
    rng = np.random.Generator(np.random.PCG64(seed))
    f = rng.normal(0.0, sigma, mesh.n_dof)
    return np.where(on & (fixed == 0), f, 0.0)


def element_young_loguniform(n_elem: int, lo: float, hi: float, seed: int) -> np.ndarray:
    """Per-element E = 10^U(log10 lo, log10 hi) from numpy PCG64(seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return 10.0 ** rng.uniform(np.log10(lo), np.log10(hi), n_elem)
