"""Python mirror of smearFEM's FEM entry points, backed by the gfx950 library.

Same names, argument meaning and error behaviour as the reference's
``src/fem.jl``:

* ``gaussian_quadrature(a, b, nGaussPoints=2)``       -- src/fem.jl:21
* ``basis_function(xi, eta=None, zeta=None, FunctionClass="Q1")`` -- :48
* ``assemble_system(ne, NodeList, IEN, ndim, FunctionClass="Q1", nDof=1,
  ID=None, Young=1, nu=0.3)``                         -- :135

Arrays use the reference's shapes: ``NodeList`` is ndim x nNodes, ``IEN`` is
nElements x nLocalNodes, ``ID`` is nNodes x nDof, all ids 1-based.  The
assembled matrix comes back as a ``scipy.sparse.csc_matrix`` (the analogue of
``SparseMatrixCSC{Float64,Int64}``), or as a device-resident
``StiffnessMatrix`` with ``device=True``.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from .errors import raise_for

FUNCTION_CLASSES = {"Q1": 1, "Q2": 2}


class Context:
    """A device, a HIP stream and a caching allocator (``sfem_ctx``)."""

    def __init__(self, device: int = 0):
        self._lib = _lib.load()
        h = C.c_void_p()
        st = self._lib.sfem_ctx_create(device, C.byref(h))
        raise_for(st, f"sfem_ctx_create(device={device}) failed (status {st})")
        self.handle = h
        self.device = device

    def check(self, status: int) -> None:
        if status:
            msg = self._lib.sfem_last_error(self.handle)
            raise_for(status, msg.decode() if msg else "")

    def device_info(self) -> dict:
        name = C.create_string_buffer(256)
        ncu = C.c_int()
        hbm = C.c_int64()
        self.check(self._lib.sfem_device_info(self.handle, name, 256, C.byref(ncu), C.byref(hbm)))
        return {"name": name.value.decode(), "n_cu": ncu.value, "hbm_bytes": hbm.value}

    def device_bytes(self) -> int:
        return int(self._lib.sfem_ctx_device_bytes(self.handle))

    def synchronize(self) -> None:
        self.check(self._lib.sfem_ctx_synchronize(self.handle))

    def close(self) -> None:
        if getattr(self, "handle", None):
            self._lib.sfem_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        if _lib.finalizing():  # the HIP runtime may be gone: leave it to process exit
            return
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None or _default_ctx.handle is None:
        _default_ctx = Context(0)
    return _default_ctx


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


# ------------------------------------------------------------ utilities --
def gaussian_quadrature(a, b, nGaussPoints=2):
    """Nodes and weights of the 2- or 3-point Gauss rule on [a, b]."""
    lib = _lib.load()
    n = int(nGaussPoints)
    xi = np.zeros(max(n, 1))
    w = np.zeros(max(n, 1))
    raise_for(lib.sfem_gaussian_quadrature(float(a), float(b), n, _ptr(xi, C.c_double), _ptr(w, C.c_double)),
              f"gaussian_quadrature: nGaussPoints={n} not in (2, 3)")
    return list(xi[:n]), list(w[:n])


def basis_function(xi, eta=None, zeta=None, FunctionClass="Q1"):
    """(N, dN) of the master element; dN is nfun x ndim (1D: the 1x2 row)."""
    lib = _lib.load()
    nargs = 1 if eta is None else (2 if zeta is None else 3)
    cls = FUNCTION_CLASSES.get(FunctionClass, 0)
    N = np.zeros(9)
    dN = np.zeros(27)
    nf, nr, ncol = C.c_int(), C.c_int(), C.c_int()
    st = lib.sfem_basis_function(nargs, float(xi), float(eta or 0.0), float(zeta or 0.0), cls,
                                 _ptr(N, C.c_double), _ptr(dN, C.c_double), C.byref(nf), C.byref(nr),
                                 C.byref(ncol))
    raise_for(st, f"basis_function: FunctionClass={FunctionClass!r} with {nargs} coordinates")
    return N[: nf.value].copy(), dN[: nr.value * ncol.value].reshape(ncol.value, nr.value).T.copy()


# ------------------------------------------------------------- assembly --
class StiffnessMatrix:
    """Device-resident assembled K (``sfem_csr``)."""

    def __init__(self, ctx: Context, handle: C.c_void_p):
        self.ctx = ctx
        self.handle = handle
        m, n, nnz, bs = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int()
        ctx.check(ctx._lib.sfem_csr_info(handle, C.byref(m), C.byref(n), C.byref(nnz), C.byref(bs)))
        self.shape = (m.value, n.value)
        self.nnz = nnz.value
        self.block = bs.value

    @property
    def grid(self):
        """Elements per direction of the grid K was assembled on
        (``sfem_csr_grid``), or None."""
        nd = C.c_int()
        ne = np.zeros(3, dtype=np.int64)
        self.ctx.check(self.ctx._lib.sfem_csr_grid(self.handle, C.byref(nd), _ptr(ne, C.c_int64)))
        return tuple(int(v) for v in ne[: nd.value]) if nd.value else None

    def grid_values(self) -> dict:
        """How K's values were formed on its grid (``sfem_csr_grid_values``):
        ``from_scales`` (summed from s_e Ke(0)) and the measured
        ``ke_deviation``."""
        fs, dev = C.c_int(), C.c_double()
        self.ctx.check(self.ctx._lib.sfem_csr_grid_values(self.handle, C.byref(fs), C.byref(dev)))
        return {"from_scales": bool(fs.value), "ke_deviation": dev.value}

    def stage_ms(self) -> dict:
        ms = np.zeros(6)
        self.ctx.check(self.ctx._lib.sfem_csr_stage_ms(self.handle, _ptr(ms, C.c_double)))
        return dict(zip(("validate", "element", "keys", "sort", "reduce", "total"), ms.tolist()))

    def csc_arrays(self):
        """1-based (colptr, rowval, nzval) as SparseMatrixCSC holds them."""
        n = self.shape[1]
        colptr = np.zeros(n + 1, dtype=np.int64)
        rowval = np.zeros(self.nnz, dtype=np.int64)
        nzval = np.zeros(self.nnz, dtype=np.float64)
        self.ctx.check(self.ctx._lib.sfem_export_csc1(self.ctx.handle, self.handle, _ptr(colptr, C.c_int64),
                                                      _ptr(rowval, C.c_int64), _ptr(nzval, C.c_double)))
        return colptr, rowval, nzval

    def apply(self, x) -> np.ndarray:
        """y = K x on the device (``sfem_csr_apply``)."""
        xv = np.ascontiguousarray(x, dtype=np.float64)
        if xv.size != self.shape[1]:
            raise ValueError(f"x has {xv.size} entries, K has {self.shape[1]} columns")
        y = np.zeros(self.shape[0])
        self.ctx.check(self.ctx._lib.sfem_csr_apply(self.ctx.handle, self.handle, _ptr(xv, C.c_double),
                                                    _ptr(y, C.c_double)))
        return y

    def to_scipy(self):
        import scipy.sparse as sp

        colptr, rowval, nzval = self.csc_arrays()
        return sp.csc_matrix((nzval, rowval - 1, colptr - 1), shape=self.shape)

    def close(self) -> None:
        if self.handle:
            self.ctx._lib.sfem_csr_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        if _lib.finalizing():  # the HIP runtime may be gone: leave it to process exit
            return
        try:
            self.close()
        except Exception:
            pass


def _mesh_arrays(NodeList, IEN, ID, Young, n_elem):
    nodes = np.asfortranarray(np.asarray(NodeList, dtype=np.float64))
    if nodes.ndim != 2:
        raise ValueError("NodeList must be ndim x nNodes")
    ien = np.asfortranarray(np.asarray(IEN, dtype=np.int64))
    if ien.ndim != 2:
        raise ValueError("IEN must be nElements x nLocalNodes")
    idm = None if ID is None else np.asfortranarray(np.asarray(ID, dtype=np.int64))
    young = np.ascontiguousarray(np.atleast_1d(np.asarray(Young, dtype=np.float64)))
    return nodes, ien, idm, young


class Mesh:
    """Mesh and material inputs uploaded to HBM once (``sfem_mesh``)."""

    def __init__(self, ctx: Context, n_elem, NodeList, IEN, ndim, FunctionClass="Q1", nDof=1, ID=None,
                 Young=1, nu=0.3):
        self.ctx = ctx
        nodes, ien, idm, young = _mesh_arrays(NodeList, IEN, ID, Young, n_elem)
        cls = FUNCTION_CLASSES.get(FunctionClass, 0)
        h = C.c_void_p()
        idp = _ptr(idm, C.c_int64) if idm is not None else None
        id_rows = idm.shape[0] if idm is not None else 0
        ctx.check(ctx._lib.sfem_mesh_create(
            ctx.handle, int(ndim), cls, int(nDof), _ptr(nodes, C.c_double), nodes.shape[1],
            _ptr(ien, C.c_int64), ien.shape[0], ien.shape[1], int(n_elem), idp, id_rows,
            _ptr(young, C.c_double), young.size, float(nu), C.byref(h)))
        self.handle = h

    def set_grid(self, dims) -> None:
        """Declare the mesh the structured grid of the meshgrid numbering with
        ``dims`` elements per direction (``sfem_mesh_set_grid``, checked on
        the device); assembly then takes node pairs from the grid instead of
        sorting keys: the same pattern, and the sorted path's values bit for
        bit, or, when every element matrix is s_e Ke(0) to within 1e-13,
        values summed from the scales (``assemble(exact_sums=True)`` keeps
        the element sums)."""
        d = np.ascontiguousarray(dims, dtype=np.int64).ravel()
        self.ctx.check(self.ctx._lib.sfem_mesh_set_grid(self.ctx.handle, self.handle, int(d.size),
                                                        _ptr(d, C.c_int64)))

    def detect_grid(self):
        """Declare the mesh a structured grid if it is one, with the element
        counts read from element 1's corners (``sfem_mesh_detect_grid``).
        Returns the element counts, or None (the mesh is then unchanged)."""
        nd = C.c_int()
        ne = np.zeros(3, dtype=np.int64)
        self.ctx.check(self.ctx._lib.sfem_mesh_detect_grid(self.ctx.handle, self.handle, C.byref(nd),
                                                           _ptr(ne, C.c_int64)))
        return tuple(int(v) for v in ne[: nd.value]) if nd.value else None

    def assemble(self, force_scalar: bool = False, exact_sums: bool = False) -> StiffnessMatrix:
        """``sfem_assemble_mesh``; ``exact_sums``: on a declared grid, K from
        the element matrices even when they are all s_e Ke(0)
        (``SFEM_ASM_EXACT_SUMS``)."""
        h = C.c_void_p()
        flags = (1 if force_scalar else 0) | (2 if exact_sums else 0)
        self.ctx.check(self.ctx._lib.sfem_assemble_mesh(self.ctx.handle, self.handle, flags, C.byref(h)))
        return StiffnessMatrix(self.ctx, h)

    def close(self) -> None:
        if self.handle:
            self.ctx._lib.sfem_mesh_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        if _lib.finalizing():  # the HIP runtime may be gone: leave it to process exit
            return
        try:
            self.close()
        except Exception:
            pass


def assemble_system(ne, NodeList, IEN, ndim, FunctionClass="Q1", nDof=1, ID=None, Young=1, nu=0.3, *,
                    ctx: Optional[Context] = None, device: bool = False, force_scalar: bool = False,
                    n_elem: Optional[int] = None, detect_grid: bool = True, exact_sums: bool = False):
    """Assemble the global stiffness matrix on the GPU.

    Mirrors ``assemble_system`` (src/fem.jl:135): element count ``ne^ndim``
    (``n_elem`` overrides it for non-cubic meshes), ``Young`` scalar or one
    value per element.  Returns ``scipy.sparse.csc_matrix`` (or the device
    matrix with ``device=True``).  As ``sfem_assemble``, a mesh in the
    meshgrid numbering is assembled from its grid (the same pattern; the
    device matrix then knows its grid, ``K.grid``); when every element
    matrix is s_e Ke(0) to within 1e-13, K's values are summed from the
    scales (``K.grid_values()``; ``exact_sums=True``: the element sums,
    bitwise the sort-based K); ``detect_grid=False`` keeps the sort-based
    pipeline.
    """
    ctx = ctx or default_context()
    if n_elem is None:
        n_elem = int(ne) ** int(ndim) if int(ndim) >= 0 else 0
    mesh = Mesh(ctx, n_elem, NodeList, IEN, ndim, FunctionClass, nDof, ID, Young, nu)
    try:
        if detect_grid and not force_scalar:  # as sfem_assemble
            mesh.detect_grid()
        K = mesh.assemble(force_scalar=force_scalar, exact_sums=exact_sums)
    finally:
        mesh.close()
    if device:
        return K
    try:
        return K.to_scipy()
    finally:
        K.close()
