"""CPU oracle for smearFEM's assemble-and-solve path.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of bench.py, as the checker -- never by the product.

ctypes wrapper over ``sfem_oracle.c`` (a C restatement of the reference, see
its header for file:line citations) plus small pure-Python restatements:

* ``meshgrid_ref``  -- examples/vector3D.jl:10-130 (cube meshes, as written,
                       optionally with the top / bottom face connectivity)
* ``inflate_sphere_ref`` -- src/PostProcess.jl:30-46
* ``setboundaryCond_ref`` -- examples/vector3D.jl:133-173
* ``sparse_ref``    -- Julia ``sparse(I,J,V)`` semantics with pure-Python
                       loops (tiny inputs only), an independent check of the
                       C combine
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_SO = HERE / "libsfem_oracle.so"


class OracleError(Exception):
    def __init__(self, code, where=None):
        super().__init__(f"oracle status {code} at element {where}")
        self.code = code
        self.where = where


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return _SO


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not _SO.exists():
            build()
        L = C.CDLL(os.fspath(_SO))
        d, i64, i32, vp = C.c_double, C.c_int64, C.c_int32, C.c_void_p
        P = C.POINTER
        L.or_gaussian_quadrature.argtypes = [d, d, C.c_int, P(d), P(d)]
        L.or_basis_function.argtypes = [C.c_int, d, d, d, C.c_int, P(d), P(d), P(C.c_int), P(C.c_int),
                                        P(C.c_int)]
        L.or_element_stiffness.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, P(d), d, d, P(d)]
        L.or_coo_length.restype = i64
        L.or_coo_length.argtypes = [C.c_int, C.c_int, C.c_int, i64]
        L.or_assemble_coo.argtypes = [C.c_int, C.c_int, C.c_int, P(d), i64, P(i64), i64, C.c_int, i64, P(i64),
                                      i64, P(d), i64, d, P(i64), P(i64), P(d), P(i64)]
        L.or_sparse_dims.argtypes = [i64, P(i64), P(i64), P(i64), P(i64)]
        L.or_sparse_csc.argtypes = [i64, P(i64), P(i64), P(d), i64, i64, P(i64), P(i64), P(d), P(i64)]
        L.or_eliminate_nnz.restype = i64
        L.or_eliminate_nnz.argtypes = [i64, P(i64), P(i64), P(i64)]
        L.or_eliminate.argtypes = [i64, P(i64), P(i64), P(d), P(i64), P(d), P(d), P(i64), P(i32), P(d), P(d)]
        L.or_spmv.argtypes = [i64, P(i64), P(i32), P(d), P(d), P(d)]
        L.or_dense_spd_solve.argtypes = [i64, P(d), P(d), P(d)]
        L.or_pcg.argtypes = [i64, P(i64), P(i32), P(d), P(d), P(d), d, i64, P(i64), P(d)]
        L.or_pcg_mt.argtypes = [i64, P(i64), P(i32), P(d), P(d), P(d), d, i64, C.c_int, P(i64), P(d)]
        L.or_surface_coo.argtypes = [P(d), i64, P(i64), P(i64), i64, C.c_int, C.c_int, C.c_int, P(i64), i64,
                                     P(i64), P(i64), P(d), P(i64)]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


# ------------------------------------------------------------ L1 --
def gaussian_quadrature(a, b, n=2):
    xi = np.zeros(3)
    w = np.zeros(3)
    st = lib().or_gaussian_quadrature(float(a), float(b), int(n), _p(xi, C.c_double), _p(w, C.c_double))
    if st:
        raise OracleError(st)
    return list(xi[:n]), list(w[:n])


def basis_function(xi, eta=None, zeta=None, FunctionClass="Q1"):
    nargs = 1 if eta is None else (2 if zeta is None else 3)
    cls = {"Q1": 1, "Q2": 2}.get(FunctionClass, 0)
    N = np.zeros(9)
    dN = np.zeros(27)
    nf, nr, nc = C.c_int(), C.c_int(), C.c_int()
    st = lib().or_basis_function(nargs, float(xi), float(eta or 0), float(zeta or 0), cls, _p(N, C.c_double),
                                 _p(dN, C.c_double), C.byref(nf), C.byref(nr), C.byref(nc))
    if st:
        raise OracleError(st)
    return N[: nf.value].copy(), dN[: nr.value * nc.value].reshape(nc.value, nr.value).T.copy()


def element_stiffness(ndim, FunctionClass, nDof, coords, E=1.0, nu=0.3):
    """Ke summed over the Gauss loop; coords: ndim x nen.  Row-major L x L."""
    coords = np.asfortranarray(coords, dtype=np.float64)
    nen = coords.shape[1]
    L = nen * nDof
    Ke = np.zeros(L * L)
    st = lib().or_element_stiffness(ndim, {"Q1": 1, "Q2": 2}[FunctionClass], nDof, nen, _p(coords, C.c_double),
                                    float(E), float(nu), _p(Ke, C.c_double))
    if st:
        raise OracleError(st)
    return Ke.reshape(L, L)


# ------------------------------------------------------------ L2 --
def assemble_coo(n_elem, NodeList, IEN, ndim, FunctionClass="Q1", nDof=1, ID=None, Young=1, nu=0.3):
    nodes = np.asfortranarray(NodeList, dtype=np.float64)
    ien = np.asfortranarray(IEN, dtype=np.int64)
    young = np.ascontiguousarray(np.atleast_1d(np.asarray(Young, dtype=np.float64)))
    idm = None if ID is None else np.asfortranarray(ID, dtype=np.int64)
    nen = ien.shape[1]
    ncoo = lib().or_coo_length(ndim, nDof, nen, n_elem)
    E = np.zeros(max(ncoo, 0), dtype=np.int64)
    J = np.zeros(max(ncoo, 0), dtype=np.int64)
    V = np.zeros(max(ncoo, 0), dtype=np.float64)
    err = C.c_int64(-1)
    st = lib().or_assemble_coo(ndim, {"Q1": 1, "Q2": 2}.get(FunctionClass, 0), nDof, _p(nodes, C.c_double),
                               nodes.shape[1], _p(ien, C.c_int64), ien.shape[0], nen, n_elem,
                               _p(idm, C.c_int64) if idm is not None else None,
                               idm.shape[0] if idm is not None else 0, _p(young, C.c_double), young.size,
                               float(nu), _p(E, C.c_int64), _p(J, C.c_int64), _p(V, C.c_double), C.byref(err))
    if st:
        raise OracleError(st, err.value)
    return E, J, V


def sparse_csc(I, J, V):
    """Julia sparse(I,J,V) -> (m, n, colptr, rowval, nzval), 1-based arrays."""
    I = np.ascontiguousarray(I, dtype=np.int64)
    J = np.ascontiguousarray(J, dtype=np.int64)
    V = np.ascontiguousarray(V, dtype=np.float64)
    m, n = C.c_int64(0), C.c_int64(0)
    st = lib().or_sparse_dims(I.size, _p(I, C.c_int64), _p(J, C.c_int64), C.byref(m), C.byref(n))
    if st:
        raise OracleError(st)
    colptr = np.zeros(n.value + 1, dtype=np.int64)
    rowval = np.zeros(max(I.size, 1), dtype=np.int64)
    nzval = np.zeros(max(I.size, 1), dtype=np.float64)
    nnz = C.c_int64(0)
    st = lib().or_sparse_csc(I.size, _p(I, C.c_int64), _p(J, C.c_int64), _p(V, C.c_double), m.value, n.value,
                             _p(colptr, C.c_int64), _p(rowval, C.c_int64), _p(nzval, C.c_double), C.byref(nnz))
    if st:
        raise OracleError(st)
    if n.value == 0:
        colptr[0] = 1
    return m.value, n.value, colptr, rowval[: nnz.value].copy(), nzval[: nnz.value].copy()


def assemble_system(ne, NodeList, IEN, ndim, FunctionClass="Q1", nDof=1, ID=None, Young=1, nu=0.3,
                    n_elem=None):
    """The reference's assemble_system on the CPU -> scipy csc_matrix."""
    import scipy.sparse as sp

    if n_elem is None:
        n_elem = int(ne) ** int(ndim)
    E, J, V = assemble_coo(n_elem, NodeList, IEN, ndim, FunctionClass, nDof, ID, Young, nu)
    m, n, colptr, rowval, nzval = sparse_csc(E, J, V)
    return sp.csc_matrix((nzval, rowval - 1, colptr - 1), shape=(m, n))


def sparse_ref(I, J, V):
    """Pure-Python Julia sparse(): dict keyed by (i,j), values folded left to
    right in input order.  Returns a dict {(i, j): value} and (m, n)."""
    if any(i < 1 for i in I) or any(j < 1 for j in J):
        raise OracleError(-1)
    acc = {}
    for i, j, v in zip(I, J, V):
        k = (int(i), int(j))
        acc[k] = acc[k] + v if k in acc else v
    m = max(I) if len(I) else 0
    n = max(J) if len(J) else 0
    return acc, (int(m), int(n))


def meshgrid_ref(x0, x1, y0, y1, z0, z1, ne, ndim, with_faces=False):
    """examples/vector3D.jl:10-130 restated with plain loops.  Returns
    (NodeList, IEN, ID), or with ``with_faces`` (NodeList, IEN, ID, IEN_top,
    IEN_btm) as the reference does (:94-120; 2D: :45-56)."""
    NodeList = np.zeros((ndim, (ne + 1) ** ndim))
    IEN = np.zeros((ne ** ndim, 2 ** ndim), dtype=np.int64)
    ID = np.zeros(((ne + 1) ** ndim, ndim), dtype=np.int64)
    IEN_top = np.zeros((ne ** (ndim - 1), 2 ** (ndim - 1)), dtype=np.int64)
    IEN_btm = np.zeros((ne ** (ndim - 1), 2 ** (ndim - 1)), dtype=np.int64)
    x = np.linspace(x0, x1, ne + 1)
    y = np.linspace(y0, y1, ne + 1)
    if ndim == 2:
        m = 1
        for j in range(1, ne + 2):
            for i in range(1, ne + 2):
                NodeList[0, m - 1] = x[i - 1]
                NodeList[1, m - 1] = y[j - 1]
                for l in range(1, ndim + 1):
                    ID[m - 1, l - 1] = ndim * (m - 1) + l
                m += 1
        n = 1
        for j in range(1, ne + 1):
            for i in range(1, ne + 1):
                IEN[n - 1] = [(j - 1) * (ne + 1) + i, (j - 1) * (ne + 1) + i + 1, j * (ne + 1) + i + 1,
                              j * (ne + 1) + i]
                if j == 1:
                    IEN_btm[i - 1] = [IEN[n - 1, 0], IEN[n - 1, 1]]
                elif j == ne:
                    IEN_top[i - 1] = [IEN[n - 1, 3], IEN[n - 1, 2]]
                n += 1
    else:
        z = np.linspace(z0, z1, ne + 1)
        m = 1
        for k in range(1, ne + 2):
            for j in range(1, ne + 2):
                for i in range(1, ne + 2):
                    NodeList[:, m - 1] = [x[i - 1], y[j - 1], z[k - 1]]
                    for l in range(1, ndim + 1):
                        ID[m - 1, l - 1] = ndim * (m - 1) + l
                    m += 1
        n = 1
        nt = nb = 1
        p = (ne + 1) ** 2
        for k in range(1, ne + 1):
            for j in range(1, ne + 1):
                for i in range(1, ne + 1):
                    IEN[n - 1] = [(k - 1) * p + (j - 1) * (ne + 1) + i, (k - 1) * p + (j - 1) * (ne + 1) + i + 1,
                                  (k - 1) * p + j * (ne + 1) + i + 1, (k - 1) * p + j * (ne + 1) + i,
                                  k * p + (j - 1) * (ne + 1) + i, k * p + (j - 1) * (ne + 1) + i + 1,
                                  k * p + j * (ne + 1) + i + 1, k * p + j * (ne + 1) + i]
                    if k == 1:
                        IEN_btm[nb - 1] = IEN[n - 1, 0:4]
                        nb += 1
                    elif k == ne:
                        IEN_top[nt - 1] = IEN[n - 1, 4:8]
                        nt += 1
                    n += 1
    if with_faces:
        return NodeList, IEN, ID, IEN_top, IEN_btm
    return NodeList, IEN, ID


def inflate_sphere_ref(NodeList, x0, x1, y0, y1):
    """src/PostProcess.jl:30-46 with plain loops, on a copy (the reference
    overwrites its argument in place, so in examples/vector3D.jl:292 NodeList
    itself becomes the cylinder).  The (x, y) of each node is moved to radius
    max(|x - cx|, |y - cy|) along its direction from the centre; the centre is
    not added back, so the result is centred at the origin; a node at the
    centre goes to (0, 0) (``scale ≈ 0`` holds only for scale == 0)."""
    out = np.array(NodeList, dtype=np.float64, copy=True)
    cx, cy = 0.5 * (x0 + x1), 0.5 * (y0 + y1)
    for i in range(out.shape[1]):
        dx, dy = out[0, i] - cx, out[1, i] - cy
        scale = max(abs(dx), abs(dy))
        if scale == 0.0:
            out[0, i], out[1, i] = 0.0, 0.0
        else:
            r = np.sqrt(dx * dx + dy * dy)
            out[0, i], out[1, i] = scale * dx / r, scale * dy / r
    return out


def setboundaryCond_ref(NodeList, ne, ndim, d, nDof=3):
    """examples/vector3D.jl:133-173: z dof of nodes on z == 0 fixed to 0, on
    z == 1 fixed to -d (exact comparisons, first match).  Returns (q_d, fixed)
    with fixed the mask of the columns removed from C (rCol)."""
    n = nDof * (ne + 1) ** ndim
    q_d = np.zeros(n)
    fixed = np.zeros(n, dtype=np.uint8)
    for node in range(NodeList.shape[1]):
        z = NodeList[2, node]
        if z == 0:
            q_d[3 * node + 2] = 0.0
            fixed[3 * node + 2] = 1
        elif z == 1:
            q_d[3 * node + 2] = -d
            fixed[3 * node + 2] = 1
    return q_d, fixed


def surface_coo(NodeList, IEN_btm, IEN_top, ID, nDof=3):
    """apply_boundary_conditions (examples/vector3D.jl:175-264) up to sparse():
    (E, J, V), bottom faces first, then top."""
    nodes = np.asfortranarray(NodeList, dtype=np.float64)
    top = np.asfortranarray(IEN_top, dtype=np.int64)
    btm = np.asfortranarray(IEN_btm, dtype=np.int64)
    idm = np.asfortranarray(ID, dtype=np.int64)
    nf, nfn = btm.shape
    L = nDof * nfn
    n = 2 * nf * L * L
    E = np.zeros(n, dtype=np.int64)
    J = np.zeros(n, dtype=np.int64)
    V = np.zeros(n, dtype=np.float64)
    err = C.c_int64(-1)
    st = lib().or_surface_coo(_p(nodes, C.c_double), nodes.shape[1], _p(btm, C.c_int64), _p(top, C.c_int64), nf,
                              nfn, nodes.shape[0], nDof, _p(idm, C.c_int64), idm.shape[0], _p(E, C.c_int64),
                              _p(J, C.c_int64), _p(V, C.c_double), C.byref(err))
    if st:
        raise OracleError(st, err.value)
    return E, J, V


def apply_boundary_conditions(ne, NodeList, IEN, IEN_top, IEN_btm, ndim, FunctionClass="Q1", ID=None, nDof=3):
    """The reference's surface penalty b = sparse(E, J, V) -> scipy csc_matrix."""
    import scipy.sparse as sp

    E, J, V = surface_coo(NodeList, IEN_btm, IEN_top, ID, nDof)
    m, n, colptr, rowval, nzval = sparse_csc(E, J, V)
    return sp.csc_matrix((nzval, rowval - 1, colptr - 1), shape=(m, n))


# ---------------------------------------------------- elimination / solve --
def eliminate(K, fixed, f=None, q_d=None):
    """(K_ff as scipy csc, rhs) following examples/vector3D.jl:315-316,320."""
    import scipy.sparse as sp

    K = K.tocsc()
    n = K.shape[0]
    colptr = (K.indptr + 1).astype(np.int64)
    rowval = (K.indices + 1).astype(np.int64)
    nzval = K.data.astype(np.float64)
    fixed = np.asarray(fixed, dtype=bool)
    free_map = np.where(~fixed, np.cumsum(~fixed) - 1, -1).astype(np.int64)
    nf = int((~fixed).sum())
    nnz = lib().or_eliminate_nnz(n, _p(colptr, C.c_int64), _p(rowval, C.c_int64), _p(free_map, C.c_int64))
    ffp = np.zeros(nf + 1, dtype=np.int64)
    ffr = np.zeros(max(nnz, 1), dtype=np.int32)
    ffv = np.zeros(max(nnz, 1), dtype=np.float64)
    rhs = np.zeros(nf)
    fv = None if f is None else np.ascontiguousarray(f, dtype=np.float64)
    qv = None if q_d is None else np.ascontiguousarray(np.where(fixed, q_d, 0.0), dtype=np.float64)
    st = lib().or_eliminate(n, _p(colptr, C.c_int64), _p(rowval, C.c_int64), _p(nzval, C.c_double),
                            _p(free_map, C.c_int64), _p(fv, C.c_double) if fv is not None else None,
                            _p(qv, C.c_double) if qv is not None else None, _p(ffp, C.c_int64),
                            _p(ffr, C.c_int32), _p(ffv, C.c_double), _p(rhs, C.c_double))
    if st:
        raise OracleError(st)
    Kff = sp.csc_matrix((ffv[:nnz], ffr[:nnz], ffp), shape=(nf, nf))
    return Kff, rhs


def dense_solve(Kff, rhs):
    A = np.ascontiguousarray(Kff.toarray(), dtype=np.float64)
    x = np.zeros_like(rhs)
    b = np.ascontiguousarray(rhs, dtype=np.float64)
    st = lib().or_dense_spd_solve(A.shape[0], _p(A, C.c_double), _p(b, C.c_double), _p(x, C.c_double))
    if st:
        raise OracleError(st)
    return x


def pcg(Kff, rhs, rtol=1e-10, max_iter=100000, x0=None):
    """FP64 Jacobi-PCG, single thread; returns (x, iterations, relres)."""
    A = Kff.tocsc()
    ptr = A.indptr.astype(np.int64)
    idx = A.indices.astype(np.int32)
    val = A.data.astype(np.float64)
    b = np.ascontiguousarray(rhs, dtype=np.float64)
    x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, dtype=np.float64).copy()
    it = C.c_int64(0)
    rr = C.c_double(0)
    st = lib().or_pcg(b.size, _p(ptr, C.c_int64), _p(idx, C.c_int32), _p(val, C.c_double), _p(b, C.c_double),
                      _p(x, C.c_double), float(rtol), int(max_iter), C.byref(it), C.byref(rr))
    if st not in (0, -6):
        raise OracleError(st)
    return x, it.value, rr.value


def pcg_threads(Kff, rhs, nthreads, rtol=1e-10, max_iter=100000):
    """pcg with the loops over nthreads OpenMP threads (CPU baseline timing)."""
    A = Kff.tocsc()
    ptr = A.indptr.astype(np.int64)
    idx = A.indices.astype(np.int32)
    val = A.data.astype(np.float64)
    b = np.ascontiguousarray(rhs, dtype=np.float64)
    x = np.zeros_like(b)
    it = C.c_int64(0)
    rr = C.c_double(0)
    st = lib().or_pcg_mt(b.size, _p(ptr, C.c_int64), _p(idx, C.c_int32), _p(val, C.c_double), _p(b, C.c_double),
                         _p(x, C.c_double), float(rtol), int(max_iter), int(nthreads), C.byref(it), C.byref(rr))
    if st not in (0, -6):
        raise OracleError(st)
    return x, it.value, rr.value


def recombine(fixed, q_d, x):
    """q = q_d + C q_f (examples/vector3D.jl:322)."""
    fixed = np.asarray(fixed, dtype=bool)
    q = np.where(fixed, 0.0 if q_d is None else q_d, 0.0).astype(np.float64)
    q[~fixed] = x
    return q


# ------------------------------------------------------------------ specs --
# Restatement of the device problem generator (sfem_problem_spec) with numpy:
# mesh as meshgrid (examples/vector3D.jl:10-130, generalised to nx x ny x nz
# elements on a box), clamp as setboundaryCond's fixed face (:133-173), and the
# consistent loads and counter-based random draws the header describes (loads
# are not in the reference: they are pinned by closed forms in the tests).
_U64 = np.uint64
_SM1, _SM2 = _U64(0xBF58476D1CE4E5B9), _U64(0x94D049BB133111EB)
_SC1, _SC2 = _U64(0xD1B54A32D192ED03), _U64(0x9E3779B97F4A7C15)


def spec_u01(seed, stream, index):
    """u in (0, 1) for (seed, stream, index array): splitmix64 finaliser."""
    idx = np.asarray(index, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (_U64(seed) ^ (_U64(stream) * _SC1) ^ (idx * _SC2)) + _SC2
        z = (z ^ (z >> _U64(30))) * _SM1
        z = (z ^ (z >> _U64(27))) * _SM2
        z = z ^ (z >> _U64(31))
    return ((z >> _U64(12)).astype(np.float64) + 0.5) * 2.0 ** -52


def _lin(x0, x1, ne, i):
    i = np.asarray(i)
    v = i.astype(np.float64) * ((x1 - x0) / float(ne)) + x0
    return np.where(i == ne, x1, v)


_Q1SX = np.array([-1, 1, 1, -1, -1, 1, 1, -1], dtype=np.float64)
_Q1SY = np.array([-1, -1, 1, 1, -1, -1, 1, 1], dtype=np.float64)
_Q1SZ = np.array([-1, -1, -1, -1, 1, 1, 1, 1], dtype=np.float64)


def _gauss_points(ndim):
    s = 2.0 / (2.0 * np.sqrt(3.0))
    return [(_Q1SX[g] * s, _Q1SY[g] * s, _Q1SZ[g] * s) for g in range(1 << ndim)]


def _cell_integral(X, ndim):
    """int N_a dV (dA in 2D) per corner a of cells with corner coordinates X
    [cells, nen, ndim], 2-point Gauss, the device's order of roundings."""
    nen = 1 << ndim
    out = np.zeros((X.shape[0], nen))
    for (xi, eta, zeta) in _gauss_points(ndim):
        J = np.zeros((X.shape[0], ndim, ndim))
        for b in range(nen):
            if ndim == 3:
                fx, fy, fz = 1 + _Q1SX[b] * xi, 1 + _Q1SY[b] * eta, 1 + _Q1SZ[b] * zeta
                dn = (_Q1SX[b] * (fy * fz) / 8, _Q1SY[b] * (fx * fz) / 8, _Q1SZ[b] * (fx * fy) / 8)
            else:
                dn = (_Q1SX[b] * (1 + _Q1SY[b] * eta) / 4, _Q1SY[b] * (1 + _Q1SX[b] * xi) / 4)
            for i in range(ndim):
                for j in range(ndim):
                    J[:, i, j] = J[:, i, j] + X[:, b, i] * dn[j]
        if ndim == 3:
            det = (J[:, 0, 0] * (J[:, 1, 1] * J[:, 2, 2] - J[:, 1, 2] * J[:, 2, 1]) +
                   J[:, 0, 1] * (J[:, 1, 2] * J[:, 2, 0] - J[:, 1, 0] * J[:, 2, 2]) +
                   J[:, 0, 2] * (J[:, 1, 0] * J[:, 2, 1] - J[:, 1, 1] * J[:, 2, 0]))
        else:
            det = J[:, 0, 0] * J[:, 1, 1] - J[:, 0, 1] * J[:, 1, 0]
        for a in range(nen):
            if ndim == 3:
                N = (1 + _Q1SX[a] * xi) * (1 + _Q1SY[a] * eta) * (1 + _Q1SZ[a] * zeta) / 8
            else:
                N = (1 + _Q1SX[a] * xi) * (1 + _Q1SY[a] * eta) / 4
            out[:, a] = out[:, a] + N * np.abs(det)
    return out


def _face_integral(X):
    """int N_a dA per corner a of bilinear faces X [faces, 4, 3], 2x2 Gauss."""
    out = np.zeros((X.shape[0], 4))
    for (xi, eta, _) in _gauss_points(2):
        t1 = np.zeros((X.shape[0], 3))
        t2 = np.zeros((X.shape[0], 3))
        for b in range(4):
            d0 = _Q1SX[b] * (1 + _Q1SY[b] * eta) / 4
            d1 = _Q1SY[b] * (1 + _Q1SX[b] * xi) / 4
            for d in range(3):
                t1[:, d] = t1[:, d] + X[:, b, d] * d0
                t2[:, d] = t2[:, d] + X[:, b, d] * d1
        c0 = t1[:, 1] * t2[:, 2] - t1[:, 2] * t2[:, 1]
        c1 = t1[:, 2] * t2[:, 0] - t1[:, 0] * t2[:, 2]
        c2 = t1[:, 0] * t2[:, 1] - t1[:, 1] * t2[:, 0]
        area = np.sqrt(c0 * c0 + c1 * c1 + c2 * c2)
        for a in range(4):
            out[:, a] = out[:, a] + (1 + _Q1SX[a] * xi) * (1 + _Q1SY[a] * eta) / 4 * area
    return out


def _edge_integral(X):
    """int N_a ds per end a of straight edges X [edges, 2, 2], 2-point Gauss."""
    s = 2.0 / (2.0 * np.sqrt(3.0))
    t0 = (X[:, 1, 0] - X[:, 0, 0]) / 2
    t1 = (X[:, 1, 1] - X[:, 0, 1]) / 2
    length = np.sqrt(t0 * t0 + t1 * t1)
    out = np.zeros((X.shape[0], 2))
    for xi in (-s, s):
        out[:, 0] = out[:, 0] + (1 - xi) / 2 * length
        out[:, 1] = out[:, 1] + (1 + xi) / 2 * length
    return out


def spec_inputs(spec):
    """(NodeList ndim x n, IEN n_elem x nen 1-based, ID n x ndim, young,
    nu, f, fixed) of a problem spec: a dict with the fields of
    sfem_problem_spec (ndim, dims, box, young_lo, young_hi, nu, seed,
    clamp_axis, clamp_side, load_axis, load_side, traction, body, noise)."""
    nd = int(spec["ndim"])
    ne = [int(spec["dims"][d]) if d < nd else 1 for d in range(3)]
    nn = [ne[d] + 1 if d < nd else 1 for d in range(3)]
    box = list(spec["box"])
    n_nodes = nn[0] * nn[1] * nn[2]
    n_elem = ne[0] * ne[1] * ne[2]
    m = np.arange(n_nodes)
    idx = np.stack([m % nn[0], (m // nn[0]) % nn[1], m // (nn[0] * nn[1])])
    coord = [_lin(box[2 * d], box[2 * d + 1], ne[d], np.arange(nn[d])) for d in range(nd)]
    NodeList = np.stack([coord[d][idx[d]] for d in range(nd)])
    # meshgrid connectivity in the reference's corner order
    e = np.arange(n_elem)
    ci = np.stack([e % ne[0], (e // ne[0]) % ne[1], e // (ne[0] * ne[1])])
    nen = 1 << nd
    IEN = np.zeros((n_elem, nen), dtype=np.int64)
    for a in range(nen):
        ox, oy, oz = int(_Q1SX[a] > 0), int(_Q1SY[a] > 0), int(_Q1SZ[a] > 0) if nd == 3 else 0
        IEN[:, a] = (ci[0] + ox) + nn[0] * ((ci[1] + oy) + nn[1] * (ci[2] + oz)) + 1
    ID = (nd * m[:, None] + np.arange(nd)[None, :] + 1).astype(np.int64)
    lo, hi = float(spec["young_lo"]), float(spec["young_hi"])
    if hi == lo:
        young = np.full(n_elem, lo)
    else:
        young = np.exp(np.log(lo) + spec_u01(spec["seed"], 1, e) * (np.log(hi) - np.log(lo)))
    nu = np.full(n_elem, float(spec["nu"]))
    ca, cs = int(spec["clamp_axis"]), int(spec["clamp_side"])
    la, ls = int(spec["load_axis"]), int(spec["load_side"])
    clamped = idx[ca] == (nn[ca] - 1 if cs else 0)
    loaded = idx[la] == (nn[la] - 1 if ls else 0)
    load = np.zeros((n_nodes, nd))
    body = np.array(spec["body"][:nd], dtype=np.float64)
    if np.any(body != 0):
        Xc = NodeList[:, IEN - 1].transpose(1, 2, 0)  # [cells, nen, nd]
        w_cell = _cell_integral(Xc, nd)
        w = np.zeros(n_nodes)
        for oz in ((-1, 0) if nd == 3 else (0,)):
            for oy in (-1, 0):
                for ox in (-1, 0):
                    c = [idx[0] + ox, idx[1] + oy, idx[2] + oz]
                    ok = np.ones(n_nodes, dtype=bool)
                    for d in range(3):
                        ok &= (c[d] >= 0) & (c[d] < ne[d])
                    a = ((-ox) ^ (-oy)) + 2 * (-oy) + (4 * (-oz) if nd == 3 else 0)
                    cell = c[0] + ne[0] * (c[1] + ne[1] * c[2])
                    w = np.where(ok, w + w_cell[np.where(ok, cell, 0), a], w)
        load = load + body[None, :] * w[:, None]
    trac = np.array(spec["traction"][:nd], dtype=np.float64)
    if np.any(trac != 0):
        w = np.zeros(n_nodes)
        if nd == 3:
            du, dv = (1, 2) if la == 0 else ((0, 2) if la == 1 else (0, 1))
            for ov in (-1, 0):
                for ou in (-1, 0):
                    iu, iv = idx[du] + ou, idx[dv] + ov
                    ok = loaded & (iu >= 0) & (iu < ne[du]) & (iv >= 0) & (iv < ne[dv])
                    X = np.zeros((n_nodes, 4, 3))
                    for b in range(4):
                        ii = [None, None, None]
                        ii[la] = idx[la]
                        ii[du] = iu + int(_Q1SX[b] > 0)
                        ii[dv] = iv + int(_Q1SY[b] > 0)
                        for d in range(3):
                            X[:, b, d] = _lin(box[2 * d], box[2 * d + 1], ne[d], np.clip(ii[d], 0, ne[d]))
                    a = ((-ou) ^ (-ov)) + 2 * (-ov)
                    w = np.where(ok, w + _face_integral(X)[:, a], w)
        else:
            du = 1 - la
            for ou in (-1, 0):
                iu = idx[du] + ou
                ok = loaded & (iu >= 0) & (iu < ne[du])
                X = np.zeros((n_nodes, 2, 2))
                for b in range(2):
                    ii = [None, None]
                    ii[la] = idx[la]
                    ii[du] = iu + b
                    for d in range(2):
                        X[:, b, d] = _lin(box[2 * d], box[2 * d + 1], ne[d], np.clip(ii[d], 0, ne[d]))
                w = np.where(ok, w + _edge_integral(X)[:, -ou], w)
        load = load + trac[None, :] * w[:, None]
    f = load.reshape(-1).copy()
    fixed = np.repeat(clamped, nd).astype(np.uint8)
    sigma = float(spec["noise"])
    if sigma != 0.0:
        k = np.arange(n_nodes * nd, dtype=np.uint64)
        u1 = spec_u01(spec["seed"], 2, 2 * k)
        u2 = spec_u01(spec["seed"], 2, 2 * k + 1)
        noise = sigma * (np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2))
        on = np.repeat(loaded & ~clamped, nd)
        f = np.where(on, f + noise, f)
    return {"NodeList": NodeList, "IEN": IEN, "ID": ID, "young": young, "nu": nu, "f": f, "fixed": fixed,
            "n_elem": n_elem}
