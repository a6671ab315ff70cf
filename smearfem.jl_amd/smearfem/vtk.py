"""VTK output of displacement fields: ``write_vtk`` and ``write_scene``.

Mirrors src/PostProcess.jl:258-304, which write through WriteVTK.jl:

* ``write_vtk(fileName, fieldName, NodeList, IEN, ne, ndim, q)`` -- one
  unstructured grid (``fileName.vtu``) with the ne^ndim cells of IEN (lines,
  quads or hexahedra by ndim) and the point field ``fieldName``;
* ``write_scene(fileName, NodeList, IEN, ne, ndim, fields)`` -- a ParaView
  collection ``fileName + "displacement.pvd"`` whose time step i-1 is
  ``fileName + f"timestep_{i}.vtu"`` with the point field "u" = fields[i]
  (hexahedra, as the reference writes).

Files are VTK XML UnstructuredGrid, version 1.0, with UInt64 headers; data
are appended raw binary (``encoding="raw"``, the compact default) or ASCII.
Connectivity is written 0-based, as VTK expects; the reference's IEN node
order (bottom face counter-clockwise, then top) is VTK's hexahedron order.
A field of 3 x nNodes (or nNodes*3 node-interleaved values) is a 3-component
point vector, one value per node a scalar.  Host code; fields come from the
device solutions.
"""
from __future__ import annotations

import os
from typing import Sequence

import numpy as np

_CELL_TYPES = {1: 3, 2: 9, 3: 12}  # VTK_LINE, VTK_QUAD, VTK_HEXAHEDRON


def _point_field(q, n_points):
    a = np.asarray(q, dtype=np.float64)
    if a.ndim == 2 and a.shape[1] == n_points:       # ncomp x nNodes (the reference's motion)
        return np.ascontiguousarray(a.T)
    if a.ndim == 2 and a.shape[0] == n_points:
        return np.ascontiguousarray(a)
    a = a.ravel()
    if a.size == n_points:
        return a
    if a.size % n_points == 0:                      # node-interleaved dofs, ID[m,l] = ncomp(m-1)+l
        return a.reshape(n_points, a.size // n_points)
    raise ValueError(f"field of {a.size} values does not fit {n_points} points")


class _Writer:
    def __init__(self, encoding):
        if encoding not in ("raw", "ascii"):
            raise ValueError("encoding must be 'raw' or 'ascii'")
        self.encoding = encoding
        self.blobs = []
        self.offset = 0

    def array(self, name, data, vtk_type, ncomp=1):
        attrs = f'type="{vtk_type}"' + (f' Name="{name}"' if name else "") + \
                (f' NumberOfComponents="{ncomp}"' if ncomp > 1 else "")
        if self.encoding == "ascii":
            flat = np.asarray(data).ravel()
            text = " ".join(repr(float(v)) if vtk_type.startswith("Float") else str(int(v)) for v in flat)
            return f'<DataArray {attrs} format="ascii">{text}</DataArray>'
        raw = np.ascontiguousarray(data).tobytes()
        blob = np.uint64(len(raw)).tobytes() + raw
        tag = f'<DataArray {attrs} format="appended" offset="{self.offset}"/>'
        self.blobs.append(blob)
        self.offset += len(blob)
        return tag

    def finish(self, body):
        head = '<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" ' \
               'header_type="UInt64">\n'
        out = head.encode() + body.encode()
        if self.encoding == "raw":
            out += b'<AppendedData encoding="raw">\n_' + b"".join(self.blobs) + b"\n</AppendedData>\n"
        return out + b"</VTKFile>\n"


def _grid(NodeList, IEN, n_cells, ndim, point_data, encoding):
    nodes = np.asarray(NodeList, dtype=np.float64)
    if nodes.shape[0] != ndim:
        raise ValueError("NodeList must be ndim x nNodes")
    n_points = nodes.shape[1]
    pts = np.zeros((n_points, 3))
    pts[:, :ndim] = nodes.T
    ien = np.asarray(IEN, dtype=np.int64)[:n_cells]
    if ien.shape[0] < n_cells:
        raise IndexError("IEN has fewer rows than ne^ndim cells")
    conn = (ien - 1).ravel()
    if conn.size and (conn.min() < 0 or conn.max() >= n_points):
        raise IndexError("IEN refers to nodes outside NodeList")
    offsets = np.arange(1, n_cells + 1, dtype=np.int64) * ien.shape[1]
    types = np.full(n_cells, _CELL_TYPES[ndim], dtype=np.uint8)
    w = _Writer(encoding)
    pd = "".join(w.array(name, arr, "Float64", 1 if arr.ndim == 1 else arr.shape[1]) for name, arr in point_data)
    body = (f'<UnstructuredGrid>\n<Piece NumberOfPoints="{n_points}" NumberOfCells="{n_cells}">\n'
            f'<Points>{w.array(None, pts, "Float64", 3)}</Points>\n'
            f'<Cells>{w.array("connectivity", conn, "Int64")}{w.array("offsets", offsets, "Int64")}'
            f'{w.array("types", types, "UInt8")}</Cells>\n'
            f'<PointData>{pd}</PointData>\n</Piece>\n</UnstructuredGrid>\n')
    return w.finish(body)


def write_vtk(fileName, fieldName, NodeList, IEN, ne, ndim, q, encoding="raw") -> str:
    """src/PostProcess.jl:258-274: ``fileName.vtu`` with point field ``fieldName``."""
    if ndim not in _CELL_TYPES:
        raise ValueError("ndim must be 1, 2 or 3")
    n_points = np.asarray(NodeList).shape[1]
    path = f"{fileName}.vtu"
    with open(path, "wb") as fh:
        fh.write(_grid(NodeList, IEN, int(ne) ** int(ndim), ndim, [(fieldName, _point_field(q, n_points))],
                       encoding))
    return path


def write_scene(fileName, NodeList, IEN, ne, ndim, fields: Sequence, encoding="raw") -> str:
    """src/PostProcess.jl:289-304: one hexahedral grid per field ("u"), time
    steps 0, 1, ... collected in ``fileName + "displacement.pvd"``."""
    n_points = np.asarray(NodeList).shape[1]
    entries = []
    for i, q in enumerate(fields, start=1):
        path = f"{fileName}timestep_{i}.vtu"
        with open(path, "wb") as fh:
            fh.write(_grid(NodeList, IEN, int(ne) ** int(ndim), 3, [("u", _point_field(q, n_points))], encoding))
        entries.append(f'<DataSet timestep="{i - 1}" part="0" file="{os.path.basename(path)}"/>')
    pvd = f"{fileName}displacement.pvd"
    with open(pvd, "w") as fh:
        fh.write('<?xml version="1.0"?>\n<VTKFile type="Collection" version="1.0" byte_order="LittleEndian">\n'
                 '<Collection>\n' + "\n".join(entries) + "\n</Collection>\n</VTKFile>\n")
    return pvd


def write_solution(fileName, request, q, fieldName="u", encoding="raw") -> str:
    """``fileName.vtu`` of a solved problem: a ``Problem`` (its mesh) or a
    ``ProblemSpec`` (the meshgrid of its box and element counts, the mesh the
    device generated), with the displacement q (n_dof values, node-interleaved,
    as a server returns it) as the point vector ``fieldName``."""
    mesh = getattr(request, "mesh", None)
    if mesh is None:  # a spec: the same numbering and coordinates as the device's meshgrid
        from .mesh import structured_mesh

        nd = request.ndim
        box = request.box
        mesh = structured_mesh(request.dims[:nd], lengths=[box[2 * d + 1] - box[2 * d] for d in range(nd)],
                               origin=[box[2 * d] for d in range(nd)])
    n_points = mesh.NodeList.shape[1]
    path = f"{fileName}.vtu"
    with open(path, "wb") as fh:
        fh.write(_grid(mesh.NodeList, mesh.IEN, mesh.IEN.shape[0], mesh.ndim,
                       [(fieldName, _point_field(q, n_points))], encoding))
    return path


def read_vtu(path) -> dict:
    """Points, connectivity, offsets, types and point fields of a file written
    here (for checks; raw or ASCII)."""
    import re
    import xml.etree.ElementTree as ET

    data = open(path, "rb").read()
    app = data.find(b"<AppendedData")
    xml_part = data if app < 0 else data[:app] + b"</VTKFile>\n"
    blob = b""
    if app >= 0:
        start = data.index(b"_", app) + 1
        end = data.rindex(b"</AppendedData>")
        blob = data[start:end].rstrip(b"\n")
    root = ET.fromstring(xml_part)
    types = {"Float64": np.float64, "Int64": np.int64, "UInt8": np.uint8}

    def read(el):
        dt = types[el.get("type")]
        nc = int(el.get("NumberOfComponents", "1"))
        if el.get("format") == "ascii":
            arr = np.array(re.split(r"\s+", el.text.strip()), dtype=dt) if el.text and el.text.strip() else \
                np.zeros(0, dt)
        else:
            off = int(el.get("offset"))
            n = int(np.frombuffer(blob[off:off + 8], dtype=np.uint64)[0])
            arr = np.frombuffer(blob[off + 8:off + 8 + n], dtype=dt)
        return arr.reshape(-1, nc) if nc > 1 else arr

    piece = root.find("UnstructuredGrid/Piece")
    out = {"points": read(piece.find("Points/DataArray"))}
    for el in piece.find("Cells"):
        out[el.get("Name")] = read(el)
    out["fields"] = {el.get("Name"): read(el) for el in piece.find("PointData")}
    return out
