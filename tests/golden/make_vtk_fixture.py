"""Golden VTK files: the exact bytes the writers must produce in ASCII mode.

The text is spelled out here from the VTK XML file format (UnstructuredGrid
with Points, Cells = connectivity / offsets / types and PointData; a
Collection of DataSets for the ParaView time series), the layout WriteVTK.jl
gives the files of src/PostProcess.jl:258-304: one Hex8 cell (the reference's
IEN order = VTK_HEXAHEDRON, type 12, 0-based connectivity) with a point
vector "u", and the collection "<name>displacement.pvd" listing
"<name>timestep_<i>.vtu" at time steps 0, 1.  It is not produced by the
writer under test.

Run:  python tests/golden/make_vtk_fixture.py
"""
from pathlib import Path

HERE = Path(__file__).resolve().parent

# unit cube, nodes in meshgrid order (x fastest), u = (x, 2y, -z)
POINTS = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (1, 1, 1)]
IEN = [1, 2, 4, 3, 5, 6, 8, 7]


def f(v):
    return repr(float(v))


def vtu():
    pts = " ".join(f(c) for p in POINTS for c in p)
    conn = " ".join(str(i - 1) for i in IEN)
    u = " ".join(f(c) for (x, y, z) in POINTS for c in (float(x), 2.0 * y, -float(z)))  # -0.0 at z = 0
    return ('<?xml version="1.0"?>\n'
            '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n'
            '<UnstructuredGrid>\n'
            '<Piece NumberOfPoints="8" NumberOfCells="1">\n'
            f'<Points><DataArray type="Float64" NumberOfComponents="3" format="ascii">{pts}</DataArray></Points>\n'
            f'<Cells><DataArray type="Int64" Name="connectivity" format="ascii">{conn}</DataArray>'
            '<DataArray type="Int64" Name="offsets" format="ascii">8</DataArray>'
            '<DataArray type="UInt8" Name="types" format="ascii">12</DataArray></Cells>\n'
            f'<PointData><DataArray type="Float64" Name="u" NumberOfComponents="3" format="ascii">{u}</DataArray>'
            '</PointData>\n'
            '</Piece>\n'
            '</UnstructuredGrid>\n'
            '</VTKFile>\n')


def pvd(name):
    return ('<?xml version="1.0"?>\n<VTKFile type="Collection" version="1.0" byte_order="LittleEndian">\n'
            '<Collection>\n'
            f'<DataSet timestep="0" part="0" file="{name}timestep_1.vtu"/>\n'
            f'<DataSet timestep="1" part="0" file="{name}timestep_2.vtu"/>\n'
            '</Collection>\n</VTKFile>\n')


if __name__ == "__main__":
    (HERE / "vtk_hex1.vtu").write_text(vtu())
    (HERE / "vtk_scene.pvd").write_text(pvd("scene_"))
    print("wrote vtk_hex1.vtu, vtk_scene.pvd")
