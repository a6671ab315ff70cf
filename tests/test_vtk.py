"""VTK output (src/PostProcess.jl:258-304): files parsed back.

Checks: points, 0-based connectivity in the reference's IEN order, offsets,
cell types (hexahedron 12, quad 9), the displacement field as 3-component
point vectors from both the reference's 3 x nNodes motion and a
node-interleaved solution vector, and the ParaView collection's time steps.
"""
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import oracle as O
from smearfem.vtk import read_vtu, write_scene, write_vtk


@pytest.mark.parametrize("encoding", ["raw", "ascii"])
def test_write_vtk_hex_roundtrip(tmp_path, encoding):
    ne = 3
    NL, IEN, ID = O.meshgrid_ref(0, 1, 0, 1, 0, 1, ne, 3)
    q = np.arange(NL.shape[1] * 3, dtype=np.float64) * 0.5
    path = write_vtk(str(tmp_path / "cube"), "q", NL, IEN, ne, 3, q, encoding=encoding)
    g = read_vtu(path)
    assert np.array_equal(g["points"], NL.T)
    assert np.array_equal(g["connectivity"].reshape(-1, 8), IEN - 1)
    assert np.array_equal(g["offsets"], 8 * np.arange(1, ne ** 3 + 1))
    assert np.all(g["types"] == 12)
    assert np.array_equal(g["fields"]["q"], q.reshape(-1, 3))


def test_write_vtk_quad_scalar(tmp_path):
    ne = 4
    NL, IEN, ID = O.meshgrid_ref(0, 2, 0, 1, 0, 0, ne, 2)
    T = NL[0] ** 2 + NL[1]
    g = read_vtu(write_vtk(str(tmp_path / "sq"), "T", NL, IEN, ne, 2, T))
    assert np.array_equal(g["points"][:, :2], NL.T) and np.all(g["points"][:, 2] == 0)
    assert np.all(g["types"] == 9)
    assert np.array_equal(g["fields"]["T"], T)


def test_write_scene_collection(tmp_path):
    ne = 2
    NL, IEN, ID = O.meshgrid_ref(0, 1, 0, 1, 0, 1, ne, 3)
    fields = [np.vstack([NL[0] * s, NL[1] * s, -NL[2] * s]) for s in (0.0, 0.1, 0.2)]  # 3 x nNodes motions
    pvd = write_scene(str(tmp_path / "squeeze_flow"), NL, IEN, ne, 3, fields)
    assert pvd.endswith("squeeze_flowdisplacement.pvd")
    ds = ET.parse(pvd).getroot().findall("Collection/DataSet")
    assert [d.get("timestep") for d in ds] == ["0", "1", "2"]
    for d, f in zip(ds, fields):
        g = read_vtu(str(tmp_path / d.get("file")))
        assert np.array_equal(g["fields"]["u"], f.T)


def test_write_vtk_rejects_bad_cells(tmp_path):
    NL, IEN, ID = O.meshgrid_ref(0, 1, 0, 1, 0, 1, 2, 3)
    bad = IEN.copy()
    bad[3, 5] = NL.shape[1] + 1
    with pytest.raises(IndexError):
        write_vtk(str(tmp_path / "bad"), "q", NL, bad, 2, 3, np.zeros(NL.shape[1]))
    with pytest.raises(ValueError):
        write_vtk(str(tmp_path / "bad"), "q", NL, IEN, 2, 3, np.zeros(NL.shape[1] + 1))


def test_layout_matches_golden_files(tmp_path):
    """The ASCII files byte for byte against tests/golden (spelled out from the
    VTK XML format by make_vtk_fixture.py, not by this writer): one Hex8
    cell, 0-based connectivity in the reference's IEN order, offsets, type
    12, the point vector "u"; the collection's file names and time steps."""
    from pathlib import Path

    gold = Path(__file__).resolve().parent / "golden"
    NL = np.array([[0, 1, 0, 1, 0, 1, 0, 1], [0, 0, 1, 1, 0, 0, 1, 1], [0, 0, 0, 0, 1, 1, 1, 1]], dtype=float)
    IEN = np.array([[1, 2, 4, 3, 5, 6, 8, 7]])
    u = np.vstack([NL[0], 2 * NL[1], -NL[2]])
    path = write_vtk(str(tmp_path / "hex1"), "u", NL, IEN, 1, 3, u, encoding="ascii")
    assert open(path, "rb").read() == (gold / "vtk_hex1.vtu").read_bytes()
    pvd = write_scene(str(tmp_path / "scene_"), NL, IEN, 1, 3, [u, u], encoding="ascii")
    assert open(pvd, "rb").read() == (gold / "vtk_scene.pvd").read_bytes()
    assert open(str(tmp_path / "scene_timestep_1.vtu"), "rb").read() == (gold / "vtk_hex1.vtu").read_bytes()


def test_write_solution_of_a_spec(tmp_path):
    """A served solution (node-interleaved dofs) of a spec request: the file's
    grid is the spec's meshgrid (the oracle's restatement of the device
    generator), the field the solution."""
    from smearfem.spec import cantilever_spec
    from smearfem.vtk import write_solution

    sp = cantilever_spec((4, 3, 2), lengths=(2.0, 1.0, 0.5))
    ref = O.spec_inputs(sp.as_dict())
    q = np.arange(sp.n_dof, dtype=float) / 7.0
    g = read_vtu(write_solution(str(tmp_path / "req"), sp, q))
    assert np.allclose(g["points"], ref["NodeList"].T, rtol=1e-15, atol=1e-15)
    assert np.array_equal(g["connectivity"].reshape(-1, 8), ref["IEN"] - 1)
    assert np.array_equal(g["fields"]["u"], q.reshape(-1, 3))
