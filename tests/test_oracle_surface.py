"""Oracle pins for the surface penalty, the cylinder mesh and the z-face
constraints of the reference example (examples/vector3D.jl:10-264,
src/PostProcess.jl:30-46).

What fixes each function without running Julia:
* surface COO / sparse(): the consistent Q4 mass matrix of a rectangle,
  (hx*hy/36) [4 2 1 2; 2 4 2 1; 1 2 4 2; 2 1 2 4] per dof (2x2 Gauss is exact
  for bilinear x bilinear on rectangles); sum of all entries per dof = area
  of the faces (sum_ab N_a N_b = 1), checked against the shoelace area of the
  (planar, inflated) face quads; symmetry, positive semi-definiteness,
  no coupling between dof components; COO order bottom faces then top.
* meshgrid faces: hand-derived connectivity for ne = 2 (:94-120).
* inflate_sphere: radius = max(|dx|, |dy|), direction kept, centre -> origin.
* setboundaryCond: node counts and values on z = 0 / z = 1.
"""
import numpy as np
import pytest

import oracle as O

RECT_MASS = np.array([[4, 2, 1, 2], [2, 4, 2, 1], [1, 2, 4, 2], [2, 1, 2, 4]], dtype=np.float64) / 36.0


def one_face_mesh(hx, hy, z_btm=0.0, z_top=1.0, x0=0.3, y0=-0.2):
    """Two parallel hx x hy rectangles as one bottom and one top face."""
    pts = []
    for z in (z_btm, z_top):
        pts += [(x0, y0, z), (x0 + hx, y0, z), (x0 + hx, y0 + hy, z), (x0, y0 + hy, z)]
    NL = np.array(pts, dtype=np.float64).T
    ID = np.array([[3 * m + l + 1 for l in range(3)] for m in range(8)], dtype=np.int64)
    btm = np.array([[1, 2, 3, 4]], dtype=np.int64)
    top = np.array([[5, 6, 7, 8]], dtype=np.int64)
    return NL, ID, btm, top


@pytest.mark.parametrize("hx,hy", [(1.0, 1.0), (0.25, 0.125), (0.3, 0.7)])
def test_rectangle_face_is_the_consistent_mass_matrix(hx, hy):
    NL, ID, btm, top = one_face_mesh(hx, hy)
    b = O.apply_boundary_conditions(1, NL, None, top, btm, 3, "Q1", ID).toarray()
    assert b.shape == (24, 24)
    for face in (0, 1):
        for i in range(3):
            for j in range(3):
                blk = b[np.ix_(12 * face + 3 * np.arange(4) + i, 12 * face + 3 * np.arange(4) + j)]
                want = hx * hy * RECT_MASS if i == j else np.zeros((4, 4))
                assert np.allclose(blk, want, rtol=1e-14, atol=0), (face, i, j)
    assert np.all(b[:12, 12:] == 0) and np.all(b[12:, :12] == 0)


def test_coo_order_bottom_then_top():
    NL, IEN, ID, top, btm = O.meshgrid_ref(0, 1, 0, 1, 0, 1, 2, 3, with_faces=True)
    E, J, V = O.surface_coo(NL, btm, top, ID)
    nf = btm.shape[0]
    first = E[: nf * 144]
    assert set(((first - 1) // 3 + 1).tolist()) == set(btm.ravel().tolist())
    assert set(((E[nf * 144:] - 1) // 3 + 1).tolist()) == set(top.ravel().tolist())
    # [iNode][jNode][iDof][jDof] inside a face
    e0 = E[:144].reshape(4, 4, 3, 3)
    j0 = J[:144].reshape(4, 4, 3, 3)
    for a in range(4):
        for bb in range(4):
            for i in range(3):
                for j in range(3):
                    assert e0[a, bb, i, j] == ID[btm[0, a] - 1, i]
                    assert j0[a, bb, i, j] == ID[btm[0, bb] - 1, j]


def test_meshgrid_faces_known_answer():
    _, _, _, top, btm = O.meshgrid_ref(0, 1, 0, 1, 0, 1, 2, 3, with_faces=True)
    assert btm.tolist() == [[1, 2, 5, 4], [2, 3, 6, 5], [4, 5, 8, 7], [5, 6, 9, 8]]
    assert top.tolist() == [[19, 20, 23, 22], [20, 21, 24, 23], [22, 23, 26, 25], [23, 24, 27, 26]]
    _, IEN2, _, top2, btm2 = O.meshgrid_ref(0, 1, 0, 1, 0, 0, 3, 2, with_faces=True)
    assert btm2.tolist() == [[1, 2], [2, 3], [3, 4]]
    assert top2.tolist() == [[13, 14], [14, 15], [15, 16]]  # [IEN[n,4], IEN[n,3]] of row j = ne


def shoelace(xy):
    x, y = xy[0], xy[1]
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


@pytest.mark.parametrize("ne", [2, 4, 7])
def test_inflated_faces_sum_to_their_area(ne):
    NL, IEN, ID, top, btm = O.meshgrid_ref(0, 1, 0, 1, 0, 1, ne, 3, with_faces=True)
    cyl = O.inflate_sphere_ref(NL, 0, 1, 0, 1)
    b = O.apply_boundary_conditions(ne, cyl, IEN, top, btm, 3, "Q1", ID)
    area = sum(shoelace(cyl[:2, f - 1]) for f in np.vstack([btm, top]))
    for d in range(3):
        idx = np.arange(d, b.shape[0], 3)
        assert b[idx][:, idx].sum() == pytest.approx(area, rel=1e-13)
    # the inscribed polygon of the radius-0.5 circle: area below pi/4 per face
    assert area / 2 < np.pi / 4
    dense = b.toarray()
    assert np.array_equal(dense, dense.T)
    assert np.linalg.eigvalsh(dense).min() > -1e-14
    # no coupling between dof components
    comp = (np.arange(b.shape[0]) % 3)
    assert np.all(dense[comp[:, None] != comp[None, :]] == 0)


def test_inflate_sphere_geometry():
    NL, _, _ = O.meshgrid_ref(0, 1, 0, 1, 0, 1, 4, 3)
    cyl = O.inflate_sphere_ref(NL, 0, 1, 0, 1)
    dx, dy = NL[0] - 0.5, NL[1] - 0.5
    scale = np.maximum(np.abs(dx), np.abs(dy))
    r = np.hypot(cyl[0], cyl[1])
    assert np.allclose(r, scale, rtol=1e-15, atol=0)
    moved = scale > 0
    assert np.allclose(np.arctan2(cyl[1], cyl[0])[moved], np.arctan2(dy, dx)[moved], atol=1e-15)
    assert np.all(cyl[:2, ~moved] == 0)
    assert np.array_equal(cyl[2], NL[2])
    # the input is not modified
    assert NL[0].max() == 1.0


def test_setboundaryCond_z_faces():
    ne = 3
    NL, _, _ = O.meshgrid_ref(0, 1, 0, 1, 0, 1, ne, 3)
    q_d, fixed = O.setboundaryCond_ref(NL, ne, 3, 0.25)
    assert fixed.sum() == 2 * (ne + 1) ** 2
    assert np.all(fixed[0::3] == 0) and np.all(fixed[1::3] == 0)
    z = NL[2]
    assert np.all(q_d[2::3][z == 1] == -0.25) and np.all(q_d[2::3][z == 0] == 0)
    assert np.all(fixed[2::3][(z > 0) & (z < 1)] == 0)


def test_surface_dimension_errors():
    NL, IEN, ID, top, btm = O.meshgrid_ref(0, 1, 0, 1, 0, 0, 2, 2, with_faces=True)
    with pytest.raises(O.OracleError) as e:
        O.surface_coo(NL, btm, top, ID, nDof=2)
    assert e.value.code == -3
