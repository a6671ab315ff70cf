"""The oracle's restatement of the device problem generator (specs), pinned by
closed forms and known answers -- CPU only.

* the counter-based draw: at stream 0, index 0 it is the first output of
  SplitMix64 seeded with `seed` (published value for seed 1234567);
* the mesh is meshgrid's (examples/vector3D.jl:10-130 restated by
  ``meshgrid_ref``) for the unit cube;
* consistent loads of a uniform traction / body force on box meshes: every
  cell integrates N_a to V/8 (V/4 in 2D), every face to A/4 (edges L/2), so
  the loads are t h_u h_v x {1, 1/2, 1/4} on the face and b V_cell x
  {1, 1/2, 1/4, 1/8}, and sum to t x area and b x volume;
* the moduli are log-uniform in [lo, hi], the noise N(0, sigma^2) on the free
  dofs of the loaded face only, the clamp exactly the chosen face.
"""
import numpy as np
import pytest

import oracle as O


def spec(**kw):
    base = dict(ndim=3, dims=(6, 4, 5), box=(0.0, 3.0, 0.0, 2.0, 0.0, 2.5), young_lo=0.1, young_hi=10.0,
                nu=0.3, seed=99, clamp_axis=0, clamp_side=0, load_axis=0, load_side=1,
                traction=(0.0, 0.0, 0.0), body=(0.0, 0.0, 0.0), noise=0.0)
    base.update(kw)
    return base


def test_draw_is_splitmix64():
    # SplitMix64 (Steele, Lea, Flood 2014) seeded with 1234567: first output
    z = 6457827717110365317
    u = O.spec_u01(1234567, 0, np.array([0], dtype=np.uint64))[0]
    assert u == ((z >> 12) + 0.5) * 2.0 ** -52
    v = O.spec_u01(5, 1, np.arange(100000, dtype=np.uint64))
    assert v.min() > 0 and v.max() < 1 and abs(v.mean() - 0.5) < 0.005


def test_mesh_is_meshgrid():
    ne = 3
    r = O.spec_inputs(spec(dims=(ne, ne, ne), box=(0, 1, 0, 1, 0, 1)))
    NL, IEN, ID = O.meshgrid_ref(0, 1, 0, 1, 0, 1, ne, 3)
    assert np.array_equal(r["NodeList"], NL) and np.array_equal(r["IEN"], IEN) and np.array_equal(r["ID"], ID)
    r2 = O.spec_inputs(spec(ndim=2, dims=(ne, ne, 0), box=(0, 1, 0, 1, 0, 0)))
    NL, IEN, ID = O.meshgrid_ref(0, 1, 0, 1, 0, 0, ne, 2)
    assert np.array_equal(r2["NodeList"], NL) and np.array_equal(r2["IEN"], IEN)


@pytest.mark.parametrize("load_axis,load_side", [(0, 1), (1, 0), (2, 1)])
def test_traction_closed_form_3d(load_axis, load_side):
    t = np.array([0.3, -1.5, 2.0])
    sp = spec(load_axis=load_axis, load_side=load_side, traction=tuple(t), young_lo=1.0, young_hi=1.0)
    r = O.spec_inputs(sp)
    dims, box = sp["dims"], sp["box"]
    h = [(box[2 * d + 1] - box[2 * d]) / dims[d] for d in range(3)]
    other = [d for d in range(3) if d != load_axis]
    area = np.prod([box[2 * d + 1] - box[2 * d] for d in other])
    f = r["f"].reshape(-1, 3)
    assert np.allclose(f.sum(0), t * area, rtol=1e-14, atol=0)
    nl = r["NodeList"]
    on = np.isclose(nl[load_axis], box[2 * load_axis + load_side])
    assert np.all(f[~on] == 0)
    # per node: t h_u h_v times 1 (interior), 1/2 (edge), 1/4 (corner)
    share = np.ones(nl.shape[1])
    for d in other:
        edge = np.isclose(nl[d], box[2 * d]) | np.isclose(nl[d], box[2 * d + 1])
        share = share * np.where(edge, 0.5, 1.0)
    expect = np.where(on, share * h[other[0]] * h[other[1]], 0.0)[:, None] * t[None, :]
    assert np.allclose(f, expect, rtol=1e-13, atol=1e-16)


def test_body_force_closed_form():
    b = np.array([0.0, 2.0, -1.0])
    sp = spec(body=tuple(b), young_lo=1.0, young_hi=1.0)
    r = O.spec_inputs(sp)
    dims, box = sp["dims"], sp["box"]
    vol = np.prod([box[2 * d + 1] - box[2 * d] for d in range(3)])
    vcell = vol / np.prod(dims)
    f = r["f"].reshape(-1, 3)
    assert np.allclose(f.sum(0), b * vol, rtol=1e-14)
    nl = r["NodeList"]
    cells = np.ones(nl.shape[1])  # cells around each node
    for d in range(3):
        edge = np.isclose(nl[d], box[2 * d]) | np.isclose(nl[d], box[2 * d + 1])
        cells = cells * np.where(edge, 1.0, 2.0)
    assert np.allclose(f, (cells * vcell / 8)[:, None] * b[None, :], rtol=1e-13)


def test_2d_edge_traction_and_body():
    sp = spec(ndim=2, dims=(5, 4, 0), box=(0.0, 2.0, 0.0, 1.0, 0, 0), load_axis=1, load_side=1,
              traction=(0.5, -1.0, 0.0), body=(0.0, 3.0, 0.0), young_lo=2.0, young_hi=2.0)
    r = O.spec_inputs(sp)
    f = r["f"].reshape(-1, 2)
    # traction on y = 1 (length 2) plus body over the 2 x 1 rectangle
    assert np.allclose(f.sum(0), np.array([0.5, -1.0]) * 2.0 + np.array([0.0, 3.0]) * 2.0, rtol=1e-14)
    assert np.all(r["young"] == 2.0)


def test_young_clamp_and_noise():
    sp = spec(dims=(10, 40, 40), box=(0, 2, 0, 1, 0, 1), noise=1.5, seed=123)
    r = O.spec_inputs(sp)
    y = r["young"]
    assert y.min() >= 0.1 and y.max() <= 10.0
    ly = np.log10(y)
    assert abs(ly.mean()) < 0.05 and abs(ly.std() - 2 / np.sqrt(12)) < 0.02  # log-uniform on [-1, 1]
    nl = r["NodeList"]
    clamped = np.repeat(nl[0] == 0.0, 3)
    assert np.array_equal(r["fixed"] == 1, clamped)
    on = np.repeat(nl[0] == 2.0, 3)
    f = r["f"]
    assert np.all(f[~on] == 0.0)
    # 5,043 draws: 4 standard errors
    assert abs(f[on].std() - 1.5) < 0.06 and abs(f[on].mean()) < 0.085
    # another seed draws other values; the same seed the same
    r2 = O.spec_inputs(dict(sp, seed=124))
    assert not np.array_equal(r2["young"], y)
    assert np.array_equal(O.spec_inputs(sp)["f"], f)
