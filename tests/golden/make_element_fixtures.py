"""Write exact element-stiffness fixtures with sympy.

The stiffness of a rectangle / box element is integrated exactly (sympy
``integrate`` of the bilinear form over the element), starting from the
textbook Lagrange shape functions, which sympy differentiates itself.  For Q4
and Hex8 on rectangles / boxes the reference's 2x2(x2) Gauss rule is exact,
so these are the values the reference must produce up to rounding.  For the
9-node Q2 (Laplace only) the reference's 2x2 rule is not exact, so that
fixture is the exact 2x2 Gauss sum in sympy's rational/sqrt arithmetic.

Run:  python tests/golden/make_element_fixtures.py   (writes element_fixtures.json)
"""
from __future__ import annotations

import json
from pathlib import Path

import sympy as sp

xi, eta, zeta = sp.symbols("xi eta zeta")
OUT = Path(__file__).resolve().parent / "element_fixtures.json"


def q1_shapes(ndim):
    # vertex order: counter-clockwise (-1,-1),(1,-1),(1,1),(-1,1), bottom then top
    sq = [(-1, -1), (1, -1), (1, 1), (-1, 1)]
    if ndim == 2:
        return [(1 + sx * xi) * (1 + sy * eta) / 4 for sx, sy in sq], [(sx, sy) for sx, sy in sq]
    verts = [(sx, sy, sz) for sz in (-1, 1) for sx, sy in sq]
    return [(1 + sx * xi) * (1 + sy * eta) * (1 + sz * zeta) / 8 for sx, sy, sz in verts], verts


def q2_shapes():
    # 1D quadratic Lagrange on nodes -1, 0, 1
    l = {-1: xi * (xi - 1) / 2, 0: (1 - xi) * (1 + xi), 1: xi * (xi + 1) / 2}
    m = {-1: eta * (eta - 1) / 2, 0: (1 - eta) * (1 + eta), 1: eta * (eta + 1) / 2}
    nodes = [(-1, -1), (1, -1), (1, 1), (-1, 1), (0, -1), (1, 0), (0, 1), (-1, 0), (0, 0)]
    return [sp.expand(l[a] * m[b]) for a, b in nodes], nodes


def grad_phys(N, ndim, h):
    # box element [0,h0]x[0,h1](x[0,h2]); x_d = h_d (xi_d + 1)/2 -> d/dx_d = 2/h_d d/dxi_d
    vars_ = [xi, eta, zeta][:ndim]
    return [[sp.diff(Na, v) * 2 / h[d] for d, v in enumerate(vars_)] for Na in N]


def integrate_ref(expr, ndim, h):
    det = sp.Mul(*[hd / 2 for hd in h])
    res = expr * det
    for v in [xi, eta, zeta][:ndim]:
        res = sp.integrate(res, (v, -1, 1))
    return sp.nsimplify(res)


def gauss2_ref(expr, ndim, h):
    g = 1 / sp.sqrt(3)
    det = sp.Mul(*[hd / 2 for hd in h])
    pts = [(-g, g)] * ndim
    tot = 0
    import itertools

    for p in itertools.product(*pts):
        tot += expr.subs(dict(zip([xi, eta, zeta][:ndim], p)))
    return sp.nsimplify(sp.simplify(tot * det))


def laplace_ke(N, ndim, h, rule):
    G = grad_phys(N, ndim, h)
    n = len(N)
    K = [[None] * n for _ in range(n)]
    for a in range(n):
        for b in range(a, n):
            e = sp.expand(sum(G[a][k] * G[b][k] for k in range(ndim)))
            K[a][b] = K[b][a] = rule(e, ndim, h)
    return K


def elastic_ke(N, ndim, h, E, nu):
    G = grad_phys(N, ndim, h)
    n = len(N)
    L = n * ndim
    if ndim == 2:  # plane stress, Voigt [exx, eyy, gxy]
        c = E / (1 - nu ** 2)
        D = sp.Matrix([[c, nu * c, 0], [nu * c, c, 0], [0, 0, E / (2 * (1 + nu))]])
        B = sp.zeros(3, L)
        for a in range(n):
            gx, gy = G[a]
            B[0, 2 * a] = gx
            B[1, 2 * a + 1] = gy
            B[2, 2 * a] = gy
            B[2, 2 * a + 1] = gx
    else:  # Voigt [exx, eyy, ezz, gyz, gxz, gxy]
        c = E / ((1 + nu) * (1 - 2 * nu))
        D = sp.zeros(6, 6)
        for i in range(3):
            for j in range(3):
                D[i, j] = c * ((1 - nu) if i == j else nu)
        for i in range(3, 6):
            D[i, i] = c * (1 - 2 * nu) / 2
        B = sp.zeros(6, L)
        for a in range(n):
            gx, gy, gz = G[a]
            B[0, 3 * a] = gx
            B[1, 3 * a + 1] = gy
            B[2, 3 * a + 2] = gz
            B[3, 3 * a + 1] = gz
            B[3, 3 * a + 2] = gy
            B[4, 3 * a] = gz
            B[4, 3 * a + 2] = gx
            B[5, 3 * a] = gy
            B[5, 3 * a + 1] = gx
    integrand = (B.T * D * B).applyfunc(sp.expand)
    K = [[None] * L for _ in range(L)]
    for i in range(L):
        for j in range(i, L):
            K[i][j] = K[j][i] = integrate_ref(integrand[i, j], ndim, h)
    return K


def box_coords(verts, h):
    # physical coordinates of the vertices, ndim x nen
    ndim = len(h)
    return [[float(h[d] * (v[d] + 1) / 2) for v in verts] for d in range(ndim)]


def main():
    fixtures = []
    E = sp.Rational(5, 2)
    nu = sp.Rational(3, 10)
    h2 = (sp.Rational(1, 2), sp.Rational(3, 4))
    h3 = (sp.Rational(1, 2), sp.Integer(1), sp.Rational(3, 4))

    N, verts = q1_shapes(2)
    fixtures.append({"name": "q4_plane_stress_rect", "ndim": 2, "cls": "Q1", "ndof": 2, "E": float(E),
                     "nu": float(nu), "coords": box_coords(verts, h2),
                     "Ke": [[float(v) for v in r] for r in elastic_ke(N, 2, h2, E, nu)]})
    fixtures.append({"name": "q4_laplace_rect", "ndim": 2, "cls": "Q1", "ndof": 1, "E": 1.0, "nu": 0.3,
                     "coords": box_coords(verts, h2),
                     "Ke": [[float(v) for v in r] for r in laplace_ke(N, 2, h2, integrate_ref)]})
    N, verts = q1_shapes(3)
    fixtures.append({"name": "hex8_elastic_box", "ndim": 3, "cls": "Q1", "ndof": 3, "E": float(E),
                     "nu": float(nu), "coords": box_coords(verts, h3),
                     "Ke": [[float(v) for v in r] for r in elastic_ke(N, 3, h3, E, nu)]})
    fixtures.append({"name": "hex8_laplace_box", "ndim": 3, "cls": "Q1", "ndof": 1, "E": 1.0, "nu": 0.3,
                     "coords": box_coords(verts, h3),
                     "Ke": [[float(v) for v in r] for r in laplace_ke(N, 3, h3, integrate_ref)]})
    N, verts = q2_shapes()
    fixtures.append({"name": "q2_laplace_rect_gauss2", "ndim": 2, "cls": "Q2", "ndof": 1, "E": 1.0, "nu": 0.3,
                     "coords": box_coords(verts, h2),
                     "Ke": [[float(v) for v in r] for r in laplace_ke(N, 2, h2, gauss2_ref)]})
    OUT.write_text(json.dumps({"generator": "tests/golden/make_element_fixtures.py", "fixtures": fixtures},
                              indent=1))
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
