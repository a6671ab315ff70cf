// basis.hpp -- Q1 (Q4 / Hex8) and Q2 (9-node) master-element basis functions and
// the 2-point Gauss rule, shared by the host utilities and the element kernel so
// both evaluate identical expressions.
//
// Formulas and node order: smearFEM src/fem.jl:53-110 (basis_function).
// Gauss points and their order: src/fem.jl:21-31 and :158-176.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "smearfem.h"

namespace sfem {

#define SFEM_HD __host__ __device__ __forceinline__

// Corner sign pattern of the Q1 nodes: 2D CCW (-1,-1),(1,-1),(1,1),(-1,1); 3D the
// same square on zeta=-1 then on zeta=+1.
SFEM_HD double q1_sx(int a) { return (a & 3) == 1 || (a & 3) == 2 ? 1.0 : -1.0; }
SFEM_HD double q1_sy(int a) { return (a & 3) >= 2 ? 1.0 : -1.0; }
SFEM_HD double q1_sz(int a) { return a >= 4 ? 1.0 : -1.0; }

// 2-point rule on [-1,1]: -(b-a)/(2*sqrt(3)) + (b+a)/2 with a=-1, b=1.
SFEM_HD double gauss2_abscissa() { return 2.0 / (2.0 * sqrt(3.0)); }

// Gauss point g (0-based) of the tensor rule in the reference's order:
// 2D (x1,y1),(x2,y1),(x2,y2),(x1,y2); 3D the same four at z1 then at z2.
// All weights are 1.0.
SFEM_HD void gauss_point(int g, double& xi, double& eta, double& zeta) {
    const double s = gauss2_abscissa();
    xi = q1_sx(g) * s;
    eta = q1_sy(g) * s;
    zeta = q1_sz(g) * s;
}

// ---- Q1 ----------------------------------------------------------------
SFEM_HD double q1_2d_N(int a, double x, double y) {
    return (1 + q1_sx(a) * x) * (1 + q1_sy(a) * y) / 4;
}
SFEM_HD void q1_2d_dN(int a, double x, double y, double* d) {
    d[0] = q1_sx(a) * (1 + q1_sy(a) * y) / 4;
    d[1] = q1_sy(a) * (1 + q1_sx(a) * x) / 4;
}
SFEM_HD double q1_3d_N(int a, double x, double y, double z) {
    return (1 + q1_sx(a) * x) * (1 + q1_sy(a) * y) * (1 + q1_sz(a) * z) / 8;
}
SFEM_HD void q1_3d_dN(int a, double x, double y, double z, double* d) {
    const double fx = 1 + q1_sx(a) * x, fy = 1 + q1_sy(a) * y, fz = 1 + q1_sz(a) * z;
    d[0] = q1_sx(a) * (fy * fz) / 8;
    d[1] = q1_sy(a) * (fx * fz) / 8;
    d[2] = q1_sz(a) * (fx * fy) / 8;
}

// ---- Q2 (9 nodes: corners CCW, mid-sides bottom,right,top,left, centre) --
SFEM_HD double q2_2d_N(int a, double x, double y) {
    switch (a) {
        case 0: return (1 - x) * x * (1 - y) * y / 4;
        case 1: return -x * (1 + x) * (1 - y) * y / 4;
        case 2: return x * (1 + x) * y * (1 + y) / 4;
        case 3: return -(1 - x) * x * y * (1 + y) / 4;
        case 4: return -(1 - x) * (1 + x) * (1 - y) * y / 2;
        case 5: return x * (1 + x) * (1 - y) * (1 + y) / 2;
        case 6: return (1 - x) * (1 + x) * y * (1 + y) / 2;
        case 7: return -(1 - x) * x * (1 - y) * (1 + y) / 2;
        default: return (1 - x) * (1 + x) * (1 - y) * (1 + y);
    }
}
SFEM_HD void q2_2d_dN(int a, double x, double y, double* d) {
    switch (a) {
        case 0: d[0] = (1 - 2 * x) * (1 - y) * y / 4;      d[1] = (1 - x) * x * (1 - 2 * y) / 4; break;
        case 1: d[0] = -(1 + 2 * x) * (1 - y) * y / 4;     d[1] = -x * (1 + x) * (1 - 2 * y) / 4; break;
        case 2: d[0] = (1 + 2 * x) * y * (1 + y) / 4;      d[1] = x * (1 + x) * (1 + 2 * y) / 4; break;
        case 3: d[0] = -(1 - 2 * x) * y * (1 + y) / 4;     d[1] = -(1 - x) * x * (1 + 2 * y) / 4; break;
        case 4: d[0] = x * (1 - y) * y;                    d[1] = -(1 - x) * (1 + x) * (1 - 2 * y) / 2; break;
        case 5: d[0] = (1 + 2 * x) * (1 - y) * (1 + y) / 2; d[1] = -x * (1 + x) * y; break;
        case 6: d[0] = -x * y * (1 + y);                   d[1] = (1 - x) * (1 + x) * (1 + 2 * y) / 2; break;
        case 7: d[0] = -(1 - 2 * x) * (1 - y) * (1 + y) / 2; d[1] = (1 - x) * x * y; break;
        default: d[0] = -2 * x * (1 - y) * (1 + y);        d[1] = -(1 - x) * (1 + x) * 2 * y; break;
    }
}

// Number of basis functions the reference returns for (class, coordinates
// given); 0 = the reference leaves N undefined (UndefVarError).
SFEM_HD int basis_count(int cls, int nargs) {
    if (cls == SFEM_Q1) return nargs >= 3 ? 8 : (nargs == 2 ? 4 : 2);
    if (cls == SFEM_Q2) return nargs >= 2 ? 9 : 0;
    return 0;
}

// dN/dxi_j of node a of the element kind <ndim,nen> at (x,y,z); d gets ndim values.
template <int NDIM, int NEN>
SFEM_HD void dshape(int a, double x, double y, double z, double* d) {
    if constexpr (NDIM == 3) {
        q1_3d_dN(a, x, y, z, d);
    } else if constexpr (NEN == 9) {
        q2_2d_dN(a, x, y, d);
    } else {
        q1_2d_dN(a, x, y, d);
    }
}

}  // namespace sfem
