// block_ops.hpp -- per-value and per-node-block helpers shared by the
// single-mesh elimination and the batch builder: the 16-bit correction of
// the split operator format and the inverse of a node's diagonal block.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sfem {

// 16-bit correction of the FP32 rounding h of v: round((v - h) / (ulp(h) 2^-16)),
// so that h + q ulp(h) 2^-16 is v to 2^-39 relative.  0 where h is 0, not
// finite, or too small for the scale to be a normal float (|h| < 2^-87).
__device__ __forceinline__ int16_t split_correction(double v) {
    const float h = (float)v;
    const int e = (__float_as_int(h) >> 23) & 0xFF;
    if (e <= 39 || e == 255) return 0;
    const float scale = __int_as_float((__float_as_int(h) & 0x7F800000) - (39 << 23));
    const double t = rint((v - (double)h) / (double)scale);
    return (int16_t)fmin(fmax(t, -32767.0), 32767.0);
}

// Inverse of a node's diagonal block M (M[j][i] = m[j*BS+i]), symmetric by
// construction (upper triangle mirrored).  A singular block falls back to
// the inverse of its diagonal (1 where that is 0).
template <int BS>
__device__ inline void invert_block(const double* m, double* inv) {
    double det = 0.0;
    if (BS == 3) {
        const double a = m[0], b = m[1], c = m[2], e = m[4], f = m[5], i = m[8];
        const double A = e * i - f * f, B = c * f - b * i, C = b * f - c * e;
        det = a * A + b * B + c * C;
        if (det != 0.0 && isfinite(det)) {
            const double r = 1.0 / det;
            inv[0] = A * r;
            inv[1] = inv[3] = B * r;
            inv[2] = inv[6] = C * r;
            inv[4] = (a * i - c * c) * r;
            inv[5] = inv[7] = (b * c - a * f) * r;
            inv[8] = (a * e - b * b) * r;
            return;
        }
    } else if (BS == 2) {
        const double a = m[0], b = m[1], d = m[3];
        det = a * d - b * b;
        if (det != 0.0 && isfinite(det)) {
            const double r = 1.0 / det;
            inv[0] = d * r;
            inv[1] = inv[2] = -b * r;
            inv[3] = a * r;
            return;
        }
    }
    for (int j = 0; j < BS; ++j)
        for (int i = 0; i < BS; ++i) inv[j * BS + i] = (i == j && m[j * BS + j] != 0.0) ? 1.0 / m[j * BS + j] : 0.0;
    for (int j = 0; j < BS; ++j)
        if (inv[j * BS + j] == 0.0) inv[j * BS + j] = 1.0;
}

}  // namespace sfem
