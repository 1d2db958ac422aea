// Element frame of the constant-strain membrane triangle, shared by the element kernels (reference
// geometry) and the updated-geometry kernels (per-sample geometry X + u):
//   e1 = (X2-X1)/|X2-X1|, e2 = normalised (X0-X2) minus its e1 component, e3 = e1 x e2
//   (optimizedSolidAssemble.pyx:82-114, :606-651); local coordinates x_a = e1.X_a, y_a = e2.X_a
//   and the area from the 3x3 determinant (pyx:274-284).
#pragma once

#include "common.h"

namespace cvfes {

struct Frame {
    double t[3][3];   // rows e1, e2, e3
    double lx[3], ly[3];
    double area;
};

// RECIP: normalise with one reciprocal per vector (x * (1/|v|)) instead of three divisions; the
// components then differ from x / |v| by at most an ulp.
template <bool RECIP = false>
__device__ __forceinline__ Frame frame_from_points(const double (&X)[3][3]) {
    Frame f;
    double e0[3], e1[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        e0[k] = X[2][k] - X[1][k];
        e1[k] = X[0][k] - X[2][k];
    }
    double nrm = sqrt(e0[0] * e0[0] + e0[1] * e0[1] + e0[2] * e0[2]);
    if constexpr (RECIP) {
        const double inv = 1.0 / nrm;
#pragma unroll
        for (int k = 0; k < 3; ++k) f.t[0][k] = e0[k] * inv;
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) f.t[0][k] = e0[k] / nrm;
    }
    const double proj = e1[0] * f.t[0][0] + e1[1] * f.t[0][1] + e1[2] * f.t[0][2];
#pragma unroll
    for (int k = 0; k < 3; ++k) f.t[1][k] = e1[k] - proj * f.t[0][k];
    nrm = sqrt(f.t[1][0] * f.t[1][0] + f.t[1][1] * f.t[1][1] + f.t[1][2] * f.t[1][2]);
    if constexpr (RECIP) {
        const double inv = 1.0 / nrm;
#pragma unroll
        for (int k = 0; k < 3; ++k) f.t[1][k] = f.t[1][k] * inv;
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) f.t[1][k] = f.t[1][k] / nrm;
    }
    f.t[2][0] = f.t[0][1] * f.t[1][2] - f.t[0][2] * f.t[1][1];
    f.t[2][1] = f.t[0][2] * f.t[1][0] - f.t[0][0] * f.t[1][2];
    f.t[2][2] = f.t[0][0] * f.t[1][1] - f.t[0][1] * f.t[1][0];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        f.lx[a] = X[a][0] * f.t[0][0] + X[a][1] * f.t[0][1] + X[a][2] * f.t[0][2];
        f.ly[a] = X[a][0] * f.t[1][0] + X[a][1] * f.t[1][1] + X[a][2] * f.t[1][2];
    }
    f.area = ((f.lx[1] * f.ly[2] - f.lx[2] * f.ly[1]) - (f.lx[0] * f.ly[2] - f.lx[2] * f.ly[0]) +
              (f.lx[0] * f.ly[1] - f.lx[1] * f.ly[0])) * 0.5;
    return f;
}

__device__ __forceinline__ Frame element_frame(const double *__restrict__ nodes, int n0, int n1, int n2) {
    double X[3][3];
    const int ids[3] = {n0, n1, n2};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) X[a][k] = nodes[3 * static_cast<int64_t>(ids[a]) + k];
    return frame_from_points(X);
}

}  // namespace cvfes
