// gt_ftpatch.h -- per-patch culling of the Fourier grid evaluation.
//
// An FT-space atom contributes I exp(-|U^T k|^2 / 2) exp(-i x . k) where
// |U^T k|^2 < 30 (GaussDictCode/bin/FTGaussian3D.py:95-110, 151-164): its
// support is an ellipsoid about k = 0.  ft_patch_min bounds |U^T k|^2 from
// below over a rectangle of (k1, k2) at fixed k0, so that a patch of samples
// the support cannot reach is skipped.  Host-callable for the brute-force
// check in tests/native.
#pragma once

#include <hip/hip_runtime.h>

namespace gt {

struct FtL {  // the staged part of an FtRec (64 bytes)
    float amp, xh[3], xl[3], u[6], iso, pad[2];
};
static_assert(sizeof(FtL) == 64, "FtL must stay 64 bytes");

// Smallest scaled |U^T k|^2 over k0 fixed and (k1, k2) in [a1, b1] x [a2, b2]
// (U upper triangular in the slot order of FtRec.u, or iso: u[0] / r).
__host__ __device__ inline float ft_patch_min(const FtL &r, float k0, float a1, float b1, float a2, float b2)
{
    if (r.iso != 0.0f) {
        const float m1 = (a1 <= 0.0f && b1 >= 0.0f) ? 0.0f : fminf(a1 * a1, b1 * b1);
        const float m2 = (a2 <= 0.0f && b2 >= 0.0f) ? 0.0f : fminf(a2 * a2, b2 * b2);
        return (fmaf(k0, k0, m1) + m2) * (r.u[0] * r.u[0]);
    }
    // v0 = U00 k0, v1 = U01 k0 + U11 k1, v2 = U02 k0 + U12 k1 + U22 k2
    const float u00 = r.u[0], u11 = r.u[1], u22 = r.u[2], u01 = r.u[3], u12 = r.u[4], u02 = r.u[5];
    const float v0 = u00 * k0, al1 = u01 * k0, al2 = u02 * k0;
    if (!(u11 != 0.0f) || !(u22 != 0.0f)) return 0.0f;  // degenerate: keep
    const float k1s = -al1 / u11, k2s = -fmaf(u12, k1s, al2) / u22;
    if (k1s >= a1 && k1s <= b1 && k2s >= a2 && k2s <= b2) return v0 * v0;
    float best = 3.0e38f;
    const float A = fmaf(u11, u11, u12 * u12);
#pragma unroll
    for (int e = 0; e < 2; e++) {
        // edge k1 = c: v1 fixed, v2 = w + U22 k2 minimised over k2
        const float c = e ? b1 : a1;
        const float v1 = fmaf(u11, c, al1), w = fmaf(u12, c, al2);
        const float k2 = fminf(fmaxf(-w / u22, a2), b2);
        const float v2 = fmaf(u22, k2, w);
        best = fminf(best, fmaf(v1, v1, v2 * v2));
        // edge k2 = d: (al1 + U11 k1)^2 + (w2 + U12 k1)^2 minimised over k1
        const float d = e ? b2 : a2;
        const float w2 = fmaf(u22, d, al2);
        const float k1 = fminf(fmaxf(-fmaf(al1, u11, w2 * u12) / A, a1), b1);
        const float p1 = fmaf(u11, k1, al1), p2 = fmaf(u12, k1, w2);
        best = fminf(best, fmaf(p1, p1, p2 * p2));
    }
    return fmaf(v0, v0, best);
}

}  // namespace gt
